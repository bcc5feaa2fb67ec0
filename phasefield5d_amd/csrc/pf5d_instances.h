// The large kernels are compiled in translation units of their own
// (pf5d_kern_fft32.hip, pf5d_kern_fft64.hip, pf5d_kern_stencil.hip), in
// parallel with the host unit.
// The host unit sees these instances as explicit instantiation declarations
// (PF5D_INST = extern template), so it neither compiles nor inlines them; the
// kernel units define them (PF5D_INST = template).
#pragma once
#include "pf5d_fft.h"
#include "pf5d_fftplane.h"
#include "pf5d_fftreg.h"
#include "pf5d_fused.h"
#include "pf5d_split.h"

#ifndef PF5D_INST
#define PF5D_INST extern template
#endif

namespace pf5d {

#define PF5D_FFT_INSTANCE(T, NT)                                                                               \
  PF5D_INST __global__ void k_fft<T, NT>(const FftItem* __restrict__, int, const LinePlanDev* __restrict__, \
                                         const cx<T>* __restrict__, const cx<T>* __restrict__,                 \
                                         const SimState* __restrict__);
#define PF5D_FFTP_INSTANCE(T, NT, N, P)                                                                    \
  PF5D_INST __global__ void k_fftp_fwd<T, NT, N, P>(const FftPlaneItem* __restrict__, int,                 \
                                                    const LinePlanDev* __restrict__, const cx<T>* __restrict__, \
                                                    const SimState* __restrict__);                        \
  PF5D_INST __global__ void k_fftp_inv<T, NT, N, P>(const FftPlaneItem* __restrict__, int,                 \
                                                    const LinePlanDev* __restrict__, const cx<T>* __restrict__, \
                                                    const SimState* __restrict__);
#define PF5D_FRP_INSTANCE(T, NW, P)                                                                      \
  PF5D_INST __global__ void k_frp_fwd<T, NW, P>(const FftPlaneItem* __restrict__, int,                  \
                                                const LinePlanDev* __restrict__, const cx<T>* __restrict__, \
                                                const SimState* __restrict__);                         \
  PF5D_INST __global__ void k_frp_inv<T, NW, P>(const FftPlaneItem* __restrict__, int,                  \
                                                const LinePlanDev* __restrict__, const cx<T>* __restrict__, \
                                                const SimState* __restrict__);
#define PF5D_FRX_INSTANCE(T, NW)                                                                           \
  PF5D_INST __global__ void k_frx<T, NW>(const FrXItem* __restrict__, int, const LinePlanDev* __restrict__, \
                                         const cx<T>* __restrict__, const SimState* __restrict__);
// the wave-resident passes of 128-point planes and x lines: (waves per block, parts per plane)
#define PF5D_FR_INSTANCES         \
  PF5D_FRP_INSTANCE(float, 8, 2)  \
  PF5D_FRP_INSTANCE(float, 9, 2)  \
  PF5D_FRP_INSTANCE(double, 4, 2) \
  PF5D_FRX_INSTANCE(float, 1)     \
  PF5D_FRX_INSTANCE(double, 1)
#define PF5D_POTENTIALS_INSTANCE(T, NT, MINW)                                                                   \
  PF5D_INST __global__ void k_potentials<T, NT, MINW, false>(const SimDev<T>* __restrict__,                       \
                                                             const SimState* __restrict__, const K1Item* __restrict__, \
                                                             int);                                               \
  PF5D_INST __global__ void k_potentials<T, NT, MINW, true>(const SimDev<T>* __restrict__,                        \
                                                            const SimState* __restrict__, const K1Item* __restrict__,  \
                                                            int);
#define PF5D_FLUXM_INSTANCE(T, NT, CPL, MINW)                                                              \
  PF5D_INST __global__ void k_fluxmarch<T, NT, CPL, MINW>(const SimDev<T>* __restrict__, SimState* __restrict__, \
                                                    T* __restrict__, const K1Item* __restrict__, int,       \
                                                    const double* __restrict__, const int* __restrict__, int, \
                                                    const int* __restrict__, unsigned* __restrict__);
#define PF5D_SMARCH_INSTANCE(T, NT, MINW)                                                                   \
  PF5D_INST __global__ void k_stencilmarch<T, NT, MINW>(const SimDev<T>* __restrict__, SimState* __restrict__,  \
                                                        T* __restrict__, const K1Item* __restrict__, int,       \
                                                        const double* __restrict__, const int* __restrict__, int, \
                                                        const int* __restrict__, unsigned* __restrict__);
#define PF5D_FLUX_INSTANCE(T, NT, LOADS)                                                                       \
  PF5D_INST __global__ void k_fluxdiv<T, NT, LOADS>(const SimDev<T>* __restrict__, const SimState* __restrict__, \
                                                    T* __restrict__, const K1Item* __restrict__, int);

#define PF5D_FFT32_INSTANCES      \
  PF5D_FFT_INSTANCE(float, 64)    \
  PF5D_FFT_INSTANCE(float, 128)   \
  PF5D_FFT_INSTANCE(float, 256)   \
  PF5D_FFT_INSTANCE(float, 512)
#define PF5D_FFT64_INSTANCES      \
  PF5D_FFT_INSTANCE(double, 64)   \
  PF5D_FFT_INSTANCE(double, 128)  \
  PF5D_FFT_INSTANCE(double, 256)  \
  PF5D_FFT_INSTANCE(double, 512)

// plane passes: block sizes, plane lengths (128; 64 and 96 with fewer
// settings) and part counts (the FP64 inverse of 128 with 2 parts needs more LDS
// than a block has; the host then runs the line passes)
#define PF5D_FFTP_INSTANCES_T(T)     \
  PF5D_FFTP_INSTANCE(T, 256, 128, 2)  \
  PF5D_FFTP_INSTANCE(T, 256, 128, 4)  \
  PF5D_FFTP_INSTANCE(T, 256, 128, 8)  \
  PF5D_FFTP_INSTANCE(T, 512, 128, 2)  \
  PF5D_FFTP_INSTANCE(T, 512, 128, 4)  \
  PF5D_FFTP_INSTANCE(T, 512, 128, 8)  \
  PF5D_FFTP_INSTANCE(T, 512, 64, 2)   \
  PF5D_FFTP_INSTANCE(T, 512, 64, 4)   \
  PF5D_FFTP_INSTANCE(T, 1024, 64, 2)  \
  PF5D_FFTP_INSTANCE(T, 1024, 64, 4)  \
  PF5D_FFTP_INSTANCE(T, 512, 96, 2)   \
  PF5D_FFTP_INSTANCE(T, 512, 96, 4)   \
  PF5D_FFTP_INSTANCE(T, 1024, 96, 2)  \
  PF5D_FFTP_INSTANCE(T, 1024, 96, 4)
// (128-point planes with 1024 threads in FP32 only: held to 128 VGPRs, the FP64
// ones spill)
#define PF5D_FFTP_INSTANCES                 \
  PF5D_FFTP_INSTANCES_T(float)              \
  PF5D_FFTP_INSTANCES_T(double)             \
  PF5D_FFTP_INSTANCE(float, 1024, 128, 2)   \
  PF5D_FFTP_INSTANCE(float, 1024, 128, 4)   \
  PF5D_FFTP_INSTANCE(float, 1024, 128, 8)

// the stencil forms the engine offers the autotuner
#define PF5D_STENCIL_INSTANCES                    \
  PF5D_POTENTIALS_INSTANCE(double, 256, 2)        \
  PF5D_POTENTIALS_INSTANCE(float, 256, 3)         \
  PF5D_POTENTIALS_INSTANCE(float, 256, 4)         \
  PF5D_FLUX_INSTANCE(double, 256, 0)              \
  PF5D_FLUX_INSTANCE(double, 256, 1)              \
  PF5D_FLUX_INSTANCE(double, 256, 2)              \
  PF5D_FLUX_INSTANCE(float, 256, 0)               \
  PF5D_FLUX_INSTANCE(float, 256, 1)               \
  PF5D_FLUX_INSTANCE(float, 256, 2)               \
  PF5D_FLUXM_INSTANCE(double, 256, 1, 1)          \
  PF5D_FLUXM_INSTANCE(double, 256, 1, 3)          \
  PF5D_FLUXM_INSTANCE(float, 256, 1, 1)           \
  PF5D_FLUXM_INSTANCE(float, 256, 2, 1)           \
  PF5D_FLUXM_INSTANCE(float, 256, 2, 3)           \
  PF5D_FLUXM_INSTANCE(float, 512, 2, 1)           \
  PF5D_SMARCH_INSTANCE(float, 512, 2)

}  // namespace pf5d
