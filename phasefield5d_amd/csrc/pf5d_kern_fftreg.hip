// Instances of the wave-resident 128-point FFT passes (see pf5d_instances.h)
#define PF5D_INST template
#include "pf5d_instances.h"

namespace pf5d {
PF5D_FR_INSTANCES
}  // namespace pf5d
