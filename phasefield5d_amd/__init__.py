"""phasefield5d_amd -- MI355X (gfx950) explicit Cahn-Hilliard step for phasefield5d.

Host setup mirrors the reference package layout (solver, kinetics,
thermodynamics, elasticity, materials, utils); the per-step arithmetic runs in
libpf5d.so (hand-written HIP kernels behind a C ABI, include/pf5d.h).
"""
__version__ = "0.1.0"
