"""pannde_amd: an MI355X-native DeviceSolver for PanNDE's rotated-staggered-grid (RSG)
elastodynamic time step.

The compute path is libpannde_rsg.so (hand-written HIP for gfx950 behind the C ABI in
include/pannde_rsg.h); ``pannde_amd.rsg`` binds it, ``pannde_amd.cases`` builds jobs.
"""
__all__ = ["rsg", "cases", "build"]
