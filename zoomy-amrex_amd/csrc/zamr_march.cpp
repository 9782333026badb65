// The per-variant half of a system library: the CFL-law constants that
// zoomy_amrex/_constants.py emits into MarchConstants.H (c_cfl_dimension =
// mesh dimension d, c_cfl_degree_factor = 2k+1), consumed by ComputeDt as
// dt = cfl / (d (2k+1) max|lambda|/dx) (Source/ZoomyAmr.cpp:501-513).
#include "MarchConstants.H"

extern "C" void zamr_internal_march(int* dim, int* deg) {
    if (dim) *dim = ZoomyConstants::c_cfl_dimension;
    if (deg) *deg = ZoomyConstants::c_cfl_degree_factor;
}
