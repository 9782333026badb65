#pragma once
// MarchConstants.H (_constants.py:59-88) includes <AMReX_REAL.H> for amrex::Real.
#include "AMReX_SmallMatrix.H"
