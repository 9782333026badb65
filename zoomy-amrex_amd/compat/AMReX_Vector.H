#pragma once
// Host-side std::vector is what Model::get_boundary_tags / parameter_names /
// default_parameters return (to_amrex.py:358-360).
#include <vector>
#include "AMReX_SmallMatrix.H"
