#pragma once
// TEST SYSTEM: the swe2d Riemann solver over the fallback test model.
#include "Model.H"
#include "../../../zoomy-amrex_amd/systems/swe2d/Numerics.H"
