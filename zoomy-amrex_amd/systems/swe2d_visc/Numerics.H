#pragma once
// The swe2d Riemann solver (../swe2d/Numerics.H) over the viscous model: the
// diffusive flux enters the right-hand side through make_rhs.H's
// add_diffusive_flux, not through the numerical flux.
#include "Model.H"
#include "../swe2d/Numerics.H"
