#pragma once
// The printer's Model.H includes <AMReX_Array4.H>; the model bodies only need
// the scalar and SmallMatrix names, which live in AMReX_SmallMatrix.H.
#include "AMReX_SmallMatrix.H"
