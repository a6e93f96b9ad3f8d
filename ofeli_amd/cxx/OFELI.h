// OFELI.h -- umbrella header of the drop-in C++ layer (counterpart:
// include/OFELI.h).  Link with libofeli_hip.so; every solve-path
// computation runs on the GPU.
#pragma once

#include "Equation.h"
#include "Mesh.h"
#include "OFELI_Config.h"
#include "Solvers.h"
#include "SpMatrix.h"
#include "Vect.h"
