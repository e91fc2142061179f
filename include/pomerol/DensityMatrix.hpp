// DensityMatrix (DensityMatrix.hpp): the reference's header name for a class this library declares in Hamiltonian.hpp.
#pragma once

#include "Hamiltonian.hpp"
