// Umbrella header of the pomerol-compatible C++ API (reference: include/pomerol.hpp).
#pragma once

#include "pomerol/Misc.hpp"
#include "pomerol/Operators.hpp"
#include "pomerol/LatticePresets.hpp"
#include "pomerol/HilbertSpace.hpp"
#include "pomerol/Hamiltonian.hpp"
#include "pomerol/MonomialOperator.hpp"
#include "pomerol/GreensFunction.hpp"
#include "pomerol/TwoParticleGF.hpp"
#include "pomerol/ThreePointSusceptibility.hpp"
#include "pomerol/Vertex4.hpp"
#include "pomerol/Batch.hpp"
#include "pomerol/Tangent.hpp"
