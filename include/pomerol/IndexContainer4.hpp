// IndexCombination4 and the 2PGF container (IndexContainer4.hpp): the reference's header name for a class this library declares in TwoParticleGF.hpp.
#pragma once

#include "TwoParticleGF.hpp"
