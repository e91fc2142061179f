// TwoParticleGFContainer (TwoParticleGFContainer.hpp): the reference's header name for a class this library declares in TwoParticleGF.hpp.
#pragma once

#include "TwoParticleGF.hpp"
