// ThreePointSusceptibilityContainer (ThreePointSusceptibilityContainer.hpp): the reference's header name for a class this library declares in ThreePointSusceptibility.hpp.
#pragma once

#include "ThreePointSusceptibility.hpp"
