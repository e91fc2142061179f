// GreensFunctionPart (GreensFunctionPart.hpp): the reference's header name for a class this library declares in GreensFunction.hpp.
#pragma once

#include "GreensFunction.hpp"
