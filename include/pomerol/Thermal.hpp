// Thermal (Thermal.hpp): the reference's header name for a class this library declares in Misc.hpp.
#pragma once

#include "Misc.hpp"
