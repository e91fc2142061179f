// StatesClassification (StatesClassification.hpp): the reference's header name for a class this library declares in HilbertSpace.hpp.
#pragma once

#include "HilbertSpace.hpp"
