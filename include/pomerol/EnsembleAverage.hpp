// EnsembleAverage (EnsembleAverage.hpp): the reference's header name for a class this library declares in MonomialOperator.hpp.
#pragma once

#include "MonomialOperator.hpp"
