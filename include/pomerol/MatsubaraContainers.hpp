// MatsubaraContainer4 (MatsubaraContainers.hpp): the reference's header name for a class this library declares in Vertex4.hpp.
#pragma once

#include "Vertex4.hpp"
