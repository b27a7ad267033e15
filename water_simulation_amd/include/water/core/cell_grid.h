// Reference include path kept for drop-in builds; everything lives in water/wsim.hpp.
#pragma once
#include "water/wsim.hpp"
