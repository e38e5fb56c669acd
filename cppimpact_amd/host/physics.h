// NeohookeanPhysics lives in tetrahedral.h on this backend (a type tag);
// this header keeps the reference's include list (test.cu:6-12) valid.
#pragma once
#include "tetrahedral.h"
