// common.hpp -- scalar types of the BLCO C++ API (include/common.hpp:36-49 of
// the reference): IType is the index type, FType the value type of the host
// API.  The device precision is chosen at run time (cpals options), not by
// recompiling with another FType.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>

typedef unsigned long long IType;
typedef double FType;

#ifndef MAX_NUM_MODES
#define MAX_NUM_MODES 5
#endif

// Aligned host allocation (AlignedMalloc / AlignedFree).
void* AlignedMalloc(size_t size);
void AlignedFree(void* addr);
