// Drop-in include name of the reference's include/simple_slab_alloc.hpp (hyoungjook/IndexinGPU):
// the MI355X allocators, in the global namespace under the reference's names.
#pragma once
#include "indexingpu_amd/slab_alloc.hpp"
#include "indexingpu_amd/device_context.hpp"
