// Drop-in include name of the reference's include/simple_dummy_reclaim.hpp (hyoungjook/IndexinGPU):
// the MI355X reclaimers, in the global namespace under the reference's names.
#pragma once
#include "indexingpu_amd/reclaim.hpp"
#include "indexingpu_amd/device_context.hpp"
