// Drop-in include name of the reference's include/device_context.hpp
// (hyoungjook/IndexinGPU): device_allocator_context / device_reclaimer_context in the
// global namespace.
#pragma once
#include "indexingpu_amd/device_context.hpp"
