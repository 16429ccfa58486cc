// Drop-in include name of the reference's include/gpu_masstree.hpp
// (hyoungjook/IndexinGPU): GpuMasstree::gpu_masstree for MI355X, plus the allocators,
// reclaimers and device contexts that header brings in (its lines 34-38).
#pragma once
#include "indexingpu_amd/gpu_masstree.hpp"
