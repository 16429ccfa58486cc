"""MI355X-native GPU Masstree (B+tree of tries over u32 key slices).

The index itself is HIP C++ (include/indexingpu_amd/, built into libgpu_masstree.so
with a C ABI declared in include/gpu_masstree_c.h); this package is its Python host
binding and workload generators.
"""
from .masstree import (DEVICE_ERRORS, REQUEST_ERASE, REQUEST_FIND, REQUEST_INSERT,
                       REQUEST_UPDATE, GpuMasstree, MasstreeError, load_library)

__all__ = ["GpuMasstree", "MasstreeError", "load_library", "REQUEST_INSERT", "REQUEST_ERASE",
           "REQUEST_FIND", "REQUEST_UPDATE", "DEVICE_ERRORS"]
