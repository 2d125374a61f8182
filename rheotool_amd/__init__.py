"""rheotool_amd: the log-conformation stress step of rheoTool on MI355X (gfx950).

Native parts (built in-tree by ``rheotool_amd.build``):
  lib/librheo_gpu.so   HIP kernels + the C-ABI of include/rheo_gpu.h
  lib/librheo_case.so  synthetic case generator (meshes, seeded fields)

Python entry points: ``rheotool_amd.abi`` (ctypes over the C-ABI) and
``rheotool_amd.cases`` (workload generation).
"""
__all__ = ["abi", "cases", "build"]
