"""MI355X-native chunk extraction for BinaryMeshFitting-style voxel terrain.

The compute path is the C-ABI library ``libbmf.so`` (hand-written HIP
kernels for gfx950, declared in include/bmf.h); ``capi`` binds it with
ctypes.  ``workloads`` builds the seeded synthetic chunk sets.
"""
