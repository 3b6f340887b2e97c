"""MI355X-native field engine for packed Cartesian multipoles: a drop-in for
the `field_direct` / `field_fmm` calls of peter-reinholdt/fmm.

The compute lives in libfmm_hip.so (HIP kernels for gfx950 behind the C ABI
of include/fmm_hip.h).  `fmm_amd.octree` mirrors the f2py module `octree`.
"""
__all__ = ["octree", "synth"]
