"""lbmcl_amd -- MI355X (gfx950) device layer and host tools for the D3Q19
lid-driven-cavity solver LBMCL.

Native parts (built in-tree by ``python -m lbmcl_amd.build``):
  liblbm_hip.so   the HIP kernels behind the C ABI of include/lbm_hip.h
  bin/lbmcl       the lbmcl command line (same options and outputs as the reference)

Python parts:
  capi       ctypes binding of the C ABI (no CPU fallback)
  vti        numpy-only VTI reader
  verify     the verify.py counterpart (MSE / max-abs error per dumped iteration)
  accuracy   the storage/precision forms against an FP64 run
"""
__all__ = ["capi", "vti", "verify", "accuracy", "build"]
