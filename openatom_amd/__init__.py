"""openatom_amd -- MI355X-native CP electronic-state step.

The compute lives in liboa_step.so (hand-written gfx950 HIP kernels behind the
C ABI of include/oa_step.h).  This package is the host-side view:

  _lib      ctypes binding of the C ABI (loads the library, no fallback)
  step      CPContext: one device context, reference-named stages
  system    seeded synthetic water systems, coefficients and V(r)
"""
from ._lib import LIB_PATH, OaError, load  # noqa: F401

__all__ = ["LIB_PATH", "OaError", "load"]
