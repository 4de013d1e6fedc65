"""MI355X-native AVPL radiance/antiradiance gather and Karras LBVH.

The compute lives in ``libarc.so`` (hand-written HIP for gfx950) behind the
C ABI of ``include/arc.h``; ``ffi`` declares it, ``runtime`` wraps it in host
objects mirroring ``CudaGather`` / ``Bvh`` / ``AvplBvh`` / ``SimpleBvh``, and
``scene`` generates the synthetic workloads.
"""
from . import ffi  # noqa: F401

__version__ = "0.1.0"
