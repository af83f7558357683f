"""MI355X-native ray-cast and path-tracing hot path of libYafaRay.

The compute path is libymx.so (hand-written HIP kernels behind the C ABI in
include/); this package holds its Python binding, the scene generators of the
benchmark workloads and the build script.
"""
__all__ = ["native", "accel", "scenes"]
