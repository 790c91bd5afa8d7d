"""rex: MI355X-native wavefront path-tracing core (flattened SAH BVH traversal,
wavefront PBR integrator, ReSTIR DI) behind a C ABI (include/rex.h).

The compute path is librex.so: host C++ plus hand-written HIP kernels for
gfx950, built in-tree by ``python -m renderingengine_amd.build``.  This Python
package is a thin host-side mirror of the reference's interfaces used by the
tests and the benchmark.
"""
from . import _lib  # noqa: F401

__all__ = ["_lib"]
