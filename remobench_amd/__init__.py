"""MI355X-native RemoLOD progressive-LOD path: HIP kernels behind a C ABI.

The native library is remobench_amd/_lib/libremolod.so (include/remolod.h);
`remobench_amd.lib` binds it and `remobench_amd.pipeline` mirrors the
reference's ILodPipeline plugin interface on top.
"""
from . import lib  # noqa: F401

__all__ = ["lib"]
