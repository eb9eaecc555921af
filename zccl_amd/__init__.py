"""MI355X-native SZp fastmode1 codec and hZCCL homomorphic sum.

The compute path is the HIP library ``zccl_amd/libzccl_amd.so`` (C ABI in
``include/zccl_amd.h``); this package is its Python host mirror.  Importing it
does not touch the GPU; the first call that needs the library loads it and
raises ``LibraryMissing`` if it has not been built.
"""
from ._lib import LibraryMissing, ZcclError, lib  # noqa: F401

__all__ = ["LibraryMissing", "ZcclError", "lib", "codec"]


def __getattr__(name):
    if name in ("codec", "allreduce", "batch", "serve", "autotune", "synth"):
        import importlib

        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)
