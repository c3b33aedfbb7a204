"""emu_amd: an MI355X (gfx950) implementation of Emu's particle-in-cell
flavor-evolution step behind a C ABI (include/emu_amd.h).

The HIP kernels live in libemu_amd.so (built by ``python -m emu_amd.build``);
this package binds them and mirrors the reference's call surface.
"""
from . import capi, configs  # noqa: F401
from .params import TestParams  # noqa: F401

__all__ = ["capi", "configs", "TestParams", "core"]


def __getattr__(name):
    # core imports torch; load it on first use
    if name == "core":
        import importlib
        return importlib.import_module(".core", __name__)
    raise AttributeError(name)
