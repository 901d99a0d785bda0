"""MI355X homotopy path tracker for Stephenson II finite root generation.

The tracker runs as hand-written HIP kernels for gfx950 in libfrg.so, reached through the C ABI
of include/frg.h; ``capi`` is its Python binding and ``tasks`` holds the target definitions.
"""
from . import tasks  # noqa: F401

__all__ = ["tasks", "capi"]
