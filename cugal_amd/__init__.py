"""MI355X (gfx950) implementation of the FUGAL graph-alignment pipeline with
the public surface of cuGAL (cugal/__init__.py:6-7): ``Config``,
``SinkhornMethod``, ``HungarianMethod`` and ``cugal``.

The numerical work runs in hand-written HIP kernels (libcugal_hip.so, C ABI in
include/cugal_hip.h); torch supplies device memory and streams.
"""

from .config import Config, HungarianMethod, SinkhornMethod  # noqa: F401
from .pred import cugal  # noqa: F401
