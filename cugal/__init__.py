"""The reference's package name (cugal/__init__.py:6-7) for code written
against it: ``import cugal``, ``from cugal.pred import cugal``,
``from cugal.config import Config``, ``from cugal.profile import Profile,
TimeStamp`` and the other modules of the reference's layout resolve to the
modules of cugal_amd (the same module objects, registered under the
``cugal.`` names), so the HIP path runs behind the reference's imports."""

import sys

from cugal_amd import (  # noqa: F401
    adjacency, config, cuda_kernels, feature_extraction, greedy_lap, pred, profile, sinkhorn,
)

for _module in (adjacency, config, feature_extraction, greedy_lap, pred, profile, sinkhorn):
    sys.modules[f"{__name__}.{_module.__name__.rsplit('.', 1)[1]}"] = _module
del _module

from cugal_amd.config import Config, HungarianMethod, SinkhornMethod  # noqa: E402,F401
from cugal_amd.pred import cugal  # noqa: E402,F401
