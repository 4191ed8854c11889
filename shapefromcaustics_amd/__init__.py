"""MI355X-native photon-differential splatting (PDS) for Shape from Caustics.

Drop-in for the reference's ``PhotonDifferentialSplatting`` extension and its
``PDS`` autograd function (``model/photon_differential.py:7-19``), built on
hand-written HIP kernels for gfx950 behind the C ABI of ``include/pds.h``.

    from shapefromcaustics_amd import PDS
    grid = PDS.apply(Ep, xp, Mp, cp, radius, (C, H, W), max_pixel_radius)

``install_alias()`` registers the module under the reference's import name so
``from PhotonDifferentialSplatting import pds_forward, pds_backward`` works.
"""
from __future__ import annotations

import sys

__version__ = "0.1.0"


def install_alias() -> None:
    """Make ``import PhotonDifferentialSplatting`` resolve to this package's module."""
    from . import PhotonDifferentialSplatting as mod
    sys.modules.setdefault("PhotonDifferentialSplatting", mod)


def __getattr__(name):
    # lazy: importing the package does not import torch or load the library
    if name in ("PDS", "pds"):
        from .photon_differential import PDS
        return PDS
    if name in ("pds_forward", "pds_backward"):
        from . import PhotonDifferentialSplatting as mod
        return getattr(mod, name)
    raise AttributeError(name)
