"""gmvs_amd -- the gmvs PatchMatch MVS hot path, native to AMD MI355X (gfx950).

Mirrors the reference package layout (gmvs/__init__.py:2-7):
``run_patch_match`` / ``run_fusion`` (consistency), ``initialization``
(planar_init), ``PatchMatchParams`` and ``Problem`` (primitives); the
compiled boundary lives in ``gmvs_amd.src`` and
``gmvs_amd.lib.build.planar_initer``. The HIP library is loaded lazily by
those modules, so importing the package itself needs neither a GPU nor
the built library.
"""
from .version import __version__

__all__ = ["__version__", "run_patch_match", "run_fusion", "initialization",
           "PatchMatchParams", "Problem"]


def __getattr__(name):
    if name in ("run_patch_match", "run_fusion"):
        from . import consistency

        return getattr(consistency, name)
    if name == "initialization":
        from .planar_init import initialization

        return initialization
    if name in ("PatchMatchParams", "Problem"):
        from . import primitives

        return getattr(primitives, name)
    raise AttributeError(name)
