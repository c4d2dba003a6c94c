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
           "PatchMatchParams", "Problem", "install_as_gmvs"]

# The reference's module names (gmvs/__init__.py, gmvs/consistency.py:34,
# gmvs/planar_init.py:54) and the modules of this package that replace them.
_REFERENCE_MODULES = (
    ("gmvs", ""),
    ("gmvs.src", ".src"),
    ("gmvs.lib", ".lib"),
    ("gmvs.lib.build", ".lib.build"),
    ("gmvs.lib.build.planar_initer", ".lib.build.planar_initer"),
    ("gmvs.consistency", ".consistency"),
    ("gmvs.planar_init", ".planar_init"),
    ("gmvs.primitives", ".primitives"),
    ("gmvs.scripts", ".scripts"),
    ("gmvs.scripts.launch", ".scripts.launch"),
)


def install_as_gmvs() -> None:
    """Registers this package under the reference's import names, so code
    written against the reference (``import gmvs.src as _C``, ``from
    gmvs.lib.build import planar_initer``, ``from gmvs import
    run_patch_match``) runs on this implementation unchanged. Raises
    ImportError if another package is already imported as ``gmvs``."""
    import importlib
    import sys

    cur = sys.modules.get("gmvs")
    if cur is not None and cur is not sys.modules[__name__]:
        raise ImportError(f"a different gmvs package is already imported ({cur!r})")
    for ref_name, rel in _REFERENCE_MODULES:
        mod = sys.modules[__name__] if not rel else importlib.import_module(rel, __name__)
        sys.modules[ref_name] = mod


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
