"""MI355X force engine for AstroGen2 (Philip-Spaeth/AstroGenesis2.0).

Python side of the host: seeded initial conditions, the reference's file
formats, and a ctypes binding of the C ABI declared in include/ag2.h.  The
directory name is not a Python identifier, so callers load it by path::

    import importlib.util, sys
    spec = importlib.util.spec_from_file_location(
        "ag2_amd", "astrogenesis2.0_amd/__init__.py",
        submodule_search_locations=["astrogenesis2.0_amd"])
    ag2 = importlib.util.module_from_spec(spec); sys.modules["ag2_amd"] = ag2
    spec.loader.exec_module(ag2)
"""
import importlib
import os
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))

from . import units  # noqa: E402,F401
from . import ics  # noqa: E402,F401
from . import ageio  # noqa: E402,F401


def __getattr__(name):
    if name in ("engine", "serving", "build"):
        mod = importlib.import_module("." + name, __name__)
        setattr(sys.modules[__name__], name, mod)
        return mod
    raise AttributeError(name)
