"""pocketsdr_amd - MI355X (gfx950) PocketSDR search / correlate hot path.

The compute lives in ``libpocketsdr_amd.so`` (hand-written HIP, C ABI declared
in ``include/pocket_sdr_amd.h``).  This package is a thin ctypes binding in the
style of the reference's own Python binding of ``libsdr.so``
(python/sdr_func.py:30-45): numpy arrays in, numpy arrays out.

Importing the package loads the shared library; if it has not been built the
import fails loudly.  There is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpocketsdr_amd.so")


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libpocketsdr_amd.so is not built (run __graft_entry__.build() or "
            "make -C pocketsdr_amd)")
    # The library's own symbols stay local: its drop-in names equal the
    # reference libsdr.so's, which another component of the process may load.
    h = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_LOCAL)
    # Its HIP runtime, though, is made global, so that a framework loaded
    # later with a bundled HIP runtime of its own (PyTorch's wheels carry one)
    # binds to this one instead of starting a second runtime in the process,
    # which then finds no device.
    try:
        ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL | os.RTLD_NOLOAD)
    except OSError:
        pass
    return h


lib = _load()

from . import sdr_func  # noqa: E402,F401
