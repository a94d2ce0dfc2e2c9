"""pfpa -- host side of the MI355X-native KKS phase-field step.

libpfpa.so (HIP for gfx950, C-ABI in include/pfpa.h) does all the work; this
package binds it (ctypes), mirrors the reference's host API (api.py), holds
the model parameters and initial conditions (model.py) and a native stepping
loop (sim.py).
"""
from ._lib import LIB_PATH, PfError, lib  # noqa: F401
from . import model  # noqa: F401
