"""
dadi_amd -- MI355X-native backend for dadi's hot path: ADI time stepping of
the allele-frequency density phi in 1-5 populations and the phi -> spectrum
sampling integral, behind dadi's own API (Integration.one_pop .. five_pops,
Spectrum.from_phi, cuda_enabled).

All integration and sampling runs in libdadi_hip.so (hand-written HIP for
gfx950) through a ctypes C ABI; there is no CPU integrator in this package.
"""
import os

import numpy

# Independent work runs on separate HIP streams (each launch class of the
# serving loop, the grids of a batched extrapolated evaluation).  HIP maps
# streams onto GPU_MAX_HW_QUEUES hardware queues, 4 by default, and streams
# that share a queue run one after another.  With 16 the serving loop's
# classes overlap on the device (seeded 200-request mix on MI355X: p50 79 ->
# 53 ms, p99 1.14 -> 0.64 s), so a smaller value is raised to
# DADI_AMD_HW_QUEUES (default 16, at most 32; 0 leaves the variable alone).
# HIP reads the variable once, when its runtime initialises: this takes
# effect only if nothing in the process has used the GPU yet.
def _hw_queues():
    try:
        want = min(int(os.environ.get("DADI_AMD_HW_QUEUES", "16")), 32)
        have = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
    except ValueError:
        return
    if want > have:
        os.environ["GPU_MAX_HW_QUEUES"] = str(want)


_hw_queues()

from . import _lib
from . import Demes, Misc, Numerics, PhiManip, Integration, Inference
from . import Spectrum_mod
from .Spectrum_mod import Spectrum
from . import Demographics1D, Demographics2D, Demographics3D
from . import Godambe
from . import DFE

# dadi ignores floating-point warnings globally (dadi/__init__.py:28)
numpy.seterr(all='ignore')

__version__ = "0.1.0"


def cuda_enabled(toggle=None):
    """Query or switch the GPU backend (dadi/__init__.py:30-48).

    None returns the current state; True initialises the HIP library on
    device 0 and returns True, or prints why it cannot and returns False;
    False disables it (integration calls then raise, since this package has
    no CPU integrator).
    """
    if toggle is None:
        return Integration.cuda_enabled
    elif toggle is True or toggle == True:  # noqa: E712  (dadi accepts truthy 1)
        try:
            _lib.init(0)
        except Exception as e:  # mirror dadi: report and return False
            print("Failed to initialise the HIP backend: %s" % e)
            Integration.cuda_enabled = False
            return False
        Integration.cuda_enabled = True
        return True
    elif toggle is False or toggle == False:  # noqa: E712
        Integration.cuda_enabled = False
        return False
    else:
        raise ValueError("toggle must be True, False, or None")


def pts_to_RAM(pts, P):
    """Approximate memory (GiB) of an integration (dadi/__init__.py:52-59)."""
    return 8 * 4 * pts ** P / 1024 ** 3


def RAM_to_pts(RAM, P):
    """Largest grid size for a memory budget (dadi/__init__.py:61-68)."""
    return int((RAM * 1024 ** 3 / (8 * 4)) ** (1. / P))
