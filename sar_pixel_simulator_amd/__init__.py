"""MI355X-native FDTD engine for the fdtdsarsim SAR-pixel simulator.

Drop-in for the reference's user API (``import fdtdsarsim as fdtd``):
``load_from_json``, ``set_geom_params``, ``ParameterCollection``,
``materials``, ``run_experiment`` and the ``'hip'`` processing type.  The
time loop runs in hand-written HIP kernels (libfdtd_hip.so) reached through
a C-ABI loaded with ctypes.
"""
from .parameters import *  # noqa: F401,F403
from .parameters import (ParameterCollection, load_from_json, set_geom_params,  # noqa: F401
                         inc_filename, results_filename)
from .experiment import get_fdtd_simulator, run_experiment  # noqa: F401
from .incidentbeam import IncidentBeam  # noqa: F401
from . import materials  # noqa: F401
from . import near2far  # noqa: F401
from . import processresults  # noqa: F401
from .utils import round_to_odd  # noqa: F401
