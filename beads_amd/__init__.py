"""beads_amd: an MI355X-native engine behind the BEADS Python API.

``import beads_amd as beads`` gives the same builders the reference exposes
(python/beads/__init__.py:1-16): System, ForceField, Dynamics, Simulation,
build_info and the ``io`` helpers. Runs execute in the gfx950 engine
(libbeads_amd.so) through the C ABI declared in include/beads_amd.h.
"""

from ._runtime import share_hip_runtime

share_hip_runtime()  # before anything loads libbeads_amd.so

from . import io  # noqa: E402
from ._build_config import build_info  # noqa: E402
from .forcefield import Dynamics, ForceField  # noqa: E402
from .simulation import Simulation  # noqa: E402
from .system import System  # noqa: E402

__version__ = "0.0.1"

__all__ = ["Dynamics", "ForceField", "Simulation", "System", "__version__", "build_info", "io"]
