"""Radial LDMD propagation on the MI355X backend."""

from pyticc_amd.propagation.config import Propagation, PropagationMode
from pyticc_amd.propagation.device import ResolvedDevice, resolve_device
from pyticc_amd.propagation.grid import RadialSector, build_radial_sectors, iter_radial_windows
from pyticc_amd.propagation.logd import initialize_logD_inelastic, propagate_logD, propagate_logD_sector

__all__ = [
    "Propagation",
    "PropagationMode",
    "RadialSector",
    "ResolvedDevice",
    "build_radial_sectors",
    "initialize_logD_inelastic",
    "iter_radial_windows",
    "propagate_logD",
    "propagate_logD_sector",
    "resolve_device",
]
