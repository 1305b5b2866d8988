"""pyticc_amd -- MI355X (gfx950) backend of PyTICC's coupled-channel path.

Drop-in surface of the reference's ``import pyticc as ticc`` for this path:
build_ScattSystem, prepare_potential, PotentialGrid, Propagation, solve
(exact CC, and the CS / NNCC K-block approximations).
The hot path (W(R) contraction, LDMD sweep, K->S matching) runs in the HIP
kernels of libticc_hip.so; there is no CPU fallback.
"""

from pyticc_amd.basis.channel import Channel, ChannelBasis, build_ChannelBasis
from pyticc_amd.basis.kblock import KBlock, build_cs_blocks, build_nncc_blocks
from pyticc_amd.basis.monomer import AtomSpec, DiatomBasis, DiatomSpec, rigid_rotor
from pyticc_amd.basis.rovib import RovibBasis
from pyticc_amd.constants import (
    AMU2AU,
    ANG2AU,
    AU2AMU,
    AU2ANG,
    AU2CM,
    AU2EV,
    AU2GHZ,
    AU2HZ,
    AU2K,
    AU2KHZ,
    AU2MHZ,
    CM2AU,
    EV2AU,
    GHZ2AU,
    HZ2AU,
    K2AU,
    KHZ2AU,
    MHZ2AU,
)
from pyticc_amd.crosssection import PartialCrossSections, partial_cross_sections
from pyticc_amd.energy import get_Etot
from pyticc_amd.pes import PESWrapper
from pyticc_amd.pes.native import load_native_pes
from pyticc_amd.propagation import Propagation
from pyticc_amd.result import CoupledStatesResult, KBlockResult, ScatteringResult, Timing
from pyticc_amd.scattering import PotentialGrid, ScattHamiltonian, prepare_potential, solve
from pyticc_amd import report
from pyticc_amd.input import run
from pyticc_amd.basis.monomer import prepare_Diatom
from pyticc_amd.system import Approx, ChannelSpec, MolInnerState, ScatteringType, ScattSystem, build_ScattSystem, element_mass_au, reduced_mass

__all__ = [
    "AMU2AU", "ANG2AU", "AU2AMU", "AU2ANG", "AU2CM", "AU2EV", "AU2GHZ", "AU2HZ", "AU2K", "AU2KHZ", "AU2MHZ",
    "Approx", "AtomSpec", "CM2AU", "Channel", "ChannelBasis", "ChannelSpec", "CoupledStatesResult", "DiatomBasis",
    "DiatomSpec", "EV2AU", "GHZ2AU", "HZ2AU", "K2AU", "KBlock", "KBlockResult", "KHZ2AU", "MHZ2AU", "MolInnerState", "PartialCrossSections", "PESWrapper", "PotentialGrid", "Propagation",
    "RovibBasis", "ScatteringResult", "ScatteringType", "ScattHamiltonian", "ScattSystem", "Timing",
    "build_ChannelBasis", "build_ScattSystem", "build_cs_blocks", "build_nncc_blocks", "element_mass_au", "get_Etot", "load_native_pes", "partial_cross_sections", "prepare_potential", "reduced_mass",
    "prepare_Diatom", "report", "rigid_rotor", "run", "solve",
]
