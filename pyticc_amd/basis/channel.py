"""Field-free BF channel basis (src/pyticc/basis/channel.py of the reference).

Channel 30-58, ChannelBasis 125-187, K=0 parity rule 428-436, state selection
457-484, build_ChannelBasis 491-576 (stable sort by threshold, 569).
"""

from __future__ import annotations

from collections.abc import Sequence
from dataclasses import dataclass, field
from functools import lru_cache

import numpy as np

from pyticc_amd._log import fail
from pyticc_amd.energy import get_Etot
from pyticc_amd.system import ChannelSpec, MolInnerState, MonomerType


@lru_cache
def set_Kmax(j_couple: int, Jtot: int, Kcut: int | None = None) -> int:
    Kmax = min(j_couple, Jtot)
    return Kmax if Kcut is None else min(Kmax, Kcut)


@dataclass(frozen=True)
class Channel:
    mis_X: MolInnerState
    mis_Y: MolInnerState
    j_couple: int
    K: int
    E_int: float


@dataclass(frozen=True)
class OpenClosedChannels:
    open_mask: np.ndarray

    @property
    def n_open(self) -> np.ndarray:
        return np.asarray(np.sum(self.open_mask, axis=1), dtype=np.int64)

    @property
    def n_closed(self) -> np.ndarray:
        return np.asarray(self.open_mask.shape[1] - self.n_open, dtype=np.int64)


@dataclass(frozen=True)
class ChannelBasis(Sequence):
    channels: tuple
    Jtot: int
    system_parity: int
    channel_spec: ChannelSpec = field(default_factory=ChannelSpec)
    exchange: object = None

    @property
    def molecule_exchange(self) -> int:
        return 0

    @property
    def n_channel(self) -> int:
        return len(self.channels)

    @property
    def E_int(self) -> np.ndarray:
        return np.asarray([c.E_int for c in self.channels], dtype=np.float64)

    def __len__(self) -> int:
        return self.n_channel

    def __getitem__(self, index):
        return self.channels[index]

    def open_closed(self, total_energies) -> OpenClosedChannels:
        energies = get_Etot(total_energies)
        return OpenClosedChannels(open_mask=self.E_int[None, :] < energies[:, None])


def _allow_closed_shell_K0(mis_X, mis_Y, j_couple, Jtot, system_parity) -> bool:
    return system_parity * (-1) ** (mis_X.j + mis_Y.j + j_couple + Jtot) == 1


def _state_values(values, n_state: int, name: str):
    selected = (values,) * n_state if isinstance(values, int) else values
    if len(selected) != n_state:
        raise fail(ValueError, f"{name} must provide one value per electronic state; expected {n_state}, got {len(selected)}")
    return selected


def _selected_states(monomer, E_cut, vmin, exchange_parity):
    states = tuple(monomer.mis_iter(E_cut))
    if monomer.type is not MonomerType.DIATOM:
        return states
    vmin_v = _state_values(vmin, 1, "vmin")[0]
    par = _state_values(exchange_parity, 1, "exchange_parity")[0]
    out = []
    for s in states:
        if s.v is None:
            raise fail(ValueError, "Diatomic channel state requires a vibrational quantum number")
        if s.v < vmin_v:
            continue
        if par != 0 and (-1) ** s.j != par:
            continue
        out.append(s)
    return tuple(out)


def build_ChannelBasis(system, channel: ChannelSpec | None = None) -> ChannelBasis:
    """Enumerate (mis_X, mis_Y, j12, K) channels, K=0 parity rule, energy order."""
    if system.Jtot is None or system.system_parity is None:
        raise fail(ValueError, "Field-free channel construction requires Jtot and system_parity")
    spec = channel if channel is not None else (system.channel_spec or ChannelSpec())
    mX, mY = system.monomer_X, system.monomer_Y
    states_X = _selected_states(mX, spec.E_X_cut, spec.vmin_X, spec.exchange_parity_X)
    states_Y = _selected_states(mY, spec.E_Y_cut, spec.vmin_Y, spec.exchange_parity_Y)
    channels = []
    for sx in states_X:
        for sy in states_Y:
            for j12 in range(abs(sx.j - sy.j), sx.j + sy.j + 1):
                for K in range(set_Kmax(j12, system.Jtot, spec.K_cut) + 1):
                    if not mX.allows_K(sx, K) or not mY.allows_K(sy, K):
                        continue
                    if K == 0 and not _allow_closed_shell_K0(sx, sy, j12, system.Jtot, system.system_parity):
                        continue
                    channels.append(Channel(sx, sy, j12, K, float(mX.energy(sx, K) + mY.energy(sy, K))))
    channels.sort(key=lambda c: c.E_int)
    return ChannelBasis(channels=tuple(channels), Jtot=system.Jtot, system_parity=system.system_parity, channel_spec=spec)
