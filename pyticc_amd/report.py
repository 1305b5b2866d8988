"""Text reports of channel bases and scattering results (src/pyticc/report.py:27-41,
43-66, 137-290, 293-353, 358-427, 779-985, 988-1247), for the field-free atom-diatom
and diatom-diatom results this backend produces.

Every report is a right-aligned table: a header line, a dashed separator and one
line per row, columns separated by two spaces.  Energies are in cm-1; S-matrix
elements are written as S[out, in] with 16 digits after the decimal point, one
line per (energy, incoming, outgoing) triple with the incoming channel outermost.
"""

from __future__ import annotations

from collections.abc import Sequence

import numpy as np

from pyticc_amd._log import fail
from pyticc_amd.basis.kblock import KBlock
from pyticc_amd.constants import AU2CM
from pyticc_amd.energy import get_Etot
from pyticc_amd.result import CoupledStatesResult, KBlockResult, ScatteringResult


def _table(headers: Sequence[str], rows: Sequence[Sequence[str]]) -> str:
    """Right-aligned columns two spaces apart under a dashed rule (report.py:27-38)."""
    widths = [len(h) for h in headers]
    for row in rows:
        for k, cell in enumerate(row):
            widths[k] = max(widths[k], len(cell))
    lines = ["  ".join(h.rjust(w) for h, w in zip(headers, widths)), "  ".join("-" * w for w in widths)]
    lines.extend("  ".join(cell.rjust(w) for cell, w in zip(row, widths)) for row in rows)
    return "\n".join(lines)


def _is_atom(state) -> bool:
    return state.v is None and state.t is None and state.electronic_state is None and state.j == 0


def _state_columns(states, suffix: str = "") -> list:
    """(header, attribute) for the quantum numbers present among `states` (report.py:112-125)."""
    cols = []
    if any(s.electronic_state is not None for s in states):
        cols.append((f"state{suffix}", "electronic_state"))
    if any(s.v is not None for s in states):
        cols.append((f"v{suffix}", "v"))
    if any(s.t is not None for s in states):
        cols.append((f"t{suffix}", "t"))
    cols.append((f"j{suffix}", "j"))
    return cols


def _cell(state, attribute: str) -> str:
    value = getattr(state, attribute)
    return "-" if value is None else str(value)


def rovib_levels(basis) -> str:
    """Retained diatomic levels sorted by internal energy (report.py:43-66)."""
    states = sorted(basis.mis_iter(float("inf")), key=lambda s: s.Eint)
    return _table(("v", "j", "E_int/cm-1"), [[str(s.v), str(s.j), f"{s.Eint * AU2CM:.6f}"] for s in states])


def channels(basis) -> str:
    """Channel quantum numbers of the active monomers, K and E_int (report.py:252-290)."""
    if basis.n_channel == 0:
        return _table(("n", "K", "E_int/cm-1"), ())
    states_X = [c.mis_X for c in basis]
    states_Y = [c.mis_Y for c in basis]
    active_X = not all(_is_atom(s) for s in states_X)
    active_Y = not all(_is_atom(s) for s in states_Y)
    if active_X and active_Y:
        cols = [(h, "X", a) for h, a in _state_columns(states_X, "_X")] + [(h, "Y", a) for h, a in _state_columns(states_Y, "_Y")]
        with_j12 = True
    else:
        side = "X" if active_X else "Y"
        cols = [(h, side, a) for h, a in _state_columns(states_X if active_X else states_Y)]
        with_j12 = any(c.j_couple != (c.mis_X.j if active_X else c.mis_Y.j) for c in basis)
    headers = ["n", *(h for h, _, _ in cols)] + (["j_couple"] if with_j12 else []) + ["K", "E_int/cm-1"]
    rows = []
    for n, ch in enumerate(basis, start=1):
        row = [str(n)] + [_cell(ch.mis_X if side == "X" else ch.mis_Y, a) for _, side, a in cols]
        if with_j12:
            row.append(str(ch.j_couple))
        rows.append(row + [str(ch.K), f"{ch.E_int * AU2CM:.6f}"])
    return _table(headers, rows)


def open_closed(basis, energies) -> str:
    """Open and closed channel counts at each total energy (report.py:293-323)."""
    values = get_Etot(energies)
    counts = basis.open_closed(values)
    rows = [[str(n), f"{E * AU2CM:.8f}", str(o), str(c)]
            for n, (E, o, c) in enumerate(zip(values, counts.n_open, counts.n_closed), start=1)]
    return _table(("n", "Etot/cm-1", "open", "closed"), rows)


def k_blocks(blocks) -> str:
    """Membership and ownership of CS / NNCC K blocks (report.py:328-353)."""
    rows = []
    for item in blocks:
        blk = item.block if isinstance(item, KBlockResult) else item
        if not isinstance(blk, KBlock):
            raise fail(TypeError, f"k_blocks expects KBlock or KBlockResult items, got {type(item).__name__}")
        rows.append([str(blk.index), str(blk.center_K), ",".join(str(k) for k in blk.K_values), str(len(blk.channel_indices)),
                     ",".join(str(k) for k in blk.owned_K_values), str(len(blk.owned_channel_indices))])
    return _table(("block", "center_K", "K_values", "channels", "owned_K", "owned_channels"), rows)


def _selection(value, name: str):
    """None, or the set of integers an int / range / integer sequence selects (report.py:358-375)."""
    if value is None:
        return None
    if isinstance(value, bool) or isinstance(value, (str, bytes)):
        raise fail(TypeError, f"{name} must be an integer, range, integer sequence, or None")
    if isinstance(value, int):
        return {value}
    try:
        items = tuple(value)
    except TypeError as error:
        raise fail(TypeError, f"{name} must be an integer, range, integer sequence, or None") from error
    if any(not isinstance(x, int) or isinstance(x, bool) for x in items):
        raise fail(TypeError, f"{name} must contain only integers")
    return set(items)


def _energy_indices(selection, size: int) -> tuple:
    """Indices an int / slice / integer sequence selects out of `size` energies (report.py:380-396)."""
    available = np.arange(size, dtype=np.int64)
    if selection is None:
        return tuple(int(i) for i in available)
    if isinstance(selection, bool) or isinstance(selection, (str, bytes)):
        raise fail(TypeError, "energy_indices must be an integer, slice, integer sequence, or None")
    if isinstance(selection, int):
        selection = [selection]
    try:
        chosen = np.atleast_1d(available[selection])
    except (IndexError, TypeError) as error:
        raise type(error)(f"Invalid energy_indices {selection!r}: {error}") from error
    return tuple(int(i) for i in chosen)


def _matches(state, electronic, v, j) -> bool:
    return (electronic is None or state.electronic_state in electronic) and (v is None or state.v in v) and (j is None or state.j in j)


def _integral_L(value: float) -> str:
    """L as an integer; a non-integral L (a CS block at Jtot > 0) cannot be labelled (report.py:409-413)."""
    rounded = round(value)
    if abs(value - rounded) > 1.0e-8:
        raise fail(ValueError, f"Orbital angular momentum must be integral, but got L={value}")
    return str(rounded)


def _reject(context: str, filters) -> None:
    chosen = [name for name, value in filters if value is not None]
    if chosen:
        raise fail(ValueError, f"Filters {', '.join(chosen)} do not apply to {context} results")


def _field_free_data(result, block_index):
    """(S matrices, open channel indices, their L) of an exact result or one K block (report.py:779-804)."""
    if isinstance(result, ScatteringResult):
        if block_index is not None:
            raise fail(ValueError, "block_index is only valid for coupled-states results")
        open_idx = result.open_channel_indices
        return result.Smat, open_idx, tuple(result.L[idx] for idx in open_idx)
    if block_index is None:
        if len(result.blocks) != 1:
            raise fail(ValueError, "block_index is required when a coupled-states result has more than one K block")
        block_index = 0
    try:
        br = result.blocks[block_index]
    except IndexError as error:
        raise IndexError(f"block_index {block_index} is outside the available K blocks") from error
    local = {g: i for i, g in enumerate(br.block.channel_indices)}
    Ls = tuple(np.asarray([br.L[local[int(g)]] for g in idx]) for idx in br.open_channel_indices)
    return br.Smat_asymptotic, br.open_channel_indices, Ls


def _smatrix_rows(result, energy_indices, block_index, initial_ok, final_ok, labels):
    """Rows [E, labels(in), labels(out), Re, Im] of the selected S[out, in] elements."""
    mats, open_idx, Ls = _field_free_data(result, block_index)
    rows = []
    for e in _energy_indices(energy_indices, result.Etot.size):
        idx, L, S = open_idx[e], Ls[e], mats[e]
        for a, ga in enumerate(idx):
            if not initial_ok(int(ga)):
                continue
            for b, gb in enumerate(idx):
                if not final_ok(int(gb)):
                    continue
                value = S[b, a]
                rows.append([f"{result.Etot[e] * AU2CM:.8f}", *labels(int(ga), float(L[a])), *labels(int(gb), float(L[b])),
                             f"{value.real:.16E}", f"{value.imag:.16E}"])
    return rows


def _smatrix_atom_diatom(result, energy_indices, block_index, state, v, j, state_prime, v_prime, j_prime) -> str:
    """Atom-diatom table: (state) v j L -> (state') v' j' L' (report.py:809-882)."""
    internal = []
    for ch in result.basis:
        active = [s for s in (ch.mis_X, ch.mis_Y) if not _is_atom(s)]
        if len(active) != 1 or active[0].v is None:
            raise fail(ValueError, "The atom-diatom S-matrix report requires one diatomic internal state per channel")
        internal.append(active[0])
    f_in = (_selection(state, "state"), _selection(v, "v"), _selection(j, "j"))
    f_out = (_selection(state_prime, "state_prime"), _selection(v_prime, "v_prime"), _selection(j_prime, "j_prime"))
    if not any(_matches(s, *f_in) for s in internal):
        raise fail(ValueError, "The initial-state selection does not match any channel")
    if not any(_matches(s, *f_out) for s in internal):
        raise fail(ValueError, "The final-state selection does not match any channel")
    elec = any(s.electronic_state is not None for s in internal)

    def labels(g, L):
        s = internal[g]
        return ([str(s.electronic_state)] if elec else []) + [str(s.v), str(s.j), _integral_L(L)]

    headers = ["Etot/cm-1"] + (["state"] if elec else []) + ["v", "j", "L"] + (["state'"] if elec else []) + ["v'", "j'", "L'", "Re(S)", "Im(S)"]
    rows = _smatrix_rows(result, energy_indices, block_index, lambda g: _matches(internal[g], *f_in),
                         lambda g: _matches(internal[g], *f_out), labels)
    return _table(headers, rows)


def _smatrix_diatom_diatom(result, energy_indices, block_index, sel) -> str:
    """Diatom-diatom table: v_X j_X v_Y j_Y j_couple L -> primed (report.py:887-982)."""
    basis = result.basis
    if any(c.mis_X.v is None or c.mis_Y.v is None for c in basis):
        raise fail(ValueError, "The diatom-diatom S-matrix report requires two diatomic internal states per channel")
    fX = tuple(_selection(sel[k], k) for k in ("state_X", "v_X", "j_X"))
    fY = tuple(_selection(sel[k], k) for k in ("state_Y", "v_Y", "j_Y"))
    fJ = _selection(sel["j_couple"], "j_couple")
    fXp = tuple(_selection(sel[k], k) for k in ("state_X_prime", "v_X_prime", "j_X_prime"))
    fYp = tuple(_selection(sel[k], k) for k in ("state_Y_prime", "v_Y_prime", "j_Y_prime"))
    fJp = _selection(sel["j_couple_prime"], "j_couple_prime")

    def ok(ch, fx, fy, fj):
        return _matches(ch.mis_X, *fx) and _matches(ch.mis_Y, *fy) and (fj is None or ch.j_couple in fj)

    if not any(ok(c, fX, fY, fJ) for c in basis):
        raise fail(ValueError, "The initial diatom-diatom selection does not match any channel")
    if not any(ok(c, fXp, fYp, fJp) for c in basis):
        raise fail(ValueError, "The final diatom-diatom selection does not match any channel")
    eX = any(c.mis_X.electronic_state is not None for c in basis)
    eY = any(c.mis_Y.electronic_state is not None for c in basis)

    def head(p):
        return (["state_X" + p] if eX else []) + ["v_X" + p, "j_X" + p] + (["state_Y" + p] if eY else []) + ["v_Y" + p, "j_Y" + p]

    def labels(g, L):
        ch = basis[g]
        out = ([str(ch.mis_X.electronic_state)] if eX else []) + [str(ch.mis_X.v), str(ch.mis_X.j)]
        out += ([str(ch.mis_Y.electronic_state)] if eY else []) + [str(ch.mis_Y.v), str(ch.mis_Y.j)]
        return out + [str(ch.j_couple), _integral_L(L)]

    headers = ["Etot/cm-1", *head(""), "j_couple", "L", *head("'"), "j_couple'", "L'", "Re(S)", "Im(S)"]
    rows = _smatrix_rows(result, energy_indices, block_index, lambda g: ok(basis[g], fX, fY, fJ),
                         lambda g: ok(basis[g], fXp, fYp, fJp), labels)
    return _table(headers, rows)


_DIATOM_DIATOM_KEYS = ("state_X", "v_X", "j_X", "state_Y", "v_Y", "j_Y", "j_couple", "state_X_prime", "v_X_prime",
                       "j_X_prime", "state_Y_prime", "v_Y_prime", "j_Y_prime", "j_couple_prime")


def smatrix(result, *, energy_indices=None, state=None, v=None, j=None, state_prime=None, v_prime=None, j_prime=None,
            state_X=None, v_X=None, j_X=None, state_Y=None, v_Y=None, j_Y=None, j_couple=None, state_X_prime=None,
            v_X_prime=None, j_X_prime=None, state_Y_prime=None, v_Y_prime=None, j_Y_prime=None, j_couple_prime=None,
            block_index: int | None = None) -> str:
    """Selected S-matrix elements as a text table (report.py:988-1247, field-free families).

    Initial-state filters select columns and final-state filters select rows of
    S[out, in]; each filter is an int, a range, an integer sequence or None (all).
    A coupled-states result reports one K block (`block_index`, optional when
    there is only one).
    """
    if not isinstance(result, (ScatteringResult, CoupledStatesResult)):
        raise fail(TypeError, f"smatrix expects a ScatteringResult or CoupledStatesResult, got {type(result).__name__}")
    dd = dict(zip(_DIATOM_DIATOM_KEYS, (state_X, v_X, j_X, state_Y, v_Y, j_Y, j_couple, state_X_prime, v_X_prime,
                                        j_X_prime, state_Y_prime, v_Y_prime, j_Y_prime, j_couple_prime)))
    basis = result.basis
    active_X = any(not _is_atom(c.mis_X) for c in basis)
    active_Y = any(not _is_atom(c.mis_Y) for c in basis)
    if active_X and active_Y:
        _reject("diatom-diatom", (("state", state), ("v", v), ("j", j), ("state_prime", state_prime), ("v_prime", v_prime),
                                  ("j_prime", j_prime)))
        return _smatrix_diatom_diatom(result, energy_indices, block_index, dd)
    _reject("atom-diatom", tuple(dd.items()))
    return _smatrix_atom_diatom(result, energy_indices, block_index, state, v, j, state_prime, v_prime, j_prime)
