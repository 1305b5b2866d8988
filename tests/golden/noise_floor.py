"""FP64 noise floor of the N = 800 end-to-end S-matrix (CPU, about 7 minutes).

Runs the oracle with explicit inverses in the LDMD sector step instead of linear
solves -- an equally valid FP64 restatement of the same algorithm -- at two of the
energies of tests/golden/smat_rotor_rotor_N800.npz and prints the largest |dS| against the
fixture.  Measured: 4.2e-11 (energy index 31) and 4.2e-10 (energy index 63).

    python tests/golden/noise_floor.py
"""
import sys
import time
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parents[2]))
import numpy as np
from oracle import ldmd
from oracle import pipeline as opipe
from oracle import matching as omatch
from pyticc_amd import synthetic
from tests.helpers import oracle_problem

def sector_step_inv(Y, h, ref, Qs, Qm, Qe):
    y1, y2 = ldmd.reference_values(h, ref)
    n = Y.shape[-1]; idx = np.arange(n)
    A1 = Y + Qs; A1[idx, idx] += y1
    X1 = np.linalg.inv(A1) * y2[None, :]
    Ymid = Qm.copy(); Ymid[idx, idx] += y1; Ymid -= y2[:, None] * X1
    A2 = Ymid + Qm; A2[idx, idx] += y1
    X2 = np.linalg.inv(A2) * y2[None, :]
    Yend = Qe.copy(); Yend[idx, idx] += y1; Yend -= y2[:, None] * X2
    return 0.5 * (Yend + Yend.T)

ldmd.sector_step = sector_step_inv
d = np.load(Path(__file__).resolve().parent / 'smat_rotor_rotor_N800.npz')
w = synthetic.rotor_rotor()
_, _, prob = oracle_problem(w)
idx = [int(d['energy_index'][1]), int(d['energy_index'][2])]
t = time.time()
_, _, _, S = opipe.solve(prob, w.energies[idx])
for k, i in enumerate((1, 2)):
    cols = d[f'S{i}_cols']
    print(f"energy {idx[k]}: max|dS| inv-vs-solve oracle = {np.max(np.abs(S[k][:, :cols.shape[1]] - cols)):.3e}, diag {np.max(np.abs(np.diag(S[k]) - d[f'S{i}_diag'])):.3e}, defect {omatch.unitarity_defect(S[k]):.2e}", flush=True)
print(f"{time.time()-t:.0f} s")
