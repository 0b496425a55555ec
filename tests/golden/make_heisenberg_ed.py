"""Writes heisenberg_ed.json: exact ground-state energies of the open
spin-1/2 Heisenberg chain H = sum_i S_i.S_{i+1} (the model of
sample/Heisenberg.h, J = Jz = 1) by Lanczos (scipy eigsh) in the Sz = 0
sector.  These pin the oracle's DMRG energies (config: N = 20)."""
import json
import os
from itertools import combinations

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import eigsh


def sector_states(N):
    states = []
    for ups in combinations(range(N), N // 2):
        s = 0
        for u in ups:
            s |= 1 << u
        states.append(s)
    return np.array(sorted(states), dtype=np.int64)


def heisenberg_sz0(N):
    st = sector_states(N)
    dim = st.size
    diag = np.zeros(dim)
    rows, cols, vals = [], [], []
    for i in range(N - 1):
        bi = (st >> i) & 1
        bj = (st >> (i + 1)) & 1
        same = bi == bj
        diag += np.where(same, 0.25, -0.25)
        flip = st[~same] ^ ((1 << i) | (1 << (i + 1)))
        src = np.nonzero(~same)[0]
        dst = np.searchsorted(st, flip)
        rows.append(dst); cols.append(src); vals.append(np.full(src.size, 0.5))
    rows = np.concatenate(rows + [np.arange(dim)])
    cols = np.concatenate(cols + [np.arange(dim)])
    vals = np.concatenate(vals + [diag])
    return sp.csr_matrix((vals, (rows, cols)), shape=(dim, dim))


def ground_energy(N):
    H = heisenberg_sz0(N)
    if H.shape[0] < 50:
        return float(np.linalg.eigvalsh(H.toarray())[0])
    w = eigsh(H, k=1, which="SA", tol=1e-14, maxiter=100000, v0=np.ones(H.shape[0]))[0]
    return float(w[0])


if __name__ == "__main__":
    out = {"model": "open spin-1/2 Heisenberg chain, J=Jz=1, Sz=0 sector, Lanczos (scipy eigsh)",
           "energies": {str(N): ground_energy(N) for N in (4, 6, 8, 10, 12, 20)}}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "heisenberg_ed.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))
