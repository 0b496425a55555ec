"""Fermion models on the CPU (oracle; test infrastructure only): the square
lattice of the reference, and exact many-body Hamiltonians of the Hubbard
model and of the spin-1/2 Heisenberg model built from first principles.

Restates:
  squareLattice (bond list) ...... mps/lattice/square.h:24-48
  Electron local basis ........... mps/sites/electron.h:34-52 (Emp, Up, Dn, UpDn)

The Hubbard Hamiltonian here is built from 2N spinless fermion modes in the
order (1 up, 1 dn, 2 up, 2 dn, ...) with the Jordan-Wigner form
c_m = Z_1 ... Z_{m-1} a_m, Z = (-1)^n -- the textbook construction, not the
MPO's site-operator algebra -- and written in the site basis of Electron:
the local state of site i is n_up + 2 n_dn (Emp 0, Up 1, Dn 2, UpDn 3), and
site 1 is the fastest-running (column-major) digit.  Pinned by
tests/test_oracle_hubbard.py: two-site closed form (U=0 band energies, the
t-U dimer ground energy (U - sqrt(U^2 + 16 t^2))/2), particle-hole and
spin-flip symmetry, hermiticity, and the non-interacting limit against the
single-particle spectrum of the hopping matrix.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def square_lattice(Nx: int, Ny: int, yperiodic: bool = False):
    """Bonds (i, j) of the Nx x Ny square lattice, sites n = (x-1)*Ny + y,
    in the reference's order (x bond, y bond, periodic y bond per site);
    periodicity in y only for Ny > 2 (square.h:28-30)."""
    yperiodic = yperiodic and Ny > 2
    bonds = []
    for n in range(1, Nx * Ny + 1):
        x = (n - 1) // Ny + 1
        y = (n - 1) % Ny + 1
        if x < Nx:
            bonds.append((n, n + Ny))
        if Ny > 1:
            if y < Ny:
                bonds.append((n, n + 1))
            if yperiodic and y == 1:
                bonds.append((n, n + Ny - 1))
    return bonds


def _mode_ops(nmodes: int):
    """Annihilators c_m (sparse CSR, 2^nmodes), mode 0 the least significant
    bit, with Jordan-Wigner strings over the lower modes."""
    I2 = sp.identity(2, format="csr")
    Z = sp.diags([1.0, -1.0], format="csr")
    a = sp.csr_matrix(np.array([[0.0, 1.0], [0.0, 0.0]]))  # |0><1|
    ops = []
    for m in range(nmodes):
        M = None
        for k in range(nmodes - 1, -1, -1):  # most significant mode first
            f = a if k == m else (Z if k < m else I2)
            M = f if M is None else sp.kron(M, f, format="csr")
        ops.append(M)
    return ops


def hubbard_hamiltonian(N: int, bonds, t: float, U: float):
    """H = -t sum_<ij>,s (c+_is c_js + h.c.) + U sum_i n_i,up n_i,dn on N
    sites (sparse 4^N x 4^N, site basis as in the module docstring)."""
    c = _mode_ops(2 * N)
    H = sp.csr_matrix((4 ** N, 4 ** N))
    for (i, j) in bonds:
        for s in (0, 1):
            ci, cj = c[2 * (i - 1) + s], c[2 * (j - 1) + s]
            hop = ci.T @ cj
            H = H - t * (hop + hop.T)
    for i in range(1, N + 1):
        nu = c[2 * (i - 1)].T @ c[2 * (i - 1)]
        nd = c[2 * (i - 1) + 1].T @ c[2 * (i - 1) + 1]
        H = H + U * (nu @ nd)
    return H.tocsr()


def electron_sector(N: int, nf: int, sz2: int) -> np.ndarray:
    """Basis indices with sum n = nf and 2 Sz = sz2."""
    idx = np.arange(4 ** N)
    nu = np.zeros(4 ** N, dtype=int)
    nd = np.zeros(4 ** N, dtype=int)
    for i in range(N):
        loc = (idx // 4 ** i) % 4
        nu += loc & 1
        nd += (loc >> 1) & 1
    return np.nonzero((nu + nd == nf) & (nu - nd == sz2))[0]


def ground_energy(H, sector=None) -> float:
    """Lowest eigenvalue (dense solver on the sector block)."""
    if sector is not None:
        H = H[sector][:, sector]
    H = H.toarray() if sp.issparse(H) else H
    return float(np.linalg.eigvalsh(H)[0])


def heisenberg_hamiltonian(N: int, bonds, J: float = 1.0) -> np.ndarray:
    """H = J sum_<ij> S_i.S_j on spin-1/2 sites (Up = 0, Dn = 1; site 1 the
    fastest digit)."""
    sz = np.diag([0.5, -0.5])
    sp = np.array([[0.0, 1.0], [0.0, 0.0]])  # S+ |Dn> = |Up> (row = output)
    sm = sp.T

    def site_op(o, i):
        M = np.eye(1)
        for k in range(N, 0, -1):  # most significant (site N) first
            M = np.kron(M, o if k == i else np.eye(2))
        return M

    H = np.zeros((2 ** N, 2 ** N))
    for (i, j) in bonds:
        H += J * (site_op(sz, i) @ site_op(sz, j)
                  + 0.5 * (site_op(sp, i) @ site_op(sm, j) + site_op(sm, i) @ site_op(sp, j)))
    return H
