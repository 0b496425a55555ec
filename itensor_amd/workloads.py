"""Synthetic workloads of the DMRG H*psi path, built through the host layer.

The shapes of the benchmark configurations and of the tests that run them
at full size:
  * dense two-site H*psi: operand dims and labels of the four stages of
    LocalOp::product (mps/localop.h:346-364) and of the fused chain the
    host layer runs (MPO pair pre-contracted, W12 = W1*W2);
  * block-sparse H*psi at chi = 1024 over 31 Sz sectors (Heisenberg chain
    midpoint);
  * the Sz-conserving Heisenberg chain (MPO of sample/Heisenberg.h:76-116
    and the Neel product state) and the (Nf, Sz)-conserving Hubbard cylinder
    (MPO from the host AutoMPO, half-filled product state);
  * the maxdim schedules of the DMRG runs.
"""
from __future__ import annotations

import numpy as np

def stage_dims(chi: int, d: int = 2, k: int = 5):
    psi = (chi, d, d, chi)
    Ld = (chi, k, chi)
    W = (k, d, d, k)
    T1 = (d, d, chi, k, chi)
    T2 = (d, chi, chi, d, k)
    T3 = (chi, chi, d, d, k)
    return [(psi, Ld, T1), (T1, W, T2), (T2, W, T3), (T3, Ld, psi)]


# labels of the four stages as ITensor's computeLabels assigns them
STAGES = [
    ((-1, 2, 3, 4), (-1, 5, 6), (2, 3, 4, 5, 6)),
    ((-1, 3, 4, -2, 5), (-2, -1, 6, 7), (3, 4, 5, 6, 7)),
    ((-1, 3, 4, 5, -2), (-2, -1, 6, 7), (3, 4, 5, 6, 7)),
    ((-1, 3, 4, 5, -2), (-1, -2, 6), (3, 4, 5, 6)),
]
# fused chain of LocalOp::product: S1, then T1 x W12 with W12 = W1*W2
# (h0,s1,s1',s2,s2',h2), then S4 -- same result layout and flop count
FUSED = [
    STAGES[0],
    ((-1, -2, 3, -3, 4), (-3, -1, 5, -2, 6, 7), (3, 4, 5, 6, 7)),
    STAGES[3],
]


def fused_dims(chi: int, d: int = 2, k: int = 5):
    psi = (chi, d, d, chi)
    Ld = (chi, k, chi)
    T1 = (d, d, chi, k, chi)
    W12 = (k, d, d, d, d, k)
    T3 = (chi, chi, d, d, k)
    return [(psi, Ld, T1), (T1, W12, T3), (T3, Ld, psi)]


def heisenberg_w() -> np.ndarray:
    """Bulk Heisenberg MPO tensor W(h_in, s, s', h_out) (J = Jz = 1)."""
    Id = np.eye(2)
    Sz = np.diag([0.5, -0.5])
    Sp = np.zeros((2, 2)); Sp[1, 0] = 1.0
    Sm = np.zeros((2, 2)); Sm[0, 1] = 1.0
    W = np.zeros((5, 2, 2, 5))
    W[0, :, :, 0] = Id; W[1, :, :, 1] = Id
    W[2, :, :, 0] = Sz; W[1, :, :, 2] = Sz
    W[3, :, :, 0] = Sm; W[1, :, :, 3] = 0.5 * Sp
    W[4, :, :, 0] = Sp; W[1, :, :, 4] = 0.5 * Sm
    return W


def hpsi_flops(chi: int, d: int = 2, k: int = 5) -> float:
    return 4.0 * d * d * k * chi ** 3 + 4.0 * d ** 3 * k * k * chi ** 2


# bond sectors q = 2 Sz in {-30..30} of the chi = 1024 Sz-conserving
# Heisenberg bond at the chain midpoint (binomial weights, max 120, min 1)
SECTORS_1024 = [1, 1, 1, 1, 1, 1, 3, 7, 14, 25, 41, 60, 81, 101, 114, 120, 114, 101, 81, 60, 41, 25, 14, 7, 3,
                1, 1, 1, 1, 1, 1]


def qn_operands(H, sectors, seed):
    """Block-sparse two-site operands: phi(dag(l), s1, s2, r), L(l, hl,
    dag(l)'), R(dag(r), dag(hr), r'), W(dag(s), s', dag(h_in), h_out);
    site {+1: 1, -1: 1}, MPO link {0: 3, -2: 1, +2: 1}; flux-0 blocks of
    phi, L, R filled with N(0,1)."""
    def sz(v):
        return [("Sz", int(v))]

    ids = iter(range(7001, 7100))

    def idx(qns, tags):
        return H.Index(next(ids), sum(d for _, d in qns), 1, 0, tags, qns)

    bond = [(sz(q), d) for q, d in zip(range(-30, 31, 2), sectors)]
    l, r = idx(bond, "Link,l=49"), idx(bond, "Link,l=51")
    site = [(sz(1), 1), (sz(-1), 1)]
    s1, s2 = idx(site, "Site,n=50"), idx(site, "Site,n=51")
    hq = [(sz(0), 3), (sz(-2), 1), (sz(2), 1)]
    hl, hm, hr = idx(hq, "Link,h=49"), idx(hq, "Link,h=50"), idx(hq, "Link,h=51")
    W = heisenberg_w().transpose(1, 2, 0, 3)  # (s, s', h_in, h_out)
    W1 = H.from_dense([s1.dag(), s1.prime(), hl.dag(), hm], W)
    W2 = H.from_dense([s2.dag(), s2.prime(), hm.dag(), hr], W)
    phi = H.random_itensor([l.dag(), s1, s2, r], [], seed)
    L = H.random_itensor([l, hl, l.dag().prime()], [], seed + 1)
    R = H.random_itensor([r.dag(), hr.dag(), r.prime()], [], seed + 2)
    return phi, L, W1, W2, R


def heisenberg_qn_chain(H, N: int):
    """Sz-conserving spin-1/2 Heisenberg chain (J = Jz = 1, open ends):
    site indices, the MPO (bulk W(dag(s), s', dag(h_in), h_out), first and
    last site contracted with the start / end states of the MPO link,
    sample/Heisenberg.h:76-116) and the Neel product state with link QNs
    that give every site tensor flux 0."""
    def sz(v):
        return [("Sz", int(v))]

    ids = iter(range(100001, 100001 + 4 * N + 8))

    def idx(qns, tags):
        return H.Index(next(ids), sum(d for _, d in qns), 1, 0, tags, qns)

    sites = [idx([(sz(1), 1), (sz(-1), 1)], f"Site,S=1/2,n={n}") for n in range(1, N + 1)]
    hq = [(sz(0), 3), (sz(-2), 1), (sz(2), 1)]
    hl = [idx(hq, f"Link,H,l={n}") for n in range(N + 1)]
    Wb = heisenberg_w().transpose(1, 2, 0, 3)  # (s, s', h_in, h_out)
    mpo = []
    for n in range(1, N + 1):
        s, row, col = sites[n - 1], hl[n - 1].dag(), hl[n]
        if n == 1:
            W, inds = Wb[:, :, 1, :], [s.dag(), s.prime(), col]
        elif n == N:
            W, inds = Wb[:, :, :, 0], [s.dag(), s.prime(), row]
        else:
            W, inds = Wb, [s.dag(), s.prime(), row, col]
        mpo.append(H.from_dense(inds, np.ascontiguousarray(W)))
    # Neel state: link l_n carries -(sum of the site charges 1..n)
    A, q, prev = [], 0, None
    for n in range(1, N + 1):
        st = (n - 1) % 2  # 0 = Up (Sz +1), 1 = Dn
        q -= 1 if st == 0 else -1
        s = sites[n - 1]
        v = np.zeros(2)
        v[st] = 1.0
        if n < N:
            link = idx([(sz(q), 1)], f"Link,l={n}")
        if n == 1:
            A.append(H.from_dense([s, link], v.reshape(2, 1)))
        elif n == N:
            A.append(H.from_dense([prev.dag(), s], v.reshape(1, 2)))
        else:
            A.append(H.from_dense([prev.dag(), s, link], v.reshape(1, 2, 1)))
        prev = link if n < N else None
    return mpo, A


def hubbard_cylinder(H, Nx: int, Ny: int, t: float = 1.0, U: float = 4.0, yperiodic: bool = True):
    """Hubbard model on the Nx x Ny square lattice (YPeriodic cylinder, site
    order of squareLattice) with (Nf, Sz) conserved: the MPO from the host
    layer's AutoMPO (-t c+_i,s c_j,s + h.c. on every bond, U n_up n_dn on
    every site) and a half-filled product state (Up, Dn, Up, ...) whose link
    QNs give every site tensor but the last flux 0; the last one carries the
    state's total charge (Nf = N, Sz = 0)."""
    N = Nx * Ny
    sites = H.electron_sites(N)
    ampo = H.AutoMPO(sites)
    for (i, j) in H.square_lattice(Nx, Ny, yperiodic):
        for c, a in (("Cdagup", "Cup"), ("Cdagdn", "Cdn")):
            ampo.add(-t, c, i, a, j)
            ampo.add(-t, c, j, a, i)
    for i in range(1, N + 1):
        ampo.add(U, "Nupdn", i)
    mpo = ampo.to_mpo()
    ids = iter(range(300001, 300001 + N + 8))
    A, nf, sz, prev = [], 0, 0, None
    for n in range(1, N + 1):
        st = 1 if (n - 1) % 2 == 0 else 2  # Up, Dn
        nf += 1
        sz += 1 if st == 1 else -1
        s = sites(n)
        v = np.zeros(4)
        v[st] = 1.0
        if n < N:
            q = [(k, x) for k, x in (("Nf", -nf), ("Sz", -sz)) if x]
            link = H.Index(next(ids), 1, 1, 0, f"Link,l={n}", [(q, 1)])
        if n == 1:
            A.append(H.from_dense([s, link], v.reshape(4, 1)))
        elif n == N:
            tot = [(k, x) for k, x in (("Nf", nf), ("Sz", sz)) if x]
            A.append(H.from_dense([prev.dag(), s], v.reshape(1, 4), tot))
        else:
            A.append(H.from_dense([prev.dag(), s, link], v.reshape(1, 4, 1)))
        prev = link if n < N else None
    return mpo, A


DMRG_MAXDIM = [10, 20, 50, 100, 200, 400, 800, 1200, 1600, 2048]
# MinDim = MaxDim ramp of the benchmark's large-bond leg: the bond dimension
# follows it instead of saturating at the cutoff (about 110 on this chain);
# it stops at 1024 so that the leg fits the benchmark's time budget
DMRG_BIG_MAXDIM = [10, 20, 50, 100, 200, 400, 600, 800, 1024]
HUBBARD_MAXDIM = [50, 100, 200, 400, 400, 400, 400, 400]  # ramp, then converged at 400
# the benchmark's large-bond Hubbard leg: the maxdim ramp toward 4096, stopped at
# 2500 to fit the benchmark's time budget
HUBBARD_BIG_MAXDIM = [50, 100, 200, 400, 800, 1200, 1600, 2000, 2500]
