"""DMRG hot path on the CPU (oracle): davidson, truncate, diag_hermitian,
denmatDecomp, LocalOp/LocalMPO, the two-site DMRG sweep, and the model
builders used by the tests and benchmarks.

Restates:
  davidson ...................... iterativesolvers.h:99-451
  truncate ...................... decomp.cc:305-463
  diag_hermitian (dense, QN) .... hermitian.cc:40-426, tensor/algs_impl.h:91-139
  denmatDecomp .................. decomp.h:316-418
  MPS::svdBond (both routes) .... mps/mps_impl.h:21-99
  LocalOp::product .............. mps/localop.h:324-365
  LocalMPO position/makeL/makeR . mps/localmpo.h:443-618
  DMRGWorker / sweepnext ........ mps/dmrg.h:336-467, mps/sweeps.h:489-500
  Heisenberg MPO ................ sample/Heisenberg.h:55-116
  SpinHalf sites / operators .... mps/sites/spinhalf.h:39-166
  random circuit MPS layout ..... mps/mps.cc:222-260 (own seeded RNG)
Pinned by tests/test_oracle_dmrg.py: closed-form energies of the
reference's tests (iterativesolvers_test.cc:45-65 bond problem
-(1/4+1/sqrt 2); 4-site chain -1.6160254038; mpo_test.cc:541-602 critical
TFIM), exact diagonalisation of the N=20 chain (tests/golden), the
reference's truncate cases (decomp_test.cc:59-138) and reconstruction
identities of diag_hermitian / denmatDecomp.
"""
from __future__ import annotations

import math

import numpy as np

from .itensor import (ITensor, Index, QN, Out, In, add_into, combiner, contract, dense, dot,
                      index_position, norm, qdense)

REAL_EPSILON = np.finfo(float).eps


# ------------------------------------------------------------- RNG ------
class SplitMix64:
    """Counter-based generator (SplitMix64) + Box-Muller normals; the same
    stream is produced by the C++ host layer."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.s = seed & self.MASK

    def next_u64(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & self.MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return ((self.next_u64() >> 11) + 0.5) * (1.0 / 9007199254740992.0)

    def normal(self) -> float:
        u1 = self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def normals(self, n: int) -> np.ndarray:
        return np.array([self.normal() for _ in range(n)])


def random_orth_rows(r: int, c: int, rng: SplitMix64) -> np.ndarray:
    """r x c matrix (r <= c) with orthonormal rows: Gaussian rows (row-major
    draw order) orthonormalised by modified Gram-Schmidt."""
    X = rng.normals(r * c).reshape(r, c)
    for i in range(r):
        for j in range(i):
            X[i] -= (X[j] @ X[i]) * X[j]
        X[i] /= np.linalg.norm(X[i])
    return X


# ------------------------------------------------------------ sites ------
def spin_half_sites(N: int, conserve_qns: bool):
    """SpinHalf site indices (sites/spinhalf.h:39-60): states Up, Dn with
    QN Sz = +1, -1 (units of 2 Sz) when conserving."""
    if conserve_qns:
        return [Index([(QN(("Sz", 1)), 1), (QN(("Sz", -1)), 1)], f"Site,S=1/2,n={n}", Out)
                for n in range(1, N + 1)]
    return [Index(2, f"Site,S=1/2,n={n}") for n in range(1, N + 1)]


def spin_half_op(name: str) -> np.ndarray:
    """op(s, s') = <s'|O|s> (stored as ITensor(dag(s), prime(s)))."""
    if name == "Id":
        return np.eye(2)
    if name == "Sz":
        return np.diag([0.5, -0.5])
    o = np.zeros((2, 2))
    if name == "Sp":
        o[1, 0] = 1.0
    elif name == "Sm":
        o[0, 1] = 1.0
    elif name == "Sx":
        o[0, 1] = o[1, 0] = 0.5
    else:
        raise ValueError(name)
    return o


def heisenberg_mpo(sites, J=1.0, Jz=None):
    """Heisenberg MPO (sample/Heisenberg.h:76-116), W_n indices
    (dag(s), s', row, col); edges contracted with the start/end vectors."""
    if Jz is None:
        Jz = J
    N = len(sites)
    qn = sites[0].has_qns()
    if qn:
        links = [Index([(QN(("Sz", 0)), 3), (QN(("Sz", -2)), 1), (QN(("Sz", 2)), 1)], f"Link,l={l}", Out)
                 for l in range(N + 1)]
    else:
        links = [Index(5, f"Link,l={l}") for l in range(N + 1)]
    terms = [("Id", 0, 0, 1.0), ("Id", 1, 1, 1.0), ("Sz", 2, 0, 1.0), ("Sz", 1, 2, Jz),
             ("Sm", 3, 0, 1.0), ("Sp", 1, 3, J / 2), ("Sp", 4, 0, 1.0), ("Sm", 1, 4, J / 2)]
    H = []
    for n in range(1, N + 1):
        s = sites[n - 1]
        row = links[n - 1].dag()
        col = links[n]
        W = np.zeros((2, 2, 5, 5))
        for op, r, c, f in terms:
            W[:, :, r, c] += f * spin_half_op(op)
        inds = [s.dag(), s.prime(), row, col]
        if n == 1:
            W = W[:, :, 1, :]  # *= setElt(links[0](2))
            inds = [s.dag(), s.prime(), col]
        elif n == N:
            W = W[:, :, :, 0]  # *= setElt(dag(last)(1))
            inds = [s.dag(), s.prime(), row]
        if N == 1:
            raise ValueError("N >= 2")
        H.append(_from_dense(inds, W, qn))
    return H


def tfim_mpo(sites, h: float):
    """Transverse-field Ising chain H = -sum Sx_j Sx_{j+1} - h sum Sz_j, the
    model of mpo_test.cc:541-602, as a bond-3 MPO (same start/end
    convention as the Heisenberg MPO)."""
    N = len(sites)
    links = [Index(3, f"Link,l={l}") for l in range(N + 1)]
    H = []
    for n in range(1, N + 1):
        s = sites[n - 1]
        W = np.zeros((2, 2, 3, 3))
        W[:, :, 0, 0] = np.eye(2)
        W[:, :, 1, 1] = np.eye(2)
        W[:, :, 2, 0] = spin_half_op("Sx")
        W[:, :, 1, 2] = -1.0 * spin_half_op("Sx")
        W[:, :, 1, 0] = -h * spin_half_op("Sz")
        inds = [s.dag(), s.prime(), links[n - 1].dag(), links[n]]
        if n == 1:
            W = W[:, :, 1, :]
            inds = [s.dag(), s.prime(), links[n]]
        elif n == N:
            W = W[:, :, :, 0]
            inds = [s.dag(), s.prime(), links[n - 1].dag()]
        H.append(_from_dense(inds, W, False))
    return H


def _from_dense(inds, arr, qn: bool, div: QN = QN()) -> ITensor:
    if not qn:
        return dense([i.remove_qns() for i in inds], arr)
    T = qdense(inds, div)
    # every flux-allowed block gets its slice of the dense array
    for b, _ in T.offsets:
        sl = tuple(slice(I.block_start(bb), I.block_start(bb) + I.blocksize(bb)) for I, bb in zip(inds, b))
        T.block(b)[...] = arr[sl]
    full = T.to_dense_array()
    if not np.allclose(full, arr):
        raise ValueError("array has elements outside the flux-0 blocks")
    return T


class MPS:
    """Site tensors A[1..N] (stored 0-based), with orthogonality limits."""

    def __init__(self, A, sites, links, leftlim=0, rightlim=2):
        self.A = A
        self.sites = sites
        self.links = links
        self.leftlim = leftlim
        self.rightlim = rightlim

    @property
    def N(self):
        return len(self.A)

    def __call__(self, j):  # 1-based, like psi(j)
        return self.A[j - 1]

    def set(self, j, T):
        self.A[j - 1] = T

    def copy(self):
        return MPS([a.copy() for a in self.A], self.sites, list(self.links), self.leftlim, self.rightlim)


def random_mps(sites, m: int, seed: int) -> MPS:
    """Right-orthogonal random MPS with the layout of randomCircuitMPS
    (mps.cc:222-260): A_1 (s1,l1), A_j (l_{j-1}, s_j, l_j), A_N (l_{N-1}, s_N);
    leftLim 0, rightLim 2.  Dense sites only."""
    rng = SplitMix64(seed)
    N = len(sites)
    d = [s.dim for s in sites]
    links = [None] * (N + 1)
    A = [None] * N
    chi = min(m, d[N - 1])
    links[N - 1] = Index(chi, f"Link,l={N - 1}")
    O = random_orth_rows(chi, d[N - 1], rng)
    A[N - 1] = dense([links[N - 1], sites[N - 1]], O)
    for j in range(N - 1, 1, -1):
        prev = chi
        chi = min(m, chi * d[j - 1])
        links[j - 1] = Index(chi, f"Link,l={j - 1}")
        O = random_orth_rows(chi, prev * d[j - 1], rng)
        # combined index value c = s + d*l  (combiner(s_j, l_j), s fastest)
        T = O.reshape(chi, d[j - 1], prev, order="F")
        A[j - 1] = dense([links[j - 1], sites[j - 1], links[j]], T)
    O = random_orth_rows(1, d[0] * chi, rng)
    A[0] = dense([sites[0], links[1]], O.reshape(d[0], chi, order="F"))
    return MPS(A, sites, links, 0, 2)


def product_mps(sites, states) -> MPS:
    """Product state (bond dim 1), MPS(InitState) layout (mps.cc:558-590):
    links a_i = Index(qa_i, 1) (Out) with qa_i = QN of sites i+1..N, tensors
    A_1 (s1, a1), A_i (dag(a_{i-1}), s_i, a_i), A_N (dag(a_{N-1}), s_N); A_1
    carries the total QN as its flux.  Orthogonality limits (0, 2)."""
    N = len(sites)
    qn = sites[0].has_qns()
    links = [None] * (N + 1)
    if qn:
        q = [sites[i].qn(states[i]) for i in range(N)]
        total = QN()
        for x in q:
            total = total + x
        qa = total
        for j in range(1, N):
            qa = qa - q[j - 1]
            links[j] = Index([(qa, 1)], f"Link,l={j}", Out)
    else:
        for j in range(1, N):
            links[j] = Index(1, f"Link,l={j}")
    A = []
    for j in range(1, N + 1):
        s = sites[j - 1]
        if j == 1:
            inds = [s, links[1]]
        elif j == N:
            inds = [links[N - 1].dag(), s]
        else:
            inds = [links[j - 1].dag(), s, links[j]]
        arr = np.zeros(tuple(i.dim for i in inds))
        idx = [0] * len(inds)
        idx[[k for k, i in enumerate(inds) if i == s][0]] = states[j - 1]
        arr[tuple(idx)] = 1.0
        div = QN()
        if qn and j == 1:
            div = sites[0].qn(states[0]) + links[1].qn(0)
        A.append(_from_dense(inds, arr, qn, div))
    return MPS(A, sites, links, 0, 2)


# -------------------------------------------------------- truncate ------
def truncate(P: np.ndarray, maxdim: int, mindim: int, cutoff: float, absolute_cutoff=False,
             do_rel_cutoff=True, respect_degenerate=False):
    """truncate (decomp.cc:305-463). P is descending.  Returns
    (P_kept, truncerr, docut_lower, docut_upper, ndegen_below)."""
    P = np.array(P, dtype=np.float64)
    origm = P.size
    n = origm - 1
    if P[0] == 0.0:
        return P[:1].copy(), 0.0, -REAL_EPSILON, REAL_EPSILON, 1
    if origm == 1:
        dc = 1e-3 * P[0]
        return P.copy(), 0.0, P[0] - dc, P[0] + dc, 1
    P0 = P[0]
    P = P / P0
    if absolute_cutoff or not do_rel_cutoff:
        cutoff = cutoff / P0
    for zn in range(n, -1, -1):
        if P[zn] >= 0:
            break
        P[zn] = 0
    truncerr = 0.0
    while n >= maxdim:
        truncerr += P[n]
        n -= 1
    if absolute_cutoff:
        while P[n] <= cutoff and n >= mindim:
            truncerr += P[n]
            n -= 1
    else:
        scale = 1.0
        if do_rel_cutoff:
            scale = float(np.sum(P))
            if scale == 0.0:
                scale = 1.0
        while truncerr + P[n] <= cutoff * scale and n >= mindim:
            truncerr += P[n]
            n -= 1
        truncerr = 0 if scale == 0 else truncerr / scale
    if n < 0:
        n = 0
    m = n + 1
    degen_cutoff = 1e-3 * P[n]
    if P[n] == 0.0:
        degen_cutoff = REAL_EPSILON
    docut_lower = P[n] - degen_cutoff
    docut_upper = P[n] + degen_cutoff
    ndegen_below = 1
    if n > 0:
        while n - ndegen_below >= 0 and abs(P[n - ndegen_below] - P[n]) < degen_cutoff:
            ndegen_below += 1
    ndegen_above = 0
    if n + 1 < origm:
        while n + 1 + ndegen_above < origm and abs(P[n + 1 + ndegen_above] - P[n]) < degen_cutoff:
            ndegen_above += 1
    if respect_degenerate and (ndegen_below + ndegen_above > 1):
        if maxdim < m + ndegen_above:
            m -= ndegen_below
            ndegen_above += ndegen_below
            ndegen_below = 0
        else:
            m += ndegen_above
            ndegen_below += ndegen_above
            ndegen_above = 0
    P = P[:m] * P0
    return P, truncerr * P0, docut_lower * P0, docut_upper * P0, ndegen_below


def eigh_desc(M: np.ndarray):
    """diagHermitian: dsyev on -M, eigenvalues descending (algs_impl.h:122-135)."""
    w, U = np.linalg.eigh(-M)
    return -w, U


# -------------------------------------------------- diag_hermitian ------
class Args(dict):
    def get(self, k, default=None):
        return super().get(k, default)


def diag_hermitian(rho: ITensor, args: dict):
    """diag_hermitian (hermitian.cc:40-440), real case.
    Returns (U, D_values, spec) with spec = dict(eigs, truncerr)."""
    tags = args.get("Tags", "Link")
    i1, i2 = rho.inds[0], rho.inds[-1]
    ai = i1 if i1.plev < i2.plev else i2
    origdim = ai.dim
    cutoff = args.get("Cutoff", 0.0)
    maxdim = args.get("MaxDim", origdim)
    mindim = args.get("MinDim", 1)
    do_truncate = args.get("Truncate", ("Cutoff" in args) or ("MaxDim" in args))
    do_rel = args.get("DoRelCutoff", True)
    absc = args.get("AbsoluteCutoff", False)
    respect = args.get("RespectDegenerate", False)
    if not do_truncate and "MaxDim" in args:
        maxdim = origdim
    if not rho.is_qn:
        # matrix M[i][j] = rho(ai=i, ai'=j)
        pa = index_position(rho.inds, ai)
        M = rho.d if pa == 0 else rho.d.T
        DD, UU = eigh_desc(M)
        truncerr = 0.0
        if do_truncate:
            DD, truncerr, _, _, _ = truncate(DD, maxdim, mindim, cutoff, absc, do_rel, respect)
            UU = UU[:, :DD.size]
        newmid = Index(DD.size, tags)
        U = dense([ai, newmid], UU)
        return U, DD, {"eigs": DD.copy(), "truncerr": truncerr, "newmid": newmid}
    # QN: per-block diagonalisation, global truncation (hermitian.cc:180-426)
    pa = index_position(rho.inds, ai)
    blocks = []
    for b, _ in rho.offsets:
        M = rho.block(b)
        if pa != 0:
            M = M.T
        blocks.append((b[pa], M))
    dvecs, umats = [], []
    alleig = []
    for sec, M in blocks:
        d, U = eigh_desc(M)
        dvecs.append(d)
        umats.append(U)
        alleig.extend(d.tolist())
    probs = np.array(sorted(alleig, reverse=True))
    truncerr = 0.0
    docut_lower = docut_upper = -1.0
    ndegen = 0
    m = probs.size
    if do_truncate:
        probs, truncerr, docut_lower, docut_upper, ndegen = truncate(probs, maxdim, mindim, cutoff, absc,
                                                                     do_rel, respect)
        m = probs.size
    iq = []
    kept = []
    total_m = 0
    for (sec, M), d, U in zip(blocks, dvecs, umats):
        this_m = 0
        if do_truncate:
            while this_m < d.size and total_m < m and d[this_m] > docut_upper:
                this_m += 1
                total_m += 1
            while ndegen > 0 and this_m < d.size and total_m < m and d[this_m] > docut_lower:
                this_m += 1
                total_m += 1
                ndegen -= 1
        else:
            this_m = d.size
            total_m += this_m
        if this_m == 0:
            kept.append(None)
            continue
        kept.append((sec, d[:this_m], U[:, :this_m]))
        iq.append((ai.qn(sec), this_m))
    if not iq:
        iq.append((ai.qn(blocks[0][0]), 1))
    dind = Index(iq, tags, dir=-ai.dir)
    Uis = [ai.dag(), dind.dag()]
    Ut = qdense(Uis, QN())
    n = 0
    for k in kept:
        if k is None:
            continue
        sec, d, U = k
        Ut.block((sec, n))[...] = U
        n += 1
    return Ut, probs, {"eigs": probs.copy(), "truncerr": truncerr, "newmid": dind}


MIN_CUT = 1e-15  # global.h:61
MAX_DIM = 5000   # global.h:62


def _svd_desc(M: np.ndarray):
    """Thin SVD M = U diag(s) V^T, s descending (numpy's LAPACK gesdd)."""
    U, s, Vt = np.linalg.svd(M, full_matrices=False)
    return U, s, Vt.T


def svd(AA: ITensor, uinds, args: dict):
    """svd (decomp.cc:121-248) -> svdOrd2 -> svdImpl (svd.cc:39-423), real
    case.  uinds: the indices of AA that go to U (ITensor U(uinds) of the
    reference).  Returns (U, D, V, spec): U(uinds, dag(L)) ... as the
    reference lays them out, D over (L, R), spec = dict(eigs, truncerr)."""
    args = dict(args)
    Linds = [I for I in AA.inds if index_position(list(uinds), I) >= 0]
    Rinds = [I for I in AA.inds if index_position(list(uinds), I) < 0]
    AAc, Ucomb, Vcomb, ui, vi = AA, None, None, None, None
    if Linds:
        Ucomb, ui = combiner(Linds)
        AAc = AAc * Ucomb
    if Rinds:
        Vcomb, vi = combiner(Rinds)
        AAc = AAc * Vcomb
    do_truncate = args.get("Truncate", ("Cutoff" in args) or ("MaxDim" in args))
    cutoff = args.get("Cutoff", MIN_CUT)
    maxdim = args.get("MaxDim", MAX_DIM)
    mindim = args.get("MinDim", 1)
    do_rel = args.get("DoRelCutoff", True)
    absc = args.get("AbsoluteCutoff", False)
    respect = args.get("RespectDegenerate", False)
    ltags = args.get("LeftTags", "Link,U")
    rtags = args.get("RightTags", "Link,V")
    pu = index_position(AAc.inds, ui)
    if not AAc.is_qn:
        M = AAc.d if pu == 0 else AAc.d.T
        UU, DD, VV = _svd_desc(M)
        probs = DD ** 2
        truncerr = 0.0
        if do_truncate:
            probs, truncerr, _, _, _ = truncate(probs, maxdim, mindim, cutoff, absc, do_rel, respect)
        m = probs.size
        uL, vL = Index(m, ltags), Index(m, rtags)
        U = dense([ui, uL], UU[:, :m])
        V = dense([vi, vL], VV[:, :m])
        D = dense([uL, vL], np.diag(DD[:m]))
        spec = {"eigs": DD[:m] ** 2, "truncerr": truncerr}
    else:
        blocks = []
        for b, _ in AAc.offsets:
            M = AAc.block(b)
            if pu != 0:
                M = M.T
            blocks.append((b[pu], b[1 - pu], M))
        dvecs, umats, vmats, alleig = [], [], [], []
        for b1, b2, M in blocks:
            UU, d, VV = _svd_desc(M)
            dvecs.append(d)
            umats.append(UU)
            vmats.append(VV)
            alleig.extend(d.tolist())
        probs = np.array(sorted((x * x for x in alleig), reverse=True))
        truncerr, docut_lower, docut_upper, ndegen = 0.0, -1.0, -1.0, 1
        m = probs.size
        if do_truncate:
            probs, truncerr, docut_lower, docut_upper, ndegen = truncate(probs, maxdim, mindim, cutoff, absc,
                                                                         do_rel, respect)
            m = probs.size
        Liq, Riq, kept = [], [], []
        total_m = 0
        for (b1, b2, M), d in zip(blocks, dvecs):
            this_m = 0
            if do_truncate:
                while this_m < d.size and total_m < m and d[this_m] ** 2 > docut_upper:
                    this_m += 1
                    total_m += 1
                while ndegen > 0 and this_m < d.size and total_m < m and d[this_m] ** 2 > docut_lower:
                    this_m += 1
                    total_m += 1
                    ndegen -= 1
            else:
                this_m = d.size
                total_m += this_m
            kept.append(this_m)
            if this_m:
                Liq.append((ui.qn(b1), this_m))
                Riq.append((vi.qn(b2), this_m))
        L = Index(Liq, ltags, dir=ui.dir)
        R = Index(Riq, rtags, dir=vi.dir)
        U = qdense([ui, L.dag()], QN())
        V = qdense([vi, R.dag()], QN())
        # D carries AA's divergence (the reference's QDiagReal keeps the
        # diagonal blocks whatever the flux; U and V are flux-free)
        D = qdense([L, R], AAc.div())
        n = 0
        for (b1, b2, M), d, UU, VV, k in zip(blocks, dvecs, umats, vmats, kept):
            if not k:
                continue
            U.block((b1, n))[...] = UU[:, :k]
            V.block((b2, n))[...] = VV[:, :k]
            D.block((n, n))[...] = np.diag(d[:k])
            n += 1
        spec = {"eigs": probs.copy(), "truncerr": truncerr}
    if Ucomb is not None:
        U = Ucomb.dag() * U
    if Vcomb is not None:
        V = V * Vcomb.dag()
    return U, D, V, spec


def factor(T: ITensor, ainds, args: dict):
    """factor (decomp.cc:538-576): svd with LeftTags = Tags (default
    "Link"), then sqrt(D) absorbed into both factors, which end up sharing
    the left link index.  Returns (A, B, spec)."""
    args = dict(args)
    args["LeftTags"] = args.get("Tags", "Link")
    A, D, B, spec = svd(T, ainds, args)
    L, R = D.inds
    if D.is_qn:
        Dsq = D.copy()
        Dsq.store = np.sqrt(np.abs(Dsq.store))
    else:
        Dsq = dense(D.inds, np.sqrt(np.abs(D.d)))
    # A(.., dag(L)) * Z(L, dag(L)'), noprime; B = Dsq(L, R) * B(.., dag(R))
    if D.is_qn:
        Z = qdense([L, L.dag().prime()], QN())
        for b, _ in Z.offsets:
            if b[0] == b[1]:
                Z.block(b)[...] = Dsq.block(b)
    else:
        Z = dense([L, L.prime()], np.sqrt(np.abs(D.d)))
    A = (A * Z).noprime()
    B = Dsq * B
    return A, B, spec


def common_index(A: ITensor, B: ITensor):
    for I in A.inds:
        if index_position(B.inds, I) >= 0:
            return I
    return None


def denmat_decomp(AA: ITensor, A: ITensor, B: ITensor, fromleft: bool, args: dict):
    """denmatDecomp (decomp.h:316-418), noise = 0.  Returns (A_new, B_new, spec)."""
    args = dict(args)
    args.setdefault("Tags", "Link")
    to_orth, newoc = (A, B) if fromleft else (B, A)
    active = to_orth.inds
    cinds = [I for I in active if index_position(newoc.inds, I) < 0]
    cmb, ci = combiner(cinds, args["Tags"])
    AAc = cmb * AA
    rho = AAc * AAc.prime(1, ci).dag()
    U, D, spec = diag_hermitian(rho, args)
    cmb = cmb.dag()
    to_orth_new = cmb * U.dag()
    newoc_new = U * AAc
    if fromleft:
        return to_orth_new, newoc_new, spec
    return newoc_new, to_orth_new, spec


def svd_bond(A: ITensor, B: ITensor, phi: ITensor, fromleft: bool, args: dict):
    """MPS::svdBond (mps/mps_impl.h:21-99) for the pair (A, B) = (psi(b),
    psi(b+1)), noise 0.  With UseSVD, or Cutoff < 1e-12 and no noise
    (:50), svd(phi, A, D, B) with D (normalised under DoNormalize) pushed
    into the new centre (:52-65); otherwise denmatDecomp with the centre
    renormalised (:66-80).  Returns (A_new, B_new, spec); spec["route"] is
    "svd" or "denmat"."""
    noise = args.get("Noise", 0.0)
    cutoff = args.get("Cutoff", MIN_CUT)
    if args.get("UseSVD", False) or (noise == 0 and cutoff < 1e-12):
        if noise > 0:
            raise ValueError("svd: noise term not implemented (decomp.cc:183-184)")
        U, D, V, spec = svd(phi, A.inds, args)
        if args.get("DoNormalize", False):
            D = D * (1.0 / norm(D))
        if fromleft:
            V = V * D
        else:
            U = U * D
        spec = dict(spec, route="svd")
        return U, V, spec
    Anew, Bnew, spec = denmat_decomp(phi, A, B, fromleft, args)
    if args.get("DoNormalize", False):
        oc = Bnew if fromleft else Anew
        nrm = norm(oc)
        if nrm > 1e-16:
            oc = oc * (1.0 / nrm)
        if fromleft:
            Bnew = oc
        else:
            Anew = oc
    spec = dict(spec, route="denmat")
    return Anew, Bnew, spec


# ---------------------------------------------------------- davidson ----
def davidson(product, phi: ITensor, args: dict, size: int | None = None, rng_seed: int = 12345):
    """davidson (iterativesolvers.h:99-451) for one eigenpair.
    product(x) returns A*x as an ITensor.  Returns (energy, phi, n_products)."""
    maxiter = args.get("MaxIter", 2)
    errgoal = args.get("ErrGoal", 1e-14)
    miniter = args.get("MinIter", 1)
    Approx0 = 1e-12
    nrm = norm(phi)
    if nrm == 0.0:
        raise ValueError("zero initial vector")
    phi = phi * (1.0 / nrm)
    maxsize = size if size is not None else int(np.prod([i.dim for i in phi.inds]))
    actual_maxiter = min(maxiter, maxsize - 1)
    V = [None] * (actual_maxiter + 2)
    AV = [None] * (actual_maxiter + 2)
    M = np.full((actual_maxiter + 2, actual_maxiter + 2), np.nan)
    last_lambda = 1000.0
    V[0] = phi
    AV[0] = product(V[0])
    nprod = 1
    initEn = dot(V[0], AV[0])
    lam = initEn
    phi_t = phi
    rng = SplitMix64(rng_seed)
    for ii in range(actual_maxiter + 1):
        ni = ii + 1
        if ii == 0:
            lam = initEn
            M[0, 0] = lam
            q = AV[0] - V[0] * lam
        else:
            Mref = M[:ii + 1, :ii + 1]
            D, U = eigh_desc(-Mref)
            D = -D
            lam = D[0]
            phi_t = V[0] * U[0, 0]
            q = AV[0] * U[0, 0]
            for k in range(1, ii + 1):
                phi_t = add_into(phi_t, V[k] * U[k, 0], 1.0)
                q = add_into(q, AV[k] * U[k, 0], 1.0)
            q = add_into(q, phi_t, -lam)
            if U[0, 0] < 0:
                phi_t = phi_t * -1.0
                q = q * -1.0
        qnorm = norm(q)
        converged = (qnorm < errgoal and abs(lam - last_lambda) < errgoal) or qnorm < max(Approx0, errgoal * 1e-3)
        last_lambda = lam
        if qnorm < 1e-20 or (converged and ii >= miniter) or ii == actual_maxiter:
            break
        # one classical Gram-Schmidt pass (:305-357)
        Npass, passn, tot_pass = 1, 1, 0
        done = False
        while passn <= Npass:
            tot_pass += 1
            Vq = [dot(V[k], q) for k in range(ni)]
            for k in range(ni):
                q = add_into(q, V[k], -Vq[k])
            qnrm = norm(q)
            if qnrm < 1e-10:
                q = V[ni - 1].copy()
                _randomize(q, rng)
                qnrm = norm(q)
                passn -= 1
                if ni >= maxsize or tot_pass > Npass * 3:
                    done = True
                    break
            q = q * (1.0 / qnrm)
            passn += 1
        if done:
            break
        V[ni] = q
        AV[ni] = product(V[ni])
        nprod += 1
        for k in range(ni + 1):
            M[k, ni] = dot(V[k], AV[ni])
            M[ni, k] = M[k, ni]
    return lam, phi_t, nprod


def _randomize(T: ITensor, rng: SplitMix64):
    if T.is_qn:
        T.store = rng.normals(T.store.size)
    else:
        T.d = rng.normals(T.d.size).reshape(T.d.shape, order="F")


# ---------------------------------------------------- LocalOp / MPO -----
def local_op_product(phi: ITensor, L, W1, W2, R) -> ITensor:
    """LocalOp::product, two centre sites (mps/localop.h:324-365)."""
    if L is None:
        x = phi
        if R is not None:
            x = x * R
        x = x * W2
        x = x * W1
    else:
        x = phi * L
        x = x * W1
        x = x * W2
        if R is not None:
            x = x * R
    return x.noprime()


class LocalMPO:
    """Environments PH[0..N+1] (mps/localmpo.h:46-264, 443-618)."""

    def __init__(self, H):
        self.H = H
        self.N = len(H)
        self.PH = [None] * (self.N + 2)
        self.LHlim = 0
        self.RHlim = self.N + 1
        self.b = None
        self.env_flops = 0.0

    def _env_step(self, E, A, W):
        x = A if E is None else E * A
        x = x * W
        return x * A.prime().dag()

    def position(self, b: int, psi: MPS):
        while self.LHlim < b - 1:
            ll = self.LHlim
            self.PH[ll + 1] = self._env_step(self.PH[ll], psi(ll + 1), self.H[ll])
            self.LHlim = ll + 1
        while self.RHlim > b + 2:
            rl = self.RHlim
            self.PH[rl - 1] = self._env_step(self.PH[rl], psi(rl - 1), self.H[rl - 2])
            self.RHlim = rl - 1
        self.LHlim = b - 1
        self.RHlim = b + 2
        self.b = b

    def L(self):
        return self.PH[self.b - 1]

    def R(self):
        return self.PH[self.b + 2]

    def product(self, phi: ITensor) -> ITensor:
        return local_op_product(phi, self.L(), self.H[self.b - 1], self.H[self.b], self.R())


def sweepnext(b: int, ha: int, N: int):
    inc = 1 if ha == 1 else -1
    b += inc
    if b == (N if ha == 1 else 0):
        b -= inc
        ha += 1
    return b, ha


def dmrg(H, psi: MPS, sweeps, quiet=True, record=None):
    """Two-site DMRG (mps/dmrg.h:336-467) with the density-matrix bond step.
    sweeps: list of dicts with MaxDim, MinDim, Cutoff, Niter.  record(dict) is
    called after every bond step.  Returns (energy, psi)."""
    psi = psi.copy()
    N = psi.N
    PH = LocalMPO(H)
    energy = float("nan")
    for sw, s in enumerate(sweeps, start=1):
        args = {"Cutoff": s.get("Cutoff", 1e-8), "MinDim": s.get("MinDim", 1), "MaxDim": s["MaxDim"],
                "MaxIter": s.get("Niter", 2), "RespectDegenerate": True, "Tags": "Link",
                "DoNormalize": True, "Noise": s.get("Noise", 0.0)}
        if args["Noise"] > 0:
            raise NotImplementedError("oracle dmrg: the noise term is not restated")
        b, ha = 1, 1
        while ha <= 2:
            PH.position(b, psi)
            phi = psi(b) * psi(b + 1)
            energy, phi, nprod = davidson(PH.product, phi, args)
            fromleft = ha == 1
            Anew, Bnew, spec = svd_bond(psi(b), psi(b + 1), phi, fromleft, args)
            psi.set(b, Anew)
            psi.set(b + 1, Bnew)
            if record is not None:
                record({"sweep": sw, "b": b, "ha": ha, "energy": energy, "truncerr": spec["truncerr"],
                        "dim": common_index(Anew, Bnew).dim, "nprod": nprod,
                        "route": spec["route"]})
            b, ha = sweepnext(b, ha, N)
    return energy, psi


def mps_norm(psi: MPS) -> float:
    E = None
    for j in range(1, psi.N + 1):
        A = psi(j)
        E = A * A.dag().prime(1, *[i for i in A.inds if "Link" in i.tags]) if E is None else \
            (E * A) * A.dag().prime(1, *[i for i in A.inds if "Link" in i.tags])
    return math.sqrt(abs(E.elt()))
