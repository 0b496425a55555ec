"""CPU: pin the oracle's DMRG path (truncate, davidson, diag_hermitian,
denmatDecomp, LocalMPO, sweeps) against the reference's closed forms and
known answers, exact diagonalisation and reconstruction identities."""
import json
import math
import os

import numpy as np
import pytest

from oracle.dmrg import (LocalMPO, SplitMix64, davidson, denmat_decomp, diag_hermitian, dmrg, heisenberg_mpo,
                         product_mps, random_mps, spin_half_sites, tfim_mpo, truncate)
from oracle.itensor import QN, Index, contract, dense, norm, qdense
from oracle import dmrg as OD

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "heisenberg_ed.json")))


def _truncate_p():
    """The spectrum of decomp_test.cc:59-76 (quickran replaced by our RNG)."""
    rng = SplitMix64(7)
    p = np.array([math.exp(-2.0 * n) * (1 + 0.1 * rng.uniform()) for n in range(20)])
    p[1] = p[2]
    p[3] = p[5]
    p[4] = p[5]
    return p / p.sum()


def test_truncate_reference_cases():
    """decomp_test.cc:78-137, cases 0-4."""
    p = _truncate_p()
    assert truncate(p, 40, 1, 0.0)[0].size == 20            # case 0
    assert truncate(p, 10, 1, 0.0)[0].size == 10            # case 1
    assert truncate(p, 8, 1, 0.0)[0].size == 8              # case 2
    assert truncate(p, 40, 10, 0.01)[0].size == 10          # case 3
    P, te, *_ = truncate(p, 40, 1, 1e-5)                    # case 4
    m = P.size
    assert abs(te - p[m:].sum()) < 1e-5 and te < 1e-5


def test_truncate_respects_degeneracy():
    p = np.array([0.5, 0.2, 0.1, 0.1, 0.1])
    # maxdim falls inside the degenerate triple: drop it entirely
    assert truncate(p, 3, 1, 0.0, respect_degenerate=True)[0].size == 2
    # cutoff lands inside it but maxdim allows it: keep all three
    assert truncate(p, 10, 1, 0.25, respect_degenerate=True)[0].size == 5
    assert truncate(p, 10, 1, 0.25, respect_degenerate=False)[0].size == 3


def _neel_bond_problem(qn):
    """iterativesolvers_test.cc:43-65: N=4 Heisenberg, Neel product state,
    Davidson on bond (2,3) with MaxIter=9 -> -(1/4 + 1/sqrt 2)."""
    N = 4
    sites = spin_half_sites(N, qn)
    H = heisenberg_mpo(sites)
    psi = product_mps(sites, [0 if (i % 2 == 1) else 1 for i in range(1, N + 1)])
    PH = LocalMPO(H)
    PH.position(2, psi)
    phi = psi(2) * psi(3)
    return davidson(PH.product, phi, {"MaxIter": 9})[0]


@pytest.mark.parametrize("qn", [False, True])
def test_davidson_closed_form(qn):
    assert abs(_neel_bond_problem(qn) - (-(0.25 + 1 / math.sqrt(2)))) < 1e-10


def test_davidson_custom_linear_map():
    """iterativesolvers_test.cc:168-184: symmetric map, residual ~ 0."""
    rng = np.random.default_rng(3)
    a1, a2, a3 = Index(3, "a1"), Index(4, "a2"), Index(5, "a3")
    M = rng.standard_normal((60, 60))
    M = 0.5 * (M + M.T)
    A = dense([a1.prime(), a2.prime(), a3.prime(), a1, a2, a3], M.reshape((3, 4, 5, 3, 4, 5), order="F"))
    x = dense([a1, a2, a3], rng.standard_normal((3, 4, 5)))
    lam, v, _ = davidson(lambda t: (A * t).noprime(), x, {"MaxIter": 40, "ErrGoal": 1e-14})
    r = (A * v).noprime() - v * lam
    assert norm(r) / norm(v) < 1e-6
    assert abs(lam - np.linalg.eigvalsh(M)[0]) < 1e-9


def test_diag_hermitian_reconstruction_dense_and_qn():
    rng = np.random.default_rng(4)
    i = Index(12, "i")
    X = rng.standard_normal((12, 12))
    rho = dense([i, i.prime()], X @ X.T)
    U, D, spec = diag_hermitian(rho, {})
    R = U.d @ np.diag(D) @ U.d.T
    assert np.max(np.abs(R - rho.d)) < 1e-12 * np.max(np.abs(rho.d))
    assert np.all(np.diff(D) <= 0)
    # QN block-diagonal density matrix
    j = Index([(QN(0), 3), (QN(1), 4), (QN(-1), 2)], "j")
    a = qdense([j, Index([(QN(0), 2), (QN(-1), 3), (QN(1), 2)], "k").dag()], QN(0), rng=rng)
    rho = a * a.dag().prime(1, j)
    U, D, spec = diag_hermitian(rho, {})
    Ud = U.to_dense_array()
    # U^T rho U is diagonal with the kept eigenvalues (in block order)
    rd = rho.to_dense_array()
    proj = Ud.T @ rd @ Ud
    assert np.allclose(proj, np.diag(np.diag(proj)), atol=1e-12)
    assert np.allclose(sorted(np.diag(proj), reverse=True), D, atol=1e-12)


def test_qn_denmat_decomp_reconstructs():
    """decomp_test.cc:585-625: QN denmatDecomp reconstructs AA to 1e-11 with
    non-negative kept eigenvalues."""
    rng = np.random.default_rng(5)
    S1 = Index([(QN(1), 1), (QN(-1), 1)], "S1")
    S2 = Index([(QN(1), 1), (QN(-1), 1)], "S2")
    L1 = Index([(QN(2), 3), (QN(1), 4), (QN(0), 8), (QN(-1), 4), (QN(-2), 2)], "L1")
    L2 = Index([(QN(2), 4), (QN(1), 6), (QN(0), 10), (QN(-1), 4), (QN(-2), 3)], "L2")
    L3 = Index([(QN(2), 2), (QN(0), 4), (QN(-2), 2)], "L3")
    A1 = qdense([L1, S1, L2], QN(), rng=rng)
    A2 = qdense([L2.dag(), S2, L3], QN(), rng=rng)
    AA = A1 * A2
    AA = AA * (-1.0 / norm(AA))
    nA1, nA2, spec = denmat_decomp(AA, A1, A2, True, {"Truncate": True})
    AAres = nA1 * nA2
    diff = AA - AAres
    assert norm(diff) < 1e-11
    assert np.all(spec["eigs"] >= 0)


def test_four_site_dmrg_energy():
    """Exact 4-site energy -1.6160254038 (iterativesolvers_test.cc:45)."""
    sites = spin_half_sites(4, False)
    H = heisenberg_mpo(sites)
    psi0 = random_mps(sites, 4, seed=1)
    sw = [{"MaxDim": 10, "Cutoff": 1e-12}] * 4
    E, _ = dmrg(H, psi0, sw)
    assert abs(E - (-1.6160254038)) < 1e-9
    assert abs(E - GOLD["energies"]["4"]) < 1e-10


def test_four_site_qn_dmrg_energy():
    sites = spin_half_sites(4, True)
    H = heisenberg_mpo(sites)
    psi0 = product_mps(sites, [0, 1, 0, 1])
    E, _ = dmrg(H, psi0, [{"MaxDim": 10, "Cutoff": 1e-12}] * 4)
    assert abs(E - GOLD["energies"]["4"]) < 1e-10


def test_heisenberg_n20_config1_vs_ed():
    """Config 1: N=20, random MPS chi=10 (seed), sweeps maxdim
    [10,20,100,100,200], cutoff 1e-10, niter 2; vs exact diagonalisation."""
    sites = spin_half_sites(20, False)
    H = heisenberg_mpo(sites)
    psi0 = random_mps(sites, 10, seed=20261202 + 1)
    sw = [{"MaxDim": m, "Cutoff": 1e-10, "Niter": 2} for m in (10, 20, 100, 100, 200)]
    trace = []
    E, psi = dmrg(H, psi0, sw, record=trace.append)
    assert len(trace) == 5 * 2 * 19
    assert abs(E - GOLD["energies"]["20"]) < 1e-8
    assert all(t["nprod"] == 3 for t in trace if t["sweep"] > 1)


def test_tfim_critical_n32_closed_form():
    """mpo_test.cc:541-570: critical TFIM, N=32, sweeps maxdim 10,20,30,
    cutoff 1e-12; exact E = (1 - 1/sin(pi/(2(2N+1))))/4."""
    N = 32
    sites = spin_half_sites(N, False)
    H = tfim_mpo(sites, 0.5)
    psi0 = product_mps(sites, [0] * N)
    sw = [{"MaxDim": m, "Cutoff": 1e-12} for m in (10, 20, 30, 30, 30)]
    E, _ = dmrg(H, psi0, sw)
    exact = (1.0 - 1.0 / math.sin(math.pi / (2 * (2 * N + 1)))) / 4
    assert abs((E - exact) / exact) < 1e-5  # CHECK_CLOSE: |x| < 1e-5


# ---- svd / factor (decomp.cc:121-248, 538-576; svd.cc:39-423) ----------
def _degenerate_case():
    """decomp_test.cc:881-901: blocks QN(0..3) of dim 5 holding
    diag(2,1,1,1,0.5), normalised."""
    vals = [2.0, 1.0, 1.0, 1.0, 0.5]
    iq = Index([(QN(b), 5) for b in range(4)], "i")
    A = qdense([iq, iq.prime().dag()], QN())
    for b, _ in A.offsets:
        if b[0] == b[1]:
            A.block(b)[...] = np.diag(vals)
    A.store = A.store / np.linalg.norm(A.store)
    return iq, A


@pytest.mark.parametrize("args,qn_dim,dense_dim", [
    ({"MaxDim": 9, "Cutoff": 0.0}, 9, 9),                               # decomp_test.cc:905-940
    ({"MaxDim": 9, "Cutoff": 0.0, "RespectDegenerate": True}, 4, 4),
    ({"MinDim": 9, "Cutoff": 1.0}, 9, 9),                               # :945-985
    ({"MinDim": 9, "Cutoff": 1.0, "RespectDegenerate": True}, 16, 16),
    ({"Cutoff": 0.25}, 10, 10),                                          # :990-1025
    ({"Cutoff": 0.25, "RespectDegenerate": True}, 16, 16),
])
def test_oracle_svd_degenerate_truncation_known_answers(args, qn_dim, dense_dim):
    from oracle.itensor import remove_qns

    iq, A = _degenerate_case()
    U, D, V, spec = OD.svd(A, [iq], args)
    assert D.inds[0].dim == D.inds[1].dim == qn_dim
    Ad = remove_qns(A)
    U, D, V, spec = OD.svd(Ad, [Ad.inds[0]], args)
    assert D.inds[0].dim == D.inds[1].dim == dense_dim


def test_oracle_svd_special_cases():
    """decomp_test.cc:756-800: all-zero blocks and a dimension-one index
    keep one singular value with Cutoff 1."""
    iq = Index([(QN(b), 2) for b in range(4)], "i")
    A = qdense([iq, iq.prime().dag()], QN())
    U, D, V, spec = OD.svd(A, [iq], {"Cutoff": 1.0})
    assert D.inds[0].dim == 1 and D.inds[1].dim == 1
    i1 = Index([(QN(), 1)], "i")
    A = qdense([i1, i1.prime().dag()], QN())
    A.store = np.array([0.7])
    U, D, V, spec = OD.svd(A, [i1], {"Cutoff": 1.0})
    assert D.inds[0].dim == 1


def test_oracle_svd_reconstruction_and_factor():
    """decomp_test.cc:140-225: U*D*V and A*B reconstruct T (dense cases 1-3
    and the QN regression tests); U, V have orthonormal columns."""
    from oracle.itensor import norm

    rng = np.random.default_rng(4)
    i, j, k = Index(3, "i"), Index(4, "j"), Index(5, "k")
    T = dense([i, j, k], rng.standard_normal((3, 4, 5)))
    for uinds in ([i, j], [i, k]):
        U, D, V, spec = OD.svd(T, uinds, {})
        assert norm(U * D * V - T) < 1e-12
        A, B, _ = OD.factor(T, uinds, {})
        assert norm(A * B - T) < 1e-12
    T3 = dense([i, k, i.prime()], rng.standard_normal((3, 5, 3)))
    U, D, V, _ = OD.svd(T3, [i, i.prime()], {})
    assert norm(U * D * V - T3) < 1e-12
    u = Index([(QN(2), 1), (QN(0), 1), (QN(-2), 1)], "u")
    v = Index([(QN(2), 1), (QN(0), 1), (QN(-2), 1)], "v")
    S = qdense([u, v], QN())
    S.store = rng.standard_normal(S.store.size)
    U, D, V, _ = OD.svd(S, [u], {})
    assert norm(U * D * V - S) < 1e-12
    # singular values equal numpy's on the dense expansion
    from oracle.itensor import remove_qns

    M = rng.standard_normal((6, 9))
    a, b = Index(6, "a"), Index(9, "b")
    U, D, V, spec = OD.svd(dense([a, b], M), [a], {})
    assert np.allclose(np.sqrt(spec["eigs"]), np.linalg.svd(M, compute_uv=False), atol=1e-13)


# ---- MPS::svdBond routes (mps/mps_impl.h:21-99) ------------------------
def _bond_problem(qn: bool, seed: int = 5):
    N = 6
    sites = OD.spin_half_sites(N, qn)
    psi = OD.random_mps(sites, 4, seed=seed) if not qn else OD.product_mps(sites, [i % 2 for i in range(N)])
    if qn:
        # grow the links of the product state with two DMRG sweeps
        Hm = heisenberg_mpo(sites)
        _, psi = dmrg(Hm, psi, [{"MaxDim": 8, "Cutoff": 1e-10}, {"MaxDim": 8, "Cutoff": 1e-10}])
    b = 3
    return psi(b), psi(b + 1), psi(b) * psi(b + 1)


@pytest.mark.parametrize("qn", [False, True])
@pytest.mark.parametrize("args,route", [
    ({"Cutoff": 1e-13, "MaxDim": 100}, "svd"),
    ({"Cutoff": 1e-10, "MaxDim": 100}, "denmat"),
    ({"Cutoff": 1e-10, "MaxDim": 100, "UseSVD": True}, "svd"),
    ({"MaxDim": 100}, "svd"),  # Cutoff defaults to MIN_CUT
])
def test_svd_bond_route_and_identities(qn, args, route):
    """The route follows mps_impl.h:50 (UseSVD, or noise 0 and cutoff below
    1e-12).  Both routes reconstruct phi exactly when nothing is truncated,
    leave the side being orthogonalised an isometry, and (DoNormalize) give a
    unit-norm centre."""
    A, B, phi = _bond_problem(qn)
    a = dict(args, RespectDegenerate=True, DoNormalize=True)
    for fromleft in (True, False):
        An, Bn, spec = OD.svd_bond(A, B, phi, fromleft, a)
        assert spec["route"] == route
        AB = An * Bn
        # phi has unit norm up to the ordering of the product: compare the
        # normalised product with the normalised phi
        from oracle.itensor import add_into
        diff = add_into(AB * (1.0 / norm(AB)), phi * (1.0 / norm(phi)), -1.0)
        assert norm(diff) < 1e-12
        ortho, centre = (An, Bn) if fromleft else (Bn, An)
        assert abs(norm(centre) - 1.0) < 1e-13
        link = OD.common_index(An, Bn)
        G = ortho * ortho.prime(1, link).dag()
        from oracle.itensor import remove_qns
        g = remove_qns(G).d if G.is_qn else G.d
        assert np.allclose(g, np.eye(link.dim), atol=1e-12)


def test_svd_bond_route_keeps_small_singular_values():
    """Cutoff below 1e-12 with no noise takes the svd route, which resolves
    singular values below sqrt(machine epsilon): a matrix with spectrum
    10^-k (k = 0..9) keeps all ten, and the truncation error of dropping the
    smallest at MaxDim 9 is 1e-18 relative (the density-matrix eigenvalues
    of those values sit at the rounding level of the largest one)."""
    i, j = Index(10, "i"), Index(12, "j")
    rng = np.random.default_rng(0)
    Q1, _ = np.linalg.qr(rng.standard_normal((10, 10)))
    Q2, _ = np.linalg.qr(rng.standard_normal((12, 10)))
    s = 10.0 ** -np.arange(10)
    M = (Q1 * s) @ Q2.T
    phi = dense([i, j], M)
    l = Index(1, "Link,l=1")
    A = dense([i, l], np.ones((10, 1)))
    B = dense([l, j], np.ones((1, 12)))
    An, Bn, spec = OD.svd_bond(A, B, phi, True, {"Cutoff": 1e-17, "MaxDim": 9, "RespectDegenerate": True})
    assert spec["route"] == "svd"
    assert OD.common_index(An, Bn).dim == 9
    assert abs(spec["truncerr"] - 1e-18 / np.sum(s ** 2)) < 1e-24
    assert np.allclose(np.sqrt(spec["eigs"]), s[:9], rtol=1e-8, atol=0)


def test_qn_dmrg_svd_route_vs_ed():
    """N=10 QN Heisenberg at cutoff 1e-13 and no noise: every bond step takes
    the svd route and the energy reaches exact diagonalisation."""
    N = 10
    sites = OD.spin_half_sites(N, True)
    Hm = heisenberg_mpo(sites)
    psi0 = OD.product_mps(sites, [i % 2 for i in range(N)])
    sw = [{"MaxDim": m, "Cutoff": 1e-13, "Niter": 2} for m in (10, 20, 40, 40, 40)]
    trace = []
    E, _ = dmrg(Hm, psi0, sw, record=trace.append)
    assert {t["route"] for t in trace} == {"svd"}
    assert abs(E - GOLD["energies"]["10"]) < 1e-10


@pytest.mark.parametrize("flux", [1, -2])
def test_oracle_svd_nonzero_flux(flux):
    """A QN tensor whose divergence is not zero (the orthogonality centre of
    a particle-number MPS): U and V come out flux-free, D carries the
    divergence on its diagonal blocks (the reference's QDiagReal, svd.cc:
    321-410), U*D*V reconstructs the tensor and the singular values equal
    numpy's on the dense expansion."""
    from oracle.itensor import remove_qns

    rng = np.random.default_rng(40 + flux)
    L = Index([(QN(2), 3), (QN(1), 4), (QN(0), 5), (QN(-1), 4), (QN(-2), 2)], "L")
    s = Index([(QN(1), 1), (QN(-1), 1)], "s")
    R = Index([(QN(2), 2), (QN(1), 3), (QN(0), 6), (QN(-1), 3), (QN(-2), 3)], "R")
    T = qdense([L, s, R.dag()], QN(flux), rng=rng)
    U, D, V, spec = OD.svd(T, [L, s], {})
    assert U.div() == QN() and V.div() == QN()
    assert D.div() == QN(flux)
    assert all(b[0] == b[1] for b, _ in D.offsets)
    assert norm(U * D * V - T) < 1e-12 * norm(T)
    Td = remove_qns(T).d.reshape(L.dim * s.dim, R.dim, order="F")
    sv = np.linalg.svd(Td, compute_uv=False)
    assert np.allclose(np.sort(spec["eigs"])[::-1], (sv ** 2)[: len(spec["eigs"])], atol=1e-12)
