"""GPU parity of the ITensor-compatible C++ host layer (which calls libitx
through the C ABI) against the oracle: contraction layouts and QN block
patterns bit-exact, values to 1e-12, Davidson / denmatDecomp / DMRG
energies against closed forms, exact diagonalisation and the oracle's
own trajectory."""
import json
import math
import os

import numpy as np
import pytest

from oracle import dmrg as OD
from oracle.itensor import QN, Index, add_into, combiner, contract, norm, qdense, dense
from tests.bridge import cidx, cten, oten, same_inds

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "heisenberg_ed.json")))


@pytest.fixture(scope="module")
def H(torch_dev):
    from itensor_amd import host

    return host


def _rel(a, b):
    den = max(np.max(np.abs(b)) if np.size(b) else 0.0, 1e-300)
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / den) if np.size(b) else 0.0


def _qn_pair(seed):
    rng = np.random.default_rng(seed)

    def rind(tag):
        nb = int(rng.integers(1, 5))
        return Index([(QN(int(q)), int(rng.integers(1, 6))) for q in rng.integers(-2, 3, size=nb)], tag)

    a, b, c, d = rind("a"), rind("b"), rind("c"), rind("d")
    A = qdense([a, b.dag(), c], QN(int(rng.integers(-1, 2))), rng=rng)
    B = qdense([c.dag(), d, b], QN(int(rng.integers(-1, 2))), rng=rng)
    return A, B


@pytest.mark.parametrize("seed", range(8))
def test_qn_contraction_layout_bit_exact(H, seed):
    A, B = _qn_pair(seed)
    ref = contract(A, B)
    C = cten(A) * cten(B)
    assert same_inds(C, cten(ref))
    assert [(tuple(b), o) for b, o in C.offsets] == [(b, o) for b, o in ref.offsets]
    assert _rel(C.store(), ref.store) < 1e-12


def test_dense_contraction_and_scalar(H):
    rng = np.random.default_rng(3)
    i, j, k = Index(7, "i"), Index(5, "j"), Index(9, "k")
    A = dense([i, j, k], rng.standard_normal((7, 5, 9)))
    B = dense([k, i], rng.standard_normal((9, 7)))
    ref = contract(A, B)
    C = cten(A) * cten(B)
    assert same_inds(C, cten(ref))
    assert _rel(C.store(), ref.d.ravel(order="F")) < 1e-12
    full = contract(A, A.dag())
    assert abs((cten(A) * cten(A).dag()).elt() - full.elt()) < 1e-10


@pytest.mark.parametrize("seed", range(4))
def test_qn_add_union_and_permuted(H, seed):
    rng = np.random.default_rng(50 + seed)
    i = Index([(QN(0), 2), (QN(1), 3), (QN(2), 4), (QN(1), 5), (QN(3), 6)], "i")
    l = Index([(QN(0), 3)], "l")
    a = qdense([i, l], QN(1), rng=rng)
    A = contract(a, a.dag().prime(1, i))          # block deficient (4 blocks)
    T = qdense([i.prime().dag(), i], QN(0), rng=rng)  # full, permuted index order
    ref = add_into(A.copy(), T, -0.5)
    got = cten(A) + (cten(T) * -0.5)
    assert got.nnzblocks() == ref.nnzblocks() == 7
    assert [(tuple(b), o) for b, o in got.offsets] == [(b, o) for b, o in ref.offsets]
    assert _rel(got.store(), ref.store) < 1e-14


def test_dense_add_permuted(H):
    rng = np.random.default_rng(9)
    i, j, k = Index(3, "i"), Index(4, "j"), Index(5, "k")
    A = dense([i, j, k], rng.standard_normal((3, 4, 5)))
    B = dense([k, i, j], rng.standard_normal((5, 3, 4)))
    ref = add_into(A.copy(), B, 2.0)
    got = cten(A) + cten(B) * 2.0
    assert _rel(got.store(), ref.d.ravel(order="F")) < 1e-14


def test_qn_combiner_matches_oracle(H):
    rng = np.random.default_rng(5)
    l = Index([(QN(0), 2), (QN(1), 3), (QN(-1), 1)], "l")
    s = Index([(QN(1), 1), (QN(-1), 1)], "s")
    r = Index([(QN(0), 2), (QN(1), 2)], "r")
    T = qdense([l, s, r.dag()], QN(0), rng=rng)
    Co, co = combiner([l, s])
    ref = Co * T
    Cc, cc = H.combiner([cidx(l), cidx(s)], "Link,CMB")
    assert [(list(q), m) for q, m in cc.qns] == [(list(q.q), m) for q, m in co.qns]
    got = Cc * cten(T)
    assert [(tuple(b), o) for b, o in got.offsets] == [(b, o) for b, o in ref.offsets]
    assert _rel(got.store(), ref.store) < 1e-15
    back = Cc.dag() * got
    refb = Co.dag() * ref
    assert [(tuple(b), o) for b, o in back.offsets] == [(b, o) for b, o in refb.offsets]
    assert _rel(back.store(), refb.store) < 1e-15


def test_dense_combiner_permuting(H):
    rng = np.random.default_rng(6)
    i, j, k = Index(2, "i"), Index(3, "j"), Index(4, "k")
    T = dense([i, j, k], rng.standard_normal((2, 3, 4)))
    Co, co = combiner([k, i])
    ref = Co * T
    Cc, cc = H.combiner([cidx(k), cidx(i)], "Link,CMB")
    got = Cc * cten(T)
    assert [x.dim for x in got.inds] == [x.dim for x in ref.inds]
    assert _rel(got.store(), ref.d.ravel(order="F")) < 1e-15


def test_diag_hermitian_qn_sectors_exact(H):
    rng = np.random.default_rng(4)
    j = Index([(QN(0), 3), (QN(1), 4), (QN(-1), 2)], "j")
    a = qdense([j, Index([(QN(0), 2), (QN(-1), 3), (QN(1), 2)], "k").dag()], QN(0), rng=rng)
    rho = a * a.dag().prime(1, j)
    args = {"MaxDim": 6, "Cutoff": 1e-12, "RespectDegenerate": True}
    Uo, Do, so = OD.diag_hermitian(rho, dict(args))
    U, D, eigs, te = H.diag_hermitian(cten(rho), args)
    assert np.allclose(eigs, so["eigs"], rtol=1e-12, atol=1e-14)
    newo = Uo.inds[1]
    newc = U.inds[1]
    assert [(list(q), m) for q, m in newc.qns] == [(list(q.q), m) for q, m in newo.qns]
    assert [(tuple(b), o) for b, o in U.offsets] == [(b, o) for b, o in Uo.offsets]
    # columns agree up to sign
    Ug, Ur = oten(U).to_dense_array(), Uo.to_dense_array()
    for c in range(Ur.shape[1]):
        assert min(np.max(np.abs(Ug[:, c] - Ur[:, c])), np.max(np.abs(Ug[:, c] + Ur[:, c]))) < 1e-10


def test_qn_denmat_decomp_reconstructs(H):
    """decomp_test.cc:585-625 on the GPU path."""
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
    cA, cB, eigs, te = H.denmat_decomp(cten(AA), cten(A1), cten(A2), True, {"Truncate": True})
    res = cA * cB
    diff = cten(AA) - res
    assert H.norm(diff) < 1e-11
    assert np.all(np.asarray(eigs) >= 0)


@pytest.mark.parametrize("qn", [False, True])
def test_davidson_closed_form_gpu(H, qn):
    """iterativesolvers_test.cc:43-65 through the C++ LocalMPO + davidson."""
    sites = OD.spin_half_sites(4, qn)
    Hm = OD.heisenberg_mpo(sites)
    psi = OD.product_mps(sites, [0, 1, 0, 1])
    e, phi, nprod = H.davidson_bond([cten(w) for w in Hm], [cten(a) for a in psi.A], 2, {"MaxIter": 9})
    assert abs(e - (-(0.25 + 1 / math.sqrt(2)))) < 1e-10


@pytest.mark.parametrize("fuse", [False, True])
@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_local_product_matches_oracle_dense(H, fuse, mode):
    """LocalOp::product (four-stage chain, or with the MPO pair fused; every
    execution mode) vs the oracle's four-stage chain, at both chain ends and
    in the bulk: same result index order, values 1e-12."""
    sites = OD.spin_half_sites(8, False)
    Hm = OD.heisenberg_mpo(sites)
    psi = OD.random_mps(sites, 16, seed=11)
    PH = OD.LocalMPO(Hm)
    for b in (1, 4, 7):
        PH.position(b, psi)
        phi = psi(b) * psi(b + 1)
        ref = PH.product(phi)
        got = H.local_product([cten(w) for w in Hm], [cten(a) for a in psi.A], b, cten(phi), 0, fuse, mode)
        assert same_inds(got, cten(ref))
        assert _rel(got.store(), ref.d.ravel(order="F")) < 1e-12


@pytest.mark.parametrize("mode", [0, 2, 3])
def test_local_product_qn_matches_oracle(H, mode):
    """QN LocalOp::product at both chain ends and in the bulk (eager, graph
    and per-block modes): block list bit-exact, values 1e-12."""
    N = 10
    sites = OD.spin_half_sites(N, True)
    Hm = OD.heisenberg_mpo(sites)
    psi0 = OD.product_mps(sites, [i % 2 for i in range(N)])
    _, psi = OD.dmrg(Hm, psi0, [{"MaxDim": 20, "Cutoff": 1e-12}] * 2)
    PH = OD.LocalMPO(Hm)
    for b in (1, 5, 9):
        PH.position(b, psi)
        phi = psi(b) * psi(b + 1)
        ref = PH.product(phi)
        for fuse in (False, True):
            got = H.local_product([cten(w) for w in Hm], [cten(a) for a in psi.A], b, cten(phi), 0, fuse, mode)
            assert same_inds(got, cten(ref))
            assert [(tuple(x), o) for x, o in got.offsets] == [(x, o) for x, o in ref.offsets]
            assert _rel(got.store(), ref.store) < 1e-12


def test_dmrg_config1_trajectory_matches_oracle(H):
    """Config 1 (N=20 Heisenberg, random MPS chi=10, maxdim
    [10,20,100,100,200], cutoff 1e-10, niter 2): GPU FP64 vs the oracle on
    the same initial state: per-bond energies to 1e-10, the same kept
    dimensions, final energy vs exact diagonalisation."""
    N = 20
    sites = OD.spin_half_sites(N, False)
    Hm = OD.heisenberg_mpo(sites)
    psi0 = OD.random_mps(sites, 10, seed=20261202 + 1)
    sw = [{"MaxDim": m, "Cutoff": 1e-10, "Niter": 2} for m in (10, 20, 100, 100, 200)]
    trace = []
    Eo, _ = OD.dmrg(Hm, psi0, sw, record=trace.append)
    Eg, recs, tens, secs = H.dmrg([cten(w) for w in Hm], [cten(a) for a in psi0.A], sw, {})
    assert len(recs) == len(trace) == 5 * 2 * (N - 1)
    assert [r["dim"] for r in recs] == [t["dim"] for t in trace]
    assert max(abs(r["energy"] - t["energy"]) for r, t in zip(recs, trace)) < 1e-10
    assert abs(Eg - Eo) < 1e-10
    assert abs(Eg - GOLD["energies"]["20"]) < 1e-8
    print(f"config-1 sec/sweep (GPU FP64): {secs}")


@pytest.mark.parametrize("mode", [0, 2])
def test_qn_dmrg_matches_oracle_and_ed(H, mode):
    """QN DMRG N=12 (eager products, and the planned chain replayed as a HIP
    graph): per-bond energies and kept dims as the oracle, energy vs ED."""
    N = 12
    sites = OD.spin_half_sites(N, True)
    Hm = OD.heisenberg_mpo(sites)
    psi0 = OD.product_mps(sites, [i % 2 for i in range(N)])
    sw = [{"MaxDim": m, "Cutoff": 1e-12, "Niter": 2} for m in (10, 20, 40, 60, 60)]
    trace = []
    Eo, _ = OD.dmrg(Hm, psi0, sw, record=trace.append)
    Eg, recs, tens, secs = H.dmrg([cten(w) for w in Hm], [cten(a) for a in psi0.A], sw, {"HpsiMode": mode})
    assert [r["dim"] for r in recs] == [t["dim"] for t in trace]
    assert max(abs(r["energy"] - t["energy"]) for r, t in zip(recs, trace)) < 1e-10
    assert abs(Eg - GOLD["energies"]["12"]) < 1e-9


def test_tfim_n32_closed_form_gpu(H):
    N = 32
    sites = OD.spin_half_sites(N, False)
    Hm = OD.tfim_mpo(sites, 0.5)
    psi0 = OD.product_mps(sites, [0] * N)
    sw = [{"MaxDim": m, "Cutoff": 1e-12} for m in (10, 20, 30, 30, 30)]
    E, recs, tens, secs = H.dmrg([cten(w) for w in Hm], [cten(a) for a in psi0.A], sw, {})
    exact = (1.0 - 1.0 / math.sin(math.pi / (2 * (2 * N + 1)))) / 4
    assert abs((E - exact) / exact) < 1e-5


# ---- svd / factor (decomp.cc:121-248, 538-576; svd.cc:39-436) ---------
def _recon(H, U, D, V):
    return H.remove_qns(U * D * V).store() if U.is_qn else (U * D * V).store()


def test_svd_dense_cases_match_oracle(H):
    """decomp_test.cc:140-181 (cases 1-3): U*D*V reconstructs T; U holds the
    requested indices, V the rest; singular values as the oracle's."""
    rng = np.random.default_rng(21)
    i, j, k = Index(3, "i"), Index(4, "j"), Index(5, "k")
    T = dense([i, j, k], rng.standard_normal((3, 4, 5)))
    T3 = dense([i, k, i.prime()], rng.standard_normal((3, 5, 3)))
    for TT, uinds in ((T, [i, j]), (T, [i, k]), (T3, [i, i.prime()])):
        Uo, Do, Vo, so = OD.svd(TT, uinds, {})
        U, D, V, eigs, terr = H.svd_spec(cten(TT), [cidx(x) for x in uinds], {})
        ref = cten(TT)
        assert H.norm(U * D * V - ref) < 1e-12 * H.norm(ref)
        assert np.allclose(eigs, so["eigs"], rtol=0, atol=1e-12 * so["eigs"][0])
        assert {x.id for x in U.inds} >= {x.id for x in uinds}
        A, B, _, _ = H.factor(cten(TT), [cidx(x) for x in uinds], {})
        assert H.norm(A * B - ref) < 1e-12 * H.norm(ref)


def test_svd_qn_regressions_and_structure(H):
    """decomp_test.cc:184-225 (QN regression tests) plus a random flux-0
    block tensor: reconstruction, and L/R sector lists, arrows and the U,
    V, D block lists identical to the oracle's."""
    from oracle.itensor import QN as OQN

    rng = np.random.default_rng(8)
    u = Index([(OQN(2), 1), (OQN(0), 1), (OQN(-2), 1)], "u")
    v = Index([(OQN(2), 1), (OQN(0), 1), (OQN(-2), 1)], "v")
    S = qdense([u, v], OQN())
    S.store = rng.standard_normal(S.store.size)
    a = Index([(OQN(1), 3), (OQN(-1), 4), (OQN(3), 2)], "a")
    b = Index([(OQN(0), 2), (OQN(2), 3), (OQN(-2), 3)], "b")
    c = Index([(OQN(1), 2), (OQN(-1), 2)], "c")
    T = qdense([a, b.dag(), c], OQN(0))
    T.store = rng.standard_normal(T.store.size)
    assert T.nnzblocks() > 4
    for TT, uinds, args in ((S, [u], {}), (T, [a, c], {}), (T, [a, c], {"MaxDim": 5, "Cutoff": 1e-3}),
                            (T, [b.dag()], {"Cutoff": 1e-2})):
        Uo, Do, Vo, so = OD.svd(TT, uinds, args)
        U, D, V, eigs, terr = H.svd_spec(cten(TT), [cidx(x) for x in uinds], args)
        for got, ref in ((U, Uo), (D, Do), (V, Vo)):
            assert [(x.dim, x.dir, x.plev) for x in got.inds] == [(x.dim, x.dir, x.plev) for x in ref.inds]
            assert [(tuple(x), o) for x, o in got.offsets] == [(x, o) for x, o in ref.offsets]
        assert D.inds[0].qns == cidx(Do.inds[0]).qns and D.inds[1].qns == cidx(Do.inds[1]).qns
        assert np.allclose(eigs, so["eigs"], rtol=0, atol=1e-12 * so["eigs"][0])
        assert abs(terr - so["truncerr"]) < 1e-12
        if "MaxDim" not in args and "Cutoff" not in args:
            ref = cten(TT)
            assert H.norm(U * D * V - ref) < 1e-12 * H.norm(ref)
        A, B, _, _ = H.factor(cten(TT), [cidx(x) for x in uinds], args)
        Ao, Bo, _ = OD.factor(TT, uinds, args)
        AB = (A * B).store()
        ABo = cten(Ao * Bo).store()
        assert np.max(np.abs(AB - ABo)) < 1e-12 * max(1.0, np.max(np.abs(ABo)))


@pytest.mark.parametrize("args,dim", [
    ({"MaxDim": 9, "Cutoff": 0.0}, 9), ({"MaxDim": 9, "Cutoff": 0.0, "RespectDegenerate": True}, 4),
    ({"MinDim": 9, "Cutoff": 1.0}, 9), ({"MinDim": 9, "Cutoff": 1.0, "RespectDegenerate": True}, 16),
    ({"Cutoff": 0.25}, 10), ({"Cutoff": 0.25, "RespectDegenerate": True}, 16),
])
def test_svd_degenerate_truncation_known_answers(H, args, dim):
    """decomp_test.cc:881-1030: kept dimension of svd with and without QNs."""
    from oracle.itensor import QN as OQN, remove_qns

    vals = [2.0, 1.0, 1.0, 1.0, 0.5]
    iq = Index([(OQN(b), 5) for b in range(4)], "i")
    A = qdense([iq, iq.prime().dag()], OQN())
    for bb, _ in A.offsets:
        if bb[0] == bb[1]:
            A.block(bb)[...] = np.diag(vals)
    A.store = A.store / np.linalg.norm(A.store)
    for TT, ind in ((A, iq), (remove_qns(A), remove_qns(A).inds[0])):
        U, D, V = H.svd(cten(TT), [cidx(ind)], args)
        assert D.inds[0].dim == dim and D.inds[1].dim == dim


# ---- AutoMPO / Electron (config 5's model) on the device ------------------
@pytest.mark.parametrize("Nx,Ny,U", [(3, 2, 4.0), (2, 3, 4.0), (3, 2, 0.0)])
def test_hubbard_dmrg_matches_ed(H, Nx, Ny, U):
    """(Nf, Sz)-conserving DMRG of the Hubbard model on small clusters (the
    MPO from the host AutoMPO, YPeriodic where Ny > 2) against exact
    diagonalisation of the Jordan-Wigner Hamiltonian in the half-filled
    Sz = 0 sector; 6 sites are exact at bond dimension 64.  The first
    sweeps carry the density-matrix noise term (as sample/dmrg.cc does),
    which moves weight into the sectors a product start state lacks."""
    from itensor_amd import workloads as bench
    from oracle import hubbard as OH

    N = Nx * Ny
    mpo, A = bench.hubbard_cylinder(H, Nx, Ny, 1.0, U, True)
    noise = (1e-4, 1e-5, 1e-6, 1e-7) + (0,) * 10
    sw = [{"MaxDim": m, "Cutoff": 1e-13, "Niter": 6, "Noise": z}
          for m, z in zip((16, 32) + (64,) * 12, noise)]
    E, recs, tens, secs = H.dmrg(mpo, A, sw, {})
    Hed = OH.hubbard_hamiltonian(N, OH.square_lattice(Nx, Ny, True), 1.0, U)
    Eed = OH.ground_energy(Hed, OH.electron_sector(N, N, 0))
    assert abs(E - Eed) < 1e-9, (E, Eed)


def test_autompo_device_mpo_matches_dense(H):
    """to_mpo's QN site tensors (block-sparse on the device) expand to the
    builder's dense tensors: boundary row 'start', column 'done'."""
    s = H.electron_sites(4)
    a = H.AutoMPO(s)
    for i, j in [(1, 2), (2, 3), (3, 4), (1, 4)]:
        a.add(-1.0, "Cdagup", i, "Cup", j)
        a.add(-1.0, "Cdagup", j, "Cup", i)
        a.add(-1.0, "Cdagdn", i, "Cdn", j)
        a.add(-1.0, "Cdagdn", j, "Cdn", i)
    for i in range(1, 5):
        a.add(4.0, "Nupdn", i)
    W, dims, links = a.dense()
    mpo = a.to_mpo()
    for n, (w, t) in enumerate(zip(W, mpo)):
        got = H.remove_qns(t).store()
        ref = w[:, :, 1, :] if n == 0 else (w[:, :, :, 0] if n == len(W) - 1 else w)
        assert np.array_equal(np.asarray(got).reshape(ref.shape, order="F"), ref), n


@pytest.mark.parametrize("fromleft", [True, False])
def test_delta_rho_matches_numpy(H, fromleft):
    """LocalOp.deltaRho (localop.h:375-406) on dense tensors: drho is the
    symmetrised X X^T of X = combine * noPrime(AA*L*W1) (or AA*R*W2), so
    its spectrum and trace equal those of the numpy product."""
    rng = np.random.default_rng(55 if fromleft else 56)
    dl, dr, d, k = 6, 5, 2, 3
    ids = iter(range(700001, 700100))

    def ix(n, tag, plev=0, iid=None):
        return H.Index(iid or next(ids), n, 1, plev, tag)

    l, s1, s2, r = ix(dl, "l"), ix(d, "s1"), ix(d, "s2"), ix(dr, "r")
    h0, h1, h2 = ix(k, "h0"), ix(k, "h1"), ix(k, "h2")
    AAa = rng.standard_normal((dl, d, d, dr))
    La, Ra = rng.standard_normal((dl, k, dl)), rng.standard_normal((dr, k, dr))
    W1a, W2a = rng.standard_normal((d, d, k, k)), rng.standard_normal((d, d, k, k))

    def mk(inds, a):
        return H.dense(inds, np.asarray(a).ravel(order="F"))

    AA = mk([l, s1, s2, r], AAa)
    L, R = mk([l, h0, l.prime()], La), mk([r, h2, r.prime()], Ra)
    W1, W2 = mk([s1, s1.prime(), h0, h1], W1a), mk([s2, s2.prime(), h1, h2], W2a)
    op = H.LocalOp(W1, W2, L, R, False)
    if fromleft:
        X = np.einsum("lsur,lhm,stha->mtura", AAa, La, W1a).reshape(dl * d, -1, order="F")
        C, _ = H.combiner([l, s1])
    else:
        X = np.einsum("lsur,rhm,uvah->vmlsa", AAa, Ra, W2a).reshape(d * dr, -1, order="F")
        C, _ = H.combiner([s2, r])
    drho, tr = op.delta_rho(AA, C, fromleft)
    n = X.shape[0]
    D = np.asarray(drho.store()).reshape(n, n, order="F")
    ref = X @ X.T
    assert np.max(np.abs(D - D.T)) < 1e-12
    assert np.max(np.abs(np.sort(np.linalg.eigvalsh(D)) - np.sort(np.linalg.eigvalsh(ref)))) < 1e-10 * np.max(ref)
    assert abs(tr - np.trace(ref)) < 1e-10 * np.trace(ref)


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("fuse", [False, True])
def test_local_product_full_block_layout(H, mode, fuse):
    """LocalOp::setCoverOutput: for a phi holding every flux-allowed block,
    in any index order, the product comes out in phi's index order and
    block layout, with the oracle's values (and exact zeros in the blocks
    no block pair reaches)."""
    N, b = 10, 5
    sites = OD.spin_half_sites(N, True)
    Hm = OD.heisenberg_mpo(sites)
    psi0 = OD.product_mps(sites, [i % 2 for i in range(N)])
    _, psi = OD.dmrg(Hm, psi0, [{"MaxDim": 20, "Cutoff": 1e-12}] * 2)
    PH = OD.LocalMPO(Hm)
    PH.position(b, psi)
    phi = psi(b) * psi(b + 1)
    ref = PH.product(phi)
    refD = ref.to_dense_array()
    phiD = phi.to_dense_array()
    div = [(n, v) for n, v in phi.div().q]
    op = H.LocalOp(cten(Hm[b - 1]), cten(Hm[b]), cten(PH.PH[b - 1]), cten(PH.PH[b + 2]), fuse)
    op.set_mode(mode)
    op.set_cover_output(True)
    for perm in ([0, 1, 2, 3], [3, 2, 1, 0], [1, 3, 0, 2]):
        inds = [phi.inds[p] for p in perm]
        full = H.from_dense([cidx(i) for i in inds], np.transpose(phiD, perm), div)
        assert len(full.offsets) >= len(phi.offsets)
        for rep in range(2):
            got = op.product(full)
            assert same_inds(got, full)
            assert [(tuple(x), o) for x, o in got.offsets] == [(tuple(x), o) for x, o in full.offsets]
            # the oracle's result in phi's order
            rperm = [[ (i.id, i.plev) for i in ref.inds].index((j.id, 0)) for j in inds]
            want = np.transpose(refD, rperm)
            assert _rel(oten(got).to_dense_array(), want) < 1e-12


def test_full_block_krylov_same_trajectory(H):
    """DMRG with the Krylov vectors on phi's full flux-allowed block set
    (default) and with the reference's growing block sets (FullBlockKrylov
    off) on the 4x2 Hubbard cylinder: the same energy after every bond step
    to 1e-10 and the same bond dimensions and truncation errors; the added
    blocks are exact zeros, so only the summation order of the grouped
    contractions can differ."""
    from itensor_amd import workloads as bench

    runs = []
    for full in (True, False):
        mpo, A = bench.hubbard_cylinder(H, 4, 2, 1.0, 4.0, True)
        sw = [{"MaxDim": m, "Cutoff": 1e-10, "Niter": 3, "Noise": z}
              for m, z in ((16, 1e-5), (32, 1e-6), (48, 0.0), (48, 0.0))]
        runs.append(H.dmrg(mpo, A, sw, {"FullBlockKrylov": full}))
    (E1, r1, _, _), (E2, r2, _, _) = runs
    assert abs(E1 - E2) < 1e-10, (E1, E2)
    assert len(r1) == len(r2)
    for a, b in zip(r1, r2):
        assert a["dim"] == b["dim"], (a, b)
        assert abs(a["energy"] - b["energy"]) < 1e-10 * max(1.0, abs(b["energy"])), (a, b)
        assert abs(a["truncerr"] - b["truncerr"]) < 1e-12 + 1e-6 * b["truncerr"], (a, b)


@pytest.mark.parametrize("full", [True, False])
def test_qn_dmrg_mps_block_lists_match_oracle(H, full):
    """QN DMRG N=12 with the Krylov vectors on phi's full block set (the
    default) and without: after five sweeps every MPS tensor has the
    oracle's block list (blocks and offsets), and every bond the oracle's
    kept dimension."""
    N = 12
    sites = OD.spin_half_sites(N, True)
    Hm = OD.heisenberg_mpo(sites)
    psi0 = OD.product_mps(sites, [i % 2 for i in range(N)])
    sw = [{"MaxDim": m, "Cutoff": 1e-12, "Niter": 2} for m in (10, 20, 40, 60, 60)]
    trace = []
    Eo, psio = OD.dmrg(Hm, psi0, sw, record=trace.append)
    Eg, recs, tens, secs = H.dmrg([cten(w) for w in Hm], [cten(a) for a in psi0.A], sw, {"FullBlockKrylov": full})
    assert [r["dim"] for r in recs] == [t["dim"] for t in trace]
    got = [[(tuple(b), o) for b, o in t.offsets] for t in tens]
    want = [[(tuple(b), o) for b, o in a.offsets] for a in psio.A]
    assert got == want
    assert abs(Eg - Eo) < 1e-10


def test_qn_dmrg_svd_route_matches_oracle(H):
    """Cutoff 1e-13 with no noise: every bond step of the reference takes
    the svd route (mps_impl.h:50-65); the host layer follows the oracle's
    per-bond energies to 1e-10 and its kept dimensions exactly."""
    N = 10
    sites = OD.spin_half_sites(N, True)
    Hm = OD.heisenberg_mpo(sites)
    psi0 = OD.product_mps(sites, [i % 2 for i in range(N)])
    sw = [{"MaxDim": m, "Cutoff": 1e-13, "Niter": 2} for m in (10, 20, 40, 40, 40)]
    trace = []
    Eo, _ = OD.dmrg(Hm, psi0, sw, record=trace.append)
    assert {t["route"] for t in trace} == {"svd"}
    Eg, recs, tens, secs = H.dmrg([cten(w) for w in Hm], [cten(a) for a in psi0.A], sw, {})
    assert [r["dim"] for r in recs] == [t["dim"] for t in trace]
    assert max(abs(r["energy"] - t["energy"]) for r, t in zip(recs, trace)) < 1e-10
    assert abs(Eg - GOLD["energies"]["10"]) < 1e-10


@pytest.mark.parametrize("qn", [False, True])
@pytest.mark.parametrize("fromleft", [True, False])
def test_svd_bond_routes_match_oracle(H, qn, fromleft):
    """svdBond on a two-site wavefunction in both routes: the kept spectrum
    and truncation error as the oracle's, the product of the new site
    tensors equal to the oracle's, the centre normalised."""
    N = 6
    sites = OD.spin_half_sites(N, qn)
    Hm = OD.heisenberg_mpo(sites)
    psi0 = OD.product_mps(sites, [i % 2 for i in range(N)]) if qn else OD.random_mps(sites, 4, seed=3)
    _, psi = OD.dmrg(Hm, psi0, [{"MaxDim": 8, "Cutoff": 1e-10}] * 2)
    b = 3
    phi = psi(b) * psi(b + 1)
    for cutoff, route in ((1e-13, "svd"), (1e-10, "denmat")):
        args = {"Cutoff": cutoff, "MaxDim": 6, "RespectDegenerate": True, "DoNormalize": True}
        Ao, Bo, so = OD.svd_bond(psi(b), psi(b + 1), phi, fromleft, args)
        assert so["route"] == route
        Ag, Bg, eigs, terr = H.svd_bond([cten(a) for a in psi.A], b, cten(phi), fromleft, args)
        assert np.allclose(np.sort(eigs)[::-1], np.sort(so["eigs"])[::-1], rtol=0, atol=1e-12)
        assert abs(terr - so["truncerr"]) < 1e-14
        prod_o = Ao * Bo
        prod_g = oten(Ag * Bg)
        perm = [[(i.id, i.plev) for i in prod_g.inds].index((j.id, j.plev)) for j in prod_o.inds]
        assert _rel(np.transpose(prod_g.to_dense_array(), perm), prod_o.to_dense_array()) < 1e-12
        centre = Bg if fromleft else Ag
        assert abs(H.norm(centre) - 1.0) < 1e-13


def test_position_orthogonalises_without_changing_the_state(H):
    """MPS::position(1) (orthMPS route) on an MPS that is not orthogonal
    (rightLim = N+1): the state is unchanged (to 1e-12), sites 2..N become
    right-orthogonal, and the limits become (0, 2)."""
    N = 6
    sites = OD.spin_half_sites(N, False)
    psi = OD.random_mps(sites, 4, seed=9)
    rng = np.random.default_rng(4)
    A = [a.copy() for a in psi.A]
    for a in A:  # break the orthogonality, keep the structure
        a.d = a.d * (1.0 + 0.3 * rng.standard_normal(a.d.shape))

    def full_state(ts):
        x = ts[0]
        for t in ts[1:]:
            x = x * t
        return x

    ref = full_state(A)
    tens, ll, rl = H.position([cten(a) for a in A], 1, 0, N + 1)
    assert (ll, rl) == (0, 2)
    got = oten(full_state(tens))
    perm = [[(i.id, i.plev) for i in got.inds].index((j.id, j.plev)) for j in ref.inds]
    assert _rel(np.transpose(got.to_dense_array(), perm), ref.to_dense_array()) < 1e-12
    for j in range(1, N):
        T = oten(tens[j])
        lk = [I for I in T.inds if "Link" in I.tags and any(I.id == K.id for K in oten(tens[j - 1]).inds)][0]
        G = T * T.prime(1, lk).dag()
        assert np.allclose(G.to_dense_array(), np.eye(lk.dim), atol=1e-12)


@pytest.mark.parametrize("qn", [False, True])
def test_environment_offload_write_dim(H, qn):
    """Args WriteDim (dmrg.h:390-402, localmpo.h:620-681): offload starts
    with the first sweep whose maxdim reaches it; environments built before
    then leave device memory as the limits pass them (as the reference's
    setLHlim / setRHlim write them out), so from the next sweep on only the
    two environments in use stay in device memory and the rest sit in
    pinned host memory.  The copies are exact: every bond energy equals that
    of the run without offload."""
    N = 14
    sites = OD.spin_half_sites(N, qn)
    Hm = OD.heisenberg_mpo(sites)
    psi0 = OD.product_mps(sites, [i % 2 for i in range(N)]) if qn else OD.random_mps(sites, 4, seed=2)
    sw = [{"MaxDim": m, "Cutoff": 1e-12, "Niter": 2} for m in (8, 16, 32, 32)]
    runs = []
    for wd in (None, 16):
        args = {} if wd is None else {"WriteDim": wd}
        runs.append(H.dmrg([cten(w) for w in Hm], [cten(a) for a in psi0.A], sw, args))
    (E0, r0, _, _), (E1, r1, _, _) = runs
    assert [r["energy"] for r in r0] == [r["energy"] for r in r1]
    assert [r["dim"] for r in r0] == [r["dim"] for r in r1]
    on = [r for r in r1 if r["sweep"] >= 3]
    print("[offload] device envs", [r["envs_device"] for r in r1], "host envs", [r["envs_host"] for r in r1],
          flush=True)
    assert max(r["envs_device"] for r in on) <= 2
    assert min(r["envs_host"] for r in on) >= N - 4
    assert max(r["envs_device"] for r in r0) > 2
