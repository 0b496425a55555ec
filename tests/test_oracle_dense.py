"""CPU: pin oracle/dense.py -- contract() against explicit loops over the
reference's own contraction cases (unittest/contract_test.cc:161-486:
Case 1-4, NM1-NM5, M1-M4, Regression Test 1), and the H*psi stage labels
(HPSI_STAGES, hpsi_dense) against a brute-force sum of the two-site
effective Hamiltonian and against the ITensor-level chain of
oracle.dmrg.local_op_product (mps/localop.h:324-365), stage by stage."""
import itertools

import numpy as np
import pytest

from oracle import dmrg as OD
from oracle.dense import HPSI_STAGES, contract, heisenberg_w, hpsi_dense, hpsi_flops, hpsi_stage_shapes
from oracle.itensor import Index, dense
from oracle.itensor import contract as icontract

# (A labels, B labels, C labels, operands swapped in the call) of
# contract_test.cc; a label is also its extent, labels on both operands are
# contracted
CASES = {
    "Case 1": ((2, 3, 4), (3, 7, 2), (4, 7)),
    "Case 2": ((2, 3, 4), (3, 7, 2), (7, 4)),
    "Case 3": ((2, 4, 3), (3, 7, 2), (7, 4)),
    "Case 4": ((2, 4, 3), (3, 7, 2), (4, 7)),
    "NM1": ((2, 3, 4), (7, 3, 2), (4, 7)),
    "NM2": ((2, 3, 4, 5), (7, 6, 3, 2), (5, 4, 6, 7)),
    "NM3": ((7, 6, 3, 2), (2, 3, 4, 5), (5, 4, 6, 7)),
    "NM4": ((8, 7, 5, 6, 9), (2, 3, 4, 5, 6, 7), (2, 8, 4, 3, 9)),
    "NM5": ((8, 7, 5, 6, 9), (2, 3, 4, 5, 6, 7), (4, 9, 2, 3, 8)),
    "M1": ((2, 3, 4), (7, 2, 3), (4, 7)),
    "M2": ((4, 2, 3), (7, 2, 3), (4, 7)),
    "M3": ((4, 2, 3), (2, 3, 7), (4, 7)),
    "M4": ((2, 3, 4), (2, 3, 7), (4, 7)),
    "Regression Test 1": ((4, 3, 2), (5, 4, 3, 2), (5,)),
}


def _signed(la, lb):
    """tensor-level labels -> computeLabels convention (contracted < 0)."""
    common = set(la) & set(lb)
    return [-l if l in common else l for l in la], [-l if l in common else l for l in lb]


def _brute(A, la, B, lb, lc):
    common = [l for l in la if l in lb]
    out = np.zeros([l for l in lc])
    for u in itertools.product(*[range(l) for l in lc]):
        env = dict(zip(lc, u))
        s = 0.0
        for c in itertools.product(*[range(l) for l in common]):
            env.update(zip(common, c))
            s += A[tuple(env[l] for l in la)] * B[tuple(env[l] for l in lb)]
        out[u] = s
    return out


@pytest.mark.parametrize("name", list(CASES))
def test_contract_matches_loops_of_reference_cases(name):
    la, lb, lc = CASES[name]
    rng = np.random.default_rng(abs(hash(name)) % 2**32)
    A = rng.standard_normal(la)
    B = rng.standard_normal(lb)
    sa, sb = _signed(la, lb)
    got = contract(A, sa, B, sb, lc)
    assert got.shape == tuple(lc)
    ref = _brute(A, la, B, lb, lc)
    assert np.max(np.abs(got - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_contract_alpha_beta_and_default_order():
    rng = np.random.default_rng(7)
    A = rng.standard_normal((2, 3, 4))
    B = rng.standard_normal((3, 7, 2))
    C = rng.standard_normal((4, 7))
    T = _brute(A, (2, 3, 4), B, (3, 7, 2), (4, 7))
    got = contract(A, (-2, -3, 4), B, (-3, 7, -2), (4, 7), alpha=2.0, beta=-0.5, C=C)
    assert np.allclose(got, 2.0 * T - 0.5 * C, atol=1e-12)
    # default C order: uncontracted A labels, then B's (contractIS)
    assert contract(A, (-2, -3, 4), B, (-3, 7, -2)).shape == (4, 7)
    with pytest.raises(ValueError):
        contract(A, (-2, -3, 4), B, (-3, 7, -5))


def _hpsi_brute(psi, L, W1, W2, R):
    """H psi (l', s1', s2', r') = sum L(l,h0,l') W1(h0,s1,s1',h1)
    W2(h1,s2,s2',h2) R(r,h2,r') psi(l,s1,s2,r), written out term by term."""
    chi, d = psi.shape[0], psi.shape[1]
    k = L.shape[1]
    out = np.zeros(psi.shape)
    for lp, a, b, rp in itertools.product(range(chi), range(d), range(d), range(chi)):
        s = 0.0
        for l, s1, s2, r in itertools.product(range(chi), range(d), range(d), range(chi)):
            p = psi[l, s1, s2, r]
            if p == 0.0:
                continue
            for h0, h1, h2 in itertools.product(range(k), range(k), range(k)):
                s += L[l, h0, lp] * W1[h0, s1, a, h1] * W2[h1, s2, b, h2] * R[r, h2, rp] * p
        out[lp, a, b, rp] = s
    return out


def _operands(chi, seed, d=2, k=5, heis=False):
    rng = np.random.default_rng(seed)
    psi = rng.standard_normal((chi, d, d, chi))
    L = rng.standard_normal((chi, k, chi))
    R = rng.standard_normal((chi, k, chi))
    if heis:
        W1 = W2 = heisenberg_w()
    else:
        W1 = rng.standard_normal((k, d, d, k))
        W2 = rng.standard_normal((k, d, d, k))
    return psi, L, W1, W2, R


@pytest.mark.parametrize("heis", [False, True])
def test_hpsi_dense_matches_brute_force(heis):
    psi, L, W1, W2, R = _operands(2, 3, heis=heis)
    got = hpsi_dense(psi, L, W1, W2, R)
    ref = _hpsi_brute(psi, L, W1, W2, R)
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))


def test_hpsi_stages_follow_itensor_chain():
    """Each HPSI_STAGES entry gives the index order ITensor's operator*
    gives (contractIS), and the chain equals local_op_product on ITensors
    with the named indices, including the final (l, s1, s2, r) order."""
    chi, d, k = 3, 2, 5
    psi, L, W1, W2, R = _operands(chi, 11)
    l, s1, s2, r = Index(chi, "l"), Index(d, "Site,s1"), Index(d, "Site,s2"), Index(chi, "r")
    h0, h1, h2 = Index(k, "h0"), Index(k, "h1"), Index(k, "h2")
    T = {
        "psi": dense([l, s1, s2, r], psi),
        "L": dense([l, h0, l.prime()], L),
        "W1": dense([h0, s1, s1.prime(), h1], W1),
        "W2": dense([h1, s2, s2.prime(), h2], W2),
        "R": dense([r, h2, r.prime()], R),
    }
    x, xt = psi, T["psi"]
    for (la, lb, lc), name, arr in zip(HPSI_STAGES, ("L", "W1", "W2", "R"), (L, W1, W2, R)):
        x = contract(x, la, arr, lb, lc)
        xt = icontract(xt, T[name])
        assert x.shape == tuple(I.dim for I in xt.inds)
        assert np.max(np.abs(x - xt.d)) < 1e-12 * np.max(np.abs(xt.d))
    ref = OD.local_op_product(T["psi"], T["L"], T["W1"], T["W2"], T["R"])
    assert ref.inds == [l, s1, s2, r]
    assert np.max(np.abs(hpsi_dense(psi, L, W1, W2, R) - ref.d)) < 1e-12 * np.max(np.abs(ref.d))


def test_hpsi_shapes_and_flops():
    """Stage shapes agree with the chain; flops are 80 chi^3 + 800 chi^2 at
    d=2, k=5 and equal the sum of 2*m*n*k over the four stage GEMMs."""
    chi = 4
    psi, L, W1, W2, R = _operands(chi, 1)
    x = psi
    for (la, lb, lc), op, (da, db, dc) in zip(HPSI_STAGES, (L, W1, W2, R), hpsi_stage_shapes(chi)):
        assert x.shape == da and op.shape == db
        x = contract(x, la, op, lb, lc)
        assert x.shape == dc
    for c in (256, 512, 1024):
        assert hpsi_flops(c) == 80 * c ** 3 + 800 * c ** 2
    gemm = 0
    for (la, lb, lc), (da, db, dc) in zip(HPSI_STAGES, hpsi_stage_shapes(chi)):
        kk = int(np.prod([n for n, lab in zip(da, la) if lab < 0]))
        gemm += 2 * np.prod(dc) * kk
    assert gemm == hpsi_flops(chi)
