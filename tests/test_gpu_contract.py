"""GPU parity of the dense contraction engine (itx_contract_dense / plans)
against the oracle's contract() (tests call through the C ABI)."""
import itertools

import numpy as np
import pytest

from oracle.dense import contract, hpsi_dense, HPSI_STAGES, hpsi_flops, heisenberg_w
from tests.conftest import relerr

pytestmark = pytest.mark.gpu

TOL64 = 1e-12   # FP64 MFMA vs FP64 BLAS: relative to max |C|
TOL32 = 2e-5    # FP32 MFMA form
TOLBF = 1e-4    # BF16x3 form (operands carried to 16-17 bits, FP32 accumulate)
TOLS = {"f64": TOL64, "f32": TOL32, "bf16x3": TOLBF}


def _run(ctx, dev, A, labA, B, labB, labC, prec="f64", alpha=1.0, beta=0.0, C0=None):
    ref = contract(A, labA, B, labB, labC, alpha=alpha, beta=beta, C=C0)
    dA, dB = dev.put(A), dev.put(B)
    dC = dev.put(C0) if C0 is not None else dev.empty(ref.size)
    ctx.contract_dense(dA.data_ptr(), A.shape, labA, dB.data_ptr(), B.shape, labB,
                       dC.data_ptr(), ref.shape, labC, alpha, beta, prec)
    return dev.get(dC, ref.shape), ref


# Cases modelled on the reference's contract_test.cc permutation cases
# (NM1-NM5 non-matrix, M1-M4 matrix-like, rank-0 outputs), with dims chosen
# so that tile edges are exercised.
CASES = [
    # (dimsA, labA, dimsB, labB, labC)
    ((7, 5), (1, -1), (5, 9), (-1, 2), (1, 2)),               # plain GEMM
    ((5, 7), (-1, 1), (5, 9), (-1, 2), (1, 2)),               # A^T B
    ((7, 5), (1, -1), (9, 5), (2, -1), (1, 2)),               # A B^T
    ((7, 5), (1, -1), (5, 9), (-1, 2), (2, 1)),               # C^T
    ((3, 4, 5), (1, -1, 2), (4, 6), (-1, 3), (2, 3, 1)),      # NM: mid contracted
    ((3, 4, 5, 2), (-1, 1, -2, 2), (5, 3, 7), (-2, -1, 3), (3, 1, 2)),
    ((2, 3, 4, 5), (1, 2, -1, -2), (5, 4, 3), (-2, -1, 3), (1, 2, 3)),
    ((6, 2, 2, 6), (-1, 2, 3, 4), (6, 5, 6), (-1, 5, 6), (2, 3, 4, 5, 6)),
    ((33, 17), (-1, -2), (17, 33), (-2, -1), ()),             # full contraction -> scalar
    ((9,), (1,), (4,), (2,), (1, 2)),                          # outer product
    ((130, 70), (1, -1), (70, 129), (-1, 2), (1, 2)),          # > one 128 tile
    ((40, 300), (-1, 1), (40, 257), (-1, 2), (2, 1)),
]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("prec", ["f64", "f32", "bf16x3"])
def test_dense_cases(ctx, dev, case, prec):
    dA, labA, dB, labB, labC = CASES[case]
    rng = np.random.default_rng(100 + case)
    A = rng.standard_normal(dA)
    B = rng.standard_normal(dB)
    got, ref = _run(ctx, dev, A, labA, B, labB, labC, prec)
    err = relerr(got, ref)
    print(f"[dense case] {case} {prec} err={err:.2e}", flush=True)
    assert err < TOLS[prec], (case, prec)


def test_random_permutations(ctx, dev):
    rng = np.random.default_rng(7)
    for trial in range(40):
        rA = int(rng.integers(1, 5))
        rB = int(rng.integers(1, 5))
        nc = int(rng.integers(0, min(rA, rB) + 1))
        dims_c = [int(x) for x in rng.integers(1, 9, size=nc)]
        da = [int(x) for x in rng.integers(1, 9, size=rA - nc)]
        db = [int(x) for x in rng.integers(1, 9, size=rB - nc)]
        la = [-(i + 1) for i in range(nc)] + [10 + i for i in range(rA - nc)]
        lb = [-(i + 1) for i in range(nc)] + [20 + i for i in range(rB - nc)]
        dimsA = dims_c + da
        dimsB = dims_c + db
        pa = rng.permutation(rA)
        pb = rng.permutation(rB)
        dimsA = [dimsA[i] for i in pa]; la = [la[i] for i in pa]
        dimsB = [dimsB[i] for i in pb]; lb = [lb[i] for i in pb]
        labC = [l for l in la if l > 0] + [l for l in lb if l > 0]
        labC = [labC[i] for i in rng.permutation(len(labC))]
        A = rng.standard_normal(dimsA)
        B = rng.standard_normal(dimsB)
        got, ref = _run(ctx, dev, A, la, B, lb, labC)
        assert relerr(got, ref) < TOL64, (trial, dimsA, la, dimsB, lb, labC)


def test_alpha_beta(ctx, dev):
    rng = np.random.default_rng(3)
    A = rng.standard_normal((20, 30))
    B = rng.standard_normal((30, 25))
    C0 = rng.standard_normal((20, 25))
    got, ref = _run(ctx, dev, A, (1, -1), B, (-1, 2), (1, 2), "f64", 0.5, -2.0, C0)
    assert relerr(got, ref) < TOL64


@pytest.mark.parametrize("chi", [16, 64, 200])
def test_hpsi_chain(ctx, dev, chi):
    rng = np.random.default_rng(chi)
    psi = rng.standard_normal((chi, 2, 2, chi))
    L = rng.standard_normal((chi, 5, chi))
    R = rng.standard_normal((chi, 5, chi))
    W = heisenberg_w()
    ref = hpsi_dense(psi, L, W, W, R)
    x = dev.put(psi)
    dims = psi.shape
    for (la, lb, lc), op in zip(HPSI_STAGES, (L, W, W, R)):
        # output dims follow from labels
        outd = [d for d, l in zip(dims, la) if l > 0] + [d for d, l in zip(op.shape, lb) if l > 0]
        order = [l for l in la if l > 0] + [l for l in lb if l > 0]
        outd = [outd[order.index(l)] for l in lc]
        y = dev.empty(int(np.prod(outd)))
        dop = dev.put(op)
        ctx.contract_dense(x.data_ptr(), dims, la, dop.data_ptr(), op.shape, lb, y.data_ptr(), outd, lc)
        x, dims = y, tuple(outd)
    got = dev.get(x, dims)
    assert relerr(got, ref) < TOL64


def test_plan_reuse_and_flops(ctx, dev):
    rng = np.random.default_rng(11)
    la, lb, lc = HPSI_STAGES[0]
    chi = 48
    psi = rng.standard_normal((chi, 2, 2, chi))
    L = rng.standard_normal((chi, 5, chi))
    ref = contract(psi, la, L, lb, lc)
    plan = ctx.plan_dense(psi.shape, la, L.shape, lb, ref.shape, lc)
    assert plan.flops == pytest.approx(2.0 * (4 * chi) * (5 * chi) * chi)
    dA, dB = dev.put(psi), dev.put(L)
    for _ in range(3):
        dC = dev.empty(ref.size)
        plan.execute(dA.data_ptr(), dB.data_ptr(), dC.data_ptr())
        assert relerr(dev.get(dC, ref.shape), ref) < TOL64
    plan.close()


def test_bad_labels_raise(ctx, dev):
    from itensor_amd.itx import ItxError

    A = dev.empty(6, 0.0)
    with pytest.raises(ItxError):
        ctx.contract_dense(A.data_ptr(), (2, 3), (1, -1), A.data_ptr(), (2, 3), (-2, 2), A.data_ptr(), (2, 3), (1, 2))
    with pytest.raises(ItxError):  # contracted dims differ
        ctx.contract_dense(A.data_ptr(), (2, 3), (1, -1), A.data_ptr(), (2, 3), (-1, 2), A.data_ptr(), (2, 3), (1, 2))


def test_autotune_picks_a_tile_and_persists(ctx, dev, tmp_path):
    rng = np.random.default_rng(21)
    la, lb, lc = HPSI_STAGES[0]
    chi = 96
    psi = rng.standard_normal((chi, 2, 2, chi))
    L = rng.standard_normal((chi, 5, chi))
    ref = contract(psi, la, L, lb, lc)
    ctx.set_autotune(True)
    try:
        n0 = ctx.tune_size()
        plan = ctx.plan_dense(psi.shape, la, L.shape, lb, ref.shape, lc)
        assert ctx.tune_size() == n0 + 1
        dA, dB, dC = dev.put(psi), dev.put(L), dev.empty(ref.size)
        plan.execute(dA.data_ptr(), dB.data_ptr(), dC.data_ptr())
        assert relerr(dev.get(dC, ref.shape), ref) < TOL64
        plan.close()
        path = str(tmp_path / "tune.txt")
        ctx.tune_save(path)
        ctx.tune_load(path)
        assert ctx.tune_size() == n0 + 1
    finally:
        ctx.set_autotune(False)


def test_chain_graph_replay_matches_direct(ctx, dev):
    """itx_chain: the four H*psi stages as one HIP graph.  Replays with new
    operand pointers (two psi inputs, fresh output buffers), with eager
    work and host syncs between them, are bitwise equal to the directly
    executed plans, and match the CPU chain to 1e-12."""
    import torch

    chi = 40
    rng = np.random.default_rng(5)
    L = rng.standard_normal((chi, 5, chi))
    R = rng.standard_normal((chi, 5, chi))
    W = heisenberg_w()
    psis = [rng.standard_normal((chi, 2, 2, chi)) for _ in range(2)]
    ops = [dev.put(L), dev.put(W), dev.put(W), dev.put(R)]
    dims, plans, mids = psis[0].shape, [], []
    for (la, lb, lc), op in zip(HPSI_STAGES, (L, W, W, R)):
        outd = [d for d, l in zip(dims, la) if l > 0] + [d for d, l in zip(op.shape, lb) if l > 0]
        order = [l for l in la if l > 0] + [l for l in lb if l > 0]
        outd = tuple(outd[order.index(l)] for l in lc)
        plans.append(ctx.plan_dense(dims, la, op.shape, lb, outd, lc))
        mids.append(dev.empty(int(np.prod(outd))))
        dims = outd
    chain = ctx.chain(plans)
    try:
        for rep in range(4):
            k = rep % 2
            x = dev.put(psis[k])
            out = dev.empty(x.numel(), float("nan"))
            A = [x.data_ptr()] + [m.data_ptr() for m in mids[:-1]]
            C = [m.data_ptr() for m in mids[:-1]] + [out.data_ptr()]
            chain.launch(A, [o.data_ptr() for o in ops], C)
            torch.cuda.synchronize()
            # eager reference through the same plans
            ref_out = dev.empty(x.numel())
            Cr = [m.data_ptr() for m in mids[:-1]] + [ref_out.data_ptr()]
            for p, a, b, c in zip(plans, A, [o.data_ptr() for o in ops], Cr):
                p.execute(a, b, c)
            torch.cuda.synchronize()
            assert torch.equal(out, ref_out), f"replay {rep}"
            ref = hpsi_dense(psis[k], L, W, W, R)
            assert relerr(dev.get(out, psis[k].shape), ref) < TOL64
    finally:
        chain.close()
        for p in plans:
            p.close()


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16x3"])
def test_small_block_kernel_general_shapes(ctx, dev, monkeypatch, prec):
    """The small-block kernels forced onto every case shape (K above one
    64-wide chunk, N above one 32-wide unit, rank-0 results, permuted
    operands), with alpha and beta: FP64 on the VALU, FP32 and BF16x3 on
    the matrix cores, each to its form's tolerance."""
    monkeypatch.setenv("ITX_FORCE_TILE", "10")
    rng = np.random.default_rng(55)
    for case, (dA, labA, dB, labB, labC) in enumerate(CASES):
        A = rng.standard_normal(dA)
        B = rng.standard_normal(dB)
        ref0 = contract(A, labA, B, labB, labC)
        C0 = rng.standard_normal(ref0.shape) if ref0.shape else np.array(rng.standard_normal())
        ref = contract(A, labA, B, labB, labC, alpha=0.7, beta=0.3, C=C0)
        plan = ctx.plan_dense(A.shape, labA, B.shape, labB, ref.shape, labC, prec)
        try:
            d_A, d_B, d_C = dev.put(A), dev.put(B), dev.put(C0)
            plan.execute(d_A.data_ptr(), d_B.data_ptr(), d_C.data_ptr(), 0.7, 0.3)
            got = dev.get(d_C, ref.shape)
            info = plan.info
        finally:
            plan.close()
        assert info["tile"] == 10 and info["form"] == prec
        # reduced forms: rounding error bounded by the form's unit times the
        # sum of |terms| (a rank-0 result cancels far below its terms)
        scale = contract(np.abs(A), labA, np.abs(B), labB, labC, alpha=0.7, beta=0.3, C=np.abs(C0))
        err = np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), np.max(scale) if prec != "f64" else 0.0)
        assert err < TOLS[prec], case


SPLIT_CASES = [
    # few small output blocks, long contracted extent: split over k chunks
    ((20, 5000), (1, -1), (5000, 20), (-1, 2), (1, 2)),
    ((3, 700, 7), (1, -1, 2), (700, 5), (-1, 3), (3, 1, 2)),
    ((3000, 64), (-1, 1), (3000, 33), (-1, 2), (2, 1)),
    ((40, 40, 9), (-1, -2, 1), (40, 40, 11), (-2, -1, 2), (1, 2)),
]


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16x3"])
@pytest.mark.parametrize("case", range(len(SPLIT_CASES)))
def test_split_small_plans(ctx, dev, case, prec):
    """Small output blocks with a long contracted extent: the plan splits
    the k chunks over several workgroups and a reduction pass writes C
    (alpha, beta honoured); the result matches the oracle to 1e-12 and is
    bitwise reproducible across executions."""
    dA, labA, dB, labB, labC = SPLIT_CASES[case]
    rng = np.random.default_rng(300 + case)
    A = rng.standard_normal(dA)
    B = rng.standard_normal(dB)
    C0 = rng.standard_normal(contract(A, labA, B, labB, labC).shape)
    ref = contract(A, labA, B, labB, labC, alpha=1.3, beta=-0.4, C=C0)
    plan = ctx.plan_dense(A.shape, labA, B.shape, labB, ref.shape, labC, prec)
    try:
        d_A, d_B = dev.put(A), dev.put(B)
        outs = []
        for _ in range(2):
            d_C = dev.put(C0)
            plan.execute(d_A.data_ptr(), d_B.data_ptr(), d_C.data_ptr(), 1.3, -0.4)
            outs.append(dev.get(d_C, ref.shape))
    finally:
        plan.close()
    assert relerr(outs[0], ref) < TOLS[prec]
    assert np.array_equal(outs[0], outs[1])


def test_chain_with_split_plan(ctx, dev):
    """A two-stage chain whose first plan is split (contraction + reduction
    nodes) replayed with new operands matches direct execution bitwise."""
    import torch

    rng = np.random.default_rng(9)
    A = rng.standard_normal((20, 4000))
    B = rng.standard_normal((4000, 20))
    D = rng.standard_normal((20, 30))
    p1 = ctx.plan_dense((20, 4000), (1, -1), (4000, 20), (-1, 2), (20, 20), (1, 2))
    p2 = ctx.plan_dense((20, 20), (1, -1), (20, 30), (-1, 2), (20, 30), (1, 2))
    chain = ctx.chain([p1, p2])
    try:
        dB, dD = dev.put(B), dev.put(D)
        mid = dev.empty(400)
        for rep in range(3):
            Ar = A * (rep + 1)
            dA = dev.put(Ar)
            out = dev.empty(600)
            chain.launch([dA.data_ptr(), mid.data_ptr()], [dB.data_ptr(), dD.data_ptr()],
                         [mid.data_ptr(), out.data_ptr()])
            torch.cuda.synchronize()
            ref_out = dev.empty(600)
            p1.execute(dA.data_ptr(), dB.data_ptr(), mid.data_ptr())
            p2.execute(mid.data_ptr(), dD.data_ptr(), ref_out.data_ptr())
            torch.cuda.synchronize()
            assert torch.equal(out, ref_out), rep
            assert relerr(dev.get(out, (20, 30)), (Ar @ B) @ D) < TOL64
    finally:
        chain.close()
        p1.close()
        p2.close()


@pytest.mark.parametrize("prec", ["f32", "bf16x3"])
@pytest.mark.parametrize("chi", [64, 200])
def test_hpsi_chain_reduced_precision(ctx, dev, chi, prec):
    """The four H*psi stages in the FP32 and BF16x3 forms: error against the
    FP64 CPU chain within the form's tolerance (relative to max |H psi|)."""
    rng = np.random.default_rng(chi + 7)
    psi = rng.standard_normal((chi, 2, 2, chi))
    L = rng.standard_normal((chi, 5, chi))
    R = rng.standard_normal((chi, 5, chi))
    W = heisenberg_w()
    ref = hpsi_dense(psi, L, W, W, R)
    x = dev.put(psi)
    dims = psi.shape
    for (la, lb, lc), op in zip(HPSI_STAGES, (L, W, W, R)):
        outd = [d for d, l in zip(dims, la) if l > 0] + [d for d, l in zip(op.shape, lb) if l > 0]
        order = [l for l in la if l > 0] + [l for l in lb if l > 0]
        outd = [outd[order.index(l)] for l in lc]
        y = dev.empty(int(np.prod(outd)))
        dop = dev.put(op)
        ctx.contract_dense(x.data_ptr(), dims, la, dop.data_ptr(), op.shape, lb, y.data_ptr(), outd, lc,
                           1.0, 0.0, prec)
        x, dims = y, tuple(outd)
    err = relerr(dev.get(x, dims), ref)
    print(f"[hpsi chain] chi={chi} {prec} err={err:.2e}", flush=True)
    assert err < TOLS[prec]
