"""GPU parity of the register-pipelined tiles (k_dgemm, tile 14, and its
two-workgroups-per-CU variant k_dgemm2, tile 15).

Both serve FP64 and FP32-form plans whose operands are both contiguous
along the contracted index in aligned pairs.  Each is forced (ITX_FORCE_TILE)
on:

* plain contractions of every pipeline depth (K = 2 .. 1030: one stage,
  the two- to five-stage prologue/tail paths, and the steady-state loop),
  with M and N off the 128 tile edges, alpha and beta, against the oracle's
  contract() at 1e-12 relative to max |C| (FP64) and 2e-5 (FP32 form);
* the S1 and S4 stages of the dense H*psi chain (S4: a two-digit
  contracted index whose digits are apart in the operand);
* operands whose base is not 16-byte aligned (the plan falls back to its
  scalar kernel with the same tile and must give the same result);
* block-sparse H*psi at chi = 1024 over 31 sectors (grouped tasks: several
  k blocks summed into one output tile), against local_op_product;
* a plan that does not meet the pair condition (odd K): the forced tile is
  refused and a general tile runs instead.

Tile 15 sums in the k order of tile 14: their results are equal bit for bit.
"""
import numpy as np
import pytest

from oracle.dense import contract
from oracle.dmrg import local_op_product
from tests.bridge import cten
from tests.conftest import relerr
from tests.hpsi_cases import SECTORS_1024, hpsi_operands

pytestmark = pytest.mark.gpu

TILE = 14
TILES = [(14, "f64"), (14, "f32"), (15, "f64"), (15, "f32")]
SHAPES = [
    # (M, N, K): K / 16 stages = 1, 1, 2, 3, 4, 5, 6, 7, 65
    (1, 1, 2),
    (5, 7, 16),
    (128, 128, 32),
    (129, 257, 48),
    (300, 77, 64),
    (64, 200, 80),
    (200, 64, 96),
    (130, 140, 112),
    (513, 129, 1030),
]


TOL = {"f64": 1e-12, "f32": 2e-5}


def _plan_run(ctx, dev, A, labA, B, labB, labC, alpha=1.0, beta=0.0, C0=None, shift=0, prec="f64"):
    ref = contract(A, labA, B, labB, labC, alpha=alpha, beta=beta, C=C0)
    plan = ctx.plan_dense(A.shape, labA, B.shape, labB, ref.shape, labC, prec)
    try:
        if shift:
            # operands one element into a buffer: 8-byte, not 16-byte, aligned
            bA = dev.put(np.concatenate([[0.0], A.ravel(order="F")]))
            bB = dev.put(np.concatenate([[0.0], B.ravel(order="F")]))
            pA, pB = bA.data_ptr() + 8, bB.data_ptr() + 8
        else:
            bA, bB = dev.put(A), dev.put(B)
            pA, pB = bA.data_ptr(), bB.data_ptr()
        dC = dev.put(C0) if C0 is not None else dev.empty(ref.size)
        plan.execute(pA, pB, dC.data_ptr(), alpha, beta)
        got = dev.get(dC, ref.shape)
        info = plan.info
    finally:
        plan.close()
    return got, ref, info


@pytest.mark.parametrize("tile,prec", TILES)
@pytest.mark.parametrize("shape", SHAPES)
def test_dgemm_tile_every_depth(ctx, dev, monkeypatch, shape, tile, prec):
    monkeypatch.setenv("ITX_FORCE_TILE", str(tile))
    M, N, K = shape
    rng = np.random.default_rng(9000 + M + N + K)
    A = rng.standard_normal((K, M))
    B = rng.standard_normal((K, N))
    got, ref, info = _plan_run(ctx, dev, A, (-1, 1), B, (-1, 2), (1, 2), prec=prec)
    assert info["tile"] == tile and info["form"] == prec
    assert relerr(got, ref) < TOL[prec], shape
    C0 = rng.standard_normal((M, N))
    got, ref, _ = _plan_run(ctx, dev, A, (-1, 1), B, (-1, 2), (1, 2), 0.7, -0.3, C0, prec=prec)
    assert relerr(got, ref) < TOL[prec], shape


@pytest.mark.parametrize("tile,prec", TILES)
@pytest.mark.parametrize("stage", ["S1", "S4"])
def test_dgemm_tile_hpsi_stages(ctx, dev, monkeypatch, stage, tile, prec):
    """The S1 stage (psi x L over one index) and the S4 stage (T3 x L over
    a two-digit index, its digits apart in T3) of the dense H*psi chain at
    chi = 256, with the labels the chain uses."""
    monkeypatch.setenv("ITX_FORCE_TILE", str(tile))
    chi = 256
    rng = np.random.default_rng(77)
    L = rng.standard_normal((chi, 5, chi))
    if stage == "S1":
        psi = rng.standard_normal((chi, 2, 2, chi))
        got, ref, info = _plan_run(ctx, dev, psi, (-1, 2, 3, 4), L, (-1, 5, 6), (2, 3, 4, 5, 6), prec=prec)
    else:
        T3 = rng.standard_normal((chi, chi, 2, 2, 5))
        got, ref, info = _plan_run(ctx, dev, T3, (-1, 3, 4, 5, -2), L, (-1, -2, 6), (3, 4, 5, 6), prec=prec)
    assert info["tile"] == tile
    assert relerr(got, ref) < TOL[prec]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("shape", [(129, 257, 48), (513, 129, 1030)])
def test_dgemm_two_per_cu_equals_tile14(ctx, dev, monkeypatch, shape, prec):
    M, N, K = shape
    rng = np.random.default_rng(404 + K)
    A = rng.standard_normal((K, M))
    B = rng.standard_normal((K, N))
    out = {}
    for tile in (14, 15):
        monkeypatch.setenv("ITX_FORCE_TILE", str(tile))
        out[tile], _, info = _plan_run(ctx, dev, A, (-1, 1), B, (-1, 2), (1, 2), prec=prec)
        assert info["tile"] == tile
    assert np.array_equal(out[14], out[15])


@pytest.mark.parametrize("tile", [14, 15])
@pytest.mark.parametrize("shape", [(129, 257, 48), (200, 64, 96)])
def test_dgemm_tile_unaligned_operands(ctx, dev, monkeypatch, shape, tile):
    monkeypatch.setenv("ITX_FORCE_TILE", str(tile))
    M, N, K = shape
    rng = np.random.default_rng(31 + K)
    A = rng.standard_normal((K, M))
    B = rng.standard_normal((K, N))
    got, ref, info = _plan_run(ctx, dev, A, (-1, 1), B, (-1, 2), (1, 2), shift=1)
    assert info["tile"] == tile
    assert relerr(got, ref) < 1e-12


def test_dgemm_tile_refused_without_pairs(ctx, dev, monkeypatch):
    monkeypatch.setenv("ITX_FORCE_TILE", str(TILE))
    rng = np.random.default_rng(5)
    A = rng.standard_normal((33, 70))
    B = rng.standard_normal((33, 90))
    got, ref, info = _plan_run(ctx, dev, A, (-1, 1), B, (-1, 2), (1, 2))
    assert info["tile"] != TILE
    assert relerr(got, ref) < 1e-12


@pytest.fixture(scope="module")
def H(torch_dev):
    from itensor_amd import host

    return host


@pytest.mark.parametrize("tile", [14, 15])
def test_dgemm_tile_block_sparse_hpsi(H, monkeypatch, tile):
    phi, L, W1, W2, R = hpsi_operands(SECTORS_1024, seed=20261202 + 3)
    ref = local_op_product(phi, L, W1, W2, R)
    monkeypatch.setenv("ITX_FORCE_TILE", str(tile))
    op = H.LocalOp(cten(W1), cten(W2), cten(L), cten(R), True)
    op.set_mode(H.HPSI_PLANNED)
    op.set_precision(0)
    got = op.product(cten(phi))
    assert [(tuple(b), o) for b, o in got.offsets] == [(b, o) for b, o in ref.offsets]
    err = float(np.max(np.abs(got.store() - ref.store)) / np.max(np.abs(ref.store)))
    print(f"[dgemm qn] rel err {err:.2e}", flush=True)
    assert err < 1e-12
