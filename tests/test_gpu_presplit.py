"""GPU: pre-split BF16x3 operands (itx_split_bf16, itx_plan_set_split).

* The split words equal a numpy restatement of the split the BF16x3
  kernels form at tile load: x rounded to FP32, hi = bf16(x) and
  lo = bf16(x - hi), each rounded to nearest even.
* A BF16x3 plan reading pre-split operands gives the plan's result without
  pre-splitting bit for bit (same terms, same MFMA order), through the dense
  H*psi chain (psi*L and *R pre-split, the MPO stage in FP64) at chi where
  the 16-byte quad loads apply (chi = 256, 1024) and where they do not
  (chi = 250), and with every MFMA tile forced.
"""
import os

import numpy as np
import pytest

from itensor_amd.workloads import FUSED, fused_dims
from oracle.dense import heisenberg_w, hpsi_dense

pytestmark = pytest.mark.gpu


def _bf16_rne(u):
    u = u.astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32) & 0xFFFF


def split_ref(x):
    f = x.astype(np.float32)
    h = _bf16_rne(f.view(np.uint32))
    rem = f - (h.astype(np.uint32) << 16).view(np.float32)
    lo = _bf16_rne(rem.view(np.uint32))
    return h | (lo << 16)


def test_split_words_match_the_restated_rounding(ctx):
    import torch

    rng = np.random.default_rng(5)
    x = np.concatenate([rng.standard_normal(100000) * 10.0 ** rng.integers(-30, 30, 100000),
                        [0.0, -0.0, 1.0, -1.5, 3.0e38, 65504.0, 1.0 / 3.0]])
    dx = torch.from_numpy(x).to("cuda:0")
    dy = torch.zeros(x.size, dtype=torch.int32, device="cuda:0")
    ctx.split_bf16(dx.data_ptr(), dy.data_ptr(), x.size)
    torch.cuda.synchronize()
    got = dy.cpu().numpy().view(np.uint32)
    assert np.array_equal(got, split_ref(x))


def _chain(ctx, chi, split, tile=None):
    import torch

    rng = np.random.default_rng(4400 + chi)
    psi = rng.standard_normal((chi, 2, 2, chi))
    L = rng.standard_normal((chi, 5, chi))
    R = rng.standard_normal((chi, 5, chi))
    W = heisenberg_w()
    W12 = np.einsum("aijb,bklc->aijklc", W, W)
    dims, labels = fused_dims(chi), FUSED
    forms = ["bf16x3", "f64", "bf16x3"]
    old = os.environ.get("ITX_FORCE_TILE")
    if tile is not None:
        os.environ["ITX_FORCE_TILE"] = str(tile)
    try:
        plans = [ctx.plan_dense(da, la, db, lb, dc, lc, f)
                 for (da, db, dc), (la, lb, lc), f in zip(dims, labels, forms)]
    finally:
        if tile is not None:
            if old is None:
                del os.environ["ITX_FORCE_TILE"]
            else:
                os.environ["ITX_FORCE_TILE"] = old
    dev = torch.device("cuda:0")

    def put(a):
        return torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).to(dev)

    def words(t):
        w = torch.empty(t.numel(), dtype=torch.int32, device=dev)
        ctx.split_bf16(t.data_ptr(), w.data_ptr(), t.numel())
        return w

    ops = [put(L), put(W12), put(R)]
    x = put(psi)
    try:
        for i, (p, op, (_, _, dc)) in enumerate(zip(plans, ops, dims)):
            y = torch.empty(int(np.prod(dc)), dtype=torch.float64, device=dev)
            if split and i != 1:
                p.set_split(True)
                xa, xb = words(x), words(op)
                p.execute(xa.data_ptr(), xb.data_ptr(), y.data_ptr())
            else:
                p.execute(x.data_ptr(), op.data_ptr(), y.data_ptr())
            x = y
        torch.cuda.synchronize()
        out = x.cpu().numpy().reshape((chi, 2, 2, chi), order="F")
        infos = [p.info for p in plans]
    finally:
        for p in plans:
            p.close()
    return out, infos, (psi, L, W, R)


@pytest.mark.parametrize("chi", [250, 256, 1024])
def test_presplit_chain_equals_plain_bf16x3(ctx, chi):
    plain, infos, ops = _chain(ctx, chi, False)
    pre, _, _ = _chain(ctx, chi, True)
    print(f"[presplit] chi={chi} tiles={[i['tile'] for i in infos]}", flush=True)
    assert np.array_equal(plain, pre)
    ref = hpsi_dense(ops[0], ops[1], ops[2], ops[2], ops[3])
    assert np.max(np.abs(pre - ref)) / np.max(np.abs(ref)) < 1e-4


@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4, 6, 7, 8, 9, 11, 12, 13])
def test_presplit_every_tile(ctx, tile):
    plain, infos, _ = _chain(ctx, 256, False, tile)
    pre, _, _ = _chain(ctx, 256, True, tile)
    assert infos[0]["tile"] == tile
    assert np.array_equal(plain, pre)


def test_set_split_refuses_other_plans(ctx):
    from itensor_amd.itx import ItxError

    p = ctx.plan_dense((8, 9), (1, -1), (9, 7), (-1, 2), (8, 7), (1, 2), "f64")
    try:
        with pytest.raises(ItxError, match="pre-split"):
            p.set_split(True)
    finally:
        p.close()
