"""GPU parity of LocalOp::product in every execution mode (eager
operator*, planned chain, chain replayed as a HIP graph, per-block-pair
launches) against the oracle's four-stage chain: the Sz-conserving
midpoint layout over 31 sectors (block list bit-exact, values 1e-12) and
the dense layout."""
import numpy as np
import pytest

from oracle.dmrg import local_op_product
from tests.bridge import cten, same_inds
from tests.hpsi_cases import drop_blocks, hpsi_operands, scaled_sectors

pytestmark = pytest.mark.gpu
TOL64 = 1e-12


@pytest.fixture(scope="module")
def H(torch_dev):
    from itensor_amd import host

    return host


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def _modes(H):
    return [("eager", H.HPSI_EAGER), ("planned", H.HPSI_PLANNED), ("graph", H.HPSI_GRAPH),
            ("perblock", H.HPSI_PER_BLOCK)]


@pytest.mark.parametrize("fuse", [False, True])
@pytest.mark.parametrize("total", [141, 300])
def test_qn_hpsi_all_modes_match_oracle(H, fuse, total):
    phi, L, W1, W2, R = hpsi_operands(scaled_sectors(total), seed=total)
    ref = local_op_product(phi, L, W1, W2, R)
    cref = cten(ref)
    outs = {}
    for name, mode in _modes(H):
        op = H.LocalOp(cten(W1), cten(W2), cten(L), cten(R), fuse)
        op.set_mode(mode)
        got = op.product(cten(phi))
        print(f"[qn hpsi] total={total} fuse={fuse} mode={name} fused={op.fused}", flush=True)
        assert same_inds(got, cref), name
        assert [(tuple(x), o) for x, o in got.offsets] == [(x, o) for x, o in ref.offsets], name
        assert _rel(got.store(), ref.store) < TOL64, name
        outs[name] = got.store()
    # the graph replays exactly the kernels the planned chain runs
    assert np.array_equal(outs["graph"], outs["planned"])


def test_qn_hpsi_graph_replays_and_structure_change(H):
    """Repeated products on one LocalOp (new phi data each time, a phi with
    a different block list in between): every result matches the oracle."""
    phi, L, W1, W2, R = hpsi_operands(scaled_sectors(200), seed=3)
    op = H.LocalOp(cten(W1), cten(W2), cten(L), cten(R), True)
    op.set_mode(H.HPSI_GRAPH)
    rng = np.random.default_rng(9)
    for rep in range(5):
        x = phi.copy()
        x.store = rng.standard_normal(x.store.size)
        if rep == 2:
            x = drop_blocks(x, 4)
        ref = local_op_product(x, L, W1, W2, R)
        got = op.product(cten(x))
        print(f"[qn hpsi replay] rep={rep} blocks={x.nnzblocks()}", flush=True)
        assert [(tuple(b), o) for b, o in got.offsets] == [(b, o) for b, o in ref.offsets]
        assert _rel(got.store(), ref.store) < TOL64


def test_qn_hpsi_flops_count_block_pairs(H):
    """LocalOp.flops = sum over the planned block pairs of 2 m n k, equal
    in the grouped and per-block forms, and below the dense count."""
    secs = scaled_sectors(141)
    phi, L, W1, W2, R = hpsi_operands(secs, seed=1)
    f = {}
    for name, mode in (("graph", H.HPSI_GRAPH), ("perblock", H.HPSI_PER_BLOCK)):
        op = H.LocalOp(cten(W1), cten(W2), cten(L), cten(R), False)
        op.set_mode(mode)
        f[name] = op.flops(cten(phi))
    chi = sum(secs)
    assert f["graph"] == f["perblock"] > 0
    assert f["graph"] < 80.0 * chi ** 3 + 800.0 * chi ** 2


@pytest.mark.parametrize("fuse", [False, True])
def test_dense_hpsi_all_modes_match_oracle(H, fuse):
    phi, L, W1, W2, R = hpsi_operands(None, chi=48, seed=2)
    ref = local_op_product(phi, L, W1, W2, R)
    outs = {}
    for name, mode in _modes(H):
        op = H.LocalOp(cten(W1), cten(W2), cten(L), cten(R), fuse)
        op.set_mode(mode)
        got = op.product(cten(phi))
        assert same_inds(got, cten(ref)), name
        assert _rel(got.store(), ref.d.ravel(order="F")) < TOL64, name
        outs[name] = got.store()
    assert np.array_equal(outs["graph"], outs["planned"])
    chi = 48
    op = H.LocalOp(cten(W1), cten(W2), cten(L), cten(R), fuse)
    assert op.flops(cten(phi)) == pytest.approx(80.0 * chi ** 3 + 800.0 * chi ** 2)
