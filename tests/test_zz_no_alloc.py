"""GPU: itx_contract_qn_noalloc, the block-sparse contraction that allocates
no device memory (its task list, offset tables and split-K partial tiles go
into the library's static scratch array), against itx_contract_qn on the same
inputs -- bit for bit, since both run the same kernels on the same plan --
and against the oracle's block-sparse contraction."""
import numpy as np
import pytest

from oracle import itensor as OI
from oracle.itensor import QN, Index, qdense

pytestmark = pytest.mark.gpu
TOL = {"f64": 1e-12, "f32": 2e-5, "bf16x3": 1e-4}


def _index(tag, sizes, scale, dirn=1):
    qs = range(len(sizes) // 2, -(len(sizes) // 2) - 1, -1)
    return Index([(QN(q), max(1, d * scale)) for q, d in zip(qs, sizes)], tag, dir=dirn)


def _task_list(A, B):
    """Labels, C block layout and the block-pair task list of A*B (the
    reference's getContractedOffsets, restated by the oracle)."""
    labA, labB, _ = OI.compute_labels(A.inds, B.inds)
    Cis, labC = OI.contract_is(A.inds, labA, B.inds, labB)
    coffs, csize, triples = OI.get_contracted_offsets(A, B, Cis)
    aoff = {tuple(b): int(o) for b, o in A.offsets}
    boff = {tuple(b): int(o) for b, o in B.offsets}
    coff = {tuple(b): int(o) for b, o in coffs}
    offs, bda, bdb = [], [], []
    for ab, bb, cb in triples:
        offs += [aoff[tuple(ab)], boff[tuple(bb)], coff[tuple(cb)]]
        bda += [I.qns[b][1] for I, b in zip(A.inds, ab)]
        bdb += [I.qns[b][1] for I, b in zip(B.inds, bb)]
    return labA, labB, labC, offs, bda, bdb, csize


def _three_site_case(seed, scale):
    """T(L, s, R) * W(R, t, M): the QN link-site-link contractions of an MPS
    bond, sector dims proportional to `scale`."""
    rng = np.random.default_rng(seed)
    L = _index("L", (3, 5, 7, 5, 2), scale)
    s = Index([(QN(1), 1), (QN(-1), 1)], "s")
    R = _index("R", (2, 4, 6, 4, 3), scale)
    t = Index([(QN(1), 1), (QN(-1), 1)], "t")
    M = _index("M", (1, 3, 5, 3, 1), scale)
    A = qdense([L, s, R.dag()], QN(), rng=rng)
    B = qdense([R, t, M.dag()], QN(), rng=rng)
    return A, B


def _run_both(ctx, A, B, prec):
    import torch

    labA, labB, labC, offs, bda, bdb, csize = _task_list(A, B)
    dev = torch.device("cuda:0")
    dA = torch.from_numpy(np.ascontiguousarray(A.store)).to(dev)
    dB = torch.from_numpy(np.ascontiguousarray(B.store)).to(dev)
    out = []
    for noalloc in (False, True):
        dC = torch.full((max(1, csize),), float("nan"), dtype=torch.float64, device=dev)
        ctx.contract_qn(dA.data_ptr(), labA, dB.data_ptr(), labB, dC.data_ptr(), labC, offs, bda, bdb,
                        1.0, prec, noalloc=noalloc)
        out.append(dC)
    torch.cuda.synchronize()
    ref = (A * B).store
    return out[0], out[1], ref


def _check(name, ref_t, got_t, ref, prec):
    import torch

    got = got_t.cpu().numpy()[: ref.size]
    err = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    print(f"[noalloc] {name} form={prec} elements={ref.size} rel err vs oracle {err:.2e}", flush=True)
    assert torch.equal(ref_t, got_t), name
    assert err < TOL[prec], name


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16x3"])
def test_noalloc_matches_contract_qn_two_calls(ctx, prec):
    """Two calls in a row on one context, the second smaller than the first
    (the scratch array keeps the first call's tables): each result equals
    itx_contract_qn's bit for bit and the oracle's to the form's tolerance."""
    for step, scale in enumerate((20, 2)):
        A, B = _three_site_case(700 + step, scale)
        ref_t, got_t, ref = _run_both(ctx, A, B, prec)
        _check(f"call {step + 1} (scale {scale})", ref_t, got_t, ref, prec)


def test_noalloc_split_plan(ctx):
    """Few small C blocks with a long contracted extent: the plan splits the
    k chunks over workgroups, whose partial tiles sit in the scratch array,
    and a reduction pass writes C."""
    rng = np.random.default_rng(711)
    a = Index([(QN(1), 12), (QN(-1), 9)], "a")
    k = Index([(QN(0), 2600), (QN(2), 1900)], "k")
    b = Index([(QN(1), 7), (QN(-1), 11)], "b")
    A = qdense([a, k.dag()], QN(1), rng=rng)
    B = qdense([k, b.dag()], QN(-1), rng=rng)
    ref_t, got_t, ref = _run_both(ctx, A, B, "f64")
    _check("split", ref_t, got_t, ref, "f64")


def test_noalloc_rejects_plans_beyond_the_scratch(ctx):
    """A plan larger than the 32 MiB scratch array (a 2 x 2 result over a
    contracted extent of 3e6: offset tables and split partial tiles far
    beyond it) fails with ITX_ERR_UNSUPPORTED and a message naming the
    sizes; the context keeps working."""
    import torch

    from itensor_amd.itx import ItxError

    K = 3_000_000
    labA, labB, labC = (1, -1), (-1, 2), (1, 2)
    dev = torch.device("cuda:0")
    dA = torch.zeros(2 * K, dtype=torch.float64, device=dev)
    dB = torch.zeros(2 * K, dtype=torch.float64, device=dev)
    dC = torch.zeros(4, dtype=torch.float64, device=dev)
    with pytest.raises(ItxError, match="scratch") as ei:
        ctx.contract_qn(dA.data_ptr(), labA, dB.data_ptr(), labB, dC.data_ptr(), labC, [0, 0, 0], [2, K], [K, 2],
                        1.0, "f64", noalloc=True)
    assert "(4)" in str(ei.value)  # ITX_ERR_UNSUPPORTED
    A, B = _three_site_case(712, 1)
    ref_t, got_t, ref = _run_both(ctx, A, B, "f64")
    _check("after the rejected call", ref_t, got_t, ref, "f64")
