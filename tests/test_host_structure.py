"""CPU tests of the host layer's structure-only QN bookkeeping (no device
memory is touched): the block-pair task list of a contraction against the
oracle's getContractedOffsets restatement (pair order, result blocks and
offsets, first-touch flags, block dims)."""
import numpy as np
import pytest

from oracle import itensor as OI

H_ = pytest.importorskip("itensor_amd.host", reason="host layer not built")


def _cidx(I):
    qns = [] if I.qns is None else [(list(q.q), m) for q, m in I.qns]
    return H_.Index(I.id, I.dim, I.dir, I.plev, ",".join(sorted(I.tags)), qns)


def _rand_index(rng, tag, names):
    n = int(rng.integers(1, 6))
    secs = [(OI.QN({nm: int(rng.integers(-2, 3)) for nm in names}), int(rng.integers(1, 4))) for _ in range(n)]
    return OI.Index(secs, tag, dir=int(rng.choice([-1, 1])))


@pytest.mark.parametrize("seed", range(12))
def test_qn_task_list_matches_oracle(seed):
    rng = np.random.default_rng(100 + seed)
    names = ("Sz",) if seed % 2 else ("Nf", "Sz")
    nc = 1 + seed % 3  # contracted indices
    cs = [_rand_index(rng, f"c{k}", names) for k in range(nc)]
    ua = [_rand_index(rng, f"a{k}", names) for k in range(1 + seed % 2)]
    ub = [_rand_index(rng, f"b{k}", names) for k in range(1 + (seed // 2) % 2)]
    Ainds = ua + cs
    Binds = [c.dag() for c in cs] + ub
    rng.shuffle(Ainds)
    rng.shuffle(Binds)
    def some_div(inds):  # the flux of one random block: never an empty tensor
        return OI.calc_div(inds, tuple(int(rng.integers(0, len(I.qns))) for I in inds))

    divA, divB = some_div(Ainds), some_div(Binds)
    A = OI.qdense(Ainds, divA)
    B = OI.qdense(Binds, divB)
    if not A.offsets or not B.offsets:
        pytest.skip("empty operand for this draw")
    Lind, Rind, _ = OI.compute_labels(A.inds, B.inds)
    Nis, _ = OI.contract_is(A.inds, Lind, B.inds, Rind)
    offs, size, triples = OI.get_contracted_offsets(A, B, Nis)
    got = H_.qn_task_list([_cidx(i) for i in A.inds], [(list(b), int(o)) for b, o in A.offsets],
                          [_cidx(i) for i in B.inds], [(list(b), int(o)) for b, o in B.offsets])
    assert [(i.id, i.plev, i.dir) for i in got["inds"]] == [(i.id, i.plev, i.dir) for i in Nis]
    assert [(tuple(b), o) for b, o in got["coffs"]] == [(tuple(b), int(o)) for b, o in offs]
    assert got["csize"] == size
    aoff = {tuple(b): int(o) for b, o in A.offsets}
    boff = {tuple(b): int(o) for b, o in B.offsets}
    coff = {tuple(b): int(o) for b, o in offs}
    want = []
    for ab, bb, cb in triples:
        want += [aoff[tuple(ab)], boff[tuple(bb)], coff[tuple(cb)]]
    assert list(got["offs"]) == want
    seen, first = set(), []
    for _, _, cb in triples:
        first.append(int(tuple(cb) not in seen))
        seen.add(tuple(cb))
    assert list(got["cfirst"]) == first
    rA, rB = len(A.inds), len(B.inds)
    bda = [A.inds[i].qns[ab[i]][1] for ab, _, _ in triples for i in range(rA)]
    bdb = [B.inds[i].qns[bb[i]][1] for _, bb, _ in triples for i in range(rB)]
    assert list(got["bda"]) == bda and list(got["bdb"]) == bdb
