"""CPU: pin the oracle's ITensor layer (contraction, QN block bookkeeping,
+=, combiners) against brute force and the reference's known answers."""
import itertools

import numpy as np
import pytest

from oracle.itensor import (QN, Index, In, Out, ITensor, add_into, combiner, contract, dense, get_block_offsets,
                            norm, qdense, remove_qns)


def brute_contract(A: ITensor, B: ITensor):
    """Explicit loops: C[u] = sum_c A[..] B[..] by index identity."""
    Ad, Bd = A.to_dense_array(), B.to_dense_array()
    common = [I for I in A.inds if I in B.inds]
    ua = [I for I in A.inds if I not in B.inds]
    ub = [I for I in B.inds if I not in A.inds]
    out = np.zeros([I.dim for I in ua + ub])
    for uvals in itertools.product(*[range(I.dim) for I in ua + ub]):
        s = 0.0
        for cvals in itertools.product(*[range(I.dim) for I in common]):
            env = dict(zip(ua + ub, uvals))
            env.update(zip(common, cvals))
            s += Ad[tuple(env[I] for I in A.inds)] * Bd[tuple(env[I] for I in B.inds)]
        out[uvals] = s
    return out, ua + ub


def test_dense_contraction_brute_force():
    rng = np.random.default_rng(0)
    i, j, k, l, m = (Index(d, t) for d, t in [(2, "i"), (3, "j"), (4, "k"), (2, "l"), (3, "m")])
    cases = [([i, j], [j, k]), ([i, j, k], [k, i, l]), ([k, i, j, l], [j, m, l]), ([i, j], [k, l]),
             ([i, j, k], [k, j, i])]
    for ia, ib in cases:
        A = dense(ia, rng.standard_normal([x.dim for x in ia]))
        B = dense(ib, rng.standard_normal([x.dim for x in ib]))
        C = contract(A, B)
        ref, order = brute_contract(A, B)
        assert C.inds == order  # contractIS: uncontracted L then R, unsorted
        assert np.allclose(C.d, ref, atol=1e-13)


def test_get_block_offsets_flux_rule_and_order():
    i = Index([(QN(0), 2), (QN(1), 3)], "i")
    j = Index([(QN(0), 1), (QN(1), 2), (QN(2), 2)], "j")
    offs, tot = get_block_offsets([i, j.dag()], QN(0))
    # blocks (bi,bj) with q_i - q_j == 0, first index fastest, offsets = running sizes
    assert [b for b, _ in offs] == [(0, 0), (1, 1)]
    assert [o for _, o in offs] == [0, 2]
    assert tot == 2 * 1 + 3 * 2


def test_block_deficient_known_answer():
    """itensor_test.cc:2734-2778: A = a*prime(dag(a),i) has 4 blocks / 64 nnz,
    widening to 5/68, 6/84, 7/120 when elements of new blocks are set."""
    rng = np.random.default_rng(1)
    i = Index([(QN(0), 2), (QN(1), 3), (QN(2), 4), (QN(1), 5), (QN(3), 6)], "i")
    l = Index([(QN(0), 3)], "l")
    a = qdense([i, l], QN(1), rng=rng)
    A = contract(a, a.dag().prime(1, i))
    assert A.nnzblocks() == 4 and A.nnz() == 64
    ip = i.prime()
    for (ival, ipval), (nb, nz) in zip([(2, 1), (7, 8), (16, 18)], [(5, 68), (6, 84), (7, 120)]):
        # copyA.set(i=ival, ip=ipval, v): add a one-element tensor of the same flux
        bi = _block_of(i, ival - 1)
        bj = _block_of(i, ipval - 1)
        one = ITensor([i, ip.dag()], offsets=[((bi, bj), 0)],
                      store=np.zeros(i.blocksize(bi) * i.blocksize(bj)))
        A = add_into(A, one, 1.0)
        assert A.nnzblocks() == nb and A.nnz() == nz


def _block_of(I, v):
    s = 0
    for b in range(I.nblock):
        if v < s + I.blocksize(b):
            return b
        s += I.blocksize(b)
    raise ValueError


def test_block_deficient_add_commutes():
    rng = np.random.default_rng(2)
    i = Index([(QN(0), 2), (QN(1), 3), (QN(2), 4), (QN(1), 5), (QN(3), 6)], "i")
    l = Index([(QN(0), 3)], "l")
    a = qdense([i, l], QN(1), rng=rng)
    A = contract(a, a.dag().prime(1, i))
    T = qdense([i.prime().dag(), i], QN(0), rng=rng)
    AT = A + T
    TA = T + A
    assert norm(AT - TA) == 0


def test_contraction_with_no_output_blocks():
    """itensor_test.cc:2849-2865."""
    s = Index([(QN(("Sz", 1)), 1), (QN(("Sz", -1)), 1)], "n=10,Site,S=1/2")
    lA = Index([(QN(("Sz", 1)), 1)], "l=9,Link")
    lB = Index([(QN(("Sz", -1)), 1)], "l=9,Link")
    A = qdense([lA.dag(), s], QN(("Sz", 0)), fill=1.0)
    B = qdense([lB.dag(), s], QN(("Sz", 0)), fill=1.0)
    C = contract(A, B.dag())
    assert C.nnz() == 0 and C.nnzblocks() == 0


@pytest.mark.parametrize("seed", range(6))
def test_qn_contraction_matches_dense_expansion(seed):
    rng = np.random.default_rng(seed)
    def rind(tag):
        nb = int(rng.integers(1, 4))
        return Index([(QN(int(q)), int(rng.integers(1, 4))) for q in rng.integers(-2, 3, size=nb)], tag)
    a, b, c, d = rind("a"), rind("b"), rind("c"), rind("d")
    A = qdense([a, b.dag(), c], QN(int(rng.integers(-1, 2))), rng=rng)
    B = qdense([c.dag(), d, b], QN(int(rng.integers(-1, 2))), rng=rng)
    C = contract(A, B)
    ref, order = brute_contract(A, B)
    assert C.inds == order
    assert np.allclose(C.to_dense_array(), ref, atol=1e-12)
    # block pattern: every stored block is reachable, sorted reverse-lex,
    # offsets are running sums
    keys = [tuple(reversed(bk)) for bk, _ in C.offsets]
    assert keys == sorted(keys) and len(set(keys)) == len(keys)
    run = 0
    for bk, off in C.offsets:
        assert off == run
        run += int(np.prod([I.blocksize(x) for I, x in zip(C.inds, bk)]))


def test_arrow_check():
    i = Index([(QN(0), 1), (QN(1), 1)], "i")
    A = qdense([i, i.prime()], QN(1))
    with pytest.raises(ValueError):
        contract(A, qdense([i, i.prime(2)], QN(1)))


def test_dense_combiner_round_trip():
    rng = np.random.default_rng(4)
    i, j, k = Index(2, "i"), Index(3, "j"), Index(4, "k")
    T = dense([i, j, k], rng.standard_normal((2, 3, 4)))
    for sel in ([i, j], [j, k], [k, i], [j]):
        C, c = combiner(sel)
        Tc = C * T
        assert c in Tc.inds
        back = C * Tc
        assert sorted(x.id for x in back.inds) == sorted(x.id for x in T.inds)
        # compare by index identity
        perm = [back.inds.index(x) for x in T.inds]
        assert np.allclose(np.transpose(back.d, perm), T.d)


def test_qn_combiner_sector_order_and_round_trip():
    rng = np.random.default_rng(5)
    l = Index([(QN(0), 2), (QN(1), 3), (QN(-1), 1)], "l")
    s = Index([(QN(1), 1), (QN(-1), 1)], "s")
    r = Index([(QN(0), 2), (QN(1), 2)], "r")
    T = qdense([l, s, r.dag()], QN(0), rng=rng)
    C, c = combiner([l, s])
    # combined sectors in order of first appearance over (l-block, s-block),
    # l fastest: (0,0)->1, (1,0)->2, (2,0)->0, (0,1)->-1, (1,1)->0, (2,1)->-2
    assert [q.val("") for q, _ in c.qns] == [1, 2, 0, -1, -2]
    assert [m for _, m in c.qns] == [2, 3, 1 + 3, 2, 1]
    Tc = C * T
    assert Tc.inds[0] == c
    back = C.dag() * Tc
    perm = [back.inds.index(x) for x in T.inds]
    assert np.allclose(np.transpose(back.to_dense_array(), perm), T.to_dense_array())
    # combined data equals the dense combine of the expanded tensor, up to the
    # sector permutation of the combined index
    assert np.isclose(norm(Tc), norm(T))
