"""Two-site H*psi operands on the oracle side: the bulk Heisenberg MPO pair
and random environments / wavefunctions, dense or Sz-conserving.

The QN layout is the one of a Heisenberg chain at its midpoint: bond
sectors q = 2 Sz in {-30, ..., +30} (step 2), site {+1: 1, -1: 1}, MPO link
{0: 3, -2: 1, +2: 1} (sample/Heisenberg.h:64-72, zero sector first as
AutoMPO orders it, autompo.cc:1232-1242).  Index arrows follow the DMRG
contraction: phi(dag(l), s1, s2, r), L(l, hl, dag(l)'), R(dag(r), dag(hr), r'),
W(dag(s), s', dag(h_in), h_out).
"""
import numpy as np

from oracle import dmrg as OD
from oracle.itensor import QN, Index, ITensor, dense, qdense

# sector dims of the chi = 1024 bond over 31 Sz sectors (binomial weights)
SECTORS_1024 = [1, 1, 1, 1, 1, 1, 3, 7, 14, 25, 41, 60, 81, 101, 114, 120, 114, 101, 81, 60, 41, 25, 14, 7, 3,
                1, 1, 1, 1, 1, 1]


def scaled_sectors(total: int) -> list[int]:
    """SECTORS_1024 rescaled to sum ~total (floor 1)."""
    return [max(1, int(round(d * total / 1024))) for d in SECTORS_1024]


def _sz(v):
    return QN(("Sz", int(v)))


def heisenberg_bulk_w(s_in: Index, s_out: Index, h_in: Index, h_out: Index, qn: bool) -> ITensor:
    """Bulk Heisenberg W (J = Jz = 1) on (dag(s), s', dag(h_in), h_out)."""
    terms = [("Id", 0, 0, 1.0), ("Id", 1, 1, 1.0), ("Sz", 2, 0, 1.0), ("Sz", 1, 2, 1.0),
             ("Sm", 3, 0, 1.0), ("Sp", 1, 3, 0.5), ("Sp", 4, 0, 1.0), ("Sm", 1, 4, 0.5)]
    W = np.zeros((2, 2, 5, 5))
    for op, r, c, f in terms:
        W[:, :, r, c] += f * OD.spin_half_op(op)
    inds = [s_in.dag(), s_in.prime(), h_in.dag(), h_out]
    return OD._from_dense(inds, W, qn)


def hpsi_operands(sectors: list[int] | None, chi: int = 0, seed: int = 7, tag: str = "c"):
    """(phi, L, W1, W2, R) with N(0,1) entries; QN when sectors is given."""
    rng = np.random.default_rng(seed)
    qn = sectors is not None
    if qn:
        bond = [(_sz(q), d) for q, d in zip(range(-30, 31, 2), sectors)]
        l = Index(bond, f"Link,l=49,{tag}")
        r = Index(bond, f"Link,l=51,{tag}")
        site = [(_sz(1), 1), (_sz(-1), 1)]
        s1, s2 = Index(site, f"Site,n=50,{tag}"), Index(site, f"Site,n=51,{tag}")
        hq = [(_sz(0), 3), (_sz(-2), 1), (_sz(2), 1)]
        hl, hm, hr = Index(hq, f"Link,h=49,{tag}"), Index(hq, f"Link,h=50,{tag}"), Index(hq, f"Link,h=51,{tag}")
    else:
        l, r = Index(chi, f"Link,l=49,{tag}"), Index(chi, f"Link,l=51,{tag}")
        s1, s2 = Index(2, f"Site,n=50,{tag}"), Index(2, f"Site,n=51,{tag}")
        hl, hm, hr = Index(5, f"Link,h=49,{tag}"), Index(5, f"Link,h=50,{tag}"), Index(5, f"Link,h=51,{tag}")

    def rand(inds):
        if qn:
            T = qdense(inds, QN())
            T.store = rng.standard_normal(T.store.size)
            return T
        return dense(inds, rng.standard_normal([i.dim for i in inds]))

    phi = rand([l.dag(), s1, s2, r])
    L = rand([l, hl, l.dag().prime()])
    R = rand([r.dag(), hr.dag(), r.prime()])
    W1 = heisenberg_bulk_w(s1, s1, hl, hm, qn)
    W2 = heisenberg_bulk_w(s2, s2, hm, hr, qn)
    return phi, L, W1, W2, R


def drop_blocks(T: ITensor, every: int) -> ITensor:
    """T without every `every`-th block (a different block structure)."""
    keep, cur, parts = [], 0, []
    for k, (b, o) in enumerate(T.offsets):
        if k % every == every - 1:
            continue
        blk = T.block(b)
        keep.append((b, cur))
        parts.append(np.asarray(blk).ravel(order="F"))
        cur += blk.size
    return ITensor(T.inds, offsets=keep, store=np.concatenate(parts))
