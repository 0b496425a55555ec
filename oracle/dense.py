"""Dense contraction and the dense two-site H*psi chain (oracle).

contract() restates contract(A,ai,B,bi,C,ci,alpha,beta)
(tensor/contract.h:26-33, tensor/contract.cc:843-866) with the label
convention of computeLabels (tensor/contract.h:155-202): negative labels
are contracted, positive labels name the uncontracted indices, and the
result order is the given C labels (default: uncontracted A then B, as
contractIS, indexset_impl.h:79-127).  The permute->GEMM->permute plan of
CProps is the same arithmetic as numpy.tensordot + transpose.

Pinned by tests/test_oracle_dense.py: contract() against explicit loops
over the reference's own cases (unittest/contract_test.cc:161-486, Case 1-4,
NM1-NM5, M1-M4, Regression Test 1); hpsi_dense against a term-by-term sum
of the effective Hamiltonian; each HPSI_STAGES entry against the index
order and values of the ITensor-level chain (oracle.dmrg.local_op_product).
"""
from __future__ import annotations

import numpy as np


def contract(A: np.ndarray, labA, B: np.ndarray, labB, labC=None, alpha=1.0, beta=0.0, C=None):
    labA, labB = list(labA), list(labB)
    ca = [i for i, l in enumerate(labA) if l < 0]
    cb = []
    for i in ca:
        if labB.count(labA[i]) != 1:
            raise ValueError("contracted label must appear once in B")
        cb.append(labB.index(labA[i]))
    for j, l in enumerate(labB):
        if l < 0 and l not in labA:
            raise ValueError("contracted label of B missing in A")
    T = np.tensordot(A, B, axes=(ca, cb))
    order = [l for l in labA if l > 0] + [l for l in labB if l > 0]
    if labC is not None:
        labC = list(labC)
        if sorted(labC) != sorted(order):
            raise ValueError("C labels must be the uncontracted labels")
        T = np.transpose(T, [order.index(l) for l in labC])
    T = alpha * T
    if beta != 0.0:
        T = T + beta * C
    return T


# ---- the two-site effective Hamiltonian product ------------------------

# Index order of the synthetic H*psi operands (names of ITensor indices):
#   psi (l, s1, s2, r)        dims (chi, d, d, chi)
#   L   (l, h0, l')           dims (chi, k, chi)
#   W1  (h0, s1, s1', h1)     dims (k, d, d, k)
#   W2  (h1, s2, s2', h2)     dims (k, d, d, k)
#   R   (r, h2, r')           dims (chi, k, chi)
# Labels per stage are what computeLabels assigns for ITensor's operator*
# (contracted indices numbered in L order, uncontracted ncont+1,... ).
HPSI_STAGES = [
    # phip = phi * L      (localop.h:348)  contracted: l
    ((-1, 2, 3, 4), (-1, 5, 6), (2, 3, 4, 5, 6)),
    # phip *= Op1         (localop.h:352)  T1=(s1,s2,r,h0,l') x W1=(h0,s1,s1',h1): s1,h0
    ((-1, 3, 4, -2, 5), (-2, -1, 6, 7), (3, 4, 5, 6, 7)),
    # phip *= Op2         (localop.h:353)  T2=(s2,r,l',s1',h1) x W2=(h1,s2,s2',h2): s2,h1
    ((-1, 3, 4, 5, -2), (-2, -1, 6, 7), (3, 4, 5, 6, 7)),
    # phip *= R           (localop.h:362)  T3=(r,l',s1',s2',h2) x R=(r,h2,r'): r,h2
    ((-1, 3, 4, 5, -2), (-1, -2, 6), (3, 4, 5, 6)),
]


def hpsi_dense(psi, L, W1, W2, R):
    """H*psi = noPrime(((psi*L)*W1)*W2*R)  (mps/localop.h:324-365, bulk case).

    Returns the result with indices (l, s1, s2, r) after noPrime.
    """
    x = psi
    for (la, lb, lc), op in zip(HPSI_STAGES, (L, W1, W2, R)):
        x = contract(x, la, op, lb, lc)
    return x


def hpsi_stage_shapes(chi: int, d: int = 2, k: int = 5):
    """Operand dims of the four stages (A dims, B dims, C dims)."""
    psi = (chi, d, d, chi)
    Ld = (chi, k, chi)
    W = (k, d, d, k)
    T1 = (d, d, chi, k, chi)
    T2 = (d, chi, chi, d, k)
    T3 = (chi, chi, d, d, k)
    out = (chi, d, d, chi)
    return [(psi, Ld, T1), (T1, W, T2), (T2, W, T3), (T3, Ld, out)]


def hpsi_flops(chi: int, d: int = 2, k: int = 5) -> float:
    """Algorithmic flops of one dense H*psi: 4 d^2 k chi^3 + 4 d^3 k^2 chi^2."""
    return 4.0 * d * d * k * chi ** 3 + 4.0 * d ** 3 * k * k * chi ** 2


def heisenberg_w(J: float = 1.0, Jz: float | None = None) -> np.ndarray:
    """Bulk Heisenberg MPO tensor W(h_in, s, s', h_out), dims (5,2,2,5).

    Follows sample/Heisenberg.h:84-94 (row = incoming link, col = outgoing,
    1-based states 1..5): (1,1)=Id, (2,2)=Id, (3,1)=Sz, (2,3)=Jz*Sz,
    (4,1)=S-, (2,4)=J/2*S+, (5,1)=S+, (2,5)=J/2*S-.  Site basis: 0=Up, 1=Dn
    (sites/spinhalf.h), operator element op(s, s') = <s'|O|s> as ITensor
    stores op(dag(s), prime(s)).
    """
    if Jz is None:
        Jz = J
    Id = np.eye(2)
    Sz = np.diag([0.5, -0.5])
    # <s'|S+|s>: S+ |Dn> = |Up>  -> element (s=Dn, s'=Up)
    Sp = np.zeros((2, 2)); Sp[1, 0] = 1.0
    Sm = np.zeros((2, 2)); Sm[0, 1] = 1.0
    W = np.zeros((5, 2, 2, 5))
    W[0, :, :, 0] = Id
    W[1, :, :, 1] = Id
    W[2, :, :, 0] = Sz
    W[1, :, :, 2] = Jz * Sz
    W[3, :, :, 0] = Sm
    W[1, :, :, 3] = J / 2 * Sp
    W[4, :, :, 0] = Sp
    W[1, :, :, 4] = J / 2 * Sm
    return W
