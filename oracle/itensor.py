"""ITensor semantics on the CPU (oracle): Index, QN, dense and block-sparse
(QDense) ITensors, contraction, +=, norm, combiners.

Restates (file:line in the reference, ITensor v3):
  Index equality = id + tags + prime level ........ index.cc:154-157
  QN arithmetic (named U(1) charges) ............... qn.h:37-180
  computeLabels (contracted = negative) ............ tensor/contract.h:155-202
  contractIS result order (L then R, unsorted) ..... indexset_impl.h:79-127
  operator*= (scalar cases, arrows, removeQNs) ..... itensor.cc:935-981
  Dense contraction ................................ itdata/dense.cc:262-326
  getBlockOffsets (flux rule, block order) ......... itdata/qdense.cc:133-173
  reverse-lexicographic Block order ................ itdata/qdense.cc:60-65
  getContractedOffsets + block loop ................ itdata/qutil.h:93-269
  QDense contraction (first-touch beta) ............ itdata/qdense.cc:671-743
  PlusEQ dense / QDense (block-set union) .......... itdata/dense.cc:362-415,
                                                     itdata/qdense.cc:506-668
  combiner (dense, QN sector first appearance) ..... itensor.cc:1468-1560,
                                                     itdata/combiner.cc:55-150,
                                                     itdata/qcombiner.cc:125-300
Pinned by tests/test_oracle_itensor.py (brute force on tiny tensors, the
reference's QN known answers itensor_test.cc:2734-2865, and dense expansion
of every QN result).
"""
from __future__ import annotations

import itertools
import numpy as np

Out, In = 1, -1
_ids = itertools.count(1)


# ---------------------------------------------------------------- QN ----
class QN:
    """Up to 4 named U(1) charges (qn.h:37-180).  Missing names count as 0,
    so QN(("Sz", 0)) == QN().  The unnamed charge of QN(int) has name ""."""

    __slots__ = ("q",)

    def __init__(self, *args):
        d = {}
        if len(args) == 1 and isinstance(args[0], (int, np.integer)):
            d[""] = int(args[0])
        elif len(args) == 1 and isinstance(args[0], dict):
            d = dict(args[0])
        else:
            for n, v in args:
                d[n] = d.get(n, 0) + int(v)
        self.q = tuple(sorted((n, v) for n, v in d.items() if v != 0))

    def _d(self):
        return dict(self.q)

    def __add__(self, o):
        d = self._d()
        for n, v in o.q:
            d[n] = d.get(n, 0) + v
        return QN(d)

    def __sub__(self, o):
        return self + (-o)

    def __neg__(self):
        return QN({n: -v for n, v in self.q})

    def __mul__(self, arrow: int):
        return self if arrow == 1 else -self

    __rmul__ = __mul__

    def __eq__(self, o):
        return isinstance(o, QN) and self.q == o.q

    def __hash__(self):
        return hash(self.q)

    def __repr__(self):
        return "QN(" + ",".join(f"{n}:{v}" for n, v in self.q) + ")"

    def val(self, name: str) -> int:
        return self._d().get(name, 0)


# ------------------------------------------------------------- Index ----
class Index:
    """Index{id, dim, dir, tags, prime level, QN sectors} (index.h:78-282).

    Ids come from a counter (the reference draws random 64-bit ids,
    index.h:31-46; only identity matters)."""

    __slots__ = ("id", "dim", "tags", "plev", "dir", "qns")

    def __init__(self, spec, tags: str = "", dir: int = Out, *, plev: int = 0, id=None):
        if isinstance(spec, (int, np.integer)):
            self.qns = None
            self.dim = int(spec)
        else:
            self.qns = tuple((q, int(m)) for q, m in spec)
            self.dim = sum(m for _, m in self.qns)
        self.tags = frozenset(t for t in tags.split(",") if t) if isinstance(tags, str) else frozenset(tags)
        self.plev = plev
        self.dir = dir
        self.id = next(_ids) if id is None else id

    def _copy(self, **kw):
        n = object.__new__(Index)
        for s in Index.__slots__:
            setattr(n, s, kw.get(s, getattr(self, s)))
        return n

    def __eq__(self, o):
        return isinstance(o, Index) and self.id == o.id and self.plev == o.plev and self.tags == o.tags

    def __hash__(self):
        return hash((self.id, self.plev, self.tags))

    def __repr__(self):
        t = ",".join(sorted(self.tags))
        return f"({self.dim}|{t}|id{self.id}{chr(39) * self.plev}{'' if self.qns is None else ('>' if self.dir == Out else '<')})"

    def prime(self, inc: int = 1) -> "Index":
        return self._copy(plev=self.plev + inc)

    def noprime(self) -> "Index":
        return self._copy(plev=0)

    def setprime(self, p: int) -> "Index":
        return self._copy(plev=p)

    def dag(self) -> "Index":
        return self._copy(dir=-self.dir)

    def settags(self, tags) -> "Index":
        ts = frozenset(t for t in tags.split(",") if t) if isinstance(tags, str) else frozenset(tags)
        return self._copy(tags=ts)

    def has_qns(self) -> bool:
        return self.qns is not None

    @property
    def nblock(self) -> int:
        return len(self.qns) if self.qns is not None else 1

    def blocksize(self, b: int) -> int:
        return self.qns[b][1]

    def qn(self, b: int) -> QN:
        return self.qns[b][0]

    def block_start(self, b: int) -> int:
        return sum(m for _, m in self.qns[:b])

    def remove_qns(self) -> "Index":
        return self._copy(qns=None)


def dims_of(inds):
    return tuple(i.dim for i in inds)


def _block_less(a, b):
    """Block order: lexicographic compare from the LAST index (qdense.cc:60-65)."""
    return tuple(reversed(a)) < tuple(reversed(b))


def _block_key(b):
    return tuple(reversed(b))


# ----------------------------------------------------------- ITensor ----
class ITensor:
    """Dense (self.d: ndarray with shape = dims, logical index order) or
    block-sparse (self.offsets: list of (block, offset) sorted in block
    order, self.store: flat column-major block storage), or a combiner
    (self.cmb set)."""

    def __init__(self, inds, d=None, offsets=None, store=None, cmb=None):
        self.inds = list(inds)
        self.d = d
        self.offsets = offsets
        self.store = store
        self.cmb = cmb

    # -- kinds
    @property
    def is_qn(self) -> bool:
        return self.offsets is not None

    @property
    def order(self) -> int:
        return len(self.inds)

    def copy(self) -> "ITensor":
        return ITensor(list(self.inds),
                       None if self.d is None else self.d.copy(),
                       None if self.offsets is None else list(self.offsets),
                       None if self.store is None else self.store.copy(),
                       self.cmb)

    # -- blocks
    def block_dims(self, block):
        return tuple(I.blocksize(b) for I, b in zip(self.inds, block))

    def _bmap(self):
        m = getattr(self, "_bm", None)
        if m is None or m[0] is not self.offsets:
            m = (self.offsets, {b: (k, off) for k, (b, off) in enumerate(self.offsets)})
            self._bm = m
        return m[1]

    def block(self, block):
        """View of a block (column-major), or None if absent (the lookup of
        offsetOf, qdense.cc:214-221)."""
        hit = self._bmap().get(tuple(block))
        if hit is None:
            return None
        bd = self.block_dims(block)
        n = int(np.prod(bd))
        return self.store[hit[1]:hit[1] + n].reshape(bd, order="F")

    def block_loc(self, block) -> int:
        return self._bmap()[tuple(block)][0]

    def nnz(self) -> int:
        return int(self.store.size) if self.is_qn else int(self.d.size)

    def nnzblocks(self) -> int:
        return len(self.offsets) if self.is_qn else 1

    def to_dense_array(self) -> np.ndarray:
        """removeQNs data (qdense.cc RemoveQNs, :921-952)."""
        if not self.is_qn:
            return self.d
        out = np.zeros(dims_of(self.inds))
        for b, _ in self.offsets:
            sl = tuple(slice(I.block_start(bb), I.block_start(bb) + I.blocksize(bb))
                       for I, bb in zip(self.inds, b))
            out[sl] = self.block(b)
        return out

    def div(self) -> QN:
        """calcDiv of the first block (qdense.cc:86-108)."""
        if not self.is_qn or not self.offsets:
            return QN()
        return calc_div(self.inds, self.offsets[0][0])

    # -- index ops (return new tensors sharing data where harmless)
    def prime(self, inc=1, *which) -> "ITensor":
        t = self.copy()
        t.inds = [i.prime(inc) if (not which or i in which) else i for i in t.inds]
        return t

    def noprime(self) -> "ITensor":
        t = self.copy()
        t.inds = [i.noprime() for i in t.inds]
        return t

    def dag(self) -> "ITensor":
        t = self.copy()
        t.inds = [i.dag() for i in t.inds]
        return t

    def __mul__(self, o):
        if isinstance(o, ITensor):
            return contract(self, o)
        t = self.copy()
        if t.is_qn:
            t.store *= o
        else:
            t.d = t.d * o
        return t

    __rmul__ = __mul__

    def __add__(self, o):
        t = self.copy()
        return add_into(t, o, 1.0)

    def __sub__(self, o):
        t = self.copy()
        return add_into(t, o, -1.0)

    def elt(self):
        if self.order != 0:
            raise ValueError("elt of non-scalar")
        return float(self.store[0]) if self.is_qn else float(np.asarray(self.d).reshape(-1)[0])


def calc_div(inds, block) -> QN:
    q = QN()
    for I, b in zip(inds, block):
        q = q + I.qn(b) * I.dir
    return q


def has_qns(inds) -> bool:
    return any(i.has_qns() for i in inds)


def get_block_offsets(inds, div: QN):
    """getBlockOffsets (qdense.cc:133-173): all blocks, first index fastest,
    kept iff sum dir*qn == div; offsets are running sums of block sizes."""
    if not inds:
        return [((), 0)], 1
    offs = []
    tot = 0
    ranges = [range(I.nblock) for I in inds]
    for rev in itertools.product(*reversed(ranges)):
        blk = tuple(reversed(rev))
        if calc_div(inds, blk) == div:
            offs.append((blk, tot))
            tot += int(np.prod([I.blocksize(b) for I, b in zip(inds, blk)]))
    return offs, tot


def qdense(inds, div: QN = QN(), fill=None, rng=None) -> ITensor:
    offs, tot = get_block_offsets(inds, div)
    if rng is not None:
        store = rng.standard_normal(tot)
    else:
        store = np.zeros(tot) if fill is None else np.full(tot, float(fill))
    return ITensor(inds, offsets=offs, store=store)


def dense(inds, d) -> ITensor:
    d = np.asarray(d, dtype=np.float64)
    assert d.shape == dims_of(inds)
    return ITensor(inds, d=d.copy())


def remove_qns(T: ITensor) -> ITensor:
    return ITensor([i.remove_qns() for i in T.inds], d=T.to_dense_array().copy())


def index_position(inds, I) -> int:
    for k, J in enumerate(inds):
        if J == I:
            return k
    return -1


def compute_labels(Lis, Ris):
    """computeLabels (tensor/contract.h:155-202)."""
    Lind = [0] * len(Lis)
    Rind = [0] * len(Ris)
    ncont = 0
    for i, I in enumerate(Lis):
        for j, J in enumerate(Ris):
            if I == J:
                Lind[i] = -(1 + ncont)
                Rind[j] = -(1 + ncont)
                ncont += 1
                break
    u = ncont
    for j in range(len(Lis)):
        if Lind[j] == 0:
            u += 1
            Lind[j] = u
    for j in range(len(Ris)):
        if Rind[j] == 0:
            u += 1
            Rind[j] = u
    return Lind, Rind, ncont


def contract_is(Lis, Lind, Ris, Rind):
    """contractIS without sorting: uncontracted L then R (indexset_impl.h:79-127)."""
    Nis = [I for I, l in zip(Lis, Lind) if l > 0] + [J for J, l in zip(Ris, Rind) if l > 0]
    Cind = [l for l in Lind if l > 0] + [l for l in Rind if l > 0]
    return Nis, Cind


def _block_contract(a, la, b, lb, lc):
    ca = [i for i, l in enumerate(la) if l < 0]
    cb = [lb.index(la[i]) for i in ca]
    T = np.tensordot(a, b, axes=(ca, cb))
    order = [l for l in la if l > 0] + [l for l in lb if l > 0]
    if list(lc) != order:
        T = np.transpose(T, [order.index(l) for l in lc])
    return T


def check_arrows(Lis, Ris):
    """checkArrows (itensor.cc:891-920): contracted QN indices must have
    opposite arrows."""
    if has_qns(Lis) and has_qns(Ris):
        for I in Lis:
            for J in Ris:
                if I == J and I.dir == J.dir:
                    raise ValueError(f"Mismatched QN Index arrows {I} {J}")


def get_contracted_offsets(A: ITensor, B: ITensor, Cis):
    """getContractedOffsets (qutil.h:93-242): block triples in (A block,
    B block) loop order, C blocks sorted (reverse-lex), unique, prefix-sum
    offsets."""
    Ais, Bis = A.inds, B.inds
    rA, rB = len(Ais), len(Bis)
    AtoB = [-1] * rA
    AtoC = [-1] * rA
    BtoC = [-1] * rB
    for ic, I in enumerate(Cis):
        j = index_position(Ais, I)
        if j >= 0:
            AtoC[j] = ic
        else:
            BtoC[index_position(Bis, I)] = ic
    for ia in range(rA):
        for ib in range(rB):
            if Ais[ia] == Bis[ib]:
                AtoB[ia] = ib
                break
    triples = []
    csizes = []
    for ablk, _ in A.offsets:
        cblk = [0] * len(Cis)
        for ia in range(rA):
            if AtoC[ia] != -1:
                cblk[AtoC[ia]] = ablk[ia]
        for bblk, _ in B.offsets:
            if any(AtoB[ia] != -1 and ablk[ia] != bblk[AtoB[ia]] for ia in range(rA)):
                continue
            for ib in range(rB):
                if BtoC[ib] != -1:
                    cblk[BtoC[ib]] = bblk[ib]
            c = tuple(cblk)
            triples.append((ablk, bblk, c))
            csizes.append((c, int(np.prod([I.blocksize(b) for I, b in zip(Cis, c)]))))
    csizes.sort(key=lambda x: _block_key(x[0]))
    uniq = []
    for c, s in csizes:
        if not uniq or uniq[-1][0] != c:
            uniq.append((c, s))
    offs = []
    cur = 0
    for c, s in uniq:
        offs.append((c, cur))
        cur += s
    return offs, cur, triples


def contract(L: ITensor, R: ITensor) -> ITensor:
    """ITensor operator* (itensor.cc:935-981, 1090-1093)."""
    if L.cmb is not None or R.cmb is not None:
        return _combiner_contract(L, R)
    if L.order == 0:
        return R * L.elt()
    if R.order == 0:
        return L * R.elt()
    check_arrows(L.inds, R.inds)
    if L.is_qn and not R.is_qn:
        L = remove_qns(L)
    elif R.is_qn and not L.is_qn:
        R = remove_qns(R)
    Lind, Rind, _ = compute_labels(L.inds, R.inds)
    Nis, Cind = contract_is(L.inds, Lind, R.inds, Rind)
    if not L.is_qn:
        d = _block_contract(L.d, Lind, R.d, Rind, Cind)
        return ITensor(Nis, d=np.asarray(d, dtype=np.float64))
    offs, size, triples = get_contracted_offsets(L, R, Nis)
    C = ITensor(Nis, offsets=offs, store=np.full(size, np.nan))
    betas = [0.0] * len(offs)
    for ablk, bblk, cblk in triples:
        a = L.block(ablk)
        b = R.block(bblk)
        loc = C.block_loc(cblk)
        cref = C.block(cblk)
        val = _block_contract(a, Lind, b, Rind, Cind)
        if betas[loc] == 0.0:
            cref[...] = val
        else:
            cref[...] += val
        betas[loc] = 1.0
    return C


def calc_perm(is_from, is_to):
    """perm[i] = position in is_from of is_to[i] (calcPerm semantics)."""
    p = []
    for I in is_to:
        k = index_position(is_from, I)
        if k < 0:
            raise ValueError("different index structure in +=")
        p.append(k)
    return p


def add_into(A: ITensor, B: ITensor, alpha: float) -> ITensor:
    """A += alpha*B (daxpy, itensor.cc:1207-1245; PlusEQ dense.cc:362-415,
    qdense.cc:506-668).  Returns A (possibly with widened block set)."""
    if A.order != B.order:
        raise ValueError("+=: different number of indices")
    if A.is_qn != B.is_qn:
        raise ValueError("+=: QN and dense mixed")
    if B.is_qn and not B.offsets:
        return A
    perm = calc_perm(B.inds, A.inds)  # A's index i is B's index perm[i]
    if not A.is_qn:
        A.d = A.d + alpha * np.transpose(B.d, perm)
        return A
    if A.order == 0:
        A.store = A.store + alpha * B.store
        return A
    # QDense: block-set union when B has blocks A lacks (qdense.cc:582-663)
    bblocks_in_A = sorted((tuple(bb[perm[i]] for i in range(A.order)) for bb, _ in B.offsets),
                          key=_block_key)
    ablocks = [b for b, _ in A.offsets]
    merged = []
    ia = ib = 0
    while ia < len(ablocks) and ib < len(bblocks_in_A):
        a, b = ablocks[ia], bblocks_in_A[ib]
        if _block_less(b, a):
            merged.append(b); ib += 1
        elif _block_less(a, b):
            merged.append(a); ia += 1
        else:
            merged.append(a); ia += 1; ib += 1
    merged += ablocks[ia:] + bblocks_in_A[ib:]
    if len(ablocks) < len(merged):
        offs = []
        cur = 0
        for b in merged:
            offs.append((b, cur))
            cur += int(np.prod([I.blocksize(x) for I, x in zip(A.inds, b)]))
        N = ITensor(A.inds, offsets=offs, store=np.zeros(cur))
        for b, _ in A.offsets:
            N.block(b)[...] += A.block(b)
        A.offsets, A.store = N.offsets, N.store
    for ablk, _ in A.offsets:
        bblk = [0] * A.order
        for i in range(A.order):
            bblk[perm[i]] = ablk[i]
        bb = B.block(tuple(bblk))
        if bb is None:
            continue
        A.block(ablk)[...] += alpha * np.transpose(bb, perm)
    return A


def norm(T: ITensor) -> float:
    """norm = dnrm2 of the storage (itensor.cc:754-766, dense.cc:150-156,
    qdense.cc:409-415)."""
    return float(np.linalg.norm(T.store if T.is_qn else T.d))


def dot(A: ITensor, B: ITensor) -> float:
    """eltC(dag(A)*B) for same index sets (real data)."""
    return (A.dag() * B).elt()


# ----------------------------------------------------------- combiner ----
class _QCmb:
    """QCombiner block ranges: for each combination of sector numbers of the
    combined indices, the combined sector and [start, end) inside it."""

    def __init__(self, inds, ranges, nsectors):
        self.inds = inds
        self.ranges = ranges  # dict: tuple(sector ids) -> (csector, start, end)
        self.nsectors = nsectors


def combiner(inds, tags="Link,CMB", dir=None):
    """combiner(inds) (itensor.cc:1468-1560): returns (C, cind) where C has
    indices (cind, dag(inds)...) for QN, (cind, inds...) dense."""
    inds = list(inds)
    if not has_qns(inds):
        rm = int(np.prod([i.dim for i in inds]))
        c = Index(rm, tags)
        return ITensor([c] + inds, cmb=("dense", None)), c
    cdir = Out
    if dir is not None:
        cdir = dir
    elif all(i.dir == In for i in inds):
        cdir = In
    qms = []  # list of [QN, m]
    ranges = {}
    ranges_list = [range(I.nblock) for I in inds]
    for rev in itertools.product(*reversed(ranges_list)):
        I = tuple(reversed(rev))  # C.range(): first index fastest
        q = QN()
        m = 1
        for J, b in zip(inds, I):
            q = q + J.qn(b) * J.dir * cdir
            m *= J.blocksize(b)
        n = next((k for k, x in enumerate(qms) if x[0] == q), None)
        if n is not None:
            ranges[I] = (n, qms[n][1], qms[n][1] + m)
            qms[n][1] += m
        else:
            ranges[I] = (len(qms), 0, m)
            qms.append([q, m])
    c = Index([(q, m) for q, m in qms], tags, dir=cdir)
    C = ITensor([c] + [i.dag() for i in inds], cmb=("qn", _QCmb(inds, ranges, len(qms))))
    return C, c


def _combiner_contract(L: ITensor, R: ITensor) -> ITensor:
    if L.cmb is not None and R.cmb is not None:
        raise ValueError("combiner times combiner")
    C, T = (L, R) if L.cmb is not None else (R, L)
    kind, q = C.cmb
    cind = C.inds[0]
    jc = index_position(T.inds, cind)
    comb = C.inds[1:]
    if kind == "dense" or not T.is_qn:
        if T.is_qn:
            T = remove_qns(T)
        if jc >= 0:  # uncombine (combiner.cc:63-80): replace cind by comb inds in place
            newinds = T.inds[:jc] + list(comb) + T.inds[jc + 1:]
            shape = [i.dim for i in newinds]
            d = T.d.reshape(shape, order="F")
            return ITensor([i.remove_qns() for i in newinds], d=d.copy())
        # combine (combiner.cc:81-150)
        pos = [index_position(T.inds, i) for i in comb]
        if any(p < 0 for p in pos):
            raise ValueError("combiner: missing index")
        contig = all(pos[k] == pos[0] + k for k in range(len(pos)))
        if contig:
            j1 = pos[0]
            newinds = T.inds[:j1] + [cind] + T.inds[j1 + len(comb):]
            shape = [i.dim for i in newinds]
            return ITensor(newinds, d=T.d.reshape(shape, order="F").copy())
        rest = [k for k in range(T.order) if k not in pos]
        perm = pos + rest
        d = np.transpose(T.d, perm)
        newinds = [cind] + [T.inds[k] for k in rest]
        shape = [i.dim for i in newinds]
        return ITensor(newinds, d=np.reshape(d, shape, order="F").copy())
    # QN combine / uncombine (qcombiner.cc:125-300)
    if jc < 0:
        return _qcombine(T, C, q)
    return _quncombine(T, C, q, jc)


def _qcombine(T: ITensor, C: ITensor, q: _QCmb) -> ITensor:
    """combine (qcombiner.cc:125-213): new indices (cind, uncombined...)."""
    cind = C.inds[0]
    comb = C.inds[1:]  # dag of the combined indices
    dis = T.inds
    dr = len(dis)
    ncomb = len(comb)
    dperm = []
    dest = ncomb
    for i in range(dr):
        jc = index_position(comb, dis[i])
        if jc >= 0:
            dperm.append(jc)
        else:
            dperm.append(dest)
            dest += 1
    combined = [dperm[i] < ncomb for i in range(dr)]
    Nis = [cind] + [dis[i] for i in range(dr) if not combined[i]]
    N = qdense(Nis, T.div())
    for io, _ in T.offsets:
        dref = T.block(io)
        # permute combined indices to front (in combiner order), rest after
        src_axes = [0] * dr
        for i in range(dr):
            src_axes[dperm[i]] = i
        P = np.transpose(dref, src_axes)
        cshape = P.shape[:ncomb]
        g = P.reshape((int(np.prod(cshape)),) + P.shape[ncomb:], order="F")
        cblock = [0] * ncomb
        nblock = [0] * len(Nis)
        nu = 1
        for i in range(dr):
            if combined[i]:
                cblock[dperm[i]] = io[i]
            else:
                nblock[nu] = io[i]
                nu += 1
        sec, start, end = q.ranges[tuple(cblock)]
        nblock[0] = sec
        nb = N.block(tuple(nblock))
        nb[start:end, ...] = g
    return N


def _quncombine(T: ITensor, C: ITensor, q: _QCmb, jc: int) -> ITensor:
    """uncombine (qcombiner.cc:216-300): cind at position jc is replaced by
    the combiner's other indices, exactly as they appear on the combiner."""
    comb = list(C.inds[1:])
    Nis = T.inds[:jc] + comb + T.inds[jc + 1:]
    N = qdense(Nis, T.div())
    inv = {}
    for secs, (sec, start, end) in q.ranges.items():
        inv.setdefault(sec, []).append((secs, start, end))
    for io, _ in T.offsets:
        tb = T.block(io)
        sec = io[jc]
        for secs, start, end in inv.get(sec, []):
            nblock = tuple(io[:jc]) + tuple(secs) + tuple(io[jc + 1:])
            nb = N.block(nblock)
            if nb is None:
                continue
            sub = np.take(tb, range(start, end), axis=jc)
            shape = tb.shape[:jc] + tuple(I.blocksize(b) for I, b in zip(comb, secs)) + tb.shape[jc + 1:]
            nb[...] = sub.reshape(shape, order="F")
    return N
