// ITensor operations of the host layer.  Bookkeeping follows the reference
// (file:line cited per function); every arithmetic step is a libitx call.
#include <algorithm>
#include <cmath>
#include <numeric>

#include <map>
#include "itensor.h"

namespace itensor {

namespace {

std::shared_ptr<Storage> mkDense(size_t n) {
  auto s = std::make_shared<Storage>();
  s->type = SType::Dense;
  s->size = n;
  s->data = devAlloc(n);
  return s;
}

std::shared_ptr<Storage> mkQDense(BlockOffsets offs, size_t n, bool zero) {
  auto s = std::make_shared<Storage>();
  s->type = SType::QDense;
  s->size = n;
  s->offsets = std::move(offs);
  s->data = zero ? devZeros(n) : devAlloc(n);
  return s;
}

std::vector<int64_t> dimsOf(IndexSet const &is) {
  std::vector<int64_t> d(is.size());
  for (size_t i = 0; i < is.size(); ++i) d[i] = is[i].dim();
  return d;
}

std::vector<int64_t> blockDims(IndexSet const &is, Block const &b) {
  std::vector<int64_t> d(is.size());
  for (size_t i = 0; i < is.size(); ++i) d[i] = is[i].blocksize0(b[i]);
  return d;
}

int64_t prod(std::vector<int64_t> const &v) {
  int64_t p = 1;
  for (auto x : v) p *= x;
  return p;
}

std::vector<int64_t> colStrides(std::vector<int64_t> const &d) {
  std::vector<int64_t> s(d.size());
  int64_t x = 1;
  for (size_t i = 0; i < d.size(); ++i) {
    s[i] = x;
    x *= d[i];
  }
  return s;
}

struct BlockCmp {
  bool operator()(BlOf const &a, Block const &b) const { return blockLess(a.block, b); }
};

// offsetOf: binary search in block order (itdata/qdense.cc:214-221)
int64_t offsetOf(BlockOffsets const &offs, Block const &b) {
  auto it = std::lower_bound(offs.begin(), offs.end(), b, BlockCmp());
  if (it != offs.end() && it->block == b) return it->offset;
  return -1;
}

// computeLabels (tensor/contract.h:155-202)
void computeLabels(IndexSet const &L, IndexSet const &R, std::vector<int32_t> &Lind, std::vector<int32_t> &Rind) {
  Lind.assign(L.size(), 0);
  Rind.assign(R.size(), 0);
  int ncont = 0;
  for (size_t i = 0; i < L.size(); ++i)
    for (size_t j = 0; j < R.size(); ++j)
      if (L[i] == R[j]) {
        Lind[i] = -(1 + ncont);
        Rind[j] = -(1 + ncont);
        ++ncont;
        break;
      }
  int u = ncont;
  for (auto &x : Lind)
    if (x == 0) x = ++u;
  for (auto &x : Rind)
    if (x == 0) x = ++u;
}

// contractIS, unsorted: uncontracted L then R (indexset_impl.h:79-127)
void contractIS(IndexSet const &L, std::vector<int32_t> const &Lind, IndexSet const &R,
                std::vector<int32_t> const &Rind, IndexSet &N, std::vector<int32_t> &Cind) {
  N.clear();
  Cind.clear();
  for (size_t i = 0; i < L.size(); ++i)
    if (Lind[i] > 0) {
      N.push_back(L[i]);
      Cind.push_back(Lind[i]);
    }
  for (size_t j = 0; j < R.size(); ++j)
    if (Rind[j] > 0) {
      N.push_back(R[j]);
      Cind.push_back(Rind[j]);
    }
}

// checkArrows (itensor.cc:891-920)
void checkArrows(IndexSet const &a, IndexSet const &b, bool shouldMatch = false) {
  if (!hasQNs(a) || !hasQNs(b)) return;
  for (auto &I : a)
    for (auto &J : b)
      if (I == J && (shouldMatch ^ (I.dir() == J.dir())))
        throw ITError("Mismatched QN Index arrows: " + I.str() + " " + J.str());
}

ITensor contractDense(ITensor const &A, ITensor const &B) {
  HostTimer ht(HT_CONTRACT_DENSE);
  std::vector<int32_t> la, lb, lc;
  computeLabels(A.inds(), B.inds(), la, lb);
  IndexSet N;
  contractIS(A.inds(), la, B.inds(), lb, N, lc);
  auto dA = dimsOf(A.inds()), dB = dimsOf(B.inds()), dC = dimsOf(N);
  auto st = mkDense(size_t(prod(dC)));
  check(itx_contract_dense(ctx(), A.store().data->p, int(dA.size()), dA.data(), la.data(), B.store().data->p,
                           int(dB.size()), dB.data(), lb.data(), st->data->p, int(dC.size()), dC.data(),
                           lc.data(), 1.0, 0.0, contractPrecision()),
        "itx_contract_dense");
  return ITensor(N, st);
}

}  // namespace

// getContractedOffsets (itdata/qutil.h:93-242): labels, result indices, the
// reachable block pairs (A block x B block with equal contracted sectors,
// :155-202) and the result block list, sorted in block order and
// deduplicated, with prefix-sum offsets (:221-239).  Structure only.
QNTaskList qnTaskList(IndexSet const &Ais, BlockOffsets const &Aoffs, IndexSet const &Bis,
                      BlockOffsets const &Boffs) {
  HostTimer ht(HT_QN_TASKLIST);
  QNTaskList tl;
  computeLabels(Ais, Bis, tl.la, tl.lb);
  contractIS(Ais, tl.la, Bis, tl.lb, tl.N, tl.lc);
  IndexSet const &N = tl.N;
  const int rA = int(Ais.size()), rB = int(Bis.size()), rC = int(N.size());
  std::vector<int> AtoB(rA, -1), AtoC(rA, -1), BtoC(rB, -1);
  for (int ic = 0; ic < rC; ++ic) {
    int j = indexPosition(Ais, N[ic]);
    if (j >= 0)
      AtoC[j] = ic;
    else
      BtoC[indexPosition(Bis, N[ic])] = ic;
  }
  for (int ia = 0; ia < rA; ++ia)
    for (int ib = 0; ib < rB; ++ib)
      if (Ais[ia] == Bis[ib]) {
        AtoB[ia] = ib;
        break;
      }
  // B blocks ordered by their contracted block numbers (in A's index
  // order): each A block then finds its partners as one equal range, and
  // the pairs come out in the order of the reference's double loop (A
  // blocks outer, B blocks in storage order inner)
  std::vector<int> cA, cB;
  for (int ia = 0; ia < rA; ++ia)
    if (AtoB[ia] != -1) {
      cA.push_back(ia);
      cB.push_back(AtoB[ia]);
    }
  const int nc = int(cA.size());
  std::vector<int> border(Boffs.size());
  std::iota(border.begin(), border.end(), 0);
  auto bkey_less = [&](int x, int y) {
    Block const &bx = Boffs[size_t(x)].block, &by = Boffs[size_t(y)].block;
    for (int q = 0; q < nc; ++q)
      if (bx[cB[q]] != by[cB[q]]) return bx[cB[q]] < by[cB[q]];
    return x < y;  // storage order within a key
  };
  std::sort(border.begin(), border.end(), bkey_less);
  struct Pair {
    int a, b;
  };
  std::vector<Pair> pairs;
  for (int x = 0; x < int(Aoffs.size()); ++x) {
    Block const &ab = Aoffs[size_t(x)].block;
    auto cmpAB = [&](int y, int) {  // B block y's key < A block's key
      Block const &by = Boffs[size_t(y)].block;
      for (int q = 0; q < nc; ++q)
        if (by[cB[q]] != ab[cA[q]]) return by[cB[q]] < ab[cA[q]];
      return false;
    };
    auto cmpBA = [&](int, int y) {  // A block's key < B block y's key
      Block const &by = Boffs[size_t(y)].block;
      for (int q = 0; q < nc; ++q)
        if (ab[cA[q]] != by[cB[q]]) return ab[cA[q]] < by[cB[q]];
      return false;
    };
    auto lo = std::lower_bound(border.begin(), border.end(), 0, cmpAB);
    auto hi = std::upper_bound(lo, border.end(), 0, cmpBA);
    for (auto it = lo; it != hi; ++it) pairs.push_back({x, *it});
  }
  // C blocks: sorted (reference order), deduplicated, prefix-summed sizes.
  // A C block is coded as the mixed-radix number sum_c block[c] * R_c with
  // R_c = nblock(N_0) ... nblock(N_{c-1}): the last index is the most
  // significant digit, so numeric order is the reverse-lexicographic block
  // order, and the code of a pair is the A part's plus the B part's.
  std::vector<int64_t> R(size_t(rC) + 1, 1);
  bool fits = true;
  for (int c = 0; c < rC; ++c) {
    const int64_t nb = N[c].nblock();
    fits = fits && R[size_t(c)] <= (int64_t(1) << 62) / std::max<int64_t>(1, nb);
    R[size_t(c) + 1] = fits ? R[size_t(c)] * nb : 0;
  }
  if (!fits) throw ITError("contraction: too many result blocks to enumerate");
  std::vector<int64_t> aCode(Aoffs.size(), 0), bCode(Boffs.size(), 0);
  for (size_t x = 0; x < Aoffs.size(); ++x)
    for (int ia = 0; ia < rA; ++ia)
      if (AtoC[ia] != -1) aCode[x] += Aoffs[x].block[size_t(ia)] * R[size_t(AtoC[ia])];
  for (size_t y = 0; y < Boffs.size(); ++y)
    for (int ib = 0; ib < rB; ++ib)
      if (BtoC[ib] != -1) bCode[y] += Boffs[y].block[size_t(ib)] * R[size_t(BtoC[ib])];
  const int nt = int(pairs.size());
  std::vector<int64_t> pkey(pairs.size(), 0);
  for (int t = 0; t < nt; ++t) pkey[size_t(t)] = aCode[size_t(pairs[size_t(t)].a)] + bCode[size_t(pairs[size_t(t)].b)];
  std::vector<int64_t> ukey(pkey);
  std::sort(ukey.begin(), ukey.end());
  ukey.erase(std::unique(ukey.begin(), ukey.end()), ukey.end());
  {
    int64_t cur = 0;
    tl.coffs.reserve(ukey.size());
    for (int64_t code : ukey) {
      Block cb(size_t(rC), 0);
      int64_t sz = 1;
      for (int c = 0; c < rC; ++c) {
        cb[size_t(c)] = int((code / R[size_t(c)]) % N[c].nblock());
        sz *= N[c].blocksize0(cb[size_t(c)]);
      }
      tl.coffs.push_back({std::move(cb), cur});
      cur += sz;
    }
    tl.csize = cur;
  }
  // block dims of every A and B block, once
  std::vector<int64_t> adims(Aoffs.size() * size_t(rA)), bdims(Boffs.size() * size_t(rB));
  for (size_t x = 0; x < Aoffs.size(); ++x)
    for (int i = 0; i < rA; ++i) adims[x * rA + i] = Ais[i].blocksize0(Aoffs[x].block[size_t(i)]);
  for (size_t x = 0; x < Boffs.size(); ++x)
    for (int i = 0; i < rB; ++i) bdims[x * rB + i] = Bis[i].blocksize0(Boffs[x].block[size_t(i)]);
  tl.nt = nt;
  tl.offs.resize(3 * size_t(nt));
  tl.bda.resize(size_t(nt) * rA);
  tl.bdb.resize(size_t(nt) * rB);
  tl.cfirst.assign(size_t(nt), 0);
  std::vector<char> seen(tl.coffs.size(), 0);
  for (int t = 0; t < nt; ++t) {
    Pair const &pr = pairs[size_t(t)];
    const int ci = int(std::lower_bound(ukey.begin(), ukey.end(), pkey[size_t(t)]) - ukey.begin());
    tl.offs[3 * size_t(t)] = Aoffs[size_t(pr.a)].offset;
    tl.offs[3 * size_t(t) + 1] = Boffs[size_t(pr.b)].offset;
    tl.offs[3 * size_t(t) + 2] = tl.coffs[size_t(ci)].offset;
    if (!seen[size_t(ci)]) tl.cfirst[size_t(t)] = seen[size_t(ci)] = 1;
    std::copy_n(adims.begin() + ptrdiff_t(size_t(pr.a) * rA), rA, tl.bda.begin() + ptrdiff_t(size_t(t) * rA));
    std::copy_n(bdims.begin() + ptrdiff_t(size_t(pr.b) * rB), rB, tl.bdb.begin() + ptrdiff_t(size_t(t) * rB));
  }
  return tl;
}

namespace {

// the block loop of doTask(Contract,QDense,QDense) (itdata/qdense.cc:671-743)
// as one grouped launch
ITensor contractQN(ITensor const &A, ITensor const &B) {
  HostTimer ht(HT_CONTRACT_QN);
  QNTaskList tl = qnTaskList(A.inds(), A.store().offsets, B.inds(), B.store().offsets);
  auto st = mkQDense(tl.coffs, size_t(tl.csize), false);
  if (tl.nt > 0) {
    check(itx_contract_qn(ctx(), A.store().data->p, A.order(), tl.la.data(), B.store().data->p, B.order(),
                          tl.lb.data(), st->data->p, int(tl.N.size()), tl.lc.data(), tl.nt, tl.offs.data(),
                          tl.bda.data(), tl.bdb.data(), 1.0, contractPrecision()),
          "itx_contract_qn");
  }
  return ITensor(tl.N, st);
}

// -------------------------------------------------------------- block copy --
struct CopyTasks {
  int r;
  std::vector<int64_t> desc;
  explicit CopyTasks(int r_) : r(r_) {}
  void add(int64_t src, int64_t dst, std::vector<int64_t> const &dims, std::vector<int64_t> const &ss,
           std::vector<int64_t> const &ds) {
    desc.push_back(src);
    desc.push_back(dst);
    desc.insert(desc.end(), dims.begin(), dims.end());
    desc.insert(desc.end(), ss.begin(), ss.end());
    desc.insert(desc.end(), ds.begin(), ds.end());
  }
  int count() const { return int(desc.size() / size_t(2 + 3 * r)); }
  void run(double const *X, double *Y, double alpha, double beta) const {
    if (count() == 0) return;
    check(itx_block_copy(ctx(), X, Y, count(), r, desc.data(), alpha, beta), "itx_block_copy");
  }
};

// ------------------------------------------------------------ combiners --
ITensor combineDense(ITensor const &C, ITensor const &T) {
  Index const &cind = C.inds()[0];
  IndexSet comb(C.inds().begin() + 1, C.inds().end());
  int jc = indexPosition(T.inds(), cind);
  if (jc >= 0) {  // uncombine: replace cind in place (itdata/combiner.cc:63-80)
    IndexSet N;
    for (int j = 0; j < T.order(); ++j) {
      if (j == jc)
        for (auto &I : comb) N.push_back(I);
      else
        N.push_back(T.inds()[j]);
    }
    return ITensor(N, T.storePtr());
  }
  std::vector<int> pos;
  for (auto &I : comb) {
    int p = indexPosition(T.inds(), I);
    if (p < 0) throw ITError("combiner: missing index " + I.str());
    pos.push_back(p);
  }
  bool contig = true;
  for (size_t k = 0; k < pos.size(); ++k)
    if (pos[k] != pos[0] + int(k)) contig = false;
  if (contig) {  // no data movement (combiner.cc:98-112)
    IndexSet N(T.inds().begin(), T.inds().begin() + pos[0]);
    N.push_back(cind);
    N.insert(N.end(), T.inds().begin() + pos[0] + long(pos.size()), T.inds().end());
    return ITensor(N, T.storePtr());
  }
  // permute combined indices to the front in combiner order (combiner.cc:113-148)
  std::vector<int32_t> perm;  // Y index j <- X index perm[j]
  for (int p : pos) perm.push_back(p);
  IndexSet N{cind};
  for (int j = 0; j < T.order(); ++j)
    if (std::find(pos.begin(), pos.end(), j) == pos.end()) {
      perm.push_back(j);
      N.push_back(T.inds()[j]);
    }
  std::vector<int64_t> dY;
  for (int p : perm) dY.push_back(T.inds()[p].dim());
  auto st = mkDense(T.store().size);
  check(itx_permute(ctx(), T.store().data->p, st->data->p, int(perm.size()), dY.data(), perm.data(), 1.0, 0.0),
        "itx_permute");
  return ITensor(N, st);
}

// combine (itdata/qcombiner.cc:125-213)
ITensor combineQN(ITensor const &C, ITensor const &T) {
  auto const &qc = *C.store().qc;
  Index const &cind = C.inds()[0];
  IndexSet comb(C.inds().begin() + 1, C.inds().end());
  IndexSet const &dis = T.inds();
  const int dr = int(dis.size()), ncomb = int(comb.size());
  std::vector<int> dperm(dr);
  int dest = ncomb;
  for (int i = 0; i < dr; ++i) {
    int j = indexPosition(comb, dis[i]);
    dperm[i] = j >= 0 ? j : dest++;
  }
  IndexSet N{cind};
  for (int i = 0; i < dr; ++i)
    if (dperm[i] >= ncomb) N.push_back(dis[i]);
  auto [noffs, nsize] = getBlockOffsets(N, div(T));
  auto st = mkQDense(noffs, size_t(nsize), true);
  CopyTasks tasks(dr);
  for (auto const &io : T.store().offsets) {
    auto bd = blockDims(dis, io.block);
    auto bs = colStrides(bd);
    // iteration order: combined indices (combiner order) then the others
    std::vector<int> src_axis(dr);
    for (int i = 0; i < dr; ++i) src_axis[dperm[i]] = i;
    Block cblock(ncomb), nblock(N.size());
    int nu = 1;
    for (int i = 0; i < dr; ++i) {
      if (dperm[i] < ncomb)
        cblock[dperm[i]] = io.block[i];
      else
        nblock[nu++] = io.block[i];
    }
    auto it = qc.ranges.find(cblock);
    if (it == qc.ranges.end()) throw ITError("QCombiner: unknown sector combination");
    auto [sec, start, end] = it->second;
    nblock[0] = sec;
    int64_t noff = offsetOf(st->offsets, nblock);
    if (noff < 0) throw ITError("QCombiner: target block missing");
    auto nd = blockDims(N, nblock);
    auto ns = colStrides(nd);
    std::vector<int64_t> dims(dr), ss(dr), ds(dr);
    int64_t cstride = 1;
    for (int k = 0; k < dr; ++k) {
      int i = src_axis[k];
      dims[k] = bd[i];
      ss[k] = bs[i];
      if (k < ncomb) {
        ds[k] = cstride * ns[0];
        cstride *= bd[i];
      } else {
        ds[k] = ns[k - ncomb + 1];
      }
    }
    tasks.add(io.offset, noff + int64_t(start) * ns[0], dims, ss, ds);
  }
  tasks.run(T.store().data->p, st->data->p, 1.0, 0.0);
  return ITensor(N, st);
}

// uncombine (itdata/qcombiner.cc:216-300)
ITensor uncombineQN(ITensor const &C, ITensor const &T, int jc) {
  auto const &qc = *C.store().qc;
  IndexSet comb(C.inds().begin() + 1, C.inds().end());
  IndexSet const &dis = T.inds();
  const int dr = int(dis.size()), ncomb = int(comb.size());
  IndexSet N;
  for (int n = 0; n < dr; ++n) {
    if (n == jc)
      for (auto &I : comb) N.push_back(I);
    else
      N.push_back(dis[n]);
  }
  auto [noffs, nsize] = getBlockOffsets(N, div(T));
  auto st = mkQDense(noffs, size_t(nsize), true);
  const int nr = int(N.size());
  CopyTasks tasks(nr);
  for (auto const &io : T.store().offsets) {
    auto bd = blockDims(dis, io.block);
    auto bs = colStrides(bd);
    const int sec = io.block[jc];
    for (auto const &kv : qc.ranges) {
      auto [s, start, end] = kv.second;
      if (s != sec) continue;
      Block nblock;
      for (int m = 0; m < jc; ++m) nblock.push_back(io.block[m]);
      for (int v : kv.first) nblock.push_back(v);
      for (int m = jc + 1; m < dr; ++m) nblock.push_back(io.block[m]);
      int64_t noff = offsetOf(st->offsets, nblock);
      if (noff < 0) continue;
      auto nd = blockDims(N, nblock);
      auto ns = colStrides(nd);
      std::vector<int64_t> ss(nr);
      int64_t sub = bs[jc];
      for (int m = 0; m < nr; ++m) {
        if (m < jc) {
          ss[m] = bs[m];
        } else if (m < jc + ncomb) {
          ss[m] = sub;
          sub *= nd[m];
        } else {
          ss[m] = bs[m - ncomb + 1];
        }
      }
      tasks.add(io.offset + int64_t(start) * bs[jc], noff, nd, ss, ns);
    }
  }
  tasks.run(T.store().data->p, st->data->p, 1.0, 0.0);
  return ITensor(N, st);
}

ITensor combinerContract(ITensor const &C, ITensor const &T) {
  if (C.store().type == SType::Combiner || !T.isQN()) {
    if (T.isQN()) return combineDense(C, removeQNs(T));
    return combineDense(C, T);
  }
  int jc = indexPosition(T.inds(), C.inds()[0]);
  if (jc < 0) return combineQN(C, T);
  return uncombineQN(C, T, jc);
}

bool sameIndexSet(IndexSet const &a, IndexSet const &b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); ++i)
    if (!(a[i] == b[i])) return false;
  return true;
}

bool sameOffsets(BlockOffsets const &a, BlockOffsets const &b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); ++i)
    if (a[i].offset != b[i].offset || a[i].block != b[i].block) return false;
  return true;
}

}  // namespace

// ------------------------------------------------------------ ContractPlan --
ContractPlan::ContractPlan(ITensor const &A, ITensor const &B, int prec, bool perBlock, Storage const *cover,
                           std::vector<int> const *cperm) : perBlock_(perBlock) {
  if (!A || !B) throw ITError("ContractPlan: default constructed ITensor");
  if (A.isCombiner() || B.isCombiner()) throw ITError("ContractPlan: combiner operand");
  if (A.isQN() != B.isQN()) throw ITError("ContractPlan: QN and dense operands");
  checkArrows(A.inds(), B.inds());
  qn_ = A.isQN();
  if (!qn_) {
    std::vector<int32_t> la, lb, lc;
    computeLabels(A.inds(), B.inds(), la, lb);
    contractIS(A.inds(), la, B.inds(), lb, N_, lc);
    auto dA = dimsOf(A.inds()), dB = dimsOf(B.inds()), dC = dimsOf(N_);
    size_ = size_t(prod(dC));
    HostTimer ht(HT_PLAN_DENSE);
    check(itx_plan_dense(ctx(), int(dA.size()), dA.data(), la.data(), int(dB.size()), dB.data(), lb.data(),
                         int(dC.size()), dC.data(), lc.data(), prec, &plan_),
          "itx_plan_dense");
    account(plan_);
    return;
  }
  QNTaskList tl = qnTaskList(A.inds(), A.store().offsets, B.inds(), B.store().offsets);
  N_ = tl.N;
  if (cover && (cperm || cover->offsets.size() > tl.coffs.size())) {
    // re-address every task's C block in the covering layout (result
    // indices in the order cperm names); a produced block missing from it
    // leaves the natural layout in place
    const int rC = int(tl.N.size());
    std::vector<int64_t> noff(tl.coffs.size());
    bool ok = !cperm || int(cperm->size()) == rC;
    Block pb(size_t(rC), 0);
    for (size_t c = 0; c < tl.coffs.size() && ok; ++c) {
      Block const &nb = tl.coffs[c].block;
      for (int q = 0; q < rC; ++q) pb[size_t(q)] = nb[size_t(cperm ? (*cperm)[size_t(q)] : q)];
      noff[c] = offsetOf(cover->offsets, pb);
      ok = noff[c] >= 0;
    }
    if (ok) {
      for (int t = 0; t < tl.nt; ++t) {
        int64_t &co = tl.offs[3 * size_t(t) + 2];
        auto it = std::lower_bound(tl.coffs.begin(), tl.coffs.end(), co,
                                   [](BlOf const &x, int64_t v) { return x.offset < v; });
        co = noff[size_t(it - tl.coffs.begin())];
      }
      if (cperm) {
        IndexSet N2;
        std::vector<int32_t> lc2;
        for (int q = 0; q < rC; ++q) {
          N2.push_back(tl.N[size_t((*cperm)[size_t(q)])]);
          lc2.push_back(tl.lc[size_t((*cperm)[size_t(q)])]);
        }
        tl.N = N2;
        tl.lc = lc2;
        N_ = tl.N;
      }
      zeroOut_ = cover->offsets.size() > tl.coffs.size();
      tl.coffs = cover->offsets;
      tl.csize = int64_t(cover->size);
    }
  }
  coffs_ = tl.coffs;
  size_ = size_t(tl.csize);
  const int rA = A.order(), rB = B.order(), rC = int(N_.size());
  if (tl.nt == 0) return;
  if (!perBlock) {
    HostTimer ht(HT_PLAN_QN);
    check(itx_plan_qn(ctx(), rA, tl.la.data(), rB, tl.lb.data(), rC, tl.lc.data(), tl.nt, tl.offs.data(),
                      tl.bda.data(), tl.bdb.data(), prec, &plan_),
          "itx_plan_qn");
    account(plan_);
    return;
  }
  try {
    for (int t = 0; t < tl.nt; ++t) {
      itx_plan *p = nullptr;
      check(itx_plan_qn(ctx(), rA, tl.la.data(), rB, tl.lb.data(), rC, tl.lc.data(), 1, tl.offs.data() + 3 * t,
                        tl.bda.data() + size_t(t) * rA, tl.bdb.data() + size_t(t) * rB, prec, &p),
            "itx_plan_qn");
      blockPlans_.push_back(p);
      blockBeta_.push_back(tl.cfirst[size_t(t)] ? 0.0 : 1.0);
      account(p);
    }
  } catch (...) {
    for (auto *p : blockPlans_) itx_plan_destroy(ctx(), p);
    throw;
  }
}

void ContractPlan::account(itx_plan const *p) {
  const double f = itx_plan_flops(p);
  int form = ITX_PREC_F64;
  check(itx_plan_info(p, nullptr, &form, nullptr, nullptr), "itx_plan_info");
  flops_ += f;
  formFlops_[size_t(form)] += f;
}

ContractPlan::~ContractPlan() {
  // after shutdown() the context has released every plan already
  itx_ctx *c = liveCtx();
  if (!c) return;
  if (plan_) itx_plan_destroy(c, plan_);
  for (auto *p : blockPlans_) itx_plan_destroy(c, p);
}

std::shared_ptr<Storage> ContractPlan::newOutput() const {
  return qn_ ? mkQDense(coffs_, size_, zeroOut_) : mkDense(size_);
}

void ContractPlan::exec(double const *A, double const *B, double *C) const {
  if (plan_) {
    check(itx_plan_execute(ctx(), plan_, A, B, C, 1.0, 0.0), "itx_plan_execute");
    return;
  }
  for (size_t t = 0; t < blockPlans_.size(); ++t)
    check(itx_plan_execute(ctx(), blockPlans_[t], A, B, C, 1.0, blockBeta_[t]), "itx_plan_execute");
}

bool ContractPlan::sameStructure(ITensor const &a, IndexSet const &is, bool qn, BlockOffsets const &offs) {
  if (!a || a.isCombiner() || a.isQN() != qn) return false;
  IndexSet const &x = a.inds();
  if (x.size() != is.size()) return false;
  for (size_t i = 0; i < x.size(); ++i)
    if (!(x[i] == is[i]) || x[i].dim() != is[i].dim() || x[i].dir() != is[i].dir()) return false;
  if (!qn) return true;
  BlockOffsets const &o = a.store().offsets;
  if (o.size() != offs.size()) return false;
  for (size_t i = 0; i < o.size(); ++i)
    if (o[i].offset != offs[i].offset || o[i].block != offs[i].block) return false;
  return true;
}

// ---------------------------------------------------------------- ITensor --
void ITensor::detach() {
  if (!store_ || store_.use_count() == 1) return;
  auto s = std::make_shared<Storage>(*store_);
  if (store_->data) {
    s->data = devAlloc(store_->size);
    if (store_->size)
      check(itx_memcpy_dtod(ctx(), s->data->p, store_->data->p, sizeof(double) * store_->size), "itx_memcpy_dtod");
  }
  store_ = s;
}

ITensor &ITensor::operator*=(ITensor const &R) {
  ITensor &L = *this;
  if (!L || !R) throw ITError("Default constructed ITensor in product");
  if (L.isCombiner() || R.isCombiner()) {
    if (L.isCombiner() && R.isCombiner()) throw ITError("product of two combiners");
    *this = L.isCombiner() ? combinerContract(L, R) : combinerContract(R, L);
    return *this;
  }
  if (L.order() == 0) {
    Real z = L.elt();
    *this = R * z;
    return *this;
  }
  if (R.order() == 0) return operator*=(R.elt());
  checkArrows(L.inds(), R.inds());
  const bool hqL = L.isQN(), hqR = R.isQN();
  if (hqL && !hqR) {
    *this = contractDense(removeQNs(L), R);
  } else if (!hqL && hqR) {
    *this = contractDense(L, removeQNs(R));
  } else if (hqL) {
    *this = contractQN(L, R);
  } else {
    *this = contractDense(L, R);
  }
  return *this;
}

ITensor &ITensor::operator*=(Real r) {
  if (!*this) throw ITError("scaling a default constructed ITensor");
  if (isCombiner()) throw ITError("scaling a combiner");
  if (!store_->size) return *this;
  if (store_.use_count() > 1) {
    // shared storage (copy-on-write): write the scaled copy directly
    auto s = std::make_shared<Storage>(*store_);
    s->data = devAlloc(store_->size);
    const double *V[1] = {store_->data->p};
    check(itx_lincomb(ctx(), V, &r, 1, s->data->p, store_->size), "itx_lincomb");
    store_ = s;
    return *this;
  }
  check(itx_scal(ctx(), r, store_->data->p, store_->size), "itx_scal");
  return *this;
}

ITensor &ITensor::operator+=(ITensor const &R) {
  if (!*this || !store_) return (*this = R);
  if (!R) throw ITError("Right-hand-side of ITensor += is default constructed");
  if (this == &R) return operator*=(2.0);
  daxpy(*this, R, 1.0);
  return *this;
}

ITensor &ITensor::operator-=(ITensor const &R) {
  if (!*this || !store_) {
    *this = R;
    return operator*=(-1.0);
  }
  if (this == &R) return operator*=(0.0);
  daxpy(*this, R, -1.0);
  return *this;
}

// daxpy / PlusEQ (itensor.cc:1207-1245; itdata/dense.cc:362-415;
// itdata/qdense.cc:506-668 incl. the block-set union)
void daxpy(ITensor &L, ITensor const &R, Real alpha) {
  if (L.order() != R.order()) throw ITError("ITensor::operator+=: different number of indices");
  if (R.isQN() && R.store().offsets.empty()) return;
  if (L.isQN() != R.isQN()) throw ITError("ITensor::operator+=: QN and dense tensors mixed");
  const int r = L.order();
  std::vector<int32_t> perm(r);  // L index i <- R index perm[i]
  bool ident = true;
  for (int i = 0; i < r; ++i) {
    int p = indexPosition(R.inds(), L.inds()[i]);
    if (p < 0) throw ITError("ITensor::operator+=: different index structure");
    perm[i] = p;
    if (p != i) ident = false;
  }
  checkArrows(L.inds(), R.inds(), true);
  if (!L.isQN()) {
    if (ident && L.storePtr().use_count() > 1) {
      auto st = mkDense(L.store().size);
      const double *V[2] = {L.store().data->p, R.store().data->p};
      const double c[2] = {1.0, alpha};
      if (st->size) check(itx_lincomb(ctx(), V, c, 2, st->data->p, st->size), "itx_lincomb");
      L = ITensor(L.inds(), st);
      return;
    }
    L.detach();
    auto &Ls = L.store();
    if (ident) {
      check(itx_axpy(ctx(), alpha, R.store().data->p, Ls.data->p, Ls.size), "itx_axpy");
    } else {
      auto dY = dimsOf(L.inds());
      check(itx_permute(ctx(), R.store().data->p, Ls.data->p, r, dY.data(), perm.data(), alpha, 1.0), "itx_permute");
    }
    return;
  }
  if (r == 0) {
    L.detach();
    check(itx_axpy(ctx(), alpha, R.store().data->p, L.store().data->p, 1), "itx_axpy");
    return;
  }
  if (ident && sameOffsets(L.store().offsets, R.store().offsets)) {
    if (L.storePtr().use_count() > 1) {
      // shared storage: one kernel writes L + alpha R into a new block
      auto st = mkQDense(L.store().offsets, L.store().size, false);
      const double *V[2] = {L.store().data->p, R.store().data->p};
      const double c[2] = {1.0, alpha};
      if (st->size) check(itx_lincomb(ctx(), V, c, 2, st->data->p, st->size), "itx_lincomb");
      L = ITensor(L.inds(), st);
      return;
    }
    check(itx_axpy(ctx(), alpha, R.store().data->p, L.store().data->p, L.store().size), "itx_axpy");
    return;
  }
  // union of block sets (qdense.cc:582-663)
  std::vector<Block> bps;
  for (auto const &bo : R.store().offsets) {
    Block b(r);
    for (int i = 0; i < r; ++i) b[i] = bo.block[perm[i]];
    bps.push_back(b);
  }
  std::sort(bps.begin(), bps.end(), blockLess);
  auto const &Aoffs = L.store().offsets;
  std::vector<Block> merged;
  size_t ia = 0, ib = 0;
  while (ia < Aoffs.size() && ib < bps.size()) {
    if (blockLess(bps[ib], Aoffs[ia].block)) {
      merged.push_back(bps[ib++]);
    } else if (blockLess(Aoffs[ia].block, bps[ib])) {
      merged.push_back(Aoffs[ia++].block);
    } else {
      merged.push_back(Aoffs[ia].block);
      ++ia;
      ++ib;
    }
  }
  while (ia < Aoffs.size()) merged.push_back(Aoffs[ia++].block);
  while (ib < bps.size()) merged.push_back(bps[ib++]);
  if (Aoffs.size() < merged.size()) {
    BlockOffsets noffs;
    int64_t cur = 0;
    for (auto const &b : merged) {
      noffs.push_back({b, cur});
      cur += prod(blockDims(L.inds(), b));
    }
    auto st = mkQDense(noffs, size_t(cur), true);
    CopyTasks t(r);
    for (auto const &bo : Aoffs) {
      auto bd = blockDims(L.inds(), bo.block);
      auto s = colStrides(bd);
      t.add(bo.offset, offsetOf(noffs, bo.block), bd, s, s);
    }
    t.run(L.store().data->p, st->data->p, 1.0, 0.0);
    L = ITensor(L.inds(), st);
  } else {
    L.detach();
  }
  CopyTasks t(r);
  for (auto const &ao : L.store().offsets) {
    Block bb(r);
    for (int i = 0; i < r; ++i) bb[perm[i]] = ao.block[i];
    int64_t boff = offsetOf(R.store().offsets, bb);
    if (boff < 0) continue;
    auto ad = blockDims(L.inds(), ao.block);
    auto bd = blockDims(R.inds(), bb);
    auto bs = colStrides(bd);
    std::vector<int64_t> ss(r);
    for (int i = 0; i < r; ++i) ss[i] = bs[perm[i]];
    t.add(boff, ao.offset, ad, ss, colStrides(ad));
  }
  t.run(R.store().data->p, L.store().data->p, alpha, 1.0);
}

ITensor &ITensor::noPrime() {
  for (auto &I : is_) I.noPrime();
  return *this;
}
ITensor &ITensor::prime(int inc) {
  for (auto &I : is_) I.prime(inc);
  return *this;
}
ITensor &ITensor::prime(Index const &J, int inc) {
  for (auto &I : is_)
    if (I == J) I.prime(inc);
  return *this;
}
ITensor &ITensor::dag() {
  for (auto &I : is_) I.dag();
  return *this;
}
ITensor &ITensor::setTags(std::string const &tags, Index const &J) {
  for (auto &I : is_)
    if (I == J) I.setTags(tags);
  return *this;
}
ITensor &ITensor::replaceInd(Index const &from, Index const &to) {
  for (auto &I : is_)
    if (I == from) I = to;
  return *this;
}

Real ITensor::elt() const {
  if (order() != 0) throw ITError("elt: ITensor is not a scalar");
  if (isQN() && store_->offsets.empty()) return 0.0;
  double v = 0;
  check(itx_memcpy_dtoh(ctx(), &v, store_->data->p, sizeof(double)), "itx_memcpy_dtoh");
  return v;
}

size_t ITensor::nnz() const { return store_ ? store_->size : 0; }
long ITensor::nnzblocks() const { return isQN() ? long(store_->offsets.size()) : 1; }

ITensor operator*(ITensor L, ITensor const &R) {
  L *= R;
  return L;
}
ITensor operator*(ITensor T, Real r) {
  T *= r;
  return T;
}
ITensor operator*(Real r, ITensor T) {
  T *= r;
  return T;
}
ITensor operator+(ITensor L, ITensor const &R) {
  L += R;
  return L;
}
ITensor operator-(ITensor L, ITensor const &R) {
  L -= R;
  return L;
}

Real norm(ITensor const &T) {
  HostTimer ht(HT_NORM_DOT);
  if (!T) throw ITError("norm of default constructed ITensor");
  if (T.store().size == 0) return 0.0;
  double r = 0;
  check(itx_nrm2(ctx(), T.store().data->p, T.store().size, &r), "itx_nrm2");
  return r;
}

Real dotDag(ITensor const &A, ITensor const &B) {
  HostTimer ht(HT_NORM_DOT);
  if (sameIndexSet(A.inds(), B.inds()) &&
      (!A.isQN() || sameOffsets(A.store().offsets, B.store().offsets)) && A.isQN() == B.isQN()) {
    if (A.store().size == 0) return 0.0;
    double r = 0;
    check(itx_dot(ctx(), A.store().data->p, B.store().data->p, A.store().size, &r), "itx_dot");
    return r;
  }
  return (dag(A) * B).elt();
}

ITensor dag(ITensor T) {
  T.dag();
  return T;
}
ITensor prime(ITensor T, int inc) {
  T.prime(inc);
  return T;
}
ITensor prime(ITensor T, Index const &I, int inc) {
  T.prime(I, inc);
  return T;
}
ITensor noPrime(ITensor T) {
  T.noPrime();
  return T;
}

// removeQNs (itdata/qdense.cc:921-952)
ITensor removeQNs(ITensor const &T) {
  IndexSet N = T.inds();
  for (auto &I : N) I.removeQNs();
  if (!T.isQN()) return ITensor(N, T.storePtr());
  auto full = dimsOf(T.inds());
  auto fs = colStrides(full);
  auto st = mkDense(size_t(prod(full)));
  if (st->size) check(itx_fill(ctx(), st->data->p, st->size, 0.0), "itx_fill");
  const int r = T.order();
  CopyTasks t(r);
  for (auto const &bo : T.store().offsets) {
    auto bd = blockDims(T.inds(), bo.block);
    int64_t dst = 0;
    for (int i = 0; i < r; ++i) dst += T.inds()[i].blockStart(bo.block[i]) * fs[i];
    t.add(bo.offset, dst, bd, colStrides(bd), fs);
  }
  t.run(T.store().data->p, st->data->p, 1.0, 0.0);
  return ITensor(N, st);
}

Index commonIndex(ITensor const &A, ITensor const &B) {
  for (auto const &I : A.inds())
    if (indexPosition(B.inds(), I) >= 0) return I;
  return Index();
}

bool hasIndex(ITensor const &T, Index const &I) { return indexPosition(T.inds(), I) >= 0; }

QN div(ITensor const &T) {
  if (!T.isQN() || T.store().offsets.empty()) return QN();
  return calcDiv(T.inds(), T.store().offsets.front().block);
}

// combiner (itensor.cc:1468-1560)
std::tuple<ITensor, Index> combiner(IndexSet const &inds, Args const &args) {
  if (inds.empty()) throw ITError("No indices passed to combiner");
  std::string tags = args.getString("Tags", "Link,CMB");
  if (!hasQNs(inds)) {
    long rm = dim(inds);
    Index c(rm, tags);
    IndexSet is{c};
    is.insert(is.end(), inds.begin(), inds.end());
    auto st = std::make_shared<Storage>();
    st->type = SType::Combiner;
    return {ITensor(is, st), c};
  }
  Arrow cdir = Out;
  if (args.defined("IndexDir")) {
    cdir = Arrow(args.getInt("IndexDir"));
  } else {
    bool allin = true;
    for (auto &i : inds)
      if (i.dir() != In) allin = false;
    if (allin) cdir = In;
  }
  auto qc = std::make_shared<QCombInfo>();
  qc->inds = inds;
  std::vector<std::pair<QN, long>> qms;
  const int r = int(inds.size());
  Block I(r, 0);
  while (true) {
    QN q;
    long m = 1;
    for (int j = 0; j < r; ++j) {
      q = q + inds[j].qn0(I[j]) * inds[j].dir() * cdir;
      m *= inds[j].blocksize0(I[j]);
    }
    size_t n = 0;
    for (; n < qms.size(); ++n)
      if (qms[n].first == q) break;
    if (n < qms.size()) {
      qc->ranges[I] = {int(n), qms[n].second, qms[n].second + m};
      qms[n].second += m;
    } else {
      qc->ranges[I] = {int(qms.size()), 0, m};
      qms.push_back({q, m});
    }
    int j = 0;
    for (; j < r; ++j) {
      if (++I[j] < inds[j].nblock()) break;
      I[j] = 0;
    }
    if (j == r) break;
  }
  qc->nsectors = int(qms.size());
  Index c(qms, cdir, tags);
  IndexSet is{c};
  for (auto const &i : inds) is.push_back(dag(i));
  auto st = std::make_shared<Storage>();
  st->type = SType::QCombiner;
  st->qc = qc;
  return {ITensor(is, st), c};
}

ITensor denseFromHost(IndexSet const &is, std::vector<double> const &data) {
  if (long(data.size()) != dim(is)) throw ITError("denseFromHost: size mismatch");
  auto st = mkDense(data.size());
  if (!data.empty())
    check(itx_memcpy_htod(ctx(), st->data->p, data.data(), sizeof(double) * data.size()), "itx_memcpy_htod");
  return ITensor(is, st);
}

ITensor qdenseFromHost(IndexSet const &is, BlockOffsets const &offs, std::vector<double> const &store) {
  int64_t cur = 0;
  for (size_t k = 0; k < offs.size(); ++k) {
    if (offs[k].offset != cur) throw ITError("qdenseFromHost: offsets are not running block sizes");
    if (k && !blockLess(offs[k - 1].block, offs[k].block)) throw ITError("qdenseFromHost: blocks not sorted");
    cur += prod(blockDims(is, offs[k].block));
  }
  if (cur != int64_t(store.size())) throw ITError("qdenseFromHost: store size mismatch");
  auto st = mkQDense(offs, store.size(), false);
  if (!store.empty())
    check(itx_memcpy_htod(ctx(), st->data->p, store.data(), sizeof(double) * store.size()), "itx_memcpy_htod");
  return ITensor(is, st);
}

ITensor qdenseZeros(IndexSet const &is, QN const &dv) {
  auto [offs, n] = getBlockOffsets(is, dv);
  return ITensor(is, mkQDense(offs, size_t(n), true));
}

std::vector<double> storageToHost(ITensor const &T) {
  std::vector<double> h(T.store().size);
  if (!h.empty())
    check(itx_memcpy_dtoh(ctx(), h.data(), T.store().data->p, sizeof(double) * h.size()), "itx_memcpy_dtoh");
  return h;
}

// SplitMix64 + Box-Muller (the oracle draws the same stream)
void randomize(ITensor &T, uint64_t seed) {
  T.detach();
  std::vector<double> h(T.store().size);
  uint64_t s = seed;
  auto next = [&s]() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  auto unif = [&]() { return (double(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); };
  for (auto &x : h) {
    double u1 = unif(), u2 = unif();
    x = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
  if (!h.empty())
    check(itx_memcpy_htod(ctx(), T.store().data->p, h.data(), sizeof(double) * h.size()), "itx_memcpy_htod");
}

// ---------------------------------------------- batched vector helpers --
// Same index order (with arrows) and, for QDense, the same block list:
// the storages are then aligned element by element.
static bool alignedStorage(ITensor const &a, ITensor const &b) {
  if (!a || !b || a.isCombiner() || b.isCombiner() || a.isQN() != b.isQN()) return false;
  if (a.order() != b.order() || a.store().size != b.store().size) return false;
  for (int i = 0; i < a.order(); ++i) {
    Index const &x = a.inds()[size_t(i)], &y = b.inds()[size_t(i)];
    if (!(x == y) || x.dir() != y.dir()) return false;
  }
  return !a.isQN() || sameOffsets(a.store().offsets, b.store().offsets);
}

ITensor lincomb(std::vector<ITensor const *> const &V, std::vector<Real> const &c) {
  if (V.empty() || V.size() != c.size()) throw ITError("lincomb: empty or mismatched input");
  bool aligned = V.size() <= 16;
  for (size_t k = 1; k < V.size() && aligned; ++k) aligned = alignedStorage(*V[0], *V[k]);
  if (!aligned) {
    ITensor y = c[0] * (*V[0]);
    for (size_t k = 1; k < V.size(); ++k) y += c[k] * (*V[k]);
    return y;
  }
  auto st = std::make_shared<Storage>(V[0]->store());
  st->data = devAlloc(st->size);
  std::vector<const double *> p(V.size());
  for (size_t k = 0; k < V.size(); ++k) p[k] = V[k]->store().data->p;
  if (st->size) check(itx_lincomb(ctx(), p.data(), c.data(), int(V.size()), st->data->p, st->size), "itx_lincomb");
  return ITensor(V[0]->inds(), st);
}

std::vector<Real> dotDagMany(std::vector<ITensor const *> const &V, ITensor const &q) {
  HostTimer ht(HT_NORM_DOT);
  std::vector<Real> out(V.size(), 0.0);
  bool aligned = V.size() <= 16;
  for (size_t k = 0; k < V.size() && aligned; ++k) aligned = alignedStorage(*V[k], q);
  if (!aligned) {
    for (size_t k = 0; k < V.size(); ++k) out[k] = dotDag(*V[k], q);
    return out;
  }
  if (V.empty() || q.store().size == 0) return out;
  std::vector<const double *> p(V.size());
  for (size_t k = 0; k < V.size(); ++k) p[k] = V[k]->store().data->p;
  check(itx_multi_dot(ctx(), p.data(), int(V.size()), q.store().data->p, q.store().size, out.data()),
        "itx_multi_dot");
  return out;
}

// randomITensor(QN, IndexSet) / randomITensor(IndexSet) (itensor.h:969-975):
// every flux-allowed block (or the dense tensor) filled with N(0,1) from
// the seeded SplitMix64 + Box-Muller stream
ITensor randomITensor(QN const &div, IndexSet const &is, uint64_t seed) {
  ITensor T = hasQNs(is) ? qdenseZeros(is, div) : ITensor(is, mkDense(size_t(dim(is))));
  randomize(T, seed);
  return T;
}

// QDense tensor of flux div holding the entries of a dense column-major
// array; entries outside the flux-allowed blocks must be zero
ITensor qdenseFromDense(IndexSet const &is, std::vector<double> const &colmajor, QN const &div) {
  if (long(colmajor.size()) != dim(is)) throw ITError("qdenseFromDense: size mismatch");
  auto [offs, n] = getBlockOffsets(is, div);
  std::vector<double> store(size_t(n), 0.0);
  std::vector<char> used(colmajor.size(), 0);
  const size_t r = is.size();
  auto strides = colStrides(dimsOf(is));
  for (auto const &bo : offs) {
    auto bd = blockDims(is, bo.block);
    std::vector<int64_t> start(r);
    for (size_t j = 0; j < r; ++j) start[j] = is[j].blockStart(bo.block[j]);
    const int64_t nb = prod(bd);
    std::vector<int64_t> i(r, 0);
    for (int64_t e = 0; e < nb; ++e) {
      int64_t src = 0;
      for (size_t j = 0; j < r; ++j) src += (start[j] + i[j]) * strides[j];
      store[size_t(bo.offset + e)] = colmajor[size_t(src)];
      used[size_t(src)] = 1;
      for (size_t j = 0; j < r; ++j) {
        if (++i[j] < bd[j]) break;
        i[j] = 0;
      }
    }
  }
  for (size_t k = 0; k < colmajor.size(); ++k)
    if (!used[k] && colmajor[k] != 0.0) throw ITError("qdenseFromDense: nonzero entry outside the flux-allowed blocks");
  return qdenseFromHost(is, offs, store);
}

}  // namespace itensor
