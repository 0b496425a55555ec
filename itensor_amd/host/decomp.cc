// truncate, diag_hermitian and denmatDecomp of the host layer.  The
// eigen-decompositions run on the device (itx_syev / itx_syev_batched);
// the truncation decision is host-side integer/scalar work, as in the
// reference.
#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

#include "itensor.h"

namespace itensor {

namespace {
constexpr double REAL_EPSILON = std::numeric_limits<double>::epsilon();

std::vector<int64_t> bdims(IndexSet const &is, Block const &b) {
  std::vector<int64_t> d(is.size());
  for (size_t i = 0; i < is.size(); ++i) d[i] = is[i].blocksize0(b[i]);
  return d;
}
}  // namespace

// truncate (decomp.cc:305-463)
std::tuple<Real, Real, Real, int> truncate(std::vector<Real> &P, long maxdim, long mindim, Real cutoff,
                                           bool absoluteCutoff, bool doRelCutoff, Args const &args) {
  const bool respectDegenerate = args.getBool("RespectDegenerate", false);
  const long origm = long(P.size());
  long n = origm - 1;
  if (P[0] == 0.0) {
    P.resize(1);
    return {0.0, -REAL_EPSILON, REAL_EPSILON, 1};
  }
  if (origm == 1) {
    const double dc = 1e-3 * P[0];
    return {0.0, P[0] - dc, P[0] + dc, 1};
  }
  const double P0 = P[0];
  for (auto &x : P) x /= P0;
  if (absoluteCutoff || !doRelCutoff) cutoff /= P0;
  for (long zn = n; zn >= 0; --zn) {
    if (P[zn] >= 0) break;
    P[zn] = 0;
  }
  double truncerr = 0;
  while (n >= maxdim) {
    truncerr += P[n];
    --n;
  }
  if (absoluteCutoff) {
    for (; P[n] <= cutoff && n >= mindim; --n) truncerr += P[n];
  } else {
    double scale = 1.0;
    if (doRelCutoff) {
      scale = std::accumulate(P.begin(), P.end(), 0.0);
      if (scale == 0.0) scale = 1.0;
    }
    while (truncerr + P[n] <= cutoff * scale && n >= mindim) {
      truncerr += P[n];
      --n;
    }
    truncerr = (scale == 0 ? 0 : truncerr / scale);
  }
  if (n < 0) n = 0;
  long m = n + 1;
  double degen_cutoff = 1e-3 * P[n];
  if (P[n] == 0.0) degen_cutoff = REAL_EPSILON;
  double docut_lower = P[n] - degen_cutoff, docut_upper = P[n] + degen_cutoff;
  int ndegen_below = 1;
  if (n > 0)
    while ((n - ndegen_below >= 0) && (std::abs(P[n - ndegen_below] - P[n]) < degen_cutoff)) ++ndegen_below;
  int ndegen_above = 0;
  if (n + 1 < origm)
    while ((n + 1 + ndegen_above < origm) && (std::abs(P[n + 1 + ndegen_above] - P[n]) < degen_cutoff))
      ++ndegen_above;
  if (respectDegenerate && (ndegen_below + ndegen_above > 1)) {
    if (maxdim < m + ndegen_above) {
      m -= ndegen_below;
      ndegen_above += ndegen_below;
      ndegen_below = 0;
    } else {
      m += ndegen_above;
      ndegen_below += ndegen_above;
      ndegen_above = 0;
    }
  }
  P.resize(size_t(m));
  for (auto &x : P) x *= P0;
  return {truncerr * P0, docut_lower * P0, docut_upper * P0, ndegen_below};
}

// diag_hermitian (hermitian.cc:40-440), real symmetric input
Spectrum diag_hermitian(ITensor H, ITensor &U, ITensor &D, Args const &args_in) {
  HostTimer ht(HT_DIAG_HERMITIAN);
  Args args = args_in;
  if (H.order() != 2) throw ITError("diag_hermitian requires an order 2 tensor");
  Index i1 = H.inds().front(), i2 = H.inds().back();
  Index ai = (i1.primeLevel() < i2.primeLevel()) ? i1 : i2;
  const int pdiff = std::abs(i1.primeLevel() - i2.primeLevel());
  const long origdim = ai.dim();
  const double cutoff = args.getReal("Cutoff", 0.0);
  long maxdim = args.getInt("MaxDim", origdim);
  const long mindim = args.getInt("MinDim", 1);
  const bool def_do_trunc = args.defined("Cutoff") || args.defined("MaxDim");
  const bool do_truncate = args.getBool("Truncate", def_do_trunc);
  const bool doRelCutoff = args.getBool("DoRelCutoff", true);
  const bool absoluteCutoff = args.getBool("AbsoluteCutoff", false);
  const std::string tags = args.getString("Tags", "Link");
  if (!do_truncate && args.defined("MaxDim")) maxdim = origdim;

  if (!H.isQN()) {
    const long n = ai.dim();
    // dsyev on -M: the device solver returns M's eigenvalues descending
    // (tensor/algs_impl.h:122-135).  H is symmetric, so its storage order
    // (ai first or ai' first) gives the same matrix.
    auto W = devAlloc(size_t(n * n));
    auto w = devAlloc(size_t(n));
    check(itx_memcpy_dtod(ctx(), W->p, H.store().data->p, sizeof(double) * size_t(n * n)), "itx_memcpy_dtod");
    check(itx_syev(ctx(), int(n), W->p, w->p), "itx_syev");
    std::vector<double> DD(static_cast<size_t>(n));
    check(itx_memcpy_dtoh(ctx(), DD.data(), w->p, sizeof(double) * size_t(n)), "itx_memcpy_dtoh");
    Real truncerr = 0;
    if (do_truncate) truncerr = std::get<0>(truncate(DD, maxdim, mindim, cutoff, absoluteCutoff, doRelCutoff, args));
    const long m = long(DD.size());
    if (m > maxdim) throw ITError("diag_hermitian: m > maxdim");
    Index newmid(m, tags);
    auto st = std::make_shared<Storage>();
    st->type = SType::Dense;
    st->size = size_t(n * m);
    st->data = devAlloc(st->size);
    check(itx_memcpy_dtod(ctx(), st->data->p, W->p, sizeof(double) * st->size), "itx_memcpy_dtod");
    U = ITensor({ai, newmid}, st);
    std::vector<double> dd(size_t(m * m), 0.0);
    for (long k = 0; k < m; ++k) dd[size_t(k * m + k)] = DD[size_t(k)];
    D = denseFromHost({prime(newmid, pdiff), newmid}, dd);
    return Spectrum{DD, truncerr};
  }

  // --- QN: per-block eigen-decompositions, global truncation ---
  const int pa = indexPosition(H.inds(), ai);
  auto const &offs = H.store().offsets;
  const int nb = int(offs.size());
  std::vector<int32_t> ns(nb);
  std::vector<int64_t> aoffs(nb), woffs(nb);
  std::vector<int> sec(nb);
  int64_t wtot = 0;
  for (int b = 0; b < nb; ++b) {
    auto d = bdims(H.inds(), offs[b].block);
    if (d[0] != d[1]) throw ITError("diag_hermitian: non-square QN block");
    ns[b] = int32_t(d[0]);
    aoffs[b] = offs[b].offset;
    woffs[b] = wtot;
    wtot += d[0];
    sec[b] = offs[b].block[pa];
  }
  auto W = devAlloc(H.store().size);
  auto w = devAlloc(size_t(std::max<int64_t>(wtot, 1)));
  if (H.store().size)
    check(itx_memcpy_dtod(ctx(), W->p, H.store().data->p, sizeof(double) * H.store().size), "itx_memcpy_dtod");
  // eigenvalues straight to the host: the truncation below does not wait
  // for the eigenvector kernels, which finish on the stream meanwhile
  std::vector<double> all(static_cast<size_t>(wtot));
  if (nb) {
    HostTimer ht(HT_SYEV);
    check(itx_syev_batched_host(ctx(), nb, ns.data(), aoffs.data(), W->p, woffs.data(), w->p, all.data()),
          "itx_syev_batched_host");
  }
  std::vector<double> probs = all;
  std::sort(probs.begin(), probs.end(), std::greater<double>());
  Real truncerr = 0, docut_lower = -1, docut_upper = -1;
  int ndegen = 0;
  long m = long(probs.size());
  if (do_truncate && !probs.empty()) {
    std::tie(truncerr, docut_lower, docut_upper, ndegen) =
        truncate(probs, maxdim, mindim, cutoff, absoluteCutoff, doRelCutoff, args);
    m = long(probs.size());
  }
  if (m > maxdim) throw ITError("diag_hermitian: m > maxdim");
  QNBlocks iq;
  std::vector<long> keep(nb, 0);
  long total_m = 0;
  for (int b = 0; b < nb; ++b) {
    long this_m = 0;
    const double *d = all.data() + woffs[b];
    if (do_truncate) {
      while (this_m < ns[b] && total_m < m && d[this_m] > docut_upper) {
        ++this_m;
        ++total_m;
      }
      while (ndegen > 0 && this_m < ns[b] && total_m < m && d[this_m] > docut_lower) {
        ++this_m;
        ++total_m;
        --ndegen;
      }
    } else {
      this_m = ns[b];
      total_m += this_m;
    }
    keep[b] = this_m;
    if (this_m > 0) iq.push_back({ai.qn0(sec[b]), this_m});
  }
  if (iq.empty()) {
    if (nb == 0) throw ITError("No blocks in diag_hermitian");
    iq.push_back({ai.qn0(sec[0]), 1});
  }
  Index dind(iq, Arrow(-int(ai.dir())), tags);
  IndexSet Uis{dag(ai), dag(dind)};
  auto [uoffs, usize] = getBlockOffsets(Uis, QN());
  auto ust = std::make_shared<Storage>();
  ust->type = SType::QDense;
  ust->offsets = uoffs;
  ust->size = size_t(usize);
  ust->data = devZeros(ust->size);
  std::vector<int64_t> desc;
  int n = 0, ntask = 0;
  std::vector<double> dkept;
  for (int b = 0; b < nb; ++b) {
    if (keep[b] == 0) continue;
    Block ub{sec[b], n};
    auto it = std::lower_bound(uoffs.begin(), uoffs.end(), ub,
                               [](BlOf const &x, Block const &y) { return blockLess(x.block, y); });
    if (it == uoffs.end() || it->block != ub) throw ITError("diag_hermitian: U block missing");
    const int64_t cnt = int64_t(ns[b]) * keep[b];
    desc.insert(desc.end(), {aoffs[b], it->offset, cnt, 1, 1});
    ++ntask;
    for (long k = 0; k < keep[b]; ++k) dkept.push_back(all[size_t(woffs[b] + k)]);
    ++n;
  }
  if (ntask) check(itx_block_copy(ctx(), W->p, ust->data->p, ntask, 1, desc.data(), 1.0, 0.0), "itx_block_copy");
  U = ITensor(Uis, ust);
  // D as a block-diagonal QDense over (prime(d,pdiff), dag(d))
  IndexSet Dis{prime(dind, pdiff), dag(dind)};
  auto [doffs, dsize] = getBlockOffsets(Dis, QN());
  std::vector<double> dh(size_t(dsize), 0.0);
  size_t kk = 0;
  for (auto const &bo : doffs) {
    if (bo.block[0] != bo.block[1]) continue;
    const long mm = dind.blocksize0(bo.block[0]);
    for (long k = 0; k < mm; ++k) dh[size_t(bo.offset + k * mm + k)] = dkept[kk++];
  }
  D = qdenseFromHost(Dis, doffs, dh);
  return Spectrum{probs, truncerr};
}

// ---------------------------------------------------------------- svd --
namespace {
constexpr double MIN_CUT = 1e-15;  // global.h:61
constexpr long MAX_DIM = 5000;    // global.h:62

// Keep counts per block from the global truncation of all singular values
// (svd.cc:281-330; the same rule as diag_hermitian, on squared values).
std::vector<long> keepPerBlock(std::vector<std::vector<double>> const &s, bool do_truncate, long m,
                               double docut_lower, double docut_upper, int ndegen) {
  std::vector<long> keep(s.size(), 0);
  long total_m = 0;
  for (size_t b = 0; b < s.size(); ++b) {
    long this_m = 0;
    const long nb = long(s[b].size());
    if (do_truncate) {
      while (this_m < nb && total_m < m && s[b][size_t(this_m)] * s[b][size_t(this_m)] > docut_upper) {
        ++this_m;
        ++total_m;
      }
      while (ndegen > 0 && this_m < nb && total_m < m &&
             s[b][size_t(this_m)] * s[b][size_t(this_m)] > docut_lower) {
        ++this_m;
        ++total_m;
        --ndegen;
      }
    } else {
      this_m = nb;
      total_m += this_m;
    }
    keep[b] = this_m;
  }
  return keep;
}

// svdOrd2 -> svdImpl (svd.cc:39-436) for an order-2 tensor over (ui, vi)
Spectrum svdOrd2(ITensor const &A, Index const &ui, Index const &vi, ITensor &U, ITensor &D, ITensor &V,
                 Args const &args) {
  if (A.order() != 2) throw ITError("svd: A must be matrix-like (order 2)");
  const bool do_truncate = args.getBool("Truncate", args.defined("Cutoff") || args.defined("MaxDim"));
  const double cutoff = args.getReal("Cutoff", MIN_CUT);
  const long maxdim = args.getInt("MaxDim", MAX_DIM);
  const long mindim = args.getInt("MinDim", 1);
  const bool doRelCutoff = args.getBool("DoRelCutoff", true);
  const bool absoluteCutoff = args.getBool("AbsoluteCutoff", false);
  const std::string ltags = args.getString("LeftTags", "Link,U");
  const std::string rtags = args.getString("RightTags", "Link,V");
  if (ltags == rtags) throw ITError("svd: LeftTags and RightTags must differ");
  const int pu = indexPosition(A.inds(), ui);
  if (pu < 0 || indexPosition(A.inds(), vi) < 0) throw ITError("svd: indices not found");

  if (!A.isQN()) {
    const long m = ui.dim(), n = vi.dim(), k = std::min(m, n);
    auto Ub = devAlloc(size_t(m * k)), Vb = devAlloc(size_t(n * k)), Sb = devAlloc(size_t(std::max(k, 1L)));
    // storage is (ui, vi) column-major when pu == 0, else the transpose
    if (pu == 0)
      check(itx_gesvd(ctx(), int(m), int(n), A.store().data->p, Ub->p, Sb->p, Vb->p), "itx_gesvd");
    else
      check(itx_gesvd(ctx(), int(n), int(m), A.store().data->p, Vb->p, Sb->p, Ub->p), "itx_gesvd");
    std::vector<double> S(static_cast<size_t>(k));
    if (k) check(itx_memcpy_dtoh(ctx(), S.data(), Sb->p, sizeof(double) * size_t(k)), "itx_memcpy_dtoh");
    std::vector<double> probs(S.size());
    for (size_t j = 0; j < S.size(); ++j) probs[j] = S[j] * S[j];
    Real truncerr = 0;
    if (do_truncate && !probs.empty())
      truncerr = std::get<0>(truncate(probs, maxdim, mindim, cutoff, absoluteCutoff, doRelCutoff, args));
    const long mk = long(probs.size());
    Index uL(mk, ltags), vL(mk, rtags);
    auto cols = [&](DevPtr const &src, long rows) {
      auto st = std::make_shared<Storage>();
      st->type = SType::Dense;
      st->size = size_t(rows * mk);
      st->data = devAlloc(st->size);
      if (st->size)
        check(itx_memcpy_dtod(ctx(), st->data->p, src->p, sizeof(double) * st->size), "itx_memcpy_dtod");
      return st;
    };
    U = ITensor({ui, uL}, cols(Ub, m));
    V = ITensor({vi, vL}, cols(Vb, n));
    std::vector<double> dd(size_t(mk * mk), 0.0);
    for (long j = 0; j < mk; ++j) dd[size_t(j * mk + j)] = S[size_t(j)];
    D = denseFromHost({uL, vL}, dd);
    return Spectrum{probs, truncerr};
  }

  // QN: one SVD per block, global truncation (svd.cc:157-423)
  auto const &offs = A.store().offsets;
  const int nb = int(offs.size());
  if (nb == 0) throw ITError("svd: QN ITensor has no blocks");
  std::vector<int32_t> bm(nb), bn(nb);
  std::vector<int64_t> ao(nb), po(nb), qo(nb), so(nb);
  std::vector<int> b1(nb), b2(nb);
  int64_t ptot = 0, qtot = 0, stot = 0;
  for (int b = 0; b < nb; ++b) {
    auto const &blk = offs[b].block;
    b1[b] = blk[pu];
    b2[b] = blk[1 - pu];
    bm[b] = int32_t(A.inds()[0].blocksize0(blk[0]));
    bn[b] = int32_t(A.inds()[1].blocksize0(blk[1]));
    const int64_t k = std::min(bm[b], bn[b]);
    ao[b] = offs[b].offset;
    po[b] = ptot;
    qo[b] = qtot;
    so[b] = stot;
    ptot += int64_t(bm[b]) * k;
    qtot += int64_t(bn[b]) * k;
    stot += k;
  }
  auto P = devAlloc(size_t(std::max<int64_t>(ptot, 1))), Q = devAlloc(size_t(std::max<int64_t>(qtot, 1)));
  auto Sb = devAlloc(size_t(std::max<int64_t>(stot, 1)));
  check(itx_gesvd_batched(ctx(), nb, bm.data(), bn.data(), ao.data(), A.store().data->p, po.data(), P->p,
                          so.data(), Sb->p, qo.data(), Q->p),
        "itx_gesvd_batched");
  std::vector<double> all(static_cast<size_t>(stot));
  if (stot) check(itx_memcpy_dtoh(ctx(), all.data(), Sb->p, sizeof(double) * size_t(stot)), "itx_memcpy_dtoh");
  std::vector<std::vector<double>> sv(static_cast<size_t>(nb));
  std::vector<double> probs;
  for (int b = 0; b < nb; ++b) {
    const int64_t k = std::min(bm[b], bn[b]);
    sv[size_t(b)].assign(all.begin() + so[b], all.begin() + so[b] + k);
    for (double x : sv[size_t(b)]) probs.push_back(x * x);
  }
  std::sort(probs.begin(), probs.end(), std::greater<double>());
  Real truncerr = 0, docut_lower = -1, docut_upper = -1;
  int ndegen = 1;
  long m = long(probs.size());
  if (do_truncate && !probs.empty()) {
    std::tie(truncerr, docut_lower, docut_upper, ndegen) =
        truncate(probs, maxdim, mindim, cutoff, absoluteCutoff, doRelCutoff, args);
    m = long(probs.size());
  }
  auto keep = keepPerBlock(sv, do_truncate, m, docut_lower, docut_upper, ndegen);
  QNBlocks Liq, Riq;
  for (int b = 0; b < nb; ++b)
    if (keep[size_t(b)] > 0) {
      Liq.push_back({ui.qn0(b1[b]), keep[size_t(b)]});
      Riq.push_back({vi.qn0(b2[b]), keep[size_t(b)]});
    }
  if (Liq.empty()) throw ITError("svd: new index of U is empty");
  Index L(Liq, ui.dir(), ltags), R(Riq, vi.dir(), rtags);
  IndexSet Uis{ui, dag(L)}, Vis{vi, dag(R)}, Dis{L, R};
  auto build = [&](IndexSet const &is, bool isU) {
    auto [bo, size] = getBlockOffsets(is, QN());
    auto st = std::make_shared<Storage>();
    st->type = SType::QDense;
    st->offsets = bo;
    st->size = size_t(size);
    st->data = devZeros(st->size);
    // the factor of our U (V) is P when ui (vi) is the storage's first
    // index, else Q; kept columns are a contiguous prefix of each block
    const bool fromP = (isU == (pu == 0));
    std::vector<int64_t> desc;
    int ntask = 0;
    for (int b = 0, nk = 0; b < nb; ++b) {
      if (keep[size_t(b)] == 0) continue;
      Block want{isU ? b1[b] : b2[b], nk++};
      auto it = std::lower_bound(bo.begin(), bo.end(), want,
                                 [](BlOf const &x, Block const &y) { return blockLess(x.block, y); });
      if (it == bo.end() || it->block != want) throw ITError("svd: factor block missing");
      const int64_t rows = fromP ? bm[b] : bn[b];
      desc.insert(desc.end(), {fromP ? po[b] : qo[b], it->offset, rows * keep[size_t(b)], 1, 1});
      ++ntask;
    }
    if (ntask)
      check(itx_block_copy(ctx(), (fromP ? P : Q)->p, st->data->p, ntask, 1, desc.data(), 1.0, 0.0),
            "itx_block_copy");
    return ITensor(is, st);
  };
  U = build(Uis, true);
  V = build(Vis, false);
  // D holds A's divergence (the reference's QDiagReal over (L, R) has the
  // diagonal blocks whatever the flux; with U and V flux-free they are
  // exactly the blocks of flux div(A))
  auto [doffs, dsize] = getBlockOffsets(Dis, div(A));
  std::vector<double> dh(size_t(dsize), 0.0);
  for (int b = 0, nk = 0; b < nb; ++b) {
    if (keep[size_t(b)] == 0) continue;
    Block want{nk, nk};
    auto it = std::lower_bound(doffs.begin(), doffs.end(), want,
                               [](BlOf const &x, Block const &y) { return blockLess(x.block, y); });
    if (it == doffs.end() || it->block != want) throw ITError("svd: D block missing");
    const long mm = keep[size_t(b)];
    for (long j = 0; j < mm; ++j) dh[size_t(it->offset + j * mm + j)] = sv[size_t(b)][size_t(j)];
    ++nk;
  }
  D = qdenseFromHost(Dis, doffs, dh);
  return Spectrum{probs, truncerr};
}
}  // namespace

// svd (decomp.cc:121-248): the indices of U (or, if U has none, of V)
// are combined into ui, the rest into vi; U and V come back uncombined
Spectrum svd(ITensor const &AA, ITensor &U, ITensor &D, ITensor &V, Args args) {
  if (args.getReal("Noise", 0) > 0) throw ITError("svd: noise term not implemented");
  ITensor const &Lt = U.order() > 0 ? U : V;
  if (Lt.order() == 0) throw ITError("svd: U and V carry no indices");
  IndexSet Uinds, Vinds;
  for (auto const &I : AA.inds()) {
    const bool inL = indexPosition(Lt.inds(), I) >= 0;
    ((inL == (U.order() > 0)) ? Uinds : Vinds).push_back(I);
  }
  ITensor AAc = AA, Ucomb, Vcomb;
  Index ui, vi;
  if (!Uinds.empty()) {
    std::tie(Ucomb, ui) = combiner(Uinds);
    AAc *= Ucomb;
  }
  if (!Vinds.empty()) {
    std::tie(Vcomb, vi) = combiner(Vinds);
    AAc *= Vcomb;
  }
  if (!ui || !vi) throw ITError("svd: U or V would carry no index of AA");
  auto spec = svdOrd2(AAc, ui, vi, U, D, V, args);
  U = dag(Ucomb) * U;
  V = V * dag(Vcomb);
  return spec;
}

std::tuple<ITensor, ITensor, ITensor> svd(ITensor const &AA, IndexSet const &Uis, Args const &args) {
  ITensor U(Uis, nullptr), D, V;
  svd(AA, U, D, V, args);
  return {U, D, V};
}

// factor (decomp.cc:538-576): svd with LeftTags = Tags (default "Link"),
// sqrt(D) absorbed into both factors, which then share the left link
Spectrum factor(ITensor const &T, ITensor &A, ITensor &B, Args args) {
  args.add("LeftTags", args.getString("Tags", "Link"));
  ITensor D;
  auto spec = svd(T, A, D, B, args);
  Index L = D.inds()[0];
  auto h = storageToHost(D);
  for (auto &x : h) x = std::sqrt(std::abs(x));
  ITensor Dsq = D.isQN() ? qdenseFromHost(D.inds(), D.store().offsets, h) : denseFromHost(D.inds(), h);
  // Z(L, dag(L)') = diag(sqrt s): A(.., dag(L)) * Z, then L' back to L
  Index Lp = prime(dag(L));
  IndexSet Zis{L, Lp};
  ITensor Z;
  if (D.isQN()) {
    auto [zoffs, zsize] = getBlockOffsets(Zis, QN());
    std::vector<double> zh(size_t(zsize), 0.0);
    for (auto const &zb : zoffs) {
      if (zb.block[0] != zb.block[1]) continue;
      Block db{zb.block[0], zb.block[0]};
      auto it = std::lower_bound(D.store().offsets.begin(), D.store().offsets.end(), db,
                                 [](BlOf const &x, Block const &y) { return blockLess(x.block, y); });
      if (it == D.store().offsets.end() || it->block != db) continue;
      const long mm = L.blocksize0(zb.block[0]);
      for (long j = 0; j < mm; ++j) zh[size_t(zb.offset + j * mm + j)] = h[size_t(it->offset + j * mm + j)];
    }
    Z = qdenseFromHost(Zis, zoffs, zh);
  } else {
    Z = denseFromHost(Zis, h);
  }
  A = A * Z;
  A.prime(Lp, -1);
  B = Dsq * B;
  return spec;
}

std::tuple<ITensor, ITensor> factor(ITensor const &T, IndexSet const &Ais, Args const &args) {
  ITensor A(Ais, nullptr), B;
  factor(T, A, B, args);
  return {A, B};
}

// denmatDecomp (decomp.h:316-418)
Spectrum denmatDecomp(ITensor const &AA, ITensor &A, ITensor &B, Direction dir, Args args,
                      DeltaRhoFn const &deltaRho) {
  HostTimer ht(HT_DENMAT);
  const double noise = args.getReal("Noise", 0.0);
  if (noise > 0 && !deltaRho) throw ITError("denmatDecomp: non-zero noise needs the local Hamiltonian");
  if (!args.defined("Tags")) args.add("Tags", std::string("Link"));
  Index mid = commonIndex(A, B);
  if (dir == NoDir) dir = (mid.dir() == Out ? Fromright : Fromleft);
  ITensor &to_orth = (dir == Fromleft ? A : B);
  ITensor &newoc = (dir == Fromleft ? B : A);
  IndexSet const &active = (to_orth ? to_orth : AA).inds();
  IndexSet cinds;
  for (auto const &I : active)
    if (!hasIndex(newoc, I)) cinds.push_back(I);
  auto [cmb, ci] = combiner(cinds, args);
  ITensor AAc = cmb * AA;
  ITensor rho = AAc * dag(prime(AAc, ci));
  if (noise > 0) {
    // rho += noise * deltaRho, renormalised to unit trace; the trace of
    // AAc AAc^dag is |AAc|^2 and deltaRho reports its own
    auto [drho, trd] = deltaRho(AA, cmb, dir);
    rho += noise * drho;
    const double nrm = norm(AAc), tr = nrm * nrm + noise * trd;
    if (tr > 1e-16) rho *= 1.0 / tr;
  }
  ITensor U, D;
  auto spec = diag_hermitian(rho, U, D, args);
  cmb.dag();
  to_orth = cmb * dag(U);
  newoc = U * AAc;
  return spec;
}

}  // namespace itensor
