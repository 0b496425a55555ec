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
  if (nb) check(itx_syev_batched(ctx(), nb, ns.data(), aoffs.data(), W->p, woffs.data(), w->p), "itx_syev_batched");
  std::vector<double> all(static_cast<size_t>(wtot));
  if (wtot) check(itx_memcpy_dtoh(ctx(), all.data(), w->p, sizeof(double) * size_t(wtot)), "itx_memcpy_dtoh");
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

// denmatDecomp, noise = 0 (decomp.h:316-418)
Spectrum denmatDecomp(ITensor const &AA, ITensor &A, ITensor &B, Direction dir, Args args) {
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
  ITensor U, D;
  auto spec = diag_hermitian(rho, U, D, args);
  cmb.dag();
  to_orth = cmb * dag(U);
  newoc = U * AAc;
  return spec;
}

}  // namespace itensor
