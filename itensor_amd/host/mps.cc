// LocalOp::size, LocalMPO::position, svdBond, the DMRG driver and the small
// host eigensolver used by davidson's Rayleigh-Ritz step.
#include <cmath>

#include "mps.h"

namespace itensor {

namespace detail {
// cyclic Jacobi for a tiny real symmetric matrix (n <= ~8)
void smallSymEig(int n, std::vector<double> A, std::vector<double> &w, std::vector<double> &V) {
  V.assign(size_t(n * n), 0.0);
  for (int i = 0; i < n; ++i) V[size_t(i * n + i)] = 1.0;
  auto a = [&](int r, int c) -> double & { return A[size_t(r * n + c)]; };
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += a(p, q) * a(p, q);
    if (off < 1e-300) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        if (a(p, q) == 0.0) continue;
        const double theta = (a(q, q) - a(p, p)) / (2.0 * a(p, q));
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::abs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  // A <- A J
          const double akp = a(k, p), akq = a(k, q);
          a(k, p) = c * akp - s * akq;
          a(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // A <- J^T A
          const double apk = a(p, k), aqk = a(q, k);
          a(p, k) = c * apk - s * aqk;
          a(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {  // V <- V J
          const double vkp = V[size_t(k * n + p)], vkq = V[size_t(k * n + q)];
          V[size_t(k * n + p)] = c * vkp - s * vkq;
          V[size_t(k * n + q)] = s * vkp + c * vkq;
        }
      }
  }
  // sort ascending
  std::vector<int> idx(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) idx[size_t(i)] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return a(x, x) < a(y, y); });
  w.resize(size_t(n));
  std::vector<double> Vs(static_cast<size_t>(n * n));
  for (int j = 0; j < n; ++j) {
    w[size_t(j)] = a(idx[size_t(j)], idx[size_t(j)]);
    for (int k = 0; k < n; ++k) Vs[size_t(k * n + j)] = V[size_t(k * n + idx[size_t(j)])];
  }
  V.swap(Vs);
}
}  // namespace detail

// LocalOp::size (localop.h:488-530): dim of the first primed index of L
// and of R, times the dims of the unprimed site indices of Op1 and Op2
long LocalOp::size() const {
  long s = 1;
  auto firstPrimed = [&](ITensor const *T) {
    for (auto const &I : T->inds())
      if (I.primeLevel() > 0) {
        s *= I.dim();
        return;
      }
  };
  auto site0 = [&](ITensor const *T) {
    for (auto const &I : T->inds())
      if (I.primeLevel() == 0 && I.hasTag("Site")) {
        s *= I.dim();
        return;
      }
    throw ITError("LocalOp::size: operator has no unprimed Site index");
  };
  if (!LIsNull()) firstPrimed(L_);
  if (!RIsNull()) firstPrimed(R_);
  site0(Op1_);
  site0(Op2_);
  return s;
}

namespace {
// result index order of A*B (computeLabels + contractIS), no data
IndexSet productInds(IndexSet const &A, IndexSet const &B) {
  IndexSet N;
  std::vector<bool> bc(B.size(), false);
  std::vector<bool> ac(A.size(), false);
  for (size_t i = 0; i < A.size(); ++i)
    for (size_t j = 0; j < B.size(); ++j)
      if (!bc[j] && A[i] == B[j]) {
        ac[i] = bc[j] = true;
        break;
      }
  for (size_t i = 0; i < A.size(); ++i)
    if (!ac[i]) N.push_back(A[i]);
  for (size_t j = 0; j < B.size(); ++j)
    if (!bc[j]) N.push_back(B[j]);
  return N;
}
bool sameOrder(IndexSet const &a, IndexSet const &b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); ++i)
    if (!(a[i] == b[i])) return false;
  return true;
}
}  // namespace

bool LocalOp::setFused(ITensor const *W12, IndexSet const &phiInds) {
  W12_ = nullptr;
  if (!W12 || LIsNull() || RIsNull()) return false;
  IndexSet t1 = productInds(phiInds, L_->inds());
  IndexSet ref = productInds(productInds(productInds(t1, Op1_->inds()), Op2_->inds()), R_->inds());
  IndexSet fus = productInds(productInds(t1, W12->inds()), R_->inds());
  if (!sameOrder(ref, fus)) return false;
  W12_ = W12;
  return true;
}

void LocalMPO::fuseOps(IndexSet const &phiInds) {
  if (!fuse_ || !lop_.isBulk()) return;
  auto const &H = *H_;
  W12_ = H[size_t(b_ - 1)] * H[size_t(b_)];
  if (!lop_.setFused(&W12_, phiInds)) W12_ = ITensor();
}

void LocalMPO::position(int b, MPS const &psi) {
  auto t0 = std::chrono::steady_clock::now();
  auto const &H = *H_;
  while (LHlim_ < b - 1) {
    const int ll = LHlim_;
    ITensor E = PH_[size_t(ll)] ? PH_[size_t(ll)] * psi(ll + 1) : psi(ll + 1);
    E *= H[size_t(ll)];
    E *= dag(prime(psi(ll + 1)));
    PH_[size_t(ll + 1)] = E;
    LHlim_ = ll + 1;
  }
  while (RHlim_ > b + 2) {
    const int rl = RHlim_;
    ITensor E = PH_[size_t(rl)] ? PH_[size_t(rl)] * psi(rl - 1) : psi(rl - 1);
    E *= H[size_t(rl - 2)];
    E *= dag(prime(psi(rl - 1)));
    PH_[size_t(rl - 1)] = E;
    RHlim_ = rl - 1;
  }
  LHlim_ = b - 1;
  RHlim_ = b + 2;
  lop_.update(&H[size_t(b - 1)], &H[size_t(b)], &PH_[size_t(b - 1)], &PH_[size_t(b + 2)]);
  b_ = b;
  env_seconds_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

Spectrum svdBond(MPS &psi, int b, ITensor const &phi, Direction dir, Args const &args_in) {
  Args args = args_in;
  args.add("RespectDegenerate", args.getBool("RespectDegenerate", true));
  Index oldlink = commonIndex(psi(b), psi(b + 1));
  std::string tags = oldlink ? oldlink.tags() : std::string("Link");
  ITensor A = psi(b), B = psi(b + 1);
  auto spec = denmatDecomp(phi, A, B, dir, args);
  if (args.getBool("DoNormalize", false)) {
    ITensor &oc = (dir == Fromleft ? B : A);
    const double nrm = norm(oc);
    if (nrm > 1e-16) oc *= 1.0 / nrm;
  }
  // put the old tags back onto the new link (mps_impl.h:84-88)
  Index lb = commonIndex(A, B);
  A.setTags(tags, lb);
  B.setTags(tags, lb);
  psi.ref(b) = A;
  psi.ref(b + 1) = B;
  if (dir == Fromleft) {
    psi.leftLim(b);
    if (psi.rightLim() < b + 2) psi.rightLim(b + 2);
  } else {
    if (psi.leftLim() > b - 1) psi.leftLim(b - 1);
    psi.rightLim(b + 1);
  }
  return spec;
}

Real dmrg(MPS &psi, MPO const &H, Sweeps const &sweeps, Args const &args_in,
          std::function<void(BondRecord const &)> obs, std::vector<double> *sweep_seconds) {
  Args args = args_in;
  args.add("RespectDegenerate", args.getBool("RespectDegenerate", true));
  args.add("DoNormalize", true);
  const int N = psi.length();
  if (int(H.size()) != N) throw ITError("dmrg: MPO and MPS lengths differ");
  LocalMPO PH(H);
  PH.setPrecision(int(args.getInt("HpsiPrecision", ITX_PREC_F64)));
  PH.setFusion(args.getBool("FuseMPO", true));
  Real energy = NAN;
  for (size_t sw = 0; sw < sweeps.size(); ++sw) {
    auto const &S = sweeps[sw];
    auto t_sweep = std::chrono::steady_clock::now();
    args.add("Sweep", long(sw + 1));
    args.add("Cutoff", S.cutoff);
    args.add("MinDim", S.mindim);
    args.add("MaxDim", S.maxdim);
    args.add("Noise", S.noise);
    args.add("MaxIter", S.niter);
    if (S.noise != 0) throw ITError("dmrg: noise term (deltaRho) is not part of this build");
    for (int b = 1, ha = 1; ha <= 2; sweepnext(b, ha, N)) {
      auto t0 = std::chrono::steady_clock::now();
      PH.position(b, psi);
      ITensor phi = psi(b) * psi(b + 1);
      PH.fuseOps(phi.inds());
      DavidsonStats st;
      energy = davidson(PH, phi, args, &st);
      auto spec = svdBond(psi, b, phi, ha == 1 ? Fromleft : Fromright, args);
      if (obs) {
        BondRecord r{int(sw + 1), b, ha, energy, spec.truncerr, commonIndex(psi(b), psi(b + 1)).dim(), st.products,
                     std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count()};
        obs(r);
      }
    }
    check(itx_synchronize(ctx()), "itx_synchronize");
    if (sweep_seconds)
      sweep_seconds->push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t_sweep).count());
  }
  return energy;
}

}  // namespace itensor
