// LocalOp::size, LocalMPO::position, svdBond, the DMRG driver and the small
// host eigensolver used by davidson's Rayleigh-Ritz step.
#include <algorithm>
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
  progs_.clear();
  W12_ = nullptr;
  if (!W12 || LIsNull() || RIsNull()) return false;
  IndexSet t1 = productInds(phiInds, L_->inds());
  IndexSet ref = productInds(productInds(productInds(t1, Op1_->inds()), Op2_->inds()), R_->inds());
  IndexSet fus = productInds(productInds(t1, W12->inds()), R_->inds());
  if (!sameOrder(ref, fus)) return false;
  W12_ = W12;
  return true;
}

// ------------------------------------------------------ planned products --
// One product chain planned for one phi structure: the stage plans, the
// intermediate buffers (reused by every product) and, in Graph mode, the
// chain replayed as a HIP graph.
struct LocalOp::Program {
  IndexSet phiInds;
  bool qn = false;
  int prec = ITX_PREC_F64;
  BlockOffsets phiOffs;
  std::vector<ITensor const *> ops;
  std::vector<std::unique_ptr<ContractPlan>> stages;
  std::vector<DevPtr> tmp;  // outputs of every stage but the last
  IndexSet outInds;
  itx_chain *chain = nullptr;
  double flops = 0;
  std::array<double, 3> formFlops{};
  ~Program() {
    // after shutdown() the context has released the chain already
    if (chain && liveCtx()) itx_chain_destroy(liveCtx(), chain);
  }
};

std::vector<ITensor const *> LocalOp::chainOps() const {
  std::vector<ITensor const *> ops;
  if (W12_ && isBulk()) return {L_, W12_, R_};
  if (LIsNull()) {
    if (!RIsNull()) ops.push_back(R_);
    ops.push_back(Op2_);
    ops.push_back(Op1_);
  } else {
    ops = {L_, Op1_, Op2_};
    if (!RIsNull()) ops.push_back(R_);
  }
  return ops;
}

LocalOp::Program *LocalOp::program(ITensor const &phi) const {
  if (mode_ == HpsiMode::Eager || !phi || phi.isCombiner() || phi.order() == 0) return nullptr;
  for (auto it = progs_.rbegin(); it != progs_.rend(); ++it)
    if ((*it)->prec == prec_ && ContractPlan::sameStructure(phi, (*it)->phiInds, (*it)->qn, (*it)->phiOffs))
      return it->get();
  auto ops = chainOps();
  for (auto *o : ops)
    if (!o || !*o || o->isCombiner() || o->order() == 0 || o->isQN() != phi.isQN()) return nullptr;
  HostTimer ht(HT_PROGRAM_BUILD);
  auto prg = std::make_shared<Program>();
  prg->phiInds = phi.inds();
  prg->qn = phi.isQN();
  prg->prec = prec_;
  if (prg->qn) prg->phiOffs = phi.store().offsets;
  prg->ops = ops;
  // structure-only stand-in for each stage's left operand
  ITensor cur = phi;
  for (size_t i = 0; i < ops.size(); ++i) {
    Storage const *cover = nullptr;
    std::vector<int> cperm;
    if (cover_ && prg->qn && i + 1 == ops.size() && mode_ != HpsiMode::PerBlock) {
      // result of the chain in phi's index order and block layout
      IndexSet o = productInds(cur.inds(), ops[i]->inds());
      for (auto &I : o) I.noPrime();
      if (o.size() == phi.inds().size()) {
        for (auto const &I : phi.inds()) cperm.push_back(indexPosition(o, I));
        bool ok = true;
        for (size_t q = 0; q < cperm.size(); ++q) ok = ok && cperm[q] >= 0 && o[size_t(cperm[q])].dir() == phi.inds()[q].dir();
        if (ok) cover = &phi.store();
      }
    }
    prg->stages.push_back(std::make_unique<ContractPlan>(cur, *ops[i], prec_, mode_ == HpsiMode::PerBlock, cover,
                                                         cover ? &cperm : nullptr));
    ContractPlan const &cp = *prg->stages.back();
    prg->flops += cp.flops();
    for (int f = 0; f < 3; ++f) prg->formFlops[size_t(f)] += cp.flopsByForm()[size_t(f)];
    auto st = std::make_shared<Storage>();
    st->type = cp.isQN() ? SType::QDense : SType::Dense;
    st->offsets = cp.offsets();
    st->size = cp.size();
    if (i + 1 < ops.size()) {
      st->data = devAlloc(cp.size());
      prg->tmp.push_back(st->data);
    }
    cur = ITensor(cp.inds(), st);
  }
  cur.noPrime();
  prg->outInds = cur.inds();
  if (mode_ == HpsiMode::Graph) {
    std::vector<itx_plan *> plans;
    for (auto &cp : prg->stages)
      if (cp->plan()) plans.push_back(cp->plan());
    if (plans.size() == prg->stages.size()) {
      HostTimer hc(HT_CHAIN_CREATE);
      check(itx_chain_create(ctx(), int(plans.size()), plans.data(), &prg->chain), "itx_chain_create");
    }
  }
  if (progs_.size() >= 4) progs_.erase(progs_.begin());
  progs_.push_back(prg);
  return prg.get();
}

void LocalOp::product(ITensor const &phi, ITensor &phip) const {
  HostTimer ht(HT_PRODUCT);
  PrecisionScope ps(prec_);
  if (Program *prg = program(phi)) {
    const size_t n = prg->stages.size();
    auto out = prg->stages.back()->newOutput();
    std::vector<const double *> A(n), B(n);
    std::vector<double *> C(n);
    for (size_t i = 0; i < n; ++i) {
      A[i] = i == 0 ? phi.store().data->p : prg->tmp[i - 1]->p;
      B[i] = prg->ops[i]->store().data->p;
      C[i] = i + 1 == n ? out->data->p : prg->tmp[i]->p;
    }
    if (prg->chain) {
      check(itx_chain_launch(ctx(), prg->chain, A.data(), B.data(), C.data(), nullptr, nullptr),
            "itx_chain_launch");
    } else {
      for (size_t i = 0; i < n; ++i) prg->stages[i]->exec(A[i], B[i], C[i]);
    }
    phip = ITensor(prg->outInds, out);
    return;
  }
  auto ops = chainOps();
  phip = phi;
  for (auto *o : ops) phip *= *o;
  phip.noPrime();
}

double LocalOp::flops(ITensor const &phi) const {
  PrecisionScope ps(prec_);
  Program *prg = program(phi);
  if (!prg) throw ITError("LocalOp::flops: chain cannot be planned for this phi (Eager mode or mixed storage)");
  return prg->flops;
}

std::array<double, 3> LocalOp::flopsByForm(ITensor const &phi) const {
  PrecisionScope ps(prec_);
  Program *prg = program(phi);
  if (!prg) return {};
  return prg->formFlops;
}

std::pair<ITensor, double> LocalOp::deltaRho(ITensor const &AA, ITensor const &combine, Direction dir) const {
  if (!Op1_ || !Op2_) throw ITError("LocalOp::deltaRho: operators not set");
  ITensor X = AA;
  if (dir == Fromleft) {
    if (!LIsNull()) X *= *L_;
    X *= *Op1_;
  } else {
    if (!RIsNull()) X *= *R_;
    X *= *Op2_;
  }
  X = combine * noPrime(X);
  const Index ci = commonIndex(combine, X);
  const double nx = norm(X);
  ITensor drho = X * dag(prime(X, ci));
  // (drho + dag(drho with the primes of ci swapped)) / 2
  Index c1 = ci, c2 = ci;
  c1.prime(1);
  c2.prime(2);
  ITensor t = prime(prime(prime(drho, c1, 1), ci, 1), c2, -2);
  drho += dag(t);
  drho *= 0.5;
  return {drho, nx * nx};
}

void LocalMPO::fuseOps(IndexSet const &phiInds) {
  if (!fuse_ || !lop_.isBulk()) return;
  auto const &H = *H_;
  W12_ = H[size_t(b_ - 1)] * H[size_t(b_)];
  if (!lop_.setFused(&W12_, phiInds)) W12_ = ITensor();
}

// An environment held in pinned host memory: its structure and the block
// its values were copied into.
struct LocalMPO::HostEnv {
  IndexSet inds;
  SType type = SType::Dense;
  BlockOffsets offsets;
  size_t size = 0;
  void *host = nullptr;
  ~HostEnv() {
    if (host && liveCtx()) itx_host_free(liveCtx(), host);
  }
};

void LocalMPO::stash(int j) {
  ITensor &E = PH_[size_t(j)];
  if (!E) return;
  auto h = std::make_shared<HostEnv>();
  h->inds = E.inds();
  h->type = E.store().type;
  h->offsets = E.store().offsets;
  h->size = E.store().size;
  const size_t bytes = sizeof(double) * std::max<size_t>(h->size, 1);
  check(itx_host_alloc(ctx(), bytes, &h->host), "itx_host_alloc");
  if (h->size)
    check(itx_memcpy_async(ctx(), h->host, E.store().data->p, sizeof(double) * h->size), "itx_memcpy_async");
  stash_[size_t(j)] = h;
  E = ITensor();  // the device block returns to the stream-ordered pool
}

void LocalMPO::restore(int j) {
  auto &h = stash_[size_t(j)];
  if (PH_[size_t(j)] || !h) return;
  auto st = std::make_shared<Storage>();
  st->type = h->type;
  st->offsets = h->offsets;
  st->size = h->size;
  st->data = devAlloc(h->size);
  if (h->size) check(itx_memcpy_async(ctx(), st->data->p, h->host, sizeof(double) * h->size), "itx_memcpy_async");
  PH_[size_t(j)] = ITensor(h->inds, st);
  h.reset();  // the pinned block returns to its pool in stream order
}

// setLHlim / setRHlim (localmpo.h:620-668): with doWrite the environment
// at the old limit leaves device memory and the one at the new limit is
// brought back
void LocalMPO::setLHlim(int val) {
  if (do_write_ && LHlim_ != val && PH_[size_t(LHlim_)]) stash(LHlim_);
  LHlim_ = val;
  if (do_write_ && LHlim_ >= 1) restore(LHlim_);
}

void LocalMPO::setRHlim(int val) {
  if (do_write_ && RHlim_ != val && PH_[size_t(RHlim_)]) stash(RHlim_);
  RHlim_ = val;
  if (do_write_ && RHlim_ <= int(H_->size())) restore(RHlim_);
}

int LocalMPO::residentEnvs() const {
  int n = 0;
  for (auto const &E : PH_) n += bool(E) ? 1 : 0;
  return n;
}

int LocalMPO::hostEnvs() const {
  int n = 0;
  for (auto const &h : stash_) n += h ? 1 : 0;
  return n;
}

void LocalMPO::position(int b, MPS const &psi) {
  HostTimer ht(HT_ENV);
  auto t0 = std::chrono::steady_clock::now();
  auto const &H = *H_;
  while (LHlim_ < b - 1) {
    const int ll = LHlim_;
    ITensor E = PH_[size_t(ll)] ? PH_[size_t(ll)] * psi(ll + 1) : psi(ll + 1);
    E *= H[size_t(ll)];
    E *= dag(prime(psi(ll + 1)));
    PH_[size_t(ll + 1)] = E;
    stash_[size_t(ll + 1)].reset();  // superseded
    setLHlim(ll + 1);
  }
  while (RHlim_ > b + 2) {
    const int rl = RHlim_;
    ITensor E = PH_[size_t(rl)] ? PH_[size_t(rl)] * psi(rl - 1) : psi(rl - 1);
    E *= H[size_t(rl - 2)];
    E *= dag(prime(psi(rl - 1)));
    PH_[size_t(rl - 1)] = E;
    stash_[size_t(rl - 1)].reset();
    setRHlim(rl - 1);
  }
  setLHlim(b - 1);
  setRHlim(b + 2);
  lop_.update(&H[size_t(b - 1)], &H[size_t(b)], &PH_[size_t(b - 1)], &PH_[size_t(b + 2)]);
  b_ = b;
  env_seconds_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

Spectrum svdBond(MPS &psi, int b, ITensor const &phi, Direction dir, Args const &args_in, LocalMPO const *PH) {
  Args args = args_in;
  const double noise = args.getReal("Noise", 0.0);
  const double cutoff = args.getReal("Cutoff", 1e-15);  // MIN_CUT, global.h:61
  const bool usesvd = args.getBool("UseSVD", false);
  args.add("RespectDegenerate", args.getBool("RespectDegenerate", true));
  Index oldlink = commonIndex(psi(b), psi(b + 1));
  std::string tags = oldlink ? oldlink.tags() : std::string("Link");
  ITensor A = psi(b), B = psi(b + 1);
  Spectrum spec;
  if (usesvd || (noise == 0 && cutoff < 1e-12)) {
    // the accurate route (mps_impl.h:50-65): A keeps psi(b)'s indices of
    // phi as U, B becomes V, and D goes into the new centre
    ITensor D;
    spec = svd(phi, A, D, B, args);
    if (args.getBool("DoNormalize", false)) D *= 1.0 / norm(D);
    if (dir == Fromleft)
      B *= D;
    else
      A *= D;
  } else {
    DeltaRhoFn drho;
    if (PH && noise > 0)
      drho = [PH](ITensor const &AA, ITensor const &cmb, Direction d) { return PH->deltaRho(AA, cmb, d); };
    spec = denmatDecomp(phi, A, B, dir, args, drho);
    if (args.getBool("DoNormalize", false)) {
      ITensor &oc = (dir == Fromleft ? B : A);
      const double nrm = norm(oc);
      if (nrm > 1e-16) oc *= 1.0 / nrm;
    }
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

Spectrum orthMPS(ITensor &A1, ITensor &A2, Direction dir, Args const &args) {
  ITensor &L = (dir == Fromleft ? A1 : A2);
  ITensor &R = (dir == Fromleft ? A2 : A1);
  Index bnd = commonIndex(L, R);
  if (!bnd) return Spectrum();
  ITensor A, B(IndexSet{bnd}, nullptr), D;
  auto spec = svd(L, A, D, B, args);
  L = A;
  R *= (D * B);
  return spec;
}

void position(MPS &psi, int i, Args const &args) {
  const int N = psi.length();
  if (N == 0) throw ITError("position: MPS is empty");
  auto linkTags = [&](int b) {
    Index l = commonIndex(psi(b), psi(b + 1));
    return l ? l.tags() : std::string("Link");
  };
  while (psi.leftLim() < i - 1) {
    if (psi.leftLim() < 0) psi.leftLim(0);
    const int b = psi.leftLim() + 1;
    Args a = args;
    a.add("LeftTags", linkTags(b));
    orthMPS(psi.ref(b), psi.ref(b + 1), Fromleft, a);
    psi.leftLim(psi.leftLim() + 1);
    if (psi.rightLim() < psi.leftLim() + 2) psi.rightLim(psi.leftLim() + 2);
  }
  while (psi.rightLim() > i + 1) {
    if (psi.rightLim() > N + 1) psi.rightLim(N + 1);
    const int b = psi.rightLim() - 2;
    Args a = args;
    const std::string t = linkTags(b);
    // LeftTags equal to svd's default RightTags would give U and V the
    // same tags (mps.cc:720-725)
    a.add("LeftTags", t == "Link,V" ? std::string("Link") : t);
    orthMPS(psi.ref(b), psi.ref(b + 1), Fromright, a);
    psi.rightLim(psi.rightLim() - 1);
    if (psi.leftLim() > psi.rightLim() - 2) psi.leftLim(psi.rightLim() - 2);
  }
}

Real dmrg(MPS &psi, MPO const &H, Sweeps const &sweeps, Args const &args_in,
          std::function<void(BondRecord const &)> obs, std::vector<double> *sweep_seconds) {
  Args args = args_in;
  args.add("RespectDegenerate", args.getBool("RespectDegenerate", true));
  args.add("DoNormalize", true);
  const int N = psi.length();
  if (int(H.size()) != N) throw ITError("dmrg: MPO and MPS lengths differ");
  position(psi, 1);  // dmrg.h:373
  LocalMPO PH(H);
  PH.setPrecision(int(args.getInt("HpsiPrecision", ITX_PREC_F64)));
  PH.setFusion(args.getBool("FuseMPO", true));
  PH.setMode(HpsiMode(args.getInt("HpsiMode", long(HpsiMode::Graph))));
  // QN: davidson starts from phi widened to every flux-allowed block
  // (zeros where psi(b)*psi(b+1) has none) and the products keep that
  // layout, so the Krylov vectors of a bond share one block set.  The
  // values are those of the reference's growing block sets; the added
  // blocks are exact zeros, whose density-matrix weight is zero.
  const bool fullBlocks = args.getBool("FullBlockKrylov", true);
  PH.setCoverOutput(fullBlocks);
  const bool syncTiming = args.getBool("SyncTiming", false);
  const long finalFP64 = args.getInt("FinalFP64Sweeps", 0);
  const long writeDim = args.getInt("WriteDim", -1);
  const bool energyFP64 = args.getBool("EnergyFP64", true);
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
    // precision schedule: the last FinalFP64Sweeps sweeps run the H*psi
    // contractions in FP64 whatever HpsiPrecision says
    if (long(sweeps.size() - sw) <= finalFP64) PH.setPrecision(ITX_PREC_F64);
    // dmrg.h:390-402: past WriteDim only the environments in use stay on
    // the device
    if (!PH.doWrite() && writeDim >= 0 && S.maxdim >= writeDim) PH.doWrite(true);
    for (int b = 1, ha = 1; ha <= 2; sweepnext(b, ha, N)) {
      using clk = std::chrono::steady_clock;
      auto secs = [](clk::time_point a, clk::time_point b2) { return std::chrono::duration<double>(b2 - a).count(); };
      auto mark = [&]() {
        if (syncTiming) check(itx_synchronize(ctx()), "itx_synchronize");
        return clk::now();
      };
      auto t0 = clk::now();
      PH.position(b, psi);
      auto t1 = mark();
      ITensor phi = psi(b) * psi(b + 1);
      if (fullBlocks && phi.isQN()) {
        ITensor full = qdenseZeros(phi.inds(), div(phi));
        full += phi;
        phi = full;
      }
      PH.fuseOps(phi.inds());
      DavidsonStats st;
      energy = davidson(PH, phi, args, &st);
      const double ritz = energy;
      std::array<double, 3> hf{};
      if (obs) {
        hf = PH.flopsByForm(phi);
        for (auto &x : hf) x *= st.products;
      }
      // Reduced-precision products leave an error of the order of the
      // form's rounding times |H| in the Ritz value, while the Ritz vector
      // is accurate to second order.  The energy a sweep reports is then
      // the FP64 Rayleigh quotient of its last Ritz vector: one FP64
      // product per sweep.
      double e64 = NAN;
      const bool lastStep = ha == 2 && b == 1;
      if (energyFP64 && lastStep && PH.precision() != ITX_PREC_F64) {
        const int p = PH.precision();
        PH.setPrecision(ITX_PREC_F64);
        ITensor Hphi;
        PH.product(phi, Hphi);
        PH.setPrecision(p);
        e64 = dotDag(phi, Hphi) / dotDag(phi, phi);
        energy = e64;
        if (obs) hf[0] += PH.flops(phi);  // the FP64 product's flops
      }
      auto t2 = mark();
      auto spec = svdBond(psi, b, phi, ha == 1 ? Fromleft : Fromright, args, &PH);
      auto t3 = mark();
      if (obs) {
        BondRecord r{int(sw + 1), b, ha, ritz, spec.truncerr, commonIndex(psi(b), psi(b + 1)).dim(), st.products,
                     secs(t0, t3), secs(t0, t1), secs(t1, t2), secs(t2, t3), hf, PH.residentEnvs(),
                     PH.hostEnvs(), e64};
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
