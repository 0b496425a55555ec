// MPS / MPO containers, the projected operator (LocalOp, LocalMPO), Sweeps
// and the two-site DMRG driver of the host layer.
//
// Reference behaviour: mps/localop.h:324-365 (product), mps/localmpo.h
// (position :443-470, makeL/makeR :542-618), mps/mps_impl.h:21-99
// (svdBond, density-matrix route), mps/sweeps.h (Sweeps :42-145,
// sweepnext :489-500), mps/dmrg.h:336-467 (DMRGWorker).
#pragma once

#include <chrono>
#include <functional>

#include "iterativesolvers.h"
#include "itensor.h"

namespace itensor {

class MPS {
 public:
  MPS() = default;
  explicit MPS(std::vector<ITensor> A, int leftLim = 0, int rightLim = 2)
      : A_(std::move(A)), llim_(leftLim), rlim_(rightLim) {}
  int length() const { return int(A_.size()); }
  ITensor const &operator()(int j) const { return A_.at(size_t(j - 1)); }
  ITensor &ref(int j) { return A_.at(size_t(j - 1)); }
  int leftLim() const { return llim_; }
  int rightLim() const { return rlim_; }
  void leftLim(int x) { llim_ = x; }
  void rightLim(int x) { rlim_ = x; }
  std::vector<ITensor> const &tensors() const { return A_; }

 private:
  std::vector<ITensor> A_;
  int llim_ = 0, rlim_ = 2;
};

using MPO = std::vector<ITensor>;  // W_1..W_N (0-based storage)

// MPS::write / read (mps/mps.cc:353-378): length, the N+2 tensors of the
// reference's 1-based store, the orthogonality limits
void write(std::ostream &s, MPS const &psi);
MPS readMPS(std::istream &s);
void writeToFile(std::string const &fname, MPS const &psi);
MPS readMPSFromFile(std::string const &fname);

// How LocalOp::product runs its contraction chain.
//   Eager    : ITensor operator* per stage (planned on every call, plans
//              cached by structure inside libitx)
//   Planned  : the chain is planned once per phi structure (result
//              structures, device task lists, intermediate buffers) and
//              its stage plans are executed directly
//   Graph    : as Planned, replayed as one HIP graph (itx_chain)
//   PerBlock : as Planned, but each QN block pair is its own launch (the
//              per-block loop of doTask(Contract,QDense,QDense),
//              itdata/qdense.cc:700-731) -- for comparison
enum class HpsiMode { Eager = 0, Planned = 1, Graph = 2, PerBlock = 3 };

// The two-site projected Hamiltonian acting on phi (mps/localop.h).
class LocalOp {
 public:
  LocalOp() = default;
  void update(ITensor const *Op1, ITensor const *Op2, ITensor const *L, ITensor const *R) {
    Op1_ = Op1;
    Op2_ = Op2;
    L_ = L;
    R_ = R;
    W12_ = nullptr;
    progs_.clear();
  }
  bool LIsNull() const { return L_ == nullptr || !(*L_); }
  bool isBulk() const { return !LIsNull() && !RIsNull(); }
  bool RIsNull() const { return R_ == nullptr || !(*R_); }
  // product (localop.h:324-365): phip = noPrime(phi*L*Op1*Op2*R) for the
  // two centre sites (the L-null and R-null ends drop that factor); the
  // contractions run in the precision form `prec` (ITX_PREC_*)
  void product(ITensor const &phi, ITensor &phip) const;
  long size() const;  // localop.h:488-530
  // programs are kept per precision form, so switching forms (e.g. one
  // FP64 product among reduced-precision ones) rebuilds nothing
  void setPrecision(int p) { prec_ = p; }
  int precision() const { return prec_; }
  void setMode(HpsiMode m) {
    if (m != mode_) progs_.clear();
    mode_ = m;
  }
  HpsiMode mode() const { return mode_; }
  // Full block set: when phi (QN) holds every block its product can reach
  // and the product has phi's index order, the product is laid out with
  // phi's blocks (those no block pair reaches are zero), so that the
  // Krylov vectors of davidson share one structure -- one planned chain
  // and aligned vector operations per bond instead of one per new block set
  void setCoverOutput(bool on) {
    if (on != cover_) progs_.clear();
    cover_ = on;
  }
  // Use W12 = Op1*Op2 in the bulk case when ((phi*L)*W12)*R yields exactly
  // the index order of ((phi*L)*Op1)*Op2*R; returns whether it does.
  bool setFused(ITensor const *W12, IndexSet const &phiInds);
  // algorithmic flops (sum of 2mnk over every block pair) of one product
  // on a tensor with phi's structure (plans the chain if needed)
  double flops(ITensor const &phi) const;
  // the same flops split by the arithmetic form the kernels compute in
  // (ITX_PREC_* index)
  std::array<double, 3> flopsByForm(ITensor const &phi) const;
  // noise correction to the density matrix of the bond (localop.h:375-406):
  // X = combine * noPrime(AA*L*Op1) (Fromleft) or noPrime(AA*R*Op2)
  // (Fromright), drho = (X X^dag + its transpose) / 2; returns drho and its
  // trace |X|^2
  std::pair<ITensor, double> deltaRho(ITensor const &AA, ITensor const &combine, Direction dir) const;

 private:
  struct Program;
  std::vector<ITensor const *> chainOps() const;
  Program *program(ITensor const &phi) const;
  ITensor const *Op1_ = nullptr, *Op2_ = nullptr, *L_ = nullptr, *R_ = nullptr;
  ITensor const *W12_ = nullptr;
  int prec_ = ITX_PREC_F64;
  HpsiMode mode_ = HpsiMode::Graph;
  bool cover_ = false;
  mutable std::vector<std::shared_ptr<Program>> progs_;  // most recent last
};

class LocalMPO {
 public:
  explicit LocalMPO(MPO const &H)
      : H_(&H), PH_(H.size() + 2), stash_(H.size() + 2), LHlim_(0), RHlim_(int(H.size()) + 1) {}
  // position (localmpo.h:443-470) with makeL / makeR (:542-618)
  void position(int b, MPS const &psi);
  void product(ITensor const &phi, ITensor &phip) const { lop_.product(phi, phip); }
  long size() const { return lop_.size(); }
  ITensor const &L() const { return PH_.at(size_t(LHlim_)); }
  ITensor const &R() const { return PH_.at(size_t(RHlim_)); }
  void setPrecision(int p) { lop_.setPrecision(p); }
  int precision() const { return lop_.precision(); }
  void setMode(HpsiMode m) { lop_.setMode(m); }
  void setCoverOutput(bool on) { lop_.setCoverOutput(on); }
  double flops(ITensor const &phi) const { return lop_.flops(phi); }
  std::array<double, 3> flopsByForm(ITensor const &phi) const { return lop_.flopsByForm(phi); }
  std::pair<ITensor, double> deltaRho(ITensor const &AA, ITensor const &combine, Direction dir) const {
    return lop_.deltaRho(AA, combine, dir);
  }
  double envSeconds() const { return env_seconds_; }
  // pre-contract the two MPO tensors of the current bond for phi's index set
  void fuseOps(IndexSet const &phiInds);
  void setFusion(bool on) { fuse_ = on; }
  // doWrite (localmpo.h:620-681): only the two environments in use stay in
  // device memory; the others move to pinned host memory when the limits
  // pass them and come back when a limit returns (the reference writes
  // them to files instead).  Copies are ordered on the stream and never
  // waited for.
  bool doWrite() const { return do_write_; }
  void doWrite(bool on) { do_write_ = on; }
  // environments now held in device / in host memory (diagnostics)
  int residentEnvs() const;
  int hostEnvs() const;

 private:
  struct HostEnv;
  void setLHlim(int val);
  void setRHlim(int val);
  void stash(int j);
  void restore(int j);
  MPO const *H_;
  std::vector<ITensor> PH_;
  std::vector<std::shared_ptr<HostEnv>> stash_;
  bool do_write_ = false;
  int LHlim_, RHlim_;
  LocalOp lop_;
  ITensor W12_;
  int b_ = 0;
  bool fuse_ = true;
  double env_seconds_ = 0;
};

struct SweepParams {
  long maxdim = 1, mindim = 1, niter = 2;
  double cutoff = 1e-8, noise = 0;
};
// Sweeps (mps/sweeps.h:42-145): one entry per sweep
using Sweeps = std::vector<SweepParams>;

inline void sweepnext(int &b, int &ha, int N) {
  const int inc = (ha == 1 ? +1 : -1);
  b += inc;
  if (b == (ha == 1 ? N : 0)) {
    b -= inc;
    ++ha;
  }
}

struct BondRecord {
  int sweep, b, ha;
  double energy, truncerr;
  long dim;
  int products;
  double seconds;
  // wall time of the environment update (position), davidson and svdBond;
  // device work is attributed to its phase only with Args "SyncTiming"
  double t_env = 0, t_davidson = 0, t_svd = 0;
  // algorithmic flops of the bond's H*psi products by the arithmetic form
  // their kernels computed in (ITX_PREC_* index; zero in Eager mode)
  std::array<double, 3> hpsi_flops{};
  // environments held in device and in host memory after the step
  int envs_device = 0, envs_host = 0;
  // the FP64 Rayleigh quotient <phi|H|phi> of the Ritz vector, computed
  // with one FP64 product when the step's products ran in a reduced form
  // and it ends a sweep (NaN otherwise)
  double energy_fp64 = NAN;
};

// MPS::svdBond (mps/mps_impl.h:21-99).  With Args "UseSVD", or with no
// noise and Cutoff < 1e-12, the bond is split by svd and the singular
// values go into the new orthogonality centre; otherwise by the density
// matrix (denmatDecomp), where with PH and Args "Noise" > 0 the noise term
// of PH's deltaRho is added.
Spectrum svdBond(MPS &psi, int b, ITensor const &phi, Direction dir, Args const &args,
                 LocalMPO const *PH = nullptr);

// orthMPS (mps/mps.cc:641-663): A1 (Fromleft) or A2 (Fromright) becomes
// the U factor of its svd over the shared link, the other tensor absorbs
// D*V
Spectrum orthMPS(ITensor &A1, ITensor &A2, Direction dir, Args const &args);
// MPS::position (mps/mps.cc:667-735, orthMPS route): moves the
// orthogonality limits so that every site left of i is left-orthogonal and
// every site right of i right-orthogonal
void position(MPS &psi, int i, Args const &args = Args());

// dmrg (mps/dmrg.h:51-76, 336-467).  obs (optional) receives one record
// per bond step; sweep_seconds (optional) the wall time of each sweep.
// Precision Args (no counterpart in the reference): "HpsiPrecision"
// (ITX_PREC_*) runs davidson's H*psi products in a reduced form;
// "FinalFP64Sweeps" n runs the last n sweeps in FP64; "EnergyFP64"
// (default true) makes the energy of a sweep run in a reduced form the FP64
// Rayleigh quotient of its last Ritz vector (one FP64 product per sweep).
Real dmrg(MPS &psi, MPO const &H, Sweeps const &sweeps, Args const &args,
          std::function<void(BondRecord const &)> obs = nullptr, std::vector<double> *sweep_seconds = nullptr);

}  // namespace itensor
