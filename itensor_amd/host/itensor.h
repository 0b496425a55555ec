// ITensor-compatible host layer of the MI355X DMRG engine.
//
// The public names and semantics follow ITensor v3's C++ API for the DMRG
// hot path (Index / QN / ITensor operator* and +=, norm, combiner,
// diag_hermitian, denmatDecomp, davidson, LocalOp / LocalMPO, Sweeps,
// dmrg), so that driver code written against the reference compiles
// against this header with the same meaning.  Tensor data lives in device
// memory; every arithmetic step is a call into libitx through the C ABI of
// include/itx.h.  Host code only does the integer bookkeeping the reference
// does on the CPU (labels, QN block lists, truncation decisions).
#pragma once

#include <array>
#include <cstdint>
#include <functional>
#include <iosfwd>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <variant>
#include <vector>

#include "itx.h"

namespace itensor {

using Real = double;

enum Arrow { In = -1, Out = 1 };
enum Direction { Fromleft, Fromright, NoDir };

// Errors of the library are thrown (the reference aborts in Error(),
// util/error.h:26,54-75; a library must not abort its host process).
class ITError : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

// ---------------------------------------------------------------- Args --
// Named-argument bag (util/args.h:50-160), with the keys of the DMRG path.
class Args {
 public:
  using Val = std::variant<bool, long, double, std::string>;
  Args() = default;
  Args(std::initializer_list<std::pair<const std::string, Val>> il) : m_(il) {}
  bool defined(std::string const &k) const { return m_.count(k) > 0; }
  Args &add(std::string const &k, Val v) {
    m_[k] = std::move(v);
    return *this;
  }
  long getInt(std::string const &k, long def) const;
  long getInt(std::string const &k) const;
  double getReal(std::string const &k, double def) const;
  double getReal(std::string const &k) const;
  bool getBool(std::string const &k, bool def) const;
  std::string getString(std::string const &k, std::string const &def) const;

 private:
  std::map<std::string, Val> m_;
};

// ------------------------------------------------------------------ QN --
// Up to four named U(1) charges (qn.h:37-180); missing names count as 0.
class QN {
 public:
  QN() = default;
  explicit QN(int v) { set("", v); }
  QN(std::string const &name, int v) { set(name, v); }
  QN(std::string const &n1, int v1, std::string const &n2, int v2) {
    set(n1, v1);
    add(n2, v2);
  }
  int val(std::string const &name) const;
  QN operator+(QN const &o) const;
  QN operator-(QN const &o) const { return *this + (-o); }
  QN operator-() const;
  bool operator==(QN const &o) const { return q_ == o.q_; }
  bool operator!=(QN const &o) const { return !(*this == o); }
  std::vector<std::pair<std::string, int>> const &store() const { return q_; }
  std::string str() const;

 private:
  void set(std::string const &n, int v);
  void add(std::string const &n, int v);
  std::vector<std::pair<std::string, int>> q_;  // sorted by name, no zeros
};
inline QN operator*(QN const &q, Arrow a) { return a == Out ? q : -q; }

// --------------------------------------------------------------- Index --
using QNBlocks = std::vector<std::pair<QN, long>>;

// Index{id, dim, dir, tags, prime level, QN sectors} (index.h:78-282).
// Equality compares id, tags and prime level (index.cc:154-157).
class Index {
 public:
  Index() = default;
  explicit Index(long dim, std::string const &tags = "");
  Index(QNBlocks qns, Arrow dir, std::string const &tags = "");
  // full description (used by bindings and tests)
  Index(uint64_t id, long dim, Arrow dir, int plev, std::string const &tags, QNBlocks qns);

  uint64_t id() const { return id_; }
  long dim() const { return dim_; }
  Arrow dir() const { return dir_; }
  int primeLevel() const { return plev_; }
  std::string const &tags() const { return tags_; }
  bool hasTag(std::string const &t) const;
  bool hasQNs() const { return !qns_.empty(); }
  long nblock() const { return hasQNs() ? long(qns_.size()) : 1; }
  long blocksize0(long b) const { return hasQNs() ? qns_[b].second : dim_; }
  QN const &qn0(long b) const { return qns_.at(b).first; }  // 0-based sector
  long blockStart(long b) const;
  QNBlocks const &qnBlocks() const { return qns_; }
  explicit operator bool() const { return id_ != 0; }

  Index &prime(int inc = 1) {
    plev_ += inc;
    return *this;
  }
  Index &setPrime(int p) {
    plev_ = p;
    return *this;
  }
  Index &noPrime() { return setPrime(0); }
  Index &dag() {
    dir_ = Arrow(-int(dir_));
    return *this;
  }
  Index &setTags(std::string const &t);
  Index &removeQNs() {
    qns_.clear();
    return *this;
  }
  std::string str() const;

 private:
  uint64_t id_ = 0;
  long dim_ = 0;
  Arrow dir_ = Out;
  int plev_ = 0;
  std::string tags_;  // canonical: sorted, comma separated
  QNBlocks qns_;
};
bool operator==(Index const &a, Index const &b);
inline bool operator!=(Index const &a, Index const &b) { return !(a == b); }
inline Index prime(Index i, int inc = 1) { return i.prime(inc); }
inline Index dag(Index i) { return i.dag(); }
inline Index noPrime(Index i) { return i.noPrime(); }
inline long dim(Index const &i) { return i.dim(); }
uint64_t newIndexId();

using IndexSet = std::vector<Index>;
long dim(IndexSet const &is);
int indexPosition(IndexSet const &is, Index const &I);
bool hasQNs(IndexSet const &is);

using Block = std::vector<int>;
// reverse-lexicographic block order (itdata/qdense.cc:60-65)
bool blockLess(Block const &a, Block const &b);
struct BlOf {
  Block block;
  int64_t offset;
};
using BlockOffsets = std::vector<BlOf>;
QN calcDiv(IndexSet const &is, Block const &b);
// getBlockOffsets (itdata/qdense.cc:133-173)
std::pair<BlockOffsets, int64_t> getBlockOffsets(IndexSet const &is, QN const &div);

// ------------------------------------------------------------- device --
// The process-wide libitx context used by the host layer (device 0, the
// null stream unless setStream is called).
itx_ctx *ctx();
// the context if one is open (null before the first use and after
// shutdown); destructors use it so that they make no HIP call once the
// context is gone
itx_ctx *liveCtx();
// Releases the context and everything it holds (plans, chains, the
// device pool, pinned buffers) while the HIP runtime is still up.  Device
// buffers, plans and chains destroyed later are dropped without a HIP
// call, and ctx() throws from then on.  Registered to run at interpreter
// exit by the Python module; a C++ program calls it before returning from
// main.
void shutdown();
void setStream(void *stream);
void setDevice(int device);
void check(int rc, const char *what);
// contraction precision of the H*psi products (ITX_PREC_*); everything
// else on the path always runs in FP64
int contractPrecision();
void setContractPrecision(int prec);
struct PrecisionScope {
  int saved;
  explicit PrecisionScope(int p) : saved(contractPrecision()) { setContractPrecision(p); }
  ~PrecisionScope() { setContractPrecision(saved); }
};

struct DevBuf {
  double *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t n_);
  ~DevBuf();
  DevBuf(DevBuf const &) = delete;
  DevBuf &operator=(DevBuf const &) = delete;
};
using DevPtr = std::shared_ptr<DevBuf>;
DevPtr devAlloc(size_t n);
DevPtr devZeros(size_t n);

// ------------------------------------------------------------ storage --
enum class SType { None, Dense, QDense, Combiner, QCombiner };

struct QCombInfo {
  IndexSet inds;  // the combined indices (original arrows)
  // sector numbers of the combined indices -> (combined sector, start, end)
  std::map<Block, std::tuple<int, long, long>> ranges;
  int nsectors = 0;
};

struct Storage {
  SType type = SType::None;
  DevPtr data;  // Dense / QDense values
  size_t size = 0;
  BlockOffsets offsets;  // QDense, sorted in block order
  std::shared_ptr<QCombInfo> qc;
};

// -------------------------------------------------------------- ITensor --
class ITensor {
 public:
  ITensor() = default;
  ITensor(IndexSet is, std::shared_ptr<Storage> st) : is_(std::move(is)), store_(std::move(st)) {}

  explicit operator bool() const { return store_ && store_->type != SType::None; }
  IndexSet const &inds() const { return is_; }
  int order() const { return int(is_.size()); }
  Storage const &store() const { return *store_; }
  std::shared_ptr<Storage> const &storePtr() const { return store_; }
  bool isQN() const { return store_ && store_->type == SType::QDense; }
  bool isCombiner() const {
    return store_ && (store_->type == SType::Combiner || store_->type == SType::QCombiner);
  }

  // contraction (itensor.cc:935-981)
  ITensor &operator*=(ITensor const &R);
  ITensor &operator*=(Real r);
  ITensor &operator/=(Real r) { return operator*=(1.0 / r); }
  // daxpy / += / -= (itensor.cc:1207-1264)
  ITensor &operator+=(ITensor const &R);
  ITensor &operator-=(ITensor const &R);

  // index manipulation
  ITensor &noPrime();
  ITensor &prime(int inc = 1);
  ITensor &prime(Index const &I, int inc = 1);
  ITensor &dag();
  ITensor &setTags(std::string const &tags, Index const &I);
  ITensor &replaceInd(Index const &from, Index const &to);

  // order-0 value (eltC of a scalar ITensor)
  Real elt() const;
  // make storage exclusive (copy-on-write, itdata/itdata.h:126-328)
  void detach();

  size_t nnz() const;
  long nnzblocks() const;

 private:
  IndexSet is_;
  std::shared_ptr<Storage> store_;
};

ITensor operator*(ITensor L, ITensor const &R);
ITensor operator*(ITensor T, Real r);
ITensor operator*(Real r, ITensor T);
ITensor operator+(ITensor L, ITensor const &R);
ITensor operator-(ITensor L, ITensor const &R);
void daxpy(ITensor &L, ITensor const &R, Real alpha);

Real norm(ITensor const &T);
// sum_k c[k] * V[k] (one kernel when the storages are aligned: same index
// order and block list; else through +=)
ITensor lincomb(std::vector<ITensor const *> const &V, std::vector<Real> const &c);
// eltC(dag(V[k]) * q) for every k (one reduction launch and one wait when
// the storages are aligned)
std::vector<Real> dotDagMany(std::vector<ITensor const *> const &V, ITensor const &q);
// eltC(dag(A)*B) for real tensors
Real dotDag(ITensor const &A, ITensor const &B);
ITensor dag(ITensor T);
ITensor prime(ITensor T, int inc = 1);
ITensor prime(ITensor T, Index const &I, int inc = 1);
ITensor noPrime(ITensor T);
ITensor removeQNs(ITensor const &T);
Index commonIndex(ITensor const &A, ITensor const &B);
bool hasIndex(ITensor const &T, Index const &I);
QN div(ITensor const &T);

// combiner (itensor.cc:1468-1560)
std::tuple<ITensor, Index> combiner(IndexSet const &inds, Args const &args = Args());

// construction / transfer
ITensor denseFromHost(IndexSet const &is, std::vector<double> const &colmajor);
// QDense with the given blocks (sorted) and flat block store
ITensor qdenseFromHost(IndexSet const &is, BlockOffsets const &offs, std::vector<double> const &store);
ITensor qdenseZeros(IndexSet const &is, QN const &div);
std::vector<double> storageToHost(ITensor const &T);
void randomize(ITensor &T, uint64_t seed);
// randomITensor(QN, IndexSet) (itensor.h:969-975): N(0,1) entries in every
// flux-allowed block (dense when the indices carry no QNs)
ITensor randomITensor(QN const &div, IndexSet const &is, uint64_t seed);
// QDense of flux div from a dense column-major array (zero elsewhere)
ITensor qdenseFromDense(IndexSet const &is, std::vector<double> const &colmajor, QN const &div);

// ------------------------------------------------- planned contractions --
// Block-pair task list of a QN contraction, from the operands' structure
// (getContractedOffsets, itdata/qutil.h:93-242).  Task t contracts the A
// block at offs[3t] with the B block at offs[3t+1] into the C block at
// offs[3t+2]; cfirst[t] marks the first task of its C block.
// Host-side phase timers (wall clock, accumulated per process); active when
// the environment variable ITX_HOST_TIMERS is set, reported to stderr at
// exit.  Diagnostics only: they add no synchronisation.
enum HostTimerId {
  HT_QN_TASKLIST,
  HT_PLAN_QN,
  HT_PLAN_DENSE,
  HT_CONTRACT_QN,
  HT_CONTRACT_DENSE,
  HT_PROGRAM_BUILD,
  HT_PRODUCT,
  HT_NORM_DOT,
  HT_DIAG_HERMITIAN,
  HT_DENMAT,
  HT_ENV,
  HT_CHAIN_CREATE,
  HT_SYEV,
  HT_COUNT
};
class HostTimer {
 public:
  explicit HostTimer(HostTimerId id);
  ~HostTimer();

 private:
  HostTimerId id_;
  bool on_;
  double t0_;
};

struct QNTaskList {
  IndexSet N;
  std::vector<int32_t> la, lb, lc;
  BlockOffsets coffs;
  int64_t csize = 0;
  int nt = 0;
  std::vector<int64_t> offs, bda, bdb;
  std::vector<int> cfirst;
};
QNTaskList qnTaskList(IndexSet const &Ais, BlockOffsets const &Aoffs, IndexSet const &Bis,
                      BlockOffsets const &Boffs);

// A contraction A*B planned on the structure of its operands only (index
// sets, storage kind, block lists).  The result structure and the device
// plan are built once; exec() runs on any operands of the same structures
// and writes every element of the result (first-touch).  With perBlock,
// a QN contraction runs as one launch per block pair with the betas of
// doTask(Contract,QDense,QDense) (itdata/qdense.cc:700-731) -- the
// reference's per-block loop -- instead of one grouped launch.
class ContractPlan {
 public:
  // cover (QN only, optional): a block layout of the result holding every
  // block the contraction produces; the result is then laid out as cover,
  // with the blocks no block pair reaches left at zero.  cperm (optional,
  // with cover): result index q of that layout is index cperm[q] of the
  // natural result order (the result is written in the permuted order).
  ContractPlan(ITensor const &A, ITensor const &B, int prec, bool perBlock = false,
               Storage const *cover = nullptr, std::vector<int> const *cperm = nullptr);
  ~ContractPlan();
  ContractPlan(ContractPlan const &) = delete;
  ContractPlan &operator=(ContractPlan const &) = delete;
  IndexSet const &inds() const { return N_; }
  bool isQN() const { return qn_; }
  BlockOffsets const &offsets() const { return coffs_; }
  size_t size() const { return size_; }
  double flops() const { return flops_; }
  // flops per arithmetic form its kernels compute in (ITX_PREC_* index)
  std::array<double, 3> const &flopsByForm() const { return formFlops_; }
  itx_plan *plan() const { return plan_; }  // null in per-block mode
  bool perBlock() const { return !blockPlans_.empty() || (perBlock_ && !plan_); }
  std::shared_ptr<Storage> newOutput() const;
  void exec(double const *A, double const *B, double *C) const;
  // structure test of an operand against the planned one
  static bool sameStructure(ITensor const &a, IndexSet const &is, bool qn, BlockOffsets const &offs);

 private:
  IndexSet N_;
  bool qn_ = false, perBlock_ = false, zeroOut_ = false;
  BlockOffsets coffs_;
  size_t size_ = 0;
  double flops_ = 0;
  std::array<double, 3> formFlops_{};
  void account(itx_plan const *p);
  itx_plan *plan_ = nullptr;
  std::vector<itx_plan *> blockPlans_;
  std::vector<double> blockBeta_;
};

// ------------------------------------------------------------ file I/O --
// Binary stream format of the reference's write/read (index.cc:711-745,
// itensor.cc:768-806, itdata/qdense.h:236-248): real Dense and QDense
// storage, U(1) quantum numbers, tags of at most 7 characters.
void write(std::ostream &s, Index const &I);
Index readIndex(std::istream &s);
void write(std::ostream &s, ITensor const &T);
ITensor readITensor(std::istream &s);
void read(std::istream &s, ITensor &T);
// writeToFile / readFromFile (util/readwrite.h:336-370)
void writeToFile(std::string const &fname, ITensor const &T);
ITensor readITensorFromFile(std::string const &fname);

// ------------------------------------------------------------- decomps --
struct Spectrum {
  std::vector<Real> eigs;
  Real truncerr = 0;
  Real truncerr_() const { return truncerr; }
};

// truncate (decomp.cc:305-463): P descending, resized in place
std::tuple<Real, Real, Real, int> truncate(std::vector<Real> &P, long maxdim, long mindim, Real cutoff,
                                           bool absoluteCutoff, bool doRelCutoff, Args const &args);
// diag_hermitian (hermitian.cc:40-440)
Spectrum diag_hermitian(ITensor rho, ITensor &U, ITensor &D, Args const &args);
// svd (decomp.cc:121-248 -> svd.cc:39-436): U (or V) names the indices of
// AA that go to U; Args Cutoff, MaxDim, MinDim, Truncate, LeftTags ("Link,U"),
// RightTags ("Link,V"), RespectDegenerate, DoRelCutoff, AbsoluteCutoff
Spectrum svd(ITensor const &AA, ITensor &U, ITensor &D, ITensor &V, Args args = Args());
std::tuple<ITensor, ITensor, ITensor> svd(ITensor const &AA, IndexSet const &Uis, Args const &args = Args());
// factor (decomp.cc:538-576): T = A * B with sqrt(D) in both
Spectrum factor(ITensor const &T, ITensor &A, ITensor &B, Args args = Args());
std::tuple<ITensor, ITensor> factor(ITensor const &T, IndexSet const &Ais, Args const &args = Args());
// denmatDecomp (decomp.h:316-418).  The noise term needs the local
// Hamiltonian: deltaRho(AA, combiner, dir) returns the correction to the
// density matrix and its trace (LocalOp::deltaRho, localop.h:375-406);
// Args "Noise" > 0 without one is an error, as in the reference.
using DeltaRhoFn = std::function<std::pair<ITensor, double>(ITensor const &, ITensor const &, Direction)>;
Spectrum denmatDecomp(ITensor const &AA, ITensor &A, ITensor &B, Direction dir, Args args,
                      DeltaRhoFn const &deltaRho = DeltaRhoFn());

}  // namespace itensor
