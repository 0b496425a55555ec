// Internal declarations shared by the libitx translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/itx.h"

namespace itx {

// One GEMM of a contraction: C[m,n] (+)= sum_k A[m,k] B[k,n], where the
// element of A at logical (m,k) lives at a_off + tab[tAM+m] + tab[tAK+k]
// (and likewise for B and C).  The offset tables carry the index
// permutations, so no operand is ever permuted in memory.
struct GemmTask {
  int64_t a_off, b_off, c_off;
  int32_t M, N, K, pad0;
  int64_t tAM, tAK, tBK, tBN, tCM, tCN;
};

// One output tile of one C block: the tasks [tbeg,tend) all write the
// same C block and are summed into the tile in registers.  With slot >= 0
// the unit computes a partial tile (split over the task list) into slot
// `slot` of the plan's workspace, and a reduction pass writes C.
struct WorkUnit {
  int32_t tbeg, tend, m0, n0;
  int32_t slot, pad;
};

// The argument list shared by every contraction kernel (k_gemm variants
// and k_skinny), in declaration order.
struct KernelArgs {
  const double *A, *B;
  double *C;
  const GemmTask *tasks;
  const WorkUnit *units;
  int nunits;
  const int64_t *tab;
  double alpha, beta;
  static constexpr int kCount = 9;
  void pointers(void **a) {
    a[0] = &A; a[1] = &B; a[2] = &C; a[3] = &tasks; a[4] = &units;
    a[5] = &nunits; a[6] = &tab; a[7] = &alpha; a[8] = &beta;
  }
};

}  // namespace itx

struct itx_plan {
  itx_ctx *ctx = nullptr;  // owner of the device block (its caching pool)
  int prec = ITX_PREC_F64;
  int tile = 0;        // index into the tile-shape table of the engine
  bool a_kcontig = false, b_kcontig = false;
  int ntasks = 0, nunits = 0;
  int grid = 0;
  double flops = 0;
  void *dev = nullptr;  // one device block: tasks | units | tables
  size_t dev_bytes = 0;
  const itx::GemmTask *d_tasks = nullptr;
  const itx::WorkUnit *d_units = nullptr;
  const int64_t *d_tab = nullptr;
  const void *fn = nullptr;  // kernel entry point and workgroup size
  const void *fn_scalar = nullptr;  // the same tile without vector loads
  int vec = 0;  // fn uses 16-byte operand loads (operands must be 16-byte aligned)
  int split = 0;  // BF16x3 with pre-split operands (itx_plan_set_split)
  int block = 0;
  // split plans: partial tiles into ws, then a reduction pass over
  // nred reduction units (d_red) writes C
  int nred = 0, grid_red = 0;
  const itx::WorkUnit *d_red = nullptr;
  const void *fn_red = nullptr;
  double *ws = nullptr;
  bool own_dev = true;  // dev came from the pool (false: the caller's memory)
};

// A fixed sequence of plan executions replayed as one HIP graph (one or
// two kernel nodes per plan: the contraction and, for split plans, its
// reduction pass).
struct itx_chain {
  std::vector<const itx_plan *> plans;
  std::vector<itx::KernelArgs> args;  // arguments the graph nodes hold now
  std::vector<hipGraphNode_t> nodes;  // 2 per stage; null where no launch
  std::vector<const void *> funcs;    // the kernel each node runs
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

struct itx_ctx {
  int device = 0;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // caching allocator: size -> free blocks; live block -> size
  std::multimap<size_t, void *> free_blocks;
  std::unordered_map<void *, size_t> live;
  size_t cached_bytes = 0;
  // pinned host blocks (itx_host_alloc): stream-ordered reuse like the
  // device pool, released at itx_destroy
  std::multimap<size_t, void *> host_free;
  std::unordered_map<void *, size_t> host_live;
  // reduction scratch (partials + results)
  double *red = nullptr;
  // device address of this device's copy of the library's static scratch
  // array (itx_contract_qn_noalloc), from hipGetSymbolAddress at creation
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  // rocblas handle of the large-block eigensolver (created on first use)
  void *rocblas = nullptr;
  // pinned host memory: reduction results written by the device
  // (hres, device view hres_dev; coherent, with a completion flag after
  // the HRES_SLOTS result slots) and the staging arena of uploads
  double *hres = nullptr, *hres_dev = nullptr;
  uint64_t hres_seq = 0;  // completion flag value of the last host reduction
  char *stage = nullptr;
  size_t stage_cap = 0, stage_off = 0;
  // the arena is used as two halves; stage_ev[h] is recorded on the stream
  // when the uploads move out of half h (stage_ev_used[h]: recorded at least
  // once), so that reusing a half waits only for the copies issued from it
  int stage_half = 0;
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  bool stage_ev_used[2] = {false, false};
  // plan cache for the one-shot contraction entry points
  std::unordered_map<std::string, itx_plan *> plan_cache;
  std::vector<std::string> plan_lru;
  // tile choice per contraction size class (autotuner results)
  int autotune = 0;
  std::map<std::string, int> tune;
  // every live plan and chain of the context: itx_destroy releases those
  // the caller has not (so no HIP object outlives the context)
  std::unordered_set<itx_plan *> plans;
  std::unordered_set<itx_chain *> chains;
};

namespace itx {

int set_err(itx_ctx *ctx, int code, const std::string &msg);
// releases the eigensolver's library handle (itx_destroy)
void eig_release(itx_ctx *ctx);
int hip_err(itx_ctx *ctx, hipError_t e, const char *where);

// Device allocation through the context's caching pool.
int pool_alloc(itx_ctx *ctx, size_t bytes, void **p);
int pool_free(itx_ctx *ctx, void *p);
// Host-to-device copy ordered on the context's stream that returns without
// waiting: the bytes go through the pinned staging arena (moving into one of
// its halves waits for the copies last issued from that half; copies larger
// than half the arena are synchronous).
int stage_upload(itx_ctx *ctx, void *dst, const void *src, size_t bytes);
// One upload of `total` bytes assembled from pieces (source, offset in the
// image, bytes) directly in the staging arena: no intermediate host image.
// Bytes of the image that no piece covers are left unspecified.
struct UploadPart {
  const void *src;
  size_t off, bytes;
};
int stage_upload_parts(itx_ctx *ctx, void *dst, const UploadPart *parts, int nparts, size_t total);
constexpr int HRES_SLOTS = 64;  // pinned result slots (multi-dot up to 16)

// Build a contraction plan (dense = one task).  Shared by the dense and
// QN entry points.  blk*: per-task block dims (ntasks x r).  With `into`
// (device memory of into_cap bytes) the plan's device block -- task and
// unit lists, offset tables, split-K partial tiles -- is laid out there
// instead of in the caching pool, and the tile comes from the tuning table
// or the cost model (the autotuner, which allocates, is skipped); a plan
// that does not fit fails with ITX_ERR_UNSUPPORTED.
int build_plan(itx_ctx *ctx, int rA, const int32_t *labA, int rB, const int32_t *labB,
               int rC, const int32_t *labC, int ntasks, const int64_t *offs,
               const int64_t *blkA, const int64_t *blkB, const int64_t *blkC_or_null,
               int prec, itx_plan **out, void *into = nullptr, size_t into_cap = 0);
int launch_plan(itx_ctx *ctx, const itx_plan *p, const double *A, const double *B, double *C,
                double alpha, double beta);
// Arguments of launch `which` of a plan (0: the contraction, 1: the
// reduction pass of a split plan) for operands A, B, C.
// the kernel a plan launches for operands A, B (vector loads when aligned)
const void *plan_fn(const itx_plan *p, const double *A, const double *B);
KernelArgs plan_args(const itx_plan *p, int which, const double *A, const double *B, double *C, double alpha,
                     double beta);
constexpr int SM_T_RED = 256;  // workgroup size of the reduction pass

constexpr int RED_SLOTS = 2048;  // partial-sum slots in ctx->red
constexpr int kSkinnyTile = 5;   // tile index of the skinny (FP64 VALU) path

}  // namespace itx

#define ITX_CHECK_HIP(ctx, call)                                   \
  do {                                                             \
    hipError_t _e = (call);                                        \
    if (_e != hipSuccess) return itx::hip_err((ctx), _e, #call);   \
  } while (0)
