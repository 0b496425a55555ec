// Context, stream and caching device allocator of libitx.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include <vector>
#include "itx_internal.h"

namespace {
std::mutex g_create_mu;
std::string g_create_err;  // error text of the last failed itx_create
}  // namespace

namespace itx {

int set_err(itx_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int hip_err(itx_ctx *ctx, hipError_t e, const char *where) {
  std::string m = std::string(where) + ": " + hipGetErrorString(e);
  if (ctx) ctx->err = m;
  return e == hipErrorOutOfMemory ? ITX_ERR_NOMEM : ITX_ERR_HIP;
}

int pool_alloc(itx_ctx *ctx, size_t bytes, void **p) {
  if (bytes == 0) bytes = 8;
  // round to 256 B so that blocks of nearby sizes are interchangeable
  bytes = (bytes + 255) & ~size_t(255);
  auto it = ctx->free_blocks.lower_bound(bytes);
  if (it != ctx->free_blocks.end() && it->first <= bytes + bytes / 4 + 4096) {
    *p = it->second;
    ctx->live[*p] = it->first;
    ctx->cached_bytes -= it->first;
    ctx->free_blocks.erase(it);
    return ITX_OK;
  }
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) {
    // release the cache and retry once
    (void)hipGetLastError();
    hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->free_blocks) hipFree(kv.second);
    ctx->free_blocks.clear();
    ctx->cached_bytes = 0;
    e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return hip_err(ctx, e, "hipMalloc");
  }
  ctx->live[q] = bytes;
  *p = q;
  return ITX_OK;
}

int pool_free(itx_ctx *ctx, void *p) {
  if (!p) return ITX_OK;
  auto it = ctx->live.find(p);
  if (it == ctx->live.end()) return set_err(ctx, ITX_ERR_ARG, "itx_free: pointer not from itx_malloc");
  // Stream order makes reuse safe: every later user of this block is
  // enqueued on the same stream after the work that used it before.
  ctx->free_blocks.emplace(it->second, p);
  ctx->cached_bytes += it->second;
  ctx->live.erase(it);
  return ITX_OK;
}

// Offset in the arena for an upload of `bytes` (at most half the arena):
// the next 256-byte aligned place in the current half, or the start of the
// other half once that half's last copies are complete.
static int stage_reserve(itx_ctx *ctx, size_t bytes, size_t *out) {
  const size_t half = ctx->stage_cap / 2;
  size_t off = (ctx->stage_off + 255) & ~size_t(255);
  if (off + bytes > size_t(ctx->stage_half) * half + half) {
    const int h = ctx->stage_half, nh = h ^ 1;
    if (ctx->stage_ev[0] && ctx->stage_ev[1]) {
      ITX_CHECK_HIP(ctx, hipEventRecord(ctx->stage_ev[h], ctx->stream));
      ctx->stage_ev_used[h] = true;
      if (ctx->stage_ev_used[nh]) ITX_CHECK_HIP(ctx, hipEventSynchronize(ctx->stage_ev[nh]));
    } else {
      // every earlier copy out of the arena is complete once the stream is
      ITX_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    ctx->stage_half = nh;
    off = size_t(nh) * half;
  }
  *out = off;
  return ITX_OK;
}

int stage_upload(itx_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return ITX_OK;
  if (!ctx->stage || bytes > ctx->stage_cap / 2) {
    ITX_CHECK_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    ITX_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return ITX_OK;
  }
  size_t off = 0;
  int rc = stage_reserve(ctx, bytes, &off);
  if (rc) return rc;
  std::memcpy(ctx->stage + off, src, bytes);
  ITX_CHECK_HIP(ctx, hipMemcpyAsync(dst, ctx->stage + off, bytes, hipMemcpyHostToDevice, ctx->stream));
  ctx->stage_off = off + bytes;
  return ITX_OK;
}

int stage_upload_parts(itx_ctx *ctx, void *dst, const UploadPart *parts, int nparts, size_t total) {
  if (total == 0) return ITX_OK;
  if (!ctx->stage || total > ctx->stage_cap / 2) {
    std::vector<char> img(total);
    for (int i = 0; i < nparts; ++i)
      if (parts[i].bytes) std::memcpy(img.data() + parts[i].off, parts[i].src, parts[i].bytes);
    return stage_upload(ctx, dst, img.data(), total);
  }
  size_t off = 0;
  int rc = stage_reserve(ctx, total, &off);
  if (rc) return rc;
  for (int i = 0; i < nparts; ++i)
    if (parts[i].bytes) std::memcpy(ctx->stage + off + parts[i].off, parts[i].src, parts[i].bytes);
  ITX_CHECK_HIP(ctx, hipMemcpyAsync(dst, ctx->stage + off, total, hipMemcpyHostToDevice, ctx->stream));
  ctx->stage_off = off + total;
  return ITX_OK;
}

}  // namespace itx

namespace {
constexpr size_t kStageBytes = size_t(64) << 20;  // plan tables of large QN contractions stay asynchronous

// Device memory of itx_contract_qn_noalloc: one copy per device for the
// life of the process, shared by every call on that device.  In host code
// the name stands for a host-side shadow, not a device address: the address
// comes from hipGetSymbolAddress (itx_create, with the device current) and
// only that address is ever handed to kernels or copies.
constexpr size_t kScratchBytes = size_t(32) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

__global__ void k_fill(double *x, size_t n, double v) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t stride = size_t(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) x[i] = v;
}
}  // namespace

extern "C" {

int itx_create(int device, itx_ctx **out) {
  std::lock_guard<std::mutex> lk(g_create_mu);
  if (!out) return ITX_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) {
    g_create_err = "itx_create: no HIP device available";
    return ITX_ERR_HIP;
  }
  if (device < 0 || device >= ndev) {
    g_create_err = "itx_create: device ordinal out of range";
    return ITX_ERR_ARG;
  }
  e = hipSetDevice(device);
  if (e != hipSuccess) {
    g_create_err = std::string("hipSetDevice: ") + hipGetErrorString(e);
    return ITX_ERR_HIP;
  }
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) {
    g_create_err = std::string("hipGetDeviceProperties: ") + hipGetErrorString(e);
    return ITX_ERR_HIP;
  }
  auto *ctx = new itx_ctx();
  ctx->device = device;
  ctx->num_cus = prop.multiProcessorCount;
  ctx->stream = nullptr;
  // partial sums, then 64 doubles' worth of arrival counters (zero between
  // reductions: the last block of each reduction resets its own)
  e = hipMalloc(&ctx->red, sizeof(double) * (itx::RED_SLOTS + 64));
  if (e == hipSuccess) e = hipMemset(ctx->red + itx::RED_SLOTS, 0, sizeof(double) * 64);
  if (e != hipSuccess) {
    g_create_err = std::string("hipMalloc(scratch): ") + hipGetErrorString(e);
    delete ctx;
    return ITX_ERR_NOMEM;
  }
  void *hp = nullptr;
  // coherent: device writes become visible to the host without a stream
  // synchronisation (the completion flag is polled)
  e = hipHostMalloc(&hp, sizeof(double) * (itx::HRES_SLOTS + 1), hipHostMallocMapped | hipHostMallocCoherent);
  if (e == hipSuccess) std::memset(hp, 0, sizeof(double) * (itx::HRES_SLOTS + 1));
  void *hd = nullptr;
  if (e == hipSuccess) e = hipHostGetDevicePointer(&hd, hp, 0);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void **>(&ctx->stage), kStageBytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    g_create_err = std::string("hipHostMalloc: ") + hipGetErrorString(e);
    if (hp) hipHostFree(hp);
    hipFree(ctx->red);
    delete ctx;
    return ITX_ERR_NOMEM;
  }
  ctx->hres = static_cast<double *>(hp);
  ctx->hres_dev = static_cast<double *>(hd);
  ctx->stage_cap = kStageBytes;
  // without the events the arena falls back to a stream synchronisation
  if (hipEventCreateWithFlags(&ctx->stage_ev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->stage_ev[1], hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    if (ctx->stage_ev[0]) hipEventDestroy(ctx->stage_ev[0]);
    ctx->stage_ev[0] = ctx->stage_ev[1] = nullptr;
  }
  // this device's copy of the static scratch array (the device is current)
  void *sp = nullptr;
  e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !sp) {
    g_create_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    for (auto &ev : ctx->stage_ev)
      if (ev) hipEventDestroy(ev);
    hipHostFree(ctx->stage);
    hipHostFree(hp);
    hipFree(ctx->red);
    delete ctx;
    return ITX_ERR_HIP;
  }
  ctx->scratch = sp;
  ctx->scratch_bytes = kScratchBytes;
  // ITX_AUTOTUNE=1 turns the tile autotuner on; ITX_TUNE_FILE names a
  // tuning table to load (see itx_tune_load)
  if (const char *a = std::getenv("ITX_AUTOTUNE")) ctx->autotune = std::atoi(a) ? 1 : 0;
  if (const char *f = std::getenv("ITX_TUNE_FILE")) itx_tune_load(ctx, f);
  *out = ctx;
  return ITX_OK;
}

int itx_destroy(itx_ctx *ctx) {
  if (!ctx) return ITX_OK;
  hipStreamSynchronize(ctx->stream);
  // chains first (their graphs name the plans' kernels), then every plan
  // still alive (the one-shot cache's and the caller's), then the pool
  while (!ctx->chains.empty()) itx_chain_destroy(ctx, *ctx->chains.begin());
  ctx->plan_cache.clear();
  while (!ctx->plans.empty()) itx_plan_destroy(ctx, *ctx->plans.begin());
  itx::eig_release(ctx);
  for (auto &kv : ctx->free_blocks) hipFree(kv.second);
  for (auto &kv : ctx->live) hipFree(kv.first);
  hipFree(ctx->red);
  for (auto &kv : ctx->host_free) hipHostFree(kv.second);
  for (auto &kv : ctx->host_live) hipHostFree(kv.first);
  if (ctx->hres) hipHostFree(ctx->hres);
  if (ctx->stage) hipHostFree(ctx->stage);
  for (auto &ev : ctx->stage_ev)
    if (ev) hipEventDestroy(ev);
  delete ctx;
  return ITX_OK;
}

int itx_set_stream(itx_ctx *ctx, void *stream) {
  if (!ctx) return ITX_ERR_ARG;
  // uploads and frees of the pool are ordered on the old stream: drain it
  if (reinterpret_cast<hipStream_t>(stream) != ctx->stream) ITX_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->stream = reinterpret_cast<hipStream_t>(stream);
  return ITX_OK;
}

void *itx_get_stream(itx_ctx *ctx) { return ctx ? reinterpret_cast<void *>(ctx->stream) : nullptr; }

int itx_synchronize(itx_ctx *ctx) {
  if (!ctx) return ITX_ERR_ARG;
  ITX_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ITX_OK;
}

const char *itx_last_error(itx_ctx *ctx) {
  if (!ctx) return g_create_err.c_str();
  return ctx->err.c_str();
}

int itx_device_cus(itx_ctx *ctx) { return ctx ? ctx->num_cus : 0; }

int itx_malloc(itx_ctx *ctx, size_t bytes, void **ptr) {
  if (!ctx || !ptr) return ITX_ERR_ARG;
  return itx::pool_alloc(ctx, bytes, ptr);
}

int itx_free(itx_ctx *ctx, void *ptr) {
  if (!ctx) return ITX_ERR_ARG;
  return itx::pool_free(ctx, ptr);
}

int itx_trim(itx_ctx *ctx) {
  if (!ctx) return ITX_ERR_ARG;
  ITX_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (auto &kv : ctx->free_blocks) hipFree(kv.second);
  ctx->free_blocks.clear();
  ctx->cached_bytes = 0;
  return ITX_OK;
}

int itx_host_alloc(itx_ctx *ctx, size_t bytes, void **ptr) {
  if (!ctx || !ptr) return ITX_ERR_ARG;
  if (bytes == 0) bytes = 8;
  bytes = (bytes + 4095) & ~size_t(4095);
  auto it = ctx->host_free.lower_bound(bytes);
  if (it != ctx->host_free.end() && it->first <= bytes + bytes / 4 + 65536) {
    *ptr = it->second;
    ctx->host_live[*ptr] = it->first;
    ctx->host_free.erase(it);
    return ITX_OK;
  }
  void *q = nullptr;
  ITX_CHECK_HIP(ctx, hipHostMalloc(&q, bytes, hipHostMallocDefault));
  ctx->host_live[q] = bytes;
  *ptr = q;
  return ITX_OK;
}

int itx_host_free(itx_ctx *ctx, void *ptr) {
  if (!ctx) return ITX_ERR_ARG;
  if (!ptr) return ITX_OK;
  auto it = ctx->host_live.find(ptr);
  if (it == ctx->host_live.end()) return itx::set_err(ctx, ITX_ERR_ARG, "itx_host_free: pointer not from itx_host_alloc");
  // stream order makes reuse safe: copies into or out of the block are
  // enqueued on the context's stream, after the copies that used it before
  ctx->host_free.emplace(it->second, ptr);
  ctx->host_live.erase(it);
  return ITX_OK;
}

int itx_memcpy_async(itx_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return ITX_ERR_ARG;
  if (bytes == 0) return ITX_OK;
  ITX_CHECK_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream));
  return ITX_OK;
}

int itx_memcpy_htod(itx_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return ITX_ERR_ARG;
  if (bytes == 0) return ITX_OK;
  ITX_CHECK_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  ITX_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ITX_OK;
}

int itx_memcpy_dtoh(itx_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return ITX_ERR_ARG;
  if (bytes == 0) return ITX_OK;
  ITX_CHECK_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  ITX_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ITX_OK;
}

int itx_memcpy_dtod(itx_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return ITX_ERR_ARG;
  if (bytes == 0) return ITX_OK;
  ITX_CHECK_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return ITX_OK;
}

int itx_fill(itx_ctx *ctx, double *x, size_t n, double v) {
  if (!ctx) return ITX_ERR_ARG;
  if (n == 0) return ITX_OK;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_fill, dim3(unsigned(blocks)), dim3(256), 0, ctx->stream, x, n, v);
  ITX_CHECK_HIP(ctx, hipGetLastError());
  return ITX_OK;
}

}  // extern "C"
