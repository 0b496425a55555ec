// FP64 vector kernels of libitx: dot / nrm2 / multi-dot (two-pass,
// deterministic wavefront reductions), axpy, scal, linear combinations and
// the permuted add used by ITensor's PlusEQ.
//
// Reference call sites: ddot/dnrm2/daxpy/dscal wrappers
// (tensor/lapack_wrap.cc:26-75,353) reached from norm (itensor.cc:754-766,
// itdata/dense.cc:150-156), += (itensor.cc:1207-1264, itdata/dense.cc:362-415)
// and the Davidson dots eltC(dag(V)*q) (iterativesolvers.h:169,315,393).
#include <cstring>

#include <atomic>
#include <chrono>
#include "itx_internal.h"

namespace itx {
namespace {

constexpr int VT = 256;
constexpr int MAXV = 16;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// block-wide sum; result valid in thread 0
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double part[VT / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) part[w] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < VT / 64; ++i) r += part[i];
  __syncthreads();
  return r;
}

struct VPtrs {
  const double *p[MAXV];
};
struct VCoef {
  double c[MAXV];
};

// partial[j*nbx + bx] = sum over this block's chunk of V_j . y
__global__ __launch_bounds__(VT) void k_dot_partial(VPtrs V, const double *__restrict__ y, size_t n,
                                                    double *__restrict__ partial) {
  const int j = blockIdx.y;
  const double *__restrict__ x = V.p[j];
  double s = 0;
  const size_t stride = size_t(gridDim.x) * VT;
  for (size_t i = size_t(blockIdx.x) * VT + threadIdx.x; i < n; i += stride) s += x[i] * y[i];
  s = block_sum(s);
  if (threadIdx.x == 0) partial[size_t(j) * gridDim.x + blockIdx.x] = s;
}

// out[j] = f(sum_b partial[j*nb + b]), f = identity or sqrt
__global__ __launch_bounds__(VT) void k_dot_final(const double *__restrict__ partial, int nb,
                                                  double *__restrict__ out, int do_sqrt) {
  const int j = blockIdx.x;
  double s = 0;
  for (int b = threadIdx.x; b < nb; b += VT) s += partial[size_t(j) * nb + b];
  s = block_sum(s);
  if (threadIdx.x == 0) out[j] = do_sqrt ? sqrt(s) : s;
}

// Single-launch form for results read by the host: blocks write partials,
// the last block to arrive for vector j (arrival counter, fence) sums them
// in the same fixed order as k_dot_final and resets the counter; the last
// vector to finish sets the host-visible completion flag to seq after a
// system-scope fence, so the host polls one word instead of synchronising
// the stream.
__global__ __launch_bounds__(VT) void k_dot_host(VPtrs V, const double *__restrict__ y, size_t n,
                                                 double *partial, unsigned *cnt, double *out, int do_sqrt,
                                                 unsigned long long *flag, unsigned long long seq) {
  __shared__ int last;
  const int j = blockIdx.y, nb = gridDim.x, nv = gridDim.y;
  const double *__restrict__ x = V.p[j];
  double s = 0;
  const size_t stride = size_t(nb) * VT;
  for (size_t i = size_t(blockIdx.x) * VT + threadIdx.x; i < n; i += stride) s += x[i] * y[i];
  s = block_sum(s);
  if (threadIdx.x == 0) {
    partial[size_t(j) * nb + blockIdx.x] = s;
    __threadfence();
    last = (atomicAdd(&cnt[j], 1u) == unsigned(nb - 1));
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  const volatile double *pv = partial + size_t(j) * nb;
  double t = 0;
  for (int b = threadIdx.x; b < nb; b += VT) t += pv[b];
  t = block_sum(t);
  if (threadIdx.x == 0) {
    out[j] = do_sqrt ? sqrt(t) : t;
    cnt[j] = 0u;
    __threadfence_system();
    if (nv == 1) {  // the only result is visible to the host: raise the flag
      *reinterpret_cast<volatile unsigned long long *>(flag) = seq;
    } else if (atomicAdd(&cnt[MAXV], 1u) == unsigned(nv - 1)) {
      cnt[MAXV] = 0u;
      __threadfence_system();
      *reinterpret_cast<volatile unsigned long long *>(flag) = seq;
    }
  }
}

__global__ __launch_bounds__(VT) void k_axpy(double a, const double *__restrict__ x, double *__restrict__ y,
                                             size_t n) {
  const size_t stride = size_t(gridDim.x) * VT;
  for (size_t i = size_t(blockIdx.x) * VT + threadIdx.x; i < n; i += stride) y[i] += a * x[i];
}

__global__ __launch_bounds__(VT) void k_scal(double a, double *__restrict__ x, size_t n) {
  const size_t stride = size_t(gridDim.x) * VT;
  for (size_t i = size_t(blockIdx.x) * VT + threadIdx.x; i < n; i += stride) x[i] *= a;
}

__global__ __launch_bounds__(VT) void k_lincomb(VPtrs V, VCoef c, int nv, double *__restrict__ y, size_t n) {
  const size_t stride = size_t(gridDim.x) * VT;
  for (size_t i = size_t(blockIdx.x) * VT + threadIdx.x; i < n; i += stride) {
    double s = 0;
    for (int j = 0; j < nv; ++j) s += c.c[j] * V.p[j][i];
    y[i] = s;
  }
}

struct PermDesc {
  int r;
  int64_t dimY[ITX_MAX_RANK];
  int64_t sX[ITX_MAX_RANK];  // stride in X of Y's index j
};

__global__ __launch_bounds__(VT) void k_permute(const double *__restrict__ X, double *__restrict__ Y,
                                                PermDesc d, size_t n, double alpha, double beta) {
  const size_t stride = size_t(gridDim.x) * VT;
  for (size_t i = size_t(blockIdx.x) * VT + threadIdx.x; i < n; i += stride) {
    size_t rem = i;
    int64_t xo = 0;
    for (int j = 0; j < d.r; ++j) {
      const int64_t dj = d.dimY[j];
      const int64_t q = int64_t(rem / size_t(dj));
      xo += (int64_t(rem) - q * dj) * d.sX[j];
      rem = size_t(q);
    }
    const double v = alpha * X[xo];
    Y[i] = (beta == 0.0) ? v : v + beta * Y[i];
  }
}

// one task of itx_block_copy, decoded from the int64 descriptor array
__global__ __launch_bounds__(VT) void k_block_copy(const double *__restrict__ X, double *__restrict__ Y,
                                                   const int64_t *__restrict__ desc, const int64_t *__restrict__ start,
                                                   int ntasks, int r, double alpha, double beta) {
  // blockIdx.x walks a flattened (task, chunk) space: start[t] = first chunk of t
  const int64_t chunk = blockIdx.x;
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {  // last t with start[t] <= chunk
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= chunk) lo = mid; else hi = mid - 1;
  }
  const int t = lo;
  const int64_t *d = desc + int64_t(t) * (2 + 3 * r);
  const int64_t *dims = d + 2, *ss = d + 2 + r, *ds = d + 2 + 2 * r;
  int64_t n = 1;
  for (int j = 0; j < r; ++j) n *= dims[j];
  const int64_t i = (chunk - start[t]) * VT + threadIdx.x;
  if (i >= n) return;
  int64_t rem = i, so = d[0], dof = d[1];
  for (int j = 0; j < r; ++j) {
    const int64_t q = rem / dims[j];
    const int64_t x = rem - q * dims[j];
    so += x * ss[j];
    dof += x * ds[j];
    rem = q;
  }
  const double v = alpha * X[so];
  Y[dof] = (beta == 0.0) ? v : v + beta * Y[dof];
}

unsigned grid_for(size_t n) {
  size_t b = (n + VT * 4 - 1) / (VT * 4);
  if (b < 1) b = 1;
  if (b > 4096) b = 4096;
  return unsigned(b);
}

int multi_dot_dev(itx_ctx *ctx, const double *const *V, int nv, const double *y, size_t n, double *dout,
                  int do_sqrt) {
  if (nv <= 0) return ITX_OK;
  if (nv > MAXV) return set_err(ctx, ITX_ERR_UNSUPPORTED, "multi_dot: more than 16 vectors");
  VPtrs vp{};
  for (int j = 0; j < nv; ++j) vp.p[j] = V[j];
  int nbx = int(grid_for(n));
  const int cap = RED_SLOTS / nv;
  if (nbx > cap) nbx = cap;
  hipLaunchKernelGGL(k_dot_partial, dim3(nbx, nv), dim3(VT), 0, ctx->stream, vp, y, n, ctx->red);
  hipLaunchKernelGGL(k_dot_final, dim3(nv), dim3(VT), 0, ctx->stream, ctx->red, nbx, dout, do_sqrt);
  ITX_CHECK_HIP(ctx, hipGetLastError());
  return ITX_OK;
}

// V_j . y (or |y| with do_sqrt) for j < nv into the host: one launch into
// the coherent result slots, then a poll of the completion flag.  A poll
// that runs past 50 ms falls back to synchronising the stream, which also
// reports a failed launch.
int dots_to_host(itx_ctx *ctx, const double *const *V, int nv, const double *y, size_t n, double *hdst,
                 int do_sqrt) {
  if (nv <= 0) return ITX_OK;
  if (nv > MAXV) return set_err(ctx, ITX_ERR_UNSUPPORTED, "multi_dot: more than 16 vectors");
  VPtrs vp{};
  for (int j = 0; j < nv; ++j) vp.p[j] = V[j];
  int nbx = int(grid_for(n));
  const int cap = RED_SLOTS / nv;
  if (nbx > cap) nbx = cap;
  const unsigned long long seq = ++ctx->hres_seq;
  volatile unsigned long long *hflag = reinterpret_cast<volatile unsigned long long *>(ctx->hres + HRES_SLOTS);
  auto *dflag = reinterpret_cast<unsigned long long *>(ctx->hres_dev + HRES_SLOTS);
  auto *cnt = reinterpret_cast<unsigned *>(ctx->red + RED_SLOTS);
  hipLaunchKernelGGL(k_dot_host, dim3(nbx, nv), dim3(VT), 0, ctx->stream, vp, y, n, ctx->red, cnt, ctx->hres_dev,
                     do_sqrt, dflag, seq);
  ITX_CHECK_HIP(ctx, hipGetLastError());
  const auto t0 = std::chrono::steady_clock::now();
  for (long it = 0; *hflag != seq; ++it) {
    if ((it & 1023) == 1023 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) {
      ITX_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
      if (*hflag != seq) return set_err(ctx, ITX_ERR_HIP, "reduction: completion flag not set");
      break;
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  for (int j = 0; j < nv; ++j) hdst[j] = ctx->hres[j];
  return ITX_OK;
}

}  // namespace
}  // namespace itx

extern "C" {

int itx_dot_dev(itx_ctx *ctx, const double *x, const double *y, size_t n, double *dresult) {
  if (!ctx) return ITX_ERR_ARG;
  const double *V[1] = {x};
  return itx::multi_dot_dev(ctx, V, 1, y, n, dresult, 0);
}

int itx_dot(itx_ctx *ctx, const double *x, const double *y, size_t n, double *result) {
  if (!ctx || !result) return ITX_ERR_ARG;
  const double *V[1] = {x};
  return itx::dots_to_host(ctx, V, 1, y, n, result, 0);
}

int itx_nrm2(itx_ctx *ctx, const double *x, size_t n, double *result) {
  if (!ctx || !result) return ITX_ERR_ARG;
  const double *V[1] = {x};
  return itx::dots_to_host(ctx, V, 1, x, n, result, 1);
}

int itx_multi_dot(itx_ctx *ctx, const double *const *V, int nv, const double *y, size_t n, double *out) {
  if (!ctx || !out || nv < 0 || nv > 16) return ITX_ERR_ARG;
  if (nv == 0) return ITX_OK;
  return itx::dots_to_host(ctx, V, nv, y, n, out, 0);
}

int itx_axpy(itx_ctx *ctx, double a, const double *x, double *y, size_t n) {
  if (!ctx) return ITX_ERR_ARG;
  if (n == 0) return ITX_OK;
  hipLaunchKernelGGL(itx::k_axpy, dim3(itx::grid_for(n)), dim3(itx::VT), 0, ctx->stream, a, x, y, n);
  ITX_CHECK_HIP(ctx, hipGetLastError());
  return ITX_OK;
}

int itx_scal(itx_ctx *ctx, double a, double *x, size_t n) {
  if (!ctx) return ITX_ERR_ARG;
  if (n == 0) return ITX_OK;
  hipLaunchKernelGGL(itx::k_scal, dim3(itx::grid_for(n)), dim3(itx::VT), 0, ctx->stream, a, x, n);
  ITX_CHECK_HIP(ctx, hipGetLastError());
  return ITX_OK;
}

int itx_lincomb(itx_ctx *ctx, const double *const *V, const double *c, int nv, double *y, size_t n) {
  if (!ctx || nv < 0 || nv > itx::MAXV) return ITX_ERR_ARG;
  if (n == 0) return ITX_OK;
  itx::VPtrs vp{};
  itx::VCoef vc{};
  for (int j = 0; j < nv; ++j) {
    vp.p[j] = V[j];
    vc.c[j] = c[j];
  }
  hipLaunchKernelGGL(itx::k_lincomb, dim3(itx::grid_for(n)), dim3(itx::VT), 0, ctx->stream, vp, vc, nv, y, n);
  ITX_CHECK_HIP(ctx, hipGetLastError());
  return ITX_OK;
}

int itx_permute(itx_ctx *ctx, const double *X, double *Y, int r, const int64_t *dimsY, const int32_t *perm,
                double alpha, double beta) {
  if (!ctx || r < 0 || r > ITX_MAX_RANK) return ITX_ERR_ARG;
  itx::PermDesc d{};
  d.r = r;
  // X dims: dimsX[perm[j]] = dimsY[j]; strides of X are column-major in X's order
  int64_t dimsX[ITX_MAX_RANK];
  std::vector<int> seen(r, 0);
  for (int j = 0; j < r; ++j) {
    if (perm[j] < 0 || perm[j] >= r || seen[perm[j]]) return itx::set_err(ctx, ITX_ERR_ARG, "itx_permute: bad perm");
    seen[perm[j]] = 1;
    dimsX[perm[j]] = dimsY[j];
  }
  int64_t sX[ITX_MAX_RANK];
  int64_t s = 1;
  size_t n = 1;
  for (int i = 0; i < r; ++i) {
    sX[i] = s;
    s *= dimsX[i];
  }
  for (int j = 0; j < r; ++j) {
    d.dimY[j] = dimsY[j];
    d.sX[j] = sX[perm[j]];
    n *= size_t(dimsY[j]);
  }
  if (n == 0) return ITX_OK;
  hipLaunchKernelGGL(itx::k_permute, dim3(itx::grid_for(n)), dim3(itx::VT), 0, ctx->stream, X, Y, d, n, alpha,
                     beta);
  ITX_CHECK_HIP(ctx, hipGetLastError());
  return ITX_OK;
}

}  // extern "C"

extern "C" int itx_block_copy(itx_ctx *ctx, const double *X, double *Y, int ntasks, int r, const int64_t *desc,
                              double alpha, double beta) {
  if (!ctx || ntasks < 0 || r < 0 || r > ITX_MAX_RANK) return ITX_ERR_ARG;
  if (ntasks == 0) return ITX_OK;
  const size_t per = size_t(2 + 3 * r);
  std::vector<int64_t> start(ntasks);
  int64_t chunks = 0;
  for (int t = 0; t < ntasks; ++t) {
    int64_t n = 1;
    for (int j = 0; j < r; ++j) {
      const int64_t dj = desc[t * per + 2 + j];
      if (dj < 0) return itx::set_err(ctx, ITX_ERR_ARG, "itx_block_copy: negative dim");
      n *= dj;
    }
    start[t] = chunks;
    chunks += (n + itx::VT - 1) / itx::VT;
  }
  if (chunks == 0) return ITX_OK;
  if (chunks > 0x7fffffff) return itx::set_err(ctx, ITX_ERR_UNSUPPORTED, "itx_block_copy: too large");
  void *buf = nullptr;
  const size_t dbytes = sizeof(int64_t) * per * ntasks;
  const size_t sbytes = sizeof(int64_t) * ntasks;
  int rc = itx::pool_alloc(ctx, dbytes + sbytes, &buf);
  if (rc) return rc;
  std::vector<int64_t> host(per * ntasks + ntasks);
  std::memcpy(host.data(), desc, dbytes);
  std::memcpy(host.data() + per * ntasks, start.data(), sbytes);
  rc = itx::stage_upload(ctx, buf, host.data(), dbytes + sbytes);
  if (rc) {
    itx::pool_free(ctx, buf);
    return rc;
  }
  const int64_t *dd = static_cast<const int64_t *>(buf);
  hipLaunchKernelGGL(itx::k_block_copy, dim3(unsigned(chunks)), dim3(itx::VT), 0, ctx->stream, X, Y, dd,
                     dd + per * ntasks, ntasks, r, alpha, beta);
  hipError_t e = hipGetLastError();
  itx::pool_free(ctx, buf);  // stream-ordered: reused only by later work
  if (e != hipSuccess) return itx::hip_err(ctx, e, "itx_block_copy");
  return ITX_OK;
}
