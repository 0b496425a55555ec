// FP64 symmetric eigensolver of libitx (one workgroup per matrix).
//
// Replaces dsyev('V','U') as called by diagHermitian on -M
// (tensor/algs_impl.h:91-139, tensor/algs.cc:34-41): eigenvalues come out
// in descending order with orthonormal eigenvectors as columns.
//
// Method (textbook, not LAPACK's code): Householder reduction to
// tridiagonal form T = Q^T A Q with explicit accumulation of Q, then the
// implicit QL iteration with Wilkinson-type shift on T (Bowdler, Martin,
// Reinsch & Wilkinson, Numer. Math. 11 (1968) 293, "tql2"), whose plane
// rotations are applied to Q.  Thread 0 runs the scalar QL recurrence of
// one sweep and records the rotations in LDS; all threads then apply them
// to their rows of Q.  Batched: block b of the grid solves matrix b.
#include <cfloat>
#include <cstring>

#include "itx_internal.h"

namespace itx {
namespace {

constexpr int ET = 512;       // threads per matrix
constexpr int EMAX = 2048;    // largest n held in LDS (d, e, rotations)
constexpr int QL_MAXIT = 60;  // QL sweeps per eigenvalue before giving up

struct EigTask {
  int32_t n, pad;
  int64_t a_off, w_off, z_off;  // A (in/out, col-major), w, workspace Z (n*n, row-major)
};

__device__ __forceinline__ double bsum(double v, double *red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0;
  for (int i = 0; i < ET / 64; ++i) s += red[i];
  return s;
}

__global__ __launch_bounds__(ET) void k_syev(const EigTask *__restrict__ tasks, double *__restrict__ Abase,
                                             double *__restrict__ Wbase, double *__restrict__ Zbase,
                                             int *__restrict__ status) {
  __shared__ double d[EMAX], e[EMAX];
  __shared__ double rc[EMAX], rs[EMAX];
  __shared__ int ri[EMAX];
  __shared__ double red[ET / 64];
  __shared__ int sh_m, sh_nrot, sh_fail;

  const EigTask tk = tasks[blockIdx.x];
  const int n = tk.n;
  const int tid = threadIdx.x;
  double *__restrict__ A = Abase + tk.a_off;
  double *__restrict__ W = Wbase + tk.w_off;
  double *__restrict__ Z = Zbase + tk.z_off;
  if (n <= 0) return;

  // ---- 1. Householder tridiagonalisation (full symmetric update) ----
  for (int k = 0; k + 2 < n; ++k) {
    const int L = n - k - 1;
    double *col = A + size_t(k) * n + (k + 1);  // x = A[k+1:, k]
    double s = 0;
    for (int i = 1 + tid; i < L; i += ET) s += col[i] * col[i];
    const double sigma = bsum(s, red);
    const double x0 = col[0];
    double tau = 0, alpha = x0;
    if (sigma > 0) {
      const double mu = sqrt(x0 * x0 + sigma);
      const double v0 = (x0 <= 0) ? x0 - mu : x0 + mu;
      alpha = (x0 <= 0) ? mu : -mu;
      tau = 2.0 / (v0 * v0 + sigma);
      __syncthreads();
      if (tid == 0) col[0] = v0;  // col now holds v
    }
    if (tid == 0) {
      e[k] = alpha;
      rc[k] = tau;  // rc doubles as tau storage during this phase
    }
    if (tau == 0) {
      __syncthreads();
      continue;
    }
    __syncthreads();
    // p = tau * S v  with S = A[k+1:, k+1:]; p stored in W[0:L] temporarily
    for (int i = tid; i < L; i += ET) {
      double acc = 0;
      const double *Si = A + size_t(k + 1) * n + (k + 1 + i);
      for (int j = 0; j < L; ++j) acc += Si[size_t(j) * n] * col[j];
      W[i] = tau * acc;
    }
    __syncthreads();
    s = 0;
    for (int i = tid; i < L; i += ET) s += col[i] * W[i];
    const double Kc = 0.5 * tau * bsum(s, red);
    for (int i = tid; i < L; i += ET) W[i] -= Kc * col[i];  // w = p - K v
    __syncthreads();
    // S -= v w^T + w v^T
    for (int i = tid; i < L; i += ET) {
      const double vi = col[i], wi = W[i];
      double *Si = A + size_t(k + 1) * n + (k + 1 + i);
      for (int j = 0; j < L; ++j) Si[size_t(j) * n] -= vi * W[j] + wi * col[j];
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += ET) d[i] = A[size_t(i) * n + i];
  if (tid == 0) {
    if (n >= 2) e[n - 2] = A[size_t(n - 2) * n + (n - 1)];
    e[n - 1] = 0;
    sh_fail = 0;
  }
  __syncthreads();

  // ---- 2. Q = H_0 H_1 ... H_{n-3}, accumulated backwards, Z row-major ----
  for (size_t x = tid; x < size_t(n) * n; x += ET) Z[x] = ((x / n) == (x % n)) ? 1.0 : 0.0;
  __syncthreads();
  for (int k = n - 3; k >= 0; --k) {
    const double tau = rc[k];
    if (tau == 0) continue;
    const int L = n - k - 1;
    const double *v = A + size_t(k) * n + (k + 1);
    // Z_sub -= tau v (v^T Z_sub); columns c handled by threads
    for (int c = tid; c < L; c += ET) {
      double sc = 0;
      for (int i = 0; i < L; ++i) sc += v[i] * Z[size_t(k + 1 + i) * n + (k + 1 + c)];
      sc *= tau;
      for (int i = 0; i < L; ++i) Z[size_t(k + 1 + i) * n + (k + 1 + c)] -= sc * v[i];
    }
    __syncthreads();
  }

  // ---- 3. implicit QL on (d, e), rotations applied to Z's columns ----
  // Deflation test of tql2: |e_m| small against the largest |d|+|e| seen so
  // far (tst1), not only against its neighbours, so eigenvalues near zero
  // (rank-deficient density matrices) converge to absolute accuracy.
  int failed = 0;
  double tst1 = 0;  // used by thread 0 only
  for (int l = 0; l < n && !failed; ++l) {
    int iter = 0;
    if (tid == 0) tst1 = fmax(tst1, fabs(d[l]) + fabs(e[l]));
    while (true) {
      if (tid == 0) {
        int m = l;
        for (; m < n - 1; ++m) {
          const double dd = fabs(d[m]) + fabs(d[m + 1]);
          if (fabs(e[m]) <= DBL_EPSILON * fmax(dd, tst1)) break;
        }
        sh_m = m;
        sh_nrot = 0;
        if (m != l) {
          if (iter >= QL_MAXIT) {
            sh_fail = 1;
          } else {
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + copysign(r, g));
            double s = 1, c = 1, p = 0;
            int nrot = 0;
            bool early = false;
            for (int i = m - 1; i >= l; --i) {
              const double f = s * e[i], b = c * e[i];
              r = hypot(f, g);
              e[i + 1] = r;
              if (r == 0) {
                d[i + 1] -= p;
                e[m] = 0;
                early = true;
                break;
              }
              s = f / r;
              c = g / r;
              g = d[i + 1] - p;
              r = (d[i] - g) * s + 2.0 * c * b;
              p = s * r;
              d[i + 1] = g + p;
              g = c * r - b;
              rc[nrot] = c;
              rs[nrot] = s;
              ri[nrot] = i;
              ++nrot;
            }
            if (!early) {
              d[l] -= p;
              e[l] = g;
              e[m] = 0;
            }
            sh_nrot = nrot;
          }
        }
      }
      __syncthreads();
      const int m = sh_m, nrot = sh_nrot, fail = sh_fail;
      __syncthreads();  // every thread has read sh_* before thread 0 rewrites them
      failed = fail;
      if (m == l || fail) break;
      for (int r = tid; r < n; r += ET) {
        double *Zr = Z + size_t(r) * n;
        for (int q = 0; q < nrot; ++q) {
          const int i = ri[q];
          const double f = Zr[i + 1], z = Zr[i];
          Zr[i + 1] = rs[q] * z + rc[q] * f;
          Zr[i] = rc[q] * z - rs[q] * f;
        }
      }
      __syncthreads();
      ++iter;
    }
  }
  __syncthreads();
  if (sh_fail) {
    if (tid == 0) status[blockIdx.x] = 1;
    return;
  }

  // ---- 4. sort descending (stable rank), write eigenpairs ----
  for (int c = tid; c < n; c += ET) {
    const double dc = d[c];
    int rank = 0;
    for (int j = 0; j < n; ++j) rank += (d[j] > dc) || (d[j] == dc && j < c);
    ri[c] = rank;
  }
  __syncthreads();
  for (int c = tid; c < n; c += ET) W[ri[c]] = d[c];
  for (size_t x = tid; x < size_t(n) * n; x += ET) {
    const size_t r = x / n, c = x % n;
    A[size_t(ri[c]) * n + r] = Z[x];
  }
  if (tid == 0) status[blockIdx.x] = 0;
}

}  // namespace

int syev_batched(itx_ctx *ctx, int nmat, const int32_t *ns, const int64_t *aoffs, double *Abase,
                 const int64_t *woffs, double *Wbase) {
  if (nmat <= 0) return ITX_OK;
  std::vector<EigTask> tasks(nmat);
  int64_t zsize = 0;
  for (int b = 0; b < nmat; ++b) {
    if (ns[b] < 0 || ns[b] > EMAX) return set_err(ctx, ITX_ERR_UNSUPPORTED, "itx_syev: n above 2048");
    tasks[b].n = ns[b];
    tasks[b].a_off = aoffs[b];
    tasks[b].w_off = woffs[b];
    tasks[b].z_off = zsize;
    zsize += int64_t(ns[b]) * ns[b];
  }
  void *ws = nullptr;
  size_t tbytes = sizeof(EigTask) * nmat;
  size_t sbytes = sizeof(int) * nmat;
  size_t zoff = (tbytes + sbytes + 255) & ~size_t(255);
  int rc = pool_alloc(ctx, zoff + sizeof(double) * size_t(zsize), &ws);
  if (rc) return rc;
  char *wsb = static_cast<char *>(ws);
  std::vector<char> hbuf(tbytes + sbytes, 0);
  std::memcpy(hbuf.data(), tasks.data(), tbytes);
  hipError_t e = hipMemcpyAsync(wsb, hbuf.data(), tbytes + sbytes, hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) {
    pool_free(ctx, ws);
    return hip_err(ctx, e, "itx_syev upload");
  }
  hipLaunchKernelGGL(k_syev, dim3(nmat), dim3(ET), 0, ctx->stream, reinterpret_cast<const EigTask *>(wsb), Abase,
                     Wbase, reinterpret_cast<double *>(wsb + zoff), reinterpret_cast<int *>(wsb + tbytes));
  e = hipGetLastError();
  std::vector<int> st(nmat, 0);
  if (e == hipSuccess)
    e = hipMemcpyAsync(st.data(), wsb + tbytes, sbytes, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  pool_free(ctx, ws);
  if (e != hipSuccess) return hip_err(ctx, e, "itx_syev");
  for (int b = 0; b < nmat; ++b)
    if (st[b] != 0) return set_err(ctx, ITX_ERR_NOCONV, "itx_syev: QL iteration did not converge");
  return ITX_OK;
}

}  // namespace itx

extern "C" int itx_syev(itx_ctx *ctx, int n, double *A, double *w) {
  if (!ctx || n < 0) return ITX_ERR_ARG;
  if (n == 0) return ITX_OK;
  const int32_t ns[1] = {n};
  const int64_t ao[1] = {0}, wo[1] = {0};
  return itx::syev_batched(ctx, 1, ns, ao, A, wo, w);
}

extern "C" int itx_syev_batched(itx_ctx *ctx, int nmat, const int32_t *n, const int64_t *a_offs, double *A,
                                const int64_t *w_offs, double *w) {
  if (!ctx || nmat < 0) return ITX_ERR_ARG;
  if (nmat == 0) return ITX_OK;
  return itx::syev_batched(ctx, nmat, n, a_offs, A, w_offs, w);
}
