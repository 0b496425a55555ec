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
// to their rows of Q.  Batched: block b of the grid solves matrix b;
// matrices up to 128 stay in LDS for the whole solve.
#include <cfloat>
#include <cstring>

#include "itx_internal.h"

namespace itx {
namespace {

constexpr int ET = 512;       // threads per matrix
constexpr int EMAX = 2048;    // largest n (d, e held in LDS)
constexpr int ROTG = 6144;    // rotation buffer of the global-memory kernel
constexpr int QL_MAXIT = 60;  // QL sweeps per eigenvalue before giving up

struct EigTask {
  int32_t n, pad;
  int64_t a_off, w_off, z_off;  // A (in/out, col-major), w, workspace Z (n*n, row-major)
};

// Givens pair of (f, g): r = sqrt(f^2 + g^2), s = f / r, c = g / r, with
// one reciprocal square root instead of hypot and two divisions (the
// operands of tql2's recurrence are bounded by the matrix norm, so the
// unscaled sum of squares cannot overflow).
__device__ __forceinline__ void givens(double f, double g, double &r, double &s, double &c) {
  const double t = fma(f, f, g * g);
  if (t == 0.0) {
    r = 0.0;
    s = 0.0;
    c = 1.0;
    return;
  }
  const double ri = rsqrt(t);
  r = t * ri;
  s = f * ri;
  c = g * ri;
}

// One implicit QL sweep of tql2 on (d, e)[l..m].  Executed by all lanes of
// wave 0 with identical values (uniform control flow; LDS stores of equal
// values); the plane rotations are appended to (rc, rs, ri).  Operands of
// the next step are loaded before the current step's dependency chain.
// Returns the number of rotations (negative - 1 on an early split).
__device__ int ql_sweep(double *d, double *e, int l, int m, double *rc, double *rs, int *ri) {
  double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
  double r = sqrt(fma(g, g, 1.0));
  g = d[m] - d[l] + e[l] / (g + copysign(r, g));
  double s = 1, c = 1, p = 0;
  int nrot = 0;
  double ei = e[m - 1], di = d[m - 1], di1 = d[m];
  for (int i = m - 1; i >= l; --i) {
    const int in = i > 0 ? i - 1 : 0;
    const double ei_next = e[in], di_next = d[in];
    const double f = s * ei, b = c * ei;
    double cs, sn;
    givens(f, g, r, sn, cs);
    e[i + 1] = r;
    if (r == 0) {
      d[i + 1] = di1 - p;
      e[m] = 0;
      return -(nrot + 1);  // early split: d[l], e[l] stay
    }
    s = sn;
    c = cs;
    g = di1 - p;
    r = (di - g) * s + 2.0 * c * b;
    p = s * r;
    d[i + 1] = g + p;
    g = c * r - b;
    rc[nrot] = c;
    rs[nrot] = s;
    ri[nrot] = i;
    ++nrot;
    di1 = di;
    di = di_next;
    ei = ei_next;
  }
  d[l] -= p;
  e[l] = g;
  e[m] = 0;
  return nrot;
}

// Implicit QL on the tridiagonal (d, e) with the deflation test of tql2
// (|e_m| against the largest |d|+|e| seen so far, so eigenvalues near zero
// converge to absolute accuracy), rotations applied to the columns of Z
// (element (r, c) at Z[r * zr + c * zc]).  Wave 0 runs the scalar
// recurrence -- it never depends on Z -- and fills the rotation buffer
// (cap entries) sweep by sweep; then every thread applies the buffered
// rotations to its rows of Z, carrying the updated element of one
// rotation into the next one in a register.  Returns 1 when an eigenvalue
// needs more than QL_MAXIT sweeps.
template <int NT>
__device__ int ql_phase(int n, double *d, double *e, double *rc, double *rs, int *ri, int cap, int *sh,
                        double *Z, int zr, int zc) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // wave 0's recurrence state (uniform across its lanes)
  int l = 0, iter = 0;
  bool newl = true;
  double tst1 = 0;
  while (true) {
    if (wave == 0) {
      int cnt = 0, done = 0, fail = 0;
      while (cnt + n <= cap) {
        if (l >= n) {
          done = 1;
          break;
        }
        if (newl) {
          tst1 = fmax(tst1, fabs(d[l]) + fabs(e[l]));
          iter = 0;
          newl = false;
        }
        int m = n - 1;
        for (int base = l; base < n - 1; base += 64) {
          const int mm = base + lane;
          const bool neg = mm < n - 1 && fabs(e[mm]) <= DBL_EPSILON * fmax(fabs(d[mm]) + fabs(d[mm + 1]), tst1);
          const unsigned long long bal = __ballot(neg);
          if (bal) {
            m = base + __ffsll(bal) - 1;
            break;
          }
        }
        if (m == l) {
          ++l;
          newl = true;
          continue;
        }
        if (iter >= QL_MAXIT) {
          fail = 1;
          break;
        }
        const int nr = ql_sweep(d, e, l, m, rc + cnt, rs + cnt, ri + cnt);
        cnt += nr < 0 ? -nr - 1 : nr;
        ++iter;
      }
      if (lane == 0) {
        sh[0] = cnt;
        sh[1] = done;
        sh[2] = fail;
      }
    }
    __syncthreads();
    const int cnt = sh[0], done = sh[1], fail = sh[2];
    for (int r = tid; r < n; r += NT) {
      double *zrow = Z + size_t(r) * zr;
      int prev = -2;
      double carry = 0;
      for (int q = 0; q < cnt; ++q) {
        const int i = ri[q];
        const double c = rc[q], sn = rs[q];
        const double f = (i + 1 == prev) ? carry : zrow[size_t(i + 1) * zc];
        const double z = zrow[size_t(i) * zc];
        zrow[size_t(i + 1) * zc] = sn * z + c * f;
        carry = c * z - sn * f;
        zrow[size_t(i) * zc] = carry;
        prev = i;
      }
    }
    __syncthreads();  // the buffer and sh are read before wave 0 rewrites them
    if (done || fail) return fail;
  }
}

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
  __shared__ double rc[ROTG], rs[ROTG];
  __shared__ int ri[ROTG];
  __shared__ double red[ET / 64];
  __shared__ int sh[4];

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
  if (ql_phase<ET>(n, d, e, rc, rs, ri, ROTG, sh, Z, n, 1)) {
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

// ---- LDS-resident form for n <= NMAX (<= 128) ----------------------
// The matrix lives in LDS for the whole solve: Householder reduction, Q
// formed in place from the stored reflectors (backward accumulation, the
// row and column of each new trailing index reset to the unit vector), the
// QL sweeps applied to Q's columns, then the sorted eigenpairs written out.
constexpr int ES = 256;

__device__ __forceinline__ double bsum_s(double v, double *red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0;
  for (int i = 0; i < ES / 64; ++i) s += red[i];
  return s;
}

template <int NMAX, int ROT>
__global__ __launch_bounds__(ES) void k_syev_lds(const EigTask *__restrict__ tasks, double *__restrict__ Abase,
                                                 double *__restrict__ Wbase, int *__restrict__ status) {
  static_assert(ROT >= 2 * NMAX, "rotation buffer below two sweeps");
  constexpr int LD = NMAX + 1;
  __shared__ double S[NMAX * LD];
  __shared__ double d[NMAX], e[NMAX], tau[NMAX], pv[NMAX];
  __shared__ double rc[ROT], rs[ROT];
  __shared__ int ri[ROT];
  __shared__ double red[ES / 64];
  __shared__ int sh[4];

  const EigTask tk = tasks[blockIdx.x];
  const int n = tk.n;
  const int tid = threadIdx.x;
  double *__restrict__ A = Abase + tk.a_off;
  double *__restrict__ W = Wbase + tk.w_off;
  if (n <= 0 || n > NMAX) {
    if (tid == 0) status[blockIdx.x] = (n <= 0) ? 0 : 2;
    return;
  }
  for (int x = tid; x < n * n; x += ES) {
    const int c = x / n, r = x - c * n;
    S[c * LD + r] = A[x];
  }
  __syncthreads();

  // ---- 1. Householder tridiagonalisation ----
  for (int k = 0; k + 2 < n; ++k) {
    const int L = n - k - 1;
    double *col = S + k * LD + (k + 1);
    double s = 0;
    for (int i = 1 + tid; i < L; i += ES) s += col[i] * col[i];
    const double sigma = bsum_s(s, red);
    const double x0 = col[0];
    double t = 0, alpha = x0, v0 = x0;
    if (sigma > 0) {
      const double mu = sqrt(x0 * x0 + sigma);
      v0 = (x0 <= 0) ? x0 - mu : x0 + mu;
      alpha = (x0 <= 0) ? mu : -mu;
      t = 2.0 / (v0 * v0 + sigma);
    }
    __syncthreads();  // col[0] read by all before it becomes v0
    if (tid == 0) {
      col[0] = v0;
      e[k] = alpha;
      tau[k] = t;
    }
    __syncthreads();
    if (t == 0) continue;
    for (int i = tid; i < L; i += ES) {
      double acc = 0;
      const double *Si = S + (k + 1) * LD + (k + 1 + i);
      for (int j = 0; j < L; ++j) acc += Si[j * LD] * col[j];
      pv[i] = t * acc;
    }
    __syncthreads();
    s = 0;
    for (int i = tid; i < L; i += ES) s += col[i] * pv[i];
    const double Kc = 0.5 * t * bsum_s(s, red);
    for (int i = tid; i < L; i += ES) pv[i] -= Kc * col[i];
    __syncthreads();
    for (int x = tid; x < L * L; x += ES) {
      const int j = x / L, i = x - j * L;
      S[(k + 1 + j) * LD + (k + 1 + i)] -= col[i] * pv[j] + pv[i] * col[j];
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += ES) d[i] = S[i * LD + i];
  if (tid == 0) {
    if (n >= 2) e[n - 2] = S[(n - 2) * LD + (n - 1)];
    e[n - 1] = 0;
    S[(n - 1) * LD + (n - 1)] = 1.0;
  }
  __syncthreads();

  // ---- 2. Q = H_0 ... H_{n-3} in place (trailing block grows backwards) ----
  for (int k = n - 3; k >= 0; --k) {
    const int j0 = k + 1;
    for (int i = j0 + tid; i < n; i += ES) {
      S[j0 * LD + i] = (i == j0) ? 1.0 : 0.0;  // column j0 (held v_{k+1})
      S[i * LD + j0] = (i == j0) ? 1.0 : 0.0;  // row j0
    }
    __syncthreads();
    const double t = tau[k];
    if (t != 0) {
      const double *v = S + k * LD + j0;  // v_k, first entry at row j0
      const int L = n - j0;
      for (int c = tid; c < L; c += ES) {
        double *Tc = S + (j0 + c) * LD + j0;
        double sc = 0;
        for (int i = 0; i < L; ++i) sc += v[i] * Tc[i];
        sc *= t;
        for (int i = 0; i < L; ++i) Tc[i] -= sc * v[i];
      }
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += ES) {
    S[i] = (i == 0) ? 1.0 : 0.0;       // column 0
    S[i * LD] = (i == 0) ? 1.0 : 0.0;  // row 0
  }
  __syncthreads();

  // ---- 3. implicit QL (tql2 deflation), rotations on Q's columns ----
  if (ql_phase<ES>(n, d, e, rc, rs, ri, ROT, sh, S, 1, LD)) {
    if (tid == 0) status[blockIdx.x] = 1;
    return;
  }

  // ---- 4. sort descending (stable rank), write eigenpairs ----
  for (int c = tid; c < n; c += ES) {
    const double dc = d[c];
    int rank = 0;
    for (int j = 0; j < n; ++j) rank += (d[j] > dc) || (d[j] == dc && j < c);
    ri[c] = rank;
  }
  __syncthreads();
  for (int c = tid; c < n; c += ES) W[ri[c]] = d[c];
  for (int x = tid; x < n * n; x += ES) {
    const int c = x / n, r = x - c * n;
    A[ri[c] * n + r] = S[c * LD + r];
  }
  if (tid == 0) status[blockIdx.x] = 0;
}

}  // namespace

int syev_batched(itx_ctx *ctx, int nmat, const int32_t *ns, const int64_t *aoffs, double *Abase,
                 const int64_t *woffs, double *Wbase) {
  if (nmat <= 0) return ITX_OK;
  // three size classes: LDS-resident up to 64 and up to 128, else the
  // global-memory kernel (with its n x n workspace per matrix)
  std::vector<EigTask> t64, t128, tg;
  int64_t zsize = 0;
  for (int b = 0; b < nmat; ++b) {
    if (ns[b] < 0 || ns[b] > EMAX) return set_err(ctx, ITX_ERR_UNSUPPORTED, "itx_syev: n above 2048");
    if (ns[b] == 0) continue;
    EigTask t{};
    t.n = ns[b];
    t.a_off = aoffs[b];
    t.w_off = woffs[b];
    if (ns[b] <= 64) {
      t64.push_back(t);
    } else if (ns[b] <= 128) {
      t128.push_back(t);
    } else {
      t.z_off = zsize;
      zsize += int64_t(ns[b]) * ns[b];
      tg.push_back(t);
    }
  }
  const int ntask = int(t64.size() + t128.size() + tg.size());
  if (ntask == 0) return ITX_OK;
  std::vector<EigTask> all;
  all.insert(all.end(), t64.begin(), t64.end());
  all.insert(all.end(), t128.begin(), t128.end());
  all.insert(all.end(), tg.begin(), tg.end());
  const size_t tbytes = sizeof(EigTask) * size_t(ntask);
  const size_t sbytes = sizeof(int) * size_t(ntask);
  const size_t zoff = (tbytes + sbytes + 255) & ~size_t(255);
  void *ws = nullptr;
  int rc = pool_alloc(ctx, zoff + sizeof(double) * size_t(zsize), &ws);
  if (rc) return rc;
  char *wsb = static_cast<char *>(ws);
  rc = stage_upload(ctx, wsb, all.data(), tbytes);
  if (rc) {
    pool_free(ctx, ws);
    return rc;
  }
  const EigTask *dt = reinterpret_cast<const EigTask *>(wsb);
  int *dst = reinterpret_cast<int *>(wsb + tbytes);
  const int n64 = int(t64.size()), n128 = int(t128.size()), ng = int(tg.size());
  if (n64)
    hipLaunchKernelGGL((k_syev_lds<64, 4096>), dim3(n64), dim3(ES), 0, ctx->stream, dt, Abase, Wbase, dst);
  if (n128)
    hipLaunchKernelGGL((k_syev_lds<128, 1024>), dim3(n128), dim3(ES), 0, ctx->stream, dt + n64, Abase, Wbase, dst + n64);
  if (ng)
    hipLaunchKernelGGL(k_syev, dim3(ng), dim3(ET), 0, ctx->stream, dt + n64 + n128, Abase, Wbase,
                       reinterpret_cast<double *>(wsb + zoff), dst + n64 + n128);
  hipError_t e = hipGetLastError();
  std::vector<int> st(size_t(ntask), 0);
  if (e == hipSuccess) e = hipMemcpyAsync(st.data(), dst, sbytes, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  pool_free(ctx, ws);
  if (e != hipSuccess) return hip_err(ctx, e, "itx_syev");
  for (int b = 0; b < ntask; ++b) {
    if (st[size_t(b)] == 1) return set_err(ctx, ITX_ERR_NOCONV, "itx_syev: QL iteration did not converge");
    if (st[size_t(b)] != 0) return set_err(ctx, ITX_ERR_ARG, "itx_syev: size class mismatch");
  }
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
