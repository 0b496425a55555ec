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
// rotations are applied to Q.  The work is split the way hybrid symmetric
// eigensolvers split it: the O(n^3) parts on the device (reduction and Q:
// one workgroup per matrix with the matrix resident in LDS up to 128, one
// 16-wave workgroup per matrix over global memory up to 512, one
// cooperative launch over the whole device per larger matrix; the
// rotations applied to Q's rows in parallel in a second launch),
// the O(n^2) scalar QL recurrence -- a chain of dependent flops that one
// host core runs far faster than one GPU lane -- on the host in between.
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <type_traits>
#include <vector>

#include <hip/hip_cooperative_groups.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "itx_internal.h"

namespace itx {
namespace {

constexpr int ET = 512;       // threads per matrix of the batched global-memory reduction
constexpr int EMAX = 2048;    // largest n
constexpr int QL_MAXIT = 60;  // QL sweeps per eigenvalue before giving up

struct EigTask {
  int32_t n, pad;
  int64_t a_off, w_off;  // A (in/out, col-major), w
  int64_t z_off;         // workspace: Q, column-major n x (n+1) (the last column unused)
  int64_t de_off;        // workspace: (d, e) handed to the host
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

// Reduction and Q for n > 128: A in global memory (overwritten by the
// reflectors), W as scratch for p, Q into the workspace.
// One workgroup per matrix over global memory (n > 128): used when a batch
// holds several such matrices that are cheaper side by side than one after
// the other through the cooperative kernel.
__global__ __launch_bounds__(ET) void k_syev(const EigTask *__restrict__ tasks, double *__restrict__ Abase,
                                             double *__restrict__ Wbase, double *__restrict__ Zbase) {
  __shared__ double d[EMAX], e[EMAX], tau_[EMAX];
  __shared__ double red[ET / 64];

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
      tau_[k] = tau;
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

  // ---- 2. Q = H_0 H_1 ... H_{n-3}, accumulated backwards, Z column-major ----
  // (one wave per column: the lanes read and update it coalesced)
  for (size_t x = tid; x < size_t(n) * n; x += ET) Z[x] = ((x / n) == (x % n)) ? 1.0 : 0.0;
  __syncthreads();
  {
    const int lane = tid & 63, wv = tid >> 6;
    for (int k = n - 3; k >= 0; --k) {
      const double tau = tau_[k];
      if (tau == 0) continue;
      const int L = n - k - 1;
      const double *v = A + size_t(k) * n + (k + 1);
      // Z_sub -= tau v (v^T Z_sub)
      for (int c = wv; c < L; c += ET / 64) {
        double *zc = Z + size_t(k + 1 + c) * n + (k + 1);
        double sc = 0;
        for (int i = lane; i < L; i += 64) sc += v[i] * zc[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) sc += __shfl_xor(sc, o, 64);
        sc *= tau;
        for (int i = lane; i < L; i += 64) zc[i] -= sc * v[i];
      }
      __syncthreads();
    }
  }

  // ---- 3. (d, e) to the host, which runs the QL recurrence ----
  double *DE = Zbase + tk.de_off;
  for (int i = tid; i < n; i += ET) {
    DE[i] = d[i];
    DE[n + i] = e[i];
  }
}

// ---- one workgroup of 16 waves per matrix, 128 < n <= 64 RA -------------
// The matrix stays in global memory (it fits the L2); the reduction makes
// one pass over the trailing block per Householder step.  Step k: wave 0
// brings column k up to date with the pending rank-2 update of step k-1
// (v', w'), forms the reflector (v, tau) from it and stores v in the
// column; then every thread walks its part of the trailing block
// [k+1, n)^2 -- rows lane + 64 a, columns = wave (mod 16) -- applying the
// pending update, storing the element (column k+1 also to LDS, where wave 0
// reads it for the next reflector) and accumulating p = S v with the
// updated value; the 16 partial sums of each row go through LDS, and wave 0
// forms K = tau/2 v.p and w = p - K v, the next pending update.  Each
// element is read and written once per step and every thread has up to
// RA x (its columns) independent loads in flight.  Q is accumulated
// backwards with the same column ownership: a wave owns whole columns of
// Q, so the reflections need no barrier between waves.
constexpr int MT = 1024, MW = MT / 64;

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int RA, int UP, int UQ, int QG>
__global__ __launch_bounds__(MT) void k_syev_mid(const EigTask *__restrict__ tasks, double *__restrict__ Abase,
                                                 double *__restrict__ Zbase) {
  constexpr int NM = 64 * RA;
  __shared__ double ppart[MW][NM];
  __shared__ double sv[NM], svp[NM], swp[NM];
  __shared__ double scol[NM];  // column k+1 as the pass of step k stored it
  __shared__ double sd[NM], se[NM], stau[NM];

  const EigTask tk = tasks[blockIdx.x];
  const int n = tk.n;
  if (n <= 0 || n > NM) return;
  double *__restrict__ A = Abase + tk.a_off;
  double *__restrict__ Z = Zbase + tk.z_off;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int i = tid; i < NM; i += MT) {
    svp[i] = 0.0;
    swp[i] = 0.0;
  }
  __syncthreads();

  for (int k = 0; k + 2 < n; ++k) {
    if (wave == 0) {
      // column k with the pending update; rows i = k + lane + 64 q
      const double wpk = swp[k], vpk = svp[k];
      double c[RA];
      double sig = 0.0;
#pragma unroll
      for (int q = 0; q < RA; ++q) {
        const int i = k + lane + 64 * q;
        const double a0 = (i < n) ? (k == 0 ? A[size_t(k) * n + i] : scol[i]) : 0.0;
        c[q] = (i < n) ? a0 - (svp[i] * wpk + swp[i] * vpk) : 0.0;
        if (i > k + 1 && i < n) sig += c[q] * c[q];
      }
      sig = wsum(sig);
      const double dk = __shfl(c[0], 0, 64), x0 = __shfl(c[0], 1, 64);
      double tau = 0.0, alpha = x0, v0 = x0;
      if (sig > 0.0) {
        const double mu = sqrt(x0 * x0 + sig);
        v0 = (x0 <= 0) ? x0 - mu : x0 + mu;
        alpha = (x0 <= 0) ? mu : -mu;
        tau = 2.0 / (v0 * v0 + sig);
      }
#pragma unroll
      for (int q = 0; q < RA; ++q) {
        const int i = k + lane + 64 * q;
        if (i > k && i < n) {
          const double vi = (tau == 0.0) ? 0.0 : ((i == k + 1) ? v0 : c[q]);
          sv[i] = vi;
          A[size_t(k) * n + i] = vi;
        }
      }
      if (lane == 0) {
        sd[k] = dk;
        se[k] = alpha;
        stau[k] = tau;
      }
    }
    __syncthreads();
    const double tau = stau[k];
    const int r0 = k + 1;
    double pacc[RA], vpi[RA], wpi[RA];
#pragma unroll
    for (int a = 0; a < RA; ++a) {
      const int i = r0 + lane + 64 * a;
      pacc[a] = 0.0;
      vpi[a] = (i < n) ? svp[i] : 0.0;
      wpi[a] = (i < n) ? swp[i] : 0.0;
    }
    // groups of UP columns: every load of a group is issued before its
    // stores, so that a thread keeps UP x RA loads in flight
    for (int j0 = r0 + wave; j0 < n; j0 += MW * UP) {
      double x[UP][RA];
#pragma unroll
      for (int u = 0; u < UP; ++u) {
        const int j = j0 + MW * u;
#pragma unroll
        for (int a = 0; a < RA; ++a) {
          const int i = r0 + lane + 64 * a;
          x[u][a] = (j < n && i < n) ? A[size_t(j) * n + i] : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < UP; ++u) {
        const int j = j0 + MW * u;
        if (j < n) {
          const double vj = sv[j], vpj = svp[j], wpj = swp[j];
#pragma unroll
          for (int a = 0; a < RA; ++a) {
            const int i = r0 + lane + 64 * a;
            if (i < n) {
              const double y = x[u][a] - (vpi[a] * wpj + wpi[a] * vpj);
              A[size_t(j) * n + i] = y;
              if (j == r0) scol[i] = y;  // the next step's column, without a round trip
              pacc[a] += y * vj;
            }
          }
        }
      }
    }
#pragma unroll
    for (int a = 0; a < RA; ++a) {
      const int i = r0 + lane + 64 * a;
      if (i < n) ppart[wave][i] = pacc[a];
    }
    __syncthreads();
    // wave 0 alone sums the partials into p, forms K and the next pending
    // update and goes on to the next step's column: everything it reads
    // there it wrote itself, so the step needs no further barrier (the
    // other waves read svp / swp / sv only after the next barrier)
    if (wave == 0) {
      double pv[RA];
      double sm = 0.0;
#pragma unroll
      for (int q = 0; q < RA; ++q) {
        const int i = r0 + lane + 64 * q;
        double ps = 0.0;
        if (i < n) {
#pragma unroll
          for (int w = 0; w < MW; ++w) ps += ppart[w][i];
        }
        pv[q] = tau * ps;
        if (i < n) sm += sv[i] * pv[q];
      }
      const double K = 0.5 * tau * wsum(sm);
#pragma unroll
      for (int q = 0; q < RA; ++q) {
        const int i = r0 + lane + 64 * q;
        if (i < n) {
          swp[i] = pv[q] - K * sv[i];
          svp[i] = sv[i];
        }
      }
    }
  }
  // the last 2 x 2 block with the pending update
  if (tid == 0) {
    if (n == 1) {
      sd[0] = A[0];
      se[0] = 0.0;
    } else {
      const int k = n - 2;
      auto upd = [&](int i, int j) { return A[size_t(j) * n + i] - (svp[i] * swp[j] + swp[i] * svp[j]); };
      sd[k] = upd(k, k);
      se[k] = upd(k + 1, k);
      sd[k + 1] = upd(k + 1, k + 1);
      se[k + 1] = 0.0;
    }
  }
  for (size_t x = tid; x < size_t(n) * n; x += MT) Z[x] = ((x / n) == (x % n)) ? 1.0 : 0.0;
  __syncthreads();

  // Q = H_0 ... H_{n-3}, backwards; wave w owns the columns c = w (mod 16).
  // QG reflectors per pass: a column is loaded once, H_kh, ..., H_kl are
  // applied to it in registers (each reflector masked to its rows >= k+1,
  // so a column it does not reach gets an exact zero dot), stored once.
  for (int kh = n - 3; kh >= 0; kh -= QG) {
    const int kl = max(kh - QG + 1, 0);
    const int r0 = kl + 1;
    double vg[QG][RA], tg[QG];
#pragma unroll
    for (int g = 0; g < QG; ++g) {
      const int k = kh - g;
      tg[g] = (k >= kl) ? stau[k] : 0.0;
#pragma unroll
      for (int a = 0; a < RA; ++a) {
        const int i = r0 + lane + 64 * a;
        vg[g][a] = (k >= kl && i < n && i > k) ? A[size_t(k) * n + i] : 0.0;
      }
    }
    const int cstart = r0 + (((wave - r0) % MW) + MW) % MW;
    for (int c0 = cstart; c0 < n; c0 += MW * UQ) {
      double z[UQ][RA];
#pragma unroll
      for (int u = 0; u < UQ; ++u) {
        const int c = c0 + MW * u;
#pragma unroll
        for (int a = 0; a < RA; ++a) {
          const int i = r0 + lane + 64 * a;
          z[u][a] = (c < n && i < n) ? Z[size_t(c) * n + i] : 0.0;
        }
      }
#pragma unroll
      for (int g = 0; g < QG; ++g) {
        if (tg[g] == 0.0) continue;  // uniform
        double sm[UQ];
#pragma unroll
        for (int u = 0; u < UQ; ++u) {
          sm[u] = 0.0;
#pragma unroll
          for (int a = 0; a < RA; ++a) sm[u] += vg[g][a] * z[u][a];
        }
#pragma unroll
        for (int u = 0; u < UQ; ++u) {
          const double f = tg[g] * wsum(sm[u]);
#pragma unroll
          for (int a = 0; a < RA; ++a) z[u][a] -= f * vg[g][a];
        }
      }
#pragma unroll
      for (int u = 0; u < UQ; ++u) {
        const int c = c0 + MW * u;
        if (c < n) {
#pragma unroll
          for (int a = 0; a < RA; ++a) {
            const int i = r0 + lane + 64 * a;
            if (i < n) Z[size_t(c) * n + i] = z[u][a];
          }
        }
      }
    }
  }
  // (d, e) to the host, which runs the QL recurrence
  double *DE = Zbase + tk.de_off;
  for (int i = tid; i < n; i += MT) {
    DE[i] = sd[i];
    DE[n + i] = se[i];
  }
}

// ---- multi-workgroup reduction for n > 128 (cooperative launch) ---------
// The same Householder reduction and backward accumulation of Q, spread
// over the whole device: the trailing block is cut into 64 x 64 tiles that
// the workgroups walk (persistent loop), and grid-wide barriers separate
// the phases of a step.  Householder step: every workgroup forms the
// reflector itself (same data, same reduction order, hence the same branch
// everywhere); tiles write partial products of p = tau S v per column tile
// [barrier]; row tiles sum them into p and partial dots v.p [barrier];
// every workgroup forms K = tau/2 v.p and updates its tiles
// S -= v w^T + w v^T, w = p - K v [barrier].  Q step: partial column dots
// v^T Z per row tile [barrier]; each tile sums the 64 column dots it needs
// and applies the rank-1 update [barrier].  Every workgroup passes every
// barrier.
constexpr int CT = 256;  // threads per workgroup (4 waves)
constexpr int CB = 64;   // tile edge

__device__ __forceinline__ double csum(double v, double *red) {  // block sum, CT threads
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0;
  for (int i = 0; i < CT / 64; ++i) s += red[i];
  __syncthreads();
  return s;
}

// workspace doubles of k_syev_coop for order n
inline size_t coop_ws(int n) {
  const size_t nt = size_t((n + CB - 1) / CB);
  return nt * size_t(n) + 3 * size_t(n) + nt;
}

// A (n x n, column-major; the reflectors are left in it), Z: Q column-major
// (leading dimension n), DE: (d, e), wk: coop_ws(n) doubles
__global__ __launch_bounds__(CT) void k_syev_coop(double *__restrict__ A, int n, double *__restrict__ Z,
                                                  double *__restrict__ DE, double *__restrict__ wk) {
  namespace cg = cooperative_groups;
  cg::grid_group grid = cg::this_grid();
  __shared__ double red[CT / 64];
  __shared__ double sh[CT];
  __shared__ double scol[CB];
  const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nt = (n + CB - 1) / CB;
  double *part = wk;                     // nt x n partial sums
  double *pvec = wk + size_t(nt) * n;    // n
  double *tauv = pvec + n;               // n
  double *ev = tauv + n;                 // n
  double *dotp = ev + n;                 // nt

  // ---- 1. Householder tridiagonalisation ----
  for (int k = 0; k + 2 < n; ++k) {
    const int L = n - k - 1, o = k + 1;
    const int ntl = (L + CB - 1) / CB;
    double *col = A + size_t(k) * n + o;  // x = A(k+1:, k)
    double sg = 0;
    for (int i = 1 + tid; i < L; i += CT) sg += col[i] * col[i];
    const double sigma = csum(sg, red);
    const double x0 = col[0];
    double t = 0, alpha = x0, v0 = x0;
    if (sigma > 0) {
      const double mu = sqrt(x0 * x0 + sigma);
      v0 = (x0 <= 0) ? x0 - mu : x0 + mu;
      alpha = (x0 <= 0) ? mu : -mu;
      t = 2.0 / (v0 * v0 + sigma);
    }
    if (t == 0) {  // the same on every workgroup
      if (g == 0 && tid == 0) {
        ev[k] = alpha;
        tauv[k] = 0;
      }
      continue;
    }
    // partial products over column tiles: part[tj * n + i]
    for (int tt = g; tt < ntl * ntl; tt += G) {
      const int ti = tt % ntl, tj = tt / ntl;
      const int i = ti * CB + lane;
      double acc = 0;
      if (i < L)
        for (int jj = wv; jj < CB; jj += CT / 64) {
          const int j = tj * CB + jj;
          if (j < L) acc += A[size_t(o + j) * n + o + i] * (j == 0 ? v0 : col[j]);
        }
      sh[tid] = acc;
      __syncthreads();
      if (wv == 0 && i < L) part[size_t(tj) * n + i] = (sh[lane] + sh[64 + lane]) + (sh[128 + lane] + sh[192 + lane]);
      __syncthreads();
    }
    grid.sync();
    // p on row tiles, with the partial dots v.p
    for (int ti = g; ti < ntl; ti += G) {
      double dv = 0;
      if (tid < CB) {
        const int i = ti * CB + tid;
        if (i < L) {
          double pi = 0;
          for (int q = 0; q < ntl; ++q) pi += part[size_t(q) * n + i];
          pi *= t;
          pvec[i] = pi;
          dv = (i == 0 ? v0 : col[i]) * pi;
        }
      }
      dv = csum(dv, red);
      if (tid == 0) dotp[ti] = dv;
    }
    grid.sync();
    double ks = 0;
    for (int q = tid; q < ntl; q += CT) ks += dotp[q];
    const double Kc = 0.5 * t * csum(ks, red);
    // S -= v w^T + w v^T on this workgroup's tiles
    for (int tt = g; tt < ntl * ntl; tt += G) {
      const int ti = tt % ntl, tj = tt / ntl;
      const int i = ti * CB + lane;
      if (i >= L) continue;
      const double vi = (i == 0 ? v0 : col[i]), wi = pvec[i] - Kc * vi;
      for (int jj = wv; jj < CB; jj += CT / 64) {
        const int j = tj * CB + jj;
        if (j >= L) break;
        const double vj = (j == 0 ? v0 : col[j]), wj = pvec[j] - Kc * vj;
        A[size_t(o + j) * n + o + i] -= vi * wj + wi * vj;
      }
    }
    if (g == 0 && tid == 0) {
      ev[k] = alpha;
      tauv[k] = t;
    }
    grid.sync();  // every workgroup is done with col[0] before it becomes v0
    if (g == 0 && tid == 0) col[0] = v0;
  }
  grid.sync();
  for (int i = g * CT + tid; i < n; i += G * CT) {
    DE[i] = A[size_t(i) * n + i];
    DE[n + i] = (i + 2 < n) ? ev[i] : (i == n - 2 ? A[size_t(n - 2) * n + (n - 1)] : 0.0);
  }

  // ---- 2. Q = H_0 ... H_{n-3}, accumulated backwards in Z ----
  for (size_t x = size_t(g) * CT + tid; x < size_t(n) * n; x += size_t(G) * CT)
    Z[x] = ((x / n) == (x % n)) ? 1.0 : 0.0;
  grid.sync();
  for (int k = n - 3; k >= 0; --k) {
    const double t = tauv[k];
    if (t == 0) continue;  // the same on every workgroup
    const int j0 = k + 1, L = n - j0;
    const int ntl = (L + CB - 1) / CB;
    const double *v = A + size_t(k) * n + j0;
    // part[ti * n + c] = sum over row tile ti of v_i Z(j0+i, j0+c)
    for (int tt = g; tt < ntl * ntl; tt += G) {
      const int ti = tt % ntl, tj = tt / ntl;
      const int i = ti * CB + lane;
      const double vi = i < L ? v[i] : 0.0;
      for (int cc = wv; cc < CB; cc += CT / 64) {
        const int c = tj * CB + cc;
        if (c >= L) break;
        double a = i < L ? vi * Z[size_t(j0 + c) * n + j0 + i] : 0.0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
        if (lane == 0) part[size_t(ti) * n + c] = a;
      }
    }
    grid.sync();
    for (int tt = g; tt < ntl * ntl; tt += G) {
      const int ti = tt % ntl, tj = tt / ntl;
      if (tid < CB) {  // the 64 column dots of this tile
        const int c = tj * CB + tid;
        double sc = 0;
        if (c < L)
          for (int q = 0; q < ntl; ++q) sc += part[size_t(q) * n + c];
        scol[tid] = t * sc;
      }
      __syncthreads();
      const int i = ti * CB + lane;
      if (i < L) {
        const double vi = v[i];
        for (int cc = wv; cc < CB; cc += CT / 64) {
          const int c = tj * CB + cc;
          if (c >= L) break;
          Z[size_t(j0 + c) * n + j0 + i] -= scol[cc] * vi;
        }
      }
      __syncthreads();
    }
    grid.sync();
  }
}

// ---- LDS-resident reduction for n <= NMAX (<= 128) -------------------
// The matrix lives in LDS: Householder reduction, then Q formed in place
// from the stored reflectors (backward accumulation, the row and column of
// each new trailing index reset to the unit vector), written out with
// (d, e) for the host.
constexpr int ES = 512;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Householder reduction of the symmetric n x n matrix S (element (r, c) at
// S[c * LD + r], LDS or global memory) to tridiagonal form (d, e), then Q
// formed in place in S.  ES threads; PL = rows per lane (n <= 64 PL).
template <int PL>
__device__ __forceinline__ void reduce_form_q(double *__restrict__ S, int LD, int n, double *d, double *e,
                                              double *tau, double *pv, double *part) {
  constexpr int NMAX = 64 * PL;
  const int tid = threadIdx.x;
  // ---- 1. Householder tridiagonalisation ----
  // Every wave computes the reflector redundantly (lanes over rows, shuffle
  // sums), the product p = tau S v is split over the four waves by column
  // residue mod 4, and the rank-2 update runs waves over columns, lanes
  // over rows: three barriers per step.
  constexpr int NW = ES / 64;
  const int lane = tid & 63, wv = tid >> 6;
  for (int k = 0; k + 2 < n; ++k) {
    const int L = n - k - 1;
    double *col = S + k * LD + (k + 1);  // x = S(k+1:, k)
    double xr[PL];
    double sg = 0;
#pragma unroll
    for (int u = 0; u < PL; ++u) {
      const int i = lane + 64 * u;
      xr[u] = i < L ? col[i] : 0.0;
      if (i >= 1) sg += xr[u] * xr[u];
    }
    const double sigma = wave_sum(sg);
    const double x0 = __shfl(xr[0], 0, 64);
    double t = 0, alpha = x0, v0 = x0;
    if (sigma > 0) {
      const double mu = sqrt(x0 * x0 + sigma);
      v0 = (x0 <= 0) ? x0 - mu : x0 + mu;
      alpha = (x0 <= 0) ? mu : -mu;
      t = 2.0 / (v0 * v0 + sigma);
    }
    if (lane == 0) xr[0] = v0;  // this lane's copy of v
    if (t == 0) {
      if (tid == 0) {
        e[k] = alpha;
        tau[k] = 0;
      }
      continue;  // S unchanged; nothing to order
    }
    // partial products over columns j = wv mod NW
    double acc[PL];
#pragma unroll
    for (int u = 0; u < PL; ++u) acc[u] = 0;
    for (int j = wv; j < L; j += NW) {
      const double vj = (j == 0) ? v0 : col[j];
      const double *Sj = S + (k + 1 + j) * LD + (k + 1);
#pragma unroll
      for (int u = 0; u < PL; ++u) {
        const int i = lane + 64 * u;
        if (i < L) acc[u] += Sj[i] * vj;
      }
    }
#pragma unroll
    for (int u = 0; u < PL; ++u) {
      const int i = lane + 64 * u;
      if (i < L) part[wv * NMAX + i] = acc[u];
    }
    __syncthreads();
    double pr[PL], dv = 0;
#pragma unroll
    for (int u = 0; u < PL; ++u) {
      const int i = lane + 64 * u;
      double pi = 0;
      if (i < L) {
#pragma unroll
        for (int q = 0; q < NW; ++q) pi += part[q * NMAX + i];
      }
      pr[u] = t * pi;
      dv += xr[u] * pr[u];
    }
    const double Kc = 0.5 * t * wave_sum(dv);
#pragma unroll
    for (int u = 0; u < PL; ++u) pr[u] -= Kc * xr[u];  // w = p - K v
    if (wv == 0) {
#pragma unroll
      for (int u = 0; u < PL; ++u) {
        const int i = lane + 64 * u;
        if (i < L) pv[i] = pr[u];
      }
      if (lane == 0) {
        col[0] = v0;  // col now holds v
        e[k] = alpha;
        tau[k] = t;
      }
    }
    __syncthreads();
    // S -= v w^T + w v^T
    for (int j = wv; j < L; j += NW) {
      const double vj = col[j], wj = pv[j];
      double *Sj = S + (k + 1 + j) * LD + (k + 1);
#pragma unroll
      for (int u = 0; u < PL; ++u) {
        const int i = lane + 64 * u;
        if (i < L) Sj[i] -= xr[u] * wj + pr[u] * vj;
      }
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
  // s_c = tau v^T Z(:, c) split over the waves by row residue mod 4 (lanes
  // over columns), then Z -= v s^T with waves over columns, lanes over rows.
  for (int k = n - 3; k >= 0; --k) {
    const int j0 = k + 1;
    for (int i = j0 + tid; i < n; i += ES) {
      S[j0 * LD + i] = (i == j0) ? 1.0 : 0.0;  // column j0 (held v_{k+1})
      S[i * LD + j0] = (i == j0) ? 1.0 : 0.0;  // row j0
    }
    __syncthreads();
    const double t = tau[k];
    if (t == 0) continue;  // block-uniform; the next step's reset is ordered by its own barrier
    const double *v = S + k * LD + j0;  // v_k, first entry at row j0
    const int L = n - j0;
    double acc[PL];
#pragma unroll
    for (int u = 0; u < PL; ++u) acc[u] = 0;
    for (int i = wv; i < L; i += NW) {
      const double vi = v[i];
#pragma unroll
      for (int u = 0; u < PL; ++u) {
        const int c = lane + 64 * u;
        if (c < L) acc[u] += vi * S[(j0 + c) * LD + j0 + i];
      }
    }
#pragma unroll
    for (int u = 0; u < PL; ++u) {
      const int c = lane + 64 * u;
      if (c < L) part[wv * NMAX + c] = acc[u];
    }
    __syncthreads();
    double vr[PL];
#pragma unroll
    for (int u = 0; u < PL; ++u) {
      const int i = lane + 64 * u;
      vr[u] = i < L ? v[i] : 0.0;
    }
    for (int c = wv; c < L; c += NW) {
      double sc = 0;
#pragma unroll
      for (int q = 0; q < NW; ++q) sc += part[q * NMAX + c];
      sc *= t;
      double *Tc = S + (j0 + c) * LD + j0;
#pragma unroll
      for (int u = 0; u < PL; ++u) {
        const int i = lane + 64 * u;
        if (i < L) Tc[i] -= sc * vr[u];
      }
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += ES) {
    S[i] = (i == 0) ? 1.0 : 0.0;       // column 0
    S[i * LD] = (i == 0) ? 1.0 : 0.0;  // row 0
  }
  __syncthreads();

}

template <int NMAX>
__global__ __launch_bounds__(ES) void k_syev_lds(const EigTask *__restrict__ tasks, const double *__restrict__ Abase,
                                                 double *__restrict__ Zbase) {
  constexpr int LD = NMAX + 1;
  __shared__ double S[NMAX * LD];
  __shared__ double d[NMAX], e[NMAX], tau[NMAX], pv[NMAX];
  __shared__ double part[(ES / 64) * NMAX];

  const EigTask tk = tasks[blockIdx.x];
  const int n = tk.n;
  const int tid = threadIdx.x;
  const double *__restrict__ A = Abase + tk.a_off;
  if (n <= 0 || n > NMAX) return;  // the host sorts matrices into size classes
  // staged through registers, eight loads in flight per thread
  for (int x0 = 0; x0 < n * n; x0 += 8 * ES) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int x = x0 + u * ES + tid;
      v[u] = x < n * n ? A[x] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int x = x0 + u * ES + tid;
      if (x < n * n) {
        const int c = x / n, r = x - c * n;
        S[c * LD + r] = v[u];
      }
    }
  }
  __syncthreads();
  reduce_form_q<NMAX / 64>(S, LD, n, d, e, tau, pv, part);

  // ---- 3. Q (column-major) and (d, e) out; the host runs the QL recurrence ----
  double *Z = Zbase + tk.z_off;
  for (int x = tid; x < n * n; x += ES) {
    const int c = x / n, r = x - c * n;
    Z[x] = S[c * LD + r];
  }
  double *DE = Zbase + tk.de_off;
  for (int i = tid; i < n; i += ES) {
    DE[i] = d[i];
    DE[n + i] = e[i];
  }
}

// ---- hybrid form: rotations recorded by the host, applied here ------------
// Every QL sweep is a chain of rotations on consecutive index pairs
// (i, i+1), i = top, top-1, ..., top-cnt+1.  The host sends one stream of
// double pairs per matrix: a header (top, cnt) per sweep followed by that
// sweep's (c, s) pairs in application order.
struct RotMeta {
  int64_t rot_off, ev_off;      // the stream; eigenvalues then ranks
  int64_t a_off, w_off, z_off;  // A, w; Q (column-major, n x (n+1)) in the workspace
  int32_t n, nrot;              // nrot: stream length (headers included)
};
constexpr int AT = 64;    // rows per apply block (one wave, one row per lane)
constexpr int AK = 16;    // rotations per straight-line chunk (the longest)
constexpr int RB = 1024;  // rotations staged in LDS per batch

// Applies the rotation stream to this lane's row of Q (element (row, i) at
// zc[i * ld]).  Within a sweep position i is read once, before the rotation
// on (i, i+1), and the rotation on (i+1, i+2) has left only i+1 changed, so
// a chunk loads its positions together ahead of its rotations; the updated
// element i travels to the next rotation in a register and each rotation
// stores position i+1, which the chain does not touch again.  A sweep runs
// as chunks of 16 rotations, its remainder as chunks of 8, 4, 2 and 1, so
// no chunk reaches past the sweep and none needs a mask.
template <int LDC>  // LDC > 0: Q rows in LDS with that leading dimension (32-bit offsets)
__device__ __forceinline__ void apply_stream(double *zc, int ld, int n, const RotMeta &mt,
                                             const double2 *__restrict__ R, double2 *Rs) {
  auto at = [&](int p) -> double & {
    if constexpr (LDC > 0)
      return zc[p * LDC];
    else
      return zc[size_t(p) * ld];
  };
  int segrem = 0, i = 0;
  double carry = 0;
  for (int b0 = 0; b0 < mt.nrot; b0 += RB) {
    const int qn = min(RB, mt.nrot - b0);
    double2 ld_[RB / AT];  // all loads of the batch in flight at once
#pragma unroll
    for (int u = 0; u < RB / AT; ++u) {
      const int x = u * AT + threadIdx.x;
      ld_[u] = x < qn ? R[b0 + x] : make_double2(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < RB / AT; ++u) Rs[u * AT + threadIdx.x] = ld_[u];
    __syncthreads();
    int q = 0;
    // C consecutive rotations of the current sweep, positions i .. i-C+1
    // (all inside the sweep, so no clamp or mask): every load ahead of the
    // stores, offsets from the chunk base compile-time constants
    auto chunk = [&](auto cc) {
      constexpr int C = decltype(cc)::value;
      double *zb = &at(i - (C - 1));  // position i - t is zb[(C - 1 - t) * step]
      double z[C];
      double2 g[C];
#pragma unroll
      for (int t = 0; t < C; ++t) {
        if constexpr (LDC > 0)
          z[t] = zb[(C - 1 - t) * LDC];
        else
          z[t] = zb[size_t(C - 1 - t) * ld];
        g[t] = Rs[q + t];
      }
#pragma unroll
      for (int t = 0; t < C; ++t) {
        const double cz = g[t].x * z[t];
        const double up = fma(g[t].y, z[t], g[t].x * carry);
        carry = fma(-g[t].y, carry, cz);
        if constexpr (LDC > 0)
          zb[(C - t) * LDC] = up;  // position i - t + 1
        else
          zb[size_t(C - t) * ld] = up;
      }
      q += C;
      segrem -= C;
      i -= C;
    };
    while (q < qn) {
      if (segrem == 0) {  // a sweep header
        const double2 h = Rs[q++];
        i = int(h.x);
        segrem = int(h.y);
        carry = at(i + 1);
        continue;  // the header may end the batch
      }
      // the sweep's rotations in this batch, in chunks of 16, then 8, 4, 2, 1
      const int avail = min(segrem, qn - q);
      if (avail >= AK)
        chunk(std::integral_constant<int, AK>{});
      else if (avail >= 8)
        chunk(std::integral_constant<int, 8>{});
      else if (avail >= 4)
        chunk(std::integral_constant<int, 4>{});
      else if (avail >= 2)
        chunk(std::integral_constant<int, 2>{});
      else
        chunk(std::integral_constant<int, 1>{});
      if (segrem == 0) at(i + 1) = carry;
    }
  }
}

// Q <- Q G_1 G_2 ..., then this block's rows of the sorted eigenvectors
// (column rank[c] of A is column c of Q) and, from the first block of a
// matrix, the sorted eigenvalues.  LC > 0: the block's 64 rows of Q are
// staged in LDS (n <= LC); LC == 0: they are updated in the workspace.
template <int LC>
__global__ __launch_bounds__(AT) void k_syev_apply(const RotMeta *__restrict__ meta,
                                                   const double2 *__restrict__ rots, const double *__restrict__ evr, double *__restrict__ Abase,
                                                   double *__restrict__ Wbase, double *__restrict__ Zbase) {
  __shared__ double2 Rs[RB + AK];
  __shared__ double Zs[LC > 0 ? (LC + 1) * AT : 1];
  const RotMeta mt = meta[blockIdx.x];
  const int n = mt.n;
  const int r0 = blockIdx.y * AT;
  if (r0 >= n) return;
  const int lane = threadIdx.x;
  const int r = min(r0 + lane, n - 1);  // lanes past the last row redo row n-1 (same values)
  const bool act = r0 + lane < n;
  const double *Zg = Zbase + mt.z_off + r;
  const double2 *R = rots + mt.rot_off;
  double *zc;
  int ld;
  if constexpr (LC > 0) {
    zc = Zs + lane;
    ld = AT;
    for (int c0 = 0; c0 < n; c0 += 16) {  // sixteen loads in flight
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) v[u] = c0 + u < n ? Zg[size_t(c0 + u) * n] : 0.0;
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (c0 + u < n) zc[(c0 + u) * AT] = v[u];
    }
  } else {
    zc = Zbase + mt.z_off + r;
    ld = n;
  }
  apply_stream<(LC > 0 ? AT : 0)>(zc, ld, n, mt, R, Rs);
  // ranks to LDS (the rotation buffer is free now), then the sorted output
  const double *ev = evr + mt.ev_off, *rk = ev + n;
  int *rks = reinterpret_cast<int *>(Rs);
  __syncthreads();
  for (int c = lane; c < n; c += AT) {
    const int k = int(rk[c]);
    rks[c] = k;
    if (blockIdx.y == 0) Wbase[mt.w_off + k] = ev[c];
  }
  __syncthreads();
  if (!act) return;
  double *__restrict__ A = Abase + mt.a_off + r;
  for (int c = 0; c < n; ++c) A[size_t(rks[c]) * n] = zc[size_t(c) * ld];
}

// sqrt(a^2 + b^2) without overflow or underflow of the squares
inline double pyth(double a, double b) {
  a = std::fabs(a);
  b = std::fabs(b);
  const double w = std::max(a, b), v = std::min(a, b);
  if (w == 0) return 0;
  const double q = v / w;
  return w * std::sqrt(1.0 + q * q);
}

// Implicit QL with the tql2 deflation test on the host (the recurrence is
// a serial chain of a few dependent flops per rotation, which one CPU core
// runs an order of magnitude faster than one GPU lane).  Appends a header
// (top, cnt) and the (c, s) pairs of every sweep to the stream; d ends up
// holding the eigenvalues.
bool host_ql(int n, double *d, double *e, std::vector<double2> &rots) {
  double tst1 = 0;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    tst1 = std::max(tst1, std::fabs(d[l]) + std::fabs(e[l]));
    while (true) {
      int m = l;
      for (; m < n - 1; ++m)
        if (std::fabs(e[m]) <= DBL_EPSILON * std::max(std::fabs(d[m]) + std::fabs(d[m + 1]), tst1)) break;
      if (m == l) break;
      if (iter++ >= QL_MAXIT) return false;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = pyth(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
      double s = 1, c = 1, p = 0;
      bool early = false;
      const size_t hdr = rots.size();
      rots.push_back(make_double2(0.0, 0.0));
      for (int i = m - 1; i >= l; --i) {
        const double f = s * e[i], b = c * e[i];
        r = pyth(f, g);
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
        rots.push_back(make_double2(c, s));
      }
      const size_t cnt = rots.size() - hdr - 1;
      if (cnt)
        rots[hdr] = make_double2(double(m - 1), double(cnt));
      else
        rots.pop_back();
      if (!early) {
        d[l] -= p;
        e[l] = g;
        e[m] = 0;
      }
    }
  }
  return true;
}

// A fixed set of host threads for independent per-matrix work (the QL
// recurrences of a batch): run(n, f) calls f(i) for i < n, on the calling
// thread and the workers, and returns when all are done.  Created on first
// use and kept for the life of the process (the workers wait on a
// condition variable between jobs).
class HostPool {
 public:
  static HostPool &get() {
    static HostPool *pool = new HostPool();  // never destroyed: workers park until the process ends
    return *pool;
  }
  int threads() const { return int(workers_.size()) + 1; }
  template <class F>
  void run(int n, F &&f) {
    if (n <= 0) return;
    std::unique_lock<std::mutex> job(job_mu_);  // one job at a time
    std::function<void(int)> fn(std::forward<F>(f));
    {
      std::lock_guard<std::mutex> lk(mu_);
      fn_ = &fn;
      count_ = n;
      next_.store(0);
      active_ = int(workers_.size());
      ++gen_;
    }
    cv_.notify_all();
    drain();
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return active_ == 0; });
    fn_ = nullptr;
  }

 private:
  HostPool() {
    int nt = int(std::thread::hardware_concurrency());
    if (const char *o = std::getenv("OMP_NUM_THREADS")) nt = std::atoi(o);
    nt = std::max(1, std::min(nt, 16));
    for (int i = 1; i < nt; ++i) workers_.emplace_back([this] { loop(); });
    for (auto &w : workers_) w.detach();
  }
  void drain() {
    for (int i = next_.fetch_add(1); i < count_; i = next_.fetch_add(1)) (*fn_)(i);
  }
  void loop() {
    uint64_t seen = 0;
    while (true) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
      }
      drain();
      std::lock_guard<std::mutex> lk(mu_);
      if (--active_ == 0) done_cv_.notify_one();
    }
  }
  std::vector<std::thread> workers_;
  std::mutex job_mu_, mu_;
  std::condition_variable cv_, done_cv_;
  std::function<void(int)> *fn_ = nullptr;
  int count_ = 0, active_ = 0;
  std::atomic<int> next_{0};
  uint64_t gen_ = 0;
};

}  // namespace

// Our kernels for every matrix of the call (n <= EMAX).  `between` runs once
// the tridiagonal forms are on the host, before the QL recurrences: work it
// enqueues runs on the device while the host iterates.
int syev_own(itx_ctx *ctx, int nmat, const int32_t *ns, const int64_t *aoffs, double *Abase, const int64_t *woffs,
             double *Wbase, double *whost, const std::function<int()> &between) {
  if (nmat <= 0) return between();
  // The kernels reduce every matrix to tridiagonal form and build Q; the
  // host runs the QL recurrences; k_syev_apply applies the recorded
  // rotations and writes the sorted eigenpairs.  Reduction size classes:
  // LDS-resident up to 64 and up to 128, one workgroup per matrix over
  // global memory up to 256 and up to 512, else one cooperative launch per
  // matrix over the whole device.
  // Workspace: every (d, e) first (one copy to the host), then every Q.
  std::vector<EigTask> t64, t128, t256, t512, tg;
  int64_t desize = 0, qsize = 0;
  for (int b = 0; b < nmat; ++b) {
    if (ns[b] < 0 || ns[b] > EMAX) return set_err(ctx, ITX_ERR_UNSUPPORTED, "itx_syev: n above 2048");
    if (ns[b] == 0) continue;
    EigTask t{};
    t.n = ns[b];
    t.a_off = aoffs[b];
    t.w_off = woffs[b];
    t.de_off = desize;
    desize += 2 * int64_t(ns[b]);
    t.z_off = qsize;
    qsize += int64_t(ns[b]) * (ns[b] + 1);
    if (ns[b] <= 64)
      t64.push_back(t);
    else if (ns[b] <= 128)
      t128.push_back(t);
    else if (ns[b] <= 256)
      t256.push_back(t);
    else if (ns[b] <= 512)
      t512.push_back(t);
    else
      tg.push_back(t);
  }
  std::vector<EigTask> all;
  for (auto *v : {&t64, &t128, &t256, &t512, &tg}) all.insert(all.end(), v->begin(), v->end());
  const int ntask = int(all.size());
  if (ntask == 0) return between();
  for (auto &t : all) t.z_off += desize;
  const int64_t zsize = desize + qsize;
  const size_t tbytes = sizeof(EigTask) * size_t(ntask);
  const size_t zoff = (tbytes + 255) & ~size_t(255);
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
  double *Z = reinterpret_cast<double *>(wsb + zoff);
  {
    const EigTask *q = dt;
    if (!t64.empty()) hipLaunchKernelGGL(k_syev_lds<64>, dim3(unsigned(t64.size())), dim3(ES), 0, ctx->stream, q, Abase, Z);
    q += t64.size();
    if (!t128.empty())
      hipLaunchKernelGGL(k_syev_lds<128>, dim3(unsigned(t128.size())), dim3(ES), 0, ctx->stream, q, Abase, Z);
    q += t128.size();
    if (!t256.empty())
      hipLaunchKernelGGL((k_syev_mid<4, 4, 2, 8>), dim3(unsigned(t256.size())), dim3(MT), 0, ctx->stream, q, Abase, Z);
    q += t256.size();
    if (!t512.empty())
      hipLaunchKernelGGL((k_syev_mid<8, 1, 2, 4>), dim3(unsigned(t512.size())), dim3(MT), 0, ctx->stream, q, Abase, Z);
    q += t512.size();
    // above 512: side by side, one workgroup each (time ~ the largest
    // n^3), or one after the other, each a cooperative launch over the
    // whole device (time ~ the sum of n^2) -- whichever the model of the
    // measured costs (n = 256: 14 ms / 8.4 ms) says is shorter
    double t_side = 0, t_coop = 0;
    for (const EigTask &t : tg) {
      const double r = double(t.n) / 256.0;
      t_side = std::max(t_side, 14.0 * r * r * r);
      t_coop += 8.4 * r * r;
    }
    if (!tg.empty() && t_side < t_coop) {
      hipLaunchKernelGGL(k_syev, dim3(unsigned(tg.size())), dim3(ET), 0, ctx->stream, q, Abase, Wbase, Z);
      tg.clear();
    }
    static int coop_blocks = 0;
    if (!tg.empty() && coop_blocks == 0) {
      int occ = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void *>(k_syev_coop), CT, 0) !=
              hipSuccess ||
          occ < 1)
        occ = 1;
      coop_blocks = std::max(1, ctx->num_cus * std::min(occ, 2));
    }
    for (size_t b = all.size() - tg.size(); b < all.size(); ++b) {  // (offsets as uploaded)
      const EigTask &t = all[b];
      const int n = t.n, nt = (n + CB - 1) / CB;
      double *wk = nullptr;
      int rc2 = pool_alloc(ctx, sizeof(double) * coop_ws(n), reinterpret_cast<void **>(&wk));
      if (rc2) {
        pool_free(ctx, ws);
        return rc2;
      }
      double *Ap = Abase + t.a_off, *Zp = Z + t.z_off, *DEp = Z + t.de_off;
      int nn = n;
      void *args[] = {&Ap, &nn, &Zp, &DEp, &wk};
      const unsigned G = unsigned(std::min<long>(coop_blocks, long(nt) * nt));
      hipError_t e2 = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k_syev_coop), dim3(G), dim3(CT), args,
                                                 0, ctx->stream);
      pool_free(ctx, wk);  // stream-ordered
      if (e2 != hipSuccess) {
        pool_free(ctx, ws);
        return hip_err(ctx, e2, "itx_syev (cooperative reduction)");
      }
    }
  }
  hipError_t e = hipGetLastError();
  // the tridiagonals of every matrix to the host (one wait)
  std::vector<double> de(static_cast<size_t>(desize));
  if (e == hipSuccess)
    e = hipMemcpyAsync(de.data(), Z, sizeof(double) * size_t(desize), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    pool_free(ctx, ws);
    return hip_err(ctx, e, "itx_syev");
  }
  if (int rb = between()) {
    pool_free(ctx, ws);
    return rb;
  }
  // The QL recurrences, one matrix per task on the host threads, largest
  // first; every task writes its own rotations, sweep records, eigenvalues
  // and descending ranks, concatenated in matrix order afterwards.
  struct QlOut {
    std::vector<double2> rots;  // the rotation stream (a header per sweep)
    std::vector<double> evr;    // eigenvalues (QL order), then ranks
    bool ok = true;
  };
  std::vector<QlOut> outs(static_cast<size_t>(ntask));
  std::vector<int> order(static_cast<size_t>(ntask));
  double work = 0;
  for (int b = 0; b < ntask; ++b) {
    order[size_t(b)] = b;
    work += double(all[size_t(b)].n) * all[size_t(b)].n;
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return all[size_t(a)].n > all[size_t(b)].n; });
  auto ql_task = [&](int j) {
    const int b = order[size_t(j)];
    const EigTask &t = all[size_t(b)];
    const int n = t.n;
    double *d = de.data() + t.de_off, *ee = d + n;
    QlOut &o = outs[size_t(b)];
    o.ok = host_ql(n, d, ee, o.rots);
    if (!o.ok) return;
    // stable descending rank: larger first, equal values in QL order
    std::vector<int> idx(static_cast<size_t>(n));
    for (int c = 0; c < n; ++c) idx[size_t(c)] = c;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int c) { return d[a] > d[c]; });
    o.evr.assign(d, d + n);
    o.evr.resize(2 * size_t(n));
    for (int k = 0; k < n; ++k) {
      o.evr[size_t(n + idx[size_t(k)])] = double(k);
      if (whost) whost[t.w_off + k] = d[idx[size_t(k)]];  // the values the apply writes to w
    }
  };
  // threads pay off from a few milliseconds of recurrence (about 64^2 x 8)
  if (ntask > 1 && work > 32768.0)
    HostPool::get().run(ntask, ql_task);
  else
    for (int j = 0; j < ntask; ++j) ql_task(j);
  // apply classes: Q rows staged in LDS for n <= 64, 128, 256; else global
  std::vector<RotMeta> cls[4];
  int maxn[4] = {0, 0, 0, 0};
  size_t nrot = 0, nev = 0;
  for (int b = 0; b < ntask; ++b) {
    const QlOut &o = outs[size_t(b)];
    if (!o.ok) {
      pool_free(ctx, ws);
      return set_err(ctx, ITX_ERR_NOCONV, "itx_syev: QL iteration did not converge");
    }
    const EigTask &t = all[size_t(b)];
    RotMeta mt{};
    mt.rot_off = int64_t(nrot);
    mt.nrot = int32_t(o.rots.size());
    mt.ev_off = int64_t(nev);
    mt.a_off = t.a_off;
    mt.w_off = t.w_off;
    mt.z_off = t.z_off;
    mt.n = t.n;
    nrot += o.rots.size();
    nev += o.evr.size();
    const int k = t.n <= 64 ? 0 : t.n <= 128 ? 1 : t.n <= 256 ? 2 : 3;
    cls[k].push_back(mt);
    maxn[k] = std::max(maxn[k], t.n);
  }
  std::vector<RotMeta> meta;
  for (auto &c : cls) meta.insert(meta.end(), c.begin(), c.end());
  const size_t mb = sizeof(RotMeta) * meta.size(), rb = sizeof(double2) * nrot, eb = sizeof(double) * nev;
  const size_t o_r = (mb + 255) & ~size_t(255), o_e = (o_r + rb + 255) & ~size_t(255);
  std::vector<char> host(o_e + eb);
  std::memcpy(host.data(), meta.data(), mb);
  {
    size_t c = 0, v = 0;
    for (const QlOut &o : outs) {
      std::memcpy(host.data() + o_r + c, o.rots.data(), sizeof(double2) * o.rots.size());
      std::memcpy(host.data() + o_e + v, o.evr.data(), sizeof(double) * o.evr.size());
      c += sizeof(double2) * o.rots.size();
      v += sizeof(double) * o.evr.size();
    }
  }
  void *ub = nullptr;
  rc = pool_alloc(ctx, host.size(), &ub);
  if (rc) {
    pool_free(ctx, ws);
    return rc;
  }
  char *u = static_cast<char *>(ub);
  rc = stage_upload(ctx, u, host.data(), host.size());
  if (!rc) {
    const RotMeta *dm = reinterpret_cast<const RotMeta *>(u);
    const double2 *dr = reinterpret_cast<const double2 *>(u + o_r);
    const double *dv = reinterpret_cast<const double *>(u + o_e);
    using Kfn = void (*)(const RotMeta *, const double2 *, const double *, double *, double *, double *);
    const Kfn fns[4] = {k_syev_apply<64>, k_syev_apply<128>, k_syev_apply<256>, k_syev_apply<0>};
    size_t first = 0;
    for (int k = 0; k < 4; ++k) {
      const size_t cnt = cls[k].size();
      if (cnt)
        hipLaunchKernelGGL(fns[k], dim3(unsigned(cnt), unsigned((maxn[k] + AT - 1) / AT)), dim3(AT), 0, ctx->stream,
                           dm + first, dr, dv, Abase, Wbase, Z);
      first += cnt;
    }
    e = hipGetLastError();
    if (e != hipSuccess) rc = hip_err(ctx, e, "itx_syev apply");
  }
  // stream-ordered frees: reused only by later work on the stream
  pool_free(ctx, ub);
  pool_free(ctx, ws);
  return rc;
}

// ---- blocks above kOwnMax rows: rocSOLVER's divide-and-conquer dsyevd ----
// ROCm's LAPACK in place of the reference's dsyev on -M: each block is
// negated into a slot of its size class (N x N, column-major), whose
// padding is c I with c = 1 + 2 n max|a_ij|, above every eigenvalue of -A; the
// tridiagonal form then splits exactly at the block edge, so the first n
// ascending eigenpairs of the slot are those of -A: A's eigenvalues in
// descending order (equal ones in LAPACK's order, as in the reference),
// their eigenvectors zero outside the block.  dsyevd is bound by the
// latency of its serial phases -- measured on MI355X: about 2 ms + 20 us
// per row for one matrix, and a batch of 8 costs little more than one --
// so every block of a call goes into as few strided-batched calls as
// possible: one, padded to the largest block, unless the slots would pass
// 1 GiB.  When a call holds a block above kOwnMax rows, the blocks above
// kJoinMin rows join that call (their own-kernel time would add to it);
// otherwise every block stays on our kernels.
constexpr int kOwnMax = 256;
constexpr int kJoinMin = 128;
constexpr int EMAX_LARGE = 8192;

struct PadTask {
  int64_t a_off, p_off, w_off, d_off;  // A and w of the block; its slot and eigenvalues in the workspace
  int32_t n, N;
};

// Kernels over (block, chunk of PCOL columns): max |a_ij| of each block
// (an integer maximum of the non-negative bit patterns: order-independent),
// then the negated block in its slot with c = 1 + 2 n max|a_ij| (above
// every |eigenvalue|) on the padding's diagonal, then the results back.
constexpr int PCOL = 16;

__global__ __launch_bounds__(256) void k_block_absmax(const PadTask *__restrict__ tasks,
                                                      const double *__restrict__ Abase,
                                                      unsigned long long *__restrict__ maxbits) {
  const PadTask t = tasks[blockIdx.x];
  const int c0 = blockIdx.y * PCOL;
  if (c0 >= t.n) return;
  const double *__restrict__ A = Abase + t.a_off;
  double m = 0;
  for (int c = c0; c < min(t.n, c0 + PCOL); ++c)
    for (int r = threadIdx.x; r < t.n; r += 256) m = fmax(m, fabs(A[r + int64_t(c) * t.n]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(&maxbits[blockIdx.x], (unsigned long long)__double_as_longlong(m));
}

__global__ __launch_bounds__(256) void k_pad_neg(const PadTask *__restrict__ tasks, const double *__restrict__ Abase,
                                                 const unsigned long long *__restrict__ maxbits,
                                                 double *__restrict__ Pbase) {
  const PadTask t = tasks[blockIdx.x];
  const int c0 = blockIdx.y * PCOL;
  if (c0 >= t.N) return;
  const double c = 1.0 + 2.0 * double(t.n) * __longlong_as_double((long long)maxbits[blockIdx.x]);
  const double *__restrict__ A = Abase + t.a_off;
  double *__restrict__ P = Pbase + t.p_off;
  for (int col = c0; col < min(t.N, c0 + PCOL); ++col)
    for (int r = threadIdx.x; r < t.N; r += 256)
      P[r + int64_t(col) * t.N] = (r < t.n && col < t.n) ? -A[r + int64_t(col) * t.n] : (r == col ? c : 0.0);
}

__global__ __launch_bounds__(256) void k_unpad(const PadTask *__restrict__ tasks, const double *__restrict__ Pbase,
                                               const double *__restrict__ Dbase, double *__restrict__ Abase,
                                               double *__restrict__ Wbase) {
  const PadTask t = tasks[blockIdx.x];
  const int c0 = blockIdx.y * PCOL;
  if (c0 >= t.n) return;
  const double *__restrict__ P = Pbase + t.p_off;
  double *__restrict__ A = Abase + t.a_off;
  for (int col = c0; col < min(t.n, c0 + PCOL); ++col)
    for (int r = threadIdx.x; r < t.n; r += 256) A[r + int64_t(col) * t.n] = P[r + int64_t(col) * t.N];
  if (blockIdx.y == 0)
    for (int k = threadIdx.x; k < t.n; k += 256) Wbase[t.w_off + k] = -Dbase[t.d_off + k];
}

struct LargeCall {
  void *ws = nullptr;
  int *info = nullptr;
  int count = 0;
  std::vector<PadTask> tasks;
};

rocblas_handle solver_handle(itx_ctx *ctx) {
  if (!ctx->rocblas) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
    ctx->rocblas = h;
  }
  rocblas_handle h = static_cast<rocblas_handle>(ctx->rocblas);
  rocblas_set_stream(h, ctx->stream);
  return h;
}

int syev_large_launch(itx_ctx *ctx, std::vector<PadTask> tasks, double *Abase, double *Wbase, LargeCall &lc) {
  if (tasks.empty()) return ITX_OK;
  rocblas_handle h = solver_handle(ctx);
  if (!h) return set_err(ctx, ITX_ERR_HIP, "itx_syev: rocblas_create_handle failed");
  // size classes, largest first
  std::stable_sort(tasks.begin(), tasks.end(), [](const PadTask &a, const PadTask &b) { return a.n > b.n; });
  struct Cls {
    int N, first, count;
  };
  std::vector<Cls> cls;
  int64_t pbytes = 0, dsz = 0;
  for (int i = 0; i < int(tasks.size());) {
    const int N = (tasks[size_t(i)].n + 31) & ~31;
    const int per_call = int(std::max<int64_t>(1, (int64_t(1) << 27) / (int64_t(N) * N)));  // 1 GiB of slots
    int j = i;
    while (j < int(tasks.size()) && j - i < per_call) {
      tasks[size_t(j)].N = N;
      tasks[size_t(j)].p_off = pbytes + int64_t(j - i) * N * N;
      tasks[size_t(j)].d_off = dsz + int64_t(j - i) * N;
      ++j;
    }
    cls.push_back({N, i, j - i});
    pbytes += int64_t(j - i) * N * N;
    dsz += int64_t(j - i) * N;
    i = j;
  }
  // workspace: tasks | slots | eigenvalues | off-diagonals | info
  const size_t tb = sizeof(PadTask) * tasks.size();
  const size_t o_p = (tb + 255) & ~size_t(255);
  const size_t o_d = o_p + sizeof(double) * size_t(pbytes);
  const size_t o_e = o_d + sizeof(double) * size_t(dsz);
  const size_t o_i = o_e + sizeof(double) * size_t(dsz);
  const size_t o_m = (o_i + sizeof(int) * tasks.size() + 255) & ~size_t(255);
  int rc = pool_alloc(ctx, o_m + sizeof(double) * tasks.size(), &lc.ws);
  if (rc) return rc;
  char *w = static_cast<char *>(lc.ws);
  rc = stage_upload(ctx, w, tasks.data(), tb);
  if (rc) return rc;
  const PadTask *dt = reinterpret_cast<const PadTask *>(w);
  double *P = reinterpret_cast<double *>(w + o_p), *D = reinterpret_cast<double *>(w + o_d),
         *E = reinterpret_cast<double *>(w + o_e);
  lc.info = reinterpret_cast<int *>(w + o_i);
  lc.count = int(tasks.size());
  unsigned long long *maxbits = reinterpret_cast<unsigned long long *>(w + o_m);
  const unsigned nchunk = unsigned((tasks.front().N + PCOL - 1) / PCOL);  // the largest slot
  rc = itx_fill(ctx, reinterpret_cast<double *>(maxbits), tasks.size(), 0.0);  // bit pattern 0
  if (rc) return rc;
  hipLaunchKernelGGL(k_block_absmax, dim3(unsigned(tasks.size()), nchunk), dim3(256), 0, ctx->stream, dt, Abase,
                     maxbits);
  hipLaunchKernelGGL(k_pad_neg, dim3(unsigned(tasks.size()), nchunk), dim3(256), 0, ctx->stream, dt, Abase, maxbits,
                     P);
  for (const Cls &c : cls) {
    const PadTask &t0 = tasks[size_t(c.first)];
    const rocblas_status st = rocsolver_dsyevd_strided_batched(
        h, rocblas_evect_original, rocblas_fill_lower, c.N, P + t0.p_off, c.N, rocblas_stride(c.N) * c.N,
        D + t0.d_off, c.N, E + t0.d_off, c.N, lc.info + c.first, c.count);
    if (st != rocblas_status_success) return set_err(ctx, ITX_ERR_HIP, "itx_syev: rocsolver_dsyevd failed");
  }
  hipLaunchKernelGGL(k_unpad, dim3(unsigned(tasks.size()), nchunk), dim3(256), 0, ctx->stream, dt, P, D, Abase, Wbase);
  lc.tasks = std::move(tasks);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? ITX_OK : hip_err(ctx, e, "itx_syev (large blocks)");
}

// waits for the large blocks: convergence flags, and their eigenvalues to
// the host when asked for
int syev_large_finish(itx_ctx *ctx, LargeCall &lc, const double *Wbase, double *whost) {
  if (!lc.count) return ITX_OK;
  std::vector<int> info(size_t(lc.count));
  hipError_t e = hipMemcpyAsync(info.data(), lc.info, sizeof(int) * info.size(), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess && whost)
    for (const PadTask &t : lc.tasks) {
      e = hipMemcpyAsync(whost + t.w_off, Wbase + t.w_off, sizeof(double) * size_t(t.n), hipMemcpyDeviceToHost,
                         ctx->stream);
      if (e != hipSuccess) break;
    }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  pool_free(ctx, lc.ws);
  lc.ws = nullptr;
  if (e != hipSuccess) return hip_err(ctx, e, "itx_syev (large blocks)");
  for (int x : info)
    if (x != 0) return set_err(ctx, ITX_ERR_NOCONV, "itx_syev: dsyevd did not converge");
  return ITX_OK;
}

// Every matrix of a call: small blocks (all of them with ITX_EIG_OWN=1, up
// to EMAX) through our kernels, the others through rocSOLVER, enqueued
// while the host runs the QL recurrences of the small ones.
int syev_batched(itx_ctx *ctx, int nmat, const int32_t *ns, const int64_t *aoffs, double *Abase,
                 const int64_t *woffs, double *Wbase, double *whost) {
  if (nmat <= 0) return ITX_OK;
  static const bool own_only = [] {
    const char *v = std::getenv("ITX_EIG_OWN");
    return v && std::atoi(v) != 0;
  }();
  std::vector<int32_t> on;
  std::vector<int64_t> oa, ow;
  std::vector<PadTask> large;
  bool any_large = false;
  for (int b = 0; b < nmat; ++b) any_large = any_large || ns[b] > kOwnMax;
  const int own_max = own_only ? EMAX : any_large ? kJoinMin : kOwnMax;
  for (int b = 0; b < nmat; ++b) {
    const int n = ns[b];
    if (n < 0 || n > (own_only ? EMAX : EMAX_LARGE))
      return set_err(ctx, ITX_ERR_UNSUPPORTED,
                     own_only ? "itx_syev: n above 2048 (ITX_EIG_OWN)" : "itx_syev: n above 8192");
    if (n == 0) continue;
    if (n <= own_max) {
      on.push_back(n);
      oa.push_back(aoffs[b]);
      ow.push_back(woffs[b]);
    } else {
      PadTask t{};
      t.a_off = aoffs[b];
      t.w_off = woffs[b];
      t.n = n;
      large.push_back(t);
    }
  }
  LargeCall lc;
  int rc = syev_own(ctx, int(on.size()), on.data(), oa.data(), Abase, ow.data(), Wbase, whost,
                    [&] { return syev_large_launch(ctx, std::move(large), Abase, Wbase, lc); });
  const int rc2 = syev_large_finish(ctx, lc, Wbase, whost);
  return rc ? rc : rc2;
}

void eig_release(itx_ctx *ctx) {
  if (ctx->rocblas) rocblas_destroy_handle(static_cast<rocblas_handle>(ctx->rocblas));
  ctx->rocblas = nullptr;
}

}  // namespace itx

extern "C" int itx_syev(itx_ctx *ctx, int n, double *A, double *w) {
  if (!ctx || n < 0) return ITX_ERR_ARG;
  if (n == 0) return ITX_OK;
  const int32_t ns[1] = {n};
  const int64_t ao[1] = {0}, wo[1] = {0};
  return itx::syev_batched(ctx, 1, ns, ao, A, wo, w, nullptr);
}

extern "C" int itx_syev_batched(itx_ctx *ctx, int nmat, const int32_t *n, const int64_t *a_offs, double *A,
                                const int64_t *w_offs, double *w) {
  if (!ctx || nmat < 0) return ITX_ERR_ARG;
  if (nmat == 0) return ITX_OK;
  return itx::syev_batched(ctx, nmat, n, a_offs, A, w_offs, w, nullptr);
}

extern "C" int itx_syev_batched_host(itx_ctx *ctx, int nmat, const int32_t *n, const int64_t *a_offs, double *A,
                                     const int64_t *w_offs, double *w, double *w_host) {
  if (!ctx || nmat < 0 || !w_host) return ITX_ERR_ARG;
  if (nmat == 0) return ITX_OK;
  return itx::syev_batched(ctx, nmat, n, a_offs, A, w_offs, w, w_host);
}
