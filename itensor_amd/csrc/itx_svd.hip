// FP64 singular value decomposition of libitx (one workgroup per matrix).
//
// Replaces the SVD(M,U,D,V) call of svdImpl (svd.cc:87,206-213, through
// SVDRef, tensor/algs.cc:586-631, i.e. dgesdd or the density-matrix method
// of SVDRefImpl): A = U diag(S) V^T with S descending, U and V with
// orthonormal columns, for every block of a QN tensor in one launch.
//
// Method (textbook, not LAPACK's code): one-sided Jacobi (Hestenes, J. SIAM
// 6 (1958) 51).  The k = min(m,n) columns of X (X = A^T when m <= n, else
// A; columns of length max(m,n)) are rotated pairwise until every pair is
// orthogonal to a relative tolerance; the rotations accumulate in J.  Then
// the singular values are the column norms of X, the normalised columns
// of X are the singular vectors of the long side and J holds those of the
// short side.  One-sided Jacobi keeps the small singular values and their
// vectors accurate without the refinement step the density-matrix method
// needs.  Pairs are scheduled round-robin (k/2 disjoint pairs per round,
// k-1 rounds per sweep), each pair reduced by a group of lanes.
#include <cfloat>
#include <cstring>

#include "itx_internal.h"

namespace itx {
namespace {

constexpr int ST = 256;             // threads per matrix
constexpr int SVD_MAX_SWEEPS = 60;  // sweeps before reporting no convergence

struct SvdTask {
  int32_t m, n;
  int64_t a_off, u_off, s_off, v_off;
  int64_t x_off;  // workspace: X (C x R), J (R x R), sigma (R), rank (R)
};

__device__ __forceinline__ double group_sum(double v, int G) {
  for (int o = G >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(ST) void k_gesvd(const SvdTask *__restrict__ tasks, const double *__restrict__ Abase,
                                              double *__restrict__ Ubase, double *__restrict__ Sbase,
                                              double *__restrict__ Vbase, double *__restrict__ Wbase,
                                              int *__restrict__ status) {
  __shared__ int sh_rot;
  const SvdTask tk = tasks[blockIdx.x];
  const int m = tk.m, n = tk.n;
  const int R = min(m, n), C = max(m, n);
  const int tid = threadIdx.x;
  if (R <= 0) {
    if (tid == 0) status[blockIdx.x] = 0;
    return;
  }
  const double *__restrict__ A = Abase + tk.a_off;
  double *__restrict__ X = Wbase + tk.x_off;
  double *__restrict__ J = X + size_t(C) * R;
  double *__restrict__ sig = J + size_t(R) * R;
  double *__restrict__ rnk = sig + R;
  const bool wide = m <= n;  // X = A^T

  for (size_t x = tid; x < size_t(C) * R; x += ST) {
    const size_t c = x % C, r = x / C;
    X[x] = wide ? A[r + size_t(m) * c] : A[c + size_t(m) * r];
  }
  for (size_t x = tid; x < size_t(R) * R; x += ST) J[x] = (x % R == x / R) ? 1.0 : 0.0;
  if (tid == 0) sh_rot = 0;
  __syncthreads();

  // round-robin schedule over Rp (even) slots; slot R is a dummy when R is odd
  const int Rp = R + (R & 1), P = Rp / 2;
  int G = 64;
  while (G > 1 && G * P > ST) G >>= 1;
  const int NG = ST / G, grp = tid / G, g = tid % G;
  const double tol = 4.0 * DBL_EPSILON * sqrt(double(C));

  int sweep = 0;
  for (; sweep < SVD_MAX_SWEEPS; ++sweep) {
    for (int k = 0; k < Rp - 1; ++k) {
      for (int p = grp; p < P; p += NG) {
        auto pos = [&](int x) { return x == 0 ? 0 : 1 + (x - 1 + k) % (Rp - 1); };
        int a = pos(p), b = pos(Rp - 1 - p);
        if (a >= R || b >= R) continue;  // group-uniform
        if (a > b) {
          const int t = a;
          a = b;
          b = t;
        }
        double *xa = X + size_t(a) * C, *xb = X + size_t(b) * C;
        double al = 0, be = 0, ga = 0;
        for (int c = g; c < C; c += G) {
          const double u = xa[c], v = xb[c];
          al = fma(u, u, al);
          be = fma(v, v, be);
          ga = fma(u, v, ga);
        }
        al = group_sum(al, G);
        be = group_sum(be, G);
        ga = group_sum(ga, G);
        if (ga == 0.0 || al == 0.0 || be == 0.0 || fabs(ga) <= tol * sqrt(al) * sqrt(be)) continue;
        const double zeta = (be - al) / (2.0 * ga);
        const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(fma(zeta, zeta, 1.0)));
        const double cs = 1.0 / sqrt(fma(t, t, 1.0)), sn = cs * t;
        for (int c = g; c < C; c += G) {
          const double u = xa[c], v = xb[c];
          xa[c] = cs * u - sn * v;
          xb[c] = sn * u + cs * v;
        }
        double *ja = J + size_t(a) * R, *jb = J + size_t(b) * R;
        for (int r = g; r < R; r += G) {
          const double u = ja[r], v = jb[r];
          ja[r] = cs * u - sn * v;
          jb[r] = sn * u + cs * v;
        }
        if (g == 0) sh_rot = 1;
      }
      __syncthreads();
    }
    const int rot = sh_rot;
    __syncthreads();  // every thread has read the flag before it is reset
    if (!rot) break;
    if (tid == 0) sh_rot = 0;
    __syncthreads();
  }
  if (sweep == SVD_MAX_SWEEPS) {
    if (tid == 0) status[blockIdx.x] = 1;
    return;
  }

  // singular values (column norms) and their descending stable ranks
  for (int r = grp; r < R; r += NG) {
    const double *xr = X + size_t(r) * C;
    double s = 0;
    for (int c = g; c < C; c += G) s = fma(xr[c], xr[c], s);
    s = group_sum(s, G);
    if (g == 0) sig[r] = sqrt(s);
  }
  __syncthreads();
  for (int r = tid; r < R; r += ST) {
    const double sr = sig[r];
    int rank = 0;
    for (int j = 0; j < R; ++j) rank += (sig[j] > sr) || (sig[j] == sr && j < r);
    rnk[r] = double(rank);
  }
  __syncthreads();
  double *__restrict__ U = Ubase + tk.u_off;
  double *__restrict__ V = Vbase + tk.v_off;
  double *__restrict__ S = Sbase + tk.s_off;
  double *__restrict__ Long = wide ? V : U;    // C x R: normalised columns of X
  double *__restrict__ Short = wide ? U : V;   // R x R: J
  for (int r = tid; r < R; r += ST) S[int(rnk[r])] = sig[r];
  for (size_t x = tid; x < size_t(C) * R; x += ST) {
    const size_t c = x % C, r = x / C;
    const double sr = sig[r];
    Long[size_t(rnk[r]) * C + c] = sr > 0.0 ? X[x] / sr : 0.0;
  }
  for (size_t x = tid; x < size_t(R) * R; x += ST) {
    const size_t i = x % R, r = x / R;
    Short[size_t(rnk[r]) * R + i] = J[x];
  }
  if (tid == 0) status[blockIdx.x] = 0;
}

}  // namespace
}  // namespace itx

extern "C" int itx_gesvd_batched(itx_ctx *ctx, int nmat, const int32_t *m, const int32_t *n, const int64_t *a_offs,
                                 const double *A, const int64_t *u_offs, double *U, const int64_t *s_offs, double *S,
                                 const int64_t *v_offs, double *V) {
  if (!ctx || nmat < 0 || (nmat > 0 && (!m || !n || !a_offs || !u_offs || !s_offs || !v_offs))) return ITX_ERR_ARG;
  if (nmat == 0) return ITX_OK;
  std::vector<itx::SvdTask> tasks(static_cast<size_t>(nmat));
  int64_t wsize = 0;
  for (int b = 0; b < nmat; ++b) {
    if (m[b] < 0 || n[b] < 0) return itx::set_err(ctx, ITX_ERR_ARG, "itx_gesvd: negative dimension");
    const int64_t R = std::min(m[b], n[b]), C = std::max(m[b], n[b]);
    auto &t = tasks[size_t(b)];
    t.m = m[b];
    t.n = n[b];
    t.a_off = a_offs[b];
    t.u_off = u_offs[b];
    t.s_off = s_offs[b];
    t.v_off = v_offs[b];
    t.x_off = wsize;
    wsize += C * R + R * R + 2 * R;
  }
  const size_t tbytes = sizeof(itx::SvdTask) * size_t(nmat), sbytes = sizeof(int) * size_t(nmat);
  const size_t woff = (tbytes + sbytes + 255) & ~size_t(255);
  void *ws = nullptr;
  int rc = itx::pool_alloc(ctx, woff + sizeof(double) * size_t(std::max<int64_t>(wsize, 1)), &ws);
  if (rc) return rc;
  char *wsb = static_cast<char *>(ws);
  rc = itx::stage_upload(ctx, wsb, tasks.data(), tbytes);
  if (rc) {
    itx::pool_free(ctx, ws);
    return rc;
  }
  int *dst = reinterpret_cast<int *>(wsb + tbytes);
  hipLaunchKernelGGL(itx::k_gesvd, dim3(unsigned(nmat)), dim3(itx::ST), 0, ctx->stream,
                     reinterpret_cast<const itx::SvdTask *>(wsb), A, U, S, V, reinterpret_cast<double *>(wsb + woff),
                     dst);
  hipError_t e = hipGetLastError();
  std::vector<int> st(static_cast<size_t>(nmat), 0);
  if (e == hipSuccess) e = hipMemcpyAsync(st.data(), dst, sbytes, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  itx::pool_free(ctx, ws);
  if (e != hipSuccess) return itx::hip_err(ctx, e, "itx_gesvd");
  for (int b = 0; b < nmat; ++b)
    if (st[size_t(b)] != 0) return itx::set_err(ctx, ITX_ERR_NOCONV, "itx_gesvd: Jacobi sweeps did not converge");
  return ITX_OK;
}

extern "C" int itx_gesvd(itx_ctx *ctx, int m, int n, const double *A, double *U, double *S, double *V) {
  if (!ctx || m < 0 || n < 0) return ITX_ERR_ARG;
  const int32_t mm[1] = {m}, nn[1] = {n};
  const int64_t z[1] = {0};
  return itx_gesvd_batched(ctx, 1, mm, nn, z, A, z, U, z, S, z, V);
}
