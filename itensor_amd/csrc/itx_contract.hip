// Contraction engine of libitx: dense and block-sparse (QN) tensor
// contractions as one persistent MFMA GEMM kernel with the index
// permutations folded into offset-table driven tile loads and stores.
//
// Reference behaviour replaced:
//   contract() / CProps (tensor/contract.cc:137-866) reduces a binary
//   contraction to permute(A), permute(B) -> dgemm -> permute(C).  Here the
//   permutations never touch memory: the plan records, for every logical
//   row m, column n and contracted position k of the GEMM, the element
//   offset inside each operand (tables aM,aK,bK,bN,cM,cN), so the kernel
//   gathers tiles straight from the original layouts.
//   doTask(Contract,QDense,QDense) (itdata/qdense.cc:671-743) runs one
//   contract() per (A block, B block) pair; here all pairs of one
//   contraction are tasks of a single launch, grouped by C block so that
//   all pairs feeding one C tile are summed in registers (the reference's
//   OpenMP path groups by output block for the same reason,
//   itdata/qutil.h:271-349).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <numeric>
#include <queue>
#include <sstream>
#include <type_traits>

#include <unordered_map>
#include "itx_internal.h"

namespace itx {
namespace {

constexpr int BK = 16;   // K depth of one LDS stage (default)

template <typename T>
struct MF;
template <>
struct MF<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // swizzle of LDS columns per K row (keeps 8-byte accesses conflict free)
  static __device__ __forceinline__ int sw(int k) { return (k >> 1) & 7; }
  // output row of accumulator element r held by `lane` (measured on gfx950:
  // the f64 16x16x4 result interleaves rows, row = lane/16 + 4r)
  static __device__ __forceinline__ int orow(int lane, int r) { return (lane >> 4) + 4 * r; }
  template <int BMX>
  static constexpr int pad() { return 16; }
};
template <>
struct MF<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int sw(int k) { return (k >> 2) & 3; }
  // f32 16x16x4 result: blocked rows, row = 4*(lane/16) + r
  static __device__ __forceinline__ int orow(int lane, int r) { return 4 * (lane >> 4) + r; }
  template <int BMX>
  static constexpr int pad() { return BMX == 32 ? 48 : 16; }
};

// Loader of one operand tile (rows x BK) into LDS laid out [k][row].
// KC: the operand is contiguous along k (lanes walk k), else along rows.
// The k-offset table entries of a stage are fetched one stage ahead
// (fetch_k) so that the data loads of the next stage never wait on a table
// read, and every load is issued unconditionally from a clamped address
// (out-of-range elements read element 0 of the block and are zeroed), which
// keeps the load stream free of exec-mask branches.
// Offsets stay 64-bit: 32-bit offsets measured 1.6x slower on gfx950.
// With V = 2 (k-contiguous operands whose k runs come in aligned pairs,
// checked when the plan is built) a lane loads two consecutive k of one
// row as one 16-byte load and fetches one table entry per pair.
template <typename T, int BR, bool KC, int NTH, int BK = 16, int V = 1>
struct TileLoader {
  static constexpr int PER = BR * BK / NTH;  // elements per thread
  static_assert(PER >= 1, "tile too small for the thread count");
  static_assert(V == 1 || (KC && PER % V == 0 && BK % V == 0), "vector loads need a k-contiguous operand");
  static constexpr int NV = PER / V;         // loads per thread
  static constexpr int LK = BK / V;          // lanes along k (KC)
  static constexpr int LD = BR + MF<T>::template pad<BR>();
  static constexpr int NK = KC ? 1 : PER;  // k offsets per thread and stage
  double r[PER];
  bool rok[NV];  // element in range (applied at store, after the MFMAs)
  int64_t roff[KC ? NV : 1];
  bool rv[KC ? NV : 1];
  int64_t ko[NK];
  bool kv[NK];
  int lr, lk;  // thread's base row / base k

  __device__ __forceinline__ void init(int tid, int row0, int R, const int64_t *__restrict__ tabR) {
    if constexpr (KC) {
      lk = (tid % LK) * V;
      lr = tid / LK;
#pragma unroll
      for (int p = 0; p < NV; ++p) {
        int row = row0 + lr + (NTH / LK) * p;
        rv[p] = row < R;
        roff[p] = tabR[rv[p] ? row : 0];
      }
    } else {
      lr = tid % BR;
      lk = tid / BR;
      int row = row0 + lr;
      rv[0] = row < R;
      roff[0] = tabR[rv[0] ? row : 0];
    }
  }
  // table reads for the stage starting at k0
  __device__ __forceinline__ void fetch_k(int k0, int K, const int64_t *__restrict__ tabK) {
#pragma unroll
    for (int q = 0; q < NK; ++q) {
      const int k = k0 + lk + (KC ? 0 : (NTH / BR) * q);
      kv[q] = k < K;
      ko[q] = tabK[kv[q] ? k : 0];
    }
  }
  // data loads of the stage whose offsets fetch_k last read
  __device__ __forceinline__ void load(const double *__restrict__ base) {
#pragma unroll
    for (int p = 0; p < NV; ++p) {
      const bool ok = KC ? (kv[0] && rv[p]) : (kv[p] && rv[0]);
      const int64_t off = (KC ? roff[p] + ko[0] : roff[0] + ko[p]);
      if constexpr (V == 2) {
        const double2 v = *reinterpret_cast<const double2 *>(base + (ok ? off : 0));
        r[2 * p] = v.x;
        r[2 * p + 1] = v.y;
      } else {
        r[p] = base[ok ? off : 0];
      }
      rok[p] = ok;
    }
  }
  __device__ __forceinline__ void store(T *__restrict__ s) const {
    if constexpr (KC) {
#pragma unroll
      for (int p = 0; p < NV; ++p) {
        int row = lr + (NTH / LK) * p;
#pragma unroll
        for (int e = 0; e < V; ++e) {
          const int k = lk + e;
          s[k * LD + (row ^ MF<T>::sw(k))] = rok[p] ? T(r[V * p + e]) : T(0);
        }
      }
    } else {
#pragma unroll
      for (int p = 0; p < PER; ++p) {
        int k = lk + (NTH / BR) * p;
        s[k * LD + (lr ^ MF<T>::sw(k))] = rok[p] ? T(r[p]) : T(0);
      }
    }
  }
};

// One workgroup = GM x GN waves computing a BM x BN tile of C; each wave
// owns a (BM/GM) x (BN/GN) sub-tile of 16x16 MFMA fragments.
template <typename T, int BM, int BN, int GM, int GN, bool AKC, bool BKC, int BK, int V>
__global__ __launch_bounds__(64 * GM * GN) void k_gemm(const double *__restrict__ A, const double *__restrict__ B,
                                                       double *__restrict__ C, const GemmTask *__restrict__ tasks,
                                                       const WorkUnit *__restrict__ units, int nunits,
                                                       const int64_t *__restrict__ tab, double alpha, double beta) {
  constexpr int NTH = 64 * GM * GN;
  using v4 = typename MF<T>::v4;
  using LA = TileLoader<T, BM, AKC, NTH, BK, V>;
  using LB = TileLoader<T, BN, BKC, NTH, BK, V>;
  constexpr int WM = BM / GM, WN = BN / GN, FM = WM / 16, FN = WN / 16;
  static_assert(FM >= 1 && FN >= 1, "wave tile below one MFMA fragment");
  __shared__ T As[2][BK * LA::LD];
  __shared__ T Bs[2][BK * LB::LD];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % GM, wn = wave / GM;

  // Units are spread so that each XCD (blocks b with equal b % 8 land on
  // the same XCD) walks a contiguous chunk of the unit list: neighbouring
  // tiles share A rows / B columns, which then stay in that XCD's L2.
  const int G = gridDim.x, b = blockIdx.x;
  int ustart, uend, ustep;
  if (G % 8 == 0 && nunits >= G) {
    const int nb = G / 8, x = b % 8, lb = b / 8;
    const int chunk = (nunits + 7) / 8;
    ustart = x * chunk + lb;
    uend = min(nunits, (x + 1) * chunk);
    ustep = nb;
  } else {
    ustart = b;
    uend = nunits;
    ustep = G;
  }

  for (int u = ustart; u < uend; u += ustep) {
    const WorkUnit wu = units[u];
    v4 acc[FN][FM];
#pragma unroll
    for (int i = 0; i < FN; ++i)
#pragma unroll
      for (int j = 0; j < FM; ++j) acc[i][j] = v4{0, 0, 0, 0};

    for (int t = wu.tbeg; t < wu.tend; ++t) {
      const GemmTask tk = tasks[t];
      if (tk.K <= 0) continue;
      LA la;
      LB lb;
      la.init(tid, wu.m0, tk.M, tab + tk.tAM);
      lb.init(tid, wu.n0, tk.N, tab + tk.tBN);
      const double *Ab = A + tk.a_off;
      const double *Bb = B + tk.b_off;
      const int64_t *tAK = tab + tk.tAK;
      const int64_t *tBK = tab + tk.tBK;
      la.fetch_k(0, tk.K, tAK);
      lb.fetch_k(0, tk.K, tBK);
      la.load(Ab);
      lb.load(Bb);
      la.fetch_k(BK, tk.K, tAK);
      lb.fetch_k(BK, tk.K, tBK);
      la.store(As[0]);
      lb.store(Bs[0]);
      __syncthreads();
      int cur = 0;
      for (int k0 = 0; k0 < tk.K; k0 += BK) {
        const bool nxt = k0 + BK < tk.K;
        if (nxt) {
          la.load(Ab);  // stage k0+BK (offsets fetched one stage earlier)
          lb.load(Bb);
          la.fetch_k(k0 + 2 * BK, tk.K, tAK);
          lb.fetch_k(k0 + 2 * BK, tk.K, tBK);
        }
        const T *as = As[cur];
        const T *bs = Bs[cur];
#pragma unroll
        for (int kq = 0; kq < BK / 4; ++kq) {
          const int kr = kq * 4 + (lane >> 4);
          T af[FM], bf[FN];
#pragma unroll
          for (int j = 0; j < FM; ++j) {
            int m = wm * WM + j * 16 + (lane & 15);
            af[j] = as[kr * LA::LD + (m ^ MF<T>::sw(kr))];
          }
#pragma unroll
          for (int i = 0; i < FN; ++i) {
            int n = wn * WN + i * 16 + (lane & 15);
            bf[i] = bs[kr * LB::LD + (n ^ MF<T>::sw(kr))];
          }
          // D^T tile: MFMA rows run over n, columns over m, so that the
          // 16 lanes of a row hold 16 consecutive m (coalesced C stores).
#pragma unroll
          for (int i = 0; i < FN; ++i)
#pragma unroll
            for (int j = 0; j < FM; ++j) acc[i][j] = MF<T>::mma(bf[i], af[j], acc[i][j]);
        }
        if (nxt) {
          la.store(As[cur ^ 1]);
          lb.store(Bs[cur ^ 1]);
        }
        __syncthreads();
        cur ^= 1;
      }
    }

    // epilogue: C = alpha*acc + beta*C through the C offset tables
    const GemmTask t0 = tasks[wu.tbeg];
    const int64_t *tCM = tab + t0.tCM;
    const int64_t *tCN = tab + t0.tCN;
    double *Cb = C + t0.c_off;
#pragma unroll
    for (int j = 0; j < FM; ++j) {
      const int m = wu.m0 + wm * WM + j * 16 + (lane & 15);
      if (m >= t0.M) continue;
      const int64_t cm = tCM[m];
#pragma unroll
      for (int i = 0; i < FN; ++i) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = wu.n0 + wn * WN + i * 16 + MF<T>::orow(lane, r);
          if (n < t0.N) {
            const int64_t ad = cm + tCN[n];
            double v = alpha * double(acc[i][j][r]);
            if (beta != 0.0) v += beta * Cb[ad];
            Cb[ad] = v;
          }
        }
      }
    }
  }
}

// ---- FP64 register-pipelined tile ----------------------------------------
// 128 x 128 C tile, four waves of 64 x 64 (4 x 4 fragments of the f64
// 16x16x4 MFMA), 16-deep k stages, for operands contiguous along k in
// aligned pairs (the plan's vector-load condition).  A thread loads 16-byte
// k pairs; an LDS row holds the 16 k of a stage contiguously (pitch 18
// doubles: the 16 rows that a quarter-wave reads land on distinct 16-byte
// bank groups), and a lane reads the 4 consecutive k of its fragment row
// with two 16-byte reads.  MFMA s (0..3) of a stage takes k = 4q + s from
// the lanes of quarter q, so every k of the stage enters once.
// Pipeline (one barrier per stage): while the 64 MFMAs of stage n run from
// registers, the operands of stage n+1 are read from LDS, stage n+2 is
// stored to LDS from the load registers, and the global loads of stage n+4
// are issued into those registers (two register sets: loads have two
// stages, about 8000 MFMA cycles, to arrive).
// FP32 form: the same pipeline with operands rounded to FP32 at the LDS
// store, LDS rows of 20 floats (80 bytes: again 16 distinct bank groups),
// one 16-byte read per fragment row and the f32 16x16x4 MFMA.
constexpr int DG_BK = 16, DG_NT = 256;
template <typename T>
struct DgCfg;
template <>
struct DgCfg<double> {
  static constexpr int P = 18;   // LDS row pitch (elements)
  static constexpr int RD = 16;  // LDS reads per stage and lane
};
template <>
struct DgCfg<float> {
  static constexpr int P = 20;
  static constexpr int RD = 8;
};

struct DgRegs {  // one stage of one operand in flight: four 16-byte pairs
  double g[8];
  bool ok[4];
};

struct DgLoader {  // one operand: 128 rows x 16 k per stage, four 16-byte loads per thread
  int64_t roff[4];
  bool rv[4];
  int64_t ko[2];  // k offsets of the next two stages to load (slot = stage & 1)
  bool kv[2];
  int lr, lc;  // base row (32 rows apart per load), k pair
  __device__ __forceinline__ void init(int tid, int row0, int R, const int64_t *__restrict__ tabR) {
    lc = tid & 7;
    lr = tid >> 3;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int row = row0 + lr + 32 * p;
      rv[p] = row < R;
      roff[p] = tabR[rv[p] ? row : 0];
    }
  }
  __device__ __forceinline__ void fetch_k(int k0, int K, const int64_t *__restrict__ tabK, int slot) {
    const int k = k0 + 2 * lc;  // K is even: the pair is in range with its first k
    kv[slot] = k < K;
    ko[slot] = tabK[kv[slot] ? k : 0];
  }
  __device__ __forceinline__ void load(const double *__restrict__ base, DgRegs &r, int slot) const {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      r.ok[p] = kv[slot] && rv[p];
      const double2 v = *reinterpret_cast<const double2 *>(base + (r.ok[p] ? roff[p] + ko[slot] : 0));
      r.g[2 * p] = v.x;
      r.g[2 * p + 1] = v.y;
    }
  }
  __device__ __forceinline__ void store(double *__restrict__ s, const DgRegs &r) const {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      double2 v;
      v.x = r.ok[p] ? r.g[2 * p] : 0.0;
      v.y = r.ok[p] ? r.g[2 * p + 1] : 0.0;
      *reinterpret_cast<double2 *>(s + (lr + 32 * p) * DgCfg<double>::P + 2 * lc) = v;
    }
  }
  __device__ __forceinline__ void store(float *__restrict__ s, const DgRegs &r) const {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      float2 v;
      v.x = r.ok[p] ? float(r.g[2 * p]) : 0.0f;
      v.y = r.ok[p] ? float(r.g[2 * p + 1]) : 0.0f;
      *reinterpret_cast<float2 *>(s + (lr + 32 * p) * DgCfg<float>::P + 2 * lc) = v;
    }
  }
};

template <typename T>
struct DgFrag;
template <>
struct DgFrag<double> {  // a lane's operands of one stage: 4 fragments x 4 k, for A and for B
  double2 a[4][2], b[4][2];
};
template <>
struct DgFrag<float> {
  float4 a[4], b[4];
};

__device__ __forceinline__ void dg_read(DgFrag<double> &f, const double *__restrict__ as,
                                        const double *__restrict__ bs, int wm, int wn, int lane) {
  const int r = lane & 15, q = lane >> 4;
  constexpr int P = DgCfg<double>::P;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const double *pa = as + (wm * 64 + j * 16 + r) * P + 4 * q;
    f.a[j][0] = *reinterpret_cast<const double2 *>(pa);
    f.a[j][1] = *reinterpret_cast<const double2 *>(pa + 2);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double *pb = bs + (wn * 64 + i * 16 + r) * P + 4 * q;
    f.b[i][0] = *reinterpret_cast<const double2 *>(pb);
    f.b[i][1] = *reinterpret_cast<const double2 *>(pb + 2);
  }
}

__device__ __forceinline__ void dg_read(DgFrag<float> &f, const float *__restrict__ as, const float *__restrict__ bs,
                                        int wm, int wn, int lane) {
  const int r = lane & 15, q = lane >> 4;
  constexpr int P = DgCfg<float>::P;
#pragma unroll
  for (int j = 0; j < 4; ++j) f.a[j] = *reinterpret_cast<const float4 *>(as + (wm * 64 + j * 16 + r) * P + 4 * q);
#pragma unroll
  for (int i = 0; i < 4; ++i) f.b[i] = *reinterpret_cast<const float4 *>(bs + (wn * 64 + i * 16 + r) * P + 4 * q);
}

using dv4 = MF<double>::v4;

__device__ __forceinline__ double dg_k(const double2 (&x)[2], int s) {
  return s == 0 ? x[0].x : s == 1 ? x[0].y : s == 2 ? x[1].x : x[1].y;
}
__device__ __forceinline__ float dg_k(const float4 &x, int s) {
  return s == 0 ? x.x : s == 1 ? x.y : s == 2 ? x.z : x.w;
}

// the 16 MFMAs of k slot s (k = 4q + s in quarter q)
template <typename T>
__device__ __forceinline__ void dg_mfma(typename MF<T>::v4 (&acc)[4][4], const DgFrag<T> &f, int s) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = MF<T>::mma(dg_k(f.b[i], s), dg_k(f.a[j], s), acc[i][j]);
}

template <typename T>
__global__ __launch_bounds__(DG_NT) void k_dgemm(const double *__restrict__ A, const double *__restrict__ B,
                                                 double *__restrict__ C, const GemmTask *__restrict__ tasks,
                                                 const WorkUnit *__restrict__ units, int nunits,
                                                 const int64_t *__restrict__ tab, double alpha, double beta) {
  using v4 = typename MF<T>::v4;
  constexpr int P = DgCfg<T>::P, RD = DgCfg<T>::RD;
  __shared__ __align__(16) T As[2][128 * P];
  __shared__ __align__(16) T Bs[2][128 * P];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  // XCD-contiguous unit chunks (as in k_gemm)
  const int G = gridDim.x, b = blockIdx.x;
  int ustart, uend, ustep;
  if (G % 8 == 0 && nunits >= G) {
    const int nb = G / 8, x = b % 8, lb = b / 8;
    const int chunk = (nunits + 7) / 8;
    ustart = x * chunk + lb;
    uend = min(nunits, (x + 1) * chunk);
    ustep = nb;
  } else {
    ustart = b;
    uend = nunits;
    ustep = G;
  }

  for (int u = ustart; u < uend; u += ustep) {
    const WorkUnit wu = units[u];
    v4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = v4{0, 0, 0, 0};

    for (int t = wu.tbeg; t < wu.tend; ++t) {
      const GemmTask tk = tasks[t];
      if (tk.K <= 0) continue;
      DgLoader la, lb;
      la.init(tid, wu.m0, tk.M, tab + tk.tAM);
      lb.init(tid, wu.n0, tk.N, tab + tk.tBN);
      const double *Ab = A + tk.a_off;
      const double *Bb = B + tk.b_off;
      const int64_t *tAK = tab + tk.tAK;
      const int64_t *tBK = tab + tk.tBK;
      const int S = (tk.K + DG_BK - 1) / DG_BK;
      DgRegs ra0, rb0, ra1, rb1;  // load register sets
// load stage st (its k offsets fetched two loads ago, slot st & 1) and
// fetch the k offsets of stage st + 2 into the same slot: a table read then
// never waits behind the data loads just issued
#define ITX_DG_LOAD(st, ra, rb, slot)                      \
  {                                                        \
    la.load(Ab, ra, slot);                                 \
    lb.load(Bb, rb, slot);                                 \
    la.fetch_k(((st) + 2) * DG_BK, tk.K, tAK, slot);       \
    lb.fetch_k(((st) + 2) * DG_BK, tk.K, tBK, slot);       \
  }
      // prologue: stages 0 and 1 into LDS, stage 0 into registers, stage 2
      // in set 1 and stage 3 in set 0
      la.fetch_k(0, tk.K, tAK, 0);
      lb.fetch_k(0, tk.K, tBK, 0);
      la.fetch_k(DG_BK, tk.K, tAK, 1);
      lb.fetch_k(DG_BK, tk.K, tBK, 1);
      ITX_DG_LOAD(0, ra0, rb0, 0)
      __syncthreads();  // the previous task's LDS reads are done
      la.store(As[0], ra0);
      lb.store(Bs[0], rb0);
      if (S > 1) ITX_DG_LOAD(1, ra0, rb0, 1)
      if (S > 2) ITX_DG_LOAD(2, ra1, rb1, 0)
      __syncthreads();
      DgFrag<T> f0, f1;
      dg_read(f0, As[0], Bs[0], wm, wn, lane);
      if (S > 1) {
        la.store(As[1], ra0);
        lb.store(Bs[1], rb0);
      }
      if (S > 3) ITX_DG_LOAD(3, ra0, rb0, 1)
      __syncthreads();
      // stage n: its MFMAs from registers; stage n+1 read from LDS[(n+1)&1]
      // at the start; after the first 16 MFMAs, stage n+2 stored from its
      // register set into LDS[n&1] (read into registers one stage ago) and
      // the loads of stage n+4 issued into that set.  The scheduling
      // barriers keep the store and the loads inside the MFMA stream and the
      // workgroup barrier behind the last MFMA (the stores have long
      // completed when the waves meet there).
#define ITX_DG_STEP(n, cur, nxt, ra, rb, slot)                                        \
  {                                                                               \
    if ((n) + 1 < S) dg_read(nxt, As[((n) + 1) & 1], Bs[((n) + 1) & 1], wm, wn, lane); \
    dg_mfma(acc, cur, 0);                                                         \
    __builtin_amdgcn_sched_barrier(0);                                            \
    if ((n) + 2 < S) {                                                            \
      la.store(As[(n) & 1], ra);                                                  \
      lb.store(Bs[(n) & 1], rb);                                                  \
    }                                                                             \
    if ((n) + 4 < S) ITX_DG_LOAD((n) + 4, ra, rb, slot)                           \
    __builtin_amdgcn_sched_barrier(0);                                            \
    dg_mfma(acc, cur, 1);                                                         \
    dg_mfma(acc, cur, 2);                                                         \
    dg_mfma(acc, cur, 3);                                                         \
    __builtin_amdgcn_sched_barrier(0);                                            \
    __syncthreads();                                                              \
  }
      // Steady state (stages n with n+4 < S): the same work without
      // conditions, in one scheduling region, with the LDS reads, the LDS
      // stores (and their selects) and the global loads (and their address
      // arithmetic) placed between consecutive MFMAs, so that the MFMA pipe
      // never waits on the wave's other instructions.
#define ITX_DG_BODY(n, cur, nxt, ra, rb, id)                                      \
  {                                                                               \
    dg_read(nxt, As[((n) + 1) & 1], Bs[((n) + 1) & 1], wm, wn, lane);            \
    dg_mfma(acc, cur, 0);                                                         \
    for (int g_ = 0; g_ < RD; ++g_) {                                             \
      __builtin_amdgcn_sched_group_barrier(0x008, 16 / RD, id);                   \
      __builtin_amdgcn_sched_group_barrier(0x100, 1, id);                         \
    }                                                                             \
    __builtin_amdgcn_sched_barrier(0);                                            \
    la.store(As[(n) & 1], ra);                                                    \
    lb.store(Bs[(n) & 1], rb);                                                    \
    dg_mfma(acc, cur, 1);                                                         \
    for (int g_ = 0; g_ < 8; ++g_) {                                              \
      __builtin_amdgcn_sched_group_barrier(0x008, 1, id);                         \
      __builtin_amdgcn_sched_group_barrier(0x006, 4, id);                         \
      __builtin_amdgcn_sched_group_barrier(0x200, 1, id);                         \
    }                                                                             \
    __builtin_amdgcn_sched_group_barrier(0x008, 8, id);                           \
    __builtin_amdgcn_sched_barrier(0);                                            \
    ITX_DG_LOAD((n) + 4, ra, rb, id)                                              \
    dg_mfma(acc, cur, 2);                                                         \
    for (int g_ = 0; g_ < 12; ++g_) {                                             \
      __builtin_amdgcn_sched_group_barrier(0x008, 1, id);                         \
      __builtin_amdgcn_sched_group_barrier(0x006, 6, id);                         \
      __builtin_amdgcn_sched_group_barrier(0x020, 1, id);                         \
    }                                                                             \
    __builtin_amdgcn_sched_group_barrier(0x008, 4, id);                           \
    __builtin_amdgcn_sched_barrier(0);                                            \
    dg_mfma(acc, cur, 3);                                                         \
    __builtin_amdgcn_sched_barrier(0);                                            \
    __syncthreads();                                                              \
  }
      int n = 0;
      for (; n + 5 < S; n += 2) {
        ITX_DG_BODY(n, f0, f1, ra1, rb1, 0)
        ITX_DG_BODY(n + 1, f1, f0, ra0, rb0, 1)
      }
      for (; n + 1 < S; n += 2) {
        ITX_DG_STEP(n, f0, f1, ra1, rb1, 0)
        ITX_DG_STEP(n + 1, f1, f0, ra0, rb0, 1)
      }
      if (n < S) ITX_DG_STEP(n, f0, f1, ra1, rb1, 0)
#undef ITX_DG_BODY
#undef ITX_DG_STEP
#undef ITX_DG_LOAD
    }

    // epilogue: C = alpha*acc + beta*C through the C offset tables (the
    // table entries of all of the lane's rows and columns are read before
    // the first store, from clamped indices)
    const GemmTask t0 = tasks[wu.tbeg];
    const int64_t *tCM = tab + t0.tCM;
    const int64_t *tCN = tab + t0.tCN;
    double *Cb = C + t0.c_off;
    int64_t cm[4], cn[4][4];
    bool mv[4], nv[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int m = wu.m0 + wm * 64 + j * 16 + (lane & 15);
      mv[j] = m < t0.M;
      cm[j] = tCM[mv[j] ? m : 0];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = wu.n0 + wn * 64 + i * 16 + MF<T>::orow(lane, r);
        nv[i][r] = n < t0.N;
        cn[i][r] = tCN[nv[i][r] ? n : 0];
      }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (mv[j] && nv[i][r]) {
            const int64_t ad = cm[j] + cn[i][r];
            double v = alpha * double(acc[i][j][r]);
            if (beta != 0.0) v += beta * Cb[ad];
            Cb[ad] = v;
          }
        }
  }
}

// ---- FP64 tile, two workgroups per CU ------------------------------------
// The tile, LDS layout and k order of k_dgemm (results bit for bit equal),
// with the register budget of two waves per SIMD (128 accumulator + at most
// 128 other registers): the operands are read from LDS one k slot ahead
// (16 MFMAs) instead of one stage ahead, and one set of load registers
// holds the next stage.  Two resident workgroups per CU cover each other's
// barrier waits, epilogues and first loads.  One barrier per stage, before
// the last slot, which reads the next stage's first slot.
template <typename T>
struct DgSlot {  // a lane's operands of one k slot: 4 A and 4 B fragments
  T a[4], b[4];
};

template <typename T>
__device__ __forceinline__ void dg_read_slot(DgSlot<T> &f, const T *__restrict__ as, const T *__restrict__ bs,
                                             int wm, int wn, int lane, int s) {
  const int r = lane & 15, q = lane >> 4;
  constexpr int P = DgCfg<T>::P;
#pragma unroll
  for (int j = 0; j < 4; ++j) f.a[j] = as[(wm * 64 + j * 16 + r) * P + 4 * q + s];
#pragma unroll
  for (int i = 0; i < 4; ++i) f.b[i] = bs[(wn * 64 + i * 16 + r) * P + 4 * q + s];
}

template <typename T>
__device__ __forceinline__ void dg_mfma_slot(typename MF<T>::v4 (&acc)[4][4], const DgSlot<T> &f) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = MF<T>::mma(f.b[i], f.a[j], acc[i][j]);
}

template <typename T>
__global__ __launch_bounds__(DG_NT, 2) void k_dgemm2(const double *__restrict__ A, const double *__restrict__ B,
                                                     double *__restrict__ C, const GemmTask *__restrict__ tasks,
                                                     const WorkUnit *__restrict__ units, int nunits,
                                                     const int64_t *__restrict__ tab, double alpha, double beta) {
  using v4 = typename MF<T>::v4;
  constexpr int P = DgCfg<T>::P;
  __shared__ __align__(16) T As[2][128 * P];
  __shared__ __align__(16) T Bs[2][128 * P];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  const int G = gridDim.x, b = blockIdx.x;
  int ustart, uend, ustep;
  if (G % 8 == 0 && nunits >= G) {
    const int nb = G / 8, x = b % 8, lb = b / 8;
    const int chunk = (nunits + 7) / 8;
    ustart = x * chunk + lb;
    uend = min(nunits, (x + 1) * chunk);
    ustep = nb;
  } else {
    ustart = b;
    uend = nunits;
    ustep = G;
  }

  for (int u = ustart; u < uend; u += ustep) {
    const WorkUnit wu = units[u];
    v4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = v4{0, 0, 0, 0};

    for (int t = wu.tbeg; t < wu.tend; ++t) {
      const GemmTask tk = tasks[t];
      if (tk.K <= 0) continue;
      DgLoader la, lb;
      la.init(tid, wu.m0, tk.M, tab + tk.tAM);
      lb.init(tid, wu.n0, tk.N, tab + tk.tBN);
      const double *Ab = A + tk.a_off;
      const double *Bb = B + tk.b_off;
      const int64_t *tAK = tab + tk.tAK;
      const int64_t *tBK = tab + tk.tBK;
      const int K = tk.K, S = (K + DG_BK - 1) / DG_BK;
      DgRegs ra, rb;
// load stage st and fetch the k offsets of stage st + 1 (one slot: with two
// workgroups per CU the table read's wait behind the data loads is covered)
#define ITX_DG2_LOAD(st, sl)                                  \
  {                                                           \
    la.load(Ab, ra, 0);                                       \
    lb.load(Bb, rb, 0);                                       \
    la.fetch_k(((st) + 1) * DG_BK, K, tAK, 0);                \
    lb.fetch_k(((st) + 1) * DG_BK, K, tBK, 0);                \
  }
      la.fetch_k(0, K, tAK, 0);
      lb.fetch_k(0, K, tBK, 0);
      ITX_DG2_LOAD(0, 0)
      __syncthreads();  // the previous task's LDS reads are done
      la.store(As[0], ra);
      lb.store(Bs[0], rb);
      if (S > 1) ITX_DG2_LOAD(1, 1)
      __syncthreads();
      DgSlot<T> cur, nxt;
      dg_read_slot(cur, As[0], Bs[0], wm, wn, lane, 0);
      // stage n (buffer n & 1, load slot sl = n & 1): stage n+1 (in the load
      // registers) into the other buffer, whose stage n-1 reads ended before
      // the previous barrier; the loads of stage n+2; the four k slots, each
      // reading the next slot's operands
#define ITX_DG2_STAGE(n, sl)                                                      \
  {                                                                               \
    const T *as = As[sl], *bs = Bs[sl];                                           \
    if ((n) + 1 < S) {                                                            \
      la.store(As[(sl) ^ 1], ra);                                                 \
      lb.store(Bs[(sl) ^ 1], rb);                                                 \
    }                                                                             \
    if ((n) + 2 < S) ITX_DG2_LOAD((n) + 2, sl)                                    \
    dg_read_slot(nxt, as, bs, wm, wn, lane, 1);                                   \
    dg_mfma_slot(acc, cur);                                                       \
    dg_read_slot(cur, as, bs, wm, wn, lane, 2);                                   \
    dg_mfma_slot(acc, nxt);                                                       \
    dg_read_slot(nxt, as, bs, wm, wn, lane, 3);                                   \
    dg_mfma_slot(acc, cur);                                                       \
    __syncthreads(); /* stage n+1 stored by every wave; stage n's reads done */   \
    if ((n) + 1 < S) dg_read_slot(cur, As[(sl) ^ 1], Bs[(sl) ^ 1], wm, wn, lane, 0); \
    dg_mfma_slot(acc, nxt);                                                       \
  }
      int n = 0;
      for (; n + 1 < S; n += 2) {
        ITX_DG2_STAGE(n, 0)
        ITX_DG2_STAGE(n + 1, 1)
      }
      if (n < S) ITX_DG2_STAGE(n, 0)
#undef ITX_DG2_STAGE
#undef ITX_DG2_LOAD
    }

    const GemmTask t0 = tasks[wu.tbeg];
    const int64_t *tCM = tab + t0.tCM;
    const int64_t *tCN = tab + t0.tCN;
    double *Cb = C + t0.c_off;
    int64_t cm[4], cn[4][4];
    bool mv[4], nv[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int m = wu.m0 + wm * 64 + j * 16 + (lane & 15);
      mv[j] = m < t0.M;
      cm[j] = tCM[mv[j] ? m : 0];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = wu.n0 + wn * 64 + i * 16 + MF<T>::orow(lane, r);
        nv[i][r] = n < t0.N;
        cn[i][r] = tCN[nv[i][r] ? n : 0];
      }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (mv[j] && nv[i][r]) {
            const int64_t ad = cm[j] + cn[i][r];
            double v = alpha * double(acc[i][j][r]);
            if (beta != 0.0) v += beta * Cb[ad];
            Cb[ad] = v;
          }
        }
  }
}

// ---- BF16x3 form ----------------------------------------------------------
// Each FP64 operand value is rounded to FP32 and split into two BF16 terms
// x ~ hi + lo (hi = bf16(x), lo = bf16(x - hi): 16-17 significant bits);
// a product is formed from the three leading terms hi*hi + hi*lo + lo*hi on
// the BF16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation (the
// dropped lo*lo term is below 2^-16 relative).  LDS tiles are [row][k]
// with k contiguous so that a lane's 8 consecutive k of one row (its MFMA
// operand) are one 16-byte read; rows are padded to 80 bytes, which keeps
// the 16 rows read by a quarter-wave on distinct banks.
constexpr int BK3 = 32;        // K depth of one LDS stage = one bf16 MFMA
constexpr int LDK3 = BK3 + 8;  // bf16 elements per LDS row
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint32_t bf16_rne(uint32_t u) {  // upper 16 bits, round to nearest even
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ void split_bf16(double x, unsigned short &hi, unsigned short &lo) {
  const float f = float(x);
  const uint32_t h = bf16_rne(__float_as_uint(f));
  const float rem = f - __uint_as_float(h << 16);
  hi = (unsigned short)h;
  lo = (unsigned short)bf16_rne(__float_as_uint(rem));
}

// The pre-split operand format: y[i] = hi | lo << 16 of split_bf16(x[i]).
__global__ __launch_bounds__(256) void k_split_bf16(const double *__restrict__ x, uint32_t *__restrict__ y, size_t n) {
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) {
    unsigned short h, l;
    split_bf16(x[i], h, l);
    y[i] = uint32_t(h) | (uint32_t(l) << 16);
  }
}

// With V = 4 (k-contiguous operands whose k runs come in aligned quads,
// checked when the plan is built) a lane loads four consecutive k of one
// row as two 16-byte loads and writes their hi and lo terms to LDS as one
// 8-byte store each.
// With SP the operand comes pre-split (itx_split_bf16): one 32-bit word per
// element, hi term in the low half and lo term in the high half, as
// split_bf16 forms them; the loader moves half the bytes and does no
// conversion (V = 4: one 16-byte load per quad).
template <int BR, bool KC, int NTH, int V = 1, bool SP = false>
struct TileLoader3 {
  static constexpr int PER = BR * BK3 / NTH;
  static_assert(PER >= 1, "tile too small for the thread count");
  static_assert(V == 1 || (KC && V == 4 && PER % V == 0), "vector loads need a k-contiguous operand");
  static constexpr int NV = PER / V;
  static constexpr int LK = BK3 / V;
  static constexpr int NK = KC ? 1 : PER;
  using RT = std::conditional_t<SP, uint32_t, double>;
  RT r[PER];
  bool rok[NV];
  int64_t roff[KC ? NV : 1];
  bool rv[KC ? NV : 1];
  int64_t ko[NK];
  bool kv[NK];
  int lr, lk;

  __device__ __forceinline__ void init(int tid, int row0, int R, const int64_t *__restrict__ tabR) {
    if constexpr (KC) {
      lk = (tid % LK) * V;
      lr = tid / LK;
#pragma unroll
      for (int p = 0; p < NV; ++p) {
        const int row = row0 + lr + (NTH / LK) * p;
        rv[p] = row < R;
        roff[p] = tabR[rv[p] ? row : 0];
      }
    } else {
      lr = tid % BR;
      lk = tid / BR;
      const int row = row0 + lr;
      rv[0] = row < R;
      roff[0] = tabR[rv[0] ? row : 0];
    }
  }
  __device__ __forceinline__ void fetch_k(int k0, int K, const int64_t *__restrict__ tabK) {
#pragma unroll
    for (int q = 0; q < NK; ++q) {
      const int k = k0 + lk + (KC ? 0 : (NTH / BR) * q);
      kv[q] = k < K;
      ko[q] = tabK[kv[q] ? k : 0];
    }
  }
  __device__ __forceinline__ void load(const void *__restrict__ base_) {
    const RT *__restrict__ base = static_cast<const RT *>(base_);
#pragma unroll
    for (int p = 0; p < NV; ++p) {
      const bool ok = KC ? (kv[0] && rv[p]) : (kv[p] && rv[0]);
      const int64_t off = (KC ? roff[p] + ko[0] : roff[0] + ko[p]);
      if constexpr (V == 4 && SP) {
        const uint4 v = *reinterpret_cast<const uint4 *>(base + (ok ? off : 0));
        r[4 * p] = v.x;
        r[4 * p + 1] = v.y;
        r[4 * p + 2] = v.z;
        r[4 * p + 3] = v.w;
      } else if constexpr (V == 4) {
        const double2 *q = reinterpret_cast<const double2 *>(base + (ok ? off : 0));
        const double2 v0 = q[0], v1 = q[1];
        r[4 * p] = v0.x;
        r[4 * p + 1] = v0.y;
        r[4 * p + 2] = v1.x;
        r[4 * p + 3] = v1.y;
      } else {
        r[p] = base[ok ? off : 0];
      }
      rok[p] = ok;
    }
  }
  __device__ __forceinline__ void store(unsigned short *__restrict__ sh, unsigned short *__restrict__ sl) const {
    if constexpr (V == 4 && SP) {
#pragma unroll
      for (int p = 0; p < NV; ++p) {
        const int idx = (lr + (NTH / LK) * p) * LDK3 + lk;
        const uint32_t z = rok[p] ? 0xFFFFFFFFu : 0u;
        const uint32_t w0 = r[4 * p] & z, w1 = r[4 * p + 1] & z, w2 = r[4 * p + 2] & z, w3 = r[4 * p + 3] & z;
        *reinterpret_cast<uint2 *>(&sh[idx]) = uint2{(w0 & 0xFFFFu) | (w1 << 16), (w2 & 0xFFFFu) | (w3 << 16)};
        *reinterpret_cast<uint2 *>(&sl[idx]) =
            uint2{(w0 >> 16) | (w1 & 0xFFFF0000u), (w2 >> 16) | (w3 & 0xFFFF0000u)};
      }
    } else if constexpr (SP) {
#pragma unroll
      for (int p = 0; p < PER; ++p) {
        const int idx = KC ? (lr + (NTH / BK3) * p) * LDK3 + lk : lr * LDK3 + lk + (NTH / BR) * p;
        const uint32_t w = rok[p] ? r[p] : 0u;
        sh[idx] = (unsigned short)(w & 0xFFFFu);
        sl[idx] = (unsigned short)(w >> 16);
      }
    } else if constexpr (V == 4) {
#pragma unroll
      for (int p = 0; p < NV; ++p) {
        const int idx = (lr + (NTH / LK) * p) * LDK3 + lk;
        unsigned short h[4] = {0, 0, 0, 0}, l[4] = {0, 0, 0, 0};
        if (rok[p]) {
#pragma unroll
          for (int e = 0; e < 4; ++e) split_bf16(r[4 * p + e], h[e], l[e]);
        }
        *reinterpret_cast<uint2 *>(&sh[idx]) =
            uint2{uint32_t(h[0]) | (uint32_t(h[1]) << 16), uint32_t(h[2]) | (uint32_t(h[3]) << 16)};
        *reinterpret_cast<uint2 *>(&sl[idx]) =
            uint2{uint32_t(l[0]) | (uint32_t(l[1]) << 16), uint32_t(l[2]) | (uint32_t(l[3]) << 16)};
      }
    } else {
#pragma unroll
      for (int p = 0; p < PER; ++p) {
        const int idx = KC ? (lr + (NTH / BK3) * p) * LDK3 + lk : lr * LDK3 + lk + (NTH / BR) * p;
        unsigned short h = 0, l = 0;
        if (rok[p]) split_bf16(r[p], h, l);
        sh[idx] = h;
        sl[idx] = l;
      }
    }
  }
};

// SP: both operands pre-split (their pointers are the split words, typed
// as the other forms' FP64 operands)
template <int BM, int BN, int GM, int GN, bool AKC, bool BKC, int V, bool SP = false>
__global__ __launch_bounds__(64 * GM * GN) void k_gemm3(const double *__restrict__ A, const double *__restrict__ B,
                                                        double *__restrict__ C, const GemmTask *__restrict__ tasks,
                                                        const WorkUnit *__restrict__ units, int nunits,
                                                        const int64_t *__restrict__ tab, double alpha, double beta) {
  constexpr int NTH = 64 * GM * GN;
  typedef float v4 __attribute__((ext_vector_type(4)));
  using LA = TileLoader3<BM, AKC, NTH, V, SP>;
  using LB = TileLoader3<BN, BKC, NTH, V, SP>;
  constexpr int WM = BM / GM, WN = BN / GN, FM = WM / 16, FN = WN / 16;
  static_assert(FM >= 1 && FN >= 1, "wave tile below one MFMA fragment");
  __shared__ __align__(16) unsigned short Ah[2][BM * LDK3], Al[2][BM * LDK3];
  __shared__ __align__(16) unsigned short Bh[2][BN * LDK3], Bl[2][BN * LDK3];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % GM, wn = wave / GM;
  const int G = gridDim.x, b = blockIdx.x;
  int ustart, uend, ustep;
  if (G % 8 == 0 && nunits >= G) {
    const int nb = G / 8, x = b % 8, lb = b / 8;
    const int chunk = (nunits + 7) / 8;
    ustart = x * chunk + lb;
    uend = min(nunits, (x + 1) * chunk);
    ustep = nb;
  } else {
    ustart = b;
    uend = nunits;
    ustep = G;
  }

  for (int u = ustart; u < uend; u += ustep) {
    const WorkUnit wu = units[u];
    v4 acc[FN][FM];
#pragma unroll
    for (int i = 0; i < FN; ++i)
#pragma unroll
      for (int j = 0; j < FM; ++j) acc[i][j] = v4{0, 0, 0, 0};

    for (int t = wu.tbeg; t < wu.tend; ++t) {
      const GemmTask tk = tasks[t];
      if (tk.K <= 0) continue;
      LA la;
      LB lb;
      la.init(tid, wu.m0, tk.M, tab + tk.tAM);
      lb.init(tid, wu.n0, tk.N, tab + tk.tBN);
      const void *Ab = SP ? static_cast<const void *>(reinterpret_cast<const uint32_t *>(A) + tk.a_off)
                          : static_cast<const void *>(A + tk.a_off);
      const void *Bb = SP ? static_cast<const void *>(reinterpret_cast<const uint32_t *>(B) + tk.b_off)
                          : static_cast<const void *>(B + tk.b_off);
      const int64_t *tAK = tab + tk.tAK;
      const int64_t *tBK = tab + tk.tBK;
      la.fetch_k(0, tk.K, tAK);
      lb.fetch_k(0, tk.K, tBK);
      la.load(Ab);
      lb.load(Bb);
      la.fetch_k(BK3, tk.K, tAK);
      lb.fetch_k(BK3, tk.K, tBK);
      la.store(Ah[0], Al[0]);
      lb.store(Bh[0], Bl[0]);
      __syncthreads();
      int cur = 0;
      for (int k0 = 0; k0 < tk.K; k0 += BK3) {
        const bool nxt = k0 + BK3 < tk.K;
        if (nxt) {
          la.load(Ab);
          lb.load(Bb);
          la.fetch_k(k0 + 2 * BK3, tk.K, tAK);
          lb.fetch_k(k0 + 2 * BK3, tk.K, tBK);
        }
        const int k8 = 8 * (lane >> 4);
        bf16x8 ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
        for (int j = 0; j < FM; ++j) {
          const int m = wm * WM + j * 16 + (lane & 15);
          ah[j] = *reinterpret_cast<const bf16x8 *>(&Ah[cur][m * LDK3 + k8]);
          al[j] = *reinterpret_cast<const bf16x8 *>(&Al[cur][m * LDK3 + k8]);
        }
#pragma unroll
        for (int i = 0; i < FN; ++i) {
          const int n = wn * WN + i * 16 + (lane & 15);
          bh[i] = *reinterpret_cast<const bf16x8 *>(&Bh[cur][n * LDK3 + k8]);
          bl[i] = *reinterpret_cast<const bf16x8 *>(&Bl[cur][n * LDK3 + k8]);
        }
        // D^T tile as in k_gemm; the two cross terms first, then hi*hi
#pragma unroll
        for (int i = 0; i < FN; ++i)
#pragma unroll
          for (int j = 0; j < FM; ++j) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[i], al[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[i], ah[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[i], ah[j], acc[i][j], 0, 0, 0);
          }
        if (nxt) {
          la.store(Ah[cur ^ 1], Al[cur ^ 1]);
          lb.store(Bh[cur ^ 1], Bl[cur ^ 1]);
        }
        __syncthreads();
        cur ^= 1;
      }
    }

    const GemmTask t0 = tasks[wu.tbeg];
    const int64_t *tCM = tab + t0.tCM;
    const int64_t *tCN = tab + t0.tCN;
    double *Cb = C + t0.c_off;
#pragma unroll
    for (int j = 0; j < FM; ++j) {
      const int m = wu.m0 + wm * WM + j * 16 + (lane & 15);
      if (m >= t0.M) continue;
      const int64_t cm = tCM[m];
#pragma unroll
      for (int i = 0; i < FN; ++i) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = wu.n0 + wn * WN + i * 16 + 4 * (lane >> 4) + r;
          if (n < t0.N) {
            const int64_t ad = cm + tCN[n];
            double v = alpha * double(acc[i][j][r]);
            if (beta != 0.0) v += beta * Cb[ad];
            Cb[ad] = v;
          }
        }
      }
    }
  }
}

// Skinny contractions (N <= 32 and K <= 64 for every task, e.g. the
// MPO-tensor stages of H*psi): HBM-bound, so one thread owns one row m of
// a 256-row unit and accumulates its N outputs in registers (FP64 FMA)
// while B (K x N) sits in LDS.  Same task/unit/table format as k_gemm.
constexpr int SK_ROWS = 256, SK_N = 32, SK_K = 64;

// V4: every task's A is contiguous along k in runs of 4 starting at even
// offsets (K a multiple of 4): a row's 8 k of a batch are two runs, read as
// four 16-byte loads instead of eight 8-byte loads.
template <bool V4>
__global__ __launch_bounds__(SK_ROWS) void k_skinny(const double *__restrict__ A, const double *__restrict__ B,
                                                    double *__restrict__ C, const GemmTask *__restrict__ tasks,
                                                    const WorkUnit *__restrict__ units, int nunits,
                                                    const int64_t *__restrict__ tab, double alpha, double beta) {
  __shared__ double Bs[SK_K * SK_N];
  const int tid = threadIdx.x;
  for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
    const WorkUnit wu = units[u];
    const int m = wu.m0 + tid;
    double acc[SK_N];
#pragma unroll
    for (int n = 0; n < SK_N; ++n) acc[n] = 0.0;
    for (int t = wu.tbeg; t < wu.tend; ++t) {
      const GemmTask tk = tasks[t];
      const int N = tk.N, K = tk.K;
      __syncthreads();  // the previous task's B tile is no longer read
      for (int x = tid; x < K * N; x += SK_ROWS) {
        const int k = x / N, n = x - k * N;
        Bs[k * SK_N + n] = B[tk.b_off + tab[tk.tBK + k] + tab[tk.tBN + n]];
      }
      __syncthreads();
      if (m < tk.M) {
        const double *Ar = A + tk.a_off + tab[tk.tAM + m];
        const int64_t *tAK = tab + tk.tAK;
        for (int k0 = 0; k0 < K; k0 += 8) {
          double av[8];  // issue 8 independent loads before using any
          if constexpr (V4) {
            const bool two = k0 + 4 < K;  // K % 4 == 0: one or two runs left
            const double2 *r0 = reinterpret_cast<const double2 *>(Ar + tAK[k0]);
            const double2 *r1 = reinterpret_cast<const double2 *>(Ar + tAK[two ? k0 + 4 : k0]);
            const double2 x0 = r0[0], x1 = r0[1], x2 = r1[0], x3 = r1[1];
            av[0] = x0.x;
            av[1] = x0.y;
            av[2] = x1.x;
            av[3] = x1.y;
            av[4] = two ? x2.x : 0.0;
            av[5] = two ? x2.y : 0.0;
            av[6] = two ? x3.x : 0.0;
            av[7] = two ? x3.y : 0.0;
          } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) av[j] = (k0 + j < K) ? Ar[tAK[k0 + j]] : 0.0;
          }
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            if (k0 + j < K) {
#pragma unroll
              for (int n = 0; n < SK_N; ++n)
                if (n < N) acc[n] += av[j] * Bs[(k0 + j) * SK_N + n];
            }
          }
        }
      }
    }
    const GemmTask t0 = tasks[wu.tbeg];
    if (m < t0.M) {
      double *Cm = C + t0.c_off + tab[t0.tCM + m];
      const int64_t *tCN = tab + t0.tCN;
#pragma unroll
      for (int n = 0; n < SK_N; ++n) {
        if (n < t0.N) {
          double v = alpha * acc[n];
          if (beta != 0.0) v += beta * Cm[tCN[n]];
          Cm[tCN[n]] = v;
        }
      }
    }
  }
}

// The skinny path when every task has K <= 8 and N <= 8 (the block-pair
// tasks of a block-sparse MPO stage: site and MPO-link blocks of a few
// rows): a unit then sums many tiny tasks, and the per-task load round trip
// and LDS barriers of k_skinny dominate.  Here the B tiles of as many tasks
// as fit are staged in LDS at once (rows stacked), and the A elements of
// SKQ_G tasks are loaded together before any is used.  Same accumulation
// order as k_skinny (tasks in order, k in order), so the same results.
// Each thread computes two rows, SK_ROWS apart (units of 2 * SK_ROWS rows):
// the per-unit task reads, B staging and barriers are shared by twice the
// rows, and a thread has the loads of two rows in flight.
constexpr int SKQ_K = 8, SKQ_N = 8, SKQ_G = 2, SKQ_R = 2;

__global__ __launch_bounds__(SK_ROWS) void k_skinny_q(const double *__restrict__ A, const double *__restrict__ B,
                                                      double *__restrict__ C, const GemmTask *__restrict__ tasks,
                                                      const WorkUnit *__restrict__ units, int nunits,
                                                      const int64_t *__restrict__ tab, double alpha, double beta) {
  __shared__ double Bs[SK_K * SKQ_N];
  const int tid = threadIdx.x;
  for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
    const WorkUnit wu = units[u];
    const GemmTask t0 = tasks[wu.tbeg];
    const int N = t0.N, M = t0.M;
    int mr[SKQ_R];
    bool in[SKQ_R];
#pragma unroll
    for (int h = 0; h < SKQ_R; ++h) {
      mr[h] = wu.m0 + tid + h * SK_ROWS;
      in[h] = mr[h] < M;
    }
    double acc[SKQ_R][SKQ_N];
#pragma unroll
    for (int h = 0; h < SKQ_R; ++h)
#pragma unroll
      for (int n = 0; n < SKQ_N; ++n) acc[h][n] = 0.0;
    for (int tc = wu.tbeg; tc < wu.tend;) {
      // the next tasks whose stacked B rows fit the LDS tile
      int te = tc, kt = 0;
      while (te < wu.tend && kt + tasks[te].K <= SK_K) kt += tasks[te++].K;
      __syncthreads();  // the previous chunk's B rows are no longer read
      for (int t = tc, kb = 0; t < te; ++t) {
        const GemmTask tk = tasks[t];
        for (int x = tid; x < tk.K * N; x += SK_ROWS) {
          const int k = x / N, n = x - k * N;
          Bs[(kb + k) * SKQ_N + n] = B[tk.b_off + tab[tk.tBK + k] + tab[tk.tBN + n]];
        }
        kb += tk.K;
      }
      __syncthreads();
      if (in[0]) {
        for (int t = tc, kb = 0; t < te; t += SKQ_G) {
          double av[SKQ_R][SKQ_G][SKQ_K];
          int kk[SKQ_G];
#pragma unroll
          for (int q = 0; q < SKQ_G; ++q) {
            kk[q] = 0;
            if (t + q < te) {
              const GemmTask tk = tasks[t + q];
              kk[q] = tk.K;
              const int64_t *tAK = tab + tk.tAK;
#pragma unroll
              for (int h = 0; h < SKQ_R; ++h) {
                const double *Ar = A + tk.a_off + tab[tk.tAM + (in[h] ? mr[h] : mr[0])];
#pragma unroll
                for (int j = 0; j < SKQ_K; ++j) av[h][q][j] = (j < tk.K && in[h]) ? Ar[tAK[j]] : 0.0;
              }
            }
          }
#pragma unroll
          for (int q = 0; q < SKQ_G; ++q) {
#pragma unroll
            for (int j = 0; j < SKQ_K; ++j) {
              if (j < kk[q]) {
#pragma unroll
                for (int n = 0; n < SKQ_N; ++n)
                  if (n < N) {
                    const double bv = Bs[(kb + j) * SKQ_N + n];
#pragma unroll
                    for (int h = 0; h < SKQ_R; ++h) acc[h][n] += av[h][q][j] * bv;
                  }
              }
            }
            kb += kk[q];
          }
        }
      }
      tc = te;
    }
#pragma unroll
    for (int h = 0; h < SKQ_R; ++h) {
      if (in[h]) {
        double *Cm = C + t0.c_off + tab[t0.tCM + mr[h]];
        const int64_t *tCN = tab + t0.tCN;
#pragma unroll
        for (int n = 0; n < SKQ_N; ++n) {
          if (n < N) {
            double v = alpha * acc[h][n];
            if (beta != 0.0) v += beta * Cm[tCN[n]];
            Cm[tCN[n]] = v;
          }
        }
      }
    }
  }
}

// Small-block contractions (C blocks of a few thousand elements: the QN
// blocks of DMRG at moderate bond dimension): latency-bound, so FP64 VALU
// arithmetic in every form.  A unit is a 64 x 32 tile of one C block; for
// each task and 64-wide k chunk the A tile and the B tile are staged
// cooperatively in LDS -- one round of independent loads, lanes packed
// along the operand's unit stride -- then each thread accumulates 8
// outputs of one row.  Same task/unit/table format as k_gemm.
constexpr int SM_M = 64, SM_N = 32, SM_K = 64, SM_T = 256, SM_NPT = 8;
constexpr int SM_LDA = SM_M + 1;  // pad: the k-contiguous staging stores hit distinct banks

template <bool AKC, bool BKC>
__global__ __launch_bounds__(SM_T) void k_small(const double *__restrict__ A, const double *__restrict__ B,
                                                double *__restrict__ C, const GemmTask *__restrict__ tasks,
                                                const WorkUnit *__restrict__ units, int nunits,
                                                const int64_t *__restrict__ tab, double alpha, double beta) {
  __shared__ double As[SM_K * SM_LDA];
  __shared__ double Bs[SM_K * SM_N];
  const int tid = threadIdx.x;
  const int mi = tid & (SM_M - 1), ng = tid >> 6;  // output row, group of 8 columns
  for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
    const WorkUnit wu = units[u];
    double acc[SM_NPT];
#pragma unroll
    for (int j = 0; j < SM_NPT; ++j) acc[j] = 0.0;
    for (int t = wu.tbeg; t < wu.tend; ++t) {
      const GemmTask tk = tasks[t];
      const int rows = min(SM_M, tk.M - wu.m0), cols = min(SM_N, tk.N - wu.n0);
      for (int k0 = 0; k0 < tk.K; k0 += SM_K) {
        const int kc = min(SM_K, tk.K - k0);
        __syncthreads();  // the previous chunk's tiles are no longer read
        // A tile (rows x kc) -> As[k][m]; consecutive lanes follow A's
        // unit stride (k when A is k-contiguous, else m), packed densely
        {
          const int na = rows * kc, lead = AKC ? kc : rows;
          for (int e = tid; e < na; e += SM_T) {
            const int q = e / lead, r = e - q * lead;
            const int m = AKC ? q : r, k = AKC ? r : q;
            As[k * SM_LDA + m] = A[tk.a_off + tab[tk.tAM + wu.m0 + m] + tab[tk.tAK + k0 + k]];
          }
        }
        // B tile (kc x cols) -> Bs[k][n], likewise
        {
          const int nb = kc * cols, lead = BKC ? kc : cols;
          for (int e = tid; e < nb; e += SM_T) {
            const int q = e / lead, r = e - q * lead;
            const int n = BKC ? q : r, k = BKC ? r : q;
            Bs[k * SM_N + n] = B[tk.b_off + tab[tk.tBK + k0 + k] + tab[tk.tBN + wu.n0 + n]];
          }
          // columns beyond the block are read by the last column group
          for (int e = tid; e < kc * (SM_N - cols); e += SM_T) {
            const int k = e / (SM_N - cols), n = cols + (e - k * (SM_N - cols));
            Bs[k * SM_N + n] = 0.0;
          }
        }
        __syncthreads();
        const double *bp = Bs + ng * SM_NPT;
        for (int k = 0; k < kc; ++k) {
          const double a = As[k * SM_LDA + mi];
#pragma unroll
          for (int j = 0; j < SM_NPT; ++j) acc[j] = fma(a, bp[k * SM_N + j], acc[j]);
        }
      }
    }
    if (wu.slot >= 0) {  // partial tile of a split unit (C is the workspace)
      double *P = C + size_t(wu.slot) * (SM_M * SM_N);
#pragma unroll
      for (int j = 0; j < SM_NPT; ++j) P[(ng * SM_NPT + j) * SM_M + mi] = acc[j];
      continue;
    }
    const GemmTask t0 = tasks[wu.tbeg];
    const int m = wu.m0 + mi;
    if (m < t0.M) {
      double *Cm = C + t0.c_off + tab[t0.tCM + m];
#pragma unroll
      for (int j = 0; j < SM_NPT; ++j) {
        const int n = wu.n0 + ng * SM_NPT + j;
        if (n < t0.N && ng * SM_NPT + j < SM_N) {
          double *cp = Cm + tab[t0.tCN + n];
          double v = alpha * acc[j];
          if (beta != 0.0) v += beta * (*cp);
          *cp = v;
        }
      }
    }
  }
}

// Small-block contractions on the matrix cores: the FP32 and BF16x3 forms
// of the small-block path (same 64 x 32 unit tiles, task/unit/table format,
// split units and partial-tile layout as k_small).  Each k chunk of the A
// and B tiles is staged in LDS zero-padded to the MFMA depth (and to the
// full tile: rows and columns outside the block are zeros, so the padded
// fragments add nothing); wave w computes rows 16w..16w+15 of the tile as
// two 16 x 16 MFMA fragments.
// FP32: operands rounded to FP32 at staging, v_mfma_f32_16x16x4f32.
constexpr int SMF_LDA = SM_M + 17, SMF_LDB = SM_N + 17;  // odd pitches: k-strided staging hits distinct banks

template <bool AKC, bool BKC>
__global__ __launch_bounds__(SM_T) void k_small_f32(const double *__restrict__ A, const double *__restrict__ B,
                                                    double *__restrict__ C, const GemmTask *__restrict__ tasks,
                                                    const WorkUnit *__restrict__ units, int nunits,
                                                    const int64_t *__restrict__ tab, double alpha, double beta) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  __shared__ float As[SM_K * SMF_LDA];
  __shared__ float Bs[SM_K * SMF_LDB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
    const WorkUnit wu = units[u];
    v4 acc[2] = {v4{0, 0, 0, 0}, v4{0, 0, 0, 0}};
    for (int t = wu.tbeg; t < wu.tend; ++t) {
      const GemmTask tk = tasks[t];
      const int rows = min(SM_M, tk.M - wu.m0), cols = min(SM_N, tk.N - wu.n0);
      for (int k0 = 0; k0 < tk.K; k0 += SM_K) {
        const int kc = min(SM_K, tk.K - k0), kcp = (kc + 3) & ~3;
        __syncthreads();  // the previous chunk's tiles are no longer read
        {
          const int lead = AKC ? kcp : SM_M;
          for (int e = tid; e < SM_M * kcp; e += SM_T) {
            const int q = e / lead, r = e - q * lead;
            const int m = AKC ? q : r, k = AKC ? r : q;
            float v = 0.f;
            if (m < rows && k < kc) v = float(A[tk.a_off + tab[tk.tAM + wu.m0 + m] + tab[tk.tAK + k0 + k]]);
            As[k * SMF_LDA + m] = v;
          }
        }
        {
          const int lead = BKC ? kcp : SM_N;
          for (int e = tid; e < SM_N * kcp; e += SM_T) {
            const int q = e / lead, r = e - q * lead;
            const int n = BKC ? q : r, k = BKC ? r : q;
            float v = 0.f;
            if (n < cols && k < kc) v = float(B[tk.b_off + tab[tk.tBK + k0 + k] + tab[tk.tBN + wu.n0 + n]]);
            Bs[k * SMF_LDB + n] = v;
          }
        }
        __syncthreads();
        const int kl = lane >> 4, ml = 16 * wave + (lane & 15);
        for (int kk = 0; kk < kcp; kk += 4) {
          const float a = As[(kk + kl) * SMF_LDA + ml];
          const float b0 = Bs[(kk + kl) * SMF_LDB + (lane & 15)];
          const float b1 = Bs[(kk + kl) * SMF_LDB + 16 + (lane & 15)];
          // D^T fragments: MFMA rows run over n, columns over m
          acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(b0, a, acc[0], 0, 0, 0);
          acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(b1, a, acc[1], 0, 0, 0);
        }
      }
    }
    const int ml = 16 * wave + (lane & 15);
    if (wu.slot >= 0) {  // partial tile of a split unit (C is the workspace)
      double *P = C + size_t(wu.slot) * (SM_M * SM_N);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) P[(16 * i + 4 * (lane >> 4) + r) * SM_M + ml] = double(acc[i][r]);
      continue;
    }
    const GemmTask t0 = tasks[wu.tbeg];
    const int m = wu.m0 + ml;
    if (m < t0.M) {
      double *Cm = C + t0.c_off + tab[t0.tCM + m];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = wu.n0 + 16 * i + 4 * (lane >> 4) + r;
          if (n < t0.N) {
            double *cp = Cm + tab[t0.tCN + n];
            double v = alpha * double(acc[i][r]);
            if (beta != 0.0) v += beta * (*cp);
            *cp = v;
          }
        }
    }
  }
}

// BF16x3: each operand value split at staging into x ~ hi + lo (the split
// of k_gemm3), tiles [row][k] with k contiguous so that a lane's 8
// consecutive k are one 16-byte LDS read; hi*lo + lo*hi + hi*hi on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation.  k chunks are padded to
// the MFMA depth of 32.
constexpr int SM3_LD = SM_K + 8;  // bf16 elements per LDS row (16-byte aligned rows)

template <bool AKC, bool BKC>
__global__ __launch_bounds__(SM_T) void k_small_bf16x3(const double *__restrict__ A, const double *__restrict__ B,
                                                       double *__restrict__ C, const GemmTask *__restrict__ tasks,
                                                       const WorkUnit *__restrict__ units, int nunits,
                                                       const int64_t *__restrict__ tab, double alpha,
                                                       double beta) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  __shared__ __align__(16) unsigned short Ah[SM_M * SM3_LD], Al[SM_M * SM3_LD];
  __shared__ __align__(16) unsigned short Bh[SM_N * SM3_LD], Bl[SM_N * SM3_LD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
    const WorkUnit wu = units[u];
    v4 acc[2] = {v4{0, 0, 0, 0}, v4{0, 0, 0, 0}};
    for (int t = wu.tbeg; t < wu.tend; ++t) {
      const GemmTask tk = tasks[t];
      const int rows = min(SM_M, tk.M - wu.m0), cols = min(SM_N, tk.N - wu.n0);
      for (int k0 = 0; k0 < tk.K; k0 += SM_K) {
        const int kc = min(SM_K, tk.K - k0), kcp = (kc + 31) & ~31;
        __syncthreads();
        {
          const int lead = AKC ? kcp : SM_M;
          for (int e = tid; e < SM_M * kcp; e += SM_T) {
            const int q = e / lead, r = e - q * lead;
            const int m = AKC ? q : r, k = AKC ? r : q;
            unsigned short h = 0, l = 0;
            if (m < rows && k < kc) split_bf16(A[tk.a_off + tab[tk.tAM + wu.m0 + m] + tab[tk.tAK + k0 + k]], h, l);
            Ah[m * SM3_LD + k] = h;
            Al[m * SM3_LD + k] = l;
          }
        }
        {
          const int lead = BKC ? kcp : SM_N;
          for (int e = tid; e < SM_N * kcp; e += SM_T) {
            const int q = e / lead, r = e - q * lead;
            const int n = BKC ? q : r, k = BKC ? r : q;
            unsigned short h = 0, l = 0;
            if (n < cols && k < kc) split_bf16(B[tk.b_off + tab[tk.tBK + k0 + k] + tab[tk.tBN + wu.n0 + n]], h, l);
            Bh[n * SM3_LD + k] = h;
            Bl[n * SM3_LD + k] = l;
          }
        }
        __syncthreads();
        const int k8 = 8 * (lane >> 4), ml = 16 * wave + (lane & 15);
        for (int kk = 0; kk < kcp; kk += 32) {
          const bf16x8 ah = *reinterpret_cast<const bf16x8 *>(&Ah[ml * SM3_LD + kk + k8]);
          const bf16x8 al = *reinterpret_cast<const bf16x8 *>(&Al[ml * SM3_LD + kk + k8]);
#pragma unroll
          for (int i = 0; i < 2; ++i) {
            const int nl = 16 * i + (lane & 15);
            const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(&Bh[nl * SM3_LD + kk + k8]);
            const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(&Bl[nl * SM3_LD + kk + k8]);
            acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, al, acc[i], 0, 0, 0);
            acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl, ah, acc[i], 0, 0, 0);
            acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, ah, acc[i], 0, 0, 0);
          }
        }
      }
    }
    const int ml = 16 * wave + (lane & 15);
    if (wu.slot >= 0) {
      double *P = C + size_t(wu.slot) * (SM_M * SM_N);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) P[(16 * i + 4 * (lane >> 4) + r) * SM_M + ml] = double(acc[i][r]);
      continue;
    }
    const GemmTask t0 = tasks[wu.tbeg];
    const int m = wu.m0 + ml;
    if (m < t0.M) {
      double *Cm = C + t0.c_off + tab[t0.tCM + m];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = wu.n0 + 16 * i + 4 * (lane >> 4) + r;
          if (n < t0.N) {
            double *cp = Cm + tab[t0.tCN + n];
            double v = alpha * double(acc[i][r]);
            if (beta != 0.0) v += beta * (*cp);
            *cp = v;
          }
        }
    }
  }
}

static_assert(SM_T_RED == SM_T, "reduction pass block size");

// Reduction pass of split plans: unit u = {task (for C's tables), first
// slot, m0, n0, end slot}; C tile = alpha * (sum of the slots in order)
// + beta * C.  A is the workspace; deterministic (fixed summation order).
__global__ __launch_bounds__(SM_T) void k_reduce_tiles(const double *__restrict__ A,
                                                       const double *__restrict__ B, double *__restrict__ C,
                                                       const GemmTask *__restrict__ tasks,
                                                       const WorkUnit *__restrict__ units, int nunits,
                                                       const int64_t *__restrict__ tab, double alpha,
                                                       double beta) {
  (void)B;
  // one item = one tile x one slice of SM_T of its SM_M * SM_N elements;
  // four independent partial sums keep several loads in flight
  constexpr int PARTS = SM_M * SM_N / SM_T;
  for (int w = blockIdx.x; w < nunits * PARTS; w += gridDim.x) {
    const WorkUnit ru = units[w / PARTS];
    const GemmTask t0 = tasks[ru.tbeg];
    const int e = (w % PARTS) * SM_T + threadIdx.x;
    const int n = e / SM_M, m = e - n * SM_M;
    if (ru.m0 + m >= t0.M || ru.n0 + n >= t0.N) continue;
    const double *src = A + e;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int q = ru.tend;
    for (; q + 4 <= ru.slot; q += 4) {
      s0 += src[size_t(q) * (SM_M * SM_N)];
      s1 += src[size_t(q + 1) * (SM_M * SM_N)];
      s2 += src[size_t(q + 2) * (SM_M * SM_N)];
      s3 += src[size_t(q + 3) * (SM_M * SM_N)];
    }
    for (; q < ru.slot; ++q) s0 += src[size_t(q) * (SM_M * SM_N)];
    const double s = (s0 + s1) + (s2 + s3);
    double *cp = C + t0.c_off + tab[t0.tCM + ru.m0 + m] + tab[t0.tCN + ru.n0 + n];
    double v = alpha * s;
    if (beta != 0.0) v += beta * (*cp);
    *cp = v;
  }
}

// ---- host side --------------------------------------------------------

struct Digit {
  int64_t dim, s1, s2;
};
// A digit set in place (at most ITX_MAX_RANK digits): planning handles one
// per operand side and task, so it must not allocate
struct Digits {
  Digit d[ITX_MAX_RANK];
  int n = 0;
  void push_back(Digit x) { d[n++] = x; }
  bool empty() const { return n == 0; }
  int size() const { return n; }
  const Digit &operator[](int i) const { return d[i]; }
  Digit &back() { return d[n - 1]; }
  const Digit *begin() const { return d; }
  const Digit *end() const { return d + n; }
};

// Remove unit digits and fuse neighbours that are contiguous in both
// operands (the same simplification CProps makes before its GEMM).
void simplify(Digits &d) {
  Digits o;
  for (auto &x : d) {
    if (x.dim == 1) continue;
    if (!o.empty() && o.back().s1 * o.back().dim == x.s1 && o.back().s2 * o.back().dim == x.s2) {
      o.back().dim *= x.dim;
    } else {
      o.push_back(x);
    }
  }
  d = o;
}

int64_t digits_size(const Digits &d) {
  int64_t s = 1;
  for (auto &x : d) s *= x.dim;
  return s;
}

// Offset tables of a plan: every entry is written by ColumnCache::emit, so growing
// the vector leaves new entries uninitialised (default- instead of value-
// initialised: no zeroing pass over the megabytes of a large QN plan).
template <class T>
struct NoInitAlloc : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <class U>
  NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
  template <class U>
  void construct(U *p) noexcept {
    ::new (static_cast<void *>(p)) U;
  }
  template <class U, class... Args>
  void construct(U *p, Args &&...args) {
    ::new (static_cast<void *>(p)) U(std::forward<Args>(args)...);
  }
};
using TabVec = std::vector<int64_t, NoInitAlloc<int64_t>>;

// Offset tables are emitted one column at a time (the A / C row offsets of
// the m digits, B / C of the n digits, A / B of the k digits), each from
// the column's own digit list: the digits with more than one value, merged
// where the column is contiguous by itself (stride of one = dim x stride of
// the previous).  The merged list enumerates the same offsets in the same
// order as the joint digits (first digit fastest) and determines the table,
// so equal columns share one table, and a column that is one run of stride
// 1 (most block rows and k runs) is a prefix of a single 0, 1, 2, ... table.
struct ColDigit {
  int64_t dim, s;
};
int column_digits(const Digits &d, bool first, ColDigit *c) {
  int nc = 0;
  for (auto &x : d) {
    if (x.dim == 1) continue;
    const int64_t st = first ? x.s1 : x.s2;
    if (nc > 0 && c[nc - 1].s * c[nc - 1].dim == st)
      c[nc - 1].dim *= x.dim;
    else
      c[nc++] = ColDigit{x.dim, st};
  }
  return nc;
}

struct ColumnCache {
  struct Key {
    int n;
    ColDigit c[ITX_MAX_RANK];
    bool operator<(const Key &o) const {
      if (n != o.n) return n < o.n;
      for (int i = 0; i < n; ++i) {
        if (c[i].dim != o.c[i].dim) return c[i].dim < o.c[i].dim;
        if (c[i].s != o.c[i].s) return c[i].s < o.c[i].s;
      }
      return false;
    }
  };
  std::map<Key, int64_t> at;
  int64_t iota_start = -1, iota_len = 0;
  // start of the table of column `first` (s1) or not (s2) of d in tab
  int64_t emit(const Digits &d, bool first, TabVec &tab) {
    Key k;
    k.n = column_digits(d, first, k.c);
    int64_t n = 1;
    for (int i = 0; i < k.n; ++i) n *= k.c[i].dim;
    if (n == 0) return int64_t(tab.size());  // an empty extent reads nothing
    if (k.n == 0 || (k.n == 1 && k.c[0].s == 1)) {
      if (iota_len < n) {  // a longer 0, 1, 2, ... table (earlier ones stay valid)
        iota_len = std::max<int64_t>(n, 2 * iota_len);
        iota_start = int64_t(tab.size());
        tab.resize(tab.size() + size_t(iota_len));
        int64_t *t = tab.data() + iota_start;
        for (int64_t x = 0; x < iota_len; ++x) t[x] = x;
      }
      return iota_start;
    }
    auto it = at.find(k);
    if (it != at.end()) return it->second;
    const int64_t start = int64_t(tab.size());
    tab.resize(tab.size() + size_t(n));
    int64_t *t = tab.data() + start;
    // runs along the fastest digit, the other digits advanced once per run
    const int64_t d0 = k.c[0].dim, a0 = k.c[0].s;
    int64_t ctr[ITX_MAX_RANK] = {};
    int64_t o = 0;
    for (int64_t x = 0; x < n; x += d0) {
      for (int64_t i = 0; i < d0; ++i) t[x + i] = o + i * a0;
      for (int i = 1; i < k.n; ++i) {
        ctr[i]++;
        o += k.c[i].s;
        if (ctr[i] < k.c[i].dim) break;
        o -= k.c[i].s * k.c[i].dim;
        ctr[i] = 0;
      }
    }
    at.emplace(k, start);
    return start;
  }
};
// largest offset either table of a digit set holds (strides are positive)
void digits_max(const Digits &d, int64_t &m1, int64_t &m2) {
  m1 = m2 = 0;
  for (auto &x : d) {
    m1 += (x.dim - 1) * x.s1;
    m2 += (x.dim - 1) * x.s2;
  }
}

struct LabelMap {
  // for A index i: contracted partner in B (or -1) / position in C (or -1)
  std::vector<int> a2b, a2c, b2a, b2c;
};

int map_labels(itx_ctx *ctx, int rA, const int32_t *labA, int rB, const int32_t *labB, int rC,
               const int32_t *labC, LabelMap &lm) {
  if (rA < 0 || rB < 0 || rC < 0 || rA > ITX_MAX_RANK || rB > ITX_MAX_RANK || rC > ITX_MAX_RANK)
    return set_err(ctx, ITX_ERR_ARG, "contraction: rank out of range");
  lm.a2b.assign(rA, -1);
  lm.a2c.assign(rA, -1);
  lm.b2a.assign(rB, -1);
  lm.b2c.assign(rB, -1);
  std::vector<int> cused(rC, 0);
  for (int i = 0; i < rA; ++i) {
    if (labA[i] < 0) {
      for (int j = 0; j < rB; ++j)
        if (labB[j] == labA[i]) {
          if (lm.a2b[i] >= 0) return set_err(ctx, ITX_ERR_ARG, "contraction: repeated label in B");
          lm.a2b[i] = j;
          lm.b2a[j] = i;
        }
      if (lm.a2b[i] < 0) return set_err(ctx, ITX_ERR_ARG, "contraction: contracted label of A missing in B");
    } else {
      for (int c = 0; c < rC; ++c)
        if (labC[c] == labA[i]) {
          if (lm.a2c[i] >= 0) return set_err(ctx, ITX_ERR_ARG, "contraction: repeated label in C");
          lm.a2c[i] = c;
          cused[c]++;
        }
      if (lm.a2c[i] < 0) return set_err(ctx, ITX_ERR_ARG, "contraction: label of A missing in C");
    }
  }
  for (int j = 0; j < rB; ++j) {
    if (labB[j] < 0) {
      if (lm.b2a[j] < 0) return set_err(ctx, ITX_ERR_ARG, "contraction: contracted label of B missing in A");
    } else {
      for (int c = 0; c < rC; ++c)
        if (labC[c] == labB[j]) {
          if (lm.b2c[j] >= 0) return set_err(ctx, ITX_ERR_ARG, "contraction: repeated label in C");
          lm.b2c[j] = c;
          cused[c]++;
        }
      if (lm.b2c[j] < 0) return set_err(ctx, ITX_ERR_ARG, "contraction: label of B missing in C");
    }
  }
  for (int c = 0; c < rC; ++c)
    if (cused[c] != 1) return set_err(ctx, ITX_ERR_ARG, "contraction: label of C not found exactly once in A,B");
  return ITX_OK;
}

void colmajor_strides(int r, const int64_t *d, int64_t *s) {
  int64_t x = 1;
  for (int i = 0; i < r; ++i) {
    s[i] = x;
    x *= d[i];
  }
}

// MFMA tiles (BM x BN with GM x GN waves) and the small-block VALU path.
struct TileCfg {
  int bm, bn, gm, gn;
  double eff64, eff32;  // relative per-CU efficiency (MI355X measurements)
};
constexpr TileCfg kTiles[] = {
    {128, 128, 2, 2, 0.55, 0.70},  // 0
    {128, 64, 2, 2, 0.92, 1.00},   // 1
    {64, 128, 2, 2, 0.92, 0.95},   // 2
    {64, 64, 2, 2, 1.00, 0.95},    // 3
    {32, 32, 2, 2, 0.50, 0.45},    // 4
    {SK_ROWS, SK_N, 0, 0, 1.0, 1.0},  // 5: skinny (FP64 VALU, large M)
    {128, 128, 2, 4, 0.80, 0.85},  // 6: 8 waves
    {256, 64, 4, 2, 0.70, 0.55},   // 7: 8 waves
    {64, 256, 2, 4, 0.95, 1.00},   // 8: 8 waves
    {256, 128, 4, 2, 0.60, 0.50},  // 9: 8 waves, wave tile 64x64
    {SM_M, SM_N, 0, 0, 1.0, 1.0},  // 10: small blocks (FP64 VALU)
    {128, 128, 2, 4, 0.80, 0.85},  // 11: as 6 with 32-deep k stages
    {128, 64, 2, 2, 0.90, 0.95},   // 12: as 1 with 32-deep k stages
    {64, 128, 2, 2, 0.90, 0.95},   // 13: as 2 with 32-deep k stages
    {128, 128, 2, 2, 1.15, 0.85},  // 14: register-pipelined (k_dgemm), k-contiguous pairs only
    {128, 128, 2, 2, 1.10, 0.85},  // 15: as 14, two workgroups per CU (k_dgemm2)
};
constexpr int kNumTiles = 16;
constexpr int kDgemm = 14, kDgemm2 = 15;
constexpr int kSkinny = 5;
static_assert(kSkinny == kSkinnyTile, "tile index of the skinny path");
constexpr int kSmall = 10;
constexpr int kSkinnyOcc = 4;

// Kernel instantiation of a tile shape, operand layout and form.  All
// kernels of the engine share one argument list (see KernelArgs).
// vec: both operands k-contiguous with aligned runs of 2 (FP64 / FP32
// loaders) or 4 (BF16x3 loader) consecutive k -- the vector-load variants
template <typename T, int BM, int BN, int GM, int GN, int BKS = BK>
const void *gemm_fn(bool akc, bool bkc, bool vec) {
  if (akc && bkc && vec) return reinterpret_cast<const void *>(&k_gemm<T, BM, BN, GM, GN, true, true, BKS, 2>);
  if (akc)
    return bkc ? reinterpret_cast<const void *>(&k_gemm<T, BM, BN, GM, GN, true, true, BKS, 1>)
               : reinterpret_cast<const void *>(&k_gemm<T, BM, BN, GM, GN, true, false, BKS, 1>);
  return bkc ? reinterpret_cast<const void *>(&k_gemm<T, BM, BN, GM, GN, false, true, BKS, 1>)
             : reinterpret_cast<const void *>(&k_gemm<T, BM, BN, GM, GN, false, false, BKS, 1>);
}

template <int BM, int BN, int GM, int GN>
const void *gemm3_fn(bool akc, bool bkc, bool vec, bool split = false) {
  if (split)  // pre-split operands: k-contiguous layouts only
    return vec ? reinterpret_cast<const void *>(&k_gemm3<BM, BN, GM, GN, true, true, 4, true>)
               : reinterpret_cast<const void *>(&k_gemm3<BM, BN, GM, GN, true, true, 1, true>);
  if (akc && bkc && vec) return reinterpret_cast<const void *>(&k_gemm3<BM, BN, GM, GN, true, true, 4>);
  if (akc)
    return bkc ? reinterpret_cast<const void *>(&k_gemm3<BM, BN, GM, GN, true, true, 1>)
               : reinterpret_cast<const void *>(&k_gemm3<BM, BN, GM, GN, true, false, 1>);
  return bkc ? reinterpret_cast<const void *>(&k_gemm3<BM, BN, GM, GN, false, true, 1>)
             : reinterpret_cast<const void *>(&k_gemm3<BM, BN, GM, GN, false, false, 1>);
}

const void *gemm3_fn_tile(int tile, bool akc, bool bkc, bool vec, bool split) {
  switch (tile) {
    case 0: return gemm3_fn<128, 128, 2, 2>(akc, bkc, vec, split);
    case 1: return gemm3_fn<128, 64, 2, 2>(akc, bkc, vec, split);
    case 2: return gemm3_fn<64, 128, 2, 2>(akc, bkc, vec, split);
    case 3: return gemm3_fn<64, 64, 2, 2>(akc, bkc, vec, split);
    case 6: return gemm3_fn<128, 128, 2, 4>(akc, bkc, vec, split);
    case 7: return gemm3_fn<256, 64, 4, 2>(akc, bkc, vec, split);
    case 8: return gemm3_fn<64, 256, 2, 4>(akc, bkc, vec, split);
    case 9: return gemm3_fn<256, 128, 4, 2>(akc, bkc, vec, split);
    case 11: return gemm3_fn<128, 128, 2, 4>(akc, bkc, vec, split);
    case 12: return gemm3_fn<128, 64, 2, 2>(akc, bkc, vec, split);
    case 13: return gemm3_fn<64, 128, 2, 2>(akc, bkc, vec, split);
    case 14:
    case 15: return gemm3_fn<128, 128, 2, 2>(akc, bkc, vec, split);  // not chosen for BF16x3
    default: return gemm3_fn<32, 32, 2, 2>(akc, bkc, vec, split);
  }
}

template <typename T>
const void *gemm_fn_tile(int tile, bool akc, bool bkc, bool vec) {
  switch (tile) {
    case 0: return gemm_fn<T, 128, 128, 2, 2>(akc, bkc, vec);
    case 1: return gemm_fn<T, 128, 64, 2, 2>(akc, bkc, vec);
    case 2: return gemm_fn<T, 64, 128, 2, 2>(akc, bkc, vec);
    case 3: return gemm_fn<T, 64, 64, 2, 2>(akc, bkc, vec);
    case 6: return gemm_fn<T, 128, 128, 2, 4>(akc, bkc, vec);
    case 7: return gemm_fn<T, 256, 64, 4, 2>(akc, bkc, vec);
    case 8: return gemm_fn<T, 64, 256, 2, 4>(akc, bkc, vec);
    case 9: return gemm_fn<T, 256, 128, 4, 2>(akc, bkc, vec);
    case 11: return gemm_fn<T, 128, 128, 2, 4, 32>(akc, bkc, vec);
    case 12: return gemm_fn<T, 128, 64, 2, 2, 32>(akc, bkc, vec);
    case 13: return gemm_fn<T, 64, 128, 2, 2, 32>(akc, bkc, vec);
    case 14:  // aligned k pairs: k_dgemm; otherwise (unaligned operands) tile 0's kernel
      if (akc && bkc && vec) return reinterpret_cast<const void *>(&k_dgemm<T>);
      return gemm_fn<T, 128, 128, 2, 2>(akc, bkc, false);
    case 15:
      if (akc && bkc && vec) return reinterpret_cast<const void *>(&k_dgemm2<T>);
      return gemm_fn<T, 128, 128, 2, 2>(akc, bkc, false);
    default: return gemm_fn<T, 32, 32, 2, 2>(akc, bkc, vec);
  }
}

// Entry point and workgroup size of a plan's kernel.
const void *kernel_of(int tile, int prec, bool akc, bool bkc, int *block, int vec = 0, bool split = false,
                      bool tiny = false) {
  if (tile == kSkinny) {  // HBM-bound: FP64 arithmetic in every form
    *block = SK_ROWS;
    // tiny: every task has K <= 8 and N <= 8; vec bit 3: A in aligned k runs of 4
    if (tiny) return reinterpret_cast<const void *>(&k_skinny_q);
    return (vec & 8) ? reinterpret_cast<const void *>(&k_skinny<true>)
                     : reinterpret_cast<const void *>(&k_skinny<false>);
  }
  if (tile == kSmall) {  // FP64 on the VALU, FP32 / BF16x3 on the matrix cores
    *block = SM_T;
    if (prec == ITX_PREC_F32) {
      if (akc) return bkc ? reinterpret_cast<const void *>(&k_small_f32<true, true>)
                          : reinterpret_cast<const void *>(&k_small_f32<true, false>);
      return bkc ? reinterpret_cast<const void *>(&k_small_f32<false, true>)
                 : reinterpret_cast<const void *>(&k_small_f32<false, false>);
    }
    if (prec == ITX_PREC_BF16X3) {
      if (akc) return bkc ? reinterpret_cast<const void *>(&k_small_bf16x3<true, true>)
                          : reinterpret_cast<const void *>(&k_small_bf16x3<true, false>);
      return bkc ? reinterpret_cast<const void *>(&k_small_bf16x3<false, true>)
                 : reinterpret_cast<const void *>(&k_small_bf16x3<false, false>);
    }
    if (akc) return bkc ? reinterpret_cast<const void *>(&k_small<true, true>)
                        : reinterpret_cast<const void *>(&k_small<true, false>);
    return bkc ? reinterpret_cast<const void *>(&k_small<false, true>)
               : reinterpret_cast<const void *>(&k_small<false, false>);
  }
  *block = 64 * kTiles[tile].gm * kTiles[tile].gn;
  // vec: bit 0 = runs of 2 usable (FP64 / FP32), bit 1 = runs of 4 (BF16x3),
  // bit 2 = runs of 4 on 4-element boundaries (BF16x3 pre-split words)
  if (prec == ITX_PREC_BF16X3)
    return gemm3_fn_tile(tile, akc, bkc, split ? (vec & 4) != 0 : (vec & 2) != 0, split);
  return prec == ITX_PREC_F64 ? gemm_fn_tile<double>(tile, akc, bkc, (vec & 1) != 0)
                              : gemm_fn_tile<float>(tile, akc, bkc, (vec & 1) != 0);
}

// Resident workgroups per CU of one kernel instantiation (registers and
// LDS), from the runtime's occupancy calculator; cached per variant.
int tile_occupancy(int tile, int prec, bool akc, bool bkc) {
  static std::map<int, int> cache;
  if (tile == kSkinny) return kSkinnyOcc;
  const int key = tile * 16 + prec * 4 + (akc ? 2 : 0) + (bkc ? 1 : 0);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int block = 0, nb = 0;
  const void *f = kernel_of(tile, prec, akc, bkc, &block, (tile == kDgemm || tile == kDgemm2) ? 1 : 0);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, block, 0) != hipSuccess || nb < 1) {
    (void)hipGetLastError();
    nb = 1;
  }
  cache[key] = nb;
  return nb;
}

// Makespan of units (costs) on `slots` identical workers, heaviest first
// onto the least loaded worker (what the persistent loop approximates).
double lpt_makespan(std::vector<double> costs, int64_t slots) {
  if (costs.empty()) return 0;
  if (int64_t(costs.size()) <= slots) return *std::max_element(costs.begin(), costs.end());
  if (int64_t(costs.size()) > 4 * slots) {
    // many more units than slots: LPT ends within one unit of the load
    // bound, which is what the comparison between tiles needs
    double sum = 0, mx = 0;
    for (double c : costs) {
      sum += c;
      mx = std::max(mx, c);
    }
    return std::max(mx, sum / double(slots) + 0.5 * mx);
  }
  std::sort(costs.begin(), costs.end(), std::greater<double>());
  std::priority_queue<double, std::vector<double>, std::greater<double>> q;
  for (int64_t i = 0; i < slots; ++i) q.push(0.0);
  double mk = 0;
  for (double c : costs) {
    double t = q.top() + c;
    q.pop();
    q.push(t);
    mk = std::max(mk, t);
  }
  return mk;
}

// Size class of a contraction for the tuning table: precision, operand
// contiguity, and log2 buckets of the largest C block, summed K, #blocks.
std::string tune_key(int prec, bool akc, bool bkc, int maxM, int maxN, double maxKsum, int ngroups) {
  auto lg = [](double x) { return int(std::ceil(std::log2(std::max(1.0, x)))); };
  char b[96];
  std::snprintf(b, sizeof b, "p%d a%d b%d M%d N%d K%d G%d", prec, int(akc), int(bkc), lg(maxM), lg(maxN),
                lg(maxKsum), lg(ngroups));
  return b;
}

}  // namespace

int build_plan(itx_ctx *ctx, int rA, const int32_t *labA, int rB, const int32_t *labB, int rC,
               const int32_t *labC, int ntasks, const int64_t *offs, const int64_t *blkA,
               const int64_t *blkB, const int64_t *blkC, int prec, itx_plan **out, void *into,
               size_t into_cap) {
  *out = nullptr;
  if (prec != ITX_PREC_F64 && prec != ITX_PREC_F32 && prec != ITX_PREC_BF16X3)
    return set_err(ctx, ITX_ERR_ARG, "contraction: unknown precision form");
  if (ntasks < 0) return set_err(ctx, ITX_ERR_ARG, "contraction: negative task count");
  LabelMap lm;
  int rc = map_labels(ctx, rA, labA, rB, labB, rC, labC, lm);
  if (rc) return rc;

  std::vector<GemmTask> tasks;
  TabVec tab;
  ColumnCache ccache;
  tasks.reserve(ntasks);
  // operand extents (elements) for the autotuner's scratch buffers
  int64_t extA = 1, extB = 1, extC = 1;
  // A is read along k if its unit-stride digit is a contracted one (else
  // along m); likewise B.  Decided on the first non-empty task.
  bool akc = false, bkc = false, kc_set = false;

  double flops = 0;
  // Everything a task needs besides its offsets follows from the block
  // dims of its A and B blocks; QN task lists repeat few distinct dims, so
  // the analysis (GEMM digits, offset tables, extents) is done once per
  // distinct (dims A, dims B) and looked up for the other tasks.
  struct DimInfo {
    int64_t dC[ITX_MAX_RANK];
    int32_t M, N, K;
    int64_t tAM, tCM, tBN, tCN, tAK, tBK;
    int64_t am, cm, bn, cn, ak, bk;
    bool akc, bkc, empty;
    int vec;  // bit 0: k runs of 2 usable by both operands, bit 1: runs of 4 (bits 2, 3: below)
  };
  std::vector<DimInfo> infos;
  std::vector<int64_t> ikeys;  // the dims of each entry (rA + rB per entry)
  std::unordered_map<uint64_t, int> icache;
  const int rAB = rA + rB;
  auto analyze = [&](const int64_t *dA, const int64_t *dB, DimInfo &fi) -> int {
    int64_t sA[ITX_MAX_RANK], sB[ITX_MAX_RANK], sC[ITX_MAX_RANK];
    int64_t *dC = fi.dC;
    for (int i = 0; i < rA; ++i) {
      if (dA[i] < 0) return set_err(ctx, ITX_ERR_ARG, "contraction: negative dimension");
      if (lm.a2b[i] >= 0 && dB[lm.a2b[i]] != dA[i])
        return set_err(ctx, ITX_ERR_ARG, "contraction: contracted dimensions differ");
      if (lm.a2c[i] >= 0) dC[lm.a2c[i]] = dA[i];
    }
    for (int j = 0; j < rB; ++j) {
      if (dB[j] < 0) return set_err(ctx, ITX_ERR_ARG, "contraction: negative dimension");
      if (lm.b2c[j] >= 0) dC[lm.b2c[j]] = dB[j];
    }
    colmajor_strides(rA, dA, sA);
    colmajor_strides(rB, dB, sB);
    colmajor_strides(rC, dC, sC);
    Digits dm, dn, dk;
    for (int i = 0; i < rA; ++i) {
      if (lm.a2c[i] >= 0) dm.push_back({dA[i], sA[i], sC[lm.a2c[i]]});
      else dk.push_back({dA[i], sA[i], sB[lm.a2b[i]]});
    }
    for (int j = 0; j < rB; ++j)
      if (lm.b2c[j] >= 0) dn.push_back({dB[j], sB[j], sC[lm.b2c[j]]});
    simplify(dm);
    simplify(dn);
    simplify(dk);
    const int64_t M = digits_size(dm), N = digits_size(dn), K = digits_size(dk);
    if (M > INT32_MAX || N > INT32_MAX || K > INT32_MAX)
      return set_err(ctx, ITX_ERR_UNSUPPORTED, "contraction: GEMM extent exceeds 2^31");
    fi.M = int32_t(M);
    fi.N = int32_t(N);
    fi.K = int32_t(K);
    fi.empty = (M == 0 || N == 0);  // empty C block: nothing to write
    if (fi.empty) return ITX_OK;
    {
      int64_t na = 1, nb = 1, nc = 1;
      for (int i = 0; i < rA; ++i) na *= dA[i];
      for (int j = 0; j < rB; ++j) nb *= dB[j];
      for (int c = 0; c < rC; ++c) nc *= dC[c];
      if (na > INT32_MAX || nb > INT32_MAX || nc > INT32_MAX)
        return set_err(ctx, ITX_ERR_UNSUPPORTED, "contraction: a block of 2^31 or more elements");
    }
    fi.akc = !dk.empty() && dk[0].s1 == 1;
    fi.bkc = !dk.empty() && dk[0].s2 == 1;
    fi.tAM = ccache.emit(dm, true, tab);
    fi.tCM = ccache.emit(dm, false, tab);
    fi.tBN = ccache.emit(dn, true, tab);
    fi.tCN = ccache.emit(dn, false, tab);
    fi.tAK = ccache.emit(dk, true, tab);
    fi.tBK = ccache.emit(dk, false, tab);
    digits_max(dm, fi.am, fi.cm);
    digits_max(dn, fi.bn, fi.cn);
    digits_max(dk, fi.ak, fi.bk);
    // vector loads: both operands k-contiguous, every aligned group of V
    // consecutive k one run of V consecutive elements starting at an even
    // offset, every row offset even (16-byte aligned given an aligned base)
    fi.vec = 0;
    // Decided from the digits instead of scanning the tables: one column of
    // a digit set (s1 or s2) re-merged where it is contiguous by itself;
    // every offset of that table is a multiple of `al` iff every stride is,
    // and the table comes in runs of V consecutive offsets starting at
    // multiples of `al` iff V divides its size, the fastest digit has
    // stride 1 and a multiple of V entries, and the other strides are
    // multiples of `al` (checked against the table scan on random digit
    // sets, both columns, V = 2, 4 and al = 2, 4)
    auto column = [](const Digits &d, bool first, Digit *c) {
      int nc = 0;
      for (auto &x : d) {
        if (x.dim == 1) continue;
        const int64_t st = first ? x.s1 : x.s2;
        if (nc > 0 && c[nc - 1].s1 * c[nc - 1].dim == st)
          c[nc - 1].dim *= x.dim;
        else
          c[nc++] = Digit{x.dim, st, 0};
      }
      return nc;
    };
    auto aligned = [&](const Digits &d, bool first, int al) {
      Digit c[ITX_MAX_RANK];
      const int nc = column(d, first, c);
      for (int i = 0; i < nc; ++i)
        if (c[i].s1 % al) return false;
      return true;
    };
    auto runs = [&](const Digits &d, bool first, int64_t n, int V, int al) {
      if (n % V) return false;
      Digit c[ITX_MAX_RANK];
      const int nc = column(d, first, c);
      if (nc == 0 || c[0].s1 != 1 || c[0].dim % V) return false;
      for (int i = 1; i < nc; ++i)
        if (c[i].s1 % al) return false;
      return true;
    };
    // bit 3: A alone in runs of 4 at even offsets (the skinny path reads
    // only A from memory)
    if (fi.akc && aligned(dm, true, 2) && runs(dk, true, K, 4, 2)) fi.vec |= 8;
    if (fi.akc && fi.bkc) {
      if (aligned(dm, true, 2) && aligned(dn, true, 2)) {
        if (runs(dk, true, K, 2, 2) && runs(dk, false, K, 2, 2)) fi.vec |= 1;
        if (runs(dk, true, K, 4, 2) && runs(dk, false, K, 4, 2)) fi.vec |= 2;
      }
      if (aligned(dm, true, 4) && aligned(dn, true, 4) && runs(dk, true, K, 4, 4) && runs(dk, false, K, 4, 4))
        fi.vec |= 4;
    }
    return ITX_OK;
  };
  int vecbits = 15;  // AND over the tasks
  for (int t = 0; t < ntasks; ++t) {
    const int64_t *dA = blkA + size_t(t) * rA;
    const int64_t *dB = blkB + size_t(t) * rB;
    uint64_t h = 1469598103934665603ull;  // FNV-1a over the dims
    for (int i = 0; i < rA; ++i) h = (h ^ uint64_t(dA[i])) * 1099511628211ull;
    for (int j = 0; j < rB; ++j) h = (h ^ uint64_t(dB[j])) * 1099511628211ull;
    const DimInfo *fp = nullptr;
    auto it = icache.find(h);
    if (it != icache.end()) {
      const int64_t *kd = ikeys.data() + size_t(it->second) * rAB;
      if (std::equal(dA, dA + rA, kd) && std::equal(dB, dB + rB, kd + rA)) fp = &infos[size_t(it->second)];
    }
    DimInfo tmp;
    if (!fp) {
      int rc2 = analyze(dA, dB, tmp);
      if (rc2) return rc2;
      if (it == icache.end()) {  // first dims with this hash: remember them
        icache.emplace(h, int(infos.size()));
        ikeys.insert(ikeys.end(), dA, dA + rA);
        ikeys.insert(ikeys.end(), dB, dB + rB);
        infos.push_back(tmp);
        fp = &infos.back();
      } else {
        fp = &tmp;
      }
    }
    const DimInfo &fi = *fp;
    if (blkC) {
      for (int c = 0; c < rC; ++c)
        if (blkC[size_t(t) * rC + c] != fi.dC[c])
          return set_err(ctx, ITX_ERR_ARG, "contraction: dimensions of C do not match A and B");
    }
    if (fi.empty) continue;
    if (!kc_set) {
      kc_set = true;
      akc = fi.akc;
      bkc = fi.bkc;
    }
    vecbits &= fi.vec;
    if (offs[3 * size_t(t)] % 2) vecbits = 0;
    if (offs[3 * size_t(t) + 1] % 2) vecbits &= 8;
    if (offs[3 * size_t(t)] % 4 || offs[3 * size_t(t) + 1] % 4) vecbits &= ~4;
    GemmTask g{};
    g.a_off = offs[3 * size_t(t)];
    g.b_off = offs[3 * size_t(t) + 1];
    g.c_off = offs[3 * size_t(t) + 2];
    g.M = fi.M;
    g.N = fi.N;
    g.K = fi.K;
    g.tAM = fi.tAM;
    g.tCM = fi.tCM;
    g.tBN = fi.tBN;
    g.tCN = fi.tCN;
    g.tAK = fi.tAK;
    g.tBK = fi.tBK;
    extA = std::max(extA, g.a_off + fi.am + fi.ak + 1);
    extB = std::max(extB, g.b_off + fi.bk + fi.bn + 1);
    extC = std::max(extC, g.c_off + fi.cm + fi.cn + 1);
    flops += 2.0 * double(fi.M) * double(fi.N) * double(fi.K);
    tasks.push_back(g);
  }

  // group tasks by C block, in C offset order and stable within a block
  // (keeps the planner's pair order): a counting sort over the distinct
  // C offsets
  std::vector<GemmTask> st(tasks.size());
  {
    // C offset -> bucket (first-appearance numbering) through an
    // open-addressing table, then buckets renumbered in offset order
    size_t cap = 16;
    while (cap < 2 * tasks.size()) cap <<= 1;
    std::vector<int64_t> hk(cap, -1);
    std::vector<int> hv(cap, 0), bucket(tasks.size());
    std::vector<int64_t> first;  // distinct C offsets, first-appearance order
    for (size_t i = 0; i < tasks.size(); ++i) {
      const int64_t c = tasks[i].c_off;
      size_t h = size_t((uint64_t(c) * 0x9E3779B97F4A7C15ull) >> 20) & (cap - 1);
      while (hk[h] != -1 && hk[h] != c) h = (h + 1) & (cap - 1);
      if (hk[h] == -1) {
        hk[h] = c;
        hv[h] = int(first.size());
        first.push_back(c);
      }
      bucket[i] = hv[h];
    }
    std::vector<int> ord(first.size()), rank(first.size());
    std::iota(ord.begin(), ord.end(), 0);
    std::sort(ord.begin(), ord.end(), [&](int x, int y) { return first[size_t(x)] < first[size_t(y)]; });
    for (size_t r = 0; r < ord.size(); ++r) rank[size_t(ord[r])] = int(r);
    std::vector<int> cnt(first.size() + 1, 0);
    for (size_t i = 0; i < tasks.size(); ++i) {
      bucket[i] = rank[size_t(bucket[i])];
      ++cnt[size_t(bucket[i]) + 1];
    }
    for (size_t b = 1; b < cnt.size(); ++b) cnt[b] += cnt[b - 1];
    for (size_t i = 0; i < tasks.size(); ++i) st[size_t(cnt[size_t(bucket[i])]++)] = tasks[i];
  }
  struct Grp {
    int tb, te;
    int M, N;
    double ksum;
  };
  std::vector<Grp> groups;
  for (size_t i = 0; i < st.size();) {
    size_t e = i + 1;
    while (e < st.size() && st[e].c_off == st[i].c_off) {
      if (st[e].M != st[i].M || st[e].N != st[i].N)
        return set_err(ctx, ITX_ERR_ARG, "contraction: tasks writing one C block disagree on its shape");
      ++e;
    }
    double ks = 0;
    for (size_t q = i; q < e; ++q) ks += st[q].K;
    groups.push_back({int(i), int(e), st[i].M, st[i].N, ks});
    i = e;
  }

  const int cus = std::max(1, ctx->num_cus);
  int maxN = 0, maxK = 0, maxM = 0;
  double maxKsum = 0;
  for (auto &t : st) {
    maxN = std::max(maxN, t.N);
    maxK = std::max(maxK, t.K);
    maxM = std::max(maxM, t.M);
  }
  for (auto &g : groups) maxKsum = std::max(maxKsum, g.ksum);

  // Build the device plan for one tile choice.
  // the register-pipelined FP64 tile needs k-contiguous operands in aligned pairs
  const bool dg_ok = (prec == ITX_PREC_F64 || prec == ITX_PREC_F32) && akc && bkc && (vecbits & 1);
  // skinny plans whose tasks all have K, N <= 8 run k_skinny_q: two rows per thread
  const bool tinyq = maxN <= SKQ_N && maxK <= SKQ_K;
  auto finalize = [&](int best, itx_plan **pout) -> int {
    *pout = nullptr;
    if ((best == kDgemm || best == kDgemm2) && !dg_ok) return ITX_ERR_UNSUPPORTED;
    const int tbm = (best == kSkinny && tinyq) ? SKQ_R * SK_ROWS : kTiles[best].bm;
    const int tbn = (best == kSkinny) ? std::max(1, maxN) : kTiles[best].bn;
    std::vector<WorkUnit> units, red;
    std::vector<double> ucost;
    std::vector<GemmTask> vt;  // device task list (st, or its k-chunks when split)
    int nslots = 0;
    bool split = false;
    if (best == kSmall) {
      // Split the k-chunk sequence of heavy tiles over several units when
      // the tiles alone cannot fill the device (few C blocks, long task
      // lists): chunk budget c per unit from the total work.
      double total = 0, maxw = 0;
      for (auto &g : groups) {
        double w = 0;
        for (int t = g.tb; t < g.te; ++t) w += double((st[t].K + SM_K - 1) / SM_K);
        const double tiles = double((g.M + tbm - 1) / tbm) * double((g.N + tbn - 1) / tbn);
        total += w * tiles;
        maxw = std::max(maxw, w);
      }
      const double target = 2.0 * cus * tile_occupancy(kSmall, prec, akc, bkc);
      const double c = std::max(1.0, std::ceil(total / target));
      split = maxw > 2.0 * c && total * (SM_M * SM_N) * 8.0 / c < double(size_t(1) << 30);
      if (split) {
        // at least two k-chunks per unit: a one-chunk unit writes a
        // partial tile for as many bytes as it reads
        const int per = std::max(2, int(c));
        for (auto &g : groups) {
          const int vb = int(vt.size());
          for (int t = g.tb; t < g.te; ++t)
            for (int k0 = 0; k0 < st[t].K; k0 += SM_K) {
              GemmTask x = st[t];
              x.K = std::min(SM_K, st[t].K - k0);
              x.tAK += k0;
              x.tBK += k0;
              vt.push_back(x);
            }
          const int ve = int(vt.size());
          for (int m0 = 0; m0 < g.M; m0 += tbm)
            for (int n0 = 0; n0 < g.N; n0 += tbn) {
              const int s0 = nslots;
              for (int a = vb; a < ve; a += per) {
                units.push_back({a, std::min(ve, a + per), m0, n0, nslots++, 0});
                ucost.push_back(double(std::min(ve, a + per) - a));
              }
              red.push_back({vb, s0, m0, n0, nslots, 0});
            }
        }
      }
    }
    if (!split) {
      vt = st;
      // Tiles of a C block in super-tiles of SUPM x SUPN: the 32 workgroups
      // an XCD runs side by side then share 4 row blocks of A and 8 column
      // blocks of B in that XCD's L2 (m-major order shares one row block
      // with 32 column blocks, re-reading B once per row block).
      const int SUPM = 4, SUPN = 8;
      for (auto &g : groups) {
        const int tm = (g.M + tbm - 1) / tbm, tn = (g.N + tbn - 1) / tbn;
        for (int sm = 0; sm < tm; sm += SUPM)
          for (int sn = 0; sn < tn; sn += SUPN)
            for (int im = sm; im < std::min(tm, sm + SUPM); ++im)
              for (int in = sn; in < std::min(tn, sn + SUPN); ++in) {
                units.push_back({g.tb, g.te, im * tbm, in * tbn, -1, 0});
                ucost.push_back(g.ksum);
              }
      }
    }
    // heaviest units first (the persistent loop then balances like LPT);
    // stable so equal-cost tiles keep their m-major order
    {
      std::vector<std::pair<double, int>> ui(units.size());
      bool sorted = true;
      for (size_t i = 0; i < units.size(); ++i) {
        ui[i] = {-ucost[i], int(i)};  // (cost descending, position ascending): the stable order
        sorted = sorted && (i == 0 || ucost[i] <= ucost[i - 1]);
      }
      if (!sorted) {
        std::sort(ui.begin(), ui.end());
        std::vector<WorkUnit> su(units.size());
        for (size_t i = 0; i < ui.size(); ++i) su[i] = units[size_t(ui[i].second)];
        units.swap(su);
      }
    }
    auto *p = new itx_plan();
    p->prec = prec;
    p->tile = best;
    p->a_kcontig = akc;
    p->b_kcontig = bkc;
    p->ntasks = int(vt.size());
    p->nunits = int(units.size());
    p->flops = flops;
    p->grid = int(std::min<int64_t>(int64_t(units.size()), int64_t(cus) * tile_occupancy(best, prec, akc, bkc)));
    p->vec = (akc && bkc) ? vecbits : akc ? (vecbits & 8) : 0;
    const bool tiny = best == kSkinny && tinyq;
    p->fn = kernel_of(best, prec, akc, bkc, &p->block, p->vec, false, tiny);
    p->fn_scalar = kernel_of(best, prec, akc, bkc, &p->block, 0, false, tiny);
    p->nred = int(red.size());
    if (p->nred) {
      p->fn_red = reinterpret_cast<const void *>(&k_reduce_tiles);
      p->grid_red = int(std::min<int64_t>(int64_t(red.size()) * (SM_M * SM_N / SM_T), int64_t(cus) * 8));
    }
    if (std::getenv("ITX_PLAN_LOG"))
      std::fprintf(stderr, "[itx plan] tasks=%d units=%d tile=%dx%d%s grid=%d akc=%d bkc=%d flops=%.3e split=%d\n",
                   p->ntasks, p->nunits, tbm, tbn, best == kSkinny ? "(skinny)" : best == kSmall ? "(small)" : "",
                   p->grid, int(akc), int(bkc), flops, p->nred);
    if (p->nunits > 0) {
      size_t bt_bytes = vt.size() * sizeof(GemmTask);
      size_t bu_bytes = units.size() * sizeof(WorkUnit);
      size_t br_bytes = red.size() * sizeof(WorkUnit);
      size_t o_units = (bt_bytes + 255) & ~size_t(255);
      size_t o_red = (o_units + bu_bytes + 255) & ~size_t(255);
      size_t o_tab = (o_red + br_bytes + 255) & ~size_t(255);
      size_t o_ws = (o_tab + tab.size() * sizeof(int64_t) + 255) & ~size_t(255);
      size_t total = o_ws + sizeof(double) * size_t(nslots) * (SM_M * SM_N);
      // the device image: tasks | units | reduction units | tables, each
      // 256-byte aligned, assembled in the staging arena
      const UploadPart parts[4] = {{vt.data(), 0, bt_bytes},
                                   {units.data(), o_units, bu_bytes},
                                   {red.data(), o_red, br_bytes},
                                   {tab.data(), o_tab, tab.size() * sizeof(int64_t)}};
      int rc3 = ITX_OK;
      if (into) {
        if (total > into_cap) {
          delete p;
          return set_err(ctx, ITX_ERR_UNSUPPORTED,
                         "contraction: the plan needs " + std::to_string(total) + " bytes of device memory, more than the " +
                             std::to_string(into_cap) + "-byte scratch of the allocation-free entry point");
        }
        p->dev = into;
        p->own_dev = false;
      } else {
        rc3 = pool_alloc(ctx, total, &p->dev);
        if (rc3) {
          delete p;
          return rc3;
        }
      }
      p->ctx = ctx;
      rc3 = stage_upload_parts(ctx, p->dev, parts, 4, o_ws);
      if (rc3) {
        if (p->own_dev) pool_free(ctx, p->dev);
        delete p;
        return rc3;
      }
      p->dev_bytes = total;
      p->d_tasks = reinterpret_cast<const GemmTask *>(p->dev);
      p->d_units = reinterpret_cast<const WorkUnit *>(static_cast<char *>(p->dev) + o_units);
      p->d_tab = reinterpret_cast<const int64_t *>(static_cast<char *>(p->dev) + o_tab);
      if (p->nred) {
        p->d_red = reinterpret_cast<const WorkUnit *>(static_cast<char *>(p->dev) + o_red);
        p->ws = reinterpret_cast<double *>(static_cast<char *>(p->dev) + o_ws);
      }
    }
    ctx->plans.insert(p);
    *pout = p;
    return ITX_OK;
  };

  // ---- tile selection ------------------------------------------------
  // Skinny tasks (all N <= 32, K <= 64) take the VALU path.  Otherwise, in
  // order: ITX_FORCE_TILE, the tuning table entry of this size class, a
  // measurement of every MFMA tile when autotuning is on (result stored in
  // the table), else the cost model: units (one per C tile, cost ~ tile
  // area x summed K) scheduled LPT on CUs x occupancy slots.
  int best = 0;
  const std::string key = tune_key(prec, akc, bkc, maxM, maxN, maxKsum, int(groups.size()));
  const char *force = std::getenv("ITX_FORCE_TILE");
  // Narrow tasks (all N <= 32, K <= 64) over large C blocks take the
  // skinny VALU path; C blocks of at most 8192 elements (small QN blocks)
  // the small-block VALU path; everything else an MFMA tile.
  double rows = 0;
  for (auto &g : groups) rows += g.M;
  const double avg_rows = groups.empty() ? 0.0 : rows / double(groups.size());
  const bool narrow = maxN <= SK_N && maxK <= SK_K;
  if (!st.empty() && !force && narrow && avg_rows >= 256) {
    best = kSkinny;
  } else if (!st.empty() && !force && double(maxM) * double(maxN) <= 8192.0) {
    best = kSmall;
  } else if (force) {
    best = std::atoi(force);
    if (best < 0 || best >= kNumTiles || (best == kSkinny && !narrow) || ((best == kDgemm || best == kDgemm2) && !dg_ok))
      best = 0;
  } else if (ctx->tune.count(key)) {
    best = ctx->tune[key];
  } else if (ctx->autotune && !into && !st.empty()) {
    void *bufA = nullptr, *bufB = nullptr, *bufC = nullptr;
    int rc2 = pool_alloc(ctx, sizeof(double) * size_t(extA), &bufA);
    if (!rc2) rc2 = pool_alloc(ctx, sizeof(double) * size_t(extB), &bufB);
    if (!rc2) rc2 = pool_alloc(ctx, sizeof(double) * size_t(extC), &bufC);
    if (rc2) return rc2;
    itx_fill(ctx, static_cast<double *>(bufA), size_t(extA), 0.0);
    itx_fill(ctx, static_cast<double *>(bufB), size_t(extB), 0.0);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    float best_ms = 1e30f;
    for (int ti = 0; ti < kNumTiles; ++ti) {
      if (ti == kSkinny || ti == kSmall) continue;
      itx_plan *pc = nullptr;
      if (finalize(ti, &pc)) continue;
      const double *A = static_cast<const double *>(bufA), *B = static_cast<const double *>(bufB);
      double *Cp = static_cast<double *>(bufC);
      launch_plan(ctx, pc, A, B, Cp, 1.0, 0.0);
      (void)hipEventRecord(e0, ctx->stream);
      for (int r = 0; r < 3; ++r) launch_plan(ctx, pc, A, B, Cp, 1.0, 0.0);
      (void)hipEventRecord(e1, ctx->stream);
      (void)hipEventSynchronize(e1);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms < best_ms) {
        best_ms = ms;
        best = ti;
      }
      itx_plan_destroy(ctx, pc);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    pool_free(ctx, bufA);
    pool_free(ctx, bufB);
    pool_free(ctx, bufC);
    ctx->tune[key] = best;
  } else {
    double best_t = 1e300;
    // the 32-deep variants (11-13) are left to measurement (autotuner,
    // ITX_FORCE_TILE): they only pay off for long K on large tiles
    for (int ti : {0, 1, 2, 3, 4, 6, 7, 8, 9, kDgemm}) {
      if (ti == kDgemm && !dg_ok) continue;
      const TileCfg ts = kTiles[ti];
      std::vector<double> costs;
      for (auto &g : groups) {
        const int64_t tiles = int64_t((g.M + ts.bm - 1) / ts.bm) * ((g.N + ts.bn - 1) / ts.bn);
        const double c = double(ts.bm) * ts.bn * (g.ksum + BK * (g.te - g.tb) + 8);
        for (int64_t q = 0; q < tiles; ++q) costs.push_back(c);
      }
      if (costs.empty()) continue;
      // workgroups resident per CU: for the reduced forms the tile's
      // occupancy, or fewer when the units do not fill that many slots (a CU
      // running one workgroup does not run it `occupancy` times slower;
      // without this the FP32 form picked a 64 x 256 tile for 150 units of
      // the chi = 1024 QN chain, 86 us against 40 us); the FP64 efficiencies
      // were calibrated against the plain occupancy
      const int occ_max = tile_occupancy(ti, prec, akc, bkc);
      const int occ = prec == ITX_PREC_F64
                          ? occ_max
                          : int(std::max<int64_t>(1, std::min<int64_t>(occ_max, (int64_t(costs.size()) + cus - 1) / cus)));
      const double eff = (prec == ITX_PREC_F64) ? ts.eff64 : ts.eff32;
      const double tt = lpt_makespan(costs, int64_t(cus) * occ) * occ / eff;
      if (tt < best_t * 0.97) {
        best_t = tt;
        best = ti;
      }
    }
  }
  return finalize(best, out);
}

const void *plan_fn(const itx_plan *p, const double *A, const double *B) {
  // the vector-load kernel reads 16-byte pairs (or quads of split words):
  // its operands must be 16-byte aligned (caching-pool blocks are 256-byte
  // aligned)
  const bool aligned = ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(B)) & 15) == 0;
  return (p->vec && aligned) ? p->fn : p->fn_scalar;
}

KernelArgs plan_args(const itx_plan *p, int which, const double *A, const double *B, double *C, double alpha,
                     double beta) {
  if (which == 0)  // the contraction (into the workspace when split)
    return KernelArgs{A, B, p->nred ? p->ws : C, p->d_tasks, p->d_units, p->nunits, p->d_tab, alpha, beta};
  return KernelArgs{p->ws, nullptr, C, p->d_tasks, p->d_red, p->nred, p->d_tab, alpha, beta};
}

int launch_plan(itx_ctx *ctx, const itx_plan *p, const double *A, const double *B, double *C, double alpha,
                double beta) {
  if (!p) return set_err(ctx, ITX_ERR_ARG, "null plan");
  if (p->nunits == 0) return ITX_OK;
  void *args[KernelArgs::kCount];
  KernelArgs ka = plan_args(p, 0, A, B, C, alpha, beta);
  ka.pointers(args);
  ITX_CHECK_HIP(ctx, hipLaunchKernel(plan_fn(p, A, B), dim3(p->grid), dim3(p->block), args, 0, ctx->stream));
  if (p->nred) {
    KernelArgs kr = plan_args(p, 1, A, B, C, alpha, beta);
    kr.pointers(args);
    ITX_CHECK_HIP(ctx, hipLaunchKernel(p->fn_red, dim3(p->grid_red), dim3(SM_T_RED), args, 0, ctx->stream));
  }
  return ITX_OK;
}

namespace {

std::string plan_key(const char *kind, int rA, const int32_t *labA, int rB, const int32_t *labB, int rC,
                     const int32_t *labC, int ntasks, const int64_t *offs, const int64_t *blkA,
                     const int64_t *blkB, int prec) {
  std::string k;
  k.reserve(64 + 8 * (rA + rB + rC) + size_t(ntasks) * 8 * (3 + rA + rB));
  auto put = [&](const void *d, size_t n) { k.append(static_cast<const char *>(d), n); };
  put(kind, std::strlen(kind));
  put(&rA, sizeof rA);
  put(&rB, sizeof rB);
  put(&rC, sizeof rC);
  put(&prec, sizeof prec);
  put(&ntasks, sizeof ntasks);
  put(labA, sizeof(int32_t) * rA);
  put(labB, sizeof(int32_t) * rB);
  put(labC, sizeof(int32_t) * rC);
  put(offs, sizeof(int64_t) * 3 * size_t(ntasks));
  put(blkA, sizeof(int64_t) * rA * size_t(ntasks));
  put(blkB, sizeof(int64_t) * rB * size_t(ntasks));
  return k;
}

constexpr size_t kPlanCacheMax = 96;

int cached_plan(itx_ctx *ctx, const std::string &key, itx_plan **out,
                const std::function<int(itx_plan **)> &make) {
  auto it = ctx->plan_cache.find(key);
  if (it != ctx->plan_cache.end()) {
    *out = it->second;
    return ITX_OK;
  }
  itx_plan *p = nullptr;
  int rc = make(&p);
  if (rc) return rc;
  if (ctx->plan_cache.size() >= kPlanCacheMax) {
    // evict the oldest entry (FIFO); its device block returns to the
    // stream-ordered pool, so kernels already enqueued still read it intact
    const std::string old = ctx->plan_lru.front();
    ctx->plan_lru.erase(ctx->plan_lru.begin());
    auto jt = ctx->plan_cache.find(old);
    if (jt != ctx->plan_cache.end()) {
      itx_plan_destroy(ctx, jt->second);
      ctx->plan_cache.erase(jt);
    }
  }
  ctx->plan_cache[key] = p;
  ctx->plan_lru.push_back(key);
  *out = p;
  return ITX_OK;
}

}  // namespace
}  // namespace itx

extern "C" {

int itx_plan_dense(itx_ctx *ctx, int rA, const int64_t *dimsA, const int32_t *labA, int rB,
                   const int64_t *dimsB, const int32_t *labB, int rC, const int64_t *dimsC,
                   const int32_t *labC, int prec, itx_plan **out) {
  if (!ctx || !out) return ITX_ERR_ARG;
  const int64_t offs[3] = {0, 0, 0};
  return itx::build_plan(ctx, rA, labA, rB, labB, rC, labC, 1, offs, dimsA, dimsB, dimsC, prec, out);
}

int itx_plan_qn(itx_ctx *ctx, int rA, const int32_t *labA, int rB, const int32_t *labB, int rC,
                const int32_t *labC, int ntasks, const int64_t *offs, const int64_t *blkdimsA,
                const int64_t *blkdimsB, int prec, itx_plan **out) {
  if (!ctx || !out) return ITX_ERR_ARG;
  return itx::build_plan(ctx, rA, labA, rB, labB, rC, labC, ntasks, offs, blkdimsA, blkdimsB, nullptr, prec,
                         out);
}

int itx_plan_execute(itx_ctx *ctx, const itx_plan *plan, const double *A, const double *B, double *C,
                     double alpha, double beta) {
  if (!ctx) return ITX_ERR_ARG;
  return itx::launch_plan(ctx, plan, A, B, C, alpha, beta);
}

double itx_plan_flops(const itx_plan *plan) { return plan ? plan->flops : 0.0; }

int itx_split_bf16(itx_ctx *ctx, const double *x, uint32_t *y, size_t n) {
  if (!ctx || (n && (!x || !y))) return ITX_ERR_ARG;
  if (n == 0) return ITX_OK;
  const unsigned grid = unsigned(std::min<size_t>((n + 255) / 256, size_t(ctx->num_cus) * 16));
  hipLaunchKernelGGL(itx::k_split_bf16, dim3(grid), dim3(256), 0, ctx->stream, x, y, n);
  ITX_CHECK_HIP(ctx, hipGetLastError());
  return ITX_OK;
}

int itx_plan_set_split(itx_ctx *ctx, itx_plan *plan, int on) {
  if (!ctx || !plan) return ITX_ERR_ARG;
  const bool want = on != 0;
  if (want == bool(plan->split)) return ITX_OK;
  if (want && (plan->prec != ITX_PREC_BF16X3 || plan->tile == itx::kSkinnyTile || plan->tile == 10 ||
               !plan->a_kcontig || !plan->b_kcontig || plan->nred))
    return itx::set_err(ctx, ITX_ERR_UNSUPPORTED,
                        "itx_plan_set_split: pre-split operands need a BF16x3 MFMA plan with both operands "
                        "contiguous along the contracted index");
  plan->split = want ? 1 : 0;
  if (plan->nunits > 0) {
    const bool akc = plan->a_kcontig, bkc = plan->b_kcontig;
    if (want) {
      plan->fn = itx::kernel_of(plan->tile, plan->prec, akc, bkc, &plan->block, plan->vec & 4 ? 4 : 0, true);
      plan->fn_scalar = itx::kernel_of(plan->tile, plan->prec, akc, bkc, &plan->block, 0, true);
    } else {
      plan->fn = itx::kernel_of(plan->tile, plan->prec, akc, bkc, &plan->block, plan->vec);
      plan->fn_scalar = itx::kernel_of(plan->tile, plan->prec, akc, bkc, &plan->block, 0);
    }
  }
  return ITX_OK;
}

int itx_plan_info(const itx_plan *plan, int *tile, int *form, int *ntasks, int *nunits) {
  if (!plan) return ITX_ERR_ARG;
  // the skinny VALU path is HBM-bound and computes in FP64 in every form
  const int f = plan->tile == itx::kSkinnyTile ? ITX_PREC_F64 : plan->prec;
  if (tile) *tile = plan->tile;
  if (form) *form = f;
  if (ntasks) *ntasks = plan->ntasks;
  if (nunits) *nunits = plan->nunits;
  return ITX_OK;
}

int itx_plan_destroy(itx_ctx *ctx, itx_plan *plan) {
  if (!plan) return ITX_OK;
  // the device block goes back to the context's stream-ordered pool (no
  // wait); without a context it stays with the pool until itx_destroy
  if (ctx) {
    if (plan->dev && plan->own_dev) itx::pool_free(ctx, plan->dev);
    ctx->plans.erase(plan);
  }
  delete plan;
  return ITX_OK;
}

int itx_chain_create(itx_ctx *ctx, int nstages, itx_plan *const *plans, itx_chain **out) {
  if (!ctx || !out || nstages <= 0 || !plans) return ITX_ERR_ARG;
  *out = nullptr;
  for (int i = 0; i < nstages; ++i)
    if (!plans[i]) return itx::set_err(ctx, ITX_ERR_ARG, "itx_chain_create: null plan");
  auto *ch = new itx_chain();
  ch->plans.assign(plans, plans + nstages);
  ch->args.resize(2 * size_t(nstages));
  ch->nodes.assign(2 * size_t(nstages), nullptr);
  ch->funcs.assign(2 * size_t(nstages), nullptr);
  ctx->chains.insert(ch);
  *out = ch;
  return ITX_OK;
}

int itx_chain_launch(itx_ctx *ctx, itx_chain *ch, const double *const *A, const double *const *B,
                     double *const *C, const double *alpha, const double *beta) {
  if (!ctx || !ch || !A || !B || !C) return ITX_ERR_ARG;
  const int n = int(ch->plans.size());
  // node 2i: the contraction of stage i; node 2i+1: its reduction pass
  auto args_of = [&](int x) {
    const int i = x / 2;
    return itx::plan_args(ch->plans[size_t(i)], x % 2, A[i], B[i], C[i], alpha ? alpha[i] : 1.0,
                          beta ? beta[i] : 0.0);
  };
  auto active = [&](int x) {
    const itx_plan *p = ch->plans[size_t(x / 2)];
    return p->nunits > 0 && (x % 2 == 0 || p->nred > 0);
  };
  auto fn_of = [&](int x) {
    const itx_plan *p = ch->plans[size_t(x / 2)];
    return x % 2 ? p->fn_red : itx::plan_fn(p, A[x / 2], B[x / 2]);
  };
  if (ch->exec) {
    // a changed operand alignment changes a stage's kernel: rebuild
    bool same = true;
    for (int x = 0; x < 2 * n && same; ++x)
      if (ch->nodes[size_t(x)] && fn_of(x) != ch->funcs[size_t(x)]) same = false;
    if (!same) {
      hipGraphExecDestroy(ch->exec);
      hipGraphDestroy(ch->graph);
      ch->exec = nullptr;
      ch->graph = nullptr;
      ch->nodes.assign(ch->nodes.size(), nullptr);
    }
  }
  auto params_of = [&](int x, void **argv) {
    hipKernelNodeParams kp{};
    const itx_plan *p = ch->plans[size_t(x / 2)];
    ch->funcs[size_t(x)] = fn_of(x);
    kp.func = const_cast<void *>(ch->funcs[size_t(x)]);
    kp.gridDim = dim3(unsigned(x % 2 ? p->grid_red : p->grid));
    kp.blockDim = dim3(unsigned(x % 2 ? itx::SM_T_RED : p->block));
    kp.sharedMemBytes = 0;
    ch->args[size_t(x)].pointers(argv);
    kp.kernelParams = argv;
    kp.extra = nullptr;
    return kp;
  };
  if (!ch->exec) {
    // first launch: one kernel node per launch, each after the previous
    // one (the stream order of the eager path)
    ITX_CHECK_HIP(ctx, hipGraphCreate(&ch->graph, 0));
    hipGraphNode_t prev = nullptr;
    for (int x = 0; x < 2 * n; ++x) {
      ch->args[size_t(x)] = args_of(x);
      if (!active(x)) continue;
      void *argv[itx::KernelArgs::kCount];
      hipKernelNodeParams kp = params_of(x, argv);
      ITX_CHECK_HIP(ctx, hipGraphAddKernelNode(&ch->nodes[size_t(x)], ch->graph, prev ? &prev : nullptr,
                                               prev ? 1 : 0, &kp));
      prev = ch->nodes[size_t(x)];
    }
    ITX_CHECK_HIP(ctx, hipGraphInstantiate(&ch->exec, ch->graph, nullptr, nullptr, 0));
  } else {
    // replay: patch the nodes whose operands or scalars changed
    for (int x = 0; x < 2 * n; ++x) {
      if (!ch->nodes[size_t(x)]) continue;
      const itx::KernelArgs a = args_of(x);
      itx::KernelArgs &c = ch->args[size_t(x)];
      if (a.A == c.A && a.B == c.B && a.C == c.C && a.alpha == c.alpha && a.beta == c.beta) continue;
      c = a;
      void *argv[itx::KernelArgs::kCount];
      hipKernelNodeParams kp = params_of(x, argv);
      ITX_CHECK_HIP(ctx, hipGraphExecKernelNodeSetParams(ch->exec, ch->nodes[size_t(x)], &kp));
    }
  }
  ITX_CHECK_HIP(ctx, hipGraphLaunch(ch->exec, ctx->stream));
  return ITX_OK;
}

int itx_chain_destroy(itx_ctx *ctx, itx_chain *ch) {
  if (!ch) return ITX_OK;
  if (ctx) {
    hipStreamSynchronize(ctx->stream);
    ctx->chains.erase(ch);
  }
  if (ch->exec) hipGraphExecDestroy(ch->exec);
  if (ch->graph) hipGraphDestroy(ch->graph);
  delete ch;
  return ITX_OK;
}

int itx_contract_dense(itx_ctx *ctx, const double *A, int rA, const int64_t *dimsA, const int32_t *labA,
                       const double *B, int rB, const int64_t *dimsB, const int32_t *labB, double *C, int rC,
                       const int64_t *dimsC, const int32_t *labC, double alpha, double beta, int prec) {
  if (!ctx) return ITX_ERR_ARG;
  const int64_t offs[3] = {0, 0, 0};
  std::string key = itx::plan_key("D", rA, labA, rB, labB, rC, labC, 1, offs, dimsA, dimsB, prec);
  key.append(reinterpret_cast<const char *>(dimsC), sizeof(int64_t) * rC);
  itx_plan *p = nullptr;
  int rc = itx::cached_plan(ctx, key, &p, [&](itx_plan **o) {
    return itx::build_plan(ctx, rA, labA, rB, labB, rC, labC, 1, offs, dimsA, dimsB, dimsC, prec, o);
  });
  if (rc) return rc;
  return itx::launch_plan(ctx, p, A, B, C, alpha, beta);
}

int itx_contract_qn(itx_ctx *ctx, const double *A, int rA, const int32_t *labA, const double *B, int rB,
                    const int32_t *labB, double *C, int rC, const int32_t *labC, int ntasks,
                    const int64_t *offs, const int64_t *blkdimsA, const int64_t *blkdimsB, double alpha,
                    int prec) {
  if (!ctx) return ITX_ERR_ARG;
  std::string key = itx::plan_key("Q", rA, labA, rB, labB, rC, labC, ntasks, offs, blkdimsA, blkdimsB, prec);
  itx_plan *p = nullptr;
  int rc = itx::cached_plan(ctx, key, &p, [&](itx_plan **o) {
    return itx::build_plan(ctx, rA, labA, rB, labB, rC, labC, ntasks, offs, blkdimsA, blkdimsB, nullptr, prec,
                           o);
  });
  if (rc) return rc;
  return itx::launch_plan(ctx, p, A, B, C, alpha, 0.0);
}

int itx_contract_qn_noalloc(itx_ctx *ctx, const double *A, int rA, const int32_t *labA, const double *B, int rB,
                            const int32_t *labB, double *C, int rC, const int32_t *labC, int ntasks,
                            const int64_t *offs, const int64_t *blkdimsA, const int64_t *blkdimsB, double alpha,
                            int prec) {
  if (!ctx) return ITX_ERR_ARG;
  if (!ctx->scratch) return itx::set_err(ctx, ITX_ERR_UNSUPPORTED, "itx_contract_qn_noalloc: no scratch address");
  // the plan lives in host memory for this call only; its device block is
  // the scratch array, which the stream reuses in order for the next call
  itx_plan *p = nullptr;
  int rc = itx::build_plan(ctx, rA, labA, rB, labB, rC, labC, ntasks, offs, blkdimsA, blkdimsB, nullptr, prec, &p,
                           ctx->scratch, ctx->scratch_bytes);
  if (rc) return rc;
  rc = itx::launch_plan(ctx, p, A, B, C, alpha, 0.0);
  itx_plan_destroy(ctx, p);
  return rc;
}

}  // extern "C"

extern "C" {

int itx_set_autotune(itx_ctx *ctx, int on) {
  if (!ctx) return ITX_ERR_ARG;
  ctx->autotune = on ? 1 : 0;
  return ITX_OK;
}

int itx_tune_save(itx_ctx *ctx, const char *path) {
  if (!ctx || !path) return ITX_ERR_ARG;
  std::ofstream f(path);
  if (!f) return itx::set_err(ctx, ITX_ERR_ARG, std::string("itx_tune_save: cannot write ") + path);
  for (auto &kv : ctx->tune) f << kv.first << '|' << kv.second << '\n';
  return ITX_OK;
}

int itx_tune_load(itx_ctx *ctx, const char *path) {
  if (!ctx || !path) return ITX_ERR_ARG;
  std::ifstream f(path);
  if (!f) return itx::set_err(ctx, ITX_ERR_ARG, std::string("itx_tune_load: cannot read ") + path);
  std::string line;
  while (std::getline(f, line)) {
    auto bar = line.find('|');
    if (bar == std::string::npos) continue;
    ctx->tune[line.substr(0, bar)] = std::atoi(line.c_str() + bar + 1);
  }
  return ITX_OK;
}

int itx_tune_size(itx_ctx *ctx) { return ctx ? int(ctx->tune.size()) : 0; }

}  // extern "C"
