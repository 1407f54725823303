// solver.hip -- SpMV and (preconditioned) conjugate gradients on the device.
//
// The reference delegates this to Hypre / PETSc (femutils/HypreDoFLinearSystem.cc:596-769,
// femutils/PetscDoFLinearSystem.cc:427-572): PCG with a relative 2-norm test and a
// Jacobi preconditioner (femutils/PetscDoFLinearSystemFactory.axl:5-10).
//
// One PCG iteration k is three kernels (four for the mixed solver):
//   spmv     q = A p, per-block partials of p.q
//   update   alpha = rho_{k-1} / p.q; x += alpha p, r -= alpha q, z = M r;
//            partials of r.z (parity buffer k&1) and r.r
//  [reliable update, mixed only: r64 = b - A64 x when the FP32 residual has
//            dropped by reliable_delta or reached the tolerance]
//   pupdate  beta = rho_k / rho_{k-1}; convergence test; p = z + beta p
// Reductions never use fences or atomics: a producing kernel writes one
// partial per block, and every block of a consuming kernel re-reduces all
// partials in the same order (deterministic, and identical in every block).
// Values read across kernels are parity-indexed by the host-supplied
// iteration index k, so no kernel reads what another block of the same
// kernel writes.  The mixed variant keeps r, z, p, q, A and M in FP32 and x
// in FP64; only the FP64 residual of a reliable update can declare
// convergence.  Dots are accumulated in FP64.
//
// SpMV formats: CSR with G lanes per row (wavefront-group row reductions) and
// SELL-64 (slices of 64 rows stored column-major, one lane per row, fully
// coalesced value/column streams), scalar or with 3x3 blocks.  The first
// solve on a matrix times the candidates and keeps the fastest.
#include "common.h"
#include "device_utils.h"
#include "linsys.h"

#include <chrono>
#include <type_traits>
#include <vector>
#include <cmath>

namespace afem {

constexpr int kBlock = 256;
constexpr int kMaxBlocks = 1024;

struct State
{
  double rtol2, atol2, delta2, tol2, bb, rr_out, rr_true;
  double rr_ref[2];
  int ru[2];
  int max_iter, iter, done, converged, breakdown, nb_ru;
  int natural; // stopping norm: 0 r.r, 1 r.z (natural), 2 z.z (preconditioned; rr holds z.z, bb (M b).(M b))
};

// per-block partial sums, kMaxBlocks entries each
struct Parts
{
  double* pq;
  double* rz[2];
  double* rzru[2];
  double* rr;
  double* rr64;
  double* bb;
  double* zq; // flexible PCG: z_{k+1}.q_k
};

// ---- block reductions --------------------------------------------------------------------

// Sum of n partials, identical in every block of BS threads; all threads get
// the result.  The order of the additions depends on BS only through the
// final per-wave sums, which are added in a fixed order.
template <int BS = kBlock>
__device__ __forceinline__ double sum_parts(const double* __restrict__ p, int n)
{
  __shared__ double lds[BS / kWave];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += BS)
    s += p[i];
  s = wave_reduce_sum(s);
  if ((threadIdx.x & (kWave - 1)) == 0)
    lds[threadIdx.x / kWave] = s;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < BS / kWave; ++w)
    t += lds[w];
  __syncthreads();
  return t;
}

// K partial arrays reduced in one pass (out[j] = sum_parts(p[j], n)): the
// loads of all arrays are in flight together and one barrier round serves
// them all; each sum is formed in exactly the order sum_parts uses.
template <int K, int BS = kBlock>
__device__ __forceinline__ void sum_parts_n(const double* const (&p)[K], int n, double (&out)[K])
{
  __shared__ double lds[K][BS / kWave];
  double s[K];
#pragma unroll
  for (int j = 0; j < K; ++j)
    s[j] = 0.0;
  for (int i = threadIdx.x; i < n; i += BS)
#pragma unroll
    for (int j = 0; j < K; ++j)
      s[j] += p[j][i];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    s[j] = wave_reduce_sum(s[j]);
    if ((threadIdx.x & (kWave - 1)) == 0)
      lds[j][threadIdx.x / kWave] = s[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < K; ++j) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < BS / kWave; ++w)
      t += lds[j][w];
    out[j] = t;
  }
  __syncthreads();
}

// Writes the sum of v over the BS threads of the block to out[blockIdx.x].
template <int BS = kBlock>
__device__ __forceinline__ void put_part(double v, double* __restrict__ out)
{
  __shared__ double lds[BS / kWave];
  v = wave_reduce_sum(v);
  if ((threadIdx.x & (kWave - 1)) == 0)
    lds[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < BS / kWave; ++w)
      t += lds[w];
    out[blockIdx.x] = t;
  }
  __syncthreads(); // the next call reuses lds
}

// Contiguous chunk of tiles owned by this block (XCD-aware order).
__device__ __forceinline__ void chunk_of(int64_t ntiles, int64_t& t0, int64_t& t1)
{
  const int64_t chunk = (ntiles + gridDim.x - 1) / gridDim.x;
  const int64_t L = xcd_tile(blockIdx.x, gridDim.x);
  t0 = L * chunk;
  t1 = t0 + chunk < ntiles ? t0 + chunk : ntiles;
}

// ---- gates -------------------------------------------------------------------------------

enum GateMode
{
  kAlways = 0,
  kUntilDone = 1,
  kReliable = 2 // runs only when iteration k decides a reliable update
};

__device__ __forceinline__ bool reliable_decision(const State* st, double rr, int k)
{
  return rr <= st->delta2 * st->rr_ref[(k - 1) & 1] || rr <= st->tol2 || k + 1 >= st->max_iter || !isfinite(rr);
}

struct Gate
{
  const State* st;
  int mode;
  int k;
  const double* rr_parts;
  int nparts;
  // evaluated by every thread of a block of BS threads (uses __syncthreads
  // through sum_parts)
  template <int BS = kBlock>
  __device__ bool open() const
  {
    if (mode == kAlways)
      return true;
    if (st->done || st->breakdown)
      return false;
    if (mode == kReliable)
      return reliable_decision(st, sum_parts<BS>(rr_parts, nparts), k);
    return true;
  }
};

// ---- SpMV ----------------------------------------------------------------------------------

enum SpmvMode
{
  kPlain = 0,
  kPQ = 1,      // y = A x; partial x.y
  kResidual = 2 // y = b - A x; partial y.y
};

template <class VT>
struct Acc
{
  using type = double;
};
template <>
struct Acc<float>
{
  using type = float;
};

template <int MODE, class XT>
__device__ __forceinline__ void spmv_epilogue(int64_t row, double sum, const XT* __restrict__ x, XT* __restrict__ y,
                                              const double* __restrict__ b, double& part)
{
  if constexpr (MODE == kResidual) {
    const double r = b[row] - sum;
    y[row] = (XT)r;
    part += r * r;
  }
  else {
    y[row] = (XT)sum;
    if constexpr (MODE == kPQ)
      part += sum * (double)x[row];
  }
}

// CSR, G lanes per row, 256/G rows per tile.
// DIFF (scalar FP32 copies in difference form, see k_to_float_rowsum): the
// diagonal slot holds the row sum s_i and multiplies x_i, every other slot
// multiplies x_j - x_i, so y_i = s_i x_i + sum_{j != i} a_ij (x_j - x_i).
template <int G, class VT, class XT, int MODE, bool DIFF = false>
__global__ void __launch_bounds__(kBlock)
k_spmv_csr(int32_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const VT* __restrict__ v,
           const XT* __restrict__ x, XT* __restrict__ y, const double* __restrict__ b, Gate gate,
           double* __restrict__ part_out)
{
  if (!gate.open())
    return;
  using A = typename Acc<VT>::type;
  constexpr int RPT = kBlock / G;
  int64_t t0, t1;
  chunk_of(((int64_t)n + RPT - 1) / RPT, t0, t1);
  const int lane = threadIdx.x & (G - 1);
  const int sub = threadIdx.x / G;
  double part = 0.0;
  for (int64_t t = t0; t < t1; ++t) {
    const int64_t row = t * RPT + sub;
    if (row < n) {
      const int32_t s = rp[row], e = rp[row + 1];
      A sum = 0;
      if constexpr (DIFF) {
        const A xi = (A)x[row];
        for (int32_t k = s + lane; k < e; k += G) {
          const int32_t c = ci[k];
          sum += (A)v[k] * ((A)x[c] - (c == row ? (A)0 : xi));
        }
      }
      else
        for (int32_t k = s + lane; k < e; k += G)
          sum += (A)v[k] * (A)x[ci[k]];
      sum = group_reduce_sum<G>(sum);
      if (lane == 0)
        spmv_epilogue<MODE>(row, (double)sum, x, y, b, part);
    }
  }
  if constexpr (MODE != kPlain)
    put_part(part, part_out);
}

// SELL-64: slice s holds rows [64s, 64s+64); its slots (j, lane) live at
// (off[s] + j) * 64 + lane for j < off[s+1] - off[s]; padding slots carry a
// zero value and the row's own column.
template <class VT, class XT, int MODE, bool DIFF = false>
__global__ void __launch_bounds__(kBlock)
k_spmv_sell(int32_t n, const int32_t* __restrict__ off, const int32_t* __restrict__ col, const VT* __restrict__ v,
            const XT* __restrict__ x, XT* __restrict__ y, const double* __restrict__ b, Gate gate,
            double* __restrict__ part_out)
{
  if (!gate.open())
    return;
  using A = typename Acc<VT>::type;
  int64_t t0, t1;
  chunk_of(((int64_t)n + kBlock - 1) / kBlock, t0, t1);
  double part = 0.0;
  for (int64_t t = t0; t < t1; ++t) {
    const int64_t row = t * kBlock + threadIdx.x;
    if (row < n) {
      const int64_t s = row >> 6;
      const int lane = (int)(row & 63);
      const int64_t base = (int64_t)off[s] * 64 + lane;
      const int w = off[s + 1] - off[s];
      const A xi = DIFF ? (A)x[row] : (A)0;
      auto gx = [&](int32_t c) { return DIFF ? (A)x[c] - (c == row ? (A)0 : xi) : (A)x[c]; };
      A s0 = 0, s1 = 0;
      int j = 0;
      for (; j + 1 < w; j += 2) {
        const int64_t i0 = base + (int64_t)j * 64, i1 = i0 + 64;
        s0 += (A)v[i0] * gx(col[i0]);
        s1 += (A)v[i1] * gx(col[i1]);
      }
      if (j < w) {
        const int64_t i0 = base + (int64_t)j * 64;
        s0 += (A)v[i0] * gx(col[i0]);
      }
      spmv_epilogue<MODE>(row, (double)(s0 + s1), x, y, b, part);
    }
  }
  if constexpr (MODE != kPlain)
    put_part(part, part_out);
}

// SELL-64 with 16-bit column offsets: slot (j, lane) of slice s stores
// col - row at (off[s] + j) * 64 + lane (padding slots: 0 with a zero value).
// The value and offset streams are read once, so they are loaded non-temporal
// to leave the L2 to the gathered vector.  BS threads per workgroup (256 or
// 512: the grid is capped at kMaxBlocks workgroups, so 512 doubles the waves
// in flight without more reduction partials).
template <class VT, class XT, int MODE, int UNR = 4, int BS = kBlock, bool DIFF = false>
__global__ void __launch_bounds__(BS)
k_spmv_sell16(int32_t n, const int32_t* __restrict__ off, const int16_t* __restrict__ dcol,
              const VT* __restrict__ v, const XT* __restrict__ x, XT* __restrict__ y, const double* __restrict__ b,
              Gate gate, double* __restrict__ part_out)
{
  if (!gate.open<BS>())
    return;
  using A = typename Acc<VT>::type;
  int64_t t0, t1;
  chunk_of(((int64_t)n + BS - 1) / BS, t0, t1);
  double part = 0.0;
  for (int64_t t = t0; t < t1; ++t) {
    const int64_t row = t * BS + threadIdx.x;
    if (row < n) {
      const int64_t s = row >> 6;
      const int64_t base = (int64_t)off[s] * 64 + (row & 63);
      const int w = off[s + 1] - off[s];
      const XT* xr = x + row;
      // offset 0 is the diagonal slot (and the padding, whose value is 0)
      const A xi = DIFF ? (A)xr[0] : (A)0;
      auto gx = [&](int d) { return DIFF ? (A)xr[d] - (d != 0 ? xi : (A)0) : (A)xr[d]; };
      A s0 = 0, s1 = 0, s2 = 0, s3 = 0;
      int j = 0;
      if constexpr (UNR == 8) {
        // all offsets and values of 8 slots first, then the 8 gathers
        for (; j + 7 < w; j += 8) {
          const int64_t i0 = base + (int64_t)j * 64;
          int d[8];
          VT m[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            d[q] = __builtin_nontemporal_load(dcol + i0 + 64 * q);
            m[q] = __builtin_nontemporal_load(v + i0 + 64 * q);
          }
          s0 += (A)m[0] * gx(d[0]) + (A)m[4] * gx(d[4]);
          s1 += (A)m[1] * gx(d[1]) + (A)m[5] * gx(d[5]);
          s2 += (A)m[2] * gx(d[2]) + (A)m[6] * gx(d[6]);
          s3 += (A)m[3] * gx(d[3]) + (A)m[7] * gx(d[7]);
        }
      }
      for (; j + 3 < w; j += 4) {
        const int64_t i0 = base + (int64_t)j * 64;
        const int d0 = __builtin_nontemporal_load(dcol + i0);
        const int d1 = __builtin_nontemporal_load(dcol + i0 + 64);
        const int d2 = __builtin_nontemporal_load(dcol + i0 + 128);
        const int d3 = __builtin_nontemporal_load(dcol + i0 + 192);
        s0 += (A)__builtin_nontemporal_load(v + i0) * gx(d0);
        s1 += (A)__builtin_nontemporal_load(v + i0 + 64) * gx(d1);
        s2 += (A)__builtin_nontemporal_load(v + i0 + 128) * gx(d2);
        s3 += (A)__builtin_nontemporal_load(v + i0 + 192) * gx(d3);
      }
      if (j < w) { // the last 1-3 slots, loads issued together
        const int64_t i0 = base + (int64_t)j * 64;
        const int rem = w - j;
        int d[3] = { 0, 0, 0 };
        VT m[3] = { 0, 0, 0 };
#pragma unroll
        for (int u = 0; u < 3; ++u)
          if (u < rem) {
            d[u] = __builtin_nontemporal_load(dcol + i0 + 64 * u);
            m[u] = __builtin_nontemporal_load(v + i0 + 64 * u);
          }
        s0 += (A)m[0] * gx(d[0]);
        s1 += (A)m[1] * gx(d[1]);
        s2 += (A)m[2] * gx(d[2]);
      }
      spmv_epilogue<MODE>(row, (double)((s0 + s1) + (s2 + s3)), x, y, b, part);
    }
  }
  if constexpr (MODE != kPlain)
    put_part<BS>(part, part_out);
}

// BSR 3x3 blocks stored per block (9 contiguous values), G lanes per node row.
template <int G, class VT, class XT, int MODE, bool DIFF = false>
__global__ void __launch_bounds__(kBlock)
k_spmv_bsr3(int32_t nb_node, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const VT* __restrict__ v,
            const XT* __restrict__ x, XT* __restrict__ y, const double* __restrict__ b, Gate gate,
            double* __restrict__ part_out)
{
  if (!gate.open())
    return;
  using A = typename Acc<VT>::type;
  constexpr int RPT = kBlock / G;
  int64_t t0, t1;
  chunk_of(((int64_t)nb_node + RPT - 1) / RPT, t0, t1);
  const int lane = threadIdx.x & (G - 1);
  const int sub = threadIdx.x / G;
  double part = 0.0;
  for (int64_t t = t0; t < t1; ++t) {
    const int64_t a = t * RPT + sub;
    if (a < nb_node) {
      const int32_t s = rp[a], e = rp[a + 1];
      // DIFF: the diagonal block holds the block row sum and multiplies x_a,
      // every other block x_c - x_a
      const A xa0 = DIFF ? (A)x[3 * a] : (A)0, xa1 = DIFF ? (A)x[3 * a + 1] : (A)0,
              xa2 = DIFF ? (A)x[3 * a + 2] : (A)0;
      A s0 = 0, s1 = 0, s2 = 0;
      for (int32_t k = s + lane; k < e; k += G) {
        const VT* blk = v + 9 * (int64_t)k;
        const int32_t cn = ci[k];
        const int64_t c = 3 * (int64_t)cn;
        const bool off = DIFF && cn != a;
        const A x0 = (A)x[c] - (off ? xa0 : (A)0), x1 = (A)x[c + 1] - (off ? xa1 : (A)0),
                x2 = (A)x[c + 2] - (off ? xa2 : (A)0);
        s0 += (A)blk[0] * x0 + (A)blk[1] * x1 + (A)blk[2] * x2;
        s1 += (A)blk[3] * x0 + (A)blk[4] * x1 + (A)blk[5] * x2;
        s2 += (A)blk[6] * x0 + (A)blk[7] * x1 + (A)blk[8] * x2;
      }
      s0 = group_reduce_sum<G>(s0);
      s1 = group_reduce_sum<G>(s1);
      s2 = group_reduce_sum<G>(s2);
      if (lane == 0) {
        spmv_epilogue<MODE>(3 * a, (double)s0, x, y, b, part);
        spmv_epilogue<MODE>(3 * a + 1, (double)s1, x, y, b, part);
        spmv_epilogue<MODE>(3 * a + 2, (double)s2, x, y, b, part);
      }
    }
  }
  if constexpr (MODE != kPlain)
    put_part(part, part_out);
}

// SELL-64 with 3x3 blocks: slot (j, lane) of slice s has its column at
// (off[s]+j)*64 + lane and value q (row-major in the block) at
// ((off[s]+j)*9 + q)*64 + lane.
template <class VT, class XT, int MODE, bool DIFF = false>
__global__ void __launch_bounds__(kBlock)
k_spmv_sell3(int32_t nb_node, const int32_t* __restrict__ off, const int32_t* __restrict__ col,
             const VT* __restrict__ v, const XT* __restrict__ x, XT* __restrict__ y, const double* __restrict__ b,
             Gate gate, double* __restrict__ part_out)
{
  if (!gate.open())
    return;
  using A = typename Acc<VT>::type;
  int64_t t0, t1;
  chunk_of(((int64_t)nb_node + kBlock - 1) / kBlock, t0, t1);
  double part = 0.0;
  for (int64_t t = t0; t < t1; ++t) {
    const int64_t a = t * kBlock + threadIdx.x;
    if (a < nb_node) {
      const int64_t s = a >> 6;
      const int lane = (int)(a & 63);
      const int64_t o = off[s];
      const int w = off[s + 1] - off[s];
      const A xa0 = DIFF ? (A)x[3 * a] : (A)0, xa1 = DIFF ? (A)x[3 * a + 1] : (A)0,
              xa2 = DIFF ? (A)x[3 * a + 2] : (A)0;
      A s0 = 0, s1 = 0, s2 = 0;
      for (int j = 0; j < w; ++j) {
        const int64_t slot = o + j;
        const int32_t cn = col[slot * 64 + lane];
        const int64_t c = 3 * (int64_t)cn;
        const VT* vb = v + slot * 9 * 64 + lane;
        const bool offd = DIFF && cn != a; // padding slots: the own column and a zero block
        const A x0 = (A)x[c] - (offd ? xa0 : (A)0), x1 = (A)x[c + 1] - (offd ? xa1 : (A)0),
                x2 = (A)x[c + 2] - (offd ? xa2 : (A)0);
        s0 += (A)vb[0] * x0 + (A)vb[64] * x1 + (A)vb[128] * x2;
        s1 += (A)vb[192] * x0 + (A)vb[256] * x1 + (A)vb[320] * x2;
        s2 += (A)vb[384] * x0 + (A)vb[448] * x1 + (A)vb[512] * x2;
      }
      spmv_epilogue<MODE>(3 * a, (double)s0, x, y, b, part);
      spmv_epilogue<MODE>(3 * a + 1, (double)s1, x, y, b, part);
      spmv_epilogue<MODE>(3 * a + 2, (double)s2, x, y, b, part);
    }
  }
  if constexpr (MODE != kPlain)
    put_part(part, part_out);
}

// ---- SELL construction -------------------------------------------------------------------------

__global__ void k_sell_width(int32_t nrows, const int32_t* __restrict__ rp, int32_t* __restrict__ width)
{
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t nslice = (nrows + 63) / 64;
  if (s >= nslice)
    return;
  int32_t w = 0;
  for (int32_t r = 64 * s; r < 64 * s + 64 && r < nrows; ++r)
    w = max(w, rp[r + 1] - rp[r]);
  width[s] = w;
}

// Columns and FP64/FP32 values of the SELL copy; NB = 1 (scalar) or 3.
template <int NB>
__global__ void k_sell_fill(int32_t nrows, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                            const double* __restrict__ v, const int32_t* __restrict__ off, int32_t* __restrict__ col,
                            double* __restrict__ v64, float* __restrict__ v32, int fill_columns)
{
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nslots_rows = ((int64_t)nrows + 63) / 64 * 64;
  if (r >= nslots_rows)
    return;
  const int64_t s = r >> 6;
  const int lane = (int)(r & 63);
  const int64_t o = off[s];
  const int w = off[s + 1] - off[s];
  const int32_t len = (r < nrows) ? rp[r + 1] - rp[r] : 0;
  const int32_t b0 = (r < nrows) ? rp[r] : 0;
  for (int j = 0; j < w; ++j) {
    const int64_t slot = o + j;
    const bool real = j < len;
    if (fill_columns)
      col[slot * 64 + lane] = real ? ci[b0 + j] : (int32_t)(r < nrows ? r : 0);
    for (int q = 0; q < NB * NB; ++q) {
      const double val = real ? v[(int64_t)(b0 + j) * NB * NB + q] : 0.0;
      const int64_t idx = (NB == 1) ? slot * 64 + lane : (slot * 9 + q) * 64 + lane;
      if (v64)
        v64[idx] = val;
      if (v32)
        v32[idx] = (float)val;
    }
  }
}

// 16-bit offsets col - row of the SELL columns (one workgroup of 64 per
// slice); *overflow = 1 when an offset does not fit.
__global__ void k_sell_col16(int32_t nrows, const int32_t* __restrict__ off, const int32_t* __restrict__ col,
                             int16_t* __restrict__ dcol, int32_t* __restrict__ overflow)
{
  const int64_t s = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t row = s * 64 + lane;
  const int w = off[s + 1] - off[s];
  bool bad = false;
  for (int j = 0; j < w; ++j) {
    const int64_t e = ((int64_t)off[s] + j) * 64 + lane;
    int64_t d = (row < nrows) ? (int64_t)col[e] - row : 0;
    if (d < -32768 || d > 32767) {
      bad = true;
      d = 0;
    }
    dcol[e] = (int16_t)d;
  }
  if (bad)
    atomicOr(overflow, 1);
}

// ---- preconditioner set-up ----------------------------------------------------------------

__global__ void k_diag_inverse(int32_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                               const double* __restrict__ v, double* __restrict__ d64, float* __restrict__ d32)
{
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n)
    return;
  double d = 0.0;
  for (int32_t k = rp[r]; k < rp[r + 1]; ++k)
    if (ci[k] == r)
      d = v[k];
  const double inv = (d != 0.0) ? 1.0 / d : 1.0;
  d64[r] = inv;
  if (d32)
    d32[r] = (float)inv;
}

// 3x3 diagonal-block inverse per node (adjugate / determinant).
__global__ void k_block_inverse3(int32_t nb_node, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                 const double* __restrict__ v, double* __restrict__ d64, float* __restrict__ d32)
{
  const int32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nb_node)
    return;
  double m[9] = { 1, 0, 0, 0, 1, 0, 0, 0, 1 };
  for (int32_t k = rp[a]; k < rp[a + 1]; ++k)
    if (ci[k] == a) {
      for (int q = 0; q < 9; ++q)
        m[q] = v[9 * (int64_t)k + q];
      break;
    }
  double inv[9];
  inv[0] = m[4] * m[8] - m[5] * m[7];
  inv[1] = m[2] * m[7] - m[1] * m[8];
  inv[2] = m[1] * m[5] - m[2] * m[4];
  inv[3] = m[5] * m[6] - m[3] * m[8];
  inv[4] = m[0] * m[8] - m[2] * m[6];
  inv[5] = m[2] * m[3] - m[0] * m[5];
  inv[6] = m[3] * m[7] - m[4] * m[6];
  inv[7] = m[1] * m[6] - m[0] * m[7];
  inv[8] = m[0] * m[4] - m[1] * m[3];
  const double det = m[0] * inv[0] + m[1] * inv[3] + m[2] * inv[6];
  if (det != 0.0 && isfinite(det)) {
    const double id = 1.0 / det;
    for (int q = 0; q < 9; ++q)
      inv[q] *= id;
  }
  else {
    for (int q = 0; q < 9; ++q)
      inv[q] = (q % 4 == 0) ? 1.0 : 0.0;
  }
  for (int q = 0; q < 9; ++q) {
    d64[9 * (int64_t)a + q] = inv[q];
    if (d32)
      d32[9 * (int64_t)a + q] = (float)inv[q];
  }
}

__global__ void k_to_float(int64_t n, const double* __restrict__ in, float* __restrict__ out)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (float)in[i];
}

// FP32 copy of a scalar matrix in difference form: the off-diagonal entries
// rounded to FP32, the diagonal entry replaced by the row sum s_i taken in
// FP64.  The FP32 SpMV then forms A x as s_i x_i + sum_{j != i} a_ij (x_j - x_i):
// for a smooth x (the slow modes of a diffusion operator, whose rows nearly
// sum to zero) the terms are small differences instead of a large diagonal
// term cancelling the off-diagonal ones, so rounding A to FP32 no longer
// perturbs the low end of its spectrum.  *missing = 1 when a non-empty row
// has no diagonal entry (the form needs one slot per row for s_i).
__global__ void k_to_float_rowsum(int32_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                  const double* __restrict__ v, float* __restrict__ out, int32_t* __restrict__ missing)
{
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n)
    return;
  const int32_t e = rp[r + 1];
  double sum = 0.0;
  int32_t dpos = -1;
  for (int32_t k = rp[r]; k < e; ++k) {
    const double a = v[k];
    sum += a;
    if (ci[k] == r)
      dpos = k;
    else
      out[k] = (float)a;
  }
  if (dpos >= 0)
    out[dpos] = (float)sum;
  else if (rp[r] < e)
    *missing = 1;
}

// x0: zero except eliminated rows, which start at their value (keeps a
// row-eliminated system on its symmetric reduced part).
__global__ void k_initial_guess(int32_t n, const uint8_t* __restrict__ elim, const double* __restrict__ elim_value,
                                double* __restrict__ x)
{
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n)
    x[r] = (elim[r] == kEliminateRow || elim[r] == kEliminateRowColumn) ? elim_value[r] : 0.0;
}

// ---- vector kernels ------------------------------------------------------------------------

template <int NB, class T>
__device__ __forceinline__ void apply_jacobi(int64_t u, const T* __restrict__ dinv, const T* rv, T* zv, int precond)
{
  if (!precond) {
#pragma unroll
    for (int i = 0; i < NB; ++i)
      zv[i] = rv[i];
  }
  else if constexpr (NB == 1) {
    zv[0] = dinv[u] * rv[0];
  }
  else {
    const T* d = dinv + 9 * u;
    zv[0] = d[0] * rv[0] + d[1] * rv[1] + d[2] * rv[2];
    zv[1] = d[3] * rv[0] + d[4] * rv[1] + d[5] * rv[2];
    zv[2] = d[6] * rv[0] + d[7] * rv[1] + d[8] * rv[2];
  }
}

// rho_{k-1}, the r.z of iteration k-1 or of the reliable update that ended
// it: the partials it is summed from.
__device__ __forceinline__ const double* rho_prev_parts(const State* st, const Parts& P, int k)
{
  const int q = (k - 1) & 1;
  return st->ru[q] ? P.rzru[q] : P.rz[q];
}

// r0 = b - A x0 is in r64.  z = M r, p = z; partials r.z (parity 1), r.r, b.b.
template <int NB, class T>
__global__ void __launch_bounds__(kBlock)
k_init(int64_t nunits, const double* r64, const double* __restrict__ b, T* r, T* __restrict__ z, T* __restrict__ p,
       const T* __restrict__ dinv, int precond, int pnorm, Parts P)
{
  int64_t t0, t1;
  chunk_of((nunits + kBlock - 1) / kBlock, t0, t1);
  double rz = 0.0, rr = 0.0, bb = 0.0;
  for (int64_t t = t0; t < t1; ++t) {
    const int64_t u = t * kBlock + threadIdx.x;
    if (u >= nunits)
      continue;
    T rv[NB], zv[NB], bv[NB], mb[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const double rd = r64[u * NB + i];
      rv[i] = (T)rd;
      r[u * NB + i] = rv[i];
      bv[i] = (T)b[u * NB + i];
    }
    apply_jacobi<NB, T>(u, dinv, rv, zv, precond);
    if (pnorm)
      apply_jacobi<NB, T>(u, dinv, bv, mb, precond);
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      z[u * NB + i] = zv[i];
      p[u * NB + i] = zv[i];
      rz += (double)rv[i] * (double)zv[i];
      rr += pnorm ? (double)zv[i] * (double)zv[i] : r64[u * NB + i] * r64[u * NB + i];
      bb += pnorm ? (double)mb[i] * (double)mb[i] : b[u * NB + i] * b[u * NB + i];
    }
  }
  put_part(rz, P.rz[1]);
  put_part(rr, P.rr);
  put_part(bb, P.bb);
}

__global__ void __launch_bounds__(kBlock) k_init_finalize(Parts P, int nparts, State* st)
{
  const double rz = sum_parts(P.rz[1], nparts);
  const double rr = sum_parts(P.rr, nparts);
  const double bb = sum_parts(P.bb, nparts);
  if (threadIdx.x == 0) {
    st->bb = bb;
    st->tol2 = st->natural == 1 ? st->rtol2 * rz : fmax(st->rtol2 * bb, st->atol2);
    st->rr_ref[1] = rr;
    st->ru[1] = 0;
    st->rr_out = rr;
    st->iter = 0;
    st->done = ((st->natural == 1 ? rz : rr) <= st->tol2 || rz == 0.0) ? 1 : 0;
    st->converged = st->done;
  }
}

// x += alpha p, r -= alpha q, z = M r; partials r.z (parity k&1) and r.r.
template <int NB, class T>
__global__ void __launch_bounds__(kBlock)
k_update(int k, int64_t nunits, double* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
         const T* __restrict__ p, const T* __restrict__ q, const T* __restrict__ dinv, int precond, Parts P,
         int nparts, State* st)
{
  if (st->done || st->breakdown)
    return;
  const double* const arrs[2] = { P.pq, rho_prev_parts(st, P, k) };
  double sums[2];
  sum_parts_n<2>(arrs, nparts, sums);
  const double pq = sums[0], rho = sums[1];
  if (!(pq > 0.0) || !isfinite(pq)) { // the operator is not positive definite on p
    if (blockIdx.x == 0 && threadIdx.x == 0)
      st->breakdown = 1;
    return;
  }
  const double alpha = rho / pq;
  const T alpha_t = (T)alpha;
  const bool pnorm = st->natural == 2;
  int64_t t0, t1;
  chunk_of((nunits + kBlock - 1) / kBlock, t0, t1);
  double rz = 0.0, rr = 0.0;
  for (int64_t t = t0; t < t1; ++t) {
    const int64_t u = t * kBlock + threadIdx.x;
    if (u >= nunits)
      continue;
    T rv[NB], zv[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int64_t row = u * NB + i;
      x[row] += alpha * (double)p[row];
      rv[i] = r[row] - alpha_t * q[row];
      r[row] = rv[i];
    }
    apply_jacobi<NB, T>(u, dinv, rv, zv, precond);
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      z[u * NB + i] = zv[i];
      rz += (double)rv[i] * (double)zv[i];
      rr += pnorm ? (double)zv[i] * (double)zv[i] : (double)rv[i] * (double)rv[i];
    }
  }
  put_part(rz, P.rz[k & 1]);
  put_part(rr, P.rr);
}

// Convergence test and p = z + beta p.  Mixed solver: after a reliable
// update r and z are replaced by the FP64 residual and its preconditioned
// value, and the new r.z goes to the rzru parity buffer.
template <int NB, class T>
__global__ void __launch_bounds__(kBlock)
k_pupdate(int k, int mixed, int64_t nunits, T* __restrict__ p, T* __restrict__ z, T* __restrict__ r,
          const double* __restrict__ r64, const T* __restrict__ dinv, int precond, Parts P, int nparts, State* st)
{
  if (st->done)
    return;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (st->breakdown) {
    if (lead) {
      st->done = 1;
      st->iter = k;
    }
    return;
  }
  const double* const arrs[4] = { P.rz[k & 1], P.rr, rho_prev_parts(st, P, k), P.rr64 };
  double sums[4];
  sum_parts_n<4>(arrs, nparts, sums);
  const double rz = sums[0], rr = sums[1], rho = sums[2];
  const double beta = (rho != 0.0) ? rz / rho : 0.0;
  const int q = k & 1;
  bool ru = false;
  if (mixed) {
    ru = reliable_decision(st, rr, k);
    if (ru) {
      const double rr64 = sums[3]; // written by the reliable-update residual of this iteration
      if (rr64 <= st->tol2 || k + 1 >= st->max_iter || !isfinite(rr64)) {
        if (lead) {
          st->done = 1;
          st->converged = rr64 <= st->tol2;
          st->iter = k + 1;
          st->rr_out = rr64;
          st->nb_ru += 1;
        }
        return;
      }
      if (lead) {
        st->ru[q] = 1;
        st->rr_ref[q] = rr64;
        st->rr_out = rr64;
        st->nb_ru += 1;
      }
    }
    else if (lead) {
      st->ru[q] = 0;
      st->rr_ref[q] = st->rr_ref[(k - 1) & 1];
      st->rr_out = rr;
    }
  }
  else {
    const double crit = st->natural == 1 ? rz : rr;
    if (crit <= st->tol2 || k + 1 >= st->max_iter || !isfinite(rr)) {
      if (lead) {
        st->done = 1;
        st->converged = crit <= st->tol2;
        st->iter = k + 1;
        st->rr_out = rr;
      }
      return;
    }
    if (lead) {
      st->ru[q] = 0;
      st->rr_out = rr;
    }
  }
  if (lead)
    st->iter = k + 1;
  const T beta_t = (T)beta;
  int64_t t0, t1;
  chunk_of((nunits + kBlock - 1) / kBlock, t0, t1);
  double rzn = 0.0;
  for (int64_t t = t0; t < t1; ++t) {
    const int64_t u = t * kBlock + threadIdx.x;
    if (u >= nunits)
      continue;
    T zv[NB];
    if (ru) {
      T rv[NB];
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        rv[i] = (T)r64[u * NB + i];
        r[u * NB + i] = rv[i];
      }
      apply_jacobi<NB, T>(u, dinv, rv, zv, precond);
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        z[u * NB + i] = zv[i];
        rzn += (double)rv[i] * (double)zv[i];
      }
    }
    else {
#pragma unroll
      for (int i = 0; i < NB; ++i)
        zv[i] = z[u * NB + i];
    }
#pragma unroll
    for (int i = 0; i < NB; ++i)
      p[u * NB + i] = zv[i] + beta_t * p[u * NB + i];
  }
  if (ru)
    put_part(rzn, P.rzru[q]);
}

__global__ void __launch_bounds__(kBlock) k_true_residual_finalize(Parts P, int nparts, State* st)
{
  const double rr = sum_parts(P.rr64, nparts);
  if (threadIdx.x == 0)
    st->rr_true = rr;
}

// ---- flexible PCG with a sparse preconditioner matrix ------------------------------------------
//
// z = M r is an SpMV with the preconditioner's values (FP32, pattern of A),
// so an iteration is: q = A p (p.q); x += alpha p, r -= alpha q (r.r);
// [reliable update: r64 = b - A64 x]; z = M r (r.z, z.q); p = z + beta p
// with the Polak-Ribiere beta = z_{k+1}.(r_{k+1} - r_k) / rho_k
// = -alpha z_{k+1}.q_k / rho_k, which keeps CG's convergence when M is only
// approximately symmetric.  After a reliable update r is replaced and
// beta = rho_{k+1} / rho_k.

// Input of z = M r on iteration k: the FP32 residual, or the FP64 one on an
// iteration that takes a reliable update (always on the first call).
struct PrecIn
{
  const float* r32;
  const double* r64;
  bool from64;
  __device__ __forceinline__ float at(int64_t c) const { return from64 ? (float)r64[c] : r32[c]; }
};

template <int INIT>
__device__ __forceinline__ void prec_row(int64_t row, float zrow, const PrecIn& in, float* __restrict__ r32,
                                         float* __restrict__ z, float* __restrict__ p, const float* __restrict__ q,
                                         const double* __restrict__ b, double& rz, double& zq, double& rr,
                                         double& bb)
{
  const float rin = in.at(row);
  z[row] = zrow;
  if (in.from64)
    r32[row] = rin; // no other thread reads r32 on such an iteration
  rz += (double)zrow * (double)rin;
  if constexpr (INIT) {
    p[row] = zrow;
    const double bv = b[row], r64v = in.r64[row];
    bb += bv * bv;
    rr += r64v * r64v;
  }
  else {
    zq += (double)zrow * (double)q[row];
  }
}

template <int INIT, int BS = kBlock>
__device__ __forceinline__ bool prec_open(const State* st, const Parts& P, int nparts, int k, PrecIn& in)
{
  if constexpr (INIT) {
    in.from64 = true;
    return true;
  }
  else {
    if (st->done || st->breakdown)
      return false;
    in.from64 = reliable_decision(st, sum_parts<BS>(P.rr, nparts), k);
    return true;
  }
}

template <int INIT, int BS = kBlock>
__device__ __forceinline__ void prec_parts(const Parts& P, int k, double rz, double zq, double rr, double bb)
{
  put_part<BS>(rz, P.rz[k & 1]);
  if constexpr (INIT) {
    put_part<BS>(rr, P.rr);
    put_part<BS>(bb, P.bb);
  }
  else {
    put_part<BS>(zq, P.zq);
  }
}

template <int G, int INIT>
__global__ void __launch_bounds__(kBlock)
k_prec_csr(int k, int32_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
           const float* __restrict__ mv, PrecIn in, float* __restrict__ r32, float* __restrict__ z,
           float* __restrict__ p, const float* __restrict__ q, const double* __restrict__ b, Parts P, int nparts,
           const State* st)
{
  if (!prec_open<INIT>(st, P, nparts, k, in))
    return;
  constexpr int RPT = kBlock / G;
  int64_t t0, t1;
  chunk_of(((int64_t)n + RPT - 1) / RPT, t0, t1);
  const int lane = threadIdx.x & (G - 1);
  const int sub = threadIdx.x / G;
  double rz = 0.0, zq = 0.0, rr = 0.0, bb = 0.0;
  for (int64_t t = t0; t < t1; ++t) {
    const int64_t row = t * RPT + sub;
    if (row < n) {
      float sum = 0.f;
      for (int32_t e = rp[row] + lane; e < rp[row + 1]; e += G)
        sum += mv[e] * in.at(ci[e]);
      sum = group_reduce_sum<G>(sum);
      if (lane == 0)
        prec_row<INIT>(row, sum, in, r32, z, p, q, b, rz, zq, rr, bb);
    }
  }
  prec_parts<INIT>(P, k, rz, zq, rr, bb);
}

template <int INIT>
__global__ void __launch_bounds__(kBlock)
k_prec_sell(int k, int32_t n, const int32_t* __restrict__ off, const int32_t* __restrict__ col,
            const float* __restrict__ mv, PrecIn in, float* __restrict__ r32, float* __restrict__ z,
            float* __restrict__ p, const float* __restrict__ q, const double* __restrict__ b, Parts P, int nparts,
            const State* st)
{
  if (!prec_open<INIT>(st, P, nparts, k, in))
    return;
  int64_t t0, t1;
  chunk_of(((int64_t)n + kBlock - 1) / kBlock, t0, t1);
  double rz = 0.0, zq = 0.0, rr = 0.0, bb = 0.0;
  for (int64_t t = t0; t < t1; ++t) {
    const int64_t row = t * kBlock + threadIdx.x;
    if (row < n) {
      const int64_t s = row >> 6;
      const int64_t base = (int64_t)off[s] * 64 + (row & 63);
      const int w = off[s + 1] - off[s];
      float s0 = 0.f, s1 = 0.f;
      int j = 0;
      for (; j + 1 < w; j += 2) {
        const int64_t i0 = base + (int64_t)j * 64, i1 = i0 + 64;
        s0 += mv[i0] * in.at(col[i0]);
        s1 += mv[i1] * in.at(col[i1]);
      }
      if (j < w) {
        const int64_t i0 = base + (int64_t)j * 64;
        s0 += mv[i0] * in.at(col[i0]);
      }
      prec_row<INIT>(row, s0 + s1, in, r32, z, p, q, b, rz, zq, rr, bb);
    }
  }
  prec_parts<INIT>(P, k, rz, zq, rr, bb);
}

// Slot i of the preconditioner in SELL order: its column offset and value.
// FP32 values come with the separate 16-bit offset stream; BF16 values are
// packed with the offset in one 32-bit word (value in the high half, offset
// in the low half), so a slot costs one load instead of two.
__device__ __forceinline__ void prec_slot(const int16_t* __restrict__ dcol, const float* __restrict__ m, int64_t i,
                                          int& d, float& v)
{
  d = __builtin_nontemporal_load(dcol + i);
  v = __builtin_nontemporal_load(m + i);
}
__device__ __forceinline__ void prec_slot(const int16_t* __restrict__, const uint32_t* __restrict__ m, int64_t i,
                                          int& d, float& v)
{
  const uint32_t w = __builtin_nontemporal_load(m + i);
  d = (int)(int16_t)(w & 0xffffu);
  v = __uint_as_float(w & 0xffff0000u);
}

// z = M r_in over the rows of this workgroup with the input fixed at compile
// time (FROM64: the FP64 residual of a reliable update, converted).
template <int INIT, class MT, bool FROM64, int BS>
__device__ __forceinline__ void prec_sell16_rows(int32_t n, const int32_t* __restrict__ off,
                                                 const int16_t* __restrict__ dcol, const MT* __restrict__ mv,
                                                 const PrecIn& in, float* __restrict__ r32, float* __restrict__ z,
                                                 float* __restrict__ p, const float* __restrict__ q,
                                                 const double* __restrict__ b, double& rz, double& zq, double& rr,
                                                 double& bb)
{
  auto xin = [&](int64_t c) -> float { return FROM64 ? (float)in.r64[c] : in.r32[c]; };
  int64_t t0, t1;
  chunk_of(((int64_t)n + BS - 1) / BS, t0, t1);
  for (int64_t t = t0; t < t1; ++t) {
    const int64_t row = t * BS + threadIdx.x;
    if (row >= n)
      continue;
    const int64_t s = row >> 6;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    if constexpr (std::is_same<MT, uint32_t>::value) {
      // BF16 words, 4 slots per 16-byte load (off holds the 4-slot layout)
      typedef unsigned v4u __attribute__((ext_vector_type(4)));
      const int64_t base = (int64_t)off[s] * 64 + 4 * (row & 63);
      const int w4 = off[s + 1] - off[s];
      for (int j = 0; j < w4; j += 8) {
        const v4u a = __builtin_nontemporal_load(reinterpret_cast<const v4u*>(mv + base + (int64_t)j * 64));
        v4u c = { 0u, 0u, 0u, 0u };
        if (j + 4 < w4)
          c = __builtin_nontemporal_load(reinterpret_cast<const v4u*>(mv + base + (int64_t)(j + 4) * 64));
        const uint32_t wd[8] = { a.x, a.y, a.z, a.w, c.x, c.y, c.z, c.w };
        int d[8];
        float m[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          d[u] = (int)(int16_t)(wd[u] & 0xffffu);
          m[u] = __uint_as_float(wd[u] & 0xffff0000u);
        }
        s0 += m[0] * xin(row + d[0]) + m[4] * xin(row + d[4]);
        s1 += m[1] * xin(row + d[1]) + m[5] * xin(row + d[5]);
        s2 += m[2] * xin(row + d[2]) + m[6] * xin(row + d[6]);
        s3 += m[3] * xin(row + d[3]) + m[7] * xin(row + d[7]);
      }
    }
    else {
    const int64_t base = (int64_t)off[s] * 64 + (row & 63);
    const int w = off[s + 1] - off[s];
    int j = 0;
    for (; j + 7 < w; j += 8) {
      const int64_t i0 = base + (int64_t)j * 64;
      int d[8];
      float m[8];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        prec_slot(dcol, mv, i0 + 64 * u, d[u], m[u]);
      s0 += m[0] * xin(row + d[0]) + m[4] * xin(row + d[4]);
      s1 += m[1] * xin(row + d[1]) + m[5] * xin(row + d[5]);
      s2 += m[2] * xin(row + d[2]) + m[6] * xin(row + d[6]);
      s3 += m[3] * xin(row + d[3]) + m[7] * xin(row + d[7]);
    }
    for (; j + 3 < w; j += 4) {
      const int64_t i0 = base + (int64_t)j * 64;
      int d[4];
      float m[4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        prec_slot(dcol, mv, i0 + 64 * u, d[u], m[u]);
      s0 += m[0] * xin(row + d[0]);
      s1 += m[1] * xin(row + d[1]);
      s2 += m[2] * xin(row + d[2]);
      s3 += m[3] * xin(row + d[3]);
    }
    if (j < w) { // the last 1-3 slots, loads issued together
      const int64_t i0 = base + (int64_t)j * 64;
      const int rem = w - j;
      int d[3] = { 0, 0, 0 };
      float m[3] = { 0.f, 0.f, 0.f };
#pragma unroll
      for (int u = 0; u < 3; ++u)
        if (u < rem)
          prec_slot(dcol, mv, i0 + 64 * u, d[u], m[u]);
      s0 += m[0] * xin(row + d[0]);
      s1 += m[1] * xin(row + d[1]);
      s2 += m[2] * xin(row + d[2]);
    }
    }
    const float zrow = (s0 + s1) + (s2 + s3);
    const float rin = xin(row);
    z[row] = zrow;
    if (FROM64)
      r32[row] = rin; // no other thread reads r32 on such an iteration
    rz += (double)zrow * (double)rin;
    if constexpr (INIT) {
      p[row] = zrow;
      const double bv = b[row], r64v = in.r64[row];
      bb += bv * bv;
      rr += r64v * r64v;
    }
    else {
      zq += (double)zrow * (double)q[row];
    }
  }
}

template <int INIT, class MT, int BS = kBlock>
__global__ void __launch_bounds__(BS)
k_prec_sell16(int k, int32_t n, const int32_t* __restrict__ off, const int16_t* __restrict__ dcol,
              const MT* __restrict__ mv, PrecIn in, float* __restrict__ r32, float* __restrict__ z,
              float* __restrict__ p, const float* __restrict__ q, const double* __restrict__ b, Parts P, int nparts,
              const State* st)
{
  if (!prec_open<INIT, BS>(st, P, nparts, k, in))
    return;
  double rz = 0.0, zq = 0.0, rr = 0.0, bb = 0.0;
  if (in.from64)
    prec_sell16_rows<INIT, MT, true, BS>(n, off, dcol, mv, in, r32, z, p, q, b, rz, zq, rr, bb);
  else if constexpr (!INIT)
    prec_sell16_rows<INIT, MT, false, BS>(n, off, dcol, mv, in, r32, z, p, q, b, rz, zq, rr, bb);
  prec_parts<INIT, BS>(P, k, rz, zq, rr, bb);
}

// BF16 (round to nearest even) of the FP32 preconditioner values, packed
// with the 16-bit column offsets: word = bf16 << 16 | (uint16_t)offset.  The
// words of a lane's slots 4q .. 4q + 3 are adjacent -- slot j of lane l of
// slice s at off4[s] * 64 + 256 (j / 4) + 4 l + j % 4, widths rounded up to
// multiples of 4 with zero words (value 0, offset 0) -- so one 16-byte load
// reads four slots.  One 64-thread workgroup per slice.
__global__ void __launch_bounds__(64)
k_pack_bf16(const int32_t* __restrict__ off, const int32_t* __restrict__ off4, const float* __restrict__ in,
            const int16_t* __restrict__ dcol, uint32_t* __restrict__ out)
{
  const int64_t s = blockIdx.x;
  const int l = threadIdx.x;
  const int w = off[s + 1] - off[s], w4 = off4[s + 1] - off4[s];
  for (int j = 0; j < w4; ++j) {
    uint32_t word = 0u;
    if (j < w) {
      const int64_t i = ((int64_t)off[s] + j) * 64 + l;
      const uint32_t u = __float_as_uint(in[i]);
      const uint32_t h = (u & 0x7fffffffu) > 0x7f800000u ? 0x7fc0u : (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
      word = (h << 16) | (uint32_t)(uint16_t)dcol[i];
    }
    out[(int64_t)off4[s] * 64 + 256 * (j >> 2) + 4 * l + (j & 3)] = word;
  }
}

// x += alpha p, r -= alpha q; partial r.r.
__global__ void __launch_bounds__(kBlock)
k_update_flex(int k, int32_t n, double* __restrict__ x, float* __restrict__ r, const float* __restrict__ p,
              const float* __restrict__ q, Parts P, int nparts, State* st)
{
  if (st->done || st->breakdown)
    return;
  const double* const arrs[2] = { P.pq, P.rz[(k - 1) & 1] };
  double sums[2];
  sum_parts_n<2>(arrs, nparts, sums);
  const double pq = sums[0], rho = sums[1];
  if (!(pq > 0.0) || !isfinite(pq) || !(rho > 0.0)) {
    if (blockIdx.x == 0 && threadIdx.x == 0)
      st->breakdown = 1;
    return;
  }
  const double alpha = rho / pq;
  const float alpha_t = (float)alpha;
  int64_t t0, t1;
  chunk_of(((int64_t)n + kBlock - 1) / kBlock, t0, t1);
  double rr = 0.0;
  for (int64_t t = t0; t < t1; ++t) {
    const int64_t row = t * kBlock + threadIdx.x;
    if (row >= n)
      continue;
    x[row] += alpha * (double)p[row];
    const float rv = r[row] - alpha_t * q[row];
    r[row] = rv;
    rr += (double)rv * (double)rv;
  }
  put_part(rr, P.rr);
}

// Convergence test (FP64 residual of a reliable update), beta, p = z + beta p.
__global__ void __launch_bounds__(kBlock)
k_pupdate_flex(int k, int32_t n, float* __restrict__ p, const float* __restrict__ z, Parts P, int nparts, State* st)
{
  if (st->done)
    return;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (st->breakdown) {
    if (lead) {
      st->done = 1;
      st->iter = k;
    }
    return;
  }
  const double* const arrs[6] = { P.rr, P.rz[(k - 1) & 1], P.pq, P.rz[k & 1], P.rr64, P.zq };
  double sums[6];
  sum_parts_n<6>(arrs, nparts, sums);
  const double rr = sums[0], rho = sums[1], pq = sums[2], rz = sums[3];
  const int q = k & 1;
  const bool ru = reliable_decision(st, rr, k);
  double beta;
  if (ru) {
    const double rr64 = sums[4];
    if (rr64 <= st->tol2 || k + 1 >= st->max_iter || !isfinite(rr64)) {
      if (lead) {
        st->done = 1;
        st->converged = rr64 <= st->tol2;
        st->iter = k + 1;
        st->rr_out = rr64;
        st->nb_ru += 1;
      }
      return;
    }
    if (lead) {
      st->rr_ref[q] = rr64;
      st->rr_out = rr64;
      st->nb_ru += 1;
    }
    beta = rz / rho;
  }
  else {
    const double zq = sums[5];
    if (lead) {
      st->rr_ref[q] = st->rr_ref[(k - 1) & 1];
      st->rr_out = rr;
    }
    beta = -(rho / pq) * zq / rho;
  }
  if (!(rz > 0.0) || !isfinite(rz) || !isfinite(beta)) { // M is not positive on this residual
    if (lead) {
      st->breakdown = 1;
      st->done = 1;
      st->iter = k + 1;
    }
    return;
  }
  if (lead) {
    st->ru[q] = 0;
    st->iter = k + 1;
  }
  const float beta_t = (float)beta;
  int64_t t0, t1;
  chunk_of(((int64_t)n + kBlock - 1) / kBlock, t0, t1);
  for (int64_t t = t0; t < t1; ++t) {
    const int64_t row = t * kBlock + threadIdx.x;
    if (row < n)
      p[row] = z[row] + beta_t * p[row];
  }
}

// k_to_float_rowsum for 3x3 blocks: the diagonal block holds the block row
// sum (taken in FP64), the other blocks are rounded; rigid translations then
// meet the block row sums exactly in the FP32 product.
__global__ void k_to_float_rowsum3(int32_t nb_node, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                   const double* __restrict__ v, float* __restrict__ out, int32_t* __restrict__ missing)
{
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nb_node)
    return;
  const int32_t e = rp[a + 1];
  double sum[9] = { 0, 0, 0, 0, 0, 0, 0, 0, 0 };
  int32_t dpos = -1;
  for (int32_t k = rp[a]; k < e; ++k) {
    const double* blk = v + 9 * (int64_t)k;
    const bool diag = ci[k] == a;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      sum[q] += blk[q];
      if (!diag)
        out[9 * (int64_t)k + q] = (float)blk[q];
    }
    if (diag)
      dpos = k;
  }
  if (dpos >= 0) {
#pragma unroll
    for (int q = 0; q < 9; ++q)
      out[9 * (int64_t)dpos + q] = (float)sum[q];
  }
  else if (rp[a] < e)
    *missing = 1;
}

// SELL copy of FP32 3x3 blocks given in CSR block order (slot layout of
// k_sell_fill<3>).
__global__ void k_sell_fill_f32_b3(int32_t nrows, const int32_t* __restrict__ rp, const float* __restrict__ v,
                                   const int32_t* __restrict__ off, float* __restrict__ out)
{
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t rows_pad = ((int64_t)nrows + 63) / 64 * 64;
  if (r >= rows_pad)
    return;
  const int64_t s = r >> 6;
  const int lane = (int)(r & 63);
  const int w = off[s + 1] - off[s];
  const int32_t len = (r < nrows) ? rp[r + 1] - rp[r] : 0;
  const int32_t b0 = (r < nrows) ? rp[r] : 0;
  for (int j = 0; j < w; ++j)
    for (int q = 0; q < 9; ++q)
      out[(((int64_t)off[s] + j) * 9 + q) * 64 + lane] = (j < len) ? v[9 * ((int64_t)b0 + j) + q] : 0.f;
}

// SELL copy of FP32 values given in CSR order (same slice layout as A's).
__global__ void k_sell_fill_f32(int32_t nrows, const int32_t* __restrict__ rp, const float* __restrict__ v,
                                const int32_t* __restrict__ off, float* __restrict__ out)
{
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t rows_pad = ((int64_t)nrows + 63) / 64 * 64;
  if (r >= rows_pad)
    return;
  const int64_t s = r >> 6;
  const int lane = (int)(r & 63);
  const int w = off[s + 1] - off[s];
  const int32_t len = (r < nrows) ? rp[r + 1] - rp[r] : 0;
  const int32_t b0 = (r < nrows) ? rp[r] : 0;
  for (int j = 0; j < w; ++j)
    out[((int64_t)off[s] + j) * 64 + lane] = (j < len) ? v[b0 + j] : 0.f;
}

} // namespace afem

// ---- workspace --------------------------------------------------------------------------------

enum SpmvFormat
{
  kFmtCsr = 0,
  kFmtSell = 1,
  kFmtSell16 = 2, // SELL-64 with 16-bit column offsets
  kFmtSell16x8 = 3 // the same, 8 slots per step
};

struct afem_solver_ws
{
  int32_t n = 0;
  int block = 1;
  const int32_t* rp = nullptr; // pattern the derived data was built from
  const int32_t* ci = nullptr;
  int64_t nnz = 0;
  afem::DBuf<double> r64, z64, p64, q64, dinv64;
  afem::DBuf<float> r32, z32, p32, q32, dinv32, vals32;
  afem::DBuf<double> parts;
  afem::DBuf<afem::State> st;
  afem::State* host_st = nullptr; // pinned
  // SELL copy
  afem::DBuf<int32_t> sell_off, sell_col;
  afem::DBuf<int16_t> sell_dcol; // col - row, when every offset fits 16 bits
  bool has_dcol = false;
  afem::DBuf<double> sell_v64;
  afem::DBuf<float> sell_v32;
  bool has_fp32 = false;
  bool diff32 = false; // the FP32 copies are in difference form
  afem::DBuf<int32_t> missing; // set by the difference-form copy when a row lacks its diagonal entry
  bool values_valid = false;
  bool pattern_valid = false;
  int format = -1; // chosen SpMV format
  int lanes = 16;
  // SELL16 kernels: workgroup size of the FP64 SpMV, and slots per step and
  // workgroup size of the FP32 SpMV and the preconditioner (unr32 = 0: not
  // tuned, as the FP64 choice)
  int bs64 = 256, unr32 = 0, bs32 = 256;
  bool forced_format = false;
  // sparse preconditioner matrix (pattern of A, FP32), CSR and SELL order
  afem::DBuf<float> spai32, spai_sell32;
  afem::DBuf<uint32_t> spai_sell_pk; // BF16 values packed with the 16-bit column offsets, 4 slots per lane together
  size_t pk_words = 0;                // its size: the 4-slot layout's slots times 64 lanes
  afem::DBuf<int32_t> sell_off4;     // slice offsets of that layout (widths rounded up to multiples of 4)
  bool has_spai = false;
  bool spai_sell_valid = false;
  bool spai_bf16_valid = false;
  ~afem_solver_ws()
  {
    if (host_st)
      (void)hipHostFree(host_st);
  }
};

namespace afem {

void system_release_workspace(afem_system* s)
{
  delete s->ws;
  s->ws = nullptr;
}

static int solver_blocks(int64_t units)
{
  int64_t tiles = (units + kBlock - 1) / kBlock;
  if (tiles < 1)
    tiles = 1;
  return (int)(tiles < kMaxBlocks ? tiles : kMaxBlocks);
}

static Parts parts_of(afem_solver_ws* w)
{
  double* b = w->parts.p;
  Parts P;
  P.pq = b;
  P.rz[0] = b + kMaxBlocks;
  P.rz[1] = b + 2 * kMaxBlocks;
  P.rzru[0] = b + 3 * kMaxBlocks;
  P.rzru[1] = b + 4 * kMaxBlocks;
  P.rr = b + 5 * kMaxBlocks;
  P.rr64 = b + 6 * kMaxBlocks;
  P.bb = b + 7 * kMaxBlocks;
  P.zq = b + 8 * kMaxBlocks;
  return P;
}

static afem_solver_ws* workspace(afem_system* s)
{
  system_require_matrix(s);
  if (!s->ws)
    s->ws = new afem_solver_ws();
  afem_solver_ws* w = s->ws;
  const int32_t n = s->nb_dof;
  if (w->n != n || w->block != s->block) {
    w->n = n;
    w->block = s->block;
    w->r64.alloc(n);
    w->z64.alloc(n);
    w->p64.alloc(n);
    w->q64.alloc(n);
    w->dinv64.alloc((size_t)n * s->block);
    w->dinv32.alloc((size_t)n * s->block);
    w->has_fp32 = false;
    w->values_valid = false;
    w->pattern_valid = false;
    w->format = -1;
    w->bs64 = 256;
    w->unr32 = 0;
    w->has_spai = false;
    w->spai_sell_valid = false;
  }
  if (!w->parts.p) {
    w->parts.alloc((size_t)9 * kMaxBlocks);
    w->parts.zero();
    w->st.alloc(1);
    AFEM_HIP(hipHostMalloc(reinterpret_cast<void**>(&w->host_st), sizeof(State), hipHostMallocDefault));
  }
  if (w->rp != s->row_index || w->ci != s->columns || w->nnz != s->nnz || s->pattern_changed) {
    s->pattern_changed = false;
    w->rp = s->row_index;
    w->ci = s->columns;
    w->nnz = s->nnz;
    w->pattern_valid = false;
    w->values_valid = false;
    w->format = -1;
    w->bs64 = 256;
    w->unr32 = 0;
    w->has_spai = false;
    w->spai_sell_valid = false;
  }
  if (s->matrix_changed) {
    w->values_valid = false;
    s->matrix_changed = false;
  }
  return w;
}

// SELL-64 pattern (slice offsets and columns) from the system's CSR/BSR view.
static void build_sell_pattern(afem_system* s, afem_solver_ws* w)
{
  const int32_t nrows = s->nb_dof / s->block;
  const int32_t nslice = (nrows + 63) / 64;
  DBuf<int32_t> width((size_t)nslice);
  hipLaunchKernelGGL(k_sell_width, dim3(grid_for(nslice, 256)), dim3(256), 0, stream(), nrows, s->row_index, width.p);
  check_launch("sell_width");
  w->sell_off.alloc((size_t)nslice + 1);
  w->sell_off4.release(); // the 4-slot layout follows the slices
  const int64_t cols = exclusive_scan_i32(width.p, w->sell_off.p, nslice);
  AFEM_REQUIRE(cols * 64 * s->block * s->block < (1LL << 31), AFEM_ERR_ARGUMENT, "matrix too large for SELL offsets");
  w->sell_col.alloc((size_t)cols * 64);
  w->sell_v64.alloc((size_t)cols * 64 * s->block * s->block);
  w->sell_v32.release();
}

static void fill_sell(afem_system* s, afem_solver_ws* w, bool columns)
{
  const int32_t nrows = s->nb_dof / s->block;
  const int64_t rows_pad = ((int64_t)nrows + 63) / 64 * 64;
  // the difference-form FP32 values come from their CSR-order copy
  float* v32 = (w->has_fp32 && !w->diff32) ? w->sell_v32.p : nullptr;
  if (w->has_fp32 && w->diff32) {
    if (s->block == 3)
      hipLaunchKernelGGL(k_sell_fill_f32_b3, dim3(grid_for(rows_pad, 256)), dim3(256), 0, stream(), nrows,
                         s->row_index, w->vals32.p, w->sell_off.p, w->sell_v32.p);
    else
      hipLaunchKernelGGL(k_sell_fill_f32, dim3(grid_for(rows_pad, 256)), dim3(256), 0, stream(), nrows,
                         s->row_index, w->vals32.p, w->sell_off.p, w->sell_v32.p);
    check_launch("sell_fill_f32");
  }
  if (s->block == 3)
    hipLaunchKernelGGL(k_sell_fill<3>, dim3(grid_for(rows_pad, 256)), dim3(256), 0, stream(), nrows, s->row_index,
                       s->columns, s->values, w->sell_off.p, w->sell_col.p, w->sell_v64.p, v32, columns ? 1 : 0);
  else
    hipLaunchKernelGGL(k_sell_fill<1>, dim3(grid_for(rows_pad, 256)), dim3(256), 0, stream(), nrows, s->row_index,
                       s->columns, s->values, w->sell_off.p, w->sell_col.p, w->sell_v64.p, v32, columns ? 1 : 0);
  check_launch("sell_fill");
}

// Diagonal (or block-diagonal) inverse, FP32 copies and the SELL copy of
// the current matrix values.
static void ensure_derived(afem_system* s, afem_solver_ws* w, bool need_fp32)
{
  const int32_t n = s->nb_dof;
  if (need_fp32 && !w->has_fp32) {
    w->r32.alloc(n);
    w->z32.alloc(n);
    w->p32.alloc(n);
    w->q32.alloc(n);
    w->vals32.alloc((size_t)s->nnz);
    if (w->sell_v64.p)
      w->sell_v32.alloc(w->sell_v64.n);
    w->has_fp32 = true;
    w->values_valid = false;
  }
  if (w->values_valid)
    return;
  if (s->block == 3) {
    const int32_t nb_node = n / 3;
    hipLaunchKernelGGL(k_block_inverse3, dim3(grid_for(nb_node, 256)), dim3(256), 0, stream(), nb_node, s->row_index,
                       s->columns, s->values, w->dinv64.p, w->dinv32.p);
  }
  else {
    hipLaunchKernelGGL(k_diag_inverse, dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, s->row_index, s->columns,
                       s->values, w->dinv64.p, w->dinv32.p);
  }
  check_launch("jacobi_setup");
  if (w->has_fp32) {
    w->diff32 = false;
    if (s->block == 1 || s->block == 3) {
      if (!w->missing.p)
        w->missing.alloc(1);
      DBuf<int32_t>& missing = w->missing;
      missing.zero();
      if (s->block == 1)
        hipLaunchKernelGGL(k_to_float_rowsum, dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, s->row_index,
                           s->columns, s->values, w->vals32.p, missing.p);
      else
        hipLaunchKernelGGL(k_to_float_rowsum3, dim3(grid_for(n / 3, 256)), dim3(256), 0, stream(), n / 3,
                           s->row_index, s->columns, s->values, w->vals32.p, missing.p);
      check_launch("to_float_rowsum");
      int32_t h = 1;
      missing.download(&h, 1);
      w->diff32 = h == 0;
    }
    if (!w->diff32) {
      hipLaunchKernelGGL(k_to_float, dim3(2048), dim3(256), 0, stream(), (int64_t)s->nnz, s->values, w->vals32.p);
      check_launch("to_float");
    }
  }
  if (w->pattern_valid && w->sell_v64.p) {
    if (w->has_fp32 && !w->sell_v32.p)
      w->sell_v32.alloc(w->sell_v64.n);
    fill_sell(s, w, false);
  }
  w->values_valid = true;
}

static void ensure_sell(afem_system* s, afem_solver_ws* w)
{
  if (w->pattern_valid && w->sell_v64.p)
    return;
  build_sell_pattern(s, w);
  if (w->has_fp32)
    w->sell_v32.alloc(w->sell_v64.n);
  fill_sell(s, w, true);
  w->pattern_valid = true;
  // 16-bit column offsets (scalar matrices whose bandwidth fits)
  w->has_dcol = false;
  w->sell_dcol.release();
  if (s->block == 1) {
    const int32_t nslice = (s->nb_dof + 63) / 64;
    w->sell_dcol.alloc(w->sell_col.n);
    DBuf<int32_t> overflow(1);
    overflow.zero();
    hipLaunchKernelGGL(k_sell_col16, dim3(nslice), dim3(64), 0, stream(), s->nb_dof, w->sell_off.p, w->sell_col.p,
                       w->sell_dcol.p, overflow.p);
    check_launch("sell_col16");
    int32_t h = 0;
    overflow.download(&h, 1);
    w->has_dcol = h == 0;
    if (!w->has_dcol)
      w->sell_dcol.release();
  }
}

// ---- SpMV dispatch ---------------------------------------------------------------------------

template <class VT, class XT, int MODE>
static void spmv_launch(afem_system* s, afem_solver_ws* w, bool fp32, const XT* x, XT* y, const double* b, Gate gate,
                        double* part_out, int nblk, const char* prof_name)
{
  ProfScope prof(prof_name);
  const VT* csr_vals;
  const VT* sell_vals;
  if constexpr (sizeof(VT) == 4) {
    csr_vals = w->vals32.p;
    sell_vals = w->sell_v32.p;
  }
  else {
    csr_vals = s->values;
    sell_vals = w->sell_v64.p;
  }
  (void)fp32;
  // FP32 values in difference form take the difference kernels
  const bool diff = sizeof(VT) == 4 && w->diff32;
  if (s->block == 3) {
    const int32_t nb_node = s->nb_dof / 3;
    if (w->format == kFmtSell) {
      if (diff)
        hipLaunchKernelGGL((k_spmv_sell3<VT, XT, MODE, true>), dim3(nblk), dim3(kBlock), 0, stream(), nb_node,
                           w->sell_off.p, w->sell_col.p, sell_vals, x, y, b, gate, part_out);
      else
        hipLaunchKernelGGL((k_spmv_sell3<VT, XT, MODE>), dim3(nblk), dim3(kBlock), 0, stream(), nb_node,
                           w->sell_off.p, w->sell_col.p, sell_vals, x, y, b, gate, part_out);
    }
    else if (diff)
      hipLaunchKernelGGL((k_spmv_bsr3<16, VT, XT, MODE, true>), dim3(nblk), dim3(kBlock), 0, stream(), nb_node,
                         s->row_index, s->columns, csr_vals, x, y, b, gate, part_out);
    else
      hipLaunchKernelGGL((k_spmv_bsr3<16, VT, XT, MODE>), dim3(nblk), dim3(kBlock), 0, stream(), nb_node,
                         s->row_index, s->columns, csr_vals, x, y, b, gate, part_out);
  }
  else if (w->format == kFmtSell) {
    if (diff)
      hipLaunchKernelGGL((k_spmv_sell<VT, XT, MODE, true>), dim3(nblk), dim3(kBlock), 0, stream(), s->nb_dof,
                         w->sell_off.p, w->sell_col.p, sell_vals, x, y, b, gate, part_out);
    else
      hipLaunchKernelGGL((k_spmv_sell<VT, XT, MODE>), dim3(nblk), dim3(kBlock), 0, stream(), s->nb_dof,
                         w->sell_off.p, w->sell_col.p, sell_vals, x, y, b, gate, part_out);
  }
  else if (w->format == kFmtSell16 || w->format == kFmtSell16x8) {
    const int fmt_unr = w->format == kFmtSell16x8 ? 8 : 4;
    const int unr = (sizeof(VT) == 4 && w->unr32) ? w->unr32 : fmt_unr;
    const int bs = (sizeof(VT) == 4 && w->unr32) ? w->bs32 : w->bs64;
    auto go = [&](auto utag, auto btag) {
      constexpr int U = decltype(utag)::value, BS = decltype(btag)::value;
      if (diff)
        hipLaunchKernelGGL((k_spmv_sell16<VT, XT, MODE, U, BS, true>), dim3(nblk), dim3(BS), 0, stream(),
                           s->nb_dof, w->sell_off.p, w->sell_dcol.p, sell_vals, x, y, b, gate, part_out);
      else
        hipLaunchKernelGGL((k_spmv_sell16<VT, XT, MODE, U, BS>), dim3(nblk), dim3(BS), 0, stream(), s->nb_dof,
                           w->sell_off.p, w->sell_dcol.p, sell_vals, x, y, b, gate, part_out);
    };
    using I4 = std::integral_constant<int, 4>;
    using I8 = std::integral_constant<int, 8>;
    using B256 = std::integral_constant<int, 256>;
    using B512 = std::integral_constant<int, 512>;
    if (unr == 8)
      bs == 512 ? go(I8{}, B512{}) : go(I8{}, B256{});
    else
      bs == 512 ? go(I4{}, B512{}) : go(I4{}, B256{});
  }
  else {
    auto go = [&](auto gtag) {
      constexpr int G = decltype(gtag)::value;
      if (diff)
        hipLaunchKernelGGL((k_spmv_csr<G, VT, XT, MODE, true>), dim3(nblk), dim3(kBlock), 0, stream(), s->nb_dof,
                           s->row_index, s->columns, csr_vals, x, y, b, gate, part_out);
      else
        hipLaunchKernelGGL((k_spmv_csr<G, VT, XT, MODE>), dim3(nblk), dim3(kBlock), 0, stream(), s->nb_dof,
                           s->row_index, s->columns, csr_vals, x, y, b, gate, part_out);
    };
    switch (w->lanes) {
    case 4: go(std::integral_constant<int, 4>{}); break;
    case 8: go(std::integral_constant<int, 8>{}); break;
    case 32: go(std::integral_constant<int, 32>{}); break;
    default: go(std::integral_constant<int, 16>{}); break;
    }
  }
  check_launch(prof_name);
}

static int spmv_blocks(afem_system* s)
{
  return solver_blocks(s->nb_dof / s->block);
}

// Timing of the SpMV candidates: the first of kTuneReps runs warms up, the
// best of the others counts; each run times kTuneLaunches back-to-back
// launches, so a single launch's jitter (a few us against 50-80 us at a
// million rows) does not pick a slower shape.
constexpr int kTuneReps = 5;
constexpr int kTuneLaunches = 3;

// Picks the SpMV format: times FP64 SpMV with each candidate (best of the
// timed runs) and keeps the fastest.  A positive spmv_lanes forces CSR with that
// many lanes per row, -1 forces SELL-64.
static void choose_format(afem_system* s, afem_solver_ws* w, int forced)
{
  if (forced != 0)
    w->forced_format = true; // until the next autotuning
  if (forced > 0) {
    w->format = kFmtCsr;
    w->lanes = forced;
    return;
  }
  ensure_sell(s, w);
  if (forced <= -2 && forced >= -5 && w->has_dcol) {
    w->format = (forced == -2 || forced == -4) ? kFmtSell16 : kFmtSell16x8;
    w->bs64 = forced <= -4 ? 512 : 256;
    w->unr32 = 0;
    return;
  }
  if (forced < 0) {
    w->format = kFmtSell;
    return;
  }
  if (w->format >= 0)
    return;
  struct Cand
  {
    int format, lanes, bs;
  };
  std::vector<Cand> cands;
  if (s->block == 3)
    cands = { { kFmtCsr, 16, 256 }, { kFmtSell, 0, 256 } };
  else
    cands = { { kFmtCsr, 8, 256 }, { kFmtCsr, 16, 256 }, { kFmtSell, 0, 256 } };
  if (s->block == 1 && w->has_dcol) {
    for (int bs : { 256, 512 }) {
      cands.push_back({ kFmtSell16, 0, bs });
      cands.push_back({ kFmtSell16x8, 0, bs });
    }
  }
  hipEvent_t e0, e1;
  AFEM_HIP(hipEventCreate(&e0));
  AFEM_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  Cand pick = cands.back();
  const Gate always{ w->st.p, kAlways, 0, nullptr, 0 };
  w->unr32 = 0;
  w->forced_format = false;
  for (const Cand& c : cands) {
    w->format = c.format;
    w->lanes = c.lanes;
    w->bs64 = c.bs;
    float t_best = 1e30f;
    for (int rep = 0; rep < kTuneReps; ++rep) {
      AFEM_HIP(hipEventRecord(e0, stream()));
      for (int l = 0; l < kTuneLaunches; ++l)
        spmv_launch<double, double, kPlain>(s, w, false, w->p64.p, w->q64.p, nullptr, always, nullptr,
                                            spmv_blocks(s), "spmv_tune");
      AFEM_HIP(hipEventRecord(e1, stream()));
      AFEM_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      AFEM_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (rep > 0 && ms < t_best)
        t_best = ms;
    }
    if (t_best < best) {
      best = t_best;
      pick = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  w->format = pick.format;
  w->lanes = pick.lanes;
  w->bs64 = pick.bs;
}

// The FP32 SELL16 SpMV (mixed solvers, preconditioner) is tuned on its own:
// slots per step and workgroup size, timed like choose_format.  Needs the FP32
// values; after a forced format (the last choose_format call) the FP32 kernels
// keep the FP64 shape.
static void choose_fp32(afem_system* s, afem_solver_ws* w)
{
  if (w->forced_format || w->unr32 || (w->format != kFmtSell16 && w->format != kFmtSell16x8) || !w->sell_v32.p)
    return;
  hipEvent_t e0, e1;
  AFEM_HIP(hipEventCreate(&e0));
  AFEM_HIP(hipEventCreate(&e1));
  const Gate always{ w->st.p, kAlways, 0, nullptr, 0 };
  float best = 1e30f;
  int pick_u = 4, pick_b = 256;
  for (int u : { 4, 8 })
    for (int bs : { 256, 512 }) {
      w->unr32 = u;
      w->bs32 = bs;
      float t_best = 1e30f;
      for (int rep = 0; rep < kTuneReps; ++rep) {
        AFEM_HIP(hipEventRecord(e0, stream()));
        for (int l = 0; l < kTuneLaunches; ++l)
          spmv_launch<float, float, kPlain>(s, w, true, w->p32.p, w->q32.p, nullptr, always, nullptr,
                                            spmv_blocks(s), "spmv_tune_fp32");
        AFEM_HIP(hipEventRecord(e1, stream()));
        AFEM_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        AFEM_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (rep > 0 && ms < t_best)
          t_best = ms;
      }
      if (t_best < best) {
        best = t_best;
        pick_u = u;
        pick_b = bs;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  w->unr32 = pick_u;
  w->bs32 = pick_b;
}

// ---- solve ----------------------------------------------------------------------------------

template <int NB>
static void launch_init(afem_system* s, afem_solver_ws* w, bool mixed, int precond, int pnorm, const Parts& P,
                        int nblk)
{
  const int64_t nunits = s->nb_dof / NB;
  if (mixed)
    hipLaunchKernelGGL((k_init<NB, float>), dim3(nblk), dim3(kBlock), 0, stream(), nunits, w->r64.p, s->rhs.p,
                       w->r32.p, w->z32.p, w->p32.p, w->dinv32.p, precond, pnorm, P);
  else
    hipLaunchKernelGGL((k_init<NB, double>), dim3(nblk), dim3(kBlock), 0, stream(), nunits, w->r64.p, s->rhs.p,
                       w->r64.p, w->z64.p, w->p64.p, w->dinv64.p, precond, pnorm, P);
  check_launch("pcg_init");
  hipLaunchKernelGGL(k_init_finalize, dim3(1), dim3(kBlock), 0, stream(), P, nblk, w->st.p);
  check_launch("pcg_init_finalize");
}

template <int NB>
static void launch_iteration(afem_system* s, afem_solver_ws* w, int k, bool mixed, int precond, const Parts& P,
                             int nblk)
{
  const int64_t nunits = s->nb_dof / NB;
  const Gate until_done{ w->st.p, kUntilDone, k, nullptr, nblk };
  if (!mixed) {
    spmv_launch<double, double, kPQ>(s, w, false, w->p64.p, w->q64.p, nullptr, until_done, P.pq, nblk, "spmv_fp64");
    {
      ProfScope prof("pcg_update_fp64");
      hipLaunchKernelGGL((k_update<NB, double>), dim3(nblk), dim3(kBlock), 0, stream(), k, nunits, s->x.p, w->r64.p,
                         w->z64.p, w->p64.p, w->q64.p, w->dinv64.p, precond, P, nblk, w->st.p);
      check_launch("pcg_update");
    }
    {
      ProfScope prof("pcg_pupdate_fp64");
      hipLaunchKernelGGL((k_pupdate<NB, double>), dim3(nblk), dim3(kBlock), 0, stream(), k, 0, nunits, w->p64.p,
                         w->z64.p, w->r64.p, (const double*)nullptr, w->dinv64.p, precond, P, nblk, w->st.p);
      check_launch("pcg_pupdate");
    }
  }
  else {
    spmv_launch<float, float, kPQ>(s, w, true, w->p32.p, w->q32.p, nullptr, until_done, P.pq, nblk, "spmv_fp32");
    {
      ProfScope prof("pcg_update_fp32");
      hipLaunchKernelGGL((k_update<NB, float>), dim3(nblk), dim3(kBlock), 0, stream(), k, nunits, s->x.p, w->r32.p,
                         w->z32.p, w->p32.p, w->q32.p, w->dinv32.p, precond, P, nblk, w->st.p);
      check_launch("pcg_update_mixed");
    }
    const Gate reliable{ w->st.p, kReliable, k, P.rr, nblk };
    spmv_launch<double, double, kResidual>(s, w, false, s->x.p, w->r64.p, s->rhs.p, reliable, P.rr64, nblk,
                                           "spmv_residual_fp64");
    {
      ProfScope prof("pcg_pupdate_fp32");
      hipLaunchKernelGGL((k_pupdate<NB, float>), dim3(nblk), dim3(kBlock), 0, stream(), k, 1, nunits, w->p32.p,
                         w->z32.p, w->r32.p, (const double*)w->r64.p, w->dinv32.p, precond, P, nblk, w->st.p);
      check_launch("pcg_pupdate_mixed");
    }
  }
}

static void read_state(afem_solver_ws* w)
{
  AFEM_HIP(hipMemcpyAsync(w->host_st, w->st.p, sizeof(State), hipMemcpyDeviceToHost, stream()));
  AFEM_HIP(hipStreamSynchronize(stream()));
}

// ||b - A x||^2 with the FP64 matrix, into st->rr_true (r64 overwritten).
static double true_rr(afem_system* s, afem_solver_ws* w, const Parts& P, int nblk)
{
  const Gate always{ w->st.p, kAlways, 0, nullptr, nblk };
  spmv_launch<double, double, kResidual>(s, w, false, s->x.p, w->r64.p, s->rhs.p, always, P.rr64, nblk,
                                         "spmv_residual_fp64");
  hipLaunchKernelGGL(k_true_residual_finalize, dim3(1), dim3(kBlock), 0, stream(), P, nblk, w->st.p);
  check_launch("true_residual_finalize");
  read_state(w);
  return w->host_st->rr_true;
}

static void ensure_spai_sell(afem_system* s, afem_solver_ws* w, bool bf16)
{
  if (w->format != kFmtSell && w->format != kFmtSell16 && w->format != kFmtSell16x8)
    return;
  if (!w->spai_sell_valid) {
    if (w->spai_sell32.n != w->sell_col.n) // a new M on the same pattern reuses the buffers
      w->spai_sell32.alloc(w->sell_col.n);
    const int64_t rows_pad = ((int64_t)s->nb_dof + 63) / 64 * 64;
    hipLaunchKernelGGL(k_sell_fill_f32, dim3(grid_for(rows_pad, 256)), dim3(256), 0, stream(), s->nb_dof,
                       s->row_index, w->spai32.p, w->sell_off.p, w->spai_sell32.p);
    check_launch("spai_sell_fill");
    w->spai_sell_valid = true;
    w->spai_bf16_valid = false;
  }
  if (bf16 && !w->spai_bf16_valid) {
    AFEM_REQUIRE(w->has_dcol, AFEM_ERR_STATE, "BF16 preconditioner values need the 16-bit offset layout");
    const int64_t nslice = ((int64_t)s->nb_dof + 63) / 64;
    if (w->sell_off4.n != (size_t)nslice + 1) { // the 4-slot layout's slice offsets, once per pattern
      std::vector<int32_t> o((size_t)nslice + 1), o4((size_t)nslice + 1, 0);
      w->sell_off.download(o.data(), o.size());
      for (int64_t k = 0; k < nslice; ++k)
        o4[k + 1] = o4[k] + (o[k + 1] - o[k] + 3) / 4 * 4;
      w->sell_off4.alloc(o4.size());
      w->sell_off4.upload(o4.data(), o4.size());
      w->pk_words = (size_t)o4[(size_t)nslice] * 64;
    }
    if (w->spai_sell_pk.n != w->pk_words)
      w->spai_sell_pk.alloc(w->pk_words);
    hipLaunchKernelGGL(k_pack_bf16, dim3((unsigned)nslice), dim3(64), 0, stream(), w->sell_off.p, w->sell_off4.p,
                       w->spai_sell32.p, w->sell_dcol.p, w->spai_sell_pk.p);
    check_launch("spai_to_bf16");
    w->spai_bf16_valid = true;
  }
}

template <int INIT>
static void prec_launch(afem_system* s, afem_solver_ws* w, int k, const Parts& P, int nblk, bool bf16)
{
  ProfScope prof("fpcg_precond_spmv");
  const PrecIn in{ w->r32.p, w->r64.p, true };
  if (w->format == kFmtSell16 || w->format == kFmtSell16x8) {
    // workgroup size measured at 128^3 in the learned solve: the packed BF16
    // words run faster with 512 threads (104 against 108 ms), the FP32
    // values with 256 (123 against 131 ms)
    const int bs = bf16 ? 512 : 256;
    auto go = [&](auto btag) {
      constexpr int BS = decltype(btag)::value;
      if (bf16) // the packed words carry their offsets, in the 4-slot layout
        hipLaunchKernelGGL((k_prec_sell16<INIT, uint32_t, BS>), dim3(nblk), dim3(BS), 0, stream(), k, s->nb_dof,
                           w->sell_off4.p, w->sell_dcol.p, w->spai_sell_pk.p, in, w->r32.p, w->z32.p, w->p32.p,
                           w->q32.p, s->rhs.p, P, nblk, w->st.p);
      else
        hipLaunchKernelGGL((k_prec_sell16<INIT, float, BS>), dim3(nblk), dim3(BS), 0, stream(), k, s->nb_dof,
                           w->sell_off.p, w->sell_dcol.p, w->spai_sell32.p, in, w->r32.p, w->z32.p, w->p32.p,
                           w->q32.p, s->rhs.p, P, nblk, w->st.p);
    };
    if (bs == 512)
      go(std::integral_constant<int, 512>{});
    else
      go(std::integral_constant<int, 256>{});
  }
  else if (w->format == kFmtSell) {
    hipLaunchKernelGGL((k_prec_sell<INIT>), dim3(nblk), dim3(kBlock), 0, stream(), k, s->nb_dof, w->sell_off.p,
                       w->sell_col.p, w->spai_sell32.p, in, w->r32.p, w->z32.p, w->p32.p, w->q32.p, s->rhs.p, P, nblk,
                       w->st.p);
  }
  else {
    auto go = [&](auto gtag) {
      constexpr int G = decltype(gtag)::value;
      hipLaunchKernelGGL((k_prec_csr<G, INIT>), dim3(nblk), dim3(kBlock), 0, stream(), k, s->nb_dof, s->row_index,
                         s->columns, w->spai32.p, in, w->r32.p, w->z32.p, w->p32.p, w->q32.p, s->rhs.p, P, nblk,
                         w->st.p);
    };
    switch (w->lanes) {
    case 4: go(std::integral_constant<int, 4>{}); break;
    case 8: go(std::integral_constant<int, 8>{}); break;
    case 32: go(std::integral_constant<int, 32>{}); break;
    default: go(std::integral_constant<int, 16>{}); break;
    }
  }
  check_launch("fpcg_precond_spmv");
}

static void launch_flex_init(afem_system* s, afem_solver_ws* w, const Parts& P, int nblk, bool bf16)
{
  prec_launch<1>(s, w, -1, P, nblk, bf16);
  hipLaunchKernelGGL(k_init_finalize, dim3(1), dim3(kBlock), 0, stream(), P, nblk, w->st.p);
  check_launch("fpcg_init_finalize");
}

static void launch_flex_iteration(afem_system* s, afem_solver_ws* w, int k, const Parts& P, int nblk, bool bf16)
{
  const Gate until_done{ w->st.p, kUntilDone, k, nullptr, nblk };
  spmv_launch<float, float, kPQ>(s, w, true, w->p32.p, w->q32.p, nullptr, until_done, P.pq, nblk, "spmv_fp32");
  {
    ProfScope prof("fpcg_update");
    hipLaunchKernelGGL(k_update_flex, dim3(nblk), dim3(kBlock), 0, stream(), k, s->nb_dof, s->x.p, w->r32.p,
                       w->p32.p, w->q32.p, P, nblk, w->st.p);
    check_launch("fpcg_update");
  }
  const Gate reliable{ w->st.p, kReliable, k, P.rr, nblk };
  spmv_launch<double, double, kResidual>(s, w, false, s->x.p, w->r64.p, s->rhs.p, reliable, P.rr64, nblk,
                                         "spmv_residual_fp64");
  prec_launch<0>(s, w, k, P, nblk, bf16);
  {
    ProfScope prof("fpcg_pupdate");
    hipLaunchKernelGGL(k_pupdate_flex, dim3(nblk), dim3(kBlock), 0, stream(), k, s->nb_dof, w->p32.p, w->z32.p, P,
                       nblk, w->st.p);
    check_launch("fpcg_pupdate");
  }
}

static int solve(afem_system* s, const afem_solver_options& o, afem_solve_result* res)
{
  AFEM_REQUIRE(o.method == AFEM_SOLVER_PCG_FP64 || o.method == AFEM_SOLVER_PCG_MIXED ||
                 o.method == AFEM_SOLVER_CG_FP64 || o.method == AFEM_SOLVER_FPCG_SPAI,
               AFEM_ERR_ARGUMENT, "unknown solver method");
  AFEM_REQUIRE(o.preconditioner == AFEM_PRECOND_JACOBI || o.preconditioner == AFEM_PRECOND_NONE, AFEM_ERR_ARGUMENT,
               "unknown preconditioner");
  AFEM_REQUIRE(o.rtol >= 0.0 && o.atol >= 0.0 && o.max_iter >= 0, AFEM_ERR_ARGUMENT, "bad tolerances");
  AFEM_REQUIRE(o.norm >= 0 && o.norm <= 2, AFEM_ERR_ARGUMENT, "unknown stopping norm");
  AFEM_REQUIRE(o.norm == 0 || (o.method != AFEM_SOLVER_PCG_MIXED && o.method != AFEM_SOLVER_FPCG_SPAI),
               AFEM_ERR_NOT_IMPLEMENTED,
               "the natural-norm and preconditioned-norm stopping tests are available for the FP64 solvers");
  // The reference's row-column elimination never touches column 0
  // (femutils/CsrDoFLinearSystemImpl.cc:111), so with DoF 0 eliminated the
  // matrix keeps A[i,0] in every row while the RHS is lifted by A[0,i] g
  // (femutils/DoFLinearSystemImplBase.cc:65-87): a non-symmetric system
  // that no conjugate-gradient method solves.
  AFEM_REQUIRE(!s->dof0_row_column, AFEM_ERR_NOT_IMPLEMENTED,
               "DoF 0 is row-column eliminated: the transformed matrix keeps column 0 and is not symmetric, "
               "so PCG does not apply (use RowElimination or Penalty for that DoF)");
  auto t_start = std::chrono::steady_clock::now();
  afem_solver_ws* w = workspace(s);
  const bool flex = o.method == AFEM_SOLVER_FPCG_SPAI;
  if (flex) {
    AFEM_REQUIRE(s->block == 1, AFEM_ERR_NOT_IMPLEMENTED, "the sparse preconditioner is for scalar systems");
    AFEM_REQUIRE(w->has_spai, AFEM_ERR_STATE,
                 "no preconditioner values: call afem_system_set_spai or afem_spai_apply after the matrix transformation");
  }
  const bool mixed = o.method == AFEM_SOLVER_PCG_MIXED || flex;
  const int precond = (o.method == AFEM_SOLVER_CG_FP64 || o.preconditioner == AFEM_PRECOND_NONE) ? 0 : 1;
  const int check_every = o.check_every > 0 ? o.check_every : 16;
  const double delta = o.reliable_delta > 0.0 ? o.reliable_delta : 0.1;
  ensure_derived(s, w, mixed);
  choose_format(s, w, o.spmv_lanes);
  ensure_derived(s, w, mixed);
  if (mixed)
    choose_fp32(s, w);
  // BF16 preconditioner values need a SELL layout; CSR keeps FP32
  const bool bf16_used = flex && o.precond_bf16 && (w->format == kFmtSell16 || w->format == kFmtSell16x8);
  if (flex)
    ensure_spai_sell(s, w, bf16_used);
  const Parts P = parts_of(w);
  const int nblk = spmv_blocks(s);

  if (!o.use_initial_guess) {
    hipLaunchKernelGGL(k_initial_guess, dim3(grid_for(s->nb_dof, 256)), dim3(256), 0, stream(), s->nb_dof,
                       s->elim_info.p, s->elim_value.p, s->x.p);
    check_launch("initial_guess");
  }
  State init{};
  init.rtol2 = o.rtol * o.rtol;
  init.atol2 = o.atol * o.atol;
  init.delta2 = delta * delta;
  init.natural = o.norm;
  int total_iter = 0, restarts = 0, nb_ru = 0;
  bool converged = false, breakdown = false;
  double relres = 0.0;
  double bnorm = 0.0;
  for (;;) {
    init.max_iter = o.max_iter - total_iter;
    *w->host_st = init;
    AFEM_HIP(hipMemcpyAsync(w->st.p, w->host_st, sizeof(State), hipMemcpyHostToDevice, stream()));
    const Gate always{ w->st.p, kAlways, 0, nullptr, nblk };
    spmv_launch<double, double, kResidual>(s, w, false, s->x.p, w->r64.p, s->rhs.p, always, P.rr64, nblk,
                                           "spmv_residual_fp64");
    if (flex)
      launch_flex_init(s, w, P, nblk, bf16_used);
    else if (s->block == 3)
      launch_init<3>(s, w, mixed, precond, o.norm == 2, P, nblk);
    else
      launch_init<1>(s, w, mixed, precond, o.norm == 2, P, nblk);
    read_state(w);
    int k = 0;
    while (!w->host_st->done && k < init.max_iter) {
      for (int j = 0; j < check_every && k < init.max_iter; ++j, ++k) {
        if (flex)
          launch_flex_iteration(s, w, k, P, nblk, bf16_used);
        else if (s->block == 3)
          launch_iteration<3>(s, w, k, mixed, precond, P, nblk);
        else
          launch_iteration<1>(s, w, k, mixed, precond, P, nblk);
      }
      read_state(w);
    }
    total_iter += w->host_st->iter;
    nb_ru += w->host_st->nb_ru;
    bnorm = std::sqrt(w->host_st->bb);
    const bool device_converged = w->host_st->converged != 0;
    breakdown = w->host_st->breakdown != 0;
    const double rr = true_rr(s, w, P, nblk);
    relres = bnorm > 0.0 ? std::sqrt(rr) / bnorm : std::sqrt(rr);
    if (o.norm == 2)
      relres = bnorm > 0.0 ? std::sqrt(w->host_st->rr_out) / bnorm : std::sqrt(w->host_st->rr_out);
    if (o.norm != 0) {
      converged = device_converged && std::isfinite(relres);
      break; // the natural and preconditioned tests have no true-residual restart
    }
    converged = relres <= o.rtol || std::sqrt(rr) <= o.atol;
    if (converged || total_iter >= o.max_iter || restarts >= 8 || !std::isfinite(relres) ||
        w->host_st->iter == 0 || (flex && breakdown))
      break;
    // the recursive residual met the tolerance but the true one does not:
    // restart from the current iterate
    ++restarts;
  }
  int fallback = 0;
  if (flex && !converged && breakdown && total_iter < o.max_iter && std::isfinite(relres)) {
    // the preconditioner is not positive on some residual: finish with
    // Jacobi mixed-precision PCG from the current iterate
    afem_solver_options jo = o;
    jo.method = AFEM_SOLVER_PCG_MIXED;
    jo.preconditioner = AFEM_PRECOND_JACOBI;
    jo.use_initial_guess = 1;
    jo.max_iter = o.max_iter - total_iter;
    afem_solve_result jr{};
    const int rc = solve(s, jo, &jr);
    if (rc != AFEM_OK && rc != AFEM_ERR_NOT_CONVERGED)
      return rc;
    total_iter += jr.iterations;
    nb_ru += jr.reliable_updates;
    relres = jr.relres;
    converged = jr.converged != 0;
    fallback = 1;
  }
  AFEM_HIP(hipStreamSynchronize(stream()));
  auto t_end = std::chrono::steady_clock::now();
  if (res) {
    res->iterations = total_iter;
    res->converged = converged ? 1 : 0;
    res->reliable_updates = nb_ru;
    res->relres = relres;
    res->solve_ms = std::chrono::duration<double, std::milli>(t_end - t_start).count();
    res->fallback = fallback;
  }
  return converged ? AFEM_OK : AFEM_ERR_NOT_CONVERGED;
}

// ---- allocation-free FP64 PCG ---------------------------------------------------------------
//
// The FP64 (Jacobi or unpreconditioned) CG of solve() for callers that may not
// allocate: the SpMV runs straight on the system's matrix view (CSR with 16
// lanes per row, or 3x3 BSR with 16 lanes per node row), and r, z, p, q, the
// Jacobi inverse, the reduction partials and the solver state live in
// `scratch`, a fixed device array of this file.  Each device holds its own
// copy of the array until the process ends; every call on a device shares it
// (calls that overlap on one device race).  The host reads the state with
// plain copies into stack variables, so no pinned memory is needed either.

constexpr int64_t kScratchMaxDof = int64_t(1) << 20;
constexpr int64_t kScratchStateDoubles = 32;
static_assert(sizeof(State) <= kScratchStateDoubles * sizeof(double), "State does not fit its scratch slot");
// r, z, p, q (one double per DoF each), the Jacobi inverse (up to 3 doubles
// per DoF for 3x3 blocks), 9 partial arrays, the state
constexpr int64_t kScratchDoubles = 7 * kScratchMaxDof + 9 * kMaxBlocks + kScratchStateDoubles;
static __device__ __align__(256) double scratch[kScratchDoubles]; // 56 MiB per device

constexpr int kScratchMaxDevices = 64;

// Device address of this device's scratch array, looked up once per device
// with that device current (the array's name is only a host-side stand-in).
static double* scratch_on_device()
{
  ensure_init();
  static double* ptr[kScratchMaxDevices] = {};
  const int dev = rt().device;
  AFEM_REQUIRE(dev >= 0 && dev < kScratchMaxDevices, AFEM_ERR_STATE, "device index out of range for the scratch table");
  if (!ptr[dev]) {
    int cur = -1;
    AFEM_HIP(hipGetDevice(&cur));
    AFEM_REQUIRE(cur == dev, AFEM_ERR_STATE, "the library's device is not the current device");
    void* p = nullptr;
    AFEM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    AFEM_REQUIRE(p != nullptr, AFEM_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    ptr[dev] = static_cast<double*>(p);
  }
  return ptr[dev];
}

__global__ void k_fill_zero(int64_t n, double* __restrict__ out)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = 0.0;
}

__global__ void k_state_init(State init, State* __restrict__ st)
{
  if (blockIdx.x == 0 && threadIdx.x == 0)
    *st = init;
}

struct NoAllocWs
{
  double *r, *z, *p, *q, *dinv;
  State* st;
  Parts P;
};

static NoAllocWs noalloc_layout(double* base)
{
  NoAllocWs w;
  w.r = base;
  w.z = base + kScratchMaxDof;
  w.p = base + 2 * kScratchMaxDof;
  w.q = base + 3 * kScratchMaxDof;
  w.dinv = base + 4 * kScratchMaxDof;
  double* parts = base + 7 * kScratchMaxDof;
  w.P.pq = parts;
  w.P.rz[0] = parts + kMaxBlocks;
  w.P.rz[1] = parts + 2 * kMaxBlocks;
  w.P.rzru[0] = parts + 3 * kMaxBlocks;
  w.P.rzru[1] = parts + 4 * kMaxBlocks;
  w.P.rr = parts + 5 * kMaxBlocks;
  w.P.rr64 = parts + 6 * kMaxBlocks;
  w.P.bb = parts + 7 * kMaxBlocks;
  w.P.zq = parts + 8 * kMaxBlocks;
  w.st = reinterpret_cast<State*>(parts + 9 * kMaxBlocks);
  return w;
}

template <int MODE>
static void noalloc_spmv(const afem_system* s, const double* x, double* y, const double* b, Gate gate,
                         double* part_out, int nblk)
{
  if (s->block == 3)
    hipLaunchKernelGGL((k_spmv_bsr3<16, double, double, MODE>), dim3(nblk), dim3(kBlock), 0, stream(),
                       s->nb_dof / 3, s->row_index, s->columns, s->values, x, y, b, gate, part_out);
  else
    hipLaunchKernelGGL((k_spmv_csr<16, double, double, MODE>), dim3(nblk), dim3(kBlock), 0, stream(), s->nb_dof,
                       s->row_index, s->columns, s->values, x, y, b, gate, part_out);
  check_launch("noalloc_spmv");
}

static State noalloc_read_state(const NoAllocWs& w)
{
  State h{};
  AFEM_HIP(hipMemcpyAsync(&h, w.st, sizeof(State), hipMemcpyDeviceToHost, stream()));
  AFEM_HIP(hipStreamSynchronize(stream()));
  return h;
}

template <int NB>
static void noalloc_init(const afem_system* s, const NoAllocWs& w, int precond, int pnorm, int nblk)
{
  hipLaunchKernelGGL((k_init<NB, double>), dim3(nblk), dim3(kBlock), 0, stream(), (int64_t)(s->nb_dof / NB), w.r,
                     s->rhs.p, w.r, w.z, w.p, w.dinv, precond, pnorm, w.P);
  check_launch("noalloc_init");
  hipLaunchKernelGGL(k_init_finalize, dim3(1), dim3(kBlock), 0, stream(), w.P, nblk, w.st);
  check_launch("noalloc_init_finalize");
}

template <int NB>
static void noalloc_iteration(const afem_system* s, const NoAllocWs& w, int k, int precond, int nblk)
{
  const int64_t nunits = s->nb_dof / NB;
  const Gate until_done{ w.st, kUntilDone, k, nullptr, nblk };
  noalloc_spmv<kPQ>(s, w.p, w.q, nullptr, until_done, w.P.pq, nblk);
  hipLaunchKernelGGL((k_update<NB, double>), dim3(nblk), dim3(kBlock), 0, stream(), k, nunits, s->x.p, w.r, w.z,
                     w.p, w.q, w.dinv, precond, w.P, nblk, w.st);
  check_launch("noalloc_update");
  hipLaunchKernelGGL((k_pupdate<NB, double>), dim3(nblk), dim3(kBlock), 0, stream(), k, 0, nunits, w.p, w.z, w.r,
                     (const double*)nullptr, w.dinv, precond, w.P, nblk, w.st);
  check_launch("noalloc_pupdate");
}

static int solve_noalloc(afem_system* s, const afem_solver_options& o, afem_solve_result* res)
{
  AFEM_REQUIRE(o.method == AFEM_SOLVER_PCG_FP64 || o.method == AFEM_SOLVER_CG_FP64, AFEM_ERR_NOT_IMPLEMENTED,
               "the allocation-free solve runs the FP64 methods (AFEM_SOLVER_PCG_FP64, AFEM_SOLVER_CG_FP64)");
  AFEM_REQUIRE(o.preconditioner == AFEM_PRECOND_JACOBI || o.preconditioner == AFEM_PRECOND_NONE, AFEM_ERR_ARGUMENT,
               "unknown preconditioner");
  AFEM_REQUIRE(o.rtol >= 0.0 && o.atol >= 0.0 && o.max_iter >= 0, AFEM_ERR_ARGUMENT, "bad tolerances");
  AFEM_REQUIRE(o.norm >= 0 && o.norm <= 2, AFEM_ERR_ARGUMENT, "unknown stopping norm");
  AFEM_REQUIRE((int64_t)s->nb_dof <= kScratchMaxDof, AFEM_ERR_ARGUMENT,
               "the allocation-free solve takes at most 1048576 DoFs (its fixed scratch array holds 7 vectors of "
               "that length); use afem_system_solve for larger systems");
  system_require_matrix(s);
  AFEM_REQUIRE(s->block == 1 || s->block == 3, AFEM_ERR_NOT_IMPLEMENTED, "block size must be 1 or 3");
  AFEM_REQUIRE(!s->dof0_row_column, AFEM_ERR_NOT_IMPLEMENTED,
               "DoF 0 is row-column eliminated: the transformed matrix keeps column 0 and is not symmetric, "
               "so PCG does not apply (use RowElimination or Penalty for that DoF)");
  auto t_start = std::chrono::steady_clock::now();
  const NoAllocWs w = noalloc_layout(scratch_on_device());
  const int precond = (o.method == AFEM_SOLVER_CG_FP64 || o.preconditioner == AFEM_PRECOND_NONE) ? 0 : 1;
  const int check_every = o.check_every > 0 ? o.check_every : 16;
  const int nblk = spmv_blocks(s);
  // the array keeps the previous call's contents: clear the partials and the
  // state (a fill kernel, no memset node in a captured stream)
  hipLaunchKernelGGL(k_fill_zero, dim3(40), dim3(256), 0, stream(), (int64_t)(9 * kMaxBlocks + kScratchStateDoubles),
                     w.P.pq);
  check_launch("noalloc_clear");
  if (precond) {
    if (s->block == 3)
      hipLaunchKernelGGL(k_block_inverse3, dim3(grid_for(s->nb_dof / 3, 256)), dim3(256), 0, stream(), s->nb_dof / 3,
                         s->row_index, s->columns, s->values, w.dinv, (float*)nullptr);
    else
      hipLaunchKernelGGL(k_diag_inverse, dim3(grid_for(s->nb_dof, 256)), dim3(256), 0, stream(), s->nb_dof,
                         s->row_index, s->columns, s->values, w.dinv, (float*)nullptr);
    check_launch("noalloc_jacobi_setup");
  }
  if (!o.use_initial_guess) {
    hipLaunchKernelGGL(k_initial_guess, dim3(grid_for(s->nb_dof, 256)), dim3(256), 0, stream(), s->nb_dof,
                       s->elim_info.p, s->elim_value.p, s->x.p);
    check_launch("noalloc_initial_guess");
  }
  State init{};
  init.rtol2 = o.rtol * o.rtol;
  init.atol2 = o.atol * o.atol;
  init.delta2 = 0.01;
  init.natural = o.norm;
  int total_iter = 0, restarts = 0;
  bool converged = false;
  double relres = 0.0;
  for (;;) {
    init.max_iter = o.max_iter - total_iter;
    hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, stream(), init, w.st);
    check_launch("noalloc_state_init");
    const Gate always{ w.st, kAlways, 0, nullptr, nblk };
    noalloc_spmv<kResidual>(s, s->x.p, w.r, s->rhs.p, always, w.P.rr64, nblk);
    if (s->block == 3)
      noalloc_init<3>(s, w, precond, o.norm == 2, nblk);
    else
      noalloc_init<1>(s, w, precond, o.norm == 2, nblk);
    State h = noalloc_read_state(w);
    int k = 0;
    while (!h.done && k < init.max_iter) {
      for (int j = 0; j < check_every && k < init.max_iter; ++j, ++k) {
        if (s->block == 3)
          noalloc_iteration<3>(s, w, k, precond, nblk);
        else
          noalloc_iteration<1>(s, w, k, precond, nblk);
      }
      h = noalloc_read_state(w);
    }
    total_iter += h.iter;
    const double bnorm = std::sqrt(h.bb);
    // ||b - A x||^2 with the FP64 matrix (r overwritten)
    noalloc_spmv<kResidual>(s, s->x.p, w.r, s->rhs.p, always, w.P.rr64, nblk);
    hipLaunchKernelGGL(k_true_residual_finalize, dim3(1), dim3(kBlock), 0, stream(), w.P, nblk, w.st);
    check_launch("noalloc_true_residual");
    const State t = noalloc_read_state(w);
    const double rr = t.rr_true;
    relres = bnorm > 0.0 ? std::sqrt(rr) / bnorm : std::sqrt(rr);
    if (o.norm == 2)
      relres = bnorm > 0.0 ? std::sqrt(h.rr_out) / bnorm : std::sqrt(h.rr_out);
    if (o.norm != 0) {
      converged = h.converged != 0 && std::isfinite(relres);
      break;
    }
    converged = relres <= o.rtol || std::sqrt(rr) <= o.atol;
    if (converged || total_iter >= o.max_iter || restarts >= 8 || !std::isfinite(relres) || h.iter == 0)
      break;
    ++restarts;
  }
  AFEM_HIP(hipStreamSynchronize(stream()));
  auto t_end = std::chrono::steady_clock::now();
  if (res) {
    res->iterations = total_iter;
    res->converged = converged ? 1 : 0;
    res->reliable_updates = 0;
    res->relres = relres;
    res->solve_ms = std::chrono::duration<double, std::milli>(t_end - t_start).count();
    res->fallback = 0;
  }
  return converged ? AFEM_OK : AFEM_ERR_NOT_CONVERGED;
}

} // namespace afem

using namespace afem;

extern "C" {

AFEM_API void afem_solver_default_options(afem_solver_options* opt)
{
  if (!opt)
    return;
  opt->method = AFEM_SOLVER_PCG_FP64;
  opt->preconditioner = AFEM_PRECOND_JACOBI;
  opt->rtol = 1e-8;
  opt->atol = 0.0;
  opt->max_iter = 100000;
  opt->use_initial_guess = 0;
  opt->check_every = 16;
  opt->reliable_delta = 0.1;
  opt->spmv_lanes = 0;
  opt->norm = 0;
  opt->precond_bf16 = 0;
}

AFEM_API int afem_system_solve(afem_system_t sys, const afem_solver_options* opt, afem_solve_result* result)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(sys, AFEM_ERR_ARGUMENT, "null handle");
  afem_solver_options o;
  afem_solver_default_options(&o);
  if (opt)
    o = *opt;
  int rc = solve(sys, o, result);
  if (rc != AFEM_OK) {
    set_last_error("solver reached max_iter before the tolerance");
    return rc;
  }
  AFEM_GUARD_END
}

AFEM_API int afem_system_solve_noalloc(afem_system_t sys, const afem_solver_options* opt, afem_solve_result* result)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(sys, AFEM_ERR_ARGUMENT, "null handle");
  afem_solver_options o;
  afem_solver_default_options(&o);
  if (opt)
    o = *opt;
  int rc = solve_noalloc(sys, o, result);
  if (rc != AFEM_OK) {
    set_last_error("solver reached max_iter before the tolerance");
    return rc;
  }
  AFEM_GUARD_END
}

AFEM_API int afem_system_spmv(afem_system_t sys, const double* x, double* y)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(sys && x && y, AFEM_ERR_ARGUMENT, "null argument");
  afem_solver_ws* w = workspace(sys);
  ensure_derived(sys, w, false);
  choose_format(sys, w, 0);
  ensure_derived(sys, w, false);
  const Gate always{ w->st.p, kAlways, 0, nullptr, 0 };
  spmv_launch<double, double, kPlain>(sys, w, false, x, y, nullptr, always, nullptr, spmv_blocks(sys), "spmv_fp64");
  AFEM_GUARD_END
}

AFEM_API int afem_system_spmv_fp32(afem_system_t sys, const float* x, float* y)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(sys && x && y, AFEM_ERR_ARGUMENT, "null argument");
  afem_solver_ws* w = workspace(sys);
  ensure_derived(sys, w, true);
  choose_format(sys, w, 0);
  ensure_derived(sys, w, true);
  choose_fp32(sys, w);
  const Gate always{ w->st.p, kAlways, 0, nullptr, 0 };
  spmv_launch<float, float, kPlain>(sys, w, true, x, y, nullptr, always, nullptr, spmv_blocks(sys), "spmv_fp32");
  AFEM_GUARD_END
}

AFEM_API int afem_system_set_spai(afem_system_t sys, const float* values)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(sys && values, AFEM_ERR_ARGUMENT, "null argument");
  afem_solver_ws* w = workspace(sys);
  AFEM_REQUIRE(sys->block == 1, AFEM_ERR_NOT_IMPLEMENTED, "the sparse preconditioner is for scalar systems");
  if (w->spai32.n != (size_t)sys->nnz)
    w->spai32.alloc((size_t)sys->nnz);
  AFEM_HIP(hipMemcpyAsync(w->spai32.p, values, sizeof(float) * (size_t)sys->nnz, hipMemcpyDeviceToDevice, stream()));
  AFEM_HIP(hipStreamSynchronize(stream()));
  w->has_spai = true;
  w->spai_sell_valid = false;
  AFEM_GUARD_END
}

AFEM_API int afem_system_spmv_format(afem_system_t sys, int* format, int* lanes)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(sys && format && lanes, AFEM_ERR_ARGUMENT, "null argument");
  afem_solver_ws* w = sys->ws;
  *format = w ? w->format : -1;
  *lanes = w ? w->lanes : 0;
  AFEM_GUARD_END
}

AFEM_API int afem_system_spmv_tuning(afem_system_t sys, int* fp64_slots, int* fp64_block, int* fp32_slots,
                                     int* fp32_block)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(sys && fp64_slots && fp64_block && fp32_slots && fp32_block, AFEM_ERR_ARGUMENT, "null argument");
  afem_solver_ws* w = sys->ws;
  const bool sell16 = w && (w->format == kFmtSell16 || w->format == kFmtSell16x8);
  *fp64_slots = sell16 ? (w->format == kFmtSell16x8 ? 8 : 4) : 0;
  *fp64_block = sell16 ? w->bs64 : 0;
  *fp32_slots = sell16 ? (w->unr32 ? w->unr32 : *fp64_slots) : 0;
  *fp32_block = sell16 ? (w->unr32 ? w->bs32 : w->bs64) : 0;
  AFEM_GUARD_END
}

} // extern "C"
