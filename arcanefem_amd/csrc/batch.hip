// batch.hip -- many independent scalar systems solved together.
//
// No counterpart in the reference, which solves one system per module
// instance (femutils/DoFLinearSystem.cc:154-156 and the Hypre / PETSc
// back-ends).  The systems are the diagonal blocks of one block-diagonal
// matrix -- typically the Laplace matrix of the disjoint union of several
// meshes, so the existing pattern and assembly kernels build all of them in
// one launch each.  This file solves them with per-system scalars:
// mixed-precision Jacobi PCG (FP32 matrix, r, z, p, q; FP64 x, FP64
// reliable-update residual r = b - A x, which alone declares convergence).
//
// Each workgroup owns rows of exactly one system (a segment), so the
// fence-free reduction scheme of solver.hip carries over per segment:
// producers write one partial per workgroup, every consumer workgroup of a
// segment re-reduces that segment's partials in the same order.  Kernels
// take the iteration parity as an argument and read the absolute iteration
// number from the segment state (parity-indexed, written one iteration
// ahead), so a chunk of iterations can be captured once in a hipGraph and
// replayed; converged segments make their workgroups return at once.
//
// Tuning: lanes per row (G) and row tiles per workgroup (T) are timed on
// the batch and the fastest pair is kept; G = 1 selects a SELL-64 copy of
// the matrix whose slices start at every segment's first row (16-bit column
// offsets, one thread per row).  Between chunks the grid is cut down to the
// workgroups of the segments still running (through a workgroup list), so
// the tail of the longest solves does not pay for the finished ones.
#include "common.h"
#include "device_utils.h"
#include "linsys.h"

#include <algorithm>
#include <type_traits>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace afem {
namespace batch {

constexpr int kBT = 256;

struct SegState
{
  double rtol2, atol2, delta2, tol2, bb, rr_out, rr_true;
  double rr_ref[2];
  int iter_par[2]; // iteration number seen by kernels of parity q
  int ru[2];
  int max_iter, iter, done, converged, breakdown, nb_ru, active;
};

struct BParts
{
  double* pq;
  double* rz[2];
  double* rzru[2];
  double* rr;
  double* rr64;
  double* bb;
};

struct Seg
{
  const int32_t* rows;   // nb_seg + 1 row offsets
  const int32_t* blocks; // nb_seg + 1 workgroup offsets
  const int32_t* owner;  // segment of each workgroup
  int rows_per_block;
  const int32_t* list;   // launched workgroup -> workgroup of the layout (nullptr: the same)
  // segment-aligned SELL-64: first slice of each segment, slot offset of each
  // slice (nb_slice + 1), column offsets col - row
  const int32_t* sslice;
  const int32_t* soff;
  const int16_t* sdcol;
};

struct Ctx
{
  int s;          // segment
  int gb;         // workgroup of the layout
  int b0, b1;     // the segment's workgroups
  int32_t r0, r1; // rows of this workgroup
};

__device__ __forceinline__ Ctx context(const Seg& g)
{
  Ctx c;
  c.gb = g.list ? g.list[blockIdx.x] : (int)blockIdx.x;
  c.s = g.owner[c.gb];
  c.b0 = g.blocks[c.s];
  c.b1 = g.blocks[c.s + 1];
  const int32_t sr0 = g.rows[c.s], sr1 = g.rows[c.s + 1];
  c.r0 = sr0 + (int32_t)(c.gb - c.b0) * g.rows_per_block;
  c.r1 = min(sr1, c.r0 + g.rows_per_block);
  return c;
}

// sum of the segment's partials p[b0..b1), identical in every workgroup
__device__ __forceinline__ double seg_sum(const double* __restrict__ p, int b0, int b1)
{
  __shared__ double lds[kBT / kWave];
  double s = 0.0;
  for (int i = b0 + threadIdx.x; i < b1; i += kBT)
    s += p[i];
  s = wave_reduce_sum(s);
  if ((threadIdx.x & (kWave - 1)) == 0)
    lds[threadIdx.x / kWave] = s;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < kBT / kWave; ++w)
    t += lds[w];
  __syncthreads();
  return t;
}

// out[gb] = the sum of v over the workgroup
__device__ __forceinline__ void put(double v, double* __restrict__ out, int gb)
{
  __shared__ double lds[kBT / kWave];
  v = wave_reduce_sum(v);
  if ((threadIdx.x & (kWave - 1)) == 0)
    lds[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kBT / kWave; ++w)
      t += lds[w];
    out[gb] = t;
  }
  __syncthreads();
}

__device__ __forceinline__ bool running(const SegState& st) { return st.active && !st.done && !st.breakdown; }

__device__ __forceinline__ bool reliable(const SegState& st, double rr, int k)
{
  return rr <= st.delta2 * st.rr_ref[(k - 1) & 1] || rr <= st.tol2 || k + 1 >= st.max_iter || !isfinite(rr);
}

// Row product of G lanes: returns the row sum in lane 0 of the group.  The
// FP32 values are in difference form (k_prepare): the diagonal slot holds the
// row sum s_i and multiplies x_i, every other slot multiplies x_j - x_i.
template <int G, class VT, class XT>
__device__ __forceinline__ double row_dot(const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                          const VT* __restrict__ v, const XT* __restrict__ x, int32_t row, int lane)
{
  constexpr bool DIFF = sizeof(VT) == 4;
  using A = typename std::conditional<DIFF, float, double>::type;
  const A xi = DIFF ? (A)x[row] : (A)0;
  A sum = 0;
  for (int32_t k = rp[row] + lane; k < rp[row + 1]; k += G) {
    const int32_t c = ci[k];
    sum += (A)v[k] * (DIFF ? (A)x[c] - (c == row ? (A)0 : xi) : (A)x[c]);
  }
  return (double)group_reduce_sum<G>(sum);
}

// Row r of the segment-aligned SELL-64 copy (4 slots per step; the last 1-3
// slots loaded together).
template <class VT, class XT>
__device__ __forceinline__ double sell_row(const Seg& g, const VT* __restrict__ v, const XT* __restrict__ x,
                                           int32_t r, int32_t sr0, int32_t sl0)
{
  constexpr bool DIFF = sizeof(VT) == 4; // FP32 values in difference form (offset 0: the diagonal or padding)
  using A = typename std::conditional<DIFF, float, double>::type;
  const int32_t local = r - sr0;
  const int32_t sl = sl0 + (local >> 6);
  const int64_t base = (int64_t)g.soff[sl] * 64 + (local & 63);
  const int w = g.soff[sl + 1] - g.soff[sl];
  const XT* xr = x + r;
  const A xi = DIFF ? (A)xr[0] : (A)0;
  auto gx = [&](int d) { return DIFF ? (A)xr[d] - (d != 0 ? xi : (A)0) : (A)xr[d]; };
  A s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  int j = 0;
  for (; j + 3 < w; j += 4) {
    const int64_t i0 = base + (int64_t)j * 64;
    int d[4];
    VT m[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      d[u] = __builtin_nontemporal_load(g.sdcol + i0 + 64 * u);
      m[u] = __builtin_nontemporal_load(v + i0 + 64 * u);
    }
    s0 += (A)m[0] * gx(d[0]);
    s1 += (A)m[1] * gx(d[1]);
    s2 += (A)m[2] * gx(d[2]);
    s3 += (A)m[3] * gx(d[3]);
  }
  if (j < w) {
    const int64_t i0 = base + (int64_t)j * 64;
    const int rem = w - j;
    int d[3] = { 0, 0, 0 };
    VT m[3] = { 0, 0, 0 };
#pragma unroll
    for (int u = 0; u < 3; ++u)
      if (u < rem) {
        d[u] = __builtin_nontemporal_load(g.sdcol + i0 + 64 * u);
        m[u] = __builtin_nontemporal_load(v + i0 + 64 * u);
      }
    s0 += (A)m[0] * gx(d[0]);
    s1 += (A)m[1] * gx(d[1]);
    s2 += (A)m[2] * gx(d[2]);
  }
  return (double)((s0 + s1) + (s2 + s3));
}

// q = A32 p on the segment-aligned SELL-64 copy; partial p.q
__global__ void __launch_bounds__(kBT)
k_spmv_pq_sell(Seg g, const float* __restrict__ v, const float* __restrict__ p, float* __restrict__ q,
               const SegState* __restrict__ st, BParts P)
{
  const Ctx c = context(g);
  if (!running(st[c.s]))
    return;
  const int32_t sr0 = g.rows[c.s], sl0 = g.sslice[c.s];
  double part = 0.0;
  for (int32_t r = c.r0 + threadIdx.x; r < c.r1; r += kBT) {
    const double y = sell_row(g, v, p, r, sr0, sl0);
    q[r] = (float)y;
    part += y * (double)p[r];
  }
  put(part, P.pq, c.gb);
}

// q = A32 p; partial p.q
template <int G>
__global__ void __launch_bounds__(kBT)
k_spmv_pq(Seg g, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const float* __restrict__ v,
          const float* __restrict__ p, float* __restrict__ q, const SegState* __restrict__ st, BParts P)
{
  const Ctx c = context(g);
  if (!running(st[c.s]))
    return;
  const int lane = threadIdx.x & (G - 1);
  double part = 0.0;
  for (int32_t r = c.r0 + threadIdx.x / G; r < c.r1; r += kBT / G) {
    const double y = row_dot<G>(rp, ci, v, p, r, lane);
    if (lane == 0) {
      q[r] = (float)y;
      part += y * (double)p[r];
    }
  }
  put(part, P.pq, c.gb);
}

__device__ __forceinline__ double rho_prev(const SegState& st, const BParts& P, const Ctx& c, int par)
{
  const int q = par ^ 1;
  return st.ru[q] ? seg_sum(P.rzru[q], c.b0, c.b1) : seg_sum(P.rz[q], c.b0, c.b1);
}

// x += alpha p; r -= alpha q; z = D^-1 r; partials r.z (parity par) and r.r
__global__ void __launch_bounds__(kBT)
k_update(Seg g, double* __restrict__ x, float* __restrict__ r, float* __restrict__ z, const float* __restrict__ p,
         const float* __restrict__ q, const float* __restrict__ dinv, SegState* __restrict__ st, BParts P, int par)
{
  const Ctx c = context(g);
  if (!running(st[c.s]))
    return;
  const double pq = seg_sum(P.pq, c.b0, c.b1);
  const double rho = rho_prev(st[c.s], P, c, par);
  if (!(pq > 0.0) || !isfinite(pq)) {
    if (c.gb == c.b0 && threadIdx.x == 0)
      st[c.s].breakdown = 1;
    return;
  }
  const double alpha = rho / pq;
  const float af = (float)alpha;
  double rz = 0.0, rr = 0.0;
  for (int32_t i = c.r0 + threadIdx.x; i < c.r1; i += kBT) {
    x[i] += alpha * (double)p[i];
    const float rv = r[i] - af * q[i];
    r[i] = rv;
    const float zv = dinv[i] * rv;
    z[i] = zv;
    rz += (double)rv * (double)zv;
    rr += (double)rv * (double)rv;
  }
  put(rz, P.rz[par], c.gb);
  put(rr, P.rr, c.gb);
}

// r64 = b - A64 x on the segments that take a reliable update this
// iteration (or all active segments when always != 0); partial r64.r64.
template <int G>
__global__ void __launch_bounds__(kBT)
k_residual(Seg g, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const double* __restrict__ v,
           const double* __restrict__ x, const double* __restrict__ b, double* __restrict__ r64,
           const SegState* __restrict__ st, BParts P, int par, int always)
{
  const Ctx c = context(g);
  const SegState& S = st[c.s];
  if (always) {
    if (!S.active)
      return;
  }
  else {
    if (!running(S))
      return;
    if (!reliable(S, seg_sum(P.rr, c.b0, c.b1), S.iter_par[par]))
      return;
  }
  const int lane = threadIdx.x & (G - 1);
  double part = 0.0;
  for (int32_t r = c.r0 + threadIdx.x / G; r < c.r1; r += kBT / G) {
    const double y = row_dot<G>(rp, ci, v, x, r, lane);
    if (lane == 0) {
      const double rv = b[r] - y;
      r64[r] = rv;
      part += rv * rv;
    }
  }
  put(part, P.rr64, c.gb);
}

// k_residual on the segment-aligned SELL-64 copy (FP64 values)
__global__ void __launch_bounds__(kBT)
k_residual_sell(Seg g, const double* __restrict__ v, const double* __restrict__ x, const double* __restrict__ b,
                double* __restrict__ r64, const SegState* __restrict__ st, BParts P, int par, int always)
{
  const Ctx c = context(g);
  const SegState& S = st[c.s];
  if (always) {
    if (!S.active)
      return;
  }
  else {
    if (!running(S))
      return;
    if (!reliable(S, seg_sum(P.rr, c.b0, c.b1), S.iter_par[par]))
      return;
  }
  const int32_t sr0 = g.rows[c.s], sl0 = g.sslice[c.s];
  double part = 0.0;
  for (int32_t r = c.r0 + threadIdx.x; r < c.r1; r += kBT) {
    const double rv = b[r] - sell_row(g, v, x, r, sr0, sl0);
    r64[r] = rv;
    part += rv * rv;
  }
  put(part, P.rr64, c.gb);
}

// Slot count (longest row) of each slice of the segment-aligned SELL copy.
__global__ void k_sell_width(int32_t nslice, const int32_t* __restrict__ slice_row, const int32_t* __restrict__ rp,
                             int32_t* __restrict__ width)
{
  const int32_t sl = blockIdx.x * blockDim.x + threadIdx.x;
  if (sl >= nslice)
    return;
  int32_t w = 0;
  for (int32_t r = slice_row[sl]; r < slice_row[sl + 1]; ++r)
    w = max(w, rp[r + 1] - rp[r]);
  width[sl] = w;
}

// Column offsets of the SELL copy (one thread per slice lane); *overflow = 1
// when an offset does not fit in 16 bits.  Padding slots: offset 0.
__global__ void k_sell_columns(int32_t nslice, const int32_t* __restrict__ slice_row,
                               const int32_t* __restrict__ soff, const int32_t* __restrict__ rp,
                               const int32_t* __restrict__ ci, int16_t* __restrict__ dcol, int32_t* __restrict__ overflow)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nslice * 64)
    return;
  const int32_t sl = (int32_t)(t >> 6);
  const int lane = (int)(t & 63);
  const int32_t r = slice_row[sl] + lane;
  const bool real_row = r < slice_row[sl + 1];
  const int32_t len = real_row ? rp[r + 1] - rp[r] : 0;
  bool bad = false;
  for (int j = 0; j < soff[sl + 1] - soff[sl]; ++j) {
    int64_t d = 0;
    if (j < len) {
      d = (int64_t)ci[rp[r] + j] - r;
      if (d < -32768 || d > 32767) {
        bad = true;
        d = 0;
      }
    }
    dcol[((int64_t)soff[sl] + j) * 64 + lane] = (int16_t)d;
  }
  if (bad)
    atomicOr(overflow, 1);
}

// FP64 and FP32 values of the SELL copy (padding slots: 0); the FP32 ones
// from the difference-form CSR copy.
__global__ void k_sell_values(int32_t nslice, const int32_t* __restrict__ slice_row, const int32_t* __restrict__ soff,
                              const int32_t* __restrict__ rp, const double* __restrict__ v,
                              const float* __restrict__ v32csr, double* __restrict__ v64, float* __restrict__ v32)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nslice * 64)
    return;
  const int32_t sl = (int32_t)(t >> 6);
  const int lane = (int)(t & 63);
  const int32_t r = slice_row[sl] + lane;
  const int32_t len = r < slice_row[sl + 1] ? rp[r + 1] - rp[r] : 0;
  for (int j = 0; j < soff[sl + 1] - soff[sl]; ++j) {
    const int64_t e = ((int64_t)soff[sl] + j) * 64 + lane;
    v64[e] = j < len ? v[rp[r] + j] : 0.0;
    v32[e] = j < len ? v32csr[rp[r] + j] : 0.f;
  }
}

// Convergence, reliable-update bookkeeping and p = z + beta p.
__global__ void __launch_bounds__(kBT)
k_pupdate(Seg g, float* __restrict__ p, float* __restrict__ z, float* __restrict__ r, const double* __restrict__ r64,
          const float* __restrict__ dinv, SegState* __restrict__ st, BParts P, int par)
{
  const Ctx c = context(g);
  SegState& S = st[c.s];
  if (!S.active || S.done)
    return;
  const bool lead = c.gb == c.b0 && threadIdx.x == 0;
  const int k = S.iter_par[par];
  if (S.breakdown) {
    if (lead) {
      S.done = 1;
      S.iter = k;
    }
    return;
  }
  const double rz = seg_sum(P.rz[par], c.b0, c.b1);
  const double rr = seg_sum(P.rr, c.b0, c.b1);
  const double rho = rho_prev(S, P, c, par);
  const double beta = (rho != 0.0) ? rz / rho : 0.0;
  const bool ru = reliable(S, rr, k);
  if (ru) {
    const double rr64 = seg_sum(P.rr64, c.b0, c.b1);
    if (rr64 <= S.tol2 || k + 1 >= S.max_iter || !isfinite(rr64)) {
      if (lead) {
        S.done = 1;
        S.converged = rr64 <= S.tol2;
        S.iter = k + 1;
        S.rr_out = rr64;
        S.nb_ru += 1;
      }
      return;
    }
    if (lead) {
      S.ru[par] = 1;
      S.rr_ref[par] = rr64;
      S.rr_out = rr64;
      S.nb_ru += 1;
    }
  }
  else if (lead) {
    S.ru[par] = 0;
    S.rr_ref[par] = S.rr_ref[par ^ 1];
    S.rr_out = rr;
  }
  if (lead) {
    S.iter = k + 1;
    S.iter_par[par ^ 1] = k + 1;
  }
  const float bt = (float)beta;
  double rzn = 0.0;
  for (int32_t i = c.r0 + threadIdx.x; i < c.r1; i += kBT) {
    float zv;
    if (ru) {
      const float rv = (float)r64[i];
      r[i] = rv;
      zv = dinv[i] * rv;
      z[i] = zv;
      rzn += (double)rv * (double)zv;
    }
    else {
      zv = z[i];
    }
    p[i] = zv + bt * p[i];
  }
  if (ru)
    put(rzn, P.rzru[par], c.gb);
}

// r = r64; z = D^-1 r; p = z; partials r.z (parity 1), r.r, b.b
__global__ void __launch_bounds__(kBT)
k_init(Seg g, const double* __restrict__ r64, const double* __restrict__ b, float* __restrict__ r,
       float* __restrict__ z, float* __restrict__ p, const float* __restrict__ dinv, const SegState* __restrict__ st,
       BParts P)
{
  const Ctx c = context(g);
  if (!st[c.s].active)
    return;
  double rz = 0.0, rr = 0.0, bb = 0.0;
  for (int32_t i = c.r0 + threadIdx.x; i < c.r1; i += kBT) {
    const double rd = r64[i];
    const float rv = (float)rd;
    const float zv = dinv[i] * rv;
    r[i] = rv;
    z[i] = zv;
    p[i] = zv;
    rz += (double)rv * (double)zv;
    rr += rd * rd;
    bb += b[i] * b[i];
  }
  put(rz, P.rz[1], c.gb);
  put(rr, P.rr, c.gb);
  put(bb, P.bb, c.gb);
}

// one workgroup per segment
__global__ void __launch_bounds__(kBT)
k_init_finalize(Seg g, SegState* __restrict__ st, BParts P)
{
  const int s = blockIdx.x;
  SegState& S = st[s];
  if (!S.active)
    return;
  const int b0 = g.blocks[s], b1 = g.blocks[s + 1];
  const double rz = seg_sum(P.rz[1], b0, b1);
  const double rr = seg_sum(P.rr, b0, b1);
  const double bb = seg_sum(P.bb, b0, b1);
  if (threadIdx.x == 0) {
    S.bb = bb;
    S.tol2 = fmax(S.rtol2 * bb, S.atol2);
    S.rr_ref[1] = rr;
    S.ru[1] = 0;
    S.rr_out = rr;
    S.iter = 0;
    S.iter_par[0] = 0;
    S.iter_par[1] = 0;
    S.done = (rr <= S.tol2 || rz == 0.0) ? 1 : 0;
    S.converged = S.done;
  }
}

__global__ void __launch_bounds__(kBT)
k_true_finalize(Seg g, SegState* __restrict__ st, BParts P)
{
  const int s = blockIdx.x;
  SegState& S = st[s];
  if (!S.active)
    return;
  const double rr = seg_sum(P.rr64, g.blocks[s], g.blocks[s + 1]);
  if (threadIdx.x == 0)
    S.rr_true = rr;
}

__global__ void k_initial_guess(int32_t n, const uint8_t* __restrict__ elim, const double* __restrict__ elim_value,
                                double* __restrict__ x)
{
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n)
    x[r] = (elim[r] == 1 || elim[r] == 2) ? elim_value[r] : 0.0;
}

// FP32 values in difference form (off-diagonal entries rounded, the
// diagonal slot holding the row sum taken in FP64: the FP32 product is then
// s_i x_i + sum_{j != i} a_ij (x_j - x_i), free of the cancellation between the
// diagonal and the off-diagonal terms on smooth vectors) and the Jacobi
// inverse.  *missing = 1 when a non-empty row has no diagonal entry.
__global__ void k_prepare(int32_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                          const double* __restrict__ v, float* __restrict__ v32, float* __restrict__ dinv,
                          int32_t* __restrict__ missing)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t r = t; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    double d = 0.0, sum = 0.0;
    int32_t dpos = -1;
    for (int32_t k = rp[r]; k < rp[r + 1]; ++k) {
      const double a = v[k];
      sum += a;
      if (ci[k] == r) {
        d = a;
        dpos = k;
      }
      else
        v32[k] = (float)a;
    }
    if (dpos >= 0)
      v32[dpos] = (float)sum;
    else if (rp[r] < rp[r + 1])
      *missing = 1;
    dinv[r] = (float)(d != 0.0 ? 1.0 / d : 1.0);
  }
}

// ---- persistent solve ----------------------------------------------------------------------
//
// One launch runs the whole mixed PCG of every active segment.  Segment s
// owns a team of consecutive workgroups of the launch; workgroup j of the
// team owns a strip of the segment's rows (a multiple of 64 rows, so a wave
// of the SELL product is one slice).  The workgroups of a team synchronize
// through a counter in global memory: every wave arrives with an agent-scope
// release add, the first thread polls the counter and, once every wave of the
// team has arrived, takes an agent-scope acquire fence.  The grid never
// exceeds the number of workgroups the device holds at once, and every wait
// is bounded: after spin_limit polls, or when the host raises its cancel flag
// (pinned host memory), the waiting workgroup sets the abort flag, and every
// workgroup that sees it leaves the kernel; the host then reports an error.
//
// The p update is folded into the product, so an iteration has two barriers:
//   A  beta = rz / rho; p_new = z + beta p_old for the own rows (p is
//      double-buffered by parity; every gathered column forms the same
//      z + beta p_old with the same fma); q = A32 p_new; partial p.q
//   -- barrier --
//   B  alpha = rho / p.q; x += alpha p_new; r -= alpha q; z = D r; partials r.z, r.r
//   -- barrier --
// An iteration that takes a reliable update adds two: r64 = b - A64 x
// (partial r64.r64) -- barrier -- r = r64, z = D r (partial r.z) -- barrier.
// All the workgroups of a team reduce the same partials in the same order, so
// they take the same decisions and leave the loop together.

struct Team
{
  const int32_t* seg_of; // segment of each launched workgroup (-1: idle)
  const int32_t* slot_of; // team-contiguous index of each launched workgroup (partial sums)
  const int32_t* first;  // first team index of each segment's team (nb_seg + 1)
  const int32_t* rows;   // segment row offsets (nb_seg + 1)
  const int32_t* strip;  // rows per workgroup of each segment's team (multiple of 64)
  // SELL copy (sell != 0) or CSR arrays
  int sell;
  const int32_t* sslice;
  const int32_t* soff;
  const int16_t* sdcol;
  const int32_t* rp;
  const int32_t* ci;
  const float* v32;
  const double* v64;
  unsigned* cnt;           // per-segment barrier counters, zero at the launch
  int* abort_flag;         // zero at the launch
  const int* host_cancel;  // pinned host flag (may be null)
  long long spin_limit;
  int xcc[8];              // XCD-local layout: XCD id of launched workgroups g with g mod 8 = c
};

struct TeamVec
{
  double* x;
  const double* b;
  const float* dinv;
  float* r;
  float* z;
  float* p[2];
  float* q;
  double* r64;
};

// one entry per launched workgroup; written by one workgroup, read by its
// team after a barrier (plain pointers: never cached as read-only)
struct TeamParts
{
  double* pq;
  double* rz;
  double* rr;
  double* rr64;
  double* rzru;
  double* bb;
};

__device__ __forceinline__ bool team_abort(const Team& T)
{
  return __hip_atomic_load(T.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// The vectors workgroups of a team exchange (z, both p buffers, x, the
// partial sums) are written with agent-scope atomic stores and read with
// agent-scope atomic loads, which go to the device's coherence point instead
// of the per-CU L1 and per-XCD L2; everything else (the matrix, D^-1, b, the
// rows a workgroup alone reads and writes) uses ordinary cached accesses.  So
// the barrier needs neither an L2 write-back (an agent-scope release) nor an
// L2 invalidation (an agent-scope acquire, which would also drop the cached
// matrix of every other team on the XCD): every wave waits until its stores
// are acknowledged, the workgroup synchronizes, and its first thread arrives
// with one relaxed add and polls the counter; the loads that follow are
// issued only after the poll has seen every arrival.
__device__ __forceinline__ void publish(float* p, float v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void publish(double* p, double v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float fetch(const float* p)
{
  return __hip_atomic_load(const_cast<float*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double fetch(const double* p)
{
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exchanged vectors: XCD-local teams (LOCAL) use ordinary cached accesses --
// every workgroup of the team shares the XCD's L2, and the L1 is written
// through and invalidated at each barrier -- others the coherence point.
template <bool LOCAL, class V>
__device__ __forceinline__ void xstore(V* p, V v)
{
  if constexpr (LOCAL)
    *p = v;
  else
    publish(p, v);
}
template <bool LOCAL, class V>
__device__ __forceinline__ V xload(const V* p)
{
  if constexpr (LOCAL)
    return *p;
  else
    return fetch(p);
}

// Barrier of the team of segment s (nw workgroups); `phase` counts this
// workgroup's barriers.  False: give up (abort).
template <bool LOCAL>
__device__ bool team_barrier(const Team& T, int s, int nw, unsigned& phase)
{
  __shared__ int s_ok;
  ++phase;
  const unsigned target = phase * (unsigned)nw;
  __atomic_signal_fence(__ATOMIC_SEQ_CST); // keep the phase's stores above the wait
  __builtin_amdgcn_s_waitcnt(0);           // this wave's stores are acknowledged
  __syncthreads();                         // ... and so are every wave's
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(T.cnt + s, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int ok = 1;
    long long n = 0;
    while (__hip_atomic_load(T.cnt + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      ++n;
      bool give_up = n >= T.spin_limit;
      if (!give_up && (n & 31) == 0) // the flags are polled less often than the counter
        give_up = team_abort(T) || (T.host_cancel && __hip_atomic_load(const_cast<int*>(T.host_cancel),
                                                                       __ATOMIC_RELAXED,
                                                                       __HIP_MEMORY_SCOPE_SYSTEM) != 0);
      if (give_up) {
        __hip_atomic_fetch_max(T.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    if constexpr (LOCAL) // drop the CU's L1 lines: the team's data is current in the XCD's L2
      asm volatile("buffer_inv sc1" ::: "memory");
    s_ok = ok;
  }
  __syncthreads();
  const bool ok = s_ok != 0;
  __syncthreads(); // s_ok is rewritten by the next barrier
  return ok;
}

// sum of a team's partials p[t0..t1), identical in every workgroup of the team
__device__ __forceinline__ double team_sum(const double* p, int t0, int t1)
{
  __shared__ double lds[kBT / kWave];
  double s = 0.0;
  for (int i = t0 + threadIdx.x; i < t1; i += kBT)
    s += __hip_atomic_load(p + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  s = wave_reduce_sum(s);
  if ((threadIdx.x & (kWave - 1)) == 0)
    lds[threadIdx.x / kWave] = s;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < kBT / kWave; ++w)
    t += lds[w];
  __syncthreads();
  return t;
}

// p_new of column c: the same fma at the owner and at every gather
template <bool LOCAL>
__device__ __forceinline__ float p_of(const float* z, const float* pold, float beta, int64_t c)
{
  return __fmaf_rn(beta, xload<LOCAL>(pold + c), xload<LOCAL>(z + c));
}

// Row r of A32 p_new (p_new formed from z and p_old on the fly), the FP32
// values in difference form: the diagonal slot (offset 0, also the padding,
// whose value is 0) multiplies p_i, every other slot p_j - p_i.
template <bool LOCAL>
__device__ __forceinline__ float prod_row32(const Team& T, const float* z, const float* pold, float beta, int32_t r,
                                            int32_t sr0, int32_t sl0)
{
  float s0 = 0.f, s1 = 0.f;
  const float pi = p_of<LOCAL>(z, pold, beta, (int64_t)r);
  if (T.sell) {
    const int32_t local = r - sr0;
    const int32_t sl = sl0 + (local >> 6);
    const int64_t base = (int64_t)T.soff[sl] * 64 + (local & 63);
    const int w = T.soff[sl + 1] - T.soff[sl];
    if (w <= 8) { // every slot's loads in one round, then every gather in one round
      int d[8];
      float m[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        d[u] = u < w ? T.sdcol[base + 64 * u] : 0;
        m[u] = u < w ? T.v32[base + 64 * u] : 0.f;
      }
      float g[8];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        g[u] = u < w ? p_of<LOCAL>(z, pold, beta, (int64_t)r + d[u]) - (d[u] != 0 ? pi : 0.f) : 0.f;
#pragma unroll
      for (int u = 0; u < 8; u += 2) {
        s0 += m[u] * g[u];
        s1 += m[u + 1] * g[u + 1];
      }
      return s0 + s1;
    }
    float s2 = 0.f, s3 = 0.f;
    int j = 0;
    for (; j + 3 < w; j += 4) { // four slots' loads in flight
      const int64_t i0 = base + (int64_t)j * 64;
      int d[4];
      float m[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        d[u] = T.sdcol[i0 + 64 * u];
        m[u] = T.v32[i0 + 64 * u];
      }
      s0 += m[0] * (p_of<LOCAL>(z, pold, beta, (int64_t)r + d[0]) - (d[0] != 0 ? pi : 0.f));
      s1 += m[1] * (p_of<LOCAL>(z, pold, beta, (int64_t)r + d[1]) - (d[1] != 0 ? pi : 0.f));
      s2 += m[2] * (p_of<LOCAL>(z, pold, beta, (int64_t)r + d[2]) - (d[2] != 0 ? pi : 0.f));
      s3 += m[3] * (p_of<LOCAL>(z, pold, beta, (int64_t)r + d[3]) - (d[3] != 0 ? pi : 0.f));
    }
    for (; j < w; ++j) {
      const int64_t i0 = base + (int64_t)j * 64;
      const int d0 = T.sdcol[i0];
      s0 += T.v32[i0] * (p_of<LOCAL>(z, pold, beta, (int64_t)r + d0) - (d0 != 0 ? pi : 0.f));
    }
    s0 += s2;
    s1 += s3;
  }
  else {
    const int32_t e = T.rp[r + 1];
    int32_t k = T.rp[r];
    auto g = [&](int32_t c) { return p_of<LOCAL>(z, pold, beta, c) - (c != r ? pi : 0.f); };
    for (; k + 1 < e; k += 2) {
      s0 += T.v32[k] * g(T.ci[k]);
      s1 += T.v32[k + 1] * g(T.ci[k + 1]);
    }
    if (k < e)
      s0 += T.v32[k] * g(T.ci[k]);
  }
  return s0 + s1;
}

// Row r of b - A64 x.
template <bool LOCAL>
__device__ __forceinline__ double residual_row64(const Team& T, const double* x, const double* b, int32_t r,
                                                 int32_t sr0, int32_t sl0)
{
  double s0 = 0.0, s1 = 0.0;
  if (T.sell) {
    const int32_t local = r - sr0;
    const int32_t sl = sl0 + (local >> 6);
    const int64_t base = (int64_t)T.soff[sl] * 64 + (local & 63);
    const int w = T.soff[sl + 1] - T.soff[sl];
    int j = 0;
    for (; j + 1 < w; j += 2) {
      const int64_t i0 = base + (int64_t)j * 64, i1 = i0 + 64;
      s0 += T.v64[i0] * xload<LOCAL>(x + (int64_t)r + T.sdcol[i0]);
      s1 += T.v64[i1] * xload<LOCAL>(x + (int64_t)r + T.sdcol[i1]);
    }
    if (j < w) {
      const int64_t i0 = base + (int64_t)j * 64;
      s0 += T.v64[i0] * xload<LOCAL>(x + (int64_t)r + T.sdcol[i0]);
    }
  }
  else {
    const int32_t e = T.rp[r + 1];
    int32_t k = T.rp[r];
    for (; k + 1 < e; k += 2) {
      s0 += T.v64[k] * xload<LOCAL>(x + T.ci[k]);
      s1 += T.v64[k + 1] * xload<LOCAL>(x + T.ci[k + 1]);
    }
    if (k < e)
      s0 += T.v64[k] * xload<LOCAL>(x + T.ci[k]);
  }
  return b[r] - (s0 + s1);
}

// out[gb] = the sum of v over the workgroup
__device__ __forceinline__ void team_put(double v, double* out, int gb)
{
  __shared__ double lds[kBT / kWave];
  v = wave_reduce_sum(v);
  if ((threadIdx.x & (kWave - 1)) == 0)
    lds[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kBT / kWave; ++w)
      t += lds[w];
    publish(out + gb, t);
  }
  __syncthreads();
}

// LOCAL: every team's workgroups run on one XCD (the device dispatches
// launched workgroup g to XCD g mod 8; a probe launch recorded which XCD id
// each class gets); each workgroup checks its XCD id and, on a mismatch,
// aborts the launch with abort flag 2, after which the host restores x and
// runs the solve again with the coherent exchange.  A team with a missing
// member never passes its first barrier, so no result of it is kept.
template <bool LOCAL>
__global__ void __launch_bounds__(kBT, 4) // 4 waves per SIMD without spills: 4 workgroups per CU
k_persistent(Team T, TeamVec V, TeamParts P, SegState* st)
{
  const int gb = blockIdx.x;
  const int s = T.seg_of[gb];
  if (s < 0)
    return; // an idle slot of the XCD-local layout
  if constexpr (LOCAL) {
    uint32_t xcc;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
    if ((int)xcc != T.xcc[gb & 7]) {
      if (threadIdx.x == 0)
        __hip_atomic_fetch_max(T.abort_flag, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
  }
  const int slot = T.slot_of[gb];
  const int t0 = T.first[s], t1 = T.first[s + 1], nw = t1 - t0;
  const int32_t sr0 = T.rows[s], sr1 = T.rows[s + 1];
  const int32_t r0 = sr0 + (slot - t0) * T.strip[s];
  const int32_t r1 = min(sr1, r0 + T.strip[s]);
  const int32_t sl0 = T.sell ? T.sslice[s] : 0;
  const bool lead = slot == t0 && threadIdx.x == 0;
  SegState& S = st[s];
  if (!S.active || team_abort(T))
    return;
  const double tol_r2 = S.rtol2, tol_a2 = S.atol2, delta2 = S.delta2;
  const int max_iter = S.max_iter;
  unsigned phase = 0;
  // init: r64 = b - A64 x; r = r64, z = D r, p_old = 0; partials r.z, r.r, b.b
  double rz = 0.0, rr = 0.0, bb = 0.0;
  for (int32_t i = r0 + threadIdx.x; i < r1; i += kBT) {
    const double rd = residual_row64<LOCAL>(T, V.x, V.b, i, sr0, sl0);
    const float rv = (float)rd;
    const float zv = V.dinv[i] * rv;
    V.r[i] = rv;
    xstore<LOCAL>(V.z + i, zv);
    xstore<LOCAL>(V.p[1] + i, 0.f);
    rz += (double)rv * (double)zv;
    rr += rd * rd;
    bb += V.b[i] * V.b[i];
  }
  team_put(rz, P.rz, slot);
  team_put(rr, P.rr, slot);
  team_put(bb, P.bb, slot);
  if (!team_barrier<LOCAL>(T, s, nw, phase))
    return;
  rz = team_sum(P.rz, t0, t1);
  rr = team_sum(P.rr, t0, t1);
  bb = team_sum(P.bb, t0, t1);
  const double tol2 = fmax(tol_r2 * bb, tol_a2);
  double rho = rz, rr_ref = rr, rr_out = rr;
  double beta = 0.0;
  int k = 0, nb_ru = 0, converged = 0, breakdown = 0;
  bool done = rr <= tol2 || rz == 0.0;
  converged = done ? 1 : 0;
  while (!done) {
    // A: p_new = z + beta p_old; q = A32 p_new; partial p.q
    const int par = k & 1;
    float* pnew = V.p[par];
    const float* pold = V.p[par ^ 1];
    const float bt = (float)beta;
    double pq = 0.0;
    // two rows per step, so that their loads are in flight together
    for (int32_t i = r0 + threadIdx.x; i < r1; i += 2 * kBT) {
      const int32_t i2 = i + kBT;
      const bool two = i2 < r1;
      const float y = prod_row32<LOCAL>(T, V.z, pold, bt, i, sr0, sl0);
      const float y2 = two ? prod_row32<LOCAL>(T, V.z, pold, bt, i2, sr0, sl0) : 0.f;
      const float pi = p_of<LOCAL>(V.z, pold, bt, i);
      const float pi2 = two ? p_of<LOCAL>(V.z, pold, bt, i2) : 0.f;
      V.q[i] = y;
      xstore<LOCAL>(pnew + i, pi); // other gathers of this phase read pold, a different buffer
      pq += (double)y * (double)pi;
      if (two) {
        V.q[i2] = y2;
        xstore<LOCAL>(pnew + i2, pi2);
        pq += (double)y2 * (double)pi2;
      }
    }
    team_put(pq, P.pq, slot);
    if (!team_barrier<LOCAL>(T, s, nw, phase))
      return;
    // B: alpha; x, r, z; partials r.z, r.r
    pq = team_sum(P.pq, t0, t1);
    if (!(pq > 0.0) || !isfinite(pq)) {
      breakdown = 1;
      break;
    }
    const double alpha = rho / pq;
    const float af = (float)alpha;
    rz = 0.0;
    rr = 0.0;
#pragma unroll 4
    for (int32_t i = r0 + threadIdx.x; i < r1; i += kBT) {
      xstore<LOCAL>(V.x + i, xload<LOCAL>(V.x + i) + alpha * (double)xload<LOCAL>(pnew + i));
      const float rv = V.r[i] - af * V.q[i];
      V.r[i] = rv;
      const float zv = V.dinv[i] * rv;
      xstore<LOCAL>(V.z + i, zv);
      rz += (double)rv * (double)zv;
      rr += (double)rv * (double)rv;
    }
    team_put(rz, P.rz, slot);
    team_put(rr, P.rr, slot);
    if (!team_barrier<LOCAL>(T, s, nw, phase))
      return;
    rz = team_sum(P.rz, t0, t1);
    rr = team_sum(P.rr, t0, t1);
    beta = (rho != 0.0) ? rz / rho : 0.0;
    const bool ru = rr <= delta2 * rr_ref || rr <= tol2 || k + 1 >= max_iter || !isfinite(rr);
    if (ru) {
      double part = 0.0;
      for (int32_t i = r0 + threadIdx.x; i < r1; i += kBT) {
        const double rd = residual_row64<LOCAL>(T, V.x, V.b, i, sr0, sl0);
        V.r64[i] = rd;
        part += rd * rd;
      }
      team_put(part, P.rr64, slot);
      if (!team_barrier<LOCAL>(T, s, nw, phase))
        return;
      const double rr64 = team_sum(P.rr64, t0, t1);
      ++nb_ru;
      rr_out = rr64;
      if (rr64 <= tol2 || k + 1 >= max_iter || !isfinite(rr64)) {
        converged = rr64 <= tol2 ? 1 : 0;
        ++k;
        break;
      }
      double rzn = 0.0;
      for (int32_t i = r0 + threadIdx.x; i < r1; i += kBT) {
        const float rv = (float)V.r64[i];
        V.r[i] = rv;
        const float zv = V.dinv[i] * rv;
        xstore<LOCAL>(V.z + i, zv);
        rzn += (double)rv * (double)zv;
      }
      team_put(rzn, P.rzru, slot);
      if (!team_barrier<LOCAL>(T, s, nw, phase))
        return;
      rho = team_sum(P.rzru, t0, t1);
      rr_ref = rr64;
    }
    else {
      rho = rz;
      rr_out = rr;
    }
    ++k;
  }
  if (lead) {
    S.done = 1;
    S.converged = converged;
    S.breakdown = breakdown;
    S.iter = k;
    S.nb_ru = nb_ru;
    S.bb = bb;
    S.tol2 = tol2;
    S.rr_out = rr_out;
  }
}

// XCD id of each launched workgroup.
__global__ void k_xcc_probe(int* out)
{
  if (threadIdx.x == 0) {
    uint32_t xcc;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
    out[blockIdx.x] = (int)xcc;
  }
}

__global__ void k_persistent_clear(unsigned* cnt, int nb_seg, int* abort_flag)
{
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nb_seg; i += gridDim.x * blockDim.x)
    cnt[i] = 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0)
    *abort_flag = 0;
}

} // namespace batch
} // namespace afem

using namespace afem;
using namespace afem::batch;

struct afem_batch
{
  afem_system* sys = nullptr;
  int nb_seg = 0;
  std::vector<int32_t> seg_rows;
  DBuf<int32_t> d_rows, d_blocks, d_owner;
  int lanes = 8, tiles = 1, nblk = 0;
  std::vector<int32_t> blocks_host;
  // launched workgroups: nlaunch > 0 -> the first nlaunch entries of d_list
  DBuf<int32_t> d_list;
  int nlaunch = 0;
  // segment-aligned SELL-64 copy (built with the first solve of a pattern)
  const int32_t* sell_rp = nullptr;
  bool sell_ok = false;
  DBuf<int32_t> d_sslice, d_soff, d_slice_row;
  DBuf<int16_t> sdcol;
  DBuf<double> sv64;
  DBuf<float> sv32;
  int nslice = 0;
  DBuf<float> v32, dinv, r32, z32, p32, q32;
  DBuf<int32_t> missing; // set by k_prepare when a non-empty row lacks its diagonal entry
  DBuf<double> r64, parts;
  DBuf<SegState> st;
  SegState* host_st = nullptr;
  bool prepared = false;
  hipGraphExec_t graph = nullptr;
  int64_t graph_key = -1;
  // persistent solve: team layout, second p buffer, partials, barrier
  // counters, abort flag, host cancel flag (pinned), its captured graph
  int pt_budget = 0; // requested workgroups (0 = default)
  int pt_grid = 0;   // workgroups of the layout
  int pt_layout_budget = -1;
  int pt_layout_local = -1; // layout kind built: 1 XCD-local teams, 0 contiguous
  int pt_local_ok = -1;     // XCD-local teams usable: -1 not probed yet
  int pt_xcc[8] = {};
  int pt_last_mode = 0;     // 1 XCD-local, 2 coherent, 0 none yet
  std::vector<int32_t> pt_first_host, pt_strip_host;
  DBuf<int32_t> pt_seg_of, pt_slot_of, pt_first, pt_strip;
  DBuf<double> pt_xsave; // x at the launch of an XCD-local solve
  DBuf<float> p32b;
  DBuf<double> pt_parts;
  DBuf<unsigned> pt_cnt;
  DBuf<int> pt_abort;
  int* pt_cancel = nullptr; // pinned, mapped
  hipGraphExec_t pt_graph = nullptr;
  uint64_t pt_graph_key = 0;
  ~afem_batch()
  {
    if (graph)
      (void)hipGraphExecDestroy(graph);
    if (pt_graph)
      (void)hipGraphExecDestroy(pt_graph);
    if (host_st)
      (void)hipHostFree(host_st);
    if (pt_cancel)
      (void)hipHostFree(pt_cancel);
  }
};

namespace {

// rows per workgroup: `tiles` strips of 256 rows (the SpMV kernels give each
// row `lanes` lanes and walk the strip 256 / lanes rows at a time)
int rows_per_block(const afem_batch* b) { return b->tiles * kBT; }

void drop_graph(afem_batch* b)
{
  if (b->graph) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  b->graph_key = -1;
}

// lanes = 1 runs the SELL kernels when the copy exists, else CSR with 8 lanes
bool use_sell(const afem_batch* b) { return b->lanes == 1 && b->sell_ok; }

int grid(const afem_batch* b) { return b->nlaunch > 0 ? b->nlaunch : b->nblk; }

void layout(afem_batch* b)
{
  const int rpb = rows_per_block(b);
  b->blocks_host.assign(b->nb_seg + 1, 0);
  for (int s = 0; s < b->nb_seg; ++s) {
    const int32_t rows = b->seg_rows[s + 1] - b->seg_rows[s];
    b->blocks_host[s + 1] = b->blocks_host[s] + std::max(1, (rows + rpb - 1) / rpb);
  }
  b->nblk = b->blocks_host[b->nb_seg];
  std::vector<int32_t> owner(b->nblk);
  for (int s = 0; s < b->nb_seg; ++s)
    for (int k = b->blocks_host[s]; k < b->blocks_host[s + 1]; ++k)
      owner[k] = s;
  b->d_blocks.alloc(b->nb_seg + 1);
  b->d_blocks.upload(b->blocks_host.data(), b->blocks_host.size());
  b->d_owner.alloc(b->nblk);
  b->d_owner.upload(owner.data(), owner.size());
  b->d_list.alloc(b->nblk);
  b->nlaunch = 0;
  b->parts.alloc((size_t)8 * b->nblk);
  drop_graph(b);
}

Seg seg_of(const afem_batch* b)
{
  return Seg{ b->d_rows.p,  b->d_blocks.p, b->d_owner.p, rows_per_block(b), b->nlaunch > 0 ? b->d_list.p : nullptr,
              b->d_sslice.p, b->d_soff.p,   b->sdcol.p };
}

// Launches only the workgroups of the segments still running (host states
// just read) once they are at most half of the current launch.
void shrink_grid(afem_batch* b)
{
  std::vector<int32_t> list;
  for (int k = 0; k < b->nb_seg; ++k)
    if (b->host_st[k].active && !b->host_st[k].done && !b->host_st[k].breakdown)
      for (int w = b->blocks_host[k]; w < b->blocks_host[k + 1]; ++w)
        list.push_back(w);
  if (list.empty() || 2 * list.size() > (size_t)grid(b))
    return;
  b->d_list.upload(list.data(), list.size());
  b->nlaunch = (int)list.size();
}

// The segment-aligned SELL-64 layout of the pattern: slices of 64 rows
// starting at every segment's first row, 16-bit column offsets.
void build_sell(afem_batch* b)
{
  afem_system* s = b->sys;
  b->sell_rp = s->row_index;
  b->sell_ok = false;
  std::vector<int32_t> slice_row, sslice(b->nb_seg + 1, 0);
  for (int k = 0; k < b->nb_seg; ++k) {
    sslice[k] = (int32_t)slice_row.size();
    for (int32_t r = b->seg_rows[k]; r < b->seg_rows[k + 1]; r += 64)
      slice_row.push_back(r);
  }
  sslice[b->nb_seg] = (int32_t)slice_row.size();
  slice_row.push_back(s->nb_dof);
  b->nslice = sslice[b->nb_seg];
  b->d_sslice.alloc(sslice.size());
  b->d_sslice.upload(sslice.data(), sslice.size());
  b->d_slice_row.alloc(slice_row.size());
  b->d_slice_row.upload(slice_row.data(), slice_row.size());
  DBuf<int32_t> width(b->nslice);
  hipLaunchKernelGGL(k_sell_width, dim3(grid_for(b->nslice, 256)), dim3(256), 0, stream(), b->nslice,
                     b->d_slice_row.p, s->row_index, width.p);
  check_launch("batch_sell_width");
  std::vector<int32_t> w(b->nslice), off(b->nslice + 1, 0);
  width.download(w.data(), w.size());
  for (int k = 0; k < b->nslice; ++k)
    off[k + 1] = off[k] + w[k];
  b->d_soff.alloc(off.size());
  b->d_soff.upload(off.data(), off.size());
  const size_t slots = (size_t)off[b->nslice] * 64;
  b->sdcol.alloc(std::max<size_t>(slots, 1));
  b->sv64.alloc(std::max<size_t>(slots, 1));
  b->sv32.alloc(std::max<size_t>(slots, 1));
  DBuf<int32_t> overflow(1);
  overflow.zero();
  hipLaunchKernelGGL(k_sell_columns, dim3(grid_for((int64_t)b->nslice * 64, 256)), dim3(256), 0, stream(),
                     b->nslice, b->d_slice_row.p, b->d_soff.p, s->row_index, s->columns, b->sdcol.p, overflow.p);
  check_launch("batch_sell_columns");
  int32_t h = 0;
  overflow.download(&h, 1);
  b->sell_ok = h == 0;
}

BParts parts_of(afem_batch* b)
{
  double* p = b->parts.p;
  const int n = b->nblk;
  return BParts{ p, { p + n, p + 2 * n }, { p + 3 * n, p + 4 * n }, p + 5 * n, p + 6 * n, p + 7 * n };
}

template <class F>
void with_lanes(int lanes, F&& f)
{
  switch (lanes) {
  case 4: f(std::integral_constant<int, 4>{}); break;
  case 16: f(std::integral_constant<int, 16>{}); break;
  default: f(std::integral_constant<int, 8>{}); break;
  }
}

void launch_residual(afem_batch* b, int par, int always)
{
  afem_system* s = b->sys;
  const Seg g = seg_of(b);
  const BParts P = parts_of(b);
  if (use_sell(b)) {
    hipLaunchKernelGGL(k_residual_sell, dim3(grid(b)), dim3(kBT), 0, stream(), g, b->sv64.p, s->x.p, s->rhs.p,
                       b->r64.p, b->st.p, P, par, always);
    return;
  }
  with_lanes(b->lanes, [&](auto tag) {
    constexpr int G = decltype(tag)::value;
    hipLaunchKernelGGL((k_residual<G>), dim3(grid(b)), dim3(kBT), 0, stream(), g, s->row_index, s->columns,
                       s->values, s->x.p, s->rhs.p, b->r64.p, b->st.p, P, par, always);
  });
}

void launch_iteration(afem_batch* b, int par)
{
  afem_system* s = b->sys;
  const Seg g = seg_of(b);
  const BParts P = parts_of(b);
  if (use_sell(b)) {
    hipLaunchKernelGGL(k_spmv_pq_sell, dim3(grid(b)), dim3(kBT), 0, stream(), g, b->sv32.p, b->p32.p, b->q32.p,
                       b->st.p, P);
  }
  else {
    with_lanes(b->lanes, [&](auto tag) {
      constexpr int G = decltype(tag)::value;
      hipLaunchKernelGGL((k_spmv_pq<G>), dim3(grid(b)), dim3(kBT), 0, stream(), g, s->row_index, s->columns,
                         b->v32.p, b->p32.p, b->q32.p, b->st.p, P);
    });
  }
  hipLaunchKernelGGL(k_update, dim3(grid(b)), dim3(kBT), 0, stream(), g, s->x.p, b->r32.p, b->z32.p, b->p32.p,
                     b->q32.p, b->dinv.p, b->st.p, P, par);
  launch_residual(b, par, 0);
  hipLaunchKernelGGL(k_pupdate, dim3(grid(b)), dim3(kBT), 0, stream(), g, b->p32.p, b->z32.p, b->r32.p, b->r64.p,
                     b->dinv.p, b->st.p, P, par);
}

void read_states(afem_batch* b)
{
  AFEM_HIP(hipMemcpyAsync(b->host_st, b->st.p, sizeof(SegState) * b->nb_seg, hipMemcpyDeviceToHost, stream()));
  AFEM_HIP(hipStreamSynchronize(stream()));
}

void write_states(afem_batch* b)
{
  AFEM_HIP(hipMemcpyAsync(b->st.p, b->host_st, sizeof(SegState) * b->nb_seg, hipMemcpyHostToDevice, stream()));
}

void prepare(afem_batch* b)
{
  afem_system* s = b->sys;
  const int32_t n = s->nb_dof;
  if (b->v32.n != (size_t)s->nnz) {
    drop_graph(b);
    b->v32.alloc((size_t)s->nnz);
  }
  if (b->r32.n != (size_t)n) {
    drop_graph(b);
    b->dinv.alloc(n);
    b->r32.alloc(n);
    b->z32.alloc(n);
    b->p32.alloc(n);
    b->q32.alloc(n);
    b->r64.alloc(n);
  }
  if (!b->missing.p)
    b->missing.alloc(1);
  DBuf<int32_t>& missing = b->missing;
  missing.zero();
  hipLaunchKernelGGL(k_prepare, dim3(1024), dim3(256), 0, stream(), n, s->row_index, s->columns, s->values, b->v32.p,
                     b->dinv.p, missing.p);
  check_launch("batch_prepare");
  int32_t h = 0;
  missing.download(&h, 1);
  AFEM_REQUIRE(h == 0, AFEM_ERR_ARGUMENT,
               "the batched solver needs a diagonal entry in every non-empty row (its FP32 product is in "
               "difference form)");
  if (b->sell_rp != s->row_index) {
    drop_graph(b);
    build_sell(b);
  }
  if (b->sell_ok) {
    hipLaunchKernelGGL(k_sell_values, dim3(grid_for((int64_t)b->nslice * 64, 256)), dim3(256), 0, stream(),
                       b->nslice, b->d_slice_row.p, b->d_soff.p, s->row_index, s->values, b->v32.p, b->sv64.p,
                       b->sv32.p);
    check_launch("batch_sell_values");
  }
  b->prepared = true;
}

// the iterations of one chunk, captured once and replayed
void run_chunk(afem_batch* b, int chunk, bool use_graph)
{
  if (!use_graph) {
    for (int j = 0; j < chunk; ++j)
      launch_iteration(b, j & 1);
    check_launch("batch_iteration");
    return;
  }
  const int64_t key = ((int64_t)grid(b) << 32) | (chunk * 1000 + b->lanes * 10 + b->tiles);
  if (!b->graph || b->graph_key != key) {
    drop_graph(b);
    hipGraph_t gr;
    AFEM_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
    for (int j = 0; j < chunk; ++j)
      launch_iteration(b, j & 1);
    AFEM_HIP(hipStreamEndCapture(stream(), &gr));
    const hipError_t e = hipGraphInstantiate(&b->graph, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    AFEM_HIP(e);
    b->graph_key = key;
  }
  AFEM_HIP(hipGraphLaunch(b->graph, stream()));
}

struct RunStats
{
  int chunks = 0;
};

// Solves the active segments (host_st[s].active) from the current x.
RunStats run(afem_batch* b, const afem_batch_options& o, bool use_graph)
{
  RunStats rs;
  b->nlaunch = 0;
  const Seg g = seg_of(b);
  const BParts P = parts_of(b);
  afem_system* s = b->sys;
  launch_residual(b, 0, 1);
  hipLaunchKernelGGL(k_init, dim3(b->nblk), dim3(kBT), 0, stream(), g, b->r64.p, s->rhs.p, b->r32.p, b->z32.p,
                     b->p32.p, b->dinv.p, b->st.p, P);
  hipLaunchKernelGGL(k_init_finalize, dim3(b->nb_seg), dim3(kBT), 0, stream(), g, b->st.p, P);
  check_launch("batch_init");
  const int chunk = o.chunk > 0 ? (o.chunk + 1) / 2 * 2 : 32;
  int done_iters = 0;
  for (;;) {
    read_states(b);
    bool all = true;
    for (int k = 0; k < b->nb_seg; ++k)
      if (b->host_st[k].active && !b->host_st[k].done)
        all = false;
    if (all || done_iters >= o.max_iter)
      break;
    shrink_grid(b);
    run_chunk(b, chunk, use_graph);
    done_iters += chunk;
    rs.chunks += 1;
  }
  b->nlaunch = 0; // the full layout again (true residual, next run)
  return rs;
}

// ---- persistent solve: host side ----

void drop_persistent_graph(afem_batch* b)
{
  if (b->pt_graph) {
    (void)hipGraphExecDestroy(b->pt_graph);
    b->pt_graph = nullptr;
  }
  b->pt_graph_key = 0;
}

// Workgroups of k_persistent the device holds at once.
int persistent_capacity()
{
  int per_cu = 0, per_cu_local = 0;
  AFEM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_persistent<false>, kBT, 0));
  AFEM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_local, k_persistent<true>, kBT, 0));
  return std::max(0, std::min(per_cu, per_cu_local)) * rt().num_cu;
}

// Teams: segment s gets m_s workgroups in proportion to rows^1.5 (its work
// per iteration times the growth of its iteration count with its size), at
// least one and at most one per 256 rows; strips are multiples of 64 rows.
// Team slots (the partial-sum index of each member) are contiguous per team.
// Contiguous layout: launched workgroup g is slot g.  XCD-local layout: the
// teams are packed onto the 8 XCDs (largest first, onto the least loaded XCD,
// at most cap/8 workgroups each) and the j-th workgroup of XCD x is launched
// as g = 8 j + x; unused g are idle.  False: the XCD-local packing does not
// fit (the caller uses the contiguous layout).
bool persistent_layout(afem_batch* b, int budget, int cap, bool local)
{
  const int ns = b->nb_seg;
  const int per_xcd = cap / 8;
  if (local && (per_xcd < 1 || ns > 8 * per_xcd))
    return false;
  std::vector<double> w(ns);
  double wsum = 0.0;
  for (int k = 0; k < ns; ++k) {
    const double r = (double)(b->seg_rows[k + 1] - b->seg_rows[k]);
    w[k] = r * std::sqrt(r);
    wsum += w[k];
  }
  std::vector<int32_t> m(ns), strip(ns);
  auto assign = [&](double scale) {
    int64_t tot = 0;
    for (int k = 0; k < ns; ++k) {
      const int32_t rows = b->seg_rows[k + 1] - b->seg_rows[k];
      int32_t mk = (int32_t)std::min<double>(std::max(1, (rows + 255) / 256),
                                             std::max(1.0, std::floor(scale * w[k] / wsum)));
      if (local)
        mk = std::min(mk, per_xcd);
      int32_t st = (rows + mk - 1) / mk;
      st = (st + 63) / 64 * 64;
      strip[k] = st;
      m[k] = (rows + st - 1) / st;
      tot += m[k];
    }
    return tot;
  };
  // XCD packing of the current m: xcd_of[k], members placed per XCD
  std::vector<int> xcd_of(ns, 0), order(ns);
  int load[8];
  auto pack = [&]() {
    for (int k = 0; k < ns; ++k)
      order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int i, int j) { return m[i] > m[j]; });
    for (int x = 0; x < 8; ++x)
      load[x] = 0;
    for (int k : order) {
      int best = 0;
      for (int x = 1; x < 8; ++x)
        if (load[x] < load[best])
          best = x;
      if (load[best] + m[k] > per_xcd)
        return false;
      xcd_of[k] = best;
      load[best] += m[k];
    }
    return true;
  };
  double scale = budget;
  while (assign(scale) > budget && scale > 1.0)
    scale *= 0.95;
  if (local) {
    while (!pack()) {
      if (scale <= 1.0)
        return false;
      scale *= 0.9;
      assign(scale);
    }
  }
  b->pt_first_host.assign(ns + 1, 0);
  b->pt_strip_host.assign(strip.begin(), strip.end());
  for (int k = 0; k < ns; ++k)
    b->pt_first_host[k + 1] = b->pt_first_host[k] + m[k];
  const int members = b->pt_first_host[ns];
  std::vector<int32_t> seg_of, slot_of;
  if (local) {
    int depth = 0;
    for (int x = 0; x < 8; ++x)
      depth = std::max(depth, load[x]);
    b->pt_grid = 8 * depth;
    seg_of.assign(b->pt_grid, -1);
    slot_of.assign(b->pt_grid, -1);
    int used[8] = {};
    for (int k = 0; k < ns; ++k)
      for (int j = 0; j < m[k]; ++j) {
        const int x = xcd_of[k];
        const int g = 8 * used[x]++ + x;
        seg_of[g] = k;
        slot_of[g] = b->pt_first_host[k] + j;
      }
  }
  else {
    b->pt_grid = members;
    seg_of.resize(members);
    slot_of.resize(members);
    for (int k = 0; k < ns; ++k)
      for (int g = b->pt_first_host[k]; g < b->pt_first_host[k + 1]; ++g) {
        seg_of[g] = k;
        slot_of[g] = g;
      }
  }
  b->pt_seg_of.alloc(seg_of.size());
  b->pt_seg_of.upload(seg_of.data(), seg_of.size());
  b->pt_slot_of.alloc(slot_of.size());
  b->pt_slot_of.upload(slot_of.data(), slot_of.size());
  b->pt_first.alloc(ns + 1);
  b->pt_first.upload(b->pt_first_host.data(), ns + 1);
  b->pt_strip.alloc(ns);
  b->pt_strip.upload(b->pt_strip_host.data(), ns);
  b->pt_parts.alloc((size_t)6 * members);
  b->pt_cnt.alloc(ns);
  b->pt_abort.alloc(1);
  if (!b->pt_cancel) {
    AFEM_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->pt_cancel), sizeof(int),
                           hipHostMallocMapped | hipHostMallocCoherent));
  }
  *b->pt_cancel = 0;
  b->pt_layout_budget = budget;
  b->pt_layout_local = local ? 1 : 0;
  drop_persistent_graph(b);
  return true;
}

// XCD-local teams need the device to dispatch launched workgroup g to XCD
// g mod 8 (8 XCDs of equal size): one probe launch of the layout's grid
// records each workgroup's XCD id, and every class g mod 8 must show one id.
bool probe_xcd_dispatch(afem_batch* b)
{
  if (rt().num_cu <= 0 || rt().num_cu % 8 != 0)
    return false;
  const int n = std::max(b->pt_grid, 64);
  DBuf<int> out(n);
  hipLaunchKernelGGL(k_xcc_probe, dim3(n), dim3(64), 0, stream(), out.p);
  check_launch("batch_xcc_probe");
  std::vector<int> h(n);
  AFEM_HIP(hipMemcpyAsync(h.data(), out.p, sizeof(int) * n, hipMemcpyDeviceToHost, stream()));
  AFEM_HIP(hipStreamSynchronize(stream()));
  for (int c = 0; c < 8; ++c)
    b->pt_xcc[c] = h[c];
  for (int g = 0; g < n; ++g)
    if (h[g] != h[g & 7])
      return false;
  for (int c = 0; c < 8; ++c)
    for (int d = 0; d < c; ++d)
      if (h[c] == h[d])
        return false;
  return true;
}

struct PersistentArgs
{
  Team T;
  TeamVec V;
  TeamParts P;
};

PersistentArgs persistent_args(afem_batch* b, long long spin_limit)
{
  afem_system* s = b->sys;
  PersistentArgs a;
  std::memset(&a, 0, sizeof(a)); // padding too: the graph key hashes the bytes
  Team& T = a.T;
  T.seg_of = b->pt_seg_of.p;
  T.slot_of = b->pt_slot_of.p;
  T.first = b->pt_first.p;
  T.rows = b->d_rows.p;
  T.strip = b->pt_strip.p;
  T.sell = b->sell_ok ? 1 : 0;
  T.sslice = b->d_sslice.p;
  T.soff = b->d_soff.p;
  T.sdcol = b->sdcol.p;
  T.rp = s->row_index;
  T.ci = s->columns;
  T.v32 = b->sell_ok ? b->sv32.p : b->v32.p;
  T.v64 = b->sell_ok ? b->sv64.p : s->values;
  T.cnt = b->pt_cnt.p;
  T.abort_flag = b->pt_abort.p;
  int* dev_cancel = nullptr;
  AFEM_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&dev_cancel), b->pt_cancel, 0));
  T.host_cancel = dev_cancel;
  T.spin_limit = spin_limit;
  for (int c = 0; c < 8; ++c)
    T.xcc[c] = b->pt_xcc[c];
  a.V = TeamVec{ s->x.p, s->rhs.p, b->dinv.p, b->r32.p, b->z32.p, { b->p32.p, b->p32b.p }, b->q32.p, b->r64.p };
  const int n = b->pt_first_host[b->nb_seg];
  double* pp = b->pt_parts.p;
  a.P = TeamParts{ pp, pp + n, pp + 2 * n, pp + 3 * n, pp + 4 * n, pp + 5 * n };
  return a;
}

// FNV-1a over the launch arguments: a captured graph is replayed only with
// exactly the buffers and settings it was captured with.
uint64_t persistent_key(const PersistentArgs& a, int grid, int local)
{
  uint64_t h = 1469598103934665603ULL;
  auto mix = [&h](const void* p, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) {
      h ^= c[i];
      h *= 1099511628211ULL;
    }
  };
  mix(&a, sizeof(a));
  mix(&grid, sizeof(grid));
  mix(&local, sizeof(local));
  return h;
}

void launch_persistent(afem_batch* b, const PersistentArgs& a, bool local)
{
  hipLaunchKernelGGL(k_persistent_clear, dim3(1), dim3(256), 0, stream(), b->pt_cnt.p, b->nb_seg, b->pt_abort.p);
  if (local)
    hipLaunchKernelGGL(k_persistent<true>, dim3(b->pt_grid), dim3(kBT), 0, stream(), a.T, a.V, a.P, b->st.p);
  else
    hipLaunchKernelGGL(k_persistent<false>, dim3(b->pt_grid), dim3(kBT), 0, stream(), a.T, a.V, a.P, b->st.p);
}

// Runs the persistent launch once (captured in a hipGraph and replayed); the
// host waits for it with a bounded wall time and raises the cancel flag when
// that passes.  Returns the abort flag (0: finished).
int persistent_once(afem_batch* b, const afem_batch_options& o, bool local, bool& cancelled)
{
  const long long spin_limit = o.spin_limit > 0 ? o.spin_limit : (1LL << 21);
  *b->pt_cancel = 0;
  const PersistentArgs args = persistent_args(b, spin_limit);
  if (o.use_graph) {
    const uint64_t key = persistent_key(args, b->pt_grid, local ? 1 : 0);
    if (!b->pt_graph || b->pt_graph_key != key) {
      drop_persistent_graph(b);
      hipGraph_t gr;
      AFEM_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
      launch_persistent(b, args, local);
      AFEM_HIP(hipStreamEndCapture(stream(), &gr));
      const hipError_t e = hipGraphInstantiate(&b->pt_graph, gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      AFEM_HIP(e);
      b->pt_graph_key = key;
    }
    AFEM_HIP(hipGraphLaunch(b->pt_graph, stream()));
  }
  else {
    launch_persistent(b, args, local);
  }
  check_launch("batch_persistent");
  const auto t0 = std::chrono::steady_clock::now();
  const double limit_ms = o.timeout_ms > 0.0 ? o.timeout_ms : 60000.0;
  cancelled = false;
  for (;;) {
    const hipError_t q = hipStreamQuery(stream());
    if (q == hipSuccess)
      break;
    AFEM_REQUIRE(q == hipErrorNotReady, AFEM_ERR_HIP, std::string("persistent solve: ") + hipGetErrorString(q));
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (!cancelled && ms > limit_ms) {
      __atomic_store_n(b->pt_cancel, 1, __ATOMIC_SEQ_CST);
      cancelled = true;
    }
    std::this_thread::sleep_for(std::chrono::microseconds(50));
  }
  int aborted = 0;
  AFEM_HIP(hipMemcpyAsync(&aborted, b->pt_abort.p, sizeof(int), hipMemcpyDeviceToHost, stream()));
  AFEM_HIP(hipStreamSynchronize(stream()));
  *b->pt_cancel = 0;
  return aborted;
}

// Solves the active segments in one persistent launch.  persistent = 1 uses
// XCD-local teams when the device's dispatch allows it (probed once per
// batch) and the packing fits, else the coherent exchange; persistent = 2
// always uses the coherent exchange.
void run_persistent(afem_batch* b, const afem_batch_options& o)
{
  const int cap = persistent_capacity();
  AFEM_REQUIRE(cap >= b->nb_seg, AFEM_ERR_CAPACITY,
               "the persistent batched solve needs one resident workgroup per system at least");
  const int want = o.workgroups > 0 ? o.workgroups : cap;
  const int budget = std::max(b->nb_seg, std::min(want, cap));
  bool local = o.persistent == 1 && b->pt_local_ok != 0;
  if (b->p32b.n != b->p32.n) {
    b->p32b.alloc(b->p32.n);
    drop_persistent_graph(b);
  }
  if (b->pt_layout_budget != budget || b->pt_layout_local != (local ? 1 : 0) ||
      (size_t)b->pt_first_host.size() != (size_t)b->nb_seg + 1) {
    if (local && !persistent_layout(b, budget, cap, true))
      local = false;
    if (!local)
      persistent_layout(b, budget, cap, false);
  }
  if (local && b->pt_local_ok < 0) {
    b->pt_local_ok = probe_xcd_dispatch(b) ? 1 : 0;
    local = b->pt_local_ok == 1;
    if (!local)
      persistent_layout(b, budget, cap, false);
  }
  afem_system* s = b->sys;
  if (local) {
    b->pt_xsave.reserve((size_t)s->nb_dof);
    AFEM_HIP(hipMemcpyAsync(b->pt_xsave.p, s->x.p, sizeof(double) * s->nb_dof, hipMemcpyDeviceToDevice, stream()));
  }
  bool cancelled = false;
  int aborted = persistent_once(b, o, local, cancelled);
  if (local && aborted == 2 && !cancelled) {
    // a workgroup ran on another XCD than probed: restore x and the states,
    // and solve with the coherent exchange from now on
    b->pt_local_ok = 0;
    AFEM_HIP(hipMemcpyAsync(s->x.p, b->pt_xsave.p, sizeof(double) * s->nb_dof, hipMemcpyDeviceToDevice, stream()));
    write_states(b);
    persistent_layout(b, budget, cap, false);
    aborted = persistent_once(b, o, false, cancelled);
  }
  b->pt_last_mode = aborted == 0 && local && b->pt_local_ok == 1 ? 1 : 2;
  AFEM_REQUIRE(!aborted && !cancelled, AFEM_ERR_DEVICE,
               cancelled ? "persistent batched solve cancelled after timeout_ms"
                         : "persistent batched solve gave up a barrier wait (spin_limit)");
}

} // namespace

extern "C" {

AFEM_API void afem_batch_default_options(afem_batch_options* o)
{
  if (!o)
    return;
  o->rtol = 1e-8;
  o->atol = 0.0;
  o->max_iter = 100000;
  o->reliable_delta = 0.1;
  o->chunk = 32;
  o->use_graph = 1;
  o->use_initial_guess = 0;
  o->persistent = 0;
  o->workgroups = 0;
  o->spin_limit = 0;
  o->timeout_ms = 0.0;
}

AFEM_API int afem_batch_create(afem_system_t sys, int nb_segment, const int32_t* segment_rows, afem_batch_t* out)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(sys && segment_rows && out, AFEM_ERR_ARGUMENT, "null argument");
  AFEM_REQUIRE(nb_segment >= 1, AFEM_ERR_ARGUMENT, "at least one segment");
  AFEM_REQUIRE(segment_rows[0] == 0 && segment_rows[nb_segment] == sys->nb_dof, AFEM_ERR_ARGUMENT,
               "segment offsets must start at 0 and end at the number of DoFs");
  for (int k = 0; k < nb_segment; ++k)
    AFEM_REQUIRE(segment_rows[k + 1] > segment_rows[k], AFEM_ERR_ARGUMENT, "empty or decreasing segment");
  auto* b = new afem_batch();
  b->sys = sys;
  b->nb_seg = nb_segment;
  b->seg_rows.assign(segment_rows, segment_rows + nb_segment + 1);
  b->d_rows.alloc(nb_segment + 1);
  b->d_rows.upload(b->seg_rows.data(), b->seg_rows.size());
  b->st.alloc(nb_segment);
  AFEM_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->host_st), sizeof(SegState) * nb_segment, hipHostMallocDefault));
  layout(b);
  *out = b;
  AFEM_GUARD_END
}

AFEM_API int afem_batch_persistent_mode(afem_batch_t b, int* mode)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(b && mode, AFEM_ERR_ARGUMENT, "null argument");
  *mode = b->pt_last_mode;
  AFEM_GUARD_END
}

AFEM_API int afem_batch_destroy(afem_batch_t b)
{
  AFEM_GUARD_BEGIN
  delete b;
  AFEM_GUARD_END
}

AFEM_API int afem_batch_set_tiles(afem_batch_t b, int lanes, int tiles)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(b, AFEM_ERR_ARGUMENT, "null argument");
  AFEM_REQUIRE(lanes == 1 || lanes == 4 || lanes == 8 || lanes == 16, AFEM_ERR_ARGUMENT,
               "lanes must be 1 (SELL), 4, 8 or 16");
  AFEM_REQUIRE(tiles >= 1 && tiles <= 16, AFEM_ERR_ARGUMENT, "tiles must be in [1, 16]");
  b->lanes = lanes;
  b->tiles = tiles;
  layout(b);
  AFEM_GUARD_END
}

AFEM_API int afem_batch_solve(afem_batch_t b, const afem_batch_options* opt, afem_batch_result* results,
                              double* total_ms)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(b && opt, AFEM_ERR_ARGUMENT, "null argument");
  afem_system* s = b->sys;
  system_require_matrix(s);
  AFEM_REQUIRE(s->block == 1, AFEM_ERR_NOT_IMPLEMENTED, "batched solves are for scalar systems");
  AFEM_REQUIRE(opt->rtol >= 0.0 && opt->atol >= 0.0 && opt->max_iter >= 0, AFEM_ERR_ARGUMENT, "bad tolerances");
  AFEM_REQUIRE(opt->persistent >= 0 && opt->persistent <= 2, AFEM_ERR_ARGUMENT, "persistent must be 0, 1 or 2");
  const auto t0 = std::chrono::steady_clock::now();
  prepare(b);
  if (!opt->use_initial_guess) {
    hipLaunchKernelGGL(k_initial_guess, dim3(grid_for(s->nb_dof, 256)), dim3(256), 0, stream(), s->nb_dof,
                       s->elim_info.p, s->elim_value.p, s->x.p);
    check_launch("batch_initial_guess");
  }
  const double delta = opt->reliable_delta > 0.0 ? opt->reliable_delta : 0.1;
  std::vector<int> iters(b->nb_seg, 0), nru(b->nb_seg, 0), conv(b->nb_seg, 0);
  std::vector<double> relres(b->nb_seg, 0.0);
  std::vector<char> want(b->nb_seg, 1);
  for (int round = 0; round < 5; ++round) {
    bool any = false;
    for (int k = 0; k < b->nb_seg; ++k) {
      SegState z{};
      z.rtol2 = opt->rtol * opt->rtol;
      z.atol2 = opt->atol * opt->atol;
      z.delta2 = delta * delta;
      z.max_iter = std::max(0, opt->max_iter - iters[k]);
      z.active = (want[k] && z.max_iter > 0) ? 1 : 0;
      any = any || z.active;
      b->host_st[k] = z;
    }
    if (!any)
      break;
    write_states(b);
    if (opt->persistent)
      run_persistent(b, *opt);
    else
      run(b, *opt, opt->use_graph != 0);
    // FP64 true residual of every segment that ran
    read_states(b);
    std::vector<SegState> after(b->host_st, b->host_st + b->nb_seg);
    launch_residual(b, 0, 1);
    hipLaunchKernelGGL(k_true_finalize, dim3(b->nb_seg), dim3(kBT), 0, stream(), seg_of(b), b->st.p, parts_of(b));
    check_launch("batch_true_residual");
    read_states(b);
    bool again = false;
    for (int k = 0; k < b->nb_seg; ++k) {
      if (!after[k].active)
        continue;
      iters[k] += after[k].iter;
      nru[k] += after[k].nb_ru;
      const double bn = std::sqrt(after[k].bb);
      const double tr = std::sqrt(b->host_st[k].rr_true);
      relres[k] = bn > 0.0 ? tr / bn : tr;
      conv[k] = (relres[k] <= opt->rtol || tr <= opt->atol) ? 1 : 0;
      // the recursive residual met the tolerance but the true one does not:
      // run that segment again from its current iterate
      want[k] = (!conv[k] && after[k].converged && after[k].iter > 0 && std::isfinite(relres[k])) ? 1 : 0;
      again = again || want[k];
    }
    if (!again)
      break;
  }
  AFEM_HIP(hipStreamSynchronize(stream()));
  const auto t1 = std::chrono::steady_clock::now();
  bool all = true;
  for (int k = 0; k < b->nb_seg; ++k) {
    if (results) {
      results[k].iterations = iters[k];
      results[k].converged = conv[k];
      results[k].reliable_updates = nru[k];
      results[k].relres = relres[k];
    }
    all = all && conv[k];
  }
  if (total_ms)
    *total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  return all ? AFEM_OK : AFEM_ERR_NOT_CONVERGED;
  AFEM_GUARD_END
}

// Times one chunk of iterations for each (lanes, tiles) candidate from the
// initial state and keeps the fastest; x is left at the initial guess.
AFEM_API int afem_batch_autotune(afem_batch_t b, int* lanes, int* tiles, double* best_ms)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(b, AFEM_ERR_ARGUMENT, "null argument");
  afem_system* s = b->sys;
  system_require_matrix(s);
  prepare(b);
  struct Cand
  {
    int lanes, tiles;
  };
  std::vector<Cand> cands = { { 4, 1 }, { 4, 2 }, { 8, 1 }, { 8, 2 }, { 8, 4 }, { 16, 1 }, { 16, 2 } };
  if (b->sell_ok)
    for (int t : { 1, 2, 4 })
      cands.push_back({ 1, t });
  hipEvent_t e0, e1;
  AFEM_HIP(hipEventCreate(&e0));
  AFEM_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  Cand pick = cands[2];
  afem_batch_options o;
  afem_batch_default_options(&o);
  for (const Cand& c : cands) {
    b->lanes = c.lanes;
    b->tiles = c.tiles;
    layout(b);
    float t_best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      hipLaunchKernelGGL(k_initial_guess, dim3(grid_for(s->nb_dof, 256)), dim3(256), 0, stream(), s->nb_dof,
                         s->elim_info.p, s->elim_value.p, s->x.p);
      for (int k = 0; k < b->nb_seg; ++k) {
        SegState z{};
        z.rtol2 = 0.0; // run every iteration of the chunk
        z.delta2 = 0.01;
        z.max_iter = 1 << 30;
        z.active = 1;
        b->host_st[k] = z;
      }
      write_states(b);
      const Seg g = seg_of(b);
      const BParts P = parts_of(b);
      launch_residual(b, 0, 1);
      hipLaunchKernelGGL(k_init, dim3(b->nblk), dim3(kBT), 0, stream(), g, b->r64.p, s->rhs.p, b->r32.p, b->z32.p,
                         b->p32.p, b->dinv.p, b->st.p, P);
      hipLaunchKernelGGL(k_init_finalize, dim3(b->nb_seg), dim3(kBT), 0, stream(), g, b->st.p, P);
      AFEM_HIP(hipEventRecord(e0, stream()));
      run_chunk(b, 16, true);
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
  b->lanes = pick.lanes;
  b->tiles = pick.tiles;
  layout(b);
  hipLaunchKernelGGL(k_initial_guess, dim3(grid_for(s->nb_dof, 256)), dim3(256), 0, stream(), s->nb_dof,
                     s->elim_info.p, s->elim_value.p, s->x.p);
  AFEM_HIP(hipStreamSynchronize(stream()));
  if (lanes)
    *lanes = pick.lanes;
  if (tiles)
    *tiles = pick.tiles;
  if (best_ms)
    *best_ms = best / 16.0;
  AFEM_GUARD_END
}

} // extern "C"
