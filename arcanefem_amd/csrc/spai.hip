// spai.hip -- learned sparse approximate inverse (SPAI) preconditioner.
//
// No counterpart in the reference, whose PCG preconditioners are Hypre's /
// PETSc's (femutils/HypreDoFLinearSystem.cc:596-769,
// femutils/PetscDoFLinearSystemFactory.axl:5-10).  A small message-passing
// network on the matrix graph predicts a preconditioner M with the sparsity
// pattern of A; it is trained without labels on the systems the FP64
// assembly produces and applied inside the flexible PCG of solver.hip.
//
// Scaling.  The network sees the Jacobi-scaled matrix
//   S = D^-1/2 A D^-1/2   (D = diag A),
// with every Dirichlet row (eliminated or penalised) and its column replaced
// by the identity, so S is symmetric with unit diagonal.  It predicts Mh for
// S; the solver applies M = D^-1/2 Mh D^-1/2, and M A is similar to Mh S.
//
// Model (F = 16 input features, 8 used; H = 32; 3 layers):
//   h0 = relu(x W0 + b0)
//   h_l = h_{l-1} + relu([h_{l-1}, a_l] W_l + b_l),  a_l,i = sum_{j != i} |s_ij| h_{l-1},j
//   u  = h3 Wu
//   Mh_ii = softplus(h3_i . wd + bd)                    (1 on Dirichlet rows)
//   Mh_ij = s_ij (c + u_i . u_j)                        (j != i)
// Mh is symmetric by construction, and with Wu = 0, c = 0, wd = 0 and
// bd = log(e - 1) it is the identity: the untrained model is Jacobi.
//
// Loss (8 Rademacher probes z_p per system, y_p = S z_p):
//   L = 1/(8 N) sum_p || z_p - Mh y_p ||^2,
// i.e. ||(I - Mh S) z||^2 for the scaled system.
//
// Precision: weights are kept in FP32 (Adam) and used as BF16; activations
// h, a, u and pre-activations are stored as BF16; the dense node layers run
// on MFMA (v_mfma_f32_32x32x8_bf16, FP32 accumulation); everything else
// computes in FP32 and reduces in FP64.
#include "common.h"
#include "device_utils.h"

#include <cmath>
#include <vector>

namespace afem {
namespace spai {

constexpr int kF = 16; // input features (MFMA K step multiple)
constexpr int kH = 32; // hidden width
constexpr int kL = 3;  // message-passing layers
constexpr int kP = 8;  // probes
constexpr int kUsedFeatures = 8;

// FP32 parameter layout
constexpr int oW0 = 0;                        // kF x kH  (row k = input feature)
constexpr int oB0 = oW0 + kF * kH;            // kH
constexpr int kLayerSize = 2 * kH * kH + kH;  // W_l (2kH x kH) then b_l
constexpr int oLayer = oB0 + kH;              // kL layers
constexpr int oWu = oLayer + kL * kLayerSize; // kH x kH
constexpr int oWd = oWu + kH * kH;            // kH
constexpr int oBd = oWd + kH;                 // 1
constexpr int oC = oBd + 1;                   // 1
constexpr int kNumParams = oC + 1;

constexpr int layer_w(int l) { return oLayer + (l - 1) * kLayerSize; }
constexpr int layer_b(int l) { return layer_w(l) + 2 * kH * kH; }

// BF16 weight copies: W as stored ([K][32]) and transposed ([32][K])
constexpr int bW0 = 0, bW0t = bW0 + kF * kH;
constexpr int bLayer = bW0t + kF * kH; // per layer: W [64][32], Wt [32][64]
constexpr int bWu = bLayer + kL * 2 * (2 * kH * kH);
constexpr int bWut = bWu + kH * kH;
constexpr int kNumBf16 = bWut + kH * kH;
constexpr int bf_w(int l) { return bLayer + (l - 1) * 2 * (2 * kH * kH); }
constexpr int bf_wt(int l) { return bf_w(l) + 2 * kH * kH; }

constexpr int kHeadParts = kH + 2; // dwd[32], dbd, dc
constexpr int kDwBlocks = 256;
constexpr int kMaxDwK = 2 * kH;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// ---- BF16 helpers (round to nearest even) ------------------------------------------------------

__host__ __device__ __forceinline__ uint16_t f2bf(float f)
{
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u)
    return 0x7fc0; // NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__host__ __device__ __forceinline__ float bf2f(uint16_t h)
{
  const uint32_t u = (uint32_t)h << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

__device__ __forceinline__ bf16x4 load_bf4(const uint16_t* p)
{
  const uint2 w = *reinterpret_cast<const uint2*>(p);
  bf16x4 r;
  r[0] = (short)(w.x & 0xffffu);
  r[1] = (short)(w.x >> 16);
  r[2] = (short)(w.y & 0xffffu);
  r[3] = (short)(w.y >> 16);
  return r;
}

__device__ __forceinline__ float softplus(float x) { return x > 20.f ? x : log1pf(expf(x)); }
__device__ __forceinline__ float sigmoid(float x) { return 1.f / (1.f + expf(-x)); }

// Rademacher probe signs of row j: bit p of the hash is probe p.
__device__ __forceinline__ uint32_t probe_bits(uint64_t seed, int32_t j)
{
  uint64_t z = seed * 0x9E3779B97F4A7C15ULL + (uint64_t)j + 0x632BE59BD9B4E019ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return (uint32_t)(z & 0xffu);
}
__device__ __forceinline__ float probe(uint32_t bits, int p) { return ((bits >> p) & 1u) ? 1.f : -1.f; }

// ---- block reduction of several values ----------------------------------------------------------

// Sums v[0..NV) over the 256 threads of the block; thread 0 writes out[0..NV).
template <int NV>
__device__ void block_sum_write(const double* v, double* __restrict__ out)
{
  __shared__ double lds[NV][4];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const double s = wave_reduce_sum(v[q]);
    if ((threadIdx.x & 63) == 0)
      lds[q][threadIdx.x >> 6] = s;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < NV; q += blockDim.x)
    out[q] = lds[q][0] + lds[q][1] + lds[q][2] + lds[q][3];
}

// ---- graph features ---------------------------------------------------------------------------------

// s_i = 1/sqrt(A_ii); Dirichlet rows (eliminated, penalised, or a
// non-positive diagonal) are marked and get s = 1.
__global__ void k_scale(int32_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                        const double* __restrict__ v, const uint8_t* __restrict__ elim,
                        const uint8_t* __restrict__ forced, float* __restrict__ s, uint8_t* __restrict__ ident)
{
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  double d = 0.0;
  for (int32_t k = rp[i]; k < rp[i + 1]; ++k)
    if (ci[k] == i)
      d = v[k];
  const bool id = elim[i] != 0 || forced[i] != 0 || !(d > 0.0);
  ident[i] = id ? 1 : 0;
  s[i] = id ? 1.f : (float)(1.0 / sqrt(d));
}

// Scaled entries s_ij and the node features (stored as BF16):
//   0: 1   1: sum_j s_ij   2: sum_j |s_ij|   3: max_j |s_ij|   4: sum_j s_ij^2
//   5: (number of j with s_ij != 0) / 16   6: Dirichlet flag   7: min_j |s_ij| over s_ij != 0
// (sums over the off-diagonal entries).
__global__ void k_features(int32_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                           const double* __restrict__ v, const float* __restrict__ s,
                           const uint8_t* __restrict__ ident, float* __restrict__ shat, uint16_t* __restrict__ x)
{
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const bool idi = ident[i] != 0;
  const double si = s[i];
  float f1 = 0.f, f2 = 0.f, f3 = 0.f, f4 = 0.f, cnt = 0.f, f7 = 0.f;
  const int32_t e = rp[i + 1];
  // four entries at a time: their loads, then their neighbours' scales, in flight together
  for (int32_t k0 = rp[i]; k0 < e; k0 += 4) {
    int32_t j[4];
    double vv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool in = k0 + u < e;
      j[u] = in ? ci[k0 + u] : i;
      vv[u] = in ? v[k0 + u] : 0.0;
    }
    float sj[4];
    bool idj[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      sj[u] = s[j[u]];
      idj[u] = ident[j[u]] != 0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (k0 + u >= e)
        break;
      float a;
      if (j[u] == i)
        a = idi ? 1.f : (float)(vv[u] * si * si);
      else
        a = (idi || idj[u]) ? 0.f : (float)(vv[u] * si * (double)sj[u]);
      shat[k0 + u] = a;
      if (j[u] != i && a != 0.f) {
        const float m = fabsf(a);
        f1 += a;
        f2 += m;
        f3 = fmaxf(f3, m);
        f4 += a * a;
        f7 = (cnt == 0.f) ? m : fminf(f7, m);
        cnt += 1.f;
      }
    }
  }
  const float f[kF] = { 1.f, f1, f2, f3, f4, cnt * (1.f / 16.f), idi ? 1.f : 0.f, f7, 0, 0, 0, 0, 0, 0, 0, 0 };
  // the row's 16 BF16 features as two 16-byte stores
  uint32_t w[kF / 2];
#pragma unroll
  for (int q = 0; q < kF / 2; ++q)
    w[q] = (uint32_t)f2bf(f[2 * q]) | ((uint32_t)f2bf(f[2 * q + 1]) << 16);
  uint4* xo = reinterpret_cast<uint4*>(x + (int64_t)i * kF);
  xo[0] = make_uint4(w[0], w[1], w[2], w[3]);
  xo[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

// ---- MFMA dense layers --------------------------------------------------------------------------
//
// v_mfma_f32_32x32x8_bf16: D(32x32) += A(32x8) B(8x32).  Lane l supplies
// A[l%32][4(l/32)+v] and B[4(l/32)+v][l%32] (v = 0..3) and holds
// D[8(r/4) + 4(l/32) + r%4][l%32] in accumulator r (r = 0..15).

__device__ __forceinline__ int acc_row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// out = [relu](X W + b) [+ resid] for a tile of 32 nodes per wave.
//   K = 16: X1 is N x 16.  K = 32: X1 is N x 32.  K = 64: X = [X1 | X2], both N x 32.
//   Wt: BF16 [32][K] (transposed weights).  bias may be null.
//   pre (BF16, may be null) receives X W + b before the activation.
template <int K>
__global__ void __launch_bounds__(256)
k_gemm_fwd(int32_t n, const uint16_t* __restrict__ X1, const uint16_t* __restrict__ X2,
           const uint16_t* __restrict__ Wt, const float* __restrict__ bias, const uint16_t* __restrict__ resid,
           uint16_t* __restrict__ pre, uint16_t* __restrict__ out, int relu)
{
  constexpr int LD1 = (K == 64) ? 32 : K;
  const int lane = threadIdx.x & 63;
  const int64_t row0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 32;
  if (row0 >= n)
    return;
  const int64_t arow = row0 + (lane & 31);
  const bool in = arow < n;
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r)
    acc[r] = 0.f;
#pragma unroll
  for (int k0 = 0; k0 < K; k0 += 8) {
    const int kk = k0 + 4 * (lane >> 5);
    bf16x4 a = { 0, 0, 0, 0 };
    if (in)
      a = (K == 64 && kk >= 32) ? load_bf4(X2 + arow * 32 + (kk - 32)) : load_bf4(X1 + arow * LD1 + kk);
    const bf16x4 b = load_bf4(Wt + (lane & 31) * K + kk);
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
  }
  const int col = lane & 31;
  const float bcol = bias ? bias[col] : 0.f;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int64_t row = row0 + acc_row(r, lane);
    if (row >= n)
      continue;
    float val = acc[r] + bcol;
    if (pre)
      pre[row * 32 + col] = f2bf(val);
    if (relu)
      val = fmaxf(val, 0.f);
    if (resid)
      val += bf2f(resid[row * 32 + col]);
    out[row * 32 + col] = f2bf(val);
  }
}

// dY masked by pre > 0 (when pre is given), as four BF16 for an MFMA operand.
__device__ __forceinline__ bf16x4 grad4(const float* __restrict__ dY, const uint16_t* __restrict__ pre, int64_t row,
                                        int j0)
{
  const float4 g = *reinterpret_cast<const float4*>(dY + row * 32 + j0);
  float gv[4] = { g.x, g.y, g.z, g.w };
  if (pre) {
    const uint2 w = *reinterpret_cast<const uint2*>(pre + row * 32 + j0);
    const uint16_t pv[4] = { (uint16_t)(w.x & 0xffffu), (uint16_t)(w.x >> 16), (uint16_t)(w.y & 0xffffu),
                             (uint16_t)(w.y >> 16) };
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (!(bf2f(pv[v]) > 0.f))
        gv[v] = 0.f;
  }
  bf16x4 r;
#pragma unroll
  for (int v = 0; v < 4; ++v)
    r[v] = (short)f2bf(gv[v]);
  return r;
}

// dX = (dY o [pre > 0]) W^T for 32 nodes per wave; W is BF16 [K][32].
// Output columns [0, 32) go to out0, [32, 64) (K = 64) to out1.  With base0
// given, out0 = base0 + result (the residual branch); otherwise out0 = result
// (accumulate0 = 0) or out0 += result (accumulate0 = 1).  out1 is written.
template <int K>
__global__ void __launch_bounds__(256)
k_gemm_dx(int32_t n, const float* __restrict__ dY, const uint16_t* __restrict__ pre, const uint16_t* __restrict__ W,
          const float* __restrict__ base0, float* __restrict__ out0, int accumulate0, float* __restrict__ out1)
{
  const int lane = threadIdx.x & 63;
  const int64_t row0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 32;
  if (row0 >= n)
    return;
  const int64_t arow = row0 + (lane & 31);
  const bool in = arow < n;
  bf16x4 a[4];
#pragma unroll
  for (int st = 0; st < 4; ++st) {
    const int j0 = 8 * st + 4 * (lane >> 5);
    a[st] = in ? grad4(dY, pre, arow, j0) : bf16x4{ 0, 0, 0, 0 };
  }
#pragma unroll
  for (int kt = 0; kt < K / 32; ++kt) {
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r)
      acc[r] = 0.f;
#pragma unroll
    for (int st = 0; st < 4; ++st) {
      const int j0 = 8 * st + 4 * (lane >> 5);
      // B[j][k] = W[k][j]: four consecutive j at column k
      const bf16x4 b = load_bf4(W + (int64_t)(kt * 32 + (lane & 31)) * 32 + j0);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a[st], b, acc, 0, 0, 0);
    }
    const int col = lane & 31;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int64_t row = row0 + acc_row(r, lane);
      if (row >= n)
        continue;
      const int64_t o = row * 32 + col;
      if (kt == 0) {
        if (base0)
          out0[o] = base0[o] + acc[r];
        else if (accumulate0)
          out0[o] += acc[r];
        else
          out0[o] = acc[r];
      }
      else {
        out1[o] = acc[r];
      }
    }
  }
}

// Partial dW = X^T (dY o [pre > 0]) and db = sum (dY o [pre > 0]) over a
// contiguous range of nodes per wave; one partial (K*32 + 32 floats) per
// block.  X as in k_gemm_fwd.
template <int K>
__global__ void __launch_bounds__(256)
k_gemm_dw(int32_t n, int32_t nodes_per_wave, const uint16_t* __restrict__ X1, const uint16_t* __restrict__ X2,
          const float* __restrict__ dY, const uint16_t* __restrict__ pre, float* __restrict__ partial)
{
  constexpr int LD1 = (K == 64) ? 32 : K;
  constexpr int KT = (K + 31) / 32;
  __shared__ float red[4][KT * 1024 + 32];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t wid = (int64_t)blockIdx.x * 4 + wave;
  const int64_t n0 = wid * nodes_per_wave;
  const int64_t n1 = min((int64_t)n, n0 + nodes_per_wave);
  f32x16 acc[KT];
#pragma unroll
  for (int kt = 0; kt < KT; ++kt)
#pragma unroll
    for (int r = 0; r < 16; ++r)
      acc[kt][r] = 0.f;
  float dbias = 0.f;
  const int col = lane & 31;
  for (int64_t base = n0; base < n1; base += 8) {
    const int64_t nn = base + 4 * (lane >> 5); // 4 nodes per lane half
    // B[n][j] = dYm[n][j]
    bf16x4 b;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t node = nn + v;
      float g = 0.f;
      if (node < n1) {
        g = dY[node * 32 + col];
        if (pre && !(bf2f(pre[node * 32 + col]) > 0.f))
          g = 0.f;
      }
      dbias += g;
      b[v] = (short)f2bf(g);
    }
#pragma unroll
    for (int kt = 0; kt < KT; ++kt) {
      // A[k][n] = X[n][k], k = kt*32 + lane%32
      const int k = kt * 32 + (lane & 31);
      bf16x4 a;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t node = nn + v;
        uint16_t xv = 0;
        if (node < n1 && k < K)
          xv = (K == 64 && k >= 32) ? X2[node * 32 + (k - 32)] : X1[node * LD1 + k];
        a[v] = (short)xv;
      }
      acc[kt] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc[kt], 0, 0, 0);
    }
  }
  // wave partials to LDS: D[k][j] of tile kt
#pragma unroll
  for (int kt = 0; kt < KT; ++kt)
#pragma unroll
    for (int r = 0; r < 16; ++r)
      red[wave][kt * 1024 + acc_row(r, lane) * 32 + col] = acc[kt][r];
  // the bias gradient: lanes l and l+32 hold halves of column l%32
  dbias += __shfl_xor(dbias, 32, kWave);
  if (lane < 32)
    red[wave][KT * 1024 + lane] = dbias;
  __syncthreads();
  float* out = partial + (int64_t)blockIdx.x * (K * 32 + 32);
  for (int q = threadIdx.x; q < K * 32; q += 256)
    out[q] = red[0][q] + red[1][q] + red[2][q] + red[3][q];
  if (threadIdx.x < 32) {
    const int q = KT * 1024 + threadIdx.x;
    out[K * 32 + threadIdx.x] = red[0][q] + red[1][q] + red[2][q] + red[3][q];
  }
}

// out[q] (+)= sum_b partial[b * stride + q] for q < len: 32 outputs per
// workgroup, 8 groups of lanes take every 8th partial, and the 8 group sums
// are added in a fixed order (deterministic).
__global__ void __launch_bounds__(256)
k_reduce_partials(int nparts, int stride, int len, const float* __restrict__ partial, float* __restrict__ out,
                  int accumulate)
{
  __shared__ double red[8][32];
  const int ql = threadIdx.x & 31, g = threadIdx.x >> 5;
  const int q = blockIdx.x * 32 + ql;
  double s = 0.0;
  if (q < len)
    for (int b = g; b < nparts; b += 8)
      s += partial[(int64_t)b * stride + q];
  red[g][ql] = s;
  __syncthreads();
  if (g == 0 && q < len) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      t += red[k][ql];
    out[q] = accumulate ? out[q] + (float)t : (float)t;
  }
}

// ---- message passing and heads ------------------------------------------------------------------

// a_i = sum_{j != i} |s_ij| h_j (BF16 in, FP32 accumulation, BF16 out).
// Four lanes per row, eight features (one 16-byte load) per lane; the row's
// entries are taken four at a time, so four gathers are in flight per lane
// (the entries are added in row order, as one lane per feature would).
constexpr int kAggLanes = 4;

__global__ void __launch_bounds__(256)
k_aggregate(int32_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
            const float* __restrict__ shat, const uint16_t* __restrict__ h, uint16_t* __restrict__ a)
{
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kAggLanes;
  const int q = threadIdx.x & (kAggLanes - 1); // features 8q .. 8q + 7
  if (i >= n)
    return;
  float acc[8];
#pragma unroll
  for (int t = 0; t < 8; ++t)
    acc[t] = 0.f;
  const int32_t e = rp[i + 1];
  for (int32_t k = rp[i]; k < e; k += 4) {
    int32_t j[4];
    float w[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool in = k + u < e;
      j[u] = in ? ci[k + u] : (int32_t)i;
      w[u] = (in && j[u] != i) ? fabsf(shat[k + u]) : 0.f;
    }
    uint4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      v[u] = *reinterpret_cast<const uint4*>(h + (int64_t)j[u] * 32 + 8 * q);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t ws[4] = { v[u].x, v[u].y, v[u].z, v[u].w };
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        acc[2 * t] += w[u] * __uint_as_float(ws[t] << 16);
        acc[2 * t + 1] += w[u] * __uint_as_float(ws[t] & 0xffff0000u);
      }
    }
  }
  uint32_t o[4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
    o[t] = (uint32_t)f2bf(acc[2 * t]) | ((uint32_t)f2bf(acc[2 * t + 1]) << 16);
  *reinterpret_cast<uint4*>(a + i * 32 + 8 * q) = make_uint4(o[0], o[1], o[2], o[3]);
}

// g_out_i += sum_{j != i} |s_ij| g_in_j in FP32 (|S| off the diagonal is
// symmetric, so this is also the transposed product).  Eight lanes per row,
// four features (one 16-byte load) per lane, four neighbours in flight; the
// entries are added in row order.
constexpr int kAggGradLanes = 8;

__global__ void __launch_bounds__(256)
k_aggregate_grad(int32_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                 const float* __restrict__ shat, const float* __restrict__ g_in, float* __restrict__ g_out)
{
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kAggGradLanes;
  const int q = threadIdx.x & (kAggGradLanes - 1); // features 4q .. 4q + 3
  if (i >= n)
    return;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  const int32_t e = rp[i + 1];
  for (int32_t k = rp[i]; k < e; k += 4) {
    int32_t j[4];
    float w[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool in = k + u < e;
      j[u] = in ? ci[k + u] : (int32_t)i;
      w[u] = (in && j[u] != i) ? fabsf(shat[k + u]) : 0.f;
    }
    float4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      v[u] = *reinterpret_cast<const float4*>(g_in + (int64_t)j[u] * 32 + 4 * q);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      acc.x += w[u] * v[u].x;
      acc.y += w[u] * v[u].y;
      acc.z += w[u] * v[u].z;
      acc.w += w[u] * v[u].w;
    }
  }
  float4* out = reinterpret_cast<float4*>(g_out + i * 32 + 4 * q);
  float4 o = *out;
  o.x += acc.x;
  o.y += acc.y;
  o.z += acc.z;
  o.w += acc.w;
  *out = o;
}

__device__ __forceinline__ void load_vec32(const uint16_t* __restrict__ p, float* v)
{
  const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const uint4 t = q[w];
    const uint32_t ws[4] = { t.x, t.y, t.z, t.w };
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      v[8 * w + 2 * e] = bf2f((uint16_t)(ws[e] & 0xffffu));
      v[8 * w + 2 * e + 1] = bf2f((uint16_t)(ws[e] >> 16));
    }
  }
}

// Mh (FP32, pattern of A) and the preconditioner M = D^-1/2 Mh D^-1/2
// (FP32) that the solver applies; pre_d keeps the diagonal pre-activation.
// Four lanes per row, eight features per lane (one 16-byte load of u_j);
// the row's entries are taken four at a time (four gathers in flight per
// lane), the partial dot products summed across the four lanes by a
// butterfly (every lane ends with the same bits, and u_i . u_j and u_j . u_i
// are formed from the same products in the same order, so Mh stays exactly
// symmetric), and lane u writes entry k + u.
constexpr int kHeadLanes = 4;

__device__ __forceinline__ void unpack8(uint4 v, float* f)
{
  const uint32_t ws[4] = { v.x, v.y, v.z, v.w };
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    f[2 * t] = __uint_as_float(ws[t] << 16);
    f[2 * t + 1] = __uint_as_float(ws[t] & 0xffff0000u);
  }
}

__device__ __forceinline__ float sum4(float v)
{
  v += __shfl_xor(v, 1, kHeadLanes);
  v += __shfl_xor(v, 2, kHeadLanes);
  return v;
}

__global__ void __launch_bounds__(256)
k_head(int32_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const float* __restrict__ shat,
       const float* __restrict__ s, const uint8_t* __restrict__ ident, const uint16_t* __restrict__ h3,
       const uint16_t* __restrict__ u, const float* __restrict__ params, float* __restrict__ mhat,
       float* __restrict__ mval, float* __restrict__ pre_d)
{
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kHeadLanes;
  const int q = threadIdx.x & (kHeadLanes - 1); // features 8q .. 8q + 7
  if (i >= n) // the four lanes of a row leave together
    return;
  const bool idi = ident[i] != 0;
  float hv[8];
  unpack8(*reinterpret_cast<const uint4*>(h3 + i * 32 + 8 * q), hv);
  float pd = 0.f;
#pragma unroll
  for (int t = 0; t < 8; ++t)
    pd += hv[t] * params[oWd + 8 * q + t];
  pd = params[oBd] + sum4(pd);
  if (q == 0)
    pre_d[i] = pd;
  float ui[8];
  unpack8(*reinterpret_cast<const uint4*>(u + i * 32 + 8 * q), ui);
  const float c = params[oC];
  const float si = s[i];
  const float diag = idi ? 1.f : softplus(pd);
  const int32_t e = rp[i + 1];
  for (int32_t k = rp[i]; k < e; k += 4) {
    int32_t j[4];
    float a[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const bool in = k + v < e;
      j[v] = in ? ci[k + v] : (int32_t)i;
      a[v] = in ? shat[k + v] : 0.f;
    }
    uint4 w[4];
#pragma unroll
    for (int v = 0; v < 4; ++v)
      w[v] = *reinterpret_cast<const uint4*>(u + (int64_t)j[v] * 32 + 8 * q);
    float d[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      float uj[8];
      unpack8(w[v], uj);
      float t = 0.f;
#pragma unroll
      for (int f = 0; f < 8; ++f)
        t += ui[f] * uj[f];
      d[v] = sum4(t);
    }
    // lane q writes entry k + q
    const int32_t jq = q == 0 ? j[0] : q == 1 ? j[1] : q == 2 ? j[2] : j[3];
    const float aq = q == 0 ? a[0] : q == 1 ? a[1] : q == 2 ? a[2] : a[3];
    const float dq = q == 0 ? d[0] : q == 1 ? d[1] : q == 2 ? d[2] : d[3];
    if (k + q < e) {
      const float m = (jq == i) ? diag : (aq != 0.f ? aq * (c + dq) : 0.f);
      mhat[k + q] = m;
      mval[k + q] = si * m * s[jq];
    }
  }
}

// y_p = S z_p for the 8 probes (N x 8, FP32).
__global__ void k_probe_y(int32_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                          const float* __restrict__ shat, uint64_t seed, float* __restrict__ y)
{
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  float acc[kP] = {};
  for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
    const float a = shat[k];
    const uint32_t bits = probe_bits(seed, ci[k]);
#pragma unroll
    for (int p = 0; p < kP; ++p)
      acc[p] += a * probe(bits, p);
  }
#pragma unroll
  for (int p = 0; p < kP; ++p)
    y[(int64_t)i * kP + p] = acc[p];
}

// r_p = z_p - Mh y_p; per-block partial of sum r^2.
__global__ void __launch_bounds__(256)
k_probe_r(int32_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const float* __restrict__ mhat,
          const float* __restrict__ y, uint64_t seed, float* __restrict__ r, double* __restrict__ part)
{
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  double sq = 0.0;
  if (i < n) {
    float acc[kP] = {};
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
      const float m = mhat[k];
      const float* yj = y + (int64_t)ci[k] * kP;
#pragma unroll
      for (int p = 0; p < kP; ++p)
        acc[p] += m * yj[p];
    }
    const uint32_t bits = probe_bits(seed, i);
#pragma unroll
    for (int p = 0; p < kP; ++p) {
      const float rv = probe(bits, p) - acc[p];
      r[(int64_t)i * kP + p] = rv;
      sq += (double)rv * rv;
    }
  }
  double v[1] = { sq };
  block_sum_write<1>(v, part + blockIdx.x);
}

// Backward of the heads.  With G_ij = dL/dMh_ij = -2/(8N) sum_p r_p,i y_p,j:
//   diagonal: dpre_d = G_ii sigmoid(pre_d); dh3_i = dpre_d wd; dwd += dpre_d h3_i; dbd += dpre_d
//   off-diagonal: dc += G_ij s_ij; du_i = sum_j (G_ij + G_ji) s_ij u_j
// (Mh_ij and Mh_ji are the same function of the pair, hence G_ij + G_ji).
// Per-block partials of (dwd[32], dbd, dc) in FP64.
__global__ void __launch_bounds__(256)
k_head_backward(int32_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                const float* __restrict__ shat, const uint8_t* __restrict__ ident, const uint16_t* __restrict__ h3,
                const uint16_t* __restrict__ u, const float* __restrict__ pre_d, const float* __restrict__ params,
                const float* __restrict__ y, const float* __restrict__ r, float coef, float* __restrict__ du,
                float* __restrict__ dh3, double* __restrict__ part)
{
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  double pv[kHeadParts] = {};
  if (i < n) {
    float dui[32] = {};
    float dh[32] = {};
    if (!ident[i]) {
      float ri[kP], yi[kP];
#pragma unroll
      for (int p = 0; p < kP; ++p) {
        ri[p] = r[(int64_t)i * kP + p];
        yi[p] = y[(int64_t)i * kP + p];
      }
      float hv[32];
      load_vec32(h3 + (int64_t)i * 32, hv);
      float dc = 0.f;
      for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
        const int32_t j = ci[k];
        const float* yj = y + (int64_t)j * kP;
        const float* rj = r + (int64_t)j * kP;
        float gij = 0.f, gji = 0.f;
#pragma unroll
        for (int p = 0; p < kP; ++p) {
          gij += ri[p] * yj[p];
          gji += rj[p] * yi[p];
        }
        gij *= coef;
        gji *= coef;
        if (j == i) {
          const float dpre = gij * sigmoid(pre_d[i]);
#pragma unroll
          for (int q = 0; q < 32; ++q) {
            dh[q] = dpre * params[oWd + q];
            pv[q] += (double)dpre * hv[q];
          }
          pv[kH] += dpre;
        }
        else {
          const float a = shat[k];
          if (a != 0.f) {
            dc += gij * a;
            float uj[32];
            load_vec32(u + (int64_t)j * 32, uj);
            const float w = (gij + gji) * a;
#pragma unroll
            for (int q = 0; q < 32; ++q)
              dui[q] += w * uj[q];
          }
        }
      }
      pv[kH + 1] = dc;
    }
    // the row's 32 + 32 values as 16-byte stores
    float4* duo = reinterpret_cast<float4*>(du + (int64_t)i * 32);
    float4* dho = reinterpret_cast<float4*>(dh3 + (int64_t)i * 32);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      duo[q] = make_float4(dui[4 * q], dui[4 * q + 1], dui[4 * q + 2], dui[4 * q + 3]);
      dho[q] = make_float4(dh[4 * q], dh[4 * q + 1], dh[4 * q + 2], dh[4 * q + 3]);
    }
  }
  block_sum_write<kHeadParts>(pv, part + (int64_t)blockIdx.x * kHeadParts);
}

// one workgroup per head parameter q
__global__ void __launch_bounds__(256)
k_reduce_head(int nparts, const double* __restrict__ part, float* __restrict__ grad)
{
  __shared__ double lds[4];
  const int q = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nparts; b += 256)
    s += part[(int64_t)b * kHeadParts + q];
  s = wave_reduce_sum(s);
  if ((threadIdx.x & 63) == 0)
    lds[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t = lds[0] + lds[1] + lds[2] + lds[3];
    if (q < kH)
      grad[oWd + q] = (float)t;
    else if (q == kH)
      grad[oBd] = (float)t;
    else
      grad[oC] = (float)t;
  }
}

__global__ void k_sum_double(int nparts, const double* __restrict__ part, double* __restrict__ out)
{
  __shared__ double lds[4];
  double s = 0.0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x)
    s += part[b];
  s = wave_reduce_sum(s);
  if ((threadIdx.x & 63) == 0)
    lds[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0)
    *out = lds[0] + lds[1] + lds[2] + lds[3];
}

// ---- optimiser and weight copies ----------------------------------------------------------------

__global__ void k_adam(int n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n)
    return;
  const float gq = g[q];
  const float mq = b1 * m[q] + (1.f - b1) * gq;
  const float vq = b2 * v[q] + (1.f - b2) * gq * gq;
  m[q] = mq;
  v[q] = vq;
  w[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

__global__ void k_refresh_bf16(const float* __restrict__ p, uint16_t* __restrict__ bf)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  // W0 [16][32]
  if (t < kF * kH) {
    const int k = t / kH, j = t % kH;
    const uint16_t b = f2bf(p[oW0 + t]);
    bf[bW0 + t] = b;
    bf[bW0t + j * kF + k] = b;
  }
  // layers [64][32]
  for (int l = 1; l <= kL; ++l)
    if (t < 2 * kH * kH) {
      const int k = t / kH, j = t % kH;
      const uint16_t b = f2bf(p[layer_w(l) + t]);
      bf[bf_w(l) + t] = b;
      bf[bf_wt(l) + j * (2 * kH) + k] = b;
    }
  // Wu [32][32]
  if (t < kH * kH) {
    const int k = t / kH, j = t % kH;
    const uint16_t b = f2bf(p[oWu + t]);
    bf[bWu + t] = b;
    bf[bWut + j * kH + k] = b;
  }
}

} // namespace spai
} // namespace afem

// ---- model handle -----------------------------------------------------------------------------------

using namespace afem;
using namespace afem::spai;

struct afem_spai
{
  DBuf<float> params, grad, adam_m, adam_v;
  DBuf<uint16_t> wbf;
  int64_t adam_t = 0;
  // per-system work arrays (grown on demand)
  int32_t cap_n = 0;
  int64_t cap_nnz = 0;
  DBuf<float> s, shat, mhat, mval, pre_d, y, r;
  DBuf<uint8_t> ident;
  DBuf<uint16_t> x, h[kL + 1], a[kL + 1], pre[kL + 1], u;
  DBuf<float> g0, g1, du, da, dwp;
  DBuf<double> part, loss;
  bool forward_valid = false;
};

namespace {

void host_init_params(uint64_t seed, std::vector<float>& p)
{
  p.assign(kNumParams, 0.f);
  uint64_t state = seed * 0x9E3779B97F4A7C15ULL + 0x1234567ULL;
  auto uni = [&]() { // uniform in [-1, 1)
    state += 0x9E3779B97F4A7C15ULL;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  };
  // He-uniform for the ReLU layers (only the used input features)
  const double a0 = std::sqrt(6.0 / kUsedFeatures);
  for (int k = 0; k < kUsedFeatures; ++k)
    for (int j = 0; j < kH; ++j)
      p[oW0 + k * kH + j] = (float)(a0 * uni());
  const double al = 0.5 * std::sqrt(6.0 / (2 * kH));
  for (int l = 1; l <= kL; ++l)
    for (int t = 0; t < 2 * kH * kH; ++t)
      p[layer_w(l) + t] = (float)(al * uni());
  // small head weights: the untrained model is (close to) Jacobi
  const double au = 0.1 / std::sqrt((double)kH);
  for (int t = 0; t < kH * kH; ++t)
    p[oWu + t] = (float)(au * uni());
  p[oBd] = (float)std::log(std::exp(1.0) - 1.0); // softplus(bd) = 1
  p[oC] = 0.f;
}

void refresh(afem_spai* m)
{
  hipLaunchKernelGGL(k_refresh_bf16, dim3((2 * kH * kH + 255) / 256), dim3(256), 0, stream(), m->params.p, m->wbf.p);
  check_launch("spai_refresh_bf16");
}

void ensure_capacity(afem_spai* m, int32_t n, int64_t nnz)
{
  if (n > m->cap_n) {
    m->s.alloc(n);
    m->pre_d.alloc(n);
    m->ident.alloc(n);
    m->y.alloc((size_t)n * kP);
    m->r.alloc((size_t)n * kP);
    m->x.alloc((size_t)n * kF);
    for (int l = 0; l <= kL; ++l) {
      m->h[l].alloc((size_t)n * kH);
      m->pre[l].alloc((size_t)n * kH);
      if (l > 0)
        m->a[l].alloc((size_t)n * kH);
    }
    m->u.alloc((size_t)n * kH);
    m->g0.alloc((size_t)n * kH);
    m->g1.alloc((size_t)n * kH);
    m->du.alloc((size_t)n * kH);
    m->da.alloc((size_t)n * kH);
    m->part.alloc((size_t)grid_for(n, 256) * kHeadParts);
    m->cap_n = n;
  }
  if (nnz > m->cap_nnz) {
    m->shat.alloc((size_t)nnz);
    m->mhat.alloc((size_t)nnz);
    m->mval.alloc((size_t)nnz);
    m->cap_nnz = nnz;
  }
}

void require_scalar_system(afem_system* s)
{
  AFEM_REQUIRE(s->has_matrix, AFEM_ERR_STATE, "no matrix: call setCSRValues first");
  AFEM_REQUIRE(s->block == 1, AFEM_ERR_NOT_IMPLEMENTED, "the learned preconditioner is for scalar systems");
  AFEM_REQUIRE(s->nb_dof > 0, AFEM_ERR_ARGUMENT, "empty system");
}

unsigned tiles32(int32_t n) { return (unsigned)(((int64_t)n + 127) / 128); }

void forward(afem_spai* m, afem_system* s)
{
  const int32_t n = s->nb_dof;
  ensure_capacity(m, n, s->nnz);
  const unsigned g = grid_for(n, 256);
  hipLaunchKernelGGL(k_scale, dim3(g), dim3(256), 0, stream(), n, s->row_index, s->columns, s->values,
                     s->elim_info.p, s->forced_info.p, m->s.p, m->ident.p);
  hipLaunchKernelGGL(k_features, dim3(g), dim3(256), 0, stream(), n, s->row_index, s->columns, s->values, m->s.p,
                     m->ident.p, m->shat.p, m->x.p);
  check_launch("spai_features");
  const uint16_t* W = m->wbf.p;
  const float* P = m->params.p;
  hipLaunchKernelGGL(k_gemm_fwd<kF>, dim3(tiles32(n)), dim3(256), 0, stream(), n, m->x.p, nullptr, W + bW0t,
                     P + oB0, nullptr, m->pre[0].p, m->h[0].p, 1);
  check_launch("spai_gemm_fwd0");
  const unsigned ga = (unsigned)(((int64_t)n * kAggLanes + 255) / 256);
  for (int l = 1; l <= kL; ++l) {
    hipLaunchKernelGGL(k_aggregate, dim3(ga), dim3(256), 0, stream(), n, s->row_index, s->columns, m->shat.p,
                       m->h[l - 1].p, m->a[l].p);
    hipLaunchKernelGGL(k_gemm_fwd<2 * kH>, dim3(tiles32(n)), dim3(256), 0, stream(), n, m->h[l - 1].p, m->a[l].p,
                       W + bf_wt(l), P + layer_b(l), m->h[l - 1].p, m->pre[l].p, m->h[l].p, 1);
    check_launch("spai_layer");
  }
  hipLaunchKernelGGL(k_gemm_fwd<kH>, dim3(tiles32(n)), dim3(256), 0, stream(), n, m->h[kL].p, nullptr, W + bWut,
                     nullptr, nullptr, nullptr, m->u.p, 0);
  hipLaunchKernelGGL(k_head, dim3((unsigned)(((int64_t)n * kHeadLanes + 255) / 256)), dim3(256), 0, stream(), n,
                     s->row_index, s->columns, m->shat.p, m->s.p,
                     m->ident.p, m->h[kL].p, m->u.p, P, m->mhat.p, m->mval.p, m->pre_d.p);
  check_launch("spai_head");
  m->forward_valid = true;
}

double probe_loss(afem_spai* m, afem_system* s, uint64_t seed)
{
  const int32_t n = s->nb_dof;
  const unsigned g = grid_for(n, 256);
  hipLaunchKernelGGL(k_probe_y, dim3(g), dim3(256), 0, stream(), n, s->row_index, s->columns, m->shat.p, seed,
                     m->y.p);
  hipLaunchKernelGGL(k_probe_r, dim3(g), dim3(256), 0, stream(), n, s->row_index, s->columns, m->mhat.p, m->y.p,
                     seed, m->r.p, m->part.p);
  hipLaunchKernelGGL(k_sum_double, dim3(1), dim3(256), 0, stream(), (int)g, m->part.p, m->loss.p);
  check_launch("spai_loss");
  double sum = 0.0;
  m->loss.download(&sum, 1);
  return sum / ((double)kP * n);
}

// dW (and db) of one dense layer into grad (w_off, b_off; b_off < 0: none).
template <int K>
void weight_grad(afem_spai* m, int32_t n, const uint16_t* X1, const uint16_t* X2, const float* dY,
                 const uint16_t* pre, int w_off, int b_off)
{
  const int waves = kDwBlocks * 4;
  int32_t per = (int32_t)(((int64_t)n + waves - 1) / waves);
  per = (per + 7) / 8 * 8;
  if (per < 8)
    per = 8;
  hipLaunchKernelGGL(k_gemm_dw<K>, dim3(kDwBlocks), dim3(256), 0, stream(), n, per, X1, X2, dY, pre, m->dwp.p);
  check_launch("spai_gemm_dw");
  const int stride = K * 32 + 32;
  hipLaunchKernelGGL(k_reduce_partials, dim3((K * 32 + 31) / 32), dim3(256), 0, stream(), kDwBlocks, stride,
                     K * 32, m->dwp.p, m->grad.p + w_off, 0);
  if (b_off >= 0)
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, stream(), kDwBlocks, stride, 32,
                       m->dwp.p + K * 32, m->grad.p + b_off, 0);
  check_launch("spai_reduce_dw");
}

void backward(afem_spai* m, afem_system* s)
{
  const int32_t n = s->nb_dof;
  const unsigned g = grid_for(n, 256);
  const uint16_t* W = m->wbf.p;
  const float coef = (float)(-2.0 / ((double)kP * n));
  // heads: du and dh3 (from the diagonal), head parameter gradients
  float* dh = m->g0.p;
  float* dh_next = m->g1.p;
  hipLaunchKernelGGL(k_head_backward, dim3(g), dim3(256), 0, stream(), n, s->row_index, s->columns, m->shat.p,
                     m->ident.p, m->h[kL].p, m->u.p, m->pre_d.p, m->params.p, m->y.p, m->r.p, coef, m->du.p, dh,
                     m->part.p);
  hipLaunchKernelGGL(k_reduce_head, dim3(kHeadParts), dim3(256), 0, stream(), (int)g, m->part.p, m->grad.p);
  check_launch("spai_head_backward");
  // u = h3 Wu
  weight_grad<kH>(m, n, m->h[kL].p, nullptr, m->du.p, nullptr, oWu, -1);
  hipLaunchKernelGGL(k_gemm_dx<kH>, dim3(tiles32(n)), dim3(256), 0, stream(), n, m->du.p, nullptr, W + bWu,
                     nullptr, dh, 1, nullptr);
  check_launch("spai_gemm_dx_u");
  const unsigned ga = (unsigned)(((int64_t)n * kAggGradLanes + 255) / 256);
  for (int l = kL; l >= 1; --l) {
    weight_grad<2 * kH>(m, n, m->h[l - 1].p, m->a[l].p, dh, m->pre[l].p, layer_w(l), layer_b(l));
    // dh_next = dh + (dpre W^T)[:, :32];  da = (dpre W^T)[:, 32:]
    hipLaunchKernelGGL(k_gemm_dx<2 * kH>, dim3(tiles32(n)), dim3(256), 0, stream(), n, dh, m->pre[l].p,
                       W + bf_w(l), dh, dh_next, 0, m->da.p);
    check_launch("spai_gemm_dx");
    hipLaunchKernelGGL(k_aggregate_grad, dim3(ga), dim3(256), 0, stream(), n, s->row_index, s->columns, m->shat.p,
                       m->da.p, dh_next);
    check_launch("spai_aggregate_grad");
    float* t = dh;
    dh = dh_next;
    dh_next = t;
  }
  weight_grad<kF>(m, n, m->x.p, nullptr, dh, m->pre[0].p, oW0, oB0);
}

} // namespace

extern "C" {

AFEM_API int afem_spai_num_params(void) { return kNumParams; }

AFEM_API int afem_spai_create(uint64_t seed, afem_spai_t* out)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(out, AFEM_ERR_ARGUMENT, "null argument");
  auto* m = new afem_spai();
  std::vector<float> p;
  host_init_params(seed, p);
  m->params.alloc(kNumParams);
  m->params.upload(p.data(), p.size());
  m->grad.alloc(kNumParams);
  m->grad.zero();
  m->adam_m.alloc(kNumParams);
  m->adam_m.zero();
  m->adam_v.alloc(kNumParams);
  m->adam_v.zero();
  m->wbf.alloc(kNumBf16);
  m->dwp.alloc((size_t)kDwBlocks * (kMaxDwK * 32 + 32));
  m->loss.alloc(1);
  refresh(m);
  *out = m;
  AFEM_GUARD_END
}

AFEM_API int afem_spai_destroy(afem_spai_t m)
{
  AFEM_GUARD_BEGIN
  delete m;
  AFEM_GUARD_END
}

AFEM_API int afem_spai_get_params(afem_spai_t m, float* host, int n)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(m && host, AFEM_ERR_ARGUMENT, "null argument");
  AFEM_REQUIRE(n == kNumParams, AFEM_ERR_ARGUMENT, "parameter count mismatch");
  m->params.download(host, kNumParams);
  AFEM_GUARD_END
}

AFEM_API int afem_spai_set_params(afem_spai_t m, const float* host, int n)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(m && host, AFEM_ERR_ARGUMENT, "null argument");
  AFEM_REQUIRE(n == kNumParams, AFEM_ERR_ARGUMENT, "parameter count mismatch");
  m->params.upload(host, kNumParams);
  m->adam_m.zero();
  m->adam_v.zero();
  m->adam_t = 0;
  refresh(m);
  m->forward_valid = false;
  AFEM_GUARD_END
}

AFEM_API int afem_spai_loss(afem_spai_t m, afem_system_t sys, uint64_t probe_seed, double* loss)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(m && sys && loss, AFEM_ERR_ARGUMENT, "null argument");
  require_scalar_system(sys);
  forward(m, sys);
  *loss = probe_loss(m, sys, probe_seed);
  AFEM_GUARD_END
}

AFEM_API int afem_spai_loss_grad(afem_spai_t m, afem_system_t sys, uint64_t probe_seed, double* loss, float* grad)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(m && sys && loss, AFEM_ERR_ARGUMENT, "null argument");
  require_scalar_system(sys);
  forward(m, sys);
  *loss = probe_loss(m, sys, probe_seed);
  m->grad.zero();
  backward(m, sys);
  if (grad)
    m->grad.download(grad, kNumParams);
  AFEM_GUARD_END
}

AFEM_API int afem_spai_adam_step(afem_spai_t m, double lr)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(m, AFEM_ERR_ARGUMENT, "null argument");
  AFEM_REQUIRE(lr > 0.0, AFEM_ERR_ARGUMENT, "learning rate must be positive");
  const float b1 = 0.9f, b2 = 0.999f;
  m->adam_t += 1;
  const float c1 = (float)(1.0 - std::pow((double)b1, (double)m->adam_t));
  const float c2 = (float)(1.0 - std::pow((double)b2, (double)m->adam_t));
  hipLaunchKernelGGL(k_adam, dim3((kNumParams + 255) / 256), dim3(256), 0, stream(), kNumParams, m->params.p,
                     m->grad.p, m->adam_m.p, m->adam_v.p, (float)lr, b1, b2, 1e-8f, c1, c2);
  check_launch("spai_adam");
  refresh(m);
  m->forward_valid = false;
  AFEM_GUARD_END
}

AFEM_API int afem_spai_train_step(afem_spai_t m, afem_system_t sys, uint64_t probe_seed, double lr, double* loss)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(m && sys, AFEM_ERR_ARGUMENT, "null argument");
  double l = 0.0;
  const int rc = afem_spai_loss_grad(m, sys, probe_seed, &l, nullptr);
  if (rc != AFEM_OK)
    return rc;
  if (loss)
    *loss = l;
  return afem_spai_adam_step(m, lr);
  AFEM_GUARD_END
}

AFEM_API int afem_spai_apply(afem_spai_t m, afem_system_t sys)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(m && sys, AFEM_ERR_ARGUMENT, "null argument");
  require_scalar_system(sys);
  forward(m, sys);
  const int rc = afem_system_set_spai(sys, m->mval.p);
  if (rc != AFEM_OK)
    return rc;
  AFEM_GUARD_END
}

// Intermediate arrays of the last forward pass, for tests: 0 = S values
// (nnz), 1 = node features (N x 16), 2 = h3 (N x 32), 3 = u (N x 32),
// 4 = Mh (nnz), 5 = M (nnz), 6 = scale s (N), 7 = h0 (N x 32).
AFEM_API int afem_spai_inspect(afem_spai_t m, int what, float* host, int64_t count)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(m && host, AFEM_ERR_ARGUMENT, "null argument");
  AFEM_REQUIRE(m->forward_valid, AFEM_ERR_STATE, "no forward pass since the last parameter change");
  auto bf = [&](const DBuf<uint16_t>& b, int64_t cnt) {
    AFEM_REQUIRE(cnt <= (int64_t)b.n, AFEM_ERR_ARGUMENT, "count exceeds the array");
    std::vector<uint16_t> t((size_t)cnt);
    b.download(t.data(), t.size());
    for (int64_t q = 0; q < cnt; ++q)
      host[q] = bf2f(t[q]);
  };
  auto f32 = [&](const DBuf<float>& b, int64_t cnt) {
    AFEM_REQUIRE(cnt <= (int64_t)b.n, AFEM_ERR_ARGUMENT, "count exceeds the array");
    b.download(host, (size_t)cnt);
  };
  switch (what) {
  case 0: f32(m->shat, count); break;
  case 1: bf(m->x, count); break;
  case 2: bf(m->h[kL], count); break;
  case 3: bf(m->u, count); break;
  case 4: f32(m->mhat, count); break;
  case 5: f32(m->mval, count); break;
  case 6: f32(m->s, count); break;
  case 7: bf(m->h[0], count); break;
  default: fail(AFEM_ERR_ARGUMENT, "unknown array");
  }
  AFEM_GUARD_END
}

} // extern "C"
