// solver.hip -- SpMV and (preconditioned) conjugate gradients on the device.
//
// The reference delegates this to Hypre / PETSc (femutils/HypreDoFLinearSystem.cc:596-769,
// femutils/PetscDoFLinearSystem.cc:427-572): PCG with a relative 2-norm test and a
// Jacobi preconditioner (femutils/PetscDoFLinearSystemFactory.axl:5-10).
//
// Layout of one PCG iteration (all scalars live on the device, so the host
// only polls every check_every iterations):
//   spmv   q = A p, fused dot p.q               -> last block: alpha = rho / p.q
//   update x += alpha p, r -= alpha q, z = M r, dots r.z, r.r
//                                                -> last block: beta, rho, convergence
//   pupd   p = z + beta p
// The mixed variant keeps r, z, p, q, A and M in FP32 and x in FP64; when the
// FP32 residual has dropped by reliable_delta (or reaches the tolerance) a
// reliable update recomputes r = b - A x with the FP64 matrix, and only that
// FP64 residual can declare convergence.  Dots are always accumulated in FP64.
//
// Reductions: each block writes its partial sums, the last block to finish
// (atomic ticket) adds the partials in block order -- deterministic -- and
// updates the scalars.
#include "common.h"
#include "device_utils.h"
#include "linsys.h"

#include <chrono>
#include <cmath>

namespace afem {

struct Scalars
{
  double rho, alpha, beta, rr, bb, tol2, rr_ref, delta2, rtol2, atol2;
  int iter, max_iter, done, converged, ru, nb_ru;
};

struct RedCtx
{
  double* partials;
  unsigned* counter;
  Scalars* sc;
};

constexpr int kBlock = 256;

// Reduction of NV per-thread values over the grid; the last block calls fin.
template <int NV, class Fin>
__device__ __forceinline__ void grid_reduce(double (&v)[NV], const RedCtx& ctx, Fin fin)
{
  __shared__ double red[NV][kBlock / kWave];
  __shared__ int is_last;
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = wave_reduce_sum(v[k]);
    if (lane == 0)
      red[k][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int w = 0; w < kBlock / kWave; ++w)
        s += red[k][w];
      __hip_atomic_store(&ctx.partials[(int64_t)blockIdx.x * NV + k], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __threadfence();
    const unsigned prev = atomicAdd(ctx.counter, 1u);
    is_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last)
    return;
  __threadfence();
  double t[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k)
    t[k] = 0.0;
  for (unsigned b = threadIdx.x; b < gridDim.x; b += kBlock)
#pragma unroll
    for (int k = 0; k < NV; ++k)
      t[k] += __hip_atomic_load(&ctx.partials[(int64_t)b * NV + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = wave_reduce_sum(t[k]);
    if (lane == 0)
      red[k][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      s[k] = 0.0;
      for (int w = 0; w < kBlock / kWave; ++w)
        s[k] += red[k][w];
    }
    fin(ctx.sc, s);
    *ctx.counter = 0u;
    __threadfence();
  }
}

// Gates: which iterations a kernel runs in.
enum Gate
{
  kAlways = 0,
  kUntilDone = 1,
  kReliableUpdate = 2
};

__device__ __forceinline__ bool gate_closed(const Scalars* sc, int gate)
{
  if (gate == kAlways)
    return false;
  if (sc->done)
    return true;
  return gate == kReliableUpdate && !sc->ru;
}

// ---- finalizers (run by one thread of the last block) ---------------------------

struct FinAlpha
{
  __device__ void operator()(Scalars* sc, const double* s) const
  {
    const double pq = s[0];
    if (pq > 0.0 && isfinite(pq)) {
      sc->alpha = sc->rho / pq;
    }
    else { // breakdown: the operator is not positive definite on p
      sc->alpha = 0.0;
      sc->done = 1;
    }
  }
};

struct FinTrueResidual
{
  __device__ void operator()(Scalars* sc, const double* s) const { sc->rr = s[0]; }
};

// After a reliable update: the FP64 residual is the judge.
struct FinReliable
{
  __device__ void operator()(Scalars* sc, const double* s) const
  {
    sc->rr = s[0];
    sc->rr_ref = s[0];
    sc->nb_ru += 1;
    if (s[0] <= sc->tol2) {
      sc->done = 1;
      sc->converged = 1;
    }
    else if (sc->iter >= sc->max_iter) {
      sc->done = 1;
    }
  }
};

struct FinInit
{
  __device__ void operator()(Scalars* sc, const double* s) const
  {
    sc->rho = s[0];
    sc->rr = s[1];
    sc->rr_ref = s[1];
    sc->bb = s[2];
    sc->tol2 = fmax(sc->rtol2 * s[2], sc->atol2);
    if (s[1] <= sc->tol2 || s[0] == 0.0) {
      sc->done = 1;
      sc->converged = 1;
    }
  }
};

// update step of the FP64 solver: the recursive residual decides.
struct FinUpdate
{
  __device__ void operator()(Scalars* sc, const double* s) const
  {
    const double rho_new = s[0];
    sc->beta = (sc->rho != 0.0) ? rho_new / sc->rho : 0.0;
    sc->rho = rho_new;
    sc->rr = s[1];
    sc->iter += 1;
    if (s[1] <= sc->tol2) {
      sc->done = 1;
      sc->converged = 1;
    }
    else if (sc->iter >= sc->max_iter || !isfinite(s[1])) {
      sc->done = 1;
    }
  }
};

// update step of the mixed solver: the FP32 residual only requests a
// reliable update.
struct FinUpdateMixed
{
  __device__ void operator()(Scalars* sc, const double* s) const
  {
    const double rho_new = s[0];
    sc->beta = (sc->rho != 0.0) ? rho_new / sc->rho : 0.0;
    sc->rho = rho_new;
    sc->rr = s[1];
    sc->iter += 1;
    if (s[1] <= sc->delta2 * sc->rr_ref || s[1] <= sc->tol2 || sc->iter >= sc->max_iter || !isfinite(s[1]))
      sc->ru = 1;
  }
};

struct FinRho
{
  __device__ void operator()(Scalars* sc, const double* s) const
  {
    sc->rho = s[0];
    sc->ru = 0;
  }
};

// Contiguous chunk of work units for this block (XCD-aware).
__device__ __forceinline__ void block_chunk(int64_t nunits, int64_t chunk, int64_t& u0, int64_t& u1)
{
  const int64_t L = xcd_tile(blockIdx.x, gridDim.x);
  u0 = L * chunk;
  u1 = u0 + chunk;
  if (u1 > nunits)
    u1 = nunits;
}

// ---- SpMV ------------------------------------------------------------------------------

enum SpmvMode
{
  kPlain = 0,
  kPQ = 1,      // y = A x and dot(x, y) -> alpha
  kResidual = 2 // y = b - A x and dot(y, y); finalizer chosen by template
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

// G lanes per row, 256/G rows per tile; each block owns a contiguous chunk of
// tiles.  Lanes of a group read consecutive entries of their row, so a
// wavefront streams 64/G consecutive rows with coalesced value/column loads.
template <int G, class VT, class XT, int MODE, class Fin>
__global__ void __launch_bounds__(kBlock)
k_spmv_csr(int32_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const VT* __restrict__ v,
           const XT* __restrict__ x, XT* __restrict__ y, const double* __restrict__ b, int64_t chunk, int gate,
           RedCtx ctx, Fin fin)
{
  if (MODE != kPlain && gate_closed(ctx.sc, gate))
    return;
  using A = typename Acc<VT>::type;
  constexpr int RPT = kBlock / G;
  const int64_t ntiles = ((int64_t)n + RPT - 1) / RPT;
  int64_t t0, t1;
  block_chunk(ntiles, chunk, t0, t1);
  const int lane = threadIdx.x & (G - 1);
  const int sub = threadIdx.x / G;
  double acc = 0.0;
  for (int64_t t = t0; t < t1; ++t) {
    const int64_t row = t * RPT + sub;
    if (row < n) {
      const int32_t s = rp[row], e = rp[row + 1];
      A sum = 0;
      for (int32_t k = s + lane; k < e; k += G)
        sum += (A)v[k] * (A)x[ci[k]];
      sum = group_reduce_sum<G>(sum);
      if (lane == 0) {
        if constexpr (MODE == kResidual) {
          const double r = b[row] - (double)sum;
          y[row] = (XT)r;
          acc += r * r;
        }
        else {
          y[row] = (XT)sum;
          if constexpr (MODE == kPQ)
            acc += (double)sum * (double)x[row];
        }
      }
    }
  }
  if constexpr (MODE != kPlain) {
    double vv[1] = { acc };
    grid_reduce<1>(vv, ctx, fin);
  }
}

// BSR with 3x3 blocks stored per block (9 contiguous values, row-major).
template <int G, class VT, class XT, int MODE, class Fin>
__global__ void __launch_bounds__(kBlock)
k_spmv_bsr3(int32_t nb_node, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const VT* __restrict__ v,
            const XT* __restrict__ x, XT* __restrict__ y, const double* __restrict__ b, int64_t chunk, int gate,
            RedCtx ctx, Fin fin)
{
  if (MODE != kPlain && gate_closed(ctx.sc, gate))
    return;
  using A = typename Acc<VT>::type;
  constexpr int RPT = kBlock / G;
  const int64_t ntiles = ((int64_t)nb_node + RPT - 1) / RPT;
  int64_t t0, t1;
  block_chunk(ntiles, chunk, t0, t1);
  const int lane = threadIdx.x & (G - 1);
  const int sub = threadIdx.x / G;
  double acc = 0.0;
  for (int64_t t = t0; t < t1; ++t) {
    const int64_t a = t * RPT + sub;
    if (a < nb_node) {
      const int32_t s = rp[a], e = rp[a + 1];
      A s0 = 0, s1 = 0, s2 = 0;
      for (int32_t k = s + lane; k < e; k += G) {
        const VT* blk = v + 9 * (int64_t)k;
        const int64_t c = 3 * (int64_t)ci[k];
        const A x0 = (A)x[c], x1 = (A)x[c + 1], x2 = (A)x[c + 2];
        s0 += (A)blk[0] * x0 + (A)blk[1] * x1 + (A)blk[2] * x2;
        s1 += (A)blk[3] * x0 + (A)blk[4] * x1 + (A)blk[5] * x2;
        s2 += (A)blk[6] * x0 + (A)blk[7] * x1 + (A)blk[8] * x2;
      }
      s0 = group_reduce_sum<G>(s0);
      s1 = group_reduce_sum<G>(s1);
      s2 = group_reduce_sum<G>(s2);
      if (lane == 0) {
        const A sv[3] = { s0, s1, s2 };
        for (int i = 0; i < 3; ++i) {
          const int64_t row = 3 * a + i;
          if constexpr (MODE == kResidual) {
            const double r = b[row] - (double)sv[i];
            y[row] = (XT)r;
            acc += r * r;
          }
          else {
            y[row] = (XT)sv[i];
            if constexpr (MODE == kPQ)
              acc += (double)sv[i] * (double)x[row];
          }
        }
      }
    }
  }
  if constexpr (MODE != kPlain) {
    double vv[1] = { acc };
    grid_reduce<1>(vv, ctx, fin);
  }
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
    d32[9 * (int64_t)a + q] = (float)inv[q];
  }
}

__global__ void k_to_float(int64_t n, const double* __restrict__ in, float* __restrict__ out)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (float)in[i];
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

// z = M r for the unit u (a row when NB = 1, a node when NB = 3).
template <int NB, class T>
__device__ __forceinline__ void apply_jacobi(int64_t u, const T* __restrict__ dinv, const T* rv, T* zv)
{
  if constexpr (NB == 1) {
    zv[0] = dinv[u] * rv[0];
  }
  else {
    const T* d = dinv + 9 * u;
    zv[0] = d[0] * rv[0] + d[1] * rv[1] + d[2] * rv[2];
    zv[1] = d[3] * rv[0] + d[4] * rv[1] + d[5] * rv[2];
    zv[2] = d[6] * rv[0] + d[7] * rv[1] + d[8] * rv[2];
  }
}

// x += alpha p, r -= alpha q, z = M r; partial dots (r.z, r.r).
template <int NB, class T, class Fin>
__global__ void __launch_bounds__(kBlock)
k_update(int64_t nunits, int64_t chunk, double* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
         const T* __restrict__ p, const T* __restrict__ q, const T* __restrict__ dinv, int precond, RedCtx ctx, Fin fin)
{
  if (gate_closed(ctx.sc, kUntilDone))
    return;
  const double alpha = ctx.sc->alpha;
  const T alpha_t = (T)alpha;
  int64_t u0, u1;
  block_chunk((nunits + kBlock - 1) / kBlock, chunk, u0, u1);
  double rz = 0.0, rr = 0.0;
  for (int64_t t = u0; t < u1; ++t) {
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
    if (precond)
      apply_jacobi<NB, T>(u, dinv, rv, zv);
    else
#pragma unroll
      for (int i = 0; i < NB; ++i)
        zv[i] = rv[i];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      z[u * NB + i] = zv[i];
      rz += (double)rv[i] * (double)zv[i];
      rr += (double)rv[i] * (double)rv[i];
    }
  }
  double vv[2] = { rz, rr };
  grid_reduce<2>(vv, ctx, fin);
}

// p = z + beta p.  After a reliable update (mixed solver) r and z are first
// replaced by the FP64 residual and its preconditioned value, and rho = r.z is
// recomputed.
template <int NB, class T>
__global__ void __launch_bounds__(kBlock)
k_pupdate(int64_t nunits, int64_t chunk, T* __restrict__ p, T* __restrict__ z, T* __restrict__ r,
          const double* __restrict__ r64, const T* __restrict__ dinv, int precond, RedCtx ctx)
{
  if (gate_closed(ctx.sc, kUntilDone))
    return;
  const T beta = (T)ctx.sc->beta;
  const bool ru = ctx.sc->ru != 0;
  int64_t u0, u1;
  block_chunk((nunits + kBlock - 1) / kBlock, chunk, u0, u1);
  double rz = 0.0;
  for (int64_t t = u0; t < u1; ++t) {
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
      if (precond)
        apply_jacobi<NB, T>(u, dinv, rv, zv);
      else
#pragma unroll
        for (int i = 0; i < NB; ++i)
          zv[i] = rv[i];
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        z[u * NB + i] = zv[i];
        rz += (double)rv[i] * (double)zv[i];
      }
    }
    else {
#pragma unroll
      for (int i = 0; i < NB; ++i)
        zv[i] = z[u * NB + i];
    }
#pragma unroll
    for (int i = 0; i < NB; ++i)
      p[u * NB + i] = zv[i] + beta * p[u * NB + i];
  }
  if (ru) {
    double vv[1] = { rz };
    grid_reduce<1>(vv, ctx, FinRho{});
  }
}

// Initial z = M r, p = z from the residual r64 = b - A x0 (already computed);
// dots (r.z, r.r, b.b).  r may alias r64 (FP64 solver).
template <int NB, class T>
__global__ void __launch_bounds__(kBlock)
k_init(int64_t nunits, int64_t chunk, const double* r64, const double* __restrict__ b, T* r, T* __restrict__ z,
       T* __restrict__ p, const T* __restrict__ dinv, int precond, RedCtx ctx)
{
  int64_t u0, u1;
  block_chunk((nunits + kBlock - 1) / kBlock, chunk, u0, u1);
  double rz = 0.0, rr = 0.0, bb = 0.0;
  for (int64_t t = u0; t < u1; ++t) {
    const int64_t u = t * kBlock + threadIdx.x;
    if (u >= nunits)
      continue;
    T rv[NB], zv[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const double rd = r64[u * NB + i];
      rv[i] = (T)rd;
      r[u * NB + i] = rv[i];
      rr += rd * rd;
      const double bv = b[u * NB + i];
      bb += bv * bv;
    }
    if (precond)
      apply_jacobi<NB, T>(u, dinv, rv, zv);
    else
#pragma unroll
      for (int i = 0; i < NB; ++i)
        zv[i] = rv[i];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      z[u * NB + i] = zv[i];
      p[u * NB + i] = zv[i];
      rz += (double)rv[i] * (double)zv[i];
    }
  }
  double vv[3] = { rz, rr, bb };
  grid_reduce<3>(vv, ctx, FinInit{});
}

struct NoFin
{
  __device__ void operator()(Scalars*, const double*) const {}
};

} // namespace afem

// ---- workspace --------------------------------------------------------------------------------

struct afem_solver_ws
{
  int32_t n = 0;
  int64_t nnz = 0;
  int block = 1;
  afem::DBuf<double> r64, z64, p64, q64, dinv64;
  afem::DBuf<float> r32, z32, p32, q32, dinv32, vals32;
  afem::DBuf<double> partials;
  afem::DBuf<unsigned> counter;
  afem::DBuf<afem::Scalars> sc;
  afem::Scalars* host_sc = nullptr; // pinned
  bool has_fp32 = false;
  bool derived_valid = false;
  ~afem_solver_ws()
  {
    if (host_sc)
      (void)hipHostFree(host_sc);
  }
};

namespace afem {

void system_release_workspace(afem_system* s)
{
  delete s->ws;
  s->ws = nullptr;
}

static int grid_blocks()
{
  ensure_init();
  int cu = rt().num_cu > 0 ? rt().num_cu : 256;
  return cu * 8;
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
    w->has_fp32 = false;
    w->derived_valid = false;
  }
  if (!w->partials.p) {
    w->partials.alloc((size_t)grid_blocks() * 4);
    w->counter.alloc(1);
    w->counter.zero();
    w->sc.alloc(1);
    AFEM_HIP(hipHostMalloc(reinterpret_cast<void**>(&w->host_sc), sizeof(Scalars), hipHostMallocDefault));
  }
  if (s->matrix_changed) {
    w->derived_valid = false;
    s->matrix_changed = false;
  }
  return w;
}

// Diagonal (or block-diagonal) inverse and the FP32 copy of the matrix.
static void ensure_derived(afem_system* s, afem_solver_ws* w, bool need_fp32)
{
  const int32_t n = s->nb_dof;
  if (need_fp32 && !w->has_fp32) {
    w->r32.alloc(n);
    w->z32.alloc(n);
    w->p32.alloc(n);
    w->q32.alloc(n);
    w->has_fp32 = true;
    w->derived_valid = false;
  }
  if (w->derived_valid && w->nnz == s->nnz)
    return;
  w->dinv32.reserve((size_t)n * s->block);
  if (w->has_fp32)
    w->vals32.reserve((size_t)s->nnz);
  if (s->block == 3) {
    const int32_t nb_node = s->nb_dof / 3;
    hipLaunchKernelGGL(k_block_inverse3, dim3(grid_for(nb_node, 256)), dim3(256), 0, stream(), nb_node, s->row_index,
                       s->columns, s->values, w->dinv64.p, w->dinv32.p);
  }
  else {
    hipLaunchKernelGGL(k_diag_inverse, dim3(grid_for(s->nb_dof, 256)), dim3(256), 0, stream(), s->nb_dof,
                       s->row_index, s->columns, s->values, w->dinv64.p, w->dinv32.p);
  }
  check_launch("jacobi_setup");
  if (w->has_fp32) {
    hipLaunchKernelGGL(k_to_float, dim3(2048), dim3(256), 0, stream(), (int64_t)s->nnz, s->values, w->vals32.p);
    check_launch("to_float");
  }
  w->nnz = s->nnz;
  w->derived_valid = true;
}

// ---- SpMV dispatch ---------------------------------------------------------------------------

template <class VT, class XT, int MODE, class Fin>
static void spmv_launch(const afem_system* s, const VT* vals, const XT* x, XT* y, const double* b, int gate,
                        const RedCtx& ctx, Fin fin, int lanes, const char* prof_name)
{
  const int nblk = grid_blocks();
  ProfScope prof(prof_name);
  if (s->block == 3) {
    const int32_t nb_node = s->nb_dof / 3;
    constexpr int G = 16;
    const int64_t ntiles = ((int64_t)nb_node + kBlock / G - 1) / (kBlock / G);
    const int64_t grid = ntiles < nblk ? ntiles : nblk;
    const int64_t chunk = (ntiles + grid - 1) / grid;
    hipLaunchKernelGGL((k_spmv_bsr3<G, VT, XT, MODE, Fin>), dim3((unsigned)grid), dim3(kBlock), 0, stream(), nb_node,
                       s->row_index, s->columns, vals, x, y, b, chunk, gate, ctx, fin);
  }
  else {
    auto go = [&](auto gtag) {
      constexpr int G = decltype(gtag)::value;
      const int64_t ntiles = ((int64_t)s->nb_dof + kBlock / G - 1) / (kBlock / G);
      const int64_t grid = ntiles < nblk ? ntiles : nblk;
      const int64_t chunk = (ntiles + grid - 1) / grid;
      hipLaunchKernelGGL((k_spmv_csr<G, VT, XT, MODE, Fin>), dim3((unsigned)grid), dim3(kBlock), 0, stream(),
                         s->nb_dof, s->row_index, s->columns, vals, x, y, b, chunk, gate, ctx, fin);
    };
    switch (lanes) {
    case 4: go(std::integral_constant<int, 4>{}); break;
    case 8: go(std::integral_constant<int, 8>{}); break;
    case 32: go(std::integral_constant<int, 32>{}); break;
    case 64: go(std::integral_constant<int, 64>{}); break;
    default: go(std::integral_constant<int, 16>{}); break;
    }
  }
  check_launch(prof_name);
}

// Lanes per row from the mean row length (power of two, 4..64).
static int auto_lanes(const afem_system* s)
{
  const double mean = (double)s->nnz / (double)(s->nb_dof > 0 ? s->nb_dof : 1);
  int g = 4;
  while (g < 64 && g < mean)
    g <<= 1;
  return g;
}

// ---- solve ----------------------------------------------------------------------------------

template <int NB>
static void launch_init(afem_system* s, afem_solver_ws* w, bool mixed, int precond, const RedCtx& ctx)
{
  const int64_t nunits = s->nb_dof / NB;
  const int64_t tiles = (nunits + kBlock - 1) / kBlock;
  const int64_t grid = tiles < grid_blocks() ? tiles : grid_blocks();
  const int64_t chunk = (tiles + grid - 1) / grid;
  if (mixed)
    hipLaunchKernelGGL((k_init<NB, float>), dim3((unsigned)grid), dim3(kBlock), 0, stream(), nunits, chunk, w->r64.p,
                       s->rhs.p, w->r32.p, w->z32.p, w->p32.p, w->dinv32.p, precond, ctx);
  else
    hipLaunchKernelGGL((k_init<NB, double>), dim3((unsigned)grid), dim3(kBlock), 0, stream(), nunits, chunk, w->r64.p,
                       s->rhs.p, w->r64.p, w->z64.p, w->p64.p, w->dinv64.p, precond, ctx);
  check_launch("pcg_init");
}

template <int NB>
static void launch_iteration(afem_system* s, afem_solver_ws* w, bool mixed, int precond, int lanes, const RedCtx& ctx)
{
  const int64_t nunits = s->nb_dof / NB;
  const int64_t tiles = (nunits + kBlock - 1) / kBlock;
  const int64_t grid = tiles < grid_blocks() ? tiles : grid_blocks();
  const int64_t chunk = (tiles + grid - 1) / grid;
  if (!mixed) {
    spmv_launch<double, double, kPQ>(s, s->values, w->p64.p, w->q64.p, nullptr, kUntilDone, ctx, FinAlpha{}, lanes,
                                     "spmv_fp64");
    {
      ProfScope prof("pcg_update_fp64");
      hipLaunchKernelGGL((k_update<NB, double, FinUpdate>), dim3((unsigned)grid), dim3(kBlock), 0, stream(), nunits,
                         chunk, s->x.p, w->r64.p, w->z64.p, w->p64.p, w->q64.p, w->dinv64.p, precond, ctx, FinUpdate{});
      check_launch("pcg_update");
    }
    {
      ProfScope prof("pcg_pupdate_fp64");
      hipLaunchKernelGGL((k_pupdate<NB, double>), dim3((unsigned)grid), dim3(kBlock), 0, stream(), nunits, chunk,
                         w->p64.p, w->z64.p, w->r64.p, (const double*)nullptr, w->dinv64.p, precond, ctx);
      check_launch("pcg_pupdate");
    }
  }
  else {
    spmv_launch<float, float, kPQ>(s, w->vals32.p, w->p32.p, w->q32.p, nullptr, kUntilDone, ctx, FinAlpha{}, lanes,
                                   "spmv_fp32");
    {
      ProfScope prof("pcg_update_fp32");
      hipLaunchKernelGGL((k_update<NB, float, FinUpdateMixed>), dim3((unsigned)grid), dim3(kBlock), 0, stream(),
                         nunits, chunk, s->x.p, w->r32.p, w->z32.p, w->p32.p, w->q32.p, w->dinv32.p, precond, ctx,
                         FinUpdateMixed{});
      check_launch("pcg_update_mixed");
    }
    spmv_launch<double, double, kResidual>(s, s->values, s->x.p, w->r64.p, s->rhs.p, kReliableUpdate, ctx,
                                           FinReliable{}, lanes, "spmv_residual_fp64");
    {
      ProfScope prof("pcg_pupdate_fp32");
      hipLaunchKernelGGL((k_pupdate<NB, float>), dim3((unsigned)grid), dim3(kBlock), 0, stream(), nunits, chunk,
                         w->p32.p, w->z32.p, w->r32.p, (const double*)w->r64.p, w->dinv32.p, precond, ctx);
      check_launch("pcg_pupdate_mixed");
    }
  }
}

static void read_scalars(afem_solver_ws* w)
{
  AFEM_HIP(hipMemcpyAsync(w->host_sc, w->sc.p, sizeof(Scalars), hipMemcpyDeviceToHost, stream()));
  AFEM_HIP(hipStreamSynchronize(stream()));
}

static double true_relres(afem_system* s, afem_solver_ws* w, const RedCtx& ctx, int lanes)
{
  spmv_launch<double, double, kResidual>(s, s->values, s->x.p, w->r64.p, s->rhs.p, kAlways, ctx, FinTrueResidual{},
                                         lanes, "spmv_residual_fp64");
  read_scalars(w);
  const double bb = w->host_sc->bb;
  const double rr = w->host_sc->rr;
  return bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
}

static int solve(afem_system* s, const afem_solver_options& o, afem_solve_result* res)
{
  AFEM_REQUIRE(o.method == AFEM_SOLVER_PCG_FP64 || o.method == AFEM_SOLVER_PCG_MIXED || o.method == AFEM_SOLVER_CG_FP64,
               AFEM_ERR_ARGUMENT, "unknown solver method");
  AFEM_REQUIRE(o.preconditioner == AFEM_PRECOND_JACOBI || o.preconditioner == AFEM_PRECOND_NONE, AFEM_ERR_ARGUMENT,
               "unknown preconditioner");
  AFEM_REQUIRE(o.rtol >= 0.0 && o.atol >= 0.0 && o.max_iter >= 0, AFEM_ERR_ARGUMENT, "bad tolerances");
  afem_solver_ws* w = workspace(s);
  const bool mixed = o.method == AFEM_SOLVER_PCG_MIXED;
  const int precond = (o.method == AFEM_SOLVER_CG_FP64 || o.preconditioner == AFEM_PRECOND_NONE) ? 0 : 1;
  const int lanes = o.spmv_lanes > 0 ? o.spmv_lanes : auto_lanes(s);
  const int check_every = o.check_every > 0 ? o.check_every : 16;
  const double delta = o.reliable_delta > 0.0 ? o.reliable_delta : 0.1;
  ensure_derived(s, w, mixed);
  RedCtx ctx{ w->partials.p, w->counter.p, w->sc.p };

  auto t_start = std::chrono::steady_clock::now();
  if (!o.use_initial_guess) {
    hipLaunchKernelGGL(k_initial_guess, dim3(grid_for(s->nb_dof, 256)), dim3(256), 0, stream(), s->nb_dof,
                       s->elim_info.p, s->elim_value.p, s->x.p);
    check_launch("initial_guess");
  }
  Scalars init{};
  init.rtol2 = o.rtol * o.rtol;
  init.atol2 = o.atol * o.atol;
  init.delta2 = delta * delta;
  init.max_iter = o.max_iter;
  int total_iter = 0;
  int restarts = 0;
  int nb_ru = 0;
  bool converged = false;
  double relres = 0.0;
  for (;;) {
    init.iter = 0;
    init.max_iter = o.max_iter - total_iter;
    *w->host_sc = init;
    AFEM_HIP(hipMemcpyAsync(w->sc.p, w->host_sc, sizeof(Scalars), hipMemcpyHostToDevice, stream()));
    // r64 = b - A x0
    spmv_launch<double, double, kResidual>(s, s->values, s->x.p, w->r64.p, s->rhs.p, kAlways, ctx, FinTrueResidual{},
                                           lanes, "spmv_residual_fp64");
    if (s->block == 3)
      launch_init<3>(s, w, mixed, precond, ctx);
    else
      launch_init<1>(s, w, mixed, precond, ctx);
    read_scalars(w);
    while (!w->host_sc->done && w->host_sc->iter < init.max_iter) {
      for (int k = 0; k < check_every; ++k) {
        if (s->block == 3)
          launch_iteration<3>(s, w, mixed, precond, lanes, ctx);
        else
          launch_iteration<1>(s, w, mixed, precond, lanes, ctx);
      }
      read_scalars(w);
    }
    total_iter += w->host_sc->iter;
    nb_ru += w->host_sc->nb_ru;
    relres = true_relres(s, w, ctx, lanes);
    const double bnorm = std::sqrt(w->host_sc->bb);
    converged = relres <= o.rtol || relres * bnorm <= o.atol;
    if (converged || total_iter >= o.max_iter || restarts >= 8 || !std::isfinite(relres))
      break;
    // the recursive residual drifted below the tolerance while the true one
    // did not: restart from the current iterate
    ++restarts;
  }
  AFEM_HIP(hipStreamSynchronize(stream()));
  auto t_end = std::chrono::steady_clock::now();
  if (res) {
    res->iterations = total_iter;
    res->converged = converged ? 1 : 0;
    res->reliable_updates = nb_ru;
    res->relres = relres;
    res->solve_ms = std::chrono::duration<double, std::milli>(t_end - t_start).count();
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

AFEM_API int afem_system_spmv(afem_system_t sys, const double* x, double* y)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(sys && x && y, AFEM_ERR_ARGUMENT, "null argument");
  afem_solver_ws* w = workspace(sys);
  RedCtx ctx{ w->partials.p, w->counter.p, w->sc.p };
  spmv_launch<double, double, kPlain>(sys, sys->values, x, y, nullptr, kAlways, ctx, NoFin{}, auto_lanes(sys),
                                      "spmv_fp64");
  AFEM_GUARD_END
}

AFEM_API int afem_system_spmv_fp32(afem_system_t sys, const float* x, float* y)
{
  AFEM_GUARD_BEGIN
  AFEM_REQUIRE(sys && x && y, AFEM_ERR_ARGUMENT, "null argument");
  afem_solver_ws* w = workspace(sys);
  ensure_derived(sys, w, true);
  RedCtx ctx{ w->partials.p, w->counter.p, w->sc.p };
  spmv_launch<float, float, kPlain>(sys, w->vals32.p, x, y, nullptr, kAlways, ctx, NoFin{}, auto_lanes(sys),
                                    "spmv_fp32");
  AFEM_GUARD_END
}

} // extern "C"
