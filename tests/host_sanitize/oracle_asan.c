/* Host-only harness: the CPU oracle (test infrastructure) built with
 * AddressSanitizer and UndefinedBehaviorSanitizer by
 * tests/test_host_sanitize.py.  It runs the oracle's pattern, assembly (Laplace,
 * heat, elasticity in both value orders), right-hand sides, transformations,
 * toCsr, value lookup, SpMV and Jacobi-PCG on small structured meshes and
 * checks their return codes and a few identities; any out-of-bounds access
 * or undefined behaviour aborts the run. */
#include "afem_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define CHECK(c)                                                        \
  do {                                                                  \
    if (!(c)) {                                                         \
      fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c);  \
      return 1;                                                         \
    }                                                                   \
  } while (0)

/* n^3 cubes, 6 tets each around the (0,0,0)-(1,1,1) diagonal, or n^2 squares
 * with 2 triangles each; cells with stride 4 (triangles: slot 3 = -1). */
static void box(int dim, int n, double** X, int32_t* nn, int32_t** C, int32_t* nc)
{
  const int s = n + 1;
  *nn = dim == 3 ? s * s * s : s * s;
  *X = calloc((size_t)*nn * 3, sizeof(double));
  for (int32_t v = 0; v < *nn; ++v) {
    (*X)[3 * v] = (double)(v % s) / n;
    (*X)[3 * v + 1] = (double)((v / s) % s) / n;
    (*X)[3 * v + 2] = dim == 3 ? (double)(v / (s * s)) / n : 0.0;
  }
  *nc = dim == 3 ? 6 * n * n * n : 2 * n * n;
  *C = malloc((size_t)*nc * 4 * sizeof(int32_t));
  int32_t c = 0;
  for (int k = 0; k < (dim == 3 ? n : 1); ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const int32_t o = i + s * (j + s * k);
        if (dim == 2) {
          const int32_t t[2][3] = { { o, o + 1, o + s + 1 }, { o, o + s + 1, o + s } };
          for (int q = 0; q < 2; ++q, ++c) {
            for (int a = 0; a < 3; ++a)
              (*C)[4 * c + a] = t[q][a];
            (*C)[4 * c + 3] = -1;
          }
        }
        else {
          const int32_t v[8] = { o, o + 1, o + s, o + s + 1, o + s * s, o + s * s + 1, o + s * s + s,
                                 o + s * s + s + 1 };
          const int paths[6][2] = { { 1, 3 }, { 1, 5 }, { 2, 3 }, { 2, 6 }, { 4, 5 }, { 4, 6 } };
          for (int q = 0; q < 6; ++q, ++c) {
            const int a = paths[q][0], b = paths[q][1];
            (*C)[4 * c] = v[0];
            (*C)[4 * c + 1] = v[a];
            (*C)[4 * c + 2] = v[a | b];
            (*C)[4 * c + 3] = v[7];
          }
        }
      }
}

static int run(int dim, int n)
{
  double* X;
  int32_t* C;
  int32_t N, T;
  box(dim, n, &X, &N, &C, &T);
  const int npc = dim + 1;
  int32_t* row = malloc(sizeof(int32_t) * (N + 1));
  int64_t cap = (int64_t)N * 32;
  int32_t* col = malloc(sizeof(int32_t) * cap);
  const int64_t nnz = afo_pattern(N, T, npc, C, row, col, cap);
  CHECK(nnz > 0 && nnz == row[N]);
  double* val = calloc(nnz, sizeof(double));
  double* lam = malloc(sizeof(double) * T);
  for (int32_t t = 0; t < T; ++t)
    lam[t] = 1.0 + 0.25 * (t % 3);
  CHECK(afo_assemble(dim, N, X, T, npc, C, row, col, AFO_LAPLACE, 1, 1, lam, 0.0, 0.0, val) == 0);
  for (int32_t r = 0; r < N; ++r) { /* rows of the Laplace matrix sum to zero */
    double s = 0.0, m = 0.0;
    for (int32_t k = row[r]; k < row[r + 1]; ++k) {
      s += val[k];
      m = fmax(m, fabs(val[k]));
    }
    CHECK(fabs(s) <= 1e-12 * m);
  }
  double* rhs = calloc(N, sizeof(double));
  afo_rhs_constant_source(dim, X, T, npc, C, NULL, 1.0, 1, 0, rhs);
  double vol = 0.0;
  for (int32_t i = 0; i < N; ++i)
    vol += rhs[i];
  CHECK(fabs(vol - 1.0) < 1e-12);
  /* row elimination on the x = 0 face, then Jacobi-PCG */
  uint8_t* finfo = calloc(N, 1);
  double* fval = calloc(N, sizeof(double));
  uint8_t* einfo = calloc(N, 1);
  double* evalue = calloc(N, sizeof(double));
  double* x = calloc(N, sizeof(double));
  for (int32_t i = 0; i < N; ++i)
    if (X[3 * i] == 0.0)
      einfo[i] = AFO_ELIMINATE_ROW;
  CHECK(afo_apply_transformations(N, row, col, val, rhs, finfo, fval, einfo, evalue) == 0);
  double relres = 1.0;
  CHECK(afo_pcg_jacobi(N, row, col, val, rhs, x, 1e-10, 10000, &relres) >= 0 && relres <= 1e-10);
  double* y = calloc(N, sizeof(double));
  afo_spmv(N, row, col, val, x, y);
  CHECK(afo_count_differences(N, rhs, y, 1e-6, 1e-12) == 0);
  /* heat element */
  CHECK(afo_assemble(dim, N, X, T, npc, C, row, col, AFO_HEAT, 1, 1, lam, 0.1, 0.0, val) == 0);
  if (dim == 3) {
    /* elasticity, both value orders, toCsr and findValueIndex */
    double lame_l, lame_m;
    afo_lame(21e5, 0.28, &lame_l, &lame_m);
    double* v3 = calloc((size_t)nnz * 9, sizeof(double));
    for (int order = 0; order < 2; ++order)
      CHECK(afo_assemble(3, N, X, T, 4, C, row, col, AFO_ELASTICITY, 3, order, NULL, lame_l, lame_m, v3) == 0);
    int32_t* crow = malloc(sizeof(int32_t) * (3 * (size_t)N + 1));
    int32_t* ccol = malloc(sizeof(int32_t) * 9 * (size_t)nnz);
    afo_bsr_to_csr(N, 3, row, col, crow, ccol);
    CHECK(crow[3 * N] == 9 * nnz);
    int32_t* cnt = malloc(sizeof(int32_t) * N);
    for (int32_t r = 0; r < N; ++r)
      cnt[r] = row[r + 1] - row[r];
    for (int32_t r = 0; r < N; r += 7) /* DoF ids: node * 3 + component */
      for (int order = 0; order < 2; ++order)
        CHECK(afo_find_value_index(N, 3, order, row, nnz, col, cnt, 3 * r + 1, 3 * col[row[r + 1] - 1] + 2) >= 0);
    CHECK(afo_find_value_index(N, 3, 0, row, nnz, col, cnt, 0, 3 * (N - 1)) == -1);
    double* b3 = calloc(3 * (size_t)N, sizeof(double));
    const double f[3] = { 0.0, 0.0, -1.0 };
    afo_rhs_body_force(X, T, C, f, b3);
    free(v3);
    free(crow);
    free(ccol);
    free(cnt);
    free(b3);
  }
  free(X);
  free(C);
  free(row);
  free(col);
  free(val);
  free(lam);
  free(rhs);
  free(finfo);
  free(fval);
  free(einfo);
  free(evalue);
  free(x);
  free(y);
  printf("dim %d n %d: N=%d T=%d nnz=%lld ok\n", dim, n, N, T, (long long)nnz);
  return 0;
}

int main(void)
{
  if (run(2, 9) || run(3, 5))
    return 1;
  return 0;
}
