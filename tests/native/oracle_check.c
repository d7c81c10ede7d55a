/* oracle_check.c -- every entry of the CPU oracle (oracle/sg_oracle.c) on
 * small inputs, built with the address and undefined-behaviour sanitizers.
 * Besides running clean, results are checked against what the mathematics
 * fixes: partition of unity of the basis, the constant spline, the
 * transpose identity <M x, y> = <x, M^T y> of the grid map and of the
 * refinement products, the Flag scatter against its definition, and the
 * overwrite / gather pair.  Prints "oracle_check ok ..." and exits 0, or
 * names the first failure. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sg_oracle.h"

static int g_checks = 0;

static void require(int ok, const char* what) {
  ++g_checks;
  if (!ok) {
    fprintf(stderr, "oracle_check FAILED: %s\n", what);
    exit(1);
  }
}

static unsigned long long g_state = 0x5EEDULL;
static double urand(void) { /* SplitMix64 -> [0, 1) */
  unsigned long long z = (g_state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

typedef struct {
  int n, p, m, D;
  double* knots;
  double* t;
  int32_t* idx;
  double* tab; /* [m, p+1, D+1] */
  float* tab32;
} Dim;

/* clamped knots (some interior values doubled), samples: uniform over the
 * extent plus every knot and two points outside it */
static void make_dim(Dim* d, int n, int p, int m, int D) {
  d->n = n, d->p = p, d->m = m, d->D = D;
  const int k = n - p + 1;
  double* vals = malloc(sizeof(double) * k);
  int32_t* mults = malloc(sizeof(int32_t) * k);
  for (int i = 0; i < k; ++i) vals[i] = (double)i / (k - 1), mults[i] = 1;
  mults[0] = mults[k - 1] = p + 1;
  if (p >= 1 && k > 4) { /* one interior knot of multiplicity 2 */
    mults[1] = 2;
    mults[k - 2] = 0;
  }
  int kk = 0, pos = 0;
  for (int i = 0; i < k; ++i)
    if (mults[i] > 0) vals[kk] = vals[i], mults[kk++] = mults[i];
  d->knots = malloc(sizeof(double) * (n + p + 1));
  pos = sgo_expand_knots(vals, mults, kk, d->knots);
  require(pos == n + p + 1, "expand_knots length");
  d->t = malloc(sizeof(double) * m);
  for (int i = 0; i < m; ++i) d->t[i] = m > 1 ? (double)i / (m - 1) : 0.0;
  for (int i = 0; i < m && i < n + p + 1; i += 3) d->t[i] = d->knots[i];
  if (m > 4) d->t[1] = -0.25, d->t[m - 2] = 1.25;
  d->idx = malloc(sizeof(int32_t) * m);
  sgo_sample_indices(d->t, m, d->knots, n + p + 1, p, d->idx);
  for (int j = 0; j < m; ++j) require(d->idx[j] >= p + 1 && d->idx[j] <= n, "span in [p+1, n]");
  const size_t cnt = (size_t)m * (p + 1) * (D + 1);
  d->tab = malloc(sizeof(double) * cnt);
  sgo_basis(d->knots, d->t, d->idx, m, p, D, d->tab);
  d->tab32 = malloc(sizeof(float) * cnt);
  for (size_t i = 0; i < cnt; ++i) d->tab32[i] = (float)d->tab[i];
  for (int j = 0; j < m; ++j) {
    double s = 0.0;
    for (int a = 0; a <= p; ++a) s += d->tab[j + (size_t)m * a];
    /* inside the extent the basis sums to one */
    if (d->t[j] >= 0.0 && d->t[j] <= 1.0) require(fabs(s - 1.0) < 1e-12, "partition of unity");
  }
  free(vals);
  free(mults);
}

static void free_dim(Dim* d) {
  free(d->knots), free(d->t), free(d->idx), free(d->tab), free(d->tab32);
}

static void grid_checks(int nin, const int* n, const int* p, const int* m, int nout, int nurbs) {
  Dim dims[3];
  int32_t nn[3], mm[3], pp[3], DD[3], deriv[3] = {0, 0, 0};
  const double* tabs[3];
  const float* tabs32[3];
  const int32_t* idx[3];
  long long ncp = 1, ns = 1;
  for (int d = 0; d < nin; ++d) {
    make_dim(&dims[d], n[d], p[d], m[d], p[d] > 0 ? 1 : 0);
    nn[d] = n[d], mm[d] = m[d], pp[d] = p[d], DD[d] = dims[d].D;
    tabs[d] = dims[d].tab, tabs32[d] = dims[d].tab32, idx[d] = dims[d].idx;
    ncp *= n[d], ns *= m[d];
  }
  double* P = malloc(sizeof(double) * ncp * nout);
  double* W = nurbs ? malloc(sizeof(double) * ncp) : NULL;
  double* E = malloc(sizeof(double) * ns * nout);
  double* Y = malloc(sizeof(double) * ns * nout);
  double* Pb = malloc(sizeof(double) * ncp * nout);
  float* P32 = malloc(sizeof(float) * ncp * nout);
  float* E32 = malloc(sizeof(float) * ns * nout);
  float* Pb32 = malloc(sizeof(float) * ncp * nout);
  for (long long i = 0; i < ncp * nout; ++i) P[i] = 2.0 * urand() - 1.0, P32[i] = (float)P[i];
  if (W)
    for (long long i = 0; i < ncp; ++i) W[i] = 0.5 + 1.5 * urand();
  /* the constant spline: all control points 1 gives 1 where samples are
   * inside the extent */
  for (long long i = 0; i < ncp * nout; ++i) Pb[i] = 1.0;
  sgo_eval_direct_f64(nin, nout, nn, mm, pp, DD, tabs, idx, deriv, Pb, W, E, 0, mm[nin - 1], 1);
  long long inside = 0;
  for (long long q = 0; q < ns; ++q) {
    long long r = q;
    int in = 1;
    for (int d = 0; d < nin; ++d) {
      const double t = dims[d].t[r % m[d]];
      r /= m[d];
      in = in && t >= 0.0 && t <= 1.0;
    }
    if (in) require(fabs(E[q] - 1.0) < 1e-12, "constant spline"), ++inside;
  }
  require(inside > 0, "some samples inside");
  /* forward, and both threadings agree */
  sgo_eval_direct_f64(nin, nout, nn, mm, pp, DD, tabs, idx, deriv, P, W, E, 0, mm[nin - 1], 1);
  sgo_eval_direct_f64(nin, nout, nn, mm, pp, DD, tabs, idx, deriv, P, W, Y, 0, mm[nin - 1], 4);
  require(memcmp(E, Y, sizeof(double) * ns * nout) == 0, "threads agree");
  /* a slab of the slowest dimension */
  if (mm[nin - 1] > 2) {
    const long long slab = ns / mm[nin - 1];
    double* S = malloc(sizeof(double) * slab * nout);
    sgo_eval_direct_f64(nin, nout, nn, mm, pp, DD, tabs, idx, deriv, P, W, S, 1, 2, 1);
    for (int c = 0; c < nout; ++c)
      for (long long i = 0; i < slab; ++i)
        require(S[i + slab * c] == E[slab + i + ns * c], "slab equals the whole grid");
    free(S);
  }
  /* points */
  {
    const int np = 17;
    int64_t lin[17];
    double out[17 * 8];
    for (int i = 0; i < np; ++i) lin[i] = (int64_t)(urand() * ns) % ns;
    sgo_eval_points_f64(nin, nout, nn, mm, pp, DD, tabs, idx, deriv, P, W, lin, np, out, 1);
    for (int i = 0; i < np; ++i)
      for (int c = 0; c < nout; ++c)
        require(out[i + np * c] == E[lin[i] + ns * c], "points equal the grid");
  }
  if (!nurbs) {
    /* transpose identity, FP64 and FP32 */
    for (long long i = 0; i < ns * nout; ++i) Y[i] = urand() - 0.5, E32[i] = (float)Y[i];
    sgo_adjoint_direct_f64(nin, nout, nn, mm, pp, DD, tabs, idx, deriv, Y, Pb);
    double lhs = 0.0, rhs = 0.0, scale = 0.0;
    for (long long i = 0; i < ns * nout; ++i) lhs += E[i] * Y[i], scale += fabs(E[i] * Y[i]);
    for (long long i = 0; i < ncp * nout; ++i) rhs += P[i] * Pb[i];
    require(fabs(lhs - rhs) <= 1e-12 * (scale + 1.0), "transpose identity");
    sgo_adjoint_direct_f32(nin, nout, nn, mm, pp, DD, tabs32, idx, deriv, E32, Pb32, 2);
    for (long long i = 0; i < ncp * nout; ++i)
      require(fabs(Pb32[i] - Pb[i]) <= 1e-4 * (1.0 + fabs(Pb[i])), "FP32 adjoint");
    /* derivative orders */
    for (int d = 0; d < nin; ++d) deriv[d] = DD[d];
    sgo_eval_direct_f64(nin, nout, nn, mm, pp, DD, tabs, idx, deriv, P, NULL, E, 0, mm[nin - 1], 1);
    float* E2 = malloc(sizeof(float) * ns * nout);
    sgo_eval_direct_f32(nin, nout, nn, mm, pp, DD, tabs32, idx, deriv, P32, NULL, E2, 0,
                        mm[nin - 1], 2);
    double mx = 0.0;
    for (long long i = 0; i < ns * nout; ++i) mx = fmax(mx, fabs(E[i]));
    for (long long i = 0; i < ns * nout; ++i)
      require(fabs(E2[i] - E[i]) <= 1e-4 * (mx + 1.0), "FP32 forward");
    free(E2);
  }
  free(P), free(W), free(E), free(Y), free(Pb), free(P32), free(E32), free(Pb32);
  for (int d = 0; d < nin; ++d) free_dim(&dims[d]);
}

/* a random banded refinement matrix (m x n, m >= n), 1-based storage */
typedef struct {
  int m, n;
  int32_t* rp;
  int32_t* cs;
  double* nz;
  int nnz;
} Mat;

static Mat make_mat(int m, int n) {
  Mat a;
  a.m = m, a.n = n;
  a.rp = malloc(sizeof(int32_t) * m);
  a.cs = malloc(sizeof(int32_t) * m);
  a.nz = malloc(sizeof(double) * m * 3);
  a.nnz = 0;
  for (int i = 0; i < m; ++i) {
    int c0 = (int)((long long)i * n / m) + 1;
    int c1 = c0 + (i % 2);
    if (c1 > n) c1 = n;
    a.rp[i] = a.nnz + 1;
    a.cs[i] = c0;
    for (int c = c0; c <= c1; ++c) a.nz[a.nnz++] = 0.1 + urand();
  }
  return a;
}

static void refmat_checks(void) {
  const int shape_b[4] = {5, 4, 3, 2};
  Mat A = make_mat(8, 5), B = make_mat(6, 3);
  sgo_refmat mats[2] = {{A.m, A.n, A.nnz, A.rp, A.cs, A.nz}, {B.m, B.n, B.nnz, B.rp, B.cs, B.nz}};
  const int32_t refmat_of[4] = {1, 0, 2, 0}; /* dimensions 1 and 3 refined */
  const int32_t sb[4] = {5, 4, 3, 2}, sy[4] = {8, 4, 6, 2};
  const int nb = 5 * 4 * 3 * 2, ny = 8 * 4 * 6 * 2;
  double* Xb = malloc(sizeof(double) * nb);
  double* Yv = malloc(sizeof(double) * ny);
  double* Zy = malloc(sizeof(double) * ny);
  double* Bt = malloc(sizeof(double) * nb);
  for (int i = 0; i < nb; ++i) Xb[i] = urand() - 0.5;
  for (int i = 0; i < ny; ++i) Zy[i] = urand() - 0.5;
  sgo_refmat_mult(4, sy, sb, mats, refmat_of, Xb, Yv);
  sgo_refmat_mult_adjoint(4, sy, sb, mats, refmat_of, Zy, Bt);
  double lhs = 0.0, rhs = 0.0;
  for (int i = 0; i < ny; ++i) lhs += Yv[i] * Zy[i];
  for (int i = 0; i < nb; ++i) rhs += Xb[i] * Bt[i];
  require(fabs(lhs - rhs) <= 1e-12 * (fabs(lhs) + 1.0), "refinement transpose identity");
  /* the Flag scatter marks exactly the columns some flagged row holds */
  uint8_t* fy = malloc(ny);
  uint8_t* fb = malloc(nb);
  for (int i = 0; i < ny; ++i) fy[i] = urand() < 0.2;
  sgo_refmat_mult_adjoint_flags(4, sy, sb, mats, refmat_of, fy, fb);
  for (int i = 0; i < ny; ++i) Zy[i] = fy[i];
  sgo_refmat_mult_adjoint(4, sy, sb, mats, refmat_of, Zy, Bt);
  for (int i = 0; i < nb; ++i) require(fb[i] == (Bt[i] != 0.0), "Flag scatter");
  /* overwrite then gather returns the values and zeroes the entries */
  const int32_t shape[2] = {6, 5};
  double cp[6 * 5 * 2];
  for (int i = 0; i < 60; ++i) cp[i] = urand();
  const int32_t ind[3 * 2] = {1, 6, 3, 1, 5, 2}; /* [3, 2] column-major: (1,1) (6,5) (3,2) */
  const double vals[3 * 2] = {10, 11, 12, 20, 21, 22};
  double back[3 * 2];
  sgo_local_refinement_overwrite(2, 2, shape, 3, ind, vals, cp);
  sgo_local_refinement_gather(2, 2, shape, 3, ind, back, cp);
  require(memcmp(back, vals, sizeof(vals)) == 0, "gather returns the overwritten values");
  require(cp[0] == 0.0 && cp[5 + 6 * 4] == 0.0 && cp[2 + 6 * 1 + 30] == 0.0, "gather zeroes");
  (void)shape_b;
  free(Xb), free(Yv), free(Zy), free(Bt), free(fy), free(fb);
  free(A.rp), free(A.cs), free(A.nz), free(B.rp), free(B.cs), free(B.nz);
}

int main(void) {
  int grids = 0;
  for (int p = 0; p <= 4; ++p) {
    const int n2[2] = {p + 3, p + 5}, p2[2] = {p, p > 0 ? p - 1 : 0}, m2[2] = {11, 9};
    grid_checks(2, n2, p2, m2, 2, 0), ++grids;
    grid_checks(2, n2, p2, m2, 3, 1), ++grids;
    const int n3[3] = {p + 2, 6, p + 4}, p3[3] = {p, 2, p > 1 ? 1 : p}, m3[3] = {7, 6, 5};
    grid_checks(3, n3, p3, m3, 2, 0), ++grids;
    grid_checks(3, n3, p3, m3, 1, 1), ++grids;
  }
  const int n1[1] = {7}, p1[1] = {3}, m1[1] = {19};
  grid_checks(1, n1, p1, m1, 2, 0), ++grids;
  refmat_checks();
  printf("oracle_check ok: %d grids, %d checks, %d threads\n", grids, g_checks,
         sgo_max_threads());
  return 0;
}
