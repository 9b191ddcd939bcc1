/*
 * sanitize_driver.c - runs the C restatement (o2d_oracle.c) under AddressSanitizer and
 * UndefinedBehaviorSanitizer. TEST INFRASTRUCTURE ONLY: built by "make -C oracle sanitize",
 * run by tests/test_oracle_sanitize.py.
 *
 * It calls every routine on small seeded inputs, including the edge cases the tests cover
 * (empty inputs, one element, coincident source and target, out-of-range arguments), with
 * every buffer allocated at exactly its stated length so that any read or write past the end
 * is reported. It prints "sanitize ok" and exits 0 when no sanitizer fired; a sanitizer report
 * ends the process with a non-zero status.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "o2d_oracle.h"

static uint64_t g_state = 0x9E3779B97F4A7C15ull;

static double urand(double a, double b) {
  g_state ^= g_state << 13;
  g_state ^= g_state >> 7;
  g_state ^= g_state << 17;
  return a + (b - a) * (double)(g_state >> 11) * (1.0 / 9007199254740992.0);
}

static double* arr(int64_t n, double a, double b) {
  double* p = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  for (int64_t i = 0; i < n; ++i) p[i] = urand(a, b);
  return p;
}

static double* zeros(int64_t n) { return (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

static int g_fail = 0;
#define CHECK(cond)                                               \
  do {                                                            \
    if (!(cond)) {                                                \
      fprintf(stderr, "check failed: %s (line %d)\n", #cond, __LINE__); \
      g_fail = 1;                                                 \
    }                                                             \
  } while (0)

/* a closed polygon of np panels on a circle of radius r, nodes clockwise */
static void circle(int64_t np, double r, double** nx, double** ny, uint32_t** idx) {
  *nx = (double*)malloc((size_t)np * sizeof(double));
  *ny = (double*)malloc((size_t)np * sizeof(double));
  *idx = (uint32_t*)malloc((size_t)(2 * np) * sizeof(uint32_t));
  for (int64_t i = 0; i < np; ++i) {
    const double t = 2.0 * 3.14159265358979323846 * (double)i / (double)np;
    (*nx)[i] = r * cos(t);
    (*ny)[i] = -r * sin(t);
    (*idx)[2 * i] = (uint32_t)i;
    (*idx)[2 * i + 1] = (uint32_t)((i + 1) % np);
  }
}

static void points(void) {
  const int64_t ns = 97, nt = 53;
  double *sx = arr(ns, -1, 1), *sy = arr(ns, -1, 1), *ss = arr(ns, -1, 1), *sr = arr(ns, 0.01, 0.1);
  double *tx = arr(nt, -1, 1), *ty = arr(nt, -1, 1), *tr = arr(nt, 0.01, 0.1);
  /* a coincident pair: the first target sits on the first source */
  tx[0] = sx[0];
  ty[0] = sy[0];
  for (int core = 1; core <= 5; ++core) {
    const int res[3] = {1, 4, 5};
    for (int k = 0; k < 3; ++k) {
      for (int point = 0; point < 2; ++point) {
        double *u = zeros(nt), *v = zeros(nt), *w = zeros(nt), *e = zeros(nt);
        const int rc = oc_points_on_points(core, res[k], ns, sx, sy, ss, sr, nt, tx, ty,
                                           point ? NULL : tr, u, v, res[k] == 4 ? w : NULL,
                                           res[k] == 5 ? e : NULL);
        CHECK(rc == 0);
        for (int64_t i = 0; i < nt; ++i) CHECK(isfinite(u[i]) && isfinite(v[i]));
        free(u); free(v); free(w); free(e);
      }
    }
    /* the adjoint against point targets */
    double *gu = arr(nt, -1, 1), *gv = arr(nt, -1, 1);
    double *gx = zeros(ns), *gy = zeros(ns), *gs = zeros(ns), *gr = zeros(ns);
    CHECK(oc_points_on_points_adjoint(core, ns, sx, sy, ss, sr, nt - 1, tx + 1, ty + 1, gu, gv,
                                      gx, gy, gs, gr) == 0);
    free(gu); free(gv); free(gx); free(gy); free(gs); free(gr);
  }
  /* empty inputs and bad arguments return without touching memory */
  double *one = zeros(1), *eu = zeros(nt), *ev = zeros(nt);
  /* no sources: every target is written (+= 0) and nothing is read from the sources */
  CHECK(oc_points_on_points(2, 1, 0, NULL, NULL, NULL, NULL, nt, tx, ty, NULL, eu, ev, NULL, NULL) == 0);
  for (int64_t i = 0; i < nt; ++i) CHECK(eu[i] == 0.0 && ev[i] == 0.0);
  free(eu); free(ev);
  CHECK(oc_points_on_points(2, 1, ns, sx, sy, ss, sr, 0, tx, ty, NULL, one, one, NULL, NULL) == 0);
  CHECK(oc_points_on_points(9, 1, ns, sx, sy, ss, sr, nt, tx, ty, NULL, one, one, NULL, NULL) < 0);
  CHECK(oc_points_on_points(2, 7, ns, sx, sy, ss, sr, nt, tx, ty, NULL, one, one, NULL, NULL) < 0);
  CHECK(oc_points_on_points(2, 1, -1, sx, sy, ss, sr, nt, tx, ty, NULL, one, one, NULL, NULL) < 0);
  /* Gaussian vorticity of particles (src/InfluenceVort.h) */
  double* w = zeros(nt);
  CHECK(oc_points_vorticity(ns, sx, sy, ss, sr, nt, tx, ty, w) == 0);
  free(w);
  /* Morton keys */
  uint32_t* keys = (uint32_t*)malloc((size_t)ns * sizeof(uint32_t));
  CHECK(oc_morton_keys(ns, sx, sy, -1.0, -1.0, 0.5, keys) == 0);
  free(keys);
  free(one);
  free(sx); free(sy); free(ss); free(sr); free(tx); free(ty); free(tr);
}

static void panels_and_bem(void) {
  const int64_t np = 24, nt = 41;
  double *nx, *ny;
  uint32_t* idx;
  circle(np, 0.5, &nx, &ny, &idx);
  double *vs = arr(np, -1, 1), *ss = arr(np, -1, 1);
  double *tx = arr(nt, -1, 1), *ty = arr(nt, -1, 1);
  double *u = zeros(nt), *v = zeros(nt);
  CHECK(oc_panels_on_points(np, nx, ny, np, idx, vs, ss, nt, tx, ty, u, v) == 0);
  CHECK(oc_panels_on_points(np, nx, ny, np, idx, vs, NULL, nt, tx, ty, u, v) == 0);
  double *px = arr(nt, -1, 1), *py = arr(nt, -1, 1), *ps = arr(nt, -1, 1);
  double *pu = zeros(np), *pv = zeros(np);
  CHECK(oc_points_on_panels(nt, px, py, ps, np, nx, ny, np, idx, pu, pv) == 0);
  double *tgx = zeros(np), *tgy = zeros(np), *nrx = zeros(np), *nry = zeros(np), *len = zeros(np);
  double *area = zeros(np);
  CHECK(oc_panel_geometry(np, nx, ny, np, idx, tgx, tgy, nrx, nry, len) == 0);
  double vol = 0.0, cx = 0.0, cy = 0.0;
  CHECK(oc_body_volume(np, nx, ny, np, idx, &vol, &cx, &cy) == 0);
  double *urs = zeros(np), *urv = zeros(np);
  CHECK(oc_unit_rot_strengths(np, nx, ny, np, idx, cx, cy, urs, urv) == 0);
  for (int nunk = 1; nunk <= 2; ++nunk) {
    for (int aug = 0; aug < 2; ++aug) {
      const int64_t n = nunk * np + aug;
      double* A = zeros(n * n);
      double *ru = arr(np, -1, 1), *rv = arr(np, -1, 1);
      CHECK(oc_bem_assemble(np, nx, ny, np, idx, nunk, aug, ru, rv, vol, A, n) == 0);
      double* rhs = zeros(nunk * np);
      CHECK(oc_vels_to_rhs(np, nunk, pu, pv, tgx, tgy, NULL, rhs) == 0);
      free(A); free(ru); free(rv); free(rhs);
    }
  }
  /* one body on another: a second, smaller circle */
  double *mx, *my;
  uint32_t* midx;
  const int64_t mp = 10;
  circle(mp, 0.2, &mx, &my, &midx);
  for (int64_t i = 0; i < mp; ++i) mx[i] += 2.0;
  for (int nunk = 1; nunk <= 2; ++nunk) {
    const int64_t nr = nunk * mp + 1, nc = nunk * np + 1;
    double* B = zeros(nr * nc);
    double *ru = arr(mp, -1, 1), *rv = arr(mp, -1, 1);
    CHECK(oc_bem_assemble_pair(np, nx, ny, np, idx, mp, mx, my, mp, midx, nunk, 1, 1, ru, rv, B,
                               nr) == 0);
    free(B); free(ru); free(rv);
  }
  /* finalize and the RK combine of 1, 2 and 3 stages */
  double *x = arr(nt, -1, 1), *y = arr(nt, -1, 1);
  double* us[3] = {arr(nt, -1, 1), arr(nt, -1, 1), arr(nt, -1, 1)};
  double* vv[3] = {arr(nt, -1, 1), arr(nt, -1, 1), arr(nt, -1, 1)};
  const double wts[3][3] = {{1.0, 0, 0}, {0.25, 0.75, 0}, {2.0 / 9, 3.0 / 9, 4.0 / 9}};
  for (int st = 1; st <= 3; ++st)
    CHECK(oc_rk_combine(nt, st, wts[st - 1], (const double* const*)us, (const double* const*)vv,
                        0.01, x, y) == 0);
  CHECK(oc_finalize(nt, u, v, 1.0, 0.0) == 0);
  /* proximity: nearest distances, reflection, clearing with both methods */
  double* dist = zeros(nt);
  CHECK(oc_nearest_distances(np, nx, ny, np, idx, nt, tx, ty, dist) == 0);
  double *qx = arr(nt, -0.6, 0.6), *qy = arr(nt, -0.6, 0.6);
  int64_t nrefl = 0;
  CHECK(oc_reflect(np, nx, ny, np, idx, nrx, nry, nt, qx, qy, &nrefl) == 0);
  double pos[256], sm[256], ds[256];
  const int rows = oc_cut_table(0.1, 256, pos, sm, ds);
  CHECK(rows > 0 && rows <= 256);
  for (int method = 0; method <= 1; ++method) {
    double *cx2 = arr(nt, -0.7, 0.7), *cy2 = arr(nt, -0.7, 0.7), *cs = arr(nt, -1, 1);
    double* cr = arr(nt, 0.01, 0.02);
    double circ = 0.0;
    int64_t cropped = 0, unsettled = 0;
    CHECK(oc_clear_inner(method, np, nx, ny, np, idx, nrx, nry, 1.0, 0.01, nt, cx2, cy2, cs, cr,
                         64, &circ, &cropped, &unsettled) == 0);
    CHECK(unsettled == 0);
    free(cx2); free(cy2); free(cs); free(cr);
  }
  free(dist); free(qx); free(qy);
  free(x); free(y);
  for (int k = 0; k < 3; ++k) { free(us[k]); free(vv[k]); }
  free(mx); free(my); free(midx);
  free(urs); free(urv); free(tgx); free(tgy); free(nrx); free(nry); free(len); free(area);
  free(px); free(py); free(ps); free(pu); free(pv);
  free(u); free(v); free(tx); free(ty); free(vs); free(ss);
  free(nx); free(ny); free(idx);
}

int main(void) {
  points();
  panels_and_bem();
  if (g_fail) return 1;
  printf("sanitize ok\n");
  return 0;
}
