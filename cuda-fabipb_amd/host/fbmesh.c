/*
 * fbmesh: deterministic synthetic molecule generator.
 *
 * Writes the three files the fabipb input reader consumes for one molecule:
 *   <base>.vert  MSMS vertex file: 2 comment lines, "nv natm den prob",
 *                then "x y z nx ny nz 0 atom 2" per vertex
 *   <base>.face  MSMS face file:   2 comment lines, "nf natm den prob",
 *                then "i j k 1 1" per face (1-based, outward CCW)
 *   <base>.pqr   ATOM records "... x y z charge radius"
 * which is the format src/input.c:320-821 (loadPanel) of the reference reads
 * when FABIPB_REUSE_MESH=1 is set.
 *
 * Every number is printed with %.17g so the reader recovers the exact binary
 * value, and every random draw comes from one named generator: SplitMix64
 * seeding xoshiro256** (Blackman & Vigna), uniform doubles from the top 53
 * bits, normals from Box-Muller. The same seed gives bit-identical files on
 * any x86-64 host.
 *
 * Shapes:
 *   sphere   level-L icosphere of radius R (standard 12-vertex icosahedron,
 *            4-way midpoint subdivision, vertices projected to radius R),
 *            radial vertex normals.
 *   globule  the same icosphere with each vertex direction u moved to
 *            r(u) = R (1 + amp * sum_{l=1..lmax} sum_m a_lm Y_lm(u)),
 *            a_lm ~ N(0,1)/l, real orthonormal Y_lm (no Condon-Shortley
 *            phase); vertex normals are the analytic surface normals,
 *            evaluated by central differences of the parametrisation.
 *   shell    a hollow shell: an outer globule surface of radius R and an
 *            inner one of radius --inner (both modulated with --amp, --lmax,
 *            separate coefficients, outer drawn first), each a level-L
 *            icosphere; the inner surface's faces are reversed and its
 *            normals point toward the centre, i.e. out of the wall.  Its
 *            radius function uses recurrences for P_l^m and cos/sin(m phi)
 *            (the same function, evaluated faster for lmax ~ 12).
 * Charges:
 *   ball     nq charges uniform in the ball |x| <= rq, q ~ U[-1,1]
 *   center   one charge +q0 at the origin
 *   inside   nq charges uniform inside the star-shaped surface at least
 *            `depth` from every mesh triangle (exact point-triangle
 *            distance), q ~ N(0, sigma), shifted so the total is the
 *            nearest integer.
 *   wall     (shell) nq charges uniform in the wall between the two
 *            surfaces, at least `depth` from every triangle, q ~ N(0, sigma),
 *            total shifted to the nearest integer.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ---------------- PRNG: SplitMix64 -> xoshiro256** ---------------- */
typedef struct { uint64_t s[4]; } rng_t;

static uint64_t splitmix64(uint64_t *x) {
  uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static void rng_seed(rng_t *r, uint64_t seed) {
  uint64_t x = seed;
  for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}
static uint64_t rng_next(rng_t *r) {
  uint64_t *s = r->s;
  uint64_t result = rotl(s[1] * 5, 7) * 9;
  uint64_t t = s[1] << 17;
  s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
  s[2] ^= t; s[3] = rotl(s[3], 45);
  return result;
}
/* uniform in [0,1) from the top 53 bits */
static double rng_uniform(rng_t *r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }
/* standard normal, Box-Muller (one draw per call, the sine branch is dropped) */
static double rng_normal(rng_t *r) {
  double u1 = rng_uniform(r), u2 = rng_uniform(r);
  if (u1 < 1e-300) u1 = 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* ---------------- icosphere ---------------- */
typedef struct {
  int nv, nf, capv, capf;
  double *v;   /* 3*nv unit directions */
  int *f;      /* 3*nf */
} mesh_t;

static void mesh_push_v(mesh_t *m, double x, double y, double z) {
  if (m->nv == m->capv) {
    m->capv = m->capv ? 2 * m->capv : 64;
    m->v = realloc(m->v, sizeof(double) * 3 * (size_t)m->capv);
  }
  double n = sqrt(x * x + y * y + z * z);
  m->v[3 * m->nv] = x / n; m->v[3 * m->nv + 1] = y / n; m->v[3 * m->nv + 2] = z / n;
  m->nv++;
}
static void mesh_push_f(mesh_t *m, int a, int b, int c) {
  if (m->nf == m->capf) {
    m->capf = m->capf ? 2 * m->capf : 64;
    m->f = realloc(m->f, sizeof(int) * 3 * (size_t)m->capf);
  }
  m->f[3 * m->nf] = a; m->f[3 * m->nf + 1] = b; m->f[3 * m->nf + 2] = c;
  m->nf++;
}

/* open-addressing map from an undirected edge (a<b) to its midpoint vertex */
typedef struct { uint64_t *key; int *val; size_t cap; } emap_t;
static int emap_get(emap_t *e, mesh_t *m, int a, int b) {
  if (a > b) { int t = a; a = b; b = t; }
  uint64_t k = ((uint64_t)a << 32) | (uint64_t)b;
  size_t h = (size_t)((k * 0x9E3779B97F4A7C15ULL) >> 17) & (e->cap - 1);
  while (e->key[h] != UINT64_MAX) {
    if (e->key[h] == k) return e->val[h];
    h = (h + 1) & (e->cap - 1);
  }
  const double *va = &m->v[3 * a], *vb = &m->v[3 * b];
  mesh_push_v(m, va[0] + vb[0], va[1] + vb[1], va[2] + vb[2]);
  e->key[h] = k;
  e->val[h] = m->nv - 1;
  return m->nv - 1;
}

static void icosphere(mesh_t *m, int level) {
  const double p = (1.0 + sqrt(5.0)) / 2.0;
  static const int F[20][3] = {
      {0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11},
      {1, 5, 9}, {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
      {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9},
      {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
  memset(m, 0, sizeof(*m));
  mesh_push_v(m, -1, p, 0); mesh_push_v(m, 1, p, 0);
  mesh_push_v(m, -1, -p, 0); mesh_push_v(m, 1, -p, 0);
  mesh_push_v(m, 0, -1, p); mesh_push_v(m, 0, 1, p);
  mesh_push_v(m, 0, -1, -p); mesh_push_v(m, 0, 1, -p);
  mesh_push_v(m, p, 0, -1); mesh_push_v(m, p, 0, 1);
  mesh_push_v(m, -p, 0, -1); mesh_push_v(m, -p, 0, 1);
  for (int i = 0; i < 20; i++) mesh_push_f(m, F[i][0], F[i][1], F[i][2]);
  for (int l = 0; l < level; l++) {
    emap_t e;
    e.cap = 1;
    while (e.cap < (size_t)m->nf * 4) e.cap <<= 1;
    e.key = malloc(sizeof(uint64_t) * e.cap);
    e.val = malloc(sizeof(int) * e.cap);
    for (size_t i = 0; i < e.cap; i++) e.key[i] = UINT64_MAX;
    int nf0 = m->nf;
    int *f0 = malloc(sizeof(int) * 3 * (size_t)nf0);
    memcpy(f0, m->f, sizeof(int) * 3 * (size_t)nf0);
    m->nf = 0;
    for (int i = 0; i < nf0; i++) {
      int a = f0[3 * i], b = f0[3 * i + 1], c = f0[3 * i + 2];
      int ab = emap_get(&e, m, a, b), bc = emap_get(&e, m, b, c), ca = emap_get(&e, m, c, a);
      mesh_push_f(m, a, ab, ca);
      mesh_push_f(m, b, bc, ab);
      mesh_push_f(m, c, ca, bc);
      mesh_push_f(m, ab, bc, ca);
    }
    free(f0); free(e.key); free(e.val);
  }
}

/* ---------------- real spherical harmonics ---------------- */
typedef struct { int lmax; double amp, R; double *a; int fast; double *N; } shape_t; /* a, N: [l*l + l + m] */

/* associated Legendre P_l^m(x) without Condon-Shortley phase, m >= 0 */
static double plm(int l, int m, double x) {
  double pmm = 1.0, s = sqrt(fmax(0.0, 1.0 - x * x)), fact = 1.0;
  for (int i = 1; i <= m; i++) { pmm *= fact * s; fact += 2.0; }
  if (l == m) return pmm;
  double pmmp1 = x * (2 * m + 1) * pmm;
  if (l == m + 1) return pmmp1;
  double pll = 0.0;
  for (int ll = m + 2; ll <= l; ll++) {
    pll = (x * (2 * ll - 1) * pmmp1 - (ll + m - 1) * pmm) / (ll - m);
    pmm = pmmp1; pmmp1 = pll;
  }
  return pll;
}
static double ylm(int l, int m, const double *u) {
  int am = m < 0 ? -m : m;
  double lf = 1.0;  /* (l-|m|)!/(l+|m|)! */
  for (int i = l - am + 1; i <= l + am; i++) lf /= i;
  double N = sqrt((2 * l + 1) / (4.0 * M_PI) * lf);
  double ph = atan2(u[1], u[0]);
  double P = plm(l, am, u[2]);
  if (m == 0) return N * P;
  if (m > 0) return sqrt(2.0) * N * P * cos(m * ph);
  return sqrt(2.0) * N * P * sin(am * ph);
}
static double shape_r(const shape_t *s, const double *u) {
  if (s->lmax <= 0) return s->R;
  double sum = 0.0;
  for (int l = 1; l <= s->lmax; l++)
    for (int m = -l; m <= l; m++) sum += s->a[l * l + l + m] * ylm(l, m, u);
  return s->R * (1.0 + s->amp * sum);
}
/* normalisations sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) of the real Y_lm, m >= 0 */
static double *ylm_norms(int lmax) {
  double *N = calloc((size_t)(lmax + 1) * (lmax + 1), sizeof(double));
  for (int l = 0; l <= lmax; l++)
    for (int m = 0; m <= l; m++) {
      double lf = 1.0;
      for (int i = l - m + 1; i <= l + m; i++) lf /= i;
      N[l * l + l + m] = sqrt((2 * l + 1) / (4.0 * M_PI) * lf);
    }
  return N;
}

/* the same r(u) with the Legendre functions by recurrence in l and the
 * azimuthal factors by recurrence in m: O(lmax^2) per direction */
static double shape_r_fast(const shape_t *s, const double *u) {
  if (s->lmax <= 0) return s->R;
  enum { LM = 32 };
  const int L = s->lmax < LM - 1 ? s->lmax : LM - 1;
  const double x = u[2], st = sqrt(fmax(0.0, 1.0 - x * x)), ph = atan2(u[1], u[0]);
  double cm[LM], sm[LM];
  const double c1 = cos(ph), s1 = sin(ph);
  cm[0] = 1.0; sm[0] = 0.0;
  for (int m = 1; m <= L; m++) { cm[m] = cm[m - 1] * c1 - sm[m - 1] * s1; sm[m] = sm[m - 1] * c1 + cm[m - 1] * s1; }
  double sum = 0.0, pmm = 1.0, fact = 1.0;
  for (int m = 0; m <= L; m++) {
    if (m > 0) { pmm *= fact * st; fact += 2.0; }
    double pl2 = 0.0, pl1 = pmm;   /* P_{l-2}^m, P_{l-1}^m */
    for (int l = m; l <= L; l++) {
      double p;
      if (l == m) p = pmm;
      else if (l == m + 1) p = x * (2 * m + 1) * pmm;
      else p = (x * (2 * l - 1) * pl1 - (l + m - 1) * pl2) / (l - m);
      if (l > m) { pl2 = pl1; pl1 = p; }
      if (l >= 1) {
        const double N = s->N[l * l + l + m];
        if (m == 0) sum += s->a[l * l + l] * N * p;
        else sum += sqrt(2.0) * N * p * (s->a[l * l + l + m] * cm[m] + s->a[l * l + l - m] * sm[m]);
      }
    }
  }
  return s->R * (1.0 + s->amp * sum);
}

static void normalize3(double *x) {
  double n = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  x[0] /= n; x[1] /= n; x[2] /= n;
}
static void shape_point(const shape_t *s, const double *u, double *p) {
  double r = s->fast ? shape_r_fast(s, u) : shape_r(s, u);
  p[0] = r * u[0]; p[1] = r * u[1]; p[2] = r * u[2];
}
/* outward surface normal at direction u, by central differences of p(u) */
static void shape_normal(const shape_t *s, const double *u, double *n) {
  if (s->lmax <= 0) { n[0] = u[0]; n[1] = u[1]; n[2] = u[2]; return; }
  double t1[3], t2[3], ax[3] = {1, 0, 0};
  if (fabs(u[0]) > 0.9) { ax[0] = 0; ax[1] = 1; }
  t1[0] = u[1] * ax[2] - u[2] * ax[1]; t1[1] = u[2] * ax[0] - u[0] * ax[2]; t1[2] = u[0] * ax[1] - u[1] * ax[0];
  normalize3(t1);
  t2[0] = u[1] * t1[2] - u[2] * t1[1]; t2[1] = u[2] * t1[0] - u[0] * t1[2]; t2[2] = u[0] * t1[1] - u[1] * t1[0];
  const double h = 1e-5;
  double d1[3], d2[3], q[3], pp[3], pm[3];
  for (int k = 0; k < 3; k++) q[k] = u[k] + h * t1[k];
  normalize3(q); shape_point(s, q, pp);
  for (int k = 0; k < 3; k++) q[k] = u[k] - h * t1[k];
  normalize3(q); shape_point(s, q, pm);
  for (int k = 0; k < 3; k++) d1[k] = pp[k] - pm[k];
  for (int k = 0; k < 3; k++) q[k] = u[k] + h * t2[k];
  normalize3(q); shape_point(s, q, pp);
  for (int k = 0; k < 3; k++) q[k] = u[k] - h * t2[k];
  normalize3(q); shape_point(s, q, pm);
  for (int k = 0; k < 3; k++) d2[k] = pp[k] - pm[k];
  n[0] = d1[1] * d2[2] - d1[2] * d2[1];
  n[1] = d1[2] * d2[0] - d1[0] * d2[2];
  n[2] = d1[0] * d2[1] - d1[1] * d2[0];
  normalize3(n);
  if (n[0] * u[0] + n[1] * u[1] + n[2] * u[2] < 0) { n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2]; }
}

/* ---------------- point-triangle distance (Ericson, RTCD 5.1.5) ---------------- */
static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static double pt_tri_dist2(const double *p, const double *a, const double *b, const double *c) {
  double ab[3], ac[3], ap[3], bp[3], cp[3], q[3];
  for (int k = 0; k < 3; k++) { ab[k] = b[k] - a[k]; ac[k] = c[k] - a[k]; ap[k] = p[k] - a[k]; }
  double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  if (d1 <= 0 && d2 <= 0) { for (int k = 0; k < 3; k++) q[k] = a[k]; goto done; }
  for (int k = 0; k < 3; k++) bp[k] = p[k] - b[k];
  double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  if (d3 >= 0 && d4 <= d3) { for (int k = 0; k < 3; k++) q[k] = b[k]; goto done; }
  double vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) {
    double v = d1 / (d1 - d3);
    for (int k = 0; k < 3; k++) q[k] = a[k] + v * ab[k];
    goto done;
  }
  for (int k = 0; k < 3; k++) cp[k] = p[k] - c[k];
  double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
  if (d6 >= 0 && d5 <= d6) { for (int k = 0; k < 3; k++) q[k] = c[k]; goto done; }
  double vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) {
    double w = d2 / (d2 - d6);
    for (int k = 0; k < 3; k++) q[k] = a[k] + w * ac[k];
    goto done;
  }
  double va = d3 * d6 - d5 * d4;
  if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
    double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    for (int k = 0; k < 3; k++) q[k] = b[k] + w * (c[k] - b[k]);
    goto done;
  }
  {
    double denom = 1.0 / (va + vb + vc);
    double v = vb * denom, w = vc * denom;
    for (int k = 0; k < 3; k++) q[k] = a[k] + ab[k] * v + ac[k] * w;
  }
done:
  return (p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) + (p[2] - q[2]) * (p[2] - q[2]);
}

/* ---------------- merging of small faces ---------------- */
/* Edge collapse of faces below area amin: for each such face, in index order,
 * its edges shortest first are collapsed to the surface point over the mean of
 * the two directions.  A collapse is skipped when it would break the closed
 * manifold (the two end vertices must share exactly the two opposite
 * vertices), leave a vertex of valence below 4, or flip or flatten a face.  A
 * collapse locks the vertices of every face it touched for the rest of the
 * pass; passes repeat until one makes no collapse.  Fully deterministic. */
static void tri_normal(const double *a, const double *b, const double *c, double *n) {
  double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  n[0] = e1[1] * e2[2] - e1[2] * e2[1];
  n[1] = e1[2] * e2[0] - e1[0] * e2[2];
  n[2] = e1[0] * e2[1] - e1[1] * e2[0];
}
static double tri_area(const double *a, const double *b, const double *c) {
  double n[3];
  tri_normal(a, b, c, n);
  return 0.5 * sqrt(dot3(n, n));
}

static int merge_small(mesh_t *m, const shape_t *s, double amin, double *pos) {
  int nv = m->nv, nf = m->nf, collapsed = 0;
  char *alive = malloc(nf), *lock = malloc(nv);
  int *off = malloc(sizeof(int) * (size_t)(nv + 1)), *adj = malloc(sizeof(int) * 3 * (size_t)nf);
  memset(alive, 1, nf);
  for (int pass = 0; pass < 100; pass++) {
    int done = 0;
    memset(lock, 0, nv);
    memset(off, 0, sizeof(int) * (size_t)(nv + 1));
    for (int f = 0; f < nf; f++)
      if (alive[f])
        for (int k = 0; k < 3; k++) off[m->f[3 * f + k] + 1]++;
    for (int v = 0; v < nv; v++) off[v + 1] += off[v];
    int *fill = malloc(sizeof(int) * (size_t)nv);
    memcpy(fill, off, sizeof(int) * (size_t)nv);
    for (int f = 0; f < nf; f++)
      if (alive[f])
        for (int k = 0; k < 3; k++) adj[fill[m->f[3 * f + k]]++] = f;
    free(fill);
    for (int f = 0; f < nf; f++) {
      if (!alive[f]) continue;
      int *F = &m->f[3 * f];
      if (tri_area(&pos[3 * F[0]], &pos[3 * F[1]], &pos[3 * F[2]]) >= amin) continue;
      int order[3] = {0, 1, 2};
      double len[3];
      for (int k = 0; k < 3; k++) {
        const double *a = &pos[3 * F[k]], *b = &pos[3 * F[(k + 1) % 3]];
        double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
        len[k] = dot3(d, d);
      }
      for (int i = 0; i < 3; i++)
        for (int j = i + 1; j < 3; j++)
          if (len[order[j]] < len[order[i]]) { int t = order[i]; order[i] = order[j]; order[j] = t; }
      for (int oi = 0; oi < 3; oi++) {
        const int u = F[order[oi]], v = F[(order[oi] + 1) % 3];
        if (lock[u] || lock[v]) continue;
        /* faces on the edge and the opposite vertices */
        int ef[2], ne = 0, opp[2];
        for (int q = off[u]; q < off[u + 1]; q++) {
          const int g = adj[q], *G = &m->f[3 * g];
          if (!alive[g]) continue;
          if (G[0] == v || G[1] == v || G[2] == v) {
            if (ne < 2) { ef[ne] = g; opp[ne] = G[0] != u && G[0] != v ? G[0] : G[1] != u && G[1] != v ? G[1] : G[2]; }
            ne++;
          }
        }
        if (ne != 2) continue;
        /* link condition: common neighbours of u and v are exactly opp[0], opp[1] */
        int nbu[64], nu = 0, bad = 0, common = 0;
        for (int q = off[u]; q < off[u + 1] && !bad; q++) {
          const int *G = &m->f[3 * adj[q]];
          if (!alive[adj[q]]) continue;
          for (int k = 0; k < 3; k++) {
            int w = G[k], seen = 0;
            if (w == u) continue;
            for (int t = 0; t < nu; t++) seen |= nbu[t] == w;
            if (!seen) { if (nu == 64) { bad = 1; break; } nbu[nu++] = w; }
          }
        }
        int nvv = 0, seenv[64];
        for (int q = off[v]; q < off[v + 1] && !bad; q++) {
          const int *G = &m->f[3 * adj[q]];
          if (!alive[adj[q]]) continue;
          for (int k = 0; k < 3; k++) {
            int w = G[k], seen = 0;
            if (w == v) continue;
            for (int t = 0; t < nvv; t++) seen |= seenv[t] == w;
            if (seen) continue;
            if (nvv == 64) { bad = 1; break; }
            seenv[nvv++] = w;
            if (w == u) continue;
            for (int t = 0; t < nu; t++)
              if (nbu[t] == w) { common++; if (w != opp[0] && w != opp[1]) bad = 1; }
          }
        }
        if (bad || common != 2 || nu + nvv - 4 < 4) continue;
        int valOpp = 1;
        for (int t = 0; t < 2; t++) {
          int cnt = 0;
          for (int q = off[opp[t]]; q < off[opp[t] + 1]; q++) cnt += alive[adj[q]];
          if (cnt - 1 < 4) valOpp = 0;
        }
        if (!valOpp) continue;
        /* new point and the face checks */
        double dir[3] = {m->v[3 * u] + m->v[3 * v], m->v[3 * u + 1] + m->v[3 * v + 1], m->v[3 * u + 2] + m->v[3 * v + 2]};
        normalize3(dir);
        double pn[3];
        shape_point(s, dir, pn);
        int okf = 1;
        for (int side = 0; side < 2 && okf; side++) {
          const int x = side ? v : u;
          for (int q = off[x]; q < off[x + 1] && okf; q++) {
            const int g = adj[q];
            if (!alive[g] || g == ef[0] || g == ef[1]) continue;
            const int *G = &m->f[3 * g];
            const double *P[3], *Q[3];
            for (int k = 0; k < 3; k++) {
              P[k] = &pos[3 * G[k]];
              Q[k] = (G[k] == u || G[k] == v) ? pn : P[k];
            }
            double n0[3], n1[3];
            tri_normal(P[0], P[1], P[2], n0);
            tri_normal(Q[0], Q[1], Q[2], n1);
            double c = dot3(n0, n1), l0 = sqrt(dot3(n0, n0)), l1 = sqrt(dot3(n1, n1));
            if (!(c > 0.5 * l0 * l1) || !(l1 > 1e-8)) okf = 0;
          }
        }
        if (!okf) continue;
        /* collapse v into u */
        for (int k = 0; k < 3; k++) { m->v[3 * u + k] = dir[k]; pos[3 * u + k] = pn[k]; }
        alive[ef[0]] = alive[ef[1]] = 0;
        for (int side = 0; side < 2; side++) {
          const int x = side ? v : u;
          for (int q = off[x]; q < off[x + 1]; q++) {
            const int g = adj[q];
            int *G = &m->f[3 * g];
            for (int k = 0; k < 3; k++) {
              if (alive[g] && G[k] == v) G[k] = u;
              lock[G[k]] = 1;
            }
          }
        }
        lock[u] = lock[v] = 1;
        done++;
        break;
      }
    }
    collapsed += done;
    if (!done) break;
  }
  /* compact: live faces, used vertices, in their original order */
  int *map = malloc(sizeof(int) * (size_t)nv), nf2 = 0, nv2 = 0;
  for (int v = 0; v < nv; v++) map[v] = -1;
  for (int f = 0; f < nf; f++)
    if (alive[f]) {
      for (int k = 0; k < 3; k++) map[m->f[3 * f + k]] = 0;
      for (int k = 0; k < 3; k++) m->f[3 * nf2 + k] = m->f[3 * f + k];
      nf2++;
    }
  for (int v = 0; v < nv; v++)
    if (map[v] == 0) {
      map[v] = nv2;
      for (int k = 0; k < 3; k++) { m->v[3 * nv2 + k] = m->v[3 * v + k]; pos[3 * nv2 + k] = pos[3 * v + k]; }
      nv2++;
    }
  for (int i = 0; i < 3 * nf2; i++) m->f[i] = map[m->f[i]];
  m->nv = nv2;
  m->nf = nf2;
  free(map); free(alive); free(lock); free(off); free(adj);
  return collapsed;
}

/* ---------------- output ---------------- */
static FILE *open_out(const char *base, const char *ext) {
  char path[4096];
  snprintf(path, sizeof(path), "%s%s", base, ext);
  FILE *fp = fopen(path, "w");
  if (!fp) { fprintf(stderr, "fbmesh: cannot write %s\n", path); exit(1); }
  return fp;
}

static void usage(void);

/* vertices, normals and faces of both surfaces, then charges in the wall */
static int make_shell(const char *base, int level, double R, double Rin, double amp, int lmax, const char *charges,
                      int nq, double sigma, double depth, rng_t *rng, uint64_t seed) {
  if (!(Rin > 0.0 && Rin < R) || strcmp(charges, "wall") != 0 || lmax > 30) usage();
  shape_t so = {lmax, amp, R, calloc((size_t)(lmax + 1) * (lmax + 1), sizeof(double)), 1, ylm_norms(lmax)};
  shape_t si = {lmax, amp, Rin, calloc((size_t)(lmax + 1) * (lmax + 1), sizeof(double)), 1, ylm_norms(lmax)};
  for (int l = 1; l <= lmax; l++)
    for (int m = -l; m <= l; m++) so.a[l * l + l + m] = rng_normal(rng) / l;
  for (int l = 1; l <= lmax; l++)
    for (int m = -l; m <= l; m++) si.a[l * l + l + m] = rng_normal(rng) / l;
  mesh_t m;
  icosphere(&m, level);
  const int nv1 = m.nv, nf1 = m.nf, nv = 2 * nv1, nf = 2 * nf1;
  double *pos = malloc(sizeof(double) * 3 * (size_t)nv), *nrm = malloc(sizeof(double) * 3 * (size_t)nv);
  int *fc = malloc(sizeof(int) * 3 * (size_t)nf);
  for (int i = 0; i < nv1; i++) {
    const double *u = &m.v[3 * i];
    shape_point(&so, u, &pos[3 * i]);
    shape_normal(&so, u, &nrm[3 * i]);
    shape_point(&si, u, &pos[3 * (nv1 + i)]);
    shape_normal(&si, u, &nrm[3 * (nv1 + i)]);
    for (int k = 0; k < 3; k++) nrm[3 * (nv1 + i) + k] = -nrm[3 * (nv1 + i) + k];
  }
  for (int f = 0; f < nf1; f++) {
    for (int k = 0; k < 3; k++) fc[3 * f + k] = m.f[3 * f + k];
    fc[3 * (nf1 + f)] = nv1 + m.f[3 * f];          /* reversed: the normal points inward */
    fc[3 * (nf1 + f) + 1] = nv1 + m.f[3 * f + 2];
    fc[3 * (nf1 + f) + 2] = nv1 + m.f[3 * f + 1];
  }
  /* charges: rejection in the bounding cube, radially inside the wall, then
     the exact point-triangle distance against a uniform grid of centroids */
  double rmax = 0.0, emax = 0.0;
  for (int i = 0; i < nv; i++) rmax = fmax(rmax, sqrt(dot3(&pos[3 * i], &pos[3 * i])));
  double *cen = malloc(sizeof(double) * 3 * (size_t)nf);
  for (int f = 0; f < nf; f++) {
    const double *a = &pos[3 * fc[3 * f]], *b = &pos[3 * fc[3 * f + 1]], *c = &pos[3 * fc[3 * f + 2]];
    for (int k = 0; k < 3; k++) cen[3 * f + k] = (a[k] + b[k] + c[k]) / 3.0;
    for (int k = 0; k < 3; k++) {
      const double *v = (k == 0) ? a : (k == 1) ? b : c;
      double d2 = 0.0;
      for (int d = 0; d < 3; d++) d2 += (v[d] - cen[3 * f + d]) * (v[d] - cen[3 * f + d]);
      emax = fmax(emax, sqrt(d2));
    }
  }
  const double reach = depth + emax, reach2 = reach * reach, depth2 = depth * depth;
  const int G = (int)fmin(256.0, fmax(1.0, floor(2.0 * rmax / reach)));
  const double cell = 2.0 * rmax / G;
  const size_t G3 = (size_t)G * G * G;
  int *gstart = calloc(G3 + 1, sizeof(int)), *gface = malloc(sizeof(int) * (size_t)nf), *gcell = malloc(sizeof(int) * (size_t)nf);
  for (int f = 0; f < nf; f++) {
    int ci[3];
    for (int k = 0; k < 3; k++) {
      ci[k] = (int)floor((cen[3 * f + k] + rmax) / cell);
      ci[k] = ci[k] < 0 ? 0 : ci[k] >= G ? G - 1 : ci[k];
    }
    gcell[f] = (ci[0] * G + ci[1]) * G + ci[2];
    gstart[gcell[f] + 1]++;
  }
  for (size_t i = 0; i < G3; i++) gstart[i + 1] += gstart[i];
  {
    int *fill = malloc(sizeof(int) * G3);
    memcpy(fill, gstart, sizeof(int) * G3);
    for (int f = 0; f < nf; f++) gface[fill[gcell[f]]++] = f;
    free(fill);
  }
  double *qx = malloc(sizeof(double) * 3 * (size_t)nq), *qq = malloc(sizeof(double) * (size_t)nq);
  long tries = 0;
  for (int j = 0; j < nq;) {
    if (++tries > 4000000000L) { fprintf(stderr, "fbmesh: charge placement failed\n"); return 1; }
    double p[3], d[3];
    for (int k = 0; k < 3; k++) p[k] = rmax * (2.0 * rng_uniform(rng) - 1.0);
    double rho = sqrt(dot3(p, p));
    if (rho < 1e-12) continue;
    for (int k = 0; k < 3; k++) d[k] = p[k] / rho;
    if (rho > shape_r_fast(&so, d) - depth * 0.5 || rho < shape_r_fast(&si, d) + depth * 0.5) continue;
    int ok = 1, pc[3];
    for (int k = 0; k < 3; k++) pc[k] = (int)floor((p[k] + rmax) / cell);
    for (int i0 = pc[0] - 1; i0 <= pc[0] + 1 && ok; i0++)
      for (int i1 = pc[1] - 1; i1 <= pc[1] + 1 && ok; i1++)
        for (int i2 = pc[2] - 1; i2 <= pc[2] + 1 && ok; i2++) {
          if (i0 < 0 || i1 < 0 || i2 < 0 || i0 >= G || i1 >= G || i2 >= G) continue;
          const size_t cc = ((size_t)i0 * G + i1) * G + i2;
          for (int q = gstart[cc]; q < gstart[cc + 1] && ok; q++) {
            const int f = gface[q];
            double dx = p[0] - cen[3 * f], dy = p[1] - cen[3 * f + 1], dz = p[2] - cen[3 * f + 2];
            if (dx * dx + dy * dy + dz * dz > reach2) continue;
            if (pt_tri_dist2(p, &pos[3 * fc[3 * f]], &pos[3 * fc[3 * f + 1]], &pos[3 * fc[3 * f + 2]]) < depth2) ok = 0;
          }
        }
    if (!ok) continue;
    for (int k = 0; k < 3; k++) qx[3 * j + k] = p[k];
    qq[j] = sigma * rng_normal(rng);
    j++;
  }
  double sq = 0.0;
  for (int j = 0; j < nq; j++) sq += qq[j];
  const double shift = (nearbyint(sq) - sq) / nq;
  for (int j = 0; j < nq; j++) qq[j] += shift;

  FILE *fp = open_out(base, ".vert");
  fprintf(fp, "# fbmesh shell level=%d R=%.17g Rin=%.17g amp=%.17g lmax=%d seed=%llu\n", level, R, Rin, amp, lmax,
          (unsigned long long)seed);
  fprintf(fp, "#faces  #sphere density probe_r\n");
  fprintf(fp, "%d %d %.2f %.2f\n", nv, nq, 1.0, 1.4);
  for (int i = 0; i < nv; i++)
    fprintf(fp, "%.17g %.17g %.17g %.17g %.17g %.17g 0 %d 2\n", pos[3 * i], pos[3 * i + 1], pos[3 * i + 2],
            nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2], 1);
  fclose(fp);
  fp = open_out(base, ".face");
  fprintf(fp, "# fbmesh shell level=%d\n", level);
  fprintf(fp, "#faces  #sphere density probe_r\n");
  fprintf(fp, "%d %d %.2f %.2f\n", nf, nq, 1.0, 1.4);
  for (int f = 0; f < nf; f++) fprintf(fp, "%d %d %d 1 1\n", fc[3 * f] + 1, fc[3 * f + 1] + 1, fc[3 * f + 2] + 1);
  fclose(fp);
  fp = open_out(base, ".pqr");
  for (int j = 0; j < nq; j++)
    fprintf(fp, "ATOM %6d  C   CHG A%4d    %.17g %.17g %.17g %.17g 1.5000\n", j + 1, j % 10000, qx[3 * j],
            qx[3 * j + 1], qx[3 * j + 2], qq[j]);
  fprintf(fp, "END\n");
  fclose(fp);
  printf("fbmesh: %s vertices=%d faces=%d charges=%d (tries %ld)\n", base, nv, nf, nq, tries);
  free(pos); free(nrm); free(fc); free(cen); free(gstart); free(gface); free(gcell); free(qx); free(qq);
  free(so.a); free(si.a); free(so.N); free(si.N);
  return 0;
}

static void usage(void) {
  fprintf(stderr,
          "usage: fbmesh <sphere|globule|shell> <out-base> [--level L] [--radius R] [--inner Rin]\n"
          "              [--amp A] [--lmax L] [--charges ball|center|inside|wall]\n"
          "              [--nq N] [--rq r] [--q0 q] [--sigma s] [--depth d] [--seed S]\n"
          "              [--merge A]   (collapse faces below area A, square Angstrom)\n");
  exit(2);
}

int main(int argc, char **argv) {
  if (argc < 3) usage();
  const char *kind = argv[1], *base = argv[2];
  int level = 3, lmax = 6, nq = 1;
  double R = 10.0, amp = 0.08, rq = 8.0, q0 = 1.0, sigma = 0.4, depth = 1.5, merge = 0.0, inner = 0.0;
  const char *charges = "center";
  uint64_t seed = 1;
  for (int i = 3; i < argc; i++) {
    const char *a = argv[i];
    if (i + 1 >= argc) usage();
    const char *v = argv[++i];
    if (!strcmp(a, "--level")) level = atoi(v);
    else if (!strcmp(a, "--radius")) R = atof(v);
    else if (!strcmp(a, "--amp")) amp = atof(v);
    else if (!strcmp(a, "--lmax")) lmax = atoi(v);
    else if (!strcmp(a, "--charges")) charges = v;
    else if (!strcmp(a, "--nq")) nq = atoi(v);
    else if (!strcmp(a, "--rq")) rq = atof(v);
    else if (!strcmp(a, "--q0")) q0 = atof(v);
    else if (!strcmp(a, "--sigma")) sigma = atof(v);
    else if (!strcmp(a, "--depth")) depth = atof(v);
    else if (!strcmp(a, "--seed")) seed = strtoull(v, NULL, 10);
    else if (!strcmp(a, "--merge")) merge = atof(v);
    else if (!strcmp(a, "--inner")) inner = atof(v);
    else usage();
  }
  if (level < 0 || level > 9 || nq < 1) usage();

  rng_t rng;
  rng_seed(&rng, seed);
  shape_t shape;
  shape.R = R; shape.amp = amp; shape.lmax = 0; shape.a = NULL; shape.fast = 0; shape.N = NULL;
  if (!strcmp(kind, "shell")) return make_shell(base, level, R, inner, amp, lmax, charges, nq, sigma, depth, &rng, seed);
  if (!strcmp(kind, "globule")) {
    shape.lmax = lmax;
    shape.a = calloc((size_t)(lmax + 1) * (lmax + 1), sizeof(double));
    for (int l = 1; l <= lmax; l++)
      for (int m = -l; m <= l; m++) shape.a[l * l + l + m] = rng_normal(&rng) / l;
  } else if (strcmp(kind, "sphere") != 0) {
    usage();
  }

  mesh_t m;
  icosphere(&m, level);
  double *pos = malloc(sizeof(double) * 3 * (size_t)m.nv);
  double *nrm = malloc(sizeof(double) * 3 * (size_t)m.nv);
  for (int i = 0; i < m.nv; i++) shape_point(&shape, &m.v[3 * i], &pos[3 * i]);
  if (merge > 0.0) {
    int nf0 = m.nf, k = merge_small(&m, &shape, merge, pos);
    printf("fbmesh: merged faces below %g A^2: %d collapses, %d -> %d faces\n", merge, k, nf0, m.nf);
  }
  for (int i = 0; i < m.nv; i++) shape_normal(&shape, &m.v[3 * i], &nrm[3 * i]);

  /* charges */
  double *qx = malloc(sizeof(double) * 3 * (size_t)nq), *qq = malloc(sizeof(double) * (size_t)nq);
  if (!strcmp(charges, "center")) {
    nq = 1; qx[0] = qx[1] = qx[2] = 0.0; qq[0] = q0;
  } else if (!strcmp(charges, "ball")) {
    for (int j = 0; j < nq; j++) {
      double d[3];
      for (int k = 0; k < 3; k++) d[k] = rng_normal(&rng);
      normalize3(d);
      double r = rq * cbrt(rng_uniform(&rng));
      for (int k = 0; k < 3; k++) qx[3 * j + k] = r * d[k];
      qq[j] = 2.0 * rng_uniform(&rng) - 1.0;
    }
  } else if (!strcmp(charges, "inside")) {
    double rmax = 0.0, lmaxEdge = 0.0;
    for (int i = 0; i < m.nv; i++) rmax = fmax(rmax, sqrt(dot3(&pos[3 * i], &pos[3 * i])));
    double *cen = malloc(sizeof(double) * 3 * (size_t)m.nf);
    for (int f = 0; f < m.nf; f++) {
      const double *a = &pos[3 * m.f[3 * f]], *b = &pos[3 * m.f[3 * f + 1]], *c = &pos[3 * m.f[3 * f + 2]];
      for (int k = 0; k < 3; k++) cen[3 * f + k] = (a[k] + b[k] + c[k]) / 3.0;
      for (int k = 0; k < 3; k++) {
        const double *v = (k == 0) ? a : (k == 1) ? b : c;
        double d2 = (v[0] - cen[3 * f]) * (v[0] - cen[3 * f]) + (v[1] - cen[3 * f + 1]) * (v[1] - cen[3 * f + 1]) +
                    (v[2] - cen[3 * f + 2]) * (v[2] - cen[3 * f + 2]);
        lmaxEdge = fmax(lmaxEdge, sqrt(d2));
      }
    }
    double reach = depth + lmaxEdge, reach2 = reach * reach, depth2 = depth * depth;
    /* uniform grid of face centroids, cell = reach: a point's candidates lie in
       the 27 cells around it */
    const int G = (int)fmin(256.0, fmax(1.0, floor(2.0 * rmax / reach)));
    const double cell = 2.0 * rmax / G;
    int *gstart = calloc((size_t)G * G * G + 1, sizeof(int)), *gface = malloc(sizeof(int) * (size_t)m.nf);
    int *gcell = malloc(sizeof(int) * (size_t)m.nf);
    for (int f = 0; f < m.nf; f++) {
      int ci[3];
      for (int k = 0; k < 3; k++) {
        ci[k] = (int)floor((cen[3 * f + k] + rmax) / cell);
        ci[k] = ci[k] < 0 ? 0 : ci[k] >= G ? G - 1 : ci[k];
      }
      gcell[f] = (ci[0] * G + ci[1]) * G + ci[2];
      gstart[gcell[f] + 1]++;
    }
    for (int i = 0; i < G * G * G; i++) gstart[i + 1] += gstart[i];
    {
      int *fill = malloc(sizeof(int) * (size_t)G * G * G);
      memcpy(fill, gstart, sizeof(int) * (size_t)G * G * G);
      for (int f = 0; f < m.nf; f++) gface[fill[gcell[f]]++] = f;
      free(fill);
    }
    long tries = 0;
    for (int j = 0; j < nq;) {
      if (++tries > 1000000000L) { fprintf(stderr, "fbmesh: charge placement failed\n"); return 1; }
      double p[3], d[3];
      for (int k = 0; k < 3; k++) p[k] = rmax * (2.0 * rng_uniform(&rng) - 1.0);
      double rho = sqrt(dot3(p, p));
      if (rho < 1e-12) continue;
      for (int k = 0; k < 3; k++) d[k] = p[k] / rho;
      if (rho > shape_r(&shape, d) - depth * 0.5) continue;  /* outside, or too close radially */
      int ok = 1, pc[3];
      for (int k = 0; k < 3; k++) pc[k] = (int)floor((p[k] + rmax) / cell);
      for (int i0 = pc[0] - 1; i0 <= pc[0] + 1 && ok; i0++)
        for (int i1 = pc[1] - 1; i1 <= pc[1] + 1 && ok; i1++)
          for (int i2 = pc[2] - 1; i2 <= pc[2] + 1 && ok; i2++) {
            if (i0 < 0 || i1 < 0 || i2 < 0 || i0 >= G || i1 >= G || i2 >= G) continue;
            const int cc = (i0 * G + i1) * G + i2;
            for (int q = gstart[cc]; q < gstart[cc + 1] && ok; q++) {
              const int f = gface[q];
              double dx = p[0] - cen[3 * f], dy = p[1] - cen[3 * f + 1], dz = p[2] - cen[3 * f + 2];
              if (dx * dx + dy * dy + dz * dz > reach2) continue;
              if (pt_tri_dist2(p, &pos[3 * m.f[3 * f]], &pos[3 * m.f[3 * f + 1]], &pos[3 * m.f[3 * f + 2]]) < depth2)
                ok = 0;
            }
          }
      if (!ok) continue;
      for (int k = 0; k < 3; k++) qx[3 * j + k] = p[k];
      qq[j] = sigma * rng_normal(&rng);
      j++;
    }
    double s = 0.0;
    for (int j = 0; j < nq; j++) s += qq[j];
    double shift = (nearbyint(s) - s) / nq;
    for (int j = 0; j < nq; j++) qq[j] += shift;
    free(cen); free(gstart); free(gface); free(gcell);
  } else {
    usage();
  }

  FILE *fp = open_out(base, ".vert");
  fprintf(fp, "# fbmesh %s level=%d R=%.17g amp=%.17g lmax=%d seed=%llu\n", kind, level, R, amp,
          shape.lmax, (unsigned long long)seed);
  fprintf(fp, "#faces  #sphere density probe_r\n");
  fprintf(fp, "%d %d %.2f %.2f\n", m.nv, nq, 1.0, 1.4);
  for (int i = 0; i < m.nv; i++)
    fprintf(fp, "%.17g %.17g %.17g %.17g %.17g %.17g 0 %d 2\n", pos[3 * i], pos[3 * i + 1], pos[3 * i + 2],
            nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2], 1);
  fclose(fp);
  fp = open_out(base, ".face");
  fprintf(fp, "# fbmesh %s level=%d\n", kind, level);
  fprintf(fp, "#faces  #sphere density probe_r\n");
  fprintf(fp, "%d %d %.2f %.2f\n", m.nf, nq, 1.0, 1.4);
  for (int f = 0; f < m.nf; f++) fprintf(fp, "%d %d %d 1 1\n", m.f[3 * f] + 1, m.f[3 * f + 1] + 1, m.f[3 * f + 2] + 1);
  fclose(fp);
  fp = open_out(base, ".pqr");
  for (int j = 0; j < nq; j++)
    fprintf(fp, "ATOM %6d  C   CHG A%4d    %.17g %.17g %.17g %.17g 1.5000\n", j + 1, j + 1, qx[3 * j],
            qx[3 * j + 1], qx[3 * j + 2], qq[j]);
  fprintf(fp, "END\n");
  fclose(fp);
  printf("fbmesh: %s vertices=%d faces=%d charges=%d\n", base, m.nv, m.nf, nq);
  return 0;
}
