/*
 * fb_host.c - mesh input, octree, interaction lists and flattened layouts,
 * charge octree.  Host C; its outputs are uploaded once and stay in HBM.
 *
 * The tree must match the reference's exactly (same leaves, same neighbour
 * and interaction lists, same cube centres), because the FMM result depends
 * on them.  So the construction keeps the reference's arithmetic: the same
 * midplane tests, the same enclosing-box centres and radii, and the same
 * separation test, in the same order.  The data structures are flat arrays
 * with integer links rather than the reference's pointer-linked cubes.
 */
#define _POSIX_C_SOURCE 200809L
#include "fb_host.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static char gErr[512];
const char *fb_host_error(void) { return gErr; }
static void setErr(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(gErr, sizeof(gErr), fmt, ap);
  va_end(ap);
}

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) { fprintf(stderr, "fabipb: out of memory (%zu x %zu)\n", n, s); exit(1); }
  return p;
}
#define NEW(T, n) ((T *)xcalloc((size_t)(n), sizeof(T)))
#define SQR(x) ((x) * (x))

/* ================================================================ input */

/* PQR ATOM/HETATM record: the last five fields are x y z charge radius
 * (src/input.c:38-95). Returns 1 parsed, 0 not an atom, -1 malformed. */
static int pqrLine(const char *line, double *v) {
  char copy[512], *tok[32], *save = NULL, *end;
  int nt = 0;
  snprintf(copy, sizeof(copy), "%s", line);
  for (char *t = strtok_r(copy, " \t\r\n", &save); t && nt < 32; t = strtok_r(NULL, " \t\r\n", &save)) tok[nt++] = t;
  if (nt <= 0) return 0;
  if (strcmp(tok[0], "ATOM") && strcmp(tok[0], "HETATM")) return 0;
  if (nt < 10) return -1;
  for (int k = 0; k < 5; k++) {
    v[k] = strtod(tok[nt - 5 + k], &end);
    if (*end) return -1;
  }
  return 1;
}

static int skipLine(FILE *fp) {
  int ch;
  while ((ch = getc(fp)) != '\n')
    if (ch == EOF) return 0;
  return 1;
}

/* Heron's formula, src/input.c:253-267 */
static double heron(double v[3][3]) {
  double a[3], b[3], c[3];
  for (int i = 0; i < 3; i++) { a[i] = v[0][i] - v[1][i]; b[i] = v[0][i] - v[2][i]; c[i] = v[1][i] - v[2][i]; }
  double aa = sqrt(SQR(a[0]) + SQR(a[1]) + SQR(a[2]));
  double bb = sqrt(SQR(b[0]) + SQR(b[1]) + SQR(b[2]));
  double cc = sqrt(SQR(c[0]) + SQR(c[1]) + SQR(c[2]));
  double ss = 0.5 * (aa + bb + cc);
  return sqrt(ss * (ss - aa) * (ss - bb) * (ss - cc));
}

static int geomNormal(const double *v0, const double *v1, const double *v2, double *n) {
  double e1[3], e2[3];
  for (int k = 0; k < 3; k++) { e1[k] = v1[k] - v0[k]; e2[k] = v2[k] - v0[k]; }
  n[0] = e1[1] * e2[2] - e1[2] * e2[1];
  n[1] = e1[2] * e2[0] - e1[0] * e2[2];
  n[2] = e1[0] * e2[1] - e1[1] * e2[0];
  double len = sqrt(SQR(n[0]) + SQR(n[1]) + SQR(n[2]));
  if (len <= 0.0) return 0;
  for (int k = 0; k < 3; k++) n[k] /= len;
  return 1;
}

void fb_mesh_free(fb_mesh *m) {
  if (!m) return;
  free(m->vtx); free(m->nrm); free(m->area); free(m->face); free(m->pos); free(m->chr); free(m);
}

/* The rest of fp from its current position, NUL-terminated; *len bytes. */
static char *slurp(FILE *fp, size_t *len) {
  long at = ftell(fp);
  if (at < 0 || fseek(fp, 0, SEEK_END) != 0) return NULL;
  long end = ftell(fp);
  if (end < at || fseek(fp, at, SEEK_SET) != 0) return NULL;
  size_t n = (size_t)(end - at);
  char *b = malloc(n + 1);
  if (!b) return NULL;
  if (fread(b, 1, n, fp) != n) { free(b); return NULL; }
  b[n] = 0;
  *len = n;
  return b;
}

/* starts of the first `want` lines of b (want + 1 entries: the last is the end
 * of line want - 1, past its '\n'); 0 if b has fewer lines */
static int lineStarts(const char *b, size_t len, int want, size_t *st) {
  size_t at = 0;
  for (int i = 0; i < want; i++) {
    if (at >= len) return 0;
    st[i] = at;
    const char *nl = memchr(b + at, '\n', len - at);
    at = nl ? (size_t)(nl - b) + 1 : len;
  }
  st[want] = at;
  return 1;
}

/* the whitespace-separated tokens of b[s, e): doubles first (nd of them, each
 * consumed whole by strtod), then plain integers (ni: [+-]digits, by strtol);
 * 1 if the line is exactly that, which is where fscanf's token stream and this
 * per-line reading agree */
static int lineTokens(const char *b, size_t s, size_t e, int nd, double *d, int ni, long *iv) {
  const char *p = b + s, *end = b + e;
  for (int k = 0; k < nd + ni; k++) {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) p++;
    const char *t = p;
    while (p < end && !(*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) p++;
    if (p == t) return 0;
    char tok[128];
    if ((size_t)(p - t) >= sizeof(tok)) return 0;
    memcpy(tok, t, (size_t)(p - t));
    tok[p - t] = 0;
    char *q;
    if (k < nd) {
      d[k] = strtod(tok, &q);
    } else {
      const char *c = tok + (tok[0] == '+' || tok[0] == '-');
      if (!*c) return 0;
      for (; *c; c++)
        if (*c < '0' || *c > '9') return 0;
      iv[k - nd] = strtol(tok, &q, 10);
    }
    if (*q) return 0;
  }
  while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) p++;
  return p == end;
}

/* src/input.c:320-821 (loadPanel) with the mesh taken from existing files */
fb_mesh *fb_mesh_load(const char *base, int useGeom) {
  char path[4096], line[512];
  FILE *fp;
  double v[5];
  fb_mesh *m = NEW(fb_mesh, 1);

  snprintf(path, sizeof(path), "%s.pqr", base);
  if (!(fp = fopen(path, "r"))) { setErr("cannot open PQR file '%s'", path); fb_mesh_free(m); return NULL; }
  int cap = 1024;
  m->pos = NEW(double, 3 * cap);
  m->chr = NEW(double, cap);
  /* every line parsed on its own, in parallel, when no line is long enough
   * for fgets to split it; the records kept in file order either way */
  size_t blen = 0, nl = 0, scap = 0;
  size_t *st = NULL;
  char *buf = slurp(fp, &blen);
  int par = buf != NULL;
  for (size_t at = 0; par && at < blen;) {   /* line starts; no line longer than fgets reads at once */
    const char *e = memchr(buf + at, '\n', blen - at);
    size_t next = e ? (size_t)(e - buf) + 1 : blen;
    if (next - at > sizeof(line) - 1) par = 0;
    if (nl == scap) {
      scap = scap ? 2 * scap : 4096;
      st = realloc(st, sizeof(size_t) * (scap + 1));
    }
    st[nl++] = at;
    at = next;
  }
  if (par) {
    st[nl] = blen;
    signed char *ok = NEW(signed char, nl + 1);
    double *vv = NEW(double, 5 * (nl + 1));
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)nl; i++) {
      char l2[512];
      size_t w = st[i + 1] - st[i];
      memcpy(l2, buf + st[i], w);
      l2[w] = 0;
      ok[i] = (signed char)pqrLine(l2, &vv[5 * i]);
    }
    for (size_t i = 0; i < nl; i++) {
      if (ok[i] != 1) continue;
      if (m->nChar == cap) {
        cap *= 2;
        m->pos = realloc(m->pos, sizeof(double) * 3 * cap);
        m->chr = realloc(m->chr, sizeof(double) * cap);
      }
      m->pos[3 * m->nChar] = vv[5 * i]; m->pos[3 * m->nChar + 1] = vv[5 * i + 1]; m->pos[3 * m->nChar + 2] = vv[5 * i + 2];
      m->chr[m->nChar] = vv[5 * i + 3];
      m->nChar++;
    }
    free(ok); free(vv);
  } else {
    rewind(fp);
    while (fgets(line, sizeof(line), fp)) {
      if (pqrLine(line, v) != 1) continue;
      if (m->nChar == cap) {
        cap *= 2;
        m->pos = realloc(m->pos, sizeof(double) * 3 * cap);
        m->chr = realloc(m->chr, sizeof(double) * cap);
      }
      m->pos[3 * m->nChar] = v[0]; m->pos[3 * m->nChar + 1] = v[1]; m->pos[3 * m->nChar + 2] = v[2];
      m->chr[m->nChar] = v[3];
      m->nChar++;
    }
  }
  free(st); free(buf);
  fclose(fp);
  if (m->nChar <= 0) { setErr("no ATOM/HETATM records in '%s'", path); fb_mesh_free(m); return NULL; }

  snprintf(path, sizeof(path), "%s.vert", base);
  if (!(fp = fopen(path, "r"))) { setErr("cannot open vertices file '%s'", path); fb_mesh_free(m); return NULL; }
  int nspt = 0, natm, ierr;
  double den, prb;
  if (!skipLine(fp) || !skipLine(fp) || fscanf(fp, "%d %d %lf %lf ", &nspt, &natm, &den, &prb) != 4 || nspt <= 0) {
    setErr("malformed vertex header in '%s'", path); fclose(fp); fb_mesh_free(m); return NULL;
  }
  double *sp = NEW(double, 3 * (size_t)nspt), *sn = NEW(double, 3 * (size_t)nspt);
  {
    /* one vertex per line (6 doubles, 3 integers) read in parallel; any other
     * layout goes through fscanf's token stream as before */
    long at = ftell(fp);
    size_t blen = 0;
    char *buf = slurp(fp, &blen);
    size_t *st = buf ? NEW(size_t, (size_t)nspt + 1) : NULL;
    int par = buf && lineStarts(buf, blen, nspt, st), bad = 0;
    if (par) {
#pragma omp parallel for schedule(static) reduction(|:bad)
      for (long i = 0; i < nspt; i++) {
        double a[6];
        long iv[3];
        if (!lineTokens(buf, st[i], st[i + 1], 6, a, 3, iv)) { bad |= 1; continue; }
        for (int k = 0; k < 3; k++) { sp[3 * i + k] = a[k]; sn[3 * i + k] = a[3 + k]; }
      }
    }
    free(st); free(buf);
    if (!par || bad) {
      fseek(fp, at, SEEK_SET);
      for (int i = 0; i < nspt; i++) {
        double a[6];
        int i1, i2, i3;
        ierr = fscanf(fp, "%lf %lf %lf %lf %lf %lf %d %d %d", &a[0], &a[1], &a[2], &a[3], &a[4], &a[5], &i1, &i2, &i3);
        if (ierr < 6) { setErr("'%s' truncated at vertex %d", path, i + 1); fclose(fp); free(sp); free(sn); fb_mesh_free(m); return NULL; }
        for (int k = 0; k < 3; k++) { sp[3 * i + k] = a[k]; sn[3 * i + k] = a[3 + k]; }
      }
    }
  }
  fclose(fp);

  snprintf(path, sizeof(path), "%s.face", base);
  if (!(fp = fopen(path, "r"))) { setErr("cannot open faces file '%s'", path); free(sp); free(sn); fb_mesh_free(m); return NULL; }
  int nface = 0;
  if (!skipLine(fp) || !skipLine(fp) || fscanf(fp, "%d %d %lf %lf ", &nface, &natm, &den, &prb) != 4 || nface <= 0) {
    setErr("malformed face header in '%s'", path); fclose(fp); free(sp); free(sn); fb_mesh_free(m); return NULL;
  }
  int *nv = NEW(int, 3 * (size_t)nface);
  {
    /* one face per line (5 integers, vertex numbers in range) in parallel; any
     * other layout, or a bad face, goes through fscanf as before */
    long at = ftell(fp);
    size_t blen = 0;
    char *buf = slurp(fp, &blen);
    size_t *st = buf ? NEW(size_t, (size_t)nface + 1) : NULL;
    int par = buf && lineStarts(buf, blen, nface, st), bad = 0;
    if (par) {
#pragma omp parallel for schedule(static) reduction(|:bad)
      for (long i = 0; i < nface; i++) {
        long iv[5];
        if (!lineTokens(buf, st[i], st[i + 1], 0, NULL, 5, iv) || iv[0] < 1 || iv[0] > nspt || iv[1] < 1 ||
            iv[1] > nspt || iv[2] < 1 || iv[2] > nspt) { bad |= 1; continue; }
        nv[3 * i] = (int)iv[0]; nv[3 * i + 1] = (int)iv[1]; nv[3 * i + 2] = (int)iv[2];
      }
    }
    free(st); free(buf);
    if (!par || bad) {
      fseek(fp, at, SEEK_SET);
      for (int i = 0; i < nface; i++) {
        int j1, j2, j3, i1, i2;
        ierr = fscanf(fp, "%d %d %d %d %d", &j1, &j2, &j3, &i1, &i2);
        if (ierr < 3 || j1 < 1 || j1 > nspt || j2 < 1 || j2 > nspt || j3 < 1 || j3 > nspt) {
          setErr("'%s' bad or truncated face %d", path, i + 1); fclose(fp); free(sp); free(sn); free(nv); fb_mesh_free(m); return NULL;
        }
        nv[3 * i] = j1; nv[3 * i + 1] = j2; nv[3 * i + 2] = j3;
      }
    }
  }
  fclose(fp);
  m->nVert = nspt;
  m->nFace = nface;

  /* face filter (src/input.c:686-725): drop area < 1e-5 and centroids within
   * squared distance 1e-5 of any of the 10 preceding faces (the window only
   * applies from the 11th face on, as there) */
  m->vtx = NEW(double, 9 * (size_t)nface);
  m->nrm = NEW(double, 3 * (size_t)nface);
  m->area = NEW(double, nface);
  m->face = NEW(int, nface);
  /* each face's outcome depends on the input alone (its area and the 10
   * faces before it in the file, kept or not), so the faces are decided in
   * parallel, then numbered in file order and written in parallel */
  double *far = NEW(double, nface);
  int *slot = NEW(int, (size_t)nface + 1);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < nface; i++) {
    double face[3][3], xx[3] = {0, 0, 0};
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) {
        face[k][j] = sp[3 * (nv[3 * i + k] - 1) + j];
        xx[j] += 1.0 / 3.0 * (face[k][j]);
      }
    double ar = heron(face);
    int alert = 0;
    for (int j = i - 10; j >= 0 && j < i; j++) {
      double yy[3] = {0, 0, 0};
      for (int k = 0; k < 3; k++)
        for (int ii = 0; ii < 3; ii++) yy[k] += 1.0 / 3.0 * (sp[3 * (nv[3 * j + ii] - 1) + k]);
      double d = 0.0;
      for (int k = 0; k < 3; k++) d += (xx[k] - yy[k]) * (xx[k] - yy[k]);
      if (d < 1.0e-5) alert = 1;
    }
    far[i] = ar;
    slot[i] = ar >= 1.0e-5 && alert == 0;
  }
  for (int i = 0; i < nface; i++) {   /* kept faces numbered in file order */
    int keep = slot[i];
    slot[i] = keep ? m->n++ : -1;
  }
#pragma omp parallel for schedule(static)
  for (long i = 0; i < nface; i++) {
    if (slot[i] < 0) continue;
    int p = slot[i];
    double ar = far[i];
    double vn[3][3];
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) {
        m->vtx[9 * p + 3 * j + k] = sp[3 * (nv[3 * i + j] - 1) + k];
        vn[j][k] = sn[3 * (nv[3 * i + j] - 1) + k];
      }
    double *nr = &m->nrm[3 * p];
    /* panel normal from the vertex normals, src/input.c:750-786 */
    double w[3];
    for (int k = 0; k < 3; k++) w[k] = vn[0][k] / 2. + (vn[1][k] + vn[2][k]) / 4.;
    double len = sqrt(SQR(w[0]) + SQR(w[1]) + SQR(w[2]));
    if (useGeom) {
      double g[3];
      if (len > 0.0)
        for (int k = 0; k < 3; k++) w[k] /= len;
      if (!geomNormal(&m->vtx[9 * p], &m->vtx[9 * p + 3], &m->vtx[9 * p + 6], g))
        for (int k = 0; k < 3; k++) g[k] = w[k];
      if (g[0] * w[0] + g[1] * w[1] + g[2] * w[2] < 0.0)
        for (int k = 0; k < 3; k++) g[k] = -g[k];
      for (int k = 0; k < 3; k++) nr[k] = g[k];
    } else if (len > 0.0) {
      for (int k = 0; k < 3; k++) nr[k] = w[k] / len;
    } else if (!geomNormal(&m->vtx[9 * p], &m->vtx[9 * p + 3], &m->vtx[9 * p + 6], nr)) {
      nr[0] = 0.0; nr[1] = 0.0; nr[2] = 1.0;
    }
    m->area[p] = ar;
    m->face[p] = (int)i;
  }
  free(far); free(slot);
  free(sp); free(sn); free(nv);
  if (m->n <= 0) { setErr("no panels survived the face filter"); fb_mesh_free(m); return NULL; }
  return m;
}

/* ================================================================ octree */

typedef struct {
  int level, i, j, k;
  int kids[8], nKids, parent, nPnls;
  int head, tail;           /* panel list (leaves), or charge count (charge tree) */
  int next;                 /* next cube at the same level (DFS order) */
  double x[3], lo[3], up[3], eRad;
  int nNbrs, n2Nbrs, *nbrs;
  int flat;                 /* FMM index, or charge-tree index */
} cube_t;

typedef struct {
  cube_t *c;
  int nc, cap;
  double minx, miny, minz, *side;
  int maxLev;
  int *first, *last;        /* per level */
} otree_t;

static int newCube(otree_t *T, int level, int i, int j, int k, int parent) {
  if (T->nc == T->cap) {
    T->cap = T->cap ? 2 * T->cap : 1024;
    T->c = realloc(T->c, sizeof(cube_t) * (size_t)T->cap);
  }
  cube_t *c = &T->c[T->nc];
  memset(c, 0, sizeof(*c));
  c->level = level; c->i = i; c->j = j; c->k = k; c->parent = parent;
  for (int q = 0; q < 8; q++) c->kids[q] = -1;
  c->head = c->tail = -1; c->next = -1; c->flat = -1;
  return T->nc++;
}

/* src/gkSetup.c:142-183 (goDown): strict '>' against the cube midplanes */
static int goDown(otree_t *T, const double *p) {
  int cb = 0;
  while (T->c[cb].level < T->maxLev) {
    cube_t *c = &T->c[cb];
    int ix = 0, iy = 0, iz = 0;
    if (p[0] > T->minx + (c->i + 0.5) * T->side[c->level]) ix = 1;
    if (p[1] > T->miny + (c->j + 0.5) * T->side[c->level]) iy = 1;
    if (p[2] > T->minz + (c->k + 0.5) * T->side[c->level]) iz = 1;
    int q = ix + 2 * iy + 4 * iz;
    if (c->kids[q] < 0) {
      int lev = c->level + 1, ci = 2 * c->i + ix, cj = 2 * c->j + iy, ck = 2 * c->k + iz;
      int kid = newCube(T, lev, ci, cj, ck, cb);
      T->c[cb].kids[q] = kid;   /* T->c may have moved */
    }
    cb = T->c[cb].kids[q];
  }
  return cb;
}

/* src/gkSetup.c:194-236 (linkcubes): per-level DFS lists, compacted kids */
static void linkCubes(otree_t *T, int cb) {
  /* explicit stack: pre-order, kids in compacted order */
  int *stack = NEW(int, T->nc + 1), sp = 0;
  stack[sp++] = cb;
  while (sp) {
    int c = stack[--sp];
    cube_t *C = &T->c[c];
    int lev = C->level;
    if (T->last[lev] < 0) T->first[lev] = T->last[lev] = c;
    else { T->c[T->last[lev]].next = c; T->last[lev] = c; }
    if (lev < T->maxLev) {
      int q2 = 0;
      for (int q = 0; q < 8; q++)
        if (C->kids[q] >= 0) C->kids[q2++] = C->kids[q];
      C->nKids = q2;
      for (int q = q2; q < 8; q++) C->kids[q] = -1;
      for (int q = C->nKids - 1; q >= 0; q--) stack[sp++] = C->kids[q];
    }
  }
  free(stack);
}

/* leaf counts propagate up; a leaf's count is already set */
static void countUp(otree_t *T) {
  for (int lev = T->maxLev - 1; lev >= 0; lev--)
    for (int c = T->first[lev]; c >= 0; c = T->c[c].next) {
      T->c[c].nPnls = 0;
      for (int q = 0; q < T->c[c].nKids; q++) T->c[c].nPnls += T->c[T->c[c].kids[q]].nPnls;
    }
}

/* parents' boxes from their kids', src/gkSetup.c:303-321 */
static void parentBoxes(otree_t *T, int height) {
  for (int lev = T->maxLev - 1; lev >= height; lev--)
    for (int c = T->first[lev]; c >= 0; c = T->c[c].next) {
      cube_t *C = &T->c[c];
      cube_t *K = &T->c[C->kids[0]];
      for (int k = 0; k < 3; k++) { C->up[k] = K->up[k]; C->lo[k] = K->lo[k]; }
      for (int q = 1; q < C->nKids; q++) {
        K = &T->c[C->kids[q]];
        for (int k = 0; k < 3; k++) {
          C->up[k] = K->up[k] > C->up[k] ? K->up[k] : C->up[k];
          C->lo[k] = K->lo[k] < C->lo[k] ? K->lo[k] : C->lo[k];
        }
      }
      for (int k = 0; k < 3; k++) C->x[k] = 0.5 * (C->lo[k] + C->up[k]);
      C->eRad = 0.5 * sqrt(SQR(C->up[0] - C->lo[0]) + SQR(C->up[1] - C->lo[1]) + SQR(C->up[2] - C->lo[2]));
    }
}

/* src/gkSetup.c:332-349 (near): 2 same, 1 neighbours, 0 well separated */
static int nearCubes(const cube_t *a, const cube_t *b, double eta0) {
  if (a == b) return 2;
  double d = SQR(a->x[0] - b->x[0]) + SQR(a->x[1] - b->x[1]) + SQR(a->x[2] - b->x[2]);
  d = sqrt(d);
  if (d < 1e-15) return 1;
  return ((a->eRad + b->eRad) / d < eta0) ? 0 : 1;
}

/* src/gkSetup.c:363-437 (getNbrs): nbrs[0] = self, [1,nNbrs) near,
 * [nNbrs,n2Nbrs) interaction list */
/* the cubes of level lev in list order; *cnt of them */
static int *levelCubes(const otree_t *T, int lev, int *cnt) {
  int n = 0;
  for (int c = T->first[lev]; c >= 0; c = T->c[c].next) n++;
  int *a = NEW(int, n + 1);
  n = 0;
  for (int c = T->first[lev]; c >= 0; c = T->c[c].next) a[n++] = c;
  *cnt = n;
  return a;
}

/* each cube's lists depend only on its parent's (the level above), so the
 * cubes of a level are done in parallel */
static void neighbours(otree_t *T, int height, int depth, double eta0) {
  int nTop = 0;
  int *top = levelCubes(T, height, &nTop);
#pragma omp parallel for schedule(dynamic, 16)
  for (int ci = 0; ci < nTop; ci++) {
    int c = top[ci];
    cube_t *C = &T->c[c];
    int nN = 0;
    for (int b = T->first[height]; b >= 0; b = T->c[b].next) if (nearCubes(C, &T->c[b], eta0)) nN++;
    C->nbrs = NEW(int, nTop);
    int i1 = 1, i2 = nN;
    for (int b = T->first[height]; b >= 0; b = T->c[b].next) {
      int f = nearCubes(C, &T->c[b], eta0);
      if (f == 1) C->nbrs[i1++] = b;
      else if (f == 2) C->nbrs[0] = b;
      else C->nbrs[i2++] = b;
    }
    C->nNbrs = i1; C->n2Nbrs = i2;
  }
  free(top);
  for (int lev = height + 1; lev <= depth; lev++) {
    int nl = 0;
    int *cubes = levelCubes(T, lev, &nl);
#pragma omp parallel for schedule(dynamic, 64)
    for (int ci = 0; ci < nl; ci++) {
      int c = cubes[ci];
      cube_t *C = &T->c[c];
      cube_t *P = &T->c[C->parent];
      int nN = 0, n2 = 0;
      for (int q = 0; q < P->nNbrs; q++) {
        cube_t *NP = &T->c[P->nbrs[q]];
        for (int k = 0; k < NP->nKids; k++) {
          if (nearCubes(&T->c[NP->kids[k]], C, eta0)) nN++;
          n2++;
        }
      }
      C->nbrs = NEW(int, n2);
      int i1 = 1, i2 = nN;
      for (int q = 0; q < P->nNbrs; q++) {
        cube_t *NP = &T->c[P->nbrs[q]];
        for (int k = 0; k < NP->nKids; k++) {
          int b = NP->kids[k];
          int f = nearCubes(C, &T->c[b], eta0);
          if (f == 1) C->nbrs[i1++] = b;
          else if (f == 2) C->nbrs[0] = b;
          else C->nbrs[i2++] = b;
        }
      }
      C->nNbrs = i1; C->n2Nbrs = i2;
    }
    free(cubes);
  }
}

static void otreeFree(otree_t *T) {
  for (int c = 0; c < T->nc; c++) free(T->c[c].nbrs);
  free(T->c); free(T->side); free(T->first); free(T->last);
}

void fb_tree_free(fb_tree *t) {
  if (!t) return;
  free(t->ordMom); free(t->ordM2L); free(t->perm); free(t->vtx); free(t->nrm); free(t->area); free(t->face);
  free(t->levStart); free(t->levCount); free(t->cubeLevel); free(t->kidStart); free(t->nKids); free(t->parent);
  free(t->cubeX); free(t->eRad); free(t->leafStart); free(t->leafCount); free(t->nearOff); free(t->nearSrc);
  free(t->m2lSrc); free(t->m2lDst); free(t->m2lOff);
  free(t);
}

static fb_tree *finishTree(otree_t T, fb_tree *t, int depth, int height, int order, int orderMom, double eta0);

/* src/gkSetup.c:446-508 (gkInit) and src/fmm.c:234-524 (setupFMM layouts) */
fb_tree *fb_tree_build(const fb_mesh *m, int depth, int height, int order, int orderMom, double eta0) {
  if (depth < 1 || height < 0 || height > depth || order == 0) {
    setErr("bad tree parameters depth=%d height=%d order=%d", depth, height, order);
    return NULL;
  }
  int n = m->n;
  otree_t T;
  memset(&T, 0, sizeof(T));
  T.maxLev = depth;

  /* bounding box and level side lengths, src/gkSetup.c:90-128 */
  double maxx, maxy, maxz;
  T.minx = maxx = m->vtx[0]; T.miny = maxy = m->vtx[1]; T.minz = maxz = m->vtx[2];
  double *cen = NEW(double, 3 * (size_t)n);
  for (int p = 0; p < n; p++) {
    const double *v = &m->vtx[9 * p];
    for (int k = 0; k < 3; k++) {
      /* MAX(v, max) = v > max ? v : max; MIN(v, min) = v > min ? min : v (gkGlobal.h:33-38) */
      maxx = v[3 * k] > maxx ? v[3 * k] : maxx; T.minx = v[3 * k] > T.minx ? T.minx : v[3 * k];
      maxy = v[3 * k + 1] > maxy ? v[3 * k + 1] : maxy; T.miny = v[3 * k + 1] > T.miny ? T.miny : v[3 * k + 1];
      maxz = v[3 * k + 2] > maxz ? v[3 * k + 2] : maxz; T.minz = v[3 * k + 2] > T.minz ? T.minz : v[3 * k + 2];
    }
    for (int k = 0; k < 3; k++) cen[3 * p + k] = 0.33333333333333333333333333 * (v[k] + v[3 + k] + v[6 + k]);
  }
  T.side = NEW(double, depth + 1);
  double len = (maxx - T.minx) > (maxy - T.miny) ? (maxx - T.minx) : (maxy - T.miny);
  len = (maxz - T.minz) > len ? (maxz - T.minz) : len;
  for (int lev = 0; lev <= depth; lev++) { T.side[lev] = len; len *= 0.5; }

  /* insert panels (input order) into leaves */
  int *nextC = NEW(int, n);
  newCube(&T, 0, 0, 0, 0, -1);
  for (int p = 0; p < n; p++) {
    int leaf = goDown(&T, &cen[3 * p]);
    cube_t *L = &T.c[leaf];
    nextC[p] = -1;
    if (L->head < 0) L->head = p; else nextC[L->tail] = p;
    L->tail = p;
    L->nPnls++;
  }
  free(cen);
  T.first = NEW(int, depth + 1);
  T.last = NEW(int, depth + 1);
  for (int l = 0; l <= depth; l++) T.first[l] = T.last[l] = -1;
  linkCubes(&T, 0);
  countUp(&T);

  fb_tree *t = NEW(fb_tree, 1);
  t->n = n; t->depth = depth; t->height = height; t->nChar = m->nChar;
  t->perm = NEW(int, n);
  int idx = 0, nLeaf = 0;
  for (int c = T.first[depth]; c >= 0; c = T.c[c].next) {
    nLeaf++;
    for (int p = T.c[c].head; p >= 0; p = nextC[p]) t->perm[idx++] = p;
  }
  free(nextC);

  /* leaf boxes over their panels' vertices (src/gkSetup.c:264-301), idx order */
  {
    int i0 = 0;
    for (int c = T.first[depth]; c >= 0; c = T.c[c].next) {
      cube_t *C = &T.c[c];
      const double *v0 = &m->vtx[9 * t->perm[i0]];
      for (int k = 0; k < 3; k++) C->up[k] = C->lo[k] = v0[k];
      for (int q = 0; q < C->nPnls; q++) {
        const double *v = &m->vtx[9 * t->perm[i0 + q]];
        for (int j = 0; j < 3; j++)
          for (int k = 0; k < 3; k++) {
            double xv = v[3 * j + k];
            C->up[k] = xv > C->up[k] ? xv : C->up[k];
            C->lo[k] = xv < C->lo[k] ? xv : C->lo[k];
          }
      }
      for (int k = 0; k < 3; k++) C->x[k] = 0.5 * (C->lo[k] + C->up[k]);
      C->eRad = 0.5 * sqrt(SQR(C->up[0] - C->lo[0]) + SQR(C->up[1] - C->lo[1]) + SQR(C->up[2] - C->lo[2]));
      i0 += C->nPnls;
    }
  }
  /* panels in idx order */
  t->vtx = NEW(double, 9 * (size_t)n); t->nrm = NEW(double, 3 * (size_t)n);
  t->area = NEW(double, n); t->face = NEW(int, n);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    int p = t->perm[i];
    memcpy(&t->vtx[9 * i], &m->vtx[9 * p], 9 * sizeof(double));
    memcpy(&t->nrm[3 * i], &m->nrm[3 * p], 3 * sizeof(double));
    t->area[i] = m->area[p];
    t->face[i] = m->face[p];
  }
  return finishTree(T, t, depth, height, order, orderMom, eta0);
}

/* The same tree from its leaves, as a device pass finds them (fbh_tree_panels):
 * panels sorted by leaf (the octant digits of each level, coarsest first: the
 * leaves' depth-first order) and, within a leaf, in input order (perm), the
 * panel arrays in that order, and per leaf its digit key, panel count and
 * vertex box (lo, up).  Cubes, links, boxes and lists follow as in
 * fb_tree_build. */
fb_tree *fb_tree_build_leaves(const fb_mesh *m, int depth, int height, int order, int orderMom, double eta0,
                              const int *perm, const double *vtx, const double *nrm, const double *area,
                              const int *face, int nLeaf, const unsigned *leafKey, const int *leafCount,
                              const double *leafBox) {
  if (depth < 1 || depth > 10 || height < 0 || height > depth || order == 0 || nLeaf < 1) {
    setErr("bad tree parameters depth=%d height=%d order=%d leaves=%d", depth, height, order, nLeaf);
    return NULL;
  }
  int n = m->n;
  otree_t T;
  memset(&T, 0, sizeof(T));
  T.maxLev = depth;
  T.side = NEW(double, depth + 1);
  newCube(&T, 0, 0, 0, 0, -1);
  long long total = 0;
  for (int L = 0; L < nLeaf; L++) {
    if (L > 0 && leafKey[L] <= leafKey[L - 1]) { setErr("leaf keys not increasing at %d", L); otreeFree(&T); return NULL; }
    int cb = 0;
    for (int lev = 1; lev <= depth; lev++) {
      int q = (int)((leafKey[L] >> (3 * (depth - lev))) & 7u);
      if (T.c[cb].kids[q] < 0) {
        cube_t *c = &T.c[cb];
        int ci = 2 * c->i + (q & 1), cj = 2 * c->j + ((q >> 1) & 1), ck = 2 * c->k + ((q >> 2) & 1);
        int kid = newCube(&T, lev, ci, cj, ck, cb);
        T.c[cb].kids[q] = kid;   /* T.c may have moved */
      }
      cb = T.c[cb].kids[q];
    }
    cube_t *C = &T.c[cb];
    C->nPnls = leafCount[L];
    total += leafCount[L];
    for (int k = 0; k < 3; k++) { C->lo[k] = leafBox[6 * L + k]; C->up[k] = leafBox[6 * L + 3 + k]; }
    for (int k = 0; k < 3; k++) C->x[k] = 0.5 * (C->lo[k] + C->up[k]);
    C->eRad = 0.5 * sqrt(SQR(C->up[0] - C->lo[0]) + SQR(C->up[1] - C->lo[1]) + SQR(C->up[2] - C->lo[2]));
  }
  if (total != n) { setErr("leaf counts add to %lld, not %d panels", total, n); otreeFree(&T); return NULL; }
  T.first = NEW(int, depth + 1);
  T.last = NEW(int, depth + 1);
  for (int l = 0; l <= depth; l++) T.first[l] = T.last[l] = -1;
  linkCubes(&T, 0);
  countUp(&T);
  fb_tree *t = NEW(fb_tree, 1);
  t->n = n; t->depth = depth; t->height = height; t->nChar = m->nChar;
  t->perm = NEW(int, n);
  t->vtx = NEW(double, 9 * (size_t)n); t->nrm = NEW(double, 3 * (size_t)n);
  t->area = NEW(double, n); t->face = NEW(int, n);
  memcpy(t->perm, perm, (size_t)n * sizeof(int));
  memcpy(t->vtx, vtx, 9 * (size_t)n * sizeof(double));
  memcpy(t->nrm, nrm, 3 * (size_t)n * sizeof(double));
  memcpy(t->area, area, (size_t)n * sizeof(double));
  memcpy(t->face, face, (size_t)n * sizeof(int));
  return finishTree(T, t, depth, height, order, orderMom, eta0);
}

/* boxes above the leaves, lists, orders and the flat layouts; frees T */
static fb_tree *finishTree(otree_t T, fb_tree *t, int depth, int height, int order, int orderMom, double eta0) {
  int idx, nLeaf = 0;
  for (int c = T.first[depth]; c >= 0; c = T.c[c].next) nLeaf++;
  parentBoxes(&T, height);
  neighbours(&T, height, depth, eta0);

  /* orders, src/gkSetup.c:42-81 */
  t->ordMom = NEW(int, depth + 1);
  t->ordM2L = NEW(int, depth + 1);
  if (order > 0) {
    for (int l = 0; l <= depth; l++) t->ordM2L[l] = t->ordMom[l] = order;
    t->maxOrder = order;
  } else {
    t->ordM2L[depth] = -order;
    for (int l = depth - 1; l >= height; l--) t->ordM2L[l] = t->ordM2L[l + 1] + 1;
    t->maxOrder = t->ordM2L[height];
    for (int l = depth; l >= height; l--) {
      int o = orderMom >= 0 ? t->ordM2L[l] + orderMom : t->maxOrder;
      t->ordMom[l] = o < t->maxOrder ? o : t->maxOrder;
    }
  }

  /* FMM cube layout, src/fmm.c:295-324 */
  int nCube = 0;
  for (int l = depth; l >= height; l--)
    for (int c = T.first[l]; c >= 0; c = T.c[c].next) nCube++;
  t->nCube = nCube;
  t->levStart = NEW(int, depth + 1); t->levCount = NEW(int, depth + 1);
  t->cubeLevel = NEW(int, nCube); t->kidStart = NEW(int, nCube); t->nKids = NEW(int, nCube);
  t->parent = NEW(int, nCube); t->cubeX = NEW(double, 3 * (size_t)nCube); t->eRad = NEW(double, nCube);
  idx = 0;
  for (int l = depth; l >= height; l--) {
    t->levStart[l] = idx;
    for (int c = T.first[l]; c >= 0; c = T.c[c].next, idx++) T.c[c].flat = idx;
    t->levCount[l] = idx - t->levStart[l];
  }
  for (int l = depth; l >= height; l--)
    for (int c = T.first[l]; c >= 0; c = T.c[c].next) {
      cube_t *C = &T.c[c];
      int f = C->flat;
      t->cubeLevel[f] = l;
      t->nKids[f] = l < depth ? C->nKids : 0;
      t->kidStart[f] = (l < depth && C->nKids) ? T.c[C->kids[0]].flat : -1;
      t->parent[f] = (l > height) ? T.c[C->parent].flat : -1;
      for (int k = 0; k < 3; k++) t->cubeX[3 * f + k] = C->x[k];
      t->eRad[f] = C->eRad;
    }

  /* leaves and near pairs, src/fmm.c:234-268 */
  t->nLeaf = nLeaf;
  t->leafStart = NEW(int, nLeaf); t->leafCount = NEW(int, nLeaf);
  t->nearOff = NEW(int, nLeaf + 1);
  int nNear = 0, i0 = 0;
  for (int c = T.first[depth]; c >= 0; c = T.c[c].next) {
    int f = T.c[c].flat;
    t->leafStart[f] = i0;
    t->leafCount[f] = T.c[c].nPnls;
    if (T.c[c].nPnls > t->maxLeafPanels) t->maxLeafPanels = T.c[c].nPnls;
    i0 += T.c[c].nPnls;
    nNear += T.c[c].nNbrs;
  }
  t->nNear = nNear;
  t->nearSrc = NEW(int, nNear);
  idx = 0;
  for (int c = T.first[depth]; c >= 0; c = T.c[c].next) {
    t->nearOff[T.c[c].flat] = idx;
    for (int q = 0; q < T.c[c].nNbrs; q++) t->nearSrc[idx++] = T.c[T.c[c].nbrs[q]].flat;
  }
  t->nearOff[nLeaf] = idx;

  /* M2L pairs, src/fmm.c:326-404: destination-major, sources in reverse list order */
  int nM2L = 0;
  for (int l = depth; l >= height; l--)
    for (int c = T.first[l]; c >= 0; c = T.c[c].next) nM2L += T.c[c].n2Nbrs - T.c[c].nNbrs;
  t->nM2L = nM2L;
  t->m2lSrc = NEW(int, nM2L); t->m2lDst = NEW(int, nM2L); t->m2lOff = NEW(int, nCube + 1);
  idx = 0;
  for (int l = depth; l >= height; l--)
    for (int c = T.first[l]; c >= 0; c = T.c[c].next) {
      cube_t *C = &T.c[c];
      t->m2lOff[C->flat] = idx;
      if (C->n2Nbrs > C->nNbrs) t->nM2LGroups++;
      for (int q = C->n2Nbrs - 1; q >= C->nNbrs; q--, idx++) {
        t->m2lSrc[idx] = T.c[C->nbrs[q]].flat;
        t->m2lDst[idx] = C->flat;
      }
    }
  t->m2lOff[nCube] = idx;

  /* block-preconditioner cluster stats at level depth-1 (src/precond_fmm.c:553-651) */
  if (depth - 1 >= 0)
    for (int c = T.first[depth - 1]; c >= 0; c = T.c[c].next) {
      t->nPrecondBlocks++;
      if (2 * T.c[c].nPnls > t->maxPrecondBlock) t->maxPrecondBlock = 2 * T.c[c].nPnls;
    }
  otreeFree(&T);
  return t;
}

/* ============================================================ charge tree */

void fb_ctree_free(fb_ctree *t) {
  if (!t) return;
  free(t->levStart); free(t->levCount); free(t->level); free(t->kidStart); free(t->nKids);
  free(t->x); free(t->eRad); free(t->chgStart); free(t->nChgs); free(t->chgIdx); free(t);
}

/* src/chargeTree.c:228-307 (buildChargeTree) with :136-226 (linkChgCubes, getChgEnclBoxs) */
fb_ctree *fb_ctree_build(const fb_mesh *m, int depth) {
  otree_t T;
  memset(&T, 0, sizeof(T));
  T.maxLev = depth;
  const double *pos = m->pos;
  double cmaxx, cmaxy, cmaxz;
  T.minx = cmaxx = pos[0]; T.miny = cmaxy = pos[1]; T.minz = cmaxz = pos[2];
  for (int j = 1; j < m->nChar; j++) {
    double x = pos[3 * j], y = pos[3 * j + 1], z = pos[3 * j + 2];
    T.minx = T.minx > x ? x : T.minx; cmaxx = cmaxx > x ? cmaxx : x;
    T.miny = T.miny > y ? y : T.miny; cmaxy = cmaxy > y ? cmaxy : y;
    T.minz = T.minz > z ? z : T.minz; cmaxz = cmaxz > z ? cmaxz : z;
  }
  double len = (cmaxx - T.minx) > (cmaxy - T.miny) ? (cmaxx - T.minx) : (cmaxy - T.miny);
  len = (cmaxz - T.minz) > len ? (cmaxz - T.minz) : len;
  if (len <= 0.0) len = 1.0;
  T.side = NEW(double, depth + 1);
  for (int l = 0; l <= depth; l++) { T.side[l] = len; len *= 0.5; }
  newCube(&T, 0, 0, 0, 0, -1);
  int *leafOf = NEW(int, m->nChar);
  for (int j = 0; j < m->nChar; j++) {
    leafOf[j] = goDown(&T, &pos[3 * j]);
    T.c[leafOf[j]].nPnls++;
  }
  T.first = NEW(int, depth + 1);
  T.last = NEW(int, depth + 1);
  for (int l = 0; l <= depth; l++) T.first[l] = T.last[l] = -1;
  linkCubes(&T, 0);
  countUp(&T);

  fb_ctree *t = NEW(fb_ctree, 1);
  t->depth = depth;
  int nc = 0;
  for (int l = 0; l <= depth; l++)
    for (int c = T.first[l]; c >= 0; c = T.c[c].next) T.c[c].flat = nc++;
  t->nc = nc;
  t->levStart = NEW(int, depth + 1); t->levCount = NEW(int, depth + 1);
  t->level = NEW(int, nc); t->kidStart = NEW(int, nc); t->nKids = NEW(int, nc);
  t->x = NEW(double, 3 * (size_t)nc); t->eRad = NEW(double, nc);
  t->chgStart = NEW(int, nc); t->nChgs = NEW(int, nc); t->chgIdx = NEW(int, m->nChar);
  for (int l = 0; l <= depth; l++) {
    t->levStart[l] = T.first[l] >= 0 ? T.c[T.first[l]].flat : nc;
    for (int c = T.first[l]; c >= 0; c = T.c[c].next) t->levCount[l]++;
  }
  /* leaf charge lists: charges in increasing index within a leaf (chargeTree.c:289-296) */
  int cur = 0;
  for (int c = T.first[depth]; c >= 0; c = T.c[c].next) {
    t->chgStart[T.c[c].flat] = cur;
    T.c[c].head = cur;
    cur += T.c[c].nPnls;
  }
  for (int j = 0; j < m->nChar; j++) t->chgIdx[T.c[leafOf[j]].head++] = j;
  free(leafOf);
  /* leaf boxes over their charges, parents from kids (chargeTree.c:174-226) */
  for (int c = T.first[depth]; c >= 0; c = T.c[c].next) {
    cube_t *C = &T.c[c];
    const int *ci = &t->chgIdx[t->chgStart[C->flat]];
    for (int k = 0; k < 3; k++) C->up[k] = C->lo[k] = pos[3 * ci[0] + k];
    for (int i = 1; i < C->nPnls; i++)
      for (int k = 0; k < 3; k++) {
        double v = pos[3 * ci[i] + k];
        C->up[k] = v > C->up[k] ? v : C->up[k];
        C->lo[k] = v < C->lo[k] ? v : C->lo[k];
      }
    for (int k = 0; k < 3; k++) C->x[k] = 0.5 * (C->lo[k] + C->up[k]);
    C->eRad = 0.5 * sqrt(SQR(C->up[0] - C->lo[0]) + SQR(C->up[1] - C->lo[1]) + SQR(C->up[2] - C->lo[2]));
  }
  /* the reference builds parent boxes down to `height` only; coarser cubes
   * keep the grid centre goDown gave leaves or zero, and are never accepted */
  for (int l = 0; l <= depth; l++)
    for (int c = T.first[l]; c >= 0; c = T.c[c].next) {
      cube_t *C = &T.c[c];
      int f = C->flat;
      t->level[f] = l;
      t->nKids[f] = l < depth ? C->nKids : 0;
      t->kidStart[f] = (l < depth && C->nKids) ? T.c[C->kids[0]].flat : -1;
      t->nChgs[f] = C->nPnls;
      if (l < depth) t->chgStart[f] = 0;
    }
  /* boxes: the caller's height decides which levels are valid; fill all
   * levels >= 1 from kids so any height works, level 0 included */
  parentBoxes(&T, 0);
  for (int l = 0; l <= depth; l++)
    for (int c = T.first[l]; c >= 0; c = T.c[c].next) {
      for (int k = 0; k < 3; k++) t->x[3 * T.c[c].flat + k] = T.c[c].x[k];
      t->eRad[T.c[c].flat] = T.c[c].eRad;
    }
  otreeFree(&T);
  return t;
}

/* ================================================================ policy */

/* src/fabipb_options.c:15-16, 174-183 */
int fb_should_use_tree_rhs(int nPnls, int nChar, int gpuMode) {
  const char *f = getenv("FABIPB_FORCE_TREE_RHS");
  if (f && !strcmp(f, "1")) return 1;
  unsigned long long pairs = (unsigned long long)nPnls * (unsigned long long)nChar;
  unsigned long long lim = gpuMode > 0 ? 200000000000ULL : 5000000000ULL;
  const char *e = getenv(gpuMode > 0 ? "FABIPB_MAX_GPU_DIRECT_RHS_PAIRS" : "FABIPB_MAX_DIRECT_RHS_PAIRS");
  if (e && *e) lim = strtoull(e, NULL, 10);
  const char *a = getenv("FABIPB_ALLOW_LARGE_DIRECT_RHS");
  return pairs > lim && !(a && !strcmp(a, "1"));
}

/* src/chargeTree.c:117-129: min(max(ordMom[level], 4), maxOrder) unless overridden */
static unsigned long long envULL(const char *name, unsigned long long dflt) {
  const char *e = getenv(name);
  return (e && *e) ? strtoull(e, NULL, 10) : dflt;
}

/* src/chargeTree.c:117-129 with the capsid-scale policy of
 * src/fabipb_options.c:90-97, 220-229: more than 5e6 charges
 * (FABIPB_HUGE_CAPSID_ATOMS) or panels (FABIPB_HUGE_CAPSID_PANELS) fix the
 * order at 3; FABIPB_CHARGE_TREE_ORDER overrides both */
int fb_charge_tree_order(const fb_tree *t, int level) {
  const char *e = getenv("FABIPB_CHARGE_TREE_ORDER");
  int o;
  const char *ea = getenv("FABIPB_HUGE_CAPSID_ATOMS");
  if (!ea || !*ea) ea = getenv("FABIPB_Q2M_HUGE_CAPSID_ATOMS");
  const unsigned long long atoms = (ea && *ea) ? strtoull(ea, NULL, 10) : 5000000ULL;
  const int capsid = (unsigned long long)t->nChar > atoms ||
                     (unsigned long long)t->n > envULL("FABIPB_HUGE_CAPSID_PANELS", 5000000ULL);
  if (e && atoi(e) >= 1 && atoi(e) <= 32) o = atoi(e);
  else if (capsid) o = 3;
  else o = t->ordMom[level] > 4 ? t->ordMom[level] : 4;
  return o < t->maxOrder ? o : t->maxOrder;
}

/* ============================================================== accessors */
#define ARR(nm, ptr, cnt, dbl) if (!strcmp(name, nm)) { *count = (cnt); *isDouble = (dbl); return (ptr); }

const void *fb_mesh_array(const fb_mesh *m, const char *name, int *count, int *isDouble) {
  ARR("vtx", m->vtx, 9 * m->n, 1); ARR("nrm", m->nrm, 3 * m->n, 1); ARR("area", m->area, m->n, 1);
  ARR("face", m->face, m->n, 0); ARR("pos", m->pos, 3 * m->nChar, 1); ARR("chr", m->chr, m->nChar, 1);
  *count = 0; return NULL;
}

const void *fb_tree_array(const fb_tree *t, const char *name, int *count, int *isDouble) {
  ARR("ordMom", t->ordMom, t->depth + 1, 0); ARR("ordM2L", t->ordM2L, t->depth + 1, 0);
  ARR("perm", t->perm, t->n, 0); ARR("vtx", t->vtx, 9 * t->n, 1); ARR("nrm", t->nrm, 3 * t->n, 1);
  ARR("area", t->area, t->n, 1); ARR("face", t->face, t->n, 0);
  ARR("levStart", t->levStart, t->depth + 1, 0); ARR("levCount", t->levCount, t->depth + 1, 0);
  ARR("cubeLevel", t->cubeLevel, t->nCube, 0); ARR("kidStart", t->kidStart, t->nCube, 0);
  ARR("nKids", t->nKids, t->nCube, 0); ARR("parent", t->parent, t->nCube, 0);
  ARR("cubeX", t->cubeX, 3 * t->nCube, 1); ARR("eRad", t->eRad, t->nCube, 1);
  ARR("leafStart", t->leafStart, t->nLeaf, 0); ARR("leafCount", t->leafCount, t->nLeaf, 0);
  ARR("nearOff", t->nearOff, t->nLeaf + 1, 0); ARR("nearSrc", t->nearSrc, t->nNear, 0);
  ARR("m2lSrc", t->m2lSrc, t->nM2L, 0); ARR("m2lDst", t->m2lDst, t->nM2L, 0);
  ARR("m2lOff", t->m2lOff, t->nCube + 1, 0);
  *count = 0; return NULL;
}

const void *fb_ctree_array(const fb_ctree *t, const char *name, int *count, int *isDouble) {
  ARR("levStart", t->levStart, t->depth + 1, 0); ARR("levCount", t->levCount, t->depth + 1, 0);
  ARR("level", t->level, t->nc, 0); ARR("kidStart", t->kidStart, t->nc, 0); ARR("nKids", t->nKids, t->nc, 0);
  ARR("x", t->x, 3 * t->nc, 1); ARR("eRad", t->eRad, t->nc, 1); ARR("chgStart", t->chgStart, t->nc, 0);
  ARR("nChgs", t->nChgs, t->nc, 0);
  { int tot = 0; for (int c = 0; c < t->nc; c++) if (t->level[c] == t->depth) tot += t->nChgs[c];
    ARR("chgIdx", t->chgIdx, tot, 0); }
  *count = 0; return NULL;
}
