/*
 * ref_harness: records fixtures from the reference FABIPB CPU path.
 *
 * Linked (by oracle/Makefile.ref) against the reference's own objects, with
 * the reference driver src/fabipb.c compiled with its main() renamed, so every
 * number written here is computed by the reference code itself.  This file
 * only sequences the reference's public functions the way src/fabipb.c:804-1229
 * does and dumps their inputs and outputs; it restates no arithmetic.
 *
 * usage: ref_harness <mesh-base> <outdir> [-t= -p= -S= -q= -k= -eps1= -eps2= -H=]
 *                    [-seed=N] [-pairs=N] [-energy=0|1]
 * Needs FABIPB_REUSE_MESH=1 (src/input.c:428) so the mesh files are read and kept.
 *
 * Writes <outdir>/<name>.bin (raw little-endian arrays) and <outdir>/manifest.txt
 * ("name dtype count" per line).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "gkGlobal.h"
#include "gk.h"

/* reference globals (src/fabipb.c:33-44) and functions */
extern int orderMom;
extern double kappa, epsilon, epsilon1, epsilon2;
extern ssystem *sys;
extern void (*kernel)(double *x, double *y);
extern void (*kernelD)(double r, int p, double *G0, double *Gk);
extern double **dG0, **dGk;
extern int ***idx3;
extern double **Q2M0, **Q2M1;

panel *loadPanel(char *panelfile, const char *meshParam, int *numSing, ssystem *sys,
                 const char *meshControlName, double meshControlValue,
                 const char *backendParamName, double backendParamValue);
void gkInit(ssystem *sys, panel *pnlList, int order, int orderMom);
void setupFMM(ssystem *sys);
void applyFMM(ssystem *sys, double *alpha, double *sgm, double *beta, double *pot);
void applyNearfield1(ssystem *sys, double *alpha, double *sgm, double *beta, double *pot);
void setupPreconditioning(ssystem *sys);
void setupRHS(ssystem *sys, double *sgm);
int MtVmain(double *alpha, double *sgm, double *beta, double *pot);
int PtVmain(double *pot, double *sgm);
double *panelIA0(panel *pnlX, panel *pnlY);
int nrCommonVtx(panel *p, panel *q, int *idxX, int *idxY);
void kernelKER4(double *x, double *y);
void kernelDG0DGk(double r, int p, double *G0, double *Gk);
void setupDerivs(int order, double *x);
void buildChargeTree(ssystem *sys);
void computeChgMoments(ssystem *sys);
void applyPanelChargeTreeEnergy(ssystem *sys, double *sgm, double *pot);
int rhsChargeExpansionOrder(ssystem *sys, int level);
void fabipb_resolve_auto_solver_policy(ssystem *sys, int q2mExplicit, int precondExplicit,
                                       int sepExplicit, int nPnls);

static const char *gOut;
static FILE *gMan;

static void dump(const char *name, const char *dtype, const void *data, size_t count) {
  char path[1024];
  size_t esz = (!strcmp(dtype, "f8") || !strcmp(dtype, "i8")) ? 8 : 4;
  snprintf(path, sizeof(path), "%s/%s.bin", gOut, name);
  FILE *fp = fopen(path, "wb");
  if (!fp) { fprintf(stderr, "cannot write %s\n", path); exit(1); }
  if (count) fwrite(data, esz, count, fp);
  fclose(fp);
  fprintf(gMan, "%s %s %zu\n", name, dtype, count);
}

/* SplitMix64 stream for test vectors: x = 2*u - 1, u from the top 53 bits */
static uint64_t gState;
static double urand(void) {
  uint64_t z = (gState += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return 2.0 * ((double)(z >> 11) * 0x1.0p-53) - 1.0;
}

static const char *optv(const char *a, const char *name) {
  size_t n = strlen(name);
  if (a[0] == '-' && !strncmp(a + 1, name, n) && a[1 + n] == '=') return a + n + 2;
  return NULL;
}

int main(int argc, char **argv) {
  char panelfile[80];
  int order = 6, nPnls, i, k, depth = 3, height = 2, qOrd = 1, pairsPerCase = 8, doEnergy = 1;
  double sep = 0.8;
  uint64_t seed = 12345;
  const char *v;

  if (argc < 3) { fprintf(stderr, "usage: ref_harness <mesh-base> <outdir> [opts]\n"); return 2; }
  snprintf(panelfile, sizeof(panelfile), "%s", argv[1]);
  gOut = argv[2];
  kappa = 0.1257; epsilon1 = 1.0; epsilon2 = 80.0;
  for (i = 3; i < argc; i++) {
    if ((v = optv(argv[i], "t"))) depth = atoi(v);
    else if ((v = optv(argv[i], "H"))) height = atoi(v);
    else if ((v = optv(argv[i], "p"))) order = atoi(v);
    else if ((v = optv(argv[i], "S"))) sep = atof(v);
    else if ((v = optv(argv[i], "q"))) qOrd = atoi(v);
    else if ((v = optv(argv[i], "k"))) kappa = atof(v);
    else if ((v = optv(argv[i], "eps1"))) epsilon1 = atof(v);
    else if ((v = optv(argv[i], "eps2"))) epsilon2 = atof(v);
    else if ((v = optv(argv[i], "seed"))) seed = strtoull(v, NULL, 10);
    else if ((v = optv(argv[i], "pairs"))) pairsPerCase = atoi(v);
    else if ((v = optv(argv[i], "energy"))) doEnergy = atoi(v);
    else { fprintf(stderr, "unknown option %s\n", argv[i]); return 2; }
  }
  epsilon = epsilon2 / epsilon1;
  {
    char path[1024];
    snprintf(path, sizeof(path), "%s/manifest.txt", gOut);
    gMan = fopen(path, "w");
    if (!gMan) { fprintf(stderr, "cannot write %s\n", path); return 1; }
  }

  /* defaults of src/fabipb.c:834-849 */
  CALLOC(sys, 1, ssystem);
  sys->height = height; sys->maxSepRatio = sep; sys->maxQuadOrder = qOrd; sys->nKerl = 4;
  sys->depth = depth; sys->mesh_flag = 1; sys->gpuMode = 0; sys->matvecMode = 0;
  sys->gpuQ2MMode = 0; sys->gpuNearfieldMode = -1; sys->precondCacheMode = 3;

  panel *inputLst = loadPanel(panelfile, "1", &nPnls, sys, "mesh_resolution", 1.0, "msms_density", 1.0);
  fabipb_resolve_auto_solver_policy(sys, 1, 1, 1, nPnls);
  sys->pnlOLst = inputLst;
  gkInit(sys, inputLst, order, orderMom);
  setupFMM(sys);
  setupPreconditioning(sys);

  int n = sys->nPnls;
  int params[8] = {n, sys->nLeafCubesFlat, sys->nNearPairsFlat, sys->nFmmCubesFlat,
                   sys->nM2LPairsFlat, sys->nM2LDstGroups, sys->maxOrder, sys->nChar};
  dump("counts", "i4", params, 8);
  double fparams[4] = {kappa, epsilon1, epsilon2, sep};
  dump("params", "f8", fparams, 4);

  /* panels in idx order */
  {
    double *geo = malloc(sizeof(double) * 13 * (size_t)n);
    int *face = malloc(sizeof(int) * (size_t)n);
    for (i = 0; i < n; i++) {
      panel *p = sys->panelByIdx[i];
      for (k = 0; k < 9; k++) geo[13 * i + k] = p->vtx[k / 3][k % 3];
      for (k = 0; k < 3; k++) geo[13 * i + 9 + k] = p->normal[k];
      geo[13 * i + 12] = p->area;
      face[i] = p->nSurf - 1;
    }
    dump("panel_geom", "f8", geo, 13 * (size_t)n);
    dump("panel_face", "i4", face, n);
    free(geo); free(face);
  }
  dump("pos", "f8", sys->pos, 3 * (size_t)sys->nChar);
  dump("chr", "f8", sys->chr, sys->nChar);

  /* flattened layout (src/fmm.c:234-404) */
  dump("leaf_start", "i4", sys->leafPanelStart, sys->nLeafCubesFlat);
  dump("leaf_count", "i4", sys->leafPanelCount, sys->nLeafCubesFlat);
  dump("near_src", "i4", sys->nearPairSrc, sys->nNearPairsFlat);
  dump("near_dst", "i4", sys->nearPairDst, sys->nNearPairsFlat);
  dump("near_off", "i4", sys->nearLeafPairOffset, sys->nLeafCubesFlat + 1);
  dump("m2l_src", "i4", sys->m2lPairSrc, sys->nM2LPairsFlat);
  dump("m2l_dst", "i4", sys->m2lPairDst, sys->nM2LPairsFlat);
  dump("m2l_order", "i4", sys->m2lPairOrder, sys->nM2LPairsFlat);
  dump("lev_start", "i4", sys->fmmLevelStart, sys->depth + 1);
  dump("lev_count", "i4", sys->fmmLevelCount, sys->depth + 1);
  dump("ord_mom", "i4", sys->ordMom, sys->depth + 1);
  dump("ord_m2l", "i4", sys->ordM2L, sys->depth + 1);
  {
    int nc = sys->nFmmCubesFlat;
    double *cx = malloc(sizeof(double) * 4 * (size_t)nc);
    int *ci = malloc(sizeof(int) * 12 * (size_t)nc);
    for (i = 0; i < nc; i++) {
      cube *c = sys->fmmCubeByIdx[i];
      for (k = 0; k < 3; k++) cx[4 * i + k] = c->x[k];
      cx[4 * i + 3] = c->eRad;
      ci[12 * i] = c->level;
      ci[12 * i + 1] = c->nKids;
      ci[12 * i + 2] = (c->parent && c->level > sys->height) ? c->parent->flatIdx : -1;
      ci[12 * i + 3] = c->nPnls;
      for (k = 0; k < 8; k++) ci[12 * i + 4 + k] = (c->level < sys->depth && k < c->nKids) ? c->kids[k]->flatIdx : -1;
    }
    dump("cube_x", "f8", cx, 4 * (size_t)nc);
    dump("cube_i", "i4", ci, 12 * (size_t)nc);
    free(cx); free(ci);
  }

  /* single-kernel values: KER4 at a few points with panel normals */
  {
    extern __thread double *nrmX, *nrmY;
    double out[5 * 7];
    double pts[5][3] = {{1.0, 0.5, -0.25}, {0.1, -0.2, 0.05}, {7.0, 3.0, 2.0}, {-3.0, 0.4, 11.0}, {0.01, 0.02, -0.03}};
    nrmX = sys->panelByIdx[0]->normal;
    nrmY = sys->panelByIdx[n - 1]->normal;
    for (i = 0; i < 5; i++) {
      kernelKER4(pts[i], &out[7 * i]);
      out[7 * i + 4] = pts[i][0]; out[7 * i + 5] = pts[i][1]; out[7 * i + 6] = pts[i][2];
    }
    dump("ker4", "f8", out, 35);
  }

  /* derivative tensors (src/expan.c:138-209) at a few translation vectors */
  {
    int p = sys->maxOrder, nm = sys->nMom[p];
    double rs[4][3] = {{3.0, -1.0, 2.0}, {0.7, 0.2, -0.4}, {-12.0, 5.0, 1.0}, {0.0, 0.0, 4.0}};
    double *buf = malloc(sizeof(double) * 2 * (size_t)nm * 4);
    for (i = 0; i < 4; i++) {
      setupDerivs(p, rs[i]);
      memcpy(&buf[2 * nm * i], dG0[0], sizeof(double) * nm);
      memcpy(&buf[2 * nm * i + nm], dGk[0], sizeof(double) * nm);
    }
    dump("derivs", "f8", buf, 2 * (size_t)nm * 4);
    free(buf);
  }

  /* leaf Q2M matrices of the first leaves (src/moments.c:216-236) */
  {
    int nm = sys->nMom[sys->ordMom[sys->depth]], nl = sys->nLeafCubesFlat < 4 ? sys->nLeafCubesFlat : 4, tot = 0;
    for (i = 0; i < nl; i++) tot += sys->leafPanelCount[i] * nm;
    double *q0 = malloc(sizeof(double) * (size_t)tot), *q1 = malloc(sizeof(double) * (size_t)tot);
    int off = 0;
    for (i = 0; i < nl; i++) {
      memcpy(&q0[off], Q2M0[i], sizeof(double) * sys->leafPanelCount[i] * nm);
      memcpy(&q1[off], Q2M1[i], sizeof(double) * sys->leafPanelCount[i] * nm);
      off += sys->leafPanelCount[i] * nm;
    }
    dump("q2m0", "f8", q0, tot);
    dump("q2m1", "f8", q1, tot);
    free(q0); free(q1);
  }

  /* panelIA0 per geometric case (src/numQuad.c:629-697), sampled from near pairs */
  {
    int want[5] = {0, 1, 2, -2, 3}, got[5] = {0, 0, 0, 0, 0}, idxX[3], idxY[3];
    int cap = 5 * pairsPerCase, cnt = 0;
    int *pi = malloc(sizeof(int) * 3 * (size_t)cap);
    double *pv = malloc(sizeof(double) * 4 * (size_t)cap);
    kernel = kernelKER4;
    for (int L = 0; L < sys->nLeafCubesFlat && cnt < cap; L++) {
      for (int pr = sys->nearLeafPairOffset[L]; pr < sys->nearLeafPairOffset[L + 1] && cnt < cap; pr++) {
        int s = sys->nearPairSrc[pr];
        for (int a = 0; a < sys->leafPanelCount[L] && cnt < cap; a++)
          for (int b = 0; b < sys->leafPanelCount[s] && cnt < cap; b++) {
            int x = sys->leafPanelStart[L] + a, y = sys->leafPanelStart[s] + b;
            int c = nrCommonVtx(sys->panelByIdx[x], sys->panelByIdx[y], idxX, idxY), w;
            for (w = 0; w < 5 && want[w] != c; w++) ;
            if (w == 5 || got[w] >= pairsPerCase) continue;
            double *K = panelIA0(sys->panelByIdx[x], sys->panelByIdx[y]);
            pi[3 * cnt] = x; pi[3 * cnt + 1] = y; pi[3 * cnt + 2] = c;
            for (k = 0; k < 4; k++) pv[4 * cnt + k] = K[k];
            got[w]++; cnt++;
          }
      }
    }
    dump("ia0_idx", "i4", pi, 3 * (size_t)cnt);
    dump("ia0_val", "f8", pv, 4 * (size_t)cnt);
    free(pi); free(pv);
  }

  /* one matvec: y = (D - K) x at a seeded x, plus the nearfield alone */
  {
    double *x = malloc(sizeof(double) * 2 * (size_t)n), *y = calloc(2 * (size_t)n, sizeof(double));
    double *yn = calloc(2 * (size_t)n, sizeof(double)), *yp = calloc(2 * (size_t)n, sizeof(double));
    double one = 1.0, zero = 0.0;
    gState = seed;
    for (i = 0; i < 2 * n; i++) x[i] = urand();
    dump("mv_x", "f8", x, 2 * (size_t)n);
    MtVmain(&one, x, &zero, y);
    dump("mv_y", "f8", y, 2 * (size_t)n);
    kernel = kernelKER4;
    applyNearfield1(sys, &one, x, &zero, yn);
    dump("mv_near", "f8", yn, 2 * (size_t)n);
    /* moments and local expansions left in the cubes by the last applyFMM */
    applyFMM(sys, &one, x, &zero, yp);
    {
      int nc = sys->nFmmCubesFlat, nm = sys->nMom[sys->maxOrder];
      double *mom = calloc((size_t)nc * 2 * nm, sizeof(double)), *lec = calloc((size_t)nc * 4 * nm, sizeof(double));
      for (i = 0; i < nc; i++) {
        cube *c = sys->fmmCubeByIdx[i];
        int m = sys->nMom[sys->ordMom[c->level]];
        memcpy(&mom[(size_t)i * 2 * nm], c->mom_pot, sizeof(double) * m);
        memcpy(&mom[(size_t)i * 2 * nm + nm], c->mom_dpdn, sizeof(double) * m);
        memcpy(&lec[(size_t)i * 4 * nm], c->lec_k1, sizeof(double) * m);
        memcpy(&lec[(size_t)i * 4 * nm + nm], c->lec_k2, sizeof(double) * m);
        memcpy(&lec[(size_t)i * 4 * nm + 2 * nm], c->lec_k3, sizeof(double) * m);
        memcpy(&lec[(size_t)i * 4 * nm + 3 * nm], c->lec_k4, sizeof(double) * m);
      }
      dump("mv_mom", "f8", mom, (size_t)nc * 2 * nm);
      dump("mv_lec", "f8", lec, (size_t)nc * 4 * nm);
      free(mom); free(lec);
    }
    dump("mv_fmm", "f8", yp, 2 * (size_t)n);
    /* diagonal preconditioner solve of the same x (src/precond_fmm.c:664-728) */
    PtVmain(yp, x);
    dump("pc_y", "f8", yp, 2 * (size_t)n);
    free(x); free(y); free(yn); free(yp);
  }

  /* right-hand side (src/fabipb.c:744-800) */
  {
    double *b = calloc(2 * (size_t)n, sizeof(double));
    setupRHS(sys, b);
    dump("rhs", "f8", b, 2 * (size_t)n);
    free(b);
  }

  /* panel-tree energy (src/treecode.c:927-1017) of a seeded density, and the charge tree */
  if (doEnergy) {
    double *s = malloc(sizeof(double) * 2 * (size_t)n), e = 0.0;
    gState = seed + 1;
    for (i = 0; i < 2 * n; i++) s[i] = urand();
    buildChargeTree(sys);
    computeChgMoments(sys);
    applyPanelChargeTreeEnergy(sys, s, &e);
    dump("en_sigma", "f8", s, 2 * (size_t)n);
    dump("en_raw", "f8", &e, 1);
    free(s);
    /* charge tree, all levels 0..depth in per-level DFS order (src/chargeTree.c:136-307) */
    int lev, nc = 0, nm = sys->nMom[sys->maxOrder], nidx = 0;
    for (lev = 0; lev <= sys->chgDepth; lev++)
      for (cube *c = sys->chgCubeList[lev]; c; c = c->next) nc++;
    double *cx = calloc((size_t)nc * (4 + nm), sizeof(double));
    int *ci = calloc((size_t)nc * 3, sizeof(int));
    int *cidx = calloc((size_t)sys->nChar, sizeof(int));
    int j = 0;
    for (lev = 0; lev <= sys->chgDepth; lev++)
      for (cube *c = sys->chgCubeList[lev]; c; c = c->next, j++) {
        for (k = 0; k < 3; k++) cx[(size_t)j * (4 + nm) + k] = c->x[k];
        cx[(size_t)j * (4 + nm) + 3] = c->eRad;
        if (lev >= sys->height && c->mom_chr)
          memcpy(&cx[(size_t)j * (4 + nm) + 4], c->mom_chr, sizeof(double) * sys->nMom[rhsChargeExpansionOrder(sys, lev)]);
        ci[3 * j] = lev; ci[3 * j + 1] = c->nKids; ci[3 * j + 2] = c->nChgs;
        if (lev == sys->chgDepth)
          for (k = 0; k < c->nChgs; k++) cidx[nidx++] = c->chgIdx[k];
      }
    dump("ctree_x", "f8", cx, (size_t)nc * (4 + nm));
    dump("ctree_i", "i4", ci, (size_t)nc * 3);
    dump("ctree_idx", "i4", cidx, nidx);
    free(cx); free(ci); free(cidx);
  }
  fclose(gMan);
  printf("ref_harness: n=%d leaves=%d near-pairs=%d cubes=%d m2l=%d\n", n, sys->nLeafCubesFlat,
         sys->nNearPairsFlat, sys->nFmmCubesFlat, sys->nM2LPairsFlat);
  return 0;
}
