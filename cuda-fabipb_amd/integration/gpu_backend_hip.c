/*
 * gpu_backend_hip.c - the reference's GPU backend ABI (the 17 functions of
 * include/gpu_backend.h in Jiahuic/CUDA-FABIPB) implemented on
 * libfabipb_hip.so.
 *
 * This is the binding a maintainer adds to the reference's own build in
 * place of src/gpu_backend_cuda.cu: it is compiled against the reference's
 * gk.h (struct ssystem / cube / panel) and gpu_backend.h, reads the host
 * layout the reference already flattened (src/fmm.c:234-524), and forwards
 * each stage call to the flat C ABI.  Return conventions are the reference's:
 * 1 = done, 0 = not done (the caller then runs its CPU path and prints the
 * message from gpuNearfieldLastError / gpuM2LLastError).
 *
 * One operator is cached per (sys, kappa, eps1, eps2, quadrature order), as
 * the reference caches per sys pointer (src/gpu_backend_cuda.cu:342-351).
 * Not re-entrant, like the reference's backend.
 *
 * The two rhs-reuse entry points (the rhs-reuse energy mode, out of scope
 * here) return 0, which the reference host treats as "use the CPU path".
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gk.h"
#include "gpu_backend.h"
#include "fabipb_hip.h"

extern double kappa, epsilon, epsilon1, epsilon2;   /* src/fabipb.c:34 */
int rhsChargeExpansionOrder(ssystem *sys, int level);  /* src/chargeTree.c:117 */
double energyTreeTheta(void);                          /* src/treecode.c:133 */
double rhsTreeTheta(void);                             /* src/treecode.c:75 */

static fbh_ctx *gCtx;
static ssystem *gSys;
static double gKappa, gEps1, gEps2;
static int gQ, gNm;
static double *gMom, *gLec;
static char gMsg[600];

static int nmom(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }

/* FABIPB_HIP_TRACE=1: one line on stderr per completed hook call that has a
 * CPU fallback in the reference, so a run can show which work the device did */
static void trace(const char *hook, long long count) {
  static int on = -1;
  if (on < 0) {
    const char *e = getenv("FABIPB_HIP_TRACE");
    on = e && *e && *e != '0';
  }
  if (on) fprintf(stderr, "gpu_backend_hip: %s done (%lld)\n", hook, count);
}

static int fail(const char *what) {
  snprintf(gMsg, sizeof(gMsg), "%s: %s", what, fbh_last_error());
  return 0;
}

int gpuBackendAvailable(void) { return fbh_available(); }
const char *gpuNearfieldLastError(void) { return gMsg[0] ? gMsg : fbh_last_error(); }
const char *gpuM2LLastError(void) { return gMsg[0] ? gMsg : fbh_last_error(); }

void gpuReleaseMatvecCaches(void) {
  fbh_destroy(gCtx);
  gCtx = NULL;
  gSys = NULL;
  free(gMom); free(gLec);
  gMom = gLec = NULL;
}
void gpuReleaseChargeTreeCache(void) {}

/* build the device operator from the reference's flattened layout */
static int ensure(ssystem *sys) {
  if (gCtx && gSys == sys && gKappa == kappa && gEps1 == epsilon1 && gEps2 == epsilon2 && gQ == sys->maxQuadOrder)
    return 1;
  gpuReleaseMatvecCaches();
  if (!fbh_available()) return fail("no gfx950 device");
  int n = sys->nPnls, depth = sys->depth, height = sys->height, nCube = sys->nFmmCubesFlat, i, k;
  double *vtx = malloc(sizeof(double) * 9 * (size_t)n), *nrm = malloc(sizeof(double) * 3 * (size_t)n);
  double *area = malloc(sizeof(double) * (size_t)n), *cx = malloc(sizeof(double) * 3 * (size_t)nCube);
  int *level = malloc(sizeof(int) * (size_t)nCube), *kidStart = malloc(sizeof(int) * (size_t)nCube);
  int *nKids = malloc(sizeof(int) * (size_t)nCube), *parent = malloc(sizeof(int) * (size_t)nCube);
  int *m2lOff = calloc((size_t)nCube + 1, sizeof(int));
  int ok = vtx && nrm && area && cx && level && kidStart && nKids && parent && m2lOff;
  for (i = 0; ok && i < n; i++) {
    panel *p = sys->panelByIdx[i];
    for (k = 0; k < 9; k++) vtx[9 * i + k] = p->vtx[k / 3][k % 3];
    for (k = 0; k < 3; k++) nrm[3 * i + k] = p->normal[k];
    area[i] = p->area;
  }
  for (i = 0; ok && i < nCube; i++) {
    cube *cb = sys->fmmCubeByIdx[i];
    level[i] = cb->level;
    for (k = 0; k < 3; k++) cx[3 * i + k] = cb->x[k];
    nKids[i] = cb->level < depth ? cb->nKids : 0;
    kidStart[i] = nKids[i] ? cb->kids[0]->flatIdx : -1;
    for (k = 1; k < nKids[i]; k++)
      if (cb->kids[k]->flatIdx != kidStart[i] + k) ok = 0;   /* kids are contiguous in the DFS layout */
    parent[i] = cb->level > height ? cb->parent->flatIdx : -1;
  }
  for (i = 0; ok && i < sys->nM2LPairsFlat; i++) m2lOff[sys->m2lPairDst[i] + 1]++;
  for (i = 0; ok && i < nCube; i++) m2lOff[i + 1] += m2lOff[i];
  for (i = 1; ok && i < sys->nM2LPairsFlat; i++)
    if (sys->m2lPairDst[i] < sys->m2lPairDst[i - 1]) ok = 0;   /* grouped by destination, fmm.c:356-372 */
  if (ok) {
    gCtx = fbh_create(kappa, epsilon1, epsilon2, sys->maxQuadOrder, n, vtx, nrm, area, depth, height, sys->ordMom,
                      sys->ordM2L, nCube, sys->fmmLevelStart, sys->fmmLevelCount, level, cx, kidStart, nKids, parent,
                      sys->nLeafCubesFlat, sys->leafPanelStart, sys->leafPanelCount, sys->nearLeafPairOffset,
                      sys->nearPairSrc, m2lOff, sys->m2lPairSrc, NULL);
  } else {
    snprintf(gMsg, sizeof(gMsg), "layout not in the flattened form this backend reads");
  }
  free(vtx); free(nrm); free(area); free(cx); free(level); free(kidStart); free(nKids); free(parent); free(m2lOff);
  if (!gCtx) return ok ? fail("operator setup") : 0;
  long long st[16];
  fbh_stats(gCtx, st);
  gNm = (int)st[3];
  gMom = calloc((size_t)nCube * 2 * gNm, sizeof(double));
  gLec = calloc((size_t)nCube * 4 * gNm, sizeof(double));
  gSys = sys; gKappa = kappa; gEps1 = epsilon1; gEps2 = epsilon2; gQ = sys->maxQuadOrder;
  gMsg[0] = 0;
  return gMom && gLec;
}

int gpuNearfieldApply(ssystem *sys, double alpha, const double *sgm, double *pot) {
  if (!ensure(sys)) return 0;
  return fbh_near_apply(gCtx, alpha, sgm, pot) ? 1 : fail("nearfield");
}

int gpuQ2MApply(ssystem *sys, const double *sgm) {
  if (!ensure(sys)) return 0;
  if (!fbh_q2m_host(gCtx, sgm, gMom)) return fail("Q2M");
  int m = nmom(sys->ordMom[sys->depth]);
  for (int l = 0; l < sys->nLeafCubesFlat; l++) {
    cube *cb = sys->fmmCubeByIdx[l];
    memcpy(cb->mom_pot, &gMom[(size_t)l * 2 * gNm], sizeof(double) * m);
    memcpy(cb->mom_dpdn, &gMom[(size_t)l * 2 * gNm + gNm], sizeof(double) * m);
  }
  return 1;
}

int gpuM2LApply(ssystem *sys) {
  if (!ensure(sys)) return 0;
  for (int c = 0; c < sys->nFmmCubesFlat; c++) {
    cube *cb = sys->fmmCubeByIdx[c];
    int m = nmom(sys->ordMom[cb->level]);
    memset(&gMom[(size_t)c * 2 * gNm], 0, sizeof(double) * 2 * gNm);
    memcpy(&gMom[(size_t)c * 2 * gNm], cb->mom_pot, sizeof(double) * m);
    memcpy(&gMom[(size_t)c * 2 * gNm + gNm], cb->mom_dpdn, sizeof(double) * m);
  }
  if (!fbh_m2l_host(gCtx, gMom, gLec)) return fail("M2L");
  for (int c = 0; c < sys->nFmmCubesFlat; c++) {   /* overwrite, as src/gpu_backend_cuda.cu:622-681 */
    cube *cb = sys->fmmCubeByIdx[c];
    int m = nmom(sys->ordMom[cb->level]);
    double *dst[4] = {cb->lec_k1, cb->lec_k2, cb->lec_k3, cb->lec_k4};
    for (int k = 0; k < 4; k++) memcpy(dst[k], &gLec[((size_t)c * 4 + k) * gNm], sizeof(double) * m);
  }
  return 1;
}

int gpuL2PApply(ssystem *sys, double alpha, double beta, double *pot) {
  if (!ensure(sys)) return 0;
  int m = nmom(sys->ordMom[sys->depth]);
  for (int l = 0; l < sys->nLeafCubesFlat; l++) {
    cube *cb = sys->fmmCubeByIdx[l];
    double *src[4] = {cb->lec_k1, cb->lec_k2, cb->lec_k3, cb->lec_k4};
    for (int k = 0; k < 4; k++) {
      memset(&gLec[((size_t)l * 4 + k) * gNm], 0, sizeof(double) * gNm);
      memcpy(&gLec[((size_t)l * 4 + k) * gNm], src[k], sizeof(double) * m);
    }
  }
  return fbh_l2p_host(gCtx, gLec, alpha, beta, pot) ? 1 : fail("L2P");
}

int gpuSetupRHS(ssystem *sys, int qOrder, double fac, double *sgm) {
  if (qOrder != sys->maxQuadOrder || !ensure(sys) || !fbh_set_rhs_factor(gCtx, fac)) return 0;
  if (!fbh_rhs(gCtx, sys->nChar, sys->pos, sys->chr, sgm)) return fail("setupRHS");
  trace("gpuSetupRHS", sys->nPnls);
  return 1;
}

/* the reference's charge tree (sys->chgCubeList, src/chargeTree.c:228-307) as flat arrays */
typedef struct {
  int nc, *level, *kidStart, *nKids, *chgStart, *nChgs, *chgIdx, *ord;
  double *x, *eRad;
} flatCT;

static void freeCT(flatCT *t) {
  free(t->level); free(t->kidStart); free(t->nKids); free(t->chgStart); free(t->nChgs); free(t->chgIdx);
  free(t->ord); free(t->x); free(t->eRad);
}

static int flattenCT(ssystem *sys, flatCT *t) {
  int depth = sys->chgDepth, lev, i, j = 0, cur = 0;
  memset(t, 0, sizeof(*t));
  for (lev = 0; lev <= depth; lev++)
    for (cube *c = sys->chgCubeList[lev]; c; c = c->next) t->nc++;
  int nc = t->nc;
  t->level = malloc(sizeof(int) * (size_t)nc); t->kidStart = malloc(sizeof(int) * (size_t)nc);
  t->nKids = malloc(sizeof(int) * (size_t)nc); t->chgStart = calloc((size_t)nc, sizeof(int));
  t->nChgs = malloc(sizeof(int) * (size_t)nc); t->chgIdx = malloc(sizeof(int) * (size_t)sys->nChar);
  t->ord = malloc(sizeof(int) * (size_t)(depth + 1));
  t->x = malloc(sizeof(double) * 3 * (size_t)nc); t->eRad = malloc(sizeof(double) * (size_t)nc);
  if (!t->level || !t->kidStart || !t->nKids || !t->chgStart || !t->nChgs || !t->chgIdx || !t->ord || !t->x ||
      !t->eRad) {
    freeCT(t);
    return 0;
  }
  for (lev = 0; lev <= depth; lev++) {
    t->ord[lev] = rhsChargeExpansionOrder(sys, lev);
    for (cube *c = sys->chgCubeList[lev]; c; c = c->next, j++) {
      t->level[j] = lev;
      for (i = 0; i < 3; i++) t->x[3 * j + i] = c->x[i];
      t->eRad[j] = c->eRad;
      t->nKids[j] = lev < depth ? c->nKids : 0;
      t->nChgs[j] = c->nChgs;
      if (lev == depth) {
        t->chgStart[j] = cur;
        for (i = 0; i < c->nChgs; i++) t->chgIdx[cur++] = c->chgIdx[i];
      }
    }
  }
  /* kids of consecutive parents are consecutive in the next level's list */
  j = 0;
  for (lev = 0; lev <= depth; lev++) {
    int kidNext = j;
    for (cube *c = sys->chgCubeList[lev]; c; c = c->next) kidNext++;
    for (cube *c = sys->chgCubeList[lev]; c; c = c->next, j++) {
      t->kidStart[j] = t->nKids[j] ? kidNext : -1;
      kidNext += t->nKids[j];
    }
  }
  return 1;
}

int gpuPanelChargeTreeEnergy(ssystem *sys, const double *sgm, double *pot) {
  if (!sys->chgCubeList || !ensure(sys)) return 0;
  flatCT t;
  if (!flattenCT(sys, &t)) return 0;
  int rc = 0, n2 = 2 * sys->nPnls;
  double *d = fbh_dev_alloc(gCtx, n2);
  if (d && fbh_copy_h2d(gCtx, d, sgm, n2) &&
      fbh_energy_dev(gCtx, sys->nChar, sys->pos, sys->chr, t.nc, t.level, t.x, t.eRad, t.kidStart, t.nKids,
                     t.chgStart, t.nChgs, t.chgIdx, t.ord, energyTreeTheta(), d, pot))
    rc = 1;
  else
    fail("panel-tree energy");
  fbh_dev_free(gCtx, d);
  freeCT(&t);
  return rc;
}

/* sgm = the charge-tree right-hand side (setupRHSTree, src/fabipb.c:733-742) scaled
 * by fac; qOrder the operator's own */
int gpuChargeTreeRHS(ssystem *sys, int qOrder, double fac, double *sgm) {
  if (!sys->chgCubeList || !ensure(sys) || qOrder != sys->maxQuadOrder || !fbh_set_rhs_factor(gCtx, fac)) return 0;
  flatCT t;
  if (!flattenCT(sys, &t)) return 0;
  int rc = 0, n2 = 2 * sys->nPnls;
  double *d = fbh_dev_alloc(gCtx, n2);
  if (d && fbh_rhs_tree_dev(gCtx, sys->nChar, sys->pos, sys->chr, t.nc, t.level, t.x, t.eRad, t.kidStart,
                            t.nKids, t.chgStart, t.nChgs, t.chgIdx, t.ord, rhsTreeTheta(), d) &&
      fbh_copy_d2h(gCtx, sgm, d, n2)) {
    rc = 1;
    trace("gpuChargeTreeRHS", sys->nPnls);
  } else {
    fail("tree right-hand side");
  }
  fbh_dev_free(gCtx, d);
  freeCT(&t);
  return rc;
}

/* pot = beta*pot + alpha*K_direct*sgm, the dense operator of -r=1 (src/fabipb.c:1251-1258) */
int gpuDirectApply(ssystem *sys, double alpha, double beta, const double *sgm, double *pot) {
  if (!ensure(sys)) return 0;
  return fbh_direct_apply(gCtx, alpha, beta, sgm, pot) ? 1 : fail("direct operator");
}
int gpuRhsReuseCoulombDot(ssystem *sys, const double *rhs, const double *sgm, double *dot) {
  (void)sys; (void)rhs; (void)sgm; (void)dot;
  return 0;
}
int gpuRhsReusePanelChargeTreeEnergy(ssystem *sys, const double *rhs, const double *sgm, double *pot) {
  (void)sys; (void)rhs; (void)sgm; (void)pot;
  return 0;
}
/* the disjoint-pair coefficients of a batch of preconditioner blocks
 * (src/precond_fmm.c:453-552): k[p] = panelIA0(X = pairSrc[p], Y = pairDst[p]), pnl->idx
 * indices; the caller keeps the disjoint pairs' values */
int gpuBuildPrecondPairsBatched(ssystem *sys, const int *pairSrc, const int *pairDst, long long nPairs, double *k0,
                                double *k1, double *k2, double *k3) {
  if (!sys || !pairSrc || !pairDst || nPairs <= 0 || !k0 || !k1 || !k2 || !k3 || !ensure(sys)) return 0;
  if (!fbh_pair_coeffs(gCtx, nPairs, pairSrc, pairDst, k0, k1, k2, k3)) return fail("preconditioner pairs");
  trace("gpuBuildPrecondPairsBatched", nPairs);
  return 1;
}

/* one block's disjoint pairs (src/precond_fmm.c:223-240, used at -q=1): KER = panelIA0(X =
 * panels[srcLocal], Y = panels[dstLocal]) written into the row-major 2m x 2m block as the
 * reference's own loop writes it (-KER1 | -KER0 over -KER3 | -KER2) */
int gpuBuildPrecondDisjointBlock(panel **panels, int nPanels, const int *dstLocal, const int *srcLocal, int nPairs,
                                 double *block) {
  if (!panels || nPanels <= 0 || !dstLocal || !srcLocal || nPairs <= 0 || !block) return 0;
  double *vtx = malloc(sizeof(double) * 9 * (size_t)nPanels), *nrm = malloc(sizeof(double) * 3 * (size_t)nPanels);
  double *area = malloc(sizeof(double) * (size_t)nPanels), *k = malloc(sizeof(double) * 4 * (size_t)nPairs);
  int ok = vtx && nrm && area && k;
  for (int i = 0; ok && i < nPanels; i++) {
    for (int j = 0; j < 9; j++) vtx[9 * i + j] = panels[i]->vtx[j / 3][j % 3];
    for (int j = 0; j < 3; j++) nrm[3 * i + j] = panels[i]->normal[j];
    area[i] = panels[i]->area;
  }
  if (ok)
    ok = fbh_panel_pair_coeffs(nPanels, vtx, nrm, area, kappa, epsilon, 1, nPairs, srcLocal, dstLocal, k,
                               k + nPairs, k + 2 * (size_t)nPairs, k + 3 * (size_t)nPairs);
  if (ok) {
    const int hm = nPanels, ms = 2 * nPanels;
    for (int p = 0; p < nPairs; p++) {
      const int d = dstLocal[p], s = srcLocal[p];
      block[(size_t)d * ms + s] = -k[nPairs + p];
      block[(size_t)d * ms + s + hm] = -k[p];
      block[(size_t)(d + hm) * ms + s] = -k[3 * (size_t)nPairs + p];
      block[(size_t)(d + hm) * ms + s + hm] = -k[2 * (size_t)nPairs + p];
    }
    trace("gpuBuildPrecondDisjointBlock", nPairs);
  } else {
    fail("preconditioner block");
  }
  free(vtx); free(nrm); free(area); free(k);
  return ok;
}
