/*
 * fbo.c - CPU oracle for the FABIPB Galerkin BIE matvec (see fbo.h).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never part of it.
 *
 * Each function restates one reference routine, in the same arithmetic order
 * where the order is observable, and cites the reference file:line it follows.
 */
#define _POSIX_C_SOURCE 200809L
#include "fbo.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define OMAX 16                        /* largest total degree handled */
#define NMOM(p) (((p) + 1) * ((p) + 2) * ((p) + 3) / 6)

static const double kFourPiI = 1.0 / (4 * M_PI);   /* src/gkGlobal.c:44 */

/* ------------------------------------------------------------------ tables */
/* Gauss-Legendre abscissas/weights on [0,1], the 15-digit table of
 * src/numQuad.c:221-294 (orders 1..10).  Values, not recomputed: the table's
 * last-digit asymmetries are part of the reference's arithmetic. */
static const double kT[11][10] = {
    {0},
    {0.500000000000000},
    {0.211324865405187, 0.788675134594813},
    {0.112701665379258, 0.500000000000000, 0.887298334620742},
    {0.069431844202974, 0.330009478207572, 0.669990521792428, 0.930568155797026},
    {0.046910077030668, 0.230765344947158, 0.500000000000000, 0.769234655052841, 0.953089922969332},
    {0.033765242898424, 0.169395306766868, 0.380690406958402, 0.619309593041598, 0.830604693233132,
     0.966234757101576},
    {0.025446043828621, 0.129234407200303, 0.297077424311301, 0.500000000000000, 0.702922575688699,
     0.870765592799697, 0.974553956171379},
    {0.019855071751232, 0.101666761293187, 0.237233795041836, 0.408282678752175, 0.591717321247825,
     0.762766204958164, 0.898333238706813, 0.980144928248768},
    {0.015919880246187, 0.081984446336682, 0.193314283649705, 0.337873288298095, 0.500000000000000,
     0.662126711701904, 0.806685716350295, 0.918015553663318, 0.984080119753813},
    {0.013046735741414, 0.067468316655508, 0.160295215850488, 0.283302302935376, 0.425562830509184,
     0.574437169490816, 0.716697697064624, 0.839704784149512, 0.932531683344492, 0.986953264258586}};
static const double kW[11][10] = {
    {0},
    {1.000000000000000},
    {0.500000000000000, 0.500000000000000},
    {0.277777777777778, 0.444444444444444, 0.277777777777778},
    {0.173927422568727, 0.326072577431273, 0.326072577431273, 0.173927422568727},
    {0.118463442528095, 0.239314335249683, 0.284444444444444, 0.239314335249683, 0.118463442528095},
    {0.085662246189585, 0.180380786524070, 0.233956967286346, 0.233956967286346, 0.180380786524069,
     0.085662246189585},
    {0.064742483084435, 0.139852695744638, 0.190915025252560, 0.208979591836735, 0.190915025252559,
     0.139852695744638, 0.064742483084435},
    {0.050614268145188, 0.111190517226687, 0.156853322938944, 0.181341891689181, 0.181341891689181,
     0.156853322938944, 0.111190517226687, 0.050614268145188},
    {0.040637194180787, 0.090324080347429, 0.130305348201467, 0.156173538520002, 0.165119677500630,
     0.156173538520002, 0.130305348201468, 0.090324080347429, 0.040637194180787},
    {0.033335672154344, 0.074725674575290, 0.109543181257991, 0.134633359654998, 0.147762112357376,
     0.147762112357376, 0.134633359654998, 0.109543181257991, 0.074725674575290, 0.033335672154344}};

/* multi-index tables, src/expan.c:39-58 (idx3) and :68-104 (fact, ifact3, sgn3):
 * graded order, within a degree i1 slowest then i2. */
static int gIdx3[OMAX + 2][OMAX + 2][OMAX + 2];
static double gFact[OMAX + 5], gIfact[OMAX + 5], gIfact3[NMOM(OMAX + 1)];
static int gSgn3[NMOM(OMAX + 1)];
static pthread_once_t gOnce = PTHREAD_ONCE_INIT;

static void initTables(void) {
  int i = 0, n, i1, i2, sgn = 1;
  gFact[0] = gIfact[0] = 1.0;
  for (int k = 1; k < OMAX + 5; k++) { gFact[k] = gFact[k - 1] * k; gIfact[k] = 1.0 / gFact[k]; }
  for (n = 0; n <= OMAX + 1; n++, sgn *= -1)
    for (i1 = 0; i1 <= n; i1++)
      for (i2 = 0; i2 <= n - i1; i2++, i++) {
        int i3 = n - i1 - i2;
        gIdx3[i1][i2][i3] = i;
        gIfact3[i] = 1.0 / (gFact[i1] * gFact[i2] * gFact[i3]);
        gSgn3[i] = sgn;
      }
}
static void tables(void) { pthread_once(&gOnce, initTables); }
#define IDX(a, b, c) gIdx3[a][b][c]

int fbo_nmom(int order) { return NMOM(order); }

void fbo_gl_table(int q, double *t, double *w) {
  for (int i = 0; i < q; i++) { t[i] = kT[q][i]; w[i] = kW[q][i]; }
}

/* --------------------------------------------------------------- kernels */
/* src/kernel.c:109-147 (kernelKER4); nrmX = target normal, nrmY = source */
void fbo_ker4(double kappa, double eps, const double *x, const double *nrmX, const double *nrmY, double *y) {
  double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
  double r = sqrt(r2);
  double expKa = exp(-kappa * r);
  double ip0 = nrmX[0] * nrmY[0] + nrmX[1] * nrmY[1] + nrmX[2] * nrmY[2];
  double ipX = nrmX[0] * x[0] + nrmX[1] * x[1] + nrmX[2] * x[2];
  double ipY = nrmY[0] * x[0] + nrmY[1] * x[1] + nrmY[2] * x[2];
  double G0 = 1.0 / r * kFourPiI;
  double Gk = expKa * G0;
  double coef = (kappa * r + 1.0) * expKa;
  double dG0dy = ipY * G0 / r2;
  double dGkdy = coef * dG0dy;
  double dG0dx = ipX * G0 / r2;
  double dGkdx = coef * dG0dx;
  double ddG0 = (ip0 * G0 - dG0dy * ipX * 3.0) / r2;
  double ddGk = coef * ddG0 - kappa * kappa * expKa * ipX * dG0dy;
  y[0] = G0 - Gk;
  y[1] = eps * dGkdy - dG0dy;
  y[2] = dGkdx / eps - dG0dx;
  y[3] = ddGk - ddG0;
}

/* src/kernel.c:84-104 (kernelDG0DGk): radial derivative functions */
static void radial(double kappa, double r, int p, double *G0, double *Gk) {
  double kappa2 = kappa * kappa, rexp = exp(-kappa * r), r1 = 1.0 / r, r2 = r1 * r1, nr2 = -r2;
  G0[0] = r1 * kFourPiI;
  Gk[0] = rexp * G0[0];
  G0[1] = nr2 * G0[0];
  Gk[1] = -Gk[0] * (r2 + kappa * r1);
  for (int k = 1; k < p; k++) {
    G0[k + 1] = (2 * k + 1) * nr2 * G0[k];
    Gk[k + 1] = r2 * (kappa2 * Gk[k - 1] - (2 * k + 1) * Gk[k]);
  }
}

/* src/expan.c:138-209 (setupDerivs); returns dG[0][*] up to degree `order` */
void fbo_derivs(double kappa, int order, const double *x, double *g0out, double *gkout) {
  double G0v[OMAX + 3], Gkv[OMAX + 3];
  static __thread double d0[OMAX + 2][NMOM(OMAX + 1)], dk[OMAX + 2][NMOM(OMAX + 1)];
  int p, p1, iRow, iRow1, i1, i2, i3, idx, idx1, idx2;
  tables();
  double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  radial(kappa, r, order, G0v, Gkv);
  for (p = 0; p <= order; p++) { d0[p][0] = G0v[p]; dk[p][0] = Gkv[p]; }
  for (p = 0; p < order; p++) {
    p1 = p + 1;
    d0[p][1] = d0[p1][0] * x[2]; d0[p][2] = d0[p1][0] * x[1]; d0[p][3] = d0[p1][0] * x[0];
    dk[p][1] = dk[p1][0] * x[2]; dk[p][2] = dk[p1][0] * x[1]; dk[p][3] = dk[p1][0] * x[0];
  }
  for (iRow = 2; iRow <= order; iRow++) {
    for (p = 0; p <= order - iRow; p++) {
      p1 = p + 1;
      idx = IDX(0, 0, iRow);
      iRow1 = iRow - 1;
      idx1 = IDX(0, 0, iRow1);
      idx2 = IDX(0, 0, iRow - 2);
      d0[p][idx] = d0[p1][idx1] * x[2] + d0[p1][idx2] * iRow1;
      dk[p][idx] = dk[p1][idx1] * x[2] + dk[p1][idx2] * iRow1;
      idx++;
      idx1 = IDX(0, 0, iRow1);
      d0[p][idx] = d0[p1][idx1] * x[1];
      dk[p][idx] = dk[p1][idx1] * x[1];
      idx++;
      for (i2 = 2; i2 <= iRow; i2++, idx++) {
        i3 = iRow - i2;
        idx1 = IDX(0, i2 - 1, i3);
        idx2 = IDX(0, i2 - 2, i3);
        d0[p][idx] = d0[p1][idx1] * x[1] + d0[p1][idx2] * (i2 - 1);
        dk[p][idx] = dk[p1][idx1] * x[1] + dk[p1][idx2] * (i2 - 1);
      }
      for (i2 = 0; i2 <= iRow1; i2++, idx++) {
        i3 = iRow1 - i2;
        idx1 = IDX(0, i2, i3);
        d0[p][idx] = d0[p1][idx1] * x[0];
        dk[p][idx] = dk[p1][idx1] * x[0];
      }
      for (i1 = 2; i1 <= iRow; i1++)
        for (i2 = 0; i2 <= iRow - i1; i2++, idx++) {
          i3 = iRow - i1 - i2;
          idx1 = IDX(i1 - 1, i2, i3);
          idx2 = IDX(i1 - 2, i2, i3);
          d0[p][idx] = d0[p1][idx1] * x[0] + d0[p1][idx2] * (i1 - 1);
          dk[p][idx] = dk[p1][idx1] * x[0] + dk[p1][idx2] * (i1 - 1);
        }
    }
  }
  memcpy(g0out, d0[0], sizeof(double) * NMOM(order));
  memcpy(gkout, dk[0], sizeof(double) * NMOM(order));
}

/* ------------------------------------------------------- Galerkin quadrature */
typedef struct {
  double kappa, eps;
  const double *nX, *nY;
} kctx;

static inline void K4(const kctx *k, const double *r, double *f) { fbo_ker4(k->kappa, k->eps, r, k->nX, k->nY, f); }

static void edges(const double *v, double a[3][3]) {   /* src/input.c:743-747 */
  for (int j = 0; j < 3; j++) {
    a[0][j] = v[6 + j] - v[3 + j];
    a[1][j] = v[j] - v[6 + j];
    a[2][j] = v[3 + j] - v[j];
  }
}

/* src/numQuad.c:312-353 (nrCommonVtx): exact vertex coordinate equality */
int fbo_common_vtx(const double *p, const double *q, int *idxX, int *idxY) {
  int n1, n2 = 0, flag = 0, cnt = 0;
  idxX[0] = idxY[0] = 0; idxX[1] = idxY[1] = 1; idxX[2] = idxY[2] = 2;
  if (p == q) return 3;
  for (n1 = 0; n1 < 3; n1++) {
    flag = 0;
    for (n2 = 0; n2 < 3; n2++)
      if (p[3 * n1] == q[3 * n2] && p[3 * n1 + 1] == q[3 * n2 + 1] && p[3 * n1 + 2] == q[3 * n2 + 2]) {
        flag = 1; cnt++; break;
      }
    if (flag) break;
  }
  if (!flag) return 0;
  idxX[0] = n1 % 3; idxX[1] = (n1 + 1) % 3; idxX[2] = (n1 + 2) % 3;
  idxY[0] = n2 % 3; idxY[1] = (n2 + 1) % 3; idxY[2] = (n2 + 2) % 3;
  const double *x1 = &p[3 * idxX[1]], *x2 = &p[3 * idxX[2]], *y1 = &q[3 * idxY[1]], *y2 = &q[3 * idxY[2]];
  if (x1[0] == y2[0] && x1[1] == y2[1] && x1[2] == y2[2]) cnt = 2;
  if (x2[0] == y1[0] && x2[1] == y1[1] && x2[2] == y1[2]) cnt = -2;
  return cnt;
}

/* src/numQuad.c:401-447 (pnlThr0): self panel, Sauter-Schwab */
static void pnlThr0(const kctx *K, int q, const double *a2, const double *a0, double *intgr) {
  const double *t = kT[q], *w = kW[q];
  double r1[3], r2[3], r3[3], r4[3], r5[3], r6[3], f1[4], f2[4], f3[4], f4[4], f5[4], f6[4], iy[4];
  for (int i = 0; i < 4; i++) intgr[i] = 0.0;
  for (int a = 0; a < q; a++)
    for (int b = 0; b < q; b++) {
      double omg1 = t[a] * (1.0 - t[b]);
      for (int c = 0; c < q; c++) {
        for (int i = 0; i < 4; i++) iy[i] = 0.0;
        for (int d = 0; d < q; d++) {
          double eta1 = t[d], fta1 = 1.0 - eta1;
          for (int k = 0; k < 3; k++) {
            r1[k] = omg1 * (eta1 * a2[k] + a0[k]);
            r2[k] = omg1 * (fta1 * a0[k] - eta1 * a2[k]);
            r3[k] = omg1 * (a2[k] + eta1 * a0[k]);
            r4[k] = -omg1 * (eta1 * a2[k] + a0[k]);
            r5[k] = omg1 * (eta1 * a2[k] - fta1 * a0[k]);
            r6[k] = -omg1 * (a2[k] + eta1 * a0[k]);
          }
          K4(K, r1, f1); K4(K, r2, f2); K4(K, r3, f3); K4(K, r4, f4); K4(K, r5, f5); K4(K, r6, f6);
          for (int i = 0; i < 4; i++) iy[i] += (f1[i] + f3[i] + f4[i] + f2[i] + f5[i] + f6[i]) * w[d];
        }
        for (int i = 0; i < 4; i++) intgr[i] += iy[i] * omg1 * (t[a] * t[a]) * t[b] * w[a] * w[b] * w[c];
      }
    }
}

/* src/numQuad.c:470-518 (pnlTwo0): common edge */
static void pnlTwo0(const kctx *K, int q, const double *a, const double *b, const double *c, double *intgr) {
  const double *t = kT[q], *w = kW[q];
  double r1[3], r2[3], r3[3], r4[3], r5[3], r6[3], f1[4], f2[4], f3[4], f4[4], f5[4], f6[4], iy[4];
  for (int i = 0; i < 4; i++) intgr[i] = 0.0;
  for (int ix = 0; ix < q; ix++)
    for (int jx = 0; jx < q; jx++) {
      double omg1 = t[ix] * t[jx];
      for (int i = 0; i < 4; i++) iy[i] = 0.0;
      for (int iyy = 0; iyy < q; iyy++) {
        double eta1 = t[iyy], fta1 = 1.0 - eta1;
        for (int jy = 0; jy < q; jy++) {
          double eta2 = t[jy], eta3 = eta1 * eta2, fta2 = 1.0 - eta2, fta3 = 1.0 - eta3;
          for (int k = 0; k < 3; k++) {
            r1[k] = omg1 * (eta1 * (a[k] + eta2 * b[k]) + fta1 * c[k]);
            r2[k] = omg1 * (eta1 * (b[k] + eta2 * a[k]) + fta3 * c[k]);
            r3[k] = omg1 * (fta1 * a[k] + b[k] + eta3 * c[k]);
            r4[k] = omg1 * (eta1 * (eta2 * b[k] - fta2 * a[k]) + c[k]);
            r5[k] = omg1 * (eta3 * b[k] - fta3 * a[k] + eta1 * c[k]);
            r6[k] = omg1 * (eta1 * b[k] - fta1 * a[k] + eta3 * c[k]);
          }
          K4(K, r1, f1); K4(K, r2, f2); K4(K, r3, f3); K4(K, r4, f4); K4(K, r5, f5); K4(K, r6, f6);
          for (int i = 0; i < 4; i++) {
            double intgrd = (f1[i] + f3[i] + f4[i] + f2[i] + f5[i] + f6[i]) * eta1;
            iy[i] += intgrd * w[iyy] * w[jy];
          }
        }
      }
      for (int i = 0; i < 4; i++) intgr[i] += iy[i] * t[jx] * (omg1 * omg1) * w[jx] * w[ix];
    }
}

/* src/numQuad.c:535-571 (pnlOne0): common vertex */
static void pnlOne0(const kctx *K, int q, const double *a2, const double *a0, const double *b2, const double *b0,
                    double *intgr) {
  const double *t = kT[q], *w = kW[q];
  double r1[3], r2[3], f1[4], f2[4], iy[4];
  for (int i = 0; i < 4; i++) intgr[i] = 0.0;
  for (int ix = 0; ix < q; ix++) {
    double omg = t[ix];
    for (int i = 0; i < 4; i++) iy[i] = 0.0;
    for (int a = 0; a < q; a++) {
      double eta1 = t[a];
      for (int b = 0; b < q; b++) {
        double eta2 = t[b], wij = w[a] * w[b] * eta1;
        for (int c = 0; c < q; c++) {
          double eta3 = t[c];
          for (int k = 0; k < 3; k++) {
            r1[k] = omg * (a2[k] + eta2 * a0[k] - eta1 * (b2[k] + eta3 * b0[k]));
            r2[k] = omg * (eta1 * (a2[k] + eta3 * a0[k]) - (b2[k] + eta2 * b0[k]));
          }
          K4(K, r1, f1); K4(K, r2, f2);
          for (int i = 0; i < 4; i++) iy[i] += (f1[i] + f2[i]) * w[c] * wij;
        }
      }
    }
    for (int i = 0; i < 4; i++) intgr[i] += iy[i] * (omg * omg) * omg * w[ix];
  }
}

/* src/numQuad.c:586-618 (pnlNil0): disjoint panels, Duffy squares */
static void pnlNil0(const kctx *K, int q, const double *r0, const double *ax2, const double *ax0, const double *ay1,
                    const double *ay2, double *intgr) {
  const double *t = kT[q], *w = kW[q];
  double r[3], r1[3], f[4], iy[4];
  for (int i = 0; i < 4; i++) intgr[i] = 0.0;
  for (int ix = 0; ix < q; ix++)
    for (int jx = 0; jx < q; jx++) {
      for (int k = 0; k < 3; k++) r1[k] = r0[k] + t[ix] * (ax2[k] + t[jx] * ax0[k]);
      for (int i = 0; i < 4; i++) iy[i] = 0.0;
      for (int iy0 = 0; iy0 < q; iy0++)
        for (int jy = 0; jy < q; jy++) {
          for (int k = 0; k < 3; k++) r[k] = r1[k] - t[iy0] * (ay1[k] + t[jy] * ay2[k]);
          K4(K, r, f);
          for (int i = 0; i < 4; i++) iy[i] += f[i] * t[iy0] * w[iy0] * w[jy];
        }
      for (int i = 0; i < 4; i++) intgr[i] += iy[i] * t[ix] * w[ix] * w[jx];
    }
}

/* src/numQuad.c:629-697 (panelIA0) */
void fbo_panel_ia0(const fbo_layout *L, int x, int y, double *out) {
  const double *vx = &L->vtx[9 * x], *vy = &L->vtx[9 * y];
  double ax[3][3], ay[3][3], in[4], a0[3], b0[3], r0[3];
  int idxX[3], idxY[3], nv;
  kctx K = {L->kappa, L->eps, &L->nrm[3 * x], &L->nrm[3 * y]};
  int q = L->qorder;
  edges(vx, ax);
  edges(vy, ay);
  nv = (x == y) ? 3 : fbo_common_vtx(vx, vy, idxX, idxY);
  if (x == y) { idxX[0] = idxY[0] = 0; idxX[1] = idxY[1] = 1; idxX[2] = idxY[2] = 2; }
  if (nv == 0) {
    for (int k = 0; k < 3; k++) r0[k] = vx[k] - vy[k];
    pnlNil0(&K, q, r0, ax[2], ax[0], ay[2], ay[0], in);
  } else if (nv == 1) {
    pnlOne0(&K, q, ax[idxX[2]], ax[idxX[0]], ay[idxY[2]], ay[idxY[0]], in);
  } else if (nv == 2) {
    pnlTwo0(&K, q, ax[idxX[2]], ax[idxX[0]], ay[idxY[0]], in);
  } else if (nv == -2) {
    for (int k = 0; k < 3; k++) { b0[k] = -ay[idxY[0]][k]; a0[k] = -ax[idxX[0]][k]; }
    pnlTwo0(&K, q, ax[idxX[1]], b0, a0, in);
  } else {
    pnlThr0(&K, q, ax[2], ax[0], in);
  }
  for (int i = 0; i < 4; i++) out[i] = 4.0 * L->area[x] * L->area[y] * in[i];
}

/* -------------------------------------------------------- moments (Q2M) */
/* src/moments.c:83-103 (setupConVect) */
static void conVect(int order, const double *tr, double *cv) {
  double f1[OMAX + 2], f2[OMAX + 2], f3[OMAX + 2];
  f1[0] = f2[0] = f3[0] = 1.0;
  for (int i = 1; i <= order; i++) { f1[i] = f1[i - 1] * tr[0]; f2[i] = f2[i - 1] * tr[1]; f3[i] = f3[i - 1] * tr[2]; }
  int i = 0;
  for (int n = 0; n <= order; n++)
    for (int i1 = 0; i1 <= n; i1++)
      for (int i2 = 0; i2 <= n - i1; i2++, i++) {
        int i3 = n - i1 - i2;
        cv[i] = f1[i1] * f2[i2] * f3[i3];
        cv[i] *= gIfact3[i];
      }
}

/* src/moments.c:56-78 (convolution): c_a = sum_{b+g=a} a_g b_b, overwrite */
static void convolve(int order, const double *a, const double *b, double *c) {
  int i = 0;
  for (int n = 0; n <= order; n++)
    for (int i1 = 0; i1 <= n; i1++)
      for (int i2 = 0; i2 <= n - i1; i2++, i++) {
        int i3 = n - i1 - i2;
        double tmp = 0.0;
        for (int k1 = 0; k1 <= i1; k1++)
          for (int k2 = 0; k2 <= i2; k2++)
            for (int k3 = 0; k3 <= i3; k3++) tmp += a[IDX(i1 - k1, i2 - k2, i3 - k3)] * b[IDX(k1, k2, k3)];
        c[i] = tmp;
      }
}

/* src/moments.c:148-200 (calcOneMoment0) and :115-133 (calcOneNoment) */
static void oneMoment(int order, const double *v, const double *nrm, double area, double *Mom, int job) {
  double c1[NMOM(OMAX)], c2[NMOM(OMAX)], buf[NMOM(OMAX)], a1[3], a2[3];
  int order1 = job == 0 ? order : order - 1;
  double area2 = 2.0 * area;
  for (int i = 0; i < 3; i++) { a1[i] = v[3 + i] - v[i]; a2[i] = v[6 + i] - v[i]; }
  conVect(order1, a1, c1);
  conVect(order1, a2, c2);
  int i = 0;
  for (int n = 0; n <= order1; n++)
    for (int i1 = 0; i1 <= n; i1++)
      for (int i2 = 0; i2 <= n - i1; i2++, i++) { c1[i] *= gFact[n]; c2[i] *= gFact[n]; }
  /* the reference convolves to `order` even for job 1; the degree-`order`
   * entries it forms there are never read, so stopping at order1 is identical */
  double *M = job == 0 ? Mom : buf;
  convolve(order1, c1, c2, M);
  i = 0;
  for (int n = 0; n <= order1; n++)
    for (int i1 = 0; i1 <= n; i1++)
      for (int i2 = 0; i2 <= n - i1; i2++, i++) M[i] *= area2 * gIfact[n + 2];
  if (job == 0) return;
  for (i = 0; i < NMOM(order); i++) Mom[i] = 0.0;
  i = 0;
  for (int n = 0; n < order; n++)
    for (int i1 = 0; i1 <= n; i1++)
      for (int i2 = 0; i2 <= n - i1; i2++, i++) {
        int i3 = n - i1 - i2;
        Mom[IDX(i1 + 1, i2, i3)] += nrm[0] * M[i];
        Mom[IDX(i1, i2 + 1, i3)] += nrm[1] * M[i];
        Mom[IDX(i1, i2, i3 + 1)] += nrm[2] * M[i];
      }
}

/* src/moments.c:216-236 (calcMoments0): column per panel, translated to the leaf centre */
static void leafMoments(const fbo_layout *L, int leaf, int order, int job, double *Q) {
  int nm = NMOM(order);
  double loc[NMOM(OMAX)], cv[NMOM(OMAX)], tr[3];
  const double *ctr = &L->cubeX[3 * leaf];
  for (int j = 0; j < L->leafCount[leaf]; j++) {
    int p = L->leafStart[leaf] + j;
    const double *v = &L->vtx[9 * p];
    oneMoment(order, v, &L->nrm[3 * p], L->area[p], loc, job);
    tr[0] = v[0] - ctr[0]; tr[1] = v[1] - ctr[1]; tr[2] = v[2] - ctr[2];
    conVect(order, tr, cv);
    convolve(order, cv, loc, &Q[(size_t)j * nm]);
  }
}

/* ----------------------------------------------------------- context */
typedef struct {
  int start, m;              /* first panel (idx order) and panel count of the cube */
  double *a;                 /* 2m x 2m, LU factors in LAPACK's column-major view */
  int *piv;
} pblock;

struct fbo_ctx {
  fbo_layout L;
  int nmMax;                 /* nMom(maxOrder) */
  double **q2m0, **q2m1;     /* per leaf, nMom(ordMom[depth]) x count, column per panel */
  double *mom, *lec;         /* per FMM cube: 2 / 4 arrays of nmMax */
  int precMode;              /* 3 diagonal, 2 block LU */
  double *dense;             /* -r=1/2: KER[4] of every (d, s) pair, row-major, or NULL */
  int nPB;
  pblock *pb;
};

fbo_ctx *fbo_create(const fbo_layout *L) {
  tables();
  if (L->maxOrder < 0 || L->maxOrder > OMAX - 2 || L->qorder < 1 || L->qorder > 10) return NULL;
  fbo_ctx *c = calloc(1, sizeof(*c));
  c->L = *L;
  if (c->L.nThreads < 1) c->L.nThreads = 1;
  c->nmMax = NMOM(L->maxOrder);
  c->q2m0 = calloc((size_t)L->nLeaf, sizeof(double *));
  c->q2m1 = calloc((size_t)L->nLeaf, sizeof(double *));
  int ord = L->ordMom[L->depth], nm = NMOM(ord);
  for (int l = 0; l < L->nLeaf; l++) {
    c->q2m0[l] = malloc(sizeof(double) * (size_t)nm * (size_t)(L->leafCount[l] ? L->leafCount[l] : 1));
    c->q2m1[l] = malloc(sizeof(double) * (size_t)nm * (size_t)(L->leafCount[l] ? L->leafCount[l] : 1));
    leafMoments(L, l, ord, 0, c->q2m0[l]);
    leafMoments(L, l, ord, 1, c->q2m1[l]);
  }
  c->mom = calloc((size_t)L->nCube * 2 * c->nmMax, sizeof(double));
  c->lec = calloc((size_t)L->nCube * 4 * c->nmMax, sizeof(double));
  c->precMode = 3;
  return c;
}

static void block_free(fbo_ctx *c);

void fbo_destroy(fbo_ctx *c) {
  if (!c) return;
  block_free(c);
  free(c->dense);
  for (int l = 0; l < c->L.nLeaf; l++) { free(c->q2m0[l]); free(c->q2m1[l]); }
  free(c->q2m0); free(c->q2m1); free(c->mom); free(c->lec); free(c);
}

const double *fbo_q2m(fbo_ctx *c, int leaf, int job) { return job ? c->q2m1[leaf] : c->q2m0[leaf]; }

void fbo_expansions(fbo_ctx *c, double *mom, double *lec) {
  if (mom) memcpy(mom, c->mom, sizeof(double) * (size_t)c->L.nCube * 2 * c->nmMax);
  if (lec) memcpy(lec, c->lec, sizeof(double) * (size_t)c->L.nCube * 4 * c->nmMax);
}

#define MOM(c, cube, k) (&(c)->mom[((size_t)(cube) * 2 + (k)) * (c)->nmMax])
#define LEC(c, cube, k) (&(c)->lec[((size_t)(cube) * 4 + (k)) * (c)->nmMax])

/* src/expan.c:342-382 (transM2MWs) with convM2M (:223-247): kid -> parent */
static void m2m(fbo_ctx *c, int kid, int par) {
  const fbo_layout *L = &c->L;
  int ordIn = L->ordMom[L->cubeLevel[kid]], ordOut = L->ordMom[L->cubeLevel[par]];
  int nIn = NMOM(ordIn), nOut = NMOM(ordOut);
  double tr[3], R[NMOM(OMAX)], b1[NMOM(OMAX)], b2[NMOM(OMAX)];
  for (int k = 0; k < 3; k++) tr[k] = L->cubeX[3 * kid + k] - L->cubeX[3 * par + k];
  conVect(ordOut, tr, R);
  memcpy(b1, MOM(c, kid, 0), sizeof(double) * nIn);
  memcpy(b2, MOM(c, kid, 1), sizeof(double) * nIn);
  for (int i = nIn; i < nOut; i++) b1[i] = b2[i] = 0.0;
  double *o1 = MOM(c, par, 0), *o2 = MOM(c, par, 1);
  int i = 0;
  for (int n = 0; n <= ordOut; n++)
    for (int i1 = 0; i1 <= n; i1++)
      for (int i2 = 0; i2 <= n - i1; i2++, i++) {
        int i3 = n - i1 - i2;
        double t1 = 0.0, t2 = 0.0;
        for (int k1 = 0; k1 <= i1; k1++)
          for (int k2 = 0; k2 <= i2; k2++)
            for (int k3 = 0; k3 <= i3; k3++) {
              double a = R[IDX(i1 - k1, i2 - k2, i3 - k3)];
              t1 += a * b1[IDX(k1, k2, k3)];
              t2 += a * b2[IDX(k1, k2, k3)];
            }
        o1[i] += t1;
        o2[i] += t2;
      }
}

/* src/expan.c:403-452 (transL2LWs) with convL2L (:295-326): parent -> kid */
static void l2l(fbo_ctx *c, int par, int kid) {
  const fbo_layout *L = &c->L;
  int ordIn = L->ordMom[L->cubeLevel[par]], ordOut = L->ordMom[L->cubeLevel[kid]];
  double tr[3], R[NMOM(OMAX)];
  for (int k = 0; k < 3; k++) tr[k] = L->cubeX[3 * kid + k] - L->cubeX[3 * par + k];
  conVect(ordIn, tr, R);
  for (int s = 0; s < 4; s++) {
    const double *In = LEC(c, par, s);
    double *Out = LEC(c, kid, s);
    int i = 0;
    for (int n = 0; n <= ordOut; n++)
      for (int i1 = 0; i1 <= n; i1++)
        for (int i2 = 0; i2 <= n - i1; i2++, i++) {
          int i3 = n - i1 - i2, j = 0;
          double t = 0.0;
          for (int m = 0; m <= ordIn - n; m++)
            for (int j1 = 0; j1 <= m; j1++)
              for (int j2 = 0; j2 <= m - j1; j2++, j++) {
                int j3 = m - j1 - j2;
                t += R[j] * In[IDX(i1 + j1, i2 + j2, i3 + j3)];
              }
          Out[i] += t;
        }
  }
}

/* src/fmm.c:774-787 + src/expan.c:471-494 (transM2L) with convM2L (:258-290) */
static void m2l(fbo_ctx *c, int src, int dst) {
  const fbo_layout *L = &c->L;
  int order = L->ordM2L[L->cubeLevel[dst]], nm = NMOM(order);
  double r[3], G0[NMOM(OMAX)], Gk[NMOM(OMAX)], c1[NMOM(OMAX)], c2[NMOM(OMAX)];
  double v1[NMOM(OMAX)], v2[NMOM(OMAX)], v3[NMOM(OMAX)];
  for (int k = 0; k < 3; k++) r[k] = L->cubeX[3 * dst + k] - L->cubeX[3 * src + k];
  fbo_derivs(L->kappa, order, r, G0, Gk);
  const double *mp = MOM(c, src, 0), *md = MOM(c, src, 1);
  for (int i = 0; i < nm; i++) {
    c1[i] = gSgn3[i] * mp[i];
    c2[i] = gSgn3[i] * md[i];
    v1[i] = G0[i] - Gk[i];
    v2[i] = L->eps * Gk[i] - G0[i];
    v3[i] = G0[i] - Gk[i] / L->eps;
  }
  double *l1 = LEC(c, dst, 0), *l2 = LEC(c, dst, 1), *l3 = LEC(c, dst, 2), *l4 = LEC(c, dst, 3);
  int i = 0;
  for (int n = 0; n <= order; n++)
    for (int i1 = 0; i1 <= n; i1++)
      for (int i2 = 0; i2 <= n - i1; i2++, i++) {
        int i3 = n - i1 - i2, j = 0;
        double t1 = 0.0, t2 = 0.0, t3 = 0.0, t4 = 0.0;
        for (int m = 0; m <= order - n; m++)
          for (int j1 = 0; j1 <= m; j1++)
            for (int j2 = 0; j2 <= m - j1; j2++, j++) {
              int j3 = m - j1 - j2, kk = IDX(i1 + j1, i2 + j2, i3 + j3);
              t1 += c2[j] * v1[kk];
              t2 += c1[j] * v2[kk];
              t3 += c2[j] * v3[kk];
              t4 += c1[j] * (-v1[kk]);
            }
        l1[i] += t1; l2[i] += t2; l3[i] += t3; l4[i] += t4;
      }
}

/* near apply over destination leaves [b, e): src/fmm.c:554-593 */
typedef struct { fbo_ctx *c; double alpha; const double *sgm; double *pot; int b, e; } nearTask;

static void *nearWork(void *arg) {
  nearTask *t = arg;
  const fbo_layout *L = &t->c->L;
  int n = L->n;
  for (int leaf = t->b; leaf < t->e; leaf++)
    for (int i = 0; i < L->leafCount[leaf]; i++) {
      int x = L->leafStart[leaf] + i;
      double sumPot = 0.0, sumDpdn = 0.0, K[4];
      for (int pr = L->nearOff[leaf]; pr < L->nearOff[leaf + 1]; pr++) {
        int s = L->nearSrc[pr];
        for (int j = 0; j < L->leafCount[s]; j++) {
          int y = L->leafStart[s] + j;
          double xp = t->sgm[y], xd = t->sgm[y + n];
          fbo_panel_ia0(L, x, y, K);
          sumPot += (K[0] * xd + K[1] * xp) * t->alpha;
          sumDpdn += (K[2] * xd + K[3] * xp) * t->alpha;
        }
      }
      t->pot[x] += sumPot;
      t->pot[x + n] += sumDpdn;
    }
  return NULL;
}

/* src/fmm.c:710-838 (applyFMM): pot = beta*pot + alpha*K*sgm */
void fbo_apply(fbo_ctx *c, double alpha, const double *sgm, double beta, double *pot, int parts) {
  const fbo_layout *L = &c->L;
  int n = L->n, depth = L->depth, height = L->height;
  memset(c->mom, 0, sizeof(double) * (size_t)L->nCube * 2 * c->nmMax);
  memset(c->lec, 0, sizeof(double) * (size_t)L->nCube * 4 * c->nmMax);
  int nm = NMOM(L->ordMom[depth]);
  if (parts & 1) {
    /* Q2M: mom_pot = Q2M1 sgm, mom_dpdn = Q2M0 sgm_n (fmm.c:745-753) */
    for (int l = 0; l < L->nLeaf; l++) {
      const double *Q1 = c->q2m1[l], *Q0 = c->q2m0[l];
      double *mp = MOM(c, l, 0), *md = MOM(c, l, 1);
      for (int j = 0; j < L->leafCount[l]; j++) {
        int p = L->leafStart[l] + j;
        for (int a = 0; a < nm; a++) {
          mp[a] += Q1[(size_t)j * nm + a] * sgm[p];
          md[a] += Q0[(size_t)j * nm + a] * sgm[p + n];
        }
      }
    }
    for (int lev = depth - 1; lev >= height; lev--)   /* M2M, fmm.c:758-764 */
      for (int cb = L->levStart[lev]; cb < L->levStart[lev] + L->levCount[lev]; cb++)
        for (int k = 0; k < L->nKids[cb]; k++) m2m(c, L->kidStart[cb] + k, cb);
    for (int pr = 0; pr < L->nM2L; pr++) m2l(c, L->m2lSrc[pr], L->m2lDst[pr]);
    for (int lev = height; lev < depth; lev++)        /* L2L, fmm.c:792-798 */
      for (int cb = L->levStart[lev]; cb < L->levStart[lev] + L->levCount[lev]; cb++)
        for (int k = 0; k < L->nKids[cb]; k++) l2l(c, cb, L->kidStart[cb] + k);
  }
  /* L2P (fmm.c:815-828): the first dgemv carries beta (beta == 0 clears) */
  for (int l = 0; l < L->nLeaf; l++) {
    const double *Q1 = c->q2m1[l], *Q0 = c->q2m0[l];
    const double *l1 = LEC(c, l, 0), *l2 = LEC(c, l, 1), *l3 = LEC(c, l, 2), *l4 = LEC(c, l, 3);
    for (int j = 0; j < L->leafCount[l]; j++) {
      int p = L->leafStart[l] + j;
      double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
      if (parts & 1)
        for (int a = 0; a < nm; a++) {
          s1 += Q0[(size_t)j * nm + a] * l1[a];
          s2 += Q0[(size_t)j * nm + a] * l2[a];
          s3 += Q1[(size_t)j * nm + a] * l3[a];
          s4 += Q1[(size_t)j * nm + a] * l4[a];
        }
      double y0 = beta == 0.0 ? 0.0 : beta * pot[p];
      double y1 = beta == 0.0 ? 0.0 : beta * pot[p + n];
      pot[p] = y0 + alpha * s1 + alpha * s2;
      pot[p + n] = y1 + alpha * s3 + alpha * s4;
    }
  }
  if (parts & 2) {   /* nearfield by contiguous destination-leaf ranges, fmm.c:630-677 */
    int T = L->nThreads < L->nLeaf ? L->nThreads : L->nLeaf;
    if (T < 1) T = 1;
    nearTask *tk = calloc((size_t)T, sizeof(nearTask));
    pthread_t *th = calloc((size_t)T, sizeof(pthread_t));
    for (int t = 0; t < T; t++) {
      tk[t].c = c; tk[t].alpha = alpha; tk[t].sgm = sgm; tk[t].pot = pot;
      tk[t].b = (int)(((long long)L->nLeaf * t) / T);
      tk[t].e = (int)(((long long)L->nLeaf * (t + 1)) / T);
    }
    if (T == 1) nearWork(&tk[0]);
    else {
      for (int t = 0; t < T; t++) pthread_create(&th[t], NULL, nearWork, &tk[t]);
      for (int t = 0; t < T; t++) pthread_join(th[t], NULL);
    }
    free(tk); free(th);
  }
}

/* -r=1/2: the dense direct operator (src/direct_backend.c:57-78 builds every
 * pair's panelIA0, :172-208 applies it row by row in panel order) */
int fbo_set_direct(fbo_ctx *c, int on) {
  const fbo_layout *L = &c->L;
  free(c->dense);
  c->dense = NULL;
  if (!on) return 0;
  const size_t n = (size_t)L->n;
  c->dense = malloc(sizeof(double) * 4 * n * n);
  if (!c->dense) return -1;
  for (size_t d = 0; d < n; d++)
    for (size_t s = 0; s < n; s++) fbo_panel_ia0(L, (int)d, (int)s, &c->dense[4 * (d * n + s)]);
  return 0;
}

static void direct_apply(fbo_ctx *c, double alpha, const double *sgm, double beta, double *pot) {
  const int n = c->L.n;
  for (int d = 0; d < n; d++) {
    const double *k = &c->dense[4 * (size_t)d * n];
    double sp = 0.0, sd = 0.0;
    for (int s = 0; s < n; s++) {
      sp += k[4 * s] * sgm[s + n] + k[4 * s + 1] * sgm[s];
      sd += k[4 * s + 2] * sgm[s + n] + k[4 * s + 3] * sgm[s];
    }
    pot[d] = beta * pot[d] + alpha * sp;
    pot[d + n] = beta * pot[d + n] + alpha * sd;
  }
}

/* src/fabipb.c:1237-1287 (MtVmain) */
void fbo_matvec(fbo_ctx *c, double alpha, const double *x, double beta, double *y) {
  const fbo_layout *L = &c->L;
  int n = L->n;
  double s1 = (1.0 + L->eps) / 2.0 * alpha, s2 = (1.0 + 1.0 / L->eps) / 2.0 * alpha;
  if (c->dense) direct_apply(c, alpha, x, -beta, y);
  else fbo_apply(c, alpha, x, -beta, y, 3);
  for (int i = 0; i < n; i++) {
    y[i] = (s1 * L->area[i] * x[i] - y[i]);
    y[i + n] = s2 * L->area[i] * x[i + n] - y[i + n];
  }
}

/* src/precond_fmm.c:664-728 (PtVfmmDiagonal); in-place safe */
void fbo_psolve_diag(fbo_ctx *c, double *out, const double *in) {
  const fbo_layout *L = &c->L;
  int n = L->n;
  double s1 = (1.0 + L->eps) / 2.0, s2 = (1.0 + 1.0 / L->eps) / 2.0;
  for (int i = 0; i < n; i++) {
    out[i] = in[i] / (s1 * L->area[i]);
    out[n + i] = in[n + i] / (s2 * L->area[i]);
  }
}

/* ----------------------------------------------- block preconditioner */
/* src/precond_fmm.c:553-652 (setupPreconditioning), :147-268 (block assembly),
 * :861-953 (PtVfmmCachedLU).  One dense block per cube at level depth-1 over
 * its m panels, unknowns [sigma(m) | sigma_n(m)].  The reference fills the
 * array B row-major, B[i][j] from KER = panelIA0(X = panel j, Y = panel i):
 *   B[i][j] = -KER1, B[i][j+m] = -KER0, B[i+m][j] = -KER3, B[i+m][j+m] = -KER2,
 *   B[i][i] += s1 area_i, B[i+m][i+m] += s2 area_i,
 * and hands it to dgetrf/dgetrs ('N'), which read it column-major: the
 * operator solved is A(r, c) = B[c][r].  LU with partial pivoting as dgetf2
 * (first largest pivot, row swaps over the whole row). */
static void block_free(fbo_ctx *c) {
  for (int b = 0; b < c->nPB; b++) { free(c->pb[b].a); free(c->pb[b].piv); }
  free(c->pb);
  c->pb = NULL;
  c->nPB = 0;
}

int fbo_set_precond(fbo_ctx *c, int mode) {
  const fbo_layout *L = &c->L;
  block_free(c);
  if (mode == 3) { c->precMode = 3; return 0; }
  if (mode < 0 || mode > 2) return -1;
  const int lev = L->depth - 1;
  if (lev < L->height) return -1;
  const double s1 = (1.0 + L->eps) / 2.0, s2 = (1.0 + 1.0 / L->eps) / 2.0;
  c->nPB = L->levCount[lev];
  c->pb = calloc((size_t)c->nPB, sizeof(pblock));
  for (int b = 0; b < c->nPB; b++) {
    const int cb = L->levStart[lev] + b, k0 = L->kidStart[cb];
    pblock *P = &c->pb[b];
    P->start = L->leafStart[k0];
    P->m = 0;
    for (int k = 0; k < L->nKids[cb]; k++) P->m += L->leafCount[k0 + k];
    const int m = P->m, M = 2 * m;
    P->a = calloc((size_t)M * (size_t)M, sizeof(double));
    P->piv = calloc((size_t)M, sizeof(int));
    double *B = P->a, ker[4];
    for (int i = 0; i < m; i++) {
      for (int j = 0; j < m; j++) {
        fbo_panel_ia0(L, P->start + j, P->start + i, ker);
        B[(size_t)i * M + j] = -ker[1];
        B[(size_t)i * M + j + m] = -ker[0];
        B[(size_t)(i + m) * M + j] = -ker[3];
        B[(size_t)(i + m) * M + j + m] = -ker[2];
      }
      B[(size_t)i * M + i] += s1 * L->area[P->start + i];
      B[(size_t)(i + m) * M + i + m] += s2 * L->area[P->start + i];
    }
    /* LU of the column-major view A(r, c) = B[c*M + r] */
#define A_(r, cc) B[(size_t)(cc) * M + (r)]
    for (int j = 0; j < M; j++) {
      int p = j;
      double best = fabs(A_(j, j));
      for (int r = j + 1; r < M; r++)
        if (fabs(A_(r, j)) > best) { best = fabs(A_(r, j)); p = r; }
      P->piv[j] = p;
      if (A_(p, j) == 0.0) return -2;
      if (p != j)
        for (int cc = 0; cc < M; cc++) { double t = A_(j, cc); A_(j, cc) = A_(p, cc); A_(p, cc) = t; }
      const double inv = 1.0 / A_(j, j);
      for (int r = j + 1; r < M; r++) A_(r, j) *= inv;
      for (int cc = j + 1; cc < M; cc++) {
        const double u = A_(j, cc);
        if (u != 0.0)
          for (int r = j + 1; r < M; r++) A_(r, cc) -= A_(r, j) * u;
      }
    }
#undef A_
  }
  c->precMode = 2;
  return 0;
}

/* the block solves of PtVfmmCachedLU; in-place safe */
static void psolve_block(fbo_ctx *c, double *out, const double *in) {
  const int n = c->L.n;
  for (int b = 0; b < c->nPB; b++) {
    const pblock *P = &c->pb[b];
    const int m = P->m, M = 2 * m;
    const double *B = P->a;
    double *x = malloc(sizeof(double) * (size_t)M);
    for (int i = 0; i < m; i++) { x[i] = in[P->start + i]; x[i + m] = in[n + P->start + i]; }
#define A_(r, cc) B[(size_t)(cc) * M + (r)]
    for (int j = 0; j < M; j++)
      if (P->piv[j] != j) { double t = x[j]; x[j] = x[P->piv[j]]; x[P->piv[j]] = t; }
    for (int j = 0; j < M; j++)
      if (x[j] != 0.0)
        for (int r = j + 1; r < M; r++) x[r] -= x[j] * A_(r, j);
    for (int j = M - 1; j >= 0; j--) {
      if (x[j] != 0.0) {
        x[j] /= A_(j, j);
        for (int r = 0; r < j; r++) x[r] -= x[j] * A_(r, j);
      }
    }
#undef A_
    for (int i = 0; i < m; i++) { out[P->start + i] = x[i]; out[n + P->start + i] = x[i + m]; }
    free(x);
  }
}

void fbo_psolve(fbo_ctx *c, double *out, const double *in) {
  if (c->precMode == 2) psolve_block(c, out, in);
  else fbo_psolve_diag(c, out, in);
}

/* ------------------------------------------------------------- GMRES */
static double dot(int n, const double *a, const double *b) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}
static double nrm2(int n, const double *a) { return sqrt(dot(n, a, a)); }
/* BLAS drotg: (a, b) -> c, s with r = sign(roe) * sqrt(a^2 + b^2) */
static void rotg(double a, double b, double *c, double *s) {
  double roe = fabs(a) > fabs(b) ? a : b, scale = fabs(a) + fabs(b);
  if (scale == 0.0) { *c = 1.0; *s = 0.0; return; }
  double r = scale * sqrt((a / scale) * (a / scale) + (b / scale) * (b / scale));
  if (roe < 0) r = -r;
  *c = a / r; *s = b / r;
}
static void rot(double *x, double *y, double c, double s) {
  double t = c * *x + s * *y;
  *y = c * *y - s * *x;
  *x = t;
}

/* src/gmres.c:153-373 and :527-621 (serial path; n < 2^20 there) */
int fbo_gmres(fbo_ctx *c, const double *b, double *x, int restrt, double tol, int maxit, int *iters,
              double *resid, double *hist) {
  int n = 2 * c->L.n, ldh = restrt + 1, i, k, nh = 0;
  double *V = malloc(sizeof(double) * (size_t)n * (size_t)(restrt + 1));
  double *r = malloc(sizeof(double) * (size_t)n), *w = malloc(sizeof(double) * (size_t)n);
  double *H = calloc((size_t)ldh * (size_t)(restrt + 2), sizeof(double));
  double *s = calloc((size_t)restrt + 1, sizeof(double)), *cs = calloc((size_t)restrt, sizeof(double));
  double *sn = calloc((size_t)restrt, sizeof(double)), *y = calloc((size_t)restrt + 1, sizeof(double));
  int info = 0;
  *iters = 0;
  memcpy(w, b, sizeof(double) * n);
  if (nrm2(n, x) != 0.0) fbo_matvec(c, -1.0, x, 1.0, w);
  fbo_psolve(c, r, w);
  double bnrm2 = nrm2(n, b);
  if (bnrm2 == 0.0) bnrm2 = 1.0;
  *resid = nrm2(n, r) / bnrm2;
  if (*resid < tol) goto out;
  for (;;) {
    i = 0;
    double rn = nrm2(n, r);
    if (!(rn > 0.0) || rn != rn) { *resid = rn != rn ? rn : 0.0; info = rn != rn ? 2 : 0; goto out; }
    for (k = 0; k < n; k++) V[k] = r[k] * (1.0 / rn);
    for (k = 0; k <= restrt; k++) s[k] = 0.0;
    s[0] = rn;
    for (;;) {
      ++i; ++*iters;
      fbo_matvec(c, 1.0, &V[(size_t)(i - 1) * n], 0.0, w);
      fbo_psolve(c, w, w);
      double *h = &H[(size_t)(i - 1) * ldh];
      for (k = 0; k < i; k++) {          /* modified Gram-Schmidt */
        h[k] = dot(n, w, &V[(size_t)k * n]);
        for (int q = 0; q < n; q++) w[q] += -h[k] * V[(size_t)k * n + q];
      }
      h[i] = nrm2(n, w);
      memcpy(&V[(size_t)i * n], w, sizeof(double) * n);
      if (h[i] > 0.0) {
        double sc = 1.0 / h[i];
        for (int q = 0; q < n; q++) V[(size_t)i * n + q] *= sc;
      }
      for (k = 0; k < i - 1; k++) rot(&h[k], &h[k + 1], cs[k], sn[k]);
      rotg(h[i - 1], h[i], &cs[i - 1], &sn[i - 1]);
      rot(&h[i - 1], &h[i], cs[i - 1], sn[i - 1]);
      rot(&s[i - 1], &s[i], cs[i - 1], sn[i - 1]);
      *resid = fabs(s[i]) / bnrm2;
      if (hist) hist[nh++] = *resid;
      if (*resid <= tol || *iters == maxit || i >= restrt) break;
    }
    /* update: y = H(1:i,1:i) \ s, x += V y (gmres.c:364-373) */
    for (k = i - 1; k >= 0; k--) {
      double t = s[k];
      for (int q = k + 1; q < i; q++) t -= H[(size_t)q * ldh + k] * y[q];
      y[k] = t / H[(size_t)k * ldh + k];
    }
    for (k = 0; k < i; k++)
      for (int q = 0; q < n; q++) x[q] += y[k] * V[(size_t)k * n + q];
    if (*resid <= tol) goto out;
    memcpy(w, b, sizeof(double) * n);
    fbo_matvec(c, -1.0, x, 1.0, w);
    fbo_psolve(c, r, w);
    *resid = nrm2(n, r) / bnrm2;
    if (hist) hist[nh++] = *resid;
    if (*resid <= tol) goto out;
    if (*iters == maxit) { info = 1; goto out; }
  }
out:
  free(V); free(r); free(w); free(H); free(s); free(cs); free(sn); free(y);
  return info;
}

/* ------------------------------------------------------------- RHS */
/* src/fabipb.c:789-798 + src/numQuad.c:786-819 (panelRHS) + kernel.c:152-167 */
void fbo_rhs_direct(fbo_ctx *c, int nChar, const double *pos, const double *chr, double *b) {
  const fbo_layout *L = &c->L;
  int n = L->n, q = L->qorder;
  const double *t = kT[q], *w = kW[q];
  double fac = kFourPiI / L->eps1;
  for (int i = 0; i < n; i++) {
    const double *v = &L->vtx[9 * i], *nx = &L->nrm[3 * i];
    double a[3][3];
    edges(v, a);
    b[i] = 0.0; b[n + i] = 0.0;
    for (int j = 0; j < nChar; j++) {
      double r0[3], r[3], I0 = 0.0, I1 = 0.0;
      for (int k = 0; k < 3; k++) r0[k] = v[k] - pos[3 * j + k];
      for (int ix = 0; ix < q; ix++)
        for (int jx = 0; jx < q; jx++) {
          for (int k = 0; k < 3; k++) r[k] = r0[k] + t[ix] * (a[2][k] + t[jx] * a[0][k]);
          double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2], ri = 1.0 / sqrt(r2), r3i = ri / r2;
          double ip = nx[0] * r[0] + nx[1] * r[1] + nx[2] * r[2];
          I0 += ri * t[ix] * w[ix] * w[jx];
          I1 += (-ip * r3i) * t[ix] * w[ix] * w[jx];
        }
      I0 *= 2.0 * L->area[i];
      I1 *= 2.0 * L->area[i];
      b[i] += chr[j] * I0;
      b[n + i] += chr[j] * I1;
    }
    b[i] *= fac;
    b[n + i] *= fac;
  }
}

/* ------------------------------------------------------------- energy */
/* src/chargeTree.c:117-129 (rhsChargeExpansionOrder), default policy */
static int ctOrder(const fbo_layout *L, int lev) {
  int o = L->ordMom[lev] > 4 ? L->ordMom[lev] : 4;
  return o < L->maxOrder ? o : L->maxOrder;
}

/* src/chargeTree.c:313-416 (computeChgLeafMoments, transM2Mchr, computeChgMoments) */
void fbo_ctree_moments(fbo_ctx *c, const fbo_ctree *T, double *mom) {
  const fbo_layout *L = &c->L;
  int nm = c->nmMax;
  double cv[NMOM(OMAX)], R[NMOM(OMAX)], tr[3];
  memset(mom, 0, sizeof(double) * (size_t)T->nc * nm);
  for (int cb = T->levStart[T->depth]; cb < T->levStart[T->depth] + T->levCount[T->depth]; cb++) {
    int order = ctOrder(L, T->depth), m = NMOM(order);
    for (int k = 0; k < T->nChgs[cb]; k++) {
      int j = T->chgIdx[T->chgStart[cb] + k];
      for (int d = 0; d < 3; d++) tr[d] = T->pos[3 * j + d] - T->x[3 * cb + d];
      conVect(order, tr, cv);
      for (int i = 0; i < m; i++) mom[(size_t)cb * nm + i] += T->chr[j] * cv[i];
    }
  }
  for (int lev = T->depth - 1; lev >= L->height; lev--)
    for (int cb = T->levStart[lev]; cb < T->levStart[lev] + T->levCount[lev]; cb++)
      for (int q = 0; q < T->nKids[cb]; q++) {
        int kid = T->kidStart[cb] + q;
        int ordIn = ctOrder(L, T->level[kid]), ordOut = ctOrder(L, lev), nIn = NMOM(ordIn);
        for (int d = 0; d < 3; d++) tr[d] = T->x[3 * kid + d] - T->x[3 * cb + d];
        conVect(ordOut, tr, R);
        const double *mi = &mom[(size_t)kid * nm];
        double *mo = &mom[(size_t)cb * nm];
        int i = 0;
        for (int n = 0; n <= ordOut; n++)
          for (int i1 = 0; i1 <= n; i1++)
            for (int i2 = 0; i2 <= n - i1; i2++, i++) {
              int i3 = n - i1 - i2;
              double tmp = 0.0;
              for (int k1 = 0; k1 <= i1; k1++)
                for (int k2 = 0; k2 <= i2; k2++)
                  for (int k3 = 0; k3 <= i3; k3++) {
                    int bi = IDX(k1, k2, k3);
                    if (bi < nIn) tmp += R[IDX(i1 - k1, i2 - k2, i3 - k3)] * mi[bi];
                  }
              mo[i] += tmp;
            }
      }
}

typedef struct {
  fbo_ctx *c;
  const fbo_ctree *T;
  const double *mom;
  double g0[NMOM(OMAX)], gk[NMOM(OMAX)];
} ewalk;

/* src/treecode.c:617-661 (energyTreeWalk) with :581-615 (chgClusterEnergyEval) */
static void energyWalk(ewalk *W, int cb, const double *qp, const double *nrm, double *y) {
  const fbo_layout *L = &W->c->L;
  const fbo_ctree *T = W->T;
  double r[3];
  for (int k = 0; k < 3; k++) r[k] = qp[k] - T->x[3 * cb + k];
  double dist = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  int lev = T->level[cb];
  if (T->eRad[cb] < T->theta * dist && lev >= L->height) {
    int order = ctOrder(L, lev);
    fbo_derivs(L->kappa, order + 1, r, W->g0, W->gk);
    const double *mom = &W->mom[(size_t)cb * W->c->nmMax];
    double y0 = 0.0, y1 = 0.0;
    int i = 0;
    for (int n = 0; n <= order; n++)
      for (int i1 = 0; i1 <= n; i1++)
        for (int i2 = 0; i2 <= n - i1; i2++, i++) {
          int i3 = n - i1 - i2;
          y0 += gSgn3[i] * mom[i] * (W->g0[i] - W->gk[i]);
          double dnG0 = nrm[0] * W->g0[IDX(i1 + 1, i2, i3)] + nrm[1] * W->g0[IDX(i1, i2 + 1, i3)] +
                        nrm[2] * W->g0[IDX(i1, i2, i3 + 1)];
          double dnGk = nrm[0] * W->gk[IDX(i1 + 1, i2, i3)] + nrm[1] * W->gk[IDX(i1, i2 + 1, i3)] +
                        nrm[2] * W->gk[IDX(i1, i2, i3 + 1)];
          y1 += gSgn3[i] * mom[i] * (L->eps * dnGk - dnG0);
        }
    y[0] += y0;
    y[1] += y1;
    return;
  }
  if (lev == T->depth) {
    for (int i = 0; i < T->nChgs[cb]; i++) {
      int j = T->chgIdx[T->chgStart[cb] + i];
      double x[3];
      for (int k = 0; k < 3; k++) x[k] = qp[k] - T->pos[3 * j + k];
      double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2], rn = sqrt(r2);
      double G0 = kFourPiI / rn, Gk = exp(-L->kappa * rn) * G0;
      double coef = (L->kappa * rn + 1.0) * exp(-L->kappa * rn);
      double ip = nrm[0] * x[0] + nrm[1] * x[1] + nrm[2] * x[2];
      double dG0dn = -ip * G0 / r2, dGkdn = coef * dG0dn;
      y[0] += T->chr[j] * (G0 - Gk);
      y[1] += T->chr[j] * (L->eps * dGkdn - dG0dn);
    }
    return;
  }
  for (int k = 0; k < T->nKids[cb]; k++) energyWalk(W, T->kidStart[cb] + k, qp, nrm, y);
}

/* src/treecode.c:381-416 (rhsTreeWalkLocal) with :349-379 (chgClusterEvalLocal):
 * the Coulomb kernel only; the derivative tensors of 1/(4 pi r) are those of
 * setupCoulombDerivsLocal (:283-347), the G0 half of setupDerivs */
static void rhsWalk(ewalk *W, int cb, const double *qp, const double *nrm, double theta, double *y) {
  const fbo_layout *L = &W->c->L;
  const fbo_ctree *T = W->T;
  double r[3];
  for (int k = 0; k < 3; k++) r[k] = qp[k] - T->x[3 * cb + k];
  double dist = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  int lev = T->level[cb];
  if (T->eRad[cb] < theta * dist && lev >= L->height) {
    int order = ctOrder(L, lev);
    fbo_derivs(L->kappa, order + 1, r, W->g0, W->gk);
    const double *mom = &W->mom[(size_t)cb * W->c->nmMax];
    double y0 = 0.0, y1 = 0.0;
    int i = 0;
    for (int n = 0; n <= order; n++)
      for (int i1 = 0; i1 <= n; i1++)
        for (int i2 = 0; i2 <= n - i1; i2++, i++) {
          int i3 = n - i1 - i2;
          y0 += gSgn3[i] * mom[i] * W->g0[i];
          double dn = nrm[0] * W->g0[IDX(i1 + 1, i2, i3)] + nrm[1] * W->g0[IDX(i1, i2 + 1, i3)] +
                      nrm[2] * W->g0[IDX(i1, i2, i3 + 1)];
          y1 += gSgn3[i] * mom[i] * dn;
        }
    y[0] += y0 / kFourPiI;
    y[1] += y1 / kFourPiI;
    return;
  }
  if (lev == T->depth) {
    for (int i = 0; i < T->nChgs[cb]; i++) {
      int j = T->chgIdx[T->chgStart[cb] + i];
      double x[3];
      for (int k = 0; k < 3; k++) x[k] = qp[k] - T->pos[3 * j + k];
      double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2], ri = 1.0 / sqrt(r2), r3i = ri / r2;
      double ip = nrm[0] * x[0] + nrm[1] * x[1] + nrm[2] * x[2];
      y[0] += T->chr[j] * ri;
      y[1] += T->chr[j] * (-ip * r3i);
    }
    return;
  }
  for (int k = 0; k < T->nKids[cb]; k++) rhsWalk(W, T->kidStart[cb] + k, qp, nrm, theta, y);
}

/* src/fabipb.c:633-649 (rhsTreeSetupWorker) with src/treecode.c:418-450
 * (panelRHSTreeWorkspace): b = fac * (2 A sum_q t w w y), fac = 1/(4 pi eps1) */
void fbo_rhs_tree(fbo_ctx *c, const fbo_ctree *T, double theta, double *b) {
  const fbo_layout *L = &c->L;
  int n = L->n, q = L->qorder;
  const double *t = kT[q], *w = kW[q];
  const double fac = kFourPiI / L->eps1;
  ewalk *W = calloc(1, sizeof(ewalk));
  double *mom = malloc(sizeof(double) * (size_t)T->nc * c->nmMax);
  fbo_ctree_moments(c, T, mom);
  W->c = c; W->T = T; W->mom = mom;
  for (int p = 0; p < n; p++) {
    const double *v = &L->vtx[9 * p];
    double a[3][3], qp[3], y[2], o0 = 0.0, o1 = 0.0;
    edges(v, a);
    for (int ix = 0; ix < q; ix++)
      for (int jx = 0; jx < q; jx++) {
        for (int k = 0; k < 3; k++) qp[k] = v[k] + t[ix] * (a[2][k] + t[jx] * a[0][k]);
        y[0] = y[1] = 0.0;
        rhsWalk(W, 0, qp, &L->nrm[3 * p], theta, y);
        o0 += y[0] * t[ix] * w[ix] * w[jx];
        o1 += y[1] * t[ix] * w[ix] * w[jx];
      }
    o0 *= 2.0 * L->area[p];
    o1 *= 2.0 * L->area[p];
    b[p] = fac * o0;
    b[p + n] = fac * o1;
  }
  free(mom);
  free(W);
}

/* src/treecode.c:663-695 (panelEnergyTree) and :869-1017 (the panel sum) */
double fbo_energy_tree(fbo_ctx *c, const fbo_ctree *T, const double *sgm) {
  const fbo_layout *L = &c->L;
  int n = L->n, q = L->qorder;
  const double *t = kT[q], *w = kW[q];
  ewalk *W = calloc(1, sizeof(ewalk));
  double *mom = malloc(sizeof(double) * (size_t)T->nc * c->nmMax);
  fbo_ctree_moments(c, T, mom);
  W->c = c; W->T = T; W->mom = mom;
  double sum = 0.0;
  for (int p = 0; p < n; p++) {
    const double *v = &L->vtx[9 * p];
    double a[3][3], qp[3], y[2], o0 = 0.0, o1 = 0.0;
    edges(v, a);
    for (int ix = 0; ix < q; ix++)
      for (int jx = 0; jx < q; jx++) {
        for (int k = 0; k < 3; k++) qp[k] = v[k] + t[ix] * (a[2][k] + t[jx] * a[0][k]);
        y[0] = y[1] = 0.0;
        energyWalk(W, 0, qp, &L->nrm[3 * p], y);
        o0 += y[0] * t[ix] * w[ix] * w[jx];
        o1 += y[1] * t[ix] * w[ix] * w[jx];
      }
    o0 *= 2.0 * L->area[p];
    o1 *= 2.0 * L->area[p];
    sum += o0 * sgm[p + n] + o1 * sgm[p];
  }
  free(mom);
  free(W);
  return sum;
}
