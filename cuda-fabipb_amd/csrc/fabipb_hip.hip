// fabipb_hip.hip - MI355X (gfx950) implementation of FABIPB's Galerkin BIE
// matvec, right-hand side, solvation energy and GMRES, behind the flat C ABI
// of include/fabipb_hip.h.
//
// Device layout (HBM, structure of arrays, panels in leaf depth-first order):
//   pnl[n]          22 doubles per panel: vertices, edges a0=v2-v1 a1=v0-v2
//                   a2=v1-v0 (src/input.c:743-747), normal, area
//   Q0,Q1[n][nmL]   leaf Q2M matrices, one contiguous column per panel
//                   (Q2M0 moments, Q2M1 noments; L2P uses the same, fmm.c:457)
//   mom[nCube][2][nm], lec[nCube][4][nm]   multipole / local expansions
//   coef            nearfield Galerkin coefficients KER[4] per (dst, src)
//                   panel pair, row-major per destination leaf: for leaf D the
//                   rows are its panels, the columns the concatenated panels
//                   of D's near leaves (srcList), so a row is one contiguous
//                   stream of 32-byte records
//
// Kernels (FP64 throughout):
//   k_q2m_mats   setup: leaf moment/noment columns, one wave per panel
//   k_near_build setup: KER4 Galerkin quadrature per pair (disjoint Duffy rule
//                and the Sauter-Schwab vertex/edge/self rules), one block per
//                destination panel
//   k_near_apply one block per 16-row tile of a destination leaf; the
//                sources' densities are staged in LDS once per tile, each wave
//                streams four rows of coefficients (HBM-bound)
//   k_q2m        four waves split a leaf's panels, lanes over multi-indices
//   k_m2m, k_l2l thread per multi-index, translation tables in LDS
//   k_m2l_cached per destination cube, (alpha, beta) terms balanced over the
//                lanes by a host plan, derivative tensors cached per pair
//                (k_m2l recomputes them when the cache does not fit)
//   k_l2p        eight lanes per panel
//   the nearfield runs on a second stream beside the far-field chain
//   k_combine    y = alpha*(D x - far - near) + beta*y
//   k_rhs        direct right-hand side, thread per panel, charges in LDS
//   k_energy     panel-tree energy, one wave per panel walking the charge tree
//   GMRES helpers: fixed-grid deterministic dot products, MGS update,
//                normalisation, diagonal preconditioner, x += V y
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/fabipb_hip.h"

#define NMOM(p) (((p) + 1) * ((p) + 2) * ((p) + 3) / 6)
#define MAXP 12           // largest expansion order supported (nMom(13) = 560)
#define NEAR_ROWS 16      // rows per nearfield apply tile
#define NEAR_CHUNK 2048   // source columns staged in LDS per pass (32 KiB)
#define RED_BLOCKS 256    // fixed grid of the deterministic reductions
#define M2L_CACHE_BYTES (16.0 * 1024 * 1024 * 1024)   // cache M2L derivative tensors up to 16 GiB

static char gErr[1024];
extern "C" const char *fbh_last_error(void) { return gErr; }
static void setErr(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(gErr, sizeof(gErr), fmt, ap);
  va_end(ap);
}
#define HCK(call)                                                                  \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      setErr("%s failed at %s:%d: %s", #call, __FILE__, __LINE__, hipGetErrorString(e_)); \
      return 0;                                                                    \
    }                                                                              \
  } while (0)

static const double kFourPiI = 1.0 / (4 * M_PI);

// Gauss-Legendre table on [0,1], orders 1..10 (the reference's 15-digit table,
// src/numQuad.c:221-294).
static const double hT[11][10] = {
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
static const double hW[11][10] = {
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

struct Pnl {
  double v[9];   // v0 v1 v2
  double a[9];   // a0 a1 a2
  double nr[3];
  double area;
};

struct QTab {     // quadrature table for one order, passed by value
  double t[10], w[10];
};

// ------------------------------------------------------------------ kernels

// The Galerkin quadrature, the moment columns, the right-hand side and the
// energy walk keep the reference's operation order: no fused multiply-add
// contraction there (FMA rounding shows up amplified in the cancellations of
// the self-panel and near-singular terms).  The streaming kernels below
// (nearfield apply, FMM passes) contract freely.
#pragma clang fp contract(off)

// src/kernel.c:109-147 (kernelKER4): nX target normal, nY source normal
__device__ __forceinline__ void ker4(double kappa, double eps, const double *x, const double *nX,
                                     const double *nY, double *y) {
  double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
  double r = sqrt(r2);
  double expKa = exp(-kappa * r);
  double ip0 = nX[0] * nY[0] + nX[1] * nY[1] + nX[2] * nY[2];
  double ipX = nX[0] * x[0] + nX[1] * x[1] + nX[2] * x[2];
  double ipY = nY[0] * x[0] + nY[1] * x[1] + nY[2] * x[2];
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

// d/dkappa of the four kernelKER4 components (the Galerkin rules are linear
// in the kernel, so the same rules give d KER / d kappa):
//   d(G0 - Gk) = e^{-kappa r} / (4 pi),  d coef = -kappa r^2 e^{-kappa r},
//   d y1 = eps dcoef dG0dy,  d y2 = dcoef dG0dx / eps,
//   d y3 = dcoef ddG0 - kappa e^{-kappa r} (2 - kappa r) (n_x.r) dG0dy
__device__ __forceinline__ void ker4dk(double kappa, double eps, const double *x, const double *nX,
                                       const double *nY, double *y) {
  double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
  double r = sqrt(r2);
  double expKa = exp(-kappa * r);
  double ip0 = nX[0] * nY[0] + nX[1] * nY[1] + nX[2] * nY[2];
  double ipX = nX[0] * x[0] + nX[1] * x[1] + nX[2] * x[2];
  double ipY = nY[0] * x[0] + nY[1] * x[1] + nY[2] * x[2];
  double G0 = 1.0 / r * kFourPiI;
  double dG0dy = ipY * G0 / r2;
  double dG0dx = ipX * G0 / r2;
  double ddG0 = (ip0 * G0 - dG0dy * ipX * 3.0) / r2;
  double dcoef = -kappa * r2 * expKa;
  y[0] = expKa * kFourPiI;
  y[1] = eps * dcoef * dG0dy;
  y[2] = dcoef * dG0dx / eps;
  y[3] = dcoef * ddG0 - kappa * expKa * (2.0 - kappa * r) * ipX * dG0dy;
}

struct KC {
  double kappa, eps;
  const double *nX, *nY;
  bool dk;   // integrate d ker4 / d kappa instead of ker4
};

__device__ __forceinline__ void kerK(const KC &K, const double *x, double *y) {
  if (K.dk) ker4dk(K.kappa, K.eps, x, K.nX, K.nY, y);
  else ker4(K.kappa, K.eps, x, K.nX, K.nY, y);
}

// src/numQuad.c:586-618 (pnlNil0): disjoint panels, Duffy squares
__device__ void dNil0(const KC &K, const QTab &Q, int q, const double *r0, const double *ax2, const double *ax0,
                      const double *ay1, const double *ay2, double *in) {
  double r[3], r1[3], f[4], iy[4];
  for (int i = 0; i < 4; i++) in[i] = 0.0;
  for (int ix = 0; ix < q; ix++)
    for (int jx = 0; jx < q; jx++) {
      for (int k = 0; k < 3; k++) r1[k] = r0[k] + Q.t[ix] * (ax2[k] + Q.t[jx] * ax0[k]);
      for (int i = 0; i < 4; i++) iy[i] = 0.0;
      for (int a = 0; a < q; a++)
        for (int b = 0; b < q; b++) {
          for (int k = 0; k < 3; k++) r[k] = r1[k] - Q.t[a] * (ay1[k] + Q.t[b] * ay2[k]);
          kerK(K, r, f);
          for (int i = 0; i < 4; i++) iy[i] += f[i] * Q.t[a] * Q.w[a] * Q.w[b];
        }
      for (int i = 0; i < 4; i++) in[i] += iy[i] * Q.t[ix] * Q.w[ix] * Q.w[jx];
    }
}

// src/numQuad.c:535-571 (pnlOne0): one common vertex
__device__ void dOne0(const KC &K, const QTab &Q, int q, const double *a2, const double *a0, const double *b2,
                      const double *b0, double *in) {
  double r1[3], r2[3], f1[4], f2[4], iy[4];
  for (int i = 0; i < 4; i++) in[i] = 0.0;
  for (int ix = 0; ix < q; ix++) {
    double omg = Q.t[ix];
    for (int i = 0; i < 4; i++) iy[i] = 0.0;
    for (int a = 0; a < q; a++) {
      double eta1 = Q.t[a];
      for (int b = 0; b < q; b++) {
        double eta2 = Q.t[b], wij = Q.w[a] * Q.w[b] * eta1;
        for (int c = 0; c < q; c++) {
          double eta3 = Q.t[c];
          for (int k = 0; k < 3; k++) {
            r1[k] = omg * (a2[k] + eta2 * a0[k] - eta1 * (b2[k] + eta3 * b0[k]));
            r2[k] = omg * (eta1 * (a2[k] + eta3 * a0[k]) - (b2[k] + eta2 * b0[k]));
          }
          kerK(K, r1, f1);
          kerK(K, r2, f2);
          for (int i = 0; i < 4; i++) iy[i] += (f1[i] + f2[i]) * Q.w[c] * wij;
        }
      }
    }
    for (int i = 0; i < 4; i++) in[i] += iy[i] * (omg * omg) * omg * Q.w[ix];
  }
}

// src/numQuad.c:470-518 (pnlTwo0): one common edge
__device__ void dTwo0(const KC &K, const QTab &Q, int q, const double *a, const double *b, const double *c,
                      double *in) {
  double r[6][3], f[4], iy[4];
  for (int i = 0; i < 4; i++) in[i] = 0.0;
  for (int ix = 0; ix < q; ix++)
    for (int jx = 0; jx < q; jx++) {
      double omg1 = Q.t[ix] * Q.t[jx];
      for (int i = 0; i < 4; i++) iy[i] = 0.0;
      for (int ia = 0; ia < q; ia++) {
        double eta1 = Q.t[ia], fta1 = 1.0 - eta1;
        for (int jy = 0; jy < q; jy++) {
          double eta2 = Q.t[jy], eta3 = eta1 * eta2, fta2 = 1.0 - eta2, fta3 = 1.0 - eta3;
          for (int k = 0; k < 3; k++) {
            r[0][k] = omg1 * (eta1 * (a[k] + eta2 * b[k]) + fta1 * c[k]);
            r[1][k] = omg1 * (eta1 * (b[k] + eta2 * a[k]) + fta3 * c[k]);
            r[2][k] = omg1 * (fta1 * a[k] + b[k] + eta3 * c[k]);
            r[3][k] = omg1 * (eta1 * (eta2 * b[k] - fta2 * a[k]) + c[k]);
            r[4][k] = omg1 * (eta3 * b[k] - fta3 * a[k] + eta1 * c[k]);
            r[5][k] = omg1 * (eta1 * b[k] - fta1 * a[k] + eta3 * c[k]);
          }
          // fcn1+fcn3+fcn4+fcn2+fcn5+fcn6, the reference's summation order
          double s[4] = {0, 0, 0, 0};
          const int ord[6] = {0, 2, 3, 1, 4, 5};
          for (int m = 0; m < 6; m++) {
            kerK(K, r[ord[m]], f);
            for (int i = 0; i < 4; i++) s[i] = (m == 0) ? f[i] : s[i] + f[i];
          }
          for (int i = 0; i < 4; i++) iy[i] += (s[i] * eta1) * Q.w[ia] * Q.w[jy];
        }
      }
      for (int i = 0; i < 4; i++) in[i] += iy[i] * Q.t[jx] * (omg1 * omg1) * Q.w[jx] * Q.w[ix];
    }
}

// src/numQuad.c:401-447 (pnlThr0): the panel with itself
__device__ void dThr0(const KC &K, const QTab &Q, int q, const double *a2, const double *a0, double *in) {
  double r[6][3], f[4], iy[4];
  for (int i = 0; i < 4; i++) in[i] = 0.0;
  for (int a = 0; a < q; a++)
    for (int b = 0; b < q; b++) {
      double omg1 = Q.t[a] * (1.0 - Q.t[b]);
      for (int c = 0; c < q; c++) {
        for (int i = 0; i < 4; i++) iy[i] = 0.0;
        for (int d = 0; d < q; d++) {
          double eta1 = Q.t[d], fta1 = 1.0 - eta1;
          for (int k = 0; k < 3; k++) {
            r[0][k] = omg1 * (eta1 * a2[k] + a0[k]);
            r[1][k] = omg1 * (fta1 * a0[k] - eta1 * a2[k]);
            r[2][k] = omg1 * (a2[k] + eta1 * a0[k]);
            r[3][k] = -omg1 * (eta1 * a2[k] + a0[k]);
            r[4][k] = omg1 * (eta1 * a2[k] - fta1 * a0[k]);
            r[5][k] = -omg1 * (a2[k] + eta1 * a0[k]);
          }
          double s[4] = {0, 0, 0, 0};
          const int ord[6] = {0, 2, 3, 1, 4, 5};
          for (int m = 0; m < 6; m++) {
            kerK(K, r[ord[m]], f);
            for (int i = 0; i < 4; i++) s[i] = (m == 0) ? f[i] : s[i] + f[i];
          }
          for (int i = 0; i < 4; i++) iy[i] += s[i] * Q.w[d];
        }
        for (int i = 0; i < 4; i++) in[i] += iy[i] * omg1 * (Q.t[a] * Q.t[a]) * Q.t[b] * Q.w[a] * Q.w[b] * Q.w[c];
      }
    }
}

__device__ __forceinline__ bool veq(const double *a, const double *b) {
  return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}

// src/numQuad.c:312-353 (nrCommonVtx): exact coordinate equality
__device__ int commonVtx(const double *p, const double *q, int *ix, int *iy) {
  int n1, n2 = 0, cnt = 0;
  bool flag = false;
  ix[0] = iy[0] = 0; ix[1] = iy[1] = 1; ix[2] = iy[2] = 2;
  for (n1 = 0; n1 < 3; n1++) {
    flag = false;
    for (n2 = 0; n2 < 3; n2++)
      if (veq(&p[3 * n1], &q[3 * n2])) { flag = true; cnt++; break; }
    if (flag) break;
  }
  if (!flag) return 0;
  ix[0] = n1 % 3; ix[1] = (n1 + 1) % 3; ix[2] = (n1 + 2) % 3;
  iy[0] = n2 % 3; iy[1] = (n2 + 1) % 3; iy[2] = (n2 + 2) % 3;
  if (veq(&p[3 * ix[1]], &q[3 * iy[2]])) cnt = 2;
  if (veq(&p[3 * ix[2]], &q[3 * iy[1]])) cnt = -2;
  return cnt;
}

// src/numQuad.c:629-697 (panelIA0): KER = 4 A_x A_y * (rule for the case)
__device__ void panelIA0(const Pnl &X, const Pnl &Y, bool same, double kappa, double eps, const QTab &Q, int q,
                         double *out, bool dk = false) {
  KC K = {kappa, eps, X.nr, Y.nr, dk};
  int ix[3], iy[3];
  int nv = same ? 3 : commonVtx(X.v, Y.v, ix, iy);
  double in[4];
  if (nv == 0) {
    double r0[3] = {X.v[0] - Y.v[0], X.v[1] - Y.v[1], X.v[2] - Y.v[2]};
    dNil0(K, Q, q, r0, &X.a[6], &X.a[0], &Y.a[6], &Y.a[0], in);
  } else if (nv == 1) {
    dOne0(K, Q, q, &X.a[3 * ix[2]], &X.a[3 * ix[0]], &Y.a[3 * iy[2]], &Y.a[3 * iy[0]], in);
  } else if (nv == 2) {
    dTwo0(K, Q, q, &X.a[3 * ix[2]], &X.a[3 * ix[0]], &Y.a[3 * iy[0]], in);
  } else if (nv == -2) {
    double b0[3], a0[3];
    for (int k = 0; k < 3; k++) { b0[k] = -Y.a[3 * iy[0] + k]; a0[k] = -X.a[3 * ix[0] + k]; }
    dTwo0(K, Q, q, &X.a[3 * ix[1]], b0, a0, in);
  } else {
    dThr0(K, Q, q, &X.a[6], &X.a[0], in);
  }
  for (int i = 0; i < 4; i++) out[i] = 4.0 * X.area * Y.area * in[i];
}

// ----------------------------------------------------------- nearfield

// one block per destination panel; threads over its source columns
__global__ __launch_bounds__(256) void k_near_build(const Pnl *__restrict__ pnl, const int *__restrict__ leafOf,
                                                    const int *__restrict__ leafStart, const int *__restrict__ srcOff,
                                                    const int *__restrict__ ncolArr,
                                                    const long long *__restrict__ coefOff,
                                                    const int *__restrict__ srcList, double4 *__restrict__ coef,
                                                    double kappa, double eps, QTab Q, int q) {
  int x = blockIdx.x;
  int D = leafOf[x];
  int row = x - leafStart[D];
  int ncol = ncolArr[D];
  const Pnl X = pnl[x];
  double4 *out = coef + coefOff[D] + (long long)row * ncol;
  const int *src = srcList + srcOff[D];
  for (int j = threadIdx.x; j < ncol; j += blockDim.x) {
    int y = src[j];
    const Pnl Y = pnl[y];
    double k[4];
    panelIA0(X, Y, x == y, kappa, eps, Q, q, k);
    out[j] = make_double4(k[0], k[1], k[2], k[3]);
  }
}

// the stored coefficient blocks of the streamed tiles (tileCoef[t] >= 0):
// one block per (tile, row), threads over the row's source columns
// TR: the transposed store, (KER3, KER1, KER2, KER0) of panelIA0(X = y, Y = x)
template <bool TR = false>
__global__ __launch_bounds__(256) void k_near_build_tiles(const Pnl *__restrict__ pnl, const int2 *__restrict__ tiles,
                                                          const int *__restrict__ leafStart,
                                                          const int *__restrict__ leafCount,
                                                          const int *__restrict__ srcOff,
                                                          const int *__restrict__ ncolArr,
                                                          const long long *__restrict__ tileCoef,
                                                          const int *__restrict__ srcList, double4 *__restrict__ coef,
                                                          double kappa, double eps, QTab Q, int q) {
  const int t = blockIdx.x, row = blockIdx.y;
  const long long off = tileCoef[t];
  if (off < 0) return;
  const int2 tl = tiles[t];
  const int D = tl.x, r0 = tl.y;
  if (row >= min(NEAR_ROWS, leafCount[D] - r0)) return;
  const int x = leafStart[D] + r0 + row, ncol = ncolArr[D];
  const Pnl X = pnl[x];
  double4 *out = coef + off + (long long)row * ncol;
  const int *src = srcList + srcOff[D];
  for (int j = threadIdx.x; j < ncol; j += blockDim.x) {
    const int y = src[j];
    const Pnl Y = pnl[y];
    double k[4];
    if (TR) {
      panelIA0(Y, X, x == y, kappa, eps, Q, q, k);
      out[j] = make_double4(k[3], k[1], k[2], k[0]);
    } else {
      panelIA0(X, Y, x == y, kappa, eps, Q, q, k);
      out[j] = make_double4(k[0], k[1], k[2], k[3]);
    }
  }
}

// one pair's coefficients (a query of the stage ABI)
__global__ void k_near_pair(const Pnl *__restrict__ pnl, int x, int y, double kappa, double eps, QTab Q, int q,
                            double *__restrict__ out) {
  double k[4];
  panelIA0(pnl[x], pnl[y], x == y, kappa, eps, Q, q, k);
  for (int i = 0; i < 4; i++) out[i] = k[i];
}

// Galerkin coefficients of listed panel pairs, one thread per pair:
// (k0..k3)[p] = panelIA0(X = pnl[xi[p]], Y = pnl[yi[p]]) with the rule its
// geometric case selects (the assembly hooks of src/precond_fmm.c:223, 496)
__global__ __launch_bounds__(256) void k_pair_coeffs(const Pnl *__restrict__ pnl, long long nPairs,
                                                     const int *__restrict__ xi, const int *__restrict__ yi,
                                                     double kappa, double eps, QTab Q, int q, double *__restrict__ k0,
                                                     double *__restrict__ k1, double *__restrict__ k2,
                                                     double *__restrict__ k3) {
  const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (p >= nPairs) return;
  const int x = xi[p], y = yi[p];
  double k[4];
  panelIA0(pnl[x], pnl[y], x == y, kappa, eps, Q, q, k);
  k0[p] = k[0];
  k1[p] = k[1];
  k2[p] = k[2];
  k3[p] = k[3];
}

// ------------------------------------------------- block preconditioner
// Leaf-parent blocks (src/precond_fmm.c:147-268, 553-652): block b covers the
// m panels [start, start+m) of one cube at level depth-1, unknowns
// [sigma | sigma_n].  The reference fills B[i][j] from panelIA0(X = j, Y = i)
// and LAPACK factors B read column-major, i.e. the operator
//   M(r, c) = B[c][r]:  M(r, c) = -KER1(r, c), M(r, c+m) = -KER3(r, c),
//   M(r+m, c) = -KER0(r, c), M(r+m, c+m) = -KER2(r, c), KER = panelIA0(X = r, Y = c),
//   plus s1 area_r, s2 area_r on the diagonal.
// k_bp_assemble writes M row-major into a group buffer padded to Mp with the
// identity; the batched LU and inverse read that column-major (as M^T), so the
// inverse they leave, read row-major, is M^-1.
__global__ __launch_bounds__(256) void k_bp_assemble(const Pnl *__restrict__ pnl, const int *__restrict__ grpBlock,
                                                     const int *__restrict__ bStart, const int *__restrict__ bM,
                                                     int Mp, double *__restrict__ T, double kappa, double eps,
                                                     double s1, double s2, QTab Q, int q) {
  const int g = blockIdx.x, r = blockIdx.y, b = grpBlock[g];
  const int m = bM[b], start = bStart[b];
  double *A = T + (long long)g * Mp * Mp;
  if (r >= m) {   // padding rows (r >= 2m is reached through r - m below)
    return;
  }
  const Pnl X = pnl[start + r];
  for (int c = threadIdx.x; c < m; c += blockDim.x) {
    const Pnl Y = pnl[start + c];
    double k[4];
    panelIA0(X, Y, r == c, kappa, eps, Q, q, k);
    A[(long long)r * Mp + c] = -k[1] + (r == c ? s1 * X.area : 0.0);
    A[(long long)r * Mp + c + m] = -k[3];
    A[(long long)(r + m) * Mp + c] = -k[0];
    A[(long long)(r + m) * Mp + c + m] = -k[2] + (r == c ? s2 * X.area : 0.0);
  }
}

// identity on the padding of every block of a group (rows and columns >= 2m)
__global__ void k_bp_pad(const int *__restrict__ grpBlock, const int *__restrict__ bM, int Mp, double *__restrict__ T) {
  const int g = blockIdx.x, M = 2 * bM[grpBlock[g]];
  double *A = T + (long long)g * Mp * Mp;
  for (int r = M + threadIdx.x; r < Mp; r += blockDim.x) A[(long long)r * Mp + r] = 1.0;
}

// the M x M inverse out of the padded group buffer into the packed store
__global__ void k_bp_unpad(const int *__restrict__ grpBlock, const int *__restrict__ bM,
                           const long long *__restrict__ bOff, int Mp, const double *__restrict__ T,
                           double *__restrict__ inv) {
  const int g = blockIdx.x, b = grpBlock[g], M = 2 * bM[b];
  const double *A = T + (long long)g * Mp * Mp;
  double *o = inv + bOff[b];
  for (int r = blockIdx.y; r < M; r += gridDim.y)
    for (int c = threadIdx.x; c < M; c += blockDim.x) o[(long long)r * M + c] = A[(long long)r * Mp + c];
}

// the transposed inverse of every block (for the adjoint solve)
__global__ void k_bp_transpose(const int *__restrict__ bM, const long long *__restrict__ bOff,
                               const double *__restrict__ inv, double *__restrict__ invT) {
  const int b = blockIdx.x, M = 2 * bM[b];
  const double *a = inv + bOff[b];
  double *t = invT + bOff[b];
  for (int r = blockIdx.y; r < M; r += gridDim.y)
    for (int c = threadIdx.x; c < M; c += blockDim.x) t[(long long)c * M + r] = a[(long long)r * M + c];
}

// y = M^-1 x per block, row-major inverse streamed once: tile = (block, first
// row), 4 waves x 4 rows, the block's x gathered into LDS
__global__ __launch_bounds__(256) void k_bp_apply(const int2 *__restrict__ tiles, const int *__restrict__ bStart,
                                                  const int *__restrict__ bM, const long long *__restrict__ bOff,
                                                  const double *__restrict__ inv, const double *__restrict__ x,
                                                  double *__restrict__ y, int n) {
  __shared__ double sx[NEAR_CHUNK];
  const int2 t = tiles[blockIdx.x];
  const int b = t.x, r0 = t.y, m = bM[b], M = 2 * m, start = bStart[b];
  const int rows = min(16, M - r0);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double acc[4] = {0, 0, 0, 0};
  for (int c0 = 0; c0 < M; c0 += NEAR_CHUNK) {
    const int cn = min(NEAR_CHUNK, M - c0);
    __syncthreads();
    for (int j = threadIdx.x; j < cn; j += 256) {
      const int cc = c0 + j;
      sx[j] = cc < m ? x[start + cc] : x[n + start + cc - m];
    }
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
      const int row = w + 4 * rr;
      if (row < rows) {
        const double *a = inv + bOff[b] + (long long)(r0 + row) * M + c0;
        double s = 0.0;
        for (int j = lane; j < cn; j += 64) s += a[j] * sx[j];
        acc[rr] += s;
      }
    }
  }
#pragma unroll
  for (int rr = 0; rr < 4; rr++) {
    const int row = w + 4 * rr;
    double v = acc[rr];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (row < rows && lane == 0) {
      const int r = r0 + row;
      y[r < m ? start + r : n + start + r - m] = v;
    }
  }
}

extern __shared__ __align__(16) unsigned char dynLds[];   // dynamic LDS of the templated kernels

#pragma clang fp contract(fast)
__device__ __forceinline__ double waveSum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <typename T>
__device__ __forceinline__ T waveSumT(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <typename R> struct Vec4;
template <> struct Vec4<double> { typedef double4 type; };
template <> struct Vec4<float> { typedef float4 type; };

// tile = (destination leaf, first row); 4 waves x 4 rows; out[x] = sum over
// sources of K0*sgm_n + K1*sgm, out[x+n] = K2*sgm_n + K3*sgm.  R is the
// arithmetic of the coefficients, the staged densities and the sums.  The
// tile's coefficient rows start at tileCoef[t] (row stride = the leaf's
// column count).
template <typename R>
__device__ __forceinline__ void nearStreamTile(int t, const int2 *__restrict__ tiles,
                                               const int *__restrict__ leafStart, const int *__restrict__ leafCount,
                                               const int *__restrict__ srcOff, const int *__restrict__ ncolArr,
                                               const long long *__restrict__ tileCoef,
                                               const int *__restrict__ srcList,
                                               const typename Vec4<R>::type *__restrict__ coef,
                                               const double *__restrict__ x, R *__restrict__ out, int n, R *sS,
                                               R *sN) {
  typedef typename Vec4<R>::type R4;
  const int2 tl = tiles[t];
  const int D = tl.x, r0 = tl.y;
  const int rows = min(NEAR_ROWS, leafCount[D] - r0);
  const int ncol = ncolArr[D];
  const int *src = srcList + srcOff[D];
  const R4 *cbase = coef + tileCoef[t];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  R a0[4] = {0, 0, 0, 0}, a1[4] = {0, 0, 0, 0};
  for (int c0 = 0; c0 < ncol; c0 += NEAR_CHUNK) {
    const int cn = min(NEAR_CHUNK, ncol - c0);
    __syncthreads();
    for (int j = threadIdx.x; j < cn; j += 256) {
      int y = src[c0 + j];
      sS[j] = (R)x[y];
      sN[j] = (R)x[y + n];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = w + 4 * r;
      if (row < rows) {
        const R4 *cr = cbase + (long long)row * ncol + c0;
        R s0 = 0, s1 = 0;
        int j = lane;
        for (; j + 64 < cn; j += 128) {
          R4 c = cr[j], d = cr[j + 64];
          R p = sS[j], m = sN[j], p2 = sS[j + 64], m2 = sN[j + 64];
          s0 += c.x * m + c.y * p;
          s1 += c.z * m + c.w * p;
          s0 += d.x * m2 + d.y * p2;
          s1 += d.z * m2 + d.w * p2;
        }
        if (j < cn) {
          R4 c = cr[j];
          R p = sS[j], m = sN[j];
          s0 += c.x * m + c.y * p;
          s1 += c.z * m + c.w * p;
        }
        a0[r] += s0;
        a1[r] += s1;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int row = w + 4 * r;
    R s0 = waveSumT(a0[r]), s1 = waveSumT(a1[r]);
    if (row < rows && lane == 0) {
      int p = leafStart[D] + r0 + row;
      out[p] = s0;
      out[p + n] = s1;
    }
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_near_apply(const int2 *__restrict__ tiles, const int *__restrict__ leafStart,
                                                    const int *__restrict__ leafCount,
                                                    const int *__restrict__ srcOff, const int *__restrict__ ncolArr,
                                                    const long long *__restrict__ tileCoef,
                                                    const int *__restrict__ srcList,
                                                    const typename Vec4<R>::type *__restrict__ coef,
                                                    const double *__restrict__ x, R *__restrict__ out, int n) {
  __shared__ R sS[NEAR_CHUNK], sN[NEAR_CHUNK];
  nearStreamTile<R>(blockIdx.x, tiles, leafStart, leafCount, srcOff, ncolArr, tileCoef, srcList, coef, x, out, n,
                    sS, sN);
}

// ------------------------------------------- FP64 nearfield computed on the fly
// The plain form's nearfield without the coefficient stream: at q = 1 every
// source but the panel itself by KER = A_x A_y ker4(p_x - p_y) in FP64, plus
// the stored corrections (rule minus formula) of the self / shared-vertex
// pairs.  Same tiles as k_near_apply; sources (p, A sigma), (n, A sigma_n)
// staged in LDS.  The values differ from the stored coefficients only by the
// rounding of the formula's operation order.
#define FLY64_CHUNK 512
#define FLY64_RB 2   // rows per pass over the staged sources
#define FLY64_OWNSPLIT 1   // the self-pair selects only on the tile's own leaf columns
// TR: the transposed operator's row x, i.e. KER(y, x) read transposed: with
// r, n_x, n_y as below, y0 and y3 are symmetric in (x, y) and
//   y1(y,x) = -(eps coef - 1) dG0dx,  y2(y,x) = -(coef/eps - 1) dG0dy,
//   out[x] += y3 A sgm_n + y1(y,x) A sgm,  out[x+n] += y2(y,x) A sgm_n + y0 A sgm
// 1/sqrt(x) for x > 0 (a normal FP64 number): the hardware estimate (relative
// error 2^-24 measured on gfx950), then one third-order step
// y <- y (1 + e/2 + 3e^2/8), e = 1 - x y^2, which cubes the error (measured
// worst case 1.4e-16 over 4M arguments from 1e-4 to 1e4, as two Newton steps).
__device__ __forceinline__ double rsq64(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-(x * y), y, 1.0);
  return fma(y, e * fma(e, 0.375, 0.5), y);
}

// e^x for x <= 0 by the table-driven method of Tang (ACM TOMS 15 (1989) 144-157):
// x = (256 m + j) ln2/256 + r with |r| <= ln2/512 (ln2/256 split in two so that
// k ln2/256 is exact), e^x = 2^m 2^(j/256) e^r, e^r from its Taylor polynomial
// through r^4 (truncation r^5/120 < 4e-17), 2^(j/256) from the table T.  x is
// clamped at -700, so the result stays a normal number and nothing overflows.
__device__ __forceinline__ double expNeg(double x, const double *__restrict__ T) {
  x = fmax(x, -700.0);
  const double kf = rint(x * 369.3299304675746);                       // 256 / ln2
  const double r = fma(-kf, -0x1.718432a1b0e26p-43, fma(-kf, 0x1.62e42ff000000p-9, x));
  const int k = (int)kf;
  const double q = fma(r, fma(r, fma(r, fma(r, 1.0 / 24.0, 1.0 / 6.0), 0.5), 1.0), 1.0);
  return ldexp(T[k & 255] * q, k >> 8);
}

// 2^(j/256), j = 0..255, the table of expNeg: filled once per device at set-up
// (k_exp2_tab) and copied into LDS by every nearfield block
static __device__ double gExp2Tab[256];
__global__ void k_exp2_tab() {
  const int j = threadIdx.x;
  if (j < 256) gExp2Tab[j] = exp2((double)j / 256.0);
}

// One pair of the on-the-fly FP64 nearfield (kernelKER4 at the one-point rule,
// src/kernel.c:109-147): target point/normal X, XN against staged source
// P = (point, A sigma), N = (normal, A sigma_n).  skip: the panel itself (its
// weights are zeroed and r2 replaced by 1 so nothing divides by zero).
// CHK: the source may be the panel itself (the tile's own leaf columns).
// Every term is linear in G0 = 1/(4 pi r), so the sums are kept with G0
// replaced by 1/r and the row's total is scaled by 1/(4 pi) once; with
// a = ipX dG0dy shared by ddG0 and the kappa^2 term, and the accumulations as
// FMA chains, a pair is 60 FP64/integer VALU instructions.
template <bool TR, bool CHK>
__device__ __forceinline__ void fly64Pair(const double4 &X, const double4 &XN, const double4 &P, const double4 &N,
                                          bool skip, double kappa, double kap2, double eps, double ieps,
                                          const double *__restrict__ T, double &s0, double &s1) {
  const double rx = X.x - P.x, ry = X.y - P.y, rz = X.z - P.z;
  const double r2 = CHK && skip ? 1.0 : rx * rx + ry * ry + rz * rz;
  const double ri = rsq64(r2), rn = r2 * ri, ri2 = ri * ri, ri3 = ri * ri2;
  const double kr = kappa * rn;
  const double e = expNeg(-kr, T);
  const double ip0 = XN.x * N.x + XN.y * N.y + XN.z * N.z;
  const double ipX = XN.x * rx + XN.y * ry + XN.z * rz;
  const double ipY = N.x * rx + N.y * ry + N.z * rz;
  // 4 pi times: G0, dG0/dn_y, dG0/dn_x, d2G0/dn_x dn_y
  const double coef = (kr + 1.0) * e;
  const double dy = ipY * ri3, dx = ipX * ri3;
  const double a = ipX * dy;
  const double ddG = fma(a * ri2, -3.0, ip0 * ri3);
  const double y0 = fma(-ri, e, ri);
  const double y3 = fma(coef, ddG, -(kap2 * e) * a) - ddG;
  const double pw = CHK && skip ? 0.0 : P.w, nw = CHK && skip ? 0.0 : N.w;
  if constexpr (TR) {
    const double y1 = fma(-eps, coef * dx, dx);
    const double y2 = fma(-(coef * dy), ieps, dy);
    s0 = fma(y3, nw, fma(y1, pw, s0));
    s1 = fma(y2, nw, fma(y0, pw, s1));
  } else {
    const double y1 = fma(eps, coef * dy, -dy);
    const double y2 = fma(coef * dx, ieps, -dx);
    s0 = fma(y0, nw, fma(y1, pw, s0));
    s1 = fma(y2, nw, fma(y3, pw, s1));
  }
}

// Staged columns [j0, j1) against the wave's rows h .. h + FLY64_RB - 1, one
// column per lane (rows past the tile skipped wave-uniformly).  CHK: the
// columns may hold the row's own panel.  The own leaf's columns run as a
// range of their own so that no pass executes both forms under a split
// lane mask.
template <bool TR, bool CHK>
__device__ __forceinline__ void fly64Cols(int h, int j0, int j1, int lane, int w, int rows, const double4 *sP,
                                          const double4 *sN, const int *sY, const double4 (&X)[FLY64_RB],
                                          const double4 (&XN)[FLY64_RB], const int (&xi)[4], double kappa,
                                          double kap2, double eps, double ieps, const double *__restrict__ T,
                                          double (&a0)[4], double (&a1)[4]) {
  for (int j = j0 + lane; j < j1; j += 64) {
    const double4 P = sP[j], N = sN[j];
    const int y = CHK ? sY[j] : -1;
#pragma unroll
    for (int r = 0; r < FLY64_RB; r++)
      if (w + 4 * (h + r) < rows)
        fly64Pair<TR, CHK>(X[r], XN[r], P, N, CHK && y == xi[h + r], kappa, kap2, eps, ieps, T, a0[h + r],
                           a1[h + r]);
  }
}

template <bool TR>
__device__ __forceinline__ void nearFly64Tile(int t, const int2 *__restrict__ tiles,
                                              const int *__restrict__ leafStart, const int *__restrict__ leafCount,
                                              const int *__restrict__ srcOff, const int *__restrict__ ncolArr,
                                              const int *__restrict__ srcList, const double4 *__restrict__ PA,
                                              const double4 *__restrict__ PN, const double *__restrict__ x,
                                              const long long *__restrict__ specOff, const int *__restrict__ specCol,
                                              const double4 *__restrict__ specVal, double *__restrict__ out, int n,
                                              double kappa, double eps, double4 *sP, double4 *sN, int *sY,
                                              double *sE, const int *__restrict__ ownOff) {
  const int2 tl = tiles[t];
  const int D = tl.x, r0 = tl.y;
  const int oc = ownOff[D], nOwn = oc >= 0 ? leafCount[D] : 0;   // D's own columns: [oc, oc + nOwn)
  const int rows = min(NEAR_ROWS, leafCount[D] - r0);
  const int ncol = ncolArr[D];
  const int *src = srcList + srcOff[D];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const double ieps = 1.0 / eps, kap2 = kappa * kappa;
  // wave w owns rows w, w + 4, w + 8, w + 12 and takes each staged source once
  // for FLY64_RB of them at a time, held in registers (rows past the tile are
  // skipped wave-uniformly)
  int xi[4];
#pragma unroll
  for (int r = 0; r < 4; r++) xi[r] = leafStart[D] + r0 + min(w + 4 * r, rows - 1);
  double a0[4] = {0, 0, 0, 0}, a1[4] = {0, 0, 0, 0};
  for (int c0 = 0; c0 < ncol; c0 += FLY64_CHUNK) {
    const int cn = min(FLY64_CHUNK, ncol - c0);
    __syncthreads();
    for (int j = threadIdx.x; j < cn; j += 256) {
      const int y = src[c0 + j];
      const double4 p = PA[y], nn = PN[y];
      sP[j] = make_double4(p.x, p.y, p.z, p.w * x[y]);
      sN[j] = make_double4(nn.x, nn.y, nn.z, p.w * x[y + n]);
      sY[j] = y;
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 4; h += FLY64_RB) {
      if (w + 4 * h >= rows) break;
      double4 X[FLY64_RB], XN[FLY64_RB];
#pragma unroll
      for (int r = 0; r < FLY64_RB; r++) {
        X[r] = PA[xi[h + r]];
        XN[r] = PN[xi[h + r]];
      }
      if (FLY64_OWNSPLIT) {   // D's own columns [lo, hi) of this chunk apart, each range a loop of its own;
        // the checked range is extended to whole passes of 64 columns (the check
        // is exact for any column), so it leaves no partly idle pass behind
        const int lo = min(max(oc - c0, 0), cn);
        const int hi = min(lo + ((min(max(oc + nOwn - c0, 0), cn) - lo + 63) & ~63), cn);
        fly64Cols<TR, false>(h, 0, lo, lane, w, rows, sP, sN, sY, X, XN, xi, kappa, kap2, eps, ieps, sE, a0, a1);
        fly64Cols<TR, true>(h, lo, hi, lane, w, rows, sP, sN, sY, X, XN, xi, kappa, kap2, eps, ieps, sE, a0, a1);
        fly64Cols<TR, false>(h, hi, cn, lane, w, rows, sP, sN, sY, X, XN, xi, kappa, kap2, eps, ieps, sE, a0, a1);
      } else {
        fly64Cols<TR, true>(h, 0, cn, lane, w, rows, sP, sN, sY, X, XN, xi, kappa, kap2, eps, ieps, sE, a0, a1);
      }
    }
  }
  for (int r = 0; r < 4; r++) {
    const int row = w + 4 * r;
    const double A = PA[xi[r]].w * kFourPiI;
    double t0 = A * a0[r], t1 = A * a1[r];
    if (row < rows)
      for (long long e = specOff[xi[r]] + lane; e < specOff[xi[r] + 1]; e += 64) {
        const int y = specCol[e];
        const double4 d = specVal[e];
        t0 += d.x * x[y + n] + d.y * x[y];
        t1 += d.z * x[y + n] + d.w * x[y];
      }
    t0 = waveSumT(t0);
    t1 = waveSumT(t1);
    if (row < rows && lane == 0) {
      out[xi[r]] = t0;
      out[xi[r] + n] = t1;
    }
  }
}

// The mixed nearfield: tiles with a stored coefficient block (tileCoef >= 0)
// stream it (HBM-bound), the others compute their pairs on the fly
// (VALU-bound).  The host interleaves the two kinds in block order, so every
// CU holds both at once and the stream and the arithmetic overlap.
#define MIX64_LDS (2 * FLY64_CHUNK * 32 + FLY64_CHUNK * 4)
static_assert(MIX64_LDS >= 2 * NEAR_CHUNK * 8, "the stream tiles' LDS must fit the union");
template <bool TR = false>
__global__ __launch_bounds__(256) void k_near_mix64(const int2 *__restrict__ tiles, const int *__restrict__ leafStart,
                                                    const int *__restrict__ leafCount,
                                                    const int *__restrict__ srcOff, const int *__restrict__ ncolArr,
                                                    const long long *__restrict__ tileCoef,
                                                    const int *__restrict__ srcList,
                                                    const double4 *__restrict__ coef, const double4 *__restrict__ PA,
                                                    const double4 *__restrict__ PN, const double *__restrict__ x,
                                                    const long long *__restrict__ specOff,
                                                    const int *__restrict__ specCol,
                                                    const double4 *__restrict__ specVal, double *__restrict__ out,
                                                    int n, double kappa, double eps, int nTiles,
                                                    const int *__restrict__ order, const int *__restrict__ ownOff) {
  __shared__ __align__(16) unsigned char lds[MIX64_LDS];
  __shared__ double sE[256];   // 2^(j/256) for expNeg
  for (int j = threadIdx.x; j < 256; j += blockDim.x) sE[j] = gExp2Tab[j];
  // tiles in `order`, costliest first (the default grid has one block per tile;
  // a smaller grid strides over them)
  for (int k = blockIdx.x; k < nTiles; k += gridDim.x) {
    const int t = order[k];
    __syncthreads();
    if (tileCoef[t] >= 0) {
      double *sS = reinterpret_cast<double *>(lds);
      nearStreamTile<double>(t, tiles, leafStart, leafCount, srcOff, ncolArr, tileCoef, srcList, coef, x, out, n, sS,
                             sS + NEAR_CHUNK);
    } else {
      double4 *sP = reinterpret_cast<double4 *>(lds);
      nearFly64Tile<TR>(t, tiles, leafStart, leafCount, srcOff, ncolArr, srcList, PA, PN, x, specOff, specCol,
                        specVal, out, n, kappa, eps, sP, sP + FLY64_CHUNK, reinterpret_cast<int *>(sP + 2 * FLY64_CHUNK),
                        sE, ownOff);
    }
  }
}

__global__ void k_fly64_panels(const Pnl *__restrict__ pnl, int n, double t, double4 *__restrict__ PA,
                               double4 *__restrict__ PN) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Pnl &P = pnl[i];
  double q[3];
  for (int k = 0; k < 3; k++) q[k] = P.v[k] + t * (P.a[6 + k] + t * P.a[k]);
  PA[i] = make_double4(q[0], q[1], q[2], P.area);
  PN[i] = make_double4(P.nr[0], P.nr[1], P.nr[2], 0.0);
}

// ------------------------------------------- FP32 nearfield computed on the fly
// Precision form at q = 1: a disjoint pair's Galerkin coefficient is one
// kernel evaluation, KER = A_x A_y ker4(p_x - p_y) with p = v0 + a2/2 + a0/4
// (src/numQuad.c:586-618 at one Gauss point), cheaper to compute in FP32 than
// to stream.  Every source column of a destination's near list is evaluated
// that way except the panel itself; the pairs that share a vertex or an edge
// or are the panel itself get a stored correction KER_exact - KER_formula
// (FP64 at setup, applied in FP32).
// The quadrature point is kept as an FP32 pair hi + lo (hi = fl(p), lo =
// fl(p - hi)): r = (hi_x - hi_y) + (lo_x - lo_y) then carries FP32 relative
// accuracy in |r| itself, where one FP32 copy of p would carry an absolute
// error of ulp(|p|) (1e-5 of r for 0.4 A panels 40 A from the origin).
__global__ void k_fly_panels(const Pnl *__restrict__ pnl, int n, double t, float4 *__restrict__ PA,
                             float4 *__restrict__ PN, float4 *__restrict__ PL) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Pnl &P = pnl[i];
  double q[3];
  float h[3], l[3];
  for (int k = 0; k < 3; k++) {
    q[k] = P.v[k] + t * (P.a[6 + k] + t * P.a[k]);
    h[k] = (float)q[k];
    l[k] = (float)(q[k] - (double)h[k]);
  }
  PA[i] = make_float4(h[0], h[1], h[2], (float)P.area);
  PN[i] = make_float4((float)P.nr[0], (float)P.nr[1], (float)P.nr[2], 0.f);
  PL[i] = make_float4(l[0], l[1], l[2], 0.f);
}

// the corrected pairs of each destination panel: itself, and the near sources
// that share a vertex with it (commonVtx != 0)
__device__ inline bool flySpecial(const Pnl &X, const Pnl &Y, int x, int y) {
  int ix[3], iy[3];
  return x == y || commonVtx(X.v, Y.v, ix, iy) != 0;
}

__global__ __launch_bounds__(256) void k_spec_count(const Pnl *__restrict__ pnl, const int *__restrict__ leafOf,
                                                    const int *__restrict__ srcOff, const int *__restrict__ ncolArr,
                                                    const int *__restrict__ srcList, int *__restrict__ cnt) {
  __shared__ int part[256];
  const int x = blockIdx.x, D = leafOf[x], nc = ncolArr[D];
  const int *src = srcList + srcOff[D];
  const Pnl X = pnl[x];
  int k = 0;
  for (int j = threadIdx.x; j < nc; j += blockDim.x) k += flySpecial(X, pnl[src[j]], x, src[j]) ? 1 : 0;
  part[threadIdx.x] = k;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) cnt[x] = part[0];
}

// the column indices of each row's corrected pairs, in near-list order
__global__ __launch_bounds__(256) void k_spec_index(const Pnl *__restrict__ pnl, const int *__restrict__ leafOf,
                                                    const int *__restrict__ srcOff, const int *__restrict__ ncolArr,
                                                    const int *__restrict__ srcList,
                                                    const long long *__restrict__ specOff, int *__restrict__ specCol) {
  __shared__ int scan[256];
  const int x = blockIdx.x, D = leafOf[x], nc = ncolArr[D];
  const int *src = srcList + srcOff[D];
  const Pnl X = pnl[x];
  long long base = specOff[x];
  for (int j0 = 0; j0 < nc; j0 += 256) {   // ordered compaction, 256 columns at a time
    const int j = j0 + threadIdx.x;
    const int y = j < nc ? src[j] : 0;
    const bool sp = j < nc && flySpecial(X, pnl[y], x, y);
    scan[threadIdx.x] = sp ? 1 : 0;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {   // inclusive scan
      int v = threadIdx.x >= o ? scan[threadIdx.x - o] : 0;
      __syncthreads();
      scan[threadIdx.x] += v;
      __syncthreads();
    }
    if (sp) specCol[base + scan[threadIdx.x] - 1] = y;
    base += scan[255];
    __syncthreads();
  }
}

// correction of each listed pair: the Galerkin rule minus the one-point
// formula at q = 1 (the whole rule for the panel itself); DK: of d/dkappa.
// One block per row; the result goes to v32 (FP32) or v64.
// TR: of the transposed operator, (KER3, KER1, KER2, KER0) of the pair (y, x)
template <bool DK, bool TR = false>
__global__ __launch_bounds__(64) void k_spec_vals(const Pnl *__restrict__ pnl, const long long *__restrict__ specOff,
                                                  const int *__restrict__ specCol, float4 *__restrict__ v32,
                                                  double4 *__restrict__ v64, double kappa, double eps, QTab Q,
                                                  int q) {
  const int x = blockIdx.x;
  const Pnl X = pnl[x];
  const double t = Q.t[0];
  double px[3];
  for (int k = 0; k < 3; k++) px[k] = X.v[k] + t * (X.a[6 + k] + t * X.a[k]);
  for (long long e = specOff[x] + threadIdx.x; e < specOff[x + 1]; e += 64) {
    const int y = specCol[e];
    const Pnl Y = pnl[y];
    double k[4], f[4] = {0, 0, 0, 0};
    if (TR) panelIA0(Y, X, x == y, kappa, eps, Q, q, k, DK);
    else panelIA0(X, Y, x == y, kappa, eps, Q, q, k, DK);
    if (x != y) {
      double r[3];
      for (int d = 0; d < 3; d++) r[d] = px[d] - (Y.v[d] + t * (Y.a[6 + d] + t * Y.a[d]));
      if (TR)
        for (int d = 0; d < 3; d++) r[d] = -r[d];
      const double *na = TR ? Y.nr : X.nr, *nb = TR ? X.nr : Y.nr;
      if (DK) ker4dk(kappa, eps, r, na, nb, f);
      else ker4(kappa, eps, r, na, nb, f);
      for (int i = 0; i < 4; i++) f[i] *= X.area * Y.area;
    }
    double d[4];
    for (int i = 0; i < 4; i++) d[i] = k[i] - f[i];
    if (TR) {
      const double d0 = d[0];
      d[0] = d[3];
      d[3] = d0;
    }
    if (v32) v32[e] = make_float4((float)d[0], (float)d[1], (float)d[2], (float)d[3]);
    else v64[e] = make_double4(d[0], d[1], d[2], d[3]);
  }
}

#define FLY_CHUNK 512
#define FLY_RB 4   // rows per pass over the staged sources
// A quadrature point in the frame of a destination leaf: (hi - O) + lo with O
// the hi point of the leaf's first panel.  hi - O is exact (Sterbenz) for the
// points of a leaf and its neighbours away from the coordinate planes and at
// worst rounded at their small magnitude, so a difference of two local points
// carries an absolute error of a few 1e-7 of the neighbourhood's extent.
__device__ __forceinline__ float4 flyLocal(const float4 &hi, const float4 &lo, const float4 &O) {
  return make_float4((hi.x - O.x) + lo.x, (hi.y - O.y) + lo.y, (hi.z - O.z) + lo.z, 0.f);
}
// One FP32 pair (kernelKER4 at the one-point rule, src/kernel.c:109-147, with
// G0 = 1/(4 pi r) replaced by 1/r; the row's total is scaled by 1/(4 pi) once):
// r from the two points in their destination leaf's frame (flyLocal), 1/r from
// the hardware reciprocal square root (r2 is never denormal: the closest
// distinct quadrature points are 0.1 A apart, the panel itself has r2 set to
// 1), e^{-kappa r} from the hardware 2^x with kappa log2(e) folded into one
// constant nk = -kappa / ln 2.  1 - e and (1 + kappa r) e - 1 are formed
// directly: their absolute error (~1e-7) stays inside the FP32 operator's error
// budget, and the pairs closest together (a shared vertex or edge, or the
// panel itself) carry stored corrections taken against this same FP32 formula
// on the same local points.  42 VALU + 2 transcendental instructions.
// Returns the four kernels y0..y3 (times 4 pi, unweighted) of the pair.
__device__ __forceinline__ float4 fly32Ker(const float4 &XQ, const float4 &XN, const float4 &Q, const float4 &N,
                                           bool skip, float kappa, float nk, float kap2, float eps, float ieps) {
  const float rx = XQ.x - Q.x, ry = XQ.y - Q.y, rz = XQ.z - Q.z;
  const float r2 = skip ? 1.f : fmaf(rx, rx, fmaf(ry, ry, rz * rz));
  const float ri = __builtin_amdgcn_rsqf(r2), rr = r2 * ri;
  const float e = __builtin_amdgcn_exp2f(rr * nk), t = kappa * rr;
  const float ip0 = XN.x * N.x + XN.y * N.y + XN.z * N.z;
  const float ipX = XN.x * rx + XN.y * ry + XN.z * rz;
  const float ipY = N.x * rx + N.y * ry + N.z * rz;
  const float coef = fmaf(t, e, e), cm1 = coef - 1.f, ri2 = ri * ri, ri3 = ri * ri2;
  const float dy = ipY * ri3, dx = ipX * ri3, a = ipX * dy;
  const float ddG = fmaf(a * ri2, -3.f, ip0 * ri3);
  return make_float4(ri * (1.f - e), dy * fmaf(eps, coef, -1.f), dx * fmaf(coef, ieps, -1.f),
                     fmaf(cm1, ddG, -(kap2 * e) * a));
}

// Staged columns [j0, j1) against the wave's rows h .. h + FLY_RB - 1, one
// column per lane (FP32).  CHK: the columns may hold the row's own panel.
template <bool CHK>
__device__ __forceinline__ void fly32Cols(int h, int j0, int j1, int lane, int w, int rows, const float4 *sP,
                                          const float4 *sN, const int *sY, const float4 (&XQ)[4],
                                          const float4 (&XN)[4], const int (&xi)[4], float kappa, float nk,
                                          float kap2, float eps, float ieps, float (&a0)[4], float (&a1)[4]) {
  for (int j = j0 + lane; j < j1; j += 64) {
    const float4 P = sP[j], N = sN[j];
    const int y = CHK ? sY[j] : -1;
#pragma unroll
    for (int r = h; r < h + FLY_RB; r++) {
      if (w + 4 * r >= rows) continue;
      const bool skip = CHK && y == xi[r];   // the panel itself: weights zeroed, r2 replaced by 1
      const float4 k = fly32Ker(XQ[r], XN[r], P, N, skip, kappa, nk, kap2, eps, ieps);
      const float pw = skip ? 0.f : P.w, nw = skip ? 0.f : N.w;
      a0[r] = fmaf(k.x, nw, fmaf(k.y, pw, a0[r]));
      a1[r] = fmaf(k.z, nw, fmaf(k.w, pw, a1[r]));
    }
  }
}

// FP32 corrections of the listed pairs: the Galerkin rule (FP64, panelIA0)
// minus what an FP32 on-the-fly tile adds for the pair (fly32Ker on the same
// FP32 panel data, times the tile's own scale A_x / (4 pi) and A_y), so a
// corrected pair carries the rule's value up to the rounding of the stored
// difference.  One block per row.
__global__ __launch_bounds__(64) void k_spec_vals32(const Pnl *__restrict__ pnl, const float4 *__restrict__ PA,
                                                    const float4 *__restrict__ PN, const float4 *__restrict__ PL,
                                                    const int *__restrict__ leafOf,
                                                    const int *__restrict__ leafStart,
                                                    const long long *__restrict__ specOff,
                                                    const int *__restrict__ specCol, float4 *__restrict__ v32,
                                                    double kappa, double eps, QTab Q, int q) {
  const int x = blockIdx.x;
  const Pnl X = pnl[x];
  const float4 O = PA[leafStart[leafOf[x]]];   // the frame of x's leaf, as its tiles use it
  const float4 XP = PA[x], XN = PN[x], XQ = flyLocal(XP, PL[x], O);
  const float kf = (float)kappa, ef = (float)eps, nk = -kf * (float)M_LOG2E, kap2 = kf * kf, ieps = 1.f / ef;
  const double sx = (double)(XP.w * (float)kFourPiI);
  for (long long e = specOff[x] + threadIdx.x; e < specOff[x + 1]; e += 64) {
    const int y = specCol[e];
    double k[4], f[4] = {0, 0, 0, 0};
    panelIA0(X, pnl[y], x == y, kappa, eps, Q, q, k, false);
    if (x != y) {
      const float4 P = PA[y];
      const float4 kk = fly32Ker(XQ, XN, flyLocal(P, PL[y], O), PN[y], false, kf, nk, kap2, ef, ieps);
      const double s = sx * (double)P.w;
      f[0] = kk.x * s;
      f[1] = kk.y * s;
      f[2] = kk.z * s;
      f[3] = kk.w * s;
    }
    v32[e] = make_float4((float)(k[0] - f[0]), (float)(k[1] - f[1]), (float)(k[2] - f[2]), (float)(k[3] - f[3]));
  }
}

// LDS: (p in the destination leaf's frame, A sigma), (n, A sigma_n), source panel index; FLY_CHUNK each
__device__ __forceinline__ void nearFly32Tile(int t, const int2 *__restrict__ tiles,
                                              const int *__restrict__ leafStart, const int *__restrict__ leafCount,
                                              const int *__restrict__ srcOff, const int *__restrict__ ncolArr,
                                              const int *__restrict__ srcList, const float4 *__restrict__ PA,
                                              const float4 *__restrict__ PN, const float4 *__restrict__ PL,
                                              const double *__restrict__ x, const long long *__restrict__ specOff,
                                              const int *__restrict__ specCol, const float4 *__restrict__ specVal,
                                              float *__restrict__ out, int n, float kappa, float eps, float4 *sP,
                                              float4 *sN, int *sY, const int *__restrict__ ownOff) {
  const int2 tl = tiles[t];
  const int D = tl.x, r0 = tl.y;
  const int oc = ownOff[D], nOwn = oc >= 0 ? leafCount[D] : 0;   // D's own columns: [oc, oc + nOwn)
  const int rows = min(NEAR_ROWS, leafCount[D] - r0);
  const int ncol = ncolArr[D];
  const int *src = srcList + srcOff[D];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const float k4pi = (float)kFourPiI, kap2 = kappa * kappa, ieps = 1.f / eps, nk = -kappa * (float)M_LOG2E;
  const float4 O = PA[leafStart[D]];   // the leaf's frame
  float4 XQ[4], XN[4];
  float XA[4];
  int xi[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int row = w + 4 * r;
    xi[r] = leafStart[D] + r0 + min(row, rows - 1);
    const float4 p = PA[xi[r]];
    XA[r] = p.w;
    XQ[r] = flyLocal(p, PL[xi[r]], O);
    XN[r] = PN[xi[r]];
  }
  float a0[4] = {0, 0, 0, 0}, a1[4] = {0, 0, 0, 0};
  for (int c0 = 0; c0 < ncol; c0 += FLY_CHUNK) {
    const int cn = min(FLY_CHUNK, ncol - c0);
    __syncthreads();
    for (int j = threadIdx.x; j < cn; j += 256) {
      const int y = src[c0 + j];
      const float4 p = PA[y], nn = PN[y], q = flyLocal(p, PL[y], O);
      sP[j] = make_float4(q.x, q.y, q.z, p.w * (float)x[y]);
      sN[j] = make_float4(nn.x, nn.y, nn.z, p.w * (float)x[y + n]);
      sY[j] = y;
    }
    __syncthreads();
    // each staged source is read once for FLY_RB rows at a time
#pragma unroll
    for (int h = 0; h < 4; h += FLY_RB) {
      if (w + 4 * h >= rows) break;
      // D's own columns, padded to whole passes of 64, are checked for the panel itself
      const int lo = min(max(oc - c0, 0), cn);
      const int hi = min(lo + ((min(max(oc + nOwn - c0, 0), cn) - lo + 63) & ~63), cn);
      fly32Cols<false>(h, 0, lo, lane, w, rows, sP, sN, sY, XQ, XN, xi, kappa, nk, kap2, eps, ieps, a0, a1);
      fly32Cols<true>(h, lo, hi, lane, w, rows, sP, sN, sY, XQ, XN, xi, kappa, nk, kap2, eps, ieps, a0, a1);
      fly32Cols<false>(h, hi, cn, lane, w, rows, sP, sN, sY, XQ, XN, xi, kappa, nk, kap2, eps, ieps, a0, a1);
    }
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int row = w + 4 * r;
    const float A = XA[r] * k4pi;
    float t0 = A * a0[r], t1 = A * a1[r];
    if (row < rows) {   // the stored corrections of this row
      for (long long e = specOff[xi[r]] + lane; e < specOff[xi[r] + 1]; e += 64) {
        const int y = specCol[e];
        const float4 d = specVal[e];
        const float sg = (float)x[y], sn = (float)x[y + n];
        t0 += d.x * sn + d.y * sg;
        t1 += d.z * sn + d.w * sg;
      }
    }
    t0 = waveSumT(t0);
    t1 = waveSumT(t1);
    if (row < rows && lane == 0) {
      out[xi[r]] = t0;
      out[xi[r] + n] = t1;
    }
  }
}

#define MIX32_LDS (2 * FLY_CHUNK * 16 + FLY_CHUNK * 4)
static_assert(MIX32_LDS >= 2 * NEAR_CHUNK * 4, "the stream tiles' LDS must fit the union");
__global__ __launch_bounds__(256) void k_near_mix32(const int2 *__restrict__ tiles, const int *__restrict__ leafStart,
                                                    const int *__restrict__ leafCount,
                                                    const int *__restrict__ srcOff, const int *__restrict__ ncolArr,
                                                    const long long *__restrict__ tileCoef,
                                                    const int *__restrict__ srcList,
                                                    const float4 *__restrict__ coef, const float4 *__restrict__ PA,
                                                    const float4 *__restrict__ PN, const float4 *__restrict__ PL,
                                                    const double *__restrict__ x,
                                                    const long long *__restrict__ specOff,
                                                    const int *__restrict__ specCol,
                                                    const float4 *__restrict__ specVal, float *__restrict__ out, int n,
                                                    float kappa, float eps, const int *__restrict__ order,
                                                    const int *__restrict__ ownOff) {
  __shared__ __align__(16) unsigned char lds[MIX32_LDS];
  const int t = order[blockIdx.x];   // costliest tiles first
  if (tileCoef[t] >= 0) {
    float *sS = reinterpret_cast<float *>(lds);
    nearStreamTile<float>(t, tiles, leafStart, leafCount, srcOff, ncolArr, tileCoef, srcList, coef, x, out, n, sS,
                          sS + NEAR_CHUNK);
  } else {
    float4 *sP = reinterpret_cast<float4 *>(lds);
    nearFly32Tile(t, tiles, leafStart, leafCount, srcOff, ncolArr, srcList, PA, PN, PL, x, specOff, specCol, specVal,
                  out, n, kappa, eps, sP, sP + FLY_CHUNK, reinterpret_cast<int *>(sP + 2 * FLY_CHUNK), ownOff);
  }
}

// FP32 copies of FP64 data (the same layout, element by element)
__global__ void k_to_f32(const double *__restrict__ in, float *__restrict__ out, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] = (float)in[i];
}

// -------------------------------------------------- multi-index helpers
#pragma clang fp contract(off)

// mi: .x = i1 | i2<<8 | i3<<16 | deg<<24, .y = axis | fac<<8, .z = j1, .w = j2 (-1: none);
// j1/j2 are local indices of alpha-e_axis / alpha-2e_axis inside their degree
struct DevTabs {
  const int4 *mi;      // nMom(P)
  const int4 *ep;      // nMom(maxOrder): idx(a+e1), idx(a+e2), idx(a+e3), sgn
  const double *ifact3;
  const int *convStart, *conv;   // (beta | (alpha+beta) << 16) per alpha
  const int *m2mStart, *m2m;     // (alpha-k | k << 16) per alpha
};

__device__ __forceinline__ int deg_of(int4 m) { return (m.x >> 24) & 0xff; }

// src/kernel.c:84-104 (kernelDG0DGk)
__device__ void radial(double kappa, double r, int p, double *G0, int s0, double *Gk, int sk) {
  double kappa2 = kappa * kappa, rexp = exp(-kappa * r), r1 = 1.0 / r, r2 = r1 * r1, nr2 = -r2;
  double g0m, g0 = r1 * kFourPiI, gkm, gk = rexp * g0;
  G0[0] = g0; Gk[0] = gk;
  if (p < 1) return;
  g0m = g0; gkm = gk;
  g0 = nr2 * g0m;
  gk = -gkm * (r2 + kappa * r1);
  G0[s0] = g0; Gk[sk] = gk;
  for (int k = 1; k < p; k++) {
    double g0n = (2 * k + 1) * nr2 * g0;
    double gkn = r2 * (kappa2 * gkm - (2 * k + 1) * gk);
    gkm = gk; g0 = g0n; gk = gkn;
    G0[(k + 1) * s0] = g0; Gk[(k + 1) * sk] = gk;
  }
}

// Derivative tensors of G0 and Gk at r up to total degree p (src/expan.c:138-209),
// computed by the whole block degree by degree: block d holds entries
// (q, j), q = 0..p-d, j < (d+1)(d+2)/2, at offset off(d) + q*len(d) + j.
// Leaves dG[0][alpha] in o0/ok.  Ends with a barrier.
// dk: o0 receives d/dkappa of the Gk tensors instead of the G0 tensors.  The
// radial seeds f_q = ((1/r) d/dr)^q Gk satisfy d f_q / d kappa = kappa f_(q-1)
// for q >= 1 (d/dkappa commutes with (1/r) d/dr and (1/r) d/dr e^{-kappa r}
// = -kappa e^{-kappa r}/r), and d f_0 / d kappa = -e^{-kappa r}/(4 pi).
__device__ void derivs_block(double kappa, int p, const double *r, double *s0, double *sk, double *o0, double *ok,
                             const int4 *mi, int tid, int nt, bool dk = false) {
  if (tid == 0) {
    double rn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    radial(kappa, rn, p, s0, 1, sk, 1);
    if (dk) {
      for (int q = p; q >= 1; q--) s0[q] = kappa * sk[q - 1];
      s0[0] = -exp(-kappa * rn) * kFourPiI;
    }
    o0[0] = s0[0];
    ok[0] = sk[0];
  }
  __syncthreads();
  int offPrev2 = 0, offPrev = 0, off = p + 1;   // block offsets of d-2, d-1, d
  int lenPrev2 = 0, lenPrev = 1;
  for (int d = 1; d <= p; d++) {
    int len = (d + 1) * (d + 2) / 2, cnt = (p - d + 1) * len, base = NMOM(d - 1);
    for (int e = tid; e < cnt; e += nt) {
      int q = e / len, j = e - q * len;
      int4 m = mi[base + j];
      int axis = m.y & 0xff, fac = m.y >> 8;
      double v0 = s0[offPrev + (q + 1) * lenPrev + m.z] * r[axis];
      double vk = sk[offPrev + (q + 1) * lenPrev + m.z] * r[axis];
      if (m.w >= 0) {
        v0 = v0 + s0[offPrev2 + (q + 1) * lenPrev2 + m.w] * fac;
        vk = vk + sk[offPrev2 + (q + 1) * lenPrev2 + m.w] * fac;
      }
      s0[off + q * len + j] = v0;
      sk[off + q * len + j] = vk;
      if (q == 0) { o0[base + j] = v0; ok[base + j] = vk; }
    }
    __syncthreads();
    offPrev2 = offPrev; lenPrev2 = lenPrev;
    offPrev = off; lenPrev = len;
    off += cnt;
  }
}
__host__ __device__ constexpr int derivScratch(int p) {   // sum_d (p-d+1)(d+1)(d+2)/2
  return (p + 1) * (p + 2) * (p + 3) * (p + 4) / 24;
}

// src/moments.c:83-103 (setupConVect) for one multi-index: t^a / a!
__device__ __forceinline__ double conVect1(const double *tr, int4 m, double ifact3) {
  int i1 = m.x & 0xff, i2 = (m.x >> 8) & 0xff, i3 = (m.x >> 16) & 0xff;
  double f1 = 1.0, f2 = 1.0, f3 = 1.0;
  for (int i = 0; i < i1; i++) f1 = f1 * tr[0];
  for (int i = 0; i < i2; i++) f2 = f2 * tr[1];
  for (int i = 0; i < i3; i++) f3 = f3 * tr[2];
  double v = f1 * f2 * f3;
  return v * ifact3;
}

// ------------------------------------------------------------- Q2M matrices

// src/moments.c:148-236 (calcOneMoment0, calcOneNoment, calcMoments0): one
// wave per panel, lanes over multi-indices; columns translated to the leaf centre
__global__ __launch_bounds__(64) void k_q2m_mats(const Pnl *__restrict__ pnl, const int *__restrict__ leafOf,
                                                 const double *__restrict__ cubeX, DevTabs T, int order,
                                                 const double *__restrict__ fact, double *__restrict__ Q0,
                                                 double *__restrict__ Q1) {
  extern __shared__ double sm[];
  const int nm = NMOM(order);
  double *c1 = sm, *c2 = sm + nm, *M = sm + 2 * nm, *cv = sm + 3 * nm, *loc = sm + 4 * nm;
  const int p = blockIdx.x, lane = threadIdx.x;
  const Pnl P = pnl[p];
  const double *ctr = &cubeX[3 * leafOf[p]];
  double a1[3], a2[3], tr[3];
  for (int i = 0; i < 3; i++) { a1[i] = P.v[3 + i] - P.v[i]; a2[i] = P.v[6 + i] - P.v[i]; }
  for (int i = 0; i < 3; i++) tr[i] = P.v[i] - ctr[i];
  const double area2 = 2.0 * P.area;
  for (int job = 0; job < 2; job++) {
    const int order1 = job == 0 ? order : order - 1, nm1 = NMOM(order1);
    for (int a = lane; a < nm1; a += 64) {
      int4 m = T.mi[a];
      int d = deg_of(m);
      c1[a] = conVect1(a1, m, T.ifact3[a]) * fact[d];
      c2[a] = conVect1(a2, m, T.ifact3[a]) * fact[d];
      cv[a] = conVect1(tr, m, T.ifact3[a]);
    }
    for (int a = nm1 + lane; a < nm; a += 64) cv[a] = conVect1(tr, T.mi[a], T.ifact3[a]);
    __syncthreads();
    // convolution (moments.c:56-78) of c1, c2 up to order1, times 2A/(n+2)!
    for (int a = lane; a < nm1; a += 64) {
      double tmp = 0.0;
      for (int e = T.m2mStart[a]; e < T.m2mStart[a + 1]; e++) {
        int pr = T.m2m[e];
        tmp += c1[pr & 0xffff] * c2[pr >> 16];
      }
      M[a] = tmp * (area2 * fact[MAXP + 3 + deg_of(T.mi[a]) + 2]);   // fact[MAXP+3+k] holds 1/k!
    }
    __syncthreads();
    if (job == 0) {
      for (int a = lane; a < nm; a += 64) loc[a] = M[a];
    } else {
      // noments (moments.c:115-133) as a gather: N[b] = n0 M[b-e1] + n1 M[b-e2] + n2 M[b-e3]
      for (int b = lane; b < nm; b += 64) {
        int4 m = T.mi[b];
        int i1 = m.x & 0xff, i2 = (m.x >> 8) & 0xff, i3 = (m.x >> 16) & 0xff;
        double v = 0.0;
        bool any = false;
        if (i1 > 0) { v = P.nr[0] * M[T.ep[b].x]; any = true; }   // ep[b].x holds idx(b - e1) here
        if (i2 > 0) { double t = P.nr[1] * M[T.ep[b].y]; v = any ? v + t : t; any = true; }
        if (i3 > 0) { double t = P.nr[2] * M[T.ep[b].z]; v = any ? v + t : t; any = true; }
        loc[b] = v;
      }
    }
    __syncthreads();
    // translation to the leaf centre: column = conv(cv, loc) (moments.c:228-233)
    double *Qo = job == 0 ? Q0 : Q1;
    for (int a = lane; a < nm; a += 64) {
      double tmp = 0.0;
      for (int e = T.m2mStart[a]; e < T.m2mStart[a + 1]; e++) {
        int pr = T.m2m[e];
        tmp += cv[pr & 0xffff] * loc[pr >> 16];
      }
      Qo[(long long)p * nm + a] = tmp;
    }
    __syncthreads();
  }
}

#pragma clang fp contract(fast)
// ------------------------------------------------------------- FMM passes

// mom_pot = Q2M1 sgm, mom_dpdn = Q2M0 sgm_n (src/fmm.c:745-753): one block per
// leaf, four waves split the leaf's panels (a contiguous column each), lanes own
// multi-indices; the wave partials are added in wave order
// Q2M (src/fmm.c:745-753, moments.c:216-236): leaf moments
// mom_pot = Q2M1 sgm, mom_dpdn = Q2M0 sgm_n.  One block per leaf; the block
// reads the leaf's matrices (panel-major columns of nm, contiguous over the
// leaf's panels) as rows of nm / V threads, each thread loading V consecutive
// slots at once (16 bytes where nm allows), 256 / (nm / V) panels per pass, so
// every load is one coalesced run over several columns; the passes are
// unrolled so that several loads per thread are in flight.  The rows' partial
// sums are added in row order at the end.
// swap: the transposed L2P of the adjoint operator (Q1 x_n, Q0 x).
template <typename R, int V>
__global__ __launch_bounds__(256) void k_q2m(const R *__restrict__ Q0, const R *__restrict__ Q1,
                                             const int *__restrict__ leafStart, const int *__restrict__ leafCount,
                                             const double *__restrict__ x, R *__restrict__ mom, int n, int nm,
                                             int nmMax, int swap) {
  typedef R VT __attribute__((ext_vector_type(V)));
  R *sm = reinterpret_cast<R *>(dynLds);   // [2][256 V] partial sums: (row, slot)
  const int L = blockIdx.x, t = threadIdx.x, G = nm / V;
  const int rows = max(1, 256 / G), g = t % G, r = t / G;
  const int s = leafStart[L], c = leafCount[L];
  const int oP = swap ? n : 0, oN = swap ? 0 : n;
  VT m0 = (R)0, m1 = (R)0;
  if (r < rows) {
#pragma unroll 4
    for (int j = r; j < c; j += rows) {
      const long long col = (long long)(s + j) * nm + g * V;
      m0 += *reinterpret_cast<const VT *>(Q1 + col) * (R)x[s + j + oP];
      m1 += *reinterpret_cast<const VT *>(Q0 + col) * (R)x[s + j + oN];
    }
    for (int v = 0; v < V; v++) {
      sm[r * nm + g * V + v] = m0[v];
      sm[256 * V + r * nm + g * V + v] = m1[v];
    }
  }
  __syncthreads();
  for (int k = t; k < nm; k += 256) {
    R v0 = 0, v1 = 0;
    for (int q = 0; q < rows; q++) {
      v0 += sm[q * nm + k];
      v1 += sm[256 * V + q * nm + k];
    }
    mom[(long long)L * 2 * nmMax + k] = v0;
    mom[(long long)L * 2 * nmMax + nmMax + k] = v1;
  }
}

// LDS writes of a wave become visible to all its lanes (no block barrier)
__device__ __forceinline__ void waveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// src/expan.c:342-382 (transM2MWs) + convM2M (:223-247): one block per parent,
// one wave per child (at most eight), lanes over multi-indices.  Each wave
// forms its child's translation monomials and moments in LDS and convolves
// them; the children's contributions are added in child order at the end.
// The (alpha-k, k) table sits in LDS.
template <typename R>
__global__ __launch_bounds__(512) void k_m2m(int levStart, const int *__restrict__ kidStart,
                                             const int *__restrict__ nKids, const int *__restrict__ cubeLevel,
                                             const int *__restrict__ ordMom, const double *__restrict__ cubeX,
                                             R *__restrict__ mom, int nmMax, DevTabs T) {
  R *sm = reinterpret_cast<R *>(dynLds);
  const int par = levStart + blockIdx.x, tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int nk = nKids[par], ordOut = ordMom[cubeLevel[par]], nOut = NMOM(ordOut);
  R *Rt = sm + 3 * w * nOut, *b1 = Rt + nOut, *b2 = Rt + 2 * nOut;   // this wave's child
  int *ts = (int *)(sm + 24 * nOut), *tab = ts + nOut + 1;
  const int nT = T.m2mStart[nOut];
  for (int i = tid; i <= nOut; i += blockDim.x) ts[i] = T.m2mStart[i];
  for (int i = tid; i < nT; i += blockDim.x) tab[i] = T.m2m[i];
  if (w < nk) {
    const int kid = kidStart[par] + w;
    const int nIn = NMOM(ordMom[cubeLevel[kid]]);
    double tr[3];
    for (int d = 0; d < 3; d++) tr[d] = cubeX[3 * kid + d] - cubeX[3 * par + d];
    for (int a = lane; a < nOut; a += 64) {
      Rt[a] = (R)conVect1(tr, T.mi[a], T.ifact3[a]);
      b1[a] = a < nIn ? mom[(long long)kid * 2 * nmMax + a] : (R)0;
      b2[a] = a < nIn ? mom[(long long)kid * 2 * nmMax + nmMax + a] : (R)0;
    }
  }
  __syncthreads();
  R t1[4] = {0, 0, 0, 0}, t2[4] = {0, 0, 0, 0};
  if (w < nk) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int a = lane + 64 * q;
      if (a < nOut)
        for (int e = ts[a]; e < ts[a + 1]; e++) {
          const int pr = tab[e];
          const R rv = Rt[pr & 0xffff];
          t1[q] += rv * b1[pr >> 16];
          t2[q] += rv * b2[pr >> 16];
        }
    }
  }
  __syncthreads();
  if (w < nk) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int a = lane + 64 * q;
      if (a < nOut) {
        b1[a] = t1[q];
        b2[a] = t2[q];
      }
    }
  }
  __syncthreads();
  for (int a = tid; a < nOut; a += blockDim.x) {
    R s1 = 0, s2 = 0;
    for (int k = 0; k < nk; k++) {
      s1 += sm[3 * k * nOut + nOut + a];
      s2 += sm[3 * k * nOut + 2 * nOut + a];
    }
    mom[(long long)par * 2 * nmMax + a] = s1;
    mom[(long long)par * 2 * nmMax + nmMax + a] = s2;
  }
}

// src/fmm.c:774-787, src/expan.c:471-494 (transM2L) + convM2L (:258-290):
// one wave per destination cube, sources of its interaction list in order
__global__ __launch_bounds__(64) void k_m2l(const int *__restrict__ m2lOff, const int *__restrict__ m2lSrc,
                                            const int *__restrict__ cubeLevel, const int *__restrict__ ordM2L,
                                            const int *__restrict__ ordMom, const double *__restrict__ cubeX,
                                            const double *__restrict__ mom, double *__restrict__ lec, int nmMax,
                                            double kappa, double eps, DevTabs T) {
  extern __shared__ double sm[];
  const int c = blockIdx.x, tid = threadIdx.x;
  const int p = ordM2L[cubeLevel[c]], nm = NMOM(p), dsz = derivScratch(p);
  double *s0 = sm, *sk = s0 + dsz, *G0 = sk + dsz, *Gk = G0 + nm, *c1 = Gk + nm, *c2 = c1 + nm;
  double *v1 = c2 + nm, *v2 = v1 + nm, *v3 = v2 + nm;
  double acc[4][4] = {};   // up to 4 multi-indices per lane (nm <= 256)
  for (int pr = m2lOff[c]; pr < m2lOff[c + 1]; pr++) {
    const int src = m2lSrc[pr];
    double r[3];
    for (int d = 0; d < 3; d++) r[d] = cubeX[3 * c + d] - cubeX[3 * src + d];
    derivs_block(kappa, p, r, s0, sk, G0, Gk, T.mi, tid, 64);
    const double *mp = mom + (long long)src * 2 * nmMax, *md = mp + nmMax;
    for (int i = tid; i < nm; i += 64) {
      int sg = T.ep[i].w;
      c1[i] = sg * mp[i];
      c2[i] = sg * md[i];
      v1[i] = G0[i] - Gk[i];
      v2[i] = eps * Gk[i] - G0[i];
      v3[i] = G0[i] - Gk[i] / eps;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const int a = tid + 64 * s;
      if (a < nm) {
        const int cnt = NMOM(p - deg_of(T.mi[a]));
        const int *cv = T.conv + T.convStart[a];
        double t1 = 0.0, t2 = 0.0, t3 = 0.0, t4 = 0.0;
        for (int e = 0; e < cnt; e++) {
          int pe = cv[e], j = pe & 0xffff, kk = pe >> 16;
          t1 += c2[j] * v1[kk];
          t2 += c1[j] * v2[kk];
          t3 += c2[j] * v3[kk];
          t4 += c1[j] * (-v1[kk]);
        }
        acc[s][0] += t1; acc[s][1] += t2; acc[s][2] += t3; acc[s][3] += t4;
      }
    }
    __syncthreads();
  }
  const int nmc = NMOM(ordMom[cubeLevel[c]]);
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const int a = tid + 64 * s;
    if (a < nmMax)
      for (int k = 0; k < 4; k++) lec[((long long)c * 4 + k) * nmMax + a] = (a < nm && a < nmc) ? acc[s][k] : 0.0;
  }
}

// Derivative tensors D^a G0, D^a Gk at r = x_dst - x_src for every M2L pair,
// once per operator (src/expan.c:138-209): one wave per pair, stride 2*nm.
__global__ __launch_bounds__(64) void k_m2l_derivs(const int *__restrict__ pairDst, const int *__restrict__ m2lSrc,
                                                   const int *__restrict__ cubeLevel, const int *__restrict__ ordM2L,
                                                   const double *__restrict__ cubeX, double *__restrict__ G,
                                                   int nmMax, double kappa, DevTabs T, bool dk = false) {
  extern __shared__ double sm[];
  const long long pr = blockIdx.x;
  const int c = pairDst[pr], src = m2lSrc[pr];
  const int p = ordM2L[cubeLevel[c]], dsz = derivScratch(p);
  double r[3];
  for (int d = 0; d < 3; d++) r[d] = cubeX[3 * c + d] - cubeX[3 * src + d];
  double *g0 = G + pr * 2 * nmMax, *gk = g0 + nmMax;
  derivs_block(kappa, p, r, sm, sm + dsz, g0, gk, T.mi, threadIdx.x, 64, dk);   // dk: g0 = dGk/dkappa
}

// M2L with cached derivative tensors (src/fmm.c:774-787, src/expan.c:258-290,
// 471-494): one block of nw waves (four unless LDS is short) per destination
// cube; wave w takes pairs w, w + nw, ... of the cube's interaction list, on
// its own (wave barriers only, next pair prefetched into registers).  The
// (alpha, beta) terms of the convolution are cut into items (short runs of one alpha)
// that the host spreads over the 64 lanes by term count, so every lane walks a
// list of about 1/64 of the terms instead of one lane carrying alpha = 0's long
// beta list.  The last term of an item carries the item number: there the lane
// adds its running sums into the item's slot in LDS (its own, per wave).  The
// slots are summed per alpha at the end, in item then wave order.
#define M2L_FLUSH 0x80000000u
template <typename R> struct Vec2;
template <> struct Vec2<double> { typedef double2 type; };
template <> struct Vec2<float> { typedef float2 type; };

// MAXT > 0: every lane's term list (at most MAXT entries, the same for every
// pair) is held in registers and the contraction is unrolled, so the LDS reads
// of different terms are in flight together; MAXT = 0 reads the list from LDS
// TR: the transposed pass of the adjoint operator.  mom then holds the
// adjoint locals (b = Q1 y_n | a = Q0 y, carried up the tree) and lec receives
// the adjoint moments (M_pot | M_dpdn), two arrays per cube:
//   M_pot += a*(eps Gk - G0) + b*(Gk - G0),  M_dpdn += a*(G0 - Gk) + b*(G0 - Gk/eps)
// with the convolutions taken over the cube's own list (the lists are
// symmetric and D^a G(x_S - x_D) = sgn_a D^a G(x_D - x_S)).
template <typename R, int MAXT, bool TR = false>
__global__ __launch_bounds__(256) void k_m2l_cached(const int *__restrict__ m2lOff, const int *__restrict__ m2lSrc,
                                                    const int *__restrict__ cubeLevel,
                                                    const int *__restrict__ ordM2L, const int *__restrict__ ordMom,
                                                    const R *__restrict__ G, const R *__restrict__ mom,
                                                    R *__restrict__ lec, int nmMax, R eps, DevTabs T,
                                                    const unsigned *__restrict__ planTerms,
                                                    const int *__restrict__ planIdx, const int4 *__restrict__ planOff) {
  typedef typename Vec2<R>::type R2;
  R *sm = reinterpret_cast<R *>(dynLds);
  const int c = blockIdx.x, tid = threadIdx.x, w = tid >> 6, lane = tid & 63, nw = blockDim.x >> 6;
  const int p = ordM2L[cubeLevel[c]], nm = NMOM(p);
  // plan of order p: x = first term, y = term count, z = first index entry, w = item count;
  // index entries: 65 lane starts into the terms, then nm + 1 item starts per alpha
  const int4 po = planOff[p];
  const int nTerms = po.y, nItems = po.w;
  const int nmP = (nm + 1) & ~1;                       // even, so every slot starts 16-byte aligned
  R *slots = sm;                                       // per wave: nmP x (c1 c2), nmP x (a0 ak)
  R *red = sm + nw * 4 * nmP;                          // [wave][item][4]
  unsigned *tab = (unsigned *)(red + nw * 4 * nItems);
  int *idx = (int *)(tab + nTerms);
  for (int i = tid; i < nTerms; i += blockDim.x) tab[i] = planTerms[po.x + i];
  for (int i = tid; i < 65 + nm + 1; i += blockDim.x) idx[i] = planIdx[po.z + i];
  for (int i = tid; i < nw * 4 * nItems; i += blockDim.x) red[i] = 0;
  __syncthreads();
  // each wave walks its own pairs without block barriers: the next pair's
  // tensors and moments are loaded into registers while the current one is
  // contracted out of this wave's LDS slot
  R2 *__restrict__ cc = reinterpret_cast<R2 *>(slots + w * 4 * nmP);          // c1 = sgn M_pot, c2 = sgn M_dpdn
  R2 *__restrict__ gg = reinterpret_cast<R2 *>(slots + w * 4 * nmP + 2 * nmP);  // a0 = D G0, ak = D Gk
  R *__restrict__ myRed = red + w * 4 * nItems;
  const int e0 = idx[lane], e1 = idx[lane + 1];
  unsigned pe[MAXT > 0 ? MAXT : 1];
  if constexpr (MAXT > 0) {
#pragma unroll
    for (int k = 0; k < MAXT; k++) pe[k] = e0 + k < e1 ? tab[e0 + k] : 0u;
  }
  R sg[4];
#pragma unroll
  for (int q = 0; q < 4; q++) sg[q] = lane + 64 * q < nm ? (R)T.ep[lane + 64 * q].w : (R)0;
  R ga[4], gb[4], ma[4], mb[4];
  auto fetch = [&](int pr) {
    const int src = m2lSrc[pr];
    const R *g0 = G + (long long)pr * 2 * nmMax, *gk = g0 + nmMax;
    const R *mp = mom + (long long)src * 2 * nmMax, *md = mp + nmMax;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int i = lane + 64 * q;
      if (i < nm) {
        ga[q] = g0[i];
        gb[q] = gk[i];
        ma[q] = mp[i];
        mb[q] = md[i];
      }
    }
  };
  // four products per (alpha, beta): c1 a0, c1 ak, c2 a0, c2 ak; an item's
  // sums combine into the four locals when it ends:
  //   l1 += c2*(a0 - ak), l2 += c1*(eps ak - a0), l3 += c2*(a0 - ak/eps), l4 += c1*(ak - a0)
  auto term = [&](unsigned e, R &t1, R &t2, R &t3, R &t4) {
    const R2 x = cc[e & 0xff];
    const R2 g = gg[(e >> 8) & 0xff];
    t1 += x.x * g.x;
    t2 += x.x * g.y;
    t3 += x.y * g.x;
    t4 += x.y * g.y;
    if (e & M2L_FLUSH) {
      R *r = myRed + 4 * ((e >> 16) & 0x7fff);
      if constexpr (TR) {   // x = (sgn b, sgn a): t1 = b*G0, t2 = b*Gk, t3 = a*G0, t4 = a*Gk
        r[0] += eps * t4 - t3 + t2 - t1;
        r[1] += t3 - t4 + t1 - t2 / eps;
      } else {
        r[0] += t3 - t4;
        r[1] += eps * t2 - t1;
        r[2] += t3 - t4 / eps;
        r[3] += t2 - t1;
      }
      t1 = t2 = t3 = t4 = 0;
    }
  };
  const int b1 = m2lOff[c + 1];
  int pr = m2lOff[c] + w;
  if (pr < b1) fetch(pr);
  for (; pr < b1; pr += nw) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int i = lane + 64 * q;
      if (i < nm) {
        R2 x, g;
        x.x = sg[q] * ma[q];
        x.y = sg[q] * mb[q];
        cc[i] = x;
        g.x = ga[q];
        g.y = gb[q];
        gg[i] = g;
      }
    }
    waveSync();
    if (pr + nw < b1) fetch(pr + nw);
    R t1 = 0, t2 = 0, t3 = 0, t4 = 0;
    if constexpr (MAXT > 0) {
#pragma unroll
      for (int k = 0; k < MAXT; k++)
        if (e0 + k < e1) term(pe[k], t1, t2, t3, t4);
    } else {
      for (int e = e0; e < e1; e++) term(tab[e], t1, t2, t3, t4);
    }
    waveSync();
  }
  __syncthreads();
  const int nmc = NMOM(ordMom[cubeLevel[c]]);
  const int *itemStart = idx + 65;
  constexpr int NO = TR ? 2 : 4;
  for (int i = tid; i < NO * nmMax; i += blockDim.x) {
    const int k = i / nmMax, a = i - k * nmMax;
    R v = 0;
    if (a < nm && a < nmc)
      for (int it = itemStart[a]; it < itemStart[a + 1]; it++)
        for (int ww = 0; ww < nw; ww++) v += red[(ww * nItems + it) * 4 + k];
    lec[((long long)c * NO + k) * nmMax + a] = v;
  }
}

// M2L on a fixed term grid (src/fmm.c:774-787, src/expan.c:258-290, 471-494):
// the host lays the (alpha, beta) terms of one pair on K steps x 64 lanes; a
// lane keeps one alpha for each run of L consecutive steps, so run r's four
// running sums live in registers (acc[r]) for the whole interaction list of the
// cube, and no lane stores anything per pair.  At each step the lanes' beta and
// alpha+beta slots are chosen on the host to spread over the LDS banks; an idle
// lane reads the zero slot of the moments.  After the list, the waves add their
// registers into one [run][lane] table in wave order and one thread per alpha
// sums the cells that hold it (a fixed list), then combines the four sums
//   t1 = sum c1 a0, t2 = sum c1 ak, t3 = sum c2 a0, t4 = sum c2 ak
// into the locals (TR: into the adjoint moments), as k_m2l_cached does per item.
// DK: G holds (dGk/dkappa, Gk) and the output is the raw pair (t1, t3), the
// moments against dGk/dkappa (the kappa derivative of the far operator).
template <typename R, int K, int L, bool TR = false, bool DK = false>
__global__ __launch_bounds__(256) void k_m2l_grid(const int *__restrict__ m2lOff, const int *__restrict__ m2lSrc,
                                                  const int *__restrict__ cubeLevel, const int *__restrict__ ordM2L,
                                                  const int *__restrict__ ordMom, const R *__restrict__ G,
                                                  const R *__restrict__ mom, R *__restrict__ lec, int nmMax, R eps,
                                                  DevTabs T, const unsigned *__restrict__ gridTerms,
                                                  const int *__restrict__ gridOff, const int *__restrict__ redStart,
                                                  const unsigned short *__restrict__ redCell, int slotZ, int nccP,
                                                  int nggP, const int *__restrict__ cubeOrder) {
  typedef typename Vec2<R>::type R2;
  constexpr int NR = K / L;
  constexpr int NQ = K <= 8 ? 1 : K <= 32 ? 2 : 3;   // moments per lane: nMom(5) = 56, nMom(7) = 120, nMom(8) = 165
  // cubes with the longest interaction lists first, so that none starts last
  const int c = cubeOrder[blockIdx.x], tid = threadIdx.x, w = tid >> 6, lane = tid & 63, nw = blockDim.x >> 6;
  const int p = ordM2L[cubeLevel[c]], nm = NMOM(p);
  R2 *__restrict__ cc = reinterpret_cast<R2 *>(dynLds) + w * (nccP + nggP);   // c1 = sgn M_pot, c2 = sgn M_dpdn
  R2 *__restrict__ gg = cc + nccP;                                            // a0 = D G0, ak = D Gk
  unsigned pe[K];
  const unsigned *gt = gridTerms + gridOff[2 * p];
#pragma unroll
  for (int k = 0; k < K; k++) pe[k] = gt[k * 64 + lane];
  if (lane == 0) {
    R2 z;
    z.x = z.y = 0;
    cc[slotZ] = z;
  }
  R sg[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) sg[q] = lane + 64 * q < nm ? (R)T.ep[lane + 64 * q].w : (R)0;
  R acc[NR][4];
#pragma unroll
  for (int r = 0; r < NR; r++) acc[r][0] = acc[r][1] = acc[r][2] = acc[r][3] = 0;
  R ga[NQ], gb[NQ], ma[NQ], mb[NQ];
  auto fetch = [&](int pr) {
    const int src = m2lSrc[pr];
    const R *g0 = G + (long long)pr * 2 * nmMax, *gk = g0 + nmMax;
    const R *mp = mom + (long long)src * 2 * nmMax, *md = mp + nmMax;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const int i = lane + 64 * q;
      if (i < nm) {
        ga[q] = g0[i];
        gb[q] = gk[i];
        ma[q] = mp[i];
        mb[q] = md[i];
      }
    }
  };
  const int b1 = m2lOff[c + 1];
  int pr = m2lOff[c] + w;
  if (pr < b1) fetch(pr);
  for (; pr < b1; pr += nw) {
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const int i = lane + 64 * q;
      if (i < nm) {
        R2 x, g;
        x.x = sg[q] * ma[q];
        x.y = sg[q] * mb[q];
        cc[i] = x;
        g.x = ga[q];
        g.y = gb[q];
        gg[i] = g;
      }
    }
    waveSync();
    if (pr + nw < b1) fetch(pr + nw);
#pragma unroll
    for (int k = 0; k < K; k++) {
      const R2 x = cc[pe[k] & 0xffffu];
      const R2 g = gg[pe[k] >> 16];
      R *a = acc[k / L];
      a[0] += x.x * g.x;
      a[1] += x.x * g.y;
      a[2] += x.y * g.x;
      a[3] += x.y * g.y;
    }
    waveSync();
  }
  __syncthreads();
  R *red = reinterpret_cast<R *>(dynLds);   // [run][lane][4], over the slots
  for (int ww = 0; ww < nw; ww++) {
    if (w == ww) {
#pragma unroll
      for (int r = 0; r < NR; r++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          R *d = red + (r * 64 + lane) * 4 + q;
          *d = ww == 0 ? acc[r][q] : *d + acc[r][q];
        }
    }
    __syncthreads();
  }
  const int nmc = NMOM(ordMom[cubeLevel[c]]);
  const int *rs = redStart + gridOff[2 * p + 1];
  constexpr int NO = TR || DK ? 2 : 4;
  for (int a = tid; a < nmMax; a += blockDim.x) {
    R v[NO];
#pragma unroll
    for (int k = 0; k < NO; k++) v[k] = 0;
    if (a < nm && a < nmc) {
      R t1 = 0, t2 = 0, t3 = 0, t4 = 0;
      for (int i = rs[a]; i < rs[a + 1]; i++) {
        const R *q = red + 4 * redCell[i];
        t1 += q[0];
        t2 += q[1];
        t3 += q[2];
        t4 += q[3];
      }
      if constexpr (DK) {
        v[0] = t1;
        v[1] = t3;
      } else if constexpr (TR) {   // x = (sgn b, sgn a): t1 = b*G0, t2 = b*Gk, t3 = a*G0, t4 = a*Gk
        v[0] = eps * t4 - t3 + t2 - t1;
        v[1] = t3 - t4 + t1 - t2 / eps;
      } else {
        v[0] = t3 - t4;
        v[1] = eps * t2 - t1;
        v[2] = t3 - t4 / eps;
        v[3] = t2 - t1;
      }
    }
#pragma unroll
    for (int k = 0; k < NO; k++) lec[((long long)c * NO + k) * nmMax + a] = v[k];
  }
}

// src/expan.c:403-452 (transL2LWs) + convL2L (:295-326): one block per child,
// thread per output multi-index, (beta, alpha+beta) table in LDS
// NA = 4: the local expansions l1..l4; NA = 2: the transposed M2M of the
// adjoint operator (same translation, two arrays)
template <typename R, int NA = 4>
__global__ __launch_bounds__(512) void k_l2l(int levStart, const int *__restrict__ parent,
                                             const int *__restrict__ cubeLevel, const int *__restrict__ ordMom,
                                             const double *__restrict__ cubeX, R *__restrict__ lec, int nmMax,
                                             DevTabs T) {
  R *sm = reinterpret_cast<R *>(dynLds);
  const int kid = levStart + blockIdx.x, par = parent[kid];
  const int ordIn = ordMom[cubeLevel[par]], ordOut = ordMom[cubeLevel[kid]];
  const int nIn = NMOM(ordIn), nOut = NMOM(ordOut);
  R *Rt = sm, *In = sm + nIn;   // In[4][nIn]
  int *cst = (int *)(sm + 5 * nIn), *tab = cst + nOut + 1;
  const int nT = T.convStart[nOut];
  for (int i = threadIdx.x; i <= nOut; i += blockDim.x) cst[i] = T.convStart[i];
  for (int i = threadIdx.x; i < nT; i += blockDim.x) tab[i] = T.conv[i];
  double tr[3];
  for (int d = 0; d < 3; d++) tr[d] = cubeX[3 * kid + d] - cubeX[3 * par + d];
  for (int a = threadIdx.x; a < nIn; a += blockDim.x) {
    Rt[a] = (R)conVect1(tr, T.mi[a], T.ifact3[a]);
    for (int k = 0; k < NA; k++) In[k * nIn + a] = lec[((long long)par * NA + k) * nmMax + a];
  }
  __syncthreads();
  for (int a = threadIdx.x; a < nOut; a += blockDim.x) {
    const int cnt = NMOM(ordIn - deg_of(T.mi[a]));
    const int *cv = tab + cst[a];
    R t[NA] = {};
    for (int e = 0; e < cnt; e++) {
      int pe = cv[e], j = pe & 0xffff, kk = pe >> 16;
      R rv = Rt[j];
      for (int k = 0; k < NA; k++) t[k] += rv * In[k * nIn + kk];
    }
    for (int k = 0; k < NA; k++) lec[((long long)kid * NA + k) * nmMax + a] += t[k];
  }
}

// far[p] = Q2M0^T (l1 + l2), far[p+n] = Q2M1^T (l3 + l4) (src/fmm.c:815-828):
// one block per leaf, its panels in chunks of L2P_CH.  As in k_q2m, rows of
// nm / V threads read the columns V slots at a time in coalesced runs; each
// thread leaves its V-slot partial dot product in LDS and one thread per panel
// adds the row's partials in slot order.
// TR: the transposed Q2M of the adjoint operator, from the adjoint moments
// (M_pot | M_dpdn): far[p] = Q2M1^T M_pot, far[p+n] = Q2M0^T M_dpdn
#define L2P_CH 32
template <typename R, int V, bool TR = false>
__global__ __launch_bounds__(256) void k_l2p(const R *__restrict__ Q0, const R *__restrict__ Q1,
                                             const int *__restrict__ leafStart, const int *__restrict__ leafCount,
                                             const R *__restrict__ lec, R *__restrict__ far, int n, int nm,
                                             int nmMax) {
  typedef R VT __attribute__((ext_vector_type(V)));
  R *sm = reinterpret_cast<R *>(dynLds);
  R *L12 = sm, *L34 = sm + nm, *P0 = sm + 2 * nm;   // P0/P1: [L2P_CH][G] partial dot products
  const int L = blockIdx.x, t = threadIdx.x, G = nm / V;
  R *P1 = P0 + L2P_CH * G;
  if constexpr (TR) {
    const R *l = lec + (long long)L * 2 * nmMax;
    for (int a = t; a < nm; a += blockDim.x) {
      L12[a] = l[nmMax + a];
      L34[a] = l[a];
    }
  } else {
    const R *l = lec + (long long)L * 4 * nmMax;
    for (int a = t; a < nm; a += blockDim.x) {
      L12[a] = l[a] + l[nmMax + a];
      L34[a] = l[2 * nmMax + a] + l[3 * nmMax + a];
    }
  }
  __syncthreads();
  const int rows = max(1, 256 / G), g = t % G, r = t / G, s = leafStart[L], c = leafCount[L];
  VT l12, l34;
  for (int v = 0; v < V; v++) {
    l12[v] = L12[g * V + v];
    l34[v] = L34[g * V + v];
  }
  for (int j0 = 0; j0 < c; j0 += L2P_CH) {
    const int cn = min(L2P_CH, c - j0);
    if (r < rows) {
#pragma unroll 2
      for (int jj = r; jj < cn; jj += rows) {
        const long long col = (long long)(s + j0 + jj) * nm + g * V;
        const VT q0 = *reinterpret_cast<const VT *>(Q0 + col) * l12;
        const VT q1 = *reinterpret_cast<const VT *>(Q1 + col) * l34;
        R t0 = q0[0], t1 = q1[0];
        for (int v = 1; v < V; v++) {
          t0 += q0[v];
          t1 += q1[v];
        }
        P0[jj * G + g] = t0;
        P1[jj * G + g] = t1;
      }
    }
    __syncthreads();
    for (int jj = t; jj < cn; jj += blockDim.x) {
      R t0 = 0, t1 = 0;
      for (int k = 0; k < G; k++) {
        t0 += P0[jj * G + k];
        t1 += P1[jj * G + k];
      }
      far[s + j0 + jj + (TR ? n : 0)] = t0;
      far[s + j0 + jj + (TR ? 0 : n)] = t1;
    }
    __syncthreads();
  }
}

// L2L down one leaf's own ancestry and L2P, in one launch (src/expan.c:403-452,
// convL2L :295-326; src/fmm.c:815-828): one block per leaf.  The block starts
// from the M2L locals of its ancestor at the top FMM level and translates them
// down to the leaf one level at a time, adding each cube's own M2L locals on
// the way; siblings repeat the steps above their parent (arithmetic the
// level-by-level launches share) in exchange for no launch per level and no
// wait between levels.  A step L'[a] = sum_b t^b/b! L[a+b] (|a+b| <= p_in) is
// done one axis at a time, z, then y, then x, each pass
//   M[g] = sum_{k <= p_in - |g|} t_d^k/k! L[g + k e_d],
// which gives the same sum (the three bounds add up to |a| + |b| <= p_in) in
// at most p_in + 1 terms per slot instead of up to nMom(p_in); only the
// rounding differs from the level-by-level L2L.  The global locals are only
// read; the intermediate locals stay in LDS.  Then the leaf's panels as k_l2p.
#define L2LP_MAXUP 16
struct LevOrd {   // expansion order of the levels height .. depth
  int p[L2LP_MAXUP + 1];
};
__device__ __forceinline__ int idx3(int i1, int i2, int i3) {   // graded: degree, then i1, then i2
  const int d = i1 + i2 + i3;
  return NMOM(d - 1) + i1 * (d + 1) - i1 * (i1 - 1) / 2 + i2;
}
template <typename R, int V>
__global__ __launch_bounds__(256) void k_l2lp(const R *__restrict__ Q0, const R *__restrict__ Q1,
                                              const int *__restrict__ leafStart, const int *__restrict__ leafCount,
                                              const int *__restrict__ parent, LevOrd O,
                                              const double *__restrict__ cubeX, const R *__restrict__ lec,
                                              R *__restrict__ far, int n, int nm, int nmMax, int nUp, DevTabs T) {
  typedef R VT __attribute__((ext_vector_type(V)));
  typedef typename Vec4<R>::type R4;
  R *sm = reinterpret_cast<R *>(dynLds);
  __shared__ int chain[L2LP_MAXUP + 1];   // chain[nUp] = the leaf, chain[s - 1] = parent of chain[s]
  __shared__ R co[3][MAXP + 1];           // t_d^k / k! of the current step
  const int L = blockIdx.x, t = threadIdx.x;
  // locals [slot][l1..l4]: two working buffers, then the M2L locals of chain[1..nUp]
  R4 *cur = reinterpret_cast<R4 *>(sm), *nxt = cur + nmMax, *own = nxt + nmMax;
  if (t == 0) {
    int cb = L;
    for (int s = nUp; s >= 0; s--) {
      chain[s] = cb;
      if (s > 0) cb = parent[cb];
    }
  }
  __syncthreads();
  for (int s = 0; s <= nUp; s++) {   // every M2L local the block needs, loads all in flight at once
    const R *l = lec + (long long)chain[s] * 4 * nmMax;
    R4 *dst = s == 0 ? cur : own + (s - 1) * nmMax;
    for (int a = t; a < nmMax; a += blockDim.x) {
      R4 v;
      v.x = l[a];
      v.y = l[nmMax + a];
      v.z = l[2 * nmMax + a];
      v.w = l[3 * nmMax + a];
      dst[a] = v;
    }
  }
  for (int s = 1; s <= nUp; s++) {
    const int kid = chain[s], par = chain[s - 1];
    const int pin = O.p[s - 1], nIn = NMOM(pin), nOut = NMOM(O.p[s]);
    if (t < 3) {
      const double tr = cubeX[3 * kid + t] - cubeX[3 * par + t];
      double f = 1.0;
      for (int k = 0; k <= pin; k++) {
        co[t][k] = (R)f;
        f = f * tr / (k + 1);
      }
    }
    __syncthreads();
    for (int d = 2; d >= 0; d--) {   // z, y, x
      const R *cd = co[d];
      const R4 *lk = own + (s - 1) * nmMax;
      const int nO = d == 0 ? nOut : nIn;
      for (int a = t; a < nO; a += blockDim.x) {
        const int4 m = T.mi[a];
        const int i1 = m.x & 0xff, i2 = (m.x >> 8) & 0xff, i3 = (m.x >> 16) & 0xff, kmax = pin - deg_of(m);
        R4 u;
        u.x = u.y = u.z = u.w = 0;
        for (int k = 0; k <= kmax; k++) {
          const R4 v = cur[idx3(i1 + (d == 0 ? k : 0), i2 + (d == 1 ? k : 0), i3 + (d == 2 ? k : 0))];
          const R c = cd[k];
          u.x += c * v.x;
          u.y += c * v.y;
          u.z += c * v.z;
          u.w += c * v.w;
        }
        if (d == 0) {   // the kid's own M2L locals
          const R4 w = lk[a];
          u.x += w.x;
          u.y += w.y;
          u.z += w.z;
          u.w += w.w;
        }
        nxt[a] = u;
      }
      __syncthreads();
      R4 *sw = cur;
      cur = nxt;
      nxt = sw;
    }
  }
  __syncthreads();
  // L2P from the leaf's locals (k_l2p's arithmetic)
  const int G = nm / V, rows = max(1, 256 / G), g = t % G, r = t / G, s0 = leafStart[L], c = leafCount[L];
  VT l12, l34;
  for (int v = 0; v < V; v++) {
    const R4 w = cur[g * V + v];
    l12[v] = w.x + w.y;
    l34[v] = w.z + w.w;
  }
  __syncthreads();   // the partial sums below overlay the locals
  R *P0 = sm, *P1 = sm + L2P_CH * G;
  for (int j0 = 0; j0 < c; j0 += L2P_CH) {
    const int cn = min(L2P_CH, c - j0);
    if (r < rows) {
#pragma unroll 2
      for (int jj = r; jj < cn; jj += rows) {
        const long long col = (long long)(s0 + j0 + jj) * nm + g * V;
        const VT q0 = *reinterpret_cast<const VT *>(Q0 + col) * l12;
        const VT q1 = *reinterpret_cast<const VT *>(Q1 + col) * l34;
        R t0 = q0[0], t1 = q1[0];
        for (int v = 1; v < V; v++) {
          t0 += q0[v];
          t1 += q1[v];
        }
        P0[jj * G + g] = t0;
        P1[jj * G + g] = t1;
      }
    }
    __syncthreads();
    for (int jj = t; jj < cn; jj += blockDim.x) {
      R t0 = 0, t1 = 0;
      for (int k = 0; k < G; k++) {
        t0 += P0[jj * G + k];
        t1 += P1[jj * G + k];
      }
      far[s0 + j0 + jj] = t0;
      far[s0 + j0 + jj + n] = t1;
    }
    __syncthreads();
  }
}

// M2M of one level (src/expan.c:342-382, convM2M :223-247): one block per
// parent, one wave per child.  A translation M'[a] = sum_{b <= a} t^b/b! M[a-b]
// runs one axis at a time (z, y, x; at most p + 1 terms per slot instead of up
// to nMom(p) for the full convolution of k_m2m, so the block's chain of
// dependent LDS reads is short); only the rounding differs from k_m2m.  The
// children's contributions are added in child order.
template <typename R>
__global__ __launch_bounds__(512) void k_m2m_ax(int levStart, int pOut, int nIn, const int *__restrict__ kidStart,
                                                const int *__restrict__ nKids, const double *__restrict__ cubeX,
                                                R *__restrict__ mom, int nmMax, DevTabs T) {
  typedef typename Vec2<R>::type R2;
  R *sm = reinterpret_cast<R *>(dynLds);
  const int par = levStart + blockIdx.x, t = threadIdx.x, w = t >> 6, lane = t & 63, nw = blockDim.x >> 6;
  const int nk = nKids[par], k0 = kidStart[par], nOut = NMOM(pOut);
  R2 *kidT = reinterpret_cast<R2 *>(sm);   // [8][nmMax] translated children
  R2 *wa = kidT + 8 * nmMax + 2 * w * nmMax, *wb = wa + nmMax;
  R *co = reinterpret_cast<R *>(kidT + 8 * nmMax + 2 * nw * nmMax) + w * 3 * (MAXP + 1);
  for (int k = w; k < nk; k += nw) {
    const int kid = k0 + k;
    const R *mk = mom + (long long)kid * 2 * nmMax;
    for (int a = lane; a < nOut; a += 64) {   // children of a lower order: zeros above it
      R2 v;
      const int ai = a < nIn ? a : 0;
      const R x0 = mk[ai], x1 = mk[nmMax + ai];
      v.x = a < nIn ? x0 : (R)0;
      v.y = a < nIn ? x1 : (R)0;
      wa[a] = v;
    }
    if (lane < 3) {
      const double tr = cubeX[3 * kid + lane] - cubeX[3 * par + lane];
      double f = 1.0;
      for (int j = 0; j <= pOut; j++) {
        co[lane * (MAXP + 1) + j] = (R)f;
        f = f * tr / (j + 1);
      }
    }
    waveSync();
    R2 *src = wa, *dst = wb;
    for (int d = 2; d >= 0; d--) {   // z, y, x
      R2 *out = d == 0 ? kidT + k * nmMax : dst;
      const R *cd = co + d * (MAXP + 1);
      for (int a = lane; a < nOut; a += 64) {
        const int4 m = T.mi[a];
        const int i1 = m.x & 0xff, i2 = (m.x >> 8) & 0xff, i3 = (m.x >> 16) & 0xff;
        const int kmax = d == 0 ? i1 : d == 1 ? i2 : i3;
        R2 u;
        u.x = u.y = 0;
        for (int j = 0; j <= kmax; j++) {
          const R2 v = src[idx3(i1 - (d == 0 ? j : 0), i2 - (d == 1 ? j : 0), i3 - (d == 2 ? j : 0))];
          const R c = cd[j];
          u.x += c * v.x;
          u.y += c * v.y;
        }
        out[a] = u;
      }
      waveSync();
      R2 *sw = src;
      src = dst;
      dst = sw;
    }
  }
  __syncthreads();
  R *mp = mom + (long long)par * 2 * nmMax;
  for (int a = t; a < nOut; a += blockDim.x) {
    R s1 = 0, s2 = 0;
    for (int k = 0; k < nk; k++) {
      const R2 v = kidT[k * nmMax + a];
      s1 += v.x;
      s2 += v.y;
    }
    mp[a] = s1;
    mp[nmMax + a] = s2;
  }
}

// y = alpha*(D x - far - near) + beta*y (src/fabipb.c:1271-1279)
template <typename R>
__global__ void k_combine(const double *__restrict__ area, const double *__restrict__ x, const R *__restrict__ far,
                          const R *__restrict__ nearv, double *__restrict__ y, int n, double alpha, double beta,
                          double s1, double s2, int useFar, int useNear) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * n) return;
  int p = i < n ? i : i - n;
  double s = i < n ? s1 : s2;
  double k = (useFar ? (double)far[i] : 0.0) + (useNear ? (double)nearv[i] : 0.0);
  double v = alpha * (s * area[p] * x[i] - k);
  y[i] = beta == 0.0 ? v : v + beta * y[i];
}

// pot = beta*pot + alpha*(far + near) on the selected parts (stage API)
__global__ void k_accum(const double *__restrict__ far, const double *__restrict__ nearv, double *__restrict__ y,
                        int n2, double alpha, double beta, int useFar, int useNear) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n2) return;
  double k = (useFar ? far[i] : 0.0) + (useNear ? nearv[i] : 0.0);
  y[i] = (beta == 0.0 ? 0.0 : beta * y[i]) + alpha * k;
}

#pragma clang fp contract(off)
// --------------------------------------------------------------- RHS

// src/fabipb.c:789-798 with src/numQuad.c:786-819 (panelRHS), kernelRHS (kernel.c:152-167)
__global__ __launch_bounds__(256) void k_rhs(const Pnl *__restrict__ pnl, int n, int nChar,
                                             const double *__restrict__ pos, const double *__restrict__ chr,
                                             double *__restrict__ b, double fac, QTab Q, int q) {
  __shared__ double sp[256][4];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  Pnl P;
  if (i < n) P = pnl[i];
  double b0 = 0.0, b1 = 0.0;
  for (int j0 = 0; j0 < nChar; j0 += 256) {
    __syncthreads();
    int j = j0 + threadIdx.x;
    if (j < nChar) { sp[threadIdx.x][0] = pos[3 * j]; sp[threadIdx.x][1] = pos[3 * j + 1]; sp[threadIdx.x][2] = pos[3 * j + 2]; sp[threadIdx.x][3] = chr[j]; }
    __syncthreads();
    if (i < n) {
      int cnt = min(256, nChar - j0);
      for (int jj = 0; jj < cnt; jj++) {
        double r0[3] = {P.v[0] - sp[jj][0], P.v[1] - sp[jj][1], P.v[2] - sp[jj][2]};
        double I0 = 0.0, I1 = 0.0;
        for (int ix = 0; ix < q; ix++)
          for (int jx = 0; jx < q; jx++) {
            double r[3];
            for (int k = 0; k < 3; k++) r[k] = r0[k] + Q.t[ix] * (P.a[6 + k] + Q.t[jx] * P.a[k]);
            double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2], ri = 1.0 / sqrt(r2), r3i = ri / r2;
            double ip = P.nr[0] * r[0] + P.nr[1] * r[1] + P.nr[2] * r[2];
            I0 += ri * Q.t[ix] * Q.w[ix] * Q.w[jx];
            I1 += (-ip * r3i) * Q.t[ix] * Q.w[ix] * Q.w[jx];
          }
        I0 *= 2.0 * P.area;
        I1 *= 2.0 * P.area;
        b0 += sp[jj][3] * I0;
        b1 += sp[jj][3] * I1;
      }
    }
  }
  if (i < n) {
    b[i] = b0 * fac;
    b[i + n] = b1 * fac;
  }
}

#pragma clang fp contract(fast)
// the direct right-hand side at q = 1: one quadrature point per panel, so the
// point and its weight 2 A t w w leave the charge loop; 1/r from rsqrt
__global__ __launch_bounds__(256) void k_rhs_q1(const Pnl *__restrict__ pnl, int n, int nChar,
                                                const double *__restrict__ pos, const double *__restrict__ chr,
                                                double *__restrict__ b, double fac, double t, double ww) {
  __shared__ double4 sp[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double xq[3] = {0, 0, 0}, nr[3] = {0, 0, 0}, W = 0.0;
  if (i < n) {
    const Pnl &P = pnl[i];
    for (int k = 0; k < 3; k++) {
      xq[k] = P.v[k] + t * (P.a[6 + k] + t * P.a[k]);
      nr[k] = P.nr[k];
    }
    W = ww * 2.0 * P.area * fac;
  }
  double b0 = 0.0, b1 = 0.0;
  for (int j0 = 0; j0 < nChar; j0 += 256) {
    __syncthreads();
    const int j = j0 + threadIdx.x;
    if (j < nChar) sp[threadIdx.x] = make_double4(pos[3 * j], pos[3 * j + 1], pos[3 * j + 2], chr[j]);
    __syncthreads();
    const int cnt = min(256, nChar - j0);
#pragma unroll 4
    for (int jj = 0; jj < cnt; jj++) {
      const double4 c = sp[jj];
      const double rx = xq[0] - c.x, ry = xq[1] - c.y, rz = xq[2] - c.z;
      const double r2 = rx * rx + ry * ry + rz * rz, ri = rsqrt(r2);
      const double ip = nr[0] * rx + nr[1] * ry + nr[2] * rz;
      b0 += c.w * ri;
      b1 -= c.w * ip * (ri * ri * ri);
    }
  }
  if (i < n) {
    b[i] = b0 * W;
    b[i + n] = b1 * W;
  }
}

// Precision form: the direct right-hand side at q = 1 with the panel's point
// and the charges moved in FP64 into the frame of the block's first panel and
// rounded to FP32 there (r = x - x_j then carries an absolute error of a few
// 1e-7 of the charges' distance from that panel, well inside the FP32
// operator's error budget) and the kernel in FP32: 1/r from the hardware
// reciprocal square root (r2 is not denormal: the charges sit at least 1.5 A
// inside the surface, and 0 gives inf as the FP64 kernel does), sums in FP32
// over each staged block of 256 charges and in FP64 across blocks.
__global__ __launch_bounds__(256) void k_rhs32_q1(const Pnl *__restrict__ pnl, int n, int nChar,
                                                  const double *__restrict__ pos, const double *__restrict__ chr,
                                                  double *__restrict__ b, double fac, double t, double ww) {
  // FP32 coordinates in the frame of the block's first panel (its 256 panels are
  // consecutive in leaf order, so close together): the charges are moved into it
  // once per block as they are staged, and a pair's separation is then three
  // FP32 subtractions
  __shared__ float4 sp[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x, i0 = blockIdx.x * blockDim.x;
  double O[3];
  {
    const Pnl &P0 = pnl[i0];
    for (int k = 0; k < 3; k++) O[k] = P0.v[k] + t * (P0.a[6 + k] + t * P0.a[k]);
  }
  float xl[3] = {0, 0, 0}, nr[3] = {0, 0, 0};
  double W = 0.0;
  if (i < n) {
    const Pnl &P = pnl[i];
    for (int k = 0; k < 3; k++) {
      xl[k] = (float)(P.v[k] + t * (P.a[6 + k] + t * P.a[k]) - O[k]);
      nr[k] = (float)P.nr[k];
    }
    W = ww * 2.0 * P.area * fac;
  }
  double b0 = 0.0, b1 = 0.0;
  for (int j0 = 0; j0 < nChar; j0 += 256) {
    __syncthreads();
    const int j = j0 + threadIdx.x;
    if (j < nChar)
      sp[threadIdx.x] = make_float4((float)(pos[3 * j] - O[0]), (float)(pos[3 * j + 1] - O[1]),
                                    (float)(pos[3 * j + 2] - O[2]), (float)chr[j]);
    __syncthreads();
    const int cnt = min(256, nChar - j0);
    float s0 = 0.f, s1 = 0.f;
#pragma unroll 4
    for (int jj = 0; jj < cnt; jj++) {
      const float4 c = sp[jj];
      const float rx = xl[0] - c.x, ry = xl[1] - c.y, rz = xl[2] - c.z;
      const float r2 = fmaf(rx, rx, fmaf(ry, ry, rz * rz)), ri = __builtin_amdgcn_rsqf(r2);
      const float ip = nr[0] * rx + nr[1] * ry + nr[2] * rz;
      s0 = fmaf(c.w, ri, s0);
      s1 = fmaf(-c.w * ip, ri * ri * ri, s1);
    }
    b0 += s0;
    b1 += s1;
  }
  if (i < n) {
    b[i] = b0 * W;
    b[i + n] = b1 * W;
  }
}

// ------------------------------------------------ differentiation form
// Direct energy weights: the panel-tree energy (src/treecode.c:581-695,
// enDirect below) with every charge taken directly, as a vector,
//   E_raw = g . [sgm | sgm_n],  g[p] = 2A sum_pts t w w y1,  g[p+n] = 2A sum_pts t w w y0,
// y0 = sum_j q_j (G0 - Gk), y1 = sum_j q_j (eps dGk/dn - dG0/dn); gk = dg/dkappa:
//   d(G0 - Gk)/dkappa = e^{-kappa r}/(4 pi),  d(eps dGk/dn)/dkappa = eps dG0/dn (-kappa r^2 e^{-kappa r})
__global__ __launch_bounds__(256) void k_energy_g(const Pnl *__restrict__ pnl, int n, int nChar,
                                                  const double *__restrict__ pos, const double *__restrict__ chr,
                                                  double *__restrict__ g, double *__restrict__ gk, double kappa,
                                                  double eps, QTab Q, int q) {
  __shared__ double4 sp[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const Pnl P = pnl[min(i, n - 1)];
  double o0 = 0.0, o1 = 0.0, d0 = 0.0, d1 = 0.0;
  for (int j0 = 0; j0 < nChar; j0 += 256) {
    __syncthreads();
    const int j = j0 + threadIdx.x;
    if (j < nChar) sp[threadIdx.x] = make_double4(pos[3 * j], pos[3 * j + 1], pos[3 * j + 2], chr[j]);
    __syncthreads();
    const int cnt = min(256, nChar - j0);
    for (int ix = 0; ix < q; ix++)
      for (int jx = 0; jx < q; jx++) {
        double qp[3];
        for (int k = 0; k < 3; k++) qp[k] = P.v[k] + Q.t[ix] * (P.a[6 + k] + Q.t[jx] * P.a[k]);
        const double wq = Q.t[ix] * Q.w[ix] * Q.w[jx];
        double y0 = 0.0, y1 = 0.0, e0 = 0.0, e1 = 0.0;
        for (int jj = 0; jj < cnt; jj++) {
          const double4 c = sp[jj];
          const double x0 = qp[0] - c.x, x1 = qp[1] - c.y, x2 = qp[2] - c.z;
          const double r2 = x0 * x0 + x1 * x1 + x2 * x2, rn = sqrt(r2);
          const double G0 = kFourPiI / rn, ek = exp(-kappa * rn), Gk = ek * G0;
          const double coef = (kappa * rn + 1.0) * ek;
          const double ip = P.nr[0] * x0 + P.nr[1] * x1 + P.nr[2] * x2;
          const double dG0dn = -ip * G0 / r2, dGkdn = coef * dG0dn;
          y0 += c.w * (G0 - Gk);
          y1 += c.w * (eps * dGkdn - dG0dn);
          e0 += c.w * (ek * kFourPiI);
          e1 += c.w * (eps * dG0dn * (-kappa * r2 * ek));
        }
        o0 += wq * y0;
        o1 += wq * y1;
        d0 += wq * e0;
        d1 += wq * e1;
      }
  }
  if (i < n) {
    const double a2 = 2.0 * P.area;
    g[i] = a2 * o1;
    g[i + n] = a2 * o0;
    if (gk) {
      gk[i] = a2 * d1;
      gk[i + n] = a2 * d0;
    }
  }
}

// dE/dq_j = (G^T u)_j + (R^T lam)_j: the energy weights of charge j against
// the solution u = [sgm | sgm_n] plus the direct right-hand side
// (src/numQuad.c:786-819, kernelRHS src/kernel.c:152-167) transposed against
// the adjoint solution lam.  Block (charge tile, panel chunk); partial sums
// part[chunk][j] are added in chunk order by k_grad_sum.
#define GRAD_PCH 2048
// GP false: the right-hand side's term R^T lam alone (u is not read)
template <bool GP = true>
__global__ __launch_bounds__(256) void k_grad_q(const Pnl *__restrict__ pnl, int n, int nChar,
                                                const double *__restrict__ pos, const double *__restrict__ u,
                                                const double *__restrict__ lam, double *__restrict__ part,
                                                double kappa, double eps, double fac, QTab Q, int q) {
  __shared__ double sv[256][16];   // v0, a2, a0, n, 2A sgm, 2A sgm_n, fac 2A lam, fac 2A lam_n
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int p0 = blockIdx.y * GRAD_PCH, p1 = min(n, p0 + GRAD_PCH);
  double c[3] = {0, 0, 0};
  if (j < nChar)
    for (int k = 0; k < 3; k++) c[k] = pos[3 * j + k];
  double acc = 0.0;
  for (int b0 = p0; b0 < p1; b0 += 256) {
    __syncthreads();
    const int pi = b0 + threadIdx.x;
    if (pi < p1) {
      const Pnl &P = pnl[pi];
      double *v = sv[threadIdx.x];
      for (int k = 0; k < 3; k++) {
        v[k] = P.v[k];
        v[3 + k] = P.a[6 + k];
        v[6 + k] = P.a[k];
        v[9 + k] = P.nr[k];
      }
      const double a2 = 2.0 * P.area;
      v[12] = GP ? a2 * u[pi] : 0.0;
      v[13] = GP ? a2 * u[pi + n] : 0.0;
      v[14] = fac * a2 * lam[pi];
      v[15] = fac * a2 * lam[pi + n];
    }
    __syncthreads();
    const int cnt = min(256, p1 - b0);
    for (int jj = 0; jj < cnt; jj++) {
      const double *v = sv[jj];
      for (int ix = 0; ix < q; ix++)
        for (int jx = 0; jx < q; jx++) {
          const double wq = Q.t[ix] * Q.w[ix] * Q.w[jx];
          double x[3];
          for (int k = 0; k < 3; k++) x[k] = v[k] + Q.t[ix] * (v[3 + k] + Q.t[jx] * v[6 + k]) - c[k];
          const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2], rn = sqrt(r2), ri = 1.0 / rn;
          const double ip = v[9] * x[0] + v[10] * x[1] + v[11] * x[2];
          if constexpr (GP) {
            const double G0 = kFourPiI * ri, ek = exp(-kappa * rn), Gk = ek * G0;
            const double coef = (kappa * rn + 1.0) * ek;
            const double dG0dn = -ip * G0 / r2, dGkdn = coef * dG0dn;
            acc += wq * ((G0 - Gk) * v[13] + (eps * dGkdn - dG0dn) * v[12] + ri * v[14] - ip * ri * ri * ri * v[15]);
          } else {
            acc += wq * (ri * v[14] - ip * ri * ri * ri * v[15]);
          }
        }
    }
  }
  if (j < nChar) part[(long long)blockIdx.y * nChar + j] = acc;
}

// (R^T lam)_j at q = 1 (the transposed k_rhs_q1): block (charge tile, panel
// chunk); each panel's point, normal and two weights fac 2A t w w (lam,
// lam_n) staged once per chunk; 1/r from rsqrt; part[chunk][j] as k_grad_q
__global__ __launch_bounds__(256) void k_gradR_q1(const Pnl *__restrict__ pnl, int n, int nChar,
                                                  const double *__restrict__ pos, const double *__restrict__ lam,
                                                  double *__restrict__ part, double fac, double t, double ww) {
  __shared__ double4 sp[256], sn[256];   // (point, fac W lam), (normal, fac W lam_n)
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int p0 = blockIdx.y * GRAD_PCH, p1 = min(n, p0 + GRAD_PCH);
  double c[3] = {0, 0, 0};
  if (j < nChar)
    for (int k = 0; k < 3; k++) c[k] = pos[3 * j + k];
  double acc = 0.0;
  for (int b0 = p0; b0 < p1; b0 += 256) {
    __syncthreads();
    const int pi = b0 + threadIdx.x;
    if (pi < p1) {
      const Pnl &P = pnl[pi];
      double x[3];
      for (int k = 0; k < 3; k++) x[k] = P.v[k] + t * (P.a[6 + k] + t * P.a[k]);
      const double W = ww * 2.0 * P.area * fac;
      sp[threadIdx.x] = make_double4(x[0], x[1], x[2], W * lam[pi]);
      sn[threadIdx.x] = make_double4(P.nr[0], P.nr[1], P.nr[2], W * lam[pi + n]);
    }
    __syncthreads();
    const int cnt = min(256, p1 - b0);
#pragma unroll 4
    for (int jj = 0; jj < cnt; jj++) {
      const double4 a = sp[jj], m = sn[jj];
      const double rx = a.x - c[0], ry = a.y - c[1], rz = a.z - c[2];
      const double r2 = rx * rx + ry * ry + rz * rz, ri = rsqrt(r2);
      const double ip = m.x * rx + m.y * ry + m.z * rz;
      acc += a.w * ri - m.w * ip * (ri * ri * ri);
    }
  }
  if (j < nChar) part[(long long)blockIdx.y * nChar + j] = acc;
}

__global__ void k_grad_sum(const double *__restrict__ part, int nChunk, int nChar, double *__restrict__ out,
                           int add = 0) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nChar) return;
  double v = 0.0;
  for (int k = 0; k < nChunk; k++) v += part[(long long)k * nChar + j];
  out[j] = add ? out[j] + v : v;
}

// near part of lam^T (dK/dkappa) u at q = 1: every source of the row's near
// list but the panel itself by the one-point rule A_x A_y dker4(p_x - p_y),
// FP64, plus the stored corrections of the self / shared-vertex pairs.
// Tile as k_near_apply; part[x] = A_x (lam_x s0 + lam_{x+n} s1) + corrections.
__global__ __launch_bounds__(256) void k_near_dk(const int2 *__restrict__ tiles, const int *__restrict__ leafStart,
                                                 const int *__restrict__ leafCount, const int *__restrict__ srcOff,
                                                 const int *__restrict__ ncolArr, const int *__restrict__ srcList,
                                                 const Pnl *__restrict__ pnl, const double *__restrict__ u,
                                                 const double *__restrict__ lam,
                                                 const long long *__restrict__ specOff,
                                                 const int *__restrict__ specCol,
                                                 const double4 *__restrict__ specDk, double *__restrict__ part,
                                                 int n, double kappa, double eps, double t) {
  __shared__ double4 sP[256], sN[256];
  __shared__ int sY[256];
  __shared__ double T[256];   // 2^(j/256) for expNeg
  T[threadIdx.x] = gExp2Tab[threadIdx.x];
  const double ieps = 1.0 / eps;
  const int2 tl = tiles[blockIdx.x];
  const int D = tl.x, r0 = tl.y;
  const int rows = min(NEAR_ROWS, leafCount[D] - r0);
  const int ncol = ncolArr[D];
  const int *src = srcList + srcOff[D];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double a0[4] = {0, 0, 0, 0}, a1[4] = {0, 0, 0, 0};
  for (int c0 = 0; c0 < ncol; c0 += 256) {
    const int cn = min(256, ncol - c0);
    __syncthreads();
    if (threadIdx.x < cn) {
      const int y = src[c0 + threadIdx.x];
      const Pnl &Y = pnl[y];
      double q[3];
      for (int k = 0; k < 3; k++) q[k] = Y.v[k] + t * (Y.a[6 + k] + t * Y.a[k]);
      sP[threadIdx.x] = make_double4(q[0], q[1], q[2], Y.area * u[y]);
      sN[threadIdx.x] = make_double4(Y.nr[0], Y.nr[1], Y.nr[2], Y.area * u[y + n]);
      sY[threadIdx.x] = y;
    }
    __syncthreads();
    for (int r = 0; r < 4; r++) {
      const int row = w + 4 * r;
      if (row >= rows) continue;
      const int xi = leafStart[D] + r0 + row;
      const Pnl &X = pnl[xi];
      double px[3];
      for (int k = 0; k < 3; k++) px[k] = X.v[k] + t * (X.a[6 + k] + t * X.a[k]);
      double s0 = 0.0, s1 = 0.0;
      for (int j = lane; j < cn; j += 64) {
        if (sY[j] == xi) continue;
        const double4 P = sP[j], N = sN[j];
        // d ker4 / d kappa (ker4dk) with 1/r from rsq64 and e^{-kappa r} from expNeg
        const double rx = px[0] - P.x, ry = px[1] - P.y, rz = px[2] - P.z;
        const double r2 = rx * rx + ry * ry + rz * rz, ri = rsq64(r2), rn = r2 * ri, ri2 = ri * ri;
        const double ek = expNeg(-kappa * rn, T);
        const double ip0 = X.nr[0] * N.x + X.nr[1] * N.y + X.nr[2] * N.z;
        const double ipX = X.nr[0] * rx + X.nr[1] * ry + X.nr[2] * rz;
        const double ipY = N.x * rx + N.y * ry + N.z * rz;
        const double G0 = ri * kFourPiI, dG0dy = ipY * G0 * ri2, dG0dx = ipX * G0 * ri2;
        const double ddG0 = (ip0 * G0 - dG0dy * ipX * 3.0) * ri2;
        const double dcoef = -kappa * r2 * ek;
        const double f0 = ek * kFourPiI, f1 = eps * dcoef * dG0dy, f2 = dcoef * dG0dx * ieps;
        const double f3 = dcoef * ddG0 - kappa * ek * (2.0 - kappa * rn) * ipX * dG0dy;
        s0 += f0 * N.w + f1 * P.w;
        s1 += f2 * N.w + f3 * P.w;
      }
      a0[r] += s0;
      a1[r] += s1;
    }
  }
  for (int r = 0; r < 4; r++) {
    const int row = w + 4 * r;
    const int xi = leafStart[D] + r0 + min(row, rows - 1);
    const double A = pnl[xi].area;
    double t0 = A * a0[r], t1 = A * a1[r];
    if (row < rows)
      for (long long e = specOff[xi] + lane; e < specOff[xi + 1]; e += 64) {
        const int y = specCol[e];
        const double4 d = specDk[e];
        t0 += d.x * u[y + n] + d.y * u[y];
        t1 += d.z * u[y + n] + d.w * u[y];
      }
    const double v = waveSumT(lam[xi] * t0 + lam[xi + n] * t1);
    if (row < rows && lane == 0) part[xi] = v;
  }
}

// the same over every near pair with the full Galerkin rules (any q)
__global__ __launch_bounds__(256) void k_near_dk_rules(const Pnl *__restrict__ pnl, const int *__restrict__ leafOf,
                                                       const int *__restrict__ srcOff,
                                                       const int *__restrict__ ncolArr,
                                                       const int *__restrict__ srcList, const double *__restrict__ u,
                                                       const double *__restrict__ lam, double *__restrict__ part,
                                                       int n, double kappa, double eps, QTab Q, int q) {
  __shared__ double red[256];
  const int x = blockIdx.x, D = leafOf[x], nc = ncolArr[D];
  const int *src = srcList + srcOff[D];
  const Pnl X = pnl[x];
  double v = 0.0;
  for (int j = threadIdx.x; j < nc; j += blockDim.x) {
    const int y = src[j];
    const Pnl Y = pnl[y];
    double k[4];
    panelIA0(X, Y, x == y, kappa, eps, Q, q, k, true);
    v += lam[x] * (k[0] * u[y + n] + k[1] * u[y]) + lam[x + n] * (k[2] * u[y + n] + k[3] * u[y]);
  }
  red[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[x] = red[0];
}

// far part of lam^T (dK/dkappa) u: per M2L pair (D <- S) the Gk terms of the
// pair differentiated in kappa, contracted with the adjoint locals of D
// (adjL = b | a) and the moments of S:
//   sum_a (eps a + b)[a] (c1 * dGk)[a] - (a + b/eps)[a] (c2 * dGk)[a],
// c1 = sgn M_pot, c2 = sgn M_dpdn, (c * T)[a] = sum_b c[b] T[a+b].  One wave
// per destination cube; part[D] is its sum.
__global__ __launch_bounds__(64) void k_m2l_dk(const int *__restrict__ m2lOff, const int *__restrict__ m2lSrc,
                                               const int *__restrict__ cubeLevel, const int *__restrict__ ordM2L,
                                               const int *__restrict__ ordMom, const double *__restrict__ cubeX,
                                               const double *__restrict__ mom, const double *__restrict__ adjL,
                                               double *__restrict__ part, int nmMax, double kappa, double eps,
                                               DevTabs T) {
  extern __shared__ double sm[];
  const int c = blockIdx.x, tid = threadIdx.x;
  const int p = ordM2L[cubeLevel[c]], nm = NMOM(p), dsz = derivScratch(p);
  const int nmc = NMOM(ordMom[cubeLevel[c]]);
  double *s0 = sm, *sk = s0 + dsz, *Tk = sk + dsz, *Gk = Tk + nm, *c1 = Gk + nm, *c2 = c1 + nm;
  double *A1 = c2 + nm, *A2 = A1 + nm;
  for (int i = tid; i < nm; i += 64) {
    const double b = adjL[(long long)c * 2 * nmMax + i], a = adjL[(long long)c * 2 * nmMax + nmMax + i];
    A1[i] = i < nmc ? eps * a + b : 0.0;
    A2[i] = i < nmc ? a + b / eps : 0.0;
  }
  double acc = 0.0;
  for (int pr = m2lOff[c]; pr < m2lOff[c + 1]; pr++) {
    const int src = m2lSrc[pr];
    double r[3];
    for (int d = 0; d < 3; d++) r[d] = cubeX[3 * c + d] - cubeX[3 * src + d];
    derivs_block(kappa, p, r, s0, sk, Tk, Gk, T.mi, tid, 64, true);
    for (int i = tid; i < nm; i += 64) {
      const double sg = T.ep[i].w;
      c1[i] = sg * mom[(long long)src * 2 * nmMax + i];
      c2[i] = sg * mom[(long long)src * 2 * nmMax + nmMax + i];
    }
    __syncthreads();
    for (int a = tid; a < nm; a += 64) {
      const int cnt = NMOM(p - deg_of(T.mi[a]));
      const int *cv = T.conv + T.convStart[a];
      double t2 = 0.0, t4 = 0.0;
      for (int e = 0; e < cnt; e++) {
        const int pe = cv[e], j = pe & 0xffff, kk = pe >> 16;
        t2 += c1[j] * Tk[kk];
        t4 += c2[j] * Tk[kk];
      }
      acc += A1[a] * t2 - A2[a] * t4;
    }
    __syncthreads();
  }
  acc = waveSumT(acc);
  if (tid == 0) part[c] = acc;
}

// the same sum from k_m2l_grid<DK>'s (t1, t3) per cube: part[c] =
// sum_a (eps a + b)[a] t1[a] - (a + b / eps)[a] t3[a], (b, a) the adjoint locals
__global__ __launch_bounds__(64) void k_dk_far_dot(const int *__restrict__ cubeLevel, const int *__restrict__ ordMom,
                                                   const double *__restrict__ adjL, const double *__restrict__ t13,
                                                   double *__restrict__ part, int nmMax, double eps) {
  const int c = blockIdx.x, tid = threadIdx.x;
  const int nmc = NMOM(ordMom[cubeLevel[c]]);
  double acc = 0.0;
  for (int i = tid; i < nmc; i += 64) {
    const double b = adjL[(long long)c * 2 * nmMax + i], a = adjL[(long long)c * 2 * nmMax + nmMax + i];
    acc += (eps * a + b) * t13[(long long)c * 2 * nmMax + i] - (a + b / eps) * t13[(long long)c * 2 * nmMax + nmMax + i];
  }
  acc = waveSumT(acc);
  if (tid == 0) part[c] = acc;
}
#pragma clang fp contract(off)

// ------------------------------------------------------------ panel tree, the panel pass
// (src/gkSetup.c:90-183 bounding box and goDown, :264-301 leaf boxes).  No
// contraction here: the comparisons against the cube midplanes must see the
// same doubles as the host's.

// per block: min (0..2) and max (3..5) of the panels' vertex coordinates
__global__ __launch_bounds__(256) void k_tree_bbox(const double *__restrict__ vtx, int n, double *__restrict__ part) {
  __shared__ double sm[6][256];
  double lo[3], up[3];
  for (int k = 0; k < 3; k++) {
    lo[k] = __builtin_inf();
    up[k] = -__builtin_inf();
  }
  for (int p = blockIdx.x * 256 + threadIdx.x; p < n; p += gridDim.x * 256)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) {
        const double v = vtx[9 * (long long)p + 3 * j + k];
        lo[k] = v < lo[k] ? v : lo[k];
        up[k] = v > up[k] ? v : up[k];
      }
  for (int k = 0; k < 3; k++) {
    sm[k][threadIdx.x] = lo[k];
    sm[3 + k][threadIdx.x] = up[k];
  }
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int k = 0; k < 3; k++) {
        const double a = sm[k][threadIdx.x + o], b = sm[3 + k][threadIdx.x + o];
        sm[k][threadIdx.x] = a < sm[k][threadIdx.x] ? a : sm[k][threadIdx.x];
        sm[3 + k][threadIdx.x] = b > sm[3 + k][threadIdx.x] ? b : sm[3 + k][threadIdx.x];
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = sm[threadIdx.x][0];
}

struct TreeSides {
  double s[11];   // cube side per level 0..depth (depth <= 10)
};

// each panel's leaf: its centroid against the midplanes of the cube it is in,
// level by level (strict >), the octant digits ix + 2 iy + 4 iz concatenated
// coarsest first; idx = the panel
__global__ __launch_bounds__(256) void k_tree_keys(const double *__restrict__ vtx, int n, double minx, double miny,
                                                   double minz, TreeSides side, int depth,
                                                   unsigned *__restrict__ key, int *__restrict__ idx) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const double *v = vtx + 9 * (long long)p;
  double c[3];
  for (int k = 0; k < 3; k++) c[k] = 0.33333333333333333333333333 * (v[k] + v[3 + k] + v[6 + k]);
  int i = 0, j = 0, l = 0;
  unsigned kk = 0;
  for (int lev = 0; lev < depth; lev++) {
    const int ix = c[0] > minx + (i + 0.5) * side.s[lev] ? 1 : 0;
    const int iy = c[1] > miny + (j + 0.5) * side.s[lev] ? 1 : 0;
    const int iz = c[2] > minz + (l + 0.5) * side.s[lev] ? 1 : 0;
    kk = (kk << 3) | (unsigned)(ix + 2 * iy + 4 * iz);
    i = 2 * i + ix;
    j = 2 * j + iy;
    l = 2 * l + iz;
  }
  key[p] = kk;
  idx[p] = p;
}

// the panel arrays in leaf order
__global__ __launch_bounds__(256) void k_tree_gather(const int *__restrict__ perm, int n, const double *__restrict__ vtx,
                                                     const double *__restrict__ nrm, const double *__restrict__ area,
                                                     const int *__restrict__ face, double *__restrict__ vtxR,
                                                     double *__restrict__ nrmR, double *__restrict__ areaR,
                                                     int *__restrict__ faceR) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const long long p = perm[i];
  for (int k = 0; k < 9; k++) vtxR[9 * (long long)i + k] = vtx[9 * p + k];
  for (int k = 0; k < 3; k++) nrmR[3 * (long long)i + k] = nrm[3 * p + k];
  areaR[i] = area[p];
  faceR[i] = face[p];
}

// per leaf (a run of equal keys): lo and up of its panels' vertices
__global__ __launch_bounds__(64) void k_tree_leafbox(const int *__restrict__ start, const int *__restrict__ count,
                                                     const double *__restrict__ vtxR, double *__restrict__ box) {
  __shared__ double sm[6][64];
  const int L = blockIdx.x, s0 = start[L], s1 = s0 + count[L];
  double lo[3], up[3];
  for (int k = 0; k < 3; k++) {
    lo[k] = __builtin_inf();
    up[k] = -__builtin_inf();
  }
  for (int i = s0 + threadIdx.x; i < s1; i += 64)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) {
        const double v = vtxR[9 * (long long)i + 3 * j + k];
        lo[k] = v < lo[k] ? v : lo[k];
        up[k] = v > up[k] ? v : up[k];
      }
  for (int k = 0; k < 3; k++) {
    sm[k][threadIdx.x] = lo[k];
    sm[3 + k][threadIdx.x] = up[k];
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    double r = sm[k][0];
    for (int t = 1; t < 64; t++) r = k < 3 ? (sm[k][t] < r ? sm[k][t] : r) : (sm[k][t] > r ? sm[k][t] : r);
    box[6 * (long long)L + k] = r;
  }
}

// ------------------------------------------------------------ charge tree

// leaf moments of the charges (src/chargeTree.c:313-337): block per leaf
__global__ void k_ct_leaf(int levStart, const int *__restrict__ chgStart, const int *__restrict__ nChgs,
                          const int *__restrict__ chgIdx, const double *__restrict__ ctx_,
                          const double *__restrict__ pos, const double *__restrict__ chr, int order,
                          double *__restrict__ cmom, int nmMax, DevTabs T) {
  const int c = levStart + blockIdx.x, nm = NMOM(order);
  for (int a = threadIdx.x; a < nm; a += blockDim.x) {
    double m = 0.0;
    int4 mi = T.mi[a];
    for (int k = 0; k < nChgs[c]; k++) {
      int j = chgIdx[chgStart[c] + k];
      double tr[3] = {pos[3 * j] - ctx_[3 * c], pos[3 * j + 1] - ctx_[3 * c + 1], pos[3 * j + 2] - ctx_[3 * c + 2]};
      m += chr[j] * conVect1(tr, mi, T.ifact3[a]);
    }
    cmom[(long long)c * nmMax + a] = m;
  }
}

// upward pass of the charge moments (src/chargeTree.c:344-395, transM2Mchr)
__global__ void k_ct_m2m(int levStart, const int *__restrict__ kidStart, const int *__restrict__ nKids,
                         const double *__restrict__ ctx_, const int *__restrict__ ordOf, int lev,
                         double *__restrict__ cmom, int nmMax, DevTabs T) {
  extern __shared__ double sm[];
  const int par = levStart + blockIdx.x;
  const int ordOut = ordOf[lev], nOut = NMOM(ordOut), nIn = NMOM(ordOf[lev + 1]);
  double *R = sm, *b = sm + nOut;
  // the (alpha - k, k) pair table of the outputs, staged once: the inner loop
  // then reads LDS instead of waiting on global loads term by term
  int *tS = reinterpret_cast<int *>(b + nOut), *tab = tS + nOut + 1;
  const int t0 = T.m2mStart[0];
  for (int a = threadIdx.x; a <= nOut; a += blockDim.x) tS[a] = T.m2mStart[a] - t0;
  for (int e = threadIdx.x; e < T.m2mStart[nOut] - t0; e += blockDim.x) tab[e] = T.m2m[t0 + e];
  double acc[4] = {0, 0, 0, 0};
  for (int k = 0; k < nKids[par]; k++) {
    const int kid = kidStart[par] + k;
    double tr[3];
    for (int d = 0; d < 3; d++) tr[d] = ctx_[3 * kid + d] - ctx_[3 * par + d];
    __syncthreads();
    for (int a = threadIdx.x; a < nOut; a += blockDim.x) {
      R[a] = conVect1(tr, T.mi[a], T.ifact3[a]);
      b[a] = a < nIn ? cmom[(long long)kid * nmMax + a] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const int a = threadIdx.x + s * blockDim.x;
      if (a < nOut) {
        double t = 0.0;
        for (int e = tS[a]; e < tS[a + 1]; e++) {
          const int pr = tab[e];
          if ((pr >> 16) < nIn) t += R[pr & 0xffff] * b[pr >> 16];
        }
        acc[s] += t;
      }
    }
  }
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const int a = threadIdx.x + s * blockDim.x;
    if (a < nOut) cmom[(long long)par * nmMax + a] = acc[s];
  }
}

#define EN_STACK 512

// Panel-tree energy (src/treecode.c:617-695 energyTreeWalk/panelEnergyTree,
// :581-615 chgClusterEnergyEval): one wave per panel walks the charge tree
// depth first, kids in order, accepting clusters with eRad < theta*dist at
// levels >= height.  part[p] = I0*sgm_n[p] + I1*sgm[p].
__global__ __launch_bounds__(64) void k_energy(const Pnl *__restrict__ pnl, int n, const double *__restrict__ sgm,
                                               const int *__restrict__ level, const double *__restrict__ cx,
                                               const double *__restrict__ eRad, const int *__restrict__ kidStart,
                                               const int *__restrict__ nKids, const int *__restrict__ chgStart,
                                               const int *__restrict__ nChgs, const int *__restrict__ chgIdx,
                                               const double *__restrict__ pos, const double *__restrict__ chr,
                                               const int *__restrict__ ordOf, const double *__restrict__ cmom,
                                               int nmMax, int depth, int height, double theta, double kappa,
                                               double eps, QTab Q, int q, DevTabs T, double *__restrict__ part) {
  extern __shared__ double sm[];
  __shared__ int stack[EN_STACK];
  const int p = blockIdx.x, lane = threadIdx.x;
  const int P1 = ordOf[depth + 1];   // largest derivative order used
  const int dsz = derivScratch(P1), nmP = NMOM(P1);
  double *s0 = sm, *sk = s0 + dsz, *g0 = sk + dsz, *gk = g0 + nmP;
  const Pnl X = pnl[p];
  double o0 = 0.0, o1 = 0.0;
  for (int ix = 0; ix < q; ix++)
    for (int jx = 0; jx < q; jx++) {
      double qp[3];
      for (int k = 0; k < 3; k++) qp[k] = X.v[k] + Q.t[ix] * (X.a[6 + k] + Q.t[jx] * X.a[k]);
      double y0 = 0.0, y1 = 0.0;
      int sp = 0;
      stack[0] = 0;
      sp = 1;
      __syncthreads();
      while (sp > 0) {
        const int cb = stack[sp - 1];
        sp--;
        double r[3] = {qp[0] - cx[3 * cb], qp[1] - cx[3 * cb + 1], qp[2] - cx[3 * cb + 2]};
        double dist = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        const int lev = level[cb];
        if (eRad[cb] < theta * dist && lev >= height) {
          const int order = ordOf[lev];
          derivs_block(kappa, order + 1, r, s0, sk, g0, gk, T.mi, lane, 64);
          const double *mom = cmom + (long long)cb * nmMax;
          double t0 = 0.0, t1 = 0.0;
          for (int a = lane; a < NMOM(order); a += 64) {
            int4 e = T.ep[a];
            double sm_ = e.w * mom[a];
            t0 += sm_ * (g0[a] - gk[a]);
            double dnG0 = X.nr[0] * g0[e.x] + X.nr[1] * g0[e.y] + X.nr[2] * g0[e.z];
            double dnGk = X.nr[0] * gk[e.x] + X.nr[1] * gk[e.y] + X.nr[2] * gk[e.z];
            t1 += sm_ * (eps * dnGk - dnG0);
          }
          y0 += waveSum(t0);
          y1 += waveSum(t1);
          __syncthreads();
        } else if (lev == depth) {
          double t0 = 0.0, t1 = 0.0;
          for (int i = lane; i < nChgs[cb]; i += 64) {
            int j = chgIdx[chgStart[cb] + i];
            double x[3] = {qp[0] - pos[3 * j], qp[1] - pos[3 * j + 1], qp[2] - pos[3 * j + 2]};
            double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2], rn = sqrt(r2);
            double G0 = kFourPiI / rn, ek = exp(-kappa * rn), Gk = ek * G0;
            double coef = (kappa * rn + 1.0) * ek;
            double ip = X.nr[0] * x[0] + X.nr[1] * x[1] + X.nr[2] * x[2];
            double dG0dn = -ip * G0 / r2, dGkdn = coef * dG0dn;
            t0 += chr[j] * (G0 - Gk);
            t1 += chr[j] * (eps * dGkdn - dG0dn);
          }
          y0 += waveSum(t0);
          y1 += waveSum(t1);
        } else {
          const int nk = nKids[cb], k0 = kidStart[cb];
          if (sp + nk > EN_STACK) { sp = 0; y0 = y1 = __builtin_nan(""); break; }   // stack bound: poison, never fault
          for (int k = nk - 1; k >= 0; k--) stack[sp++] = k0 + k;   // pushed in reverse: kids visited in order
        }
      }
      o0 += y0 * Q.t[ix] * Q.w[ix] * Q.w[jx];
      o1 += y1 * Q.t[ix] * Q.w[ix] * Q.w[jx];
    }
  o0 *= 2.0 * X.area;
  o1 *= 2.0 * X.area;
  if (lane == 0) part[p] = o0 * sgm[p + n] + o1 * sgm[p];
}

// one charge of a leaf the walk did not accept (src/treecode.c:581-615): the
// terms of the panel-tree energy for the point qp with panel normal nr
__device__ inline void enDirect(const double *qp, const double *nr, const double *pj, double qj, double kappa,
                                double eps, double &y0, double &y1) {
  double x[3] = {qp[0] - pj[0], qp[1] - pj[1], qp[2] - pj[2]};
  double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2], rn = sqrt(r2);
  double G0 = kFourPiI / rn, ek = exp(-kappa * rn), Gk = ek * G0;
  double coef = (kappa * rn + 1.0) * ek;
  double ip = nr[0] * x[0] + nr[1] * x[1] + nr[2] * x[2];
  double dG0dn = -ip * G0 / r2, dGkdn = coef * dG0dn;
  y0 += qj * (G0 - Gk);
  y1 += qj * (eps * dGkdn - dG0dn);
}

// d/dkappa of enDirect: d(G0 - Gk)/dkappa = e^{-kappa r}/(4 pi),
// d(eps dGk/dn - dG0/dn)/dkappa = eps dG0/dn (-kappa r^2 e^{-kappa r})
__device__ inline void enDirectDk(const double *qp, const double *nr, const double *pj, double qj, double kappa,
                                  double eps, double &y0, double &y1) {
  double x[3] = {qp[0] - pj[0], qp[1] - pj[1], qp[2] - pj[2]};
  double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2], rn = sqrt(r2);
  double ek = exp(-kappa * rn), G0 = kFourPiI / rn;
  double ip = nr[0] * x[0] + nr[1] * x[1] + nr[2] * x[2];
  double dG0dn = -ip * G0 / r2;
  y0 += qj * (kFourPiI * ek);
  y1 += qj * (eps * dG0dn * (-kappa * r2 * ek));
}

// one charge of an unaccepted leaf in the tree right-hand side
// (src/treecode.c:398-411): y0 += q / r, y1 += -q (n.r) / r^3
__device__ inline void rhsDirect(const double *qp, const double *nr, const double *pj, double qj, double &y0,
                                 double &y1) {
  double x[3] = {qp[0] - pj[0], qp[1] - pj[1], qp[2] - pj[2]};
  double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2], ri = 1.0 / sqrt(r2), r3i = ri / r2;
  double ip = nr[0] * x[0] + nr[1] * x[1] + nr[2] * x[2];
  y0 += qj * ri;
  y1 += qj * (-ip * r3i);
}

#pragma clang fp contract(fast)
// ------------------------------------- panel-tree energy, wave-coherent walk
//
// The far term of a charge cluster with moments M (order p) at r = x - x_c is
// sum_a s_a M_a d^a G(r) (and its normal derivative), s_a = (-1)^|a|.  With the
// closed form of the Cartesian derivatives of a radial function,
//   d^a f(|r|) = sum_{2b <= a} a! / ((a-2b)! b! 2^|b|) r^(a-2b) f_(|a|-|b|),
//   f_k = ((1/r) d/dr)^k f   (the values radial() returns),
// the far term is a polynomial in r whose coefficients combine the f_k:
//   sum_g r^g sum_j W[g][j] (G0 - Gk)_(|g|+j)
//   sum_g r^g sum_j (n . V[g][j]) (eps Gk - G0)_(|g|+j)
// W and V depend on the cluster only, so they are formed once per cluster
// (k_ct_poly) and every point evaluates a fixed, unrolled polynomial.
// Entry order (k_ct_poly and enFar agree on it): W then V; g1 descending,
// g2 descending, g3 descending (the Horner order), then j; a V entry holds
// three values.

__host__ __device__ constexpr int enPolyW(int p) {
  int s = 0;
  for (int d = 0; d <= p; d++) s += (d + 1) * (d + 2) / 2 * ((p - d) / 2 + 1);
  return s;
}
__host__ __device__ constexpr int enPolyV(int p) {
  int s = 0;
  for (int d = 0; d <= p + 1; d++) s += (d + 1) * (d + 2) / 2 * ((p + 1 - d) / 2 + 1);
  return s;
}
__host__ __device__ constexpr int enPolySize(int p) { return enPolyW(p) + 3 * enPolyV(p); }

__device__ inline int gidx3(int i1, int i2, int i3) {   // graded index: degree, then i1, then i2 ascending
  const int d = i1 + i2 + i3;
  return (d > 0 ? NMOM(d - 1) : 0) + i1 * (d + 1) - i1 * (i1 - 1) / 2 + i2;
}

__device__ inline double factd(int k) {
  double f = 1.0;
  for (int i = 2; i <= k; i++) f *= i;
  return f;
}

// per cluster at levels >= height: the W and V coefficients from the moments.
// desc[e] = (g1 | g2 << 8 | g3 << 16, j, kind 0 = W / 1 = V, output offset)
__global__ __launch_bounds__(256) void k_ct_poly(int c0, const double *__restrict__ cmom, int nmMax,
                                                 const int4 *__restrict__ desc, int nDesc, int size,
                                                 double *__restrict__ poly) {
  const int c = c0 + blockIdx.x;
  const double *M = cmom + (long long)c * nmMax;
  double *out = poly + (long long)blockIdx.x * size;
  for (int e = threadIdx.x; e < nDesc; e += blockDim.x) {
    const int4 ds = desc[e];
    const int g[3] = {ds.x & 0xff, (ds.x >> 8) & 0xff, (ds.x >> 16) & 0xff};
    const int j = ds.y, kind = ds.z, o = ds.w;
    const double gsgn = ((g[0] + g[1] + g[2]) & 1) ? -1.0 : 1.0;
    const double igf = 1.0 / (factd(g[0]) * factd(g[1]) * factd(g[2]) * (double)(1 << j));
    double v[3] = {0.0, 0.0, 0.0};
    for (int b1 = 0; b1 <= j; b1++)
      for (int b2 = 0; b2 <= j - b1; b2++) {
        const int b[3] = {b1, b2, j - b1 - b2};
        const int ap[3] = {g[0] + 2 * b[0], g[1] + 2 * b[1], g[2] + 2 * b[2]};
        const double C = factd(ap[0]) * factd(ap[1]) * factd(ap[2]) * igf / (factd(b[0]) * factd(b[1]) * factd(b[2]));
        if (kind == 0) {
          v[0] += gsgn * M[gidx3(ap[0], ap[1], ap[2])] * C;
        } else {
          for (int i = 0; i < 3; i++) {
            if (ap[i] < 1) continue;
            int a[3] = {ap[0], ap[1], ap[2]};
            a[i] -= 1;
            v[i] += -gsgn * M[gidx3(a[0], a[1], a[2])] * C;
          }
        }
      }
    out[o] = v[0];
    if (kind) {
      out[o + 1] = v[1];
      out[o + 2] = v[2];
    }
  }
}

#ifndef EN_CHUNK
#define EN_CHUNK -1   // doubles of walk coefficients per scalar-load chunk (0: no chunking; -1: per order)
#endif
// per order: the largest chunk that leaves the scalar registers unspilled
// (counted in the compiled code: 32 at orders 6 and 7, 8 elsewhere)
__host__ __device__ constexpr int enChunk(int p) { return EN_CHUNK >= 0 ? EN_CHUNK : (p == 6 || p == 7) ? 32 : 8; }
// MODE 0: the energy's kernels G0 - Gk and eps Gk - G0 (src/treecode.c:581-615);
// MODE 1: the right-hand side's Coulomb kernel 1/r (:283-379), both halves;
// MODE 3: d/dkappa of MODE 0's kernels; MODE 4: MODE 0 into (y0, y1) and
// MODE 3 into (z0, z1) from one pass over the coefficients
// R: the type of the coefficients and of the polynomial's arithmetic (float for
// the precision form's energy: half the scalar-load bytes per cluster; the
// radial values are formed in FP64 and rounded, the sums return to FP64)
template <int P, int MODE, typename R = double>
__device__ __forceinline__ void enFar(const R *__restrict__ co0, double rxd, double ryd, double rzd, double nxd,
                                      double nyd, double nzd, double kappa, double eps, double &y0, double &y1,
                                      double &z0, double &z1) {
  constexpr int NS = MODE == 4 ? 2 : 1;   // radial sequence pairs evaluated together
  const double rho = sqrt(rxd * rxd + ryd * ryd + rzd * rzd);
  const R rx = (R)rxd, ry = (R)ryd, rz = (R)rzd, nx = (R)nxd, ny = (R)nyd, nz = (R)nzd;
  double h[NS][P + 2], u[NS][P + 2];
  if constexpr (MODE == 1) {
    const double nr2 = -1.0 / (rho * rho);
    h[0][0] = 1.0 / rho;
#pragma unroll
    for (int k = 0; k <= P; k++) h[0][k + 1] = (2 * k + 1) * nr2 * h[0][k];
#pragma unroll
    for (int k = 0; k <= P + 1; k++) u[0][k] = h[0][k];
  } else {
    double g0[P + 2], gk[P + 2];
    radial(kappa, rho, P + 1, g0, 1, gk, 1);
    if constexpr (MODE == 0 || MODE == 4) {
#pragma unroll
      for (int k = 0; k <= P + 1; k++) {
        h[0][k] = g0[k] - gk[k];
        u[0][k] = eps * gk[k] - g0[k];
      }
    }
    if constexpr (MODE == 3 || MODE == 4) {   // d f_k / dkappa = kappa f_(k-1) (d/dkappa commutes with
      constexpr int t = NS - 1;               // (1/r) d/dr), d f_0 / dkappa = -e^{-kappa r} / (4 pi) for
      h[t][0] = rho * gk[0];                  // f = Gk; G0 is free of kappa
      u[t][0] = -eps * rho * gk[0];
#pragma unroll
      for (int k = 1; k <= P + 1; k++) {
        h[t][k] = -kappa * gk[k - 1];
        u[t][k] = eps * kappa * gk[k - 1];
      }
    }
  }
  // nested Horner in x (outer), y, z (inner); a coefficient of r^g is the
  // j-sum over the radial values of degree |g| + j.  The coefficients are
  // wave-uniform and come in through scalar loads in chunks of (a, b) rows of
  // at least enChunk(P) doubles: after a chunk an empty asm takes the table
  // pointer and the running sums, so the next chunk's loads cannot be issued
  // before this chunk is consumed (hoisted all at once, they overflow the
  // scalar registers and are spilled).
  using CD = const __attribute__((address_space(4))) R;
  CD *co = (CD *)co0;
  constexpr int CHUNK = enChunk(P) * (int)(sizeof(double) / sizeof(R));   // the same bytes per chunk
  R hr[NS][P + 2], ur[NS][P + 2];
#pragma unroll
  for (int t = 0; t < NS; t++)
#pragma unroll
    for (int k = 0; k <= P + 1; k++) {
      hr[t][k] = (R)h[t][k];
      ur[t][k] = (R)u[t][k];
    }
  int o = 0, since = 0;   // since: coefficients read since the last chunk boundary
  R s0[NS], s1[NS];
#pragma unroll
  for (int t = 0; t < NS; t++) s0[t] = s1[t] = 0;
#pragma unroll
  for (int a = P; a >= 0; a--) {
    R qa[NS];
#pragma unroll
    for (int t = 0; t < NS; t++) qa[t] = 0;
#pragma unroll
    for (int b = P - a; b >= 0; b--) {
      R rb[NS];
#pragma unroll
      for (int t = 0; t < NS; t++) rb[t] = 0;
#pragma unroll
      for (int c = P - a - b; c >= 0; c--) {
        const int d = a + b + c;
#pragma unroll
        for (int t = 0; t < NS; t++) {
          R cw = 0;
#pragma unroll
          for (int j = 0; j <= (P - d) / 2; j++) cw += hr[t][d + j] * co[o + j];
          rb[t] = rb[t] * rz + cw;
        }
        o += (P - d) / 2 + 1;
        since += (P - d) / 2 + 1;
      }
#pragma unroll
      for (int t = 0; t < NS; t++) qa[t] = qa[t] * ry + rb[t];
      if (CHUNK > 0 && since >= CHUNK) {   // folds at compile time
        if constexpr (NS == 1)
          asm volatile("" : "+s"(co), "+v"(qa[0]));
        else
          asm volatile("" : "+s"(co), "+v"(qa[0]), "+v"(qa[1]));
        since = 0;
      }
    }
#pragma unroll
    for (int t = 0; t < NS; t++) s0[t] = s0[t] * rx + qa[t];
  }
#pragma unroll
  for (int a = P + 1; a >= 0; a--) {
    R qa[NS];
#pragma unroll
    for (int t = 0; t < NS; t++) qa[t] = 0;
#pragma unroll
    for (int b = P + 1 - a; b >= 0; b--) {
      R rb[NS];
#pragma unroll
      for (int t = 0; t < NS; t++) rb[t] = 0;
#pragma unroll
      for (int c = P + 1 - a - b; c >= 0; c--) {
        const int d = a + b + c;
#pragma unroll
        for (int t = 0; t < NS; t++) {
          R cv = 0;
#pragma unroll
          for (int j = 0; j <= (P + 1 - d) / 2; j++)
            cv += ur[t][d + j] * (nx * co[o + 3 * j] + ny * co[o + 3 * j + 1] + nz * co[o + 3 * j + 2]);
          rb[t] = rb[t] * rz + cv;
        }
        o += 3 * ((P + 1 - d) / 2 + 1);
        since += 3 * ((P + 1 - d) / 2 + 1);
      }
#pragma unroll
      for (int t = 0; t < NS; t++) qa[t] = qa[t] * ry + rb[t];
      if (CHUNK > 0 && since >= CHUNK) {   // folds at compile time
        if constexpr (NS == 1)
          asm volatile("" : "+s"(co), "+v"(qa[0]));
        else
          asm volatile("" : "+s"(co), "+v"(qa[0]), "+v"(qa[1]));
        since = 0;
      }
    }
#pragma unroll
    for (int t = 0; t < NS; t++) s1[t] = s1[t] * rx + qa[t];
  }
  y0 += (double)s0[0];
  y1 += (double)s1[0];
  if constexpr (NS == 2) {
    z0 += (double)s0[1];
    z1 += (double)s1[1];
  }
}

// Panel-tree energy (src/treecode.c:617-695, 581-615; MODE 0), its weight
// vector dE/dsgm (MODE 2), its kappa derivative at fixed sgm (MODE 3; MODE 4:
// MODE 2 into part and MODE 3 into part2 in one walk) and the charge-tree
// right-hand side (src/treecode.c:381-450; MODE 1) with one lane per panel:
// the 64 panels of a wave (consecutive in leaf order, so close together) walk
// the charge tree together, depth first with kids in order, carrying the mask
// of lanes still descending.  A lane's contributions arrive in the order of its
// own walk.  Clusters at levels >= height with eRad < theta*dist are accepted;
// unaccepted leaves are summed directly.  Cluster data are wave-uniform, so the
// coefficient tables come in through the scalar cache.
#define EN_WSTACK 256
#ifndef EN_WPB
#define EN_WPB 4   // waves per block of the wave walk
#endif
#ifndef EN_TARGET_WAVES
#define EN_TARGET_WAVES 8192   // slices are added until the walk has this many waves
#endif
// waves per SIMD the walk's registers must allow: 4 for the energy's modes (a
// few spilled registers; 2.5 against 2.8 ms with three waves at 317k panels),
// the compiler's choice for the right-hand side (five waves) and MODE 4
__host__ __device__ constexpr int enMinWaves(int mode) { return mode == 0 || mode == 2 || mode == 3 ? 4 : 1; }
template <int P, int MODE, typename R = double>
__global__ __launch_bounds__(64 * EN_WPB, enMinWaves(MODE)) void k_energy_wave(const Pnl *__restrict__ pnl, int n, const double *__restrict__ sgm,
                                                     const int *__restrict__ level, const double *__restrict__ cx,
                                                     const double *__restrict__ eRad,
                                                     const int *__restrict__ kidStart, const int *__restrict__ nKids,
                                                     const int *__restrict__ chgStart, const int *__restrict__ nChgs,
                                                     const int *__restrict__ chgIdx, const double *__restrict__ pos,
                                                     const double *__restrict__ chr, const R *__restrict__ poly,
                                                     int c0, int depth, int height, double theta, double kappa,
                                                     double eps, QTab Q, int q, int nH, int S,
                                                     double *__restrict__ part, double *__restrict__ part2) {
  __shared__ int stC[EN_WPB][EN_WSTACK];
  __shared__ unsigned long long stM[EN_WPB][EN_WSTACK];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // S waves per 64 panels: slice s walks the clusters c0 + s, c0 + s + S, ... of
  // level `height` (no cluster above it can be accepted, so every lane reaches
  // all of them) and writes its own slice of part.  The EN_WPB waves of a block
  // take one slice for EN_WPB consecutive panel groups: they walk nearly the
  // same clusters in the same order, so the coefficient tables one wave brings
  // into the CU's scalar cache serve the others.  Blocks are numbered so that
  // each XCD (blocks b, b + 8, ... in dispatch order) gets a contiguous range.
  const int nb = gridDim.x, xb = blockIdx.x, per = nb / 8, rem = nb % 8, xcd = xb % 8;
  const int b = xcd < rem ? xcd * (per + 1) + xb / 8 : rem * (per + 1) + (xcd - rem) * per + xb / 8;
  const int sl = b % S;
  const int p = ((b / S) * EN_WPB + w) * 64 + lane;
  const bool live = p < n;
  const unsigned long long all = __ballot(live);
  if (all == 0ull) return;   // whole wave past the end (uniform)
  const int pp = live ? p : n - 1;
  double v0[3], e2[3], e0[3], nr[3];
  for (int k = 0; k < 3; k++) {
    v0[k] = pnl[pp].v[k];
    e2[k] = pnl[pp].a[6 + k];
    e0[k] = pnl[pp].a[k];
    nr[k] = pnl[pp].nr[k];
  }
  const double area = pnl[pp].area;
  constexpr int SZ = enPolySize(P);
  double o0 = 0.0, o1 = 0.0, o2 = 0.0, o3 = 0.0;
  for (int ix = 0; ix < q; ix++)
    for (int jx = 0; jx < q; jx++) {
      double qp[3];
      for (int k = 0; k < 3; k++) qp[k] = v0[k] + Q.t[ix] * (e2[k] + Q.t[jx] * e0[k]);
      double y0 = 0.0, y1 = 0.0, z0 = 0.0, z1 = 0.0;
      int sp = 0;
      for (int k = sl + ((nH - 1 - sl) / S) * S; k >= sl; k -= S, sp++) {   // pushed in reverse: visited in order
        stC[w][sp] = c0 + k;
        stM[w][sp] = all;
      }
      while (sp > 0) {
        sp--;
        const int cb = __builtin_amdgcn_readfirstlane(stC[w][sp]);
        const unsigned long long mraw = stM[w][sp];
        const unsigned long long m =
            ((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(mraw >> 32)) << 32) |
            (unsigned)__builtin_amdgcn_readfirstlane((int)(mraw & 0xffffffffu));
        const bool on = (m >> lane) & 1ull;
        const double rx = qp[0] - cx[3 * cb], ry = qp[1] - cx[3 * cb + 1], rz = qp[2] - cx[3 * cb + 2];
        const double dist = sqrt(rx * rx + ry * ry + rz * rz);
        const int lev = level[cb];
        const bool acc = on && eRad[cb] < theta * dist && lev >= height;
        const bool dir = on && !acc && lev == depth;
        const unsigned long long down = __ballot(on && !acc && lev < depth);
        if (acc)
          enFar<P, MODE == 2 ? 0 : MODE, R>(poly + (long long)(cb - c0) * SZ, rx, ry, rz, nr[0], nr[1], nr[2], kappa, eps,
                                         y0, y1, z0, z1);
        if (__ballot(dir) != 0ull) {
          const int s = chgStart[cb], nq = nChgs[cb];
          for (int k = 0; k < nq; k++) {
            const int j = chgIdx[s + k];
            if (dir) {
              if constexpr (MODE == 3) {
                enDirectDk(qp, nr, pos + 3 * j, chr[j], kappa, eps, y0, y1);
              } else if constexpr (MODE == 4) {
                enDirect(qp, nr, pos + 3 * j, chr[j], kappa, eps, y0, y1);
                enDirectDk(qp, nr, pos + 3 * j, chr[j], kappa, eps, z0, z1);
              } else if constexpr (MODE != 1) {
                enDirect(qp, nr, pos + 3 * j, chr[j], kappa, eps, y0, y1);
              } else {
                rhsDirect(qp, nr, pos + 3 * j, chr[j], y0, y1);
              }
            }
          }
        }
        if (down != 0ull) {
          const int nk = nKids[cb], k0 = kidStart[cb];
          if (sp + nk > EN_WSTACK) {   // stack bound: poison, never fault
            y0 = y1 = z0 = z1 = __builtin_nan("");
            break;
          }
          for (int k = nk - 1; k >= 0; k--) {   // pushed in reverse: kids visited in order
            stC[w][sp] = k0 + k;
            stM[w][sp] = down;
            sp++;
          }
        }
      }
      o0 += y0 * Q.t[ix] * Q.w[ix] * Q.w[jx];
      o1 += y1 * Q.t[ix] * Q.w[ix] * Q.w[jx];
      if constexpr (MODE == 4) {
        o2 += z0 * Q.t[ix] * Q.w[ix] * Q.w[jx];
        o3 += z1 * Q.t[ix] * Q.w[ix] * Q.w[jx];
      }
    }
  o0 *= 2.0 * area;
  o1 *= 2.0 * area;
  o2 *= 2.0 * area;
  o3 *= 2.0 * area;
  if (live) {
    if constexpr (MODE == 0 || MODE == 3) {   // 3: dE/dkappa at fixed sgm
      part[(long long)sl * n + p] = o0 * sgm[p + n] + o1 * sgm[p];
    } else if constexpr (MODE == 2 || MODE == 4) {   // part is the energy's weight vector g[2n]: E = g . [sgm | sgm_n]
      part[(long long)sl * 2 * n + p] = o1;
      part[(long long)sl * 2 * n + p + n] = o0;
      if constexpr (MODE == 4) part2[(long long)sl * n + p] = o2 * sgm[p + n] + o3 * sgm[p];   // and MODE 3's terms
    } else {   // part is the right-hand side b[2n], fac = 1/(4 pi eps1) passed as eps
      part[(long long)sl * 2 * n + p] = eps * o0;
      part[(long long)sl * 2 * n + p + n] = eps * o1;
    }
  }
}

// ------------------------- transposed panel-tree walk (exact dE/dq of the tree energy)
//
// The panel-tree energy (and the charge-tree right-hand side) is linear in
// the charges at fixed densities: E = sum_j q_j phi_j.  A far cluster A with
// moments M_a = sum_{j in A} q_j (x_j - c_A)^a / a! enters a panel point's sum
// as sum_a M_a K_a(r), r = x - c_A, with
//   K_a = (-1)^|a| [ wh d^a h(r) + wu d^a (n . grad u)(r) ],
// h and u the two radial kernels (G0 - Gk and eps Gk - G0 for the energy,
// 1/r twice for the right-hand side) and wh, wu the point's weights
// (2A t w w times sgm_n, sgm; or fac 2A t w w times lam, lam_n).  The closed
// form of the Cartesian derivatives of a radial function (k_energy_wave) gives
// each d^a f, factorised axis by axis (adjFar).  So phi_j = sum over the clusters A above
// and at j's leaf of sum_a Lam_A[a] (x_j - c_A)^a / a!, with
// Lam_A[a] = sum over the points accepting A of K_a, plus the direct terms of
// the points that reached j's leaf without accepting it.  k_adj_walk walks
// exactly as k_energy_wave (same acceptance test, same order) and stores each
// accepting wave's K (summed over its lanes in a fixed order) and each direct
// wave's per-charge sums as events keyed by the cluster.  A stable sort by
// cluster then lets one block per cluster add its events in the walk's order:
// no atomics, and the same sums in the same order on every run.  (Adding the
// events straight into Lam_A with atomics serialises on the top clusters,
// which every wave accepts.)

// k! (a loop, not a recursion: with constant k it folds to a constant)
__host__ __device__ constexpr double cfactc(int k) {
  double f = 1.0;
  for (int i = 2; i <= k; i++) f *= i;
  return f;
}
__host__ __device__ constexpr int gidx3c(int i1, int i2, int i3) {
  return (i1 + i2 + i3 > 0 ? (i1 + i2 + i3) * (i1 + i2 + i3 + 1) * (i1 + i2 + i3 + 2) / 6 : 0) +
         i1 * (i1 + i2 + i3 + 1) - i1 * (i1 - 1) / 2 + i2;
}
#define ADJ_LAYER(P) (((P) + 1) * ((P) + 2) / 2)
// a! / ((a - 2b)! b! 2^b): the one-axis factor of the closed form
__host__ __device__ constexpr double hcoef(int a, int b) {
  return cfactc(a) / (cfactc(a - 2 * b) * cfactc(b) * (double)(1 << b));
}
template <int A, int B>
struct HCoef {
  static constexpr double v = hcoef(A, B);
};

// f(integral_constant<int, 0>), ..., f(integral_constant<int, N - 1>): a loop
// unrolled by construction, its index a constant expression in the body
template <class F, int... I>
__device__ __forceinline__ void sforImpl(F &f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void sfor(F &&f) {
  if constexpr (N > 0) sforImpl(f, std::make_integer_sequence<int, N>{});
}

// One accepting event: every lane forms K_a of its point (zero weights off the
// event) for all |a| <= P.  With n . grad commuting with d^a,
//   K_a = (-1)^|a| [ d^a (wh h) + sum_i n_i d^(a+e_i) (wu u) ],
// and the closed form factorises axis by axis: with F_n[a3] = sum_b3 c(a3,b3)
// z^(a3-2b3) f_(n+a3-b3), E_m[a2,a3] = sum_b2 c(a2,b2) y^(a2-2b2) F_(m+a2-b2)[a3]
// and d^a f = sum_b1 c(a1,b1) x^(a1-2b1) E_(a1-b1)[a2,a3] (c = hcoef), so each
// slab a3 needs two F rows (h with the z-neighbour of u folded in, and u) and
// each (a2, a3) two E rows (Q: h plus the y- and z-neighbours of u; Eu: u for
// the x-neighbour).  Every loop is unrolled at compile time (sfor) and every
// coefficient is a constant.  A slab's (P - a3 + 1)(P - a3 + 2)/2 values (a2,
// then a1, ascending) pass through LDS (one row of 64 lanes per value, rows 65
// apart); the 64 lanes are summed in a fixed order (ng groups of 64 / ng
// lanes, ng the largest power of two with ng m <= 64, one lane per group and
// value, then the groups' sums combined by lane exchange) into out, slab
// after slab.
template <int P, int KIND>
__device__ __forceinline__ void adjFar(double rx, double ry, double rz, double nx, double ny, double nz, double wh,
                                       double wu, double kappa, double eps, double *__restrict__ out, double *sl,
                                       int lane) {
  const double rho = sqrt(rx * rx + ry * ry + rz * rz);
  double fh[P + 1], fu[P + 2];   // the radial sequences of wh h and wu u
  if constexpr (KIND == 0) {
    double g0[P + 2], gk[P + 2];
    radial(kappa, rho, P + 1, g0, 1, gk, 1);
    sfor<P + 2>([&](auto K) {
      constexpr int k = decltype(K)::value;
      if constexpr (k <= P) fh[k] = wh * (g0[k] - gk[k]);
      fu[k] = wu * (eps * gk[k] - g0[k]);
    });
  } else {
    const double nr2 = -1.0 / (rho * rho);
    double c = 1.0 / rho;
    sfor<P + 2>([&](auto K) {
      constexpr int k = decltype(K)::value;
      if constexpr (k <= P) fh[k] = wh * c;
      fu[k] = wu * c;
      c = (2 * k + 1) * nr2 * c;
    });
  }
  double px[P + 2], py[P + 2], pz[P + 2];
  px[0] = py[0] = pz[0] = 1.0;
  sfor<P + 1>([&](auto K) {
    constexpr int k = decltype(K)::value + 1;
    px[k] = px[k - 1] * rx;
    py[k] = py[k - 1] * ry;
    pz[k] = pz[k - 1] * rz;
  });
  int off = 0;
  sfor<P + 1>([&](auto A3) {
    constexpr int a3 = decltype(A3)::value;
    double FhZ[P + 1 - a3], Fu[P + 2 - a3];
    sfor<P + 1 - a3>([&](auto N) {
      constexpr int n = decltype(N)::value;
      double s = 0.0, t = 0.0;
      sfor<a3 / 2 + 1>([&](auto B) {
        constexpr int b = decltype(B)::value;
        s = fma(HCoef<a3, b>::v * pz[a3 - 2 * b], fh[n + a3 - b], s);
      });
      sfor<(a3 + 1) / 2 + 1>([&](auto B) {
        constexpr int b = decltype(B)::value;
        t = fma(HCoef<a3 + 1, b>::v * pz[a3 + 1 - 2 * b], fu[n + a3 + 1 - b], t);
      });
      FhZ[n] = fma(nz, t, s);
    });
    sfor<P + 2 - a3>([&](auto N) {
      constexpr int n = decltype(N)::value;
      double s = 0.0;
      sfor<a3 / 2 + 1>([&](auto B) {
        constexpr int b = decltype(B)::value;
        s = fma(HCoef<a3, b>::v * pz[a3 - 2 * b], fu[n + a3 - b], s);
      });
      Fu[n] = s;
    });
    sfor<P + 1 - a3>([&](auto A2) {
      constexpr int a2 = decltype(A2)::value;
      constexpr int k0 = a2 * (P - a3 + 1) - a2 * (a2 - 1) / 2;   // the slab index of (0, a2)
      double Q[P + 1 - a2 - a3], Eu[P + 2 - a2 - a3];
      sfor<P + 1 - a2 - a3>([&](auto M) {
        constexpr int m = decltype(M)::value;
        double s = 0.0, t = 0.0;
        sfor<a2 / 2 + 1>([&](auto B) {
          constexpr int b = decltype(B)::value;
          s = fma(HCoef<a2, b>::v * py[a2 - 2 * b], FhZ[m + a2 - b], s);
        });
        sfor<(a2 + 1) / 2 + 1>([&](auto B) {
          constexpr int b = decltype(B)::value;
          t = fma(HCoef<a2 + 1, b>::v * py[a2 + 1 - 2 * b], Fu[m + a2 + 1 - b], t);
        });
        Q[m] = fma(ny, t, s);
      });
      sfor<P + 2 - a2 - a3>([&](auto M) {
        constexpr int m = decltype(M)::value;
        double s = 0.0;
        sfor<a2 / 2 + 1>([&](auto B) {
          constexpr int b = decltype(B)::value;
          s = fma(HCoef<a2, b>::v * py[a2 - 2 * b], Fu[m + a2 - b], s);
        });
        Eu[m] = s;
      });
      sfor<P + 1 - a2 - a3>([&](auto A1) {
        constexpr int a1 = decltype(A1)::value;
        double s = 0.0, t = 0.0;
        sfor<a1 / 2 + 1>([&](auto B) {
          constexpr int b = decltype(B)::value;
          s = fma(HCoef<a1, b>::v * px[a1 - 2 * b], Q[a1 - b], s);
        });
        sfor<(a1 + 1) / 2 + 1>([&](auto B) {
          constexpr int b = decltype(B)::value;
          t = fma(HCoef<a1 + 1, b>::v * px[a1 + 1 - 2 * b], Eu[a1 + 1 - b], t);
        });
        const double v = fma(nx, t, s);
        sl[(k0 + a1) * 65 + lane] = ((a1 + a2 + a3) & 1) ? -v : v;
      });
    });
    waveSync();
    constexpr int m = (P - a3 + 1) * (P - a3 + 2) / 2;
    constexpr int W = m <= 1 ? 1 : m <= 2 ? 2 : m <= 4 ? 4 : m <= 8 ? 8 : m <= 16 ? 16 : m <= 32 ? 32 : 64;
    const int a = lane & (W - 1), r0 = lane & ~(W - 1);   // this lane's value and first lane to add
    double s = 0.0;
    if (a < m) {
      const double *row = sl + a * 65 + r0;
#pragma unroll 8
      for (int l = 0; l < W; l++) s += row[l];
    }
    sfor<6>([&](auto O) {
      constexpr int o = 1 << decltype(O)::value;
      if constexpr (o >= W) s += __shfl_xor(s, o, 64);
    });
    if (lane < m) out[off + lane] = s;
    off += m;
    waveSync();
  });
}

// the multi-index of value t of an event (slabs a3 = 0..p, then a2, a1
// ascending) as its graded index (gidx3c)
__device__ inline int slabGidx(int p, int t) {
  for (int a3 = 0; a3 <= p; a3++)
    for (int a2 = 0; a2 <= p - a3; a2++) {
      const int len = p - a3 - a2 + 1;
      if (t < len) return gidx3c(t, a2, a3);
      t -= len;
    }
  return 0;
}

// one direct charge: the kernels' values at x = point - charge (k0: G0 - Gk or
// 1/r; k1: eps dGk/dn - dG0/dn or -n.x/r^3, n the point's normal)
template <int KIND>
__device__ __forceinline__ double adjDirect(const double *qp, const double *nr, const double *pj, double wh,
                                            double wu, double kappa, double eps) {
  const double x[3] = {qp[0] - pj[0], qp[1] - pj[1], qp[2] - pj[2]};
  const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
  const double ip = nr[0] * x[0] + nr[1] * x[1] + nr[2] * x[2];
  if constexpr (KIND == 0) {
    const double rn = sqrt(r2), G0 = kFourPiI / rn, ek = exp(-kappa * rn), Gk = ek * G0;
    const double coef = (kappa * rn + 1.0) * ek, dG0dn = -ip * G0 / r2, dGkdn = coef * dG0dn;
    return wh * (G0 - Gk) + wu * (eps * dGkdn - dG0dn);
  } else {
    const double ri = 1.0 / sqrt(r2), r3i = ri / r2;
    return wh * ri + wu * (-ip * r3i);
  }
}

// The events of a transposed walk: an accepting event (a wave's lanes that
// accept one cluster at one quadrature point) carries the NMOM(P) sums over
// those lanes of K_a; a direct event (the lanes that reach a leaf without
// accepting it) carries one sum per charge of the leaf.  Per block of the walk:
// cnt (3 values: accepting events, direct events, direct values) in the count
// pass; base (3 values per block and one past the last, prefix sums of cnt)
// places its events in the emit pass, minus the batch's first offsets bA, bD, bV.
struct AdjEvents {
  long long *cnt;
  const long long *base;
  long long bA, bD, bV;
  int *accKey, *accIdx, *dirKey, *dirIdx, *dirOff, *err;
  double *accVals, *dirVals;
};

// The transposed walk (see above): KIND 0 the panel-tree energy with weights
// 2A t w w (sgm_n, sgm); KIND 1 the charge-tree right-hand side with weights
// fac 2A t w w (lam, lam_n).  Walks exactly as k_energy_wave (same acceptance
// test, same order); EMIT false counts the events of each block, EMIT true
// writes them (keys = the cluster, values in the walk's order), blocks g0 ...
template <int P, int KIND, bool EMIT>
__global__ __launch_bounds__(64) void k_adj_walk(const Pnl *__restrict__ pnl, int n, const double *__restrict__ w,
                                                 const int *__restrict__ level, const double *__restrict__ cx,
                                                 const double *__restrict__ eRad,
                                                 const int *__restrict__ kidStart, const int *__restrict__ nKids,
                                                 const int *__restrict__ chgStart, const int *__restrict__ nChgs,
                                                 const int *__restrict__ chgIdx, const double *__restrict__ pos,
                                                 int c0, int depth, int height, double theta, double kappa,
                                                 double eps, double fac, QTab Q, int q, int nH, int S, int g0,
                                                 AdjEvents ev) {
  __shared__ int stC[EN_WSTACK];
  __shared__ unsigned long long stM[EN_WSTACK];
  __shared__ double sl[EMIT ? 65 * ADJ_LAYER(P) : 1];   // value-major rows, padded against bank conflicts
  __shared__ double pt[8 * 64];   // per lane: the quadrature point, its normal and its two weights
  const int lane = threadIdx.x;
  const int g = g0 + blockIdx.x, sl_ = g % S;
  const int p = (g / S) * 64 + lane;
  const bool live = p < n;
  const unsigned long long all = __ballot(live);
  long long nA = 0, nD = 0, nV = 0;   // wave-uniform event counts
  long long eA = 0, eD = 0, eV = 0, endA = 0, endD = 0, endV = 0;
  if constexpr (EMIT) {
    eA = ev.base[3 * g] - ev.bA;
    eD = ev.base[3 * g + 1] - ev.bD;
    eV = ev.base[3 * g + 2] - ev.bV;
    endA = ev.base[3 * g + 3] - ev.bA;
    endD = ev.base[3 * g + 4] - ev.bD;
    endV = ev.base[3 * g + 5] - ev.bV;
  }
  bool bad = false;
  if (all != 0ull) {
    const int pp = live ? p : n - 1;
    for (int ix = 0; ix < q && !bad; ix++)
      for (int jx = 0; jx < q && !bad; jx++) {
        // this quadrature point of each lane, staged in LDS (the walk reads it
        // back per event, so it holds no registers across the walk)
        {
          double qp[3];
          for (int k = 0; k < 3; k++) qp[k] = pnl[pp].v[k] + Q.t[ix] * (pnl[pp].a[6 + k] + Q.t[jx] * pnl[pp].a[k]);
          // the point weights: energy (sgm_n, sgm), right-hand side fac (lam, lam_n)
          const double wq = 2.0 * pnl[pp].area * Q.t[ix] * Q.w[ix] * Q.w[jx];
          double wa = 0.0, wb = 0.0;
          if constexpr (EMIT) {
            wa = live ? (KIND == 0 ? w[pp + n] : fac * w[pp]) : 0.0;
            wb = live ? (KIND == 0 ? w[pp] : fac * w[pp + n]) : 0.0;
          }
          waveSync();
          for (int k = 0; k < 3; k++) {
            pt[k * 64 + lane] = qp[k];
            pt[(3 + k) * 64 + lane] = pnl[pp].nr[k];
          }
          pt[6 * 64 + lane] = wq * wa;
          pt[7 * 64 + lane] = wq * wb;
          waveSync();
        }
        int sp = 0;
        for (int k = sl_ + ((nH - 1 - sl_) / S) * S; k >= sl_; k -= S, sp++) {   // pushed in reverse: visited in order
          stC[sp] = c0 + k;
          stM[sp] = all;
        }
        while (sp > 0) {
          sp--;
          const int cb = __builtin_amdgcn_readfirstlane(stC[sp]);
          const unsigned long long mraw = stM[sp];
          const unsigned long long m =
              ((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(mraw >> 32)) << 32) |
              (unsigned)__builtin_amdgcn_readfirstlane((int)(mraw & 0xffffffffu));
          const bool on = (m >> lane) & 1ull;
          const double rx = pt[lane] - cx[3 * cb], ry = pt[64 + lane] - cx[3 * cb + 1],
                       rz = pt[128 + lane] - cx[3 * cb + 2];
          const double dist = sqrt(rx * rx + ry * ry + rz * rz);
          const int lev = level[cb];
          const bool acc = on && eRad[cb] < theta * dist && lev >= height;
          const bool dir = on && !acc && lev == depth;
          const unsigned long long down = __ballot(on && !acc && lev < depth);
          if (__ballot(acc) != 0ull) {   // every lane takes part; those off the event add zeros
            if constexpr (EMIT) {
              if (eA + nA >= endA) { bad = true; break; }
              const double ox = acc ? rx : 1.0, oy = acc ? ry : 0.0, oz = acc ? rz : 0.0;
              adjFar<P, KIND>(ox, oy, oz, pt[192 + lane], pt[256 + lane], pt[320 + lane], acc ? pt[384 + lane] : 0.0,
                              acc ? pt[448 + lane] : 0.0, kappa, eps, ev.accVals + (eA + nA) * NMOM(P), sl, lane);
              if (lane == 0) {
                ev.accKey[eA + nA] = cb;
                ev.accIdx[eA + nA] = (int)(eA + nA);
              }
            }
            nA++;
          }
          if (__ballot(dir) != 0ull) {
            const int s = chgStart[cb], nq = nChgs[cb];
            if constexpr (EMIT) {
              if (eD + nD >= endD || eV + nV + nq > endV) { bad = true; break; }
              const double qp[3] = {pt[lane], pt[64 + lane], pt[128 + lane]};
              const double nr[3] = {pt[192 + lane], pt[256 + lane], pt[320 + lane]};
              const double wh = pt[384 + lane], wu = pt[448 + lane];
              for (int k = 0; k < nq; k++) {
                const int j = chgIdx[s + k];
                double v = dir ? adjDirect<KIND>(qp, nr, pos + 3 * j, wh, wu, kappa, eps) : 0.0;
                for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
                if (lane == 0) ev.dirVals[eV + nV + k] = v;
              }
              if (lane == 0) {
                ev.dirKey[eD + nD] = cb;
                ev.dirIdx[eD + nD] = (int)(eD + nD);
                ev.dirOff[eD + nD] = (int)(eV + nV);
              }
            }
            nD++;
            nV += nq;
          }
          if (down != 0ull) {
            const int nk = nKids[cb], k0 = kidStart[cb];
            if (sp + nk > EN_WSTACK) { bad = true; break; }   // stack bound (checked on the host): stop, never fault
            for (int k = nk - 1; k >= 0; k--) {
              stC[sp] = k0 + k;
              stM[sp] = down;
              sp++;
            }
          }
        }
      }
  }
  if constexpr (EMIT) {
    if (!bad && (eA + nA != endA || eD + nD != endD || eV + nV != endV)) bad = true;
  } else if (lane == 0) {
    ev.cnt[3 * g] = nA;
    ev.cnt[3 * g + 1] = nD;
    ev.cnt[3 * g + 2] = nV;
  }
  if (bad && lane == 0) ev.err[0] = 1;
}

// the run bounds of sorted keys: s0[k] .. s1[k] (the caller zeroes both)
__global__ void k_seg_bounds(const int *__restrict__ key, int n, int *__restrict__ s0, int *__restrict__ s1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = key[i];
  if (i == 0 || key[i - 1] != k) s0[k] = i;
  if (i == n - 1 || key[i + 1] != k) s1[k] = i + 1;
}

__global__ void k_fill_int(int *__restrict__ y, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = v;
}

// Lam_A += the sums of cluster A's accepting events, in event order (one block
// per cluster; ns strided partial sums of each multi-index, then added in
// stride order)
#define ADJ_RED_THREADS 1024
__global__ __launch_bounds__(ADJ_RED_THREADS) void k_adj_red_acc(const int *__restrict__ s0,
                                                                 const int *__restrict__ s1,
                                                                 const int *__restrict__ idx,
                                                                 const double *__restrict__ vals, int p, int nmMax,
                                                                 double *__restrict__ lam) {
  __shared__ double part[ADJ_RED_THREADS];
  const int nm = NMOM(p);
  const int A = blockIdx.x, i0 = s0[A], i1 = s1[A];
  if (i1 <= i0) return;
  const int ns = ADJ_RED_THREADS / nm, t = threadIdx.x, a = t % nm, st = t / nm;
  double v = 0.0;
  if (st < ns)
    for (int i = i0 + st; i < i1; i += ns) v += vals[(long long)idx[i] * nm + a];
  part[t] = v;
  __syncthreads();
  if (t < nm) {
    double s = part[t];
    for (int k = 1; k < ns; k++) s += part[k * nm + t];
    lam[(long long)A * nmMax + slabGidx(p, t)] += s;
  }
}

// phi_j += the sums of leaf A's direct events for its charges, in event order
__global__ __launch_bounds__(64) void k_adj_red_dir(int levStart, const int *__restrict__ s0,
                                                    const int *__restrict__ s1, const int *__restrict__ idx,
                                                    const int *__restrict__ off, const double *__restrict__ vals,
                                                    const int *__restrict__ chgStart, const int *__restrict__ nChgs,
                                                    const int *__restrict__ chgIdx, double *__restrict__ phi) {
  const int A = levStart + blockIdx.x, i0 = s0[A], i1 = s1[A];
  if (i1 <= i0) return;
  for (int k = threadIdx.x; k < nChgs[A]; k += blockDim.x) {
    double v = 0.0;
    for (int i = i0; i < i1; i++) v += vals[off[idx[i]] + k];
    phi[chgIdx[chgStart[A] + k]] += v;
  }
}

// dE/dq_j = phi_j + sum over the clusters A from j's leaf up to level height of
// sum_a Lam_A[a] (x_j - c_A)^a / a! (the moments' monomials); one thread per
// charge of a leaf, in the leaf's charge order
__global__ void k_adj_eval(int levStart, const int *__restrict__ chgStart, const int *__restrict__ nChgs,
                           const int *__restrict__ chgIdx, const int *__restrict__ parent,
                           const int *__restrict__ level, const double *__restrict__ cx,
                           const double *__restrict__ pos, const double *__restrict__ lam, const double *phi,
                           int height, int P, int nmMax, DevTabs T, double *__restrict__ out) {
  const int leaf = levStart + blockIdx.x, nm = NMOM(P);
  for (int k = threadIdx.x; k < nChgs[leaf]; k += blockDim.x) {
    const int j = chgIdx[chgStart[leaf] + k];
    double v = phi[j];
    for (int A = leaf; A >= 0 && level[A] >= height; A = parent[A]) {
      const double tr[3] = {pos[3 * j] - cx[3 * A], pos[3 * j + 1] - cx[3 * A + 1], pos[3 * j + 2] - cx[3 * A + 2]};
      const double *L = lam + (long long)A * nmMax;
      double s = 0.0;
      for (int a = 0; a < nm; a++) s = fma(L[a], conVect1(tr, T.mi[a], T.ifact3[a]), s);
      v += s;
    }
    out[j] = v;
  }
}

// out[i] = sum over the S slices of a (slice stride m), in slice order
__global__ void k_sum_slices(const double *__restrict__ a, int S, long long m, double *__restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double v = a[i];
  for (int s = 1; s < S; s++) v += a[(long long)s * m + i];
  out[i] = v;
}

// block partials of sum a[i] over a fixed grid (deterministic order)
__global__ __launch_bounds__(256) void k_sum_part(const double *__restrict__ a, int n, double *__restrict__ part) {
  __shared__ double s[256];
  double v = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) v += a[i];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

#pragma clang fp contract(fast)
// ------------------------------------------------- reductions and GMRES

// The reductions' grid is one block per CU, so a thread's grid-stride loop
// takes about ten trips at 300k panels; the loads of RED_U trips (those inside
// n) are issued before their arithmetic, one memory round trip per RED_U trips
// instead of one per trip, and each thread still adds its terms in index order.
#define RED_U 16

// block partial of sum a[i]*b[i] over a fixed grid (deterministic)
__global__ __launch_bounds__(256) void k_dot_part(const double *__restrict__ a, const double *__restrict__ b, int n,
                                                  double *__restrict__ part) {
  __shared__ double s[256];
  const int st = gridDim.x * 256;
  double v = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += RED_U * st) {
    double av[RED_U], bv[RED_U];
#pragma unroll
    for (int u = 0; u < RED_U; u++) {
      const bool in = i + u * st < n;
      av[u] = in ? a[i + u * st] : 0.0;
      bv[u] = in ? b[i + u * st] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < RED_U; u++)
      if (i + u * st < n) v += av[u] * bv[u];
  }
  s[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__device__ double reduceParts(const double *part, int np) {   // same order in every block
  __shared__ double s[256];
  s[threadIdx.x] = threadIdx.x < np ? part[threadIdx.x] : 0.0;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  double v = s[0];
  __syncthreads();
  return v;
}

// The diagonal preconditioner (k_psolve, in place) fused with the first MGS
// dot product of an Arnoldi step: w[i] = w[i] / (s area), then the block
// partials of <w, v> over the reductions' grid.  The same expressions and
// reduction as k_psolve followed by k_dot_part, so the same bits in one pass.
__global__ __launch_bounds__(256) void k_psolve_dot(double *__restrict__ w, const double *__restrict__ v,
                                                    const double *__restrict__ area, int np, double s1, double s2,
                                                    double *__restrict__ part) {
  __shared__ double s[256];
  const int n = 2 * np, st = gridDim.x * 256;
  double acc = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += RED_U * st) {
    double wv[RED_U], vv[RED_U], av[RED_U];
#pragma unroll
    for (int u = 0; u < RED_U; u++) {
      const int j = i + u * st;
      const bool in = j < n;
      wv[u] = in ? w[j] : 0.0;
      vv[u] = in ? v[j] : 0.0;
      av[u] = in ? area[j < np ? j : j - np] : 1.0;
    }
#pragma unroll
    for (int u = 0; u < RED_U; u++) {
      const int j = i + u * st;
      if (j < n) {
        const double wp = wv[u] / ((j < np ? s1 : s2) * av[u]);
        w[j] = wp;
        acc += wp * vv[u];
      }
    }
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// MGS step (src/gmres.c:606-609) fused with the next step's dot product:
// h = <w, v> from the partials of the previous launch, w -= h v, then the block
// partials of <w, vnext> (vnext = w for the closing norm).  Element order, FMA
// use and the reduction tree are those of an update pass followed by
// k_dot_part, so the results are the same bit for bit with half the launches.
__global__ __launch_bounds__(256) void k_mgs_dot(double *__restrict__ w, const double *__restrict__ v,
                                                 const double *vnext, int n, const double *__restrict__ partIn,
                                                 double *__restrict__ h, double *__restrict__ partOut) {
  __shared__ double s[256];
  const double hk = reduceParts(partIn, RED_BLOCKS);
  if (blockIdx.x == 0 && threadIdx.x == 0) *h = hk;
  const bool self = vnext == w;
  const int st = gridDim.x * 256;
  double acc = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += RED_U * st) {
    double wv[RED_U], vv[RED_U], nv[RED_U];
#pragma unroll
    for (int u = 0; u < RED_U; u++) {   // every load of the batch before its stores
      const bool in = i + u * st < n;
      wv[u] = in ? w[i + u * st] : 0.0;
      vv[u] = in ? v[i + u * st] : 0.0;
      nv[u] = in && !self ? vnext[i + u * st] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < RED_U; u++)
      if (i + u * st < n) {
        const double wi = wv[u] + -hk * vv[u];
        w[i + u * st] = wi;
        acc += wi * (self ? wi : nv[u]);
      }
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partOut[blockIdx.x] = s[0];
}

// h = ||w||; v = w / h if h > 0 (src/gmres.c:612-620)
__global__ __launch_bounds__(256) void k_normalize(const double *__restrict__ w, double *__restrict__ v, int n,
                                                   const double *__restrict__ part, double *__restrict__ h) {
  double hh = sqrt(reduceParts(part, RED_BLOCKS));
  if (blockIdx.x == 0 && threadIdx.x == 0) *h = hh;
  double sc = hh > 0.0 ? 1.0 / hh : 1.0;
  const int st = gridDim.x * 256;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += RED_U * st) {
    double wv[RED_U];
#pragma unroll
    for (int u = 0; u < RED_U; u++) wv[u] = i + u * st < n ? w[i + u * st] : 0.0;
#pragma unroll
    for (int u = 0; u < RED_U; u++)
      if (i + u * st < n) v[i + u * st] = hh > 0.0 ? wv[u] * sc : wv[u];
  }
}

__global__ __launch_bounds__(256) void k_sum_parts(const double *__restrict__ part, int np, double *__restrict__ out) {
  if (blockIdx.x == 0) {
    double v = reduceParts(part, np);
    if (threadIdx.x == 0) *out = v;
  }
}

// x += sum_k y_k V_k (src/gmres.c:364-373, dgemv 'N')
__global__ void k_update(double *__restrict__ x, const double *__restrict__ V, const double *__restrict__ y, int m,
                         int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int k = 0; k < m; k++) s += y[k] * V[(long long)k * n + i];
  x[i] += s;
}

__global__ void k_psolve(const double *__restrict__ in, double *__restrict__ out, const double *__restrict__ area,
                         int n, double s1, double s2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * n) return;
  int p = i < n ? i : i - n;
  out[i] = in[i] / ((i < n ? s1 : s2) * area[p]);
}

__global__ void k_scale(const double *__restrict__ in, double *__restrict__ out, int n, double s) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i] * s;
}

// ================================================================ host side

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

// a charge set staged on the device (stageCharges): inputs and walk scratch in one arena
struct ChargeSet {
  void *arena = nullptr;
  size_t arenaBytes = 0;
  std::vector<unsigned char> stage;
  bool ready = false;
  // the cluster moments (and walk coefficients) of the staged charges are built:
  // they depend on the charges alone, so later walks of the same set reuse them
  mutable bool momReady = false;
  int nChar = 0, nc = 0, c0 = -1, P = 0, S = 1, SZ = 0, nClus = 0, nDesc = 0;
  bool uni = false;
  std::vector<int> lvStart, lvCount, ord;
  int *dLevel = nullptr, *dKid = nullptr, *dNK = nullptr, *dCS = nullptr, *dNC = nullptr, *dCI = nullptr,
      *dOrd = nullptr, *dPar = nullptr;
  double *dX = nullptr, *dR = nullptr, *dPos = nullptr, *dChr = nullptr, *dMom = nullptr, *dPart = nullptr,
         *dOut = nullptr, *dRed = nullptr, *dPoly = nullptr;
  float *dPoly32 = nullptr;   // the walk coefficients rounded to FP32 (the precision form's energy)
  const int4 *dDesc = nullptr;
};

struct fbh_ctx {
  int dev = 0;
  hipStream_t stream = nullptr;
  bool ownStream = false;
  hipStream_t auxStream = nullptr;   // the tree gradient's second stream (shared per device, on first use)
  double kappa, eps1, eps2, eps;
  double rhsFac = 0.0;   // right-hand side factor; 0: 1/(4 pi eps1), src/fabipb.c:755
  int qorder;
  QTab Q;
  int n, depth, height, maxOrder, P, nmMax, nmLeaf;
  int nCube, nLeaf;
  std::vector<int> levStart, levCount, ordMom, ordM2L, cubeLevel;
  std::vector<DevBuf> bufs;
  size_t bytes = 0;
  // device arrays
  Pnl *pnl;
  double *area, *cubeX, *Q0, *Q1, *mom, *lec, *far, *nearv, *xs, *ys, *fact;
  int *leafOf, *leafStart, *leafCount, *cubeLevelD, *kidStart, *nKids, *parent, *m2lOff, *m2lSrc, *ordMomD,
      *ordM2LD;
  long long nInter;
  int *srcOff, *ncol, *srcList;
  long long *tileCoef = nullptr, *dTileCoef = nullptr;   // per tile
  long long nStored = 0;       // interactions of the streamed tiles (the coefficient store)
  int nFlyTiles = 0;           // tiles computed on the fly
  double4 *coef;
  int2 *tiles;
  int nTiles;
  long long nLinks, nM2L;
  double *m2lG;   // per pair D^a G0, D^a Gk (2*nmMax)
  double *m2lGdk = nullptr;   // per pair D^a dGk/dkappa, D^a Gk (the gradient's kappa term; built on first use)
  int *m2lDst;
  DevTabs T;    // contraction table: ep = (a+e1, a+e2, a+e3, sign)
  DevTabs Tm;   // moment table: ep = (a-e1, a-e2, a-e3)
  int nConv = 0;   // entries of the (beta, alpha+beta) table
  std::vector<int> hConvStart, hM2mStart;   // host copies of the table offsets (LDS sizing)
  std::vector<int> hConv;
  // balanced M2L plan per order (cached-tensor path)
  unsigned *planTerms = nullptr;
  int *planIdx = nullptr;
  int4 *planOff = nullptr;
  size_t m2lLds = 0;
  int m2lWaves = 0;
  int m2lMaxT = 0;   // register-resident term lists (16, 32) or 0
  // FP64 mixed nearfield launch: tiles costliest first, grid-stride over them
  int *tileOrder = nullptr;
  int nearGrid = 0;
  int *ownOff = nullptr;   // per leaf: the column of its own first panel in its column list (-1: none)
  // M2L term grid (k_m2l_grid): K steps x 64 lanes per order, runs of L steps
  int gridK = 0, gridL = 0, gridSlotZ = 0, gridNcc = 0, gridNgg = 0, gridWaves = 4;
  unsigned *gridTerms = nullptr;
  int *m2lOrder = nullptr;     // cubes by interaction-list length, longest first (k_m2l_grid)
  std::vector<int> hM2lOff;    // host copy of m2lOff
  int *gridOff = nullptr, *gridRedStart = nullptr;
  unsigned short *gridRedCell = nullptr;
  size_t gridLds = 0, gridLds32 = 0;
  int ldsMax = 65536;
  // FP32 operator (precision form): copies of the cached operator in FP32
  bool fp32 = false;
  size_t near32Pad = 0;   // extra LDS per FP32 nearfield block (FABIPB_NEAR32_PAD_KB): fewer blocks per CU
  float *Q0f = nullptr, *Q1f = nullptr, *momf = nullptr, *lecf = nullptr, *farf = nullptr, *nearf = nullptr,
        *m2lGf = nullptr;
  float4 *coef32 = nullptr;   // streamed FP32 coefficients (q > 1)
  size_t m2lLds32 = 0;
  // q = 1: nearfield evaluated on the fly from FP32 panel data, plus the
  // stored corrections of the self / shared-vertex pairs (CSR per panel)
  bool fly = false;
  float4 *flyPA = nullptr, *flyPN = nullptr, *flyPL = nullptr, *specVal = nullptr;
  long long *specOff = nullptr;
  int *specCol = nullptr;
  long long nSpec = 0;
  // FP64 nearfield on the fly (q = 1): panel data and the corrections in FP64
  bool fly64 = false;
  double4 *fly64PA = nullptr, *fly64PN = nullptr, *specVal64 = nullptr, *specValT64 = nullptr;
  double *ro = nullptr, *dx = nullptr;   // refinement residual and correction
  ChargeSet csCall;   // charge set of the calls that pass host charge arrays
  ChargeSet csRes;    // resident charge set (fbh_set_charges)
  // preconditioner: 3 diagonal, 2 leaf-parent block LU (explicit inverses)
  int precMode = 3;
  std::vector<int> hKidStart, hNKids, hLeafStart, hLeafCount;
  int nPB = 0, nPBTiles = 0;
  int *pbStart = nullptr, *pbM = nullptr;
  long long *pbOff = nullptr;
  int2 *pbTiles = nullptr;
  double *pbInv = nullptr, *pbTmp = nullptr, *pbInvT = nullptr;
  size_t pbBytes = 0;
  // adjoint operator A^T: transposed near store, adjoint locals / moments
  bool nearSym = false, m2lSym = false;
  double4 *coefT = nullptr;
  double *adjL = nullptr, *adjM = nullptr;
  double4 *specDk = nullptr;   // d/dkappa corrections of the special pairs
  rocblas_handle blas = nullptr;
  // dense direct operator (-r=1/2): every panel pair through the nearfield store
  bool direct = false;
  int *dSrcOff = nullptr, *dNcol = nullptr, *dSrcList = nullptr;
  long long *dCoefOff = nullptr;
  double4 *dCoef = nullptr;
  // nearfield on a second stream, overlapped with the far-field chain
  hipStream_t side = nullptr;
  hipEvent_t evFork = nullptr, evJoin = nullptr;
  hipEvent_t evX = nullptr;   // the refined solve's x, ready for the energy walk beside its residual
  // nearfield launch timing (fbh_near_timing): an event pair around every
  // nearfield launch on the side stream, read back after the timed region
  bool ntOn = false;
  int ntUsed = 0;
  std::vector<hipEvent_t> ntEv;
  std::vector<unsigned char> ntKind;   // 0 FP64, 1 FP32
  // GMRES workspace
  int gmRestart = 0;
  double *V = nullptr, *w = nullptr, *r = nullptr, *part = nullptr, *hcol = nullptr, *ydev = nullptr, *scal = nullptr;
  double *hpin = nullptr;   // pinned host staging of the small readbacks (norms, Hessenberg columns, y)
  int hpinLen = 0;
  hipEvent_t evDone = nullptr;   // marks a readback; the host spins on it
};

template <class T>
static T *dalloc(fbh_ctx *c, size_t count) {
  DevBuf b;
  b.bytes = std::max<size_t>(count, 1) * sizeof(T);
  if (hipMalloc(&b.p, b.bytes) != hipSuccess) {
    setErr("hipMalloc of %zu bytes failed", b.bytes);
    return nullptr;
  }
  c->bufs.push_back(b);
  c->bytes += b.bytes;
  return (T *)b.p;
}

template <class T>
static T *dupload(fbh_ctx *c, const T *h, size_t count) {
  T *d = dalloc<T>(c, count);
  if (!d) return nullptr;
  if (count && hipMemcpy(d, h, count * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
    setErr("upload of %zu bytes failed", count * sizeof(T));
    return nullptr;
  }
  return d;
}

// The panel pass of the FMM tree on the device (src/gkSetup.c:90-301): each
// panel's leaf by the host's own comparisons, a stable sort by leaf (so the
// panels of a leaf keep their input order), the panel arrays in that order,
// and each leaf's digit key, panel count and vertex box.  The rest of the tree
// (cubes, boxes above the leaves, lists, layouts) is fb_tree_build_leaves.
// Host arrays in and out; the leaf arrays hold up to n entries.
extern "C" int fbh_tree_panels(int n, const double *vtx, const double *nrm, const double *area, const int *face,
                               int depth, int *perm, double *vtxR, double *nrmR, double *areaR, int *faceR,
                               int *nLeaf, unsigned *leafKey, int *leafCount, double *leafBox) {
  if (n <= 0 || depth < 1 || depth > 10 || !vtx || !nrm || !area || !face || !perm || !vtxR || !nrmR || !areaR ||
      !faceR || !nLeaf || !leafKey || !leafCount || !leafBox) {
    setErr("fbh_tree_panels: bad arguments (n=%d depth=%d)", n, depth);
    return 0;
  }
  hipStream_t st = nullptr;
  HCK(hipStreamCreate(&st));
  struct Guard {
    hipStream_t st;
    std::vector<void *> p;
    ~Guard() {
      for (void *q : p) hipFree(q);
      hipStreamDestroy(st);
    }
  } G{st, {}};
  auto get = [&](size_t bytes) -> void * {
    void *q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
    G.p.push_back(q);
    return q;
  };
  const size_t N = (size_t)n;
  const int nb = std::min(1024, (n + 255) / 256);
  double *dV = (double *)get(9 * N * 8), *dN = (double *)get(3 * N * 8), *dA = (double *)get(N * 8),
         *dVR = (double *)get(9 * N * 8), *dNR = (double *)get(3 * N * 8), *dAR = (double *)get(N * 8),
         *dPart = (double *)get(6 * (size_t)nb * 8), *dBox = (double *)get(6 * N * 8);
  int *dF = (int *)get(N * 4), *dFR = (int *)get(N * 4), *dIdx = (int *)get(N * 4), *dPerm = (int *)get(N * 4),
      *dCnt = (int *)get(N * 4), *dStart = (int *)get(N * 4), *dNRun = (int *)get(4);
  unsigned *dKey = (unsigned *)get(N * 4), *dKeyS = (unsigned *)get(N * 4), *dUniq = (unsigned *)get(N * 4);
  if (!dV || !dN || !dA || !dVR || !dNR || !dAR || !dPart || !dBox || !dF || !dFR || !dIdx || !dPerm || !dCnt ||
      !dStart || !dNRun || !dKey || !dKeyS || !dUniq) {
    setErr("fbh_tree_panels: device allocation failed (%d panels)", n);
    return 0;
  }
  HCK(hipMemcpyAsync(dV, vtx, 9 * N * 8, hipMemcpyHostToDevice, st));
  HCK(hipMemcpyAsync(dN, nrm, 3 * N * 8, hipMemcpyHostToDevice, st));
  HCK(hipMemcpyAsync(dA, area, N * 8, hipMemcpyHostToDevice, st));
  HCK(hipMemcpyAsync(dF, face, N * 4, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_tree_bbox, dim3(nb), dim3(256), 0, st, dV, n, dPart);
  std::vector<double> part(6 * (size_t)nb);
  HCK(hipMemcpyAsync(part.data(), dPart, part.size() * 8, hipMemcpyDeviceToHost, st));
  HCK(hipStreamSynchronize(st));
  double lo[3], up[3];
  for (int k = 0; k < 3; k++) {
    lo[k] = part[k];
    up[k] = part[3 + k];
  }
  for (int b = 1; b < nb; b++)
    for (int k = 0; k < 3; k++) {
      lo[k] = part[6 * b + k] < lo[k] ? part[6 * b + k] : lo[k];
      up[k] = part[6 * b + 3 + k] > up[k] ? part[6 * b + 3 + k] : up[k];
    }
  // side lengths as the host builds them (src/gkSetup.c:90-128)
  TreeSides side;
  double len = (up[0] - lo[0]) > (up[1] - lo[1]) ? (up[0] - lo[0]) : (up[1] - lo[1]);
  len = (up[2] - lo[2]) > len ? (up[2] - lo[2]) : len;
  for (int lev = 0; lev <= depth; lev++) {
    side.s[lev] = len;
    len *= 0.5;
  }
  hipLaunchKernelGGL(k_tree_keys, dim3((n + 255) / 256), dim3(256), 0, st, dV, n, lo[0], lo[1], lo[2], side, depth,
                     dKey, dIdx);
  size_t tb = 0, tb2 = 0, tb3 = 0;
  HCK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, dKey, dKeyS, dIdx, dPerm, n, 0, 3 * depth, st));
  HCK(hipcub::DeviceRunLengthEncode::Encode(nullptr, tb2, dKeyS, dUniq, dCnt, dNRun, n, st));
  HCK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb3, dCnt, dStart, n, st));
  void *tmp = get(std::max(tb, std::max(tb2, tb3)));
  if (!tmp) { setErr("fbh_tree_panels: sort storage allocation failed"); return 0; }
  tb = std::max(tb, std::max(tb2, tb3));
  size_t t1 = tb;
  HCK(hipcub::DeviceRadixSort::SortPairs(tmp, t1, dKey, dKeyS, dIdx, dPerm, n, 0, 3 * depth, st));
  hipLaunchKernelGGL(k_tree_gather, dim3((n + 255) / 256), dim3(256), 0, st, dPerm, n, dV, dN, dA, dF, dVR, dNR, dAR,
                     dFR);
  t1 = tb;
  HCK(hipcub::DeviceRunLengthEncode::Encode(tmp, t1, dKeyS, dUniq, dCnt, dNRun, n, st));
  int nRun = 0;
  HCK(hipMemcpyAsync(&nRun, dNRun, 4, hipMemcpyDeviceToHost, st));
  HCK(hipStreamSynchronize(st));
  if (nRun < 1 || nRun > n) { setErr("fbh_tree_panels: %d leaves", nRun); return 0; }
  t1 = tb;
  HCK(hipcub::DeviceScan::ExclusiveSum(tmp, t1, dCnt, dStart, nRun, st));
  hipLaunchKernelGGL(k_tree_leafbox, dim3(nRun), dim3(64), 0, st, dStart, dCnt, dVR, dBox);
  HCK(hipGetLastError());
  HCK(hipMemcpyAsync(perm, dPerm, N * 4, hipMemcpyDeviceToHost, st));
  HCK(hipMemcpyAsync(vtxR, dVR, 9 * N * 8, hipMemcpyDeviceToHost, st));
  HCK(hipMemcpyAsync(nrmR, dNR, 3 * N * 8, hipMemcpyDeviceToHost, st));
  HCK(hipMemcpyAsync(areaR, dAR, N * 8, hipMemcpyDeviceToHost, st));
  HCK(hipMemcpyAsync(faceR, dFR, N * 4, hipMemcpyDeviceToHost, st));
  HCK(hipMemcpyAsync(leafKey, dUniq, (size_t)nRun * 4, hipMemcpyDeviceToHost, st));
  HCK(hipMemcpyAsync(leafCount, dCnt, (size_t)nRun * 4, hipMemcpyDeviceToHost, st));
  HCK(hipMemcpyAsync(leafBox, dBox, 6 * (size_t)nRun * 8, hipMemcpyDeviceToHost, st));
  HCK(hipStreamSynchronize(st));
  *nLeaf = nRun;
  return 1;
}

extern "C" int fbh_available(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    setErr("no HIP device");
    return 0;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 0;
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    setErr("device 0 is %s, this build targets gfx950", prop.gcnArchName);
    return 0;
  }
  return 1;
}

// Streams on hardware queues of their own.  HIP hands out a small pool of
// hardware queues to the process's streams, and streams that land on one queue
// run one after the other: the nearfield on the side stream would then wait
// for the whole far chain (or the reverse) instead of sharing the CUs with it.
// A stream created with a CU mask (here every CU) gets a queue outside the
// pool.  Two such streams per device -- the nearfield's side stream and the
// tree gradient's second stream -- are shared by every context of the process
// and live until it ends: queues of their own per context would add up, with
// many contexts alive, until the hardware has to time-slice them.
#define FBH_MAX_DEV 64
static hipStream_t sharedQueueStream(int which) {
  static std::mutex mu;
  static hipStream_t made[FBH_MAX_DEV][2] = {};
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= FBH_MAX_DEV) return nullptr;
  std::lock_guard<std::mutex> g(mu);
  if (!made[dev][which]) {
    hipStream_t st = nullptr;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
      return nullptr;
    std::vector<uint32_t> mask((ncu + 31) / 32, 0xffffffffu);
    if (ncu % 32) mask.back() = (1u << (ncu % 32)) - 1u;
    if (hipExtStreamCreateWithCUMask(&st, (uint32_t)mask.size(), mask.data()) != hipSuccess) return nullptr;
    made[dev][which] = st;
  }
  return made[dev][which];
}

extern "C" void fbh_destroy(fbh_ctx *c) {
  if (!c) return;
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->auxStream) hipStreamSynchronize(c->auxStream);   // shared: not destroyed
  for (auto &b : c->bufs) hipFree(b.p);
  if (c->V) hipFree(c->V);
  if (c->w) hipFree(c->w);
  if (c->r) hipFree(c->r);
  if (c->part) hipFree(c->part);
  if (c->hcol) hipFree(c->hcol);
  if (c->ydev) hipFree(c->ydev);
  if (c->scal) hipFree(c->scal);
  if (c->hpin) hipHostFree(c->hpin);
  if (c->evDone) hipEventDestroy(c->evDone);
  if (c->ro) hipFree(c->ro);
  for (ChargeSet *cs : {&c->csCall, &c->csRes})
    if (cs->arena) hipFree(cs->arena);
  for (void *q : {(void *)c->pbStart, (void *)c->pbM, (void *)c->pbOff, (void *)c->pbTiles, (void *)c->pbInv,
                  (void *)c->pbTmp, (void *)c->pbInvT})
    if (q) hipFree(q);
  if (c->blas) rocblas_destroy_handle(c->blas);
  for (void *q : {(void *)c->dSrcOff, (void *)c->dNcol, (void *)c->dSrcList, (void *)c->dCoefOff, (void *)c->dCoef,
                  (void *)c->dTileCoef})
    if (q) hipFree(q);
  if (c->dx) hipFree(c->dx);
  if (c->side) hipStreamSynchronize(c->side);   // shared: not destroyed
  for (hipEvent_t e : c->ntEv) hipEventDestroy(e);
  if (c->evFork) hipEventDestroy(c->evFork);
  if (c->evJoin) hipEventDestroy(c->evJoin);
  if (c->evX) hipEventDestroy(c->evX);
  if (c->ownStream && c->stream) hipStreamDestroy(c->stream);
  delete c;
}

static int idx3of(int i1, int i2, int i3) {   // graded index: degree, then i1, then i2
  int d = i1 + i2 + i3, base = d > 0 ? NMOM(d - 1) : 0, off = 0;
  for (int a = 0; a < i1; a++) off += d - a + 1;
  return base + off + i2;
}

static bool buildTables(fbh_ctx *c) {
  const int P = c->P, nmP = NMOM(P), nmM = NMOM(c->maxOrder);
  std::vector<int4> mi(nmP), ep(nmP);
  std::vector<double> ifact3(nmP), fact(2 * (MAXP + 3) + 8);
  double f[2 * MAXP + 8];
  f[0] = 1.0;
  for (int k = 1; k < 2 * MAXP + 8; k++) f[k] = f[k - 1] * k;
  for (int k = 0; k < MAXP + 3; k++) fact[k] = f[k];
  for (int k = 0; k < MAXP + 3 + 5 && MAXP + 3 + k < (int)fact.size(); k++) fact[MAXP + 3 + k] = 1.0 / f[k];
  int g = 0, sgn = 1;
  for (int d = 0; d <= P; d++, sgn = -sgn)
    for (int i1 = 0; i1 <= d; i1++)
      for (int i2 = 0; i2 <= d - i1; i2++, g++) {
        int i3 = d - i1 - i2, axis = 0, fac = 0, j1 = -1, j2 = -1;
        auto loc = [&](int a, int b, int cc) { return idx3of(a, b, cc) - ((a + b + cc) > 0 ? NMOM(a + b + cc - 1) : 0); };
        if (d >= 1) {
          if (i1 == 0 && i2 == 0) { axis = 2; j1 = loc(0, 0, d - 1); if (d >= 2) { j2 = loc(0, 0, d - 2); fac = d - 1; } }
          else if (i1 == 0 && i2 == 1) { axis = 1; j1 = loc(0, 0, d - 1); }
          else if (i1 == 0) { axis = 1; j1 = loc(0, i2 - 1, i3); j2 = loc(0, i2 - 2, i3); fac = i2 - 1; }
          else if (i1 == 1) { axis = 0; j1 = loc(0, i2, i3); }
          else { axis = 0; j1 = loc(i1 - 1, i2, i3); j2 = loc(i1 - 2, i2, i3); fac = i1 - 1; }
        }
        mi[g] = make_int4(i1 | (i2 << 8) | (i3 << 16) | (d << 24), axis | (fac << 8), j1, j2);
        ifact3[g] = 1.0 / (f[i1] * f[i2] * f[i3]);
        if (d < P) ep[g] = make_int4(idx3of(i1 + 1, i2, i3), idx3of(i1, i2 + 1, i3), idx3of(i1, i2, i3 + 1), sgn);
        else ep[g] = make_int4(0, 0, 0, sgn);
      }
  // for the noment gather the moment kernel reads b-e_k from ep of the target
  // index; keep that as a separate table appended after the contraction one
  std::vector<int4> em(nmP);
  g = 0;
  for (int d = 0; d <= P; d++)
    for (int i1 = 0; i1 <= d; i1++)
      for (int i2 = 0; i2 <= d - i1; i2++, g++) {
        int i3 = d - i1 - i2;
        em[g] = make_int4(i1 > 0 ? idx3of(i1 - 1, i2, i3) : 0, i2 > 0 ? idx3of(i1, i2 - 1, i3) : 0,
                          i3 > 0 ? idx3of(i1, i2, i3 - 1) : 0, 0);
      }
  // M2L/L2L pairs (beta, alpha+beta), beta graded, |alpha|+|beta| <= maxOrder
  std::vector<int> cs(nmM + 1), cl, ms(nmP + 1), ml;
  g = 0;
  for (int d = 0; d <= c->maxOrder; d++)
    for (int i1 = 0; i1 <= d; i1++)
      for (int i2 = 0; i2 <= d - i1; i2++, g++) {
        int i3 = d - i1 - i2;
        cs[g] = (int)cl.size();
        for (int m = 0; m <= c->maxOrder - d; m++)
          for (int j1 = 0; j1 <= m; j1++)
            for (int j2 = 0; j2 <= m - j1; j2++) {
              int j3 = m - j1 - j2;
              cl.push_back(idx3of(j1, j2, j3) | (idx3of(i1 + j1, i2 + j2, i3 + j3) << 16));
            }
      }
  cs[nmM] = (int)cl.size();
  c->nConv = (int)cl.size();
  // convolution pairs (alpha-k, k), k <= alpha in (k1,k2,k3) loop order
  g = 0;
  for (int d = 0; d <= P; d++)
    for (int i1 = 0; i1 <= d; i1++)
      for (int i2 = 0; i2 <= d - i1; i2++, g++) {
        int i3 = d - i1 - i2;
        ms[g] = (int)ml.size();
        for (int k1 = 0; k1 <= i1; k1++)
          for (int k2 = 0; k2 <= i2; k2++)
            for (int k3 = 0; k3 <= i3; k3++)
              ml.push_back(idx3of(i1 - k1, i2 - k2, i3 - k3) | (idx3of(k1, k2, k3) << 16));
      }
  ms[nmP] = (int)ml.size();
  int4 *dmi = dupload(c, mi.data(), mi.size());
  int4 *dep = dupload(c, ep.data(), ep.size());
  int4 *dem = dupload(c, em.data(), em.size());
  double *dif = dupload(c, ifact3.data(), ifact3.size());
  int *dcs = dupload(c, cs.data(), cs.size()), *dcl = dupload(c, cl.data(), cl.size());
  int *dms = dupload(c, ms.data(), ms.size()), *dml = dupload(c, ml.data(), ml.size());
  c->fact = dupload(c, fact.data(), fact.size());
  if (!dmi || !dep || !dem || !dif || !dcs || !dcl || !dms || !dml || !c->fact) return false;
  c->hConvStart = cs;
  c->hConv = cl;
  c->hM2mStart = ms;
  c->T = DevTabs{dmi, dep, dif, dcs, dcl, dms, dml};
  c->Tm = DevTabs{dmi, dem, dif, dcs, dcl, dms, dml};
  return true;
}

// Balanced M2L plan (k_m2l_cached) for every order 0..maxOrder: alpha's beta
// list is cut into items of `chunk` terms, items go to the least loaded of the
// 64 lanes, longest first; the last term of an item is flagged with its number.
// Returns the dynamic LDS the kernel needs with `waves` waves, at the largest
// order used by a level.
static size_t m2lPlanLds(const fbh_ctx *c, const std::vector<int4> &off, int waves, size_t elem) {
  size_t best = 0;
  for (int l = c->height; l <= c->depth; l++) {
    const int p = c->ordM2L[l], nm = NMOM(p);
    const int4 po = off[p];
    size_t b = ((size_t)waves * 4 * ((nm + 1) & ~1) + (size_t)waves * 4 * po.w) * elem + (size_t)po.y * sizeof(unsigned) +
               (size_t)(65 + nm + 1) * sizeof(int);
    best = std::max(best, b);
  }
  return best;
}

static bool buildM2LPlan(fbh_ctx *c) {
  std::vector<unsigned> terms;
  std::vector<int> idx;
  std::vector<int4> off(c->maxOrder + 1);
  std::vector<int> laneTerms(c->maxOrder + 1, 0);
  std::vector<int> deg(NMOM(c->maxOrder));
  for (int d = 0, g = 0; d <= c->maxOrder; d++)
    for (int k = 0; k < (d + 1) * (d + 2) / 2; k++) deg[g++] = d;
  for (int p = 0; p <= c->maxOrder; p++) {
    const int nm = NMOM(p);
    long long total = 0;
    for (int a = 0; a < nm; a++) total += NMOM(p - deg[a]);
    const int chunk = std::max(8, (int)((total + 127) / 128));
    struct Item { int alpha, first, len; };
    std::vector<Item> items;
    std::vector<int> itemStart(nm + 1);
    for (int a = 0; a < nm; a++) {
      itemStart[a] = (int)items.size();
      const int cnt = NMOM(p - deg[a]);
      for (int f = 0; f < cnt; f += chunk) items.push_back({a, f, std::min(chunk, cnt - f)});
    }
    itemStart[nm] = (int)items.size();
    if (items.size() > 0x7fff) { setErr("M2L plan: %zu items at order %d", items.size(), p); return false; }
    std::vector<int> order(items.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return items[x].len > items[y].len; });
    std::vector<std::vector<int>> lane(64);
    std::vector<long long> load(64, 0);
    for (int it : order) {
      int best = 0;
      for (int l = 1; l < 64; l++)
        if (load[l] < load[best]) best = l;
      lane[best].push_back(it);
      load[best] += items[it].len;
    }
    int4 po;
    po.x = (int)terms.size();
    po.z = (int)idx.size();
    po.w = (int)items.size();
    std::vector<int> laneStart(65);
    for (int l = 0; l < 64; l++) {
      laneStart[l] = (int)terms.size() - po.x;
      for (int it : lane[l]) {
        const Item &I = items[it];
        const int base = c->hConvStart[I.alpha] + I.first;
        for (int e = 0; e < I.len; e++) {
          const int pe = c->hConv[base + e];
          unsigned t = (unsigned)(pe & 0xffff) | ((unsigned)(pe >> 16) << 8);
          if (e == I.len - 1) t |= M2L_FLUSH | ((unsigned)it << 16);
          terms.push_back(t);
        }
      }
    }
    laneStart[64] = (int)terms.size() - po.x;
    po.y = laneStart[64];
    for (int l = 0; l < 64; l++) laneTerms[p] = std::max(laneTerms[p], laneStart[l + 1] - laneStart[l]);
    idx.insert(idx.end(), laneStart.begin(), laneStart.end());
    idx.insert(idx.end(), itemStart.begin(), itemStart.end());
    off[p] = po;
  }
  c->m2lWaves = 0;
  for (int wv = 4; wv >= 1; wv /= 2)
    if (m2lPlanLds(c, off, wv, sizeof(double)) <= (size_t)c->ldsMax) { c->m2lWaves = wv; break; }
  if (!c->m2lWaves) return true;   // no plan fits: the recompute kernel is used
  c->m2lLds = m2lPlanLds(c, off, c->m2lWaves, sizeof(double));
  c->m2lLds32 = m2lPlanLds(c, off, c->m2lWaves, sizeof(float));
  int mt = 0;   // longest lane term list at the orders the levels use
  for (int l = c->height; l <= c->depth; l++) mt = std::max(mt, laneTerms[c->ordM2L[l]]);
  c->m2lMaxT = mt <= 16 ? 16 : mt <= 32 ? 32 : 0;
  if (!(c->planTerms = dupload(c, terms.data(), terms.size())) || !(c->planIdx = dupload(c, idx.data(), idx.size())) ||
      !(c->planOff = dupload(c, off.data(), off.size())))
    return false;
  return true;
}

// Term grid of k_m2l_grid for order p on K steps x 64 lanes, runs of L steps.
// The terms of one alpha are cut into cells of L (the last one shorter); the
// cells, longest first, fill the lanes lane by lane (cell i: run i % R, lane
// i / R).  Then, step by step, each busy lane takes the next beta of its
// alpha (lanes whose alpha has fewest terms left choose first) among the ones
// whose moment slot and tensor slot share no bank with a different slot of the
// lanes already placed in its 8- and 16-lane group (16-byte slots; else the
// fewest such clashes).  Idle lanes read the zero moment slot and a tensor slot
// another lane of the group reads.  Returns false when the cells do not fit.
static bool planGrid(const fbh_ctx *c, int p, int K, int L, int slotZ, std::vector<unsigned> &terms,
                     std::vector<int> &redStart, std::vector<unsigned short> &redCell) {
  const int nm = NMOM(p), R = K / L;
  std::vector<int> deg(nm);
  for (int d = 0, g = 0; d <= p; d++)
    for (int k = 0; k < (d + 1) * (d + 2) / 2; k++) deg[g++] = d;
  std::vector<std::vector<std::pair<int, int>>> pool(nm);
  struct Cell { int alpha, cnt; };
  std::vector<Cell> cells;
  for (int a = 0; a < nm; a++) {
    const int cnt = NMOM(p - deg[a]);
    for (int e = 0; e < cnt; e++) {
      const int pe = c->hConv[c->hConvStart[a] + e];
      pool[a].push_back({pe & 0xffff, pe >> 16});
    }
    for (int f = 0; f < cnt; f += L) cells.push_back({a, std::min(L, cnt - f)});
  }
  if ((int)cells.size() > R * 64) return false;
  std::stable_sort(cells.begin(), cells.end(), [](const Cell &x, const Cell &y) { return x.cnt > y.cnt; });
  std::vector<int> place((size_t)R * 64, -1);
  for (size_t i = 0; i < cells.size(); i++) place[(i % R) * 64 + i / R] = (int)i;
  std::vector<int> cs(64), gs(64);   // this step's slots per lane (-1: not placed yet)
  terms.assign((size_t)K * 64, 0);
  auto clash = [&](const std::vector<int> &v, int l, int s) {
    int n = 0;
    for (int G = 8; G <= 16; G *= 2)
      for (int l2 = l / G * G; l2 < l / G * G + G; l2++)
        if (l2 != l && v[l2] >= 0 && v[l2] != s && v[l2] % G == s % G) n++;
    return n;
  };
  for (int r = 0; r < R; r++)
    for (int t = 0; t < L; t++) {
      std::fill(cs.begin(), cs.end(), -1);
      std::fill(gs.begin(), gs.end(), -1);
      std::vector<int> busy;
      for (int l = 0; l < 64; l++) {
        const int ci = place[r * 64 + l];
        if (ci >= 0 && cells[ci].cnt > t) busy.push_back(l);
      }
      std::stable_sort(busy.begin(), busy.end(), [&](int x, int y) {
        return pool[cells[place[r * 64 + x]].alpha].size() < pool[cells[place[r * 64 + y]].alpha].size();
      });
      for (int l : busy) {
        auto &pl = pool[cells[place[r * 64 + l]].alpha];
        size_t best = 0;
        int bc = 1 << 30;
        for (size_t i = 0; i < pl.size() && bc > 0; i++) {
          const int k = clash(cs, l, pl[i].first) + clash(gs, l, pl[i].second);
          if (k < bc) { bc = k; best = i; }
        }
        cs[l] = pl[best].first;
        gs[l] = pl[best].second;
        pl.erase(pl.begin() + best);
      }
      for (int l = 0; l < 64; l++)
        if (cs[l] < 0) {
          int g = -1;
          for (int G = 8; G <= 16 && g < 0; G *= 2)
            for (int l2 = l / G * G; l2 < l / G * G + G && g < 0; l2++)
              if (gs[l2] >= 0 && cs[l2] != slotZ) g = gs[l2];
          cs[l] = slotZ;
          gs[l] = g < 0 ? 0 : g;
        }
      for (int l = 0; l < 64; l++) terms[(size_t)(r * L + t) * 64 + l] = (unsigned)cs[l] | ((unsigned)gs[l] << 16);
    }
  for (int a = 0; a < nm; a++)
    if (!pool[a].empty()) return false;
  redStart.assign(nm + 1, 0);
  std::vector<std::vector<int>> byAlpha(nm);
  for (int cell = 0; cell < R * 64; cell++)
    if (place[cell] >= 0) byAlpha[cells[place[cell]].alpha].push_back(cell);
  redCell.clear();
  for (int a = 0; a < nm; a++) {
    redStart[a] = (int)redCell.size();
    for (int cell : byAlpha[a]) redCell.push_back((unsigned short)cell);
  }
  redStart[nm] = (int)redCell.size();
  return true;
}

// The term grid for every order the levels use, on the smallest compiled shape
// (K, L) that holds the largest of them; none above order 8 (k_m2l_cached then).
static bool buildM2LGrid(fbh_ctx *c) {
  int pU = 0;
  for (int l = c->height; l <= c->depth; l++) pU = std::max(pU, c->ordM2L[l]);
  static const int shapes[][3] = {{3, 4, 1}, {5, 8, 1}, {6, 16, 2}, {7, 32, 4}, {8, 64, 8}};   // max order, K, L
  int K = 0, L = 0;
  for (auto &sh : shapes)
    if (pU <= sh[0]) { K = sh[1]; L = sh[2]; break; }
  const char *em = getenv("FABIPB_M2L");   // "items": the item-list kernel only (check path)
  if (!K || (em && !strcmp(em, "items"))) return true;
  const int nmU = NMOM(pU);
  std::vector<unsigned> allTerms;
  std::vector<int> off(2 * (pU + 1), 0), allStart;
  std::vector<unsigned short> allCell;
  std::vector<bool> used(pU + 1, false);
  for (int l = c->height; l <= c->depth; l++) used[c->ordM2L[l]] = true;
  for (int p = 0; p <= pU; p++) {
    if (!used[p]) continue;
    std::vector<unsigned> t;
    std::vector<int> rs;
    std::vector<unsigned short> rc;
    if (!planGrid(c, p, K, L, nmU, t, rs, rc)) return true;   // does not fit: k_m2l_cached
    off[2 * p] = (int)allTerms.size();
    off[2 * p + 1] = (int)allStart.size();
    for (int &v : rs) v += (int)allCell.size();
    allTerms.insert(allTerms.end(), t.begin(), t.end());
    allStart.insert(allStart.end(), rs.begin(), rs.end());
    allCell.insert(allCell.end(), rc.begin(), rc.end());
  }
  // waves per block (one cube): four when there are few cubes, two from 4,096
  // cubes on, where more cubes in flight per CU pay more than more waves per
  // cube (317k panels: 212 -> 184 us; 81,920: 63 us either way)
  const int ncc = (nmU + 2) & ~1, ngg = (nmU + 1) & ~1, nw = c->nCube >= 4096 ? 2 : 4;
  auto lds = [&](size_t elem) {
    return std::max((size_t)nw * (ncc + ngg) * 2 * elem, (size_t)(K / L) * 64 * 4 * elem);
  };
  if (lds(sizeof(double)) > (size_t)c->ldsMax) return true;
  std::vector<int> order(c->nCube);   // cubes by interaction-list length, longest first
  for (int i = 0; i < c->nCube; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return c->hM2lOff[a + 1] - c->hM2lOff[a] > c->hM2lOff[b + 1] - c->hM2lOff[b];
  });
  if (!(c->m2lOrder = dupload(c, order.data(), order.size()))) return false;
  if (!(c->gridTerms = dupload(c, allTerms.data(), allTerms.size())) ||
      !(c->gridOff = dupload(c, off.data(), off.size())) ||
      !(c->gridRedStart = dupload(c, allStart.data(), allStart.size())) ||
      !(c->gridRedCell = dupload(c, allCell.data(), allCell.size())))
    return false;
  c->gridK = K;
  c->gridL = L;
  c->gridSlotZ = nmU;
  c->gridNcc = ncc;
  c->gridNgg = ngg;
  c->gridLds = lds(sizeof(double));
  c->gridLds32 = lds(sizeof(float));
  c->gridWaves = nw;
  return true;
}

static inline int blk64(int n, int cap) {
  int b = ((n + 63) / 64) * 64;
  return std::max(64, std::min(b, cap));
}

static int enableFly64(fbh_ctx *c);
static int resolveScratch(int dev);

extern "C" fbh_ctx *fbh_create(double kappa, double eps1, double eps2, int qorder, int n, const double *vtx,
                               const double *nrm, const double *area, int depth, int height, const int *ordMom,
                               const int *ordM2L, int nCube, const int *levStart, const int *levCount,
                               const int *cubeLevel, const double *cubeX, const int *kidStart, const int *nKids,
                               const int *parent, int nLeaf, const int *leafStart, const int *leafCount,
                               const int *nearOff, const int *nearSrc, const int *m2lOff, const int *m2lSrc,
                               void *stream) {
  if (!fbh_available()) return nullptr;
  if (n <= 0 || nLeaf <= 0 || nCube < nLeaf || depth < height || qorder < 1 || qorder > 10 || !(eps1 > 0) ||
      !(eps2 > 0)) {
    setErr("fbh_create: bad arguments (n=%d nLeaf=%d nCube=%d depth=%d height=%d q=%d)", n, nLeaf, nCube, depth,
           height, qorder);
    return nullptr;
  }
  int maxOrder = 0;
  for (int l = height; l <= depth; l++) {
    if (ordMom[l] < 0 || ordM2L[l] < 0) {
      setErr("fbh_create: negative expansion order at level %d", l);
      return nullptr;
    }
    maxOrder = std::max(maxOrder, std::max(ordMom[l], ordM2L[l]));
  }
  if (maxOrder < 1 || maxOrder > 9) {
    setErr("fbh_create: expansion order %d outside 1..9", maxOrder);
    return nullptr;
  }
  // host-side validation of every index the kernels dereference
  for (int l = 0; l < nLeaf; l++)
    if (leafStart[l] < 0 || leafCount[l] <= 0 || leafStart[l] + leafCount[l] > n) {
      setErr("fbh_create: leaf %d range [%d,+%d) outside %d panels", l, leafStart[l], leafCount[l], n);
      return nullptr;
    }
  for (int c = 0; c < nCube; c++) {
    if (cubeLevel[c] < height || cubeLevel[c] > depth || (c < nLeaf) != (cubeLevel[c] == depth) ||
        nKids[c] < 0 || nKids[c] > 8 || (nKids[c] > 0 && (kidStart[c] < 0 || kidStart[c] + nKids[c] > nCube)) ||
        (cubeLevel[c] > height && (parent[c] < 0 || parent[c] >= nCube))) {
      setErr("fbh_create: cube %d has an invalid level/kid/parent record", c);
      return nullptr;
    }
    if (m2lOff[c + 1] < m2lOff[c]) { setErr("fbh_create: m2lOff not monotone at %d", c); return nullptr; }
  }
  for (int k = 0; k < m2lOff[nCube]; k++)
    if (m2lSrc[k] < 0 || m2lSrc[k] >= nCube) { setErr("fbh_create: m2l source %d out of range", k); return nullptr; }
  for (int l = 0; l < nLeaf; l++) {
    if (nearOff[l + 1] < nearOff[l]) { setErr("fbh_create: nearOff not monotone at %d", l); return nullptr; }
    for (int k = nearOff[l]; k < nearOff[l + 1]; k++)
      if (nearSrc[k] < 0 || nearSrc[k] >= nLeaf) { setErr("fbh_create: near source %d out of range", k); return nullptr; }
  }

  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) { setErr("fbh_create: no current device"); return nullptr; }
  if (!resolveScratch(dev)) return nullptr;   // the no-allocation staging array of this device
  fbh_ctx *c = new fbh_ctx();
  c->dev = dev;
  c->kappa = kappa; c->eps1 = eps1; c->eps2 = eps2; c->eps = eps2 / eps1; c->qorder = qorder;
  for (int i = 0; i < 10; i++) { c->Q.t[i] = hT[qorder][i]; c->Q.w[i] = hW[qorder][i]; }
  c->n = n; c->depth = depth; c->height = height; c->nCube = nCube; c->nLeaf = nLeaf;
  c->maxOrder = maxOrder; c->P = maxOrder + 1; c->nmMax = NMOM(maxOrder); c->nmLeaf = NMOM(ordMom[depth]);
  c->levStart.assign(levStart, levStart + depth + 1);
  c->levCount.assign(levCount, levCount + depth + 1);
  c->ordMom.assign(ordMom, ordMom + depth + 1);
  c->ordM2L.assign(ordM2L, ordM2L + depth + 1);
  c->cubeLevel.assign(cubeLevel, cubeLevel + nCube);
  c->hKidStart.assign(kidStart, kidStart + nCube);
  c->hNKids.assign(nKids, nKids + nCube);
  c->hLeafStart.assign(leafStart, leafStart + nLeaf);
  c->hLeafCount.assign(leafCount, leafCount + nLeaf);
  c->hM2lOff.assign(m2lOff, m2lOff + nCube + 1);
  if (stream) {
    c->stream = (hipStream_t)stream;
  } else {
    if (hipStreamCreate(&c->stream) != hipSuccess) { setErr("hipStreamCreate failed"); delete c; return nullptr; }
    c->ownStream = true;
  }
#define FAIL() do { fbh_destroy(c); return nullptr; } while (0)
  if (!(c->side = sharedQueueStream(0)) ||
      hipEventCreateWithFlags(&c->evFork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->evJoin, hipEventDisableTiming) != hipSuccess) {
    setErr("stream/event creation failed");
    FAIL();
  }
  {
    int dev = 0, lds = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
      c->ldsMax = std::min(lds, 65536);
  }
  if (!buildTables(c)) FAIL();
  // panels
  std::vector<Pnl> hp(n);
  for (int i = 0; i < n; i++) {
    const double *v = &vtx[9 * i];
    for (int k = 0; k < 9; k++) hp[i].v[k] = v[k];
    for (int j = 0; j < 3; j++) {   // src/input.c:743-747
      hp[i].a[j] = v[6 + j] - v[3 + j];
      hp[i].a[3 + j] = v[j] - v[6 + j];
      hp[i].a[6 + j] = v[3 + j] - v[j];
    }
    for (int k = 0; k < 3; k++) hp[i].nr[k] = nrm[3 * i + k];
    hp[i].area = area[i];
  }
  std::vector<int> leafOf(n);
  for (int l = 0; l < nLeaf; l++)
    for (int j = 0; j < leafCount[l]; j++) leafOf[leafStart[l] + j] = l;
  if (!(c->pnl = dupload(c, hp.data(), n)) || !(c->area = dupload(c, area, n)) ||
      !(c->leafOf = dupload(c, leafOf.data(), n)) || !(c->leafStart = dupload(c, leafStart, nLeaf)) ||
      !(c->leafCount = dupload(c, leafCount, nLeaf)) || !(c->cubeX = dupload(c, cubeX, 3 * (size_t)nCube)) ||
      !(c->cubeLevelD = dupload(c, cubeLevel, nCube)) || !(c->kidStart = dupload(c, kidStart, nCube)) ||
      !(c->nKids = dupload(c, nKids, nCube)) || !(c->parent = dupload(c, parent, nCube)) ||
      !(c->m2lOff = dupload(c, m2lOff, nCube + 1)) || !(c->m2lSrc = dupload(c, m2lSrc, std::max(1, m2lOff[nCube]))) ||
      !(c->ordMomD = dupload(c, ordMom, depth + 1)) || !(c->ordM2LD = dupload(c, ordM2L, depth + 1)))
    FAIL();
  c->nM2L = m2lOff[nCube];
  c->nLinks = 0;
  for (int l = height; l < depth; l++)
    for (int cb = levStart[l]; cb < levStart[l] + levCount[l]; cb++) c->nLinks += nKids[cb];

  // expansions and work vectors
  if (!(c->Q0 = dalloc<double>(c, (size_t)n * c->nmLeaf)) || !(c->Q1 = dalloc<double>(c, (size_t)n * c->nmLeaf)) ||
      !(c->mom = dalloc<double>(c, (size_t)nCube * 2 * c->nmMax)) ||
      !(c->lec = dalloc<double>(c, (size_t)nCube * 4 * c->nmMax)) || !(c->far = dalloc<double>(c, 2 * (size_t)n)) ||
      !(c->nearv = dalloc<double>(c, 2 * (size_t)n)) || !(c->xs = dalloc<double>(c, 2 * (size_t)n)) ||
      !(c->ys = dalloc<double>(c, 2 * (size_t)n)))
    FAIL();

  // expansion entries beyond a level's order are never written by the passes
  // and never read; zero them once so the buffers hold no stale values
  if (hipMemsetAsync(c->mom, 0, (size_t)nCube * 2 * c->nmMax * sizeof(double), c->stream) != hipSuccess ||
      hipMemsetAsync(c->lec, 0, (size_t)nCube * 4 * c->nmMax * sizeof(double), c->stream) != hipSuccess) {
    setErr("clearing expansions failed");
    FAIL();
  }

  // leaf Q2M/L2P matrices on the device
  {
    size_t lds = 5 * (size_t)c->nmLeaf * sizeof(double);
    hipLaunchKernelGGL(k_q2m_mats, dim3(n), dim3(64), lds, c->stream, c->pnl, c->leafOf, c->cubeX, c->Tm,
                       ordMom[depth], c->fact, c->Q0, c->Q1);
    if (hipGetLastError() != hipSuccess) { setErr("k_q2m_mats launch failed"); FAIL(); }
  }

  // nearfield layout: per destination leaf, its source columns and coefficient rows
  std::vector<int> srcOff(nLeaf + 1), ncol(nLeaf);
  std::vector<long long> coefOff(nLeaf + 1);
  std::vector<int> srcList;
  std::vector<int2> tiles;
  srcOff[0] = 0; coefOff[0] = 0;
  std::vector<int> ownOff(nLeaf, -1);
  for (int l = 0; l < nLeaf; l++) {
    int nc = 0;
    for (int k = nearOff[l]; k < nearOff[l + 1]; k++) {
      int s = nearSrc[k];
      if (s == l && ownOff[l] < 0) ownOff[l] = nc;
      for (int j = 0; j < leafCount[s]; j++) srcList.push_back(leafStart[s] + j);
      nc += leafCount[s];
    }
    ncol[l] = nc;
    srcOff[l + 1] = srcOff[l] + nc;
    coefOff[l + 1] = coefOff[l] + (long long)leafCount[l] * nc;
    for (int r = 0; r < leafCount[l]; r += NEAR_ROWS) tiles.push_back(make_int2(l, r));
  }
  c->nInter = coefOff[nLeaf];
  c->nTiles = (int)tiles.size();
  {   // tiles costliest first (rows x columns) for the grid-stride launch
    std::vector<int> order(tiles.size());
    std::vector<long long> cost(tiles.size());
    for (size_t t = 0; t < tiles.size(); t++) {
      order[t] = (int)t;
      cost[t] = (long long)std::min(NEAR_ROWS, leafCount[tiles[t].x] - tiles[t].y) * ncol[tiles[t].x];
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    if (!(c->tileOrder = dupload(c, order.data(), order.size())) || !(c->ownOff = dupload(c, ownOff.data(), ownOff.size())))
      FAIL();
    // blocks of the FP64 nearfield launch (0: one per tile, the fastest here;
    // resident grids of 2-4 blocks per CU were slower beside the far chain)
    const char *eg = getenv("FABIPB_NEAR_GRID");
    const int g = eg ? atoi(eg) : 0;
    c->nearGrid = g > 0 ? std::min(g, c->nTiles) : c->nTiles;
  }
  {   // the adjoint operator relies on symmetric near and M2L lists
    auto symmetric = [](int cnt, const int *off, const int *lst) {
      std::vector<std::pair<int, int>> e;
      for (int a = 0; a < cnt; a++)
        for (int k = off[a]; k < off[a + 1]; k++) e.push_back({a, lst[k]});
      std::sort(e.begin(), e.end());
      for (auto &pr : e)
        if (!std::binary_search(e.begin(), e.end(), std::make_pair(pr.second, pr.first))) return false;
      return true;
    };
    c->nearSym = symmetric(nLeaf, nearOff, nearSrc);
    c->m2lSym = symmetric(nCube, m2lOff, m2lSrc);
    for (int cb = 0; cb < nCube && c->m2lSym; cb++)   // pairs at one level share their orders
      for (int k = m2lOff[cb]; k < m2lOff[cb + 1]; k++)
        if (cubeLevel[m2lSrc[k]] != cubeLevel[cb]) { c->m2lSym = false; break; }
  }
  // Tiles that stream stored coefficients and tiles computed on the fly (q = 1
  // only), interleaved in tile order: a tile streams while the streamed share
  // of the interactions so far stays within the fraction f (FABIPB_NEAR_STREAM;
  // FABIPB_NEAR=stream or fly for all of one kind).  The default computes every
  // tile: the on-the-fly tiles (two target rows per pass over the staged
  // sources) outrun the 32-byte coefficient stream, and the far chain's
  // memory-bound passes overlap them better than they overlap the stream.
  double f = 0.0;
  {
    const char *fs = getenv("FABIPB_NEAR_STREAM"), *fm = getenv("FABIPB_NEAR");
    if (fs) f = std::min(1.0, std::max(0.0, atof(fs)));
    if (fm && !strcmp(fm, "stream")) f = 1.0;
    if (fm && !strcmp(fm, "fly")) f = 0.0;
    if (qorder != 1) f = 1.0;
    // the store is capped at a fifth of the free device memory (the FP32 copy
    // and the adjoint's transposed store come on top): at capsid scale more of
    // the nearfield is computed
    size_t freeB = 0, totalB = 0;
    if (qorder == 1 && hipMemGetInfo(&freeB, &totalB) == hipSuccess && c->nInter > 0)
      f = std::min(f, 0.2 * (double)freeB / (32.0 * (double)c->nInter));
  }
  std::vector<long long> tileCoef(tiles.size());
  {
    long long seen = 0;
    c->nStored = 0;
    c->nFlyTiles = 0;
    for (size_t t = 0; t < tiles.size(); t++) {
      const int D = tiles[t].x, r0 = tiles[t].y, rows = std::min(NEAR_ROWS, leafCount[D] - r0);
      const long long inter = (long long)rows * ncol[D];
      if ((double)(c->nStored + inter) <= f * (double)(seen + inter)) {
        tileCoef[t] = c->nStored;
        c->nStored += inter;
      } else {
        tileCoef[t] = -1;
        c->nFlyTiles++;
      }
      seen += inter;
    }
  }
  if (!(c->srcOff = dupload(c, srcOff.data(), nLeaf + 1)) || !(c->ncol = dupload(c, ncol.data(), nLeaf)) ||
      !(c->srcList = dupload(c, srcList.data(), srcList.size())) ||
      !(c->tiles = dupload(c, tiles.data(), tiles.size())) ||
      !(c->tileCoef = dupload(c, tileCoef.data(), tiles.size())) ||
      !(c->coef = dalloc<double4>(c, (size_t)std::max<long long>(c->nStored, 1))))
    FAIL();
  if (c->nStored > 0) {
    hipLaunchKernelGGL(k_near_build_tiles<false>, dim3(c->nTiles, NEAR_ROWS), dim3(256), 0, c->stream, c->pnl, c->tiles,
                       c->leafStart, c->leafCount, c->srcOff, c->ncol, c->tileCoef, c->srcList, c->coef, kappa,
                       c->eps, c->Q, qorder);
    if (hipGetLastError() != hipSuccess) { setErr("k_near_build_tiles launch failed"); FAIL(); }
  }
  // M2L derivative tensors per pair, cached when they fit (1344 B per pair at p=6)
  c->m2lG = nullptr;
  if (!buildM2LPlan(c) || !buildM2LGrid(c)) FAIL();
  if (c->m2lWaves > 0 && c->nM2L > 0 && (double)c->nM2L * 2 * c->nmMax * sizeof(double) <= M2L_CACHE_BYTES) {
    std::vector<int> pd((size_t)c->nM2L);
    for (int cb = 0; cb < nCube; cb++)
      for (int k = m2lOff[cb]; k < m2lOff[cb + 1]; k++) pd[k] = cb;
    if (!(c->m2lDst = dupload(c, pd.data(), pd.size())) ||
        !(c->m2lG = dalloc<double>(c, (size_t)c->nM2L * 2 * c->nmMax)))
      FAIL();
    hipLaunchKernelGGL(k_m2l_derivs, dim3((unsigned)c->nM2L), dim3(64),
                       2 * (size_t)derivScratch(c->maxOrder) * sizeof(double), c->stream, c->m2lDst, c->m2lSrc,
                       c->cubeLevelD, c->ordM2LD, c->cubeX, c->m2lG, c->nmMax, kappa, c->T);
    if (hipGetLastError() != hipSuccess) { setErr("k_m2l_derivs launch failed"); FAIL(); }
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) { setErr("setup kernels failed: %s", hipGetErrorString(hipGetLastError())); FAIL(); }
  if (c->nFlyTiles > 0 && !enableFly64(c)) FAIL();
#undef FAIL
  return c;
}

// the operator's buffers in one arithmetic: FP64 (plain form) or FP32 (precision form)
template <typename R>
struct OpBufs {
  const R *Q0, *Q1, *G;
  R *mom, *lec, *far, *nearv;
  const typename Vec4<R>::type *coef;
  size_t m2lLds;
};
static OpBufs<double> bufs64(fbh_ctx *c) {
  return {c->Q0, c->Q1, c->m2lG, c->mom, c->lec, c->far, c->nearv, c->coef, c->m2lLds};
}
static OpBufs<float> bufs32(fbh_ctx *c) {
  return {c->Q0f, c->Q1f, c->m2lGf, c->momf, c->lecf, c->farf, c->nearf, c->coef32, c->m2lLds32};
}

template <typename R, int MT>
static void m2lLaunch(fbh_ctx *c, const OpBufs<R> &B, hipStream_t st) {
  hipLaunchKernelGGL((k_m2l_cached<R, MT>), dim3(c->nCube), dim3(64 * c->m2lWaves), B.m2lLds, st, c->m2lOff,
                     c->m2lSrc, c->cubeLevelD, c->ordM2LD, c->ordMomD, B.G, B.mom, B.lec, c->nmMax, (R)c->eps, c->T,
                     c->planTerms, c->planIdx, c->planOff);
}

template <typename R, int K, int L, bool TR = false, bool DK = false>
static void m2lGridLaunch(fbh_ctx *c, const R *G, const R *mom, R *lec, size_t lds, hipStream_t st) {
  hipLaunchKernelGGL((k_m2l_grid<R, K, L, TR, DK>), dim3(c->nCube), dim3(64 * c->gridWaves), lds, st, c->m2lOff,
                     c->m2lSrc,
                     c->cubeLevelD, c->ordM2LD, c->ordMomD, G, mom, lec, c->nmMax, (R)c->eps, c->T, c->gridTerms,
                     c->gridOff, c->gridRedStart, c->gridRedCell, c->gridSlotZ, c->gridNcc, c->gridNgg,
                     c->m2lOrder);
}

// k_m2l_grid at the shape the plan was built for; false when there is no grid
template <typename R, bool TR = false, bool DK = false>
static bool m2lGrid(fbh_ctx *c, const R *G, const R *mom, R *lec, hipStream_t st) {
  if (!c->gridK || !G) return false;
  const size_t lds = sizeof(R) == sizeof(double) ? c->gridLds : c->gridLds32;
  switch (c->gridK) {
    case 4: m2lGridLaunch<R, 4, 1, TR, DK>(c, G, mom, lec, lds, st); return true;
    case 8: m2lGridLaunch<R, 8, 1, TR, DK>(c, G, mom, lec, lds, st); return true;
    case 16: m2lGridLaunch<R, 16, 2, TR, DK>(c, G, mom, lec, lds, st); return true;
    case 32: m2lGridLaunch<R, 32, 4, TR, DK>(c, G, mom, lec, lds, st); return true;
    case 64: m2lGridLaunch<R, 64, 8, TR, DK>(c, G, mom, lec, lds, st); return true;
  }
  return false;
}

template <typename R>
static void m2lPass(fbh_ctx *c, const OpBufs<R> &B, hipStream_t st) {
  if (m2lGrid<R>(c, B.G, B.mom, B.lec, st)) return;
  if (B.G) {
    if (c->m2lMaxT == 16) m2lLaunch<R, 16>(c, B, st);
    else if (c->m2lMaxT == 32) m2lLaunch<R, 32>(c, B, st);
    else m2lLaunch<R, 0>(c, B, st);
    return;
  }
  if constexpr (std::is_same<R, double>::value) {   // the FP32 operator always has the cache
    const int pm = c->maxOrder, nm = NMOM(pm);
    size_t lds = (2 * (size_t)derivScratch(pm) + 7 * (size_t)nm) * sizeof(double);
    hipLaunchKernelGGL(k_m2l, dim3(c->nCube), dim3(64), lds, st, c->m2lOff, c->m2lSrc, c->cubeLevelD, c->ordM2LD,
                       c->ordMomD, c->cubeX, B.mom, B.lec, c->nmMax, c->kappa, c->eps, c->T);
  }
}

// the Q2M / L2P launches: 16-byte vectors where nm allows
template <typename R>
static int vecWidth(int nm) {
  const int v = 16 / (int)sizeof(R);
  return nm % v == 0 ? v : nm % 2 == 0 ? 2 : 1;
}
template <typename R>
static void launchQ2M(const fbh_ctx *c, const R *Q0, const R *Q1, const double *x, R *mom, int swap, hipStream_t st) {
  const int V = vecWidth<R>(c->nmLeaf);
  const size_t lds = 512 * (size_t)V * sizeof(R);
  if (V == 4)
    hipLaunchKernelGGL((k_q2m<R, 4>), dim3(c->nLeaf), dim3(256), lds, st, Q0, Q1, c->leafStart, c->leafCount, x, mom,
                       c->n, c->nmLeaf, c->nmMax, swap);
  else if (V == 2)
    hipLaunchKernelGGL((k_q2m<R, 2>), dim3(c->nLeaf), dim3(256), lds, st, Q0, Q1, c->leafStart, c->leafCount, x, mom,
                       c->n, c->nmLeaf, c->nmMax, swap);
  else
    hipLaunchKernelGGL((k_q2m<R, 1>), dim3(c->nLeaf), dim3(256), lds, st, Q0, Q1, c->leafStart, c->leafCount, x, mom,
                       c->n, c->nmLeaf, c->nmMax, swap);
}
template <typename R, bool TR = false>
static void launchL2P(const fbh_ctx *c, const R *Q0, const R *Q1, const R *lec, R *far, hipStream_t st) {
  const int V = vecWidth<R>(c->nmLeaf), G = c->nmLeaf / V;
  const size_t lds = (2 * (size_t)c->nmLeaf + 2 * (size_t)L2P_CH * G) * sizeof(R);
  if (V == 4)
    hipLaunchKernelGGL((k_l2p<R, 4, TR>), dim3(c->nLeaf), dim3(256), lds, st, Q0, Q1, c->leafStart, c->leafCount, lec,
                       far, c->n, c->nmLeaf, c->nmMax);
  else if (V == 2)
    hipLaunchKernelGGL((k_l2p<R, 2, TR>), dim3(c->nLeaf), dim3(256), lds, st, Q0, Q1, c->leafStart, c->leafCount, lec,
                       far, c->n, c->nmLeaf, c->nmMax);
  else
    hipLaunchKernelGGL((k_l2p<R, 1, TR>), dim3(c->nLeaf), dim3(256), lds, st, Q0, Q1, c->leafStart, c->leafCount, lec,
                       far, c->n, c->nmLeaf, c->nmMax);
}

template <typename R>
static void q2mPass(fbh_ctx *c, const OpBufs<R> &B, const double *x, hipStream_t st) {
  launchQ2M<R>(c, B.Q0, B.Q1, x, B.mom, 0, st);
}

template <typename R>   // children at lev + 1 into parents at lev
static void m2mPass(fbh_ctx *c, const OpBufs<R> &B, int lev, hipStream_t st) {
  const int nOut = NMOM(c->ordMom[lev]);
  if (c->levCount[lev] <= 0) return;
  size_t lds = 24 * (size_t)nOut * sizeof(R) + (size_t)(nOut + 1 + c->hM2mStart[nOut]) * sizeof(int);
  hipLaunchKernelGGL(k_m2m<R>, dim3(c->levCount[lev]), dim3(512), lds, st, c->levStart[lev],
                     c->kidStart, c->nKids, c->cubeLevelD, c->ordMomD, c->cubeX, B.mom, c->nmMax, c->T);
}

template <typename R>   // parents at lev into children at lev + 1
static void l2lPass(fbh_ctx *c, const OpBufs<R> &B, int lev, hipStream_t st) {
  const int nIn = NMOM(c->ordMom[lev]), nOut = NMOM(c->ordMom[lev + 1]);
  if (c->levCount[lev + 1] <= 0) return;
  size_t lds = 5 * (size_t)nIn * sizeof(R) + (size_t)(nOut + 1 + c->hConvStart[nOut]) * sizeof(int);
  hipLaunchKernelGGL((k_l2l<R, 4>), dim3(c->levCount[lev + 1]), dim3(blk64(nOut, 512)), lds, st, c->levStart[lev + 1],
                     c->parent, c->cubeLevelD, c->ordMomD, c->cubeX, B.lec, c->nmMax, c->T);
}

template <typename R>
static void l2pPass(fbh_ctx *c, const OpBufs<R> &B, hipStream_t st) {
  launchL2P<R>(c, B.Q0, B.Q1, B.lec, B.far, st);
}

// L2L (every level) and L2P in one launch; the locals in B.lec stay the M2L ones.
// False (nothing launched) when the block's tables do not fit in LDS.
template <typename R>
static size_t l2lpLds(const fbh_ctx *c) {
  const int V = vecWidth<R>(c->nmLeaf), G = c->nmLeaf / V;
  return std::max(4 * (size_t)(c->depth - c->height + 2) * c->nmMax * sizeof(R),
                  2 * (size_t)L2P_CH * G * sizeof(R));
}
template <typename R>
static bool l2lpFits(const fbh_ctx *c) {
  return l2lpLds<R>(c) <= (size_t)c->ldsMax && c->depth - c->height <= L2LP_MAXUP;
}
template <typename R>
static bool l2lpPass(fbh_ctx *c, const OpBufs<R> &B, hipStream_t st) {
  if (!l2lpFits<R>(c)) return false;
  const int V = vecWidth<R>(c->nmLeaf), nmx = c->nmMax, nUp = c->depth - c->height;
  const size_t lds = l2lpLds<R>(c);
  LevOrd O{};
  for (int s = 0; s <= nUp; s++) O.p[s] = c->ordMom[c->height + s];
  if (V == 4)
    hipLaunchKernelGGL((k_l2lp<R, 4>), dim3(c->nLeaf), dim3(256), lds, st, B.Q0, B.Q1, c->leafStart, c->leafCount,
                       c->parent, O, c->cubeX, B.lec, B.far, c->n, c->nmLeaf, nmx, nUp, c->T);
  else if (V == 2)
    hipLaunchKernelGGL((k_l2lp<R, 2>), dim3(c->nLeaf), dim3(256), lds, st, B.Q0, B.Q1, c->leafStart, c->leafCount,
                       c->parent, O, c->cubeX, B.lec, B.far, c->n, c->nmLeaf, nmx, nUp, c->T);
  else
    hipLaunchKernelGGL((k_l2lp<R, 1>), dim3(c->nLeaf), dim3(256), lds, st, B.Q0, B.Q1, c->leafStart, c->leafCount,
                       c->parent, O, c->cubeX, B.lec, B.far, c->n, c->nmLeaf, nmx, nUp, c->T);
  return true;
}

// M2M of one level by axis-wise translation (k_m2m_ax); false when its LDS does not fit
template <typename R>
static bool m2mAxPass(fbh_ctx *c, const OpBufs<R> &B, int lev, hipStream_t st) {
  const size_t lds = (24 * (size_t)c->nmMax * 2 + 8 * 3 * (MAXP + 1)) * sizeof(R);
  if (lds > (size_t)c->ldsMax) return false;
  if (c->levCount[lev] > 0)
    hipLaunchKernelGGL(k_m2m_ax<R>, dim3(c->levCount[lev]), dim3(512), lds, st, c->levStart[lev], c->ordMom[lev],
                       NMOM(c->ordMom[lev + 1]), c->kidStart, c->nKids, c->cubeX, B.mom, c->nmMax, c->T);
  return true;
}

// far field into B.far: Q2M, M2M (finest parent level up), M2L, L2L (down), L2P.
// staged: L2L level by level into B.lec and a separate L2P (the stage ABI reads
// the locals of every cube afterwards); else L2L and L2P in one launch.
template <typename R>
static void farPass(fbh_ctx *c, const OpBufs<R> &B, const double *x, hipStream_t st, bool staged = false) {
  q2mPass(c, B, x, st);
  for (int lev = c->depth - 1; lev >= c->height; lev--)
    if (staged || !m2mAxPass(c, B, lev, st)) m2mPass(c, B, lev, st);
  m2lPass(c, B, st);
  if (!staged && l2lpPass(c, B, st)) return;
  for (int lev = c->height; lev < c->depth; lev++) l2lPass(c, B, lev, st);
  l2pPass(c, B, st);
}

template <typename R>
static void nearPass(fbh_ctx *c, const OpBufs<R> &B, const double *x, hipStream_t st) {
  if (c->nFlyTiles > 0) {
    if constexpr (std::is_same<R, double>::value)
      hipLaunchKernelGGL(k_near_mix64<false>, dim3(c->nearGrid), dim3(256), 0, st, c->tiles, c->leafStart,
                         c->leafCount, c->srcOff, c->ncol, c->tileCoef, c->srcList, B.coef, c->fly64PA, c->fly64PN, x,
                         c->specOff, c->specCol, c->specVal64, B.nearv, c->n, c->kappa, c->eps, c->nTiles,
                         c->tileOrder, c->ownOff);
    else
      hipLaunchKernelGGL(k_near_mix32, dim3(c->nTiles), dim3(256), c->near32Pad, st, c->tiles, c->leafStart, c->leafCount,
                         c->srcOff, c->ncol, c->tileCoef, c->srcList, B.coef, c->flyPA, c->flyPN, c->flyPL, x,
                         c->specOff, c->specCol, c->specVal, B.nearv, c->n, (float)c->kappa, (float)c->eps, c->tileOrder,
                         c->ownOff);
    return;
  }
  hipLaunchKernelGGL(k_near_apply<R>, dim3(c->nTiles), dim3(256), 0, st, c->tiles, c->leafStart, c->leafCount,
                     c->srcOff, c->ncol, c->tileCoef, c->srcList, B.coef, x, B.nearv, c->n);
}

// the dense direct operator: the nearfield kernel over every panel pair
static void directPass(fbh_ctx *c, const double *x, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_near_apply<double>, dim3(c->nTiles), dim3(256), 0, st, c->tiles, c->leafStart, c->leafCount,
                     c->dSrcOff, c->dNcol, c->dTileCoef, c->dSrcList, c->dCoef, x, out, c->n);
}

// y = alpha*(D - K) x + beta*y; the nearfield runs on the side stream beside
// the far-field chain (both read x only) and the combine waits for both
template <typename R>
static int matvecT(fbh_ctx *c, const OpBufs<R> &B, double alpha, const double *x, double beta, double *y) {
  if (c->direct) {   // MtVmain with -r=1/2 (src/fabipb.c:1251-1266): no far field
    if (!std::is_same<R, double>::value) { setErr("the FP32 operator has no direct (-r) form"); return 0; }
    directPass(c, x, c->nearv, c->stream);
    const double s1 = (1.0 + c->eps) / 2.0, s2 = (1.0 + 1.0 / c->eps) / 2.0;
    const int n2 = 2 * c->n;
    hipLaunchKernelGGL(k_combine<double>, dim3((n2 + 255) / 256), dim3(256), 0, c->stream, c->area, x, c->far,
                       c->nearv, y, c->n, alpha, beta, s1, s2, 0, 1);
    HCK(hipGetLastError());
    return 1;
  }
  HCK(hipEventRecord(c->evFork, c->stream));
  hipStream_t ns = c->side;
  HCK(hipStreamWaitEvent(ns, c->evFork, 0));
  const bool timed = c->ntOn && 2 * (size_t)c->ntUsed + 1 < c->ntEv.size();
  if (timed) HCK(hipEventRecord(c->ntEv[2 * c->ntUsed], ns));
  nearPass(c, B, x, ns);
  if (timed) {
    HCK(hipEventRecord(c->ntEv[2 * c->ntUsed + 1], ns));
    c->ntKind[c->ntUsed++] = sizeof(R) == 4;
  }
  HCK(hipEventRecord(c->evJoin, ns));
  farPass(c, B, x, c->stream);
  HCK(hipStreamWaitEvent(c->stream, c->evJoin, 0));
  double s1 = (1.0 + c->eps) / 2.0, s2 = (1.0 + 1.0 / c->eps) / 2.0;
  int n2 = 2 * c->n;
  hipLaunchKernelGGL(k_combine<R>, dim3((n2 + 255) / 256), dim3(256), 0, c->stream, c->area, x, B.far, B.nearv, y,
                     c->n, alpha, beta, s1, s2, 1, 1);
  HCK(hipGetLastError());
  return 1;
}

extern "C" int fbh_matvec_dev(fbh_ctx *c, double alpha, const double *x, double beta, double *y) {
  if (!c || !x || !y) { setErr("fbh_matvec_dev: null argument"); return 0; }
  return matvecT(c, bufs64(c), alpha, x, beta, y);
}

// ---------------------------------------------------------- adjoint operator
// A^T = D - K_near^T - K_far^T.  K_far = L2P L2L M2L M2M Q2M, so
// K_far^T = Q2M^T M2M^T M2L^T L2L^T L2P^T, and each transposed pass is one of
// the forward kernels: L2P^T is a Q2M of (y_n, y) (adjoint locals
// l1 = l2 = Q0 y, l3 = l4 = Q1 y_n), L2L^T is an M2M of those two arrays,
// M2L^T the cached M2L with the transposed combination, M2M^T an L2L of two
// arrays, Q2M^T an L2P from the adjoint moments.
static int ensureAdjoint(fbh_ctx *c) {
  if (c->coefT) return 1;
  if (!c->nearSym || !c->m2lSym) {
    setErr("the adjoint operator needs symmetric near and M2L interaction lists");
    return 0;
  }
  if (!c->m2lG) { setErr("the adjoint operator needs the cached M2L tensors"); return 0; }
  const size_t ne = (size_t)c->nCube * 2 * c->nmMax;
  if (!(c->coefT = dalloc<double4>(c, (size_t)std::max<long long>(c->nStored, 1))) ||
      !(c->adjL = dalloc<double>(c, ne)) || !(c->adjM = dalloc<double>(c, ne)))
    return 0;
  HCK(hipMemsetAsync(c->adjL, 0, ne * sizeof(double), c->stream));
  HCK(hipMemsetAsync(c->adjM, 0, ne * sizeof(double), c->stream));
  // the streamed tiles' transposed blocks, the others' transposed corrections
  if (c->nStored > 0)
    hipLaunchKernelGGL(k_near_build_tiles<true>, dim3(c->nTiles, NEAR_ROWS), dim3(256), 0, c->stream, c->pnl,
                       c->tiles, c->leafStart, c->leafCount, c->srcOff, c->ncol, c->tileCoef, c->srcList, c->coefT,
                       c->kappa, c->eps, c->Q, c->qorder);
  if (c->nFlyTiles > 0) {
    if (!(c->specValT64 = dalloc<double4>(c, (size_t)c->nSpec))) return 0;
    hipLaunchKernelGGL((k_spec_vals<false, true>), dim3(c->n), dim3(64), 0, c->stream, c->pnl, c->specOff,
                       c->specCol, (float4 *)nullptr, c->specValT64, c->kappa, c->eps, c->Q, c->qorder);
  }
  HCK(hipGetLastError());
  HCK(hipStreamSynchronize(c->stream));
  return 1;
}

// the transposed far chain from y into c->far (adjL: adjoint locals, adjM: adjoint moments)
static void farPassT(fbh_ctx *c, const double *y, hipStream_t st) {
  launchQ2M<double>(c, c->Q0, c->Q1, y, c->adjL, 1, st);
  OpBufs<double> B = bufs64(c);
  B.mom = c->adjL;
  for (int lev = c->depth - 1; lev >= c->height; lev--) m2mPass(c, B, lev, st);
  auto launch = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(c->nCube), dim3(64 * c->m2lWaves), c->m2lLds, st, c->m2lOff, c->m2lSrc,
                       c->cubeLevelD, c->ordM2LD, c->ordMomD, (const double *)c->m2lG, (const double *)c->adjL,
                       c->adjM, c->nmMax, c->eps, c->T, c->planTerms, c->planIdx, c->planOff);
  };
  if (m2lGrid<double, true>(c, c->m2lG, c->adjL, c->adjM, st)) {
  } else if (c->m2lMaxT == 16) launch(k_m2l_cached<double, 16, true>);
  else if (c->m2lMaxT == 32) launch(k_m2l_cached<double, 32, true>);
  else launch(k_m2l_cached<double, 0, true>);
  for (int lev = c->height; lev < c->depth; lev++) {
    const int nIn = NMOM(c->ordMom[lev]), nOut = NMOM(c->ordMom[lev + 1]);
    if (c->levCount[lev + 1] <= 0) continue;
    size_t lds = 5 * (size_t)nIn * sizeof(double) + (size_t)(nOut + 1 + c->hConvStart[nOut]) * sizeof(int);
    hipLaunchKernelGGL((k_l2l<double, 2>), dim3(c->levCount[lev + 1]), dim3(blk64(nOut, 512)), lds, st,
                       c->levStart[lev + 1], c->parent, c->cubeLevelD, c->ordMomD, c->cubeX, c->adjM, c->nmMax,
                       c->T);
  }
  launchL2P<double, true>(c, c->Q0, c->Q1, c->adjM, c->far, st);
}

// y = alpha*(D - K)^T x + beta*y
static int matvecTr(fbh_ctx *c, double alpha, const double *x, double beta, double *y) {
  if (c->direct) { setErr("the adjoint operator has no direct (-r) form"); return 0; }
  if (!ensureAdjoint(c)) return 0;
  HCK(hipEventRecord(c->evFork, c->stream));
  HCK(hipStreamWaitEvent(c->side, c->evFork, 0));
  if (c->nFlyTiles > 0)
    hipLaunchKernelGGL(k_near_mix64<true>, dim3(c->nearGrid), dim3(256), 0, c->side, c->tiles, c->leafStart,
                       c->leafCount, c->srcOff, c->ncol, c->tileCoef, c->srcList, c->coefT, c->fly64PA, c->fly64PN, x,
                       c->specOff, c->specCol, c->specValT64, c->nearv, c->n, c->kappa, c->eps, c->nTiles,
                       c->tileOrder, c->ownOff);
  else
    hipLaunchKernelGGL(k_near_apply<double>, dim3(c->nTiles), dim3(256), 0, c->side, c->tiles, c->leafStart,
                       c->leafCount, c->srcOff, c->ncol, c->tileCoef, c->srcList, c->coefT, x, c->nearv, c->n);
  HCK(hipEventRecord(c->evJoin, c->side));
  farPassT(c, x, c->stream);
  HCK(hipStreamWaitEvent(c->stream, c->evJoin, 0));
  const double s1 = (1.0 + c->eps) / 2.0, s2 = (1.0 + 1.0 / c->eps) / 2.0;
  const int n2 = 2 * c->n;
  hipLaunchKernelGGL(k_combine<double>, dim3((n2 + 255) / 256), dim3(256), 0, c->stream, c->area, x, c->far,
                     c->nearv, y, c->n, alpha, beta, s1, s2, 1, 1);
  HCK(hipGetLastError());
  return 1;
}

extern "C" int fbh_matvec_t_dev(fbh_ctx *c, double alpha, const double *x, double beta, double *y) {
  if (!c || !x || !y) { setErr("fbh_matvec_t_dev: null argument"); return 0; }
  return matvecTr(c, alpha, x, beta, y);
}

// the self / shared-vertex pairs of every row's near list (CSR, near-list order)
static int ensureSpec(fbh_ctx *c) {
  if (c->specCol) return 1;
  const size_t n = c->n;
  int *cnt = dalloc<int>(c, n);
  if (!cnt || !(c->specOff = dalloc<long long>(c, n + 1))) return 0;
  hipLaunchKernelGGL(k_spec_count, dim3(n), dim3(256), 0, c->stream, c->pnl, c->leafOf, c->srcOff, c->ncol,
                     c->srcList, cnt);
  HCK(hipGetLastError());
  std::vector<int> hc(n);
  std::vector<long long> ho(n + 1, 0);
  HCK(hipMemcpyAsync(hc.data(), cnt, n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < n; i++) ho[i + 1] = ho[i] + hc[i];
  c->nSpec = ho[n];
  if (!(c->specCol = dalloc<int>(c, (size_t)std::max<long long>(c->nSpec, 1)))) return 0;
  HCK(hipMemcpyAsync(c->specOff, ho.data(), (n + 1) * sizeof(long long), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_spec_index, dim3(n), dim3(256), 0, c->stream, c->pnl, c->leafOf, c->srcOff, c->ncol,
                     c->srcList, c->specOff, c->specCol);
  HCK(hipGetLastError());
  HCK(hipStreamSynchronize(c->stream));
  return 1;
}

// panel data and special-pair corrections of the tiles computed on the fly
static int enableFly64(fbh_ctx *c) {
  const size_t n = c->n;
  if (!(c->fly64PA = dalloc<double4>(c, n)) || !(c->fly64PN = dalloc<double4>(c, n)) || !ensureSpec(c) ||
      !(c->specVal64 = dalloc<double4>(c, (size_t)c->nSpec)))
    return 0;
  hipLaunchKernelGGL(k_fly64_panels, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->pnl, (int)n, c->Q.t[0],
                     c->fly64PA, c->fly64PN);
  hipLaunchKernelGGL(k_exp2_tab, dim3(1), dim3(256), 0, c->stream);
  hipLaunchKernelGGL(k_spec_vals<false>, dim3(n), dim3(64), 0, c->stream, c->pnl, c->specOff, c->specCol,
                     (float4 *)nullptr, c->specVal64, c->kappa, c->eps, c->Q, c->qorder);
  HCK(hipGetLastError());
  HCK(hipStreamSynchronize(c->stream));
  c->fly64 = true;
  return 1;
}

extern "C" int fbh_enable_fp32(fbh_ctx *c) {
  if (!c) { setErr("fbh_enable_fp32: null context"); return 0; }
  if (c->fp32) return 1;
  if (!c->m2lG && c->nM2L > 0) { setErr("fbh_enable_fp32: the FP32 operator needs the cached M2L tensors"); return 0; }
  const size_t n = c->n, nq = n * c->nmLeaf, ne = (size_t)c->nCube * c->nmMax, ng = (size_t)c->nM2L * 2 * c->nmMax;
  if (!(c->Q0f = dalloc<float>(c, nq)) || !(c->Q1f = dalloc<float>(c, nq)) || !(c->momf = dalloc<float>(c, 2 * ne)) ||
      !(c->lecf = dalloc<float>(c, 4 * ne)) || !(c->farf = dalloc<float>(c, 2 * n)) ||
      !(c->nearf = dalloc<float>(c, 2 * n)) || (c->m2lG && !(c->m2lGf = dalloc<float>(c, ng))))
    return 0;
  auto cvt = [&](const double *in, float *out, size_t cnt) {
    hipLaunchKernelGGL(k_to_f32, dim3(2048), dim3(256), 0, c->stream, in, out, (long long)cnt);
  };
  cvt(c->Q0, c->Q0f, nq);
  cvt(c->Q1, c->Q1f, nq);
  if (c->m2lG) cvt(c->m2lG, c->m2lGf, ng);   // no far field (no M2L pairs): the locals stay zero
  // the streamed tiles read FP32 copies of their stored coefficients, the
  // others compute in FP32 on the fly
  c->fly = c->nFlyTiles > 0;
  if (c->nStored > 0) {
    if (!(c->coef32 = dalloc<float4>(c, (size_t)c->nStored))) return 0;
    cvt((const double *)c->coef, (float *)c->coef32, 4 * (size_t)c->nStored);
  }
  if (c->fly) {
    if (!(c->flyPA = dalloc<float4>(c, n)) || !(c->flyPN = dalloc<float4>(c, n)) ||
        !(c->flyPL = dalloc<float4>(c, n)) || !ensureSpec(c) ||
        !(c->specVal = dalloc<float4>(c, (size_t)c->nSpec)))
      return 0;
    hipLaunchKernelGGL(k_fly_panels, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->pnl, (int)n, c->Q.t[0],
                       c->flyPA, c->flyPN, c->flyPL);
    hipLaunchKernelGGL(k_spec_vals32, dim3(n), dim3(64), 0, c->stream, c->pnl, c->flyPA, c->flyPN, c->flyPL,
                       c->leafOf, c->leafStart, c->specOff, c->specCol, c->specVal, c->kappa, c->eps, c->Q,
                       c->qorder);
  }
  HCK(hipMemsetAsync(c->momf, 0, 2 * ne * sizeof(float), c->stream));
  HCK(hipMemsetAsync(c->lecf, 0, 4 * ne * sizeof(float), c->stream));
  HCK(hipGetLastError());
  HCK(hipStreamSynchronize(c->stream));
  {
    const char *ep = getenv("FABIPB_NEAR32_PAD_KB");   // occupancy sweeps: LDS held per block beyond its 24 KiB
    const int kb = ep ? atoi(ep) : 0;
    c->near32Pad = kb > 0 && kb <= 40 ? (size_t)kb * 1024 : 0;
  }
  c->fp32 = true;
  return 1;
}

extern "C" int fbh_matvec32_dev(fbh_ctx *c, double alpha, const double *x, double beta, double *y) {
  if (!c || !x || !y) { setErr("fbh_matvec32_dev: null argument"); return 0; }
  if (!c->fp32) { setErr("fbh_matvec32_dev: call fbh_enable_fp32 first"); return 0; }
  return matvecT(c, bufs32(c), alpha, x, beta, y);
}

extern "C" int fbh_matvec(fbh_ctx *c, double alpha, const double *x, double beta, double *y) {
  if (!c || !x || !y) { setErr("fbh_matvec: null argument"); return 0; }
  size_t b = 2 * (size_t)c->n * sizeof(double);
  HCK(hipMemcpyAsync(c->xs, x, b, hipMemcpyHostToDevice, c->stream));
  if (beta != 0.0) HCK(hipMemcpyAsync(c->ys, y, b, hipMemcpyHostToDevice, c->stream));
  if (!fbh_matvec_dev(c, alpha, c->xs, beta, c->ys)) return 0;
  HCK(hipMemcpyAsync(y, c->ys, b, hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  return 1;
}

static int stageApply(fbh_ctx *c, double alpha, const double *sgm, double beta, double *pot, int useFar,
                      int useNear) {
  if (!c || !sgm || !pot) { setErr("stage apply: null argument"); return 0; }
  size_t b = 2 * (size_t)c->n * sizeof(double);
  HCK(hipMemcpyAsync(c->xs, sgm, b, hipMemcpyHostToDevice, c->stream));
  HCK(hipMemcpyAsync(c->ys, pot, b, hipMemcpyHostToDevice, c->stream));
  if (useFar) farPass(c, bufs64(c), c->xs, c->stream, true);
  if (useNear) nearPass(c, bufs64(c), c->xs, c->stream);
  int n2 = 2 * c->n;
  hipLaunchKernelGGL(k_accum, dim3((n2 + 255) / 256), dim3(256), 0, c->stream, c->far, c->nearv, c->ys, n2, alpha,
                     beta, useFar, useNear);
  HCK(hipGetLastError());
  HCK(hipMemcpyAsync(pot, c->ys, b, hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  return 1;
}

// the dense direct operator (cpuDirectApply, src/direct_backend.c:93-208;
// gpuDirectApply, gpu_backend.h:15): every pair's panelIA0 in one store of
// n*n KER[4] records, rows contiguous and summed in panel order
extern "C" int fbh_set_direct(fbh_ctx *c, int on) {
  if (!c) { setErr("fbh_set_direct: null context"); return 0; }
  if (!on) {
    for (void **q : {(void **)&c->dSrcOff, (void **)&c->dNcol, (void **)&c->dSrcList, (void **)&c->dCoefOff,
                     (void **)&c->dCoef, (void **)&c->dTileCoef}) {
      if (*q) hipFree(*q);
      *q = nullptr;
    }
    c->direct = false;
    return 1;
  }
  if (c->direct) return 1;
  const size_t n = c->n, need = n * n * sizeof(double4);
  size_t freeB = 0, totalB = 0;
  if (hipMemGetInfo(&freeB, &totalB) != hipSuccess || need + ((size_t)1 << 30) > freeB) {
    setErr("fbh_set_direct: the dense operator needs %.1f GB of device memory, %.1f GB free", need / 1e9,
           freeB / 1e9);
    return 0;
  }
  std::vector<int> srcOff(c->nLeaf + 1, 0), ncol(c->nLeaf, c->n), srcList(n);
  std::vector<long long> coefOff(c->nLeaf + 1);
  for (int l = 0; l < c->nLeaf; l++) coefOff[l] = (long long)c->hLeafStart[l] * (long long)n;
  coefOff[c->nLeaf] = (long long)(n * n);
  for (size_t i = 0; i < n; i++) srcList[i] = (int)i;
  std::vector<long long> tileOff;   // the tiles of fbh_create, in its order
  for (int l = 0; l < c->nLeaf; l++)
    for (int r = 0; r < c->hLeafCount[l]; r += NEAR_ROWS) tileOff.push_back(coefOff[l] + (long long)r * (long long)n);
  auto fail = [&](const char *what) {
    setErr("fbh_set_direct: %s", what);
    fbh_set_direct(c, 0);
    return 0;
  };
  if (tileOff.size() != (size_t)c->nTiles) return fail("tile layout mismatch");
  if (hipMalloc(&c->dSrcOff, srcOff.size() * sizeof(int)) != hipSuccess ||
      hipMalloc(&c->dNcol, ncol.size() * sizeof(int)) != hipSuccess ||
      hipMalloc(&c->dSrcList, n * sizeof(int)) != hipSuccess ||
      hipMalloc(&c->dCoefOff, coefOff.size() * sizeof(long long)) != hipSuccess ||
      hipMalloc(&c->dTileCoef, tileOff.size() * sizeof(long long)) != hipSuccess ||
      hipMalloc(&c->dCoef, need) != hipSuccess)
    return fail("device allocation failed");
  if (hipMemcpy(c->dSrcOff, srcOff.data(), srcOff.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->dNcol, ncol.data(), ncol.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->dSrcList, srcList.data(), n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->dCoefOff, coefOff.data(), coefOff.size() * sizeof(long long), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(c->dTileCoef, tileOff.data(), tileOff.size() * sizeof(long long), hipMemcpyHostToDevice) !=
          hipSuccess)
    return fail("upload failed");
  hipLaunchKernelGGL(k_near_build, dim3(c->n), dim3(256), 0, c->stream, c->pnl, c->leafOf, c->leafStart, c->dSrcOff,
                     c->dNcol, c->dCoefOff, c->dSrcList, c->dCoef, c->kappa, c->eps, c->Q, c->qorder);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(hipGetErrorString(e));
  c->direct = true;
  return 1;
}

// pot = beta*pot + alpha*K_direct*sgm on host arrays (gpuDirectApply)
extern "C" int fbh_direct_apply(fbh_ctx *c, double alpha, double beta, const double *sgm, double *pot) {
  if (!c || !sgm || !pot) { setErr("fbh_direct_apply: null argument"); return 0; }
  if (!c->direct && !fbh_set_direct(c, 1)) return 0;
  size_t b = 2 * (size_t)c->n * sizeof(double);
  HCK(hipMemcpyAsync(c->xs, sgm, b, hipMemcpyHostToDevice, c->stream));
  HCK(hipMemcpyAsync(c->ys, pot, b, hipMemcpyHostToDevice, c->stream));
  directPass(c, c->xs, c->nearv, c->stream);
  int n2 = 2 * c->n;
  hipLaunchKernelGGL(k_accum, dim3((n2 + 255) / 256), dim3(256), 0, c->stream, c->far, c->nearv, c->ys, n2, alpha,
                     beta, 0, 1);
  HCK(hipGetLastError());
  HCK(hipMemcpyAsync(pot, c->ys, b, hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  return 1;
}

extern "C" int fbh_near_apply(fbh_ctx *c, double alpha, const double *sgm, double *pot) {
  return stageApply(c, alpha, sgm, 1.0, pot, 0, 1);
}
extern "C" int fbh_far_apply(fbh_ctx *c, double alpha, const double *sgm, double beta, double *pot) {
  return stageApply(c, alpha, sgm, beta, pot, 1, 0);
}

extern "C" int fbh_expansions(fbh_ctx *c, double *mom, double *lec) {
  if (!c) return 0;
  HCK(hipStreamSynchronize(c->stream));
  if (mom) HCK(hipMemcpy(mom, c->mom, (size_t)c->nCube * 2 * c->nmMax * sizeof(double), hipMemcpyDeviceToHost));
  if (lec) HCK(hipMemcpy(lec, c->lec, (size_t)c->nCube * 4 * c->nmMax * sizeof(double), hipMemcpyDeviceToHost));
  return 1;
}

extern "C" int fbh_q2m_matrices(fbh_ctx *c, double *q0, double *q1) {
  if (!c) return 0;
  HCK(hipStreamSynchronize(c->stream));
  size_t b = (size_t)c->n * c->nmLeaf * sizeof(double);
  if (q0) HCK(hipMemcpy(q0, c->Q0, b, hipMemcpyDeviceToHost));
  if (q1) HCK(hipMemcpy(q1, c->Q1, b, hipMemcpyDeviceToHost));
  return 1;
}

extern "C" int fbh_near_coeff(fbh_ctx *c, int x, int y, double *ker4) {
  if (!c || x < 0 || x >= c->n || y < 0 || y >= c->n) { setErr("fbh_near_coeff: bad index"); return 0; }
  HCK(hipStreamSynchronize(c->stream));
  int D = 0;
  std::vector<int> ls(c->nLeaf), lc(c->nLeaf), so(c->nLeaf + 1), nc(c->nLeaf);
  HCK(hipMemcpy(ls.data(), c->leafStart, c->nLeaf * sizeof(int), hipMemcpyDeviceToHost));
  HCK(hipMemcpy(lc.data(), c->leafCount, c->nLeaf * sizeof(int), hipMemcpyDeviceToHost));
  HCK(hipMemcpy(so.data(), c->srcOff, (c->nLeaf + 1) * sizeof(int), hipMemcpyDeviceToHost));
  HCK(hipMemcpy(nc.data(), c->ncol, c->nLeaf * sizeof(int), hipMemcpyDeviceToHost));
  while (D < c->nLeaf && !(x >= ls[D] && x < ls[D] + lc[D])) D++;
  std::vector<int> sl(nc[D]);
  HCK(hipMemcpy(sl.data(), c->srcList + so[D], nc[D] * sizeof(int), hipMemcpyDeviceToHost));
  for (int j = 0; j < nc[D]; j++)
    if (sl[j] == y) {
      double *d = nullptr;
      HCK(hipMalloc(&d, 4 * sizeof(double)));
      hipLaunchKernelGGL(k_near_pair, dim3(1), dim3(1), 0, c->stream, c->pnl, x, y, c->kappa, c->eps, c->Q,
                         c->qorder, d);
      hipError_t e = hipMemcpyAsync(ker4, d, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      hipFree(d);
      if (e != hipSuccess) { setErr("fbh_near_coeff: %s", hipGetErrorString(e)); return 0; }
      return 1;
    }
  setErr("fbh_near_coeff: panels %d and %d are not near neighbours", x, y);
  return 0;
}

// panelIA0 of listed pairs of the panels pnl[0..nPanels) (device), host index
// and result arrays, in chunks; the device buffers live for the call only
static int pairCoeffs(const Pnl *pnl, int nPanels, long long nPairs, const int *xi, const int *yi, double kappa,
                      double eps, const QTab &Q, int q, double *k0, double *k1, double *k2, double *k3,
                      hipStream_t st) {
  for (long long p = 0; p < nPairs; p++)
    if (xi[p] < 0 || xi[p] >= nPanels || yi[p] < 0 || yi[p] >= nPanels) {
      setErr("pair %lld: panel index outside 0..%d", p, nPanels - 1);
      return 0;
    }
  const long long chunk = std::min<long long>(nPairs, 1 << 22);
  int *dx = nullptr, *dy = nullptr;
  double *dk = nullptr;
  bool ok = hipMalloc(&dx, chunk * sizeof(int)) == hipSuccess && hipMalloc(&dy, chunk * sizeof(int)) == hipSuccess &&
            hipMalloc(&dk, 4 * chunk * sizeof(double)) == hipSuccess;
  double *out[4] = {k0, k1, k2, k3};
  for (long long p0 = 0; ok && p0 < nPairs; p0 += chunk) {
    const long long m = std::min(chunk, nPairs - p0);
    ok = hipMemcpyAsync(dx, xi + p0, m * sizeof(int), hipMemcpyHostToDevice, st) == hipSuccess &&
         hipMemcpyAsync(dy, yi + p0, m * sizeof(int), hipMemcpyHostToDevice, st) == hipSuccess;
    if (!ok) break;
    hipLaunchKernelGGL(k_pair_coeffs, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, pnl, m, dx, dy, kappa, eps,
                       Q, q, dk, dk + chunk, dk + 2 * chunk, dk + 3 * chunk);
    ok = hipGetLastError() == hipSuccess;
    for (int i = 0; ok && i < 4; i++)
      ok = hipMemcpyAsync(out[i] + p0, dk + i * chunk, m * sizeof(double), hipMemcpyDeviceToHost, st) == hipSuccess;
    if (ok) ok = hipStreamSynchronize(st) == hipSuccess;
  }
  if (!ok) setErr("pair coefficients: %s", hipGetErrorString(hipGetLastError()));
  hipStreamSynchronize(st);
  if (dx) hipFree(dx);
  if (dy) hipFree(dy);
  if (dk) hipFree(dk);
  return ok ? 1 : 0;
}

extern "C" int fbh_pair_coeffs(fbh_ctx *c, long long nPairs, const int *xIdx, const int *yIdx, double *k0,
                               double *k1, double *k2, double *k3) {
  if (!c || nPairs < 0 || (nPairs > 0 && (!xIdx || !yIdx || !k0 || !k1 || !k2 || !k3))) {
    setErr("fbh_pair_coeffs: bad arguments");
    return 0;
  }
  if (nPairs == 0) return 1;
  HCK(hipStreamSynchronize(c->stream));
  return pairCoeffs(c->pnl, c->n, nPairs, xIdx, yIdx, c->kappa, c->eps, c->Q, c->qorder, k0, k1, k2, k3, c->stream);
}

extern "C" int fbh_panel_pair_coeffs(int nPanels, const double *vtx, const double *nrm, const double *area,
                                     double kappa, double eps, int qorder, long long nPairs, const int *xLoc,
                                     const int *yLoc, double *k0, double *k1, double *k2, double *k3) {
  if (nPanels <= 0 || !vtx || !nrm || !area || qorder < 1 || qorder > 10 || !(eps > 0) || nPairs < 0 ||
      (nPairs > 0 && (!xLoc || !yLoc || !k0 || !k1 || !k2 || !k3))) {
    setErr("fbh_panel_pair_coeffs: bad arguments");
    return 0;
  }
  if (nPairs == 0) return 1;
  if (!fbh_available()) return 0;
  std::vector<Pnl> hp(nPanels);
  for (int i = 0; i < nPanels; i++) {
    const double *v = &vtx[9 * i];
    for (int k = 0; k < 9; k++) hp[i].v[k] = v[k];
    for (int j = 0; j < 3; j++) {   // src/input.c:743-747
      hp[i].a[j] = v[6 + j] - v[3 + j];
      hp[i].a[3 + j] = v[j] - v[6 + j];
      hp[i].a[6 + j] = v[3 + j] - v[j];
    }
    for (int k = 0; k < 3; k++) hp[i].nr[k] = nrm[3 * i + k];
    hp[i].area = area[i];
  }
  QTab Q;
  for (int i = 0; i < 10; i++) { Q.t[i] = hT[qorder][i]; Q.w[i] = hW[qorder][i]; }
  Pnl *dp = nullptr;
  if (hipMalloc(&dp, nPanels * sizeof(Pnl)) != hipSuccess ||
      hipMemcpy(dp, hp.data(), nPanels * sizeof(Pnl), hipMemcpyHostToDevice) != hipSuccess) {
    if (dp) hipFree(dp);
    setErr("fbh_panel_pair_coeffs: uploading %d panels failed", nPanels);
    return 0;
  }
  const int ok = pairCoeffs(dp, nPanels, nPairs, xLoc, yLoc, kappa, eps, Q, qorder, k0, k1, k2, k3, nullptr);
  hipFree(dp);
  return ok;
}

extern "C" int fbh_psolve_diag_dev(fbh_ctx *c, double *out, const double *in) {
  int n2 = 2 * c->n;
  hipLaunchKernelGGL(k_psolve, dim3((n2 + 255) / 256), dim3(256), 0, c->stream, in, out, c->area, c->n,
                     (1.0 + c->eps) / 2.0, (1.0 + 1.0 / c->eps) / 2.0);
  HCK(hipGetLastError());
  return 1;
}

static void freePB(fbh_ctx *c) {
  for (void **q : {(void **)&c->pbStart, (void **)&c->pbM, (void **)&c->pbOff, (void **)&c->pbTiles,
                   (void **)&c->pbInv, (void **)&c->pbTmp, (void **)&c->pbInvT}) {
    if (*q) hipFree(*q);
    *q = nullptr;
  }
  c->nPB = c->nPBTiles = 0;
  c->pbBytes = 0;
}

// the preconditioner of GMRES (PtVmain, src/fabipb.c:1289-1309): 3 is the
// diagonal; 0, 1 and 2 are the reference's three ways to apply the same
// leaf-parent block LU (recomputed, cached blocks, cached factors), here built
// once as explicit block inverses: batched LU and inverse per group of blocks
// padded to the same size (rocSOLVER getrf/getri), then one HBM-bound pass
// over the packed inverses per application
extern "C" int fbh_set_precond(fbh_ctx *c, int mode) {
  if (!c) { setErr("fbh_set_precond: null context"); return 0; }
  if (mode == 3 || mode == -1) { freePB(c); c->precMode = mode; return 1; }
  if (mode < -1 || mode > 3) { setErr("fbh_set_precond: mode %d not in -1..3", mode); return 0; }
  if (c->precMode == 2 && c->pbInv) return 1;
  const int lev = c->depth - 1;
  if (lev < c->height) {
    setErr("fbh_set_precond: the block preconditioner needs a level depth-1 >= height (depth %d, height %d)",
           c->depth, c->height);
    return 0;
  }
  freePB(c);
  const int nb = c->levCount[lev], lv0 = c->levStart[lev];
  std::vector<int> st(nb), mm(nb);
  std::vector<long long> off(nb + 1, 0);
  std::vector<int2> tiles;
  for (int b = 0; b < nb; b++) {
    const int cb = lv0 + b, k0 = c->hKidStart[cb];
    st[b] = c->hLeafStart[k0];
    mm[b] = 0;
    for (int k = 0; k < c->hNKids[cb]; k++) mm[b] += c->hLeafCount[k0 + k];
    const int M = 2 * mm[b];
    off[b + 1] = off[b] + (long long)M * M;
    for (int r = 0; r < M; r += 16) tiles.push_back(make_int2(b, r));
  }
  // groups of blocks padded to the same size (multiples of 64)
  std::vector<std::vector<int>> groups;
  std::vector<int> gMp;
  {
    std::vector<int> order(nb);
    for (int b = 0; b < nb; b++) order[b] = b;
    auto pad = [&](int b) { return ((2 * mm[b] + 63) / 64) * 64; };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return pad(x) < pad(y); });
    for (int b : order) {
      if (gMp.empty() || gMp.back() != pad(b)) { gMp.push_back(pad(b)); groups.emplace_back(); }
      groups.back().push_back(b);
    }
  }
  size_t tmpMax = 0;
  int cntMax = 0;
  for (size_t g = 0; g < groups.size(); g++) {
    tmpMax = std::max(tmpMax, groups[g].size() * (size_t)gMp[g] * gMp[g]);
    cntMax = std::max(cntMax, (int)groups[g].size());
  }
  double *T = nullptr;
  int *ipiv = nullptr, *info = nullptr, *grp = nullptr;
  const size_t n2 = 2 * (size_t)c->n;
  auto fail = [&](const char *what) {
    setErr("fbh_set_precond: %s", what);
    if (T) hipFree(T);
    if (ipiv) hipFree(ipiv);
    if (info) hipFree(info);
    if (grp) hipFree(grp);
    freePB(c);
    return 0;
  };
  if (hipMalloc(&c->pbStart, nb * sizeof(int)) != hipSuccess || hipMalloc(&c->pbM, nb * sizeof(int)) != hipSuccess ||
      hipMalloc(&c->pbOff, (nb + 1) * sizeof(long long)) != hipSuccess ||
      hipMalloc(&c->pbTiles, tiles.size() * sizeof(int2)) != hipSuccess ||
      hipMalloc(&c->pbInv, std::max<long long>(off[nb], 1) * sizeof(double)) != hipSuccess ||
      hipMalloc(&c->pbTmp, n2 * sizeof(double)) != hipSuccess || hipMalloc(&T, tmpMax * sizeof(double)) != hipSuccess ||
      hipMalloc(&ipiv, (size_t)cntMax * gMp.back() * sizeof(int)) != hipSuccess ||
      hipMalloc(&info, cntMax * sizeof(int)) != hipSuccess || hipMalloc(&grp, nb * sizeof(int)) != hipSuccess)
    return fail("device allocation failed");
  c->pbBytes = off[nb] * sizeof(double);
  if (hipMemcpy(c->pbStart, st.data(), nb * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->pbM, mm.data(), nb * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->pbOff, off.data(), (nb + 1) * sizeof(long long), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->pbTiles, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess)
    return fail("upload failed");
  if (!c->blas && rocblas_create_handle(&c->blas) != rocblas_status_success) return fail("rocblas handle");
  if (rocblas_set_stream(c->blas, c->stream) != rocblas_status_success) return fail("rocblas stream");
  const double s1 = (1.0 + c->eps) / 2.0, s2 = (1.0 + 1.0 / c->eps) / 2.0;
  std::vector<int> hinfo(cntMax);
  for (size_t g = 0; g < groups.size(); g++) {
    const int cnt = (int)groups[g].size(), Mp = gMp[g];
    int mMax = 0;
    for (int b : groups[g]) mMax = std::max(mMax, mm[b]);
    if (hipMemcpyAsync(grp, groups[g].data(), cnt * sizeof(int), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemsetAsync(T, 0, (size_t)cnt * Mp * Mp * sizeof(double), c->stream) != hipSuccess)
      return fail("group staging failed");
    hipLaunchKernelGGL(k_bp_pad, dim3(cnt), dim3(64), 0, c->stream, grp, c->pbM, Mp, T);
    hipLaunchKernelGGL(k_bp_assemble, dim3(cnt, mMax), dim3(256), 0, c->stream, c->pnl, grp, c->pbStart, c->pbM, Mp,
                       T, c->kappa, c->eps, s1, s2, c->Q, c->qorder);
    if (rocsolver_dgetrf_strided_batched(c->blas, Mp, Mp, T, Mp, (rocblas_stride)Mp * Mp, ipiv, Mp, info, cnt) !=
            rocblas_status_success ||
        rocsolver_dgetri_strided_batched(c->blas, Mp, T, Mp, (rocblas_stride)Mp * Mp, ipiv, Mp, info, cnt) !=
            rocblas_status_success)
      return fail("batched LU/inverse failed");
    if (hipMemcpyAsync(hinfo.data(), info, cnt * sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      return fail("LU status read failed");
    for (int k = 0; k < cnt; k++)
      if (hinfo[k] != 0) {
        setErr("fbh_set_precond: block %d is singular (info %d)", groups[g][k], hinfo[k]);
        return fail("singular block");
      }
    hipLaunchKernelGGL(k_bp_unpad, dim3(cnt, 64), dim3(256), 0, c->stream, grp, c->pbM, c->pbOff, Mp, T, c->pbInv);
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(T);
  hipFree(ipiv);
  hipFree(info);
  hipFree(grp);
  T = nullptr; ipiv = info = grp = nullptr;
  if (e != hipSuccess) return fail(hipGetErrorString(e));
  c->nPB = nb;
  c->nPBTiles = (int)tiles.size();
  c->precMode = 2;
  return 1;
}

// the current preconditioner; out may equal in
extern "C" int fbh_psolve_dev(fbh_ctx *c, double *out, const double *in) {
  if (!c || !out || !in) { setErr("fbh_psolve_dev: null argument"); return 0; }
  if (c->precMode == -1) {   // PtVmain with the preconditioner disabled: a copy
    if (out != in) HCK(hipMemcpyAsync(out, in, 2 * (size_t)c->n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return 1;
  }
  if (c->precMode != 2) return fbh_psolve_diag_dev(c, out, in);
  HCK(hipMemcpyAsync(c->pbTmp, in, 2 * (size_t)c->n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(k_bp_apply, dim3(c->nPBTiles), dim3(256), 0, c->stream, c->pbTiles, c->pbStart, c->pbM, c->pbOff,
                     c->pbInv, c->pbTmp, out, c->n);
  HCK(hipGetLastError());
  return 1;
}

static int ensureGmres(fbh_ctx *c, int restart) {
  if (c->V && c->gmRestart >= restart) return 1;
  if (c->V) { hipFree(c->V); c->V = nullptr; }
  size_t n2 = 2 * (size_t)c->n;
  HCK(hipMalloc(&c->V, n2 * (restart + 1) * sizeof(double)));
  if (!c->w) HCK(hipMalloc(&c->w, n2 * sizeof(double)));
  if (!c->r) HCK(hipMalloc(&c->r, n2 * sizeof(double)));
  if (!c->part) HCK(hipMalloc(&c->part, 2 * RED_BLOCKS * sizeof(double)));   // two sets (ping-pong)
  if (c->hcol) hipFree(c->hcol);
  if (c->ydev) hipFree(c->ydev);
  HCK(hipMalloc(&c->hcol, (restart + 2) * sizeof(double)));
  HCK(hipMalloc(&c->ydev, (restart + 2) * sizeof(double)));
  if (!c->scal) HCK(hipMalloc(&c->scal, 4 * sizeof(double)));
  if (c->hpinLen < restart + 8) {
    if (c->hpin) hipHostFree(c->hpin);
    c->hpin = nullptr;
    c->hpinLen = 0;
    HCK(hipHostMalloc((void **)&c->hpin, (restart + 8) * sizeof(double), hipHostMallocDefault));
    c->hpinLen = restart + 8;
  }
  if (!c->evDone) HCK(hipEventCreateWithFlags(&c->evDone, hipEventDisableTiming));
  c->gmRestart = restart;
  return 1;
}

// waits for the context stream by spinning on an event recorded at its end:
// the host round trip of every GMRES step (readback, Givens rotation, next
// launch) costs a sleeping wait's wake-up latency otherwise
static int streamSpin(fbh_ctx *c) {
  HCK(hipEventRecord(c->evDone, c->stream));
  hipError_t e;
  while ((e = hipEventQuery(c->evDone)) == hipErrorNotReady) {
  }
  if (e != hipSuccess) { setErr("device work failed: %s", hipGetErrorString(e)); return 0; }
  return 1;
}

// 2-norms of k <= 4 device vectors of 2n: k deterministic reductions (as
// k_dot_part + k_sum_parts do for one), one pinned readback, one wait
static int devNorms(fbh_ctx *c, int k, const double *const *v, double *out) {
  const int n2 = 2 * c->n;
  for (int q = 0; q < k; q++) {
    hipLaunchKernelGGL(k_dot_part, dim3(RED_BLOCKS), dim3(256), 0, c->stream, v[q], v[q], n2, c->part);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, c->stream, c->part, RED_BLOCKS, c->scal + q);
  }
  HCK(hipMemcpyAsync(c->hpin, c->scal, k * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (!streamSpin(c)) return 0;
  for (int q = 0; q < k; q++) out[q] = sqrt(c->hpin[q]);
  return 1;
}
static int devNorm(fbh_ctx *c, const double *v, double *out) { return devNorms(c, 1, &v, out); }

// BLAS drotg / drot as used by src/gmres.c:280-293
static void rotg(double a, double b, double *cs, double *sn) {
  double roe = fabs(a) > fabs(b) ? a : b, scale = fabs(a) + fabs(b);
  if (scale == 0.0) { *cs = 1.0; *sn = 0.0; return; }
  double r = scale * sqrt((a / scale) * (a / scale) + (b / scale) * (b / scale));
  if (roe < 0) r = -r;
  *cs = a / r; *sn = b / r;
}
static void rot(double *x, double *y, double c, double s) {
  double t = c * *x + s * *y;
  *y = c * *y - s * *x;
  *x = t;
}

enum OpKind { OP64 = 0, OP32 = 1, OPT64 = 2 };   // FP64 A, FP32 A, FP64 A^T

static int applyOp(fbh_ctx *c, int kind, double alpha, const double *x, double beta, double *y) {
  if (kind == OPT64) return matvecTr(c, alpha, x, beta, y);
  return kind == OP32 ? matvecT(c, bufs32(c), alpha, x, beta, y) : matvecT(c, bufs64(c), alpha, x, beta, y);
}

// the preconditioner of the operator kind: P for A, P^T for A^T
static int precApply(fbh_ctx *c, int kind, double *out, const double *in) {
  if (kind != OPT64 || c->precMode != 2) return fbh_psolve_dev(c, out, in);   // none and diagonal: symmetric
  if (!c->pbInvT) {
    long long tot = 0;
    HCK(hipMemcpy(&tot, c->pbOff + c->nPB, sizeof(long long), hipMemcpyDeviceToHost));
    HCK(hipMalloc(&c->pbInvT, std::max<long long>(tot, 1) * sizeof(double)));
    hipLaunchKernelGGL(k_bp_transpose, dim3(c->nPB, 64), dim3(256), 0, c->stream, c->pbM, c->pbOff, c->pbInv,
                       c->pbInvT);
    HCK(hipGetLastError());
  }
  HCK(hipMemcpyAsync(c->pbTmp, in, 2 * (size_t)c->n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(k_bp_apply, dim3(c->nPBTiles), dim3(256), 0, c->stream, c->pbTiles, c->pbStart, c->pbM, c->pbOff,
                     c->pbInvT, c->pbTmp, out, c->n);
  HCK(hipGetLastError());
  return 1;
}

// restarted left-preconditioned GMRES on the FP64, the FP32 or the transposed
// operator; the Krylov basis, Gram-Schmidt and Givens rotations are FP64 in all.
// bKnown >= 0: ||b|| is already known (the refinement's residual norm);
// xZero: x is known to be zero on entry (no initial product).  One host wait
// per Arnoldi step (the Hessenberg column) and per restart.
static int gmresCore(fbh_ctx *c, int kind, const double *b, double *x, int restrt, double tol, int maxit,
                     int *iters, double *resid, double *hist, int *info, double bKnown = -1.0, bool xZero = false) {
  if (!ensureGmres(c, restrt)) return 0;
  const int n2 = 2 * c->n, ldh = restrt + 1, g = (n2 + 255) / 256;
  const size_t vb = (size_t)n2 * sizeof(double);
  std::vector<double> H((size_t)ldh * (restrt + 1), 0.0), s(restrt + 2), cs(restrt + 1), sn(restrt + 1),
      y(restrt + 1);
  int nh = 0;
  *info = 0;
  *iters = 0;
  double bnrm2, rn;
  if (!xZero) {   // x != 0 needs the initial residual product
    double xn;
    if (!devNorm(c, x, &xn)) return 0;
    xZero = xn == 0.0;
  }
  HCK(hipMemcpyAsync(c->w, b, vb, hipMemcpyDeviceToDevice, c->stream));
  if (!xZero && !applyOp(c, kind, -1.0, x, 1.0, c->w)) return 0;
  if (!precApply(c, kind, c->r, c->w)) return 0;
  if (bKnown >= 0.0) {
    bnrm2 = bKnown;
    if (!devNorm(c, c->r, &rn)) return 0;
  } else {
    const double *v2[2] = {b, c->r};
    double nn[2];
    if (!devNorms(c, 2, v2, nn)) return 0;
    bnrm2 = nn[0];
    rn = nn[1];
  }
  if (bnrm2 == 0.0) bnrm2 = 1.0;
  *resid = rn / bnrm2;
  if (*resid < tol) return 1;
  for (;;) {   // rn = ||r|| of this cycle's start
    int i = 0;
    if (!(rn > 0.0) || rn != rn) {
      *resid = rn != rn ? rn : 0.0;
      *info = rn != rn ? 2 : 0;
      return 1;
    }
    hipLaunchKernelGGL(k_scale, dim3(g), dim3(256), 0, c->stream, c->r, c->V, n2, 1.0 / rn);
    std::fill(s.begin(), s.end(), 0.0);
    s[0] = rn;
    for (;;) {
      ++i;
      ++*iters;
      double *vi = c->V + (size_t)(i - 1) * n2, *vn = c->V + (size_t)i * n2;
      if (!applyOp(c, kind, 1.0, vi, 0.0, c->w)) return 0;
      // modified Gram-Schmidt: <w, v0> (with the diagonal preconditioner applied
      // in the same pass), then per k the update fused with the next dot product
      double *pa = c->part, *pb = c->part + RED_BLOCKS;
      if (c->precMode == 3) {   // the diagonal, as fbh_psolve_diag_dev
        hipLaunchKernelGGL(k_psolve_dot, dim3(RED_BLOCKS), dim3(256), 0, c->stream, c->w, c->V, c->area, c->n,
                           (1.0 + c->eps) / 2.0, (1.0 + 1.0 / c->eps) / 2.0, pa);
      } else {
        if (!precApply(c, kind, c->w, c->w)) return 0;
        hipLaunchKernelGGL(k_dot_part, dim3(RED_BLOCKS), dim3(256), 0, c->stream, c->w, c->V, n2, pa);
      }
      for (int k = 0; k < i; k++) {
        const double *vk = c->V + (size_t)k * n2, *vnext = k + 1 < i ? vk + n2 : c->w;
        hipLaunchKernelGGL(k_mgs_dot, dim3(RED_BLOCKS), dim3(256), 0, c->stream, c->w, vk, vnext, n2, pa,
                           c->hcol + k, pb);
        std::swap(pa, pb);
      }
      hipLaunchKernelGGL(k_normalize, dim3(RED_BLOCKS), dim3(256), 0, c->stream, c->w, vn, n2, pa, c->hcol + i);
      HCK(hipGetLastError());
      HCK(hipMemcpyAsync(c->hpin, c->hcol, (i + 1) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      if (!streamSpin(c)) return 0;
      double *h = &H[(size_t)(i - 1) * ldh];
      for (int k = 0; k <= i; k++) h[k] = c->hpin[k];
      for (int k = 0; k < i - 1; k++) rot(&h[k], &h[k + 1], cs[k], sn[k]);
      rotg(h[i - 1], h[i], &cs[i - 1], &sn[i - 1]);
      rot(&h[i - 1], &h[i], cs[i - 1], sn[i - 1]);
      rot(&s[i - 1], &s[i], cs[i - 1], sn[i - 1]);
      *resid = fabs(s[i]) / bnrm2;
      if (hist) hist[nh++] = *resid;
      if (*resid <= tol || *iters == maxit || i >= restrt) break;
    }
    for (int k = i - 1; k >= 0; k--) {   // y = H \ s (dtrsv upper)
      double t = s[k];
      for (int q = k + 1; q < i; q++) t -= H[(size_t)q * ldh + k] * y[q];
      y[k] = t / H[(size_t)k * ldh + k];
    }
    for (int k = 0; k < i; k++) c->hpin[k] = y[k];
    HCK(hipMemcpyAsync(c->ydev, c->hpin, i * sizeof(double), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_update, dim3(g), dim3(256), 0, c->stream, x, c->V, c->ydev, i, n2);
    if (*resid <= tol) return streamSpin(c);
    HCK(hipMemcpyAsync(c->w, b, vb, hipMemcpyDeviceToDevice, c->stream));
    if (!applyOp(c, kind, -1.0, x, 1.0, c->w)) return 0;
    if (!precApply(c, kind, c->r, c->w)) return 0;
    if (!devNorm(c, c->r, &rn)) return 0;   // also waits for the update's copy out of hpin
    *resid = rn / bnrm2;
    if (hist) hist[nh++] = *resid;
    if (*resid <= tol) return 1;
    if (*iters == maxit) { *info = 1; return 1; }
  }
}

extern "C" int fbh_gmres_dev(fbh_ctx *c, const double *b, double *x, int restrt, double tol, int maxit, int *iters,
                             double *resid, double *hist, int *info) {
  if (!c || !b || !x || restrt < 1 || maxit < 1) { setErr("fbh_gmres_dev: bad arguments"); return 0; }
  return gmresCore(c, OP64, b, x, restrt, tol, maxit, iters, resid, hist, info);
}

// adjoint solve A^T x = b (left-preconditioned with P^T), same stopping rule
extern "C" int fbh_gmres_t_dev(fbh_ctx *c, const double *b, double *x, int restrt, double tol, int maxit,
                               int *iters, double *resid, double *hist, int *info) {
  if (!c || !b || !x || restrt < 1 || maxit < 1) { setErr("fbh_gmres_t_dev: bad arguments"); return 0; }
  if (!ensureAdjoint(c)) return 0;
  return gmresCore(c, OPT64, b, x, restrt, tol, maxit, iters, resid, hist, info);
}

__global__ void k_fill(double *__restrict__ y, int n, double v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = v;
}

__global__ void k_axpy(double *__restrict__ y, const double *__restrict__ x, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += x[i];
}

// Precision form: iterative refinement around GMRES on the FP32 operator.
// Each correction solves A32 d = r with d0 = 0 to a relative residual of
// 0.2 tol measured against ||b||; x += d; convergence is decided on
// r = b - A64 x in FP64 with the same test as the plain solve,
// ||M^-1 r|| / ||b|| <= tol.  The first residual (x = x0) is taken with the
// FP32 operator.  After a small correction (an FP64 residual already at hand
// and within 100 tol ||b||) the residual is updated as r - A32 d, and a pass
// of the test on that update is confirmed with b - A64 x before returning.
// iters counts inner iterations (FP32 products), outer the corrections, mv64
// the FP64 products.
static int walkCharges(fbh_ctx *c, const ChargeSet &s, int mode, double theta, const double *sgm_dev, double *b_dev,
                       double *energy_raw, bool f32, hipStream_t st);
static hipStream_t sharedQueueStream(int which);

// energyRaw: also the precision form's energy of the resident charges for the
// solution (theta the walk's MAC): each correction's walk runs on a second
// stream, from a second host thread, beside the FP64 residual product that
// decides whether it is the last; the walk of a correction that is not the
// last is discarded
static int refinedSolve(fbh_ctx *c, const double *b, double *x, int restrt, double tol, int maxit, int maxOuter,
                        int *iters, double *resid, double *hist, int *info, int *outer, int *mv64, double theta,
                        double *energyRaw) {
  if (!c || !b || !x || restrt < 1 || maxit < 1 || maxOuter < 1) {
    setErr("fbh_gmres_refined_dev: bad arguments");
    return 0;
  }
  if (energyRaw) {
    if (!c->csRes.ready) { setErr("fbh_gmres_refined_energy_dev: no resident charge set (fbh_set_charges)"); return 0; }
    if (!c->auxStream && !(c->auxStream = sharedQueueStream(1))) { setErr("no second stream"); return 0; }
    if (!c->evX) HCK(hipEventCreateWithFlags(&c->evX, hipEventDisableTiming));
  }
  if (!c->fp32 && !fbh_enable_fp32(c)) return 0;
  if (!ensureGmres(c, restrt)) return 0;
  const int n2 = 2 * c->n, g = (n2 + 255) / 256;
  const size_t vb = (size_t)n2 * sizeof(double);
  if (!c->ro) HCK(hipMalloc(&c->ro, vb));
  if (!c->dx) HCK(hipMalloc(&c->dx, vb));
  *iters = 0;
  *outer = 0;
  *mv64 = 0;
  *info = 0;
  int nh = 0;
  double bnrm2, rr;
  {
    const double *v2[2] = {b, x};
    double nn[2];
    if (!devNorms(c, 2, v2, nn)) return 0;
    bnrm2 = nn[0];
    HCK(hipMemcpyAsync(c->ro, b, vb, hipMemcpyDeviceToDevice, c->stream));
    if (nn[1] != 0.0) {
      if (!applyOp(c, OP32, -1.0, x, 1.0, c->ro)) return 0;
      if (!devNorm(c, c->ro, &rr)) return 0;
    } else {
      rr = bnrm2;   // x = 0: ro = b
    }
  }
  if (bnrm2 == 0.0) bnrm2 = 1.0;
  const double braw = bnrm2;
  // the FP64 residual of the current x: ro = b - A64 x, w = M^-1 ro, and their norms
  auto residual64 = [&](double *rn) {
    HCK(hipMemcpyAsync(c->ro, b, vb, hipMemcpyDeviceToDevice, c->stream));
    if (!applyOp(c, OP64, -1.0, x, 1.0, c->ro)) return 0;
    ++*mv64;
    if (!fbh_psolve_dev(c, c->w, c->ro)) return 0;
    const double *v2[2] = {c->w, c->ro};
    double nn[2];
    if (!devNorms(c, 2, v2, nn)) return 0;
    *rn = nn[0];
    rr = nn[1];
    return 1;
  };
  for (;;) {
    if (!(rr > 0.0)) break;
    hipLaunchKernelGGL(k_fill, dim3(g), dim3(256), 0, c->stream, c->dx, n2, 0.0);
    int it = 0, inf = 0;
    double ir;
    // inner target 0.2 tol: the FP64 check then passes after one correction with a
    // residual well inside tol, which keeps the energy within ~1e-7 of the converged one
    if (!gmresCore(c, OP32, c->ro, c->dx, restrt, 0.2 * tol * bnrm2 / rr, maxit - *iters, &it, &ir, nullptr, &inf,
                   rr, true))
      return 0;
    *iters += it;
    ++*outer;
    hipLaunchKernelGGL(k_axpy, dim3(g), dim3(256), 0, c->stream, x, c->dx, n2);
    std::thread walk;
    int okWalk = 1;
    double eWalk = 0.0;
    struct Join {
      std::thread &t;
      ~Join() {
        if (t.joinable()) t.join();
      }
    } joinWalk{walk};
    if (energyRaw) {   // the walk of this x beside its FP64 residual
      HCK(hipEventRecord(c->evX, c->stream));
      walk = std::thread([&] {
        if (hipSetDevice(c->dev) != hipSuccess || hipStreamWaitEvent(c->auxStream, c->evX, 0) != hipSuccess) {
          okWalk = 0;
          return;
        }
        okWalk = walkCharges(c, c->csRes, 0, theta, x, nullptr, &eWalk, true, c->auxStream);
      });
    }
    // x is final for this correction when the loop leaves: the walk's energy is x's
    auto leave = [&](int rc) {
      if (walk.joinable()) walk.join();
      if (rc && energyRaw) {
        if (!okWalk) return 0;
        *energyRaw = eWalk;
      }
      return rc;
    };
    double rn;
    bool updated = false;
    if (*mv64 > 0 && rr <= 100.0 * tol * braw) {
      // ro held b - A x in FP64 before this small correction: ro -= A32 dx,
      // whose error (the FP32 operator's relative error times |A dx| ~ |ro|)
      // is small against tol |b|; a pass is confirmed in FP64 below
      if (!applyOp(c, OP32, -1.0, c->dx, 1.0, c->ro)) return 0;
      if (!fbh_psolve_dev(c, c->w, c->ro)) return 0;
      const double *v2[2] = {c->w, c->ro};
      double nn[2];
      if (!devNorms(c, 2, v2, nn)) return 0;
      rn = nn[0];
      rr = nn[1];
      updated = true;
    } else if (!residual64(&rn)) {
      return 0;
    }
    *resid = rn / bnrm2;
    if (rn != rn) { if (hist) hist[nh++] = *resid; *info = 2; return leave(1); }
    if (updated && *resid <= tol) {   // the FP64 residual decides
      if (!residual64(&rn)) return 0;
      *resid = rn / bnrm2;
    }
    if (hist) hist[nh++] = *resid;
    if (rn != rn) { *info = 2; return leave(1); }
    if (*resid <= tol) return leave(1);
    if (*iters >= maxit || *outer >= maxOuter) { *info = 1; return leave(1); }
  }
  *resid = 0.0;
  if (energyRaw) return walkCharges(c, c->csRes, 0, theta, x, nullptr, energyRaw, true, nullptr);
  return 1;
}

extern "C" int fbh_gmres_refined_dev(fbh_ctx *c, const double *b, double *x, int restrt, double tol, int maxit,
                                     int maxOuter, int *iters, double *resid, double *hist, int *info, int *outer,
                                     int *mv64) {
  return refinedSolve(c, b, x, restrt, tol, maxit, maxOuter, iters, resid, hist, info, outer, mv64, 0.0, nullptr);
}

extern "C" int fbh_gmres_refined_energy_dev(fbh_ctx *c, const double *b, double *x, int restrt, double tol,
                                            int maxit, int maxOuter, double theta, int *iters, double *resid,
                                            double *hist, int *info, int *outer, int *mv64, double *energy_raw) {
  if (!energy_raw) { setErr("fbh_gmres_refined_energy_dev: bad arguments"); return 0; }
  return refinedSolve(c, b, x, restrt, tol, maxit, maxOuter, iters, resid, hist, info, outer, mv64, theta,
                      energy_raw);
}

// the factor every right-hand side (direct, tree, and the gradient's R^T) is scaled by
static double rhsFactor(const fbh_ctx *c) { return c->rhsFac != 0.0 ? c->rhsFac : kFourPiI / c->eps1; }

extern "C" int fbh_set_rhs_factor(fbh_ctx *c, double fac) {
  if (!c || !(fac == fac)) { setErr("fbh_set_rhs_factor: bad arguments"); return 0; }
  c->rhsFac = fac;
  return 1;
}

// device copies of the charges (freed by the caller)
static int uploadCharges(int nChar, const double *pos, const double *chr, double **dp, double **dc) {
  *dp = *dc = nullptr;
  if (hipMalloc(dp, 3 * (size_t)nChar * sizeof(double)) != hipSuccess ||
      hipMalloc(dc, (size_t)nChar * sizeof(double)) != hipSuccess ||
      hipMemcpy(*dp, pos, 3 * (size_t)nChar * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(*dc, chr, (size_t)nChar * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    if (*dp) hipFree(*dp);
    if (*dc) hipFree(*dc);
    setErr("uploading %d charges failed", nChar);
    return 0;
  }
  return 1;
}

// No-allocation staging for fbh_rhs_dev_noalloc: a file-scope device array
// of FBH_NOALLOC_MAX_CHARGES charges (x, y, z, then the charges; 8 MiB), one
// copy per GPU held until the process ends.  On the host `scratch` names a
// stand-in, not the device copy, so host code never passes it to a kernel or a
// copy: fbh_create resolves each device's address once with
// hipGetSymbolAddress (that device current) and keeps it in gScratchDev.
static __device__ __align__(256) double scratch[4 * FBH_NOALLOC_MAX_CHARGES];
#define FBH_MAX_DEVICES 64
static std::mutex gScratchMu;
static double *gScratchDev[FBH_MAX_DEVICES];

static int resolveScratch(int dev) {
  if (dev < 0 || dev >= FBH_MAX_DEVICES) { setErr("device %d beyond the staging table", dev); return 0; }
  std::lock_guard<std::mutex> lock(gScratchMu);
  if (gScratchDev[dev]) return 1;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    setErr("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    return 0;
  }
  gScratchDev[dev] = (double *)p;
  return 1;
}

// the direct right-hand side (src/fabipb.c:789-798) from device copies of the
// charges; fp32: the precision form's kernel (q = 1)
static int rhsDirect(fbh_ctx *c, int nChar, const double *dp, const double *dc, double *b_dev, bool fp32 = false) {
  const double fac = rhsFactor(c);
  if (fp32 && c->qorder == 1)
    hipLaunchKernelGGL(k_rhs32_q1, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, c->pnl, c->n, nChar, dp, dc,
                       b_dev, fac, c->Q.t[0], c->Q.t[0] * c->Q.w[0] * c->Q.w[0]);
  else if (c->qorder == 1)
    hipLaunchKernelGGL(k_rhs_q1, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, c->pnl, c->n, nChar, dp, dc,
                       b_dev, fac, c->Q.t[0], c->Q.t[0] * c->Q.w[0] * c->Q.w[0]);
  else
    hipLaunchKernelGGL(k_rhs, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, c->pnl, c->n, nChar, dp, dc, b_dev,
                       fac, c->Q, c->qorder);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) { setErr("k_rhs failed: %s", hipGetErrorString(e)); return 0; }
  return 1;
}

extern "C" int fbh_rhs_dev(fbh_ctx *c, int nChar, const double *pos, const double *chr, double *b_dev) {
  if (!c || nChar <= 0 || !pos || !chr || !b_dev) { setErr("fbh_rhs: bad arguments"); return 0; }
  double *dp = nullptr, *dc = nullptr;
  if (!uploadCharges(nChar, pos, chr, &dp, &dc)) return 0;
  const int ok = rhsDirect(c, nChar, dp, dc, b_dev);
  hipFree(dp);
  hipFree(dc);
  return ok;
}

// fbh_rhs_dev with the charges staged in the static array instead of device
// memory allocated per call: no hipMalloc / hipFree.  Calls that overlap on one
// GPU share its array and race.
extern "C" int fbh_rhs_dev_noalloc(fbh_ctx *c, int nChar, const double *pos, const double *chr, double *b_dev) {
  if (!c || nChar <= 0 || !pos || !chr || !b_dev) { setErr("fbh_rhs_dev_noalloc: bad arguments"); return 0; }
  if (nChar > FBH_NOALLOC_MAX_CHARGES) {
    setErr("fbh_rhs_dev_noalloc: %d charges exceed the static staging array's %d", nChar, FBH_NOALLOC_MAX_CHARGES);
    return 0;
  }
  double *sp = c->dev >= 0 && c->dev < FBH_MAX_DEVICES ? gScratchDev[c->dev] : nullptr;
  if (!sp) { setErr("fbh_rhs_dev_noalloc: staging array of device %d not resolved", c->dev); return 0; }
  HCK(hipMemcpyAsync(sp, pos, 3 * (size_t)nChar * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HCK(hipMemcpyAsync(sp + 3 * (size_t)nChar, chr, (size_t)nChar * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return rhsDirect(c, nChar, sp, sp + 3 * (size_t)nChar, b_dev);
}

// ------------------------------------------------ differentiation form
static int devDot(fbh_ctx *c, const double *a, const double *b, int m, double *out) {
  hipLaunchKernelGGL(k_dot_part, dim3(RED_BLOCKS), dim3(256), 0, c->stream, a, b, m, c->part);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, c->stream, c->part, RED_BLOCKS, c->scal);
  HCK(hipMemcpyAsync(out, c->scal, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  return 1;
}
static int devSum(fbh_ctx *c, const double *a, int m, double *out) {
  hipLaunchKernelGGL(k_sum_part, dim3(RED_BLOCKS), dim3(256), 0, c->stream, a, m, c->part);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, c->stream, c->part, RED_BLOCKS, c->scal);
  HCK(hipMemcpyAsync(out, c->scal, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  return 1;
}


extern "C" int fbh_energy_direct_dev(fbh_ctx *c, int nChar, const double *pos, const double *chr,
                                     const double *sgm_dev, double *g_dev, double *energy_raw) {
  if (!c || nChar <= 0 || !pos || !chr || !sgm_dev || !energy_raw) {
    setErr("fbh_energy_direct_dev: bad arguments");
    return 0;
  }
  if (!ensureGmres(c, 1)) return 0;
  double *dp, *dc, *g = g_dev;
  if (!uploadCharges(nChar, pos, chr, &dp, &dc)) return 0;
  if (!g && hipMalloc(&g, 2 * (size_t)c->n * sizeof(double)) != hipSuccess) {
    hipFree(dp);
    hipFree(dc);
    setErr("fbh_energy_direct_dev: allocation failed");
    return 0;
  }
  hipLaunchKernelGGL(k_energy_g, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, c->pnl, c->n, nChar, dp, dc, g,
                     (double *)nullptr, c->kappa, c->eps, c->Q, c->qorder);
  int ok = hipGetLastError() == hipSuccess && devDot(c, g, sgm_dev, 2 * c->n, energy_raw);
  hipStreamSynchronize(c->stream);
  if (!g_dev) hipFree(g);
  hipFree(dp);
  hipFree(dc);
  if (!ok) setErr("fbh_energy_direct_dev: %s", gErr[0] ? gErr : "kernel failed");
  return ok;
}

static int kappaOperatorTerms(fbh_ctx *c, const double *sgm_dev, const double *lam_dev, double *part, double *eNear,
                              double *eFar);

// dE/dq and dE/dkappa of E_raw = g(q, kappa) . u with A(kappa) u = b(q):
//   A^T lam = g,  dE/dq = G^T u + R^T lam,  dE/dkappa = (dg/dkappa) . u + lam^T (dK/dkappa) u
// (b = R q does not depend on kappa, A = D - K with D free of kappa)
extern "C" int fbh_gradient_dev(fbh_ctx *c, int nChar, const double *pos, const double *chr, const double *sgm_dev,
                                double *lam_dev, int restrt, double tol, int maxit, double *dEdq,
                                double *dEdkappa, int *iters, double *resid, int *info) {
  if (!c || nChar <= 0 || !pos || !chr || !sgm_dev || !lam_dev || !dEdq || !dEdkappa || !iters || !resid ||
      !info || restrt < 1 || maxit < 1) {
    setErr("fbh_gradient_dev: bad arguments");
    return 0;
  }
  if (!ensureAdjoint(c) || !ensureGmres(c, restrt)) return 0;
  const int n = c->n, n2 = 2 * n;
  const int nChunk = (n + GRAD_PCH - 1) / GRAD_PCH;
  double *dp = nullptr, *dc = nullptr, *g = nullptr, *gk = nullptr, *part = nullptr, *dq = nullptr;
  if (!uploadCharges(nChar, pos, chr, &dp, &dc)) return 0;
  size_t partN = std::max<size_t>((size_t)nChunk * nChar, std::max(n, c->nCube));
  int ok = hipMalloc(&g, n2 * sizeof(double)) == hipSuccess && hipMalloc(&gk, n2 * sizeof(double)) == hipSuccess &&
           hipMalloc(&part, partN * sizeof(double)) == hipSuccess &&
           hipMalloc(&dq, (size_t)nChar * sizeof(double)) == hipSuccess;
  if (!ok) setErr("fbh_gradient_dev: allocation failed");
  double eK = 0.0, eNear = 0.0, eFar = 0.0;
  if (ok) {
    hipLaunchKernelGGL(k_energy_g, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->pnl, n, nChar, dp, dc, g, gk,
                       c->kappa, c->eps, c->Q, c->qorder);
    ok = hipGetLastError() == hipSuccess && gmresCore(c, OPT64, g, lam_dev, restrt, tol, maxit, iters, resid,
                                                      nullptr, info);
  }
  if (ok) {   // dE/dq
    const double fac = rhsFactor(c);
    hipLaunchKernelGGL(k_grad_q<true>, dim3((nChar + 255) / 256, nChunk), dim3(256), 0, c->stream, c->pnl, n, nChar, dp,
                       sgm_dev, lam_dev, part, c->kappa, c->eps, fac, c->Q, c->qorder);
    hipLaunchKernelGGL(k_grad_sum, dim3((nChar + 255) / 256), dim3(256), 0, c->stream, part, nChunk, nChar, dq, 0);
    ok = hipGetLastError() == hipSuccess &&
         hipMemcpyAsync(dEdq, dq, (size_t)nChar * sizeof(double), hipMemcpyDeviceToHost, c->stream) == hipSuccess &&
         hipStreamSynchronize(c->stream) == hipSuccess;
    if (!ok) setErr("fbh_gradient_dev: charge gradient kernels failed");
  }
  if (ok) ok = devDot(c, gk, sgm_dev, n2, &eK);
  if (ok) ok = kappaOperatorTerms(c, sgm_dev, lam_dev, part, &eNear, &eFar);
  *dEdkappa = eK + eNear + eFar;
  hipStreamSynchronize(c->stream);
  for (double *q : {dp, dc, g, gk, part, dq})
    if (q) hipFree(q);
  return ok;
}

// lam^T (dK/dkappa) u, near and far parts (D is free of kappa); part: a device
// buffer of max(n, nCube) doubles
static int kappaOperatorTerms(fbh_ctx *c, const double *sgm_dev, const double *lam_dev, double *part, double *eNear,
                              double *eFar) {
  const int n = c->n;
  int ok = 1;
  if (ok) {   // near part of lam^T dK/dkappa u
    const char *er = getenv("FABIPB_DK_RULES");   // "1": the full rules for every pair (check path)
    if (c->qorder == 1 && !(er && er[0] == '1')) {
      if (!c->specDk) {
        ok = ensureSpec(c) && (c->specDk = dalloc<double4>(c, (size_t)c->nSpec)) != nullptr;
        if (ok)
          hipLaunchKernelGGL(k_spec_vals<true>, dim3(n), dim3(64), 0, c->stream, c->pnl, c->specOff, c->specCol,
                             (float4 *)nullptr, c->specDk, c->kappa, c->eps, c->Q, c->qorder);
      }
      if (ok) {
        // expNeg's table, also when no tile of the operator is computed on the fly
        hipLaunchKernelGGL(k_exp2_tab, dim3(1), dim3(256), 0, c->stream);
        hipLaunchKernelGGL(k_near_dk, dim3(c->nTiles), dim3(256), 0, c->stream, c->tiles, c->leafStart,
                           c->leafCount, c->srcOff, c->ncol, c->srcList, c->pnl, sgm_dev, lam_dev, c->specOff,
                           c->specCol, c->specDk, part, n, c->kappa, c->eps, c->Q.t[0]);
      }
    } else {
      hipLaunchKernelGGL(k_near_dk_rules, dim3(n), dim3(256), 0, c->stream, c->pnl, c->leafOf, c->srcOff, c->ncol,
                         c->srcList, sgm_dev, lam_dev, part, n, c->kappa, c->eps, c->Q, c->qorder);
    }
    ok = ok && hipGetLastError() == hipSuccess && devSum(c, part, n, eNear);
  }
  if (ok) {   // far part: moments of u, adjoint locals of lam, the kappa-differentiated M2L between them
    OpBufs<double> B = bufs64(c);
    q2mPass(c, B, sgm_dev, c->stream);
    for (int lev = c->depth - 1; lev >= c->height; lev--) m2mPass(c, B, lev, c->stream);
    launchQ2M<double>(c, c->Q0, c->Q1, lam_dev, c->adjL, 1, c->stream);
    B.mom = c->adjL;
    for (int lev = c->depth - 1; lev >= c->height; lev--) m2mPass(c, B, lev, c->stream);
    if (c->gridK && c->m2lG && !c->m2lGdk) {   // (dGk/dkappa, Gk) per pair, cached as the M2L tensors are
      if (!(c->m2lGdk = dalloc<double>(c, (size_t)c->nM2L * 2 * c->nmMax))) return 0;
      hipLaunchKernelGGL(k_m2l_derivs, dim3((unsigned)c->nM2L), dim3(64),
                         2 * (size_t)derivScratch(c->maxOrder) * sizeof(double), c->stream, c->m2lDst, c->m2lSrc,
                         c->cubeLevelD, c->ordM2LD, c->cubeX, c->m2lGdk, c->nmMax, c->kappa, c->T, true);
    }
    if (c->m2lGdk && m2lGrid<double, false, true>(c, c->m2lGdk, c->mom, c->lec, c->stream)) {
      hipLaunchKernelGGL(k_dk_far_dot, dim3(c->nCube), dim3(64), 0, c->stream, c->cubeLevelD, c->ordMomD, c->adjL,
                         c->lec, part, c->nmMax, c->eps);
    } else {
      const size_t lds = (2 * (size_t)derivScratch(c->maxOrder) + 6 * (size_t)NMOM(c->maxOrder)) * sizeof(double);
      hipLaunchKernelGGL(k_m2l_dk, dim3(c->nCube), dim3(64), lds, c->stream, c->m2lOff, c->m2lSrc, c->cubeLevelD,
                         c->ordM2LD, c->ordMomD, c->cubeX, c->mom, c->adjL, part, c->nmMax, c->kappa, c->eps, c->T);
    }
    ok = hipGetLastError() == hipSuccess && devSum(c, part, c->nCube, eFar);
    if (!ok) setErr("fbh_gradient_dev: kappa-derivative kernels failed");
  }
  return ok;
}


extern "C" int fbh_rhs(fbh_ctx *c, int nChar, const double *pos, const double *chr, double *b) {
  if (!fbh_rhs_dev(c, nChar, pos, chr, c->ys)) return 0;
  HCK(hipMemcpy(b, c->ys, 2 * (size_t)c->n * sizeof(double), hipMemcpyDeviceToHost));
  return 1;
}

// the charge-tree walks: mode 0 the panel-tree energy of sgm_dev into
// *energy_raw, mode 1 the tree right-hand side into b_dev, mode 2 the energy's
// weight vector (dE/dsgm) into b_dev
// A charge set on the device: the charges, their charge tree (the arrays of
// src/chargeTree.c:228-307), the walk's coefficient descriptors and its
// scratch (moments, per-cluster polynomial tables, slice partials), in one
// arena.  stageCharges validates every index the walks dereference and copies
// the inputs in one transfer; walkCharges runs the moments and one walk on
// what is staged, so a set staged once (fbh_set_charges) serves any number
// of right-hand sides and energies without host traffic.
static int stageCharges(fbh_ctx *c, ChargeSet &s, int nChar, const double *pos, const double *chr, int nc,
                        const int *ctLevel, const double *ctX, const double *ctERad, const int *ctKidStart,
                        const int *ctNKids, const int *ctChgStart, const int *ctNChgs, const int *ctChgIdx,
                        const int *ctOrder) {
  s.ready = false;
  s.momReady = false;
  if (!c || nChar <= 0 || nc <= 0 || !pos || !chr || !ctLevel || !ctX || !ctERad || !ctKidStart || !ctNKids ||
      !ctChgStart || !ctNChgs || !ctChgIdx || !ctOrder) {
    setErr("charge set: bad arguments");
    return 0;
  }
  const int depth = c->depth;
  std::vector<int> lvStart(depth + 2, -1), lvCount(depth + 1, 0);
  int nLeafChg = 0, maxOrd = 0;
  for (int i = 0; i < nc; i++) {
    int l = ctLevel[i];
    if (l < 0 || l > depth || (i > 0 && l < ctLevel[i - 1])) { setErr("fbh_energy: bad charge-tree levels"); return 0; }
    if (lvStart[l] < 0) lvStart[l] = i;
    lvCount[l]++;
    if (l < depth && ctNKids[i] > 0 && (ctKidStart[i] <= i || ctKidStart[i] + ctNKids[i] > nc || ctNKids[i] > 8)) {
      setErr("fbh_energy: bad kid range at cube %d", i);
      return 0;
    }
    if (l == depth) {
      if (ctChgStart[i] < 0 || ctNChgs[i] < 0 || ctChgStart[i] + ctNChgs[i] > nChar) { setErr("fbh_energy: bad charge range"); return 0; }
      nLeafChg += ctNChgs[i];
    }
  }
  if (nLeafChg != nChar) { setErr("fbh_energy: leaves hold %d of %d charges", nLeafChg, nChar); return 0; }
  for (int k = 0; k < nChar; k++)
    if (ctChgIdx[k] < 0 || ctChgIdx[k] >= nChar) { setErr("fbh_energy: bad charge index"); return 0; }
  std::vector<int> ord(depth + 2, 0);
  for (int l = 0; l <= depth; l++) {
    ord[l] = ctOrder[l];
    if (l >= c->height) maxOrd = std::max(maxOrd, ord[l]);
    if (ord[l] < 0 || ord[l] > c->maxOrder) { setErr("fbh_energy: charge-tree order %d beyond %d", ord[l], c->maxOrder); return 0; }
  }
  ord[depth + 1] = maxOrd + 1;
  std::vector<int> par(nc, -1);   // each cluster's parent (the transposed walk climbs from the leaves)
  for (int i = 0; i < nc; i++)
    if (ctLevel[i] < depth)
      for (int k = 0; k < ctNKids[i]; k++) par[ctKidStart[i] + k] = i;
  const int nmMax = c->nmMax;
  // one expansion order over the levels the walk can accept: the wave walk
  // with per-cluster polynomial tables; otherwise the wave-per-panel walk
  const int c0 = lvStart[c->height], P = ord[c->height];
  bool uni = c0 >= 0 && P >= 1 && P <= 9;
  for (int l = c->height; l <= depth && uni; l++) uni = ord[l] == P;
  std::vector<int4> dsc;
  const int SZ = uni ? enPolySize(P) : 0, nClus = uni ? nc - c0 : 0;
  // waves per 64 panels (the wave walk): enough waves to fill the device when
  // the panels alone do not, at most one per cluster of level `height`, and
  // few enough start clusters per wave for the walk's stack
  int S = 1;
  if (uni) {
    const int nH = lvCount[c->height], panelWaves = (c->n + 63) / 64;
    S = std::max(1, std::min(nH, (EN_TARGET_WAVES + panelWaves - 1) / panelWaves));
    while ((nH + S - 1) / S + 8 * (depth - c->height + 1) > EN_WSTACK) S++;
  }
  if (uni) {
    int o = 0;
    for (int a = P; a >= 0; a--)   // the order enFar reads them in
      for (int b = P - a; b >= 0; b--)
        for (int g3 = P - a - b; g3 >= 0; g3--)
          for (int j = 0; j <= (P - a - b - g3) / 2; j++) dsc.push_back(make_int4(a | b << 8 | g3 << 16, j, 0, o++));
    for (int a = P + 1; a >= 0; a--)
      for (int b = P + 1 - a; b >= 0; b--)
        for (int g3 = P + 1 - a - b; g3 >= 0; g3--)
          for (int j = 0; j <= (P + 1 - a - b - g3) / 2; j++, o += 3)
            dsc.push_back(make_int4(a | b << 8 | g3 << 16, j, 1, o));
    if (o != SZ) { setErr("fbh_energy: coefficient table has %d of %d entries", o, SZ); return 0; }
  }
  // inputs staged in one host block and copied once into the set's arena;
  // scratch follows the inputs in the same arena
  struct Seg { const void *src; size_t bytes; size_t off; };
  std::vector<Seg> seg;
  size_t top = 0;
  auto add = [&](const void *src, size_t bytes) {
    seg.push_back({src, bytes, top});
    top += (bytes + 255) & ~(size_t)255;
    return seg.size() - 1;
  };
  const size_t iLevel = add(ctLevel, nc * sizeof(int)), iKid = add(ctKidStart, nc * sizeof(int)),
               iNK = add(ctNKids, nc * sizeof(int)), iCS = add(ctChgStart, nc * sizeof(int)),
               iNC = add(ctNChgs, nc * sizeof(int)), iCI = add(ctChgIdx, nChar * sizeof(int)),
               iOrd = add(ord.data(), ord.size() * sizeof(int)), iX = add(ctX, 3 * (size_t)nc * sizeof(double)),
               iR = add(ctERad, nc * sizeof(double)), iPos = add(pos, 3 * (size_t)nChar * sizeof(double)),
               iChr = add(chr, nChar * sizeof(double)), iDesc = add(dsc.data(), dsc.size() * sizeof(int4)),
               iPar = add(par.data(), nc * sizeof(int));
  const size_t inBytes = top;
  const size_t iMom = add(nullptr, (size_t)nc * nmMax * sizeof(double)),
               iPart = add(nullptr, (size_t)S * 3 * c->n * sizeof(double)), iOut = add(nullptr, sizeof(double)),
               iRed = add(nullptr, RED_BLOCKS * sizeof(double)),
               iPoly = add(nullptr, (size_t)nClus * SZ * sizeof(double)),
               iPoly32 = add(nullptr, (size_t)nClus * SZ * sizeof(float));
  if (s.arenaBytes < top) {
    if (s.arena) {
      HCK(hipStreamSynchronize(c->stream));   // a walk on the old arena may still run
      hipFree(s.arena);
    }
    s.arena = nullptr;
    s.arenaBytes = 0;
    if (hipMalloc(&s.arena, top) != hipSuccess) { setErr("fbh_energy: device arena of %zu bytes", top); return 0; }
    s.arenaBytes = top;
  }
  s.stage.resize(inBytes);
  for (const Seg &g : seg)
    if (g.src && g.bytes) memcpy(s.stage.data() + g.off, g.src, g.bytes);
  unsigned char *A = (unsigned char *)s.arena;
  HCK(hipMemcpyAsync(A, s.stage.data(), inBytes, hipMemcpyHostToDevice, c->stream));
  s.dLevel = (int *)(A + seg[iLevel].off);
  s.dKid = (int *)(A + seg[iKid].off);
  s.dNK = (int *)(A + seg[iNK].off);
  s.dCS = (int *)(A + seg[iCS].off);
  s.dNC = (int *)(A + seg[iNC].off);
  s.dCI = (int *)(A + seg[iCI].off);
  s.dOrd = (int *)(A + seg[iOrd].off);
  s.dPar = (int *)(A + seg[iPar].off);
  s.dX = (double *)(A + seg[iX].off);
  s.dR = (double *)(A + seg[iR].off);
  s.dPos = (double *)(A + seg[iPos].off);
  s.dChr = (double *)(A + seg[iChr].off);
  s.dMom = (double *)(A + seg[iMom].off);
  s.dPart = (double *)(A + seg[iPart].off);
  s.dOut = (double *)(A + seg[iOut].off);
  s.dRed = (double *)(A + seg[iRed].off);
  s.dPoly = (double *)(A + seg[iPoly].off);
  s.dPoly32 = (float *)(A + seg[iPoly32].off);
  s.dDesc = (const int4 *)(A + seg[iDesc].off);
  s.nChar = nChar;
  s.nc = nc;
  s.lvStart = lvStart;
  s.lvCount = lvCount;
  s.ord = ord;
  s.c0 = c0;
  s.P = P;
  s.uni = uni;
  s.S = S;
  s.SZ = SZ;
  s.nClus = nClus;
  s.nDesc = (int)dsc.size();
  s.ready = true;
  return 1;
}

// the charge moments and one walk of a staged set: mode 0 the panel-tree
// energy (src/treecode.c:927-1017) into *energy_raw, 1 the charge-tree
// right-hand side (src/fabipb.c:633-742) into b_dev, 2 the energy's weights
// dE/dsgm into b_dev, 3 the energy's kappa derivative at fixed sgm into
// *energy_raw
// st: the stream to walk on (the context's stream when null)
static int walkCharges(fbh_ctx *c, const ChargeSet &s, int mode, double theta, const double *sgm_dev, double *b_dev,
                       double *energy_raw, bool f32 = false, hipStream_t st = nullptr) {
  const hipStream_t strm = st ? st : c->stream;
  if (!s.ready) { setErr("fbh_energy/fbh_rhs_tree: no charge set staged"); return 0; }
  const bool scalar = mode == 0 || mode == 3 || mode == 4;   // 4: the weights into b_dev as well
  if ((scalar && (!energy_raw || !sgm_dev)) || (!(mode == 0 || mode == 3) && !b_dev)) {
    setErr("fbh_energy/fbh_rhs_tree: bad arguments");
    return 0;
  }
  const int depth = c->depth, nmMax = c->nmMax, P = s.P, S = s.S;
  const std::vector<int> &lvStart = s.lvStart, &lvCount = s.lvCount, &ord = s.ord;
  // charge moments: leaves, then upward to `height` (src/chargeTree.c:397-416),
  // and the walk coefficients, once per staged set
  if (!s.momReady) {
    if (lvCount[depth] > 0)
      hipLaunchKernelGGL(k_ct_leaf, dim3(lvCount[depth]), dim3(blk64(NMOM(ord[depth]), 512)), 0, strm,
                         lvStart[depth], s.dCS, s.dNC, s.dCI, s.dX, s.dPos, s.dChr, ord[depth], s.dMom, nmMax, c->T);
    for (int l = depth - 1; l >= c->height; l--) {
      if (lvCount[l] <= 0) continue;
      int nOut = NMOM(ord[l]);
      const size_t lds =
          2 * nOut * sizeof(double) + (size_t)(nOut + 1 + c->hM2mStart[nOut] - c->hM2mStart[0]) * sizeof(int);
      hipLaunchKernelGGL(k_ct_m2m, dim3(lvCount[l]), dim3(blk64((nOut + 3) / 4, 256)), lds, strm, lvStart[l],
                         s.dKid, s.dNK, s.dX, s.dOrd, l, s.dMom, nmMax, c->T);
    }
    if (s.uni) {
      hipLaunchKernelGGL(k_ct_poly, dim3(s.nClus), dim3(256), 0, strm, s.c0, s.dMom, nmMax, s.dDesc, s.nDesc,
                         s.SZ, s.dPoly);
      const long long nP = (long long)s.nClus * s.SZ;
      if (nP > 0)
        hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((nP + 255) / 256)), dim3(256), 0, strm, s.dPoly, s.dPoly32,
                           nP);
    }
    if (hipGetLastError() != hipSuccess) { setErr("charge moment kernels failed"); return 0; }
    s.momReady = true;
  }
  if (s.uni) {
    const int blocks = ((c->n + 64 * EN_WPB - 1) / (64 * EN_WPB)) * S;
#define EN_WAVE(PP, MM)                                                                                           \
  case PP:                                                                                                        \
    hipLaunchKernelGGL((k_energy_wave<PP, MM>), dim3(blocks), dim3(64 * EN_WPB), 0, strm, c->pnl, c->n,       \
                       sgm_dev,                                                                                   \
                       s.dLevel, s.dX, s.dR, s.dKid, s.dNK, s.dCS, s.dNC, s.dCI, s.dPos, s.dChr, s.dPoly, s.c0,     \
                       depth, c->height, theta, c->kappa, MM == 1 ? rhsFactor(c) : c->eps, c->Q, c->qorder,       \
                       lvCount[c->height], S, (MM && S == 1) ? b_dev : s.dPart, s.dPart + 2 * (size_t)S * c->n);  \
    break;
    if (mode == 0 && f32) {
#define EN_WAVE32(PP)                                                                                             \
  case PP:                                                                                                        \
    hipLaunchKernelGGL((k_energy_wave<PP, 0, float>), dim3(blocks), dim3(64 * EN_WPB), 0, strm, c->pnl,      \
                       c->n, sgm_dev, s.dLevel, s.dX, s.dR, s.dKid, s.dNK, s.dCS, s.dNC, s.dCI, s.dPos, s.dChr,    \
                       s.dPoly32, s.c0, depth, c->height, theta, c->kappa, c->eps, c->Q, c->qorder,              \
                       lvCount[c->height], S, s.dPart, s.dPart + 2 * (size_t)S * c->n);                           \
    break;
      switch (P) { EN_WAVE32(1) EN_WAVE32(2) EN_WAVE32(3) EN_WAVE32(4) EN_WAVE32(5) EN_WAVE32(6) EN_WAVE32(7) EN_WAVE32(8) EN_WAVE32(9) }
#undef EN_WAVE32
    } else if (mode == 0) {
      switch (P) { EN_WAVE(1, 0) EN_WAVE(2, 0) EN_WAVE(3, 0) EN_WAVE(4, 0) EN_WAVE(5, 0) EN_WAVE(6, 0) EN_WAVE(7, 0) EN_WAVE(8, 0) EN_WAVE(9, 0) }
    } else if (mode == 3) {
      switch (P) { EN_WAVE(1, 3) EN_WAVE(2, 3) EN_WAVE(3, 3) EN_WAVE(4, 3) EN_WAVE(5, 3) EN_WAVE(6, 3) EN_WAVE(7, 3) EN_WAVE(8, 3) EN_WAVE(9, 3) }
    } else if (mode == 2) {
      switch (P) { EN_WAVE(1, 2) EN_WAVE(2, 2) EN_WAVE(3, 2) EN_WAVE(4, 2) EN_WAVE(5, 2) EN_WAVE(6, 2) EN_WAVE(7, 2) EN_WAVE(8, 2) EN_WAVE(9, 2) }
    } else if (mode == 4) {
      switch (P) { EN_WAVE(1, 4) EN_WAVE(2, 4) EN_WAVE(3, 4) EN_WAVE(4, 4) EN_WAVE(5, 4) EN_WAVE(6, 4) EN_WAVE(7, 4) EN_WAVE(8, 4) EN_WAVE(9, 4) }
    } else {
      switch (P) { EN_WAVE(1, 1) EN_WAVE(2, 1) EN_WAVE(3, 1) EN_WAVE(4, 1) EN_WAVE(5, 1) EN_WAVE(6, 1) EN_WAVE(7, 1) EN_WAVE(8, 1) EN_WAVE(9, 1) }
    }
#undef EN_WAVE
    if ((mode == 1 || mode == 2 || mode == 4) && S > 1)
      hipLaunchKernelGGL(k_sum_slices, dim3((2 * c->n + 255) / 256), dim3(256), 0, strm, s.dPart, S,
                         2 * (long long)c->n, b_dev);
  } else if (mode != 0) {
    setErr("fbh_rhs_tree / tree gradient: the accepted levels must share one expansion order");
    return 0;
  } else {
    int P1 = ord[depth + 1];
    size_t lds = (2 * (size_t)derivScratch(P1) + 2 * (size_t)NMOM(P1)) * sizeof(double);
    hipLaunchKernelGGL(k_energy, dim3(c->n), dim3(64), lds, strm, c->pnl, c->n, sgm_dev, s.dLevel, s.dX, s.dR,
                       s.dKid, s.dNK, s.dCS, s.dNC, s.dCI, s.dPos, s.dChr, s.dOrd, s.dMom, nmMax, depth, c->height,
                       theta, c->kappa, c->eps, c->Q, c->qorder, c->T, s.dPart);
  }
  hipError_t e = hipSuccess;
  if (mode == 0 || mode == 3 || mode == 4) {
    hipLaunchKernelGGL(k_sum_part, dim3(RED_BLOCKS), dim3(256), 0, strm,
                       mode == 4 ? s.dPart + 2 * (size_t)S * c->n : s.dPart, S * c->n, s.dRed);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, strm, s.dRed, RED_BLOCKS, s.dOut);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(energy_raw, s.dOut, sizeof(double), hipMemcpyDeviceToHost, strm);
  } else {
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(strm);
  if (e != hipSuccess) { setErr("energy kernels failed: %s", hipGetErrorString(e)); return 0; }
  return 1;
}

static int chargeTreeWalk(fbh_ctx *c, int mode, int nChar, const double *pos, const double *chr, int nc,
                          const int *ctLevel, const double *ctX, const double *ctERad, const int *ctKidStart,
                          const int *ctNKids, const int *ctChgStart, const int *ctNChgs, const int *ctChgIdx,
                          const int *ctOrder, double theta, const double *sgm_dev, double *b_dev,
                          double *energy_raw) {
  if (!c || (mode == 0 && (!energy_raw || !sgm_dev)) || (mode != 0 && !b_dev)) {
    setErr("fbh_energy/fbh_rhs_tree: bad arguments");
    return 0;
  }
  return stageCharges(c, c->csCall, nChar, pos, chr, nc, ctLevel, ctX, ctERad, ctKidStart, ctNKids, ctChgStart,
                      ctNChgs, ctChgIdx, ctOrder) &&
         walkCharges(c, c->csCall, mode, theta, sgm_dev, b_dev, energy_raw);
}

extern "C" int fbh_energy_dev(fbh_ctx *c, int nChar, const double *pos, const double *chr, int nc,
                              const int *ctLevel, const double *ctX, const double *ctERad, const int *ctKidStart,
                              const int *ctNKids, const int *ctChgStart, const int *ctNChgs, const int *ctChgIdx,
                              const int *ctOrder, double theta, const double *sgm_dev, double *energy_raw) {
  return chargeTreeWalk(c, 0, nChar, pos, chr, nc, ctLevel, ctX, ctERad, ctKidStart, ctNKids, ctChgStart, ctNChgs,
                        ctChgIdx, ctOrder, theta, sgm_dev, nullptr, energy_raw);
}

extern "C" int fbh_rhs_tree_dev(fbh_ctx *c, int nChar, const double *pos, const double *chr, int nc,
                                const int *ctLevel, const double *ctX, const double *ctERad, const int *ctKidStart,
                                const int *ctNKids, const int *ctChgStart, const int *ctNChgs, const int *ctChgIdx,
                                const int *ctOrder, double theta, double *b_dev) {
  return chargeTreeWalk(c, 1, nChar, pos, chr, nc, ctLevel, ctX, ctERad, ctKidStart, ctNKids, ctChgStart, ctNChgs,
                        ctChgIdx, ctOrder, theta, nullptr, b_dev, nullptr);
}

// resident charge set: staged once, then right-hand sides and energies with no host traffic
extern "C" int fbh_set_charges(fbh_ctx *c, int nChar, const double *pos, const double *chr, int nc,
                               const int *ctLevel, const double *ctX, const double *ctERad, const int *ctKidStart,
                               const int *ctNKids, const int *ctChgStart, const int *ctNChgs, const int *ctChgIdx,
                               const int *ctOrder) {
  if (!stageCharges(c, c->csRes, nChar, pos, chr, nc, ctLevel, ctX, ctERad, ctKidStart, ctNKids, ctChgStart,
                    ctNChgs, ctChgIdx, ctOrder))
    return 0;
  HCK(hipStreamSynchronize(c->stream));
  return 1;
}

extern "C" int fbh_rhs_res_dev(fbh_ctx *c, double *b_dev) {
  if (!c || !b_dev) { setErr("fbh_rhs_res_dev: bad arguments"); return 0; }
  if (!c->csRes.ready) { setErr("fbh_rhs_res_dev: no resident charge set (fbh_set_charges)"); return 0; }
  return rhsDirect(c, c->csRes.nChar, c->csRes.dPos, c->csRes.dChr, b_dev);
}

extern "C" int fbh_rhs32_res_dev(fbh_ctx *c, double *b_dev) {
  if (!c || !b_dev) { setErr("fbh_rhs32_res_dev: bad arguments"); return 0; }
  if (!c->csRes.ready) { setErr("fbh_rhs32_res_dev: no resident charge set (fbh_set_charges)"); return 0; }
  return rhsDirect(c, c->csRes.nChar, c->csRes.dPos, c->csRes.dChr, b_dev, true);
}

extern "C" int fbh_rhs_tree_res_dev(fbh_ctx *c, double theta, double *b_dev) {
  if (!c || !b_dev) { setErr("fbh_rhs_tree_res_dev: bad arguments"); return 0; }
  return walkCharges(c, c->csRes, 1, theta, nullptr, b_dev, nullptr);
}

extern "C" int fbh_energy_res_dev(fbh_ctx *c, double theta, const double *sgm_dev, double *energy_raw) {
  if (!c || !sgm_dev || !energy_raw) { setErr("fbh_energy_res_dev: bad arguments"); return 0; }
  return walkCharges(c, c->csRes, 0, theta, sgm_dev, nullptr, energy_raw);
}

// the same walk with the far polynomials in FP32 (precision form)
extern "C" int fbh_energy32_res_dev(fbh_ctx *c, double theta, const double *sgm_dev, double *energy_raw) {
  if (!c || !sgm_dev || !energy_raw) { setErr("fbh_energy32_res_dev: bad arguments"); return 0; }
  return walkCharges(c, c->csRes, 0, theta, sgm_dev, nullptr, energy_raw, true);
}

// ------------------------------------------ exact charge gradients of the tree forms
// device buffers of one tree-gradient call, freed together
struct GradBufs {
  std::vector<void *> p;
  ~GradBufs() {
    for (void *q : p)
      if (q) hipFree(q);
  }
  template <class T>
  T *get(size_t count) {
    void *q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return nullptr;
    p.push_back(q);
    return (T *)q;
  }
};

// event storage of one batch of the transposed walk
#define ADJ_BATCH_BYTES (2ll << 30)

// The transposed walk of a staged set (k_adj_walk): kind 0 adds d/dq of the
// panel-tree energy at fixed sgm (w = sgm), kind 1 d/dq of lam . b_tree, the
// charge-tree right-hand side transposed (w = lam), into lamBuf (nc x nmMax,
// the accepted clusters' local sums) and phi (the charges' direct terms).
// A count pass sizes each block's events; the blocks then run in batches of at
// most ADJ_BATCH_BYTES of events: emit, sort by cluster, add per cluster.
static int adjWalk(fbh_ctx *c, const ChargeSet &s, int kind, double theta, const double *w, double *lamBuf,
                   double *phi, hipStream_t st) {
  if (!s.uni) { setErr("tree gradient: the accepted levels must share one expansion order"); return 0; }
  // slices: each wave's walk is split over S waves (one per share of the
  // clusters at level height), enough for ~16k waves in flight
  const int nWave = (c->n + 63) / 64, nH = s.lvCount[c->height];
  const int S = std::max(s.S, std::min(nH, (16384 + nWave - 1) / nWave));
  const int blocks = nWave * S, nm = NMOM(s.P), nc = s.nc;
  const double fac = rhsFactor(c);
  GradBufs B;
  AdjEvents ev{};
  ev.cnt = B.get<long long>(3 * (size_t)blocks + 3);
  ev.err = B.get<int>(1);
  if (!ev.cnt || !ev.err) { setErr("tree gradient: allocation failed"); return 0; }
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, ev.err, 1, 0);
#define ADJ_WALK(PP, KK, EE, G, G0)                                                                                \
  case PP:                                                                                                         \
    hipLaunchKernelGGL((k_adj_walk<PP, KK, EE>), dim3(G), dim3(64), 0, st, c->pnl, c->n, w, s.dLevel, s.dX,  \
                       s.dR, s.dKid, s.dNK, s.dCS, s.dNC, s.dCI, s.dPos, s.c0, c->depth, c->height, theta, c->kappa, \
                       c->eps, fac, c->Q, c->qorder, nH, S, G0, ev);                                                \
    break;
#define ADJ_SWITCH(KK, EE, G, G0)                                                                                  \
  switch (s.P) {                                                                                                   \
    ADJ_WALK(1, KK, EE, G, G0) ADJ_WALK(2, KK, EE, G, G0) ADJ_WALK(3, KK, EE, G, G0) ADJ_WALK(4, KK, EE, G, G0)    \
    ADJ_WALK(5, KK, EE, G, G0) ADJ_WALK(6, KK, EE, G, G0) ADJ_WALK(7, KK, EE, G, G0) ADJ_WALK(8, KK, EE, G, G0)    \
    ADJ_WALK(9, KK, EE, G, G0)                                                                                     \
  }
  if (kind == 0) {
    ADJ_SWITCH(0, false, blocks, 0)
  } else {
    ADJ_SWITCH(1, false, blocks, 0)
  }
  HCK(hipGetLastError());
  std::vector<long long> cnt(3 * (size_t)blocks), base(3 * (size_t)blocks + 3, 0);
  HCK(hipMemcpyAsync(cnt.data(), ev.cnt, cnt.size() * sizeof(long long), hipMemcpyDeviceToHost, st));
  HCK(hipStreamSynchronize(st));
  for (int g = 0; g < blocks; g++)
    for (int k = 0; k < 3; k++) base[3 * (g + 1) + k] = base[3 * g + k] + cnt[3 * g + k];
  // batches of whole blocks: accepting events nm values and four ints, direct events three ints,
  // direct values one double
  auto bytesOf = [&](int g0, int g1) {
    const long long a = base[3 * g1] - base[3 * g0], d = base[3 * g1 + 1] - base[3 * g0 + 1],
                    v = base[3 * g1 + 2] - base[3 * g0 + 2];
    return a * (8ll * nm + 16) + d * 16 + v * 8;
  };
  std::vector<int> cut{0};
  while (cut.back() < blocks) {
    int g = cut.back() + 1;
    while (g < blocks && bytesOf(cut.back(), g + 1) <= ADJ_BATCH_BYTES) g++;
    cut.push_back(g);
  }
  long long mA = 1, mD = 1, mV = 1;
  for (size_t i = 0; i + 1 < cut.size(); i++) {
    mA = std::max(mA, base[3 * cut[i + 1]] - base[3 * cut[i]]);
    mD = std::max(mD, base[3 * cut[i + 1] + 1] - base[3 * cut[i] + 1]);
    mV = std::max(mV, base[3 * cut[i + 1] + 2] - base[3 * cut[i] + 2]);
  }
  if (mA > INT32_MAX || mD > INT32_MAX || mV > INT32_MAX) { setErr("tree gradient: too many events in one block"); return 0; }
  long long *dBase = B.get<long long>(base.size());
  ev.accKey = B.get<int>(mA);
  ev.accIdx = B.get<int>(mA);
  ev.accVals = B.get<double>((size_t)mA * nm);
  ev.dirKey = B.get<int>(mD);
  ev.dirIdx = B.get<int>(mD);
  ev.dirOff = B.get<int>(mD);
  ev.dirVals = B.get<double>(mV);
  int *keyS = B.get<int>(std::max(mA, mD)), *idxS = B.get<int>(std::max(mA, mD)), *s0 = B.get<int>(nc),
      *s1 = B.get<int>(nc);
  int bits = 1;
  while ((1ll << bits) < nc) bits++;
  size_t tmpA = 0, tmpD = 0;
  HCK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmpA, ev.accKey, keyS, ev.accIdx, idxS, (int)mA, 0, bits, st));
  HCK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmpD, ev.dirKey, keyS, ev.dirIdx, idxS, (int)mD, 0, bits, st));
  size_t tmpBytes = std::max(tmpA, tmpD);
  void *tmp = B.get<unsigned char>(tmpBytes);
  if (!dBase || !ev.accKey || !ev.accIdx || !ev.accVals || !ev.dirKey || !ev.dirIdx || !ev.dirOff || !ev.dirVals ||
      !keyS || !idxS || !s0 || !s1 || !tmp) {
    setErr("tree gradient: event buffers (%lld accepting, %lld direct) allocation failed", mA, mD);
    return 0;
  }
  HCK(hipMemcpyAsync(dBase, base.data(), base.size() * sizeof(long long), hipMemcpyHostToDevice, st));
  ev.base = dBase;
  const int depth = c->depth;
  for (size_t i = 0; i + 1 < cut.size(); i++) {
    const int g0 = cut[i], g1 = cut[i + 1];
    ev.bA = base[3 * g0];
    ev.bD = base[3 * g0 + 1];
    ev.bV = base[3 * g0 + 2];
    const int nA = (int)(base[3 * g1] - ev.bA), nD = (int)(base[3 * g1 + 1] - ev.bD);
    if (kind == 0) {
      ADJ_SWITCH(0, true, g1 - g0, g0)
    } else {
      ADJ_SWITCH(1, true, g1 - g0, g0)
    }
    HCK(hipGetLastError());
    if (nA > 0) {
      size_t tb = tmpBytes;
      HCK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, ev.accKey, keyS, ev.accIdx, idxS, nA, 0, bits, st));
      hipLaunchKernelGGL(k_fill_int, dim3((nc + 255) / 256), dim3(256), 0, st, s0, nc, 0);
      hipLaunchKernelGGL(k_fill_int, dim3((nc + 255) / 256), dim3(256), 0, st, s1, nc, 0);
      hipLaunchKernelGGL(k_seg_bounds, dim3((nA + 255) / 256), dim3(256), 0, st, keyS, nA, s0, s1);
      hipLaunchKernelGGL(k_adj_red_acc, dim3(nc), dim3(ADJ_RED_THREADS), 0, st, s0, s1, idxS, ev.accVals, s.P,
                         c->nmMax, lamBuf);
    }
    if (nD > 0 && s.lvCount[depth] > 0) {
      size_t tb = tmpBytes;
      HCK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, ev.dirKey, keyS, ev.dirIdx, idxS, nD, 0, bits, st));
      hipLaunchKernelGGL(k_fill_int, dim3((nc + 255) / 256), dim3(256), 0, st, s0, nc, 0);
      hipLaunchKernelGGL(k_fill_int, dim3((nc + 255) / 256), dim3(256), 0, st, s1, nc, 0);
      hipLaunchKernelGGL(k_seg_bounds, dim3((nD + 255) / 256), dim3(256), 0, st, keyS, nD, s0, s1);
      hipLaunchKernelGGL(k_adj_red_dir, dim3(s.lvCount[depth]), dim3(64), 0, st, s.lvStart[depth], s0, s1,
                         idxS, ev.dirOff, ev.dirVals, s.dCS, s.dNC, s.dCI, phi);
    }
    HCK(hipGetLastError());
  }
#undef ADJ_SWITCH
#undef ADJ_WALK
  int err = 0;
  HCK(hipMemcpyAsync(&err, ev.err, sizeof(int), hipMemcpyDeviceToHost, st));
  HCK(hipStreamSynchronize(st));   // the buffers are freed on return
  if (err) { setErr("tree gradient: the transposed walk's passes disagree or its stack overflowed"); return 0; }
  return 1;
}

// dq (device, nChar) = phi + the clusters' local sums at the charges
static int adjEval(fbh_ctx *c, const ChargeSet &s, const double *lamBuf, const double *phi, double *dq) {
  const int depth = c->depth;
  if (s.lvCount[depth] > 0)
    hipLaunchKernelGGL(k_adj_eval, dim3(s.lvCount[depth]), dim3(64), 0, c->stream, s.lvStart[depth], s.dCS, s.dNC,
                       s.dCI, s.dPar, s.dLevel, s.dX, s.dPos, lamBuf, phi, c->height, s.P, c->nmMax, c->T, dq);
  HCK(hipGetLastError());
  return 1;
}

extern "C" int fbh_tree_adjoint_dev(fbh_ctx *c, int nChar, const double *pos, const double *chr, int nc,
                                    const int *ctLevel, const double *ctX, const double *ctERad,
                                    const int *ctKidStart, const int *ctNKids, const int *ctChgStart,
                                    const int *ctNChgs, const int *ctChgIdx, const int *ctOrder, double theta,
                                    int kind, const double *w_dev, double *dq_host) {
  if (!c || !w_dev || !dq_host || (kind != 0 && kind != 1)) { setErr("fbh_tree_adjoint_dev: bad arguments"); return 0; }
  ChargeSet &s = c->csCall;
  if (!stageCharges(c, s, nChar, pos, chr, nc, ctLevel, ctX, ctERad, ctKidStart, ctNKids, ctChgStart, ctNChgs,
                    ctChgIdx, ctOrder))
    return 0;
  GradBufs B;
  double *lamBuf = B.get<double>((size_t)nc * c->nmMax), *phi = B.get<double>(nChar), *dq = B.get<double>(nChar);
  if (!lamBuf || !phi || !dq) { setErr("fbh_tree_adjoint_dev: allocation failed"); return 0; }
  hipLaunchKernelGGL(k_fill, dim3((nc * c->nmMax + 255) / 256), dim3(256), 0, c->stream, lamBuf, nc * c->nmMax, 0.0);
  hipLaunchKernelGGL(k_fill, dim3((nChar + 255) / 256), dim3(256), 0, c->stream, phi, nChar, 0.0);
  if (!adjWalk(c, s, kind, theta, w_dev, lamBuf, phi, c->stream) || !adjEval(c, s, lamBuf, phi, dq)) return 0;
  HCK(hipMemcpyAsync(dq_host, dq, (size_t)nChar * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  return 1;
}

// The exact gradient of the panel-tree energy the plain solve reports
// (src/treecode.c:581-695, 927-1017): E_raw = W(q, kappa) . u with A u = b(q),
// b the direct right-hand side (rhsTree 0) or the charge-tree one (1).
//   A^T lam = W q            (the energy's weights: the walk's mode 2)
//   dE/dq = (transposed energy walk of u) + R^T lam   (R direct sums, or the
//           transposed charge-tree walk of lam)
//   dE/dkappa = dE/dkappa at fixed u (the walk's mode 3) + lam^T (dK/dkappa) u
extern "C" int fbh_gradient_tree_dev(fbh_ctx *c, int nChar, const double *pos, const double *chr, int nc,
                                     const int *ctLevel, const double *ctX, const double *ctERad,
                                     const int *ctKidStart, const int *ctNKids, const int *ctChgStart,
                                     const int *ctNChgs, const int *ctChgIdx, const int *ctOrder, double theta,
                                     int rhsTree, const double *sgm_dev, double *lam_dev, int lamFromRhs,
                                     int restrt, double tol, int maxit, double *dEdq, double *dEdkappa, int *iters,
                                     double *resid, int *info) {
  if (!c || !sgm_dev || !lam_dev || !dEdq || !dEdkappa || !iters || !resid || !info || restrt < 1 || maxit < 1 ||
      !(theta > 0.0)) {
    setErr("fbh_gradient_tree_dev: bad arguments");
    return 0;
  }
  if (!ensureAdjoint(c) || !ensureGmres(c, restrt)) return 0;
  ChargeSet &s = c->csCall;
  if (!stageCharges(c, s, nChar, pos, chr, nc, ctLevel, ctX, ctERad, ctKidStart, ctNKids, ctChgStart, ctNChgs,
                    ctChgIdx, ctOrder))
    return 0;
  const int n = c->n, nChunk = (n + GRAD_PCH - 1) / GRAD_PCH;
  GradBufs B;
  const size_t partN = std::max<size_t>(rhsTree ? 0 : (size_t)nChunk * nChar, std::max(n, c->nCube));
  double *g = B.get<double>(2 * (size_t)n), *lamBuf = B.get<double>((size_t)nc * c->nmMax),
         *phi = B.get<double>(nChar), *dq = B.get<double>(nChar), *part = B.get<double>(partN);
  if (!g || !lamBuf || !phi || !dq || !part) { setErr("fbh_gradient_tree_dev: allocation failed"); return 0; }
  // the adjoint: A^T lam = dE/dsgm (the walk's mode 2), with dE/dkappa at
  // fixed sgm (mode 3) from the same walk
  double eK = 0.0;
  if (!walkCharges(c, s, 4, theta, sgm_dev, g, &eK)) return 0;   // ends with the stream synchronized
  // dE/dq at fixed sgm (the transposed energy walk) does not need lam: a second
  // host thread runs it on a second stream while this one drives the adjoint GMRES
  if (!c->auxStream && !(c->auxStream = sharedQueueStream(1))) {
    setErr("fbh_gradient_tree_dev: no second stream");
    return 0;
  }
  int okAux = 1;
  std::thread aux([&] {
    if (hipSetDevice(c->dev) != hipSuccess) { okAux = 0; return; }
    hipStream_t st = c->auxStream;
    hipLaunchKernelGGL(k_fill, dim3((nc * c->nmMax + 255) / 256), dim3(256), 0, st, lamBuf, nc * c->nmMax, 0.0);
    hipLaunchKernelGGL(k_fill, dim3((nChar + 255) / 256), dim3(256), 0, st, phi, nChar, 0.0);
    okAux = adjWalk(c, s, 0, theta, sgm_dev, lamBuf, phi, st);   // ends with st synchronized
  });
  int ok = 1;
  if (lamFromRhs && hipMemcpyAsync(lam_dev, g, 2 * (size_t)n * sizeof(double), hipMemcpyDeviceToDevice,
                                   c->stream) != hipSuccess) {
    setErr("fbh_gradient_tree_dev: initial guess copy failed");
    ok = 0;
  }
  ok = ok && gmresCore(c, OPT64, g, lam_dev, restrt, tol, maxit, iters, resid, nullptr, info);
  aux.join();
  if (!ok || !okAux) return 0;
  // dE/dq: the right-hand side's part needs lam
  if (rhsTree && !adjWalk(c, s, 1, theta, lam_dev, lamBuf, phi, c->stream)) return 0;
  if (!adjEval(c, s, lamBuf, phi, dq)) return 0;
  if (!rhsTree) {   // R^T lam by the direct sums
    if (c->qorder == 1)
      hipLaunchKernelGGL(k_gradR_q1, dim3((nChar + 255) / 256, nChunk), dim3(256), 0, c->stream, c->pnl, n, nChar,
                         s.dPos, lam_dev, part, rhsFactor(c), c->Q.t[0], c->Q.t[0] * c->Q.w[0] * c->Q.w[0]);
    else
      hipLaunchKernelGGL(k_grad_q<false>, dim3((nChar + 255) / 256, nChunk), dim3(256), 0, c->stream, c->pnl, n,
                         nChar, s.dPos, (const double *)nullptr, lam_dev, part, c->kappa, c->eps, rhsFactor(c), c->Q,
                         c->qorder);
    hipLaunchKernelGGL(k_grad_sum, dim3((nChar + 255) / 256), dim3(256), 0, c->stream, part, nChunk, nChar, dq, 1);
  }
  HCK(hipGetLastError());
  HCK(hipMemcpyAsync(dEdq, dq, (size_t)nChar * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  // dE/dkappa
  double eNear = 0.0, eFar = 0.0;
  if (!kappaOperatorTerms(c, sgm_dev, lam_dev, part, &eNear, &eFar)) return 0;
  *dEdkappa = eK + eNear + eFar;
  return 1;
}
// per-stage device times of one operator application, stages run back to back
// on one stream (near, Q2M, M2M, M2L, L2L, L2P, combine, total); where L2L
// and L2P are one launch, its time is the L2P entry and L2L is zero.  Then the
// far chain alone with no events between its launches (ms[8]).
template <typename R>
static int profileT(fbh_ctx *c, const OpBufs<R> &B, const double *x, double *y, int iters, double *ms) {
  hipEvent_t ev[9];
  for (int i = 0; i < 9; i++) HCK(hipEventCreate(&ev[i]));
  for (int i = 0; i < 8; i++) ms[i] = 0.0;
  double s1 = (1.0 + c->eps) / 2.0, s2 = (1.0 + 1.0 / c->eps) / 2.0;
  int n2 = 2 * c->n;
  for (int it = 0; it < iters; it++) {
    hipEventRecord(ev[0], c->stream);
    nearPass(c, B, x, c->stream);
    hipEventRecord(ev[1], c->stream);
    q2mPass(c, B, x, c->stream);
    hipEventRecord(ev[2], c->stream);
    for (int lev = c->depth - 1; lev >= c->height; lev--)
      if (!m2mAxPass(c, B, lev, c->stream)) m2mPass(c, B, lev, c->stream);
    hipEventRecord(ev[3], c->stream);
    m2lPass(c, B, c->stream);
    hipEventRecord(ev[4], c->stream);
    if (l2lpFits<R>(c)) {   // one launch: all of it counts as L2P
      hipEventRecord(ev[5], c->stream);
      l2lpPass(c, B, c->stream);
    } else {
      for (int lev = c->height; lev < c->depth; lev++) l2lPass(c, B, lev, c->stream);
      hipEventRecord(ev[5], c->stream);
      l2pPass(c, B, c->stream);
    }
    hipEventRecord(ev[6], c->stream);
    hipLaunchKernelGGL(k_combine<R>, dim3((n2 + 255) / 256), dim3(256), 0, c->stream, c->area, x, B.far, B.nearv, y,
                       c->n, 1.0, 0.0, s1, s2, 1, 1);
    hipEventRecord(ev[7], c->stream);
    HCK(hipEventSynchronize(ev[7]));
    for (int i = 0; i < 7; i++) {
      float t = 0.f;
      hipEventElapsedTime(&t, ev[i], ev[i + 1]);
      ms[i] += t;
    }
    float t = 0.f;
    hipEventElapsedTime(&t, ev[0], ev[7]);
    ms[7] += t;
  }
  for (int i = 0; i < 8; i++) ms[i] /= iters;
  // the far chain alone, back to back: events only around the whole chain
  ms[8] = 0.0;
  for (int it = 0; it < iters; it++) {
    hipEventRecord(ev[0], c->stream);
    farPass(c, B, x, c->stream);
    hipEventRecord(ev[1], c->stream);
    HCK(hipEventSynchronize(ev[1]));
    float t = 0.f;
    hipEventElapsedTime(&t, ev[0], ev[1]);
    ms[8] += t;
  }
  ms[8] /= iters;
  for (int i = 0; i < 9; i++) hipEventDestroy(ev[i]);
  return 1;
}

extern "C" int fbh_profile(fbh_ctx *c, const double *x, double *y, int iters, double *ms) {
  if (!c || !x || !y || !ms || iters < 1 || c->direct) { setErr("fbh_profile: bad arguments"); return 0; }
  return profileT(c, bufs64(c), x, y, iters, ms);
}

extern "C" int fbh_profile32(fbh_ctx *c, const double *x, double *y, int iters, double *ms) {
  if (!c || !x || !y || !ms || iters < 1 || c->direct) { setErr("fbh_profile32: bad arguments"); return 0; }
  if (!c->fp32) { setErr("fbh_profile32: call fbh_enable_fp32 first"); return 0; }
  return profileT(c, bufs32(c), x, y, iters, ms);
}

#define NT_MAX 2048   // nearfield launches timed per fbh_near_timing(ctx, 1)
extern "C" int fbh_near_timing(fbh_ctx *c, int on) {
  if (!c) { setErr("fbh_near_timing: null context"); return 0; }
  if (on && c->ntEv.empty()) {
    c->ntEv.resize(2 * NT_MAX, nullptr);
    c->ntKind.resize(NT_MAX, 0);
    for (auto &e : c->ntEv) HCK(hipEventCreate(&e));
  }
  c->ntOn = on != 0;
  if (on) c->ntUsed = 0;
  return 1;
}

extern "C" int fbh_near_timing_read(fbh_ctx *c, long long *count, double *ms) {
  if (!c || !count || !ms) { setErr("fbh_near_timing_read: bad arguments"); return 0; }
  count[0] = count[1] = 0;
  ms[0] = ms[1] = 0.0;
  if (c->ntUsed == 0) return 1;
  HCK(hipEventSynchronize(c->ntEv[2 * c->ntUsed - 1]));
  for (int k = 0; k < c->ntUsed; k++) {
    float t = 0.f;
    HCK(hipEventElapsedTime(&t, c->ntEv[2 * k], c->ntEv[2 * k + 1]));
    count[c->ntKind[k]]++;
    ms[c->ntKind[k]] += t;
  }
  return 1;
}

extern "C" int fbh_stats(fbh_ctx *c, long long *out) {
  if (!c) return 0;
  out[0] = c->n; out[1] = c->nInter; out[2] = c->nM2L; out[3] = c->nmMax; out[4] = c->nmLeaf;
  out[5] = (long long)c->bytes; out[6] = c->nLinks; out[7] = c->nCube;
  out[8] = c->nStored; out[9] = c->nFlyTiles; out[10] = c->nTiles; out[11] = c->nSpec;
  out[12] = c->gridK;
  return 1;
}

extern "C" void *fbh_stream(fbh_ctx *c) { return c ? (void *)c->stream : nullptr; }

extern "C" double *fbh_dev_alloc(fbh_ctx *c, long long count) {
  double *p = nullptr;
  if (!c || count <= 0 || hipMalloc(&p, (size_t)count * sizeof(double)) != hipSuccess) {
    setErr("fbh_dev_alloc of %lld doubles failed", count);
    return nullptr;
  }
  return p;
}
extern "C" void fbh_dev_free(fbh_ctx *c, double *p) {
  if (c && c->stream) hipStreamSynchronize(c->stream);
  if (p) hipFree(p);
}
static int copyOn(fbh_ctx *c, double *dst, const double *src, long long count, hipMemcpyKind k) {
  if (!c || count < 0) { setErr("copy: bad arguments"); return 0; }
  HCK(hipMemcpyAsync(dst, src, (size_t)count * sizeof(double), k, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  return 1;
}
extern "C" int fbh_copy_h2d(fbh_ctx *c, double *d, const double *s, long long n) { return copyOn(c, d, s, n, hipMemcpyHostToDevice); }
extern "C" int fbh_copy_d2h(fbh_ctx *c, double *d, const double *s, long long n) { return copyOn(c, d, s, n, hipMemcpyDeviceToHost); }
extern "C" int fbh_copy_d2d(fbh_ctx *c, double *d, const double *s, long long n) { return copyOn(c, d, s, n, hipMemcpyDeviceToDevice); }

extern "C" int fbh_q2m_host(fbh_ctx *c, const double *sgm, double *mom) {
  if (!c || !sgm || !mom) { setErr("fbh_q2m_host: null argument"); return 0; }
  HCK(hipMemcpyAsync(c->xs, sgm, 2 * (size_t)c->n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  q2mPass(c, bufs64(c), c->xs, c->stream);
  HCK(hipGetLastError());
  HCK(hipMemcpyAsync(mom, c->mom, (size_t)c->nLeaf * 2 * c->nmMax * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  return 1;
}

extern "C" int fbh_m2l_host(fbh_ctx *c, const double *mom, double *lec) {
  if (!c || !mom || !lec) { setErr("fbh_m2l_host: null argument"); return 0; }
  HCK(hipMemcpyAsync(c->mom, mom, (size_t)c->nCube * 2 * c->nmMax * sizeof(double), hipMemcpyHostToDevice, c->stream));
  m2lPass(c, bufs64(c), c->stream);
  HCK(hipGetLastError());
  HCK(hipMemcpyAsync(lec, c->lec, (size_t)c->nCube * 4 * c->nmMax * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  return 1;
}

extern "C" int fbh_l2p_host(fbh_ctx *c, const double *lec, double alpha, double beta, double *pot) {
  if (!c || !lec || !pot) { setErr("fbh_l2p_host: null argument"); return 0; }
  size_t b = 2 * (size_t)c->n * sizeof(double);
  HCK(hipMemcpyAsync(c->lec, lec, (size_t)c->nLeaf * 4 * c->nmMax * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HCK(hipMemcpyAsync(c->ys, pot, b, hipMemcpyHostToDevice, c->stream));
  l2pPass(c, bufs64(c), c->stream);
  int n2 = 2 * c->n;
  hipLaunchKernelGGL(k_accum, dim3((n2 + 255) / 256), dim3(256), 0, c->stream, c->far, c->nearv, c->ys, n2, alpha, beta,
                     1, 0);
  HCK(hipGetLastError());
  HCK(hipMemcpyAsync(pot, c->ys, b, hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  return 1;
}
