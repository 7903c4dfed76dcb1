/*
 * fbo.h - the CPU oracle: a plain-C restatement of the reference FABIPB
 * Galerkin BIE matvec and its neighbours (kernels, Galerkin quadrature,
 * Taylor moments and translations, FMM apply, diagonal and block-LU
 * preconditioners,
 * restarted GMRES, direct right-hand side, panel-tree solvation energy).
 *
 * TEST INFRASTRUCTURE ONLY.  It is the checker the HIP path is compared with;
 * nothing in the product links, loads or calls it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * The oracle consumes a tree layout (it does not build trees): the flattened
 * arrays of src/fmm.c:234-404 and gk.h:120-146 in the reference.  Tests feed
 * it either the layout the reference harness recorded or the one the product
 * builds, so both the product tree and the product kernels are checked.
 */
#ifndef FBO_H
#define FBO_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double kappa, eps, eps1;     /* eps = eps_out/eps_in, eps1 = eps_in */
  int qorder;                  /* Gauss-Legendre order, 1..10 */
  int n;                       /* panels, in leaf-DFS (pnl->idx) order */
  const double *vtx;           /* 9n: v0 v1 v2 */
  const double *nrm;           /* 3n: panel normal */
  const double *area;          /* n */
  int depth, height, maxOrder;
  const int *ordMom, *ordM2L;  /* depth+1 */
  int nCube;                   /* FMM cubes: level depth first, then up to height */
  const int *levStart, *levCount;  /* depth+1 */
  const int *cubeLevel;        /* nCube */
  const double *cubeX;         /* 3*nCube, enclosing-box centres */
  const int *kidStart, *nKids; /* nCube: kids are contiguous in the next level */
  int nLeaf;                   /* leaves are FMM cubes 0..nLeaf-1 */
  const int *leafStart, *leafCount;
  const int *nearOff;          /* nLeaf+1 */
  const int *nearSrc;          /* near source leaf per pair, destination-major */
  int nM2L;
  const int *m2lSrc, *m2lDst;  /* FMM cube indices, grouped by destination */
  int nThreads;                /* worker threads for the nearfield apply (>=1) */
} fbo_layout;

typedef struct {
  int nc;                      /* charge-tree cubes, all levels 0..depth, DFS per level */
  int depth;
  const int *levStart, *levCount;  /* depth+1 */
  const int *level;            /* nc */
  const double *x;             /* 3*nc */
  const double *eRad;          /* nc */
  const int *kidStart, *nKids; /* nc */
  const int *chgStart, *nChgs; /* nc: leaves index chgIdx[chgStart..+nChgs) */
  const int *chgIdx;
  int nChar;
  const double *pos, *chr;
  double theta;                /* acceptance ratio, reference default 0.8 */
} fbo_ctree;

typedef struct fbo_ctx fbo_ctx;

/* single-function restatements */
void fbo_ker4(double kappa, double eps, const double *r, const double *nx, const double *ny, double *out);
int fbo_common_vtx(const double *vx, const double *vy, int *idxX, int *idxY);
void fbo_panel_ia0(const fbo_layout *L, int x, int y, double *out);
void fbo_derivs(double kappa, int order, const double *r, double *g0, double *gk);
int fbo_nmom(int order);
void fbo_gl_table(int q, double *t, double *w);   /* the Gauss-Legendre table in use */

/* context: owns the leaf Q2M matrices and per-cube expansions */
fbo_ctx *fbo_create(const fbo_layout *L);
void fbo_destroy(fbo_ctx *c);
const double *fbo_q2m(fbo_ctx *c, int leaf, int job);   /* job 0: moments, 1: noments */
/* pot = beta*pot + alpha*K*sgm; parts: 1 = far field, 2 = nearfield, 3 = both */
void fbo_apply(fbo_ctx *c, double alpha, const double *sgm, double beta, double *pot, int parts);
/* moments (2*nMomMax per cube) and locals (4*nMomMax per cube) after the last apply */
void fbo_expansions(fbo_ctx *c, double *mom, double *lec);
/* y = alpha*((D - K) x) + beta*y, D = diag((1+eps)/2*area, (1+1/eps)/2*area) */
void fbo_matvec(fbo_ctx *c, double alpha, const double *x, double beta, double *y);
void fbo_psolve_diag(fbo_ctx *c, double *out, const double *in);
/* preconditioner for fbo_psolve / fbo_gmres: 3 diagonal (default); 0, 1, 2 the
 * leaf-parent block LU (src/precond_fmm.c: the three modes solve the same
 * blocks).  Returns 0, -1 if the layout has no level depth-1, -2 on a zero pivot. */
int fbo_set_precond(fbo_ctx *c, int mode);
void fbo_psolve(fbo_ctx *c, double *out, const double *in);
/* -r=1/2: fbo_matvec / fbo_gmres use the dense direct operator (every pair's
 * panelIA0, n*n*32 bytes) instead of the FMM; 0 turns it off.  Returns 0, -1 if
 * the table cannot be allocated. */
int fbo_set_direct(fbo_ctx *c, int on);
/* restarted left-preconditioned GMRES; returns info; hist (maxit+1) optional */
int fbo_gmres(fbo_ctx *c, const double *b, double *x, int restart, double tol, int maxit,
              int *iters, double *resid, double *hist);
void fbo_rhs_direct(fbo_ctx *c, int nChar, const double *pos, const double *chr, double *b);
/* raw panel-tree energy (multiply by 2*pi*332.0716 for kcal/mol) */
double fbo_energy_tree(fbo_ctx *c, const fbo_ctree *T, const double *sgm);
/* charge-tree right-hand side (src/fabipb.c:633-742, src/treecode.c:283-450):
 * b[2n], the same walk as the energy with the Coulomb kernel, acceptance theta */
void fbo_rhs_tree(fbo_ctx *c, const fbo_ctree *T, double theta, double *b);
/* charge-tree moments as computed for the walk: nMom(maxOrder) per cube */
void fbo_ctree_moments(fbo_ctx *c, const fbo_ctree *T, double *mom);

#ifdef __cplusplus
}
#endif
#endif
