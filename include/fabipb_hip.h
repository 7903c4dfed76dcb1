/*
 * fabipb_hip.h - the flat C ABI of libfabipb_hip.so, the MI355X (gfx950)
 * implementation of FABIPB's Galerkin BIE matvec and its neighbours.
 *
 * Plain pointers and sizes only.  Host arrays are read during the call and
 * may be freed afterwards; device state lives in HBM inside an opaque
 * context until fbh_destroy.  Every entry point returns 1 on success and 0 on
 * failure with a message in fbh_last_error(); nothing falls back to a CPU
 * path.  Calls on one context are not re-entrant (like the reference's
 * file-scope caches); different contexts are independent.
 *
 * Each entry cites the reference interface it replaces
 * (include/gpu_backend.h of Jiahuic/CUDA-FABIPB and the host routines
 * behind it).
 */
#ifndef FABIPB_HIP_H
#define FABIPB_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fbh_ctx fbh_ctx;

/* replaces gpuBackendAvailable (include/gpu_backend.h:11): 1 if a gfx950 device is usable */
int fbh_available(void);

/* replaces gpuNearfieldLastError / gpuM2LLastError (include/gpu_backend.h:12-13) */
const char *fbh_last_error(void);

/*
 * The panel pass of the FMM tree on the device, in place of the host's
 * bounding box, goDown insertion and leaf boxes (src/gkSetup.c:90-183,
 * 264-301): each panel's leaf by the same midplane comparisons (strict >), the
 * panels stably sorted by leaf (depth-first leaf order, input order within a
 * leaf) and gathered in that order, and per leaf its octant-digit key
 * (coarsest digit first, 3 bits per level), panel count and vertex box
 * (lo[3], up[3]).  Host arrays: vtx[9n], nrm[3n], area[n], face[n] in; perm[n]
 * (input panel of each position), vtxR/nrmR/areaR/faceR out; leafKey,
 * leafCount[n], leafBox[6n] hold *nLeaf entries.  depth 1..10.  The cubes and
 * lists above the leaves: fb_tree_build_leaves (cuda-fabipb_amd/host/fb_host.h).
 */
int fbh_tree_panels(int n, const double *vtx, const double *nrm, const double *area, const int *face, int depth,
                    int *perm, double *vtxR, double *nrmR, double *areaR, int *faceR, int *nLeaf,
                    unsigned *leafKey, int *leafCount, double *leafBox);

/*
 * Build the device-resident operator for one mesh and tree.  Replaces the
 * per-sys-pointer caches built lazily by gpuNearfieldApply / gpuM2LApply /
 * gpuQ2MApply (src/gpu_backend_cuda.cu:328-798): geometry, layouts, leaf
 * Q2M/L2P matrices and nearfield coefficients are built on the device here.
 *   kappa, eps1, eps2   Debye parameter and dielectrics (eps = eps2/eps1)
 *   qorder              Gauss-Legendre order 1..10 (the -q option)
 *   n, vtx[9n], nrm[3n], area[n]  panels in leaf depth-first order
 *   depth, height, ordMom[depth+1], ordM2L[depth+1]
 *   nCube, levStart/levCount[depth+1], cubeLevel/cubeX[3]/kidStart/nKids/parent[nCube]
 *                       FMM cubes, finest level first (leaves are cubes 0..nLeaf-1)
 *   nLeaf, leafStart/leafCount[nLeaf], nearOff[nLeaf+1], nearSrc[nearOff[nLeaf]]
 *   m2lOff[nCube+1], m2lSrc[m2lOff[nCube]]   interaction lists by destination cube
 * stream: hipStream_t the context launches on (NULL: a stream of its own,
 * default priority).
 */
fbh_ctx *fbh_create(double kappa, double eps1, double eps2, int qorder,
                    int n, const double *vtx, const double *nrm, const double *area,
                    int depth, int height, const int *ordMom, const int *ordM2L,
                    int nCube, const int *levStart, const int *levCount,
                    const int *cubeLevel, const double *cubeX, const int *kidStart,
                    const int *nKids, const int *parent,
                    int nLeaf, const int *leafStart, const int *leafCount,
                    const int *nearOff, const int *nearSrc,
                    const int *m2lOff, const int *m2lSrc, void *stream);

/* replaces gpuReleaseMatvecCaches (include/gpu_backend.h:19) */
void fbh_destroy(fbh_ctx *ctx);

/*
 * y = alpha*[(D - K) x] + beta*y on device pointers (2n doubles each,
 * [potential | normal derivative]), D = diag((1+eps)/2 area, (1+1/eps)/2 area).
 * Replaces MtVmain (src/fabipb.c:1237-1287) -> applyFMM (src/fmm.c:710-838)
 * with all six stages on the device.  Asynchronous on the context stream.
 */
int fbh_matvec_dev(fbh_ctx *ctx, double alpha, const double *x, double beta, double *y);

/* the same on host arrays (copies in and out, synchronous) */
int fbh_matvec(fbh_ctx *ctx, double alpha, const double *x, double beta, double *y);

/*
 * Stage calls on host arrays, for callers that sequence the stages
 * themselves as src/fmm.c:710-838 does:
 *   fbh_near_apply  pot += alpha*K_near*sgm      (gpuNearfieldApply, gpu_backend.h:14)
 *   fbh_far_apply   pot = beta*pot + alpha*K_far*sgm, Q2M+M2M+M2L+L2L+L2P
 *                   (gpuQ2MApply + gpuM2LApply + gpuL2PApply, gpu_backend.h:16-18,
 *                   with M2M/L2L moved onto the device as well)
 */
int fbh_near_apply(fbh_ctx *ctx, double alpha, const double *sgm, double *pot);
int fbh_far_apply(fbh_ctx *ctx, double alpha, const double *sgm, double beta, double *pot);

/*
 * The far field one stage at a time on host arrays, for a host that runs
 * M2M and L2L itself between them, as src/fmm.c:737-831 does:
 *   fbh_q2m_host  leaf moments of sgm -> mom[nLeaf][2][nMom(maxOrder)]
 *                 (gpuQ2MApply, gpu_backend.h:17: fills cube->mom_pot/mom_dpdn)
 *   fbh_m2l_host  all FMM cubes' moments mom[nCube][2][nm] -> locals
 *                 lec[nCube][4][nm], overwritten (gpuM2LApply, gpu_backend.h:16)
 *   fbh_l2p_host  leaf locals lec[nLeaf][4][nm] -> pot = beta*pot + alpha*L2P(lec)
 *                 (gpuL2PApply, gpu_backend.h:18)
 * nm = nMom(maxOrder) is the stride; entries beyond a cube's own order are ignored.
 */
int fbh_q2m_host(fbh_ctx *ctx, const double *sgm, double *mom);
int fbh_m2l_host(fbh_ctx *ctx, const double *mom, double *lec);
int fbh_l2p_host(fbh_ctx *ctx, const double *lec, double alpha, double beta, double *pot);

/*
 * The dense direct operator of -r=1/2 (cpuDirectApply, src/direct_backend.c:
 * 93-208; gpuDirectApply, gpu_backend.h:15): every panel pair's panelIA0 kept in
 * one n*n store of KER[4] records (32 n^2 bytes; refused when it would not fit)
 * and applied row by row in panel order.  While on, fbh_matvec_dev and the
 * solvers use it in place of the FMM (no far field).  fbh_direct_apply is
 * gpuDirectApply on host arrays: pot = beta*pot + alpha*K*sgm (turns it on).
 */
int fbh_set_direct(fbh_ctx *ctx, int on);
int fbh_direct_apply(fbh_ctx *ctx, double alpha, double beta, const double *sgm, double *pot);

/* device expansions after the last far-field pass: nCube*2*nMom(maxOrder)
 * moments and nCube*4*nMom(maxOrder) locals (the reference's cube->mom_*, lec_k*) */
int fbh_expansions(fbh_ctx *ctx, double *mom, double *lec);
/* leaf Q2M matrices, panel-major: n*nMom(ordMom[depth]) each (Q2M0 moments, Q2M1 noments) */
int fbh_q2m_matrices(fbh_ctx *ctx, double *q0, double *q1);
/* nearfield coefficient of destination panel x and source panel y (KER[4], numQuad.c:629-697) */
int fbh_near_coeff(fbh_ctx *ctx, int x, int y, double *ker4);

/* diagonal (Jacobi) preconditioner on device pointers, in place allowed
 * (PtVfmmDiagonal, src/precond_fmm.c:664-728) */
int fbh_psolve_diag_dev(fbh_ctx *ctx, double *out, const double *in);

/*
 * The preconditioner GMRES uses, as the -P option picks it (PtVmain,
 * src/fabipb.c:1289-1309): 3 diagonal (the default of a new context), -1
 * none (a copy); 0, 1 and 2 the leaf-parent block LU of
 * setupPreconditioning / PtVfmm / PtVfmmCached / PtVfmmCachedLU
 * (src/precond_fmm.c:553-953; the three modes solve the same blocks).  One dense 2m x 2m block per cube at level depth-1,
 * built on the device and inverted once (batched LU); applying it streams the
 * inverses.  Replaces gpuBuildPrecondPairsBatched / gpuBuildPrecondDisjointBlock
 * (gpu_backend.h:38-45), which only assembled the blocks.  Fails if depth-1 <
 * height or a block is singular (the reference exits there).
 * fbh_psolve_dev applies the current one (out may equal in).
 */
int fbh_set_precond(fbh_ctx *ctx, int mode);
int fbh_psolve_dev(fbh_ctx *ctx, double *out, const double *in);

/*
 * Galerkin coefficients of listed panel pairs for a host that assembles the
 * preconditioner blocks itself (setupPreconditioning, src/precond_fmm.c:147-268):
 * (k0..k3)[p] = panelIA0(X = panel xIdx[p], Y = panel yIdx[p]) (numQuad.c:629-697,
 * whichever rule the pair's geometric case selects, the context's quadrature
 * order).  Host arrays, nPairs entries each.
 * fbh_pair_coeffs: panels of the context (leaf order, the reference's pnl->idx);
 *   replaces gpuBuildPrecondPairsBatched (gpu_backend.h:38-41; its pairSrc is X,
 *   pairDst is Y, src/precond_fmm.c:342-349).
 * fbh_panel_pair_coeffs: panels given here (vtx[9m], nrm[3m], area[m], edges from
 *   the vertices as src/input.c:743-747), no context; replaces
 *   gpuBuildPrecondDisjointBlock (gpu_backend.h:42-44; srcLocal is X, dstLocal Y).
 */
int fbh_pair_coeffs(fbh_ctx *ctx, long long nPairs, const int *xIdx, const int *yIdx,
                    double *k0, double *k1, double *k2, double *k3);
int fbh_panel_pair_coeffs(int nPanels, const double *vtx, const double *nrm, const double *area,
                          double kappa, double eps, int qorder, long long nPairs, const int *xLoc,
                          const int *yLoc, double *k0, double *k1, double *k2, double *k3);

/*
 * Restarted left-preconditioned GMRES with the preconditioner of
 * fbh_set_precond, all vectors in HBM (src/gmres.c:153-373, modified Gram-Schmidt :527-621;
 * convergence on ||M^-1 r|| / ||b||).  b, x are device pointers; x holds the
 * initial guess on entry.  hist (optional, host, maxit+restarts entries)
 * receives the residual history.  Returns info as the reference does
 * (0 converged, 1 iteration limit, 2 NaN).
 */
int fbh_gmres_dev(fbh_ctx *ctx, const double *b, double *x, int restart, double tol,
                  int maxit, int *iters, double *resid, double *hist, int *info);

/*
 * Precision form (no counterpart in the reference, whose solve is FP64 only).
 * fbh_enable_fp32 makes FP32 copies of the cached operator: nearfield
 * coefficients (16 bytes per interaction), leaf Q2M/L2P matrices and M2L
 * derivative tensors; expansions, translations and the nearfield sums then
 * run in FP32.  Needs the M2L tensor cache (it is kept whenever it fits).
 * fbh_matvec32_dev is fbh_matvec_dev on that operator (x, y stay FP64).
 * fbh_gmres_refined_dev solves to the same criterion as fbh_gmres_dev
 * (||M^-1 (b - A x)|| / ||b|| <= tol, A the FP64 operator) by iterative
 * refinement: corrections from GMRES on the FP32 operator (FP64 Krylov basis,
 * Gram-Schmidt and rotations), residuals in FP64.  iters = inner iterations,
 * outer = corrections (at most maxOuter), mv64 = FP64 products; hist (optional,
 * maxOuter entries) receives the outer residuals; info as for fbh_gmres_dev.
 */
int fbh_enable_fp32(fbh_ctx *ctx);
int fbh_matvec32_dev(fbh_ctx *ctx, double alpha, const double *x, double beta, double *y);
int fbh_gmres_refined_dev(fbh_ctx *ctx, const double *b, double *x, int restart, double tol,
                          int maxit, int maxOuter, int *iters, double *resid, double *hist,
                          int *info, int *outer, int *mv64);
/* fbh_gmres_refined_dev followed by fbh_energy32_res_dev on the solution (the
 * resident charge set, theta the walk's MAC), with each correction's energy
 * walk run on a second stream beside the FP64 residual product that decides
 * whether the correction is the last: the energy of the returned x.  No
 * counterpart in the reference (its solve and energy are FP64 and in turn). */
int fbh_gmres_refined_energy_dev(fbh_ctx *ctx, const double *b, double *x, int restart, double tol,
                                 int maxit, int maxOuter, double theta, int *iters, double *resid,
                                 double *hist, int *info, int *outer, int *mv64, double *energy_raw);

/* direct right-hand side on host arrays: b[2n] = fac * sum_q ... (setupRHS direct loop,
 * src/fabipb.c:789-798; gpuSetupRHS, gpu_backend.h:21), fac = 1/(4 pi eps1) unless
 * fbh_set_rhs_factor set another */
int fbh_rhs(fbh_ctx *ctx, int nChar, const double *pos, const double *chr, double *b);
int fbh_rhs_dev(fbh_ctx *ctx, int nChar, const double *pos, const double *chr, double *b_dev);
/* fbh_rhs_dev without any device allocation: no counterpart in the reference;
 * for callers that cannot allocate (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  The charges are staged in a static device
 * array of FBH_NOALLOC_MAX_CHARGES charges (8 MiB, one per GPU, held until the
 * process ends); larger sets are refused.  Calls that overlap on one GPU share
 * that array and race. */
#define FBH_NOALLOC_MAX_CHARGES 262144
int fbh_rhs_dev_noalloc(fbh_ctx *ctx, int nChar, const double *pos, const double *chr, double *b_dev);
/* the factor of every right-hand side of the context (direct, charge-tree, and the
 * R^T of the gradients): the fac argument of gpuSetupRHS / gpuChargeTreeRHS
 * (gpu_backend.h:21, 36), which src/fabipb.c:755 passes as 1/(4 pi eps1).
 * 0 restores that default. */
int fbh_set_rhs_factor(fbh_ctx *ctx, double fac);

/*
 * Panel-tree solvation energy (raw, before the 2*pi*332.0716 scale) of the
 * density sgm (device pointer, 2n) with the charge octree of the charges
 * (applyPanelChargeTreeEnergy, src/treecode.c:927-1017; gpuPanelChargeTreeEnergy,
 * gpu_backend.h:27).  Charge tree arrays as built by buildChargeTree
 * (src/chargeTree.c:228-307): nc cubes over levels 0..depth in per-level DFS
 * order; theta is the acceptance ratio (0.8).
 */
int fbh_energy_dev(fbh_ctx *ctx, int nChar, const double *pos, const double *chr,
                   int nc, const int *ctLevel, const double *ctX, const double *ctERad,
                   const int *ctKidStart, const int *ctNKids, const int *ctChgStart,
                   const int *ctNChgs, const int *ctChgIdx, const int *ctOrder,
                   double theta, const double *sgm_dev, double *energy_raw);

/*
 * Charge-tree right-hand side into b_dev (2n, device): the walk of the energy
 * with the Coulomb kernel (setupRHSTree, src/fabipb.c:633-742, and
 * rhsTreeWalkLocal / chgClusterEvalLocal, src/treecode.c:283-450; replaces
 * gpuChargeTreeRHS, gpu_backend.h:36).  Same charge-tree arrays as
 * fbh_energy_dev; theta is FABIPB_RHS_TREE_THETA's value (0.8).  The levels
 * the walk can accept must share one expansion order (true for fixed -p).
 */
int fbh_rhs_tree_dev(fbh_ctx *ctx, int nChar, const double *pos, const double *chr,
                     int nc, const int *ctLevel, const double *ctX, const double *ctERad,
                     const int *ctKidStart, const int *ctNKids, const int *ctChgStart,
                     const int *ctNChgs, const int *ctChgIdx, const int *ctOrder,
                     double theta, double *b_dev);

/*
 * Resident charge set: the charges and their charge tree (the arrays of
 * fbh_energy_dev) validated and copied to the device once, kept by the context
 * until the next fbh_set_charges or fbh_destroy.  The calls after it are
 * fbh_rhs_dev, fbh_rhs_tree_dev and fbh_energy_dev on that set with no host
 * traffic (and the precision form's FP32 variants); the set's cluster moments
 * and walk coefficients are built on its first walk and reused until it is
 * staged again: in the reference the charges live in host memory for the whole
 * solve (setupRHS / setupRHSTree, src/fabipb.c:633-800, and
 * applyPanelChargeTreeEnergy, src/treecode.c:927-1017, read them in place).
 */
int fbh_set_charges(fbh_ctx *ctx, int nChar, const double *pos, const double *chr,
                    int nc, const int *ctLevel, const double *ctX, const double *ctERad,
                    const int *ctKidStart, const int *ctNKids, const int *ctChgStart,
                    const int *ctNChgs, const int *ctChgIdx, const int *ctOrder);
int fbh_rhs_res_dev(fbh_ctx *ctx, double *b_dev);
/* the precision form's direct right-hand side of the resident set (q = 1): the
 * points moved in FP64 into the frame of each 256-panel block and rounded to
 * FP32, the kernel in FP32, sums in FP32 per 256 charges and FP64 across them;
 * other quadrature orders take the FP64 kernel */
int fbh_rhs32_res_dev(fbh_ctx *ctx, double *b_dev);
int fbh_rhs_tree_res_dev(fbh_ctx *ctx, double theta, double *b_dev);
int fbh_energy_res_dev(fbh_ctx *ctx, double theta, const double *sgm_dev, double *energy_raw);
/* the precision form's panel-tree energy of the resident set: the same walk with
 * the clusters' far polynomials rounded to FP32 and evaluated in FP32 (radial
 * values formed in FP64, each cluster's sums returned to FP64), the direct
 * leaves and every sum above them in FP64; no counterpart in the reference,
 * whose energy is FP64 only (src/treecode.c:617-695) */
int fbh_energy32_res_dev(fbh_ctx *ctx, double theta, const double *sgm_dev, double *energy_raw);

/* Differentiation form (no counterpart in the reference, which has no
 * adjoint; it differentiates what MtVmain, setupRHS and the panel-tree energy
 * compute: src/fabipb.c:1237-1287, 744-800, src/treecode.c:581-695).
 *
 * fbh_matvec_t_dev: y = alpha*A^T x + beta*y for the operator of
 * fbh_matvec_dev: the transposed nearfield (its own coefficient store, built
 * on first use) and the transposed FMM chain (Q2M^T M2M^T M2L^T L2L^T L2P^T).
 * Needs symmetric near and M2L lists and the M2L tensor cache; not the -r form.
 * fbh_gmres_t_dev: GMRES on A^T with the transposed preconditioner, the
 * stopping rule of fbh_gmres_dev.
 * fbh_energy_direct_dev: E_raw = g . sgm with the direct energy weights
 * g (2n; the panel-tree energy with every charge taken directly), g_dev may be
 * NULL.  Scale by 2*pi*332.0716 for kcal/mol, as the panel-tree energy.
 * fbh_gradient_dev: for E_raw(q, kappa) = g(q, kappa) . sgm, A sgm = b(q)
 * (direct right-hand side), solves A^T lam = g (lam_dev: initial guess in,
 * solution out) and returns dE_raw/dq (host array, nChar) and dE_raw/dkappa.
 */
int fbh_matvec_t_dev(fbh_ctx *ctx, double alpha, const double *x, double beta, double *y);
int fbh_gmres_t_dev(fbh_ctx *ctx, const double *b, double *x, int restart, double tol, int maxit,
                    int *iters, double *resid, double *hist, int *info);
int fbh_energy_direct_dev(fbh_ctx *ctx, int nChar, const double *pos, const double *chr,
                          const double *sgm_dev, double *g_dev, double *energy_raw);
int fbh_gradient_dev(fbh_ctx *ctx, int nChar, const double *pos, const double *chr,
                     const double *sgm_dev, double *lam_dev, int restart, double tol, int maxit,
                     double *dEdq, double *dEdkappa, int *iters, double *resid, int *info);
/* The exact gradient of the panel-tree energy the plain solve reports
 * (applyPanelChargeTreeEnergy, src/treecode.c:927-1017, with fbh_energy_dev's
 * charge tree and theta): E_raw = W(q, kappa) . u with A u = b(q), the
 * right-hand side direct (rhsTree 0) or the charge-tree one (rhsTree 1, theta
 * as the energy's).  The adjoint's right-hand side is W q (the walk's weights);
 * dE/dq comes from the transposed walk (fbh_tree_adjoint_dev of u, plus R^T lam
 * by direct sums or by the transposed right-hand-side walk of lam); dE/dkappa
 * from the walk's kappa derivative at fixed u plus lam^T (dK/dkappa) u.  Needs
 * one expansion order over the levels the walk can accept.  lam_dev: the
 * adjoint solution out; its initial guess is lam_dev itself, or with
 * lamFromRhs 1 the adjoint's right-hand side (as the plain solve starts from
 * x0 = b). */
int fbh_gradient_tree_dev(fbh_ctx *ctx, int nChar, const double *pos, const double *chr,
                          int nc, const int *ctLevel, const double *ctX, const double *ctERad,
                          const int *ctKidStart, const int *ctNKids, const int *ctChgStart,
                          const int *ctNChgs, const int *ctChgIdx, const int *ctOrder, double theta,
                          int rhsTree, const double *sgm_dev, double *lam_dev, int lamFromRhs, int restart,
                          double tol, int maxit, double *dEdq, double *dEdkappa, int *iters, double *resid,
                          int *info);
/* The transposed walk alone: dq_host[j] = d/dq_j at fixed densities of the
 * panel-tree energy of w_dev (kind 0, w = sgm) or of lam . (charge-tree
 * right-hand side) (kind 1, w = lam).  Exact for the walk's own truncation:
 * each far cluster contributes through its moments' monomials at the charge,
 * each unaccepted leaf directly.  Its sums over panels use atomic adds, so
 * repeated calls agree to rounding, not bit for bit. */
int fbh_tree_adjoint_dev(fbh_ctx *ctx, int nChar, const double *pos, const double *chr,
                         int nc, const int *ctLevel, const double *ctX, const double *ctERad,
                         const int *ctKidStart, const int *ctNKids, const int *ctChgStart,
                         const int *ctNChgs, const int *ctChgIdx, const int *ctOrder, double theta,
                         int kind, const double *w_dev, double *dq_host);

/* per-stage device timing over `iters` matvecs (HIP events on the context stream,
 * stages back to back): ms[0..7] = near, q2m, m2m, m2l, l2l, l2p, combine, whole
 * matvec (where L2L runs inside the L2P launch, l2l is the empty gap between two
 * events); ms[8] = the far chain (Q2M .. L2P) alone with events only around it.
 * ms holds 9 doubles. */
int fbh_profile(fbh_ctx *ctx, const double *x_dev, double *y_dev, int iters, double *ms);
/* the same for the FP32 operator (after fbh_enable_fp32) */
int fbh_profile32(fbh_ctx *ctx, const double *x_dev, double *y_dev, int iters, double *ms);

/* Nearfield launch timing inside ordinary work (no counterpart in the
 * reference; it serves measurement): on = 1 clears and starts recording an
 * event pair around each nearfield launch of every product that follows (up to
 * 2048 launches), on = 0 stops.  fbh_near_timing_read returns the launches
 * recorded since the start and their summed device time, [0] FP64, [1] FP32. */
int fbh_near_timing(fbh_ctx *ctx, int on);
int fbh_near_timing_read(fbh_ctx *ctx, long long *count, double *ms);

/* sizes (out holds 13): out[0]=n, [1]=nearfield interactions, [2]=M2L pairs,
 * [3]=nMom(maxOrder), [4]=nMom(ordMom[depth]), [5]=device bytes held, [6]=M2M/L2L
 * links, [7]=nCube, [8]=interactions with stored coefficients (streamed tiles),
 * [9]=nearfield tiles computed on the fly, [10]=nearfield tiles, [11]=corrected
 * special pairs (0 until the on-the-fly tiles need them), [12]=steps of the M2L
 * term grid (0: the M2L runs from per-lane item lists) */
int fbh_stats(fbh_ctx *ctx, long long *out);

/* the context's stream (hipStream_t), for callers that order their own work on it */
void *fbh_stream(fbh_ctx *ctx);

/* device vectors for C callers without a HIP toolchain (the fabipb driver):
 * allocate/free `count` doubles and copy between host and device, ordered on
 * the context stream; the copies complete before they return */
double *fbh_dev_alloc(fbh_ctx *ctx, long long count);
void fbh_dev_free(fbh_ctx *ctx, double *p);
int fbh_copy_h2d(fbh_ctx *ctx, double *dst, const double *src, long long count);
int fbh_copy_d2h(fbh_ctx *ctx, double *dst, const double *src, long long count);
int fbh_copy_d2d(fbh_ctx *ctx, double *dst, const double *src, long long count);

#ifdef __cplusplus
}
#endif
#endif
