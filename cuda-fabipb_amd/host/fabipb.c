/*
 * fabipb - command line driver of the MI355X build, option-compatible with
 * the reference driver (src/fabipb.c:804-1229 of Jiahuic/CUDA-FABIPB).
 *
 *   fabipb [options] <panel-base-or-pqr-path>
 *
 * Sequence (as src/fabipb.c): read <base>.pqr/.vert/.face, build the octree
 * and interaction lists on the host, build the device operator, set up the
 * right-hand side, run GMRES with the -P preconditioner (diagonal, block LU
 * or none), then the panel-tree solvation energy.  Every numeric stage runs on
 * the GPU through include/fabipb_hip.h; the CPU path of the reference (-g=0)
 * is not part of this build, nor is meshing (MSMS/NanoShaper).  -r=1/2 (the
 * dense direct operator) both run it on the device.  Every option is parsed
 * and validated exactly as the reference does; then
 *   -g=0 (the CPU path), -c=1 and -C=1 (the reference's CPU-against-GPU
 *   applyFMM and preconditioner debug compares, src/fabipb.c:103-201, which need
 *   its CPU path) are refused with a message, so scripts fail loudly;
 *   -R and -d (meshing parameters), -G (the reference CUDA backend's nearfield
 *   storage strategy) and -Q (whether its Q2M/L2P run on the GPU) have no
 *   counterpart here (the mesh is read from existing files, there is one
 *   nearfield path and every stage runs on the device): they are accepted and
 *   a note on stderr says that they are ignored.
 */
#define _POSIX_C_SOURCE 200809L
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "fabipb_hip.h"
#include "fb_host.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static double now(void) {
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}

/* option helpers with the reference's strict validation (src/fabipb.c:220-283) */
static const char *optValue(const char *arg, const char *name) {
  size_t n = strlen(name);
  if (arg[0] != '-' || strncmp(arg + 1, name, n) != 0 || arg[1 + n] != '=') return NULL;
  return arg + 1 + n + 1;
}
static double optDouble(const char *arg, const char *v, double lo, double hi) {
  char *end;
  errno = 0;
  double out = strtod(v, &end);
  if (errno || end == v || *end) { fprintf(stderr, "Error: bad option '%s': expected a number after '='\n", arg); exit(1); }
  if (!(out >= lo && out <= hi)) {
    fprintf(stderr, "Error: bad option '%s': %g is outside the accepted range [%g, %g]\n", arg, out, lo, hi);
    exit(1);
  }
  return out;
}
static int optInt(const char *arg, const char *v, int lo, int hi) {
  char *end;
  errno = 0;
  long out = strtol(v, &end, 10);
  if (errno || end == v || *end) { fprintf(stderr, "Error: bad option '%s': expected an integer after '='\n", arg); exit(1); }
  if (out < lo || out > hi) {
    fprintf(stderr, "Error: bad option '%s': %ld is outside the accepted range [%d, %d]\n", arg, out, lo, hi);
    exit(1);
  }
  return (int)out;
}

static void usage(const char *prog) {
  printf("Usage: %s [options] <panel-base-or-pqr-path>\n", prog);
  printf("  -g=1       GPU (the only path of this build)\n");
  printf("  -B=0|1     quiet or benchmark/profiling output\n");
  printf("  -t=<lev>   tree depth          -H=<lev> coarsest FMM level (default 2)\n");
  printf("  -p=<ord>   FMM order (>0 fixed, <0 variable)  -pm=<ord> moment order offset\n");
  printf("  -S=<val>   separation ratio    -q=<ord> panel quadrature order (1..10)\n");
  printf("  -k=<val>   Debye kappa         -eps1=<val> -eps2=<val> dielectrics (-e1/-e2)\n");
  printf("  -o=<val>   GMRES tolerance     -a=<n> restart   -i=<n> max iterations\n");
  printf("  -P=-1..3   preconditioner: -1 none, 0/1/2 leaf-parent block LU, 3 diagonal (default auto)\n");
  printf("Environment: FABIPB_GMRES_INITIAL=rhs|zero FABIPB_GMRES_LOG_RESID=1 FABIPB_WRITE_SOLUTION=<csv>\n");
  printf("             FABIPB_ENERGY_MODE=panel-tree|compare FABIPB_STOP_AFTER_RHS FABIPB_STOP_AFTER_GMRES\n");
}

static void die(const char *what) {
  fprintf(stderr, "Error: %s: %s\n", what, fbh_last_error());
  exit(1);
}

/* acceptance ratio of the charge-tree walks: FABIPB_RHS_TREE_THETA /
 * FABIPB_ENERGY_TREE_THETA in (0, 1], else 0.8 (src/treecode.c:75-158) */
static double envTheta(const char *name) {
  const char *e = getenv(name);
  if (e && *e) {
    char *end = NULL;
    double v = strtod(e, &end);
    if (end != e && *end == '\0' && v > 0.0 && v <= 1.0) return v;
    fprintf(stderr, "Warning: ignoring invalid %s='%s'\n", name, e);
  }
  return 0.8;
}

int main(int argc, char **argv) {
  char panelfile[80] = {0};
  int order = -1, orderMom = 0, numItr = 100, arnoldiSz = 30, depth = 5, height = 2, qOrd = 1;
  int gpuMode = -1, bench = 0, meshFlag = 1, meshOnly = 0, matvecMode = 0, precond = 3, precondExplicit = 0;
  double sep = 0.8, tol = 1.0e-4, kappa = 0.1257, eps1 = 1.0, eps2 = 80.0, para = 332.0716;
  int sepExplicit = 0, dbgApply = 0, dbgPrecond = 0;
  const char *ignored[8];
  int nIgnored = 0;
  for (int i = 1; i < argc; i++) {
    const char *a = argv[i], *v;
    if (a[0] != '-') {
      if (strlen(a) >= sizeof(panelfile)) { fprintf(stderr, "Error: input path exceeds %zu characters: %s\n", sizeof(panelfile) - 1, a); exit(1); }
      strcpy(panelfile, a);
      continue;
    }
    if (!strcmp(a, "-h")) { usage(argv[0]); return 0; }
    if ((v = optValue(a, "S"))) { sep = optDouble(a, v, 1e-12, 10.0); sepExplicit = 1; }
    else if ((v = optValue(a, "o"))) tol = optDouble(a, v, 1e-300, 1.0);
    else if ((v = optValue(a, "a"))) arnoldiSz = optInt(a, v, 1, 100000);
    else if ((v = optValue(a, "i"))) numItr = optInt(a, v, 1, 1000000);
    else if ((v = optValue(a, "pm"))) orderMom = optInt(a, v, -32, 32);
    else if ((v = optValue(a, "p"))) order = optInt(a, v, -32, 32);
    else if ((v = optValue(a, "q"))) qOrd = optInt(a, v, 1, 10);
    else if ((v = optValue(a, "t"))) depth = optInt(a, v, -1, 64);
    else if ((v = optValue(a, "H"))) height = optInt(a, v, 0, 64);
    else if ((v = optValue(a, "R"))) { (void)optDouble(a, v, 1e-12, 1e6); if (nIgnored < 8) ignored[nIgnored++] = a; }
    else if ((v = optValue(a, "eps1")) || (v = optValue(a, "e1"))) eps1 = optDouble(a, v, 1e-12, 1e6);
    else if ((v = optValue(a, "eps2")) || (v = optValue(a, "e2"))) eps2 = optDouble(a, v, 1e-12, 1e6);
    else if ((v = optValue(a, "k"))) kappa = optDouble(a, v, 0.0, 1e6);
    else if ((v = optValue(a, "m"))) meshFlag = optInt(a, v, 1, 2);
    else if ((v = optValue(a, "M"))) meshOnly = optInt(a, v, 0, 1);
    else if ((v = optValue(a, "B"))) bench = optInt(a, v, 0, 1);
    else if ((v = optValue(a, "g"))) gpuMode = optInt(a, v, -1, 1);
    else if ((v = optValue(a, "c"))) dbgApply = optInt(a, v, 0, 1);
    else if ((v = optValue(a, "C"))) dbgPrecond = optInt(a, v, 0, 1);
    else if ((v = optValue(a, "r"))) matvecMode = optInt(a, v, 0, 2);
    else if ((v = optValue(a, "Q"))) { (void)optInt(a, v, 0, 1); if (nIgnored < 8) ignored[nIgnored++] = a; }
    else if ((v = optValue(a, "G"))) { (void)optInt(a, v, 0, 2); if (nIgnored < 8) ignored[nIgnored++] = a; }
    else if ((v = optValue(a, "P"))) { precond = optInt(a, v, -1, 3); precondExplicit = 1; }
    else if ((v = optValue(a, "d"))) { (void)optDouble(a, v, 1e-12, 1e9); if (nIgnored < 8) ignored[nIgnored++] = a; }
    else {
      fprintf(stderr, "Error: unknown option '%s'\nRun '%s -h' for the list of supported options.\n", a, argv[0]);
      exit(1);
    }
  }
  (void)meshFlag;
  if (!panelfile[0]) { fprintf(stderr, "Error: no input given\n"); usage(argv[0]); return 1; }
  if (depth < 1 || height < 0 || height > depth) { printf("Bad FMM level range: height=%d depth=%d\n", height, depth); return 0; }
  if (gpuMode == 0) { fprintf(stderr, "Error: -g=0 (the CPU path) is not part of this build; it runs on the GPU only\n"); return 1; }
  if (dbgApply || dbgPrecond) {
    fprintf(stderr, "Error: %s compares the reference's CPU %s with its GPU one; this build has no CPU path "
            "(its parity checks against the reference are the tests)\n", dbgApply ? "-c=1" : "-C=1",
            dbgApply ? "applyFMM" : "block preconditioner");
    return 1;
  }
  for (int k = 0; k < nIgnored; k++)
    fprintf(stderr, "Note: %s ignored: %s\n", ignored[k],
            ignored[k][1] == 'R' || ignored[k][1] == 'd'
                ? "a meshing parameter; the mesh is read from the existing .vert/.face files"
                : ignored[k][1] == 'G' ? "the reference CUDA backend's nearfield strategy; this build has one nearfield path"
                                       : "Q2M and L2P always run on the device here");
  if (!fbh_available()) die("GPU backend unavailable");

  /* inputs: <base>.pqr/.vert/.face; a path without '/' is looked up in test_proteins/ (src/input.c:24-36) */
  char base[256];
  if (strchr(panelfile, '/')) snprintf(base, sizeof(base), "%s", panelfile);
  else snprintf(base, sizeof(base), "test_proteins/%s", panelfile);
  double t0 = now(), tStage = t0;
  const char *pn = getenv("FABIPB_PANEL_NORMAL");
  fb_mesh *mesh = fb_mesh_load(base, pn && (!strcmp(pn, "geometric") || !strcmp(pn, "face")));
  if (!mesh) { fprintf(stderr, "Error: %s (meshing is not part of this build; write the mesh with bin/fbmesh)\n", fb_host_error()); return 1; }
  printf("Mesh input: panel=%s mesh_atoms=%d charge_atoms=%d\n", panelfile, mesh->nChar, mesh->nChar);
  printf("Mesh raw counts: vertices=%d faces=%d\n", mesh->nVert, mesh->nFace);
  {
    double s = 0.0;
    for (int i = 0; i < mesh->n; i++) s += mesh->area[i];
    printf("Mesh filtered panels: kept=%d area=%f\n", mesh->n, s);
  }
  if (meshOnly) { printf("Mesh-only mode: stopping after mesh generation.\n"); return 0; }
  /* preconditioner policy (src/fabipb_options.c:85-88, 198-265): capsid-scale
   * charge counts take the diagonal, high dielectric contrast the block LU */
  {
    const char *ev = getenv("FABIPB_HUGE_CAPSID_ATOMS");
    if (!ev || !*ev) ev = getenv("FABIPB_Q2M_HUGE_CAPSID_ATOMS");
    unsigned long long huge = (ev && *ev) ? strtoull(ev, NULL, 10) : 5000000ULL;
    int hugeCapsid = (unsigned long long)mesh->nChar > huge;
    if (!precondExplicit) {
      int hi = eps1 > 0.0 && eps1 <= 2.0 && eps2 / eps1 >= 40.0;
      precond = hugeCapsid ? 3 : hi ? 2 : 3;
    }
    if (!sepExplicit && hugeCapsid) sep = 1.2;   /* HUGE_CAPSID_SEP_RATIO, src/fabipb_options.c:216-218 */
  }
  double tLoad = now() - tStage;
  if (bench) {
    printf("----------------------------\n");
    printf("FMM variables: nLev=%d height=%d ord=%d SepRat=%lg qOrd=%d\n", depth, height, order, sep, qOrd);
    printf("GMRES variables: tol=%1.e arnoldiSz=%d maxIt=%d\n", tol, arnoldiSz, numItr);
    printf("kappa=%f, eps1=%f, eps2=%f\n", kappa, eps1, eps2);
    printf("GPU mode=1 (0=CPU, 1=GPU)\n");
    if (precondExplicit)
      printf("Preconditioner mode=%d (-1=disabled, 0=original, 1=cached-blocks, 2=cached-LU, 3=diagonal)\n", precond);
    else
      printf("Preconditioner mode=auto (-1=disabled, 0=original, 1=cached-blocks, 2=cached-LU, 3=diagonal)\n");
  }

  tStage = now();
  fb_tree *t = fb_tree_build(mesh, depth, height, order, orderMom, sep);
  if (!t) { fprintf(stderr, "Error: %s\n", fb_host_error()); return 1; }
  double tTree = now() - tStage;
  tStage = now();
  fbh_ctx *ctx = fbh_create(kappa, eps1, eps2, qOrd, t->n, t->vtx, t->nrm, t->area, depth, height, t->ordMom, t->ordM2L,
                            t->nCube, t->levStart, t->levCount, t->cubeLevel, t->cubeX, t->kidStart, t->nKids,
                            t->parent, t->nLeaf, t->leafStart, t->leafCount, t->nearOff, t->nearSrc, t->m2lOff,
                            t->m2lSrc, NULL);
  if (!ctx) die("operator setup failed");
  double tSetup = now() - tStage;
  if (bench) {
    long long st[16];
    fbh_stats(ctx, st);
    printf("Flattened apply layout: leaf-cubes=%d near-pairs=%d\n", t->nLeaf, t->nNear);
    printf("Flattened interaction layout: m2l-pairs=%d dst-groups=%d\n", t->nM2L, t->nM2LGroups);
    printf("Flattened FMM cube layout: cubes=%d\n", t->nCube);
    printf("Maximum number of elements in finest cluster: %d\n", t->maxPrecondBlock);
    printf("Device operator: nearfield interactions=%lld device bytes=%lld\n", st[1], st[5]);
    printf("----------------------------\n");
  }

  int n = t->n, n2 = 2 * n;
  double *b = fbh_dev_alloc(ctx, n2), *x = fbh_dev_alloc(ctx, n2);
  if (!b || !x) die("device vectors");
  tStage = now();
  /* charge octree for the tree right-hand side and the energy (src/chargeTree.c:228-307) */
  fb_ctree *ct = fb_ctree_build(mesh, depth);
  if (!ct) die("charge tree");
  int *ord = malloc(sizeof(int) * (size_t)(depth + 1));
  for (int l = 0; l <= depth; l++) ord[l] = fb_charge_tree_order(t, l);
  const int treeRhs = fb_should_use_tree_rhs(n, mesh->nChar, 1);
  const double rhsTheta = envTheta("FABIPB_RHS_TREE_THETA"), energyTheta = envTheta("FABIPB_ENERGY_TREE_THETA");
  if (bench) {
    printf("setupRHS direct pairs: panels=%d charges=%d pairs=%llu mode=%s", n, mesh->nChar,
           (unsigned long long)n * (unsigned long long)mesh->nChar, treeRhs ? "tree" : "direct");
    if (treeRhs) printf(" theta=%g", rhsTheta);
    printf("\n");
  }
  if (treeRhs) {   /* src/fabipb.c:733-742 (setupRHSTree) */
    if (!fbh_rhs_tree_dev(ctx, mesh->nChar, mesh->pos, mesh->chr, ct->nc, ct->level, ct->x, ct->eRad, ct->kidStart,
                          ct->nKids, ct->chgStart, ct->nChgs, ct->chgIdx, ord, rhsTheta, b))
      die("setupRHS (tree)");
  } else if (!fbh_rhs_dev(ctx, mesh->nChar, mesh->pos, mesh->chr, b)) {
    die("setupRHS");
  }
  double tRhs = now() - tStage;
  if (getenv("FABIPB_STOP_AFTER_RHS")) {
    printf("FABIPB_STOP_AFTER_RHS set: stopping after setupRHS.\n");
    return 0;
  }
  const char *init = getenv("FABIPB_GMRES_INITIAL");
  if (init && !strcmp(init, "zero")) {
    double *z = calloc((size_t)n2, sizeof(double));
    fbh_copy_h2d(ctx, x, z, n2);
    free(z);
  } else if (!fbh_copy_d2d(ctx, x, b, n2)) die("initial guess");
  int iters = 0, info = 0;
  double resid = 0.0, *hist = calloc((size_t)2 * numItr + 2, sizeof(double));
  tStage = now();
  if (!fbh_set_precond(ctx, precond)) die("preconditioner setup");
  if (matvecMode != 0 && !fbh_set_direct(ctx, 1)) die("dense direct operator");   /* -r=1/2 */
  if (!fbh_gmres_dev(ctx, b, x, arnoldiSz, tol, numItr, &iters, &resid, hist, &info)) die("GMRES");
  double tGmres = now() - tStage;
  if (getenv("FABIPB_GMRES_LOG_RESID") && atoi(getenv("FABIPB_GMRES_LOG_RESID")) > 0)
    for (int i = 0; i < iters; i++) printf("GMRES residual: iter=%d resid=%e\n", i + 1, hist[i]);
  printf("GMRES status: info=%d iterations=%d final-residual=%e\n", info, iters, resid);

  double *xs = malloc(sizeof(double) * (size_t)n2);
  fbh_copy_d2h(ctx, xs, x, n2);
  const char *sol = getenv("FABIPB_WRITE_SOLUTION");
  if (sol && *sol) {   /* src/fabipb.c:448-473 */
    FILE *fp = fopen(sol, "w");
    if (fp) {
      fprintf(fp, "idx,mesh_face_idx,x,y,z,area,component0,component1\n");
      for (int i = 0; i < n; i++) {
        const double *v = &t->vtx[9 * i];
        double c[3];
        for (int k = 0; k < 3; k++) c[k] = 0.33333333333333333333333333 * (v[k] + v[3 + k] + v[6 + k]);
        fprintf(fp, "%d,%d,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g\n", i, t->face[i], c[0], c[1], c[2], t->area[i],
                xs[i], xs[n + i]);
      }
      fclose(fp);
    } else fprintf(stderr, "Warning: cannot write solution dump '%s'\n", sol);
  }
  if (getenv("FABIPB_STOP_AFTER_GMRES")) {
    printf("FABIPB_STOP_AFTER_GMRES set: skipping post-GMRES treecode energy.\n");
    printf("ttl time: %f, gmres-its=%d\n", now() - t0, iters);
    printf("solvation energy: skipped\n");
    return 0;
  }
  tStage = now();
  double raw = 0.0;
  if (!fbh_energy_dev(ctx, mesh->nChar, mesh->pos, mesh->chr, ct->nc, ct->level, ct->x, ct->eRad, ct->kidStart,
                      ct->nKids, ct->chgStart, ct->nChgs, ct->chgIdx, ord, energyTheta, x, &raw))
    die("panel-tree energy");
  double tEnergy = now() - tStage;
  const char *em = getenv("FABIPB_ENERGY_MODE");
  if (em && !strcmp(em, "compare"))
    printf("Energy compare: panel-tree raw=%.17g panel-tree-time=%.6f\n", raw, tEnergy);
  double ptl = raw * 2.0 * M_PI * para;
  printf("ttl time: %f, gmres-its=%d\n", now() - t0, iters);
  printf("solvation energy: %f\n", ptl);
  if (bench)
    printf("Top-level stage times (s): loadPanel=%.6f tree=%.6f setupDevice=%.6f setupRHS=%.6f gmres=%.6f treecode=%.6f\n",
           tLoad, tTree, tSetup, tRhs, tGmres, tEnergy);
  fbh_dev_free(ctx, b);
  fbh_dev_free(ctx, x);
  fbh_destroy(ctx);
  fb_ctree_free(ct);
  fb_tree_free(t);
  fb_mesh_free(mesh);
  free(hist); free(xs); free(ord);
  return 0;
}
