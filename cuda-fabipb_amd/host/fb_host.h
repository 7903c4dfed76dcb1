/*
 * fb_host.h - host-side data model of the fabipb MI355X build.
 *
 * Everything here is flat arrays (structure of arrays, integer links), laid
 * out the way the device consumes it: panels in leaf depth-first order
 * (the reference's pnl->idx order, gkSetup.c:498), FMM cubes level by level
 * from the finest level up (fmm.c:295-324), near pairs destination-major
 * (fmm.c:234-268) and M2L pairs grouped by destination (fmm.c:326-404).
 */
#ifndef FB_HOST_H
#define FB_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

/* panels and charges as read from <base>.vert/.face/.pqr (src/input.c:320-821) */
typedef struct {
  int n;            /* panels kept by the face filter */
  double *vtx;      /* 9n, input order */
  double *nrm;      /* 3n panel normals */
  double *area;     /* n */
  int *face;        /* n: 0-based index of the face in the .face file */
  int nChar;
  double *pos;      /* 3 nChar */
  double *chr;      /* nChar */
  int nVert, nFace; /* raw counts */
} fb_mesh;

/* octree, interaction lists and flattened layouts (src/gkSetup.c, src/fmm.c:234-524) */
typedef struct {
  int n, depth, height, maxOrder;
  int *ordMom, *ordM2L;          /* depth+1 */
  int *perm;                      /* n: input panel of idx i */
  double *vtx, *nrm, *area;       /* panels in idx order (9n, 3n, n) */
  int *face;                      /* n */
  int nCube;                      /* FMM cubes, levels depth..height */
  int *levStart, *levCount;       /* depth+1 */
  int *cubeLevel, *kidStart, *nKids, *parent;
  double *cubeX, *eRad;           /* 3 nCube, nCube */
  int nLeaf;
  int *leafStart, *leafCount;     /* nLeaf */
  int nNear;
  int *nearOff, *nearSrc;         /* nLeaf+1, nNear (destination-major) */
  int nM2L, nM2LGroups;
  int *m2lSrc, *m2lDst;           /* nM2L */
  int *m2lOff;                    /* nCube+1: pairs of destination cube c */
  int maxLeafPanels;
  int nPrecondBlocks, maxPrecondBlock;
  int nChar;                      /* charges of the mesh (capsid-scale policy) */
} fb_tree;

/* charge octree (src/chargeTree.c:228-307): all levels 0..depth, DFS per level */
typedef struct {
  int nc, depth;
  int *levStart, *levCount;       /* depth+1 */
  int *level, *kidStart, *nKids;  /* nc */
  double *x, *eRad;               /* 3 nc, nc */
  int *chgStart, *nChgs;          /* nc (leaves only) */
  int *chgIdx;                    /* nChar */
} fb_ctree;

/* input (src/input.c): reads <base>.pqr/.vert/.face; returns NULL and sets fb_host_error() */
fb_mesh *fb_mesh_load(const char *base, int useGeometricNormals);
void fb_mesh_free(fb_mesh *m);
const char *fb_host_error(void);

/* tree: order > 0 fixed order, order < 0 variable order (gkSetup.c:42-81) */
fb_tree *fb_tree_build(const fb_mesh *m, int depth, int height, int order, int orderMom, double sepRatio);
/* the same tree from the leaves of a device panel pass (fbh_tree_panels): perm and the panel arrays in
 * leaf order, and per leaf (nLeaf, increasing keys) its octant-digit key, panel count and box (lo, up) */
fb_tree *fb_tree_build_leaves(const fb_mesh *m, int depth, int height, int order, int orderMom, double sepRatio,
                              const int *perm, const double *vtx, const double *nrm, const double *area,
                              const int *face, int nLeaf, const unsigned *leafKey, const int *leafCount,
                              const double *leafBox);
void fb_tree_free(fb_tree *t);

fb_ctree *fb_ctree_build(const fb_mesh *m, int depth);
void fb_ctree_free(fb_ctree *t);

/* policy helpers (src/fabipb_options.c) */
int fb_should_use_tree_rhs(int nPnls, int nChar, int gpuMode);
int fb_charge_tree_order(const fb_tree *t, int level);

/* accessors for foreign-function callers: name -> pointer, *count elements */
const void *fb_tree_array(const fb_tree *t, const char *name, int *count, int *isDouble);
const void *fb_ctree_array(const fb_ctree *t, const char *name, int *count, int *isDouble);
const void *fb_mesh_array(const fb_mesh *m, const char *name, int *count, int *isDouble);

#ifdef __cplusplus
}
#endif
#endif
