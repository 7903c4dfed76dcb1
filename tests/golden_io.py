"""Loader for the fixtures recorded from the reference (tests/golden/*).

A fixture directory holds <name>.bin raw little-endian arrays and a
manifest.txt with "name dtype count" lines, as written by oracle/ref_harness.c.
"""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

_DT = {"f8": np.float64, "i4": np.int32, "i8": np.int64}


def load_dir(path):
    out = {}
    with open(os.path.join(path, "manifest.txt")) as fp:
        for line in fp:
            name, dt, cnt = line.split()
            a = np.fromfile(os.path.join(path, name + ".bin"), dtype=_DT[dt])
            assert a.size == int(cnt), (name, a.size, cnt)
            out[name] = a
    return out


def layout_from_fixture(F, qorder=1, depth=None, height=2):
    """Oracle layout dict from a harness fixture (the reference's own tree)."""
    counts = F["counts"]
    n, nleaf, nnear, ncube, nm2l, ngroup, maxorder, nchar = (int(v) for v in counts)
    kappa, eps1, eps2, sep = F["params"]
    geo = F["panel_geom"].reshape(n, 13)
    ci = F["cube_i"].reshape(ncube, 12)
    cx = F["cube_x"].reshape(ncube, 4)
    lev_start = F["lev_start"]
    depth = len(lev_start) - 1 if depth is None else depth
    kid_start = np.where(ci[:, 1] > 0, ci[:, 4], -1).astype(np.int32)
    for c in range(ncube):  # kids are contiguous in the next level (DFS layout)
        k = ci[c, 1]
        if k:
            assert np.all(ci[c, 4:4 + k] == ci[c, 4] + np.arange(k))
    return dict(
        kappa=float(kappa), eps=float(eps2 / eps1), eps1=float(eps1), qorder=qorder,
        vtx=geo[:, :9].copy(), nrm=geo[:, 9:12].copy(), area=geo[:, 12].copy(),
        depth=depth, height=height, maxOrder=maxorder,
        ordMom=F["ord_mom"], ordM2L=F["ord_m2l"],
        levStart=lev_start, levCount=F["lev_count"],
        cubeLevel=ci[:, 0].copy(), cubeX=cx[:, :3].copy(), kidStart=kid_start,
        nKids=ci[:, 1].copy(),
        leafStart=F["leaf_start"], leafCount=F["leaf_count"],
        nearOff=F["near_off"], nearSrc=F["near_src"],
        m2lSrc=F["m2l_src"], m2lDst=F["m2l_dst"],
    )


def ctree_from_fixture(F, depth):
    """Charge tree dict (all levels 0..depth, per-level DFS order)."""
    ci = F["ctree_i"].reshape(-1, 3)
    nc = ci.shape[0]
    nm = (F["ctree_x"].size // nc) - 4
    cx = F["ctree_x"].reshape(nc, 4 + nm)
    level = ci[:, 0].astype(np.int32)
    nkids = ci[:, 1].astype(np.int32)
    nchg = ci[:, 2].astype(np.int32)
    lev_start = np.zeros(depth + 1, np.int32)
    lev_count = np.zeros(depth + 1, np.int32)
    for lev in range(depth + 1):
        idx = np.nonzero(level == lev)[0]
        lev_start[lev] = idx[0] if idx.size else 0
        lev_count[lev] = idx.size
    kid_start = np.full(nc, -1, np.int32)
    for lev in range(depth):  # kids of consecutive parents are consecutive
        pos = lev_start[lev + 1]
        for c in range(lev_start[lev], lev_start[lev] + lev_count[lev]):
            kid_start[c] = pos
            pos += nkids[c]
    chg_start = np.zeros(nc, np.int32)
    pos = 0
    for c in range(lev_start[depth], lev_start[depth] + lev_count[depth]):
        chg_start[c] = pos
        pos += nchg[c]
    return dict(depth=depth, levStart=lev_start, levCount=lev_count, level=level,
                x=cx[:, :3].copy(), eRad=cx[:, 3].copy(), kidStart=kid_start, nKids=nkids,
                chgStart=chg_start, nChgs=nchg, chgIdx=F["ctree_idx"], mom=cx[:, 4:].copy())
