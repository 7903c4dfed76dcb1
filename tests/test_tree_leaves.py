"""The FMM tree built from its leaves (fb_tree_build_leaves, the host half of the
device tree build) against the host's own build (fb_tree_build), on the CPU.

The leaves come from a numpy restatement of the device panel pass
(fbh_tree_panels): each panel's centroid against the cube midplanes level by
level with the host's arithmetic (numpy evaluates a + b * c as two rounded
operations, as the host and the device kernel do), a stable sort by the octant
digits, and per leaf its key, panel count and vertex box.  Every array of the
two trees must be identical.
"""
import numpy as np
import pytest

from conftest import CASES


def panel_pass(vtx, depth):
    v = vtx.reshape(-1, 3, 3)
    lo, up = vtx.reshape(-1, 3).min(axis=0), vtx.reshape(-1, 3).max(axis=0)
    cen = 0.33333333333333333333333333 * (v[:, 0, :] + v[:, 1, :] + v[:, 2, :])
    ln = (up[0] - lo[0]) if (up[0] - lo[0]) > (up[1] - lo[1]) else (up[1] - lo[1])
    ln = (up[2] - lo[2]) if (up[2] - lo[2]) > ln else ln
    side = []
    for _ in range(depth + 1):
        side.append(ln)
        ln *= 0.5
    ijk = np.zeros((len(cen), 3), np.int64)
    key = np.zeros(len(cen), np.uint32)
    for lev in range(depth):
        bit = (cen > lo + (ijk + 0.5) * side[lev]).astype(np.int64)
        key = (key << np.uint32(3)) | (bit[:, 0] + 2 * bit[:, 1] + 4 * bit[:, 2]).astype(np.uint32)
        ijk = 2 * ijk + bit
    perm = np.argsort(key, kind="stable").astype(np.int32)
    ks = key[perm]
    uniq, start, count = np.unique(ks, return_index=True, return_counts=True)
    vs = v[perm]
    box = np.empty((len(uniq), 6))
    for L, (s0, c) in enumerate(zip(start, count)):
        pts = vs[s0:s0 + c].reshape(-1, 3)
        box[L, :3], box[L, 3:] = pts.min(axis=0), pts.max(axis=0)
    return perm, uniq.astype(np.uint32), count.astype(np.int32), box.ravel()


@pytest.mark.parametrize("name", ["sph3_ball", "glob4", "cfg1_sphere", "cfg4_grad"])
def test_tree_from_leaves_matches_host_build(fa, meshdir, name):
    import ctypes as C
    L = fa.host_lib()
    case = CASES["cases"][name]
    fl = dict(CASES["flags"], **case.get("flags", {}))
    mesh = fa.Mesh(str(meshdir / name))
    depth, height, order, sep = case["t"], fl["H"], fl["p"], fl["S"]
    host = fa.Tree(mesh, depth, order=order, height=height, sep=sep)
    arr = lambda k: fa._arr(L.fb_mesh_array, mesh.h, k)
    vtx, nrm, area, face = arr("vtx"), arr("nrm"), arr("area"), arr("face")
    perm, key, cnt, box = panel_pass(vtx, depth)
    h = L.fb_tree_build_leaves(mesh.h, depth, height, order, 0, sep, fa._i(perm),
                               fa._d(np.ascontiguousarray(vtx.reshape(-1, 9)[perm].ravel())),
                               fa._d(np.ascontiguousarray(nrm.reshape(-1, 3)[perm].ravel())),
                               fa._d(np.ascontiguousarray(area[perm])), fa._i(np.ascontiguousarray(face[perm])),
                               len(key), key.ctypes.data_as(C.c_void_p), fa._i(cnt), fa._d(box))
    assert h, L.fb_host_error().decode()
    try:
        for k in fa.TREE_KEYS:
            got = fa._arr(L.fb_tree_array, h, k)
            assert np.array_equal(got, host.a[k]), k
    finally:
        L.fb_tree_free(h)


def test_tree_from_leaves_rejects_bad_leaves(fa, meshdir):
    import ctypes as C
    L = fa.host_lib()
    mesh = fa.Mesh(str(meshdir / "sph3_ball"))
    arr = lambda k: fa._arr(L.fb_mesh_array, mesh.h, k)
    vtx, nrm, area, face = arr("vtx"), arr("nrm"), arr("area"), arr("face")
    perm, key, cnt, box = panel_pass(vtx, 3)
    cnt = cnt.copy()
    cnt[0] += 1   # counts no longer add up to the panels
    h = L.fb_tree_build_leaves(mesh.h, 3, 2, 6, 0, 0.8, fa._i(perm), fa._d(vtx), fa._d(nrm), fa._d(area),
                               fa._i(face), len(key), key.ctypes.data_as(C.c_void_p), fa._i(cnt), fa._d(box))
    assert not h and b"leaf counts" in L.fb_host_error()
