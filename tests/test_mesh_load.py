"""The mesh reader's layouts: the parallel line-by-line path (one vertex or
face per line, no PQR line longer than one fgets read) and the fscanf / fgets
token-stream path it falls back to give the same panels and charges."""
import os
import subprocess

import numpy as np

from conftest import ROOT


def _rewrite(src, dst):
    with open(src + ".vert") as f:
        lines = f.readlines()
    with open(dst + ".vert", "w") as f:   # every vertex split over two lines, tabs between tokens
        f.writelines(lines[:3])
        for ln in lines[3:]:
            t = ln.split()
            if t:
                f.write("\t".join(t[:4]) + "\n  " + "  ".join(t[4:]) + "\n")
    with open(src + ".face") as f:
        lines = f.readlines()
    with open(dst + ".face", "w") as f:   # two faces per line
        f.writelines(lines[:3])
        body = [" ".join(ln.split()) for ln in lines[3:] if ln.strip()]
        for i in range(0, len(body), 2):
            f.write("   ".join(body[i:i + 2]) + "\n")
    with open(src + ".pqr") as f:
        lines = f.readlines()
    with open(dst + ".pqr", "w") as f:   # a remark longer than one fgets read up front
        f.write("REMARK " + "x" * 700 + "\n")
        f.writelines(lines)


def test_irregular_layouts_read_the_same(fa, tmp_path):
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    subprocess.check_call([os.path.join(ROOT, "cuda-fabipb_amd/bin/fbmesh"), "globule", a, "--level", "3",
                           "--radius", "12", "--charges", "inside", "--nq", "300", "--seed", "9"],
                          stdout=subprocess.DEVNULL)
    _rewrite(a, b)
    L = fa.host_lib()
    for geom in (False, True):
        ma, mb = fa.Mesh(a, geom), fa.Mesh(b, geom)
        for k in ("vtx", "nrm", "area", "face", "pos", "chr"):
            x, y = fa._arr(L.fb_mesh_array, ma.h, k), fa._arr(L.fb_mesh_array, mb.h, k)
            assert x.shape == y.shape and np.array_equal(x, y), k
