"""Host side of the product, on the CPU: mesh reader, octree and interaction
lists, charge tree, the generator, the C-ABI library's exports and the
command line driver's option handling (no GPU compute here)."""
import ctypes as C
import filecmp
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import CASES, ROOT, fixture, harness_cases
import golden_io

FL = CASES["flags"]


@pytest.mark.parametrize("name", harness_cases())
def test_mesh_and_tree_match_reference(fa, meshdir, name):
    """Panels (filter, normals, Heron areas), leaf order, near and M2L lists, cube
    centres: bit-identical to what the reference built from the same files."""
    F = fixture(name)
    mesh = fa.Mesh(str(meshdir / name))
    tree = fa.Tree(mesh, CASES["cases"][name]["t"], order=FL["p"], height=FL["H"], sep=FL["S"])
    n = int(F["counts"][0])
    assert tree.n == n and mesh.n == n
    if "panel_geom" in F:
        geo = F["panel_geom"].reshape(n, 13)
        assert np.array_equal(tree.a["vtx"].reshape(n, 9), geo[:, :9])
        assert np.array_equal(tree.a["nrm"].reshape(n, 3), geo[:, 9:12])
        assert np.array_equal(tree.a["area"], geo[:, 12])
    assert np.array_equal(tree.a["face"], F["panel_face"])
    for k, fk in [("leafStart", "leaf_start"), ("leafCount", "leaf_count"), ("nearOff", "near_off"),
                  ("nearSrc", "near_src"), ("m2lSrc", "m2l_src"), ("m2lDst", "m2l_dst"),
                  ("levStart", "lev_start"), ("levCount", "lev_count"), ("ordMom", "ord_mom"),
                  ("ordM2L", "ord_m2l")]:
        assert np.array_equal(tree.a[k], F[fk]), k
    ci = F["cube_i"].reshape(-1, 12)
    cx = F["cube_x"].reshape(-1, 4)
    assert np.array_equal(tree.a["cubeX"].reshape(-1, 3), cx[:, :3])
    assert np.array_equal(tree.a["eRad"], cx[:, 3])
    assert np.array_equal(tree.a["cubeLevel"], ci[:, 0]) and np.array_equal(tree.a["nKids"], ci[:, 1])
    assert np.array_equal(tree.a["parent"], ci[:, 2])
    assert np.array_equal(mesh.pos.ravel(), F["pos"]) and np.array_equal(mesh.chr, F["chr"])


@pytest.mark.parametrize("name", harness_cases())
def test_tree_layout_of_the_fmm_passes(fa, meshdir, name):
    """The cube numbering the device passes index by: the leaves are cubes
    0 .. nLeaf-1, all at the deepest level; every cube of the FMM levels has at
    most eight kids, numbered contiguously on the next level, each with it as
    parent, the kid ranges of a level tiling the next one; each leaf reaches the
    top FMM level in depth - height parent steps (the fused L2L + L2P launch
    walks exactly that chain, the M2M launches take one wave per kid)."""
    mesh = fa.Mesh(str(meshdir / name))
    t0 = CASES["cases"][name]["t"]
    for depth in sorted({max(FL["H"], t0 - 1), t0, t0 + 1}):
        tree = fa.Tree(mesh, depth, order=FL["p"], height=FL["H"], sep=FL["S"])
        a, D, H = tree.a, tree.depth, tree.height
        lev, ls, lc = a["cubeLevel"], a["levStart"], a["levCount"]
        ks, nk, par = a["kidStart"], a["nKids"], a["parent"]
        assert ls[D] == 0 and lc[D] == tree.nleaf
        assert np.all(lev[: tree.nleaf] == D)
        for l in range(H, D + 1):
            assert np.all(lev[ls[l]: ls[l] + lc[l]] == l)
        for l in range(H, D):
            cubes = np.arange(ls[l], ls[l] + lc[l])
            assert np.all((nk[cubes] >= 1) & (nk[cubes] <= 8))
            assert ks[cubes[0]] == ls[l + 1]
            assert np.array_equal(ks[cubes[1:]], ks[cubes[:-1]] + nk[cubes[:-1]])
            assert ks[cubes[-1]] + nk[cubes[-1]] == ls[l + 1] + lc[l + 1]
            assert np.array_equal(par[ls[l + 1]: ls[l + 1] + lc[l + 1]], np.repeat(cubes, nk[cubes]))
        c = np.arange(tree.nleaf)
        for _ in range(D - H):
            c = par[c]
        assert np.all(lev[c] == H)


@pytest.mark.parametrize("name", harness_cases())
def test_charge_tree_matches_reference(fa, meshdir, name):
    F = fixture(name)
    mesh = fa.Mesh(str(meshdir / name))
    tree = fa.Tree(mesh, CASES["cases"][name]["t"], order=FL["p"], height=FL["H"], sep=FL["S"])
    ct = tree.charge_tree()
    ref = golden_io.ctree_from_fixture(F, tree.depth)
    assert np.array_equal(ct["level"], ref["level"]) and np.array_equal(ct["nKids"], ref["nKids"])
    assert np.array_equal(ct["nChgs"], ref["nChgs"]) and np.array_equal(ct["chgIdx"], ref["chgIdx"])
    sel = ct["level"] >= FL["H"]   # coarser cubes carry no box in the reference and are never accepted
    assert np.array_equal(ct["x"][sel], ref["x"][sel]) and np.array_equal(ct["eRad"][sel], ref["eRad"][sel])
    assert np.array_equal(ct["order"], np.minimum(np.maximum(tree.a["ordMom"], 4), tree.max_order))


def test_counts_match_reference_solves():
    """Layout counts the reference printed (-B=1) for every recorded solve, the bench size included."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("fabipb_amd", os.path.join(ROOT, "cuda-fabipb_amd", "__init__.py"))
    fa = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(fa)
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        for name, case in CASES["cases"].items():
            F = fixture(name)
            if "cli" not in F:
                continue
            base = os.path.join(d, name)
            subprocess.check_call([os.path.join(ROOT, "cuda-fabipb_amd/bin/fbmesh"), case["mesh"][0], base] +
                                  case["mesh"][1:], stdout=subprocess.DEVNULL)
            mesh = fa.Mesh(base)
            tree = fa.Tree(mesh, case["t"], order=FL["p"], height=FL["H"], sep=FL["S"])
            cli = F["cli"]
            assert tree.nleaf == cli["leaf_cubes"] and len(tree.a["nearSrc"]) == cli["near_pairs"], name
            assert len(tree.a["m2lSrc"]) == cli["m2l_pairs"] and tree.ncube == cli["cubes"], name
            assert tree.n == cli["kept_panels"], name


def test_generator_is_deterministic(built, tmp_path):
    exe = os.path.join(ROOT, "cuda-fabipb_amd/bin/fbmesh")
    args = ["globule", "--level", "3", "--radius", "12", "--charges", "inside", "--nq", "30", "--seed", "9"]
    subprocess.check_call([exe, args[0], str(tmp_path / "a")] + args[1:], stdout=subprocess.DEVNULL)
    subprocess.check_call([exe, args[0], str(tmp_path / "b")] + args[1:], stdout=subprocess.DEVNULL)
    for ext in (".vert", ".face", ".pqr"):
        assert filecmp.cmp(tmp_path / ("a" + ext), tmp_path / ("b" + ext), shallow=False)
    # charges sit inside, total charge is an integer
    q = np.array([float(l.split()[-2]) for l in open(tmp_path / "a.pqr") if l.startswith("ATOM")])
    assert len(q) == 30 and abs(q.sum() - round(q.sum())) < 1e-12


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fbh_\w+)\s*\(", txt)))


def test_c_abi_exports_every_declared_symbol(built):
    """libfabipb_hip.so loads without a GPU and exports what include/fabipb_hip.h declares."""
    lib = C.CDLL(os.path.join(ROOT, "cuda-fabipb_amd/lib/libfabipb_hip.so"))
    names = _declared("fabipb_hip.h")
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "cuda-fabipb_amd/lib/libfabipb_hip.so")],
                         stdout=subprocess.PIPE, text=True).stdout
    exported = set(re.findall(r" T (fbh_\w+)", out))
    assert exported == set(names), exported ^ set(names)


def test_reference_binding_builds(built):
    """cuda-fabipb_amd/integration/gpu_backend_hip.c implements all 17 entry points of the
    reference's include/gpu_backend.h and links the reference's host code against
    libfabipb_hip.so (built by oracle/Makefile.ref where the reference is present)."""
    src = open(os.path.join(ROOT, "cuda-fabipb_amd/integration/gpu_backend_hip.c")).read()
    body = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = set(re.findall(r"^(?:int|void|const char \*)\s*(gpu\w+)\s*\(", body, flags=re.M))
    assert len(names) == 17, sorted(names)
    exe = os.path.join(ROOT, "oracle/_ref/fabipb_ref_hip")
    if not os.path.isdir("/root/reference/src") and not os.path.exists(exe):
        pytest.skip("reference sources absent and no prebuilt drop-in binary")
    assert os.path.exists(exe)
    out = subprocess.run(["nm", exe], stdout=subprocess.PIPE, text=True).stdout
    for n in names:
        assert re.search(r" T " + n + r"$", out, flags=re.M), n


def test_no_gpu_reports_unavailable(built):
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    lib = C.CDLL(os.path.join(ROOT, "cuda-fabipb_amd/lib/libfabipb_hip.so"))
    lib.fbh_available.restype = C.c_int
    lib.fbh_last_error.restype = C.c_char_p
    assert lib.fbh_available() == 0 and lib.fbh_last_error()


def test_cli_options(built, meshdir):
    exe = os.path.join(ROOT, "cuda-fabipb_amd/bin/fabipb")
    r = subprocess.run([exe, "-h"], stdout=subprocess.PIPE, text=True)
    assert r.returncode == 0 and "Usage" in r.stdout
    r = subprocess.run([exe, "-kappa=0.1", "x/y"], stderr=subprocess.PIPE, text=True)
    assert r.returncode == 1 and "unknown option" in r.stderr
    r = subprocess.run([exe, "-q=0", "x/y"], stderr=subprocess.PIPE, text=True)
    assert r.returncode == 1 and "outside the accepted range" in r.stderr
    r = subprocess.run([exe, "-g=0", str(meshdir / "sph3_ball")], stderr=subprocess.PIPE, text=True)
    assert r.returncode == 1 and "CPU path" in r.stderr
    r = subprocess.run([exe, "-P=4", str(meshdir / "sph3_ball")], stderr=subprocess.PIPE, stdout=subprocess.PIPE, text=True)
    assert r.returncode == 1 and "outside the accepted range" in r.stderr


def test_shell_generator(built, tmp_path):
    """fbmesh shell: two level-L surfaces, the inner one reversed (faces and normals point
    toward the centre, out of the wall); charges in the wall with an integer total; the
    fast radius function (recurrences) gives the globule generator's outer surface."""
    import numpy as np
    exe = os.path.join(ROOT, "cuda-fabipb_amd", "bin", "fbmesh")
    sh, gl = str(tmp_path / "sh"), str(tmp_path / "gl")
    subprocess.check_call([exe, "shell", sh, "--level", "3", "--radius", "250", "--inner", "210", "--amp", "0.03",
                           "--lmax", "12", "--charges", "wall", "--nq", "3000", "--sigma", "0.3", "--depth", "1.5",
                           "--seed", "5"], stdout=subprocess.DEVNULL)
    subprocess.check_call([exe, "globule", gl, "--level", "3", "--radius", "250", "--amp", "0.03", "--lmax", "12",
                           "--charges", "center", "--seed", "5"], stdout=subprocess.DEVNULL)
    v = np.loadtxt(sh + ".vert", skiprows=3, usecols=range(6))
    f = np.loadtxt(sh + ".face", skiprows=3, usecols=range(3), dtype=int) - 1
    g = np.loadtxt(gl + ".vert", skiprows=3, usecols=range(6))
    nv, nf = len(v) // 2, len(f) // 2
    assert len(v) == 2 * len(g) and nf == 1280
    assert np.max(np.abs(v[:nv, :3] - g[:, :3])) <= 1e-12 * 250 and np.max(np.abs(v[:nv, 3:] - g[:, 3:])) <= 1e-9
    P, N = v[:, :3], v[:, 3:]
    a, b, c = P[f[:, 0]], P[f[:, 1]], P[f[:, 2]]
    s = np.einsum("ij,ij->i", np.cross(b - a, c - a), (a + b + c) / 3)
    assert np.all(s[:nf] > 0) and np.all(s[nf:] < 0)
    assert np.all(np.einsum("ij,ij->i", N[:nv], P[:nv]) > 0) and np.all(np.einsum("ij,ij->i", N[nv:], P[nv:]) < 0)
    q = np.array([l.split()[-5:] for l in open(sh + ".pqr") if l.startswith("ATOM")], float)
    r = np.linalg.norm(q[:, :3], axis=1)
    rin, rout = np.linalg.norm(P[nv:], axis=1), np.linalg.norm(P[:nv], axis=1)
    assert len(q) == 3000 and rin.min() < r.min() and r.max() < rout.max()
    assert abs(q[:, 3].sum() - round(q[:, 3].sum())) < 1e-9 and 0.25 < q[:, 3].std() < 0.35


def test_capsid_charge_tree_order(built, tmp_path, monkeypatch):
    """src/fabipb_options.c:90-97, 220-229: above FABIPB_HUGE_CAPSID_ATOMS charges (or
    FABIPB_HUGE_CAPSID_PANELS panels) the charge-tree order is 3; below, max(ordMom, 4)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("fa_host", os.path.join(ROOT, "cuda-fabipb_amd", "__init__.py"))
    fa = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(fa)
    base = str(tmp_path / "g")
    subprocess.check_call([os.path.join(ROOT, "cuda-fabipb_amd", "bin", "fbmesh"), "globule", base, "--level", "3",
                           "--radius", "15", "--charges", "inside", "--nq", "40", "--seed", "3"], stdout=subprocess.DEVNULL)
    mesh = fa.Mesh(base)
    tree = fa.Tree(mesh, 3, order=8)
    L = fa.host_lib()
    assert [L.fb_charge_tree_order(tree.h, l) for l in range(4)] == [8, 8, 8, 8]
    monkeypatch.setenv("FABIPB_HUGE_CAPSID_ATOMS", "10")
    assert L.fb_charge_tree_order(tree.h, 3) == 3
    monkeypatch.delenv("FABIPB_HUGE_CAPSID_ATOMS")
    monkeypatch.setenv("FABIPB_HUGE_CAPSID_PANELS", "100")
    assert L.fb_charge_tree_order(tree.h, 3) == 3
    monkeypatch.setenv("FABIPB_CHARGE_TREE_ORDER", "5")
    assert L.fb_charge_tree_order(tree.h, 3) == 5
