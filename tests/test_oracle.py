"""The CPU oracle pinned before it is trusted.

1. Against the fixtures recorded from the reference CPU path (tests/golden):
   kernels, every panelIA0 geometric case, derivative tensors, leaf Q2M
   matrices, nearfield, applyFMM, the full matvec, RHS, diagonal
   preconditioner, panel-tree energy, and whole GMRES solves.
2. Against what the mathematics fixes, independently of the reference:
   closed-form derivatives of 1/r and exp(-kappa r)/r, the null operator at
   eps_in = eps_out and kappa = 0, the symmetry of the disjoint KER0/KER3
   blocks, Gauss-Legendre moments, panel moments by direct quadrature, the
   FMM against the dense (all-near) operator, and the analytic Born/Kirkwood
   energies.
"""
import math
import os

import numpy as np
import pytest

from conftest import CASES, cli_cases, fixture, harness_cases, rel, splitmix_vec
import golden_io
import kirkwood
import oracle_binding as O

FL = CASES["flags"]


@pytest.fixture(scope="module")
def oracles(built):
    cache = {}

    def get(name):
        if name not in cache:
            F = fixture(name)
            if "panel_geom" not in F:
                return None
            cache[name] = (F, O.Oracle(golden_io.layout_from_fixture(F)))
        return cache[name]
    return get


SMALL = [c for c in harness_cases() if c != "cfg1_sphere"]


@pytest.mark.parametrize("name", SMALL)
def test_kernel_and_derivs_exact(oracles, name):
    F, o = oracles(name)
    lay = golden_io.layout_from_fixture(F)
    k = F["ker4"].reshape(5, 7)
    for i in range(5):
        got = O.ker4(lay["kappa"], lay["eps"], k[i, 4:], lay["nrm"][0], lay["nrm"][-1])
        assert np.array_equal(got, k[i, :4])
    p = lay["maxOrder"]
    nm = O.nmom(p)
    D = F["derivs"].reshape(4, 2, nm)
    for i, r in enumerate([[3.0, -1.0, 2.0], [0.7, 0.2, -0.4], [-12.0, 5.0, 1.0], [0.0, 0.0, 4.0]]):
        g0, gk = O.derivs(lay["kappa"], p, r)
        assert np.array_equal(g0, D[i, 0]) and np.array_equal(gk, D[i, 1])


@pytest.mark.parametrize("name", SMALL)
def test_panel_ia0_every_case_exact(oracles, name):
    F, o = oracles(name)
    idx = F["ia0_idx"].reshape(-1, 3)
    val = F["ia0_val"].reshape(-1, 4)
    assert {0, 1, 2, -2, 3} <= set(int(c) for c in idx[:, 2])
    for (x, y, c), v in zip(idx, val):
        assert np.array_equal(o.panel_ia0(int(x), int(y)), v), (x, y, c)


@pytest.mark.parametrize("name", SMALL)
def test_q2m_exact(oracles, name):
    F, o = oracles(name)
    lay = golden_io.layout_from_fixture(F)
    p = lay["ordMom"][lay["depth"]]
    nl = min(4, len(lay["leafStart"]))
    q0 = np.concatenate([o.q2m(l, 0, int(lay["leafCount"][l]), p) for l in range(nl)])
    q1 = np.concatenate([o.q2m(l, 1, int(lay["leafCount"][l]), p) for l in range(nl)])
    assert np.array_equal(q0, F["q2m0"]) and np.array_equal(q1, F["q2m1"])


@pytest.mark.parametrize("name", SMALL)
def test_matvec_pieces(oracles, name):
    F, o = oracles(name)
    x = F["mv_x"]
    assert np.array_equal(o.apply(x, parts=2), F["mv_near"])        # nearfield: same order, bit-exact
    assert rel(o.apply(x, parts=3), F["mv_fmm"]) <= 1e-14           # BLAS summation order only
    assert rel(o.matvec(x), F["mv_y"]) <= 1e-14
    if "pc_y" in F:
        assert np.array_equal(o.psolve(x), F["pc_y"])
    if "mv_mom" in F:
        mom, lec = o.expansions(len(F["cube_i"]) // 12)
        assert np.max(np.abs(mom - F["mv_mom"])) <= 1e-14 * np.max(np.abs(F["mv_mom"]))
        assert np.max(np.abs(lec - F["mv_lec"])) <= 1e-14 * np.max(np.abs(F["mv_lec"]))


@pytest.mark.parametrize("name", SMALL)
def test_rhs_and_energy(oracles, name):
    F, o = oracles(name)
    pos, chr_ = F["pos"].reshape(-1, 3), F["chr"]
    assert np.array_equal(o.rhs(pos, chr_), F["rhs"])
    depth = len(F["lev_start"]) - 1
    ct = golden_io.ctree_from_fixture(F, depth)
    mom = o.ctree_moments(ct, pos, chr_).reshape(len(ct["level"]), -1)
    sel = ct["level"] >= FL["H"]
    m = min(mom.shape[1], ct["mom"].shape[1])
    assert np.max(np.abs(mom[sel, :m] - ct["mom"][sel, :m])) <= 1e-14 * np.max(np.abs(ct["mom"][sel]))
    e = o.energy(ct, pos, chr_, F["en_sigma"])
    assert abs(e - F["en_raw"][0]) <= 1e-13 * abs(F["en_raw"][0])


@pytest.mark.parametrize("name", SMALL)
def test_gmres_solve_matches_reference(oracles, name):
    """The oracle's GMRES reproduces the reference's iterations, residuals and energy."""
    F, o = oracles(name)
    cli = F["cli"]
    b = F["rhs"]
    x, info, its, res, hist = o.gmres(b, b.copy(), restart=FL["a"], tol=FL["o"], maxit=FL["i"])
    assert info == cli["info"] and its == cli["iterations"]
    h = np.array(cli["resid_history"])
    assert np.allclose(hist[:len(h)], h, rtol=1e-5, atol=0)   # the reference prints %e (7 digits)
    depth = len(F["lev_start"]) - 1
    ct = golden_io.ctree_from_fixture(F, depth)
    e = o.energy(ct, F["pos"].reshape(-1, 3), F["chr"], x)
    assert abs(e - cli["panel_tree_raw"]) <= 1e-11 * abs(cli["panel_tree_raw"])


@pytest.mark.parametrize("name", ["glob4_a3", "glob4_a3_i5"])
def test_gmres_restart_and_limit_match_reference(oracles, name):
    """GMRES(3) with restarts, and the iteration limit (info 1), against the reference's
    log on the glob4 mesh: iterations, info, the residual after each step and after each
    restart (src/gmres.c:295-340), and the energy of the returned iterate."""
    from conftest import CASES, case_flags
    F, o = oracles(CASES["cases"][name]["same_as"])
    cli = fixture(name)["cli"]
    fl = case_flags(name)
    b = F["rhs"]
    x, info, its, res, hist = o.gmres(b, b.copy(), restart=fl["a"], tol=fl["o"], maxit=fl["i"])
    assert info == cli["info"] and its == cli["iterations"]
    h = np.array(cli["resid_history"])
    assert len(h) > its   # restarts add the recomputed residual to the log
    assert np.allclose(hist[:len(h)], h, rtol=1e-5, atol=0)
    depth = len(F["lev_start"]) - 1
    ct = golden_io.ctree_from_fixture(F, depth)
    e = o.energy(ct, F["pos"].reshape(-1, 3), F["chr"], x)
    assert abs(e - cli["panel_tree_raw"]) <= 1e-11 * abs(cli["panel_tree_raw"])


@pytest.mark.parametrize("name", ["sph3_ball_p2", "glob4_p2"])
def test_gmres_block_precond_matches_reference(oracles, name):
    """-P=2: GMRES left-preconditioned with the leaf-parent block LU (src/precond_fmm.c:
    553-652, 861-953) against the reference's log: iterations, info, residual history
    (the residual test uses ||M^-1 r||, so it pins the block solve itself) and energy."""
    from conftest import CASES, case_flags
    F, o = oracles(CASES["cases"][name]["same_as"])
    cli = fixture(name)["cli"]
    fl = case_flags(name)
    o.set_precond(fl["P"])
    try:
        b = F["rhs"]
        x, info, its, res, hist = o.gmres(b, b.copy(), restart=fl["a"], tol=fl["o"], maxit=fl["i"])
    finally:
        o.set_precond(3)
    assert info == cli["info"] and its == cli["iterations"]
    h = np.array(cli["resid_history"])
    assert np.allclose(hist[:len(h)], h, rtol=1e-5, atol=0)
    depth = len(F["lev_start"]) - 1
    ct = golden_io.ctree_from_fixture(F, depth)
    e = o.energy(ct, F["pos"].reshape(-1, 3), F["chr"], x)
    assert abs(e - cli["panel_tree_raw"]) <= 1e-11 * abs(cli["panel_tree_raw"])


def test_gmres_direct_operator_matches_reference(oracles):
    """-r=2: GMRES on the dense direct operator (src/direct_backend.c) against the
    reference's log: iterations, residual history and energy."""
    from conftest import CASES, case_flags
    name = "sph3_ball_r2"
    F, o = oracles(CASES["cases"][name]["same_as"])
    cli = fixture(name)["cli"]
    fl = case_flags(name)
    o.set_direct(True)
    try:
        b = F["rhs"]
        x, info, its, res, hist = o.gmres(b, b.copy(), restart=fl["a"], tol=fl["o"], maxit=fl["i"])
    finally:
        o.set_direct(False)
    assert info == cli["info"] and its == cli["iterations"]
    h = np.array(cli["resid_history"])
    assert np.allclose(hist[:len(h)], h, rtol=1e-5, atol=0)
    depth = len(F["lev_start"]) - 1
    ct = golden_io.ctree_from_fixture(F, depth)
    e = o.energy(ct, F["pos"].reshape(-1, 3), F["chr"], x)
    assert abs(e - cli["panel_tree_raw"]) <= 1e-11 * abs(cli["panel_tree_raw"])
    assert rel(x, fixture(name)["solution"]) <= 1e-10


@pytest.mark.parametrize("name", ["sph3_ball_tr", "glob4_tr"])
def test_tree_rhs_matches_reference(oracles, name):
    """The charge-tree right-hand side (FABIPB_FORCE_TREE_RHS=1) row by row against the
    reference's RHS sample, and GMRES + energy from it against the reference solve."""
    from conftest import CASES, case_flags
    F, o = oracles(CASES["cases"][name]["same_as"])
    G = fixture(name)
    cli = G["cli"]
    fl = case_flags(name)
    depth = len(F["lev_start"]) - 1
    ct = golden_io.ctree_from_fixture(F, depth)
    pos, chr_ = F["pos"].reshape(-1, 3), F["chr"]
    b = o.rhs_tree(ct, pos, chr_)
    ref = np.fromfile(os.path.join(os.path.dirname(__file__), "golden", name, "rhs_tree.bin"))
    assert np.max(np.abs(b - ref)) <= 1e-13 * np.max(np.abs(ref))
    assert rel(b, F["rhs"]) > 1e-4   # really the tree, not the direct sum
    x, info, its, res, hist = o.gmres(b, b.copy(), restart=fl["a"], tol=fl["o"], maxit=fl["i"])
    assert info == cli["info"] and its == cli["iterations"]
    e = o.energy(ct, pos, chr_, x)
    assert abs(e - cli["panel_tree_raw"]) <= 1e-11 * abs(cli["panel_tree_raw"])


def test_block_precond_inverts_its_blocks(oracles):
    """The block solve inverts the operator restricted to each leaf-parent cube: for x
    supported on one cube, M^-1 (M x) = x, with M x formed from panelIA0 directly."""
    F, o = oracles("glob4")
    lay = golden_io.layout_from_fixture(F)
    o.set_precond(2)
    try:
        n = len(lay["area"])
        lev = lay["depth"] - 1
        cb = lay["levStart"][lev]
        k0 = lay["kidStart"][cb]
        s0 = lay["leafStart"][k0]
        m = int(sum(lay["leafCount"][k0 + k] for k in range(lay["nKids"][cb])))
        rng = np.random.default_rng(5)
        xs, xn = rng.standard_normal(m), rng.standard_normal(m)
        s1, s2 = (1 + lay["eps"]) / 2, (1 + 1 / lay["eps"]) / 2
        y = np.zeros(2 * n)
        for r in range(m):   # row r: sum over c of the operator the reference factors, A(r, c) = B[c][r]
            for c in range(m):
                k = o.panel_ia0(s0 + r, s0 + c)
                y[s0 + r] += -k[1] * xs[c] - k[3] * xn[c]
                y[n + s0 + r] += -k[0] * xs[c] - k[2] * xn[c]
            y[s0 + r] += s1 * lay["area"][s0 + r] * xs[r]
            y[n + s0 + r] += s2 * lay["area"][s0 + r] * xn[r]
        z = o.psolve_any(y)
        assert np.allclose(z[s0:s0 + m], xs, rtol=0, atol=1e-10 * np.abs(xs).max())
        assert np.allclose(z[n + s0:n + s0 + m], xn, rtol=0, atol=1e-10 * np.abs(xn).max())
        mask = np.ones(2 * n, bool)
        mask[s0:s0 + m] = mask[n + s0:n + s0 + m] = False
        assert np.all(z[mask] == 0.0)
    finally:
        o.set_precond(3)


def test_cfg1_matvec(fa, meshdir):
    """20,480-panel sphere: oracle nearfield and matvec on the product's tree, whose lists
    and panel order equal the reference's, against the recorded vectors."""
    F = fixture("cfg1_sphere")
    mesh = fa.Mesh(str(meshdir / "cfg1_sphere"))
    tree = fa.Tree(mesh, 3, order=6)
    for k, fk in [("leafStart", "leaf_start"), ("nearSrc", "near_src"), ("m2lSrc", "m2l_src"), ("face", "panel_face")]:
        assert np.array_equal(tree.a[k], F[fk]), k
    o = O.Oracle(tree.oracle_layout(0.1257, 1.0, 80.0, 1))
    assert np.array_equal(o.apply(F["mv_x"], parts=2), F["mv_near"])
    assert rel(o.matvec(F["mv_x"]), F["mv_y"]) <= 1e-14
    assert np.array_equal(o.rhs(mesh.pos, mesh.chr), F["rhs"])


# ------------------------------------------------------------ mathematics

def test_gauss_legendre_table():
    """The oracle's 15-digit table: sum w t^k = 1/(k+1) for k < 2q, and the nodes and
    weights agree with numpy's Gauss-Legendre rule mapped to [0,1]."""
    for q in range(1, 11):
        t, w = O.gl_table(q)
        x, wx = np.polynomial.legendre.leggauss(q)
        assert np.allclose(t, (x + 1) / 2, rtol=0, atol=2e-15)
        assert np.allclose(w, wx / 2, rtol=0, atol=2e-15)
        for k in range(2 * q):
            assert abs(np.sum(w * t ** k) - 1.0 / (k + 1)) <= 2e-14, (q, k)


def test_derivatives_closed_form():
    """D^a of G0 = 1/(4 pi r) and Gk = exp(-k r)/(4 pi r) against closed forms."""
    kappa = 0.37
    r = np.array([0.8, -1.3, 2.1])
    R = np.linalg.norm(r)
    g0, gk = O.derivs(kappa, 2, r)
    # index order: degree 0: (0,0,0); degree 1: (0,0,1) (0,1,0) (1,0,0); degree 2: (0,0,2) (0,1,1) (0,2,0) (1,0,1) (1,1,0) (2,0,0)
    fp = 1 / (4 * np.pi)
    assert abs(g0[0] - fp / R) <= 1e-16
    assert abs(gk[0] - fp * np.exp(-kappa * R) / R) <= 1e-16
    grad0 = -fp * r / R ** 3
    assert np.allclose([g0[3], g0[2], g0[1]], grad0, rtol=1e-14, atol=0)
    gradk = -fp * r * (1 + kappa * R) * np.exp(-kappa * R) / R ** 3
    assert np.allclose([gk[3], gk[2], gk[1]], gradk, rtol=1e-14, atol=0)
    # Laplacian of 1/r is zero away from 0: d2/dz2 + d2/dy2 + d2/dx2
    assert abs(g0[4] + g0[6] + g0[9]) <= 1e-15 * abs(g0[4])
    # Yukawa: (Laplacian - kappa^2) Gk = 0
    assert abs(gk[4] + gk[6] + gk[9] - kappa ** 2 * gk[0]) <= 1e-14 * abs(gk[4])
    # mixed second derivative d2/dxdy = 3 x y / r^5 (times 1/4pi)
    assert abs(g0[8] - fp * 3 * r[0] * r[1] / R ** 5) <= 1e-14 * abs(g0[8])


def test_null_operator(fa, meshdir):
    """eps_in = eps_out and kappa = 0 make G0 = Gk, so K = 0 and (D-K) x = area * x.
    The nearfield cancels exactly (kernel.c:138-141 subtracts equal numbers); the
    far field to roundoff, because the radial recurrences of G0 and Gk
    (kernel.c:84-104) round differently."""
    mesh = fa.Mesh(str(meshdir / "sph3_ball"))
    tree = fa.Tree(mesh, 3, order=6)
    o = O.Oracle(tree.oracle_layout(0.0, 2.0, 2.0, 1))
    x = splitmix_vec(99, 2 * tree.n)
    assert np.array_equal(o.apply(x, parts=2), np.zeros(2 * tree.n))
    y = o.matvec(x)
    area = np.concatenate([tree.a["area"], tree.a["area"]])
    assert rel(y, area * x) <= 1e-16


def test_disjoint_blocks_symmetric(fa, meshdir):
    """KER0 and KER3 of two disjoint panels are symmetric under x <-> y at q=1."""
    mesh = fa.Mesh(str(meshdir / "glob4"))
    tree = fa.Tree(mesh, 3, order=6)
    lay = tree.oracle_layout(0.1257, 1.0, 80.0, 1)
    o = O.Oracle(lay)
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(200):
        x, y = rng.integers(0, tree.n, 2)
        iX, iY = (np.zeros(3, np.int32) for _ in range(2))
        if x == y or O.lib().fbo_common_vtx(O.dptr(lay["vtx"][x].copy()), O.dptr(lay["vtx"][y].copy()),
                                            O.iptr(iX), O.iptr(iY)) != 0:
            continue
        a, b = o.panel_ia0(int(x), int(y)), o.panel_ia0(int(y), int(x))
        assert abs(a[0] - b[0]) <= 1e-15 * abs(a[0]) and abs(a[3] - b[3]) <= 1e-13 * abs(a[3])
        checked += 1
    assert checked > 100


def test_panel_moments_by_quadrature(fa, meshdir):
    """Leaf Q2M columns = (1/a!) int_panel (x - c)^a dS, by a 10-point Duffy rule."""
    mesh = fa.Mesh(str(meshdir / "sph3_ball"))
    tree = fa.Tree(mesh, 3, order=4)
    o = O.Oracle(tree.oracle_layout(0.1257, 1.0, 80.0, 1))
    p = 4
    nm = O.nmom(p)
    idx = [(i1, i2, d - i1 - i2) for d in range(p + 1) for i1 in range(d + 1) for i2 in range(d - i1 + 1)]
    x10, w10 = np.polynomial.legendre.leggauss(10)
    t, w = (x10 + 1) / 2, w10 / 2
    leaf = 0
    cnt = int(tree.a["leafCount"][leaf])
    q0 = o.q2m(leaf, 0, cnt, p).reshape(cnt, nm)
    q1 = o.q2m(leaf, 1, cnt, p).reshape(cnt, nm)
    c = tree.a["cubeX"].reshape(-1, 3)[leaf]
    for j in range(cnt):
        pnl = int(tree.a["leafStart"][leaf]) + j
        v = tree.a["vtx"].reshape(-1, 9)[pnl].reshape(3, 3)
        nrm = tree.a["nrm"].reshape(-1, 3)[pnl]
        A = tree.a["area"][pnl]
        mom = np.zeros(nm)
        nom = np.zeros(nm)
        for a_, (u, wu) in enumerate(zip(t, w)):
            for vv, wv in zip(t, w):
                # Duffy: x = v0 + u (a2 + v a0) with jacobian 2A u
                xq = v[0] + u * ((v[1] - v[0]) + vv * (v[2] - v[1]))
                d = xq - c
                for k, (i1, i2, i3) in enumerate(idx):
                    f = d[0] ** i1 * d[1] ** i2 * d[2] ** i3 / (math.factorial(i1) * math.factorial(i2) * math.factorial(i3))
                    mom[k] += 2 * A * u * wu * wv * f
                    # noments: sum_k n_k (d^(a-e_k) / (a-e_k)!)
                    g = 0.0
                    for ax, e in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
                        b1, b2, b3 = i1 - e[0], i2 - e[1], i3 - e[2]
                        if min(b1, b2, b3) >= 0:
                            g += nrm[ax] * d[0] ** b1 * d[1] ** b2 * d[2] ** b3 / (
                                math.factorial(b1) * math.factorial(b2) * math.factorial(b3))
                    nom[k] += 2 * A * u * wu * wv * g
        assert np.allclose(q0[j], mom, rtol=1e-12, atol=1e-14 * np.abs(mom).max())
        assert np.allclose(q1[j], nom, rtol=1e-12, atol=1e-14 * np.abs(nom).max())


def test_fmm_converges_to_dense(fa, meshdir):
    """With every leaf in every near list (separation ratio 1e-9: (r1+r2)/d is never
    below it, gkSetup.c:342-348) the apply is the dense Galerkin operator.  At
    quadrature order 3, where the far pairs are integrated accurately, the FMM at
    S=0.8 converges to it geometrically in the Taylor order."""
    mesh = fa.Mesh(str(meshdir / "sph3_ball"))
    dense_tree = fa.Tree(mesh, 3, order=2, sep=1e-9)
    assert len(dense_tree.a["m2lSrc"]) == 0
    od = O.Oracle(dense_tree.oracle_layout(0.1257, 1.0, 80.0, 3))
    x = splitmix_vec(5, 2 * dense_tree.n)
    yd = od.apply(x)
    errs = []
    for p in (2, 4, 6, 8):
        tr = fa.Tree(mesh, 3, order=p, sep=0.8)
        assert np.array_equal(tr.a["perm"], dense_tree.a["perm"])
        of = O.Oracle(tr.oracle_layout(0.1257, 1.0, 80.0, 3))
        errs.append(rel(of.apply(x), yd))
    assert errs[-1] < 5e-4 and all(b < 0.3 * a for a, b in zip(errs, errs[1:])), errs


def test_born_and_kirkwood(built):
    """Reference energies on spheres sit at discretisation distance from the analytic ones."""
    e_born = fixture("sph3_born")["cli"]["panel_tree_raw"] * 2 * math.pi * 332.0716
    assert abs(e_born - kirkwood.born(1.0, 10.0, 1.0, 80.0, 0.1257)) <= 1e-2 * abs(e_born)
    F = fixture("cfg1_sphere")
    e = F["cli"]["panel_tree_raw"] * 2 * math.pi * 332.0716
    ek = kirkwood.energy(F["pos"].reshape(-1, 3), F["chr"], 10.0, 1.0, 80.0, 0.1257)
    assert abs(e - ek) <= 2e-3 * abs(ek)


def test_kirkwood_limits():
    """The series itself: Born+Debye for one central charge; kappa=0 bracket."""
    assert abs(kirkwood.energy([[0, 0, 0]], [1.0], 10, 1, 80, 0.1257) - kirkwood.born(1, 10, 1, 80, 0.1257)) < 1e-13
    # kappa = 0, off-centre charge: image-series closed form with the Kirkwood bracket
    b, a, ei, eo = 3.0, 10.0, 1.0, 80.0
    s = sum((n + 1) * (ei - eo) / ((n + 1) * eo + n * ei) * (b / a) ** (2 * n) / a for n in range(200))
    assert abs(kirkwood.energy([[0, 0, b]], [1.0], a, ei, eo, 0.0) - 0.5 * 332.0716 * s / ei) < 1e-12
