"""GPU parity: the HIP path (through the C ABI of libfabipb_hip.so) against the
reference fixtures recorded from the reference CPU path and against the oracle.

Bars (plain FP64): per-matvec relative L2 difference <= 1e-13; RHS <= 1e-13;
nearfield coefficients <= 1e-12 (see that test); solvation energy relative error <= 1e-9 and
GMRES iterations within +-1 of the reference solve; Kirkwood analytic energies
at the discretisation level.
"""
import os
import sys

import numpy as np
import pytest

from conftest import CASES, case_flags, cli_cases, fixture, harness_cases, rel, splitmix_vec
import golden_io
import kirkwood

pytestmark = pytest.mark.gpu
FL = CASES["flags"]

_ops = {}


def case_op(fa, meshdir, name):
    if name not in _ops:
        print(f"\n[gpu] building operator for {name}", flush=True)
        mesh = fa.Mesh(str(meshdir / name))
        tree = fa.Tree(mesh, CASES["cases"][name]["t"], order=FL["p"], height=FL["H"], sep=FL["S"])
        op = fa.Operator(tree, FL["k"], FL["eps1"], FL["eps2"], FL["q"])
        _ops[name] = (mesh, tree, op)
    return _ops[name]


@pytest.mark.parametrize("name", harness_cases())
def test_tree_matches_reference(fa, meshdir, name):
    mesh, tree, op = case_op(fa, meshdir, name)
    F = fixture(name)
    for k, fk in [("leafStart", "leaf_start"), ("leafCount", "leaf_count"), ("nearOff", "near_off"),
                  ("nearSrc", "near_src"), ("m2lSrc", "m2l_src"), ("m2lDst", "m2l_dst")]:
        assert np.array_equal(tree.a[k], F[fk]), k


@pytest.mark.parametrize("name", harness_cases())
def test_q2m_matrices(fa, meshdir, name):
    mesh, tree, op = case_op(fa, meshdir, name)
    F = fixture(name)
    q0, q1 = op.q2m_matrices()
    m = len(F["q2m0"])
    for got, ref in ((q0[:m], F["q2m0"]), (q1[:m], F["q2m1"])):
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("name", harness_cases())
def test_near_coefficients_every_case(fa, meshdir, name):
    """KER4 of panelIA0 for disjoint, vertex, edge (+-2) and self pairs.

    Bar 1e-12 of the record's largest entry: y3 = d2Gk - d2G0 cancels by a
    factor ~1/(kappa r)^2 ~ 1e3 on a self panel (r ~ 0.3 A), which amplifies the
    one-ulp difference between the device's and the host's exp().
    """
    mesh, tree, op = case_op(fa, meshdir, name)
    F = fixture(name)
    idx = F["ia0_idx"].reshape(-1, 3)
    val = F["ia0_val"].reshape(-1, 4)
    seen = set()
    for (x, y, c), v in zip(idx, val):
        k = op.near_coeff(x, y)
        assert np.max(np.abs(k - v)) <= 1e-12 * np.max(np.abs(v)), (x, y, c, k, v)
        seen.add(int(c))
    assert {0, 1, 2, -2, 3} <= seen


@pytest.mark.parametrize("name", harness_cases())
def test_pair_coeffs_every_case(fa, meshdir, name):
    """fbh_pair_coeffs (the preconditioner-assembly entry) on the recorded panelIA0
    pairs of every geometric case, in one batch, and with both index orders: the
    reference's values (bar as above) and bit-identical to the nearfield's own."""
    mesh, tree, op = case_op(fa, meshdir, name)
    F = fixture(name)
    idx = F["ia0_idx"].reshape(-1, 3)
    val = F["ia0_val"].reshape(-1, 4)
    k = op.pair_coeffs(idx[:, 0], idx[:, 1])
    for (x, y, c), v, got in zip(idx, val, k):
        assert np.max(np.abs(got - v)) <= 1e-12 * np.max(np.abs(v)), (x, y, c)
        assert np.array_equal(got, op.near_coeff(x, y))
    kt = op.pair_coeffs(idx[:, 1], idx[:, 0])   # X and Y exchanged: y0, y3 symmetric for disjoint pairs
    dis = idx[:, 2] == 0
    assert np.allclose(kt[dis][:, [0, 3]], k[dis][:, [0, 3]], rtol=1e-12, atol=0)


@pytest.mark.parametrize("name", harness_cases())
def test_near_apply(fa, meshdir, name):
    mesh, tree, op = case_op(fa, meshdir, name)
    F = fixture(name)
    assert rel(op.near_apply(F["mv_x"]), F["mv_near"]) <= 1e-13


@pytest.mark.parametrize("name", harness_cases())
def test_fmm_apply(fa, meshdir, name):
    """Far field + nearfield = the reference's applyFMM (alpha=1, beta=0)."""
    mesh, tree, op = case_op(fa, meshdir, name)
    F = fixture(name)
    far = op.far_apply(F["mv_x"])
    assert rel(far + op.near_apply(F["mv_x"]), F["mv_fmm"]) <= 1e-13


def test_expansions_match_reference(fa, meshdir):
    """Moments after Q2M+M2M and locals after M2L+L2L, cube by cube."""
    mesh, tree, op = case_op(fa, meshdir, "sph3_born")
    F = fixture("sph3_born")
    op.far_apply(F["mv_x"])
    mom, lec = op.expansions()
    assert np.max(np.abs(mom - F["mv_mom"])) <= 1e-13 * np.max(np.abs(F["mv_mom"]))
    assert np.max(np.abs(lec - F["mv_lec"])) <= 1e-13 * np.max(np.abs(F["mv_lec"]))


@pytest.mark.parametrize("name", harness_cases())
def test_matvec(fa, meshdir, name):
    mesh, tree, op = case_op(fa, meshdir, name)
    F = fixture(name)
    assert rel(op.matvec(F["mv_x"]), F["mv_y"]) <= 1e-13


@pytest.mark.parametrize("name", ["sph3_ball", "glob4"])
def test_matvec_alpha_beta_against_oracle(fa, meshdir, name):
    """y = alpha (D-K) x + beta y at a fresh seed, oracle on the product's own tree."""
    import oracle_binding as O
    mesh, tree, op = case_op(fa, meshdir, name)
    lay = tree.oracle_layout(FL["k"], FL["eps1"], FL["eps2"], FL["q"])
    ora = O.Oracle(lay)
    x = splitmix_vec(777, 2 * tree.n)
    y0 = splitmix_vec(778, 2 * tree.n)
    got = op.matvec(x, alpha=-0.7, beta=1.3, y=y0)
    ref = ora.matvec(x, alpha=-0.7, beta=1.3, y=y0)
    assert rel(got, ref) <= 1e-13


@pytest.mark.parametrize("name", harness_cases())
def test_rhs(fa, meshdir, name):
    mesh, tree, op = case_op(fa, meshdir, name)
    F = fixture(name)
    assert rel(op.rhs(mesh.pos, mesh.chr), F["rhs"]) <= 1e-13


@pytest.mark.parametrize("name", ["sph3_ball", "glob4"])
def test_near_modes_agree(fa, meshdir, monkeypatch, name):
    """The nearfield launch can mix tiles that stream stored Galerkin coefficients with
    tiles that evaluate the one-point formula on the fly (q = 1) plus stored
    corrections for the self / shared-vertex pairs.  All-streamed, a 45 % mix and the
    default (every tile on the fly) give the same matvec to rounding, in FP64 and FP32."""
    import torch
    mesh, tree, op = case_op(fa, meshdir, name)
    assert op.stats()["near_stored"] == 0
    g = torch.Generator(device="cpu").manual_seed(12)
    x = (torch.rand(2 * tree.n, generator=g, dtype=torch.float64) * 2 - 1).cuda()
    ys, y32 = {}, {}
    for frac in ("0.45", "1", None):
        if frac is None:
            o = op
        else:
            monkeypatch.setenv("FABIPB_NEAR_STREAM", frac)
            o = fa.Operator(tree, FL["k"], FL["eps1"], FL["eps2"], FL["q"])
        o.enable_fp32()
        ys[frac], y32[frac] = torch.empty_like(x), torch.empty_like(x)
        o.matvec_dev(x, ys[frac])
        o.matvec32_dev(x, y32[frac])
    for frac in ("0.45", "1"):
        assert rel(ys[frac].cpu().numpy(), ys[None].cpu().numpy()) <= 1e-14
        assert rel(y32[frac].cpu().numpy(), ys[None].cpu().numpy()) <= 1e-5
    assert rel(y32[None].cpu().numpy(), ys[None].cpu().numpy()) <= 1e-5


def test_large_leaves_all_near(fa, meshdir, monkeypatch):
    """Depth 1: eight leaves of 560-735 panels and no far field, so every pair is near
    and a leaf's own columns run past the first 512-column chunk the on-the-fly
    nearfield stages at a time.  The default (every pair on the fly, the own leaf's
    columns as a range of their own) matches the oracle and the all-streamed path."""
    import torch
    import oracle_binding as O
    mesh = fa.Mesh(str(meshdir / "glob4"))
    tree = fa.Tree(mesh, 1, order=FL["p"], height=1, sep=FL["S"])
    assert len(tree.a["m2lSrc"]) == 0 and tree.a["leafCount"].max() > 512
    op = fa.Operator(tree, FL["k"], FL["eps1"], FL["eps2"], FL["q"])
    assert op.stats()["near_stored"] == 0
    ora = O.Oracle(tree.oracle_layout(FL["k"], FL["eps1"], FL["eps2"], FL["q"]))
    x = splitmix_vec(779, 2 * tree.n)
    y = op.matvec(x)
    assert rel(y, ora.matvec(x)) <= 1e-13
    monkeypatch.setenv("FABIPB_NEAR_STREAM", "1")
    os_ = fa.Operator(tree, FL["k"], FL["eps1"], FL["eps2"], FL["q"])
    assert os_.stats()["near_stored"] > 0
    xd = torch.from_numpy(x).cuda()
    ya, yb = torch.empty_like(xd), torch.empty_like(xd)
    op.matvec_dev(xd, ya)
    os_.matvec_dev(xd, yb)
    assert rel(ya.cpu().numpy(), yb.cpu().numpy()) <= 1e-14
    assert rel(ya.cpu().numpy(), y) <= 1e-14
    # the FP32 on-the-fly nearfield over the same long own-column ranges
    op.enable_fp32()
    y32 = torch.empty_like(xd)
    op.matvec32_dev(xd, y32)
    assert rel(y32.cpu().numpy(), y) <= 1e-5


@pytest.mark.parametrize("name", ["sph3_ball", "glob4"])
def test_m2l_paths_agree(fa, meshdir, monkeypatch, name):
    """The M2L on the fixed term grid (registers per run, host-placed LDS slots) and the
    M2L over per-lane item lists give the same matvec, FP32 matvec and transposed matvec
    to rounding (they sum the same terms in different orders)."""
    import torch
    mesh, tree, op = case_op(fa, meshdir, name)
    assert op.stats()["m2l_grid_steps"] > 0
    monkeypatch.setenv("FABIPB_M2L", "items")
    oi = fa.Operator(tree, FL["k"], FL["eps1"], FL["eps2"], FL["q"])
    assert oi.stats()["m2l_grid_steps"] == 0
    g = torch.Generator(device="cpu").manual_seed(13)
    x = (torch.rand(2 * tree.n, generator=g, dtype=torch.float64) * 2 - 1).cuda()
    out = {}
    for key, o in (("grid", op), ("items", oi)):
        o.enable_fp32()
        y, y32, yt = torch.empty_like(x), torch.empty_like(x), torch.empty_like(x)
        o.matvec_dev(x, y)
        o.matvec32_dev(x, y32)
        o.matvec_t_dev(x, yt)
        out[key] = [v.cpu().numpy() for v in (y, y32, yt)]
    d = [rel(a, b) for a, b in zip(out["grid"], out["items"])]
    print(f"\n[{name}] grid vs items: fp64 {d[0]:.2e} fp32 {d[1]:.2e} transposed {d[2]:.2e}", flush=True)
    assert d[0] <= 1e-14 and d[2] <= 1e-14 and d[1] <= 1e-6


@pytest.mark.parametrize("name", harness_cases())
def test_energy_walk(fa, meshdir, name):
    import torch
    mesh, tree, op = case_op(fa, meshdir, name)
    F = fixture(name)
    ct = tree.charge_tree()
    s = torch.tensor(F["en_sigma"], dtype=torch.float64, device="cuda")
    e = op.energy_dev(ct, mesh.pos, mesh.chr, s)
    assert abs(e - F["en_raw"][0]) <= 1e-12 * abs(F["en_raw"][0])


@pytest.mark.parametrize("name", ["sph3_ball", "glob4"])
def test_resident_charges(fa, meshdir, name):
    """The resident charge set (fbh_set_charges) gives what the per-call entries give
    from host arrays, bit for bit (same kernels on the same data), on repeated calls and
    after it is replaced by another charge set."""
    import torch
    mesh, tree, op = case_op(fa, meshdir, name)
    ct = tree.charge_tree()
    s = torch.from_numpy(splitmix_vec(4242, 2 * tree.n)).cuda()
    b1, b2 = torch.empty_like(s), torch.empty_like(s)
    for q in (mesh.chr, -0.5 * mesh.chr[::-1].copy()):
        op.set_charges(ct, mesh.pos, q)
        for _ in range(2):
            op.rhs_dev(mesh.pos, q, b1)
            op.rhs_res_dev(b2)
            assert torch.equal(b1, b2)
            op.rhs_tree_dev(ct, mesh.pos, q, b1)
            op.rhs_tree_res_dev(b2)
            assert torch.equal(b1, b2)
            e64 = op.energy_dev(ct, mesh.pos, q, s)
            assert e64 == op.energy_res_dev(s)
            # the precision form's walk (far polynomials in FP32): repeatable, near the FP64 walk
            e32 = op.energy_res_dev(s, precision="fp32")
            assert e32 == op.energy_res_dev(s, precision="fp32")
            assert abs(e32 - e64) <= 1e-5 * abs(e64)


@pytest.mark.parametrize("name", ["sph3_ball", "glob4"])
def test_refined_energy_overlap(fa, meshdir, name):
    """The refined solve with the energy walk beside its last FP64 residual
    (gmres_refined_energy_dev) returns the x and the energy of the two calls in turn."""
    import torch
    mesh, tree, op = case_op(fa, meshdir, name)
    ct = tree.charge_tree()
    op.enable_fp32()
    op.set_charges(ct, mesh.pos, mesh.chr)
    b = torch.empty(2 * tree.n, dtype=torch.float64, device="cuda")
    op.rhs32_res_dev(b)
    xa, xb = torch.zeros_like(b), torch.zeros_like(b)
    ga = op.gmres_refined_dev(b, xa, restart=10, tol=1e-6, maxit=100)
    ea = op.energy_res_dev(xa, precision="fp32")
    for _ in range(2):
        xb.zero_()
        gb = op.gmres_refined_energy_dev(b, xb, restart=10, tol=1e-6, maxit=100)
        assert gb["info"] == ga["info"] == 0 and gb["iters"] == ga["iters"] and gb["outer"] == ga["outer"]
        assert torch.equal(xa, xb) and gb["energy_raw"] == ea


@pytest.mark.parametrize("name", [c for c in cli_cases() if c != "cfg2_globule"])
def test_solve_matches_reference(fa, meshdir, name):
    """Full plain-FP64 solve: energy rel. error <= 1e-9, GMRES iterations within +-1, the
    same info (the cases include GMRES(3) restarts, an iteration limit, the -P=2
    block-LU preconditioner and the -r=2 dense direct operator)."""
    F = fixture(name)
    cli = F["cli"]
    fl = case_flags(name)
    r = fa.solve(str(meshdir / name), CASES["cases"][name]["t"], order=fl["p"], sep=fl["S"], kappa=fl["k"],
                 eps1=fl["eps1"], eps2=fl["eps2"], qorder=fl["q"], tol=fl["o"], restart=fl["a"],
                 maxit=fl["i"], height=fl["H"], precond=fl["P"], direct=fl.get("r", 0) > 0,
                 rhs="tree" if CASES["cases"][name].get("tree_rhs") else "direct")
    print(f"\n[{name}] its={r['iters']} (ref {cli['iterations']}) raw={r['energy_raw']!r} "
          f"(ref {cli['panel_tree_raw']!r})", flush=True)
    assert abs(r["iters"] - cli["iterations"]) <= 1 and r["info"] == cli["info"]
    assert abs(r["energy_raw"] - cli["panel_tree_raw"]) <= 1e-9 * abs(cli["panel_tree_raw"])
    if "solution" in F and r["iters"] == cli["iterations"]:
        assert rel(r["x"], F["solution"]) <= 1e-9


def test_solve_cfg2_globule(fa, meshdir):
    """The bench configuration (81,920 panels, 2,000 charges, t=4) against the reference solve."""
    name = "cfg2_globule"
    F = fixture(name)
    cli = F["cli"]
    r = fa.solve(str(meshdir / name), CASES["cases"][name]["t"], order=FL["p"], sep=FL["S"], kappa=FL["k"],
                 eps1=FL["eps1"], eps2=FL["eps2"], qorder=FL["q"], tol=FL["o"], restart=FL["a"],
                 maxit=FL["i"], height=FL["H"])
    print(f"\n[{name}] its={r['iters']} (ref {cli['iterations']}) raw={r['energy_raw']!r} "
          f"(ref {cli['panel_tree_raw']!r}) gmres={r['t_gmres']:.3f}s", flush=True)
    assert abs(r["iters"] - cli["iterations"]) <= 1
    assert abs(r["energy_raw"] - cli["panel_tree_raw"]) <= 1e-9 * abs(cli["panel_tree_raw"])


def test_kirkwood_sphere(fa, meshdir):
    """cfg1-style sphere with 50 charges: energy against the analytic Kirkwood series."""
    name = "cfg1_sphere"
    r = fa.solve(str(meshdir / name), CASES["cases"][name]["t"], order=FL["p"], sep=FL["S"], kappa=FL["k"],
                 eps1=FL["eps1"], eps2=FL["eps2"], qorder=FL["q"], tol=FL["o"], height=FL["H"])
    mesh = r["mesh"]
    ek = kirkwood.energy(mesh.pos, mesh.chr, 10.0, FL["eps1"], FL["eps2"], FL["k"])
    print(f"\n[kirkwood] fabipb={r['energy']:.6f} analytic={ek:.6f}", flush=True)
    assert abs(r["energy"] - ek) <= 2e-3 * abs(ek)


@pytest.mark.parametrize("name", ["sph3_ball", "glob4"])
@pytest.mark.parametrize("order", [6, -4])
def test_orders_against_oracle(fa, meshdir, name, order):
    """Fixed order 6 and the variable-order mode (finest level 4, one more per level up,
    src/gkSetup.c:42-81): matvec and panel-tree energy against the oracle on the product's
    own tree.  Order -4 gives M2L orders that differ by level and charge-tree orders 4/5,
    so the energy takes the wave-per-panel walk; order 6 takes the wave walk."""
    import torch
    import oracle_binding as O
    mesh = fa.Mesh(str(meshdir / name))
    tree = fa.Tree(mesh, CASES["cases"][name]["t"], order=order, height=FL["H"], sep=FL["S"])
    op = fa.Operator(tree, FL["k"], FL["eps1"], FL["eps2"], FL["q"])
    ora = O.Oracle(tree.oracle_layout(FL["k"], FL["eps1"], FL["eps2"], FL["q"]))
    x = splitmix_vec(4242, 2 * tree.n)
    assert rel(op.matvec(x), ora.matvec(x)) <= 1e-13
    ct = tree.charge_tree()
    if order < 0:
        assert len(set(ct["order"][tree.height:].tolist())) > 1
    sg = splitmix_vec(4243, 2 * tree.n)
    e = op.energy_dev(ct, mesh.pos, mesh.chr, torch.tensor(sg, dtype=torch.float64, device="cuda"))
    ref = ora.energy(ct, mesh.pos, mesh.chr, sg)
    print(f"\n[{name} p={order}] energy {e!r} oracle {ref!r}", flush=True)
    assert abs(e - ref) <= 1e-12 * abs(ref)


@pytest.mark.parametrize("order,steps", [(3, 4), (5, 8), (7, 32), (8, 64)])
def test_m2l_grid_shapes(fa, meshdir, order, steps):
    """Each compiled shape of the M2L term grid (steps x runs): matvec against the oracle
    on the product's own tree, and the transpose identity of the adjoint operator."""
    import torch
    import oracle_binding as O
    name = "sph3_ball"
    mesh = fa.Mesh(str(meshdir / name))
    tree = fa.Tree(mesh, CASES["cases"][name]["t"], order=order, height=FL["H"], sep=FL["S"])
    op = fa.Operator(tree, FL["k"], FL["eps1"], FL["eps2"], FL["q"])
    assert op.stats()["m2l_grid_steps"] == steps
    ora = O.Oracle(tree.oracle_layout(FL["k"], FL["eps1"], FL["eps2"], FL["q"]))
    x = splitmix_vec(4244, 2 * tree.n)
    d = rel(op.matvec(x), ora.matvec(x))
    g = torch.Generator(device="cpu").manual_seed(14)
    u = (torch.rand(2 * tree.n, generator=g, dtype=torch.float64) * 2 - 1).cuda()
    v = (torch.rand(2 * tree.n, generator=g, dtype=torch.float64) * 2 - 1).cuda()
    Av, ATu = torch.empty_like(u), torch.empty_like(u)
    op.matvec_dev(v, Av)
    op.matvec_t_dev(u, ATu)
    ident = abs(float(torch.dot(ATu, v)) - float(torch.dot(u, Av))) / float(ATu.norm() * v.norm())
    print(f"\n[{name} p={order}] grid {steps} steps: matvec vs oracle {d:.2e}, transpose identity {ident:.2e}",
          flush=True)
    assert d <= 1e-13 and ident <= 1e-11


@pytest.mark.parametrize("name", ["glob4", "cfg1_sphere"])
def test_fp32_operator(fa, meshdir, name):
    """The FP32 operator of the precision form against the FP64 one.  Tolerance 1e-5:
    FP32 coefficients and expansions (unit roundoff 6e-8) summed over ~1e3 near
    sources per row, with the cancellation of D x - K x."""
    mesh, tree, op = case_op(fa, meshdir, name)
    op.enable_fp32()
    for seed in (91, 92):
        x = splitmix_vec(seed, 2 * tree.n)
        y64 = op.matvec(x)
        y32 = op.matvec32(x)
        print(f"\n[{name}] fp32 matvec rel {rel(y32, y64):.3e}", flush=True)
        assert rel(y32, y64) <= 1e-5


@pytest.mark.parametrize("name", ["glob4", "cfg1_sphere", "cfg2_globule"])
def test_fp32_refined_solve(fa, meshdir, name):
    """Precision form: FP64-refined solve on the FP32 operator reaches the plain solve's
    residual test in FP64 and its energy within 1e-6 relative of the discrete solution
    (an FP64 solve at tol 1e-10) and of the reference's own recorded 17-digit panel-tree
    energy (its CPU solve at tol 1e-6); the plain solve's own distance is printed beside."""
    mesh, tree, op = case_op(fa, meshdir, name)
    kw = dict(kappa=FL["k"], eps1=FL["eps1"], eps2=FL["eps2"], qorder=FL["q"],
              restart=FL["a"], maxit=400, height=FL["H"], op=op, tree=tree, mesh=mesh)
    ref = fa.solve(None, tree.depth, tol=1e-10, **kw)["energy_raw"]
    r64 = fa.solve(None, tree.depth, tol=FL["o"], **kw)
    r32 = fa.solve(None, tree.depth, precision="fp32", tol=FL["o"], **kw)
    err = abs(r32["energy_raw"] - ref) / abs(ref)
    err64 = abs(r64["energy_raw"] - ref) / abs(ref)
    print(f"\n[{name}] fp64 its={r64['iters']} fp32 its={r32['iters']} outer={r32['outer']} "
          f"mv64={r32['mv64']} resid={r32['resid']:.3e} energy rel err fp32={err:.3e} fp64={err64:.3e}",
          flush=True)
    assert r32["info"] == 0 and r32["resid"] <= FL["o"]
    assert err <= 1e-6
    cli = fixture(name)["cli"]
    err_ref = abs(r32["energy_raw"] - cli["panel_tree_raw"]) / abs(cli["panel_tree_raw"])
    print(f"[{name}] fp32-refined energy against the reference's: {err_ref:.3e}", flush=True)
    assert err_ref <= 1e-6


@pytest.mark.parametrize("name", ["sph3_ball_tr", "glob4_tr"])
def test_tree_rhs_matches_reference(fa, meshdir, name):
    """The charge-tree right-hand side on the device against the reference's RHS sample
    (FABIPB_FORCE_TREE_RHS=1), row by row."""
    import torch
    mesh, tree, op = case_op(fa, meshdir, CASES["cases"][name]["same_as"])
    ref = np.fromfile(os.path.join(os.path.dirname(__file__), "golden", name, "rhs_tree.bin"))
    b = torch.empty(2 * tree.n, dtype=torch.float64, device="cuda")
    op.rhs_tree_dev(tree.charge_tree(), mesh.pos, mesh.chr, b)
    got = b.cpu().numpy()
    print(f"\n[{name}] tree rhs max rel {np.max(np.abs(got - ref)) / np.max(np.abs(ref)):.2e}", flush=True)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("name", ["sph3_ball", "glob4"])
def test_block_precond_against_oracle(fa, meshdir, name):
    """The device block preconditioner (explicit inverses of the leaf-parent blocks)
    against the oracle's LU solves on the product's own tree, on a fresh vector."""
    import torch
    import oracle_binding as O
    mesh, tree, op = case_op(fa, meshdir, name)
    ora = O.Oracle(tree.oracle_layout(FL["k"], FL["eps1"], FL["eps2"], FL["q"]))
    ora.set_precond(2)
    op.set_precond(2)
    try:
        v = splitmix_vec(515, 2 * tree.n)
        x = torch.tensor(v, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        op.psolve_dev(x, y)
        op.psolve_dev(x, x)   # in place gives the same
        got, inplace = y.cpu().numpy(), x.cpu().numpy()
    finally:
        op.set_precond(3)
    ref = ora.psolve_any(v)
    assert rel(got, ref) <= 1e-11
    assert np.array_equal(got, inplace)


def test_matvec_linearity_full_size(fa, meshdir):
    """Size-independent property at the bench size: A(a x + b z) = a A x + b A z."""
    mesh, tree, op = case_op(fa, meshdir, "cfg2_globule")
    n2 = 2 * tree.n
    rng = np.random.default_rng(5)
    x, z = rng.standard_normal(n2), rng.standard_normal(n2)
    lhs = op.matvec(0.3 * x - 1.7 * z)
    rhs = 0.3 * op.matvec(x) - 1.7 * op.matvec(z)
    assert rel(lhs, rhs) <= 1e-13
    # repeatability: the apply has no atomics, results are bit-identical run to run
    assert np.array_equal(op.matvec(x), op.matvec(x))


# the reference host re-executes itself unless all five are set (src/fabipb_options.c:134-163)
THREAD_ENV = dict(OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1", VECLIB_MAXIMUM_THREADS="1",
                  BLIS_NUM_THREADS="1")


@pytest.mark.parametrize("name", ["sph3_ball", "glob4"])
def test_reference_host_with_hip_backend(meshdir, name):
    """Drop-in: the reference's own host code (oracle/_ref/fabipb_ref_hip, built with
    cuda-fabipb_amd/integration/gpu_backend_hip.c in place of its CUDA backend) runs
    -g=1 on the GPU.  Its built-in -c=1 check compares one CPU applyFMM with the
    GPU stages; its solve must reproduce the reference's CPU solve."""
    import re
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle/_ref/fabipb_ref_hip")
    if not os.path.exists(exe):
        pytest.skip("drop-in binary not built (needs the reference sources)")
    F = fixture(name)
    cli = F["cli"]
    env = dict(os.environ, **THREAD_ENV, FABIPB_REUSE_MESH="1", FABIPB_ENERGY_MODE="compare",
               FABIPB_NEARFIELD_APPLY_THREADS="8", FABIPB_SETUP_THREADS="8")
    r = subprocess.run([exe, "-g=1", "-c=1", "-B=1"] + cli["flags"] + [str(meshdir / name)], env=env,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=240)
    print(r.stdout[-1500:], flush=True)
    assert r.returncode == 0
    assert "unavailable" not in r.stdout and "fallback" not in r.stdout.lower()
    m = re.search(r"applyFMM debug compare: max_abs=(\S+) rel_l2=(\S+)", r.stdout)
    assert m and float(m[2]) <= 1e-13
    m = re.search(r"GMRES status: info=(\d+) iterations=(\d+)", r.stdout)
    assert abs(int(m[2]) - cli["iterations"]) <= 1
    raw = float(re.search(r"panel-tree raw=(\S+)", r.stdout)[1])
    assert abs(raw - cli["panel_tree_raw"]) <= 1e-9 * abs(cli["panel_tree_raw"])


@pytest.mark.parametrize("name", ["sph3_ball_r2", "glob4_r2"])
def test_reference_host_direct_on_gpu(meshdir, name):
    """Drop-in, dense direct operator: the reference host with -r=1 calls gpuDirectApply
    (the binding's fbh_direct_apply) every matvec; its solve must reproduce the
    reference's own -r=2 CPU solve (same operator, same order of summation)."""
    import re
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle/_ref/fabipb_ref_hip")
    if not os.path.exists(exe):
        pytest.skip("drop-in binary not built (needs the reference sources)")
    cli = fixture(name)["cli"]
    flags = [f if f != "-r=2" else "-r=1" for f in cli["flags"]]
    env = dict(os.environ, **THREAD_ENV, FABIPB_REUSE_MESH="1", FABIPB_ENERGY_MODE="compare")
    r = subprocess.run([exe, "-g=1", "-B=1"] + flags + [str(meshdir / name)], env=env,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=240)
    print(r.stdout[-1200:], flush=True)
    assert r.returncode == 0 and "Direct GPU matvec unavailable" not in r.stdout
    m = re.search(r"GMRES status: info=(\d+) iterations=(\d+)", r.stdout)
    assert int(m[1]) == cli["info"] and abs(int(m[2]) - cli["iterations"]) <= 1
    raw = float(re.search(r"panel-tree raw=(\S+)", r.stdout)[1])
    assert abs(raw - cli["panel_tree_raw"]) <= 1e-9 * abs(cli["panel_tree_raw"])


def test_reference_host_tree_rhs_on_gpu(meshdir):
    """Drop-in, charge-tree right-hand side: the reference host with
    FABIPB_FORCE_TREE_RHS=1 and -g=1 calls gpuChargeTreeRHS (the binding's
    fbh_rhs_tree_dev); its solve must reproduce the reference's CPU tree-RHS solve."""
    import re
    import subprocess
    name = "glob4_tr"
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle/_ref/fabipb_ref_hip")
    if not os.path.exists(exe):
        pytest.skip("drop-in binary not built (needs the reference sources)")
    cli = fixture(name)["cli"]
    env = dict(os.environ, **THREAD_ENV, FABIPB_REUSE_MESH="1", FABIPB_ENERGY_MODE="compare",
               FABIPB_FORCE_TREE_RHS="1", FABIPB_HIP_TRACE="1")
    r = subprocess.run([exe, "-g=1", "-B=1"] + cli["flags"] + [str(meshdir / name)], env=env,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=240)
    print(r.stdout[-1200:], flush=True)
    assert r.returncode == 0 and "mode=tree" in r.stdout
    assert "gpu_backend_hip: gpuChargeTreeRHS done" in r.stdout
    m = re.search(r"GMRES status: info=(\d+) iterations=(\d+)", r.stdout)
    assert int(m[1]) == cli["info"] and abs(int(m[2]) - cli["iterations"]) <= 1
    raw = float(re.search(r"panel-tree raw=(\S+)", r.stdout)[1])
    assert abs(raw - cli["panel_tree_raw"]) <= 1e-9 * abs(cli["panel_tree_raw"])


@pytest.mark.parametrize("name,batched", [("sph3_ball_p2", False), ("glob4_p2", False), ("glob4_p2", True)])
def test_reference_host_precond_on_gpu(meshdir, name, batched):
    """Drop-in, block preconditioner: the reference host with -P=2 -g=1 assembles its
    leaf-parent blocks from gpuBuildPrecondDisjointBlock (one call per block) or, with
    FABIPB_PRECOND_BATCH_MIN_BLOCKS=0, from gpuBuildPrecondPairsBatched (the binding's
    fbh_panel_pair_coeffs / fbh_pair_coeffs); its solve must reproduce the reference's
    own -P=2 CPU solve."""
    import re
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle/_ref/fabipb_ref_hip")
    if not os.path.exists(exe):
        pytest.skip("drop-in binary not built (needs the reference sources)")
    cli = fixture(name)["cli"]
    env = dict(os.environ, **THREAD_ENV, FABIPB_REUSE_MESH="1", FABIPB_ENERGY_MODE="compare", FABIPB_HIP_TRACE="1")
    if batched:
        env["FABIPB_PRECOND_BATCH_MIN_BLOCKS"] = "0"
    r = subprocess.run([exe, "-g=1", "-B=1"] + cli["flags"] + [str(meshdir / name)], env=env,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=240)
    print(r.stdout[-1500:], flush=True)
    assert r.returncode == 0
    hook = "gpuBuildPrecondPairsBatched" if batched else "gpuBuildPrecondDisjointBlock"
    assert f"gpu_backend_hip: {hook} done" in r.stdout
    m = re.search(r"GMRES status: info=(\d+) iterations=(\d+)", r.stdout)
    assert int(m[1]) == cli["info"] and abs(int(m[2]) - cli["iterations"]) <= 1
    raw = float(re.search(r"panel-tree raw=(\S+)", r.stdout)[1])
    assert abs(raw - cli["panel_tree_raw"]) <= 1e-9 * abs(cli["panel_tree_raw"])


@pytest.mark.parametrize("name", ["glob4", "glob4_p2", "glob4_a3", "glob4_r2", "glob4_tr"])
def test_fabipb_cli_solve(meshdir, name):
    """bin/fabipb (the C host driver) on the GPU with the reference's recorded flags
    (diagonal, -P=2 block LU, GMRES(3), -r=2 dense direct, tree RHS): iterations and
    17-digit energy."""
    import re
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cli = fixture(name)["cli"]
    env = dict(os.environ, FABIPB_ENERGY_MODE="compare")
    if CASES["cases"][name].get("tree_rhs"):
        env["FABIPB_FORCE_TREE_RHS"] = "1"
    r = subprocess.run([os.path.join(root, "cuda-fabipb_amd/bin/fabipb"), "-g=1", "-B=1"] + cli["flags"] +
                       [str(meshdir / name)], env=env,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=240)
    print(r.stdout[-1200:], flush=True)
    assert r.returncode == 0
    m = re.search(r"GMRES status: info=(\d+) iterations=(\d+)", r.stdout)
    assert int(m[1]) == cli["info"] and abs(int(m[2]) - cli["iterations"]) <= 1
    raw = float(re.search(r"panel-tree raw=(\S+)", r.stdout)[1])
    assert abs(raw - cli["panel_tree_raw"]) <= 1e-9 * abs(cli["panel_tree_raw"])
    assert f"leaf-cubes={cli['leaf_cubes']} near-pairs={cli['near_pairs']}" in r.stdout
