"""GPU: the three large workloads at their full size.

- Precision form, 317,112 panels (R = 40 A level-7 star-shaped globule with the
  faces under 0.05 A^2 merged, 8,000 charges, depth 5): the plain FP64 solve
  against the reference's own CPU solve of the same mesh (tests/golden/
  cfg3_precision: iterations +-1, 17-digit panel-tree energy <= 1e-9), and the
  FP32-refined solve: FP64 residual test met, energy <= 1e-6 from the reference.
- Differentiation form, 81,920 panels with 5,000 charges: the plain solve against
  the reference (tests/golden/cfg4_grad), dE/dq against central differences
  (h = 1e-4 e) on 20 seeded charges at GMRES tol 1e-10 (E is quadratic in q, so
  the quotient is exact to roughly tol |q| / h), dE/dkappa against a central
  difference, and the adjoint gradient's cost against one forward solve (<= 1.5x).
- Capsid-scale shell, 2,621,440 panels with 1,000,000 charges (order 8, MAC 0.7,
  depth 6), where no reference run fits: size-independent properties -- the
  FP32-refined energy within 1e-6 of the FP64 one, the transpose identity of the
  whole operator, linearity of the matvec, the charge-tree right-hand side
  against the direct sum at the walk's truncation level, the device-built tree
  against the host's, and the exact gradient's transposed walks against forward
  walks of unit charges.
"""
import os
import subprocess
import time

import numpy as np
import pytest

from conftest import CASES, ROOT, fixture, rel

pytestmark = pytest.mark.gpu
FL = CASES["flags"]


def _solve(fa, torch, op, mesh, ct, n, precision, tol=1e-6, maxit=100, x0=None):
    b = torch.empty(2 * n, dtype=torch.float64, device="cuda")
    op.rhs_dev(mesh.pos, mesh.chr, b)
    x = (torch.zeros_like(b) if precision == "fp32" else b.clone()) if x0 is None else x0
    if precision == "fp32":
        g = op.gmres_refined_dev(b, x, restart=FL["a"], tol=tol, maxit=maxit)
    else:
        g = op.gmres_dev(b, x, restart=FL["a"], tol=tol, maxit=maxit)
    return x, g, op.energy_dev(ct, mesh.pos, mesh.chr, x)


def test_precision_form_317k(fa, meshdir):
    import torch
    name = "cfg3_precision"
    cli = fixture(name)["cli"]
    print(f"\n[gpu] {name}: building the operator", flush=True)
    mesh = fa.Mesh(str(meshdir / name))
    tree = fa.Tree(mesh, CASES["cases"][name]["t"], order=FL["p"], height=FL["H"], sep=FL["S"])
    assert tree.n == cli["kept_panels"] and tree.nleaf == cli["leaf_cubes"]
    assert len(tree.a["m2lSrc"]) == cli["m2l_pairs"] and len(tree.a["nearSrc"]) == cli["near_pairs"]
    op = fa.Operator(tree, FL["k"], FL["eps1"], FL["eps2"], FL["q"])
    ct = tree.charge_tree()
    ref = cli["panel_tree_raw"]
    _, g64, e64 = _solve(fa, torch, op, mesh, ct, tree.n, "fp64")
    print(f"[{name}] fp64: its {g64['iters']} (reference {cli['iterations']}) energy {e64!r} "
          f"(reference {ref!r}, rel {abs(e64 - ref) / abs(ref):.2e})", flush=True)
    assert g64["info"] == 0 and abs(g64["iters"] - cli["iterations"]) <= 1
    assert abs(e64 - ref) <= 1e-9 * abs(ref)
    op.enable_fp32()
    _, g32, e32 = _solve(fa, torch, op, mesh, ct, tree.n, "fp32")
    print(f"[{name}] fp32-refined: its {g32['iters']} outer {g32['outer']} fp64 products {g32['mv64']} "
          f"resid {g32['resid']:.2e} energy rel {abs(e32 - ref) / abs(ref):.2e}", flush=True)
    assert g32["info"] == 0 and g32["resid"] <= 1e-6
    assert abs(e32 - ref) <= 1e-6 * abs(ref)
    # the precision form's energy walk (far polynomials in FP32) against the FP64
    # walk on the same density, and the reference's energy
    x32, _, _ = _solve(fa, torch, op, mesh, ct, tree.n, "fp32")
    op.set_charges(ct, mesh.pos, mesh.chr)
    w64, w32 = op.energy_res_dev(x32), op.energy_res_dev(x32, precision="fp32")
    print(f"[{name}] energy walk with FP32 far polynomials: rel {abs(w32 - w64) / abs(w64):.2e} from the FP64 walk, "
          f"{abs(w32 - ref) / abs(ref):.2e} from the reference", flush=True)
    assert abs(w32 - w64) <= 3e-7 * abs(w64) and abs(w32 - ref) <= 1e-6 * abs(ref)
    # the refined solve with the walk beside its last FP64 residual: the same x and energy
    b = torch.empty(2 * tree.n, dtype=torch.float64, device="cuda")
    op.rhs_dev(mesh.pos, mesh.chr, b)
    xf = torch.zeros_like(b)
    gf = op.gmres_refined_energy_dev(b, xf, restart=FL["a"], tol=1e-6, maxit=100)
    assert gf["info"] == 0 and gf["outer"] == g32["outer"] and gf["iters"] == g32["iters"]
    assert torch.equal(xf, x32) and gf["energy_raw"] == w32


_cfg4 = {}


def cfg4_problem(fa, meshdir, tol):
    if tol not in _cfg4:
        print(f"\n[gpu] cfg4_grad problem tol={tol}", flush=True)
        mesh = fa.Mesh(str(meshdir / "cfg4_grad"))
        _cfg4[tol] = fa.Problem(mesh, CASES["cases"]["cfg4_grad"]["t"], order=FL["p"], sep=FL["S"], eps1=FL["eps1"],
                                eps2=FL["eps2"], qorder=FL["q"], height=FL["H"], tol=tol, restart=FL["a"],
                                maxit=400, precond=FL["P"])
    return _cfg4[tol]


def test_differentiation_form_solve(fa, meshdir):
    """The 5,000-charge globule's plain FP64 solve against the reference's."""
    import torch
    cli = fixture("cfg4_grad")["cli"]
    P = cfg4_problem(fa, meshdir, 1e-6)
    op = P.operator(FL["k"])
    _, g, e = _solve(fa, torch, op, P.mesh, P.tree.charge_tree(), P.tree.n, "fp64")
    ref = cli["panel_tree_raw"]
    print(f"\n[cfg4_grad] its {g['iters']} (reference {cli['iterations']}) energy rel {abs(e - ref) / abs(ref):.2e}",
          flush=True)
    assert g["info"] == 0 and abs(g["iters"] - cli["iterations"]) <= 1
    assert abs(e - ref) <= 1e-9 * abs(ref)


def test_differentiation_form_gradient_q(fa, meshdir):
    P = cfg4_problem(fa, meshdir, 1e-10)
    q0 = P.mesh.chr.astype(np.float64).copy()
    e0, dq, _ = P.gradient(q0, FL["k"])
    pick = np.random.default_rng(44).choice(len(q0), size=20, replace=False)
    h = 1e-4
    fd = np.empty(len(pick))
    for i, j in enumerate(pick):
        qp, qm = q0.copy(), q0.copy()
        qp[j] += h
        qm[j] -= h
        fd[i] = (P.energy(qp, FL["k"]) - P.energy(qm, FL["k"])) / (2 * h)
    err = rel(fd, dq[pick])
    worst = float(np.max(np.abs(fd - dq[pick]) / np.abs(dq[pick])))
    print(f"\n[cfg4_grad] E={e0:.10f} kcal/mol, 20 charges: |FD - adjoint| / |adjoint| = {err:.2e} "
          f"(worst charge {worst:.2e})", flush=True)
    assert err <= 1e-5


def test_differentiation_form_gradient_kappa(fa, meshdir):
    P = cfg4_problem(fa, meshdir, 1e-10)
    q0 = P.mesh.chr.astype(np.float64)
    k = FL["k"]
    _, _, dk = P.gradient(q0, k)
    h = 1e-4
    fd = (P.energy(q0, k + h) - P.energy(q0, k - h)) / (2 * h)
    print(f"\n[cfg4_grad] dE/dkappa adjoint {dk!r} central difference {fd!r} rel {abs(dk - fd) / abs(fd):.2e}",
          flush=True)
    assert abs(dk - fd) <= 1e-5 * abs(fd)


def test_differentiation_form_cost(fa, meshdir):
    """The adjoint gradient (dE/dq and dE/dkappa) against one forward solve (right-hand
    side, GMRES and energy), best of 3; the gradient reuses the forward solution and energy."""
    import torch
    P = cfg4_problem(fa, meshdir, 1e-6)
    q0 = P.mesh.chr.astype(np.float64)
    fwd, adj = [], []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        u, e, _ = P.solve(q0, FL["k"])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        P.gradient(q0, FL["k"], u=u, e=e)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        fwd.append(t1 - t0)
        adj.append(t2 - t1)
    ratio = min(adj) / min(fwd)
    print(f"\n[cfg4_grad] forward {min(fwd) * 1e3:.2f} ms, adjoint gradient {min(adj) * 1e3:.2f} ms ({ratio:.2f}x)",
          flush=True)
    assert ratio <= 1.5


CAPSID = ["shell", "--level", "8", "--radius", "250", "--inner", "210", "--amp", "0.03", "--lmax", "12",
          "--charges", "wall", "--nq", "1000000", "--sigma", "0.3", "--depth", "1.5", "--seed", "5"]


def test_capsid_shell_properties(fa, tmp_path):
    import torch
    free, _ = torch.cuda.mem_get_info()
    if free < 64e9:
        pytest.skip(f"{free / 1e9:.0f} GB free; the shell needs ~35 GB")
    base = str(tmp_path / "capsid")
    print("\n[gpu] capsid shell: writing the mesh", flush=True)
    subprocess.check_call([os.path.join(ROOT, "cuda-fabipb_amd/bin/fbmesh"), CAPSID[0], base] + CAPSID[1:],
                          stdout=subprocess.DEVNULL)
    mesh = fa.Mesh(base)
    t0 = time.perf_counter()
    tree = fa.Tree(mesh, 6, order=8, sep=0.7)
    t1 = time.perf_counter()
    dtree = fa.Tree(mesh, 6, order=8, sep=0.7, device=True)
    t2 = time.perf_counter()
    print(f"[capsid] tree: host {t1 - t0:.3f} s, device panel pass {t2 - t1:.3f} s", flush=True)
    for k in fa.TREE_KEYS:
        assert np.array_equal(dtree.a[k], tree.a[k]), k
    del dtree
    assert tree.n == 2621440 and len(mesh.chr) == 1000000
    op = fa.Operator(tree, FL["k"], FL["eps1"], FL["eps2"], FL["q"])
    op.set_precond(3)
    ct = tree.charge_tree()
    n2 = 2 * tree.n
    print(f"[capsid] operator ready: {op.stats()['near_interactions']} near interactions", flush=True)

    # the charge-tree right-hand side against the direct sum: its error is the
    # walk's truncation (order ct["order"], 3 by the capsid policy), which falls
    # with the acceptance ratio theta
    bt = torch.empty(n2, dtype=torch.float64, device="cuda")
    bd = torch.empty_like(bt)
    op.rhs_dev(mesh.pos, mesh.chr, bd)
    e_rhs = {}
    for theta in (0.5, 0.8):
        op.rhs_tree_dev(ct, mesh.pos, mesh.chr, bt, theta=theta)
        e_rhs[theta] = float((bt - bd).norm() / bd.norm())
    print(f"[capsid] tree right-hand side (order {int(ct['order'][-1])}) vs direct: rel L2 {e_rhs[0.8]:.2e} at "
          f"theta 0.8, {e_rhs[0.5]:.2e} at 0.5", flush=True)
    assert e_rhs[0.8] <= 2e-3 and e_rhs[0.5] <= 0.3 * e_rhs[0.8]

    # FP64 and FP32-refined solves (tree right-hand side, as the reference's policy picks)
    res = {}
    for prec in ("fp64", "fp32"):
        if prec == "fp32":
            op.enable_fp32()
        x = torch.zeros_like(bt) if prec == "fp32" else bt.clone()
        g = (op.gmres_refined_dev if prec == "fp32" else op.gmres_dev)(bt, x, restart=FL["a"], tol=1e-6, maxit=200)
        assert g["info"] == 0 and g["resid"] <= 1e-6
        res[prec] = (x, op.energy_dev(ct, mesh.pos, mesh.chr, x))
        print(f"[capsid] {prec}: its {g['iters']} energy {res[prec][1]!r}", flush=True)
    e64, e32 = res["fp64"][1], res["fp32"][1]
    assert abs(e32 - e64) <= 1e-6 * abs(e64)
    op.set_charges(ct, mesh.pos, mesh.chr)   # the precision form's walk (order 8, FP32 far polynomials)
    w64, w32 = op.energy_res_dev(res["fp32"][0]), op.energy_res_dev(res["fp32"][0], precision="fp32")
    print(f"[capsid] energy walk with FP32 far polynomials: rel {abs(w32 - w64) / abs(w64):.2e}", flush=True)
    assert abs(w32 - w64) <= 3e-7 * abs(w64)

    # the transpose identity and linearity of the whole operator
    gen = torch.Generator(device="cpu").manual_seed(11)
    x = (torch.rand(n2, generator=gen, dtype=torch.float64) * 2 - 1).cuda()
    y = (torch.rand(n2, generator=gen, dtype=torch.float64) * 2 - 1).cuda()
    Ay, ATx, Ax, Az = (torch.empty_like(x) for _ in range(4))
    op.matvec_dev(y, Ay)
    op.matvec_t_dev(x, ATx)
    lhs, rhs = float(torch.dot(ATx, y)), float(torch.dot(x, Ay))
    scale = float(ATx.norm() * y.norm())
    print(f"[capsid] <A^T x, y> - <x, A y> = {abs(lhs - rhs) / scale:.2e} of the scale", flush=True)
    assert abs(lhs - rhs) <= 1e-11 * scale
    op.matvec_dev(x, Ax)
    op.matvec_dev(0.5 * x - 2.0 * y, Az)
    assert float((Az - (0.5 * Ax - 2.0 * Ay)).norm() / Az.norm()) <= 1e-13

    # the exact tree gradient's transposed walks at full size, against forward walks of unit charges
    q = np.ascontiguousarray(mesh.chr, dtype=np.float64)
    u = res["fp64"][0]
    for kind in (0, 1):
        dq = op.tree_adjoint_dev(ct, mesh.pos, q, u, kind=kind)
        worst = 0.0
        for j in np.random.default_rng(8 + kind).choice(len(q), size=2, replace=False):
            e = np.zeros_like(q)
            e[j] = 1.0
            if kind == 0:
                ref = op.energy_dev(ct, mesh.pos, e, u)
            else:
                op.rhs_tree_dev(ct, mesh.pos, e, bd)
                ref = float(torch.dot(u, bd))
            worst = max(worst, abs(dq[j] - ref))
        worst /= np.max(np.abs(dq))
        print(f"[capsid] transposed walk (kind {kind}) vs unit charges: {worst:.2e} of max |dq|", flush=True)
        assert worst <= 1e-11
