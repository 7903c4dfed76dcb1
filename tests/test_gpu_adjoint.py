"""GPU: the differentiation form.  The adjoint operator, the energies and the
adjoint gradients dE/dq, dE/dkappa, checked against what mathematics fixes:

- <A^T x, y> = <x, A y> for the whole operator (transposed nearfield and the
  transposed FMM chain) at 1e-11;
- the direct energy against the oracle's panel-tree energy at theta = 0 (every
  charge taken directly) at 1e-12;
- the transposed walks: at fixed densities the panel-tree energy and the
  charge-tree right-hand side are linear in the charges, so d/dq_j equals the
  energy (or w . right-hand side) of the unit charge vector e_j, evaluated by the
  forward walks (1e-12);
- dE/dq of the panel-tree energy (the default) and of the direct energy against
  central differences on 20 seeded charges: E is quadratic in q, so with GMRES at
  tol 1e-11 the difference quotient is exact to ~1e-8; with the direct and with the
  charge-tree right-hand side;
- dE/dkappa against a central difference in kappa (operators rebuilt at kappa +- h);
- the q = 1 kappa-derivative nearfield (one-point formula plus stored special-pair
  corrections) against the full Galerkin rules over every pair;
- the torch.autograd wrapper returns the adjoint gradient.
"""
import numpy as np
import pytest

from conftest import CASES, rel

pytestmark = pytest.mark.gpu
FL = CASES["flags"]
_probs = {}


def problem(fa, meshdir, name, tol=1e-11, precond=3, energy="tree", rhs="direct", theta=0.8):
    key = (name, tol, precond, energy, rhs, theta)
    if key not in _probs:
        print(f"\n[gpu] problem {name} tol={tol} P={precond} energy={energy} rhs={rhs} theta={theta}", flush=True)
        mesh = fa.Mesh(str(meshdir / name))
        _probs[key] = fa.Problem(mesh, CASES["cases"][name]["t"], order=FL["p"], sep=FL["S"], eps1=FL["eps1"],
                                 eps2=FL["eps2"], qorder=FL["q"], height=FL["H"], tol=tol, restart=FL["a"],
                                 maxit=400, precond=precond, energy=energy, rhs=rhs, theta=theta)
    return _probs[key]


@pytest.mark.parametrize("name", ["sph3_ball", "glob4", "cfg1_sphere"])
def test_adjoint_identity(fa, meshdir, name):
    import torch
    P = problem(fa, meshdir, name)
    op = P.operator(FL["k"])
    n2 = 2 * P.tree.n
    g = torch.Generator(device="cpu").manual_seed(5)
    x = (torch.rand(n2, generator=g, dtype=torch.float64) * 2 - 1).cuda()
    y = (torch.rand(n2, generator=g, dtype=torch.float64) * 2 - 1).cuda()
    Ay, ATx = torch.empty_like(x), torch.empty_like(x)
    op.matvec_dev(y, Ay)
    op.matvec_t_dev(x, ATx)
    lhs, rhs = float(torch.dot(ATx, y)), float(torch.dot(x, Ay))
    scale = float(ATx.norm() * y.norm())
    print(f"\n[{name}] <A^T x, y> = {lhs!r}  <x, A y> = {rhs!r}  diff/scale {abs(lhs - rhs) / scale:.2e}",
          flush=True)
    assert abs(lhs - rhs) <= 1e-11 * scale
    # and A^T is not A again (the diagonal dominates; the rest is far from symmetric)
    Ax = torch.empty_like(x)
    op.matvec_dev(x, Ax)
    assert rel(ATx.cpu().numpy(), Ax.cpu().numpy()) > 1e-3


@pytest.mark.parametrize("precond", [3, 2])
def test_adjoint_solve(fa, meshdir, precond):
    import torch
    P = problem(fa, meshdir, "glob4", tol=1e-10, precond=precond)
    op = P.operator(FL["k"])
    n2 = 2 * P.tree.n
    b = (torch.rand(n2, generator=torch.Generator().manual_seed(9), dtype=torch.float64) * 2 - 1).cuda()
    lam = torch.zeros_like(b)
    info = op.gmres_t_dev(b, lam, restart=30, tol=1e-10, maxit=400)
    r = torch.empty_like(b)
    op.matvec_t_dev(lam, r)
    res = float((r - b).norm() / b.norm())
    print(f"\n[glob4 P={precond}] adjoint GMRES {info}, true residual {res:.2e}", flush=True)
    assert info["info"] == 0 and res <= 1e-8


@pytest.mark.parametrize("name", ["sph3_ball", "glob4"])
def test_energy_direct_against_oracle(fa, meshdir, name):
    import torch
    import oracle_binding as O
    P = problem(fa, meshdir, name)
    op = P.operator(FL["k"])
    tree, mesh = P.tree, P.mesh
    ora = O.Oracle(tree.oracle_layout(FL["k"], FL["eps1"], FL["eps2"], FL["q"]))
    s = np.random.default_rng(3).uniform(-1, 1, 2 * tree.n)
    e = op.energy_direct_dev(mesh.pos, mesh.chr, torch.tensor(s, device="cuda"))
    ref = ora.energy(tree.charge_tree(), mesh.pos, mesh.chr, s, theta=0.0)
    tre = ora.energy(tree.charge_tree(), mesh.pos, mesh.chr, s, theta=0.8)
    print(f"\n[{name}] direct {e!r} oracle(theta=0) {ref!r} panel-tree(0.8) {tre!r}", flush=True)
    assert abs(e - ref) <= 1e-12 * abs(ref)


@pytest.mark.parametrize("name", ["sph3_ball", "glob4", "cfg1_sphere"])
@pytest.mark.parametrize("kind", [0, 1])
def test_tree_adjoint_unit_charges(fa, meshdir, name, kind):
    """d/dq of the panel-tree energy at fixed sgm (kind 0) and of w . (charge-tree
    right-hand side) (kind 1) by the transposed walk, against the forward walks of
    unit charges."""
    import torch
    P = problem(fa, meshdir, name)
    op = P.operator(FL["k"])
    mesh, ct = P.mesh, P.ct
    n2 = 2 * P.tree.n
    w = torch.tensor(np.random.default_rng(31 + kind).uniform(-1, 1, n2), device="cuda")
    q0 = mesh.chr.astype(np.float64)
    dq = op.tree_adjoint_dev(ct, mesh.pos, q0, w, kind=kind)
    pick = np.random.default_rng(32).choice(len(q0), size=min(10, len(q0)), replace=False)
    b = torch.empty_like(w)
    worst = 0.0
    for j in pick:
        e = np.zeros_like(q0)
        e[j] = 1.0
        if kind == 0:
            ref = op.energy_dev(ct, mesh.pos, e, w)
        else:
            op.rhs_tree_dev(ct, mesh.pos, e, b)
            ref = float(torch.dot(w, b))
        worst = max(worst, abs(dq[j] - ref))
    worst /= np.max(np.abs(dq))
    print(f"\n[{name} kind {kind}] transposed walk vs unit charges: worst {worst:.2e} of max |dq|", flush=True)
    assert worst <= 1e-12


@pytest.mark.parametrize("name,energy,rhs", [("sph3_ball", "tree", "direct"), ("glob4", "tree", "direct"),
                                             ("glob4", "tree", "tree"), ("glob4", "direct", "direct")])
def test_gradient_q_against_differences(fa, meshdir, name, energy, rhs):
    P = problem(fa, meshdir, name, energy=energy, rhs=rhs)
    q0 = P.mesh.chr.astype(np.float64).copy()
    e0, dq, dk = P.gradient(q0, FL["k"])
    pick = np.random.default_rng(20).choice(len(q0), size=min(20, len(q0)), replace=False)
    h = 1e-4
    worst = 0.0
    for j in pick:
        qp, qm = q0.copy(), q0.copy()
        qp[j] += h
        qm[j] -= h
        fd = (P.energy(qp, FL["k"]) - P.energy(qm, FL["k"])) / (2 * h)
        worst = max(worst, abs(fd - dq[j]) / np.max(np.abs(dq)))
    print(f"\n[{name} {energy} energy, {rhs} rhs] E={e0:.10f} |dE/dq|max={np.max(np.abs(dq)):.4e} "
          f"worst FD diff {worst:.2e}", flush=True)
    assert worst <= 1e-6


@pytest.mark.parametrize("name,energy,rhs", [("sph3_ball", "tree", "direct"), ("glob4", "tree", "direct"),
                                             ("glob4", "tree", "tree"), ("glob4", "direct", "direct")])
def test_gradient_kappa_against_difference(fa, meshdir, name, energy, rhs):
    P = problem(fa, meshdir, name, energy=energy, rhs=rhs)
    q0 = P.mesh.chr.astype(np.float64)
    k = FL["k"]
    _, _, dk = P.gradient(q0, k)
    h = 1e-4
    fd = (P.energy(q0, k + h) - P.energy(q0, k - h)) / (2 * h)
    print(f"\n[{name}] dE/dkappa adjoint {dk!r} central difference {fd!r} rel {abs(dk - fd) / abs(fd):.2e}",
          flush=True)
    assert abs(dk - fd) <= 1e-5 * abs(fd)


def test_kappa_nearfield_formula_matches_rules(fa, meshdir, monkeypatch):
    """q = 1: the one-point formula with stored corrections equals the full rules."""
    P = problem(fa, meshdir, "sph3_ball")
    q0 = P.mesh.chr.astype(np.float64)
    u, _, _ = P.solve(q0, FL["k"])
    _, _, dk_fast = P.gradient(q0, FL["k"], u=u)
    monkeypatch.setenv("FABIPB_DK_RULES", "1")
    _, _, dk_rules = P.gradient(q0, FL["k"], u=u)
    print(f"\n[sph3_ball] dE/dkappa formula {dk_fast!r} rules {dk_rules!r}", flush=True)
    assert abs(dk_fast - dk_rules) <= 1e-11 * abs(dk_rules)


def test_autograd_wrapper(fa, meshdir):
    import torch
    P = problem(fa, meshdir, "sph3_ball")
    q = torch.tensor(P.mesh.chr, dtype=torch.float64, requires_grad=True)
    k = torch.tensor(FL["k"], dtype=torch.float64, requires_grad=True)
    E = fa.solvation_energy(P, q, k)
    E.backward()
    e0, dq, dk = P.gradient(P.mesh.chr, FL["k"])
    assert abs(float(E) - e0) <= 1e-12 * abs(e0)
    assert np.max(np.abs(q.grad.numpy() - dq)) <= 1e-12 * np.max(np.abs(dq))
    assert abs(float(k.grad) - dk) <= 1e-12 * abs(dk)


@pytest.mark.parametrize("frac", ["0.45", "1"])
def test_adjoint_identity_near_modes(fa, meshdir, monkeypatch, frac):
    """The transpose identity with 45 % of the nearfield interactions streaming stored
    coefficients and with every tile streaming (1); the default computes every tile."""
    import torch
    monkeypatch.setenv("FABIPB_NEAR_STREAM", frac)
    mesh = fa.Mesh(str(meshdir / "glob4"))
    P = fa.Problem(mesh, CASES["cases"]["glob4"]["t"], order=FL["p"], sep=FL["S"], eps1=FL["eps1"],
                   eps2=FL["eps2"], qorder=FL["q"], height=FL["H"])
    op = P.operator(FL["k"])
    st = op.stats()
    assert (st["near_fly_tiles"] == 0) == (frac == "1") and st["near_stored"] > 0
    n2 = 2 * P.tree.n
    g = torch.Generator(device="cpu").manual_seed(6)
    x = (torch.rand(n2, generator=g, dtype=torch.float64) * 2 - 1).cuda()
    y = (torch.rand(n2, generator=g, dtype=torch.float64) * 2 - 1).cuda()
    Ay, ATx = torch.empty_like(x), torch.empty_like(x)
    op.matvec_dev(y, Ay)
    op.matvec_t_dev(x, ATx)
    lhs, rhs = float(torch.dot(ATx, y)), float(torch.dot(x, Ay))
    assert abs(lhs - rhs) <= 1e-11 * float(ATx.norm() * y.norm())


@pytest.mark.parametrize("name", ["sph3_ball", "glob4"])
def test_tree_gradient_approaches_direct(fa, meshdir, name):
    """The panel-tree energy's gradient approaches the direct energy's as theta shrinks
    (the walk takes more charges directly)."""
    out = {}
    for energy, theta in (("direct", 0.8), ("tree", 0.3), ("tree", 0.8)):
        P = problem(fa, meshdir, name, energy=energy, theta=theta)
        out[(energy, theta)] = P.gradient(P.mesh.chr.astype(np.float64), FL["k"])[1]
    ref = out[("direct", 0.8)]
    e3 = np.max(np.abs(out[("tree", 0.3)] - ref)) / np.max(np.abs(ref))
    e8 = np.max(np.abs(out[("tree", 0.8)] - ref)) / np.max(np.abs(ref))
    print(f"\n[{name}] tree dE/dq from the direct energy's: theta 0.3 {e3:.2e}, 0.8 {e8:.2e}", flush=True)
    assert e3 < e8 and e3 <= 1e-4
