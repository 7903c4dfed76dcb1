"""GPU: the no-allocation right-hand side (fbh_rhs_dev_noalloc, through the C ABI)
against fbh_rhs_dev on the same charges: bit-identical (the same kernel on the
same data, staged in the static device array instead of per-call allocations),
on two calls in a row, the second with fewer charges than the first; a set
beyond the static array's size is refused with an error before any device work."""
import numpy as np
import pytest

from conftest import CASES

pytestmark = pytest.mark.gpu
FL = CASES["flags"]


def test_rhs_noalloc_matches_rhs(fa, meshdir):
    import torch
    name = "glob4"
    print(f"\n[gpu] no-allocation right-hand side on {name}", flush=True)
    mesh = fa.Mesh(str(meshdir / name))
    tree = fa.Tree(mesh, CASES["cases"][name]["t"], order=FL["p"], height=FL["H"], sep=FL["S"])
    op = fa.Operator(tree, FL["k"], FL["eps1"], FL["eps2"], FL["q"])
    b_ref = torch.empty(2 * tree.n, dtype=torch.float64, device="cuda")
    b_na = torch.full_like(b_ref, float("nan"))
    nq = len(mesh.chr)
    assert nq >= 2
    for m in (nq, nq // 2):   # the second call stages fewer charges over the first call's
        pos, q = mesh.pos[:m], mesh.chr[:m]
        op.rhs_dev(pos, q, b_ref)
        op.rhs_dev_noalloc(pos, q, b_na)
        assert torch.equal(b_ref, b_na), m


def test_rhs_noalloc_refuses_oversize(fa, meshdir):
    import torch
    name = "glob4"
    mesh = fa.Mesh(str(meshdir / name))
    tree = fa.Tree(mesh, CASES["cases"][name]["t"], order=FL["p"], height=FL["H"], sep=FL["S"])
    op = fa.Operator(tree, FL["k"], FL["eps1"], FL["eps2"], FL["q"])
    b = torch.zeros(2 * tree.n, dtype=torch.float64, device="cuda")
    m = 262144 + 1
    with pytest.raises(fa.FabipbError, match="exceed the static staging array"):
        op.rhs_dev_noalloc(np.zeros((m, 3)), np.zeros(m), b)
    assert float(b.abs().sum()) == 0.0
