"""GPU: the FMM tree with its panel pass on the device (fbh_tree_panels +
fb_tree_build_leaves) against the host build, array by array, on every recorded
case and on the 317,112-panel precision-form mesh (depth 5).  The capsid-scale
shell (2.6M panels, depth 6) is compared in test_gpu_configs.py."""
import numpy as np
import pytest

from conftest import CASES

pytestmark = pytest.mark.gpu


def _same(fa, a, b):
    for k in fa.TREE_KEYS:
        assert np.array_equal(a.a[k], b.a[k]), k


@pytest.mark.parametrize("name", sorted(CASES["cases"]))
def test_device_tree_matches_host(fa, meshdir, name):
    case = CASES["cases"][name]
    fl = dict(CASES["flags"], **case.get("flags", {}))
    mesh = fa.Mesh(str(meshdir / name))
    kw = dict(order=fl["p"], height=fl["H"], sep=fl["S"])
    _same(fa, fa.Tree(mesh, case["t"], device=True, **kw), fa.Tree(mesh, case["t"], **kw))


def test_device_tree_deep(fa, meshdir):
    """Depths 1 and 7 on one mesh: one digit per key, and leaves of a few panels."""
    mesh = fa.Mesh(str(meshdir / "glob4"))
    for depth, height in ((1, 0), (7, 2)):
        _same(fa, fa.Tree(mesh, depth, height=height, device=True), fa.Tree(mesh, depth, height=height))
