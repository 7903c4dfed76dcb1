import importlib.util
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.join(ROOT, "tests")
sys.path.insert(0, TESTS)

import golden_io  # noqa: E402

CASES = json.load(open(os.path.join(TESTS, "golden", "cases.json")))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _built():
    need = ["cuda-fabipb_amd/lib/libfabipb_hip.so", "cuda-fabipb_amd/lib/libfabipb_host.so",
            "cuda-fabipb_amd/bin/fbmesh", "oracle/_build/liboracle.so"]
    return all(os.path.exists(os.path.join(ROOT, p)) for p in need)


@pytest.fixture(scope="session")
def built():
    """Libraries and tools are built in-tree; build them only when missing."""
    if not _built():
        subprocess.check_call(["make", "-s", "-C", "cuda-fabipb_amd", "-j8"], cwd=ROOT)
        subprocess.check_call(["make", "-s", "-f", "oracle/Makefile"], cwd=ROOT)
    return True


@pytest.fixture(scope="session")
def fa(built):
    spec = importlib.util.spec_from_file_location("fabipb_amd", os.path.join(ROOT, "cuda-fabipb_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def meshdir(tmp_path_factory, built):
    """Writes every golden case's mesh with the seeded generator."""
    d = tmp_path_factory.mktemp("meshes")
    for name, case in CASES["cases"].items():
        subprocess.check_call([os.path.join(ROOT, "cuda-fabipb_amd/bin/fbmesh"), case["mesh"][0],
                               str(d / name)] + case["mesh"][1:], stdout=subprocess.DEVNULL)
    return d


def splitmix_vec(seed, n):
    """The harness's SplitMix64 test-vector stream: 2*u - 1 from the top 53 bits."""
    M = (1 << 64) - 1
    out = np.empty(n)
    s = seed
    for i in range(n):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        out[i] = 2.0 * ((z >> 11) * 2.0 ** -53) - 1.0
    return out


_fix = {}


def fixture(name):
    """Golden arrays of a case, with the seeded vectors regenerated."""
    if name not in _fix:
        d = os.path.join(TESTS, "golden", name)
        F = golden_io.load_dir(d) if os.path.exists(os.path.join(d, "manifest.txt")) else {}
        if "counts" in F:
            n = int(F["counts"][0])
            if "mv_x" not in F:
                F["mv_x"] = splitmix_vec(12345, 2 * n)
            if "en_sigma" not in F and "en_raw" in F:
                F["en_sigma"] = splitmix_vec(12346, 2 * n)
        cli = os.path.join(TESTS, "golden", name, "cli.json")
        if os.path.exists(cli):
            F["cli"] = json.load(open(cli))
        sol = os.path.join(TESTS, "golden", name, "solution.bin")
        if os.path.exists(sol):
            F["solution"] = np.fromfile(sol)
        _fix[name] = F
    return _fix[name]


def harness_cases():
    return [k for k, v in CASES["cases"].items() if v.get("harness")]


def cli_cases():
    return [k for k, v in CASES["cases"].items() if v.get("cli")]


def case_flags(name):
    """The common solve flags with the case's own overrides (GMRES restart -a, limit -i)."""
    return dict(CASES["flags"], **CASES["cases"][name].get("flags", {}))


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))
