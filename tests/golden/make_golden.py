#!/usr/bin/env python3
"""Record the golden fixtures of tests/golden/ from the reference FABIPB CPU path.

Runs in the development container only (it needs /root/reference):
  1. builds the reference binary and the fixture harness (oracle/Makefile.ref)
     and the mesh generator (cuda-fabipb_amd/bin/fbmesh);
  2. for every case of cases.json writes the mesh, then
     - runs oracle/_ref/ref_harness: layout, kernels, one matvec, RHS, energy
       of a seeded density, charge tree -> tests/golden/<case>/*.bin
     - runs oracle/_ref/fabipb_ref (-g=0 -B=1, FABIPB_ENERGY_MODE=compare,
       FABIPB_GMRES_LOG_RESID=1, FABIPB_WRITE_SOLUTION; with "tree_rhs" also
       FABIPB_FORCE_TREE_RHS=1 and the RHS sample of every row) and keeps the
       iteration count, residual history, 17-digit panel-tree energy, layout
       counts, matvec timings and (where asked) the solution vector
       -> tests/golden/<case>/cli.json (+ solution.bin)
Only numbers are written; no reference source leaves /root/reference.

usage: python tests/golden/make_golden.py [case ...]
"""
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"


def run(cmd, env=None, cwd=None):
    r = subprocess.run(cmd, env=env, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        print(r.stdout)
        raise SystemExit(f"command failed: {cmd}")
    return r.stdout


def main():
    if not os.path.isdir(REF):
        raise SystemExit("make_golden.py needs the reference at /root/reference")
    cfg = json.load(open(os.path.join(HERE, "cases.json")))
    run(["make", "-s", "-f", "oracle/Makefile.ref", "-j8"], cwd=ROOT)
    run(["make", "-s", "-C", "cuda-fabipb_amd", "bin/fbmesh"], cwd=ROOT)
    want = sys.argv[1:] or list(cfg["cases"])
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1",
               VECLIB_MAXIMUM_THREADS="1", BLIS_NUM_THREADS="1", FABIPB_REUSE_MESH="1")
    tmp = tempfile.mkdtemp(prefix="fbgold_")
    try:
        for name in want:
            case = cfg["cases"][name]
            fl = dict(cfg["flags"], **case.get("flags", {}))   # per-case overrides (restart, iteration limit)
            base = os.path.join(tmp, name)
            run([os.path.join(ROOT, "cuda-fabipb_amd/bin/fbmesh"), case["mesh"][0], base] + case["mesh"][1:])
            out = os.path.join(HERE, name)
            os.makedirs(out, exist_ok=True)
            common = [f"-t={case['t']}", f"-p={fl['p']}", f"-S={fl['S']}", f"-q={fl['q']}",
                      f"-k={fl['k']}", f"-eps1={fl['eps1']}", f"-eps2={fl['eps2']}", f"-H={fl['H']}"]
            if case.get("harness"):
                print(f"[{name}] harness", flush=True)
                run([os.path.join(ROOT, "oracle/_ref/ref_harness"), base, out] + common, env=env)
                drop = list(case.get("drop", [])) + ([] if case.get("expansions") else ["mv_mom", "mv_lec"])
                for k in drop:
                    os.remove(os.path.join(out, k + ".bin"))
                man = [l for l in open(os.path.join(out, "manifest.txt")) if l.split()[0] not in drop]
                open(os.path.join(out, "manifest.txt"), "w").writelines(man)
            if case.get("cli"):
                print(f"[{name}] reference solve", flush=True)
                sol = os.path.join(tmp, name + "_sol.csv")
                e2 = dict(env, FABIPB_ENERGY_MODE="compare", FABIPB_GMRES_LOG_RESID="1",
                          FABIPB_WRITE_SOLUTION=sol)
                rhsCsv = os.path.join(tmp, name + "_rhs.csv")
                if case.get("tree_rhs"):   # the charge-tree right-hand side, recorded row by row
                    e2.update(FABIPB_FORCE_TREE_RHS="1", FABIPB_RHS_SAMPLE_PATH=rhsCsv, FABIPB_RHS_SAMPLE_STRIDE="1")
                solve = [f"-o={fl['o']}", f"-P={fl['P']}", f"-a={fl['a']}", f"-i={fl['i']}"]
                if "r" in fl:   # dense direct operator instead of the FMM
                    solve.append(f"-r={fl['r']}")
                txt = run([os.path.join(ROOT, "oracle/_ref/fabipb_ref"), "-g=0", "-B=1"] + solve + common + [base],
                          env=e2)
                res = {"flags": common + solve}
                m = re.search(r"GMRES status: info=(\d+) iterations=(\d+) final-residual=(\S+)", txt)
                res["info"], res["iterations"], res["final_residual"] = int(m[1]), int(m[2]), float(m[3])
                res["resid_history"] = [float(v) for v in re.findall(r"GMRES residual: iter=\d+ resid=(\S+)", txt)]
                res["panel_tree_raw"] = float(re.search(r"panel-tree raw=(\S+)", txt)[1])
                m = re.search(r"leaf-cubes=(\d+) near-pairs=(\d+)", txt)
                res["leaf_cubes"], res["near_pairs"] = int(m[1]), int(m[2])
                m = re.search(r"m2l-pairs=(\d+) dst-groups=(\d+)", txt)
                res["m2l_pairs"], res["dst_groups"] = int(m[1]), int(m[2])
                res["cubes"] = int(re.search(r"FMM cube layout: cubes=(\d+)", txt)[1])
                res["max_cluster"] = int(re.search(r"finest cluster: (\d+)", txt)[1])
                res["kept_panels"] = int(re.search(r"kept=(\d+)", txt)[1])
                m = re.search(r"GMRES call counts: matvec=(\d+)", txt)
                res["matvecs"] = int(m[1])
                m = re.search(r"FMM stage avg/call \(ms\): (.*)", txt)
                if m:   # FMM runs only (not the dense direct operator)
                    res["stage_ms"] = {k: float(v) for k, v in re.findall(r"(\w+)=([\d.]+)", m[1])}
                m = re.search(r"FMM matvec stats: calls=(\d+) total=([\d.]+) s", txt)
                if m:
                    res["matvec_calls"], res["matvec_total_s"] = int(m[1]), float(m[2])
                res["gmres_s"] = float(re.search(r"gmres=([\d.]+)", txt)[1])
                res["host_threads"] = os.cpu_count()
                if case.get("solution"):
                    a = np.loadtxt(sol, delimiter=",", skiprows=1)
                    order = np.argsort(a[:, 0])
                    a = a[order]
                    res["solution_face_order"] = "idx"
                    np.concatenate([a[:, 6], a[:, 7]]).astype(np.float64).tofile(os.path.join(out, "solution.bin"))
                    a[:, 1].astype(np.int32).tofile(os.path.join(out, "solution_face.bin"))
                if case.get("tree_rhs"):
                    res["rhs_mode"] = re.search(r"setupRHS direct pairs: .* mode=(\w+)", txt)[1]
                    a = np.loadtxt(rhsCsv, delimiter=",", skiprows=1)
                    a = a[np.argsort(a[:, 0])]
                    np.concatenate([a[:, 9], a[:, 10]]).astype(np.float64).tofile(os.path.join(out, "rhs_tree.bin"))
                json.dump(res, open(os.path.join(out, "cli.json"), "w"), indent=1)
                print(f"[{name}] its={res['iterations']} raw={res['panel_tree_raw']!r}", flush=True)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main()
