#!/usr/bin/env python3
"""Benchmark: FABIPB's PB solve on one MI355X, precision form against plain FP64.

Headline workload (the metric's configuration): the precision form on the
synthetic star-shaped globule of R = 40 A (level-7 icosphere, r = R (1 + 0.08
sum_{l<=6} a_lm Y_lm), faces under 0.05 A^2 merged: 317,112 flat panels), 8,000
charges at least 1.5 A inside, eps_in = 1, eps_out = 80, kappa = 0.1257, Taylor
order 6, MAC 0.8, tree depth 5, q = 1, GMRES tol 1e-6 (restart 30), diagonal
preconditioner.  Inputs come from the seeded generator (cuda-fabipb_amd/bin/
fbmesh, seed 3), so the reference run recorded in tests/golden/cfg3_precision
saw the same mesh and charges.

A step is one complete PB solve of the precision form with every input resident
in HBM: the direct right-hand side (r in FP64, its kernel in FP32), iterative
refinement around GMRES on the FP32 operator (FP64 Krylov basis, Gram-Schmidt,
Givens rotations and the FP64 residual test b - A64 x), and the panel-tree
solvation energy in FP64.
`value` is solves per second over the timed steps.  The same number of plain
FP64 solves (GMRES on the FP64 operator, x0 = b as the reference) is timed the
same way beside it; the line reports the wall-time ratio and the energy errors
(FP32-refined against the FP64 solve, against the reference's own recorded
17-digit energy, and against an FP64 solve converged to 1e-10).

`roofline` is the dominant kernel of the timed solves, the FP32 nearfield
(k_near_mix32): 88 algorithmic flops per interaction (kernelKER4 with sqrt,
exp and each divide counted as one) over its average launch time, measured with
HIP events on its own stream around every launch of the timed region
(fbh_near_timing), against the FP32 peak.  `matvec` adds the whole-operator
roofline position of both operators (the algorithmic time of every stage over
the measured back-to-back matvec time) and per-stage device times.
`cpu_baseline` is the reference CPU path (oracle/_ref/fabipb_ref, built from the
reference sources) on the host cores available to this process, timed on a
bounded sample of the same workload (two GMRES matvecs).

Further blocks: the 81,920-panel globule (level 6, R = 22 A, 2,000 charges,
depth 4: FP64 and FP32 matvec rates, both solves), the differentiation form (the
level-6 globule with 5,000 charges: the exact adjoint gradient dE/dq and dE/dkappa
of the panel-tree energy against one forward solve) and the capsid-scale shell (2,621,440 panels, 1,000,000
charges, order 8, MAC 0.7, depth 6; skipped with less than 64 GB free).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-globule] [--no-grad]
                       [--no-capsid] [--calib-copy]
"""
import argparse
import hashlib
import importlib.util
import json
import math
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "PB solve wall time & matvec/s at 1e-6 energy err, FP32-refined vs FP64 plain"
KAPPA, EPS1, EPS2, Q, TOL, RESTART, MAXIT, ORDER, MAC = 0.1257, 1.0, 80.0, 1, 1e-6, 30, 100, 6, 0.8
# headline: R = 40 A level-7 globule, faces under 0.05 A^2 merged, 8000 charges, depth 5
MESH_LARGE = ["globule", "--level", "7", "--radius", "40", "--merge", "0.05", "--charges", "inside", "--nq", "8000",
              "--sigma", "0.4", "--depth", "1.5", "--seed", "3"]
DEPTH_LARGE = 5
# the level-6 globule: R = 22 A, 81,920 panels, 2000 charges, depth 4
MESH = ["globule", "--level", "6", "--radius", "22", "--charges", "inside", "--nq", "2000", "--sigma", "0.4",
        "--depth", "1.5", "--seed", "2"]
DEPTH = 4
# capsid-scale shell: outer R = 250 A and inner R = 210 A level-8 surfaces (0.03 Y_lm noise, l <= 12),
# 2,621,440 panels, 1,000,000 charges in the wall, Taylor order 8, MAC 0.7, depth 6 (84 panels per leaf)
MESH_CAPSID = ["shell", "--level", "8", "--radius", "250", "--inner", "210", "--amp", "0.03", "--lmax", "12",
               "--charges", "wall", "--nq", "1000000", "--sigma", "0.3", "--depth", "1.5", "--seed", "5"]
DEPTH_CAPSID, ORDER_CAPSID, MAC_CAPSID = 6, 8, 0.7
# differentiation form: the level-6 globule with 5000 charges
MESH_GRAD = ["globule", "--level", "6", "--radius", "22", "--charges", "inside", "--nq", "5000", "--sigma", "0.4",
             "--depth", "1.5", "--seed", "4"]
HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
FP64_PEAK = 78.6e12        # MI355X FP64 vector (= matrix), flop/s
FP32_PEAK = 157.3e12       # MI355X FP32 vector, flop/s (the rate of packed FP32; one FP32 op per lane runs at 78.6)
F_NEAR = 88                # nearfield on the fly: flops per pair (82 + sqrt + exp + 4 div, counted 1 each)
CPU_SHARE = 16             # host cores per GPU on the MI355X boxes
SRC = os.path.join(ROOT, "cuda-fabipb_amd", "csrc", "fabipb_hip.hip")


def load_pkg():
    spec = importlib.util.spec_from_file_location("fabipb_amd", os.path.join(ROOT, "cuda-fabipb_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def host_cores():
    """Cores this process may run on, capped at the host share of one GPU."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, CPU_SHARE))


def make_mesh(spec, tmp, name):
    base = os.path.join(tmp, name)
    subprocess.check_call([os.path.join(ROOT, "cuda-fabipb_amd", "bin", "fbmesh"), spec[0], base] + spec[1:],
                          stdout=subprocess.DEVNULL)
    return base


def cpu_baseline(base, depth, panels, timeout_s=300):
    """The reference CPU path (oracle/_ref/fabipb_ref, -g=0) on the host cores: GMRES
    stopped after one iteration (two matvecs: the initial residual and one Krylov
    step); matvecs per second from its own -B=1 stage accounting.  The charge-tree
    right-hand side is forced so that the sample is not dominated by the direct
    right-hand side (68 s on 8 cores at 317k panels); it does not enter the matvec."""
    exe = os.path.join(ROOT, "oracle", "_ref", "fabipb_ref")
    if not os.path.exists(exe):
        return None
    cores = host_cores()
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1",
               VECLIB_MAXIMUM_THREADS="1", BLIS_NUM_THREADS="1", FABIPB_REUSE_MESH="1",
               FABIPB_NEARFIELD_APPLY_THREADS=str(cores), FABIPB_SETUP_THREADS=str(cores),
               FABIPB_GMRES_STOP_AFTER_ITER="1", FABIPB_STOP_AFTER_GMRES="1", FABIPB_FORCE_TREE_RHS="1")
    cmd = [exe, "-g=0", "-B=1", f"-t={depth}", f"-p={ORDER}", f"-S={MAC}", f"-q={Q}", f"-k={KAPPA}",
           f"-eps1={EPS1}", f"-eps2={EPS2}", "-H=2", f"-o={TOL}", "-P=3", f"-a={RESTART}", f"-i={MAXIT}", base]
    print(f"[bench] cpu baseline: reference CPU path, {cores} threads, 2 matvecs at {panels} panels",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    try:
        r = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           timeout=timeout_s)
    except subprocess.TimeoutExpired:
        return None
    wall = time.perf_counter() - t0
    m = re.search(r"FMM matvec stats: calls=(\d+) total=([\d.]+) s", r.stdout)
    if not m:
        print(r.stdout[-2000:], file=sys.stderr)
        return None
    calls, total = int(m[1]), float(m[2])
    near = re.search(r"FMM stage avg/call \(ms\):.*Near=([\d.]+)", r.stdout)
    return {"value": calls / total, "unit": "matvec/s", "cores": cores, "kind": "reference",
            "sample": f"{calls} FP64 matvecs of the same {panels}-panel workload (GMRES stopped after one "
                      f"iteration), {total:.1f} s of matvec time, {wall:.1f} s wall",
            "near_ms_per_matvec": float(near[1]) if near else None}


def matvec_roofline(st, n, nmom, ms, precision):
    """Whole-operator roofline position: the sum of every stage's algorithmic time at
    the MI355X peaks over the measured matvec time.  Nearfield: F_NEAR flops per
    interaction; M2L: 8 C(p+6,6) flops per pair plus 6 C(p+4,4) for its derivative
    tensors; Q2M + L2P: both moment matrices read once (4 n nMom words); M2M + L2L:
    12 C(p+6,6) flops per parent-child link; the combine: six 2n-vectors of FP64."""
    p = ORDER
    P = FP32_PEAK if precision == "fp32" else FP64_PEAK
    w = 4 if precision == "fp32" else 8
    parts = {
        "near": st["near_interactions"] * F_NEAR / P,
        "m2l": st["m2l_pairs"] * (8 * math.comb(p + 6, 6) + 6 * math.comb(p + 4, 4)) / P,
        "q2m_l2p": 4 * n * nmom * w / HBM_PEAK,
        "m2m_l2l": st["links"] * 12 * math.comb(p + 6, 6) / P,
        "combine": 2 * n * 8 * 6 / HBM_PEAK,
    }
    t_roof = sum(parts.values())
    return {"t_roof_ms": t_roof * 1e3, "t_measured_ms": ms, "achieved": t_roof * 1e3 / ms,
            "parts_ms": {k: v * 1e3 for k, v in parts.items()}, "peak_tflops": P / 1e12}


def traffic_from_profile(kernel, workload):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary of this kernel and workload, only when it was taken of the current
    kernel source (its src_sha equals this tree's); None otherwise."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic_{workload}.md")
    if not os.path.exists(path):
        return None, None
    sha = hashlib.sha256(open(SRC, "rb").read()).hexdigest()[:16]
    for line in open(path):
        if line.startswith("{") and '"fetch_kib"' in line:
            t = json.loads(line)
            if t.get("src_sha") != sha:
                return None, f"{os.path.relpath(path, ROOT)} is of another kernel source ({t.get('src_sha')})"
            return (t["fetch_kib"] * 1024 * t["fetch_scale"] + t["write_kib"] * 1024 * t["write_scale"],
                    os.path.relpath(path, ROOT))
    return None, None


class Solver:
    """One solve with every input resident: the charge set staged once (fbh_set_charges),
    the right-hand side, x and the solution in preallocated device tensors."""

    def __init__(self, fa, torch, op, tree, mesh, tree_rhs=False, maxit=MAXIT):
        self.torch, self.op, self.n2, self.maxit = torch, op, 2 * tree.n, maxit
        self.ct = tree.charge_tree()
        op.set_charges(self.ct, mesh.pos, mesh.chr)
        self.b = torch.empty(self.n2, dtype=torch.float64, device="cuda")
        self.x = torch.empty_like(self.b)
        self.tree_rhs = tree_rhs

    def rhs(self, precision):
        if self.tree_rhs:
            self.op.rhs_tree_res_dev(self.b)
        else:   # the precision form's direct right-hand side has its kernel in FP32
            (self.op.rhs32_res_dev if precision == "fp32" else self.op.rhs_res_dev)(self.b)

    def gmres(self, precision, tol=TOL, maxit=None):
        maxit = maxit or self.maxit
        if precision == "fp32":
            self.x.zero_()   # the first correction solves A d = b
            return self.op.gmres_refined_dev(self.b, self.x, restart=RESTART, tol=tol, maxit=maxit)
        self.x.copy_(self.b)   # x0 = b, the reference's default
        return self.op.gmres_dev(self.b, self.x, restart=RESTART, tol=tol, maxit=maxit)

    def energy(self, precision="fp64"):   # the precision form's walk has its far polynomials in FP32
        return self.op.energy_res_dev(self.x, precision=precision)

    def solve(self, precision, tol=TOL, maxit=None):
        self.rhs(precision)
        if precision == "fp32":   # the energy walk runs beside the refinement's last FP64 residual
            self.x.zero_()
            g = self.op.gmres_refined_energy_dev(self.b, self.x, restart=RESTART, tol=tol, maxit=maxit or self.maxit)
            return g, g["energy_raw"]
        g = self.gmres(precision, tol, maxit)
        return g, self.energy(precision)

    def staged(self, precision):
        """One solve with a synchronisation between the stages, for the breakdown."""
        sync, pc = self.torch.cuda.synchronize, time.perf_counter
        sync()
        t0 = pc()
        self.rhs(precision)
        sync()
        t1 = pc()
        g = self.gmres(precision)
        sync()
        t2 = pc()
        e = self.energy(precision)
        sync()
        t3 = pc()
        r = {"solve_ms": (t3 - t0) * 1e3, "rhs_ms": (t1 - t0) * 1e3, "gmres_ms": (t2 - t1) * 1e3,
             "energy_ms": (t3 - t2) * 1e3, "iterations": g["iters"], "final_residual": g["resid"], "info": g["info"],
             "energy_raw": e, "energy_kcal_per_mol": e * 2 * math.pi * 332.0716}
        if precision == "fp32":
            r.update(outer=g["outer"], fp64_matvecs=g["mv64"], fp32_matvecs=g["iters"])
        else:
            r["matvecs"] = g["iters"] + 1
        return r


def timed_solves(torch, S, op, precision, steps, warmup):
    for _ in range(warmup):
        S.solve(precision)
    torch.cuda.synchronize()
    op.near_timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        S.solve(precision)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    op.near_timing(False)
    return dt, op.near_timing_read()


def calib_copy(torch, mib=1024):
    """A plain device copy of known size, for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    a = torch.ones(mib * 131072, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"[bench] calibration copy: {mib} MiB read + {mib} MiB written per launch", file=sys.stderr, flush=True)


def matvec_rate(torch, f, x, y, steps, warmup):
    for _ in range(warmup):
        f(x, y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        f(x, y)
    torch.cuda.synchronize()
    return 1e3 * (time.perf_counter() - t0) / steps


def headline(fa, torch, args, tmp):
    """The precision form at 317,112 panels: FP32-refined solves (the timed steps), FP64
    plain solves beside them, energies, matvec rates and rooflines."""
    base = make_mesh(MESH_LARGE, tmp, "glob7")
    t0 = time.perf_counter()
    mesh = fa.Mesh(base)
    tree = fa.Tree(mesh, DEPTH_LARGE, order=ORDER, sep=MAC)
    t_tree = time.perf_counter() - t0
    op = fa.Operator(tree, KAPPA, EPS1, EPS2, Q)
    op.enable_fp32()
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0 - t_tree
    st = op.stats()
    n = tree.n
    print(f"[bench] operator ready: n={n} near={st['near_interactions']} m2l={st['m2l_pairs']} "
          f"setup={t_setup:.3f}s", file=sys.stderr, flush=True)
    S = Solver(fa, torch, op, tree, mesh)

    dt32, nt32 = timed_solves(torch, S, op, "fp32", args.steps, args.warmup)
    print(f"[bench] timed {args.steps} FP32-refined solves: {dt32 / args.steps * 1e3:.2f} ms/solve",
          file=sys.stderr, flush=True)
    dt64, nt64 = timed_solves(torch, S, op, "fp64", args.steps, args.warmup)
    print(f"[bench] timed {args.steps} FP64 solves: {dt64 / args.steps * 1e3:.2f} ms/solve", file=sys.stderr,
          flush=True)
    r32, r64 = S.staged("fp32"), S.staged("fp64")
    # the discrete solution's energy: an FP64 solve converged to 1e-10
    g_conv, e_conv = S.solve("fp64", tol=1e-10, maxit=400)
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "cfg3_precision", "cli.json")))
    e_ref = ref["panel_tree_raw"]
    pf = {"fp32_refined": r32, "fp64_plain": dict(r64, reference_iterations=ref["iterations"]),
          "fp32_refined_ms_per_solve": dt32 / args.steps * 1e3, "fp64_plain_ms_per_solve": dt64 / args.steps * 1e3,
          "wall_speedup": dt64 / dt32,
          "energy_rel_err_vs_fp64": abs(r32["energy_raw"] - r64["energy_raw"]) / abs(r64["energy_raw"]),
          "energy_rel_err_vs_reference": abs(r32["energy_raw"] - e_ref) / abs(e_ref),
          "fp64_energy_rel_err_vs_reference": abs(r64["energy_raw"] - e_ref) / abs(e_ref),
          "energy_rel_err_vs_converged": abs(r32["energy_raw"] - e_conv) / abs(e_conv),
          "fp64_energy_rel_err_vs_converged": abs(r64["energy_raw"] - e_conv) / abs(e_conv),
          "converged_iterations": g_conv["iters"],
          "reference_cpu_gmres_s_8threads": ref["gmres_s"]}
    print(f"[bench] precision form: fp64 {pf['fp64_plain_ms_per_solve']:.2f} ms, fp32-refined "
          f"{pf['fp32_refined_ms_per_solve']:.2f} ms ({pf['wall_speedup']:.2f}x), energy err vs fp64 "
          f"{pf['energy_rel_err_vs_fp64']:.2e}, vs reference {pf['energy_rel_err_vs_reference']:.2e}",
          file=sys.stderr, flush=True)

    # matvec rates, stage times and whole-operator rooflines
    g = torch.Generator(device="cuda").manual_seed(0)
    x = torch.rand(2 * n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    y = torch.empty_like(x)
    ms64 = matvec_rate(torch, op.matvec_dev, x, y, 100, 20)
    ms32 = matvec_rate(torch, op.matvec32_dev, x, y, 100, 20)
    y64 = torch.empty_like(x)
    op.matvec_dev(x, y64)
    op.matvec32_dev(x, y)
    mv = {"fp64_per_s": 1e3 / ms64, "fp32_per_s": 1e3 / ms32, "fp64_ms": ms64, "fp32_ms": ms32,
          "fp32_operator_rel_err": float((y - y64).norm() / y64.norm()),
          "stages64_ms": op.profile(x, y, iters=10), "stages32_ms": op.profile(x, y, iters=10, precision="fp32"),
          "roofline_fp64": matvec_roofline(st, n, st["nmom_leaf"], ms64, "fp64"),
          "roofline_fp32": matvec_roofline(st, n, st["nmom_leaf"], ms32, "fp32")}

    # the dominant kernel of the timed solves: the FP32 nearfield, timed live
    launches, tot = nt32["fp32"]
    k_ms = tot / launches
    flops = st["near_interactions"] * F_NEAR
    traffic, tsrc = traffic_from_profile("k_near_mix32", "globule317112")
    roof = {"bound": "fp32-valu", "achieved": flops / (k_ms * 1e-3) / 1e12, "peak": FP32_PEAK / 1e12,
            "unit": "TFLOP/s", "frac": flops / FP32_PEAK / (k_ms * 1e-3), "traffic": traffic,
            "traffic_source": tsrc, "kernel": "k_near_mix32", "kernel_ms": k_ms, "launches": launches,
            "algorithmic_flops": flops, "algorithmic_bytes": 2 * n * 8 + 2 * n * 4,
            "fp64_launches": nt32["fp64"][0],
            "fp64_kernel_ms": nt32["fp64"][1] / nt32["fp64"][0] if nt32["fp64"][0] else None,
            "fp64_plain_kernel_ms": nt64["fp64"][1] / nt64["fp64"][0] if nt64["fp64"][0] else None,
            "kernel_ms_alone": mv["stages32_ms"]["near"],
            "frac_alone": flops / FP32_PEAK / (mv["stages32_ms"]["near"] * 1e-3),
            "note": "kernel_ms: the launches inside the timed solves, beside the far chain on the other "
                    "stream; kernel_ms_alone: the same launch with the GPU to itself (stage profile)"}
    cfg = {"workload": "precision form: FP32-refined PB solve against FP64 plain, star-shaped globule R=40 A "
                       "level-7 icosphere with faces under 0.05 A^2 merged (317,112 panels), 8,000 charges, "
                       "eps 1/80, kappa 0.1257, Taylor order 6, MAC 0.8, depth 5, q=1, GMRES tol 1e-6 restart 30, "
                       "diagonal preconditioner, inputs resident in HBM",
           "panels": n, "charges": int(len(mesh.chr)), "depth": DEPTH_LARGE, "order": ORDER, "mac": MAC,
           "near_interactions": st["near_interactions"], "m2l_pairs": st["m2l_pairs"]}
    out = {"dt": dt32, "config": cfg, "roofline": roof, "precision_form": pf, "matvec": mv,
           "setup_s": {"tree": t_tree, "device_operator": t_setup}, "base": base, "panels": n}
    del S, op
    torch.cuda.empty_cache()
    return out


def globule(fa, torch, args, tmp):
    """The level-6 globule (81,920 panels): matvec rates and both solves."""
    base = make_mesh(MESH, tmp, "glob6")
    mesh = fa.Mesh(base)
    tree = fa.Tree(mesh, DEPTH, order=ORDER, sep=MAC)
    op = fa.Operator(tree, KAPPA, EPS1, EPS2, Q)
    op.enable_fp32()
    st = op.stats()
    x = torch.rand(2 * tree.n, dtype=torch.float64, device="cuda") * 2 - 1
    y = torch.empty_like(x)
    ms64 = matvec_rate(torch, op.matvec_dev, x, y, 200, 50)
    ms32 = matvec_rate(torch, op.matvec32_dev, x, y, 200, 50)
    S = Solver(fa, torch, op, tree, mesh)
    for prec in ("fp64", "fp32"):
        S.solve(prec)
    r64, r32 = S.staged("fp64"), S.staged("fp32")
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "cfg2_globule", "cli.json")))
    out = {"panels": tree.n, "charges": int(len(mesh.chr)), "depth": DEPTH,
           "near_interactions": st["near_interactions"], "m2l_pairs": st["m2l_pairs"],
           "fp64_matvec_per_s": 1e3 / ms64, "fp32_matvec_per_s": 1e3 / ms32, "fp64_matvec_ms": ms64,
           "fp32_matvec_ms": ms32, "stages64_ms": op.profile(x, y, iters=10),
           "stages32_ms": op.profile(x, y, iters=10, precision="fp32"),
           "roofline_fp64": matvec_roofline(st, tree.n, st["nmom_leaf"], ms64, "fp64"),
           "fp64": dict(r64, reference_iterations=ref["iterations"],
                        energy_rel_err_vs_reference=abs(r64["energy_raw"] - ref["panel_tree_raw"])
                        / abs(ref["panel_tree_raw"])),
           "fp32_refined": dict(r32, energy_rel_err_vs_reference=abs(r32["energy_raw"] - ref["panel_tree_raw"])
                                / abs(ref["panel_tree_raw"])),
           "wall_speedup": r64["solve_ms"] / r32["solve_ms"]}
    print(f"[bench] globule {tree.n}: fp64 matvec {ms64:.3f} ms, fp32 {ms32:.3f} ms; solves {r64['solve_ms']:.2f} / "
          f"{r32['solve_ms']:.2f} ms", file=sys.stderr, flush=True)
    del S, op
    torch.cuda.empty_cache()
    return out


def capsid(fa, torch, tmp):
    """The capsid-scale shell: charge-tree RHS and energy (2.6e12 panel-charge pairs), FP64 and
    FP32-refined solves with the diagonal preconditioner, the adjoint dE/dq by tree walks and
    by the direct sums."""
    free, _ = torch.cuda.mem_get_info()
    if free < 64e9:
        return {"skipped": f"{free / 1e9:.0f} GB free, the shell needs ~35 GB"}
    base = make_mesh(MESH_CAPSID, tmp, "capsid")
    t0 = time.perf_counter()
    mesh = fa.Mesh(base)
    t_load = time.perf_counter() - t0
    tree = fa.Tree(mesh, DEPTH_CAPSID, order=ORDER_CAPSID, sep=MAC_CAPSID, device=True)
    t_tree = time.perf_counter() - t0 - t_load
    op = fa.Operator(tree, KAPPA, EPS1, EPS2, Q)
    op.set_precond(3)
    op.enable_fp32()
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0 - t_load - t_tree
    t1 = time.perf_counter()
    fa.Tree(mesh, DEPTH_CAPSID, order=ORDER_CAPSID, sep=MAC_CAPSID)   # the host build, for comparison
    t_tree_host = time.perf_counter() - t1
    S = Solver(fa, torch, op, tree, mesh, tree_rhs=True, maxit=200)
    out = {prec: S.staged(prec) for prec in ("fp64", "fp32")}
    xs = S.x
    import numpy as np
    q = np.ascontiguousarray(mesh.chr, dtype=np.float64)
    # the exact gradient of the reported (panel-tree) energy, charge-tree right-hand side
    for rep in range(2):   # the first call builds the transposed operator's stores
        lam_t = torch.zeros_like(xs)   # each call solves the adjoint from zero
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dqt, dkt, tinfo = op.gradient_tree_dev(S.ct, mesh.pos, q, xs, lam_t, theta=0.8, rhs_tree=True,
                                               restart=RESTART, tol=TOL, maxit=200)
        torch.cuda.synchronize()
        t_gtree = time.perf_counter() - t0
    # exactness at scale: the transposed energy walk against forward walks of unit charges
    dq0 = op.tree_adjoint_dev(S.ct, mesh.pos, q, xs, kind=0)
    worst = 0.0
    for j in np.random.default_rng(7).choice(len(q), size=3, replace=False):
        e = np.zeros_like(q)
        e[j] = 1.0
        worst = max(worst, abs(dq0[j] - op.energy_dev(S.ct, mesh.pos, e, xs)))
    worst /= np.max(np.abs(dq0))
    x = torch.rand(2 * tree.n, dtype=torch.float64, device="cuda") * 2 - 1
    y = torch.empty_like(x)
    st = op.stats()
    res = {"panels": tree.n, "charges": int(len(mesh.chr)), "depth": DEPTH_CAPSID, "order": ORDER_CAPSID,
           "mac": MAC_CAPSID, "near_interactions": st["near_interactions"], "near_stored": st["near_stored"],
           "m2l_pairs": st["m2l_pairs"], "device_gb": st["device_bytes"] / 1e9, "mesh_load_s": t_load, "tree_s": t_tree, "tree_host_s": t_tree_host,
           "setup_s": t_setup, "charge_tree_order": int(S.ct["order"][-1]),
           "stages64_ms": op.profile(x, y, iters=2), "stages32_ms": op.profile(x, y, iters=2, precision="fp32"),
           "fp64": out["fp64"], "fp32_refined": out["fp32"],
           "wall_speedup": out["fp64"]["solve_ms"] / out["fp32"]["solve_ms"],
           "energy_rel_err": abs(out["fp32"]["energy_raw"] - out["fp64"]["energy_raw"]) / abs(out["fp64"]["energy_raw"]),
           "adjoint_tree": {"gradient_s": t_gtree, "over_fp64_solve": t_gtree * 1e3 / out["fp64"]["solve_ms"],
                            "adjoint_iterations": tinfo["iters"],
                            "dEdkappa_kcal_per_mol_A": dkt * fa.ENERGY_SCALE,
                            "dEdq_norm_kcal_per_mol_e": float(np.linalg.norm(dqt)) * fa.ENERGY_SCALE,
                            "transposed_walk_vs_unit_charges": worst}}
    print(f"[bench] capsid shell {tree.n} panels, {len(mesh.chr)} charges: fp64 {out['fp64']['solve_ms']:.1f} ms, "
          f"fp32-refined {out['fp32']['solve_ms']:.1f} ms ({res['wall_speedup']:.2f}x), err {res['energy_rel_err']:.2e}; "
          f"exact tree gradient {t_gtree:.3f} s ({res['adjoint_tree']['over_fp64_solve']:.2f}x the FP64 solve; "
          f"transposed walk vs unit charges {worst:.1e})",
          file=sys.stderr, flush=True)
    del S, op
    torch.cuda.empty_cache()
    return res


def differentiation(fa, torch, np, tmp):
    """Adjoint gradient dE/dq (every charge) and dE/dkappa of the panel-tree energy against one
    forward solve (right-hand side, GMRES, energy)."""
    base = make_mesh(MESH_GRAD, tmp, "glob6q")
    mesh = fa.Mesh(base)
    P = fa.Problem(mesh, DEPTH, order=ORDER, sep=MAC, eps1=EPS1, eps2=EPS2, qorder=Q, tol=TOL, restart=RESTART,
                   maxit=MAXIT)
    op = P.operator(KAPPA)
    q = np.ascontiguousarray(mesh.chr, dtype=np.float64)

    def forward():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        u, e, g = P.solve(q, KAPPA)
        torch.cuda.synchronize()
        return u, e, g, time.perf_counter() - t0

    def adjoint(u, e):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, dq, dk = P.gradient(q, KAPPA, u=u, e=e)
        torch.cuda.synchronize()
        return dq, dk, time.perf_counter() - t0

    u, e, _, _ = forward()
    adjoint(u, e)   # first call builds the transposed store and the special-pair corrections
    torch.cuda.synchronize()
    u, e, g, t_fwd = forward()
    dq, dk, t_adj = adjoint(u, e)
    out = {"panels": P.tree.n, "charges": int(len(q)), "energy": "panel tree, theta 0.8",
           "energy_kcal_per_mol": e * fa.ENERGY_SCALE, "forward_s": t_fwd, "forward_iterations": g["iters"],
           "adjoint_s": t_adj, "adjoint_over_forward": t_adj / t_fwd, "dEdkappa_kcal_per_mol_A": dk,
           "dEdq_norm_kcal_per_mol_e": float(np.linalg.norm(dq))}
    print(f"[bench] differentiation at {P.tree.n} panels, {len(q)} charges: forward {t_fwd * 1e3:.2f} ms, adjoint "
          f"gradient {t_adj * 1e3:.2f} ms ({out['adjoint_over_forward']:.2f}x)", file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-globule", action="store_true")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--no-capsid", action="store_true")
    ap.add_argument("--calib-copy", action="store_true", help="a 1 GiB device copy for counter calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU (one MI355X per box); --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    fa = load_pkg()
    tmp = tempfile.mkdtemp(prefix="fbbench_")
    h = headline(fa, torch, args, tmp)
    glob = None if args.no_globule else globule(fa, torch, args, tmp)
    grad = None if args.no_grad else differentiation(fa, torch, np, tmp)
    shell = None if args.no_capsid else capsid(fa, torch, tmp)
    if args.calib_copy:
        calib_copy(torch)
    cpu = None if args.no_cpu else cpu_baseline(h["base"], DEPTH_LARGE, h["panels"])
    if cpu:
        cpu["gpu_fp64_matvec_per_s"] = h["matvec"]["fp64_per_s"]

    out = {
        "metric": METRIC,
        "value": args.steps / h["dt"],
        "unit": "solve/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * h["dt"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64",
        "data": "synthetic (seeded generator, identical input to the reference run)",
        "config": h["config"],
        "roofline": h["roofline"],
        "cpu_baseline": cpu,
        "precision_form": h["precision_form"],
        "matvec": h["matvec"],
        "setup_s": h["setup_s"],
        "globule81920": glob,
        "differentiation_form": grad,
        "capsid_shell": shell,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
