#!/usr/bin/env python3
"""Benchmark: FABIPB Galerkin BIE matvec on one MI355X.

Workload: the synthetic star-shaped globule of 81,920 flat panels (level-6
icosphere, r = 22 A (1 + 0.08 sum_{l<=6} a_lm Y_lm)), 2,000 charges at least
1.5 A inside, eps_in=1, eps_out=80, kappa=0.1257, Taylor order 6, MAC 0.8,
tree depth 4 (mean 74 panels per leaf), q=1, GMRES tol 1e-6 (restart 30),
diagonal preconditioner.  Inputs come from the seeded generator
(cuda-fabipb_amd/bin/fbmesh, seed 2), so the reference and this build see
the same mesh and charges.

A step is one matvec y = (D - K) x of the plain FP64 operator with x, y
resident in HBM.  Its nearfield is one launch of 16-row tiles that evaluate
every interaction on the fly (FP64 VALU; a share of the tiles can stream
stored Galerkin coefficients instead, FABIPB_NEAR_STREAM); the roofline entry
reports the launch against the bound of its model time (the larger of
streamed bytes / 8 TB/s and computed flops / 78.6 TFLOP/s, 88 flops per
computed interaction with sqrt, exp and each divide counted as one).  `value`
is matvecs per second over the timed steps.  The
JSON line also reports the full solve (RHS + GMRES + energy) wall time, its
iteration count and energy against the reference's recorded solve, the
precision form on the same input (FP32 operator inside FP64 iterative
refinement: its solve wall time, speedup over the plain solve and energy
error against it), per-stage device times (HIP events on the launch stream),
the nearfield kernel's roofline position, and the reference CPU path timed on
the host cores.

Two further blocks: the precision form at ~317k panels (R = 40 A level-7
globule, 8000 charges, depth 5) and the differentiation form (the level-6
globule with 5000 charges: adjoint gradient dE/dq, dE/dkappa against one
forward solve).

The last block is the capsid-scale shell (2,621,440 panels, 1,000,000 charges,
order 8, MAC 0.7, depth 6): charge-tree RHS and energy, FP64 and FP32-refined
solves (skipped with less than 64 GB of free device memory; it holds ~31 GB).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-large] [--no-grad] [--no-capsid]
                       [--calib-copy]
"""
import argparse
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
MESH = ["globule", "--level", "6", "--radius", "22", "--charges", "inside", "--nq", "2000", "--sigma", "0.4",
        "--depth", "1.5", "--seed", "2"]
DEPTH, ORDER, MAC, KAPPA, EPS1, EPS2, Q, TOL, RESTART, MAXIT = 4, 6, 0.8, 0.1257, 1.0, 80.0, 1, 1e-6, 30, 100
# precision form at scale: R = 40 A level-7 globule, faces under 0.05 A^2 merged, 8000 charges, depth 5
MESH_LARGE = ["globule", "--level", "7", "--radius", "40", "--merge", "0.05", "--charges", "inside", "--nq", "8000",
              "--sigma", "0.4", "--depth", "1.5", "--seed", "3"]
DEPTH_LARGE = 5
# capsid-scale shell: outer R = 250 A and inner R = 210 A level-8 surfaces (0.03 Y_lm noise, l <= 12),
# 2,621,440 panels, 1,000,000 charges in the wall, Taylor order 8, MAC 0.7, depth 6 (84 panels per leaf)
MESH_CAPSID = ["shell", "--level", "8", "--radius", "250", "--inner", "210", "--amp", "0.03", "--lmax", "12",
               "--charges", "wall", "--nq", "1000000", "--sigma", "0.3", "--depth", "1.5", "--seed", "5"]
DEPTH_CAPSID, ORDER_CAPSID, MAC_CAPSID = 6, 8, 0.7
# differentiation form: the level-6 globule with 5000 charges
MESH_GRAD = ["globule", "--level", "6", "--radius", "22", "--charges", "inside", "--nq", "5000", "--sigma", "0.4",
             "--depth", "1.5", "--seed", "4"]
HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
CPU_CORES = 16             # the GPU box's host share for one GPU
COEF_BYTES = 32            # nearfield: KER[4] doubles streamed per (dst, src) panel pair
F_NEAR = 88                # nearfield on the fly: flops per pair (82 + sqrt + exp + 4 div, counted 1 each)
FP64_PEAK = 78.6e12        # MI355X FP64 vector, flop/s
VALU_ISSUE_PEAK = 256 * 4 * 16 * 2.4e9   # lane-instructions/s: 256 CUs x 4 SIMD16 at 2.4 GHz


def load_pkg():
    spec = importlib.util.spec_from_file_location("fabipb_amd", os.path.join(ROOT, "cuda-fabipb_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def cpu_baseline(base, deadline_s=40.0):
    """The reference CPU path (oracle/_ref/fabipb_ref, -g=0) on CPU_CORES host
    threads, stopped after 2 GMRES iterations (3 matvecs): matvecs per second
    from its own -B=1 stage accounting."""
    exe = os.path.join(ROOT, "oracle", "_ref", "fabipb_ref")
    if not os.path.exists(exe):
        return None
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1",
               VECLIB_MAXIMUM_THREADS="1", BLIS_NUM_THREADS="1", FABIPB_REUSE_MESH="1",
               FABIPB_NEARFIELD_APPLY_THREADS=str(CPU_CORES), FABIPB_SETUP_THREADS=str(CPU_CORES),
               FABIPB_GMRES_STOP_AFTER_ITER="2", FABIPB_STOP_AFTER_GMRES="1")
    cmd = [exe, "-g=0", "-B=1", f"-t={DEPTH}", f"-p={ORDER}", f"-S={MAC}", f"-q={Q}", f"-k={KAPPA}",
           f"-eps1={EPS1}", f"-eps2={EPS2}", "-H=2", f"-o={TOL}", "-P=3", f"-a={RESTART}", f"-i={MAXIT}", base]
    print(f"[bench] cpu baseline: reference CPU path, {CPU_CORES} threads, 3 matvecs", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    r = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=deadline_s * 4)
    wall = time.perf_counter() - t0
    m = re.search(r"FMM matvec stats: calls=(\d+) total=([\d.]+) s", r.stdout)
    if not m:
        print(r.stdout[-2000:], file=sys.stderr)
        return None
    calls, total = int(m[1]), float(m[2])
    return {"value": calls / total, "unit": "matvec/s", "cores": CPU_CORES, "kind": "reference",
            "sample": f"{calls} matvecs of the same 81,920-panel workload (GMRES stopped after 2 iterations), "
                      f"{total:.1f} s of matvec time, {wall:.1f} s wall"}


def calib_copy(torch, mib=1024):
    """A plain device copy of known size, for scaling the FETCH_SIZE/WRITE_SIZE counters."""
    a = torch.ones(mib * 131072, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"[bench] calibration copy: {mib} MiB read + {mib} MiB written per launch", file=sys.stderr, flush=True)


def timed_solve(fa, torch, op, n, mesh, ct, precision):
    """RHS + GMRES + panel-tree energy, each stage synchronised and timed."""
    b = torch.empty(2 * n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ts = time.perf_counter()
    op.rhs_dev(mesh.pos, mesh.chr, b)
    xs = torch.zeros_like(b) if precision == "fp32" else b.clone()   # x0: 0 refined, b plain (reference)
    torch.cuda.synchronize()
    tr = time.perf_counter()
    if precision == "fp32":
        gm = op.gmres_refined_dev(b, xs, restart=RESTART, tol=TOL, maxit=MAXIT)
    else:
        gm = op.gmres_dev(b, xs, restart=RESTART, tol=TOL, maxit=MAXIT)
    torch.cuda.synchronize()
    tg = time.perf_counter()
    raw = op.energy_dev(ct, mesh.pos, mesh.chr, xs)
    torch.cuda.synchronize()
    te = time.perf_counter()
    return {"wall_s": te - ts, "rhs_s": tr - ts, "gmres_s": tg - tr, "energy_s": te - tg, "iterations": gm["iters"],
            "final_residual": gm["resid"], "energy_raw": raw, "outer": gm.get("outer"), "fp64_matvecs": gm.get("mv64")}


def precision_large(fa, torch, tmp):
    """The precision form at ~317k panels: FP32-refined against plain FP64 solve wall time."""
    base = os.path.join(tmp, "glob7")
    subprocess.check_call([os.path.join(ROOT, "cuda-fabipb_amd", "bin", "fbmesh"), MESH_LARGE[0], base] + MESH_LARGE[1:],
                          stdout=subprocess.DEVNULL)
    mesh = fa.Mesh(base)
    tree = fa.Tree(mesh, DEPTH_LARGE, order=ORDER, sep=MAC)
    op = fa.Operator(tree, KAPPA, EPS1, EPS2, Q)
    op.enable_fp32()
    ct = tree.charge_tree()
    r = {}
    for prec in ("fp64", "fp32"):
        timed_solve(fa, torch, op, tree.n, mesh, ct, prec)
        r[prec] = timed_solve(fa, torch, op, tree.n, mesh, ct, prec)
    # the discrete solution's energy: an FP64 solve converged to 1e-10
    b = torch.empty(2 * tree.n, dtype=torch.float64, device="cuda")
    op.rhs_dev(mesh.pos, mesh.chr, b)
    xt = b.clone()
    op.gmres_dev(b, xt, restart=RESTART, tol=1e-10, maxit=400)
    e_ref = op.energy_dev(ct, mesh.pos, mesh.chr, xt)
    st = op.stats()
    out = {"panels": tree.n, "charges": int(len(mesh.chr)), "depth": DEPTH_LARGE,
           "near_interactions": st["near_interactions"], "m2l_pairs": st["m2l_pairs"],
           "fp64": r["fp64"], "fp32_refined": r["fp32"],
           "wall_speedup": r["fp64"]["wall_s"] / r["fp32"]["wall_s"],
           "energy_rel_err": abs(r["fp32"]["energy_raw"] - r["fp64"]["energy_raw"]) / abs(r["fp64"]["energy_raw"]),
           "energy_rel_err_vs_converged": abs(r["fp32"]["energy_raw"] - e_ref) / abs(e_ref),
           "fp64_energy_rel_err_vs_converged": abs(r["fp64"]["energy_raw"] - e_ref) / abs(e_ref)}
    print(f"[bench] precision form at {tree.n} panels: fp64 {r['fp64']['wall_s'] * 1e3:.2f} ms, fp32-refined "
          f"{r['fp32']['wall_s'] * 1e3:.2f} ms ({out['wall_speedup']:.2f}x), err {out['energy_rel_err']:.2e}",
          file=sys.stderr, flush=True)
    del op
    torch.cuda.empty_cache()
    return out


def capsid(fa, torch, tmp):
    """The capsid-scale shell: charge-tree RHS and energy (2.6e12 panel-charge pairs), FP64 and
    FP32-refined solves with the diagonal preconditioner."""
    free, _ = torch.cuda.mem_get_info()
    if free < 64e9:
        return {"skipped": f"{free / 1e9:.0f} GB free, the shell needs ~35 GB"}
    base = os.path.join(tmp, "capsid")
    subprocess.check_call([os.path.join(ROOT, "cuda-fabipb_amd", "bin", "fbmesh"), MESH_CAPSID[0], base] + MESH_CAPSID[1:],
                          stdout=subprocess.DEVNULL)
    t0 = time.perf_counter()
    mesh = fa.Mesh(base)
    tree = fa.Tree(mesh, DEPTH_CAPSID, order=ORDER_CAPSID, sep=MAC_CAPSID)
    t_tree = time.perf_counter() - t0
    op = fa.Operator(tree, KAPPA, EPS1, EPS2, Q)
    op.set_precond(3)
    op.enable_fp32()
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0 - t_tree
    ct = tree.charge_tree()
    n2 = 2 * tree.n
    b = torch.empty(n2, dtype=torch.float64, device="cuda")
    out = {}
    for prec in ("fp64", "fp32"):
        torch.cuda.synchronize()
        ts = time.perf_counter()
        op.rhs_tree_dev(ct, mesh.pos, mesh.chr, b)
        xs = torch.zeros_like(b) if prec == "fp32" else b.clone()
        torch.cuda.synchronize()
        tr = time.perf_counter()
        gm = (op.gmres_refined_dev if prec == "fp32" else op.gmres_dev)(b, xs, restart=RESTART, tol=TOL, maxit=200)
        torch.cuda.synchronize()
        tg = time.perf_counter()
        raw = op.energy_dev(ct, mesh.pos, mesh.chr, xs)
        torch.cuda.synchronize()
        te = time.perf_counter()
        out[prec] = {"wall_s": te - ts, "rhs_s": tr - ts, "gmres_s": tg - tr, "energy_s": te - tg,
                     "iterations": gm["iters"], "final_residual": gm["resid"], "energy_raw": raw}
    # the adjoint gradient with the direct energy (every panel x charge pair; exact, O(n m))
    import numpy as np
    q = np.ascontiguousarray(mesh.chr, dtype=np.float64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e_dir = op.energy_direct_dev(mesh.pos, q, xs)
    torch.cuda.synchronize()
    t_edir = time.perf_counter() - t0
    lam = torch.zeros_like(xs)
    t0 = time.perf_counter()
    dq, dk, ginfo = op.gradient_dev(mesh.pos, q, xs, lam, restart=RESTART, tol=TOL, maxit=200)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    # the same gradient by tree walks (panel-tree energy weights, panel FMM tree to the charges)
    lam_t = torch.zeros_like(xs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dqt, tinfo = op.gradient_tree_dev(ct, mesh.pos, q, xs, lam_t, theta=0.8, restart=RESTART, tol=TOL, maxit=200)
    torch.cuda.synchronize()
    t_gtree = time.perf_counter() - t0
    x = torch.rand(n2, dtype=torch.float64, device="cuda") * 2 - 1
    y = torch.empty_like(x)
    st = op.stats()
    res = {"panels": tree.n, "charges": int(len(mesh.chr)), "depth": DEPTH_CAPSID, "order": ORDER_CAPSID,
           "mac": MAC_CAPSID, "near_interactions": st["near_interactions"], "near_stored": st["near_stored"],
           "m2l_pairs": st["m2l_pairs"], "device_gb": st["device_bytes"] / 1e9, "tree_s": t_tree,
           "setup_s": t_setup, "charge_tree_order": int(ct["order"][-1]),
           "stages64_ms": op.profile(x, y, iters=2), "stages32_ms": op.profile(x, y, iters=2, precision="fp32"),
           "fp64": out["fp64"], "fp32_refined": out["fp32"],
           "wall_speedup": out["fp64"]["wall_s"] / out["fp32"]["wall_s"],
           "energy_kcal_per_mol": out["fp64"]["energy_raw"] * fa.ENERGY_SCALE,
           "energy_rel_err": abs(out["fp32"]["energy_raw"] - out["fp64"]["energy_raw"]) / abs(out["fp64"]["energy_raw"]),
           "adjoint_tree": {"gradient_s": t_gtree, "over_fp64_solve": t_gtree / out["fp64"]["wall_s"],
                            "adjoint_iterations": tinfo["iters"],
                            "max_rel_diff_vs_direct": float(np.max(np.abs(dqt - dq)) / np.max(np.abs(dq)))},
           "adjoint_direct": {"energy_direct_s": t_edir, "gradient_s": t_grad, "adjoint_iterations": ginfo["iters"],
                              "direct_vs_tree_energy_rel": abs(e_dir - out["fp32"]["energy_raw"]) / abs(e_dir),
                              "dEdkappa_kcal_per_mol_A": dk * fa.ENERGY_SCALE,
                              "dEdq_norm_kcal_per_mol_e": float(np.linalg.norm(dq)) * fa.ENERGY_SCALE}}
    print(f"[bench] capsid shell {tree.n} panels, {len(mesh.chr)} charges: fp64 {out['fp64']['wall_s']:.3f} s, "
          f"fp32-refined {out['fp32']['wall_s']:.3f} s ({res['wall_speedup']:.2f}x), err {res['energy_rel_err']:.2e}; "
          f"tree dE/dq {t_gtree:.3f} s (direct {t_grad:.1f} s, diff {res['adjoint_tree']['max_rel_diff_vs_direct']:.1e})",
          file=sys.stderr, flush=True)
    del op
    torch.cuda.empty_cache()
    return res


def differentiation(fa, torch, np, tmp):
    """Adjoint gradient dE/dq (every charge) and dE/dkappa against one forward solve (direct energy)."""
    base = os.path.join(tmp, "glob6q")
    subprocess.check_call([os.path.join(ROOT, "cuda-fabipb_amd", "bin", "fbmesh"), MESH_GRAD[0], base] + MESH_GRAD[1:],
                          stdout=subprocess.DEVNULL)
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

    def adjoint(u):
        lam = torch.zeros_like(u)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dq, dk, info = op.gradient_dev(mesh.pos, q, u, lam, restart=RESTART, tol=TOL, maxit=MAXIT)
        torch.cuda.synchronize()
        return dq, dk, info, time.perf_counter() - t0

    u, _, _, _ = forward()
    adjoint(u)   # first call builds the transposed store and the special-pair corrections
    torch.cuda.synchronize()
    u, e, g, t_fwd = forward()
    dq, dk, info, t_adj = adjoint(u)
    out = {"panels": P.tree.n, "charges": int(len(q)), "energy_kcal_per_mol": e * fa.ENERGY_SCALE,
           "forward_s": t_fwd, "forward_iterations": g["iters"], "adjoint_s": t_adj,
           "adjoint_iterations": info["iters"], "adjoint_over_forward": t_adj / t_fwd,
           "dEdkappa_kcal_per_mol_A": dk * fa.ENERGY_SCALE,
           "dEdq_norm_kcal_per_mol_e": float(np.linalg.norm(dq)) * fa.ENERGY_SCALE}
    print(f"[bench] differentiation at {P.tree.n} panels, {len(q)} charges: forward {t_fwd * 1e3:.2f} ms, adjoint "
          f"gradient {t_adj * 1e3:.2f} ms ({out['adjoint_over_forward']:.2f}x)", file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--calib-copy", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--no-capsid", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU (one MI355X per box); --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    fa = load_pkg()
    tmp = tempfile.mkdtemp(prefix="fbbench_")
    base = os.path.join(tmp, "glob6")
    subprocess.check_call([os.path.join(ROOT, "cuda-fabipb_amd", "bin", "fbmesh"), MESH[0], base] + MESH[1:],
                          stdout=subprocess.DEVNULL)
    print("[bench] mesh written", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    mesh = fa.Mesh(base)
    tree = fa.Tree(mesh, DEPTH, order=ORDER, sep=MAC)
    t_tree = time.perf_counter() - t0
    op = fa.Operator(tree, KAPPA, EPS1, EPS2, Q)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0 - t_tree
    st = op.stats()
    n2 = 2 * tree.n
    print(f"[bench] operator ready: n={tree.n} near={st['near_interactions']} m2l={st['m2l_pairs']} "
          f"setup={t_setup:.3f}s", file=sys.stderr, flush=True)

    g = torch.Generator(device="cuda").manual_seed(0)
    x = torch.rand(n2, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        op.matvec_dev(x, y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        op.matvec_dev(x, y)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_step = 1e3 * dt / args.steps
    print(f"[bench] timed {args.steps} matvecs: {ms_step:.3f} ms/matvec", file=sys.stderr, flush=True)

    # per-stage device times with HIP events on the launch stream
    stages = op.profile(x, y, iters=max(3, min(args.steps, 10)))
    near_ms = stages["near"]
    # the nearfield launch: stored tiles stream 32 B per interaction, the others
    # compute theirs (F_NEAR flops each); the bound is the larger model time
    stored, fly = st["near_stored"], st["near_interactions"] - st["near_stored"]
    near_kernel = "k_near_mix64" if st["near_fly_tiles"] > 0 else "k_near_apply"
    near_bytes = stored * COEF_BYTES + 2 * n2 * 8
    near_flops = fly * F_NEAR
    achieved = near_bytes / (near_ms * 1e-3)

    def run_solve(precision):
        b = torch.empty(n2, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        ts = time.perf_counter()
        op.rhs_dev(mesh.pos, mesh.chr, b)
        xs = torch.zeros_like(b) if precision == "fp32" else b.clone()
        torch.cuda.synchronize()
        tr = time.perf_counter()
        if precision == "fp32":
            gm = op.gmres_refined_dev(b, xs, restart=RESTART, tol=TOL, maxit=MAXIT)
        else:
            gm = op.gmres_dev(b, xs, restart=RESTART, tol=TOL, maxit=MAXIT)
        torch.cuda.synchronize()
        tg = time.perf_counter()
        raw = op.energy_dev(ct, mesh.pos, mesh.chr, xs)
        torch.cuda.synchronize()
        te = time.perf_counter()
        r = {"wall_s": te - ts, "rhs_s": tr - ts, "gmres_s": tg - tr, "energy_s": te - tg,
             "iterations": gm["iters"], "final_residual": gm["resid"], "energy_raw": raw,
             "energy_kcal_per_mol": raw * fa.ENERGY_SCALE}
        if precision == "fp32":
            r.update(outer=gm["outer"], fp64_matvecs=gm["mv64"], fp32_matvecs=gm["iters"] + 1)
        else:
            r["matvecs"] = gm["iters"] + 1
        return r

    solve = precision = None
    if not args.no_solve:
        ct = tree.charge_tree()
        run_solve("fp64")   # first call allocates the GMRES workspace; time the second
        solve = run_solve("fp64")
        raw = solve["energy_raw"]
        ref = os.path.join(ROOT, "tests", "golden", "cfg2_globule", "cli.json")
        if os.path.exists(ref):
            R = json.load(open(ref))
            solve["reference_iterations"] = R["iterations"]
            solve["energy_rel_err_vs_reference"] = abs(raw - R["panel_tree_raw"]) / abs(R["panel_tree_raw"])
            solve["reference_cpu_gmres_s_8threads"] = R["gmres_s"]
        print(f"[bench] fp64 solve: {solve['wall_s'] * 1e3:.2f} ms its={solve['iterations']}", file=sys.stderr,
              flush=True)

        # precision form: FP32 operator, FP64 refinement
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.enable_fp32()
        torch.cuda.synchronize()
        t_fp32 = time.perf_counter() - t0
        for _ in range(args.warmup):
            op.matvec32_dev(x, y)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            op.matvec32_dev(x, y)
        torch.cuda.synchronize()
        ms32 = 1e3 * (time.perf_counter() - t0) / args.steps
        run_solve("fp32")
        precision = run_solve("fp32")
        precision.update(
            energy_rel_err_vs_fp64=abs(precision["energy_raw"] - raw) / abs(raw),
            wall_speedup_vs_fp64=solve["wall_s"] / precision["wall_s"],
            gmres_speedup_vs_fp64=solve["gmres_s"] / precision["gmres_s"],
            fp32_matvec_ms=ms32, fp32_matvec_per_s=1e3 / ms32, fp32_setup_s=t_fp32)
        print(f"[bench] fp32-refined solve: {precision['wall_s'] * 1e3:.2f} ms its={precision['iterations']} "
              f"outer={precision['outer']} err={precision['energy_rel_err_vs_fp64']:.2e} "
              f"matvec32={ms32:.3f} ms", file=sys.stderr, flush=True)

    large = grad = shell = None
    if not args.no_solve and not args.no_large:
        large = precision_large(fa, torch, tmp)
    if not args.no_solve and not args.no_grad:
        grad = differentiation(fa, torch, np, tmp)
    if not args.no_solve and not args.no_capsid:
        del op
        torch.cuda.empty_cache()
        shell = capsid(fa, torch, tmp)

    if args.calib_copy:
        calib_copy(torch)

    cpu = None if args.no_cpu else cpu_baseline(base)

    # measured HBM traffic (rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled by a 1 GiB copy) and
    # VALU instructions per launch on this workload, from the committed counter summary
    traffic, valu = None, None
    tpath = os.path.join(ROOT, "profiles", f"{near_kernel}_traffic_globule81920.md")
    if os.path.exists(tpath):
        for line in open(tpath):
            if line.startswith("{") and '"fetch_kib"' in line:
                t = json.loads(line)
                if t["interactions"] == st["near_interactions"] and t.get("stored", stored) == stored:
                    traffic = t["fetch_kib"] * 1024 * t["fetch_scale"] + t["write_kib"] * 1024 * t["write_scale"]
                    valu = t.get("valu_insts")
    hbm_bound = near_bytes / HBM_PEAK >= near_flops / FP64_PEAK
    if hbm_bound:
        roof = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": near_bytes / HBM_PEAK / (near_ms * 1e-3)}
    else:   # FP64 VALU: algorithmic flops of the computed interactions per second
        roof = {"bound": "fp64-valu", "achieved": near_flops / (near_ms * 1e-3) / 1e12, "peak": FP64_PEAK / 1e12,
                "unit": "TFLOP/s", "frac": near_flops / FP64_PEAK / (near_ms * 1e-3)}

    out = {
        "metric": METRIC,
        "value": args.steps / dt,
        "unit": "matvec/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, identical input to the reference run)",
        "config": {"workload": "FP64 plain matvec, star-shaped globule level-6 icosphere R=22 A, 81920 panels, "
                               "2000 charges, eps 1/80, kappa 0.1257, Taylor order 6, MAC 0.8, depth 4, q=1",
                   "panels": tree.n, "charges": int(len(mesh.chr)), "depth": DEPTH, "order": ORDER, "mac": MAC,
                   "near_interactions": st["near_interactions"], "m2l_pairs": st["m2l_pairs"]},
        "roofline": dict(roof, kernel=near_kernel, traffic=traffic, algorithmic_bytes=near_bytes,
                         algorithmic_flops=near_flops, kernel_ms=near_ms, streamed_interactions=stored,
                         computed_interactions=fly, hbm_gbps=achieved / 1e9,
                         valu_insts_per_interaction=(valu * 64 / st["near_interactions"] if valu else None),
                         valu_issue_frac=(valu * 64 / (near_ms * 1e-3) / VALU_ISSUE_PEAK if valu else None)),
        "cpu_baseline": cpu,
        "stages_ms": stages,
        "setup_s": {"tree": t_tree, "device_operator": t_setup},
        "solve": solve,
        "precision_form": precision,
        "precision_form_large": large,
        "differentiation_form": grad,
        "capsid_shell": shell,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
