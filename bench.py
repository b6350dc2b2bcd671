"""Frequency points solved per second on the 20480-triangle pulsating sphere
(icosphere level 5, radius 0.5 m, 1000 Hz, ka = 9.16; leaf 64, eta 2, ACA
1e-5, GMRES 1e-8, Burton-Miller), one MI355X.

A step is one frequency point: Phi assembly, b = s + Phi f, DPhi assembly,
GMRES to tolerance, boundary-condition reconstruction and sound pressure
(boundaryelementsolver.cpp:526-744, 2251-2261).  The per-mesh cluster tree,
partition and near-singular geometry are set up once, before the timed
region, and reported as setup.

Prints one JSON line (rank 0).  --gpus N > 1 is refused: this build runs on
one GPU.
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# algorithmic flop counts per entry (each sqrt, sin, cos, div counts as one)
F_REG = 16 * 95 + 40
F_SELF = 3 * 32 * 12 + 60


def f_near(nsub):
    return nsub * 25 * 95 + 150


def workload(level):
    bem = importlib.import_module("3d_bem_amd")
    tri = bem.meshes.icosphere(level, 0.5)
    bc = bem.meshes.robin_bc(len(tri), 0.0, 1.0, 1.0)
    return bem, tri, bc


def run_cpu_baseline(tri, bc, args):
    """The CPU restatement (oracle/) on the same workload, all host threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    P = oracle.Problem(tri, bc, leaf=args.leaf, eta=args.eta)
    P.set_frequency(args.freq, 343.0, 1)
    _, _, pr, st = P.solve(args.aca_tol, gmres_tol=args.gmres_tol, seed=args.seed, rho=1.21, freq=args.freq,
                           nthreads=threads)
    t = time.perf_counter() - t0
    return pr, t, threads, st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--level", type=int, default=5)
    ap.add_argument("--freq", type=float, default=1000.0)
    ap.add_argument("--leaf", type=int, default=64)
    ap.add_argument("--eta", type=float, default=2.0)
    ap.add_argument("--aca-tol", type=float, default=1e-5)
    ap.add_argument("--gmres-tol", type=float, default=1e-8)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (no multi-GPU path)"}))
        sys.exit(2)

    bem, tri, bc = workload(args.level)
    n = len(tri)
    prec = bem.FP64 if args.precision == "fp64" else bem.FP32
    t0 = time.perf_counter()
    S = bem.Solver(tri, bc, leaf=args.leaf, eta=args.eta)
    setup_s = time.perf_counter() - t0
    _, blocks = S.layout()
    ctx_h = bem.lib().bem_solver_context(S.h)

    def solve():
        return S.solve(args.freq, c=343.0, rho=1.21, coupling=1, aca_tol=args.aca_tol, gmres_tol=args.gmres_tol,
                       seed=args.seed, precision=prec)

    for _ in range(args.warmup):
        solve()
    prof = np.zeros(12)
    bem.lib().bem_profile(ctx_h, 1, prof.ctypes.data_as(__import__("ctypes").c_void_p))
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, _, pr, st = solve()
        stats.append(st)
    elapsed = time.perf_counter() - t0
    bem.lib().bem_profile(ctx_h, 0, prof.ctypes.data_as(__import__("ctypes").c_void_p))
    avg_ms, cnt = prof[:6], prof[6:]
    families = ["dense_assembly", "near_singular_fixup", "aca", "mvm_dense", "mvm_lowrank", "mvm_gather"]
    tot = avg_ms * cnt
    dom = int(np.argmax(tot))
    st = stats[-1]

    # algorithmic work per launch of each family
    dense = blocks[blocks[:, 4] == 0]
    n_dense = int((dense[:, 1] * dense[:, 3]).sum())
    nb = [len(blocks)]
    rl = {}
    if dom in (0, 1, 2):
        if dom == 0:
            flops = F_REG * (n_dense - n)  # regular entries of the dense blocks (self/near done by the fix-up)
        elif dom == 2:
            ev = (st["evaluated_phi"] + st["evaluated_dphi"]) / 2 - n_dense
            flops = F_REG * ev
        else:
            flops = F_SELF * n + f_near(2.5) * 12 * n
        peak = PEAK_FP64_TFLOPS if prec == bem.FP64 else PEAK_FP32_TFLOPS
        ach = flops / (avg_ms[dom] * 1e-3) / 1e12
        rl = {"bound": "mfma", "achieved": round(ach, 3), "peak": peak, "unit": "TFLOP/s",
              "frac": round(ach / peak, 4), "traffic": None}
    else:
        s = 16 if prec == bem.FP64 else 8
        if dom == 3:
            byts = s * n_dense + 16 * n_dense / 64
        elif dom == 4:
            byts = st["dphi_bytes"] - s * n_dense
        else:
            byts = 16 * 2 * n
        ach = byts / (avg_ms[dom] * 1e-3) / 1e9
        rl = {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
              "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": None}
    rl["kernel"] = families[dom]
    rl["avg_ms"] = round(float(avg_ms[dom]), 4)

    out = {
        "metric": "freq-points solved/s at 20k tris (FP32+refine vs FP64); rel-L2 err vs FP64 oracle",
        "value": round(args.steps / elapsed, 4),
        "unit": "freq-points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == bem.FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"pulsating sphere, icosphere level {args.level} ({n} triangles), r=0.5 m, "
                               f"{args.freq:g} Hz, Burton-Miller, leaf {args.leaf}, eta {args.eta:g}, "
                               f"ACA {args.aca_tol:g}, GMRES {args.gmres_tol:g}, {args.precision}",
                   "triangles": n, "setup_s": round(setup_s, 3),
                   "gmres_iters": st["gmres_iters"], "residual": st["residual"],
                   "max_rank_phi": st["max_rank_phi"], "max_rank_dphi": st["max_rank_dphi"],
                   "phase_ms": {k: round(st[k], 2) for k in ("t_phi_asm", "t_rhs_mvm", "t_dphi_asm", "t_gmres")},
                   "kernel_ms": {f: [round(float(a), 4), int(c)] for f, a, c in zip(families, avg_ms, cnt)}},
        "roofline": rl,
    }
    if not args.no_cpu_baseline:
        pr_cpu, t_cpu, threads, st_cpu = run_cpu_baseline(tri, bc, args)
        out["cpu_baseline"] = {"value": round(1.0 / t_cpu, 5), "unit": "freq-points/s", "cores": threads,
                               "kind": "port",
                               "sample": f"one full frequency point of the same workload ({n} triangles) with the "
                                         f"CPU restatement, {t_cpu:.1f} s"}
        out["rel_l2_vs_cpu"] = float(np.linalg.norm(pr - pr_cpu) / np.linalg.norm(pr_cpu))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
