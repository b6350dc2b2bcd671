"""Frequency points solved per second on the 20480-triangle pulsating sphere
(icosphere level 5, radius 0.5 m, ka = 9.16 at 1000 Hz; leaf 64, eta 2, ACA
1e-5, GMRES 1e-8, Burton-Miller), one MI355X.

A step is one batch of --batch frequency points spread over 960-1040 Hz
(1000 Hz included).  Each point is the full per-frequency path: Phi
assembly, b = s + Phi f, DPhi assembly, GMRES to tolerance, boundary-
condition reconstruction and sound pressure (boundaryelementsolver.cpp:
526-744, 2251-2261).  The per-mesh cluster tree, partition and near-singular
geometry are set up once, before the timed region, and reported as setup.

The headline is the mixed-precision batched form (--precision fp32,
default): regular far-field quadrature, ACA and block storage in complex
FP32; self and near-singular entries and (with --refine 1, the default) the
whole near field in FP64; GMRES in FP64 arithmetic iterating on the FP32
operator inside an iterative refinement whose residuals b - A x are taken in
FP64 with the FP64 near field and the stored low-rank factors (--refine 2:
with FP64 H-matrices; 0: no refinement).  The batch's points go in chunks
of --sweep-batch (default 2): Phi and DPhi of a chunk's frequencies from
one batched assembly each (one ACA launch per size class over the
flattened frequency x block items, the near field of both operators from
one quadrature), and one lockstep GMRES over the chunk whose Arnoldi
columns replay a captured HIP graph with every operator's product fused
into one launch per kernel; the chunks run concurrently on up to --lanes
solver lanes (four for the default 8 points), with the ACA schedule tuned on
the first chunk during the warm-up.  The earlier scheduling, one point per
lane on eight lanes, is timed beside it (lanes_leg); the benchmark runs as
a child process with GPU_MAX_HW_QUEUES=16 so that lanes do not share
hardware queues.  The plain
FP64 path (one point after the other) is timed beside it on the same points
(fp64_value, speedup_vs_fp64), the unrefined FP32 form too (fp32_norefine);
the 1000 Hz surface pressures are compared with the FP64 H-matrix path, with
an FP64 dense-assembly solve (bem_solver_solve_dense: every entry of the
dense operators, GMRES to 1e-12) and with the FP64 CPU restatement run on
the host (cpu_baseline leg, all host threads and one thread).

--plane D puts a rigid impedance plane D metres below the sphere's centre
(the plane z = -D, the single-plane branch of hMatrixSolve,
boundaryelementsolver.cpp:549-557, 594-604): every entry then carries the
mirror image's regular quadrature as well.  --field NOBS times the field
evaluation at NOBS exterior points instead, --mvm the H-MVM of the assembled
DPhi operator (FP32 and FP64 storage, against the HBM peak).

Prints one JSON line.  --gpus N > 1 is refused: the path runs on one GPU.
"""
import argparse
import ctypes
import importlib
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# Each solver lane launches on three streams; with HIP's default of four
# hardware queues per process the streams of different lanes share queues
# and serialise.  The HIP runtime takes the queue count from the environment
# the process starts with, so unless it already is 16, main() runs the
# benchmark as a child process with GPU_MAX_HW_QUEUES=16 (measured: 11 % more
# points/s than 4 with eight lanes; 32 was slower).  The parent touches no
# GPU; the child's output and exit status pass through.
HW_QUEUES = "16"

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}   # MI355X vector peaks
PEAK_HBM_GBS = 8000.0
# algorithmic flop count of one regular entry (16 points x 95 + 40; sqrt,
# sin, cos and div count as one flop each)
F_REG = 16 * 95 + 40
FAMILIES = ["dense_near_field", "singular_table", "aca", "mvm_dense", "mvm_lowrank", "mvm_gather", "field"]
# algorithmic flop count of one (observation point, element) pair of the
# field evaluation: 7 points x (15 to form the point + 38 kernel) + 4 (area)
# + 14 (M phi - L dphi)
F_FIELD = 7 * (15 + 38) + 18


def workload(level):
    bem = importlib.import_module("3d_bem_amd")
    tri = bem.meshes.icosphere(level, 0.5)
    bc = bem.meshes.robin_bc(len(tri), 0.0, 1.0, 1.0)
    return bem, tri, bc


def plane_of(args):
    """(point, normal) of the rigid plane z = -D, or None"""
    if args.plane is None:
        return None
    return np.array([0.0, 0.0, -args.plane]), np.array([0.0, 0.0, 1.0])


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def run_cpu_baseline(tri, bc, args, threads=None):
    """The CPU restatement (oracle/) on the same workload, all host threads
    (or `threads`): one full frequency point."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    if threads is None:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    P = oracle.Problem(tri, bc, leaf=args.leaf, eta=args.eta)
    P.set_frequency(args.freq, 343.0, 1)
    if plane_of(args) is not None:
        P.set_plane(*plane_of(args))
    _, _, pr, st = P.solve(args.aca_tol, gmres_tol=args.gmres_tol, seed=args.seed, rho=1.21, freq=args.freq,
                           nthreads=threads)
    t = time.perf_counter() - t0
    return pr, t, threads, st


def run_field(args):
    """--field NOBS: the field evaluation (bem_solver_field: the direct sum
    of calculateFieldSolutionRegular on the device) of the solved workload at
    NOBS seeded exterior points, --steps timed; k_field against the FP64
    vector peak; the CPU restatement's field (oracle/, all host threads) on a
    bounded sample of the same points with the same boundary solution."""
    bem, tri, bc = workload(args.level)
    n = len(tri)
    S = bem.Solver(tri, bc, leaf=args.leaf, eta=args.eta)
    if plane_of(args) is not None:
        S.set_impedance_plane(*plane_of(args))
    prec = bem.FP32 if args.precision == "fp32" else bem.FP64
    phi, dphi, _, _ = S.solve(args.freq, c=343.0, rho=1.21, coupling=1, aca_tol=args.aca_tol,
                              gmres_tol=args.gmres_tol, seed=args.seed, precision=prec)
    rng = np.random.default_rng(args.seed)
    d = rng.standard_normal((args.field, 3))
    obs = d / np.linalg.norm(d, axis=1)[:, None] * rng.uniform(0.6, 3.0, (args.field, 1))
    ctx_h = bem.lib().bem_solver_context(S.h)
    for _ in range(args.warmup):
        S.field(obs)
    prof = np.zeros(14)
    bem.lib().bem_profile(ctx_h, 1, prof.ctypes.data_as(ctypes.c_void_p))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        f = S.field(obs)
    elapsed = time.perf_counter() - t0
    bem.lib().bem_profile(ctx_h, 0, prof.ctypes.data_as(ctypes.c_void_p))
    avg_ms, cnt = float(prof[6]), int(prof[13])
    flops = float(F_FIELD) * args.field * n
    ach = flops / (avg_ms * 1e-3) / 1e12
    out = {
        "metric": "field points evaluated/s at 20k tris (direct sum, FP64)",
        "value": round(args.field * args.steps / elapsed, 1),
        "unit": "field-points/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"field of the pulsating sphere, icosphere level {args.level} ({n} triangles), "
                               f"{args.freq:g} Hz boundary solution ({args.precision}), {args.field} points at "
                               f"0.6-3 m (seeded)" + ("" if args.plane is None else f", rigid plane {args.plane:g} m "
                                                      "below the centre"),
                   "triangles": n, "points": args.field},
        "roofline": {"bound": "valu", "achieved": round(ach, 3), "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
                     "frac": round(ach / PEAK_TFLOPS["fp64"], 4), "traffic": None, "kernel": "field",
                     "avg_ms": round(avg_ms, 4), "launches": cnt,
                     "algorithmic_per_launch": f"{flops:.4g} flop ({args.field} x {n} pairs x {F_FIELD})"},
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        P = oracle.Problem(tri, bc, leaf=args.leaf, eta=args.eta)
        P.set_frequency(args.freq, 343.0, 1)
        if plane_of(args) is not None:
            P.set_plane(*plane_of(args))
        m = min(args.field, 2000)
        t0 = time.perf_counter()
        fc = P.field(obs[:m], phi, dphi)
        t_cpu = time.perf_counter() - t0
        fc = fc * (1j * 1.21 * 2 * np.pi * args.freq)
        out["cpu_baseline"] = {"value": round(m / t_cpu, 1), "unit": "field-points/s", "cores": threads,
                               "kind": "port",
                               "sample": f"the first {m} of the points with the same boundary solution, CPU "
                                         f"restatement: {t_cpu:.2f} s"}
        out["rel_l2_vs_cpu"] = float(np.linalg.norm(f[:m] - fc) / np.linalg.norm(fc))
    print(json.dumps(out))


def run_mvm(args):
    """--mvm: H-MVM throughput of the DPhi operator at 1000 Hz, FP32 and FP64
    storage: one single-lane solve per precision (--steps times, after
    --warmup) with the kernel families timed by HIP events on the context
    stream; an MVM = near-field tiles + low-rank phases + gather (the GMRES
    MVMs dominate the counts; the one right-hand-side MVM with Phi is of the
    same size).  Bytes per MVM: the operator's storage plus three n-vectors
    of complex FP64 (x, y and the partial sums are re-read from cache)."""
    bem, tri, bc = workload(args.level)
    n = len(tri)
    S = bem.Solver(tri, bc, leaf=args.leaf, eta=args.eta)
    ctx_h = bem.lib().bem_solver_context(S.h)
    res = {}
    for name, prec in (("fp32", bem.FP32), ("fp64", bem.FP64)):
        def solve():
            return S.solve(args.freq, c=343.0, rho=1.21, coupling=1, aca_tol=args.aca_tol,
                           gmres_tol=args.gmres_tol, seed=args.seed, precision=prec)
        for _ in range(args.warmup):
            solve()
        prof = np.zeros(14)
        bem.lib().bem_profile(ctx_h, 1, prof.ctypes.data_as(ctypes.c_void_p))
        for _ in range(args.steps):
            _, _, _, st = solve()
        bem.lib().bem_profile(ctx_h, 0, prof.ctypes.data_as(ctypes.c_void_p))
        avg, cnt = prof[:7], prof[7:]
        ms = float(avg[3] + avg[4] + avg[5])   # dense tiles, low-rank phases, gather
        nbytes = float(st["dphi_bytes"]) + 3.0 * 16.0 * n
        res[name] = {"ms_per_mvm": ms, "mvms": int(cnt[5]), "bytes": nbytes,
                     "gbs": nbytes / (ms * 1e-3) / 1e9, "family_ms": [round(float(avg[q]), 4) for q in (3, 4, 5)]}
    r = res["fp32"]
    out = {
        "metric": "H-MVMs/s of the DPhi operator at 20k tris (FP32 storage)",
        "value": round(1e3 / r["ms_per_mvm"], 1), "unit": "MVM/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_mvm"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c64 storage, f64 vectors",
        "data": "synthetic",
        "config": {"workload": f"DPhi H-matrix of the pulsating sphere, icosphere level {args.level} ({n} triangles), "
                               f"{args.freq:g} Hz, leaf {args.leaf}, eta {args.eta:g}, ACA {args.aca_tol:g}, "
                               "one solve per step, one lane",
                   "triangles": n, "mvms_timed": r["mvms"],
                   "family_ms_dense_lowrank_gather": r["family_ms"],
                   "fp64": {"mvm_per_s": round(1e3 / res["fp64"]["ms_per_mvm"], 1),
                            "ms_per_mvm": round(res["fp64"]["ms_per_mvm"], 4),
                            "bytes_per_mvm": res["fp64"]["bytes"], "gbs": round(res["fp64"]["gbs"], 1),
                            "family_ms_dense_lowrank_gather": res["fp64"]["family_ms"]}},
        "roofline": {"bound": "hbm", "achieved": round(r["gbs"], 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(r["gbs"] / PEAK_HBM_GBS, 4), "traffic": None, "kernel": "mvm",
                     "avg_ms": round(r["ms_per_mvm"], 4),
                     "algorithmic_per_launch": f"{r['bytes']:.4g} B (DPhi storage + 3 n-vectors)"},
    }
    print(json.dumps(out), flush=True)


def batch_freqs(center, batch):
    """batch points over [0.96, 1.04] x center with center itself included"""
    if batch == 1:
        return np.array([center])
    f = center * (0.96 + 0.08 * np.arange(batch) / (batch - 1))
    f[np.argmin(np.abs(f - center))] = center
    return f


def timed_run(bem, S, ctx_h, args, prec, freqs, lanes, sweep_batch=0):
    """warmup + timed steps (one sweep of the batch each); kernel family
    times from HIP events recorded around every launch of the timed steps on
    the launching streams of all the solver's lanes (bem_solver_profile)"""
    S.set_sweep_batch(sweep_batch)

    def step():
        return S.sweep(freqs, c=343.0, rho=1.21, coupling=1, aca_tol=args.aca_tol, gmres_tol=args.gmres_tol,
                       seed=args.seed, precision=prec, lanes=lanes)

    for _ in range(args.warmup):
        step()
    prof = np.zeros(14)
    bem.lib().bem_solver_profile(S.h, 1, prof.ctypes.data_as(ctypes.c_void_p))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pr, its, stats = step()
    elapsed = time.perf_counter() - t0
    bem.lib().bem_solver_profile(S.h, 0, prof.ctypes.data_as(ctypes.c_void_p))
    return pr, its, stats, elapsed, prof[:7].copy(), prof[7:].copy()


def roofline(avg_ms, cnt, st, blocks, n, prec_name, points0, traffic_sched, quads=1, points_per_s=None):
    """Dominant kernel family (largest total device time) against its bound.
    Assembly families: algorithmic flops (F_REG per regular quadrature,
    quads of them per regular entry: 2 with an impedance plane) over the
    family's mean launch duration, vector-ALU peak of the precision.  MVM
    families: algorithmic bytes of the stored operator over the duration."""
    tot = avg_ms * cnt
    dom = int(np.argmax(tot))
    s = 16 if prec_name == "fp64" else 8
    dense = blocks[blocks[:, 4] == 0]
    n_dense = int((dense[:, 1] * dense[:, 3]).sum())
    launches = cnt[dom] / points0        # launches of this family per frequency point
    if dom in (0, 2):
        if dom == 0:
            entries = 2 * (n_dense - n)  # regular entries of the Phi and DPhi near field
        else:
            entries = st["evaluated_phi"] + st["evaluated_dphi"] - 2 * n_dense
        flops = quads * F_REG * entries / launches
        ach = flops / (avg_ms[dom] * 1e-3) / 1e12
        peak = PEAK_TFLOPS[prec_name]
        rl = {"bound": "valu", "achieved": round(ach, 3), "peak": peak, "unit": "TFLOP/s",
              "frac": round(ach / peak, 4), "traffic": None,
              "algorithmic_per_launch": f"{flops:.4g} flop ({entries / launches:.4g} regular entries x "
                                        f"{quads} x {F_REG})"}
        if points_per_s:
            # the lanes' launches overlap: the family's flops per point times
            # the points per second is the chip-wide rate
            rl["achieved_chip_wide"] = round(flops * launches * points_per_s / 1e12, 3)
    elif dom == 1:
        rl = {"bound": "valu", "achieved": None, "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s", "frac": None,
              "traffic": None}
    else:
        if dom == 3:
            byts = s * n_dense
        elif dom == 4:
            byts = st["dphi_bytes"] - s * n_dense
        else:
            byts = 16 * 2 * n
        ach = byts / (avg_ms[dom] * 1e-3) / 1e9
        rl = {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
              "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": None,
              "algorithmic_per_launch": f"{byts:.4g} B"}
    if dom == 2 and prec_name == "fp32" and traffic_sched:
        rl["traffic"] = measured_traffic("aca", "fp32", traffic_sched)
    rl["kernel"] = FAMILIES[dom]
    rl["avg_ms"] = round(float(avg_ms[dom]), 4)
    rl["launches_per_point"] = round(float(launches), 2)
    return rl


def measured_traffic(family, prec_name, sched):
    """HBM bytes per launch of a kernel family under one scheduling (sched:
    e.g. "batch4 s384 h2048 q0") from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (scaled by the copy calibration), or None"""
    key = f"traffic_bytes_per_launch {family} {prec_name} {sched} ="
    path = os.path.join(ROOT, "profiles", f"{family}_{prec_name}_sphere20480_traffic.md")
    try:
        for line in open(path):
            if line.startswith(key):
                return float(line.split("=")[1])
    except OSError:
        pass
    return None


def copy_calibration(mib):
    """hipMemcpy device-to-device of mib MiB, five times: the runtime's plain
    copy kernel reads and writes exactly mib MiB per launch"""
    bem = importlib.import_module("3d_bem_amd")
    hip = ctypes.CDLL("libamdhip64.so")
    ctx = bem.Context(0)
    nbytes = mib << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (src, dst):
        if bem.lib().bem_malloc(ctx.h, ctypes.c_size_t(nbytes), ctypes.byref(p)) != 0:
            raise RuntimeError("bem_malloc failed")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    for _ in range(5):
        rc = hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3)   # hipMemcpyDeviceToDevice
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed: {rc}")
    for p in (src, dst):
        bem.lib().bem_free(ctx.h, p)
    print(json.dumps({"copy_calibration_bytes": nbytes, "launches": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=8, help="frequency points per step")
    ap.add_argument("--lanes", type=int, default=8, help="concurrent solver lanes of the FP32 form")
    ap.add_argument("--sweep-batch", type=int, default=2, metavar="B",
                    help="FP32 form: frequency chunks of B points, each with batched assembly and one lockstep GMRES, "
                         "the chunks on up to --lanes lanes (0: one point per lane)")
    ap.add_argument("--aca-schedule", default=None, metavar="S,H,Q",
                    help="a fixed ACA schedule instead of tuning: one-wavefront bound S, largest-class bound H "
                         "(0: none), work queue Q (0 / 1)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched sweeps: keep the built-in ACA schedule (default: tuned on the first chunk, in the "
                         "warm-up)")
    ap.add_argument("--leg-batch", type=int, default=8, metavar="B",
                    help="chunk size of the batched comparison leg (when the headline runs on lanes)")
    ap.add_argument("--no-batch-leg", action="store_true",
                    help="skip the comparison run of the other FP32 scheduling (lanes or batched chunks)")
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--level", type=int, default=5)
    ap.add_argument("--freq", type=float, default=1000.0)
    ap.add_argument("--leaf", type=int, default=64)
    ap.add_argument("--eta", type=float, default=2.0)
    ap.add_argument("--aca-tol", type=float, default=1e-5)
    ap.add_argument("--gmres-tol", type=float, default=1e-8)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--plane", type=float, default=None, metavar="D",
                    help="rigid impedance plane D m below the sphere's centre (default: none)")
    ap.add_argument("--field", type=int, default=0, metavar="NOBS",
                    help="time the field evaluation at NOBS points instead of the sweep")
    ap.add_argument("--mvm", action="store_true",
                    help="time the H-MVM of the assembled DPhi operator (FP32 and FP64 storage) instead of the sweep")
    ap.add_argument("--refine", type=int, choices=[0, 1, 2], default=1,
                    help="FP32 form: 0 no refinement, 1 FP64 residuals with the FP64 near field and the low-rank "
                         "factors (default), 2 FP64 residuals with FP64 H-matrices")
    ap.add_argument("--no-fp64-leg", action="store_true", help="skip the plain FP64 comparison run")
    ap.add_argument("--no-norefine-leg", action="store_true", help="skip the unrefined FP32 comparison run")
    ap.add_argument("--no-dense", action="store_true", help="skip the FP64 dense-assembly comparison solve")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cpu-1thread", action="store_true", help="skip the one-thread CPU restatement leg")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (the plain copy kernel that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if os.environ.get("GPU_MAX_HW_QUEUES") != HW_QUEUES and not args.copy_calibration:
        env = dict(os.environ, GPU_MAX_HW_QUEUES=HW_QUEUES)
        sys.exit(subprocess.run([sys.executable, os.path.abspath(__file__)] + sys.argv[1:], env=env).returncode)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "the path runs on one GPU; --gpus must be 1 (no multi-GPU path)"}))
        sys.exit(2)
    if args.field:
        run_field(args)
        return
    if args.mvm:
        run_mvm(args)
        return

    bem, tri, bc = workload(args.level)
    n = len(tri)
    t0 = time.perf_counter()
    S = bem.Solver(tri, bc, leaf=args.leaf, eta=args.eta)
    if plane_of(args) is not None:
        S.set_impedance_plane(*plane_of(args))
    setup_s = time.perf_counter() - t0
    _, blocks = S.layout()
    ctx_h = bem.lib().bem_solver_context(S.h)
    prec = bem.FP32 if args.precision == "fp32" else bem.FP64
    freqs = batch_freqs(args.freq, args.batch)
    ic = int(np.argmin(np.abs(freqs - args.freq)))
    batched = prec == bem.FP32 and args.sweep_batch > 1
    lanes = min(args.lanes, -(-args.batch // args.sweep_batch)) if batched else args.lanes if prec == bem.FP32 else 1
    sb = args.sweep_batch if batched else 0
    refine = args.refine if prec == bem.FP32 else 0
    S.set_refinement(refine)
    S.set_autotune(batched and not args.no_autotune and not args.aca_schedule)
    if batched and args.aca_schedule:
        sm, hu, qu = (int(v) for v in args.aca_schedule.split(","))
        S.set_aca_schedule(prec, sm, hu if hu > 0 else 1 << 40, qu)

    pr, its, stats, elapsed, avg_ms, cnt = timed_run(bem, S, ctx_h, args, prec, freqs, lanes, sb)
    rstats = S.refine_stats()
    points = args.steps * len(freqs)
    sched = f"batches of {sb} on {lanes} lanes" if batched else f"{lanes} lanes"
    # the scheduling the committed traffic measurement must match: workload,
    # chunking and the ACA schedule in use (tuned, else the built-in one)
    acs = S.aca_schedule(prec) or {"small_bound": 384, "huge_bound": 2048, "queue": 0}
    huge = acs["huge_bound"] if acs["huge_bound"] < (1 << 40) else "none"
    traffic_sched = None
    if (args.level, args.leaf, args.eta, args.aca_tol, args.plane, args.batch) == (5, 64, 2.0, 1e-5, None, 8):
        traffic_sched = (f"batch{sb}" if batched else f"lanes{lanes}") + \
            f" s{acs['small_bound']} h{huge} q{acs['queue']}"
    print(f"# {args.precision}: {1e3 * elapsed / points:.2f} ms/point, {sched}", file=sys.stderr, flush=True)
    st = stats[ic]
    points0 = args.steps * len(freqs)   # points whose launches the profile window records (every lane)
    out = {
        "metric": "freq-points solved/s at 20k tris (FP32+refine vs FP64); rel-L2 err vs FP64 oracle",
        "value": round(points / elapsed, 4),
        "unit": "freq-points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64+f64" if prec == bem.FP32 else "c128",
        "data": "synthetic",
        "config": {"workload": f"pulsating sphere, icosphere level {args.level} ({n} triangles), r=0.5 m, "
                               f"{len(freqs)} frequency points {freqs[0]:.0f}-{freqs[-1]:.0f} Hz per step "
                               f"(incl. {args.freq:g} Hz), Burton-Miller, leaf {args.leaf}, eta {args.eta:g}, "
                               f"ACA {args.aca_tol:g}, GMRES {args.gmres_tol:g}, {args.precision}, {sched}"
                               + ("" if args.plane is None else f", rigid plane {args.plane:g} m below the centre"),
                   "triangles": n, "points_per_step": len(freqs), "lanes": lanes, "sweep_batch": sb, "setup_s": round(setup_s, 3),
                   "refine": refine, "refine_stats_last_point_lane0": rstats,
                   "aca_schedule": S.aca_schedule(prec),
                   "ms_per_point": round(1e3 * elapsed / points, 3),
                   "gmres_iters": int(its[ic, 0]), "residual_estimate": st["residual"],
                   "max_rank_phi": st["max_rank_phi"], "max_rank_dphi": st["max_rank_dphi"],
                   "evaluated_entries": [st["evaluated_phi"], st["evaluated_dphi"]],
                   "phase_ms": {k: round(st[k], 2) for k in ("t_phi_asm", "t_rhs_mvm", "t_dphi_asm", "t_gmres")},
                   "kernel_ms_all_lanes": {f: [round(float(a), 4), int(c)] for f, a, c in zip(FAMILIES, avg_ms, cnt)}},
        "roofline": roofline(avg_ms, cnt, st, blocks, n, args.precision, points0, traffic_sched,
                             quads=1 if args.plane is None else 2, points_per_s=points / elapsed),
    }
    out["roofline"]["duration_source"] = ("HIP events around every launch of the family on all lanes' launching "
                                          "streams during the timed steps (a launch's duration includes time it "
                                          "shares the chip with the other lanes' kernels)")
    p_head = pr[ic, 0]
    p64 = None
    pr0 = None
    if prec == bem.FP32 and not args.no_batch_leg:
        # the other scheduling of the same FP32 points: lanes, or batched
        # chunks of --leg-batch on as many lanes as there are chunks
        ob = 0 if batched else args.leg_batch
        ol = args.lanes if batched else min(args.lanes, -(-args.batch // args.leg_batch))
        prb, _, stb, elb, _, _ = timed_run(bem, S, ctx_h, args, prec, freqs, ol, ob)
        key = "lanes_leg" if batched else "sweep_batch_leg"
        print(f"# {key}: {1e3 * elb / points:.2f} ms/point", file=sys.stderr, flush=True)
        out[key] = {"value": round(points / elb, 4), "ms_per_point": round(1e3 * elb / points, 3),
                    "lanes": ol, "sweep_batch": ob,
                    "phase_ms": {k: round(stb[ic][k], 2) for k in ("t_phi_asm", "t_rhs_mvm", "t_dphi_asm", "t_gmres")},
                    "max_rel_diff_vs_headline": float(max(np.linalg.norm(prb[i, 0] - pr[i, 0]) /
                                                          np.linalg.norm(pr[i, 0]) for i in range(len(freqs))))}
    if refine and not args.no_norefine_leg:
        S.set_refinement(0)
        pr0, _, _, el0, _, _ = timed_run(bem, S, ctx_h, args, prec, freqs, lanes, sb)
        S.set_refinement(refine)
        print(f"# fp32 without refinement: {1e3 * el0 / points:.2f} ms/point", file=sys.stderr, flush=True)
        out["fp32_norefine"] = {"value": round(points / el0, 4), "ms_per_point": round(1e3 * el0 / points, 3)}
    if not args.no_fp64_leg and prec != bem.FP64:
        pr64, _, _, el64, _, _ = timed_run(bem, S, ctx_h, args, bem.FP64, freqs, 1)
        print(f"# fp64: {1e3 * el64 / points:.2f} ms/point", file=sys.stderr, flush=True)
        p64 = pr64[ic, 0]
        out["fp64_value"] = round(points / el64, 4)
        out["fp64_ms_per_point"] = round(1e3 * el64 / points, 3)
        out["speedup_vs_fp64"] = round(el64 / elapsed, 3)
        out["rel_l2_vs_fp64_gpu"] = float(np.linalg.norm(p_head - p64) / np.linalg.norm(p64))
        if pr0 is not None:
            out["fp32_norefine"]["rel_l2_vs_fp64_gpu"] = float(np.linalg.norm(pr0[ic, 0] - p64) / np.linalg.norm(p64))
    if not args.no_dense:
        t0 = time.perf_counter()
        _, _, pd, std = S.solve_dense(args.freq, c=343.0, rho=1.21, coupling=1, gmres_tol=1e-12)
        t_dense = time.perf_counter() - t0
        print(f"# dense FP64 solve: {t_dense:.2f} s, {std['gmres_iters']:.0f} iterations", file=sys.stderr, flush=True)
        rd = lambda p: float(np.linalg.norm(p - pd) / np.linalg.norm(pd))
        out["rel_l2_vs_dense_fp64"] = rd(p_head)
        out["dense_fp64"] = {"seconds": round(t_dense, 3), "gmres_iters": int(std["gmres_iters"]),
                             "gmres_residual": std["residual"]}
        if p64 is not None:
            out["fp64_rel_l2_vs_dense_fp64"] = rd(p64)
        if pr0 is not None:
            out["fp32_norefine"]["rel_l2_vs_dense_fp64"] = rd(pr0[ic, 0])
    if not args.no_cpu_baseline:
        pr_cpu, t_cpu, threads, st_cpu = run_cpu_baseline(tri, bc, args)
        out["cpu_baseline"] = {"value": round(1.0 / t_cpu, 5), "unit": "freq-points/s", "cores": threads,
                               "kind": "port", "cpu_model": cpu_model(),
                               "sample": f"one full frequency point ({args.freq:g} Hz) of the same workload ({n} "
                                         f"triangles, FP64, set-up included) with the CPU restatement: {t_cpu:.1f} s",
                               "phase_ms": {k: round(st_cpu[k], 1) for k in ("t_phi_asm", "t_rhs_mvm", "t_dphi_asm",
                                                                            "t_gmres")}}
        out["rel_l2_vs_cpu_fp64"] = float(np.linalg.norm(p_head - pr_cpu) / np.linalg.norm(pr_cpu))
        if p64 is not None:
            out["fp64_rel_l2_vs_cpu_fp64"] = float(np.linalg.norm(p64 - pr_cpu) / np.linalg.norm(pr_cpu))
        if not args.no_cpu_1thread:
            _, t1, _, st1 = run_cpu_baseline(tri, bc, args, threads=1)
            out["cpu_baseline"]["one_thread"] = {"value": round(1.0 / t1, 5), "cores": 1, "seconds": round(t1, 1),
                                                 "phase_ms": {k: round(st1[k], 1) for k in
                                                              ("t_phi_asm", "t_rhs_mvm", "t_dphi_asm", "t_gmres")}}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
