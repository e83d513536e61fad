"""Benchmark: 3D elastic RTM on one MI355X (Gcell-updates/s).

Workload (the 3D single-shot configuration): 256x256x128 model, d = 10 m,
20-cell C-PML on every face (296x296x168 padded grid; --boundary damp selects
the reference's damping instead), 8th-order
staggered FD (radius 4), dt = 0.8 ms, nt = 2500, 9 wavefield components plus
the vector-P split (tp, vxp, vyp, vzp), FP32.  Model: five dipping layers
vp 1800-4200 m/s, vs = vp/1.732, density 230*vp^0.25, three random Gaussian
anomalies (seed 2); migration model = the same smoothed by a 7-point box
filter.  15 Hz Ricker at the centre of the surface, 128x128 receivers.

One benchmark step = one migrated shot: the source pass (nt steps with
excitation tracking) + the receiver pass (nt steps with record injection and
PP/PS/tp excitation-amplitude imaging).  The receiver record is modelled once
before timing and stays in HBM (inputs resident).  value = updated cells x
time steps / wall time.

Also reported: the roofline of the dominant kernel (algorithmic bytes per
launch / average launch duration from HIP events around every launch on the
context's stream, against 8 TB/s), HBM traffic per launch from the committed
rocprofv3 PMC summary (profiles/), and an FP32 CPU baseline (the oracle's
FP32 instantiation, OpenMP) on a bounded sample of the same grid.
"""
import argparse
import ctypes
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
PKG = "2d_and_3d_elastic_reverse_time_migration_amd"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def smooth_box(a, passes=4):
    a = a.astype(np.float32)
    for _ in range(passes):
        q = np.pad(a, 1, mode="edge")
        a = (q[1:-1, 1:-1, 1:-1] * 2 + q[2:, 1:-1, 1:-1] + q[:-2, 1:-1, 1:-1] + q[1:-1, 2:, 1:-1]
             + q[1:-1, :-2, 1:-1] + q[1:-1, 1:-1, 2:] + q[1:-1, 1:-1, :-2]) / 8.0
    return np.ascontiguousarray(a, dtype=np.float32)


def make_params(erm, nt, boundary):
    p = erm.default_params()
    p.update(nx=256, ny=256, nz=128, dx=10.0, dy=10.0, dz=10.0, dt=0.8, lt=nt, lt_forward=nt, freq=15.0,
             fd_order=8, boundary=boundary, coe_attenuation=220.0,
             bu=20, bd=20, bl=20, br=20, bf=20, bb=20,
             shot_start_x=128, shot_start_y=128, shot_start_z=0,
             receiver_num_x=128, receiver_num_y=128, receiver_num_z=1,
             receiver_start_x=0, receiver_interval_x=2, receiver_start_y=0, receiver_interval_y=2,
             receiver_start_z=0, add_receiver_bool=1, RTM=1, GPU_N=1)
    return p


def host_cpu():
    """(cores this process may use, nproc, CPU model name) of the host."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return aff, os.cpu_count() or 1, model


def cpu_baseline(p, vp_s, vs_s, rho_s, budget_s):
    """FP32 oracle (OpenMP over z) on the host: source-pass steps over the full
    grid, median of 3 runs.  Threads: OMP_NUM_THREADS when set (the GPU box
    grants one GPU a share of 16 host cores and sets it so), else every core
    this process may run on."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle3d
    aff, nproc, model = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS") or aff)
    oracle3d.set_threads(threads)
    o = oracle3d.Oracle3D(p, 32)
    o.load_model(1, vp_s, vs_s, rho_s)
    o.step(1, 0, 1)  # warm the pages
    R = 4
    cells = (p.nx + p.bl + p.br - 2 * R) * (p.ny + p.bf + p.bb - 2 * R) * (p.nz + p.bu + p.bd - 2 * R)
    rates, it, steps_run = [], 1, 0
    for run in range(3):
        steps, t0 = 0, time.perf_counter()
        while True:
            o.step(1, it, 2)
            it += 2
            steps += 2
            el = time.perf_counter() - t0
            if el > budget_s / 3 or steps >= 100:
                break
        rates.append(cells * steps / el / 1e9)
        steps_run += steps
    o.close()
    med = float(np.median(rates))
    shot_s = cells * 2 * p.lt / (med * 1e9)
    return {"value": round(med, 4), "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
            "nproc": nproc, "affinity_cores": aff, "cpu_model": model,
            "runs_gcells": [round(r, 4) for r in rates],
            "projected_s_per_shot": round(shot_s, 1),
            "sample": f"median of 3 runs ({steps_run} source-pass time steps in all: stress+velocity+vP+"
                      f"excitation, C-PML) of the same 296x296x168 grid, FP32 oracle, OpenMP over z, "
                      f"{threads} threads; steady-state rate, not a full shot"}


def batch_windows(n=24, nx=600, ny=600, seed=33, lo=192, hi=320):
    """Random surface shots, each with a lateral window of uneven size centred on it."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        wx, wy = (int(v) for v in rng.integers(lo, hi + 1, size=2))
        sx, sy = (int(v) for v in rng.integers(0, nx, size=2))
        x0 = min(max(sx - wx // 2, 0), nx - wx)
        y0 = min(max(sy - wy // 2, 0), ny - wy)
        out.append((sx, sy, 0, x0, y0, wx, wy))
    return out


def batch_setup(erm, nt, boundary, lo=192, hi=320):
    """The batched workload's parameters, true and migration models and shot
    windows: 600x600x250 model (seed 3), 24 windows of lo..hi cells (192..320)."""
    p = erm.default_params()
    p.update(nx=600, ny=600, nz=250, dx=10.0, dy=10.0, dz=10.0, dt=0.8, lt=nt, lt_forward=nt, freq=15.0,
             fd_order=8, boundary=boundary, coe_attenuation=220.0, bu=20, bd=20, bl=20, br=20, bf=20, bb=20,
             receiver_interval_x=2, receiver_interval_y=2, receiver_start_z=0, add_receiver_bool=1, RTM=1, GPU_N=1)
    vp, vs, rho = erm.make_model3d(1, p.nx, p.ny, p.nz, p.dx, p.dy, p.dz, 3)
    vp_s = smooth_box(vp, 2)
    vs_s = (vp_s / 1.732).astype(np.float32)
    rho_s = (230.0 * vp_s.astype(np.float64) ** 0.25).astype(np.float32)
    return p, (vp, vs, rho), (vp_s, vs_s, rho_s), batch_windows(lo=lo, hi=hi)


def run_batch(erm, args):
    """Batched workload: 24 shots over a 600x600x250 model (seed 3), windows
    192..320 cells laterally, full depth, 20-cell C-PML, 8th order, dt 0.8 ms,
    15 Hz, receivers every 2 cells over each window.  One step = one migration
    of all 24 shots (source pass + receiver pass, nt steps each); records are
    modelled once (batched) before timing.  The launch shape is autotuned."""
    nt = args.nt
    lo, hi = args.batch_window
    p, (vp, vs, rho), (vp_s, vs_s, rho_s), wins = batch_setup(
        erm, nt, erm.BOUNDARY_CPML if args.boundary == "cpml" else erm.BOUNDARY_DAMP, lo, hi)
    log(f"bench batch: {len(wins)} shots, nt={nt}, creating window contexts")
    b = erm.Batch3D(p, 0)
    b.load_model(0, vp, vs, rho)
    b.load_model(1, vp_s, vs_s, rho_s)
    del vp, vs, rho
    b.set_shots(wins)
    b.set_progress(lambda ps, ms: log(f"bench batch: {('model', 'source', 'receiver')[ps]} pass of all shots done "
                                      f"({ms / 1e3:.1f} s)"))
    tune = b.autotune(steps=3)
    log(f"bench batch: autotuned {tune}")
    t0 = time.perf_counter()
    b.run(want_images=False)  # models the records (batched) and runs one migration
    t = b.timing()
    log(f"bench batch: records + first migration in {time.perf_counter() - t0:.1f} s ({t})")
    for w in range(max(0, args.warmup - 1)):
        b.run(want_images=False)
    el, passes_ms = 0.0, 0.0
    t0 = time.perf_counter()
    for s in range(args.steps):
        b.run(want_images=False)
        t = b.timing()
        passes_ms += t["ms_source"] + t["ms_receiver"]
        log(f"bench batch: timed migration {s + 1}/{args.steps}: {t['ms_source']:.0f} + {t['ms_receiver']:.0f} ms")
    el = time.perf_counter() - t0
    cells = t["cells_per_step"]
    updates = cells * (t["steps_source"] + t["steps_receiver"]) * args.steps
    value = updates / el / 1e9
    batched_pass_value = updates / (passes_ms * 1e-3) / 1e9
    # one shot per launch on the same windows: each window's full source and
    # receiver passes through the one-shot path (its own launches, same launch
    # shape), timed by the same pass events as the batched passes
    seq_ms, seq_updates = 0.0, 0.0
    for s in range(len(wins)):
        v = b.shot(s)
        v.set_profiling(False)
        v.zero_excitation()
        v.forward_smoothed()
        seq_ms += v.timing()["pass_ms"]
        v.backward_smoothed()
        seq_ms += v.timing()["pass_ms"]
        nnz, nny, nnx = v.shape
        seq_updates += (nnx - 8) * (nny - 8) * (nnz - 8) * 2.0 * nt
    seq_value = seq_updates / (seq_ms * 1e-3) / 1e9
    log(f"bench batch: pass-event rates: batched {batched_pass_value:.3f}, one shot per launch {seq_value:.3f} "
        f"Gcell-updates/s")
    # CPU baseline: the FP32 oracle on the window closest to 256x256 (a representative size)
    cpu = None
    if not args.no_cpu_baseline:
        k = int(np.argmin([abs(w[5] - 256) + abs(w[6] - 256) for w in wins]))
        sx, sy, sz, x0, y0, wx, wy = wins[k]
        q = erm.default_params()
        ctypes.memmove(ctypes.addressof(q), ctypes.addressof(p), ctypes.sizeof(q))
        q.nx, q.ny = wx, wy
        q.shot_start_x, q.shot_start_y, q.shot_start_z = sx - x0, sy - y0, sz
        q.receiver_num_x = q.receiver_num_y = 0
        cut = [np.ascontiguousarray(a[x0:x0 + wx, y0:y0 + wy, :]) for a in (vp_s, vs_s, rho_s)]
        log(f"bench batch: CPU baseline (FP32 oracle) on window {k}, {wx}x{wy}x{p.nz}")
        cpu = cpu_baseline(q, *cut, args.cpu_budget)
        cpu["sample"] = f"window {k} ({wx}x{wy}x{p.nz} + 20-cell C-PML): " + cpu["sample"].replace(
            "of the same 296x296x168 grid", "of that window's grid")
    # per-step roofline of the batched step (both kernels): algorithmic bytes / autotuned step time
    step_bytes = cells * (76.0 + 84.0)
    roof = {"bound": "hbm", "achieved": round(step_bytes / (tune["ms_per_step"] * 1e-3) / 1e9, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(step_bytes / (tune["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "k_stress_batch + k_velocity_batch (one source-pass step, excluding C-PML memory traffic)",
            "algorithmic_bytes_per_launch": step_bytes, "avg_launch_ms": round(tune["ms_per_step"], 4)}
    b.close()
    line = {
        "metric": "Gcell-updates/s 3D elastic RTM (batched shots)",
        "value": round(value, 3), "unit": "Gcell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded dipping-layer + anomaly model, random shot layout)",
        "config": {"workload": "3D elastic RTM, 24 shots, 600x600x250 model, windows %d-%d cells laterally "
                               "(full depth) + 20-cell %s, 8th-order FD, dt=0.8 ms, nt=%d per pass, 15 Hz; "
                               "step = source + receiver pass of all shots in one batched graph-captured launch "
                               "per kernel per step" % (lo, hi, "C-PML" if args.boundary == "cpml" else "damping", nt),
                   "cells_per_time_step": cells, "autotuned": tune, "parallelism": "single GPU"},
        "batched_pass_Gcells": round(batched_pass_value, 3),
        "one_shot_per_launch_Gcells": round(seq_value, 3),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def visco_model(nx, nz, seed):
    """Layered viscoelastic model with smooth random Q (Qp 30..200, Qs = 0.6 Qp):
    background (smoothed) model and the relative perturbation dm (dvp/vp, dvs/vs)."""
    learned = importlib.import_module(PKG + ".learned")
    bvp, bvs, brho, dm = learned.random_model(nx, nz, seed, smooth_sigma=8.0)
    rng = np.random.default_rng(seed + 1)
    q = learned._smooth(rng.standard_normal((nx, nz)), 25.0)
    q = (q - q.min()) / max(q.max() - q.min(), 1e-12)
    qp = (30.0 + 170.0 * q).astype(np.float32)
    qs = (0.6 * qp).astype(np.float32)
    return bvp, bvs, brho, qp, qs, dm


LEARNED_QUICK = dict(train_models=48, train_shape=(64, 48), train_shots=4, train_lt=700, test_shape=(128, 48),
                     test_shots=8, test_lt=700, test_iters=600, tol=1e-2, epochs=30, batch=16, warm=0)


def learned_case(erm, quick, overrides=None):
    """LSRTM seconds to data misfit 1e-2 on a held-out model: plain CG, CG with
    the illumination diagonal, and the learned preconditioner (flexible CG
    alternating z = net(w g) and z = w g), from the whole learned workflow of
    learned.py (training pairs from CG on random models, U-Net training on the
    GPU with its own BF16 MFMA kernels, held-out comparison).  Full: 160
    training models 128x96 (6 shots), held-out 256x96 with 12 shots (about 12
    minutes).  quick: 48 models 64x48 (4 shots), held-out 128x48 with 8 shots
    (about 2 minutes)."""
    learned = importlib.import_module(PKG + ".learned")
    kw = dict(LEARNED_QUICK) if quick else dict(warm=3)
    kw.update(overrides or {})
    t0 = time.perf_counter()
    res = learned.run(log=log, **kw)
    if kw.get("plain_only"):
        cfg, pl = res["config"], res["plain"]
        return {"tol": cfg["tol"], "plain_only": True, "plain": {"iters": pl["iters_to_tol"], "s": pl["s_to_tol"]},
                "model_err_vs_true": res["model_err_vs_true"], "plain_file": kw.get("plain_file"),
                "scale": "held-out %dx%d with %d shots, nt %d" % (*cfg["test_shape"], cfg["test_shots"], cfg["test_lt"]),
                "learned_vs_plain_time": None}
    if kw.get("test_iters", 1) <= 0:  # training only (parameters saved)
        return {"tol": res["config"]["tol"], "trained_only": True, "params_out": kw.get("params_out"),
                "train_s": res.get("train_s"), "datagen_s": res.get("datagen_s"), "scale": "training only",
                "learned_vs_plain_time": None}
    el = time.perf_counter() - t0
    cfg = res["config"]
    p, d, l = res["plain"], res["illumination"], res["schedule_alt2"]
    ratio = lambda a, b: round(a / b, 4) if (a and b) else None
    out = {
        "tol": cfg["tol"],
        "scale": "%d training models %dx%d (%d shots), held-out %dx%d with %d shots, nt %d%s"
                 % (cfg["train_models"], *cfg["train_shape"], cfg["train_shots"], *cfg["test_shape"],
                    cfg["test_shots"], cfg["test_lt"],
                    " (reduced case)" if quick else " (scaled case; the configured one is 200 models 400x200 and "
                                                   "a held-out 800x300 with 80 shots)")
                 + (" (net loaded from %s)" % kw["params_in"] if kw.get("params_in") else ""),
        "plain": {"iters": p["iters_to_tol"], "s": p["s_to_tol"]},
        "illumination_diagonal": {"iters": d["iters_to_tol"], "s": d["s_to_tol"]},
        "learned": {"iters": l["iters_to_tol"], "s": l["s_to_tol"], "schedule": "net and diagonal alternating"},
        "learned_vs_plain_time": ratio(l["s_to_tol"], p["s_to_tol"]),
        "learned_vs_diagonal_time": ratio(l["s_to_tol"], d["s_to_tol"]),
        "model_err_vs_true": res["model_err_vs_true"],
        "learned_model_err_vs_plain": l["model_err_vs_plain"],
        "unet_train_mfma_tflops": res.get("train_mfma_tflops"),
        "workflow_s": round(el, 1),
    }
    if "warm_start" in res:
        w = res["warm_start"]
        out["learned_warm_start"] = {"learned_iters": w["warm_iters"], "iters": w["iters_to_tol"], "s": w["s_to_tol"]}
    return out


def run_learned(erm, args):
    """LSRTM seconds to data misfit 1e-2, learned preconditioner vs plain CG
    (learned_case); one step = one whole workflow; value = learned seconds /
    plain seconds."""
    t0 = time.perf_counter()
    ov = {}
    if args.learned_test_shape:
        ov["test_shape"] = tuple(args.learned_test_shape)
    if args.learned_train_shape:
        ov["train_shape"] = tuple(args.learned_train_shape)
    for k, v in (("train_models", args.learned_train_models), ("train_lt", args.learned_train_lt),
                 ("train_shots", args.learned_train_shots), ("plain_file", args.learned_plain_file)):
        if v is not None:
            ov[k] = v
    if args.learned_plain_only:
        ov["plain_only"] = True
    for k, v in (("test_shots", args.learned_test_shots), ("test_lt", args.learned_test_lt),
                 ("test_iters", args.learned_test_iters), ("params_in", args.learned_params_in),
                 ("params_out", args.learned_params_out)):
        if v is not None:
            ov[k] = v
    if args.learned_params_in:
        ov["warm"] = 0
    out = learned_case(erm, args.quick, ov)
    if out.get("trained_only") or out.get("plain_only"):
        print(json.dumps({"learned_training" if out.get("trained_only") else "learned_plain_cg": out}), flush=True)
        return
    el = time.perf_counter() - t0
    line = {
        "metric": "LSRTM seconds to data misfit %g, learned vs plain CG" % out["tol"],
        "value": out["learned_vs_plain_time"],
        "unit": "x plain CG time",
        "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(el * 1e3, 1),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 operators, bf16 U-Net (fp32 accumulate)",
        "data": "synthetic (seeded random layered-plus-anomaly models; Born data)",
        "config": {"workload": "learned-preconditioner LSRTM (2D elastic, dvp/vp and dvs/vs), " + out["scale"]},
        **{k: v for k, v in out.items() if k not in ("tol", "scale")},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def rtm2d_case(erm):
    """The 2D elastic RTM shot: 400x200 model, d = 10 m, three flat layers (vp
    2000 / 2800 / 3500 m/s at depth thirds) plus a Gaussian vp anomaly (+300 m/s,
    sigma 200 m, at x = 2000 m, z = 1000 m), vs = vp/1.732, density
    230 vp^0.25; 8th-order FD, dt = 1 ms, nt = 2000, 40-cell C-PML, 20 Hz
    Ricker at (2000 m, 20 m), 400 receivers at the surface; migration model =
    the true model Gaussian-smoothed (sigma 6 cells)."""
    learned = importlib.import_module(PKG + ".learned")
    nx, nz = 400, 200
    p = learned.survey_params(nx, nz, 2000, freq=20.0, d=10.0, dt=1.0, b=40)
    p.boundary, p.cpml_r = 1, 1e-3
    p.shot_x, p.shot_z = 200, 2
    p.nrec, p.rec_x0, p.rec_dx, p.rec_z = 400, 0, 1, 0
    x = np.arange(nx)[:, None] * 10.0
    z = np.arange(nz)[None, :] * 10.0
    vp = np.where(z < 670.0, 2000.0, np.where(z < 1330.0, 2800.0, 3500.0)) * np.ones((nx, 1))
    vp = vp + 300.0 * np.exp(-((x - 2000.0) ** 2 + (z - 1000.0) ** 2) / (2 * 200.0 ** 2))
    true = [vp.astype(np.float32), (vp / 1.732).astype(np.float32), (230.0 * vp ** 0.25).astype(np.float32)]
    vps = learned._smooth(vp, 6.0)
    mig = [vps.astype(np.float32), (vps / 1.732).astype(np.float32), (230.0 * vps ** 0.25).astype(np.float32)]
    return p, true, mig


def rtm2d_errors(p, ref, et_ref, ea_ref, imgs, et):
    """rel L2 of tp / PP / PS images against FP64 over cells with ex_amp > 1e-3 max
    whose excitation step agrees (model region), and the mismatch fraction."""
    cut = lambda a: a[p.bl:p.bl + p.nx, p.bu:p.bu + p.nz]
    exc = cut(ea_ref > 1e-3 * ea_ref.max())
    same = exc & (cut(et) == cut(et_ref))
    out = {"ex_time_mismatch_frac": round(float(1 - same.sum() / max(exc.sum(), 1)), 6)}
    for name, a, b in zip(("tp", "pp", "ps"), imgs, ref):
        out["img_%s_rel_l2" % name] = float(np.linalg.norm((a - b)[same]) / max(np.linalg.norm(b[same]), 1e-300))
    return out


def run_rtm2d(erm, args):
    """2D elastic RTM of one shot (the CPU-runnable configuration): GPU shot
    time (model pass in the true model + source pass + receiver pass with
    excitation imaging), the FP32 oracle's time for the same shot on the
    host cores (median of 3), and both images' error against the FP64 oracle."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle2d
    p, true, mig = rtm2d_case(erm)
    nt = p.lt
    nnx, nnz = p.padded_shape
    cells = (nnx - 8) * (nnz - 8)
    g = erm.Elastic2D(p)
    g.load_model(0, *true)
    g.load_model(1, *mig)
    for w in range(max(args.warmup, 1)):
        rec = g.forward()
        g.rtm(rec)
    times = []
    for s in range(args.steps):
        t0 = time.perf_counter()
        rec = g.forward()
        imgs = g.rtm(rec)
        times.append(time.perf_counter() - t0)
        log(f"bench rtm2d: GPU shot {s + 1}/{args.steps}: {times[-1] * 1e3:.1f} ms")
    et_g = g.get_field("ex_time")
    g.close()
    gpu_s = float(np.median(times))

    def oracle_shot(prec):
        o = oracle2d.Oracle2D(p, prec)
        o.load_model(0, *true)
        o.load_model(1, *mig)
        t0 = time.perf_counter()
        o.step(0, 0, nt)
        o.zero(7)
        o.step(1, 0, nt)
        o.zero(1)
        o.step(2, nt - 1, nt)
        el = time.perf_counter() - t0
        return o, el

    aff, nproc, model = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS") or aff)
    oracle2d.oracle3d.set_threads(threads)
    o64, _ = oracle_shot(64)
    ref = o64.images()
    et_ref, ea_ref = o64.get_field("ex_time"), o64.get_field("ex_amp")
    cpu_times, o32 = [], None
    for r in range(3):
        o32, el = oracle_shot(32)
        cpu_times.append(el)
        log(f"bench rtm2d: FP32 oracle shot {r + 1}/3: {el:.2f} s on {threads} threads")
    cpu_s = float(np.median(cpu_times))
    err_gpu = rtm2d_errors(p, ref, et_ref, ea_ref, imgs, et_g)
    err_cpu = rtm2d_errors(p, ref, et_ref, ea_ref, o32.images(), o32.get_field("ex_time"))
    updates = cells * 3.0 * nt
    line = {
        "metric": "2D elastic RTM, one shot (s per shot)",
        "value": round(gpu_s, 4), "unit": "s/shot", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(gpu_s * 1e3, 2), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (three flat layers + Gaussian anomaly)",
        "config": {"workload": "2D elastic RTM, one shot: 400x200 model + 40-cell C-PML (480x280), d=10 m, 8th-order "
                               "staggered FD, dt=1 ms, nt=2000, 20 Hz Ricker at (2000 m, 20 m), 400 surface "
                               "receivers; step = model pass + source pass + receiver pass with excitation imaging",
                   "cells_per_time_step": cells, "time_steps_per_shot": 3 * nt,
                   "gpu_Gcell_updates_per_s": round(updates / gpu_s / 1e9, 3), "parallelism": "single GPU"},
        "error_vs_fp64": {"gpu": err_gpu, "cpu_fp32": err_cpu},
        "roofline": None,
        "cpu_baseline": {"value": round(cpu_s, 3), "unit": "s/shot", "cores": threads, "kind": "port",
                         "nproc": nproc, "affinity_cores": aff, "cpu_model": model,
                         "runs_s": [round(t, 3) for t in cpu_times],
                         "Gcell_updates_per_s": round(updates / cpu_s / 1e9, 4),
                         "sample": "the full shot (3 x 2000 steps), FP32 oracle, OpenMP over x, median of 3"},
    }
    print(json.dumps(line), flush=True)


def run_rtm2d_xcorr(erm, args):
    """2D elastic RTM of the same shot with cross-correlation imaging, one
    line per source-wavefield strategy (erm2d_rtm_xcorr: stored history,
    reversed propagation, boundary strips + reversal): GPU time per shot
    (source pass + receiver pass with imaging; the model pass runs once
    beforehand), the HBM the strategy holds, and each image's error against
    the FP64 oracle's history images."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle2d
    p, true, mig = rtm2d_case(erm)
    nt = p.lt
    g = erm.Elastic2D(p)
    g.load_model(0, *true)
    g.load_model(1, *mig)
    rec = g.forward()
    o = oracle2d.Oracle2D(p, 64)
    o.load_model(1, *mig)
    t0 = time.perf_counter()
    ref = o.rtm_xcorr(rec, 0)
    log(f"bench rtm2d xcorr: FP64 oracle history images in {time.perf_counter() - t0:.1f} s")
    names = ("pp", "ps", "ppd", "psc", "psc_polarity")
    mn = p.nx * p.nz
    ring = sum(1 for i in range(max(p.bl - 4, 0), min(p.bl + p.nx + 4, p.padded_shape[0]))
               for k in range(max(p.bu - 4, 0), min(p.bu + p.nz + 4, p.padded_shape[1]))
               if not (p.bl + 1 <= i < p.bl + p.nx - 1 and p.bu + 1 <= k < p.bu + p.nz - 1))
    hbm = {0: 4.0 * nt * 3 * mn, 1: 0.0, 2: 4.0 * nt * 8 * ring}
    res = {}
    for strat, label in ((0, "history"), (1, "reverse"), (2, "strips")):
        g.rtm_xcorr(rec, strat)  # warm-up (allocations)
        times = []
        for _ in range(max(args.steps, 1)):
            t0 = time.perf_counter()
            img = g.rtm_xcorr(rec, strat)
            times.append(time.perf_counter() - t0)
        err = {n: round(float(np.linalg.norm(img[k] - ref[k]) / np.linalg.norm(ref[k])), 7) for k, n in enumerate(names)}
        res[label] = {"s_per_shot": round(float(np.median(times)), 4), "hbm_GB": round(hbm[strat] / 1e9, 3),
                      "rel_l2_vs_fp64_history": err}
        log(f"bench rtm2d xcorr: {label}: {res[label]}")
    g.close()
    line = {
        "metric": "2D elastic RTM with cross-correlation imaging, one shot (s per shot), by source-wavefield strategy",
        "value": res["history"]["s_per_shot"], "unit": "s/shot", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": round(res["history"]["s_per_shot"] * 1e3, 2), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (three flat layers + Gaussian anomaly)",
        "config": {"workload": "2D elastic RTM, one shot: 400x200 model + 40-cell C-PML (480x280), d=10 m, 8th-order "
                               "staggered FD, dt=1 ms, nt=2000, 20 Hz Ricker, 400 surface receivers; step = source "
                               "pass + receiver pass with cross-correlation imaging (PP, PS, div-div, div-curl, "
                               "div-curl with the PS polarity correction)",
                   "strategies": res, "parallelism": "single GPU"},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def ring_cells(p, R):
    """Cells of the strips ring of a 2D context (the library's xcorr_ring):
    outside the model region shrunk by one cell, within R + 1 of it."""
    nnx, nnz = p.nx + p.bl + p.br, p.nz + p.bu + p.bd
    i0, i1, k0, k1 = p.bl + 1, p.bl + p.nx - 1, p.bu + 1, p.bu + p.nz - 1
    ia, ib = max(i0 - R - 1, 0), min(i1 + R + 1, nnx)
    ka, kb = max(k0 - R - 1, 0), min(k1 + R + 1, nnz)
    return (ib - ia) * (kb - ka) - (i1 - i0) * (k1 - k0)


def run_gradient(erm, args):
    """LSRTM gradient (the exact-transpose adjoint) with the background from
    the stored E/F history against the background stepped back from boundary
    strips (erm2d_params.adjoint_source), on the learned workload's held-out
    problem: 800x300 model, d = 10 m, 12 shots in one batched launch per
    kernel, nt = 2000 (1 ms), 15 Hz, Born data.  Times one adjoint
    application (median of 3) and a few CG iterations with each; states the
    HBM each holds for the background and the gradient difference."""
    learned = importlib.import_module(PKG + ".learned")
    nx, nz, nshots, lt, iters = 800, 300, 12, 2000, 10
    bvp, bvs, brho, dm = learned.random_model(nx, nz, 5000)
    shots = learned.shot_positions(nx, nshots)
    out, grads, d = {}, {}, None
    for name, src in (("history", erm.SRC_HISTORY), ("strips", erm.SRC_STRIPS)):
        p = learned.survey_params(nx, nz, lt)
        p.adjoint_source = src
        g = erm.Elastic2D(p)
        g.load_model(1, bvp, bvs, brho)
        g.set_batch(nshots)
        if d is None:
            d = learned.born_data(g, shots, dm)
        g.adjoint_shots(shots, d)  # warm-up
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            grads[name] = g.adjoint_shots(shots, d)
            ts.append(time.perf_counter() - t0)
        R = 4
        held = (nshots * lt * 5 * nx * nz * 4.0 if name == "history"
                else nshots * lt * 8 * ring_cells(p, R) * 4.0)
        m, mis, sec, _ = g.lsrtm_cg(shots, d, iters, pmask=3)
        g.close()
        out[name] = {"adjoint_s": round(float(np.median(ts)), 4), "background_hbm_gb": round(held / 1e9, 3),
                     "cg_s_per_iteration": round(float(sec[-1] / iters), 4),
                     "misfit": [round(float(v), 6) for v in mis]}
        log(f"bench gradient: {name}: adjoint {out[name]['adjoint_s']} s, background {out[name]['background_hbm_gb']} GB, "
            f"CG misfit after {iters}: {mis[-1]:.4f}")
    diff = [float(np.linalg.norm(grads["strips"][c] - grads["history"][c]) / np.linalg.norm(grads["history"][c]))
            for c in range(3)]
    line = {
        "metric": "LSRTM gradient (exact-transpose adjoint) s per application, background from boundary strips",
        "value": out["strips"]["adjoint_s"], "unit": "s", "n_gpus": 1, "steps": 3, "warmup": 1,
        "ms_per_step": round(out["strips"]["adjoint_s"] * 1e3, 1), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded layered-plus-anomaly model, Born data)",
        "config": {"workload": f"2D elastic LSRTM adjoint, {nx}x{nz}, d=10 m, {nshots} shots batched, nt={lt}, "
                               f"15 Hz, 20-cell damping, radius 4; {iters} CG iterations per strategy",
                   "parallelism": "single GPU"},
        "history": out["history"], "strips": out["strips"],
        "strips_vs_history_gradient_rel_l2": [round(v, 8) for v in diff],
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def run_visco(erm, args):
    """2D viscoelastic LSRTM with plain CG: 600x300 model, d = 8 m, 3 SLS
    mechanisms, 60 shots, 600 receivers, 20 Hz, nt = 3000 (dt 0.8 ms), 40-cell
    damping boundary, Born data of a seeded layered-plus-anomaly perturbation;
    20 CG iterations (dvp/vp, dvs/vs).  One step = one CG iteration."""
    learned = importlib.import_module(PKG + ".learned")
    nx, nz = 600, 300
    p = learned.survey_params(nx, nz, args.nt if args.nt != 2500 else 3000, freq=20.0, d=8.0, dt=0.8, b=40)
    p.nsls = 3
    p.nrec, p.rec_x0 = 600, 0
    bvp, bvs, brho, qp, qs, dm = visco_model(nx, nz, 4)
    g = erm.Elastic2D(p)
    g.load_model(1, bvp, bvs, brho, qp, qs)
    if args.batch > 1:
        g.set_batch(args.batch)  # shots per launch (each holds its own E/F history: ~10.8 GB here)
    else:
        g.set_workers(args.workers)
    shots = learned.shot_positions(nx, args.shots)
    t0 = time.perf_counter()
    d = learned.born_data(g, shots, dm)
    log(f"bench visco: Born data of {len(shots)} shots in {time.perf_counter() - t0:.1f} s")

    def mon(it, m, grad, mis):
        log(f"bench visco: CG iteration {it}: misfit {mis:.4f}")

    t0 = time.perf_counter()
    m, mis, sec, it = g.lsrtm_cg(shots, d, args.iters, pmask=3, monitor=mon)
    el = time.perf_counter() - t0
    err = float(np.linalg.norm(m[:2] - dm[:2]) / np.linalg.norm(dm[:2]))
    nnx, nnz = g.shape
    g.close()
    oracle_check = None
    if args.oracle_check:
        oracle_check = visco_oracle_check(erm, learned, p, (bvp, bvs, brho, qp, qs), dm)
    cpu = None
    if not args.no_cpu_baseline:
        log("bench visco: CPU baseline (FP32 oracle, one shot)")
        cpu = visco_cpu_baseline(learned, p, (bvp, bvs, brho, qp, qs), dm, len(shots), args.cpu_budget * 3)
    # operator throughput: per shot and iteration one Born (2 states) and one adjoint (forward + backward)
    updates = float(nnx * nnz) * p.lt * 4 * len(shots) * it
    line = {
        "metric": "2D viscoelastic LSRTM (plain CG): s per CG iteration",
        "value": round(el / max(it, 1), 3), "unit": "s/iteration", "n_gpus": 1, "steps": it, "warmup": 0,
        "ms_per_step": round(el / max(it, 1) * 1e3, 1), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded layered model, smooth random Q, Born data)",
        "config": {"workload": f"2D visco LSRTM, {nx}x{nz}, d=8 m, L=3 SLS, {len(shots)} shots, 600 receivers, "
                               f"20 Hz, nt={p.lt}, dt=0.8 ms, 40-cell damping, {args.iters} CG iterations",
                   "workers": args.workers if args.batch <= 1 else None,
                   "shots_per_launch": args.batch if args.batch > 1 else 1, "parallelism": "single GPU"},
        "misfit": [round(float(v), 5) for v in mis], "seconds": [round(float(v), 2) for v in sec],
        "model_error_vs_true": round(err, 4),
        "operator_Gcell_updates_per_s": round(updates / el / 1e9, 3),
        "oracle_check": oracle_check,
        "roofline": None, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def visco_cpu_baseline(learned, p, bg, dm, nshots, budget_s, nt_sample=200):
    """FP32 oracle on the host (OpenMP, threads as the 3D baseline): one Born
    and one adjoint application for one shot of the same 600x300 L = 3 problem
    over nt_sample steps, median of 3, projected to the full nt and to the
    shots of one CG iteration (one Born + one adjoint per shot)."""
    import ctypes
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle3d
    import oracle2d
    aff, nproc, model = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS") or aff)
    oracle3d.set_threads(threads)
    q = learned.Params2D()
    ctypes.memmove(ctypes.addressof(q), ctypes.addressof(p), ctypes.sizeof(q))
    q.lt = q.lt_forward = min(nt_sample, p.lt)
    o = oracle2d.Oracle2D(q, 32)
    o.load_model(1, *bg)
    r = np.random.default_rng(3).standard_normal((2, q.lt, q.nrec)).astype(np.float32)
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        o.born(dm)
        o.adjoint(r)
        runs.append(time.perf_counter() - t0)
        if sum(runs) > budget_s:
            break
    o.close()
    per_shot = float(np.median(runs)) * p.lt / q.lt
    return {"value": round(per_shot, 2), "unit": "s per shot (Born + adjoint, full nt)", "cores": threads,
            "kind": "port", "nproc": nproc, "affinity_cores": aff, "cpu_model": model,
            "runs_s": [round(v, 3) for v in runs], "projected_s_per_cg_iteration": round(per_shot * nshots, 1),
            "sample": f"median of {len(runs)} runs of one Born + one adjoint application of one shot over "
                      f"{q.lt} of the {p.lt} steps, the same 600x300 L = 3 grid, FP32 oracle, {threads} threads; "
                      f"projected linearly to {p.lt} steps"}


def visco_oracle_check(erm, learned, p, bg, dm, iters=2, nt=500):
    """GPU CG against CGLS on the FP64 oracle's operators for one shot, on the
    full grid with nt shortened: relative model difference per iteration."""
    import ctypes
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle2d
    q = type(p)()
    ctypes.memmove(ctypes.addressof(q), ctypes.addressof(p), ctypes.sizeof(q))
    q.lt = q.lt_forward = nt
    shot = (p.nx // 2, 2)
    g = erm.Elastic2D(q)
    g.load_model(1, *bg)
    g.set_shot(*shot)
    d = g.born(dm)
    m_gpu, mis_gpu, _, _ = g.lsrtm_cg([shot], d[None], iters, pmask=3)
    g.close()
    o = oracle2d.Oracle2D(q, 64)
    o.load_model(1, *bg)
    o.set_shot(*shot)
    mask = np.zeros((3, 1, 1), np.float32)
    mask[:2] = 1
    x = np.zeros((3, p.nx, p.nz), np.float64)
    r = d.astype(np.float64)
    s = o.adjoint(r.astype(np.float32)).astype(np.float64) * mask
    pv = s.copy()
    gam = float(np.sum(s * s))
    mis = [1.0]
    log("bench visco: oracle check: FP64 adjoint done")
    for k in range(iters):
        qd = o.born(pv.astype(np.float32)).astype(np.float64)
        log(f"bench visco: oracle check: FP64 Born of iteration {k + 1} done")
        al = gam / float(np.sum(qd * qd))
        x += al * pv
        r -= al * qd
        mis.append(float(np.linalg.norm(r) / np.linalg.norm(d)))
        s = o.adjoint(r.astype(np.float32)).astype(np.float64) * mask
        gn = float(np.sum(s * s))
        pv = s + gn / gam * pv
        gam = gn
    o.close()
    e = float(np.linalg.norm(m_gpu[:2] - x[:2]) / np.linalg.norm(x[:2]))
    return {"shot": shot, "nt": nt, "iterations": iters, "model_rel_diff_gpu_vs_fp64": e,
            "misfit_gpu": [float(v) for v in mis_gpu], "misfit_fp64": mis}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nt", type=int, default=2500, help="time steps per pass (profiling runs shorten it)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the CPU baseline and the LSRTM comparison (profiling runs)")
    ap.add_argument("--no-kernel-events", action="store_true")
    ap.add_argument("--tile-x", type=int, default=0, choices=[0, 32, 64],
                    help="tile width (0: library default, 32 at radius 4); 32-wide tiles take tile_y 8 or 16")
    ap.add_argument("--tile-y", type=int, default=0, help="launch shape override (0: library default)")
    ap.add_argument("--zchunk", type=int, default=0, help="z planes per workgroup (0: library default)")
    ap.add_argument("--boundary", choices=["cpml", "damp"], default="cpml")
    ap.add_argument("--decim", type=int, default=4,
                    help="--imaging history: source steps per stored history slot")
    ap.add_argument("--imaging", choices=["excitation", "xcorr", "history"], default="excitation",
                    help="shot workload: excitation-amplitude imaging (default), or zero-lag cross-correlation "
                         "with the source wavefield reconstructed backward from boundary strips")
    ap.add_argument("--cpml-split", type=int, choices=[0, 1], default=0,
                    help="C-PML launches: 0 plain + C-PML tile lists (default), 1 one C-PML launch over all tiles")
    ap.add_argument("--region-streams", type=int, choices=[0, 1], default=0,
                    help="run the plain-tile and C-PML-tile launches of a step concurrently on two streams")
    ap.add_argument("--xcd-map", type=int, choices=[0, 1, 2], default=2,
                    help="tile order: 0 plain, 1 XCD-contiguous chunks, 2 XCD y bands (default)")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="only run the plain copy kernel over GIB GiB (HBM counter calibration) and exit")
    ap.add_argument("--shots", type=int, default=60, help="visco workload: shots")
    ap.add_argument("--batch-window", type=int, nargs=2, default=(192, 320), metavar=("LO", "HI"),
                    help="batch workload: lateral window sizes (cells; the configured 192 320)")
    ap.add_argument("--iters", type=int, default=20, help="visco workload: CG iterations")
    ap.add_argument("--workers", type=int, default=4, help="visco workload: concurrent shot contexts")
    ap.add_argument("--batch", type=int, default=0, help="visco workload: shots per batched launch (0: workers)")
    ap.add_argument("--oracle-check", action="store_true", help="visco workload: FP64 oracle CG check (1 shot)")
    ap.add_argument("--workload", choices=["shot", "batch", "visco", "learned", "rtm2d", "gradient"], default="shot",
                    help="shot: the one-shot 3D configuration (default); batch: 24 windowed shots over a "
                         "600x600x250 model in one batched, graph-captured launch per kernel per step; "
                         "learned: LSRTM time to misfit, learned preconditioner vs plain CG; gradient: the LSRTM "
                         "adjoint from the stored E/F history vs from boundary strips")
    ap.add_argument("--quick", action="store_true", help="learned workload: the reduced test case")
    ap.add_argument("--learned-test-shape", type=int, nargs=2, default=None, metavar=("NX", "NZ"),
                    help="learned workload: held-out model size")
    ap.add_argument("--learned-test-shots", type=int, default=None)
    ap.add_argument("--learned-test-lt", type=int, default=None)
    ap.add_argument("--learned-test-iters", type=int, default=None,
                    help="learned workload: CG iteration cap (0: train and save only)")
    ap.add_argument("--learned-params-out", default=None, help="learned workload: save the trained net (.npy)")
    ap.add_argument("--learned-train-shape", type=int, nargs=2, default=None, metavar=("NX", "NZ"))
    ap.add_argument("--learned-train-models", type=int, default=None)
    ap.add_argument("--learned-train-lt", type=int, default=None)
    ap.add_argument("--learned-train-shots", type=int, default=None)
    ap.add_argument("--learned-plain-file", default=None,
                    help="learned workload: plain-CG results file (.npz), written with --learned-plain-only, "
                         "read otherwise (plain CG is then not rerun)")
    ap.add_argument("--learned-plain-only", action="store_true", help="learned workload: run plain CG only")
    ap.add_argument("--learned-params-in", default=None,
                    help="learned workload: evaluate a saved net instead of generating data and training")
    ap.add_argument("--no-learned", action="store_true",
                    help="shot workload: skip the reduced learned-vs-plain LSRTM comparison")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (the whole grid fits one MI355X; shots would shard across GPUs, "
              f"one process each); --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    erm = importlib.import_module(PKG)
    if args.copy_calibration > 0:
        nbytes = int(args.copy_calibration * (1 << 30))
        gbs, ms = erm.copy_bandwidth(nbytes, reps=10)
        print(json.dumps({"copy_bytes": nbytes, "ms_per_copy": round(ms, 4), "GB/s": round(gbs, 1)}), flush=True)
        return
    if args.workload == "batch":
        return run_batch(erm, args)
    if args.workload == "visco":
        return run_visco(erm, args)
    if args.workload == "learned":
        return run_learned(erm, args)
    if args.workload == "rtm2d":
        return run_rtm2d_xcorr(erm, args) if args.imaging == "xcorr" else run_rtm2d(erm, args)
    if args.workload == "gradient":
        return run_gradient(erm, args)
    nt = args.nt
    p = make_params(erm, nt, erm.BOUNDARY_CPML if args.boundary == "cpml" else erm.BOUNDARY_DAMP)
    p.tile_x, p.tile_y, p.zchunk = args.tile_x, args.tile_y, args.zchunk
    p.xcd_map = args.xcd_map
    p.region_streams = args.region_streams
    p.cpml_split = args.cpml_split
    vp, vs, rho = erm.make_model3d(1, p.nx, p.ny, p.nz, p.dx, p.dy, p.dz, 2)
    vp_s = smooth_box(vp)
    vs_s = (vp_s / 1.732).astype(np.float32)
    rho_s = (230.0 * vp_s.astype(np.float64) ** 0.25).astype(np.float32)
    log(f"bench: grid {tuple(p.padded_shape)} nt={nt} radius 4, building context")
    g = erm.RTM3D(p, 0)
    g.load_model(0, vp, vs, rho)
    g.load_model(1, vp_s, vs_s, rho_s)
    t0 = time.perf_counter()
    g.forward_real(want_record=False)  # the record stays on the device
    log(f"bench: modeled the receiver record in {time.perf_counter() - t0:.2f} s")
    tune = None
    if not (args.tile_x or args.tile_y or args.zchunk):
        tune = g.autotune(steps=3)  # launch shape; results do not depend on it
        log(f"bench: autotuned {tune}")

    R = 4
    nnz, nny, nnx = g.shape
    cells = (nnx - 2 * R) * (nny - 2 * R) * (nnz - 2 * R)
    sb, vb_src = g.algorithmic_bytes(1)
    _, vb_rcv = g.algorithmic_bytes(2)

    xc = args.imaging != "excitation"
    if args.imaging == "xcorr":
        fwd, bwd = g.forward_smoothed_strips, g.backward_smoothed_xcorr
        strip_gb = g.strip_bytes() / 1e9
        log(f"bench: cross-correlation imaging, boundary strips {strip_gb:.1f} GB in HBM")
    elif args.imaging == "history":
        fwd, bwd = (lambda: g.forward_smoothed_history(args.decim)), g.backward_smoothed_xcorr_history
        strip_gb = g.history_bytes(args.decim) / 1e9
        log(f"bench: cross-correlation imaging, source history every {args.decim} steps, {strip_gb:.1f} GB in HBM")
    else:
        fwd, bwd = g.forward_smoothed, g.backward_smoothed
        strip_gb = None
    for w in range(args.warmup):
        fwd()
        bwd()
        log(f"bench: warmup shot {w + 1} done")
    g.set_profiling(not args.no_kernel_events)
    acc = {"stress_ms": 0.0, "stress_n": 0, "vel_ms": 0.0, "vel_n": 0, "vel_bytes": 0.0, "pass_ms": 0.0,
           "src_ms": 0.0, "rcv_ms": 0.0}
    g.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        for fn, vb, key in ((fwd, vb_src, "src_ms"), (bwd, vb_rcv, "rcv_ms")):
            fn()
            t = g.timing()
            acc["pass_ms"] += t["pass_ms"]
            acc[key] += t["pass_ms"]
            acc["stress_ms"] += t["stress_ms"]
            acc["stress_n"] += t["stress_launches"]
            acc["vel_ms"] += t["velocity_ms"]
            acc["vel_n"] += t["velocity_launches"]
            acc["vel_bytes"] += vb * cells * t["velocity_launches"]
        log(f"bench: timed shot {s + 1}/{args.steps} done")
    g.synchronize()
    el = time.perf_counter() - t0
    g.set_profiling(False)

    updates = cells * 2 * nt * args.steps
    value = updates / el / 1e9
    roof = None
    extra = {}
    if acc["stress_n"] > 0 and acc["vel_n"] > 0:
        st_avg = acc["stress_ms"] / acc["stress_n"]
        ve_avg = acc["vel_ms"] / acc["vel_n"]
        st_ach = sb * cells / (st_avg * 1e-3) / 1e9
        ve_bytes = acc["vel_bytes"] / acc["vel_n"]
        ve_ach = ve_bytes / (ve_avg * 1e-3) / 1e9
        dom = "k_velocity" if acc["vel_ms"] >= acc["stress_ms"] else "k_stress"
        ach, per_launch, avg = (ve_ach, ve_bytes, ve_avg) if dom == "k_velocity" else (st_ach, sb * cells, st_avg)
        traffic = None
        # the committed PMC traffic figures cover the default workload only
        # (296x296x168 grid, C-PML, excitation imaging)
        tf = os.path.join(ROOT, "profiles", "traffic_k_stress_k_velocity_296x296x168_cpml.json")
        profiled = args.boundary == "cpml" and tuple(g.shape) == (168, 296, 296) and not xc
        if profiled and os.path.exists(tf):
            with open(tf) as f:
                traffic = json.load(f).get(dom, {}).get("hbm_bytes_per_launch")
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom,
                "algorithmic_bytes_per_launch": per_launch, "avg_launch_ms": round(avg, 5)}
        extra = {"k_stress": {"avg_ms": round(st_avg, 5), "GB/s": round(st_ach, 1),
                              "bytes_per_cell": sb},
                 "k_velocity": {"avg_ms": round(ve_avg, 5), "GB/s": round(ve_ach, 1),
                                "bytes_per_cell": round(ve_bytes / cells, 2)}}
    g.close()

    # error against the FP64 oracle: the same shot cut to 500 steps per pass,
    # with the timed launch shape, against the committed oracle fixture
    fp64 = None
    fx_path = os.path.join(ROOT, "tests", "golden", "config2_nt500.npz")
    if args.boundary == "cpml" and not xc and tuple(g.shape) == (168, 296, 296) and os.path.exists(fx_path):
        sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
        import config2
        fx = np.load(fx_path)
        shape = {"tile_x": tune["tile_x"] if tune else args.tile_x,
                 "tile_y": tune["tile_y"] if tune else args.tile_y, "zchunk": tune["zchunk"] if tune else args.zchunk,
                 "xcd_map": args.xcd_map}
        pe, et, ea, imgs, vz = config2.gpu_shot(erm, int(fx["nt"]), **shape)
        fp64 = config2.compare(fx, pe, et, ea, imgs, vz)
        fp64 = {k: (round(v, 9) if isinstance(v, float) else v) for k, v in fp64.items()}
        log(f"bench: error vs FP64 oracle (nt={fp64['nt']}): {fp64}")

    cpu = None
    if not args.no_cpu_baseline:
        log("bench: CPU baseline (FP32 oracle)")
        cpu = cpu_baseline(p, vp_s, vs_s, rho_s, args.cpu_budget)

    # the metric's second half, LSRTM seconds to misfit 1e-2 learned vs plain,
    # on the reduced case (the scaled case: --workload learned)
    lsrtm = None
    if not (args.no_learned or args.no_cpu_baseline):
        log("bench: learned-preconditioner LSRTM, reduced case")
        try:
            lsrtm = learned_case(erm, quick=True)
        except Exception as e:  # the 3D line stands on its own
            lsrtm = {"error": repr(e)}
        log(f"bench: LSRTM learned vs plain: {lsrtm}")

    if args.imaging == "xcorr":
        imaging = ("zero-lag cross-correlation imaging, source wavefield reconstructed backward from %.0f GB of "
                   "boundary strips (the reconstruction's extra steps are not counted as updates)" % strip_gb)
    elif args.imaging == "history":
        imaging = ("zero-lag cross-correlation imaging against a source history of tp and vP stored every %d "
                   "steps (%.0f GB in HBM)" % (args.decim, strip_gb))
    else:
        imaging = "excitation-amplitude imaging"
    line = {
        "metric": "Gcell-updates/s 3D elastic RTM; LSRTM s to 1e-2 misfit, learned vs plain",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded dipping-layer + anomaly model)",
        "config": {"workload": "3D elastic RTM, one shot: 256x256x128 model + 20-cell %s "
                               "(296x296x168), 8th-order staggered FD, dt=0.8 ms, nt=%d, 15 Hz Ricker, "
                               "128x128 receivers; step = source pass + receiver pass; %s"
                               % ("C-PML" if args.boundary == "cpml" else "damping boundary", nt, imaging),
                   "cells_per_time_step": cells, "time_steps_per_shot": 2 * nt,
                   "pass_ms_per_shot": {"source": round(acc["src_ms"] / args.steps, 1),
                                        "receiver": round(acc["rcv_ms"] / args.steps, 1)},
                   "autotuned": tune, "parallelism": "single GPU"},
        "roofline": roof,
        "kernels": extra,
        "error_vs_fp64": fp64,
        "cpu_baseline": cpu,
        "lsrtm_learned_vs_plain": lsrtm,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
