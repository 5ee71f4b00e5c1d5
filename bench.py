"""Benchmark: the FALL3D ADS step (Kurganov-Tadmor sweeps, Sources/mod_ADS.F90)
on one MI355X.

The headline workload is the large single domain: 1024 x 768 x 80 cells,
20 bins (settling velocities log-spaced 0.005-5 m/s), the puff problem's
wind profile rotated by a seeded angle, 3 seeded line sources, dx = 5 km,
dz = 250 m, minmod + RK4, CFL 0.9, FP32 state with FP64 budgets. One step =
one F3D_time_step: source injection, x and y sweeps (order by iteration
parity), z sweep with ground deposition, FP64 flux/budget reductions; inputs
resident in HBM before the timed region.

The default run prints ONE JSON line: the headline's Gcell*bin/s, the
roofline of its dominant kernel (algorithmic bytes / average launch time from
HIP events on the library's stream) and the FP64 CPU oracle timed on a
bounded sample, with sub-results under "sub":
  operational -- the 360 x 240 x 60 x 10 operational domain (Fourier wind,
                 central line source);
  ensemble    -- 64 members of uneven sizes (96-256 per horizontal axis, 40-60
                 levels, 6 bins) advanced in batched launches;
  large_fp64  -- the large domain with the FP64 state (the reference's
                 arithmetic in double precision; truth runs);
  learned     -- the learned coarse-grid correction on 512 x 384 x 80 x 20
                 against the plain fine run, at equal simulated time; errors
                 against the FP64 fine run.
Each sub-result times --steps steps after --warmup warm-up steps.
`--workload large|operational|ensemble|learned` runs one workload alone
(the profiled commands use these).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROF_EVERY = 10
CPU_THREADS = 16

# the large single domain (dx = dy = 5 km, dz = 250 m, 20 km top)
LARGE = (1024, 768, 80, 20)
LARGE_KW = {"dx": 5000.0, "dz": 250.0, "nsources": 3, "puff": False, "vmin": 0.005, "vmax": 5.0}
LARGE_SEED = 1234


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--state", default="fp32", choices=["fp32", "fp32x2", "fp64"])
    ap.add_argument("--workload", default="all", choices=["all", "large", "operational", "ensemble", "learned"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle samples")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="rocprofv3 PMC summary (FETCH/WRITE KiB per dispatch) for the traffic fields")
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--tiling", default="",
                    help="x_cells_per_lane,y_bins_per_group[/bins_per_thread],z_bins_per_group[/bins_per_thread]: "
                         "fixed tile shapes instead of the autotuner (profiling runs)")
    ap.add_argument("--eval-steps", type=int, default=100,
                    help="learned workload: coarse steps of the accuracy evaluation (fine: twice as many)")
    ap.add_argument("--train-seeds", type=int, default=32, help="learned workload: training seeds (512x512x40x4)")
    ap.add_argument("--train-steps", type=int, default=2000, help="learned workload: Adam steps of 65536 rows")
    ap.add_argument("--gains", default="1.0,0.5", help="learned workload: correction gains evaluated")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB with torch and with the library's "
                         "16-byte buffer-load copy (calibrates the FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--hbm-reference", type=int, default=0, metavar="MIB",
                    help="only time torch's streaming kernels (add, scale, copy, sum) on MIB MiB arrays, "
                         "as a reference for the bandwidth HBM delivers to plain streams, and exit")
    return ap.parse_args()


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
def hbm_reference(mib, reps=20):
    """Bandwidth of torch's elementwise streams on MIB MiB FP32 arrays, bytes
    read + written per call / average call time (torch events around reps
    calls after 3 untimed ones): the rate HBM delivers to a plain stream."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.rand(n, device="cuda:0")
    b = torch.rand(n, device="cuda:0")
    c = torch.empty_like(a)

    def rate(f, nbytes):
        for _ in range(3):
            f()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            f()
        e1.record()
        torch.cuda.synchronize()
        return round(nbytes / (e0.elapsed_time(e1) * 1e-3 / reps) / 1e9, 1)

    out = {"hbm_reference_mib": mib, "unit": "GB/s (bytes read + written per call / time)",
           "add": rate(lambda: torch.add(a, b, out=c), 3 * 4 * n),
           "scale": rate(lambda: torch.mul(a, 2.0, out=c), 2 * 4 * n),
           "copy": rate(lambda: c.copy_(a), 2 * 4 * n),
           "sum": rate(lambda: a.sum(), 4 * n)}
    print(json.dumps(out))


def copy_probe(mib):
    """Known-traffic references for the HBM counters: each copy reads and
    writes exactly mib MiB (torch's copy kernel, the library's copy with
    16-byte buffer loads and stores, the access width of the FP32 sweeps, and
    its copy with 8-byte loads and stores, that of the FP64 sweeps)."""
    import ctypes

    import torch

    from fall3d_amd import ads

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    lib = ads.load_library()
    for _ in range(5):
        rc = lib.fall3d_ads_copy_probe(ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()),
                                       ctypes.c_size_t(n), None)
        assert rc == 0
    torch.cuda.synchronize()
    assert bool((dst == src).all())
    src64 = src.view(torch.float64)  # the same bytes as n/2 doubles
    dst.zero_()
    for _ in range(5):
        rc = lib.fall3d_ads_copy_probe64(ctypes.c_void_p(src64.data_ptr()), ctypes.c_void_p(dst.data_ptr()),
                                         ctypes.c_size_t(n // 2), None)
        assert rc == 0
    torch.cuda.synchronize()
    assert bool((dst == src).all())
    print(json.dumps({"copy_probe_mib": mib, "bytes_read_per_copy": n * 4, "bytes_written_per_copy": n * 4}))


def _traffic(args, key, alg_bytes):
    """Calibrated HBM bytes per launch from the committed PMC summary, only
    when that entry was measured on a launch with these algorithmic bytes."""
    if not os.path.exists(args.traffic):
        return None
    try:
        tr = json.load(open(args.traffic))
        ent = tr.get(key)
        if ent and abs(ent.get("algorithmic_bytes", -1) - alg_bytes) <= 1e-6 * alg_bytes:
            return ent.get("bytes_per_launch")
    except Exception:
        return None
    return None


def _cpu_sample(prob, dt, steps, label):
    """The FP64 oracle (OpenMP over lines) on a bounded sample of a workload."""
    from oracle import ads_oracle
    from oracle.ads_oracle import Oracle

    ads_oracle.set_num_threads(CPU_THREADS)
    o = Oracle(prob, "minmod", "rk")
    c, acum = prob.new_c(), prob.new_acum()
    t0 = time.perf_counter()
    for it in range(steps):
        o.step(dt, it, c, acum)
    t1 = time.perf_counter()
    cb = prob.nx * prob.ny * prob.nz * prob.nb * steps
    return {"value": round(cb / (t1 - t0) / 1e9, 5), "unit": "Gcell*bin/s", "cores": CPU_THREADS, "kind": "port",
            "sample": f"{label} ({t1 - t0:.1f} s, FP64, OpenMP over lines, {CPU_THREADS} threads)"}


# ---------------------------------------------------------------------------
def single_domain(args, which):
    """One resident domain: autotune, warm up, time args.steps steps with HIP
    events on one step in PROF_EVERY, budget, CPU sample."""
    import numpy as np
    import torch

    from fall3d_amd import ads
    from fall3d_amd.problem import config1_wind_problem, operational_problem

    t_setup = time.perf_counter()
    if which == "large":
        nx, ny, nz, nb = LARGE
        p = config1_wind_problem(nx, ny, nz, nb, seed=LARGE_SEED, dtype=np.float32, **LARGE_KW)
        kind = 4
        desc = (f"large single domain {nx}x{ny}x{nz}, {nb} bins (v_s 0.005-5 m/s), 3 line sources, "
                "dx 5 km, dz 250 m, minmod RK4, CFL 0.9")
    else:
        p = operational_problem(360, 240, 60, 10)
        kind = 8
        desc = "operational single domain 360x240x60, 10 bins, minmod RK4, CFL 0.9"
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", kind, args.state)
    g.set_problem(p)
    dt = g.time_step(0.0, 2, 0.9)
    if args.state == "fp64":
        tiling = (0, 0, 0)  # one thread per line: no tile shapes
    elif args.tiling:
        sx, sy, sz = args.tiling.split(",")
        tx = int(sx)
        ty, by = (int(v) for v in (sy + "/0").split("/")[:2])
        tz, bz = (int(v) for v in (sz + "/0").split("/")[:2])
        g.set_tiling(0, cells_per_thread=tx)
        g.set_tiling(1, cells_per_thread=by, bins_per_group=ty)
        g.set_tiling(2, cells_per_thread=bz, bins_per_group=tz)
        tiling = (tx, ty, tz)
    else:
        tiling, _ = g.autotune(dt, 0.0, 3)
    setup_s = time.perf_counter() - t_setup
    cells = p.nx * p.ny * p.nz
    cell_bins = cells * p.nb
    m0 = g.budget()["mass_cells"]

    g.run(dt, args.warmup, 0)
    g.synchronize()
    torch.cuda.synchronize()
    g.set_profiling(PROF_EVERY)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.run(dt, args.steps, args.warmup)
    g.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ktimes = {k: g.kernel_time(k) for k in ("source", "sweep_x", "sweep_y", "sweep_z", "reduce")}
    g.set_profiling(False)
    b = g.budget()
    value = cell_bins * args.steps / elapsed / 1e9
    sweeps = {"sweep_x": 0, "sweep_y": 1, "sweep_z": 2}
    dom = max(sweeps, key=lambda k: ktimes[k][0])
    ms, n = ktimes[dom]
    avg_s = ms / 1e3 / max(n, 1)
    alg_bytes = g.sweep_bytes(sweeps[dom])
    achieved = alg_bytes / avg_s / 1e9
    step_alg_bytes = cells * (28 * p.nb + 36)  # state r/w x3 sweeps + v_s + 9 coefficient fields
    closure = abs(b["mass_cells"] + float(np.sum(b["ledger"])) + b.get("residual", 0.0) - b["injected"] - m0) / \
        max(b["injected"] + m0, 1e-300)
    btiles = {"x_cells_per_lane": tiling[0], "y_bins_per_group": tiling[1], "z_bins_per_group": tiling[2],
              "y_bins_per_thread": g.tiling(1)[0], "z_bins_per_thread": g.tiling(2)[0]}
    g.release()
    del g

    cpu = None
    if not args.no_cpu:
        if which == "large":
            # a 256 x 192 horizontal window of the same generator (1/16 of the domain)
            ps = config1_wind_problem(256, 192, LARGE[2], LARGE[3], seed=LARGE_SEED, **LARGE_KW)
            cpu = _cpu_sample(ps, dt, 10, f"10 steps of the same generator on a 256x192x{LARGE[2]}x{LARGE[3]} "
                                          "window (1/16 of the horizontal domain, same spacing, winds, bins)")
        else:
            cpu = _cpu_sample(p, dt, 30, f"30 steps of the same {p.nx}x{p.ny}x{p.nz}x{p.nb} workload")
    del p

    return {
        "metric": ("Gcell*bin updates/s (ADS KT step, FP64 state and fields)" if args.state == "fp64" else
                   "Gcell*bin updates/s (ADS KT step, FP32 bulk, FP64 budgets)"),
        "value": round(value, 3),
        "unit": "Gcell*bin/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "fp32x2": "f32x2", "fp64": "f64"}[args.state],
        "data": "synthetic (seeded winds and line sources)",
        "config": {
            "workload": desc,
            "cell_bins": cell_bins,
            "state": args.state,
            "dt_s": dt,
            "setup_s": round(setup_s, 1),
            "step_hbm_roofline_frac": round(step_alg_bytes * args.steps / elapsed / 1e9 / HBM_PEAK_GBS, 4),
            "kernel_ms": {k: round(v[0] / max(v[1], 1), 5) for k, v in ktimes.items()},
            "kernel_timing": f"HIP events on the library stream, 1 step in {PROF_EVERY} of the timed region",
            "autotuned_tiling": btiles,
            "mass_closure": closure,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": _traffic(args, f"{which}:{args.state}:{dom}", alg_bytes),
            "algorithmic_bytes_per_launch": alg_bytes,
        },
        "cpu_baseline": cpu,
    }


# ---------------------------------------------------------------------------
def ensemble(args):
    """64 members, nx, ny ~ U{96..256}, nz ~ U{40..60}, 6 bins, perturbed
    column height / mass rate / wind modes; autotuned tile shapes; args.steps
    steps of every member as one replayed graph of batched launches."""
    import numpy as np
    import torch

    from fall3d_amd import ads
    from fall3d_amd.problem import ensemble_members

    specs = ensemble_members(args.members, nb=6, seed=3)
    ens = ads.Ensemble([(m.nx, m.ny, m.nz) for m in specs], 6, "minmod", "rk", 8, args.state)
    dts = []
    t0 = time.perf_counter()
    for g, sp in zip(ens.members, specs):
        p = sp.problem()
        g.set_problem(p)
        dts.append(g.time_step(0.0, 2, 0.9))
        del p
    setup_s = time.perf_counter() - t0
    dts = np.array(dts)
    if args.tiling:  # x cells per lane, y and z bins per workgroup (profiling runs)
        tx, ty, tz = (int(v.split("/")[0]) for v in args.tiling.split(","))
        ens.set_tiling(tx, ty, tz)
        tiling = ens.tiling()
        ms = [1.0, 1.0, 1.0]
    else:
        tiling, ms = ens.autotune(dts, reps=3)
    m0 = [g.budget()["mass_cells"] for g in ens.members]
    ens.run(dts, args.warmup, 0)
    ens.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ens.run(dts, args.steps, args.warmup)
    ens.synchronize()
    torch.cuda.synchronize()
    el = time.perf_counter() - t1
    cb = sum(m.nx * m.ny * m.nz * m.nb for m in specs)
    cells = sum(m.nx * m.ny * m.nz for m in specs)
    alg = [cells * (6 * 8 + 12), cells * (6 * 8 + 12), cells * (6 * 8 + 12) + cells * 6 * 4]
    d = int(np.argmax(ms))
    ach = alg[d] / (ms[d] / 1e3) / 1e9
    closure = 0.0
    for g, mm in zip(ens.members, m0):
        b = g.budget()
        closure = max(closure, abs(b["mass_cells"] + float(np.sum(b["ledger"])) + b.get("residual", 0.0)
                                   - b["injected"] - mm) / max(b["injected"] + mm, 1e-300))
    ens.release()
    cpu = None
    if not args.no_cpu:
        p0 = specs[0].problem()
        cpu = _cpu_sample(p0, dts[0], 10, f"10 steps of member 0 ({p0.nx}x{p0.ny}x{p0.nz}x{p0.nb}), "
                                          "the unperturbed member")
    return {
        "metric": "Gcell*bin updates/s (ADS KT step, ensemble batch)",
        "value": round(cb * args.steps / el / 1e9, 3), "unit": "Gcell*bin/s", "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 4),
        "config": {"workload": f"ensemble {args.members} members, nx,ny U(96..256), nz U(40..60), 6 bins",
                   "cell_bins_per_step": cb, "autotuned_tiling": list(tiling),
                   "sweep_ms": [round(v, 4) for v in ms], "setup_s": round(setup_s, 1),
                   "mass_closure_max": closure},
        "roofline": {"bound": "hbm", "kernel": "sweep_" + "xyz"[d], "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": _traffic(args, f"ensemble:{args.state}:sweep_{'xyz'[d]}", alg[d]),
                     "algorithmic_bytes_per_launch": alg[d], "timing": "autotuner's per-sweep launch times"},
        "cpu_baseline": cpu,
    }


# ---------------------------------------------------------------------------
def learned_compare(args):
    """Train the flux correction on fine/coarse runs (512x512x40, 4 bins), then
    on the large domain compare the learned coarse run (512x384x80x20, the
    coarse CFL step dt_c) with the plain fine run at its own CFL step dt_f:
    throughput at equal simulated time, and the column-load and deposit L1
    errors of plain and learned coarse runs against the fine run of the
    FP64 state after args.eval_steps coarse steps."""
    import numpy as np
    import torch

    from fall3d_amd import ads, learned
    from fall3d_amd.problem import config1_wind_problem

    t0 = time.perf_counter()
    tr = (512, 512, 40, 4)
    kw_tr = {"dx": 5000.0, "dz": 500.0, "nsources": 2, "puff": True, "vmin": 0.005, "vmax": 5.0}
    net = ads.Network(capacity=int(args.train_seeds * 8_000_000 + 1_000_000))
    rows = learned.collect(net, range(args.train_seeds), tr, nwarm=40, nsamples=8, gap=60, **kw_tr)
    t1 = time.perf_counter()
    loss = net.train(args.train_steps, batch=65536, lr=1e-3, seed=1)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    log(f"learned: {rows} rows in {t1 - t0:.1f} s, trained in {t2 - t1:.1f} s")

    nx, ny, nz, nb = LARGE
    kw = dict(LARGE_KW)
    f32 = np.float32
    pf = config1_wind_problem(nx, ny, nz, nb, seed=LARGE_SEED, dtype=f32, **kw)
    gf = ads.ADS(nx, ny, nz, nb, "minmod", "rk", 4, "fp32")
    gf.set_problem(pf)
    pc = config1_wind_problem(nx, ny, nz, nb, seed=LARGE_SEED, coarsen=2, dtype=f32, **kw)
    gc = ads.ADS(pc.nx, pc.ny, pc.nz, nb, "minmod", "rk", 4, "fp32")
    gc.set_problem(pc)
    dt_f = gf.time_step(0.0, 2, 0.9)                 # the fine grid's own CFL step
    dt_c = min(gc.time_step(0.0, 2, 0.9), 2.0 * dt_f)  # the coarse step the correction is trained for
    tiles = {"fine": list(gf.autotune(dt_f)[0]), "coarse": list(gc.autotune(dt_c)[0])}
    gc.set_learned(net, 1e-10)
    k = args.steps
    # the fine run covers the coarse run's simulated time k dt_c at its own step
    kf = max(1, int(round(k * dt_c / dt_f)))
    for g, d in ((gf, dt_f), (gc, dt_c)):
        g.run(d, args.warmup, 0)
        g.synchronize()
    torch.cuda.synchronize()
    ta = time.perf_counter()
    gf.run(dt_f, kf, args.warmup)
    gf.synchronize()
    tb = time.perf_counter()
    gc.run(dt_c, k, args.warmup)
    gc.synchronize()
    tc = time.perf_counter()
    gf.release()
    gc.release()
    fine_ms, coarse_ms = (tb - ta) / kf * 1e3, (tc - tb) / k * 1e3
    # fine cell*bin updates per simulated second over wall time per simulated second
    fine_rate = nx * ny * nz * nb / dt_f              # fine cell*bin updates per simulated second
    thr_fine = fine_rate * (kf * dt_f) / (tb - ta) / 1e9
    thr_learned = fine_rate * (k * dt_c) / (tc - tb) / 1e9
    speedup = thr_learned / thr_fine
    log(f"learned: fine {fine_ms:.3f} ms/step at dt {dt_f:.2f} s, learned coarse {coarse_ms:.3f} ms/step at "
        f"dt {dt_c:.2f} s -> {speedup:.2f}x at equal simulated time")
    del pf, pc
    res = learned.evaluate(net, LARGE_SEED, LARGE, args.eval_steps, smin=1e-10, truth_state="fp64",
                           fine_fp32=True, gains=tuple(float(g) for g in args.gains.split(",")), dtype=f32,
                           real_kind=4,
                           progress=lambda m: log(f"learned: {m}"), **kw)
    net.release()
    log(f"learned: evaluated {res}")
    return {
        "metric": "fine-equivalent Gcell*bin/s of the learned coarse step at equal simulated time",
        "value": round(thr_learned, 3), "unit": "Gcell*bin/s", "steps": k, "warmup": args.warmup,
        "ms_per_step": round(coarse_ms, 4),
        "config": {"workload": f"learned coarse {nx // 2}x{ny // 2}x{nz}x{nb} vs fine {nx}x{ny}x{nz}x{nb}",
                   "dt_fine_s": dt_f, "dt_coarse_s": dt_c, "fine_steps_timed": kf, "fine_ms_per_step":
                   round(fine_ms, 4), "fine_plain_gcbs": round(thr_fine, 3),
                   "speedup_equal_sim_time": round(speedup, 3),
                   "training": {"grid": list(tr), "seeds": args.train_seeds, "rows": rows, "adam_steps": args.train_steps,
                                "collect_s": round(t1 - t0, 1), "train_s": round(t2 - t1, 2),
                                "loss_first": float(loss[:10].mean()), "loss_last": float(loss[-50:].mean())},
                   "eval_coarse_steps": args.eval_steps, "errors_vs_fine_fp64": res,
                   "autotuned_tiling": tiles},
        "roofline": None, "cpu_baseline": None,
    }


# ---------------------------------------------------------------------------
def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU decomposition is a design (see the sharding notes), not a run.", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return
    if args.hbm_reference > 0:
        hbm_reference(args.hbm_reference)
        return
    if args.workload == "ensemble":
        out = ensemble(args)
    elif args.workload == "learned":
        out = learned_compare(args)
    elif args.workload in ("large", "operational"):
        out = single_domain(args, args.workload)
    else:
        out = single_domain(args, "large")
        log(f"large: {out['value']} Gcell*bin/s, roofline {out['roofline']['frac']}")
        sub = {}
        sub["operational"] = single_domain(args, "operational")
        log(f"operational: {sub['operational']['value']} Gcell*bin/s")
        sub["ensemble"] = ensemble(args)
        log(f"ensemble: {sub['ensemble']['value']} Gcell*bin/s")
        # the FP64 state on the large domain (truth runs): a few steps, no CPU sample
        a64 = argparse.Namespace(**dict(vars(args), state="fp64", no_cpu=True, tiling="",
                                        steps=max(2, args.steps // 4), warmup=1))
        sub["large_fp64"] = single_domain(a64, "large")
        log(f"large fp64: {sub['large_fp64']['value']} Gcell*bin/s")
        sub["learned"] = learned_compare(args)
        out["sub"] = sub
    print(json.dumps(out))


if __name__ == "__main__":
    main()
