"""Benchmark: the FALL3D ADS step (Kurganov-Tadmor sweeps, Sources/mod_ADS.F90)
on one MI355X.

Workload: the operational single domain, 360x240x60 cells, 10 tephra bins
(settling velocities log-spaced 0.01-3 m/s), seeded 16-mode Fourier wind
(15 m/s) + 2 m/s/km shear, vertical line source 0-15 km (1e6 kg/s over the
bins), minmod + RK4, CFL 0.9. One step = one F3D_time_step: source
injection, x and y sweeps (order by iteration parity), z sweep with ground
deposition, FP64 flux/budget reductions; inputs resident in HBM.

Prints ONE JSON line: Gcell*bin/s for the whole job, the roofline of the
dominant kernel (algorithmic bytes / average launch time, HIP events on
the library's stream) and the FP64 CPU oracle timed on a bounded sample.
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--state", default="fp32", choices=["fp32", "fp32x2"])
    ap.add_argument("--nx", type=int, default=360)
    ap.add_argument("--ny", type=int, default=240)
    ap.add_argument("--nz", type=int, default=60)
    ap.add_argument("--nb", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="rocprofv3 PMC summary (FETCH/WRITE KiB per dispatch) for the traffic field")
    ap.add_argument("--workload", default="operational", choices=["operational", "ensemble", "learned"],
                    help="operational: the single 360x240x60x10 domain (default, the headline line); "
                         "ensemble: 64 members of uneven sizes in batched launches; learned: training and "
                         "evaluation of the coarse-grid flux correction")
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--tiling", default="",
                    help="x_cells_per_lane,y_bins_per_group[/bins_per_thread],z_bins_per_group[/bins_per_thread]: "
                         "fixed tile shapes instead of the autotuner (profiling runs)")
    ap.add_argument("--scale", type=float, default=1.0,
                    help="learned workload: linear scale of the grids (1.0 = 512x512x40 training, "
                         "1024x768x80x20 evaluation)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (calibrates the FETCH_SIZE / "
                         "WRITE_SIZE counters against a kernel of known traffic) and exit")
    return ap.parse_args()


def copy_probe(mib):
    """Known-traffic reference for the HBM counters: each copy reads and
    writes exactly mib MiB."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_mib": mib, "bytes_read_per_copy": n * 4, "bytes_written_per_copy": n * 4}))


def bench_ensemble(args):
    """64 members, nx, ny ~ U{96..256}, nz ~ U{40..60}, 6 bins, perturbed
    column height / mass rate / wind modes; autotuned tile shapes; 300 steps
    (args.steps) of every member as one replayed graph of batched launches."""
    import numpy as np
    import torch

    from fall3d_amd import ads
    from fall3d_amd.problem import ensemble_members

    specs = ensemble_members(args.members, nb=6, seed=3)
    ens = ads.Ensemble([(m.nx, m.ny, m.nz) for m in specs], 6, "minmod", "rk", 8, args.state)
    dts = []
    t0 = time.perf_counter()
    for g, sp in zip(ens.members, specs):
        p = sp.problem(dtype_c=np.float32)
        g.set_grid(p.grid)
        g.set_meteo(p.u1, p.v1, p.w1, p.k1, p.k2, p.k3, p.vs, p.u2, p.v2, p.w2)
        g.set_source(p.s)
        g.set_state(None)
        dts.append(g.time_step(0.0, 2, 0.9))
        del p
    setup_s = time.perf_counter() - t0
    dts = np.array(dts)
    tiling, ms = ens.autotune(dts, reps=3)
    ens.run(dts, max(args.warmup, 2), 0)
    ens.synchronize()
    torch.cuda.synchronize()
    steps = args.steps if args.steps != 500 else 300
    t1 = time.perf_counter()
    ens.run(dts, steps, max(args.warmup, 2))
    ens.synchronize()
    torch.cuda.synchronize()
    el = time.perf_counter() - t1
    cb = sum(m.nx * m.ny * m.nz * m.nb for m in specs)
    cells = sum(m.nx * m.ny * m.nz for m in specs)
    alg = [cells * (6 * 8 + 12), cells * (6 * 8 + 12), cells * (6 * 8 + 12) + cells * 6 * 4]
    d = int(np.argmax(ms))
    ach = alg[d] / (ms[d] / 1e3) / 1e9
    mass = [g.budget() for g in ens.members]
    closure = max(abs(b["mass_cells"] + float(np.sum(b["ledger"])) - b["injected"]) / max(b["injected"], 1e-300)
                  for b in mass)
    print(json.dumps({
        "metric": "Gcell*bin updates/s (ADS KT step, ensemble batch)",
        "value": round(cb * steps / el / 1e9, 3), "unit": "Gcell*bin/s", "n_gpus": 1, "steps": steps,
        "warmup": max(args.warmup, 2), "ms_per_step": round(el / steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if args.state == "fp32" else "f32x2",
        "data": "synthetic (perturbed operational members)",
        "config": {"workload": f"ensemble {args.members} members, nx,ny U(96..256), nz U(40..60), 6 bins",
                   "cell_bins_per_step": cb, "autotuned_tiling": list(tiling),
                   "sweep_ms": [round(v, 4) for v in ms], "setup_s": round(setup_s, 1),
                   "mass_closure_max": closure},
        "roofline": {"bound": "hbm", "kernel": "sweep_" + "xyz"[d], "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_launch": alg[d]},
        "cpu_baseline": None,
    }))
    ens.release()


def bench_learned(args):
    """Train the flux correction on fine/coarse runs (512x512x40, 4 bins, 32
    seeds at scale 1, 8 samples 60 fine steps apart per seed) and compare plain coarse, learned coarse and plain fine
    (1024x768x80, 20 bins, 3 line sources at scale 1) against the fine run of
    the compensated state; throughput in fine-equivalent cell*bin/s."""
    import numpy as np
    import torch

    from fall3d_amd import ads, learned
    from fall3d_amd.problem import config1_wind_problem

    sc = args.scale

    def even(v):
        return max(8, int(round(v * sc / 8)) * 8)

    tr = (even(512), even(512), max(8, int(40 * sc)), 4)
    kw_tr = {"dx": 5000.0, "dz": 500.0, "nsources": 2, "puff": True, "vmin": 0.005, "vmax": 5.0}
    net = ads.Network(capacity=int(120_000_000 * min(1.0, sc * sc) + 1_000_000))
    t0 = time.perf_counter()
    # samples spread over ~900 fine steps of each run: the evaluation runs 1000
    # fine-equivalent steps, so the rows cover early and late plumes alike
    rows = learned.collect(net, range(int(max(4, 32 * min(1.0, sc)))), tr, nwarm=40, nsamples=8, gap=60, **kw_tr)
    t1 = time.perf_counter()
    print(f"bench: {rows} training rows in {t1 - t0:.1f} s", file=sys.stderr, flush=True)
    loss = net.train(2000, batch=65536, lr=1e-3, seed=1)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    print(f"bench: trained in {t2 - t1:.1f} s, loss {loss[:10].mean():.4e} -> {loss[-50:].mean():.4e}",
          file=sys.stderr, flush=True)
    ev = (even(1024), even(768), max(8, int(80 * sc)), 20)
    kw_ev = {"dx": 5000.0, "dz": 250.0, "nsources": 3, "puff": False, "vmin": 0.005, "vmax": 5.0}
    nsteps = max(10, int(args.steps if args.steps != 500 else 500))  # coarse steps (fine: twice as many)
    res = learned.evaluate(net, 1234, ev, nsteps, smin=1e-10, truth_state="fp32x2", fine_fp32=True,
                           gains=(1.0, 0.5), **kw_ev)
    print(f"bench: evaluated {res}", file=sys.stderr, flush=True)
    # throughput: plain fine vs learned coarse, same simulated time
    nx, ny, nz, nb = ev
    pf = config1_wind_problem(nx, ny, nz, nb, 5000.0, 250.0, 1234, nsources=3, puff=False, vmin=0.005, vmax=5.0)
    gf = ads.ADS(nx, ny, nz, nb, "minmod", "rk", 8, "fp32")
    gf.set_problem(pf)
    pc = config1_wind_problem(nx, ny, nz, nb, 5000.0, 250.0, 1234, coarsen=2, nsources=3, puff=False,
                              vmin=0.005, vmax=5.0)
    del pf
    gc = ads.ADS(pc.nx, pc.ny, pc.nz, nb, "minmod", "rk", 8, "fp32")
    gc.set_problem(pc)
    dt = min(gc.time_step(0.0, 2, 0.9), 2.0 * gf.time_step(0.0, 2, 0.9))
    # both runs with their autotuned tile shapes (the autotuner restores the state)
    tiles = {"fine": list(gf.autotune(0.5 * dt)[0]), "coarse": list(gc.autotune(dt)[0])}
    gc.set_learned(net, 1e-10)
    k = 20
    for g, d, n in ((gf, 0.5 * dt, 2 * k), (gc, dt, k)):
        g.run(d, 4, 0)
        g.synchronize()
    torch.cuda.synchronize()
    ta = time.perf_counter()
    gf.run(0.5 * dt, 2 * k, 4)
    gf.synchronize()
    tb = time.perf_counter()
    gc.run(dt, k, 4)
    gc.synchronize()
    tc = time.perf_counter()
    fine_cb = nx * ny * nz * nb * 2 * k
    thr_fine = fine_cb / (tb - ta) / 1e9
    thr_learned = fine_cb / (tc - tb) / 1e9
    print(json.dumps({
        "metric": "fine-equivalent Gcell*bin/s of the learned coarse step; deposit rel L1 vs fine",
        "value": round(thr_learned, 3), "unit": "Gcell*bin/s", "n_gpus": 1, "steps": k, "warmup": 4,
        "ms_per_step": round((tc - tb) / k * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (bf16 network)", "data": "synthetic (config-1 wind, line sources)",
        "config": {"workload": f"learned coarse {ev[0] // 2}x{ev[1] // 2}x{ev[2]}x{ev[3]} vs fine "
                               f"{ev[0]}x{ev[1]}x{ev[2]}x{ev[3]}",
                   "training_grid": list(tr), "training_rows": rows, "collect_s": round(t1 - t0, 1),
                   "train_s": round(t2 - t1, 2), "loss_first": float(loss[:10].mean()),
                   "loss_last": float(loss[-50:].mean()), "eval_coarse_steps": nsteps, "errors": res,
                   "fine_plain_gcbs": round(thr_fine, 3), "speedup_vs_fine": round(thr_learned / thr_fine, 3),
                   "autotuned_tiling": tiles},
        "roofline": None, "cpu_baseline": None,
    }))
    gf.release()
    gc.release()
    net.release()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU decomposition is a design (see the sharding notes), not a run.", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return
    if args.workload == "ensemble":
        bench_ensemble(args)
        return
    if args.workload == "learned":
        bench_learned(args)
        return
    import numpy as np
    import torch

    from fall3d_amd import ads
    from fall3d_amd.problem import operational_problem

    p = operational_problem(args.nx, args.ny, args.nz, args.nb)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, args.state)
    g.set_problem(p)
    dt = g.time_step(0.0, 2, 0.9)
    if args.tiling:
        sx, sy, sz = args.tiling.split(",")
        tx = int(sx)
        ty, by = (int(v) for v in (sy + "/0").split("/")[:2])
        tz, bz = (int(v) for v in (sz + "/0").split("/")[:2])
        g.set_tiling(0, cells_per_thread=tx)
        g.set_tiling(1, cells_per_thread=by, bins_per_group=ty)
        g.set_tiling(2, cells_per_thread=bz, bins_per_group=tz)
        tiling = (tx, ty, tz)
    else:
        tiling, tune_ms = g.autotune(dt, 0.0, 3)
    cells = p.nx * p.ny * p.nz
    cell_bins = cells * p.nb

    g.run(dt, args.warmup, 0)
    g.synchronize()
    torch.cuda.synchronize()
    # per-kernel HIP events on one step in PROF_EVERY of the timed region (the
    # steps in between are graph replays: events on every step would cost ~8%)
    g.set_profiling(PROF_EVERY)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.run(dt, args.steps, args.warmup)
    g.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ktimes = {k: g.kernel_time(k) for k in ("source", "sweep_x", "sweep_y", "sweep_z", "reduce")}
    g.set_profiling(False)
    b = g.budget()

    value = cell_bins * args.steps / elapsed / 1e9
    # dominant kernel by total time
    sweeps = {"sweep_x": 0, "sweep_y": 1, "sweep_z": 2}
    dom = max(sweeps, key=lambda k: ktimes[k][0])
    ms, n = ktimes[dom]
    avg_s = ms / 1e3 / max(n, 1)
    alg_bytes = g.sweep_bytes(sweeps[dom])
    achieved = alg_bytes / avg_s / 1e9
    step_alg_bytes = cells * (28 * p.nb + 36)  # state r/w x3 sweeps + v_s + 9 coefficient fields
    traffic = None
    if os.path.exists(args.traffic):
        try:
            tr = json.load(open(args.traffic))
            ent = tr.get(f"{args.state}:{dom}") or tr.get(dom)
            # measured for one workload: only when this launch moves the same algorithmic bytes
            if ent and abs(ent.get("algorithmic_bytes", -1) - alg_bytes) <= 1e-6 * alg_bytes:
                traffic = ent.get("bytes_per_launch")
        except Exception:
            traffic = None

    cpu = None
    if not args.no_cpu and args.cpu_steps > 0:
        from oracle import ads_oracle
        from oracle.ads_oracle import Oracle

        nthr = min(os.cpu_count() or 1, 16)
        ads_oracle.set_num_threads(nthr)
        o = Oracle(p, "minmod", "rk")
        c, acum = p.new_c(), p.new_acum()
        tc0 = time.perf_counter()
        for it in range(args.cpu_steps):
            o.step(dt, it, c, acum)
        tc1 = time.perf_counter()
        cpu = {
            "value": round(cell_bins * args.cpu_steps / (tc1 - tc0) / 1e9, 5),
            "unit": "Gcell*bin/s",
            "cores": nthr,
            "kind": "port",
            "sample": f"{args.cpu_steps} FP64 steps of the same {p.nx}x{p.ny}x{p.nz}x{p.nb} workload "
                      f"({tc1 - tc0:.1f} s, OpenMP over lines)",
        }

    out = {
        "metric": "Gcell*bin updates/s (ADS KT step, FP32 bulk, FP64 budgets)",
        "value": round(value, 3),
        "unit": "Gcell*bin/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.state == "fp32" else "f32x2",
        "data": "synthetic (seeded Fourier wind, line source)",
        "config": {
            "workload": f"operational single domain {p.nx}x{p.ny}x{p.nz}, {p.nb} bins, minmod RK4, CFL 0.9",
            "state": args.state,
            "dt_s": dt,
            "step_hbm_roofline_frac": round(step_alg_bytes * args.steps / elapsed / 1e9 / HBM_PEAK_GBS, 4),
            "kernel_ms": {k: round(v[0] / max(v[1], 1), 5) for k, v in ktimes.items()},
            "kernel_timing": f"HIP events on the library stream, 1 step in {PROF_EVERY} of the timed region",
            "autotuned_tiling": {"x_cells_per_lane": tiling[0], "y_bins_per_group": tiling[1],
                                 "z_bins_per_group": tiling[2], "y_bins_per_thread": g.tiling(1)[0],
                                 "z_bins_per_thread": g.tiling(2)[0]},
            "mass_closure": abs(b["mass_cells"] + float(np.sum(b["ledger"])) - b["injected"]) / max(b["injected"],
                                                                                                     1e-300),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": alg_bytes,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
