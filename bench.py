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
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="rocprofv3 PMC summary (FETCH/WRITE KiB per dispatch) for the traffic field")
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


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU decomposition is a design (see the sharding notes), not a run.", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return
    import numpy as np
    import torch

    from fall3d_amd import ads
    from fall3d_amd.problem import operational_problem

    p = operational_problem(args.nx, args.ny, args.nz, args.nb)
    g = ads.ADS(p.nx, p.ny, p.nz, p.nb, "minmod", "rk", 8, args.state)
    g.set_problem(p)
    dt = g.time_step(0.0, 2, 0.9)
    cells = p.nx * p.ny * p.nz
    cell_bins = cells * p.nb

    g.run(dt, args.warmup, 0)
    g.synchronize()
    torch.cuda.synchronize()
    g.set_profiling(True)
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
            if ent:
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
