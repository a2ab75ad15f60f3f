"""Benchmark: FP64 RK3 steps/s of the channel step at 192 x 129 x 192.

Workload (the headline turbulent-channel configuration): Re_c = 3300
(Re_tau = 180), box 4pi x 2 x 4pi/3, 192 x 129 x 192 Fourier/grid points,
3/2-dealiased to 288 x 288, tanh-stretched y (2.75), dt = 0.005, constant
mass flux.  Synthetic seeded input: Reichardt mean profile plus solenoidal
k^-5/3 fluctuations of rms 0.1.  One "step" = boundary + three
(nonlinear, rhs_solve) substeps, all device resident.

Prints one JSON line.  `value` is RK3 steps/s of the whole job (one GPU).
`roofline` is computed for the dominant part of the step from HIP events
recorded on the library's stream; `cpu_baseline` times the numpy oracle on a
bounded sample (2 steps) of the same workload on one host core.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "3d-navier-stokes-thesis-paper_amd")

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def load_pkg():
    if "chanflow" in sys.modules:
        return sys.modules["chanflow"]
    spec = importlib.util.spec_from_file_location(
        "chanflow", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["chanflow"] = mod
    spec.loader.exec_module(mod)
    return mod


HEADLINE = dict(nx=192, ny=129, nz=192, nx2=288, nz2=288, lx_over_pi=4.0, aspect=3.0, re=3300.0,
                dt=0.005, stretch=2.75, seed=0, rms=0.1, re_tau=180.0)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary
    (FETCH_SIZE and WRITE_SIZE runs scaled by a plain copy kernel's counts);
    None when the summary does not list it."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_192x129x192.md")
    if not os.path.exists(path):
        return None, None
    txt = open(path).read()
    if "```json" not in txt:
        return None, None
    block = txt.split("```json", 1)[1].split("```", 1)[0]
    data = json.loads(block)
    v = data.get("bytes_per_launch", {}).get(kernel)
    return (float(v), "profiles/hbm_traffic_192x129x192.md") if v is not None else (None, None)


def make_channel(cf, c):
    import math
    pi = math.atan2(0.0, -1.0)
    lx = c["lx_over_pi"] * pi
    lz = lx / c["aspect"]
    ch = cf.Channel(c["nx"], c["ny"], c["nz"], c["nx2"], c["nz2"], lx, lz, c["re"], fix_mass=1, dt_flag=0)
    y, yh = ch.build_operators(c["stretch"])
    return ch, y, yh, lx, lz


def cpu_baseline(c, U0, nsteps=2):
    """Time the oracle (numpy, one core) on nsteps RK3 steps of the same workload."""
    sys.path.insert(0, ROOT)
    from oracle import chan_oracle as O
    g = O.make_grid(c["nx"], c["ny"], c["nz"], c["nx2"], c["nz2"], c["lx_over_pi"], c["aspect"], c["re"],
                    stretch=c["stretch"], fix_mass=1)
    st = O.initial_state(g, U0, dt=c["dt"])
    t0 = time.perf_counter()
    for s in range(1, nsteps + 1):
        O.rk3_step(g, st, s)
    t1 = time.perf_counter()
    return nsteps / (t1 - t0), t1 - t0


LES_CFG = dict(nx=64, ny=65, nz=64, nx2=96, nz2=96, dt=0.03)


def les_rate(cf, synth, c, steps):
    """RK3 steps/s of the 64 x 65 x 64 LES (same box and Reynolds number) with
    the trained learned closure and its dynamic-Smagorinsky gate, and the
    fraction of points that fell back at the last evaluation."""
    import math
    pi = math.atan2(0.0, -1.0)
    lx = c["lx_over_pi"] * pi
    lz = lx / c["aspect"]
    L = LES_CFG
    ch = cf.Channel(L["nx"], L["ny"], L["nz"], L["nx2"], L["nz2"], lx, lz, c["re"], fix_mass=1, dt_flag=0)
    y, yh = ch.build_operators(c["stretch"])
    params = np.load(os.path.join(PKG_DIR, "les_net_64x65x64.npy"), allow_pickle=False)
    net = cf.Mlp(params)
    U0 = synth.turbulent_like(L["nx"], L["ny"], L["nz"], yh, y, seed=c["seed"], rms=c["rms"], re_tau=c["re_tau"])
    ch.set_state(data1=U0)
    ch.set_scalars(dt=L["dt"], t_total=0.0, p_grad=-2.0 / c["re"])
    ch.les_configure("learned", net=net, beta=1.0, every_substep=False)
    ch.initialize_flow()
    ens = cf.Ensemble([ch], nstreams=1)   # the step replayed as one captured graph
    ens.step(5)
    ch.sync()
    t0 = time.perf_counter()
    ens.step(steps)
    ch.sync()
    t1 = time.perf_counter()
    g, n = ch.les_gated_points()
    ens.close()
    ch.close()
    net.close()
    return steps / (t1 - t0), g / n


def ensemble_rate(cf, synth, steps, n=48, seed=3):
    """The batched ensemble: n independent channels drawn from a seed
    (Nx, Nz in {32, 48, 64, 96}, Ny in {49, 65, 97}, Re_c uniform in
    [2000, 3500], box 2pi x 2 x pi, dt 0.005, laminar + 1e-2 perturbation),
    advanced together by one captured graph per step over 8 streams.
    Returns (ensemble steps/s, grid-point steps/s summed over the members)."""
    import math
    pi = math.atan2(0.0, -1.0)
    rng = np.random.default_rng(seed)
    mem, pts = [], 0
    for i in range(n):
        nx = int(rng.choice([32, 48, 64, 96]))
        nz = int(rng.choice([32, 48, 64, 96]))
        ny = int(rng.choice([49, 65, 97]))
        re = float(rng.uniform(2000.0, 3500.0))
        ch = cf.Channel(nx, ny, nz, 3 * nx // 2, 3 * nz // 2, 2.0 * pi, pi, re, fix_mass=1, dt_flag=0)
        y, yh = ch.build_operators(1.5)
        U0 = synth.laminar_perturbed(nx, ny, nz, yh, y, seed=seed * 1000 + i, amp=1e-2, kmax=4.0)
        ch.set_state(data1=U0)
        ch.set_scalars(dt=0.005, t_total=0.0, p_grad=-2.0 / re)
        ch.initialize_flow()
        mem.append(ch)
        pts += nx * ny * nz
    ens = cf.Ensemble(mem, nstreams=8)
    ens.step(3)
    for ch in mem:
        ch.sync()
    t0 = time.perf_counter()
    ens.step(steps)
    for ch in mem:
        ch.sync()
    t1 = time.perf_counter()
    ens.close()
    for ch in mem:
        ch.close()
    sps = steps / (t1 - t0)
    return sps, sps * pts, pts


def les_eval_results():
    """Errors of the committed filtered-DNS evaluation (les_eval.py)."""
    path = os.path.join(ROOT, "profiles", "les_eval_64x65x64.md")
    if not os.path.exists(path):
        return None
    txt = open(path).read()
    if "```json" not in txt:
        return None
    return json.loads(txt.split("```json", 1)[1].split("```", 1)[0])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=50,
                    help="steps of the 48-member batched ensemble timed (0: skip)")
    ap.add_argument("--les-steps", type=int, default=200,
                    help="RK3 steps of the learned-closure LES timed for the LES speed-up (0: skip)")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="only run warmup + this many steps (for rocprof), print no JSON")
    ap.add_argument("--les-profile-steps", type=int, default=0,
                    help="only run this many steps of the learned-closure LES (for rocprof), print no JSON")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run 5 device-to-device copies of this many MiB (a plain copy kernel "
                         "of known size, used to calibrate rocprof's FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_mib:
        import torch
        n = args.copy_mib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(f"copied {args.copy_mib} MiB x 5", flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                          "(the multi-GPU sharding is a design note, not implemented)"}))
        sys.exit(2)

    cf = load_pkg()
    import chanflow.synth as synth
    c = HEADLINE
    if args.les_profile_steps:
        sps, fb = les_rate(cf, synth, c, args.les_profile_steps)
        print(f"LES {args.les_profile_steps} steps, {sps:.1f} steps/s, fallback fraction {fb:.3f}", flush=True)
        return
    ch, y, yh, lx, lz = make_channel(cf, c)
    U0 = synth.turbulent_like(c["nx"], c["ny"], c["nz"], yh, y, seed=c["seed"], rms=c["rms"], re_tau=c["re_tau"])
    ch.set_state(data1=U0)
    ch.set_scalars(dt=c["dt"], t_total=0.0, p_grad=-2.0 / c["re"])
    ch.initialize_flow()
    ch.step(args.warmup, step0=1)
    ch.sync()
    if args.profile_steps:
        ch.step(args.profile_steps, step0=args.warmup + 1)
        ch.sync()
        print(f"profiled {args.profile_steps} steps", flush=True)
        return

    # ---- timed region: exactly K steps, synchronised on both sides
    ch.sync()
    t0 = time.perf_counter()
    ch.step(args.steps, step0=args.warmup + 1)
    ch.sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    steps_per_s = args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    div = ch.divergence_max()

    # ---- per-kernel durations with HIP events on the library's stream
    nmeas = max(5, min(20, args.steps // 10))
    ch.set_timing(True)
    ch.step(nmeas, step0=args.warmup + args.steps + 1)
    kt = ch.kernel_times()
    ch.set_timing(False)

    nyp = c["ny"] + 1
    nxh = c["nx"] // 2 + 1
    n_c = c["nz"] * nxh * nyp                      # complex coefficients per component
    alg_step = 960.0 * n_c                         # bytes per RK3 step (state traffic)
    per_launch = {k: (ms / cnt if cnt else 0.0) for k, (ms, cnt) in kt.items()}
    # dominant kernel: the longest launch; the x-pass moves only intermediates
    # (no algorithmic bytes), so when it dominates the NL pass is taken as a unit
    alg = {"z-inverse": 48.0 * n_c, "z-forward": 48.0 * n_c, "solve": 224.0 * n_c}
    dom = max(("z-inverse", "x-pass", "z-forward", "solve"), key=lambda k: per_launch[k])
    if dom == "x-pass":
        name = "nl_pass (z-inverse + x-pass + z-forward)"
        dur_ms = per_launch["z-inverse"] + per_launch["x-pass"] + per_launch["z-forward"]
        bytes_launch = 96.0 * n_c
    else:
        name, dur_ms, bytes_launch = dom, per_launch[dom], alg[dom]
    achieved = bytes_launch / (dur_ms * 1e-3) / 1e9
    step_gbs = alg_step * steps_per_s / 1e9
    traffic, traffic_src = measured_traffic(name)

    les = None
    if args.les_steps > 0:
        les_sps, fb = les_rate(cf, synth, c, args.les_steps)
        dns_tu = steps_per_s * c["dt"]
        les_tu = les_sps * LES_CFG["dt"]
        ev = les_eval_results()
        les = {"metric": "learned-LES speed-up (simulated time per wall second, LES / DNS)",
               "config": "LES 64x65x64 (96x96 dealiased), dt 0.03 (CFL about 0.7), learned closure with the dynamic-"
                         "Smagorinsky gate (beta 1), stress evaluated once per step, step replayed as a "
                         "captured graph; vs the 192x129x192 DNS above, same box and Re",
               "les_steps_per_s": les_sps, "les_time_units_per_s": les_tu, "dns_time_units_per_s": dns_tu,
               "speedup": les_tu / dns_tu, "fallback_fraction": fb,
               "error": None if ev is None else ev.get("learned_error"),
               "error_source": None if ev is None else "profiles/les_eval_64x65x64.md (les_eval.py, "
                                                       "a posteriori against the filtered DNS)"}

    ensemble = None
    if args.ensemble_steps > 0:
        e_sps, e_pts, npts = ensemble_rate(cf, synth, args.ensemble_steps)
        ensemble = {"metric": "RK3 steps/s of a 48-member batched ensemble (one captured graph per step)",
                    "config": "48 channels from seed 3: Nx, Nz in {32,48,64,96}, Ny in {49,65,97}, Re_c "
                              "uniform in [2000,3500], box 2pi x 2 x pi, 3/2 dealiased, dt 0.005, 8 streams",
                    "ensemble_steps_per_s": e_sps, "grid_point_steps_per_s": e_pts, "grid_points": npts,
                    "steps": args.ensemble_steps}

    cpu = None
    if not args.no_cpu_baseline:
        sps, secs = cpu_baseline(c, U0, args.cpu_steps)
        cpu = {"value": sps, "unit": "RK3 steps/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_steps} RK3 steps of the 192x129x192 workload by the numpy oracle "
                         f"({secs:.1f} s, single-threaded numpy FFT)"}

    out = {
        "metric": "FP64 RK3 steps/s at 192x129x192 (%HBM roofline)",
        "value": steps_per_s,
        "unit": "RK3 steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Reichardt profile + k^-5/3 solenoidal fluctuations)",
        "config": {"workload": "turbulent channel Re_c=3300 (Re_tau=180), box 4pi x 2 x 4pi/3, "
                               "192x129x192, 3/2 dealiased to 288x288, dt=0.005, constant mass flux",
                   "nx": 192, "ny": 129, "nz": 192, "nx2": 288, "nz2": 288, "dt": 0.005},
        "grid_point_steps_per_s": steps_per_s * c["nx"] * c["ny"] * c["nz"],
        "step_roofline": {"bound": "hbm", "achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": step_gbs / HBM_PEAK_GBS, "bytes_per_step": alg_step},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src, "bytes_per_launch": bytes_launch, "avg_launch_ms": dur_ms},
        "kernels_ms_per_launch": per_launch,
        "kernel_launches_measured": {k: v[1] for k, v in kt.items()},
        "max_divergence": div,
        "cpu_baseline": cpu,
        "les": les,
        "ensemble": ensemble,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
