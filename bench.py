"""Benchmark: FP64 RK3 steps/s of the channel step at 192 x 129 x 192.

Workload (the headline turbulent-channel configuration): Re_c = 3300
(Re_tau = 180), box 4pi x 2 x 4pi/3, 192 x 129 x 192 Fourier/grid points,
3/2-dealiased to 288 x 288, tanh-stretched y (2.75), dt = 0.005, constant
mass flux.  Synthetic seeded input: Reichardt mean profile plus solenoidal
k^-5/3 fluctuations of rms 0.1.  One "step" = boundary + three
(nonlinear, rhs_solve) substeps, all device resident.

Prints one JSON line.  `value` is RK3 steps/s of the whole job (one GPU).
`roofline` is computed for the dominant part of the step from HIP events
recorded on the library's stream; `fft_fp64` is the step's FFT flop rate
against the FP64 vector peak; `counters` quotes the committed rocprofv3
counter summary; `cpu_baseline` times the numpy oracle on a bounded sample
(3 steps) of the same workload on one host core.  Beside the headline: the
learned-closure LES speed-up with a live a-posteriori comparison, the
48-member batched ensemble, the reference's own driver over the library,
the jet-control run and the Arnoldi matvec of the linearized-stability
configuration (progress lines on stderr).
"""
import argparse
import math
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "3d-navier-stokes-thesis-paper_amd")

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 vector


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


def measured_counters():
    """FP64 VALU, VALU-busy, LDS-busy and MFMA figures of the step kernels and
    the closure network from the committed rocprofv3 counter summary (None
    when absent)."""
    path = os.path.join(ROOT, "profiles", "counters_192x129x192.md")
    if not os.path.exists(path):
        return None
    txt = open(path).read()
    if "```json" not in txt:
        return None
    data = json.loads(txt.split("```json", 1)[1].split("```", 1)[0])
    data["source"] = "profiles/counters_192x129x192.md"
    return data


def note(msg):
    """Progress line on stderr (the JSON line is the only output on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


LES_CFG = dict(nx=64, ny=65, nz=64, nx2=96, nz2=96, dt=0.01)


def les_rate(cf, synth, c, steps, dt=None):
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
    ch.set_scalars(dt=L["dt"] if dt is None else dt, t_total=0.0, p_grad=-2.0 / c["re"])
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
    advanced together by one captured graph per step whose every phase is one
    launch per transform length over all members, its launch plan chosen by
    chan_ensemble_autotune.  Returns (ensemble steps/s,
    grid-point steps/s summed over the members, grid points, algorithmic
    bytes per ensemble step = sum over members of 960 N_c)."""
    import math
    pi = math.atan2(0.0, -1.0)
    mem, pts, byts = [], 0, 0
    for d in synth.ensemble_draw(n, seed):
        nx, ny, nz, re = d["nx"], d["ny"], d["nz"], d["re"]
        ch = cf.Channel(nx, ny, nz, d["nx2"], d["nz2"], 2.0 * pi, pi, re, fix_mass=1, dt_flag=0)
        y, yh = ch.build_operators(1.5)
        U0 = synth.laminar_perturbed(nx, ny, nz, yh, y, seed=d["seed"], amp=1e-2, kmax=4.0)
        ch.set_state(data1=U0)
        ch.set_scalars(dt=0.005, t_total=0.0, p_grad=-2.0 / re)
        ch.initialize_flow()
        mem.append(ch)
        pts += nx * ny * nz
        byts += 960 * nz * (nx // 2 + 1) * (ny + 1)
    ens = cf.Ensemble(mem)
    ens.autotune(steps=3)            # the x-pass plane-range plan, every member's state restored
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
    return sps, sps * pts, pts, byts


def dropin_rate(cf, c, U0, steps=400, window=200):
    """The reference's own driver (PFC/main_mpi.f90 and its kept modules,
    built from the reference's sources against the replacement modules and
    libchanflow.so; absent unless built) on the headline workload: its own
    WTIME line over the last `window` steps of `steps` (each window includes
    the one host copy of the state the driver takes at its output step).
    The driver starts from its own dt rule, 0.25 lx/nx (PFC/setup.f90:1376),
    0.0164 at this box, which is above the CFL limit of this field, so the
    run uses dt_flag = 1: dt adapted to CFL 0.95 at every first substep
    (PFC/pass1.f90:202-250), which also makes every step wait for its CFL
    number, as the reference does.
    Returns (steps/s, seconds of the window), None when the driver is not
    built, or a dict naming the failure."""
    import resource
    import subprocess
    import tempfile
    exe = os.path.join(PKG_DIR, "fsrc", "_build", "ref", "mpiflow")
    if not os.path.exists(exe):
        return None
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import write_flow_dat
    with tempfile.TemporaryDirectory() as d:
        cf.savevel_write(os.path.join(d, "savevel.out"), U0, c["lx_over_pi"] * np.pi, c["lx_over_pi"] * np.pi /
                         c["aspect"], c["re"], c["dt"], 0.0)
        write_flow_dat(os.path.join(d, "flow.dat"), c["re"], c["lx_over_pi"], c["aspect"], c["nx"], c["ny"],
                       c["nz"], c["nx2"], c["nz2"], tend=steps, output_step=window, save_step=10 * steps,
                       l2_step=10 * steps, read_vel=1, save_vel=0, dt_flag=1)

        def big_stack():
            resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
        r = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=300, preexec_fn=big_stack)
    w = [float(l.split("=")[1].split()[0]) for l in r.stdout.splitlines() if l.strip().startswith("WTIME")]
    if r.returncode != 0 or not w:
        return {"error": f"driver exit status {r.returncode}, {len(w)} WTIME lines",
                "stdout_tail": r.stdout[-1500:], "stderr_tail": r.stderr[-1500:]}
    return window / w[-1], w[-1]


def arnoldi_matvec_rate(cf, nmv=20):
    """The linearized-stability configuration's Arnoldi matvec
    (chan_linearized_matvec, compute_time_deriv of PFC/arnoldi.f90:485-567:
    tend linearized RK3 steps about the laminar profile from a scaled start
    field, host field in and out): 64 x 97 x 64 (96 x 96 dealiased, box
    2pi x 2 x pi, stretch 2.75), Re_c 2800, dt = 0.95 (lx/nx2)/32 and
    tend = max(int(0.1/dt), 10) as init_baseflow and arnoldi.f90 set them
    (PFC/pass1_linearized.f90:46, PFC/arnoldi.f90:36).  Returns (matvecs/s,
    linearized steps/s, tend)."""
    pi = np.pi
    ch = cf.Channel(64, 97, 64, 96, 96, 2.0 * pi, pi, 2800.0, fix_mass=0, dt_flag=0)
    y, yh = ch.build_operators(2.75)
    base = np.zeros((3, 98, 96, 96))
    base[0] = (1.0 - np.asarray(yh) ** 2)[:, None, None]    # laminar branch of init_baseflow
    ch.set_baseflow(base)
    dt = 0.95 * (2.0 * pi / 96.0) / 32.0
    tend = max(int(0.1 / dt), 10)
    ch.set_scalars(dt=dt)
    v = 1.0e-6 * np.random.default_rng(9).standard_normal((3, 98, 64, 64))
    ch.linearized_matvec(v, tend)            # untimed
    t0 = time.perf_counter()
    for _ in range(nmv):
        w, _t = ch.linearized_matvec(v, tend)
    t1 = time.perf_counter()
    ch.close()
    mps = nmv / (t1 - t0)
    return mps, mps * tend, tend


def jet_control_rate(cf, synth, steps=1000):
    """The jet-control run of the linearized-stability configuration:
    64 x 97 x 64 (96 x 96 dealiased, box 2pi x 2 x pi, stretch 2.75), Re_c
    2800, a steady Gaussian jet (radius 0.1, peak 0.2) at the lower wall and
    the wall-shear feedback with gains 0.5, dt 0.002, from a perturbed laminar
    profile.  Returns (steps/s, max divergence, finite)."""
    pi = np.pi
    ch = cf.Channel(64, 97, 64, 96, 96, 2.0 * pi, pi, 2800.0, fix_mass=1, dt_flag=0, ctrlu=0.5, ctrlw=0.5)
    y, yh = ch.build_operators(2.75)
    U0 = synth.laminar_perturbed(64, 97, 64, yh, y, seed=23, amp=1e-2, kmax=4.0)
    ch.set_state(data1=U0)
    ch.set_scalars(dt=0.002, t_total=0.0, p_grad=-2.0 / 2800.0)
    ch.jet_configure(0, 0.2, 0.1, 0.25 * 2.0 * pi, 0.5 * pi, 0.0, 0.0)
    ch.initialize_flow()
    ch.step(5, step0=1)
    ch.sync()
    t0 = time.perf_counter()
    ch.step(steps, step0=6)
    ch.sync()
    t1 = time.perf_counter()
    div = ch.divergence_max()
    fin = bool(np.isfinite(ch.get_data1()).all())
    ch.close()
    return steps / (t1 - t0), div, fin


def les_aposteriori_live(cf, dns, spinup_tu, horizon_tu):
    """A-posteriori errors measured in this run: the headline DNS (after the
    timed steps) continued `spinup_tu` time units, then filtered on the device
    onto the 64 x 65 x 64 LES grid; LES runs at dt 0.01 (no model, dynamic
    Smagorinsky, the committed learned closure with its gate) against the DNS
    continued over `horizon_tu` time units (les_eval.aposteriori)."""
    import chanflow.les_eval as E
    dens = cf.Ensemble([dns])
    E.spinup(dns, dens, spinup_tu, spinup_tu, 1.0, E.DNS["dt"], log=lambda m: None)
    les, _, _ = E.channel(cf, E.LES)
    les.les_filter_from(dns, tau=False)
    Ul0 = les.get_data1()
    net = cf.Mlp(np.load(os.path.join(PKG_DIR, "les_net_64x65x64.npy"), allow_pickle=False))
    res = E.aposteriori(cf, dns, dens, les, net, Ul0, horizon_tu, 1.0, E.DNS["dt"], LES_CFG["dt"], 1.0,
                        log=lambda m: None)
    dens.close()
    les.close()
    net.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=50,
                    help="steps of the 48-member batched ensemble timed (0: skip)")
    ap.add_argument("--dropin-steps", type=int, default=400,
                    help="steps of the reference's own driver on the headline workload (0: skip)")
    ap.add_argument("--arnoldi-matvecs", type=int, default=20,
                    help="timed Arnoldi matvecs of the linearized-stability configuration (0: skip)")
    ap.add_argument("--jet-steps", type=int, default=1000,
                    help="steps of the 64x97x64 jet-control run (0: skip)")
    ap.add_argument("--les-spinup", type=float, default=40.0,
                    help="time units the DNS runs on before the live a-posteriori LES comparison")
    ap.add_argument("--les-horizon", type=float, default=60.0,
                    help="time units of the live a-posteriori comparison (0: skip)")
    ap.add_argument("--les-steps", type=int, default=200,
                    help="RK3 steps of the learned-closure LES timed for the LES speed-up (0: skip)")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="only run warmup + this many steps (for rocprof), print no JSON")
    ap.add_argument("--les-profile-steps", type=int, default=0,
                    help="only run this many steps of the learned-closure LES (for rocprof), print no JSON")
    ap.add_argument("--jet-profile-steps", type=int, default=0,
                    help="run only this many steps of the jet-control run (for a profiler) and exit")
    ap.add_argument("--ensemble-profile-steps", type=int, default=0,
                    help="only run this many steps of the 48-member batched ensemble (for rocprof), print no JSON")
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
        print(json.dumps({"error": "chanflow runs on one GPU only; --gpus must be 1 "
                          "(the multi-GPU sharding is a design note, not implemented)"}))
        sys.exit(2)

    cf = load_pkg()
    import chanflow.synth as synth
    c = HEADLINE
    if args.ensemble_profile_steps:
        e_sps, _, _, _ = ensemble_rate(cf, synth, args.ensemble_profile_steps)
        print(f"ensemble {args.ensemble_profile_steps} steps, {e_sps:.1f} steps/s", flush=True)
        return
    if args.les_profile_steps:
        sps, fb = les_rate(cf, synth, c, args.les_profile_steps)
        print(f"LES {args.les_profile_steps} steps, {sps:.1f} steps/s, fallback fraction {fb:.3f}", flush=True)
        return
    if args.jet_profile_steps:
        jsps, _, _ = jet_control_rate(cf, synth, args.jet_profile_steps)
        print(f"jet control {args.jet_profile_steps} steps, {jsps:.1f} steps/s", flush=True)
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
    note(f"headline: {args.warmup} warm-up steps done, timing {args.steps}")
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
    # FFT flops of a substep, the FP64 utilisation figure: 12 real x-transforms
    # of length nx2 per z-row and 9 complex z-transforms of length nz2 per kx
    # column, per plane, at 2.5 N log2 N (real) and 5 N log2 N (complex)
    nx2, nz2 = c["nx2"], c["nz2"]
    fft_sub = nyp * (12 * nz2 * 2.5 * nx2 * math.log2(nx2) + 9 * nxh * 5 * nz2 * math.log2(nz2))
    fft_tflops = 3 * fft_sub * steps_per_s / 1e12
    traffic, traffic_src = measured_traffic(name)

    note(f"headline: {steps_per_s:.1f} steps/s")
    les = None
    if args.les_steps > 0:
        note("LES speed at dt 0.01 and 0.03")
        les_sps, fb = les_rate(cf, synth, c, args.les_steps)
        les_sps3, _ = les_rate(cf, synth, c, args.les_steps, dt=0.03)
        dns_tu = steps_per_s * c["dt"]
        les_tu = les_sps * LES_CFG["dt"]
        les = {"metric": "learned-LES speed-up (simulated time per wall second, LES / DNS)",
               "config": "LES 64x65x64 (96x96 dealiased), dt 0.01, learned closure with the dynamic-"
                         "Smagorinsky gate (beta 1), stress evaluated once per step, step replayed as a "
                         "captured graph; vs the 192x129x192 DNS above, same box and Re",
               "les_steps_per_s": les_sps, "les_time_units_per_s": les_tu, "dns_time_units_per_s": dns_tu,
               "speedup": les_tu / dns_tu, "fallback_fraction": fb,
               "at_dt_0.03": {"les_steps_per_s": les_sps3, "speedup": les_sps3 * 0.03 / dns_tu}}
        if args.les_horizon > 0:
            note(f"LES a posteriori: {args.les_spinup:g} + {args.les_horizon:g} time units")
            ap_ = les_aposteriori_live(cf, ch, args.les_spinup, args.les_horizon)
            les["aposteriori"] = {
                "source": f"measured in this run: the DNS above continued {args.les_spinup:g} time units, then "
                          f"{args.les_horizon:g} time units of DNS and of each LES (dt 0.01) from the same "
                          f"filtered field, profiles sampled every time unit (les_eval.aposteriori)",
                "errors": {m: {k: v for k, v in ap_[m].items() if k.startswith("rel_err") or
                               k in ("reynolds_stress_max", "failed")}
                           for m in ("off", "dynamic", "learned") if m in ap_}}
            le = ap_.get("learned", {})
            if "rel_err_mean_profile" in le:
                les["error"] = {"mean_profile": le["rel_err_mean_profile"],
                                "reynolds_stress_max": le["reynolds_stress_max"]}

    ensemble = None
    if args.ensemble_steps > 0:
        note("48-member ensemble")
        e_sps, e_pts, npts, e_bytes = ensemble_rate(cf, synth, args.ensemble_steps)
        e_gbs = e_bytes * e_sps / 1e9
        ensemble = {"metric": "RK3 steps/s of a 48-member batched ensemble (one captured graph per step)",
                    "config": "48 channels from seed 3: Nx, Nz in {32,48,64,96}, Ny in {49,65,97}, Re_c "
                              "uniform in [2000,3500], box 2pi x 2 x pi, 3/2 dealiased, dt 0.005; every phase "
                              "of a substage one launch per transform length over all members; launch plan "
                              "autotuned (chan_ensemble_autotune)",
                    "ensemble_steps_per_s": e_sps, "grid_point_steps_per_s": e_pts, "grid_points": npts,
                    "roofline": {"bound": "hbm", "achieved": e_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": e_gbs / HBM_PEAK_GBS, "bytes_per_step": e_bytes},
                    "steps": args.ensemble_steps}

    dropin = None
    if args.dropin_steps > 0:
        note("the reference's own driver")
        dr = dropin_rate(cf, c, U0, args.dropin_steps, args.dropin_steps // 2)
        if isinstance(dr, dict):
            dropin = dr
        elif dr is not None:
            dropin = {"metric": "RK3 steps/s of the reference's own driver (main_mpi.f90 and its kept modules over "
                                "the replacement modules), by its WTIME line",
                      "config": "the headline field and grid, dt_flag 1 (CFL-adapted dt, the driver's fixed "
                                "initial dt 0.25 lx/nx = 0.0164 being above the CFL limit of this field)",
                      "steps_per_s": dr[0], "window_s": dr[1], "window_steps": args.dropin_steps // 2,
                      "note": "the window includes the one copy of data1/data2/data4 to the host that the "
                              "driver's output step takes"}
    jet = None
    if args.jet_steps > 0:
        note("jet-control run")
        jsps, jdiv, jfin = jet_control_rate(cf, synth, args.jet_steps)
        jet = {"config": "64x97x64, Re_c 2800, Gaussian jet radius 0.1 peak 0.2 at the lower wall, feedback "
                         "gains 0.5, dt 0.002", "steps": args.jet_steps, "steps_per_s": jsps,
               "max_divergence": jdiv, "finite": jfin}

    arn = None
    if args.arnoldi_matvecs > 0:
        note("Arnoldi matvecs")
        mps, lsps, tend = arnoldi_matvec_rate(cf, args.arnoldi_matvecs)
        arn = {"config": "64x97x64 linearized about the laminar profile, Re_c 2800, box 2pi x 2 x pi, "
                         f"dt 0.95 lx/nx2/32, tend {tend} steps per matvec (the Arnoldi operator of "
                         "arnoldi.f90), host field in and out", "matvecs": args.arnoldi_matvecs,
               "matvecs_per_s": mps, "linearized_steps_per_s": lsps}

    cpu = None
    if not args.no_cpu_baseline:
        note("CPU baseline (oracle)")
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
        "fft_fp64": {"flops_per_step": 3 * fft_sub, "achieved": fft_tflops, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": fft_tflops / FP64_PEAK_TFLOPS},
        "step_roofline": {"bound": "hbm", "achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": step_gbs / HBM_PEAK_GBS, "bytes_per_step": alg_step},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src, "bytes_per_launch": bytes_launch, "avg_launch_ms": dur_ms},
        "kernels_ms_per_launch": per_launch,
        "counters": measured_counters(),
        "kernel_launches_measured": {k: v[1] for k, v in kt.items()},
        "max_divergence": div,
        "cpu_baseline": cpu,
        "les": les,
        "ensemble": ensemble,
        "dropin": dropin,
        "jet_control": jet,
        "arnoldi": arn,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
