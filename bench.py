#!/usr/bin/env python3
"""Throughput of the PM N-body step on one MI355X: particle-updates/s at 512^3 particles on a 1024^3
mesh (box 1000 Mpc/h, TSC with interlacing, 6th-order gradient and Laplace kernels, FastPM 20 steps
from z = 19 to z = 0), FP32 mesh / FFT / forces with FP64 step coefficients.

A step is one FastPM DKD step over all particles (two interlaced force passes).  The initial
conditions are synthetic: the N-GenIC white noise of seed 42, the Eisenstein-Hu P(k) of the
workload's cosmology, 2LPT at z = 19, generated on the device (outside the timed region).  The
particle state is resident in HBM when the timed region starts.

Reported beside the throughput:
  timing         median of 3 runs of the 20 steps from the initial conditions (state reset between
                 runs), after a warm-up run of --warmup steps; --steps K < 20 times the first K steps
  accuracy       max |P32(k)/P64(k) - 1| for kf/2 <= k < k_Nyq/2 and rms(Psi32 - Psi64)/(L/n) of the
                 FP32 run against the CPU oracle's FP64 run of the same workload from the same ICs
                 (tests/golden/pm512_mesh1024_fastpm20_fp64.npz), and the force at the ICs against it
  roofline       the dominant kernel's algorithmic bytes per launch / its mean launch time (HIP
                 events on the library's stream, in one more run of the same steps) against 8 TB/s
  cpu_baseline   the CPU oracle (the C restatement of the reference path, FP32, OpenMP) on the host
                 cores: the force of one step (two interlacing passes) at the ICs; the GPU force at
                 the ICs is compared with it
  extra lines    256^3 particles on a 256^3 mesh, BullFrog 10 (particle-updates/s) and the density-loss
                 gradient at 256^3 (gradients/s)

usage: python bench.py [--gpus 1] [--steps 20] [--warmup 2] [--no-accuracy] [--no-cpu] [--no-extra] [--small]
       python bench.py --workload gradient|learned|studies  (the differentiation, learned-correction and
                                                        discreteness-suppressed workloads; one JSON line each)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "discodj.jl_amd"))

HBM_PEAK_GBS = 8000.0
COSMO = dict(Omega_c=0.2621, Omega_b=0.049, h=0.6766, sigma8=0.8102, n_s=0.9665)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the 256^3 throughput and gradient lines")
    ap.add_argument("--small", action="store_true", help="128^3 particles on a 256^3 mesh (quick check)")
    ap.add_argument("--workload", choices=["pm", "gradient", "learned", "studies"], default="pm")
    ap.add_argument("--copy-rate", action="store_true", help="only time a 1 GiB read + 1 GiB write copy kernel")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy kernel three times (the reference point that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters of a profiling run to bytes)")
    args = ap.parse_args()
    if args.copy_rate:
        import torch

        a = torch.ones(2**28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.neg(a, out=b)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(20):
            torch.neg(a, out=b)
        e1.record()
        e1.synchronize()
        print(json.dumps({"copy_GBs": round(20 * 2 * 2**30 / (e0.elapsed_time(e1) * 1e-3) / 1e9, 1)}), flush=True)
        return 0
    if args.copy_calibration:
        import torch

        a = torch.ones(2**28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.neg(a, out=b)  # one elementwise kernel: reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        log("copy calibration: 3 launches of a 1 GiB read + 1 GiB write elementwise kernel")
        return 0
    if args.workload == "gradient":
        return bench_gradient(args)
    if args.workload == "learned":
        return bench_learned(args)
    if args.workload == "studies":
        return bench_studies(args)
    if args.gpus != 1:
        print(f"bench.py measures one GPU (one MI355X); --gpus {args.gpus} is not run: the multi-GPU slab/pencil "
              f"decomposition of the mesh is designed, not implemented", file=sys.stderr)
        return 2

    import numpy as np

    import discodj_amd as D

    metric = "PM particle-updates/s at 512³ on 1024³ mesh, ≤1e-3 P(k) err vs FP64; %HBM roofline"
    n, N, L = (128, 256, 250.0) if args.small else (512, 1024, 1000.0)
    a_ini, a_end, seed, n_steps = 0.05, 1.0, 42, 20
    W, K = args.warmup, args.steps
    if not 1 <= K <= n_steps or W < 0:
        print(f"--steps must be in 1..{n_steps} (the workload is the {n_steps}-step FastPM schedule from z = 19 "
              f"to 0; a timed run is its first K steps from the initial conditions) and --warmup >= 0",
              file=sys.stderr)
        return 2
    cfg = D.PMConfig(res_pm=N, n_part=n, boxsize=L, worder=3, antialias=1, grad_order=6, lap_order=6)
    c = D.compute_timetables(D.Cosmology(**COSMO))
    k_tab, pk_tab = D.linear_power_spectrum(c)
    co = D.stepper_coefficients(c, "fastpm", a_ini, a_end, n_steps, "D")
    Dg = c.Dplus(a_ini)
    coef = [co["ddrift1"], co["alpha"], co["beta"], co["ddrift2"]]

    log(f"workload {n}^3 particles, {N}^3 mesh, L={L}, FastPM {n_steps} steps; ICs on the device")
    S = D.PMSolverHIP(cfg, precision=32)
    t = time.perf_counter()
    S.ics_lpt(k_tab, pk_tab, Dg, order=2, seed=seed)
    S.synchronize()
    log(f"2LPT ICs in {time.perf_counter() - t:.1f}s")
    ic_psi, ic_pi = S.get_state()

    # warm-up: the first W steps of the schedule, then back to the initial conditions; each timed run
    # steps the schedule's first K steps (all of them by default) from the initial conditions
    if W:
        S.run(*(a[:W] for a in coef))
        S.synchronize()
    times = []
    for rep in range(3):
        S.set_state(ic_psi, ic_pi)
        S.synchronize()
        t0 = time.perf_counter()
        S.run(*(a[:K] for a in coef))
        S.synchronize()
        times.append(time.perf_counter() - t0)
        log(f"timed run {rep}: {K} steps in {times[-1]:.4f} s ({times[-1] / K * 1e3:.2f} ms/step)")
    dt = float(np.median(times))
    value = n**3 * K / dt
    log(f"median {dt / K * 1e3:.2f} ms/step, {value:.3e} particle-updates/s")
    psi32 = S.get_state()[0] if K == n_steps else None

    # per-kernel times: one more run of the same steps with HIP events around every launch (on the
    # library's stream), outside the timed runs
    S.set_state(ic_psi, ic_pi)
    S.set_profiling(True)
    S.reset_stats()
    S.run(*(a[:K] for a in coef))
    S.synchronize()
    stats = S.kernel_stats()
    S.set_profiling(False)
    kernels = {}
    for name, s in stats.items():
        if s["calls"]:
            ms = s["ms_total"] / s["calls"]
            gbs = s["bytes_per_call"] / (ms * 1e-3) / 1e9
            kernels[name] = {"calls": s["calls"], "ms_per_call": round(ms, 4), "ms_total": round(s["ms_total"], 3),
                             "algorithmic_GB": round(s["bytes_per_call"] / 1e9, 4), "achieved_GBs": round(gbs, 1),
                             "frac": round(gbs / HBM_PEAK_GBS, 4)}
    if "scatter" in kernels:
        # unit-mass atomic adds of one deposit: W^3 per particle (TSC: 27), as the reference's scatter makes
        kernels["scatter"]["atomic_adds_per_s"] = 27 * n**3 / (kernels["scatter"]["ms_per_call"] * 1e-3)
    dom = max(kernels, key=lambda k: kernels[k]["ms_total"])
    roof = {"kernel": dom, "bound": "hbm", "achieved": kernels[dom]["achieved_GBs"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": kernels[dom]["frac"], "traffic": None,
            "algorithmic_bytes_per_launch": int(kernels[dom]["algorithmic_GB"] * 1e9)}
    # HBM bytes per launch of the same kernel: not measured in this run, but taken from rocprofv3
    # FETCH_SIZE / WRITE_SIZE passes over this workload, scaled by a copy kernel of known size
    tpath = os.path.join(ROOT, "profiles", "pm_step_c3_traffic.json")
    if not args.small and os.path.exists(tpath):
        tk = json.load(open(tpath))["kernels"].get(dom)
        if tk:
            roof["traffic"] = int(tk["traffic_bytes"])
            roof["traffic_source"] = ("rocprofv3 counters of a separate profiling run of this workload, "
                                      "profiles/pm_step_c3_counters.md (" + tk["rocprof_kernel"] + ")")

    # a plain streaming kernel of known size (1 GiB read + 1 GiB written, torch.neg) on the same device,
    # in a child process (torch brings its own HIP runtime): the bandwidth a kernel of this chip reaches
    # without any computation, beside the 8 TB/s peak
    import subprocess

    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-rate"], capture_output=True, text=True,
                       timeout=300)
    if r.returncode == 0 and r.stdout.strip():
        roof["copy_GBs"] = json.loads(r.stdout.strip().splitlines()[-1])["copy_GBs"]
        roof["frac_of_copy"] = round(roof["achieved"] / roof["copy_GBs"], 4)
        log(f"streaming copy {roof['copy_GBs']} GB/s")

    accuracy = None
    acc_ini32 = None
    if not args.no_accuracy:
        acc_ini32 = S.pm_acceleration(ic_psi)
        accuracy = accuracy_vs_oracle(np, S, psi32, acc_ini32, c, n, N, L) if not args.small else None
    S.close()

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        ocfg = O.pm_config(res_pm=N, n_part=n, boxsize=L, worder=3, antialias=1, grad_order=6, lap_order=6)
        log(f"CPU oracle: the FP32 force of the workload (two interlacing passes) at the ICs on {threads} threads")
        t = time.perf_counter()
        acc_cpu = O.pm_acceleration(ic_psi, ocfg, "solver")
        t_force = time.perf_counter() - t
        cpu = {"value": n**3 / t_force, "unit": "particle-updates/s", "cores": threads, "kind": "port",
               "sample": f"the force of one step (two interlacing passes: deposit, FFT Poisson solve, FD "
                         f"gradient, gather) of the {n}^3 / {N}^3 FP32 workload at the z=19 ICs, {t_force:.1f} s; "
                         f"a step's kick/drift is not counted"}
        log(f"CPU force {t_force:.1f}s -> {cpu['value']:.3e} particle-updates/s")
        if acc_ini32 is not None:
            e = float(np.max(np.abs(acc_ini32.astype(np.float64) - acc_cpu)) / np.max(np.abs(acc_cpu)))
            cpu["gpu_force_vs_this_force"] = e
            if accuracy is not None:
                accuracy["force_ini_fp32_vs_oracle_fp32"] = e
            log(f"GPU FP32 force at the ICs vs the oracle's FP32 force: max rel err {e:.2e}")
        del acc_cpu

    extra = {}
    if not args.small and not args.no_extra:
        extra["pm256_mesh256_bullfrog10"] = bench_pm256(np, D, c, k_tab, pk_tab)
        extra["gradient256"] = bench_gradient(args, as_dict=True)

    out = {"metric": metric, "value": value, "unit": "particle-updates/s", "n_gpus": 1, "steps": K, "warmup": W,
           "ms_per_step": dt / K * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"PM N-body, {n}^3 particles on a {N}^3 mesh, L={L} Mpc/h, TSC, interlacing "
                                  f"(2 passes), grad/Laplace order 6, FastPM {n_steps} steps z=19->0, 2LPT ICs "
                                  f"(N-GenIC seed {seed}), FP32 state and mesh, FP64 coefficients",
                      "particles": n**3, "mesh": N**3, "boxsize": L, "stepper": "fastpm", "n_steps": n_steps,
                      "timed_steps": K, "timing": f"median of 3 runs of the first {K} steps from the ICs, after "
                                                  f"a {W}-step warm-up run",
                      "run_seconds": [round(x, 5) for x in times]},
           "roofline": roof, "kernels": kernels, "accuracy": accuracy, "cpu_baseline": cpu}
    out.update(extra)
    print(json.dumps(out), flush=True)
    return 0


REF512 = os.path.join(ROOT, "tests", "golden", "pm512_mesh1024_fastpm20_fp64.npz")


def accuracy_vs_oracle(np, S, psi32, acc_ini32, c, n, N, L):
    """The FP32 device run against the CPU oracle's FP64 run of the same workload (same ICs: seed 42, 2LPT),
    stored by tests/golden/make_pm512_mesh1024_reference.py: P(k) of the final positions (CIC, deconvolved,
    255 linear bins kf/2 .. k_Nyq/2), the displacement and the force at the ICs on the stored subsample."""
    if psi32 is None or not os.path.exists(REF512):
        log("no oracle reference for the accuracy (partial run or missing fixture)")
        return None
    ref = np.load(REF512)
    s = int(ref["sub"])
    sub = (slice(None),) + (slice(None, None, s),) * 3
    _, p32, nm = S.power_spectrum(psi=psi32, W=2)
    ok = ref["nmodes"] > 0
    rel = np.abs(p32[ok] / ref["pk"][ok] - 1)
    disp = float(np.sqrt(np.mean((psi32[sub].astype(np.float64) - ref["psi_fin"]) ** 2)) / (L / n))
    e_acc = float(np.max(np.abs(acc_ini32[sub].astype(np.float64) - ref["acc_ini"])) / np.max(np.abs(ref["acc_ini"])))
    log(f"vs the oracle's FP64 run: P(k) max rel err {rel.max():.2e}, displacement rms / spacing {disp:.2e}, "
        f"force at the ICs {e_acc:.2e}")
    return {"pk_max_rel_err": float(rel.max()), "pk_bins": int(ok.sum()),
            "pk_k_range": [float(ref["k"][ok][0]), float(ref["k"][ok][-1])],
            "disp_rms_over_spacing": disp, "disp_sample": f"{(n // s) ** 3} particles (every {s}th along each axis)",
            "force_ini_max_rel_err": e_acc,
            "reference": "CPU oracle (FP64) run of the same workload from the same ICs, "
                         "tests/golden/pm512_mesh1024_fastpm20_fp64.npz"}


def bench_pm256(np, D, c, k_tab, pk_tab):
    """256^3 particles on a 256^3 mesh, L = 500, CIC, interlacing, gradient order 4, BullFrog 10 (FP32):
    particle-updates/s over the 10 steps from the 2LPT ICs of seed 42, median of 3 after a warm-up run."""
    n, L = 256, 500.0
    cfg = D.PMConfig(res_pm=n, n_part=n, boxsize=L, worder=2, antialias=1, grad_order=4)
    co = D.stepper_coefficients(c, "bullfrog", 0.05, 1.0, 10, "D")
    coef = [co["ddrift1"], co["alpha"], co["beta"], co["ddrift2"]]
    with D.PMSolverHIP(cfg, precision=32) as S:
        S.ics_lpt(k_tab, pk_tab, c.Dplus(0.05), order=2, seed=42)
        p0, m0 = S.get_state()
        times = []
        for rep in range(4):
            S.set_state(p0, m0)
            S.synchronize()
            t0 = time.perf_counter()
            S.run(*coef)
            S.synchronize()
            if rep:
                times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    log(f"256^3 / 256^3 BullFrog 10: {dt / 10 * 1e3:.3f} ms/step")
    return {"value": n**3 * 10 / dt, "unit": "particle-updates/s", "ms_per_step": dt / 10 * 1e3,
            "workload": "256^3 particles, 256^3 mesh, L=500, CIC, interlacing, grad order 4, BullFrog 10, FP32",
            "timing": "median of 3 runs of the 10 steps from the ICs, after a warm-up run"}


def bench_gradient(args, as_dict=False):
    """Gradient of sum (delta - target)^2 on the 256^3 mesh with respect to the 256^3 white noise through
    2LPT and 10 BullFrog CIC steps (grad order 4), FP32: seconds per gradient, split into the forward run
    with step checkpoints and the reverse pass (as_dict: returned instead of printed).  The white noise, the
    target and the white-noise cotangent stay in device memory and the loss is formed on the device
    (dj_density_loss_device); the same gradient with host arrays is timed beside it."""
    import numpy as np

    import discodj_amd as D

    n, L = (64, 125.0) if args.small else (256, 500.0)
    c = D.compute_timetables(D.Cosmology(**COSMO))
    k_tab, pk_tab = D.linear_power_spectrum(c)
    white = np.random.default_rng(42).normal(size=(n, n, n))
    target = np.zeros((n, n, n))
    co = D.stepper_coefficients(c, "bullfrog", 0.05, 1.0, 10, "D")
    cfg = D.PMConfig(res_pm=n, n_part=n, boxsize=L, worder=2, grad_order=4)
    rho = 1.0
    with D.PMSolverHIP(cfg, precision=32) as S:
        w_d = D.DeviceArray.from_numpy(white)
        t_d = D.DeviceArray.from_numpy(target)
        wb_d = D.DeviceArray(white.shape, np.float64)
        times = []
        for rep in range(4):  # the first is a warm-up (it allocates the IC and reverse-pass buffers)
            S.synchronize()
            t0 = time.perf_counter()
            S.ics_lpt_device(k_tab, pk_tab, c.Dplus(0.05), w_d, order=2)
            S.run_checkpointed(co["ddrift1"], co["alpha"], co["beta"], co["ddrift2"])
            loss = S.density_loss_device(t_d)
            t1 = time.perf_counter()
            S.vjp_run()
            S.vjp_ics_lpt_device(k_tab, pk_tab, c.Dplus(0.05), w_d, wb_d, order=2)
            S.synchronize()
            t2 = time.perf_counter()
            if rep:
                times.append((t1 - t0, t2 - t1))
            log(f"gradient {rep}: forward {t1 - t0:.3f}s, reverse {t2 - t1:.3f}s, loss {loss:.6e}")
        log(f"|wbar| {np.linalg.norm(wb_d.numpy()):.4e}")
        # the same gradient with host arrays (white noise up, mesh down, loss on the host, cotangent up,
        # white-noise cotangent down)
        th = time.perf_counter()
        S.ics_lpt(k_tab, pk_tab, c.Dplus(0.05), order=2, white=white)
        S.run_checkpointed(co["ddrift1"], co["alpha"], co["beta"], co["ddrift2"])
        delta = S.deposit().astype(np.float64) / rho - 1.0
        S.vjp_deposit(None, 2.0 * (delta - target) / rho, resident=True)
        S.vjp_run()
        wb_h = S.vjp_ics_lpt(k_tab, pk_tab, c.Dplus(0.05), order=2, white=white)
        S.synchronize()
        t_host = time.perf_counter() - th
        log(f"host arrays: {t_host:.3f}s per gradient, |wbar| {np.linalg.norm(wb_h):.4e}")
        for a in (w_d, t_d, wb_d):
            a.free()
    fw, bw = np.median([t[0] for t in times]), np.median([t[1] for t in times])
    # the same gradient at 512^3 particles on a 512^3 mesh (L = 1000): checkpoints and recompute keep it in
    # HBM (device arrays, FP32; one warm-up, median of 2)
    g512 = None
    if not args.small:
        n5, L5 = 512, 1000.0
        cfg5 = D.PMConfig(res_pm=n5, n_part=n5, boxsize=L5, worder=2, grad_order=4)
        w5 = D.DeviceArray.from_numpy(np.random.default_rng(43).normal(size=(n5, n5, n5)))
        wb5 = D.DeviceArray((n5, n5, n5), np.float64)
        t5 = []
        with D.PMSolverHIP(cfg5, precision=32) as S:
            for rep in range(3):
                S.synchronize()
                t0 = time.perf_counter()
                S.ics_lpt_device(k_tab, pk_tab, c.Dplus(0.05), w5, order=2)
                S.run_checkpointed(co["ddrift1"], co["alpha"], co["beta"], co["ddrift2"])
                S.density_loss_device(None)
                S.vjp_run()
                S.vjp_ics_lpt_device(k_tab, pk_tab, c.Dplus(0.05), w5, wb5, order=2)
                S.synchronize()
                if rep:
                    t5.append(time.perf_counter() - t0)
                log(f"512^3 gradient {rep}: {time.perf_counter() - t0:.3f}s")
        w5.free()
        wb5.free()
        g512 = {"gradients_per_s": 1.0 / float(np.median(t5)), "s_per_gradient": float(np.median(t5)),
                "workload": f"{n5}^3 particles, {n5}^3 mesh, L={L5}, CIC, grad order 4, BullFrog 10, FP32"}
    out = {"metric": "gradients/s of the density loss w.r.t. the white noise (2LPT + 10 BullFrog steps)",
           "value": 1.0 / (fw + bw), "unit": "gradients/s", "forward_s": fw, "reverse_s": bw,
           "host_arrays_s_per_gradient": t_host, "at_512": g512, "dtype": "f32", "data": "synthetic", "n_gpus": 1,
           "config": {"workload": f"{n}^3 particles, {n}^3 mesh, L={L}, CIC, grad order 4, BullFrog 10",
                      "median_of": 3, "arrays": "white noise, target, loss and cotangents in device memory"}}
    if as_dict:
        return out
    print(json.dumps(out), flush=True)
    return 0


def bench_studies(args):
    """The discreteness-suppressed PM of the reference's shell-crossing study (studies/eft_shell_crossing/
    config/fiducial.toml:17-30): 256^3 particles on a 512^3 mesh, L = 500 Mpc/h, PCS with window
    deconvolution, spectral ik gradient, interlacing, 2^3 Fourier sheet resampling, BullFrog 200 steps uniform
    in D from D = 0.02 to 1.  ICs: 2LPT of the N-GenIC seed 42 field (the study starts from 3LPT), FP32.
    Seconds per step over the 200 steps (median of 3 runs from the ICs), beside the reference's 6.1 min per
    job of 200 steps on an RTX A6000 (about 1.8 s per step, diagnostics included)."""
    import numpy as np

    import discodj_amd as D

    n, N, L, steps = (64, 128, 125.0, 20) if args.small else (256, 512, 500.0, 200)
    c = D.compute_timetables(D.Cosmology(**COSMO))
    k_tab, pk_tab = D.linear_power_spectrum(c)
    a_ini = c.a_of_Dplus(0.02)
    co = D.stepper_coefficients(c, "bullfrog", a_ini, 1.0, steps, "D")
    coef = [co["ddrift1"], co["alpha"], co["beta"], co["ddrift2"]]
    cfg = D.PMConfig(res_pm=N, n_part=n, boxsize=L, worder=4, deconvolve=True, antialias=1, grad_order=0,
                     n_resample=2, resampling="fourier")
    with D.PMSolverHIP(cfg, precision=32) as S:
        S.ics_lpt(k_tab, pk_tab, c.Dplus(a_ini), order=2, seed=42)
        p0, m0 = S.get_state()
        S.run(*(a[:2] for a in coef))  # warm-up
        times = []
        for rep in range(3):
            S.set_state(p0, m0)
            S.synchronize()
            t0 = time.perf_counter()
            S.run(*coef)
            S.synchronize()
            times.append(time.perf_counter() - t0)
            log(f"studies run {rep}: {steps} steps in {times[-1]:.3f} s")
        S.set_state(p0, m0)
        S.set_profiling(True)
        S.reset_stats()
        S.run(*(a[:20] for a in coef))
        S.synchronize()
        stats = S.kernel_stats()
        S.set_profiling(False)
    dt = float(np.median(times)) / steps
    kernels = {k: {"calls": v["calls"], "ms_per_call": round(v["ms_total"] / v["calls"], 4)}
               for k, v in stats.items() if v["calls"]}
    print(json.dumps({"metric": "seconds per step of the discreteness-suppressed PM (shell-crossing study settings)",
                      "value": dt, "unit": "s/step", "higher_is_better": False,
                      "particle_updates_per_s": n**3 / dt, "reference": {"value": 1.8, "unit": "s/step",
                      "hardware": "1x RTX A6000, diagnostics included (studies/eft_shell_crossing/README.md:1023)"},
                      "kernels_first_20_steps": kernels, "dtype": "f32", "data": "synthetic", "n_gpus": 1,
                      "config": {"workload": f"{n}^3 particles, {N}^3 mesh, L={L}, PCS, deconvolution, ik gradient, "
                                             f"interlacing, 2^3 Fourier sheet resampling, BullFrog {steps} steps "
                                             f"D = 0.02 -> 1, 2LPT ICs", "timing": "median of 3 full runs"}}),
          flush=True)
    return 0


def bench_learned(args):
    """The learned correction at 256^3 particles on a 256^3 mesh, BullFrog 10 steps, against the 1024^3-mesh
    run (4x mesh): the transfer function of the Poisson kernel fitted on two seeds (1000, 1001) on top of
    window deconvolution, its P(k) error on a held-out seed (42) and its step time relative to plain PM;
    and the step time of the residual conv net (untrained weights: the kernels' cost does not depend on
    their values)."""
    import numpy as np

    import discodj_amd as D

    n, L = (64, 125.0) if args.small else (256, 500.0)
    c = D.compute_timetables(D.Cosmology(**COSMO))
    work = D.learned.Workload(c, n=n, N=n, L=L, ratio=4, n_steps=10, deconvolve=True)
    plain_work = D.learned.Workload(c, n=n, N=n, L=L, ratio=4, n_steps=10)
    fit_seeds, eval_seed = [1000, 1001], 42
    t = time.perf_counter()
    fine = D.learned.fine_states(work, fit_seeds + [eval_seed])
    log(f"fine-mesh runs ({4 * n}^3 mesh): {time.perf_counter() - t:.1f} s")
    t = time.perf_counter()
    g, kmax = D.learned.fit_transfer(work, fit_seeds, psi_fine=fine[:2], log=log)
    t_fit = time.perf_counter() - t
    pk_plain, _, _ = D.learned.power_spectrum_errors(plain_work, eval_seed, psi_fine=fine[2])
    pk_dec, pk_corr, _ = D.learned.power_spectrum_errors(work, eval_seed, kgain=(g, kmax), psi_fine=fine[2])
    log(f"held-out P(k) max rel err to k_Nyq/2: plain {pk_plain:.4f}, deconvolved {pk_dec:.4f}, "
        f"corrected {pk_corr:.4f}")

    def timed(S, setup):
        best = None
        for _ in range(2):
            setup(S)
            S.ics_lpt(work.k_tab, work.pk_tab, work.D_ini, order=2, seed=eval_seed)
            S.synchronize()
            S.set_profiling(True)
            S.reset_stats()
            t0 = time.perf_counter()
            S.run(*work.coef)
            S.synchronize()
            dt = (time.perf_counter() - t0) / len(work.coef[0])
            stats = S.kernel_stats()
            S.set_profiling(False)
            best = dt if best is None else min(best, dt)
        return best, stats

    with D.PMSolverHIP(plain_work.cfg_lo, precision=32) as S:
        t_plain, _ = timed(S, lambda S: None)
        S.nn_set_params(D.learned.init_params(0))
        S.nn_set_schedule(work.a_force)

        def net_on(S):
            S.nn_enable(True, 1.0)
        t_net, stats = timed(S, net_on)
    with D.PMSolverHIP(work.cfg_lo, precision=32) as S:
        t_corr, _ = timed(S, lambda S: S.set_kgain(g, kmax))
    log(f"ms/step: plain {t_plain * 1e3:.3f}, corrected {t_corr * 1e3:.3f}, with the conv net {t_net * 1e3:.3f}")
    nn = stats.get("nn_correction", {})
    nn_ms = nn["ms_total"] / nn["calls"] if nn.get("calls") else None
    # MFMA work per 16 voxels: 27 16x16x16 products in each of the two 16-channel layers, 4 in the
    # packed two-channel first layer and in the one-channel last layer (8 taps x 2 channels each)
    flops = (2 * 27 + 2 * 4) * 2 * 16 * 16 * 16 * (n**3 / 16)
    print(json.dumps({"metric": "step time of the learned-correction PM relative to plain PM",
                      "value": t_corr / t_plain, "unit": "ratio", "higher_is_better": False,
                      "ms_per_step_plain": t_plain * 1e3, "ms_per_step_corrected": t_corr * 1e3,
                      "pk_max_rel_err": {"plain": pk_plain, "deconvolved": pk_dec, "corrected": pk_corr,
                                         "k_range": "kf to k_Nyq/2", "reference": f"same seed on a {4 * n}^3 mesh",
                                         "fit_seeds": fit_seeds, "eval_seed": eval_seed},
                      "fit_seconds": t_fit,
                      "conv_net": {"ms_per_step": t_net * 1e3, "ratio": t_net / t_plain, "nn_ms_per_call": nn_ms,
                                   "nn_mfma_tflops": (flops / (nn_ms * 1e-3) / 1e12) if nn_ms else None},
                      "dtype": "f32 PM, f64 fit; bf16 conv net", "data": "synthetic", "n_gpus": 1,
                      "config": {"workload": f"{n}^3 particles, {n}^3 mesh (fine: {4 * n}^3), L={L}, CIC, grad "
                                             f"order 4, BullFrog 10, correction: window deconvolution + fitted "
                                             f"transfer function"}}),
          flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
