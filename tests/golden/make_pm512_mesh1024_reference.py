"""Writes pm512_mesh1024_fastpm20_fp64.npz: the CPU oracle's FP64 run of the precision-at-scale workload,
the reference against which the FP32 device run of the same workload is judged.

Workload: 512^3 particles on a 1024^3 mesh, box 1000 Mpc/h, TSC, interlacing (antialias = 1), gradient
and Laplace kernels of order 6, FastPM 20 steps from a = 0.05 to 1 in the time variable D; 2LPT initial
conditions from the N-GenIC white noise of seed 42 with the Eisenstein-Hu P(k) of Omega_c = 0.2621,
Omega_b = 0.049, h = 0.6766, sigma8 = 0.8102, n_s = 0.9665.

The oracle (oracle/, FP64, OpenMP) computes everything: white noise, 2LPT (compute_core on the 3n/2
grid), the force at the initial conditions, the 20 steps and the P(k) of the final positions (CIC on the
1024^3 mesh, window-deconvolved, linear bins of width kf from kf/2 to k_Nyq/2).  Stored (about 0.6 MB):
  k, pk, nmodes                   the final P(k) (255 bins)
  sub                             the subsample of particles kept: i, j, k all multiples of 32 (16^3)
  psi_ini, pi_ini                 the initial state on the subsample, (3, 16, 16, 16) [c, k, j, i]
  acc_ini                         the force at the initial conditions on the subsample
  psi_fin, p_fin                  the final displacement and canonical momentum on the subsample
  psi_rms_fin                     rms of the final displacement over all particles
Run it in two processes (the 2LPT stage alone needs ~45 GB, the run ~50 GB):
  python tests/golden/make_pm512_mesh1024_reference.py ics   WORKDIR
  python tests/golden/make_pm512_mesh1024_reference.py run   WORKDIR
It takes about an hour on 8 cores.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402

COSMO = (0.2621, 0.049, 0.6766, 0.8102, 0.9665)
n, N, L, SEED, A_INI, A_END, STEPS = 512, 1024, 1000.0, 42, 0.05, 1.0, 20
SUB = 32
OUT = os.path.join(HERE, "pm512_mesh1024_fastpm20_fp64.npz")


def log(msg):
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


def ics(work):
    oc = O.Cosmology(*COSMO)
    k, pk = O.linear_power_spectrum(oc)
    t = time.time()
    white = O.ngenic_wnoise_real(SEED, n)
    fphi = O.white_noise_to_fphi(O.ic_operator(n, L, k, pk), white)
    del white
    log(f"white noise and phi(k): {time.time() - t:.0f} s")
    psi1, psi2 = O.compute_core2(fphi, L)
    del fphi
    log(f"2LPT displacement fields: {time.time() - t:.0f} s")
    psi, pi = O.nbody_ics_lpt2(oc, psi1, psi2, A_INI)
    del psi1, psi2
    np.save(os.path.join(work, "psi_ini.npy"), psi)
    np.save(os.path.join(work, "pi_ini.npy"), pi)
    log(f"initial conditions written: {time.time() - t:.0f} s")


def run(work):
    oc = O.Cosmology(*COSMO)
    psi0 = np.load(os.path.join(work, "psi_ini.npy"), mmap_mode="r")
    pi0 = np.load(os.path.join(work, "pi_ini.npy"), mmap_mode="r")
    s = (slice(None),) + (slice(None, None, SUB),) * 3
    cfg = O.pm_config(N, n, L, worder=3, antialias=1, grad_order=6, lap_order=6)
    t = time.time()
    acc = O.pm_acceleration(psi0, cfg, "solver")
    acc_ini = np.ascontiguousarray(acc[s])
    del acc
    log(f"force at the initial conditions: {time.time() - t:.0f} s")
    t = time.time()
    psi, p = O.run_nbody(oc, psi0, pi0, cfg, A_INI, A_END, STEPS, "fastpm", "D",
                         callback=lambda k, a: log(f"step {k} done (a = {a:.4f}), {time.time() - t:.0f} s"))
    t_run = time.time() - t
    X = O.lagrangian_positions(psi, L)
    mesh = O.pm_scatter(X, N, L, 2)
    del X
    mesh *= N**3 / n**3
    mesh -= 1.0
    kk, pk, nm = O.power_spectrum_lin(mesh, L, W=2, nbins=N // 4 - 1)
    del mesh
    np.savez_compressed(OUT, k=kk, pk=pk, nmodes=nm, sub=np.int64(SUB),
                        psi_ini=np.ascontiguousarray(psi0[s]), pi_ini=np.ascontiguousarray(pi0[s]),
                        acc_ini=acc_ini, psi_fin=np.ascontiguousarray(psi[s]), p_fin=np.ascontiguousarray(p[s]),
                        psi_rms_fin=np.sqrt(np.mean(psi.astype(np.float64) ** 2)),
                        run_seconds=t_run, threads=np.int64(int(os.environ.get("OMP_NUM_THREADS", "0"))))
    log(f"written {OUT}; the 20 steps took {t_run:.0f} s")


if __name__ == "__main__":
    stage, work = sys.argv[1], sys.argv[2]
    os.makedirs(work, exist_ok=True)
    {"ics": ics, "run": run}[stage](work)
