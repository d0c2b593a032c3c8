"""The C-ABI library without a device: it loads, exports every function include/discodj_hip.h
declares, rejects the configurations the reference rejects, and its host-side pieces (N-GenIC noise,
FP64 cosmology and stepper coefficients) match the oracle and the golden vectors."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import golden as G
import oracle as O
import discodj_amd as D
from discodj_amd import _lib


def test_library_exports_every_header_symbol():
    names = D.header_exports()
    assert len(names) >= 20
    lib = C.CDLL(D.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert D.lib().dj_version().startswith(b"discodj_hip")


@pytest.mark.parametrize("kw, code", [
    (dict(worder=5), -1), (dict(grad_order=3), -1), (dict(lap_order=1), -1), (dict(antialias=4), -1),
    (dict(antialias=1, lean=1), -1), (dict(grad_order=0, lean=1), -1), (dict(precision=16), -1),
    (dict(n_resample=5), -3), (dict(resampling=2), -1), (dict(n_resample=0), -1), (dict(res_pm=40), None)])
def test_invalid_configs_are_rejected_before_any_device_work(kw, code):
    base = dict(res_pm=32, n_part=16, boxsize=100.0, worder=2, deconvolve=0, antialias=0, grad_order=4,
                lap_order=0, n_resample=1, resampling=0, precision=32, deposit=0, fixed_frac_bits=0,
                lean=0, device=0)
    base.update(kw)
    cfg = _lib.DJConfig(**base)
    ptr = C.c_void_p()
    rc = D.lib().dj_create(C.byref(cfg), C.byref(ptr))
    if code is None:
        # a valid option set; without a device this fails in the runtime, never by a crash
        assert rc in (0, -3, -5)
        if rc == 0:
            D.lib().dj_destroy(ptr)
    else:
        assert rc == code
        assert D.lib().dj_last_error(None)


def test_ngenic_host_noise():
    # the ranlxd1 stream is integer-exact arithmetic in doubles: bit-exact
    u = np.empty(5000)
    D.lib().dj_ranlxd1_uniforms(42, 5000, u.ctypes.data)
    assert np.array_equal(u, O.ranlxd1_uniforms(42, 5000))
    assert np.max(np.abs(u[:5] - np.array(G.KNOWN["ranlxd1_seed42_first5"]))) <= G.KNOWN["ranlxd1_atol"]
    # the field applies sqrt(-log u) and cos/sin: same zero pattern, values within 2 ulp (libm's
    # sin/cos/sincos may round differently by one ulp; the reference uses Julia's own libm)
    for seed, res in ((42, 8), (7, 16), (123, 24)):
        a, b = D.ngenic_field(seed, res), O.ngenic_field_gsl(seed, res)
        assert np.array_equal(a == 0, b == 0)
        assert np.max(np.abs(a - b)) <= 2 * np.finfo(np.float64).eps * np.max(np.abs(b))
    f = D.ngenic_field(42, 8)
    assert abs(np.sqrt(np.sum(np.abs(f) ** 2)) - G.KNOWN["ngenic_field_gsl_42_8_norm"]) < 1e-3


def test_host_cosmology_matches_golden():
    c = D.Cosmology.named("Planck18EEBAOSN")
    assert G.rel(c.tables["D1"], G.arr("cosmo/Dplus")) < 1e-13
    assert G.rel(c.tables["Dplusda"], G.arr("cosmo/Dplusda")) < 1e-13
    assert G.rel(c.tables["D2"], G.arr("cosmo/D2plus")) < 1e-13
    assert G.rel(c.tables["D2plusda"], G.arr("cosmo/D2plusda")) < 1e-13
    assert G.rel(c.tables["superconft"], G.arr("cosmo/superconft")) < 1e-13


@pytest.mark.parametrize("name", G.CONFIGS)
def test_host_stepper_coefficients_match_golden(name):
    kw = G.kwargs(name)
    c = D.Cosmology.named("Planck18EEBAOSN")
    tv = kw["time_var"]
    tv = np.array(tv, np.float64) if isinstance(tv, list) else tv
    co = D.stepper_coefficients(c, kw["stepper"], kw["a_ini"], kw["a_end"], kw["n_steps"], tv)
    for k in ("alpha", "beta", "ddrift1", "ddrift2", "all_a"):
        assert G.rel(co[k], G.arr(f"{name}/{k}")) < 1e-12, k


def test_host_linear_power_spectrum_matches_oracle():
    c = D.compute_timetables(D.Cosmology(0.2621, 0.049, 0.6766, 0.8102, 0.9665))
    oc = O.Cosmology(0.2621, 0.049, 0.6766, 0.8102, 0.9665)
    k, pk = D.linear_power_spectrum(c)
    ko, pko = O.linear_power_spectrum(oc)
    assert G.rel(k, ko) < 1e-14 and G.rel(pk, pko) < 1e-12


# ------------------------------------------------------------ FFT pass kernels, emulated on the host
_EMU = None


def _emu():
    global _EMU
    if _EMU is None:
        src = os.path.join(os.path.dirname(D.LIB_PATH), "..", "csrc", "emu", "emu_fft.cpp")
        out = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"libemu_fft_{os.getpid()}.so")
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Wno-unknown-pragmas",
                        "-o", out, src], check=True)
        _EMU = C.CDLL(out)
        _EMU.emu_fft.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                 C.c_double, C.c_double]
    return _EMU


@pytest.mark.parametrize("N", [8, 12, 16, 24, 32, 48, 64, 96, -64])
@pytest.mark.parametrize("prec", [32, 64])
def test_fft_pass_kernels_emulated(N, prec):
    """The device kernels' phase code run workgroup by workgroup on the host: the x-halved B-layout
    spectrum equals numpy's rfftn, and the inverse passes return the field.  N = -64 runs the
    wave-local schedule (each wave transforms only its own lines) on a 64^3 mesh."""
    code = N
    N = abs(N)
    dt, ct = (np.float64, np.complex128) if prec == 64 else (np.float32, np.complex64)
    H = N // 2 + 1
    x = np.random.default_rng(N).normal(size=(N, N, N)).astype(dt)
    B = np.zeros(N * H * N, ct)
    assert _emu().emu_fft(code, prec, 0, x.ctypes.data, B.ctypes.data, None, 0, 1.0, 1.0) == 0
    ref = np.fft.rfftn(x.astype(np.float64), axes=(0, 1, 2))
    got = B.reshape(N, H, N).transpose(2, 0, 1)
    tol = 1e-14 if prec == 64 else 1e-6
    assert np.max(np.abs(got - ref)) < tol * np.max(np.abs(ref))
    Bc = np.ascontiguousarray(ref.transpose(1, 2, 0)).astype(ct)
    out = np.zeros((N, N, N), dt)
    assert _emu().emu_fft(code, prec, 1, Bc.ctypes.data, out.ctypes.data, None, 0, 1.0, 1.0) == 0
    assert np.max(np.abs(out / N**3 - x)) < tol * np.max(np.abs(x))


@pytest.mark.parametrize("code", [32, -64])
def test_fft_poisson_emulated_vs_oracle(code):
    """Poisson solve (XF, YF, ZM with K_phi, YI, XI) against the oracle's FP64 FFT and multiplier, with
    the barrier schedule (32^3) and the wave-local schedule (64^3)."""
    N, L, n = abs(code), 100.0, 16
    H = N // 2 + 1
    x = np.random.default_rng(1).normal(size=(N, N, N))
    dk = 2 * np.pi / L
    kf = np.array([(m if m < N // 2 else m - N) * dk for m in range(N)])
    kh = np.arange(H) * dk
    ktab = np.concatenate([kf**2, kh**2, np.zeros(N), np.zeros(H)])
    out = np.zeros((N, N, N))
    rho = n**3 / N**3
    assert _emu().emu_fft(code, 64, 2, x.ctypes.data, out.ctypes.data, ktab.ctypes.data, 0, rho, 1.0) == 0
    # reference: brfft(rfft(x) * -1/k^2 / rho / N^3)
    F = O.rfft3(x)
    kz = kh[:, None, None]
    k2 = kz**2 + kf[None, :, None] ** 2 + kf[None, None, :] ** 2
    k2[0, 0, 0] = 1
    K = -1 / k2 / rho / N**3
    K[0, 0, 0] = 0
    ref = O.brfft3(F * K)
    assert G.rel(out, ref) < 1e-13


def test_snapshot_roundtrip_and_layout(tmp_path):
    from discodj_amd import snapshots as SN
    from discodj_amd.nbody import add_q
    rng = np.random.default_rng(0)
    n, L = 8, 50.0
    psi = rng.normal(size=(3, n, n, n)).astype(np.float32)
    P = rng.normal(size=(3, n, n, n)).astype(np.float32)
    # remove_q inverts add_q (the X that run_nbody's callback receives)
    assert np.allclose(SN.remove_q(add_q(psi, L), L), psi, atol=1e-5)
    w = SN.SnapshotWriter(str(tmp_path), [0.5, 1.0], L)
    assert w.when(3, 0.5) and w.when(9, 1.0) and not w.when(4, 0.6)
    w(3, 0.5, add_q(psi, L), P)
    w(4, 0.6, add_q(psi, L), P)  # not a target: nothing written
    assert len(w.written) == 1
    s = SN.read_snapshot(w.written[0])
    assert s["a"] == 0.5 and s["step"] == 3 and s["n"] == n and s["boxsize"] == L
    assert np.allclose(s["psi"], psi, atol=1e-5) and np.array_equal(s["P"], P)
    # the reference's column-major (n,n,n,3) [i,j,k,c] array has the same memory layout
    ref = np.asfortranarray(s["psi"].transpose(3, 2, 1, 0))
    assert ref[2, 5, 7, 1] == s["psi"][1, 7, 5, 2]
    assert np.array_equal(ref.ravel(order="F"), s["psi"].ravel(order="C"))


def test_missing_library_fails_loudly():
    """No CPU fallback: without libdiscodj_hip.so every entry point raises with build instructions."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r)\n"
            "import discodj_amd._lib as L\n"
            "L.LIB_PATH = '/nonexistent/libdiscodj_hip.so'; L._lib = None\n"
            "import discodj_amd as D\n"
            "try:\n"
            "    D.PMSolverHIP(D.PMConfig(res_pm=32, n_part=16, boxsize=100.0))\n"
            "except RuntimeError as e:\n"
            "    print('raised:', e)\n") % os.path.dirname(os.path.dirname(os.path.abspath(D.__file__)))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert "raised:" in r.stdout and "build it" in r.stdout, r.stdout + r.stderr


class _FakeSolver:
    """Stands in for a PMSolverHIP (no device here): the attributes check_solver reads."""

    def __init__(self, cfg, precision=32, deposit_mode="float", lean=False):
        self.cfg, self.precision, self.deposit_mode, self.lean = cfg, precision, deposit_mode, lean


def test_a_passed_solver_must_match_the_requested_model():
    from discodj_amd.pm import check_solver
    cfg = D.PMConfig(res_pm=32, n_part=16, boxsize=100.0, worder=3, grad_order=4)
    check_solver(_FakeSolver(cfg), D.PMConfig(res_pm=32, n_part=16, boxsize=100, worder=3, grad_order=4), 32)
    for bad in (dict(worder=2), dict(antialias=1), dict(grad_order=6), dict(deconvolve=True)):
        other = D.PMConfig(**{**cfg.__dict__, **bad})
        with pytest.raises(ValueError):
            check_solver(_FakeSolver(other), cfg, 32)
    with pytest.raises(ValueError):
        check_solver(_FakeSolver(cfg, precision=64), cfg, 32)
    with pytest.raises(ValueError):
        check_solver(_FakeSolver(cfg, deposit_mode="fixed64"), cfg, 32, deposit="float")
    # the drivers check before any device work
    c = D.Cosmology.named("Planck18EEBAOSN")
    n = 16
    with pytest.raises(ValueError):
        D.density_loss_gradient(c, np.zeros((n, n, n)), np.zeros((32, 32, 32)), boxsize=100.0, res_pm=32,
                                precision=32, solver=_FakeSolver(D.PMConfig(res_pm=32, n_part=n, boxsize=100.0,
                                                                             worder=2, grad_order=4),
                                                                 deposit_mode="fixed64"))
    with pytest.raises(ValueError):
        D.run_nbody(c, np.zeros((3, n, n, n), np.float32), np.zeros((3, n, n, n), np.float32), boxsize=100.0,
                    a_ini=0.05, a_end=1.0, n_steps=2, res_pm=32, worder=3,
                    solver=_FakeSolver(D.PMConfig(res_pm=32, n_part=n, boxsize=100.0, worder=2)))
