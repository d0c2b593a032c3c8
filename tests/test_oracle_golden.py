"""The CPU oracle against the reference's golden vectors (JAX DISCO-DJ, FP64, 16^3 / L = 100).

Mirrors NR/test/nbody_parity_tests.jl:13-91 with the same tolerances: 1e-13 timetables,
1e-12 ICs / coefficients / acc_ini, 1e-11 full runs.
"""
import numpy as np
import pytest

import golden as G
import oracle as O


@pytest.fixture(scope="module")
def cosmo():
    return O.Cosmology.named("Planck18EEBAOSN")


def test_timetables(cosmo):
    # nbody_parity_tests.jl:17-24
    assert G.rel(cosmo.table("a"), G.arr("cosmo/a")) < 1e-14
    assert G.rel(cosmo.table("D1"), G.arr("cosmo/Dplus")) < 1e-13
    assert G.rel(cosmo.table("Dplusda"), G.arr("cosmo/Dplusda")) < 1e-13
    assert G.rel(cosmo.table("D2"), G.arr("cosmo/D2plus")) < 1e-13
    assert G.rel(cosmo.table("D2plusda"), G.arr("cosmo/D2plusda")) < 1e-13
    assert G.rel(cosmo.table("superconft"), G.arr("cosmo/superconft")) < 1e-13
    ref = G.MANIFEST["groups"]["cosmo"]["attrs"]
    assert cosmo.D1_unnormed_at_1 == pytest.approx(ref["Dplus_unnormed_at_1"], rel=1e-13)
    assert cosmo.Omega_m() == pytest.approx(ref["Omega_m"], rel=1e-15)


def test_growth_eds_limits(cosmo):
    # runtests.jl:32-55: EdS limits D2/D1^2 -> -3/7, D3a/D1^3 -> 1/3, D3b -> -10/21, D3c -> 1/7 at early times
    i = 100
    D1 = cosmo.table("D1")[i]
    assert cosmo.table("D2")[i] / D1**2 == pytest.approx(-3 / 7, rel=1e-6)
    assert cosmo.table("D3a")[i] / D1**3 == pytest.approx(1 / 3, rel=1e-6)
    assert cosmo.table("D3b")[i] / D1**3 == pytest.approx(-10 / 21, rel=1e-6)
    assert cosmo.table("D3c")[i] / D1**3 == pytest.approx(1 / 7, rel=1e-6)


def _time_var(cfg):
    kw = G.kwargs(cfg)
    tv = kw["time_var"]
    return np.array(tv, np.float64) if isinstance(tv, list) else tv


@pytest.mark.parametrize("cfg", G.CONFIGS)
def test_stepper_coefficients(cosmo, cfg):
    kw = G.kwargs(cfg)
    co = O.stepper_coefficients(cosmo, kw["stepper"], kw["a_ini"], kw["a_end"], kw["n_steps"], _time_var(cfg))
    for k in ("alpha", "beta", "ddrift1", "ddrift2", "all_a"):
        assert G.rel(co[k], G.arr(f"{cfg}/{k}")) < 1e-12, k


def test_lpt_ics(cosmo):
    # nbody_parity_tests.jl:28-34
    psi1, psi2 = O.compute_core2(G.fphi(), G.BOXSIZE)
    psi, pi = O.nbody_ics_lpt2(cosmo, psi1, psi2, 0.05)
    assert G.rel(psi, G.particles("ics/psi_ini")) < 1e-12
    assert G.rel(pi, G.particles("ics/pi_ini")) < 1e-12
    # psi1 of the file is the 1LPT shape field (D-derivative of the 1st order at a = 1)
    assert G.rel(psi1, G.particles("ics/psi1")) < 1e-12


def _cfg(cfg):
    kw = G.kwargs(cfg)
    return O.pm_config(res_pm=kw["res_pm"], n_part=G.N_PART, boxsize=G.BOXSIZE, worder=kw["worder"],
                       deconvolve=kw["deconvolve"], antialias=kw["antialias"],
                       grad_order=kw["grad_kernel_order"], lap_order=kw["laplace_kernel_order"],
                       n_resample=kw["n_resample"], resampling=kw["resampling_method"])


@pytest.mark.parametrize("form", ["reference", "solver"])
@pytest.mark.parametrize("cfg", G.CONFIGS)
def test_acc_ini(cfg, form):
    # nbody_parity_tests.jl:47-51
    acc = O.pm_acceleration(G.particles("ics/psi_ini"), _cfg(cfg), form=form)
    assert G.rel(acc, G.particles(f"{cfg}/acc_ini")) < 1e-12


@pytest.mark.parametrize("cfg", G.CONFIGS)
def test_run_nbody(cosmo, cfg):
    # nbody_parity_tests.jl:52-61
    kw = G.kwargs(cfg)
    if kw["ic_method"] == "bullfrog":
        p0, m0 = O.bullfrog_ics(cosmo, G.particles("ics/psi1"), kw["a_ini"], G.BOXSIZE, kw["res_pm"])
    else:
        p0, m0 = G.particles("ics/psi_ini"), G.particles("ics/pi_ini")
    psi, p = O.run_nbody(cosmo, p0, m0, _cfg(cfg), kw["a_ini"], kw["a_end"], kw["n_steps"],
                         stepper=kw["stepper"], time_var=_time_var(cfg))
    assert G.rel(psi, G.particles(f"{cfg}/psi_out")) < 1e-11
    assert G.rel(p, G.particles(f"{cfg}/p_out")) < 1e-11
