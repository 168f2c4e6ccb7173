"""Pins of the oracle's implicit collision source (flag_pip_imp = 1,
PIP_rot.f90:1333-1358) and of the temperature/drift-dependent collision
coefficients (flag_col 2..7, PIP_rot.f90:44-97), by closed forms, limits and
scaling laws that do not restate the formulas.

Implicit source, uniform two-fluid state (no fluxes), stage length
dt_c = dt/2 (set_dt_coll, Solver_rot.f90:83-93):
  * the drift w = v_n - v_p is multiplied by e = exp(-alpha (rho_n+rho_p) dt_c)
    per stage, so one SSP-RK2 step gives w e (1 + e) / 2;
  * with no drift, D = T_n - T_p is multiplied by exp(-kappa dt_c) per stage,
    kappa = 3 (gm-1) alpha (mu_p rho_n + mu_n rho_p): D e_k (1 + e_k) / 2 per
    step.  (The explicit source relaxes D at kappa/2; the reference's two
    branches differ by this factor, and the oracle keeps both as written.)
  * as dt -> 0 the drift heating of the implicit branch tends to 3/4 of the
    explicit branch's in the centre-of-mass frame (the branch's stage
    structure carries 3/4 of every exchange per SSP-RK2 step);
  * total mass, momentum and energy are conserved.
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from pip_amd import problems
from pip_amd.config import PIPConfig


def _uniform(xi_n=0.5, col=20.0, T_n=2.0, T_p=0.5, v0=0.0, imp=1, flag_col=1, gm=5.0 / 3.0):
    cfg, uh, um, _ = problems.pn_coupling(nx=8, xi_n=xi_n, col=col, T_n=T_n, T_p=T_p, v0=v0, gm=gm)
    cfg = PIPConfig(**{**cfg.__dict__, "flag_pip_imp": imp, "flag_col": flag_col})
    o = orc.Oracle(cfg, uh, um, prologue=False)
    o.pipbnd()
    return cfg, o


def _step(o, cfg, dt):
    o.set_scalars(dt, 0.0)
    o.run_solver()


def _prims(o, cfg):
    uh, um = o.state()
    gm = cfg.gm
    rn, rp = uh[0, 0, 5], um[0, 0, 5]
    vn, vp = uh[1, 0, 5] / rn, um[1, 0, 5] / rp
    pn = (gm - 1) * (uh[4, 0, 5] - 0.5 * rn * vn * vn)
    pp = (gm - 1) * (um[4, 0, 5] - 0.5 * rp * vp * vp)
    return dict(rn=rn, rp=rp, w=vn - vp, Tn=gm * pn / rn, Tp=0.5 * gm * pp / rp,
                m=uh[1, 0, 5] + um[1, 0, 5], E=uh[4, 0, 5] + um[4, 0, 5], rho=rn + rp)


def test_implicit_drift_relaxation_closed_form():
    col, dt = 40.0, 0.01
    cfg, o = _uniform(xi_n=0.3, col=col, v0=0.4)
    p0 = _prims(o, cfg)
    e = math.exp(-col * p0["rho"] * dt / 2)
    w = p0["w"]
    for _ in range(5):
        _step(o, cfg, dt)
        w *= e * (1 + e) / 2
    p = _prims(o, cfg)
    assert p["w"] == pytest.approx(w, rel=1e-12)
    assert p["m"] == pytest.approx(p0["m"], abs=1e-15)
    assert p["E"] == pytest.approx(p0["E"], rel=1e-14)
    assert p["rho"] == p0["rho"]


def test_implicit_temperature_relaxation_closed_form():
    col, dt, xi_n = 15.0, 0.02, 0.6
    cfg, o = _uniform(xi_n=xi_n, col=col, T_n=2.0, T_p=0.5)
    p0 = _prims(o, cfg)
    kappa = 3 * (cfg.gm - 1) * col * (0.5 * p0["rn"] + 1.0 * p0["rp"])
    ek = math.exp(-kappa * dt / 2)
    D = p0["Tn"] - p0["Tp"]
    for _ in range(4):
        _step(o, cfg, dt)
        D *= ek * (1 + ek) / 2
    p = _prims(o, cfg)
    assert p["Tn"] - p["Tp"] == pytest.approx(D, rel=1e-11)
    assert p["E"] == pytest.approx(p0["E"], rel=1e-14)


def test_implicit_drift_heating_tends_to_explicit():
    """Equal temperatures, centre-of-mass frame with rho_n != rho_p: the total
    energy moved to the neutrals per unit time (the dS5 term alone) tends, as
    dt -> 0, to 3/4 of the explicit branch's.  The 3/4 is the stage
    structure of the implicit branch: it updates the stage base U0 by dt/2
    in each stage, and SSP-RK2 weights the second update by c0 = 1/2 (the
    same factor the drift closed form above shows: e(1+e)/2 ~ 1 - 3/4 z)."""
    col, w = 5.0, 0.3
    rho_n, rho_p = 0.3, 0.7
    gm = 5.0 / 3.0
    cfg0 = PIPConfig(ndim=1, nx=8, xmin=0.0, xmax=2.0, col=col, gm=gm, flag_bnd=(1, 1, 1, 1))
    z = np.zeros((cfg0.jx, cfg0.ix))
    vn, vp = rho_p * w / (rho_n + rho_p), -rho_n * w / (rho_n + rho_p)
    T = 1.0
    uh = problems.pv2cq_hd(cfg0, z + rho_n, z + vn, z, z, z + rho_n * T / gm)
    um = problems.pv2cq_mhd(cfg0, z + rho_p, z + vp, z, z, z + 2 * rho_p * T / gm, z, z, z)
    rates = {}
    for imp in (0, 1):
        cfg = PIPConfig(**{**cfg0.__dict__, "flag_pip_imp": imp})
        for dt in (1e-3, 1e-4, 1e-5):
            o = orc.Oracle(cfg, uh, um, prologue=False)
            o.pipbnd()
            _step(o, cfg, dt)
            uh1, _ = o.state()
            rates[(imp, dt)] = (uh1[4, 0, 5] - uh[4, 0, 5]) / dt
    # the explicit rate is ac rho_n rho_p (v_n^2 - v_p^2) / 2 (T_n = T_p)
    assert rates[(0, 1e-5)] == pytest.approx(-col * rho_n * rho_p * 0.5 * (vn ** 2 - vp ** 2), rel=1e-3)
    for dt in (1e-3, 1e-4, 1e-5):
        rel = abs(rates[(1, dt)] / rates[(0, dt)] / 0.75 - 1)
        assert rel < 5 * col * dt, (dt, rates[(1, dt)], rates[(0, dt)])


def test_implicit_conservation_on_a_2d_run():
    cfg, uh, um, _ = problems.orszag_tang(24)
    cfg = PIPConfig(**{**cfg.__dict__, "flag_pip_imp": 1})
    o = orc.Oracle(cfg, uh, um)
    h0, m0 = o.state()
    tot0 = [h0[n, 2:-2, 2:-2].sum() + m0[n, 2:-2, 2:-2].sum() for n in range(5)]
    o.run(nsteps=10)
    h1, m1 = o.state()
    tot1 = [h1[n, 2:-2, 2:-2].sum() + m1[n, 2:-2, 2:-2].sum() for n in range(5)]
    scale = [abs(t) + h0[n, 2:-2, 2:-2].__abs__().sum() for n, t in enumerate(tot0)]
    for n in range(5):
        assert abs(tot1[n] - tot0[n]) <= 1e-13 * scale[n], n


def test_implicit_cfl_limit():
    """cfl_pn_col (Scheme_rot.f90:305-313): pip_imp_factor / max(rho*alpha)."""
    cfg, o = _uniform(xi_n=0.4, col=1000.0, imp=1)
    dt = o.cfl()
    assert dt == pytest.approx(cfg.pip_imp_factor / (0.6 * 1000.0), rel=1e-15)
    cfg, o = _uniform(xi_n=0.4, col=1000.0, imp=0)
    assert o.cfl() == pytest.approx(min(cfg.safety, 0.3) / (0.6 * 1000.0), rel=1e-15)


# ----------------------------------------------------- set_collisional
def _alpha(flag_col, T_n, T_p, v0, col=7.0, beta=0.1):
    cfg, o = _uniform(xi_n=0.5, col=col, T_n=T_n, T_p=T_p, v0=v0, imp=0, flag_col=flag_col)
    cfg.beta = beta
    o2 = orc.Oracle(PIPConfig(**{**cfg.__dict__, "beta": beta}), *o.state(), prologue=False)
    orc.lib().orc_set_coefficients(o2.h, 1)
    ac = np.empty((cfg.jx, cfg.ix))
    orc.lib().orc_get_ac(o2.h, orc._p(ac))
    return ac[0, 5]


def test_collisional_sqrt_temperature_scaling():
    """col_type 1, 2 (flag_col 2, 3): hard-sphere rate, alpha ~ sqrt(T)."""
    for fc in (2, 3):
        a1 = _alpha(fc, 1.0, 0.6, 0.0)
        a4 = _alpha(fc, 4.0, 2.4, 0.0)
        assert a4 == pytest.approx(2 * a1, rel=1e-14)
    # col_type 1 differs from 2 only by the constant 1/sqrt(beta gm / 2)
    for beta in (0.1, 0.4):
        r = _alpha(2, 1.3, 0.7, 0.0, beta=beta) / _alpha(3, 1.3, 0.7, 0.0, beta=beta)
        assert r == pytest.approx(1 / math.sqrt(beta / 2 * 5 / 3), rel=1e-14)


def test_collisional_drift_correction():
    """col_type 3, 4 (flag_col 4, 5): the drift factor sqrt(1 + c v_d^2/T) --
    equal to col_type 1, 2 with no drift; (alpha/alpha_0)^2 - 1 is
    proportional to v_d^2."""
    for fc, base in ((4, 2), (5, 3)):
        assert _alpha(fc, 1.0, 0.5, 0.0) == pytest.approx(_alpha(base, 1.0, 0.5, 0.0), rel=1e-15)
        a0 = _alpha(fc, 1.0, 0.5, 0.0)
        g1 = (_alpha(fc, 1.0, 0.5, 0.2) / a0) ** 2 - 1
        g2 = (_alpha(fc, 1.0, 0.5, 0.4) / a0) ** 2 - 1
        assert g2 == pytest.approx(4 * g1, rel=1e-6)
        assert g1 > 0


def test_collisional_eighth_power_laws():
    """col_type 5, 6 (flag_col 6, 7) multiply col_type 3, 4 by an eighth
    power of a temperature-like argument: with no drift, scaling T by 16
    moves the extra factor by 16^(-1/8) (col_type 5) and 16^(-1/8) (6)."""
    r5a = _alpha(6, 1.0, 0.5, 0.0) / _alpha(4, 1.0, 0.5, 0.0)
    r5b = _alpha(6, 16.0, 8.0, 0.0) / _alpha(4, 16.0, 8.0, 0.0)
    assert r5b / r5a == pytest.approx(16 ** (-0.125), rel=1e-12)
    r6a = _alpha(7, 1.0, 0.5, 0.0) / _alpha(5, 1.0, 0.5, 0.0)
    r6b = _alpha(7, 16.0, 8.0, 0.0) / _alpha(5, 16.0, 8.0, 0.0)
    assert r6b / r6a == pytest.approx(16 ** (-0.125), rel=1e-12)
