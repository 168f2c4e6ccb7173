"""Independent pins of the oracle pieces that the textbook-solver, closed-form
and symmetry tests of test_oracle.py leave open:

  * the HLLC-G fallback of the HLLD solver (HLL_rot.f90:775-807) against
    Li's (2005) HLLC for ideal MHD on pairs that force the fallback;
  * the smoothed upwind switches of the HLLC solver inside their band
    |S| <= epsi/2 (HLL_rot.f90:679-685), written piecewise;
  * the ionisation-potential / arb_heat energy terms of set_IR type 0
    (PIP_rot.f90:219-225, 1382-1397) by total-energy budgets with frozen
    rates, and cfl_pip_ir's ion_pot limit (Scheme_rot.f90:320-323) in
    closed form;
  * the constants of set_collisional (PIP_rot.f90:50-95): 9 pi/64 of the
    drift factor against the hard-sphere momentum-transfer rate of two
    drifting Maxwellians, pi/16 of the eighth-power factors against their
    mean relative speed, and the 1/sqrt(beta gm/2) normalisation at the
    reference temperature.
Each check uses a published formula or a limit, not the oracle's own
expression, and fails on a plausible restatement mistake (a wrong side,
sign, factor or constant).
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from pip_amd import problems
from pip_amd.config import PIPConfig
from tests import textbook as tb

GM = 5.0 / 3.0
K_B_EV = 8.6173e-5  # Boltzmann constant, eV/K
CHI_H = 13.6        # hydrogen ionisation potential, eV


# ------------------------------------------------------- HLLC-G fallback
def _strong_field_state(rng, bx):
    """Low-beta states with a dominant normal field: the Alfven and fast
    speeds nearly coincide, which is where the fallback is taken."""
    return np.array([rng.uniform(0.5, 1.5), rng.normal(0, 0.05), rng.normal(0, 0.05), rng.normal(0, 0.05),
                     rng.uniform(0.005, 0.2), bx, rng.normal(0, 0.05), rng.normal(0, 0.05)])


def test_hllc_g_fallback_matches_li_hllc_mhd():
    """Where the fallback condition holds, the oracle's HLLD flux equals
    Li's (2005) HLLC-MHD flux (intermediate states with the HLL transverse
    velocity and field); the full HLLD flux of Miyoshi & Kusano differs
    there, so the fallback branch is what is compared."""
    rng = np.random.default_rng(2005)
    n = far = 0
    sides = set()
    while n < 200:
        bx = rng.choice([-1.0, 1.0]) * rng.uniform(0.5, 3.0)
        vl, vr = _strong_field_state(rng, bx), _strong_field_state(rng, bx)
        if not tb.hlld_fallback_triggered(vl, vr, GM):
            continue
        f = orc.hlld_flux(vl, vr, GM)
        ref = tb.li_hllc_mhd(vl, vr, GM)
        scale = np.abs(ref).max()
        np.testing.assert_allclose(f, ref, rtol=0, atol=1e-12 * scale)
        mk, _ = tb.mk_hlld(vl, vr, GM)
        far += np.abs(mk - ref).max() > 1e-6 * scale
        cf = max(tb.fast_speed(vl, GM), tb.fast_speed(vr, GM))
        sl, sr = min(vl[1], vr[1]) - cf, max(vl[1], vr[1]) + cf
        Uh = (sr * tb.mhd_cons(vr, GM) - sl * tb.mhd_cons(vl, GM) - tb.mhd_flux(vr, GM) + tb.mhd_flux(vl, GM)) / (sr - sl)
        sides.add(Uh[1] >= 0)
        n += 1
    assert far > 150  # the two solvers differ on most of these pairs
    assert sides == {True, False}  # both intermediate states sampled


def test_hllc_g_fallback_threshold_factor():
    """The fallback starts at the factor 1.001 (HLL_rot.f90:701): pairs whose
    condition holds only for factors in (1.0, 1.001] take Li's HLLC-MHD,
    pairs that need a factor in (1.001, 1.05] keep the full HLLD of
    Miyoshi & Kusano."""
    rng = np.random.default_rng(1001)
    below = above = 0
    while below < 40 or above < 120:
        bx = rng.choice([-1.0, 1.0]) * rng.uniform(0.5, 3.0)
        vl, vr = _strong_field_state(rng, bx), _strong_field_state(rng, bx)
        f = tb.hlld_fallback_margin(vl, vr, GM)
        if not 1.0 < f <= 1.05:
            continue
        got = orc.hlld_flux(vl, vr, GM)
        li = tb.li_hllc_mhd(vl, vr, GM)
        mk, _ = tb.mk_hlld(vl, vr, GM)
        scale = np.abs(mk).max()
        if f <= 1.001:
            np.testing.assert_allclose(got, li, rtol=0, atol=1e-12 * scale)
            below += 1
        else:
            np.testing.assert_allclose(got, mk, rtol=0, atol=1e-9 * scale)
            above += np.abs(mk - li).max() > 1e-6 * scale
    assert above >= 120 and below >= 40


# ------------------------------------------------- HLLC switch band
def _hd(r, u, vt, wt, p):
    return np.array([r, u, vt, wt, p])


def test_hllc_contact_switch_band():
    """Contact speed S_M in (0, epsi/2): the flux is the th_M = 2 S_M/epsi
    blend of the two star fluxes, not the upwind star flux."""
    eps = 1.0e-10
    for u0, (rl, rr), (vl_, vr_) in ((2.5e-11, (1.0, 0.4), (0.3, -0.2)), (1.0e-11, (0.7, 1.3), (-0.5, 0.1)),
                                     (4.0e-11, (2.0, 0.5), (0.0, 0.8))):
        vl, vr = _hd(rl, u0, vl_, 0.1, 1.0), _hd(rr, u0, vr_, -0.3, 1.0)
        f = orc.hllc_flux(vl, vr, GM)
        ref, thl, thm, thr = tb.hllc_blend(vl, vr, GM, eps)
        assert thl == 0.0 and thr == 0.0 and 0.05 < thm < 0.95
        np.testing.assert_allclose(f, ref, rtol=0, atol=1e-14)
        sharp = tb.toro_hllc(vl, vr, GM)
        assert np.abs(sharp - ref).max() > 1e-3 * np.abs(u0)  # the band changes the flux


def test_hllc_outer_switch_bands():
    """S_L in (0, epsi/2) (flow to the right, just supersonic) and S_R in
    (0, epsi/2) (flow to the left): th_L = 2 S_L/epsi, th_R = 1 - 2 S_R/epsi."""
    eps, acc = 1.0e-10, 1.03
    for side in ("left", "right"):
        for d in (1.0e-11, 2.5e-11, 4.0e-11):
            rl, pl, rr, pr = 1.0, 1.0, 0.5, 0.8
            cmax = max(math.sqrt(GM * pl / rl), math.sqrt(GM * pr / rr))
            cmin = min(math.sqrt(GM * pl / rl), math.sqrt(GM * pr / rr))
            u = (cmax + d / acc) if side == "left" else (-cmax + d / acc)
            vl, vr = _hd(rl, u, 0.2, 0.0, pl), _hd(rr, u, -0.1, 0.3, pr)
            f = orc.hllc_flux(vl, vr, GM)
            ref, thl, thm, thr = tb.hllc_blend(vl, vr, GM, eps, acc)
            if side == "left":
                assert 0.05 < thl < 0.95 and thr == 0.0
            else:
                assert 0.05 < thr < 0.95 and thl == 0.0
            np.testing.assert_allclose(f, ref, rtol=0, atol=1e-13)
            sharp = tb.toro_hllc(vl, vr, GM)
            assert np.abs(sharp - ref).max() > 1e-3 * d
            del cmin


# -------------------------------------------- IR type 0 energy terms
def _uniform_ir(T=10.0, xi_n=0.8, col=1.0, gr=0.2, gi=0.3, arb=3.0, flag_col=1):
    cfg, uh, um, _ = problems.pn_coupling(nx=8, xi_n=xi_n, col=col, T_n=T, T_p=T)
    cfg = PIPConfig(**{**cfg.__dict__, "flag_ir": 3, "flag_ir_type": 0, "flag_col": flag_col})
    o = orc.Oracle(cfg, uh, um, prologue=False)
    o.fixed_rates(True)
    shape = (cfg.jx, cfg.ix)
    o.set_rates(np.full(shape, gr), np.full(shape, gi))
    o.set_arb_heat(np.full(shape, arb))
    o.pipbnd()
    return cfg, o


def _heat_factor(cfg):
    """Ionisation energy per unit of ionised mass in the code's units: chi_H
    / (k_B T0) / gm for odd flag_col, chi_H beta / (2 k_B T0) for even."""
    if cfg.flag_col % 2 == 1:
        return CHI_H / (K_B_EV * cfg.T0) / cfg.gm
    return CHI_H * cfg.beta / (2.0 * K_B_EV * cfg.T0)


def _totals(uh, um):
    return uh[0, 0, 5] + um[0, 0, 5], uh[4, 0, 5] + um[4, 0, 5]


@pytest.mark.parametrize("flag_col", [1, 2])
def test_ir_type0_energy_budget_with_frozen_rates(flag_col):
    """Uniform periodic box, frozen rates G_rec, G_ion and a given arb_heat A:
    the exchange terms cancel between the fluids, so the total energy of a
    cell changes only by the source (A - ion_pot), ion_pot = G_ion rho_n h:
      stage 1: dt (A - h G_ion rho_n^0),
      step:    dt A - dt/2 h G_ion (rho_n^0 + rho_n^(1)),
    rho_n^(1) = rho_n^0 + dt (G_rec rho_p^0 - G_ion rho_n^0); total mass is
    conserved."""
    gr, gi, arb = 0.2, 0.3, 3.0
    cfg, o = _uniform_ir(gr=gr, gi=gi, arb=arb, flag_col=flag_col)
    dt = o.cfl()
    uh0, um0 = o.state()
    M0, E0 = _totals(uh0, um0)
    rn0, rp0 = uh0[0, 0, 5], um0[0, 0, 5]
    s1 = o.run_solver_stage1()
    uh1, um1 = o.state()
    M2, E2 = _totals(uh1, um1)
    E1 = s1[4, 0, 5] + s1[cfg.nvar_m + 4, 0, 5]
    h = _heat_factor(cfg)
    ip0 = h * gi * rn0
    rn1 = rn0 + dt * (gr * rp0 - gi * rn0)
    assert abs(ip0 * dt) > 1e-3  # the source is not negligible
    # the reference's k_B is a single-precision literal (relative 3e-8)
    assert E1 - E0 == pytest.approx(dt * (arb - ip0), rel=1e-6)
    assert E2 - E0 == pytest.approx(dt * arb - 0.5 * dt * h * gi * (rn0 + rn1), rel=1e-6)
    assert M2 == pytest.approx(M0, rel=1e-15)
    _, _, _, ion_pot = o.rates()
    assert ion_pot[0, 5] == pytest.approx(h * gi * rn1, rel=1e-6)  # ion_pot of the last stage


def test_ir_type0_first_stage_after_set_ir_is_balanced():
    """set_IR fixes arb_heat = G_ion rho_n h on the first call of a fresh
    run (PIP_rot.f90:219-225), so in the first stage the ionisation
    potential loss and arb_heat cancel: the total energy of every cell is
    unchanged to round-off, with ion_pot dt well above it."""
    cfg, uh, um, _ = problems.pn_coupling(nx=8, xi_n=0.8, col=1.0, T_n=1.2, T_p=1.2)
    cfg = PIPConfig(**{**cfg.__dict__, "flag_ir": 3, "flag_ir_type": 0, "t_ir": 0.3})
    o = orc.Oracle(cfg, uh, um)  # prologue: set_IR fixes arb_heat
    gr, gi, arb, _ = o.rates()
    uhp, _ = o.state()
    h = _heat_factor(cfg)
    np.testing.assert_allclose(arb[0, 2:-2], gi[0, 2:-2] * uhp[0, 0, 2:-2] * h, rtol=1e-6)
    o.set_ir()
    dt = o.cfl()
    uh0, um0 = o.state()
    _, E0 = _totals(uh0, um0)
    s1 = o.run_solver_stage1()
    E1 = s1[4, 0, 5] + s1[cfg.nvar_m + 4, 0, 5]
    ip_dt = h * gi[0, 5] * uh0[0, 0, 5] * dt
    assert ip_dt > 1e6 * math.ulp(E0)  # the two terms are far above round-off
    assert abs(E1 - E0) <= 1e-10 * ip_dt + 4 * math.ulp(E0)


def test_cfl_pip_ir_ion_pot_limit_closed_form():
    """cfl_pip_ir (Scheme_rot.f90:320-323): once source_PIP has defined
    ion_pot, dt <= safety / max(ion_pot / (p_p/(gm-1))).  Uniform state with
    frozen rates chosen so this limit binds: after one step, the next dt is
    safety (p_p/(gm-1)) / (G_ion rho_n^(1) h), with rho_n^(1) the neutral
    density of the step's second stage (where ion_pot was last computed)
    and p_p the plasma pressure of the current state."""
    gr, gi = 0.5, 0.5
    cfg, o = _uniform_ir(T=1.0, col=1.0e-3, gr=gr, gi=gi, arb=0.0)
    dt0 = o.cfl()
    uh0, um0 = o.state()
    rn0, rp0 = uh0[0, 0, 5], um0[0, 0, 5]
    o.run_solver()
    dt1 = o.cfl()
    _, um = o.state()
    r, mx, my, mz, E = um[0:5, 0, 5]
    b2 = (um[5:8, 0, 5] ** 2).sum()
    p_p = (cfg.gm - 1.0) * (E - 0.5 * ((mx * mx + my * my + mz * mz) / r + b2))
    rn1 = rn0 + dt0 * (gr * rp0 - gi * rn0)
    expect = cfg.safety * (p_p / (cfg.gm - 1.0)) / (gi * rn1 * _heat_factor(cfg))
    cfl_limit = cfg.safety * cfg.dx / math.sqrt(cfg.gm * p_p / r)
    assert expect < 0.5 * min(cfl_limit, min(cfg.safety, 0.1) / max(gr, gi))  # this limit binds
    assert dt1 == pytest.approx(expect, rel=1e-6)


# ------------------------------------------------ set_collisional
def _alpha(flag_col, T_n, T_p, v0, col=7.0, beta=0.1):
    cfg, uh, um, _ = problems.pn_coupling(nx=8, xi_n=0.5, col=col, T_n=T_n, T_p=T_p, v0=v0)
    cfg = PIPConfig(**{**cfg.__dict__, "flag_col": flag_col, "beta": beta})
    o = orc.Oracle(cfg, uh, um, prologue=False)
    o.pipbnd()
    orc.lib().orc_set_coefficients(o.h, 1)
    ac = np.empty((cfg.jx, cfg.ix))
    orc.lib().orc_get_ac(o.h, orc._p(ac))
    return ac[0, 5]


def _thermal_speed(T_n, T_p):
    """a with a^2 = 2 k (T_n/m_n + T_p/m_p): the code's temperatures are
    T = gm k T / m_H for both fluids (get_Te_HD, get_Te_MHD)."""
    return math.sqrt(2.0 * (T_n + T_p) / GM)


def test_kinetic_rates_closed_forms_match_quadrature():
    for x in (0.05, 0.5, 1.0, 2.0, 6.0):
        assert tb.drift_drag_rate(x) == pytest.approx(tb.drift_drag_rate_quad(x), rel=1e-9)
        assert tb.mean_relative_speed(x) == pytest.approx(tb.mean_relative_speed_quad(x), rel=1e-9)


def test_collisional_drift_factor_is_the_hard_sphere_drag_interpolation():
    """col_type 4 (flag_col 5): alpha(v_d)/alpha(0) is the interpolation
    sqrt(1 + (3 sqrt(pi)/8)^2 x^2) of the hard-sphere momentum-transfer rate
    of two drifting Maxwellians, x = v_d / a: exact in both limits (x -> 0:
    1; x -> infinity: the rate's asymptote 3 sqrt(pi)/8 x) and within 1.5 %
    in between.  A constant other than 9 pi/64, or a missing gm/2 in v_d^2 /
    a^2, fails the large-drift limit."""
    T_n, T_p = 1.3, 0.7
    a = _thermal_speed(T_n, T_p)
    a0 = _alpha(5, T_n, T_p, 0.0)
    r0 = tb.drift_drag_rate(0.0)
    for x, tol in ((1e-3, 1e-6), (0.3, 0.015), (1.0, 0.015), (3.0, 0.015), (10.0, 0.015), (1e3, 1e-5)):
        ratio = _alpha(5, T_n, T_p, x * a) / a0
        exact = tb.drift_drag_rate(x) / r0
        assert ratio == pytest.approx(exact, rel=tol), (x, ratio, exact)


def test_collisional_normalisation_at_the_reference_temperature():
    """col_type 1 and 3 (flag_col 2, 4) divide by sqrt(beta gm/2): alpha = col
    exactly when the mean temperature (T_n + T_p)/2 is the normalisation
    temperature gm p0/rho0 = gm beta/2 (B0 = rho0 = 1), with no drift."""
    for beta in (0.1, 0.5, 2.0):
        T = GM * beta / 2.0
        for fc in (2, 4):
            assert _alpha(fc, T, T, 0.0, col=7.0, beta=beta) == pytest.approx(7.0, rel=1e-14)


def test_collisional_eighth_power_uses_the_mean_relative_speed():
    """col_type 6 (flag_col 7) multiplies col_type 4 by
    (T_tot/2 + gm pi/16 v_d^2)^(-1/8) = ((gm pi/16) <|w|>^2)^(-1/8) with
    <|w|>^2 interpolated as (4 a^2/pi)(1 + (pi/4) x^2) -- exact in both
    limits of the mean relative speed of two drifting Maxwellians (2a/sqrt(pi)
    at rest, v_d at large drift) and within 6 % in between.  col_type 5
    (flag_col 6) carries the same drift term: at large drift the two eighth
    powers cancel."""
    T_n, T_p = 1.3, 0.7
    a = _thermal_speed(T_n, T_p)
    for x, tol in ((1e-3, 1e-5), (0.3, 0.06), (1.0, 0.06), (3.0, 0.06), (1e3, 1e-5)):
        v = x * a
        f6 = _alpha(7, T_n, T_p, v) / _alpha(5, T_n, T_p, v)
        w2 = (a * tb.mean_relative_speed(x)) ** 2
        assert f6 ** -8 == pytest.approx(GM * math.pi / 16.0 * w2, rel=tol), x
    v = 1e3 * a
    f5 = _alpha(6, T_n, T_p, v) / _alpha(4, T_n, T_p, v)
    f6 = _alpha(7, T_n, T_p, v) / _alpha(5, T_n, T_p, v)
    assert f5 * f6 == pytest.approx(1.0, rel=1e-6)
