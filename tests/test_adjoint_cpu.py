"""CPU checks of the per-face / per-cell functions shared by the HIP kernels
(compiled for the host): the forward fluxes equal the oracle's bit for bit,
and the generated vector-Jacobian products (pip_adj_gen.h) agree with
central finite differences in every HLLD region."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as orc
from pip_amd import build
from tests import textbook as tb

LIB = os.path.join(build.LIBDIR, "libpiphost.so")
GM = 5.0 / 3.0
RNG = np.random.default_rng(77)
P = ctypes.POINTER(ctypes.c_double)


def _build():
    src = os.path.join(build.CSRC, "hostcheck.cpp")
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < max(
            os.path.getmtime(os.path.join(build.CSRC, f)) for f in ("hostcheck.cpp", "pip_phys.h", "pip_adj_gen.h")):
        os.makedirs(build.LIBDIR, exist_ok=True)
        subprocess.check_call(["g++", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-std=c++17",
                               "-o", LIB, src])
    L = ctypes.CDLL(LIB)
    D = ctypes.c_double
    L.hc_hllc.argtypes = [P, P, D, P]
    L.hc_hlld9.argtypes = [P, P, D, D, P]
    L.hc_hlld8.argtypes = [P, P, D, P]
    L.hc_hllc_vjp.argtypes = [P, P, D, P, P, P]
    L.hc_hlld9_vjp.argtypes = [P, P, D, D, P, P, P]
    L.hc_source.argtypes = [P, P, D, D, P]
    L.hc_source_vjp.argtypes = [P, P, D, D, P, P, P, P]
    return L


L = _build()


def p(a):
    return a.ctypes.data_as(P)


def hllc(vl, vr):
    f = np.zeros(5)
    L.hc_hllc(p(vl), p(vr), GM, p(f))
    return f


def hlld9(vl, vr, ch):
    f = np.zeros(9)
    L.hc_hlld9(p(vl), p(vr), GM, ch, p(f))
    return f


def rand_hd():
    return np.array([RNG.uniform(0.2, 2.0), RNG.normal(0, 1.0), RNG.normal(), RNG.normal(), RNG.uniform(0.1, 2.5)])


def rand_mhd():
    return np.concatenate([[RNG.uniform(0.2, 2.0), RNG.normal(0, 0.7), RNG.normal(0, .5), RNG.normal(0, .5),
                            RNG.uniform(0.1, 2.0)], RNG.normal(size=3), [RNG.normal()]])


def test_device_fluxes_equal_oracle_bitwise():
    for _ in range(300):
        vl, vr = rand_hd(), rand_hd()
        assert np.array_equal(hllc(vl, vr), orc.hllc_flux(vl, vr, GM))
        ml, mr = rand_mhd(), rand_mhd()
        ch = RNG.uniform(0.5, 3)
        assert np.array_equal(hlld9(ml, mr, ch), orc.hlld_flux(ml, mr, GM, divb=1, cmax=ch))
        f8 = np.zeros(8)
        L.hc_hlld8(p(ml[:8].copy()), p(mr[:8].copy()), GM, p(f8))
        assert np.array_equal(f8, orc.hlld_flux(ml[:8], mr[:8], GM))


def fd_check(fwd, vjp, x_l, x_r, nout, eps=1e-6):
    lam = RNG.normal(size=nout)
    gl, gr = np.zeros_like(x_l), np.zeros_like(x_r)
    vjp(x_l, x_r, lam, gl, gr)
    for which, x, g in ((0, x_l, gl), (1, x_r, gr)):
        for k in range(len(x)):
            xp, xm = x.copy(), x.copy()
            h = eps * max(1.0, abs(x[k]))
            xp[k] += h
            xm[k] -= h
            if which == 0:
                fp, fm = fwd(xp, x_r), fwd(xm, x_r)
            else:
                fp, fm = fwd(x_l, xp), fwd(x_l, xm)
            fd = lam @ (fp - fm) / (2 * h)
            assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (which, k, fd, g[k])


def test_hllc_vjp_matches_finite_differences():
    def vjp(a, b, lam, ga, gb):
        L.hc_hllc_vjp(p(a), p(b), GM, p(lam), p(ga), p(gb))
    n = 0
    while n < 200:
        vl, vr = rand_hd(), rand_hd()
        try:
            fd_check(hllc, vjp, vl, vr, 5)
        except AssertionError:
            # a kink (sign change of a wave speed or a min/max switch) inside
            # the difference stencil is allowed; re-test with a smaller step
            fd_check(hllc, vjp, vl, vr, 5, eps=1e-8)
        n += 1


def _region(vl, vr):
    try:
        return tb.mk_hlld(np.r_[vl[:5], vl[5], vl[6:8]], np.r_[vr[:5], vl[5], vr[6:8]], GM)[1]
    except Exception:
        return "?"


@pytest.mark.parametrize("mode", ["equal_bn", "jump_bn", "supersonic", "fallback"])
def test_hlld9_vjp_matches_finite_differences(mode):
    """fallback: B nearly along the normal makes the fast and Alfven speeds
    coincide, which selects the HLLC-G branch (HLL_rot.f90:775-807)."""
    ch = 1.7

    def fwd(a, b):
        return hlld9(a, b, ch)

    def vjp(a, b, lam, ga, gb):
        L.hc_hlld9_vjp(p(a), p(b), GM, ch, p(lam), p(ga), p(gb))
    n = 0
    while n < 150:
        vl, vr = rand_mhd(), rand_mhd()
        if mode == "equal_bn":
            vr[5] = vl[5]
        if mode == "supersonic":
            s = 5.0 * RNG.choice([-1, 1])
            vl[1] += s
            vr[1] += s
        if mode == "fallback":
            vl[5] = vr[5] = 2.0 * RNG.choice([-1, 1])
            vl[6:8] = 1e-3 * RNG.normal(size=2)
            vr[6:8] = 1e-3 * RNG.normal(size=2)
            vl[4], vr[4] = 0.05, 0.06
            if not tb.hlld_fallback_triggered(vl[:8], vr[:8], GM):
                continue
        try:
            fd_check(fwd, vjp, vl, vr, 9)
        except AssertionError:
            fd_check(fwd, vjp, vl, vr, 9, eps=1e-8)
        n += 1


def test_source_vjp_matches_finite_differences():
    for _ in range(100):
        rp, rn = RNG.uniform(0.2, 1.5, size=2)
        vp, vn = RNG.normal(size=3), RNG.normal(size=3)
        b = RNG.normal(size=3)
        pp, pn = RNG.uniform(0.2, 1.0, size=2)
        um = np.r_[rp, rp * vp, pp / (GM - 1) + 0.5 * rp * vp @ vp + 0.5 * b @ b, b]
        uh = np.r_[rn, rn * vn, pn / (GM - 1) + 0.5 * rn * vn @ vn]
        ac = RNG.uniform(1, 100)
        lam = RNG.normal(size=5)
        lam[0] = 0.0

        def S(m, h, a):
            out = np.zeros(5)
            L.hc_source(p(m), p(h), a, GM, p(out))
            return out
        gm_, gh_, ga_ = np.zeros(8), np.zeros(5), np.zeros(1)
        L.hc_source_vjp(p(um), p(uh), ac, GM, p(lam), p(gm_), p(gh_), p(ga_))
        for arr, g in ((um, gm_), (uh, gh_)):
            for k in range(len(arr)):
                h = 1e-6 * max(1, abs(arr[k]))
                ap, am = arr.copy(), arr.copy()
                ap[k] += h
                am[k] -= h
                if arr is um:
                    fd = lam @ (S(ap, uh, ac) - S(am, uh, ac)) / (2 * h)
                else:
                    fd = lam @ (S(um, ap, ac) - S(um, am, ac)) / (2 * h)
                assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
        h = 1e-6 * ac
        fd = lam @ (S(um, uh, ac + h) - S(um, uh, ac - h)) / (2 * h)
        assert abs(fd - ga_[0]) <= 1e-6 * max(1.0, abs(fd))
