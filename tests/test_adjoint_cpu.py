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
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hostcheck.cpp")
    deps = [src] + [os.path.join(build.CSRC, f) for f in ("pip_phys.h", "pip_adj_gen.h", "pip_adj_point.h",
                                                          "pip_bnd.h", "pip_adj_items.h")]
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < max(os.path.getmtime(f) for f in deps):
        os.makedirs(build.LIBDIR, exist_ok=True)
        subprocess.check_call(["g++", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-std=c++17",
                               "-I" + build.CSRC, "-o", LIB, src])
    L = ctypes.CDLL(LIB)
    D = ctypes.c_double
    L.hc_hllc.argtypes = [P, P, D, P]
    L.hc_hlld9.argtypes = [P, P, D, D, P]
    L.hc_hlld8.argtypes = [P, P, D, P]
    L.hc_hllc_vjp.argtypes = [P, P, D, P, P, P]
    L.hc_hlld9_vjp.argtypes = [P, P, D, D, P, P, P]
    L.hc_source.argtypes = [P, P, D, D, P]
    L.hc_source_vjp.argtypes = [P, P, D, D, P, P, P, P]
    L.hc_minmod.argtypes = [D, D, D]
    L.hc_minmod.restype = D
    L.hc_recon_grad.argtypes = [D, D, D, ctypes.c_int, P]
    for f in ("hc_u2v9", "hc_u2v_hd"):
        getattr(L, f).argtypes = [P, D, D, D, P]
    for f in ("hc_u2v9_vjp", "hc_u2v_hd_vjp"):
        getattr(L, f).argtypes = [P, P, D, D, D, P]
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


def test_minmod_equals_oracle_and_routing_matches_differences():
    """minmod (Util_rot.f90:59-76) bitwise vs the oracle; the face-state
    derivative routing (left: v1 + mm/2, right: v1 - mm/2) vs central
    differences away from kinks, and zero slope where a difference vanishes."""
    for _ in range(2000):
        v = RNG.normal(size=3)
        if RNG.uniform() < 0.2:
            v[RNG.integers(3)] = v[1]  # a vanishing difference
        assert L.hc_minmod(*v) == orc.minmod(*v)
        for right in (0, 1):
            g = np.zeros(3)
            L.hc_recon_grad(v[0], v[1], v[2], right, p(g))
            sgn = -0.5 if right else 0.5
            X, Y = v[2] - v[1], v[1] - v[0]
            if X * Y <= 0.0:
                assert np.array_equal(g, [0.0, 1.0, 0.0])
                continue
            if abs(abs(X) - abs(Y)) < 1e-6:
                continue  # kink between the two slopes
            for k in range(3):
                vp, vm = v.copy(), v.copy()
                vp[k] += 1e-8
                vm[k] -= 1e-8
                fd = ((vp[1] + sgn * L.hc_minmod(*vp)) - (vm[1] + sgn * L.hc_minmod(*vm))) / 2e-8
                assert abs(fd - g[k]) < 1e-6, (v, right, k, fd, g[k])


@pytest.mark.parametrize("floors", [False, True])
def test_u2v_vjp_matches_finite_differences(floors):
    """Transposes of u2v_mhd / u2v_hd (HLL_rot.f90:1136-1186); with floors
    active the floored outputs carry no derivative."""
    ro_lim, pr_lim = (0.3, 0.2) if floors else (1e-8, 1e-8)
    for _ in range(200):
        um = rand_mhd()
        uh = rand_hd()
        um = np.r_[um[0], um[0] * um[1:4], um[4] / (GM - 1) + 0.5 * um[0] * um[1:4] @ um[1:4]
                   + 0.5 * um[5:8] @ um[5:8], um[5:9]]
        uh = np.r_[uh[0], uh[0] * uh[1:4], uh[4] / (GM - 1) + 0.5 * uh[0] * uh[1:4] @ uh[1:4]]
        for u, n, fwd, vjp in ((um, 9, L.hc_u2v9, L.hc_u2v9_vjp), (uh, 5, L.hc_u2v_hd, L.hc_u2v_hd_vjp)):
            lv = RNG.normal(size=n)
            g = np.zeros(n)
            vjp(p(u), p(lv), GM, ro_lim, pr_lim, p(g))

            def V(x):
                out = np.zeros(n)
                fwd(p(x), GM, ro_lim, pr_lim, p(out))
                return out
            for k in range(n):
                h = 1e-7 * max(1.0, abs(u[k]))
                up, um_ = u.copy(), u.copy()
                up[k] += h
                um_[k] -= h
                fd = lv @ (V(up) - V(um_)) / (2 * h)
                if floors and (abs(u[0] - ro_lim) < 1e-5 or abs(V(u)[4] - pr_lim) < 1e-5):
                    continue  # kink at the floor inside the stencil
                assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_generated_vjps_are_the_generators_output(tmp_path):
    """pip_adj_gen.h is what gen_adjoint.py writes (the kernels and these
    checks compile the committed file)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("gen_adjoint", os.path.join(build.CSRC, "gen_adjoint.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    out = tmp_path / "pip_adj_gen.h"
    gen.main(str(out))
    assert out.read_text() == open(os.path.join(build.CSRC, "pip_adj_gen.h")).read()


@pytest.mark.parametrize("sm_target", [3.0e-11, 1.0e-9, 1.0e-4])
def test_hllc_switch_derivative_at_small_contact_speeds(sm_target):
    """The smoothed switch thm = (|S_M| + S_M) / max(2 S_M, 1e-10)
    (HLL_rot.f90:679-681) has derivative 2e10 inside its band and zero
    above it.  States whose contact speed S_M lies in the band or just above
    it: the VJP equals central differences with steps small enough to stay
    on one side of the band edge (to 1e-4: the steps are 1e-11 there)."""
    vl = np.array([1.0, 0.0, 0.2, -0.1, 1.0])
    vr = np.array([0.8, 0.0, -0.1, 0.3, 0.8])

    def sm_of(vr1):
        # HLLC contact speed of the pair (HLL_rot.f90:656-660 in the
        # primitive form of the generator), the wave speeds from the
        # sound speeds with the factor 1.03
        w = vr.copy()
        w[1] = vr1
        cl, cr = np.sqrt(GM * vl[4] / vl[0]), np.sqrt(GM * w[4] / w[0])
        sl = 1.03 * min(vl[1] - cl, w[1] - cr)
        sr = 1.03 * max(vl[1] + cl, w[1] + cr)
        num = w[0] * w[1] * (sr - w[1]) - vl[0] * vl[1] * (sl - vl[1]) + vl[4] - w[4]
        den = w[0] * (sr - w[1]) - vl[0] * (sl - vl[1])
        return num / den
    lo, hi = -1.0, 1.0  # bisect the right state's velocity for S_M = sm_target
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if sm_of(mid) < sm_target else (lo, mid)
    vr[1] = 0.5 * (lo + hi)
    assert abs(sm_of(vr[1]) - sm_target) <= 1e-3 * sm_target
    lam = RNG.normal(size=5)
    lvl, lvr = np.zeros(5), np.zeros(5)
    L.hc_hllc_vjp(p(vl), p(vr), GM, p(lam), p(lvl), p(lvr))
    # steps that keep S_M on its side of the band edge (5e-11); the
    # differences then carry round-off of ~1e-16 / h
    h = 1e-11 if sm_target < 1e-8 else 1e-8
    for side, v, g in (("l", vl, lvl), ("r", vr, lvr)):
        for k in range(5):
            vp, vm = v.copy(), v.copy()
            vp[k] += h
            vm[k] -= h
            fp = hllc(vp, vr) if side == "l" else hllc(vl, vp)
            fm = hllc(vm, vr) if side == "l" else hllc(vl, vm)
            fd = float(lam @ (fp - fm)) / (2 * h)
            assert abs(fd - g[k]) <= 1e-4 * max(1.0, abs(fd)), (side, k, fd, g[k])


def test_face_chain_vjp_matches_finite_differences():
    """The per-face chain of the adjoint face kernel: primitives of the four
    cells around a face -> minmod face states -> density floor -> HLLD (GLM)
    and HLLC fluxes, pulled back by the VJPs and the routing; checked against
    central differences of the composed forward map."""
    ch, ro_lim = 1.3, 1e-8

    def states(V):  # V: (4, 14) primitives of cells LL, L, R, RR
        vl = V[1] + 0.5 * np.array([L.hc_minmod(V[0, n], V[1, n], V[2, n]) for n in range(14)])
        vr = V[2] - 0.5 * np.array([L.hc_minmod(V[1, n], V[2, n], V[3, n]) for n in range(14)])
        vl[0], vr[0] = max(vl[0], ro_lim), max(vr[0], ro_lim)
        vl[9], vr[9] = max(vl[9], ro_lim), max(vr[9], ro_lim)
        return vl, vr

    def F(V):
        vl, vr = states(V)
        return np.r_[hlld9(vl[:9], vr[:9], ch), hllc(vl[9:], vr[9:])]

    checked = 0
    while checked < 60:
        base_m, base_h = rand_mhd(), rand_hd()
        V = np.array([np.r_[base_m, base_h] + 0.05 * RNG.normal(size=14) for _ in range(4)])
        V[:, [0, 4, 9, 13]] = np.abs(V[:, [0, 4, 9, 13]]) + 0.1
        lam = RNG.normal(size=14)
        vl, vr = states(V)
        lvl, lvr = np.zeros(14), np.zeros(14)
        L.hc_hlld9_vjp(p(vl[:9].copy()), p(vr[:9].copy()), GM, ch, p(lam[:9].copy()), p(lvl[:9]), p(lvr[:9]))
        a, b = np.zeros(5), np.zeros(5)
        L.hc_hllc_vjp(p(vl[9:].copy()), p(vr[9:].copy()), GM, p(lam[9:].copy()), p(a), p(b))
        lvl[9:], lvr[9:] = a, b
        G = np.zeros((4, 14))
        for n in range(14):
            gl, gr = np.zeros(3), np.zeros(3)
            L.hc_recon_grad(V[0, n], V[1, n], V[2, n], 0, p(gl))
            L.hc_recon_grad(V[1, n], V[2, n], V[3, n], 1, p(gr))
            G[0, n] += lvl[n] * gl[0]
            G[1, n] += lvl[n] * gl[1] + lvr[n] * gr[0]
            G[2, n] += lvl[n] * gl[2] + lvr[n] * gr[1]
            G[3, n] += lvr[n] * gr[2]
        ok = True
        for c in range(4):
            for n in range(14):
                h = 1e-7 * max(1.0, abs(V[c, n]))
                Vp, Vm = V.copy(), V.copy()
                Vp[c, n] += h
                Vm[c, n] -= h
                fd = lam @ (F(Vp) - F(Vm)) / (2 * h)
                if abs(fd - G[c, n]) > 1e-5 * max(1.0, abs(fd)):
                    ok = False  # a kink (limiter switch, wave-speed min/max) in the stencil
        if ok:
            checked += 1
        else:
            # the same point must pass with a much smaller step
            for c in range(4):
                for n in range(14):
                    h = 1e-10 * max(1.0, abs(V[c, n]))
                    Vp, Vm = V.copy(), V.copy()
                    Vp[c, n] += h
                    Vm[c, n] -= h
                    fd = lam @ (F(Vp) - F(Vm)) / (2 * h)
                    assert abs(fd - G[c, n]) <= 1e-4 * max(1.0, abs(fd)), (c, n, fd, G[c, n])
            checked += 1


def test_ghost_image_lists_invert_the_boundary_map():
    """The packed ghost-image lists used by the stage kernels' stores and by
    the adjoint gather are the exact inverse of the PIPbnd source map
    (Boundary_rot.f90:323-402) for every boundary type and for padded sizes
    up to the largest allowed, including corners in 2-D."""
    I = ctypes.c_int
    L.hc_src_index.argtypes = [I, I, I, I, ctypes.POINTER(I)]
    L.hc_src_index.restype = I
    L.hc_images.argtypes = [I, I, I, I, ctypes.POINTER(I), ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(I)]
    L.hc_images.restype = I
    sym = (ctypes.c_uint * 4)(0b01, 0b10, 0b100, 0b1000)
    out = (I * 96)()
    fl = I()
    for ix, jx in ((8, 8), (131, 9), (300, 260), (4100, 1), (32767, 12)):
        for tx in ((1, 1), (2, 2), (10, 10), (2, 10), (10, 1)):
            for ty in (((1, 1), (2, 10)) if jx > 1 else ((0, 0),)):
                types = (I * 4)(tx[0], tx[1], ty[0], ty[1])
                # expected: every ghost (g_i, g_j) -> source, composed x then y
                want = {}
                for gj in (range(jx) if jx > 1 else [0]):
                    for gi in range(ix):
                        gx = gi < 2 or gi >= ix - 2
                        gy = jx > 1 and (gj < 2 or gj >= jx - 2)
                        if not (gx or gy):
                            continue
                        m = 0
                        si, sj = gi, gj
                        if gx:
                            si = L.hc_src_index(gi, ix, tx[0], tx[1], ctypes.byref(fl))
                            m ^= sym[0 if gi < 2 else 1] if fl.value else 0
                        if gy:
                            sj = L.hc_src_index(gj, jx, ty[0], ty[1], ctypes.byref(fl))
                            m ^= sym[2 if gj < 2 else 3] if fl.value else 0
                        want.setdefault((si, sj), set()).add((gi, gj, m))
                got = {}
                cells = {c for c in want}
                # interior cells far from the edges have no images
                cells |= {(ix // 2, max(jx // 2, 0))} if ix > 8 else set()
                for (ci, cj) in cells:
                    n = L.hc_images(ci, cj, ix, jx, types, sym, out)
                    assert n <= 32
                    got[(ci, cj)] = {(out[3 * k], out[3 * k + 1], out[3 * k + 2]) for k in range(n)}
                for c in cells:
                    assert got.get(c, set()) == want.get(c, set()), (ix, jx, tx, ty, c)


def _smooth_state(cfg, seed):
    """Smooth two-fluid state with O(0.1) waves in every variable."""
    from pip_amd import problems
    rng = np.random.default_rng(seed)
    X, Y = problems._grid(cfg)
    kx, ky = 2 * np.pi / (cfg.xmax - cfg.xmin), 2 * np.pi / (cfg.ymax - cfg.ymin)

    def wave(a):
        ph = rng.uniform(0, 2 * np.pi, size=2)
        return a * (np.sin(kx * X + ph[0]) * np.cos(ky * Y + ph[1]))
    rp, rn = 0.4 + wave(0.05), 0.6 + wave(0.05)
    v = [wave(0.2) + 0.1 for _ in range(3)]
    vn = [vv + wave(0.05) for vv in v]
    B = [1.0 + wave(0.1), 0.5 + wave(0.1), 0.3 + wave(0.1)]
    pp, pn = 0.3 + wave(0.03), 0.4 + wave(0.03)
    um = problems.pv2cq_mhd(cfg, rp, *v, pp, *B, wave(0.01))
    uh = problems.pv2cq_hd(cfg, rn, *vn, pn)
    ac = 30.0 * np.exp(wave(0.5))
    return uh, um, ac


@pytest.mark.parametrize("bnd", [(1, 1, 1, 1), (2, 2, 1, 1), (1, 1, 3, 2)])
def test_adjoint_step_matches_oracle_differences(bnd):
    """One SSP-RK2 step: the host build of the adjoint work items (the code
    the HIP kernels run) against central differences of the oracle's
    run_solver with the same fixed dt and cmax, for every interior state
    entry and every alpha_c entry; periodic and mirror boundaries."""
    from pip_amd.config import PIPConfig
    cfg = PIPConfig(ndim=2, nx=10, ny=8, xmin=0.0, xmax=1.0, ymin=0.0, ymax=0.8, flag_divb=1,
                    ac_field=True, flag_bnd=bnd, col=30.0)
    uh, um, ac = _smooth_state(cfg, 3)
    o = orc.Oracle(cfg, uh, um, ac, prologue=False)
    o.pipbnd()
    uh, um = o.state()
    dt = 2.0e-3
    cmax = cfg.safety * min(cfg.dx, cfg.dy) / dt
    nvt, N = 14, cfg.ix * cfg.jx
    rng = np.random.default_rng(9)
    W = np.zeros((nvt, cfg.jx, cfg.ix))
    W[:, 2:-2, 2:-2] = rng.normal(size=(nvt, cfg.ny, cfg.nx))

    def fwd(uh_, um_, ac_, stage1=False):
        o.set_state(uh_, um_)
        lib_ac = np.ascontiguousarray(ac_)
        orc.lib().orc_set_ac(o.h, lib_ac.ctypes.data_as(P))
        o.pipbnd()
        o.set_scalars(dt, cmax)
        U1 = o.run_solver_stage1()
        h, m = o.state()
        return np.concatenate([m, h]), U1

    Unp1, U1 = fwd(uh, um, ac)
    J0 = float((W * Unp1).sum())
    Un = np.ascontiguousarray(np.concatenate([um, uh]))
    lam = W.copy()
    lam_ac = np.zeros((cfg.jx, cfg.ix))
    I = ctypes.c_int
    L.hc_adj_step.argtypes = [I, I, P, ctypes.POINTER(I), P, P, P, P, P]
    prm = np.array([cfg.gm, cfg.ro_lim, cfg.pr_lim, cfg.dx, cfg.dy, cfg.col, cfg.safety, cfg.db_clean, dt, cmax])
    fb = (I * 4)(*bnd)
    L.hc_adj_step(cfg.ix, cfg.jx, p(prm), fb, p(Un), p(np.ascontiguousarray(U1)), p(ac), p(lam), p(lam_ac))
    assert J0 == J0
    def central(n, j, i, h):
        Up, Um = Un.copy(), Un.copy()
        Up[n, j, i] += h
        Um[n, j, i] -= h
        Jp = float((W * fwd(Up[9:], Up[:9], ac)[0]).sum())
        Jm = float((W * fwd(Um[9:], Um[:9], ac)[0]).sum())
        return (Jp - Jm) / (2 * h)

    kinks = 0
    for n in range(nvt):
        for j in range(2, cfg.jx - 2):
            for i in range(2, cfg.ix - 2):
                h = 1e-6 * max(1.0, abs(Un[n, j, i]))
                fd = central(n, j, i, h)
                if abs(fd - lam[n, j, i]) < 1e-6 * max(1.0, abs(fd)):
                    continue
                # a limiter or wave-speed switch inside the stencil: the same
                # entry must agree once the step is below the switch distance
                kinks += 1
                fd = central(n, j, i, h / 10)
                assert abs(fd - lam[n, j, i]) < 1e-6 * max(1.0, abs(fd)), (n, j, i, fd, lam[n, j, i])
    assert kinks <= 3
    for j in range(2, cfg.jx - 2):
        for i in range(2, cfg.ix - 2):
            h = 1e-6 * ac[j, i]
            ap, am = ac.copy(), ac.copy()
            ap[j, i] += h
            am[j, i] -= h
            fd = (float((W * fwd(uh, um, ap)[0]).sum()) - float((W * fwd(uh, um, am)[0]).sum())) / (2 * h)
            assert abs(fd - lam_ac[j, i]) <= 1e-6 * max(abs(fd), 1e-3), (j, i, fd, lam_ac[j, i])
    # gradient support: interior only
    assert np.all(lam[:, :2] == 0) and np.all(lam[:, :, :2] == 0)
