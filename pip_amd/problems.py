"""Initial states for the benchmark configurations (synthetic, seeded).

Each builder returns ``(cfg, U_h, U_m, ac)`` with the state in the reference
layout: numpy arrays of shape ``(nvar, jx, ix)`` in C order, which is the
memory order of the Fortran ``U(ix,jx,kx=1,nvar)`` (Globalvar.f90:17).
Only interior cells are meaningful; ghost cells are filled by the boundary
map when the state is loaded (pip_init / the oracle prologue).

Primitive -> conserved conversion follows ``pv2cq_mhd``/``pv2cq_hd``
(Scheme_rot.f90:30-79) and the neutral/plasma split of the reference ICs
(shock_tube.f90:27-38).
"""
from __future__ import annotations

import math
from typing import Optional, Tuple

import numpy as np

from .config import PIPConfig, split_fractions

Arrays = Tuple[PIPConfig, np.ndarray, np.ndarray, Optional[np.ndarray]]


def pv2cq_mhd(cfg: PIPConfig, de, vx, vy, vz, pr, bx, by, bz, psi=None) -> np.ndarray:
    """pv2cq_mhd, flag_b_stg=0 branch (Scheme_rot.f90:68-77)."""
    gm = cfg.gm
    U = np.zeros((cfg.nvar_m, cfg.jx, cfg.ix))
    U[0] = de
    U[1] = de * vx
    U[2] = de * vy
    U[3] = de * vz
    U[4] = (pr / (gm - 1.0) + (0.5 * de) * ((vx * vx + vy * vy) + vz * vz)) \
        + 0.5 * ((bx * bx + by * by) + bz * bz)
    U[5] = bx
    U[6] = by
    U[7] = bz
    if cfg.nvar_m == 9:
        U[8] = 0.0 if psi is None else psi
    return U


def pv2cq_hd(cfg: PIPConfig, de, vx, vy, vz, pr) -> np.ndarray:
    """pv2cq_hd (Scheme_rot.f90:30-40)."""
    gm = cfg.gm
    U = np.zeros((5, cfg.jx, cfg.ix))
    U[0] = de
    U[1] = de * vx
    U[2] = de * vy
    U[3] = de * vz
    U[4] = pr / (gm - 1.0) + (0.5 * de) * ((vx * vx + vy * vy) + vz * vz)
    return U


def _grid(cfg: PIPConfig):
    x = cfg.x_centers()
    y = cfg.y_centers()
    X = np.broadcast_to(x[None, :], (cfg.jx, cfg.ix)).copy()
    Y = np.broadcast_to(y[:, None], (cfg.jx, cfg.ix)).copy()
    return X, Y


def _two_fluid(cfg, xi_n, rho, p, vx, vy, vz, bx, by, bz, psi=None):
    f_n, f_p, f_p_n, f_p_p = split_fractions(xi_n)
    U_h = pv2cq_hd(cfg, f_n * rho, vx, vy, vz, f_p_n * p)
    U_m = pv2cq_mhd(cfg, f_p * rho, vx, vy, vz, f_p_p * p, bx, by, bz, psi)
    return U_h, U_m


# --------------------------------------------------------------------------
# SplitMix64 -> Box-Muller (project RNG; the Fortran intrinsic RNG used by the
# reference ICs is compiler-specific, KH.f90:28-35)
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, n: int) -> np.ndarray:
    with np.errstate(over="ignore"):
        state = np.uint64(seed) + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
        z = state
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int) -> np.ndarray:
    """Uniform in (0,1): top 53 bits of SplitMix64, offset by half an ulp."""
    z = splitmix64(seed, n)
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def normal(seed: int, n: int) -> np.ndarray:
    m = (n + 1) // 2
    u = uniform01(seed, 2 * m)
    u1, u2 = u[:m], u[m:]
    r = np.sqrt(-2.0 * np.log(u1))
    g = np.concatenate([r * np.cos(2.0 * math.pi * u2), r * np.sin(2.0 * math.pi * u2)])
    return g[:n]


# --------------------------------------------------------------------------
def brio_wu(nx: int = 1024, xi_n: float = 0.5, col: float = 100.0,
            p_left: float = 1.0, p_right: float = 0.1,
            rho_left: float = 1.0, rho_right: float = 0.125) -> Arrays:
    """1-D two-fluid Brio-Wu shock tube on [0,1], step at x=0.5, Bx=0.75,
    By=+1/-1, outflow (flag_bnd=10) both sides."""
    cfg = PIPConfig(ndim=1, nx=nx, xmin=0.0, xmax=1.0, col=col,
                    flag_bnd=(10, 10, 10, 10))
    X, _ = _grid(cfg)
    left = X < 0.5
    rho = np.where(left, rho_left, rho_right)
    p = np.where(left, p_left, p_right)
    by = np.where(left, 1.0, -1.0)
    z = np.zeros_like(X)
    U_h, U_m = _two_fluid(cfg, xi_n, rho, p, z, z, z, z + 0.75, by, z)
    return cfg, U_h, U_m, None


def orszag_tang(n: int = 1024, xi_n: float = 0.9, col: float = 1.0e3) -> Arrays:
    """2-D two-fluid Orszag-Tang on the periodic box [0,2pi]^2:
    rho=gm^2, p=gm, v=(-sin y, sin x), B=(-sin y, sin 2x), psi=0."""
    cfg = PIPConfig(ndim=2, nx=n, ny=n, xmin=0.0, xmax=2.0 * math.pi,
                    ymin=0.0, ymax=2.0 * math.pi, col=col, flag_divb=1,
                    flag_bnd=(1, 1, 1, 1))
    X, Y = _grid(cfg)
    gm = cfg.gm
    z = np.zeros_like(X)
    U_h, U_m = _two_fluid(cfg, xi_n, z + gm * gm, z + gm, -np.sin(Y), np.sin(X), z,
                          -np.sin(Y), np.sin(2.0 * X), z)
    return cfg, U_h, U_m, None


def kelvin_helmholtz(n: int = 2048, xi_n: float = 0.8, col: float = 1.0e2,
                     seed: int = 20241202, width: float = 0.1) -> Arrays:
    """2-D two-fluid KH with set_IR case 3 rates, flag_IR_type=0:
    x in [0,1] periodic, y in [-0.5,0.5] reflecting (flag_bnd=2);
    vx=0.5 tanh(y/0.05), Bx=0.1, rho=1, p=1/gm, vy = 1e-3 N(0,1) exp(-y^2/2w^2)."""
    cfg = PIPConfig(ndim=2, nx=n, ny=n, xmin=0.0, xmax=1.0, ymin=-0.5, ymax=0.5,
                    col=col, flag_divb=1, flag_ir=3, flag_ir_type=0, flag_col=1,
                    flag_bnd=(1, 1, 2, 2))
    X, Y = _grid(cfg)
    z = np.zeros_like(X)
    g = normal(seed, cfg.ix * cfg.jx).reshape(cfg.jx, cfg.ix)
    vy = 1.0e-3 * g * np.exp(-(Y * Y) / (2.0 * width * width))
    vx = 0.5 * np.tanh(Y / 0.05)
    U_h, U_m = _two_fluid(cfg, xi_n, z + 1.0, z + 1.0 / cfg.gm, vx, vy, z,
                          z + 0.1, z, z)
    return cfg, U_h, U_m, None


def smooth_field(cfg: PIPConfig, seed: int, filt: float = 8.0) -> np.ndarray:
    """Unit-variance, periodic Gaussian-filtered white noise on the interior,
    returned on the padded grid with periodic ghosts."""
    nx, ny = cfg.nx, cfg.ny
    w = normal(seed, nx * ny).reshape(ny, nx)
    kx = np.fft.fftfreq(nx)[None, :]
    ky = np.fft.fftfreq(ny)[:, None]
    sig = filt / (2.0 * math.pi)  # filter width in cells
    h = np.exp(-0.5 * (2.0 * math.pi) ** 2 * sig ** 2 * (kx ** 2 + ky ** 2))
    g = np.real(np.fft.ifft2(np.fft.fft2(w) * h))
    g = (g - g.mean()) / g.std()
    out = np.zeros((cfg.jx, cfg.ix))
    out[2:-2, 2:-2] = g
    return periodic_fill(out)


def periodic_fill(a: np.ndarray) -> np.ndarray:
    """Periodic ghost fill of a padded 2-D plane (Boundary_rot.f90:342-357)."""
    a = a.copy()
    a[:, 0:2] = a[:, -4:-2]
    a[:, -2:] = a[:, 2:4]
    a[0:2, :] = a[-4:-2, :]
    a[-2:, :] = a[2:4, :]
    return a


def alfven_damping(n: int = 512, xi_n: float = 0.5, amp: float = 1.0e-3,
                   ac_mean: float = 50.0, ac_sigma: float = 0.5, seed: int = 11,
                   phase: float = 0.3) -> Arrays:
    """2-D linear Alfven wave on periodic [0,1]^2, rho=1, p=0.1, B0=(1,0.5),
    polarised along z, with an oblique wave vector k = 2 pi (1, 1) so that
    no state variable of the wave is uniform along a grid direction (an
    exactly uniform direction puts every minmod limiter of that sweep at its
    non-differentiable point X = Y = 0), and a phase offset so that the zero
    crossings do not fall midway between cell centres (there |X| = |Y|
    exactly, a tie between the two one-sided slopes); per-cell log-normal
    alpha_c field
    ac = exp(ln(mean) - sigma^2/2 + sigma*g)."""
    cfg = PIPConfig(ndim=2, nx=n, ny=n, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0,
                    col=ac_mean, flag_divb=1, ac_field=True, flag_bnd=(1, 1, 1, 1))
    X, Y = _grid(cfg)
    z = np.zeros_like(X)
    f_n, f_p, _, _ = split_fractions(xi_n)
    s = np.sin(2.0 * math.pi * (X + Y) + phase)
    dvz = amp * s
    dbz = -amp * math.sqrt(f_p * 1.0) * s
    U_h, U_m = _two_fluid(cfg, xi_n, z + 1.0, z + 0.1, z, z, dvz, z + 1.0, z + 0.5, dbz)
    g = smooth_field(cfg, seed)
    ac = np.exp(math.log(ac_mean) - 0.5 * ac_sigma ** 2 + ac_sigma * g)
    return cfg, U_h, U_m, ac


def pn_coupling(nx: int = 16, xi_n: float = 0.5, col: float = 10.0, T_n: float = 2.0,
                T_p: float = 0.5, v0: float = 0.0, gm: float = 5.0 / 3.0) -> Arrays:
    """Uniform 0-D temperature relaxation (PNcoupling.f90:109-129): T_n=2,
    T_p=0.5, optional drift v_n=v0/2, v_p=-v0/2, B=0, periodic 1-D box."""
    cfg = PIPConfig(ndim=1, nx=nx, xmin=0.0, xmax=2.0, col=col, gm=gm,
                    flag_bnd=(1, 1, 1, 1))
    z = np.zeros((cfg.jx, cfg.ix))
    ro_h = z + 1.0 * xi_n
    ro_m = z + 1.0 * (1.0 - xi_n)
    p_h = ro_h * T_n / gm
    p_m = 2.0 * ro_m * T_p / gm
    U_h = pv2cq_hd(cfg, ro_h, z + 0.5 * v0, z, z, p_h)
    U_m = pv2cq_mhd(cfg, ro_m, z - 0.5 * v0, z, z, p_m, z, z, z)
    return cfg, U_h, U_m, None


def ensemble_members(n_members: int = 256, seed: int = 7):
    """Member parameters of the batched 1-D shock-tube ensemble:
    nx = floor(256*16^u), xi_n ~ U[0.1,0.99], ac = 10^(4u), p_L/p_R =
    10^(log10 2 + u log10 50) with p_L = 1."""
    u = uniform01(seed, 4 * n_members).reshape(4, n_members)
    nx = np.floor(256.0 * 16.0 ** u[0]).astype(int)
    nx = np.clip(nx, 256, 4096)
    xi_n = 0.1 + 0.89 * u[1]
    ac = 10.0 ** (4.0 * u[2])
    ratio = 10.0 ** (math.log10(2.0) + u[3] * math.log10(50.0))
    return [dict(nx=int(nx[m]), xi_n=float(xi_n[m]), col=float(ac[m]),
                 p_left=1.0, p_right=float(1.0 / ratio[m])) for m in range(n_members)]


def ensemble_member(params) -> Arrays:
    return brio_wu(nx=params["nx"], xi_n=params["xi_n"], col=params["col"],
                   p_left=params["p_left"], p_right=params["p_right"])


def ensemble_members_2d(n_members: int = 64, seed: int = 9):
    """Member parameters of a batched 2-D ensemble (Orszag-Tang states with
    psi): n x n cells with n = floor(32*8^u) in [32, 256], xi_n ~ U[0.1,0.99],
    ac = 10^(4u) -- a sweep over neutral fraction and collision coefficient
    on uneven grids."""
    u = uniform01(seed, 3 * n_members).reshape(3, n_members)
    n = np.clip(np.floor(32.0 * 8.0 ** u[0]).astype(int), 32, 256)
    xi_n = 0.1 + 0.89 * u[1]
    ac = 10.0 ** (4.0 * u[2])
    return [dict(n=int(n[m]), xi_n=float(xi_n[m]), col=float(ac[m])) for m in range(n_members)]


def ensemble_member_2d(params) -> Arrays:
    return orszag_tang(params["n"], xi_n=params["xi_n"], col=params["col"])
