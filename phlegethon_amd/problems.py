"""Initial-condition generators for the benchmark configurations and the
validation problems.

All random perturbations come from a counter-based generator
(splitmix64 of seed + zone index), so a CPU run and a GPU run of the same
problem start from identical bytes.  Geometry follows the reference's
Cartesian-uniform grid (source/source.F90:16271-16371): cell centres at
x_l + (i + 1/2) dx and nodes/faces at x_l + i dx (0-based i).

Each generator returns a ``Problem`` carrying the configuration, the
interior primitive state, the face-centred magnetic field (MHD) and the
plane-parallel static profiles (equilibrium, gravity, potential, heating).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from .config import (BC_PERIODIC, BC_REFLECTIVE, BC_OUTFLOW, EOS_HELMHOLTZ, EOS_PIG, HLLC_FLUX,
                     LHLLC_FLUX, HLLD_FLUX, LHLLD_FLUX, LIM2ND_REC, PPH_REC, PhlxConfig)

SEED = 20261202
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser on uint64 counters (Steele, Lea & Flood 2014)."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, stream: int, n: int) -> np.ndarray:
    """n doubles in [0,1) from counters (seed, stream, zone index)."""
    base = np.uint64((seed * 0x100000001B3 + stream * 0x9E3779B1) & 0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        ctr = (np.arange(n, dtype=np.uint64) + base) & _M64
    bits = splitmix64(ctr)
    return (bits >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclass
class Problem:
    name: str
    cfg: PhlxConfig
    prim: np.ndarray
    b_faces: Optional[List[np.ndarray]] = None
    profiles: Dict[str, np.ndarray] = field(default_factory=dict)
    temp: Optional[np.ndarray] = None      # initial T (Helmholtz warm start)
    helm: Optional[object] = None          # helmholtz.HelmTable
    kappa: Optional[np.ndarray] = None     # opacity on cells -1..n (one ghost layer)

    def copy(self) -> "Problem":
        return Problem(self.name, self.cfg, self.prim.copy(),
                       None if self.b_faces is None else
                       [None if b is None else b.copy() for b in self.b_faces],
                       {k: v.copy() for k, v in self.profiles.items()},
                       None if self.temp is None else self.temp.copy(), self.helm,
                       None if self.kappa is None else self.kappa.copy())


def centres(cfg: PhlxConfig, d: int) -> np.ndarray:
    lo = (cfg.x1l, cfg.x2l, cfg.x3l)[d]
    n = (cfg.nx1, cfg.nx2, cfg.nx3)[d]
    return lo + (np.arange(n, dtype=np.float64) + 0.5) * cfg.dx[d]


def nodes(cfg: PhlxConfig, d: int) -> np.ndarray:
    lo = (cfg.x1l, cfg.x2l, cfg.x3l)[d]
    n = (cfg.nx1, cfg.nx2, cfg.nx3)[d]
    return lo + np.arange(n + 1, dtype=np.float64) * cfg.dx[d]


def mesh(cfg: PhlxConfig):
    """Cell-centre coordinate arrays shaped (nz, ny, nx)."""
    x = centres(cfg, 0)
    y = centres(cfg, 1)
    z = centres(cfg, 2) if cfg.sdims == 3 else np.array([0.0])
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


# --------------------------------------------------------------------- C1
def hot_bubble(n: int = 256, recon: int = LIM2ND_REC, flux: int = LHLLC_FLUX,
               etot: bool = False, seed: int = SEED, amp: float = 1e-3,
               noise: float = 1e-6, rk_stages: int = 2) -> Problem:
    """2D hydrostatic polytrope (index 1.5, rho0 = p0 = 1, g = -1 in x2) with a
    Gaussian entropy bump (amplitude amp, radius 0.1 at (0.5, 0.3)) and
    seeded uniform velocity noise; periodic x1, reflecting x2, WB gravity.
    Pattern of tests/hotbubble/app.F90:27-180 on the unit square."""
    cfg = PhlxConfig(sdims=2, nx1=n, nx2=n, nx3=1, recon=recon, flux=flux,
                     rk_stages=rk_stages, cfl=0.8, mach_cutoff=1e-4,
                     bc=(BC_PERIODIC, BC_PERIODIC, BC_REFLECTIVE, BC_REFLECTIVE, 0, 0),
                     use_wb=True, use_gravity=True, evolve_etot=etot).validate()
    gm = cfg.gamma
    yc = centres(cfg, 1)
    yf = nodes(cfg, 1)
    rho_cc = (1.0 - 0.4 * yc) ** 1.5
    p_cc = rho_cc ** gm
    rho_f = (1.0 - 0.4 * yf) ** 1.5
    p_f = rho_f ** gm
    X, Y, _ = mesh(cfg)
    r2 = (X - 0.5) ** 2 + (Y - 0.3) ** 2
    A = 1.0 + amp * np.exp(-r2 / 0.01)
    p = np.broadcast_to(p_cc[None, :, None], X.shape).copy()
    rho = (p / A) ** (1.0 / gm)
    N = X.size
    prim = np.empty((cfg.nvars,) + cfg.shape)
    prim[0] = rho
    prim[1] = (2.0 * uniform01(seed, 1, N) - 1.0).reshape(X.shape) * noise
    prim[2] = (2.0 * uniform01(seed, 2, N) - 1.0).reshape(X.shape) * noise
    prim[3] = p
    prof = dict(eq_rho_cc=rho_cc, eq_p_cc=p_cc, eq_rho_f=rho_f, eq_p_f=p_f,
                grav_cc=np.full(cfg.nx2, -1.0), phi_cc=yc.copy(), phi_f=yf.copy())
    return Problem("hot_bubble", cfg, prim, None, prof)


def hydrostatic_rest(n: int = 64, flux: int = LHLLC_FLUX, recon: int = LIM2ND_REC,
                     etot: bool = False, sdims: int = 2) -> Problem:
    """The C1 polytrope exactly at rest: the well-balanced scheme must keep it
    bitwise (face p_eq cancels in the momentum flux, rho' = 0)."""
    pb = hot_bubble(n=n, recon=recon, flux=flux, etot=etot, amp=0.0, noise=0.0)
    if sdims == 3:
        cfg = pb.cfg.with_(sdims=3, nx3=max(4, n // 8),
                           bc=(BC_PERIODIC, BC_PERIODIC, BC_REFLECTIVE, BC_REFLECTIVE,
                               BC_PERIODIC, BC_PERIODIC)).validate()
        prim = np.empty((cfg.nvars,) + cfg.shape)
        prim[0] = pb.prim[0][0][None]
        prim[1:4] = 0.0
        prim[4] = pb.prim[3][0][None]
        return Problem("hydrostatic_rest_3d", cfg, prim, None, pb.profiles)
    return Problem("hydrostatic_rest", pb.cfg, pb.prim, None, pb.profiles)


# --------------------------------------------------------------------- C2
def orszag_tang_3d(n: int = 256, flux: int = LHLLD_FLUX, recon: int = LIM2ND_REC,
                   seed: int = SEED, noise: float = 1e-4, rk_stages: int = 2,
                   shape: Optional[tuple] = None) -> Problem:
    """3D Orszag-Tang on the periodic unit cube: rho = 25/(36 pi)(1 + noise U),
    p = 5/(12 pi), v = (-sin 2pi y, sin 2pi x, 0.1 sin 2pi z); face B from the
    nodal vector potential A_z = B0/(2pi) cos 2pi y + B0/(4pi) cos 4pi x,
    B0 = 1/sqrt(4 pi) (tests/orszag-tang/app.F90:13-79), so div B = 0 to
    round-off."""
    nx, ny, nz = shape if shape is not None else (n, n, n)
    cfg = PhlxConfig(sdims=3, nx1=nx, nx2=ny, nx3=nz, recon=recon, flux=flux,
                     rk_stages=rk_stages, cfl=0.8, mach_cutoff=1e-5, mach_ct=1e-6,
                     bc=(0, 0, 0, 0, 0, 0)).validate()
    pi = math.pi
    B0 = 1.0 / math.sqrt(4.0 * pi)
    X, Y, Z = mesh(cfg)
    prim = np.empty((cfg.nvars,) + cfg.shape)
    u = uniform01(seed, 3, X.size).reshape(X.shape)
    prim[0] = 25.0 / (36.0 * pi) * (1.0 + noise * (2.0 * u - 1.0))
    prim[1] = -np.sin(2.0 * pi * Y)
    prim[2] = np.sin(2.0 * pi * X)
    prim[3] = 0.1 * np.sin(2.0 * pi * Z)
    prim[4] = 5.0 / (12.0 * pi)
    xn = nodes(cfg, 0)
    yn = nodes(cfg, 1)
    Az = (B0 / (2.0 * pi)) * np.cos(2.0 * pi * yn)[:, None] + \
         (B0 / (4.0 * pi)) * np.cos(4.0 * pi * xn)[None, :]          # (ny+1, nx+1)
    idx1, idx2 = 1.0 / cfg.dx[0], 1.0 / cfg.dx[1]
    bx = (Az[1:, :] - Az[:-1, :]) * idx2                               # (ny, nx+1)
    by = -(Az[:, 1:] - Az[:, :-1]) * idx1                              # (ny+1, nx)
    b1 = np.broadcast_to(bx[None], cfg.face_shape(0)).copy()
    b2 = np.broadcast_to(by[None], cfg.face_shape(1)).copy()
    b3 = np.zeros(cfg.face_shape(2))
    return Problem("orszag_tang_3d", cfg, prim, [b1, b2, b3], {})


def orszag_tang_2d(n: int = 128, flux: int = LHLLD_FLUX, recon: int = LIM2ND_REC,
                   rk_stages: int = 2) -> Problem:
    """2D Orszag-Tang (tests/orszag-tang/app.F90) on the periodic unit square."""
    cfg = PhlxConfig(sdims=2, nx1=n, nx2=n, recon=recon, flux=flux, rk_stages=rk_stages,
                     cfl=0.8, mach_ct=1e-6).validate()
    pi = math.pi
    B0 = 1.0 / math.sqrt(4.0 * pi)
    X, Y, _ = mesh(cfg)
    prim = np.empty((cfg.nvars,) + cfg.shape)
    prim[0] = 25.0 / (36.0 * pi)
    prim[1] = -np.sin(2.0 * pi * Y)
    prim[2] = np.sin(2.0 * pi * X)
    prim[3] = 5.0 / (12.0 * pi)
    xn, yn = nodes(cfg, 0), nodes(cfg, 1)
    Az = (B0 / (2.0 * pi)) * np.cos(2.0 * pi * yn)[:, None] + \
         (B0 / (4.0 * pi)) * np.cos(4.0 * pi * xn)[None, :]
    bx = (Az[1:, :] - Az[:-1, :]) / cfg.dx[1]
    by = -(Az[:, 1:] - Az[:, :-1]) / cfg.dx[0]
    return Problem("orszag_tang_2d", cfg, prim,
                   [bx[None].copy(), by[None].copy(), None], {})


# ------------------------------------------------------ validation problems
def field_loop(n: int = 64, vel=(2.0, 1.0), A0: float = 1e-3, R: float = 0.3,
               flux: int = LHLLD_FLUX, recon: int = LIM2ND_REC, mach_ct: float = 1e-6,
               tmax: float = 1e30) -> Problem:
    """Field-loop advection (Gardiner & Stone 2005, section 5.1): periodic
    [-1, 1] x [-0.5, 0.5] with 2n x n cells, rho = p = 1, uniform velocity
    `vel`, face B from the nodal A_z = A0 (R - r) inside r < R (zero
    outside), so div B = 0 to round-off; with vel = (2, 1) the loop is back
    at its start at t = 1."""
    cfg = PhlxConfig(sdims=2, nx1=2 * n, nx2=n, x1l=-1.0, x1u=1.0, x2l=-0.5, x2u=0.5,
                     recon=recon, flux=flux, cfl=0.8, mach_ct=mach_ct, tmax=tmax,
                     gamma=5.0 / 3.0).validate()
    prim = np.empty((cfg.nvars,) + cfg.shape)
    prim[0] = 1.0
    prim[1] = vel[0]
    prim[2] = vel[1]
    prim[3] = 1.0
    xn, yn = nodes(cfg, 0), nodes(cfg, 1)
    r = np.sqrt(xn[None, :] ** 2 + yn[:, None] ** 2)
    Az = np.where(r < R, A0 * (R - r), 0.0)
    bx = (Az[1:, :] - Az[:-1, :]) / cfg.dx[1]
    by = -(Az[:, 1:] - Az[:, :-1]) / cfg.dx[0]
    return Problem("field_loop", cfg, prim, [bx[None].copy(), by[None].copy(), None], {})


def sod_tube(n: int = 200, flux: int = HLLC_FLUX, recon: int = LIM2ND_REC,
             ny: int = 4, axis: int = 0) -> Problem:
    """Sod shock tube along x1 (or x2 with axis=1): (1, 0, 1) | (0.125, 0, 0.1),
    gamma = 1.4, outflow ends, periodic across."""
    nx1, nx2 = (n, ny) if axis == 0 else (ny, n)
    bc = [BC_PERIODIC] * 6
    bc[2 * axis] = bc[2 * axis + 1] = BC_OUTFLOW
    cfg = PhlxConfig(sdims=2, nx1=nx1, nx2=nx2, gamma=1.4, recon=recon, flux=flux,
                     cfl=0.8, mach_cutoff=1e-5, bc=tuple(bc)).validate()
    X, Y, _ = mesh(cfg)
    s = X if axis == 0 else Y
    left = s < 0.5
    prim = np.zeros((cfg.nvars,) + cfg.shape)
    prim[0] = np.where(left, 1.0, 0.125)
    prim[3] = np.where(left, 1.0, 0.1)
    return Problem("sod", cfg, prim, None, {})


def blast(cfg: PhlxConfig, seed: int = SEED, p_ratio: float = 20.0,
          radius: float = 0.2) -> Problem:
    """Pressure blast for the shock-flattening cases: p = p_ratio inside a
    sphere (circle) of the given radius around the box centre, 1 outside;
    rho = 1 + 5 % noise, small random velocities; a random divergence-free
    field of strength ~0.3 under MHD."""
    X, Y, Z = mesh(cfg)
    N = X.size
    u = lambda s: (2.0 * uniform01(seed, s, N) - 1.0).reshape(X.shape)
    c = [0.5 * (cfg.x1l + cfg.x1u), 0.5 * (cfg.x2l + cfg.x2u), 0.5 * (cfg.x3l + cfg.x3u)]
    r2 = (X - c[0]) ** 2 + (Y - c[1]) ** 2
    if cfg.sdims == 3:
        r2 = r2 + (Z - c[2]) ** 2
    prim = np.zeros((cfg.nvars,) + cfg.shape)
    prim[0] = 1.0 + 0.05 * u(1)
    for d in range(cfg.sdims):
        prim[1 + d] = 0.05 * u(2 + d)
    prim[cfg.sdims + 1] = np.where(r2 < radius * radius, p_ratio, 1.0)
    for s in range(cfg.sdims + 2, cfg.nvars):
        prim[s] = 0.5 + 0.2 * u(10 + s)
    bf = _random_div_free(cfg, seed) if cfg.mhd else None
    return Problem("blast", cfg, prim, bf, {})


def sound_wave(n: int = 64, amp: float = 1e-6, flux: int = HLLC_FLUX,
               recon: int = LIM2ND_REC, rk_stages: int = 2) -> Problem:
    """Right-going linear sound wave along x1 on a periodic box; rho0 = 1,
    p0 = 1/gamma (c = 1).  After t = 1 the exact solution equals the IC."""
    cfg = PhlxConfig(sdims=2, nx1=n, nx2=4, recon=recon, flux=flux, cfl=0.8,
                     rk_stages=rk_stages, tmax=1.0).validate()
    X, _, _ = mesh(cfg)
    d = amp * np.sin(2.0 * math.pi * X)
    prim = np.zeros((cfg.nvars,) + cfg.shape)
    prim[0] = 1.0 + d
    prim[1] = d
    prim[3] = 1.0 / cfg.gamma + d
    return Problem("sound_wave", cfg, prim, None, {})


def alfven_wave(n: int = 32, flux: int = HLLD_FLUX, recon: int = LIM2ND_REC,
                rk_stages: int = 2, amp: float = 0.1) -> Problem:
    """Circularly polarised Alfven wave along x1 (exact nonlinear solution):
    rho = 1, p = 0.1, B = (1, amp sin 2pi x, amp cos 2pi x),
    v = (0, -amp sin, -amp cos); travels at v_A = 1, period 1."""
    cfg = PhlxConfig(sdims=3, nx1=n, nx2=4, nx3=4, recon=recon, flux=flux, cfl=0.8,
                     rk_stages=rk_stages, tmax=1.0).validate()
    X, _, _ = mesh(cfg)
    s = np.sin(2.0 * math.pi * X)
    c = np.cos(2.0 * math.pi * X)
    prim = np.zeros((cfg.nvars,) + cfg.shape)
    prim[0] = 1.0
    prim[2] = -amp * s
    prim[3] = -amp * c
    prim[4] = 0.1
    b1 = np.ones(cfg.face_shape(0))
    xc = centres(cfg, 0)
    b2 = np.broadcast_to(amp * np.sin(2.0 * math.pi * xc)[None, None, :], cfg.face_shape(1)).copy()
    b3 = np.broadcast_to(amp * np.cos(2.0 * math.pi * xc)[None, None, :], cfg.face_shape(2)).copy()
    return Problem("alfven_wave", cfg, prim, [b1, b2, b3], {})


def random_smooth(cfg: PhlxConfig, seed: int = SEED, amp: float = 0.2) -> Problem:
    """Random smooth-ish state for parity tests (any dims / options): density
    and pressure positive, velocities O(amp), divergence-free face field
    from a random nodal vector potential when MHD."""
    X, Y, Z = mesh(cfg)
    N = X.size
    sh = X.shape
    u = lambda s: (2.0 * uniform01(seed, s, N) - 1.0).reshape(sh)
    tp = 2.0 * math.pi
    prim = np.zeros((cfg.nvars,) + cfg.shape)
    prim[0] = 1.0 + 0.3 * np.sin(tp * X) * np.cos(tp * Y) + 0.05 * u(1)
    for d in range(cfg.sdims):
        prim[1 + d] = amp * (np.sin(tp * (Y if d == 0 else X) + d) + 0.2 * u(2 + d))
    ip = cfg.sdims + 1
    prim[ip] = 1.0 + 0.2 * np.cos(tp * (X + Y)) + 0.05 * u(7)
    for s in range(cfg.sdims + 2, cfg.nvars):
        prim[s] = 0.5 + 0.2 * u(10 + s)
    bf = None
    if cfg.mhd:
        bf = _random_div_free(cfg, seed)
    return Problem("random_smooth", cfg, prim, bf, {})


def _random_div_free(cfg: PhlxConfig, seed: int, amp: float = 0.3):
    """Face field b = curl A with A on cell edges (discrete div b == 0).
    The modes are functions of coordinates normalised by each domain length,
    so they are periodic on any box; the differences are divided by the cell
    widths in units of the x1 length (one common scale, so the discrete
    divergence cancels on boxes of any aspect; an exact no-op on the unit
    box and on cubes)."""
    nz, ny, nx = cfg.shape
    lo = (cfg.x1l, cfg.x2l, cfg.x3l)
    ln = (cfg.x1u - cfg.x1l, cfg.x2u - cfg.x2l, cfg.x3u - cfg.x3l)
    dx1, dx2, dx3 = (cfg.dx[d] / ln[0] for d in range(3))
    tp = 2.0 * math.pi

    def nodes(cfg, d):  # noqa: F811  (normalised node / centre coordinates)
        return (_nodes(cfg, d) - lo[d]) / ln[d]

    def centres(cfg, d):  # noqa: F811
        return (_centres(cfg, d) - lo[d]) / ln[d]

    xn, yn = nodes(cfg, 0), nodes(cfg, 1)
    xc, yc = centres(cfg, 0), centres(cfg, 1)
    if cfg.sdims == 2:
        # A_z on nodes, periodic-compatible Fourier modes
        Az = amp * (np.cos(tp * yn)[:, None] / tp + 0.5 * np.sin(2 * tp * xn)[None, :] / tp
                    + 0.2 * np.sin(tp * (xn[None, :] + 2 * yn[:, None])) / tp)
        bx = (Az[1:, :] - Az[:-1, :]) / dx2
        by = -(Az[:, 1:] - Az[:, :-1]) / dx1
        return [bx[None].copy(), by[None].copy(), None]
    zn, zc = nodes(cfg, 2), centres(cfg, 2)
    # edge-centred potentials: Ax at (xc, yn, zn), Ay at (xn, yc, zn), Az at (xn, yn, zc)
    Ax = amp * np.sin(tp * (yn[None, :, None] + zn[:, None, None])) * \
        np.cos(tp * xc[None, None, :]) / tp
    Ay = amp * np.cos(tp * (zn[:, None, None] - xn[None, None, :])) * \
        (1 + 0 * yc[None, :, None]) / tp
    Az = amp * (np.cos(tp * yn)[None, :, None] / tp + 0.5 * np.sin(2 * tp * xn)[None, None, :] / tp
                + 0 * zc[:, None, None])
    # b = curl A
    b1 = (Az[:, 1:, :] - Az[:, :-1, :]) / dx2 - (Ay[1:, :, :] - Ay[:-1, :, :]) / dx3
    b2 = (Ax[1:, :, :] - Ax[:-1, :, :]) / dx3 - (Az[:, :, 1:] - Az[:, :, :-1]) / dx1
    b3 = (Ay[:, :, 1:] - Ay[:, :, :-1]) / dx1 - (Ax[:, 1:, :] - Ax[:, :-1, :]) / dx2
    return [np.ascontiguousarray(b1), np.ascontiguousarray(b2), np.ascontiguousarray(b3)]


_nodes, _centres = nodes, centres

# --------------------------------------------------------------------- C3
SLAB_L = 1.0e8  # cm


def slab_profiles(y: np.ndarray, L: float = SLAB_L):
    """log10 rho linear 6 -> 2, log10 T linear 9 -> 7 over [0, L]; abar a
    tanh step 4 -> 16 centred at L/2 (width L/10); zbar = abar / 2."""
    s = y / L
    rho = 10.0 ** (6.0 - 4.0 * s)
    T = 10.0 ** (9.0 - 2.0 * s)
    abar = 10.0 + 6.0 * np.tanh((s - 0.5) / 0.1)
    return rho, T, abar


def helm_slab(eos, table, n: int = 256, ny: Optional[int] = None, sdims: int = 3,
              recon: int = LIM2ND_REC, flux: int = LHLLC_FLUX, seed: int = SEED,
              noise: float = 1e-4, mhd_beta: Optional[float] = None,
              rk_stages: int = 2, fasteos: bool = True) -> Problem:
    """Stratified Helmholtz slab: plane-parallel equilibrium rho(y), T(y),
    abar(y) with P and e from the EoS, gravity from the discrete balance
    (p_f[j+1] - p_f[j]) / dy = rho_cc[j] g_j and the potential integrated
    from it; periodic x1/x3, reflecting x2; USE_FASTEOS + USE_WB +
    EVOLVE_ETOT + ADVECT_YE_IABAR.  Zones get a seeded density perturbation
    rho (1 + noise U(-1,1)) at fixed P (the step-0 T(rho, P) solve finds T).
    Pattern of tests/hotbubble-helmholtz/app.F90:33-191.

    ``eos`` evaluates the table (an object with ``full(rho, T, abar, zbar)``
    returning P, E, cs, cv): the GPU ``runtime.HelmEoS`` or, in CPU-only
    tests, the CPU restatement.  With ``mhd_beta`` set, LHLLD MHD with a weak
    random divergence-free field of plasma beta ~ mhd_beta.  ``fasteos``
    False: no USE_FASTEOS (the Riemann solvers make a T(rho, P) Newton per
    face state from the mean T of the two cells)."""
    ny = ny or n
    mhd = mhd_beta is not None
    cfg = PhlxConfig(sdims=sdims, nx1=n, nx2=ny, nx3=n if sdims == 3 else 1, nas=2,
                     x1l=0.0, x1u=SLAB_L * n / ny, x2l=0.0, x2u=SLAB_L,
                     x3l=0.0, x3u=SLAB_L * n / ny,
                     recon=recon, flux=(LHLLD_FLUX if mhd else flux), rk_stages=rk_stages,
                     cfl=0.8, mach_cutoff=1e-4,
                     bc=(BC_PERIODIC, BC_PERIODIC, BC_REFLECTIVE, BC_REFLECTIVE,
                         BC_PERIODIC, BC_PERIODIC),
                     use_wb=True, use_gravity=True, evolve_etot=True,
                     eos=EOS_HELMHOLTZ, fasteos=fasteos).validate()
    yc = centres(cfg, 1)
    yf = nodes(cfg, 1)
    dy = cfg.dx[1]
    rho_cc, T_cc, abar_cc = slab_profiles(yc)
    rho_f, T_f, abar_f = slab_profiles(yf)
    P_cc, E_cc, _, _ = (np.asarray(a) for a in eos.full(rho_cc, T_cc, abar_cc, abar_cc / 2))
    P_f, E_f, _, _ = (np.asarray(a) for a in eos.full(rho_f, T_f, abar_f, abar_f / 2))
    grav = (P_f[1:] - P_f[:-1]) / dy / rho_cc
    phi_f = np.zeros(ny + 1)
    phi_f[1:] = -np.cumsum(grav * dy)
    phi_cc = 0.5 * (phi_f[1:] + phi_f[:-1])
    # gamma_e = p / (rho e) + 1 (tests/hotbubble-helmholtz/app.F90:74-76)
    ge_cc = P_cc / (rho_cc * E_cc) + 1.0
    ge_f = P_f / (rho_f * E_f) + 1.0
    shape = cfg.shape
    N = int(np.prod(shape))
    u = (2.0 * uniform01(seed, 7, N) - 1.0).reshape(shape)
    prim = np.zeros((cfg.nvars,) + shape)
    prim[0] = rho_cc[None, :, None] * (1.0 + noise * u)
    ip = cfg.sdims + 1
    prim[ip] = np.broadcast_to(P_cc[None, :, None], shape)
    prim[ip + 1] = 0.5
    prim[ip + 2] = np.broadcast_to((1.0 / abar_cc)[None, :, None], shape)
    temp = np.broadcast_to(T_cc[None, :, None], shape).copy()
    prof = dict(eq_rho_cc=rho_cc, eq_p_cc=P_cc, eq_rho_f=rho_f, eq_p_f=P_f, grav_cc=grav,
                phi_cc=phi_cc, phi_f=phi_f, eq_gammae_cc=ge_cc, eq_gammae_f=ge_f)
    b_faces = None
    if mhd:
        b_faces = _random_div_free(cfg, seed, amp=1.0)
        # scale to plasma beta ~ mhd_beta at the mid-plane pressure
        bmag = max(np.sqrt(np.mean(b ** 2)) for b in b_faces if b is not None)
        scale = np.sqrt(2.0 * P_cc[ny // 2] / mhd_beta) / bmag
        b_faces = [None if b is None else b * scale for b in b_faces]
    return Problem("helm_slab", cfg, prim, b_faces, prof, temp, table)


def pig_box(eos, table, n: int = 32, sdims: int = 2, recon: int = LIM2ND_REC,
            flux: int = LHLLC_FLUX, seed: int = SEED, mhd_beta: Optional[float] = None,
            fasteos: bool = True, rho0: float = 1e-7, T0: float = 1.5e4,
            nas: int = 2) -> Problem:
    """Periodic box of partially ionised gas for PIG_EOS: rho and T vary by
    30 % across the box around (rho0, T0) (hydrogen about half ionised at
    the defaults), P and the step-0 T from the PIG table, seeded velocities
    of ~0.1 of the sound speed; with nas = 2 two advected scalars (1, 1)
    that the EoS ignores.  ``eos`` evaluates the table (``full(rho, T, abar,
    zbar)``)."""
    mhd = mhd_beta is not None
    L = 1e9
    cfg = PhlxConfig(sdims=sdims, nx1=n, nx2=n, nx3=n if sdims == 3 else 1, nas=nas,
                     x1u=L, x2u=L, x3u=L, recon=recon, flux=(LHLLD_FLUX if mhd else flux),
                     cfl=0.8, mach_cutoff=1e-4, eos=EOS_PIG, fasteos=fasteos).validate()
    X, Y, Z = mesh(cfg)
    tp = 2.0 * math.pi / L
    N = X.size
    u = lambda s: (2.0 * uniform01(seed, s, N) - 1.0).reshape(X.shape)
    rho = rho0 * (1.0 + 0.3 * np.sin(tp * X) * np.cos(tp * Y) + 0.02 * u(1))
    T = T0 * (1.0 + 0.3 * np.cos(tp * (X + Y)) + 0.02 * u(2))
    one = np.ones(N)
    P, E, cs, cv = (np.asarray(a).reshape(X.shape)
                    for a in eos.full(rho.ravel(), T.ravel(), one, one))
    prim = np.zeros((cfg.nvars,) + cfg.shape)
    prim[0] = rho
    for d in range(sdims):
        prim[1 + d] = 0.1 * cs * (np.sin(tp * (Y if d == 0 else X) + d) + 0.2 * u(3 + d))
    ip = sdims + 1
    prim[ip] = P
    if nas == 2:
        prim[ip + 1] = 1.0
        prim[ip + 2] = 1.0
    b_faces = None
    if mhd:
        b_faces = _random_div_free(cfg, seed, amp=1.0)
        bmag = max(np.sqrt(np.mean(b ** 2)) for b in b_faces if b is not None)
        scale = np.sqrt(2.0 * float(P.mean()) / mhd_beta) / bmag
        b_faces = [None if b is None else b * scale for b in b_faces]
    return Problem("pig_box", cfg, prim, b_faces, {}, temp=T.copy(), helm=table)


def state_box(pb: Problem, dex: float = 0.5) -> tuple:
    """Training box of the T(rho, e) surrogate around a Helmholtz problem's
    states: (log10 rhoYe lo, hi, log10 T lo, hi, abar lo, hi, Ye lo, hi),
    widened by ``dex`` decades in density and temperature, 10 % in abar and
    0.01 in Ye, for the states the flow may still reach."""
    ip = pb.cfg.sdims + 1
    rho, ye, iabar = pb.prim[0], pb.prim[ip + 1], pb.prim[ip + 2]
    lrs = np.log10(rho * ye)
    lt = np.log10(pb.temp)
    abar = 1.0 / iabar
    return (float(lrs.min()) - dex, float(lrs.max()) + dex,
            float(lt.min()) - dex, float(lt.max()) + dex,
            float(abar.min()) * 0.9, float(abar.max()) * 1.1,
            max(float(ye.min()) - 0.01, 0.05), min(float(ye.max()) + 0.01, 1.0))


# --------------------------------------------------------------------- C4
def lowmach_box(n: int = 384, heating: float = 1e-6, sdims: int = 3, recon: int = LIM2ND_REC,
                flux: int = LHLLC_FLUX, precision: int = 0, seed: int = SEED,
                noise: float = 1e-6, mach_cutoff: float = 1e-5,
                mach: Optional[float] = None) -> Problem:
    """Low-Mach convection box: ideal gas gamma = 5/3, constant gravity
    g = -1 in x2, the C1 polytrope rho = (1 - 0.4 y)^1.5, p = rho^gamma on
    the unit cube (periodic x1/x3, reflecting x2), heated at the bottom by
    edot = heating * sin(8 pi y) for y < 1/8 (pattern of
    tests/andrassy2022/app.F90:132-167), seeded density noise.  USE_WB +
    USE_GRAVITY + EVOLVE_ETOT + USE_EDOT, LHLLC, RK2, CFL 0.8.  The flow
    speed grows with the heating (Mach ~ heating^(1/3) once convective).

    ``mach`` starts the box in that convective state instead of at rest
    (reaching it from rest takes ~1/mach sound crossings): divergence-free
    rolls v1 = pi sin(2 pi x) cos(pi y) cos(2 pi z) A, v2 = -2 pi cos(2 pi x)
    sin(pi y) cos(2 pi z) A (v.n = 0 on the walls), A set for an RMS Mach
    number ``mach``, and heating = mach^3."""
    if mach is not None:
        heating = mach ** 3
    cfg = PhlxConfig(sdims=sdims, nx1=n, nx2=n, nx3=n if sdims == 3 else 1,
                     recon=recon, flux=flux, rk_stages=2, cfl=0.8, mach_cutoff=mach_cutoff,
                     bc=(BC_PERIODIC, BC_PERIODIC, BC_REFLECTIVE, BC_REFLECTIVE,
                         BC_PERIODIC, BC_PERIODIC),
                     use_wb=True, use_gravity=True, evolve_etot=True, use_edot=True,
                     precision=precision).validate()
    gm = cfg.gamma
    yc = centres(cfg, 1)
    yf = nodes(cfg, 1)
    rho_cc = (1.0 - 0.4 * yc) ** 1.5
    p_cc = rho_cc ** gm
    rho_f = (1.0 - 0.4 * yf) ** 1.5
    p_f = rho_f ** gm
    edot = np.where(yc < 0.125, heating * np.sin(8.0 * math.pi * yc), 0.0)
    shape = cfg.shape
    N = int(np.prod(shape))
    u = (2.0 * uniform01(seed, 11, N) - 1.0).reshape(shape)
    prim = np.zeros((cfg.nvars,) + shape)
    prim[0] = rho_cc[None, :, None] * (1.0 + noise * u)
    prim[cfg.sdims + 1] = np.broadcast_to(p_cc[None, :, None], shape)
    if mach is not None:
        # arrays are (z, y, x)
        x = centres(cfg, 0)[None, None, :]
        y = yc[None, :, None]
        z = centres(cfg, 2)[:, None, None] if sdims == 3 else np.zeros((1, 1, 1))
        cz = np.cos(2.0 * math.pi * z)
        v1 = math.pi * np.sin(2.0 * math.pi * x) * np.cos(math.pi * y) * cz
        v2 = -2.0 * math.pi * np.cos(2.0 * math.pi * x) * np.sin(math.pi * y) * cz
        v1, v2 = np.broadcast_to(v1, shape), np.broadcast_to(v2, shape)
        cs = np.sqrt(gm * prim[cfg.sdims + 1] / prim[0])
        amp = mach / np.sqrt(np.mean((v1 ** 2 + v2 ** 2) / cs ** 2))
        prim[1] = amp * v1
        prim[2] = amp * v2
    prof = dict(eq_rho_cc=rho_cc, eq_p_cc=p_cc, eq_rho_f=rho_f, eq_p_f=p_f,
                grav_cc=np.full(n, -1.0), phi_cc=yc.copy(), phi_f=yf.copy(), edot_cc=edot)
    return Problem("lowmach_box", cfg, prim, None, prof)
