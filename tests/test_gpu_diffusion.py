"""THERMAL_DIFFUSION_STS on the GPU (phlx_diffusion.hip through the C-ABI)
against the oracle: one diffusion step (eint, T, p to 1e-12 relative L1,
dt_par and the stage count exact) on ideal-gas modes and random fields with
periodic, zero-gradient and fixed-temperature walls, on the Helmholtz EoS;
and full loops (diffusion / hydro / diffusion, the hyperbolic dt after
it) in 2D, 3D, MHD and on the Helmholtz slab at the parity bar of the
hydro tests; the entry points' argument and order rules.
"""
import numpy as np
import pytest

from phlegethon_amd import helmholtz as H
from phlegethon_amd import problems as P
from phlegethon_amd.config import (BC_OUTFLOW, BC_PERIODIC, BC_REFLECTIVE, LHLLD_FLUX,
                                   PPH_REC, PhlxConfig)

from test_gpu_parity import TOL, check_parity, rel_l1

pytestmark = pytest.mark.gpu

R = 8.31446261815324e7
C = 2.99792458e10


def _kappa(cfg, field_fn=None, seed=1):
    """Opacity with one ghost layer: periodic images on periodic axes."""
    rng = np.random.default_rng(seed)
    kin = 1.0 + 0.3 * rng.uniform(0, 1, cfg.shape) if field_fn is None else field_fn
    kap = kin
    for ax, d in ((2, 0), (1, 1), (0, 2)):
        if d < cfg.sdims:
            pad = [(0, 0)] * 3
            pad[ax] = (1, 1)
            kap = np.pad(kap, pad, mode="wrap" if cfg.bc[2 * d] == BC_PERIODIC else "edge")
    return kap


def _random_T(cfg, amp=0.1, seed=3, kscale=1.0):
    rng = np.random.default_rng(seed)
    T = 1e7 * (1.0 + amp * rng.uniform(-1, 1, cfg.shape))
    rho = 1.0 + 0.2 * rng.uniform(-1, 1, cfg.shape)
    prim = np.zeros((cfg.nvars,) + cfg.shape)
    prim[0] = rho
    prim[cfg.sdims + 1] = R * rho * T / cfg.mu
    return P.Problem("randT", cfg, prim, None, {}, kappa=kscale * _kappa(cfg, seed=seed))


def _one_step(pb, ratio, oracle_mod):
    """One diffusion step at dt = ratio * dt_par on both paths."""
    from phlegethon_amd.runtime import Simulation
    r = oracle_mod.OracleRun(pb)
    r.init()
    dtp = r.parabolic_dt()
    r.set_scalar("dt", ratio * dtp)
    r.diffusion_step(1)
    ref = r.state()
    r.close()
    sim = Simulation(pb.cfg)
    sim.setup(pb)
    sim.init()
    dtp_g = sim.parabolic_dt()
    sc = sim.scalars()
    sim.set_scalars(sc["time"], ratio * dtp_g, sc["smax"], sc["step"])
    sim.diffusion_step(1)
    s = sim.diffusion_stages()
    got = sim.download()
    sim.close()
    assert dtp_g == dtp, (dtp_g, dtp)
    return got, ref, s


@pytest.mark.parametrize("sdims,bc,fix", [
    (2, (0, 0, 0, 0, 0, 0), None),
    (3, (0, 0, 0, 0, 0, 0), None),
    (3, (BC_REFLECTIVE, BC_OUTFLOW, 0, 0, BC_OUTFLOW, BC_REFLECTIVE), None),
    (2, (BC_OUTFLOW, BC_OUTFLOW, BC_REFLECTIVE, BC_REFLECTIVE, 0, 0),
     (1.3e7, None, 0.8e7, 1.1e7, None, None)),
])
@pytest.mark.parametrize("ratio", [0.7, 40.0])
def test_step_parity(oracle_mod, sdims, bc, fix, ratio):
    cfg = PhlxConfig(sdims=sdims, nx1=20, nx2=16, nx3=12 if sdims == 3 else 1, x1u=1e8,
                     x2u=8e7, x3u=6e7, bc=bc, thermal_diffusion=True, fix_temperature=fix)
    pb = _random_T(cfg)
    print(f"[diffusion step] sdims={sdims} bc={bc} fix={fix} dt/dt_par={ratio}", flush=True)
    got, ref, s = _one_step(pb, ratio, oracle_mod)
    print(f"  stages {s}", flush=True)
    assert s == max(3, 1 + int(np.floor(0.5 * (np.sqrt(9 + 16 * ratio) - 1))))
    for name in ("eint", "temp"):
        assert rel_l1(got[name], ref[name]) <= TOL, name
    assert rel_l1(got["prim"][sdims + 1], ref["prim"][sdims + 1]) <= TOL
    # the state did change
    assert rel_l1(ref["eint"], pb.prim[sdims + 1] / (cfg.gamma - 1)) > 1e-6


@pytest.fixture(scope="module")
def table(tmp_path_factory):
    p = str(tmp_path_factory.mktemp("helm_diff") / "t.bin")
    H.generate(141, 81, 0.0, 7.0, 6.0, 6.0 + 80 * 10.0 / 270.0, path=p)
    return H.load(p)


def _slab_with_kappa(table, oracle_mod, ratio_target=8.0, **kw):
    pb = P.helm_slab(oracle_mod.OracleHelm(table), table, **kw)
    pb.cfg = pb.cfg.with_(thermal_diffusion=True)
    # kappa so that dt_hyp / dt_par is about ratio_target at the start
    r = oracle_mod.OracleRun(pb)
    r.init()
    dt = r.scalar("dt")
    r.close()
    rho_min = float(pb.prim[0].min())
    inv_dl2 = sum(1.0 / pb.cfg.dx[d] ** 2 for d in range(pb.cfg.sdims))
    kap0 = 2.0 * C * inv_dl2 * (dt / ratio_target) / (3.0 * rho_min)
    pb.kappa = _kappa(pb.cfg, field_fn=np.full(pb.cfg.shape, kap0))
    return pb


@pytest.mark.parametrize("fasteos", [True, False])
def test_step_parity_helmholtz(table, oracle_mod, fasteos):
    pb = _slab_with_kappa(table, oracle_mod, n=10, ny=16, sdims=3, fasteos=fasteos)
    got, ref, s = _one_step(pb, 12.0, oracle_mod)
    print(f"helmholtz diffusion step: {s} stages", flush=True)
    for name in ("eint", "temp"):
        assert rel_l1(got[name], ref[name]) <= TOL, name
    assert rel_l1(got["prim"][4], ref["prim"][4]) <= TOL
    if fasteos:
        for c in range(2):
            assert rel_l1(got["gammaf"][c], ref["gammaf"][c]) <= TOL


def _spot(cfg, amp=2.0):
    X, Y, Z = P.mesh(cfg)
    r2 = (X - 5e7) ** 2 + (Y - 5e7) ** 2 + ((Z - 5e7) ** 2 if cfg.sdims == 3 else 0.0)
    T = 1e7 * (1.0 + amp * np.exp(-r2 / (1.5e7) ** 2))
    prim = np.zeros((cfg.nvars,) + cfg.shape)
    prim[0] = 1.0
    prim[cfg.sdims + 1] = R * T / cfg.mu
    return prim


@pytest.mark.parametrize("sdims,evaluate", [(2, True), (3, False)])
def test_run_parity_hot_spot(oracle_mod, sdims, evaluate):
    cfg = PhlxConfig(sdims=sdims, nx1=32, nx2=24, nx3=16 if sdims == 3 else 1, x1u=1e8,
                     x2u=1e8, x3u=1e8, thermal_diffusion=True, eval_parabolic_dt=evaluate,
                     recon=PPH_REC)
    pb = P.Problem("spot", cfg, _spot(cfg), None, {}, kappa=_kappa(cfg) * 1e-5)
    check_parity(pb, 8)


def test_run_parity_walls_mhd(oracle_mod):
    cfg = PhlxConfig(sdims=3, nx1=16, nx2=16, nx3=12, x1u=1e8, x2u=1e8, x3u=1e8,
                     flux=LHLLD_FLUX, thermal_diffusion=True,
                     bc=(BC_PERIODIC, BC_PERIODIC, BC_REFLECTIVE, BC_REFLECTIVE,
                         BC_PERIODIC, BC_PERIODIC),
                     fix_temperature=(None, None, 1.5e7, 0.9e7, None, None))
    pb = P.random_smooth(cfg)
    pb.prim[0] = 1.0 + 0.1 * pb.prim[0]
    pb.prim[4] = R * pb.prim[0] * 1e7 * (1.0 + 0.1 * np.sin(np.arange(cfg.nx1) / 3.0)) / cfg.mu
    pb.prim[1:4] *= 1e7
    pb.b_faces = [b * 1e7 for b in pb.b_faces]
    pb.kappa = _kappa(cfg) * 1e-5
    check_parity(pb, 6)


def test_run_parity_helmholtz_slab(table, oracle_mod):
    pb = _slab_with_kappa(table, oracle_mod, n=10, ny=16, sdims=3)
    from phlegethon_amd.runtime import Simulation
    nsteps = 4
    sim = Simulation(pb.cfg)
    sim.setup(pb)
    sim.init()
    sim.run(nsteps)
    got = sim.download()
    got.update(sim.scalars())
    sim.close()
    ref = oracle_mod.run_problem(pb, nsteps)
    assert got["step"] == ref["step"] == nsteps
    assert abs(got["time"] - ref["time"]) <= 1e-14 * abs(ref["time"])
    for v in range(pb.cfg.nvars):
        assert rel_l1(got["prim"][v], ref["prim"][v]) <= TOL, v
    for name in ("eint", "temp"):
        assert rel_l1(got[name], ref[name]) <= TOL, name


def test_entry_point_rules():
    from phlegethon_amd.runtime import Simulation
    E_CONFIG, E_ARG, E_STATE = -2, -3, -5
    cfg = PhlxConfig(sdims=2, nx1=8, nx2=8, bc=(BC_OUTFLOW, BC_OUTFLOW, 0, 0, 0, 0))
    sim = Simulation(cfg)
    lib, ctx = sim.lib, sim._ctx
    try:
        k = np.ones((1, 10, 10))
        kp = k.ctypes.data_as(__import__("ctypes").c_void_p)
        assert lib.phlx_set_kappa(ctx, kp) == E_STATE          # before enabling diffusion
        assert lib.phlx_diffusion_step(ctx, 1) == E_STATE      # before init
        bad = np.array([np.nan, np.nan, 1e7, np.nan, np.nan, np.nan])   # x2 is periodic
        assert lib.phlx_set_thermal_diffusion(
            ctx, 1, 1, bad.ctypes.data_as(__import__("ctypes").c_void_p)) == E_ARG
        assert lib.phlx_set_thermal_diffusion(ctx, 2, 1, None) == E_ARG
        assert lib.phlx_set_thermal_diffusion(ctx, 1, 1, None) == 0
        kneg = -k
        assert lib.phlx_set_kappa(ctx, kneg.ctypes.data_as(__import__("ctypes").c_void_p)) == E_ARG
        assert lib.phlx_set_kappa(ctx, kp) == 0
        assert lib.phlx_diffusion_step(ctx, 3) == E_ARG
    finally:
        sim.close()
    sim = Simulation(PhlxConfig(sdims=2, nx1=8, nx2=8, nas=2))
    try:
        assert sim.lib.phlx_set_thermal_diffusion(sim._ctx, 1, 1, None) == E_CONFIG
        assert sim.lib.phlx_parabolic_dt(sim._ctx, None) == E_CONFIG
    finally:
        sim.close()
