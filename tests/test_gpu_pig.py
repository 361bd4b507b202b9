"""PIG_EOS on the GPU against the oracle on the generated partially ionised
hydrogen table: the batched evaluations (full, T(rho, e), T(rho, P)) to
1e-12 relative (in practice bit for bit), the hydro path on the PIG box
(FASTEOS and the per-face Newton, 2D and 3D MHD, with thermal diffusion) at
the FP64 parity bar, the mixed-precision path against the FP64 one, and
the entry points' rules (table kinds, no surrogate on PIG).
"""
import numpy as np
import pytest

from phlegethon_amd import piggen
from phlegethon_amd import problems as P
from phlegethon_amd.config import EOS_PIG, LHLLC_FLUX, HLLC_FLUX, LIM2ND_REC, PPH_REC, PhlxConfig

from test_gpu_parity import TOL, rel_l1

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def table():
    return piggen.build(**piggen.DEFAULT)


@pytest.fixture(scope="module")
def eos(table):
    from phlegethon_amd.runtime import HelmEoS
    e = HelmEoS(table)
    yield e
    e.close()


@pytest.fixture(scope="module")
def orc(table, oracle_mod):
    o = oracle_mod.OracleHelm(table)
    yield o
    o.close()


def _points(n, seed):
    rng = np.random.default_rng(seed)
    return 10 ** rng.uniform(-11.5, -0.5, n), 10 ** rng.uniform(3.6, 7.9, n)


def test_batch_parity(eos, orc):
    rho, T = _points(20000, 4)
    one = np.ones_like(rho)
    got = eos.full(rho, T, one, one)
    ref = orc.full(rho, T, one, one)
    for a, b, name in zip(got, ref, ("P", "E", "cs", "cv")):
        assert np.max(np.abs(a / b - 1.0)) <= 1e-12, name
    P_, E_ = ref[0], ref[1]
    for fn_g, fn_o, target in ((eos.rhoe_given, orc.rhoe_given, E_),
                               (eos.rhoP_given, orc.rhoP_given, P_)):
        g = fn_g(rho, target, one, one, T * 1.01)
        o = fn_o(rho, target, one, one, T * 1.01)
        assert np.max(np.abs(g[0] / o[0] - 1.0)) <= 1e-12
        assert np.max(np.abs(g[1] / o[1] - 1.0)) <= 1e-12


def _run_gpu(pb, nsteps, cfg=None):
    from phlegethon_amd.runtime import Simulation
    sim = Simulation(cfg or pb.cfg)
    sim.setup(pb)
    sim.init()
    sim.run(nsteps)
    got = sim.download()
    got.update(sim.scalars())
    got["divb"] = sim.divb_max()
    sim.close()
    return got


@pytest.mark.parametrize("case", [
    dict(n=32, sdims=2, recon=LIM2ND_REC, flux=LHLLC_FLUX),
    dict(n=24, sdims=2, recon=PPH_REC, flux=HLLC_FLUX, fasteos=False),
    dict(n=12, sdims=3, recon=LIM2ND_REC, mhd_beta=10.0),
    dict(n=10, sdims=3, recon=LIM2ND_REC, flux=LHLLC_FLUX, fasteos=False),
    dict(n=24, sdims=2, recon=LIM2ND_REC, flux=LHLLC_FLUX, nas=0),
    dict(n=10, sdims=3, recon=PPH_REC, flux=HLLC_FLUX, fasteos=False, nas=0),
    dict(n=10, sdims=3, recon=LIM2ND_REC, mhd_beta=10.0, nas=0),
], ids=["2d-fasteos", "2d-face-newton", "3d-mhd", "3d-face-newton", "2d-fasteos-nas0",
        "3d-face-newton-nas0", "3d-mhd-nas0"])
def test_hydro_parity(table, orc, oracle_mod, case):
    pb = P.pig_box(orc, table, **case)
    nsteps = 6
    print(f"pig box {case} {nsteps} steps", flush=True)
    got = _run_gpu(pb, nsteps)
    ref = oracle_mod.run_problem(pb, nsteps)
    assert got["step"] == ref["step"] == nsteps
    assert abs(got["time"] - ref["time"]) <= 1e-14 * abs(ref["time"])
    for v in range(pb.cfg.nvars):
        assert rel_l1(got["prim"][v], ref["prim"][v]) <= TOL, v
    for name in ("eint", "temp"):
        assert rel_l1(got[name], ref[name]) <= TOL, name
    if pb.cfg.fasteos:
        for c in range(2):
            assert rel_l1(got["gammaf"][c], ref["gammaf"][c]) <= TOL
    if pb.cfg.mhd:
        for d in range(3):
            assert rel_l1(got[f"b_x{d+1}"], ref[f"b_x{d+1}"]) <= TOL
        assert got["divb"] <= 1e-13 * max(1.0, float(np.abs(got["b_cc"]).max()) / pb.cfg.dx[0])


def test_hydro_parity_with_diffusion(table, orc, oracle_mod):
    pb = P.pig_box(orc, table, n=24, sdims=2)
    pb.cfg = pb.cfg.with_(thermal_diffusion=True)
    r = oracle_mod.OracleRun(pb)
    r.init()
    dt = r.scalar("dt")
    r.close()
    inv_dl2 = sum(1.0 / pb.cfg.dx[d] ** 2 for d in range(2))
    kap = 2.0 * 2.99792458e10 * inv_dl2 * (dt / 6.0) / (3.0 * float(pb.prim[0].min()))
    pb.kappa = np.full((1, pb.cfg.nx2 + 2, pb.cfg.nx1 + 2), kap)
    got = _run_gpu(pb, 4)
    ref = oracle_mod.run_problem(pb, 4)
    for v in range(pb.cfg.nvars):
        assert rel_l1(got["prim"][v], ref["prim"][v]) <= TOL, v
    for name in ("eint", "temp"):
        assert rel_l1(got[name], ref[name]) <= TOL, name


def test_mixed_precision_close_to_fp64(table, orc):
    pb = P.pig_box(orc, table, n=32, sdims=2)
    fp64 = _run_gpu(pb, 6)
    mixed = _run_gpu(pb, 6, pb.cfg.with_(precision=1))
    for v in range(pb.cfg.nvars):
        err = rel_l1(mixed["prim"][v], fp64["prim"][v])
        print(f"  mixed vs FP64 prim[{v}] rel L1 {err:.2e}", flush=True)
        assert err <= 1e-6 or v in (1, 2)
    # velocities: relative to the sound speed scale
    cs = np.sqrt(fp64["prim"][3] / fp64["prim"][0]).mean()
    for v in (1, 2):
        assert np.abs(mixed["prim"][v] - fp64["prim"][v]).mean() <= 1e-6 * cs


def test_entry_point_rules(table):
    import ctypes
    from phlegethon_amd.runtime import Simulation
    E_CONFIG = -2
    f9 = np.ascontiguousarray(table.f9)
    c4 = np.ascontiguousarray(table.c4)
    args = (f9.ctypes.data_as(ctypes.c_void_p), c4.ctypes.data_as(ctypes.c_void_p),
            table.nrho, table.nT, table.ldlo, table.ldhi, table.ltlo, table.lthi)
    sim = Simulation(PhlxConfig(sdims=2, nx1=8, nx2=8, nas=2, eos=EOS_PIG, fasteos=True))
    try:
        assert sim.lib.phlx_load_helm_table(sim._ctx, *args) == E_CONFIG
        assert sim.lib.phlx_load_pig_table(sim._ctx, *args) == 0
        assert sim.lib.phlx_set_eos_guess(sim._ctx, 1, 0.25, 0) == E_CONFIG   # surrogate
        rc = sim.lib.phlx_mlp_train(sim._ctx, 1 << 13, 1, 1024, 1e-3, 1, None, None, None)
        assert rc == E_CONFIG
    finally:
        sim.close()
