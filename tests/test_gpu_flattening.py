"""USE_SHOCK_FLATTENING on the GPU (k_sf_detect / k_sf_dilate and the
flattened k_flux) against the oracle: the FP64 bar of the parity tests
(relative L1 <= 1e-12 per conserved variable, same steps and time, face
field and div B under CT) on problems whose flags cover part of the grid
(the fraction is printed), across reconstructions, flux families,
boundaries, WB gravity and the Helmholtz EoS; and the entry point's
argument rules.
"""
import numpy as np
import pytest

from phlegethon_amd import helmholtz as H
from phlegethon_amd import problems as P
from phlegethon_amd.config import (BC_DIODE, BC_OUTFLOW, BC_PERIODIC, BC_REFLECTIVE,
                                   HLLC_FLUX, HLLD_FLUX, LHLLC_FLUX, LHLLD_FLUX,
                                   LIM2ND_REC, LIM5TH_REC, PPH_REC, PhlxConfig)

from test_gpu_parity import TOL, check_parity

pytestmark = pytest.mark.gpu


def _flag_fraction(oracle_mod, pb, nsteps=0):
    """Fraction of flagged cells (ghost layer included) after nsteps."""
    r = oracle_mod.OracleRun(pb)
    try:
        if nsteps:
            r.init()
            r.run(nsteps)
        return float(r.shock_flags().mean())
    finally:
        r.close()


def _flat(pb, eps=0.25):
    q = pb.copy()
    q.cfg = pb.cfg.with_(shock_flattening=True, eps_sf=eps)
    return q


@pytest.mark.parametrize("recon,flux,axis", [(LIM5TH_REC, HLLC_FLUX, 0),
                                             (PPH_REC, LHLLC_FLUX, 1),
                                             (LIM2ND_REC, HLLC_FLUX, 0)])
def test_sod(oracle_mod, recon, flux, axis):
    pb = _flat(P.sod_tube(n=96, flux=flux, recon=recon, axis=axis))
    print(f"flag fraction {_flag_fraction(oracle_mod, pb):.3f}", flush=True)
    check_parity(pb, 30)


@pytest.mark.parametrize("recon,flux,bc", [
    (PPH_REC, LHLLC_FLUX, (BC_REFLECTIVE,) * 4 + (0, 0)),
    (LIM5TH_REC, HLLC_FLUX, (BC_PERIODIC, BC_PERIODIC, BC_OUTFLOW, BC_OUTFLOW, 0, 0)),
    (LIM2ND_REC, LHLLC_FLUX, (BC_DIODE, BC_DIODE, BC_REFLECTIVE, BC_OUTFLOW, 0, 0)),
])
def test_blast_2d(oracle_mod, recon, flux, bc):
    cfg = PhlxConfig(sdims=2, nx1=40, nx2=36, recon=recon, flux=flux, bc=bc, nas=2,
                     all_mach=(bc[0] == BC_DIODE))
    pb = _flat(P.blast(cfg))
    print(f"flag fraction {_flag_fraction(oracle_mod, pb):.3f}", flush=True)
    check_parity(pb, 15)


@pytest.mark.parametrize("recon,flux,rk", [(LIM5TH_REC, HLLC_FLUX, 2), (PPH_REC, LHLLC_FLUX, 3)])
def test_blast_3d(oracle_mod, recon, flux, rk):
    cfg = PhlxConfig(sdims=3, nx1=20, nx2=16, nx3=12, recon=recon, flux=flux, rk_stages=rk,
                     bc=(BC_REFLECTIVE, BC_REFLECTIVE, BC_PERIODIC, BC_PERIODIC,
                         BC_OUTFLOW, BC_OUTFLOW))
    pb = _flat(P.blast(cfg))
    print(f"flag fraction {_flag_fraction(oracle_mod, pb):.3f}", flush=True)
    check_parity(pb, 8)


@pytest.mark.parametrize("recon,flux", [(PPH_REC, LHLLD_FLUX), (LIM5TH_REC, HLLD_FLUX)])
def test_blast_mhd(oracle_mod, recon, flux):
    cfg = PhlxConfig(sdims=3, nx1=16, nx2=16, nx3=12, recon=recon, flux=flux,
                     bc=(BC_PERIODIC, BC_PERIODIC, BC_REFLECTIVE, BC_REFLECTIVE,
                         BC_PERIODIC, BC_PERIODIC))
    pb = _flat(P.blast(cfg))
    print(f"flag fraction {_flag_fraction(oracle_mod, pb):.3f}", flush=True)
    check_parity(pb, 8)


def test_orszag_tang_2d(oracle_mod):
    # uniform pressure at the start: the flags appear during the run
    pb = _flat(P.orszag_tang_2d(n=48, flux=HLLD_FLUX, recon=LIM5TH_REC), eps=0.3)
    frac = _flag_fraction(oracle_mod, pb, 20)
    print(f"flag fraction at the last step {frac:.3f}", flush=True)
    assert _flag_fraction(oracle_mod, pb) == 0.0 and 0.05 < frac < 0.95
    check_parity(pb, 20)


def test_hot_bubble_wb(oracle_mod):
    # the detector sees the full (equilibrium + deviation) pressure: the
    # stratification flags part of the box at this threshold
    pb = _flat(P.hot_bubble(n=48, recon=PPH_REC), eps=0.05)
    frac = _flag_fraction(oracle_mod, pb)
    print(f"flag fraction {frac:.3f}", flush=True)
    assert 0.05 < frac < 0.95
    check_parity(pb, 15)


@pytest.fixture(scope="module")
def table(tmp_path_factory):
    p = str(tmp_path_factory.mktemp("helm_sf") / "t.bin")
    H.generate(141, 81, 0.0, 7.0, 6.0, 6.0 + 80 * 10.0 / 270.0, path=p)
    return H.load(p)


@pytest.mark.parametrize("case", [
    dict(n=10, ny=24, sdims=3, recon=LIM5TH_REC, flux=LHLLC_FLUX),
    dict(n=16, ny=24, sdims=2, recon=PPH_REC, flux=HLLC_FLUX, fasteos=False),
    dict(n=8, ny=24, sdims=3, recon=LIM2ND_REC, mhd_beta=1e3),
], ids=["3d-lim5th", "2d-pph-face-newton", "3d-mhd"])
def test_helm_slab(table, oracle_mod, case):
    """The slab's stratification alone stays below eps_sf = 1.5; a band of
    five times the pressure across the middle third of x1 is flattened."""
    from phlegethon_amd.runtime import Simulation
    pb = P.helm_slab(oracle_mod.OracleHelm(table), table, **case)
    ip, nx = pb.cfg.sdims + 1, pb.cfg.nx1
    pb.prim[ip][..., nx // 3: 2 * nx // 3] *= 5.0
    pb = _flat(pb, eps=1.5)
    frac = _flag_fraction(oracle_mod, pb)
    assert 0.05 < frac < 0.95
    print(f"helm slab {case}: flag fraction {frac:.3f}", flush=True)
    nsteps = 5
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
        a, b = got["prim"][v], ref["prim"][v]
        r = np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300)
        assert r <= TOL, (v, r)
    np.testing.assert_allclose(got["temp"], ref["temp"], rtol=1e-11)


def test_entry_point_rules():
    from phlegethon_amd.runtime import PhlxError, Simulation
    E_CONFIG, E_ARG = -2, -3
    sim = Simulation(PhlxConfig(sdims=2, nx1=8, nx2=8))
    try:
        f = sim.lib.phlx_set_shock_flattening
        for enable, eps in ((1, 0.0), (1, -1.0), (1, float("nan")), (1, float("inf")),
                            (2, 0.25), (-1, 0.25)):
            assert f(sim._ctx, enable, eps) == E_ARG, (enable, eps)
        assert f(None, 1, 0.25) == E_ARG
        assert f(sim._ctx, 1, 0.3) == 0
        assert f(sim._ctx, 0, 0.0) == 0          # off needs no threshold
    finally:
        sim.close()
    # FP64 only; every axis at least 3 cells
    for cfg in (PhlxConfig(sdims=2, nx1=8, nx2=8, precision=1),
                PhlxConfig(sdims=3, nx1=8, nx2=8, nx3=2)):
        sim = Simulation(cfg)
        try:
            assert sim.lib.phlx_set_shock_flattening(sim._ctx, 1, 0.25) == E_CONFIG
            with pytest.raises(PhlxError):
                sim.set_shock_flattening(True, 0.25)
        finally:
            sim.close()
