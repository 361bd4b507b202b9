"""The fused one-pass sweep (k_sweep: a 16 x 15 x1-x2 tile plus halo staged
in LDS, reconstruction -> Riemann -> flux divergence -> RK combination in
one kernel per stage, the x3 faces from a register window) against the
three directional sweeps and the multi-kernel update it replaces, and
against the CPU oracle.

Bar: the fused and the separate path are bit for bit equal (every stored
field: primitives, eint, T, cell and face B, conserved state) after N
steps, in FP64 and in the mixed-precision path, since every face state,
face solve and divergence is the same arithmetic in the same order; and
the FP64 fused path meets the parity bar of tests/test_gpu_parity.py
against the oracle (the mixed path's own bar against the oracle is
tests/test_gpu_mixed.py's, whose 3D hydro cases with x1 a multiple of 16
run the fused sweep, the default without MHD).  Cases cover hydro and MHD,
LIM2ND and PPH, HLLC / LHLLC / HLLD / LHLLD, periodic, reflecting, outflow
and diode walls, WB + gravity + ETOT + EDOT, RK2 and RK3, the Helmholtz
FASTEOS states (with and without MHD, with the Newton-counting update), a
grid whose x3 extent is not a multiple of the x3 chunk, x2 extents that
leave a partial 15-row tile (down to 6 rows), and grids the fused path
does not take (x1 not a multiple of 16), where it must stand aside.  The C5-shaped mixed slab runs with the carried-T, the energy-
predictor and the surrogate's Newton start (the network trained in each
run: training is deterministic).
"""
import numpy as np
import pytest

from phlegethon_amd import problems as P
from phlegethon_amd.config import (BC_DIODE, BC_OUTFLOW, BC_PERIODIC, BC_REFLECTIVE,
                                   HLLC_FLUX, HLLD_FLUX, LHLLC_FLUX, LHLLD_FLUX,
                                   LIM2ND_REC, PPH_REC, PREC_MIXED, PhlxConfig)

pytestmark = pytest.mark.gpu

TOL = 1e-12


def conserved(st, cfg):
    p = st["prim"]
    rho = p[0]
    out = [rho]
    ke = np.zeros_like(rho)
    for d in range(cfg.sdims):
        out.append(rho * p[1 + d])
        ke = ke + 0.5 * rho * p[1 + d] ** 2
    E = st["eint"] + ke
    if cfg.mhd:
        E = E + 0.5 * (st["b_cc"] ** 2).sum(axis=0)
    out.append(E)
    for s in range(cfg.sdims + 2, cfg.nvars):
        out.append(rho * p[s])
    return out


def rel_l1(a, b):
    den = np.abs(b).sum()
    return np.abs(a - b).sum() / (den if den > 0 else 1.0)


def _run(pb, nsteps, fused, count=False, guess="warm"):
    from phlegethon_amd.runtime import Simulation
    sim = Simulation(pb.cfg)
    sim.setup(pb)
    sim.set_fused_sweep(fused)
    if guess == "mlp":
        # the training is deterministic: both runs get the same network
        sim.mlp_train(nsamples=1 << 18, epochs=2, batch=2048, lr=5e-3, box=P.state_box(pb))
    if count:
        sim.set_eos_guess(guess, count=True)
    sim.profile(True)
    sim.init()
    sim.run(nsteps)
    st = sim.download()
    st.update(sim.scalars())
    st["divb"] = sim.divb_max() if pb.cfg.mhd else 0.0
    st["swept"] = sim.profile_read("sweep")[1]
    st["x1_sweeps"] = sim.profile_read("flux_x1")[1]
    if count:
        st["eos"] = sim.eos_stats()
    sim.close()
    return st


def _check_same(a, b):
    for k, v in b.items():
        if isinstance(v, np.ndarray):
            assert np.array_equal(a[k], v), k
    assert a["time"] == b["time"] and a["step"] == b["step"]


def _compare(pb, nsteps, oracle_mod=None, expect_fused=True, count=False, guess="warm"):
    print(f"[sweep] {pb.name} {pb.cfg.shape} flux={pb.cfg.flux} recon={pb.cfg.recon} "
          f"bc={tuple(pb.cfg.bc)} rk={pb.cfg.rk_stages} precision={pb.cfg.precision} "
          f"steps={nsteps}", flush=True)
    fused = _run(pb, nsteps, True, count, guess)
    split = _run(pb, nsteps, False, count, guess)
    assert split["swept"] == 0 and split["x1_sweeps"] > 0
    if expect_fused:
        assert fused["swept"] == nsteps * pb.cfg.rk_stages and fused["x1_sweeps"] == 0
    else:
        assert fused["swept"] == 0
    _check_same(fused, split)
    if count:
        assert fused["eos"] == split["eos"]
    if oracle_mod is not None and pb.cfg.precision != PREC_MIXED:
        ref = oracle_mod.run_problem(pb, nsteps, nthreads=16)
        errs = [rel_l1(g, r) for g, r in zip(conserved(fused, pb.cfg), conserved(ref, pb.cfg))]
        assert max(errs) <= TOL, errs
        assert fused["step"] == ref["step"]
        if pb.cfg.mhd:
            for d in range(3):
                k = f"b_x{d+1}"
                assert rel_l1(fused[k], ref[k]) <= TOL, k
    if pb.cfg.mhd:
        scale = max(1.0, float(np.abs(fused["b_cc"]).max()) / min(pb.cfg.dx[:3]))
        assert fused["divb"] < 1e-13 * scale, fused["divb"]
    return fused


def test_orszag_tang_3d(oracle_mod):
    _compare(P.orszag_tang_3d(n=32), 6, oracle_mod)


@pytest.mark.parametrize("flux", [HLLD_FLUX, LHLLD_FLUX])
def test_mhd_walls_ragged_x3(oracle_mod, flux):
    # x3 = 40: one full 32-plane chunk and one of 8; x1 != x2
    cfg = PhlxConfig(sdims=3, nx1=32, nx2=16, nx3=40, flux=flux, recon=LIM2ND_REC,
                     all_mach=(flux == LHLLD_FLUX),
                     bc=(BC_PERIODIC, BC_PERIODIC, BC_REFLECTIVE, BC_REFLECTIVE,
                         BC_OUTFLOW, BC_OUTFLOW))
    _compare(P.random_smooth(cfg), 6, oracle_mod)


@pytest.mark.parametrize("flux,recon", [(HLLC_FLUX, PPH_REC), (LHLLC_FLUX, LIM2ND_REC)])
def test_hydro_scalars_diode_rk3(oracle_mod, flux, recon):
    cfg = PhlxConfig(sdims=3, nx1=16, nx2=32, nx3=20, nas=2, flux=flux, recon=recon,
                     rk_stages=3,
                     bc=(BC_OUTFLOW, BC_OUTFLOW, BC_DIODE, BC_DIODE, BC_PERIODIC, BC_PERIODIC))
    _compare(P.random_smooth(cfg), 6, oracle_mod)


def test_lowmach_box_gravity_heating(oracle_mod):
    # WB + gravity + ETOT + EDOT, convective rolls at Mach 1e-2
    _compare(P.lowmach_box(n=32, mach=1e-2, heating=1e-6), 8, oracle_mod)


@pytest.mark.parametrize("bc2", [BC_REFLECTIVE, BC_PERIODIC])
def test_short_x2(oracle_mod, bc2):
    # nx2 = 6: one partial x2 tile of 15 rows, the rows past the ghosts
    # clamped to the last ghost row
    cfg = PhlxConfig(sdims=3, nx1=16, nx2=6, nx3=8, flux=LHLLD_FLUX, recon=LIM2ND_REC,
                     bc=(BC_PERIODIC, BC_PERIODIC, bc2, bc2, BC_PERIODIC, BC_PERIODIC))
    _compare(P.random_smooth(cfg), 6, oracle_mod)


def test_not_taken_when_tiles_do_not_fit():
    cfg = PhlxConfig(sdims=3, nx1=24, nx2=16, nx3=16, flux=LHLLD_FLUX, recon=LIM2ND_REC)
    _compare(P.random_smooth(cfg), 3, expect_fused=False)


@pytest.fixture(scope="module")
def slab_table():
    from phlegethon_amd import helmholtz as H
    import os
    import tempfile
    p = os.path.join(tempfile.mkdtemp(), "t.bin")
    H.generate(141, 81, 0.0, 7.0, 6.0, 6.0 + 80 * 10.0 / 270.0, path=p)
    return H.load(p)


@pytest.mark.parametrize("mhd", [False, True])
def test_helmholtz_fasteos_slab(oracle_mod, slab_table, mhd):
    orc = oracle_mod.OracleHelm(slab_table)
    # nx1 a multiple of 64: the Newton-counting update needs full blocks
    pb = P.helm_slab(orc, slab_table, n=64, ny=16, sdims=3,
                     mhd_beta=(1e5 if mhd else None))
    st = _compare(pb, 6, oracle_mod, count=True)
    its = st["eos"]["newton_iters"] / st["eos"]["solves"]
    print(f"[sweep] Helmholtz mhd={mhd}: Newton iterations/solve {its:.3f}", flush=True)


def _mixed(pb):
    pb.cfg = pb.cfg.with_(precision=PREC_MIXED)
    return pb


def test_mixed_orszag_tang_3d():
    _compare(_mixed(P.orszag_tang_3d(n=32)), 6)


def test_mixed_lowmach_box():
    _compare(_mixed(P.lowmach_box(n=32, mach=1e-3, heating=1e-9)), 8)


@pytest.mark.parametrize("guess", ["warm", "pred", "mlp"])
def test_mixed_helmholtz_mhd_slab(oracle_mod, slab_table, guess):
    orc = oracle_mod.OracleHelm(slab_table)
    pb = _mixed(P.helm_slab(orc, slab_table, n=64, ny=16, sdims=3, mhd_beta=1e5))
    st = _compare(pb, 6, count=True, guess=guess)
    its = st["eos"]["newton_iters"] / st["eos"]["solves"]
    print(f"[sweep] mixed Helmholtz MHD, {guess} start: Newton iterations/solve {its:.3f}",
          flush=True)


def test_default_choice():
    """The library's choice (mode -1): fused without MHD (FP64 and mixed),
    split with MHD."""
    from phlegethon_amd.runtime import Simulation
    cases = [(_mixed(P.lowmach_box(n=16, mach=1e-3, heating=1e-9)), True),
             (_mixed(P.orszag_tang_3d(n=16)), False),
             (P.lowmach_box(n=16, mach=1e-3, heating=1e-9), True),
             (P.orszag_tang_3d(n=16), False)]
    for pb, fused in cases:
        sim = Simulation(pb.cfg)
        sim.setup(pb)
        sim.set_fused_sweep(None)
        sim.profile(True)
        sim.init()
        sim.run(1)
        swept = sim.profile_read("sweep")[1]
        sim.close()
        assert (swept > 0) == fused, (pb.name, pb.cfg.precision, swept)
