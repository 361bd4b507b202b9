"""GPU parity of the batched Helmholtz EoS (phlx_helm_* through the C-ABI)
against the CPU oracle on the same table and the same seeded inputs.

Bar: the device code follows the oracle's operation order with no
contraction, so outputs agree bit for bit except where a log10 rounding
difference moves a point into the neighbouring table cell (only at cell
boundaries); the test asserts max relative difference <= 1e-12 for every
output (the FP64 bar) and reports the bitwise-equal fraction.
"""
import os

import numpy as np
import pytest

from phlegethon_amd import helmholtz as H

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def table(tmp_path_factory):
    # box of the stratified-slab workload: rhoYe 1e0..1e7, T 1e6..~1e9
    p = str(tmp_path_factory.mktemp("helm_gpu") / "t.bin")
    H.generate(141, 81, 0.0, 7.0, 6.0, 6.0 + 80 * 10.0 / 270.0, path=p)
    return H.load(p)


@pytest.fixture(scope="module")
def eos(table):
    from phlegethon_amd.runtime import HelmEoS
    e = HelmEoS(table)
    yield e
    e.close()


@pytest.fixture(scope="module")
def orc(table, oracle_mod):
    return oracle_mod.OracleHelm(table)


def inputs(n, seed):
    rng = np.random.default_rng(seed)
    rho = 10 ** rng.uniform(0.5, 6.5, n)
    T = 10 ** rng.uniform(6.3, 8.8, n)
    abar = rng.uniform(4.0, 16.0, n)
    zbar = abar * rng.uniform(0.45, 0.5, n)
    return rho, T, abar, zbar


def relmax(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def report(name, a, b):
    eq = np.mean(np.asarray(a) == np.asarray(b))
    print(f"{name}: bitwise-equal {eq:.4f}  max rel {relmax(a, b):.2e}", flush=True)


def test_full_batch(eos, orc):
    rho, T, abar, zbar = inputs(20000, 1)
    got = eos.full(rho, T, abar, zbar)
    ref = orc.full(rho, T, abar, zbar)
    for name, a, b in zip(("P", "E", "cs", "cv"), got, ref):
        report(name, a, b)
        assert relmax(a, b) <= TOL, name


def test_rhoe_given(eos, orc):
    rho, T, abar, zbar = inputs(20000, 2)
    _, E, _, _ = orc.full(rho, T, abar, zbar)
    T0 = T * np.random.default_rng(3).uniform(0.7, 1.4, T.size)
    for ret in (True, False):
        got = eos.rhoe_given(rho, E, abar, zbar, T0, ret=ret)
        ref = orc.rhoe_given(rho, E, abar, zbar, T0, ret=ret)
        for name, a, b in zip(("T", "P", "cs"), got[:3], ref[:3]):
            report(f"rhoe ret={ret} {name}", a, b)
            assert relmax(a, b) <= TOL, name
        np.testing.assert_array_equal(got[3], ref[3])
        np.testing.assert_allclose(got[0], T, rtol=1e-9)


def test_rhoP_given(eos, orc):
    rho, T, abar, zbar = inputs(20000, 4)
    P, _, _, _ = orc.full(rho, T, abar, zbar)
    T0 = T * np.random.default_rng(5).uniform(0.7, 1.4, T.size)
    for ret in (True, False):
        got = eos.rhoP_given(rho, P, abar, zbar, T0, ret=ret)
        ref = orc.rhoP_given(rho, P, abar, zbar, T0, ret=ret)
        for name, a, b in zip(("T", "E", "cs"), got[:3], ref[:3]):
            report(f"rhoP ret={ret} {name}", a, b)
            assert relmax(a, b) <= TOL, name
        np.testing.assert_array_equal(got[3], ref[3])


def test_reset_path_and_nodes(eos, orc, table):
    # targets beyond the table top: T resets to the start value
    th = table.T_nodes[-1]
    P, E, _, _ = orc.full(1e4, th, 12.0, 6.0)
    got = eos.rhoe_given([1e4, 1e4], [E[0] * 100.0, E[0] * 50.0], 12.0, 6.0, 1e8)
    ref = orc.rhoe_given([1e4, 1e4], [E[0] * 100.0, E[0] * 50.0], 12.0, 6.0, 1e8)
    np.testing.assert_array_equal(got[0], [1e8, 1e8])
    np.testing.assert_array_equal(got[3], ref[3])
    # points exactly on table nodes (the cell index comes from log10)
    rs = table.rho_nodes[10:130:7]
    Tn = table.T_nodes[5:75:4]
    R, TT = np.meshgrid(rs, Tn)
    rho = (R * 2.0).ravel()          # abar = 2 zbar -> rhoYe = rho / 2
    got = eos.full(rho, TT.ravel(), 12.0, 6.0)
    ref = orc.full(rho, TT.ravel(), 12.0, 6.0)
    for name, a, b in zip(("P", "E", "cs", "cv"), got, ref):
        report(f"nodes {name}", a, b)
        assert relmax(a, b) <= TOL, name


def test_entry_errors(table):
    import ctypes
    from phlegethon_amd.config import PhlxConfig
    from phlegethon_amd.runtime import Simulation
    sim = Simulation(PhlxConfig(sdims=2, nx1=2, nx2=2))
    lib = sim.lib
    z = ctypes.c_void_p(0)
    # no table loaded yet
    assert lib.phlx_helm_rhoT_full_batch(sim.ctx, 1, z, z, z, z, z, z, z, z) == -7
    sim.load_helm_table(table)
    # empty batch is a no-op; negative size and missing inputs are rejected
    assert lib.phlx_helm_rhoT_full_batch(sim.ctx, 0, z, z, z, z, z, z, z, z) == 0
    assert lib.phlx_helm_rhoT_full_batch(sim.ctx, -1, z, z, z, z, z, z, z, z) == -3
    assert lib.phlx_helm_rhoe_given_batch(sim.ctx, 4, z, z, z, z, z, z, z, 1, z) == -3
    # non-finite table entries are rejected
    bad = H.HelmTable(table.nrho, table.nT, table.ldlo, table.ldhi, table.ltlo, table.lthi,
                      table.f9.copy(), table.c4.copy())
    bad.f9[3, 4, 2] = np.nan
    with pytest.raises(Exception):
        sim.load_helm_table(bad)
    sim.close()


# ---------------------------------------------------------------- hydro path
def _slab(table, oracle_mod, **kw):
    from phlegethon_amd import problems as P
    return P.helm_slab(oracle_mod.OracleHelm(table), table, **kw)


def _gpu_run(pb, nsteps):
    from phlegethon_amd.runtime import Simulation
    sim = Simulation(pb.cfg)
    sim.setup(pb)
    sim.init()
    sim.run(nsteps)
    st = sim.download()
    st.update(sim.scalars())
    st["divb"] = sim.divb_max()
    sim.close()
    return st


def _rel_l1(a, b):
    den = np.abs(b).sum()
    return float(np.abs(a - b).sum() / (den if den > 0 else 1.0))


from phlegethon_amd.config import HLLC_FLUX, LHLLC_FLUX, LIM2ND_REC, LIM5TH_REC, PPH_REC  # noqa: E402


@pytest.mark.parametrize("case", [
    dict(n=12, sdims=3, recon=LIM2ND_REC, flux=LHLLC_FLUX),
    dict(n=12, sdims=3, recon=PPH_REC, flux=HLLC_FLUX),
    dict(n=24, ny=32, sdims=2, recon=LIM2ND_REC, flux=LHLLC_FLUX),
    dict(n=10, sdims=3, recon=LIM2ND_REC, mhd_beta=1e3),
    dict(n=10, sdims=3, recon=LIM5TH_REC, flux=LHLLC_FLUX),
    dict(n=10, sdims=3, recon=LIM2ND_REC, flux=LHLLC_FLUX, fasteos=False),
    dict(n=16, ny=24, sdims=2, recon=PPH_REC, flux=HLLC_FLUX, fasteos=False),
    dict(n=8, sdims=3, recon=LIM2ND_REC, mhd_beta=1e3, fasteos=False),
    dict(n=8, sdims=3, recon=LIM5TH_REC, flux=LHLLC_FLUX, fasteos=False),
], ids=["3d-lim2nd-lhllc", "3d-pph-hllc", "2d-lim2nd-lhllc", "3d-mhd-lhlld", "3d-lim5th",
        "3d-lhllc-face-newton", "2d-pph-hllc-face-newton", "3d-mhd-lhlld-face-newton",
        "3d-lim5th-face-newton"])
def test_helm_slab_parity(table, oracle_mod, case):
    nsteps = 6
    pb = _slab(table, oracle_mod, **case)
    print(f"helm slab {case} {nsteps} steps", flush=True)
    got = _gpu_run(pb, nsteps)
    ref = oracle_mod.run_problem(pb, nsteps)
    assert got["step"] == ref["step"] == nsteps
    assert abs(got["time"] - ref["time"]) <= 1e-14 * abs(ref["time"])
    for v in range(pb.cfg.nvars):
        r = _rel_l1(got["prim"][v], ref["prim"][v])
        print(f"  prim[{v}] rel L1 {r:.2e}")
        assert r <= TOL, v
    for name in ("eint", "temp"):
        assert _rel_l1(got[name], ref[name]) <= TOL, name
    if pb.cfg.fasteos:
        for c in range(2):
            assert _rel_l1(got["gammaf"][c], ref["gammaf"][c]) <= TOL, c
    if pb.cfg.mhd:
        for d in range(3):
            assert _rel_l1(got[f"b_x{d+1}"], ref[f"b_x{d+1}"]) <= TOL
        assert got["divb"] <= 1e-13 * max(1.0, float(np.abs(got["b_cc"]).max()) /
                                          pb.cfg.dx[0])


def test_full_length_slab_full_table(oracle_mod):
    """The Helmholtz slab at 32^3 for 100 steps with the full 541 x 271
    table (log rhoYe -12 .. 15, log T 3 .. 13)."""
    full = H.default_table()
    pb = _slab(full, oracle_mod, n=32, sdims=3, recon=LIM2ND_REC, flux=LHLLC_FLUX)
    nsteps = 100
    print(f"helm slab 32^3, full table {full.nrho}x{full.nT}, {nsteps} steps", flush=True)
    got = _gpu_run(pb, nsteps)
    ref = oracle_mod.run_problem(pb, nsteps, nthreads=16)
    assert got["step"] == ref["step"] == nsteps
    assert abs(got["time"] - ref["time"]) <= 1e-14 * abs(ref["time"])
    for v in range(pb.cfg.nvars):
        r = _rel_l1(got["prim"][v], ref["prim"][v])
        print(f"  prim[{v}] rel L1 {r:.2e}")
        assert r <= TOL, v
    for name in ("eint", "temp"):
        assert _rel_l1(got[name], ref[name]) <= TOL, name


def _conserved(st, cfg):
    p = st["prim"]
    rho = p[0]
    out = {"rho": rho}
    ke = np.zeros_like(rho)
    for d in range(cfg.sdims):
        out[f"m{d+1}"] = rho * p[1 + d]
        ke = ke + 0.5 * rho * p[1 + d] ** 2
    out["E"] = st["eint"] + ke + 0.5 * (st["b_cc"] ** 2).sum(axis=0)
    for s in range(cfg.sdims + 2, cfg.nvars):
        out[f"s{s}"] = rho * p[s]
    return out


def test_full_length_c5_slab_full_table(oracle_mod):
    """The configuration of the headline workload (MHD + Helmholtz slab,
    beta 1e5: LHLLD + CT + LIM2ND, FASTEOS + WB + ETOT + Ye / abar, RK2) on
    64 x 16 x 64 zones (x1 a multiple of 64, as the surrogate's wave-wide
    update needs) for 100 steps with the full 541 x 271 table, against one oracle
    run: the plain FP64 path at 1e-12 (face B and div B included), and the
    mixed path with the surrogate's Newton start (the network trained in the
    run) at the 1e-6 bar per conserved field and per face-B component
    (relative to the L1 of the whole face field)."""
    from phlegethon_amd import problems as P
    from phlegethon_amd.config import PREC_MIXED
    from phlegethon_amd.runtime import Simulation
    full = H.default_table()
    pb = _slab(full, oracle_mod, n=64, ny=16, sdims=3, recon=LIM2ND_REC, mhd_beta=1e5)
    nsteps = 100
    print(f"C5-shaped slab {pb.cfg.shape}, full table, {nsteps} steps", flush=True)
    ref = oracle_mod.run_problem(pb, nsteps, nthreads=16)
    got = _gpu_run(pb, nsteps)
    assert got["step"] == ref["step"] == nsteps
    assert abs(got["time"] - ref["time"]) <= 1e-14 * abs(ref["time"])
    for v in range(pb.cfg.nvars):
        assert _rel_l1(got["prim"][v], ref["prim"][v]) <= TOL, v
    for name in ("eint", "temp"):
        assert _rel_l1(got[name], ref[name]) <= TOL, name
    for d in range(3):
        assert _rel_l1(got[f"b_x{d+1}"], ref[f"b_x{d+1}"]) <= TOL, d
    assert got["divb"] <= 1e-13 * max(1.0, float(np.abs(got["b_cc"]).max()) / pb.cfg.dx[0])

    mixed = pb.copy()
    mixed.cfg = pb.cfg.with_(precision=PREC_MIXED)
    sim = Simulation(mixed.cfg)
    sim.setup(mixed)
    sim.mlp_train(nsamples=1 << 18, epochs=2, batch=2048, lr=5e-3, box=P.state_box(mixed))
    sim.set_eos_guess("mlp")
    sim.init()
    sim.run(nsteps)
    st = sim.download()
    st.update(sim.scalars())
    sim.close()
    assert st["step"] == nsteps
    cg, cr = _conserved(st, pb.cfg), _conserved(ref, pb.cfg)
    errs = {k: _rel_l1(cg[k], cr[k]) for k in cr}
    bn = sum(np.abs(ref[f"b_x{d+1}"]).sum() for d in range(3))
    for d in range(3):
        k = f"b_x{d+1}"
        errs[k] = float(np.abs(st[k] - ref[k]).sum() / bn)
    print("  mixed + surrogate:", "  ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    for k, v in errs.items():
        assert v <= 1e-6, (k, v)
