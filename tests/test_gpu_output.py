"""Grid output of a GPU run in the reference's HDF5 layout (write_output,
source/source.F90:2899-3169) against the same file written from the CPU
oracle's run of the same problem: every dataset of grid_n00000.h5 (statics
included) and of grid_nNNNNN.h5 after N steps agrees to 1e-12 relative
(h5diff -p), the files list the same datasets, and the FP64 path makes them
equal in practice."""
import os
import subprocess

import numpy as np
import pytest

from phlegethon_amd import output, problems as P

H5BIN = "/opt/conda/bin"

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skipif(not (os.path.exists(output.H5_LIB_PATH) and
                            os.path.exists(os.path.join(H5BIN, "h5diff"))),
                       reason="libphlx_h5.so or the HDF5 tools are absent"),
]


def _h5diff(a, b):
    r = subprocess.run([os.path.join(H5BIN, "h5diff"), "-r", "-p", "1e-12", a, b],
                       capture_output=True, text=True)
    return r.returncode, r.stdout + r.stderr


def _listing(path):
    r = subprocess.run([os.path.join(H5BIN, "h5ls"), "-r", path], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return sorted(line.split()[0] for line in r.stdout.splitlines() if line.startswith("/grid/"))


@pytest.mark.parametrize("case", ["ot3d", "hot_bubble_etot", "spot_diffusion"])
def test_gpu_output_matches_oracle(tmp_path, oracle_mod, case):
    from phlegethon_amd.runtime import Simulation
    kappa = None
    if case == "ot3d":
        pb = P.orszag_tang_3d(n=12)
    elif case == "hot_bubble_etot":
        pb = P.hot_bubble(n=32, etot=True)
    else:
        from phlegethon_amd.config import PhlxConfig
        R = 8.31446261815324e7
        cfg = PhlxConfig(sdims=2, nx1=24, nx2=16, x1u=1e8, x2u=1e8,
                         thermal_diffusion=True).validate()
        X, Y, _ = P.mesh(cfg)
        T = 1e7 * (1.0 + np.exp(-((X - 5e7) ** 2 + (Y - 5e7) ** 2) / (2e7) ** 2))
        prim = np.zeros((cfg.nvars,) + cfg.shape)
        prim[0] = 1.0
        prim[3] = R * T / cfg.mu
        pb = P.Problem("spot", cfg, prim, None, {})
        rng = np.random.default_rng(2)
        kappa = 1e-5 * (1.0 + 0.3 * rng.uniform(0, 1, (1, cfg.nx2 + 2, cfg.nx1 + 2)))
        pb.kappa = kappa
    nsteps = 5
    print(f"[output] {case}: {pb.cfg.shape}, {nsteps} steps", flush=True)
    dg, do = str(tmp_path / "gpu"), str(tmp_path / "oracle")
    sim = Simulation(pb.cfg)
    sim.setup(pb)
    sim.init()
    g0 = sim.write_output(dg, pb.profiles)
    sim.run(nsteps)
    gN = sim.write_output(dg, pb.profiles)
    sim.close()
    r = oracle_mod.OracleRun(pb)
    r.init()
    sc = lambda: {"step": int(r.scalar("step")), "time": r.scalar("time"), "dt": r.scalar("dt")}
    o0 = output.write_grid(do, pb.cfg, sc(), r.state(), pb.profiles, kappa)
    r.run(nsteps)
    oN = output.write_grid(do, pb.cfg, sc(), r.state(), pb.profiles, kappa)
    r.close()
    assert os.path.basename(gN) == os.path.basename(oN) == f"grid_n{nsteps:05d}.h5"
    for a, b in ((g0, o0), (gN, oN)):
        assert _listing(a) == _listing(b)
        rc, msg = _h5diff(a, b)
        assert rc == 0, msg[-2000:]
    if kappa is not None:
        assert "/grid/kappa" in _listing(g0)
