"""Checkpoint and resume, the reference's write_restart / read_restart path
(source/source.F90:2507-2616, :2761-2897) and time_loop's step /= 0
initialisation (:6413-6500): the restart file holds prim, T, the face fields
and time, dt, smax, step; on resume eint (and the FASTEOS gammas) is rebuilt
from (rho, T) or p, and dt is taken from the file.

- CPU: the oracle's resumed run equals the uninterrupted run up to the
  rounding of the rebuilt eint (ideal gas, 1e-13; Helmholtz, where eint is
  re-evaluated at the Newton-converged T, 1e-9), with identical scalars;
- GPU: checkpoint from the GPU run, resume on a fresh context; equal to the
  oracle resumed from the same checkpoint to the FP64 bar (1e-12 rel. L1),
  step counter and time continued.
"""
import numpy as np
import pytest

from phlegethon_amd import helmholtz as H
from phlegethon_amd import problems as P

N1, N2 = 3, 4  # steps before / after the checkpoint


@pytest.fixture(scope="module")
def table(tmp_path_factory):
    p = str(tmp_path_factory.mktemp("restart") / "t.bin")
    H.generate(141, 81, 0.0, 7.0, 6.0, 6.0 + 80 * 10.0 / 270.0, path=p)
    return H.load(p)


def _problems(oracle_mod, table):
    orc = oracle_mod.OracleHelm(table)
    return {
        "hot_bubble_2d": P.hot_bubble(n=24),
        "orszag_tang_3d": P.orszag_tang_3d(n=12),
        "helm_slab_fasteos": P.helm_slab(orc, table, n=8, sdims=3),
        "helm_slab_face_newton": P.helm_slab(orc, table, n=8, sdims=3, fasteos=False),
    }


def _from_checkpoint(pb, ck):
    q = pb.copy()
    q.prim = ck["prim"].copy()
    q.temp = ck["temp"].copy()
    if pb.cfg.mhd:
        q.b_faces = [ck.get(f"b_x{d+1}") for d in range(3)]
    return q


def _rel(a, b):
    den = np.abs(b).sum()
    return float(np.abs(a - b).sum() / (den if den > 0 else 1.0))


def _oracle_resume(oracle_mod, pb, ck, nsteps):
    r = oracle_mod.OracleRun(_from_checkpoint(pb, ck), 1)
    r.resume(ck["time"], ck["dt"], ck["smax"], ck["step"])
    r.run(nsteps)
    st = r.state()
    st.update(time=r.scalar("time"), step=int(r.scalar("step")))
    r.close()
    return st


def _oracle_checkpoint(oracle_mod, pb, nsteps):
    r = oracle_mod.OracleRun(pb, 1)
    r.init()
    r.run(nsteps)
    st = r.state()
    ck = {"prim": st["prim"], "temp": st["temp"], "time": r.scalar("time"),
          "dt": r.scalar("dt"), "smax": r.scalar("smax"), "step": int(r.scalar("step"))}
    for d in range(3):
        if f"b_x{d+1}" in st:
            ck[f"b_x{d+1}"] = st[f"b_x{d+1}"]
    r.close()
    return ck


@pytest.mark.parametrize("name", ["hot_bubble_2d", "orszag_tang_3d", "helm_slab_fasteos",
                                  "helm_slab_face_newton"])
def test_oracle_resume_continues_the_run(oracle_mod, table, name):
    pb = _problems(oracle_mod, table)[name]
    straight = oracle_mod.run_problem(pb, N1 + N2)
    ck = _oracle_checkpoint(oracle_mod, pb, N1)
    assert ck["step"] == N1
    res = _oracle_resume(oracle_mod, pb, ck, N2)
    assert res["step"] == N1 + N2
    tol = 1e-9 if pb.cfg.eos else 1e-13
    assert abs(res["time"] / straight["time"] - 1.0) < tol
    # velocities against the sound-speed scale sqrt(p / rho): in the
    # near-static problems they are noise-sized and carry the rounding of
    # the pressure gradients
    ip = pb.cfg.sdims + 1
    cs = np.sqrt(straight["prim"][ip] / straight["prim"][0])
    for v in range(pb.cfg.nvars):
        if 1 <= v <= pb.cfg.sdims:
            err = np.abs(res["prim"][v] - straight["prim"][v]).sum() / cs.sum()
        else:
            err = _rel(res["prim"][v], straight["prim"][v])
        assert err < tol, (v, err)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["hot_bubble_2d", "orszag_tang_3d", "helm_slab_fasteos",
                                  "helm_slab_face_newton"])
def test_gpu_resume_matches_oracle(oracle_mod, table, name):
    from phlegethon_amd.runtime import Simulation
    pb = _problems(oracle_mod, table)[name]
    print(f"resume {name}", flush=True)
    sim = Simulation(pb.cfg)
    sim.setup(pb)
    sim.init()
    sim.run(N1)
    ck = sim.checkpoint()
    sim.close()
    assert ck["step"] == N1
    q = _from_checkpoint(pb, ck)
    sim = Simulation(q.cfg)
    sim.setup(q)
    sim.resume(ck["time"], ck["dt"], ck["smax"], ck["step"])
    sim.run(N2)
    got = sim.download()
    got.update(sim.scalars())
    sim.close()
    ref = _oracle_resume(oracle_mod, pb, ck, N2)
    assert got["step"] == ref["step"] == N1 + N2
    assert abs(got["time"] - ref["time"]) <= 1e-14 * abs(ref["time"])
    for v in range(pb.cfg.nvars):
        assert _rel(got["prim"][v], ref["prim"][v]) <= 1e-12, v
    assert _rel(got["eint"], ref["eint"]) <= 1e-12
    if pb.cfg.mhd:
        for d in range(pb.cfg.sdims):
            assert _rel(got[f"b_x{d+1}"], ref[f"b_x{d+1}"]) <= 1e-12


@pytest.mark.gpu
def test_gpu_resume_arguments():
    from phlegethon_amd.runtime import PhlxError, Simulation
    pb = P.hot_bubble(n=16)
    sim = Simulation(pb.cfg)
    sim.setup(pb)
    for bad in ((0.1, 0.0, 1.0, 3), (0.1, 1e-3, -1.0, 3), (0.1, 1e-3, 1.0, -1)):
        with pytest.raises(PhlxError):
            sim.resume(*bad)
    sim.close()
