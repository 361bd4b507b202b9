"""phlx_run replaying one captured HIP graph per step (phlx_set_step_graph)
against the eager launches of the same steps.

Bar: bit for bit.  The graph holds the same kernels with the same
arguments, so every stored field (primitives, eint, T, gammas, face and
cell B) and the scalars (time, dt, step) must be identical after each run.
Each case runs in chunks with downloads and host synchronisation between
them, so the graph is replayed several times with the program's own eager
work in between; a setting changed between chunks (the sweep form) makes
the library capture the step again.  Cases: 3D MHD RK2 (the face-field
buffers alternate within the captured step), 2D hydro in mixed precision,
the fused sweep on a low-Mach WB box, the Helmholtz slab with MHD in mixed
precision, RK3, and steps past tmax.
"""
import numpy as np
import pytest

from phlegethon_amd import problems as P
from phlegethon_amd.config import PREC_MIXED

pytestmark = pytest.mark.gpu


def _chunks(pb, chunks, graph, fused=None):
    from phlegethon_amd.runtime import Simulation
    sim = Simulation(pb.cfg)
    sim.setup(pb)
    sim.set_step_graph(graph)
    sim.init()
    out = []
    for n, f in zip(chunks, fused or [None] * len(chunks)):
        sim.set_fused_sweep(f)
        sim.run(n)
        st = sim.download()
        st.update(sim.scalars())
        out.append(st)
    sim.close()
    return out


def _same(pb, chunks, fused=None):
    print(f"[graph] {pb.name} {pb.cfg.shape} precision={pb.cfg.precision} "
          f"rk={pb.cfg.rk_stages} chunks={chunks}", flush=True)
    eager = _chunks(pb, chunks, False, fused)
    graph = _chunks(pb, chunks, True, fused)
    for a, b in zip(eager, graph):
        for k, v in a.items():
            if isinstance(v, np.ndarray):
                assert np.array_equal(v, b[k]), k
            else:
                assert v == b[k], k
    return graph


def _mixed(pb):
    pb.cfg = pb.cfg.with_(precision=PREC_MIXED)
    return pb


def test_mhd_rk2():
    _same(P.orszag_tang_3d(shape=(16, 16, 16)), [3, 4, 1])


def test_mhd_rk3():
    _same(P.orszag_tang_3d(shape=(16, 16, 8), rk_stages=3), [2, 3])


def test_hydro_2d_mixed():
    _same(_mixed(P.hot_bubble(n=48)), [5, 5])


def test_fused_sweep_and_recapture():
    # the sweep form switches between chunks: fused, separate, fused
    _same(_mixed(P.lowmach_box(n=32, mach=1e-2, heating=1e-6)), [3, 3, 3],
          fused=[True, False, True])


def test_helmholtz_mhd_slab_mixed():
    from phlegethon_amd import helmholtz as H
    import oracle
    import os
    import tempfile
    p = os.path.join(tempfile.mkdtemp(), "t.bin")
    H.generate(141, 81, 0.0, 7.0, 6.0, 6.0 + 80 * 10.0 / 270.0, path=p)
    table = H.load(p)
    orc = oracle.OracleHelm(table)
    _same(_mixed(P.helm_slab(orc, table, n=32, ny=16, sdims=3, mhd_beta=1e5)), [3, 3])


def test_steps_past_tmax():
    pb = P.orszag_tang_3d(shape=(16, 12, 8))
    probe = _chunks(pb, [4], False)[0]
    pb.cfg = pb.cfg.with_(tmax=0.999 * probe["time"])
    out = _same(pb, [3, 3, 2])
    assert out[-1]["step"] == out[-2]["step"] == 4
