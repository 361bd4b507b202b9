"""phlx_compute_dt_noalloc (no allocation: scratch in a file-scope device
array) against phlx_compute_dt on the same state, bit for bit: two calls in
a row, the second on a smaller grid."""
import ctypes

import numpy as np
import pytest

from phlegethon_amd import problems as P

pytestmark = pytest.mark.gpu


def _reference_dt(pb, nsteps):
    from phlegethon_amd.runtime import Simulation
    sim = Simulation(pb.cfg)
    sim.setup(pb)
    sim.init()
    sim.run(nsteps)
    dt, smax = sim.compute_hyperbolic_dt()
    st = sim.download()
    lib = sim.lib
    sim.close()
    return lib, dt, smax, st


def test_compute_dt_noalloc_matches():
    import torch
    cases = [(P.orszag_tang_3d(n=20), 3), (P.orszag_tang_2d(n=12), 2)]
    for pb, nsteps in cases:
        lib, dt, smax, st = _reference_dt(pb, nsteps)
        cfg = pb.cfg
        prim = torch.from_numpy(np.ascontiguousarray(st["prim"])).to("cuda")
        bcc = torch.from_numpy(np.ascontiguousarray(st["b_cc"][:cfg.sdims])).to("cuda")
        out = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        lib.phlx_compute_dt_noalloc.argtypes = [ctypes.c_void_p] * 5
        lib.phlx_compute_dt_noalloc.restype = ctypes.c_int
        c = cfg.to_c()
        rc = lib.phlx_compute_dt_noalloc(ctypes.addressof(c), prim.data_ptr(), bcc.data_ptr(),
                                         out.data_ptr(), None)
        assert rc == 0
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        print(f"[noalloc] {pb.name} {cfg.shape}: dt {got[0]!r} vs {dt!r}, smax {got[1]!r} vs "
              f"{smax!r}", flush=True)
        assert got[0] == dt and got[1] == smax


def test_compute_dt_noalloc_rejects():
    from phlegethon_amd.runtime import load_library
    lib = load_library()
    lib.phlx_compute_dt_noalloc.argtypes = [ctypes.c_void_p] * 5
    lib.phlx_compute_dt_noalloc.restype = ctypes.c_int
    cfg = P.orszag_tang_3d(n=8).cfg
    c = cfg.to_c()
    assert lib.phlx_compute_dt_noalloc(ctypes.addressof(c), None, None, None, None) == -3
    huge = cfg.with_(nx1=4096, nx2=4096, nx3=4096).to_c()
    # a size check, made before any pointer is touched
    assert lib.phlx_compute_dt_noalloc(ctypes.addressof(huge), 1, 1, 1, None) == -3
