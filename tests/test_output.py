"""Grid output in the reference's HDF5 layout (write_output,
source/source.F90:2899-3169), read back with the HDF5 tools of the image:
dataset names, shapes and types as the reference's Fortran writer leaves
them (dims (nx3, nx2, nx1[, ncomp]) in C order, component fastest), the
flag attributes, the step-0 statics, and the arrays bit for bit."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from phlegethon_amd import output, problems as P
from phlegethon_amd.config import EOS_HELMHOLTZ

H5BIN = "/opt/conda/bin"

pytestmark = pytest.mark.skipif(
    not (os.path.exists(output.H5_LIB_PATH) and os.path.exists(os.path.join(H5BIN, "h5dump"))),
    reason="libphlx_h5.so or the HDF5 tools are absent")


def _tool(name, *args):
    r = subprocess.run([os.path.join(H5BIN, name), *args], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _read(path, dset, tmp):
    out = str(tmp / (dset.replace("/", "_") + ".bin"))
    _tool("h5dump", "-d", dset, "-b", "LE", "-o", out, path)
    return np.fromfile(out, dtype="<f8")


def _datasets(path):
    txt = _tool("h5ls", "-r", path)
    out = {}
    for line in txt.splitlines():
        m = re.match(r"^/grid/(\S+)\s+Dataset \{([^}]*)\}", line)
        if m:
            dims = m.group(2).strip()
            out[m.group(1)] = () if dims == "SCALAR" else tuple(int(v) for v in dims.split(","))
    return out


def _state(pb, rng):
    cfg = pb.cfg
    return {"prim": pb.prim.copy(),
            "temp": rng.uniform(1e6, 1e9, cfg.shape) if pb.temp is None else pb.temp,
            "b_cc": rng.normal(size=(3,) + cfg.shape)}


def test_mhd_3d_layout(tmp_path):
    rng = np.random.default_rng(3)
    pb = P.orszag_tang_3d(n=6)
    st = _state(pb, rng)
    p0 = output.write_grid(str(tmp_path), pb.cfg, {"step": 0, "time": 0.0, "dt": 1e-3}, st,
                           pb.profiles)
    assert os.path.basename(p0) == "grid_n00000.h5"
    ds = _datasets(p0)
    nz, ny, nx = pb.cfg.shape
    assert ds["prim"] == (nz, ny, nx, pb.cfg.nvars)
    assert ds["bfield"] == (nz, ny, nx, 3)
    assert ds["temp"] == (nz, ny, nx) and ds["vol"] == (nz, ny, nx)
    assert ds["coords"] == (nz, ny, nx, 3)
    for k in ("time", "dt", "step", "gamma_ad", "mu", "x1l", "x3u", "sdims", "nx1", "nx3"):
        assert ds[k] == (), k
    prim = _read(p0, "/grid/prim", tmp_path).reshape(nz, ny, nx, pb.cfg.nvars)
    np.testing.assert_array_equal(prim, np.moveaxis(st["prim"], 0, -1))
    b = _read(p0, "/grid/bfield", tmp_path).reshape(nz, ny, nx, 3)
    np.testing.assert_array_equal(b, np.moveaxis(st["b_cc"], 0, -1))
    xc = _read(p0, "/grid/coords", tmp_path).reshape(nz, ny, nx, 3)
    np.testing.assert_allclose(xc[0, 0, :, 0], (np.arange(nx) + 0.5) / nx, rtol=0, atol=1e-15)
    np.testing.assert_allclose(xc[:, 0, 0, 2], (np.arange(nz) + 0.5) / nz, rtol=0, atol=1e-15)
    vol = _read(p0, "/grid/vol", tmp_path)
    np.testing.assert_allclose(vol, 1.0 / (nx * ny * nz), rtol=1e-15)
    hdr = _tool("h5dump", "-A", p0)
    assert re.search(r'ATTRIBUTE "geometry-type".*?"cartesian"', hdr, re.S)
    assert re.search(r'ATTRIBUTE "use_helmholtz".*?"false"', hdr, re.S)
    # later steps carry the state and the scalars only
    p5 = output.write_grid(str(tmp_path), pb.cfg, {"step": 5, "time": 0.25, "dt": 2e-3}, st,
                           pb.profiles)
    ds5 = _datasets(p5)
    assert set(ds5) == {"time", "dt", "step", "prim", "temp", "bfield"}
    assert "0.25" in _tool("h5dump", "-d", "/grid/time", p5)
    assert re.search(r"\(0\): 5\b", _tool("h5dump", "-d", "/grid/step", p5))


def test_gravity_and_helmholtz_statics(tmp_path):
    rng = np.random.default_rng(4)
    pb = P.hot_bubble(n=8, etot=True)
    st = _state(pb, rng)
    p0 = output.write_grid(str(tmp_path), pb.cfg, {"step": 0, "time": 0.0, "dt": 1e-3}, st,
                           pb.profiles)
    ds = _datasets(p0)
    nz, ny, nx = pb.cfg.shape
    assert ds["grav"] == (nz, ny, nx, 2) and ds["phi_cc"] == (nz, ny, nx)
    assert "bfield" not in ds
    g = _read(p0, "/grid/grav", tmp_path).reshape(nz, ny, nx, 2)
    np.testing.assert_array_equal(g[0, :, 3, 1], pb.profiles["grav_cc"])
    np.testing.assert_array_equal(g[..., 0], 0.0)
    phi = _read(p0, "/grid/phi_cc", tmp_path).reshape(nz, ny, nx)
    np.testing.assert_array_equal(phi[0, :, 5], pb.profiles["phi_cc"])


def test_helmholtz_flags(tmp_path):
    from phlegethon_amd.config import PhlxConfig
    cfg = PhlxConfig(sdims=3, nx1=4, nx2=3, nx3=2, nas=2, eos=EOS_HELMHOLTZ,
                     fasteos=True).validate()
    rng = np.random.default_rng(5)
    st = {"prim": rng.uniform(1, 2, (cfg.nvars,) + cfg.shape),
          "temp": rng.uniform(1e7, 1e8, cfg.shape), "b_cc": None}
    p0 = output.write_grid(str(tmp_path), cfg, {"step": 0, "time": 0.0, "dt": 1.0}, st)
    hdr = _tool("h5dump", "-A", p0)
    for key in ("use_helmholtz", "advect_yeiabar"):
        assert re.search(rf'ATTRIBUTE "{key}".*?"true"', hdr, re.S), key
    assert re.search(r'ATTRIBUTE "use_coulomb_corrections".*?"false"', hdr, re.S)
    t = _read(p0, "/grid/temp", tmp_path).reshape(cfg.shape)
    np.testing.assert_array_equal(t, st["temp"])


def test_pig_and_kappa_flags(tmp_path):
    """PIG runs write use_pig true (advect_yeiabar with two scalars);
    THERMAL_DIFFUSION_STS writes the interior kappa."""
    from phlegethon_amd.config import EOS_PIG, PhlxConfig
    rng = np.random.default_rng(6)
    for nas, yei in ((0, "false"), (2, "true")):
        cfg = PhlxConfig(sdims=2, nx1=5, nx2=4, nas=nas, eos=EOS_PIG, fasteos=True,
                         thermal_diffusion=True).validate()
        st = {"prim": rng.uniform(1, 2, (cfg.nvars,) + cfg.shape),
              "temp": rng.uniform(1e4, 1e5, cfg.shape), "b_cc": None}
        kap = rng.uniform(1.0, 2.0, (1, cfg.nx2 + 2, cfg.nx1 + 2))
        d = tmp_path / f"nas{nas}"
        p0 = output.write_grid(str(d), cfg, {"step": 0, "time": 0.0, "dt": 1.0}, st, kappa=kap)
        hdr = _tool("h5dump", "-A", p0)
        assert re.search(r'ATTRIBUTE "use_pig".*?"true"', hdr, re.S)
        assert re.search(r'ATTRIBUTE "use_helmholtz".*?"false"', hdr, re.S)
        assert re.search(rf'ATTRIBUTE "advect_yeiabar".*?"{yei}"', hdr, re.S)
        k = _read(p0, "/grid/kappa", tmp_path).reshape(cfg.shape)
        np.testing.assert_array_equal(k, kap[:, 1:-1, 1:-1])
    # without thermal diffusion no kappa dataset
    cfg = PhlxConfig(sdims=2, nx1=5, nx2=4).validate()
    st = {"prim": rng.uniform(1, 2, (cfg.nvars,) + cfg.shape), "temp": None, "b_cc": None}
    p0 = output.write_grid(str(tmp_path / "plain"), cfg, {"step": 0, "time": 0.0, "dt": 1.0},
                           st, kappa=np.ones((1, 6, 7)))
    assert "kappa" not in _datasets(p0)
