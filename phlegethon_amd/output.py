"""Grid output in the reference's HDF5 layout (write_output,
source/source.F90:2899-3169) through libphlx_h5.so (include/phlx_h5.h).

``write_grid(directory, cfg, scalars, state, profiles)`` writes
``<directory>/grid_nNNNNN.h5`` from a downloaded state (``Simulation.download``)
and the run's scalars (``Simulation.scalars``); the static fields written at
step 0 (grav, phi_cc, edot) come from the problem's x2 profiles."""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional

import numpy as np

from .config import PhlxConfig

_HERE = os.path.dirname(os.path.abspath(__file__))
H5_LIB_PATH = os.path.join(_HERE, "_lib", "libphlx_h5.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(H5_LIB_PATH):
            raise RuntimeError(f"{H5_LIB_PATH} is not built (build.build_h5 needs HDF5)")
        lib = ctypes.CDLL(H5_LIB_PATH)
        P, D = ctypes.c_void_p, ctypes.c_double
        lib.phlx_h5_write_grid.argtypes = [ctypes.c_char_p, P, ctypes.c_int64, D, D] + [P] * 7
        lib.phlx_h5_write_grid.restype = ctypes.c_int
        _lib = lib
    return _lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def grid_path(directory: str, step: int) -> str:
    return os.path.join(directory, f"grid_n{int(step):05d}.h5")


def interior(cfg: PhlxConfig, kappa: np.ndarray) -> np.ndarray:
    """The interior cells of an array with one ghost layer per axis (the
    opacity as set with Simulation.set_kappa)."""
    if cfg.sdims == 3:
        return kappa[1:-1, 1:-1, 1:-1]
    return kappa[:, 1:-1, 1:-1]


def write_grid(directory: str, cfg: PhlxConfig, scalars: Dict, state: Dict,
               profiles: Optional[Dict] = None, kappa: Optional[np.ndarray] = None) -> str:
    """``kappa``: the opacity with its ghost layer (thermal-diffusion runs),
    written as the reference's kappa dataset."""
    lib = _load()
    os.makedirs(directory, exist_ok=True)
    c = cfg.to_c()
    prim = np.ascontiguousarray(state["prim"], dtype=np.float64)
    temp = None if state.get("temp") is None else np.ascontiguousarray(state["temp"],
                                                                       dtype=np.float64)
    bcc = None if state.get("b_cc") is None else np.ascontiguousarray(state["b_cc"],
                                                                      dtype=np.float64)
    prof = profiles or {}
    arr = {k: (None if prof.get(k) is None else np.ascontiguousarray(prof[k], dtype=np.float64))
           for k in ("grav_cc", "phi_cc", "edot_cc")}
    kap = None
    if kappa is not None and cfg.thermal_diffusion:
        kap = np.ascontiguousarray(interior(cfg, np.asarray(kappa)), dtype=np.float64)
        assert kap.shape == cfg.shape, kap.shape
    rc = lib.phlx_h5_write_grid(directory.encode(), ctypes.byref(c), int(scalars["step"]),
                                float(scalars["time"]), float(scalars["dt"]), _ptr(prim),
                                _ptr(temp), _ptr(bcc), _ptr(arr["grav_cc"]),
                                _ptr(arr["phi_cc"]), _ptr(arr["edot_cc"]), _ptr(kap))
    if rc != 0:
        raise RuntimeError(f"phlx_h5_write_grid failed ({rc})")
    return grid_path(directory, scalars["step"])
