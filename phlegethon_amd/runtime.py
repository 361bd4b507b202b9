"""Host-side mirror of the reference's hydro_step interface over the C-ABI.

``Simulation`` plays the role of the reference's ``locgrid`` + ``time_loop``
caller (source/source.F90:839-1130, :5694-6841): it uploads the initial
state, runs ``init()`` (step-0 EoS and CFL), then ``run(nsteps)`` /
``hydro_step()`` / ``compute_hyperbolic_dt()`` on the GPU, and downloads
the state.  All numerics happen in libphlx.so (phlegethon_amd/csrc); if the
library is missing this module raises instead of computing anything on
the CPU.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional

import numpy as np

from .config import EOS_HELMHOLTZ, EOS_IDEAL, PhlxConfig

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libphlx.so")

_lib = None


class PhlxError(RuntimeError):
    def __init__(self, code: int, what: str):
        msg = _lib.phlx_error_string(code).decode() if _lib is not None else str(code)
        super().__init__(f"{what} failed: {msg} ({code})")
        self.code = code


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libphlx.so (built by ``__graft_entry__.build()``).  Raises if it is
    missing: there is no fallback implementation."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} not found: build the HIP library first (python -c "
            "'import __graft_entry__ as g; g.build()')")
    # torch ships its own HIP runtime under the same soname
    # (libamdhip64.so.7) as the system one the library is linked against:
    # whichever loads first serves the whole process, and torch's device
    # code does not initialise on the system runtime.  Loading torch first
    # makes both use its runtime (the library's calls are plain HIP API).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    P, I, D = ctypes.c_void_p, ctypes.c_int32, ctypes.c_double
    DP = ctypes.POINTER(ctypes.c_double)
    lib.phlx_error_string.restype = ctypes.c_char_p
    lib.phlx_error_string.argtypes = [ctypes.c_int]
    lib.phlx_create.argtypes = [P, ctypes.POINTER(P)]
    lib.phlx_destroy.argtypes = [P]
    lib.phlx_nvars.argtypes = [P]
    lib.phlx_upload_state.argtypes = [P] + [P] * 6
    lib.phlx_download_state.argtypes = [P] + [P] * 7
    lib.phlx_upload_prim_aos.argtypes = [P, P, I]
    lib.phlx_download_prim_aos.argtypes = [P, P, I]
    lib.phlx_set_profiles.argtypes = [P] + [P] * 8
    lib.phlx_set_eos_profiles.argtypes = [P, P, P]
    lib.phlx_download_gammaf.argtypes = [P, P]
    lib.phlx_init.argtypes = [P]
    lib.phlx_hydro_step.argtypes = [P]
    lib.phlx_compute_dt.argtypes = [P, DP, DP]
    lib.phlx_run.argtypes = [P, I]
    lib.phlx_synchronize.argtypes = [P]
    lib.phlx_get_scalars.argtypes = [P, DP, DP, DP, ctypes.POINTER(ctypes.c_int64)]
    lib.phlx_set_scalars.argtypes = [P, D, D, D, ctypes.c_int64]
    lib.phlx_resume.argtypes = [P, D, D, D, ctypes.c_int64]
    lib.phlx_stage_begin.argtypes = [P, I]
    lib.phlx_sweep.argtypes = [P, I]
    lib.phlx_ct.argtypes = [P, I, D, D, D]
    lib.phlx_rk_update.argtypes = [P, I, D, D, D, I]
    lib.phlx_set_fused_sweep.argtypes = [P, I]
    lib.phlx_set_step_graph.argtypes = [P, I]
    lib.phlx_divb_max.argtypes = [P, DP]
    lib.phlx_profile_enable.argtypes = [P, I]
    lib.phlx_profile_read.argtypes = [P, I, DP, ctypes.POINTER(ctypes.c_int64)]
    lib.phlx_stream_copy.argtypes = [P, ctypes.c_int64, I, DP]
    lib.phlx_stream.argtypes = [P]
    lib.phlx_stream.restype = P
    lib.phlx_load_helm_table.argtypes = [P, P, P, I, I, D, D, D, D]
    lib.phlx_load_pig_table.argtypes = [P, P, P, I, I, D, D, D, D]
    lib.phlx_helm_rhoT_full_batch.argtypes = [P, ctypes.c_int64] + [P] * 8
    lib.phlx_helm_rhoe_given_batch.argtypes = [P, ctypes.c_int64] + [P] * 7 + [I, P]
    lib.phlx_helm_rhoP_given_batch.argtypes = [P, ctypes.c_int64] + [P] * 7 + [I, P]
    lib.phlx_helm_rhoe_given_guess_batch.argtypes = ([P, ctypes.c_int64] + [P] * 7 +
                                                     [I, I, P, P])
    lib.phlx_mlp_set_params.argtypes = [P, P, P]
    lib.phlx_mlp_get_params.argtypes = [P, P, P]
    lib.phlx_mlp_forward.argtypes = [P, ctypes.c_int64, P, P, P]
    lib.phlx_mlp_grad.argtypes = [P, ctypes.c_int64, P, P, P, P, DP]
    lib.phlx_mlp_train.argtypes = [P, ctypes.c_int64, I, I, D, ctypes.c_uint64, DP, DP, DP]
    lib.phlx_mlp_set_train_box.argtypes = [P, P]
    lib.phlx_set_eos_guess.argtypes = [P, I, D, I]
    lib.phlx_set_eos_lin_tol.argtypes = [P, D]
    lib.phlx_set_shock_flattening.argtypes = [P, I, D]
    lib.phlx_set_thermal_diffusion.argtypes = [P, I, I, P]
    lib.phlx_set_kappa.argtypes = [P, P]
    lib.phlx_parabolic_dt.argtypes = [P, DP]
    lib.phlx_diffusion_step.argtypes = [P, I]
    lib.phlx_diffusion_stages.argtypes = [P, ctypes.POINTER(ctypes.c_int32)]
    I64P = ctypes.POINTER(ctypes.c_int64)
    lib.phlx_eos_stats.argtypes = [P, I64P, I64P, I64P, I]
    _lib = lib
    return lib


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def _chk(rc: int, what: str):
    if rc != 0:
        raise PhlxError(rc, what)


MLP_NPARAM = 4545
GUESS_MODES = {"warm": 0, "mlp": 1, "bracket": 2, "pred": 3}

KERNELS = {"flux_x1": 0, "flux_x2": 1, "flux_x3": 2, "corner": 3, "faceb": 4,
           "update": 5, "misc": 6, "diffusion": 7, "sweep": 8}


class Simulation:
    """Device-resident state of one problem (one context, one HIP stream)."""

    def __init__(self, cfg: PhlxConfig):
        self.cfg = cfg.validate()
        self.lib = load_library()
        self._ctx = ctypes.c_void_p()
        c = self.cfg.to_c()
        _chk(self.lib.phlx_create(ctypes.byref(c), ctypes.byref(self._ctx)), "phlx_create")
        if self.cfg.shock_flattening:
            self.set_shock_flattening(True, self.cfg.eps_sf)
        if self.cfg.thermal_diffusion:
            self.set_thermal_diffusion(True, self.cfg.eval_parabolic_dt,
                                       self.cfg.fix_temperature)

    # -- lifecycle -------------------------------------------------------
    def close(self):
        if self._ctx:
            self.lib.phlx_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state exchange ----------------------------------------------------
    def upload(self, prim: np.ndarray, b_faces=None, eint: Optional[np.ndarray] = None,
               temp: Optional[np.ndarray] = None):
        cfg = self.cfg
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        assert prim.shape == (cfg.nvars,) + cfg.shape, prim.shape
        bs = [None, None, None]
        if b_faces is not None:
            for d in range(cfg.sdims):
                bs[d] = np.ascontiguousarray(b_faces[d], dtype=np.float64)
                assert bs[d].shape == cfg.face_shape(d), (d, bs[d].shape)
        e = None if eint is None else np.ascontiguousarray(eint, dtype=np.float64)
        t = None if temp is None else np.ascontiguousarray(temp, dtype=np.float64)
        _chk(self.lib.phlx_upload_state(self._ctx, _ptr(prim), _ptr(e), _ptr(t),
                                        _ptr(bs[0]), _ptr(bs[1]), _ptr(bs[2])),
             "phlx_upload_state")

    def set_profiles(self, **prof):
        names = ["eq_rho_cc", "eq_p_cc", "eq_rho_f", "eq_p_f", "grav_cc", "phi_cc",
                 "phi_f", "edot_cc"]
        arrs = [None if prof.get(n) is None else
                np.ascontiguousarray(prof[n], dtype=np.float64) for n in names]
        _chk(self.lib.phlx_set_profiles(self._ctx, *[_ptr(a) for a in arrs]),
             "phlx_set_profiles")
        ge = [None if prof.get(n) is None else np.ascontiguousarray(prof[n], dtype=np.float64)
              for n in ("eq_gammae_cc", "eq_gammae_f")]
        if ge[0] is not None or ge[1] is not None:
            _chk(self.lib.phlx_set_eos_profiles(self._ctx, _ptr(ge[0]), _ptr(ge[1])),
                 "phlx_set_eos_profiles")

    def download(self) -> Dict[str, np.ndarray]:
        cfg = self.cfg
        prim = np.empty((cfg.nvars,) + cfg.shape)
        eint = np.empty(cfg.shape)
        temp = np.empty(cfg.shape)
        bcc = np.empty((3,) + cfg.shape)
        bs = [np.empty(cfg.face_shape(d)) if cfg.mhd and d < cfg.sdims else None
              for d in range(3)]
        _chk(self.lib.phlx_download_state(self._ctx, _ptr(prim), _ptr(eint), _ptr(temp),
                                          _ptr(bs[0]), _ptr(bs[1]), _ptr(bs[2]), _ptr(bcc)),
             "phlx_download_state")
        out = {"prim": prim, "eint": eint, "temp": temp, "b_cc": bcc}
        for d in range(3):
            if bs[d] is not None:
                out[f"b_x{d+1}"] = bs[d]
        if cfg.eos != EOS_IDEAL:
            gf = np.empty((2,) + cfg.shape)
            _chk(self.lib.phlx_download_gammaf(self._ctx, _ptr(gf)), "phlx_download_gammaf")
            out["gammaf"] = gf
        return out

    def setup(self, problem):
        """Upload a ``problems.Problem``: state, profiles and, for Helmholtz
        runs, the table and the initial temperature."""
        if problem.helm is not None:
            self.load_helm_table(problem.helm)
        self.upload(problem.prim, problem.b_faces, temp=problem.temp)
        if problem.profiles:
            self.set_profiles(**problem.profiles)
        if getattr(problem, "kappa", None) is not None:
            self.set_kappa(problem.kappa)
        return self

    # -- the reference's entry points ------------------------------------
    def init(self):
        _chk(self.lib.phlx_init(self._ctx), "phlx_init")

    def resume(self, time: float, dt: float, smax: float, step: int):
        """Continue from a checkpoint (uploaded prim, T, faces) with its
        scalars: the reference's read_restart + step /= 0 initialisation."""
        _chk(self.lib.phlx_resume(self._ctx, float(time), float(dt), float(smax), int(step)),
             "phlx_resume")

    def checkpoint(self) -> Dict[str, np.ndarray]:
        """What the reference's write_restart stores (source/source.F90
        :2507-2616): prim, T, the face fields and time, dt, smax, step."""
        st = self.download()
        out = {"prim": st["prim"], "temp": st["temp"]}
        for d in range(3):
            if f"b_x{d+1}" in st:
                out[f"b_x{d+1}"] = st[f"b_x{d+1}"]
        out.update(self.scalars())
        return out

    def write_output(self, directory: str, profiles=None) -> str:
        """write_output of the reference: <directory>/grid_nNNNNN.h5 in its
        HDF5 layout (phlegethon_amd.output); ``profiles``: the problem's x2
        profiles for the step-0 statics (grav, phi_cc, edot)."""
        from . import output
        return output.write_grid(directory, self.cfg, self.scalars(), self.download(), profiles,
                                 getattr(self, "_kappa_host", None))

    def hydro_step(self):
        _chk(self.lib.phlx_hydro_step(self._ctx), "phlx_hydro_step")

    def compute_hyperbolic_dt(self):
        dt, smax = ctypes.c_double(), ctypes.c_double()
        _chk(self.lib.phlx_compute_dt(self._ctx, ctypes.byref(dt), ctypes.byref(smax)),
             "phlx_compute_dt")
        return dt.value, smax.value

    def run(self, nsteps: int):
        _chk(self.lib.phlx_run(self._ctx, int(nsteps)), "phlx_run")

    def synchronize(self):
        _chk(self.lib.phlx_synchronize(self._ctx), "phlx_synchronize")

    def scalars(self):
        t, dt, smax = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        st = ctypes.c_int64()
        _chk(self.lib.phlx_get_scalars(self._ctx, ctypes.byref(t), ctypes.byref(dt),
                                       ctypes.byref(smax), ctypes.byref(st)),
             "phlx_get_scalars")
        return {"time": t.value, "dt": dt.value, "smax": smax.value, "step": st.value}

    def set_scalars(self, time: float, dt: float, smax: float, step: int):
        _chk(self.lib.phlx_set_scalars(self._ctx, time, dt, smax, step), "phlx_set_scalars")

    def divb_max(self) -> float:
        v = ctypes.c_double()
        _chk(self.lib.phlx_divb_max(self._ctx, ctypes.byref(v)), "phlx_divb_max")
        return v.value

    # -- fine-grained stage API ------------------------------------------
    def stage_begin(self, irk):
        _chk(self.lib.phlx_stage_begin(self._ctx, irk), "phlx_stage_begin")

    def sweep(self, d):
        _chk(self.lib.phlx_sweep(self._ctx, d), "phlx_sweep")

    def ct(self, irk, a1, a2, a3):
        _chk(self.lib.phlx_ct(self._ctx, irk, a1, a2, a3), "phlx_ct")

    def rk_update(self, irk, a1, a2, a3, last):
        _chk(self.lib.phlx_rk_update(self._ctx, irk, a1, a2, a3, int(last)), "phlx_rk_update")

    def set_fused_sweep(self, on):
        """hydro_step with the fused one-pass sweep (True), the three
        directional sweeps (False) or the library's choice (None: fused
        without MHD); the results are the same bit for bit."""
        mode = -1 if on is None else int(bool(on))
        _chk(self.lib.phlx_set_fused_sweep(self._ctx, mode), "phlx_set_fused_sweep")

    def set_step_graph(self, on: bool):
        """run() replays one captured HIP graph per step (True) or launches
        each kernel (False, the default); the results are the same bit for
        bit."""
        _chk(self.lib.phlx_set_step_graph(self._ctx, int(bool(on))), "phlx_set_step_graph")

    def stream_copy(self, nbytes: int = 1 << 30, reps: int = 10) -> float:
        """Achieved HBM GB/s of a plain device copy (read + write)."""
        v = ctypes.c_double()
        _chk(self.lib.phlx_stream_copy(self._ctx, int(nbytes), int(reps), ctypes.byref(v)),
             "phlx_stream_copy")
        return v.value

    # -- Helmholtz EoS ---------------------------------------------------
    def load_helm_table(self, table):
        """Upload a Helmholtz table (``phlegethon_amd.helmholtz.HelmTable``:
        file column order, [T][rho][column])."""
        f9 = np.ascontiguousarray(table.f9, dtype=np.float64)
        c4 = np.ascontiguousarray(table.c4, dtype=np.float64)
        fn = "phlx_load_pig_table" if getattr(table, "kind", "helm") == "pig" else \
            "phlx_load_helm_table"
        _chk(getattr(self.lib, fn)(self._ctx, _ptr(f9), _ptr(c4), table.nrho, table.nT,
                                   table.ldlo, table.ldhi, table.ltlo, table.lthi), fn)

    @property
    def ctx(self):
        return self._ctx

    # -- T(rho, e) surrogate ---------------------------------------------
    def mlp_train(self, nsamples: int = 1 << 22, epochs: int = 4, batch: int = 4096,
                  lr: float = 3e-3, seed: int = 20261202, box=None):
        """Generate (rho, e, abar, zbar) -> T pairs with the loaded table and
        train the surrogate on the GPU; returns (final batch loss, held-out
        RMS error of log10 T, held-out RMS of the implied log10 E error).
        box: (log10 rhoYe lo, hi, log10 T lo, hi, abar lo, hi, Ye lo, hi) to
        sample from (e.g. problems.state_box); None = the table box."""
        if box is None:
            _chk(self.lib.phlx_mlp_set_train_box(self._ctx, None), "phlx_mlp_set_train_box")
        else:
            b = np.ascontiguousarray(box, dtype=np.float64)
            assert b.size == 8
            _chk(self.lib.phlx_mlp_set_train_box(self._ctx, b.ctypes.data_as(ctypes.c_void_p)),
                 "phlx_mlp_set_train_box")
        loss, rmst, rmse = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _chk(self.lib.phlx_mlp_train(self._ctx, int(nsamples), int(epochs), int(batch),
                                     float(lr), int(seed), ctypes.byref(loss),
                                     ctypes.byref(rmst), ctypes.byref(rmse)), "phlx_mlp_train")
        return loss.value, rmst.value, rmse.value

    def mlp_params(self):
        theta = np.empty(MLP_NPARAM, dtype=np.float32)
        norm = np.empty(10, dtype=np.float32)
        _chk(self.lib.phlx_mlp_get_params(self._ctx, theta.ctypes.data_as(ctypes.c_void_p),
                                          norm.ctypes.data_as(ctypes.c_void_p)),
             "phlx_mlp_get_params")
        return theta, norm

    def mlp_set_params(self, theta, norm):
        theta = np.ascontiguousarray(theta, dtype=np.float32)
        norm = np.ascontiguousarray(norm, dtype=np.float32)
        assert theta.size == MLP_NPARAM and norm.size == 10
        _chk(self.lib.phlx_mlp_set_params(self._ctx, theta.ctypes.data_as(ctypes.c_void_p),
                                          norm.ctypes.data_as(ctypes.c_void_p)),
             "phlx_mlp_set_params")

    def set_eos_guess(self, mode: str = "warm", resid_tol: float = 0.25, count: bool = False):
        """T(rho, e) start of the per-zone Newton: "warm" (the carried T, as
        the reference), "mlp" (surrogate, bracketed fallback), "bracket" or
        "pred" (the carried T moved by (e - e_prev) / c_v of the zone's
        previous solve)."""
        _chk(self.lib.phlx_set_eos_guess(self._ctx, GUESS_MODES[mode], float(resid_tol),
                                         int(count)), "phlx_set_eos_guess")

    def set_eos_lin_tol(self, lin_tol: float = 0.0):
        """Mixed precision + Helmholtz: 0 (default) converges the per-zone
        T(rho, e) Newton to the reference's tolerance; a positive value is
        the energy residual below which the solve instead takes its Newton
        step linearly after one FP64 evaluation (a shortcut)."""
        _chk(self.lib.phlx_set_eos_lin_tol(self._ctx, float(lin_tol)), "phlx_set_eos_lin_tol")

    def set_shock_flattening(self, enable: bool = True, eps_sf: float = 0.25):
        """USE_SHOCK_FLATTENING with eps_sf_make: flagged cells take LIM2ND,
        faces beside them the HLL flux (FP64 contexts)."""
        _chk(self.lib.phlx_set_shock_flattening(self._ctx, int(bool(enable)), float(eps_sf)),
             "phlx_set_shock_flattening")

    def set_thermal_diffusion(self, enable: bool = True, evaluate_dt: bool = True,
                              fix_temperature=None):
        """THERMAL_DIFFUSION_STS (RKL2 around every hydro step);
        fix_temperature: six wall temperatures or None entries."""
        ft = None
        if fix_temperature is not None:
            ft = np.array([np.nan if v is None else float(v) for v in fix_temperature],
                          dtype=np.float64)
            assert ft.size == 6
        _chk(self.lib.phlx_set_thermal_diffusion(self._ctx, int(bool(enable)),
                                                 int(bool(evaluate_dt)), _ptr(ft)),
             "phlx_set_thermal_diffusion")
        self._keep_ft = ft

    def set_kappa(self, kappa: np.ndarray):
        """Opacity on cells -1..n per axis: shape (nz+2 or 1, ny+2, nx+2)."""
        c = self.cfg
        shp = ((c.nx3 + 2) if c.sdims == 3 else 1, c.nx2 + 2, c.nx1 + 2)
        k = np.ascontiguousarray(kappa, dtype=np.float64)
        if k.shape != shp:
            raise ValueError(f"kappa shape {k.shape}, expected {shp}")
        _chk(self.lib.phlx_set_kappa(self._ctx, _ptr(k)), "phlx_set_kappa")
        self._kappa_host = k   # for write_output's kappa dataset

    def parabolic_dt(self) -> float:
        v = ctypes.c_double()
        _chk(self.lib.phlx_parabolic_dt(self._ctx, ctypes.byref(v)), "phlx_parabolic_dt")
        return v.value

    def diffusion_step(self, strang: int = 1):
        _chk(self.lib.phlx_diffusion_step(self._ctx, int(strang)), "phlx_diffusion_step")

    def diffusion_stages(self) -> int:
        v = ctypes.c_int32()
        _chk(self.lib.phlx_diffusion_stages(self._ctx, ctypes.byref(v)), "phlx_diffusion_stages")
        return v.value

    def eos_stats(self, reset: bool = False):
        it, n, fb = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _chk(self.lib.phlx_eos_stats(self._ctx, ctypes.byref(it), ctypes.byref(n),
                                     ctypes.byref(fb), int(reset)), "phlx_eos_stats")
        return {"newton_iters": it.value, "solves": n.value, "fallbacks": fb.value}

    # -- profiling -------------------------------------------------------
    def profile(self, on: bool):
        _chk(self.lib.phlx_profile_enable(self._ctx, int(on)), "phlx_profile_enable")

    def profile_read(self, name: str):
        ms, n = ctypes.c_double(), ctypes.c_int64()
        _chk(self.lib.phlx_profile_read(self._ctx, KERNELS[name], ctypes.byref(ms),
                                        ctypes.byref(n)), "phlx_profile_read")
        return ms.value, n.value


class HelmEoS:
    """Batched Helmholtz EoS on the GPU (helm_rhoT_given_full,
    helm_rhoe_given, helm_rhoP_given of source/source.F90:23686-25019).

    Inputs are host arrays (broadcast to a common length); they are staged
    in device memory through torch, the library computes on its own stream,
    and results come back as numpy arrays.  The batch entry points accept
    any context, so a minimal one is created here."""

    def __init__(self, table, device: int = 0):
        import torch
        self.torch = torch
        self.device = torch.device("cuda", device)
        cfg = PhlxConfig(sdims=2, nx1=2, nx2=2, device=device)
        self.sim = Simulation(cfg)
        self.sim.load_helm_table(table)
        self.lib = self.sim.lib

    def close(self):
        self.sim.close()

    def _dev(self, *xs):
        n = max(int(np.size(x)) for x in xs)
        out = []
        for x in xs:
            a = np.ascontiguousarray(np.broadcast_to(np.asarray(x, dtype=np.float64), (n,)))
            out.append(self.torch.from_numpy(a.copy()).to(self.device))
        return out, n

    def _empty(self, n, dtype=None):
        return self.torch.empty(n, dtype=dtype or self.torch.float64, device=self.device)

    @staticmethod
    def _p(t):
        return ctypes.c_void_p(t.data_ptr())

    def _sync_in(self):
        self.torch.cuda.synchronize(self.device)

    def _sync_out(self):
        _chk(self.lib.phlx_synchronize(self.sim.ctx), "phlx_synchronize")

    def full(self, rho, T, abar, zbar):
        (rho, T, abar, zbar), n = self._dev(rho, T, abar, zbar)
        P, E, cs, cv = (self._empty(n) for _ in range(4))
        self._sync_in()
        _chk(self.lib.phlx_helm_rhoT_full_batch(self.sim.ctx, n, *map(self._p, (
            rho, T, abar, zbar, P, E, cs, cv))), "phlx_helm_rhoT_full_batch")
        self._sync_out()
        return tuple(x.cpu().numpy() for x in (P, E, cs, cv))

    def rhoe_given(self, rho, e, abar, zbar, T0, ret=True):
        (rho, e, abar, zbar, T), n = self._dev(rho, e, abar, zbar, T0)
        P, cs = self._empty(n), self._empty(n)
        it = self._empty(n, self.torch.int32)
        self._sync_in()
        _chk(self.lib.phlx_helm_rhoe_given_batch(
            self.sim.ctx, n, *map(self._p, (rho, e, abar, zbar, T, P, cs)), int(ret),
            self._p(it)), "phlx_helm_rhoe_given_batch")
        self._sync_out()
        return tuple(x.cpu().numpy() for x in (T, P, cs, it))

    def rhoP_given(self, rho, p, abar, zbar, T0, ret=True):
        (rho, p, abar, zbar, T), n = self._dev(rho, p, abar, zbar, T0)
        E, cs = self._empty(n), self._empty(n)
        it = self._empty(n, self.torch.int32)
        self._sync_in()
        _chk(self.lib.phlx_helm_rhoP_given_batch(
            self.sim.ctx, n, *map(self._p, (rho, p, abar, zbar, T, E, cs)), int(ret),
            self._p(it)), "phlx_helm_rhoP_given_batch")
        self._sync_out()
        return tuple(x.cpu().numpy() for x in (T, E, cs, it))

    # -- surrogate checks (device staging through torch) -----------------
    def rhoe_given_guess(self, rho, e, abar, zbar, T0, mode: str = "mlp", ret=True):
        (rho, e, abar, zbar, T), n = self._dev(rho, e, abar, zbar, T0)
        P, cs = self._empty(n), self._empty(n)
        it = self._empty(n, self.torch.int32)
        fb = self._empty(n, self.torch.int32)
        self._sync_in()
        _chk(self.lib.phlx_helm_rhoe_given_guess_batch(
            self.sim.ctx, n, *map(self._p, (rho, e, abar, zbar, T, P, cs)), int(ret),
            GUESS_MODES[mode], self._p(it), self._p(fb)), "phlx_helm_rhoe_given_guess_batch")
        self._sync_out()
        return tuple(x.cpu().numpy() for x in (T, P, cs, it, fb))

    def mlp_forward(self, raw):
        raw = np.ascontiguousarray(raw, dtype=np.float32)
        n = raw.shape[0]
        r = self.torch.from_numpy(raw.copy()).to(self.device)
        y = self._empty(n, self.torch.float32)
        inbox = self._empty(n, self.torch.int32)
        self._sync_in()
        _chk(self.lib.phlx_mlp_forward(self.sim.ctx, n, self._p(r), self._p(y), self._p(inbox)),
             "phlx_mlp_forward")
        self._sync_out()
        return y.cpu().numpy(), inbox.cpu().numpy()

    def mlp_grad(self, xn, tn, wt=None):
        xn = np.ascontiguousarray(xn, dtype=np.float32)
        tn = np.ascontiguousarray(tn, dtype=np.float32)
        n = xn.shape[0]
        x = self.torch.from_numpy(xn.copy()).to(self.device)
        t = self.torch.from_numpy(tn.copy()).to(self.device)
        w = None if wt is None else self.torch.from_numpy(
            np.ascontiguousarray(wt, dtype=np.float32).copy()).to(self.device)
        g = self._empty(MLP_NPARAM, self.torch.float32)
        loss = ctypes.c_double()
        self._sync_in()
        _chk(self.lib.phlx_mlp_grad(self.sim.ctx, n, self._p(x), self._p(t),
                                    None if w is None else self._p(w), self._p(g),
                                    ctypes.byref(loss)), "phlx_mlp_grad")
        self._sync_out()
        return g.cpu().numpy(), loss.value
