"""Run-time configuration mirroring the reference's compile-time OPTS.

The reference selects every option with Makefile ``OPTS`` that become cpp
defines (Makefile.config:16-23) and are validated by Make.checks:27-305;
defaults live in source/source.F90:31-477.  Here the same names are fields
of a dataclass that is validated in Python (the subset this library
supports) and converted to the C struct ``phlx_config`` of include/phlx.h.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field, replace
from typing import Optional, Sequence

LIM2ND_REC, PPH_REC, LIM5TH_REC = 0, 1, 2
HLLC_FLUX, LHLLC_FLUX, HLLD_FLUX, LHLLD_FLUX = 0, 1, 2, 3
BC_PERIODIC, BC_REFLECTIVE, BC_OUTFLOW, BC_DIODE = 0, 1, 2, 3
EOS_IDEAL, EOS_HELMHOLTZ, EOS_PIG = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1

_RECON = {"LIM2ND_REC": LIM2ND_REC, "PPH_REC": PPH_REC, "LIM5TH_REC": LIM5TH_REC}
_FLUX = {"HLLC_FLUX": HLLC_FLUX, "LHLLC_FLUX": LHLLC_FLUX,
         "HLLD_FLUX": HLLD_FLUX, "LHLLD_FLUX": LHLLD_FLUX}
_BC = {"PERIODIC": BC_PERIODIC, "REFLECTIVE": BC_REFLECTIVE,
       "OUTFLOW": BC_OUTFLOW, "DIODE": BC_DIODE}


class ConfigError(ValueError):
    """Raised for option sets the reference's Make.checks (or this library) rejects."""


class PhlxConfigC(ctypes.Structure):
    """ctypes image of ``phlx_config`` (include/phlx.h)."""
    _fields_ = [
        ("sdims", ctypes.c_int32), ("nx1", ctypes.c_int32), ("nx2", ctypes.c_int32),
        ("nx3", ctypes.c_int32), ("nas", ctypes.c_int32),
        ("x1l", ctypes.c_double), ("x1u", ctypes.c_double), ("x2l", ctypes.c_double),
        ("x2u", ctypes.c_double), ("x3l", ctypes.c_double), ("x3u", ctypes.c_double),
        ("gamma", ctypes.c_double), ("mu", ctypes.c_double),
        ("recon", ctypes.c_int32), ("flux", ctypes.c_int32), ("all_mach", ctypes.c_int32),
        ("rk_stages", ctypes.c_int32),
        ("cfl", ctypes.c_double), ("mach_cutoff", ctypes.c_double),
        ("mach_ct", ctypes.c_double), ("tmax", ctypes.c_double),
        ("bc", ctypes.c_int32 * 6),
        ("use_wb", ctypes.c_int32), ("use_gravity", ctypes.c_int32),
        ("evolve_etot", ctypes.c_int32), ("use_edot", ctypes.c_int32),
        ("eos", ctypes.c_int32), ("fasteos", ctypes.c_int32),
        ("precision", ctypes.c_int32), ("device", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 8),
    ]


@dataclass(frozen=True)
class PhlxConfig:
    """One problem's option set.  Field names follow the reference OPTS:
    ``sdims_make``, ``nx*_make``, ``nas_make``, ``cfl_make``,
    ``Mach_cutoff_make`` (default 1e-5, source.F90:170-175), ``Mach_ct_make``
    (default 1e-6, :184-189), ``LIM2ND_REC``/``PPH_REC``,
    ``HLLC``/``LHLLC``/``HLLD``/``LHLLD_FLUX``, ``RK2``/``RK3_STEPPER``,
    ``X1L_*`` ... boundary flags, ``USE_WB``, ``USE_GRAVITY``,
    ``EVOLVE_ETOT``, ``USE_EDOT``."""

    sdims: int = 2
    nx1: int = 64
    nx2: int = 64
    nx3: int = 1
    nas: int = 0
    x1l: float = 0.0
    x1u: float = 1.0
    x2l: float = 0.0
    x2u: float = 1.0
    x3l: float = 0.0
    x3u: float = 1.0
    gamma: float = 5.0 / 3.0
    mu: float = 1.0
    recon: int = LIM2ND_REC
    flux: int = LHLLC_FLUX
    all_mach: bool = False
    rk_stages: int = 2
    cfl: float = 0.8
    mach_cutoff: float = 1e-5
    mach_ct: float = 1e-6
    tmax: float = 1e30
    bc: Sequence[int] = field(default_factory=lambda: (0, 0, 0, 0, 0, 0))
    use_wb: bool = False
    use_gravity: bool = False
    evolve_etot: bool = False
    use_edot: bool = False
    eos: int = EOS_IDEAL
    fasteos: bool = False
    precision: int = PREC_FP64
    device: int = 0
    # USE_SHOCK_FLATTENING / eps_sf_make (default 0.25, source.F90:177-181,
    # :8073-8131); applied through phlx_set_shock_flattening, not the C struct
    shock_flattening: bool = False
    eps_sf: float = 0.25
    # THERMAL_DIFFUSION_STS (RKL2 around hydro_step), EVALUATE_PARABOLIC_TIMESTEP
    # and FIX_TEMPERATURE_AT_X1L..X3U (wall temperature per side, None = the
    # zero-gradient ghost); applied through phlx_set_thermal_diffusion
    thermal_diffusion: bool = False
    eval_parabolic_dt: bool = True
    fix_temperature: Optional[Sequence[Optional[float]]] = None

    # ------------------------------------------------------------------
    @property
    def mhd(self) -> bool:
        return self.flux in (HLLD_FLUX, LHLLD_FLUX)

    @property
    def nvars(self) -> int:
        """nvars = sdims + 2 + nas (source/source.F90:472-477)."""
        return self.sdims + 2 + self.nas

    @property
    def shape(self) -> tuple:
        """Interior array shape (nz, ny, nx) of one cell-centred field."""
        return (self.nx3 if self.sdims == 3 else 1, self.nx2, self.nx1)

    def face_shape(self, d: int) -> tuple:
        nz, ny, nx = self.shape
        return (nz + (d == 2), ny + (d == 1), nx + (d == 0))

    @property
    def dx(self) -> tuple:
        """Cell widths, computed as the reference does (source.F90:1725-1728)."""
        d1 = (self.x1u - self.x1l) / float(self.nx1)
        d2 = (self.x2u - self.x2l) / float(self.nx2)
        d3 = (self.x3u - self.x3l) / float(self.nx3) if self.sdims == 3 else 1.0
        return (d1, d2, d3)

    def with_(self, **kw) -> "PhlxConfig":
        return replace(self, **kw)

    def validate(self) -> "PhlxConfig":
        """Subset of Make.checks:27-305 that applies to this library."""
        if self.sdims not in (2, 3):
            raise ConfigError("sdims_make must be 2 or 3")
        if min(self.nx1, self.nx2) < 2 or (self.sdims == 3 and self.nx3 < 2):
            raise ConfigError("each nx*_make must be >= 2")
        if self.recon not in (LIM2ND_REC, PPH_REC, LIM5TH_REC):
            raise ConfigError("exactly one of LIM2ND_REC / PPH_REC / LIM5TH_REC")
        if self.flux not in (HLLC_FLUX, LHLLC_FLUX, HLLD_FLUX, LHLLD_FLUX):
            raise ConfigError("flux must be one of HLLC/LHLLC/HLLD/LHLLD")
        if self.rk_stages not in (2, 3):
            raise ConfigError("exactly one of RK2_STEPPER / RK3_STEPPER")
        if self.nas not in (0, 2):
            raise ConfigError("nas_make must be 0 or 2 (ADVECT_YE_IABAR)")
        bc = tuple(self.bc)
        if len(bc) != 6:
            raise ConfigError("bc needs six entries X1L X1U X2L X2U X3L X3U")
        for d in range(self.sdims):
            lo, hi = bc[2 * d], bc[2 * d + 1]
            if (lo == BC_PERIODIC) != (hi == BC_PERIODIC):
                raise ConfigError(f"X{d+1}_PERIODIC must apply to both sides")
            if self.mhd and BC_DIODE in (lo, hi):
                raise ConfigError("DIODE boundaries are not supported with USE_MHD")
        if self.fasteos and self.eos == EOS_IDEAL:
            raise ConfigError("USE_FASTEOS requires a complex EoS (Make.checks:221-226)")
        if self.eos not in (EOS_IDEAL, EOS_HELMHOLTZ, EOS_PIG):
            raise ConfigError("eos must be the ideal gas, HELMHOLTZ_EOS or PIG_EOS")
        if self.eos == EOS_HELMHOLTZ and self.nas != 2:
            raise ConfigError("HELMHOLTZ_EOS runs here with ADVECT_YE_IABAR "
                              "(nas_make=2: Ye, 1/abar)")
        if self.eos == EOS_PIG and self.nas not in (0, 2):
            raise ConfigError("PIG_EOS runs with nas_make=0 or 2 (scalars unused by the EoS)")
        if self.eos != EOS_IDEAL and not self.fasteos and self.precision != PREC_FP64:
            raise ConfigError("HELMHOLTZ_EOS without USE_FASTEOS (a T(rho, P) Newton per "
                              "face state) runs in FP64 only")
        if self.use_wb and not self.use_gravity:
            raise ConfigError("USE_WB requires USE_GRAVITY")
        if self.shock_flattening:
            if self.precision != PREC_FP64:
                raise ConfigError("USE_SHOCK_FLATTENING runs in FP64 only")
            if min(self.nx1, self.nx2) < 3 or (self.sdims == 3 and self.nx3 < 3):
                raise ConfigError("USE_SHOCK_FLATTENING needs at least 3 cells per axis")
            if not self.eps_sf > 0:
                raise ConfigError("eps_sf_make must be positive")
        if self.thermal_diffusion:
            if self.eos == EOS_IDEAL and self.nas != 0:
                raise ConfigError("THERMAL_DIFFUSION_STS with the ideal gas runs with "
                                  "nas_make=0 (a constant mean molecular weight)")
            if self.fix_temperature is not None:
                if len(tuple(self.fix_temperature)) != 6:
                    raise ConfigError("fix_temperature needs six entries X1L X1U X2L X2U X3L X3U")
                for d, t in enumerate(self.fix_temperature):
                    if t is not None and (not t > 0 or tuple(self.bc)[d] == BC_PERIODIC):
                        raise ConfigError("FIX_TEMPERATURE_AT_* needs a positive wall "
                                          "temperature on a non-periodic side")
        if self.cfl <= 0 or self.gamma <= 1 or self.mu <= 0:
            raise ConfigError("cfl, gamma-1 and mu must be positive")
        return self

    def to_c(self) -> PhlxConfigC:
        c = PhlxConfigC()
        for name, _ in PhlxConfigC._fields_:
            if name in ("bc", "reserved"):
                continue
            val = getattr(self, name)
            setattr(c, name, int(val) if isinstance(val, bool) else val)
        for i, v in enumerate(tuple(self.bc)):
            c.bc[i] = int(v)
        return c

    @staticmethod
    def from_opts(opts: Sequence[str], **kw) -> "PhlxConfig":
        """Build from reference-style OPTS strings, e.g.
        ``["sdims_make=3", "LHLLD_FLUX", "X2L_REFLECTIVE", "USE_WB"]``."""
        d: dict = {}
        bc = [BC_PERIODIC] * 6
        sides = {"X1L": 0, "X1U": 1, "X2L": 2, "X2U": 3, "X3L": 4, "X3U": 5}
        for o in opts:
            o = o.strip()
            if "=" in o:
                k, v = o.split("=", 1)
                v = v.replace("_rp", "")
                key = {"sdims_make": "sdims", "nx1_make": "nx1", "nx2_make": "nx2",
                       "nx3_make": "nx3", "nas_make": "nas", "cfl_make": "cfl",
                       "Mach_cutoff_make": "mach_cutoff", "Mach_ct_make": "mach_ct",
                       "tmax_make": "tmax", "eps_sf_make": "eps_sf"}.get(k)
                if key is None:
                    continue
                d[key] = float(v) if key in ("cfl", "mach_cutoff", "mach_ct", "tmax",
                                             "eps_sf") else int(v)
            elif o in _RECON:
                d["recon"] = _RECON[o]
            elif o in _FLUX:
                d["flux"] = _FLUX[o]
            elif o == "RK2_STEPPER":
                d["rk_stages"] = 2
            elif o == "RK3_STEPPER":
                d["rk_stages"] = 3
            elif o == "FLUX_IS_ALL_MACH":
                d["all_mach"] = True
            elif o == "HELMHOLTZ_EOS":
                d["eos"] = EOS_HELMHOLTZ
            elif o == "PIG_EOS":
                d["eos"] = EOS_PIG
            elif o == "USE_FASTEOS":
                d["fasteos"] = True
            elif o == "USE_SHOCK_FLATTENING":
                d["shock_flattening"] = True
            elif o == "THERMAL_DIFFUSION_STS":
                d["thermal_diffusion"] = True
                d.setdefault("eval_parabolic_dt", False)
            elif o == "EVALUATE_PARABOLIC_TIMESTEP":
                d["eval_parabolic_dt"] = True
            elif o in ("USE_WB", "USE_GRAVITY", "EVOLVE_ETOT", "USE_EDOT"):
                d[o.lower()] = True
            elif o[:3] in sides and o[3:4] == "_" and o[4:] in _BC:
                bc[sides[o[:3]]] = _BC[o[4:]]
            elif o[:2] in ("X1", "X2", "X3") and o[2:] == "_PERIODIC":
                a = int(o[1]) - 1
                bc[2 * a] = bc[2 * a + 1] = BC_PERIODIC
        d["bc"] = tuple(bc)
        d.update(kw)
        return PhlxConfig(**d).validate()
