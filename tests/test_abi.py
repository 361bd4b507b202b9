"""CPU-side checks of the boundary: the C-ABI library loads and exports every
entry point declared in include/phlx.h; the ctypes config struct matches the
C layout; configuration validation mirrors the reference's Make.checks
rules.  No compute call is made (no GPU here)."""
import ctypes
import os
import re

import pytest

from phlegethon_amd import config as C
from phlegethon_amd.runtime import LIB_PATH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "phlx.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(phlx_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_core_entry_points():
    syms = declared_symbols()
    for s in ("phlx_create", "phlx_destroy", "phlx_upload_state", "phlx_download_state",
              "phlx_init", "phlx_hydro_step", "phlx_compute_dt", "phlx_run",
              "phlx_sweep", "phlx_ct", "phlx_rk_update", "phlx_error_string"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB_PATH):
        pytest.skip("libphlx.so not built (run __graft_entry__.build())")
    from phlegethon_amd.runtime import load_library
    lib = load_library(LIB_PATH)   # loads torch's HIP runtime first
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    lib.phlx_error_string.restype = ctypes.c_char_p
    assert lib.phlx_error_string(-2) == b"unsupported or inconsistent configuration"
    assert lib.phlx_version() >= 100


def test_config_struct_layout():
    # offsets follow the C declaration order with natural alignment
    c = C.PhlxConfigC
    assert c.x1l.offset == 24
    assert ctypes.sizeof(c) % 8 == 0
    cfg = C.PhlxConfig(sdims=3, nx1=8, nx2=6, nx3=4, bc=(1, 1, 0, 0, 2, 2))
    s = cfg.to_c()
    assert (s.sdims, s.nx1, s.nx2, s.nx3) == (3, 8, 6, 4)
    assert list(s.bc) == [1, 1, 0, 0, 2, 2]


def test_from_opts_matches_reference_flags():
    cfg = C.PhlxConfig.from_opts([
        "sdims_make=3", "nx1_make=32", "nx2_make=16", "nx3_make=8", "cfl_make=0.6_rp",
        "LHLLD_FLUX", "PPH_REC", "RK3_STEPPER", "X1_PERIODIC", "X2L_REFLECTIVE",
        "X2U_REFLECTIVE", "X3_PERIODIC", "Mach_ct_make=1e-6_rp"])
    assert cfg.sdims == 3 and cfg.nx1 == 32 and cfg.nx3 == 8
    assert cfg.flux == C.LHLLD_FLUX and cfg.recon == C.PPH_REC and cfg.rk_stages == 3
    assert tuple(cfg.bc) == (0, 0, 1, 1, 0, 0)
    assert cfg.mhd and cfg.nvars == 5
    assert abs(cfg.cfl - 0.6) < 1e-15


@pytest.mark.parametrize("bad", [
    dict(sdims=4),
    dict(rk_stages=4),
    dict(bc=(0, 1, 0, 0, 0, 0)),                      # periodic on one side only
    dict(flux=C.HLLD_FLUX, bc=(3, 3, 0, 0, 0, 0)),    # DIODE with MHD
    dict(fasteos=True),                               # FASTEOS needs a complex EoS
    dict(use_wb=True),                                # WB needs gravity
    dict(nas=1),
    dict(eos=C.EOS_HELMHOLTZ, nas=0, fasteos=True),   # Helmholtz carries Ye, 1/abar
    dict(eos=C.EOS_HELMHOLTZ, nas=2, fasteos=False,   # per-face Newton: FP64 only
         precision=C.PREC_MIXED),
])
def test_validation_rejects(bad):
    with pytest.raises(C.ConfigError):
        C.PhlxConfig(**bad).validate()
