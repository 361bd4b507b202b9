"""Deliberately broken builds of the CPU oracle, for the mutation tests.

A pin of an oracle function is only worth something if a plausible mistake
in that function fails it.  Each mutant below is a copy of the oracle's C
sources with one textual change (a dropped factor, a swapped upwind side, a
wrong threshold), compiled into its own shared library under
oracle/_mutants/; the pin tests run the same check on the real oracle (must
pass) and on the mutant (must fail).  The oracle's own sources are never
modified.
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import threading

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC = os.path.join(ROOT, "oracle")
OUT = os.path.join(ORC, "_mutants")
SOURCES = ("phlx_oracle.c", "helm_oracle.c", "oracle.h")

# name -> (file, original text, replacement); the original must occur
# exactly once in that file
MUTANTS = {
    # LHLLC: low-dissipation factor phi = chi instead of chi (2 - chi)
    "lhllc_phi_linear": ("phlx_oracle.c",
                         "double phi = chi * (2.0 - chi);",
                         "double phi = chi;"),
    # LHLLC: Mach number from the normal velocity only
    "lhllc_mach_normal": ("phlx_oracle.c",
                          "double machL = sqrt(v2L) / cL, machR = sqrt(v2R) / cR;\n"
                          "    double chi",
                          "double machL = fabs(vx1L) / cL, machR = fabs(vx1R) / cR;\n"
                          "    double chi"),
    # LHLLC: the dissipation term without its factor 1/2
    "lhllc_half": ("phlx_oracle.c",
                   "- 0.5 * phi * rhobar * cbar * (vx1R - vx1L);",
                   "- phi * rhobar * cbar * (vx1R - vx1L);"),
    # LHLLD: phi = chi instead of chi (2 - chi)
    "lhlld_phi_linear": ("phlx_oracle.c",
                         "      phi = chi * (2.0 - chi);",
                         "      phi = chi;"),
    # LHLLD: the low-Mach switch at 0.5 instead of 0.6
    "lhlld_threshold": ("phlx_oracle.c", "if (mx < 0.6) {", "if (mx < 0.5) {"),
    # HLLD: the double-star states with the sign of B_n reversed
    "hlld_sign_bn": ("phlx_oracle.c", "double sgn = copysign(1.0, bx1);",
                     "double sgn = copysign(1.0, -bx1);"),
    # HLLD: the double-star tangential velocity with swapped sqrt(rho) weights
    "hlld_weights": ("phlx_oracle.c",
                     "double vx22star = (sqL * vx2starL + sqR * vx2starR",
                     "double vx22star = (sqR * vx2starL + sqL * vx2starR"),
    # HLLD: the single-star transverse field without the B_n^2 term
    "hlld_bstar": ("phlx_oracle.c",
                   "double bx2starL = bx2L * (rhoL * dsuL * dsuL - bx1_2) * a1_inv;",
                   "double bx2starL = bx2L * (rhoL * dsuL * dsuL) * a1_inv;"),
    # CT-contact: every corner gradient taken from the downwind side
    "ct_downwind": ("phlx_oracle.c",
                    "  double dS = 0.5 * (1.0 + sS) * (eW - eSW) + 0.5 * (1.0 - sS) * (eE - eSE);\n"
                    "  double dN = 0.5 * (1.0 + sN) * (eNW - eW) + 0.5 * (1.0 - sN) * (eNE - eE);\n"
                    "  double dW = 0.5 * (1.0 + sW) * (eS - eSW) + 0.5 * (1.0 - sW) * (eN - eNW);\n"
                    "  double dE = 0.5 * (1.0 + sE) * (eSE - eS) + 0.5 * (1.0 - sE) * (eNE - eN);",
                    "  double dS = 0.5 * (1.0 - sS) * (eW - eSW) + 0.5 * (1.0 + sS) * (eE - eSE);\n"
                    "  double dN = 0.5 * (1.0 - sN) * (eNW - eW) + 0.5 * (1.0 + sN) * (eNE - eE);\n"
                    "  double dW = 0.5 * (1.0 - sW) * (eS - eSW) + 0.5 * (1.0 + sW) * (eN - eNW);\n"
                    "  double dE = 0.5 * (1.0 - sE) * (eSE - eS) + 0.5 * (1.0 + sE) * (eNE - eN);"),
    # CT-contact: no upwinding at all (the arithmetic mean of the two sides)
    "ct_no_upwind": ("phlx_oracle.c",
                     "          if (mach > g->c.mach_ct) s = 1.0;\n"
                     "          if (mach < -g->c.mach_ct) s = -1.0;",
                     "          (void)mach;"),
    # contact sign: the threshold's density without the mean's 1/2
    "sign_density": ("phlx_oracle.c",
                     "double tmp = 0.5 * (qL[(long)f * nv] + qR[(long)f * nv]) * g->smax;",
                     "double tmp = (qL[(long)f * nv] + qR[(long)f * nv]) * g->smax;"),
    # USE_EDOT: heating with the wrong sign
    "edot_sign": ("phlx_oracle.c",
                  "if (g->c.use_edot) R[ip] = R[ip] - g->edot[j];",
                  "if (g->c.use_edot) R[ip] = R[ip] + g->edot[j];"),
    # USE_EDOT: the heating profile read one row off
    "edot_row": ("phlx_oracle.c",
                 "if (g->c.use_edot) R[ip] = R[ip] - g->edot[j];",
                 "if (g->c.use_edot) R[ip] = R[ip] - g->edot[j + 1 < g->n[1] ? j + 1 : j];"),
}

_lock = threading.Lock()


def build(name: str) -> str:
    """Path of the mutant's shared library (built on first use; rebuilt when
    the oracle sources or the mutation change)."""
    fname, old, new = MUTANTS[name]
    texts = {}
    for s in SOURCES:
        with open(os.path.join(ORC, s)) as f:
            texts[s] = f.read()
    n = texts[fname].count(old)
    if n != 1:
        raise AssertionError(f"mutant {name}: the original text occurs {n} times in {fname}")
    texts[fname] = texts[fname].replace(old, new)
    h = hashlib.sha256("".join(texts[s] for s in SOURCES).encode()).hexdigest()[:16]
    d = os.path.join(OUT, name)
    lib = os.path.join(d, f"lib_{h}.so")
    with _lock:
        if os.path.exists(lib):
            return lib
        os.makedirs(d, exist_ok=True)
        for s in SOURCES:
            with open(os.path.join(d, s), "w") as f:
                f.write(texts[s])
    tmp = lib + ".tmp"
    subprocess.run(["gcc", "-O1", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
                    "-fopenmp", "-o", tmp, os.path.join(d, "phlx_oracle.c"),
                    os.path.join(d, "helm_oracle.c"), "-lm"], check=True)
    os.replace(tmp, lib)
    return lib
