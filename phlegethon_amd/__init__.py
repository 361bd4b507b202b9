"""phlegethon_amd -- MI355X-native explicit finite-volume MHD update.

Drop-in device implementation of Phlegethon's hydro_step path
(source/source.F90:6847-13701): directional sweeps with PLM/parabolic
reconstruction and (L)HLLC / (L)HLLD Riemann fluxes, well-balanced
gravity, constrained transport, RK2/RK3 stepping and the CFL time step,
behind the C-ABI of include/phlx.h.
"""
from .config import (PhlxConfig, ConfigError, LIM2ND_REC, PPH_REC, HLLC_FLUX,
                     LHLLC_FLUX, HLLD_FLUX, LHLLD_FLUX, BC_PERIODIC, BC_REFLECTIVE,
                     BC_OUTFLOW, BC_DIODE)

__all__ = ["PhlxConfig", "ConfigError", "LIM2ND_REC", "PPH_REC", "HLLC_FLUX",
           "LHLLC_FLUX", "HLLD_FLUX", "LHLLD_FLUX", "BC_PERIODIC", "BC_REFLECTIVE",
           "BC_OUTFLOW", "BC_DIODE", "Simulation"]


def __getattr__(name):
    if name == "Simulation":
        from .runtime import Simulation
        return Simulation
    raise AttributeError(name)
