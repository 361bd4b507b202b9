"""Helmholtz free-energy table of the electron-positron gas.

The reference reads Timmes' ``helm_table.dat`` (source/source.F90:23077-23156,
downloaded by data/get_data.sh); the file is not available offline, so the
table is computed by ``phlegethon_amd/helmgen/helmgen.c`` (Fermi-Dirac
integrals + truncated-Taylor derivatives) and cached under
``phlegethon_amd/_data``.  The binary layout is
``"PHLXHLM1" | int32 nrho, nT | float64 ldlo, ldhi, ltlo, lthi |
f9[nT][nrho][9] | c4[nT][nrho][4]`` with the columns in the reference file's
order (f, fd, ft, fdd, ftt, fdt, fddt, fdtt, fddtt; c, cd, ct, cdt).
"""
from __future__ import annotations

import hashlib
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GEN_SRC = os.path.join(HERE, "helmgen", "helmgen.c")
GEN_BIN = os.path.join(HERE, "_lib", "helmgen")
DATA_DIR = os.path.join(HERE, "_data")

# Defaults: 0.05 dex in rho*Ye over [-12, 15] and 10/270 dex in T over [3, 13]
# (541 x 271, the table size of the benchmark configurations).
DEFAULT = dict(nrho=541, nT=271, ldlo=-12.0, ldhi=15.0, ltlo=3.0, lthi=13.0)


@dataclass
class HelmTable:
    nrho: int
    nT: int
    ldlo: float
    ldhi: float
    ltlo: float
    lthi: float
    f9: np.ndarray      # (nT, nrho, 9), file column order
    c4: np.ndarray      # (nT, nrho, 4)
    kind: str = "helm"  # "helm": e-e+ free energy at rho*Ye; "pig": the whole gas at rho

    @property
    def rho_nodes(self) -> np.ndarray:
        d = (self.ldhi - self.ldlo) / (self.nrho - 1)
        return 10.0 ** (self.ldlo + np.arange(self.nrho) * d)

    @property
    def T_nodes(self) -> np.ndarray:
        d = (self.lthi - self.ltlo) / (self.nT - 1)
        return 10.0 ** (self.ltlo + np.arange(self.nT) * d)


def build_generator() -> str:
    os.makedirs(os.path.dirname(GEN_BIN), exist_ok=True)
    subprocess.run(["gcc", "-O2", "-fopenmp", "-o", GEN_BIN, GEN_SRC, "-lm"], check=True)
    return GEN_BIN


def _stamp(args) -> str:
    h = hashlib.sha256()
    h.update(open(GEN_SRC, "rb").read())
    h.update(repr(tuple(args)).encode())
    return h.hexdigest()


def table_path(nrho, nT, ldlo, ldhi, ltlo, lthi) -> str:
    return os.path.join(DATA_DIR, f"helm_{nrho}x{nT}_{ldlo:g}_{ldhi:g}_{ltlo:g}_{lthi:g}.bin")


def generate(nrho, nT, ldlo, ldhi, ltlo, lthi, path: str = None, ascii_path: str = None,
             force: bool = False) -> str:
    """Run the generator (cached by a stamp of its source and arguments)."""
    args = (int(nrho), int(nT), float(ldlo), float(ldhi), float(ltlo), float(lthi))
    path = path or table_path(*args)
    stamp = path + ".stamp"
    st = _stamp(args)
    if not force and os.path.exists(path) and os.path.exists(stamp) and \
            open(stamp).read().strip() == st and ascii_path is None:
        return path
    if not os.path.exists(GEN_BIN):
        build_generator()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    cmd = [GEN_BIN] + [str(a) for a in args] + [path]
    if ascii_path:
        cmd.append(ascii_path)
    subprocess.run(cmd, check=True)
    with open(stamp, "w") as fh:
        fh.write(st + "\n")
    return path


def load(path: str) -> HelmTable:
    with open(path, "rb") as fh:
        magic = fh.read(8)
        if magic not in (b"PHLXHLM1", b"PHLXPIG1"):
            raise ValueError(f"{path}: not a phlx Helmholtz or PIG table")
        nr, nt = np.frombuffer(fh.read(8), dtype=np.int32)
        box = np.frombuffer(fh.read(32), dtype=np.float64)
        f9 = np.frombuffer(fh.read(8 * 9 * nr * nt), dtype=np.float64).reshape(nt, nr, 9)
        c4 = np.frombuffer(fh.read(8 * 4 * nr * nt), dtype=np.float64).reshape(nt, nr, 4)
    return HelmTable(int(nr), int(nt), *map(float, box), f9.copy(), c4.copy(),
                     kind="pig" if magic == b"PHLXPIG1" else "helm")


def default_table() -> HelmTable:
    return load(generate(**DEFAULT))


def sha256(path: str) -> str:
    return hashlib.sha256(open(path, "rb").read()).hexdigest()
