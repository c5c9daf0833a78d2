"""Basis / density containers and their on-disk forms.

A *case* is exactly what the reference handle API consumes:
``liberi_setup(nsh, natoms, num_bas, mxgtot, ang_mom, contr_num, sh_loc,
atom_num, atom_loc, start_bas, end_bas, exponents, contr_coef_s/p/d/f, coords,
schwrz_int, ...)`` (src/interface/liberi_interface.F90:72-129 of gms_libERI)
followed by ``liberi_fock_build(density, fock, n)`` (:202-217).  All integer
arrays keep the reference's 1-based conventions (``sh_loc`` is the first
primitive, ``atom_loc`` the first basis function of a shell, ``ang_mom`` 1..4 =
s..f); coordinates are bohr; density is packed lower-triangular with
``idx(i,j) = i(i-1)/2 + j``.

Two serialisations:
  * ``.npz`` (numpy, no pickles) for committed fixtures;
  * a flat little-endian stream read by the Fortran driver of the flang
    reference build (``oracle/ref_driver.f90``).
"""
from __future__ import annotations

import dataclasses
import io
from pathlib import Path

import numpy as np

MAGIC = int.from_bytes(b"LERI", "little")
VERSION = 1

INT_FIELDS = ("ang_mom", "contr_num", "sh_loc", "atom_num", "atom_loc", "start_bas", "end_bas")
REAL_FIELDS = ("exponents", "coef_s", "coef_p", "coef_d", "coef_f", "coords", "schwrz")


@dataclasses.dataclass
class Basis:
    nsh: int
    natoms: int
    num_bas: int
    mxgtot: int
    ang_mom: np.ndarray
    contr_num: np.ndarray
    sh_loc: np.ndarray
    atom_num: np.ndarray
    atom_loc: np.ndarray
    start_bas: np.ndarray
    end_bas: np.ndarray
    exponents: np.ndarray
    coef_s: np.ndarray
    coef_p: np.ndarray
    coef_d: np.ndarray
    coef_f: np.ndarray
    coords: np.ndarray
    schwrz: np.ndarray

    def __post_init__(self):
        for f in INT_FIELDS:
            setattr(self, f, np.ascontiguousarray(getattr(self, f), dtype=np.int32))
        for f in REAL_FIELDS:
            setattr(self, f, np.ascontiguousarray(getattr(self, f), dtype=np.float64))
        self.validate()

    @property
    def ntri(self) -> int:
        return self.num_bas * (self.num_bas + 1) // 2

    def validate(self) -> None:
        nsh, m = self.nsh, self.mxgtot * self.nsh
        for f in INT_FIELDS:
            if getattr(self, f).shape != (nsh,):
                raise ValueError(f"{f} must have nsh={nsh} entries")
        for f in ("exponents", "coef_s", "coef_p", "coef_d", "coef_f"):
            if getattr(self, f).shape != (m,):
                raise ValueError(f"{f} must have mxgtot*nsh={m} entries")
        if self.coords.shape != (3 * self.natoms,):
            raise ValueError("coords must have 3*natoms entries")
        if self.schwrz.shape != (nsh * (nsh + 1) // 2,):
            raise ValueError("schwrz must have nsh*(nsh+1)/2 entries")

    def with_schwarz(self, schwrz: np.ndarray) -> "Basis":
        return dataclasses.replace(self, schwrz=np.asarray(schwrz, dtype=np.float64))

    def shell_sizes(self) -> np.ndarray:
        return np.array([(l) * (l + 1) // 2 for l in self.ang_mom + 0], dtype=np.int64)


@dataclasses.dataclass
class Case:
    name: str
    basis: Basis
    density: np.ndarray
    # optional expected packed Fock (golden) and its provenance
    fock: np.ndarray | None = None
    source: str = ""

    def save_npz(self, path: str | Path) -> None:
        arrs = {f: getattr(self.basis, f) for f in INT_FIELDS + REAL_FIELDS}
        arrs["dims"] = np.array([self.basis.nsh, self.basis.natoms, self.basis.num_bas, self.basis.mxgtot], np.int64)
        arrs["density"] = np.asarray(self.density, np.float64)
        if self.fock is not None:
            arrs["fock"] = np.asarray(self.fock, np.float64)
        arrs["name"] = np.array(self.name)
        arrs["source"] = np.array(self.source)
        np.savez_compressed(path, **arrs)

    @staticmethod
    def load_npz(path: str | Path) -> "Case":
        z = np.load(path, allow_pickle=False)
        nsh, natoms, num_bas, mxgtot = (int(v) for v in z["dims"])
        basis = Basis(nsh, natoms, num_bas, mxgtot, *(z[f] for f in INT_FIELDS + REAL_FIELDS))
        fock = z["fock"] if "fock" in z.files else None
        return Case(str(z["name"]), basis, z["density"], fock, str(z["source"]))

    def write_stream(self, path: str | Path) -> None:
        b = self.basis
        buf = io.BytesIO()
        buf.write(np.array([MAGIC, VERSION, b.nsh, b.natoms, b.num_bas, b.mxgtot], "<i4").tobytes())
        for f in INT_FIELDS:
            buf.write(getattr(b, f).astype("<i4").tobytes())
        for f in REAL_FIELDS:
            buf.write(getattr(b, f).astype("<f8").tobytes())
        buf.write(np.asarray(self.density, "<f8").tobytes())
        Path(path).write_bytes(buf.getvalue())

    @staticmethod
    def read_stream(path: str | Path, name: str = "") -> "Case":
        raw = Path(path).read_bytes()
        hdr = np.frombuffer(raw, "<i4", 6)
        if hdr[0] != MAGIC:
            raise ValueError("not a case stream")
        nsh, natoms, num_bas, mxgtot = (int(v) for v in hdr[2:6])
        off = 24
        ints = []
        for _ in INT_FIELDS:
            ints.append(np.frombuffer(raw, "<i4", nsh, off).copy())
            off += 4 * nsh
        sizes = [mxgtot * nsh] * 5 + [3 * natoms, nsh * (nsh + 1) // 2]
        reals = []
        for n in sizes:
            reals.append(np.frombuffer(raw, "<f8", n, off).copy())
            off += 8 * n
        ntri = num_bas * (num_bas + 1) // 2
        dens = np.frombuffer(raw, "<f8", ntri, off).copy()
        return Case(name, Basis(nsh, natoms, num_bas, mxgtot, *ints, *reals), dens)


def tri_index(i: int, j: int) -> int:
    """0-based packed index of (i, j), i.e. the reference's i(i-1)/2+j minus 1."""
    if i < j:
        i, j = j, i
    return i * (i + 1) // 2 + j


def pack(m: np.ndarray) -> np.ndarray:
    """Lower triangle of a symmetric matrix in the reference's packed order."""
    n = m.shape[0]
    il = np.tril_indices(n)
    order = np.lexsort((il[1], il[0]))
    return np.ascontiguousarray(m[il[0][order], il[1][order]])


def unpack(p: np.ndarray) -> np.ndarray:
    ntri = p.shape[0]
    n = int((np.sqrt(8 * ntri + 1) - 1) // 2)
    m = np.zeros((n, n))
    il = np.tril_indices(n)
    order = np.lexsort((il[1], il[0]))
    m[il[0][order], il[1][order]] = p
    return m + np.tril(m, -1).T
