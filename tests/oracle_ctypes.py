"""ctypes binding of the CPU oracle (oracle/liberi_oracle.so) -- test infrastructure."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from gms_liberi_amd.caseio import Basis

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "oracle" / "liberi_oracle.so"
NCLASS = 55


class _OBasis(C.Structure):
    _fields_ = [("nsh", C.c_int32), ("natoms", C.c_int32), ("num_bas", C.c_int32), ("mxgtot", C.c_int32)] + [
        (f, C.POINTER(C.c_int32)) for f in ("ang_mom", "contr_num", "sh_loc", "atom_num", "atom_loc")
    ] + [(f, C.POINTER(C.c_double)) for f in ("exponents", "coef_s", "coef_p", "coef_d", "coef_f", "coords", "schwrz")]


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Oracle:
    def __init__(self, path: Path = LIB):
        if not Path(path).exists():
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        self.lib = C.CDLL(str(path))
        L = self.lib
        L.oracle_class_name.restype = C.c_char_p
        L.oracle_screen_counts.argtypes = [C.POINTER(_OBasis), C.c_double, C.POINTER(C.c_int64)]
        L.oracle_fock.argtypes = [C.POINTER(_OBasis), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double,
                                  C.POINTER(C.c_int64), C.c_int64, C.c_int]
        L.oracle_fock.restype = C.c_int64
        L.oracle_fock_classes.argtypes = L.oracle_fock.argtypes + [C.POINTER(C.c_int32)]
        L.oracle_fock_classes.restype = C.c_int64
        L.oracle_fock_classes_sw.argtypes = L.oracle_fock_classes.argtypes + [C.c_int]
        L.oracle_fock_classes_sw.restype = C.c_int64
        L.oracle_schwarz.argtypes = [C.POINTER(_OBasis), C.POINTER(C.c_double)]
        L.oracle_shell_quartet.argtypes = [C.POINTER(_OBasis), C.c_int, C.c_int, C.c_int, C.c_int,
                                           C.POINTER(C.c_double)]
        L.oracle_boys.argtypes = [C.c_int, C.c_double]
        L.oracle_boys.restype = C.c_double
        self.class_names = [L.oracle_class_name(i).decode() for i in range(NCLASS)]
        L.oracle_class_def.argtypes = [C.c_int, C.POINTER(C.c_int)]
        self.class_defs = []  # (name, bra pair class, ket pair class, "gt"|"ge"|"none", "cross"|"self")
        for i in range(NCLASS):
            d = (C.c_int * 4)()
            L.oracle_class_def(i, d)
            self.class_defs.append((self.class_names[i], d[0], d[1], ("gt", "ge", "none")[d[2]],
                                    ("cross", "self")[d[3]]))

    def _basis(self, b: Basis, schwrz=None):
        keep = []
        ob = _OBasis(b.nsh, b.natoms, b.num_bas, b.mxgtot)
        for f in ("ang_mom", "contr_num", "sh_loc", "atom_num", "atom_loc"):
            a = np.ascontiguousarray(getattr(b, f), np.int32)
            keep.append(a)
            setattr(ob, f, _ptr(a, C.c_int32))
        for f in ("exponents", "coef_s", "coef_p", "coef_d", "coef_f", "coords", "schwrz"):
            a = np.ascontiguousarray(getattr(b, f) if schwrz is None or f != "schwrz" else schwrz, np.float64)
            keep.append(a)
            setattr(ob, f, _ptr(a, C.c_double))
        return ob, keep

    def screen_counts(self, b: Basis, cutoff: float = 5.00000000000000018e-11) -> np.ndarray:
        ob, keep = self._basis(b)
        out = np.zeros(NCLASS, np.int64)
        self.lib.oracle_screen_counts(C.byref(ob), cutoff, _ptr(out, C.c_int64))
        return out

    def fock(self, b: Basis, density, cutoff=5.00000000000000018e-11, stride=1, nthreads=0, classes=None,
             ref_switch=False):
        """Packed Fock matrix (exact integrals, or with ref_switch the
        reference's large-argument Boys values past its switch points), the
        per-class quartet counts and the number of quartets evaluated."""
        ob, keep = self._basis(b)
        d = np.ascontiguousarray(density, np.float64)
        f = np.zeros(b.ntri)
        counts = np.zeros(NCLASS, np.int64)
        en = None
        if classes is not None:
            en = np.array([1 if (n in classes or i in classes) else 0 for i, n in enumerate(self.class_names)],
                          np.int32)
        n = self.lib.oracle_fock_classes_sw(C.byref(ob), _ptr(d, C.c_double), _ptr(f, C.c_double), cutoff,
                                            _ptr(counts, C.c_int64), stride, nthreads,
                                            None if en is None else _ptr(en, C.c_int32), int(ref_switch))
        if n < 0:
            raise RuntimeError("oracle_fock failed")
        return f, counts, n

    def schwarz(self, b: Basis) -> np.ndarray:
        ob, keep = self._basis(b, schwrz=np.zeros(b.nsh * (b.nsh + 1) // 2))
        out = np.zeros(b.nsh * (b.nsh + 1) // 2)
        rc = self.lib.oracle_schwarz(C.byref(ob), _ptr(out, C.c_double))
        if rc != 0:
            raise RuntimeError("oracle_schwarz failed")
        return out

    def quartet(self, b: Basis, i, j, k, l) -> np.ndarray:
        ob, keep = self._basis(b)
        nc = [(b.ang_mom[s]) * (b.ang_mom[s] + 1) // 2 for s in (i, j, k, l)]
        out = np.zeros(int(np.prod(nc)))
        rc = self.lib.oracle_shell_quartet(C.byref(ob), i, j, k, l, _ptr(out, C.c_double))
        if rc != 0:
            raise RuntimeError("bad shell index")
        return out.reshape(nc)

    def boys(self, m: int, t: float) -> float:
        return self.lib.oracle_boys(m, t)
