"""Python mirror of the reference handle API over the HIP C ABI.

Same names, argument meaning and error behaviour as gms_libERI's
``liberi_interface`` (src/interface/liberi_interface.F90:23-262) and
``gamess_liberi_wrapper`` (src/interface/gms_liberi_wrapper.F90:29-146):

    h = liberi_create()
    liberi_setup(h, nsh, natoms, num_bas, mxgtot, ang_mom, contr_num, sh_loc,
                 atom_num, atom_loc, start_bas, end_bas, exponents,
                 contr_coef_s, contr_coef_p, contr_coef_d, contr_coef_f,
                 coords, schwrz_int, my_rank, num_procs)
    liberi_fock_build(h, density, fock, matrix_size)   # fock overwritten in place
    liberi_cleanup(h); liberi_destroy(h)

Where the reference stops the process (``error stop 1``: Fock build before
setup, size mismatch in the wrapper) these raise :class:`LiberiError`.
The library is loaded from the package directory; a missing library raises
at import of the handle, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from .caseio import Basis

LIB_PATH = Path(__file__).resolve().parent / "libliberi_hip.so"
NCLASS = 55

OPT_SCHWARZ_CUTOFF = 1
OPT_VALUE_CUTOFF = 2
OPT_FP32_TAU = 3
OPT_EXACT_LARGE_T = 4
OPT_CLASS_TIMING = 5
OPT_ONLY_CLASS = 6
OPT_PRIM_CUTOFF = 7
OPT_ROW_TILES = 8
OPT_AUTOTUNE = 9
OPT_GRAPH = 10
OPT_MAX_DEVICE_BYTES = 11
OPT_RANK_SPLIT = 12
OPT_FUSED = 13
OPT_TILE_LEN = 14
OPT_ROW_ORIENT = 15

_STATUS = {1: "EARG", 2: "ESTATE", 3: "ESIZE", 4: "EDEVICE", 5: "ENOMEM"}


class LiberiError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"liberi {_STATUS.get(code, code)}: {msg}")
        self.code = code


class Stats(C.Structure):
    _fields_ = [("quartets", C.c_int64 * NCLASS), ("prim_quartets", C.c_int64 * NCLASS),
                ("total_quartets", C.c_int64), ("fp32_quartets", C.c_int64), ("last_build_ms", C.c_double),
                ("device_bytes", C.c_int64), ("class_ms", C.c_double * NCLASS), ("prim_pairs", C.c_int64),
                ("prim_pairs_dropped", C.c_int64), ("class_variant", C.c_int32 * NCLASS),
                ("class_fp32", C.c_int32 * NCLASS), ("class_tile", C.c_int32 * NCLASS)]


_lib = None


def load_library(path: Path = LIB_PATH) -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not Path(path).exists():
        raise FileNotFoundError(f"{path} is not built: run `python -m gms_liberi_amd.build`")
    lib = C.CDLL(str(path))
    i32, i64, dbl, vp = C.c_int32, C.c_int64, C.c_double, C.c_void_p
    pi32, pdbl = C.POINTER(i32), C.POINTER(dbl)
    basis_args = [i32, i32, i32, i32] + [pi32] * 7 + [pdbl] * 7 + [i32, i32]
    lib.liberi_c_create.argtypes = [C.POINTER(vp)]
    lib.liberi_c_destroy.argtypes = [C.POINTER(vp)]
    lib.liberi_c_setup.argtypes = [vp] + basis_args
    lib.liberi_c_fock_build.argtypes = [vp, pdbl, pdbl, i64]
    lib.liberi_c_fock_build_device.argtypes = [vp, vp, vp, i64, vp]
    lib.liberi_c_cleanup.argtypes = [vp]
    lib.liberi_c_last_error.restype = C.c_char_p
    lib.liberi_c_set_option.argtypes = [vp, i32, dbl]
    lib.liberi_c_get_option.argtypes = [vp, i32, pdbl]
    lib.liberi_c_get_stats.argtypes = [vp, C.POINTER(Stats)]
    lib.liberi_c_class_name.argtypes = [i32]
    lib.liberi_c_class_name.restype = C.c_char_p
    lib.liberi_c_init.argtypes = basis_args
    lib.liberi_c_fock.argtypes = [pdbl, pdbl, i64]
    lib.liberi_c_finalize.argtypes = []
    lib.liberi_c_schwarz.argtypes = [i32, i32, i32, i32] + [pi32] * 5 + [pdbl] * 7
    lib.liberi_c_schwarz_noalloc.argtypes = [i32, i32, i32, i32] + [pi32] * 5 + [pdbl] * 7 + [vp]
    lib.liberi_c_fock_build_uhf.argtypes = [vp, pdbl, pdbl, pdbl, pdbl, i64]
    lib.liberi_c_fock_build_uhf_device.argtypes = [vp, vp, vp, vp, vp, i64, vp]
    _lib = lib
    return lib


def schwarz(basis: Basis, noalloc: bool = False, stream: int | None = None) -> np.ndarray:
    """Schwarz integrals sqrt(max|(ab|ab)|) per shell pair, computed on the GPU,
    in the layout of liberi_setup's schwrz_int (packed i(i-1)/2 + j, 1-based).

    noalloc=True calls liberi_c_schwarz_noalloc instead: no device allocation,
    staging in the library's fixed device array, on `stream` (a hipStream_t
    address; None = the default stream)."""
    lib = load_library()
    b = basis
    keep, args = [], [b.nsh, b.natoms, b.num_bas, b.mxgtot]
    for f in ("ang_mom", "contr_num", "sh_loc", "atom_num", "atom_loc"):
        a, p = _i32(getattr(b, f))
        keep.append(a)
        args.append(p)
    for f in ("exponents", "coef_s", "coef_p", "coef_d", "coef_f", "coords"):
        a, p = _f64(getattr(b, f))
        keep.append(a)
        args.append(p)
    out = np.zeros(b.nsh * (b.nsh + 1) // 2)
    if noalloc:
        _check(lib.liberi_c_schwarz_noalloc(*args, out.ctypes.data_as(C.POINTER(C.c_double)), stream))
    else:
        _check(lib.liberi_c_schwarz(*args, out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


def _check(rc: int) -> None:
    if rc != 0:
        raise LiberiError(rc, _lib.liberi_c_last_error().decode())


def _i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(C.POINTER(C.c_int32))


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def class_names() -> list[str]:
    lib = load_library()
    return [lib.liberi_c_class_name(i).decode() for i in range(NCLASS)]


class Handle:
    """An opaque liberi handle (liberi_handle_t, src/interface/liberi_interface.F90:30-42)."""

    def __init__(self):
        self._lib = load_library()
        self._h = C.c_void_p()
        _check(self._lib.liberi_c_create(C.byref(self._h)))
        self.num_bas = 0
        self.is_initialized = False

    # -- lifecycle
    def setup(self, basis: Basis, my_rank: int = 0, num_procs: int = 1) -> None:
        b = basis
        keep = []
        args = [b.nsh, b.natoms, b.num_bas, b.mxgtot]
        for f in ("ang_mom", "contr_num", "sh_loc", "atom_num", "atom_loc", "start_bas", "end_bas"):
            a, p = _i32(getattr(b, f))
            keep.append(a)
            args.append(p)
        for f in ("exponents", "coef_s", "coef_p", "coef_d", "coef_f", "coords", "schwrz"):
            a, p = _f64(getattr(b, f))
            keep.append(a)
            args.append(p)
        # a failed setup leaves the handle released (not set up)
        self.is_initialized = False
        _check(self._lib.liberi_c_setup(self._h, *args, my_rank, num_procs))
        self.num_bas = b.num_bas
        self.is_initialized = True

    def fock_build(self, density: np.ndarray, fock: np.ndarray | None = None) -> np.ndarray:
        d, pd = _f64(density)
        if fock is None:
            fock = np.empty(d.shape[0])
        if fock.dtype != np.float64 or not fock.flags.c_contiguous:
            raise ValueError("fock must be a contiguous float64 array")
        _check(self._lib.liberi_c_fock_build(self._h, pd, fock.ctypes.data_as(C.POINTER(C.c_double)), d.shape[0]))
        return fock

    def fock_build_uhf(self, density_a: np.ndarray, density_b: np.ndarray):
        """Unrestricted Fock matrices (Fa, Fb) = S o (J[Da+Db] - K[Da]), S o (J[Da+Db] - K[Db])."""
        da, pa = _f64(density_a)
        db, pb = _f64(density_b)
        if da.shape != db.shape:
            raise ValueError("alpha and beta densities differ in size")
        fa, fb = np.empty(da.shape[0]), np.empty(da.shape[0])
        P = C.POINTER(C.c_double)
        _check(self._lib.liberi_c_fock_build_uhf(self._h, pa, pb, fa.ctypes.data_as(P), fb.ctypes.data_as(P),
                                                 da.shape[0]))
        return fa, fb

    def fock_build_device(self, d_ptr: int, f_ptr: int, n: int, stream: int = 0) -> None:
        """Build on device-resident arrays (e.g. torch.Tensor.data_ptr()); asynchronous on `stream`."""
        _check(self._lib.liberi_c_fock_build_device(self._h, C.c_void_p(d_ptr), C.c_void_p(f_ptr), n,
                                                    C.c_void_p(stream or None)))

    def cleanup(self) -> None:
        _check(self._lib.liberi_c_cleanup(self._h))
        self.is_initialized = False

    def destroy(self) -> None:
        if self._h:
            _check(self._lib.liberi_c_destroy(C.byref(self._h)))
        self.is_initialized = False

    def __del__(self):
        try:
            if self._h:
                self._lib.liberi_c_destroy(C.byref(self._h))
        except Exception:
            pass

    # -- knobs and counters
    def set_option(self, opt: int, value: float) -> None:
        _check(self._lib.liberi_c_set_option(self._h, opt, value))

    def get_option(self, opt: int) -> float:
        v = C.c_double()
        _check(self._lib.liberi_c_get_option(self._h, opt, C.byref(v)))
        return v.value

    def stats(self) -> dict:
        s = Stats()
        _check(self._lib.liberi_c_get_stats(self._h, C.byref(s)))
        return {"quartets": np.array(s.quartets[:], np.int64), "prim_quartets": np.array(s.prim_quartets[:], np.int64),
                "total_quartets": s.total_quartets, "fp32_quartets": s.fp32_quartets,
                "last_build_ms": s.last_build_ms, "device_bytes": s.device_bytes,
                "class_ms": np.array(s.class_ms[:], np.float64), "prim_pairs": s.prim_pairs,
                "prim_pairs_dropped": s.prim_pairs_dropped,
                "class_variant": np.array(s.class_variant[:], dtype=np.int32),
                "class_fp32": np.array(s.class_fp32[:], dtype=np.int32),
                "class_tile": np.array(s.class_tile[:], dtype=np.int32)}


# ---------------------------------------------- reference-named functions
def liberi_create() -> Handle:
    return Handle()


def liberi_destroy(handle: Handle) -> None:
    handle.destroy()


def liberi_setup(handle: Handle, nsh, natoms, num_bas, mxgtot, ang_mom, contr_num, sh_loc, atom_num, atom_loc,
                 start_bas, end_bas, exponents, contr_coef_s, contr_coef_p, contr_coef_d, contr_coef_f, coords,
                 schwrz_int, my_rank=0, num_procs=1) -> None:
    b = Basis(nsh, natoms, num_bas, mxgtot, ang_mom, contr_num, sh_loc, atom_num, atom_loc, start_bas, end_bas,
              exponents, contr_coef_s, contr_coef_p, contr_coef_d, contr_coef_f, coords, schwrz_int)
    handle.setup(b, my_rank, num_procs)


def liberi_fock_build(handle: Handle, density_in: np.ndarray, fock_out: np.ndarray, matrix_size: int) -> None:
    if density_in.shape[0] != matrix_size or fock_out.shape[0] != matrix_size:
        raise LiberiError(3, "array length differs from matrix_size")
    handle.fock_build(density_in, fock_out)


def liberi_cleanup(handle: Handle) -> None:
    handle.cleanup()


# ------------------------------------------------ GAMESS direct-link wrapper
def liberi_init(natoms, num_bas, nsh, mxgtot, ang_mom, contr_num, sh_loc, atom_num, atom_loc, start_bas, end_bas,
                exponents, contr_coef_s, contr_coef_p, contr_coef_d, contr_coef_f, coords, schwrz_int, my_rank=0,
                num_procs=1) -> None:
    lib = load_library()
    b = Basis(nsh, natoms, num_bas, mxgtot, ang_mom, contr_num, sh_loc, atom_num, atom_loc, start_bas, end_bas,
              exponents, contr_coef_s, contr_coef_p, contr_coef_d, contr_coef_f, coords, schwrz_int)
    keep, args = [], [b.natoms, b.num_bas, b.nsh, b.mxgtot]
    for f in ("ang_mom", "contr_num", "sh_loc", "atom_num", "atom_loc", "start_bas", "end_bas"):
        a, p = _i32(getattr(b, f))
        keep.append(a)
        args.append(p)
    for f in ("exponents", "coef_s", "coef_p", "coef_d", "coef_f", "coords", "schwrz"):
        a, p = _f64(getattr(b, f))
        keep.append(a)
        args.append(p)
    _check(lib.liberi_c_init(*args, my_rank, num_procs))


def liberi_fock(density: np.ndarray, fock: np.ndarray, n: int) -> None:
    lib = load_library()
    d, pd = _f64(density)
    if fock.dtype != np.float64 or not fock.flags.c_contiguous:
        raise ValueError("fock must be a contiguous float64 array")
    _check(lib.liberi_c_fock(pd, fock.ctypes.data_as(C.POINTER(C.c_double)), n))


def liberi_finalize() -> None:
    _check(load_library().liberi_c_finalize())
