"""Per-class parameters and the algorithmic operation census of the Fock build.

The 55 classes are the kernels dispatched by src/liberi_driver.F90:69-158 of
gms_libERI, in that order.  ``CLASS_L`` gives (la, lb, lc, ld) with the
higher-L shell first in each pair, as the HIP kernels evaluate them.

FLOP model (FMA = 2, add/sub/mul/div/sqrt = 1), per primitive quartet:
  setup 15 + Rys rule 2*nr*(2*deg) + per root [22 + 3 G tables + 3*ne*nf]
(deg: the table's polynomial degree for nr roots, 7 for nr <= 2, else 9)
per contracted quartet:
  horizontal recurrence 2 per produced element + scale/cutoff 2*nabcd +
  digestion 12*nabcd (six block contractions).
These are the operations of the algorithm as written, not of the compiled
code, and they are what `roofline.achieved` divides by kernel time.
"""
from __future__ import annotations

NAMES = ["int0000", "int0001", "int0011", "int0111", "int0101", "int1111", "int0002", "int0012", "int0102",
         "int0112", "int0022", "int0122", "int1102", "int0202", "int2211", "int2121", "int2120", "int2111",
         "int2220", "int2221", "int2222gen", "int3333", "int0030", "int0031", "int0032", "int1030", "int3330",
         "int3011", "int3020", "int2130", "int2230", "int3030", "int3110", "int3111", "int3120", "int3121",
         "int2231", "int3130", "int3131", "int3210", "int3331", "int3211", "int3220", "int3221", "int3222",
         "int3230", "int3231", "int3232", "int3300", "int3310", "int3311", "int3320", "int3321", "int3322",
         "int3332"]

CLASS_L = [(0, 0, 0, 0), (0, 0, 1, 0), (0, 0, 1, 1), (1, 0, 1, 1), (1, 0, 1, 0), (1, 1, 1, 1), (0, 0, 2, 0),
           (0, 0, 2, 1), (1, 0, 2, 0), (1, 0, 2, 1), (0, 0, 2, 2), (1, 0, 2, 2), (1, 1, 2, 0), (2, 0, 2, 0),
           (2, 2, 1, 1), (2, 1, 2, 1), (2, 1, 2, 0), (2, 1, 1, 1), (2, 2, 2, 0), (2, 2, 2, 1), (2, 2, 2, 2),
           (3, 3, 3, 3), (0, 0, 3, 0), (0, 0, 3, 1), (0, 0, 3, 2), (1, 0, 3, 0), (3, 3, 3, 0), (3, 0, 1, 1),
           (3, 0, 2, 0), (2, 1, 3, 0), (2, 2, 3, 0), (3, 0, 3, 0), (3, 1, 1, 0), (3, 1, 1, 1), (3, 1, 2, 0),
           (3, 1, 2, 1), (2, 2, 3, 1), (3, 1, 3, 0), (3, 1, 3, 1), (3, 2, 1, 0), (3, 3, 3, 1), (3, 2, 1, 1),
           (3, 2, 2, 0), (3, 2, 2, 1), (3, 2, 2, 2), (3, 2, 3, 0), (3, 2, 3, 1), (3, 2, 3, 2), (3, 3, 0, 0),
           (3, 3, 1, 0), (3, 3, 1, 1), (3, 3, 2, 0), (3, 3, 2, 1), (3, 3, 2, 2), (3, 3, 3, 2)]

RYS_DEG = 9
RYS_DEG_SMALL, RYS_NSMALL = 7, 2  # the one- and two-root tables


def rys_deg(nr: int) -> int:
    return RYS_DEG_SMALL if nr <= RYS_NSMALL else RYS_DEG


def ncart(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def nrange(l0: int, l1: int) -> int:
    return sum(ncart(l) for l in range(l0, l1 + 1))


def class_params(c: int) -> dict:
    la, lb, lc, ld = CLASS_L[c]
    lab, lcd = la + lb, lc + ld
    nr = (lab + lcd) // 2 + 1
    ni, nk = lab + 1, lcd + 1
    ne, nf = nrange(la, lab), nrange(lc, lcd)
    na, nb, nc, nd = ncart(la), ncart(lb), ncart(lc), ncart(ld)
    return dict(la=la, lb=lb, lc=lc, ld=ld, nr=nr, ni=ni, nk=nk, ne=ne, nf=nf, nabcd=na * nb * nc * nd, na=na,
                nb=nb, nc=nc, nd=nd)


def flops_per_prim_quartet(c: int) -> int:
    p = class_params(c)
    g = 3 * (p["ni"] - 1) + 5 * p["ni"] * (p["nk"] - 1)  # one 2-D table
    return 15 + 2 * p["nr"] * 2 * rys_deg(p["nr"]) + p["nr"] * (22 + 3 * g + 3 * p["ne"] * p["nf"])


def flops_per_quartet(c: int) -> int:
    p = class_params(c)
    la, lb, lc, ld = p["la"], p["lb"], p["lc"], p["ld"]
    hrr = 0
    for j in range(1, lb + 1):
        hrr += 2 * nrange(la, la + lb - j) * ncart(j) * p["nf"]
    for j in range(1, ld + 1):
        hrr += 2 * p["na"] * p["nb"] * nrange(lc, lc + ld - j) * ncart(j)
    return hrr + 2 * p["nabcd"] + 12 * p["nabcd"]


def class_flops(c: int, quartets: int, prim_quartets: int) -> float:
    return float(quartets) * flops_per_quartet(c) + float(prim_quartets) * flops_per_prim_quartet(c)
