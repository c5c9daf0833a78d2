"""Seeded synthetic workloads for the Fock build.

Each generator returns a :class:`~gms_liberi_amd.caseio.Case` in the
reference's input conventions (liberi_setup arguments,
src/interface/liberi_interface.F90:72-129 of gms_libERI):

* ``water``   -- O + 2 H, s/p only (13 basis functions), D = 2 C C^T with 5
  occupied columns of a seeded QR;
* ``ethane``  -- C2H6 staggered, s/p/d (42 basis functions), 9 occupied;
* ``c60``     -- 60 C on a truncated icosahedron of radius 3.55 A (900 basis
  functions), 180 occupied;
* ``cn_chain``-- 20-atom planar C/N zig-zag with one f shell per atom (620
  basis functions), 70 occupied;
* ``cluster`` -- random H/C/N/O cluster for the request stream.

Contraction coefficients carry the GAMESS primitive normalisation the
reference's tests use (tests/fortran/test_h2o_ccpvtz.F90:449-475:
s (2a/pi)^3/4, p 2 sqrt(a) (2a/pi)^3/4, d 4a (2a/pi)^3/4,
f 8 a^3/2 (2a/pi)^3/4); the contraction weights are U(0,1), scaled so the
contracted function built from the component those factors normalise (x, xy,
xyz) has unit self-overlap.  Exponents are log-uniform and sorted descending.

The Schwarz integrals are not produced here: the caller passes a function
``schwarz(basis) -> array`` (the device generator of the product, or a CPU
engine in tests), or gets zeros.
"""
from __future__ import annotations

import math

import numpy as np

from .caseio import Basis, Case

ANG2BOHR = 1.8897261254578281

# shells per element: (L, contraction degree, exponent range)
HEAVY_SPD = [(0, 6, "wide"), (0, 3, "wide"), (0, 1, "wide"), (1, 3, "wide"), (1, 1, "wide"), (2, 1, "d")]
HEAVY_SP = [(0, 6, "wide"), (0, 3, "wide"), (0, 1, "wide"), (1, 3, "wide"), (1, 1, "wide")]
HEAVY_SPDF = HEAVY_SPD + [(2, 1, "d"), (3, 1, "f")]
HYDROGEN = [(0, 3, "h"), (0, 1, "h")]
RANGES = {"wide": (0.1, 5.0e3), "h": (0.1, 20.0), "d": (0.3, 1.5), "f": (0.5, 2.0)}


def _prim_norm(l: int, a: np.ndarray) -> np.ndarray:
    base = (2.0 * a / math.pi) ** 0.75
    return base * [1.0, 2.0 * np.sqrt(a), 4.0 * a, 8.0 * a * np.sqrt(a)][l] if l else base


def _contraction(rng: np.random.Generator, l: int, k: int, lo: float, hi: float):
    expo = np.sort(np.exp(rng.uniform(np.log(lo), np.log(hi), k)))[::-1]
    w = rng.uniform(0.0, 1.0, k)
    # self-overlap of sum_i w_i N_i g_i, g = (x y z)-type component of order l
    p = expo[:, None] + expo[None, :]
    n = _prim_norm(l, expo)
    s = (np.outer(w * n, w * n) * (math.pi / p) ** 1.5 * (0.5 / p) ** l).sum()
    return expo, w * n / math.sqrt(s)


def build_case(name: str, symbols, xyz_ang, shells_of, density_occ: int, seed: int) -> Case:
    rng = np.random.default_rng(seed)
    coords = np.asarray(xyz_ang, float).reshape(-1) * ANG2BOHR
    ang, ncon, shloc, atnum, atloc = [], [], [], [], []
    prims = []  # (l, expo, coef)
    bf = 0
    for ia, sym in enumerate(symbols):
        for (l, k, rname) in shells_of(sym):
            lo, hi = RANGES[rname]
            e, c = _contraction(rng, l, k, lo, hi)
            ang.append(l + 1)
            ncon.append(k)
            atnum.append(ia + 1)
            atloc.append(bf + 1)
            bf += (l + 1) * (l + 2) // 2
            prims.append((l, e, c))
    nsh = len(ang)
    mxgtot = max(ncon)
    expo = np.zeros(mxgtot * nsh)
    coefs = [np.zeros(mxgtot * nsh) for _ in range(4)]
    off = 0
    for s, (l, e, c) in enumerate(prims):
        shloc.append(off + 1)
        expo[off:off + len(e)] = e
        coefs[l][off:off + len(e)] = c
        off += len(e)
    atloc_a = np.array(atloc)
    sizes = np.array([(a) * (a + 1) // 2 for a in ang])
    basis = Basis(nsh, len(symbols), bf, mxgtot, np.array(ang), np.array(ncon), np.array(shloc), np.array(atnum),
                  atloc_a, atloc_a, atloc_a + sizes - 1, expo, *coefs, coords,
                  np.zeros(nsh * (nsh + 1) // 2))
    dens = random_density(bf, density_occ, rng)
    return Case(name, basis, dens)


def random_density(nbf: int, nocc: int, rng: np.random.Generator) -> np.ndarray:
    """Packed D = 2 C C^T with C the first nocc columns of a seeded QR."""
    q, _ = np.linalg.qr(rng.standard_normal((nbf, nbf)))
    c = q[:, :nocc]
    d = 2.0 * c @ c.T
    il = np.tril_indices(nbf)
    return np.ascontiguousarray(d[il])


def perturbed_density(d_packed: np.ndarray, nocc: int, eps: float, rng: np.random.Generator) -> np.ndarray:
    """Projector of an orthonormalised C + eps*G, as the SCF-like iterations of the request stream."""
    ntri = d_packed.shape[0]
    n = int((math.isqrt(8 * ntri + 1) - 1) // 2)
    full = np.zeros((n, n))
    il = np.tril_indices(n)
    full[il] = d_packed
    full = full + np.tril(full, -1).T
    w, v = np.linalg.eigh(full)
    c = v[:, np.argsort(w)[::-1][:nocc]]
    q, _ = np.linalg.qr(c + eps * rng.standard_normal(c.shape))
    d = 2.0 * q @ q.T
    return np.ascontiguousarray(d[il])


def _heavy(sym):
    return HYDROGEN if sym == "H" else HEAVY_SPD


def water(seed: int = 11) -> Case:
    ang = math.radians(104.5)
    r = 0.96
    xyz = [(0, 0, 0), (r, 0, 0), (r * math.cos(ang), r * math.sin(ang), 0)]
    return build_case("water", ["O", "H", "H"], xyz, lambda s: HYDROGEN if s == "H" else HEAVY_SP, 5, seed)


def ethane(seed: int = 12) -> Case:
    cc, ch = 1.54, 1.09
    th = math.acos(-1.0 / 3.0)  # tetrahedral angle
    xyz = [(0, 0, cc / 2), (0, 0, -cc / 2)]
    for k in range(3):  # top methyl, staggered against the bottom one
        phi = 2 * math.pi * k / 3
        xyz.append((ch * math.sin(th) * math.cos(phi), ch * math.sin(th) * math.sin(phi), cc / 2 - ch * math.cos(th)))
    for k in range(3):
        phi = 2 * math.pi * k / 3 + math.pi / 3
        xyz.append((ch * math.sin(th) * math.cos(phi), ch * math.sin(th) * math.sin(phi), -cc / 2 + ch * math.cos(th)))
    return build_case("ethane", ["C", "C"] + ["H"] * 6, xyz, _heavy, 9, seed)


def truncated_icosahedron(radius: float) -> np.ndarray:
    phi = (1 + 5 ** 0.5) / 2
    base = [(0, 1, 3 * phi), (1, 2 + phi, 2 * phi), (phi, 2, 2 * phi + 1)]
    pts = set()
    for b in base:
        for sx in (-1, 1):
            for sy in (-1, 1):
                for sz in (-1, 1):
                    v = (b[0] * sx, b[1] * sy, b[2] * sz)
                    for k in range(3):  # cyclic permutations
                        pts.add(tuple(round(v[(i + k) % 3], 12) for i in range(3)))
    p = np.array(sorted(pts))
    assert p.shape == (60, 3), p.shape
    return p * (radius / np.linalg.norm(p[0]))


def c60(seed: int = 13) -> Case:
    return build_case("c60", ["C"] * 60, truncated_icosahedron(3.55), _heavy, 180, seed)


def c60_fragment(natoms: int = 12, seed: int = 13) -> Case:
    """The `natoms` cage atoms nearest atom 0, with the C60 basis recipe."""
    pts = truncated_icosahedron(3.55)
    idx = np.argsort(np.linalg.norm(pts - pts[0], axis=1))[:natoms]
    return build_case(f"c{natoms}", ["C"] * natoms, pts[idx], _heavy, 3 * natoms, seed)


def cn_chain(seed: int = 14, natoms: int = 20) -> Case:
    bond, ang = 1.45, math.radians(112.0)
    dx, dy = bond * math.sin(ang / 2), bond * math.cos(ang / 2)
    xyz = [(i * dx, (i % 2) * dy, 0.0) for i in range(natoms)]
    syms = ["C" if i % 2 == 0 else "N" for i in range(natoms)]
    return build_case("cn_chain", syms, xyz, lambda s: HEAVY_SPDF, 70, seed)


def cluster(natoms: int, seed: int) -> Case:
    """Random H/C/N/O cluster (weights .5/.3/.1/.1), d_min 1.0 A, liquid density."""
    rng = np.random.default_rng(seed)
    radius = (3 * natoms / (4 * math.pi * 0.1)) ** (1 / 3)
    pts = []
    while len(pts) < natoms:
        p = rng.uniform(-radius, radius, 3)
        if np.linalg.norm(p) > radius:
            continue
        if pts and np.min(np.linalg.norm(np.array(pts) - p, axis=1)) < 1.0:
            continue
        pts.append(p)
    syms = list(rng.choice(["H", "C", "N", "O"], size=natoms, p=[0.5, 0.3, 0.1, 0.1]))
    nelec = sum({"H": 1, "C": 6, "N": 7, "O": 8}[s] for s in syms)
    return build_case(f"cluster{natoms}_{seed}", syms, pts, _heavy, max(1, nelec // 2), seed + 1000)


def cluster12(seed: int = 3) -> Case:
    """A 12-atom request of the serving stream's recipe (fixture size)."""
    case = cluster(12, seed)
    case.name = "cluster12"
    return case


GENERATORS = {"water": water, "ethane": ethane, "c60": c60, "cn_chain": cn_chain, "c12": c60_fragment,
              "cluster12": cluster12}


def make_case(name: str, schwarz=None) -> Case:
    case = GENERATORS[name]()
    if schwarz is not None:
        case.basis = case.basis.with_schwarz(schwarz(case.basis))
    return case
