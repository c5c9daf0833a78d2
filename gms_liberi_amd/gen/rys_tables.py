#!/usr/bin/env python3
"""Generate the Rys quadrature tables used by the HIP kernels.

For n = 1..7 roots, the Rys rule on [0, 1] with weight exp(-X t^2) is the
Gauss rule for the measure  d mu_X(u) = exp(-X u) u^-1/2 / 2 du  on u = t^2 in
[0, 1], whose moments are the Boys functions  int u^k d mu_X = F_k(X).
Roots u_i(X) and weights w_i(X) are computed here in high precision
(mpmath): moments F_0..F_{2n-1}, Chebyshev's moment algorithm for the
three-term recurrence, Golub-Welsch eigen-decomposition of the Jacobi matrix.

On [0, X_n) they are stored as piecewise polynomials: intervals of width H
(H_SMALL for n <= NSMALL), per interval a polynomial of degree DEG (DEG_SMALL)
in s = (X - x_mid) / (H / 2) in [-1, 1] for every u_i and w_i (fitted at Chebyshev nodes, then converted to
the power basis so the kernels evaluate with Horner).  For X >= X_n the
asymptotic rule is exact to double precision: with h_i, W_i the positive
Gauss-Hermite nodes/weights of order 2n,  u_i = h_i^2 / X,  w_i = W_i / sqrt(X).
X_n is the first multiple of H where that asymptotic form is within 1e-16
relative of the exact rule.

Output: gms_liberi_amd/csrc/rys_table.inc (C source, hex-float literals).
Run:    python -m gms_liberi_amd.gen.rys_tables   (a few minutes, all cores)
"""
from __future__ import annotations

import math
import os
import sys
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

import mpmath as mp
import numpy as np

NMAX = 7
H = 0.5          # interval width in X
DEG = 9          # polynomial degree per interval
# the one- and two-root rules (the s and p classes, whose kernels read the
# table the most) on intervals of H_SMALL with degree DEG_SMALL: 20 % fewer
# coefficients per evaluation, fitted to 2e-15 relative or better
NSMALL = 2
H_SMALL = 0.25
DEG_SMALL = 7
DPS = 60


def h_of(n: int) -> float:
    return H_SMALL if n <= NSMALL else H


def deg_of(n: int) -> int:
    return DEG_SMALL if n <= NSMALL else DEG

OUT = Path(__file__).resolve().parents[1] / "csrc" / "rys_table.inc"


def boys_mp(k: int, x):
    x = mp.mpf(x)
    if x == 0:
        return mp.mpf(1) / (2 * k + 1)
    return mp.gammainc(k + mp.mpf(1) / 2, 0, x) / (2 * x ** (k + mp.mpf(1) / 2))


def rys_rule(n: int, x) -> tuple[list, list]:
    """Gauss rule (u_i ascending, w_i) for d mu_X, in mp precision."""
    mom = [boys_mp(k, x) for k in range(2 * n)]
    # Chebyshev algorithm (Gautschi, "Orthogonal Polynomials", alg. 2.1)
    sig_prev = [mp.mpf(0)] * (2 * n)
    sig = list(mom)
    alpha = [mom[1] / mom[0]]
    beta = [mom[0]]
    for k in range(1, n):
        sig_new = [mp.mpf(0)] * (2 * n)
        for l in range(k, 2 * n - k):
            sig_new[l] = sig[l + 1] - alpha[k - 1] * sig[l] - beta[k - 1] * sig_prev[l]
        alpha.append(sig_new[k + 1] / sig_new[k] - sig[k] / sig[k - 1])
        beta.append(sig_new[k] / sig[k - 1])
        sig_prev, sig = sig, sig_new
    J = mp.matrix(n, n)
    for i in range(n):
        J[i, i] = alpha[i]
        if i + 1 < n:
            J[i, i + 1] = J[i + 1, i] = mp.sqrt(beta[i + 1])
    if n == 1:
        return [alpha[0]], [beta[0]]
    ev, vec = mp.eigsy(J)
    order = sorted(range(n), key=lambda i: ev[i])
    return [ev[i] for i in order], [beta[0] * vec[0, i] ** 2 for i in order]


def hermite_half(n: int):
    """Positive nodes/weights of 2n-point Gauss-Hermite (weight e^-s^2 on R)."""
    nodes, weights = np.polynomial.hermite.hermgauss(2 * n)
    # refine in mp: Newton on H_{2n}
    out_h, out_w = [], []
    for s in nodes[n:]:
        s = mp.mpf(s)
        for _ in range(8):
            h, dh = mp.hermite(2 * n, s), 2 * (2 * n) * mp.hermite(2 * n - 1, s)
            s = s - h / dh
        w = (2 ** (2 * n - 1) * mp.factorial(2 * n) * mp.sqrt(mp.pi)) / ((2 * n) ** 2 * mp.hermite(2 * n - 1, s) ** 2)
        out_h.append(s)
        out_w.append(w)
    return out_h, out_w


def asymptotic(n: int, x):
    hs, ws = hermite_half(n)
    x = mp.mpf(x)
    return [h * h / x for h in hs], [w / mp.sqrt(x) for w in ws]


def find_xmax(n: int) -> float:
    x = 10.0
    while True:
        u, w = rys_rule(n, x)
        ua, wa = asymptotic(n, x)
        err = max(max(abs((a - b) / b) for a, b in zip(ua, u)), max(abs((a - b) / b) for a, b in zip(wa, w)))
        if err < 1e-16:
            return x
        x += H


def fit_interval(args):
    n, j = args
    mp.mp.dps = DPS
    h, deg = h_of(n), deg_of(n)
    x0 = j * h
    xm = x0 + h / 2
    # Chebyshev nodes of the first kind on [-1, 1]
    m = deg + 1
    s = [math.cos(math.pi * (k + 0.5) / m) for k in range(m)]
    vals = np.zeros((2 * n, m))
    for k, sk in enumerate(s):
        u, w = rys_rule(n, mp.mpf(xm) + mp.mpf(sk) * h / 2)
        vals[:n, k] = [float(v) for v in u]
        vals[n:, k] = [float(v) for v in w]
    # interpolate, convert to power basis in s
    coefs = np.zeros((2 * n, m))
    V = np.polynomial.chebyshev.chebvander(np.array(s), deg)
    for r in range(2 * n):
        c = np.linalg.solve(V, vals[r])
        coefs[r] = np.polynomial.chebyshev.cheb2poly(c)
    # check at off-node points against the exact rule, evaluating as the kernel does
    err = 0.0
    for sk in np.linspace(-1, 1, 7):
        u, w = rys_rule(n, mp.mpf(xm) + mp.mpf(float(sk)) * h / 2)
        ex = [float(v) for v in u] + [float(v) for v in w]
        for r in range(2 * n):
            acc = 0.0
            for c in coefs[r][::-1]:
                acc = acc * sk + c
            err = max(err, abs(acc - ex[r]) / abs(ex[r]))
    return n, j, coefs, err


def main():
    mp.mp.dps = DPS
    xmax = {}
    for n in range(1, NMAX + 1):
        xmax[n] = find_xmax(n)
        print(f"n={n}: asymptotic from X={xmax[n]}", flush=True)
    jobs = [(n, j) for n in range(1, NMAX + 1) for j in range(int(round(xmax[n] / h_of(n))))]
    res = {}
    worst = {}
    with ProcessPoolExecutor(max_workers=os.cpu_count()) as ex:
        for n, j, coefs, err in ex.map(fit_interval, jobs, chunksize=4):
            res[(n, j)] = coefs
            worst[n] = max(worst.get(n, 0.0), err)
    for n in range(1, NMAX + 1):
        print(f"n={n}: worst relative fit error {worst[n]:.2e}")
    herm = {n: hermite_half(n) for n in range(1, NMAX + 1)}

    lines = ["// Generated by gms_liberi_amd/gen/rys_tables.py -- do not edit.",
             "// Rys roots u_i = t_i^2 and weights w_i for n = 1..%d roots." % NMAX,
             "// Piecewise: interval j covers X in [j*H, (j+1)*H); value_r(X) = sum_k c[k] s^k,",
             "// s = (X - (j+0.5)*H) / (H/2).  Per interval the 2n rows are u_1..u_n, w_1..w_n",
             "// (roots ascending), each DEG+1 coefficients c[0..DEG]; H = RYS_H and DEG = RYS_DEG,",
             "// except for n <= RYS_NSMALL: H = RYS_H_SMALL, DEG = RYS_DEG_SMALL.",
             "// For X >= rys_xmax[n]: u_i = rys_herm_u[n][i] / X, w_i = rys_herm_w[n][i] / sqrt(X).",
             f"// Worst relative fit error per n: " + ", ".join(f"{worst[n]:.1e}" for n in range(1, NMAX + 1)),
             "#ifndef RYS_CONST", "#define RYS_CONST static const", "#endif",
             f"#define RYS_NMAX {NMAX}",
             f"#define RYS_DEG {DEG}",
             f"#define RYS_H {H!r}",
             f"#define RYS_INV_H {1.0 / H!r}",
             f"#define RYS_NSMALL {NSMALL}",
             f"#define RYS_DEG_SMALL {DEG_SMALL}",
             f"#define RYS_H_SMALL {H_SMALL!r}",
             f"#define RYS_INV_H_SMALL {1.0 / H_SMALL!r}"]
    nint = {n: int(round(xmax[n] / h_of(n))) for n in range(1, NMAX + 1)}
    off, acc = {}, 0
    for n in range(1, NMAX + 1):
        off[n] = acc
        acc += nint[n] * 2 * n * (deg_of(n) + 1)
    lines.append(f"#define RYS_TABLE_SIZE {acc}")
    lines.append("RYS_CONST int rys_nint[RYS_NMAX + 1] = {0, " + ", ".join(str(nint[n]) for n in range(1, NMAX + 1)) + "};")
    lines.append("RYS_CONST int rys_offset[RYS_NMAX + 1] = {0, " + ", ".join(str(off[n]) for n in range(1, NMAX + 1)) + "};")
    lines.append("RYS_CONST double rys_xmax[RYS_NMAX + 1] = {0.0, " + ", ".join(f"{xmax[n]!r}" for n in range(1, NMAX + 1)) + "};")
    for name, idx in (("rys_herm_u", 0), ("rys_herm_w", 1)):
        rows = []
        for n in range(1, NMAX + 1):
            vals = [float(herm[n][0][i] ** 2) if idx == 0 else float(herm[n][1][i]) for i in range(n)]
            vals += [0.0] * (NMAX - n)
            rows.append("{" + ", ".join(v.hex() for v in vals) + "}")
        lines.append(f"RYS_CONST double {name}[RYS_NMAX + 1][RYS_NMAX] = {{{{0}}, " + ", ".join(rows) + "};")
    lines.append("RYS_CONST double rys_table[RYS_TABLE_SIZE] = {")
    for n in range(1, NMAX + 1):
        for j in range(nint[n]):
            for r in range(2 * n):
                lines.append("  " + ", ".join(float(c).hex() for c in res[(n, j)][r]) + ",")
    lines.append("};")
    OUT.write_text("\n".join(lines) + "\n")
    print(f"wrote {OUT} ({acc} coefficients)")


if __name__ == "__main__":
    sys.exit(main())
