"""Pins of the Rys quadrature tables the HIP kernels evaluate (no GPU).

The generated table (gms_liberi_amd/csrc/rys_table.inc) is evaluated here
exactly as the kernels do (interval lookup + Horner in double) and checked
against what the mathematics fixes: an n-root Rys rule integrates u^k exactly
for k < 2n, i.e.  sum_i w_i u_i^k = F_k(X)  (Boys function, incomplete gamma
in mpmath), for X inside the tabulated range and in the asymptotic range.
"""
import re
from pathlib import Path

import numpy as np
import pytest

INC = Path(__file__).resolve().parents[1] / "gms_liberi_amd" / "csrc" / "rys_table.inc"


def _parse():
    text = INC.read_text()
    deg = int(re.search(r"#define RYS_DEG (\d+)", text).group(1))
    h = float(re.search(r"#define RYS_H ([0-9.]+)", text).group(1))
    nsmall = int(re.search(r"#define RYS_NSMALL (\d+)", text).group(1))
    deg_s = int(re.search(r"#define RYS_DEG_SMALL (\d+)", text).group(1))
    h_s = float(re.search(r"#define RYS_H_SMALL ([0-9.]+)", text).group(1))
    degs = [deg_s if n <= nsmall else deg for n in range(8)]
    hs = [h_s if n <= nsmall else h for n in range(8)]

    def arr(name):
        body = re.search(rf"RYS_CONST \w+ {name}\[[^=]*=\s*\{{(.*?)\}};", text, re.S).group(1)
        toks = re.findall(r"-?0x[0-9a-fA-F.]+p[-+]?\d+|-?\d+\.\d*(?:e[-+]?\d+)?|-?\d+", body)
        return [float.fromhex(t) if "0x" in t else float(t) for t in toks]

    nint = [int(v) for v in arr("rys_nint")]
    off = [int(v) for v in arr("rys_offset")]
    xmax = arr("rys_xmax")
    hu = np.array(arr("rys_herm_u")[1:]).reshape(7, 7)  # leading {0} row collapses to one token
    hw = np.array(arr("rys_herm_w")[1:]).reshape(7, 7)
    table = np.array(arr("rys_table"))
    return degs, hs, nint, off, xmax, hu, hw, table


DEGS, HS, NINT, OFF, XMAX, HU, HW, TABLE = _parse()


def rule(n, X):
    if X < XMAX[n]:
        DEG, H = DEGS[n], HS[n]
        j = int(X * (1.0 / H))
        s = X * (2.0 / H) - (2 * j + 1)
        base = OFF[n] + j * (2 * n * (DEG + 1))
        vals = []
        for r in range(2 * n):
            c = TABLE[base + r * (DEG + 1): base + (r + 1) * (DEG + 1)]
            acc = c[DEG]
            for k in range(DEG - 1, -1, -1):
                acc = acc * s + c[k]
            vals.append(acc)
        return np.array(vals[:n]), np.array(vals[n:])
    return HU[n - 1, :n] / X, HW[n - 1, :n] / np.sqrt(X)


def test_table_layout():
    assert TABLE.shape[0] == sum(NINT[n] * 2 * n * (DEGS[n] + 1) for n in range(1, 8))
    for n in range(1, 8):
        assert abs(NINT[n] * HS[n] - XMAX[n]) < 1e-12


@pytest.mark.parametrize("n", range(1, 8))
def test_moments_equal_boys(n):
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    rng = np.random.default_rng(n)
    xs = list(rng.uniform(0, XMAX[n], 12)) + [0.0, 1e-9, XMAX[n] - 1e-9, XMAX[n], XMAX[n] + 3.0, 25.0, 37.0, 49.0,
                                              65.0, 140.0]
    worst = 0.0
    for X in xs:
        u, w = rule(n, X)
        assert np.all(u > 0) and np.all(u < 1) and np.all(w > 0)
        assert np.all(np.diff(u) > 0)
        for k in range(2 * n):
            got = float(np.sum(w * u ** k))
            if X == 0:
                ex = mp.mpf(1) / (2 * k + 1)
            else:
                ex = mp.gammainc(k + mp.mpf(1) / 2, 0, X) / (2 * mp.mpf(X) ** (k + mp.mpf(1) / 2))
            worst = max(worst, abs(got - float(ex)) / float(ex))
    # the highest moments amplify root errors by k * |du|/u; 1e-13 keeps
    # integrals at ~1e-14 relative
    assert worst < 1e-13, worst


def test_one_root_rule_is_boys():
    mp = pytest.importorskip("mpmath")
    for X in (0.3, 5.0, 20.0, 38.9):
        u, w = rule(1, X)
        f0 = mp.sqrt(mp.pi / (4 * X)) * mp.erf(mp.sqrt(X))
        f1 = (f0 - mp.exp(-X)) / (2 * X)
        assert abs(w[0] - float(f0)) <= 1e-15 * float(f0)
        assert abs(u[0] - float(f1 / f0)) <= 2e-15
