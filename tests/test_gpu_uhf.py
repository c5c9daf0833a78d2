"""Unrestricted Fock matrices (the reference's FOCK_B, registered in its MDI
plugin but never computed, src/interface/liberi_mdi_api.F90:404-409).

Pinned by the closed forms Fa = S o (J[Da+Db] - K[Da]),
Fb = S o (J[Da+Db] - K[Db]) with J and K from the oracle's full ERI tensor,
and by the RHF limit Da = Db = D/2.
"""
from pathlib import Path

import numpy as np
import pytest

from gms_liberi_amd import liberi
from gms_liberi_amd.caseio import Case
from tests.test_oracle import unpack

GOLD = Path(__file__).resolve().parent / "golden"
CUT = 5.00000000000000018e-11


def _spin_densities(n, rng, occ_a, occ_b):
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    il = np.tril_indices(n)
    da = (q[:, :occ_a] @ q[:, :occ_a].T)[il]
    db = (q[:, :occ_b] @ q[:, :occ_b].T)[il]
    return np.ascontiguousarray(da), np.ascontiguousarray(db)


@pytest.mark.gpu
def test_uhf_closed_form(oracle):
    case = Case.load_npz(GOLD / "cfg_water.npz")
    b = case.basis.with_schwarz(np.full(case.basis.nsh * (case.basis.nsh + 1) // 2, 1e10))  # no screening
    n = b.num_bas
    eri = np.zeros((n, n, n, n))
    loc = b.atom_loc - 1
    size = b.ang_mom * (b.ang_mom + 1) // 2
    for i in range(b.nsh):
        for j in range(b.nsh):
            for k in range(b.nsh):
                for l in range(b.nsh):
                    eri[loc[i]:loc[i] + size[i], loc[j]:loc[j] + size[j], loc[k]:loc[k] + size[k],
                        loc[l]:loc[l] + size[l]] = oracle.quartet(b, i, j, k, l)
    eri[np.abs(eri) < CUT] = 0.0
    da, db = _spin_densities(n, np.random.default_rng(8), 5, 4)
    Da, Db = unpack(da), unpack(db)
    J = np.einsum("ijkl,kl->ij", eri, Da + Db)
    S = 2.0 - np.eye(n)
    il = np.tril_indices(n)
    want_a = (S * (J - np.einsum("ikjl,kl->ij", eri, Da)))[il]
    want_b = (S * (J - np.einsum("ikjl,kl->ij", eri, Db)))[il]
    h = liberi.Handle()
    h.setup(b)
    fa, fb = h.fock_build_uhf(da, db)
    h.destroy()
    scale = max(1.0, float(np.max(np.abs(want_a))))
    assert np.max(np.abs(fa - want_a)) <= 1e-10 * scale
    assert np.max(np.abs(fb - want_b)) <= 1e-10 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["cfg_ethane.npz", "cfg_cn_chain.npz"])
def test_uhf_reduces_to_rhf(name):
    case = Case.load_npz(GOLD / name)
    h = liberi.Handle()
    h.setup(case.basis)
    f = h.fock_build(case.density)
    half = 0.5 * case.density
    fa, fb = h.fock_build_uhf(half, half)
    # an RHF build after the UHF one still uses its own (graph) path
    f2 = h.fock_build(case.density)
    h.destroy()
    scale = max(1.0, float(np.max(np.abs(f))))
    for g in (fa, fb, f2):
        assert np.max(np.abs(g - f)) <= 1e-11 * scale
