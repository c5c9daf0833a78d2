"""HIP Fock build against the golden fixtures and the oracle (GPU).

Bar: max-abs Fock difference <= 1e-10 and |dE_2e| <= 1e-9 Eh against the
reference's own output (tests/golden, written by make_reference_fixtures.py),
and per-class surviving shell-quartet counts bit-exact against the oracle's
restatement of the reference's screening.
"""
from pathlib import Path

import numpy as np
import pytest

from gms_liberi_amd.caseio import Case
from gms_liberi_amd import liberi

GOLD = Path(__file__).resolve().parent / "golden"
FOCK_TOL = 1e-10
ENERGY_TOL = 1e-9

REF_CASES = sorted(p.name for p in GOLD.glob("ref_*.npz"))
CFG_CASES = sorted(p.name for p in GOLD.glob("cfg_*.npz"))


def energy(d, f):
    return 0.5 * float(np.dot(d, f))


@pytest.fixture(scope="module")
def handle():
    h = liberi.Handle()
    yield h
    h.destroy()


@pytest.mark.gpu
@pytest.mark.parametrize("name", REF_CASES + CFG_CASES)
def test_fock_matches_reference(name, handle, oracle):
    case = Case.load_npz(GOLD / name)
    print(f"case {name} nbf={case.basis.num_bas}", flush=True)
    handle.setup(case.basis)
    fock = handle.fock_build(case.density)
    err = float(np.max(np.abs(fock - case.fock)))
    de = abs(energy(case.density, fock) - energy(case.density, case.fock))
    print(f"  max|dF|={err:.3e} |dE|={de:.3e}")
    assert err <= FOCK_TOL, f"{name}: max|dF| {err:.3e}"
    assert de <= ENERGY_TOL, f"{name}: |dE| {de:.3e}"
    # surviving quartets per class: bit-exact against the oracle's restated predicate
    counts = handle.stats()["quartets"]
    ref_counts = oracle.screen_counts(case.basis)
    bad = {oracle.class_names[i]: (int(counts[i]), int(ref_counts[i])) for i in range(55) if counts[i] != ref_counts[i]}
    assert not bad, f"quartet counts differ: {bad}"


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["ref_h2o_ccpvdz.npz", "cfg_water.npz", "cfg_ethane.npz"])
def test_fock_matches_oracle_random_density(name, handle, oracle):
    """A density unlike the fixtures' (random symmetric, non-idempotent)."""
    case = Case.load_npz(GOLD / name)
    rng = np.random.default_rng(5)
    d = rng.standard_normal(case.basis.ntri)
    handle.setup(case.basis)
    fock = handle.fock_build(d)
    f_or, _, _ = oracle.fock(case.basis, d)
    assert np.max(np.abs(fock - f_or)) <= FOCK_TOL
