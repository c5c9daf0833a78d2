"""Edge cases of the HIP Fock build (GPU), checked against the oracle.

* a basis of one s shell: only (ss|ss) runs, one shell quartet;
* one atom carrying every shell type s..f: every class whose shells exist
  runs on a single centre (A = B = C = D, all Schwarz values large);
* a Schwarz cutoff above every product x_ab x_cd: no quartet survives, F is
  exactly zero and no class launches;
* a density of zeros: F is exactly zero.
Tolerance as for the reference fixtures: max-abs 1e-10 on F (measured
7.4e-13 on the one-centre s..f case, whose Fock elements reach ~6e2), quartet
counts bit-exact.
"""
import numpy as np
import pytest

from gms_liberi_amd import configs, liberi

FOCK_TOL = 1e-10


def _case(name, symbols, xyz, shells, nocc, seed, oracle):
    case = configs.build_case(name, symbols, xyz, shells, nocc, seed)
    case.basis = case.basis.with_schwarz(oracle.schwarz(case.basis))
    return case


@pytest.fixture(scope="module")
def handle():
    h = liberi.Handle()
    yield h
    h.destroy()


def _check(handle, oracle, case, density=None):
    d = case.density if density is None else density
    print(f"case {case.name} nbf={case.basis.num_bas} nsh={case.basis.nsh}", flush=True)
    handle.setup(case.basis)
    f = handle.fock_build(d)
    fo, _, _ = oracle.fock(case.basis, d)
    err = float(np.max(np.abs(f - fo)))
    print(f"  max|dF| {err:.2e} (max|F| {np.max(np.abs(fo)):.2e})")
    assert err <= FOCK_TOL
    counts = handle.stats()["quartets"]
    assert np.array_equal(counts, oracle.screen_counts(case.basis))
    return f, counts


@pytest.mark.gpu
def test_one_s_shell(handle, oracle):
    case = _case("one_s", ["He"], [(0.0, 0.0, 0.0)], lambda s: [(0, 3, "wide")], 1, 21, oracle)
    _, counts = _check(handle, oracle, case)
    names = liberi.class_names()
    assert counts[names.index("int0000")] == 1 and counts.sum() == 1


@pytest.mark.gpu
def test_one_atom_every_shell_type(handle, oracle):
    shells = [(0, 6, "wide"), (0, 1, "wide"), (1, 3, "wide"), (1, 1, "wide"), (2, 1, "d"), (2, 1, "d"), (3, 1, "f")]
    case = _case("one_atom_spdf", ["N"], [(0.1, -0.2, 0.3)], lambda s: shells, 4, 22, oracle)
    _, counts = _check(handle, oracle, case)
    assert counts.sum() > 0


@pytest.mark.gpu
def test_everything_screened_out(oracle):
    case = configs.make_case("water", schwarz=oracle.schwarz)
    h = liberi.Handle()
    try:
        h.set_option(liberi.OPT_SCHWARZ_CUTOFF, 1e6)
        h.setup(case.basis)
        f = h.fock_build(case.density)
        assert not np.any(f)
        assert h.stats()["quartets"].sum() == 0
        assert oracle.screen_counts(case.basis, cutoff=1e6).sum() == 0
    finally:
        h.destroy()


@pytest.mark.gpu
def test_zero_density(handle, oracle):
    case = configs.make_case("ethane", schwarz=oracle.schwarz)
    handle.setup(case.basis)
    f = handle.fock_build(np.zeros(case.basis.ntri))
    assert not np.any(f)
