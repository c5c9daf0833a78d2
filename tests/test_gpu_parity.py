"""HIP Fock build against the reference's outputs and the oracle (GPU).

Bar: max-abs Fock difference <= 1e-10 and |dE_2e| <= 1e-9 Eh
against the reference's own output on the same inputs (tests/golden, written
by make_reference_fixtures.py), surviving shell-quartet counts bit-exact.

Large workloads are checked class by class (LIBERI_OPT_ONLY_CLASS) against
the reference build's per-class Fock matrices
(tests/golden/cfg_<name>_ref_classes.npz).  The build's default mode takes
the reference's large-argument Boys values past its switch points (T > 25,
src/liberi_boys.F90:22; X >= 37..65 in its Rys kernels); the oracle restates
that switch (oracle_fock_classes_sw), so the reference's remaining departure
from the oracle is its tabulated Boys values below the switch
(src/liberi_boys.F90:22-32, read e.g. at src/rhf/rot_axis/int0122.F90:465-490).
Where that departure exceeds 1e-10 (named from data by the fixture, e.g.
(sp|dd) int0122 on C60 by 1.97e-9) the class's whole vectors are kept, and the
bar is: the build equals the oracle with the reference's switch to 1e-12
(relative to the largest element) in the default mode and the exact oracle
in the exact mode, and is no farther from the reference than that oracle is.
Every other class, and all of them together over every Fock element, match
the reference to 1e-10.
"""
from pathlib import Path

import numpy as np
import pytest

from gms_liberi_amd import liberi
from gms_liberi_amd.caseio import Case

GOLD = Path(__file__).resolve().parent / "golden"
FOCK_TOL = 1e-10
ENERGY_TOL = 1e-9

REF_CASES = sorted(p.name for p in GOLD.glob("ref_*.npz"))
SMALL_CFG = ["cfg_water.npz", "cfg_ethane.npz"]


def energy(d, f):
    return 0.5 * float(np.dot(d, f))


@pytest.fixture(scope="module")
def handle():
    h = liberi.Handle()
    yield h
    h.destroy()


@pytest.mark.gpu
@pytest.mark.parametrize("name", REF_CASES + SMALL_CFG)
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
    counts = handle.stats()["quartets"]
    ref_counts = oracle.screen_counts(case.basis)
    bad = {oracle.class_names[i]: (int(counts[i]), int(ref_counts[i])) for i in range(55) if counts[i] != ref_counts[i]}
    assert not bad, f"quartet counts differ: {bad}"


CLASS_CFG = ["c60", "cn_chain", "cluster12"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", CLASS_CFG)
def test_workload_class_by_class(name):
    case = Case.load_npz(GOLD / f"cfg_{name}.npz")
    ref = np.load(GOLD / f"cfg_{name}_ref_classes.npz")
    defect = [str(k) for k in ref["defect"]]
    names = liberi.class_names()
    print(f"case {name} nbf={case.basis.num_bas}; reference defect classes {defect}", flush=True)
    h = liberi.Handle()
    h.setup(case.basis)
    full = h.fock_build(case.density)
    counts = h.stats()["quartets"]
    assert np.array_equal(counts, np.load(GOLD / f"cfg_{name}_counts.npy"))
    bad, defect_sum_gpu, defect_sum_ref = [], np.zeros_like(full), np.zeros_like(full)
    for c, k in enumerate(names):
        if counts[c] == 0:
            continue
        h.set_option(liberi.OPT_ONLY_CLASS, c)
        f = h.fock_build(case.density)
        if k in defect:
            r = ref[f"full_{k}"]
            osw = r + ref[f"oracle_sw_minus_ref_{k}"].astype(np.float64)
            d_ref, ref_osw = float(np.max(np.abs(f - r))), float(np.max(np.abs(r - osw)))
            d_osw = float(np.max(np.abs(f - osw)))
            print(f"  {k} (reference defect): |build - reference| {d_ref:.2e}, |build - oracle(switch)| "
                  f"{d_osw:.2e}, |reference - oracle(switch)| {ref_osw:.2e}", flush=True)
            if d_osw > 1e-12 * max(1.0, float(np.max(np.abs(osw)))) or d_ref > ref_osw + 1e-11:
                bad.append(f"{k}: |build - oracle(switch)| {d_osw:.2e}, |build - reference| {d_ref:.2e}")
            defect_sum_gpu += f
            defect_sum_ref += r
        else:
            idx = ref[f"idx_{k}"]
            err = float(np.max(np.abs(f[idx] - ref[f"ref_{k}"])))
            if err > FOCK_TOL:
                bad.append(f"{k}: {err:.2e}")
    h.destroy()
    # the defect classes once more, in the exact mode (read at setup), against
    # the oracle's vectors
    h = liberi.Handle()
    h.set_option(liberi.OPT_EXACT_LARGE_T, 1)
    h.setup(case.basis)
    for k in defect:
        h.set_option(liberi.OPT_ONLY_CLASS, names.index(k))
        f = h.fock_build(case.density)
        exact = ref[f"full_{k}"] + ref[f"oracle_minus_ref_{k}"].astype(np.float64)
        err = float(np.max(np.abs(f - exact)))
        print(f"  {k} exact mode: |build - oracle| {err:.2e}", flush=True)
        if err > 1e-12 * max(1.0, float(np.max(np.abs(exact)))):
            bad.append(f"{k} exact mode: |build - oracle| {err:.2e}")
    h.destroy()
    assert not bad, bad
    # every class free of a reference defect, all together, over every element
    clean = float(np.max(np.abs((full - defect_sum_gpu) - (case.fock - defect_sum_ref))))
    de = abs(energy(case.density, full) - energy(case.density, case.fock))
    print(f"  clean classes together: max|dF| {clean:.2e}; |dE| (all classes) {de:.2e}")
    assert clean <= FOCK_TOL
    assert de <= ENERGY_TOL


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


@pytest.mark.gpu
def test_fragment_classes(oracle):
    """Per class on a 12-atom piece of the cage: the default build reproduces
    the reference class by class, except int0122 where the reference is off
    the exact value; the exact-switch build matches the oracle everywhere."""
    case = Case.load_npz(GOLD / "frag_c12.npz")
    ref = np.load(GOLD / "frag_c12_ref_classes.npz")
    names = liberi.class_names()
    for mode in (0, 1):
        h = liberi.Handle()
        h.set_option(liberi.OPT_EXACT_LARGE_T, mode)
        h.setup(case.basis)
        for k in ref.files:
            h.set_option(liberi.OPT_ONLY_CLASS, names.index(k))
            f = h.fock_build(case.density)
            fo, _, _ = oracle.fock(case.basis, case.density, classes=[k])
            d_ref, d_or, ref_or = (np.max(np.abs(f - ref[k])), np.max(np.abs(f - fo)), np.max(np.abs(ref[k] - fo)))
            print(f"mode {mode} {k}: |gpu-ref| {d_ref:.2e} |gpu-oracle| {d_or:.2e} |ref-oracle| {ref_or:.2e}")
            if mode == 1:
                assert d_or <= 1e-12
            elif k == "int0122":
                assert ref_or > 1e-10  # the reference's own deviation (documented above)
            else:
                assert d_ref <= FOCK_TOL
        h.destroy()
