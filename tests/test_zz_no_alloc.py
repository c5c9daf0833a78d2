"""liberi_c_schwarz_noalloc: the Schwarz integrals without device allocation.

The entry point stages its records in the library's fixed device array
instead of hipMalloc'd buffers and runs the same kernels as liberi_c_schwarz,
so its results must equal that entry point's bit for bit, on consecutive
calls with different bases (the array keeps the previous call's contents) and
on a caller's stream.  A basis whose staging does not fit the array is
rejected before any device work, so that check runs without a GPU.
"""
from pathlib import Path

import numpy as np
import pytest

from gms_liberi_amd import hipmem, liberi
from gms_liberi_amd.caseio import Basis, Case

GOLD = Path(__file__).resolve().parent / "golden"
SCRATCH_BYTES = 64 << 20  # LIBERI_SCHWARZ_SCRATCH_BYTES


def _big_basis(nsh=150, k=10):
    """s shells of k primitives each, one per atom on a line: nsh(nsh+1)/2 * k^2
    primitive pairs (1.13M for 150 x 10, 72 MB of records)."""
    rng = np.random.default_rng(3)
    exps = np.sort(rng.uniform(0.1, 50.0, (nsh, k)), axis=1)[:, ::-1].ravel()
    coef = rng.uniform(0.1, 1.0, nsh * k)
    idx = np.arange(nsh)
    return Basis(nsh=nsh, natoms=nsh, num_bas=nsh, mxgtot=k, ang_mom=np.ones(nsh), contr_num=np.full(nsh, k),
                 sh_loc=idx * k + 1, atom_num=idx + 1, atom_loc=idx + 1, start_bas=idx + 1, end_bas=idx + 1,
                 exponents=exps, coef_s=coef, coef_p=np.zeros(nsh * k), coef_d=np.zeros(nsh * k),
                 coef_f=np.zeros(nsh * k), coords=np.stack([idx * 2.0, 0 * idx, 0 * idx], 1).ravel(),
                 schwrz=np.zeros(nsh * (nsh + 1) // 2))


def test_basis_above_the_staging_size_is_rejected():
    b = _big_basis()
    nprim = b.nsh * (b.nsh + 1) // 2 * b.mxgtot ** 2
    assert nprim * 64 > SCRATCH_BYTES
    with pytest.raises(liberi.LiberiError) as e:
        liberi.schwarz(b, noalloc=True)
    assert e.value.code == 3, e.value
    assert "staging" in str(e.value) and "liberi_c_schwarz" in str(e.value)


@pytest.mark.gpu
def test_noalloc_equals_allocating_entry_point():
    big = Case.load_npz(GOLD / "cfg_c60.npz").basis  # 64980 shell pairs, ~26 MB of records
    small = Case.load_npz(GOLD / "cfg_ethane.npz").basis
    want_big = liberi.schwarz(big)
    want_small = liberi.schwarz(small)
    print(f"c60: {big.nsh} shells; ethane: {small.nsh} shells", flush=True)
    # two calls in a row, the second with a smaller basis, on the default stream
    got_big = liberi.schwarz(big, noalloc=True)
    got_small = liberi.schwarz(small, noalloc=True)
    assert np.array_equal(got_big, want_big)
    assert np.array_equal(got_small, want_small)
    # the values are the Schwarz integrals the fixtures hold (device engine vs
    # the oracle's, relative 1e-10 where it matters for screening)
    ref = Case.load_npz(GOLD / "cfg_ethane.npz").basis.schwrz
    m = ref >= 1e-4
    assert np.max(np.abs(got_small[m] - ref[m]) / ref[m]) <= 1e-10
    # on a caller's stream, after a smaller call left its contents behind
    s = hipmem.Stream()
    got_big2 = liberi.schwarz(big, noalloc=True, stream=s.address)
    got_small2 = liberi.schwarz(small, noalloc=True, stream=s.address)
    s.destroy()
    assert np.array_equal(got_big2, want_big)
    assert np.array_equal(got_small2, want_small)
