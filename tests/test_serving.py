"""Request-stream helpers of the serving loop (CPU): percentiles, the
SCF-like densities, the stream generator and the footprint estimate."""
import numpy as np

from gms_liberi_amd import serving


def test_percentile_nearest_rank():
    x = list(range(1, 101))
    assert serving.percentile(x, 50) == 50
    assert serving.percentile(x, 99) == 99
    assert serving.percentile([3.0], 99) == 3.0


def _unpack(p):
    n = int((np.sqrt(8 * p.size + 1) - 1) // 2)
    m = np.zeros((n, n))
    m[np.tril_indices(n)] = p
    return m + np.tril(m, -1).T


def test_scf_like_densities_are_projectors():
    ds = serving.scf_like_densities(30, 7, 4, seed=3)
    assert len(ds) == 4
    for d in ds:
        m = _unpack(d) / 2.0
        assert np.allclose(m @ m, m, atol=1e-12)  # idempotent
        assert abs(np.trace(m) - 7) < 1e-12       # nocc electrons pairs
    # the perturbation shrinks from one iteration to the next
    steps = [np.max(np.abs(ds[k + 1] - ds[k])) for k in range(3)]
    assert steps[2] < steps[0]


def test_request_stream_is_seeded_and_poisson_ordered():
    a = serving.request_stream(3, rate=5.0, seed=9, iterations=2, atoms=(10, 14))
    b = serving.request_stream(3, rate=5.0, seed=9, iterations=2, atoms=(10, 14))
    assert [r.arrival for r in a] == [r.arrival for r in b]
    assert all(np.array_equal(x.densities[1], y.densities[1]) for x, y in zip(a, b))
    assert a[0].arrival == 0.0 and all(a[i].arrival < a[i + 1].arrival for i in range(2))
    for r in a:
        assert 10 <= r.basis.natoms <= 14 and len(r.densities) == 2
        assert r.densities[0].shape == (r.basis.ntri,)


def test_footprint_estimate_grows_with_the_molecule():
    small = serving.request_stream(1, 1.0, seed=1, iterations=1, atoms=(10, 10))[0].basis
    large = serving.request_stream(1, 1.0, seed=1, iterations=1, atoms=(60, 60))[0].basis
    assert serving.estimate_device_bytes(large) > 4 * serving.estimate_device_bytes(small)
