"""Serving loop on the GPU: continuous batching of Fock-build requests.

Each request's Fock matrices must be the same run alone and in the mix
(concurrent streams only change the order of the atomic additions: within
1e-11 relative to the largest element), in FP64 and in the mixed form, and
equal to a plain handle build of the same molecule.
"""
import numpy as np
import pytest

from gms_liberi_amd import liberi, serving

MIX_TOL = 1e-11


def _stream(tau=0.0):
    return serving.request_stream(4, rate=50.0, seed=21, iterations=3, tau=tau, atoms=(10, 40))


def _close(a, b):
    scale = max(1.0, float(np.max(np.abs(a))))
    return float(np.max(np.abs(a - b))) / scale


@pytest.mark.gpu
@pytest.mark.parametrize("tau", [0.0, 1e-6])
def test_alone_equals_mix(tau):
    reqs = _stream(tau)
    alone = []
    for r in reqs:  # one request at a time, nothing else on the device
        r0 = serving.Request(r.rid, 0.0, r.basis, r.densities, tau)
        alone += serving.Server(max_active=1).run([r0])
    # all four arrive together: the admission pass puts them on the device at once
    mix_reqs = [serving.Request(r.rid, 0.0, r.basis, r.densities, tau) for r in reqs]
    srv = serving.Server(max_active=4)
    mix = srv.run(mix_reqs)
    assert srv.peak_active >= 2  # the requests did overlap on the device
    for a, m in zip(alone, mix):
        assert a.rid == m.rid and len(a.focks) == len(m.focks) == 3
        for fa, fm in zip(a.focks, m.focks):
            assert _close(fa, fm) <= MIX_TOL
    s = serving.summarize(mix)
    assert s["requests"] == 4 and s["builds"] == 12
    assert s["p99_latency_s"] >= s["p50_latency_s"] > 0


@pytest.mark.gpu
def test_serving_result_equals_handle_build():
    r = _stream()[1]
    res = serving.Server(max_active=2).run([r])[0]
    h = liberi.Handle()
    h.setup(r.basis.with_schwarz(liberi.schwarz(r.basis)))
    for d, f in zip(r.densities, res.focks):
        assert _close(h.fock_build(d), f) <= MIX_TOL
    h.destroy()


@pytest.mark.gpu
def test_admission_by_memory_budget():
    reqs = _stream()
    need = [serving.estimate_device_bytes(r.basis) for r in reqs]
    # room for the largest request alone but never for two of them
    srv = serving.Server(max_active=4, hbm_budget=max(need) + min(need) - 1)
    res = srv.run(reqs)
    assert len(res) == 4
    assert srv.peak_bytes <= srv.hbm_budget or srv.peak_active == 1
    srv2 = serving.Server(max_active=4, hbm_budget=1)  # below any request: one at a time
    srv2.run(reqs[:2])
    assert srv2.peak_active == 1


@pytest.mark.gpu
def test_footprint_estimate_covers_the_handle():
    for r in _stream():
        res = serving.Server(max_active=1).run([r])[0]
        assert serving.estimate_device_bytes(r.basis) >= res.device_bytes
