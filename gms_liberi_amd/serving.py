"""Request loop for Fock builds on one GPU, with continuous batching.

A request is one molecule (basis in the liberi_setup convention) and the
densities of its SCF-like iterations.  The loop admits requests as they
arrive, while a slot and enough HBM are free (the admission test uses an
estimate of the handle's device footprint: shell pairs, primitive pairs,
quartet tables, D and F).  An admitted request gets its own handle -- its
shell pairs stay resident in HBM for all its iterations -- and its own HIP
stream, on which every iteration is enqueued at once: density upload, Fock
build, Fock download.  Requests admitted together therefore run concurrently
on the device; a request that arrives while others run is admitted as soon
as one of them retires.  Latency is measured from arrival to the moment the
last iteration's Fock matrix is in host memory.

The Schwarz integrals of each molecule are computed on the GPU at admission
(liberi_c_schwarz), so a request needs nothing from the host but its basis
and densities.
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import hipmem, liberi
from .caseio import Basis


@dataclass
class Request:
    rid: int
    arrival: float  # seconds after the start of the stream
    basis: Basis
    densities: list  # packed densities, one per iteration
    tau: float = 0.0  # mixed-precision switch (0 = FP64)


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    done: float
    nbf: int
    quartets: int
    device_bytes: int
    iterations: int
    focks: list = field(default_factory=list)

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def estimate_device_bytes(basis: Basis) -> int:
    """Upper estimate of the HBM a handle holds for this basis (bytes)."""
    L = np.asarray(basis.ang_mom) - 1
    k = np.asarray(basis.contr_num, dtype=np.int64)
    nsh = basis.nsh
    npairs = nsh * (nsh + 1) // 2
    nprim = int((np.sum(k) ** 2 + np.sum(k * k)) // 2)
    ntri = basis.num_bas * (basis.num_bas + 1) // 2
    per_l = np.bincount(L, minlength=4).astype(np.int64)
    # pairs of each pair class (la >= lb)
    pcount = {(a, b): (per_l[a] * (per_l[a] + 1) // 2 if a == b else per_l[a] * per_l[b])
              for a in range(4) for b in range(a + 1)}
    # each class keeps its two lists (4 B per pair) and, per kept row, up to a
    # few segments (16 B) and tiles (12 B); a pair class serves as the rows of at
    # most ~10 classes
    tables = sum(10 * 60 * n for n in pcount.values())
    return int(64 * (npairs + nprim) + 8 * 3 * ntri + tables + 4 * basis.num_bas + (1 << 20))


def percentile(x, q: float) -> float:
    """Nearest-rank percentile (q in 0..100)."""
    s = sorted(x)
    if not s:
        return float("nan")
    r = max(1, math.ceil(q / 100.0 * len(s)))
    return float(s[r - 1])


class _Slot:
    def __init__(self, req: Request, t_admit: float, need: int):
        self.req = req
        self.t_admit = t_admit
        self.need = need
        b = req.basis
        sw = liberi.schwarz(b)
        self.h = liberi.Handle()
        self.h.setup(b.with_schwarz(sw))
        if req.tau > 0:
            self.h.set_option(liberi.OPT_FP32_TAU, req.tau)
        self.n = b.ntri
        self.stream = hipmem.Stream()
        self.D = hipmem.DeviceArray(self.n)
        self.F = hipmem.DeviceArray(self.n)
        self.hD = []
        self.hF = []
        for d in req.densities:
            p = hipmem.PinnedArray(self.n)
            p.array[:] = d
            self.hD.append(p)
            self.hF.append(hipmem.PinnedArray(self.n))
        st = self.h.stats()
        self.quartets = int(st["total_quartets"])
        self.device_bytes = int(st["device_bytes"])
        # every iteration on the request's stream: D in, build, F out
        for pd, pf in zip(self.hD, self.hF):
            self.D.copy_from_async(pd, self.stream)
            self.h.fock_build_device(self.D.address, self.F.address, self.n, self.stream.address)
            self.F.copy_to_async(pf, self.stream)
        self.event = hipmem.Event()
        self.event.record(self.stream)

    def finish(self, keep: bool) -> list:
        focks = [pf.array.copy() for pf in self.hF] if keep else []
        self.device_bytes = max(self.device_bytes, int(self.h.stats()["device_bytes"]))
        self.h.destroy()
        self.event.destroy()
        self.stream.destroy()
        self.D.free()
        self.F.free()
        for p in self.hD + self.hF:
            p.free()
        return focks


class Server:
    """Continuous-batching request loop on the current GPU.

    max_active -- requests on the device at once (slots);
    hbm_budget -- bytes of HBM the admitted requests may hold together
                  (default: half of the free HBM when the server starts).
    """

    def __init__(self, max_active: int = 4, hbm_budget: int | None = None, keep_focks: bool = True):
        self.max_active = max_active
        if hbm_budget is None:
            free, _ = hipmem.device_memory()
            hbm_budget = free // 2
        self.hbm_budget = hbm_budget
        self.keep_focks = keep_focks
        self.peak_active = 0  # most requests on the device at once in the last run
        self.peak_bytes = 0   # most estimated HBM held at once in the last run

    def run(self, requests: list) -> list:
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue = collections.deque()
        active = []
        used = 0
        results = []
        self.peak_active = self.peak_bytes = 0
        t0 = time.perf_counter()
        while pending or queue or active:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            progressed = False
            # retire finished requests first: their memory and slots go to the queue
            for s in list(active):
                if s.event.done():
                    t_done = time.perf_counter() - t0
                    focks = s.finish(self.keep_focks)
                    results.append(Result(s.req.rid, s.req.arrival, s.t_admit, t_done, s.req.basis.num_bas,
                                          s.quartets, s.device_bytes, len(s.req.densities), focks))
                    active.remove(s)
                    used -= s.need
                    progressed = True
            # admit while a slot and the memory allow (a lone request is always admitted)
            while queue and len(active) < self.max_active:
                need = estimate_device_bytes(queue[0].basis)
                if active and used + need > self.hbm_budget:
                    break
                req = queue.popleft()
                active.append(_Slot(req, time.perf_counter() - t0, need))
                used += need
                self.peak_active = max(self.peak_active, len(active))
                self.peak_bytes = max(self.peak_bytes, used)
                progressed = True
            if not progressed:
                wait = 0.0005
                if not active and not queue and pending:
                    wait = max(0.0, pending[0].arrival - (time.perf_counter() - t0))
                time.sleep(min(wait, 0.05))
        results.sort(key=lambda r: r.rid)
        return results


def summarize(results: list, wall: float | None = None) -> dict:
    lat = [r.latency for r in results]
    builds = sum(r.iterations for r in results)
    t_first = min(r.arrival for r in results)
    t_last = max(r.done for r in results)
    span = wall if wall is not None else (t_last - t_first)
    return {"requests": len(results), "p50_latency_s": percentile(lat, 50), "p99_latency_s": percentile(lat, 99),
            "mean_latency_s": float(np.mean(lat)), "span_s": span, "builds": builds,
            "builds_per_s": builds / span, "quartets_per_s": sum(r.quartets * r.iterations for r in results) / span}


def scf_like_densities(nbf: int, nocc: int, iterations: int, seed: int, eps0: float = 0.05) -> list:
    """Densities of SCF-like iterations: D_k = 2 Q_k Q_k^T (packed), Q_k the
    orthonormalised C + eps_k G_k with C a seeded orthonormal n x nocc block and
    eps_k = eps0 / (k + 1) (the perturbation shrinks as an SCF converges)."""
    rng = np.random.default_rng(seed)
    c, _ = np.linalg.qr(rng.standard_normal((nbf, nocc)))
    il = np.tril_indices(nbf)
    out = []
    for k in range(iterations):
        q, _ = np.linalg.qr(c + eps0 / (k + 1) * rng.standard_normal(c.shape))
        out.append(np.ascontiguousarray((2.0 * q @ q.T)[il]))
    return out


def request_stream(n: int, rate: float, seed: int, iterations: int = 10, tau: float = 0.0,
                   atoms: tuple = (10, 200)) -> list:
    """Poisson arrivals (rate per second) of random H/C/N/O clusters of
    U{atoms} atoms (gms_liberi_amd.configs.cluster), each with its SCF-like
    densities (N_electrons/2 occupied)."""
    from .configs import cluster

    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(n):
        t += rng.exponential(1.0 / rate) if i else 0.0
        natoms = int(rng.integers(atoms[0], atoms[1] + 1))
        case = cluster(natoms, seed * 1000 + i)
        nocc = int(round(float(np.sum(case.density[[j * (j + 3) // 2 for j in range(case.basis.num_bas)]])) / 2))
        dens = scf_like_densities(case.basis.num_bas, nocc, iterations, seed * 1000 + i)
        reqs.append(Request(i, t, case.basis, dens, tau))
    return reqs
