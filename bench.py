#!/usr/bin/env python3
"""Benchmark: Schwarz-screened RHF Fock build on one MI355X.

One step = one `liberi_fock_build` of the C60-like cage (60 carbons on a
truncated icosahedron of radius 3.55 A, s/p/d shells per carbon, 900 basis
functions, 360 shells; seeded density with 180 occupied orbitals; Schwarz
cutoff 5e-11), FP64, with the shell pairs resident on the device and D/F
already in HBM (liberi_c_fock_build_device).

Prints ONE JSON line: builds/s (value), contracted-quartet and
primitive-quartet rates, the roofline of the dominant class kernel (FP64
VALU), the accuracy of the build against the reference's own Fock matrix,
and a CPU baseline: the reference CPU path (flang build, oracle/_ref) timed
on this host over a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Fock builds/s & ERI quartets/s, FP64 vs mixed @≤1e-6 Eh; serving p50/p99 latency"
FP64_VALU_PEAK = 78.6e12  # MI355X dense FP64 vector, FLOP/s
WORKLOADS = {"c60": "tests/golden/cfg_c60.npz", "cn_chain": "tests/golden/cfg_cn_chain.npz",
             "ethane": "tests/golden/cfg_ethane.npz"}
DESCR = {"c60": "C60-like cage, 60 C, s/p/d, 900 basis functions, FP64, Schwarz cutoff 5e-11",
         "cn_chain": "20-atom C/N zig-zag chain, s/p/d/f, 620 basis functions, FP64, Schwarz cutoff 5e-11",
         "ethane": "ethane-like C2H6, s/p/d, 42 basis functions, FP64, Schwarz cutoff 5e-11"}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def profiled_traffic(workload: str, cls: str, variant: int):
    """HBM bytes per launch of the class's kernel from the committed rocprof
    counter summary (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), or
    None when that kernel was not profiled."""
    f = ROOT / "profiles" / f"fock_build_{workload}_traffic.txt"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        p = line.split()
        if len(p) >= 4 and not line.startswith("#") and p[0] == cls and int(p[1]) == variant:
            return {"fetch_bytes": int(p[2]), "write_bytes": int(p[3]), "bytes": int(p[2]) + int(p[3]),
                    "source": f"profiles/{f.name} ({p[5] if len(p) > 5 else ''})"}
    return None


def timed_builds(h, D, F, n, stream, steps, hipmem):
    start, end = hipmem.Event(), hipmem.Event()
    stream.synchronize()
    start.record(stream)
    for _ in range(steps):
        h.fock_build_device(D.address, F.address, n, stream.address)
    end.record(stream)
    end.synchronize()
    return start.elapsed_ms(end) / steps


def gpu_leg(case, steps: int, warmup: int, tau: float, graph: bool = True):
    from gms_liberi_amd import census, hipmem, liberi

    h = liberi.Handle()
    t0 = time.time()
    h.setup(case.basis)
    setup_s = time.time() - t0
    h.set_option(liberi.OPT_GRAPH, 1 if graph else 0)
    n = case.basis.ntri
    D = hipmem.DeviceArray.from_host(case.density)
    F = hipmem.DeviceArray(n)
    stream = hipmem.Stream()
    # the first build after setup also runs the per-class autotuner; it is part
    # of setting up (like liberi_setup), never of the timed builds
    t0 = time.time()
    h.fock_build_device(D.address, F.address, n, stream.address)
    stream.synchronize()
    tune_s = time.time() - t0
    for _ in range(warmup):
        h.fock_build_device(D.address, F.address, n, stream.address)
    stream.synchronize()
    log(f"setup {setup_s:.3f}s, first build with autotuning {tune_s:.3f}s, warmup done ({warmup} builds)")

    h.set_option(liberi.OPT_CLASS_TIMING, 1)
    class_ms = np.zeros(55)
    start, end = hipmem.Event(), hipmem.Event()
    stream.synchronize()
    start.record(stream)
    for k in range(steps):
        h.fock_build_device(D.address, F.address, n, stream.address)
        class_ms += h.stats()["class_ms"]  # waits for this build's events
    end.record(stream)
    end.synchronize()
    ms_per_step = start.elapsed_ms(end) / steps
    h.set_option(liberi.OPT_CLASS_TIMING, 0)
    st = h.stats()
    fock = F.to_host()
    # plain FP64 without per-class events (the figure reported as value)
    ms_plain = timed_builds(h, D, F, n, stream, steps, hipmem)
    log(f"timed {steps} builds: {ms_plain:.3f} ms/build ({ms_per_step:.3f} with per-class events)")
    mixed = None
    if tau > 0:
        h.set_option(liberi.OPT_FP32_TAU, tau)
        h.fock_build_device(D.address, F.address, n, stream.address)
        ms_mixed = timed_builds(h, D, F, n, stream, steps, hipmem)
        fm = F.to_host()
        nf32 = h.stats()["fp32_quartets"]
        h.set_option(liberi.OPT_FP32_TAU, 0.0)
        e64 = 0.5 * float(np.dot(case.density, fock))
        mixed = dict(tau=tau, ms=ms_mixed, fp32_quartets=int(nf32), max_abs_dF=float(np.max(np.abs(fm - fock))),
                     dE=0.5 * float(np.dot(case.density, fm)) - e64)
        log(f"mixed tau={tau:g}: {ms_mixed:.3f} ms/build, {nf32} FP32 quartets, dE {mixed['dE']:.2e}")
    h.destroy()
    stream.destroy()
    ms_per_step = ms_plain

    class_ms /= steps
    q = st["quartets"].astype(np.float64)
    pq = st["prim_quartets"].astype(np.float64)
    flops = np.array([census.class_flops(c, q[c], pq[c]) for c in range(55)])
    dom = int(np.argmax(class_ms))
    achieved = flops[dom] / (class_ms[dom] * 1e-3)
    total_flops = float(flops.sum())
    per_class = {census.NAMES[c]: {"ms": round(float(class_ms[c]), 4), "quartets": int(q[c]),
                                   "tflops": round(flops[c] / max(class_ms[c], 1e-9) * 1e-9, 3)}
                 for c in np.argsort(-class_ms)[:8] if class_ms[c] > 0}
    return dict(ms_per_step=ms_per_step, fock=fock, stats=st, dom=dom, achieved=achieved,
                dom_ms=float(class_ms[dom]), dom_flops=float(flops[dom]), total_flops=total_flops,
                per_class=per_class, setup_s=setup_s, mixed=mixed)


def sampled_case(case, keep: float, seed: int = 7):
    """Same workload with a seeded fraction of its shell pairs (each class's
    largest pair always kept, so every pre-screen threshold is unchanged);
    the other pairs get Schwarz value 0 and drop out in the pre-screen."""
    import dataclasses

    b = case.basis
    rng = np.random.default_rng(seed)
    sw = b.schwrz.copy()
    mask = rng.random(sw.shape[0]) < keep
    L = b.ang_mom - 1
    i, j = np.tril_indices(b.nsh)
    order = np.lexsort((j, i))
    i, j = i[order], j[order]
    cls = np.minimum(L[i], L[j]) * 4 + np.maximum(L[i], L[j])
    for c in np.unique(cls):
        idx = np.where(cls == c)[0]
        mask[idx[np.argmax(sw[idx])]] = True
    sw[~mask] = 0.0
    return dataclasses.replace(case, basis=b.with_schwarz(sw)), float(mask.mean())


def cpu_leg(case, keep: float):
    """Reference CPU path (oracle/_ref, flang build) on a bounded sample."""
    from gms_liberi_amd import liberi

    exe = ROOT / "oracle" / "_ref" / "liberi_ref_run"
    if not exe.exists():
        log("cpu baseline skipped: oracle/_ref not built")
        return None
    samp, frac = sampled_case(case, keep)
    h = liberi.Handle()  # the product counts the sample's surviving quartets (same rules, bit-exact)
    h.setup(samp.basis)
    q_sample = h.stats()["total_quartets"]
    h.destroy()
    threads = len(os.sched_getaffinity(0))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", MALLOC_PERTURB_="255")
    with tempfile.TemporaryDirectory() as td:
        cf, ff = Path(td) / "s.case", Path(td) / "s.fock"
        samp.write_stream(cf)
        log(f"cpu baseline: reference CPU path on {frac:.0%} of the shell pairs, {q_sample} quartets, "
            f"{threads} threads")
        r = subprocess.run([str(exe), str(cf), str(ff), "1"], env=env, capture_output=True, text=True, timeout=900)
        if r.returncode != 0:
            log("cpu baseline failed: " + r.stderr[-500:])
            return None
        secs = [float(ln.split()[-1]) for ln in r.stdout.splitlines() if ln.startswith("build")]
    t = min(secs)
    return dict(seconds=t, quartets=int(q_sample), threads=threads, pair_fraction=frac)


def serving_leg(n: int, rate: float, seed: int, tau: float, slots: int):
    """Request stream (Poisson arrivals of random H/C/N/O clusters, 10-200
    atoms, 10 SCF-like iterations each), served in FP64 and in the mixed form;
    per-request energy error of the mixed form against FP64."""
    from gms_liberi_amd import serving

    t0 = time.time()
    reqs = serving.request_stream(n, rate, seed, iterations=10)
    log(f"serving: {n} requests generated in {time.time() - t0:.1f}s "
        f"(nbf {min(r.basis.num_bas for r in reqs)}..{max(r.basis.num_bas for r in reqs)})")
    out = {}
    focks = {}
    for label, t in (("fp64", 0.0), ("mixed", tau)):
        for r in reqs:
            r.tau = t
        srv = serving.Server(max_active=slots)
        res = srv.run(reqs)
        s = serving.summarize(res)
        s["peak_active"] = srv.peak_active
        out[label] = s
        focks[label] = res
        log(f"serving {label}: p50 {s['p50_latency_s']:.2f}s p99 {s['p99_latency_s']:.2f}s, "
            f"{s['builds_per_s']:.2f} builds/s, peak {srv.peak_active} requests on the GPU")
    dE = []
    for r, a, b in zip(reqs, focks["fp64"], focks["mixed"]):
        dE.append(max(abs(0.5 * float(np.dot(d, fb - fa))) for d, fa, fb in zip(r.densities, a.focks, b.focks)))
    out["mixed"]["tau"] = tau
    out["mixed"]["max_abs_dE_vs_fp64"] = max(dE)
    out["workload"] = (f"{n} requests, Poisson arrivals at {rate}/s, random H/C/N/O clusters of 10-200 atoms "
                       f"(s/p/d, nbf {min(r.basis.num_bas for r in reqs)}-{max(r.basis.num_bas for r in reqs)}), "
                       f"10 SCF-like iterations each, up to {slots} requests on the GPU at once")
    return out


def copy_calibration(mib: int) -> None:
    """One device-to-device copy of `mib` MiB: the reference point that turns
    rocprof's FETCH_SIZE / WRITE_SIZE counters into HBM bytes."""
    import ctypes

    from gms_liberi_amd import hipmem

    n = mib * 1024 * 1024 // 8
    src, dst = hipmem.DeviceArray(n), hipmem.DeviceArray(n)
    rt = hipmem.runtime()
    for _ in range(2):
        rc = rt.hipMemcpy(dst.ptr, src.ptr, ctypes.c_size_t(n * 8), 3)  # hipMemcpyDeviceToDevice
        if rc != 0:
            raise RuntimeError("hipMemcpy D2D failed")
    log(f"copied {mib} MiB device to device twice")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c60", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-pairs", type=float, default=0.3, help="shell-pair fraction of the CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tau", type=float, default=1e-6, help="FP32 switch of the mixed-precision line (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=12, help="requests of the serving stream (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=1.0, help="arrivals per second of the serving stream")
    ap.add_argument("--serve-slots", type=int, default=4, help="requests on the GPU at once")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the class kernels eagerly on one stream (profiling runs: kernel durations "
                         "then match the per-class events)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (rocprof counter calibration)")
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py: this build measures one GPU only (asked for {a.gpus}); "
              "the Fock build shards by quartet ranges with one all-reduce of F, which is not run here.",
              file=sys.stderr)
        sys.exit(2)

    from gms_liberi_amd.caseio import Case

    case = Case.load_npz(ROOT / WORKLOADS[a.workload])
    log(f"workload {a.workload}: nbf={case.basis.num_bas} nsh={case.basis.nsh}")
    g = gpu_leg(case, a.steps, a.warmup, a.tau, graph=not a.no_graph)
    st = g["stats"]
    builds_per_s = 1000.0 / g["ms_per_step"]
    dF = float(np.max(np.abs(g["fock"] - case.fock)))
    dE = float(0.5 * np.dot(case.density, g["fock"] - case.fock))

    serve = serving_leg(a.serve_requests, a.serve_rate, 5, a.tau, a.serve_slots) if a.serve_requests > 0 else None
    cpu = None if a.no_cpu else cpu_leg(case, a.cpu_pairs)
    cpu_base = None
    if cpu:
        q_total = st["total_quartets"]
        cpu_base = {"value": round(cpu["quartets"] / cpu["seconds"] / q_total, 6), "unit": "Fock builds/s",
                    "cores": cpu["threads"], "kind": "reference",
                    "sample": (f"reference CPU path (flang build of gms_libERI, OpenMP host fallback) on the same "
                               f"case restricted to a seeded {cpu['pair_fraction']:.0%} of its shell pairs: "
                               f"{cpu['quartets']} of {q_total} surviving quartets in {cpu['seconds']:.2f} s; "
                               f"builds/s = sample quartets/s / full quartet count"),
                    "quartets_per_s": round(cpu["quartets"] / cpu["seconds"], 1)}

    from gms_liberi_amd import census

    dom = g["dom"]
    traffic = profiled_traffic(a.workload, census.NAMES[dom], int(st["class_variant"][dom]))
    out = {
        "metric": METRIC,
        "value": round(builds_per_s, 4),
        "unit": "Fock builds/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(g["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded basis, geometry and density; committed fixture tests/golden)",
        "config": {"workload": DESCR[a.workload], "basis_functions": case.basis.num_bas, "shells": case.basis.nsh,
                   "schwarz_cutoff": 5e-11, "precision": "fp64"},
        "quartets_per_s": round(st["total_quartets"] * builds_per_s, 1),
        "prim_quartets_per_s": round(float(st["prim_quartets"].sum()) * builds_per_s, 1),
        "quartets_per_build": int(st["total_quartets"]),
        "total_tflops": round(g["total_flops"] / (g["ms_per_step"] * 1e-3) * 1e-12, 3),
        "accuracy_vs_reference": {"max_abs_dF": dF, "dE_2e": dE},
        "roofline": {"bound": "valu", "kernel": census.NAMES[dom], "achieved": round(g["achieved"] * 1e-12, 3),
                     "peak": FP64_VALU_PEAK * 1e-12, "unit": "TFLOP/s",
                     "frac": round(g["achieved"] / FP64_VALU_PEAK, 4),
                     "traffic": traffic["bytes"] if traffic else None,
                     "traffic_detail": traffic,
                     "kernel_ms": round(g["dom_ms"], 4), "kernel_flops": g["dom_flops"]},
        "top_kernels": g["per_class"],
        "mixed_precision": None if not g["mixed"] else {
            "tau": g["mixed"]["tau"], "value": round(1000.0 / g["mixed"]["ms"], 4), "unit": "Fock builds/s",
            "speedup_vs_fp64": round(g["ms_per_step"] / g["mixed"]["ms"], 3),
            "fp32_quartets": g["mixed"]["fp32_quartets"], "max_abs_dF_vs_fp64": g["mixed"]["max_abs_dF"],
            "dE_vs_fp64": g["mixed"]["dE"]},
        "serving": serve,
        "cpu_baseline": cpu_base,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
