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
the host-array entry point's rate (D upload and F download included), the
mixed-precision sweep over the FP32 switch tau, the time of each of 8 rank
shares, the C-N chain (f classes), ethane and water workloads timed alone, a serving stream, and a CPU
baseline: the reference CPU path (flang build, oracle/_ref) timed on this
host on the whole workload (median of 3 builds), with its thread scaling on a
12-atom piece of the cage.

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
             "ethane": "tests/golden/cfg_ethane.npz", "water": "tests/golden/cfg_water.npz"}
DESCR = {"c60": "C60-like cage, 60 C, s/p/d, 900 basis functions, FP64, Schwarz cutoff 5e-11",
         "cn_chain": "20-atom C/N zig-zag chain, s/p/d/f, 620 basis functions, FP64, Schwarz cutoff 5e-11",
         "ethane": "ethane-like C2H6, s/p/d, 42 basis functions, FP64, Schwarz cutoff 5e-11",
         "water": "water-like O + 2 H, s/p, 13 basis functions, FP64, Schwarz cutoff 5e-11"}


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


def profiled_counter_tflops(workload: str, cls: str):
    """FP64 FLOP/s of the class's kernel from the committed rocprof
    instruction-counter summary (64 lanes per FP64 instruction, FMA = 2), or
    None when it was not profiled."""
    f = ROOT / "profiles" / "fock_build_c60_f64_counters.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        p = [x.strip() for x in line.split("|")]
        if len(p) > 6 and p[1] == workload and p[3] == cls:
            return {"counter_tflops": float(p[5]), "kernel": p[2].strip("`"), "source": f"profiles/{f.name}"}
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


def gpu_leg(case, steps: int, warmup: int, taus=(), graph: bool = True, host_entry: bool = True):
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
    # the host-array entry point (liberi_c_fock_build): D upload and F
    # download over PCIe included, wall clock around each call
    host = None
    if host_entry:
        f_host = np.empty(n)
        h.fock_build(case.density, f_host)
        t0 = time.perf_counter()
        for _ in range(steps):
            h.fock_build(case.density, f_host)
        host_ms = (time.perf_counter() - t0) * 1e3 / steps
        host = dict(ms=host_ms, max_abs_dF_vs_device=float(np.max(np.abs(f_host - fock))))
        log(f"host-array entry point: {host_ms:.3f} ms/build (D upload and F download included)")
    mixed = []
    e64 = 0.5 * float(np.dot(case.density, fock))
    for tau in taus:
        h.set_option(liberi.OPT_FP32_TAU, tau)
        h.fock_build_device(D.address, F.address, n, stream.address)  # (re)tunes the classes for this tau
        ms_mixed = timed_builds(h, D, F, n, stream, steps, hipmem)
        fm = F.to_host()
        st_m = h.stats()
        nf32 = st_m["fp32_quartets"]
        mixed.append(dict(tau=tau, ms=ms_mixed, fp32_quartets=int(nf32),
                          fp32_classes=[census.NAMES[c] for c in range(55) if st_m["class_fp32"][c]],
                          max_abs_dF=float(np.max(np.abs(fm - fock))),
                          dE=0.5 * float(np.dot(case.density, fm)) - e64))
        log(f"mixed tau={tau:g}: {ms_mixed:.3f} ms/build, {nf32} FP32 quartets, dE {mixed[-1]['dE']:.2e}")
    h.set_option(liberi.OPT_FP32_TAU, 0.0)
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
                                   "tflops": round(flops[c] / max(class_ms[c], 1e-9) * 1e-9, 3),
                                   "form": int(st["class_variant"][c]), "tile": int(st["class_tile"][c])}
                 for c in np.argsort(-class_ms) if class_ms[c] > 0}
    return dict(ms_per_step=ms_per_step, fock=fock, stats=st, dom=dom, achieved=achieved,
                dom_ms=float(class_ms[dom]), dom_flops=float(flops[dom]), total_flops=total_flops,
                per_class=per_class, setup_s=setup_s, mixed=mixed, host=host)


def ref_builds(case, threads: int, reps: int):
    """Seconds of `reps` builds of the reference CPU path (oracle/_ref, flang
    build of gms_libERI, OpenMP host fallback) on `threads` threads, after its
    own setup, or None when it is not built."""
    exe = ROOT / "oracle" / "_ref" / "liberi_ref_run"
    if not exe.exists():
        return None
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", MALLOC_PERTURB_="255")
    with tempfile.TemporaryDirectory() as td:
        cf, ff = Path(td) / "c.case", Path(td) / "c.fock"
        case.write_stream(cf)
        r = subprocess.run([str(exe), str(cf), str(ff), str(reps)], env=env, capture_output=True, text=True,
                           timeout=1800)
        if r.returncode != 0:
            log("reference CPU build failed: " + r.stderr[-500:])
            return None
        secs = [float(ln.split()[-1]) for ln in r.stdout.splitlines() if ln.startswith("build")]
        fock = np.fromfile(ff, "<f8")
    return secs, fock


def cpu_leg(case, reps: int):
    """The reference CPU path on the whole workload, all host threads, median
    of `reps` builds; and its thread scaling on the 12-atom piece of the cage
    (tests/golden/frag_c12.npz), one build on 1 thread and on all of them."""
    from gms_liberi_amd.caseio import Case

    threads = len(os.sched_getaffinity(0))
    log(f"cpu baseline: reference CPU path, whole workload, {reps} builds on {threads} threads")
    r = ref_builds(case, threads, reps)
    if r is None:
        log("cpu baseline skipped: oracle/_ref not built")
        return None
    secs, fock = r
    out = dict(seconds=float(np.median(secs)), all_seconds=secs, threads=threads,
               max_abs_dF_vs_fixture=float(np.max(np.abs(fock - case.fock))))
    frag = Case.load_npz(ROOT / "tests" / "golden" / "frag_c12.npz")
    scal = {}
    for t in sorted({1, threads}):
        rr = ref_builds(frag, t, 1)
        if rr:
            scal[str(t)] = rr[0][0]
    out["scaling_frag_c12_s"] = scal
    log(f"cpu baseline: {out['seconds']:.4g} s/build (median of {secs}); c12 piece by threads {scal}")
    return out


def rank_shares(case, nprocs: int, steps: int):
    """Each of `nprocs` rank shares (my_rank/num_procs) built alone on this
    GPU, with the default cost-balanced split and with the reference's equal
    quartet counts: the time of each share, the imbalance max/mean, and the
    sum of the partial Fock matrices against the reference."""
    from gms_liberi_amd import hipmem, liberi

    n = case.basis.ntri
    D = hipmem.DeviceArray.from_host(case.density)
    F = hipmem.DeviceArray(n)
    stream = hipmem.Stream()
    out = {"num_procs": nprocs}
    for label, split in (("balanced", 0), ("equal_counts", 1)):
        ms, total = [], np.zeros(n)
        for r in range(nprocs):
            h = liberi.Handle()
            h.set_option(liberi.OPT_RANK_SPLIT, split)
            h.setup(case.basis, r, nprocs)
            h.fock_build_device(D.address, F.address, n, stream.address)  # autotuning
            h.fock_build_device(D.address, F.address, n, stream.address)
            ms.append(timed_builds(h, D, F, n, stream, steps, hipmem))
            total += F.to_host()
            h.destroy()
        mean = float(np.mean(ms))
        out[label] = dict(ms=[round(m, 3) for m in ms], max_over_mean=round(max(ms) / mean, 4),
                          sum_ms=round(float(np.sum(ms)), 3),
                          max_abs_dF_sum_vs_reference=float(np.max(np.abs(total - case.fock))))
    stream.destroy()
    return out


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


def _mixed_summary(g, tau_main):
    if not g["mixed"]:
        return None
    sweep = [{"tau": m["tau"], "value": round(1000.0 / m["ms"], 4), "speedup_vs_fp64": round(g["ms_per_step"] / m["ms"], 3),
              "fp32_quartets": m["fp32_quartets"], "fp32_classes": m["fp32_classes"],
              "max_abs_dF_vs_fp64": m["max_abs_dF"], "dE_vs_fp64": m["dE"]} for m in g["mixed"]]
    main = next((m for m in sweep if m["tau"] == tau_main), sweep[0])
    return {**main, "unit": "Fock builds/s", "sweep": sweep}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c60", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-reps", type=int, default=3, help="builds of the reference CPU path (median reported)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--taus", default="1e-4,1e-5,1e-6,1e-7",
                    help="FP32 switches of the mixed-precision sweep (comma separated; empty: skip)")
    ap.add_argument("--tau", type=float, default=1e-6, help="FP32 switch of the mixed serving line")
    ap.add_argument("--extra", default="cn_chain,ethane,water", help="further workloads timed alone (comma separated)")
    ap.add_argument("--rank-shares", type=int, default=8, help="rank shares of the workload timed one by one (0: skip)")
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
    taus = [float(t) for t in a.taus.split(",") if t.strip()]
    g = gpu_leg(case, a.steps, a.warmup, taus, graph=not a.no_graph)
    st = g["stats"]
    builds_per_s = 1000.0 / g["ms_per_step"]
    dF = float(np.max(np.abs(g["fock"] - case.fock)))
    dE = float(0.5 * np.dot(case.density, g["fock"] - case.fock))

    extra = {}
    for w in [x for x in a.extra.split(",") if x.strip() and x != a.workload]:
        ce = Case.load_npz(ROOT / WORKLOADS[w])
        log(f"extra workload {w}: nbf={ce.basis.num_bas} nsh={ce.basis.nsh}")
        ge = gpu_leg(ce, a.steps, a.warmup, (), graph=not a.no_graph, host_entry=False)
        extra[w] = {"workload": DESCR[w], "ms_per_build": round(ge["ms_per_step"], 4),
                    "builds_per_s": round(1000.0 / ge["ms_per_step"], 4),
                    "quartets_per_s": round(ge["stats"]["total_quartets"] * 1000.0 / ge["ms_per_step"], 1),
                    "total_tflops": round(ge["total_flops"] / (ge["ms_per_step"] * 1e-3) * 1e-12, 3),
                    "accuracy_vs_reference": {"max_abs_dF": float(np.max(np.abs(ge["fock"] - ce.fock))),
                                              "dE_2e": float(0.5 * np.dot(ce.density, ge["fock"] - ce.fock))},
                    "top_kernels": dict(list(ge["per_class"].items())[:8]),
                    "forms": {k: v["form"] for k, v in ge["per_class"].items()}}
        if not a.no_cpu:
            threads = len(os.sched_getaffinity(0))
            r = ref_builds(ce, threads, 1)
            if r:
                extra[w]["cpu_baseline"] = {"value": round(1.0 / r[0][0], 6), "unit": "Fock builds/s", "cores": threads,
                                            "kind": "reference", "sample": "the whole workload, one build",
                                            "max_abs_dF_vs_fixture": float(np.max(np.abs(r[1] - ce.fock)))}
                log(f"extra workload {w}: reference CPU path {r[0][0]:.2f} s/build on {threads} threads")
    shares = rank_shares(case, a.rank_shares, a.steps) if a.rank_shares > 1 else None
    if shares:
        for k in ("balanced", "equal_counts"):
            log(f"rank shares of {a.rank_shares} ({k}): {shares[k]['ms']} ms, max/mean {shares[k]['max_over_mean']}")
    serve = serving_leg(a.serve_requests, a.serve_rate, 5, a.tau, a.serve_slots) if a.serve_requests > 0 else None
    cpu = None if a.no_cpu else cpu_leg(case, a.cpu_reps)
    cpu_base = None
    if cpu:
        q_total = st["total_quartets"]
        cpu_base = {"value": round(1.0 / cpu["seconds"], 6), "unit": "Fock builds/s",
                    "cores": cpu["threads"], "kind": "reference",
                    "sample": (f"the whole workload: reference CPU path (flang build of gms_libERI, OpenMP host "
                               f"fallback) on {cpu['threads']} threads, median of {len(cpu['all_seconds'])} builds "
                               f"{[float('%.4g' % x) for x in cpu['all_seconds']]} s; max|F - fixture| "
                               f"{cpu['max_abs_dF_vs_fixture']:.1e}"),
                    "quartets_per_s": round(q_total / cpu["seconds"], 1),
                    "thread_scaling_frag_c12_s": cpu["scaling_frag_c12_s"]}

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
                     "kernel_ms": round(g["dom_ms"], 4), "kernel_flops": g["dom_flops"],
                     "counters": profiled_counter_tflops(a.workload, census.NAMES[dom])},
        "top_kernels": dict(list(g["per_class"].items())[:8]),
        "classes": g["per_class"],
        "host_entry": None if not g["host"] else {
            "value": round(1000.0 / g["host"]["ms"], 4), "unit": "Fock builds/s", "ms_per_build": round(g["host"]["ms"], 4),
            "note": "liberi_c_fock_build on host arrays: D upload and F download included"},
        "mixed_precision": _mixed_summary(g, a.tau),
        "rank_shares": shares,
        **extra,
        "serving": serve,
        "cpu_baseline": cpu_base,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
