#!/usr/bin/env python3
"""Benchmark of the asqtad Dslash-and-solver hot path on one MI355X.

Workload (asqtad 32^4): seeded random SU(3) links -> staggered phases -> fat7
(5/8, -1/16, 1/64, -1/384) + Naik (-1/24), all generated on the GPU; one
unit-variance Gaussian source on even sites; QLL mass 0.02 (m = 0.01 of
D^+D + 4m^2).

A step is one even-odd asqtad Dslash call over the 524,288 even sites (FP64,
the reference's double-precision path).  Printed as one JSON line:
  value       = FP64 EO Dslash GFLOP/s (1146 flop per output site)
  roofline    = the FP64 Dslash kernel vs the 8 TB/s HBM peak (2016 B/site
                with the long links' third rows rebuilt, 2400 read in full)
  extra       = FP32 / FP16 Dslash rates and the FP64-CG vs mixed-precision
                CG solves/s at FP64 true relative residual 1e-10
  cpu_baseline= the oracle's OpenMP Dslash (the dslash2T loop nest) on this
                host, on the same lattice with random links
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOPS_ASQTAD = 1146          # per EO output site
FLOPS_ASQTAD_QLL = 1152      # QLL's own count (72 per term), reported beside it
FLOPS_WILSON, FLOPS_WILSON_QLL = 1320, 1344
ORACLE_CFLAGS = "gcc -O3 -march=native -fopenmp -std=gnu11 (oracle.native_lib, built on this host)"
# threads of the CPU baseline: the CPUs this process may run on, at most 16
# (the host cores one GPU of a shared node is given)
CPU_THREADS_CAP = 16
# algorithmic bytes per EO output site: 16 link matrices + one input and one
# output colour vector per rhs (16-bit vectors carry a 4-byte site norm)
CPLX_BYTES = {0: 16, 1: 8, 2: 4, 3: 4}
LINK_BYTES = {p: 16 * 9 * c for p, c in CPLX_BYTES.items()}
# links whose long terms are scaled unitary matrices: the Dslash reads two of
# their three rows (8 terms of 9 complex numbers + 8 of 6)
LINK_BYTES_REBUILT = {p: (8 * 9 + 8 * 6) * c for p, c in CPLX_BYTES.items()}
VEC_BYTES = {0: 48, 1: 24, 2: 16, 3: 16}


def link_bytes(L, prec):
    return LINK_BYTES_REBUILT[prec] if L.reconstructs else LINK_BYTES[prec]
PREC_NAME = {0: "f64", 1: "f32", 2: "f16", 3: "q16"}
HBM_PEAK_GBS = 8000.0
COPY_BYTES = 1 << 28          # the PMC calibration copy: reads and writes 256 MiB
DIMS = (32, 32, 32, 32)
QLL_MASS = 0.02
RSQ = 1e-20


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-cg", action="store_true", help="skip the solver comparison")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-pmc", action="store_true", help="skip the rocprofv3 HBM-traffic passes")
    ap.add_argument("--no-graph", action="store_true", help="config4: eager launches instead of hipGraph replays")
    ap.add_argument("--no-config4", action="store_true",
                    help="skip the 48-solve heterogeneous batch against looping FP64 CG (~1 min)")
    ap.add_argument("--config4", action="store_true", help=argparse.SUPPRESS)   # default on; kept for old scripts
    ap.add_argument("--no-wilson", action="store_true", help="skip the 16-rhs Wilson solves")
    ap.add_argument("--no-cgms", action="store_true", help="skip the 12-shift multishift solves")
    ap.add_argument("--no-qll-api", action="store_true", help="skip the QLL C API (host array) timings")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--kernel-child", type=str, default=None,
                    help="KIND,PREC,NRHS[,MODE]: 20 Dslash launches of one kind (stag, stagfb or wilson) "
                         "for a profiler, after 3 copies of 256 MiB (the counters' scale); MODE is the "
                         "tuning mode (1 tuned, >= 2 forces launch shape MODE-2)")
    ap.add_argument("--pmc-scale-child", action="store_true",
                    help="kernels of known read bytes and load widths (a 16-byte-per-thread copy, then one "
                         "even-field norm each in FP64, FP32 and Q16), for checking the FETCH_SIZE scale")
    ap.add_argument("--fat7-child", action="store_true",
                    help="3 fat7 + Naik smearings of a random SU(3) field, after 3 copies of 256 MiB (the "
                         "counters' scale), for a profiler")
    ap.add_argument("--solver-child", type=str, default=None,
                    help="fp64 | mixed_q16 | mixed_q16_fb: one 32^4 solve of the bench's source, for a profiler")
    ap.add_argument("--block", type=int, default=8, help="sources in the batched mixed-precision solve")
    ap.add_argument("--dims", type=str, default=None, help="override lattice, e.g. 16,16,16,16")
    ap.add_argument("--precs", type=str, default="0,1,2,3",
                    help="Dslash precisions to time (0 f64, 1 f32, 2 f16, 3 16-bit fixed point)")
    return ap.parse_args()


def pmc_child(dims):
    """Run under rocprofv3 --pmc: a copy kernel of known size, the FP64
    Dslash, then the 8-rhs Q16 block Dslash, a few dispatches each."""
    import qll_amd as Q
    Q.init()
    Q.set_tuning(0)
    Q.stream_copy(COPY_BYTES, 3)
    lat = Q.Lattice(dims)
    L = Q.asqtad_links_device(lat, seed=1)
    yin = Q.Field(lat, Q.F64, 3, 1, Q.ODD).gaussian(5, 100)
    out = Q.Field(lat, Q.F64, 3, 1, Q.EVEN)
    Q.time_dslash(L, out, yin, 5)
    LQ = L.converted(Q.Q16)
    yb = Q.Field(lat, Q.Q16, 3, 8, Q.ODD).gaussian(5, 100)
    ob = Q.Field(lat, Q.Q16, 3, 8, Q.EVEN)
    Q.time_dslash(LQ, ob, yb, 5)
    Q.sync()


def kernel_child(dims, spec):
    """20 launches of one Dslash for a profiler run: asqtad ("stag", 16 terms,
    fat7/Naik links; "stagfb", the 8 forward links) or Wilson, precision
    PREC, NRHS rhs; the launch shape is fixed beforehand (MODE), so every
    profiled dispatch is the same kernel.  A copy kernel of known size runs
    first (the scale of the FETCH_SIZE / WRITE_SIZE counters)."""
    import qll_amd as Q
    kind, prec, nrhs = spec.split(",")[:3]
    mode = int(spec.split(",")[3]) if spec.count(",") >= 3 else 1
    prec, nrhs = int(prec), int(nrhs)
    Q.init()
    Q.stream_copy(COPY_BYTES, 3)
    lat = Q.Lattice(dims)
    if kind in ("stag", "stagfb"):
        L = Q.asqtad_links_device(lat, seed=1, kind=Q.STAG_FB if kind == "stagfb" else Q.STAG)
        nc = 3
    else:
        g = Q.Gauge(lat).random_su3(3)
        L = Q.Links(lat, Q.WILSON, Q.F64, 8).from_gauge(g)
        nc = 12
    if prec:
        L = L.converted(prec)
    yin = Q.Field(lat, prec, nc, nrhs, Q.ODD).gaussian(5, 100)
    out = Q.Field(lat, prec, nc, nrhs, Q.EVEN)
    Q.set_tuning(mode)
    Q.time_dslash(L, out, yin, 3)   # code loading; with mode 1 the tuner's trial launches
    ms = Q.time_dslash(L, out, yin, 20)
    Q.sync()
    print("%s prec %d nrhs %d mode %d: %.4f ms" % (kind, prec, nrhs, mode, ms), flush=True)


def pmc_scale_child(dims):
    """Kernels whose read bytes are known: the copy kernel (16-byte loads,
    COPY_BYTES read), then k_norm_dot over one even colour-vector field in
    FP64 (16-byte loads, 48 B/site), FP32 (8-byte, 24 B/site) and Q16 (4-byte
    components + a 4-byte norm, 16 B/site), in that dispatch order."""
    import qll_amd as Q
    Q.init()
    Q.stream_copy(COPY_BYTES, 1)
    lat = Q.Lattice(dims)
    vh = lat.V // 2
    for prec, per_site in ((Q.F64, 48), (Q.F32, 24), (Q.Q16, 16)):
        f = Q.Field(lat, prec, 3, 1, Q.EVEN).gaussian(1, 1)
        f.norm2()
        print("norm2 prec %d: %d bytes read" % (prec, per_site * vh), flush=True)


def fat7_child(dims):
    """Three fat7 + Naik smearings (the bench's coefficients) for a profiler
    run; a copy kernel of known size runs first."""
    import qll_amd as Q
    Q.init()
    Q.stream_copy(COPY_BYTES, 3)
    lat = Q.Lattice(dims)
    u = Q.Gauge(lat).random_su3(1).stag_phases()
    for _ in range(3):
        Q.sync()
        t0 = time.time()
        fl, ll = Q.fat7(u)
        Q.sync()
        print("fat7 + long links: %.3f ms" % ((time.time() - t0) * 1e3), flush=True)
        del fl, ll


def solver_child(dims, kind):
    """One solve of the bench's source and mass (as in the solver lines), for
    a profiler run: FP64 CG, or mixed CG on Q16 doublestored / forward-only
    sloppy links."""
    import qll_amd as Q
    Q.init()
    lat = Q.Lattice(dims)
    L64 = Q.asqtad_links_device(lat, seed=1)
    b = Q.Field(lat, Q.F64, 3, 1, Q.EVEN).gaussian(11, 100)
    x = Q.Field(lat, Q.F64, 3, 1, Q.EVEN)
    Ls = None
    if kind == "mixed_q16":
        Ls = L64.converted(Q.Q16)
    elif kind == "mixed_q16_fb":
        Ls = Q.asqtad_links_device(lat, seed=1, prec=Q.Q16, kind=Q.STAG_FB)
    Q.sync()
    t0 = time.time()
    info = Q.stag_cg(L64, x, b, QLL_MASS, RSQ, 50000) if Ls is None else \
        Q.stag_cg_mixed(L64, Ls, x, b, QLL_MASS, RSQ, 50000, 0.1)
    Q.sync()
    print("%s: %.4f s, %d iterations" % (kind, time.time() - t0, info["iterations"]), flush=True)


def _pmc_pass(counter, dims, tmp):
    """One rocprofv3 pass of one counter over pmc_child; returns the mean KiB
    per dispatch of the copy kernel and of the FP64 Dslash kernel."""
    import csv
    import subprocess
    d = os.path.join(tmp, counter.lower())
    cmd = ["timeout", "-s", "KILL", "90", "rocprofv3", "--pmc", counter, "--output-format", "csv",
           "-d", d, "-o", "pmc", "--", sys.executable, os.path.abspath(__file__), "--pmc-child",
           "--dims", ",".join(map(str, dims))]
    env = dict(os.environ, TMPDIR=tmp)
    r = subprocess.run(cmd, cwd=tmp, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=120)
    if r.returncode != 0:
        raise RuntimeError("rocprofv3 --pmc %s exited %d" % (counter, r.returncode))
    vals = {"copy": [], "dslash": []}
    for root, _, files in os.walk(d):
        for f in files:
            if not f.endswith("counter_collection.csv"):
                continue
            for row in csv.DictReader(open(os.path.join(root, f))):
                name = row["Kernel_Name"]
                if row["Counter_Name"] != counter:
                    continue
                if "k_stream_copy" in name:
                    vals["copy"].append(float(row["Counter_Value"]))
                elif "k_stag_dslash<0," in name.replace(" ", ""):
                    vals["dslash"].append(float(row["Counter_Value"]))
    if not vals["copy"] or not vals["dslash"]:
        raise RuntimeError("rocprofv3 --pmc %s: kernels not found in the counter output" % counter)
    return {k: sum(v) / len(v) for k, v in vals.items()}


def pmc_traffic(dims):
    """HBM bytes per FP64 Dslash launch from FETCH_SIZE and WRITE_SIZE, each
    scaled by bytes/KiB-count read off the copy kernel (which reads and
    writes COPY_BYTES)."""
    import tempfile
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        f = _pmc_pass("FETCH_SIZE", dims, tmp)
        w = _pmc_pass("WRITE_SIZE", dims, tmp)
    ff = COPY_BYTES / (f["copy"] * 1024.0)
    fw = COPY_BYTES / (w["copy"] * 1024.0)
    return {"bytes": f["dslash"] * 1024.0 * ff + w["dslash"] * 1024.0 * fw,
            "fetch_kib": f["dslash"], "write_kib": w["dslash"],
            "fetch_scale": ff, "write_scale": fw}


C4_VOLUMES = [(12, 12, 12, 12), (16, 16, 16, 16), (16, 16, 16, 32), (20, 20, 20, 20), (24, 24, 24, 24),
              (24, 24, 24, 48)]


def config4(Q, graph=True):
    """48 independent asqtad solves, 8 on each of six volumes, each with its
    own seeded links, 1-8 Gaussian rhs and m log-uniform in [0.005, 0.1]
    (QLL mass 2m): the fused batch (Q16 Dslash, FP64 reliable updates,
    hipGraph-replayed and eager) against looping plain FP64 CG over every rhs
    on the same GPU.  A solve is one rhs reaching FP64 |r|/|b| <= 1e-10."""
    import numpy as np
    rng = np.random.default_rng(2024)
    probs, meta = [], []
    t0 = time.time()
    seed = 500
    for dims in C4_VOLUMES:
        lat = Q.Lattice(dims)
        for _ in range(8):
            seed += 1
            nrhs = int(rng.integers(1, 9))
            m = float(np.exp(rng.uniform(np.log(0.005), np.log(0.1))))
            L64 = Q.asqtad_links_device(lat, seed)
            b = Q.Field(lat, Q.F64, 3, nrhs, Q.EVEN).gaussian(seed, 100)
            probs.append((L64, L64.converted(Q.Q16), Q.Field(lat, Q.F64, 3, nrhs, Q.EVEN), b, 2 * m))
            meta.append((lat, seed, nrhs))
    Q.sync()
    setup = time.time() - t0
    nsolve = sum(p[3].nrhs for p in probs)
    out = {"problems": len(probs), "solves": nsolve, "setup_s": round(setup, 2)}
    modes = (("batch_graph", True),) if graph else (("batch_eager", False),)
    for name, use_graph in modes:
        Q.sync()
        t0 = time.time()
        infos = Q.stag_cg_mixed_batch(probs, RSQ, 50000, use_graph=use_graph)
        Q.sync()
        dt = time.time() - t0
        out[name] = {"seconds": round(dt, 3), "solves_per_s": round(nsolve / dt, 2),
                     "all_converged": all(i["converged"] for i in infos),
                     "max_true_rel_resid": max(i["true_rsq"] for i in infos) ** 0.5}
        print("config4 %s: %s" % (name, out[name]), file=sys.stderr, flush=True)
    loop = 0.0
    for (L64, _, _, _, mass), (lat, sd, nrhs) in zip(probs, meta):
        for r in range(nrhs):
            b1 = Q.Field(lat, Q.F64, 3, 1, Q.EVEN).gaussian(sd, 100 + r)   # rhs r of the block
            x1 = Q.Field(lat, Q.F64, 3, 1, Q.EVEN)
            Q.sync()
            t0 = time.time()
            Q.stag_cg(L64, x1, b1, mass, RSQ, 50000)
            Q.sync()
            loop += time.time() - t0
    out["fp64_cg_loop"] = {"seconds": round(loop, 3), "solves_per_s": round(nsolve / loop, 2)}
    best = min(out[m]["seconds"] for m, _ in modes)
    out["speedup_vs_fp64_loop"] = round(loop / best, 2)
    print("config4 fp64 loop: %s" % out["fp64_cg_loop"], file=sys.stderr, flush=True)
    return out


def cpu_baseline(dims, nsec_target=10.0):
    """Oracle OpenMP restatement of QLL's dslash2T loop (256-site blocks, 16
    terms), timed on the host cores, on random SU(3) links of the same shape."""
    import tempfile
    import oracle as O
    affinity = len(os.sched_getaffinity(0))
    nth = min(affinity, CPU_THREADS_CAP)
    os.environ["OMP_NUM_THREADS"] = str(nth)
    tmp = tempfile.mkdtemp(dir="/tmp")
    NL = O.native_lib(tmp)
    t0 = time.time()
    u = O.random_su3_field(dims, 2024)
    links = O.stag_doublestore(dims, u, u)     # 16 link terms per site
    del u
    y = O.gauss_vector(dims, 7, 100, parity=1)
    gen_s = time.time() - t0
    t1, nthreads = O.time_dslash(dims, links, y, 10, NL)      # warm-up + size the sample
    nrep = max(1, min(5000, int(nsec_target / max(t1, 1e-6))))
    t, nthreads = O.time_dslash(dims, links, y, nrep, NL)
    vh = O.volume(dims) // 2
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    import shutil
    shutil.rmtree(tmp, ignore_errors=True)
    return {
        "value": round(FLOPS_ASQTAD * vh / t / 1e9, 3),
        "unit": "GFLOP/s",
        "cores": nthreads,
        "kind": "port",
        "sample": "%d EO Dslash calls at %s (%.2f s each; links generated in %.1f s)" % (
            nrep, "x".join(map(str, dims)), t, gen_s),
        "value_qll_flops": round(FLOPS_ASQTAD_QLL * vh / t / 1e9, 3),
        "cpu_model": model,
        "host_cpus": os.cpu_count(),
        "affinity_cpus": affinity,
        "build": ORACLE_CFLAGS,
    }


def qll_api_times(dims, nrep=20):
    """dslash2D ("all" and "even") and solve2D through the QLL C API on host
    arrays in QLL layout, run by qll_amd/lib/qll_api_bench.bin, with the
    host <-> device copy rates that bound them: pcie_ms is the time the
    call's host bytes need at the pinned DMA rates (half up, half down)."""
    import subprocess
    exe = os.path.join(ROOT, "qll_amd", "lib", "qll_api_bench.bin")
    if not os.path.exists(exe):
        return {"error": "qll_api_bench.bin is not built"}
    r = subprocess.run([exe] + [str(d) for d in dims] + [str(nrep), str(QLL_MASS)], capture_output=True,
                       text=True, timeout=600)
    if r.returncode != 0:
        return {"error": "qll_api_bench exited %d: %s" % (r.returncode, r.stdout[-300:] + r.stderr[-300:])}
    out = json.loads(r.stdout.strip().splitlines()[-1])
    c = out["copy_gbs"]
    for k in ("all", "even"):
        nb = out["dslash2D_%s_host_bytes" % k]
        pcie = (nb / 2) / (c["pinned_h2d"] * 1e9) + (nb / 2) / (c["pinned_d2h"] * 1e9)
        out["dslash2D_%s_pcie_ms" % k] = round(pcie * 1e3, 4)
        out["dslash2D_%s_vs_pcie" % k] = round(out["dslash2D_%s_ms" % k] / (pcie * 1e3), 3)
    print("qll api: %s" % out, file=sys.stderr, flush=True)
    return out


def cgms_solves(Q, dims=(32, 32, 32, 64)):
    """Asqtad multishift CG: 12 shifts log-spaced in [1e-4, 1] on the base
    m^2 = 0.005^2 (QLL mass 0.01; the smallest shift folded into the base
    mass, QLL's shifts[0] = 0), one Gaussian source, every shift to FP64 true
    |r|/|b| <= 1e-10: plain FP64 CGMS against the FP32 multishift with FP64
    per-shift refinement."""
    import numpy as np
    sig = np.logspace(-4, 0, 12)
    m0 = float(np.sqrt(0.01 ** 2 + sig[0]))
    shifts = [float(v) for v in sig - sig[0]]
    lat = Q.Lattice(dims)
    L64 = Q.asqtad_links_device(lat, 2)
    L32 = L64.converted(Q.F32)
    b = Q.Field(lat, Q.F64, 3, 1, Q.EVEN).gaussian(12, 100)
    xs = [Q.Field(lat, Q.F64, 3, 1, Q.EVEN) for _ in shifts]
    Ap = Q.Field(lat, Q.F64, 3, 1, Q.EVEN)
    t = Q.Field(lat, Q.F64, 3, 1, Q.ODD)
    bsq = b.norm2()

    def worst_rel_resid():
        w = 0.0
        for x, sh in zip(xs, shifts):
            Q.stag_normal(L64, Ap, x, t, float(np.sqrt(m0 ** 2 + sh)))
            Ap.axpby(1.0, b, -1.0)
            w = max(w, Ap.norm2() / bsq)
        return w ** 0.5

    res = {"lattice": list(dims), "shifts": len(shifts), "mass_qll": 0.01}
    Q.sync()
    t0 = time.time()
    info = Q.stag_cgms(L64, xs, b, m0, shifts, RSQ, 50000)
    Q.sync()
    dt = time.time() - t0
    res["fp64_cgms"] = {"seconds": round(dt, 4), "iterations": info["iterations"],
                        "max_true_rel_resid": worst_rel_resid()}
    print("cgms fp64: %s" % res["fp64_cgms"], file=sys.stderr, flush=True)
    # FP32 multishift, then per-shift refinement on FP32 links, or on Q16
    # links for the shifts heavy enough for them (FP32 below that)
    for name, Ls in (("mixed_fp32_cgms", L32), ("mixed_fp32_q16_cgms", L64.converted(Q.Q16))):
        for x in xs:
            x.zero()
        Q.sync()
        t0 = time.time()
        infos = Q.stag_cgms_mixed(L64, Ls, xs, b, m0, shifts, RSQ, 50000)
        Q.sync()
        dt = time.time() - t0
        res[name] = {"seconds": round(dt, 4), "iterations": max(i["iterations"] for i in infos),
                     "converged": all(i["converged"] for i in infos),
                     "max_true_rel_resid": worst_rel_resid(),
                     "speedup_vs_fp64": round(res["fp64_cgms"]["seconds"] / dt, 3)}
        print("cgms %s: %s" % (name, res[name]), file=sys.stderr, flush=True)
    return res


def wilson_solves(Q, dims=(32, 32, 32, 64), nrhs=16, kappa=0.124):
    """Wilson even-odd CGNE on 1 - kappa^2 H_eo H_oe for a block of 16
    Gaussian sources sharing every link load, to FP64 true |r|/|b| <= 1e-10
    per rhs: plain FP64 against FP64 defect correction around an FP32 and a
    half-precision (FP16 storage, per-site norm) inner solve."""
    lat = Q.Lattice(dims)
    g = Q.Gauge(lat).random_su3(3)
    L64 = Q.Links(lat, Q.WILSON, Q.F64, 8).from_gauge(g)
    del g
    b = Q.Field(lat, Q.F64, 12, nrhs, Q.EVEN).gaussian(13, 100)
    x = Q.Field(lat, Q.F64, 12, nrhs, Q.EVEN)
    res = {"lattice": list(dims), "rhs": nrhs, "kappa": kappa}
    for name, ls in (("fp64", None), ("mixed_fp32", Q.F32), ("mixed_fp16", Q.F16)):
        Ls = None if ls is None else L64.converted(ls)
        Q.wilson_cgne(L64, x, b, kappa, RSQ, 4, ls=Ls)   # untimed: launch-shape tuning
        Q.sync()
        t0 = time.time()
        infos = Q.wilson_cgne(L64, x, b, kappa, RSQ, 20000, ls=Ls)
        Q.sync()
        dt = time.time() - t0
        res[name] = {"seconds": round(dt, 4), "iterations": max(i["iterations"] for i in infos),
                     "converged": all(i["converged"] for i in infos),
                     "max_true_rel_resid": max(i["true_rsq"] for i in infos) ** 0.5,
                     "solves_per_s": round(nrhs / dt, 3)}
        if name != "fp64":
            res[name]["speedup_vs_fp64"] = round(res["fp64"]["seconds"] / dt, 3)
        print("wilson %s: %s" % (name, res[name]), file=sys.stderr, flush=True)
        del Ls
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (every box here has one GPU); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    dims = tuple(int(x) for x in args.dims.split(",")) if args.dims else DIMS
    if args.pmc_child:
        pmc_child(dims)
        return
    if args.kernel_child:
        kernel_child(dims, args.kernel_child)
        return
    if args.pmc_scale_child:
        pmc_scale_child(dims)
        return
    if args.fat7_child:
        fat7_child(dims)
        return
    if args.solver_child:
        solver_child(dims, args.solver_child)
        return
    traffic = None
    if not args.no_pmc:
        try:
            traffic = pmc_traffic(dims)
            print("pmc traffic: %s" % traffic, file=sys.stderr, flush=True)
        except Exception as e:   # no profiler: traffic stays null
            print("pmc traffic unavailable: %s" % e, file=sys.stderr, flush=True)
    import qll_amd as Q
    Q.init()
    lat = Q.Lattice(dims)
    vh = lat.V // 2
    t0 = time.time()
    L64 = Q.asqtad_links_device(lat, seed=1)
    L32 = L64.converted(Q.F32)
    L16 = L64.converted(Q.F16)
    LQ = L64.converted(Q.Q16)
    Q.sync()
    setup_s = time.time() - t0
    print("setup (links: SU(3)+phases+fat7+Naik+doublestore, 3 precisions): %.2f s" % setup_s,
          file=sys.stderr, flush=True)

    # measured streaming-copy bandwidth (read + write) over 2 x 1 GiB
    copy_gbs, _ = Q.stream_copy(1 << 30, 20)
    print("stream copy: %.1f GB/s" % copy_gbs, file=sys.stderr, flush=True)

    rates = {}
    prof = {}
    precs = [int(p) for p in args.precs.split(",")]
    if 0 not in precs:
        precs.insert(0, 0)
    bytes_site = {}
    for prec, L in ((0, L64), (1, L32), (2, L16), (3, LQ)):
        if prec not in precs:
            continue
        yin = Q.Field(lat, prec, 3, 1, Q.ODD).gaussian(5, 100)
        out = Q.Field(lat, prec, 3, 1, Q.EVEN)
        Q.time_dslash(L, out, yin, args.warmup)
        Q.sync()
        ms = Q.time_dslash(L, out, yin, args.steps)
        Q.sync()
        rates[prec] = ms
        bytes_site[prec] = link_bytes(L, prec) + 2 * VEC_BYTES[prec]
        gbs = bytes_site[prec] * vh / (ms * 1e-3) / 1e9
        prof[prec] = {"ms_per_call": round(ms, 5), "bytes_per_site": bytes_site[prec],
                      "GFLOP/s": round(FLOPS_ASQTAD * vh / (ms * 1e-3) / 1e9, 1),
                      "GFLOP/s_qll_count": round(FLOPS_ASQTAD_QLL * vh / (ms * 1e-3) / 1e9, 1),
                      "GB/s": round(gbs, 1),
                      "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
        print("dslash prec=%d: %s" % (prec, prof[prec]), file=sys.stderr, flush=True)

    # forward-only links (createStaggeredLinksFb): half the link storage; the
    # backward terms read the other parity's forward links at the neighbour,
    # so the algorithmic bytes per output site equal the doublestored set's
    # (16 link matrices, long ones as two rows), reused through L2 / MALL
    for prec, L in ((0, L64), (1, L32), (3, LQ)):
        if prec not in precs:
            continue
        Lf = Q.asqtad_links_device(lat, seed=1, prec=prec, kind=Q.STAG_FB)
        yin = Q.Field(lat, prec, 3, 1, Q.ODD).gaussian(5, 100)
        out = Q.Field(lat, prec, 3, 1, Q.EVEN)
        Q.time_dslash(Lf, out, yin, args.warmup)
        ms = Q.time_dslash(Lf, out, yin, args.steps)
        bs = link_bytes(Lf, prec) + 2 * VEC_BYTES[prec]
        gbs = bs * vh / (ms * 1e-3) / 1e9
        key = "%s_fb" % PREC_NAME[prec]
        prof[key] = {"ms_per_call": round(ms, 5), "bytes_per_site": bs,
                     "link_bytes_stored": Lf.nbytes, "link_bytes_stored_doublestored": L.nbytes,
                     "GFLOP/s": round(FLOPS_ASQTAD * vh / (ms * 1e-3) / 1e9, 1),
                     "GB/s": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                     "vs_doublestored": round(rates[prec] / ms, 3) if prec in rates else None}
        print("dslash %s: %s" % (key, prof[key]), file=sys.stderr, flush=True)
        del Lf, yin, out

    # the FP64 Dslash called on host arrays, as a caller of the QLL API's
    # dslash2 uses it: upload the input (lexicographic FP64 host array), one
    # launch, download the output.  PCIe-inclusive; not the bench value.
    if 0 in precs:
        yh = Q.Field(lat, Q.F64, 3, 1, Q.ODD).gaussian(5, 100).get_lex()
        yin = Q.Field(lat, Q.F64, 3, 1, Q.ODD)
        out = Q.Field(lat, Q.F64, 3, 1, Q.EVEN)
        nh = 10
        yin.set_lex(yh)
        Q.dslash(L64, out, yin)
        out.get_lex()
        Q.sync()
        t0 = time.perf_counter()
        for _ in range(nh):
            yin.set_lex(yh)
            Q.dslash(L64, out, yin)
            out.get_lex()
        Q.sync()
        ms = (time.perf_counter() - t0) * 1e3 / nh
        prof["f64_host_arrays"] = {"ms_per_call": round(ms, 3), "host_bytes_each_way": int(yh.nbytes),
                                   "GFLOP/s": round(FLOPS_ASQTAD * vh / (ms * 1e-3) / 1e9, 1)}
        print("dslash f64_host_arrays: %s" % prof["f64_host_arrays"], file=sys.stderr, flush=True)

    # the QLL C API on host arrays in QLL layout (dslash2D, solve2D), timed by
    # the C driver built beside the library; PCIe-inclusive, not the value
    qll_api = None if args.no_qll_api else qll_api_times(dims)

    # block Dslash: links loaded once for the whole block of rhs
    nbk = args.block
    for prec, L in ((1, L32), (3, LQ)):
        if prec not in precs:
            continue
        yin = Q.Field(lat, prec, 3, nbk, Q.ODD).gaussian(5, 100)
        out = Q.Field(lat, prec, 3, nbk, Q.EVEN)
        Q.time_dslash(L, out, yin, args.warmup)
        ms = Q.time_dslash(L, out, yin, max(1, args.steps // 4))
        byts = (link_bytes(L, prec) + nbk * 2 * VEC_BYTES[prec]) * vh
        key = "%s_block%d" % (PREC_NAME[prec], nbk)
        prof[key] = {"ms_per_call": round(ms, 5), "ms_per_rhs": round(ms / nbk, 5),
                     "GFLOP/s": round(nbk * FLOPS_ASQTAD * vh / (ms * 1e-3) / 1e9, 1),
                     "GB/s": round(byts / (ms * 1e-3) / 1e9, 1),
                     "hbm_frac": round(byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        print("dslash %s: %s" % (key, prof[key]), file=sys.stderr, flush=True)
    # the block CG's D_oe p with 16-bit links: FP32 input, output in the
    # links' format (bytes: links + FP32 in + 16-bit out)
    if 3 in precs:
        yin = Q.Field(lat, Q.F32, 3, nbk, Q.ODD).gaussian(5, 100)
        out = Q.Field(lat, Q.Q16, 3, nbk, Q.EVEN)
        Q.time_dslash(LQ, out, yin, args.warmup)
        ms = Q.time_dslash(LQ, out, yin, max(1, args.steps // 4))
        bps = link_bytes(LQ, 3) + nbk * (VEC_BYTES[1] + VEC_BYTES[3])
        key = "q16_f32in_block%d" % nbk
        prof[key] = {"ms_per_call": round(ms, 5), "ms_per_rhs": round(ms / nbk, 5), "bytes_per_site": bps,
                     "GFLOP/s": round(nbk * FLOPS_ASQTAD * vh / (ms * 1e-3) / 1e9, 1),
                     "GB/s": round(bps * vh / (ms * 1e-3) / 1e9, 1),
                     "hbm_frac": round(bps * vh / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        print("dslash %s: %s" % (key, prof[key]), file=sys.stderr, flush=True)
        del yin, out

    # Wilson EO Dslash: FP64 one rhs, and 16 rhs sharing every link load
    # (algorithmic bytes per site: 8 links + nrhs x (input + output spinor);
    # with one or two rhs SU(3) links are read as two rows of each term and
    # the third is rebuilt)
    WLINK = {p: 8 * 9 * c for p, c in CPLX_BYTES.items()}
    WLINK_REBUILT = {p: 8 * 6 * c for p, c in CPLX_BYTES.items()}
    WSPIN = {0: 192, 1: 96, 2: 52, 3: 52}
    gw = Q.Gauge(lat).random_su3(3)
    W64 = Q.Links(lat, Q.WILSON, Q.F64, 8).from_gauge(gw)
    del gw
    for prec, nrw in ((0, 1), (1, 16), (3, 16)):
        Lw = W64 if prec == 0 else W64.converted(prec)
        yin = Q.Field(lat, prec, 12, nrw, Q.ODD).gaussian(7, 100)
        out = Q.Field(lat, prec, 12, nrw, Q.EVEN)
        Q.time_dslash(Lw, out, yin, 10)
        ms = Q.time_dslash(Lw, out, yin, max(1, args.steps // 4))
        wl = WLINK_REBUILT[prec] if Lw.reconstructs and nrw <= 2 else WLINK[prec]   # the rule in wlaunch
        byts = (wl + nrw * 2 * WSPIN[prec]) * vh
        key = "wilson_%s_%drhs" % (PREC_NAME[prec], nrw)
        prof[key] = {"ms_per_call": round(ms, 5), "ms_per_rhs": round(ms / nrw, 5),
                     "bytes_per_site": wl + nrw * 2 * WSPIN[prec],
                     "GFLOP/s": round(nrw * FLOPS_WILSON * vh / (ms * 1e-3) / 1e9, 1),
                     "GFLOP/s_qll_count": round(nrw * FLOPS_WILSON_QLL * vh / (ms * 1e-3) / 1e9, 1),
                     "GB/s": round(byts / (ms * 1e-3) / 1e9, 1),
                     "hbm_frac": round(byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        print("dslash %s: %s" % (key, prof[key]), file=sys.stderr, flush=True)
        del Lw, yin, out
    del W64

    ms64 = rates[0]
    gflops = FLOPS_ASQTAD * vh / (ms64 * 1e-3) / 1e9
    achieved = bytes_site[0] * vh / (ms64 * 1e-3) / 1e9

    solver = None
    if not args.no_cg:
        b = Q.Field(lat, Q.F64, 3, 1, Q.EVEN).gaussian(11, 100)
        x = Q.Field(lat, Q.F64, 3, 1, Q.EVEN)
        Ap = Q.Field(lat, Q.F64, 3, 1, Q.EVEN)
        t = Q.Field(lat, Q.F64, 3, 1, Q.ODD)
        bsq = b.norm2()

        def true_rsq():
            Q.stag_normal(L64, Ap, x, t, QLL_MASS)
            Ap.axpby(1.0, b, -1.0)
            return Ap.norm2() / bsq

        solver = {}
        Q.sync()
        t0 = time.time()
        info = Q.stag_cg(L64, x, b, QLL_MASS, RSQ, 50000)
        Q.sync()
        dt = time.time() - t0
        solver["fp64_cg"] = {"seconds": round(dt, 4), "iterations": info["iterations"],
                             "true_rel_resid": true_rsq() ** 0.5, "solves_per_s": round(1 / dt, 4)}
        print("fp64 cg: %s" % solver["fp64_cg"], file=sys.stderr, flush=True)
        # the sloppy operator on forward-only Q16 links (half the storage;
        # FP64 reliable updates still on the doublestored FP64 links)
        LQfb = Q.asqtad_links_device(lat, seed=1, prec=Q.Q16, kind=Q.STAG_FB)
        for name, Ls in (("mixed_fp32", L32), ("mixed_q16", LQ), ("mixed_fp16", L16), ("mixed_q16_fb", LQfb)):
            x.zero()
            Q.sync()
            t0 = time.time()
            info = Q.stag_cg_mixed(L64, Ls, x, b, QLL_MASS, RSQ, 50000, 0.1)
            Q.sync()
            dt = time.time() - t0
            solver[name] = {"seconds": round(dt, 4), "iterations": info["iterations"],
                            "reliable_updates": info["reliable_updates"],
                            "restarts": info["restarts"], "escalated_to_fp32": info["escalations"],
                            "true_rel_resid": true_rsq() ** 0.5, "solves_per_s": round(1 / dt, 4),
                            "speedup_vs_fp64": round(solver["fp64_cg"]["seconds"] / dt, 3)}
            if info["escalations"]:
                # the 16-bit links' operator error was not small against m^2:
                # the inner solve ran on FP32 links, so this is not a 16-bit result
                solver[name]["inner_links"] = "fp32 (16-bit links too coarse for this mass)"
            print("%s: %s" % (name, solver[name]), file=sys.stderr, flush=True)
        del LQfb

        # batched: a block of sources solved together (one block-Dslash pair per
        # iteration for the whole block), solves/s = block size / wall time
        nb = args.block
        bB = Q.Field(lat, Q.F64, 3, nb, Q.EVEN).gaussian(21, 100)
        xB = Q.Field(lat, Q.F64, 3, nb, Q.EVEN)
        ApB = Q.Field(lat, Q.F64, 3, nb, Q.EVEN)
        tB = Q.Field(lat, Q.F64, 3, nb, Q.ODD)
        bBsq = bB.norm2() if nb > 1 else [bB.norm2()]

        def true_rsq_block():
            Q.stag_normal(L64, ApB, xB, tB, QLL_MASS)
            ApB.axpby(1.0, bB, -1.0)
            r2 = ApB.norm2() if nb > 1 else [ApB.norm2()]
            return max(r2[i] / bBsq[i] for i in range(nb))

        # the plain FP64 block CG on the same sources: separates what batching
        # gains from what the lower precision gains
        xB.zero()
        Q.sync()
        t0 = time.time()
        infos = Q.stag_cg_block(L64, xB, bB, QLL_MASS, RSQ, 50000)
        Q.sync()
        dt = time.time() - t0
        key64 = "fp64_block%d" % nb
        solver[key64] = {"seconds": round(dt, 4), "rhs": nb, "iterations": max(i["iterations"] for i in infos),
                         "converged": all(i["converged"] for i in infos),
                         "max_true_rel_resid": true_rsq_block() ** 0.5, "solves_per_s": round(nb / dt, 4),
                         "speedup_vs_fp64": round(nb / dt / solver["fp64_cg"]["solves_per_s"], 3)}
        print("%s: %s" % (key64, solver[key64]), file=sys.stderr, flush=True)
        for name, Ls in (("block%d_mixed_q16" % nb, LQ), ("block%d_mixed_fp32" % nb, L32)):
            Q.stag_cg_mixed_block(L64, Ls, xB, bB, QLL_MASS, RSQ, 16, 0.1)   # untimed: launch-shape tuning
            Q.sync()
            t0 = time.time()
            infos = Q.stag_cg_mixed_block(L64, Ls, xB, bB, QLL_MASS, RSQ, 50000, 0.1)
            Q.sync()
            dt = time.time() - t0
            sps = nb / dt
            solver[name] = {"seconds": round(dt, 4), "rhs": nb, "iterations": infos[0]["iterations"],
                            "reliable_updates": infos[0]["reliable_updates"], "restarts": infos[0]["restarts"],
                            "converged": all(i["converged"] for i in infos),
                            "max_true_rel_resid": true_rsq_block() ** 0.5, "solves_per_s": round(sps, 4),
                            "speedup_vs_fp64": round(sps / solver["fp64_cg"]["solves_per_s"], 3),
                            "speedup_vs_fp64_block": round(sps / solver[key64]["solves_per_s"], 3)}
            print("%s: %s" % (name, solver[name]), file=sys.stderr, flush=True)

    wil = None if args.no_wilson else wilson_solves(Q)
    cgms = None if args.no_cgms else cgms_solves(Q)
    c4 = None if args.no_config4 else config4(Q, graph=not args.no_graph)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(dims)
        print("cpu baseline: %s" % cpu, file=sys.stderr, flush=True)

    line = {
        "metric": "asqtad Dslash GFLOP/s + %HBM roofline; mixed-prec CG solves/s vs FP64 CG @1e-10",
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms64, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SU(3) + fat7/Naik generated on the GPU)",
        "config": {"workload": "asqtad even-odd Dslash, %s lattice (%d output sites), FP64, 1 rhs"
                               % ("x".join(map(str, dims)), vh),
                   "lattice": list(dims), "mass_qll": QLL_MASS, "flops_per_site": FLOPS_ASQTAD,
                   "flops_per_site_qll_count": FLOPS_ASQTAD_QLL},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": round(traffic["bytes"]) if traffic else None,
                     "traffic_per_site": round(traffic["bytes"] / vh, 1) if traffic else None,
                     "algorithmic_per_site": bytes_site[0],
                     "long_link_rows_read": 2 if L64.reconstructs else 3,
                     "measured_copy_gbs": round(copy_gbs, 1),
                     "frac_of_copy": round(achieved / copy_gbs, 4)},
        "cpu_baseline": cpu,
        "dslash": {(PREC_NAME[p] if p in PREC_NAME else p): prof[p] for p in prof},
        "qll_api": qll_api,
        "solvers": solver,
        "wilson_cgne_16rhs": wil,
        "cgms_12_shifts": cgms,
        "config4_batch": c4,
        "device": Q.device_name(),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
