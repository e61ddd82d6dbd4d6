"""Benchmark: PDDSparse Monte Carlo assembly on one MI355X.

Workload (default): plain PDDSparse at GPU scale -- divergence-form variable-coefficient
elliptic problem on [-1,1]^2 (sigma = diag(1+0.3 sin x, 1+0.3 cos y), b = (0.2y, -0.2x),
c = 1 + x^2, u = e^x cos y), 8x8 subdomains x 32 knots per interface (3437 knots, 3409 of
them Monte Carlo knots), 100,000 paths per knot, h = 1e-4, FP64 paths, sigma/b/c served
from 512x512 bicubic B-spline LUTs, Philox seed 7.  One step = one full assembly of G (CSR)
and B over every knot.  Synthetic input: the manufactured problem itself.

Prints one JSON line (rank 0).  `value` is whole-job phase-1 path-steps per second over the
K timed assemblies (inputs resident on the device).  Extra fields report the plain FP64 GPU
form, the FP32-path engine and the RMS interface error of each.
"""
import argparse
import glob
import hashlib
import json
import os
import sys
import time

# the CPU baseline's OpenMP threads: one per core, packed close together (read by the OpenMP
# runtime when the oracle library is first loaded, so set before anything loads it)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
T_START = time.perf_counter()


def progress(msg):
    """One line on stderr as each part of the run starts or ends (stdout carries only the JSON)."""
    print("[bench %7.1f s] %s" % (time.perf_counter() - T_START, msg), file=sys.stderr, flush=True)


def kernel_source_hash():
    """SHA-256 (16 hex digits) of the sources that decide the assembly kernels' memory traffic --
    the kernels, their launch and tiling code, the stencil geometry -- and the build flags: a
    traffic figure measured on another build of these is not reported for this one.  The serving
    loop, the interface solver and the disc mesh do not take part in the measured assembly."""
    h = hashlib.sha256()
    skip = {"server.cpp", "solver.hip", "solver.h", "disc_mesh.cpp"}
    files = sorted(p for p in glob.glob(os.path.join(ROOT, "pdds_amd", "csrc", "*")) if os.path.basename(p) not in skip)
    files += [os.path.join(ROOT, "pdds_amd", "Makefile")]
    for p in files:
        h.update(os.path.basename(p).encode())
        with open(p, "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def cpu_info():
    """Model, sockets, physical cores and the logical CPUs this process may run on."""
    model, phys, sockets = None, set(), set()
    cur = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in list(f) + [""]:
                if not line.strip():
                    if cur:
                        phys.add((cur.get("physical id", "0"), cur.get("core id", cur.get("processor"))))
                        sockets.add(cur.get("physical id", "0"))
                    cur = {}
                    continue
                k, _, v = line.partition(":")
                cur[k.strip()] = v.strip()
                if k.strip() == "model name" and model is None:
                    model = v.strip()
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return {"model": model, "sockets": len(sockets), "physical_cores": len(phys), "logical_cpus": os.cpu_count(),
            "cpus_available": affinity, "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "omp_places": os.environ.get("OMP_PLACES")}

# Algorithmic FP64 operation count per path-step of the LUT configuration (+,-,x = 1,
# fma = 2): LUT 264 (coordinates 6, weights 18, 5 fields x 40 + row sums), grad u 2, xi 7,
# Z 3, Y 3, X 8, Gobet 5, Box-Muller 10.  Transcendentals and Philox integer work are
# counted separately and are not in this figure.
FLOPS_PER_STEP = {"cfg2_varcoef_lut": 302.0, "cfg1_poisson": 41.0, "fixture": 41.0}
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3


def cpu_baseline(mesh, prob_name, n_paths_ref, h, seed, budget_s=15.0):
    """The FP64 C/OpenMP oracle (a restatement of the reference CPU path) on a bounded sample of
    the same workload: chunks of paths on seeded random interior knots until about budget_s
    seconds of stepping have elapsed."""
    from oracle import oracle as O
    from pdds_amd.workloads import varcoef_fields
    cpu = cpu_info()
    # one thread per physical core available to this process (OMP_NUM_THREADS if set)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(cpu["physical_cores"] or 1, cpu["cpus_available"] or 1)
    if prob_name == "varcoef_lut":
        n = 512
        xs = np.linspace(-1, 1, n)
        X, Y = np.meshgrid(xs, xs)
        coef = O.lut_prefilter(varcoef_fields(X, Y))
        d = 2.0 / (n - 1)
        lut = {"coef": coef, "NX": n + 2, "NY": n + 2, "nf": 5, "ox": -1 - d, "oy": -1 - d, "sx": 1 / d, "sy": 1 / d}
        pb = O.problem(O.PDE_VARCOEF_LUT, lut=lut)
    else:
        pb = O.problem({"monegros": O.PDE_MONEGROS, "sinsin": O.PDE_SINSIN}[prob_name])
    rng = np.random.default_rng(1234)
    interior = np.flatnonzero(~mesh.on_boundary)
    order = rng.permutation(interior)
    run = O.run_params(h, seed, threads=threads)
    # time-bounded sample: chunks of paths on seeded random knots until budget_s has elapsed
    chunk = 65536
    steps = 0
    used = []
    t = time.perf_counter()
    while time.perf_counter() - t < budget_s:
        k = int(order[len(used) % len(order)])
        st = mesh.stencil(k)
        paths = O.simulate(pb, run, mesh.x0[k], mesh.y0[k], st["rect"], k, 0, chunk)
        steps += int(paths["steps"].sum())
        used.append(k)
    t_sim = time.perf_counter() - t
    # the per-core rate: one thread on the first sampled knot
    run1 = O.run_params(h, seed, threads=1)
    k = int(order[0])
    st = mesh.stencil(k)
    t1 = time.perf_counter()
    p1 = O.simulate(pb, run1, mesh.x0[k], mesh.y0[k], st["rect"], k, 0, 4096)
    rate1 = int(p1["steps"].sum()) / (time.perf_counter() - t1)
    # the reduction of one sampled knot's paths, timed apart from the stepping (one thread): the
    # engine-order batched form on all of them, the reference's per-exit form (m RBF values and an
    # m x m product with Psi^-T per exit) on 2048
    paths = O.simulate(pb, run, mesh.x0[k], mesh.y0[k], st["rect"], k, 0, chunk)
    cs = st["cshape"]
    t2 = time.perf_counter()
    O.reduce_batched(pb, paths, st, mesh.glob, cs, st["ipsi"], 32, 6272)
    t_red = (time.perf_counter() - t2) / chunk
    sub = {q: v[:2048] for q, v in paths.items()}
    t3 = time.perf_counter()
    O.reduce_faithful(pb, sub, st, mesh.glob, cs, st["ipsi"])
    t_faith = (time.perf_counter() - t3) / 2048
    return {"value": steps / t_sim, "unit": "path-steps/s", "cores": threads, "kind": "port",
            "sample": "%d knots x %d paths (seeded random interior knots of %d), %d path-steps, phase-1 "
                      "stepping, FP64 C/OpenMP oracle, %.1f s" % (len(used), chunk, len(interior), steps, t_sim),
            "one_thread_path_steps_per_s": rate1,
            "reduction_s_per_path_one_thread": {"knot": k, "stencil_entries": int(np.sum(st["msz"])),
                                                "batched_engine_order": t_red, "per_exit_reference_form": t_faith},
            "cpu": cpu}


def copy_calibration(n=1 << 30, reps=3):
    """A plain copy of known size: what FETCH_SIZE / WRITE_SIZE read for it gives the factors that
    turn those counters into HBM bytes for the path kernel."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(n)) or hip.hipMalloc(C.byref(b), C.c_size_t(n)):
        raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(a, 1, C.c_size_t(n)):
        raise RuntimeError("hipMemset failed")
    for _ in range(reps):
        if hip.hipMemcpy(b, a, C.c_size_t(n), 3):
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration": {"bytes_read_per_copy": n, "bytes_written_per_copy": n, "copies": reps}}))


TRAFFIC_RUNS = {
    "fetch": "rocprofv3 --pmc FETCH_SIZE --output-format csv -d gpurun_out/pmc -o fetch -- "
             "python3 bench.py --steps 1 --warmup 0 --no-extras --no-cpu-baseline",
    "write": "rocprofv3 --pmc WRITE_SIZE --output-format csv -d gpurun_out/pmc -o write -- "
             "python3 bench.py --steps 1 --warmup 0 --no-extras --no-cpu-baseline",
    "cfetch": "rocprofv3 --pmc FETCH_SIZE --output-format csv -d gpurun_out/pmc -o cfetch -- "
              "python3 bench.py --copy-calibration",
    "cwrite": "rocprofv3 --pmc WRITE_SIZE --output-format csv -d gpurun_out/pmc -o cwrite -- "
              "python3 bench.py --copy-calibration",
}


def _counter_rows(pmc_dir, name, counter):
    """Per dispatch (in dispatch order): (kernel name, summed counter value) from the rocprofv3
    counter-collection CSV of run `name`."""
    import csv
    files = sorted(glob.glob(os.path.join(pmc_dir, "**", "%s*counter_collection.csv" % name), recursive=True))
    if not files:
        raise FileNotFoundError("no counter_collection.csv for run %r under %s" % (name, pmc_dir))
    per = {}
    for fn in files:
        with open(fn) as f:
            for r in csv.DictReader(f):
                if r.get("Counter_Name") != counter:
                    continue
                key = (fn, int(r["Dispatch_Id"]))
                k, v = per.get(key, (r["Kernel_Name"], 0.0))
                per[key] = (k, v + float(r["Counter_Value"]))
    return [per[k] for k in sorted(per)]


def summarize_traffic(pmc_dir, out_path):
    """HBM bytes per launch of the path and reduction kernels from the FETCH_SIZE / WRITE_SIZE runs
    of TRAFFIC_RUNS, scaled by what the same counters read for a device copy of known size, stamped
    with the hash of the sources they were measured on."""
    copy_bytes = 1 << 30

    def per_copy(name, counter):
        rows = [v for k, v in _counter_rows(pmc_dir, name, counter) if v > 0]
        return float(np.median(rows)) if rows else None

    cf, cw = per_copy("cfetch", "FETCH_SIZE"), per_copy("cwrite", "WRITE_SIZE")
    ff = copy_bytes / (cf * 1024.0)
    wf = copy_bytes / (cw * 1024.0)
    out = {"workload": "cfg2 (8x8 x 32 knots, 1e5 paths/knot, h=1e-4, 512^2 LUTs), FP64 paths, one assembly",
           "source_hash": kernel_source_hash(), "fetch_factor": ff, "write_factor": wf,
           "calibration": "1 GiB hipMemcpy device-to-device (bench.py --copy-calibration): FETCH_SIZE %.1f KiB, "
                          "WRITE_SIZE %.1f KiB per copy" % (cf, cw),
           "commands": list(TRAFFIC_RUNS.values()) + ["python3 bench.py --summarize-traffic gpurun_out/pmc"]}
    fetch = _counter_rows(pmc_dir, "fetch", "FETCH_SIZE")
    write = _counter_rows(pmc_dir, "write", "WRITE_SIZE")
    for label, pat in (("path_kernel", "path_kernel"), ("reduce_kernel", "reduce_kernel")):
        fk = [v for k, v in fetch if pat in k]
        wk = [v for k, v in write if pat in k]
        if not fk or not wk:
            continue
        rec = {"kernel": [k for k, v in fetch if pat in k][0][:120], "launches": len(fk),
               "fetch_size_kib": float(np.mean(fk)), "write_size_kib": float(np.mean(wk)),
               "fetch_bytes": float(np.mean(fk)) * 1024.0 * ff, "write_bytes": float(np.mean(wk)) * 1024.0 * wf}
        rec["bytes_per_launch"] = rec["fetch_bytes"] + rec["write_bytes"]
        out[label] = rec
    if "path_kernel" in out:
        out["bytes_per_launch"] = out["path_kernel"]["bytes_per_launch"]
    out["algorithmic_bytes_per_launch"] = 32 * 3409 * 100000
    out["algorithmic_bytes_note"] = ("exit records: a 32-byte FP64 record (x, y, Y, score) per path, 3.409e8 paths "
                                     "(steps and step-cap stops are summed per knot, 16 bytes per knot)")
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


SOLVE_PATHS = 3000      # paths per knot of a served interface solve
SOLVE_H_COEF = 1.2e-4   # a served solve steps with h = SOLVE_H_COEF * (subdomain width)^2


def request_stream(duration_s, rate=200.0, seed=23, solve_paths=SOLVE_PATHS, h_coef=SOLVE_H_COEF):
    """The serving workload: Poisson arrivals (mean `rate` per second); 90 % point evaluations at a
    uniform interior point with log-uniform 1e4..1e6 paths, 10 % interface solves of 4..16
    subdomains per side and 16..64 knots per interface with `solve_paths` paths per knot; every
    request with c = -k^2, k uniform in [0, 4].  A solve's step is h_coef times its squared
    subdomain width: the step bias of a solve's interface values grows as its subdomains shrink
    (measured on 8x8x32 and 16x16x64 meshes, about linear in h), so an RMS target of 1e-3 needs
    h in proportion to that width squared, and the paths per knot then take about the same number
    of steps on every mesh."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration_s:
            return out
        k2 = float(rng.uniform(0.0, 4.0)) ** 2
        if rng.random() < 0.9:
            x, y = rng.uniform(-0.999, 0.999, 2)
            out.append({"kind": "point", "t": t, "x": float(x), "y": float(y), "n": int(10 ** rng.uniform(4, 6)),
                        "k2": k2})
        else:
            n_sub = int(rng.integers(4, 17))
            out.append({"kind": "solve", "t": t, "n_sub": n_sub, "n_node": int(rng.integers(16, 65)),
                        "n": solve_paths, "k2": k2, "h": h_coef * (2.0 / n_sub) ** 2})


def _pct(v, q):
    return float(np.percentile(v, q)) if len(v) else None


def serving_bench(eng, duration_s=1.5, h=5e-5, seed=23, plain=True):
    """Serve the request stream with the continuous-batching server (FP32 paths, FP64 sums and
    solves) and, for comparison, one request at a time with the plain FP64 GPU kernel.  Latency =
    arrival to completion, including the host mesh build of a solve.  h = 5e-5 keeps the weak
    order-1 bias of a point estimate (about 10 h) under the 1e-3 error target; at h = 1e-3 the
    points carry an error of about 8e-3 whatever their path counts."""
    from pdds_amd.engine import Mesh, Server, bvp, default_run
    from pdds_amd.workloads import relative_fac
    dom = (-1.0, -1.0, 1.0, 1.0)
    reqs = request_stream(duration_s, seed=seed)
    fac = {}
    for r in reqs:
        if r["kind"] == "solve":
            fac[(r["n_sub"], r["n_node"])] = relative_fac(dom[:2], dom[2:], r["n_sub"], r["n_node"], 0.8)
    # two preparation threads: a solve's mesh and Psi^-1 tables are built on the host, each build
    # spread over the host cores by itself, so two workers keep the cores busy and preparation
    # close to arrival order (16 workers ran 16 builds at once, finished them all late and left
    # the device idle for a quarter of the stream: 117 against 167 G path-steps/s with 2)
    prep = 2
    # corner entries shared by two sides of a stencil are summed (quirk_last_side = 0): the
    # reference's "=+" keeps only the last side's weight, a bias of 1.5-2.5e-3 in the interface
    # solution of the larger meshes whatever the path count or h; serving has no reference
    # counterpart, and the plain baseline below uses the same rule
    srv = Server(eng, bvp("helmholtz"), default_run(h=h, seed=seed, precision="fp32", quirk_last_side=0), dom,
                 prep_threads=prep)
    t0 = time.perf_counter()
    tickets = []
    for i, r in enumerate(reqs):
        d = t0 + r["t"] - time.perf_counter()
        if d > 0:
            time.sleep(d)
        if r["kind"] == "point":
            tickets.append(srv.submit_point(r["x"], r["y"], r["n"], k2=r["k2"], stream=i + 1))
        else:
            tickets.append(srv.submit_solve(r["n_sub"], r["n_node"], fac[(r["n_sub"], r["n_node"])], r["n"],
                                            k2=r["k2"], stream=i + 1, h=r["h"]))
    progress("serving: %d requests submitted" % len(reqs))
    res = [srv.wait(t, cap=1 << 20) for t in tickets]
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    progress("serving: all requests done in %.2f s" % wall)
    lat = {"point": [x["latency_ms"] for x, r in zip(res, reqs) if r["kind"] == "point"],
           "solve": [x["latency_ms"] for x, r in zip(res, reqs) if r["kind"] == "solve"]}
    u_exact = bvp("helmholtz").u
    perr = [abs(x["u"] - u_exact(r["x"], r["y"])) for x, r in zip(res, reqs) if r["kind"] == "point"]
    meshes = {key: Mesh(dom[:2], dom[2:], key[0], key[1], f) for key, f in fac.items()}
    serr = []
    for x, r in zip(res, reqs):
        if r["kind"] == "solve" and x["status"] == 0:
            m = meshes[(r["n_sub"], r["n_node"])]
            serr.append(float(np.sqrt(np.mean((x["u_knots"] - u_exact(m.px, m.py)) ** 2))))
    out = {"workload": "Poisson arrivals 200/s for %.1f s (seed %d): %d point evaluations (1e4..1e6 paths, h=%g), "
                       "%d interface solves (4..16 subdomains, 16..64 knots/interface, %d paths/knot, h=%g x subdomain "
                       "width^2, corner entries summed), c = -k^2 with k ~ U[0,4], FP32 paths, FP64 sums and solves"
                       % (duration_s, seed, len(lat["point"]), h, len(lat["solve"]), SOLVE_PATHS, SOLVE_H_COEF),
           "requests": len(reqs), "failed": int(sum(x["status"] != 0 for x in res)), "prep_threads": prep,
           "latency_ms": {"p50": _pct(lat["point"] + lat["solve"], 50), "p99": _pct(lat["point"] + lat["solve"], 99),
                          "point_p50": _pct(lat["point"], 50), "point_p99": _pct(lat["point"], 99),
                          "solve_p50": _pct(lat["solve"], 50), "solve_p99": _pct(lat["solve"], 99)},
           "path_steps_per_s": st["path_steps"] / wall, "path_steps": st["path_steps"], "wall_s": wall,
           "rounds": st["rounds"], "device_busy_s": st["busy_ms"] / 1e3,
           "point_abs_error_rms": float(np.sqrt(np.mean(np.square(perr)))) if perr else None,
           "solve_rms_interface_error": {"mean": float(np.mean(serr)), "max": float(np.max(serr))} if serr else None,
           "rms_target": 1e-3, "solves_over_target": int(sum(e > 1e-3 for e in serr)),
           "points_over_target": int(sum(e > 1e-3 for e in perr))}
    if plain:
        # the same stream, one request at a time in arrival order, plain FP64 launches
        lat_p = {"point": [], "solve": []}
        steps = 0
        t0 = time.perf_counter()
        for i, r in enumerate(reqs):
            d = t0 + r["t"] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            run = default_run(h=r.get("h", h), seed=seed, precision="fp64", plain=1, stream=i + 1, quirk_last_side=0)
            prob = bvp("helmholtz", k2=r["k2"])
            if r["kind"] == "point":
                o = eng.point_eval(prob, run, dom, [r["x"]], [r["y"]], r["n"])
            else:
                m = meshes[(r["n_sub"], r["n_node"])]
                o = eng.assemble(m, prob, run, r["n"])
                eng.solve(m.n_knots, m.row_ptr, m.cols, o["G"], o["B"], tol=1e-10)
            steps += o["counters"]["path_steps"]
            lat_p[r["kind"]].append(1e3 * (time.perf_counter() - (t0 + r["t"])))
            if (i + 1) % 50 == 0:
                progress("serving, plain FP64 one at a time: %d of %d requests" % (i + 1, len(reqs)))
        wall_p = time.perf_counter() - t0
        out["plain_gpu"] = {"latency_ms": {"p50": _pct(lat_p["point"] + lat_p["solve"], 50),
                                           "p99": _pct(lat_p["point"] + lat_p["solve"], 99),
                                           "point_p50": _pct(lat_p["point"], 50), "point_p99": _pct(lat_p["point"], 99),
                                           "solve_p50": _pct(lat_p["solve"], 50), "solve_p99": _pct(lat_p["solve"], 99)},
                            "path_steps_per_s": steps / wall_p, "wall_s": wall_p,
                            "form": "one request at a time in arrival order, plain FP64 kernel (one path per lane)"}
    return out


def batching_bench(eng):
    """The batching configuration: disc mesh, autotuned launch shape, variance-targeted per-knot
    path counts, every round one persistent launch of uneven counts (FP64 paths)."""
    from pdds_amd.engine import default_run
    from pdds_amd import workloads as W
    t0 = time.perf_counter()
    mesh, prob, h, seed, cfg = W.build_disc(eng)
    t_mesh = time.perf_counter() - t0
    base = default_run(h=h, seed=seed)
    progress("batching: mesh built (%d knots), autotuning" % mesh.n_knots)
    best, tune = eng.autotune(mesh, prob, base)
    progress("batching: variance-targeted counts")
    t1 = time.perf_counter()
    out = eng.solve_adaptive(mesh, prob, best, **cfg)
    wall = time.perf_counter() - t1
    rep = out["report"]
    mc = ~mesh.on_boundary
    N = out["paths_per_knot"][mc]
    uex = prob.u(mesh.px, mesh.py)
    err = out["u"] - uex
    # the same number of paths spread evenly over the knots, one assembly and solve
    n_even = max(1, int(round(rep["paths"] / mc.sum())))
    progress("batching: even counts (%d paths per knot)" % n_even)
    t2 = time.perf_counter()
    ev = eng.assemble(mesh, prob, best, n_even)
    u_even = eng.solve(mesh.n_knots, mesh.row_ptr, mesh.cols, ev["G"], ev["B"], tol=1e-10)[0]
    t_even = time.perf_counter() - t2
    # the equation-variance rule at the same target (the knot equation B_k - sum_j G_kj u_j)
    progress("batching: equation-variance counts")
    t3 = time.perf_counter()
    eq = eng.solve_adaptive(mesh, prob, best, variance=1, **cfg)
    wall_eq = time.perf_counter() - t3
    req = eq["report"]
    Neq = eq["paths_per_knot"][mc]
    return {
        "workload": "disc r=1, 6 rings x 24 sectors, 8..64 knots/interface (seed 11): %d knots (%d MC), "
                    "cfg2 PDE from 512^2 LUTs, h=%g, FP64 paths, SE target %g per knot score (the reference's "
                    "variance rule), N in [%d, %d]"
                    % (mesh.n_knots, int(mc.sum()), h, cfg["se_target"], cfg["n_min"], cfg["n_max"]),
        "path_steps_per_s": rep["path_steps"] / wall, "seconds": wall, "mesh_build_s": t_mesh,
        "rounds": rep["rounds"], "paths": rep["paths"], "path_steps": rep["path_steps"],
        "paths_per_knot": {"min": int(N.min()), "median": int(np.median(N)), "p99": int(np.percentile(N, 99)),
                           "max": int(N.max())},
        "max_se": rep["max_se"], "knots_over": rep["knots_over"], "knots_capped": rep["knots_capped"],
        "kernel_ms": rep["kernel_ms"],
        "rms_interface_error": float(np.sqrt(np.mean(err ** 2))),
        "even_counts": {"paths_per_knot": n_even, "path_steps": ev["counters"]["path_steps"], "seconds": t_even,
                        "rms_interface_error": float(np.sqrt(np.mean((u_even - uex) ** 2)))},
        "adaptive_over_even": {"rms": float(np.sqrt(np.mean(err ** 2))) / float(np.sqrt(np.mean((u_even - uex) ** 2))),
                               "path_steps": rep["path_steps"] / max(ev["counters"]["path_steps"], 1),
                               "note": "above 1: the variance-targeted counts lose to the same paths spread evenly "
                                       "(the knots with the largest score variance are also those with the longest "
                                       "paths, and the score's standard error is not what limits the interface error)"},
        "equation_variance": {"seconds": wall_eq, "rounds": req["rounds"], "paths": req["paths"],
                              "path_steps": req["path_steps"], "max_se": req["max_se"],
                              "knots_over": req["knots_over"], "knots_capped": req["knots_capped"],
                              "paths_per_knot": {"min": int(Neq.min()), "median": int(np.median(Neq)),
                                                 "max": int(Neq.max())},
                              "rms_interface_error": float(np.sqrt(np.mean((eq["u"] - uex) ** 2)))},
        "autotune": {"best": {k: getattr(best, k) for k in ("lut_lds", "wg_size", "grid_wgs", "paths_per_lane",
                                                                "lut_pad")},
                     "candidates": tune["candidates"], "best_rate": tune["best_rate"], "base_rate": tune["base_rate"],
                     "ms": tune["total_ms"]},
    }


def rms_error(eng, mesh, prob, out):
    u, it, rr = eng.solve(mesh.n_knots, mesh.row_ptr, mesh.cols, out["G"], out["B"], tol=1e-12)
    ex = prob.u(mesh.px, mesh.py)
    return float(np.sqrt(np.mean((u - ex) ** 2))), it


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="cfg2_varcoef_lut")
    ap.add_argument("--paths", type=int, default=0, help="override paths per knot (0: the config's)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"],
                    help="path precision of the timed assemblies (the configuration's is FP64)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-only", action="store_true", help="print only the serving scenario")
    ap.add_argument("--no-batching", action="store_true")
    ap.add_argument("--batching-only", action="store_true", help="print only the batching (disc, adaptive) scenario")
    ap.add_argument("--serving-seconds", type=float, default=1.5)
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "traffic_cfg2.json"))
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 3 device-to-device copies of 1 GiB (the reference kernel for HBM counters)")
    ap.add_argument("--summarize-traffic", metavar="PMC_DIR",
                    help="only write --traffic-file from the counter runs of TRAFFIC_RUNS under PMC_DIR")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.summarize_traffic:
        summarize_traffic(args.summarize_traffic, args.traffic_file)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: the path engine has no multi-GPU form in this build "
              "(--gpus %d, WORLD_SIZE %d requested)" % (args.gpus, world), file=sys.stderr)
        sys.exit(2)

    from pdds_amd.engine import Engine, default_run
    from pdds_amd import workloads as W

    eng = Engine(0)
    info = eng.info()
    if args.serving_only:
        print(json.dumps({"serving": serving_bench(eng, args.serving_seconds)}))
        eng.close()
        return
    if args.batching_only:
        print(json.dumps({"batching": batching_bench(eng)}))
        eng.close()
        return
    mesh, prob, n_paths, h, seed = W.build(args.config, eng)
    if args.paths:
        n_paths = args.paths
    run = default_run(h=h, seed=seed, precision=args.precision)

    progress("%s: %d warmup + %d timed assemblies, %s paths" % (args.config, args.warmup, args.steps, args.precision))
    for _ in range(args.warmup):
        eng.assemble(mesh, prob, run, n_paths)
    steps = 0
    kern_ms = []
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        out = eng.assemble(mesh, prob, run, n_paths)
        lat.append(time.perf_counter() - ts)
        steps += out["counters"]["path_steps"]
        kern_ms.append(out["counters"]["kernel_ms"])
    wall = time.perf_counter() - t0
    value = steps / wall
    progress("timed assemblies done: %.1f ms per assembly" % (1e3 * wall / args.steps))
    rms, iters = rms_error(eng, mesh, prob, out)
    c = out["counters"]
    avg_kernel_s = float(np.mean(kern_ms)) / 1e3
    flops = FLOPS_PER_STEP.get(args.config, 302.0) * c["path_steps"]
    achieved = flops / avg_kernel_s / 1e12
    peak = PEAK_FP64_TFLOPS if args.precision == "fp64" else PEAK_FP32_TFLOPS
    # HBM bytes per launch, only if measured on these very sources (the counter runs cannot share
    # a process with the timing); null otherwise
    traffic, traffic_src = None, "none measured for this build"
    if args.precision == "fp64" and os.path.exists(args.traffic_file):
        try:
            tj = json.load(open(args.traffic_file))
            if tj.get("source_hash") == kernel_source_hash():
                traffic = tj.get("bytes_per_launch")
                traffic_src = "%s (measured on source hash %s)" % (os.path.relpath(args.traffic_file, ROOT),
                                                                   tj["source_hash"])
            else:
                traffic_src = "%s is for source hash %s, not this build's %s" % (
                    os.path.relpath(args.traffic_file, ROOT), tj.get("source_hash"), kernel_source_hash())
        except (OSError, ValueError):
            traffic = None
    result = {
        "metric": "path-steps/s at fixed RMS interface error; p50/p99 solve latency vs plain GPU",
        "value": value,
        "unit": "path-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 (FP64 sums)",
        "data": "synthetic (manufactured solution u = e^x cos y; LUTs sampled from the analytic coefficients)",
        "config": {"workload": "plain PDDSparse assembly, cfg2: 8x8 subdomains x 32 knots/interface, "
                               "3437 knots (3409 MC), %d paths/knot, h=%g, FP64 paths, 512x512 LUTs, seed %d"
                               % (n_paths, h, seed),
                   "config_name": args.config, "knots": mesh.n_knots, "g_nnz": mesh.nnz,
                   "paths_per_knot": n_paths, "h": h},
        "path_steps_per_step": c["path_steps"],
        "rms_interface_error": rms,
        "bicgstab_iters": iters,
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "launch": {"grid_wgs": c["grid_wgs"], "wg_size": c["wg_size"], "item_paths": c["item_paths"],
                   "sub_batch": c["sub_batch"], "work_items": c["work_items"]},
        "device": info,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic,
                     "kernel": ("path_kernel_tiled<%s> (LDS-tiled LUT)" % ("double" if args.precision == "fp64" else "float")
                                if c["wg_size"] != 256 else "path_kernel (persistent, LUT from global memory)"),
                     "avg_kernel_ms": 1e3 * avg_kernel_s,
                     "traffic_source": traffic_src,
                     "source_hash": kernel_source_hash(),
                     "note": ("no MFMA (no step is a dense contraction): the kernel is bound by VALU issue; peak = "
                              "MI355X %s; achieved counts %g algorithmic flops per path-step (transcendentals and "
                              "Philox integer work not counted); traffic = HBM bytes per launch from "
                              "FETCH_SIZE/WRITE_SIZE, copy-calibrated"
                              % ("FP64 vector 78.6 TFLOP/s" if args.precision == "fp64"
                                 else "FP32 vector 157.3 TFLOP/s (packed; this build issues unpacked FP32)",
                                 FLOPS_PER_STEP.get(args.config, 302.0)))},
    }
    if not args.no_extras:
        extras = {}
        # engine_fp32_corner_sum: the same FP32 assembly with corner entries shared by two sides of a
        # stencil summed (quirk_last_side = 0) instead of the reference's "=+", which keeps only the
        # last side's weight; the interface error shows what that rule costs
        for name, kw in (("plain_fp64", dict(plain=1)), ("engine_fp32", dict(precision="fp32")),
                         ("engine_fp32_corner_sum", dict(precision="fp32", quirk_last_side=0))):
            progress("form %s" % name)
            r2 = default_run(h=h, seed=seed, **kw)
            ts = time.perf_counter()
            o2 = eng.assemble(mesh, prob, r2, n_paths)
            dt = time.perf_counter() - ts
            e2, _ = rms_error(eng, mesh, prob, o2)
            extras[name] = {"path_steps_per_s": o2["counters"]["path_steps"] / dt, "seconds": dt,
                            "kernel_ms": o2["counters"]["kernel_ms"], "rms_interface_error": e2}
        extras["engine_fp64"] = {"path_steps_per_s": value, "seconds": wall / args.steps, "rms_interface_error": rms}
        extras["engine_fp32_vs_plain_fp64"] = extras["engine_fp32"]["path_steps_per_s"] / extras["plain_fp64"]["path_steps_per_s"]
        # the batching form (variance-targeted counts, pilot and rounds included) on the FP32 engine
        # and on the plain FP64 kernel; on this workload the RMS error is set by the time step and
        # the RBF interpolation (it is the same from 1e3 to 1e5 paths per knot), not by the counts
        for name, kw in (("adaptive_fp32", dict(precision="fp32")), ("adaptive_plain_fp64", dict(plain=1))):
            progress("form %s" % name)
            ts = time.perf_counter()
            ad = eng.solve_adaptive(mesh, prob, default_run(h=h, seed=seed, **kw), se_target=4e-3,
                                    n_min=1000, n_max=1000000, max_rounds=4)
            dt = time.perf_counter() - ts
            ad_rms = float(np.sqrt(np.mean((ad["u"] - prob.u(mesh.px, mesh.py)) ** 2)))
            extras[name] = {"seconds": dt, "rms_interface_error": ad_rms, "paths": ad["report"]["paths"],
                            "path_steps": ad["report"]["path_steps"], "rounds": ad["report"]["rounds"],
                            "max_se": ad["report"]["max_se"], "knots_over": ad["report"]["knots_over"],
                            "kernel_ms": ad["report"]["kernel_ms"]}
        extras["adaptive_plain_fp64_over_adaptive_fp32"] = (extras["adaptive_plain_fp64"]["seconds"]
                                                            / extras["adaptive_fp32"]["seconds"])
        result["forms"] = extras
    if not args.no_serving and not args.no_extras:
        progress("serving scenario")
        result["serving"] = serving_bench(eng, args.serving_seconds)
    if not args.no_extras and not args.no_batching:
        progress("batching scenario")
        result["batching"] = batching_bench(eng)
    if not args.no_cpu_baseline:
        progress("CPU baseline")
        result["cpu_baseline"] = cpu_baseline(mesh, {3: "varcoef_lut", 0: "monegros", 1: "sinsin"}[prob.pde],
                                              n_paths, h, seed)
    eng.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
