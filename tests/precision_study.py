"""FP32 path state against the FP64-stepped oracle on identical Philox streams.

Workload: the 2x2-subdomain Poisson mesh on [-1,1]^2 with 16 knots per interface (61 knots, 57 of
them Monte Carlo), u = sin(pi x) sin(pi y) + x^2 + y^2, sigma = sqrt2 I, Gaussian RBF shape 1.0
relative to the knot spacing, Philox seed 42.

The FP32 engine draws one Philox4x32-10 call per two Euler-Maruyama steps and turns its four
words into two normal pairs from 24-bit uniforms.  The oracle's normals mode 1 draws the same
words and steps the paths in FP64, so every difference below comes from the FP32 path state
(positions and Y in FP32, Z and xi Kahan-compensated FP32) and FP32 elementary functions; the
cross-path sums are FP64 on both sides.  Reported per h:

- paths: the share of paths whose step count is the same, and the largest deviation of the
  exit records over those paths;
- G and B: the largest deviation and the largest deviation in units of the Monte Carlo standard
  error sqrt(var / N) of the entry;
- u: the RMS difference of the interface solutions of I + G against the RMS error of the FP64
  solution against the exact u.

Used by tests/test_gpu_precision.py at sizes the oracle finishes in seconds; run as a script for
the full-size table (`python tests/precision_study.py --paths 10000 --out study.json`).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

SEED = 42
S, P = 32, 640


def _mesh():
    from pdds_amd.engine import Mesh
    from pdds_amd.workloads import relative_fac
    sw, ne = (-1.0, -1.0), (1.0, 1.0)
    fac = relative_fac(sw, ne, 2, 16, 1.0)
    return Mesh(sw, ne, 2, 16, fac), fac


def _cshape(st, fac):
    return [fac[min(max(int(k), 0), 3)] for k in st["kind"]]


def study(eng, h, n_paths, threads=8, quirk=1):
    from pdds_amd.engine import bvp, default_run
    m, fac = _mesh()
    prob = bvp("sinsin")
    pb = O.problem(O.PDE_SINSIN)
    orun = O.run_params(h, SEED, normals=1, threads=threads)
    run32 = default_run(h=h, seed=SEED, precision="fp32", sub_batch=S, item_paths=P, quirk_last_side=quirk)
    t0 = time.perf_counter()
    got = eng.assemble(m, prob, run32, n_paths)
    t_gpu = time.perf_counter() - t0

    G64 = np.zeros(m.nnz)
    Gv64 = np.zeros(m.nnz)
    B64 = np.zeros(m.n_knots)
    BB64 = np.zeros(m.n_knots)
    same = tot = 0
    dev = {"xt": 0.0, "yt": 0.0, "Y": 0.0, "Z": 0.0, "xi": 0.0}
    oracle_steps = 0
    t0 = time.perf_counter()
    for k in range(m.n_knots):
        if m.on_boundary[k]:
            B64[k] = prob.g(m.x0[k], m.y0[k])
            BB64[k] = B64[k] ** 2
            continue
        st = m.stencil(k)
        paths = O.simulate(pb, orun, m.x0[k], m.y0[k], st["rect"], k, 0, n_paths)
        oracle_steps += int(paths["steps"].sum())
        ref = O.reduce_batched(pb, paths, st, m.glob, _cshape(st, fac), st["ipsi"], S, P, quirk=quirk)
        a, b = m.row_ptr[k], m.row_ptr[k + 1]
        assert np.array_equal(ref["cols"], m.row_cols(k))
        G64[a:b] = ref["G"]
        Gv64[a:b] = ref["Gvar"]
        B64[k] = ref["B"]
        BB64[k] = ref["BB"]
        if k % 4 == 1:  # per-path records on every fourth knot
            p32 = eng.paths(prob, default_run(h=h, seed=SEED, precision="fp32"), m.x0[k], m.y0[k], st["rect"], k,
                            0, n_paths)
            sel = p32["steps"] == paths["steps"]
            same += int(sel.sum())
            tot += n_paths
            for key in dev:
                if sel.any():
                    dev[key] = max(dev[key], float(np.abs(p32[key][sel] - paths[key][sel]).max()))
    t_oracle = time.perf_counter() - t0

    def dense(G, B):
        A = np.eye(m.n_knots)
        for k in range(m.n_knots):
            a, b = m.row_ptr[k], m.row_ptr[k + 1]
            A[k, m.cols[a:b]] += G[a:b]
        return np.linalg.solve(A, B)

    u32 = dense(got["G"], got["B"])
    u64 = dense(G64, B64)
    ex = prob.u(m.px, m.py)
    mc = ~m.on_boundary
    dG = np.abs(got["G"] - G64)
    seG = np.sqrt(np.maximum(Gv64, 0) / n_paths)
    dB = np.abs(got["B"] - B64)[mc]
    seB = np.sqrt(np.maximum(BB64 - B64 ** 2, 0) / n_paths)[mc]
    rms = lambda v: float(np.sqrt(np.mean(v ** 2)))  # noqa: E731
    nz = seG > 0
    return {
        "h": h, "paths_per_knot": n_paths, "mc_knots": int(mc.sum()), "g_nnz": int(m.nnz),
        "path_steps_fp32": int(got["counters"]["path_steps"]), "path_steps_oracle": oracle_steps,
        "paths_compared": tot, "same_step_count": same / max(tot, 1),
        "max_record_dev_same_steps": dev,
        "G": {"max_abs_dev": float(dG.max()), "rms_dev": rms(dG), "max_dev_over_se": float((dG[nz] / seG[nz]).max()),
              "rms_se": rms(seG)},
        "B": {"max_abs_dev": float(dB.max()), "rms_dev": rms(dB), "max_dev_over_se": float((dB / seB).max()),
              "rms_se": rms(seB)},
        "u": {"rms_fp32_minus_fp64": rms(u32 - u64), "rms_err_fp64": rms(u64 - ex), "rms_err_fp32": rms(u32 - ex),
              "max_abs_fp32_minus_fp64": float(np.abs(u32 - u64).max())},
        "seconds": {"gpu_fp32_assembly": t_gpu, "oracle": t_oracle},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--paths", type=int, default=10000)
    ap.add_argument("--h", type=float, nargs="*", default=[1e-3, 1e-4, 1e-5])
    ap.add_argument("--threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "8")))
    ap.add_argument("--out", default="")
    args = ap.parse_args()
    from pdds_amd.engine import Engine
    eng = Engine(0)
    rows = []
    for h in args.h:
        r = study(eng, h, args.paths, threads=args.threads)
        print(json.dumps(r), flush=True)
        rows.append(r)
    eng.close()
    if args.out:
        with open(args.out, "w") as f:
            json.dump({"workload": "2x2 subdomains x 16 knots/interface on [-1,1]^2, u = sin(pi x) sin(pi y) "
                                   "+ x^2 + y^2, seed 42; FP32 engine vs FP64-stepped oracle, same Philox words",
                       "rows": rows, "command": "python3 tests/precision_study.py " + " ".join(sys.argv[1:])},
                      f, indent=1)


if __name__ == "__main__":
    main()
