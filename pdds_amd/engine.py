"""Python host mirror of the engine's C ABI (include/pdds_engine.h).

`Engine` owns one pdds_engine (one HIP device).  `Mesh` is the PDDSparse
interface/knot/stencil bookkeeping.  `bvp` selects a compiled-in PDE
functor, the device-side replacement of the reference's function-pointer
struct (BVPs/BVP.hpp:89-137).
"""
import ctypes as C

import numpy as np

from . import _capi as A
from ._capi import PddsError, ptr, f64

_PDE_NAMES = {"monegros": A.PDE_MONEGROS, "sinsin": A.PDE_SINSIN, "varcoef": A.PDE_VARCOEF,
              "varcoef_lut": A.PDE_VARCOEF_LUT, "helmholtz": A.PDE_HELMHOLTZ, "table": A.PDE_TABLE}


class bvp:
    """Boundary value problem: which compiled-in coefficient functor to use.

    Reference convention 1/2 s s^T : D2u + b.Du + c u + f = 0 (c <= 0), Dirichlet g = u.
    `u(x, y)` evaluates the manufactured solution on the host (for error reports).
    """

    def __init__(self, pde="monegros", k2=0.0, lut=None, neumann=0, varphi=0.0, rho=0.0, ulut=None, vlut=None):
        """neumann: reflecting global sides (bits SIDE_N/S/E/W of _capi), where du/dn = varphi u + psi
        (varphi <= 0; psi from the solution); rho: Lepingle step within this distance of them.
        "table": every coefficient from `lut` (8 fields: sigma11, sigma22, b1, b2, c, f, du/dx,
        du/dy) and u = g from `ulut` (1 field).  vlut: mu1, mu2, F1, F2 of the VR+CV step."""
        self.pde = _PDE_NAMES[pde] if isinstance(pde, str) else int(pde)
        self.k2 = float(k2)
        self.lut = lut
        self.ulut = ulut
        self.vlut = vlut
        self.neumann = int(neumann)
        self.varphi = float(varphi)
        self.rho = float(rho)
        if self.pde == A.PDE_VARCOEF_LUT and lut is None:
            raise ValueError("varcoef_lut needs a LUT")
        if self.pde == A.PDE_TABLE and (lut is None or ulut is None):
            raise ValueError("table needs lut and ulut")

    def _c(self):
        p = A.Problem()
        p.pde = self.pde
        p.k2 = self.k2
        p.lut = self.lut.h if self.lut is not None else None
        p.ulut = self.ulut.h if self.ulut is not None else None
        p.vlut = self.vlut.h if self.vlut is not None else None
        p.refl = self.neumann
        p.robin = self.varphi
        p.rho = self.rho
        return p

    def u(self, x, y):
        x = np.asarray(x, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        if self.pde == A.PDE_TABLE:
            if not hasattr(self, "u_exact"):
                raise ValueError("a tabulated problem has no closed-form u; set bvp.u_exact")
            return self.u_exact(x, y)
        if self.pde == A.PDE_MONEGROS:
            k1, k2, k3, k4, k5, k6, k7, sf = 1.0, 0.01, 0.02, 0.12, 0.05, 0.05, -0.12, 0.333
            return 3.0 + sf * (np.sin(np.sqrt(k1 + k2 * x * x + k3 * y * y))
                               + np.tanh(np.sin(k4 * x + k5 * y) + np.sin(k6 * x + k7 * y)))
        if self.pde in (A.PDE_SINSIN, A.PDE_HELMHOLTZ):
            return np.sin(np.pi * x) * np.sin(np.pi * y) + x * x + y * y
        return np.exp(x) * np.cos(y)

    g = u


def default_run(h=1e-3, seed=42, precision="fp64", **kw):
    r = A.Run()
    A.lib().pdds_run_defaults(C.byref(r))
    r.h = h
    r.seed = seed
    r.precision = A.FP32 if precision in ("fp32", A.FP32) else A.FP64
    for k, v in kw.items():
        if not hasattr(r, k):
            raise AttributeError(k)
        setattr(r, k, v)
    return r


class Engine:
    def __init__(self, device=0):
        h = C.c_void_p()
        st = A.lib().pdds_engine_create(device, C.byref(h))
        if st != A.OK:
            raise PddsError(st, "engine create on device %d" % device)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            A.lib().pdds_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, st, what):
        if st != A.OK:
            raise PddsError(st, "%s: %s" % (what, A.lib().pdds_last_error(self.h).decode()))

    def info(self):
        n = C.c_int32()
        name = C.create_string_buffer(128)
        self._chk(A.lib().pdds_engine_info(self.h, C.byref(n), name, 128), "info")
        return {"n_cu": n.value, "arch": name.value.decode()}

    # -- LUTs --
    def lut(self, samples, box, mode=0):
        return Lut(self, samples, box, mode)

    # -- per-path records (MCinCUDA shape) --
    def paths(self, problem, run, x0, y0, rect, knot, path0, n, phase2=False):
        o = {k: np.zeros(n) for k in ("xt", "yt", "Y", "Z", "xi")}
        o["steps"] = np.zeros(n, dtype=np.int64)
        p2 = {k: (np.zeros(n) if phase2 else None) for k in ("xt", "yt", "Y", "Z", "xi")}
        s2 = np.zeros(n, dtype=np.int64) if phase2 else None
        ctr = A.Counters()
        rect = f64(rect)
        self._chk(A.lib().pdds_paths(self.h, C.byref(problem._c()), C.byref(run), x0, y0, ptr(rect), knot, path0, n,
                                     1 if phase2 else 0, ptr(o["xt"]), ptr(o["yt"]), ptr(o["Y"]), ptr(o["Z"]),
                                     ptr(o["xi"]), ptr(o["steps"], A._i64), ptr(p2["xt"]), ptr(p2["yt"]),
                                     ptr(p2["Y"]), ptr(p2["Z"]), ptr(p2["xi"]), ptr(s2, A._i64), C.byref(ctr)),
                  "paths")
        if phase2:
            o["phase2"] = dict(p2, steps=s2)
        o["counters"] = ctr.as_dict()
        return o

    def solve_knot(self, mesh, problem, run, knot, n_paths):
        nnz = mesh.row_nnz(knot)
        G = np.zeros(max(nnz, 1))
        Gv = np.zeros(max(nnz, 1))
        B = C.c_double()
        BB = C.c_double()
        ctr = A.Counters()
        self._chk(A.lib().pdds_solve_knot(self.h, mesh.h, C.byref(problem._c()), C.byref(run), knot, n_paths,
                                          ptr(G), ptr(Gv), C.byref(B), C.byref(BB), C.byref(ctr)), "solve_knot")
        return {"cols": mesh.row_cols(knot), "G": G[:nnz], "Gvar": Gv[:nnz], "B": B.value, "BB": BB.value,
                "counters": ctr.as_dict()}

    def knot_steps(self, mesh, problem, run):
        """Step h of each knot's paths under run (its adaptive_h rule); 0 for boundary knots."""
        h = np.zeros(mesh.n_knots)
        self._chk(A.lib().pdds_knot_steps(self.h, mesh.h, C.byref(problem._c()), C.byref(run), ptr(h)), "knot_steps")
        return h

    def assemble(self, mesh, problem, run, n_paths, paths_per_knot=None):
        nnz = mesh.nnz
        G = np.zeros(max(nnz, 1))
        Gv = np.zeros(max(nnz, 1))
        B = np.zeros(mesh.n_knots)
        BB = np.zeros(mesh.n_knots)
        ppk = None if paths_per_knot is None else np.ascontiguousarray(paths_per_knot, dtype=np.int64)
        ctr = A.Counters()
        self._chk(A.lib().pdds_assemble(self.h, mesh.h, C.byref(problem._c()), C.byref(run), ptr(ppk, A._i64),
                                        n_paths, ptr(G), ptr(Gv), ptr(B), ptr(BB), C.byref(ctr)), "assemble")
        return {"G": G[:nnz], "Gvar": Gv[:nnz], "B": B, "BB": BB, "counters": ctr.as_dict()}

    def assemble_eq(self, mesh, problem, run, n_paths, u_ref, paths_per_knot=None):
        """pdds_assemble_eq: assemble() plus each knot equation's mean and per-path variance
        with reference interface values u_ref."""
        nnz = mesh.nnz
        G = np.zeros(max(nnz, 1))
        Gv = np.zeros(max(nnz, 1))
        B, BB, em, ev = (np.zeros(mesh.n_knots) for _ in range(4))
        ppk = None if paths_per_knot is None else np.ascontiguousarray(paths_per_knot, dtype=np.int64)
        ur = f64(u_ref)
        if ur.shape != (mesh.n_knots,):
            raise ValueError("u_ref needs one value per knot")
        ctr = A.Counters()
        self._chk(A.lib().pdds_assemble_eq(self.h, mesh.h, C.byref(problem._c()), C.byref(run), ptr(ppk, A._i64),
                                           n_paths, ptr(ur), ptr(G), ptr(Gv), ptr(B), ptr(BB), ptr(em), ptr(ev),
                                           C.byref(ctr)), "assemble_eq")
        return {"G": G[:nnz], "Gvar": Gv[:nnz], "B": B, "BB": BB, "eq_mean": em, "eq_var": ev,
                "counters": ctr.as_dict()}

    def solve_adaptive(self, mesh, problem, run, **cfg):
        """pdds_solve_adaptive: variance-targeted per-knot path counts (keys of pdds_adaptive)."""
        c = A.Adaptive()
        A.lib().pdds_adaptive_defaults(C.byref(c))
        for k, v in cfg.items():
            setattr(c, k, v)
        n = mesh.n_knots
        u = np.zeros(n)
        N = np.zeros(n, dtype=np.int64)
        se = np.zeros(n)
        rep = A.AdaptiveReport()
        self._chk(A.lib().pdds_solve_adaptive(self.h, mesh.h, C.byref(problem._c()), C.byref(run), C.byref(c),
                                              ptr(u), ptr(N, A._i64), ptr(se), C.byref(rep)), "solve_adaptive")
        return {"u": u, "paths_per_knot": N, "se": se, "report": rep.as_dict()}

    def autotune(self, mesh, problem, run, sample_paths=0, max_knots=0, reps=0):
        """pdds_autotune: returns (tuned run, report)."""
        cfg = A.TuneConfig(sample_paths, max_knots, reps)
        best = A.Run()
        rep = A.TuneReport()
        self._chk(A.lib().pdds_autotune(self.h, mesh.h, C.byref(problem._c()), C.byref(run), C.byref(cfg),
                                        C.byref(best), C.byref(rep)), "autotune")
        return best, rep.as_dict()

    def point_eval(self, problem, run, glob, x, y, n_paths):
        x = f64(x)
        y = f64(y)
        n = len(x)
        npth = np.ascontiguousarray(np.broadcast_to(np.asarray(n_paths, dtype=np.int64), (n,)))
        u = np.zeros(n)
        uv = np.zeros(n)
        ctr = A.Counters()
        self._chk(A.lib().pdds_point_eval(self.h, C.byref(problem._c()), C.byref(run), ptr(f64(glob)), n, ptr(x),
                                          ptr(y), ptr(npth, A._i64), ptr(u), ptr(uv), C.byref(ctr)), "point_eval")
        return {"u": u, "u_var": uv, "counters": ctr.as_dict()}

    def solve(self, n, row_ptr, cols, G, B, tol=1e-14, maxit=0):
        rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
        ci = np.ascontiguousarray(cols, dtype=np.int32)
        Gv = f64(G)
        Bv = f64(B)
        u = np.zeros(n)
        it = C.c_int32()
        rr = C.c_double()
        self._chk(A.lib().pdds_solve(self.h, n, ptr(rp, A._i64), ptr(ci, A._i32), ptr(Gv), ptr(Bv), ptr(u), tol,
                                     maxit, C.byref(it), C.byref(rr)), "solve")
        return u, it.value, rr.value


class Lut:
    """Device-resident bicubic B-spline coefficient table (see pdds_lut_create)."""

    def __init__(self, engine, samples, box, mode=0):
        s = f64(samples)
        if s.ndim == 2:
            s = s[:, :, None]
        ny, nx, nf = s.shape
        h = C.c_void_p()
        engine._chk(A.lib().pdds_lut_create(engine.h, nx, ny, nf, ptr(s), box[0], box[1], box[2], box[3], mode,
                                            C.byref(h)), "lut_create")
        self.engine = engine
        self.h = h
        self.nf = nf

    def eval(self, field, x, y):
        x = f64(x)
        y = f64(y)
        out = np.zeros(len(x))
        self.engine._chk(A.lib().pdds_lut_eval(self.engine.h, self.h, field, len(x), ptr(x), ptr(y), ptr(out)),
                         "lut_eval")
        return out

    def __del__(self):
        if getattr(self, "h", None):
            A.lib().pdds_lut_destroy(self.h)
            self.h = None


class Mesh:
    """PDDSparse interfaces, knots and stencils (PDDSparseGM geometry)."""

    def __init__(self, sw, ne, n_sub, n_node, fac, _disc=None):
        h = C.c_void_p()
        if _disc is None:
            fac = f64(fac)
            st = A.lib().pdds_mesh_create(sw[0], sw[1], ne[0], ne[1], n_sub, n_node, ptr(fac), C.byref(h))
        else:
            st = A.lib().pdds_mesh_create_disc(*_disc, C.byref(h))
        if st not in (A.OK, A.ERR_SINGULAR):
            raise PddsError(st, "mesh create")
        self.h = h
        self.singular = st == A.ERR_SINGULAR
        self.glob = np.array([sw[0], sw[1], ne[0], ne[1]], dtype=np.float64)
        self.n_knots = A.lib().pdds_mesh_n_knots(h)
        self.n_stencils = A.lib().pdds_mesh_n_stencils(h)
        self.n_node = n_node
        g = C.c_int32()
        A.lib().pdds_mesh_geometry(h, 0, C.byref(g), None)
        self.geom = g.value
        nnz = C.c_int64()
        A.lib().pdds_mesh_csr(h, None, None, C.byref(nnz))
        self.nnz = nnz.value
        self.row_ptr = np.zeros(self.n_knots + 1, dtype=np.int64)
        self.cols = np.zeros(max(self.nnz, 1), dtype=np.int32)
        A.lib().pdds_mesh_csr(h, ptr(self.row_ptr, A._i64), ptr(self.cols, A._i32), C.byref(nnz))
        self.cols = self.cols[:self.nnz]
        n = self.n_knots
        self.x0, self.y0, self.px, self.py = (np.zeros(n) for _ in range(4))
        self.on_boundary = np.zeros(n, dtype=np.int32)
        self.stencil_id = np.zeros(n, dtype=np.int32)
        A.lib().pdds_mesh_knots(h, ptr(self.x0), ptr(self.y0), ptr(self.px), ptr(self.py),
                                ptr(self.on_boundary, A._i32), ptr(self.stencil_id, A._i32))
        self.on_boundary = self.on_boundary.astype(bool)

    @classmethod
    def disc(cls, radius=1.0, n_rings=6, n_sectors=24, seed=11, kmin=8, kmax=64, shape=1.0):
        """Disc of `radius` cut into n_rings rings x n_sectors sectors, kmin..kmax knots per
        interface drawn with `seed` (see pdds_mesh_create_disc)."""
        return cls((-radius, -radius), (radius, radius), 0, 0, None,
                   _disc=(radius, n_rings, n_sectors, seed, kmin, kmax, shape))

    def __del__(self):
        if getattr(self, "h", None):
            A.lib().pdds_mesh_destroy(self.h)
            self.h = None

    def row_nnz(self, k):
        return int(self.row_ptr[k + 1] - self.row_ptr[k])

    def row_cols(self, k):
        return self.cols[self.row_ptr[k]:self.row_ptr[k + 1]].copy()

    def stencil(self, k):
        rect = np.zeros(4)
        msz = np.zeros(4, dtype=np.int32)
        kind = np.zeros(4, dtype=np.int32)
        ne = C.c_int32()
        A.lib().pdds_mesh_stencil(self.h, k, ptr(rect), ptr(msz, A._i32), ptr(kind, A._i32), 0, None, None, None,
                                  C.byref(ne))
        cap = max(ne.value, 1)
        idx = np.zeros(cap, dtype=np.int32)
        px = np.zeros(cap)
        py = np.zeros(cap)
        A.lib().pdds_mesh_stencil(self.h, k, ptr(rect), ptr(msz, A._i32), ptr(kind, A._i32), cap, ptr(idx, A._i32),
                                  ptr(px), ptr(py), C.byref(ne))
        cond = np.zeros(4)
        A.lib().pdds_mesh_cond(self.h, k, ptr(cond))
        t = ne.value
        ipsi = []
        for s in range(4):
            m = int(msz[s])
            a = np.zeros(max(m * m, 1))
            mm = C.c_int32()
            A.lib().pdds_mesh_ipsi(self.h, k, s, a.size, ptr(a), C.byref(mm))
            ipsi.append(a[:m * m].reshape(m, m))
        dom = np.zeros(6)
        A.lib().pdds_mesh_geometry(self.h, k, None, ptr(dom))
        cshape = np.zeros(4)
        A.lib().pdds_mesh_shape(self.h, k, ptr(cshape))
        cent = np.zeros(max(t, 1))
        A.lib().pdds_mesh_entry_shapes(self.h, k, len(cent), ptr(cent), None)
        return {"rect": rect, "msz": msz, "kind": kind, "idx": idx[:t], "px": px[:t], "py": py[:t], "cond": cond,
                "ipsi": ipsi, "domain": dom, "geom": self.geom, "cshape": cshape, "cent": cent[:t]}


class Server:
    """Continuous-batching request loop over an engine (pdds_server_*): submit point evaluations
    and interface solves as they arrive, collect each result by ticket."""

    def __init__(self, engine, problem, run, domain=(-1.0, -1.0, 1.0, 1.0), hbm_budget=0, max_inflight=0,
                 prep_threads=0, round_paths=0):
        cfg = A.ServerConfig()
        cfg.problem = problem._c()
        cfg.run = run
        for i in range(4):
            cfg.domain[i] = domain[i]
        cfg.hbm_budget, cfg.max_inflight, cfg.prep_threads, cfg.round_paths = hbm_budget, max_inflight, prep_threads, round_paths
        self._problem = problem          # keeps the LUT (if any) alive
        self.engine = engine
        h = C.c_void_p()
        engine._chk(A.lib().pdds_server_create(engine.h, C.byref(cfg), C.byref(h)), "server create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            A.lib().pdds_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def submit_point(self, x, y, n_paths, k2=0.0, stream=1, h=0.0):
        """h: this request's step (0: the server run's)."""
        r = A.Request()
        r.kind, r.stream, r.k2, r.n_paths, r.x, r.y, r.h = A.REQ_POINT, stream, k2, n_paths, x, y, h
        return self._submit(r)

    def submit_solve(self, n_sub, n_node, fac, n_paths, k2=0.0, stream=1, tol=1e-10, h=0.0):
        r = A.Request()
        r.kind, r.stream, r.k2, r.n_paths, r.n_sub, r.n_node, r.tol = A.REQ_SOLVE, stream, k2, n_paths, n_sub, n_node, tol
        r.h = h
        for i in range(4):
            r.fac[i] = fac[i]
        return self._submit(r)

    def _submit(self, r):
        t = C.c_uint64()
        st = A.lib().pdds_server_submit(self.h, C.byref(r), C.byref(t))
        if st != A.OK:
            raise PddsError(st, "server submit")
        return t.value

    def _take(self, fn, ticket, cap):
        res = A.Result()
        u = np.zeros(max(cap, 1))
        st = fn(self.h, ticket, C.byref(res), ptr(u), cap)
        if st == A.ERR_NOT_READY:
            return None
        if st != A.OK:
            raise PddsError(st, "server result")
        out = res.as_dict()
        if res.kind == A.REQ_SOLVE:
            out["u_knots"] = u[:min(cap, res.n_knots)].copy()
        return out

    def poll(self, ticket, cap=1 << 16):
        return self._take(A.lib().pdds_server_poll, ticket, cap)

    def wait(self, ticket, cap=1 << 16):
        return self._take(A.lib().pdds_server_wait, ticket, cap)

    def stats(self):
        v = [C.c_int64() for _ in range(4)]
        b = C.c_double()
        A.lib().pdds_server_stats(self.h, *[C.byref(x) for x in v], C.byref(b))
        return {"submitted": v[0].value, "done": v[1].value, "rounds": v[2].value, "path_steps": v[3].value,
                "busy_ms": b.value}
