"""TEST INFRASTRUCTURE: ctypes binding of the CPU oracle (oracle/liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module.  The product (pdds_amd) never does.  The oracle is a plain-C,
FP64, OpenMP restatement of the reference's PDDSparse CPU path; see
oracle/oracle.h for the function list and how each one is pinned.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

PDE_MONEGROS, PDE_SINSIN, PDE_VARCOEF, PDE_VARCOEF_LUT, PDE_HELMHOLTZ, PDE_TABLE = range(6)

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)
_i64 = C.POINTER(C.c_int64)
_u32 = C.POINTER(C.c_uint32)
_f = C.POINTER(C.c_float)


class Lut(C.Structure):
    _fields_ = [("c", _f), ("nx", C.c_int), ("ny", C.c_int), ("nf", C.c_int),
                ("ox", C.c_double), ("oy", C.c_double), ("sx", C.c_double), ("sy", C.c_double)]


class Problem(C.Structure):
    _fields_ = [("pde", C.c_int), ("k2", C.c_double), ("lut", _f),
                ("lut_nx", C.c_int), ("lut_ny", C.c_int), ("lut_nf", C.c_int),
                ("lut_ox", C.c_double), ("lut_oy", C.c_double),
                ("lut_sx", C.c_double), ("lut_sy", C.c_double), ("step", C.c_int),
                ("refl", C.c_int), ("robin", C.c_double), ("rho", C.c_double),
                ("ulut", Lut), ("vlut", Lut)]


class Run(C.Structure):
    _fields_ = [("h", C.c_double), ("beta", C.c_double), ("seed", C.c_uint32),
                ("stream", C.c_uint32), ("normals", C.c_int), ("phase2", C.c_int),
                ("max_steps", C.c_int64), ("threads", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_mesh_create.restype = C.c_void_p
        L.orc_mesh_create.argtypes = [C.c_double] * 4 + [C.c_int, C.c_int]
        L.orc_mesh_destroy.argtypes = [C.c_void_p]
        L.orc_mesh_n_knots.argtypes = [C.c_void_p]
        L.orc_mesh_n_interfaces.argtypes = [C.c_void_p]
        L.orc_mesh_knots.argtypes = [C.c_void_p, _d, _d, _d, _d, _i]
        L.orc_mesh_stencil.argtypes = [C.c_void_p, C.c_int, _d, _i, _i, _i, _d, _d]
        L.orc_stencil_ipsi.argtypes = [C.c_int, _d, _d, C.c_double, _d, _d]
        L.orc_rect_dist.restype = C.c_double
        L.orc_rect_dist.argtypes = [_d, C.c_double, C.c_double, _d, _d, _d, _d]
        L.orc_sector_dist.restype = C.c_double
        L.orc_sector_dist.argtypes = [_d, C.c_double, C.c_double, _d, _d, _d, _d, _i]
        L.orc_simulate_geom.argtypes = [C.POINTER(Problem), C.POINTER(Run), C.c_double, C.c_double, C.c_int, _d,
                                        C.c_uint32, C.c_int64, C.c_int64] + [_d] * 5 + [_i64]
        L.orc_pde_eval.argtypes = [C.POINTER(Problem), C.c_double, C.c_double, _d, _d, _d, _d, _d, _d, _d]
        L.orc_lut_eval.restype = C.c_double
        L.orc_lut_eval.argtypes = [C.POINTER(Problem), C.c_int, C.c_double, C.c_double]
        L.orc_simulate.argtypes = [C.POINTER(Problem), C.POINTER(Run), C.c_double, C.c_double, _d,
                                   C.c_uint32, C.c_int64, C.c_int64] + [_d] * 5 + [_i64] + [_d] * 5 + [_i64]
        red = [C.POINTER(Problem), C.c_int64, _d, _d, _d, _d, _d, _d, _d, _i, _i, _d, _d, _d, _d, C.c_int,
               _i, _d, _d, _d, _d, _i64]
        L.orc_reduce_faithful.argtypes = red
        L.orc_reduce_batched.argtypes = red[:2] + [C.c_int64, C.c_int64] + red[2:]
        L.orc_reduce_batched_eq.argtypes = (red[:2] + [C.c_int64, C.c_int64] + red[2:16] + [_d] + red[16:] +
                                            [_d, _d])
        L.orc_reduce_batched_geom.argtypes = (red[:2] + [C.c_int64, C.c_int64] + red[2:7] + [C.c_int] + red[7:16] +
                                              [_d, _d] + red[16:] + [_d, _d])
        L.orc_bicgstab.argtypes = [C.c_int, _i64, _i, _d, _d, _d, C.c_double, C.c_int, _d]
        L.orc_philox.argtypes = [_u32, _u32, _u32]
        L.orc_math_eval.argtypes = [C.c_int, _d, _d, C.c_int64]
        L.orc_normals.argtypes = [C.POINTER(Run), C.c_uint32, C.c_uint32, C.c_int64, C.c_int64, _d, _d]
        L.orc_lut_prefilter.argtypes = [C.c_int, C.c_int, C.c_int, _d, _f]
        L.orc_rbf_side_eval.argtypes = [_d, _d, C.c_int, C.c_double, C.c_double, C.c_double, _d]
        _LIB = L
    return _LIB


def ptr(a, t=_d):
    if a is None:
        return C.cast(None, t)
    return a.ctypes.data_as(t)


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox(ptr(c, _u32), ptr(k, _u32), ptr(o, _u32))
    return o


def math_eval(fn, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib().orc_math_eval({"exp": 0, "log": 1, "sin": 2, "cos": 3, "tanh": 4, "sinpi": 5, "cospi": 6,
                         "expm1": 7, "exp_tab": 8}[fn], ptr(x), ptr(y), x.size)
    return y


def rbf_side(px, py, x, y, ic2):
    """The reduction's Gaussian RBF values of one rectangle side's entries at X (recurrence form)."""
    px = np.ascontiguousarray(px, dtype=np.float64)
    py = np.ascontiguousarray(py, dtype=np.float64)
    out = np.zeros(len(px))
    lib().orc_rbf_side_eval(ptr(px), ptr(py), len(px), x, y, ic2, ptr(out))
    return out


def rect_dist(p, x, y):
    p = np.asarray(p, dtype=np.float64)
    o = np.zeros(4)
    d = lib().orc_rect_dist(ptr(p), x, y, ptr(o[0:]), ptr(o[1:]), ptr(o[2:]), ptr(o[3:]))
    return d, o[0], o[1], o[2], o[3]


def sector_dist(g, x, y):
    """(d, ex, ey, nx, ny, kind) for the annular sector g = (r0, r1, cos a, sin a, cos b, sin b)."""
    o = np.zeros(4)
    k = np.zeros(1, dtype=np.int32)
    d = lib().orc_sector_dist(ptr(np.ascontiguousarray(g, dtype=np.float64)), x, y, ptr(o[0:1]), ptr(o[1:2]),
                              ptr(o[2:3]), ptr(o[3:4]), ptr(k, _i))
    return (d, o[0], o[1], o[2], o[3], int(k[0]))


def problem(pde, k2=0.0, lut=None, step=0, refl=0, robin=0.0, rho=0.0, ulut=None, vlut=None):
    """step: 0 control variate, 1 plain, 2 VR, 3 VR+CV (the engine's pdds_run.step values); refl:
    reflecting sides of the stopping rectangle (bits N, S, E, W), robin: varphi there, rho: Lepingle
    band; ulut: u of PDE_TABLE, vlut: mu and F of the VR+CV step (lut_dict dictionaries)."""
    pb = Problem()
    pb.pde = pde
    pb.k2 = k2
    pb.step = step
    pb.refl, pb.robin, pb.rho = refl, robin, rho
    pb._keep = []
    if lut is not None:
        pb._keep.append(lut["coef"])
        pb.lut = ptr(lut["coef"], _f)
        pb.lut_nx, pb.lut_ny, pb.lut_nf = lut["NX"], lut["NY"], lut["nf"]
        pb.lut_ox, pb.lut_oy, pb.lut_sx, pb.lut_sy = lut["ox"], lut["oy"], lut["sx"], lut["sy"]
    for name, d in (("ulut", ulut), ("vlut", vlut)):
        if d is not None:
            pb._keep.append(d["coef"])
            setattr(pb, name, Lut(ptr(d["coef"], _f), d["NX"], d["NY"], d["nf"], d["ox"], d["oy"], d["sx"], d["sy"]))
    return pb


def lut_dict(samples, box, mode=0):
    """Oracle LUT of samples [ny][nx][nf] on box (x0, y0, x1, y1), with the engine's coordinate map:
    mode 0 prefiltered node-aligned samples (interpolating), mode 1 raw texel-centred samples."""
    s = np.asarray(samples, dtype=np.float64)
    if s.ndim == 2:
        s = s[:, :, None]
    ny, nx, nf = s.shape
    x0, y0, x1, y1 = box
    if mode == 0:
        dx, dy = (x1 - x0) / (nx - 1), (y1 - y0) / (ny - 1)
        return {"coef": lut_prefilter(s), "NX": nx + 2, "NY": ny + 2, "nf": nf, "ox": x0 - dx, "oy": y0 - dy,
                "sx": 1.0 / dx, "sy": 1.0 / dy}
    sx, sy = nx / (x1 - x0), ny / (y1 - y0)
    return {"coef": np.ascontiguousarray(s, dtype=np.float32), "NX": nx, "NY": ny, "nf": nf,
            "ox": x0 + 0.5 / sx, "oy": y0 + 0.5 / sy, "sx": sx, "sy": sy}


def run_params(h, seed, beta=-0.5826, stream=0, normals=0, phase2=0, max_steps=0, threads=0):
    r = Run()
    r.h, r.beta, r.seed, r.stream = h, beta, seed, stream
    r.normals, r.phase2, r.max_steps, r.threads = normals, phase2, max_steps, threads
    return r


def pde_eval(pb, x, y):
    o = np.zeros(12)
    lib().orc_pde_eval(C.byref(pb), x, y, ptr(o[0:]), ptr(o[4:]), ptr(o[6:]), ptr(o[7:]), ptr(o[8:]),
                       ptr(o[10:]), ptr(o[11:]))
    return dict(sigma=o[0:4].copy(), b=o[4:6].copy(), c=o[6], f=o[7], grad=o[8:10].copy(), g=o[10], u=o[11])


def normals(run, knot, path, step0, nsteps):
    n1 = np.zeros(nsteps)
    n2 = np.zeros(nsteps)
    lib().orc_normals(C.byref(run), knot, path, step0, nsteps, ptr(n1), ptr(n2))
    return n1, n2


class Mesh:
    """Oracle restatement of PDDSparseGM's interfaces, knots and stencils."""

    def __init__(self, sw, ne, n_sub, n_node):
        self.h = lib().orc_mesh_create(sw[0], sw[1], ne[0], ne[1], n_sub, n_node)
        if not self.h:
            raise ValueError("invalid mesh")
        self.n_knots = lib().orc_mesh_n_knots(self.h)
        self.n_node = n_node
        self.glob = np.array([sw[0], sw[1], ne[0], ne[1]], dtype=np.float64)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_destroy(self.h)
            self.h = None

    def knots(self):
        n = self.n_knots
        x0, y0, nx, ny = (np.zeros(n) for _ in range(4))
        onb = np.zeros(n, dtype=np.int32)
        lib().orc_mesh_knots(self.h, ptr(x0), ptr(y0), ptr(nx), ptr(ny), ptr(onb, _i))
        return x0, y0, nx, ny, onb.astype(bool)

    def stencil(self, k):
        cap = 4 * (4 * (self.n_node + 2) + 8)
        rect = np.zeros(4)
        msz = np.zeros(4, dtype=np.int32)
        kind = np.zeros(4, dtype=np.int32)
        idx = np.zeros(cap, dtype=np.int32)
        px = np.zeros(cap)
        py = np.zeros(cap)
        tot = lib().orc_mesh_stencil(self.h, k, ptr(rect), ptr(msz, _i), ptr(kind, _i), ptr(idx, _i),
                                     ptr(px), ptr(py))
        return dict(rect=rect, msz=msz, kind=kind, idx=idx[:tot].copy(), px=px[:tot].copy(),
                    py=py[:tot].copy())


def stencil_ipsi(px, py, c):
    m = len(px)
    px = np.ascontiguousarray(px, dtype=np.float64)
    py = np.ascontiguousarray(py, dtype=np.float64)
    ip = np.zeros(m * m)
    cond = np.zeros(1)
    st = lib().orc_stencil_ipsi(m, ptr(px), ptr(py), c, ptr(ip), ptr(cond))
    return st, ip.reshape(m, m), cond[0]


def simulate(pb, run, x0, y0, rect, knot, path0, npaths, phase2_out=False, geom=0):
    """Phase-1 paths (phase 2 as well on rectangles); geom 1: rect is an annular sector (6 values)."""
    rect = np.ascontiguousarray(rect, dtype=np.float64)
    if geom:
        o = {k: np.zeros(npaths) for k in ("xt", "yt", "Y", "Z", "xi")}
        o["steps"] = np.zeros(npaths, dtype=np.int64)
        lib().orc_simulate_geom(C.byref(pb), C.byref(run), x0, y0, geom, ptr(rect), knot, path0, npaths,
                                ptr(o["xt"]), ptr(o["yt"]), ptr(o["Y"]), ptr(o["Z"]), ptr(o["xi"]),
                                ptr(o["steps"], _i64))
        return o
    o = {k: np.zeros(npaths) for k in ("xt", "yt", "Y", "Z", "xi")}
    o["steps"] = np.zeros(npaths, dtype=np.int64)
    p2 = {k: (np.zeros(npaths) if phase2_out else None) for k in ("xt", "yt", "Y", "Z", "xi")}
    s2 = np.zeros(npaths, dtype=np.int64) if phase2_out else None
    lib().orc_simulate(C.byref(pb), C.byref(run), x0, y0, ptr(rect), knot, path0, npaths,
                       ptr(o["xt"]), ptr(o["yt"]), ptr(o["Y"]), ptr(o["Z"]), ptr(o["xi"]), ptr(o["steps"], _i64),
                       ptr(p2["xt"]), ptr(p2["yt"]), ptr(p2["Y"]), ptr(p2["Z"]), ptr(p2["xi"]), ptr(s2, _i64))
    if phase2_out:
        o["phase2"] = dict(p2, steps=s2)
    return o


def _reduce(fn, pb, paths, st, glob, cshape, ipsi_list, quirk, S=None, P=None, w=None, geom=None):
    msz = np.ascontiguousarray(st["msz"], dtype=np.int32)
    idx = np.ascontiguousarray(st["idx"], dtype=np.int32)
    tot = int(msz.sum())
    ipsi = np.concatenate([np.asarray(a, dtype=np.float64).ravel() for a in ipsi_list]) if tot else np.zeros(1)
    cols = np.zeros(tot + 1, dtype=np.int32)
    G = np.zeros(tot + 1)
    Gv = np.zeros(tot + 1)
    B = np.zeros(1)
    BB = np.zeros(1)
    cnt = np.zeros(6, dtype=np.int64)
    cs = np.ascontiguousarray(cshape, dtype=np.float64)
    n = len(paths["Y"])
    args = [C.byref(pb), n]
    if S is not None:
        args += [S, P if P else n]
    args += [ptr(paths["xt"]), ptr(paths["yt"]), ptr(paths["Y"]), ptr(paths["Z"]), ptr(paths["xi"])]
    dom = st["domain"] if geom else st["rect"]
    if geom is not None:
        args += [int(geom)]
    args += [ptr(np.ascontiguousarray(dom, dtype=np.float64)), ptr(np.ascontiguousarray(glob, dtype=np.float64)),
             ptr(msz, _i), ptr(idx, _i), ptr(np.ascontiguousarray(st["px"])), ptr(np.ascontiguousarray(st["py"])),
             ptr(cs), ptr(ipsi), quirk]
    eq = np.zeros(2)
    if w is not None:
        wv = np.zeros(tot + 1)
        wv[:len(w)] = w
        args += [ptr(wv)]
    if geom is not None:
        cent = st.get("cent")
        args += [ptr(None) if cent is None or len(cent) == 0 else ptr(np.ascontiguousarray(cent, dtype=np.float64))]
    args += [ptr(cols, _i), ptr(G), ptr(Gv), ptr(B), ptr(BB), ptr(cnt, _i64)]
    if w is not None:
        args += [ptr(eq[0:1]), ptr(eq[1:2])]
    nnz = fn(*args)
    out = dict(cols=cols[:nnz].copy(), G=G[:nnz].copy(), Gvar=Gv[:nnz].copy(), B=B[0], BB=BB[0], cnt=cnt)
    if w is not None:
        out["eq_mean"], out["eq_var"] = float(eq[0]), float(eq[1])
    return out


def reduce_faithful(pb, paths, st, glob, cshape, ipsi_list, quirk=1):
    return _reduce(lib().orc_reduce_faithful, pb, paths, st, glob, cshape, ipsi_list, quirk)


def reduce_batched(pb, paths, st, glob, cshape, ipsi_list, S, P=None, quirk=1):
    """Batch-means estimator with the engine's summation order (sub-batches of S in items of P)."""
    return _reduce(lib().orc_reduce_batched, pb, paths, st, glob, cshape, ipsi_list, quirk, S=S, P=P)


def reduce_batched_eq(pb, paths, st, glob, cshape, ipsi_list, S, P, u_ref, quirk=1, geom=0):
    """reduce_batched plus the knot equation's mean and batch-means per-path variance with
    reference interface values u_ref (indexed by knot), summed in the engine's order.  geom 1:
    st["domain"] is an annular sector of a disc of radius glob[2]."""
    w = None
    if u_ref is not None:
        cols = _reduce(lib().orc_reduce_batched_geom, pb, paths, st, glob, cshape, ipsi_list, quirk, S=S, P=P,
                       w=np.zeros(int(np.sum(st["msz"]))), geom=geom)["cols"]
        w = np.asarray(u_ref, dtype=np.float64)[cols]
    else:
        w = np.zeros(int(np.sum(st["msz"])))
    return _reduce(lib().orc_reduce_batched_geom, pb, paths, st, glob, cshape, ipsi_list, quirk, S=S, P=P, w=w,
                   geom=geom)


def bicgstab(rowptr, cols, vals, b, tol=1e-14, maxit=None):
    n = len(b)
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(n)
    rr = np.zeros(1)
    it = lib().orc_bicgstab(n, ptr(rowptr, _i64), ptr(cols, _i), ptr(vals), ptr(b), ptr(x), tol,
                            maxit if maxit else 2 * n, ptr(rr))
    return x, it, rr[0]


def lut_prefilter(samples):
    """samples: (ny, nx, nf) float64 node-aligned -> (ny+2, nx+2, nf) float32 coefficients."""
    s = np.ascontiguousarray(samples, dtype=np.float64)
    ny, nx, nf = s.shape
    coef = np.zeros((ny + 2, nx + 2, nf), dtype=np.float32)
    lib().orc_lut_prefilter(nx, ny, nf, ptr(s), ptr(coef, _f))
    return coef
