"""Python binding of the C ABI in include/cfd_hip.h (libcfd_hip.so).

The compute path is the HIP library; this module only marshals arrays.  It
raises if the library is not built: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

from . import _build
from . import mesh as _mesh

OK = 0
E_ARG, E_HIP, E_OOM, E_NONFINITE, E_NEGATIVE_PRESSURE, E_SINGULAR, E_HLLC, E_STATE = -1, -2, -3, -4, -5, -6, -7, -8
FLUX_ROE, FLUX_HLLC, FLUX_AUSMP = 0, 1, 2
LIM_NONE, LIM_VENKAT, LIM_BJ = 0, 1, 2
GRAD_LSQ, GRAD_GG = 0, 1
FP64, FP32 = 0, 1
ARITH_EXACT, ARITH_FAST = 0, 1
SOLVE_JAC_APPROX, SOLVE_JAC_FD = 0, 1
PREC_NONE, PREC_BJACOBI, PREC_ILU0, PREC_ILU0_MC = 0, 1, 2, 3
# mixed= of the linear solves: False/0 FP64; True/1 FP32 storage; MIXED_FP16 as 1 with the multicolour factors'
# block-SELL copies in FP16; MIXED_FP16_ALL as MIXED_FP16 with the matrix's block-SELL copy in FP16 too
MIXED_FP16, MIXED_FP16_ALL = 2, 3
SOLVER_FGMRES, SOLVER_BLOCK_GS, SOLVER_LUSGS = 0, 1, 2
JAC_FD_CENTRAL, JAC_FD_FORWARD, JAC_FD_BACKWARD, JAC_FD_ALTERNATE = 0, 1, 2, 3
FIELD_Q, FIELD_GRAD, FIELD_PHI, FIELD_RES_CONV, FIELD_RES_DISS, FIELD_DT_ACC, FIELD_DT, FIELD_JAC_B, FIELD_X = range(9)

_FLUX_NAMES = {"roe": FLUX_ROE, "hllc": FLUX_HLLC, "ausm+": FLUX_AUSMP, "ausmp": FLUX_AUSMP}


class CfdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("cfd error %d: %s" % (code, msg))
        self.code = code


class MeshDesc(ctypes.Structure):
    _fields_ = [("n_node", ctypes.c_int32), ("n_edge", ctypes.c_int32), ("n_bedge", ctypes.c_int32),
                ("edge", ctypes.c_void_p), ("edge_area", ctypes.c_void_p), ("bedge", ctypes.c_void_p),
                ("bedge_area", ctypes.c_void_p), ("bedge_type", ctypes.c_void_p), ("bedge_tag", ctypes.c_void_p),
                ("xyz", ctypes.c_void_p), ("vol", ctypes.c_void_p), ("n2n_ia", ctypes.c_void_p),
                ("n2n_ja", ctypes.c_void_p)]


class Params(ctypes.Structure):
    _fields_ = [("gamma", ctypes.c_double), ("inf_rho", ctypes.c_double), ("inf_u", ctypes.c_double),
                ("inf_v", ctypes.c_double), ("inf_w", ctypes.c_double), ("inf_p", ctypes.c_double),
                ("outflow_p", ctypes.c_double), ("flux", ctypes.c_int32), ("order", ctypes.c_int32),
                ("limiter", ctypes.c_int32), ("limiter_order", ctypes.c_int32), ("entropy_fix", ctypes.c_int32),
                ("gradient", ctypes.c_int32), ("zpg_iter", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("venkat_k", ctypes.c_double), ("cfl_min", ctypes.c_double), ("cfl_max", ctypes.c_double),
                ("cfl_ramp", ctypes.c_int32), ("arith", ctypes.c_int32)]


class LinsolveOpts(ctypes.Structure):
    _fields_ = [("solver", ctypes.c_int32), ("precond", ctypes.c_int32), ("krylov_m", ctypes.c_int32),
                ("max_restarts", ctypes.c_int32), ("rtol", ctypes.c_double), ("sweeps", ctypes.c_int32),
                ("mixed", ctypes.c_int32)]


class LinsolveStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("restarts", ctypes.c_int32), ("final_res", ctypes.c_double),
                ("b_norm", ctypes.c_double)]


class SolveOpts(ctypes.Structure):
    _fields_ = [("max_iter", ctypes.c_int32), ("tol", ctypes.c_double), ("first_order_iters", ctypes.c_int32),
                ("relaxation", ctypes.c_double), ("wall_tag", ctypes.c_int32), ("alpha_deg", ctypes.c_double),
                ("lin", LinsolveOpts), ("jacobian", ctypes.c_int32), ("autotune", ctypes.c_int32)]


class SolveResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("krylov_total", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("status", ctypes.c_int32), ("first_rms", ctypes.c_double), ("final_rms", ctypes.c_double),
                ("force", ctypes.c_double * 3), ("lift", ctypes.c_double), ("drag", ctypes.c_double),
                ("pressure_integral", ctypes.c_double), ("seconds", ctypes.c_double)]


class TuneReport(ctypes.Structure):
    _fields_ = [("gl_ms", ctypes.c_double * 2), ("krylov_ms", (ctypes.c_double * 2) * 2), ("gl_stage", ctypes.c_int32),
                ("sell", ctypes.c_int32), ("sync_free", ctypes.c_int32), ("pad_", ctypes.c_int32)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("max_inflight", ctypes.c_int32), ("lookahead", ctypes.c_int32), ("max_bypass", ctypes.c_int32),
                ("keep_resident", ctypes.c_int32), ("hbm_budget", ctypes.c_int64), ("batch_max", ctypes.c_int32),
                ("batch_nodes", ctypes.c_int32), ("batch_wait_us", ctypes.c_int32),
                ("batch_case_nodes", ctypes.c_int32)]


class Request(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int64), ("mesh_key", ctypes.c_int64), ("mesh", ctypes.POINTER(MeshDesc)),
                ("params", Params), ("opts", SolveOpts), ("q_out", ctypes.c_void_p), ("rms_hist", ctypes.c_void_p)]


class Completion(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int64), ("status", ctypes.c_int32), ("reused", ctypes.c_int32),
                ("batch_size", ctypes.c_int32), ("pad_", ctypes.c_int32), ("result", SolveResult),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double), ("footprint", ctypes.c_int64), ("device_bytes", ctypes.c_int64),
                ("message", ctypes.c_char * 160)]


class ServerStats(ctypes.Structure):
    _fields_ = [("bytes_in_use", ctypes.c_int64), ("peak_bytes", ctypes.c_int64), ("queued", ctypes.c_int32),
                ("running", ctypes.c_int32), ("resident", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("admitted", ctypes.c_int64), ("reused", ctypes.c_int64), ("evicted", ctypes.c_int64),
                ("clock", ctypes.c_double)]


def solve_opts(max_iter=200, tol=1e-10, first_order_iters=0, relaxation=1.0, wall_tag=3, alpha_deg=0.0,
               precond=PREC_ILU0, krylov_m=30, max_restarts=4, rtol=1e-3, mixed=False, solver=SOLVER_FGMRES,
               sweeps=0, jacobian=0, autotune=False):
    lin = LinsolveOpts(solver, precond, krylov_m, max_restarts, rtol, sweeps, int(mixed))
    return SolveOpts(max_iter, tol, first_order_iters, relaxation, wall_tag, alpha_deg, lin, int(jacobian),
                     int(bool(autotune)))


def mesh_desc(mesh, bedge_type=None):
    """Builds a MeshDesc for a mesh; returns (desc, arrays) -- keep the arrays alive while desc is in use."""
    c = np.ascontiguousarray
    keep = dict(
        edge=c(mesh.edge, np.int32), edge_area=c(mesh.edge_area, np.float64), bedge=c(mesh.bedge, np.int32),
        bedge_area=c(mesh.bedge_area, np.float64),
        bedge_type=c(mesh.bedge_type if bedge_type is None else bedge_type, np.int32),
        bedge_tag=c(mesh.bedge_tag, np.int32), xyz=c(mesh.xyz, np.float64), vol=c(mesh.vol, np.float64),
        n2n_ia=c(mesh.n2n_ia, np.int32), n2n_ja=c(mesh.n2n_ja, np.int32))
    k = keep
    d = MeshDesc(k["vol"].size, k["edge"].size // 2, k["bedge"].size // 2, _p(k["edge"]), _p(k["edge_area"]),
                 _p(k["bedge"]), _p(k["bedge_area"]), _p(k["bedge_type"]), _p(k["bedge_tag"]), _p(k["xyz"]),
                 _p(k["vol"]), _p(k["n2n_ia"]), _p(k["n2n_ja"]))
    return d, keep


class _PrepCounts(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("n_node", "n_tri", "n_tet", "n_edge", "n_bedge", "n_n2n")]


class PreprocessedMesh:
    """A mesh preprocessed on the GPU by cfd_preprocess: the arrays of mesh.Mesh, bit-identical to the host
    library's, from the UGRID content (xyz, triangles in UGRID winding, their tags, tetrahedra)."""

    def __init__(self, ctx, content):
        c = np.ascontiguousarray
        xyz, tri = c(content.xyz, np.float64), c(content.tri, np.int32)
        tag, tet = c(content.tri_tag, np.int32), c(content.tet, np.int32)
        h = ctypes.c_void_p()
        ctx.check(lib().cfd_preprocess(ctx.h, int(xyz.shape[0]), _p(xyz), int(tri.shape[0]), _p(tri), _p(tag),
                                       int(tet.shape[0]), _p(tet), ctypes.byref(h)))
        try:
            k = _PrepCounts()
            ctx.check(lib().cfd_prep_info(h, ctypes.byref(k)))
            n, e, b = k.n_node, k.n_edge, k.n_bedge
            self.n_node, self.n_edge, self.n_bedge, self.n_tet, self.n_tri = n, e, b, k.n_tet, k.n_tri
            self.xyz, self.tet, self.tri_tag = xyz.copy(), tet.copy(), tag.copy()
            self.tri = np.zeros((k.n_tri, 3), np.int32)
            self.edge = np.zeros((e, 2), np.int32)
            self.edge_area = np.zeros((e, 3))
            self.bedge = np.zeros((b, 2), np.int32)
            self.bedge_area = np.zeros((b, 3))
            self.bedge_tag = np.zeros(b, np.int32)
            self.vol = np.zeros(n)
            self.n2n_ia = np.zeros(n + 1, np.int32)
            self.n2n_ja = np.zeros(k.n_n2n, np.int32)
            ctx.check(lib().cfd_prep_get(h, _p(self.tri), _p(self.edge), _p(self.edge_area), _p(self.bedge),
                                         _p(self.bedge_area), _p(self.bedge_tag), _p(self.vol), _p(self.n2n_ia),
                                         _p(self.n2n_ja)))
        finally:
            lib().cfd_prep_destroy(h)
        # the bump channel's boundary conditions, as mesh.Mesh
        self.bedge_type = np.where(self.bedge_tag <= 2, _mesh.BC_FAR_FIELD, _mesh.BC_EULER_WALL).astype(np.int32)

    @property
    def n_total(self):
        return self.n_node + self.n_bedge


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = _build.lib_path("libcfd_hip.so")
        if not os.path.exists(path):
            raise RuntimeError("libcfd_hip.so is not built (run __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(path)
        L.cfd_last_error.restype = ctypes.c_char_p
        L.cfd_mesh_device_bytes.restype = ctypes.c_int64
        L.cfd_footprint_bytes.restype = ctypes.c_int64
        L.cfd_footprint_bytes.argtypes = [ctypes.c_int32] * 5
        L.cfd_mesh_stream.restype = ctypes.c_void_p
        L.cfd_bsr_structure.restype = ctypes.c_int32
        L.cfd_server_poll.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_double]
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def make_params(mach=0.5, alpha_deg=0.0, gamma=1.4, flux="roe", order=1, limiter=LIM_NONE, limiter_order=1,
                entropy_fix=0, gradient=GRAD_LSQ, precision=FP64, venkat_k=5.0, cfl_min=1.0, cfl_max=1000.0,
                cfl_ramp=50, zpg_iter=0, inf=None, arith=0):
    """Freestream of the GENERIC non-dimensionalisation (Euler3D/src/Material.cpp:112-210):
    rho = 1, T = gamma, R = 1/gamma, p = R*rho*T, c = sqrt(gamma*R*T)."""
    if inf is None:
        r = 1.0 / gamma
        rho, T = 1.0, gamma
        p = r * rho * T
        c = np.sqrt(gamma * r * (T + 0.0))
        a = alpha_deg * np.pi / 180.0
        u = mach * c * np.cos(a) * np.cos(0.0)
        v = mach * c * np.sin(a)
        w = mach * c * np.sin(0.0) * np.cos(a)
        outp = 0.935 * p
    else:
        rho, u, v, w, p, outp = inf[0], inf[1], inf[2], inf[3], inf[4], inf[10]
    f = _FLUX_NAMES[flux] if isinstance(flux, str) else flux
    return Params(gamma, rho, u, v, w, p, outp, f, order, limiter, limiter_order, entropy_fix, gradient, zpg_iter,
                  precision, venkat_k, cfl_min, cfl_max, cfl_ramp, arith)


class Context:
    def __init__(self, device=0):
        self.h = ctypes.c_void_p()
        rc = lib().cfd_ctx_create(int(device), ctypes.byref(self.h))
        if rc:
            raise CfdError(rc, "cfd_ctx_create failed (no GPU?)")

    def check(self, rc):
        if rc:
            raise CfdError(rc, lib().cfd_last_error(self.h).decode())
        return rc

    def close(self):
        if self.h:
            lib().cfd_ctx_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceMesh:
    """A mesh resident in HBM with its state, residuals and Jacobian workspace."""

    def __init__(self, ctx, mesh, params, bedge_type=None):
        self.ctx = ctx
        d, self._keep = mesh_desc(mesh, bedge_type)
        self.N, self.E, self.B = d.n_node, d.n_edge, d.n_bedge
        self.h = ctypes.c_void_p()
        ctx.check(lib().cfd_mesh_create(ctx.h, ctypes.byref(d), ctypes.byref(params), ctypes.byref(self.h)))

    @property
    def NT(self):
        return self.N + self.B

    def close(self):
        if self.h:
            lib().cfd_mesh_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_bytes(self):
        return lib().cfd_mesh_device_bytes(self.h)

    def stream(self):
        return lib().cfd_mesh_stream(self.h)

    def sync(self):
        self.ctx.check(lib().cfd_mesh_sync(self.h))

    _SHAPES = {FIELD_Q: lambda s: (5, s.NT), FIELD_GRAD: lambda s: (15, s.N), FIELD_PHI: lambda s: (5, s.N),
               FIELD_RES_CONV: lambda s: (5, s.N), FIELD_RES_DISS: lambda s: (5, s.N), FIELD_DT_ACC: lambda s: (s.N,),
               FIELD_DT: lambda s: (s.N,), FIELD_JAC_B: lambda s: (s.N, 5), FIELD_X: lambda s: (s.N, 5)}

    def set(self, field, arr):
        shape = self._SHAPES[field](self)
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(shape)
        self.ctx.check(lib().cfd_field_set(self.h, field, _p(a)))

    def get(self, field):
        out = np.zeros(self._SHAPES[field](self))
        self.ctx.check(lib().cfd_field_get(self.h, field, _p(out)))
        return out

    # hot path ------------------------------------------------------------
    def bc(self, p, it):
        self.ctx.check(lib().cfd_bc(self.h, ctypes.byref(p), int(it)))

    def gradient(self, p):
        self.ctx.check(lib().cfd_gradient(self.h, ctypes.byref(p)))

    def limiter(self, p):
        self.ctx.check(lib().cfd_limiter(self.h, ctypes.byref(p)))

    def gradient_limiter(self, p):
        self.ctx.check(lib().cfd_gradient_limiter(self.h, ctypes.byref(p)))

    def residual(self, p, stats=False):
        eig = np.zeros(6) if stats else None
        self.ctx.check(lib().cfd_residual(self.h, ctypes.byref(p), 1, _p(eig)))
        return eig

    def delta_t(self, p, it):
        cfl = ctypes.c_double()
        self.ctx.check(lib().cfd_delta_t(self.h, ctypes.byref(p), int(it), ctypes.byref(cfl)))
        return cfl.value

    def rms(self):
        r = ctypes.c_double()
        self.ctx.check(lib().cfd_residual_rms(self.h, ctypes.byref(r)))
        return r.value

    def jacobian(self, p, rhs_mode=1, mixed=False):
        self.ctx.check(lib().cfd_jacobian(self.h, ctypes.byref(p), int(rhs_mode), int(mixed)))

    def jacobian_fd(self, p, method=0, it=0, add_time=True, rhs_mode=0, mixed=False):
        """Finite-difference Roe Jacobian (cfd_jacobian_fd); method JAC_FD_CENTRAL/FORWARD/BACKWARD/ALTERNATE."""
        self.ctx.check(lib().cfd_jacobian_fd(self.h, ctypes.byref(p), int(method), int(it), int(bool(add_time)),
                                             int(rhs_mode), int(mixed)))

    def bsr_structure(self):
        nnz = lib().cfd_bsr_structure(self.h, None, None, None)
        ia = np.zeros(self.N + 1, np.int32)
        ja = np.zeros(nnz, np.int32)
        iau = np.zeros(self.N, np.int32)
        lib().cfd_bsr_structure(self.h, _p(ia), _p(ja), _p(iau))
        return ia, ja, iau

    def jacobian_get(self):
        nnz = lib().cfd_bsr_structure(self.h, None, None, None)
        a = np.zeros((nnz, 25))
        b = np.zeros((self.N, 5))
        self.ctx.check(lib().cfd_jacobian_get(self.h, _p(a), _p(b)))
        return a, b

    def jacobian_set(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        self.ctx.check(lib().cfd_jacobian_set(self.h, _p(a), _p(b)))

    def linsolve(self, solver=SOLVER_FGMRES, precond=PREC_ILU0, krylov_m=30, max_restarts=4, rtol=1e-3, sweeps=0,
                 mixed=False):
        o = LinsolveOpts(solver, precond, krylov_m, max_restarts, rtol, sweeps, int(mixed))
        st = LinsolveStats()
        self.ctx.check(lib().cfd_linsolve(self.h, ctypes.byref(o), ctypes.byref(st)))
        return st

    def autotune(self, p, precond=PREC_ILU0, krylov_m=30, max_restarts=4, rtol=1e-3, mixed=False,
                 solver=SOLVER_FGMRES, sweeps=0):
        """cfd_autotune: times the same-bits kernel variants on this mesh (the Krylov ones on the assembled
        Jacobian, if any) and keeps the fastest.  Returns the TuneReport."""
        lin = LinsolveOpts(solver, precond, krylov_m, max_restarts, rtol, sweeps, int(mixed))
        rep = TuneReport()
        self.ctx.check(lib().cfd_autotune(self.h, ctypes.byref(p), ctypes.byref(lin), ctypes.byref(rep)))
        return rep

    def linsolve_gs_noalloc(self, sweeps):
        """cfd_linsolve_gs_noalloc: the block Gauss-Seidel solve without device allocation (meshes of at most
        32,768 nodes).  Returns (status, LinsolveStats)."""
        st = LinsolveStats()
        rc = lib().cfd_linsolve_gs_noalloc(self.h, int(sweeps), ctypes.byref(st))
        return rc, st

    def update(self, relax=1.0):
        self.ctx.check(lib().cfd_update(self.h, ctypes.c_double(relax)))

    def forces(self, p, wall_tag=3, alpha_deg=0.0):
        f = (ctypes.c_double * 3)()
        lift, drag, pint = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self.ctx.check(lib().cfd_forces(self.h, ctypes.byref(p), int(wall_tag), ctypes.c_double(alpha_deg), f,
                                        ctypes.byref(lift), ctypes.byref(drag), ctypes.byref(pint)))
        return np.array(f[:]), lift.value, drag.value, pint.value

    def reset(self, p):
        self.ctx.check(lib().cfd_mesh_reset(self.h, ctypes.byref(p)))

    def solve_steady(self, p, max_iter=200, tol=1e-10, first_order_iters=0, relaxation=1.0, wall_tag=3,
                     alpha_deg=0.0, precond=PREC_ILU0, krylov_m=30, max_restarts=4, rtol=1e-3, mixed=False,
                     solver=SOLVER_FGMRES, sweeps=0, jacobian=SOLVE_JAC_APPROX, autotune=False, check=True):
        o = solve_opts(max_iter, tol, first_order_iters, relaxation, wall_tag, alpha_deg, precond, krylov_m,
                       max_restarts, rtol, mixed, solver, sweeps, jacobian, autotune)
        r = SolveResult()
        hist = np.zeros(max_iter)
        rc = lib().cfd_solve_steady(self.h, ctypes.byref(p), ctypes.byref(o), ctypes.byref(r), _p(hist))
        if check:
            self.ctx.check(rc)
        return r, hist[: r.iterations + 1]


class Server:
    """Continuous-batching solve server (cfd_server_* in the C ABI).

    submit() queues a steady solve; completions arrive through poll().  The
    arrays of a request are kept alive here until its completion is polled."""

    def __init__(self, ctx, hbm_budget, max_inflight=4, lookahead=16, max_bypass=8, keep_resident=True, batch_max=0,
                 batch_nodes=0, batch_wait_us=0, batch_case_nodes=0):
        self.ctx = ctx
        self.h = ctypes.c_void_p()
        o = ServerOpts(int(max_inflight), int(lookahead), int(max_bypass), int(bool(keep_resident)), int(hbm_budget),
                       int(batch_max), int(batch_nodes), int(batch_wait_us), int(batch_case_nodes))
        ctx.check(lib().cfd_server_create(ctx.h, ctypes.byref(o), ctypes.byref(self.h)))
        self._live = {}
        self._descs = {}

    def submit(self, rid, mesh_key, mesh, params, want_state=False, **solve_kw):
        if mesh_key not in self._descs:
            self._descs[mesh_key] = mesh_desc(mesh)
        d, _ = self._descs[mesh_key]
        o = solve_opts(**solve_kw)
        q = np.zeros((5, d.n_node + d.n_bedge)) if want_state else None
        hist = np.zeros(o.max_iter)
        req = Request(int(rid), int(mesh_key), ctypes.pointer(d), params, o, _p(q), _p(hist))
        self.ctx.check(lib().cfd_server_submit(self.h, ctypes.byref(req)))
        self._live[int(rid)] = (req, q, hist)

    def poll(self, max_n=64, wait=0.0):
        buf = (Completion * max_n)()
        n = lib().cfd_server_poll(self.h, buf, int(max_n), float(wait))
        out = []
        for k in range(n):
            c = buf[k]
            _, q, hist = self._live.pop(int(c.id))
            out.append(dict(id=int(c.id), status=int(c.status), reused=bool(c.reused), batch_size=int(c.batch_size),
                            result=c.result,
                            t_submit=c.t_submit, t_admit=c.t_admit, t_done=c.t_done, footprint=int(c.footprint),
                            device_bytes=int(c.device_bytes), message=c.message.decode(errors="replace"), q=q,
                            hist=hist[: c.result.iterations + 1]))
        return out

    def stats(self):
        st = ServerStats()
        self.ctx.check(lib().cfd_server_get_stats(self.h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_ if f != "pad_"}

    def pending(self):
        return len(self._live)

    def close(self):
        if self.h:
            lib().cfd_server_destroy(self.h)
            self.h = ctypes.c_void_p()
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_solve(ctx, meshes, params, alphas=None, **solve_kw):
    """cfd_batch_solve: solves len(meshes) independent cases as one union mesh.  params is one Params per case
    (same scheme, own freestream); solve_kw are the shared solver options; alphas the per-case angles for lift and
    drag.  Returns (status, [SolveResult])."""
    K = len(meshes)
    keep = [mesh_desc(m) for m in meshes]
    descs = (MeshDesc * K)(*[d for d, _ in keep])
    ps = (Params * K)(*params)
    alphas = alphas if alphas is not None else [0.0] * K
    kw = dict(solve_kw)
    kw.pop("alpha_deg", None)
    os_ = (SolveOpts * K)(*[solve_opts(alpha_deg=a, **kw) for a in alphas])
    res = (SolveResult * K)()
    rc = lib().cfd_batch_solve(ctx.h, K, descs, ps, os_, res)
    return rc, [res[k] for k in range(K)]


def footprint_bytes(n_node, n_edge, n_bedge, krylov_m=30, mixed=True):
    return lib().cfd_footprint_bytes(int(n_node), int(n_edge), int(n_bedge), int(krylov_m), int(mixed))


def gpu_available():
    try:
        n = ctypes.c_int()
        hip = ctypes.CDLL("libamdhip64.so")
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False
