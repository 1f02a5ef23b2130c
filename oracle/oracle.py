"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never by the product.
See oracle.h for the reference lines each function follows and its parity
status.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}
_BUILD = "parity"  # "parity": -O2 -ffp-contract=off (the checker); "speed": -O3 -march=x86-64-v3 (CPU timings)

ROE, HLLC, AUSMP = 0, 1, 2
LIM_NONE, LIM_VENKAT, LIM_BJ = 0, 1, 2
GRAD_LSQ, GRAD_GG = 0, 1
BC_EULER_WALL, BC_FAR_FIELD = 0, 1


class OrcMesh(ctypes.Structure):
    _fields_ = [("n_node", ctypes.c_int32), ("n_edge", ctypes.c_int32), ("n_bedge", ctypes.c_int32),
                ("edge", ctypes.c_void_p), ("edge_area", ctypes.c_void_p), ("bedge", ctypes.c_void_p),
                ("bedge_area", ctypes.c_void_p), ("bedge_type", ctypes.c_void_p), ("xyz", ctypes.c_void_p),
                ("vol", ctypes.c_void_p), ("n2n_ia", ctypes.c_void_p), ("n2n_ja", ctypes.c_void_p)]


class OrcParams(ctypes.Structure):
    _fields_ = [("gamma", ctypes.c_double), ("r_gas", ctypes.c_double), ("inf_rho", ctypes.c_double),
                ("inf_u", ctypes.c_double), ("inf_v", ctypes.c_double), ("inf_w", ctypes.c_double),
                ("inf_p", ctypes.c_double), ("outflow_p", ctypes.c_double), ("flux", ctypes.c_int32),
                ("order", ctypes.c_int32), ("limiter", ctypes.c_int32), ("limiter_order", ctypes.c_int32),
                ("entropy_fix", ctypes.c_int32), ("zpg_iter", ctypes.c_int32), ("venkat_k", ctypes.c_double),
                ("cfl_min", ctypes.c_double), ("cfl_max", ctypes.c_double), ("cfl_ramp", ctypes.c_int32)]


class OrcSolveOpts(ctypes.Structure):
    _fields_ = [("max_iter", ctypes.c_int32), ("tol", ctypes.c_double), ("krylov_m", ctypes.c_int32),
                ("max_restarts", ctypes.c_int32), ("krylov_rtol", ctypes.c_double), ("precond", ctypes.c_int32),
                ("gradient", ctypes.c_int32), ("first_order_iters", ctypes.c_int32), ("relaxation", ctypes.c_double)]


class OrcSolveResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("krylov_total", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("final_rms", ctypes.c_double), ("first_rms", ctypes.c_double), ("force", ctypes.c_double * 3),
                ("lift", ctypes.c_double), ("drag", ctypes.c_double), ("pressure_integral", ctypes.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "oracle"], check=True)


def use_build(kind):
    """Selects the library the functions below call: "parity" (the checker, bit-exact build) or "speed" (the same
    source built for speed, used only to time the CPU path).  Returns the previous choice."""
    global _BUILD
    if kind not in ("parity", "speed"):
        raise ValueError(kind)
    old, _BUILD = _BUILD, kind
    return old


def lib():
    if _BUILD not in _LIBS:
        path = os.path.join(_HERE, "liboracle.so" if _BUILD == "parity" else "liboracle_speed.so")
        if not os.path.exists(path):
            build()
        _LIBS[_BUILD] = ctypes.CDLL(path)
    return _LIBS[_BUILD]


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Mesh:
    """Holds numpy arrays alive and exposes the orc_mesh view."""

    def __init__(self, xyz, edge, edge_area, bedge, bedge_area, bedge_type, vol, n2n_ia, n2n_ja, bedge_tag=None):
        c = np.ascontiguousarray
        self.xyz = c(xyz, dtype=np.float64).ravel()
        self.edge = c(edge, dtype=np.int32).ravel()
        self.edge_area = c(edge_area, dtype=np.float64).ravel()
        self.bedge = c(bedge, dtype=np.int32).ravel()
        self.bedge_area = c(bedge_area, dtype=np.float64).ravel()
        self.bedge_type = c(bedge_type, dtype=np.int32).ravel()
        self.bedge_tag = None if bedge_tag is None else c(bedge_tag, dtype=np.int32).ravel()
        self.vol = c(vol, dtype=np.float64).ravel()
        self.n2n_ia = c(n2n_ia, dtype=np.int32).ravel()
        self.n2n_ja = c(n2n_ja, dtype=np.int32).ravel()
        self.n_node = self.vol.size
        self.n_edge = self.edge.size // 2
        self.n_bedge = self.bedge.size // 2
        self.s = OrcMesh(self.n_node, self.n_edge, self.n_bedge, _p(self.edge), _p(self.edge_area), _p(self.bedge),
                         _p(self.bedge_area), _p(self.bedge_type), _p(self.xyz), _p(self.vol), _p(self.n2n_ia),
                         _p(self.n2n_ja))

    @property
    def n_total(self):
        return self.n_node + self.n_bedge


def params(inf, flux=ROE, order=1, limiter=LIM_NONE, limiter_order=1, entropy_fix=0, venkat_k=5.0,
           cfl_min=1.0, cfl_max=1000.0, cfl_ramp=50, gamma=1.4, zpg_iter=0):
    """inf: (rho, u, v, w, p, ..., outflow_p at index 10) as dumped by the reference driver, or a dict."""
    if isinstance(inf, dict):
        rho, u, v, w, p, outp = inf["rho"], inf["u"], inf["v"], inf["w"], inf["p"], inf.get("outflow_p", 0.935)
    else:
        rho, u, v, w, p, outp = inf[0], inf[1], inf[2], inf[3], inf[4], inf[10]
    return OrcParams(gamma, 1.0 / gamma, rho, u, v, w, p, outp, flux, order, limiter, limiter_order, entropy_fix,
                     zpg_iter, venkat_k, cfl_min, cfl_max, cfl_ramp)


def residual(mesh, prm, q, grad=None, phi=None):
    n = mesh.n_node
    rc = np.zeros(5 * n)
    rd = np.zeros(5 * n)
    dt = np.zeros(n)
    eig = np.zeros(6)
    q = np.ascontiguousarray(q, dtype=np.float64).ravel()
    g = None if grad is None else np.ascontiguousarray(grad, dtype=np.float64).ravel()
    f = None if phi is None else np.ascontiguousarray(phi, dtype=np.float64).ravel()
    lib().orc_residual(ctypes.byref(mesh.s), ctypes.byref(prm), _p(q), _p(g), _p(f), _p(rc), _p(rd), _p(dt), _p(eig))
    return rc.reshape(5, n), rd.reshape(5, n), dt, eig


def edge_flux(prm, s, ql, qr):
    conv = np.zeros(5)
    diss = np.zeros(5)
    s = np.ascontiguousarray(s, dtype=np.float64)
    ql = np.ascontiguousarray(ql, dtype=np.float64)
    qr = np.ascontiguousarray(qr, dtype=np.float64)
    f = lib().orc_edge_flux
    f.restype = ctypes.c_double
    lam = f(ctypes.byref(prm), _p(s), _p(ql), _p(qr), _p(conv), _p(diss))
    return conv, diss, lam


def bc(mesh, prm, it, q):
    q = np.array(q, dtype=np.float64).ravel()
    bad = lib().orc_bc(ctypes.byref(mesh.s), ctypes.byref(prm), int(it), _p(q))
    return q.reshape(5, -1), bad


def lsq_gradient(mesh, q):
    g = np.zeros(15 * mesh.n_node)
    q = np.ascontiguousarray(q, dtype=np.float64).ravel()
    lib().orc_lsq_gradient(ctypes.byref(mesh.s), _p(q), _p(g))
    return g.reshape(15, mesh.n_node)


def gg_gradient(mesh, q):
    g = np.zeros(15 * mesh.n_node)
    q = np.ascontiguousarray(q, dtype=np.float64).ravel()
    lib().orc_gg_gradient(ctypes.byref(mesh.s), _p(q), _p(g))
    return g.reshape(15, mesh.n_node)


def limiter(mesh, prm, q, grad):
    phi = np.zeros(5 * mesh.n_node)
    q = np.ascontiguousarray(q, dtype=np.float64).ravel()
    grad = np.ascontiguousarray(grad, dtype=np.float64).ravel()
    lib().orc_limiter(ctypes.byref(mesh.s), ctypes.byref(prm), _p(q), _p(grad), _p(phi))
    return phi.reshape(5, mesh.n_node)


def delta_t(mesh, prm, it, dt_acc):
    dt = np.zeros(mesh.n_node)
    f = lib().orc_delta_t
    f.restype = ctypes.c_double
    cfl = f(ctypes.byref(mesh.s), ctypes.byref(prm), int(it), _p(np.ascontiguousarray(dt_acc, dtype=np.float64)), _p(dt))
    return dt, cfl


def bsr_structure(mesh):
    nnz = lib().orc_bsr_nnz(ctypes.byref(mesh.s))
    ia = np.zeros(mesh.n_node + 1, np.int32)
    ja = np.zeros(nnz, np.int32)
    iau = np.zeros(mesh.n_node, np.int32)
    lib().orc_bsr_structure(ctypes.byref(mesh.s), _p(ia), _p(ja), _p(iau))
    return ia, ja, iau


def jacobian(mesh, prm, q, dt, res_conv, res_diss, rhs_mode, ia, ja, iau):
    a = np.zeros(25 * int(ia[-1]))
    b = np.zeros(5 * mesh.n_node)
    c = lambda x: np.ascontiguousarray(x, dtype=np.float64).ravel()
    lib().orc_jacobian(ctypes.byref(mesh.s), ctypes.byref(prm), _p(c(q)), _p(c(dt)), _p(c(res_conv)),
                       _p(None if res_diss is None else c(res_diss)), int(rhs_mode), _p(ia), _p(ja), _p(iau),
                       _p(a), _p(b))
    return a.reshape(-1, 25), b.reshape(-1, 5)


FD_CENTRAL, FD_FORWARD, FD_BACKWARD, FD_ALTERNATE = 0, 1, 2, 3


def jacobian_fd(mesh, prm, q, grad, phi, dt, res_conv, res_diss, rhs_mode, method, it, add_time, ia, ja, iau):
    """Finite-difference Roe Jacobian (orc_jacobian_fd); returns (A blocks, B, negative ghost pressures)."""
    a = np.zeros(25 * int(ia[-1]))
    b = np.zeros(5 * mesh.n_node)
    c = lambda x: None if x is None else np.ascontiguousarray(x, dtype=np.float64).ravel()
    f = lib().orc_jacobian_fd
    f.restype = ctypes.c_int
    bad = f(ctypes.byref(mesh.s), ctypes.byref(prm), _p(c(q)), _p(c(grad)), _p(c(phi)), _p(c(dt)), _p(c(res_conv)),
            _p(c(res_diss)), int(rhs_mode), int(method), int(it), int(add_time), _p(ia), _p(ja), _p(iau), _p(a),
            _p(b))
    return a.reshape(-1, 25), b.reshape(-1, 5), bad


def edge_flux_matrix(prm, s, ql, qr):
    """Matrix-form Roe flux of one edge (Compute_RoeFlux): (conv, diss)."""
    conv, diss = np.zeros(5), np.zeros(5)
    c = lambda x: np.ascontiguousarray(x, dtype=np.float64)
    lib().orc_edge_flux_matrix(ctypes.byref(prm), _p(c(s)), _p(c(ql)), _p(c(qr)), _p(conv), _p(diss))
    return conv, diss


def block_gs(ia, ja, iau, a, b, sweeps):
    n = ia.size - 1
    x = np.zeros(5 * n)
    f = lib().orc_block_gs
    f.restype = ctypes.c_double
    rms = f(n, _p(ia), _p(ja), _p(iau), _p(np.ascontiguousarray(a).ravel()), _p(np.ascontiguousarray(b).ravel()),
            int(sweeps), _p(x))
    return x.reshape(n, 5), rms


def lusgs(ia, ja, iau, a, b, sweeps):
    n = ia.size - 1
    x = np.zeros(5 * n)
    lib().orc_lusgs(n, _p(ia), _p(ja), _p(iau), _p(np.ascontiguousarray(a).ravel()),
                    _p(np.ascontiguousarray(b).ravel()), int(sweeps), _p(x))
    return x.reshape(n, 5)


def ilu0(ia, ja, iau, a):
    n = ia.size - 1
    lu = np.array(a, dtype=np.float64).ravel()
    dinv = np.zeros(25 * n)
    rc = lib().orc_ilu0(n, _p(ia), _p(ja), _p(iau), _p(lu), _p(dinv))
    return lu.reshape(-1, 25), dinv.reshape(-1, 25), rc


def ilu0_apply(ia, ja, iau, lu, dinv, r):
    n = ia.size - 1
    z = np.zeros(5 * n)
    lib().orc_ilu0_apply(n, _p(ia), _p(ja), _p(iau), _p(np.ascontiguousarray(lu).ravel()),
                         _p(np.ascontiguousarray(dinv).ravel()), _p(np.ascontiguousarray(r, dtype=np.float64).ravel()),
                         _p(z))
    return z.reshape(n, 5)


def bsr_mv(ia, ja, a, x):
    n = ia.size - 1
    y = np.zeros(5 * n)
    lib().orc_bsr_mv(n, _p(ia), _p(ja), _p(np.ascontiguousarray(a).ravel()),
                     _p(np.ascontiguousarray(x, dtype=np.float64).ravel()), _p(y))
    return y.reshape(n, 5)


def fgmres(ia, ja, iau, a, b, m=30, max_restarts=4, rtol=1e-3, precond=2):
    n = ia.size - 1
    x = np.zeros(5 * n)
    fr = ctypes.c_double()
    its = lib().orc_fgmres(n, _p(ia), _p(ja), _p(iau), _p(np.ascontiguousarray(a).ravel()),
                           _p(np.ascontiguousarray(b, dtype=np.float64).ravel()), int(m), int(max_restarts),
                           ctypes.c_double(rtol), int(precond), _p(x), ctypes.byref(fr))
    return x.reshape(n, 5), its, fr.value


def solve_steady(mesh, prm, q, bedge_tag, alpha_deg=0.0, wall_tag=3, max_iter=200, tol=1e-10, krylov_m=30,
                 max_restarts=4, krylov_rtol=1e-3, precond=2, gradient=GRAD_LSQ, first_order_iters=0,
                 relaxation=1.0):
    o = OrcSolveOpts(max_iter, tol, krylov_m, max_restarts, krylov_rtol, precond, gradient, first_order_iters,
                     relaxation)
    r = OrcSolveResult()
    q = np.array(q, dtype=np.float64).ravel()
    hist = np.zeros(max_iter)
    khist = np.zeros(max_iter, np.int32)
    tag = np.ascontiguousarray(bedge_tag, dtype=np.int32)
    st = lib().orc_solve_steady(ctypes.byref(mesh.s), ctypes.byref(prm), ctypes.byref(o), ctypes.c_double(alpha_deg),
                                int(wall_tag), _p(tag), _p(q), _p(hist), _p(khist), ctypes.byref(r))
    return st, r, q.reshape(5, -1), hist[: max(r.iterations + 1, 1)], khist


def forces(mesh, prm, q, bedge_tag, wall_tag=3, alpha_deg=0.0):
    F = (ctypes.c_double * 3)()
    lift, drag, pint = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().orc_forces(ctypes.byref(mesh.s), ctypes.byref(prm), _p(np.ascontiguousarray(q, dtype=np.float64).ravel()),
                     _p(np.ascontiguousarray(bedge_tag, dtype=np.int32)), int(wall_tag), ctypes.c_double(alpha_deg),
                     F, ctypes.byref(lift), ctypes.byref(drag), ctypes.byref(pint))
    return np.array(F[:]), lift.value, drag.value, pint.value
