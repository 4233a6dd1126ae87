"""Python access to the CPU checkers: the C restatement (liboracle.so) and the
compiled reference harness (_ref/fvh_ref).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product package.

Solver configuration is a plain dict shared with the product's Python layer:
  flux   'roe_split' (shockTube/naca association) | 'roe_joined' (forwardStep)
  phi    'zero' | 'michalak' | 'venkat' | 'minmod'   (+ phi_yt for Michalak)
  dt     'fixed' | 'sum_cfl' | 'edge_min'           (+ dt_fixed, cfl)
  grad, lim, lin, gdt   the solver:: flags (explicit.h:35-40)
  K      limiter_k_value (explicit.cpp:27 fixes 7.5)
  gamma, alpha (list of RK stage coefficients)
  bcs    {physical name: (kind, rho, u, v, p)}, kind in
         'copy' | 'mirror' | 'inlet_outlet' | 'farfield'
  positivity   first-order face state where the reconstruction is not physical
  smooth_kinks (+ kink_width, kink_scale)   the smoothed-kink extension
"""
from __future__ import annotations

import ctypes as C
import os
import struct
import subprocess
from typing import Dict, List, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "fvh_ref")

FLUX = {"roe_split": 0, "roe_joined": 1, "roe_viscous": 2}
PHI = {"zero": 0, "michalak": 1, "venkat": 2, "minmod": 3}
DT = {"fixed": 0, "sum_cfl": 1, "edge_min": 2, "edge_min_viscous": 3, "sum_viscous": 4}
BC = {"copy": 0, "mirror": 1, "inlet_outlet": 2, "farfield": 3, "noslip": 4, "farfield_v": 5}
# consts:: of examples/flatPlate/main.cpp:28-35
VISCOUS_DEFAULTS = {"mu": 2e-5, "pr": 0.72, "cp": 1.0, "r_gas": 8.3145}
RK5 = [0.0695, 0.1602, 0.2898, 0.506, 1.0]


class OrcConfig(C.Structure):
    _fields_ = [
        ("flux", C.c_int), ("phi", C.c_int), ("phi_yt", C.c_double),
        ("dt_kind", C.c_int), ("dt_fixed", C.c_double), ("cfl", C.c_double),
        ("grad", C.c_int), ("lim", C.c_int), ("lin", C.c_int), ("gdt", C.c_int), ("pos", C.c_int),
        ("K", C.c_double), ("gamma", C.c_double),
        ("n_alpha", C.c_int), ("alpha", C.c_double * 8),
        ("n_bc", C.c_int), ("bc_kind", C.c_int * 16), ("bc_param", (C.c_double * 4) * 16),
        ("smooth", C.c_int),
        ("mu", C.c_double), ("pr", C.c_double), ("cp", C.c_double), ("r", C.c_double),
        ("diff", C.c_int),
        ("kinks", C.c_int), ("kw", C.c_double * 4), ("kwv", C.c_double), ("kwl", C.c_double),
    ]


def kink_widths(cfg: Dict):
    """Widths of the smoothed kinks (cfg 'smooth_kinks'), as the library forms
    them from kink_width (default 1e-2) and kink_scale (zeros: the first
    characteristic boundary's freestream rho, rho (|V| + c) twice, E, |V| + c;
    1 without one): ([w_rho, w_rhou, w_rhov, w_rhoE], w_speed, w_limiter)."""
    import math
    g = float(cfg.get("gamma", 1.4))
    sc = [1.0] * 5
    for kind, *par in cfg.get("bcs", {}).values():
        if kind not in ("inlet_outlet", "farfield", "farfield_v"):
            continue
        rho, u, v, p = (list(par) + [0.0] * 4)[:4]
        speed = math.sqrt(u * u + v * v)
        a = math.sqrt(g * p / rho)
        sc = [rho, rho * (speed + a), rho * (speed + a), p / (g - 1.0) + 0.5 * rho * speed * speed, speed + a]
        break
    w = float(cfg.get("kink_width", 0.0)) or 1e-2
    given = list(cfg.get("kink_scale", [0.0] * 5))
    use = [given[k] if given[k] > 0.0 else sc[k] for k in range(5)]
    return [w * use[k] for k in range(4)], w * use[4], w


class OrcMesh(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int32), ("x", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)),
        ("n_real", C.c_int32), ("n_cells", C.c_int32), ("n_edges", C.c_int32), ("n_bnd", C.c_int32),
        ("edge_nodes", C.POINTER(C.c_int32)), ("edge_cells", C.POINTER(C.c_int32)),
        ("edge_len", C.POINTER(C.c_double)), ("edge_nx", C.POINTER(C.c_double)),
        ("edge_ny", C.POINTER(C.c_double)), ("edge_cx", C.POINTER(C.c_double)),
        ("edge_cy", C.POINTER(C.c_double)),
        ("cell_nodes", C.POINTER(C.c_int32)), ("cell_tri", C.POINTER(C.c_uint8)),
        ("cell_area", C.POINTER(C.c_double)), ("cell_cx", C.POINTER(C.c_double)),
        ("cell_cy", C.POINTER(C.c_double)),
        ("bnd_edge", C.POINTER(C.c_int32)), ("bnd_bc", C.POINTER(C.c_int32)),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle oracle)")
        L = C.CDLL(LIB_PATH)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        L.orc_mesh_build.restype = C.POINTER(OrcMesh)
        L.orc_mesh_build.argtypes = [C.c_int32, dp, dp, C.c_int32, ip, C.c_int32, ip, ip,
                                     C.POINTER(C.c_int)]
        L.orc_mesh_free.argtypes = [C.POINTER(OrcMesh)]
        cfgp = C.POINTER(OrcConfig)
        mp = C.POINTER(OrcMesh)
        L.orc_calc_dt.argtypes = [mp, cfgp, dp, dp]
        L.orc_min_dt.argtypes = [mp, dp]
        L.orc_min_dt_soft.argtypes = [mp, dp]
        L.orc_sigma.argtypes = [cfgp, C.c_double]
        L.orc_sigma.restype = C.c_double
        L.orc_calc_gradients.argtypes = [mp, dp, dp, dp]
        L.orc_calc_limiters.argtypes = [mp, cfgp, dp, dp, dp, dp, dp, dp]
        L.orc_calc_time_derivatives.argtypes = [mp, cfgp, dp, dp, dp, dp, dp]
        L.orc_update_cells.argtypes = [C.c_int64, dp, dp, dp, dp, C.c_double]
        L.orc_update_bounds.argtypes = [mp, cfgp, dp, dp, dp, dp]
        L.orc_calc_residuals.argtypes = [mp, dp, dp]
        L.orc_smooth_residuals.argtypes = [mp, dp]
        L.orc_flux.argtypes = [cfgp, dp, dp, dp, dp]
        L.orc_flux_grad.argtypes = [cfgp, dp, dp, dp, dp, dp, dp]
        L.orc_face_gradient.argtypes = [dp, dp, dp, dp, dp, dp, dp, dp, dp, dp]
        L.orc_phi.argtypes = [cfgp, C.c_double]
        L.orc_phi.restype = C.c_double
        L.orc_bc.argtypes = [cfgp, C.c_int, dp, dp, dp]
        L.orc_run.argtypes = [mp, cfgp, dp, C.c_uint32, C.c_double, C.c_double, C.c_int32, C.c_uint32,
                              C.POINTER(C.c_uint32), C.POINTER(C.c_double), dp, dp, C.c_void_p]
        L.orc_run.restype = C.c_int
        _lib = L
    return _lib


def make_config(cfg: Dict, names: List[str]) -> OrcConfig:
    """OrcConfig from the shared dict; BC slots follow ``names`` order."""
    c = OrcConfig()
    c.flux = FLUX[cfg.get("flux", "roe_split")]
    c.phi = PHI[cfg.get("phi", "zero")]
    c.phi_yt = float(cfg.get("phi_yt", 2.0))
    c.dt_kind = DT[cfg.get("dt", "fixed")]
    c.dt_fixed = float(cfg.get("dt_fixed", 2e-5))
    c.cfl = float(cfg.get("cfl", 1.0))
    c.grad = int(bool(cfg.get("grad", False)))
    c.lim = int(bool(cfg.get("lim", False)))
    c.lin = int(bool(cfg.get("lin", False)))
    c.gdt = int(bool(cfg.get("gdt", False)))
    c.pos = int(bool(cfg.get("positivity", False)))
    c.smooth = int(bool(cfg.get("smooth", False)))
    c.diff = int(bool(cfg.get("diffusive", False)))
    c.mu = float(cfg.get("mu", VISCOUS_DEFAULTS["mu"]))
    c.pr = float(cfg.get("pr", VISCOUS_DEFAULTS["pr"]))
    c.cp = float(cfg.get("cp", VISCOUS_DEFAULTS["cp"]))
    c.r = float(cfg.get("r_gas", VISCOUS_DEFAULTS["r_gas"]))
    c.K = float(cfg.get("K", 7.5))
    c.gamma = float(cfg.get("gamma", 1.4))
    c.kinks = int(bool(cfg.get("smooth_kinks", False)))
    kw, kwv, kwl = kink_widths(cfg)
    for k in range(4):
        c.kw[k] = kw[k]
    c.kwv, c.kwl = kwv, kwl
    alpha = list(cfg.get("alpha", RK5))
    c.n_alpha = len(alpha)
    for i, a in enumerate(alpha):
        c.alpha[i] = a
    bcs = cfg["bcs"]
    c.n_bc = len(names)
    for s, nm in enumerate(names):
        if nm not in bcs:
            raise KeyError("Physical name %s not in the boundary table" % nm)
        kind, *par = bcs[nm]
        c.bc_kind[s] = BC[kind]
        par = list(par) + [0.0] * (4 - len(par))
        for k in range(4):
            c.bc_param[s][k] = float(par[k])
    return c


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Mesh:
    """Oracle mesh in the reference numbering, built from file-order arrays."""

    def __init__(self, m):
        self.arrays = m
        L = lib()
        err = C.c_int(0)
        x = np.ascontiguousarray(m.x, dtype=np.float64)
        y = np.ascontiguousarray(m.y, dtype=np.float64)
        cells = np.ascontiguousarray(m.cells, dtype=np.int32).ravel()
        bnd = np.ascontiguousarray(m.bnd, dtype=np.int32).ravel()
        tags = np.ascontiguousarray(m.bnd_tag, dtype=np.int32)
        p = L.orc_mesh_build(x.size, x, y, m.cells.shape[0], cells, m.bnd.shape[0], bnd, tags,
                             C.byref(err))
        if not p:
            raise ValueError("invalid edge ref: boundary element not on a cell edge")
        self._p = p
        s = p.contents
        self.n_real, self.n_cells, self.n_edges, self.n_bnd = s.n_real, s.n_cells, s.n_edges, s.n_bnd
        E, N = s.n_edges, s.n_cells
        self.edge_nodes = _arr(s.edge_nodes, 2 * E, np.int32).reshape(E, 2)
        self.edge_cells = _arr(s.edge_cells, 2 * E, np.int32).reshape(E, 2)
        self.edge_len = _arr(s.edge_len, E, np.float64)
        self.edge_nx = _arr(s.edge_nx, E, np.float64)
        self.edge_ny = _arr(s.edge_ny, E, np.float64)
        self.edge_cx = _arr(s.edge_cx, E, np.float64)
        self.edge_cy = _arr(s.edge_cy, E, np.float64)
        self.cell_nodes = _arr(s.cell_nodes, 4 * N, np.int32).reshape(N, 4)
        self.cell_tri = _arr(s.cell_tri, N, np.uint8)
        self.cell_area = _arr(s.cell_area, N, np.float64)
        self.cell_cx = _arr(s.cell_cx, N, np.float64)
        self.cell_cy = _arr(s.cell_cy, N, np.float64)
        self.bnd_edge = _arr(s.bnd_edge, s.n_bnd, np.int32)
        self.names = list(m.names)

    def __del__(self):
        try:
            lib().orc_mesh_free(self._p)
        except Exception:
            pass

    # -- phases ----------------------------------------------------------
    def phases(self, cfg: Dict, q: np.ndarray, alpha: float = 0.5) -> Dict[str, np.ndarray]:
        """One stage of every phase from state ``q`` (AoS, 4*n_cells):
        the 'phases' mode of the reference harness."""
        L = lib()
        c = make_config(cfg, self.names)
        q = np.ascontiguousarray(q, dtype=np.float64).copy()
        N = self.n_cells
        n = 4 * N
        out = {}
        dt = np.zeros(N)
        L.orc_calc_dt(self._p, C.byref(c), q, dt)
        out["dt_local"] = dt.copy()
        if c.gdt:
            L.orc_min_dt(self._p, dt)
        out["dt"] = dt
        gx, gy = np.zeros(n), np.zeros(n)
        if c.grad:
            L.orc_calc_gradients(self._p, q, gx, gy)
        out["gx"], out["gy"] = gx, gy
        lim, qmin, qmax = np.zeros(n), np.zeros(n), np.zeros(n)
        if c.lim:
            L.orc_calc_limiters(self._p, C.byref(c), q, gx, gy, lim, qmin, qmax)
        out["lim"], out["qmin"], out["qmax"] = lim, qmin, qmax
        qt = np.zeros(n)
        L.orc_calc_time_derivatives(self._p, C.byref(c), q, gx, gy, lim, qt)
        if c.smooth:
            # the update and the residuals take the smoothed qt (explicit.cpp:743)
            out["qt_raw"] = qt.copy()
            L.orc_smooth_residuals(self._p, qt)
        out["qt"] = qt
        qk = np.zeros(n)
        L.orc_update_cells(N, qk, q, qt, dt, float(alpha))
        out["qk_cells"] = qk.copy()
        L.orc_update_bounds(self._p, C.byref(c), qk, gx, gy, lim)
        out["qk"] = qk
        R = np.zeros(4)
        L.orc_calc_residuals(self._p, qt, R)
        out["R"] = R
        return out

    def run(self, cfg: Dict, q0: np.ndarray, max_step: int, max_time: float = 1e10,
            tol: float = 1e-16, print_interval: int = 1, stop_var: int = -1):
        """run() of explicit.cpp:657-817 from initial state q0 (all cells).
        stop_var=-1 is the reference's stop test; k >= 0 tests R[k] only.
        Returns (q, steps, time, R, R0, qt)."""
        L = lib()
        c = make_config(cfg, self.names)
        q = np.ascontiguousarray(q0, dtype=np.float64).copy()
        steps = C.c_uint32(0)
        t = C.c_double(0.0)
        R = np.zeros(4)
        R0 = np.zeros(4)
        qt = np.zeros_like(q)
        rc = L.orc_run(self._p, C.byref(c), q, int(max_step), float(max_time), float(tol),
                       int(stop_var), int(print_interval), C.byref(steps), C.byref(t), R, R0,
                       qt.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise MemoryError("orc_run failed")
        return q, int(steps.value), float(t.value), R, R0, qt


    def first_unphysical_step(self, cfg: Dict, q0: np.ndarray, max_step: int):
        """run() of explicit.cpp:657-817 stage by stage (the same phases as
        orc_run), returning the first step in which a stage update leaves a
        real cell with rho <= 0, p <= 0 or a non-finite value (what the
        library's failure detection flags), or None within max_step steps."""
        L = lib()
        c = make_config(cfg, self.names)
        N = self.n_cells
        n = 4 * N
        q = np.ascontiguousarray(q0, dtype=np.float64).copy()
        gx, gy, lim, qmin, qmax, qt = (np.zeros(n) for _ in range(6))
        dt, qk = np.zeros(N), np.zeros(n)
        L.orc_update_bounds(self._p, C.byref(c), q, gx, gy, lim)
        for step in range(max_step):
            L.orc_calc_dt(self._p, C.byref(c), q, dt)
            if c.gdt:
                L.orc_min_dt(self._p, dt)
            qk[:] = q
            for a in list(cfg.get("alpha", RK5)):
                if c.grad:
                    L.orc_calc_gradients(self._p, qk, gx, gy)
                if c.lim:
                    L.orc_calc_limiters(self._p, C.byref(c), qk, gx, gy, lim, qmin, qmax)
                L.orc_calc_time_derivatives(self._p, C.byref(c), qk, gx, gy, lim, qt)
                if c.smooth:
                    L.orc_smooth_residuals(self._p, qt)
                L.orc_update_cells(N, qk, q, qt, dt, float(a))
                r = qk.reshape(-1, 4)[:self.n_real]
                with np.errstate(all="ignore"):
                    ok = (r[:, 0] > 0) & (2 * r[:, 0] * r[:, 3] > r[:, 1] ** 2 + r[:, 2] ** 2)
                if not ok.all():
                    return step
                L.orc_update_bounds(self._p, C.byref(c), qk, gx, gy, lim)
            q[:] = qk
        return None

    def run_f32_state(self, cfg: Dict, q0: np.ndarray, max_step: int):
        """run() of explicit.cpp:657-817 with the cell states stored in FP32:
        every state that run() writes (the initial state, the initial
        boundary update, each stage's updated cells and ghost cells) is
        rounded to the nearest float, and everything else -- gradients,
        limiters, fluxes, time steps, the update itself -- is computed in
        FP64 from the rounded states, as orc_run computes it.  The library's
        FP32-state, FP64-arithmetic mode must give these bits.  Returns the
        state after max_step steps (no stop test)."""
        L = lib()
        c = make_config(cfg, self.names)
        N = self.n_cells
        n = 4 * N
        f32 = lambda a: a.astype(np.float32).astype(np.float64)
        q = f32(np.ascontiguousarray(q0, dtype=np.float64))
        gx, gy, lim, qmin, qmax, qt = (np.zeros(n) for _ in range(6))
        dt, qk = np.zeros(N), np.zeros(n)
        L.orc_update_bounds(self._p, C.byref(c), q, gx, gy, lim)
        q = f32(q)
        for _ in range(max_step):
            L.orc_calc_dt(self._p, C.byref(c), q, dt)
            if c.gdt:
                L.orc_min_dt(self._p, dt)
            qk[:] = q
            for a in list(cfg.get("alpha", RK5)):
                if c.grad:
                    L.orc_calc_gradients(self._p, qk, gx, gy)
                if c.lim:
                    L.orc_calc_limiters(self._p, C.byref(c), qk, gx, gy, lim, qmin, qmax)
                L.orc_calc_time_derivatives(self._p, C.byref(c), qk, gx, gy, lim, qt)
                if c.smooth:
                    L.orc_smooth_residuals(self._p, qt)
                L.orc_update_cells(N, qk, q, qt, dt, float(a))
                L.orc_update_bounds(self._p, C.byref(c), qk, gx, gy, lim)
                qk[:] = f32(qk)
            q[:] = qk
        return q


def flux(cfg: Dict, qi, qj, n) -> np.ndarray:
    c = make_config(dict(cfg, bcs={}), [])
    f = np.zeros(4)
    lib().orc_flux(C.byref(c), f, np.ascontiguousarray(qi, dtype=np.float64),
                   np.ascontiguousarray(qj, dtype=np.float64), np.ascontiguousarray(n, dtype=np.float64))
    return f


def flux_grad(cfg: Dict, qi, qj, gx, gy, n) -> np.ndarray:
    """Edge flux with a face gradient (the viscous kind uses it)."""
    c = make_config(dict(cfg, bcs={}), [])
    f = np.zeros(4)
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (qi, qj, gx, gy, n)]
    lib().orc_flux_grad(C.byref(c), f, *a)
    return f


def face_gradient(gxi, gyi, gxj, gyj, qi, qj, ci, cj):
    """gradient_for_diffusion (explicit.cpp:70-109): (gx, gy) of the face."""
    gx, gy = np.zeros(4), np.zeros(4)
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (gxi, gyi, gxj, gyj, qi, qj, ci, cj)]
    lib().orc_face_gradient(gx, gy, *a)
    return gx, gy


def phi(cfg: Dict, r: float) -> float:
    c = make_config(dict(cfg, bcs={}), [])
    return float(lib().orc_phi(C.byref(c), float(r)))


def bc(kind_tuple, gamma: float, qint, n) -> np.ndarray:
    c = make_config({"gamma": gamma, "bcs": {"b": kind_tuple}}, ["b"])
    b = np.zeros(4)
    lib().orc_bc(C.byref(c), 0, b, np.ascontiguousarray(qint, dtype=np.float64),
                 np.ascontiguousarray(n, dtype=np.float64))
    return b


# ---------------------------------------------------------------------------
# Compiled reference harness (_ref/fvh_ref)
# ---------------------------------------------------------------------------

def read_dump(path: str) -> Dict[str, np.ndarray]:
    out = {}
    with open(path, "rb") as f:
        data = f.read()
    p = 0
    while p < len(data):
        (ln,) = struct.unpack_from("<I", data, p)
        p += 4
        name = data[p:p + ln].decode()
        p += ln
        typ = chr(data[p])
        p += 1
        (n,) = struct.unpack_from("<Q", data, p)
        p += 8
        if typ == "d":
            out[name] = np.frombuffer(data, dtype="<f8", count=n, offset=p).copy()
            p += 8 * n
        else:
            out[name] = np.frombuffer(data, dtype="<u4", count=n, offset=p).copy()
            p += 4 * n
    return out


_cst = None


def _cstep():
    global _cst
    if _cst is None:
        path = os.path.join(HERE, "libcstep.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/libcstep.so is not built (run make -C oracle oracle)")
        L = C.CDLL(path)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        L.cst_run_f64.argtypes = [C.POINTER(OrcMesh), C.POINTER(OrcConfig), dp, C.c_int, C.c_void_p, C.c_void_p,
                                  dp, C.POINTER(C.c_double)]
        L.cst_derivative.argtypes = [C.POINTER(OrcMesh), C.POINTER(OrcConfig), dp, dp, C.c_double, C.c_int,
                                     C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.cst_derivative2.argtypes = [C.POINTER(OrcMesh), C.POINTER(OrcConfig), dp, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_double, C.c_int, C.c_void_p, C.c_void_p,
                                      C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.cst_geometry_f64.argtypes = [C.POINTER(OrcMesh), dp]
        L.cst_geometry_jvp.argtypes = [C.POINTER(OrcMesh), dp, dp, C.c_double, dp]
        _cst = L
    return _cst


def _wptr(w):
    if w is None:
        return None, None
    a = np.ascontiguousarray(w, dtype=np.float64)
    return a, a.ctypes.data_as(C.c_void_p)


def cstep_run(mesh: "Mesh", cfg: Dict, q0: np.ndarray, steps: int, wx=None, wy=None):
    """The templated restatement (fvh_cstep.cpp) in double: (q after `steps`
    steps of run(), J = sum_b p(ghost_b)(wx n_x + wy n_y) len_b)."""
    c = make_config(cfg, mesh.names)
    q0 = np.ascontiguousarray(q0, dtype=np.float64)
    q = np.zeros_like(q0)
    J = C.c_double(0.0)
    ax, px = _wptr(wx)
    ay, py = _wptr(wy)
    _cstep().cst_run_f64(mesh._p, C.byref(c), q0, int(steps), px, py, q, C.byref(J))
    return q, float(J.value)


def cstep_derivative(mesh: "Mesh", cfg: Dict, q0: np.ndarray, direction: np.ndarray, steps: int, wx=None,
                     wy=None, h: float = 1e-30):
    """Complex-step derivative of J along `direction`: (J, dJ)."""
    c = make_config(cfg, mesh.names)
    q0 = np.ascontiguousarray(q0, dtype=np.float64)
    d = np.ascontiguousarray(direction, dtype=np.float64)
    J, dJ = C.c_double(0.0), C.c_double(0.0)
    ax, px = _wptr(wx)
    ay, py = _wptr(wy)
    _cstep().cst_derivative(mesh._p, C.byref(c), q0, d, float(h), int(steps), px, py, C.byref(J), C.byref(dJ))
    return float(J.value), float(dJ.value)


def cstep_derivative_nodes(mesh: "Mesh", cfg: Dict, q0: np.ndarray, steps: int, dx=None, dy=None,
                           direction=None, wx=None, wy=None, h: float = 1e-30):
    """Complex-step derivative of J along a perturbation of the node
    coordinates (dx, dy: one value per node, reference node order) and,
    optionally, of the initial state: (J, dJ)."""
    c = make_config(cfg, mesh.names)
    q0 = np.ascontiguousarray(q0, dtype=np.float64)
    ad, pd = _wptr(direction)
    ax_, pdx = _wptr(dx)
    ay_, pdy = _wptr(dy)
    J, dJ = C.c_double(0.0), C.c_double(0.0)
    ax, px = _wptr(wx)
    ay, py = _wptr(wy)
    _cstep().cst_derivative2(mesh._p, C.byref(c), q0, pd, pdx, pdy, float(h), int(steps), px, py, C.byref(J),
                             C.byref(dJ))
    return float(J.value), float(dJ.value)


def cstep_geometry(mesh: "Mesh") -> Dict[str, np.ndarray]:
    """Geometry of the templated restatement in double (cells: cx, cy, area;
    edges: cx, cy, nx, ny, len)."""
    p = mesh._p.contents
    NC, E = p.n_cells, p.n_edges
    out = np.zeros(3 * NC + 5 * E)
    _cstep().cst_geometry_f64(mesh._p, out)
    keys = ["cell_cx", "cell_cy", "cell_area"]
    res = {k: out[i * NC:(i + 1) * NC] for i, k in enumerate(keys)}
    for i, k in enumerate(["edge_cx", "edge_cy", "edge_nx", "edge_ny", "edge_len"]):
        res[k] = out[3 * NC + i * E:3 * NC + (i + 1) * E]
    return res


def cstep_geometry_jvp(mesh: "Mesh", dx: np.ndarray, dy: np.ndarray, h: float = 1e-30) -> np.ndarray:
    """Complex-step derivative of the geometry along node displacements:
    per edge (n_x, n_y, len, Green-Gauss weight, centre x, y), then per
    cell, ghosts included, (area, centre x, y), packed."""
    p = mesh._p.contents
    out = np.zeros(6 * p.n_edges + 3 * p.n_cells)
    _cstep().cst_geometry_jvp(mesh._p, np.ascontiguousarray(dx, dtype=np.float64),
                              np.ascontiguousarray(dy, dtype=np.float64), float(h), out)
    return out


def ref_available() -> bool:
    return os.path.exists(REF_BIN)


def run_ref(mode: str, workdir: str, mesh_name: str, flags: str = "", timeout: float = 600,
            tag: str = "", **kv) -> Dict:
    """Run the reference harness in ``workdir`` (which holds mesh_name.msh).
    Returns the dump dict plus 'stdout'.  Runs that may overlap on the same
    mesh need distinct tags: the dump file is named after mesh, mode and tag."""
    out = os.path.join(workdir, mesh_name + "." + mode + ("." + tag if tag else "") + ".dump")
    args = [REF_BIN, mode, mesh_name, out] + ["%s=%s" % (k, v) for k, v in kv.items()]
    env = dict(os.environ)
    env["FVH_REF_FLAGS"] = flags
    r = subprocess.run(args, cwd=workdir, env=env, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError("fvh_ref failed (%d): %s %s" % (r.returncode, r.stdout[-2000:], r.stderr[-2000:]))
    d = read_dump(out)
    d["stdout"] = r.stdout
    return d


MPIRUN = "/opt/conda/bin/mpirun"


def mpi_available() -> bool:
    return os.path.exists(MPIRUN) and os.path.exists(REF_BIN)


def run_ref_mpi(mode: str, workdir: str, mesh_name: str, ranks: int, flags: str = "", timeout: float = 900,
                **kv) -> Dict:
    """The reference harness on `ranks` MPI ranks, each reading
    <mesh_name>_<rank+1>.msh from ``workdir``.  Returns {'stdout', 'ranks':
    [per-rank dump dicts]} (rank 0 prints the step rows)."""
    out = os.path.join(workdir, mesh_name + "." + mode + ".dump")
    args = [MPIRUN, "-n", str(ranks), REF_BIN, mode, mesh_name, out] + ["%s=%s" % (k, v) for k, v in kv.items()]
    env = dict(os.environ)
    env["FVH_REF_FLAGS"] = flags
    r = subprocess.run(args, cwd=workdir, env=env, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError("mpirun fvh_ref failed (%d): %s %s" % (r.returncode, r.stdout[-2000:], r.stderr[-2000:]))
    return {"stdout": r.stdout, "ranks": [read_dump(out + "." + str(k)) for k in range(ranks)]}


def gather_ranks(res: Dict, n_bnd: int, n_cells: int) -> np.ndarray:
    """Each rank's own (non-ghost) cells of q, placed in the single-file cell
    order (element tag - 1 - n_bnd); (n_cells, 4)."""
    q = np.full((n_cells, 4), np.nan)
    for d in res["ranks"]:
        own = d["cellsIsGhost"] == 0
        cells = d["cellsOriginal"].astype(np.int64)[own] - n_bnd
        q[cells] = d["q"].reshape(-1, 4)[:own.size][own]
    return q


def ref_args(cfg: Dict) -> Dict[str, str]:
    """Map a config dict onto the harness's plugin selection.  Only the
    combinations the reference examples provide are expressible."""
    kv = {}
    kv["flux"] = {"roe_split": "st", "roe_joined": "fs", "roe_viscous": "fp"}[cfg.get("flux", "roe_split")]
    kv["dt"] = {"fixed": "st", "sum_cfl": "fs", "edge_min": "nc", "edge_min_viscous": "fp",
                "sum_viscous": "nt"}[cfg.get("dt", "fixed")]
    for k, v in VISCOUS_DEFAULTS.items():
        if float(cfg.get(k, v)) != v:
            raise ValueError("the reference's viscous examples fix %s = %r" % (k, v))
    phi_ = cfg.get("phi", "zero")
    kv["phi"] = {"zero": "zero", "michalak": "fs", "venkat": "venk", "minmod": "minmod"}[phi_]
    if phi_ == "michalak" and float(cfg.get("phi_yt", 2.0)) != 2.0:
        raise ValueError("the reference's Michalak limiter fixes y_t = 2")
    if cfg.get("dt", "fixed") == "fixed" and float(cfg.get("dt_fixed", 2e-5)) != 2e-5:
        raise ValueError("the reference's fixed dt is 2e-5")
    if float(cfg.get("K", 7.5)) != 7.5:
        raise ValueError("the reference fixes K = 7.5")
    kv["gamma"] = repr(float(cfg.get("gamma", 1.4)))
    kv["cfl_fs"] = repr(float(cfg.get("cfl", 1.5)))
    kv["cfl_nc"] = repr(float(cfg.get("cfl", 1.0)))
    kv["cfl_fp"] = repr(float(cfg.get("cfl", 2.0)))
    kv["cfl_nt"] = repr(float(cfg.get("cfl", 1.5)))
    return kv


def ref_flags(cfg: Dict) -> str:
    names = {"diffusive": "diff"}
    return ",".join(names.get(k, k) for k in ("grad", "lim", "lin", "gdt", "smooth", "diffusive") if cfg.get(k, False))
