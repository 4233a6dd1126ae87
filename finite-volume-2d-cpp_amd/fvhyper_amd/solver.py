"""Python face of libfvh: Mesh and Solver over the C ABI.

Mirrors fvhyper's operator surface for the explicit path: the mesh object
(fvhyper/mesh.h:136-206 members, exported in reference order), the
solverOptions/run pair (fvhyper/explicit.h:152-181) and the per-phase
functions (explicit.h:88-129, here one stage of all of them at once).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, List, Optional

import numpy as np

from . import _lib
from ._lib import check, dptr, iptr

_INT_FIELDS = {"edgesNodes", "edgesCells", "cellsNodes", "cellsIsTriangle", "boundaryEdges",
               "boundaryTags", "rcmPermutation", "tilePermutation"}


class Mesh:
    """A mesh in the reference numbering (real cells, then boundary ghosts)."""

    def __init__(self, handle):
        self._h = handle
        L = _lib.load()
        c = (C.c_int64 * 6)()
        check(L.fvh_mesh_counts(self._h, c))
        self.n_nodes, self.n_real, self.n_cells, self.n_edges, self.n_bnd, n_names = list(c)
        buf = C.create_string_buffer(256)
        self.names: List[str] = []
        for k in range(n_names):
            check(L.fvh_mesh_name(self._h, k, buf, 256))
            self.names.append(buf.value.decode())

    @classmethod
    def read_msh(cls, path: str) -> "Mesh":
        h = C.c_void_p()
        check(_lib.load().fvh_mesh_read_msh(path.encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def read_su2(cls, path: str) -> "Mesh":
        """A 2-D SU2 ASCII mesh (fvh_mesh_read_su2); marker tags become the
        physical names."""
        h = C.c_void_p()
        check(_lib.load().fvh_mesh_read_su2(path.encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, m) -> "Mesh":
        """From a meshgen.MeshArrays (file-order arrays)."""
        L = _lib.load()
        x = np.ascontiguousarray(m.x, dtype=np.float64)
        y = np.ascontiguousarray(m.y, dtype=np.float64)
        cells = np.ascontiguousarray(m.cells, dtype=np.int32)
        bnd = np.ascontiguousarray(m.bnd, dtype=np.int32).reshape(-1, 2)
        tags = np.ascontiguousarray(m.bnd_tag, dtype=np.int32)
        names = (C.c_char_p * max(1, len(m.names)))(*[n.encode() for n in m.names])
        h = C.c_void_p()
        check(L.fvh_mesh_from_arrays(x.size, dptr(x), dptr(y), cells.shape[0], iptr(cells), bnd.shape[0],
                                     iptr(bnd), iptr(tags), len(m.names), names, C.byref(h)))
        return cls(h)

    def export(self, field: str) -> np.ndarray:
        L = _lib.load()
        n = check(L.fvh_mesh_export(self._h, field.encode(), None, 0))
        dt = np.int32 if field in _INT_FIELDS else np.float64
        out = np.zeros(n, dtype=dt)
        check(L.fvh_mesh_export(self._h, field.encode(), out.ctypes.data_as(C.c_void_p), n))
        return out

    def __del__(self):
        try:
            if self._h:
                _lib.load().fvh_mesh_free(self._h)
                self._h = None
        except Exception:
            pass


ORDERS = {"file": 0, "rcm": 1, "tiles": 2}
# f64: FP64 throughout (bit-identical to the reference); f32: FP32 states,
# geometry and arithmetic; f32_f64: FP32 states, FP64 geometry and arithmetic
PRECISION = {"f64": 0, "f32": 1, "f32_f64": 2}
FLUX_MODES = {"auto": 0, "cell": 1, "tile": 2, "fused": 3}


def make_config(cfg: Dict, precision: str = "f64", renumber="tiles", flux_mode: str = "auto",
                tile_cells: int = 0) -> _lib.Config:
    """fvh_config from the shared configuration dict (see oracle/oracle.py)."""
    c = _lib.Config()
    _lib.load().fvh_config_default(C.byref(c))
    c.precision = PRECISION[precision]
    c.flux = _lib.FLUX[cfg.get("flux", "roe_split")]
    c.gamma = float(cfg.get("gamma", 1.4))
    c.phi = _lib.PHI[cfg.get("phi", "zero")]
    c.phi_yt = float(cfg.get("phi_yt", 2.0))
    c.K = float(cfg.get("K", 7.5))
    c.dt_kind = _lib.DT[cfg.get("dt", "fixed")]
    c.dt_fixed = float(cfg.get("dt_fixed", 2e-5))
    c.cfl = float(cfg.get("cfl", 1.0))
    flags = 0
    if cfg.get("grad", False):
        flags |= _lib.FLAG_GRADIENTS
    if cfg.get("lim", False):
        flags |= _lib.FLAG_LIMITERS
    if cfg.get("lin", False):
        flags |= _lib.FLAG_LINEAR
    if cfg.get("gdt", False):
        flags |= _lib.FLAG_GLOBAL_DT
    if cfg.get("positivity", False):
        flags |= _lib.FLAG_POSITIVITY
    if cfg.get("smooth", False):
        flags |= _lib.FLAG_SMOOTH
    if cfg.get("diffusive", False):
        flags |= _lib.FLAG_DIFFUSIVE
    if cfg.get("smooth_kinks", False):
        flags |= _lib.FLAG_SMOOTH_KINKS
    c.flags = flags
    c.kink_width = float(cfg.get("kink_width", 0.0))
    for i, v in enumerate(cfg.get("kink_scale", [0.0] * 5)):
        c.kink_scale[i] = float(v)
    for k in ("mu", "pr", "cp", "r_gas"):
        if k in cfg:
            setattr(c, k, float(cfg[k]))
    c.stage_threads = int(cfg.get("stage_threads", 0))
    c.stage_launch = int(cfg.get("stage_launch", 0))
    alpha = list(cfg.get("alpha", _lib.RK5))
    if not 1 <= len(alpha) <= _lib.FVH_MAX_ALPHA:
        raise ValueError("1..%d RK stages supported" % _lib.FVH_MAX_ALPHA)
    c.n_alpha = len(alpha)
    for i, a in enumerate(alpha):
        c.alpha[i] = float(a)
    bcs = cfg.get("bcs", {})
    if len(bcs) > _lib.FVH_MAX_BC:
        raise ValueError("at most %d boundary conditions" % _lib.FVH_MAX_BC)
    c.n_bc = len(bcs)
    for k, (name, spec) in enumerate(bcs.items()):
        kind, *par = spec
        par = list(par) + [0.0] * (4 - len(par))
        c.bc[k].name = name.encode()
        c.bc[k].kind = _lib.BC[kind]
        c.bc[k].rho, c.bc[k].u, c.bc[k].v, c.bc[k].p = [float(v) for v in par[:4]]
    if renumber is True:
        renumber = "rcm"
    elif renumber is False:
        renumber = "file"
    c.renumber = ORDERS[renumber]
    c.flux_mode = FLUX_MODES[flux_mode]
    c.tile_cells = int(tile_cells)
    return c


class Solver:
    """Explicit FV solver on the current HIP device (one stream per handle)."""

    def __init__(self, mesh: Mesh, cfg: Dict, precision: str = "f64", renumber="tiles",
                 flux_mode: str = "auto", tile_cells: int = 0):
        self.mesh = mesh
        self.cfg = dict(cfg)
        self.precision = precision
        L = _lib.load()
        c = make_config(cfg, precision, renumber, flux_mode, tile_cells)
        h = C.c_void_p()
        check(L.fvh_solver_create(mesh._h, C.byref(c), C.byref(h)))
        self._h = h
        self._n_cells = mesh.n_cells

    @property
    def stage_info(self) -> Dict[str, int]:
        """Layout of the fused stage pass (fvh_solver_stage_info)."""
        a = (C.c_int64 * 8)()
        check(_lib.load().fvh_solver_stage_info(self._h, a))
        keys = ["fused", "tiles", "lds_bytes", "lanes", "max_local_cells", "max_tile_edges", "tile_cells",
                "mean_tile_ring1_cells"]
        return {k: int(v) for k, v in zip(keys, a)}

    def use_graphs(self, enable: bool = True) -> None:
        """Replay each RK step from a captured HIP graph."""
        check(_lib.load().fvh_solver_use_graphs(self._h, 1 if enable else 0))

    def __del__(self):
        try:
            if self._h:
                _lib.load().fvh_solver_free(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def n_state(self) -> int:
        return 4 * self._n_cells

    def set_state(self, q: np.ndarray) -> None:
        q = np.ascontiguousarray(q, dtype=np.float64).ravel()
        if q.size != self.n_state:
            raise ValueError("state holds %d values, mesh needs %d" % (q.size, self.n_state))
        check(_lib.load().fvh_solver_set_state(self._h, dptr(q)))

    def get_state(self) -> np.ndarray:
        q = np.zeros(self.n_state)
        check(_lib.load().fvh_solver_get_state(self._h, dptr(q)))
        return q

    def apply_bounds(self) -> None:
        check(_lib.load().fvh_solver_apply_bounds(self._h))

    def save(self, path: str) -> None:
        """Binary restart file: every cell's state and the run bookkeeping."""
        check(_lib.load().fvh_solver_save(self._h, os.fsencode(path)))

    def load(self, path: str) -> None:
        """Restore a restart file written by save() for the same meshes;
        run_continue() then carries on the saved run bit for bit."""
        check(_lib.load().fvh_solver_load(self._h, os.fsencode(path)))

    def phases(self, alpha: float = 0.5) -> Dict[str, np.ndarray]:
        n = self.n_state
        out = {"dt": np.zeros(self.mesh.n_cells), "gx": np.zeros(n), "gy": np.zeros(n),
               "lim": np.zeros(n), "qt": np.zeros(n), "qk": np.zeros(n)}
        check(_lib.load().fvh_solver_phases(self._h, float(alpha), dptr(out["dt"]), dptr(out["gx"]),
                                            dptr(out["gy"]), dptr(out["lim"]), dptr(out["qt"]),
                                            dptr(out["qk"])))
        return out

    # -- the phases of explicit.cpp one at a time on given arrays (FP64) --
    def _q(self, a) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        if a.size != self.n_state:
            raise ValueError("array holds %d values, mesh needs %d" % (a.size, self.n_state))
        return a

    def phase_gradients(self, q):
        """calc_gradients (explicit.cpp:112-171): (gx, gy)."""
        gx, gy = np.zeros(self.n_state), np.zeros(self.n_state)
        check(_lib.load().fvh_phase_gradients(self._h, dptr(self._q(q)), dptr(gx), dptr(gy)))
        return gx, gy

    def phase_limiters(self, q, gx, gy):
        """calc_limiters (explicit.cpp:174-258): (limiters, qmin, qmax)."""
        lim, qmin, qmax = (np.zeros(self.n_state) for _ in range(3))
        check(_lib.load().fvh_phase_limiters(self._h, dptr(self._q(q)), dptr(self._q(gx)), dptr(self._q(gy)),
                                             dptr(lim), dptr(qmin), dptr(qmax)))
        return lim, qmin, qmax

    def phase_time_derivatives(self, q, gx, gy, lim):
        """calc_time_derivatives (explicit.cpp:261-364): qt."""
        qt = np.zeros(self.n_state)
        check(_lib.load().fvh_phase_time_derivatives(self._h, dptr(self._q(q)), dptr(self._q(gx)),
                                                     dptr(self._q(gy)), dptr(self._q(lim)), dptr(qt)))
        return qt

    def phase_update_cells(self, ql, qt, dt, v: float):
        """update_cells (explicit.cpp:368-378): ql + qt dt v, dt one value per entry."""
        ql, qt, dt = (np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (ql, qt, dt))
        q = np.zeros(ql.size)
        check(_lib.load().fvh_phase_update_cells(self._h, dptr(q), dptr(ql), dptr(qt), dptr(dt), float(v), ql.size))
        return q

    def phase_update_bounds(self, q, gx, gy, lim):
        """update_bounds (explicit.cpp:380-417): q with its ghost cells updated."""
        out = self._q(q).copy()
        check(_lib.load().fvh_phase_update_bounds(self._h, dptr(out), dptr(self._q(gx)), dptr(self._q(gy)),
                                                  dptr(self._q(lim))))
        return out

    def phase_min_dt(self, dt):
        """min_dt (explicit.cpp:555-564): every value set to the minimum."""
        out = np.ascontiguousarray(dt, dtype=np.float64).ravel().copy()
        check(_lib.load().fvh_phase_min_dt(self._h, dptr(out), out.size))
        return out

    def phase_residuals(self, qt):
        """calc_residuals (explicit.cpp:481-552): R[4]."""
        R = np.zeros(4)
        check(_lib.load().fvh_phase_residuals(self._h, dptr(self._q(qt)), dptr(R)))
        return R

    def run(self, max_step: int = 100000000, max_time: float = 1e10, tol: float = 1e-16,
            print_interval: int = 1, stop_var: int = -1) -> Dict:
        """run() of explicit.h:181; stop_var=-1 is the reference's stop test
        (max over variables of R/R0), k >= 0 tests variable k alone."""
        o = _lib.RunOptions(float(max_time), int(max_step), int(print_interval), float(tol), int(stop_var), 0)
        st = _lib.Status()
        check(_lib.load().fvh_solver_run(self._h, C.byref(o), C.byref(st)))
        return self._status_dict(st)

    def run_begin(self) -> None:
        check(_lib.load().fvh_solver_run_begin(self._h))

    def run_continue(self, n: int, max_step: int = 100000000, max_time: float = 1e10, tol: float = 1e-16,
                     print_interval: int = 1, stop_var: int = -1) -> Dict:
        o = _lib.RunOptions(float(max_time), int(max_step), int(print_interval), float(tol), int(stop_var), 0)
        st = _lib.Status()
        check(_lib.load().fvh_solver_run_continue(self._h, C.byref(o), int(n), C.byref(st)))
        return self._status_dict(st)

    def run_log(self, max_step: int = 100000000, max_time: float = 1e10, tol: float = 1e-16,
                print_interval: int = 1, stop_var: int = -1):
        """run() with the reference's printed rows: returns (status, rows), a
        row (step, time before the step, R[4]) after step 0 and after every
        print_interval-th step (explicit.cpp:752-785)."""
        o = _lib.RunOptions(float(max_time), int(max_step), int(print_interval), float(tol), int(stop_var), 0)
        st = _lib.Status()
        rows = []

        def row(step, p, _user):
            rows.append((int(step), float(p.contents.time_prev), list(p.contents.R)))
        cb = _lib.ROW_CB(row)
        check(_lib.load().fvh_solver_run_log(self._h, C.byref(o), cb, None, C.byref(st)))
        return self._status_dict(st), rows

    def advance(self, n: int) -> None:
        check(_lib.load().fvh_solver_advance(self._h, int(n)))

    def synchronize(self) -> None:
        check(_lib.load().fvh_solver_synchronize(self._h))

    def status(self) -> Dict:
        st = _lib.Status()
        check(_lib.load().fvh_solver_status(self._h, C.byref(st)))
        return self._status_dict(st)

    @staticmethod
    def _status_dict(st) -> Dict:
        return {"step": int(st.step), "stopped": bool(st.stopped), "time": float(st.time),
                "time_prev": float(st.time_prev),
                "R": np.array(list(st.R)), "R0": np.array(list(st.R0))}

    def profile(self, enable: bool) -> None:
        check(_lib.load().fvh_solver_profile(self._h, 1 if enable else 0))

    def kernel_times(self) -> Dict[str, Dict[str, float]]:
        L = _lib.load()
        counts = (C.c_int64 * 16)()
        ms = (C.c_double * 16)()
        n = check(L.fvh_solver_kernel_times(self._h, 16, counts, ms))
        return {L.fvh_solver_kernel_name(k).decode(): {"count": int(counts[k]), "total_ms": float(ms[k])}
                for k in range(n)}

    def stage_bytes(self) -> float:
        return float(_lib.load().fvh_solver_stage_bytes(self._h))

    @property
    def tile_cells(self) -> int:
        """Cells per tile in use (the automatic choice resolved)."""
        return check(_lib.load().fvh_solver_tile_cells(self._h))

    def _weights(self, wx, wy):
        nb = self.mesh.n_bnd
        out = []
        for w in (wx, wy):
            if w is None:
                out.append(None)
            else:
                w = np.ascontiguousarray(w, dtype=np.float64).ravel()
                if w.size != nb:
                    raise ValueError("one weight per boundary element (%d)" % nb)
                out.append(w)
        return out

    def objective(self, q0: np.ndarray, n_steps: int, wx=None, wy=None) -> float:
        """J = sum_b p(ghost_b) (wx_b n_x + wy_b n_y) len_b after n_steps
        steps of run() from q0 (fvh_objective)."""
        q0 = np.ascontiguousarray(q0, dtype=np.float64).ravel()
        wx, wy = self._weights(wx, wy)
        J = C.c_double(0.0)
        check(_lib.load().fvh_objective(self._h, dptr(q0), int(n_steps), dptr(wx) if wx is not None else None,
                                        dptr(wy) if wy is not None else None, C.byref(J)))
        return float(J.value)

    def adjoint(self, q0: np.ndarray, n_steps: int, wx=None, wy=None, snaps: int = 0, mem_cap_bytes: float = 8e9,
                nodes: bool = False):
        """(J, dJ/dq0, stats) by the discrete adjoint (fvh_adjoint); with
        nodes=True, (J, dJ/dq0, dJ/dxy, stats) where dJ/dxy has shape
        (n_nodes, 2): the derivative with respect to each node's x and y."""
        q0 = np.ascontiguousarray(q0, dtype=np.float64).ravel()
        wx, wy = self._weights(wx, wy)
        J = C.c_double(0.0)
        grad = np.zeros_like(q0)
        gxy = np.zeros((self.mesh.n_nodes, 2)) if nodes else None
        st = _lib.AdjointStats()
        check(_lib.load().fvh_adjoint(self._h, dptr(q0), int(n_steps), dptr(wx) if wx is not None else None,
                                      dptr(wy) if wy is not None else None, int(snaps), float(mem_cap_bytes),
                                      C.byref(J), dptr(grad), dptr(gxy) if nodes else None, C.byref(st)))
        stats = {k: getattr(st, k) for k, _ in _lib.AdjointStats._fields_}
        if nodes:
            return float(J.value), grad, gxy, stats
        return float(J.value), grad, stats


class Batch(Solver):
    """Many independent meshes advanced together in the same launches
    (fvh_batch_create).  ``cfgs`` is one config dict per mesh; they may
    differ only in their boundary tables.  States are the concatenation of
    the instances' reference-layout states; ``split``/``join`` convert."""

    def __init__(self, meshes: List[Mesh], cfgs: List[Dict], precision: str = "f64", renumber="tiles",
                 flux_mode: str = "auto", tile_cells: int = 0):
        if len(meshes) != len(cfgs) or not meshes:
            raise ValueError("one config per mesh, at least one mesh")
        self.meshes = list(meshes)
        self.cfg = dict(cfgs[0])
        self.precision = precision
        L = _lib.load()
        n = len(meshes)
        cs = (_lib.Config * n)(*[make_config(c, precision, renumber, flux_mode, tile_cells) for c in cfgs])
        hs = (C.c_void_p * n)(*[m._h for m in meshes])
        h = C.c_void_p()
        check(L.fvh_batch_create(n, hs, cs, C.byref(h)))
        self._h = h
        self.mesh = meshes[0]
        info = (C.c_int64 * 3)()
        self.offsets = []
        for k in range(n):
            check(L.fvh_batch_instance(self._h, k, info))
            self.offsets.append(int(info[2]))
        self._n_cells = sum(m.n_cells for m in meshes)

    def __len__(self) -> int:
        return len(self.meshes)

    def split(self, q: np.ndarray) -> List[np.ndarray]:
        """Concatenated state -> one reference-layout state per instance."""
        return [q[4 * o:4 * (o + m.n_cells)].copy() for o, m in zip(self.offsets, self.meshes)]

    def join(self, qs: List[np.ndarray]) -> np.ndarray:
        return np.concatenate([np.ascontiguousarray(q, dtype=np.float64).ravel() for q in qs])

    def instance_status(self, k: int) -> Dict:
        st = _lib.Status()
        check(_lib.load().fvh_batch_status(self._h, int(k), C.byref(st)))
        return self._status_dict(st)

    def sums(self) -> np.ndarray:
        """FP64 conserved totals sum A q per instance, shape (n, 4)."""
        out = np.zeros(4 * len(self.meshes))
        check(_lib.load().fvh_batch_sums(self._h, dptr(out)))
        return out.reshape(-1, 4)


def stream_copy_gbs(mbytes: float = 1024.0, reps: int = 20) -> float:
    """Achievable HBM bandwidth of a plain device copy (read + write GB/s)."""
    g = C.c_double(0.0)
    check(_lib.load().fvh_stream_copy(float(mbytes), int(reps), C.byref(g)))
    return g.value


def autotune(meshes: List[Mesh], cfgs: List[Dict], q: np.ndarray, precision: str = "f64", steps: int = 10,
             renumber="tiles") -> Dict:
    """Time every stage form and tile size on these meshes (fvh_autotune)
    and return {"flux_mode", "tile_cells", "candidates": [(mode, tile, ms/step)]}."""
    L = _lib.load()
    n = len(meshes)
    cs = (_lib.Config * n)(*[make_config(c, precision, renumber) for c in cfgs])
    hs = (C.c_void_p * n)(*[m._h for m in meshes])
    q = np.ascontiguousarray(q, dtype=np.float64).ravel()
    bm, bt = C.c_int32(0), C.c_int32(0)
    rep = np.zeros(3 * 16)
    rows = check(L.fvh_autotune(n, hs, cs, dptr(q), int(steps), C.byref(bm), C.byref(bt), dptr(rep), 16))
    names = {v: k for k, v in FLUX_MODES.items()}
    cands = [(names[int(rep[3 * k])], int(rep[3 * k + 1]), float(rep[3 * k + 2])) for k in range(min(rows, 16))]
    return {"flux_mode": names[bm.value], "tile_cells": bt.value, "candidates": cands}
