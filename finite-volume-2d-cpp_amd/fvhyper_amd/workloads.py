"""The benchmark workloads: mesh, physics configuration and initial state.

Each returns ``(mesh_arrays, cfg, ic)`` where ``ic(cx, cy)`` maps cell
centres (all cells, reference order) to the AoS initial state.  Physics
kinds and constants follow the reference example each workload descends
from; ``reference_compatible=True`` switches the non-reference choices of the
workload definition (explicit Euler, K=5, ...) back to what the reference
library can express, for parity runs.
"""
from __future__ import annotations

from typing import Callable, Dict, Tuple

import numpy as np

from . import meshgen

GAMMA = 1.4
RK5 = [0.0695, 0.1602, 0.2898, 0.506, 1.0]


def sod_ic(gamma: float = GAMMA) -> Callable:
    """Sod initial state as examples/shockTube/main.cpp:40-57 writes it."""
    def ic(cx, cy):
        q = np.zeros((cx.size, 4))
        left = cx < 0.5
        q[:, 0] = np.where(left, 1.0, 0.125)
        q[:, 3] = np.where(left, 1.0 / (gamma - 1), 0.1 / (gamma - 1))
        return q.ravel()
    return ic


def uniform_ic(rho: float, u: float, v: float, p: float, gamma: float = GAMMA) -> Callable:
    """Uniform state, written as examples/forwardStep/main.cpp:48-53 does:
    E = p/(gamma-1) + 0.5*rho*(u*u + v*v)."""
    def ic(cx, cy):
        q = np.empty((cx.size, 4))
        q[:, 0] = rho
        q[:, 1] = rho * u
        q[:, 2] = rho * v
        q[:, 3] = p / (gamma - 1) + 0.5 * rho * (u * u + v * v)
        return q.ravel()
    return ic


def sod(nx: int = 400, ny: int = 40, seed: int = 20240601, reference_compatible: bool = False
        ) -> Tuple[meshgen.MeshArrays, Dict, Callable]:
    """Sod shock tube: [0,1]x[0,0.1], 400x40 split quads (32,000 triangles),
    jittered interior nodes; first-order Roe flux (shockTube spelling),
    slip walls top/bottom, transmissive ends, CFL 0.5 global time step,
    explicit Euler (the reference's RK5 when reference_compatible), to t=0.2."""
    m = meshgen.sod_tube(nx=nx, ny=ny, seed=seed)
    cfg = dict(flux="roe_split", phi="zero", dt="sum_cfl", cfl=0.5, gdt=True,
               grad=False, lim=False, lin=False, gamma=GAMMA,
               alpha=RK5 if reference_compatible else [1.0],
               bcs={"wall": ("mirror",), "transmissive": ("copy",)})
    return m, cfg, sod_ic()


def forward_step(n_per_unit: int = 440, seed: int = 20240602, reference_compatible: bool = False,
                 mach: float = 3.0) -> Tuple[meshgen.MeshArrays, Dict, Callable]:
    """Mach-3 forward-facing step: channel [0,3]x[0,1] minus the step
    [0.6,3]x[0,0.2], Delaunay-split jittered lattice (975,744 triangles at
    440/unit); second order (Green-Gauss gradients, Venkatakrishnan limiter
    with K=5, 7.5 when reference_compatible), forwardStep Roe spelling,
    5-stage RK, CFL 0.8 global time step; inflow rho=1.4, u=3, v=0, p=1.

    The reference's second-order scheme limits each conserved variable on its
    own; at the step corner that lets a face state reach negative pressure
    within ~100 steps (reproduced bit for bit by this library), after which
    the run is NaN.  The workload therefore turns on the positivity fallback
    (first-order face state where the reconstruction is not physical);
    reference_compatible turns it off again."""
    m = meshgen.forward_step(n_per_unit=n_per_unit, seed=seed)
    free = ("inlet_outlet", 1.4, mach, 0.0, 1.0)
    cfg = dict(flux="roe_joined", phi="venkat", dt="sum_cfl", cfl=0.8, gdt=True,
               grad=True, lim=True, lin=True, gamma=GAMMA,
               K=7.5 if reference_compatible else 5.0, alpha=RK5,
               positivity=not reference_compatible,
               bcs={"wall": ("mirror",), "inlet": free, "outlet": free})
    return m, cfg, uniform_ic(1.4, mach, 0.0, 1.0)


def step_face_weights(mesh) -> np.ndarray:
    """Objective weights of the forward-step workload: 1 on the boundary
    elements of the step's vertical face (x = 0.6, 0 <= y <= 0.2), 0 elsewhere,
    so that J = sum p n_x len is the pressure force on that face."""
    ex, ey = mesh.export("edgesCentersX"), mesh.export("edgesCentersY")
    be = mesh.export("boundaryEdges")
    on = (np.abs(ex[be] - 0.6) < 1e-9) & (ey[be] <= 0.2 + 1e-9)
    return on.astype(np.float64)


def bump_state(mesh, rho=1.4, u=3.0, v=0.0, p=1.0, amplitude=1e-3, n_bumps=4, seed=20240605,
               gamma: float = GAMMA) -> np.ndarray:
    """Uniform state with a seeded smooth perturbation of relative size
    `amplitude`: Gaussian bumps (centres uniform in the mesh's bounding box,
    widths 0.05-0.15, random signs) in density, the same bumps at half weight
    in pressure, and smooth background waves sin(2 pi x/Lx + a) cos(2 pi y/Ly
    + b) at quarter weight in pressure and in both velocity components.
    Density and pressure must not move together: scaled by one common factor
    they keep p/rho, and so the sound speed and every time-step wave speed,
    uniform, and neighbouring cells then tie in the time step's max(own,
    neighbour) -- a kink of the discrete map that central differences
    straddle."""
    cx, cy = mesh.export("cellsCentersX"), mesh.export("cellsCentersY")
    nr = mesh.n_real
    x0, x1, y0, y1 = cx[:nr].min(), cx[:nr].max(), cy[:nr].min(), cy[:nr].max()
    lx, ly = max(x1 - x0, 1e-30), max(y1 - y0, 1e-30)
    r = meshgen.uniform01(seed, 4 * n_bumps + 6).reshape(-1)
    ph = 2 * np.pi * r[4 * n_bumps:]
    r = r[:4 * n_bumps].reshape(n_bumps, 4)
    f = np.zeros(cx.size)
    for k in range(n_bumps):
        bx, by = x0 + lx * r[k, 0], y0 + ly * r[k, 1]
        w = 0.05 + 0.1 * r[k, 2]
        sgn = 1.0 if r[k, 3] < 0.5 else -1.0
        f += sgn * np.exp(-((cx - bx) ** 2 + (cy - by) ** 2) / (w * w))
    X, Y = 2 * np.pi * (cx - x0) / lx, 2 * np.pi * (cy - y0) / ly
    bg = [np.sin(X + ph[2 * i]) * np.cos(Y + ph[2 * i + 1]) for i in range(3)]
    speed = max(np.hypot(u, v), 1.0)
    rr = rho * (1.0 + amplitude * f)
    pp = p * (1.0 + amplitude * (0.5 * f + 0.25 * bg[0]))
    uu = u + amplitude * 0.25 * speed * bg[1]
    vv = v + amplitude * 0.25 * speed * bg[2]
    q = np.stack([rr, rr * uu, rr * vv, pp / (gamma - 1) + 0.5 * rr * (uu * uu + vv * vv)], 1)
    return q.ravel()


def naca_freestream(mach: float = 0.8, aoa_deg: float = 1.25, reference_compatible: bool = False):
    """(rho, u, v, p) of the airfoil freestream: rho=1.4, p=1 (so c=1 and
    |V| = Mach), velocity at the angle of attack.  reference_compatible gives
    the reference's hard-coded u=0.8, v=0.0175 (examples/naca/main.cpp:49-50,
    194-199), i.e. 1.2532 degrees."""
    if reference_compatible:
        return 1.4, 0.8, 0.0175, 1.0
    a = np.radians(aoa_deg)
    return 1.4, mach * float(np.cos(a)), mach * float(np.sin(a)), 1.0


def naca_config(mach: float = 0.8, aoa_deg: float = 1.25, reference_compatible: bool = False) -> Dict:
    """Physics of examples/naca/main.cpp: second order (gradients, Michalak
    limiter y_t=2, linear interpolation; :21-28, :56-65), the shockTube/naca
    Roe spelling (:71-137), slip wall on ``airfoil`` and the characteristic
    ``farfield`` (:192-272), 5-stage RK, local time stepping.

    The example's own time step (:142-183) takes its wave speed as
    (c+|u|)n_x + (c+|v|)n_y with signed normal components; that nearly
    vanishes for some edge directions, and on the stretched quads of the
    boundary band it over-estimates dt by up to the aspect ratio (the run goes
    non-finite within 60 steps from 50k cells up).  The workload therefore
    uses the forwardStep example's summed form, dt = CFL A / sum_e
    (|u.n|+c) l (forwardStep/main.cpp:140-185), per cell (global_dt off) with
    CFL 2; reference_compatible restores the example's form at CFL 1.0."""
    free = ("farfield",) + naca_freestream(mach, aoa_deg, reference_compatible)
    dt = dict(dt="edge_min", cfl=1.0) if reference_compatible else dict(dt="sum_cfl", cfl=2.0)
    return dict(flux="roe_split", phi="michalak", phi_yt=2.0, gdt=False, **dt,
                grad=True, lim=True, lin=True, gamma=GAMMA, alpha=RK5,
                bcs={"airfoil": ("mirror",), "farfield": free})


def naca_airfoil(total_cells: int = 2_000_000, mach: float = 0.8, aoa_deg: float = 1.25,
                 reference_compatible: bool = False) -> Tuple[meshgen.MeshArrays, Dict, Callable]:
    """Transonic NACA 0012 to steady state: O-mesh with 512 surface points, a
    64-layer quad band (first spacing 1e-3, growth 1.1) and triangles to a
    farfield of radius 50 chords, ~2.0M cells at the default size (smaller
    sizes coarsen surface points and layers with sqrt(cells)); Mach 0.8 at
    1.25 degrees; uniform freestream initial state (naca/main.cpp:43-53).
    Converged when the density residual has dropped 6 orders (stop_var=0,
    tolerance 1e-6)."""
    m = meshgen.naca0012_omesh() if total_cells == 2_000_000 else meshgen.naca0012_sized(total_cells)
    cfg = naca_config(mach, aoa_deg, reference_compatible)
    _, rho, u, v, p = cfg["bcs"]["farfield"]
    return m, cfg, uniform_ic(rho, u, v, p)


def flat_plate(nx: int = 1000, ny: int = 500, smooth: bool = False) -> Tuple[meshgen.MeshArrays, Dict, Callable]:
    """Laminar flat plate with the physics of examples/flatPlate/main.cpp:
    split Roe flux plus laminar terms on diffusive face gradients (mu 2e-5,
    Pr 0.72, cp 1, R 8.3145), Michalak limiter, viscous edge-minimum local
    dt at CFL 2, no-slip plate (bot1), slip wall upstream (bot0), farfield
    (rho 1.4, u 0.2, p 1) elsewhere; uniform freestream initial state
    (flatPlate/main.cpp:98-109).  nx x ny squares of [0,1.5]x[0,0.5], two
    triangles each: 1,000,000 cells at the default size.  ``smooth`` adds
    the residual smoothing of explicit.cpp:32-66."""
    m = meshgen.flat_plate(nx=nx, ny=ny)
    ff = ("farfield_v", 1.4, 0.2, 0.0, 1.0)
    cfg = dict(flux="roe_viscous", dt="edge_min_viscous", cfl=2.0, phi="michalak", grad=True, lim=True, lin=True,
               diffusive=True, gdt=False, smooth=smooth, mu=2e-5, pr=0.72, cp=1.0, r_gas=8.3145, alpha=RK5,
               bcs={"top": ff, "left": ff, "right": ff, "bot0": ("mirror",), "bot1": ("noslip",)})
    return m, cfg, uniform_ic(1.4, 0.2, 0.0, 1.0)


def naca_sweep(n: int = 256, seed: int = 20240604, min_cells: int = 5000, max_cells: int = 200000,
               mach_range=(0.5, 0.85), aoa_range=(-2.0, 4.0)):
    """Parameter sweep over coarsened airfoil meshes: n instances with cell
    counts log-uniform in [min_cells, max_cells], Mach and angle of attack
    uniform in their ranges, all drawn from one splitmix64 stream.  Returns a
    list of (mesh_arrays, cfg, ic, (cells, mach, aoa)); meshes of the same
    target size are generated once and shared."""
    u = meshgen.uniform01(seed, 3 * n).reshape(n, 3)
    lo, hi = np.log(min_cells), np.log(max_cells)
    cache = {}
    out = []
    for k in range(n):
        cells = int(round(float(np.exp(lo + (hi - lo) * u[k, 0]))))
        mach = mach_range[0] + (mach_range[1] - mach_range[0]) * float(u[k, 1])
        aoa = aoa_range[0] + (aoa_range[1] - aoa_range[0]) * float(u[k, 2])
        if cells not in cache:
            cache[cells] = meshgen.naca0012_sized(cells)
        cfg = naca_config(mach, aoa)
        _, rho, uu, vv, p = cfg["bcs"]["farfield"]
        out.append((cache[cells], cfg, uniform_ic(rho, uu, vv, p), (cells, mach, aoa)))
    return out


def initial_state(mesh, ic: Callable) -> np.ndarray:
    """Initial state over all cells of a built mesh (solver.Mesh or oracle Mesh)."""
    if hasattr(mesh, "export"):
        cx, cy = mesh.export("cellsCentersX"), mesh.export("cellsCentersY")
    else:
        cx, cy = mesh.cell_cx, mesh.cell_cy
    return ic(cx, cy)
