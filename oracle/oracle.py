"""numpy front end of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py -- never by the diffsph_amd product path.

All functions take and return numpy arrays on the host.  `build()` compiles
the shared library with oracle/Makefile when it is missing or stale.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "liboracle.so")
SRC = os.path.join(HERE, "oracle.cpp")
DEPS = [SRC, os.path.join(HERE, "..", "diffsph_amd", "csrc", "pmath.h")]


class Domain(C.Structure):
    _fields_ = [("min", C.c_float * 2), ("max", C.c_float * 2), ("periodic", C.c_int32 * 2)]


def make_domain(mn, mx, periodic) -> Domain:
    d = Domain()
    for k in range(2):
        d.min[k] = float(mn[k])
        d.max[k] = float(mx[k])
        d.periodic[k] = 1 if periodic[k] else 0
    return d


def build(force: bool = False) -> str:
    stale = not os.path.exists(LIB) or os.path.getmtime(LIB) < max(os.path.getmtime(p) for p in DEPS)
    if force or stale:
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = C.CDLL(LIB)
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle inputs must be contiguous"
    return a.ctypes.data_as(C.c_void_p)


def f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n: int) -> None:
    lib().orc_set_threads(C.c_int(n))


def max_threads() -> int:
    return int(lib().orc_max_threads())


@dataclass
class Grid:
    ncx: int
    ncy: int
    hcell: float
    cell_start: np.ndarray
    sort_index: np.ndarray
    sorted_pos: np.ndarray


def grid_dims(dom: Domain, hsearch: float):
    ncx, ncy, hc = C.c_int32(), C.c_int32(), C.c_float()
    rc = lib().orc_grid_dims(C.byref(dom), C.c_float(hsearch), C.byref(ncx), C.byref(ncy), C.byref(hc))
    if rc != 0:
        raise ValueError("grid_dims failed (periodic axis needs >= 3 cells)")
    return ncx.value, ncy.value, hc.value


def grid_build(dom: Domain, pos: np.ndarray, hsearch: float) -> Grid:
    pos = f32(pos)
    n = pos.shape[0]
    ncx, ncy, hc = grid_dims(dom, hsearch)
    cs = np.zeros(ncx * ncy + 1, np.int32)
    si = np.zeros(n, np.int32)
    sp = np.zeros((n, 2), np.float32)
    lib().orc_grid_build(C.byref(dom), C.c_int32(ncx), C.c_int32(ncy), C.c_float(hc), _p(pos), C.c_int64(n),
                         _p(cs), _p(si), _p(sp))
    return Grid(ncx, ncy, hc, cs, si, sp)


def neighbors_cell(dom: Domain, g: Grid, qpos, qh=None, h0=0.0, mode=0, sorted_h=None):
    """CSR (counts, offsets[nq+1], nbr) in canonical order; nbr are sorted slots."""
    qpos = f32(qpos)
    qh = f32(qh)
    sorted_h = f32(sorted_h)
    nq = qpos.shape[0]
    counts = np.zeros(nq, np.int32)
    L = lib()
    args = (C.byref(dom), C.c_int32(g.ncx), C.c_int32(g.ncy), C.c_float(g.hcell), _p(g.cell_start),
            _p(g.sorted_pos), _p(sorted_h), _p(qpos), _p(qh), C.c_int64(nq), C.c_float(h0), C.c_int32(mode))
    L.orc_neighbors_cell(*args, _p(counts), None, None)
    off = np.zeros(nq + 1, np.int32)
    np.cumsum(counts, out=off[1:])
    nbr = np.zeros(max(int(off[-1]), 1), np.int32)
    L.orc_neighbors_cell(*args, None, _p(off), _p(nbr))
    return counts, off, nbr[: int(off[-1])]


def neighbors_brute(dom: Domain, qpos, rpos, qh=None, rh=None, h0=0.0, mode=0):
    """Brute-force CSR (rows in query order, j ascending, original indices)."""
    qpos, rpos, qh, rh = f32(qpos), f32(rpos), f32(qh), f32(rh)
    nq, nr = qpos.shape[0], rpos.shape[0]
    counts = np.zeros(nq, np.int32)
    L = lib()
    args = (C.byref(dom), _p(qpos), _p(qh), C.c_int64(nq), _p(rpos), _p(rh), C.c_int64(nr), C.c_float(h0),
            C.c_int32(mode))
    L.orc_neighbors_brute(*args, _p(counts), None, None)
    off = np.zeros(nq + 1, np.int32)
    np.cumsum(counts, out=off[1:])
    nbr = np.zeros(max(int(off[-1]), 1), np.int32)
    L.orc_neighbors_brute(*args, None, _p(off), _p(nbr))
    return counts, off, nbr[: int(off[-1])]


def density(dom, off, nbr, qpos, rpos, rmass, qmass=None, h0=0.0, qh=None, rh=None, mode=0, dtype=np.float32):
    nq = qpos.shape[0]
    cv = f32 if dtype == np.float32 else f64
    qpos, rpos, rmass, qmass, qh, rh = cv(qpos), cv(rpos), cv(rmass), cv(qmass), cv(qh), cv(rh)
    rho = np.zeros(nq, dtype)
    vol = np.zeros(nq, dtype) if qmass is not None else None
    fn = lib().orc_density_f32 if dtype == np.float32 else lib().orc_density_f64
    hc = C.c_float if dtype == np.float32 else C.c_double
    fn(C.byref(dom), C.c_int64(nq), _p(off), _p(nbr), _p(qpos), _p(rpos), _p(qh), _p(rh), hc(h0), C.c_int32(mode),
       _p(rmass), _p(qmass), _p(rho), _p(vol))
    return rho, vol


def sph_operation(dom, off, nbr, qpos, rpos, op, gm, rm, rr, rq, width, qm=None, qr=None, qq=None, h0=0.0,
                  qh=None, rh=None, mode=0):
    nq = qpos.shape[0]
    qpos, rpos = f32(qpos), f32(rpos)
    arrs = [f32(a) for a in (qm, rm, qr, rr, qq, rq, qh, rh)]
    qm, rm, qr, rr, qq, rq, qh, rh = arrs
    if op == 0:
        out = np.zeros((nq, width), np.float32)
    elif op == 1:
        out = np.zeros((nq, width, 2), np.float32)
    else:
        out = np.zeros(nq, np.float32)
    lib().orc_sph_operation_f32(C.byref(dom), C.c_int64(nq), _p(off), _p(nbr), _p(qpos), _p(rpos), _p(qh), _p(rh),
                                C.c_float(h0), C.c_int32(mode), C.c_int32(op), C.c_int32(gm), _p(qm), _p(rm),
                                _p(qr), _p(rr), _p(qq), _p(rq), C.c_int32(width), _p(out))
    return out


@dataclass
class Visc:
    scheme: int = 0      # 0 deltaSPH_inviscid, 1 deltaSPH_viscid
    alpha: float = 0.01  # alpha, or nu for the viscid scheme
    cs: float = 10.0
    rho0: float = 1.0
    eps: float = 1e-6
    pi_switch: bool = False


def dfsph_prepare(dom, off, nbr, pos, vel, areas, vol, rho, dt, rho0, visc: Visc | None, h0, dtype=np.float32):
    n = pos.shape[0]
    cv = f32 if dtype == np.float32 else f64
    pos, vel, areas, vol, rho = cv(pos), cv(vel), cv(areas), cv(vol), cv(rho)
    ad = np.zeros(n, dtype)
    an = np.zeros(n, dtype)
    ds = np.zeros(n, dtype)
    vs = np.zeros((n, 2), dtype)
    hc = C.c_float if dtype == np.float32 else C.c_double
    fn = lib().orc_dfsph_prepare_f32 if dtype == np.float32 else lib().orc_dfsph_prepare_f64
    v = visc or Visc()
    fn(C.byref(dom), C.c_int64(n), _p(off), _p(nbr), _p(pos), None, hc(h0), C.c_int32(0), _p(vel), _p(areas),
       _p(vol), _p(rho), hc(dt), hc(rho0), C.c_int32(1 if visc else 0), C.c_int32(v.scheme), hc(v.alpha), hc(v.cs),
       hc(v.rho0), hc(v.eps), C.c_int32(1 if v.pi_switch else 0), _p(ad), _p(an), _p(ds), _p(vs))
    return ad, an, ds, vs


def dfsph_solve(dom, off, nbr, pos, h0, dt, rho0, omega, thr, min_iters, max_iters, clamp, density_src, vel, adv,
                rho, vol, areas, alpha, src0, p0, dtype=np.float32, isph=False):
    """dfsphSolve's loop (v2/modules/dfsph.py:52-109); with `isph` the
    relaxed Jacobi of solveIncompressible_relaxedJacobi
    (v2/modules/incompressible.py:308-364): 'difference' pressure gradient,
    error mean(|r|) (alpha, omega, clamp and warm start from the caller)."""
    n = pos.shape[0]
    cv = f32 if dtype == np.float32 else f64
    hc = C.c_float if dtype == np.float32 else C.c_double
    pos, vel, adv, rho, vol, areas, alpha, src0 = (cv(a) for a in (pos, vel, adv, rho, vol, areas, alpha, src0))
    p = cv(p0).copy()
    acc = np.zeros((n, 2), dtype)
    res = np.zeros(n, dtype)
    it = C.c_int32()
    err, err0 = C.c_double(), C.c_double()
    fn = lib().orc_dfsph_solve_f32 if dtype == np.float32 else lib().orc_dfsph_solve_f64
    fn(C.byref(dom), C.c_int64(n), _p(off), _p(nbr), _p(pos), None, hc(h0), C.c_int32(0), hc(dt), hc(rho0),
       hc(omega), hc(thr), C.c_int32(min_iters), C.c_int32(max_iters), C.c_int32(1 if clamp else 0),
       C.c_int32(1 if density_src else 0), _p(vel), _p(adv), _p(rho), _p(vol), _p(areas), _p(alpha), _p(src0),
       _p(p), _p(acc), _p(res), C.byref(it), C.byref(err), C.byref(err0), C.c_int32(1 if isph else 0))
    return p, acc, res, it.value, err.value, err0.value


def integrate_sie(pos, vel, a0, a1, a2, a3, g, dt):
    pos, vel = f32(pos).copy(), f32(vel).copy()
    a0, a1, a2, a3 = f32(a0), f32(a1), f32(a2), f32(a3)
    g = f32(g) if g is not None else None
    lib().orc_integrate_sie_f32(_p(pos), _p(vel), _p(a0), _p(a1), _p(a2), _p(a3), _p(g), C.c_int64(pos.shape[0]),
                                C.c_float(dt))
    return pos, vel


# ---------------------------------------------------------------------------
# Composed DFSPH step (v2/simulationSchemes/dfsph.py:148-227 with the
# semi-implicit Euler of v2/modules/integration.py:111-138).
# ---------------------------------------------------------------------------
@dataclass
class DfsphState:
    pos: np.ndarray       # [n,2] (unwrapped, as the reference keeps them)
    vel: np.ndarray
    mass: np.ndarray
    areas: np.ndarray
    uid: np.ndarray
    p_div: np.ndarray | None = None
    p_den: np.ndarray | None = None


def dfsph_step(st: DfsphState, dom: Domain, h: float, dt: float, rho0: float = 1.0, omega=0.3, thr_div=1e-2,
               thr_den=5e-4, min_iters=2, max_iters=256, visc: Visc | None = Visc(), warm=0.75, gravity=None,
               reorder=True, dtype=np.float32):
    """One DFSPH step in float32 (or float64 for the error reference).

    With ``reorder`` the state is sorted into cell order first, exactly as
    the GPU solver keeps it, so both walk every neighbor list in the same
    order.  Returns (state, info dict)."""
    g = grid_build(dom, st.pos, h)
    if reorder:
        si = g.sort_index
        st = DfsphState(st.pos[si], st.vel[si], st.mass[si], st.areas[si], st.uid[si],
                        None if st.p_div is None else st.p_div[si], None if st.p_den is None else st.p_den[si])
        g = grid_build(dom, st.pos, h)
    cnt, off, nbr_s = neighbors_cell(dom, g, st.pos, h0=h)
    # neighbors_cell returns sorted slots of g; map to row indices of st
    nbr = g.sort_index[nbr_s].astype(np.int32)
    wpos = np.empty_like(g.sorted_pos)
    wpos[g.sort_index] = g.sorted_pos
    cv = f32 if dtype == np.float32 else f64
    wpos = cv(wpos) if dtype == np.float32 else wrap_positions(dom, st.pos, np.float64)
    rho, vol = density(dom, off, nbr, wpos, wpos, st.mass, st.mass, h0=h, dtype=dtype)
    ad, an, ds, vs = dfsph_prepare(dom, off, nbr, wpos, st.vel, st.areas, vol, rho, dt, rho0, visc, h, dtype)
    if visc is None:
        vs = np.zeros_like(vs)
    n = st.pos.shape[0]
    p0 = (dtype(warm) * cv(st.p_div)) if st.p_div is not None else np.zeros(n, dtype)
    p_div, a_div, _, it_div, e_div, _ = dfsph_solve(dom, off, nbr, wpos, h, dt, rho0, omega, thr_div / rho0,
                                                    min_iters, max_iters, True, False, st.vel, None, rho, vol,
                                                    st.areas, ad, ds, p0, dtype)
    gv = np.zeros(2, dtype) if gravity is None else np.asarray(gravity, dtype)
    adv = (gv[None, :] + vs) + a_div
    p0 = (dtype(warm) * cv(st.p_den)) if st.p_den is not None else np.zeros(n, dtype)
    p_den, a_den, res, it_den, e_den, e_den0 = dfsph_solve(dom, off, nbr, wpos, h, dt, rho0, omega, thr_den / rho0,
                                                           min_iters, max_iters, True, True, st.vel, adv, rho, vol,
                                                           st.areas, an, ds, p0, dtype)
    if dtype == np.float32:
        pos, vel = integrate_sie(st.pos, st.vel, a_den, vs, a_div, None, gravity, dt)
    else:
        dudt = ((a_den + gv[None, :]) + vs) + a_div
        vel = cv(st.vel) + dudt * dt
        pos = cv(st.pos) + vel * dt
    new = DfsphState(pos, vel, st.mass, st.areas, st.uid, p_div, p_den)
    info = dict(rho=rho, vol=vol, alpha_div=ad, alpha_den=an, div_src=ds, visc=vs, a_div=a_div, a_den=a_den,
                p_div=p_div, p_den=p_den, it_div=it_div, it_den=it_den, err_div=e_div, err_den=e_den,
                err_den0=e_den0, num_pairs=int(off[-1]), uid=st.uid)
    return new, info


def wendland(q, h):
    """(W, dW/dr) of the Wendland C2 kernel in float64."""
    q = f64(np.atleast_1d(q))
    h = f64(np.broadcast_to(np.atleast_1d(h), q.shape))
    w = np.zeros_like(q)
    dw = np.zeros_like(q)
    lib().orc_wendland(_p(q), _p(h), C.c_int64(q.size), _p(w), _p(dw))
    return w, dw


# ---------------------------------------------------------------------------
# delta-SPH (v2/simulationSchemes/deltaPlus.py:176-422) with RK4
# (v2/modules/integration.py:231-307) and computeTimestep
# (v2/modules/timestep.py:8-39).
# ---------------------------------------------------------------------------
def pinv2x2(M, dtype=np.float32):
    """Closed-form 2x2 pseudo-inverse of v2/math.py:10-57 for M [n,2,2]."""
    cv = f32 if dtype == np.float32 else f64
    M = cv(np.asarray(M).reshape(-1, 4))
    n = M.shape[0]
    L = np.zeros((n, 4), dtype)
    ev = np.zeros((n, 2), dtype)
    fn = lib().orc_pinv2x2_f32 if dtype == np.float32 else lib().orc_pinv2x2_f64
    fn(_p(M), C.c_int64(n), _p(L), _p(ev))
    return L.reshape(n, 2, 2), ev


def atan2_pm(y, x):
    y, x = f32(y), f32(x)
    out = np.zeros(y.shape, np.float32)
    lib().orc_atan2_pm(_p(y), _p(x), C.c_int64(y.size), _p(out))
    return out


def sincos_pm(x):
    x = f32(x)
    s = np.zeros(x.shape, np.float32)
    c = np.zeros(x.shape, np.float32)
    lib().orc_sincos_pm(_p(x), C.c_int64(x.size), _p(s), _p(c))
    return s, c


@dataclass
class DeltaParams:
    rho0: float = 1.0
    cs: float = 10.0
    delta: float = 0.1
    kernel_scale: float = 1.0
    viscosity: int = 1          # -1 none, 0 deltaSPH_inviscid, 1 deltaSPH_viscid
    visc_coeff: float = 0.01    # nu (viscid) or alpha (inviscid)
    eps: float = 1e-6
    kappa: float = 1.3
    gravity: tuple = (0.0, 0.0)
    pi_switch: bool = False
    eos: int = 0                # 0 isoThermal, 1 Tait
    diffusion: bool = True

    def arrays(self):
        prm = np.array([self.rho0, self.cs, self.delta, self.kernel_scale, self.visc_coeff, self.eps, self.kappa,
                        self.gravity[0], self.gravity[1]], np.float64)
        if not (self.viscosity in (-1, 0, 1)):
            raise ValueError("viscosity must be -1, 0 or 1")
        iprm = np.array([self.viscosity, 1 if self.pi_switch else 0, self.eos, 1 if self.diffusion else 0], np.int32)
        return prm, iprm


class OrcBoundary(C.Structure):
    _fields_ = [("off", C.c_void_p), ("nbr", C.c_void_p), ("pos", C.c_void_p), ("mass", C.c_void_p),
                ("rho", C.c_void_p), ("vel", C.c_void_p), ("visc_coeff", C.c_double), ("eps", C.c_double),
                ("h0", C.c_double), ("viscosity", C.c_int32), ("pi_switch", C.c_int32)]


@dataclass
class WallSet:
    """Wall particles seen by the fluid rows: CSR (off over the fluid rows,
    nbr wall indices), wall positions, mass, density, velocity, support and
    the wall viscosity (boundaryDiffusion)."""
    off: np.ndarray
    nbr: np.ndarray
    pos: np.ndarray
    mass: np.ndarray
    rho: np.ndarray
    vel: np.ndarray
    h: float
    viscosity: int = 0
    visc_coeff: float = 0.01
    eps: float = 1e-6
    pi_switch: bool = False

    def struct(self, dtype):
        cv = f32 if dtype == np.float32 else f64
        self._keep = [np.ascontiguousarray(self.off, np.int32), np.ascontiguousarray(self.nbr, np.int32),
                      cv(self.pos), cv(self.mass), cv(self.rho), cv(self.vel)]
        k = self._keep
        return OrcBoundary(_p(k[0]), _p(k[1]), _p(k[2]), _p(k[3]), _p(k[4]), _p(k[5]), self.visc_coeff, self.eps,
                           self.h, self.viscosity, 1 if self.pi_switch else 0)


def deltasph_normalization(dom, off, nbr, wpos, mass, rho, h0, dtype=np.float32, wall: WallSet | None = None):
    n = wpos.shape[0]
    cv = f32 if dtype == np.float32 else f64
    wpos, mass, rho = cv(wpos), cv(mass), cv(rho)
    bs = wall.struct(dtype) if wall is not None else None
    L = np.zeros((n, 4), dtype)
    ev = np.zeros((n, 2), dtype)
    g = np.zeros((n, 2), dtype)
    fn = lib().orc_deltasph_normalization_f32 if dtype == np.float32 else lib().orc_deltasph_normalization_f64
    hc = C.c_float if dtype == np.float32 else C.c_double
    fn(C.byref(dom), C.c_int64(n), _p(off), _p(nbr), _p(wpos), None, hc(h0), C.c_int32(0), _p(mass), _p(rho), _p(L),
       _p(ev), _p(g), C.byref(bs) if bs is not None else None)
    return L, ev, g


def deltasph_rhs(dom, off, nbr, wpos, vel, mass, rho, grad_rho, prm: DeltaParams, h0, dtype=np.float32,
                 wall: WallSet | None = None):
    n = wpos.shape[0]
    bs = wall.struct(dtype) if wall is not None else None
    cv = f32 if dtype == np.float32 else f64
    wpos, vel, mass, rho, grad_rho = cv(wpos), cv(vel), cv(mass), cv(rho), cv(grad_rho)
    dudt = np.zeros((n, 2), dtype)
    drhodt = np.zeros(n, dtype)
    terms = np.zeros((n, 8), dtype)
    pa, ia = prm.arrays()
    fn = lib().orc_deltasph_rhs_f32 if dtype == np.float32 else lib().orc_deltasph_rhs_f64
    hc = C.c_float if dtype == np.float32 else C.c_double
    fn(C.byref(dom), C.c_int64(n), _p(off), _p(nbr), _p(wpos), None, hc(h0), C.c_int32(0), _p(pa), _p(ia), _p(vel),
       _p(mass), _p(rho), _p(grad_rho), _p(dudt), _p(drhodt), _p(terms), C.byref(bs) if bs is not None else None)
    return dudt, drhodt, terms


def boundary_ghosts(bpos, dist, normals, h):
    """x_g = x_b + (d + max(d, -h)) n (v2/modules/mDBC.py:9-17), float32."""
    bpos, dist, normals = f32(bpos), f32(dist), f32(normals)
    hf = np.float32(h)
    md = dist + np.maximum(dist, -hf)
    return (bpos + md[:, None] * normals).astype(np.float32)


def mdbc(dom, off, nbr, ghost_pos, fluid_pos, h0, mass, rho, bpos, rho0, support, dtype=np.float32, variant=0):
    """(wall density, Shepard density).  variant 0: mDBCDensity
    (v2/modules/mDBC.py:50-136; pinv solve, no clamp); variant 1: the C++
    mDBC op (mDBCDensityKernel, v2/cppSrc/mDBC.h:129-244; Cramer's rule,
    Shepard for <= 5 neighbors, clamped at rho0)."""
    cv = f32 if dtype == np.float32 else f64
    hc = C.c_float if dtype == np.float32 else C.c_double
    gp, fp, mass, rho, bpos = cv(ghost_pos), cv(fluid_pos), cv(mass), cv(rho), cv(bpos)
    nb = gp.shape[0]
    out = np.zeros(nb, dtype)
    shep = np.zeros(nb, dtype)
    fn = lib().orc_mdbc_f32 if dtype == np.float32 else lib().orc_mdbc_f64
    fn(C.byref(dom), C.c_int64(nb), _p(off), _p(nbr), _p(gp), _p(fp), hc(h0), _p(mass), _p(rho), _p(bpos), hc(rho0),
       hc(support), _p(out), _p(shep), C.c_int32(variant))
    return out, shep


def wrap_positions(dom, pos, dtype=np.float64):
    """getPeriodicPosition (v2/modules/neighborhood.py:407-430) in `dtype`:
    remainder(x - min, L) + min on periodic axes (numpy's remainder has
    torch.remainder's sign convention)."""
    p = np.array(pos, dtype=dtype, copy=True)
    for k in range(2):
        if dom.periodic[k]:
            mn = dtype(dom.min[k])
            L = dtype(dom.max[k]) - mn
            p[:, k] = np.remainder(p[:, k] - mn, L) + mn
    return p


def neighbors_state_order(dom, pos, h):
    """Neighbor list with rows in the state's order and entries as state
    indices (cell-scan order within a row), on wrapped positions: the grid is
    built on `pos`, queried with the wrapped positions in state order, and
    the sorted slots are mapped back through sortIndex."""
    g = grid_build(dom, pos, h)
    wpos = np.empty_like(g.sorted_pos)
    wpos[g.sort_index] = g.sorted_pos
    cnt, off, nbr_s = neighbors_cell(dom, g, wpos, h0=h)
    return off, g.sort_index[nbr_s].astype(np.int32), wpos


@dataclass
class WallState:
    """Static wall particles in their own cell order, with ghosts."""
    pos: np.ndarray
    dist: np.ndarray
    normal: np.ndarray
    mass: np.ndarray
    vel: np.ndarray
    ghost: np.ndarray
    h: float
    viscosity: int = 0
    visc_coeff: float = 0.01
    eps: float = 1e-6
    pi_switch: bool = False
    mdbc_variant: int = 0


def make_walls(dom, pos, dist, normal, area, rho0, h, viscosity=0, visc_coeff=0.01, eps=1e-6, pi_switch=False,
               mdbc_variant=0):
    g = grid_build(dom, pos, h)
    si = g.sort_index
    p = g.sorted_pos.copy()
    d, nrm = f32(dist)[si], f32(normal)[si]
    n = p.shape[0]
    mass = (np.broadcast_to(f32(area), (n,))[si] * np.float32(rho0)).astype(np.float32)
    return WallState(p, d, nrm, mass, np.zeros((n, 2), np.float32), boundary_ghosts(p, d, nrm, h), h, viscosity,
                     visc_coeff, eps, pi_switch, mdbc_variant)


def deltaplus_rhs(dom, pos, vel, rho, mass, h, prm: DeltaParams, dtype=np.float32, walls: WallState | None = None):
    """One right-hand side: (dudt, drhodt, extras); dxdt is the velocity.
    With `walls`: the fluid->wall list, the ghost->fluid list, mDBC wall
    densities and the wall terms (v2/simulationSchemes/deltaPlus.py:186-329)."""
    g = grid_build(dom, pos, h)
    wpos = np.empty_like(g.sorted_pos)
    wpos[g.sort_index] = g.sorted_pos
    _, off, nbr_s = neighbors_cell(dom, g, wpos, h0=h)
    nbr = g.sort_index[nbr_s].astype(np.int32)
    if dtype != np.float32:
        # the float64 error reference wraps in float64 (the list itself keeps
        # the float32 predicate)
        wpos = wrap_positions(dom, pos, np.float64)
    wall = None
    extras = {}
    if walls is not None:
        gw = grid_build(dom, walls.pos, h)
        _, off_b, nbr_bs = neighbors_cell(dom, gw, wpos, h0=h)
        nbr_b = gw.sort_index[nbr_bs].astype(np.int32)
        _, off_g, nbr_gs = neighbors_cell(dom, g, walls.ghost, h0=h)
        nbr_g = g.sort_index[nbr_gs].astype(np.int32)
        rb, shep = mdbc(dom, off_g, nbr_g, walls.ghost, wpos, h, mass, rho, walls.pos, prm.rho0, h, dtype,
                        walls.mdbc_variant)
        wall = WallSet(off_b, nbr_b, walls.pos, walls.mass, rb, walls.vel, h, walls.viscosity, walls.visc_coeff,
                       walls.eps, walls.pi_switch)
        extras.update(wall_rho=rb, shepard=shep)
    L, ev, grad = deltasph_normalization(dom, off, nbr, wpos, mass, rho, h, dtype, wall=wall)
    dudt, drhodt, terms = deltasph_rhs(dom, off, nbr, wpos, vel, mass, rho, grad, prm, h, dtype, wall=wall)
    extras.update(L=L, ev=ev, grad_rho=grad, terms=terms, off=off, nbr=nbr)
    return dudt, drhodt, extras


@dataclass
class TimestepParams:
    support: float
    kernel_scale: float
    cs: float
    nu: float                  # the viscosity constraint's nu (timestep.py:13-14)
    cfl: float = 1.5
    max_dt: float = 0.1
    min_dt: float = 0.001
    viscosity_constraint: bool = True
    acceleration_constraint: bool = True
    acoustic_constraint: bool = True


def compute_timestep(tp: TimestepParams, dt_prev: float, max_accel=None) -> float:
    """computeTimestep (v2/modules/timestep.py:8-39) in float32 arithmetic;
    `max_accel` = max |dudt| (None before the first step: dt_a = maxDt)."""
    F = np.float32
    h, ks = F(tp.support), F(tp.kernel_scale)
    dt_v = F(0.125) * h * h / F(tp.nu) / (ks * ks)
    dt_c = F(tp.cfl) * h / F(tp.cs) / ks
    if max_accel is None:
        dt_a = F(tp.max_dt)
    else:
        dt_a = F(0.25) * np.sqrt(h / (F(max_accel) + F(1e-7))) / ks
    new = F(dt_prev)
    if tp.viscosity_constraint:
        new = dt_v
    if tp.acceleration_constraint:
        new = min(new, dt_a)
    if tp.acoustic_constraint:
        new = min(new, dt_c)
    new = min(new, F(tp.max_dt))
    new = max(new, F(tp.min_dt))
    return float(F(new))


@dataclass
class DeltaState:
    pos: np.ndarray
    vel: np.ndarray
    rho: np.ndarray
    mass: np.ndarray
    uid: np.ndarray


def rk4_step(st: DeltaState, dom, h, prm: DeltaParams, dt: float, reorder=True, dtype=np.float32, rhs=None,
             walls: WallState | None = None):
    """One RK4 step of the delta-SPH scheme (integration.py:231-307): the
    state is first sorted into cell order (as the GPU solver keeps it), then
    k1..k4 with stage states x0 + k dt/2, x0 + k dt/2, x0 + k dt, and
    state += (k1 + 2 k2 + 2 k3 + k4)/6 dt for x (k = velocity), u, rho.
    `rhs(dom, x, u, rho, mass, h, prm, dtype) -> (dudt, drhodt, extras)`
    defaults to :func:`deltaplus_rhs`.  Returns (state, averaged dudt)."""
    if rhs is None:
        def rhs(dom_, x_, u_, r_, m_, h_, prm_, dtype_):
            return deltaplus_rhs(dom_, x_, u_, r_, m_, h_, prm_, dtype_, walls=walls)
    F = np.float32 if dtype == np.float32 else np.float64
    if reorder:
        si = grid_build(dom, st.pos, h).sort_index
        st = DeltaState(st.pos[si], st.vel[si], st.rho[si], st.mass[si], st.uid[si])
    x0, u0, r0 = st.pos.astype(F), st.vel.astype(F), st.rho.astype(F)
    dtf = F(dt)
    half = F(dtf * F(0.5))
    ks = []
    x, u, r = x0, u0, r0
    for s, c in enumerate([half, half, dtf, None]):
        dudt, drhodt, _ = rhs(dom, x, u, r, st.mass.astype(F), h, prm, dtype)
        ks.append((u.copy(), dudt, drhodt))
        if c is not None:
            x = x0 + ks[-1][0] * c
            u = u0 + dudt * c
            r = r0 + drhodt * c

    def avg(a, b, c_, d):
        return (((a + F(2) * b) + F(2) * c_) + d) / F(6)

    dx = avg(*[k[0] for k in ks])
    du = avg(*[k[1] for k in ks])
    dr = avg(*[k[2] for k in ks])
    new = DeltaState(x0 + dx * dtf, u0 + du * dtf, r0 + dr * dtf, st.mass, st.uid)
    return new, du


# ---------------------------------------------------------------------------
# ISPH (v2/modules/incompressible.py, v2/sparse.py)
# ---------------------------------------------------------------------------
def isph_laplacian(dom, off, nbr, pos, h0, mass, rho, dt):
    """computeLaplacian (incompressible.py:102-133): per-entry values in CSR
    order (the self entry holds -a_ii), a_ii, preconditioner -1/a_ii."""
    n = pos.shape[0]
    val = np.zeros(int(off[-1]), np.float32)
    aii = np.zeros(n, np.float32)
    pre = np.zeros(n, np.float32)
    c2dt = float(np.float32(2.0 * dt))
    lib().orc_isph_laplacian_f32(C.byref(dom), C.c_int64(n), _p(off), _p(nbr), _p(f32(pos)), C.c_float(h0),
                                 _p(f32(mass)), _p(f32(rho)), C.c_float(c2dt), _p(val), _p(aii), _p(pre))
    return val, aii, pre


def spmv(off, nbr, val, x):
    n = len(off) - 1
    y = np.zeros(n, np.float32)
    lib().orc_spmv_f32(C.c_int64(n), _p(off), _p(nbr), _p(f32(val)), _p(f32(x)), _p(y))
    return y


def bicgstab(off, nbr, val, pre, b, x0, x0_scale=1.0, rtol=1e-3, atol=0.0, maxiter=512, fp64=False):
    """BiCGStab of v2/sparse.py:141-230 in the GPU's arithmetic (float32
    vectors, or float64 with `fp64`; float64 dots in the GPU's reduction
    order).  Returns (x, iterations, status, |r|, |b|); status 0 |r| < atol,
    1 |s| < atol, -10 / -11 breakdowns, 2 iteration limit."""
    n = len(off) - 1
    x = f32(x0).copy()
    it = C.c_int32()
    rn, bn = C.c_double(), C.c_double()
    st = lib().orc_bicgstab_f32(C.c_int64(n), _p(off), _p(nbr), _p(f32(val)), _p(f32(pre)), _p(f32(b)), _p(x),
                                C.c_float(x0_scale), C.c_double(rtol), C.c_double(atol), C.c_int32(maxiter),
                                C.byref(it), C.byref(rn), C.byref(bn), C.c_int32(1 if fp64 else 0))
    return x, it.value, st, rn.value, bn.value


def isph_source(dom, off, nbr, pos, h0, vel, adv, mass, rho, dt, rho0):
    """computeDivergenceSourceTerm (incompressible.py:148-165), mean removed."""
    n = pos.shape[0]
    B = np.zeros(n, np.float32)
    lib().orc_isph_source_f32(C.byref(dom), C.c_int64(n), _p(off), _p(nbr), _p(f32(pos)), C.c_float(h0),
                              _p(f32(vel)), _p(None if adv is None else f32(adv)), _p(f32(mass)), _p(f32(rho)),
                              C.c_float(dt), C.c_float(rho0), _p(B))
    return B


def isph_close(dom, off, nbr, wpos, h0, mass, rho, vel, adv, pstar, pss, dt, rho0, pos):
    """Closing passes of the ISPH step (incompressible.py:398-460); returns
    (v', shift, new positions, new velocities)."""
    n = wpos.shape[0]
    vp = np.zeros((n, 2), np.float32)
    sh = np.zeros((n, 2), np.float32)
    p_out = f32(pos).copy()
    v_out = np.zeros((n, 2), np.float32)
    lib().orc_isph_close_f32(C.byref(dom), C.c_int64(n), _p(off), _p(nbr), _p(f32(wpos)), C.c_float(h0),
                             _p(f32(mass)), _p(f32(rho)), _p(f32(vel)), _p(None if adv is None else f32(adv)),
                             _p(f32(pstar)), _p(f32(pss)), C.c_float(dt), C.c_float(rho0), _p(vp), _p(sh), _p(p_out),
                             _p(v_out))
    return vp, sh, p_out, v_out


@dataclass
class IsphState:
    pos: np.ndarray
    vel: np.ndarray
    mass: np.ndarray
    areas: np.ndarray
    uid: np.ndarray
    p_star: np.ndarray | None = None
    p_ss: np.ndarray | None = None


def isph_step(st: IsphState, dom: Domain, h: float, dt: float, rho0: float = 1.0, visc: Visc | None = Visc(),
              gravity=None, rtol=1e-3, maxiter=512, fp64=False, omega=0.5, thr=5e-4, min_iters=2, max_iters=256):
    """One ISPH step (v2/modules/incompressible.py:366-460) in float32,
    sorted into cell order first as the GPU solver keeps its state.
    Returns (state, info)."""
    g = grid_build(dom, st.pos, h)
    si = g.sort_index
    take = lambda a: None if a is None else a[si]  # noqa: E731
    st = IsphState(st.pos[si], st.vel[si], st.mass[si], st.areas[si], st.uid[si], take(st.p_star), take(st.p_ss))
    g = grid_build(dom, st.pos, h)
    _, off, nbr_s = neighbors_cell(dom, g, st.pos, h0=h)
    nbr = g.sort_index[nbr_s].astype(np.int32)
    wpos = np.empty_like(g.sorted_pos)
    wpos[g.sort_index] = g.sorted_pos
    wpos = f32(wpos)
    n = st.pos.shape[0]
    F = np.float32
    rho, vol = density(dom, off, nbr, wpos, wpos, st.mass, st.mass, h0=h)
    ad, an, ds, vs = dfsph_prepare(dom, off, nbr, wpos, st.vel, st.areas, vol, rho, dt, rho0, visc, h)
    if visc is None:
        vs = np.zeros_like(vs)
    adv = (np.asarray(gravity, np.float32)[None, :] + vs) if gravity is not None else vs.copy()
    lap, aii, pre = isph_laplacian(dom, off, nbr, wpos, h, st.mass, rho, dt)
    B = isph_source(dom, off, nbr, wpos, h, st.vel, adv, st.mass, rho, dt, rho0)
    x0 = st.p_star if st.p_star is not None else np.zeros(n, np.float32)
    p_star, it_b, status, rn, bn = bicgstab(off, nbr, lap, pre, B, x0, 0.5 if st.p_star is not None else 1.0,
                                            rtol, 0.0, maxiter, fp64)
    alpha = (an / F(dt)).astype(np.float32)
    p0 = (F(0.5) * f32(st.p_ss)) if st.p_ss is not None else np.zeros(n, np.float32)
    p_ss, _, res, it_j, err_j, _ = dfsph_solve(dom, off, nbr, wpos, h, dt, rho0, omega, thr, min_iters, max_iters,
                                               False, True, st.vel, adv, rho, vol, st.areas, alpha, None, p0,
                                               isph=True)
    vp, sh, pos, vel = isph_close(dom, off, nbr, wpos, h, st.mass, rho, st.vel, adv, p_star, p_ss, dt, rho0, st.pos)
    new = IsphState(pos, vel, st.mass, st.areas, st.uid, p_star, p_ss)
    info = dict(rho=rho, lap=lap, B=B, p_star=p_star, bicg_iters=it_b, bicg_status=status, bicg_residual=rn, p_ss=p_ss,
                jacobi_iters=it_j, jacobi_error=err_j, vprime=vp, shift=sh, off=off, nbr=nbr)
    return new, info


# ---------------------------------------------------------------------------
# Drop-in ops on the reference's per-pair arrays and the compact hash map
# ---------------------------------------------------------------------------
def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _u8(a):
    return np.ascontiguousarray(np.asarray(a, dtype=bool).astype(np.uint8))


def op_viscosity(vel_i, vel_j, m_j, rho_i, rho_j, h_i, idx_j, supports, grads, dists, vecs, nn, off, cs, rho0, alpha,
                 eps, pi_switch):
    """viscosityFn (partiBench/viscosityKernel.h:98-145) over CSR rows."""
    n = len(nn)
    out = np.zeros((n, 2), np.float32)
    lib().orc_op_viscosity(_p(f32(vel_i)), _p(f32(vel_j)), _p(f32(m_j)), _p(f32(rho_i)), _p(f32(rho_j)), _p(f32(h_i)),
                           _p(_i64(idx_j)), _p(f32(supports)), _p(f32(grads)), _p(f32(dists)), _p(f32(vecs)),
                           C.c_int32(n), _p(_i32(nn)), _p(_i32(off)), C.c_float(cs), C.c_float(rho0),
                           C.c_float(alpha), C.c_float(eps), C.c_int32(1 if pi_switch else 0), _p(out))
    return out


def op_mdbc(m_j, rho_j, pos_b, pos_g, idx_j, W, grads, dists, vecs, supports, nn, off, rho0):
    """mDBCDensityKernel (v2/cppSrc/mDBC.h:129-244); returns (mDBC, Shepard)."""
    nb = len(nn)
    out = np.zeros(nb, np.float32)
    shep = np.zeros(nb, np.float32)
    lib().orc_op_mdbc(_p(f32(m_j)), _p(f32(rho_j)), _p(f32(pos_b)), _p(f32(pos_g)), C.c_int32(nb), _p(_i64(idx_j)),
                      _p(f32(W)), _p(f32(grads)), _p(f32(dists)), _p(f32(vecs)), _p(f32(supports)), _p(_i32(nn)),
                      _p(_i32(off)), C.c_float(rho0), _p(out), _p(shep))
    return out, shep


def op_neighborhood(pos_i, h_i, pos_j, h_j, idx_j, nn, off, dmin, dmax, periodic):
    """computeNeighborhood, C++ form (v2/cppSrc/computeNeighborhood.cpp:5-94):
    returns [counts, offsets, i, j, rij, xij, hij] as the reference does."""
    n = len(nn)
    args = (_p(f32(pos_i)), _p(f32(h_i)), C.c_int32(n), _p(f32(pos_j)), _p(f32(h_j)), _p(_i64(idx_j)), _p(_i32(nn)),
            _p(_i32(off)), _p(f32(dmin)), _p(f32(dmax)), _p(_u8(periodic)))
    counts = np.zeros(n, np.int32)
    lib().orc_op_neighborhood(*args, _p(counts), None, None, None, None, None, None)
    new_off = np.concatenate([[0], np.cumsum(counts, dtype=np.int64)])[:-1].astype(np.int32)
    P = int(counts.sum())
    oi, oj = np.zeros(P, np.int64), np.zeros(P, np.int64)
    rij, xij, hij = np.zeros(P, np.float32), np.zeros((P, 2), np.float32), np.zeros(P, np.float32)
    lib().orc_op_neighborhood(*args, None, _p(new_off), _p(oi), _p(oj), _p(rij), _p(xij), _p(hij))
    return [counts, new_off, oi, oj, rij, xij, hij]


def pairs_filter(dom, mode, pos_i, h_i, pos_j, h_j, idx_j, nn, off):
    """computeNeighborhood (v2/modules/neighborhood.py:433-460) on a candidate
    CSR plus Wendland C2 kernels; returns (counts, offsets, i, j, rij, xij,
    hij, W, gradW)."""
    n = len(nn)
    args = (C.byref(dom), C.c_int32(mode), _p(f32(pos_i)), _p(f32(h_i)), C.c_int32(n), _p(f32(pos_j)), _p(f32(h_j)),
            _p(_i64(idx_j)), _p(_i32(nn)), _p(_i32(off)))
    counts = np.zeros(n, np.int32)
    lib().orc_pairs_filter(*args, _p(counts), None, None, None, None, None, None, None, None)
    new_off = np.concatenate([[0], np.cumsum(counts, dtype=np.int64)])[:-1].astype(np.int32)
    P = int(counts.sum())
    oi, oj = np.zeros(P, np.int64), np.zeros(P, np.int64)
    rij, xij, hij = np.zeros(P, np.float32), np.zeros((P, 2), np.float32), np.zeros(P, np.float32)
    W, gW = np.zeros(P, np.float32), np.zeros((P, 2), np.float32)
    lib().orc_pairs_filter(*args, None, _p(new_off), _p(oi), _p(oj), _p(rij), _p(xij), _p(hij), _p(W), _p(gW))
    return counts, new_off, oi, oj, rij, xij, hij, W, gW


def cell_offsets(search_range: int, dim: int = 2) -> np.ndarray:
    """The cell offset table of neighborhood.cpp:150-163, generated by the
    reference's own counter loop (row o, column dim-d-1)."""
    rows = (1 + 2 * search_range) ** dim
    off = np.zeros((rows, dim), np.int32)
    for d in range(dim):
        itr, ctr = -search_range, 0
        for o in range(rows):
            c = o % (1 + 2 * search_range) ** d
            if c == 0 and ctr > 0:
                itr += 1
            if itr > search_range:
                itr = -search_range
            off[o, dim - d - 1] = itr
            ctr += 1
    return off


def build_compact_hash_map(x, min_domain, max_domain, h_max, hash_map_length):
    """buildCompactHashMap (neighborhoodDebug.ipynb cell 4, with
    sortReferenceParticles / hashCellIndices of cell 3) in numpy int32 /
    float32 arithmetic.  Returns dict(sortIndex, sortedPositions, hashTable
    [H,2] int32, cellTable [C,3] int64, numCells [2] int32, hCell float,
    qMin)."""
    x = f32(x)
    mn, mx = f32(min_domain), f32(max_domain)
    ext = mx - mn
    ncf = np.floor(ext / np.float32(h_max))
    h_cell = np.float32(np.max(ext / ncf))                       # compute_h
    num_cells = np.ceil(ext / h_cell).astype(np.int32)
    idx = np.floor((x - mn) / h_cell).astype(np.int32)
    lin = idx[:, 0] + num_cells[0] * idx[:, 1]
    sort_index = np.argsort(lin, kind="stable")
    sorted_lin = lin[sort_index]
    sorted_pos = x[sort_index]
    cell_ids, first, counts = np.unique(sorted_lin, return_index=True, return_counts=True)
    cum = np.concatenate([[0], np.cumsum(counts)])[:-1]
    sorted_idx = np.floor((sorted_pos - mn) / h_cell).astype(np.int32)
    grid_idx = sorted_idx[cum]
    cell_table = np.stack([cell_ids.astype(np.int64), cum.astype(np.int64), counts.astype(np.int64)], axis=1)
    with np.errstate(over="ignore"):
        hsh = (grid_idx[:, 0] * np.int32(73856093) + grid_idx[:, 1] * np.int32(19349663)).astype(np.int32)
    hashed = np.mod(hsh, hash_map_length)                          # torch % : non-negative
    order = np.argsort(hashed, kind="stable")
    hm, hfirst, hcount = np.unique(hashed[order], return_index=True, return_counts=True)
    table = np.full((hash_map_length, 2), -1, np.int32)
    table[:, 1] = 0
    table[hm, 0] = hfirst
    table[hm, 1] = hcount
    return dict(sortIndex=sort_index, sortedPositions=sorted_pos, hashTable=table,
                cellTable=np.ascontiguousarray(cell_table[order]), numCells=num_cells, hCell=float(h_cell), qMin=mn)


def hash_neighbors(qpos, qh, search_range, spos, sh, table, ncell, cells, hcell, dmin, dmax, periodic, mode):
    """countNeighbors + buildNeighborList (neighborhood.cpp:65-435) in the
    reference's CPU order; returns (counts, i, j) with j in sorted order."""
    nq = qpos.shape[0]
    offs = cell_offsets(search_range)
    mode_i = {"scatter": 0, "gather": 1, "symmetric": 2}[mode]
    args = (_p(f32(qpos)), _p(None if qh is None else f32(qh)), C.c_int32(nq), _p(offs), C.c_int32(offs.shape[0]),
            _p(f32(spos)), _p(None if sh is None else f32(sh)), _p(_i32(table)), C.c_int32(table.shape[0]),
            _p(_i32(ncell)), _p(_i64(cells)), C.c_float(hcell), _p(f32(dmin)), _p(f32(dmax)), _p(_u8(periodic)),
            C.c_int32(mode_i))
    counts = np.zeros(nq, np.int32)
    lib().orc_hash_neighbors(*args, _p(counts), None, None, None)
    list_off = np.concatenate([[0], np.cumsum(counts)])[:-1].astype(np.int32)
    P = int(counts.sum())
    oi, oj = np.zeros(P, np.int32), np.zeros(P, np.int32)
    lib().orc_hash_neighbors(*args, None, _p(list_off), _p(oi), _p(oj))
    return counts, oi, oj


def deltaplus_shift(dom, off, nbr, wpos, h0, mass, rho, vel, pos, dx, support, kernel_scale, R=0.25, n=4.0,
                    cfl=1.5, cs=10.0, threshold=0.5, mach=-1.0, wall: WallSet | None = None, dtype=np.float32):
    """deltaPlusShifting (v2/modules/shifting.py:119-147) as solveShifting
    applies it (scheme 'deltaSPH', one iteration, no free surface) and the
    runner adds it; returns (new positions, added displacement)."""
    nq = wpos.shape[0]
    cv = f32 if dtype == np.float32 else f64
    hc = C.c_float if dtype == np.float32 else C.c_double
    wpos, mass, rho, vel = cv(wpos), cv(mass), cv(rho), cv(vel)
    p = cv(pos).copy()
    sh = np.zeros((nq, 2), dtype)
    prm = np.array([dx, support, kernel_scale, R, n, cfl, cs, threshold, mach], np.float64)
    bs = wall.struct(dtype) if wall is not None else None
    fn = lib().orc_deltaplus_shift_f32 if dtype == np.float32 else lib().orc_deltaplus_shift_f64
    fn(C.byref(dom), C.c_int64(nq), _p(np.ascontiguousarray(off, np.int32)), _p(np.ascontiguousarray(nbr, np.int32)),
       _p(wpos), hc(h0), _p(prm), _p(mass), _p(rho), _p(vel), C.byref(bs) if bs is not None else None, _p(p), _p(sh))
    return p, sh
