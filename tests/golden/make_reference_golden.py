"""Write golden vectors of the reference's delta-SPH and DFSPH pair modules.

Run in a checkout that has the reference at /root/reference (not needed
by the tests, which only read the .npz this writes):

    python tests/golden/make_reference_golden.py

It loads the reference's own module files by path (src/diffSPH/v2/...),
skipping the package __init__ of v2/modules (which would import the
unavailable torchCompactRadius), builds a small periodic particle set and
its neighborDict per-pair arrays with the reference's `mod` and Wendland C2
functions (the pair list itself comes from the brute-force predicate
r/h <= 1, as v2/modules/neighborhood.py:433-460 filters), and records the
modules' outputs in float32:

  computeCovarianceMatrices + identity for < 4 neighbors + pinv2x2
  renormalizedDensityGradient, computeDensityDeltaTerm (deltaSPH)
  computeViscosity (deltaSPH_viscid and deltaSPH_inviscid)
  computeMomentumEquation, computeEOS (isoThermal), computePressureAccel
  computeTimestep
  DFSPH: computeAlpha, computeSourceTerm (divergence and density),
         computePressureAcceleration + updatePressure (one iteration)

and, in a second file, the reference's autograd gradient of the final
kinetic energy with respect to the initial velocities and positions through
a few RK4 steps of the Taylor-Green delta-SPH scheme (float64, fixed time
steps), composing those modules per stage exactly as deltaPlus.py:176-350
and integration.py:231-307 do for a periodic fluid without walls.
"""
import importlib
import os
import sys
import types

import numpy as np
import torch

REF = "/root/reference/src"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_modules_16.npz")
OUT_GRAD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_rollout_grad_16.npz")


def load():
    sys.path.insert(0, REF)
    pkg = types.ModuleType("diffSPH.v2.modules")
    pkg.__path__ = [os.path.join(REF, "diffSPH", "v2", "modules")]
    importlib.import_module("diffSPH.v2")
    sys.modules["diffSPH.v2.modules"] = pkg
    mods = {}
    for name in ["normalizationMatrices", "densityDiffusion", "viscosity", "momentumEquation", "pressureEOS",
                 "pressureForce", "timestep", "dfsph"]:
        mods[name] = importlib.import_module("diffSPH.v2.modules." + name)
    mods["math"] = importlib.import_module("diffSPH.v2.math")
    mods["W2"] = importlib.import_module("diffSPH.kernelFunctions.Wendland2")
    return mods


def main():
    torch.manual_seed(0)
    rng = np.random.default_rng(16)
    M = load()
    ns = 16
    dx = 2.0 / ns
    xs = np.linspace(-1 + dx / 2, 1 - dx / 2, ns, dtype=np.float32)
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    pos = np.stack([X.ravel(), Y.ravel()], 1).astype(np.float32)
    pos = (pos + rng.uniform(-0.15 * dx, 0.15 * dx, pos.shape)).astype(np.float32)
    n = pos.shape[0]
    h = np.float32(np.sqrt(50.0 * dx * dx / np.pi))
    vel = (np.stack([np.sin(np.pi * pos[:, 0]) * np.cos(np.pi * pos[:, 1]),
                     -np.cos(np.pi * pos[:, 0]) * np.sin(np.pi * pos[:, 1])], 1)
           + 0.1 * rng.standard_normal((n, 2))).astype(np.float32)
    rho = (1.0 + 0.01 * rng.standard_normal(n)).astype(np.float32)
    mass = np.full(n, dx * dx, np.float32)

    T = torch.from_numpy
    p = T(pos)
    # brute-force pair list (i ascending, j ascending) with the reference's
    # pair geometry (neighborhood.py:445-458)
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    i_all, j_all = T(ii.ravel()), T(jj.ravel())
    xij = p[i_all] - p[j_all]
    xij = torch.stack([M["math"].mod(xij[:, k], -1.0, 1.0) for k in range(2)], -1)
    rij = torch.linalg.norm(xij, dim=-1)
    xij = xij / (rij + 1e-7).view(-1, 1)
    hv = torch.full((n,), float(h))
    rij = rij / hv[j_all]
    keep = rij <= 1.0
    i, j = i_all[keep], j_all[keep]
    rij, xij, hij = rij[keep], xij[keep], hv[j][:]
    W = M["W2"].kernel(rij, hij, 2)
    gW = M["W2"].kernelGradient(rij, xij, hij, 2)
    numNeighbors = torch.bincount(i, minlength=n).to(torch.int32)
    nd = {"indices": (i, j), "distances": rij, "vectors": xij, "supports": hij, "kernels": W, "gradients": gW,
          "numNeighbors": numNeighbors}
    ks = torch.tensor(float(h / (2 * dx)))
    cfg = {"diffusion": {"delta": 0.1, "densityScheme": "deltaSPH", "velocityScheme": "deltaSPH_viscid",
                         "nu": 0.01, "alpha": 0.01, "eps": 1e-6, "pi-switch": False},
           "kernel": {"kernelScale": ks}, "fluid": {"rho0": 1.0, "cs": 10.0}, "domain": {"dim": 2},
           "EOS": {"type": "isoThermal"},
           "timestep": {"active": True, "dt": 0.001, "CFL": 1.5, "maxDt": 0.1, "minDt": 0.001,
                        "viscosityConstraint": True, "accelerationConstraint": True,
                        "acousticConstraint": True},
           "particle": {"support": torch.tensor(float(h))}}
    st = {"positions": p, "velocities": T(vel), "masses": T(mass), "densities": T(rho), "supports": hv,
          "numParticles": n, "numNeighbors": numNeighbors}
    out = {}
    C = M["normalizationMatrices"].computeCovarianceMatrices(st, st, nd, cfg)
    C[numNeighbors < 4] = torch.eye(2)[None]
    L, ev = M["math"].pinv2x2(C)
    st["L"] = L
    out["covariance"], out["L"], out["L_ev"] = C, L, ev
    st["gradRho^L"] = M["densityDiffusion"].renormalizedDensityGradient(st, st, nd, cfg)
    out["grad_rho"] = st["gradRho^L"]
    out["diffusion"] = M["densityDiffusion"].computeDensityDeltaTerm(st, st, nd, cfg)
    out["visc_viscid"] = M["viscosity"].computeViscosity(st, st, nd, cfg)
    cfg["diffusion"]["velocityScheme"] = "deltaSPH_inviscid"
    out["visc_inviscid"] = M["viscosity"].computeViscosity(st, st, nd, cfg)
    out["continuity"] = M["momentumEquation"].computeMomentumEquation(st, st, nd, cfg)
    st["pressures"] = M["pressureEOS"].computeEOS(st, cfg)
    out["pressure"] = st["pressures"]
    out["pressure_accel"] = M["pressureForce"].computePressureAccel(st, st, nd, cfg)
    cfg["diffusion"]["velocityScheme"] = "deltaSPH_viscid"
    dudt = out["pressure_accel"] + out["visc_viscid"]
    state = {"fluid": {"dudt": dudt}}
    out["dt_next"] = torch.tensor(float(M["timestep"].computeTimestep(state, cfg)))
    out["dt_first"] = torch.tensor(float(M["timestep"].computeTimestep({"fluid": {}}, cfg)))
    # DFSPH pieces (v2/modules/dfsph.py): alpha, sources, one Jacobi
    # iteration of each solve and the pressure acceleration after it
    D = M["dfsph"]
    dt = float(np.float32(3.0 * dx / 10.0))
    cfg["timestep"]["dt"] = dt
    cfg["dfsph"] = {"omega": 0.3, "clampPressure": True}
    adv = T((0.5 * rng.standard_normal((n, 2))).astype(np.float32))
    p0 = T((0.5 + 0.1 * rng.standard_normal(n)).astype(np.float32))
    d = dict(st)
    d["areas"] = T(mass) / 1.0
    d["actualArea"] = d["masses"] / d["densities"]
    for solve in ("div", "den"):
        dens = solve == "den"
        d["predictedVelocities"] = d["velocities"] + dt * adv if dens else d["velocities"]
        alpha = D.computeAlpha(d, d, cfg, nd, density=dens)
        if dens:
            src = 1 - d["areas"] / d["actualArea"] - D.computeSourceTerm(d, d, cfg, nd, density=True)
        else:
            src = D.computeSourceTerm(d, d, cfg, nd, density=False)
        d["alpha"], d["sourceTerm"] = alpha, src
        d["pressureB"], d["pressureA"] = p0.clone(), p0.clone()
        d["pressureAccel"] = D.computePressureAcceleration(d, d, cfg, nd)
        p1, res = D.updatePressure(d, d, cfg, nd, density=dens)
        d["pressureB"] = p1
        out[f"dfsph_{solve}_alpha"] = alpha
        out[f"dfsph_{solve}_source"] = src
        out[f"dfsph_{solve}_p1"] = p1
        out[f"dfsph_{solve}_residual"] = res
        out[f"dfsph_{solve}_accel"] = D.computePressureAcceleration(d, d, cfg, nd)
    out["dfsph_adv"] = adv
    out["dfsph_p0"] = p0
    out["dfsph_dt"] = torch.tensor(dt)
    np_out = {k: v.detach().numpy().astype(np.float32) for k, v in out.items()}
    np.savez(OUT, pos=pos, vel=vel, rho=rho, mass=mass, h=np.float32(h), dx=np.float32(dx),
             kernel_scale=np.float32(float(ks)), i=i.numpy().astype(np.int32), j=j.numpy().astype(np.int32), **np_out)
    print("wrote", OUT, {k: v.shape for k, v in np_out.items()})


def rollout_grad(steps=4):
    """Autograd gradient of E_k(T) = sum 0.5 m |u|^2 through `steps` RK4
    steps, float64."""
    M = load()
    torch.set_default_dtype(torch.float64)
    rng = np.random.default_rng(7)
    ns = 16
    dx = 2.0 / ns
    xs = np.linspace(-1 + dx / 2, 1 - dx / 2, ns)
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    pos = np.stack([X.ravel(), Y.ravel()], 1)
    pos = (pos + rng.uniform(-0.05 * dx, 0.05 * dx, pos.shape)).astype(np.float32).astype(np.float64)
    n = pos.shape[0]
    h = float(np.float32(np.sqrt(50.0 * dx * dx / np.pi)))
    vel = np.stack([np.sin(np.pi * pos[:, 0]) * np.cos(np.pi * pos[:, 1]),
                    -np.cos(np.pi * pos[:, 0]) * np.sin(np.pi * pos[:, 1])], 1).astype(np.float32).astype(np.float64)
    mass = np.full(n, dx * dx)
    ks = h / (2 * dx)
    cfg = {"diffusion": {"delta": 0.1, "densityScheme": "deltaSPH", "velocityScheme": "deltaSPH_viscid",
                         "nu": 0.01, "alpha": 0.01, "eps": 1e-6, "pi-switch": False},
           "kernel": {"kernelScale": ks}, "fluid": {"rho0": 1.0, "cs": 10.0}, "domain": {"dim": 2},
           "EOS": {"type": "isoThermal"}}
    hv = torch.full((n,), h)
    m = torch.tensor(mass)

    def rhs(x, u, rho):
        xw = torch.remainder(x + 1.0, 2.0) - 1.0
        with torch.no_grad():
            d = xw[:, None, :] - xw[None, :, :]
            d = torch.stack([M["math"].mod(d[..., k], -1.0, 1.0) for k in range(2)], -1)
            keep = torch.linalg.norm(d, dim=-1) / h <= 1.0
            i, j = torch.nonzero(keep, as_tuple=True)
        xij = xw[i] - xw[j]
        xij = torch.stack([M["math"].mod(xij[:, k], -1.0, 1.0) for k in range(2)], -1)
        rij = torch.linalg.norm(xij, dim=-1)
        xij = xij / (rij + 1e-7).view(-1, 1)
        rij = rij / hv[j]
        hij = hv[j]
        nd = {"indices": (i, j), "distances": rij, "vectors": xij, "supports": hij,
              "kernels": M["W2"].kernel(rij, hij, 2), "gradients": M["W2"].kernelGradient(rij, xij, hij, 2)}
        nn = torch.bincount(i, minlength=n)
        st = {"positions": xw, "velocities": u, "masses": m, "densities": rho, "supports": hv, "numParticles": n,
              "numNeighbors": nn}
        Cm = M["normalizationMatrices"].computeCovarianceMatrices(st, st, nd, cfg)
        Cm = torch.where((nn < 4)[:, None, None], torch.eye(2)[None], Cm)
        st["L"], _ = M["math"].pinv2x2(Cm)
        st["gradRho^L"] = M["densityDiffusion"].renormalizedDensityGradient(st, st, nd, cfg)
        D = M["densityDiffusion"].computeDensityDeltaTerm(st, st, nd, cfg)
        visc = M["viscosity"].computeViscosity(st, st, nd, cfg)
        cont = M["momentumEquation"].computeMomentumEquation(st, st, nd, cfg)
        st["pressures"] = M["pressureEOS"].computeEOS(st, cfg)
        ap = M["pressureForce"].computePressureAccel(st, st, nd, cfg)
        rhs.terms = dict(ap=ap, visc=visc, cont=cont, D=D, grad=st["gradRho^L"], L=st["L"])
        return u, ap + visc, cont + D

    x0 = torch.tensor(pos, requires_grad=True)
    u0 = torch.tensor(vel, requires_grad=True)
    r0 = torch.ones(n, requires_grad=True)
    with torch.no_grad():
        rhs(x0, u0, r0)
    first = {f"rhs0_{k}": v.detach().numpy() for k, v in rhs.terms.items()}
    dts = [0.01, 0.01, 0.008, 0.008, 0.006, 0.006, 0.006, 0.006][:steps]
    x, u, r = x0, u0, r0
    for dt in dts:
        k1 = rhs(x, u, r)
        k2 = rhs(x + dt / 2 * k1[0], u + dt / 2 * k1[1], r + dt / 2 * k1[2])
        k3 = rhs(x + dt / 2 * k2[0], u + dt / 2 * k2[1], r + dt / 2 * k2[2])
        k4 = rhs(x + dt * k3[0], u + dt * k3[1], r + dt * k3[2])
        x = x + dt * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0]) / 6
        u = u + dt * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1]) / 6
        r = r + dt * (k1[2] + 2 * k2[2] + 2 * k3[2] + k4[2]) / 6
    E = 0.5 * (m * (u * u).sum(1)).sum()
    E.backward()
    np.savez(OUT_GRAD, pos=pos, vel=vel, mass=mass, h=h, dx=dx, dts=np.array(dts), E=E.item(),
             grad_u0=u0.grad.numpy(), grad_x0=x0.grad.numpy(), grad_rho0=r0.grad.numpy(),
             u_T=u.detach().numpy(), x_T=x.detach().numpy(), rho_T=r.detach().numpy(), **first)
    print("wrote", OUT_GRAD, "E", E.item(), "|grad u0|", float(u0.grad.norm()))
    torch.set_default_dtype(torch.float32)


OUT_OPS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_sphops_16.npz")


def sph_operations_and_shifting():
    """Every mode of the reference's generic operator (v2/sphOps.py:253-282:
    interpolate, gradient naive/difference/summation/symmetric, divergence
    and divergenceConsistent in the same four modes) on scalar and vector
    fields with non-uniform masses and densities, and deltaPlusShifting
    (v2/modules/shifting.py:119-147) with and without the Mach number from
    the velocities, on a jittered periodic 16x16 set."""
    sys.path.insert(0, REF)
    ops = importlib.import_module("diffSPH.v2.sphOps")
    M = load()
    shifting_src = os.path.join(REF, "diffSPH", "v2", "modules", "shifting.py")
    # shifting.py imports v2.modules.neighborhood (torchCompactRadius,
    # unavailable) and surfaceDetection at module level; only
    # deltaPlusShifting is needed, so its source lines are executed alone,
    # with the names it uses.
    src = open(shifting_src).read().splitlines()
    a = next(k for k, l in enumerate(src) if l.startswith("def deltaPlusShifting"))
    b = next(k for k in range(a + 1, len(src)) if src[k].startswith("def "))
    ns_ = {"torch": torch, "scatter_sum": M["math"].scatter_sum,
           "record_function": torch.profiler.record_function}
    exec(compile("\n".join(src[a:b]), shifting_src, "exec"), ns_)
    deltaPlusShifting = ns_["deltaPlusShifting"]

    rng = np.random.default_rng(23)
    ns = 16
    dx = 2.0 / ns
    xs = np.linspace(-1 + dx / 2, 1 - dx / 2, ns, dtype=np.float32)
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    pos = np.stack([X.ravel(), Y.ravel()], 1).astype(np.float32)
    pos = (pos + rng.uniform(-0.2 * dx, 0.2 * dx, pos.shape)).astype(np.float32)
    n = pos.shape[0]
    h = np.float32(np.sqrt(50.0 * dx * dx / np.pi))
    mass = (dx * dx * (1.0 + 0.2 * rng.uniform(-1, 1, n))).astype(np.float32)
    rho = (1.0 + 0.05 * rng.standard_normal(n)).astype(np.float32)
    qs = (np.sin(2 * pos[:, 0]) + pos[:, 1] ** 2 + 0.1 * rng.standard_normal(n)).astype(np.float32)
    qv = np.stack([np.cos(pos[:, 1]) + 0.1 * rng.standard_normal(n), pos[:, 0] * pos[:, 1]], 1).astype(np.float32)
    vel = (0.5 * rng.standard_normal((n, 2))).astype(np.float32)

    T = torch.from_numpy
    p = T(pos)
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    i_all, j_all = T(ii.ravel()), T(jj.ravel())
    xij = p[i_all] - p[j_all]
    xij = torch.stack([M["math"].mod(xij[:, k], -1.0, 1.0) for k in range(2)], -1)
    rij = torch.linalg.norm(xij, dim=-1)
    xij = xij / (rij + 1e-7).view(-1, 1)
    hv = torch.full((n,), float(h))
    rij = rij / hv[j_all]
    keep = rij <= 1.0
    i, j = i_all[keep], j_all[keep]
    rij, xij, hij = rij[keep], xij[keep], hv[j]
    W = M["W2"].kernel(rij, hij, 2)
    gW = M["W2"].kernelGradient(rij, xij, hij, 2)
    m_, r_ = T(mass), T(rho)
    S, V = (T(qs), T(qs)), (T(qv), T(qv))
    out = {}

    def op(q, operation, gm="symmetric", dm="div"):
        return ops.sphOperation((m_, m_), (r_, r_), q, (i, j), W, gW, rij, xij, hij, n, operation=operation,
                                gradientMode=gm, divergenceMode=dm)

    out["interp_scalar"] = op(S, "interpolate")
    out["interp_vector"] = op(V, "interpolate")
    for gm in ("naive", "difference", "summation", "symmetric"):
        out[f"grad_{gm}_scalar"] = op(S, "gradient", gm)
        out[f"grad_{gm}_vector"] = op(V, "gradient", gm)
        out[f"div_{gm}"] = op(V, "divergence", gm)
        out[f"divc_{gm}"] = op(V, "divergenceConsistent", gm)
    for gm in ("naive", "difference", "summation", "symmetric"):
        out[f"curl_{gm}"] = op(V, "curl", gm)
    # 2x2 tensor field: its gradient and both divergence contractions
    Tq = torch.einsum("ni,nj->nij", T(qv), T(qv)) + 0.1 * torch.eye(2)[None]
    TT = (Tq, Tq)
    out["qt"] = Tq
    out["grad_difference_tensor"] = op(TT, "gradient", "difference")
    out["div_difference_tensor_div"] = op(TT, "divergence", "difference", "div")
    out["div_difference_tensor_dot"] = op(TT, "divergence", "difference", "dot")
    # Laplacian branches (v2/sphOps.py:191-250)
    lapk = M["W2"].kernelLaplacian(rij, hij, 2)
    out["laplacian_kernels"] = lapk

    def lap(q, gm, lk=None):
        return ops.sphOperation((m_, m_), (r_, r_), q, (i, j), W, gW, rij, xij, hij, n, operation="laplacian",
                                gradientMode=gm, kernelLaplacians=lk)

    import contextlib
    import io
    with contextlib.redirect_stdout(io.StringIO()):  # the naive branch prints
        out["lap_naive_scalar"] = lap(S, "naive", lapk)
        out["lap_naive_vector"] = lap(V, "naive", lapk)
    for gm in ("difference", "conserving", "divergenceFree", "dot"):
        out[f"lap_{gm}_scalar"] = lap(S, gm)
        out[f"lap_{gm}_vector"] = lap(V, gm)
    out["direct_laplacian_difference"] = op(S, "directLaplacian", "difference")
    st = {"masses": m_, "densities": r_, "numParticles": n}
    nd = {"indices": (i, j), "kernels": W, "gradients": gW, "distances": rij, "vectors": xij, "supports": hij}
    out["states_density"] = ops.sphOperationStates(st, st, None, nd, operation="density")
    out["states_grad_difference"] = ops.sphOperationStates(st, st, S, nd, operation="gradient",
                                                           gradientMode="difference")

    ks = float(h / (2 * dx))
    for mach in (True, False):
        cfg = {"kernel": {"function": M["W2"], "kernelScale": ks},
               "particle": {"dx": dx, "support": float(h)}, "domain": {"dim": 2},
               "shifting": {"R": 0.25, "n": 4, "CFL": 1.5, "computeMach": mach}, "fluid": {"cs": 10.0}}
        ps = {"neighborhood": {"indices": (i, j), "kernels": W, "gradients": gW}, "masses": m_, "densities": r_,
              "velocities": T(vel), "supports": hv, "positions": p}
        out[f"deltaplus_mach{int(mach)}"] = deltaPlusShifting(ps, cfg)
    np_out = {k: v.detach().numpy().astype(np.float32) for k, v in out.items()}
    np.savez(OUT_OPS, pos=pos, mass=mass, rho=rho, qs=qs, qv=qv, vel=vel, h=np.float32(h), dx=np.float32(dx),
             kernel_scale=np.float32(ks), i=i.numpy().astype(np.int32), j=j.numpy().astype(np.int32), **np_out)
    print("wrote", OUT_OPS, {k: v.shape for k, v in np_out.items()})


OUT_IPS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_ips_16.npz")


def _shifting_functions(M):
    """The IPS functions of v2/modules/shifting.py executed from their source
    lines (the module imports the unavailable torchCompactRadius and
    surfaceDetection at top level)."""
    sparse = importlib.import_module("diffSPH.v2.sparse")
    src_path = os.path.join(REF, "diffSPH", "v2", "modules", "shifting.py")
    src = open(src_path).read().splitlines()
    ns_ = {"torch": torch, "scatter_sum": M["math"].scatter_sum, "record_function": torch.profiler.record_function,
           "matvec_sparse_coo": sparse.matvec_sparse_coo, "make_id": sparse.make_id,
           "_get_tensor_eps": sparse._get_tensor_eps, "_get_atol_rtol": sparse._get_atol_rtol,
           "Optional": __import__("typing").Optional, "Tuple": __import__("typing").Tuple,
           "Callable": __import__("typing").Callable}
    for name in ("evalKernel", "deltaPlusShifting", "bicgstab_shifting", "getShiftingMatrices",
                 "multiplySparseShifting", "computeShifting"):
        a = next(k for k, l in enumerate(src) if l.startswith(f"def {name}("))
        b = next((k for k in range(a + 1, len(src)) if src[k] and not src[k][0].isspace()
                  and not src[k].startswith("#")), len(src))
        exec(compile("\n".join(src[a:b]), src_path, "exec"), ns_)
    return ns_


def ips_shifting():
    """Implicit particle shifting (computeShifting with scheme IPS,
    v2/modules/shifting.py:192-284, 361-460; the BiCGStab of :318-...) on a
    jittered periodic 16x16 set: the kernel Hessians, the assembled J and B,
    and the shift for the reference's default solver threshold (-0.5: one
    BiCGStab iteration) and for a positive one (the solve runs to rtol)."""
    sys.path.insert(0, REF)
    M = load()
    kern = importlib.import_module("diffSPH.kernels")
    fns = _shifting_functions(M)
    rng = np.random.default_rng(31)
    ns = 16
    dx = 2.0 / ns
    xs = np.linspace(-1 + dx / 2, 1 - dx / 2, ns, dtype=np.float32)
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    pos = np.stack([X.ravel(), Y.ravel()], 1).astype(np.float32)
    pos = (pos + rng.uniform(-0.2 * dx, 0.2 * dx, pos.shape)).astype(np.float32)
    n = pos.shape[0]
    h = np.float32(np.sqrt(50.0 * dx * dx / np.pi))
    T = torch.from_numpy
    p = T(pos)
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    i_all, j_all = T(ii.ravel()), T(jj.ravel())
    xij = p[i_all] - p[j_all]
    xij = torch.stack([M["math"].mod(xij[:, k], -1.0, 1.0) for k in range(2)], -1)
    rij = torch.linalg.norm(xij, dim=-1)
    xij = xij / (rij + 1e-7).view(-1, 1)
    hv = torch.full((n,), float(h))
    rij = rij / hv[j_all]
    keep = rij <= 1.0
    i, j = i_all[keep], j_all[keep]
    rij, xij, hij = rij[keep], xij[keep], hv[j]
    W = M["W2"].kernel(rij, hij, 2)
    gW = M["W2"].kernelGradient(rij, xij, hij, 2)
    area = torch.full((n,), dx * dx)
    out = {}
    wrapper = kern.KernelWrapper(M["W2"])
    K, J, H = fns["evalKernel"](rij, xij, hij, wrapper, 2)
    out["hessian"] = H
    for label, thr in (("default", -0.5), ("solve", 1.0)):
        cfg = {"shifting": {"freeSurface": False, "summationDensity": False, "initialization": "zero",
                            "tol": 1e-4, "rtol": 1e-4, "maxSolveIter": 64, "preconditioner": "Jacobi",
                            "verbose": False, "solverThreshold": thr, "solver": "BiCGStab_wJacobi",
                            "R": 0.25, "n": 4, "CFL": 1.5, "computeMach": True},
               "kernel": {"function": wrapper, "kernelScale": float(h / (2 * dx))}, "domain": {"dim": 2},
               "particle": {"dx": dx, "support": float(h)}, "fluid": {"rho0": 1.0, "cs": 10.0}}
        ps = {"numParticles": n, "positions": p, "areas": area, "masses": area * 1.0,
              "densities": torch.ones(n), "supports": hv, "boundaryMarker": torch.zeros(n, dtype=torch.int32),
              "neighborhood": {"indices": (i, j), "distances": rij, "vectors": xij, "supports": hij, "kernels": W,
                               "gradients": gW}}
        update, K_, J_, H_, B, conv, iters, residual = fns["computeShifting"](ps, cfg, computeRho=False)
        out[f"{label}_update"] = update
        out[f"{label}_J"] = J_
        out[f"{label}_B"] = B
        out[f"{label}_iters"] = torch.tensor(float(iters))
        out[f"{label}_residual"] = torch.tensor(float(residual))
    np_out = {k: v.detach().numpy().astype(np.float32) for k, v in out.items()}
    np.savez(OUT_IPS, pos=pos, h=np.float32(h), dx=np.float32(dx), i=i.numpy().astype(np.int32),
             j=j.numpy().astype(np.int32), **np_out)
    print("wrote", OUT_IPS, {k: (v.shape, v.ravel()[:1]) for k, v in np_out.items()})


OUT_NOISE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_noise.npz")


def octave_noise():
    """Perlin octave noise of v2/noiseFunctions (generateOctaveNoise with
    kind 'perlin', generator.py:87-99, over perlinNoise2D, perlin.py): the
    two functions of generator.py executed from their source lines (the
    module imports numba, unavailable here, for its simplex kinds)."""
    sys.path.insert(0, REF)
    perlin = importlib.import_module("diffSPH.v2.noiseFunctions.perlin")
    src_path = os.path.join(REF, "diffSPH", "v2", "noiseFunctions", "generator.py")
    src = open(src_path).read().splitlines()
    ns_ = {"torch": torch, "np": np, "perlinNoise1D": perlin.perlinNoise1D, "perlinNoise2D": perlin.perlinNoise2D,
           "perlinNoise3D": perlin.perlinNoise3D, "interpolant": perlin.interpolant}
    for name in ("generatePerlin", "generateOctaveNoise"):
        a = next(k for k, l in enumerate(src) if l.startswith(f"def {name}("))
        b = next((k for k in range(a + 1, len(src)) if src[k] and not src[k][0].isspace()), len(src))
        exec(compile("\n".join(src[a:b]), src_path, "exec"), ns_)
    out = {}
    import contextlib
    import io
    with contextlib.redirect_stdout(io.StringIO()):  # perlinNoise2D prints its arguments
        cases = {"a": dict(n=64, octaves=3, lacunarity=2, persistence=0.5, baseFrequency=2, tileable=True,
                           seed=1234, normalized=True),
                 "b": dict(n=48, octaves=2, lacunarity=2, persistence=0.7, baseFrequency=3, tileable=False,
                           seed=7, normalized=False),
                 "c": dict(n=32, octaves=1, lacunarity=2, persistence=0.5, baseFrequency=4, tileable=True,
                           seed=99, normalized=True)}
        for key, kw in cases.items():
            xx, yy, noise = ns_["generateOctaveNoise"](kw["n"], dim=2, octaves=kw["octaves"],
                                                       lacunarity=kw["lacunarity"], persistence=kw["persistence"],
                                                       baseFrequency=kw["baseFrequency"], tileable=kw["tileable"],
                                                       kind="perlin", seed=kw["seed"], normalized=kw["normalized"])
            out[f"{key}_noise"] = noise.numpy().astype(np.float32)
            out[f"{key}_xx"] = xx.numpy().astype(np.float32)
            out[f"{key}_yy"] = yy.numpy().astype(np.float32)
            out[f"{key}_params"] = np.array([kw["n"], kw["octaves"], kw["lacunarity"], kw["persistence"],
                                             kw["baseFrequency"], float(kw["tileable"]), kw["seed"],
                                             float(kw["normalized"])], np.float64)
    np.savez(OUT_NOISE, **out)
    print("wrote", OUT_NOISE, {k: v.shape for k, v in out.items()})


OUT_SDF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_sdf.npz")


def sdf_fixtures():
    """Signed distances of the reference's 2-D primitives
    (v2/sdfFunctionality/implicitFunctions.py, vmapped as v2/sdf.py's getSDF
    does), its operators (operators.py), filterParticlesWithSDF and the
    'regular' boundary sampling of sampleBoundaryParticles
    (v2/sampling.py:104-113, 55-74, 824-893; executed from their source
    lines: sampling.py imports scikit-image, unavailable here).  Every call
    gets its own copy of the points (some primitives modify them)."""
    sys.path.insert(0, REF)
    imf = importlib.import_module("diffSPH.v2.sdfFunctionality.implicitFunctions")
    ops = importlib.import_module("diffSPH.v2.sdfFunctionality.operators")
    fd = importlib.import_module("diffSPH.v2.finiteDifference")
    rmath = importlib.import_module("diffSPH.v2.math")
    src_path = os.path.join(REF, "diffSPH", "v2", "sampling.py")
    src = open(src_path).read().splitlines()
    ns_ = {"torch": torch, "np": np, "List": list, "centralDifferenceStencil": fd.centralDifferenceStencil,
           "continuousGradient": fd.continuousGradient, "volumeToSupport": rmath.volumeToSupport,
           "operatorDict": ops.operatorDict}
    for name in ("filterParticlesWithSDF", "sampleRegular", "sampleSDF_regular", "sampleBoundaryParticles"):
        a = next(k for k, l in enumerate(src) if l.startswith(f"def {name}("))
        b = next((k for k in range(a + 1, len(src)) if src[k] and not src[k][0].isspace()), len(src))
        exec(compile("\n".join(src[a:b]), src_path, "exec"), ns_)
    g = torch.Generator().manual_seed(5)
    pts = (torch.rand(400, 2, generator=g) * 3.0 - 1.5).to(torch.float32)
    V = lambda f, k: torch.vmap(f, in_dims=(0, *([None] * k)), out_dims=0)  # noqa: E731
    T = torch.tensor
    prim = {
        "circle": (V(imf.sdCircle, 1), (0.55,)),
        "box": (V(imf.sdBox, 1), (T([0.6, 0.35]),)),
        "roundedBox": (imf.sdRoundedBox, (T([0.6, 0.4]), T([0.1, 0.2, 0.05, 0.15]))),
        "orientedBox": (V(imf.sdOrientedBox, 3), (T([-0.5, -0.2]), T([0.4, 0.3]), 0.2)),
        "segment": (V(imf.sdSegment, 2), (T([-0.3, 0.1]), T([0.5, -0.4]))),
        "rhombus": (V(imf.sdRhombus, 1), (T([0.7, 0.4]),)),
        "trapezoid": (V(imf.sdTrapezoid, 3), (0.6, 0.3, 0.4)),
        "parallelogram": (V(imf.sdParallelogram, 3), (0.5, 0.3, 0.2)),
        "equilateralTriangle": (imf.sdEquilateralTriangle, (0.6,)),
        "triangleIsosceles": (V(imf.sdTriangleIsosceles, 1), (T([0.4, 0.9]),)),
        "triangle": (V(imf.sdTriangle, 3), (T([-0.5, -0.4]), T([0.6, -0.2]), T([0.1, 0.7]))),
        "unevenCapsule": (V(imf.sdUnevenCapsule, 3), (0.3, 0.15, 0.6)),
        "pentagon": (V(imf.sdPentagon, 1), (0.5,)),
        "hexagon": (V(imf.sdHexagon, 1), (0.5,)),
        "octogon": (V(imf.sdOctogon, 1), (0.5,)),
        "polygon": (V(imf.sdPolygon, 1), (T([[-0.6, -0.5], [0.7, -0.4], [0.5, 0.6], [-0.1, 0.2], [-0.5, 0.5]]),)),
    }
    out = {"points": pts.numpy()}
    import contextlib
    import io
    with contextlib.redirect_stdout(io.StringIO()):  # some primitives print shapes
        for name, (f, args) in prim.items():
            out[f"prim_{name}"] = f(pts.clone(), *args).numpy().astype(np.float32)
    circ = lambda p: V(imf.sdCircle, 1)(p, 0.55)  # noqa: E731
    box = lambda p: V(imf.sdBox, 1)(p, T([0.6, 0.35]))  # noqa: E731
    od = ops.operatorDict
    opc = {
        "union": od["union"](circ, box), "intersection": od["intersection"](circ, box),
        "difference": od["difference"](circ, box), "smooth_union": od["smooth_union"](circ, box, 0.2),
        "smooth_intersection": od["smooth_intersection"](circ, box, 0.2),
        "smooth_difference": od["smooth_difference"](circ, box, 0.2), "twist": od["twist"](box, 0.7),
        "bend": od["bend"](box, 0.6), "taper": od["taper"](box, 0.8), "shear": od["shear"](box, 0.3),
        "scale": od["scale"](box, 1.5), "translate": od["translate"](circ, T([0.2, -0.1])),
        "rotate": od["rotate"](box, 0.4), "mirror": od["mirror"](box, 0.1), "flip": od["flip"](box, 0.2),
        "invert": od["invert"](circ), "shell": od["shell"](circ, 0.05), "round": od["round"](box, 0.05),
        "blend": od["blend"](circ, box, 0.3), "select": od["select"](circ, box, 0.1),
    }
    for name, f in opc.items():
        out[f"op_{name}"] = f(pts.clone()).numpy().astype(np.float32)
    # lattice filtering and the regular boundary sampling of a circular body
    dx = 2.0 / 32
    lat, area = ns_["sampleRegular"](dx, 2, -1.0, 1.0, 50, False, None)
    h = float(rmath.volumeToSupport(torch.tensor(area), 50.0, 2))
    _, mask, vals, grads = ns_["filterParticlesWithSDF"](lat.clone(), od["invert"](box), h, 1e-4)
    out.update(lattice=lat.numpy(), h=np.float32(h), filter_mask=mask.numpy(), filter_values=vals.numpy(),
               filter_grads=grads.numpy())
    config = {"particle": {"dx": torch.tensor(dx), "support": torch.tensor(h)},
              "domain": {"dim": 2, "minExtent": torch.tensor([-1.0, -1.0]), "maxExtent": torch.tensor([1.0, 1.0])},
              "kernel": {"targetNeighbors": 50, "function": None}, "simulation": {"correctArea": False},
              "compute": {"device": "cpu", "dtype": torch.float32}, "fluid": {"rho0": 1.0}}
    body = lambda p: V(imf.sdCircle, 1)(p, 0.3)  # noqa: E731
    fluid = {"positions": lat.clone(), "supports": torch.full((lat.shape[0],), h)}
    bp, bv, bid, fmask = ns_["sampleBoundaryParticles"](fluid, [body], config, ["regular"])
    _, _, bvals, bgrads = ns_["filterParticlesWithSDF"](bp, od["invert"](body), h, 1e-4)
    out.update(boundary_positions=bp.numpy(), boundary_volumes=bv.numpy(), boundary_ids=bid.numpy(),
               fluid_mask=fmask.numpy(), boundary_distances=(-bvals).numpy(), boundary_normals=bgrads.numpy())
    np.savez(OUT_SDF, **out)
    print("wrote", OUT_SDF, {k: np.shape(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
    rollout_grad()
    sph_operations_and_shifting()
    ips_shifting()
    octave_noise()
    sdf_fixtures()
