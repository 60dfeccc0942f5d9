"""Initial conditions of the benchmark workloads (host side, numpy/torch).

* :func:`sample_regular` -- cell-centred lattice (v2/sampling.py:55-74),
* :func:`volume_to_support` -- h from a target neighbor count
  (v2/math.py:96-105),
* :func:`spectral_velocity` -- divergence-free band-limited velocity
  u = curl(psi), psi = sum_{0<|k|<=kmax} a_k cos(pi k.x + phi_k) on a
  periodic box of side 2 (seeded; rescaled to max|u| = umax),
* :func:`taylor_green` -- u = sin(pi x) cos(pi y), v = -cos(pi x) sin(pi y).
"""
from __future__ import annotations

import math

import numpy as np


def volume_to_support(volume: float, target_neighbors: float) -> float:
    """2-D: N_h = pi h^2 / v  ->  h = sqrt(N_h v / pi)."""
    return math.sqrt(target_neighbors * volume / math.pi)


def sample_regular(nx: int, ny: int | None = None, min_extent=(-1.0, -1.0), max_extent=(1.0, 1.0)):
    """Cell-centred lattice with spacing dx = (max-min)/nx, x-fastest order
    (torch.meshgrid(..., indexing='xy').flatten as v2/sampling.py:69-72).
    Returns (positions [n,2] float32, dx)."""
    ny = ny or nx
    dx = (max_extent[0] - min_extent[0]) / nx
    xs = np.linspace(min_extent[0] + dx / 2, max_extent[0] - dx / 2, nx, dtype=np.float32)
    ys = np.linspace(min_extent[1] + dx / 2, max_extent[1] - dx / 2, ny, dtype=np.float32)
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    p = np.stack([X.reshape(-1), Y.reshape(-1)], axis=1).astype(np.float32)
    return p, dx


def jitter(pos: np.ndarray, dx: float, width: float, seed: int) -> np.ndarray:
    """Uniform jitter U(-width dx/2, width dx/2) per axis."""
    rng = np.random.default_rng(seed)
    return (pos + rng.uniform(-0.5 * width * dx, 0.5 * width * dx, size=pos.shape)).astype(np.float32)


def spectral_modes(kmax: int, seed: int):
    """Integer wave vectors 0 < |k| <= kmax (half plane) with seeded
    amplitudes a_k ~ N(0,1)/|k|^2 and phases phi_k ~ U(0, 2pi)."""
    rng = np.random.default_rng(seed)
    ks = []
    for kx in range(0, kmax + 1):
        for ky in range(-kmax, kmax + 1):
            if kx == 0 and ky <= 0:
                continue
            if kx * kx + ky * ky <= kmax * kmax:
                ks.append((kx, ky))
    ks = np.array(ks, dtype=np.float64)
    kn = np.sqrt((ks ** 2).sum(1))
    a = rng.standard_normal(len(ks)) / kn ** 2
    phi = rng.uniform(0.0, 2.0 * math.pi, len(ks))
    return ks, a, phi


def spectral_velocity(pos: np.ndarray, kmax: int = 8, seed: int = 0, umax: float = 1.0,
                      period: float = 2.0) -> np.ndarray:
    """u = (d psi/dy, -d psi/dx), psi = sum a_k cos(w k.x + phi_k), w = 2 pi / period."""
    ks, a, phi = spectral_modes(kmax, seed)
    w = 2.0 * math.pi / period
    x = pos.astype(np.float64)
    u = np.zeros_like(x)
    # chunk over modes to bound memory
    for s in range(0, len(ks), 32):
        kk, aa, pp = ks[s:s + 32], a[s:s + 32], phi[s:s + 32]
        arg = w * (x @ kk.T) + pp[None, :]          # [n, m]
        sn = np.sin(arg) * (aa * w)[None, :]
        # d/dx cos(arg) = -sin(arg) w kx
        u[:, 0] += -(sn * kk[None, :, 1]).sum(1)     # d psi / dy
        u[:, 1] += (sn * kk[None, :, 0]).sum(1)      # -d psi / dx
    mag = np.sqrt((u ** 2).sum(1)).max()
    return (u * (umax / mag)).astype(np.float32)


def spectral_velocity_torch(pos, kmax: int = 8, seed: int = 0, umax: float = 1.0, period: float = 2.0):
    """:func:`spectral_velocity` for a torch tensor of positions, evaluated
    in float64 on the tensor's device (same modes, same formula); returns
    float32 [n,2] on that device."""
    import torch
    ks, a, phi = spectral_modes(kmax, seed)
    w = 2.0 * math.pi / period
    x = pos.double()
    kk = torch.from_numpy(ks).to(x.device)
    arg = w * (x @ kk.T) + torch.from_numpy(phi).to(x.device)[None, :]
    sn = torch.sin(arg) * torch.from_numpy(a * w).to(x.device)[None, :]
    u = torch.stack([-(sn * kk[None, :, 1]).sum(1), (sn * kk[None, :, 0]).sum(1)], dim=1)
    mag = torch.sqrt((u ** 2).sum(1)).max()
    return (u * (umax / mag)).float().contiguous()


def taylor_green(pos: np.ndarray, umag: float = 1.0) -> np.ndarray:
    x = pos[:, 0].astype(np.float64)
    y = pos[:, 1].astype(np.float64)
    u = np.stack([umag * np.sin(np.pi * x) * np.cos(np.pi * y), -umag * np.cos(np.pi * x) * np.sin(np.pi * y)], 1)
    return u.astype(np.float32)


def lattice_density(volume: float, dx: float, target_neighbors: float) -> float:
    """volume * sum over a square lattice of spacing dx of W(r, h(volume))
    (the reference's evalArea, v2/sampling.py:6-32), float64."""
    h = volume_to_support(volume, target_neighbors)
    m = int(h / dx) + 2
    a = np.arange(-m, m + 1, dtype=np.float64) * dx
    X, Y = np.meshgrid(a, a)
    r = np.sqrt(X * X + Y * Y).ravel()
    r = r[r < h]
    q = r / h
    w = (1.0 - q) ** 4 * (1.0 + 4.0 * q) * 7.0 / (math.pi * h * h)
    return float(volume * w.sum())


def corrected_area(dx: float, target_neighbors: float) -> float:
    """Particle volume whose lattice density is exactly 1
    (simulation.correctArea: optimizeArea, v2/sampling.py:36-51 and
    v2/parameters.py:178-181).  The reference runs 64 gradient steps; this
    bisects the same lattice sum to convergence in float64."""
    lo, hi = 0.8 * dx * dx, 1.2 * dx * dx
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if lattice_density(mid, dx, target_neighbors) > 1.0:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def dfsph_random_flow(n_side: int, seed: int = 0, jitter_width: float = 0.1, kmax: int = 8,
                      target_neighbors: float = 50.0, rho0: float = 1.0, correct_area: bool = True):
    """The 1M-particle DFSPH workload at any resolution: lattice on [-1,1]^2
    plus jitter, divergence-free spectral-noise velocity (max |u| = 1).
    With ``correct_area`` (as the reference's DFSPH setup,
    partiBench/DFSPH.ipynb cell 3) the particle volume makes the rest
    lattice density exactly rho0 -- otherwise a periodic box is uniformly
    over-compressed by 0.12% and no pressure field can correct it."""
    pos, dx = sample_regular(n_side)
    pos = jitter(pos, dx, jitter_width, seed)
    vel = spectral_velocity(pos, kmax=kmax, seed=seed + 1, umax=1.0)
    vol = corrected_area(dx, target_neighbors) if correct_area else dx * dx
    area = np.full(pos.shape[0], vol, np.float32)
    mass = (area * np.float32(rho0)).astype(np.float32)
    h = volume_to_support(vol, target_neighbors)
    return dict(pos=pos, vel=vel, area=area, mass=mass, dx=dx, h=h)


def box_walls(dx: float, box_min=(0.0, 0.0), box_max=(4.0, 2.0), layers: int = 3):
    """Wall particles on the cell-centred lattice of spacing dx in the
    `layers`-thick frame around the box: positions, signed distances
    (negative inside the wall, -sdf of the box) and normals (gradient of the
    box's sdf, pointing into the wall), as processBoundarySDFs
    (v2/sampling.py:900-941) sets 'distances' and 'normals'."""
    x0, y0 = box_min
    x1, y1 = box_max
    ix = np.arange(int(round((x0 - layers * dx) / dx)), int(round((x1 + layers * dx) / dx)))
    iy = np.arange(int(round((y0 - layers * dx) / dx)), int(round((y1 + layers * dx) / dx)))
    X, Y = np.meshgrid((ix + 0.5) * dx, (iy + 0.5) * dx, indexing="xy")
    p = np.stack([X.ravel(), Y.ravel()], 1)
    inside = (p[:, 0] > x0) & (p[:, 0] < x1) & (p[:, 1] > y0) & (p[:, 1] < y1)
    p = p[~inside]
    c = np.array([(x0 + x1) / 2, (y0 + y1) / 2])
    e = np.array([(x1 - x0) / 2, (y1 - y0) / 2])
    d = p - c
    q = np.abs(d) - e
    qp = np.maximum(q, 0.0)
    sdf = np.sqrt((qp ** 2).sum(1))
    nrm = qp * np.sign(d) / sdf[:, None]
    return p.astype(np.float32), (-sdf).astype(np.float32), nrm.astype(np.float32)


def dam_break(res: int = 256, column=(1.0, 2.0), box=(4.0, 2.0), target_neighbors: float = 50.0, layers: int = 3):
    """Dam break: a column [0,w]x[0,H] of water (dx = 1/res, cell-centred
    lattice) in the closed box [0,4]x[0,2] with `layers` wall rows outside
    each side.  Returns dict(fluid_pos, wall_pos, wall_dist, wall_normal,
    dx, h, area)."""
    dx = 1.0 / res
    nx, ny = int(round(column[0] * res)), int(round(column[1] * res))
    xs = (np.arange(nx) + 0.5) * dx
    ys = (np.arange(ny) + 0.5) * dx
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    fluid = np.stack([X.ravel(), Y.ravel()], 1).astype(np.float32)
    wp, wd, wn = box_walls(dx, (0.0, 0.0), box, layers)
    h = volume_to_support(dx * dx, target_neighbors)
    return dict(fluid_pos=fluid, wall_pos=wp, wall_dist=wd, wall_normal=wn, dx=dx, h=h, area=dx * dx)
