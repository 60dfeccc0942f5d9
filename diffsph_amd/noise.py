"""Noise initial conditions with the reference's interface (host side, like
the lattice sampler: a one-off set-up, not a per-step kernel).

* ``perlin_noise_2d(n, res, tileable, seed)``: gradient noise on an n x n
  lattice with res x res cells, the construction of
  src/diffSPH/v2/noiseFunctions/perlin.py (perlinNoise2D): unit gradients at
  the (res+1)^2 cell corners from angles 2 pi U[0,1) drawn by
  numpy.random.default_rng(seed) (the last row / column repeats the first
  when tileable), dot products with the offsets to the four corners,
  quintic fade t^3 (t (6 t - 15) + 10), bilinear blend, times sqrt(2).
* ``generateOctaveNoise(...)`` (noiseFunctions/generator.py:87-99) with
  kind 'perlin': octaves at frequencies base * lacunarity^k and amplitudes
  persistence^k (every octave re-seeded with the same seed, as the
  reference does), optionally normalised to [-1, 1].  The 'simplex' kinds
  of the reference are OpenSimplex tables compiled with numba; they are not
  built here.
* ``sampleNoisyParticles(noiseConfig, config)`` (v2/sampling.py:524-...):
  the regular lattice, the noise as a stream function and the velocity
  u = (-d psi/dy, d psi/dx) from the 'naive' SPH gradient of the library
  (sampleVelocityField, sampling.py:296-300), scaled to max |u| = u_mag.
"""
from __future__ import annotations

import numpy as np
import torch

_SQRT2 = np.float32(np.sqrt(2.0))


def _fade(t):
    return t * t * t * (t * (t * np.float32(6) - np.float32(15)) + np.float32(10))


def perlin_noise_2d(n: int, res: int, tileable: bool = True, seed: int = 42) -> np.ndarray:
    """[n, n] float32 gradient noise (axis 0 runs along the first lattice
    coordinate of the cells, as the reference's array does)."""
    res = int(res)
    if res < 1 or n % res != 0:
        raise ValueError(f"the lattice side {n} must be a multiple of the resolution {res}")
    per = n // res
    rng = np.random.default_rng(seed=seed)
    ang = 2 * np.pi * rng.random((res + 1, res + 1))
    grad = np.stack([np.cos(ang), np.sin(ang)], -1)
    if tileable:
        grad[-1, :] = grad[0, :]
        grad[:, -1] = grad[:, 0]
    grad = grad.astype(np.float32)
    # cell index and fractional position of every lattice node along each axis
    step = np.float32(res / n)
    u = (np.arange(n, dtype=np.float64) * float(step)).astype(np.float32) % np.float32(1.0)
    cell = np.arange(n) // per
    fx = u[:, None]
    fy = u[None, :]
    ci = cell[:, None]
    cj = cell[None, :]

    def corner(di, dj, ox, oy):
        g = grad[ci + di, cj + dj]
        return (fx - ox) * g[..., 0] + (fy - oy) * g[..., 1]

    n00 = corner(0, 0, 0, 0)
    n10 = corner(1, 0, 1, 0)
    n01 = corner(0, 1, 0, 1)
    n11 = corner(1, 1, 1, 1)
    tx = _fade(fx)
    ty = _fade(fy)
    one = np.float32(1)
    a = n00 * (one - tx) + tx * n10
    b = n01 * (one - tx) + tx * n11
    return (_SQRT2 * ((one - ty) * a + ty * b)).astype(np.float32)


def generateOctaveNoise(n, dim=2, octaves=4, lacunarity=2, persistence=0.5, baseFrequency=1, tileable=True,
                        kind="perlin", device="cpu", dtype=torch.float32, seed=12345, normalized=True):
    """generator.py:87-99; returns (xx, yy, noise) as the reference does."""
    if dim != 2:
        raise ValueError("2-D noise only")
    if kind != "perlin":
        raise NotImplementedError("the simplex noise kinds are not built (the reference's are numba kernels)")
    noise = np.zeros((n, n), np.float32)
    freq, amp = baseFrequency, 1.0
    for _ in range(int(octaves)):
        noise = (noise + np.float32(amp) * perlin_noise_2d(n, freq, tileable, seed)).astype(np.float32)
        freq *= lacunarity
        amp *= persistence
    if normalized:
        lo, hi = noise.min(), noise.max()
        noise = ((noise - lo) / (hi - lo) * np.float32(2) - np.float32(1)).astype(np.float32)
    dx = 2.0 / n
    x = torch.linspace(-1 + dx / 2, 1 - dx / 2, n, dtype=dtype)
    xx, yy = torch.meshgrid(x, x, indexing="xy")
    return xx.to(device), yy.to(device), torch.from_numpy(noise).to(device=device, dtype=dtype)


def sampleNoise(noiseConfig):
    """sampling.py:437-455 defaults (baseFrequency 3, octaves 1, persistence
    0.5, lacunarity 2, seed 234675, tileable), kind 'perlin'."""
    cfg = dict(noiseConfig)
    cfg.setdefault("baseFrequency", 3)
    cfg.setdefault("dim", 2)
    cfg.setdefault("octaves", 1)
    cfg.setdefault("persistence", 0.5)
    cfg.setdefault("lacunarity", 2.0)
    cfg.setdefault("seed", 234675)
    cfg.setdefault("tileable", True)
    cfg.setdefault("kind", "perlin")
    *grid, noise = generateOctaveNoise(**cfg)
    return grid, noise


def sampleNoisyParticles(noiseConfig, config, sdfs=(), randomizeParticles=False):
    """A fluid block with a noise stream function: returns (noiseState, None)
    with the reference's state keys.  `sdfs`: list of {'sdf': fn, 'type':
    ...}; the fluid particles inside the 'boundary' bodies are removed
    (sampleBoundaryParticles, 'regular' sampling) and, as the reference
    assigns them (sampling.py:651), 'normals' receives the union's sdf
    values and 'distances' its unit gradients.  Randomized (shifted)
    particle sets and the boundary ramp of an active boundary are not
    built."""
    from . import sampling
    from .neighborhood import neighborSearch
    from .sphops import sphOperationStates
    if randomizeParticles:
        raise NotImplementedError("randomized particle sets are not built")
    if config.get("boundary", {}).get("active", False):
        raise NotImplementedError("the potential ramp of an active boundary (generateRamp) is not built")
    dev = torch.device(config.get("compute", {}).get("device", "cuda:0"))
    dom = config["domain"]
    mn = [float(v) for v in dom["minExtent"]]
    mx = [float(v) for v in dom["maxExtent"]]
    n = int(noiseConfig["n"])
    dx = (mx[0] - mn[0]) / n
    rho0 = float(config["fluid"]["rho0"])
    # sampleRegular (sampling.py:55-74): cell-centred, x fastest
    xs = np.linspace(mn[0] + dx / 2, mx[0] - dx / 2, n, dtype=np.float32)
    ys = np.linspace(mn[1] + dx / 2, mx[1] - dx / 2, n, dtype=np.float32)
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    pos = np.stack([X.ravel(), Y.ravel()], 1).astype(np.float32)
    N = pos.shape[0]
    area = dx * dx
    h = sampling.volume_to_support(area, float(config["kernel"]["targetNeighbors"]))
    _, noise = sampleNoise(noiseConfig)
    st = {"numParticles": N, "positions": torch.from_numpy(pos).to(dev),
          "areas": torch.full((N,), area, device=dev), "pressures": torch.zeros(N, device=dev),
          "divergence": torch.zeros(N, device=dev), "supports": torch.full((N,), h, device=dev),
          "index": torch.arange(N, device=dev), "densities": torch.full((N,), rho0, device=dev),
          "velocities": torch.zeros(N, 2, device=dev), "accelerations": torch.zeros(N, 2, device=dev)}
    st["masses"] = st["areas"] * rho0
    if len(sdfs) > 0:
        st["distances"] = torch.full((N,), float("inf"), device=dev)
    st["potential"] = noise.flatten().to(dev)
    bodies = [b["sdf"] for b in sdfs if b["type"] == "boundary"]
    if bodies:
        from . import sdf as sdfm
        # the boundary lattice uses the fluid's spacing and support
        bcfg = dict(config, particle=dict(config.get("particle", {}), dx=dx, support=h),
                    compute=dict(config.get("compute", {}), device=dev))
        bcfg["kernel"] = dict(config["kernel"])
        bcfg["domain"] = dict(config["domain"], dim=2)
        _, _, _, fmask = sdfm.sampleBoundaryParticles(st, bodies, bcfg, ["regular"] * len(bodies))
        for k in list(st.keys()):
            if isinstance(st[k], torch.Tensor):
                st[k] = st[k][fmask]
        st["numParticles"] = st["positions"].shape[0]
        combined = bodies[0]
        for b in bodies[1:]:
            combined = sdfm.op_union(combined, b)
        _, _, st["normals"], st["distances"] = sdfm.filterParticlesWithSDF(st["positions"], combined, h, 1e-4)
    _, nd = neighborSearch(st, st, config)
    st["neighborhood"] = nd
    g = sphOperationStates(st, st, (st["potential"], st["potential"]), nd, operation="gradient",
                           gradientMode="naive")
    vel = torch.stack([-g[:, 1], g[:, 0]], -1)
    st["divergence"] = sphOperationStates(st, st, (vel, vel), nd, operation="divergence")
    u_max = torch.linalg.norm(vel, dim=1).max() + 1e-6
    st["velocities"] = vel * (float(config["fluid"]["u_mag"]) / u_max)
    st["numNeighbors"] = nd["numNeighbors"]
    return st, None
