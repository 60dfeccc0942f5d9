"""Signed distance functions and SDF-based initial conditions with the
reference's interface (host-side torch, any device: a one-off set-up, like
the lattice sampler).

* Primitives (v2/sdfFunctionality/implicitFunctions.py): the 2-D distance
  functions of the published collection by I. Quilez that the reference
  evaluates per point under ``torch.vmap``, written here over a whole
  [N, 2] batch.  Negative inside.  Unlike some of the reference's forms
  (trapezoid, isosceles triangle, uneven capsule, pentagon, equilateral
  triangle), they never modify their input positions.
* Operators (v2/sdfFunctionality/operators.py): union, intersection,
  difference, their smooth forms, translate, rotate, scale, mirror, flip,
  invert, shell, round, blend, select, twist, bend, taper, shear, as
  closures over SDFs (``operatorDict``).
* ``getSDF(name)`` (v2/sdf.py:5-...): {'function', 'arguments', 'sample'}.
* ``filterParticlesWithSDF`` (v2/sampling.py:104-113): values, the mask
  value <= threshold and unit gradients by the second-order central
  difference stencil (-1/2, 0, 1/2) at spacing 0.2 h
  (finiteDifference.py:3-6, 213-232).
* ``sampleRegular`` (sampling.py:55-74), ``sampleSDF_regular``
  (:824-835), ``sampleBoundaryParticles`` (:837-893, 'regular' sampling),
  ``processBoundarySDFs`` (:901-950) and ``sampleParticles`` (:384-436).
  The 'contour' sampling of the reference needs scikit-image and is not
  built.
"""
from __future__ import annotations

import math

import numpy as np
import torch


def _t(v, like):
    return torch.as_tensor(v, dtype=like.dtype, device=like.device)


def _dot(a, b):
    return a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]


def _norm(a):
    return torch.linalg.norm(a, dim=-1)


def _sign(x):
    return torch.sign(x)


# ---------------------------------------------------------------------------
# primitives
# ---------------------------------------------------------------------------

def sdCircle(p, r):
    return _norm(p) - r


def sdBox(p, b):
    q = torch.abs(p) - _t(b, p)
    return _norm(torch.clamp(q, min=0.0)) + torch.clamp(torch.max(q, dim=-1)[0], max=0.0)


def sdRoundedBox(p, b, r):
    """r = (r0, r1, r2, r3): corner radii picked by the signs of x and y as
    the reference does (x > 0: r0, r1 else r2, r3; then y > 0: first)."""
    r = _t(r, p)
    rx = torch.where(p[:, 0] > 0.0, r[0], r[2])
    ry = torch.where(p[:, 0] > 0.0, r[1], r[3])
    rr = torch.where(p[:, 1] > 0.0, rx, ry)
    q = torch.abs(p) - _t(b, p) + rr[:, None]
    return torch.clamp(torch.max(q, dim=-1)[0], max=0.0) + (_norm(torch.clamp(q, min=0.0)) - rr)


def sdOrientedBox(p, a, b, th):
    a, b = _t(a, p), _t(b, p)
    ln = torch.linalg.norm(b - a)
    d = (b - a) / ln
    q = p - (a + b) * 0.5
    q = torch.stack([d[0] * q[:, 0] - d[1] * q[:, 1], d[1] * q[:, 0] + d[0] * q[:, 1]], dim=-1)
    q = torch.abs(q) - torch.stack([ln, _t(th, p)]) * 0.5
    return _norm(torch.clamp(q, min=0.0)) + torch.clamp(torch.max(q, dim=-1)[0], max=0.0)


def sdSegment(p, a, b):
    a, b = _t(a, p), _t(b, p)
    pa = p - a
    ba = b - a
    h = torch.clamp(_dot(pa, ba) / torch.dot(ba, ba), 0.0, 1.0)
    return _norm(pa - ba * h[:, None])


def sdRhombus(p, b):
    b = _t(b, p)
    p = torch.abs(p)
    bb = b - 2.0 * p
    h = torch.clamp((bb[:, 0] * b[0] - bb[:, 1] * b[1]) / torch.dot(b, b), -1.0, 1.0)
    d = _norm(p - 0.5 * b * torch.stack([1.0 - h, 1.0 + h], dim=-1))
    return d * _sign(p[:, 0] * b[1] + p[:, 1] * b[0] - b[0] * b[1])


def sdTrapezoid(p, r1, r2, he):
    k1 = _t([r2, he], p)
    k2 = _t([r2 - r1, 2.0 * he], p)
    px = torch.abs(p[:, 0])
    p = torch.stack([px, p[:, 1]], dim=-1)
    rs = torch.where(p[:, 1] < 0.0, _t(r1, p), _t(r2, p))
    ca = torch.stack([p[:, 0] - torch.minimum(p[:, 0], rs), torch.abs(p[:, 1]) - he], dim=-1)
    cb = p - k1 + k2 * torch.clamp(_dot(k1 - p, k2) / torch.dot(k2, k2), 0.0, 1.0)[:, None]
    s = torch.where((cb[:, 0] < 0.0) & (ca[:, 1] < 0.0), -1.0, 1.0).to(p.dtype)
    return s * torch.sqrt(torch.minimum(_dot(ca, ca), _dot(cb, cb)))


def sdParallelogram(p, wi, he, sk):
    e = _t([sk, he], p)
    p = torch.where((p[:, 1] < 0.0)[:, None], -p, p)
    w = p - e
    w = torch.stack([w[:, 0] - torch.clamp(w[:, 0], -wi, wi), w[:, 1]], dim=-1)
    d0, d1 = _dot(w, w), -w[:, 1]
    s = p[:, 0] * e[1] - p[:, 1] * e[0]
    p = torch.where((s < 0.0)[:, None], -p, p)
    v = p - _t([wi, 0.0], p)
    v = v - e * torch.clamp(_dot(v, e) / torch.dot(e, e), -1.0, 1.0)[:, None]
    c0, c1 = _dot(v, v), wi * he - torch.abs(s)
    return torch.sqrt(torch.minimum(d0, c0)) * _sign(-torch.minimum(d1, c1))


def sdEquilateralTriangle(p, r):
    k = math.sqrt(3.0)
    x = torch.abs(p[:, 0]) - r
    y = p[:, 1] + r / k
    m = x + k * y > 0.0
    x, y = torch.where(m, (x - k * y) / 2.0, x), torch.where(m, (-k * x - y) / 2.0, y)
    x = x - torch.clamp(x, -2.0 * r, 0.0)
    return -_norm(torch.stack([x, y], dim=-1)) * _sign(y)


def sdTriangleIsosceles(p, q):
    q = _t(q, p)
    p = torch.stack([torch.abs(p[:, 0]), p[:, 1]], dim=-1)
    a = p - q * torch.clamp(_dot(p, q) / torch.dot(q, q), 0.0, 1.0)[:, None]
    b = p - q * torch.stack([torch.clamp(p[:, 0] / q[0], 0.0, 1.0), torch.ones_like(p[:, 0])], dim=-1)
    s = -_sign(q[1])
    d0 = torch.minimum(_dot(a, a), _dot(b, b))
    d1 = torch.minimum(s * (p[:, 0] * q[1] - p[:, 1] * q[0]), s * (p[:, 1] - q[1]))
    return -torch.sqrt(d0) * _sign(d1)


def sdTriangle(p, p0, p1, p2):
    p0, p1, p2 = _t(p0, p), _t(p1, p), _t(p2, p)
    e0, e1, e2 = p1 - p0, p2 - p1, p0 - p2
    v0, v1, v2 = p - p0, p - p1, p - p2

    def edge(v, e):
        return v - e * torch.clamp(_dot(v, e) / torch.dot(e, e), 0.0, 1.0)[:, None]

    pq0, pq1, pq2 = edge(v0, e0), edge(v1, e1), edge(v2, e2)
    s = _sign(e0[0] * e2[1] - e0[1] * e2[0])
    d0 = torch.minimum(torch.minimum(_dot(pq0, pq0), _dot(pq1, pq1)), _dot(pq2, pq2))
    d1 = torch.minimum(torch.minimum(s * (v0[:, 0] * e0[1] - v0[:, 1] * e0[0]),
                                     s * (v1[:, 0] * e1[1] - v1[:, 1] * e1[0])),
                       s * (v2[:, 0] * e2[1] - v2[:, 1] * e2[0]))
    return -torch.sqrt(d0) * _sign(d1)


def sdUnevenCapsule(p, r1, r2, h):
    p = torch.stack([torch.abs(p[:, 0]), p[:, 1]], dim=-1)
    b = (r1 - r2) / h
    a = math.sqrt(1.0 - b * b)
    k = p[:, 0] * (-b) + p[:, 1] * a
    d1 = _norm(p) - r1
    d2 = _norm(p - _t([0.0, h], p)) - r2
    d3 = p[:, 0] * a + p[:, 1] * b - r1
    return torch.where(k < 0.0, d1, torch.where(k > a * h, d2, d3))


def _reflect(p, n):
    """p - 2 min(n . p, 0) n."""
    return p - 2.0 * torch.clamp(_dot(p, n), max=0.0)[:, None] * n


def _polygon_tail(p, half, r):
    p = torch.stack([p[:, 0] - torch.clamp(p[:, 0], -half * r, half * r), p[:, 1] - r], dim=-1)
    return _norm(p) * _sign(p[:, 1])


def sdPentagon(p, r):
    k = (0.809016994, 0.587785252, 0.726542528)
    p = torch.stack([torch.abs(p[:, 0]), p[:, 1]], dim=-1)
    p = _reflect(p, _t([-k[0], k[1]], p))
    p = _reflect(p, _t([k[0], k[1]], p))
    return _polygon_tail(p, k[2], r)


def sdHexagon(p, r):
    k = (-0.866025404, 0.5, 0.577350269)
    p = _reflect(torch.abs(p), _t([k[0], k[1]], p))
    return _polygon_tail(p, k[2], r)


def sdOctogon(p, r):
    k = (-0.9238795325, 0.3826834323, 0.4142135623)
    p = torch.abs(p)
    p = _reflect(p, _t([k[0], k[1]], p))
    p = _reflect(p, _t([-k[0], k[1]], p))
    return _polygon_tail(p, k[2], r)


def sdPolygon(p, v):
    """Closed polygon with vertices v [m, 2], as the reference writes it
    (implicitFunctions.py:151-165): the sign test of the published form with
    its branches swapped, so the value is positive inside and negative
    outside (the magnitude is the distance to the boundary)."""
    v = _t(v, p)
    d = _dot(p - v[0], p - v[0])
    s = torch.full_like(d, -1.0)
    m = v.shape[0]
    for i in range(m):
        j = i - 1 if i > 0 else m - 1
        e = v[j] - v[i]
        w = p - v[i]
        b = w - e * torch.clamp(_dot(w, e) / torch.dot(e, e), 0.0, 1.0)[:, None]
        d = torch.minimum(d, _dot(b, b))
        c1 = p[:, 1] >= v[i][1]
        c2 = p[:, 1] < v[j][1]
        c3 = e[0] * w[:, 1] > e[1] * w[:, 0]
        keep = (c1 & c2 & c3) | (~c1 & ~c2 & ~c3)
        s = torch.where(keep, s, -s)
    return s * torch.sqrt(d)


functionDict = {
    "circle": sdCircle, "box": sdBox, "roundedBox": sdRoundedBox, "orientedBox": sdOrientedBox,
    "segment": sdSegment, "rhombus": sdRhombus, "trapezoid": sdTrapezoid, "parallelogram": sdParallelogram,
    "equilateralTriangle": sdEquilateralTriangle, "triangleIsosceles": sdTriangleIsosceles,
    "triangle": sdTriangle, "unevenCapsule": sdUnevenCapsule, "pentagon": sdPentagon, "hexagon": sdHexagon,
    "octogon": sdOctogon, "polygon": sdPolygon,
}

_ARGS = {
    "circle": ({"p": "batched Input Position", "r": "Radius of sphere, float"}, [1.0]),
    "box": ({"p": "batched Input Position", "b": "Size of box, torch.Tensor"}, [torch.tensor([1.0, 1.0])]),
    "roundedBox": ({"p": "batched Input Position", "b": "Size of box, torch.Tensor",
                    "r": "Radius of corners, torch.Tensor"},
                   [torch.tensor([1.0, 1.0]), torch.tensor([0.1, 0.2, 0.3, 0.4])]),
    "orientedBox": ({"p": "batched Input Position", "a": "Start of box, torch.Tensor",
                     "b": "End of box, torch.Tensor", "th": "Thickness of box, float"},
                    [torch.tensor([0.0, 0.0]), torch.tensor([1.0, 1.0]), 0.1]),
    "segment": ({"p": "batched Input Position", "a": "Start of segment, torch.Tensor",
                 "b": "End of segment, torch.Tensor"}, [torch.tensor([0.0, 0.0]), torch.tensor([1.0, 1.0])]),
    "rhombus": ({"p": "batched Input Position", "b": "Size of rhombus, torch.Tensor"}, [torch.tensor([1.0, 1.0])]),
    "trapezoid": ({"p": "batched Input Position", "r1": "Top radius, float", "r2": "Bottom radius, float",
                   "he": "Height, float"}, [1.0, 0.5, 0.5]),
    "parallelogram": ({"p": "batched Input Position", "wi": "Width, float", "he": "Height, float",
                       "sk": "Skew, float"}, [1.0, 0.5, 0.5]),
    "equilateralTriangle": ({"p": "batched Input Position", "r": "Radius of triangle, float"}, [1.0]),
    "triangleIsosceles": ({"p": "batched Input Position", "q": "Base of triangle, torch.Tensor"},
                          [torch.tensor([0.3, 1.1])]),
    "triangle": ({"p": "batched Input Position", "p0": "First point of triangle, torch.Tensor",
                  "p1": "Second point of triangle, torch.Tensor", "p2": "Third point of triangle, torch.Tensor"},
                 [torch.tensor([0.0, 0.0]), torch.tensor([1.0, 1.0]), torch.tensor([1.0, 0.0])]),
    "unevenCapsule": ({"p": "batched Input Position", "r1": "Top radius, float", "r2": "Bottom radius, float",
                       "h": "Height, float"}, [1.0, 0.5, 0.5]),
    "pentagon": ({"p": "batched Input Position", "r": "Radius of pentagon, float"}, [1.0]),
    "hexagon": ({"p": "batched Input Position", "r": "Radius of hexagon, float"}, [1.0]),
    "octogon": ({"p": "batched Input Position", "r": "Radius of octogon, float"}, [1.0]),
    "polygon": ({"p": "batched Input Position", "v": "Vertices, torch.Tensor [m, 2]"},
                [torch.tensor([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])]),
}


def getSDF(function: str):
    """v2/sdf.py:5-...: {'function', 'arguments', 'sample'} for a primitive."""
    if function not in functionDict:
        raise NotImplementedError(f"SDF primitive '{function}' is not built")
    args, sample = _ARGS[function]
    return {"function": functionDict[function], "arguments": args, "sample": sample}


# ---------------------------------------------------------------------------
# operators
# ---------------------------------------------------------------------------

def op_union(a, b):
    return lambda p: torch.min(a(p), b(p))


def op_intersection(a, b):
    return lambda p: torch.max(a(p), b(p))


def op_difference(a, b):
    return lambda p: torch.max(a(p), -b(p))


def op_smooth_union(a, b, k):
    def f(p):
        d1, d2 = a(p), b(p)
        h = torch.clamp(0.5 + 0.5 * (d2 - d1) / k, 0.0, 1.0)
        return torch.lerp(d2, d1, h) - k * h * (1.0 - h)
    return f


def op_smooth_intersection(a, b, k):
    def f(p):
        d1, d2 = a(p), b(p)
        h = torch.clamp(0.5 - 0.5 * (d2 - d1) / k, 0.0, 1.0)
        return torch.lerp(d2, d1, h) + k * h * (1.0 - h)
    return f


def op_smooth_difference(a, b, k):
    def f(p):
        d1, d2 = a(p), b(p)
        h = torch.clamp(0.5 - 0.5 * (d2 + d1) / k, 0.0, 1.0)
        return torch.lerp(d2, -d1, h) + k * h * (1.0 - h)
    return f


def _rot(p, c, s):
    return torch.stack([c * p[:, 0] - s * p[:, 1], s * p[:, 0] + c * p[:, 1]], dim=1)


def op_twist(a, k):
    def f(p):
        ang = (k * p[:, 1]).cpu().numpy()
        c = torch.as_tensor(np.cos(ang), device=p.device, dtype=p.dtype)
        s = torch.as_tensor(np.sin(ang), device=p.device, dtype=p.dtype)
        return a(_rot(p, c, s))
    return f


def op_bend(a, k):
    def f(p):
        ang = (k * p[:, 0]).cpu().numpy()
        c = torch.as_tensor(np.cos(ang), device=p.device, dtype=p.dtype)
        s = torch.as_tensor(np.sin(ang), device=p.device, dtype=p.dtype)
        return a(_rot(p, c, s))
    return f


def op_taper(a, k):
    return lambda p: a(torch.stack([p[:, 0] * k, p[:, 1]], dim=1))


def op_shear(a, k):
    return lambda p: a(torch.stack([p[:, 0] + k * p[:, 1], p[:, 1]], dim=1))


def op_scale(a, k):
    return lambda p: a(p / k) * k


def op_translate(a, t):
    return lambda p: a(p - _t(t, p))


def op_rotate(a, angle):
    c, s = np.cos(angle), np.sin(angle)
    return lambda p: a(_rot(p, c, s))


def op_mirror(a, axis):
    return lambda p: a(torch.stack([p[:, 0], axis - p[:, 1]], dim=1))


def op_flip(a, axis):
    return lambda p: a(torch.stack([axis - p[:, 0], p[:, 1]], dim=1))


def op_invert(a):
    return lambda p: -a(p)


def op_shell(a, thickness):
    return lambda p: torch.abs(a(p)) - thickness


def op_round(a, radius):
    return lambda p: a(p) - radius


def op_blend(a, b, l):
    return lambda p: torch.lerp(a(p), b(p), l)


def op_select(a, b, k):
    return lambda p: torch.where(p[:, 0] < k, a(p), b(p))


operatorDict = {
    "union": op_union, "intersection": op_intersection, "difference": op_difference,
    "smooth_union": op_smooth_union, "smooth_intersection": op_smooth_intersection,
    "smooth_difference": op_smooth_difference, "twist": op_twist, "bend": op_bend, "taper": op_taper,
    "shear": op_shear, "scale": op_scale, "translate": op_translate, "rotate": op_rotate, "mirror": op_mirror,
    "flip": op_flip, "invert": op_invert, "shell": op_shell, "round": op_round, "blend": op_blend,
    "select": op_select,
}


# ---------------------------------------------------------------------------
# sampling with SDFs
# ---------------------------------------------------------------------------

_STENCIL = (-0.5, 0.0, 0.5)   # first derivative, second order (finiteDifference.py:3-6)


def filterParticlesWithSDF(p, sdf, h, threshold=0.0):
    """(masked positions, mask, sdf values, unit sdf gradients): mask =
    value <= threshold; the gradient per axis is sum_k c_k sdf(p + c_k dx
    e_d) / dx with c = (-1/2, 0, 1/2) and dx = 0.2 h, normalised as
    torch.nn.functional.normalize does."""
    values = sdf(p.cpu()).to(p.device)
    mask = values <= threshold
    stencil = torch.tensor(_STENCIL, dtype=torch.float32, device=p.device)
    dx = 0.2 * h
    comps = []
    for d in range(p.shape[1]):
        offs = p.new_zeros(stencil.shape[0], p.shape[1])
        offs[:, d] = stencil
        q = (p[:, None, :] + offs[None, :, :] * dx).view(-1, p.shape[1])
        f = sdf(q.cpu()).to(p.device).view(p.shape[0], -1)
        comps.append(torch.sum(stencil * f, dim=-1) / dx ** 1)
    grad = torch.nn.functional.normalize(torch.stack(comps, dim=-1), dim=-1)
    return p[mask], mask, values, grad


def volumeToSupport(volume, targetNeighbors, dim=2):
    """math.py:96-105 for dim 2: h = sqrt(targetNeighbors volume / pi)."""
    if dim != 2:
        raise ValueError("2-D only")
    if isinstance(volume, torch.Tensor):
        return torch.sqrt(targetNeighbors * volume / np.pi)
    return math.sqrt(targetNeighbors * volume / np.pi)


def sampleRegular(dx=2 / 32, dim=2, minExtent=-1, maxExtent=1, targetNeighbors=50, correctedArea=False,
                  kernel=None):
    """sampling.py:55-74: cell-centred lattice (x fastest), area dx^dim."""
    if correctedArea:
        raise NotImplementedError("the corrected-area lattice (optimizeArea) is not built")

    def ext(v):
        if isinstance(v, (int, float)):
            return torch.tensor([v] * dim).to(torch.float32)
        return torch.as_tensor(v).detach().cpu().to(torch.float32)

    lo, hi = ext(minExtent), ext(maxExtent)
    area = dx ** dim
    ns = [torch.ceil((hi[i] - lo[i]) / dx).to(torch.long) for i in range(dim)]
    lins = [torch.linspace(lo[i] + dx / 2, hi[i] - dx / 2, ns[i]) for i in range(dim)]
    grid = torch.meshgrid(*lins, indexing="xy")
    p = torch.stack(grid, dim=-1).flatten(0, -1).view(-1, dim)
    return p, area


def _lattice(config):
    c = config
    return sampleRegular(c["particle"]["dx"], c["domain"]["dim"], c["domain"]["minExtent"],
                         c["domain"]["maxExtent"], c["kernel"]["targetNeighbors"],
                         c.get("simulation", {}).get("correctArea", False), c["kernel"].get("function"))


def sampleSDF_regular(sdf, config):
    """sampling.py:824-835: the lattice points inside the body (sdf <= -1e-4)."""
    dev = config["compute"]["device"]
    pts, vol = _lattice(config)
    pts = pts.to(dev)
    vol = torch.as_tensor(vol, dtype=torch.float32).to(dev)
    h = volumeToSupport(vol, config["kernel"]["targetNeighbors"], config["domain"]["dim"])
    _, mask, values, grads = filterParticlesWithSDF(pts, sdf, h, -1e-4)
    return pts[mask], vol, values[mask], grads[mask]


def sampleBoundaryParticles(particleState, sdfs, config, modes=("regular",)):
    """sampling.py:837-893: boundary particles of every body (lattice
    points inside it) and the mask of the fluid particles outside all
    bodies (invert(sdf) <= 1e-4)."""
    dev = config["compute"]["device"]
    dt = config["compute"].get("dtype", torch.float32)
    fluid = torch.ones(particleState["positions"].shape[0], dtype=torch.bool, device=dev)
    pos, vols, bodies = [], [], []
    if isinstance(modes, str):
        modes = [modes] * len(sdfs)
    for i, (sdf, mode) in enumerate(zip(sdfs, modes)):
        _, m, _, _ = filterParticlesWithSDF(particleState["positions"], op_invert(sdf),
                                            config["particle"]["support"], 1e-4)
        fluid = fluid & m
        if mode != "regular":
            raise NotImplementedError("only 'regular' boundary sampling is built")
        bp, bv, _, _ = sampleSDF_regular(sdf, config)
        bp = bp.to(dev).to(dt)
        pos.append(bp)
        vols.append(bv.to(dt) * bp.new_ones(bp.shape[0]))
        bodies.append(bp.new_ones(bp.shape[0]).to(dev) * i)
    return torch.cat(pos, 0), torch.cat(vols, 0), torch.cat(bodies, 0), fluid


def processBoundarySDFs(fluidState, config, sdfs, samplings=None):
    """sampling.py:901-950: removes the fluid particles inside the bodies
    (in place) and returns the boundary state: positions, areas, masses,
    supports, densities, zero velocities / accelerations / pressures,
    'distances' = -sdf and 'normals' = the unit sdf gradient of the union
    of the bodies inverted, body ids, and its neighbor search."""
    from .neighborhood import neighborSearch
    if samplings is None:
        samplings = ["regular"] * len(sdfs)
    if isinstance(samplings, str):
        samplings = [samplings] * len(sdfs)
    bp, bv, bid, fmask = sampleBoundaryParticles(fluidState, sdfs, config, samplings)
    for k in list(fluidState.keys()):
        if isinstance(fluidState[k], torch.Tensor):
            fluidState[k] = fluidState[k][fmask]
    fluidState["numParticles"] = fluidState["positions"].shape[0]
    combined = sdfs[0]
    for s in sdfs[1:]:
        combined = op_union(combined, s)
    _, _, values, grads = filterParticlesWithSDF(bp, op_invert(combined), config["particle"]["support"], 1e-4)
    rho0 = config["fluid"]["rho0"]
    dev = config["compute"]["device"]
    n = bp.shape[0]
    st = {
        "positions": bp, "areas": bv, "pressures": bv.new_zeros(n), "masses": bv * rho0,
        "supports": bv.new_ones(n) * config["particle"]["support"], "index": torch.arange(n, device=dev),
        "densities": bv.new_ones(n) * rho0, "velocities": bv.new_zeros(n, config["domain"]["dim"]),
        "accelerations": bv.new_zeros(n, config["domain"]["dim"]), "numParticles": n,
        "distances": -values, "normals": grads, "bodyIDs": bid,
    }
    _, st["neighborhood"] = neighborSearch(st, st, config)
    st["numNeighbors"] = st["neighborhood"]["numNeighbors"]
    return st


def sampleParticles(config, sdfs=(), minExtent=None, maxExtent=None, filter=True):
    """sampling.py:384-436: the lattice as a fluid state; with `sdfs`, the
    particles inside any of them (sdf <= -1e-4) are kept (or, without
    `filter`, the others get zero velocity) and 'distances' is the smallest
    sdf value.  Returns (state, mask)."""
    from .neighborhood import neighborSearch
    dev = config["compute"]["device"]
    c = dict(config)
    if minExtent is not None or maxExtent is not None:
        c = dict(config, domain=dict(config["domain"], minExtent=minExtent if minExtent is not None
                                     else config["domain"]["minExtent"],
                                     maxExtent=maxExtent if maxExtent is not None else config["domain"]["maxExtent"]))
    pts, vol = _lattice(c)
    pts = pts.to(dev)
    n = pts.shape[0]
    area = vol
    st = {"numParticles": n, "positions": pts, "areas": pts.new_ones(n) * area, "pressures": pts.new_zeros(n),
          "divergence": pts.new_zeros(n)}
    st["masses"] = st["areas"] * config["fluid"]["rho0"]
    st["supports"] = volumeToSupport(area, config["kernel"]["targetNeighbors"], config["domain"]["dim"]) * \
        pts.new_ones(n)
    st["index"] = torch.arange(n, device=dev)
    st["densities"] = pts.new_ones(n) * config["fluid"]["rho0"]
    st["velocities"] = pts.new_zeros(n, config["domain"]["dim"])
    st["accelerations"] = pts.new_zeros(n, config["domain"]["dim"])
    mask = torch.ones(n, dtype=torch.bool, device=dev)
    if len(sdfs) > 0:
        st["distances"] = pts.new_ones(n) * np.inf
        mask = torch.zeros(n, dtype=torch.bool, device=dev)
        for f in sdfs:
            _, m, values, _ = filterParticlesWithSDF(pts, f, st["supports"][0], -1e-4)
            mask = mask | m
            st["distances"] = torch.min(st["distances"], values)
    st["velocities"][~mask, :] = 0
    if filter:
        for k in list(st.keys()):
            if isinstance(st[k], torch.Tensor):
                st[k] = st[k][mask]
        st["numParticles"] = st["positions"].shape[0]
        st["index"] = torch.arange(st["numParticles"], device=dev)
    _, nd = neighborSearch(st, st, config)
    st["numNeighbors"] = nd["numNeighbors"]
    st["neighborhood"] = nd
    return st, mask
