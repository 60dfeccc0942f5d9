"""numpy oracle of the warm-start MLP (test infrastructure only).

Restates the network of diffsph_amd/csrc/mlp.hip: 16 -> 64 -> 64 -> 64 -> 1,
ReLU, layer inputs and weights rounded to bf16 (round to nearest even) where
the GPU feeds them to the MFMA, everything else float32/float64.

`bf16=False` gives the exact float64 network, whose backward pass is pinned
against central finite differences in tests/test_mlp_oracle.py.
"""
from __future__ import annotations

import numpy as np

OFFS = {}
_k = 0
for _name, _shape in [("W1", (16, 64)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (64, 64)),
                      ("b3", (64,)), ("w4", (64,)), ("b4", (1,))]:
    OFFS[_name] = (_k, _shape)
    _k += int(np.prod(_shape))
NPARAM = _k


def bf16(x):
    """Round float32 values to bf16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, (u | 0x400000) & 0xFFFF0000, r)
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(p):
    return {k: p[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in OFFS.items()}


def forward(p, X, bf=True):
    P = unpack(np.asarray(p, np.float64))
    r = (lambda a: bf16(a).astype(np.float64)) if bf else (lambda a: np.asarray(a, np.float64))
    x = r(X)
    z1 = x @ r(P["W1"]) + P["b1"]
    h1 = r(np.maximum(z1, 0.0))
    z2 = h1 @ r(P["W2"]) + P["b2"]
    h2 = r(np.maximum(z2, 0.0))
    z3 = h2 @ r(P["W3"]) + P["b3"]
    h3 = r(np.maximum(z3, 0.0))
    y = h3 @ r(P["w4"]) + P["b4"][0]
    return y, dict(x=x, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3, h3=h3)


def loss_grad(p, X, T, lw, bf=True):
    """sum lw (y - t)^2 and its gradient (flat, OFFS layout)."""
    P = unpack(np.asarray(p, np.float64))
    r = (lambda a: bf16(a).astype(np.float64)) if bf else (lambda a: np.asarray(a, np.float64))
    y, c = forward(p, X, bf)
    e = y - T
    loss = float(lw * (e * e).sum())
    dy = 2.0 * lw * e
    g = np.zeros(NPARAM)

    def put(name, v):
        o, s = OFFS[name]
        g[o:o + int(np.prod(s))] = np.asarray(v).reshape(-1)

    put("b4", [dy.sum()])
    put("w4", c["h3"].T @ dy)
    dz3 = np.where(c["z3"] > 0, dy[:, None] * r(P["w4"])[None, :], 0.0)
    put("b3", dz3.sum(0))
    put("W3", c["h2"].T @ r(dz3))
    dh2 = r(dz3) @ r(P["W3"]).T
    dz2 = np.where(c["z2"] > 0, dh2, 0.0)
    put("b2", dz2.sum(0))
    put("W2", c["h1"].T @ r(dz2))
    dh1 = r(dz2) @ r(P["W2"]).T
    dz1 = np.where(c["z1"] > 0, dh1, 0.0)
    put("b1", dz1.sum(0))
    put("W1", c["x"].T @ r(dz1))
    return loss, g
