"""SDF primitives, operators, SDF filtering and the regular boundary
sampling (diffsph_amd/sdf.py) against the reference's own outputs
(tests/golden/reference_sdf.npz, written by make_reference_golden.py from
v2/sdfFunctionality and v2/sampling.py), plus closed-form checks that do not
depend on the fixture."""
import math
import os

import numpy as np
import pytest
import torch

from diffsph_amd import sdf

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_sdf.npz")
T = torch.tensor

PRIM = {
    "circle": (0.55,),
    "box": (T([0.6, 0.35]),),
    "roundedBox": (T([0.6, 0.4]), T([0.1, 0.2, 0.05, 0.15])),
    "orientedBox": (T([-0.5, -0.2]), T([0.4, 0.3]), 0.2),
    "segment": (T([-0.3, 0.1]), T([0.5, -0.4])),
    "rhombus": (T([0.7, 0.4]),),
    "trapezoid": (0.6, 0.3, 0.4),
    "parallelogram": (0.5, 0.3, 0.2),
    "equilateralTriangle": (0.6,),
    "triangleIsosceles": (T([0.4, 0.9]),),
    "triangle": (T([-0.5, -0.4]), T([0.6, -0.2]), T([0.1, 0.7])),
    "unevenCapsule": (0.3, 0.15, 0.6),
    "pentagon": (0.5,),
    "hexagon": (0.5,),
    "octogon": (0.5,),
    "polygon": (T([[-0.6, -0.5], [0.7, -0.4], [0.5, 0.6], [-0.1, 0.2], [-0.5, 0.5]]),),
}


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLD))


@pytest.mark.parametrize("name", sorted(PRIM))
def test_primitive_matches_reference(gold, name):
    p = torch.from_numpy(gold["points"])
    before = p.clone()
    got = sdf.getSDF(name)["function"](p, *PRIM[name]).numpy()
    assert torch.equal(p, before)  # the input is not modified
    np.testing.assert_allclose(got, gold[f"prim_{name}"], rtol=0, atol=2e-6)


def _ops():
    circ = lambda p: sdf.sdCircle(p, 0.55)  # noqa: E731
    box = lambda p: sdf.sdBox(p, T([0.6, 0.35]))  # noqa: E731
    od = sdf.operatorDict
    return {
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


@pytest.mark.parametrize("name", sorted(_ops()))
def test_operator_matches_reference(gold, name):
    p = torch.from_numpy(gold["points"])
    np.testing.assert_allclose(_ops()[name](p).numpy(), gold[f"op_{name}"], rtol=0, atol=2e-6)


def test_filter_and_boundary_sampling_match_reference(gold):
    lat = torch.from_numpy(gold["lattice"])
    p, area = sdf.sampleRegular(2.0 / 32, 2, -1.0, 1.0, 50)
    np.testing.assert_array_equal(p.numpy(), gold["lattice"])
    h = float(gold["h"])
    box = lambda q: sdf.sdBox(q, T([0.6, 0.35]))  # noqa: E731
    _, mask, vals, grads = sdf.filterParticlesWithSDF(lat, sdf.op_invert(box), h, 1e-4)
    np.testing.assert_array_equal(mask.numpy(), gold["filter_mask"])
    np.testing.assert_allclose(vals.numpy(), gold["filter_values"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(grads.numpy(), gold["filter_grads"], rtol=0, atol=2e-5)
    config = {"particle": {"dx": 2.0 / 32, "support": h},
              "domain": {"dim": 2, "minExtent": [-1.0, -1.0], "maxExtent": [1.0, 1.0]},
              "kernel": {"targetNeighbors": 50, "function": None}, "simulation": {"correctArea": False},
              "compute": {"device": "cpu", "dtype": torch.float32}, "fluid": {"rho0": 1.0}}
    body = lambda q: sdf.sdCircle(q, 0.3)  # noqa: E731
    bp, bv, bid, fmask = sdf.sampleBoundaryParticles({"positions": lat.clone()}, [body], config, ["regular"])
    np.testing.assert_array_equal(bp.numpy(), gold["boundary_positions"])
    np.testing.assert_allclose(bv.numpy(), gold["boundary_volumes"], rtol=1e-6)
    np.testing.assert_array_equal(bid.numpy(), gold["boundary_ids"])
    np.testing.assert_array_equal(fmask.numpy(), gold["fluid_mask"])
    _, _, bvals, bgrads = sdf.filterParticlesWithSDF(bp, sdf.op_invert(body), h, 1e-4)
    np.testing.assert_allclose((-bvals).numpy(), gold["boundary_distances"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(bgrads.numpy(), gold["boundary_normals"], rtol=0, atol=2e-5)


def test_closed_forms():
    """Distances the geometry fixes: circle and box along the axes and at a
    corner, a segment's perpendicular distance, a polygon (unit square)
    inside / outside, and the unit gradient of a circle pointing outward."""
    p = T([[1.0, 0.0], [0.0, 0.2], [0.0, 0.0]])
    np.testing.assert_allclose(sdf.sdCircle(p, 0.5).numpy(), [0.5, -0.3, -0.5], atol=1e-7)
    q = T([[1.0, 0.0], [0.0, 0.0], [2.0, 2.0]])
    np.testing.assert_allclose(sdf.sdBox(q, T([0.5, 0.25])).numpy(), [0.5, -0.25, math.hypot(1.5, 1.75)],
                               atol=1e-6)
    s = sdf.sdSegment(T([[0.0, 1.0], [2.0, 0.0]]), T([-1.0, 0.0]), T([1.0, 0.0]))
    np.testing.assert_allclose(s.numpy(), [1.0, 1.0], atol=1e-7)
    # the reference's polygon keeps the sign where the published form flips
    # it: positive inside, negative outside (the magnitude is the distance)
    sq = T([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    d = sdf.sdPolygon(T([[0.5, 0.5], [0.5, 1.5], [0.5, 0.75]]), sq)
    np.testing.assert_allclose(d.numpy(), [0.5, -0.5, 0.25], atol=1e-6)
    ang = torch.linspace(0, 2 * math.pi, 17)[:-1]
    pts = torch.stack([0.8 * torch.cos(ang), 0.8 * torch.sin(ang)], -1)
    _, _, _, g = sdf.filterParticlesWithSDF(pts, lambda x: sdf.sdCircle(x, 0.5), 0.05)
    np.testing.assert_allclose(g.numpy(), (pts / 0.8).numpy(), atol=1e-5)


def test_unbuilt_primitive_is_reported():
    with pytest.raises(NotImplementedError):
        sdf.getSDF("hexagram")
