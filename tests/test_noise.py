"""Perlin octave noise (generateOctaveNoise with kind 'perlin',
v2/noiseFunctions/generator.py:87-99 over perlin.py) against the
reference's own outputs (tests/golden/reference_noise.npz), bit for bit, and
the noise initial condition on the GPU: a divergence-free velocity field
scaled to u_mag."""
import os

import numpy as np
import pytest
import torch

from diffsph_amd.noise import generateOctaveNoise, perlin_noise_2d

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_noise.npz")


@pytest.mark.parametrize("case", ["a", "b", "c"])
def test_octave_noise_matches_reference(case):
    d = dict(np.load(GOLD))
    n, o, lac, per, bf, tile, seed, norm = d[f"{case}_params"]
    xx, yy, noise = generateOctaveNoise(int(n), 2, int(o), int(lac), float(per), int(bf), bool(tile), "perlin",
                                        seed=int(seed), normalized=bool(norm))
    np.testing.assert_array_equal(noise.numpy(), d[f"{case}_noise"])
    np.testing.assert_array_equal(xx.numpy(), d[f"{case}_xx"])
    np.testing.assert_array_equal(yy.numpy(), d[f"{case}_yy"])


def test_perlin_properties():
    """Zero at every cell corner; a tileable field's opposite edges continue
    each other (the first row equals the extrapolated one past the last)."""
    z = perlin_noise_2d(64, 4, tileable=True, seed=3)
    assert np.all(z[::16, ::16] == 0.0)
    assert np.abs(z).max() <= np.sqrt(2) * 1.0 + 1e-6
    with pytest.raises(ValueError):
        perlin_noise_2d(50, 4)


@pytest.mark.gpu
def test_noise_initial_condition_is_divergence_free():
    from diffsph_amd.noise import sampleNoisyParticles
    cfg = {"domain": {"minExtent": [-1.0, -1.0], "maxExtent": [1.0, 1.0], "periodicity": [True, True], "dim": 2},
           "fluid": {"rho0": 1.0, "u_mag": 2.0}, "kernel": {"targetNeighbors": 50},
           "neighborhood": {"verletScale": 1.0}, "compute": {"device": "cuda:0"}}
    st, _ = sampleNoisyParticles({"n": 64, "baseFrequency": 2, "octaves": 2, "seed": 11}, cfg)
    v = st["velocities"]
    assert st["positions"].shape == (4096, 2) and v.shape == (4096, 2)
    assert float(torch.linalg.norm(v, dim=1).max()) == pytest.approx(2.0, rel=1e-5)
    # the SPH divergence of the SPH curl is small next to the velocity
    # gradients (|grad u| ~ 2 pi f u_mag for base frequency f)
    from diffsph_amd.sphops import sphOperationStates
    div = sphOperationStates(st, st, (v, v), st["neighborhood"], operation="divergence", gradientMode="difference")
    gu = sphOperationStates(st, st, (v, v), st["neighborhood"], operation="gradient", gradientMode="difference")
    print(f"mean |div u| {float(div.abs().mean()):.3e}, mean |grad u| {float(gu.abs().mean()):.3e}", flush=True)
    assert float(div.abs().mean()) < 0.05 * float(gu.abs().mean())


@pytest.mark.gpu
def test_noise_initial_condition_with_a_boundary_body():
    """A circular 'boundary' body: the fluid particles inside it are removed
    (the mask of sampleBoundaryParticles: sdf >= -1e-4 kept), 'normals'
    holds the body's sdf values and 'distances' its unit gradients (the
    reference's assignment), and the velocity is still scaled to u_mag."""
    from diffsph_amd import sdf
    from diffsph_amd.noise import sampleNoisyParticles
    cfg = {"domain": {"minExtent": [-1.0, -1.0], "maxExtent": [1.0, 1.0], "periodicity": [True, True], "dim": 2},
           "fluid": {"rho0": 1.0, "u_mag": 1.5}, "kernel": {"targetNeighbors": 50},
           "neighborhood": {"verletScale": 1.0}, "compute": {"device": "cuda:0"}}
    body = lambda p: sdf.sdCircle(p, 0.3)  # noqa: E731
    st, _ = sampleNoisyParticles({"n": 64, "baseFrequency": 2, "seed": 3}, cfg, sdfs=[{"sdf": body,
                                                                                        "type": "boundary"}])
    lat, _ = sdf.sampleRegular(2.0 / 64, 2, [-1.0, -1.0], [1.0, 1.0], 50)
    keep = (body(lat) >= -1e-4).numpy()
    assert st["numParticles"] == int(keep.sum()) < 4096
    np.testing.assert_array_equal(st["positions"].cpu().numpy(), lat.numpy()[keep])
    np.testing.assert_allclose(st["normals"].cpu().numpy(), body(lat).numpy()[keep], atol=1e-6)
    g = st["distances"].cpu().numpy()
    np.testing.assert_allclose(np.linalg.norm(g, axis=1), 1.0, atol=1e-5)
    assert float(torch.linalg.norm(st["velocities"], dim=1).max()) == pytest.approx(1.5, rel=1e-5)
