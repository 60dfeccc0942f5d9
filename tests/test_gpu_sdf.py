"""SDF initial conditions on the GPU (sampleParticles,
processBoundarySDFs: v2/sampling.py:384-436, 901-950): the particle sets
they keep follow from the SDFs (pinned against the reference by
tests/test_sdf.py), and the states carry the reference's keys with a neighbor
search on the device."""
import numpy as np
import pytest
import torch

from diffsph_amd import sdf

pytestmark = pytest.mark.gpu


def _config():
    dx = 2.0 / 48
    h = float(sdf.volumeToSupport(dx * dx, 50.0))
    return {"particle": {"dx": dx, "support": h},
            "domain": {"dim": 2, "minExtent": [-1.0, -1.0], "maxExtent": [1.0, 1.0], "periodicity": [False, False]},
            "kernel": {"targetNeighbors": 50, "function": None}, "simulation": {"correctArea": False},
            "compute": {"device": "cuda:0", "dtype": torch.float32}, "fluid": {"rho0": 1000.0},
            "neighborhood": {"verletScale": 1.0}}


def test_sample_particles_inside_a_fluid_region():
    cfg = _config()
    region = sdf.op_translate(lambda p: sdf.sdCircle(p, 0.5), torch.tensor([0.1, -0.2]))
    st, mask = sdf.sampleParticles(cfg, sdfs=[region])
    lat, area = sdf.sampleRegular(cfg["particle"]["dx"], 2, -1.0, 1.0, 50)
    inside = (region(lat) <= -1e-4).numpy()
    np.testing.assert_array_equal(mask.cpu().numpy(), inside)
    assert st["numParticles"] == int(inside.sum())
    np.testing.assert_array_equal(st["positions"].cpu().numpy(), lat.numpy()[inside])
    np.testing.assert_allclose(st["distances"].cpu().numpy(), region(lat).numpy()[inside], atol=1e-6)
    np.testing.assert_allclose(st["masses"].cpu().numpy(), area * 1000.0, rtol=1e-6)
    nn = st["numNeighbors"].cpu().numpy()
    print(f"{st['numParticles']} particles, neighbors {nn.min()}..{nn.max()}", flush=True)
    assert nn.max() <= 60 and nn.min() >= 10


def test_process_boundary_sdfs_removes_the_fluid_inside_the_bodies():
    cfg = _config()
    lat, area = sdf.sampleRegular(cfg["particle"]["dx"], 2, -1.0, 1.0, 50)
    dev = torch.device("cuda:0")
    n = lat.shape[0]
    fluid = {"positions": lat.to(dev), "areas": torch.full((n,), area, device=dev),
             "supports": torch.full((n,), cfg["particle"]["support"], device=dev), "numParticles": n}
    disk = lambda p: sdf.sdCircle(p, 0.25)  # noqa: E731
    bar = sdf.getSDF("orientedBox")["function"]
    bar_sdf = lambda p: bar(p, torch.tensor([-0.8, 0.6]), torch.tensor([0.2, 0.7]), 0.15)  # noqa: E731
    bst = sdf.processBoundarySDFs(fluid, cfg, [disk, bar_sdf])
    both = sdf.op_union(disk, bar_sdf)
    keep = (-both(lat) <= 1e-4).numpy()
    assert fluid["numParticles"] == int(keep.sum())
    np.testing.assert_array_equal(fluid["positions"].cpu().numpy(), lat.numpy()[keep])
    inside = [(f(lat) <= -1e-4).numpy() for f in (disk, bar_sdf)]
    expect = np.concatenate([lat.numpy()[m] for m in inside])
    np.testing.assert_array_equal(bst["positions"].cpu().numpy(), expect)
    ids = np.concatenate([np.full(m.sum(), i, np.float32) for i, m in enumerate(inside)])
    np.testing.assert_array_equal(bst["bodyIDs"].cpu().numpy(), ids)
    bp = bst["positions"].cpu()
    np.testing.assert_allclose(bst["distances"].cpu().numpy(), both(bp).numpy(), atol=1e-6)
    nrm = bst["normals"].cpu().numpy()
    np.testing.assert_allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=1e-5)
    assert bst["numParticles"] == expect.shape[0]
    assert bst["neighborhood"]["numNeighbors"].shape[0] == expect.shape[0]
