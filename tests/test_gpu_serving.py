"""Serving loop (torchfsm_amd/serving.py): continuous batching of mixed requests.  Each result is
bit-identical alone and in the mix, and matches a plain integrate() of the same request within
rel-L2 1e-5."""

import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _plain(r, dev):
    from torchfsm_amd import field, pde, serving
    from torchfsm_amd.integrator import SETDRKIntegrator
    from torchfsm_amd.mesh import MeshGrid

    dims, ch, dt, scheme, length = serving.SPECS[r.kind]
    if r.kind == "ks1d":
        op, mesh = pde.KuramotoSivashinsky(), MeshGrid([(0, r.param, r.n)], device=dev)
    elif r.kind == "ns2d":
        mesh = MeshGrid([(0, length, r.n)] * 2, device=dev)
        op = pde.NavierStokesVorticity(r.param, field.kolm_force(mesh.bc_mesh_grid()[1]))
    else:
        mesh = MeshGrid([(0, length, r.n)] * 3, device=dev)
        op = pde.Burgers(r.param)
    op.set_integrator(getattr(SETDRKIntegrator, scheme))
    return op.integrate(r.u0.to(dev), mesh=mesh, dt=dt, step=r.steps).cpu()


def test_mixed_stream_alone_vs_mixed_and_vs_integrate(dev):
    from torchfsm_amd import serving

    sizes = {"ks1d": (512, 1024), "ns2d": (64, 128), "burgers3d": (32,)}
    reqs = serving.synthetic_requests(48, rate=400.0, seed=3, sizes=sizes, steps=(20, 120))
    srv = serving.Server(dev, slots={("ks1d", 512): 8, ("ks1d", 1024): 8, ("ns2d", 64): 8, ("ns2d", 128): 4,
                                     ("burgers3d", 32): 4})
    res = srv.run(reqs)
    assert sorted(r.rid for r in res) == list(range(48))
    lat = serving.latency_summary(res)
    print(f"mixed: {lat}")
    assert lat["p50_ms"] <= lat["p99_ms"] <= lat["max_ms"]
    by = {r.rid: r for r in res}
    for r in res:
        assert torch.isfinite(r.u).all()
        assert r.completed >= r.admitted >= r.arrival
    # one request of every class, served alone: bit-identical to its result in the mix
    picked = {}
    for q in reqs:
        picked.setdefault((q.kind, q.n), q)
    for key, q in sorted(picked.items()):
        alone = srv.run([serving.Request(q.rid, q.kind, q.n, q.steps, q.param, q.u0, 0.0)])[0]
        assert torch.equal(alone.u, by[q.rid].u), key
        e = _rel(by[q.rid].u.numpy(), _plain(q, dev).numpy())
        print(f"{key}: alone == mixed; rel-L2 vs plain integrate {e:.2e}")
        assert e <= 1e-5
    srv.close()


def test_hbm_budget_releases_idle_classes(dev):
    """With a budget below two classes, the second class is created only after the first one is
    idle and released; every request is still served."""
    from torchfsm_amd import serving

    a = serving.ShapeClass.bytes_estimate("ns2d", 64, 4)
    b = serving.ShapeClass.bytes_estimate("ns2d", 128, 4)
    srv = serving.Server(dev, hbm_budget=max(a, b) + 0.5 * min(a, b), slots={("ns2d", 64): 4, ("ns2d", 128): 4})
    reqs = serving.synthetic_requests(6, rate=1000.0, seed=5, kinds=("ns2d",),
                                      sizes={"ns2d": (64, 128)}, steps=(10, 30))
    res = srv.run(reqs)
    assert len(res) == 6
    assert len(srv.classes) == 1
    srv.close()
