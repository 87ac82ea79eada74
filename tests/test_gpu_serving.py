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


def _oracle(r, dev):
    """The request solved by the oracle on the GPU in float32 (same inputs, same scheme)."""
    from oracle import refpath
    from torchfsm_amd import serving

    dims, ch, dt, scheme, length = serving.SPECS[r.kind]
    if r.kind == "ks1d":
        terms, mi = refpath.ks_terms(), [(0, r.param, r.n)]
    elif r.kind == "ns2d":
        mi = [(0, length, r.n)] * 2
        from torchfsm_amd.mesh import MeshGrid

        terms = refpath.ns_terms(r.param, MeshGrid(mi, device=dev).bc_mesh_grid()[1])
    else:
        terms, mi = refpath.burgers_terms(r.param), [(0, length, r.n)] * 3
    return refpath.integrate(terms, mi, r.u0.to(dev), dt, r.steps, scheme=scheme).cpu()


def test_serving_at_the_workload_classes_vs_oracle(dev):
    """The serving workload's classes: KS N in {512..4096}, NS 64^2..512^2 and 3D Burgers 64^3 and
    128^3, short requests (20-50 steps) arriving mixed.  Every served result is within rel-L2 1e-5
    of the oracle's float32 solve of the same request; one request per class served alone is
    bit-identical to its result in the mix."""
    from torchfsm_amd import serving

    sizes = {"ks1d": (512, 1024, 2048, 4096), "ns2d": (64, 128, 256, 512), "burgers3d": (64, 128)}
    reqs = serving.synthetic_requests(60, rate=300.0, seed=11, sizes=sizes, steps=(20, 50))
    classes = {(r.kind, r.n) for r in reqs}
    want = {(k, n) for k, ns in sizes.items() for n in ns}
    assert classes == want, want - classes
    srv = serving.Server(dev)
    res = srv.run(reqs)
    assert sorted(r.rid for r in res) == list(range(len(reqs)))
    print(f"served {len(res)} requests over {len(classes)} classes: {serving.latency_summary(res)}", flush=True)
    by = {r.rid: r for r in res}
    worst = {}
    for q in reqs:
        e = _rel(by[q.rid].u.numpy(), _oracle(q, dev).numpy())
        key = (q.kind, q.n)
        worst[key] = max(worst.get(key, 0.0), e)
        assert e <= 1e-5, (q.rid, key, e)
    for key in sorted(worst):
        print(f"{key}: worst rel-L2 vs oracle f32 {worst[key]:.2e}", flush=True)
    picked = {}
    for q in reqs:
        picked.setdefault((q.kind, q.n), q)
    for key, q in sorted(picked.items()):
        alone = srv.run([serving.Request(q.rid, q.kind, q.n, q.steps, q.param, q.u0, 0.0)])[0]
        assert torch.equal(alone.u, by[q.rid].u), key
    srv.close()


@pytest.mark.parametrize("case", ["ns2d_64", "ns2d_128_subbatch", "burgers3d_32"])
def test_per_sample_step_counts_2d_3d(dev, case):
    """tfsm_step_varying on 2D / 3D plans: sample b advances by exactly counts[b] steps (0: untouched),
    bit-identical to uniform stepping by that count, called directly and replayed from a captured
    graph whose counts change between replays."""
    from torchfsm_amd.engine import SpectralPlan, StepGraph

    g = torch.Generator().manual_seed(5)
    if case.startswith("ns2d"):
        n = 64 if case == "ns2d_64" else 128
        B = 5
        plan = SpectralPlan([(0, 2 * math.pi, n)] * 2, B, 1, nonlinear="vc", nonlinear_coef=-1.0,
                            linear=[("lap", [1e-3, 2e-3, 5e-3, 1e-2, 3e-3]), ("unit", -0.1)], device=dev)
        plan.set_scheme("SETDRK4", 1e-3)
        if case.endswith("subbatch"):
            tile, cv, rv, wg, _ = plan.tuning
            plan.set_tuning(tile, cv, rv, wg, 2)  # sub-batches of 2, 2, 1
        u0 = torch.randn(B, 1, n, n, generator=g)
    else:
        n, B = 32, 3
        plan = SpectralPlan([(0, 1, n)] * 3, B, 3, nonlinear="conv", nonlinear_coef=-1.0,
                            linear=[("lap", [0.01, 0.02, 0.05])], device=dev)
        plan.set_scheme("SETDRK2", 2e-3)
        u0 = 0.3 * torch.randn(B, 3, n, n, n, generator=g)
    uh0 = plan.forward(u0.to(dev))

    def uniform(k):
        u = uh0.clone()
        plan.step(u, k)
        return u

    counts_a = [3, 0, 7, 5, 1][:B]
    counts_b = [2, 4, 0, 1, 6][:B]
    want_a = {k: uniform(k) for k in set(counts_a)}
    want_ab = {}
    for b in range(B):
        u = want_a[counts_a[b]].clone()
        plan.step(u, counts_b[b])
        want_ab[b] = u[b]
    # direct call
    u = uh0.clone()
    cnt = torch.tensor(counts_a, dtype=torch.int32, device=dev)
    plan.step_varying(u, cnt, max(counts_a))
    for b in range(B):
        assert torch.equal(u[b], want_a[counts_a[b]][b]), (case, b)
    # graph of 8 steps, replayed twice with different counts
    u = uh0.clone()
    cnt = torch.zeros(B, dtype=torch.int32, device=dev)
    graph = StepGraph(plan, u, 8, counts=cnt)
    cnt.copy_(torch.tensor(counts_a, dtype=torch.int32))
    graph.launch()
    torch.cuda.synchronize(dev)
    for b in range(B):
        assert torch.equal(u[b], want_a[counts_a[b]][b]), (case, "graph", b)
    cnt.copy_(torch.tensor(counts_b, dtype=torch.int32))
    graph.launch()
    torch.cuda.synchronize(dev)
    for b in range(B):
        assert torch.equal(u[b], want_ab[b]), (case, "graph replay", b)
    graph.close()
    plan.close()
