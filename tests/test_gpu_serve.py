"""The continuous-batching server on the GPU: every served step of every
replica equals the same step computed alone in a fresh context (bitwise),
while other replicas' builds run concurrently and the HBM budget forces
arenas to be released and rebuilt."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.mark.parametrize("precision", ["fp64", "bf16x6"])
def test_served_equals_alone(precision):
    _gpu()
    from lio_amd import g2g, serve, systems
    reps = [serve.Replica(m, seed=50 + k)
            for k, m in enumerate(systems.molecule_batch(5, seed=31, atoms=(3, 9)))]
    arrivals = serve.poisson_arrivals(15, 200.0, len(reps), seed=2)
    srv = serve.Server(reps, precision=precision, slots=4, grid_mode="device")
    # budget of about two replicas: forces releases between steps
    srv.budget = 2 * max(srv._estimate(k) for k in range(len(reps)))
    out = srv.run(arrivals, keep_fock=True, max_seconds=120.0)
    print(precision, "served", len(out["requests"]), "p50 %.1f ms p99 %.1f ms evictions %d"
          % (out["latency_ms"]["p50"], out["latency_ms"]["p99"], out["evictions"]), flush=True)
    assert len(out["requests"]) == len(arrivals)
    for q in out["requests"][:8]:
        s = reps[q.replica].state(q.step)
        ctx = g2g.Context(s, precision=precision, grid_mode="device")
        ctx.build_grid()
        E, F, info = ctx.solve(s.rho_packed)
        ctx.close()
        print("request", q.rid, "replica", q.replica, "step", q.step, q.energy, E, flush=True)
        assert q.energy == E and q.electrons == info["electrons"]
        assert np.array_equal(q.fock, F)
    srv.close()
