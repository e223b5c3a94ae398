"""Host side of the replica server: seeded MD random walks and arrivals are
reproducible, and the density walk keeps its trace."""
import numpy as np

from lio_amd import serve, systems


def test_replica_steps_are_reproducible_and_keep_trace():
    s = systems.random_molecule(5, 7)
    r = serve.Replica(s, seed=3)
    a, b = r.state(4), r.state(4)
    assert np.array_equal(a.pos, b.pos) and np.array_equal(a.rho_packed, b.rho_packed)
    assert not np.array_equal(a.pos, r.state(3).pos)
    assert abs(np.trace(a.P) - np.trace(s.P)) < 1e-12 * abs(np.trace(s.P))
    # step 0 is the initial system
    z = r.state(0)
    assert np.array_equal(z.pos, s.pos) and np.array_equal(z.rho_packed, s.rho_packed)
    # per-step displacement has the requested spread (0.01 A per coordinate)
    d = (r.state(1).pos - s.pos) / serve.ANG
    assert 0.001 < np.std(d) < 0.05


def test_poisson_arrivals_reproducible():
    a = serve.poisson_arrivals(50, 20.0, 8, seed=1)
    b = serve.poisson_arrivals(50, 20.0, 8, seed=1)
    assert a == b
    t = np.array([x for x, _ in a])
    assert np.all(np.diff(t) > 0) and 1.0 < t[-1] < 6.0
    assert {r for _, r in a} <= set(range(8))


def test_slab_needs_a_budget():
    """A server on a device slab must be given the HBM budget to reserve; the
    check comes before any device call."""
    import pytest
    reps = [serve.Replica(systems.random_molecule(3, 5), seed=1)]
    with pytest.raises(ValueError):
        serve.Server(reps, slab=True)


def test_prefetched_state_is_the_state():
    """A step's inputs computed ahead on a worker thread are the same as the
    step computed on demand."""
    import concurrent.futures
    s = systems.random_molecule(6, 11)
    r1, r2 = serve.Replica(s, seed=9), serve.Replica(s, seed=9)
    with concurrent.futures.ThreadPoolExecutor(max_workers=2) as pool:
        r1.prefetch(3, pool)
        a = r1.take(3)
    b = r2.state(3)
    assert np.array_equal(a.pos, b.pos) and np.array_equal(a.rho_packed, b.rho_packed)
    # a prefetch for another step is ignored
    with concurrent.futures.ThreadPoolExecutor(max_workers=1) as pool:
        r1.prefetch(7, pool)
        c = r1.take(4)
    assert np.array_equal(c.pos, r2.state(4).pos)
