"""Batched XC Fock builds on the GPU: many molecules resident together.

A batch runs every solve phase as one launch over the groups of all its
molecules.  Each molecule's energy, electron count and packed Fock must be
bitwise those of the molecule solved alone (same per-molecule summation
order), in every precision mode; a captured hipGraph replayed several times,
with eager work between the replays, must give the eager result every time;
the autotuner's choice must not change results beyond FP64 round-off.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _systems():
    from lio_amd import systems
    # small uneven molecules (3 .. 14 atoms) keep the test short
    return systems.molecule_batch(5, seed=11, atoms=(3, 14))


def _alone(s, precision):
    from lio_amd import g2g
    ctx = g2g.Context(s, precision=precision)
    ctx.build_grid()
    E, F, info = ctx.solve()
    ctx.close()
    return E, F, info["electrons"]


def _batch_solve(b, torch):
    rho = torch.from_numpy(b.packed_density()).cuda()
    F = torch.zeros_like(rho)
    b.solve_device(rho.data_ptr(), F.data_ptr())
    E, N = b.energies()
    torch.cuda.synchronize()
    return E, N, F.cpu().numpy()


@pytest.mark.parametrize("precision", ["fp64", "bf16x6"])
def test_batch_equals_alone(precision):
    torch = _torch()
    from lio_amd import g2g
    sys_ = _systems()
    print("batch of", [s.natom for s in sys_], precision, flush=True)
    b = g2g.Batch(sys_, precision=precision)
    E, N, F = _batch_solve(b, torch)
    for k, s in enumerate(sys_):
        Ea, Fa, Na = _alone(s, precision)
        print("molecule", k, s.natom, "atoms", E[k], Ea, flush=True)
        assert E[k] == Ea and N[k] == Na
        assert np.array_equal(b.fock_of(F, k), Fa)
    b.close()


def test_batch_graph_replays_match_eager():
    torch = _torch()
    from lio_amd import g2g
    sys_ = _systems()
    b = g2g.Batch(sys_, precision="fp64")
    E0, N0, F0 = _batch_solve(b, torch)
    rho = torch.from_numpy(b.packed_density()).cuda()
    F = torch.zeros_like(rho)
    b.capture(rho.data_ptr(), F.data_ptr())
    for rep in range(3):
        print("replay", rep, flush=True)
        F.zero_()
        torch.cuda.synchronize()
        b.replay()
        E, N = b.energies()
        assert np.array_equal(E, E0) and np.array_equal(N, N0)
        assert np.array_equal(F.cpu().numpy(), F0)
        # eager work between replays (other buffers)
        x = torch.ones(1 << 20, device="cuda")
        torch.cuda.synchronize()
        assert float(x.sum()) == float(1 << 20)
    b.close()


def test_batch_tuner_over_density_variants():
    """Split-BF16: every density kernel gives the same result to FP32-GEMM
    round-off (the kernels add the K steps' products in different orders into
    FP32 sums: 1e-7, a tenth of the 1e-6 bar against FP64), and the tuner
    keeps one of them."""
    torch = _torch()
    from lio_amd import g2g
    sys_ = _systems()
    b = g2g.Batch(sys_, precision="bf16x6")
    rho = torch.from_numpy(b.packed_density()).cuda()
    F = torch.zeros_like(rho)
    res = {}
    for v in (0, 1, 2, 3, 5):
        b.tune(rho.data_ptr(), F.data_ptr(), slices=(512,), variants=(v,), reps=1)
        res[v] = _batch_solve(b, torch)
    E0, N0, F0 = res[0]
    for v, (E1, N1, F1) in res.items():
        print("variant", v, E1, flush=True)
        assert np.max(np.abs(E0 - E1)) < 1e-7 and np.max(np.abs(N0 - N1)) < 1e-7
        assert np.max(np.abs(F0 - F1)) < 1e-7
    best, table = b.tune(rho.data_ptr(), F.data_ptr(), slices=(512, 512), variants=(0, 1), reps=2)
    assert best in table
    b.close()


def test_batch_tuner_keeps_results():
    torch = _torch()
    from lio_amd import g2g
    sys_ = _systems()
    b = g2g.Batch(sys_, precision="fp64")
    E0, N0, F0 = _batch_solve(b, torch)
    rho = torch.from_numpy(b.packed_density()).cuda()
    F = torch.zeros_like(rho)
    best, table = b.tune(rho.data_ptr(), F.data_ptr(), slices=(128, 512, 2048), reps=2)
    print("tuned Fock slice", best, table, flush=True)
    assert best[0] in (128, 512, 2048) and all(v > 0 for v in table.values())
    E, N, F1 = _batch_solve(b, torch)
    assert np.allclose(E, E0, rtol=0, atol=1e-11)
    assert np.allclose(N, N0, rtol=0, atol=1e-11)
    assert np.max(np.abs(F1 - F0)) < 1e-11
    b.close()
