"""One persistent launch per build, size-bucket schedules and the autotuner,
on the GPU.

* The split-BF16 build of density variant 8 as one persistent launch over
  the ordered queue of every phase's items gives bitwise the results of
  variant 8's phase launches: single molecules, a batch, forces, and every
  replay of a captured graph (with eager work on other buffers between
  replays).
* A Fock slice per group size bucket changes only the order of the slice
  sums: within 1e-11 of the single-slice schedule in FP64.
* The autotuner writes its schedule as JSON and a second batch applies the
  file without timing anything.
"""
import json

import numpy as np
import pytest

from golden_io import system_for

pytestmark = pytest.mark.gpu


def _torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _solve(b, torch, forces=False):
    rho = torch.from_numpy(b.packed_density()).cuda()
    F = torch.zeros_like(rho)
    f = torch.zeros(3 * int(b.atoms()[-1]), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()  # the batch runs on its own stream
    if forces:
        b.solve_forces(rho.data_ptr(), F.data_ptr(), f.data_ptr())
    else:
        b.solve_device(rho.data_ptr(), F.data_ptr())
    E, N = b.energies()
    torch.cuda.synchronize()
    return E, N, F.cpu().numpy(), f.cpu().numpy()


@pytest.mark.parametrize("name", ["c0_pbe", "c0_lda3", "mol25_s103_pbe", "c1_pbe"])
def test_persistent_build_equals_phase_launches(name):
    torch = _torch()
    from lio_amd import g2g
    s = system_for(name)
    print("case", name, flush=True)
    out = []
    for on in (False, True):
        b = g2g.Batch([s], precision="bf16x6")
        b.set_schedule(density_variant=8)
        b.set_persistent(on)
        out.append(_solve(b, torch) + _solve(b, torch, forces=True))
        b.close()
    for x, y in zip(out[0], out[1]):
        assert np.array_equal(x, y)
    assert np.all(np.isfinite(out[1][0]))


def test_persistent_batch_and_graph_replays():
    torch = _torch()
    from lio_amd import g2g, systems
    sys_ = systems.molecule_batch(6, seed=23, atoms=(3, 30))
    b0 = g2g.Batch(sys_, precision="bf16x6")
    b0.set_schedule(density_variant=8)
    E0, N0, F0, _ = _solve(b0, torch)
    b0.close()
    b = g2g.Batch(sys_, precision="bf16x6")
    b.set_schedule(density_variant=8)
    b.set_persistent(True)
    E1, N1, F1, _ = _solve(b, torch)
    assert np.array_equal(E0, E1) and np.array_equal(N0, N1) and np.array_equal(F0, F1)
    rho = torch.from_numpy(b.packed_density()).cuda()
    F = torch.zeros_like(rho)
    torch.cuda.synchronize()
    b.capture(rho.data_ptr(), F.data_ptr())
    for k in range(4):
        F.fill_(7.0)  # eager work on the buffers between replays
        torch.cuda.synchronize()
        b.replay()
        E, N = b.energies()
        torch.cuda.synchronize()
        assert np.array_equal(E, E0) and np.array_equal(N, N0), k
        assert np.array_equal(F.cpu().numpy(), F0), k
        x = torch.ones(1 << 20, device="cuda")
        assert float(x.sum()) == float(1 << 20)
    b.close()


def test_bucket_slices_keep_results():
    torch = _torch()
    from lio_amd import g2g, systems
    sys_ = systems.molecule_batch(5, seed=24, atoms=(5, 40))
    b = g2g.Batch(sys_, precision="fp64")
    E0, N0, F0, _ = _solve(b, torch)
    m, p = b.group_sizes()
    print("groups", len(m), "m_pad range", m.min(), m.max(), flush=True)
    b.set_bucket_slices([64, 160], [256, 2048])
    E1, N1, F1, _ = _solve(b, torch)
    b.close()
    assert np.max(np.abs(E1 - E0)) < 1e-11 and np.max(np.abs(N1 - N0)) < 1e-11
    assert np.max(np.abs(F1 - F0)) < 1e-11


def test_autotune_writes_and_reapplies_its_schedule(tmp_path):
    torch = _torch()
    from lio_amd import g2g, systems
    sys_ = systems.molecule_batch(6, seed=25, atoms=(3, 25))
    path = str(tmp_path / "schedule.json")
    b = g2g.Batch(sys_, precision="bf16x6")
    rho = torch.from_numpy(b.packed_density()).cuda()
    F = torch.zeros_like(rho)
    torch.cuda.synchronize()
    sched = b.autotune(rho.data_ptr(), F.data_ptr(), path=path, reps=1)
    print(json.dumps({k: sched[k] for k in ("bucket_mhi", "bucket_slice", "persistent", "ms")}),
          flush=True)
    assert not sched["loaded"] and len(sched["bucket_slice"]) == len(sched["bucket_mhi"])
    E1, N1, F1, _ = _solve(b, torch)
    b.close()
    b2 = g2g.Batch(sys_, precision="bf16x6")
    again = b2.autotune(rho.data_ptr(), F.data_ptr(), path=path)
    assert again["loaded"] and again["bucket_slice"] == sched["bucket_slice"]
    assert again["persistent"] == sched["persistent"]
    assert b2.schedule()["density_variant"] == sched["density_variant"]
    E2, N2, F2, _ = _solve(b2, torch)
    b2.close()
    assert np.array_equal(E1, E2) and np.array_equal(F1, F2)
