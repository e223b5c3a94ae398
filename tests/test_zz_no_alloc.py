"""xcq_batch_solve_forces_static: the forces entry point that allocates no
device memory (its split-BF16 force scratch lives in the library's static
device array), called through the C ABI, against xcq_batch_solve_forces on
the same inputs: bitwise the same Fock matrix, forces, energies and electron
counts, on two calls in a row, the second with a smaller input.  A batch
whose scratch exceeds the static array is refused with a message.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _run(b, torch, static):
    """One forces build of batch b through the C ABI; returns host copies."""
    L = b.L
    rho = torch.from_numpy(b.packed_density()).cuda()
    F = torch.full_like(rho, 0.25)  # F is accumulated: start from the same values
    natom = int(b.atoms()[-1])
    f = torch.full((3 * natom,), -7.0, dtype=torch.float64, device="cuda")  # overwritten
    torch.cuda.synchronize()  # the batch runs on its own stream
    fn = L.xcq_batch_solve_forces_static if static else L.xcq_batch_solve_forces
    rc = fn(b.h, C.c_void_p(rho.data_ptr()), C.c_void_p(F.data_ptr()), C.c_void_p(f.data_ptr()))
    assert rc == 0, L.xcq_last_error().decode()
    E, N = b.energies()
    torch.cuda.synchronize()
    return E, N, F.cpu().numpy(), f.cpu().numpy()


@pytest.mark.parametrize("precision,variant", [("bf16x6", 2), ("bf16x6", 7), ("bf16x6", 10),
                                               ("fp64", None)])
def test_static_forces_equal_allocating_forces(precision, variant):
    torch = _torch()
    from lio_amd import g2g, systems
    L = g2g.load_library()
    assert L.xcq_static_scratch_bytes() == 64 << 20
    # two calls in a row, the second on a smaller input
    for mols in (systems.molecule_batch(2, seed=31, atoms=(5, 8)), [systems.config0()]):
        print("batch of", [m.natom for m in mols], "atoms", precision, variant, flush=True)
        b = g2g.Batch(mols, precision=precision)
        if variant is not None:
            b.set_schedule(density_variant=variant)
        E1, N1, F1, f1 = _run(b, torch, static=True)
        E0, N0, F0, f0 = _run(b, torch, static=False)
        E2, N2, F2, f2 = _run(b, torch, static=True)  # after the allocating call too
        b.close()
        for E, N, F, f in ((E1, N1, F1, f1), (E2, N2, F2, f2)):
            assert np.array_equal(E, E0) and np.array_equal(N, N0)
            assert np.array_equal(F, F0) and np.array_equal(f, f0)
        assert np.all(f0 != -7.0)


def test_static_forces_refuse_a_grid_beyond_the_scratch():
    torch = _torch()
    from lio_amd import g2g, systems
    b = g2g.Batch([systems.config1()], precision="bf16x6")
    rho = torch.from_numpy(b.packed_density()).cuda()
    F = torch.zeros_like(rho)
    f = torch.zeros(3 * int(b.atoms()[-1]), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(RuntimeError, match="static scratch holds 64 MiB"):
        b.solve_forces_static(rho.data_ptr(), F.data_ptr(), f.data_ptr())
    b.close()
