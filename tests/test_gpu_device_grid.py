"""Device grid rebuild (group-local Becke weights and zero-weight removal on
the GPU) against the host partition, which is pinned to LIO's own by the CPU
tests: the same groups and point counts, weights equal to FP64 round-off,
and the same energy and Fock matrix."""
import numpy as np
import pytest

from golden_io import load, system_for

pytestmark = pytest.mark.gpu


def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.mark.parametrize("name", ["c0_pbe", "mol25_s103_pbe", "mol45_s104_pbe"])
def test_device_grid_matches_host(name):
    _gpu()
    from lio_amd import g2g
    s = system_for(name)
    print("case", name, flush=True)
    host = g2g.Context(s, grid_mode="host")
    host.build_grid()
    dev = g2g.Context(s, grid_mode="device")
    dev.build_grid()
    gh, gd = host.groups(), dev.groups()
    assert len(gh) == len(gd)
    for a, b in zip(gh, gd):
        assert (a["kind"], a["npoints"], a["nfuncs"], a["fhash"]) == \
               (b["kind"], b["npoints"], b["nfuncs"], b["fhash"])
        assert abs(a["wsum"] - b["wsum"]) <= 1e-12 * max(1.0, abs(a["wsum"]))
        assert a["xsum"] == b["xsum"]
    Eh, Fh, ih = host.solve()
    Ed, Fd, idd = dev.solve()
    gold = load(name)
    assert abs(Ed - Eh) < 1e-10 and abs(idd["electrons"] - ih["electrons"]) < 1e-10
    assert np.max(np.abs(Fd - Fh)) < 1e-10
    assert abs(Ed - gold["energy"]) < 1e-8
    host.close()
    dev.close()


def test_device_grid_batch_matches_host_batch():
    _gpu()
    import torch
    from lio_amd import g2g, systems
    sys_ = systems.molecule_batch(4, seed=21, atoms=(3, 12))
    out = []
    for mode in ("host", "device"):
        b = g2g.Batch(sys_, grid_mode=mode)
        rho = torch.from_numpy(b.packed_density()).cuda()
        F = torch.zeros_like(rho)
        b.solve_device(rho.data_ptr(), F.data_ptr())
        E, N = b.energies()
        out.append((E, N, F.cpu().numpy()))
        b.close()
    (E0, N0, F0), (E1, N1, F1) = out
    assert np.max(np.abs(E0 - E1)) < 1e-10 and np.max(np.abs(N0 - N1)) < 1e-10
    assert np.max(np.abs(F0 - F1)) < 1e-10


@pytest.mark.parametrize("precision", ["fp64", "bf16x6"])
def test_device_grid_forces_and_open_shell(precision):
    """Forces (computation type 2) and the open-shell build on a device-built
    grid equal those on the host-built grid (the same groups, points and
    functions; weights to 1e-12 relative): to FP64 round-off in FP64, and in
    split-BF16 to 1e-7, a tenth of the mixed-precision bar, since a weight
    differing in its last bit can round a factor-scaled value to the
    neighbouring FP32 number."""
    _gpu()
    from lio_amd import g2g, systems
    s = system_for("mol25_s103_pbe")
    so = systems.open_shell(s)
    out = {}
    for mode in ("host", "device"):
        print("forces", precision, mode, flush=True)
        ctx = g2g.Context(s, precision=precision, grid_mode=mode)
        ctx.build_grid()
        F = np.zeros(s.M * (s.M + 1) // 2)
        E, f, info = ctx.solve_forces(fock=F)
        ctx.close()
        print("open shell", precision, mode, flush=True)
        co = g2g.Context(so, precision=precision, grid_mode=mode)
        co.build_grid()
        Eo, Fa, Fb, fo, io = co.solve_open(forces=True)
        co.close()
        out[mode] = (E, F, f, info["electrons"], Eo, Fa, Fb, fo, io["electrons"])
    h, d = out["host"], out["device"]
    tol = 1e-10 if precision == "fp64" else 1e-7
    for a, b in zip(h, d):
        diff = float(np.max(np.abs(np.asarray(a) - np.asarray(b))))
        print("host vs device %.2e" % diff, flush=True)
        assert diff < tol
