"""Device grid rebuilds against the host partition, which the CPU tests pin
to LIO's own, and against LIO's group inventories directly.

* 'device': the whole rebuild on the GPU (points, cube and sphere groups,
  significant functions, group-local Becke weights, zero-weight removal,
  compaction);
* 'staged': points, groups and functions from the host, weights and
  compaction on the GPU.

Both must give the same groups, point counts and function sets as LIO
(same_partition on the golden inventories, every fixture including the
195-atom molecule), equal to each other bitwise, and the host build's
energy and Fock matrix."""
import numpy as np
import pytest

from golden_io import fock_errors, load, same_partition, system_for

pytestmark = pytest.mark.gpu


def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


# mol195: groups of up to 190 nuclei, past the device Becke kernel's LDS
# table of the 64 nearest
@pytest.mark.parametrize("name", ["c0_pbe", "mol25_s103_pbe", "mol45_s104_pbe", "mol60_s106_pbe",
                                  "mol195_s107_pbe"])
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
        # the same weights bit for bit: the sums run in the same order
        assert a["wsum"] == b["wsum"] and a["xsum"] == b["xsum"]
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


GOLDEN = ["c0_pbe", "c0_lda1", "mol3_s101_pbe", "mol10_s102_pbe", "mol25_s103_pbe",
          "mol45_s104_pbe", "c1_pbe", "mol58_s105_pbe", "mol60_s106_pbe", "mol195_s107_pbe",
          "c0_pbe_open2", "mol25_s103_pbe_open2"]


@pytest.mark.parametrize("name", GOLDEN)
def test_device_built_partition_is_lios(name):
    """The device-built groups are LIO's (kind, points, functions, shell
    counts, nuclei, function hash, weight and x sums to 1e-12), the build
    equals the host-staged one bitwise, and the FP64 result is LIO's."""
    _gpu()
    from lio_amd import g2g
    gold = load(name)
    s = system_for(name)
    print("case", name, flush=True)
    dev = g2g.Context(s, grid_mode="device")
    info = dev.build_grid()
    gd = dev.groups()
    assert same_partition(gd, gold["groups"])
    stg = g2g.Context(s, grid_mode="staged")
    stg.build_grid()
    gs = stg.groups()
    assert [(a["npoints"], a["fhash"], a["wsum"], a["xsum"]) for a in gd] == \
           [(a["npoints"], a["fhash"], a["wsum"], a["xsum"]) for a in gs]
    if s.open:
        E, Fa, Fb, _, i1 = dev.solve_open()
        E2, Fa2, Fb2, _, i2 = stg.solve_open()
        assert E == E2 and np.array_equal(Fa, Fa2) and np.array_equal(Fb, Fb2)
        F = Fa
    else:
        E, F, i1 = dev.solve()
        E2, F2, i2 = stg.solve()
        assert E == E2 and np.array_equal(F, F2)
    assert i1["electrons"] == i2["electrons"]
    dF, drow = fock_errors(F, gold)
    print("%s: %d groups, grid frame %.1f ms, device build %.1f ms, dE %.1e dF %.1e dN %.1e" % (
        name, len(gd), 1e3 * info["grid_seconds"], 1e3 * info["upload_seconds"], E - gold["energy"],
        dF, i1["electrons"] - gold["electrons"]), flush=True)
    assert abs(E - gold["energy"]) <= 1e-8 and dF <= 1e-8 and drow <= 1e-8
    assert abs(i1["electrons"] - gold["electrons"]) <= 1e-8
    dev.close()
    stg.close()


@pytest.mark.parametrize("opts", [dict(sphere_radius=0.0), dict(little_cube_size=5.0),
                                  dict(assign_all_functions=1), dict(remove_zero_weights=0),
                                  dict(min_points_per_cube=40, max_function_exponent=7)])
def test_device_built_partition_other_options(opts):
    _gpu()
    from lio_amd import g2g
    s = system_for("mol25_s103_pbe")
    print("options", opts, flush=True)
    dev = g2g.Context(s, grid_mode="device", **opts)
    dev.build_grid()
    host = g2g.Context(s, grid_mode="host", **opts)
    host.build_grid()
    assert same_partition(dev.groups(), host.groups())
    E1, F1, _ = dev.solve()
    E0, F0, _ = host.solve()
    assert abs(E1 - E0) < 1e-10 and np.max(np.abs(F1 - F0)) < 1e-10
    dev.close()
    host.close()


def test_device_built_batch_equals_alone():
    _gpu()
    import torch
    from lio_amd import g2g, systems
    sys_ = systems.molecule_batch(6, seed=22, atoms=(3, 40))
    b = g2g.Batch(sys_, grid_mode="device")
    rho = torch.from_numpy(b.packed_density()).cuda()
    F = torch.zeros_like(rho)
    torch.cuda.synchronize()
    b.solve_device(rho.data_ptr(), F.data_ptr())
    E, N = b.energies()
    torch.cuda.synchronize()
    Fh = F.cpu().numpy()
    for k, s in enumerate(sys_):
        ctx = g2g.Context(s, grid_mode="device")
        ctx.build_grid()
        Ea, Fa, info = ctx.solve()
        ctx.close()
        assert E[k] == Ea and N[k] == info["electrons"]
        assert np.array_equal(Fh[b.offsets[k]:b.offsets[k + 1]], Fa)
    b.close()
