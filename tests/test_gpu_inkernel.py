"""Split-BF16 density with the basis evaluated inside the density kernel
(density variants 6 and 7: no value planes in HBM), on the GPU.

The same bar as the other mixed modes against LIO's FP64 golden vectors
(|dExc|, max |dF_ij|, |dN| <= 1e-6 and forces within 1e-6), closed and open
shell, and the arena must be smaller than the ten-plane layout's.  The
measured errors are printed and, when the directory named by
LIO_AMD_RESULTS_DIR exists, appended to inkernel_errors.jsonl there.
"""
import json
import os

import numpy as np
import pytest

from golden_io import check_inputs, load, system_for

pytestmark = pytest.mark.gpu

TOL = 1e-6
OUT = os.environ.get("LIO_AMD_RESULTS_DIR", "")  # where to append the error records, if set


def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _record(**rec):
    print(json.dumps(rec), flush=True)
    if OUT and os.path.isdir(OUT):
        with open(os.path.join(OUT, "inkernel_errors.jsonl"), "a") as f:
            f.write(json.dumps(rec) + "\n")


@pytest.mark.parametrize("variant", [7, 6])
@pytest.mark.parametrize("name", ["c0_pbe", "c0_lda3", "mol25_s103_pbe", "mol45_s104_pbe",
                                  "c1_pbe"])
def test_inkernel_evaluation_within_tolerance(name, variant):
    _gpu()
    from lio_amd import g2g
    gold = load(name)
    s = system_for(name)
    check_inputs(s, gold)
    print("case", name, "variant", variant, flush=True)
    ctx = g2g.Context(s, precision="bf16x6")
    planes = ctx.build_grid()
    ctx.set_schedule(density_variant=variant)
    info = ctx.build_grid()
    E, F, sinfo = ctx.solve()
    E2, F2, _ = ctx.solve()
    Ef, forces, _ = ctx.solve_forces()
    ctx.close()
    dE = E - gold["energy"]
    dF = float(np.max(np.abs(F - gold["fock"])))
    dN = sinfo["electrons"] - gold["electrons"]
    dfo = float(np.max(np.abs(forces - gold["forces"])))
    _record(case=name, variant=variant, dE=dE, dF=dF, dN=dN, dforce=dfo,
            arena_GB=info["arena_bytes"] / 1e9, arena_planes_GB=planes["arena_bytes"] / 1e9)
    assert E == E2 and np.array_equal(F, F2)
    assert abs(dE) <= TOL and dF <= TOL and abs(dN) <= TOL
    assert abs(Ef - gold["energy_forces"]) <= TOL and dfo <= TOL
    assert info["arena_bytes"] < planes["arena_bytes"]


@pytest.mark.parametrize("variant", [7, 6])
@pytest.mark.parametrize("name", ["c0_pbe_open2", "mol25_s103_pbe_open2"])
def test_inkernel_open_shell_within_tolerance(name, variant):
    _gpu()
    from lio_amd import g2g
    gold = load(name)
    s = system_for(name)
    ctx = g2g.Context(s, precision="bf16x6")
    ctx.set_schedule(density_variant=variant)
    ctx.build_grid()
    E, Fa, Fb, _, info = ctx.solve_open()
    Ef, _, _, forces, _ = ctx.solve_open(fock=False, forces=True)
    ctx.close()
    dE = E - gold["energy"]
    dFa = float(np.max(np.abs(Fa - gold["fock"])))
    dFb = float(np.max(np.abs(Fb - gold["fock_b"])))
    dN = info["electrons"] - gold["electrons"]
    dfo = float(np.max(np.abs(forces - gold["forces"])))
    _record(case=name, variant=variant, dE=dE, dFa=dFa, dFb=dFb, dN=dN, dforce=dfo)
    assert abs(dE) <= TOL and dFa <= TOL and dFb <= TOL and abs(dN) <= TOL
    assert abs(Ef - gold["energy_forces"]) <= TOL and dfo <= TOL


def test_inkernel_batch_equals_alone():
    torch = pytest.importorskip("torch")
    _gpu()
    from lio_amd import g2g, systems
    sys_ = systems.molecule_batch(5, seed=21, atoms=(3, 30))
    b = g2g.Batch(sys_, precision="bf16x6")
    b.set_schedule(density_variant=7)
    rho = torch.from_numpy(b.packed_density()).cuda()
    F = torch.zeros_like(rho)
    torch.cuda.synchronize()
    b.solve_device(rho.data_ptr(), F.data_ptr())
    E, N = b.energies()
    torch.cuda.synchronize()
    Fh = F.cpu().numpy()
    off = b.offsets
    for k, s in enumerate(sys_):
        ctx = g2g.Context(s, precision="bf16x6")
        ctx.set_schedule(density_variant=7)
        ctx.build_grid()
        Ea, Fa, info = ctx.solve()
        ctx.close()
        print("molecule", k, s.natom, "atoms", E[k], Ea, flush=True)
        assert E[k] == Ea and N[k] == info["electrons"]
        assert np.array_equal(Fh[off[k]:off[k + 1]], Fa)
    b.close()
