"""Split-BF16 density variants 10 and 11 (one block per point tile and
64-function row chunk; a second kernel adds the chunks' row-dot sums in
chunk order and evaluates the functional), on the GPU.

Variant 10 runs variant 8's products in the same K order, so its Y values are
the same bits; only the FP64 epilogue sums are grouped per chunk before they
are added.  Against variant 8 it must agree to FP64 round-off (1e-10 on Exc,
N, every F_ij and every force component: a chunk lost, counted twice or
added into the wrong point shows as a difference of the size of a chunk's
contribution, orders of magnitude more), against LIO's FP64 golden vectors
to the split-BF16 default's bar (0.5e-6), and it must be bitwise
reproducible.  Variant 11 takes the B operand's pieces from a split made
once per grid; they are the pieces variant 10 makes in its K loop, so its
results must be variant 10's bit for bit.
"""
import json

import numpy as np
import pytest

from golden_io import check_inputs, fock_errors, load, system_for

pytestmark = pytest.mark.gpu

TOL = 0.5e-6     # the split-BF16 default's bar against FP64
SAME = 1e-10     # variant 10 against variant 8: FP64 round-off of the epilogue sums


def _torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _ctx_run(s, variant, forces=True):
    from lio_amd import g2g
    ctx = g2g.Context(s, precision="bf16x6")
    ctx.set_schedule(density_variant=variant)
    ctx.build_grid()
    E, F, info = ctx.solve()
    E2, F2, _ = ctx.solve()
    out = dict(E=E, F=F, N=info["electrons"], E2=E2, F2=F2)
    if forces:
        Ef, f, _ = ctx.solve_forces()
        out.update(Ef=Ef, f=f)
    ctx.close()
    return out


@pytest.mark.parametrize("name", ["c0_pbe", "c0_lda3", "mol25_s103_pbe", "mol45_s104_pbe",
                                  "mol60_s106_pbe", "c1_pbe"])
def test_chunked_density_equals_variant8(name):
    _torch()
    gold = load(name)
    s = system_for(name)
    check_inputs(s, gold)
    print("case", name, flush=True)
    a = _ctx_run(s, 8)
    b = _ctx_run(s, 10)
    c = _ctx_run(s, 11)
    for k in ("E", "N", "Ef"):
        assert b[k] == c[k], k
    for k in ("F", "f"):
        assert np.array_equal(b[k], c[k]), k
    dF, _ = fock_errors(b["F"], gold)
    rec = dict(case=name, dE=b["E"] - gold["energy"], dF=dF, dN=b["N"] - gold["electrons"],
               dE_v8=b["E"] - a["E"], dF_v8=float(np.max(np.abs(b["F"] - a["F"]))),
               dN_v8=b["N"] - a["N"], dforce_v8=float(np.max(np.abs(b["f"] - a["f"]))))
    print(json.dumps(rec), flush=True)
    assert b["E"] == b["E2"] and np.array_equal(b["F"], b["F2"])
    assert abs(rec["dE_v8"]) <= SAME and rec["dF_v8"] <= SAME and abs(rec["dN_v8"]) <= SAME
    assert rec["dforce_v8"] <= SAME and abs(b["Ef"] - a["Ef"]) <= SAME
    assert abs(rec["dE"]) <= TOL and dF <= TOL and abs(rec["dN"]) <= TOL


def test_chunked_density_195_atoms():
    """The largest serving case (195 atoms, M = 2175: up to 35 chunks a
    tile), against the sampled golden Fock matrix and variant 8."""
    _torch()
    name = "mol195_s107_pbe"
    gold = load(name)
    s = system_for(name)
    a = _ctx_run(s, 8, forces=False)
    b = _ctx_run(s, 10, forces=False)
    c = _ctx_run(s, 11, forces=False)
    assert b["E"] == c["E"] and b["N"] == c["N"] and np.array_equal(b["F"], c["F"])
    dF, rowdev = fock_errors(b["F"], gold)
    rec = dict(case=name, dE=b["E"] - gold["energy"], dF=dF, rowdev=rowdev,
               dN=b["N"] - gold["electrons"], dE_v8=b["E"] - a["E"],
               dF_v8=float(np.max(np.abs(b["F"] - a["F"]))), dN_v8=b["N"] - a["N"])
    print(json.dumps(rec), flush=True)
    assert b["E"] == b["E2"] and np.array_equal(b["F"], b["F2"])
    assert abs(rec["dE_v8"]) <= SAME and rec["dF_v8"] <= SAME and abs(rec["dN_v8"]) <= SAME
    assert abs(rec["dE"]) <= TOL and dF <= TOL and abs(rec["dN"]) <= TOL


def test_chunked_open_shell_equals_variant8():
    _torch()
    from lio_amd import g2g
    name = "mol25_s103_pbe_open2"
    gold = load(name)
    s = system_for(name)
    res = []
    for v in (8, 10, 11):
        ctx = g2g.Context(s, precision="bf16x6")
        ctx.set_schedule(density_variant=v)
        ctx.build_grid()
        E, Fa, Fb, _, info = ctx.solve_open()
        Ef, _, _, f, _ = ctx.solve_open(fock=False, forces=True)
        ctx.close()
        res.append((E, Fa, Fb, info["electrons"], Ef, f))
    (E8, Fa8, Fb8, N8, Ef8, f8), (E, Fa, Fb, N, Ef, f), r11 = res
    assert r11[0] == E and r11[3] == N and r11[4] == Ef
    for x, y in zip(r11[1:3] + r11[5:], (Fa, Fb, f)):
        assert np.array_equal(x, y)
    print("open shell dE %.2e dN %.2e vs v8 %.2e" % (E - gold["energy"], N - gold["electrons"],
                                                    E - E8), flush=True)
    assert abs(E - E8) <= SAME and abs(N - N8) <= SAME and abs(Ef - Ef8) <= SAME
    assert np.max(np.abs(Fa - Fa8)) <= SAME and np.max(np.abs(Fb - Fb8)) <= SAME
    assert np.max(np.abs(f - f8)) <= SAME
    assert abs(E - gold["energy"]) <= TOL and abs(N - gold["electrons"]) <= TOL
    assert np.max(np.abs(Fa - gold["fock"])) <= TOL and np.max(np.abs(Fb - gold["fock_b"])) <= TOL


@pytest.mark.parametrize("variant", [10, 11])
def test_chunked_batch_equals_alone_and_graph_replays(variant):
    torch = _torch()
    from lio_amd import g2g, systems
    sys_ = systems.molecule_batch(6, seed=41, atoms=(3, 40))
    b = g2g.Batch(sys_, precision="bf16x6")
    b.set_schedule(density_variant=variant)
    rho = torch.from_numpy(b.packed_density()).cuda()
    F = torch.zeros_like(rho)
    torch.cuda.synchronize()  # the batch runs on its own stream
    b.solve_device(rho.data_ptr(), F.data_ptr())
    E0, N0 = b.energies()
    torch.cuda.synchronize()
    F0 = F.cpu().numpy()
    off = b.offsets
    for k, s in enumerate(sys_):
        ctx = g2g.Context(s, precision="bf16x6")
        ctx.set_schedule(density_variant=variant)
        ctx.build_grid()
        Ea, Fa, info = ctx.solve()
        ctx.close()
        print("molecule", k, s.natom, "atoms", E0[k], Ea, flush=True)
        assert E0[k] == Ea and N0[k] == info["electrons"]
        assert np.array_equal(F0[off[k]:off[k + 1]], Fa)
    b.capture(rho.data_ptr(), F.data_ptr())
    for k in range(4):
        F.fill_(3.0)  # eager work on the buffers between replays
        torch.cuda.synchronize()
        b.replay()
        E, N = b.energies()
        torch.cuda.synchronize()
        assert np.array_equal(E, E0) and np.array_equal(N, N0), k
        assert np.array_equal(F.cpu().numpy(), F0), k
        x = torch.ones(1 << 20, device="cuda")
        assert float(x.sum()) == float(1 << 20)
    b.close()
