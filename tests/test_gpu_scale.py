"""Batched and served builds at the sizes of the batched workload (64
molecules of 3-60 atoms in one batch) and the served workload (QM regions of
10-200 atoms), on the GPU.

* A batch of 64 molecules -- 62 from the seeded batch generator and two
  58/60-atom molecules whose results LIO's own CPU path fixed as golden
  vectors -- in FP64 and split-BF16: the golden-covered molecules within
  1e-8 (FP64) and 1e-6 (split-BF16, the bar on Exc, every F_ij and N), every
  molecule of the mixed batch within 1e-6 of the FP64 batch, and every
  molecule bitwise equal to itself solved alone.
* A 195-atom molecule (M = 2175) against its golden vectors in both modes,
  with the XC forces.
* Three replicas of 150-200 atoms served from a device slab under a budget
  that forces evictions: every served step equals the same step alone.

The measured errors are printed and, when the directory named by
LIO_AMD_RESULTS_DIR exists, appended to scale_errors.jsonl there.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from golden_io import check_inputs, fock_errors, load, system_for

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OUT = os.environ.get("LIO_AMD_RESULTS_DIR", "")  # where to append the error records, if set
GOLD_BATCH = ["mol58_s105_pbe", "mol60_s106_pbe"]
TOL = {"fp64": 1e-8, "bf16x6": 1e-6}


def _torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _record(**rec):
    print(json.dumps(rec), flush=True)
    if OUT and os.path.isdir(OUT):
        with open(os.path.join(OUT, "scale_errors.jsonl"), "a") as f:
            f.write(json.dumps(rec) + "\n")


def _batch_systems():
    from lio_amd import systems
    return systems.molecule_batch(62, seed=3) + [system_for(n) for n in GOLD_BATCH]


def _batch_solve(b, torch):
    rho = torch.from_numpy(b.packed_density()).cuda()
    F = torch.zeros_like(rho)
    torch.cuda.synchronize()  # the batch runs on its own stream
    b.solve_device(rho.data_ptr(), F.data_ptr())
    E, N = b.energies()
    torch.cuda.synchronize()
    return E, N, F.cpu().numpy()


_BATCH = {}


def _batch(precision, torch):
    if precision not in _BATCH:
        from lio_amd import g2g
        sys_ = _batch_systems()
        b = g2g.Batch(sys_, precision=precision)
        _BATCH[precision] = (sys_, b.offsets.copy()) + _batch_solve(b, torch)
        b.close()
    return _BATCH[precision]


@pytest.mark.parametrize("precision", ["fp64", "bf16x6"])
def test_batch_of_64_golden_molecules(precision):
    torch = _torch()
    sys_, off, E, N, F = _batch(precision, torch)
    print("batch of %d molecules, %d atoms, %s" % (len(sys_), sum(s.natom for s in sys_), precision),
          flush=True)
    for j, name in enumerate(GOLD_BATCH):
        k = len(sys_) - len(GOLD_BATCH) + j
        gold = load(name)
        check_inputs(sys_[k], gold)
        dE = E[k] - gold["energy"]
        dF = float(np.max(np.abs(F[off[k]:off[k + 1]] - gold["fock"])))
        dN = N[k] - gold["electrons"]
        _record(test="batch64_golden", case=name, precision=precision, dE=dE, dF=dF, dN=dN)
        assert abs(dE) <= TOL[precision] and dF <= TOL[precision]
        assert abs(dN) <= TOL[precision]


def test_batch_of_64_mixed_against_fp64_per_molecule():
    torch = _torch()
    _, off, E64, N64, F64 = _batch("fp64", torch)
    sys_, _, E, N, F = _batch("bf16x6", torch)
    dE = E - E64
    dN = N - N64
    dF = np.array([np.max(np.abs(F[off[k]:off[k + 1]] - F64[off[k]:off[k + 1]]))
                   for k in range(len(sys_))])
    _record(test="batch64_vs_fp64", max_dE=float(np.max(np.abs(dE))),
            max_dF=float(dF.max()), max_dN=float(np.max(np.abs(dN))),
            worst_molecule_atoms=int(sys_[int(np.argmax(np.abs(dE)))].natom))
    assert np.max(np.abs(dE)) <= 1e-6 and dF.max() <= 1e-6 and np.max(np.abs(dN)) <= 1e-6


@pytest.mark.parametrize("precision", ["fp64", "bf16x6"])
def test_batch_of_64_each_molecule_equals_alone(precision):
    torch = _torch()
    from lio_amd import g2g
    sys_, off, E, N, F = _batch(precision, torch)
    for k, s in enumerate(sys_):
        ctx = g2g.Context(s, precision=precision)
        ctx.build_grid()
        Ea, Fa, info = ctx.solve()
        ctx.close()
        assert E[k] == Ea and N[k] == info["electrons"], (k, s.natom)
        assert np.array_equal(F[off[k]:off[k + 1]], Fa), (k, s.natom)
    print("all %d molecules equal alone (%s)" % (len(sys_), precision), flush=True)


@pytest.mark.parametrize("precision", ["fp64", "bf16x6"])
def test_195_atom_molecule_against_golden(precision):
    _torch()
    from lio_amd import g2g
    name = "mol195_s107_pbe"
    gold = load(name)
    s = system_for(name)
    check_inputs(s, gold)
    ctx = g2g.Context(s, precision=precision)
    info = ctx.build_grid()
    E, F, sinfo = ctx.solve()
    Ef, forces, _ = ctx.solve_forces()
    ctx.close()
    dE = E - gold["energy"]
    dF, drow = fock_errors(F, gold)
    dN = sinfo["electrons"] - gold["electrons"]
    dfo = float(np.max(np.abs(forces - gold["forces"])))
    _record(test="mol195", precision=precision, M=s.M, groups=info["ngroups"],
            arena_GB=info["arena_bytes"] / 1e9, dE=dE, dF_sample=dF, drowsum=drow, dN=dN,
            dforce=dfo)
    assert info["ngroups"] == len(gold["groups"])
    # Exc, N and the forces exactly; F on 100k sampled entries (every diagonal
    # one) and through every packed row's sum, in units of sqrt(row length)
    assert abs(dE) <= TOL[precision] and abs(dN) <= TOL[precision]
    assert dF <= TOL[precision] and drow <= TOL[precision]
    assert dfo <= TOL[precision]
    if precision == "bf16x6":  # the default split-BF16 form keeps half the bar here
        assert max(abs(dE), abs(dN), dF) <= 0.5e-6


CHILD = r"""
import json, sys
sys.path.insert(0, %(root)r)
from lio_amd import g2g, serve, systems
reps = [serve.Replica(m, seed=90 + k)
        for k, m in enumerate(systems.molecule_batch(3, seed=41, atoms=(150, 200)))]
srv = serve.Server(reps, precision="bf16x6", slots=2, hbm_budget=%(slab)d, grid_mode="device",
                   slab=True)
# room for about two replicas: the third forces an eviction
srv.budget = int(2.2 * max(srv._estimate(k) for k in range(len(reps))))
res = srv.run(serve.poisson_arrivals(6, 20.0, len(reps), seed=5), max_seconds=300.0)
srv.close()
out = {"served": [[q.replica, q.step, q.energy, q.electrons] for q in res["requests"]],
       "evictions": res["evictions"], "atoms": [r.base.natom for r in reps]}
print("RESULT " + json.dumps(out), flush=True)
"""


def test_large_replicas_served_equal_alone():
    _torch()
    from lio_amd import g2g, serve, systems
    slab = 48 << 30
    print("child: 3 replicas of 150-200 atoms on a %d GiB slab" % (slab >> 30), flush=True)
    r = subprocess.run([sys.executable, "-u", "-c", CHILD % {"root": ROOT, "slab": slab}],
                       capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    line = [x for x in r.stdout.splitlines() if x.startswith("RESULT ")][-1]
    child = json.loads(line[len("RESULT "):])
    reps = [serve.Replica(m, seed=90 + k)
            for k, m in enumerate(systems.molecule_batch(3, seed=41, atoms=(150, 200)))]
    assert [x.base.natom for x in reps] == child["atoms"]
    assert len(child["served"]) == 6
    for rep, step, E, N in child["served"]:
        s = reps[rep].state(step)
        ctx = g2g.Context(s, precision="bf16x6", grid_mode="device")
        ctx.build_grid()
        E1, _, info = ctx.solve(s.rho_packed)
        ctx.close()
        print("replica", rep, "step", step, E, E1, flush=True)
        assert E == E1 and N == info["electrons"]
    print("served %d steps with %d evictions" % (len(child["served"]), child["evictions"]),
          flush=True)
