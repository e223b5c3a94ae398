"""The device slab (xcq_reserve_slab): grids carved from one up-front
reservation give bitwise the results of grids allocated one by one, blocks
return to the slab when a grid is released or outgrown, a grid larger than
the free slab fails with "out of memory" and leaves the slab usable, and the
replica server on the slab serves every step as computed alone.

The slab lasts until the process ends, so the slab side runs in a child
process; the parent computes the same builds with per-grid allocations."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, %(root)r)
from lio_amd import g2g, serve, systems
small, big = %(small)d, %(big)d
reps = [serve.Replica(m, seed=70 + k)
        for k, m in enumerate(systems.molecule_batch(4, seed=33, atoms=(3, 9)))]
g2g.reserve_slab(small)
g2g.reserve_slab(small // 2)          # a smaller second request keeps the slab
out = {"results": []}
for prec in ("fp64", "bf16x6"):
    for k, rep in enumerate(reps):
        ctx = g2g.Context(precision=prec, grid_mode="device")
        for step in range(3):        # a moving molecule: rebuilds, arenas reused or regrown
            s = rep.state(step)
            ctx.set_system(s)
            info = ctx.build_grid()
            E, F, sinfo = ctx.solve(s.rho_packed)
            out["results"].append([prec, k, step, E, sinfo["electrons"], float(np.abs(F).sum()),
                                   int(info["arena_bytes"])])
        ctx.close()
try:
    g2g.reserve_slab(big)
    out["grow"] = "accepted"
except RuntimeError as e:
    out["grow"] = str(e)
# a grid larger than the slab: refused as out of memory, the slab still serves
s = systems.molecule_batch(1, seed=5, atoms=(120, 120))[0]
ctx = g2g.Context(s, precision="fp64", grid_mode="device")
try:
    ctx.build_grid()
    out["too_big"] = "built"
except RuntimeError as e:
    out["too_big"] = str(e)
ctx.close()
s = reps[0].state(0)
ctx = g2g.Context(s, precision="fp64", grid_mode="device")
ctx.build_grid()
out["after"] = ctx.solve(s.rho_packed)[0]
ctx.close()
# the replica server on the slab: a budget of about two replicas forces
# evictions, every served step is compared with the same step alone
srv = serve.Server(reps, precision="bf16x6", slots=3, hbm_budget=small, grid_mode="device",
                   slab=True)
srv.budget = 2 * max(srv._estimate(k) for k in range(len(reps)))
res = srv.run(serve.poisson_arrivals(12, 200.0, len(reps), seed=3), max_seconds=120.0)
srv.close()
out["served"] = [[q.replica, q.step, q.energy, q.electrons] for q in res["requests"]]
out["evictions"] = res["evictions"]
print("RESULT " + json.dumps(out), flush=True)
"""


def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def test_slab_results_equal_per_grid_allocation():
    _gpu()
    from lio_amd import g2g, serve, systems
    small = 2 << 30
    code = CHILD % {"root": ROOT, "small": small, "big": 2 * small}
    print("child: slab of %d MiB" % (small >> 20), flush=True)
    r = subprocess.run([sys.executable, "-u", "-c", code], capture_output=True, text=True,
                       timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    line = [x for x in r.stdout.splitlines() if x.startswith("RESULT ")][-1]
    child = json.loads(line[len("RESULT "):])
    assert "already reserved" in child["grow"]
    assert "out of memory" in child["too_big"]
    reps = [serve.Replica(m, seed=70 + k)
            for k, m in enumerate(systems.molecule_batch(4, seed=33, atoms=(3, 9)))]
    i = 0
    for prec in ("fp64", "bf16x6"):
        for k, rep in enumerate(reps):
            ctx = g2g.Context(precision=prec, grid_mode="device")
            for step in range(3):
                s = rep.state(step)
                ctx.set_system(s)
                ctx.build_grid()
                E, F, sinfo = ctx.solve(s.rho_packed)
                got = child["results"][i]
                i += 1
                assert got[:3] == [prec, k, step]
                assert got[3] == E and got[4] == sinfo["electrons"]
                assert got[5] == float(np.abs(F).sum())
            ctx.close()
    s = reps[0].state(0)
    ctx = g2g.Context(s, precision="fp64", grid_mode="device")
    ctx.build_grid()
    assert child["after"] == ctx.solve(s.rho_packed)[0]
    ctx.close()
    assert len(child["served"]) == 12
    print("served on the slab with %d evictions" % child["evictions"], flush=True)
    for r, step, E, N in child["served"]:
        s = reps[r].state(step)
        ctx = g2g.Context(s, precision="bf16x6", grid_mode="device")
        ctx.build_grid()
        E1, _, info = ctx.solve(s.rho_packed)
        ctx.close()
        assert E == E1 and N == info["electrons"]
