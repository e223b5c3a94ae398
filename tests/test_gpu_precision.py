"""Mixed-precision contractions against LIO's FP64 golden vectors, on the GPU.

Bar for the mixed path: |dExc| <= 1e-6 Ha, max |dF_ij| <= 1e-6 and the
electron count within 1e-6 of the FP64 reference on identical inputs.  The
split-BF16 default (density variant 8) is held to half of that, 0.5e-6, on
(H2O)32 and the other cases here (and on a 195-atom molecule in
test_gpu_scale.py): headroom for larger systems.

Plain FP32 does not meet the bar on the electron count: its value planes are
FP32, and phi's FP32 rounding repeats over every point of a radial shell
(the atom's own s functions have one value there), so it does not average
out; at (H2O)32 |dN| is about 1.1e-6.  Its cases hold N to 5e-6 (TOL_N) and
Exc and F to the 1e-6 bar; the split-BF16 default is the mixed mode that
meets the bar.  The measured errors are printed
so the log records them per configuration.
"""
import json
import os

import numpy as np
import pytest

from golden_io import load, system_for

pytestmark = pytest.mark.gpu

TOL = 1e-6
TOL_MIXED_DEFAULT = 0.5e-6  # split-BF16 default: half the bar
TOL_N = {"fp32": 5e-6, "bf16x6": TOL_MIXED_DEFAULT}  # FP32 misses the bar on N (see above)
OUT = os.environ.get("LIO_AMD_RESULTS_DIR", "")  # where to append the error records, if set


def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


CASES = [(n, p) for p in ("fp32", "bf16x6")
         for n in ("c0_pbe", "c0_lda3", "mol25_s103_pbe", "mol45_s104_pbe", "c1_pbe",
                   "c0_gga4", "c0_gga6", "mol10_s102_gga4", "mol25_s103_gga7")]


@pytest.mark.parametrize("name,precision", CASES)
def test_mixed_precision_within_tolerance(name, precision):
    _gpu()
    from lio_amd import g2g
    gold = load(name)
    s = system_for(name)
    print("case", name, precision, flush=True)
    ctx = g2g.Context(s, precision=precision)
    ctx.build_grid()
    E, F, info = ctx.solve()
    dE = E - gold["energy"]
    dF = float(np.max(np.abs(F - gold["fock"])))
    dN = info["electrons"] - gold["electrons"]
    rec = dict(case=name, precision=precision, dE=dE, dF=dF, dN=dN, E=E)
    print(json.dumps(rec), flush=True)
    if OUT and os.path.isdir(OUT):
        with open(os.path.join(OUT, "precision_errors.jsonl"), "a") as f:
            f.write(json.dumps(rec) + "\n")
    tol = TOL_MIXED_DEFAULT if precision == "bf16x6" else TOL
    assert abs(dE) <= tol and dF <= tol and abs(dN) <= TOL_N[precision]
    ctx.close()


def test_precision_switch_rebuilds_values():
    """Switching precision on a built grid re-evaluates the basis: FP64 after
    FP32 and after split-BF16 reproduces the FP64 result bitwise, and each
    mixed mode is itself reproducible."""
    _gpu()
    from lio_amd import g2g
    s = system_for("mol10_s102_pbe")
    ctx = g2g.Context(s)
    ctx.build_grid()
    E64, F64, _ = ctx.solve()
    for mode in ("fp32", "bf16x6"):
        ctx.set_precision(mode)
        Em, Fm, _ = ctx.solve()
        Em2, Fm2, _ = ctx.solve()
        assert Em == Em2 and np.array_equal(Fm, Fm2)
        assert Em != E64
        ctx.set_precision("fp64")
        E64b, F64b, _ = ctx.solve()
        assert E64 == E64b and np.array_equal(F64, F64b)
    ctx.close()


@pytest.mark.parametrize("name", ["c0_pbe", "mol25_s103_pbe", "c1_pbe"])
def test_radial_hessian_layout_within_tolerance(name):
    """Split-BF16 density variant 5 keeps phi and its gradient as planes and
    rebuilds the second derivatives from per-shell radial sums: the same bar
    against FP64, with a smaller value arena than the ten-plane layout."""
    _gpu()
    from lio_amd import g2g
    gold = load(name)
    s = system_for(name)
    ctx = g2g.Context(s, precision="bf16x6")
    info10 = ctx.build_grid()
    ctx.set_schedule(density_variant=5)
    info4 = ctx.build_grid()
    E, F, info = ctx.solve()
    ctx.close()
    dE = E - gold["energy"]
    dF = float(np.max(np.abs(F - gold["fock"])))
    dN = info["electrons"] - gold["electrons"]
    print("%s radial-Hessian dE %.2e dF %.2e dN %.2e arena %.2f -> %.2f GB" %
          (name, dE, dF, dN, info10["arena_bytes"] / 1e9, info4["arena_bytes"] / 1e9), flush=True)
    assert abs(dE) <= TOL and dF <= TOL and abs(dN) <= TOL
    assert info4["arena_bytes"] < info10["arena_bytes"]
