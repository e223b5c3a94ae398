"""XC forces (g2g_solve_groups_ computation types 2 and 3) on the GPU.

LIO's XC force on nucleus A (g2g/cpu/iteration.cpp:213-268) is
sum_p w_p y2a_p d rho_p / d R_A over the grid with the basis centres of A
moved.  The FP64 path is held to 1e-8 absolute against the forces of LIO's own
CPU path (golden vectors); the reduced-precision contractions to 1e-6 against
FP64; a molecule in a batch must give bitwise the forces it gives alone; the
Fortran ABI must add the forces onto the caller's array with its leading
dimension and leave the Fock matrix alone (g2g/partition.cpp:426-591).
"""
import numpy as np
import pytest

from golden_io import check_inputs, load, system_for

pytestmark = pytest.mark.gpu

TOL = 1e-8
TOL_MIXED = 1e-6


def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _forces(s, precision="fp64"):
    from lio_amd import g2g
    ctx = g2g.Context(s, precision=precision)
    ctx.build_grid()
    E, f, info = ctx.solve_forces()
    ctx.close()
    return E, f, info


@pytest.mark.parametrize("name", ["c0_pbe", "c0_lda1", "c0_lda2", "c0_lda3", "mol3_s101_pbe",
                                  "mol10_s102_pbe", "mol25_s103_pbe", "mol45_s104_pbe", "c1_pbe",
                                  "c0_gga5", "c0_gga8", "mol10_s102_gga4", "mol25_s103_gga7"])
def test_forces_match_golden(name):
    _gpu()
    gold = load(name)
    s = system_for(name)
    check_inputs(s, gold)
    print("case", name, "natom", s.natom, flush=True)
    E, f, _ = _forces(s)
    dE = abs(E - gold["energy_forces"])
    df = float(np.max(np.abs(f - gold["forces"])))
    print("%s dE %.2e max|df| %.2e max|f| %.3f" % (name, dE, df, np.abs(gold["forces"]).max()),
          flush=True)
    assert dE <= TOL and df <= TOL


@pytest.mark.parametrize("precision", ["fp32", "bf16x6"])
@pytest.mark.parametrize("name", ["c0_pbe", "c0_lda2", "mol25_s103_pbe", "c1_pbe"])
def test_mixed_forces_within_tolerance(name, precision):
    _gpu()
    gold = load(name)
    s = system_for(name)
    print("case", name, precision, flush=True)
    E, f, _ = _forces(s, precision)
    df = float(np.max(np.abs(f - gold["forces"])))
    print("%s %s max|df| %.2e" % (name, precision, df), flush=True)
    assert df <= TOL_MIXED


def test_forces_keep_energy_and_fock():
    """Asking for forces changes neither the energy nor the Fock matrix."""
    _gpu()
    from lio_amd import g2g
    s = system_for("mol10_s102_pbe")
    ctx = g2g.Context(s)
    ctx.build_grid()
    E0, F0, i0 = ctx.solve()
    F1 = np.zeros_like(F0)
    E1, f1, i1 = ctx.solve_forces(fock=F1)
    E2, f2, _ = ctx.solve_forces()
    # a leading dimension above natom: rows natom .. ld-1 stay untouched
    ld = s.natom + 3
    out = np.full(3 * ld, 7.0)
    out.reshape(3, ld)[:, :s.natom] = 0.0
    E3, f3, _ = ctx.solve_forces(ld=ld, out=out)
    ctx.close()
    assert E0 == E1 == E2 == E3 and i0["electrons"] == i1["electrons"]
    assert np.array_equal(F0, F1) and np.array_equal(f1, f2) and np.array_equal(f1, f3)
    assert np.all(out.reshape(3, ld)[:, s.natom:] == 7.0)


def test_g2g_abi_forces():
    """Types 2 and 3 through the Fortran ABI: forces added onto the caller's
    (max_atoms x 3, column-major) array, energy assigned for type 2 (and for
    type 3 only with energy_all_iterations), Fock untouched."""
    _gpu()
    from lio_amd import g2g
    s = system_for("mol10_s102_pbe")
    gold = load("mol10_s102_pbe")
    run = g2g.G2G()
    run.setup(s, energy_all_iterations=False)
    run.fock[:] = 0.25
    n = s.natom
    base = np.linspace(-1.0, 1.0, 3 * n).reshape(3, n)
    forces = np.asfortranarray(base.T.copy())  # (natom, 3) Fortran order, ld = natom
    e2 = run.solve(2, forces)
    assert abs(e2 - gold["energy_forces"]) <= TOL
    assert np.max(np.abs((forces - base.T) - gold["forces"])) <= TOL
    assert np.all(run.fock == 0.25)
    forces[:] = 0.0
    e3 = run.solve(3, forces)
    assert e3 == 0.0
    assert np.max(np.abs(forces - gold["forces"])) <= TOL
    assert np.all(run.fock == 0.25)
    run.deinit()


@pytest.mark.parametrize("precision", ["fp64", "bf16x6"])
def test_batch_forces_equal_alone(precision):
    torch = _gpu()
    from lio_amd import g2g, systems
    sys_ = systems.molecule_batch(5, seed=11, atoms=(3, 14))
    b = g2g.Batch(sys_, precision=precision)
    first = b.atoms()
    rho = torch.from_numpy(b.packed_density()).cuda()
    F = torch.zeros_like(rho)
    d_f = torch.zeros(3 * int(first[-1]), dtype=torch.float64, device="cuda")
    b.solve_forces(rho.data_ptr(), F.data_ptr(), d_f.data_ptr())
    E, _ = b.energies()
    torch.cuda.synchronize()
    fb = d_f.cpu().numpy().reshape(-1, 3)
    for k, s in enumerate(sys_):
        Ea, fa, _ = _forces(s, precision)
        print("molecule", k, s.natom, "atoms", flush=True)
        assert E[k] == Ea
        assert np.array_equal(fb[first[k]:first[k + 1]], fa)
    b.close()
