"""The shipped C-ABI library: loads, exports every declared symbol, and builds
the reference's partition on the host (no GPU needed for these checks)."""
import os
import re
import subprocess

import pytest

from golden_io import load, same_partition, system_for
from lio_amd import g2g, systems
from pyxc import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in ("g2g.h", "xcq.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(\w+)\s*\(", text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof"):
                names.add(name)
    return names


def test_library_exports_every_declared_symbol():
    path = g2g.LIB_PATH
    assert os.path.exists(path), "libg2g.so is not built"
    g2g.load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    declared = _declared()
    assert len(declared) >= 37
    missing = declared - exported
    assert not missing, "declared but not exported: %s" % sorted(missing)


def test_abi_symbol_set_covers_lioamber_calls():
    """Every g2g_*/aint_* symbol lioamber calls (g2g/init.cpp, timer.cpp,
    analytic_integral/aint_init.cpp) is declared."""
    needed = {"g2g_init_", "g2g_set_options_", "g2g_parameter_init_", "g2g_reload_atom_positions_",
              "g2g_new_grid_", "g2g_solve_groups_", "g2g_deinit_", "g2g_timer_start_",
              "g2g_timer_stop_", "g2g_timer_pause_", "g2g_timer_sum_start_", "g2g_timer_sum_stop_",
              "g2g_timer_sum_pause_", "g2g_timer_clear_", "g2g_timer_summary_",
              "aint_query_gpu_level_", "aint_parameter_init_", "aint_deinit_", "aint_new_step_",
              "aint_qmmm_init_", "aint_coulomb_init_", "aint_qmmm_forces_", "aint_qmmm_fock_",
              "aint_coulomb_forces_", "aint_coulomb_fock_"}
    assert needed <= _declared()


@pytest.mark.parametrize("name", ["c0_pbe", "mol3_s101_pbe", "mol10_s102_pbe", "mol25_s103_pbe",
                                  "mol45_s104_pbe", "c1_pbe", "mol58_s105_pbe", "mol60_s106_pbe"])
def test_host_partition_matches_reference_groups(name):
    gold = load(name)
    s = system_for(name)
    assert same_partition(g2g.host_partition(s), gold["groups"])


@pytest.mark.parametrize("name", ["c0_pbe", "mol25_s103_pbe", "c1_pbe", "mol195_s107_pbe"])
def test_host_becke_weights_within_ulps_of_reference(name):
    """The Becke sums leave LIO's order (the other P_j nearest-first with an
    early exit below 2^-60 of the largest exact term, the sum in interleaved
    classes of j); the weights must stay within a few ulps of LIO's: every
    group's weight sum within 1e-14 relative of the golden inventory's (it
    is 2-4e-16 at most on these systems), and the zero weights LIO's (the
    same point counts)."""
    gold = load(name)
    mine = g2g.host_partition(system_for(name))
    assert same_partition(mine, gold["groups"], wtol=1e-14)
    assert sum(g["npoints"] for g in mine) == sum(g["npoints"] for g in gold["groups"])


@pytest.mark.parametrize("opts", [dict(sphere_radius=0.0), dict(little_cube_size=5.0),
                                  dict(assign_all_functions=1), dict(remove_zero_weights=0),
                                  dict(min_points_per_cube=40, max_function_exponent=7)])
def test_host_partition_matches_oracle_options(opts):
    s = systems.random_molecule(12, 51)
    o = Oracle()
    o.setup(s, **opts)
    assert same_partition(g2g.host_partition(s, **opts), o.groups())


@pytest.mark.parametrize("grid_type", [0, 1])
def test_host_partition_other_grids(grid_type):
    s = systems.random_molecule(9, 52)
    s.grid_type = grid_type
    o = Oracle()
    o.setup(s)
    assert same_partition(g2g.host_partition(s), o.groups())


def test_host_partition_rejects_bad_input():
    s = systems.config0()
    s.iexch = 12
    with pytest.raises(RuntimeError, match="Iexch"):
        g2g.host_partition(s)


def test_links_like_lioamber(tmp_path):
    """A program that references every symbol lioamber's Fortran calls into
    g2g (the list of the test above, Fortran's trailing underscore included)
    links against libg2g.so with --no-undefined, as liblio-g2g.so does
    (lioamber/Makefile.options:48,53); gfortran is absent, so the calls are
    declared from C with the same linkage names."""
    names = sorted({"g2g_init_", "g2g_set_options_", "g2g_parameter_init_",
                    "g2g_reload_atom_positions_", "g2g_new_grid_", "g2g_solve_groups_",
                    "g2g_deinit_", "g2g_timer_start_", "g2g_timer_stop_", "g2g_timer_pause_",
                    "g2g_timer_sum_start_", "g2g_timer_sum_stop_", "g2g_timer_sum_pause_",
                    "g2g_timer_clear_", "g2g_timer_summary_", "aint_query_gpu_level_",
                    "aint_parameter_init_", "aint_deinit_", "aint_new_step_", "aint_qmmm_init_",
                    "aint_coulomb_init_", "aint_qmmm_forces_", "aint_qmmm_fock_",
                    "aint_coulomb_forces_", "aint_coulomb_fock_"})
    src = tmp_path / "lio_calls.c"
    src.write_text("".join("extern void %s(void);\n" % n for n in names)
                   + "void* lio_table[] = {" + ", ".join("(void*)%s" % n for n in names) + "};\n"
                   + "int main(void) { return lio_table[0] == 0; }\n")
    lib_dir = os.path.dirname(g2g.LIB_PATH)
    out = tmp_path / "lio_calls"
    r = subprocess.run(["gcc", str(src), "-o", str(out), "-L" + lib_dir, "-lg2g",
                        "-Wl,--no-undefined", "-Wl,--allow-shlib-undefined"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
