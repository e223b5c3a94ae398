"""Fortran timer symbols and the analytic-integral stubs of the shipped library,
driven through the C ABI on the CPU.

The expectations restate LIO's timers (g2g/timer.cpp:37-112 for the Timer
start/stop/pause meaning, 125-296 for the Fortran entry points, init.cpp:427-432
for the levels): stop reports the interval since the last start, pause adds
to what the timer holds and prints "(so far)", times print as "<s>s. <us>us."
or "<us>us.", the sum timers nest under the running one and the summary
prints the tree sorted by name with each share of its parent.  Each case runs
in a child process, because the levels only ever switch timers on.
"""
import ctypes as C
import os
import re
import subprocess
import sys
import textwrap

import pytest

from lio_amd import g2g

DRIVER = textwrap.dedent(r"""
    import ctypes as C, sys, time
    L = C.CDLL(sys.argv[1])
    level = int(sys.argv[2])
    d = lambda v: C.byref(C.c_double(v))
    b = lambda v: C.byref(C.c_bool(v))
    u = lambda v: C.byref(C.c_uint(v))
    L.g2g_set_options_(d(0.0), d(8.0), d(0.6), b(False), b(True), b(True), u(1), u(10),
                       u(level), u(0))
    def call(fn, name=None):
        if name is None:
            getattr(L, fn)()
        else:
            getattr(L, fn)(name.encode(), C.c_uint(len(name)))
        C.CDLL(None).fflush(None)
    for line in sys.stdin:
        if not line.strip():
            continue
        op, *arg = line.split()
        if op == "sleep":
            time.sleep(float(arg[0]))
        else:
            call(op, arg[0] if arg else None)
""")


def drive(level, script):
    res = subprocess.run([sys.executable, "-c", DRIVER, g2g.LIB_PATH, str(level)],
                         input=textwrap.dedent(script), capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stderr
    return res.stdout.splitlines()


def seconds(text):
    m = re.fullmatch(r"(?:(\d+)s\. )?(\d+)us\.", text)
    assert m, text
    return int(m.group(1) or 0) + int(m.group(2)) * 1e-6


def test_level_zero_prints_nothing():
    out = drive(0, """
        g2g_timer_start_ SCF
        g2g_timer_stop_ SCF
        g2g_timer_sum_start_ Total
        g2g_timer_sum_stop_ Total
        g2g_timer_summary_
    """)
    assert out == []


def test_single_timers_stop_reports_last_interval_pause_accumulates():
    out = drive(1, """
        g2g_timer_start_ SCF
        sleep 0.05
        g2g_timer_pause_ SCF
        sleep 0.2
        g2g_timer_start_ SCF
        sleep 0.05
        g2g_timer_pause_ SCF
        g2g_timer_start_ SCF
        sleep 0.03
        g2g_timer_stop_ SCF
        g2g_timer_stop_ nope
        g2g_timer_sum_start_ Total
        g2g_timer_summary_
    """)
    assert len(out) == 5, out
    m = [re.fullmatch(r"TIMER \[SCF\]: (.*?)(\(so far\))?", l) for l in out[:3]]
    assert all(m), out
    assert m[0].group(2) and m[1].group(2) and not m[2].group(2)
    t1, t2, t3 = (seconds(x.group(1)) for x in m)
    assert 0.05 <= t1 < 0.2               # first interval
    assert 0.1 <= t2 < 0.25               # both intervals, not the 0.2 s gap between them
    assert 0.03 <= t3 < 0.1               # stop: only the interval since the last start
    assert "nope" in out[3]               # unknown timer is reported
    assert re.fullmatch(r"TIMER \[nope\]: .*", out[4])
    # level 1 has no sum timers: the summary prints nothing


def test_sum_timer_tree_and_summary():
    out = drive(2, """
        g2g_timer_start_ SCF
        g2g_timer_stop_ SCF
        g2g_timer_sum_start_ Total
        g2g_timer_sum_start_ SCF
        g2g_timer_sum_start_ XC
        sleep 0.04
        g2g_timer_sum_pause_ XC
        g2g_timer_sum_start_ Exchange
        sleep 0.02
        g2g_timer_sum_stop_ Exchange
        g2g_timer_sum_start_ XC
        sleep 0.04
        g2g_timer_sum_pause_ XC
        g2g_timer_sum_stop_ Total
        g2g_timer_sum_stop_ SCF
        g2g_timer_sum_stop_ Total
        g2g_timer_summary_
    """)
    # single timers are off at level 2; the mismatched stop is reported
    assert out[0] == "Error: not the current timer: (Total,SCF)"
    rule = "-" * 91
    assert out[1] == rule and out[2].strip() == "LIO TIMING INFORMATION:" and out[3] == rule
    assert out[4] == ""
    total = float(re.fullmatch(r"Total time: (\S+)s", out[5]).group(1))
    rows = out[6:-1]
    assert out[-1] == rule
    pat = re.compile(r"( *)(\S+) +(\d+\.\d{6})s \( *(\d+\.\d\d)% of (\S+)\)")
    parsed = [pat.fullmatch(r) for r in rows]
    assert all(parsed), rows
    tree = [(len(p.group(1)) // 2, p.group(2), float(p.group(3)), float(p.group(4)), p.group(5))
            for p in parsed]
    # top level then depth-first, children sorted by name
    assert [(d, n, par) for d, n, _, _, par in tree] == [
        (0, "Total", "Total"), (1, "SCF", "Total"), (2, "Exchange", "SCF"), (2, "XC", "SCF")]
    t = {n: v for _, n, v, _, _ in tree}
    assert abs(t["Total"] - total) < 2e-6
    assert 0.02 <= t["Exchange"] < 0.06
    assert 0.08 <= t["XC"] < 0.15         # two paused intervals add up
    assert t["SCF"] >= t["XC"] + t["Exchange"] - 1e-6
    for _, n, v, pct, par in tree:
        assert abs(pct - 100.0 * v / t[par]) < 0.01


def test_timer_clear_forgets_the_tree():
    out = drive(3, """
        g2g_timer_sum_start_ Total
        g2g_timer_sum_start_ A
        g2g_timer_sum_stop_ A
        g2g_timer_sum_stop_ Total
        g2g_timer_clear_
        g2g_timer_sum_start_ Total
        g2g_timer_sum_stop_ Total
        g2g_timer_summary_
    """)
    names = [l.split()[0] for l in out if re.match(r"\S+ +\d+\.\d{6}s \(", l)]
    assert names == ["Total"]


def test_aint_stubs_report_cpu_level_and_zero_fock_terms():
    """aint_query_gpu_level_ returns 0 (aint_init.cpp:342-345) so lioamber
    keeps its CPU integrals; the Fock stubs zero Es/Ens like
    aint_init.cpp:267-269, 326-327."""
    L = g2g.load_library()
    lvl = C.c_int(7)
    L.aint_query_gpu_level_(C.byref(lvl))
    assert lvl.value == 0
    es, ens = C.c_double(3.0), C.c_double(4.0)
    L.aint_qmmm_fock_(C.byref(es), C.byref(ens))
    assert es.value == 0.0 and ens.value == 0.0
    es = C.c_double(5.0)
    L.aint_coulomb_fock_(C.byref(es))
    assert es.value == 0.0
