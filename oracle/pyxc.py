"""ctypes front end for the CPU checkers (test infrastructure only).

* ``Oracle``  -> ``oracle/liboracle.so``: this repository's FP64 restatement of
  LIO's CPU g2g path (``g2g_oracle.cpp``).
* ``RefG2G``  -> ``oracle/_ref/libg2g_ref.so``: LIO's own CPU g2g sources built
  from the reference tree (only where that tree exists).

Only the tests, ``__graft_entry__.smoke()`` and the benchmark's CPU-baseline
leg use this module; the shipped library never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libg2g_ref.so")

_dp = C.POINTER(C.c_double)
_up = C.POINTER(C.c_uint)


class XcInput(C.Structure):
    _fields_ = [("natom", C.c_int), ("M", C.c_int), ("grid_type", C.c_int),
                ("iexch", C.c_int), ("normalize", C.c_int),
                ("r", _dp), ("Iz", _up), ("Rm", _dp), ("Nr", _up), ("Nr2", _up),
                ("Nuc", _up), ("ncont", _up), ("nshell", _up), ("a", _dp), ("c", _dp),
                ("rho", _dp), ("e", _dp * 3), ("wang", _dp * 3),
                ("little_cube_size", C.c_double), ("sphere_radius", C.c_double),
                ("assign_all_functions", C.c_int), ("remove_zero_weights", C.c_int),
                ("min_points_per_cube", C.c_uint), ("max_function_exponent", C.c_uint),
                ("open", C.c_int), ("rho_b", _dp)]


class XcGroupInfo(C.Structure):
    _fields_ = [("kind", C.c_int), ("npoints", C.c_int), ("nfuncs", C.c_int),
                ("s", C.c_int), ("p", C.c_int), ("d", C.c_int), ("nnuc", C.c_int),
                ("wsum", C.c_double), ("xsum", C.c_double), ("fhash", C.c_ulonglong)]


DEFAULT_OPTIONS = dict(little_cube_size=8.0, sphere_radius=0.6, assign_all_functions=0,
                       remove_zero_weights=1, min_points_per_cube=1,
                       max_function_exponent=10)


def _ptr(arr, ctype):
    return arr.ctypes.data_as(C.POINTER(ctype))


def make_input(system, **options):
    """Builds an XcInput for a ``lio_amd.systems.System``; returns (struct, keepalive)."""
    ab = system.abi_arrays()
    keep = {}

    def d(name, x):
        keep[name] = np.ascontiguousarray(x, dtype=np.float64)
        return _ptr(keep[name], C.c_double)

    def u(name, x):
        keep[name] = np.ascontiguousarray(x, dtype=np.uint32)
        return _ptr(keep[name], C.c_uint)

    opt = dict(DEFAULT_OPTIONS)
    opt.update(options)
    s = XcInput()
    s.natom, s.M, s.grid_type = ab["natom"], ab["M"], ab["grid_type"]
    s.iexch, s.normalize = ab["iexch"], ab["norm"]
    s.r = d("r", ab["r"])
    s.Iz = u("Iz", ab["Iz"])
    s.Rm = d("Rm", ab["Rm"])
    s.Nr = u("Nr", ab["Nr"])
    s.Nr2 = u("Nr2", ab["Nr2"])
    s.Nuc = u("Nuc", ab["Nuc"])
    s.ncont = u("ncont", ab["ncont"])
    s.nshell = u("nshell", ab["nshell"])
    s.a = d("a", ab["a"])
    s.c = d("c", ab["c"])
    if ab.get("open"):
        s.rho = d("rho", ab["rho_a"])
        s.rho_b = d("rho_b", ab["rho_b"])
        s.open = 1
    else:
        s.rho = d("rho", ab["rho"])
    for k, (en, wn) in enumerate((("e1", "w1"), ("e2", "w2"), ("e3", "w3"))):
        s.e[k] = d(en, ab[en])
        s.wang[k] = d(wn, ab[wn])
    for k, v in opt.items():
        setattr(s, k, v)
    return s, keep


class _Checker:
    def __init__(self, path):
        if not os.path.exists(path):
            raise FileNotFoundError(path + " is not built")
        self.lib = C.CDLL(path)
        self.M = 0

    def groups(self):
        n = self._groups(None, 0)
        arr = (XcGroupInfo * n)()
        self._groups(arr, n)
        return [dict(kind=g.kind, npoints=g.npoints, nfuncs=g.nfuncs, s=g.s, p=g.p, d=g.d,
                     nnuc=g.nnuc, wsum=g.wsum, xsum=g.xsum, fhash=g.fhash) for g in arr]


class RefG2G(_Checker):
    """LIO's CPU g2g (FP64) via the harness. Process-global state: one system at a time."""

    def __init__(self, path=REF_SO):
        super().__init__(path)
        L = self.lib
        L.ref_setup.argtypes = [C.POINTER(XcInput)]
        L.ref_setup.restype = C.c_double
        L.ref_solve.argtypes = [_dp, _dp, _dp]
        L.ref_solve.restype = C.c_double
        L.ref_solve_ct.argtypes = [C.c_uint, _dp, _dp, _dp, _dp]
        L.ref_solve_ct.restype = C.c_double
        L.ref_solve_open.argtypes = [C.c_uint, _dp, _dp, _dp, _dp, _dp, _dp]
        L.ref_solve_open.restype = C.c_double
        L.ref_electrons.restype = C.c_double
        L.ref_groups.argtypes = [C.POINTER(XcGroupInfo), C.c_int]
        L.ref_groups.restype = C.c_int

    def setup(self, system, **options):
        self._inp, self._keep = make_input(system, **options)
        self.M = system.M
        self.natom = system.natom
        return self.lib.ref_setup(C.byref(self._inp))

    def solve(self, rho=None):
        e = C.c_double(0.0)
        F = np.zeros(self.M * (self.M + 1) // 2)
        rp = None if rho is None else _ptr(np.ascontiguousarray(rho, np.float64), C.c_double)
        t = self.lib.ref_solve(rp, C.byref(e), _ptr(F, C.c_double))
        return e.value, F, t

    def solve_forces(self, rho=None, computation_type=2):
        """(energy, forces (natom, 3), seconds) from g2g_solve_groups_(2 or 3)."""
        e = C.c_double(0.0)
        f = np.zeros(3 * self.natom)
        rp = None if rho is None else _ptr(np.ascontiguousarray(rho, np.float64), C.c_double)
        t = self.lib.ref_solve_ct(computation_type, rp, C.byref(e), None, _ptr(f, C.c_double))
        return e.value, f.reshape(3, self.natom).T.copy(), t

    def solve_open(self, computation_type=0):
        """Open shell: (energy, F_alpha, F_beta, forces (natom, 3), seconds)."""
        e = C.c_double(0.0)
        P = self.M * (self.M + 1) // 2
        Fa, Fb = np.zeros(P), np.zeros(P)
        f = np.zeros(3 * self.natom)
        t = self.lib.ref_solve_open(computation_type, None, None, C.byref(e), _ptr(Fa, C.c_double),
                                    _ptr(Fb, C.c_double), _ptr(f, C.c_double))
        if t < 0:
            raise RuntimeError("ref_solve_open on a closed-shell problem")
        return e.value, Fa, Fb, f.reshape(3, self.natom).T.copy(), t

    def electrons(self):
        return self.lib.ref_electrons()

    def _groups(self, arr, n):
        return self.lib.ref_groups(arr, n)

    def teardown(self):
        self.lib.ref_teardown()


class Oracle(_Checker):
    """This repository's FP64 CPU restatement of the g2g path."""

    def __init__(self, path=ORACLE_SO):
        super().__init__(path)
        L = self.lib
        L.oracle_setup.argtypes = [C.POINTER(XcInput)]
        L.oracle_setup.restype = C.c_void_p
        L.oracle_solve.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.oracle_solve.restype = C.c_double
        L.oracle_solve_forces.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.oracle_solve_forces.restype = C.c_double
        L.oracle_solve_open.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.oracle_solve_open.restype = C.c_double
        L.oracle_solve_sample.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, _dp, _dp]
        L.oracle_solve_sample.restype = C.c_double
        L.oracle_groups.argtypes = [C.c_void_p, C.POINTER(XcGroupInfo), C.c_int]
        L.oracle_groups.restype = C.c_int
        L.oracle_group_work.argtypes = [C.c_void_p, C.POINTER(C.c_longlong)]
        L.oracle_group_work.restype = C.c_int
        L.oracle_free.argtypes = [C.c_void_p]
        L.oracle_grid_seconds.argtypes = [C.c_void_p]
        L.oracle_grid_seconds.restype = C.c_double
        L.oracle_gga_point.argtypes = [C.c_double, _dp, _dp, _dp, C.c_int, _dp, _dp, _dp]
        L.oracle_gga_point.restype = None
        L.oracle_lda_point.argtypes = [C.c_double, C.c_int, _dp, _dp, _dp]
        L.oracle_lda_point.restype = None
        L.oracle_basis_values.argtypes = [C.POINTER(XcInput), C.c_int, _dp, C.c_int, _dp]
        L.oracle_basis_values.restype = C.c_int
        L.oracle_gga_points.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.oracle_gga_points.restype = None
        L.oracle_gga_points_x.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_int, _dp, _dp, _dp]
        L.oracle_gga_points_x.restype = None
        L.oracle_gga_points_open.argtypes = [C.c_int] + [_dp] * 12
        L.oracle_gga_points_open.restype = None
        L.oracle_lda_points.argtypes = [C.c_int, _dp, C.c_int, _dp, _dp, _dp]
        L.oracle_lda_points.restype = None
        L.oracle_pw92.argtypes = [C.c_double, _dp, _dp]
        L.oracle_pw92.restype = None
        L.oracle_grid_points.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_int)]
        L.oracle_grid_points.restype = C.c_longlong
        L.oracle_group_functions.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        L.oracle_group_functions.restype = C.c_int
        self.h = None

    def setup(self, system, **options):
        self.free()
        self._inp, self._keep = make_input(system, **options)
        self.M = system.M
        self.natom = system.natom
        self.h = self.lib.oracle_setup(C.byref(self._inp))
        if not self.h:
            raise RuntimeError("oracle_setup failed")
        return self.lib.oracle_grid_seconds(self.h)

    def solve(self, rho=None):
        e = C.c_double(0.0)
        n = C.c_double(0.0)
        F = np.zeros(self.M * (self.M + 1) // 2)
        rp = None if rho is None else _ptr(np.ascontiguousarray(rho, np.float64), C.c_double)
        t = self.lib.oracle_solve(self.h, rp, C.byref(e), _ptr(F, C.c_double), C.byref(n))
        self._nelec = n.value
        return e.value, F, t

    def solve_forces(self, rho=None):
        """(energy, forces (natom, 3), seconds): computation type 2, no Fock."""
        e = C.c_double(0.0)
        n = C.c_double(0.0)
        f = np.zeros(3 * self.natom)
        rp = None if rho is None else _ptr(np.ascontiguousarray(rho, np.float64), C.c_double)
        t = self.lib.oracle_solve_forces(self.h, rp, C.byref(e), _ptr(f, C.c_double), C.byref(n))
        self._nelec = n.value
        return e.value, f.reshape(self.natom, 3), t

    def solve_open(self, forces=False, fock=True):
        """Open shell: (energy, F_alpha, F_beta, forces (natom, 3) or None, seconds)."""
        e = C.c_double(0.0)
        n = C.c_double(0.0)
        P = self.M * (self.M + 1) // 2
        Fa, Fb = np.zeros(P), np.zeros(P)
        f = np.zeros(3 * self.natom)
        t = self.lib.oracle_solve_open(self.h, None, None, C.byref(e),
                                       _ptr(Fa, C.c_double) if fock else None,
                                       _ptr(Fb, C.c_double) if fock else None,
                                       _ptr(f, C.c_double) if forces else None, C.byref(n))
        if t < 0:
            raise RuntimeError("oracle_solve_open on a closed-shell problem")
        self._nelec = n.value
        return e.value, Fa, Fb, (f.reshape(self.natom, 3) if forces else None), t

    def solve_sample(self, first, count, rho=None):
        """Solves only groups [first, first+count) of the cost-sorted list; returns
        (seconds, points covered)."""
        pts = C.c_double(0.0)
        e = C.c_double(0.0)
        rp = None if rho is None else _ptr(np.ascontiguousarray(rho, np.float64), C.c_double)
        t = self.lib.oracle_solve_sample(self.h, rp, first, count, C.byref(e), C.byref(pts))
        return t, pts.value

    def electrons(self):
        return self._nelec

    def group_work(self):
        """(n_groups, 4) int64 array: kind, npoints, nfuncs, flops (10 n m^2)."""
        n = self.lib.oracle_group_work(self.h, None)
        out = np.zeros((n, 4), np.int64)
        self.lib.oracle_group_work(self.h, out.ctypes.data_as(C.POINTER(C.c_longlong)))
        return out

    def _groups(self, arr, n):
        return self.lib.oracle_groups(self.h, arr, n)

    def free(self):
        if self.h:
            self.lib.oracle_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # point-level entry points for unit tests
    def gga_point(self, rho, grad, hess1, hess2, iexch=9):
        g = np.ascontiguousarray(grad, np.float64)
        h1 = np.ascontiguousarray(hess1, np.float64)
        h2 = np.ascontiguousarray(hess2, np.float64)
        ex, ec, v = C.c_double(), C.c_double(), C.c_double()
        self.lib.oracle_gga_point(rho, _ptr(g, C.c_double), _ptr(h1, C.c_double),
                                  _ptr(h2, C.c_double), iexch, C.byref(ex), C.byref(ec),
                                  C.byref(v))
        return ex.value, ec.value, v.value

    def gga_points(self, rho, grad, hess1, hess2, iexch=9):
        rho = np.ascontiguousarray(rho, np.float64)
        n = len(rho)
        g = np.ascontiguousarray(grad, np.float64).reshape(n, 3)
        h1 = np.ascontiguousarray(hess1, np.float64).reshape(n, 3)
        h2 = np.ascontiguousarray(hess2, np.float64).reshape(n, 3)
        ex, ec, v = np.zeros(n), np.zeros(n), np.zeros(n)
        self.lib.oracle_gga_points_x(n, _ptr(rho, C.c_double), _ptr(g, C.c_double),
                                     _ptr(h1, C.c_double), _ptr(h2, C.c_double), iexch,
                                     _ptr(ex, C.c_double), _ptr(ec, C.c_double),
                                     _ptr(v, C.c_double))
        return ex, ec, v

    def gga_points_open(self, ra, ga, h1a, h2a, rb, gb, h1b, h2b):
        """Spin-polarised PBE: (ex, ec, v_alpha, v_beta) per point."""
        ra = np.ascontiguousarray(ra, np.float64)
        n = len(ra)
        arrs = [ra] + [np.ascontiguousarray(x, np.float64).reshape(n, 3) for x in (ga, h1a, h2a)]
        arrs += [np.ascontiguousarray(rb, np.float64)]
        arrs += [np.ascontiguousarray(x, np.float64).reshape(n, 3) for x in (gb, h1b, h2b)]
        outs = [np.zeros(n) for _ in range(4)]
        self.lib.oracle_gga_points_open(n, *[_ptr(a, C.c_double) for a in arrs + outs])
        return tuple(outs)

    def lda_points(self, rho, iexch):
        rho = np.ascontiguousarray(rho, np.float64)
        n = len(rho)
        ex, ec, v = np.zeros(n), np.zeros(n), np.zeros(n)
        self.lib.oracle_lda_points(n, _ptr(rho, C.c_double), iexch, _ptr(ex, C.c_double),
                                   _ptr(ec, C.c_double), _ptr(v, C.c_double))
        return ex, ec, v

    def pw92(self, srs):
        e, d = C.c_double(), C.c_double()
        self.lib.oracle_pw92(srs, C.byref(e), C.byref(d))
        return e.value, d.value

    def grid(self):
        """(points x,y,z,w as (n,4), group index per point, function lists per group)."""
        n = self.lib.oracle_grid_points(self.h, None, None)
        xyzw = np.zeros((n, 4))
        grp = np.zeros(n, np.int32)
        self.lib.oracle_grid_points(self.h, _ptr(xyzw, C.c_double), grp.ctypes.data_as(C.POINTER(C.c_int)))
        fl = []
        for g in range(int(grp.max()) + 1 if n else 0):
            m = self.lib.oracle_group_functions(self.h, g, None)
            a = np.zeros(m, np.int32)
            self.lib.oracle_group_functions(self.h, g, a.ctypes.data_as(C.POINTER(C.c_int)))
            fl.append(a)
        return xyzw, grp, fl

    def lda_point(self, rho, iexch):
        ex, ec, v = C.c_double(), C.c_double(), C.c_double()
        self.lib.oracle_lda_point(rho, iexch, C.byref(ex), C.byref(ec), C.byref(v))
        return ex.value, ec.value, v.value

    def basis_values(self, system, points):
        """(npts, M, 10) values: phi, dx, dy, dz, xx, yy, zz, xy, xz, yz."""
        inp, keep = make_input(system)
        pts = np.ascontiguousarray(points, np.float64)
        out = np.zeros((len(pts), system.M, 10))
        rc = self.lib.oracle_basis_values(C.byref(inp), len(pts), _ptr(pts, C.c_double),
                                          system.M, _ptr(out, C.c_double))
        if rc != 0:
            raise RuntimeError("oracle_basis_values failed")
        return out
