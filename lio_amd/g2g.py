"""Python front end of ``libg2g.so`` (ctypes; no torch types cross the boundary).

Two ways in, mirroring how the library is used:

* :class:`G2G` drives the Fortran-style g2g ABI exactly as lioamber does
  (``g2g_set_options_`` -> ``g2g_init_`` -> ``g2g_parameter_init_`` ->
  ``g2g_reload_atom_positions_`` -> ``g2g_solve_groups_``), with the caller
  owning every array, energy assigned and the packed Fock accumulated.
* :class:`Context` uses the handle API (``xcq_*``) for several molecules at once
  and for timing with device-resident buffers.
* :class:`Batch` keeps many molecules resident together and builds all their
  XC Fock matrices with one launch per phase (optionally one hipGraph).

The shared library must be built (``__graft_entry__.build()`` or
``make -C lio_amd/csrc``); there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# LIO_AMD_LIBRARY names another build of the library (e.g. the host code
# built with AddressSanitizer by `make -C lio_amd/csrc asan`)
LIB_PATH = os.environ.get("LIO_AMD_LIBRARY") or os.path.join(HERE, "libg2g.so")

_dp = C.POINTER(C.c_double)
_up = C.POINTER(C.c_uint)
_bp = C.POINTER(C.c_bool)


class XcqProblem(C.Structure):
    _fields_ = [("natom", C.c_uint), ("max_atoms", C.c_uint), ("M", C.c_uint),
                ("ngaussians", C.c_uint), ("normalize", C.c_int), ("iexch", C.c_uint),
                ("r", _dp), ("Rm", _dp), ("Iz", _up), ("Nr", _up), ("Nr2", _up),
                ("Nuc", _up), ("ncont", _up), ("nshell", _up), ("c", _dp), ("a", _dp),
                ("e", _dp * 3), ("wang", _dp * 3),
                ("little_cube_size", C.c_double), ("sphere_radius", C.c_double),
                ("assign_all_functions", C.c_int), ("remove_zero_weights", C.c_int),
                ("min_points_per_cube", C.c_uint), ("max_function_exponent", C.c_uint),
                ("open", C.c_int)]


class XcqGroupInfo(C.Structure):
    _fields_ = [("kind", C.c_int), ("npoints", C.c_int), ("nfuncs", C.c_int),
                ("s", C.c_int), ("p", C.c_int), ("d", C.c_int), ("nnuc", C.c_int),
                ("wsum", C.c_double), ("xsum", C.c_double), ("fhash", C.c_ulonglong)]


class XcqGridInfo(C.Structure):
    _fields_ = [("ngroups", C.c_int), ("npoints", C.c_longlong), ("npoints_padded", C.c_longlong),
                ("flops", C.c_double), ("bytes_values", C.c_double),
                ("grid_seconds", C.c_double), ("upload_seconds", C.c_double),
                ("eval_ms", C.c_double), ("arena_bytes", C.c_ulonglong), ("density_tiles", C.c_int),
                ("fock_items", C.c_int), ("load_ms", C.c_double * 4)]


class XcqSolveInfo(C.Structure):
    _fields_ = [("energy", C.c_double), ("electrons", C.c_double), ("seconds", C.c_double),
                ("kernel_ms", C.c_double * 4)]


_LIB = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Loads libg2g.so once; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise FileNotFoundError(
            "%s is not built: run __graft_entry__.build() or make -C lio_amd/csrc" % path)
    L = C.CDLL(path)
    vp = C.c_void_p
    L.xcq_last_error.restype = C.c_char_p
    L.xcq_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.xcq_destroy.argtypes = [vp]
    L.xcq_set_problem.argtypes = [vp, C.POINTER(XcqProblem)]
    L.xcq_build_grid.argtypes = [vp, C.c_uint]
    L.xcq_set_grid_mode.argtypes = [vp, C.c_int]
    L.xcq_release_device.argtypes = [vp]
    L.xcq_solve_async.argtypes = [vp, _dp]
    L.xcq_query.argtypes = [vp]
    L.xcq_collect.argtypes = [vp, _dp, C.POINTER(XcqSolveInfo)]
    L.xcq_batch_create_ex.argtypes = [C.c_int, C.c_int, C.POINTER(XcqProblem), _up, C.c_int,
                                      C.c_int, C.POINTER(vp)]
    L.xcq_set_precision.argtypes = [vp, C.c_int]
    L.xcq_query_grid.argtypes = [vp, C.POINTER(XcqGridInfo)]
    L.xcq_groups.argtypes = [vp, C.POINTER(XcqGroupInfo), C.c_int, C.POINTER(C.c_int)]
    L.xcq_solve.argtypes = [vp, _dp, _dp, C.POINTER(XcqSolveInfo)]
    L.xcq_solve_device.argtypes = [vp, vp, vp]
    L.xcq_solve_forces.argtypes = [vp, _dp, _dp, _dp, C.c_int, C.POINTER(XcqSolveInfo)]
    L.xcq_solve_open.argtypes = [vp, _dp, _dp, _dp, _dp, _dp, C.c_int, C.POINTER(XcqSolveInfo)]
    L.xcq_batch_solve_forces.argtypes = [vp, vp, vp, vp]
    L.xcq_batch_solve_forces_static.argtypes = [vp, vp, vp, vp]
    L.xcq_static_scratch_bytes.restype = C.c_longlong
    L.xcq_batch_atoms.argtypes = [vp, C.POINTER(C.c_int)]
    L.xcq_sync_energy.argtypes = [vp, _dp, _dp]
    L.xcq_kernel_times.argtypes = [vp, _dp]
    L.xcq_stream.argtypes = [vp]
    L.xcq_stream.restype = vp
    L.xcq_host_partition.argtypes = [C.POINTER(XcqProblem), C.c_uint, C.POINTER(XcqGroupInfo),
                                     C.c_int, C.POINTER(C.c_int)]
    L.xcq_batch_create.argtypes = [C.c_int, C.c_int, C.POINTER(XcqProblem), _up, C.c_int,
                                   C.POINTER(vp)]
    L.xcq_batch_destroy.argtypes = [vp]
    L.xcq_batch_offsets.argtypes = [vp, C.POINTER(C.c_longlong)]
    L.xcq_batch_solve_device.argtypes = [vp, vp, vp]
    L.xcq_batch_capture.argtypes = [vp, vp, vp]
    L.xcq_batch_replay.argtypes = [vp]
    L.xcq_batch_energies.argtypes = [vp, _dp, _dp]
    L.xcq_batch_kernel_times.argtypes = [vp, _dp]
    L.xcq_batch_stream.argtypes = [vp]
    L.xcq_batch_stream.restype = vp
    L.xcq_batch_arena_bytes.argtypes = [vp]
    L.xcq_batch_arena_bytes.restype = C.c_longlong
    L.xcq_batch_load_ms.argtypes = [vp, _dp]
    L.xcq_batch_tune.argtypes = [vp, vp, vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                 C.c_int, _dp, C.POINTER(C.c_int)]
    L.xcq_set_schedule.argtypes = [vp, C.c_int, C.c_int]
    L.xcq_reserve_slab.argtypes = [C.c_int, C.c_ulonglong]
    L.xcq_batch_set_schedule.argtypes = [vp, C.c_int, C.c_int]
    L.xcq_set_persistent.argtypes = [vp, C.c_int]
    L.xcq_batch_set_persistent.argtypes = [vp, C.c_int]
    L.xcq_batch_get_schedule.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                         C.POINTER(C.c_int)]
    L.xcq_batch_set_bucket_slices.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.xcq_batch_group_sizes.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_longlong), C.c_int,
                                        C.POINTER(C.c_int)]
    _LIB = L
    return L


def reserve_slab(nbytes: int, device: int = -1):
    """Reserves one device slab for all grid arenas of this process (see
    xcq_reserve_slab); once per process and device."""
    L = load_library()
    _check(L.xcq_reserve_slab(device, nbytes), "xcq_reserve_slab")


def _check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError("%s failed: %s" % (what, load_library().xcq_last_error().decode()))


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


PRECISION = {"fp64": 0, "fp32": 1, "bf16x6": 2}
GRID_MODE = {"host": 0, "device": 1, "staged": 2}

DEFAULT_OPTIONS = dict(little_cube_size=8.0, sphere_radius=0.6, assign_all_functions=0,
                       remove_zero_weights=1, min_points_per_cube=1, max_function_exponent=10)


def make_problem(system, **options):
    """xcq_problem for a ``lio_amd.systems.System``; returns (struct, keepalive)."""
    ab = system.abi_arrays()
    keep = {}

    def d(k, x):
        keep[k] = np.ascontiguousarray(x, np.float64)
        return _ptr(keep[k], C.c_double)

    def u(k, x):
        keep[k] = np.ascontiguousarray(x, np.uint32)
        return _ptr(keep[k], C.c_uint)

    p = XcqProblem()
    p.natom = p.max_atoms = ab["natom"]
    p.M = p.ngaussians = ab["M"]
    p.normalize = ab["norm"]
    p.iexch = ab["iexch"]
    p.open = int(bool(ab.get("open")))
    p.r, p.Rm, p.Iz = d("r", ab["r"]), d("Rm", ab["Rm"]), u("Iz", ab["Iz"])
    p.Nr, p.Nr2 = u("Nr", ab["Nr"]), u("Nr2", ab["Nr2"])
    p.Nuc, p.ncont, p.nshell = u("Nuc", ab["Nuc"]), u("ncont", ab["ncont"]), u("nshell", ab["nshell"])
    p.c, p.a = d("c", ab["c"]), d("a", ab["a"])
    for k, (en, wn) in enumerate((("e1", "w1"), ("e2", "w2"), ("e3", "w3"))):
        p.e[k] = d(en, ab[en])
        p.wang[k] = d(wn, ab[wn])
    opt = dict(DEFAULT_OPTIONS)
    opt.update(options)
    for k, v in opt.items():
        setattr(p, k, v)
    return p, keep


def _groups_to_dicts(arr):
    return [dict(kind=g.kind, npoints=g.npoints, nfuncs=g.nfuncs, s=g.s, p=g.p, d=g.d,
                 nnuc=g.nnuc, wsum=g.wsum, xsum=g.xsum, fhash=g.fhash) for g in arr]


def host_partition(system, grid_type=None, **options):
    """Group inventory of the partition the engine builds (host only, no GPU)."""
    L = load_library()
    p, keep = make_problem(system, **options)
    gt = system.grid_type if grid_type is None else grid_type
    n = C.c_int(0)
    _check(L.xcq_host_partition(C.byref(p), gt, None, 0, C.byref(n)), "xcq_host_partition")
    arr = (XcqGroupInfo * n.value)()
    _check(L.xcq_host_partition(C.byref(p), gt, arr, n.value, C.byref(n)), "xcq_host_partition")
    return _groups_to_dicts(arr)


class Context:
    """One molecule on the device through the handle API."""

    def __init__(self, system=None, device: int = -1, precision: str = "fp64",
                 grid_mode: str = "host", **options):
        self.L = load_library()
        h = C.c_void_p()
        _check(self.L.xcq_create(device, C.byref(h)), "xcq_create")
        self.h = h
        self.M = 0
        self.set_precision(precision)
        self.set_grid_mode(grid_mode)
        if system is not None:
            self.set_system(system, **options)

    def set_precision(self, precision: str):
        """'fp64', 'fp32' or 'bf16x6' arithmetic for the two contractions."""
        _check(self.L.xcq_set_precision(self.h, PRECISION[precision]), "xcq_set_precision")
        self.precision = precision

    def set_grid_mode(self, mode: str):
        """Where the next build_grid builds grid and partition: 'host',
        'device' (points, groups, functions, weights and compaction on the
        GPU) or 'staged' (points, groups and functions on the host, weights
        and compaction on the GPU).  The result is the same."""
        _check(self.L.xcq_set_grid_mode(self.h, GRID_MODE[mode]), "xcq_set_grid_mode")
        self.grid_mode = mode

    def set_system(self, system, **options):
        self._p, self._keep = make_problem(system, **options)
        self.M = system.M
        self.system = system
        _check(self.L.xcq_set_problem(self.h, C.byref(self._p)), "xcq_set_problem")

    def build_grid(self, grid_type=None):
        gt = self.system.grid_type if grid_type is None else grid_type
        _check(self.L.xcq_build_grid(self.h, gt), "xcq_build_grid")
        return self.grid_info()

    def grid_info(self) -> dict:
        gi = XcqGridInfo()
        _check(self.L.xcq_query_grid(self.h, C.byref(gi)), "xcq_query_grid")
        out = {k: getattr(gi, k) for k, _ in XcqGridInfo._fields_}
        out["load_ms"] = list(out["load_ms"])
        return out

    def groups(self):
        n = C.c_int(0)
        _check(self.L.xcq_groups(self.h, None, 0, C.byref(n)), "xcq_groups")
        arr = (XcqGroupInfo * n.value)()
        _check(self.L.xcq_groups(self.h, arr, n.value, C.byref(n)), "xcq_groups")
        return _groups_to_dicts(arr)

    def solve(self, rho=None, fock=None):
        """One XC Fock build; returns (energy, packed F added onto ``fock``, info)."""
        rho = self.system.rho_packed if rho is None else rho
        rho = np.ascontiguousarray(rho, np.float64)
        F = np.zeros(self.M * (self.M + 1) // 2) if fock is None else fock
        info = XcqSolveInfo()
        _check(self.L.xcq_solve(self.h, _ptr(rho, C.c_double), _ptr(F, C.c_double), C.byref(info)),
               "xcq_solve")
        return info.energy, F, dict(electrons=info.electrons, seconds=info.seconds,
                                    kernel_ms=list(info.kernel_ms))

    def solve_forces(self, rho=None, fock=None, ld=None, out=None):
        """Energy and XC forces (g2g computation type 2; the Fock matrix too
        when ``fock`` is given); returns (energy, forces (natom, 3), info).
        The library adds onto a Fortran-layout array of leading dimension
        ``ld`` (default natom): ``out`` (3 * ld values) when given."""
        rho = self.system.rho_packed if rho is None else rho
        rho = np.ascontiguousarray(rho, np.float64)
        n = self.system.natom
        ld = n if ld is None else ld
        f = np.zeros(3 * ld) if out is None else out
        info = XcqSolveInfo()
        fp = None if fock is None else _ptr(fock, C.c_double)
        _check(self.L.xcq_solve_forces(self.h, _ptr(rho, C.c_double), fp, _ptr(f, C.c_double), ld,
                                       C.byref(info)), "xcq_solve_forces")
        return info.energy, f.reshape(3, ld)[:, :n].T.copy(), dict(electrons=info.electrons)

    def solve_open(self, fock=True, forces=False):
        """Open-shell build from the system's alpha and beta densities;
        returns (energy, F_alpha, F_beta, forces (natom, 3) or None, info)."""
        s = self.system
        ra = np.ascontiguousarray(s.rho_a_packed, np.float64)
        rb = np.ascontiguousarray(s.rho_b_packed, np.float64)
        P = self.M * (self.M + 1) // 2
        Fa, Fb = (np.zeros(P), np.zeros(P)) if fock else (None, None)
        n = s.natom
        f = np.zeros(3 * n) if forces else None
        info = XcqSolveInfo()
        _check(self.L.xcq_solve_open(self.h, _ptr(ra, C.c_double), _ptr(rb, C.c_double),
                                     None if Fa is None else _ptr(Fa, C.c_double),
                                     None if Fb is None else _ptr(Fb, C.c_double),
                                     None if f is None else _ptr(f, C.c_double), n, C.byref(info)),
               "xcq_solve_open")
        ff = None if f is None else f.reshape(3, n).T.copy()
        return info.energy, Fa, Fb, ff, dict(electrons=info.electrons)

    def solve_device(self, d_rho: int, d_fock: int):
        """Queues a solve on device pointers (integers); no synchronisation."""
        _check(self.L.xcq_solve_device(self.h, C.c_void_p(d_rho), C.c_void_p(d_fock)),
               "xcq_solve_device")

    def sync_energy(self):
        e, n = C.c_double(), C.c_double()
        _check(self.L.xcq_sync_energy(self.h, C.byref(e), C.byref(n)), "xcq_sync_energy")
        return e.value, n.value

    def kernel_times(self):
        """Waits for the stream; ms of [expand, density, fock, pack] of the last solve."""
        out = np.zeros(4)
        _check(self.L.xcq_kernel_times(self.h, _ptr(out, C.c_double)), "xcq_kernel_times")
        return out

    def stream(self) -> int:
        return self.L.xcq_stream(self.h) or 0

    def solve_async(self, rho=None):
        """Queues a build (density copied at once); see done() and collect()."""
        rho = self.system.rho_packed if rho is None else rho
        self._rho_async = np.ascontiguousarray(rho, np.float64)
        _check(self.L.xcq_solve_async(self.h, _ptr(self._rho_async, C.c_double)), "xcq_solve_async")

    def done(self) -> bool:
        r = self.L.xcq_query(self.h)
        if r < 0:
            _check(r, "xcq_query")
        return r == 1

    def collect(self):
        """Waits for the queued build; (energy, packed F, info)."""
        F = np.zeros(self.M * (self.M + 1) // 2)
        info = XcqSolveInfo()
        _check(self.L.xcq_collect(self.h, _ptr(F, C.c_double), C.byref(info)), "xcq_collect")
        return info.energy, F, dict(electrons=info.electrons)

    def set_schedule(self, fock_slice=0, density_variant=-1):
        """Kernel schedule for the next build_grid (0 / -1 keep a knob)."""
        _check(self.L.xcq_set_schedule(self.h, fock_slice, density_variant), "xcq_set_schedule")

    def set_persistent(self, on: bool):
        """One persistent launch per build (split-BF16 default, closed shell)
        from the next build_grid; bitwise the results of the phase launches."""
        _check(self.L.xcq_set_persistent(self.h, int(bool(on))), "xcq_set_persistent")

    def release_device(self):
        """Frees the device memory of this molecule; build_grid restores it."""
        _check(self.L.xcq_release_device(self.h), "xcq_release_device")

    def close(self):
        if getattr(self, "h", None):
            self.L.xcq_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class G2G:
    """The g2g Fortran ABI driven as lioamber drives it (process-global state)."""

    def __init__(self):
        self.L = L = load_library()
        L.g2g_set_options_.argtypes = [_dp, _dp, _dp, _bp, _bp, _bp, _up, _up, _up, _up]
        L.g2g_reload_atom_positions_.argtypes = [_up]
        L.g2g_new_grid_.argtypes = [_up]
        L.g2g_solve_groups_.argtypes = [_up, _dp, _dp]

    def setup(self, system, energy_all_iterations=True, timer_level=0, verbose=0, **options):
        L = self.L
        opt = dict(DEFAULT_OPTIONS)
        opt.update(options)
        k = self._keep = {}

        def D(v):
            a = np.array([v], np.float64)
            k[id(a)] = a
            return _ptr(a, C.c_double)

        def U(v):
            a = np.array([v], np.uint32)
            k[id(a)] = a
            return _ptr(a, C.c_uint)

        def B(v):
            a = (C.c_bool * 1)(bool(v))
            k[id(a)] = a
            return C.cast(a, _bp)

        L.g2g_set_options_(D(0.0), D(opt["little_cube_size"]), D(opt["sphere_radius"]),
                           B(opt["assign_all_functions"]), B(energy_all_iterations),
                           B(opt["remove_zero_weights"]), U(opt["min_points_per_cube"]),
                           U(opt["max_function_exponent"]), U(timer_level), U(verbose))
        L.g2g_init_()
        ab = system.abi_arrays()
        self.M = ab["M"]
        arr = self._arr = {}
        for name in ("r", "Rm", "a", "c", "rho", "e1", "e2", "e3", "w1", "w2", "w3"):
            arr[name] = np.ascontiguousarray(ab[name], np.float64)
        for name in ("Iz", "Nr", "Nr2", "Nuc", "ncont", "nshell"):
            arr[name] = np.ascontiguousarray(ab[name], np.uint32)
        packed = self.M * (self.M + 1) // 2
        arr["fock"] = np.zeros(packed)
        self.open = bool(ab.get("open"))
        arr["fockb"] = np.zeros(packed if self.open else 1)
        if self.open:
            arr["rhoa"] = np.ascontiguousarray(ab["rho_a"], np.float64)
            arr["rhob"] = np.ascontiguousarray(ab["rho_b"], np.float64)
        dp = lambda n: _ptr(arr[n], C.c_double)
        up = lambda n: _ptr(arr[n], C.c_uint)
        args = [U(ab["norm"]), U(ab["natom"]), U(ab["max_atoms"]), U(ab["ngaussians"]),
                dp("r"), dp("Rm"), up("Iz"), up("Nr"), up("Nr2"), up("Nuc"), U(ab["M"]),
                up("ncont"), up("nshell"), dp("c"), dp("a"), dp("rho"), dp("fock"), dp("fockb"),
                dp("rhoa") if self.open else None, dp("rhob") if self.open else None,
                U(ab["nco"]), B(self.open), U(2 if self.open else 0), U(0), U(ab["iexch"]),
                dp("e1"), dp("e2"), dp("e3"), dp("w1"), dp("w2"), dp("w3"), B(False), U(0), U(0)]
        L.g2g_parameter_init_(*args)
        self._grid_type = U(ab["grid_type"])
        L.g2g_reload_atom_positions_(self._grid_type)

    @property
    def rho(self) -> np.ndarray:
        """The caller-owned packed density the library reads at each solve."""
        return self._arr["rho"]

    @property
    def fock(self) -> np.ndarray:
        """The caller-owned packed Fock the library accumulates into (alpha
        in open shell)."""
        return self._arr["fock"]

    @property
    def fock_b(self) -> np.ndarray:
        """Open shell: the caller-owned packed beta Fock."""
        return self._arr["fockb"]

    def solve(self, computation_type: int = 0, forces=None) -> float:
        """g2g_solve_groups_(computation_type, energy, forces): ``forces`` is
        the caller's (max_atoms x 3, Fortran order) array for types 2 and 3."""
        e = np.zeros(1)
        ct = np.array([computation_type], np.uint32)
        fp = None if forces is None else _ptr(forces, C.c_double)
        self.L.g2g_solve_groups_(_ptr(ct, C.c_uint), _ptr(e, C.c_double), fp)
        return float(e[0])

    def deinit(self):
        self.L.g2g_deinit_()


class Batch:
    """Several molecules resident on one device (``xcq_batch_*``).

    Packed densities and Fock matrices are concatenated in molecule order;
    ``offsets[k]`` is the first packed entry of molecule k and ``offsets[-1]``
    the total.  Per-molecule results are bitwise those of the molecule alone.
    """

    def __init__(self, systems, device: int = -1, precision: str = "fp64",
                 grid_mode: str = "host", **options):
        self.L = load_library()
        self.systems = list(systems)
        n = len(self.systems)
        probs = (XcqProblem * n)()
        self._keep = []
        for k, s in enumerate(self.systems):
            p, keep = make_problem(s, **options)
            probs[k] = p
            self._keep.append(keep)
        gts = np.array([s.grid_type for s in self.systems], np.uint32)
        h = C.c_void_p()
        _check(self.L.xcq_batch_create_ex(device, n, probs, _ptr(gts, C.c_uint),
                                          PRECISION[precision], GRID_MODE[grid_mode], C.byref(h)),
               "xcq_batch_create_ex")
        self.h = h
        self.n = n
        self.precision = precision
        off = np.zeros(n + 1, np.int64)
        _check(self.L.xcq_batch_offsets(self.h, off.ctypes.data_as(C.POINTER(C.c_longlong))),
               "xcq_batch_offsets")
        self.offsets = off

    def packed_density(self):
        """The concatenated packed densities of the systems (host array); for
        open-shell systems all alpha blocks, then all beta blocks."""
        if self.systems[0].open:
            return np.concatenate([s.rho_a_packed for s in self.systems] +
                                  [s.rho_b_packed for s in self.systems])
        return np.concatenate([s.rho_packed for s in self.systems])

    def fock_of(self, F, k: int):
        """Molecule k's packed Fock out of a concatenated array."""
        return F[self.offsets[k]:self.offsets[k + 1]]

    def solve_device(self, d_rho: int, d_fock: int):
        _check(self.L.xcq_batch_solve_device(self.h, C.c_void_p(d_rho), C.c_void_p(d_fock)),
               "xcq_batch_solve_device")

    def solve_forces(self, d_rho: int, d_fock: int, d_forces: int):
        """Full build plus the forces of all batch atoms (device, atom-major)."""
        _check(self.L.xcq_batch_solve_forces(self.h, C.c_void_p(d_rho), C.c_void_p(d_fock),
                                             C.c_void_p(d_forces)), "xcq_batch_solve_forces")

    def solve_forces_static(self, d_rho: int, d_fock: int, d_forces: int):
        """solve_forces without any device allocation (the force pass's
        scratch from the library's static device array; see
        xcq_batch_solve_forces_static)."""
        _check(self.L.xcq_batch_solve_forces_static(self.h, C.c_void_p(d_rho),
                                                    C.c_void_p(d_fock), C.c_void_p(d_forces)),
               "xcq_batch_solve_forces_static")

    def atoms(self):
        """First batch atom of each molecule, then the total."""
        a = np.zeros(self.n + 1, np.int32)
        _check(self.L.xcq_batch_atoms(self.h, a.ctypes.data_as(C.POINTER(C.c_int))),
               "xcq_batch_atoms")
        return a

    def capture(self, d_rho: int, d_fock: int):
        _check(self.L.xcq_batch_capture(self.h, C.c_void_p(d_rho), C.c_void_p(d_fock)),
               "xcq_batch_capture")

    def replay(self):
        _check(self.L.xcq_batch_replay(self.h), "xcq_batch_replay")

    def energies(self):
        """Waits; (Exc, N) arrays of all molecules."""
        e, n = np.zeros(self.n), np.zeros(self.n)
        _check(self.L.xcq_batch_energies(self.h, _ptr(e, C.c_double), _ptr(n, C.c_double)),
               "xcq_batch_energies")
        return e, n

    def kernel_times(self):
        out = np.zeros(4)
        _check(self.L.xcq_batch_kernel_times(self.h, _ptr(out, C.c_double)),
               "xcq_batch_kernel_times")
        return out

    def stream(self) -> int:
        return self.L.xcq_batch_stream(self.h) or 0

    def arena_bytes(self) -> int:
        return int(self.L.xcq_batch_arena_bytes(self.h))

    def load_ms(self):
        """Host wall time (ms) of the last load's phases: point weights
        (device grids), planning, uploads, evaluation."""
        ms = np.zeros(4)
        _check(self.L.xcq_batch_load_ms(self.h, _ptr(ms, C.c_double)), "xcq_batch_load_ms")
        return ms.tolist()

    def set_schedule(self, fock_slice=0, density_variant=-1):
        """Sets the kernel schedule and reloads (0 / -1 keep a knob)."""
        _check(self.L.xcq_batch_set_schedule(self.h, fock_slice, density_variant),
               "xcq_batch_set_schedule")

    def set_persistent(self, on: bool):
        """One persistent launch per build (split-BF16 density variant 8,
        closed shell; other variants ignore it); reloads.  Bitwise the
        results of variant 8's phase launches."""
        _check(self.L.xcq_batch_set_persistent(self.h, int(bool(on))), "xcq_batch_set_persistent")

    def schedule(self) -> dict:
        """The batch's schedule knobs: fock_slice, density_variant, persistent."""
        a, b, c = C.c_int(0), C.c_int(0), C.c_int(0)
        _check(self.L.xcq_batch_get_schedule(self.h, C.byref(a), C.byref(b), C.byref(c)),
               "xcq_batch_get_schedule")
        return {"fock_slice": a.value, "density_variant": b.value, "persistent": bool(c.value)}

    def tune(self, d_rho: int, d_fock: int, slices=(256, 512, 1024, 2048), variants=None,
             reps: int = 5):
        """Autotunes the schedule on these buffers over the candidates
        (slices[c], variants[c]) -- or every slice with the current density
        variant when variants is None; returns (best (slice, variant or None),
        {candidate: ms per build})."""
        sl = np.array(slices, np.int32)
        va = None if variants is None else np.array(variants, np.int32)
        if va is not None and len(va) != len(sl):
            raise ValueError("one density variant per candidate")
        ms = np.zeros(len(sl))
        best = C.c_int(-1)
        vp_ = None if va is None else va.ctypes.data_as(C.POINTER(C.c_int))
        _check(self.L.xcq_batch_tune(self.h, C.c_void_p(d_rho), C.c_void_p(d_fock), reps,
                                     sl.ctypes.data_as(C.POINTER(C.c_int)), vp_, len(sl),
                                     _ptr(ms, C.c_double), C.byref(best)), "xcq_batch_tune")
        keys = [(int(a), None if va is None else int(b))
                for a, b in zip(sl, va if va is not None else sl)]
        return keys[best.value], dict(zip(keys, ms.tolist()))

    def group_sizes(self):
        """(padded function count, padded point count) of every group."""
        n = C.c_int(0)
        _check(self.L.xcq_batch_group_sizes(self.h, None, None, 0, C.byref(n)),
               "xcq_batch_group_sizes")
        m = np.zeros(n.value, np.int32)
        p = np.zeros(n.value, np.int64)
        _check(self.L.xcq_batch_group_sizes(self.h, m.ctypes.data_as(C.POINTER(C.c_int)),
                                            p.ctypes.data_as(C.POINTER(C.c_longlong)), n.value,
                                            C.byref(n)), "xcq_batch_group_sizes")
        return m, p

    def set_bucket_slices(self, m_hi, slices):
        """Fock slice per group size bucket (groups of at most m_hi[i] padded
        functions take slices[i]); reloads."""
        mh = np.ascontiguousarray(m_hi, np.int32)
        sl = np.ascontiguousarray(slices, np.int32)
        if len(mh) != len(sl):
            raise ValueError("one slice per bucket")
        _check(self.L.xcq_batch_set_bucket_slices(self.h, len(mh),
                                                  mh.ctypes.data_as(C.POINTER(C.c_int)),
                                                  sl.ctypes.data_as(C.POINTER(C.c_int))),
               "xcq_batch_set_bucket_slices")

    def _time_builds(self, d_rho: int, d_fock: int, reps: int) -> float:
        self.solve_device(d_rho, d_fock)  # warm-up
        self.energies()
        t0 = time.perf_counter()
        for _ in range(reps):
            self.solve_device(d_rho, d_fock)
        self.energies()
        return 1e3 * (time.perf_counter() - t0) / reps

    def autotune(self, d_rho: int, d_fock: int, path: str | None = None,
                 buckets=(64, 128, 192, 256, 384), slices=(256, 512, 1024, 2048),
                 reps: int = 3, persistent=(False, True)) -> dict:
        """Schedule of this batch per group size bucket, persisted as JSON.

        With ``path`` naming an existing schedule file written for the same
        bucket bounds and precision, the schedule is applied as it is.
        Otherwise: the Fock slice of each bucket (groups of at most
        buckets[i] padded functions; larger groups form the last bucket) is
        chosen by timing whole builds on these buffers, one bucket at a time,
        heaviest bucket (sum of n m^2) first, the others held at their best so
        far; then, in split-BF16 mode, one persistent launch (density variant
        8) against one launch per phase with the batch's density variant.
        The result is written to ``path`` (when given) and returned: bucket
        bounds and slices, the launch form and density variant, the ms per
        build of every candidate timed."""
        bounds = [int(b) for b in buckets]
        if path and os.path.exists(path):
            with open(path) as f:
                sched = json.load(f)
            if sched.get("bucket_mhi", [])[:len(bounds)] == bounds and \
                    sched.get("precision") == self.precision:
                self.set_bucket_slices(sched["bucket_mhi"], sched["bucket_slice"])
                if "density_variant" in sched:
                    self.set_schedule(density_variant=int(sched["density_variant"]))
                self.set_persistent(bool(sched.get("persistent")))
                sched["loaded"] = True
                return sched
        m, p = self.group_sizes()
        nb = len(bounds) + 1
        edges = bounds + [int(max(m.max(initial=0), bounds[-1] + 32))]
        which = np.searchsorted(np.array(edges), m, side="left")
        work = np.zeros(nb)
        np.add.at(work, which, p.astype(np.float64) * m.astype(np.float64) ** 2)
        choice = [1024] * nb
        table = {}
        self.set_persistent(False)
        variant = self.schedule()["density_variant"]
        self.set_bucket_slices(edges, choice)
        best = self._time_builds(d_rho, d_fock, reps)
        table["start"] = best
        for b in np.argsort(-work):
            if work[b] == 0:
                continue
            for sl in slices:
                if sl == choice[b]:
                    continue
                trial = list(choice)
                trial[b] = int(sl)
                self.set_bucket_slices(edges, trial)
                ms = self._time_builds(d_rho, d_fock, reps)
                table["bucket<=%d slice %d" % (edges[b], sl)] = ms
                if ms < best:
                    best, choice = ms, trial
        self.set_bucket_slices(edges, choice)
        use_persistent = False
        if self.precision == "bf16x6" and len(persistent) > 1:
            self.set_schedule(density_variant=8)
            self.set_persistent(True)
            ms = self._time_builds(d_rho, d_fock, reps)
            table["persistent"] = ms
            if ms < best:
                best, use_persistent, variant = ms, True, 8
            else:
                self.set_persistent(False)
                self.set_schedule(density_variant=variant)
        sched = {"bucket_mhi": edges, "bucket_slice": [int(c) for c in choice],
                 "persistent": use_persistent, "density_variant": variant,
                 "precision": self.precision, "ms": best,
                 "bucket_groups": np.bincount(which, minlength=nb).tolist(),
                 "bucket_work": work.tolist(), "table": table, "loaded": False}
        if path:
            os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
            with open(path, "w") as f:
                json.dump(sched, f, indent=1)
        return sched

    def close(self):
        if getattr(self, "h", None):
            self.L.xcq_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
