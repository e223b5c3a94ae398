"""Caller-side builders: the molecules, basis arrays, grid tables and density
matrices that lioamber hands to the g2g C ABI.

Everything here reproduces the *data conventions* of LIO's Fortran caller, so
the arrays can be passed unchanged to ``g2g_parameter_init_``:

* Coordinates in bohr with LIO's constant 0.529177 A/bohr
  (``lioamber/SCF_exter.f90:45``); ``r`` is column-major ``natom x 3``.
* Cartesian Gaussian basis read per element in file order, atoms of one element
  in input order (``lioamber/basis_data.f90:790-870``), primitive coefficients
  scaled by the primitive norm (2a/pi)^(3/4) * (4a)^(l/2)
  (``basis_data.f90:636-660`` with ``PI32`` from ``constants_mod``), one
  function per Cartesian component (p: x,y,z; d: xx,xy,yy,xz,yz,zz) and finally
  a stable s -> p -> d reorder (``basis_data.f90:1007-1064``).
* ``a``/``c`` are column-major ``M x 13`` (``g2g/init.cpp:144-147``), ``Nuc`` is
  1-based, ``nshell = (n_s, n_p_functions, n_d_functions)``.
* Slater radii table ``Rm2 = R_slater / (2 * 0.529177)`` and shell counts indexed
  by Z (``lioamber/grid.f90:16-37``).
* Lebedev angular grids with weights multiplied by 4*pi (``grid.f90:247-477``).
* Density ``P = 2 C_occ C_occ^T`` packed column-major lower ('L') with the
  off-diagonal elements doubled (``lioamber/SCF_aux.f90:31-43``).

The overlap matrix used to orthonormalise the seeded orbitals is computed here
with the Obara-Saika recursion for Cartesian Gaussians.
"""
from __future__ import annotations

import concurrent.futures
import functools
import json
import math
import os
import dataclasses
from dataclasses import dataclass, field

import numpy as np

ANG_PER_BOHR = 0.529177          # LIO's conversion constant
PI32 = 5.56832799683170698       # pi**1.5 as LIO's constants_mod spells it
MAX_CONTRACTIONS = 13
GRID_SIZES = (50, 116, 194)

_DATA = os.path.join(os.path.dirname(__file__), "data")

# ---------------------------------------------------------------------------
# Grid tables
# ---------------------------------------------------------------------------

# Slater radii in Angstrom for Z = 0..54 (grid.f90:16-20).
_SLATER_A = [1.00, 0.35, 0.93, 1.45, 1.05, 0.85, 0.70, 0.65, 0.60, 0.50, 0.71,
             3.60, 1.50, 1.25, 1.10, 1.00, 1.00, 1.00, 2.00, 2.20, 1.80, 1.60,
             1.40, 1.35, 1.40, 1.40, 1.40, 1.35, 1.35, 1.35, 1.35, 1.30, 1.25,
             1.15, 1.15, 1.15, 1.15, 2.35, 2.00, 1.80, 1.55, 1.45, 1.45, 1.35,
             1.30, 1.35, 1.40, 1.60, 1.55, 1.55, 1.45, 1.45, 1.40, 1.40, 1.31]


def slater_rm() -> np.ndarray:
    """Atomic size parameter per Z (bohr), ``Rm2(Z) = R_slater / (2*0.529177)``."""
    return np.array(_SLATER_A, dtype=np.float64) / (2.0 * ANG_PER_BOHR)


def shell_counts(n: int = 40) -> np.ndarray:
    """Radial shells per Z. The benchmark configurations use 40 for every Z."""
    return np.full(55, n, dtype=np.uint32)


def _orbit_a1():
    pts = []
    for ax in range(3):
        for s in (1.0, -1.0):
            v = [0.0, 0.0, 0.0]
            v[ax] = s
            pts.append(v)
    return pts


def _orbit_a2():
    h = math.sqrt(0.5)
    pts = []
    for zero in (2, 1, 0):
        idx = [k for k in range(3) if k != zero]
        for s0 in (1.0, -1.0):
            for s1 in (1.0, -1.0):
                v = [0.0, 0.0, 0.0]
                v[idx[0]] = s0 * h
                v[idx[1]] = s1 * h
                pts.append(v)
    return pts


def _orbit_a3():
    t = 1.0 / math.sqrt(3.0)
    return [[sx * t, sy * t, sz * t] for sx in (1, -1) for sy in (1, -1) for sz in (1, -1)]


def _orbit_b(l, m):
    """(l, l, m) and its permutations / sign changes: 24 points."""
    pts = []
    for mpos in (2, 1, 0):
        for sx in (1, -1):
            for sy in (1, -1):
                for sz in (1, -1):
                    v = [l, l, l]
                    v[mpos] = m
                    pts.append([sx * v[0], sy * v[1], sz * v[2]])
    return pts


def _orbit_c(p, q):
    """(p, q, 0) permutations / sign changes: 24 points."""
    pts = []
    for zero in (2, 1, 0):
        idx = [k for k in range(3) if k != zero]
        for a, b in ((p, q), (q, p)):
            for s0 in (1, -1):
                for s1 in (1, -1):
                    v = [0.0, 0.0, 0.0]
                    v[idx[0]] = s0 * a
                    v[idx[1]] = s1 * b
                    pts.append(v)
    return pts


def _orbit_d(r, u, w):
    """(r, u, w) all permutations and sign changes: 48 points."""
    import itertools
    pts = []
    for perm in itertools.permutations((r, u, w)):
        for sx in (1, -1):
            for sy in (1, -1):
                for sz in (1, -1):
                    pts.append([sx * perm[0], sy * perm[1], sz * perm[2]])
    return pts


@functools.lru_cache(maxsize=None)
def _lebedev_cached(npts: int) -> tuple[np.ndarray, np.ndarray]:
    e, w = _lebedev(npts)
    e.setflags(write=False)
    w.setflags(write=False)
    return e, w


def lebedev(npts: int) -> tuple[np.ndarray, np.ndarray]:
    """(directions (n, 3), weights x 4 pi) of the n-point Lebedev rule
    (copies of a cached table)."""
    e, w = _lebedev_cached(npts)
    return e.copy(), w.copy()


def _lebedev(npts: int) -> tuple[np.ndarray, np.ndarray]:
    """Lebedev angular rule as LIO tabulates it: unit vectors (npts x 3) and
    weights summing to 4*pi.  Orbit parameters and weights carry the 12 digits
    of LIO's tables (Lebedev, Zh. Vychisl. Mat. Mat. Fiz. 15 (1975), 16 (1976)),
    so the generated grid equals the caller's; point order is by orbit."""
    if npts == 194:
        orbits = [
            (_orbit_a1(), 0.00178234044724),
            (_orbit_a2(), 0.00571690594988),
            (_orbit_a3(), 0.00557338317884),
            (_orbit_b(0.444693317871, 0.777493219315), 0.00551877146727),
            (_orbit_b(0.289246562758, 0.912509096867), 0.00515823771181),
            (_orbit_b(0.671297344270, 0.314196994183), 0.00560870408259),
            (_orbit_b(0.129933544765, 0.982972302707), 0.00410677702817),
            (_orbit_c(0.938319218138, 0.345770219761), 0.00505184606462),
            (_orbit_d(0.836036015482, 0.159041710538, 0.525118572443), 0.00553024891623),
        ]
    elif npts == 116:
        orbits = [
            (_orbit_a2(), 0.00200918797730),
            (_orbit_a3(), 0.00988550016044),
            (_orbit_b(0.162263300152, 0.973314565209), 0.00844068048232),
            (_orbit_b(0.383386152638, 0.840255982384), 0.00987390742389),
            (_orbit_b(0.686647945709, 0.238807866929), 0.0093573216900),
            (_orbit_c(0.878158910604, 0.478369028812), 0.00969499636166),
        ]
    elif npts == 50:
        orbits = [
            (_orbit_a1(), 0.0126984126985),
            (_orbit_a2(), 0.0225749559083),
            (_orbit_a3(), 0.02109375),
            (_orbit_b(0.301511344578, 0.904534033733), 0.0201733355379),
        ]
    else:
        raise ValueError("Lebedev rule with %d points is not tabulated" % npts)
    e, w = [], []
    for pts, wt in orbits:
        e.extend(pts)
        w.extend([wt] * len(pts))
    e = np.asarray(e, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64) * (4.0 * math.pi)
    assert e.shape == (npts, 3)
    return e, w


# ---------------------------------------------------------------------------
# Basis
# ---------------------------------------------------------------------------

_BASIS_CACHE: dict = {}


def basis_table(name: str = "6-31G**") -> dict:
    if name not in _BASIS_CACHE:
        fn = {"6-31G**": "basis_6-31Gss.json"}[name]
        with open(os.path.join(_DATA, fn)) as f:
            _BASIS_CACHE[name] = json.load(f)["elements"]
    return _BASIS_CACHE[name]


# Cartesian exponents per component, in g2g's component order.
CART = {0: [(0, 0, 0)],
        1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
        2: [(2, 0, 0), (1, 1, 0), (0, 2, 0), (1, 0, 1), (0, 1, 1), (0, 0, 2)]}


def _prim_norm(a: float, l: int) -> float:
    return math.sqrt(math.sqrt(8.0 * a ** 3) * (4.0 * a) ** l / PI32)


@dataclass
class Basis:
    M: int
    nuc: np.ndarray        # (M,) uint32, 1-based atom index
    ncont: np.ndarray      # (M,) uint32
    a: np.ndarray          # (M, 13) float64 exponents
    c: np.ndarray          # (M, 13) float64 normalised coefficients
    l: np.ndarray          # (M,) angular momentum
    lxyz: np.ndarray       # (M, 3) Cartesian powers
    nshell: np.ndarray     # (3,) uint32: n_s, n_p functions, n_d functions


def build_basis(Z: np.ndarray, name: str = "6-31G**", normalize: bool = True) -> Basis:
    table = basis_table(name)
    funcs = []   # (l, comp, atom(0-based), exps, coefs)
    for zkey in sorted(table, key=int):
        z = int(zkey)
        for iat in range(len(Z)):
            if int(Z[iat]) != z:
                continue
            for sh in table[zkey]:
                l = sh["l"]
                exps = sh["exp"]
                coefs = [cc * (_prim_norm(aa, l) if normalize else 1.0)
                         for aa, cc in zip(exps, sh["coef"])]
                for comp in range(len(CART[l])):
                    funcs.append((l, comp, iat, exps, coefs))
    missing = set(int(z) for z in Z) - set(int(k) for k in table)
    if missing:
        raise ValueError("no %s basis for Z=%s" % (name, sorted(missing)))
    funcs.sort(key=lambda f: f[0])  # stable s -> p -> d
    M = len(funcs)
    b = Basis(M=M, nuc=np.zeros(M, np.uint32), ncont=np.zeros(M, np.uint32),
              a=np.zeros((M, MAX_CONTRACTIONS)), c=np.zeros((M, MAX_CONTRACTIONS)),
              l=np.zeros(M, np.int32), lxyz=np.zeros((M, 3), np.int32),
              nshell=np.zeros(3, np.uint32))
    for i, (l, comp, iat, exps, coefs) in enumerate(funcs):
        b.nuc[i] = iat + 1
        b.ncont[i] = len(exps)
        b.a[i, :len(exps)] = exps
        b.c[i, :len(exps)] = coefs
        b.l[i] = l
        b.lxyz[i] = CART[l][comp]
        b.nshell[l] += 1
    return b


def overlap(pos: np.ndarray, b: Basis, normalize: bool = True) -> np.ndarray:
    """Analytic overlap S_ij of g2g's basis functions (Obara-Saika, l <= 2).
    d functions xx, yy, zz carry the 1/sqrt(3) factor g2g applies when NORM."""
    # expand to primitives
    pa, pc, pctr, pl = [], [], [], []
    owner = []
    for i in range(b.M):
        lx, ly, lz = b.lxyz[i]
        nf = 1.0
        if normalize and b.l[i] == 2 and max(lx, ly, lz) == 2:
            nf = 1.0 / math.sqrt(3.0)
        for k in range(int(b.ncont[i])):
            pa.append(b.a[i, k])
            pc.append(b.c[i, k] * nf)
            pctr.append(pos[b.nuc[i] - 1])
            pl.append((lx, ly, lz))
            owner.append(i)
    pa = np.array(pa)
    pc = np.array(pc)
    pctr = np.array(pctr)
    pl = np.array(pl)
    owner = np.array(owner)
    A = pa[:, None]
    B = pa[None, :]
    p = A + B
    mu = A * B / p
    Sxyz = np.ones((len(pa), len(pa)))
    for d in range(3):
        Ax = pctr[:, d][:, None]
        Bx = pctr[:, d][None, :]
        Px = (A * Ax + B * Bx) / p
        XPA = Px - Ax
        XPB = Px - Bx
        s = {}
        s[(0, 0)] = np.sqrt(math.pi / p) * np.exp(-mu * (Ax - Bx) ** 2)
        inv2p = 0.5 / p
        for i in range(3):
            for j in range(3):
                if (i, j) == (0, 0):
                    continue
                if i > 0:
                    v = XPA * s[(i - 1, j)]
                    if i > 1:
                        v = v + (i - 1) * inv2p * s[(i - 2, j)]
                    if j > 0:
                        v = v + j * inv2p * s[(i - 1, j - 1)]
                else:
                    v = XPB * s[(i, j - 1)]
                    if j > 1:
                        v = v + (j - 1) * inv2p * s[(i, j - 2)]
                s[(i, j)] = v
        li = pl[:, d]
        pick = 3 * li[:, None] + li[None, :]
        Sxyz *= np.choose(pick, [s[(i, j)] for i in range(3) for j in range(3)])
    Sp = Sxyz * pc[:, None] * pc[None, :]
    # contract primitives to functions: each function's primitives are
    # consecutive, so both sums are segment sums
    starts = np.flatnonzero(np.r_[True, owner[1:] != owner[:-1]])
    S = np.add.reduceat(np.add.reduceat(Sp, starts, axis=0), starts, axis=1)
    return 0.5 * (S + S.T)


# ---------------------------------------------------------------------------
# Density
# ---------------------------------------------------------------------------

def lowdin_orbitals(S: np.ndarray, nocc: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((S.shape[0], nocc))
    O = C.T @ S @ C
    ev, U = np.linalg.eigh(O)
    C = C @ (U @ np.diag(ev ** -0.5) @ U.T)
    return C


def pack_density(P: np.ndarray) -> np.ndarray:
    """Column-major lower packing with off-diagonals doubled (g2g's rho_vec)."""
    M = P.shape[0]
    out = np.empty(M * (M + 1) // 2)
    k = 0
    for i in range(M):
        col = P[i:, i].copy()
        col[1:] *= 2.0
        out[k:k + M - i] = col
        k += M - i
    return out


def packed_index(i: int, j: int, M: int) -> int:
    """g2g's packed index for 0-based i <= j (``partition.cpp:132-134``)."""
    if i > j:
        i, j = j, i
    return i * M - (i * (i - 1)) // 2 + (j - i)


def unpack_fock(Fp: np.ndarray, M: int) -> np.ndarray:
    F = np.zeros((M, M))
    k = 0
    for i in range(M):
        F[i:, i] = Fp[k:k + M - i]
        F[i, i:] = Fp[k:k + M - i]
        k += M - i
    return F


# ---------------------------------------------------------------------------
# Geometries
# ---------------------------------------------------------------------------

WATER_OH_A = 0.9572
WATER_ANGLE_DEG = 104.52


def water_geometry_ang() -> tuple[np.ndarray, np.ndarray]:
    """O at origin, H atoms in the xy plane symmetric about +y."""
    half = math.radians(WATER_ANGLE_DEG) / 2.0
    d = WATER_OH_A
    pos = np.array([[0.0, 0.0, 0.0],
                    [d * math.sin(half), d * math.cos(half), 0.0],
                    [-d * math.sin(half), d * math.cos(half), 0.0]])
    return np.array([8, 1, 1]), pos


def _random_rotation(rng) -> np.ndarray:
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_cluster_ang(n: int = 32, box: float = 10.0, min_oo: float = 2.6, seed: int = 1):
    rng = np.random.default_rng(seed)
    Zw, pw = water_geometry_ang()
    centres = []
    tries = 0
    while len(centres) < n:
        tries += 1
        if tries > 200000:
            raise RuntimeError("could not place %d waters" % n)
        c = rng.uniform(0.0, box, 3)
        if all(np.linalg.norm(c - o) >= min_oo for o in centres):
            centres.append(c)
    Z, pos = [], []
    for c in centres:
        R = _random_rotation(rng)
        for zz, p in zip(Zw, pw):
            Z.append(zz)
            pos.append(c + R @ p)
    return np.array(Z), np.array(pos)


def random_molecule_ang(natoms: int, seed: int, frac=(0.35, 0.45, 0.10, 0.10),
                        bond=(1.0, 1.6), min_sep=0.9):
    """Seeded random molecule: elements C,H,N,O drawn with the given fractions,
    grown by attaching each atom to a random earlier one at a bonded distance,
    rejecting placements closer than ``min_sep`` to any atom.  An odd electron
    count is made even by swapping the last atom's element H<->C or N<->O,
    which changes Z by an odd amount and keeps the atom count."""
    rng = np.random.default_rng(seed)
    elems = np.array([6, 1, 7, 8])
    Z = list(rng.choice(elems, size=natoms, p=np.asarray(frac) / sum(frac)))
    if sum(Z) % 2 == 1:
        # flip the parity with one substitution that keeps the atom count:
        # H(1) <-> C(6) or N(7) <-> O(8) changes Z by an odd amount
        swap = {1: 6, 6: 1, 7: 8, 8: 7}
        Z[-1] = swap[int(Z[-1])]
    pos = [np.zeros(3)]
    for k in range(1, natoms):
        for _ in range(10000):
            parent = pos[rng.integers(0, len(pos))]
            v = rng.standard_normal(3)
            v /= np.linalg.norm(v)
            cand = parent + v * rng.uniform(*bond)
            if all(np.linalg.norm(cand - q) >= min_sep for q in pos):
                pos.append(cand)
                break
        else:
            raise RuntimeError("could not grow molecule")
    return np.array(Z, dtype=np.int64), np.array(pos)


# ---------------------------------------------------------------------------
# A full caller-side system
# ---------------------------------------------------------------------------

@dataclass
class System:
    name: str
    Z: np.ndarray                 # (natom,)
    pos: np.ndarray               # (natom, 3) bohr
    basis: Basis
    nocc: int
    rho_packed: np.ndarray        # (M(M+1)/2,)
    P: np.ndarray                 # (M, M) full symmetric density
    S: np.ndarray                 # (M, M) overlap
    grid_type: int = 2
    shells: int = 40
    iexch: int = 9
    normalize: bool = True
    meta: dict = field(default_factory=dict)
    # open shell: the alpha and beta densities (P = Pa + Pb, rho_packed its packing)
    open: bool = False
    Pa: np.ndarray | None = None
    Pb: np.ndarray | None = None
    rho_a_packed: np.ndarray | None = None
    rho_b_packed: np.ndarray | None = None

    @property
    def natom(self) -> int:
        return len(self.Z)

    @property
    def M(self) -> int:
        return self.basis.M

    def abi_arrays(self) -> dict:
        """Arrays in exactly the layouts ``g2g_parameter_init_`` receives."""
        b = self.basis
        natom = self.natom
        r = np.asfortranarray(self.pos.astype(np.float64))           # natom x 3, ld=natom
        e, w = {}, {}
        for n in GRID_SIZES:
            ee, ww = lebedev(n)
            e[n] = np.asfortranarray(ee)
            w[n] = ww
        return dict(
            natom=natom, max_atoms=natom, M=b.M, ngaussians=b.M,
            r=r.ravel(order="F"), Rm=slater_rm(), Iz=self.Z.astype(np.uint32),
            Nr=shell_counts(self.shells), Nr2=shell_counts(self.shells),
            Nuc=b.nuc.copy(), ncont=b.ncont.copy(), nshell=b.nshell.copy(),
            a=np.asfortranarray(b.a).ravel(order="F"),
            c=np.asfortranarray(b.c).ravel(order="F"),
            rho=self.rho_packed.copy(), nco=self.nocc, iexch=self.iexch,
            norm=int(self.normalize), open=bool(self.open),
            rho_a=None if not self.open else self.rho_a_packed.copy(),
            rho_b=None if not self.open else self.rho_b_packed.copy(),
            e1=e[50].ravel(order="F"), e2=e[116].ravel(order="F"), e3=e[194].ravel(order="F"),
            w1=w[50], w2=w[116], w3=w[194], grid_type=self.grid_type)


def make_system(name: str, Z, pos_ang, nocc: int | None = None, seed: int = 0,
                shells: int = 40, iexch: int = 9) -> System:
    Z = np.asarray(Z, dtype=np.int64)
    pos = np.asarray(pos_ang, dtype=np.float64) / ANG_PER_BOHR
    b = build_basis(Z)
    if nocc is None:
        nel = int(Z.sum())
        assert nel % 2 == 0, "closed shell needs an even electron count"
        nocc = nel // 2
    S = overlap(pos, b)
    C = lowdin_orbitals(S, nocc, seed)
    P = 2.0 * C @ C.T
    return System(name=name, Z=Z, pos=pos, basis=b, nocc=nocc,
                  rho_packed=pack_density(P), P=P, S=S, shells=shells, iexch=iexch,
                  meta={"seed": seed})


def open_shell(base: System, nunp: int = 2, seed: int | None = None) -> System:
    """The same molecule with nunp more alpha than beta electrons: seeded
    Loewdin-orthonormal orbitals, the first nocc + nunp/2 occupied in alpha and
    the first nocc - nunp/2 in beta."""
    if nunp % 2:
        raise ValueError("nunp must keep the electron count")
    seed = base.meta.get("seed", 0) + 7 if seed is None else seed
    na, nb = base.nocc + nunp // 2, base.nocc - nunp // 2
    C = lowdin_orbitals(base.S, na, seed)
    Pa = C @ C.T
    Pb = C[:, :nb] @ C[:, :nb].T
    P = Pa + Pb
    return dataclasses.replace(base, name=base.name + "_open%d" % nunp, P=P,
                               rho_packed=pack_density(P), open=True, Pa=Pa, Pb=Pb,
                               rho_a_packed=pack_density(Pa), rho_b_packed=pack_density(Pb),
                               meta=dict(base.meta, open_seed=seed, nunp=nunp))


def config0(seed: int = 0) -> System:
    """One water, 6-31G** (M=25), 5 seeded orbitals."""
    Z, p = water_geometry_ang()
    return make_system("h2o", Z, p, nocc=5, seed=seed)


def config1(seed: int = 1) -> System:
    """(H2O)32 in a 10 A cube, O-O >= 2.6 A, M=800, 160 seeded orbitals."""
    Z, p = water_cluster_ang(32, seed=seed)
    return make_system("h2o_32", Z, p, nocc=160, seed=seed)


def water_cluster(n: int, seed: int) -> System:
    box = 10.0 * (n / 32.0) ** (1.0 / 3.0)
    Z, p = water_cluster_ang(n, box=box, seed=seed)
    return make_system("h2o_%d" % n, Z, p, nocc=5 * n, seed=seed)


def random_molecule(natoms: int, seed: int) -> System:
    Z, p = random_molecule_ang(natoms, seed)
    return make_system("mol%d_s%d" % (natoms, seed), Z, p, seed=seed)


def molecule_batch(n: int = 64, seed: int = 3, atoms=(3, 60)) -> list[System]:
    """n independent random molecules (C:H:N:O ~ 0.35:0.45:0.10:0.10, bonded
    spacing 1.0-1.6 A, minimum separation 0.9 A, even electron count, 6-31G**,
    40 x 194 grid, PBE), each with its own seeded density: the atom counts are
    drawn uniformly from ``atoms`` by a generator seeded with ``seed`` and
    molecule k is grown with seed 1000 * seed + k."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(atoms[0], atoms[1] + 1, size=n)
    # threads: the overlap and density set-up is numpy array work, which runs
    # outside the interpreter lock; each molecule depends on its seed only
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        return list(ex.map(lambda k: random_molecule(int(sizes[k]), 1000 * seed + k), range(n)))
