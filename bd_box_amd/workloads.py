"""Seeded synthetic BD_BOX inputs and the `.prm` / `.str` file formats.

The reference reads a control file (`keyword value` lines, `#` comments;
src/input.c:273-343) and a structure file with `sub`/`bond` lines
(src/input.c:505-580).  This module writes both formats for the benchmark systems
and reads them back the way the reference does, including its one numerical quirk:
every DOUBLE keyword of the `.prm` file is parsed as `(FLOAT)atof(value)`
(src/input.c:933), i.e. rounded to float32 before it is stored as a double, while
numbers in the `.str` file keep full double precision (src/input.c:1152).

Systems
-------
`chain(N, seed)`  -- the bead-spring chain used by the N=256 / N=3000 / batched / adjoint
    workloads: a self-avoiding random walk with 3.8 A bonds and >= 4 A between
    non-bonded beads, hydrodynamic radius 1.9 A, LJ diameter 4 A (eps 0.6 kcal/mol),
    20 % of the beads charged +-1, harmonic bonds k = 10 kcal/mol/A^2 (FENE form with
    r_max = 1e10, which is harmonic to double precision; src/potentials/bond.c:43-85).
`spheres(N, seed, phi)` -- hard spheres of radius 25 A placed by random sequential
    addition in a cubic periodic box at volume fraction phi.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Dict, List, Optional, Tuple

import numpy as np

# Keywords whose registry type is DOUBLE in the reference (parsed through float32).
_DOUBLE_KEYS = {
    "dt", "visc", "vfactor", "T", "epsilon_c", "kappa_c", "cutoff_c", "gamma_c", "alpha_lj",
    "cutoff_lj", "xbox", "ybox", "zbox", "E_magn", "E_freq", "E_factor", "sphere_radius",
    "bond_lj_scale", "bond_c_scale", "ewald_alpha", "e_collision", "sboundary_A",
    "sboundary_cutoff",
}
_YESNO_KEYS = {"elec", "lj_6_term", "E_ext", "sboundary", "check_overlap", "geyer_on_the_fly"}
_ENUMS = {
    "hydro": {"none": 0, "cholesky": 1, "geyer": 2, "chebyshev": 3},
    "algorithm": {"ermak_const": 0, "ermak_var": 1, "ermak_newton": 2, "igt_const": 3,
                  "igt_var": 4, "igt_var_rev": 5},
    "bc": {"none": 0, "periodic": 1, "sphere": 2},
    "ewald_method": {"smith": 0, "beenakker": 1},
    "nb_list": {"brute": 0, "bucket_space": 1, "bucket_list": 2},
}


def f32(x: float) -> float:
    """`(FLOAT)atof(x)` as applied to every DOUBLE keyword (src/input.c:933)."""
    return float(np.float32(x))


@dataclasses.dataclass
class System:
    """Bead table in the reference's layout (src/data.c:85-109).

    coord : (N, 4) float64, columns x, y, z, hydrodynamic radius a
    Q     : (N,)  charges
    LJ    : (N, 2) (R, eps) with R = (LJ diameter column of the .str) / 2
    bonds : (B, 2) int32 bead indices; bond_parms (B, 3) = (r0, r_max, H)
    """
    coord: np.ndarray
    Q: np.ndarray
    LJ: np.ndarray
    bonds: np.ndarray
    bond_parms: np.ndarray
    names: Optional[List[str]] = None

    @property
    def N(self) -> int:
        return int(self.coord.shape[0])


def _unit_vectors(rng: np.random.Generator, k: int) -> np.ndarray:
    v = rng.normal(size=(k, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def saw_chain(N: int, seed: int, bond: float = 3.8, dmin: float = 4.0) -> np.ndarray:
    """Self-avoiding random walk: each new bead sits `bond` from the previous one and at
    least `dmin` from every earlier bead (trial directions uniform on the sphere).  A bead
    with no admissible direction among 64 x 8 trials makes the walk back up 8 beads.
    Earlier beads are found through a hash grid of cell edge bond + dmin, so only the 27
    cells around the previous bead are searched."""
    rng = np.random.default_rng(seed)
    X = np.zeros((N, 3))
    d2 = dmin * dmin
    h = bond + dmin
    grid: Dict[Tuple[int, int, int], List[int]] = {}

    def key(p):
        return (int(math.floor(p[0] / h)), int(math.floor(p[1] / h)), int(math.floor(p[2] / h)))

    def near(p, upto):
        cx, cy, cz = key(p)
        out: List[int] = []
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    out.extend(j for j in grid.get((cx + dx, cy + dy, cz + dz), ()) if j < upto)
        return out

    grid.setdefault(key(X[0]), []).append(0)
    i = 1
    backtracks = 0
    while i < N:
        placed = False
        for _ in range(8 if i >= 2 else 1):
            trials = X[i - 1] + bond * _unit_vectors(rng, 64)
            if i < 2:
                X[i] = trials[0]
                placed = True
                break
            cand = near(X[i - 1], i - 1)   # earlier non-bonded beads near the previous one
            if not cand:
                X[i] = trials[0]
                placed = True
                break
            prev = X[cand]
            dd = ((trials[:, None, :] - prev[None, :, :]) ** 2).sum(-1).min(axis=1)
            idx = np.flatnonzero(dd >= d2)
            if idx.size:
                X[i] = trials[idx[0]]
                placed = True
                break
        if placed:
            grid.setdefault(key(X[i]), []).append(i)
            i += 1
        else:
            backtracks += 1
            if backtracks > 10 * N:
                raise RuntimeError(f"self-avoiding walk stuck at bead {i}")
            back = max(1, i - 8)
            for j in range(back, i):
                grid[key(X[j])].remove(j)
            i = back
    return X - X.mean(axis=0)


def chain(N: int, seed: int, a: float = 1.9, lj_diam: float = 4.0, lj_eps: float = 0.6,
          charge_frac: float = 0.2, bond: float = 3.8, k_bond: float = 10.0) -> System:
    X = saw_chain(N, seed, bond=bond, dmin=lj_diam)
    rng = np.random.default_rng(seed + 7919)
    Q = np.zeros(N)
    nq = int(round(charge_frac * N))
    pick = rng.choice(N, size=nq, replace=False)
    Q[pick] = rng.choice([-1.0, 1.0], size=nq)
    coord = np.concatenate([X, np.full((N, 1), a)], axis=1)
    LJ = np.tile([lj_diam / 2.0, lj_eps], (N, 1))
    bonds = np.stack([np.arange(N - 1), np.arange(1, N)], axis=1).astype(np.int32)
    parms = np.tile([bond, 1e10, k_bond], (N - 1, 1))
    return System(coord, Q, LJ, bonds, parms)


def batch_sizes(count: int, seed: int, lo: int = 64, hi: int = 1024) -> List[int]:
    """Chain lengths drawn log-uniformly in [lo, hi] (256 chains average about 346 beads)."""
    rng = np.random.default_rng(seed)
    return [int(round(math.exp(u))) for u in rng.uniform(math.log(lo), math.log(hi), count)]


def chains(sizes, seed: int) -> List[System]:
    """One self-avoiding chain per size, chain k from seed + k."""
    return [chain(int(N), seed=seed + k) for k, N in enumerate(sizes)]


def spheres(N: int, seed: int, radius: float = 25.0, phi: float = 0.2,
            lj_diam: float = 50.0, lj_eps: float = 1.0) -> Tuple[System, float]:
    """Random sequential addition of N non-overlapping spheres in a cubic periodic box of
    edge L = (N (4/3) pi a^3 / phi)^(1/3); returns the system and L."""
    L = (N * 4.0 / 3.0 * math.pi * radius ** 3 / phi) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    dmin2 = (2 * radius) ** 2
    # cell grid for the overlap test during placement
    nc = max(1, int(L // (2 * radius)))
    cell = L / nc
    grid: Dict[Tuple[int, int, int], List[int]] = {}
    P = np.zeros((N, 3))
    n = 0
    attempts = 0
    while n < N:
        attempts += 1
        if attempts > 200 * N:
            raise RuntimeError("random sequential addition did not converge")
        p = rng.uniform(-L / 2, L / 2, size=3)
        c = tuple(((p + L / 2) // cell).astype(int) % nc)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in grid.get(((c[0] + dx) % nc, (c[1] + dy) % nc, (c[2] + dz) % nc), ()):
                        v = P[j] - p
                        v -= L * np.floor(v / L + 0.5)
                        if v @ v < dmin2:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            P[n] = p
            grid.setdefault(c, []).append(n)
            n += 1
    coord = np.concatenate([P, np.full((N, 1), radius)], axis=1)
    LJ = np.tile([lj_diam / 2.0, lj_eps], (N, 1))
    sysm = System(coord, np.zeros(N), LJ, np.zeros((0, 2), np.int32), np.zeros((0, 3)))
    return sysm, L


# ------------------------------------------------------------------------ file formats
def write_str(path: str, s: System) -> None:
    """`sub name id x y z a Q 2R eps mass` (src/input.c:567-580), `bond id1 id2 r0 rmax H`."""
    with open(path, "w") as f:
        for i in range(s.N):
            x, y, z, a = s.coord[i]
            f.write("sub B %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g 1\n"
                    % (i + 1, x, y, z, a, s.Q[i], 2 * s.LJ[i, 0], s.LJ[i, 1]))
        for b in range(len(s.bonds)):
            i, j = s.bonds[b]
            r0, rmax, H = s.bond_parms[b]
            f.write("bond %d %d %.17g %.17g %.17g\n" % (i + 1, j + 1, r0, rmax, H))


def read_str(path: str) -> System:
    ids, rows, Q, LJ, names = [], [], [], [], []
    bonds_raw = []
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t or t[0].startswith("#"):
                continue
            if t[0] == "sub":
                names.append(t[1])
                ids.append(int(t[2]))
                v = [float(x) for x in t[3:11]]
                rows.append(v[0:4])
                Q.append(v[4])
                LJ.append([v[5] / 2.0, v[6]])
            elif t[0] == "bond":
                bonds_raw.append((int(t[1]), int(t[2]), float(t[3]), float(t[4]), float(t[5])))
    pos = {d: k for k, d in enumerate(ids)}
    bonds = np.array([[pos[b[0]], pos[b[1]]] for b in bonds_raw], np.int32).reshape(-1, 2)
    parms = np.array([b[2:] for b in bonds_raw], float).reshape(-1, 3)
    return System(np.array(rows, float), np.array(Q, float), np.array(LJ, float), bonds,
                  parms, names)


def chain_params(**over) -> Dict[str, object]:
    """The chain recipe's control parameters (values as written to the .prm file)."""
    p: Dict[str, object] = dict(
        dt=0.01, T=293, visc=0.01, vfactor=14.4, alpha_lj=4.0, lj_6_term="yes",
        cutoff_lj=12.0, elec="yes", epsilon_c=78.5, kappa_c=0.1, gamma_c=331.842,
        cutoff_c=12.0, bond_lj_scale=0.0, bond_c_scale=0.0, check_overlap="no", bc="none",
        algorithm="ermak_const", hydro="cholesky", nb_list="brute", rand_seed=12345,
        move_attempts=1, bdsteps=2, save_xyz_freq=1000000, save_dcd_freq=1000000,
        save_rst_freq=1000000, save_enr_freq=1000000,
    )
    p.update(over)
    return p


def write_prm(path: str, params: Dict[str, object], str_path: str, workdir: str) -> None:
    """Write a control file; every output file name points into `workdir`."""
    base = os.path.join(workdir, "out")
    files = dict(str_filename=str_path, out_filename=base + ".log", xyz_filename=base + ".xyz",
                 dcd_filename=base + ".dcd", enr_filename=base + ".enr",
                 rst_filename=base + ".rst", pqr_filename=base + ".pqr")
    with open(path, "w") as f:
        for k, v in {**params, **files}.items():
            f.write(f"{k} {v}\n")


@dataclasses.dataclass
class Params:
    """Numerical control parameters after the reference's parsing rules."""
    dt: float
    T: float
    visc: float
    vfactor: float
    alpha_lj: float
    lj_6_term: int
    cutoff_lj: float
    elec: int
    epsilon_c: float
    kappa_c: float
    gamma_c: float
    cutoff_c: float
    bond_lj_scale: float
    bond_c_scale: float
    check_overlap: int
    bc: int
    hydro: int
    algorithm: int
    ewald_method: int
    ewald_real: int
    ewald_recip: int
    ewald_alpha: float
    box: float
    rand_seed: int
    nb_list: int
    geyer_on_the_fly: int
    move_attempts: int = 10000
    vel_grad_tensor: Optional[Tuple[float, ...]] = None   # G (row, column) of a flow field


_DEFAULTS = dict(  # code defaults of the reference registry (MAKE_STR_IN lines)
    dt=0.0, T=0.0, visc=0.01, vfactor=14.4, alpha_lj=4.0, lj_6_term="yes", cutoff_lj=0.0,
    elec="yes", epsilon_c=78.54, kappa_c=0.1, gamma_c=331.842, cutoff_c=0.0,
    bond_lj_scale=1.0, bond_c_scale=1.0, check_overlap="yes", bc="periodic",
    hydro="none", algorithm="ermak_newton", ewald_method="smith", ewald_real=10,
    ewald_recip=10, ewald_alpha=1.77245385090551602729, xbox=0.0, rand_seed=12345,
    nb_list="brute", geyer_on_the_fly="yes", move_attempts=10000,
)


def parse_params(params: Dict[str, object]) -> Params:
    """Apply the reference's keyword parsing to a parameter dict."""
    raw = {**_DEFAULTS, **params}

    def dbl(k):
        return f32(float(raw[k]))

    def yn(k):
        v = raw[k]
        return int(v) if isinstance(v, (int, np.integer)) else int(str(v).startswith("y"))

    def enum(k):
        v = raw[k]
        return int(v) if isinstance(v, (int, np.integer)) else _ENUMS[k][str(v)]

    hydro = enum("hydro")
    algorithm = enum("algorithm")
    g = yn("geyer_on_the_fly")
    if hydro != 2 or algorithm >= 3:  # src/input.c:497-501
        g = 0
    # vel_grad_tensor: nine doubles, row by row (input.c:1110-1125, scanned as %lg)
    G = raw.get("vel_grad_tensor")
    if G is not None and not isinstance(G, (tuple, list, np.ndarray)):
        G = str(G).split()
    if G is not None:
        G = tuple(float(x) for x in G)
        if len(G) != 9:
            raise ValueError("vel_grad_tensor: nine values")
    return Params(
        dt=dbl("dt"), T=dbl("T"), visc=dbl("visc"), vfactor=dbl("vfactor"),
        alpha_lj=dbl("alpha_lj"), lj_6_term=yn("lj_6_term"), cutoff_lj=dbl("cutoff_lj"),
        elec=yn("elec"), epsilon_c=dbl("epsilon_c"), kappa_c=dbl("kappa_c"),
        gamma_c=dbl("gamma_c"), cutoff_c=dbl("cutoff_c"), bond_lj_scale=dbl("bond_lj_scale"),
        bond_c_scale=dbl("bond_c_scale"), check_overlap=yn("check_overlap"), bc=enum("bc"),
        hydro=hydro, algorithm=algorithm, ewald_method=enum("ewald_method"),
        ewald_real=int(raw["ewald_real"]), ewald_recip=int(raw["ewald_recip"]),
        ewald_alpha=dbl("ewald_alpha"), box=dbl("xbox"), rand_seed=int(raw["rand_seed"]),
        nb_list=enum("nb_list"), geyer_on_the_fly=g, move_attempts=int(raw["move_attempts"]),
        vel_grad_tensor=G,
    )


def read_prm(path: str) -> Dict[str, str]:
    out: Dict[str, str] = {}
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if not line:
                continue
            k, _, v = line.partition(" ")
            out[k] = v.strip()
    return out
