"""ctypes binding of liboracle.so, the FP64 CPU restatement (test infrastructure only).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker
of the HIP path; never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrcParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("T", C.c_double), ("visc", C.c_double), ("vfactor", C.c_double),
        ("alpha_lj", C.c_double), ("cutoff_lj", C.c_double), ("lj_6_term", C.c_int),
        ("elec", C.c_int), ("epsilon_c", C.c_double), ("kappa_c", C.c_double),
        ("gamma_c", C.c_double), ("cutoff_c", C.c_double), ("bond_lj_scale", C.c_double),
        ("bond_c_scale", C.c_double), ("check_overlap", C.c_int), ("bc", C.c_int),
        ("box", C.c_double), ("hydro", C.c_int), ("ewald", C.c_int), ("ewald_real", C.c_int),
        ("ewald_method", C.c_int), ("ewald_alpha", C.c_double),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_prefactors.argtypes = [C.POINTER(OrcParams), _dp, _dp]
        L.orc_rpy_block.argtypes = [_dp, _dp, C.c_double, C.c_double, _dp]
        L.orc_beenakker_block.argtypes = [_dp, C.c_int, C.c_int, C.c_double, C.c_int,
                                          C.c_double, C.c_double, _dp]
        L.orc_compute_D.argtypes = [C.c_int, _dp, C.POINTER(OrcParams), _dp, C.c_int]
        L.orc_cholesky.argtypes = [C.c_int, _dp, C.c_int]
        L.orc_cholesky.restype = C.c_int
        L.orc_trmv_lower.argtypes = [C.c_int, _dp, C.c_int, _dp, _dp]
        L.orc_symv_lower.argtypes = [C.c_int, _dp, C.c_int, _dp, _dp]
        L.orc_forces.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, _ip, _dp,
                                 C.POINTER(OrcParams), _dp, _dp]
        L.orc_forces.restype = C.c_int
        L.orc_srand48.argtypes = [C.c_uint32]
        L.orc_srand48.restype = C.c_uint64
        L.orc_drand48.argtypes = [C.POINTER(C.c_uint64)]
        L.orc_drand48.restype = C.c_double
        L.orc_gauss.argtypes = [C.POINTER(C.c_uint64), _dp, C.c_int]
        L.orc_geyer_rs.argtypes = [C.c_int, _dp, C.POINTER(OrcParams), C.c_double, _dp, _dp]
        L.orc_geyer_rs_D.argtypes = [C.c_int, _dp, C.c_int, _dp, C.POINTER(OrcParams), C.c_double,
                                     _dp, _dp]
        L.orc_rf_onfly.argtypes = [C.c_int, _dp, C.POINTER(OrcParams), C.c_double, _dp, _dp]
        L.orc_sim_create.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, _ip, _dp,
                                     C.POINTER(OrcParams), C.c_uint32, C.c_int]
        L.orc_sim_create.restype = C.c_void_p
        L.orc_sim_run.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_sim_run.restype = C.c_int
        L.orc_sim_coords.argtypes = [C.c_void_p, _dp]
        L.orc_sim_forces.argtypes = [C.c_void_p, _dp]
        L.orc_sim_last.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_sim_rng_state.argtypes = [C.c_void_p]
        L.orc_sim_rng_state.restype = C.c_uint64
        L.orc_sim_destroy.argtypes = [C.c_void_p]
        L.orc_sim_set_algorithm.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_sim_set_flow.argtypes = [C.c_void_p, _dp]
        L.orc_sim_last_dt.argtypes = [C.c_void_p]
        L.orc_sim_last_dt.restype = C.c_double
        L.orc_sim_rejections.argtypes = [C.c_void_p]
        L.orc_sim_rejections.restype = C.c_longlong
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def params_from(P) -> OrcParams:
    """OrcParams from a bd_box_amd.workloads.Params (already float-rounded)."""
    return OrcParams(
        dt=P.dt, T=P.T, visc=P.visc, vfactor=P.vfactor, alpha_lj=P.alpha_lj,
        cutoff_lj=P.cutoff_lj, lj_6_term=P.lj_6_term, elec=P.elec, epsilon_c=P.epsilon_c,
        kappa_c=P.kappa_c, gamma_c=P.gamma_c, cutoff_c=P.cutoff_c,
        bond_lj_scale=P.bond_lj_scale, bond_c_scale=P.bond_c_scale,
        check_overlap=P.check_overlap, bc=P.bc, box=P.box, hydro=P.hydro,
        ewald=int(P.bc == 1 and P.hydro != 0),
        ewald_real=P.ewald_real, ewald_method=P.ewald_method, ewald_alpha=P.ewald_alpha)


def prefactors(p: OrcParams):
    a, b = C.c_double(), C.c_double()
    lib().orc_prefactors(C.byref(p), C.byref(a), C.byref(b))
    return a.value, b.value


def rpy_block(ci, cj, preDii, preDij):
    out = np.zeros(9)
    lib().orc_rpy_block(_d(np.ascontiguousarray(ci, float)), _d(np.ascontiguousarray(cj, float)),
                        preDii, preDij, _d(out))
    return out.reshape(3, 3)


def compute_D(coord, p: OrcParams, out=None):
    """Full symmetric D (n x n).  orc_compute_D writes every entry, both triangles, so a
    C-contiguous `out` of that shape is filled in place (no zeroing, no transpose)."""
    coord = np.ascontiguousarray(coord, float).reshape(-1)
    N = coord.size // 4
    n = 3 * N
    if out is not None:
        assert out.shape == (n, n) and out.dtype == np.float64 and out.flags.c_contiguous
        lib().orc_compute_D(N, _d(coord), C.byref(p), _d(out), n)
        return out
    D = np.zeros((n, n))  # column-major view below: D_cm[r + c*n] -> D.T row-major
    lib().orc_compute_D(N, _d(coord), C.byref(p), _d(D), n)
    return D.T.copy()


def cholesky(D):
    n = D.shape[0]
    A = np.asfortranarray(D, dtype=float).copy(order="F")
    info = lib().orc_cholesky(n, A.ctypes.data_as(_dp), n)
    if info:
        raise np.linalg.LinAlgError(f"not positive definite at {info}")
    return np.tril(np.asarray(A))


def forces(sysm, p: OrcParams, coord=None):
    coord = np.ascontiguousarray(sysm.coord if coord is None else coord, float).reshape(-1)
    N = coord.size // 4
    F = np.zeros(3 * N)
    E = np.zeros(3)
    bonds = np.ascontiguousarray(sysm.bonds, np.int32).reshape(-1)
    parms = np.ascontiguousarray(sysm.bond_parms, float).reshape(-1)
    rep = lib().orc_forces(N, _d(coord), _d(np.ascontiguousarray(sysm.Q, float)),
                           _d(np.ascontiguousarray(sysm.LJ, float).reshape(-1)),
                           len(sysm.bonds), bonds.ctypes.data_as(_ip), _d(parms), C.byref(p),
                           _d(F), _d(E))
    return F, E, rep


def gauss(seed: int, count: int, calls: int = 1):
    st = C.c_uint64(lib().orc_srand48(seed))
    out = []
    for _ in range(calls):
        g = np.zeros(count)
        lib().orc_gauss(C.byref(st), _d(g), count)
        out.append(g)
    return np.array(out), st.value


def geyer_rs(coord, p: OrcParams, dt, X):
    coord = np.ascontiguousarray(coord, float).reshape(-1)
    N = coord.size // 4
    X = np.array(X, float)
    Rs = np.zeros(3 * N)
    lib().orc_geyer_rs(N, _d(coord), C.byref(p), dt, _d(X), _d(Rs))
    return Rs, X


def rf_onfly(coord, p: OrcParams, dt, F):
    coord = np.ascontiguousarray(coord, float).reshape(-1)
    N = coord.size // 4
    Rf = np.zeros(3 * N)
    lib().orc_rf_onfly(N, _d(coord), C.byref(p), dt, _d(np.ascontiguousarray(F, float)), _d(Rf))
    return Rf


class Sim:
    """Stateful FP64 Ermak-McCammon integrator (orc_sim_*)."""

    def __init__(self, sysm, p: OrcParams, seed: int, refactor_every: int = 1, coord=None,
                 algorithm: int = 0, move_attempts: int = 10000, flow=None):
        self.N = sysm.N
        c = np.ascontiguousarray(sysm.coord if coord is None else coord, float).reshape(-1)
        self._keep = [c, np.ascontiguousarray(sysm.Q, float),
                      np.ascontiguousarray(sysm.LJ, float).reshape(-1),
                      np.ascontiguousarray(sysm.bonds, np.int32).reshape(-1),
                      np.ascontiguousarray(sysm.bond_parms, float).reshape(-1)]
        self.p = p
        self.h = lib().orc_sim_create(self.N, _d(self._keep[0]), _d(self._keep[1]),
                                      _d(self._keep[2]), len(sysm.bonds),
                                      self._keep[3].ctypes.data_as(_ip), _d(self._keep[4]),
                                      C.byref(self.p), seed, refactor_every)
        lib().orc_sim_set_algorithm(self.h, algorithm, move_attempts)
        if flow is not None:
            # velocity gradient tensor (row, column), vel_grad_tensor (diff_decl.c:72)
            self._G = np.ascontiguousarray(flow, float).reshape(9)
            lib().orc_sim_set_flow(self.h, _d(self._G))

    def run(self, steps: int, xi=None):
        x = None if xi is None else np.ascontiguousarray(xi, float).reshape(-1)
        rc = lib().orc_sim_run(self.h, steps, None if x is None else _d(x))
        if rc:
            raise RuntimeError(f"oracle step failed ({rc})")

    def coords(self):
        c = np.zeros(4 * self.N)
        lib().orc_sim_coords(self.h, _d(c))
        return c.reshape(-1, 4)

    def forces(self):
        F = np.zeros(3 * self.N)
        lib().orc_sim_forces(self.h, _d(F))
        return F

    def last(self):
        Rf, Rs = np.zeros(3 * self.N), np.zeros(3 * self.N)
        lib().orc_sim_last(self.h, _d(Rf), _d(Rs))
        return Rf, Rs

    def rng_state(self):
        return lib().orc_sim_rng_state(self.h)

    def last_dt(self):
        return lib().orc_sim_last_dt(self.h)

    def rejections(self):
        return lib().orc_sim_rejections(self.h)

    def __del__(self):
        try:
            lib().orc_sim_destroy(self.h)
        except Exception:
            pass


def pairs_within(xyz, cutoff: float, pbc: bool = False, box: float = 0.0) -> np.ndarray:
    """All pairs i < j with |r_j - r_i| <= cutoff, sorted, shape (P, 2), int32.

    The brute-force pair loop of calc_func.c:188-228 with the separation of
    trans_dist_vec (trans.c:111-121): d = r_j - r_i, then d -= L floor(d / L + 1/2) per
    axis when periodic; r = sqrt(dx dx + dy dy + dz dz) in FP64, summed left to right.
    """
    xyz = np.ascontiguousarray(np.asarray(xyz, float)[:, :3])
    N = len(xyz)
    out = []
    inv = 1.0 / box if pbc else 0.0
    for i in range(N - 1):
        d = xyz[i + 1:] - xyz[i]
        if pbc:
            d = d - box * np.floor(d * inv + 0.5)
        r = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
        j = np.nonzero(r <= cutoff)[0] + i + 1
        out.append(np.stack([np.full(len(j), i), j], axis=1))
    if not out:
        return np.zeros((0, 2), np.int32)
    return np.concatenate(out).astype(np.int32)
