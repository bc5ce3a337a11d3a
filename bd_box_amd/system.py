"""Host-side mirror of the reference's hot-path interface over the HIP C ABI.

`BDSystem` holds one simulated system on the GPU and exposes the reference's entry
points with the same meaning:

    compute_D()          compute_D / compute_single_D   (src/diff_alg/diff_tensor.c:472-561)
    compute_forces()     compute_forces                  (src/potentials/calc_func.c:106-186)
    compute_Rf()         compute_Rf                      (src/diff_alg/diff_equation.c:119-214)
    factor()             my_dpotrf on D                  (src/myblas.c:45-92)
    noise_product(x)     my_dtrmv                        (src/myblas.c:94-118)
    compute_hydro(k)     k x compute_hydro, ermak_const  (src/diff_alg/diff_comp.c:152-232)

Errors raise BackendError (the reference calls UNERR and exits, src/err.c:48-90).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from ._lib import BackendError  # noqa: F401  (raised by every call below)
from .workloads import Params, System

FP64 = 0
MIXED = 1
MIXED_SPLIT = 2


def _algorithm(P: Params) -> int:
    """ermak_const / ermak_var / igt_const / igt_var as given; ermak_newton takes the
    ermak_const path, which is the same step whenever no step is rejected
    (comp_step_ermak_newton, diff_comp.c:181-208), so it is accepted only without
    check_overlap."""
    if P.algorithm in (0, 1, 3, 4):
        return P.algorithm
    if P.algorithm == 2 and not P.check_overlap:
        return 0
    raise NotImplementedError("HIP backend: algorithm ermak_const / ermak_var / igt_const / "
                              "igt_var (ermak_newton only with check_overlap no)")


def make_config(P: Params, precision: int = MIXED, refactor_every: int = 1) -> _lib.Config:
    return _lib.Config(
        dt=P.dt, T=P.T, visc=P.visc, vfactor=P.vfactor, alpha_lj=P.alpha_lj,
        cutoff_lj=P.cutoff_lj, lj_6_term=P.lj_6_term, elec=P.elec, epsilon_c=P.epsilon_c,
        kappa_c=P.kappa_c, gamma_c=P.gamma_c, cutoff_c=P.cutoff_c,
        bond_lj_scale=P.bond_lj_scale, bond_c_scale=P.bond_c_scale,
        check_overlap=P.check_overlap, bc=P.bc, box=P.box, hydro=P.hydro,
        ewald_method=P.ewald_method, ewald_real=P.ewald_real, precision=precision,
        refactor_every=refactor_every, rand_seed=P.rand_seed & 0xFFFFFFFF, nb_list=P.nb_list,
        algorithm=_algorithm(P),
        move_attempts=P.move_attempts, ewald_alpha=P.ewald_alpha)


class BDSystem:
    def __init__(self, sysm: System, P: Params, precision: int = MIXED,
                 refactor_every: int = 1, coord: Optional[np.ndarray] = None, device: int = 0):
        L = _lib.load()
        _lib.check(L.bdh_init(device), "bdh_init")
        self.N = sysm.N
        self.n = 3 * self.N
        self.nbonds = len(sysm.bonds)
        self.cfg = make_config(P, precision, refactor_every)
        c = np.ascontiguousarray(sysm.coord if coord is None else coord, float).reshape(-1)
        Q = np.ascontiguousarray(sysm.Q, float)
        LJ = np.ascontiguousarray(sysm.LJ, float).reshape(-1)
        bonds = np.ascontiguousarray(sysm.bonds, np.int32).reshape(-1)
        parms = np.ascontiguousarray(sysm.bond_parms, float).reshape(-1)
        h = C.c_void_p()
        _lib.check(L.bdh_create(C.byref(self.cfg), self.N, _lib.dptr(c), _lib.dptr(Q),
                                _lib.dptr(LJ), len(sysm.bonds),
                                bonds.ctypes.data_as(C.POINTER(C.c_int)), _lib.dptr(parms),
                                C.byref(h)), "bdh_create")
        self._h = h
        if P.vel_grad_tensor is not None:
            self.set_flow(P.vel_grad_tensor)

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().bdh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ reference calls
    def compute_D(self) -> np.ndarray:
        D = np.zeros((self.n, self.n))
        _lib.check(_lib.load().bdh_compute_D(self._h, _lib.dptr(D), self.n), "compute_D")
        return D.T.copy()  # column-major buffer -> row-major array

    def compute_forces(self):
        F = np.zeros(self.n)
        E = np.zeros(3)
        rep = C.c_int()
        _lib.check(_lib.load().bdh_compute_forces(self._h, _lib.dptr(F), _lib.dptr(E),
                                                  C.byref(rep)), "compute_forces")
        return F, E, rep.value

    def neighbour_pairs(self, cutoff: float) -> np.ndarray:
        """Pairs (i < j) within `cutoff` from the device cell list, sorted, shape (P, 2)."""
        L = _lib.load()
        cap = 1 << 16
        while True:
            out = np.zeros((cap, 2), dtype=np.int32)
            cnt = C.c_longlong()
            _lib.check(L.bdh_neighbour_pairs(self._h, float(cutoff),
                                             out.ctypes.data_as(C.POINTER(C.c_int)), cap,
                                             C.byref(cnt)), "neighbour_pairs")
            if cnt.value <= cap:
                return out[:cnt.value].copy()
            cap = int(cnt.value)

    def compute_Rf(self) -> np.ndarray:
        Rf = np.zeros(self.n)
        _lib.check(_lib.load().bdh_compute_Rf(self._h, _lib.dptr(Rf)), "compute_Rf")
        return Rf

    def factor(self) -> np.ndarray:
        Lm = np.zeros((self.n, self.n))
        rc = _lib.load().bdh_factor(self._h, _lib.dptr(Lm), self.n)
        _lib.check(rc, "factor")
        return Lm.T.copy()

    def noise_product(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, float)
        y = np.zeros(self.n)
        _lib.check(_lib.load().bdh_noise_product(self._h, _lib.dptr(x), _lib.dptr(y)),
                   "noise_product")
        return y

    def compute_hydro(self, steps: int = 1, xi: Optional[np.ndarray] = None):
        x = None if xi is None else np.ascontiguousarray(xi, float).reshape(-1)
        if x is not None and x.size != steps * self.n:
            raise ValueError("xi must hold steps * 3N values")
        _lib.check(_lib.load().bdh_step(self._h, steps, None if x is None else _lib.dptr(x)),
                   "compute_hydro")

    def compute_Dij(self, i: int, j: int, kind: int = 0) -> np.ndarray:
        """3x3 block (rows of bead j, columns of bead i) at the current positions
        (bdh_compute_Dij): kind 0 the system's tensor, 1 free space, 2 Smith, 3 Beenakker."""
        t = np.zeros(9)
        _lib.check(_lib.load().bdh_compute_Dij(self._h, int(i), int(j), int(kind), _lib.dptr(t)),
                   "compute_Dij")
        return t.reshape(3, 3).T.copy()   # column-major -> row-major

    def compute_r(self) -> np.ndarray:
        """Rs of compute_r at the current positions from the next Gaussians of the stream."""
        Rs = np.zeros(self.n)
        _lib.check(_lib.load().bdh_compute_r(self._h, _lib.dptr(Rs)), "compute_r")
        return Rs

    def ermak_move(self, dt: Optional[float] = None) -> np.ndarray:
        """Positions (N, 4) of one Ermak-McCammon move from the current state; the system
        stays at its positions (bdh_ermak_move, compute_hydro_ermak)."""
        out = np.zeros(4 * self.N)
        _lib.check(_lib.load().bdh_ermak_move(self._h, float(self.cfg.dt if dt is None else dt),
                                              _lib.dptr(out)), "ermak_move")
        return out.reshape(-1, 4)

    def igt_phase(self, phase: int, dt: Optional[float] = None):
        """(positions (N, 4), result) of an IG-T phase (bdh_igt_phase): phase 0 predictor and
        corrector, 1 corrector only; result 1 / 2 when the predictor / corrector overlaps."""
        out = np.zeros(4 * self.N)
        res = C.c_int()
        _lib.check(_lib.load().bdh_igt_phase(self._h, int(phase), float(self.cfg.dt if dt is None else dt),
                                             _lib.dptr(out), C.byref(res)), "igt_phase")
        return out.reshape(-1, 4), res.value

    def tea_coefficients(self):
        """(beta, eps) of the last TEA displacement (bdh_tea_coefficients)."""
        b, e = C.c_double(), C.c_double()
        _lib.check(_lib.load().bdh_tea_coefficients(self._h, C.byref(b), C.byref(e)), "tea_coefficients")
        return b.value, e.value

    def set_ewald_split(self, alpha_scale: float = 1.0, R_real: Optional[int] = None,
                        R_recip: Optional[int] = None, r_cut: float = 0.0):
        """Beenakker split alpha = alpha_scale sqrt(pi)/L with real / reciprocal sphere radii
        around minimum-image separations (bdh_set_ewald_split); None keeps ewald_real;
        r_cut > 0 (A) also drops the real-space images farther than r_cut from the pair
        (bdh_set_ewald_cutoff); alpha_scale 0 returns to the reference's split."""
        Rr = self.cfg.ewald_real if R_real is None else int(R_real)
        Rk = self.cfg.ewald_real if R_recip is None else int(R_recip)
        L = _lib.load()
        _lib.check(L.bdh_set_ewald_split(self._h, float(alpha_scale), Rr, Rk), "set_ewald_split")
        if alpha_scale > 0.0:
            _lib.check(L.bdh_set_ewald_cutoff(self._h, float(r_cut)), "set_ewald_cutoff")

    def set_flow(self, G=None):
        """A flow field with velocity gradient G (3x3, row, column; vel_grad_tensor,
        diff_equation.c:265-285), or none for G None (bdh_set_flow)."""
        if G is None:
            _lib.check(_lib.load().bdh_set_flow(self._h, None), "set_flow")
            return
        self._G = np.ascontiguousarray(G, float).reshape(9)
        _lib.check(_lib.load().bdh_set_flow(self._h, _lib.dptr(self._G)), "set_flow")

    def set_forces(self, F):
        f = np.ascontiguousarray(F, float).reshape(-1)
        _lib.check(_lib.load().bdh_set_forces(self._h, _lib.dptr(f)), "set_forces")

    # ------------------------------------------------------------------ state access
    def coords(self) -> np.ndarray:
        c = np.zeros(4 * self.N)
        _lib.check(_lib.load().bdh_get_coords(self._h, _lib.dptr(c)), "get_coords")
        return c.reshape(-1, 4)

    def set_coords(self, coord):
        c = np.ascontiguousarray(coord, float).reshape(-1)
        _lib.check(_lib.load().bdh_set_coords(self._h, _lib.dptr(c)), "set_coords")

    def step_vjp(self, coord, xi, gbar, slot: Optional[int] = None) -> np.ndarray:
        """Cotangent of the positions before one FP64 step from coord with Gaussians xi,
        given the cotangent gbar (3N) of the positions after it (bdh_step_vjp); with slot,
        the factor at coord comes from that factor-cache slot (bdh_step_vjp_cached)."""
        c = np.ascontiguousarray(coord, float).reshape(-1)
        x = np.ascontiguousarray(xi, float).reshape(-1)
        g = np.ascontiguousarray(gbar, float).reshape(-1)
        out = np.zeros(self.n)
        L = _lib.load()
        if slot is None:
            rc = L.bdh_step_vjp(self._h, _lib.dptr(c), _lib.dptr(x), _lib.dptr(g), _lib.dptr(out))
        else:
            rc = L.bdh_step_vjp_cached(self._h, int(slot), _lib.dptr(c), _lib.dptr(x), _lib.dptr(g),
                                       _lib.dptr(out))
        _lib.check(rc, "step_vjp")
        return out

    def step_vjp_params(self, coord, xi, gbar, slot: Optional[int] = None):
        """step_vjp plus the cotangents of the force-field parameters for the same gbar
        (bdh_step_vjp_params): (rbar[3N], eps_bar[N], R_bar[N], Q_bar[N], H_bar[nbonds])."""
        c = np.ascontiguousarray(coord, float).reshape(-1)
        x = np.ascontiguousarray(xi, float).reshape(-1)
        g = np.ascontiguousarray(gbar, float).reshape(-1)
        out = np.zeros(self.n)
        eb, Rb, Qb = np.zeros(self.N), np.zeros(self.N), np.zeros(self.N)
        Hb = np.zeros(max(self.nbonds, 1))
        _lib.check(_lib.load().bdh_step_vjp_params(
            self._h, -1 if slot is None else int(slot), _lib.dptr(c), _lib.dptr(x), _lib.dptr(g),
            _lib.dptr(out), _lib.dptr(eb), _lib.dptr(Rb), _lib.dptr(Qb), _lib.dptr(Hb)),
            "step_vjp_params")
        return out, eb, Rb, Qb, Hb[:self.nbonds]

    def set_params(self, Q=None, LJ=None, bond_parms=None):
        """Replace the charges, LJ (R, eps) rows and per-bond (r0, r_max, H) rows
        (bdh_set_params); None keeps the current values."""
        def arr(a, shape):
            return None if a is None else np.ascontiguousarray(np.asarray(a, float).reshape(shape))
        q, lj, bp = arr(Q, -1), arr(LJ, -1), arr(bond_parms, -1)
        if q is not None and q.size != self.N or lj is not None and lj.size != 2 * self.N or \
                bp is not None and bp.size != 3 * self.nbonds:
            raise ValueError("set_params: Q[N], LJ[N, 2], bond_parms[nbonds, 3]")
        self._params = (q, lj, bp)   # kept alive across the call
        _lib.check(_lib.load().bdh_set_params(self._h, None if q is None else _lib.dptr(q),
                                              None if lj is None else _lib.dptr(lj),
                                              None if bp is None else _lib.dptr(bp)),
                   "set_params")

    def factor_cache(self, slots: int):
        _lib.check(_lib.load().bdh_factor_cache(self._h, int(slots)), "factor_cache")

    def save_factor(self, slot: int):
        _lib.check(_lib.load().bdh_save_factor(self._h, int(slot)), "save_factor")

    def forces(self) -> np.ndarray:
        F = np.zeros(self.n)
        _lib.check(_lib.load().bdh_get_forces(self._h, _lib.dptr(F)), "get_forces")
        return F

    def last_dt(self) -> float:
        """Length of the last step taken (below dt after ermak_var rejections)."""
        d = C.c_double()
        _lib.check(_lib.load().bdh_last_dt(self._h, C.byref(d)), "last_dt")
        return d.value

    def last_displacement(self):
        Rf, Rs = np.zeros(self.n), np.zeros(self.n)
        _lib.check(_lib.load().bdh_get_last_displacement(self._h, _lib.dptr(Rf), _lib.dptr(Rs)),
                   "get_last_displacement")
        return Rf, Rs

    def rng_state(self) -> int:
        s = C.c_uint64()
        _lib.check(_lib.load().bdh_get_rng_state(self._h, C.byref(s)), "get_rng_state")
        return s.value

    def set_rng_state(self, state: int):
        _lib.check(_lib.load().bdh_set_rng_state(self._h, state), "set_rng_state")

    def time_factor(self, reps: int = 1) -> float:
        ms = C.c_double()
        _lib.check(_lib.load().bdh_time_factor(self._h, reps, C.byref(ms)), "time_factor")
        return ms.value

    PHASES = ("noise", "assemble", "panel", "syrk", "matvec", "trmv", "move", "forces")

    def phase_timing(self, enable: bool = True):
        _lib.check(_lib.load().bdh_phase_timing(self._h, int(enable)), "phase_timing")

    def phase_times(self):
        """{phase: (total_ms, launches)} since the previous call."""
        ms = np.zeros(8)
        cnt = (C.c_longlong * 8)()
        _lib.check(_lib.load().bdh_phase_times(self._h, _lib.dptr(ms), cnt), "phase_times")
        return {p: (float(ms[i]), int(cnt[i])) for i, p in enumerate(self.PHASES)}

    def time_steps(self, steps: int) -> float:
        ms = C.c_double()
        _lib.check(_lib.load().bdh_time_steps(self._h, steps, C.byref(ms)), "time_steps")
        return ms.value


class BDBatch:
    """Independent chains of uneven sizes stepped together (one launch per phase over all of
    them).  Chain c follows the trajectory BDSystem(systems[c], P with rand_seed seeds[c])
    takes on its own; all chains share P (hydro cholesky, bc none, ermak_const, no
    check_overlap, brute-force pairs)."""

    def __init__(self, systems, P: Params, seeds, precision: int = MIXED,
                 refactor_every: int = 1, device: int = 0):
        L = _lib.load()
        _lib.check(L.bdh_init(device), "bdh_init")
        self.sizes = [s.N for s in systems]
        self.B = len(systems)
        self.Ntot = int(sum(self.sizes))
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(int)
        self.cfg = make_config(P, precision, refactor_every)
        N = np.array(self.sizes, dtype=np.int32)
        c = np.ascontiguousarray(np.concatenate([s.coord for s in systems]), float).reshape(-1)
        Q = np.ascontiguousarray(np.concatenate([s.Q for s in systems]), float)
        LJ = np.ascontiguousarray(np.concatenate([s.LJ for s in systems]), float).reshape(-1)
        nb = np.array([len(s.bonds) for s in systems], dtype=np.int32)
        bl = [np.asarray(s.bonds, np.int32).reshape(-1, 2) for s in systems]
        pl = [np.asarray(s.bond_parms, float).reshape(-1, 3) for s in systems]
        bonds = np.ascontiguousarray(np.concatenate(bl), np.int32).reshape(-1)
        parms = np.ascontiguousarray(np.concatenate(pl), float).reshape(-1)
        sd = np.array([int(x) & 0xFFFFFFFF for x in seeds], dtype=np.uint32)
        assert len(sd) == self.B
        h = C.c_void_p()
        _lib.check(L.bdh_batch_create(C.byref(self.cfg), self.B,
                                      N.ctypes.data_as(C.POINTER(C.c_int)), _lib.dptr(c),
                                      _lib.dptr(Q), _lib.dptr(LJ),
                                      nb.ctypes.data_as(C.POINTER(C.c_int)),
                                      bonds.ctypes.data_as(C.POINTER(C.c_int)), _lib.dptr(parms),
                                      sd.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(h)),
                   "bdh_batch_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().bdh_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def compute_hydro(self, steps: int = 1, xi: Optional[np.ndarray] = None):
        """steps steps of every chain; xi: (steps, 3 * sum N) Gaussians, or None for each
        chain's own stream."""
        p = None
        if xi is not None:
            xi = np.ascontiguousarray(xi, float).reshape(steps, 3 * self.Ntot)
            p = _lib.dptr(xi)
        _lib.check(_lib.load().bdh_batch_step(self._h, steps, p), "bdh_batch_step")

    def coords(self) -> np.ndarray:
        out = np.zeros((self.Ntot, 4))
        _lib.check(_lib.load().bdh_batch_get_coords(self._h, _lib.dptr(out)), "batch coords")
        return out

    def chain_coords(self, c: int) -> np.ndarray:
        return self.coords()[self.offsets[c]:self.offsets[c + 1]]

    def forces(self) -> np.ndarray:
        out = np.zeros(3 * self.Ntot)
        _lib.check(_lib.load().bdh_batch_get_forces(self._h, _lib.dptr(out)), "batch forces")
        return out

    def time_steps(self, steps: int) -> float:
        ms = C.c_double()
        _lib.check(_lib.load().bdh_batch_time_steps(self._h, steps, C.byref(ms)),
                   "batch time_steps")
        return ms.value

    def phase_timing(self, enable: bool = True):
        _lib.check(_lib.load().bdh_batch_phase_timing(self._h, int(enable)), "phase_timing")

    def phase_times(self):
        """{phase: (total_ms, launches)} since the previous call."""
        ms = np.zeros(8)
        cnt = (C.c_longlong * 8)()
        _lib.check(_lib.load().bdh_batch_phase_times(self._h, _lib.dptr(ms), cnt), "phase_times")
        return {p: (float(ms[i]), int(cnt[i])) for i, p in enumerate(BDSystem.PHASES)}

    def set_schedule(self, left_max: int):
        """Chains with padded size n_pad <= left_max factor with left-looking updates inside
        each outer panel, the others right-looking; -1 restores the single-system choice."""
        _lib.check(_lib.load().bdh_batch_set_schedule(self._h, int(left_max)), "set_schedule")

    def schedule(self) -> int:
        v = C.c_int()
        _lib.check(_lib.load().bdh_batch_get_schedule(self._h, C.byref(v)), "schedule")
        return v.value

    def autotune(self, reps: int = 3):
        """Times the refactor under each candidate schedule and keeps the fastest (before the
        first step).  Returns (chosen left_max, {left_max: ms})."""
        cand = (C.c_int * 8)()
        ms = np.zeros(8)
        n, best = C.c_int(), C.c_int()
        _lib.check(_lib.load().bdh_batch_autotune(self._h, int(reps), cand, _lib.dptr(ms),
                                                  C.byref(n), C.byref(best)), "autotune")
        return best.value, {int(cand[i]): float(ms[i]) for i in range(n.value)}

    def set_tiles(self, left_max: int, outer: int):
        _lib.check(_lib.load().bdh_batch_set_tiles(self._h, int(left_max), int(outer)), "set_tiles")

    def tiles(self):
        lm, ou = C.c_int(), C.c_int()
        _lib.check(_lib.load().bdh_batch_get_tiles(self._h, C.byref(lm), C.byref(ou)), "get_tiles")
        return lm.value, ou.value

    def autotune_tiles(self, reps: int = 3):
        """Times the refactor under every (schedule, outer panel width) candidate and keeps
        the fastest (before the first step).  Returns ((left_max, outer), {(left_max, outer):
        ms})."""
        M = 24
        cl, co = (C.c_int * M)(), (C.c_int * M)()
        ms = np.zeros(M)
        n, best = C.c_int(), C.c_int()
        _lib.check(_lib.load().bdh_batch_autotune_tiles(self._h, int(reps), cl, co, _lib.dptr(ms),
                                                        C.byref(n), C.byref(best)), "autotune_tiles")
        times = {(int(cl[i]), int(co[i])): float(ms[i]) for i in range(n.value)}
        return (int(cl[best.value]), int(co[best.value])), times


def gauss(state: int, count: int):
    """`count` Gaussians of the drand48 polar Box-Muller stream from 48-bit `state`."""
    st = C.c_uint64(state)
    out = np.zeros(count)
    _lib.check(_lib.load().bdh_gauss(C.byref(st), _lib.dptr(out), count), "gauss")
    return out, st.value


def seed_state(seed: int) -> int:
    return _lib.load().bdh_seed_state(seed & 0xFFFFFFFF)


def dpotrf(A: np.ndarray, uplo: str = "L", precision: int = FP64):
    """In-place Cholesky of a host matrix (my_dpotrf); returns (factor, info)."""
    n = A.shape[0]
    M = np.asfortranarray(A, dtype=float).copy(order="F")
    rc = _lib.load().bdh_dpotrf(uplo.encode(), M.ctypes.data_as(C.POINTER(C.c_double)), n, n,
                                precision)
    _lib.check(rc, "dpotrf")
    return np.asarray(M), rc


def dtrmv(A: np.ndarray, v: np.ndarray, uplo: str = "L", precision: int = FP64):
    n = A.shape[0]
    M = np.asfortranarray(A, dtype=float)
    v = np.ascontiguousarray(v, float)
    o = np.zeros(n)
    _lib.check(_lib.load().bdh_dtrmv(uplo.encode(), M.ctypes.data_as(C.POINTER(C.c_double)),
                                     _lib.dptr(v), _lib.dptr(o), n, n, precision), "dtrmv")
    return o


def probe_mfma(kind: int, A: np.ndarray, B: np.ndarray) -> np.ndarray:
    dt = np.float64 if kind == 2 else np.float32
    A = np.ascontiguousarray(A, dt)
    B = np.ascontiguousarray(B, dt)
    Cm = np.zeros((A.shape[0], B.shape[1]), dt)
    _lib.check(_lib.load().bdh_probe_mfma(kind, A.ctypes.data, B.ctypes.data, Cm.ctypes.data),
               "probe_mfma")
    return Cm
