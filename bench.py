#!/usr/bin/env python3
"""BD steps/s of the N=3000 free-space RPY + Cholesky Ermak-McCammon step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: one bead-spring chain of N=3000 beads (self-avoiding walk, 3.8 A bonds, >= 4 A
non-bonded separation, a = 1.9 A, harmonic bonds k = 10, LJ sigma 4 A eps 0.6 cutoff 12 A,
20 % of beads charged +-1 with Debye-Hueckel kappa^-1 10 A eps_r 78.5 cutoff 12 A,
T = 293 K, eta = 0.01 P, dt = 10 fs), free-space RPY, Cholesky noise refactored every
10 steps, Gaussians from the device drand48 stream.  One "step" = noise + (every 10th
step) assembly and factorization of the 9000 x 9000 tensor + D.F + L.xi + move + forces,
with all inputs resident in HBM.

Printed JSON line: `value` = mixed-precision steps/s with FP32 MFMA trailing updates
(FP32 storage and accumulation, FP64 diagonal blocks); the same steps on the all-FP64
path and with split-bf16 trailing updates beside it; the per-step relative displacement
error against the FP64 GPU step and against the FP64 CPU step; ||L L^T - D|| / ||D||; the
roofline of the dominant kernel (HIP events on the launch stream during a timed run); and
the CPU baselines timed on this host's cores: the reference's own binary (compiled from
its sources, OpenMP + OpenBLAS, refactor every step) and the FP64 CPU restatement at the
bench's refactor interval.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BD steps/s, N=3000 RPY+Cholesky: mixed vs FP64 GPU at rel err≤1e-5"
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"f32_mfma": 157.3, "f64_mfma": 78.6, "bf16_mfma": 2500.0}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_N3000.json")
TRAFFIC_FILE_BATCH = os.path.join(ROOT, "profiles", "traffic_batch256.json")
KB = 0.001986248532617


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload(N, seed, nb_list="brute"):
    from bd_box_amd import workloads as wl
    sysm = wl.chain(N, seed=seed)
    prm = wl.chain_params(rand_seed=20240 + seed, nb_list=nb_list)
    return sysm, prm, wl.parse_params(prm)


def periodic_workload(N, seed):
    """N hard spheres (a = 25 A) at volume fraction 0.2, random sequential addition in a
    periodic cube; WCA-like wall (sigma 50 A, eps 1, cutoff 2^(1/6) sigma); Beenakker-Ewald
    RPY with the reference's alpha = sqrt(pi)/L and lattice sphere R = 2 (first omitted
    real / reciprocal shells ~1e-8 / ~1e-7 relative); TEA noise; dt = 1 ps."""
    from bd_box_amd import workloads as wl
    sysm, L = wl.spheres(N, seed=seed, radius=25.0, phi=0.2)
    prm = wl.chain_params(bc="periodic", xbox=L, ybox=L, zbox=L, hydro="geyer",
                          ewald_method="beenakker", ewald_real=2, ewald_recip=2, elec="no",
                          alpha_lj=4.0, cutoff_lj=-1, lj_6_term="yes", dt=1.0,
                          nb_list="bucket_list", rand_seed=20240 + seed)
    return sysm, prm, wl.parse_params(prm), L


def batch_workload(count, seed):
    """`count` independent chains (the chain recipe above) with lengths log-uniform in
    [64, 1024] beads; chain k has its own drand48 seed."""
    from bd_box_amd import workloads as wl
    sizes = wl.batch_sizes(count, seed=seed)
    systems = wl.chains(sizes, seed=1000 * seed)
    prm = wl.chain_params(rand_seed=20240 + seed)
    seeds = [20240 + seed + 7919 * k for k in range(count)]
    return systems, prm, wl.parse_params(prm), seeds


# ----------------------------------------------------------------------------- roofline
def syrk_flops_per_factor(n, n_pad, nb=64):
    """Algorithmic flops of the trailing updates of one factorization (lower triangle
    only, unpadded size): sum over 64-column panels of nb * m (m + 1), m = rows below the
    panel.  The two-level schedule applies exactly these updates, grouped differently."""
    f = 0
    for c0 in range(0, n_pad, nb):
        m = max(0, n - (c0 + nb))
        f += nb * m * (m + 1)
    return f


def measured_traffic(kernel, path=TRAFFIC_FILE):
    """HBM bytes per launch of `kernel` (the kernel's symbol, e.g. "syrk_f32_kernel<false>")
    from the committed rocprofv3 counter summary, with the counter run's command."""
    try:
        with open(path) as f:
            t = json.load(f)
        cmd = t.get("command") or "; ".join(t.get("commands", []))
        for v in t["kernels"].values():
            if v.get("kernel", v.get("name")) == kernel:
                return v["hbm_bytes_per_launch"], cmd
    except (OSError, KeyError, ValueError):
        pass
    return None, None


def _counter_rows(d):
    """(kernel name, counter, value) of every dispatch of the rocprofv3 --pmc output under d:
    its SQLite database (the default output) or its CSV files (--output-format csv)."""
    import csv
    import glob
    import sqlite3
    rows = []
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                rows.append((r["Kernel_Name"], r["Counter_Name"], float(r["Counter_Value"])))
    for f in glob.glob(os.path.join(d, "**", "*.db"), recursive=True):
        with sqlite3.connect(f) as db:
            rows += [(n, c, float(v)) for n, c, v in
                     db.execute("select name, counter_name, counter_value from pmc_events")]
    return rows


def _short_kernel(name):
    """'void bdh::chol_epoch_f32_kernel<false>(float*, ...)' -> 'chol_epoch_f32_kernel<false>'."""
    name = name.strip().replace("(anonymous namespace)::", "")
    if name.startswith("void "):
        name = name[5:]
    depth, cut = 0, len(name)
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = i
            break
    name = name[:cut]
    return name[5:] if name.startswith("bdh::") else name


def summarize_traffic(fetch_dir, write_dir, out_path, command):
    """HBM bytes per launch of every kernel of a profile run from two rocprofv3 --pmc runs
    (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by what the same counter
    reads for the run's plain device copy of known size (--profile-only ends with three
    copies of 256 MiB each way)."""
    copy_bytes = (64 << 20) * 4   # bytes each way per copy
    per = {}
    for d, ctr in ((fetch_dir, "FETCH_SIZE"), (write_dir, "WRITE_SIZE")):
        for name, c, v in _counter_rows(d):
            if c != ctr:
                continue
            per.setdefault(name, {}).setdefault(ctr, []).append(v)
    copies = [k for k in per if "copy" in k.lower()]
    if not copies:
        sys.exit("summarize_traffic: no copy kernel in the counter runs")
    # the copy kernel also moves small staging buffers: the three 256 MiB copies are its three
    # largest dispatches
    cp = {c: sorted(v)[-3:] for c, v in per[copies[0]].items()}
    fetch_factor = copy_bytes / (np.mean(cp["FETCH_SIZE"]) * 1024.0)
    write_factor = copy_bytes / (np.mean(cp["WRITE_SIZE"]) * 1024.0)
    kernels = {}
    for name, c in per.items():
        if name in copies or "FETCH_SIZE" not in c or "WRITE_SIZE" not in c:
            continue
        fk, wk = float(np.mean(c["FETCH_SIZE"])), float(np.mean(c["WRITE_SIZE"]))
        short = _short_kernel(name)
        kernels[short] = {"kernel": short, "launches": len(c["FETCH_SIZE"]),
                          "fetch_kib_avg": fk, "write_kib_avg": wk,
                          "hbm_bytes_per_launch": fk * 1024.0 * fetch_factor
                          + wk * 1024.0 * write_factor}
    out = {"command": command,
           "calibration": {"copy_bytes_each_way": copy_bytes, "fetch_factor": fetch_factor,
                           "write_factor": write_factor},
           "kernels": kernels}
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k: round(v["hbm_bytes_per_launch"]) for k, v in kernels.items()}))


def update_steps(n_pad, nb=64, outer=256, left_max=4096):
    """(kb, K, t0, t1) of each update launch of one factorization: inside an outer panel
    the next panel's columns (n_pad <= left_max: left-looking, all of the outer panel's
    panels so far) or the outer panel's remaining columns (one panel); at its end the whole
    trailing matrix with the outer panel."""
    out = []
    k = 0
    while True:
        c0 = k * nb
        t0 = c0 + nb
        if t0 >= n_pad:
            return out
        P0 = c0 // outer * outer
        P1 = min(n_pad, P0 + outer)
        if t0 < P1 and n_pad <= left_max:
            out.append((P0, c0 + nb - P0, t0, t0 + nb))
        elif t0 < P1:
            out.append((c0, nb, t0, P1))
        else:
            out.append((P0, P1 - P0, t0, n_pad))
        k += 1


def update_bytes_per_factor(n_pad, word=4, left_max=4096, outer=256):
    """Algorithmic HBM bytes of one factorization's update launches: every updated C entry
    (rows t0..n_pad, columns t0..t1, lower part) read and written once, and the operand rows
    t0..n_pad x K read once per launch."""
    b = 0
    for kb, K, t0, t1 in update_steps(n_pad, outer=outer, left_max=left_max):
        rows = n_pad - t0
        w = t1 - t0
        c_entries = w * rows - w * (w - 1) // 2
        b += 2 * word * c_entries + word * rows * K
    return b


SYRK_KERNEL = {"fp64": "syrk_f64_kernel", "f32": "syrk_f32_kernel<false>",
               "split": "syrk_f32_kernel<true>"}
EPOCH_KERNEL = {"fp64": "chol_epoch_f64_kernel", "f32": "chol_epoch_f32_kernel<false>",
                "split": "chol_epoch_f32_kernel<true>"}
CHOL_EPOCH_MIN = 1024   # n_pad from which the factorization runs as one launch per outer panel


def chol_flops(n):
    """Flops of a Cholesky factorization of order n (LAPACK's dpotrf count)."""
    return n ** 3 / 3.0 + n ** 2 / 2.0 + n / 6.0


def roofline(phases, n, mode, refactors):
    """Roofline of the phase with the largest share of the timed run."""
    n_pad = -(-n // 128) * 128
    s = 8 if mode == "fp64" else 4
    dom = max(phases, key=lambda k: phases[k][0])
    ms, cnt = phases[dom]
    out = {"kernel": dom, "launches": cnt, "avg_ms": ms / max(cnt, 1)}
    if dom == "syrk":
        # the pipe the kernel issues on: FP32 MFMA (f32), FP64 MFMA (fp64), or bf16 MFMA at
        # three products per FP32 product (split: the ceiling is a third of the bf16 peak)
        peak = {"fp64": PEAK_TFLOPS["f64_mfma"], "f32": PEAK_TFLOPS["f32_mfma"],
                "split": PEAK_TFLOPS["bf16_mfma"] / 3.0}[mode]
        if n_pad >= CHOL_EPOCH_MIN:
            # one launch per 256-column outer panel does the whole factorization (diagonal
            # blocks, panel solves, trailing updates)
            kern = EPOCH_KERNEL[mode]
            fl = chol_flops(n) * refactors
            basis = (f"n={n}: n^3/3 + n^2/2 + n/6 flops per factorization (every flop of the "
                     f"factorization runs in these launches), {refactors} factorizations / "
                     f"{cnt} launches of {kern}")
        else:
            kern = SYRK_KERNEL[mode]
            fl = syrk_flops_per_factor(n, n_pad) * refactors
            basis = (f"n={n}: sum over 64-wide panels of 64 m (m+1) flops (lower-triangle "
                     f"trailing updates), {refactors} factorizations / {cnt} launches of {kern}")
        achieved = fl / (ms * 1e-3) / 1e12
        out["kernel_symbol"] = kern
        if mode == "split":
            basis += "; peak = bf16 dense MFMA peak / 3 (three bf16 products per FP32 product)"
        tr, cmd = measured_traffic(kern)
        out.update(bound="mfma", achieved=achieved, peak=peak, unit="TFLOP/s",
                   frac=achieved / peak, traffic=tr, basis=basis,
                   traffic_source=cmd and f"profiles/traffic_N3000.json: {cmd}")
    elif dom in ("trmv", "assemble"):
        per = s * n * (n + 1) / 2
        achieved = per * cnt / (ms * 1e-3) / 1e9
        out.update(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                   frac=achieved / HBM_PEAK_GBS, traffic=None,
                   basis="%d B per lower-triangle element, n(n+1)/2 elements" % s)
    else:
        out.update(bound="hbm", achieved=None, peak=HBM_PEAK_GBS, unit="GB/s", frac=None,
                   traffic=None, basis="not bandwidth or MFMA shaped")
    return out


# ----------------------------------------------------------------------------- CPU baseline
def host_cores():
    """Cores this process may use: the affinity mask, capped by OMP_NUM_THREADS where the
    host sets it (a one-GPU job on the GPU host is given 16 cores)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def cpu_baseline(N, seed, steps, threads):
    """Time the reference binary (compiled from its sources) on this host."""
    from bd_box_amd import workloads as wl
    exe = os.path.join(ROOT, "oracle", "_ref", "bd_box_lapack")
    kind = "reference"
    if not os.path.exists(exe):
        exe = os.path.join(ROOT, "oracle", "_ref", "bd_box")
        if not os.path.exists(exe):
            return None
    sysm, prm, _ = workload(N, seed)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OPENBLAS_NUM_THREADS=str(threads))
    times = {}
    with tempfile.TemporaryDirectory() as tmp:
        sp = os.path.join(tmp, "chain.str")
        wl.write_str(sp, sysm)
        for b in (1, steps + 1):
            pp = os.path.join(tmp, f"chain{b}.prm")
            wl.write_prm(pp, dict(prm, bdsteps=b), sp, tmp)
            t0 = time.perf_counter()
            r = subprocess.run([exe, pp], cwd=tmp, env=env, capture_output=True, text=True,
                               timeout=600)
            times[b] = time.perf_counter() - t0
            if r.returncode != 0:
                log(f"reference run failed: {r.stdout[-500:]} {r.stderr[-500:]}")
                return None
    dt = times[steps + 1] - times[1]
    return {"value": steps / dt, "unit": "steps/s", "cores": threads, "kind": kind,
            "sample": f"{steps} steps of the N={N} chain, reference semantics (refactor every "
                      f"step), {os.path.basename(exe)}: wall(bdsteps={steps + 1}) - "
                      f"wall(bdsteps=1) = {dt:.2f} s"}


def cpu_fp64_cycle(sysm, P, xi, R, threads):
    """The FP64 CPU restatement of the step at refactor interval R over len(xi) steps with
    the given Gaussians (the oracle's pinned pieces: compute_D, the reference's LAPACK
    dpotrf through scipy, forces; tests/compose.py is the same composition): per-step
    displacements and the timing of the run (the CPU baseline at the bench's interval)."""
    os.environ.setdefault("OPENBLAS_NUM_THREADS", str(threads))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from compose import compose_steps
    t0 = time.perf_counter()
    traj, disp = compose_steps(sysm, P, xi, R)
    dt = time.perf_counter() - t0
    K = len(xi)
    base = {"value": K / dt, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"{K} steps (one refactor cycle) of the N={sysm.N} chain, refactor every "
                      f"{R} steps, FP64 restatement (oracle compute_D and forces, OpenMP; "
                      f"LAPACK dpotrf and BLAS products of the OpenBLAS the reference's LAPACK "
                      f"build links): {dt:.2f} s"}
    return traj, disp, base


# ----------------------------------------------------------------------------- accuracy
def displacement_error(mixed, fp64, steps):
    """Per-step relative displacement error of the mixed path against the FP64 path from
    the same positions and the same Gaussians (both systems share step count and stream)."""
    errs = []
    for _ in range(steps):
        c = fp64.coords()
        mixed.set_coords(c)
        mixed.set_rng_state(fp64.rng_state())
        mixed.compute_hydro(1)
        fp64.compute_hydro(1)
        rf_m, rs_m = mixed.last_displacement()
        rf_d, rs_d = fp64.last_displacement()
        dm, dd = rf_m + rs_m, rf_d + rs_d
        errs.append(float(np.linalg.norm(dm - dd) / np.linalg.norm(dd)))
    return errs


def factor_residual(L, D, probes=6):
    """||L L^T - D||_F / ||D||_F, estimated with Gaussian probes (E ||A v||^2 = ||A||_F^2)."""
    rng = np.random.default_rng(0)
    num = den = 0.0
    for _ in range(probes):
        v = rng.normal(size=D.shape[0])
        r = L @ (L.T @ v) - D @ v
        num += r @ r
        den += (D @ v) @ (D @ v)
    return float(np.sqrt(num / den))


def vs_cpu_fp64(sysm, P, prec, R, xi, disp_cpu):
    """Per-step relative displacement error of a device step against the FP64 CPU step, from
    the same start with the same Gaussians, one step per call."""
    from bd_box_amd.system import BDSystem
    s = BDSystem(sysm, P, precision=prec, refactor_every=R)
    errs = []
    for k in range(len(xi)):
        s.compute_hydro(1, xi[k])
        rf, rs = s.last_displacement()
        errs.append(float(np.linalg.norm(rf + rs - disp_cpu[k]) / np.linalg.norm(disp_cpu[k])))
    F_dev = s.forces()
    coords = s.coords()
    s.close()
    return errs, F_dev, coords


def timed_steps(sysm, P, prec, R, warmup, steps):
    """steps/s of `steps` device-stream steps after `warmup` untimed ones; the system stays
    open for phase timing."""
    import torch
    from bd_box_amd.system import BDSystem
    s = BDSystem(sysm, P, precision=prec, refactor_every=R)
    s.compute_hydro(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.compute_hydro(steps)
    torch.cuda.synchronize()
    return s, steps / (time.perf_counter() - t0)


def phase_run(s, cyc):
    s.phase_timing(True)
    s.compute_hydro(cyc)
    ph = s.phase_times()
    s.phase_timing(False)
    return ph


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--N", type=int, default=3000)
    ap.add_argument("--refactor", type=int, default=10)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--precision", choices=("f32", "split", "fp64"), default="f32",
                    help="mixed-precision variant timed as `value`: trailing updates on FP32 "
                         "MFMA (f32, the configured step) or as split-bf16 MFMA products "
                         "(split); fp64 (with --profile-only) runs the FP64 step instead")
    ap.add_argument("--profile-only", action="store_true",
                    help="only the timed steps of --precision (for rocprofv3 runs), then a "
                         "256 MiB device copy that calibrates the HBM byte counters")
    ap.add_argument("--nb-list", choices=("brute", "bucket_list"), default="brute",
                    help="nonbonded pairs: all pairs, or the device cell list (the same pairs)")
    ap.add_argument("--workload", choices=("chain", "periodic", "batch", "adjoint"), default="chain",
                    help="chain: the N=3000 RPY + Cholesky step (default); periodic: N=8192 "
                         "spheres, Beenakker-Ewald RPY + TEA; batch: 256 independent chains "
                         "of 64-1024 beads stepped together; adjoint: parameter gradient of an "
                         "N=1024 FP64 trajectory (each its own JSON line)")
    ap.add_argument("--chains", type=int, default=256, help="chains of the batch workload")
    ap.add_argument("--ewald", choices=("converged", "reference"), default="converged",
                    help="periodic workload: the Ewald split for a 1e-6 truncation error, or the "
                         "reference's own split")
    ap.add_argument("--summarize-traffic", nargs=3, metavar=("FETCH_DIR", "WRITE_DIR", "OUT"),
                    help="write the HBM bytes per launch of every kernel from two rocprofv3 "
                         "--pmc runs of --profile-only (FETCH_SIZE, WRITE_SIZE) to OUT and exit")
    ap.add_argument("--traffic-command", default="",
                    help="with --summarize-traffic: the commands that made the counter runs")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="batch workload: keep the single-system update schedule instead of "
                         "timing the candidates before the run")
    args = ap.parse_args()
    if args.summarize_traffic:
        return summarize_traffic(*args.summarize_traffic, args.traffic_command)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the step of a "
              "single chain runs on one MI355X, and this build has no multi-GPU path)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from bd_box_amd.system import BDSystem, FP64, MIXED, MIXED_SPLIT
    assert torch.cuda.is_available(), "bench.py needs the GPU"
    torch.cuda.init()

    if args.workload == "periodic":
        return bench_periodic(args)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "adjoint":
        return bench_adjoint(args)

    sysm, prm, P = workload(args.N, args.seed, args.nb_list)
    n = 3 * sysm.N
    R = max(args.refactor, 1)
    log(f"workload: N={sysm.N} n={n} refactor_every={R}")
    if args.precision == "fp64" and not args.profile_only:
        sys.exit("bench.py: --precision fp64 is for --profile-only (the bench times FP64 beside "
                 "the mixed step itself)")
    precs = {"split": MIXED_SPLIT, "f32": MIXED, "fp64": FP64}
    prec = precs[args.precision]
    mixed, value = timed_steps(sysm, P, prec, R, args.warmup, args.steps)
    log(f"{args.precision}: {args.steps} steps, {value:.1f} steps/s")
    if args.profile_only:
        x = torch.empty(64 << 20, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"profile_only": True, "ms_per_step": 1e3 / value,
                          "copy_bytes": 2 * x.numel() * 4}))
        return

    # per-phase HIP events over a timed run of whole refactor cycles
    cyc = R * 2
    ph_mixed = phase_run(mixed, cyc)
    log(f"{args.precision} phases: " + json.dumps(ph_mixed))

    other = {}
    if args.precision == "f32":
        # the split-bf16 variant of the same step, for reference
        m2, other["split_bf16_steps_per_s"] = timed_steps(sysm, P, MIXED_SPLIT, R, args.warmup,
                                                          args.steps)
        other["split_bf16_phases_ms_per_step"] = {k: v[0] / cyc
                                                  for k, v in phase_run(m2, cyc).items()}
        m2.close()
        log(f"split-bf16 variant: {other['split_bf16_steps_per_s']:.1f} steps/s")
    elif args.precision == "split":
        m2, other["f32_mfma_steps_per_s"] = timed_steps(sysm, P, MIXED, R, args.warmup, args.steps)
        m2.close()

    fp64, fp64_rate = timed_steps(sysm, P, FP64, R, args.warmup, args.steps)
    log(f"fp64: {fp64_rate:.1f} steps/s")
    ph_fp64 = phase_run(fp64, cyc)
    log("fp64 phases: " + json.dumps(ph_fp64))

    acc = {}
    if not args.no_accuracy:
        # against the FP64 GPU step (the metric's comparator): from the same positions with
        # the same stream, starting on a refactor step of both systems
        done = args.warmup + args.steps + cyc
        pad = (-done) % R
        if pad:
            mixed.compute_hydro(pad)
            fp64.compute_hydro(pad)
        errs = displacement_error(mixed, fp64, R)
        acc["rel_disp_err_max"] = max(errs)
        acc["rel_disp_err_mean"] = float(np.mean(errs))
    mixed.close()
    fp64.close()

    threads = host_cores()
    cpu = cpu_port = None
    if not args.no_cpu_baseline:
        # CPU leg: the reference binary (refactor every step), then the FP64 CPU restatement
        # over one refactor cycle with given Gaussians, which also serves as the FP64 CPU
        # reference of the displacement and factor errors
        log(f"cpu baseline: reference binary, {threads} threads, {args.cpu_steps} steps")
        cpu = cpu_baseline(args.N, args.seed, args.cpu_steps, threads)
        xi = np.random.default_rng(args.seed).normal(size=(R, n))
        traj, disp_cpu, cpu_port = cpu_fp64_cycle(sysm, P, xi, R, threads)
        log(f"cpu restatement: {cpu_port['sample']}")
        if not args.no_accuracy:
            from oracle import pyoracle as po
            p = po.params_from(P)
            for name, pr in (("", prec), ("fp64_", FP64)):
                e, F_dev, c_dev = vs_cpu_fp64(sysm, P, pr, R, xi, disp_cpu)
                F_ref, _, _ = po.forces(sysm, p, c_dev)
                acc[f"{name}rel_disp_err_vs_cpu_fp64_max"] = max(e)
                acc[f"{name}rel_force_err_vs_cpu_fp64"] = float(
                    np.linalg.norm(F_dev - F_ref) / np.linalg.norm(F_ref))
            s = BDSystem(sysm, P, precision=prec, refactor_every=R)
            L = s.factor()
            s.close()
            D = po.compute_D(sysm.coord, p, out=np.empty((n, n)))
            acc["chol_rel_resid"] = factor_residual(L, D)
            del L, D
        log(f"accuracy: {acc}")

    refactors = cyc // R
    roof = roofline(ph_mixed, n, args.precision, refactors)
    kind = {"f32": "FP32 MFMA", "split": "split-bf16 MFMA"}[args.precision]
    line = {
        "metric": METRIC, "value": value, "unit": "steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 / value,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"N={args.N} bead-spring chain, free-space RPY + Cholesky "
                               f"refactored every {R} steps, mixed precision ({kind} trailing "
                               f"updates with FP32 accumulation, FP64 diagonal blocks)",
                   "N": args.N, "n": n, "refactor_every": R,
                   "noise": "device drand48 polar Box-Muller stream", "parallelism": "1 GPU"},
        "fp64_steps_per_s": fp64_rate,
        "speedup_vs_fp64": value / fp64_rate,
        **other,
        **acc,
        "accuracy_ok": (acc["rel_disp_err_max"] <= 1e-5) if "rel_disp_err_max" in acc else None,
        "phases_ms_per_step": {k: v[0] / cyc for k, v in ph_mixed.items()},
        "fp64_phases_ms_per_step": {k: v[0] / cyc for k, v in ph_fp64.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
        "cpu_baseline_fp64_restatement": cpu_port,
    }
    print(json.dumps(line))


def ewald_split_for(scale: float, tol: float = 1e-6):
    """Real and reciprocal sphere radii of the Beenakker split alpha = scale sqrt(pi)/L for a
    truncation error of about tol (image sums around minimum-image separations), and the
    real-space cutoff r_c in units of L: an image farther than r_c = sqrt(x) / scale L from
    the pair carries exp(-alpha^2 r^2) < exp(-pi x), and every image within r_c has
    |n| <= r_c / L + sqrt(3)/2, which fixes R_real; an omitted reciprocal vector carries
    exp(-pi |n|^2 / scale^2).  The reciprocal exponent pi x2 is taken 3 above ln(1/tol) for the
    polynomial prefactors of one term, and the real-space one pi x 6 above, since a cutoff drops the many
    images just beyond r_c at once (measured: 3 above leaves 1e-6 to 4e-6 of max |D| at
    alpha = (0.8 .. 1.5) sqrt(pi)/L, 6 above ~1.5e-7)."""
    x2 = (np.log(1.0 / tol) + 3.0) / np.pi
    rc = float(np.sqrt((np.log(1.0 / tol) + 6.0) / np.pi) / scale)
    d = rc + np.sqrt(3.0) / 2.0
    Rr = int(np.ceil(np.sqrt(max(0.0, d * d - 1.0))))
    Rk = int(np.ceil(np.sqrt(max(0.0, x2 * scale * scale - 1.0))))
    return Rr, Rk, rc


def images_within(xyz, L, Rr, rc, samples=200000, seed=0):
    """Mean number of real-space images r + n L (|n| <= Rr, r the minimum-image separation)
    within rc L of a pair, over a sample of the pairs i < j of xyz."""
    rng = np.random.default_rng(seed)
    N = len(xyz)
    i = rng.integers(0, N, size=samples)
    j = rng.integers(0, N, size=samples)
    keep = i != j
    r = (xyz[j[keep]] - xyz[i[keep]]) / L
    r -= np.rint(r)
    ns = np.array([(h, f, g) for h in range(-Rr, Rr + 1) for f in range(-Rr, Rr + 1)
                   for g in range(-Rr, Rr + 1) if h * h + f * f + g * g <= Rr * Rr], float)
    cnt = 0.0
    for c in range(0, len(r), 20000):
        d2 = ((r[c:c + 20000, None, :] + ns[None, :, :]) ** 2).sum(-1)
        cnt += ((d2 <= rc * rc) & (d2 > 0)).sum()
    return cnt / len(r)


def lattice_points(R):
    return sum(1 for h in range(-R, R + 1) for f in range(-R, R + 1) for g in range(-R, R + 1)
               if h * h + f * f + g * g <= R * R)


# FP32-equivalent VALU operations of one pair block (bdh_ewald.hpp, beenakker_block): a
# real-space image (shifted separation, inverse powers from one reciprocal square root,
# exp, erfc from that exp, the two scalar coefficients, six FP64 accumulations) and a
# reciprocal vector (FP64: the tabulated factor, the phase from stepped per-axis phasors, six
# accumulations), FP64 operations counted twice (the FP64 VALU rate is half the FP32 one).
OPS_REAL_IMAGE = 80
OPS_RECIP_VECTOR = 40
FP32_VALU_PEAK_TFLOPS = 157.3


def cpu_baseline_periodic(sysm, prm, threads, timeout=600):
    """One step of the reference binary on the periodic suspension (its own Ewald split,
    alpha = sqrt(pi)/L, lattice sphere R = 2, on-the-fly TEA): wall(bdsteps=2) -
    wall(bdsteps=1)."""
    from bd_box_amd import workloads as wl
    exe = os.path.join(ROOT, "oracle", "_ref", "bd_box_lapack")
    if not os.path.exists(exe):
        exe = os.path.join(ROOT, "oracle", "_ref", "bd_box")
        if not os.path.exists(exe):
            return None
    # the reference's start-up moves the centre of mass to the origin and stops unless every
    # bead then lies within half a box of it (trans.c:75-97); a box-filling suspension has
    # beads up to |centre of mass| beyond that, so the run takes a box just large enough
    xyz = np.asarray(sysm.coord, float)[:, :3]
    Lr = float(2.0 * np.max(np.abs(xyz - xyz.mean(axis=0))) * (1.0 + 1e-6))
    Lr = max(Lr, float(prm["xbox"]))
    # its bucket list stops with "Wrong cell" at the second step on this box
    # (potentials/bucket.c:127), so the forces take its brute-force pair loop
    prm = dict(prm, xbox=Lr, ybox=Lr, zbox=Lr, nb_list="brute")
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OPENBLAS_NUM_THREADS=str(threads))
    times = {}
    with tempfile.TemporaryDirectory() as tmp:
        sp = os.path.join(tmp, "spheres.str")
        wl.write_str(sp, sysm)
        for b in (1, 2):
            pp = os.path.join(tmp, f"s{b}.prm")
            wl.write_prm(pp, dict(prm, bdsteps=b), sp, tmp)
            t0 = time.perf_counter()
            try:
                r = subprocess.run([exe, pp], cwd=tmp, env=env, capture_output=True, text=True,
                                   timeout=timeout)
            except subprocess.TimeoutExpired:
                log(f"reference periodic run exceeded {timeout} s")
                return None
            times[b] = time.perf_counter() - t0
            if r.returncode != 0:
                log(f"reference run failed: {r.stdout[-500:]} {r.stderr[-500:]}")
                return None
    dt = times[2] - times[1]
    return {"value": 1.0 / dt, "unit": "steps/s", "cores": threads, "kind": "reference",
            "sample": f"1 step of the N={sysm.N} suspension, reference semantics (its Beenakker "
                      f"split alpha = sqrt(pi)/L, R = 2; on-the-fly TEA; brute-force pairs), box "
                      f"{Lr:.1f} A so that its centring keeps every bead inside, "
                      f"{os.path.basename(exe)}: wall(bdsteps=2) - wall(bdsteps=1) = {dt:.2f} s"}


def bench_periodic(args):
    """Steps/s of the periodic suspension with Beenakker-Ewald TEA noise, mixed (FP32 pair
    blocks, FP64 Ewald reciprocal/self sums and accumulation) and FP64.  The Ewald split is
    chosen for a 1e-6 truncation error (alpha = 1.25 sqrt(pi)/L, real and reciprocal spheres
    of radius 3 around minimum-image separations, the box's wave vectors); --ewald reference
    runs the reference's own split (alpha = sqrt(pi)/L, one sphere of radius 2 around the
    wrapped separation), whose reciprocal terms underflow."""
    import torch
    from bd_box_amd.system import BDSystem, FP64, MIXED
    N = 8192 if args.N == 3000 else args.N
    steps = min(args.steps, 50)
    warm = min(args.warmup, 2)
    sysm, prm, P, L = periodic_workload(N, args.seed)
    converged = args.ewald == "converged"
    scale = 1.25
    Rr, Rk, rc = ewald_split_for(scale) if converged else (P.ewald_real, P.ewald_real, 0.0)
    log(f"periodic workload: N={N} box={L:.1f} A, {steps} timed steps, Ewald split "
        f"{'alpha = %.2f sqrt(pi)/L, radii %d/%d, cutoff %.2f L' % (scale, Rr, Rk, rc) if converged else 'reference'}")
    out = {}
    for name, prec in (("mixed", MIXED), ("fp64", FP64)):
        s = BDSystem(sysm, P, precision=prec)
        if converged:
            s.set_ewald_split(scale, Rr, Rk, r_cut=rc * L)
        s.compute_hydro(warm)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.compute_hydro(steps)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        s.phase_timing(True)
        s.compute_hydro(2)
        ph = s.phase_times()
        s.phase_timing(False)
        out[name] = (steps / t, {k: v[0] / 2 for k, v in ph.items()})
        log(f"{name}: {steps / t:.2f} steps/s phases {out[name][1]}")
        s.close()
    # roofline of the pair sums (HIP events on the launch stream): FP32-equivalent VALU
    # operations of every unordered pair's block (and the self blocks) per step
    # real-space images per pair: those within the cutoff (a sample's mean), or the sphere
    n_real = images_within(sysm.coord[:, :3], L, Rr, rc) if converged else lattice_points(Rr)
    # reciprocal vectors evaluated per pair: half the sphere (k and -k carry the same term);
    # none for the reference's split, whose reciprocal factors underflow to zero
    n_recip = (lattice_points(Rk) - 1) // 2 if converged else 0
    pairs = N * (N + 1) // 2
    ops = pairs * (n_real * OPS_REAL_IMAGE + n_recip * OPS_RECIP_VECTOR)
    t_pairs = out["mixed"][1].get("matvec", 0.0) * 1e-3
    roof = None
    if t_pairs > 0:
        ach = ops / t_pairs / 1e12
        roof = {"kernel": "pair_sums_kernel<float, beenakker, tea>", "bound": "valu",
                "achieved": ach, "peak": FP32_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": ach / FP32_VALU_PEAK_TFLOPS, "traffic": None,
                "basis": f"{pairs} pair blocks per step x ({n_real:.1f} real-space images x "
                         f"{OPS_REAL_IMAGE} + {n_recip} reciprocal vectors x {OPS_RECIP_VECTOR}) "
                         f"FP32-equivalent VALU operations (FP64 counted twice) / the pair-sum "
                         f"phase's time per step"}
    cpu = None
    if not args.no_cpu_baseline:
        threads = os.cpu_count() or 1
        cpu = cpu_baseline_periodic(sysm, prm, threads)
    line = {
        "metric": "BD steps/s, N=8192 periodic Beenakker-Ewald RPY + TEA",
        "value": out["mixed"][0], "unit": "steps/s", "n_gpus": 1, "steps": steps,
        "warmup": warm, "ms_per_step": 1e3 / out["mixed"][0], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"N={N} spheres a=25 A, phi=0.2, periodic cube L={L:.1f} A, "
                               "WCA wall sigma 50 A eps 1, Beenakker-Ewald, TEA, dt 1 ps",
                   "ewald": (f"alpha = {scale} sqrt(pi)/L, real/reciprocal radii {Rr}/{Rk}, "
                             f"real-space cutoff {rc:.3f} L (1e-6 truncation)") if converged else
                            "reference split: alpha = sqrt(pi)/L, radius 2",
                   "N": N, "box": L, "parallelism": "1 GPU"},
        "fp64_steps_per_s": out["fp64"][0],
        "phases_ms_per_step": out["mixed"][1],
        "fp64_phases_ms_per_step": out["fp64"][1],
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def bench_batch(args):
    """Steps/s of a batch of independent chains (every chain advances one step per step),
    mixed (FP32 MFMA trailing updates, FP64 diagonal blocks) beside FP64 and split-bf16,
    each phase one launch over all chains and every full refactor cycle one graph launch;
    tile shapes (update schedule, outer panel width) chosen by the autotuner per precision."""
    import torch
    from bd_box_amd.system import BDBatch, FP64, MIXED, MIXED_SPLIT
    systems, prm, P, seeds = batch_workload(args.chains, args.seed)
    sizes = [s.N for s in systems]
    Ntot = int(sum(sizes))
    R = max(args.refactor, 1)
    steps = min(args.steps, 10 * R)
    warm = min(args.warmup, R)
    log(f"batch workload: {len(sizes)} chains, {Ntot} beads (N {min(sizes)}..{max(sizes)}), "
        f"{steps} timed steps")
    out, sched = {}, {}
    for name, prec in (("mixed", MIXED), ("fp64", FP64), ("split", MIXED_SPLIT)):
        b = BDBatch(systems, P, seeds, precision=prec, refactor_every=R)
        if args.autotune:
            best, times = b.autotune_tiles(reps=3)
            sched[name] = {"left_max": best[0], "outer": best[1],
                           "refactor_ms": {f"{k[0]}/{k[1]}": v for k, v in times.items()}}
            log(f"{name}: autotuned tiles left_max={best[0]} outer={best[1]} (refactor ms {times})")
        b.compute_hydro(warm)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.compute_hydro(steps)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        b.phase_timing(True)
        b.compute_hydro(2 * R)
        ph = b.phase_times()
        b.phase_timing(False)
        out[name] = (steps / t, ph)
        log(f"{name}: {steps / t:.2f} steps/s phases {json.dumps(ph)}")
        b.close()
    # accuracy: R steps of both precisions from the same state with the same streams
    disp = {}
    for name, prec in (("mixed", MIXED), ("fp64", FP64)):
        b = BDBatch(systems, P, seeds, precision=prec, refactor_every=R)
        if name in sched:
            b.set_tiles(sched[name]["left_max"], sched[name]["outer"])
        b.compute_hydro(R)
        disp[name] = b.coords()[:, :3]
        b.close()
    x0 = np.concatenate([s.coord for s in systems])[:, :3]
    off = np.concatenate([[0], np.cumsum(sizes)])
    errs = []
    for c in range(len(sizes)):
        dm = disp["mixed"][off[c]:off[c + 1]] - x0[off[c]:off[c + 1]]
        dd = disp["fp64"][off[c]:off[c + 1]] - x0[off[c]:off[c + 1]]
        errs.append(float(np.linalg.norm(dm - dd) / np.linalg.norm(dd)))
    ph = out["mixed"][1]
    cyc = 2 * R
    ms, cnt = ph["syrk"]
    fl = sum(syrk_flops_per_factor(3 * N, -(-3 * N // 128) * 128) for N in sizes) * 2
    lm = sched["mixed"]["left_max"] if "mixed" in sched else 4096
    ow = sched["mixed"]["outer"] if "mixed" in sched else 256
    by = sum(update_bytes_per_factor(-(-3 * N // 128) * 128, left_max=lm, outer=ow) for N in sizes) * 2
    ach = by / (ms * 1e-3) / 1e9
    tr, _ = measured_traffic("syrk_f32_batch_kernel<false>", TRAFFIC_FILE_BATCH)
    roof = {"kernel": "syrk", "launches": cnt, "avg_ms": ms / max(cnt, 1), "bound": "hbm",
            "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": tr, "mfma_tflops": fl / (ms * 1e-3) / 1e12,
            "mfma_frac": fl / (ms * 1e-3) / 1e12 / PEAK_TFLOPS["f32_mfma"],
            "basis": "sum over chains and update launches of the updated lower-trapezoid C "
                     "entries read and written (FP32) plus the K-column operand rows read once, "
                     "2 factorizations; traffic = measured HBM bytes per launch"}
    value = out["mixed"][0]
    line = {
        "metric": f"BD steps/s, batch of {len(sizes)} chains N~logU[64,1024] RPY+Cholesky",
        "value": value, "unit": "steps/s", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": 1e3 / value, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"{len(sizes)} bead-spring chains, lengths log-uniform in "
                               f"[64, 1024] (sum {Ntot} beads), free-space RPY + Cholesky "
                               f"refactored every {R} steps, FP32 MFMA trailing updates, FP64 "
                               f"diagonal blocks, each refactor cycle one graph launch",
                   "chains": len(sizes), "beads": Ntot, "refactor_every": R,
                   "update_schedule": sched or "single-system default (left_max 4096)",
                   "parallelism": "1 GPU"},
        "chain_steps_per_s": value * len(sizes),
        "bead_steps_per_s": value * Ntot,
        "fp64_steps_per_s": out["fp64"][0],
        "speedup_vs_fp64": value / out["fp64"][0],
        "split_bf16_steps_per_s": out["split"][0],
        "rel_disp_err_max": max(errs),
        "phases_ms_per_step": {k: v[0] / cyc for k, v in ph.items()},
        "fp64_phases_ms_per_step": {k: v[0] / cyc for k, v in out["fp64"][1].items()},
        "roofline": roof,
        "cpu_baseline": None,
    }
    print(json.dumps(line))


def bench_adjoint(args):
    """Gradient of the loss (mean squared radius of gyration over the K frames - 400 A^2)^2
    of an FP64 trajectory (N = 1024 chain of the chain recipe, hydrodynamic radius 1.9 A,
    given Gaussians, refactor every step) with respect to the LJ epsilon and sigma, the bond
    constant k and the 1024 per-bead charges: forward, then the checkpointed reverse pass
    (every 10 steps), timed together."""
    import torch
    from bd_box_amd import adjoint as adj
    from bd_box_amd import workloads as wl
    from bd_box_amd.system import BDSystem, FP64, MIXED
    N = 1024 if args.N == 3000 else args.N
    K = min(args.steps, 100)
    sysm = wl.chain(N, seed=args.seed)
    P = wl.parse_params(wl.chain_params(rand_seed=20240 + args.seed))
    rng = np.random.default_rng(args.seed)
    xi = rng.normal(size=(K, 3 * N))
    nb = len(sysm.bonds)
    log(f"adjoint workload: N={N}, {K} steps, checkpoints every 10")
    f64 = dict(dtype=torch.float64)
    x0 = torch.tensor(sysm.coord[:, :3].copy(), **f64)
    r0rmax = np.array(sysm.bond_parms, float)[:, :2]

    def run(precision):
        s = BDSystem(sysm, P, precision=precision, refactor_every=1)
        eps = torch.tensor(0.6, requires_grad=True, **f64)
        sig = torch.tensor(4.0, requires_grad=True, **f64)
        kb = torch.tensor(10.0, requires_grad=True, **f64)
        Q = torch.tensor(np.array(sysm.Q, float), requires_grad=True, **f64)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        frames = adj.param_trajectory(s, x0, eps * torch.ones(N, **f64), sig / 2 * torch.ones(N, **f64),
                                      Q, kb * torch.ones(nb, **f64), xi, sysm.coord[:, 3], r0rmax,
                                      checkpoint=10)
        loss = adj.rg_loss(frames)
        t1 = time.perf_counter()
        loss.backward()
        t2 = time.perf_counter()
        s.close()
        grads = {"eps": eps.grad.item(), "sigma": sig.grad.item(), "k": kb.grad.item(),
                 "Q_norm": float(Q.grad.norm()), "Q_first": Q.grad[:4].tolist()}
        qg = Q.grad.detach().numpy().copy()
        log(f"precision {precision}: forward {t1 - t0:.3f} s, backward {t2 - t1:.3f} s, "
            f"loss {loss.item():.6e}, grads {grads}")
        return t0, t1, t2, loss.item(), grads, qg

    t0, t1, t2, loss, grads, qg = run(FP64)
    m0, m1, m2, mloss, mgrads, mqg = run(MIXED)
    # the mixed-precision gradient against the FP64 one: each scalar parameter, and the
    # charges' gradient as a vector
    mixed_rel = max([abs(mgrads[k] - grads[k]) / abs(grads[k]) for k in ("eps", "sigma", "k")] +
                    [float(np.linalg.norm(mqg - qg) / np.linalg.norm(qg))])
    line = {
        "metric": f"gradient steps/s, N={N} FP64 trajectory, parameter gradient of the Rg loss "
                  "(forward + reverse)",
        "value": K / (t2 - t0), "unit": "steps/s", "n_gpus": 1, "steps": K, "warmup": 0,
        "ms_per_step": 1e3 * (t2 - t0) / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"N={N} bead-spring chain (a = 1.9 A), free-space RPY + Cholesky "
                               f"refactored every step, {K} steps with given Gaussians, "
                               "checkpoints every 10 steps, loss (mean Rg^2 - 400)^2, gradient "
                               "w.r.t. LJ eps and sigma, bond k and the per-bead charges",
                   "N": N, "parallelism": "1 GPU"},
        "forward_s": t1 - t0, "backward_s": t2 - t1, "loss": loss, "gradient": grads,
        "mixed_gradient_steps_per_s": K / (m2 - m0), "mixed_forward_s": m1 - m0,
        "mixed_backward_s": m2 - m1, "mixed_gradient_rel_diff_vs_fp64": mixed_rel,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
