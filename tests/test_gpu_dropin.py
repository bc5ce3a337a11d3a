"""The unmodified reference driver on the HIP backend.

oracle/_ref/bd_box_hip is the reference's own driver (main.c, main_loop.c, input.c,
output.c, ...) linked with bd_box_amd/csrc/bdbox_dropin.c + libbdbox_hip.so in place of
its backend TUs; oracle/_ref/bd_box is the reference itself.  Both run the same .prm/.str
and the xyz trajectories they write (%f, 6 decimals; src/output.c:64-75) must agree.
Both binaries are built where the reference sources are mounted (make -C oracle ref dropin)
and travel with the tree; the test skips when they are absent.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from bd_box_amd import workloads as wl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "bd_box")
HIP = os.path.join(ROOT, "oracle", "_ref", "bd_box_hip")

pytestmark = pytest.mark.gpu


def read_xyz(path):
    frames, cur = [], []
    with open(path) as f:
        lines = f.read().splitlines()
    i = 0
    while i < len(lines):
        n = int(lines[i].split()[0])
        block = lines[i + 2:i + 2 + n]
        frames.append([[float(v) for v in ln.split()[1:4]] for ln in block])
        i += 2 + n
    return np.array(frames)


def run(exe, sysm, prm, tmp, env_extra=None):
    sp = os.path.join(tmp, "c.str")
    pp = os.path.join(tmp, "c.prm")
    wl.write_str(sp, sysm)
    wl.write_prm(pp, prm, sp, tmp)
    env = dict(os.environ, OMP_NUM_THREADS="1", **(env_extra or {}))
    r = subprocess.run([exe, pp], cwd=tmp, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (exe, r.stdout[-2000:], r.stderr[-2000:])
    return read_xyz(os.path.join(tmp, "out.xyz"))


@pytest.mark.skipif(not (os.path.exists(REF) and os.path.exists(HIP)),
                    reason="reference / drop-in binaries not built here")
@pytest.mark.parametrize("N,hydro", [(48, "cholesky"), (30, "none")])
def test_dropin_trajectory_matches_reference(N, hydro):
    sysm = wl.chain(N, seed=N)
    prm = wl.chain_params(hydro=hydro, rand_seed=321, bdsteps=9, save_xyz_freq=1)
    with tempfile.TemporaryDirectory() as t1, tempfile.TemporaryDirectory() as t2:
        a = run(REF, sysm, prm, t1)
        b = run(HIP, sysm, prm, t2, {"BDBOX_HIP_PRECISION": "fp64"})
    assert a.shape == b.shape == (9, N, 3)
    # 6-decimal text output: allow one unit in the last printed place
    assert np.max(np.abs(a - b)) <= 1.5e-6, np.max(np.abs(a - b))


DROPIN_CASES = {
    # rejected steps redrawn on the device (diff_comp.c:152-179), dt halved by ermak_var
    "retry_var": (dict(lj_diam=3.3), dict(hydro="cholesky", algorithm="ermak_var", dt=2.0,
                                          check_overlap="yes", move_attempts=200)),
    "retry_const": (dict(lj_diam=3.3), dict(hydro="cholesky", algorithm="ermak_const", dt=0.5,
                                            check_overlap="yes", move_attempts=1000)),
    "geyer": ({}, dict(hydro="geyer")),
    # IG-T predictor-corrector (diff_comp.c:76-144), rejections redrawn on the device
    "igt_var_retry": (dict(lj_diam=3.3), dict(hydro="cholesky", algorithm="igt_var", dt=2.0,
                                              check_overlap="yes", move_attempts=200)),
    "igt_const": ({}, dict(hydro="cholesky", algorithm="igt_const")),
    "igt_geyer": ({}, dict(hydro="geyer", algorithm="igt_const")),
    # periodic chain with the Smith-Ewald tensor (diff_tensor.c:102-263)
    "smith_periodic": ({}, dict(hydro="cholesky", bc="periodic", xbox=100.0, ybox=100.0,
                                zbox=100.0, ewald_method="smith", ewald_real=2, ewald_recip=2,
                                ewald_alpha=2.5)),
    "bucket_list": ({}, dict(hydro="cholesky", nb_list="bucket_list")),
    # periodic chain over the bucket lists (bucket.c:173-379, phantom images at the faces)
    "bucket_list_periodic": ({}, dict(hydro="none", bc="periodic", xbox=60.0, ybox=60.0,
                                      zbox=60.0, nb_list="bucket_list")),
    "bucket_space_periodic": ({}, dict(hydro="none", bc="periodic", xbox=60.0, ybox=60.0,
                                       zbox=60.0, nb_list="bucket_space")),
    "bucket_space": ({}, dict(hydro="none", nb_list="bucket_space")),
    # a flow field (vel_grad_tensor, diff_equation.c:265-285, :304-312)
    "flow": ({}, dict(hydro="cholesky", vel_grad_tensor="0.1 2.0 -0.3 0.4 -0.2 0.5 -0.6 0.7 0.15")),
    "flow_igt": ({}, dict(hydro="cholesky", algorithm="igt_const",
                          vel_grad_tensor="0.1 2.0 -0.3 0.4 -0.2 0.5 -0.6 0.7 0.15")),
}


@pytest.mark.skipif(not (os.path.exists(REF) and os.path.exists(HIP)),
                    reason="reference / drop-in binaries not built here")
@pytest.mark.parametrize("name", sorted(DROPIN_CASES))
def test_dropin_cases_match_reference(name):
    sopt, popt = DROPIN_CASES[name]
    N = 24
    sysm = wl.chain(N, seed=5, **sopt)
    prm = wl.chain_params(rand_seed=11, bdsteps=8, save_xyz_freq=1, **popt)
    print(f"drop-in {name}", flush=True)
    with tempfile.TemporaryDirectory() as t1, tempfile.TemporaryDirectory() as t2:
        a = run(REF, sysm, prm, t1)
        b = run(HIP, sysm, prm, t2, {"BDBOX_HIP_PRECISION": "fp64"})
    assert a.shape == b.shape == (8, N, 3)
    assert np.max(np.abs(a - b)) <= 1.5e-6, (name, np.max(np.abs(a - b)))


def _run_dir(exe, d, prm_path, extra=(), env_extra=None):
    env = dict(os.environ, OMP_NUM_THREADS="1", **(env_extra or {}))
    r = subprocess.run([exe, prm_path, *extra], cwd=d, env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, (exe, r.stdout[-2000:], r.stderr[-2000:])


@pytest.mark.skipif(not (os.path.exists(REF) and os.path.exists(HIP)),
                    reason="reference / drop-in binaries not built here")
@pytest.mark.parametrize("writer,reader", [("hip", "ref"), ("ref", "hip")])
def test_dropin_restart_round_trip(writer, reader):
    """Restart files cross between the backends (restart.c:126-359): a run of 3 steps by
    one, continued from its restart file for 3 more by the other, writes the trajectory of
    an uninterrupted 6-step reference run (positions and the random stream are synced from
    the device on restart steps)."""
    import shutil
    exe = {"hip": HIP, "ref": REF}
    envs = {"hip": {"BDBOX_HIP_PRECISION": "fp64"}, "ref": {}}
    sysm = wl.chain(24, seed=5)
    base = wl.chain_params(hydro="cholesky", rand_seed=11, save_xyz_freq=1, save_rst_freq=3)
    print(f"restart {writer} -> {reader}", flush=True)
    with tempfile.TemporaryDirectory() as t:
        full, short, cont = (os.path.join(t, x) for x in ("full", "short", "cont"))
        for d in (full, short):
            os.makedirs(d)
            wl.write_str(os.path.join(d, "c.str"), sysm)
        wl.write_prm(os.path.join(full, "c.prm"), dict(base, bdsteps=7), os.path.join(full, "c.str"), full)
        _run_dir(REF, full, os.path.join(full, "c.prm"))
        a = read_xyz(os.path.join(full, "out.xyz"))
        wl.write_prm(os.path.join(short, "c.prm"), dict(base, bdsteps=4), os.path.join(short, "c.str"), short)
        _run_dir(exe[writer], short, os.path.join(short, "c.prm"), env_extra=envs[writer])
        shutil.copytree(short, cont)
        wl.write_prm(os.path.join(cont, "c2.prm"), dict(base, bdsteps=7), os.path.join(cont, "c.str"), cont)
        _run_dir(exe[reader], cont, os.path.join(cont, "c2.prm"),
                 ["--restart=" + os.path.join(cont, "out.rst1")], env_extra=envs[reader])
        b = read_xyz(os.path.join(cont, "out.xyz"))
    assert a.shape == b.shape == (7, 24, 3)
    assert np.max(np.abs(a - b)) <= 1.5e-6, np.max(np.abs(a - b))
