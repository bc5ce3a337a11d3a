"""Link check of the drop-in backend (CPU): a program that references every entry point and
global of the backend interface the drop-in replaces (src/diff_alg/diff_tensor.h:76-253,
src/potentials/calc_func.h:34-106, src/myblas.h:37-51, src/diff_alg/newton.h:33-37) links
against the reference's driver objects and bd_box_amd/csrc/bdbox_dropin.c.  Built where the
reference sources are mounted (make -C oracle dropin_symbols); skipped elsewhere."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(not os.path.isdir("/root/reference/src"), reason="reference sources not mounted")
def test_every_backend_symbol_links():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "dropin_symbols"], check=True)
    exe = os.path.join(ROOT, "oracle", "_ref", "dropin_symbols")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "79 symbols linked"
