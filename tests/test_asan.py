"""The C oracle under AddressSanitizer (host code only): oracle/liboracle_asan.so
(the same restatement built with -fsanitize=address) runs the evaluation, the
stages and the c(k) relayout on odd and even grids in a child process with
libasan preloaded; any invalid access aborts it with an ASan report."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

SCRIPT = r"""
import sys
sys.path[:0] = [%r, %r]
import numpy as np
from mdft import ck
from oracle import oracle as O
for n, L, mmax in (((9, 5, 15), (9.0, 5.0, 13.0), 3), ((8, 6, 4), (8.0, 6.0, 4.0), 5), ((5, 5, 5), (5.0, 5.0, 5.0), 0)):
    t = ck.synthetic_ck(mmax, 256, 0.05, seed=3, n0=0.0333)
    o = O.Oracle(n, L, mmax, 2, 2.4777, 0.0333)
    o.set_ck(t)
    rng = np.random.default_rng(1)
    xi = 1.0 + 0.2 * rng.uniform(-1, 1, o.N * o.no)
    v = rng.uniform(-2, 4, o.N * o.no)
    for st in (1, 2, 3, 4):
        o.excess_stage(xi, st)
    o.energy_and_gradient(xi, v)
    o.get_ck()
print("asan run done")
"""


def test_oracle_under_address_sanitizer():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle_asan.so"])
    libasan = subprocess.check_output(["gcc", "-print-file-name=libasan.so"], text=True).strip()
    env = dict(os.environ, LD_PRELOAD=libasan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               MDFT_ORACLE_LIB=os.path.join(ROOT, "oracle", "liboracle_asan.so"), OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-c", SCRIPT % (ROOT, os.path.join(ROOT, "mdft-dev_general_amd"))],
                       env=env, capture_output=True, text=True, timeout=600)
    assert "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0 and "asan run done" in r.stdout, (r.returncode, r.stderr[-4000:])
