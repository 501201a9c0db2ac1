"""The oracle's OpenMP loops write disjoint entries and its inner products use
a fixed blocked order, so its results do not depend on the thread count:
three config-2 SIMPLE iterations run with 1 and with 4 threads must agree bit
for bit (fields and linear-solver histories)."""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent

SNIPPET = r"""
import hashlib, sys
sys.path.insert(0, {tests!r}); sys.path.insert(0, {root!r})
import numpy as np, cases, oracle_lib as orc
mesh = cases.cavity3d_mesh(10, jitter=0.2, split=0.1, seed=11)
arrays, _, _ = mesh.prepare()
om = orc.oracle_mesh(arrays.arrays())
ub, pb = cases.cavity_bcs(two_d=False)
s = orc.OracleSimple(om, cases.c2_config(), ub, pb)
s.start()
s.iterate(3)
h = hashlib.sha256()
for a in s.fields():
    h.update(np.ascontiguousarray(a).tobytes())
for which in range(4):
    h.update(s.history(which).tobytes())
print(h.hexdigest())
"""


def run(threads: int) -> str:
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    code = SNIPPET.format(tests=str(HERE), root=str(HERE.parent))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    return out.stdout.strip().splitlines()[-1]


def test_oracle_is_thread_count_independent():
    assert run(1) == run(4)
