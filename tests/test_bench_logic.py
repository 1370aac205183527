"""Host side of bench.py: the workload it times, the profile-derived HBM traffic it reports and its
refusal of more than one GPU (no GPU needed)."""

from __future__ import annotations

import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_workload_is_the_128_cube_b4_pion_config():
    rp, p_i, p_f, F = bench.workload(L=24)  # the builder at a small size; the bench uses L = 128
    assert tuple(int(v) for v in p_i[:3]) == (24, 24, 24)
    assert int(p_i[5]) == 4 and int(p_i[6]) == 24 ** 3
    assert np.allclose(p_f[3:6], 0.15)                      # dx
    assert p_f[7] == pytest.approx(bench.COURANT * 0.15)    # dt = courant * dx
    assert p_f[8] == 1.0 and p_f[9] == 1.0 and p_f[10] == 4.0  # m, kappa, B
    Fm = F.reshape(4, -1)
    assert np.allclose(np.einsum("an,an->n", Fm, Fm), 1.0)   # on S^3 after the noise
    assert np.allclose(Fm[:, 0], [1.0, 0.0, 0.0, 0.0])       # vacuum in the boundary band


@pytest.mark.parametrize("form,s", [("fp32", 4), ("fp64", 8)])
def test_profiled_traffic_is_read_from_the_committed_summary(form, s):
    t = bench.profiled_traffic("stage4", form)
    algorithmic = 7 * 128 ** 3 * 4 * s  # F, V, l1, l2, l3 read; F', V' written
    assert t is not None and 1.0 <= t / algorithmic <= 1.5
    line = [ln for ln in open(os.path.join(ROOT, "profiles", f"stage_kernels_pion_128_b4_{form}.md"))
            if ln.startswith(f"traffic_bytes_per_launch stage4 {form}:")]
    assert len(line) == 1 and float(line[0].split(":")[1]) == t
    assert bench.profiled_traffic("stage9", form) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]
