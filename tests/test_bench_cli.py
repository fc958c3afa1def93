"""bench.py on the CPU: the one-GPU contract and the workload builders (no
device work)."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_workload_shapes():
    irvec, hr, kpts, omegas, eta = bench.workload(8, 5)
    assert kpts.shape == (8, 2) and omegas.shape == (5,)
    assert hr.shape == (irvec.shape[0], 36, 36) and irvec.shape[0] == 343
    _, hr2, kp2, om2, eta2 = bench.workload_plane288(5)
    assert kp2.shape == (25, 2) and np.array_equal(om2, [0.0]) and hr2.shape[1] == 144
    assert eta2 < eta
    sets = bench.workload_mixed(30)
    assert [(w["nw"], w["np"]) for w in sets] == [(8, 1), (16, 2), (36, 3), (72, 3)]
    for w in sets:
        assert w["k_index"].shape == (30,) and w["k_index"].max() < w["kpts"].shape[0]
        e = w["eta"] / bench.M.EV2HARTREE
        assert np.all((e >= 1e-4) & (e <= 1e-2))


def test_cpu_sample_is_strided_over_the_grid():
    kpts = np.stack([np.arange(10.0), np.zeros(10)], axis=1)
    om = np.arange(7.0)
    jobs = bench.grid_jobs(kpts, om, 0.5, 6)
    assert len(jobs) == 6
    assert jobs[0] == (0, 0.0, 0.0, 0.0, 0.5) and jobs[-1] == (0, 9.0, 0.0, 6.0, 0.5)
    flops = bench.alg_flops(2, [1, 3])
    assert flops == (56 + 24 + 56 * 3 + 24) * 8
