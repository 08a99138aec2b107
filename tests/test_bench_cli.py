"""The benchmark's contract on the CPU: one GPU only, the workload table, and the
algorithmic bytes it prices the kernels with (no GPU call)."""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import turbchannel_amd as tca

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_is_refused(n):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


@pytest.mark.parametrize("name", sorted(bench.WORKLOADS))
def test_workload_definitions(name):
    scheme, nx, ny, nz, zz, kw, dt = bench.problem(name, tca.FP64)
    collocated = scheme in (tca.SCHEME_AB3, tca.SCHEME_SRK)
    assert len(zz) == (nz if collocated else nz + 1)  # nodes (collocated) or faces (staggered)
    assert zz[0] == 0.0 and zz[-1] == 2.0 and np.all(np.diff(zz) > 0)
    assert nx % 2 == 0 and ny % 2 == 0 and dt > 0
    assert math.isclose(kw["hx"] * nx, 2 * math.pi) and math.isclose(kw["hy"] * ny, math.pi)
    # bulk units: nu = 2 / Re_b, u_tau = Re_tau * nu, forcing u_tau^2 / H
    assert math.isclose(kw["sour_dpdx"], (kw["visc"] * kw["inv_length_scale"]) ** 2)
    if scheme == tca.SCHEME_SRK:
        assert kw["cfr_ubulk"] == 0.0
    else:
        assert kw["cfr_ubulk"] == 1.0


def test_headline_workload_is_the_largest_config():
    """The default line is the collocated BDF3/EX3 DNS at 384 x 384 x 192 nodes in
    FP32-bulk, the largest single-GPU configuration; the staggered figures ride
    along as extra workloads."""
    assert bench.DEFAULT_WORKLOAD == "ab3_384x384x192"
    assert bench.DEFAULT_PRECISION[bench.DEFAULT_WORKLOAD] == "fp32"
    scheme, nx, ny, nz, zz, kw, dt = bench.problem(bench.DEFAULT_WORKLOAD, tca.FP32_BULK)
    assert (scheme, nx, ny, nz, dt) == (tca.SCHEME_AB3, 384, 384, 192, 5e-4)
    assert math.isclose(2.0 / kw["visc"], 13750.0) and kw["stab_type"] == 1
    assert bench.EXTRA_WORKLOADS == [("vv_384x384x192", "fp32"), ("vv_192x160x128", "fp64")]
    assert bench.SHIPPED_THREADS[tca.SCHEME_AB3] == 2 and bench.SHIPPED_THREADS[tca.SCHEME_VV] == 1


def test_algorithmic_reals_per_step():
    """Reals per grid point and step, each array element counted once: 40 for the
    staggered DNS, 43 for the collocated solver with StabType 1 (its explicit term
    and projection also read p_n, the projection writes p_n+1)."""
    r = bench.REALS_PER_POINT
    assert r["flux"] == 21 and r["project"] == 7 and r["zsolve"] == 2
    assert r["fft_y"] == r["ifft_y"] == 2
    assert sum(r.values()) == 40
    a = bench.REALS_PER_POINT_AB3
    assert a["flux"] == 22 and a["project"] == 9 and sum(a.values()) == 43


def test_baseline_threads_use_the_job_share():
    cores = bench.host_cores()
    assert cores["allowed"] >= 1
    t = bench.baseline_threads()
    assert 1 <= t <= min(cores["allowed"], 16) or t == int(os.environ.get("OMP_NUM_THREADS", "0"))
