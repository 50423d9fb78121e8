"""The drop-in boundary driven from C: tests/c_host/flock_host.c is a plain C
program that includes include/tay.h, links libtay_amd.so and makes only the
reference's public calls (tay/tay.h:87-117), in the order of the reference's
own harness (test/test_basic.c:34-84).  On the CPU it must link, start and
report the library's errors through tay_get_error; on the GPU its agents
after a run must equal, bit for bit, the oracle's on the same input."""
import os
import subprocess

import numpy as np
import pytest

from tay_amd import agents as A
from tay_amd import api
from tay_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "tests", "c_host", "_build", "flock_host")
AGENT_SRC = os.path.join(ROOT, "tay_amd", "agent_src")


def run_host(tmp_path, space, steps, records, timeout=120):
    src = tmp_path / "in.bin"
    dst = tmp_path / "out.bin"
    records.tofile(src)
    proc = subprocess.run([HOST, AGENT_SRC, space, str(steps), str(src), str(dst)], stdout=subprocess.PIPE,
                          stderr=subprocess.PIPE, text=True, timeout=timeout)
    out = np.fromfile(dst, dtype=A.FLOCK_AGENT) if proc.returncode == 0 else None
    return proc, out


def test_host_program_is_built():
    assert os.access(HOST, os.X_OK), "build tests/c_host/flock_host.c first (__graft_entry__.build())"


def test_cpu_space_is_reported_through_the_c_api(tmp_path):
    """A CPU space type reaches tay_run and comes back as the reference's
    TAY_ERROR_NOT_IMPLEMENTED with its message, with or without a GPU."""
    proc, _ = run_host(tmp_path, "cpu_grid", 1, W.flock_agents(500, seed=3))
    assert proc.returncode == 2, proc.stdout + proc.stderr
    assert "error %d" % api.TAY_ERROR_NOT_IMPLEMENTED in proc.stderr
    assert "CPU spaces are not part of this library" in proc.stderr


@pytest.mark.skipif(api.TayLib(api.LIBRARY).tay_gpu_available() == 1, reason="needs a host without GPU")
def test_missing_device_is_reported_through_the_c_api(tmp_path):
    proc, _ = run_host(tmp_path, "z_grid", 1, W.flock_agents(500, seed=3))
    assert proc.returncode == 2, proc.stdout + proc.stderr
    assert "error %d" % api.TAY_ERROR_OCL in proc.stderr
    assert "device not found" in proc.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("space,n", [("z_grid", 20000), ("simple", 4000)])
def test_c_host_run_bit_exact_vs_oracle(tmp_path, oracle, models, space, n):
    """Clustered flocking agents, 5 steps through the C host; the oracle runs
    the same model (W.flock_model(25.0): the contexts the C host hard-codes)
    on the same records, in the GPU Z-grid's CPU statement or in CpuSimple."""
    recs = W.flock_agents(n, seed=21, clusters=8, sigma=40.0, box=600.0)
    print("case: c host %s n=%d steps=5" % (space, len(recs)), flush=True)
    proc, got = run_host(tmp_path, space, 5, recs)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    oracle_space = api.TAY_GPU_Z_GRID if space == "z_grid" else api.TAY_CPU_SIMPLE
    with W.Simulation(oracle, W.flock_model(25.0), recs, oracle_space, funcs=models) as o:
        o.run(5)
        want = o.agents()
    assert got.dtype == want.dtype and len(got) == len(want)
    for f in want.dtype.names:
        x, y = got[f], want[f]
        if x.dtype.kind == "f":
            x, y = x.view(np.uint32), y.view(np.uint32)
        bad = np.nonzero((x != y).reshape(len(got), -1).any(axis=1))[0]
        assert len(bad) == 0, "field %s differs at %d agents, first %s" % (f, len(bad), bad[:5])
