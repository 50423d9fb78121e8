"""tay_gpu_space_fetch_words_static (the allocation-free fetch, staging in a
fixed device array) against tay_gpu_space_fetch_words (staging allocated by
the space) through the C-ABI: the same records bit for bit on two calls in a
row, the second for fewer agents and other words; a request beyond the
static array is refused with an error, and a smaller one on the same space
still works."""
import ctypes as C

import numpy as np
import pytest

from tay_amd import agents as A
from tay_amd import api
from tay_amd import workloads as W

pytestmark = pytest.mark.gpu


def fetch(lib, fn, space, first, words, count):
    out = np.zeros((count, words), np.uint32)
    assert fn(space, first, words, out.ctypes.data, count), lib.tay_gpu_last_error().decode()
    return out


def test_static_fetch_matches_allocating_fetch(gpu):
    recs = W.flock_agents(20000, seed=95, clusters=8, sigma=30.0, box=600.0)
    print("case: static fetch, 20000 flocking agents", flush=True)
    with W.Simulation(gpu, W.flock_model(25.0), recs, api.TAY_GPU_Z_GRID) as s:
        gpu.tay_set_fetch_mode(s.state, api.TAY_FETCH_ON_DEMAND)
        s.run(2)
        space = gpu.tay_get_gpu_space(s.state, s.group)
        assert space
        # positions of every agent
        want = fetch(gpu, gpu.tay_gpu_space_fetch_words, space, 2, 3, len(recs))
        got = fetch(gpu, gpu.tay_gpu_space_fetch_words_static, space, 2, 3, len(recs))
        assert np.array_equal(got, want)
        # one more step, then velocities and a count of the first third
        s.run(1)
        n = len(recs) // 3
        want = fetch(gpu, gpu.tay_gpu_space_fetch_words, space, 6, 3, n)
        got = fetch(gpu, gpu.tay_gpu_space_fetch_words_static, space, 6, 3, n)
        assert np.array_equal(got, want)
        gpu.tay_fetch_agents(s.state, None)
        a = s.agents()
        assert np.array_equal(got.view(np.float32), np.ascontiguousarray(a["v"][:n, :3]))


def test_static_fetch_refuses_requests_beyond_its_array(gpu):
    n, words = 2_000_000, 5          # 40 MB of 20-byte records: more than the 32 MiB array
    part = (C.c_float * 4)(1.0, 1.0, 1.0, 1.0)
    space = gpu.tay_gpu_space_create(api.TAY_GPU_KIND_SIMPLE, 4 * words, n, 1, 3, part, 0)
    assert space, gpu.tay_gpu_last_error().decode()
    try:
        rec = np.arange(n * words, dtype=np.uint32)
        assert gpu.tay_gpu_space_push(space, rec.ctypes.data, n)
        out = np.zeros((n, words), np.uint32)
        print("case: static fetch beyond its array", flush=True)
        assert not gpu.tay_gpu_space_fetch_words_static(space, 0, words, out.ctypes.data, n)
        assert "static staging buffer" in gpu.tay_gpu_last_error().decode()
        got = fetch(gpu, gpu.tay_gpu_space_fetch_words_static, space, 1, 2, n)
        assert np.array_equal(got, rec.reshape(n, words)[:, 1:3])
    finally:
        gpu.tay_gpu_space_destroy(space)
