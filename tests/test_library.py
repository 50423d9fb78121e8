"""CPU-side checks of libtay_amd.so: it loads and exports every function
include/*.h declares; model sources compile for gfx950 through hipRTC with no
GPU present (the reference's own OpenCL-C model strings included, when the
reference tree is available); the error paths of the reference's state
machine report the reference's error codes."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from tay_amd import agents as A
from tay_amd import api
from tay_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
REFERENCE = "/root/reference"


def declared_functions():
    names = set()
    for header in ("tay.h", "tay_gpu.h", "tay_ext.h", "tay_dflock.h"):
        text = open(os.path.join(INCLUDE, header)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
        for m in re.finditer(r"\b([a-z_][a-z0-9_]*)\s*\([^;{]*\)\s*;", text):
            names.add(m.group(1))
    return names - {"if", "while", "for", "sizeof"}


def test_library_exports_every_declared_function(tay):
    names = declared_functions()
    assert "tay_run" in names and "tay_gpu_space_sort" in names and len(names) > 40
    missing = [n for n in sorted(names) if not hasattr(tay.lib, n)]
    assert not missing, missing
    assert set(api.exported_names()) <= names


def _shape(kind, func, dtype, seen=None, seen_kind=api.TAY_GPU_KIND_Z_GRID, skip_self=1, point=1):
    seen = seen or dtype
    return api.TayGpuPassShape(kind, func.encode(), dtype.itemsize, seen.itemsize, point, point, seen_kind, 3,
                               skip_self)


def compile_check(tay, sources, shapes):
    srcs = [s.encode() for s in sources]
    arr = (C.c_char_p * len(srcs))(*srcs)
    sh = (api.TayGpuPassShape * len(shapes))(*shapes)
    log = C.create_string_buffer(1 << 16)
    size = C.c_ulonglong()
    ok = tay.tay_gpu_compile_check(arr, len(srcs), sh, len(shapes), log, len(log), C.byref(size))
    return ok, log.value.decode(errors="replace"), size.value


def test_models_compile_for_gfx950_without_gpu(tay):
    shapes = [_shape(0, "plain_see", A.PLAIN_AGENT), _shape(1, "plain_act", A.PLAIN_AGENT),
              _shape(0, "agent_see", A.AGENT), _shape(1, "agent_act", A.AGENT),
              _shape(0, "flock_see", A.FLOCK_AGENT), _shape(1, "flock_act", A.FLOCK_AGENT),
              _shape(0, "boid_see", A.BOID), _shape(1, "boid_act", A.BOID),
              _shape(0, "sph_density", A.SPH_PARTICLE, skip_self=0), _shape(1, "sph_pressure", A.SPH_PARTICLE),
              _shape(0, "sph_forces", A.SPH_PARTICLE), _shape(1, "sph_integrate", A.SPH_PARTICLE),
              _shape(0, "box_agent_see", A.BOX_AGENT, seen_kind=api.TAY_GPU_KIND_SIMPLE, point=0),
              _shape(1, "box_agent_act", A.BOX_AGENT)]
    ok, log, size = compile_check(tay, A.source_texts(), shapes)
    assert ok, log
    assert size > 1000


def test_split_passes_compile_without_gpu(tay):
    """Split see passes (2, 4 and 8 lanes per seer) of every model with clear
    and combine functions, over 1 to 3 dimensions and point and box groups;
    a split pass naming a function that is not a C identifier is refused."""
    def split(func, dtype, lanes, seen=None, dims=3, point=1, seen_point=None):
        clear, combine = W.SPLIT_FUNCS[func]
        seen = seen or dtype
        sp = point if seen_point is None else seen_point
        return api.TayGpuPassShape(0, func.encode(), dtype.itemsize, seen.itemsize, point, sp,
                                   api.TAY_GPU_KIND_Z_GRID, dims, 1 if seen is dtype else 0, 0, 0, lanes,
                                   clear.encode(), combine.encode())
    shapes = []
    for lanes in (2, 4, 8):
        shapes += [split("flock_see", A.FLOCK_AGENT, lanes), split("plain_see", A.PLAIN_AGENT, lanes),
                   split("box_agent_see", A.BOX_AGENT, lanes, point=0),
                   split("point_box_see", A.AGENT, lanes, seen=A.BOX_AGENT, seen_point=0),
                   split("box_point_see", A.BOX_AGENT, lanes, seen=A.AGENT, point=0, seen_point=1)]
        shapes += [split("agent_see", A.AGENT, lanes, dims=d) for d in (1, 2, 3)]
        shapes += [split("sph_density", A.SPH_PARTICLE, lanes), split("sph_forces", A.SPH_PARTICLE, lanes)]
    ok, log, size = compile_check(tay, A.source_texts(), shapes)
    assert ok, log
    assert size > 1000
    bad = split("plain_see", A.PLAIN_AGENT, 2)
    bad.split_combine = b"not a name"
    ok, log, _ = compile_check(tay, A.source_texts(), [bad])
    assert not ok


def test_batched_programs_compile_without_gpu(tay):
    P = api.TayGpuBatchPass
    srcs = [s.encode() for s in A.source_texts()]
    arr = (C.c_char_p * len(srcs))(*srcs)
    for dtype, passes in ((A.FLOCK_AGENT, [P(0, b"flock_see", 0, A.FLOCK_SEE_CONTEXT.itemsize),
                                           P(1, b"flock_act", 0, A.FLOCK_ACT_CONTEXT.itemsize)]),
                          (A.SPH_PARTICLE, [P(0, b"sph_density", 1, A.SPH_CONTEXT.itemsize),
                                            P(1, b"sph_pressure", 0, A.SPH_CONTEXT.itemsize),
                                            P(0, b"sph_forces", 0, A.SPH_CONTEXT.itemsize),
                                            P(1, b"sph_integrate", 0, A.SPH_CONTEXT.itemsize)])):
        ps = (P * len(passes))(*passes)
        log = C.create_string_buffer(1 << 16)
        size = C.c_ulonglong()
        ok = tay.tay_gpu_batch_compile_check(arr, len(srcs), ps, len(passes), dtype.itemsize, 3, log, len(log),
                                             C.byref(size))
        assert ok, log.value.decode(errors="replace") + tay.tay_gpu_last_error().decode()
        assert size.value > 1000
    bad = (P * 1)(P(0, b"not a name", 0, 0))
    assert not tay.tay_gpu_batch_compile_check(arr, len(srcs), bad, 1, A.FLOCK_AGENT.itemsize, 3, None, 0, None)
    assert b"identifier" in tay.tay_gpu_last_error()


def test_compile_error_reports_the_log(tay):
    broken = A.source_texts() + ["void plain_see_broken(global PlainAgent *a) { a->no_such_field = 1; }"]
    ok, log, _ = compile_check(tay, broken, [_shape(1, "plain_act", A.PLAIN_AGENT)])
    assert not ok
    assert "no_such_field" in log
    ok, log, _ = compile_check(tay, A.source_texts(), [_shape(0, "no_such_function", A.PLAIN_AGENT)])
    assert not ok and "no_such_function" in log


def test_agent_size_mismatch_is_a_compile_error(tay):
    """A model record larger than the group's agent size is caught at compile
    time rather than read out of bounds."""
    small = np.dtype([("w", "<u4", 6)])
    ok, log, _ = compile_check(tay, A.source_texts(), [_shape(0, "flock_see", small)])
    assert not ok and "larger than the group" in log


def _c_string_literals(path):
    """The string constants of one of the reference's generated
    *_ocl.c files (agentgen.py output), as the C compiler would join them."""
    text = open(path).read()
    out = {}
    for m in re.finditer(r'const char \*(\w+)\s*=\s*"((?:[^"\\]|\\.)*)";', text, flags=re.S):
        body = m.group(2).replace("\\\n", "")
        out[m.group(1)] = bytes(body, "utf-8").decode("unicode_escape")
    return out


REFERENCE_MODELS = [
    ("test", [(0, "agent_see", 80), (1, "agent_act", 80), (0, "box_agent_see", 96), (1, "box_agent_act", 96)]),
    ("m_flocking", [(0, "agent_see", 104), (1, "agent_act", 104)]),
    ("m_sph", [(0, "sph_particle_density", 96), (1, "sph_particle_pressure", 96), (0, "sph_force_terms", 96),
               (1, "sph_particle_leapfrog", 96)]),
]


@pytest.mark.skipif(not os.path.isdir(REFERENCE), reason="reference tree not present")
@pytest.mark.parametrize("project,passes", REFERENCE_MODELS)
def test_reference_model_strings_compile(tay, project, passes):
    """The reference's own OpenCL-C model sources, handed to ocl_add_source
    in the order the reference uses (test/test_basic.c:37-40), compile
    through this library's prelude for gfx950."""
    lits = _c_string_literals(os.path.join(REFERENCE, project, "agent_ocl.c"))
    sources = [lits["agent_ocl_h"], lits["taystd_ocl_h"], lits["agent_ocl_c"], lits["taystd_ocl_c"]]
    shapes = []
    for kind, func, size in passes:
        point = 0 if func.startswith("box_") else 1
        seen_kind = api.TAY_GPU_KIND_SIMPLE if not point else api.TAY_GPU_KIND_Z_GRID
        shapes.append(api.TayGpuPassShape(kind, func.encode(), size, size, point, point, seen_kind, 3, 1))
    ok, log, _ = compile_check(tay, sources, shapes)
    assert ok, log


# ---------------------------------------------------------------- errors

no_gpu = pytest.mark.skipif(api.TayLib(api.LIBRARY).tay_gpu_available() == 1, reason="needs a host without GPU")


def _state_with_group(tay, space, is_point=1, size=A.PLAIN_AGENT.itemsize):
    st = tay.tay_create_state()
    g = tay.tay_add_group(st, size, 10, is_point)
    tay.tay_configure_space(st, g, space, 3, api.float4(10, 10, 10, 10), 1)
    return st, g


def test_run_before_start_is_a_state_error_or_device_error(tay):
    st, g = _state_with_group(tay, api.TAY_GPU_Z_GRID)
    tay.tay_add_act(st, g, None, b"plain_act", None, 0)
    assert tay.tay_run(st, 1) == 0
    assert tay.tay_get_error(st) in (api.TAY_ERROR_STATE_STATUS, api.TAY_ERROR_OCL)
    tay.tay_destroy_state(st)


def test_non_point_group_in_a_grid(tay):
    """The reference's grids take points only (tay/state_compile.c:25-27),
    and so does TAY_CPU_Z_GRID here; the GPU Z-grid takes box agents too
    (tests/test_gpu_combos.py)."""
    st, g = _state_with_group(tay, api.TAY_CPU_Z_GRID, is_point=0, size=A.BOX_AGENT.itemsize)
    tay.tay_simulation_start(st)
    assert tay.tay_run(st, 1) == 0
    assert tay.tay_get_error(st) == api.TAY_ERROR_POINT_NONPOINT_MISMATCH
    tay.tay_destroy_state(st)


def test_cpu_space_types_are_not_part_of_the_library(tay):
    st, g = _state_with_group(tay, api.TAY_CPU_GRID)
    tay.tay_simulation_start(st)
    assert tay.tay_run(st, 1) == 0
    assert tay.tay_get_error(st) == api.TAY_ERROR_NOT_IMPLEMENTED
    assert "CPU spaces" in tay.error_message(st)
    tay.tay_destroy_state(st)


def test_pass_without_function_name(tay):
    st, g = _state_with_group(tay, api.TAY_GPU_Z_GRID)
    tay.tay_add_act(st, g, None, b"", None, 0)
    tay.tay_simulation_start(st)
    assert tay.tay_run(st, 1) == 0
    assert tay.tay_get_error(st) == api.TAY_ERROR_OCL
    tay.tay_destroy_state(st)


def test_context_without_size(tay):
    st, g = _state_with_group(tay, api.TAY_GPU_Z_GRID)
    ctx = np.zeros(1, A.ACT_CONTEXT)
    tay.tay_add_act(st, g, None, b"plain_act", ctx.ctypes.data, 0)
    tay.tay_simulation_start(st)
    assert tay.tay_run(st, 1) == 0
    assert tay.tay_get_error(st) == api.TAY_ERROR_OCL
    assert "context" in tay.error_message(st)
    tay.tay_destroy_state(st)


def test_capacity_and_group_slots(tay):
    st = tay.tay_create_state()
    g = tay.tay_add_group(st, A.PLAIN_AGENT.itemsize, 2, 1)
    for _ in range(2):
        assert tay.tay_get_available_agent(st, g)
        tay.tay_commit_available_agent(st, g)
    assert not tay.tay_get_available_agent(st, g)
    assert tay.tay_get_error(st) == api.TAY_ERROR_CAPACITY
    for _ in range(7):
        assert tay.tay_add_group(st, 32, 1, 1)
    assert not tay.tay_add_group(st, 32, 1, 1)
    assert tay.tay_get_error(st) == api.TAY_ERROR_GROUP_INDEX_OUT_OF_RANGE
    tay.tay_destroy_state(st)


def test_host_storage_that_cannot_be_allocated_is_an_error(tay):
    """2^31 agents of 64 KiB (128 TiB) cannot be allocated: the group is
    refused with TAY_ERROR_CAPACITY instead of handing out a null storage."""
    st = tay.tay_create_state()
    assert not tay.tay_add_group(st, 1 << 16, 1 << 31, 1)
    assert tay.tay_get_error(st) == api.TAY_ERROR_CAPACITY
    assert "cannot allocate host storage" in tay.error_message(st)
    assert tay.tay_add_group(st, 32, 4, 1)
    tay.tay_destroy_state(st)


@no_gpu
def test_no_device_is_reported(tay):
    st, g = _state_with_group(tay, api.TAY_GPU_Z_GRID)
    tay.tay_add_act(st, g, None, b"plain_act", None, 0)
    tay.tay_simulation_start(st)
    assert tay.tay_run(st, 1) == 0
    assert tay.tay_get_error(st) == api.TAY_ERROR_OCL
    assert "device not found" in tay.error_message(st)
    assert tay.tay_switch_to_ocl(st) == 0
    tay.tay_destroy_state(st)


def test_column_index_range_is_checked_at_creation(tay):
    """Columns are indexed in 32 bits: a space whose capacity x agent words
    reaches 2^32 is refused before any device memory is touched."""
    part = (C.c_float * 4)(25.0, 25.0, 25.0, 25.0)
    words = A.FLOCK_AGENT.itemsize // 4
    too_many = (1 << 32) // words + 1
    assert not tay.tay_gpu_space_create(api.TAY_GPU_KIND_Z_GRID, A.FLOCK_AGENT.itemsize, too_many, 1, 3, part, 1 << 20)
    assert b"exceeds 2^32 column words" in tay.tay_gpu_last_error()


def test_batch_slot_range_is_checked_at_creation(tay):
    """A batch whose padded slots exceed 2^30 (32-bit byte offsets in a
    column) is refused before anything is compiled or allocated."""
    P = api.TayGpuBatchPass
    srcs = [s.encode() for s in A.source_texts()]
    arr = (C.c_char_p * len(srcs))(*srcs)
    passes = (P * 1)(P(1, b"flock_act", 0, A.FLOCK_ACT_CONTEXT.itemsize))
    counts = (C.c_uint * 2)(1 << 29, (1 << 29) + 1)
    tables = (C.c_ulonglong * 2)(1 << 20, 1 << 20)
    assert not tay.tay_gpu_batch_create(arr, len(srcs), passes, 1, A.FLOCK_AGENT.itemsize, 3, 2, counts, tables, 0,
                                        None, 0)
    assert b"2^30 slots" in tay.tay_gpu_last_error()


@no_gpu
def test_page_locked_storage_needs_a_gpu(tay):
    """Without a GPU the page-locked allocator returns NULL and a group's
    storage is ordinary zeroed memory."""
    assert not tay.tay_gpu_host_alloc(1 << 20)
    st = tay.tay_create_state()
    g = tay.tay_add_group(st, 32, 8, 1)
    p = tay.tay_get_available_agent(st, g)
    assert p and bytes((C.c_char * 256).from_address(p)) == bytes(256)
    tay.tay_destroy_state(st)
