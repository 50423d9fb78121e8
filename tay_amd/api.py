"""ctypes mirror of the Tay C API (include/tay.h, include/tay_ext.h,
include/tay_gpu.h, include/tay_dflock.h).

TayLib binds one shared library that exports the reference API: the GPU
library libtay_amd.so, or (in tests) the CPU oracle, which exports the same
names.  Method names and argument meaning follow tay/tay.h:87-117; errors are
the same sticky TayError codes.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(PKG, "libtay_amd.so")

# TaySpaceType (include/tay.h)
TAY_SPACE_NONE = 0
TAY_CPU_SIMPLE = 1
TAY_CPU_KD_TREE = 2
TAY_CPU_AABB_TREE = 3
TAY_CPU_GRID = 4
TAY_CPU_Z_GRID = 5
TAY_GPU_Z_GRID = 6
TAY_GPU_SIMPLE = 7

# TayError
TAY_ERROR_NONE = 0
TAY_ERROR_POINT_NONPOINT_MISMATCH = 1
TAY_ERROR_SPACE_INDEX_OUT_OF_RANGE = 2
TAY_ERROR_GROUP_INDEX_OUT_OF_RANGE = 3
TAY_ERROR_STATE_STATUS = 4
TAY_ERROR_NOT_IMPLEMENTED = 5
TAY_ERROR_OCL = 6
TAY_ERROR_CAPACITY = 7

# TayFetchMode (include/tay_ext.h)
TAY_FETCH_EVERY_RUN = 0
TAY_FETCH_ON_DEMAND = 1
TAY_SEE_ORDER_REFERENCE = 0
TAY_SEE_ORDER_BALANCED = 1

# TayGpuSpaceKind / TayGpuPassType (include/tay_gpu.h)
TAY_GPU_KIND_Z_GRID = 1
TAY_GPU_KIND_SIMPLE = 2
TAY_GPU_PASS_SEE = 0
TAY_GPU_PASS_ACT = 1
TAY_GPU_TILE_ORDER_COST = 0
TAY_GPU_TILE_ORDER_STORAGE = 1


class float4(C.Structure):
    _pack_ = 1
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]

    def __init__(self, x=0.0, y=0.0, z=0.0, w=0.0):
        super().__init__(x, y, z, w)


class TayGpuPassShape(C.Structure):
    _fields_ = [("type", C.c_int), ("func_name", C.c_char_p),
                ("seer_agent_size", C.c_uint), ("seen_agent_size", C.c_uint),
                ("seer_is_point", C.c_int), ("seen_is_point", C.c_int),
                ("seen_kind", C.c_int), ("dims", C.c_int), ("skip_self", C.c_int), ("context_size", C.c_uint),
                ("see_order", C.c_int), ("split_lanes", C.c_uint), ("split_clear", C.c_char_p),
                ("split_combine", C.c_char_p)]


class TayGpuBatchPass(C.Structure):
    _fields_ = [("type", C.c_int), ("func_name", C.c_char_p), ("self_see", C.c_int), ("context_size", C.c_uint)]


class TayDflockParams(C.Structure):
    _fields_ = [("radius", C.c_float), ("h2", C.c_float), ("dt", C.c_float), ("damping", C.c_float),
                ("cell_size", C.c_float)]


_vp = C.c_void_p
_u = C.c_uint
_i = C.c_int

_CORE = {
    "tay_create_state": (_vp, []),
    "tay_destroy_state": (None, [_vp]),
    "tay_switch_to_ocl": (_i, [_vp]),
    "tay_switch_to_host": (_i, [_vp]),
    "tay_get_error": (_i, [_vp]),
    "tay_add_group": (_vp, [_vp, _u, _u, _i]),
    "tay_configure_space": (None, [_vp, _vp, _i, _i, float4, _i]),
    "tay_fix_space_box": (None, [_vp, _vp, float4, float4]),
    "tay_add_see": (None, [_vp, _vp, _vp, _vp, C.c_char_p, float4, _i, _vp, _u]),
    "tay_add_act": (None, [_vp, _vp, _vp, C.c_char_p, _vp, _u]),
    "tay_clear_all_agents": (None, [_vp, _vp]),
    "tay_get_available_agent": (_vp, [_vp, _vp]),
    "tay_commit_available_agent": (None, [_vp, _vp]),
    "tay_get_agent": (_vp, [_vp, _vp, _i]),
    "tay_simulation_start": (None, [_vp]),
    "tay_run": (_i, [_vp, _i]),
    "tay_simulation_end": (None, [_vp]),
    "tay_get_ms_per_step_for_last_run": (C.c_double, [_vp]),
    "ocl_add_source": (None, [_vp, C.c_char_p]),
    "tay_threads_start": (None, [_u, _u]),
    "tay_threads_stop": (None, []),
    "tay_get_number_of_threads": (_u, []),
}

# include/tay_ext.h and include/tay_gpu.h (GPU library only)
_GPU = {
    "tay_get_error_message": (C.c_char_p, [_vp]),
    "tay_commit_available_agents": (None, [_vp, _vp, _u]),
    "tay_set_pair_counting": (None, [_vp, _i]),
    "tay_get_pair_count": (_i, [_vp, _u, C.POINTER(C.c_ulonglong)]),
    "tay_set_pass_timing": (None, [_vp, _i]),
    "tay_get_pass_time": (_i, [_vp, _u, C.POINTER(C.c_double), C.POINTER(C.c_uint)]),
    "tay_get_gpu_space": (_vp, [_vp, _vp]),
    "tay_set_tile_width": (_i, [_vp, _u]),
    "tay_set_lane_walk": (_i, [_vp, _u]),
    "tay_set_tile_seers": (_i, [_vp, _u]),
    "tay_set_see_order": (_i, [_vp, _i]),
    "tay_set_see_split": (_i, [_vp, _u, _u, _vp, C.c_char_p, _vp, C.c_char_p]),
    "tay_set_order_bits": (_i, [_vp, _vp, _i]),
    "tay_set_fetch_mode": (None, [_vp, _i]),
    "tay_fetch_agents": (_i, [_vp, _vp]),
    "tay_fetch_agent_fields": (_u, [_vp, _vp, _u, _u, _vp]),
    "tay_gpu_available": (_i, []),
    "tay_gpu_last_error": (C.c_char_p, []),
    "tay_gpu_synchronize": (_i, []),
    "tay_gpu_host_alloc": (_vp, [C.c_size_t]),
    "tay_gpu_host_free": (None, [_vp]),
    "tay_gpu_copy_probe": (_i, [C.c_ulonglong, C.POINTER(C.c_double)]),
    "tay_gpu_sqrt_selftest": (_i, [C.c_ulonglong, C.c_ulonglong, C.POINTER(C.c_ulonglong), C.POINTER(C.c_uint)]),
    "tay_gpu_space_create": (_vp, [_i, _u, _u, _i, _i, C.POINTER(C.c_float), C.c_ulonglong]),
    "tay_gpu_space_destroy": (None, [_vp]),
    "tay_gpu_space_fix_box": (_i, [_vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "tay_gpu_space_unfix_box": (_i, [_vp]),
    "tay_gpu_space_set_order_bits": (_i, [_vp, _i]),
    "tay_gpu_space_push": (_i, [_vp, _vp, _u]),
    "tay_gpu_space_fetch": (_i, [_vp, _vp, _u]),
    "tay_gpu_space_fetch_words": (_i, [_vp, _u, _u, _vp, _u]),
    "tay_gpu_space_fetch_words_static": (_i, [_vp, _u, _u, _vp, _u]),
    "tay_gpu_space_count": (_u, [_vp]),
    "tay_gpu_space_sort": (_i, [_vp]),
    "tay_gpu_space_grid_info": (_i, [_vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_int),
                                     C.POINTER(C.c_ulonglong)]),
    "tay_gpu_compile_check": (_i, [C.POINTER(C.c_char_p), _u, C.POINTER(TayGpuPassShape), _u, C.c_char_p, _u,
                                   C.POINTER(C.c_ulonglong)]),
    "tay_gpu_program_create": (_vp, [C.POINTER(C.c_char_p), _u, _vp, _u, _i, C.c_char_p, _u]),
    "tay_gpu_program_destroy": (None, [_vp]),
    "tay_gpu_program_set_context": (_i, [_vp, _u, _vp]),
    "tay_gpu_program_run_pass": (_i, [_vp, _u]),
    "tay_gpu_program_pair_count": (_i, [_vp, _u, C.POINTER(C.c_ulonglong)]),
    "tay_gpu_program_reset_pair_counts": (_i, [_vp]),
    "tay_gpu_program_enable_timing": (_i, [_vp, _i]),
    "tay_gpu_program_pass_time": (_i, [_vp, _u, C.POINTER(C.c_double), C.POINTER(C.c_uint)]),
    "tay_gpu_program_set_tile_width": (_i, [_vp, _u]),
    "tay_gpu_program_set_lane_walk": (_i, [_vp, _u]),
    "tay_gpu_program_set_tile_seers": (_i, [_vp, _u]),
    "tay_gpu_batch_create": (_vp, [C.POINTER(C.c_char_p), _u, C.POINTER(TayGpuBatchPass), _u, _u, _i, _u,
                                   C.POINTER(C.c_uint), C.POINTER(C.c_ulonglong), _i, C.c_char_p, _u]),
    "tay_gpu_batch_destroy": (None, [_vp]),
    "tay_gpu_batch_compile_check": (_i, [C.POINTER(C.c_char_p), _u, C.POINTER(TayGpuBatchPass), _u, _u, _i,
                                         C.c_char_p, _u, C.POINTER(C.c_ulonglong)]),
    "tay_gpu_batch_set_part_sizes": (_i, [_vp, C.POINTER(C.c_float)]),
    "tay_gpu_batch_get_part_sizes": (_i, [_vp, C.POINTER(C.c_float)]),
    "tay_gpu_batch_density_part_sizes": (_i, [_vp, C.c_float, C.c_float, C.c_float, C.POINTER(C.c_float)]),
    "tay_gpu_batch_set_radii": (_i, [_vp, _u, C.POINTER(C.c_float)]),
    "tay_gpu_batch_set_contexts": (_i, [_vp, _u, _vp]),
    "tay_gpu_batch_set_tile_width": (_i, [_vp, _u]),
    "tay_gpu_batch_set_tile_order": (_i, [_vp, _i]),
    "tay_gpu_batch_set_lane_walk": (_i, [_vp, _u]),
    "tay_gpu_batch_set_order_bits": (_i, [_vp, _i]),
    "tay_gpu_batch_push": (_i, [_vp, _vp]),
    "tay_gpu_batch_fetch": (_i, [_vp, _vp]),
    "tay_gpu_batch_run": (_i, [_vp, _u, _i, C.POINTER(C.c_double)]),
    "tay_gpu_batch_enable_timing": (_i, [_vp, _i]),
    "tay_gpu_batch_pass_time": (_i, [_vp, _u, C.POINTER(C.c_double), C.POINTER(C.c_uint)]),
    "tay_gpu_batch_pair_count": (_i, [_vp, _u, C.POINTER(C.c_ulonglong)]),
    "tay_gpu_batch_reset_pair_counts": (_i, [_vp]),
    "tay_dflock_create": (_vp, [_u, C.POINTER(TayDflockParams), _u, C.c_ulonglong]),
    "tay_dflock_destroy": (None, [_vp]),
    "tay_dflock_forward": (_i, [_vp, _vp, _vp, _vp, _u, _vp, _vp]),
    "tay_dflock_backward": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "tay_dflock_times": (_i, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "tay_dflock_max_neighbours": (_i, [_vp, C.POINTER(C.c_uint)]),
    "tay_dflock_pair_count": (_i, [_vp, C.POINTER(C.c_ulonglong)]),
    "tay_gpu_batch_autotune": (_i, [_vp, C.POINTER(C.c_float), _u, C.POINTER(C.c_uint), _u, C.POINTER(C.c_uint), _u,
                                    _u, C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_uint),
                                    C.POINTER(C.c_uint)]),
    "tay_gpu_batch_autotune_density": (_i, [_vp, C.POINTER(C.c_float), _u, C.c_float, C.c_float,
                                            C.POINTER(C.c_uint), _u, C.POINTER(C.c_uint), _u, _u,
                                            C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                            C.POINTER(C.c_uint), C.POINTER(C.c_uint)]),
}

# test-only inspection calls of the oracle (oracle/oracle.h)
_ORACLE = {
    "oracle_get_pair_counts": (None, [_vp, _u, C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)]),
    "oracle_get_grid_info": (None, [_vp, _vp, C.POINTER(float4), C.POINTER(float4), C.POINTER(C.c_int),
                                    C.POINTER(C.c_ulonglong)]),
    "oracle_morton_key": (_u, [_i, _i, _i, _i]),
    "oracle_morton_decode": (None, [_u, _i, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "oracle_get_see_seconds": (C.c_double, [_vp]),
    "oracle_limit_seers": (None, [_vp, _u, _u]),
}


def exported_names():
    """Every symbol include/*.h declares for libtay_amd.so."""
    return sorted(list(_CORE) + list(_GPU))


class TayLib:
    """One loaded library exporting the Tay API."""

    def __init__(self, path=LIBRARY):
        if not os.path.exists(path):
            raise FileNotFoundError("%s is not built (run tay_amd.build.build())" % path)
        self.path = path
        self.lib = C.CDLL(path)
        for table in (_CORE, _GPU, _ORACLE):
            for name, (res, args) in table.items():
                fn = getattr(self.lib, name, None)
                if fn is None:
                    continue
                fn.restype = res
                fn.argtypes = args
                setattr(self, name, fn)

    def has(self, name):
        return hasattr(self, name)

    # -- helpers over the raw calls --------------------------------------

    def add_agents(self, state, group, records):
        """Append a numpy array of agent records (their dtype's itemsize must
        be the group's agent size) to the group's storage."""
        records = np.ascontiguousarray(records)
        n = len(records)
        if n == 0:
            return
        ptr = self.tay_get_available_agent(state, group)
        if not ptr:
            raise RuntimeError("group is full")
        C.memmove(ptr, records.ctypes.data, records.nbytes)
        if self.has("tay_commit_available_agents"):
            self.tay_commit_available_agents(state, group, n)
        else:
            for _ in range(n):
                self.tay_commit_available_agent(state, group)

    def get_agents(self, state, group, count, dtype):
        """Copy of the group's first count agents in current storage order."""
        dtype = np.dtype(dtype)
        if count == 0:
            return np.zeros(0, dtype)
        ptr = self.tay_get_agent(state, group, 0)
        buf = (C.c_char * (count * dtype.itemsize)).from_address(ptr)
        return np.frombuffer(buf, dtype=dtype, count=count).copy()

    def error_message(self, state):
        if self.has("tay_get_error_message"):
            return self.tay_get_error_message(state).decode(errors="replace")
        return ""

    def pair_counts(self, state, pass_index):
        """(broad, narrow) pair counts of a pass in the last run."""
        if self.has("oracle_get_pair_counts"):
            b, n = C.c_ulonglong(), C.c_ulonglong()
            self.oracle_get_pair_counts(state, pass_index, C.byref(b), C.byref(n))
            return b.value, n.value
        n = C.c_ulonglong()
        if not self.tay_get_pair_count(state, pass_index, C.byref(n)):
            raise RuntimeError(self.error_message(state))
        return None, n.value

    def grid_info(self, state, group):
        """(origin[4], cell_sizes[4], counts[4], cells) of a grid after its
        last sort."""
        if self.has("oracle_get_grid_info"):
            o, s = float4(), float4()
            c = (C.c_int * 4)()
            n = C.c_ulonglong()
            self.oracle_get_grid_info(state, group, C.byref(o), C.byref(s), c, C.byref(n))
            return (np.array([o.x, o.y, o.z, o.w], np.float32), np.array([s.x, s.y, s.z, s.w], np.float32),
                    np.array(list(c), np.int32), n.value)
        space = self.tay_get_gpu_space(state, group)
        o = (C.c_float * 4)()
        s = (C.c_float * 4)()
        c = (C.c_int * 4)()
        n = C.c_ulonglong()
        if not self.tay_gpu_space_grid_info(space, o, s, c, C.byref(n)):
            raise RuntimeError(self.tay_gpu_last_error().decode())
        return (np.array(list(o), np.float32), np.array(list(s), np.float32), np.array(list(c), np.int32), n.value)

    def pass_time(self, state, pass_index):
        ms, launches = C.c_double(), C.c_uint()
        if not self.tay_get_pass_time(state, pass_index, C.byref(ms), C.byref(launches)):
            raise RuntimeError(self.error_message(state))
        return ms.value, launches.value


def load(path=LIBRARY):
    return TayLib(path)
