"""Agent models of tay_amd/agent_src: record and context layouts as numpy
dtypes (packed, matching the C structs byte for byte) and the source texts
handed to ocl_add_source."""
import os

import numpy as np

SRC_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "agent_src")

# Declarations first, then definitions, as the reference passes its model
# sources (test/test_basic.c:37-40).
SOURCE_FILES = ["taystd.h", "test_agent.h", "flocking.h", "sph.h",
                "taystd.c", "test_agent.c", "flocking.c", "sph.c"]


def source_texts():
    """Model source strings, in the order they are added to a state."""
    out = []
    for name in SOURCE_FILES:
        with open(os.path.join(SRC_DIR, name)) as f:
            out.append(f.read())
    return out


_F4 = ("<f4", 4)

TAG = [("part_i", "<u4"), ("cell_agent_i", "<u4")]

# test/agent.h:2-10 layout
AGENT = np.dtype(TAG + [("p",) + _F4, ("v",) + _F4, ("b_buffer",) + _F4, ("b_buffer_count", "<i4"),
                        ("f_buffer",) + _F4, ("result_index", "<i4")])
BOX_AGENT = np.dtype(TAG + [("min",) + _F4, ("max",) + _F4, ("v",) + _F4, ("b_buffer",) + _F4,
                            ("b_buffer_count", "<i4"), ("f_buffer",) + _F4, ("result_index", "<i4")])
PLAIN_AGENT = np.dtype(TAG + [("p",) + _F4, ("v",) + _F4, ("nsum",) + _F4, ("ncount", "<u4"),
                              ("last_nsum",) + _F4, ("last_ncount", "<u4"), ("result_index", "<i4")])
# m_flocking/agent.h:2-13 layout
BOID = np.dtype(TAG + [("p",) + _F4, ("dir",) + _F4, ("color", "<u4"), ("speed", "<f4"),
                       ("separation",) + _F4, ("alignment",) + _F4, ("cohesion",) + _F4,
                       ("cohesion_count", "<i4"), ("separation_count", "<i4")])
FLOCK_AGENT = np.dtype(TAG + [("p",) + _F4, ("v",) + _F4, ("separation",) + _F4, ("alignment",) + _F4,
                              ("cohesion",) + _F4, ("separation_count", "<i4"), ("cohesion_count", "<i4"),
                              ("result_index", "<i4")])
SMOOTH_AGENT = np.dtype(TAG + [("p",) + _F4, ("v",) + _F4, ("separation",) + _F4, ("alignment",) + _F4,
                               ("cohesion",) + _F4, ("weight_sum", "<f4"), ("result_index", "<i4")])
# m_sph/agent.h:2-11 layout
SPH_PARTICLE = np.dtype(TAG + [("p",) + _F4, ("pressure_accum",) + _F4, ("viscosity_accum",) + _F4,
                               ("vh",) + _F4, ("v",) + _F4, ("density", "<f4"), ("pressure", "<f4")])

SEE_CONTEXT = np.dtype([("radii",) + _F4])
ACT_CONTEXT = np.dtype([("min",) + _F4, ("max",) + _F4])
PLAIN_SEE_CONTEXT = np.dtype([("r2", "<f4")])
BOID_SEE_CONTEXT = np.dtype([("r", "<f4"), ("separation_r", "<f4")])
BOID_ACT_CONTEXT = np.dtype([("dummy", "<i4")])
FLOCK_SEE_CONTEXT = np.dtype([("r", "<f4"), ("separation_r", "<f4")])
FLOCK_ACT_CONTEXT = np.dtype([("separation_w", "<f4"), ("alignment_w", "<f4"), ("cohesion_w", "<f4"),
                              ("dt", "<f4"), ("max_speed", "<f4")])
SMOOTH_SEE_CONTEXT = np.dtype([("r2", "<f4"), ("inv_r2", "<f4"), ("h2", "<f4")])
SMOOTH_ACT_CONTEXT = np.dtype([("separation_w", "<f4"), ("alignment_w", "<f4"), ("cohesion_w", "<f4"),
                               ("dt", "<f4"), ("damping", "<f4")])
SPH_CONTEXT = np.dtype([("h", "<f4"), ("dynamic_viscosity", "<f4"), ("dt", "<f4"), ("m", "<f4"),
                        ("surface_tension", "<f4"), ("surface_tension_threshold", "<f4"), ("K", "<f4"),
                        ("density", "<f4"), ("min",) + _F4, ("max",) + _F4, ("h2", "<f4"), ("poly6", "<f4"),
                        ("poly6_gradient", "<f4"), ("poly6_laplacian", "<f4"), ("spiky", "<f4"),
                        ("viscosity", "<f4")])

assert AGENT.itemsize == 80 and BOID.itemsize == 104 and SPH_PARTICLE.itemsize == 96
assert PLAIN_AGENT.itemsize == 84 and FLOCK_AGENT.itemsize == 100 and BOX_AGENT.itemsize == 96
