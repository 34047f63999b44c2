"""Host-side descriptors passed to the C-ABI (include/angelo_amd.h)."""
import ctypes

import numpy as np

from . import _native as nat


def grid_desc(n_levels, n_features, log2_hashmap_size, base_resolution, per_level_scale,
              res_override=None):
    """Level table by tcnn's rule (am_grid_init); host-only, no device needed."""
    g = nat.GridDesc()
    ro = None
    if res_override is not None:
        arr = (ctypes.c_uint32 * n_levels)(*[int(r) for r in res_override])
        ro = ctypes.cast(arr, ctypes.c_void_p)
    nat.check(nat.load().am_grid_init(ctypes.byref(g), int(n_levels), int(n_features),
                                      int(log2_hashmap_size), int(base_resolution),
                                      float(per_level_scale), ro), "am_grid_init")
    return g


def grid_arrays(g):
    L = g.n_levels
    return dict(scale=np.array(g.scale[:L], np.float32), res=np.array(g.res[:L], np.uint32),
                offset=np.array(g.offset[:L], np.uint32), size=np.array(g.size[:L], np.uint32),
                hashed=np.array([(g.hashed_mask >> l) & 1 for l in range(L)], np.uint32),
                n_entries=int(g.n_entries))


def mlp_desc(n_in, n_out, n_hidden_layers, n_neurons=64, hidden_act="relu", output_act="none",
             softplus_beta=100.0, precision="fp32"):
    m = nat.MlpDesc()
    m.precision = nat.PREC[precision]
    m.n_in, m.n_out, m.n_hidden_layers, m.n_neurons = int(n_in), int(n_out), int(n_hidden_layers), int(n_neurons)
    m.hidden_act = nat.ACT[hidden_act]
    m.output_act = nat.ACT[output_act]
    m.softplus_beta = float(softplus_beta)
    return m


def mlp_n_params(m):
    n = nat.load().am_mlp_n_params(ctypes.byref(m))
    if n < 0:
        nat.check(-1, "am_mlp_n_params")
    return int(n)
