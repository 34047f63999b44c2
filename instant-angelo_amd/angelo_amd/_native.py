"""ctypes binding of libangelo_amd.so (the C-ABI declared in include/angelo_amd.h).

The library is loaded from angelo_amd/lib/.  There is no fallback: if it is
missing, every GPU entry point raises.  Build it with
`python instant-angelo_amd/build.py` (or `__graft_entry__.build()`).
"""
import ctypes
import os

import torch

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libangelo_amd.so")
_lib = None

AM_MAX_LEVELS = 32
ACT = {"none": 0, "relu": 1, "softplus": 2, "sigmoid": 3}
PREC = {"fp32": 0, "mixed": 1}


class GridDesc(ctypes.Structure):
    _fields_ = [
        ("n_levels", ctypes.c_int32),
        ("n_features", ctypes.c_int32),
        ("log2_hashmap_size", ctypes.c_int32),
        ("base_resolution", ctypes.c_int32),
        ("per_level_scale", ctypes.c_float),
        ("n_entries", ctypes.c_uint32),
        ("hashed_mask", ctypes.c_uint32),
        ("scale", ctypes.c_float * AM_MAX_LEVELS),
        ("res", ctypes.c_uint32 * AM_MAX_LEVELS),
        ("offset", ctypes.c_uint32 * AM_MAX_LEVELS),
        ("size", ctypes.c_uint32 * AM_MAX_LEVELS),
    ]


class MlpDesc(ctypes.Structure):
    _fields_ = [
        ("n_in", ctypes.c_int32),
        ("n_out", ctypes.c_int32),
        ("n_hidden_layers", ctypes.c_int32),
        ("n_neurons", ctypes.c_int32),
        ("hidden_act", ctypes.c_int32),
        ("output_act", ctypes.c_int32),
        ("softplus_beta", ctypes.c_float),
        ("precision", ctypes.c_int32),
    ]


GRAD_TYPE = {"finite_difference": 0, "analytic": 1}


class SdfDesc(ctypes.Structure):
    _fields_ = [
        ("radius", ctypes.c_float),
        ("fd_eps", ctypes.c_float),
        ("with_grad", ctypes.c_int32),
        ("with_laplace", ctypes.c_int32),
        ("recontract", ctypes.c_int32),
        ("active_levels", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("grad_type", ctypes.c_int32),
    ]


class CompositeDesc(ctypes.Structure):
    _fields_ = [("cos_anneal_ratio", ctypes.c_float), ("kahan", ctypes.c_int32)]


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_F = ctypes.c_float
_SZ = ctypes.c_size_t

# name -> (restype, argtypes)
_SIGS = {
    "am_last_error": (ctypes.c_char_p, []),
    "am_last_workspace_need": (_SZ, []),
    "am_sdf_bwd_check": (ctypes.c_int, []),
    "am_march_bg_count": (ctypes.c_int, [_P, _I64, _P, _P, _P, _P, _I32, _F, _F, _I32, _P, _P, _P, _SZ, _P]),
    "am_march_bg_fill": (ctypes.c_int, [_P, _I64, _P, _P, _P, _P, _I32, _F, _F, _I32, _P, _P, _P, _P, _P]),
    "am_march_bg_visible": (ctypes.c_int, [_P, _P, _P, _P, _I64, _F, _P, _P, _P, _SZ, _P]),
    "am_march_staged_workspace_size": (ctypes.c_size_t, [_I64, _I32]),
    "am_morton_keys": (ctypes.c_int, [_P, _I64, _F, _P, _P]),
    "am_morton_order_workspace_size": (ctypes.c_size_t, [_I64, _I32]),
    "am_morton_order": (ctypes.c_int, [_P, _I64, _F, _I32, _P, _P, _P, _SZ, _P]),
    "am_march_count_staged": (ctypes.c_int, [_P, _I64, _P, _P, _P, _P, _P, _I32, _F, _I32, _P, _P, _P, _SZ, _P]),
    "am_march_fill_staged": (ctypes.c_int, [_I64, _I32, _P, _P, _SZ, _P, _P, _P, _P]),
    "am_sample_rays": (ctypes.c_int, [_P, _P, _P, _P, _I32, _I32, _I32, _I32, _I32, _P, _I32, _P, _P, _P, _I64, _P, _P, _P, _P, _P]),
    "am_composite_density_fwd": (ctypes.c_int, [_P, _P, _P, _P, _P, _I64, _I64, _P, _P, _P, _P, _P, _P]),
    "am_composite_density_bwd": (ctypes.c_int, [_P, _P, _P, _P, _P, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "am_version": (ctypes.c_int, []),
    "am_timer_enable": (ctypes.c_int, [_I32]),
    "am_timer_read": (ctypes.c_int, [ctypes.c_char_p, _P, _P]),
    "am_timer_units": (ctypes.c_int, [ctypes.c_char_p, _P]),
    "am_set_flag": (ctypes.c_int, [ctypes.c_char_p, _I32]),
    "am_grid_init": (ctypes.c_int, [_P, _I32, _I32, _I32, _I32, _F, _P]),
    "am_hashgrid_fwd": (ctypes.c_int, [_P, _P, _I64, _P, _I32, _P, _P]),
    "am_hashgrid_bwd_workspace_size": (_SZ, [_P, _I64, _I32]),
    "am_hashgrid_bwd": (ctypes.c_int, [_P, _P, _I64, _P, _P, _I32, _P, _P, _P, _SZ, _P]),
    "am_hashgrid_fwd_ordered": (ctypes.c_int, [_P, _P, _P, _I64, _P, _I32, _P, _P]),
    "am_hashgrid_bwd_ordered": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P, _I32, _P, _P, _P, _SZ, _P]),
    "am_mlp_n_params": (_I64, [_P]),
    "am_mlp_fwd": (ctypes.c_int, [_P, _P, _I64, _P, _P, _P]),
    "am_mlp_bwd_workspace_size": (_SZ, [_P, _I64]),
    "am_mlp_bwd": (ctypes.c_int, [_P, _P, _I64, _P, _P, _P, _P, _P, _SZ, _P]),
    "am_sh_fwd": (ctypes.c_int, [_P, _I64, _I32, _P, _P]),
    "am_ray_aabb": (ctypes.c_int, [_P, _I64, _P, _P, _P, _P]),
    "am_march_workspace_size": (_SZ, [_I64]),
    "am_march_count": (ctypes.c_int, [_P, _I64, _P, _P, _P, _P, _P, _I32, _F, _I32, _P, _P, _P, _SZ, _P]),
    "am_march_fill": (ctypes.c_int, [_P, _I64, _P, _P, _P, _P, _P, _I32, _F, _I32, _P, _P, _P, _P, _P]),
    "am_sample_points": (ctypes.c_int, [_P, _P, _P, _P, _I64, _P, _P, _P, _P]),
    "am_contract_unbounded": (ctypes.c_int, [_P, _I64, _F, _P, _P]),
    "am_contract_unbounded_bwd": (ctypes.c_int, [_P, _P, _I64, _F, _P, _P]),
    "am_sdf_fwd": (ctypes.c_int, [_P] * 5 + [_I64] + [_P] * 9 + [_P]),
    "am_sdf_store_floats": (_I64, [_P, _P, _P, _I64]),
    "am_sdf_bwd_workspace_size": (_SZ, [_P, _P, _P, _I64]),
    "am_sdf_bwd": (ctypes.c_int, [_P] * 5 + [_I64] + [_P] * 11 + [_P, _SZ, _P]),
    "am_sdf_fwd_ordered": (ctypes.c_int, [_P] * 6 + [_I64] + [_P] * 9 + [_P]),
    "am_sdf_bwd_ordered": (ctypes.c_int, [_P] * 6 + [_I64] + [_P] * 11 + [_P, _SZ, _P]),
    "am_radiance_fwd": (ctypes.c_int, [_P, _I32, _I32, _P, _I32, _P, _P, _P, _I64, _P, _P, _P]),
    "am_radiance_bwd_workspace_size": (_SZ, [_P, _I64]),
    "am_radiance_bwd": (ctypes.c_int, [_P, _I32, _I32, _P, _I32, _P, _P, _P, _I64, _P, _P, _P, _P, _P, _P, _SZ, _P]),
    "am_composite_fwd": (ctypes.c_int, [_P] * 10 + [_I64, _I64] + [_P] * 7 + [_P]),
    "am_composite_bwd_workspace_size": (_SZ, [_I64, _I64]),
    "am_composite_bwd": (ctypes.c_int, [_P] * 10 + [_I64, _I64] + [_P] * 12 + [_P, _SZ, _P]),
    "am_composite_bwd_noalloc": (ctypes.c_int, [_P] * 10 + [_I64, _I64] + [_P] * 12 + [_P]),
    "am_occ_update_workspace_size": (_SZ, [_I32, _I64]),
    "am_occ_update": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _F, _P, _P, _I64, _I32, _F, _F, _P, _P, _P, _SZ, _P]),
}


class NativeError(RuntimeError):
    pass


def lib_path():
    return _LIB_PATH


def load():
    """Loads (once) and returns the library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise NativeError(
            f"angelo_amd native library not built ({_LIB_PATH}); run "
            "`python instant-angelo_amd/build.py`")
    lib = ctypes.CDLL(_LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return list(_SIGS.keys())


def check(rc, what):
    if rc != 0:
        msg = load().am_last_error().decode(errors="replace")
        raise NativeError(f"{what} failed (status {rc}): {msg}")


def ptr(t):
    """Device pointer of a tensor (None -> NULL).  Every buffer the kernels
    read or write must be device memory: a host tensor's address handed to
    a kernel faults the GPU, so it is refused here."""
    if t is None:
        return None
    if not t.is_cuda:
        raise NativeError("angelo_amd kernels take device tensors (got a CPU tensor of shape "
                          f"{tuple(t.shape)}); move the module or data to the GPU first")
    return ctypes.c_void_p(t.data_ptr())


def sptr(s):
    return ctypes.byref(s)


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def require_cuda(*tensors):
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise NativeError("angelo_amd kernels need device tensors (got a CPU tensor); "
                              "the HIP path has no CPU fallback")


def workspace(nbytes, device):
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


_HELD = {}


def held_workspace(kind, nbytes, device, headroom=1.25):
    """A grow-only device buffer of at least nbytes for one kind of call,
    kept per (kind, device, stream): the steps of a training loop, whose
    workspace sizes follow their sample counts, then allocate nothing once
    the largest has been seen (a fresh allocation per call costs a
    hipMalloc whenever the size outgrows the allocator's cached blocks).
    Calls of one kind on one stream run in stream order, so reusing the
    buffer between them is safe.  A graph-captured call must not use one
    (a later growth would free the address the graph holds)."""
    if torch.device(device).type != "cuda":
        raise NativeError("angelo_amd kernels take device tensors (a workspace was asked for "
                          f"on {device}); the HIP path has no CPU fallback")
    key = (kind, device.index, torch.cuda.current_stream(device).cuda_stream)
    buf = _HELD.get(key)
    n = max(int(nbytes), 1)
    if buf is None or buf.numel() < n:
        _HELD[key] = None
        buf = _HELD[key] = workspace(int(n * headroom), device)
    return buf


AM_ERR_WORKSPACE = -3


class GrowingWorkspace:
    """Grow-only device workspace for a call whose need depends on the data
    (am_sdf_bwd: its table-gradient pairs).  The call runs with what is held;
    when that is short it fails before writing any output, reports its need
    (am_last_workspace_need) and runs again with that plus `headroom`.  One
    buffer per (device, stream): the call's kernels use it asynchronously."""

    def __init__(self, initial_bytes=1 << 20, headroom=1.25):
        self.initial = int(initial_bytes)
        self.headroom = float(headroom)
        self.grows = 0
        self._bufs = {}

    def held_bytes(self):
        return sum(b.numel() for b in self._bufs.values())

    def run(self, device, call, what, min_bytes=0):
        """call(ws_ptr, ws_bytes) -> status.  min_bytes: hold at least this
        many bytes before the call (e.g. a worst-case bound that spares the
        call its host read)."""
        key = (device.index, torch.cuda.current_stream(device).cuda_stream)
        buf = self._bufs.get(key)
        if buf is None or buf.numel() < min_bytes:
            self._bufs[key] = None
            buf = self._bufs[key] = workspace(max(self.initial, int(min_bytes)), device)
        for _ in range(4):   # head short, then pairs short, then fits
            rc = call(ptr(buf), buf.numel())
            if rc != AM_ERR_WORKSPACE:
                check(rc, what)
                return
            need = int(load().am_last_workspace_need())
            if need <= buf.numel():
                check(rc, what)
            self._bufs[key] = None
            buf = self._bufs[key] = workspace(int(need * self.headroom), device)
            self.grows += 1
        check(rc, what)


def sdf_bwd_check():
    """Raises if any SDF backward so far had inconsistent pair passes (the
    library reads those flags without waiting; this waits for all)."""
    check(load().am_sdf_bwd_check(), "am_sdf_bwd_check")


def timer_enable(on=True):
    check(load().am_timer_enable(1 if on else 0), "am_timer_enable")


def timer_read(name):
    """(total ms, launch count) of a timed kernel since timer_enable."""
    ms = ctypes.c_double(0.0)
    cnt = ctypes.c_int64(0)
    check(load().am_timer_read(name.encode(), ctypes.byref(ms), ctypes.byref(cnt)), "am_timer_read")
    return ms.value, cnt.value


def set_flag(name, value):
    """Selects an implementation variant (am_set_flag), e.g.
    set_flag("table_prepartition", 0)."""
    check(load().am_set_flag(name.encode(), int(value)), "am_set_flag")


def timer_units(name):
    """Work units (e.g. "table_pairs") recorded since timer_enable."""
    u = ctypes.c_int64(0)
    check(load().am_timer_units(name.encode(), ctypes.byref(u)), "am_timer_units")
    return u.value
