"""ctypes binding of libsslab_hip.so (C ABI declared in include/sslab_hip.h).

torch is imported first on purpose: the library's DT_NEEDED libamdhip64.so.7 then
resolves to the HIP runtime torch already loaded, so device pointers and stream
handles from torch are valid inside the library (one runtime per process).

There is no CPU fallback: if the library is missing or no GPU is visible, every
entry point raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import torch  # noqa: F401  (must precede loading the library, see module docstring)

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libsslab_hip.so"

_c_int = ctypes.c_int
_c_ll = ctypes.c_longlong
_c_d = ctypes.c_double
_vp = ctypes.c_void_p

# name -> argtypes (all return int status)
_SIGS = {
    "sslab_abi_version": [],
    "sslab_dgemm_strided_batched": [_c_int, _c_int, _c_int, _c_int, _c_int, _c_d,
                                    _vp, _c_int, _c_ll, _vp, _c_int, _c_ll, _c_d,
                                    _vp, _c_int, _c_ll, _c_int, _vp],
    "sslab_dgemm_grouped": [_vp, _c_int, _vp],
    "sslab_sgemm_strided_batched": [_c_int, _c_int, _c_int, _c_int, _c_int, ctypes.c_float,
                                    _vp, _c_int, _c_ll, _vp, _c_int, _c_ll, ctypes.c_float,
                                    _vp, _c_int, _c_ll, _c_int, _vp],
    "sslab_gemm_set_policy": [_vp, _c_int],
    "sslab_rb_set_backward_policies": [_vp, _c_int],
    "sslab_gemm_get_policy": [_vp, _c_int],
    "sslab_dgetrf_batched": [_c_int, _vp, _c_int, _c_ll, _vp, _c_ll, _vp, _c_int, _vp],
    "sslab_dgetrs_batched": [_c_int, _c_int, _vp, _c_int, _c_ll, _vp, _c_ll, _vp, _c_int,
                             _c_ll, _vp, _c_int, _c_ll, _vp, _c_int, _vp],
    "sslab_inv_sizes": [_c_int, _c_int, _vp, _vp],
    "sslab_dinv_batched": [_c_int, _vp, _c_int, _c_ll, _vp, _c_int, _c_ll, _c_int, _vp, _vp,
                           _vp, _c_d, _vp, _vp],
    "sslab_dgetrs_trans_batched": [_c_int, _c_int, _vp, _c_int, _c_ll, _vp, _c_ll, _vp,
                                   _c_int, _c_ll, _vp, _c_int, _c_ll, _vp, _vp, _c_int, _vp],
    "sslab_dgeqrf_batched": [_c_int, _c_int, _vp, _c_int, _c_ll, _vp, _c_int, _c_ll, _vp,
                             _c_ll, _c_int, _vp],
    "sslab_dtrsm_upper_batched": [_c_int, _c_int, _vp, _c_int, _c_ll, _vp, _c_int, _c_ll,
                                  _c_int, _vp],
    "sslab_add_diag_blocks": [_c_int, _c_int, _c_int, _vp, _vp, _c_int, _c_ll, _vp],
    "sslab_rb_sizes": [_c_int, _c_int, _vp, _vp, _vp],
    "sslab_rb_info_count": [_c_int],
    "sslab_rb_factor": [_c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "sslab_rb_solve": [_c_int, _c_int, _c_int, _vp, _vp, _c_int, _vp, _c_int, _vp],
    "sslab_rb_solve_ex": [_c_int, _c_int, _c_int, _vp, _vp, _c_int, _vp, _c_int, _c_int, _vp],
    "sslab_rb_forward_windowed": [_c_int, _c_int, _c_int, _vp, _vp, _c_int, _vp, _c_int, _vp, _vp,
                                  _vp],
    "sslab_rb_chain_status": [_vp],
    "sslab_rb_solve_multi": [_c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "sslab_cast_f64_to_f32": [_c_ll, _vp, _vp, _vp],
    "sslab_rb_factor_f32": [_c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "sslab_rb_solve_f32": [_c_int, _c_int, _c_int, _vp, _vp, _c_int, _vp, _c_int, _vp],
    "sslab_rb_residual": [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _c_int, _vp, _c_int,
                          _vp, _c_int, _vp],
    "sslab_rb_solve_mixed": [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                             _c_int, _vp],
    "sslab_rb_solve_mixed_noalloc": [_c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                                     _c_int, _vp],
    "sslab_thomas_sizes": [_c_int, _c_int, _vp, _vp, _vp],
    "sslab_thomas_factor": [_c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "sslab_thomas_solve": [_c_int, _c_int, _c_int, _vp, _vp, _c_int, _vp, _c_int, _vp],
    "sslab_hps_leaf_sizes": [_c_int, _c_int, _vp, _vp],
    "sslab_hps_leaf_dtn": [_c_int, _c_int] + [_vp] * 16 + [_vp],
    "sslab_spectral2d_sizes": [_c_int, _c_int, _c_int, _vp, _vp],
    "sslab_spectral2d_slab_maps": [_c_int, _c_int, _c_int] + [_vp] * 13 + [_vp],
    "sslab_hps3_leaf_sizes": [_c_int, _c_int, _vp, _vp],
    "sslab_hps3_leaf_dtn": [_c_int, _c_int] + [_vp] * 19 + [_vp],
    "sslab_hps_merge_sizes": [_c_int, _c_int, _c_int, _vp, _vp],
    "sslab_hps_merge": [_c_int, _c_int, _c_int, _c_int, _c_int, _vp, _c_ll, _vp, _c_ll,
                        _vp, _vp, _vp, _vp, _vp, _c_ll, _vp, _vp, _vp, _vp],
    "sslab_hps_merge_ex": [_c_int, _c_int, _c_int, _c_int, _c_int, _vp, _c_ll, _vp, _c_ll,
                           _vp, _vp, _vp, _vp, _vp, _c_ll, _vp, _vp, _vp, _vp, _c_ll, _vp],
    "sslab_hps_pair_solve": [_c_int, _c_int, _c_int, _c_int, _vp, _c_ll, _vp, _vp, _vp, _vp,
                             _vp, _vp, _vp, _vp, _vp],
    "sslab_oms_rhs": [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _c_int, _c_int, _c_int, _vp,
                      _c_int, _vp],
    "sslab_oms_rhs_rb0": [_c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                          _c_int, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp],
    "sslab_oms_fold_create": [_c_int, _c_int, _c_int, _vp, _c_int, _vp, _vp],
    "sslab_oms_fold_apply": [_vp, _vp, _c_int, _c_int, _vp, _c_int, _vp],
    "sslab_oms_fold_enable_f32": [_vp, _vp],
    "sslab_oms_fold_solve_mixed": [_vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _vp, _vp, _vp, _vp,
                                   _c_int, _vp],
    "sslab_oms_fold_bytes": [_vp],
    "sslab_oms_fold_groups": [_vp],
    "sslab_oms_fold_destroy": [_vp],
    "sslab_col_dots": [_c_int, _c_int, _c_int, _vp, _c_ll, _c_int, _vp, _c_int, _vp, _c_int,
                       _vp],
    "sslab_col_axpy": [_c_int, _c_int, _c_int, ctypes.c_double, _vp, _c_ll, _c_int, _vp,
                       _c_int, _vp, _c_int, _vp],
    "sslab_col_scale": [_c_int, _c_int, _vp, _c_int, _vp, _vp, _c_int, _vp],
    "sslab_gather2": [_c_int, _c_int, _c_int, _vp, _c_int, _c_ll, _vp, _c_int, _c_ll,
                      _vp, _vp, _vp, _vp, _vp, _c_int, _c_ll, _vp],
}

# entry points that do not return a status
_RESTYPES = {"sslab_oms_fold_bytes": ctypes.c_longlong, "sslab_oms_fold_destroy": None}

EXPORTED = tuple(_SIGS) + ("sslab_last_error",)

_lib = None


class SslabError(RuntimeError):
    """A libsslab_hip.so call returned a non-zero status."""


def load(path: str | os.PathLike | None = None):
    """Load (once) and return the ctypes library object.  Raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise SslabError(
            f"{p} not found: build it with `python -m sslablu_amd.build` "
            "(there is no CPU fallback for the HIP path)")
    lib = ctypes.CDLL(str(p))
    for name, args in _SIGS.items():
        fn = getattr(lib, name, None)
        if fn is None:  # stale build: calling this entry point raises (see call())
            continue
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, _c_int)
    lib.sslab_last_error.argtypes = []
    lib.sslab_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def call(name: str, *args) -> int:
    """Call entry point `name`; raise SslabError on a non-zero status."""
    lib = load()
    fn = getattr(lib, name, None)
    if fn is None:
        raise SslabError(f"{name} is not exported by {LIB_PATH}: rebuild with "
                         "`python -m sslablu_amd.build`")
    st = fn(*args)
    if st != 0:
        msg = lib.sslab_last_error().decode(errors="replace")
        raise SslabError(f"{name} failed (status {st}): {msg}")
    return st


def require_gpu():
    """Raise unless a ROCm GPU is visible to torch (no CPU fallback exists)."""
    if not torch.cuda.is_available():
        raise SslabError("no ROCm GPU visible: the slab solver runs only on the GPU")


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)
_cur_device = getattr(torch._C, "_cuda_getDevice", None)


def stream_handle(stream=None) -> int:
    """hipStream_t of `stream` (or torch's current stream) as an int.  The current stream
    is read without building a torch Stream object (every small launch pays this)."""
    if stream is not None:
        return int(stream.cuda_stream)
    if _raw_stream is not None and _cur_device is not None:
        return int(_raw_stream(_cur_device()))
    return int(torch.cuda.current_stream().cuda_stream)


def ptr(t: torch.Tensor) -> int:
    return int(t.data_ptr())


def sizes(kind: str, nb: int, m: int):
    """(factor_doubles, work_doubles, work_ints) for kind in {'rb', 'thomas'}."""
    f = ctypes.c_longlong()
    w = ctypes.c_longlong()
    i = ctypes.c_longlong()
    call(f"sslab_{kind}_sizes", nb, m, ctypes.addressof(f), ctypes.addressof(w),
         ctypes.addressof(i))
    return f.value, w.value, i.value


def sizes2(name: str, *args):
    """(work_doubles, work_ints) from a two-output size query `name`(*args, &w, &i)."""
    w = ctypes.c_longlong()
    i = ctypes.c_longlong()
    call(name, *args, ctypes.addressof(w), ctypes.addressof(i))
    return w.value, i.value
