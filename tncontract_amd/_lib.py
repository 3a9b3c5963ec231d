"""ctypes binding of ``libtncontract_amd.so`` (C ABI in include/tncontract_amd.h).

PyTorch-ROCm provides device buffers and the current stream only; every
arithmetic operation goes through the HIP library.  torch is imported BEFORE
the library is opened so that the library binds to the HIP runtime torch
already loaded (both carry the soname libamdhip64.so.7): one runtime per
process.  There is no fallback: if the library or a GPU is missing, the calls
raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtncontract_amd.so")

TN_OK = 0
TN_ERR_INVALID = -1
TN_ERR_HIP = -2
TN_ERR_NOCONV = -3
TN_ERR_UNSUPPORTED = -4
TN_ERR_WORKSPACE = -5
TN_MAX_MODES = 8
TN_F64, TN_F32 = 0, 1
PREC = {"native": 0, "fp64": 0, "fp32": 1, "bf16x3": 2}


class LinAlgError(np.linalg.LinAlgError):
    pass


class tn_modes(ctypes.Structure):
    _fields_ = [("nmodes", ctypes.c_int32),
                ("dims", ctypes.c_int64 * TN_MAX_MODES),
                ("strides", ctypes.c_int64 * TN_MAX_MODES)]


def modes(dims, strides):
    dims = list(dims)
    strides = list(strides)
    if len(dims) > TN_MAX_MODES:
        raise NotImplementedError("more than %d modes in one group" % TN_MAX_MODES)
    m = tn_modes()
    m.nmodes = len(dims)
    for i, (d, s) in enumerate(zip(dims, strides)):
        m.dims[i] = int(d)
        m.strides[i] = int(s)
    return m


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libtncontract_amd.so is not built; run `python -m tncontract_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, i64, i32, dbl, sz = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double, ctypes.c_size_t
    mp = ctypes.POINTER(tn_modes)
    sig = {
        "tn_version": (i32, []),
        "tn_last_error": (ctypes.c_char_p, []),
        "tn_last_error_code": (i32, []),
        "tn_device_count": (i32, [ctypes.POINTER(ctypes.c_int)]),
        "tn_permute": (i32, [vp, mp, vp, i32, vp]),
        "tn_convert": (i32, [vp, i32, vp, i32, i64, vp]),
        "tn_scale": (i32, [vp, i32, i64, dbl, vp]),
        "tn_sqrt": (i32, [vp, i32, i64, vp]),
        "tn_axpby": (i32, [vp, vp, vp, i32, i64, dbl, dbl, vp]),
        "tn_diag_embed": (i32, [vp, i64, vp, i32, vp]),
        "tn_diag_inverse_workspace": (sz, [i64]),
        "tn_diag_inverse": (i32, [vp, i64, i64, i64, vp, i32, vp, vp, sz, vp]),
        "tn_sumsq_workspace": (sz, [i64]),
        "tn_sumsq": (i32, [vp, i32, i64, vp, vp, sz, vp]),
        "tn_count_nonfinite": (i32, [vp, i32, i64, vp, vp]),
        "tn_trace": (i32, [vp, mp, i64, i64, vp, i32, vp]),
        "tn_normalize_workspace": (sz, [i64]),
        "tn_normalize": (i32, [vp, i32, i64, vp, vp, sz, vp]),
        "tn_contract": (i32, [vp, mp, mp, vp, mp, mp, vp, mp, mp, i32, i32, dbl, dbl, vp]),
        "tn_contract_workspace": (sz, [i64, i64, i64, i32]),
        "tn_contract_ws": (i32, [vp, mp, mp, vp, mp, mp, vp, mp, mp, i32, i32, dbl, dbl, vp, sz, vp]),
        "tn_qr_workspace": (sz, [i64, i64, i32]),
        "tn_qr": (i32, [vp, mp, mp, i64, i64, vp, vp, i32, vp, sz, vp]),
        "tn_cholqr_workspace": (sz, [i64, i64]),
        "tn_cholqr": (i32, [vp, mp, mp, i64, i64, vp, vp, dbl, ctypes.POINTER(ctypes.c_double),
                            vp, sz, vp]),
        "tn_prof_enable": (i32, [i32]),
        "tn_prof_collect": (i32, [i32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64),
                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), i32]),
        "tn_svd_workspace": (sz, [i64, i64, i32]),
        "tn_svd": (i32, [vp, mp, mp, i64, i64, vp, vp, vp, i32, dbl, i32,
                         ctypes.POINTER(ctypes.c_int), vp, sz, vp]),
        "tn_svd_noalloc": (i32, [vp, mp, mp, i64, i64, vp, vp, vp, i32, dbl, i32,
                                 ctypes.POINTER(ctypes.c_int), vp]),
        "tn_eigh_psd_workspace": (sz, [i64]),
        "tn_eigh_psd": (i32, [vp, mp, mp, i64, vp, vp, i32, dbl, i32,
                              ctypes.POINTER(ctypes.c_int), vp, sz, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc == TN_OK:
        return
    msg = lib().tn_last_error().decode(errors="replace")
    if rc == TN_ERR_INVALID or rc == TN_ERR_WORKSPACE:
        raise ValueError(msg)
    if rc == TN_ERR_NOCONV:
        raise LinAlgError(msg)
    if rc == TN_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)


PROF_TAGS = {"gemm_f64": 0, "gemm_f32": 1, "gemm_bf16x3": 2, "jac_gram": 3, "jac_eig_apply": 4,
             "jac_apply": 5, "qr_panel": 6, "direct": 7}


def prof_enable(on=True):
    check(lib().tn_prof_enable(1 if on else 0))


def prof_collect(tag, reset=False):
    """(total_ms, launches, flops, bytes) of the launches recorded for `tag`."""
    ms, n = ctypes.c_double(0), ctypes.c_int64(0)
    fl, by = ctypes.c_double(0), ctypes.c_double(0)
    check(lib().tn_prof_collect(PROF_TAGS[tag], ctypes.byref(ms), ctypes.byref(n), ctypes.byref(fl),
                                ctypes.byref(by), 1 if reset else 0))
    return ms.value, n.value, fl.value, by.value


def loaded_hip_runtimes():
    """Paths of libamdhip64 mapped in this process (must be exactly one)."""
    paths = set()
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64" in line:
                paths.add(line.split()[-1])
    return sorted(paths)


# ---------------------------------------------------------------- helpers
def device():
    if not torch.cuda.is_available():
        raise RuntimeError("tncontract_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def dtype_code(t):
    if t.dtype == torch.float64:
        return TN_F64
    if t.dtype == torch.float32:
        return TN_F32
    raise NotImplementedError("tncontract_amd supports float64/float32 tensors, not %s" % t.dtype)


def workspace(nbytes):
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device())


_SPLITK_WS = {}


def contract(A, am, ak, B, bk, bn, C, cm, cn, dtype_code, prec, M, N, K, alpha=1.0, beta=0.0):
    """tn_contract through the split-K variant: the partial-sum workspace is
    one buffer per device, grown on demand and reused by every call (all
    calls are ordered on the current stream)."""
    need = lib().tn_contract_workspace(int(M), int(N), int(K), dtype_code)
    ws, nb = None, 0
    if need:
        dev = torch.cuda.current_device()
        ws = _SPLITK_WS.get(dev)
        if ws is None or ws.numel() < need:
            ws = torch.empty(max(int(need), 1 << 24), dtype=torch.uint8, device=device())
            _SPLITK_WS[dev] = ws
        nb = ws.numel()
    check(lib().tn_contract_ws(ptr(A), ctypes.byref(am), ctypes.byref(ak), ptr(B), ctypes.byref(bk),
                               ctypes.byref(bn), ptr(C), ctypes.byref(cm), ctypes.byref(cn),
                               dtype_code, prec, alpha, beta, ptr(ws) if ws is not None else None,
                               nb, stream_ptr()))


def empty(shape, dtype=torch.float64):
    return torch.empty(tuple(int(s) for s in shape), dtype=dtype, device=device())


def tensor_modes(t, axes):
    """Mode group for the given axes of a (possibly strided) torch tensor."""
    return modes([t.shape[a] for a in axes], [t.stride(a) for a in axes])


def _merged_modes(shape, strides):
    """Axes of a strided view with size-1 axes dropped and neighbours whose
    strides nest (stride[i] == stride[i+1] * dim[i+1]) fused; the dense
    row-major copy of the view has the same fused axes."""
    dims, st = [], []
    for d, s in zip(shape, strides):
        if d == 1:
            continue
        if dims and st[-1] == s * d:
            dims[-1] *= d
            st[-1] = s
        else:
            dims.append(int(d))
            st.append(int(s))
    return dims, st


def _permute_into(src, dst):
    """dst (dense, same shape) <- src (strided view) through tn_permute; a view
    with more than TN_MAX_MODES unfusable axes is copied slice by slice along
    its outermost non-trivial axis."""
    dims, st = _merged_modes(src.shape, src.stride())
    if len(dims) <= TN_MAX_MODES:
        m = modes(dims, st) if dims else modes([1], [0])
        check(lib().tn_permute(ctypes.c_void_p(src.data_ptr()), ctypes.byref(m), ptr(dst),
                               dtype_code(src), stream_ptr()))
        return
    a = next(i for i, d in enumerate(src.shape) if d > 1)
    for i in range(src.shape[a]):
        _permute_into(src.select(a, i), dst.select(a, i))


def materialize(t):
    """Dense row-major copy of a strided device tensor through tn_permute
    (the storage offset is folded into data_ptr)."""
    out = torch.empty(tuple(t.shape), dtype=t.dtype, device=t.device)
    if t.numel() == 0:
        return out
    _permute_into(t, out)
    return out


def contiguous(t):
    return t if t.is_contiguous() else materialize(t)


def convert(t, dtype):
    if t.dtype == dtype:
        return t
    src = contiguous(t)
    out = torch.empty(tuple(t.shape), dtype=dtype, device=t.device)
    if t.numel():
        check(lib().tn_convert(ptr(src), dtype_code(src), ptr(out), dtype_code(out), t.numel(), stream_ptr()))
    return out


def sumsq(t):
    """FP64 sum of squares as a host float (synchronises the stream)."""
    src = contiguous(t)
    out = torch.empty(1, dtype=torch.float64, device=t.device)
    n = src.numel()
    ws = workspace(lib().tn_sumsq_workspace(n))
    check(lib().tn_sumsq(ptr(src), dtype_code(src), n, ptr(out), ptr(ws), ws.numel(), stream_ptr()))
    return float(out.item())


def count_nonfinite(t):
    """(number of infinite entries, number of NaN entries) of t (host ints)."""
    src = contiguous(t)
    out = torch.empty(2, dtype=torch.int64, device=t.device)
    check(lib().tn_count_nonfinite(ptr(src), dtype_code(src), src.numel(), ptr(out), stream_ptr()))
    h = out.cpu().numpy()
    return int(h[0]), int(h[1])


def normalize_(t, norm_out=None):
    """In place t /= ||t|| (device-side, no host sync); the norm is written to
    the device scalar norm_out if given."""
    assert t.is_contiguous()
    n = t.numel()
    ws = workspace(lib().tn_normalize_workspace(n))
    out = ctypes.c_void_p(norm_out.data_ptr()) if norm_out is not None else None
    check(lib().tn_normalize(ptr(t), dtype_code(t), n, out, ptr(ws), ws.numel(), stream_ptr()))
    return t


def scale_(t, alpha):
    """In-place t *= alpha; t must be contiguous."""
    if t.numel():
        check(lib().tn_scale(ptr(t), dtype_code(t), t.numel(), float(alpha), stream_ptr()))
    return t


def sqrt_(t):
    """In-place t = sqrt(t); t must be contiguous."""
    if t.numel():
        check(lib().tn_sqrt(ptr(t), dtype_code(t), t.numel(), stream_ptr()))
    return t


def axpby(x, y, a, b):
    x, y = contiguous(x), contiguous(y)
    out = torch.empty_like(x)
    if x.numel():
        check(lib().tn_axpby(ptr(x), ptr(y), ptr(out), dtype_code(x), x.numel(), float(a), float(b), stream_ptr()))
    return out


def diag_inverse(a):
    """(diag(1 / a_ii) as a dense device matrix, [nonzero off-diagonal count,
    zero diagonal count]) of a square strided device matrix (synchronises)."""
    n = a.shape[0]
    out = torch.empty((n, n), dtype=a.dtype, device=a.device)
    info = torch.zeros(2, dtype=torch.float64, device=a.device)
    ws = workspace(lib().tn_diag_inverse_workspace(n))
    check(lib().tn_diag_inverse(ptr(a) if n else None, n, a.stride(0) if n else 0,
                                a.stride(1) if n else 0, ptr(out) if n else None, dtype_code(a),
                                ptr(info), ptr(ws), ws.numel(), stream_ptr()))
    return out, info.cpu().numpy()


def diag_inverse_async(a):
    """diag_inverse without the host read: (inverse, device info [nonzero
    off-diagonal count, zero diagonal count]), both ordered on the current
    stream."""
    n = a.shape[0]
    out = torch.empty((n, n), dtype=a.dtype, device=a.device)
    info = torch.zeros(2, dtype=torch.float64, device=a.device)
    ws = workspace(lib().tn_diag_inverse_workspace(n))
    check(lib().tn_diag_inverse(ptr(a), n, a.stride(0), a.stride(1), ptr(out), dtype_code(a),
                                ptr(info), ptr(ws), ws.numel(), stream_ptr()))
    return out, info


def diag_embed(s, dtype=torch.float64):
    """k x k diagonal matrix from a device float64 vector s."""
    k = s.numel()
    out = torch.empty((k, k), dtype=dtype, device=s.device)
    if k:
        check(lib().tn_diag_embed(ptr(s), k, ptr(out), dtype_code(out), stream_ptr()))
    return out
