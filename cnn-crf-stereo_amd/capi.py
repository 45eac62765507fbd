"""ctypes binding of libccs.so (the C ABI declared in include/ccs.h).

Device memory and streams come from torch (plumbing only); every computation
runs in the HIP library.  A missing library raises -- there is no fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libccs.so")

FP32, BF16 = 0, 1
OK, ERR_ARG, ERR_UNSUPPORTED, ERR_HIP, ERR_ALLOC, ERR_IO, ERR_SHAPE, ERR_NAN = 0, -1, -2, -3, -4, -5, -6, -7
LAYOUT_NCHW, LAYOUT_NHWC = 0, 1
EPILOGUE_RAW, EPILOGUE_NEG_SOFTMAX = 0, 1

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_SZ = ctypes.c_size_t

# name -> (restype, argtypes); mirrors include/ccs.h.
SIGNATURES = {
    "ccs_last_error": (ctypes.c_char_p, []),
    "ccs_version": (ctypes.c_char_p, []),
    "ccs_create": (_I, [ctypes.POINTER(_P), _P]),
    "ccs_destroy": (_I, [_P]),
    "ccs_set_stream": (_I, [_P, _P]),
    "ccs_synchronize": (_I, [_P]),
    "ccs_reserve": (_I, [_P, _I, _I, _I, _I, _I]),
    "ccs_feature_stride": (_I, [_I]),
    "ccs_pad_normalize": (_I, [_P, _P, _I, _I, _I, _I, _P, _P]),
    "ccs_rgb_to_gray": (_I, [_P, _P, _I, _I, _P]),
    "ccs_conv2d_forward": (_I, [_P, _P, _I, _I, _I, _I, _P, _P, _I, _I, _I, _I, _I, _P]),
    "ccs_conv2d_forward_noalloc": (_I, [_P, _P, _I, _I, _I, _I, _P, _P, _I, _I, _I, _I, _I, _P]),
    "ccs_conv_packed_bytes": (_SZ, [_I, _I, _I, _I, _I]),
    "ccs_conv_pack_weights": (_I, [_P, _P, _I, _I, _I, _I, _I, _P]),
    "ccs_conv_forward_packed": (_I, [_P, _P, _I, _I, _I, _I, _P, _P, _I, _I, _I, _I, _I, _P]),
    "ccs_conv2d_backward": (_I, [_P, _P, _I, _I, _I, _P, _P, _I, _I, _I, _I, _P, _P, _P, _P]),
    "ccs_conv2d_backward_stored": (_I, [_P, _P, _I, _P, _I, _I, _I, _I, _P, _I, _I, _I, _I, _P, _P, _P, _P]),
    "ccs_correlation_forward": (_I, [_P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I]),
    "ccs_correlation_backward": (_I, [_P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I]),
    "ccs_correlation_volume": (_I, [_P, _P, _P, _I, _I, _I, _I, _I, _I, _I, _I, _P]),
    "ccs_correlation_volume_backward": (_I, [_P, _P, _P, _I, _I, _I, _I, _I, _I, _I, _P, _P, _P]),
    "ccs_softmax_xent": (_I, [_P, _P, _I, _I, _P, _P, _P]),
    "ccs_edge_weights": (_I, [_P, _P, _I, _I, _F, _F, _F, _F, _P, _P]),
    "ccs_slackprop_forward": (_I, [_P, _P, _P, _P, _P, _I, _I, _I, _I, _F, _F, _I, ctypes.POINTER(_F)]),
    "ccs_slackprop_forward_volume_out": (_I, [_P, _P, _P, _P, _P, _I, _I, _I, _I, _F, _F, _I, _I,
                                              ctypes.POINTER(_F)]),
    "ccs_slackprop": (_I, [_P, _P, _P, _P, _I, _I, _I, _F, _F, _I, _I, _P, _P, _P, _P]),
    "ccs_slackprop_status": (_I, [_P]),
    "ccs_slackprop_forward_pw": (_I, [_P, _P, _P, _P, _I, _I, _I, _I, _I, _I, ctypes.POINTER(_F)]),
    "ccs_pairwise_penalties": (_I, [_P, _P, _I, _I, _I, _P, _P]),
    "ccs_slackprop_pw": (_I, [_P, _P, _P, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P]),
    "ccs_slackprop_decode_gap": (_I, [_P, _P, _P, _I, _I, _I, _F, _F, _I, _P, _P]),
    "ccs_slackprop_sweep": (_I, [_P, _P, _P, _P, _P, _I, _I, _I, _F, _F, _I, _I, _P, _P, _P]),
    "ccs_crf_energy": (_I, [_P, _P, _P, _P, _I, _I, _I, _P, _F, _F, _I, _P]),
    "ccs_crf_energy_masked": (_I, [_P, _P, _P, _P, _I, _I, _I, _P, _P, _F, _F, _I, _P]),
    "ccs_crf_loss_augment": (_I, [_P, _P, _P, ctypes.c_long, _I, _F, _P]),
    "ccs_crf_ssvm_grad": (_I, [_P, _P, _P, _P, _P, _I, _I, _I, _F, _F, _I, _P, _P, _P, _P]),
    "ccs_bias_forward": (_I, [_P, _P, _I, _I, _I, _I, _I, _P]),
    "ccs_tanh_forward": (_I, [_P, _P, ctypes.c_long]),
    "ccs_pairwise_head": (_I, [_P, _P, _I, _I, _I, _I, _P, _I, _P]),
    "ccs_scale_forward": (_I, [_P, _P, ctypes.c_long, ctypes.c_float]),
    "ccs_xy_planes": (_I, [_P, _I, _I, _P]),
    "ccs_quad_fit_direct": (_I, [_P, _P, _P, _I, _I, _I, _F, _F]),
    "ccs_left_right_check": (_I, [_P, _P, _P, _I, _I, _F, _P]),
    "ccs_softmax_forward": (_I, [_P, _P, ctypes.c_long, _I, _I]),
    "ccs_neg_softmax_backward": (_I, [_P, _P, _P, ctypes.c_long, _I, _P]),
    "ccs_c4_to_planar": (_I, [_P, _P, _I, _I, ctypes.c_long, _I, _P]),
    "ccs_nhwc_to_nchw": (_I, [_P, _P, _I, _I, _I, _I, _I, _P]),
    "ccs_nchw_to_nhwc": (_I, [_P, _P, _I, _I, _I, _I, _P]),
    "ccs_argmin": (_I, [_P, _P, _I, _I, _I, _F, _F, _P]),
    "ccs_remap": (_I, [_P, _P, _I, _I, _F, _F, _P]),
}

_lib = None


class CCSError(RuntimeError):
    def __init__(self, msg, status=None):
        super().__init__(msg)
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CCSError("libccs.so is not built (run __graft_entry__.build()): %s" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                raise CCSError("libccs.so does not export %s: rebuild the library" % name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().ccs_last_error().decode()


def call(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        raise CCSError("%s failed (status %d): %s" % (name, rc, last_error()), rc)
    return rc


def ptr(t):
    """Device (or host) address of a torch tensor, None for None."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def feature_stride(C):
    return lib().ccs_feature_stride(int(C))


class Handle:
    """Owns a ccs handle bound to a HIP stream (torch's current stream by default)."""

    def __init__(self, stream=None):
        import torch

        if stream is None:
            stream = torch.cuda.current_stream()
        self._h = ctypes.c_void_p()
        call("ccs_create", ctypes.byref(self._h), ctypes.c_void_p(stream.cuda_stream))

    @property
    def h(self):
        return self._h

    def synchronize(self):
        call("ccs_synchronize", self._h)

    def close(self):
        if self._h:
            lib().ccs_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
