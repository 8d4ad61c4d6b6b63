"""ctypes binding of libroger_amd.so (the C ABI declared in include/roger_amd.h).

The library must have been built (``python -m roger_amd.build`` or
``__graft_entry__.build()``); there is no fallback path: a missing library or
a missing GPU raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libroger_amd.so")

RG_SUCCESS = 0
RG_ERR_INVALID = -1
RG_ERR_DEVICE = -2
RG_ERR_OOM = -3
RG_ERR_OVERFLOW = -4
RG_ERR_NOT_READY = -5

RG_F32, RG_I16, RG_I32 = 0, 1, 2
RG_STYLE_BYPASS, RG_STYLE_RESET, RG_STYLE_TERMALL = 0x01, 0x02, 0x04
RG_STYLE_PARALLEL, RG_STYLE_PARALLEL_BYPASS = 0x06, 0x07
RG_MAX_PASSES = 96
RG_COLOUR_NONE, RG_COLOUR_RCT, RG_COLOUR_ICT = 0, 1, 2
RG_FILTER_53, RG_FILTER_97 = 0, 1


class RogerError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (_ERRNAMES.get(code, "error"), code, msg))
        self.code = code


_ERRNAMES = {RG_ERR_INVALID: "RG_ERR_INVALID", RG_ERR_DEVICE: "RG_ERR_DEVICE", RG_ERR_OOM: "RG_ERR_OOM",
             RG_ERR_OVERFLOW: "RG_ERR_OVERFLOW", RG_ERR_NOT_READY: "RG_ERR_NOT_READY"}


class CbDesc(ctypes.Structure):
    _fields_ = [("x0", ctypes.c_int32), ("y0", ctypes.c_int32), ("w", ctypes.c_int32), ("h", ctypes.c_int32),
                ("comp", ctypes.c_int32), ("orient", ctypes.c_int32), ("band", ctypes.c_int32),
                ("res", ctypes.c_int32)]


class CbResult(ctypes.Structure):
    _fields_ = [("len", ctypes.c_int32), ("npasses", ctypes.c_int32), ("nbp", ctypes.c_int32),
                ("flags", ctypes.c_int32)]


class BatchImage(ctypes.Structure):
    _fields_ = [("d_src", ctypes.c_void_p), ("w", ctypes.c_int32), ("h", ctypes.c_int32), ("ncomp", ctypes.c_int32),
                ("filter", ctypes.c_int32), ("colour", ctypes.c_int32), ("levels", ctypes.c_int32),
                ("delta", ctypes.c_double)]


class EncodeParams(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int32), ("filter", ctypes.c_int32), ("colour", ctypes.c_int32),
                ("cbw", ctypes.c_int32), ("cbh", ctypes.c_int32), ("delta", ctypes.c_double),
                ("ll_fp16", ctypes.c_int32), ("bitdepth", ctypes.c_int32), ("style", ctypes.c_int32)]


class J2KInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("w", "h", "ncomp", "levels", "cbw", "cbh", "style", "filter", "colour",
                                              "nlayers", "bitdepth", "guard_bits", "ncb")] + \
               [("data_bytes", ctypes.c_int64), ("step", ctypes.c_float * 97)]


P = ctypes.c_void_p
I = ctypes.c_int
SZ = ctypes.c_size_t
L = ctypes.c_long

# name -> (restype, argtypes); the list is the exported surface of include/roger_amd.h
SIGNATURES = {
    "rg_error_string": (ctypes.c_char_p, [I]),
    "rg_last_error": (ctypes.c_char_p, []),
    "rg_context_create": (I, [I, P]),
    "rg_context_destroy": (I, [P]),
    "rg_context_sync": (I, [P]),
    "rg_context_device_name": (I, [P, ctypes.c_char_p, I]),
    "rg_context_query": (I, [P]),
    "rg_malloc": (I, [P, SZ, P]),
    "rg_free": (I, [P, P]),
    "rg_memcpy_h2d": (I, [P, P, P, SZ]),
    "rg_memcpy_h2d_async": (I, [P, P, P, SZ]),
    "rg_memcpy_d2h": (I, [P, P, P, SZ]),
    "rg_memset": (I, [P, P, I, SZ]),
    "rg_memcpy_d2d": (I, [P, P, P, SZ]),
    "rg_host_alloc": (I, [P, SZ, P]),
    "rg_host_free": (I, [P, P]),
    "rg_stream_copy": (I, [P, P, P, SZ]),
    "rg_timer_start": (I, [P]),
    "rg_timer_stop": (I, [P, P]),
    "rg_colour_fwd": (I, [P, P, I, I, I, I, I, P]),
    "rg_colour_inv": (I, [P, P, I, I, I, I, I, P]),
    "rg_planar_split_i16": (I, [P, P, I, I, I, P]),
    "rg_dwt_scratch_bytes": (SZ, [I, I, I, I]),
    "rg_dwt_fwd": (I, [P, I, P, I, I, I, I, P, I, P, P]),
    "rg_dwt_fwd_u8_scratch_bytes": (ctypes.c_size_t, [I, I, I, I, I]),
    "rg_dwt_fwd_u8": (I, [P, P, I, I, I, I, I, I, P, P, P]),
    "rg_dwt_inv_u8_scratch_bytes": (ctypes.c_size_t, [I, I, I, I, I]),
    "rg_dwt_inv_u8": (I, [P, P, I, I, I, I, I, I, P, P, P]),
    "rg_dwt97_fwd_float": (I, [P, P, I, I, I, I, P, P]),
    "rg_dwt_fwd_noalloc": (I, [P, I, P, I, I, I, I, P, I, P]),
    "rg_dwt_inv": (I, [P, I, P, I, I, I, I, P, P, P]),
    "rg_codeblocks": (I, [I, I, I, I, I, P, I]),
    "rg_t1_encode": (I, [P, P, I, SZ, P, I, P, I, P]),
    "rg_encoder_create": (I, [P, P]),
    "rg_encoder_destroy": (I, [P]),
    "rg_encoder_encode_u8": (I, [P, P, I, I, I, I, P]),
    "rg_encoder_run_planar": (I, [P, P, I, I, I, I, I, I, I, I, I, I, ctypes.c_double]),
    "rg_encoder_load_planar": (I, [P, P, I, I, I, I, I, I]),
    "rg_encoder_dwt": (I, [P, I, I, ctypes.c_double]),
    "rg_encoder_planar_i16": (I, [P]),
    "rg_encoder_channel_i16": (I, [P, I, P]),
    "rg_encoder_bpc": (I, [P, I, I]),
    "rg_encoder_coefficients": (I, [P, P, P, P, P]),
    "rg_encoder_float_coefficients": (I, [P, P]),
    "rg_encoder_dwt_in": (I, [P, I, P, P, P]),
    "rg_encoder_num_codeblocks": (I, [P]),
    "rg_encoder_get_codeblocks": (I, [P, P, P, P, P, L, P]),
    "rg_encoder_get_pass_lengths": (I, [P, P]),
    "rg_encoder_enable_rate_control": (I, [P, I]),
    "rg_encoder_set_t1_phases": (I, [P, I]),
    "rg_encoder_rate_control": (I, [P, L, P, P]),
    "rg_encoder_rate_control_layers": (I, [P, I, P, P, P]),
    "rg_j2k_parse": (I, [P, L, P, P, P, P, P, P]),
    "rg_decoder_decode_j2k": (I, [P, P, L, P]),
    "rg_encoder_write_j2k": (I, [P, P, L, P]),
    "rg_encoder_map_dwt_out": (I, [P, P]),
    "rg_encoder_unmap_dwt_out": (I, [P, P]),
    "rg_decoder_create": (I, [P, P]),
    "rg_decoder_destroy": (I, [P]),
    "rg_decoder_run_planar": (I, [P, P, I, I, I, I, I, I, I]),
    "rg_decoder_load_planar": (I, [P, P, I, I, I, I, I, I]),
    "rg_decoder_idwt": (I, [P, I]),
    "rg_decoder_map_out": (I, [P, P]),
    "rg_t1_decode": (I, [P, P, P, P, P, I, P, I, SZ]),
    "rg_t1_decode_style": (I, [P, P, P, P, P, I, P, I, SZ, I, P]),
    "rg_decoder_decode_codeblocks": (I, [P, P, P, P, P, L, I, I, I, I, P, P]),
    "rg_decoder_image_u8": (I, [P, P]),
    "rg_decoder_coefficients": (I, [P, P]),
    "rg_batch_create": (I, [P, P]),
    "rg_batch_destroy": (I, [P]),
    "rg_batch_plan": (I, [P, P, I, I, I, I, I, I]),
    "rg_batch_get_pass_lengths": (I, [P, I, P]),
    "rg_batch_run": (I, [P]),
    "rg_batch_num_codeblocks": (I, [P, I]),
    "rg_batch_coefficients": (I, [P, I, P]),
    "rg_batch_get_codeblocks": (I, [P, I, P, P, P, P, L, P]),
    "rg_batch_write_j2k": (I, [P, I, P, L, P]),
    "rg_synth_u8": (I, [P, P, I, I, I, P, P, P, P]),
    "rg_decoder_unmap_out": (I, [P, P]),
}

_lib = None


def lib():
    """Loads libroger_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libroger_amd.so is not built: run __graft_entry__.build() "
                               "(or python -m roger_amd.build)")
        L_ = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L_, name)
            f.restype = res
            f.argtypes = args
        _lib = L_
    return _lib


def check(rc):
    if rc < 0:
        msg = lib().rg_last_error()
        raise RogerError(rc, msg.decode() if msg else "")
    return rc
