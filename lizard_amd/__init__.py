"""lizard_amd: LIZARD's receive hot path on MI355X, through its C ABI (include/lizard_amd.h).

This module is a thin ctypes mirror of the C ABI for Python callers (tests, benches). Everything it computes runs in
the native library lizard_amd/liblizard_amd.so; it raises if that library is missing rather than computing anything
itself.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# LIZARD_AMD_LIB: another build of the same library (the host-sanitized one, lizard_amd/_san/)
LIB_PATH = os.environ.get("LIZARD_AMD_LIB") or os.path.join(HERE, "liblizard_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "lizard_amd.h")

LIZ_OK, LIZ_E_ARG, LIZ_E_NOMEM, LIZ_E_IO, LIZ_E_UNSUPPORTED, LIZ_E_FORMAT, LIZ_E_TOOBIG, LIZ_E_STATE, LIZ_E_INTERNAL = (
    0, -1, -2, -3, -4, -5, -6, -7, -8)
LIZ_BLOCK = 476
LIZ_PAYLOAD = 472
LDPC_CRC, LDPC_TRACE, LDPC_PLAIN = 1, 2, 4
SRC_CAPTURE, SRC_GRID = 0, 1
STAGE_SOFT, STAGE_BLOCKS = 0, 1
RUN_GRAPH, RUN_STREAM, RUN_FRAMES, RUN_PLAIN = 0, 1, 2, 3
MSG_TRACE_STRIDE = 210 * 128
RATE_INDEX = {"7/8": 6, "3/4": 4, "2/3": 3, "1/2": 2}

_P, _I, _F, _SZ = C.c_void_p, C.c_int, C.c_float, C.c_size_t


class LizError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d)" % (msg, code))
        self.code = code


class _Decoded(C.Structure):
    _fields_ = [("found", C.c_int), ("ring", C.c_int), ("n", C.c_int), ("word", C.c_int), ("blocks", C.c_int),
                ("fps", C.c_int), ("held_used", C.c_int), ("total", C.c_int), ("verified", C.c_int),
                ("quad", C.c_float * 8), ("pilot_blocks", C.c_int), ("pilot_r", C.c_float * 2),
                ("pilot_sd", C.c_float * 2)]


class Format(C.Structure):
    """liz_format: blocks (the size, 4..64), ring (0..2, or -1 the default), fps (1..255), codes (1 or 2)."""
    _fields_ = [("blocks", C.c_int), ("ring", C.c_int), ("fps", C.c_int), ("codes", C.c_int)]


class Geometry(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("n", "span", "pxm", "side", "width", "height", "gap", "frame_blocks")]


class Registration(C.Structure):
    _fields_ = [("H", C.c_float * 9), ("nodes", C.c_void_p)]


class Frame(C.Structure):
    """lizx_frame: one frame of a batch (device pointers)."""
    _fields_ = [("img", C.c_void_p), ("w", C.c_int), ("h", C.c_int), ("H", C.c_float * 9), ("nodes", C.c_void_p),
                ("grid", C.c_void_p)]


class Soft(C.Structure):
    _fields_ = [("llr", C.c_void_p), ("sub_est", C.c_void_p), ("sub_rms", C.c_void_p), ("coef", C.c_void_p),
                ("align", C.c_float * 2), ("blk_pilot", C.c_void_p), ("soft", C.c_void_p)]


_lib = None


def lib():
    """The native library; raises when it has not been built (python -m lizard_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("lizard_amd: %s is missing; build it with python -m lizard_amd.build" % LIB_PATH)
    L = C.CDLL(LIB_PATH)

    def sig(name, res, *args):
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = list(args)

    sig("liz_version", C.c_char_p)
    sig("liz_last_error", C.c_char_p)
    sig("liz_last_error_code", _I)
    sig("liz_decoder_new", _P, _I)
    sig("liz_decoder_free", None, _P)
    sig("liz_decoder_max_blocks", _I, _P)
    sig("liz_decode", _I, _P, _P, _I, _I, _I, _I, _P, _P, _P)
    sig("lizx_format_info", _I, _I, _P)
    sig("lizx_finish_llr", _I, _P, _I, _P, _P, _P, _P, _P, _P)
    sig("lizx_finish_grid", _I, _P, _I, _I, _P, _P, _P, _P, _P, C.POINTER(Soft))
    sig("lizx_node_lattice", _I, _I, _P)
    sig("lizx_frame_slots", _I, _I, _P, _I, _P)
    sig("lizx_modulate", _I, _P, _I, _I, _P, _I, _P, _P)
    sig("lizx_frame_slots_device", _I, _P, _I, _I, _P, _I, _P)
    sig("lizx_capture", _I, _P, _I, _I, _P, _P, _P)
    sig("lizx_capture_symbol", _I, _P, _I, _I, _I, _P, _P, _P)
    sig("lizx_read_word", _I, _I, _P, _I, _I, _I, _P, _P, _P, _P)
    sig("lizx_word_geometry", _I, _I, _P, _P)
    sig("lizx_border_kinds", _I, _I, _I, _I, _P)
    sig("liz_ring_cells", _I, _I)
    sig("lizx_replay_open", _P, C.c_char_p)
    sig("lizx_replay_close", None, _P)
    sig("lizx_replay_frames", _I, _P)
    sig("lizx_replay_frame_info", _I, _P, _I, _P, _P, _P)
    sig("lizx_replay_frame", _I, _P, _I, _P)
    sig("lizx_replay_meta", C.c_char_p, _P, _I)
    sig("lizx_replay_create", _P, C.c_char_p)
    sig("lizx_replay_add", _I, _P, _P, _I, _I, _I)
    sig("lizx_replay_finish", C.c_int64, _P, C.c_char_p, C.c_char_p)
    sig("lizx_replay_writer_free", None, _P)
    sig("liz_simd", _I)
    sig("liz_stream_fill", None, C.c_uint32, _P)
    sig("liz_layout_rects", _I, _I, _I, _I, _P)
    sig("liz_geometry_of", _I, C.POINTER(Format), C.POINTER(Geometry))
    sig("liz_room_for", C.c_double, _I, _I)
    sig("liz_blocks_for", _I, _I)
    sig("liz_pick", _I, C.c_double, C.c_double, _I, _I, _I)
    sig("liz_encoder_new", _P, C.POINTER(Format))
    sig("liz_encoder_free", None, _P)
    sig("liz_encoder_geometry", _I, _P, C.POINTER(Geometry))
    sig("liz_encoder_paint", _I, _P, _P, C.c_uint32, _P, _I, _I)
    sig("lizx_b3_chunk_cvs", _I, _P, _P, C.c_uint64, _I, _P, _P)
    sig("lizx_b3_hash_device", _I, _P, _P, C.c_uint64, _P, _P)
    sig("lizx_b3_root", _I, _P, C.c_uint32, _P)
    sig("lizx_b3_hash", _I, _P, C.c_uint64, _P)
    sig("lizx_word_encode", _I, _I, _I, _I, _P)
    sig("lizx_word_decode", _I, _I, _P, _P, _P)
    sig("lizx_format_coefs", _I, _I, _P, _P)
    sig("lizx_format_blocks", _I, _I, _P, _P, _P)
    sig("lizx_reg_nodes", _I, _I, _P, _I, _I, _P, _P)
    sig("lizx_reg_loss_grad", _I, _P, _I, _I, _P, _I, _I, _I, _P, _P, _P, _F, _F, _P, _P)
    sig("lizx_reg_loss_grad_spectral", _I, _P, _I, _I, _P, _I, _I, _I, _P, _P, _P, _P, _F, _F, _P, _P)
    sig("lizx_reg_refine_spectral", _I, _P, _I, _I, _P, _I, _I, _I, _P, _P, _P, _P, _I, _I, _P)
    sig("lizx_reg_refine", _I, _P, _I, _I, _P, _I, _I, _I, _P, _P, _P, _I, _P)
    sig("lizx_reg_refine_k1", _I, _P, _I, _I, _P, _I, _I, _I, _P, _P, _P, _I, _P)
    sig("lizx_reg_refine_nodes", _I, _P, _I, _I, _P, _I, _I, _I, _P, _P, _P, _I, _P)
    sig("lizx_warp", _I, _P, _I, _I, _P, _I, _I, _I, C.POINTER(Registration), _P)
    sig("lizx_decode_reg", _I, _P, _I, _I, _P, _I, _I, _I, C.POINTER(Registration), _P, _P, _P, _P, C.POINTER(Soft))
    sig("lizx_decoder_hold", _I, _P, _I, _I, C.POINTER(Registration))
    sig("lizx_decoder_graphs", _I, _P, _I)
    sig("lizx_decoder_shape", _I, _P, _I, _I, _I)
    sig("lizx_finder_anchors", _I, _P, _I)
    sig("lizx_hang_bound", _I, C.c_double)
    sig("lizx_anchor_find", _I, _P, _I, _P, _I, _I, _I, _P, _P, _P)
    sig("lizx_anchor_register", _I, _P, _I, _P, _I, _I, _I, _P, _F, _P, _P, _P, _P)
    sig("lizx_border_refine", _I, _P, _I, _P, _I, _I, _I, _P, _F, _P, _P)
    sig("lizx_find", _I, _P, _P, _I, _I, _I, _P, _P, _P, _P, _P)
    sig("lizx_find_batch", _I, _P, _I, _P, _I, _I, _I, _P, _P, _P, _P, _P)
    sig("lizx_batch_new", _P, _I)
    sig("lizx_batch_free", None, _P)
    sig("lizx_batch_plan", _I, _P, _I, _P, _P, _I, _I)
    sig("lizx_batch_dims", _I, _P, _P)
    sig("lizx_batch_run", _I, _P, _P, _I)
    sig("lizx_batch_sync", _I, _P, C.POINTER(C.c_float))
    sig("lizx_batch_tune", _I, _P, _P, _I, _P)
    sig("lizx_batch_shape", _I, _P, _P)
    sig("lizx_batch_tune_plans", _I, _P, _P, _I, _P)
    sig("lizx_batch_otf", _I, _P, _P)
    sig("lizx_decoder_otf", _I, _P, C.c_float)
    sig("lizx_batch_shape_plans", _I, _P, _P)
    sig("lizx_batch_blocks", _I, _P, _P, _P, _P, _P)
    sig("lizx_batch_soft", _I, _P, _I, _P, _P, _P, _P)
    sig("lizx_ldpc_dims", _I, _I, _P)
    sig("lizx_ldpc_new", _P, _I)
    sig("lizx_ldpc_free", None, _P)
    sig("lizx_ldpc_decode", _I, _P, _I, _P, _P, _P, _P, _P, _I, _I, _F, _I, _P, _P, _P, _P)
    sig("lizx_ldpc_decode_device", _I, _P, _I, _P, _P, _P, _P, _P, _I, _I, _F, _I, _P, _P)
    sig("lizx_ldpc_sync", _I, _P)
    sig("lizx_ldpc_tune", _I, _P, _I, _P, _P, _P, _P, _I, _I, _F, _I, _P)
    sig("lizx_ldpc_last_ms", _F, _P)
    sig("lizx_ldpc_shape", _I, _P, _I)
    sig("lizx_reg_border_fill", _I, _I, _P, _I, _I, _P, _P, _P, _P)
    sig("lizx_reg_lens_field", _I, _I, _P, _P, _P)
    sig("lizx_reg_anchor_fill", _I, _I, _P, _I, _I, _P, _P, _P, _P, _F, _P)
    sig("lizx_ldpc_decode_device_noalloc", _I, _P, _I, _P, _P, _P, _P, _P, _I, _I, _F, _I, _P, _P)
    sig("lizx_ldpc_encode", _I, _I, _I, _P, _P)
    sig("lizx_crc32c", C.c_uint32, _P, _SZ)
    sig("lizx_copy_probe", _I, _P, _P, _SZ, C.POINTER(C.c_float))
    sig("lizx_dev_alloc", _P, _SZ)
    sig("lizx_dev_free", None, _P)
    sig("lizx_dev_upload", _I, _P, _P, _SZ)
    sig("lizx_dev_download", _I, _P, _P, _SZ)
    sig("lizx_device_count", _I)
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def check(rc):
    if rc is None or (isinstance(rc, int) and rc < 0):
        L = lib()
        raise LizError(L.liz_last_error_code(), L.liz_last_error().decode(errors="replace"))
    return rc


def device_count():
    return lib().lizx_device_count()


def format_info(version):
    info = np.zeros(6, np.int32)
    check(lib().lizx_format_info(version, _p(info)))
    return dict(zip(("n", "subch", "blocks", "tiers", "vblocks", "used"), map(int, info)))


def ldpc_dims(rate):
    d = np.zeros(12, np.int32)
    check(lib().lizx_ldpc_dims(rate, _p(d)))
    return dict(zip(("n", "k", "m", "z", "mb", "kb", "np", "nt", "edges", "slots_total", "msg_bytes", "post_bytes"),
                    map(int, d)))


def frame_slots(version, blocks, cnt=0):
    info = format_info(version)
    blocks = np.ascontiguousarray(blocks, np.uint8)
    slots = np.zeros(info["subch"] * 640, np.uint8)
    check(lib().lizx_frame_slots(version, _p(blocks), cnt, _p(slots)))
    return slots


def format_coefs(version):
    S = format_info(version)["subch"] * 320
    u, v = np.zeros(S, np.int16), np.zeros(S, np.int16)
    check(lib().lizx_format_coefs(version, _p(u), _p(v)))
    return u.astype(np.int64), v.astype(np.int64)


def format_blocks(version):
    B = format_info(version)["blocks"]
    rate, s0, ns = np.zeros(B, np.int32), np.zeros(B, np.int32), np.zeros(B, np.int32)
    check(lib().lizx_format_blocks(version, _p(rate), _p(s0), _p(ns)))
    return rate, s0, ns


REG_PARAMS = 107


def reg_params(H, k1=0.0, anchors=None):
    """The 107 registration parameters: h0..h7 of H (normalised to h8 = 1), k1, the 7 x 7 anchors' (dx, dy)."""
    H = np.asarray(H, np.float64).reshape(3, 3)
    p = np.zeros(REG_PARAMS, np.float32)
    p[:8] = (H / H[2, 2]).ravel()[:8]
    p[8] = k1
    if anchors is not None:
        p[9:] = np.asarray(anchors, np.float32).ravel()
    return p


def reg_nodes(span, H0, w, h, params):
    nx = len(node_lattice(span))
    out = np.zeros((nx, nx, 2), np.float32)
    check(lib().lizx_reg_nodes(span, _p(np.ascontiguousarray(H0, np.float32).ravel()), w, h,
                               _p(np.ascontiguousarray(params, np.float32)), _p(out)))
    return out


def border_fill(span, H, w, h, nodes, use=None, axis=None):
    """The interior of the node lattice from its border (lizx_reg_border_fill): (nodes, K, model[15])."""
    nodes = np.array(nodes, np.float32, copy=True)
    model = np.zeros(15, np.float64)
    H = np.ascontiguousarray(H, np.float32)
    u = None if use is None else np.ascontiguousarray(use, np.uint8)
    ax = None if axis is None else np.ascontiguousarray(axis, np.float32)
    K = check(lib().lizx_reg_border_fill(span, _p(H), w, h, _p(ax), _p(nodes), _p(u), _p(model)))
    return nodes, K, model


def hang_bound(ms):
    """The hang bound for every handle's waits on its GPU work (lizx_hang_bound; 0: none)."""
    check(lib().lizx_hang_bound(float(ms)))


def anchor_fill(span, H, w, h, nodes, anchors, mpx, use=None, axis=None):
    """The node lattice from its border and the anchors (lizx_reg_anchor_fill): (nodes, K, model[15]); anchors
    49 x 5 (X, Y, dx, dy, found)."""
    nodes = np.array(nodes, np.float32, copy=True)
    model = np.zeros(15, np.float64)
    H = np.ascontiguousarray(H, np.float32)
    u = None if use is None else np.ascontiguousarray(use, np.uint8)
    ax = None if axis is None else np.ascontiguousarray(axis, np.float32)
    an = np.ascontiguousarray(anchors, np.float32)
    K = check(lib().lizx_reg_anchor_fill(span, _p(H), w, h, _p(ax), _p(nodes), _p(u), _p(an), mpx, _p(model)))
    return nodes, K, model


def lens_field(span, H, model):
    """A lens model's displacement at every node (lizx_reg_lens_field)."""
    nx = len(node_lattice(span))
    out = np.zeros((nx, nx, 2), np.float32)
    check(lib().lizx_reg_lens_field(span, _p(np.ascontiguousarray(H, np.float32)),
                                    _p(np.ascontiguousarray(model, np.float64)), _p(out)))
    return out


def node_lattice(span):
    pos = np.zeros(40, np.float32)
    n = check(lib().lizx_node_lattice(span, _p(pos)))
    return pos[:n].copy()


def registration(H, nodes):
    """A lizx_registration (keeps the nodes array alive on the returned object)."""
    nodes = np.ascontiguousarray(nodes, np.float32)
    r = Registration((C.c_float * 9)(*[float(x) for x in np.asarray(H, np.float64).ravel()]), _p(nodes))
    r._keep = nodes
    return r


def ldpc_encode(rate, data):
    """Encode count codewords: data is (count, k) bits (0/1); returns (count, n)."""
    d = ldpc_dims(rate)
    data = np.ascontiguousarray(data, np.uint8).reshape(-1, d["k"])
    cw = np.zeros((data.shape[0], d["n"]), np.uint8)
    check(lib().lizx_ldpc_encode(rate, data.shape[0], _p(data), _p(cw)))
    return cw


def crc32c(data):
    data = np.ascontiguousarray(data, np.uint8)
    return int(lib().lizx_crc32c(_p(data), data.size))


def copy_probe(dst, src, nbytes):
    ms = C.c_float(0)
    check(lib().lizx_copy_probe(dst.ptr, src.ptr, nbytes, C.byref(ms)))
    return ms.value


class Camera(C.Structure):
    """lizx_camera: w x h pixels, H (module -> pixel, 9 doubles), radial k1, blur sigma, AWGN sigma, noise seed,
    the second radial term k2."""
    _fields_ = [("w", C.c_int), ("h", C.c_int), ("H", C.c_double * 9), ("k1", C.c_double), ("blur", C.c_float),
                ("noise", C.c_float), ("seed", C.c_uint64), ("k2", C.c_double)]


class DeviceArray:
    """A device buffer owned through the library's allocator (no HIP runtime needed in Python)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = lib().lizx_dev_alloc(max(self.nbytes, 1))
        if not self.ptr:
            check(-1)

    @classmethod
    def of(cls, arr):
        arr = np.ascontiguousarray(arr)
        d = cls(arr.nbytes)
        check(lib().lizx_dev_upload(d.ptr, _p(arr), arr.nbytes))
        return d

    def put(self, arr):
        """Copies a host array into the start of this device array."""
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes:
            raise ValueError("the array is larger than the device array")
        check(lib().lizx_dev_upload(self.ptr, _p(arr), arr.nbytes))

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        check(lib().lizx_dev_download(_p(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().lizx_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class LdpcDecoder:
    """Batched layered min-sum LDPC (lizx_ldpc_*): ldpc_decode_stall over many codewords of mixed rates."""

    def __init__(self, device=-1):
        self.h = lib().lizx_ldpc_new(device)
        if not self.h:
            check(-1)

    def close(self):
        if self.h:
            lib().lizx_ldpc_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def decode(self, rates, llr, offsets, max_iter=30, stall_it=9, stall_ratio=0.95, crc=False, trace=False,
               plain=False, bits=True):
        rates = np.ascontiguousarray(rates, np.uint8)
        llr = np.ascontiguousarray(llr, np.int8)
        offsets = np.ascontiguousarray(offsets, np.int64)
        count = len(rates)
        its = np.zeros(count, np.int8)
        out_bits = np.zeros(llr.size, np.uint8) if bits else None
        flags = (LDPC_CRC if crc else 0) | (LDPC_TRACE if trace else 0) | (LDPC_PLAIN if plain else 0)
        by = np.zeros(count * LIZ_BLOCK, np.uint8) if crc else None
        ok = np.zeros(count, np.uint8) if crc else None
        post = np.zeros(llr.size, np.int16) if trace else None
        msg = np.zeros(count * MSG_TRACE_STRIDE, np.int8) if trace else None
        check(lib().lizx_ldpc_decode(self.h, count, _p(rates), _p(llr), _p(offsets), _p(out_bits), _p(its), max_iter,
                                     stall_it, stall_ratio, flags, _p(by), _p(ok), _p(post), _p(msg)))
        res = {"its": its, "bits": out_bits}
        if crc:
            res["bytes"] = by.reshape(count, LIZ_BLOCK)
            res["crc_ok"] = ok
        if trace:
            res["post"] = post
            res["msg"] = msg.reshape(count, MSG_TRACE_STRIDE)
        return res

    def decode_device(self, count, d_rate, d_llr, d_off, d_bits, d_its, max_iter=20, stall_it=0, stall_ratio=0.95,
                      crc=False, plain=False, d_bytes=None, d_crc_ok=None):
        flags = (LDPC_CRC if crc else 0) | (LDPC_PLAIN if plain else 0)
        ptr = lambda d: d.ptr if d is not None else None  # noqa: E731
        check(lib().lizx_ldpc_decode_device(self.h, count, ptr(d_rate), ptr(d_llr), ptr(d_off), ptr(d_bits),
                                            ptr(d_its), max_iter, stall_it, stall_ratio, flags, ptr(d_bytes),
                                            ptr(d_crc_ok)))

    def decode_device_noalloc(self, count, d_rate, d_llr, d_off, d_bits, d_its, max_iter=20, stall_it=0,
                              stall_ratio=0.95, crc=False, plain=False, d_bytes=None, d_crc_ok=None):
        """lizx_ldpc_decode_device_noalloc: the device decode with its scratch in the library's fixed work area."""
        flags = (LDPC_CRC if crc else 0) | (LDPC_PLAIN if plain else 0)
        ptr = lambda d: d.ptr if d is not None else None  # noqa: E731
        check(lib().lizx_ldpc_decode_device_noalloc(self.h, count, ptr(d_rate), ptr(d_llr), ptr(d_off), ptr(d_bits),
                                                    ptr(d_its), max_iter, stall_it, stall_ratio, flags, ptr(d_bytes),
                                                    ptr(d_crc_ok)))

    def sync(self):
        check(lib().lizx_ldpc_sync(self.h))
        return float(lib().lizx_ldpc_last_ms(self.h))

    def shape(self, per_cu):
        """Sets the persistent grid's workgroups a CU (lizx_ldpc_shape; 0: as many as fit); returns the count used."""
        rc = lib().lizx_ldpc_shape(self.h, per_cu)
        check(rc)
        return rc

    def tune(self, count, d_rate, d_llr, d_off, d_its, max_iter=20, stall_it=0, stall_ratio=0.95, reps=3):
        """The persistent grid's autotuner (lizx_ldpc_tune): workgroups a CU kept (0: as many as fit)."""
        pc = C.c_int(0)
        check(lib().lizx_ldpc_tune(self.h, count, d_rate.ptr, d_llr.ptr, d_off.ptr, d_its.ptr, max_iter, stall_it,
                                   stall_ratio, reps, C.byref(pc)))
        return pc.value


def word_geometry(span):
    """lizx_word_geometry: (word cells a side, codeword bytes) in the ring of span."""
    c, b = C.c_int(0), C.c_int(0)
    check(lib().lizx_word_geometry(span, C.byref(c), C.byref(b)))
    return c.value, b.value


def word_encode(span, version, fps=0):
    """lizx_word_encode: the format word's codeword (uint8) for the ring of span."""
    cw = np.zeros(64, np.uint8)
    n = check(lib().lizx_word_encode(span, version, fps, _p(cw)))
    return cw[:n]


def word_decode(span, q):
    """lizx_word_decode: soft values (4 x cells) -> (version, fps), or None."""
    q = np.ascontiguousarray(q, np.float32)
    v, f = C.c_int(0), C.c_int(0)
    got = check(lib().lizx_word_decode(span, _p(q), C.byref(v), C.byref(f)))
    return (v.value, f.value) if got else None


class Replay:
    """A .lizard replay read a frame at a time (lizx_replay_*)."""

    def __init__(self, path):
        self.h = lib().lizx_replay_open(os.fsencode(path))
        if not self.h:
            check(-1)

    def __len__(self):
        return check(lib().lizx_replay_frames(self.h))

    def info(self, i):
        w, h, ms = C.c_int(0), C.c_int(0), C.c_double(0)
        check(lib().lizx_replay_frame_info(self.h, i, C.byref(w), C.byref(h), C.byref(ms)))
        return w.value, h.value, ms.value

    def frame(self, i):
        w, h, _ = self.info(i)
        out = np.zeros((h, w), np.uint8)
        check(lib().lizx_replay_frame(self.h, i, _p(out)))
        return out

    def meta(self, which=0):
        return lib().lizx_replay_meta(self.h, which).decode()

    def close(self):
        if self.h:
            lib().lizx_replay_close(self.h)
            self.h = None


class ReplayWriter:
    """A .lizard replay written as it goes (lizx_replay_create / add / finish)."""

    def __init__(self, path):
        self.h = lib().lizx_replay_create(os.fsencode(path))
        if not self.h:
            check(-1)

    def add(self, luma):
        luma = np.asarray(luma, np.uint8)
        if luma.ndim != 2 or luma.strides[1] != 1:
            luma = np.ascontiguousarray(luma)
        return check(lib().lizx_replay_add(self.h, _p(luma), luma.shape[1], luma.shape[0], luma.strides[0]))

    def finish(self, meta, stats=""):
        n = lib().lizx_replay_finish(self.h, meta.encode(), stats.encode())
        check(int(n) if n < 0 else 0)
        lib().lizx_replay_writer_free(self.h)
        self.h = None
        return int(n)

    def close(self):
        if self.h:
            lib().lizx_replay_writer_free(self.h)
            self.h = None


def stream_fill(block_id):
    """liz_stream_fill: the test stream's 472-byte payload for an id."""
    out = np.zeros(LIZ_PAYLOAD, np.uint8)
    lib().liz_stream_fill(block_id & 0xFFFFFFFF, _p(out))
    return out


def layout_rects(w, h, layout=1):
    """liz_layout_rects: [(x, y, w, h), ...] of the frame's symbol regions."""
    r = np.zeros((2, 4), np.int32)
    k = check(lib().liz_layout_rects(w, h, layout, _p(r)))
    return [tuple(int(v) for v in r[i]) for i in range(k)]


def ring_cells(ring):
    return check(lib().liz_ring_cells(ring))


def geometry_of(blocks, ring=-1, fps=60, codes=1):
    """liz_geometry_of as a dict."""
    g = Geometry()
    check(lib().liz_geometry_of(C.byref(Format(blocks, ring, fps, codes)), C.byref(g)))
    return {k: getattr(g, k) for k, _ in Geometry._fields_}


def room_for(blocks, ring=-1):
    v = lib().liz_room_for(blocks, ring)
    if v < 0:
        check(int(v))
    return v


def blocks_for(blocks):
    return lib().liz_blocks_for(blocks)


def pick(w, h, codes=1, ring=-1, top_blocks=64):
    return check(lib().liz_pick(float(w), float(h), codes, ring, top_blocks))


class Encoder:
    """liblizard's per-frame encoder (liz_encoder_*), painting on the GPU."""

    def __init__(self, blocks, ring=-1, fps=60, codes=1):
        self.h = lib().liz_encoder_new(C.byref(Format(blocks, ring, fps, codes)))
        if not self.h:
            check(-1)
        g = Geometry()
        check(lib().liz_encoder_geometry(self.h, C.byref(g)))
        self.geometry = {k: getattr(g, k) for k, _ in Geometry._fields_}

    def paint(self, blocks, picture=0, fmt=1):
        """One frame: blocks (frame_blocks x 476 bytes) -> the page, height x width (x 4 for the 4-byte formats)."""
        g = self.geometry
        blocks = np.ascontiguousarray(blocks, np.uint8)
        bpp = 1 if fmt == 1 else 4
        out = np.zeros((g["height"], g["width"] * bpp), np.uint8)
        check(lib().liz_encoder_paint(self.h, _p(blocks), picture, _p(out), out.strides[0], fmt))
        return out if bpp == 1 else out.reshape(g["height"], g["width"], 4)

    def close(self):
        if self.h:
            lib().liz_encoder_free(self.h)
            self.h = None


def b3_hash(data):
    """lizx_b3_hash: BLAKE3 of host bytes (32 bytes)."""
    data = np.ascontiguousarray(np.frombuffer(bytes(data), np.uint8) if isinstance(data, (bytes, bytearray)) else data,
                                np.uint8)
    out = np.zeros(32, np.uint8)
    check(lib().lizx_b3_hash(_p(data) if data.size else None, data.size, _p(out)))
    return out.tobytes()


def b3_root(cvs):
    """lizx_b3_root: a file's root from its chunks' chaining values (chunks x 32 bytes)."""
    cvs = np.ascontiguousarray(cvs, np.uint8).reshape(-1, 32)
    out = np.zeros(32, np.uint8)
    check(lib().lizx_b3_root(_p(cvs), cvs.shape[0], _p(out)))
    return out.tobytes()


def border_kinds(span, version=0, fps=0):
    """lizx_border_kinds: the symbol's border as a (span + 30)^2 uint8 module map (0 picture, 1 light, 2 dark, | 8
    word cells)."""
    L = span + 30
    k = np.zeros((L, L), np.uint8)
    check(lib().lizx_border_kinds(span, version, fps, _p(k)))
    return k


def read_word(span, img, H, nodes, soft=False):
    """lizx_read_word: the format word of a capture (uint8 h x w) through a registration (H 9, nodes [ny, nx, 2]):
    (version, fps), or None where it did not decode; with soft, also the soft values read."""
    img = np.ascontiguousarray(img, np.uint8)
    reg = registration(H, nodes)
    cells, _ = word_geometry(span)
    q = np.zeros(4 * cells, np.float32)
    v, f = C.c_int(0), C.c_int(0)
    got = check(lib().lizx_read_word(span, _p(img), img.shape[1], img.shape[0], img.shape[1], C.byref(reg), _p(q),
                                     C.byref(v), C.byref(f)))
    w = (v.value, f.value) if got else None
    return (w, q) if soft else w


class Decoder:
    """liblizard's decoder handle (liz_decoder_*) and the finish seams."""

    def __init__(self, nmax=0):
        self.h = lib().liz_decoder_new(nmax)
        if not self.h:
            check(-1)

    def close(self):
        if self.h:
            lib().liz_decoder_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def max_blocks(self):
        return check(lib().liz_decoder_max_blocks(self.h))

    def finish_llr(self, version, llr, est):
        """lizard_finish_ext with soft values made elsewhere: per-block verdicts, iterations, estimates, bytes."""
        info = format_info(version)
        B = info["blocks"]
        llr = np.ascontiguousarray(llr, np.int8)
        est = np.ascontiguousarray(est, np.float32)
        blocks = np.zeros(B * LIZ_BLOCK, np.uint8)
        ok = np.zeros(B, np.uint8)
        its = np.zeros(B, np.int8)
        be = np.zeros(B, np.float32)
        good = check(lib().lizx_finish_llr(self.h, version, _p(llr), _p(est), _p(blocks), _p(ok), _p(its), _p(be)))
        return {"good": good, "blocks": blocks.reshape(B, LIZ_BLOCK), "ok": ok, "its": its, "est": be}

    def modulate(self, version, blocks, cnt=0, out=None, timed=False):
        """The sender's pictures on the GPU (lizx_modulate): blocks (frames x blocks x 476 bytes, or one frame's
        blocks x 476) -> a DeviceArray of frames x n x n float32 pictures (or into `out`, a DeviceArray that large);
        with timed, (pictures, the GPU passes' ms)."""
        info = format_info(version)
        blocks = np.ascontiguousarray(blocks, np.uint8).reshape(-1, info["blocks"] * LIZ_BLOCK)
        frames, n = blocks.shape[0], info["n"]
        d = out if out is not None else DeviceArray(frames * n * n * 4)
        ms = C.c_float(0.0)
        check(lib().lizx_modulate(self.h, version, frames, _p(blocks), cnt, d.ptr, C.byref(ms)))
        return (d, ms.value) if timed else d

    def capture(self, version, span, d_picture, w, h, H, k1=0.0, blur=0.0, noise=0.0, seed=0, out=None, fps=None,
                k2=0.0):
        """lizx_capture: a device picture through a camera -> a DeviceArray of h x w uint8 (or into `out`); with fps,
        lizx_capture_symbol: the whole symbol (border, word for (version, fps), anchors, guard) round the picture."""
        cam = Camera(int(w), int(h), (C.c_double * 9)(*np.asarray(H, np.float64).ravel()), float(k1), float(blur),
                     float(noise), int(seed), float(k2))
        d = out if out is not None else DeviceArray(int(w) * int(h))
        if fps is None:
            check(lib().lizx_capture(self.h, version, span, d_picture.ptr, C.byref(cam), d.ptr))
        else:
            check(lib().lizx_capture_symbol(self.h, version, span, int(fps), d_picture.ptr, C.byref(cam), d.ptr))
        return d

    def b3_chunk_cvs(self, d_file, length, chunk_log2, d_cvs=None, timed=False):
        """lizx_b3_chunk_cvs on a DeviceArray file: (chaining values as chunks x 32 uint8, kernel ms if timed); with
        d_cvs given they stay on the device (None returned for them)."""
        chunks = (length + (1 << chunk_log2) - 1) >> chunk_log2
        own = d_cvs is None
        dc = DeviceArray(max(32 * chunks, 32)) if own else d_cvs
        ms = C.c_float(0)
        check(lib().lizx_b3_chunk_cvs(self.h, d_file.ptr, length, chunk_log2, dc.ptr, C.byref(ms)))
        cvs = dc.get(np.uint8, 32 * chunks).reshape(chunks, 32) if own else None
        if own:
            dc.free()
        return (cvs, ms.value) if timed else cvs

    def b3_hash_device(self, d_data, length, timed=False):
        """lizx_b3_hash_device: BLAKE3 of a DeviceArray's first `length` bytes (32 bytes; and kernel ms if timed)."""
        out = np.zeros(32, np.uint8)
        ms = C.c_float(0)
        check(lib().lizx_b3_hash_device(self.h, d_data.ptr, length, _p(out), C.byref(ms)))
        return (out.tobytes(), ms.value) if timed else out.tobytes()

    def find_batch(self, imgs, words=False):
        """lizx_find_batch on same-size GREY8 frames: a list of None or dicts (span, H, nodes) as find gives; with
        words, each dict also carries "version", what the frame's format word names (None where it has none)."""
        imgs = [np.ascontiguousarray(i, np.uint8) for i in imgs]
        F = len(imgs)
        h, w = imgs[0].shape
        assert all(i.shape == (h, w) for i in imgs)
        ptrs = (C.c_void_p * F)(*[i.ctypes.data for i in imgs])
        spans, found = np.zeros(F, np.int32), np.zeros(F, np.int32)
        H = np.zeros((F, 9), np.float32)
        nodes = np.zeros((F, 40 * 40 * 2), np.float32)
        versions = np.zeros(F, np.int32)
        check(lib().lizx_find_batch(self.h, F, ptrs, w, h, w, _p(spans), _p(H), _p(nodes), _p(found),
                                    _p(versions) if words else None))
        out = []
        for f in range(F):
            if not found[f]:
                out.append(None)
                continue
            nx = len(node_lattice(int(spans[f])))
            out.append({"span": int(spans[f]), "H": H[f].copy(),
                        "nodes": nodes[f, :nx * nx * 2].reshape(nx, nx, 2).copy()})
            if words:
                out[-1]["version"] = int(versions[f]) if versions[f] >= 0 else None
        return out

    def frame_slots_device(self, version, d_blocks, frames, cnt, d_slots):
        """lizx_frame_slots_device: device blocks (frames x blocks x 476) -> device slots (frames x subch x 640)."""
        check(lib().lizx_frame_slots_device(self.h, version, frames, d_blocks.ptr, cnt, d_slots.ptr))

    def warp(self, version, span, img, H, nodes):
        img = np.ascontiguousarray(img, np.uint8)
        n = format_info(version)["n"]
        grid = np.zeros((n, n), np.float32)
        reg = registration(H, nodes)
        check(lib().lizx_warp(self.h, version, span, _p(img), img.shape[1], img.shape[0], img.shape[1],
                              C.byref(reg), _p(grid)))
        return grid

    def graphs(self, on):
        """Single-frame decodes replayed as one captured HIP graph (lizx_decoder_graphs; on by default)."""
        check(lib().lizx_decoder_graphs(self.h, 1 if on else 0))

    def shape(self, rows, cols, demap):
        """Waves a workgroup (1, 2 or 4) of the single-frame decode's row FFT, column FFT and demap
        (lizx_decoder_shape; the same bits at every shape)."""
        check(lib().lizx_decoder_shape(self.h, rows, cols, demap))

    def anchors(self, mode):
        """The finder's registration from the anchors outward (lizx_finder_anchors): 0 off, 1 (the default) over the
        finder's own border nodes, 2 with the border's nodes refined by their templates (the reference's chain)."""
        check(lib().lizx_finder_anchors(self.h, int(mode)))

    def border_refine(self, span, img, H, mpx, model=None):
        """The border's marked nodes refined on the GPU (lizx_border_refine): nodes [nx, nx, 4] (dx, dy, score, done)."""
        img = np.ascontiguousarray(img, np.uint8)
        h, w = img.shape
        H = np.ascontiguousarray(H, np.float32).ravel()
        m = None if model is None else np.ascontiguousarray(model, np.float64)
        out = np.zeros(40 * 40 * 4, np.float32)
        nx = check(lib().lizx_border_refine(self.h, span, _p(img), w, h, w, _p(H), float(mpx), _p(m), _p(out)))
        return out[:nx * nx * 4].reshape(nx, nx, 4)

    def anchor_find(self, span, img, H, model=None):
        """The anchors' search on the GPU (lizx_anchor_find): (found, records [49, 8]: X, Y, dx, dy, peak, u, v, ok)."""
        img = np.ascontiguousarray(img, np.uint8)
        h, w = img.shape
        H = np.ascontiguousarray(H, np.float32).ravel()
        m = None if model is None else np.ascontiguousarray(model, np.float64)
        rec = np.zeros((49, 8), np.float32)
        found = check(lib().lizx_anchor_find(self.h, span, _p(img), w, h, w, _p(H), _p(m), _p(rec)))
        return found, rec

    def anchor_register(self, span, img, H, mpx, quad):
        """The registration from the anchors outward (lizx_anchor_register): a dict of have, H (9), quad (8), records
        [49, 8], model (15), moved."""
        img = np.ascontiguousarray(img, np.uint8)
        h, w = img.shape
        H = np.array(H, np.float32).ravel()
        quad = np.array(quad, np.float32).ravel()
        rec = np.zeros((49, 8), np.float32)
        model = np.zeros(15, np.float64)
        moved = C.c_int(0)
        have = check(lib().lizx_anchor_register(self.h, span, _p(img), w, h, w, _p(H), float(mpx), _p(quad), _p(rec),
                                                _p(model), C.byref(moved)))
        return {"have": have, "H": H, "quad": quad, "records": rec, "model": model, "moved": moved.value}

    def hold(self, version, span, H, nodes):
        self._held = registration(H, nodes)
        check(lib().lizx_decoder_hold(self.h, version, span, C.byref(self._held)))

    def release(self):
        """Lets the held registration go: liz_decode then finds each frame's symbol itself."""
        check(lib().lizx_decoder_hold(self.h, 0, 0, None))
        self._held = None

    def find(self, img):
        """lizx_find on a capture (uint8 h x w): None, or a dict of span, H (9), nodes [nx, nx, 2], quad (4, 2), k1,
        track, contrast, rms, points, ms (the stages' host milliseconds). Rows may be padded (a view whose rows are
        contiguous is passed as it is, with its row stride)."""
        img = np.asarray(img, np.uint8)
        if img.ndim != 2:
            raise ValueError("find takes a GREY8 image (h x w)")
        if img.strides[1] != 1 or img.strides[0] < img.shape[1]:
            img = np.ascontiguousarray(img)
        h, w = img.shape[:2]
        nodes = np.zeros((40, 40, 2), np.float32)
        H = np.zeros(9, np.float32)
        span = C.c_int(0)
        quad = np.zeros(8, np.float32)
        info = np.zeros(11, np.float32)
        got = check(lib().lizx_find(self.h, _p(img), w, h, img.strides[0], C.byref(span), _p(H), _p(nodes), _p(quad),
                                    _p(info)))
        if not got:
            return None
        nx = len(node_lattice(span.value))
        return {"span": span.value, "H": H,
                "nodes": nodes.ravel()[:nx * nx * 2].reshape(nx, nx, 2).copy(), "quad": quad.reshape(4, 2),
                "k1": float(info[0]), "track": float(info[1]), "contrast": float(info[2]), "rms": float(info[3]),
                "points": int(info[4]), "ms": info[5:11].copy()}

    def decode(self, img, fmt=1, held=0):
        """liz_decode on a capture (through the held registration): verified blocks, count, liz_decoded."""
        img = np.ascontiguousarray(img, np.uint8)
        h, w = img.shape[:2]
        stride = img.strides[0]
        ver = np.zeros((self.max_blocks() * LIZ_BLOCK,), np.uint8)
        held_c = C.c_int(held)
        out = _Decoded()
        k = check(lib().liz_decode(self.h, _p(img), w, h, stride, fmt, C.byref(held_c), _p(ver), C.byref(out)))
        return k, ver[:k * LIZ_BLOCK].reshape(k, LIZ_BLOCK), out, held_c.value

    def decode_reg(self, version, span, img, H, nodes, soft=True):
        info = format_info(version)
        B, S = info["blocks"], info["subch"] * 320
        img = np.ascontiguousarray(img, np.uint8)
        blocks, ok = np.zeros(B * LIZ_BLOCK, np.uint8), np.zeros(B, np.uint8)
        its, be = np.zeros(B, np.int8), np.zeros(B, np.float32)
        out, sp = {}, None
        if soft:
            out["llr"] = np.zeros(2 * S, np.int8)
            out["sub_est"] = np.zeros(info["subch"], np.float64)
            out["sub_rms"] = np.zeros(info["subch"], np.float32)
            out["coef"] = np.zeros(2 * S, np.float32)
            out["blk_pilot"] = np.zeros(B, np.float32)
            out["soft"] = np.zeros(2 * S, np.float32)
            sp = Soft(_p(out["llr"]), _p(out["sub_est"]), _p(out["sub_rms"]), _p(out["coef"]), (C.c_float * 2)(0, 0),
                      _p(out["blk_pilot"]), _p(out["soft"]))
        reg = registration(H, nodes)
        good = check(lib().lizx_decode_reg(self.h, version, span, _p(img), img.shape[1], img.shape[0], img.shape[1],
                                           C.byref(reg), _p(blocks), _p(ok), _p(its), _p(be),
                                           C.byref(sp) if sp is not None else None))
        out.update({"good": good, "blocks": blocks.reshape(B, LIZ_BLOCK), "ok": ok, "its": its, "est": be})
        if soft:
            out["align"] = (sp.align[0], sp.align[1])
        return out

    def reg_loss_grad(self, version, span, img, H0, params, ref, alpha=1.0, beta=0.0):
        """Refinement loss and its gradient over the 107 registration parameters (lizx_reg_loss_grad)."""
        img = np.ascontiguousarray(img, np.uint8)
        loss, grad = C.c_double(0), np.zeros(REG_PARAMS, np.float64)
        check(lib().lizx_reg_loss_grad(self.h, version, span, _p(img), img.shape[1], img.shape[0], img.shape[1],
                                       _p(np.ascontiguousarray(H0, np.float32).ravel()),
                                       _p(np.ascontiguousarray(params, np.float32)),
                                       _p(np.ascontiguousarray(ref, np.float32)), alpha, beta, C.byref(loss),
                                       _p(grad)))
        return loss.value, grad

    def reg_loss_grad_spectral(self, version, span, img, H0, params, ref, mask, alpha=1.0, beta=0.0):
        """The loss on the masked coefficients and its gradient through the FFT and the warp
        (lizx_reg_loss_grad_spectral); mask: one byte a coefficient, coefficient order."""
        img = np.ascontiguousarray(img, np.uint8)
        loss, grad = C.c_double(0), np.zeros(REG_PARAMS, np.float64)
        check(lib().lizx_reg_loss_grad_spectral(self.h, version, span, _p(img), img.shape[1], img.shape[0],
                                                img.shape[1], _p(np.ascontiguousarray(H0, np.float32).ravel()),
                                                _p(np.ascontiguousarray(params, np.float32)),
                                                _p(np.ascontiguousarray(ref, np.float32)),
                                                _p(np.ascontiguousarray(mask, np.uint8)), alpha, beta, C.byref(loss),
                                                _p(grad)))
        return loss.value, grad

    def reg_refine_spectral(self, version, span, img, H0, params, ref, mask, steps=10, anchors=False):
        """Refinement of h0..h7 and k1 (and with anchors the 49 anchors' offsets) on the spectral loss
        (lizx_reg_refine_spectral): the new parameters and the loss trace."""
        img = np.ascontiguousarray(img, np.uint8)
        p = np.array(params, np.float32)
        trace = np.zeros(steps + 1, np.float32)
        check(lib().lizx_reg_refine_spectral(self.h, version, span, _p(img), img.shape[1], img.shape[0], img.shape[1],
                                             _p(np.ascontiguousarray(H0, np.float32).ravel()), _p(p),
                                             _p(np.ascontiguousarray(ref, np.float32)),
                                             _p(np.ascontiguousarray(mask, np.uint8)), steps, int(anchors),
                                             _p(trace)))
        return p, trace

    def reg_refine(self, version, span, img, H0, params, ref, steps=10, k1=False):
        """Gauss-Newton refinement of h0..h7 (lizx_reg_refine), or of h0..h7 and k1 with k1=True
        (lizx_reg_refine_k1): the new parameters and the loss trace."""
        img = np.ascontiguousarray(img, np.uint8)
        p = np.array(params, np.float32)
        trace = np.zeros(steps + 1, np.float32)
        fn = lib().lizx_reg_refine_k1 if k1 else lib().lizx_reg_refine
        check(fn(self.h, version, span, _p(img), img.shape[1], img.shape[0], img.shape[1],
                 _p(np.ascontiguousarray(H0, np.float32).ravel()), _p(p), _p(np.ascontiguousarray(ref, np.float32)),
                 steps, _p(trace)))
        return p, trace

    def reg_refine_nodes(self, version, span, img, H, nodes, ref, steps=10):
        """Gauss-Newton refinement of H with the node lattice held (lizx_reg_refine_nodes): the new H (9) and the loss
        trace."""
        img = np.ascontiguousarray(img, np.uint8)
        Hn = np.array(H, np.float32).ravel()
        Hn /= Hn[8]
        trace = np.zeros(steps + 1, np.float32)
        check(lib().lizx_reg_refine_nodes(self.h, version, span, _p(img), img.shape[1], img.shape[0], img.shape[1],
                                          _p(Hn), _p(np.ascontiguousarray(nodes, np.float32)),
                                          _p(np.ascontiguousarray(ref, np.float32)), steps, _p(trace)))
        return Hn, trace

    def otf(self, sigma):
        """lizx_decoder_otf: equalise the following decodes by a known Gaussian blur of sigma picture samples (0: none)."""
        check(lib().lizx_decoder_otf(self.h, float(sigma)))

    def finish_grid(self, version, span, grid, soft=True):
        """lizard_finish with a supplied n x n grid (row y, column x) of LIZARD-16V in the ring of this span."""
        info = format_info(version)
        B, S = info["blocks"], info["subch"] * 320
        grid = np.ascontiguousarray(grid, np.float32)
        blocks = np.zeros(B * LIZ_BLOCK, np.uint8)
        ok = np.zeros(B, np.uint8)
        its = np.zeros(B, np.int8)
        be = np.zeros(B, np.float32)
        out = {}
        sp = None
        if soft:
            out["llr"] = np.zeros(2 * S, np.int8)
            out["sub_est"] = np.zeros(info["subch"], np.float64)
            out["sub_rms"] = np.zeros(info["subch"], np.float32)
            out["coef"] = np.zeros(2 * S, np.float32)
            out["blk_pilot"] = np.zeros(B, np.float32)
            out["soft"] = np.zeros(2 * S, np.float32)
            sp = Soft(_p(out["llr"]), _p(out["sub_est"]), _p(out["sub_rms"]), _p(out["coef"]), (C.c_float * 2)(0, 0),
                      _p(out["blk_pilot"]), _p(out["soft"]))
        good = check(lib().lizx_finish_grid(self.h, version, span, _p(grid), _p(blocks), _p(ok), _p(its), _p(be),
                                            C.byref(sp) if sp is not None else None))
        out.update({"good": good, "blocks": blocks.reshape(B, LIZ_BLOCK), "ok": ok, "its": its, "est": be})
        if soft:
            out["align"] = (sp.align[0], sp.align[1])
        return out


class Batch:
    """lizx_batch_*: many frames of mixed formats through one launch sequence (a HIP graph once captured)."""

    def __init__(self, device=-1):
        self.h = lib().lizx_batch_new(device)
        if not self.h:
            check(-1)
        self.frames = 0

    def close(self):
        if self.h:
            lib().lizx_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def plan(self, versions, spans, source=SRC_CAPTURE, stages=STAGE_BLOCKS):
        v = np.ascontiguousarray(versions, np.int32)
        s = np.ascontiguousarray(spans, np.int32)
        check(lib().lizx_batch_plan(self.h, len(v), _p(v), _p(s), source, stages))
        d = np.zeros(4, np.int32)
        check(lib().lizx_batch_dims(self.h, _p(d)))
        self.frames, self.blocks, self.soft_bytes, self.formats = map(int, d)
        return self

    def run(self, frames, mode=RUN_GRAPH):
        """frames: a ctypes array of Frame (make_frames)."""
        check(lib().lizx_batch_run(self.h, frames, mode))

    def sync(self):
        ms = C.c_float(0)
        check(lib().lizx_batch_sync(self.h, C.byref(ms)))
        return ms.value

    def tune(self, frames, reps=3):
        """The autotuner (lizx_batch_tune): the launch shape kept, as (rows, cols, demap waves, LDPC per CU)."""
        ch = np.zeros(4, np.int32)
        check(lib().lizx_batch_tune(self.h, frames, reps, _p(ch)))
        return tuple(int(x) for x in ch)

    def shape(self, choice):
        """Sets the launch shape (4 ints) or, with 6, the shape and the row and column FFTs' radix plans."""
        c = np.ascontiguousarray(choice, np.int32)
        if c.size == 6:
            check(lib().lizx_batch_shape_plans(self.h, _p(c)))
        else:
            check(lib().lizx_batch_shape(self.h, _p(c)))

    def otf(self, sigma):
        """lizx_batch_otf: each frame's known blur (picture samples, 0 none) equalised by the demap; None: none."""
        if sigma is None:
            check(lib().lizx_batch_otf(self.h, None))
        else:
            check(lib().lizx_batch_otf(self.h, _p(np.ascontiguousarray(sigma, np.float32))))

    def tune_plans(self, frames, reps=3):
        """The autotuner over shapes and radix plans (lizx_batch_tune_plans): (rows, cols, demap waves, LDPC per CU,
        row radix, column radix)."""
        ch = np.zeros(6, np.int32)
        check(lib().lizx_batch_tune_plans(self.h, frames, reps, _p(ch)))
        return tuple(int(x) for x in ch)

    def blocks_out(self):
        B = self.blocks
        blocks, ok = np.zeros(B * LIZ_BLOCK, np.uint8), np.zeros(B, np.uint8)
        its, be = np.zeros(B, np.int8), np.zeros(B, np.float32)
        good = check(lib().lizx_batch_blocks(self.h, _p(blocks), _p(ok), _p(its), _p(be)))
        return {"good": good, "blocks": blocks.reshape(B, LIZ_BLOCK), "ok": ok, "its": its, "est": be}

    def soft(self, frame, version):
        info = format_info(version)
        S = info["subch"] * 320
        llr, est = np.zeros(2 * S, np.int8), np.zeros(info["subch"], np.float64)
        rms, al = np.zeros(info["subch"], np.float32), np.zeros(2, np.float32)
        check(lib().lizx_batch_soft(self.h, frame, _p(llr), _p(est), _p(rms), _p(al)))
        return {"llr": llr, "sub_est": est, "sub_rms": rms, "align": al}


def make_frames(items):
    """A ctypes array of Frame from dicts with device arrays: img (DeviceArray), w, h, H (9), nodes (DeviceArray), or
    grid (DeviceArray)."""
    arr = (Frame * len(items))()
    for i, it in enumerate(items):
        f = arr[i]
        if it.get("grid") is not None:
            f.grid = it["grid"].ptr
        if it.get("img") is not None:
            f.img = it["img"].ptr
            f.w, f.h = int(it["w"]), int(it["h"])
            H = np.asarray(it["H"], np.float32).ravel()
            for k in range(9):
                f.H[k] = float(H[k])
            f.nodes = it["nodes"].ptr
    return arr
