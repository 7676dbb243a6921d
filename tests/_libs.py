"""ctypes loaders for the checkers: the oracle (C restatement) and the compiled reference driver (oracle/_ref).

Test infrastructure only. Both are checkers; the product library is loaded through lizard_amd.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SCALAR_SO = os.path.join(ROOT, "oracle", "_ref", "libref_scalar.so")
REF_X86_SO = os.path.join(ROOT, "oracle", "_ref", "libref_x86.so")
REF_XFER_SO = os.path.join(ROOT, "oracle", "_ref", "libref_xfer.so")

P = C.c_void_p
I32 = C.c_int
F32 = C.c_float


def ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _sig(lib, name, res, *args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = list(args)


_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        lib = C.CDLL(ORACLE_SO)
        _sig(lib, "or_crc32c", C.c_uint32, P, I32)
        _sig(lib, "or_whiten", None, P, I32)
        _sig(lib, "or_perm_stride", I32, I32)
        _sig(lib, "or_perm", None, P, I32)
        _sig(lib, "or_block_slots", I32, I32)
        _sig(lib, "or_sent_bits", I32, I32)
        _sig(lib, "or_tiers_for", I32, I32, P, P)
        _sig(lib, "or_n_for", I32, I32)
        _sig(lib, "or_pos", I32, I32, I32, P)
        _sig(lib, "or_code_new", P, I32)
        _sig(lib, "or_code_free", None, P)
        _sig(lib, "or_code_dims", None, P, P)
        _sig(lib, "or_code_hash", C.c_uint32, P)
        _sig(lib, "or_encode", None, P, P, P)
        _sig(lib, "or_decode", I32, P, P, P, I32, I32, F32, P, P)
        _sig(lib, "or_format_new", P, I32, I32)
        _sig(lib, "or_format_free", None, P)
        _sig(lib, "or_format_info", None, P, P)
        _sig(lib, "or_format_blocks", None, P, P, P, P)
        _sig(lib, "or_format_pos", C.POINTER(C.c_int), P)
        _sig(lib, "or_finish_llr", I32, P, P, P, P, P, P, P)
        _sig(lib, "or_frame_slots", None, P, P, I32, P)
        _oracle = lib
    return _oracle


_refs = {}


def ref(form="scalar"):
    """The compiled reference driver, or None where oracle/_ref was not built."""
    if form in _refs:
        return _refs[form]
    path = REF_SCALAR_SO if form == "scalar" else REF_XFER_SO if form == "xfer" else REF_X86_SO
    if not os.path.exists(path):
        _refs[form] = None
        return None
    lib = C.CDLL(path)
    if form == "xfer":  # the transfer's BLAKE3 tree (liblizard/src/xfer.c) over the vendored BLAKE3
        _sig(lib, "xfer_b3_hash", None, P, C.c_size_t, P)
        _sig(lib, "xfer_chunk_cv", I32, P, C.c_size_t, C.c_uint32, I32, P)
        _sig(lib, "xfer_root", I32, P, C.c_uint32, P)
        _sig(lib, "stream_fill", None, C.c_uint32, P, C.c_uint32)
        _refs[form] = lib
        return lib
    _sig(lib, "ref_crc32c", C.c_uint32, P, I32)
    _sig(lib, "ref_ldpc_dims", I32, I32, P)
    _sig(lib, "ref_ldpc_encode", None, I32, P, P)
    _sig(lib, "ref_ldpc_decode", I32, I32, P, P, I32, I32, F32, P, P)
    _sig(lib, "ref_ldpc_decode_many", I32, I32, I32, P, P, P, I32, I32, F32)
    _sig(lib, "ref_format_new", P, I32, I32, I32)
    _sig(lib, "ref_format_free", None, P)
    _sig(lib, "ref_set_fps", I32, P, I32)
    _sig(lib, "ref_format_info", None, P, P)
    _sig(lib, "ref_format_pos", C.POINTER(C.c_int), P)
    _sig(lib, "ref_format_blocks", None, P, P, P, P)
    _sig(lib, "ref_format_tables", None, P, P)
    _sig(lib, "ref_format_erase", None, P, P, P, P)
    _sig(lib, "ref_finish_llr", I32, P, P, P, P, P, P, P)
    _sig(lib, "ref_finish_grid", I32, P, P, P, P, P, P, P, P, P, P, P, P)
    _sig(lib, "ref_encode", I32, P, P, I32, P, P)
    _sig(lib, "ref_paint_grey", None, P, P, P)
    _sig(lib, "ref_encoded_picture", I32, P, P)
    _sig(lib, "ref_rs_encode", None, P, I32, I32)
    _sig(lib, "ref_rs_decode", I32, P, I32, I32, P, I32)
    _sig(lib, "ref_fmt_encode", I32, I32, I32, P, I32)
    _sig(lib, "ref_fmt_decode", I32, P, I32, I32, P, P)
    _sig(lib, "ref_fmt_geom", I32, I32, P, P)
    _sig(lib, "ref_layout_kinds", I32, I32, P, P)
    _sig(lib, "ref_acquire", I32, P, P, I32, I32, P, P, P, P)
    _sig(lib, "ref_anchor_m0", I32, I32, I32)
    _sig(lib, "ref_anchor_side", I32, I32)
    _sig(lib, "ref_read_word", I32, P, P, I32, I32, P, P, P, P, P, P)
    _sig(lib, "ref_sample_grid", I32, P, P, I32, I32, P, P, I32, P)
    _sig(lib, "ref_node_lattice", I32, P, P, P, P)
    _sig(lib, "ref_decode_reg", I32, P, P, I32, I32, P, P, P, P, P, P, P, P)
    _sig(lib, "ref_border_fill_axis", None, P, P, I32, I32, F32, F32, P, P)
    _sig(lib, "ref_lensm_at", None, I32, P, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, P, P)
    _sig(lib, "ref_border_fill_anchors", None, P, P, I32, I32, F32, F32, P, P, P, F32)
    _refs[form] = lib
    return lib


RATES = (6, 4, 3, 2)  # rate indices of 7/8, 3/4, 2/3, 1/2


class OrCode:
    def __init__(self, rate):
        self.lib = oracle()
        self.h = self.lib.or_code_new(rate)
        d = np.zeros(10, np.int32)
        self.lib.or_code_dims(self.h, ptr(d))
        (self.n, self.k, self.m, self.z, self.mb, self.kb, self.np, self.nt, self.edges, self.slots_total) = map(int, d)
        self.rate = rate

    def __del__(self):
        try:
            self.lib.or_code_free(self.h)
        except Exception:
            pass

    def hash(self):
        return int(self.lib.or_code_hash(self.h))

    def encode(self, data):
        data = np.ascontiguousarray(data, np.uint8)
        cw = np.zeros(self.n, np.uint8)
        self.lib.or_encode(self.h, ptr(data), ptr(cw))
        return cw

    def decode(self, llr, max_iter=30, stall_it=9, ratio=0.95, trace=False):
        llr = np.ascontiguousarray(llr, np.int8)
        out = np.zeros(self.n, np.uint8)
        L = np.zeros(self.n, np.int16) if trace else None
        R = np.zeros(self.slots_total * 128, np.int8) if trace else None
        its = self.lib.or_decode(self.h, ptr(llr), ptr(out), max_iter, stall_it, ratio, ptr(L), ptr(R))
        return (its, out, L, R) if trace else (its, out)


def ref_decode(lib, rate, llr, max_iter=30, stall_it=9, ratio=0.95, trace=False):
    d = np.zeros(11, np.int32)
    lib.ref_ldpc_dims(rate, ptr(d))
    n, slots_total = int(d[0]), int(d[9])
    llr = np.ascontiguousarray(llr, np.int8)
    out = np.zeros(n, np.uint8)
    L = np.zeros(n, np.int16) if trace else None
    R = np.zeros(slots_total * 128, np.int8) if trace else None
    its = lib.ref_ldpc_decode(rate, ptr(llr), ptr(out), max_iter, stall_it, ratio, ptr(L), ptr(R))
    return (its, out, L, R) if trace else (its, out)


def awgn_llr(code_n, k, cw, ebn0_db, rate_r, rng, known_zero=None):
    """BI-AWGN at Eb/N0 (dB): y = +-1 + sigma g, sigma^2 = 1 / (2 R Eb/N0); LLR = 2 y / sigma^2; int8 =
    clamp(round_half_even(8 LLR), +-80) (liblizard/test/ldpc_ref.c:105-108, ldpc_awgn.c:64-76). Unsent/known bits
    (known_zero: a slice) are set to +127."""
    ebn0 = 10.0 ** (ebn0_db / 10.0)
    sigma = np.sqrt(1.0 / (2.0 * rate_r * ebn0))
    y = (1.0 - 2.0 * cw.astype(np.float64)) + sigma * rng.standard_normal(code_n)
    llr = (2.0 * y / (sigma * sigma)).astype(np.float32)
    q = np.clip(np.rint(8.0 * llr.astype(np.float64)), -80, 80).astype(np.int8)
    if known_zero is not None:
        q[known_zero] = 127
    return q
