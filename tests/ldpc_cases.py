"""Seeded LDPC test words: random info bits, encoded by the oracle, BI-AWGN at an Eb/N0, int8 LLRs in codeword order
(the 7/8 code's known zeros 3840..3967 at +127)."""
import numpy as np

from _libs import OrCode, awgn_llr

_codes = {}


def code(rate):
    if rate not in _codes:
        _codes[rate] = OrCode(rate)
    return _codes[rate]


def word(rate, ebn0_db, rng):
    c = code(rate)
    data = np.zeros(c.k, np.uint8)
    data[:3840] = rng.integers(0, 2, 3840)
    cw = c.encode(data)
    llr = awgn_llr(c.n, c.k, cw, ebn0_db, 3840.0 / c.nt, rng, slice(3840, c.k) if c.k > 3840 else None)
    return cw, llr


def batch(rates, dbs, seed):
    """Concatenated words: rates[i] at dbs[i]; returns (rate array, llr, offsets, codewords)."""
    rng = np.random.default_rng(seed)
    llrs, cws, offs, at = [], [], [], 0
    for r, db in zip(rates, dbs):
        cw, llr = word(r, db, rng)
        llrs.append(llr)
        cws.append(cw)
        offs.append(at)
        at += len(llr)
    return np.array(rates, np.uint8), np.concatenate(llrs), np.array(offs, np.int64), cws
