"""The oracle held to the format's own golden vectors (research/SPEC.md, the reference's test programs) and to what the
mathematics fixes. These pin the oracle without running the reference."""
import hashlib

import numpy as np
import pytest

from _libs import OrCode, RATES, oracle, ptr

# research/SPEC.md 7.4: LIZARD-64, block b = id b (uint32 LE) then 472 payload bytes j mod 256
SPEC74 = {
    "crc": [0x5DAC3EBB, 0xD5FBDE46, 0x48EF89B0, 0xC0B8694D, 0x772B50AD, 0xFF7CB050, 0x6268E7A6],
    "bits3800": ["d75dac3ebb", "d7d5fbde46", "d748ef89b0", "d7c0b8694d", "d7772b50ad", "d7ff7cb050", "d76268e7a6"],
    "parity": ["e9266e19f9d8d7cc96857e9f8fc572c9", "e6d991e606272833697a8160703a8d36",
               "e9266e19f9d8d7cc96857e9f8fc57236", "e6d991e606272833697a8160703a8dc9",
               "8e1fee07a8f1080cfcaec24e3654b527", "f19ae6dd86677815a1b9c7f83863c1ef",
               "f284e220c7972b3c6a97444d895fdcc7"],
    "white": ["00003e000ffc03e0f8ffffce000c1c03", "602e8acbffa7681694d5881cc9170bd7",
              "ab494584c3ea397d5a2dac3f3057b275", "4531d3c6175609fb2e6d3dc7950a6a68",
              "3866bc6962628a5c6783a12fef09a41c", "a60df6cf665e475b2d587d0cafb932b1",
              "3d4ee329250c1b091ef45580fdcb3e89"],
    "slots": ["75327e234757d37a0084f28264a9cb87", "111ceac8970cb08c6ceec510afa2dc43",
              "da7b05a78b41e9efaa56a17156f265f1", "3003b3c57ffdd169de5670c9f1bfbdfc",
              "2603840294ee52b278eab0c901015113", "ad2f8bcbc984de716604c8901ac3f399",
              "2b03baef986c0f08c897bc7aaeda824f"],
    "first_slot": [0, 4496, 8992, 13488, 17984, 23136, 28896],
    "tails": {0: ["5a69", "6e57", "80c1", "bc79", "c22fe70f"], 1: ["a596", "6e57", "7f3e", "bc79", "3dd018f0"],
              2: ["5a69", "91a8", "80c1", "4386", "c22fe70f"], 3: ["a596", "91a8", "7f3e", "4386", "3dd018f0"]},
    "cw_sha": ["a62fac54602b23b03e30eb47aaa7f8e983f7d19bc8c7e13713eba2640a23c088",
               "152ae7463835c31782ca4980d5c592c817b1b66a3a7b976c47fd0f5f436a6700",
               "602e5bbe31440df25be5562e371c15eb771046800dd05375348c84bf7b182e07",
               "3fda58ac7bd787b16da65e3da29ea12528f98a3a8bf41020f8fffd30517d95d4",
               "3be83a78fef345f13645e840469519501a7b01da4576f78d8a5b5713d7fd7035",
               "958111ce99624a166d811a1f75c056f3a09127a58238d7d84c3b4abe28886671",
               "2925aaea9a62cbe2c1b4b6c546071eaf2e96a5cd70b9edcb5e4de979155609b8"],
    "frame_sha": ["8cd51a40531864b685d052b1879df24dcb7bb9700e9379e33b937ee45a4ed947",
                  "e5bbfe8c31afa6ba38bfa27dbd3c621f1d51f8982f6be2db08dda65231b117b5",
                  "26d07dc9a5636e452679c3429d31b7ea3b7f0a11064dc5831dd5571552232ebe",
                  "1e08c661206a2133065b6637cd8a6df727c4d32256f35172274bd722f3a96712"],
    "padding128": "332c1ead5f960d5aaf4c1e8d9f9e7cd8",
}
BLOCK_RATES = [6, 6, 6, 6, 4, 3, 2]


def hexbits(bits):
    return np.packbits(np.asarray(bits, np.uint8)).tobytes().hex()


def spec74_blocks():
    blocks = np.zeros((7, 476), np.uint8)
    for b in range(7):
        blocks[b, 0] = b
        blocks[b, 4:] = np.arange(472) & 255
    return blocks


def test_crc_check_value():
    # RFC 3720 / SPEC 7.1 check value; liblizard/test/crc_check.c:21-22
    o = oracle()
    s = np.frombuffer(b"123456789", np.uint8).copy()
    assert o.or_crc32c(ptr(s), 9) == 0xE3069283
    z = np.zeros(1, np.uint8)
    assert o.or_crc32c(ptr(z), 0) == 0


def test_crc_linearity_against_bit_loop():
    # a table-free statement of CRC-32C on random buffers: crc(a ^ b) = crc(a) ^ crc(b) ^ crc(0) for equal lengths
    o = oracle()
    rng = np.random.default_rng(1)
    for n in (1, 7, 64, 476):
        a, b = rng.integers(0, 256, n).astype(np.uint8), rng.integers(0, 256, n).astype(np.uint8)
        z = np.zeros(n, np.uint8)
        c = np.bitwise_xor(a, b)
        assert o.or_crc32c(ptr(c), n) == o.or_crc32c(ptr(a), n) ^ o.or_crc32c(ptr(b), n) ^ o.or_crc32c(ptr(z), n)


def test_whitening_prefix():
    o = oracle()
    w = np.zeros(40960, np.uint8)
    o.or_whiten(ptr(w), w.size)
    assert hexbits(w[:128]) == SPEC74["white"][0]
    # the sequence starts with 18 zeros then 5 ones (SPEC 7.3)
    assert list(w[:23]) == [0] * 18 + [1] * 5


@pytest.mark.parametrize("nt,st,inv", [(4480, 1711, 2511), (5120, 1957, 2093), (5760, 2201, 3881), (7680, 2933, 1757)])
def test_perm_strides(nt, st, inv):
    o = oracle()
    assert o.or_perm_stride(nt) == st
    assert (st * inv) % nt == 1
    perm = np.zeros(nt, np.int32)
    o.or_perm(ptr(perm), nt)
    assert sorted(perm) == list(range(nt))


def test_perm_34_prefix():
    o = oracle()
    perm = np.zeros(5120, np.int32)
    o.or_perm(ptr(perm), 5120)
    assert list(perm[:6]) == [0, 1957, 3914, 751, 2708, 4665]


@pytest.mark.parametrize("rate,h,n,k,edges", [(6, 0x6f67223a, 4608, 3968, 14463), (4, 0x3f726d8e, 5120, 3840, 19455),
                                              (3, 0x27d3a72e, 5760, 3840, 22527), (2, 0x721499de, 7680, 3840, 26879)])
def test_code_hash(rate, h, n, k, edges):
    # adopted FNV-1a hashes (liblizard/test/ldpc_table.c:19; SPEC 8.1, 8.6)
    c = OrCode(rate)
    assert (c.n, c.k, c.edges) == (n, k, edges)
    assert c.hash() == h


def _spec74_codewords():
    o = oracle()
    blocks = spec74_blocks()
    out = []
    for b in range(7):
        c = OrCode(BLOCK_RATES[b])
        data = np.zeros(c.k, np.uint8)
        data[:3808] = np.unpackbits(blocks[b])
        crc = o.or_crc32c(ptr(blocks[b]), 476)
        data[3808:3840] = [(crc >> (31 - i)) & 1 for i in range(32)]
        out.append((c, crc, c.encode(data)))
    return out


def test_spec74_codewords():
    """CRCs and bits 3800..3839 of every block, parity p0..p127 of the 7/8 blocks. SPEC 7.4's parity rows for the
    3/4, 2/3 and 1/2 blocks and its whole-codeword SHA-256 lines disagree with its own frame digest for c = 0 (which
    covers those blocks' slots and which these codewords reproduce, test_spec74_frame_slots) and with the reference's
    own encoder (ldpc_encode, compiled in oracle/_ref); the whole codewords are held to that encoder in
    test_oracle_ref.py."""
    for b, (c, crc, cw) in enumerate(_spec74_codewords()):
        assert crc == SPEC74["crc"][b]
        assert hexbits(cw[3800:3840]) == SPEC74["bits3800"][b]
        if c.rate == 6:
            assert hexbits(cw[c.k:c.k + 128]) == SPEC74["parity"][b]
        if c.k > 3840:
            assert not cw[3840:c.k].any()


def test_codewords_satisfy_H():
    # every encoded word decodes in one iteration at full confidence with its bits unchanged (the syndrome is tested
    # after the first iteration: ldpc.c:1143-1148)
    rng = np.random.default_rng(4)
    for rate in RATES:
        c = OrCode(rate)
        data = rng.integers(0, 2, c.k).astype(np.uint8)
        cw = c.encode(data)
        its, bits = c.decode(np.where(cw, -80, 80).astype(np.int8), 5, 0)
        assert its == 1 and np.array_equal(bits, cw)


@pytest.mark.parametrize("cnt", [0, 1, 2, 3])
def test_spec74_frame_slots(cnt):
    o = oracle()
    f = o.or_format_new(384, 64)
    blocks = spec74_blocks()
    slots = np.zeros(40960, np.uint8)
    o.or_frame_slots(f, ptr(blocks), cnt, ptr(slots))
    # SPEC 7.4's frame digest for c = 0; its c = 1..3 digests are not reproduced by the reference's own encoder
    # (lizard_encode, compiled in oracle/_ref), which test_oracle_ref.py holds these slots to. Every c's pilot tails
    # below are SPEC's.
    if cnt == 0:
        assert hashlib.sha256(np.packbits(slots).tobytes()).hexdigest() == SPEC74["frame_sha"][0]
    w = np.zeros(40960, np.uint8)
    o.or_whiten(ptr(w), 40960)
    nslots = [4496] * 4 + [5152, 5760, 7680]
    nt = [4480] * 4 + [5120, 5760, 7680]
    for b in range(7):
        s0 = SPEC74["first_slot"][b]
        assert hexbits(w[s0:s0 + 128]) == SPEC74["white"][b]
        assert hexbits(slots[s0:s0 + 128]) == SPEC74["slots"][b]
        if nslots[b] > nt[b]:
            tail = slots[s0 + nt[b]:s0 + nslots[b]]
            want = SPEC74["tails"][cnt][b if b < 4 else 4]
            assert hexbits(tail) == want, (b, cnt)
    assert hexbits(slots[36576:36576 + 128]) == SPEC74["padding128"]
    o.or_format_free(f)


# SPEC 3.1's profile table: (subch, 7/8, 3/4, 2/3, 1/2, slots padded)
TIERS = [(64, 4, 1, 1, 1, 4384), (80, 4, 3, 1, 1, 4320), (128, 7, 4, 2, 2, 2960), (256, 14, 9, 4, 4, 768),
         (432, 24, 14, 7, 7, 2368), (512, 28, 15, 9, 9, 3552), (560, 31, 19, 9, 9, 176), (592, 33, 18, 10, 10, 3376),
         (656, 36, 21, 11, 11, 1952), (1024, 57, 33, 17, 17, 592)]


@pytest.mark.parametrize("subch,a,m,t,c,pad", TIERS)
def test_tiers(subch, a, m, t, c, pad):
    o = oracle()
    rate, blocks = np.zeros(4, np.int32), np.zeros(4, np.int32)
    assert o.or_tiers_for(subch, ptr(rate), ptr(blocks)) == 4
    assert list(rate) == [6, 4, 3, 2] and list(blocks) == [a, m, t, c]
    used = a * 4496 + m * 5152 + t * 5760 + c * 7680
    assert subch * 640 - used == pad


def test_tiers_ladder_facts():
    # SPEC 3.1: padding from none (LIZARD-320 and -1008) to 4976 (LIZARD-768); blocks under 2 V, never falling
    o = oracle()
    pads, last = {}, 0
    for v in range(4, 65):
        rate, blocks = np.zeros(4, np.int32), np.zeros(4, np.int32)
        assert o.or_tiers_for(16 * v, ptr(rate), ptr(blocks)) == 4
        B = int(blocks.sum())
        assert B < 2 * v and B >= last
        last = B
        pads[16 * v] = 16 * v * 640 - int(np.dot(blocks, [4496, 5152, 5760, 7680]))
    assert pads[320] == 0 and pads[1008] == 0 and max(pads.values()) == pads[768] == 4976


def test_n_for():
    o = oracle()
    want = {384: (4, 5), 512: (6, 8), 768: (9, 20), 1024: (21, 35), 1152: (36, 45), 1280: (46, 55), 1536: (56, 64)}
    for n, (lo, hi) in want.items():
        for v in range(lo, hi + 1):
            assert o.or_n_for(16 * v) == n, v


def _uv(pos, n):
    bw = min(n // 2, 64)
    v = (pos // (n * bw)) * bw + pos % bw
    uw = (pos // bw) % n
    u = np.where(uw < n // 2, uw, uw - n)
    return u, v


def test_coefficient_order():
    # SPEC 6.3: the first 24 (u, v) and the last coefficient of some sub-channel counts; SPEC 6.12 P1 decode
    o = oracle()
    first24 = [(1, 0), (0, 1), (-1, 1), (1, 1), (2, 0), (0, 2), (-2, 1), (2, 1), (-1, 2), (1, 2), (-2, 2), (2, 2),
               (3, 0), (0, 3), (-3, 1), (3, 1), (-1, 3), (1, 3), (-3, 2), (3, 2), (-2, 3), (2, 3), (4, 0), (0, 4)]
    pos = np.zeros(1024 * 320, np.int32)
    o.or_pos(1536, 1024, ptr(pos))
    u, v = _uv(pos, 1536)
    assert list(zip(u[:24].tolist(), v[:24].tolist())) == first24
    last = {1: (9, 11), 2: (9, 18), 8: (20, 35), 16: (57, 3), 64: (6, 114), 128: (145, 71), 256: (218, 68),
            512: (25, 322), 1024: (26, 456)}
    for s, uv in last.items():
        assert (int(u[320 * s - 1]), int(v[320 * s - 1])) == uv
    # block 0 of every format runs to C_2247, (37, 8)
    assert (int(u[2247]), int(v[2247])) == (37, 8)
    # the order does not depend on n (SPEC 3.2): a format's coefficients are the same at every n that holds it
    pos2 = np.zeros(64 * 320, np.int32)
    o.or_pos(384, 64, ptr(pos2))
    u2, v2 = _uv(pos2, 384)
    assert np.array_equal(u2, u[:64 * 320]) and np.array_equal(v2, v[:64 * 320])
    # sorted by radius, then v, then u, and each once
    r2 = u.astype(np.int64) ** 2 + v.astype(np.int64) ** 2
    key = r2 * 10 ** 8 + v * 10 ** 4 + (u + 5000)
    assert np.all(np.diff(key) > 0)


def test_lizard656_block_starts():
    # SPEC 7.2
    o = oracle()
    f = o.or_format_new(1152, 656)
    info = np.zeros(6, np.int32)
    o.or_format_info(f, ptr(info))
    B = int(info[2])
    rate, s0, ns = np.zeros(B, np.int32), np.zeros(B, np.int32), np.zeros(B, np.int32)
    o.or_format_blocks(f, ptr(rate), ptr(s0), ptr(ns))
    assert s0[0] == 0 and s0[35] == 157360 and s0[36] == 161856 and s0[56] == 264896
    assert s0[57] == 270048 and s0[67] == 327648 and s0[68] == 333408 and s0[78] == 410208
    assert info[5] == 417888 and 656 * 640 - 1 == 419839
    o.or_format_free(f)
