"""BLAKE3 (hash mode, 32-byte output) in plain Python, written from the published specification
(O'Connor, Aumasson, Neves, Wilcox-O'Hearn, "BLAKE3: one function, fast everywhere", 2020: the
compression function, 1024-byte chunks of 64-byte blocks, the binary tree whose left subtree takes
the largest power-of-two number of chunks).  A second, independently structured restatement
(recursive tree, no chaining-value stack) that the oracle's C BLAKE3 is checked against.  Test
infrastructure only."""
import struct

IV = [0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A, 0x510E527F, 0x9B05688C, 0x1F83D9AB, 0x5BE0CD19]
PERM = [2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8]
CHUNK_START, CHUNK_END, PARENT, ROOT = 1, 2, 4, 8
M32 = 0xFFFFFFFF


def _rotr(x, n):
    return ((x >> n) | (x << (32 - n))) & M32


def _g(s, a, b, c, d, x, y):
    s[a] = (s[a] + s[b] + x) & M32
    s[d] = _rotr(s[d] ^ s[a], 16)
    s[c] = (s[c] + s[d]) & M32
    s[b] = _rotr(s[b] ^ s[c], 12)
    s[a] = (s[a] + s[b] + y) & M32
    s[d] = _rotr(s[d] ^ s[a], 8)
    s[c] = (s[c] + s[d]) & M32
    s[b] = _rotr(s[b] ^ s[c], 7)


def compress(cv, words, counter, block_len, flags):
    s = list(cv) + IV[:4] + [counter & M32, (counter >> 32) & M32, block_len, flags]
    m = list(words)
    for r in range(7):
        _g(s, 0, 4, 8, 12, m[0], m[1]); _g(s, 1, 5, 9, 13, m[2], m[3])
        _g(s, 2, 6, 10, 14, m[4], m[5]); _g(s, 3, 7, 11, 15, m[6], m[7])
        _g(s, 0, 5, 10, 15, m[8], m[9]); _g(s, 1, 6, 11, 12, m[10], m[11])
        _g(s, 2, 7, 8, 13, m[12], m[13]); _g(s, 3, 4, 9, 14, m[14], m[15])
        if r < 6:
            m = [m[PERM[i]] for i in range(16)]
    return [s[i] ^ s[i + 8] for i in range(8)] + [s[i + 8] ^ cv[i] for i in range(8)]


def _words(block):
    return list(struct.unpack("<16I", block.ljust(64, b"\0")))


def _chunk_output(data, index):
    """(cv, words, counter, len, flags) of the chunk's last block, not yet compressed."""
    cv = IV
    blocks = [data[i:i + 64] for i in range(0, len(data), 64)] or [b""]
    for k, blk in enumerate(blocks):
        flags = (CHUNK_START if k == 0 else 0) | (CHUNK_END if k == len(blocks) - 1 else 0)
        if k == len(blocks) - 1:
            return (cv, _words(blk), index, len(blk), flags)
        cv = compress(cv, _words(blk), index, 64, flags)[:8]


def _node(data, index):
    if len(data) <= 1024:
        return _chunk_output(data, index)
    n = (len(data) + 1023) // 1024
    left = 1
    while left * 2 < n:
        left *= 2
    lo = _node(data[:1024 * left], index)
    hi = _node(data[1024 * left:], index + left)
    lcv = compress(*lo)[:8]
    rcv = compress(*hi)[:8]
    return (IV, lcv + rcv, 0, 64, PARENT)


def blake3(data):
    cv, words, counter, blen, flags = _node(bytes(data), 0)
    out = compress(cv, words, counter, blen, flags | ROOT)[:8]
    return struct.pack("<8I", *out)
