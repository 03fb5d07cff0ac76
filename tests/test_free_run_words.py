"""The decoder's free-run scan (row_free_run in csrc/brkga_decoder.hip) works
on blocks of 512 cells: each of 64 lanes summarises 8 cells (free cells at
the low and high end, longest free run from a 16-entry table per 4 cells and
the run across them, all-free flag),
a ballot of the all-free lanes joins runs across lanes, and a carry joins
blocks.  This restates those steps in Python and checks them against the
plain longest-zero-run scan (numba_utils.py:161-188) on random rows,
including rows that need several blocks."""
import random

M64 = (1 << 64) - 1


def _maxrun4(o):
    best = cur = 0
    for j in range(4):
        cur = 0 if (o >> j) & 1 else cur + 1
        best = max(best, cur)
    return best


def test_pattern_table_constants():
    """The kernel's two table words hold the longest free run of each 4-cell pattern."""
    lo = sum(_maxrun4(o) << (4 * o) for o in range(8))
    hi = sum(_maxrun4(o + 8) << (4 * o) for o in range(8))
    assert (lo, hi) == (0x11122234, 0x01121123)


def _ctz(v):
    return (v & -v).bit_length() - 1


def lane_block_scan(row):
    W, best, carry = len(row), 0, 0
    table = (0x11122234, 0x01121123)

    def lut(o4):
        return (table[o4 >> 3] >> (4 * (o4 & 7))) & 0xF

    for c0 in range(0, W, 512):
        pre, suf, b, full = [], [], [], []
        for lane in range(64):
            x = c0 + 8 * lane
            o = 0xFF
            if x < W:
                o = sum(1 << j for j in range(8) if x + j >= W or row[x + j])
            lo4, hi4 = o & 0xF, o >> 4
            pre.append(_ctz(o | 0x100))
            suf.append(8 - o.bit_length() if o else 8)
            across = (4 - lo4.bit_length() if lo4 else 4) + _ctz(hi4 | 0x10)
            b.append(max(lut(lo4), lut(hi4), across))
            full.append(o == 0)
        F = sum(1 << lane for lane in range(64) if full[lane])
        for lane in range(64):
            rest = F >> (lane + 1) if lane < 63 else 0
            c = min(_ctz(~rest & M64), 63 - lane)
            nxt = lane + 1 + c
            if not full[lane] and nxt <= 63:
                b[lane] = max(b[lane], suf[lane] + 8 * c + pre[nxt])
        assert max(b) < 1024   # the kernel's 10-bit ballot search
        best = max(best, max(b))
        if F == M64:
            carry += 512
            continue
        nf = ~F & M64
        f = _ctz(nf)
        best = max(best, carry + 8 * f + pre[f])
        last = nf.bit_length() - 1
        carry = suf[last] + 8 * (63 - last)
    return max(best, carry)


def plain_scan(row):
    best = cur = 0
    for v in row:
        cur = 0 if v else cur + 1
        best = max(best, cur)
    return best


def test_lane_block_scan_matches_plain_scan():
    rng = random.Random(2)
    for _ in range(1500):
        W = rng.choice([1, 2, 3, 7, 8, 9, 31, 63, 64, 65, 250, 300, 500, 511, 512, 513, 700, 1100])
        p = rng.random() ** 2
        row = [1 if rng.random() < p else 0 for _ in range(W)]
        if rng.random() < 0.3:
            a = rng.randrange(W)
            b = rng.randrange(a, W + 1)
            row[a:b] = [0] * (b - a)
        assert lane_block_scan(row) == plain_scan(row), (W, row)
