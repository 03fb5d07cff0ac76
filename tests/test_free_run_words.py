"""The decoder's free-run scan (rows_free_run in csrc/brkga_decoder.hip) gives
each grid row eight lanes: lane j scans the cells [j*S, j*S + S), S = 8 *
ceil(W / 64), eight cells at a time (free cells at the segment start and
end, longest run from a 16-entry table per 4 cells and the run across them),
and three ordered butterfly steps join the eight segments.  This restates
those steps in Python and checks them against the plain longest-zero-run scan
(numba_utils.py:161-188) on random rows of many widths."""
import random


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


def test_mask_bits_to_bytes_multiplier():
    """The stamp spreads four mask bits over the four bytes of a band word with
    one multiply: bit t lands in byte t and the shifted copies never overlap."""
    for b in range(16):
        assert (b * 0x00204081) & 0x01010101 == sum(((b >> t) & 1) << (8 * t) for t in range(4))


def _ctz(v):
    return (v & -v).bit_length() - 1


def _lut(o4):
    return ((0x11122234 if o4 < 8 else 0x01121123) >> (4 * (o4 & 7))) & 0xF


def segment_scan(row):
    W = len(row)
    S = 8 * ((W + 63) // 64)
    lanes = []
    for j in range(8):
        first, pre, best, run = True, 0, 0, 0
        for i in range(S // 8):
            x = j * S + 8 * i
            o = sum(1 << t for t in range(8) if x + t >= W or row[x + t])
            z = o == 0
            lead = _ctz(o | 0x100)
            trail = (8 - o.bit_length()) if o else 8
            lo4, hi4 = o & 15, o >> 4
            across = ((4 - lo4.bit_length()) if lo4 else 4) + _ctz(hi4 | 0x10)
            inner = max(_lut(lo4), _lut(hi4), across)
            if first and not z:
                pre = run + lead
            if not z:
                best = max(best, run + lead, inner)
            run = run + 8 if z else trail
            first = first and z
        if first:
            pre = run
        lanes.append((pre, run, best))
    length = S
    for d in (1, 2, 4):
        joined = []
        for i in range(8):
            a, b = (lanes[i], lanes[i ^ d]) if (i & d) == 0 else (lanes[i ^ d], lanes[i])
            pre = length + b[0] if a[0] == length else a[0]
            suf = length + a[1] if b[0] == length else b[1]
            joined.append((pre, suf, max(a[2], b[2], a[1] + b[0])))
        lanes = joined
        length *= 2
    assert len(set(lanes)) == 1   # every lane of the row ends with the same value
    return max(lanes[0])


def plain_scan(row):
    best = cur = 0
    for v in row:
        cur = 0 if v else cur + 1
        best = max(best, cur)
    return best


def test_segment_scan_matches_plain_scan():
    rng = random.Random(2)
    for _ in range(2000):
        W = rng.choice([1, 2, 3, 7, 8, 9, 31, 63, 64, 65, 250, 300, 350, 400, 500, 511, 512, 513, 700, 1100])
        p = rng.random() ** 2
        row = [1 if rng.random() < p else 0 for _ in range(W)]
        if rng.random() < 0.3:
            a = rng.randrange(W)
            b = rng.randrange(a, W + 1)
            row[a:b] = [0] * (b - a)
        assert segment_scan(row) == plain_scan(row), (W, row)
    for W in (1, 64, 300, 512):   # all free, all occupied
        assert segment_scan([0] * W) == W and segment_scan([1] * W) == 0
