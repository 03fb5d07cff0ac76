"""The decoder's bit-row correlation (csrc/brkga_decoder.hip): grids and LDS
bands hold one bit per cell, MFMA operands are expanded from 32-bit words with
the k order k = 16g + 4r + j <-> column 8j + 4g + r (lane group g, VGPR r, byte
j) in both A and B (frag_from_word, bfrag_from_bits), windows are staged as
16-byte-aligned word runs into rows of an odd stride (band_stride, test_rotation)
and every lane reads its A words at loff + t0 + j (corr_pass).  This restates
those steps in Python on random grids and masks and checks, per candidate and
position, that "overlap count == 0" matches the brute-force overlap; it also
restates the stamp's word extraction (stamp_part) and the bit-row free-run scan
(rows_free_run, longest_free_in_word) against plain scans.  The fp4 path
(nib_frag: 32x32x64 e2m1 MFMAs over 64-column chunks, a lane's word stream
and mask window offset by its k-group) is restated the same way."""
import random

import numpy as np

WORD = 32


def band_stride(tw, nc):
    return (tw + nc + 5) | 1


def qbits_row(mask_row, nc, nq=None):
    """The mask row's bits after 32 zero bits, nq (default nc + 1) words, zero
    padded (brkga_create keeps 2 nc4 + 2 words per row)."""
    q = [0] * (nq or nc + 1)
    for b, v in enumerate(mask_row):
        if v:
            q[(b + 32) // 32] |= 1 << ((b + 32) % 32)
    return q


def frag_bytes(y):
    """frag_from_word: VGPR r byte j = bit 8j + r of y, weight 2^r."""
    return [[((y >> (8 * j + r)) & 1) << r for j in range(4)] for r in range(4)]


def lane_k_values(y):
    """The lane's 16 operand values in MFMA k order (k = 4r + j within its group)."""
    fb = frag_bytes(y)
    return [fb[r][j] for r in range(4) for j in range(4)]


def mfma_tile(a_words, b_words):
    """int32 32x32x32 MFMA from the lanes' words: a_words[m][g], b_words[n][g] are
    the 32-bit words lane (row/col, group g) expands; returns D[m][n]."""
    D = np.zeros((32, 32), np.int64)
    for m in range(32):
        av = lane_k_values(a_words[m][0] >> 0) + lane_k_values(a_words[m][1])
        for n in range(32):
            bv = lane_k_values(b_words[n][0]) + lane_k_values(b_words[n][1])
            D[m, n] = sum(x * y for x, y in zip(av, bv))
    return D


def test_fragment_order_is_a_bijection_of_the_chunk_columns():
    cols = sorted(8 * j + 4 * g + r for g in range(2) for r in range(4) for j in range(4))
    assert cols == list(range(32))


def test_tile_counts_match_brute_force_overlap():
    """One 32x32 tile: A = 32 grid rows' word (the lane's word >> 4g), B = the
    64-bit window of the mask row's bits >> (32 - n + 4g); zero counts are the
    zero overlaps of the placements x0 + n, for every chunk of a wide mask."""
    rng = random.Random(3)
    for _ in range(12):
        w = rng.randint(1, 90)
        nc = (w + 62) // 32
        row = [1 if rng.random() < 0.6 else 0 for _ in range(w)]
        row[rng.randrange(w)] = 1
        q = qbits_row(row, nc)
        width = 32 * (nc + 1)
        grid = [[1 if rng.random() < rng.choice([0.02, 0.2]) else 0 for _ in range(width)] for _ in range(32)]
        total = np.zeros((32, 32), np.int64)
        for c in range(nc):
            a_words, b_words = [], []
            for m in range(32):
                wd = sum(grid[m][32 * c + i] << i for i in range(32))
                a_words.append((wd >> 0, wd >> 4))
            for n in range(32):
                win = (q[c + 1] << 32) | q[c]
                b_words.append(tuple((win >> (32 - n + 4 * g)) & 0xFFFFFFFF for g in range(2)))
            total += mfma_tile(a_words, b_words)
        for m in range(32):
            for n in range(32):
                ov = sum(row[b] * grid[m][n + b] for b in range(w))
                assert (total[m, n] == 0) == (ov == 0), (w, m, n)


def test_staged_window_addressing():
    """test_rotation's staging (words c0 = tw0 & ~3 onward, q4 vectors, odd stride)
    and corr_pass's word address loff + t0 + t + c for folded segments and
    rounded-up passes read exactly the grid word of column x0 + 32(t + c) of row
    y + a, or (past the window) a word inside the same band row."""
    rng = random.Random(7)
    for _ in range(300):
        W = rng.randint(40, 700)
        w = rng.randint(1, min(W, 200))
        nc = (w + 62) // 32
        T = (W - w + 1 + 31) // 32
        pw = max(4, T + nc - 1, (W + 31) // 32)
        pw = (pw + 3) & ~3
        TW = rng.randint(1, T)
        stride = band_stride(TW, nc)
        for tw0 in range(0, T, TW):
            tl = min(TW, T - tw0)
            c0 = tw0 & ~3
            q4 = (((tw0 + tl + nc - 1 + 3) & ~3) - c0) >> 2
            assert c0 + 4 * q4 <= pw and 4 * q4 <= stride
            for rl in range(6):
                nseg = 32 >> rl
                if nseg > tl:
                    continue
                tf = (tl + nseg - 1) >> (5 - rl)
                for seg in range(nseg):
                    segx = min(seg * tf, tl - tf) << 5
                    t0 = 0
                    while t0 < tf:
                        rem = tf - t0
                        tp = 8 if rem >= 6 else (4 if rem >= 3 else (2 if rem >= 2 else 1))
                        for t in range(tp):
                            for c in range(nc):
                                j = t + c
                                off = (tw0 - c0) + (segx >> 5) + t0 + j   # word within the band row
                                assert off < stride
                                col = (tw0 << 5) + segx + 32 * (t0 + t) + 32 * c
                                if off < 4 * q4:   # a staged word: the grid's word of that column
                                    assert c0 + off == col >> 5 and col % 32 == 0
                                else:              # past the staged words: only tiles past the window
                                    assert segx + 32 * (t0 + t) >= 32 * tl
                        t0 += tp


def test_stamp_word_extraction():
    """stamp_part: grid word q takes the mask bits (Q >> (32q - px + 32)) & 0xFFFFFFFF;
    OR-ing them into the covered words stamps exactly the mask at column px."""
    rng = random.Random(11)
    for _ in range(500):
        w = rng.randint(1, 200)
        nc = (w + 62) // 32
        nq = 2 * ((w + 31 + 63) // 64) + 2   # the kernel's row of mask words
        assert nq >= nc + 1
        row = [1 if rng.random() < 0.5 else 0 for _ in range(w)]
        W = rng.randint(w, 600)
        px = rng.randint(0, W - w)
        q = qbits_row(row, nc, nq)
        Q = sum(v << (32 * i) for i, v in enumerate(q))
        q0, q1 = px >> 5, (px + w - 1) >> 5
        words = {}
        for qq in range(q0, q1 + 1):
            off = (qq << 5) - px + 32
            wi = off >> 5
            assert 1 <= off and wi < nq
            lo = q[wi]
            hi = q[wi + 1] if wi + 1 < nq else 0
            words[qq] = (((hi << 32) | lo) >> (off & 31)) & 0xFFFFFFFF
            assert words[qq] == (Q >> off) & 0xFFFFFFFF
        cells = [0] * (32 * (q1 + 1))
        for qq, v in words.items():
            for i in range(32):
                cells[32 * qq + i] |= (v >> i) & 1
        assert cells[px:px + w] == row and sum(cells) == sum(row)


def longest_free_in_word(o):
    """The kernel's binary lifting over runs of 2, 4, 8 and 16 free cells."""
    M = 0xFFFFFFFF
    r1 = ~o & M
    r2 = r1 & (r1 >> 1)
    r4 = r2 & (r2 >> 2)
    r8 = r4 & (r4 >> 4)
    r16 = r8 & (r8 >> 8)
    if o == 0:
        return 32
    cur, L = r16, 16
    if not cur:
        cur, L = M, 0
    for rs, s in ((r8, 8), (r4, 4), (r2, 2), (r1, 1)):
        c = cur & (rs >> L)
        if c:
            cur, L = c, L + s
    return L


def choose_lg(rows, nw):
    """rows_free_run's lanes per row, 2^lg: the cheapest of rounds x (4 per word
    + 10 per butterfly step) for lg with 2^lg < 2 nw (at most 6)."""
    lg, cost = 0, 1 << 30
    for l in range(7):
        if (1 << l) >= 2 * nw:
            break
        rounds = (rows + (256 >> l) - 1) // (256 >> l)
        c = rounds * (4 * ((nw + (1 << l) - 1) >> l) + 10 * l)
        if c < cost:
            cost, lg = c, l
    return lg


def bit_row_scan(row, lg):
    W = len(row)
    nw = (W + 31) // 32
    L = 1 << lg
    wpl = (nw + L - 1) >> lg
    lanes = []
    for sub in range(L):
        pre = run = best = 0
        first = True
        for i in range(wpl):
            q = sub * wpl + i
            o = 0xFFFFFFFF
            if q < nw:
                o = sum(row[32 * q + t] << t for t in range(32) if 32 * q + t < W)
                valid = W - 32 * q
                if valid < 32:
                    o |= (0xFFFFFFFF << valid) & 0xFFFFFFFF
            z = o == 0
            lead = 32 if z else (o & -o).bit_length() - 1
            trail = 32 if z else 32 - o.bit_length()
            inner = longest_free_in_word(o)
            if first and not z:
                pre = run + lead
            if not z:
                best = max(best, run + lead, inner)
            run = run + 32 if z else trail
            first = first and z
        if first:
            pre = run
        lanes.append((pre, run, best))
    length = 32 * wpl
    d = 1
    while d < L:
        joined = []
        for i in range(L):
            a, b = (lanes[i], lanes[i ^ d]) if (i & d) == 0 else (lanes[i ^ d], lanes[i])
            pre = length + b[0] if a[0] == length else a[0]
            suf = length + a[1] if b[0] == length else b[1]
            joined.append((pre, suf, max(a[2], b[2], a[1] + b[0])))
        lanes = joined
        length *= 2
        d *= 2
    assert len(set(lanes)) == 1
    return max(lanes[0])


def plain_scan(row):
    best = cur = 0
    for v in row:
        cur = 0 if v else cur + 1
        best = max(best, cur)
    return best


def test_longest_free_in_word():
    rng = random.Random(5)
    for _ in range(20000):
        o = rng.getrandbits(32) & rng.getrandbits(32) if rng.random() < 0.5 else rng.getrandbits(32)
        if rng.random() < 0.1:
            o = 0 if rng.random() < 0.5 else 0xFFFFFFFF
        bits = [(o >> t) & 1 for t in range(32)]
        assert longest_free_in_word(o) == plain_scan(bits), hex(o)


def test_bit_row_scan_matches_plain_scan():
    rng = random.Random(2)
    for _ in range(1500):
        W = rng.choice([1, 2, 31, 32, 33, 63, 64, 65, 250, 300, 350, 400, 500, 511, 512, 513, 700, 1100,
                        2048, 2049, 4096])
        p = rng.random() ** 2
        row = [1 if rng.random() < p else 0 for _ in range(W)]
        if rng.random() < 0.3:
            a = rng.randrange(W)
            b = rng.randrange(a, W + 1)
            row[a:b] = [0] * (b - a)
        nw = (W + 31) // 32
        lgs = [l for l in range(7) if (1 << l) < 2 * nw]
        assert choose_lg(rng.randint(1, 600), nw) in lgs
        for lg in lgs:   # every lane count the row count can select
            assert bit_row_scan(row, lg) == plain_scan(row), (W, lg)
    for W in (1, 64, 300, 512, 4096):   # all free, all occupied
        for lg in (0, choose_lg(120, (W + 31) // 32)):
            assert bit_row_scan([0] * W, lg) == W and bit_row_scan([1] * W, lg) == 0
    # a 120-row C5 stamp on 500-cell rows takes one round of two lanes per row
    assert choose_lg(120, 16) == 1 and choose_lg(1, 16) == 2


# ---- fp4 (e2m1) operands: 32x32x64 MFMA, lane (row/col = l % 32, g = l // 32)
# supplies k = 32g + 8v + e (VGPR v, nibble e, low nibble first) ----

E2M1 = {0: 0.0, 1: 0.5, 2: 1.0, 4: 2.0}


def nib_frag(w):
    """nib_frag: the four VGPRs of e2m1 codes from one 32-bit word."""
    return [w & 0x11111111, w & 0x22222222, w & 0x44444444, (w >> 1) & 0x44444444]


def nib_values(w):
    """The lane's 32 operand values in MFMA k order (8v + e)."""
    out = []
    for v in nib_frag(w):
        for e in range(8):
            out.append(E2M1[(v >> (4 * e)) & 0xF])
    return out


def nib_bits(w):
    """Which bit of the word each k position holds (None when zero for every word)."""
    pos = []
    for k in range(32):
        src = [b for b in range(32) if nib_values(1 << b)[k] != 0]
        pos.append(src[0] if len(src) == 1 else None)
    return pos


def test_fp4_fragment_is_a_bijection_with_positive_values():
    pos = nib_bits(0)
    assert sorted(pos) == list(range(32))
    for b in range(32):   # one set cell -> one positive value, nothing else
        vals = nib_values(1 << b)
        assert sum(1 for v in vals if v) == 1 and max(vals) in (0.5, 1.0, 2.0)


def test_fp4_tile_counts_match_brute_force_overlap():
    """One fp4 tile per 64-column chunk c: lane (m, g) takes band word t + 2c + g,
    lane (n, g) the mask window (Q words 2c + g, 2c + g + 1) >> (32 - n); zero
    counts are the zero overlaps of the placements x0 + n."""
    rng = random.Random(9)
    for _ in range(10):
        w = rng.randint(1, 150)
        nc4 = (w + 31 + 63) // 64
        nq = 2 * nc4 + 2
        row = [1 if rng.random() < 0.6 else 0 for _ in range(w)]
        row[rng.randrange(w)] = 1
        q = qbits_row(row, None, nq)
        width = 64 * (nc4 + 1)
        grid = [[1 if rng.random() < rng.choice([0.02, 0.2]) else 0 for _ in range(width)] for _ in range(32)]
        words = [[sum(grid[m][32 * j + i] << i for i in range(32)) for j in range(width // 32)] for m in range(32)]
        total = np.zeros((32, 32))
        for c in range(nc4):
            A = np.zeros((32, 64))
            B = np.zeros((64, 32))
            for m in range(32):
                for g in range(2):
                    A[m, 32 * g:32 * g + 32] = nib_values(words[m][2 * c + g])
            for n in range(32):
                for g in range(2):
                    win = (q[2 * c + g + 1] << 32) | q[2 * c + g]
                    B[32 * g:32 * g + 32, n] = nib_values((win >> (32 - n)) & 0xFFFFFFFF)
            total += A @ B
        for m in range(32):
            for n in range(32):
                ov = sum(row[b] * grid[m][n + b] for b in range(w))
                assert (total[m, n] == 0) == (ov == 0), (w, m, n)


def test_fp4_staged_window_addressing():
    """The fp4 path's word addresses loff + t0 + j + g (j = t + 2c) and mask words
    g + 2c, g + 2c + 1 stay inside the band row's stride and the row's 2 nc4 + 2
    mask words, and read the grid word of column x0 + 64c + 32g when staged."""
    rng = random.Random(13)
    for _ in range(300):
        W = rng.randint(40, 700)
        w = rng.randint(1, min(W, 200))
        nc4 = (w + 31 + 63) // 64
        cw = 2 * nc4
        T = (W - w + 1 + 31) // 32
        pw = max(4, T + cw - 1, (W + 31) // 32)
        pw = (pw + 3) & ~3
        TW = rng.randint(1, T)
        stride = band_stride(TW, cw)
        for tw0 in range(0, T, TW):
            tl = min(TW, T - tw0)
            c0 = tw0 & ~3
            q4 = (((tw0 + tl + cw - 1 + 3) & ~3) - c0) >> 2
            assert c0 + 4 * q4 <= pw and 4 * q4 <= stride
            for rl in range(6):
                nseg = 32 >> rl
                if nseg > tl:
                    continue
                tf = (tl + nseg - 1) >> (5 - rl)
                for seg in range(nseg):
                    segx = min(seg * tf, tl - tf) << 5
                    t0 = 0
                    while t0 < tf:
                        rem = tf - t0
                        tp = 8 if rem >= 6 else (4 if rem >= 3 else (2 if rem >= 2 else 1))
                        for t in range(tp):
                            for c in range(nc4):
                                for g in range(2):
                                    assert g + 2 * c + 1 < 2 * nc4 + 2
                                    off = (tw0 - c0) + (segx >> 5) + t0 + t + 2 * c + g
                                    assert off < stride
                                    col = (tw0 << 5) + segx + 32 * (t0 + t) + 64 * c + 32 * g
                                    if off < 4 * q4:
                                        assert c0 + off == col >> 5
                                    else:
                                        assert segx + 32 * (t0 + t) >= 32 * tl
                        t0 += tp
