"""Folded correlation groups (Fold, window_passes and decode_zero in
csrc/brkga_decoder.hip): a group of nrows <= R = 2^rlog candidates fills the
32 accumulator rows with F = 32 / R segments of a window's tiles; passes of
8/4/2/1 tiles (a remainder of 6-7 tiles rounded up to 8, of 3 to 4) run over
every segment at once and the pass loop may stop early
once candidate 0 has a zero left of every tile still to come; each pass takes
its minimum key by a scalar selection over the lanes' keys (pass_min_key).
With screening (screen_window), the passes run over the runs of tiles whose
counts over a subset of the mask rows have a zero (screen_need).
The vacancy filter's compacted word list (vacancy_rotation) is restated
at the end against the plain definition of a candidate top row.
This restates that schedule in Python on random zero patterns and checks it against the
plain answer -- the smallest (candidate, x) with zero overlap -- so that a
wrong segment layout, a gap between segments or a premature exit fails."""
import random

XKEY = 8192
NONE = 0x7FFFFFFF


def fold_params(nrows, tl):
    rl = 5
    while rl > 0 and (1 << (rl - 1)) >= nrows:
        rl -= 1
    while (32 >> rl) > tl:
        rl += 1
    tf = (tl + (32 >> rl) - 1) >> (5 - rl)
    return rl, tf


def pass_min_key(lane_key, rl):
    """The decode's scalar selection of a pass's minimum key: the smallest
    candidate with a zero (lane bits folded over segments), then the lowest
    segment of that candidate with a zero, whose key is read from its lane."""
    b = sum(1 << m for m in range(32) if lane_key[m] != NONE)
    if not b:
        return NONE
    bc, sh = b, 16
    while sh >= (1 << rl):
        bc |= bc >> sh
        sh >>= 1
    rmask = 0xFFFFFFFF if rl >= 5 else (1 << (1 << rl)) - 1
    low = bc & rmask
    cmin = (low & -low).bit_length() - 1
    pat = 1 if rl >= 5 else 0xFFFFFFFF // rmask
    sel = b & ((pat << cmin) & 0xFFFFFFFF)
    return lane_key[(sel & -sel).bit_length() - 1]


def pass_tiles(rem, tpm, tpcap):
    """Tiles of the next pass: 8 for a remainder of 6 or more, 4 for 3 or
    more, 2 for 2, within the register limit tpm and the tuned cap tpcap."""
    for t, need in ((8, 6), (4, 3), (2, 2)):
        if tpm >= t and tpcap >= t and rem >= need:
            return t
    return 1


def screen_need(szero, nrows, tl):
    """screen_window: bit t of the mask when some accumulator row (candidate
    or fold segment; rows past nrows repeat candidate nrows - 1) has a zero
    subset count anywhere in tile t -- columns past the window or past xmax
    included, since the screen reports every zero it computes."""
    rl, tf = fold_params(nrows, tl)
    segx = [min((m >> rl) * tf, tl - tf) * 32 for m in range(32)]
    need = 0
    for t in range(tf):
        for m in range(32):
            c = min(m & ((1 << rl) - 1), nrows - 1)
            if any(szero(c, segx[m] + 32 * t + n) for n in range(32)):
                need |= 1 << t
                break
    return need


def window_keys(zero, nrows, tl, xbase, xmax, tpm, tpcap, early, need=None):
    """zero[c][x]: candidate c has zero overlap at window column x < 32 * tl;
    columns past the window (read by a rounded-up pass) hold garbage, which
    is modelled as zeros that the window-end limit must discard.  need: the
    screen's tile mask (None: every tile); window_passes runs the passes over
    each run of needed tiles in order."""
    rl, tf = fold_params(nrows, tl)
    if need is None:
        need = (1 << tf) - 1
    segx = [min((m >> rl) * tf, tl - tf) * 32 for m in range(32)]
    xend = min(xmax, xbase + 32 * tl - 1)
    key = NONE
    while need:
        t0 = (need & -need).bit_length() - 1
        rem = 0
        while (need >> (t0 + rem)) & 1:
            rem += 1
        tp = pass_tiles(rem, tpm, tpcap)
        lane_key = [NONE] * 32
        for m in range(32):   # lanes 0..31, one accumulator row each
            c = m & ((1 << rl) - 1)
            if c >= nrows:
                continue
            for t in range(tp):
                for n in range(32):
                    x = xbase + segx[m] + 32 * (t0 + t) + n
                    garbage = x - xbase >= 32 * tl
                    if x <= xend and (garbage or zero[c][x - xbase]):
                        lane_key[m] = min(lane_key[m], c * XKEY + x)
        k = pass_min_key(lane_key, rl)
        assert k == min(lane_key)
        key = min(key, k)
        need &= ~((1 << (t0 + tp)) - 1)
        if early and key < xbase + 32 * (t0 + tp):
            break
    return key


def test_folded_schedule_finds_the_smallest_key():
    rng = random.Random(7)
    for _ in range(3000):
        tl = rng.randint(1, 12)
        nrows = rng.randint(1, 32)
        xbase = 32 * rng.randint(0, 3)
        width = 32 * tl
        xmax = xbase + rng.randint(0, width + 64)   # may lie past the window
        p = rng.choice([0.0, 0.002, 0.01, 0.05, 0.3])
        zero = [[rng.random() < p for _ in range(width)] for _ in range(nrows)]
        want = min([c * XKEY + xbase + x for c in range(nrows) for x in range(width)
                    if zero[c][x] and xbase + x <= xmax], default=NONE)
        tpm = rng.choice([8, 4, 2, 1])
        tpcap = rng.choice([8, 4, 2, 1])
        for early in (False, True):
            assert window_keys(zero, nrows, tl, xbase, xmax, tpm, tpcap, early) == want, \
                (tl, nrows, xbase, xmax, tpm, tpcap, early)


def _screened_case(rng):
    tl = rng.randint(1, 12)
    nrows = rng.randint(1, 32)
    xbase = 32 * rng.randint(0, 3)
    width = 32 * tl
    xmax = xbase + rng.randint(0, width + 64)
    p = rng.choice([0.0, 0.002, 0.01, 0.05])
    zero = [[rng.random() < p for _ in range(width)] for _ in range(nrows)]
    # a subset of the mask rows leaves every zero of the full count and more
    extra = rng.choice([0.0, 0.001, 0.02, 0.2])
    sz = [[(x < width and zero[c][x]) or rng.random() < extra for x in range(width + 96)]
          for c in range(nrows)]
    want = min([c * XKEY + xbase + x for c in range(nrows) for x in range(width)
                if zero[c][x] and xbase + x <= xmax], default=NONE)
    return tl, nrows, xbase, xmax, zero, sz, want


def test_screened_schedule_finds_the_smallest_key():
    """Passes over the screen's runs of needed tiles (with the early exit) give
    the plain answer for any screen pattern that keeps the full count's zeros."""
    rng = random.Random(11)
    for _ in range(1500):
        tl, nrows, xbase, xmax, zero, sz, want = _screened_case(rng)
        need = screen_need(lambda c, x: sz[c][x], nrows, tl)
        tpm = rng.choice([8, 4, 2, 1])
        tpcap = rng.choice([8, 4, 2, 1])
        for early in (False, True):
            assert window_keys(zero, nrows, tl, xbase, xmax, tpm, tpcap, early, need) == want, \
                (tl, nrows, xbase, xmax, tpm, tpcap, early, need)


def test_screen_that_loses_a_zero_is_caught():
    """A screen that dropped the true answer's tile (a wrong row subset or a
    lost bit) changes the key on some case: the check above can fail."""
    rng = random.Random(5)
    caught = 0
    for _ in range(300):
        tl, nrows, xbase, xmax, zero, sz, want = _screened_case(rng)
        if want == NONE:
            continue
        rl, tf = fold_params(nrows, tl)
        need = screen_need(lambda c, x: sz[c][x], nrows, tl)
        # drop every tile that holds the answer's column in some segment
        x = want % XKEY - xbase
        segs = {min(q * tf, tl - tf) for q in range(32 >> rl)}
        for s0 in segs:
            if 32 * s0 <= x < 32 * (s0 + tf):
                need &= ~(1 << ((x - 32 * s0) // 32))
        got = window_keys(zero, nrows, tl, xbase, xmax, 8, 8, True, need)
        caught += got != want
    assert caught > 50


def test_segments_cover_the_window_without_leaving_it():
    for tl in range(1, 65):
        for nrows in range(1, 33):
            rl, tf = fold_params(nrows, tl)
            assert (1 << rl) >= nrows and (32 >> rl) <= tl
            starts = {min(q * tf, tl - tf) for q in range(32 >> rl)}
            covered = set()
            for s in starts:
                assert 0 <= s and s + tf <= tl
                covered.update(range(s, s + tf))
            assert covered == set(range(tl))


def vacancy_words(vac, den_sorted, ny, y0, nq, H):
    """vacancy_rotation: word K = 4k + s of 64 top rows holds y0 + 64K + l in
    lane l; lane l's read vq[base + a] (base = y0 + 256k + l, or y0 for a lane
    whose quad has no valid top row) packs the free runs of rows base + a +
    64s (rows past H pack as 0); returns the passing top rows in the kernel's
    order (words descending, lanes by descending row within a word)."""
    def vq(y):
        assert 0 <= y < H, y                       # every read stays below H
        return [vac[y + 64 * s] if y + 64 * s < H else 0 for s in range(4)]

    ok = {}
    for k in range(nq):
        for lane in range(64):
            t = 256 * k + lane
            base = y0 + (t if t < ny else 0)
            fields = [t + 64 * s < ny for s in range(4)]
            for d, row in den_sorted:
                f = vq(base + row)
                for s in range(4):
                    if f[s] < d:
                        fields[s] = False
            for s in range(4):
                ok[(4 * k + s, lane)] = fields[s]
    out = []
    for K in range(4 * nq - 1, -1, -1):
        out += [y0 + 64 * K + lane for lane in range(63, -1, -1) if ok[(K, lane)]]
    return out


def test_vacancy_words_list_the_candidates_descending():
    """The packed vacancy filter lists exactly the top rows y whose rows
    y..y+h-1 all have a free run at least the mask row's density, in
    descending order, over passes of 512 top rows (one quad of 256 when that
    covers the rest)."""
    rng = random.Random(13)
    for _ in range(150):
        h = rng.randint(1, 40)
        H = rng.randint(h, 1200)
        ny = H - h + 1
        yf = rng.randint(0, ny) if rng.random() < 0.3 else 0
        W = 100
        vac = [rng.choice([W, W, rng.randint(0, W)]) for _ in range(H)]
        for _ in range(rng.randint(0, 6)):
            a = rng.randrange(H)
            for r in range(a, min(H, a + rng.randint(1, 80))):
                vac[r] = rng.randint(0, 20)
        den = [rng.randint(1, 60) for _ in range(h)]
        den_sorted = sorted(((den[a], a) for a in range(h)), key=lambda t: -t[0])
        want = [y for y in range(ny - 1, yf - 1, -1) if all(vac[y + a] >= den[a] for a in range(h))]
        got = []
        if ny > yf:
            for cpass in range((ny - yf - 1) // 512, -1, -1):
                y0 = yf + 512 * cpass
                got += vacancy_words(vac, den_sorted, ny - y0, y0, 1 if ny - y0 <= 256 else 2, H)
        assert got == want
