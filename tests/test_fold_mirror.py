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


def vacancy_words(vac, den_sorted, ny, nk, head=8, compact=5):
    """vacancy_rotation: candidate words of 64 top rows; the densest `head`
    rows over every word, then (from `compact` words on) the words with
    candidates left listed and padded to 1, 2 or 4 entries with a dead word
    whose bits stay zero, and the other rows over the list only."""
    cw = []
    for k in range(nk):
        left = ny - 64 * k
        cw.append((1 << min(64, left)) - 1)

    def rows(a0, a1, idx, c):
        for a in range(a0, a1):
            d, row = den_sorted[a]
            for j, k in enumerate(idx):
                bits = 0
                for lane in range(64):
                    if vac[64 * k + lane + row] >= d:
                        bits |= 1 << lane
                c[j] &= bits
            if not any(c):
                return True
        return False

    h = len(den_sorted)
    idx = list(range(nk))
    if nk < compact:
        rows(0, h, idx, cw)
        return cw
    if rows(0, min(h, head), idx, cw) or h <= head:
        return cw
    live = [k for k in range(nk) if cw[k]]
    dead = next(k for k in range(nk) if not cw[k]) if len(live) < nk else None
    kk = 1 if len(live) <= 1 else (2 if len(live) <= 2 else (4 if len(live) <= 4 else nk))
    if kk < nk:
        lst = live + [dead] * (kk - len(live))
        c = [cw[k] for k in lst]
        rows(head, h, lst, c)
        out = [0] * nk
        for j, k in enumerate(lst[:len(live)]):
            out[k] = c[j]
        return out
    rows(head, h, idx, cw)
    return cw


def test_vacancy_word_compaction_keeps_the_candidates():
    """The compacted vacancy filter lists exactly the top rows y whose rows
    y..y+h-1 all have a free run at least the mask row's density."""
    rng = random.Random(13)
    for _ in range(300):
        h = rng.randint(1, 40)
        ny = rng.randint(1, 64 * 8)
        nk = (ny + 63) // 64
        H = ny + h - 1
        W = 100
        vac = [rng.choice([W, W, rng.randint(0, W)]) for _ in range(H)]
        # blocks of full rows so that some words die early and others live
        for _ in range(rng.randint(0, 6)):
            a = rng.randrange(H)
            for r in range(a, min(H, a + rng.randint(1, 80))):
                vac[r] = rng.randint(0, 20)
        vac += [0] * (64 * nk + h)     # rows past H read as occupied (masked by cw)
        den = [rng.randint(1, 60) for _ in range(h)]
        den_sorted = sorted(((den[a], a) for a in range(h)), key=lambda t: -t[0])
        want = [y for y in range(ny) if all(vac[y + a] >= den[a] for a in range(h))]
        cw = vacancy_words(vac, den_sorted, ny, nk)
        got = [64 * k + i for k in range(nk) for i in range(64) if (cw[k] >> i) & 1]
        assert got == want
