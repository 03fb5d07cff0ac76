"""Float64 NumPy restatement of the surrogate MLP (features, forward, gradients).

TEST INFRASTRUCTURE ONLY (the checker of the HIP surrogate kernels).  It
follows the network definition of brkga_surrogate.hip: W1 and W2 are used as
bfloat16, activations H1, H2 and the back-propagated dH2, dH1 are rounded to
bfloat16 where the kernels store them, everything else is float64 here (the
kernels accumulate in float32, so results agree to float32 rounding plus an
occasional one-ulp bfloat16 flip).
"""
import numpy as np


EMULATE_BF16 = True


def bf16(x):
    """Round float values to bfloat16 (nearest even), returned as float64.

    With EMULATE_BF16 False values pass through unrounded (exact float64
    network, used to check the backward formulas by finite differences)."""
    if not EMULATE_BF16:
        return np.asarray(x, dtype=np.float64)
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def machine_bucket(mv, thresholds):
    """Machine of each machine key (placement.py:121-127): bucket 0 for
    mv <= t[0], bucket m for t[m-1] < mv <= t[m], the last one above."""
    thr = np.asarray(thresholds, np.float64).reshape(-1)
    mv = np.asarray(mv, np.float64)
    if thr.size == 0:
        return np.zeros(mv.shape, np.int64)
    return np.searchsorted(thr, mv, side="left")


def lane_tree_sum(v):
    """float32 sum in the kernel's order: lane l (of 64) adds v[l], v[l + 64],
    ... from 0, then a butterfly adds lane l and lane l ^ off for off = 32, 16,
    ..., 1 (lane 0's value is the result)."""
    v = np.asarray(v, np.float32)
    k = -(-v.size // 64)
    pad = np.zeros(64 * max(k, 1), np.float32)
    pad[:v.size] = v
    part = np.zeros(64, np.float32)
    for row in pad.reshape(-1, 64):
        part = part + row
    idx = np.arange(64)
    for off in (32, 16, 8, 4, 2, 1):
        part = part + part[idx ^ off]
    return part[0]


def features(keys, slot_area, slot_height, fp, thresholds=(), machine_terms=None):
    """[B, fp] features: sequence position / n and machine bucket / max(M-1, 1)
    per slot, then area/height in ascending sequence-key order (ties by slot).
    With ``machine_terms`` ([M][3][n] float32) those 4n columns are scaled by
    1/8 and four aggregates per machine m follow: float32 sums of terms 0 and 1
    over the slots in bucket m (``lane_tree_sum``), their count times M / n,
    and the largest term 2 among them (0 for none)."""
    keys = np.asarray(keys, dtype=np.float32)
    B, g2 = keys.shape
    n = g2 // 2
    M = np.asarray(thresholds).size + 1
    mt = None if machine_terms is None else np.asarray(machine_terms, np.float32)
    ss = np.float32(0.125 if mt is not None else 1.0)
    x = np.zeros((B, fp), np.float64)
    for b in range(B):
        order = np.lexsort((np.arange(n), keys[b, :n]))   # stable by slot on ties
        rank = np.empty(n, np.int64)
        rank[order] = np.arange(n)
        x[b, :n] = bf16((rank.astype(np.float32) / np.float32(n)) * ss)
        mach = machine_bucket(keys[b, n:].astype(np.float64), thresholds)
        x[b, n:2 * n] = bf16((mach.astype(np.float32) / np.float32(max(M - 1, 1))) * ss)
        x[b, 2 * n:3 * n] = bf16(np.asarray(slot_area, np.float32)[order] * ss)
        x[b, 3 * n:4 * n] = bf16(np.asarray(slot_height, np.float32)[order] * ss)
        if mt is not None:
            for m in range(M):
                on = mach == m
                s0 = lane_tree_sum(np.where(on, mt[m, 0], np.float32(0)))
                s1 = lane_tree_sum(np.where(on, mt[m, 1], np.float32(0)))
                mx = mt[m, 2][on].max() if on.any() else np.float32(0)
                cnt = np.float32(int(on.sum()) * M) / np.float32(n)
                x[b, 4 * n + 4 * m:4 * n + 4 * m + 4] = bf16(np.array([s0, s1, cnt, mx], np.float32))
    return x


def unpack(params, fp, hidden):
    p = np.asarray(params, dtype=np.float64)
    o = 0
    W1 = p[o:o + fp * hidden].reshape(fp, hidden); o += fp * hidden
    b1 = p[o:o + hidden]; o += hidden
    W2 = p[o:o + hidden * hidden].reshape(hidden, hidden); o += hidden * hidden
    b2 = p[o:o + hidden]; o += hidden
    w3 = p[o:o + hidden]; o += hidden
    b3 = p[o:o + 1]; o += 1
    assert o == p.size
    return W1, b1, W2, b2, w3, b3


def forward(x, params, fp, hidden):
    W1, b1, W2, b2, w3, b3 = unpack(params, fp, hidden)
    H1 = bf16(np.maximum(x @ bf16(W1) + b1, 0.0))
    H2 = bf16(np.maximum(H1 @ bf16(W2) + b2, 0.0))
    y = H2 @ w3 + b3[0]
    return y, H1, H2


def gradients(x, target, params, fp, hidden):
    """Mean-squared-error gradient, flattened in the parameter layout."""
    W1, b1, W2, b2, w3, b3 = unpack(params, fp, hidden)
    y, H1, H2 = forward(x, params, fp, hidden)
    B = x.shape[0]
    dy = 2.0 * (y - np.asarray(target, np.float64)) / B
    dw3 = H2.T @ dy
    db3 = np.array([dy.sum()])
    dH2 = bf16(np.where(H2 > 0, dy[:, None] * w3[None, :], 0.0))
    db2 = dH2.sum(axis=0)
    dW2 = H1.T @ dH2
    dH1 = bf16(np.where(H1 > 0, dH2 @ bf16(W2).T, 0.0))
    db1 = dH1.sum(axis=0)
    dW1 = x.T @ dH1
    return np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dw3, db3])


def gradient_magnitudes(x, target, params, fp, hidden):
    """Per gradient element (same layout as ``gradients``): S, the sum of the
    absolute values of the products its sum adds (float64), and T, an upper
    bound of any single product.  A float32 accumulation of bfloat16 operands
    differs from the float64 value by at most (K u32 + 2 u16) S (K terms, u32 =
    2^-24 and u16 = 2^-8 the unit roundoffs, two stored bf16 operands per
    product off by at most one unit each), plus T for a term whose ReLU mask
    flips between the two computations."""
    W1, b1, W2, b2, w3, b3 = unpack(params, fp, hidden)
    y, H1, H2 = forward(x, params, fp, hidden)
    B = x.shape[0]
    dy = 2.0 * (y - np.asarray(target, np.float64)) / B
    dH2 = bf16(np.where(H2 > 0, dy[:, None] * w3[None, :], 0.0))
    dH1u = bf16(dH2 @ bf16(W2).T)          # dH1 without its mask: a flipped mask adds this term
    dH1 = np.where(H1 > 0, dH1u, 0.0)
    ax, aH1, aH2, ady, adH2, adH1, adH1u = (np.abs(v) for v in (x, H1, H2, dy, dH2, dH1, dH1u))
    S = np.concatenate([(ax.T @ adH1).ravel(), adH1.sum(axis=0), (aH1.T @ adH2).ravel(), adH2.sum(axis=0),
                        aH2.T @ ady, [ady.sum()]])
    dH2u = np.abs(dy)[:, None] * np.abs(w3)[None, :]
    T = np.concatenate([np.outer(ax.max(axis=0), adH1u.max(axis=0)).ravel(), adH1u.max(axis=0),
                        np.outer(aH1.max(axis=0), dH2u.max(axis=0)).ravel(), dH2u.max(axis=0),
                        aH2.max(axis=0) * ady.max(), [ady.max()]])
    return S, T


def gradient_tolerance(x, target, params, fp, hidden):
    """Element-wise bound on |GPU gradient - float64 gradient| (see
    ``gradient_magnitudes``)."""
    S, T = gradient_magnitudes(x, target, params, fp, hidden)
    K = max(x.shape[0], fp, hidden)
    return (K * 2.0 ** -23 + 2.0 ** -6) * S + T + 1e-30
