// brkga_decoder.hip -- MI355X (gfx950) BRKGA population decoder.
//
// One persistent launch decodes a whole population.  Work item = context =
// (chromosome, printer machine).  Each 256-thread workgroup pulls contexts
// from a device-side queue and runs the sequential placement procedure of
// the reference (placement.py:29-91, binClassNew.py:45-174) for it:
//
//   * sequence decode        placement.py:109-133, full_native_decoder.py:961-1002
//   * per-part feasibility   placement.py:46-53
//   * first fit over bins    placement.py:55-64 (f64 area check :57-59)
//   * per rotation: vacancy filter (numba_utils.py:131-157), then the exact
//     collision test: an int8 MFMA correlation of the part mask against the
//     bin's occupancy grid replaces rfft2/irfft2 (collision_backend.py:109-128)
//   * best rotation by density -> row -> col   binClassNew.py:101-117
//   * grid stamp + free-run ("vacancy") update binClassNew.py:147-174,
//     numba_utils.py:161-188, cuda_batch_update.py:90-194
//   * new bin at (0, H-1) with best rotation   placement.py:66-83
//   * makespan = sum over bins (pt + pth + setup), max over machines
//     placement.py:85-89,147-156
//
// Wave r of the workgroup runs the vacancy filter of rotation r (a bitset of
// the top rows whose every row has a long enough free run); the collision
// tests then run one rotation at a time on the whole workgroup, the four
// waves splitting the mask rows and AND-combining their zero masks in LDS.
//
// Correlation as a GEMM.  For a group of 32 candidate top rows y_m and an
// x-tile of 32 placements starting at x0 = 32t:
//     count[y_m][x0+n] = sum_a sum_c sum_k A_{a,c}[m][k] * B_{a,c}[k][n]
//     A_{a,c}[m][k] = G[y_m + a][x0 + 32c + k]        (grid bytes, k < 32)
//     B_{a,c}[k][n] = P[a][32c + k - n]               (banded Toeplitz of mask row a)
// issued as V_MFMA_I32_32X32X32_I8 with int32 accumulation, so counts are
// exact integers (B carries a positive weight of 1, 2, 4 or 8 per set mask
// cell, which leaves "count == 0", the only test made, unchanged).  A
// fragments are read from an LDS band of grid rows; B fragments are expanded
// in registers from the mask row's bits (one 32-bit word per 32 columns, kept
// per rotation at create time).
//
// Row pruning: a placement with top row y can only be collision-free if every
// grid row y+a has a free run >= the mask row's longest 1-run (the vacancy
// filter's condition, necessary, so results-neutral).  Candidate rows are
// listed bottom-up and evaluated 32 at a time; the first group holding a zero
// count contains the bottom-most valid row.  Dense mode evaluates every row
// of the valid region and never exits early (same results).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/brkga_mi355x.h"
#include "brkga_internal.h"

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));
// Pointers into LDS carry the LDS address space explicitly, so loads through
// them are ds_read instructions (with in-order completion the compiler can
// wait for precisely the ones it needs) rather than flat loads.
#define LDSP __attribute__((address_space(3)))
typedef LDSP const unsigned lds_cu32;
typedef LDSP unsigned lds_u32;
typedef LDSP const v4i lds_cv4i;
typedef LDSP v4i lds_v4i;
typedef LDSP unsigned long long lds_u64;
typedef LDSP const unsigned long long lds_cu64;
typedef LDSP const short lds_cs16;
typedef LDSP const int lds_ci32;

#define WG_THREADS 256
#define MAX_ROT 4
#define GROUP_ROWS 32
#define LDS_RESERVE 2048    // static LDS of the kernel + slack
#ifndef VC3_MAX
// Bins whose free runs stay in LDS at three workgroups per CU, at most: the
// rest of the LDS goes to the band (wider windows, larger row groups)
#define VC3_MAX 2
#endif
#define XKEY 8192           // key = m * XKEY + x ; plates must have W < XKEY
// Zero masks per wave and pass: up to 8 tiles (4 at three workgroups per CU)
// x 16 accumulators; a pass of the four waves holds 4 x that.
__host__ __device__ constexpr int zw_wave(int wpc) { return (wpc >= 3 ? 4 : 8) * 16; }
#define KEY_NONE 0x7fffffff

enum {
    ST_CALLS = 0, ST_CHROM, ST_CTX, ST_PLACE, ST_BINS, ST_TESTS, ST_GROUPS, ST_MFMA, ST_ALGO,
    ST_OVERFLOW, ST_PH0, ST_CTXMAX = ST_PH0 + 8, ST_CTXSUM,
    // per machine (BRKGA_MAX_MACHINES each): contexts, bins opened, most bins of one context
    ST_MCTX, ST_MBINS = ST_MCTX + BRKGA_MAX_MACHINES, ST_MPEAK = ST_MBINS + BRKGA_MAX_MACHINES,
    ST_COUNT = ST_MPEAK + BRKGA_MAX_MACHINES
};

struct DevProblem {
    int n_parts, n_machines, n_types;
    const double *thresholds;
    const int *slot_type, *type_id;
    const double *area;
    const int *nrot, *best_rot, *rot_base;
    const int *rot_h, *rot_w;
    const int *rot_nc, *rot_nc4;   // Toeplitz chunks of 32 / 64 columns: ceil((w + 31) / 32), ceil((w + 31) / 64)
    const long long *mask_off, *dens_off, *qbits_off;
    const unsigned char *masks;
    const int *dens;         // per rotation: (longest 1-run << 16) | 8 mask row, densest rows first
    const unsigned *qbits;   // per rotation: h rows of 2 nc4 + 2 dwords, mask bits after 32 zero bits, zero padded
    const int *plate_h, *plate_w;
    const double *setup, *proc_time, *pth;
    const unsigned char *fits;   // [M][n_types] rotation 0 fits the plate in some orientation
    // Tile shapes of the collision passes per (machine, Toeplitz chunk count
    // nc = ceil((w+31)/32), clamped to TUNE_NC-1): [M][TUNE_NC][TILE_ENTRY] =
    // {cap on x-tiles per pass, cap on x-tiles per staged LDS window, mask rows
    // of the screening pass (0: none), unused}.  Any values give the same
    // results; only the speed changes.
    const unsigned char *tiles;
    const int *type_cost;   // [n_types] expected work of placing one part: its mask's bounding-box area
};

struct DevWork {
    unsigned *grids;         // [n_wg][max_bins][hmax][pw] occupancy bit rows
    int *vac;                // [n_wg][max_bins][hmax]
    unsigned int *queue;
    unsigned long long *counters;
    int pw, hmax, max_bins;  // pw: 32-bit words per grid row (a multiple of 4)
    int hrot;                // max rotation height (LDS density rows)
    int q_off;               // LDS offset of the staged mask bits
    int vc_bins;             // bins whose vacancy vector is cached in LDS
    int band_off, band_words;  // LDS band for staged grid bit rows (32-bit words)
    int zw_off;                // LDS zero masks: [2 passes][4 waves][8 tiles * 16] u64, then 4 u64 screen masks
    int bins_off;              // LDS bin bookkeeping: area, pt, pth [max_bins] f64, minrow [max_bins] i32
};

struct LaunchArgs {
    const void *keys;
    int key_f64;
    long long n_chrom;
    double *ctx_ms;          // [n_chrom * M]
    int *log;                // nullable [n_chrom][n_parts][6]
    const int *order;        // nullable [n_chrom * M]: contexts in decode order
    int wpc;                 // workgroups per CU of the launch's occupancy
};

// Constant-rate (100 MHz on MI300-class parts) wall clock, read by thread 0
// only, to attribute a workgroup's time to the decode phases.
__device__ __forceinline__ unsigned long long tick() { return wall_clock64(); }

// Values read from LDS or derived from them are the same in every lane;
// saying so keeps them (and the loops they bound) in scalar registers.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ double load_key(const void *keys, int f64, long long idx) {
    return f64 ? ((const double *)keys)[idx] : (double)((const float *)keys)[idx];
}

__device__ __forceinline__ int machine_of(double mv, const double *thr, int M) {
    if (M <= 1) return 0;
    if (mv <= thr[0]) return 0;
    for (int m = 1; m < M - 1; m++)
        if (mv > thr[m - 1] && mv <= thr[m]) return m;
    return M - 1;
}

// floor(k / d) for the small loop counters of the staging and stamp loops:
// with m = floor((2^32 - 1) / d) + 1 = ceil(2^32 / d), umulhi(k, m) is exact
// whenever k * d < 2^32 (here k < 2^17, d < 2^10).  d must be uniform.
struct DivBy {
    int d;
    unsigned m;
    __device__ __forceinline__ explicit DivBy(int d_) : d(d_), m(d_ > 1 ? 0xFFFFFFFFu / (unsigned)d_ + 1u : 0u) {}
    __device__ __forceinline__ int operator()(int k) const { return d > 1 ? (int)__umulhi((unsigned)k, m) : k; }
};

// v_writelane_b32: `old` with lane `lane` (a constant) set to the scalar
// `val`.  The compiler exposes it as an intrinsic only, bound here by name.
extern "C" __device__ int lane_write(int val, int lane, int old) __asm("llvm.amdgcn.writelane.i32");

// MFMA operand fragments from occupancy bits.  Grids, bands and mask rows are
// bit rows (bit i of 32-bit word q = cell 32q + i, 1 = occupied / mask cell).
// The 32x32x32 int8 MFMA sums A[m][k] * B[k][n] over k; lane (n or m = lane&31,
// g = lane>>5) supplies the 16 k-values of group g as four VGPRs of 4 bytes.
// Any bijection between (g, VGPR r, byte j) and the 32 columns k of a chunk
// gives the same sums as long as A and B use the same one; this build uses
// k = 8j + 4g + r, so VGPR r of a lane is its 32-bit word shifted right by 4g,
// masked to bit r of every byte: a set cell becomes the byte 2^r (1, 2, 4 or
// 8), two ANDs per byte lane fewer than a per-byte expansion.  Only
// "count == 0" is ever tested and every product is non-negative, so any
// positive value per set cell gives the same decisions; counts stay below 64 *
// 4096 * 4096 < 2^31.
__device__ __forceinline__ v4i frag_from_word(unsigned y) {   // y: the lane's word >> 4g
    v4i r;
    r[0] = (int)(y & 0x01010101u);
    r[1] = (int)(y & 0x02020202u);
    r[2] = (int)(y & 0x04040404u);
    r[3] = (int)(y & 0x08080808u);
    return r;
}

// B fragment for Toeplitz chunk c of one mask row: B[k][n] = P[a][32c + k - n].
// Q is the mask row with 32 zero bits in front (Q[i] = P[a][i - 32]), held as
// words qv; lane n needs Q bits 32c + 32 - n + k, i.e. the 64-bit word (Q dword
// c+1 : Q dword c) shifted right by 32 - n, then by 4g for its k-group: one
// 64-bit shift by bs = 32 - n + 4g in [1, 36] (bits 28-31 of the result, which
// the shift may leave empty, are never used).
__device__ __forceinline__ v4i bfrag_from_bits(unsigned lo, unsigned hi, int bs) {
    return frag_from_word((unsigned)((((unsigned long long)hi << 32) | lo) >> bs));
}

// fp4 (e2m1) fragment of the 32x32x64 MFMA: a lane supplies 32 k-values, the
// 32 cells of one band or mask word, in VGPRs 0-3 (8 nibbles each, low nibble
// first; VGPRs 4-7 are unused by this format).  As for the int8 fragments, A
// and B share one bijection between (VGPR v, nibble i) and the word's bits:
// VGPR v < 3 keeps bit 4i + v of the word in place (e2m1 codes 1, 2, 4: 0.5,
// 1.0, 2.0), VGPR 3 takes bit 4i + 3 one place down (code 4; bit 3 of a
// nibble is the sign).  Three ANDs and one shift-and-AND per fragment.  Every
// product is 0 or positive, so a float32 sum of them is zero exactly when no
// mask cell meets an occupied cell, whatever the mask's size: "count == 0" is
// the same test as with integer counts.
__device__ __forceinline__ v8i nib_frag(unsigned w) {
    v8i r;
    r[0] = (int)(w & 0x11111111u);
    r[1] = (int)(w & 0x22222222u);
    r[2] = (int)(w & 0x44444444u);
    r[3] = (int)((w >> 1) & 0x44444444u);
    r[4] = 0; r[5] = 0; r[6] = 0; r[7] = 0;
    return r;
}

// One fp4 MFMA, both operands e2m1, unscaled: constant zero scale operands
// make the compiler emit V_MFMA_F32_32X32X64_F8F6F4 itself (one instruction,
// no scale load ahead of it, no register holding the scales).  The counts are
// sums of products of positive e2m1 values, exact in float32, and only their
// zero test is read.
__device__ __forceinline__ v16f mfma_f4(const v8i &a, const v8i &b, const v16f &c) {
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 4, 4, 0, 0, 0, 0);
}

// Accumulator of a 32x32 tile: int32 counts (int8 operands) or float32 (fp4).
template <bool F4> struct TileAcc { typedef v16i T; };
template <> struct TileAcc<true> { typedef v16f T; };

// Columns of one Toeplitz chunk: 32 (int8, K = 32) or 64 (fp4, K = 64).
__host__ __device__ constexpr int chunk_cols(bool f4) { return f4 ? 64 : 32; }

// Store one wave's zero masks of `tcnt` tiles: lane i < 16 of the wave
// collects accumulator i's ballot, then one 8-byte store per lane.
template <int TP, typename AT>
__device__ __forceinline__ void store_zero_masks(const AT (&acc)[TP], int tcnt, lds_u64 *zw) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int t = 0; t < TP; t++) {
        if (t < tcnt) {
            // ballot i (a scalar register pair) goes into lane i of (lo, hi)
            int lo = 0, hi = 0;
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const unsigned long long z = __ballot(acc[t][i] == 0);
                lo = lane_write((int)(unsigned)z, i, lo);
                hi = lane_write((int)(unsigned)(z >> 32), i, hi);
            }
            if (lane < 16) zw[t * 16 + lane] = ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
        }
    }
}

// Correlation pass of one wave over mask rows [a0, a1) for TP x-tiles of 32
// placements and the group's 32 candidate rows.  The four waves of the
// workgroup take disjoint mask-row ranges of the same tiles: every partial
// overlap count is a sum of non-negative products, so the total is zero iff
// every wave's partial is zero, and the waves combine their results with an
// AND of zero masks instead of adding counts: each wave stores its masks
// (zw[wave][t*16 + i], bit = lane) with plain LDS stores and the decode ANDs
// the four.  A operands come from the band's words, B operands from the mask
// bits (qs words per mask row in qbits); with NC (Toeplitz chunks of
// chunk_cols(F4) columns) known at compile time each A fragment feeds up to NC
// MFMAs and each B fragment feeds all TP tiles.
//   int8: tile t, chunk c reads band word t + c; both k-groups of a lane read
//         the same word and take its bits 4g onward (frag_from_word).
//   fp4:  tile t, chunk c is 64 columns, words t + 2c (k-group 0) and t + 2c +
//         1 (k-group 1): a lane's word stream starts g words later, and its B
//         window starts g mask words later.
template <int TP, int NC, bool F4>
__device__ __forceinline__ void corr_pass(lds_cu32 *band, int stride, int loff, int a0,
                                          int a1, lds_cu32 *qbits, int qs, int tloc, int tcnt,
                                          lds_u64 *zw) {
    constexpr int NA = F4 ? TP + 2 * NC - 1 : TP + NC - 1;   // band words per mask row
    constexpr int NQ = F4 ? 2 * NC : NC + 1;                 // mask words per row a lane reads
    const int lane = threadIdx.x & 63;
    const int g = lane >> 5;
    const int ks = g << 2;                               // int8: the lane's k-group shift, 4g
    const int bs = F4 ? 32 - (lane & 31) : 32 - (lane & 31) + ks;   // B window shift
    typedef typename TileAcc<F4>::T AT;
    AT acc[TP];
#pragma unroll
    for (int t = 0; t < TP; t++) acc[t] = (AT){};
    // the lane's band row (its candidate's row + mask row a) starts at base + a * stride
    lds_cu32 *base = band + loff + tloc + (F4 ? g : 0);
    lds_cu32 *qb = qbits + (F4 ? g : 0);
    if (a0 < a1) {
        // band words in order j = 0..NA-1 per mask row; word j + D is loaded
        // while j is used (the last D of a row load the next row's first D)
        constexpr int D = NA < 2 ? NA : 2;
        unsigned N[D];
        unsigned qv[NQ];
        lds_cu32 *r0 = base + a0 * stride;
#pragma unroll
        for (int j = 0; j < D; j++) N[j] = r0[j];
#pragma unroll
        for (int c = 0; c < NQ; c++) qv[c] = qb[a0 * qs + c];
        for (int a = a0; a < a1; a++) {
            typedef typename std::conditional<F4, v8i, v4i>::type FT;
            FT B[NC];
#pragma unroll
            for (int c = 0; c < NC; c++) {
                if constexpr (F4) B[c] = nib_frag((unsigned)((((unsigned long long)qv[2 * c + 1] << 32) | qv[2 * c]) >> bs));
                else B[c] = bfrag_from_bits(qv[c], qv[c + 1], bs);
            }
            const int an = a + 1 < a1 ? a + 1 : a;
            lds_cu32 *rc = base + a * stride;
            lds_cu32 *rn = base + an * stride;
#pragma unroll
            for (int c = 0; c < NQ; c++) qv[c] = qb[an * qs + c];
            unsigned Aw[NA];
#pragma unroll
            for (int j = 0; j < D; j++) Aw[j] = N[j];
#pragma unroll
            for (int j = 0; j < NA; j++) {
                const int p = j + D;
                if (p < NA) Aw[p] = rc[p];
                else N[p - NA] = rn[p - NA];
                if constexpr (F4) {
                    const v8i A = nib_frag(Aw[j]);
#pragma unroll
                    for (int c = 0; c < NC; c++) {
                        const int t = j - 2 * c;
                        if (t >= 0 && t < TP) acc[t] = mfma_f4(A, B[c], acc[t]);
                    }
                } else {
                    const v4i A = frag_from_word(Aw[j] >> ks);
#pragma unroll
                    for (int c = 0; c < NC; c++) {
                        const int t = j - c;
                        if (t >= 0 && t < TP)
                            acc[t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A, B[c], acc[t], 0, 0, 0);
                    }
                }
            }
        }
    }
    store_zero_masks<TP>(acc, tcnt, zw);
}

// Bottom-most then left-most zero in the combined masks of one pass, as
// c*XKEY + x (candidate c < nrows, x <= xmax), or KEY_NONE.  C layout of the
// 32x32 MFMA: lane holds column n = lane&31 and rows m = 8*(i/4) + 4*(lane/32)
// + i%4 of accumulator i, so row m lives in ballot word i = 4*(m/8) + m%4,
// half (m/4)%2.  Lanes 0..31 each scan one row m: candidate m & (2^rlog - 1)
// at x offset segx (the lane's segment of a folded group, 0 unfolded); every
// wave computes the result.
template <int TP, int ZW>
__device__ __forceinline__ int decode_zero(lds_u64 *zp, int xbase, int xmax, int nrows, int rlog, int segx) {
    const int lane = threadIdx.x & 63;
    const int m = lane & 31;
    const int i = ((m >> 3) << 2) + (m & 3);
    const int half = ((m >> 2) & 1) << 5;
    unsigned bits[TP];
#pragma unroll
    for (int t = 0; t < TP; t++) {   // all 4 x TP reads in flight together
        const int e = t * 16 + i;
        const unsigned long long z = zp[e] & zp[ZW + e] & zp[2 * ZW + e] & zp[3 * ZW + e];
        bits[t] = (unsigned)(z >> half);
    }
    int key = KEY_NONE;
    const int c = m & ((1 << rlog) - 1);
    if (lane < 32 && c < nrows) {
        const int x0 = xbase + segx;
#pragma unroll
        for (int t = TP - 1; t >= 0; t--) {   // the smallest tile with a zero wins
            const int lim = xmax - (x0 + (t << 5));   // last valid column offset in this tile
            unsigned b = lim < 0 ? 0u : (lim < 31 ? bits[t] & ((2u << lim) - 1u) : bits[t]);
            if (b) key = c * XKEY + x0 + (t << 5) + __builtin_ctz(b);
        }
    }
    // Minimum key in scalar form.  Row m = s * R + c (segment s, candidate c,
    // R = 2^rlog): the smallest candidate with a zero wins, and for it the
    // lowest segment with a zero holds the smallest x (segments start in
    // non-decreasing order and each covers every position between its start
    // and the next segment's, so a zero of a later segment left of an earlier
    // segment's minimum would lie in the earlier segment's range).
    const unsigned b = (unsigned)__ballot(key != KEY_NONE);
    if (!b) return KEY_NONE;
    unsigned bc = b;
    for (int sh = 16; sh >= (1 << rlog); sh >>= 1) bc |= bc >> sh;   // OR over segments
    const unsigned rmask = rlog >= 5 ? 0xFFFFFFFFu : ((1u << (1 << rlog)) - 1u);
    const int cmin = __builtin_ctz(bc & rmask);
    const unsigned pat = rlog >= 5 ? 1u : 0xFFFFFFFFu / rmask;   // bit s * R of every segment s
    return __builtin_amdgcn_readlane(key, __builtin_ctz(b & (pat << cmin)));
}

// A group's candidates across the 32 accumulator rows.  With R = 2^rlog >=
// nrows, row m holds candidate m & (R-1) in segment m >> rlog of the
// window's tiles: F = 32 / R segments of ceil(tl / F) tiles each, the last
// ones moved left to end at tile tl (no A read leaves the staged window; the
// overlap only repeats positions, and the decode keeps the smallest key).
// One MFMA then covers F times the x-range of an unfolded group.  loff: the
// lane's A-row byte offset in the band (row of its candidate * stride +
// segment column); segx: the segment's first column; skip: leading mask
// rows that meet only empty grid rows for every candidate of the group (they
// add nothing to any overlap count, so the waves split rows [skip, h)).
struct Fold { int loff, rlog, segx, skip; };

// Tiles per pass shrink as the chunk count grows so that the accumulators,
// the B fragments and the A fragments in flight fit the register budget: 256
// VGPRs at WPC = 2 workgroups (two waves) per SIMD, 168 at WPC = 3.  After
// the screening pass most runs of tiles to correlate are short, and 4-tile
// passes decode as fast as 8-tile ones.
#define MAX_NC 12
#define TUNE_NC BRKGA_TILE_CLASSES
#define TILE_ENTRY 4        // bytes per tile-table entry
#define SCREEN_DEFAULT 6    // mask rows of the screening pass unless set otherwise
__host__ __device__ constexpr int pass_tiles(int nc, int wpc = 2) {
    return wpc >= 3 ? (nc <= 4 ? 4 : (nc <= 8 ? 2 : 1)) : (nc <= 5 ? 8 : (nc <= 8 ? 4 : (nc <= MAX_NC ? 2 : 1)));
}

// One pass of TP tiles: correlate, combine the waves' zero masks, decode.
// Passes alternate between two mask buffers, so one barrier per pass is
// enough: a wave can only write buffer p%2 again (pass p+2) after every wave
// has passed the barrier of pass p+1, i.e. has finished decoding pass p.
template <int TP, int NC, bool F4, int ZW>
__device__ __forceinline__ int one_pass(lds_cu32 *band, int stride, const Fold &f, int a0, int a1,
                                        lds_cu32 *qbits, int qs, int t0, int xbase, int xmax, int nrows,
                                        lds_u64 *s_zw, int &parity) {
    lds_u64 *zp = s_zw + parity * 4 * ZW;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    corr_pass<TP, NC, F4>(band, stride, f.loff, a0, a1, qbits, qs, t0, TP, zp + wave * ZW);
    __syncthreads();
    parity ^= 1;
    return decode_zero<TP, ZW>(zp, xbase + (t0 << 5), xmax, nrows, f.rlog, f.segx);
}

// Screening pass of one wave over the tiles [t0, t0 + TW) of a staged window,
// with the S mask rows a_i = skip + floor(i * step / 2^16) only, step =
// floor(2^16 (h - 1 - skip) / (S - 1)) (evenly spaced over the rows [skip, h)
// that meet occupied grid rows; any subset would do): the same MFMA
// correlation as corr_pass, with the wave's own tiles instead of a share of
// the rows.  Every product is non-negative, so the count over a subset of
// the rows is zero wherever the full count is: a tile whose subset counts
// have no zero in any accumulator row (candidate or fold segment) has no
// placement, and the full passes skip it.  Returns bit t for tile t0 + t (t
// < tcnt) when some subset count of that tile is zero, uniform in the wave.
// Tiles from tcnt on are computed from clamped band words (wlast: the last
// word index the window's tiles read, before the lane's k-group offset) and
// not reported.  Zero counts past the valid region only set bits needlessly.
template <int TW, int NC, bool F4>
__device__ __forceinline__ unsigned screen_pass(lds_cu32 *band, int stride, int loff, int skip, unsigned long long step,
                                                int S, lds_cu32 *qbits, int qs, int t0, int tcnt, int wlast) {
    constexpr int NA = F4 ? TW + 2 * NC - 1 : TW + NC - 1;
    constexpr int NQ = F4 ? 2 * NC : NC + 1;
    const int lane = threadIdx.x & 63;
    const int g = lane >> 5;
    const int ks = g << 2;
    const int bs = F4 ? 32 - (lane & 31) : 32 - (lane & 31) + ks;
    typedef typename TileAcc<F4>::T AT;
    AT acc[TW];
#pragma unroll
    for (int t = 0; t < TW; t++) acc[t] = (AT){};
    lds_cu32 *base = band + loff + (F4 ? g : 0);
    lds_cu32 *qb = qbits + (F4 ? g : 0);
    // as in corr_pass: the next row's mask words are loaded over the current
    // ones once its B fragments are built, and band word j + D while word j
    // correlates (the last D of a row load the next row's first D)
    constexpr int D = NA < 2 ? NA : 2;
    unsigned N[D], qv[NQ];
    int a = skip;
#pragma unroll
    for (int j = 0; j < D; j++) N[j] = base[a * stride + min(t0 + j, wlast)];
#pragma unroll
    for (int c = 0; c < NQ; c++) qv[c] = qb[a * qs + c];
    for (int i = 0; i < S; i++) {
        typedef typename std::conditional<F4, v8i, v4i>::type FT;
        FT B[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) {
            if constexpr (F4) B[c] = nib_frag((unsigned)((((unsigned long long)qv[2 * c + 1] << 32) | qv[2 * c]) >> bs));
            else B[c] = bfrag_from_bits(qv[c], qv[c + 1], bs);
        }
        const int an = i + 1 < S ? skip + (int)((step * (unsigned)(i + 1)) >> 16) : a;
        lds_cu32 *rc = base + a * stride;
        lds_cu32 *rn = base + an * stride;
#pragma unroll
        for (int c = 0; c < NQ; c++) qv[c] = qb[an * qs + c];
        unsigned Aw[NA];
#pragma unroll
        for (int j = 0; j < D; j++) Aw[j] = N[j];
#pragma unroll
        for (int j = 0; j < NA; j++) {
            const int p = j + D;
            if (p < NA) Aw[p] = rc[min(t0 + p, wlast)];
            else N[p - NA] = rn[min(t0 + p - NA, wlast)];
            if constexpr (F4) {
                const v8i A = nib_frag(Aw[j]);
#pragma unroll
                for (int c = 0; c < NC; c++) {
                    const int t = j - 2 * c;
                    if (t >= 0 && t < TW) acc[t] = mfma_f4(A, B[c], acc[t]);
                }
            } else {
                const v4i A = frag_from_word(Aw[j] >> ks);
#pragma unroll
                for (int c = 0; c < NC; c++) {
                    const int t = j - c;
                    if (t >= 0 && t < TW) acc[t] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A, B[c], acc[t], 0, 0, 0);
                }
            }
        }
        a = an;
    }
    unsigned bits = 0;
#pragma unroll
    for (int t = 0; t < TW; t++) {
        if (t < tcnt) {
            // the counts are non-negative: one is zero iff their minimum is
            bool zero;
            if constexpr (F4) {
                float m = acc[t][0];
#pragma unroll
                for (int i = 1; i < 16; i++) m = __builtin_fminf(m, acc[t][i]);
                zero = m == 0.0f;
            } else {
                int m = acc[t][0];
#pragma unroll
                for (int i = 1; i < 16; i++) m = min(m, acc[t][i]);
                zero = m == 0;
            }
            if (__ballot(zero)) bits |= 1u << t;
        }
    }
    return bits;
}

// Tiles per screening step of one wave at most: the step's loop keeps the
// registers of a full pass of as many tiles, up to 4 (a wave's share of a
// window's tiles is short).
__host__ __device__ constexpr int screen_tiles(int nc, int wpc) {
    return pass_tiles(nc, wpc) < 4 ? pass_tiles(nc, wpc) : 4;
}

// Screen of a staged window (tiles [0, tl) of each segment) with S mask rows,
// run by the whole workgroup: the waves take consecutive quarters of the
// tiles, and the OR of their bits (through s_bits, 4 x u64 of LDS) is the
// mask of tiles the full passes must correlate.  mfma counts the wave's MFMA
// instructions.
template <int NC, bool F4, int WPC>
__device__ __attribute__((always_inline)) unsigned long long screen_window(lds_cu32 *band, int stride, const Fold &f,
                                                                           int h, int S, lds_cu32 *qbits, int qs, int tl,
                                                                           lds_u64 *s_bits, lds_u64 &mfma) {
    constexpr int TW = screen_tiles(NC, WPC);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int q = (tl + 3) >> 2;
    const int tb = min(tl, wave * q), te = min(tl, tb + q);
    const int wlast = F4 ? tl + 2 * NC - 3 : tl + NC - 2;
    const unsigned long long step = S > 1 ? ((unsigned long long)(h - 1 - f.skip) << 16) / (unsigned)(S - 1) : 0ull;
    unsigned long long bits = 0;
    for (int t0 = tb; t0 < te;) {
        // a step takes the wave's remaining tiles up to TW, so no tile is
        // computed only to be dropped
        const int r = te - t0;
        unsigned b;
        int tw;
        if (TW >= 4 && r >= 4) {
            b = screen_pass<(TW >= 4 ? 4 : 1), NC, F4>(band, stride, f.loff, f.skip, step, S, qbits, qs, t0, 4, wlast);
            tw = 4;
        } else if (TW >= 3 && r == 3) {
            b = screen_pass<(TW >= 3 ? 3 : 1), NC, F4>(band, stride, f.loff, f.skip, step, S, qbits, qs, t0, 3, wlast);
            tw = 3;
        } else if (TW >= 2 && r >= 2) {
            b = screen_pass<(TW >= 2 ? 2 : 1), NC, F4>(band, stride, f.loff, f.skip, step, S, qbits, qs, t0, 2, wlast);
            tw = 2;
        } else {
            b = screen_pass<1, NC, F4>(band, stride, f.loff, f.skip, step, S, qbits, qs, t0, 1, wlast);
            tw = 1;
        }
        bits |= (unsigned long long)b << t0;
        mfma += (unsigned long long)tw * NC * S;
        t0 += tw;
    }
    if ((threadIdx.x & 63) == 0) s_bits[wave] = bits;
    __syncthreads();
    return s_bits[0] | s_bits[1] | s_bits[2] | s_bits[3];
}

template <bool F4, int WPC>
__device__ __forceinline__ unsigned long long screen_dispatch(int nc, lds_cu32 *band, int stride, const Fold &f, int h,
                                                              int S, lds_cu32 *qbits, int qs, int tl, lds_u64 *s_bits,
                                                              lds_u64 &mfma) {
#define SCASE(K) case K: return screen_window<K, F4, WPC>(band, stride, f, h, S, qbits, qs, tl, s_bits, mfma);
    switch (nc) {
        SCASE(1) SCASE(2) SCASE(3) SCASE(4) SCASE(5) SCASE(6) SCASE(7) SCASE(8) SCASE(9) SCASE(10) SCASE(11) SCASE(12)
        default:
            return tl >= 64 ? ~0ull : (1ull << tl) - 1ull;
    }
#undef SCASE
}

// The passes over one staged x-window (tiles [0, tl) of each segment), run
// by the whole workgroup, over the tiles set in `need` (bit t: tile t; the
// others are known to hold no placement).  Each run of consecutive needed
// tiles is cut into passes of 8/4/2/1 tiles; a remainder of 6 or 7 tiles
// takes one 8-tile pass and a remainder of 3 one 4-tile pass (one or two
// tiles of MFMAs cost less than the extra pass's zero-mask exchange and
// barrier), so a pass may also cover up to two tiles after its run.  Tiles
// past the staged window read whatever the band row holds there (at most two
// words past the staged ones, inside the row's stride); the caller's xmax
// ends at the window's last column, so their counts are never decoded.
// Returns the window's minimum key; with `early`, stops once candidate 0 has
// a zero left of every needed tile still to come (a key below xbase + 32 *
// t_end; in a folded group the later passes of segment 0 come before the
// current ones of the other segments).  mfma counts the MFMA instructions
// issued.
template <int NC, bool F4, int WPC>
__device__ __attribute__((always_inline)) int window_passes(lds_cu32 *band, int stride, const Fold &f, int h,
                             lds_cu32 *qbits, int qs, unsigned long long need, int xbase, int xmax, int nrows,
                             lds_u64 *s_zw, lds_u64 &mfma, int tpcap, bool early) {
    constexpr int TPM = pass_tiles(NC, WPC);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int a0 = f.skip + (((h - f.skip) * wave) >> 2), a1 = f.skip + (((h - f.skip) * (wave + 1)) >> 2);
    int key = KEY_NONE, parity = 0;
    while (need) {
        const int t0 = (int)__builtin_ctzll(need);
        const unsigned long long rest = ~(need >> t0);
        const int rem = rest ? (int)__builtin_ctzll(rest) : 64 - t0;   // needed tiles from t0 on, in a run
        int k, tp;
        if (TPM >= 8 && tpcap >= 8 && rem >= 6) {
            k = one_pass<(TPM >= 8 ? 8 : 1), NC, F4, zw_wave(WPC)>(band, stride, f, a0, a1, qbits, qs, t0, xbase, xmax, nrows, s_zw, parity);
            tp = 8;
        } else if (TPM >= 4 && tpcap >= 4 && rem >= 3) {
            k = one_pass<(TPM >= 4 ? 4 : 1), NC, F4, zw_wave(WPC)>(band, stride, f, a0, a1, qbits, qs, t0, xbase, xmax, nrows, s_zw, parity);
            tp = 4;
        } else if (TPM >= 2 && tpcap >= 2 && rem >= 2) {
            k = one_pass<(TPM >= 2 ? 2 : 1), NC, F4, zw_wave(WPC)>(band, stride, f, a0, a1, qbits, qs, t0, xbase, xmax, nrows, s_zw, parity);
            tp = 2;
        } else {
            k = one_pass<1, NC, F4, zw_wave(WPC)>(band, stride, f, a0, a1, qbits, qs, t0, xbase, xmax, nrows, s_zw, parity);
            tp = 1;
        }
        mfma += (unsigned long long)tp * (a1 - a0) * NC;
        key = k < key ? k : key;
        const int te = t0 + tp;
        need = te >= 64 ? 0ull : need & (~0ull << te);
        if (early && key < xbase + (te << 5)) break;
    }
    return key;
}

// Generic chunk count (NC above the compiled set): one tile per pass, runtime
// chunk loop, one band word (int8) or two (fp4) per MFMA.
template <bool F4, int WPC>
__device__ int window_passes_generic(lds_cu32 *band, int stride, const Fold &f, int h, int nc,
                                     lds_cu32 *qbits, int qs, int tl, int xbase, int xmax, int nrows,
                                     lds_u64 *s_zw, lds_u64 &mfma, bool early) {
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = lane >> 5, ks = g << 2, bs = F4 ? 32 - (lane & 31) : 32 - (lane & 31) + ks;
    const int a0 = f.skip + (((h - f.skip) * wave) >> 2), a1 = f.skip + (((h - f.skip) * (wave + 1)) >> 2);
    int key = KEY_NONE, parity = 0;
    typedef typename TileAcc<F4>::T AT;
    for (int t0 = 0; t0 < tl; t0++) {
        AT acc[1] = {(AT){}};
        lds_cu32 *base = band + f.loff + t0 + (F4 ? g : 0);
        for (int a = a0; a < a1; a++) {
            lds_cu32 *rp = base + a * stride;
            lds_cu32 *q = qbits + a * qs + (F4 ? g : 0);
            for (int c = 0; c < nc; c++) {
                if constexpr (F4) {
                    const unsigned bw = (unsigned)((((unsigned long long)q[2 * c + 1] << 32) | q[2 * c]) >> bs);
                    acc[0] = mfma_f4(nib_frag(rp[2 * c]), nib_frag(bw), acc[0]);
                } else {
                    acc[0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(frag_from_word(rp[c] >> ks),
                                                                   bfrag_from_bits(q[c], q[c + 1], bs), acc[0], 0, 0, 0);
                }
            }
        }
        mfma += (unsigned long long)(a1 - a0) * nc;
        constexpr int ZW = zw_wave(WPC);
        lds_u64 *zp = s_zw + parity * 4 * ZW;
        store_zero_masks<1>(acc, 1, zp + wave * ZW);
        __syncthreads();
        parity ^= 1;
        const int k = decode_zero<1, ZW>(zp, xbase + (t0 << 5), xmax, nrows, f.rlog, f.segx);
        key = k < key ? k : key;
        if (early && key < xbase + ((t0 + 1) << 5)) break;
    }
    return key;
}

template <bool F4, int WPC>
__device__ __forceinline__ int window_dispatch(int nc, lds_cu32 *band, int stride, const Fold &f, int h,
                               lds_cu32 *qbits, int qs, int tl, unsigned long long need, int xbase, int xmax,
                               int nrows, lds_u64 *s_zw, lds_u64 &mfma, int tpcap, bool early) {
#define CASE(K) case K: return window_passes<K, F4, WPC>(band, stride, f, h, qbits, qs, need, xbase, xmax, nrows, s_zw, mfma, tpcap, early);
    switch (nc) {
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10) CASE(11) CASE(12)
        default:
            return window_passes_generic<F4, WPC>(band, stride, f, h, nc, qbits, qs, tl, xbase, xmax, nrows, s_zw, mfma, early);
    }
#undef CASE
}

// LDS row stride (32-bit words) of a window of tw x-tiles for a mask whose
// Toeplitz chunks span cw words (nc for int8 chunks of 32 columns, 2 * nc4 for
// fp4 chunks of 64).  A window of tiles [tw0, tw0 + tw) reads the words [tw0,
// tw0 + tw + cw - 1) of its rows; they are staged as the 16-byte-aligned words
// from tw0 rounded down (up to 3 words in front) to the end rounded up, and a
// pass rounded up past the window reads up to two words further: tw + cw + 5
// words cover all of it.  The stride is odd, so the 32 band rows one
// A-fragment load touches fall into 32 distinct banks.
__host__ __device__ __forceinline__ int band_stride(int tw, int cw) {
    return (tw + cw + 5) | 1;
}

// Words kept free after the band's last row.
#define BAND_SLACK 16

// Largest x-window (in tiles, <= cap) whose 32-row band fits (with
// BAND_SLACK words to spare); shrinks the window rather than the row group.
__host__ __device__ __forceinline__ int band_window(int T, int cw, int h, int band_words, int cap,
                                                    int *stride_out, int *rcap_out) {
    int tw = T < cap ? T : cap;
    int st, rc;
    for (;;) {
        st = band_stride(tw, cw);
        rc = (band_words - BAND_SLACK) / st;
        if (rc >= h + GROUP_ROWS - 1 || tw == 1) break;
        tw = (tw + 1) >> 1;
    }
    *stride_out = st;
    *rcap_out = rc;
    return tw;
}

#define WINDOW_CAP 64   // x-tiles per staged window at most

// Copy n elements (16-byte vectors by default) with the whole workgroup,
// COPY_BATCH loads in flight per thread: load(k) returns element k,
// store(k, v) puts it.
#define COPY_BATCH 8
template <typename T = v4i, typename Load, typename Store>
__device__ __forceinline__ void wg_copy(int n, Load load, Store store) {
    for (int k0 = threadIdx.x; k0 < n; k0 += COPY_BATCH * WG_THREADS) {
        T v[COPY_BATCH];
#pragma unroll
        for (int u = 0; u < COPY_BATCH; u++) {
            const int k = k0 + u * WG_THREADS;
            if (k < n) v[u] = load(k);
        }
#pragma unroll
        for (int u = 0; u < COPY_BATCH; u++) {
            const int k = k0 + u * WG_THREADS;
            if (k < n) store(k, v[u]);
        }
    }
}

// Collision test of one rotation against one bin, run by the whole workgroup.
// grid: the bin's bit rows (gpw words each).  cand: candidate top rows,
// descending.  Groups of up to 32 candidates are staged into the LDS band (bit
// rows ylo..ytop+h-1, one x-window at a time) and correlated.  Returns
// (candidate index)*XKEY + x of the bottom-most, then left-most, placement with
// zero overlap, or KEY_NONE.  Uniform across the workgroup.  Grid rows above r0
// are empty (no part reaches above the bin's top-most placement), so mask rows
// that only meet them are neither staged nor correlated, except in the dense
// mode.  nc: Toeplitz chunks of the operand format (32 columns for int8, 64
// for fp4), cw: the words they span, qs: mask words per row in qbits_g.
template <bool DENSE, bool F4, int WPC>
__device__ __attribute__((always_inline)) int test_rotation(const unsigned *grid, int gpw, int W, int h, int w, int nc,
                             int cw, int qs, const unsigned *qbits_g, const short *cand, int ncand, int r0,
                             lds_u32 *band, int band_words, lds_u32 *s_q, lds_u64 *s_zw,
                             lds_u64 &groups, lds_u64 &mfma, lds_u64 *ph,
                             int tpcap, int wcap, int screen) {
    const int tid = threadIdx.x, lane = tid & 63;
    // mask bits of the rotation -> LDS (read by every wave for every group)
    // with direct-to-LDS loads: they take no registers and are not waited
    // for here; the first band staging's wait for its own loads (issued
    // later, completed in order) and its barrier cover them
    {
        const int nq = h * qs, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
        for (int k0 = wave << 6; k0 < nq; k0 += WG_THREADS)
            if (k0 + lane < nq)
                __builtin_amdgcn_global_load_lds((const void *)(qbits_g + k0 + lane), (LDSP void *)(s_q + k0), 4, 0, 0);
    }
    const int T = (W - w + 1 + 31) >> 5;
    int stride, rcap;
    const int TW = band_window(T, cw, h, band_words, wcap, &stride, &rcap);
    const int maxrows = min(GROUP_ROWS, rcap - h + 1);
    int found = KEY_NONE;
    for (int g0 = 0; g0 < ncand;) {
        const int ytop = uni(cand[g0]);
        // the group: candidate g0 and the following ones (at most maxrows in
        // all) whose rows still fit the band; candidates descend, so these
        // form a prefix and one ballot counts them
        const int kk = g0 + lane;
        const bool take = lane == 0 ||
                          (lane < maxrows && kk < ncand && ytop + h - cand[min(kk, ncand - 1)] <= rcap);
        const int g1 = g0 + (int)__popcll(__ballot(take) & 0xffffffffull);
        const int ylo = uni(cand[g1 - 1]);
        const int rows = ytop + h - ylo;
        const int nrows = g1 - g0;
        int rl = 5;   // smallest R = 2^rl >= nrows
        while (rl > 0 && (1 << (rl - 1)) >= nrows) rl--;
        // mask rows a < r0 - ytop meet empty rows for every candidate
        const int skip = DENSE ? 0 : min(h, max(0, r0 - ytop));
        int gkey = skip == h ? 0 : KEY_NONE;   // the part lies in empty rows: candidate 0 at x = 0
        if (tid == 0) groups += 1;
        for (int tw0 = 0; tw0 < T && skip < h; tw0 += TW) {
            unsigned long long t0 = tid == 0 ? tick() : 0;
            const int tl = min(TW, T - tw0);
            // words [tw0, tw0 + tl + cw - 1) of each row, as the 16-byte vectors
            // from tw0 rounded down (c0) to the end rounded up (<= gpw)
            const int c0 = tw0 & ~3;
            const int q4 = (((tw0 + tl + cw - 1 + 3) & ~3) - c0) >> 2;
            const DivBy byq(q4);
            const unsigned *src = grid + (long long)(ylo + skip) * gpw + c0;
            lds_u32 *dst = band + skip * stride;
            wg_copy((rows - skip) * q4,
                    [&](int k) { const int r = byq(k), q = k - r * q4;
                                 return *(const v4i *)(src + (long long)r * gpw + (q << 2)); },
                    [&](int k, v4i v) { const int r = byq(k), q = k - r * q4;
                                        lds_u32 *d = dst + r * stride + (q << 2);
                                        d[0] = (unsigned)v[0]; d[1] = (unsigned)v[1];
                                        d[2] = (unsigned)v[2]; d[3] = (unsigned)v[3]; });
            __syncthreads();
            unsigned long long t1 = tid == 0 ? tick() : 0;
            if (tid == 0) ph[2] += t1 - t0;
            Fold f;
            f.rlog = rl;
            while ((32 >> f.rlog) > tl) f.rlog++;   // no more segments than tiles
            const int tf = (tl + (32 >> f.rlog) - 1) >> (5 - f.rlog);   // tiles per segment
            const int m = lane & 31;
            f.segx = min((m >> f.rlog) * tf, tl - tf) << 5;
            f.loff = (cand[g0 + min(m & ((1 << f.rlog) - 1), nrows - 1)] - ylo) * stride + (tw0 - c0) +
                     (f.segx >> 5);
            f.skip = skip;
            // screening pass (not in the dense mode): the tiles whose counts over
            // `screen` evenly spaced mask rows have no zero hold no placement
            unsigned long long need = tf >= 64 ? ~0ull : (1ull << tf) - 1ull;
            if (!DENSE && screen > 0 && h - skip > 2 * screen) {
                need = screen_dispatch<F4, WPC>(nc, band, stride, f, h, screen, s_q, qs, tf, s_zw + 2 * 4 * zw_wave(WPC), mfma);
                if (tid == 0) { const unsigned long long t = tick(); ph[4] += t - t1; t1 = t; }
            }
            const int k = need ? window_dispatch<F4, WPC>(nc, band, stride, f, h, s_q, qs, tf, need, tw0 << 5,
                                                     min(W - w, ((tw0 + tl) << 5) - 1), nrows, s_zw, mfma, tpcap,
                                                     !DENSE)
                               : KEY_NONE;
            gkey = k < gkey ? k : gkey;
            if (tid == 0) ph[3] += tick() - t1;
            if (!DENSE && gkey < XKEY) break;   // candidate 0 has a zero: nothing further left can win
        }
        if (gkey != KEY_NONE) {
            if (found == KEY_NONE) found = (g0 + gkey / XKEY) * XKEY + gkey % XKEY;
            if (!DENSE) break;
        }
        g0 = g1;
    }
    return found;
}

// Longest free run (longest run of zero cells) of `rows` grid bit rows in LDS
// (row r at band + r * rstride words); numba_utils.py:161-188 computes the same
// value sequentially per row.  L = 2^lg lanes per row (at most 64, chosen
// per call from the row count and the row width), lane j scanning the wpl =
// ceil(ceil(W/32) / L) words [j*wpl, j*wpl + wpl) (cells past W count as
// occupied): per word the
// free cells at its start (ctz of the occupied bits) and end (clz) and its
// longest inner run (longest_free_in_word); the lane keeps its free cells at
// the segment start (pre) and end (suf) and its longest run (best), and lg
// ordered butterfly steps join the L segments of the row.  put(r, v) receives
// each row's value from one lane.
__device__ __forceinline__ int longest_free_in_word(unsigned o) {   // o: occupied bits
    // R_s: bit i set iff cells i .. i+s-1 are free; R_{L+s} = R_L & (R_s >> L),
    // so binary lifting from L = 0 finds the longest run (< 32 unless o == 0)
    const unsigned r1 = ~o, r2 = r1 & (r1 >> 1), r4 = r2 & (r2 >> 2), r8 = r4 & (r4 >> 4), r16 = r8 & (r8 >> 8);
    if (!o) return 32;
    unsigned cur = r16;
    int L = 16;
    if (!cur) { cur = r1 | ~r1; L = 0; }
    unsigned c = cur & (r8 >> L);
    if (c) { cur = c; L += 8; }
    c = cur & (r4 >> L);
    if (c) { cur = c; L += 4; }
    c = cur & (r2 >> L);
    if (c) { cur = c; L += 2; }
    c = cur & (r1 >> L);
    if (c) L += 1;
    return L;
}

template <typename Put>
__device__ __forceinline__ void rows_free_run(lds_cu32 *band, int rstride, int rows, int W, Put put) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int nw = (W + 31) >> 5;   // words of a row
    // lanes per row: the fewest rounds over the rows, then the cheapest round
    // (a word costs about 4 units per lane, a butterfly step about 10: two
    // dependent cross-lane shuffles)
    int lg = 0, cost = 1 << 30;
    for (int l = 0; l <= 6 && (1 << l) < 2 * nw; l++) {
        const int rounds = (rows + (256 >> l) - 1) / (256 >> l);
        const int c = rounds * (4 * ((nw + (1 << l) - 1) >> l) + 10 * l);
        if (c < cost) { cost = c; lg = l; }
    }
    const int L = 1 << lg, wpl = (nw + L - 1) >> lg, rpw = 64 >> lg;
    const int sub = lane & (L - 1);
    for (int r0 = wave * rpw; r0 < rows; r0 += 4 * rpw) {
        const int r = r0 + (lane >> lg);
        lds_cu32 *rp = band + min(r, rows - 1) * rstride;
        int pre = 0, run = 0, best = 0;
        bool first = true;
        for (int i = 0; i < wpl; i++) {
            const int q = sub * wpl + i;
            unsigned o = 0xFFFFFFFFu;   // bit t: cell 32q + t occupied
            if (q < nw) {
                o = rp[q];
                const int valid = W - (q << 5);
                if (valid < 32) o |= 0xFFFFFFFFu << valid;
            }
            const bool z = o == 0;
            const int lead = z ? 32 : __builtin_ctz(o);
            const int trail = z ? 32 : __builtin_clz(o);
            const int inner = longest_free_in_word(o);
            if (first && !z) pre = run + lead;
            if (!z) best = max(best, max(run + lead, inner));
            run = z ? run + 32 : trail;
            first = first && z;
        }
        if (first) pre = run;   // every cell of the segment free: pre = suf = its length
        int suf = run, len = wpl << 5;
        for (int d = 1; d < L; d <<= 1) {   // join segments [.. d) and [d .. 2d)
            const int ps = __shfl_xor(pre | (suf << 16), d, 64);
            const int pb = __shfl_xor(best, d, 64);
            const int opre = ps & 0xFFFF, osuf = ps >> 16;
            const bool left = (sub & d) == 0;
            const int lpre = left ? pre : opre, lsuf = left ? suf : osuf;
            const int rpre = left ? opre : pre, rsuf = left ? osuf : suf;
            best = max(max(best, pb), lsuf + rpre);
            pre = lpre == len ? len + rpre : lpre;
            suf = rsuf == len ? len + lsuf : rsuf;
            len <<= 1;
        }
        if (sub == 0 && r < rows) put(r, max(best, max(pre, suf)));
    }
}

struct __align__(8) PlaceRes { int has, y, x, pad; };

// Context order.  A launch decodes n_chrom * M contexts on a fixed set of
// persistent workgroups; contexts of one population differ several-fold in
// work, and taking the heaviest first (longest-processing-time-first list
// scheduling) keeps a long context from starting last.  The expected work of
// a context is the summed bounding-box area of the masks it places, counted
// into ORDER_BUCKETS buckets over [0, twice the mean); a counting sort from
// the heaviest bucket down gives the order.  Any order gives the same
// results: a context's outputs do not depend on the workgroup or the time it
// runs.
#define ORDER_BUCKETS 1024

// One wave per context, four per workgroup: the lanes take the slots and a
// butterfly adds their costs; bucket = cost * ORDER_BUCKETS * M / (2 *
// cost_all), cost_all = the summed cost of every slot (so the M contexts of a
// chromosome average cost_all / M).  hist starts at zero (order_scatter_kernel
// clears it after use; it is zeroed when first allocated).
__global__ __launch_bounds__(256) void order_cost_kernel(DevProblem P, const void *keys, int key_f64,
                                                         long long n_chrom, long long cost_all, int *bucket,
                                                         int *hist) {
    const int M = P.n_machines, n = P.n_parts, lane = threadIdx.x & 63;
    const long long ctx = blockIdx.x * 4LL + (threadIdx.x >> 6);
    if (ctx >= n_chrom * M) return;   // the whole wave
    const long long chrom = ctx / M;
    const int m = (int)(ctx % M);
    long long cost = 0;
    for (int j = lane; j < n; j += 64)
        if (machine_of(load_key(keys, key_f64, chrom * 2LL * n + n + j), P.thresholds, M) == m)
            cost += P.type_cost[P.slot_type[j]];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) cost += __shfl_xor(cost, d, 64);
    if (lane == 0) {
        const int b = (int)min((long long)ORDER_BUCKETS - 1, cost * ORDER_BUCKETS * M / (2 * cost_all));
        bucket[ctx] = b;
        atomicAdd(&hist[b], 1);
    }
}

// One workgroup of ORDER_BUCKETS threads: start of each bucket in the order
// (heaviest bucket first, an exclusive scan), then every context takes the
// next slot of its bucket.  Clears the histogram and the decode's queue
// counter for the next launch.  Should the histogram not count n_ctx contexts
// (it is left dirty only if an earlier launch failed between the two order
// kernels), the order is the identity.
__global__ __launch_bounds__(ORDER_BUCKETS) void order_scatter_kernel(const int *bucket, int *hist, long long n_ctx,
                                                                      int *order, unsigned int *queue) {
    __shared__ int start[ORDER_BUCKETS];
    const int t = threadIdx.x;
    const int v = hist[ORDER_BUCKETS - 1 - t];   // position t in descending bucket order
    hist[ORDER_BUCKETS - 1 - t] = 0;
    if (t == 0) *queue = 0u;
    start[t] = v;
    __syncthreads();
    for (int d = 1; d < ORDER_BUCKETS; d <<= 1) {   // inclusive scan (Hillis-Steele)
        const int o = t >= d ? start[t - d] : 0;
        __syncthreads();
        start[t] += o;
        __syncthreads();
    }
    const int excl = start[t] - v;
    const bool counted = start[ORDER_BUCKETS - 1] == n_ctx;
    __syncthreads();
    start[ORDER_BUCKETS - 1 - t] = excl;   // indexed by bucket again
    __syncthreads();
    for (long long c = t; c < n_ctx; c += ORDER_BUCKETS) {
        if (!counted) { order[c] = (int)c; continue; }
        const int pos = atomicAdd(&start[bucket[c]], 1);
        if (pos < n_ctx) order[pos] = (int)c;   // always true: the buckets count every context once
    }
}

// Free runs of grid rows packed four to a word: vq[y] holds the free runs of
// rows y, y+64, y+128 and y+192 (16 bits each, row y in the low bits; rows
// from H on count as 0) for y < H.  get(y) returns row y's free run.  Whole
// workgroup; the caller synchronises before the vacancy filter reads vq.
template <typename Get>
__device__ __forceinline__ void vacancy_quads(lds_u64 *vq, int H, Get get) {
    for (int y = threadIdx.x; y < H; y += WG_THREADS) {
        unsigned long long q = 0;
#pragma unroll
        for (int s = 0; s < 4; s++)
            if (y + 64 * s < H) q |= (unsigned long long)(unsigned short)get(y + 64 * s) << (16 * s);
        vq[y] = q;
    }
}

// Vacancy filter of one rotation over top rows [y0, y0 + 256 NQ), run by one
// wave.  A top row y passes when every grid row y+a (a < h) has a free run >=
// the mask row's longest 1-run (numba_utils.py:131-157 applies the same test
// to decide whether the bin is tried).  The top rows form words of 64, word
// K = 4k + s holding y0 + 64K + l in lane l: for mask row a (density d), lane
// l's one 8-byte read vq[y0 + 256k + l + a] gives the free runs v_s of the
// grid rows its four top rows of quad k meet (consecutive lanes read
// consecutive words: no bank conflicts), and packed 16-bit saturating
// subtractions and three-input ORs keep acc_s = OR over the rows so far of
// max(d - v_s, 0); top row s passes while acc_s == 0 (free runs and densities
// are at most 4096: unsigned 16-bit fields).  den holds (density << 16) | 8a
// for the rotation's mask rows a (8a: the row's byte offset in vq), the
// densest first, so that a bin with no room fails within the first steps
// (the result is an OR over the rows, whatever their order); the entries of
// rows b .. b+63 are held one per lane and a step's come from registers
// (readlane).  VAC_STEP mask rows per
// step keep their reads in flight together, and after each step the wave
// stops once no top row is left.  ny: valid top rows from y0 on (H - h + 1 -
// y0); the fields of top rows past them start at 1 (failed), and a lane
// whose quad has none reads from y0 instead, so every read stays below H.
// The passing top rows go to cand in descending order; returns their count.
#ifndef VAC_STEP
#define VAC_STEP 8   // mask rows per step of the vacancy filter (their reads are in flight together)
#endif
static_assert(64 % VAC_STEP == 0 && VAC_STEP % 2 == 0, "vacancy steps tile a 64-row block in row pairs");
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
template <int NQ>
__device__ __forceinline__ int vacancy_rotation(lds_cu64 *vq, lds_ci32 *den, int h, int ny, short *cand,
                                                int y0) {
    const int lane = threadIdx.x & 63;
    u16x2 acc[NQ][2];
    LDSP const unsigned char *p[NQ];
#pragma unroll
    for (int k = 0; k < NQ; k++) {
        const int t = (k << 8) + lane;   // the lane's top row of word 4k, from y0
#pragma unroll
        for (int j = 0; j < 2; j++)
            acc[k][j] = (u16x2){(unsigned short)(t + 128 * j < ny ? 0 : 1),
                                (unsigned short)(t + 128 * j + 64 < ny ? 0 : 1)};
        p[k] = (LDSP const unsigned char *)(vq + y0 + (t < ny ? t : 0));
    }
    bool alive = true;
    for (int a0 = 0; a0 < h && alive; a0 += 64) {
        const int dl = den[min(a0 + lane, h - 1)];
        const int a1 = min(h, a0 + 64);
        for (int b = a0; b < a1; b += VAC_STEP) {
            u16x2 d[VAC_STEP];
            int ab[VAC_STEP];
#pragma unroll
            for (int j = 0; j < VAC_STEP; j++) {
                // (b - a0 + j < 64: a block is 64 rows, VAC_STEP divides 64;
                // lanes past the rotation's last row hold that row again)
                const int e = __builtin_amdgcn_readlane(dl, b - a0 + j);
                const unsigned short dv = (unsigned short)((unsigned)e >> 16);
                d[j] = (u16x2){dv, dv};
                ab[j] = e & 0xffff;
            }
#pragma unroll
            for (int j = 0; j < VAC_STEP; j += 2)   // rows in pairs: one three-input OR per pair
#pragma unroll
                for (int k = 0; k < NQ; k++) {
                    const unsigned long long q0 = *(lds_cu64 *)(p[k] + ab[j]);
                    const unsigned long long q1 = *(lds_cu64 *)(p[k] + ab[j + 1]);
                    const u16x2 l0 = __builtin_bit_cast(u16x2, (unsigned)q0), h0 = __builtin_bit_cast(u16x2, (unsigned)(q0 >> 32));
                    const u16x2 l1 = __builtin_bit_cast(u16x2, (unsigned)q1), h1 = __builtin_bit_cast(u16x2, (unsigned)(q1 >> 32));
                    acc[k][0] = acc[k][0] | __builtin_elementwise_sub_sat(d[j], l0) | __builtin_elementwise_sub_sat(d[j + 1], l1);
                    acc[k][1] = acc[k][1] | __builtin_elementwise_sub_sat(d[j], h0) | __builtin_elementwise_sub_sat(d[j + 1], h1);
                }
            // the step's reads first, then its arithmetic: the reads are in
            // flight together rather than each waited for before the next
            __builtin_amdgcn_sched_group_barrier(0x0100, VAC_STEP * NQ, 0);
            __builtin_amdgcn_sched_group_barrier(0x0002, 4 * VAC_STEP * NQ, 0);
            u16x2 m = __builtin_elementwise_min(acc[0][0], acc[0][1]);
#pragma unroll
            for (int k = 1; k < NQ; k++)
                m = __builtin_elementwise_min(m, __builtin_elementwise_min(acc[k][0], acc[k][1]));
            if (!__ballot(m.x == 0 || m.y == 0)) { alive = false; break; }
        }
    }
    if (!alive) return 0;
    // the list in scalar form, descending: bit-reversed, word K gives lane l
    // the row y0 + 64K + 63 - l (inverse ballot) and its rank among the
    // passing rows above it (mbcnt)
    int nc_ = 0;
#pragma unroll
    for (int K = 4 * NQ - 1; K >= 0; K--) {
        const u16x2 v = acc[K >> 2][(K >> 1) & 1];
        const unsigned long long c = __ballot(((K & 1) ? v.y : v.x) == 0);
        if (c) {
            const unsigned long long rv = __builtin_bitreverse64(c);
            const int pos = __builtin_amdgcn_mbcnt_hi((unsigned)(rv >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)rv, 0));
            if (__builtin_amdgcn_inverse_ballot_w64(rv)) cand[nc_ + pos] = (short)(y0 + (K << 6) + 63 - lane);
            nc_ += __popcll(c);
        }
    }
    return nc_;
}

// A part type's rotations, staged in LDS once per part (load_part): every
// vacancy filter, collision test and stamp of the part reads them there
// instead of following the problem tables in HBM (a chain of dependent loads
// per test otherwise).
struct RotInfo {
    int h, w, nc, nc4;                 // mask rows and columns; Toeplitz chunks (int8, fp4)
    int tiles, window, screen, doff;   // the machine's tile-table entry; densities' offset in s_den
    long long qoff;                    // mask bits of the rotation in P.qbits
};
typedef LDSP RotInfo lds_rot;

// Rotations rb..rb+nr-1 of a part's type into ri (thread r fills rotation r;
// tiles: the context machine's tile table, TUNE_NC entries of TILE_ENTRY
// bytes) and their densities ((density << 16) | 8 row, densest rows first)
// into s_den, rotation after rotation (rotation r at ri[r].doff).
// brkga_create appends the densities in rotation order and a type's rotations
// are consecutive, so they arrive in one contiguous copy by the whole
// workgroup.  Ends synchronised.
__device__ __forceinline__ void load_part(const DevProblem &P, const unsigned char *tiles, int rb, int nr,
                                          int *s_den, lds_rot *ri) {
    const int tid = threadIdx.x;
    // offsets from dens_off: the copy need not wait for ri
    const long long dens0 = P.dens_off[rb];
    const int tot = (int)(P.dens_off[rb + nr - 1] - dens0) + P.rot_h[rb + nr - 1];
    if (tid < nr) {
        const int rg = rb + tid;
        const int nc = P.rot_nc[rg];
        const unsigned char *tc = tiles + min(nc, TUNE_NC - 1) * TILE_ENTRY;
        const int doff = (int)(P.dens_off[rg] - dens0);
        ri[tid].h = P.rot_h[rg];
        ri[tid].w = P.rot_w[rg];
        ri[tid].nc = nc;
        ri[tid].nc4 = P.rot_nc4[rg];
        ri[tid].qoff = P.qbits_off[rg];
        ri[tid].tiles = tc[0];
        ri[tid].window = tc[1];
        ri[tid].screen = tc[2];
        ri[tid].doff = doff;
    }
    const int *src = P.dens + dens0;
    wg_copy<int>(tot, [&](int k) { return src[k]; }, [&](int k, int v) { s_den[k] = v; });
    __syncthreads();
}

// Vacancy filter of a part's rotations against one bin, wave r: rotation r.
// A top row y can only hold the part if every grid row y+a has a free run >=
// the mask row's longest 1-run (numba_utils.py:131-157 applies the same test
// to decide whether the bin is tried).  The candidate top rows, descending,
// go to s_cand[r * hmax ...] and their count to s_ncand[r]; the caller
// synchronises.  Dense mode lists every row of the valid region (same results).
// Rows above the bin's top-most occupied row (minrow) are empty, so a top row
// y < minrow - h + 1 keeps the whole part in empty rows and always passes: only
// the top rows from there on are tested, and the others follow them in the
// list (they are the highest, so the list stays descending).
template <bool DENSE>
__device__ __forceinline__ void vacancy_filter(lds_rot *ri, int nr, int H, int W, int minrow,
                                               lds_cu64 *vq, const int *s_den, short *s_cand, int hmax,
                                               int *s_ncand) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (wave < nr) {
        const int h = uni(ri[wave].h), w = uni(ri[wave].w);
        short *cand = s_cand + wave * hmax;
        const int *den = s_den + uni(ri[wave].doff);
        int nc_ = 0;
        if (h <= H && w <= W) {
            const int ny = H - h + 1;
            const int yf = DENSE ? 0 : min(ny, max(0, uni(minrow) - h + 1));   // top rows below yf pass
            lds_ci32 *dn = (lds_ci32 *)den;
            // passes of at most 512 top rows (one quad when 256 cover the
            // rest), the highest first so that the list stays descending; one
            // pass when H - h < 512
            for (int c = (ny - yf - 1) >> 9; c >= 0; c--) {
                const int y0 = yf + (c << 9);
                short *cc = cand + nc_;
                if (ny - y0 <= 256) nc_ += vacancy_rotation<1>(vq, dn, h, ny - y0, cc, y0);
                else nc_ += vacancy_rotation<2>(vq, dn, h, ny - y0, cc, y0);
            }
            for (int k = lane; k < yf; k += 64) cand[nc_ + k] = (short)(yf - 1 - k);
            nc_ += yf;
            if (DENSE && nc_ > 0) {   // full valid region, bottom-up
                nc_ = H - h + 1;
                for (int k = lane; k < nc_; k += 64) cand[k] = (short)(H - h - k);
            }
        }
        if (lane == 0) s_ncand[wave] = nc_;
    }
}

// Number of leading candidates (top rows, descending) of a rotation of height
// h that could still beat the choice (bl, brow) made so far: new enclosure
// length max(enc, H - y) smaller, or equal with a bottom row y + h - 1 not
// above brow.  Along the list the length never falls and the bottom row
// always falls, so the losers form a suffix and the first one ends the
// prefix.  Every wave computes the same count.
__device__ __forceinline__ int winning_prefix(const short *cand, int ncand, int enc, int H, int h, int bl,
                                              int brow) {
    const int lane = threadIdx.x & 63;
    for (int k0 = 0; k0 < ncand; k0 += 64) {
        const int k = k0 + lane;
        bool lose = false;
        if (k < ncand) {
            const int y = cand[k];
            const int nl = enc > H - y ? enc : H - y;
            lose = nl > bl || (nl == bl && y + h - 1 < brow);
        }
        const unsigned long long b = __ballot(lose);
        if (b) return k0 + (int)__builtin_ctzll(b);
    }
    return ncand;
}

// Collision tests of a part's rotations against one bin (one rotation at a
// time, whole workgroup) and the choice among them: smallest new enclosure
// length (the density), then larger bottom row, then smaller column, the
// first rotation on ties (binClassNew.py:101-117).  The rotation rf (the
// type's minimum-height rotation, the likeliest winner) is tested first and
// the others follow in order; once a rotation has a placement, the next ones
// test only the candidates that could still win (winning_prefix; the dense
// mode tests them all), and equal placements go to the smaller rotation
// index, so the order changes no choice.  Returns the rotation, or -1, with
// its top-left (by, bx); uniform across the workgroup.
template <bool DENSE, bool F4, int WPC>
__device__ __forceinline__ int best_placement(const DevProblem &P, const DevWork &Wk, lds_rot *ri, int nr, int rf,
                                              int H, int W, int enc, const unsigned *grid,
                                              const short *s_cand, const int *s_ncand, int hmax, lds_u32 *band,
                                              lds_u32 *s_q, lds_u64 *s_zw, lds_u64 &c_tests,
                                              lds_u64 &c_algo, lds_u64 &c_groups,
                                              lds_u64 &c_mfma, lds_u64 *ph, int &by, int &bx) {
    int best = -1, bl = 0, brow = 0;
    for (int k = 0; k < nr; k++) {
        const int r = k == 0 ? rf : (k <= rf ? k - 1 : k);
        int ncand = uni(s_ncand[r]);
        if (ncand == 0) continue;
        const int h = uni(ri[r].h), w = uni(ri[r].w);
        const short *cand = s_cand + r * hmax;
        if (!DENSE && best >= 0) {
            ncand = uni(winning_prefix(cand, ncand, enc, H, h, bl, brow));
            if (ncand == 0) continue;
        }
        if (threadIdx.x == 0) {
            c_tests += 1;
            c_algo += 2ull * h * w * (unsigned long long)(H - h + 1) * (W - w + 1);
        }
        const int ncr = uni(ri[r].nc), nc4 = uni(ri[r].nc4);
        const long long qoff = ri[r].qoff;
        const int key = test_rotation<DENSE, F4, WPC>(grid, Wk.pw, W, h, w, F4 ? nc4 : ncr, F4 ? 2 * nc4 : ncr,
                                                 2 * nc4 + 2, P.qbits + qoff, cand, ncand, H - enc,
                                             band, Wk.band_words, s_q, s_zw, c_groups, c_mfma, ph,
                                             uni(ri[r].tiles), uni(ri[r].window), uni(ri[r].screen));
        if (key == KEY_NONE) continue;
        // density (new enclosure length) -> row -> col (binClassNew.py:101-117)
        const int y = uni(cand[key / XKEY]), x = key % XKEY, row = y + h - 1;
        const int nl = enc > H - y ? enc : H - y;
        if (best < 0 || nl < bl || (nl == bl && row > brow) ||
            (nl == bl && row == brow && (x < bx || (x == bx && r < best)))) {
            best = r; bl = nl; brow = row; bx = x; by = y;
        }
    }
    return best;
}

// Stamp a rotation (h x w mask, nc4 fp4 chunks, mask bits at qbits) at
// top-left (py, px) into a bin (binClassNew.py:155-160,
// cuda_batch_update.py:90-126), whole workgroup: rows [py, py+h) pass through
// the LDS band in chunks -- the grid's bit rows (a fresh bin's rows are zero)
// and the mask bits come in one batched copy, every grid word the part covers
// takes the 32 mask bits that fall on it, those words go back to HBM, and the
// rows' free runs go to vv (HBM) and to vcache when it is not null.
__device__ __forceinline__ void stamp_part(const DevWork &Wk, int h, int w, int nc4, const unsigned *qbits, int W,
                                           unsigned *grid, int *vv, short *vcache, bool fresh, int py, int px,
                                           lds_u32 *band, lds_u32 *s_q) {
    const int tid = threadIdx.x;
    const int pw = Wk.pw, q4 = pw >> 2;         // words and 16-byte vectors per row
    const DivBy byq4(q4);
    const int chunk = min(h, Wk.band_words / pw);
    const int nq = 2 * nc4 + 2;                 // mask bit words per row
    const int q0 = px >> 5, nw = ((px + w - 1) >> 5) - q0 + 1;   // grid words per row the part covers
    const DivBy bynw(nw);
    for (int a0 = 0; a0 < h; a0 += chunk) {
        const int rows = min(chunk, h - a0);
        const unsigned *g0 = grid + (long long)(py + a0) * pw;
        const unsigned *qg = qbits + (long long)a0 * nq;
        if (fresh)
            for (int k = tid; k < rows * q4; k += WG_THREADS) *(lds_v4i *)(band + (k << 2)) = (v4i){0, 0, 0, 0};
        const int ng = fresh ? 0 : rows * q4;
        wg_copy(ng + rows * nq,
                [&](int k) -> v4i {
                    if (k < ng) {
                        const int r = byq4(k), q = k - r * q4;
                        return *(const v4i *)(g0 + (long long)r * pw + (q << 2));
                    }
                    return (v4i){(int)qg[k - ng], 0, 0, 0};
                },
                [&](int k, v4i v) {
                    if (k < ng) *(lds_v4i *)(band + (k << 2)) = v;
                    else s_q[k - ng] = (unsigned)v[0];
                });
        __syncthreads();
        // mask row bits Q[i] = P[a][i - 32] (32 zero bits in front, zero past w):
        // grid word q holds cells 32q .. 32q+31, i.e. mask columns 32q - px + t,
        // Q bits from off = 32q - px + 32 >= 1 on
        for (int k = tid; k < rows * nw; k += WG_THREADS) {
            const int r = bynw(k), q = q0 + (k - r * nw);
            const int off = (q << 5) - px + 32, wi = off >> 5;
            const unsigned lo = s_q[r * nq + wi];
            const unsigned hi = wi + 1 < nq ? s_q[r * nq + wi + 1] : 0u;
            const unsigned bits = (unsigned)((((unsigned long long)hi << 32) | lo) >> (off & 31));
            if (bits) band[r * pw + q] |= bits;
        }
        __syncthreads();
        for (int k = tid; k < rows * nw; k += WG_THREADS) {
            const int r = bynw(k), q = q0 + (k - r * nw);
            grid[(long long)(py + a0 + r) * pw + q] = band[r * pw + q];
        }
        rows_free_run(band, pw, rows, W, [&](int r, int v) {
            vv[py + a0 + r] = v;
            if (vcache) vcache[py + a0 + r] = (short)v;
        });
        __syncthreads();
    }
}

// DENSE: every placement row, no early exit (the full-shape decode); F4: fp4
// MFMA operands (else int8); WPC: workgroups per CU the register budget is
// compiled for (2: 256 VGPRs, 3: 168).  Every instance gives the same results.
template <bool DENSE, bool F4, int WPC>
__global__ __launch_bounds__(WG_THREADS, WPC) void decode_kernel(DevProblem P, DevWork Wk, LaunchArgs L) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = P.n_parts;
    const int M = P.n_machines;
    const int hmax = Wk.hmax;
    // LDS layout (sizes computed on the host, brkga_create)
    double *s_sv = (double *)smem;                          // [n]
    int *s_mach = (int *)(s_sv + n);                        // [n]
    int *s_ty = s_mach + n;                                 // [n] part type at sequence position i
    int *s_den = s_ty + n;                                  // [<= MAX_ROT * hrot] one type's densities
    short *s_cand = (short *)(s_den + MAX_ROT * Wk.hrot);   // [MAX_ROT][hmax]
    short *s_vcache = s_cand + MAX_ROT * hmax;              // [vc_bins][hmax] vacancy of bins < vc_bins
    lds_u32 *s_q = (lds_u32 *)(smem + Wk.q_off);            // [hrot * (max nc + 1)]
    lds_u32 *band = (lds_u32 *)(smem + Wk.band_off);        // [band_words]
    // [hmax] the tested bin's free runs, 4 rows a word: they live in the band,
    // which the vacancy filter does not use (it is staged after the filter)
    lds_u64 *s_vq = (lds_u64 *)band;
    lds_u64 *s_zw = (lds_u64 *)(smem + Wk.zw_off);          // [2][4][zw_wave] zero masks, 4 screen masks
    __shared__ int s_ctx, s_cnt, s_off, s_infeasible, s_nb;
    __shared__ int s_ncand[MAX_ROT];
    __shared__ RotInfo s_ri_[MAX_ROT];   // the current part's rotations
    lds_rot *s_ri = (lds_rot *)s_ri_;

    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const int lane = tid & 63;
    const long long wg = blockIdx.x;
    unsigned *grids = Wk.grids + wg * (long long)Wk.max_bins * hmax * Wk.pw;
    int *vacs = Wk.vac + wg * (long long)Wk.max_bins * hmax;
    // bin bookkeeping of the running context (a context opens at most n bins)
    LDSP double *b_area = (LDSP double *)(smem + Wk.bins_off);
    LDSP double *b_pt = b_area + Wk.max_bins;
    LDSP double *b_pth = b_pt + Wk.max_bins;
    LDSP int *b_minrow = (LDSP int *)(b_pth + Wk.max_bins);

    // Statistics, in LDS so that they take no registers for the kernel's
    // life: the workgroup's counts (thread 0 updates them) and each wave's
    // MFMA instructions (the wave's lanes update its own entry together).
    __shared__ unsigned long long s_st[8 + 4];
    lds_u64 &c_ctx = ((lds_u64 *)s_st)[0], &c_place = ((lds_u64 *)s_st)[1], &c_bins = ((lds_u64 *)s_st)[2],
            &c_tests = ((lds_u64 *)s_st)[3], &c_groups = ((lds_u64 *)s_st)[4], &c_algo = ((lds_u64 *)s_st)[5],
            &c_over = ((lds_u64 *)s_st)[6], &c_mfma = ((lds_u64 *)s_st)[8 + wave];
    if (tid < 12) s_st[tid] = 0;
    // Time attribution, thread 0 only, in LDS so that it takes no registers
    // for the kernel's life: [0..7] phase totals, [8] longest context, [9]
    // context sum, [10] this context's start, [11] this phase's start.
    __shared__ unsigned long long s_tm[12];
    lds_u64 *ph = (lds_u64 *)s_tm;
    if (tid == 0)
        for (int k = 0; k < 12; k++) ph[k] = 0;
    const long long n_ctx = L.n_chrom * M;

    for (bool first = true;; first = false) {
        if (tid == 0) { ph[10] = tick(); ph[11] = ph[10]; }
        // The first G = gridDim.x positions (G <= n_ctx) go one to each
        // workgroup, later ones come from the queue.  In a full launch the
        // workgroups b, b + G/k, b + 2G/k, ... (k per CU) share a compute unit,
        // so layer r = b / (G/k) takes its positions in snake order (ascending
        // for even r, descending for odd r): the heaviest context of the first
        // round shares its CU with the lightest of the next layer, which
        // finishes early.
        if (tid == 0) {
            const int G = (int)gridDim.x, b = (int)blockIdx.x, per = G / max(1, L.wpc);
            int pos = b;
            if (per > 0 && b < per * L.wpc) {
                const int r = b / per, c = b - r * per;
                pos = r * per + ((r & 1) ? per - 1 - c : c);
            }
            s_ctx = first ? pos : G + (int)atomicAdd(Wk.queue, 1u);
        }
        __syncthreads();
        const long long q = uni(s_ctx);
        if (q >= n_ctx) break;
        const long long ctx = L.order ? uni(L.order[q]) : q;
        if (ctx < 0 || ctx >= n_ctx) continue;   // never: the order is a permutation of [0, n_ctx)
        const long long chrom = ctx / M;
        const int mach = (int)(ctx % M);
        const int H = P.plate_h[mach], W = P.plate_w[mach];
        const unsigned char *tiles = P.tiles + (long long)mach * TUNE_NC * TILE_ENTRY;   // the machine's tile table
        const double plate_area = (double)H * (double)W;
        if (tid == 0) { s_cnt = 0; s_off = 0; s_infeasible = 0; s_nb = 0; }
        // ---- sequence decode (placement.py:109-133) ----
        for (int j = tid; j < n; j += WG_THREADS) {
            const long long kb = chrom * 2LL * n;
            s_sv[j] = load_key(L.keys, L.key_f64, kb + j);
            s_mach[j] = machine_of(load_key(L.keys, L.key_f64, kb + n + j), P.thresholds, M);
        }
        __syncthreads();
        for (int j = tid; j < n; j += WG_THREADS) {
            const int mj = s_mach[j];
            if (mj < mach) atomicAdd(&s_off, 1);
            if (mj != mach) continue;
            atomicAdd(&s_cnt, 1);
            const double kj = s_sv[j];
            int rank = 0;
            for (int k = 0; k < n; k++)
                if (s_mach[k] == mach && (s_sv[k] < kj || (s_sv[k] == kj && k < j))) rank++;
            s_ty[rank] = P.slot_type[j];
            if (!P.fits[(long long)mach * P.n_types + P.slot_type[j]]) s_infeasible = 1;
        }
        __syncthreads();
        if (tid == 0) { const unsigned long long t = tick(); ph[0] += t - ph[11]; ph[11] = t; }
        const int cnt = uni(s_cnt);
        const int log_off = uni(s_off);
        int *logp = L.log ? L.log + (chrom * n + log_off) * 6 : nullptr;
        if (s_infeasible) {   // placement.py:47-53
            if (tid == 0) {
                L.ctx_ms[ctx] = -1.0;
                atomicAdd(&Wk.counters[ST_MCTX + mach], 1ull);
            }
            if (logp)
                for (int k = tid; k < cnt * 6; k += WG_THREADS) logp[k] = -1;
            __syncthreads();
            continue;
        }
        if (tid == 0) c_ctx += 1;
        bool overflow = false;
        for (int i = 0; i < cnt; i++) {
            const int ty = uni(s_ty[i]);
            const double parea = P.area[ty];
            const int nr = P.nrot[ty];
            const int rb = P.rot_base[ty];
            const int best_rot = P.best_rot[ty];
            const int nb = uni(s_nb);
            load_part(P, tiles, rb, nr, s_den, s_ri);   // s_ri and s_den are read by every wave
            bool placed = false;
            int pb = 0, py = 0, px = 0, prot = 0;
            bool fresh = false;
            for (int b = 0; b < nb && !placed; b++) {
                if (b_area[b] + parea > plate_area) continue;   // placement.py:57-59
                const unsigned *grid = grids + (long long)b * hmax * Wk.pw;
                if (tid == 0) ph[11] = tick();
                if (b >= Wk.vc_bins) {   // vacancy of bins beyond the LDS cache comes from HBM
                    const int *vac = vacs + (long long)b * hmax;
                    vacancy_quads(s_vq, H, [&](int y) { return vac[y]; });
                } else {
                    lds_cs16 *vr = (lds_cs16 *)(s_vcache + b * hmax);
                    vacancy_quads(s_vq, H, [&](int y) { return vr[y]; });
                }
                __syncthreads();
                // ---- vacancy filter, per candidate top row (wave r: rotation r) ----
                vacancy_filter<DENSE>(s_ri, nr, H, W, b_minrow[b], s_vq, s_den, s_cand, hmax, s_ncand);
                __syncthreads();
                if (tid == 0) ph[1] += tick() - ph[11];
                // ---- collision tests, one rotation at a time (whole workgroup) ----
                int bx = 0, by = 0;
                const int best = best_placement<DENSE, F4, WPC>(P, Wk, s_ri, nr, best_rot, H, W, H - b_minrow[b], grid, s_cand,
                                                       s_ncand, hmax, band, s_q, s_zw, c_tests, c_algo, c_groups,
                                                       c_mfma, ph, by, bx);
                if (best >= 0) {
                    placed = true;
                    pb = b; prot = best; py = by; px = bx;
                }
                __syncthreads();
            }
            if (tid == 0) ph[11] = tick();
            if (!placed) {
                if (nb >= Wk.max_bins) { overflow = true; break; }
                pb = nb; prot = best_rot;
                py = H - uni(s_ri[prot].h); px = 0;
                fresh = true;
                // fresh bin: zero grid, full vacancy (full_native_decoder.py:2087-2089)
                v4i *g4 = (v4i *)(grids + (long long)pb * hmax * Wk.pw);
                const int nvec = H * Wk.pw / 4;
                for (int k = tid; k < nvec; k += WG_THREADS) g4[k] = (v4i){0, 0, 0, 0};
                int *vv = vacs + (long long)pb * hmax;
                for (int k = tid; k < H; k += WG_THREADS) vv[k] = W;
                if (pb < Wk.vc_bins)
                    for (int k = tid; k < H; k += WG_THREADS) s_vcache[pb * hmax + k] = (short)W;
                if (tid == 0) {
                    b_area[pb] = 0.0; b_pt[pb] = 0.0; b_pth[pb] = 0.0; b_minrow[pb] = H;
                    s_nb = nb + 1;
                }
                if (tid == 0) c_bins += 1;
                __syncthreads();
                if (tid == 0) { const unsigned long long t = tick(); ph[6] += t - ph[11]; ph[11] = t; }
            }
            // ---- stamp the mask (binClassNew.py:155-160, cuda_batch_update.py:90-126) ----
            {
                const int h = uni(s_ri[prot].h);
                stamp_part(Wk, h, uni(s_ri[prot].w), uni(s_ri[prot].nc4), P.qbits + s_ri[prot].qoff, W,
                           grids + (long long)pb * hmax * Wk.pw, vacs + (long long)pb * hmax,
                           pb < Wk.vc_bins ? s_vcache + pb * hmax : nullptr, fresh, py, px, band, s_q);
                if (tid == 0) {   // binClassNew.py:147-153,172-174 ; placement.py:81-82
                    const long long mp = (long long)mach * P.n_types + ty;
                    b_area[pb] = b_area[pb] + parea;
                    b_pt[pb] = b_pt[pb] + P.proc_time[mp];
                    const double pth = P.pth[mp];
                    if (pth > b_pth[pb]) b_pth[pb] = pth;
                    if (py < b_minrow[pb]) b_minrow[pb] = py;
                    if (logp) {
                        int *e = logp + i * 6;
                        e[0] = mach; e[1] = P.type_id[ty]; e[2] = pb; e[3] = px;
                        e[4] = py + h - 1; e[5] = prot;
                    }
                }
                if (tid == 0) c_place += 1;
                __syncthreads();
                if (tid == 0) ph[5] += tick() - ph[11];
            }
        }
        if (tid == 0) {
            const int nbf = uni(s_nb);
            if (overflow) {
                L.ctx_ms[ctx] = -2.0;
                c_over += 1;
            } else {
                double total = 0.0;   // placement.py:86-89
                for (int b = 0; b < nbf; b++) total = total + ((b_pt[b] + b_pth[b]) + P.setup[mach]);
                L.ctx_ms[ctx] = total;
            }
            // per-machine bin statistics (full_native_decoder.py:1405-1417)
            atomicAdd(&Wk.counters[ST_MCTX + mach], 1ull);
            atomicAdd(&Wk.counters[ST_MBINS + mach], (unsigned long long)nbf);
            atomicMax(&Wk.counters[ST_MPEAK + mach], (unsigned long long)nbf);
            const unsigned long long t = tick(), tc = t - ph[10];
            ph[7] += t - ph[11];
            ph[9] += tc;
            ph[8] = tc > ph[8] ? tc : ph[8];
        }
        __syncthreads();
    }
    // ---- statistics ----
    // ops per MFMA instruction: 32 x 32 outputs x K (32 int8, 64 fp4) x 2
    if (lane == 0) atomicAdd(&Wk.counters[ST_MFMA], c_mfma * 32ull * 32ull * (unsigned long long)chunk_cols(F4) * 2ull);
    if (tid == 0) {
        atomicAdd(&Wk.counters[ST_TESTS], c_tests);
        atomicAdd(&Wk.counters[ST_GROUPS], c_groups);
        atomicAdd(&Wk.counters[ST_ALGO], c_algo);
        atomicAdd(&Wk.counters[ST_CTX], c_ctx);
        atomicAdd(&Wk.counters[ST_PLACE], c_place);
        atomicAdd(&Wk.counters[ST_BINS], c_bins);
        atomicAdd(&Wk.counters[ST_OVERFLOW], c_over);
        for (int k = 0; k < 8; k++) atomicAdd(&Wk.counters[ST_PH0 + k], (unsigned long long)ph[k]);
        atomicMax(&Wk.counters[ST_CTXMAX], (unsigned long long)ph[8]);
        atomicAdd(&Wk.counters[ST_CTXSUM], (unsigned long long)ph[9]);
    }
}

// fitness = max over machines, 1e16 if any machine infeasible (placement.py:147-156)
// ---------------------------------------------------------------------------
// Greedy initial solution (BRKGA_alg3.py:445-603): the host runs the greedy;
// these kernels test one part against every open batch at once and stamp the
// chosen placement.  Batches live in [machine][batch] slots of grids/vac.
// ---------------------------------------------------------------------------
struct GreedyArgs {
    unsigned *grids;        // [M][max_batches][hmax][pw] bit rows
    int *vac;               // [M][max_batches][hmax]
    int max_batches;
    int ty;                 // part type under test
    const int *jobs;        // [n_jobs][3]: machine, batch, enclosure length
    int *res;               // [n_jobs][4]: found, col, bottom row, rotation
};

// One workgroup per (machine, batch): binClassInitialSol.py:41-111 can_insert
// -- per rotation the vacancy filter and the bottom-left collision test, then
// the density -> row -> col choice (no area check; the greedy has none).
template <bool F4>
__global__ __launch_bounds__(WG_THREADS, 2) void greedy_test_kernel(DevProblem P, DevWork Wk, GreedyArgs A) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = P.n_parts, hmax = Wk.hmax, tid = threadIdx.x;
    int *s_den = (int *)((double *)smem + n) + 2 * n;       // same layout as decode_kernel
    short *s_cand = (short *)(s_den + MAX_ROT * Wk.hrot);
    lds_u32 *s_q = (lds_u32 *)(smem + Wk.q_off);
    lds_u32 *band = (lds_u32 *)(smem + Wk.band_off);
    lds_u64 *s_vq = (lds_u64 *)band;   // the free runs, in the band (staged after the filter)
    lds_u64 *s_zw = (lds_u64 *)(smem + Wk.zw_off);
    __shared__ int s_ncand[MAX_ROT];
    __shared__ RotInfo s_ri_[MAX_ROT];
    lds_rot *s_ri = (lds_rot *)s_ri_;
    const int *job = A.jobs + 3 * blockIdx.x;
    const int m = job[0], b = job[1], enc = job[2], ty = A.ty;
    const int H = P.plate_h[m], W = P.plate_w[m], nr = P.nrot[ty], rb = P.rot_base[ty];
    const long long slot = (long long)m * A.max_batches + b;
    const unsigned *grid = A.grids + slot * hmax * Wk.pw;
    const int *vac = A.vac + slot * hmax;
    load_part(P, P.tiles + (long long)m * TUNE_NC * TILE_ENTRY, rb, nr, s_den, s_ri);
    vacancy_quads(s_vq, H, [&](int y) { return vac[y]; });
    __syncthreads();
    vacancy_filter<false>(s_ri, nr, H, W, H - enc, s_vq, s_den, s_cand, hmax, s_ncand);
    __syncthreads();
    // statistics and phase times (unused here; the decode's)
    __shared__ unsigned long long s_st[4 + 4], s_tm[8];
    lds_u64 *st = (lds_u64 *)s_st, *ph = (lds_u64 *)s_tm;
    int bx = 0, by = 0;
    const int best = best_placement<false, F4, 2>(P, Wk, s_ri, nr, P.best_rot[ty], H, W, enc, grid, s_cand, s_ncand, hmax, band, s_q,
                                           s_zw, st[0], st[1], st[2], st[4 + (tid >> 6)], ph, by, bx);
    if (tid == 0) {
        int *r = A.res + 4 * blockIdx.x;
        r[0] = best >= 0;
        r[1] = bx;
        r[2] = best >= 0 ? by + s_ri[best].h - 1 : 0;
        r[3] = best;
    }
}

struct GreedyInsert {
    unsigned *grid;
    int *vac;
    int rg, H, W, py, px, fresh;
};

// Stamp the chosen placement (binClassInitialSol.py:122-145); a new batch is
// zeroed first (empty grid, full free runs).
__global__ __launch_bounds__(WG_THREADS, 2) void greedy_insert_kernel(DevProblem P, DevWork Wk, GreedyInsert I) {
    extern __shared__ __align__(16) unsigned char smem[];
    lds_u32 *s_q = (lds_u32 *)(smem + Wk.q_off);
    lds_u32 *band = (lds_u32 *)(smem + Wk.band_off);
    if (I.fresh) {
        v4i *g4 = (v4i *)I.grid;
        const int nvec = I.H * Wk.pw / 4;
        for (int k = threadIdx.x; k < nvec; k += WG_THREADS) g4[k] = (v4i){0, 0, 0, 0};
        for (int k = threadIdx.x; k < I.H; k += WG_THREADS) I.vac[k] = I.W;
        __syncthreads();
    }
    stamp_part(Wk, P.rot_h[I.rg], P.rot_w[I.rg], P.rot_nc4[I.rg], P.qbits + P.qbits_off[I.rg], I.W, I.grid, I.vac,
               nullptr, I.fresh != 0, I.py, I.px, band, s_q);
}

__global__ void finalize_kernel(const double *ctx_ms, long long n_chrom, int M, double *out,
                                unsigned long long *counters) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_chrom) return;
    double worst = 0.0;
    bool infeasible = false, overflow = false;
    for (int m = 0; m < M; m++) {
        const double v = ctx_ms[c * M + m];
        if (v == -2.0) overflow = true;
        else if (v < 0.0) infeasible = true;
        else if (v > worst) worst = v;
    }
    out[c] = overflow ? NAN : (infeasible ? BRKGA_INFEASIBLE : worst);
    if (c == 0) { atomicAdd(&counters[ST_CALLS], 1ull); atomicAdd(&counters[ST_CHROM], (unsigned long long)n_chrom); }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

static thread_local std::string g_create_error;

struct LayoutInputs {
    int n_parts = 0, hrot = 0, hmax = 0, force = 0;
    size_t qwords = 0, need_min = 0, need_4 = 0, lds_cu = 0;
};

// LDS layout of one occupancy (workgroups per CU) of the decode kernel.
struct LdsLayout {
    int wpc = 0, n_wg = 0;        // n_wg = 0: this occupancy does not fit
    int vc_bins = 0, q_off = 0, zw_off = 0, bins_off = 0, band_off = 0, band_words = 0;
    size_t smem = 0;
};

struct brkga_handle {
    int device = 0;
    int n_cu = 0;
    int n_wg = 0;          // workgroup slots of the workspace (the larger layout's grid)
    int last_wg = 0;       // grid of the last decode launch
    bool force3 = false;   // options.wgs_per_cu == 3
    LayoutInputs li;       // what the LDS layouts depend on besides the bin count
    LdsLayout lay[2];      // [0]: two workgroups per CU (or the forced count), [1]: three
    int dense = 0;
    int f4 = 1;   // fp4 MFMA operands (else int8)
    std::vector<unsigned char> tiles;   // host copy of DevProblem::tiles
    unsigned char *d_tiles = nullptr;
    brkga_problem prob{};     // scalars only
    DevProblem dp{};
    DevWork dw{};
    std::vector<void *> allocs;
    // host-call staging
    void *d_keys = nullptr; size_t d_keys_bytes = 0;
    double *d_out = nullptr; size_t d_out_n = 0;
    int *d_log = nullptr; size_t d_log_bytes = 0;
    double *d_ctx = nullptr; size_t d_ctx_n = 0;
    int *d_order = nullptr, *d_bucket = nullptr;   // [d_ctx_n] context order and cost buckets
    int *d_hist = nullptr;                          // [ORDER_BUCKETS]
    long long cost_all = 1;                         // summed context cost of all slots
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool ev_pending = false;
    double kernel_ms = 0.0, last_ms = 0.0;
    size_t smem = 0;
    int max_bins_req = 0;
    // Bumped whenever a device buffer that a launch reads or writes moves
    // (workspace growth, larger per-call buffers): a captured graph holds the
    // old addresses and must be captured again.
    unsigned ws_gen = 0;
    long long growths = 0;   // workspace growths after an overflow (since the last reset)
    double wall_hz = 1e8;
    void *scratch = nullptr;   // device address of `scratch` on this handle's GPU
    std::string err;
};

#define HIPCHK(h, call)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) {                                                                \
            (h)->err = std::string(#call) + ": " + hipGetErrorString(e_);                      \
            return e_ == hipErrorOutOfMemory ? BRKGA_ENOMEM : BRKGA_EDEVICE;                   \
        }                                                                                      \
    } while (0)

template <typename T>
static int upload(brkga_handle *h, const T *src, size_t n, const T **dst) {
    void *p = nullptr;
    size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    HIPCHK(h, hipMalloc(&p, bytes));
    h->allocs.push_back(p);
    if (n) HIPCHK(h, hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
    *dst = (const T *)p;
    return 0;
}

static int alloc_dev(brkga_handle *h, size_t bytes, void **dst) {
    void *p = nullptr;
    HIPCHK(h, hipMalloc(&p, std::max<size_t>(bytes, 16)));
    h->allocs.push_back(p);
    *dst = p;
    return 0;
}

static int fail(brkga_handle *h, int code, const std::string &msg) {
    if (h) h->err = msg; else g_create_error = msg;
    return code;
}

static int validate(const brkga_problem *p, std::string &msg) {
    char buf[256];
    if (!p) { msg = "problem is NULL"; return BRKGA_EINVAL; }
    if (p->n_parts < 1 || p->n_parts > 8192) { msg = "n_parts must be in [1, 8192]"; return BRKGA_EINVAL; }
    if (p->n_machines < 1 || p->n_machines > BRKGA_MAX_MACHINES) { msg = "n_machines must be in [1, 64]"; return BRKGA_EINVAL; }
    if (p->n_types < 1 || p->n_rots < 1) { msg = "n_types and n_rots must be >= 1"; return BRKGA_EINVAL; }
    if (!p->slot_type || !p->type_id || !p->area || !p->nrot || !p->best_rot || !p->rot_base ||
        !p->rot_h || !p->rot_w || !p->mask_off || !p->masks || !p->plate_h || !p->plate_w ||
        !p->setup || !p->proc_time || !p->proc_time_height || (p->n_machines > 1 && !p->thresholds)) {
        msg = "a problem array is NULL"; return BRKGA_EINVAL;
    }
    for (int m = 0; m + 1 < p->n_machines; m++)
        if (m > 0 && !(p->thresholds[m] >= p->thresholds[m - 1])) { msg = "thresholds must be ascending"; return BRKGA_EINVAL; }
    for (int j = 0; j < p->n_parts; j++)
        if (p->slot_type[j] < 0 || p->slot_type[j] >= p->n_types) {
            snprintf(buf, sizeof buf, "slot_type[%d] out of range", j); msg = buf; return BRKGA_EINVAL;
        }
    int64_t mask_bytes = 0;
    for (int r = 0; r < p->n_rots; r++) {
        if (p->rot_h[r] < 1 || p->rot_w[r] < 1 || p->rot_h[r] > 4096 || p->rot_w[r] > 4096) {
            snprintf(buf, sizeof buf, "rotation %d has bad shape %dx%d", r, p->rot_h[r], p->rot_w[r]); msg = buf; return BRKGA_EINVAL;
        }
        if (p->mask_off[r] < 0) { msg = "negative mask offset"; return BRKGA_EINVAL; }
        mask_bytes = std::max<int64_t>(mask_bytes, p->mask_off[r] + (int64_t)p->rot_h[r] * p->rot_w[r]);
    }
    for (int64_t k = 0; k < mask_bytes; k++)
        if (p->masks[k] > 1) { msg = "masks must be binary {0,1}"; return BRKGA_EINVAL; }
    for (int t = 0; t < p->n_types; t++) {
        if (p->nrot[t] < 1 || p->nrot[t] > MAX_ROT) { snprintf(buf, sizeof buf, "type %d: nrot must be 1..4", t); msg = buf; return BRKGA_EINVAL; }
        if (p->rot_base[t] < 0 || p->rot_base[t] + p->nrot[t] > p->n_rots) { msg = "rot_base out of range"; return BRKGA_EINVAL; }
        if (p->best_rot[t] < 0 || p->best_rot[t] >= p->nrot[t]) { msg = "best_rot out of range"; return BRKGA_EINVAL; }
        if (!(p->area[t] > 0.0) || !std::isfinite(p->area[t])) { snprintf(buf, sizeof buf, "type %d: area must be finite and > 0", t); msg = buf; return BRKGA_EINVAL; }
    }
    for (int m = 0; m < p->n_machines; m++) {
        const int H = p->plate_h[m], W = p->plate_w[m];
        if (H < 1 || W < 1 || H > 4096 || W > 4096) { msg = "plate sizes must be in [1, 4096]"; return BRKGA_EINVAL; }
        for (int t = 0; t < p->n_types; t++) {
            const int r0 = p->rot_base[t];
            const int h0 = p->rot_h[r0], w0 = p->rot_w[r0];
            const bool fits0 = !((h0 > H || w0 > W) && (w0 > H || h0 > W));
            const int rb = r0 + p->best_rot[t];
            if (fits0 && (p->rot_h[rb] > H || p->rot_w[rb] > W)) {
                snprintf(buf, sizeof buf,
                         "type %d fits machine %d only in a rotation other than its best_rot; "
                         "the placement procedure cannot open a bin for it", t, m);
                msg = buf; return BRKGA_EINVAL;
            }
        }
    }
    return 0;
}

static int device_workspace(brkga_handle *h, int max_bins) {
    DevWork &w = h->dw;
    const size_t nwg = h->n_wg;
    if (w.grids) {   // free the previous workspace (only grids/vac are resized)
        for (void *p : {(void *)w.grids, (void *)w.vac}) {
            auto it = std::find(h->allocs.begin(), h->allocs.end(), p);
            if (it != h->allocs.end()) { (void)hipFree(p); h->allocs.erase(it); }
        }
    }
    w.max_bins = max_bins;
    h->ws_gen++;
    void *p;
    int rc;
    if ((rc = alloc_dev(h, nwg * max_bins * (size_t)w.hmax * w.pw * sizeof(unsigned), &p))) return rc;
    w.grids = (unsigned *)p;
    if ((rc = alloc_dev(h, nwg * max_bins * (size_t)w.hmax * sizeof(int), &p))) return rc;
    w.vac = (int *)p;
    return 0;
}

static int scratch_address(brkga_handle *h, int device, void **out);

// The two LDS layouts of the decode kernels for a workspace of mb bins per
// context (the bin bookkeeping holds mb bins), and the workspace slot count.
// lay[0] (the WPC = 2 kernels, the full-shape decode and the greedy): two
// workgroups per CU when a 32-row group of every rotation fits a 4-tile band
// beside the fixed arrays, else one, with a vacancy cache of up to 16 bins.
// lay[1] (the WPC = 3 kernels of the pruned decode): three per CU with a
// vacancy cache of at most VC3_MAX bins (later bins read their free runs from
// HBM); absent when that does not fit.  options.wgs_per_cu 1 or 2 forces
// lay[0] alone, 3 forces lay[1] for every pruned decode.  Called at create
// and again when the workspace grows.
static int compute_layouts(brkga_handle *h, int mb) {
    const LayoutInputs &li = h->li;
    DevWork &w = h->dw;
    // fixed arrays for a vacancy cache of vc bins, offsets into w
    auto layout = [&](int vc, int wpc) -> size_t {
        w.vc_bins = vc;
        size_t f = (size_t)li.n_parts * (8 + 4 + 4) + (size_t)MAX_ROT * li.hrot * 4 +
                   (size_t)MAX_ROT * li.hmax * 2 + (size_t)vc * li.hmax * 2;
        f = (f + 15) & ~(size_t)15;
        w.q_off = (int)f;
        f += (li.qwords * 4 + 15) & ~(size_t)15;
        w.zw_off = (int)f;
        f += (size_t)(2 * 4 * zw_wave(wpc) + 4) * sizeof(unsigned long long);   // + the screen's tile masks
        w.bins_off = (int)f;
        f += ((size_t)mb * (3 * sizeof(double) + sizeof(int)) + 15) & ~(size_t)15;
        return f;
    };
    const int vc_max = std::min(mb, 16);
    auto make = [&](int wpc, int vcb, size_t need, LdsLayout &L) -> bool {
        const size_t fixed = layout(vcb, wpc);
        need = std::max(need, (size_t)li.hmax * 8);   // the band also holds the packed free runs
        const size_t per_wg = li.lds_cu / wpc;
        if (per_wg < fixed + need + LDS_RESERVE) return false;
        L.wpc = wpc;
        L.n_wg = std::max(1, h->n_cu * wpc);
        L.vc_bins = w.vc_bins; L.q_off = w.q_off; L.zw_off = w.zw_off; L.bins_off = w.bins_off;
        L.band_off = (int)fixed;
        L.band_words = (int)(((per_wg - fixed - LDS_RESERVE) & ~(size_t)15) / 4);
        L.smem = fixed + (size_t)L.band_words * 4;
        return true;
    };
    const int force = li.force;
    LdsLayout l0{}, l1{};
    const int w0 = force == 1 || force == 2 ? force : (make(2, vc_max, li.need_4, l0) ? 2 : 1);
    if (!make(w0, vc_max, li.need_min, l0)) {
        char b2[200];
        const size_t fixed = layout(vc_max, w0), per_wg = li.lds_cu / w0;
        snprintf(b2, sizeof b2, "a part rotation needs %zu bytes of LDS band, %zu available per workgroup",
                 li.need_min, per_wg > fixed + LDS_RESERVE ? per_wg - fixed - LDS_RESERVE : (size_t)0);
        h->err = b2;
        return BRKGA_EINVAL;
    }
    if (force == 0 || force == 3) {
        for (int v = std::min(vc_max, VC3_MAX); v >= std::min(vc_max, 2) && !l1.n_wg; v--) make(3, v, li.need_4, l1);
        if (force == 3 && !l1.n_wg)
            for (int v = vc_max; v >= std::min(vc_max, 2) && !l1.n_wg; v--) make(3, v, li.need_min, l1);
        if (force == 3 && !l1.n_wg) {
            h->err = "three workgroups per CU do not fit the LDS for this problem";
            return BRKGA_EINVAL;
        }
    }
    h->lay[0] = l0;
    h->lay[1] = l1;
    h->force3 = force == 3;
    w.vc_bins = l0.vc_bins; w.q_off = l0.q_off; w.zw_off = l0.zw_off; w.bins_off = l0.bins_off;
    w.band_off = l0.band_off; w.band_words = l0.band_words;
    h->smem = l0.smem;
    h->n_wg = std::max(l0.n_wg, l1.n_wg);   // workspace slots for the larger grid
    for (const void *kern : {(const void *)decode_kernel<false, false, 2>, (const void *)decode_kernel<true, false, 2>,
                             (const void *)decode_kernel<false, true, 2>, (const void *)decode_kernel<true, true, 2>,
                             (const void *)greedy_test_kernel<false>, (const void *)greedy_test_kernel<true>,
                             (const void *)greedy_insert_kernel})
        if (hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l0.smem) != hipSuccess) {
            h->err = "hipFuncSetAttribute(MaxDynamicSharedMemorySize) failed";
            return BRKGA_EDEVICE;
        }
    if (l1.n_wg)
        for (const void *kern : {(const void *)decode_kernel<false, false, 3>, (const void *)decode_kernel<false, true, 3>})
            if (hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l1.smem) != hipSuccess) {
                h->err = "hipFuncSetAttribute(MaxDynamicSharedMemorySize) failed";
                return BRKGA_EDEVICE;
            }
    return BRKGA_OK;
}

extern "C" int brkga_create(const brkga_problem *p, const brkga_options *o, brkga_handle **out) {
    if (!out) return fail(nullptr, BRKGA_EINVAL, "out is NULL");
    *out = nullptr;
    std::string msg;
    int rc = validate(p, msg);
    if (rc) return fail(nullptr, rc, msg);
    brkga_handle *h = new (std::nothrow) brkga_handle();
    if (!h) return fail(nullptr, BRKGA_ENOMEM, "host allocation failed");
    brkga_options opt{};
    if (o) opt = *o;
    h->device = opt.device;
    h->dense = opt.dense ? 1 : 0;
    if (opt.mfma_format != BRKGA_MFMA_FP4 && opt.mfma_format != BRKGA_MFMA_INT8) {
        delete h;
        return fail(nullptr, BRKGA_EINVAL, "mfma_format must be BRKGA_MFMA_FP4 (0) or BRKGA_MFMA_INT8 (1)");
    }
    h->f4 = opt.mfma_format == BRKGA_MFMA_FP4;
    if (opt.screen_rows < -1 || opt.screen_rows > BRKGA_MAX_SCREEN_ROWS) {
        delete h;
        return fail(nullptr, BRKGA_EINVAL, "screen_rows must be -1 (off), 0 (default) or 1..16");
    }
    if (!(opt.tiles_per_pass == 0 || opt.tiles_per_pass == 1 || opt.tiles_per_pass == 2 ||
          opt.tiles_per_pass == 4 || opt.tiles_per_pass == 8)) {
        delete h;
        return fail(nullptr, BRKGA_EINVAL, "tiles_per_pass must be 0, 1, 2, 4 or 8");
    }
    h->prob = *p;
    auto bail = [&](int code) {
        std::string e = h->err;
        brkga_destroy(h);
        return fail(nullptr, code, e);
    };
    if (hipSetDevice(h->device) != hipSuccess) { h->err = "hipSetDevice failed"; return bail(BRKGA_EDEVICE); }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) != hipSuccess) { h->err = "hipGetDeviceProperties failed"; return bail(BRKGA_EDEVICE); }
    h->n_cu = prop.multiProcessorCount;
    {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, h->device) == hipSuccess && khz > 0)
            h->wall_hz = khz * 1e3;
    }

    const int M = p->n_machines, T = p->n_types, R = p->n_rots;
    // ---- per-rotation tables: densities (BRKGA_alg3.py:351-360) and Toeplitz operands ----
    std::vector<int> dens, nc(R), nc4(R);
    std::vector<long long> dens_off(R), qbits_off(R);
    std::vector<unsigned> qbits;
    int hmax = 0, wmax = 0, pw = 4;   // pw: 32-bit words per grid bit row
    for (int m = 0; m < M; m++) { hmax = std::max(hmax, p->plate_h[m]); wmax = std::max(wmax, p->plate_w[m]); }
    for (int r = 0; r < R; r++) {
        const int rh = p->rot_h[r], rw = p->rot_w[r];
        const uint8_t *mk = p->masks + p->mask_off[r];
        dens_off[r] = (long long)dens.size();
        {   // (density << 16) | 8 row (the row's byte offset in the packed free runs), densest rows
            // first: the filter's early exit comes sooner
            std::vector<int> dr(rh);
            for (int a = 0; a < rh; a++) {
                int best = 0, cur = 0;
                for (int b = 0; b < rw; b++) { cur = mk[a * rw + b] ? cur + 1 : 0; best = std::max(best, cur); }
                dr[a] = (best << 16) | (a << 3);   // best <= rw <= 4096 and 8a < 8 rh <= 32768 (16 bits)
            }
            std::stable_sort(dr.begin(), dr.end(), [](int x, int y) { return (x >> 16) > (y >> 16); });
            dens.insert(dens.end(), dr.begin(), dr.end());
        }
        nc[r] = (rw + 31 + 31) / 32;    // 32-column chunks a 32-placement tile spans (int8)
        nc4[r] = (rw + 31 + 63) / 64;   // 64-column chunks (fp4)
        qbits_off[r] = (long long)qbits.size();
        for (int a = 0; a < rh; a++) {   // 2 nc4 + 2 >= nc + 1 words per row, zero past the mask
            const size_t row0 = qbits.size();
            qbits.resize(row0 + 2 * nc4[r] + 2, 0u);
            for (int b = 0; b < rw; b++)
                if (mk[a * rw + b]) qbits[row0 + (b + 32) / 32] |= 1u << ((b + 32) % 32);
        }
        for (int m = 0; m < M; m++) {
            const int H = p->plate_h[m], W = p->plate_w[m];
            if (rh > H || rw > W) continue;
            const int Tt = (W - rw + 1 + 31) / 32;
            pw = std::max(pw, Tt + 2 * nc4[r] - 1);   // the widest span a staged window reads (2 nc4 >= nc)
        }
    }
    pw = std::max(pw, (wmax + 31) / 32);
    pw = (pw + 3) & ~3;   // whole 16-byte vectors; the words past W stay zero
    std::vector<unsigned char> fits((size_t)M * T);
    for (int m = 0; m < M; m++)
        for (int t = 0; t < T; t++) {
            const int r0 = p->rot_base[t];
            const int h0 = p->rot_h[r0], w0 = p->rot_w[r0];
            fits[(size_t)m * T + t] = !((h0 > p->plate_h[m] || w0 > p->plate_w[m]) &&
                                        (w0 > p->plate_h[m] || h0 > p->plate_w[m]));
            // a new bin takes the best rotation at (0, H-1) (placement.py:66-83);
            // the reference fails when that rotation is larger than the plate
            const int rb = r0 + p->best_rot[t];
            if (fits[(size_t)m * T + t] && (p->rot_h[rb] > p->plate_h[m] || p->rot_w[rb] > p->plate_w[m])) {
                char b2[160];
                snprintf(b2, sizeof b2, "part type %d fits machine %d only in a rotation other than its best "
                         "rotation, which a new bin would not hold", t, m);
                h->err = b2;
                return bail(BRKGA_EINVAL);
            }
        }
    int64_t mask_bytes = 0;
    for (int r = 0; r < R; r++) mask_bytes = std::max<int64_t>(mask_bytes, p->mask_off[r] + (int64_t)p->rot_h[r] * p->rot_w[r]);

    DevProblem &d = h->dp;
    d.n_parts = p->n_parts; d.n_machines = M; d.n_types = T;
    std::vector<double> thr(std::max(M - 1, 0));
    for (int m = 0; m + 1 < M; m++) thr[m] = p->thresholds[m];
    if ((rc = upload(h, thr.data(), thr.size(), &d.thresholds)) ||
        (rc = upload(h, p->slot_type, (size_t)p->n_parts, &d.slot_type)) ||
        (rc = upload(h, p->type_id, (size_t)T, &d.type_id)) ||
        (rc = upload(h, p->area, (size_t)T, &d.area)) ||
        (rc = upload(h, p->nrot, (size_t)T, &d.nrot)) ||
        (rc = upload(h, p->best_rot, (size_t)T, &d.best_rot)) ||
        (rc = upload(h, p->rot_base, (size_t)T, &d.rot_base)) ||
        (rc = upload(h, p->rot_h, (size_t)R, &d.rot_h)) ||
        (rc = upload(h, p->rot_w, (size_t)R, &d.rot_w)) ||
        (rc = upload(h, nc.data(), (size_t)R, &d.rot_nc)) ||
        (rc = upload(h, nc4.data(), (size_t)R, &d.rot_nc4)) ||
        (rc = upload(h, (const long long *)p->mask_off, (size_t)R, &d.mask_off)) ||
        (rc = upload(h, dens_off.data(), (size_t)R, &d.dens_off)) ||
        (rc = upload(h, qbits_off.data(), (size_t)R, &d.qbits_off)) ||
        (rc = upload(h, p->masks, (size_t)mask_bytes, &d.masks)) ||
        (rc = upload(h, dens.data(), dens.size(), &d.dens)) ||
        (rc = upload(h, qbits.data(), qbits.size(), &d.qbits)) ||
        (rc = upload(h, p->plate_h, (size_t)M, &d.plate_h)) ||
        (rc = upload(h, p->plate_w, (size_t)M, &d.plate_w)) ||
        (rc = upload(h, p->setup, (size_t)M, &d.setup)) ||
        (rc = upload(h, p->proc_time, (size_t)M * T, &d.proc_time)) ||
        (rc = upload(h, p->proc_time_height, (size_t)M * T, &d.pth)) ||
        (rc = upload(h, fits.data(), fits.size(), &d.fits)))
        return bail(rc);
    // tile table: per-width default pass size (or the option's cap), full
    // windows, the option's screening rows
    const int screen = opt.screen_rows < 0 ? 0 : (opt.screen_rows ? opt.screen_rows : SCREEN_DEFAULT);
    h->tiles.assign((size_t)M * TUNE_NC * TILE_ENTRY, 0);
    for (int m = 0; m < M; m++)
        for (int c = 0; c < TUNE_NC; c++) {
            const int def = pass_tiles(c < 1 ? 1 : c);
            unsigned char *e = &h->tiles[((size_t)m * TUNE_NC + c) * TILE_ENTRY];
            e[0] = (unsigned char)(opt.tiles_per_pass ? std::min(def, opt.tiles_per_pass) : def);
            e[1] = (unsigned char)WINDOW_CAP;
            e[2] = (unsigned char)screen;
        }
    if ((rc = upload(h, h->tiles.data(), h->tiles.size(), &d.tiles))) return bail(rc);
    // context cost for the decode order: bounding-box area of rotation 0
    std::vector<int> type_cost(T);
    for (int t = 0; t < T; t++) type_cost[t] = p->rot_h[p->rot_base[t]] * p->rot_w[p->rot_base[t]];
    h->cost_all = 0;
    for (int j = 0; j < p->n_parts; j++) h->cost_all += type_cost[p->slot_type[j]];
    h->cost_all = std::max(h->cost_all, 1LL);
    if ((rc = upload(h, type_cost.data(), type_cost.size(), &d.type_cost))) return bail(rc);
    h->d_tiles = (unsigned char *)d.tiles;
    h->prob.thresholds = nullptr;   // host pointers are not kept
    DevWork &w = h->dw;
    w.hmax = hmax;
    w.pw = pw;
    void *q;
    if ((rc = alloc_dev(h, 64, &q))) return bail(rc);
    w.queue = (unsigned int *)q;
    if ((rc = alloc_dev(h, ST_COUNT * sizeof(unsigned long long), &q))) return bail(rc);
    w.counters = (unsigned long long *)q;
    if (hipMemset(w.counters, 0, ST_COUNT * sizeof(unsigned long long)) != hipSuccess) { h->err = "hipMemset failed"; return bail(BRKGA_EDEVICE); }
    int mb = opt.max_bins > 0 ? opt.max_bins : std::min(p->n_parts, 32);
    h->max_bins_req = mb;
    // ---- LDS layout: keys/sequence, vacancy, densities, candidates, then the band ----
    int hrot = 1;
    for (int r = 0; r < R; r++) hrot = std::max(hrot, p->rot_h[r]);
    w.hrot = hrot;
    size_t qwords = 1;
    for (int r = 0; r < R; r++) qwords = std::max(qwords, (size_t)p->rot_h[r] * (2 * nc4[r] + 2));
    // band sizes in bytes: the smallest that works (one row group of one tile,
    // and one grid row for the stamp) and one holding a 32-row group of every
    // rotation with at least a 4-tile window (windows shrink at run time when
    // a row is wider)
    size_t need_min = ((size_t)pw + BAND_SLACK) * 4, need_4 = need_min;
    for (int m = 0; m < M; m++)
        for (int r = 0; r < R; r++) {
            const int H = p->plate_h[m], W = p->plate_w[m], rh = p->rot_h[r], rw = p->rot_w[r];
            if (rh > H || rw > W) continue;
            const int Tt = (W - rw + 1 + 31) / 32;
            need_min = std::max(need_min, ((size_t)rh * band_stride(1, 2 * nc4[r]) + BAND_SLACK) * 4);
            need_4 = std::max(need_4, ((size_t)(rh + GROUP_ROWS - 1) * band_stride(std::min(Tt, 4), 2 * nc4[r]) +
                                       BAND_SLACK) * 4);
        }
    if (opt.wgs_per_cu < 0 || opt.wgs_per_cu > 3) {
        h->err = "wgs_per_cu must be 0 (automatic), 1, 2 or 3";
        return bail(BRKGA_EINVAL);
    }
    LayoutInputs &li = h->li;
    li.n_parts = p->n_parts; li.hrot = hrot; li.hmax = hmax; li.force = opt.wgs_per_cu;
    li.qwords = qwords; li.need_min = need_min; li.need_4 = need_4;
    li.lds_cu = prop.maxSharedMemoryPerMultiProcessor > 0 ? (size_t)prop.maxSharedMemoryPerMultiProcessor : 65536;
    if ((rc = compute_layouts(h, mb))) return bail(rc);
    if ((rc = device_workspace(h, mb))) return bail(rc);
    if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) { h->err = "hipEventCreate failed"; return bail(BRKGA_EDEVICE); }
    if ((rc = scratch_address(h, h->device, &h->scratch))) return bail(rc);
    *out = h;
    return BRKGA_OK;
}

static void settle_events(brkga_handle *h) {
    if (!h->ev_pending) return;
    if (hipEventSynchronize(h->ev1) == hipSuccess) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, h->ev0, h->ev1) == hipSuccess) {
            h->last_ms = ms;
            h->kernel_ms += ms;
        }
    }
    h->ev_pending = false;
}

int brkga_internal_reserve(brkga_handle *h, int64_t n_chrom) {
    const int M = h->prob.n_machines;
    if ((size_t)(n_chrom * M) > h->d_ctx_n) {
        if (h->d_ctx) { (void)hipFree(h->d_ctx); h->d_ctx = nullptr; }
        if (h->d_order) { (void)hipFree(h->d_order); h->d_order = nullptr; }
        if (h->d_bucket) { (void)hipFree(h->d_bucket); h->d_bucket = nullptr; }
        h->d_ctx_n = 0;
        HIPCHK(h, hipMalloc(&h->d_ctx, (size_t)n_chrom * M * sizeof(double)));
        HIPCHK(h, hipMalloc(&h->d_order, (size_t)n_chrom * M * sizeof(int)));
        HIPCHK(h, hipMalloc(&h->d_bucket, (size_t)n_chrom * M * sizeof(int)));
        h->d_ctx_n = (size_t)n_chrom * M;
        h->ws_gen++;
    }
    if (!h->d_hist) {
        HIPCHK(h, hipMalloc(&h->d_hist, ORDER_BUCKETS * sizeof(int)));
        HIPCHK(h, hipMemset(h->d_hist, 0, ORDER_BUCKETS * sizeof(int)));   // cleared by each launch after use
    }
    return 0;
}

// Per-call device buffers of one decode launch: the contexts' makespans, their
// decode order and cost buckets, the order histogram (zero before the launch,
// cleared again by order_scatter_kernel) and the work-queue counter.
struct CallBuffers {
    double *ctx_ms;
    int *order, *bucket, *hist;
    unsigned int *queue;
};

// Enqueue the decode of n_chrom device-resident chromosomes on s with the
// given per-call buffers (sized for n_chrom * n_machines contexts); allocates
// nothing.
static int launch_decode(brkga_handle *h, const void *d_keys, int key_dtype, int64_t n_chrom, double *d_out,
                         int32_t *d_log, hipStream_t s, bool timed, const CallBuffers &cb) {
    const int M = h->prob.n_machines;
    if (timed) settle_events(h);
    LaunchArgs L;
    L.keys = d_keys; L.key_f64 = key_dtype == BRKGA_KEYS_F64; L.n_chrom = n_chrom;
    L.ctx_ms = cb.ctx_ms; L.log = d_log; L.order = cb.order;
    DevWork wk = h->dw;
    wk.queue = cb.queue;
    const long long n_ctx = n_chrom * M;
    // three workgroups per CU for a pruned decode of at least two rounds of
    // contexts at that occupancy (a single round finishes sooner at two, where
    // each context runs faster); the layout fields follow the choice
    const LdsLayout &L3 = h->lay[1];
    const bool three = !h->dense && L3.n_wg > 0 && (h->force3 || n_ctx >= 2LL * L3.n_wg);
    const LdsLayout &LL = three ? L3 : h->lay[0];
    wk.vc_bins = LL.vc_bins; wk.q_off = LL.q_off; wk.zw_off = LL.zw_off; wk.bins_off = LL.bins_off;
    wk.band_off = LL.band_off; wk.band_words = LL.band_words;
    const int grid = (int)std::min<long long>(LL.n_wg, std::max<long long>(n_ctx, 1));
    h->last_wg = grid;
    L.wpc = LL.wpc;
    hipLaunchKernelGGL(order_cost_kernel, dim3((unsigned)((n_ctx + 3) / 4)), dim3(256), 0, s, h->dp, d_keys,
                       L.key_f64, (long long)n_chrom, h->cost_all, cb.bucket, cb.hist);
    HIPCHK(h, hipGetLastError());
    hipLaunchKernelGGL(order_scatter_kernel, dim3(1), dim3(ORDER_BUCKETS), 0, s, cb.bucket, cb.hist, n_ctx,
                       cb.order, cb.queue);
    HIPCHK(h, hipGetLastError());
    if (timed) HIPCHK(h, hipEventRecord(h->ev0, s));
    auto kern = three ? (h->f4 ? decode_kernel<false, true, 3> : decode_kernel<false, false, 3>)
                      : h->f4 ? (h->dense ? decode_kernel<true, true, 2> : decode_kernel<false, true, 2>)
                              : (h->dense ? decode_kernel<true, false, 2> : decode_kernel<false, false, 2>);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(WG_THREADS), LL.smem, s, h->dp, wk, L);
    HIPCHK(h, hipGetLastError());
    if (timed) {
        HIPCHK(h, hipEventRecord(h->ev1, s));
        h->ev_pending = true;
    }
    const int fb = 256;
    hipLaunchKernelGGL(finalize_kernel, dim3((unsigned)((n_chrom + fb - 1) / fb)), dim3(fb), 0, s,
                       cb.ctx_ms, (long long)n_chrom, M, d_out, h->dw.counters);
    HIPCHK(h, hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// Allocation-free decode.  The per-call buffers of brkga_evaluate_device_noalloc
// live in `scratch`, one array per GPU for the life of the process, laid out
// as: order histogram [ORDER_BUCKETS] int32, queue counter (256 bytes), then
// for up to BRKGA_NOALLOC_MAX_CONTEXTS contexts their makespans (f64), decode
// order and cost buckets (int32).  In host code `scratch` names a host-side
// stand-in: its device address on the current GPU comes from
// hipGetSymbolAddress, resolved once per device by brkga_create.
// ---------------------------------------------------------------------------
#define NOALLOC_HIST_OFF 0
#define NOALLOC_QUEUE_OFF (ORDER_BUCKETS * 4)
#define NOALLOC_CTX_OFF (NOALLOC_QUEUE_OFF + 256)
#define NOALLOC_ORDER_OFF (NOALLOC_CTX_OFF + 8 * BRKGA_NOALLOC_MAX_CONTEXTS)
#define NOALLOC_BUCKET_OFF (NOALLOC_ORDER_OFF + 4 * BRKGA_NOALLOC_MAX_CONTEXTS)
#define NOALLOC_BYTES (NOALLOC_BUCKET_OFF + 4 * BRKGA_NOALLOC_MAX_CONTEXTS)
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

#define NOALLOC_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES];   // device address of `scratch` per GPU

// Device address of `scratch` on `device`, which must be current; resolved on
// the first call for the device and kept.
static int scratch_address(brkga_handle *h, int device, void **out) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) {
        h->err = "device ordinal out of range for the allocation-free decode's scratch table";
        return BRKGA_EINVAL;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            h->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return BRKGA_EDEVICE;
        }
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return 0;
}

// Zero the histogram and the queue counter of the scratch buffers (a kernel,
// not a memset node, so that a caller may capture the call into a graph).
__global__ __launch_bounds__(ORDER_BUCKETS) void scratch_clear_kernel(int *hist, unsigned int *queue) {
    hist[threadIdx.x] = 0;
    if (threadIdx.x == 0) *queue = 0u;
}

int brkga_internal_launch(brkga_handle *h, const void *d_keys, int key_dtype, int64_t n_chrom,
                          double *d_out, int32_t *d_log, hipStream_t s, bool timed) {
    int rc = brkga_internal_reserve(h, n_chrom);
    if (rc) return rc;
    const CallBuffers cb{h->d_ctx, h->d_order, h->d_bucket, h->d_hist, h->dw.queue};
    return launch_decode(h, d_keys, key_dtype, n_chrom, d_out, d_log, s, timed, cb);
}

int brkga_internal_n_parts(const brkga_handle *h) { return h->prob.n_parts; }
unsigned brkga_internal_ws_gen(const brkga_handle *h) { return h->ws_gen; }
int brkga_internal_overflows(brkga_handle *h, unsigned long long *out) {
    HIPCHK(h, hipMemcpy(out, h->dw.counters + ST_OVERFLOW, sizeof *out, hipMemcpyDeviceToHost));
    return 0;
}
int brkga_internal_device(const brkga_handle *h) { return h->device; }
void brkga_internal_error(brkga_handle *h, const char *msg) { h->err = msg; }

static int check_args(brkga_handle *h, const void *keys, int key_dtype, int64_t n_chrom, int64_t n_genes,
                      const void *out) {
    if (!h) return BRKGA_EINVAL;
    if (!keys || !out) return fail(h, BRKGA_EINVAL, "keys and out must not be NULL");
    if (key_dtype != BRKGA_KEYS_F32 && key_dtype != BRKGA_KEYS_F64) return fail(h, BRKGA_EINVAL, "key_dtype must be BRKGA_KEYS_F32 or BRKGA_KEYS_F64");
    if (n_chrom < 1) return fail(h, BRKGA_EINVAL, "n_chrom must be >= 1");
    if (n_genes != 2LL * h->prob.n_parts) return fail(h, BRKGA_EINVAL, "n_genes must equal 2*n_parts");
    if (n_chrom > (1LL << 31) / std::max(1, h->prob.n_machines)) return fail(h, BRKGA_EINVAL, "n_chrom too large");
    return 0;
}

static int read_overflow(brkga_handle *h, hipStream_t s, unsigned long long *ov) {
    HIPCHK(h, hipMemcpyAsync(ov, h->dw.counters + ST_OVERFLOW, sizeof *ov, hipMemcpyDeviceToHost, s));
    HIPCHK(h, hipStreamSynchronize(s));
    return 0;
}

extern "C" int brkga_evaluate_device(brkga_handle *h, const void *d_keys, int key_dtype, int64_t n_chrom,
                                     int64_t n_genes, double *d_out, int32_t *d_log, void *stream,
                                     int check_overflow) {
    int rc = check_args(h, d_keys, key_dtype, n_chrom, n_genes, d_out);
    if (rc) return rc;
    HIPCHK(h, hipSetDevice(h->device));
    hipStream_t s = (hipStream_t)stream;
    unsigned long long ov0 = 0, ov1 = 0;
    if (check_overflow && (rc = read_overflow(h, s, &ov0))) return rc;
    if ((rc = brkga_internal_launch(h, d_keys, key_dtype, n_chrom, d_out, d_log, s, true))) return rc;
    if (check_overflow) {
        if ((rc = read_overflow(h, s, &ov1))) return rc;
        settle_events(h);
        if (ov1 != ov0) return fail(h, BRKGA_EOVERFLOW, "a context needed more bins than max_bins");
    }
    return 0;
}

extern "C" int brkga_evaluate_device_noalloc(brkga_handle *h, const void *d_keys, int key_dtype, int64_t n_chrom,
                                             int64_t n_genes, double *d_out, int32_t *d_log, void *stream,
                                             int check_overflow) {
    int rc = check_args(h, d_keys, key_dtype, n_chrom, n_genes, d_out);
    if (rc) return rc;
    if (n_chrom * h->prob.n_machines > BRKGA_NOALLOC_MAX_CONTEXTS) {
        char b[160];
        snprintf(b, sizeof b, "n_chrom * n_machines = %lld exceeds the allocation-free decode's limit of %d contexts",
                 (long long)(n_chrom * h->prob.n_machines), BRKGA_NOALLOC_MAX_CONTEXTS);
        return fail(h, BRKGA_EINVAL, b);
    }
    if (!h->scratch) return fail(h, BRKGA_ESTATE, "scratch buffers not resolved");
    HIPCHK(h, hipSetDevice(h->device));
    hipStream_t s = (hipStream_t)stream;
    unsigned long long ov0 = 0, ov1 = 0;
    if (check_overflow && (rc = read_overflow(h, s, &ov0))) return rc;
    unsigned char *base = (unsigned char *)h->scratch;
    const CallBuffers cb{(double *)(base + NOALLOC_CTX_OFF), (int *)(base + NOALLOC_ORDER_OFF),
                         (int *)(base + NOALLOC_BUCKET_OFF), (int *)(base + NOALLOC_HIST_OFF),
                         (unsigned int *)(base + NOALLOC_QUEUE_OFF)};
    hipLaunchKernelGGL(scratch_clear_kernel, dim3(1), dim3(ORDER_BUCKETS), 0, s, cb.hist, cb.queue);
    HIPCHK(h, hipGetLastError());
    if ((rc = launch_decode(h, d_keys, key_dtype, n_chrom, d_out, d_log, s, false, cb))) return rc;
    if (check_overflow) {
        if ((rc = read_overflow(h, s, &ov1))) return rc;
        if (ov1 != ov0) return fail(h, BRKGA_EOVERFLOW, "a context needed more bins than max_bins");
    }
    return 0;
}

extern "C" int brkga_evaluate(brkga_handle *h, const void *keys, int key_dtype, int64_t n_chrom,
                              int64_t n_genes, double *out, int32_t *log, void *stream) {
    int rc = check_args(h, keys, key_dtype, n_chrom, n_genes, out);
    if (rc) return rc;
    HIPCHK(h, hipSetDevice(h->device));
    hipStream_t s = (hipStream_t)stream;
    const size_t kb = (size_t)n_chrom * n_genes * (key_dtype == BRKGA_KEYS_F64 ? 8 : 4);
    if (kb > h->d_keys_bytes) {
        if (h->d_keys) { (void)hipFree(h->d_keys); h->d_keys = nullptr; h->d_keys_bytes = 0; }
        HIPCHK(h, hipMalloc(&h->d_keys, kb));
        h->d_keys_bytes = kb;
    }
    if ((size_t)n_chrom > h->d_out_n) {
        if (h->d_out) { (void)hipFree(h->d_out); h->d_out = nullptr; h->d_out_n = 0; }
        HIPCHK(h, hipMalloc(&h->d_out, (size_t)n_chrom * sizeof(double)));
        h->d_out_n = (size_t)n_chrom;
    }
    const size_t lb = log ? (size_t)n_chrom * h->prob.n_parts * 6 * sizeof(int32_t) : 0;
    if (lb > h->d_log_bytes) {
        if (h->d_log) { (void)hipFree(h->d_log); h->d_log = nullptr; h->d_log_bytes = 0; }
        HIPCHK(h, hipMalloc(&h->d_log, lb));
        h->d_log_bytes = lb;
    }
    HIPCHK(h, hipMemcpyAsync(h->d_keys, keys, kb, hipMemcpyHostToDevice, s));
    for (int attempt = 0; attempt < 8; attempt++) {
        rc = brkga_evaluate_device(h, h->d_keys, key_dtype, n_chrom, n_genes, h->d_out, log ? h->d_log : nullptr, s, 1);
        if (rc != BRKGA_EOVERFLOW) break;
        // grow the bin workspace and decode again (full_native_decoder.py:2102-2153 grows likewise)
        const int nb = std::min(h->prob.n_parts, h->dw.max_bins * 2);
        if (nb == h->dw.max_bins) break;
        HIPCHK(h, hipStreamSynchronize(s));
        if ((rc = compute_layouts(h, nb))) return rc;   // the bin bookkeeping in LDS grows too
        if ((rc = device_workspace(h, nb))) return rc;
        h->growths++;
    }
    if (rc) return rc;
    HIPCHK(h, hipMemcpyAsync(out, h->d_out, (size_t)n_chrom * sizeof(double), hipMemcpyDeviceToHost, s));
    if (log) HIPCHK(h, hipMemcpyAsync(log, h->d_log, lb, hipMemcpyDeviceToHost, s));
    HIPCHK(h, hipStreamSynchronize(s));
    return 0;
}

extern "C" int brkga_set_tiles(brkga_handle *h, const int32_t *tiles_per_pass, const int32_t *window_tiles) {
    if (!h) return fail(nullptr, BRKGA_EINVAL, "handle is NULL");
    const int M = h->prob.n_machines;
    std::vector<unsigned char> t = h->tiles;
    for (int m = 0; m < M; m++)
        for (int c = 0; c < TUNE_NC; c++) {
            const size_t i = (size_t)m * TUNE_NC + c;
            if (tiles_per_pass) {
                const int v = tiles_per_pass[i];
                if (!(v == 0 || v == 1 || v == 2 || v == 4 || v == 8))
                    return fail(h, BRKGA_EINVAL, "tiles_per_pass entries must be 0, 1, 2, 4 or 8");
                t[i * TILE_ENTRY] = (unsigned char)(v ? std::min(v, pass_tiles(c < 1 ? 1 : c)) : pass_tiles(c < 1 ? 1 : c));
            }
            if (window_tiles) {
                const int v = window_tiles[i];
                if (v < 0 || v > WINDOW_CAP) return fail(h, BRKGA_EINVAL, "window_tiles entries must be in [0, 64]");
                t[i * TILE_ENTRY + 1] = (unsigned char)(v ? v : WINDOW_CAP);
            }
        }
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipDeviceSynchronize());   // no launch may be reading the table
    HIPCHK(h, hipMemcpy(h->d_tiles, t.data(), t.size(), hipMemcpyHostToDevice));
    h->tiles = t;
    return BRKGA_OK;
}

extern "C" int brkga_get_tiles(const brkga_handle *h, int32_t *tiles_per_pass, int32_t *window_tiles) {
    if (!h) return BRKGA_EINVAL;
    const size_t n = (size_t)h->prob.n_machines * TUNE_NC;
    for (size_t i = 0; i < n; i++) {
        if (tiles_per_pass) tiles_per_pass[i] = h->tiles[i * TILE_ENTRY];
        if (window_tiles) window_tiles[i] = h->tiles[i * TILE_ENTRY + 1];
    }
    return BRKGA_OK;
}

extern "C" int brkga_set_screen(brkga_handle *h, const int32_t *screen_rows) {
    if (!h) return fail(nullptr, BRKGA_EINVAL, "handle is NULL");
    if (!screen_rows) return fail(h, BRKGA_EINVAL, "screen_rows is NULL");
    const size_t n = (size_t)h->prob.n_machines * TUNE_NC;
    std::vector<unsigned char> t = h->tiles;
    for (size_t i = 0; i < n; i++) {
        const int v = screen_rows[i];
        if (v < 0 || v > BRKGA_MAX_SCREEN_ROWS) return fail(h, BRKGA_EINVAL, "screen_rows entries must be in [0, 16]");
        t[i * TILE_ENTRY + 2] = (unsigned char)v;
    }
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipDeviceSynchronize());   // no launch may be reading the table
    HIPCHK(h, hipMemcpy(h->d_tiles, t.data(), t.size(), hipMemcpyHostToDevice));
    h->tiles = t;
    return BRKGA_OK;
}

extern "C" int brkga_get_screen(const brkga_handle *h, int32_t *screen_rows) {
    if (!h || !screen_rows) return BRKGA_EINVAL;
    const size_t n = (size_t)h->prob.n_machines * TUNE_NC;
    for (size_t i = 0; i < n; i++) screen_rows[i] = h->tiles[i * TILE_ENTRY + 2];
    return BRKGA_OK;
}

// Greedy initial solution: the loop of BRKGA_alg3.py:445-603 on the host, the
// per-part tests of all open batches in one launch, the stamp in another.
extern "C" int brkga_initial_solution(brkga_handle *h, const double *type_height, const double *machine_keys,
                                      float *out_chromosome, double *out_makespan, int32_t *out_log,
                                      void *stream) {
    if (!h) return fail(nullptr, BRKGA_EINVAL, "handle is NULL");
    if (!type_height || !machine_keys || !out_chromosome || !out_makespan)
        return fail(h, BRKGA_EINVAL, "type_height, machine_keys, out_chromosome and out_makespan are required");
    const int n = h->prob.n_parts, M = h->prob.n_machines, T = h->prob.n_types, R = h->prob.n_rots;
    const int hmax = h->dw.hmax;
    const size_t gpw = (size_t)h->dw.pw;   // words per grid bit row
    hipStream_t s = (hipStream_t)stream;
    HIPCHK(h, hipSetDevice(h->device));
    // host copies of the tables the greedy's decisions read
    std::vector<int> slot_type(n), type_id(T), nrot(T), rot_base(T), best_rot(T), rot_h(R), rot_w(R), ph(M), pw(M);
    std::vector<double> area(T), setup(M), pt((size_t)M * T), pth((size_t)M * T);
    const DevProblem &d = h->dp;
    HIPCHK(h, hipMemcpy(slot_type.data(), d.slot_type, n * 4, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(type_id.data(), d.type_id, T * 4, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(nrot.data(), d.nrot, T * 4, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(rot_base.data(), d.rot_base, T * 4, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(best_rot.data(), d.best_rot, T * 4, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(rot_h.data(), d.rot_h, R * 4, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(rot_w.data(), d.rot_w, R * 4, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(ph.data(), d.plate_h, M * 4, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(pw.data(), d.plate_w, M * 4, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(area.data(), d.area, T * 8, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(setup.data(), d.setup, M * 8, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(pt.data(), d.proc_time, (size_t)M * T * 8, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(pth.data(), d.pth, (size_t)M * T * 8, hipMemcpyDeviceToHost));
    // slots by decreasing height, then decreasing area, ties in instance order
    std::vector<int> order(n);
    for (int j = 0; j < n; j++) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const int ta = slot_type[a], tb = slot_type[b];
        if (type_height[ta] != type_height[tb]) return type_height[ta] > type_height[tb];
        return area[ta] > area[tb];
    });
    // device batches: every machine may open up to n
    const size_t nslot = (size_t)M * n;
    unsigned *d_grids = nullptr;
    int *d_vac = nullptr, *d_jobs = nullptr, *d_res = nullptr;
    int rc = BRKGA_OK;
    auto cleanup = [&]() {
        if (d_grids) (void)hipFree(d_grids);
        if (d_vac) (void)hipFree(d_vac);
        if (d_jobs) (void)hipFree(d_jobs);
        if (d_res) (void)hipFree(d_res);
    };
    if (hipMalloc((void **)&d_grids, nslot * hmax * gpw * sizeof(unsigned)) != hipSuccess ||
        hipMalloc((void **)&d_vac, nslot * hmax * sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&d_jobs, nslot * 3 * sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&d_res, nslot * 4 * sizeof(int)) != hipSuccess) {
        cleanup();
        return fail(h, BRKGA_ENOMEM, "greedy workspace allocation failed");
    }
    struct Batch { double pt = 0.0, pth = 0.0; int minrow = 0; std::vector<int> parts; };
    std::vector<std::vector<Batch>> batches(M);
    std::vector<double> mk(M, 0.0);
    std::vector<int> jobs, res;
    double worst = 0.0, best_ms = 0.0;
    // one part; device errors return from here so the workspace is still freed
    auto place = [&](int k) -> int {
        const int ty = slot_type[order[k]], r0 = rot_base[ty];
        const int h0 = rot_h[r0], w0 = rot_w[r0];
        auto fits = [&](int m) { return !((h0 > ph[m] || w0 > pw[m]) && (w0 > ph[m] || h0 > pw[m])); };
        // test the part against every open batch of every machine it fits
        jobs.clear();
        for (int m = 0; m < M; m++)
            if (fits(m))
                for (int b = 0; b < (int)batches[m].size(); b++) {
                    jobs.push_back(m); jobs.push_back(b); jobs.push_back(ph[m] - batches[m][b].minrow);
                }
        const int nj = (int)jobs.size() / 3;
        res.assign((size_t)nj * 4, 0);
        if (nj) {
            GreedyArgs A;
            A.grids = d_grids; A.vac = d_vac; A.max_batches = n; A.ty = ty; A.jobs = d_jobs; A.res = d_res;
            HIPCHK(h, hipMemcpyAsync(d_jobs, jobs.data(), jobs.size() * sizeof(int), hipMemcpyHostToDevice, s));
            if (h->f4) hipLaunchKernelGGL(greedy_test_kernel<true>, dim3(nj), dim3(WG_THREADS), h->smem, s, h->dp, h->dw, A);
            else hipLaunchKernelGGL(greedy_test_kernel<false>, dim3(nj), dim3(WG_THREADS), h->smem, s, h->dp, h->dw, A);
            HIPCHK(h, hipGetLastError());
            HIPCHK(h, hipMemcpyAsync(res.data(), d_res, res.size() * sizeof(int), hipMemcpyDeviceToHost, s));
            HIPCHK(h, hipStreamSynchronize(s));
        }
        // the choice (BRKGA_alg3.py:467-507): "<=" lets later candidates win ties,
        // except a second equal existing batch of the same machine
        best_ms = INFINITY;
        int kind = 0, cm = -1, cb = -1, cx = 0, crow = 0, crot = 0, jo = 0;
        double cmm = 0.0;
        for (int m = 0; m < M; m++) {
            if (!fits(m)) continue;
            bool placed_exist = false;
            const double ptm = pt[(size_t)m * T + ty], pthm = pth[(size_t)m * T + ty];
            double mm = ((mk[m] + setup[m]) + ptm) + pthm;
            double nm = worst > mm ? worst : mm;
            if (nm <= best_ms) { best_ms = nm; kind = 1; cm = m; cmm = mm; }
            for (int b = 0; b < (int)batches[m].size(); b++, jo++) {
                const int *r = &res[(size_t)jo * 4];
                if (!r[0]) continue;
                const double bpth = batches[m][b].pth;
                mm = bpth < pthm ? ((mk[m] - bpth) + pthm) + ptm : mk[m] + ptm;
                nm = worst > mm ? worst : mm;
                if (nm <= best_ms) {
                    if (placed_exist && nm == best_ms) continue;
                    placed_exist = true; best_ms = nm; kind = 2; cm = m; cb = b;
                    cx = r[1]; crow = r[2]; crot = r[3]; cmm = mm;
                }
            }
        }
        if (kind == 0) return fail(h, BRKGA_EINVAL, "a part fits no machine");
        const int H = ph[cm];
        if (kind == 1) {
            cb = (int)batches[cm].size();
            batches[cm].emplace_back();
            batches[cm][cb].minrow = H;
            crot = best_rot[ty]; cx = 0; crow = H - 1;
        }
        const int rg = r0 + crot, hh = rot_h[rg];
        GreedyInsert I;
        const size_t slot = (size_t)cm * n + cb;
        I.grid = d_grids + slot * hmax * gpw; I.vac = d_vac + slot * hmax;
        I.rg = rg; I.H = H; I.W = pw[cm]; I.py = crow - hh + 1; I.px = cx; I.fresh = kind == 1;
        hipLaunchKernelGGL(greedy_insert_kernel, dim3(1), dim3(WG_THREADS), h->smem, s, h->dp, h->dw, I);
        HIPCHK(h, hipGetLastError());
        Batch &bt = batches[cm][cb];
        bt.pt += pt[(size_t)cm * T + ty];
        if (pth[(size_t)cm * T + ty] > bt.pth) bt.pth = pth[(size_t)cm * T + ty];
        if (I.py < bt.minrow) bt.minrow = I.py;
        bt.parts.push_back(type_id[ty]);
        mk[cm] = cmm;
        if (cmm > worst) worst = cmm;
        if (out_log) {
            int32_t *e = out_log + 6 * k;
            e[0] = cm; e[1] = type_id[ty]; e[2] = cb; e[3] = cx; e[4] = crow; e[5] = crot;
        }
        return BRKGA_OK;
    };
    for (int k = 0; k < n && rc == BRKGA_OK; k++) rc = place(k);
    if (rc == BRKGA_OK) rc = hipStreamSynchronize(s) == hipSuccess ? BRKGA_OK : fail(h, BRKGA_EDEVICE, "greedy kernels failed");
    cleanup();
    if (rc != BRKGA_OK) return rc;
    // the chromosome (BRKGA_alg3.py:555-595)
    std::vector<char> used(n, 0);
    for (int j = 0; j < 2 * n; j++) out_chromosome[j] = 0.0f;
    for (int m = 0; m < M; m++) {
        std::vector<int> pos;
        for (const Batch &bt : batches[m])
            for (int id : bt.parts)
                for (int j = 0; j < n; j++)
                    if (!used[j] && type_id[slot_type[j]] == id) { used[j] = 1; pos.push_back(j); break; }
        const double step = pos.empty() ? 0.0 : 1.0 / (double)pos.size();
        for (size_t i = 0; i < pos.size(); i++) {
            out_chromosome[n + pos[i]] = (float)machine_keys[m];
            out_chromosome[pos[i]] = (float)((double)i * step);
        }
    }
    *out_makespan = best_ms;
    return BRKGA_OK;
}

extern "C" int brkga_get_stats(const brkga_handle *hc, brkga_stats *out) {
    if (!hc || !out) return BRKGA_EINVAL;
    brkga_handle *h = const_cast<brkga_handle *>(hc);
    HIPCHK(h, hipSetDevice(h->device));
    settle_events(h);
    unsigned long long c[ST_COUNT];
    HIPCHK(h, hipMemcpy(c, h->dw.counters, sizeof c, hipMemcpyDeviceToHost));
    out->calls = (int64_t)c[ST_CALLS];
    out->chromosomes = (int64_t)c[ST_CHROM];
    out->contexts = (int64_t)c[ST_CTX];
    out->placements = (int64_t)c[ST_PLACE];
    out->bins_opened = (int64_t)c[ST_BINS];
    out->tests = (int64_t)c[ST_TESTS];
    out->row_groups = (int64_t)c[ST_GROUPS];
    out->mfma_ops = (int64_t)c[ST_MFMA];
    out->algo_ops = (int64_t)c[ST_ALGO];
    out->overflows = (int64_t)c[ST_OVERFLOW];
    out->kernel_ms = h->kernel_ms;
    out->last_kernel_ms = h->last_ms;
    const double us_per_tick = 1e6 / h->wall_hz;
    for (int k = 0; k < 8; k++) out->phase_us[k] = (double)c[ST_PH0 + k] * us_per_tick;
    out->context_us_max = (double)c[ST_CTXMAX] * us_per_tick;
    out->context_us_sum = (double)c[ST_CTXSUM] * us_per_tick;
    out->growths = h->growths;
    out->max_bins = h->dw.max_bins;
    out->n_workgroups = h->last_wg ? h->last_wg : h->lay[0].n_wg;
    for (int m = 0; m < BRKGA_MAX_MACHINES; m++) {
        const bool used = m < h->prob.n_machines;
        out->machine_contexts[m] = used ? (int64_t)c[ST_MCTX + m] : 0;
        out->machine_bins[m] = used ? (int64_t)c[ST_MBINS + m] : 0;
        out->machine_peak_bins[m] = used ? (int32_t)c[ST_MPEAK + m] : 0;
    }
    return 0;
}

extern "C" int brkga_reset_stats(brkga_handle *h) {
    if (!h) return BRKGA_EINVAL;
    HIPCHK(h, hipSetDevice(h->device));
    settle_events(h);
    HIPCHK(h, hipMemset(h->dw.counters, 0, ST_COUNT * sizeof(unsigned long long)));
    h->growths = 0;
    h->kernel_ms = 0.0;
    h->last_ms = 0.0;
    return 0;
}

extern "C" const char *brkga_last_error(const brkga_handle *h) {
    return h ? h->err.c_str() : g_create_error.c_str();
}

extern "C" void brkga_destroy(brkga_handle *h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    (void)hipDeviceSynchronize();
    for (void *p : h->allocs) (void)hipFree(p);
    if (h->d_keys) (void)hipFree(h->d_keys);
    if (h->d_out) (void)hipFree(h->d_out);
    if (h->d_log) (void)hipFree(h->d_log);
    if (h->d_ctx) (void)hipFree(h->d_ctx);
    if (h->d_order) (void)hipFree(h->d_order);
    if (h->d_bucket) (void)hipFree(h->d_bucket);
    if (h->d_hist) (void)hipFree(h->d_hist);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    delete h;
}
