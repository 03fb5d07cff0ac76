// brkga_surrogate.hip -- BF16 MLP surrogate of the decoder's makespan.
//
// Network: x (F = 4n features) -> 512 ReLU -> 512 ReLU -> 1.  Weights and
// activations are bfloat16, every product is accumulated in float32
// (V_MFMA_F32_32X32X16_BF16), the optimiser keeps float32 master weights.
// Features of a chromosome: its 2n keys, then the area and the height of the
// parts in sequence order (ascending sequence key), as the surrogate-screened
// configuration prescribes.  The reference has no surrogate; the network only
// chooses WHICH offspring are decoded exactly, never a fitness value.
//
// Kernels: featurize, a tiled GEMM with transposes and fused epilogues
// (bias, ReLU, ReLU-derivative mask, f32/bf16 outputs), the output layer and
// loss gradient, column sums for bias gradients, and Adam.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/brkga_mi355x.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint16_t bf16_t;   // storage type

__device__ __forceinline__ float bf2f(bf16_t v) { return __uint_as_float(((unsigned)v) << 16); }
// round to nearest even (finite inputs)
__device__ __forceinline__ bf16_t f2bf(float f) {
    const unsigned u = __float_as_uint(f);
    return (bf16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

#define GEMM_TILE 64
#define GEMM_KC 64
#define GEMM_LDS_K 72   // padded row length (bf16) of the LDS tiles
#define GEMM_SHARES (GEMM_TILE * GEMM_KC / 8 / 256)   // 16-byte shares per thread per operand tile
#define GEMM_SPLIT_MAX 8

enum { EPI_RELU = 1, EPI_MASK = 2 };

// Share q (0 <= q < 256 * GEMM_SHARES, 8 elements = 16 bytes) of a 64-row x
// GEMM_KC-k operand tile.  KCONT: element (r, k) sits at P[r*ld + k] (k
// contiguous); otherwise at P[k*ld + r] (rows contiguous).  Rows >= rows and
// k >= kend read as zero; the host guarantees whole vectors (K % 8 == 0 for
// KCONT, rows % 8 == 0 otherwise, ld % 8 == 0), so a vector is either wholly
// inside or wholly outside.
template <bool KCONT>
__device__ __forceinline__ uint4 stage_load(const bf16_t *P, int ld, int rows, int kend, int row0, int k0,
                                            int q) {
    uint4 v = {0u, 0u, 0u, 0u};
    if (KCONT) {
        const int r = q / (GEMM_KC / 8), kk = (q % (GEMM_KC / 8)) << 3;
        if (row0 + r < rows && k0 + kk < kend) v = *(const uint4 *)(P + (long long)(row0 + r) * ld + k0 + kk);
    } else {
        const int kk = q >> 3, r = (q & 7) << 3;
        if (k0 + kk < kend && row0 + r < rows) v = *(const uint4 *)(P + (long long)(k0 + kk) * ld + row0 + r);
    }
    return v;
}

// Store share q into the LDS tile S[row][k] (k contiguous, the MFMA operand order).
template <bool KCONT>
__device__ __forceinline__ void stage_store(bf16_t (*S)[GEMM_LDS_K], uint4 v, int q) {
    if (KCONT) {
        *(uint4 *)&S[q / (GEMM_KC / 8)][(q % (GEMM_KC / 8)) << 3] = v;
    } else {
        const int kk = q >> 3, r = (q & 7) << 3;
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            S[r + 2 * i][kk] = (bf16_t)(w[i] & 0xFFFFu);
            S[r + 2 * i + 1][kk] = (bf16_t)(w[i] >> 16);
        }
    }
}

// C[M][N] = op(A) op(B) (+ bias[n]) (ReLU) (* [mask[m][n] > 0]); f32 and/or bf16 out.
// op(A)(m,k) = TA ? A[k*lda + m] : A[m*lda + k];  op(B)(k,n) = TB ? B[n*ldb + k] : B[k*ldb + n].
// One workgroup = 4 waves = one 64x64 tile (2x2 of 32x32 MFMA tiles) over the k
// range [z*kper, (z+1)*kper) of its blockIdx.z.  With ws set, the raw partial
// product goes to ws[z][M][N] and splitk_reduce_kernel finishes it; the next
// k chunk is loaded into registers while the current one is multiplied.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(int M, int N, int K, int kper, const bf16_t *A, int lda,
                                                        const bf16_t *B, int ldb, const float *bias,
                                                        const bf16_t *mask, int ldm, int flags,
                                                        float *Cf, bf16_t *Cb, int ldc, float *ws) {
    __shared__ __align__(16) bf16_t As[GEMM_TILE][GEMM_LDS_K];
    __shared__ __align__(16) bf16_t Bs[GEMM_TILE][GEMM_LDS_K];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.y * GEMM_TILE, n0 = blockIdx.x * GEMM_TILE;
    const int wm = (wave >> 1) << 5, wn = (wave & 1) << 5;
    const int kb = blockIdx.z * kper, ke = min(K, kb + kper);
    f32x16 acc = {};
    uint4 ra[GEMM_SHARES], rb[GEMM_SHARES];
#pragma unroll
    for (int q = 0; q < GEMM_SHARES; q++) {
        ra[q] = stage_load<!TA>(A, lda, M, ke, m0, kb, tid + 256 * q);
        rb[q] = stage_load<TB>(B, ldb, N, ke, n0, kb, tid + 256 * q);
    }
    for (int k0 = kb; k0 < ke; k0 += GEMM_KC) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < GEMM_SHARES; q++) {
            stage_store<!TA>(As, ra[q], tid + 256 * q);
            stage_store<TB>(Bs, rb[q], tid + 256 * q);
        }
        __syncthreads();
        if (k0 + GEMM_KC < ke) {
#pragma unroll
            for (int q = 0; q < GEMM_SHARES; q++) {
                ra[q] = stage_load<!TA>(A, lda, M, ke, m0, k0 + GEMM_KC, tid + 256 * q);
                rb[q] = stage_load<TB>(B, ldb, N, ke, n0, k0 + GEMM_KC, tid + 256 * q);
            }
        }
#pragma unroll
        for (int ks = 0; ks < GEMM_KC; ks += 16) {
            const bf16x8 a = *(const bf16x8 *)&As[wm + (lane & 31)][ks + ((lane >> 5) << 3)];
            const bf16x8 b = *(const bf16x8 *)&Bs[wn + (lane & 31)][ks + ((lane >> 5) << 3)];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
    }
    const int n = n0 + wn + (lane & 31);
    if (n >= N) return;
    if (ws) {
        float *wz = ws + (long long)blockIdx.z * M * N;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int m = m0 + wm + ((i >> 2) << 3) + ((lane >> 5) << 2) + (i & 3);
            if (m < M) wz[(long long)m * N + n] = acc[i];
        }
        return;
    }
    const float bn = bias ? bias[n] : 0.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int m = m0 + wm + ((i >> 2) << 3) + ((lane >> 5) << 2) + (i & 3);
        if (m >= M) continue;
        float v = acc[i] + bn;
        if (flags & EPI_RELU) v = v > 0.0f ? v : 0.0f;
        if ((flags & EPI_MASK) && !(bf2f(mask[(long long)m * ldm + n]) > 0.0f)) v = 0.0f;
        if (Cf) Cf[(long long)m * ldc + n] = v;
        if (Cb) Cb[(long long)m * ldc + n] = f2bf(v);
    }
}

// C[m][n] = sum over z of ws[z][m][n], in z order (deterministic split-k).
__global__ void splitk_reduce_kernel(const float *ws, int S, int M, int N, float *C, int ldc) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)M * N) return;
    float v = 0.0f;
    for (int z = 0; z < S; z++) v += ws[(long long)z * M * N + t];
    C[(t / N) * ldc + t % N] = v;
}

// Machine of a machine key: the first threshold bucket holding it (placement.py:121-127).
__device__ __forceinline__ int machine_bucket(double mv, const double *thr, int M) {
    if (M <= 1 || mv <= thr[0]) return 0;
    for (int m = 1; m < M - 1; m++)
        if (mv <= thr[m]) return m;
    return M - 1;
}

// Features of each chromosome (one workgroup per row).  Slot features (4n):
// the keys as the decoder reads them (sequence position / n, machine bucket /
// max(M-1, 1)), then area and height of the parts ordered by ascending
// sequence key (ties by slot), each times ss.  With per-machine terms mt
// ([M][3][n]) four aggregates per machine m follow: the float32 sums of
// mt[m][0][j] and mt[m][1][j] over the slots j of machine m (64 strided lane
// sums, then a butterfly), their count times M / n, and the largest
// mt[m][2][j] among them.  Columns F..Fp-1 are 0.
__global__ __launch_bounds__(256) void featurize_kernel(const float *keys, int B, int n, int M,
                                                        const double *thr, int F, int Fp, float ss,
                                                        const float *slot_area, const float *slot_height,
                                                        const float *mt, bf16_t *x) {
    extern __shared__ float sk[];   // [n] sequence keys, then [n] machine buckets
    int *sb = (int *)(sk + n);
    const long long b = blockIdx.x;
    const float *kr = keys + b * 2LL * n;
    bf16_t *xr = x + b * Fp;
    for (int j = threadIdx.x; j < n; j += 256) {
        sk[j] = kr[j];
        sb[j] = machine_bucket((double)kr[n + j], thr, M);
    }
    for (int j = F + threadIdx.x; j < Fp; j += 256) xr[j] = 0;
    __syncthreads();
    const float inv_den = M > 1 ? (float)(M - 1) : 1.0f;
    for (int j = threadIdx.x; j < n; j += 256) {
        const float kj = sk[j];
        int rank = 0;
        for (int i = 0; i < n; i++) rank += (sk[i] < kj) || (sk[i] == kj && i < j);
        xr[j] = f2bf(((float)rank / (float)n) * ss);
        xr[n + j] = f2bf(((float)sb[j] / inv_den) * ss);
        xr[2 * n + rank] = f2bf(slot_area[j] * ss);
        xr[3 * n + rank] = f2bf(slot_height[j] * ss);
    }
    if (mt) {
        // wave w takes machines w, w + 4, ...: lane l adds the terms of slots
        // l, l + 64, ... in that order, then the 64 lane sums are added by a
        // butterfly (xor 32, 16, ..., 1); the count and the maximum are exact
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        for (int m = wave; m < M; m += 4) {
            const float *t0 = mt + (long long)m * 3 * n, *t1 = t0 + n, *t2 = t1 + n;
            float s0 = 0.0f, s1 = 0.0f, mx = 0.0f;
            int cnt = 0;
            for (int j = lane; j < n; j += 64) {
                const bool on = sb[j] == m;   // adding 0 to a sum of non-negative terms changes nothing
                s0 += on ? t0[j] : 0.0f;
                s1 += on ? t1[j] : 0.0f;
                mx = on ? fmaxf(mx, t2[j]) : mx;
                cnt += on;
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                s0 += __shfl_xor(s0, off, 64);
                s1 += __shfl_xor(s1, off, 64);
                mx = fmaxf(mx, __shfl_xor(mx, off, 64));
                cnt += __shfl_xor(cnt, off, 64);
            }
            if (lane == 0) {
                bf16_t *xa = xr + 4 * n + 4 * m;
                xa[0] = f2bf(s0);
                xa[1] = f2bf(s1);
                xa[2] = f2bf((float)(cnt * M) / (float)n);
                xa[3] = f2bf(mx);
            }
        }
    }
    (void)B;
}

// y[b] = H2[b] . w3 + b3 ; one wave per row.  With targets t (training) also
// dy[b] = 2 (y[b] - t[b]) / B and sq[b] = (y[b] - t[b])^2 / B.
__global__ __launch_bounds__(256) void output_kernel(const bf16_t *H, int B, int Hd, const float *w3,
                                                     const float *b3, float *y, const float *t, float *dy,
                                                     float *sq) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= B) return;
    float s = 0.0f;
    for (int j = lane; j < Hd; j += 64) s += bf2f(H[(long long)row * Hd + j]) * w3[j];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) {
        const float yr = s + b3[0];
        y[row] = yr;
        if (t) {
            const float d = yr - t[row];
            dy[row] = 2.0f * d / (float)B;
            sq[row] = d * d / (float)B;
        }
    }
}

// Column sums part[rb][j] = sum of Mx[b][j] over the 64-row block rb (one wave per
// 16 rows, 64 columns per workgroup); head_final_kernel adds the blocks in order.
#define COLSUM_ROWS 64
static_assert(COLSUM_ROWS == 64, "head_kernel reduces a block's rows with one 64-lane butterfly");
__global__ __launch_bounds__(256) void colsum_partial_kernel(const bf16_t *Mx, int B, int Hd, float *part) {
    __shared__ float sh[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + lane, rb = blockIdx.y;
    float s = 0.0f;
    if (j < Hd) {
#pragma unroll
        for (int i = 0; i < COLSUM_ROWS / 4; i++) {
            const int b = rb * COLSUM_ROWS + wv * (COLSUM_ROWS / 4) + i;
            if (b < B) s += bf2f(Mx[(long long)b * Hd + j]);
        }
    }
    sh[wv][lane] = s;
    __syncthreads();
    if (wv == 0 && j < Hd) part[(long long)rb * Hd + j] = (sh[0][lane] + sh[1][lane]) + (sh[2][lane] + sh[3][lane]);
}

// Backward through the output layer for one 64-row block and 64 columns
// (grid R x ceil(Hd / 64), one wave per 16 rows):
//   dH2[b][j] = dy[b] w3[j] [H2[b][j] > 0]   (bf16)
//   dw3 partial[j] = sum_b H2[b][j] dy[b],   db2 partial[j] = sum_b dH2[b][j]
// into part3 / part2 [block][Hd]; the first column slice also writes
// scal[block] = (sum_b sq[b], sum_b dy[b]) -- the loss and db3 partials (COLSUM_ROWS
// is one wave's width, so each is one butterfly).
// head_final_kernel adds the blocks in order, so every sum has a fixed order.
__global__ __launch_bounds__(256) void head_kernel(const bf16_t *H2, int B, int Hd, const float *w3,
                                                   const float *dy, const float *sq, bf16_t *dH2,
                                                   float *part3, float *part2, float *scal) {
    __shared__ float sdy[COLSUM_ROWS];
    __shared__ float s3[4][64], s2[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, blk = blockIdx.x;
    const int r0 = blk * COLSUM_ROWS;
    const int nr = min(COLSUM_ROWS, B - r0);
    if (threadIdx.x < COLSUM_ROWS) {
        const bool in = (int)threadIdx.x < nr;
        float g = in ? dy[r0 + threadIdx.x] : 0.0f;
        sdy[threadIdx.x] = g;
        if (blockIdx.y == 0) {   // loss and db3 partials: one butterfly over the block's rows
            float l = in ? sq[r0 + threadIdx.x] : 0.0f;
            for (int off = 32; off > 0; off >>= 1) {
                l += __shfl_xor(l, off, 64);
                g += __shfl_xor(g, off, 64);
            }
            if (threadIdx.x == 0) {
                scal[2 * blk] = l;
                scal[2 * blk + 1] = g;
            }
        }
    }
    __syncthreads();
    const int j = blockIdx.y * 64 + lane;
    float a3 = 0.0f, a2 = 0.0f;
    if (j < Hd) {
        const float wj = w3[j];
        float h[COLSUM_ROWS / 4];
#pragma unroll
        for (int i = 0; i < COLSUM_ROWS / 4; i++) {
            const int r = wv * (COLSUM_ROWS / 4) + i;
            h[i] = r < nr ? bf2f(H2[(long long)(r0 + r) * Hd + j]) : 0.0f;
        }
#pragma unroll
        for (int i = 0; i < COLSUM_ROWS / 4; i++) {
            const int r = wv * (COLSUM_ROWS / 4) + i;
            if (r < nr) {
                const bf16_t gh = f2bf(h[i] > 0.0f ? sdy[r] * wj : 0.0f);
                dH2[(long long)(r0 + r) * Hd + j] = gh;
                a3 += h[i] * sdy[r];
                a2 += bf2f(gh);
            }
        }
    }
    s3[wv][lane] = a3;
    s2[wv][lane] = a2;
    __syncthreads();
    if (wv == 0 && j < Hd) {
        part3[(long long)blk * Hd + j] = (s3[0][lane] + s3[1][lane]) + (s3[2][lane] + s3[3][lane]);
        part2[(long long)blk * Hd + j] = (s2[0][lane] + s2[1][lane]) + (s2[2][lane] + s2[3][lane]);
    }
}

// Sums the R row blocks in order: db1 (part1), db2 (part2), dw3 (part3), and
// thread 0 the loss and db3 (scal).
__global__ void head_final_kernel(const float *part1, const float *part2, const float *part3,
                                  const float *scal, int R, int Hd, float *db1, float *db2, float *dw3,
                                  float *db3, float *loss) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < Hd) {
        float s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
#pragma unroll 8
        for (int r = 0; r < R; r++) {
            s1 += part1[(long long)r * Hd + j];
            s2 += part2[(long long)r * Hd + j];
            s3 += part3[(long long)r * Hd + j];
        }
        db1[j] = s1;
        db2[j] = s2;
        dw3[j] = s3;
    }
    if (j == 0) {
        float l = 0.0f, d = 0.0f;
#pragma unroll 8
        for (int r = 0; r < R; r++) { l += scal[2 * r]; d += scal[2 * r + 1]; }
        loss[0] = l;
        db3[0] = d;
    }
}

__global__ void adam_kernel(float *p, const float *g, float *m, float *v, long long n, float lr, float b1,
                            float b2, float eps, float c1, float c2, bf16_t *pb, const long long *bf_map,
                            float wd) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] + wd * p[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    (void)bf_map;
    if (pb) pb[i] = f2bf(pi);
}

__global__ void to_bf16_kernel(const float *p, long long n, bf16_t *pb) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pb[i] = f2bf(p[i]);
}

struct brkga_mlp {
    int F = 0, Fp = 0, Hd = 512, n = 0, M = 1;
    double *thr = nullptr;   // [max(M-1, 1)] machine-key thresholds
    int device = 0;
    // parameters: W1 [Fp][Hd], b1 [Hd], W2 [Hd][Hd], b2 [Hd], w3 [Hd], b3 [1] (f32 master)
    long long oW1, ob1, oW2, ob2, ow3, ob3, np_;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    bf16_t *pb = nullptr;   // bf16 copy of every parameter (GEMMs read W1, W2)
    float *slot_area = nullptr, *slot_height = nullptr;
    float *mterms = nullptr;   // [M][3][n] per-machine terms (null: slot features only)
    float ss = 1.0f;           // slot feature scale
    // activations (grown on demand)
    long long cap = 0;
    bf16_t *x = nullptr, *H1 = nullptr, *H2 = nullptr, *dH2 = nullptr, *dH1 = nullptr;
    float *y = nullptr, *dy = nullptr, *sq = nullptr, *loss = nullptr;
    float *part = nullptr;   // column-sum partials [cap / 64][Hd]: db1
    float *part2 = nullptr, *part3 = nullptr;   // the same for db2, dw3
    float *scal = nullptr;   // [cap / 64][2]: loss and db3 partials
    float *ws = nullptr;     // split-k partials [GEMM_SPLIT_MAX][max(Fp, Hd)][Hd]
    long long step = 0;
    float lr = 1e-3f, wd = 0.0f;
    std::vector<void *> allocs;
    std::string err;
};

static thread_local std::string g_mlp_error;

#define MLPCHK(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            g_mlp_error = std::string(#call) + ": " + hipGetErrorString(e_);                  \
            return e_ == hipErrorOutOfMemory ? BRKGA_ENOMEM : BRKGA_EDEVICE;                  \
        }                                                                                     \
    } while (0)

static int mlp_alloc(brkga_mlp *net, size_t bytes, void **q) {
    MLPCHK(hipMalloc(q, std::max<size_t>(bytes, 16)));
    net->allocs.push_back(*q);
    return 0;
}

static int mlp_reserve(brkga_mlp *net, long long B) {
    if (B <= net->cap) return 0;
    for (void *q : {(void *)net->x, (void *)net->H1, (void *)net->H2, (void *)net->dH2, (void *)net->dH1,
                    (void *)net->y, (void *)net->dy, (void *)net->sq, (void *)net->part, (void *)net->part2,
                    (void *)net->part3, (void *)net->scal}) {
        if (!q) continue;
        auto it = std::find(net->allocs.begin(), net->allocs.end(), q);
        if (it != net->allocs.end()) { (void)hipFree(q); net->allocs.erase(it); }
    }
    void *q;
    int rc;
    if ((rc = mlp_alloc(net, (size_t)B * net->Fp * 2, &q))) return rc; net->x = (bf16_t *)q;
    if ((rc = mlp_alloc(net, (size_t)B * net->Hd * 2, &q))) return rc; net->H1 = (bf16_t *)q;
    if ((rc = mlp_alloc(net, (size_t)B * net->Hd * 2, &q))) return rc; net->H2 = (bf16_t *)q;
    if ((rc = mlp_alloc(net, (size_t)B * net->Hd * 2, &q))) return rc; net->dH2 = (bf16_t *)q;
    if ((rc = mlp_alloc(net, (size_t)B * net->Hd * 2, &q))) return rc; net->dH1 = (bf16_t *)q;
    if ((rc = mlp_alloc(net, (size_t)B * 4, &q))) return rc; net->y = (float *)q;
    if ((rc = mlp_alloc(net, (size_t)B * 4, &q))) return rc; net->dy = (float *)q;
    if ((rc = mlp_alloc(net, (size_t)B * 4, &q))) return rc; net->sq = (float *)q;
    const size_t R = (size_t)((B + COLSUM_ROWS - 1) / COLSUM_ROWS);
    if ((rc = mlp_alloc(net, R * net->Hd * 4, &q))) return rc; net->part = (float *)q;
    if ((rc = mlp_alloc(net, R * net->Hd * 4, &q))) return rc; net->part2 = (float *)q;
    if ((rc = mlp_alloc(net, R * net->Hd * 4, &q))) return rc; net->part3 = (float *)q;
    if ((rc = mlp_alloc(net, R * 2 * 4, &q))) return rc; net->scal = (float *)q;
    net->cap = B;
    return 0;
}

// Shapes the staging vectors need (see stage_load); every call here meets them.
template <bool TA, bool TB>
static bool gemm_shape_ok(int M, int N, int K, int lda, int ldb) {
    if (lda % 8 || ldb % 8) return false;
    if (!TA && K % 8) return false;
    if (TA && M % 8) return false;
    if (TB && K % 8) return false;
    if (!TB && N % 8) return false;
    return true;
}

template <bool TA, bool TB>
static int gemm(hipStream_t s, int M, int N, int K, const bf16_t *A, int lda, const bf16_t *B, int ldb,
                const float *bias, const bf16_t *mask, int ldm, int flags, float *Cf, bf16_t *Cb, int ldc,
                float *ws = nullptr) {
    if (!gemm_shape_ok<TA, TB>(M, N, K, lda, ldb)) { g_mlp_error = "gemm shape"; return BRKGA_EINVAL; }
    const int tiles = ((N + GEMM_TILE - 1) / GEMM_TILE) * ((M + GEMM_TILE - 1) / GEMM_TILE);
    int S = 1;
    if (ws && Cf && !Cb && !bias && !flags) {   // split k over the 256 CUs (weight gradients)
        S = std::max(1, std::min(GEMM_SPLIT_MAX, 256 / tiles));
        S = std::min(S, std::max(1, K / 128));
    }
    const int kper = ((K + S - 1) / S + GEMM_KC - 1) / GEMM_KC * GEMM_KC;
    S = (K + kper - 1) / kper;
    dim3 grid((N + GEMM_TILE - 1) / GEMM_TILE, (M + GEMM_TILE - 1) / GEMM_TILE, S);
    hipLaunchKernelGGL((gemm_bf16_kernel<TA, TB>), grid, dim3(256), 0, s, M, N, K, kper, A, lda, B, ldb, bias,
                       mask, ldm, flags, Cf, Cb, ldc, S > 1 ? ws : nullptr);
    if (S > 1)
        hipLaunchKernelGGL(splitk_reduce_kernel, dim3((unsigned)(((long long)M * N + 255) / 256)), dim3(256), 0, s,
                           ws, S, M, N, Cf, ldc);
    return 0;
}

// Forward pass from net->x: H1, H2 and y; with targets t also the loss
// gradient dy and the squared errors.
static int forward(brkga_mlp *net, long long B, hipStream_t s, const float *t) {
    const int Hd = net->Hd;
    int rc;
    if ((rc = gemm<false, false>(s, (int)B, Hd, net->Fp, net->x, net->Fp, net->pb + net->oW1, Hd,
                                 net->p + net->ob1, nullptr, 0, EPI_RELU, nullptr, net->H1, Hd)) ||
        (rc = gemm<false, false>(s, (int)B, Hd, Hd, net->H1, Hd, net->pb + net->oW2, Hd, net->p + net->ob2,
                                 nullptr, 0, EPI_RELU, nullptr, net->H2, Hd)))
        return rc;
    hipLaunchKernelGGL(output_kernel, dim3((unsigned)((B + 3) / 4)), dim3(256), 0, s, net->H2, (int)B, Hd,
                       net->p + net->ow3, net->p + net->ob3, net->y, t, net->dy, net->sq);
    return 0;
}

// Gradients of the mean squared error into net->g and the loss into net->loss
// (after forward with the targets; the inputs are already in net->x).
static int backward(brkga_mlp *net, long long B, hipStream_t s) {
    const int Hd = net->Hd, R = (int)((B + COLSUM_ROWS - 1) / COLSUM_ROWS);
    int rc;
    // dH2 and the partial sums of dw3, db2, db3 and the loss
    hipLaunchKernelGGL(head_kernel, dim3((unsigned)R, (unsigned)((Hd + 63) / 64)), dim3(256), 0, s, net->H2, (int)B,
                       Hd, net->p + net->ow3, net->dy, net->sq, net->dH2, net->part3, net->part2, net->scal);
    // dW2 = H1^T dH2
    if ((rc = gemm<true, false>(s, Hd, Hd, (int)B, net->H1, Hd, net->dH2, Hd, nullptr, nullptr, 0, 0,
                                net->g + net->oW2, nullptr, Hd, net->ws)))
        return rc;
    // dH1 = (dH2 W2^T) * [H1 > 0]
    if ((rc = gemm<false, true>(s, (int)B, Hd, Hd, net->dH2, Hd, net->pb + net->oW2, Hd, nullptr, net->H1, Hd,
                                EPI_MASK, nullptr, net->dH1, Hd)))
        return rc;
    hipLaunchKernelGGL(colsum_partial_kernel, dim3((Hd + 63) / 64, R), dim3(256), 0, s, net->dH1, (int)B, Hd,
                       net->part);
    // db1, db2, dw3, db3 and the loss, each summed over the row blocks in order
    hipLaunchKernelGGL(head_final_kernel, dim3((Hd + 255) / 256), dim3(256), 0, s, net->part, net->part2,
                       net->part3, net->scal, R, Hd, net->g + net->ob1, net->g + net->ob2, net->g + net->ow3,
                       net->g + net->ob3, net->loss);
    // dW1 = X^T dH1
    return gemm<true, false>(s, net->Fp, Hd, (int)B, net->x, net->Fp, net->dH1, Hd, nullptr, nullptr, 0, 0,
                             net->g + net->oW1, nullptr, Hd, net->ws);
}

extern "C" int brkga_mlp_create(int32_t n_parts, int32_t n_machines, const double *thresholds,
                                int32_t hidden, const float *slot_area, const float *slot_height,
                                const float *machine_terms, uint64_t seed, float lr, int32_t device,
                                brkga_mlp **out) {
    if (!out || n_parts < 1 || n_parts > 4096 || n_machines < 1 || n_machines > 256 ||
        (n_machines > 1 && !thresholds) || hidden < 32 || hidden % 32 || !slot_area || !slot_height) {
        g_mlp_error = "bad argument (1 <= n_parts <= 4096, 1 <= n_machines <= 256 with n_machines-1 "
                      "thresholds, hidden a positive multiple of 32, feature arrays)";
        return BRKGA_EINVAL;
    }
    for (int m = 1; m + 1 < n_machines; m++)
        if (!(thresholds[m] >= thresholds[m - 1])) { g_mlp_error = "thresholds must be ascending"; return BRKGA_EINVAL; }
    *out = nullptr;
    brkga_mlp *net = new (std::nothrow) brkga_mlp();
    if (!net) { g_mlp_error = "host allocation failed"; return BRKGA_ENOMEM; }
    auto bail = [&](int rc) { brkga_mlp_destroy(net); return rc; };
    net->device = device;
    net->n = n_parts;
    net->M = n_machines;
    // with per-machine aggregates the slot features are scaled down by 8: the
    // network then leans on the aggregates first and overfits the 4n slot
    // columns less (measured on 4-machine instances)
    net->F = 4 * n_parts + (machine_terms ? 4 * n_machines : 0);
    net->ss = machine_terms ? 0.125f : 1.0f;
    net->Fp = (net->F + GEMM_KC - 1) / GEMM_KC * GEMM_KC;
    net->Hd = hidden;
    net->lr = lr > 0 ? lr : 1e-3f;
    if (hipSetDevice(device) != hipSuccess) { g_mlp_error = "hipSetDevice failed"; return bail(BRKGA_EDEVICE); }
    const long long Fp = net->Fp, Hd = hidden;
    net->oW1 = 0; net->ob1 = Fp * Hd; net->oW2 = net->ob1 + Hd; net->ob2 = net->oW2 + Hd * Hd;
    net->ow3 = net->ob2 + Hd; net->ob3 = net->ow3 + Hd; net->np_ = net->ob3 + 1;
    // He-uniform initialisation (host, splitmix64 stream), zero biases
    std::vector<float> h((size_t)net->np_, 0.0f);
    unsigned long long z = seed * 0x9e3779b97f4a7c15ull + 12345;
    auto uni = [&]() {
        z += 0x9e3779b97f4a7c15ull;
        unsigned long long x = z;
        x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
        x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
        x ^= x >> 31;
        return (float)((double)(x >> 11) * (1.0 / 9007199254740992.0)) * 2.0f - 1.0f;
    };
    const float s1 = std::sqrt(6.0f / (float)net->F), s2 = std::sqrt(6.0f / (float)Hd);
    for (long long i = 0; i < Fp * Hd; i++) h[(size_t)(net->oW1 + i)] = (i / Hd) < net->F ? uni() * s1 : 0.0f;
    for (long long i = 0; i < Hd * Hd; i++) h[(size_t)(net->oW2 + i)] = uni() * s2;
    for (long long i = 0; i < Hd; i++) h[(size_t)(net->ow3 + i)] = uni() * s2;
    void *q;
    int rc;
    if ((rc = mlp_alloc(net, (size_t)net->np_ * 4, &q))) return bail(rc); net->p = (float *)q;
    if ((rc = mlp_alloc(net, (size_t)net->np_ * 4, &q))) return bail(rc); net->g = (float *)q;
    if ((rc = mlp_alloc(net, (size_t)net->np_ * 4, &q))) return bail(rc); net->m = (float *)q;
    if ((rc = mlp_alloc(net, (size_t)net->np_ * 4, &q))) return bail(rc); net->v = (float *)q;
    if ((rc = mlp_alloc(net, (size_t)net->np_ * 2, &q))) return bail(rc); net->pb = (bf16_t *)q;
    if ((rc = mlp_alloc(net, (size_t)n_parts * 4, &q))) return bail(rc); net->slot_area = (float *)q;
    if ((rc = mlp_alloc(net, (size_t)n_parts * 4, &q))) return bail(rc); net->slot_height = (float *)q;
    if (machine_terms) {
        if ((rc = mlp_alloc(net, (size_t)n_machines * 3 * n_parts * 4, &q))) return bail(rc);
        net->mterms = (float *)q;
        if (hipMemcpy(net->mterms, machine_terms, (size_t)n_machines * 3 * n_parts * 4, hipMemcpyHostToDevice) !=
            hipSuccess) {
            g_mlp_error = "device initialisation failed";
            return bail(BRKGA_EDEVICE);
        }
    }
    if ((rc = mlp_alloc(net, 16, &q))) return bail(rc); net->loss = (float *)q;
    if ((rc = mlp_alloc(net, (size_t)GEMM_SPLIT_MAX * std::max(Fp, Hd) * Hd * 4, &q))) return bail(rc);
    net->ws = (float *)q;
    std::vector<double> thr((size_t)std::max(n_machines - 1, 1), 0.0);
    for (int m = 0; m + 1 < n_machines; m++) thr[(size_t)m] = thresholds[m];
    if ((rc = mlp_alloc(net, thr.size() * 8, &q))) return bail(rc); net->thr = (double *)q;
    if (hipMemcpy(net->p, h.data(), (size_t)net->np_ * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(net->g, 0, (size_t)net->np_ * 4) != hipSuccess ||
        hipMemset(net->m, 0, (size_t)net->np_ * 4) != hipSuccess ||
        hipMemset(net->v, 0, (size_t)net->np_ * 4) != hipSuccess ||
        hipMemcpy(net->slot_area, slot_area, (size_t)n_parts * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(net->slot_height, slot_height, (size_t)n_parts * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(net->thr, thr.data(), thr.size() * 8, hipMemcpyHostToDevice) != hipSuccess) {
        g_mlp_error = "device initialisation failed";
        return bail(BRKGA_EDEVICE);
    }
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((net->np_ + 255) / 256)), dim3(256), 0, 0, net->p, net->np_,
                       net->pb);
    if (hipDeviceSynchronize() != hipSuccess) { g_mlp_error = "initialisation kernel failed"; return bail(BRKGA_EDEVICE); }
    *out = net;
    return BRKGA_OK;
}

static int mlp_features(brkga_mlp *net, const float *d_keys, long long B, hipStream_t s) {
    int rc = mlp_reserve(net, B);
    if (rc) return rc;
    hipLaunchKernelGGL(featurize_kernel, dim3((unsigned)B), dim3(256), (size_t)net->n * 8, s, d_keys, (int)B, net->n,
                       net->M, net->thr, net->F, net->Fp, net->ss, net->slot_area, net->slot_height, net->mterms,
                       net->x);
    MLPCHK(hipGetLastError());
    return 0;
}

extern "C" int brkga_mlp_predict(brkga_mlp *net, const float *d_keys, int64_t B, float *d_pred, void *stream) {
    if (!net || !d_keys || !d_pred || B < 1) { g_mlp_error = "bad argument"; return BRKGA_EINVAL; }
    MLPCHK(hipSetDevice(net->device));
    hipStream_t s = (hipStream_t)stream;
    int rc = mlp_features(net, d_keys, B, s);
    if (rc || (rc = forward(net, B, s, nullptr))) return rc;
    MLPCHK(hipMemcpyAsync(d_pred, net->y, (size_t)B * 4, hipMemcpyDeviceToDevice, s));
    MLPCHK(hipGetLastError());
    return 0;
}

extern "C" int brkga_mlp_gradients(brkga_mlp *net, const float *d_keys, const float *d_target, int64_t B,
                                   float *d_grad, void *stream) {
    if (!net || !d_keys || !d_target || B < 1) { g_mlp_error = "bad argument"; return BRKGA_EINVAL; }
    MLPCHK(hipSetDevice(net->device));
    hipStream_t s = (hipStream_t)stream;
    int rc = mlp_features(net, d_keys, B, s);
    if (rc || (rc = forward(net, B, s, d_target))) return rc;
    if ((rc = backward(net, B, s))) return rc;
    if (d_grad) MLPCHK(hipMemcpyAsync(d_grad, net->g, (size_t)net->np_ * 4, hipMemcpyDeviceToDevice, s));
    MLPCHK(hipGetLastError());
    return 0;
}

extern "C" int brkga_mlp_train_step(brkga_mlp *net, const float *d_keys, const float *d_target, int64_t B,
                                    float *h_loss, void *stream) {
    int rc = brkga_mlp_gradients(net, d_keys, d_target, B, nullptr, stream);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    net->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)net->step), c2 = 1.0f - std::pow(b2, (float)net->step);
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((net->np_ + 255) / 256)), dim3(256), 0, s, net->p, net->g, net->m,
                       net->v, net->np_, net->lr, b1, b2, 1e-8f, c1, c2, net->pb, nullptr, net->wd);
    MLPCHK(hipGetLastError());
    if (h_loss) {
        MLPCHK(hipMemcpyAsync(h_loss, net->loss, 4, hipMemcpyDeviceToHost, s));
        MLPCHK(hipStreamSynchronize(s));
    }
    return 0;
}

extern "C" int brkga_mlp_num_params(const brkga_mlp *net, int64_t *n_params, int32_t *n_features_padded) {
    if (!net) return BRKGA_EINVAL;
    if (n_params) *n_params = net->np_;
    if (n_features_padded) *n_features_padded = net->Fp;
    return 0;
}

extern "C" int brkga_mlp_get_params(const brkga_mlp *net, float *h_params) {
    if (!net || !h_params) return BRKGA_EINVAL;
    MLPCHK(hipSetDevice(net->device));
    MLPCHK(hipDeviceSynchronize());
    MLPCHK(hipMemcpy(h_params, net->p, (size_t)net->np_ * 4, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int brkga_mlp_set_params(brkga_mlp *net, const float *h_params) {
    if (!net || !h_params) return BRKGA_EINVAL;
    MLPCHK(hipSetDevice(net->device));
    MLPCHK(hipMemcpy(net->p, h_params, (size_t)net->np_ * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((net->np_ + 255) / 256)), dim3(256), 0, 0, net->p, net->np_,
                       net->pb);
    MLPCHK(hipDeviceSynchronize());
    return 0;
}

extern "C" int brkga_mlp_features(brkga_mlp *net, const float *d_keys, int64_t B, uint16_t *d_x, void *stream) {
    if (!net || !d_keys || !d_x || B < 1) return BRKGA_EINVAL;
    MLPCHK(hipSetDevice(net->device));
    hipStream_t s = (hipStream_t)stream;
    int rc = mlp_features(net, d_keys, B, s);
    if (rc) return rc;
    MLPCHK(hipMemcpyAsync(d_x, net->x, (size_t)B * net->Fp * 2, hipMemcpyDeviceToDevice, s));
    return 0;
}

extern "C" const char *brkga_mlp_last_error(void) { return g_mlp_error.c_str(); }

extern "C" void brkga_mlp_destroy(brkga_mlp *net) {
    if (!net) return;
    (void)hipSetDevice(net->device);
    (void)hipDeviceSynchronize();
    for (void *q : net->allocs) (void)hipFree(q);
    delete net;
}
