// brkga_ga.hip -- device-resident BRKGA generation loop.
//
// The population, its fitness and the best-so-far live in HBM.  One
// generation is five kernels plus the decode, all on one stream:
//
//   ga_sort      order the population by (fitness, index)      BRKGA_alg3.py:178-186
//   ga_breed     next population = elites ++ mutants ++ biased   BRKGA_alg3.py:188-208,
//                crossover offspring (elite gene with prob rho)  :233-244
//   decode       makespans of the mutants and offspring          (brkga_decoder.hip)
//   ga_elitefit  elites keep their fitness                       :250
//   ga_commit    next population/fitness become current
//   ga_best      best-so-far, per-generation minimum, counter++  :252-262
//
// The sequence is captured once as a hipGraph and replayed per generation;
// the generation counter lives on the device, so every replay draws fresh
// random numbers.  Random numbers come from a counter-based hash of
// (seed, generation, row, gene, stream) -- splitmix64's finaliser -- so the
// run is reproducible and independent of launch shape.  It does not draw the
// same numbers as NumPy's MT19937; the host loop (brkga_fft2_amd.brkga.BRKGA)
// keeps NumPy's stream for runs that must match the reference's draws.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/brkga_mi355x.h"
#include "brkga_internal.h"

#define GA_SORT_MAX 4096
#define GA_SORT_THREADS 1024

__device__ __forceinline__ unsigned long long ga_mix(unsigned long long z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ unsigned long long ga_hash(unsigned long long seed, unsigned gen, int row,
                                                      int gene, int stream) {
    unsigned long long z = ga_mix(seed ^ ((unsigned long long)gen << 32));
    z = ga_mix(z ^ (unsigned long long)(unsigned)row);
    return ga_mix(z ^ (((unsigned long long)(unsigned)gene << 3) | (unsigned)stream));
}

// uniform in [0, 1) on the float32 grid of 2^-24
__device__ __forceinline__ float ga_uniform(unsigned long long h) {
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}

// Bitonic sort of (fitness, index) pairs, ascending; one workgroup, N <= 4096.
__global__ __launch_bounds__(GA_SORT_THREADS) void ga_sort_kernel(const double *fit, int N, int *order) {
    __shared__ double key[GA_SORT_MAX];
    __shared__ int idx[GA_SORT_MAX];
    int P = 1;
    while (P < N) P <<= 1;
    for (int i = threadIdx.x; i < P; i += blockDim.x) {
        // NaN (a context that overflowed its bins) sorts as +inf, so the
        // comparisons below stay a total order
        const double f = i < N ? fit[i] : INFINITY;
        key[i] = f != f ? INFINITY : f;
        idx[i] = i < N ? i : 0x7fffffff;
    }
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += blockDim.x) {
                const int l = i ^ j;
                if (l > i) {
                    const bool up = (i & k) == 0;
                    const bool gt = key[i] > key[l] || (key[i] == key[l] && idx[i] > idx[l]);
                    if (gt == up) {
                        const double tk = key[i]; key[i] = key[l]; key[l] = tk;
                        const int ti = idx[i]; idx[i] = idx[l]; idx[l] = ti;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < N; i += blockDim.x) order[i] = idx[i];
}

__global__ void ga_breed_kernel(const float *pop, const int *order, int N, int G2, int ne, int nm,
                                float rho, unsigned long long seed, const unsigned *gen, float *next) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)N * G2) return;
    const int r = (int)(t / G2), g = (int)(t % G2);
    const unsigned ge = *gen;
    float v;
    if (r < ne) {
        v = pop[(long long)order[r] * G2 + g];
    } else if (r < ne + nm) {
        v = ga_uniform(ga_hash(seed, ge, r, g, 0));
    } else {
        const int e = order[(int)(ga_hash(seed, ge, r, 0, 1) % (unsigned long long)ne)];
        const int o = order[ne + (int)(ga_hash(seed, ge, r, 0, 2) % (unsigned long long)(N - ne))];
        const float u = ga_uniform(ga_hash(seed, ge, r, g, 3));
        v = u < rho ? pop[(long long)e * G2 + g] : pop[(long long)o * G2 + g];
    }
    next[t] = v;
}

__global__ void ga_elitefit_kernel(const double *fit, const int *order, int ne, double *fit_next) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < ne) fit_next[r] = fit[order[r]];
}

__global__ void ga_commit_kernel(const float *next, const double *fit_next, long long n_keys, int N,
                                 float *pop, double *fit) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n_keys) pop[t] = next[t];
    if (t < N) fit[t] = fit_next[t];
}

// Minimum of the generation (first index on ties), best-so-far, history,
// generation counter.  One workgroup.
__global__ __launch_bounds__(256) void ga_best_kernel(const double *fit, const float *pop, int N, int G2,
                                                      unsigned *gen, double *best, float *best_keys,
                                                      double *hist, int hist_cap) {
    __shared__ double sv[256];
    __shared__ int si[256];
    double v = INFINITY;
    int vi = 0x7fffffff;
    for (int i = threadIdx.x; i < N; i += 256) {
        const double f = fit[i];
        if (f < v || (f == v && i < vi)) { v = f; vi = i; }
    }
    sv[threadIdx.x] = v;
    si[threadIdx.x] = vi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double o = sv[threadIdx.x + s];
            const int oi = si[threadIdx.x + s];
            if (o < sv[threadIdx.x] || (o == sv[threadIdx.x] && oi < si[threadIdx.x])) {
                sv[threadIdx.x] = o;
                si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    const double m = sv[0];
    const int mi = si[0];
    const bool better = m < *best;
    __syncthreads();
    if (better)
        for (int g = threadIdx.x; g < G2; g += 256) best_keys[g] = pop[(long long)mi * G2 + g];
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned ge = *gen;
        if ((int)ge < hist_cap) hist[ge] = m;
        if (better) *best = m;
        *gen = ge + 1;
    }
}

__global__ void ga_init_kernel(unsigned *gen, double *best) {
    *gen = 0;
    *best = INFINITY;
}

struct brkga_ga {
    brkga_handle *dec = nullptr;
    int N = 0, G2 = 0, ne = 0, nm = 0;
    float rho = 0.7f;
    unsigned long long seed = 0;
    float *pop = nullptr, *next = nullptr, *best_keys = nullptr;
    double *fit = nullptr, *fit_next = nullptr, *best = nullptr, *hist = nullptr;
    int *order = nullptr;
    unsigned *gen = nullptr;
    int hist_cap = 0;
    int generations_done = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
    unsigned graph_ws_gen = 0;   // decoder workspace generation the graph was captured with
    std::vector<void *> allocs;
};

static thread_local std::string g_ga_error;

#define GACHK(call)                                                                           \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            g_ga_error = std::string(#call) + ": " + hipGetErrorString(e_);                   \
            return e_ == hipErrorOutOfMemory ? BRKGA_ENOMEM : BRKGA_EDEVICE;                  \
        }                                                                                     \
    } while (0)

static int ga_alloc(brkga_ga *ga, size_t bytes, void **p) {
    GACHK(hipMalloc(p, std::max<size_t>(bytes, 16)));
    ga->allocs.push_back(*p);
    return 0;
}

// One generation, enqueued on s (captured into the graph by brkga_ga_run).
static int ga_generation(brkga_ga *ga, hipStream_t s) {
    const long long nk = (long long)ga->N * ga->G2;
    hipLaunchKernelGGL(ga_sort_kernel, dim3(1), dim3(GA_SORT_THREADS), 0, s, ga->fit, ga->N, ga->order);
    hipLaunchKernelGGL(ga_breed_kernel, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, s, ga->pop,
                       ga->order, ga->N, ga->G2, ga->ne, ga->nm, ga->rho, ga->seed, ga->gen, ga->next);
    GACHK(hipGetLastError());
    int rc = brkga_internal_launch(ga->dec, ga->next + (long long)ga->ne * ga->G2, BRKGA_KEYS_F32,
                                   ga->N - ga->ne, ga->fit_next + ga->ne, nullptr, s, false);
    if (rc) { g_ga_error = brkga_last_error(ga->dec); return rc; }
    hipLaunchKernelGGL(ga_elitefit_kernel, dim3((ga->ne + 255) / 256), dim3(256), 0, s, ga->fit, ga->order,
                       ga->ne, ga->fit_next);
    hipLaunchKernelGGL(ga_commit_kernel, dim3((unsigned)((std::max<long long>(nk, ga->N) + 255) / 256)),
                       dim3(256), 0, s, ga->next, ga->fit_next, nk, ga->N, ga->pop, ga->fit);
    hipLaunchKernelGGL(ga_best_kernel, dim3(1), dim3(256), 0, s, ga->fit, ga->pop, ga->N, ga->G2, ga->gen,
                       ga->best, ga->best_keys, ga->hist, ga->hist_cap);
    GACHK(hipGetLastError());
    return 0;
}

extern "C" int brkga_ga_create(brkga_handle *dec, const brkga_ga_params *p, const float *init_population,
                               brkga_ga **out) {
    if (!dec || !p || !out) { g_ga_error = "NULL argument"; return BRKGA_EINVAL; }
    *out = nullptr;
    const int n = brkga_internal_n_parts(dec);
    if (p->num_individuals < 2 || p->num_individuals > GA_SORT_MAX) { g_ga_error = "num_individuals must be in [2, 4096]"; return BRKGA_EINVAL; }
    if (p->num_elites < 1 || p->num_mutants < 0 || p->num_elites + p->num_mutants > p->num_individuals ||
        p->num_elites >= p->num_individuals) {
        g_ga_error = "need 1 <= num_elites < num_individuals and num_elites + num_mutants <= num_individuals";
        return BRKGA_EINVAL;
    }
    if (!(p->elite_prob >= 0.0 && p->elite_prob <= 1.0)) { g_ga_error = "elite_prob must be in [0, 1]"; return BRKGA_EINVAL; }
    if (p->history_capacity < 0) { g_ga_error = "history_capacity must be >= 0"; return BRKGA_EINVAL; }
    brkga_ga *ga = new (std::nothrow) brkga_ga();
    if (!ga) { g_ga_error = "host allocation failed"; return BRKGA_ENOMEM; }
    ga->dec = dec;
    ga->N = p->num_individuals;
    ga->G2 = 2 * n;
    ga->ne = p->num_elites;
    ga->nm = p->num_mutants;
    ga->rho = (float)p->elite_prob;
    ga->seed = p->seed;
    ga->hist_cap = p->history_capacity + 1;
    auto bail = [&](int rc) { brkga_ga_destroy(ga); return rc; };
    int rc;
    if (hipSetDevice(brkga_internal_device(dec)) != hipSuccess) { g_ga_error = "hipSetDevice failed"; return bail(BRKGA_EDEVICE); }
    const size_t nk = (size_t)ga->N * ga->G2;
    void *q;
    if ((rc = ga_alloc(ga, nk * 4, &q))) return bail(rc);
    ga->pop = (float *)q;
    if ((rc = ga_alloc(ga, nk * 4, &q))) return bail(rc);
    ga->next = (float *)q;
    if ((rc = ga_alloc(ga, (size_t)ga->G2 * 4, &q))) return bail(rc);
    ga->best_keys = (float *)q;
    if ((rc = ga_alloc(ga, (size_t)ga->N * 8, &q))) return bail(rc);
    ga->fit = (double *)q;
    if ((rc = ga_alloc(ga, (size_t)ga->N * 8, &q))) return bail(rc);
    ga->fit_next = (double *)q;
    if ((rc = ga_alloc(ga, 8, &q))) return bail(rc);
    ga->best = (double *)q;
    if ((rc = ga_alloc(ga, (size_t)ga->hist_cap * 8, &q))) return bail(rc);
    ga->hist = (double *)q;
    if ((rc = ga_alloc(ga, (size_t)ga->N * 4, &q))) return bail(rc);
    ga->order = (int *)q;
    if ((rc = ga_alloc(ga, 4, &q))) return bail(rc);
    ga->gen = (unsigned *)q;
    if ((rc = brkga_internal_reserve(dec, ga->N))) { g_ga_error = brkga_last_error(dec); return bail(rc); }
    // initial population: given, or drawn on the device (generation-0 stream of the hash)
    if (init_population) {
        if (hipMemcpy(ga->pop, init_population, nk * 4, hipMemcpyHostToDevice) != hipSuccess) { g_ga_error = "hipMemcpy failed"; return bail(BRKGA_EDEVICE); }
    } else {
        std::vector<float> h(nk);
        unsigned long long z = p->seed ^ 0x5deece66dull;
        for (size_t i = 0; i < nk; i++) {   // host splitmix64 stream
            z += 0x9e3779b97f4a7c15ull;
            unsigned long long x = z;
            x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
            x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
            x ^= x >> 31;
            h[i] = (float)(x >> 40) * (1.0f / 16777216.0f);
        }
        if (hipMemcpy(ga->pop, h.data(), nk * 4, hipMemcpyHostToDevice) != hipSuccess) { g_ga_error = "hipMemcpy failed"; return bail(BRKGA_EDEVICE); }
    }
    unsigned long long ov0 = 0, ov1 = 0;
    if ((rc = brkga_internal_overflows(dec, &ov0))) { g_ga_error = brkga_last_error(dec); return bail(rc); }
    hipLaunchKernelGGL(ga_init_kernel, dim3(1), dim3(1), 0, 0, ga->gen, ga->best);
    if ((rc = brkga_internal_launch(dec, ga->pop, BRKGA_KEYS_F32, ga->N, ga->fit, nullptr, 0, false))) {
        g_ga_error = brkga_last_error(dec);
        return bail(rc);
    }
    hipLaunchKernelGGL(ga_best_kernel, dim3(1), dim3(256), 0, 0, ga->fit, ga->pop, ga->N, ga->G2, ga->gen,
                       ga->best, ga->best_keys, ga->hist, ga->hist_cap);
    if (hipDeviceSynchronize() != hipSuccess) { g_ga_error = "initial population decode failed"; return bail(BRKGA_EDEVICE); }
    if ((rc = brkga_internal_overflows(dec, &ov1))) { g_ga_error = brkga_last_error(dec); return bail(rc); }
    if (ov1 != ov0) {
        g_ga_error = "a context of the initial population needed more bins than the decoder's max_bins";
        return bail(BRKGA_EOVERFLOW);
    }
    *out = ga;
    return BRKGA_OK;
}

extern "C" int brkga_ga_run(brkga_ga *ga, int32_t generations, int use_graph, void *stream) {
    if (!ga || generations < 0) { g_ga_error = "bad argument"; return BRKGA_EINVAL; }
    GACHK(hipSetDevice(brkga_internal_device(ga->dec)));
    hipStream_t s = (hipStream_t)stream;
    unsigned long long ov0 = 0;
    GACHK(hipStreamSynchronize(s));
    if (int rc = brkga_internal_overflows(ga->dec, &ov0)) { g_ga_error = brkga_last_error(ga->dec); return rc; }
    if (use_graph) {
        if (!s) { g_ga_error = "graph replay needs a non-null stream"; return BRKGA_EINVAL; }
        // capture again when the stream changed or the decoder's buffers moved
        // (a larger population or a grown bin workspace reallocates them)
        if (!ga->exec || ga->graph_stream != s || ga->graph_ws_gen != brkga_internal_ws_gen(ga->dec)) {
            if (ga->exec) { (void)hipGraphExecDestroy(ga->exec); ga->exec = nullptr; }
            if (ga->graph) { (void)hipGraphDestroy(ga->graph); ga->graph = nullptr; }
            GACHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int rc = ga_generation(ga, s);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(s, &g);
            if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
            if (ec != hipSuccess) { g_ga_error = std::string("hipStreamEndCapture: ") + hipGetErrorString(ec); return BRKGA_EDEVICE; }
            ga->graph = g;
            GACHK(hipGraphInstantiate(&ga->exec, ga->graph, nullptr, nullptr, 0));
            ga->graph_stream = s;
            ga->graph_ws_gen = brkga_internal_ws_gen(ga->dec);
        }
        for (int i = 0; i < generations; i++) GACHK(hipGraphLaunch(ga->exec, s));
    } else {
        for (int i = 0; i < generations; i++) {
            int rc = ga_generation(ga, s);
            if (rc) return rc;
        }
    }
    GACHK(hipStreamSynchronize(s));
    ga->generations_done += generations;
    // a context that ran out of bins decoded to NaN: the run's fitness is not usable
    unsigned long long ov1 = 0;
    if (int rc = brkga_internal_overflows(ga->dec, &ov1)) { g_ga_error = brkga_last_error(ga->dec); return rc; }
    if (ov1 != ov0) {
        g_ga_error = "a context needed more bins than the decoder's max_bins (create the decoder with "
                     "max_bins = n_parts for the device loop)";
        return BRKGA_EOVERFLOW;
    }
    return 0;
}

extern "C" int brkga_ga_best(const brkga_ga *ga, double *best_fitness, float *best_keys, double *history,
                             int32_t *generations_done) {
    if (!ga) { g_ga_error = "NULL handle"; return BRKGA_EINVAL; }
    GACHK(hipSetDevice(brkga_internal_device(ga->dec)));
    GACHK(hipDeviceSynchronize());
    if (best_fitness) GACHK(hipMemcpy(best_fitness, ga->best, 8, hipMemcpyDeviceToHost));
    if (best_keys) GACHK(hipMemcpy(best_keys, ga->best_keys, (size_t)ga->G2 * 4, hipMemcpyDeviceToHost));
    const int nh = std::min(ga->generations_done + 1, ga->hist_cap);
    if (history) GACHK(hipMemcpy(history, ga->hist, (size_t)nh * 8, hipMemcpyDeviceToHost));
    if (generations_done) *generations_done = ga->generations_done;
    return 0;
}

extern "C" int brkga_ga_population(const brkga_ga *ga, float *keys, double *fitness) {
    if (!ga) { g_ga_error = "NULL handle"; return BRKGA_EINVAL; }
    GACHK(hipSetDevice(brkga_internal_device(ga->dec)));
    GACHK(hipDeviceSynchronize());
    if (keys) GACHK(hipMemcpy(keys, ga->pop, (size_t)ga->N * ga->G2 * 4, hipMemcpyDeviceToHost));
    if (fitness) GACHK(hipMemcpy(fitness, ga->fit, (size_t)ga->N * 8, hipMemcpyDeviceToHost));
    return 0;
}

// Offspring pool for a screened generation: pool_mutants fresh chromosomes,
// then pool_offspring biased-crossover children of (elite, non-elite) parent
// pairs drawn from the sorted population (same rules as ga_breed_kernel).
__global__ void ga_pool_kernel(const float *pop, const int *order, int N, int G2, int ne, int pm, int po,
                               float rho, unsigned long long seed, unsigned gen, float *out) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)(pm + po) * G2) return;
    const int r = (int)(t / G2), g = (int)(t % G2);
    float v;
    if (r < pm) {
        v = ga_uniform(ga_hash(seed, gen, r, g, 4));
    } else {
        const int e = order[(int)(ga_hash(seed, gen, r, 0, 5) % (unsigned long long)ne)];
        const int o = order[ne + (int)(ga_hash(seed, gen, r, 0, 6) % (unsigned long long)(N - ne))];
        const float u = ga_uniform(ga_hash(seed, gen, r, g, 7));
        v = u < rho ? pop[(long long)e * G2 + g] : pop[(long long)o * G2 + g];
    }
    out[t] = v;
}

extern "C" int brkga_sort_fitness(const double *d_fitness, int32_t n, int32_t *d_order, void *stream) {
    if (!d_fitness || !d_order || n < 1 || n > GA_SORT_MAX) { g_ga_error = "n must be in [1, 4096]"; return BRKGA_EINVAL; }
    hipLaunchKernelGGL(ga_sort_kernel, dim3(1), dim3(GA_SORT_THREADS), 0, (hipStream_t)stream, d_fitness, n, d_order);
    GACHK(hipGetLastError());
    return 0;
}

extern "C" int brkga_breed_pool(const float *d_population, const int32_t *d_order, int32_t n_individuals,
                                int32_t n_genes, int32_t num_elites, int32_t pool_mutants, int32_t pool_offspring,
                                double elite_prob, uint64_t seed, uint32_t generation, float *d_out, void *stream) {
    if (!d_population || !d_order || !d_out || num_elites < 1 || num_elites >= n_individuals || n_genes < 2 ||
        pool_mutants < 0 || pool_offspring < 0) {
        g_ga_error = "bad argument";
        return BRKGA_EINVAL;
    }
    const long long nk = (long long)(pool_mutants + pool_offspring) * n_genes;
    if (nk == 0) return 0;
    hipLaunchKernelGGL(ga_pool_kernel, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       d_population, d_order, n_individuals, n_genes, num_elites, pool_mutants, pool_offspring,
                       (float)elite_prob, seed, generation, d_out);
    GACHK(hipGetLastError());
    return 0;
}

extern "C" const char *brkga_ga_last_error(void) { return g_ga_error.c_str(); }

extern "C" void brkga_ga_destroy(brkga_ga *ga) {
    if (!ga) return;
    (void)hipSetDevice(brkga_internal_device(ga->dec));
    (void)hipDeviceSynchronize();
    if (ga->exec) (void)hipGraphExecDestroy(ga->exec);
    if (ga->graph) (void)hipGraphDestroy(ga->graph);
    for (void *p : ga->allocs) (void)hipFree(p);
    delete ga;
}
