/*
 * brkga_mi355x.h -- C ABI of the MI355X BRKGA population decoder.
 *
 * The library decodes a population of random-key chromosomes into makespans
 * (and optionally per-part placements) on one gfx950 GPU.  It is the drop-in
 * replacement for the reference's native decoder extension:
 *
 *   reference                                       here
 *   ---------                                       ----
 *   NativeFullDecoder(packed, nb_parts, nb_machines, device)
 *     full_native_decoder.py:408-533 (ctor), packed schema :2884-2925
 *                                                   brkga_create()
 *   NativeFullDecoder::evaluate_batch(float[N,2n]) -> double[N]
 *     full_native_decoder.py:563-641                brkga_evaluate() /
 *                                                   brkga_evaluate_device()
 *   get_bin_stats / reset_bin_stats, get_profile_summary / reset_profile
 *     full_native_decoder.py:2290-2327              brkga_get_stats() /
 *                                                   brkga_reset_stats()
 *   RuntimeError("...") on bad input               brkga_last_error()
 *     full_native_decoder.py:440-533,567-573
 *   (destructor)                                    brkga_destroy()
 *
 * Conventions: every call returns 0 on success or a negative BRKGA_E* code;
 * no C++ exception crosses the ABI; host buffers belong to the caller; device
 * workspaces belong to the handle.  A handle is not re-entrant: calls on one
 * handle must not overlap (one handle per thread), and the device work of its
 * calls must not overlap either: every launch through a handle uses the same
 * workspace, so issue them on one stream or order the streams with events.
 */
#ifndef BRKGA_MI355X_H
#define BRKGA_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BRKGA_OK 0
#define BRKGA_EINVAL (-1)      /* bad argument or problem data */
#define BRKGA_EDEVICE (-2)     /* HIP runtime error */
#define BRKGA_ENOMEM (-3)      /* device allocation failed */
#define BRKGA_EOVERFLOW (-4)   /* a context needed more bins than the workspace holds */
#define BRKGA_ESTATE (-5)      /* call not valid in the handle's state */

#define BRKGA_KEYS_F32 0
#define BRKGA_KEYS_F64 1

#define BRKGA_MAX_MACHINES 64   /* printer machines per problem at most */

/* Infeasible chromosomes decode to this makespan (placement.py:9). */
#define BRKGA_INFEASIBLE 1e16

/*
 * Problem data, structure of arrays.  Same content as the reference's packed
 * problem (full_native_decoder.py:2788-2926 _pack_problem_data) minus the
 * part spectra: part types are numbered 0..n_types-1, rotation r of type t is
 * row rot_base[t]+r of the rotation table, its mask is
 * masks[mask_off[row] .. + rot_h*rot_w] row-major {0,1} bytes.
 * Plates: plate_h rows (the reference's bin_length, L) by plate_w columns
 * (bin_width, W).  proc_time / proc_time_height are [n_machines][n_types].
 */
typedef struct brkga_problem {
    int32_t n_parts;      /* instance slots n; chromosomes have 2n keys */
    int32_t n_machines;   /* M >= 1; M == 1 puts every part on machine 0 */
    int32_t n_types;
    int32_t n_rots;
    const double *thresholds;        /* [M-1], ascending; t/M in the reference */
    const int32_t *slot_type;        /* [n_parts] type of instance slot j */
    const int32_t *type_id;          /* [n_types] external part id (for logs) */
    const double *area;              /* [n_types] part area, > 0 */
    const int32_t *nrot;             /* [n_types] 1..4 */
    const int32_t *best_rot;         /* [n_types] rotation used to open a bin */
    const int32_t *rot_base;         /* [n_types] */
    const int32_t *rot_h, *rot_w;    /* [n_rots] */
    const int64_t *mask_off;         /* [n_rots] */
    const uint8_t *masks;
    const int32_t *plate_h, *plate_w;              /* [M] */
    const double *setup;                           /* [M] setup time per bin */
    const double *proc_time, *proc_time_height;    /* [M * n_types] */
} brkga_problem;

typedef struct brkga_options {
    int32_t device;          /* HIP device ordinal */
    int32_t wgs_per_cu;      /* persistent workgroups per compute unit: 0 = automatic (a
                                decode of at least two rounds of contexts at three per CU
                                runs three per CU when the LDS allows, others two, or one
                                when two do not fit); 1 or 2 = always that; 3 = three for
                                every decode (BRKGA_EINVAL when it does not fit) */
    int32_t max_bins;        /* bins per context in the workspace (0 = min(n_parts, 32));
                                brkga_evaluate grows it when a context needs more */
    int32_t dense;           /* 1: correlation over every placement row, no early exit
                                (same results; used to measure the kernel at full shape) */
    int32_t tiles_per_pass;  /* cap on correlation x-tiles per pass for every class of the
                                tile table: 1, 2, 4 or 8 (0 = per-width default) */
    int32_t mfma_format;     /* operands of the collision correlation's MFMAs:
                                BRKGA_MFMA_FP4 (0, default) or BRKGA_MFMA_INT8 (1); the
                                counts are exact either way and the results identical */
    int32_t screen_rows;     /* mask rows of the screening pass for every class of the tile
                                table: 0 = default (6), -1 = no screening, 1..16 (see
                                brkga_set_screen); same results either way */
    int32_t reserved;
} brkga_options;

#define BRKGA_MFMA_FP4 0    /* e2m1 operands 0 / 0.5 / 1, float32 accumulation, K = 64 per MFMA */
#define BRKGA_MFMA_INT8 1   /* int8 operands, int32 accumulation, K = 32 per MFMA */

typedef struct brkga_stats {
    int64_t calls;           /* evaluate calls */
    int64_t chromosomes;     /* chromosomes decoded */
    int64_t contexts;        /* (chromosome, machine) contexts decoded */
    int64_t placements;      /* parts placed */
    int64_t bins_opened;
    int64_t tests;           /* collision tests (bin, rotation) run */
    int64_t row_groups;      /* 32-row correlation groups evaluated */
    int64_t mfma_ops;        /* int8 multiply-adds issued x 2 (MFMA work done) */
    int64_t algo_ops;        /* sum over tests of 2*h*w*(H-h+1)*(W-w+1) */
    int64_t overflows;       /* contexts that exceeded max_bins */
    double kernel_ms;        /* device time of the decode kernels (HIP events) */
    double last_kernel_ms;   /* decode kernel time of the last call */
    /* Where a workgroup's time goes, summed over workgroups, in microseconds of
     * the GPU's constant-rate wall clock: [0] sequence decode, [1] vacancy
     * filter, [2] band staging, [3] correlation (full passes), [4] screening pass,
     * [5] stamp + free-run update, [6] new-bin clearing, [7] makespan. */
    double phase_us[8];
    double context_us_max;   /* longest single context */
    double context_us_sum;   /* all contexts */
    /* Bin statistics, the reference's get_bin_stats (full_native_decoder.py:2286-2312,
     * per machine: peak bins per solution, total bins, solutions, growth events). */
    int64_t growths;         /* workspace growths: brkga_evaluate doubled max_bins after an overflow */
    int32_t max_bins;        /* bins per context the workspace holds now */
    int32_t n_workgroups;    /* persistent workgroups of the last decode launch */
    int64_t machine_contexts[BRKGA_MAX_MACHINES];    /* contexts decoded on machine m (feasible or not) */
    int64_t machine_bins[BRKGA_MAX_MACHINES];        /* bins the contexts of machine m opened */
    int32_t machine_peak_bins[BRKGA_MAX_MACHINES];   /* most bins one context of machine m opened */
} brkga_stats;

typedef struct brkga_handle brkga_handle;

/* Build a decoder for one problem on one device.  Validates the problem
 * (shapes, binary masks, positive areas, plate sizes, rotation/plate fit) and
 * uploads the masks and the per-rotation correlation operands. */
int brkga_create(const brkga_problem *problem, const brkga_options *options,
                 brkga_handle **out);

/* Decode n_chrom chromosomes held in HOST memory.  keys: [n_chrom][n_genes]
 * with n_genes == 2*n_parts, float32 or float64 (key_dtype).  out_makespan:
 * [n_chrom] host doubles.  opt_log (nullable): [n_chrom][n_parts][6] int32
 * rows (machine, part id, bin, col, row, rotation) in machine order then
 * placement order; rows of infeasible contexts are -1.  stream: hipStream_t or
 * NULL for the null stream.  Synchronous: returns after the results are in
 * host memory. */
int brkga_evaluate(brkga_handle *h, const void *keys, int key_dtype, int64_t n_chrom,
                   int64_t n_genes, double *out_makespan, int32_t *opt_log, void *stream);

/* Same with DEVICE pointers (e.g. from torch tensors).  Enqueued on `stream`;
 * returns after enqueueing unless check_overflow is nonzero, in which case it
 * synchronises the stream and returns BRKGA_EOVERFLOW if a context ran out of
 * bins (then no output may be used). */
int brkga_evaluate_device(brkga_handle *h, const void *d_keys, int key_dtype, int64_t n_chrom,
                          int64_t n_genes, double *d_out_makespan, int32_t *d_opt_log,
                          void *stream, int check_overflow);

/* Allocation-free variant of brkga_evaluate_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results; it calls no allocating routine.  Its per-call
 * buffers live in one static device array per GPU (about 1 MiB, held until
 * the process ends) shared by every handle on that GPU, so calls that overlap
 * on one GPU race: issue them one at a time (stream-ordered).  Limits:
 * n_chrom * n_machines <= BRKGA_NOALLOC_MAX_CONTEXTS, else BRKGA_EINVAL; the
 * handle's bin workspace (allocated by brkga_create) is not grown, so create
 * the handle with a max_bins large enough (n_parts always is) -- with
 * check_overflow a context that ran out of bins gives BRKGA_EOVERFLOW, without
 * it a NaN makespan.  Decode time is not added to brkga_stats.kernel_ms. */
#define BRKGA_NOALLOC_MAX_CONTEXTS 65536
int brkga_evaluate_device_noalloc(brkga_handle *h, const void *d_keys, int key_dtype, int64_t n_chrom,
                                  int64_t n_genes, double *d_out_makespan, int32_t *d_opt_log,
                                  void *stream, int check_overflow);

int brkga_get_stats(const brkga_handle *h, brkga_stats *out);
int brkga_reset_stats(brkga_handle *h);

/*
 * Tile shapes of the collision passes (the autotuner's knobs; no counterpart
 * in the reference, whose FFT has none).  Tables are [n_machines][BRKGA_TILE_CLASSES],
 * indexed by machine and the mask's Toeplitz chunk count nc = ceil((w + 31) / 32)
 * (the 32-column chunks a 32-placement tile spans; larger counts share the last
 * class):
 *   tiles_per_pass: cap on 32-column x-tiles one MFMA pass covers, 1, 2, 4 or 8
 *                   (0 = the default: 8 up to nc = 5, i.e. masks up to 129 columns,
 *                   4 up to nc = 8 (225 columns), else 2);
 *   window_tiles:   cap on x-tiles staged per LDS window, 1..64 (0 = 64).
 * Every setting gives identical results.  Either table pointer may be NULL
 * (left unchanged).  brkga_set_tiles waits for the device to be idle; it must
 * not race with launches on the handle.  brkga_get_tiles returns the
 * effective values.
 */
#define BRKGA_TILE_CLASSES 16
int brkga_set_tiles(brkga_handle *h, const int32_t *tiles_per_pass, const int32_t *window_tiles);
int brkga_get_tiles(const brkga_handle *h, int32_t *tiles_per_pass, int32_t *window_tiles);

/*
 * Screening pass of the collision test (an autotuner knob; no counterpart in
 * the reference).  Before the full correlation of a staged x-window, the
 * MFMAs correlate only `screen_rows` evenly spaced mask rows over every tile;
 * the overlap count over a subset of the rows is zero wherever the full count
 * is, so a tile with no zero there holds no placement and the full passes skip
 * it.  Table [n_machines][BRKGA_TILE_CLASSES] as for brkga_set_tiles, entries
 * 0 (no screening) to BRKGA_MAX_SCREEN_ROWS; a group whose mask meets fewer
 * than 2 * screen_rows occupied rows is not screened.  Every setting gives
 * identical results.  brkga_set_screen waits for the device to be idle.
 */
#define BRKGA_MAX_SCREEN_ROWS 16
int brkga_set_screen(brkga_handle *h, const int32_t *screen_rows);
int brkga_get_screen(const brkga_handle *h, int32_t *screen_rows);

/*
 * Greedy initial solution (replaces BRKGA_alg3.py:445-603 and the batch class
 * of binClassInitialSol.py:41-174).  Parts are taken by decreasing height
 * (type_height [n_types]), then decreasing area, ties in instance order; each
 * goes to the new or existing batch, on any machine whose plate fits it, that
 * keeps the running makespan lowest (the reference's tie rules).  Writes the
 * seed chromosome (float32 [2n]: per machine, its parts in batch order get
 * sequence keys i/k and the machine key machine_keys[m], which must lie in
 * machine m's threshold bucket -- the reference draws it with random.uniform),
 * the greedy's makespan, and optionally the placements (int32 [n][6]:
 * machine, part id, batch, col, bottom row, rotation) in greedy order.  All
 * buffers are host memory; the call is synchronous on `stream` (nullable).
 * BRKGA_EINVAL when a part fits no machine.
 */
int brkga_initial_solution(brkga_handle *h, const double *type_height, const double *machine_keys,
                           float *out_chromosome, double *out_makespan, int32_t *out_log, void *stream);

/* Message of the last failed call on h (or of the last failed brkga_create
 * when h is NULL).  Valid until the next call on the same handle. */
const char *brkga_last_error(const brkga_handle *h);

void brkga_destroy(brkga_handle *h);

/*
 * Device-resident BRKGA generation loop.  The reference runs the loop on the
 * host (BRKGA.fit, BRKGA_alg3.py:210-281: partition :178-186, mating
 * :188-205, mutants :207-208); here the population and its fitness stay in
 * HBM and each generation (sort, breed, decode, keep elites' fitness, best
 * tracking) is one captured hipGraph replay.  Random numbers come from a
 * counter-based hash of (seed, generation, row, gene), not NumPy's stream.
 * The decoder handle must have max_bins large enough for every context
 * (max_bins = n_parts is always enough): a graph replay cannot grow it, and
 * brkga_ga_create / brkga_ga_run return BRKGA_EOVERFLOW when a context ran out
 * of bins.  The captured graph is captured again when the decoder's device
 * buffers have moved since (a larger population decoded through the same
 * handle, or a grown bin workspace), so the decoder may be used between runs.
 */
typedef struct brkga_ga_params {
    int32_t num_individuals;   /* population N, 2..4096 */
    int32_t num_elites;        /* kept unchanged each generation */
    int32_t num_mutants;       /* fresh uniform chromosomes each generation */
    int32_t history_capacity;  /* generations whose minimum fitness is recorded */
    double elite_prob;         /* probability an offspring gene comes from the elite parent */
    uint64_t seed;
} brkga_ga_params;

typedef struct brkga_ga brkga_ga;

/* init_population: [N][2n] float32 host keys, or NULL to draw them.  Decodes
 * the initial population before returning. */
int brkga_ga_create(brkga_handle *decoder, const brkga_ga_params *params,
                    const float *init_population, brkga_ga **out);
/* Run `generations` generations on `stream` (graph replays when use_graph),
 * then synchronise the stream. */
int brkga_ga_run(brkga_ga *ga, int32_t generations, int use_graph, void *stream);
/* Best fitness so far, its keys [2n], per-generation minima [done+1]
 * (generation 0 = initial population); each pointer may be NULL. */
int brkga_ga_best(const brkga_ga *ga, double *best_fitness, float *best_keys, double *history,
                  int32_t *generations_done);
/* Current population [N][2n] and fitness [N]; each pointer may be NULL. */
int brkga_ga_population(const brkga_ga *ga, float *keys, double *fitness);
const char *brkga_ga_last_error(void);
void brkga_ga_destroy(brkga_ga *ga);

/* Building blocks of a screened generation (device pointers, enqueued on
 * stream).  brkga_sort_fitness: d_order = indices of d_fitness [n <= 4096]
 * ascending (ties by index), as the reference's partition orders elites
 * (BRKGA_alg3.py:178-186).  brkga_breed_pool: pool_mutants uniform
 * chromosomes followed by pool_offspring biased-crossover children
 * (BRKGA_alg3.py:188-208) of the sorted population, keyed by (seed, generation). */
int brkga_sort_fitness(const double *d_fitness, int32_t n, int32_t *d_order, void *stream);
int brkga_breed_pool(const float *d_population, const int32_t *d_order, int32_t n_individuals,
                     int32_t n_genes, int32_t num_elites, int32_t pool_mutants, int32_t pool_offspring,
                     double elite_prob, uint64_t seed, uint32_t generation, float *d_out, void *stream);

/*
 * Surrogate: BF16 MLP with FP32 accumulation predicting a chromosome's
 * (normalised) makespan, trained online on exact decodes.  No counterpart in
 * the reference; it only ranks oversized offspring pools so that fewer
 * chromosomes need an exact decode.  Input features of a chromosome: 4n slot
 * features -- its 2n keys as the decoder reads them (the sequence position of
 * each slot divided by n, and the machine bucket of each slot's machine key,
 * thresholds as in brkga_problem, divided by max(n_machines-1, 1)), then
 * slot_area and slot_height of its parts in ascending sequence-key order
 * (ties by slot).  With machine_terms ([n_machines][3][n_parts] float32,
 * nullable) the slot features are multiplied by 1/8 and 4 per machine m
 * follow: the float32 sums of machine_terms[m][0][j] and machine_terms[m][1][j]
 * over the slots j the chromosome sends to machine m (64 lane sums of slots
 * l, l + 64, ..., then a butterfly over xor 32, 16, ..., 1),
 * their count times n_machines / n_parts, and the largest
 * machine_terms[m][2][j] among them (0 for none); the Python layer passes the
 * processing time, the area over the plate area and the height-driven
 * processing time of each part on each machine, scaled to O(1).  Layers: F ->
 * hidden -> hidden -> 1, ReLU, F = 4n (+ 4 n_machines).  Parameters (float32
 * master copy) are laid out W1 [Fp][hidden] (rows >= F are zero), b1 [hidden],
 * W2 [hidden][hidden], b2 [hidden], w3 [hidden], b3 [1], with Fp = F rounded
 * up to 64.  1 <= n_parts <= 4096, n_machines <= 256.  d_* pointers are device
 * memory.
 */
typedef struct brkga_mlp brkga_mlp;
int brkga_mlp_create(int32_t n_parts, int32_t n_machines, const double *thresholds,
                     int32_t hidden, const float *slot_area, const float *slot_height,
                     const float *machine_terms, uint64_t seed, float learning_rate, int32_t device,
                     brkga_mlp **out);
/* d_keys [B][2n] float32 -> d_pred [B] float32 */
int brkga_mlp_predict(brkga_mlp *net, const float *d_keys, int64_t B, float *d_pred, void *stream);
/* Mean-squared-error gradient of one batch into d_grad [n_params] (nullable). */
int brkga_mlp_gradients(brkga_mlp *net, const float *d_keys, const float *d_target, int64_t B,
                        float *d_grad, void *stream);
/* One Adam step on one batch; *h_loss (nullable) receives the batch loss. */
int brkga_mlp_train_step(brkga_mlp *net, const float *d_keys, const float *d_target, int64_t B,
                         float *h_loss, void *stream);
int brkga_mlp_num_params(const brkga_mlp *net, int64_t *n_params, int32_t *n_features_padded);
int brkga_mlp_get_params(const brkga_mlp *net, float *h_params);
int brkga_mlp_set_params(brkga_mlp *net, const float *h_params);
/* Feature rows the network sees: d_x [B][Fp] bfloat16 bits. */
int brkga_mlp_features(brkga_mlp *net, const float *d_keys, int64_t B, uint16_t *d_x, void *stream);
const char *brkga_mlp_last_error(void);
void brkga_mlp_destroy(brkga_mlp *net);

#ifdef __cplusplus
}
#endif
#endif /* BRKGA_MI355X_H */
