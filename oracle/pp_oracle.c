/*
 * CPU oracle for the BRKGA population decoder.  TEST INFRASTRUCTURE ONLY:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker / CPU baseline, never as the
 * product path.
 *
 * It restates, in plain C, the reference's sequential placement procedure
 * ("PP", the reference's own ground truth):
 *
 *   machine assignment + sequence      placement.py:109-133 (M<=1: all parts
 *                                      on machine 0, full_native_decoder.py:961-964)
 *   per-part infeasibility             placement.py:46-53
 *   first-fit over open bins           placement.py:55-64 (area check :57-59)
 *   vacancy filter per rotation        binClassNew.py:66-88, numba_utils.py:131-157
 *   collision test, bottom-left pick   collision_backend.py:109-128 (cropped
 *                                      valid region, max row with a zero,
 *                                      min col in that row)
 *   density -> row -> col choice       binClassNew.py:101-117
 *   insert + enclosure + vacancy rows  binClassNew.py:147-174, numba_utils.py:161-188
 *   new bin at (0, H-1), best rotation placement.py:66-83
 *   makespan = sum over bins of        placement.py:85-89, 147-156
 *     (proc + proc_height + setup), max over machines
 *
 * The reference decides collisions with round(irfft2(...)) == 0; here the
 * overlap count is computed exactly.  For binary masks the count at a
 * placement is zero iff no 1-cell of the mask lands on a 1-cell of the grid,
 * which the fast path checks run by run with a "free cells to the right"
 * table; oracle_overlap_count() evaluates the plain double sum and the tests
 * cross-check the two.
 *
 * Density comparison: inside one bin the numerator (bin area + part area) is
 * fixed, so higher density <=> smaller new enclosure length; distinct integer
 * lengths give distinct float64 quotients, so comparing the integers is the
 * same decision without the division.  Ties keep the earlier rotation
 * (strict '>' in binClassNew.py:109-115).
 *
 * Sequence ties (equal keys) are broken by slot index (a stable sort); the
 * reference's np.argsort order for equal keys is not specified.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int32_t n_parts, n_machines, n_types, n_rots;
    const double *thresholds;      /* [n_machines-1] */
    const int32_t *slot_type;      /* [n_parts] */
    const int32_t *type_id;        /* [n_types] */
    const double *area;            /* [n_types] */
    const int32_t *nrot, *best_rot, *rot_base;  /* [n_types] */
    const int32_t *rot_h, *rot_w;  /* [n_rots] */
    const int64_t *mask_off;       /* [n_rots] */
    const uint8_t *masks;
    const int32_t *plate_h, *plate_w;  /* [n_machines] */
    const double *setup;           /* [n_machines] */
    const double *proc_time, *proc_time_height;  /* [n_machines * n_types] */
} pp_problem;

typedef struct {
    int64_t tests;        /* collision tests (bin, rotation) run */
    int64_t positions;    /* candidate placements examined */
    int64_t bins_opened;
    int64_t vacancy_rejects;
} pp_stats;

typedef struct { int32_t start, len; } run_t;

/* per-rotation run tables, built once per call */
typedef struct {
    run_t *runs;        /* all runs */
    int32_t *row_first; /* [sum h + n_rots] index of first run of each row */
    int64_t *row_base;  /* [n_rots] index into row_first */
    int32_t *dens;      /* [sum h] longest run per row */
    int64_t *dens_off;  /* [n_rots] */
} rot_tables;

static void build_tables(const pp_problem *p, rot_tables *t) {
    int64_t tot_h = 0, tot_runs = 0;
    for (int r = 0; r < p->n_rots; r++) {
        tot_h += p->rot_h[r];
        const uint8_t *m = p->masks + p->mask_off[r];
        for (int a = 0; a < p->rot_h[r]; a++)
            for (int b = 0; b < p->rot_w[r]; b++)
                if (m[a * p->rot_w[r] + b] && (b == 0 || !m[a * p->rot_w[r] + b - 1])) tot_runs++;
    }
    t->runs = (run_t *)malloc(sizeof(run_t) * (size_t)(tot_runs + 1));
    t->row_first = (int32_t *)malloc(sizeof(int32_t) * (size_t)(tot_h + p->n_rots + 1));
    t->row_base = (int64_t *)malloc(sizeof(int64_t) * (size_t)(p->n_rots + 1));
    t->dens = (int32_t *)malloc(sizeof(int32_t) * (size_t)(tot_h + 1));
    t->dens_off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(p->n_rots + 1));
    int64_t nr = 0, rf = 0, dh = 0;
    for (int r = 0; r < p->n_rots; r++) {
        const int h = p->rot_h[r], w = p->rot_w[r];
        const uint8_t *m = p->masks + p->mask_off[r];
        t->row_base[r] = rf;
        t->dens_off[r] = dh;
        for (int a = 0; a < h; a++) {
            t->row_first[rf++] = (int32_t)nr;
            int best = 0, cur = 0;
            for (int b = 0; b < w; b++) {
                if (m[a * w + b]) {
                    if (b == 0 || !m[a * w + b - 1]) { t->runs[nr].start = b; t->runs[nr].len = 0; nr++; }
                    t->runs[nr - 1].len++;
                    cur++;
                    if (cur > best) best = cur;
                } else {
                    cur = 0;
                }
            }
            t->dens[dh++] = best;
        }
        t->row_first[rf++] = (int32_t)nr;   /* sentinel of the last row */
    }
}

static void free_tables(rot_tables *t) {
    free(t->runs); free(t->row_first); free(t->row_base); free(t->dens); free(t->dens_off);
}

typedef struct {
    uint8_t *grid;      /* H*W */
    uint16_t *freer;    /* H*W: free cells from (y,x) to the right, inclusive */
    int32_t *vac;       /* H: longest free run of the row */
    double area, proc_time, proc_time_height;
    int32_t min_row, max_row, enc_len;
} pp_bin;

/* numba_utils.py:161-188 plus the free-run table of the same rows */
static void refresh_rows(pp_bin *b, int W, int y0, int y1) {
    for (int y = y0; y < y1; y++) {
        const uint8_t *g = b->grid + (int64_t)y * W;
        uint16_t *fr = b->freer + (int64_t)y * W;
        int run = 0, best = 0;
        for (int x = W - 1; x >= 0; x--) {
            run = g[x] ? 0 : run + 1;
            fr[x] = (uint16_t)run;
            if (run > best) best = run;
        }
        b->vac[y] = best;
    }
}

/* binClassNew.py:147-174 (insert at top-left (y0, x0)) */
static void insert_part(pp_bin *b, int H, int W, int y0, int x0, const uint8_t *m, int h, int w,
                        double part_area) {
    b->area += part_area;
    for (int a = 0; a < h; a++) {
        uint8_t *g = b->grid + (int64_t)(y0 + a) * W + x0;
        for (int c = 0; c < w; c++) g[c] = (uint8_t)(g[c] + m[a * w + c]);
    }
    if (y0 < b->min_row) b->min_row = y0;
    if (y0 + h - 1 > b->max_row) b->max_row = y0 + h - 1;
    b->enc_len = (b->min_row > b->max_row) ? 0 : H - b->min_row;
    refresh_rows(b, W, y0, y0 + h);
}

/* numba_utils.py:131-157 */
static int vacancy_fit(const int32_t *vac, int H, const int32_t *den, int h) {
    if (h > H) return 0;
    for (int s = 0; s + h <= H; s++) {
        int ok = 1;
        for (int i = 0; i < h; i++) if (vac[s + i] < den[i]) { ok = 0; break; }
        if (ok) return 1;
    }
    return 0;
}

/* collision_backend.py:109-128: bottom-most placement row with a zero
 * overlap, then the left-most column of that row.  Returns 1 and the
 * part's top-left (y, x) when found. */
static int bottom_left(const pp_bin *b, int H, int W, const rot_tables *t, int rot, int h, int w,
                       int *out_y, int *out_x, int64_t *positions) {
    const int32_t *den = t->dens + t->dens_off[rot];
    const int32_t *rowf = t->row_first + t->row_base[rot];
    for (int y = H - h; y >= 0; y--) {
        int ok = 1;   /* per-row vacancy necessary condition */
        for (int a = 0; a < h; a++) if (b->vac[y + a] < den[a]) { ok = 0; break; }
        if (!ok) continue;
        for (int x = 0; x + w <= W; x++) {
            (*positions)++;
            int clear = 1;
            for (int a = 0; a < h && clear; a++) {
                const uint16_t *fr = b->freer + (int64_t)(y + a) * W + x;
                for (int k = rowf[a]; k < rowf[a + 1]; k++) {
                    if (fr[t->runs[k].start] < t->runs[k].len) { clear = 0; break; }
                }
            }
            if (clear) { *out_y = y; *out_x = x; return 1; }
        }
    }
    return 0;
}

/* machine of a key: placement.py:121-127 (thresholds compared in float64) */
static int machine_of(double mv, const double *thr, int M) {
    if (M <= 1) return 0;
    if (mv <= thr[0]) return 0;
    for (int m = 1; m < M - 1; m++) if (mv > thr[m - 1] && mv <= thr[m]) return m;
    return M - 1;
}

typedef struct { double key; int32_t slot; } keyslot;

static int cmp_keyslot(const void *a, const void *b) {
    const keyslot *x = (const keyslot *)a, *y = (const keyslot *)b;
    if (x->key < y->key) return -1;
    if (x->key > y->key) return 1;
    return (x->slot > y->slot) - (x->slot < y->slot);
}

/* Decode one (chromosome, machine) context. Returns the machine makespan, or
 * -1 when the context is infeasible.  log rows (6 int32) are written for each
 * placed part: machine, part id, bin, col, row (bottom), rotation. */
static double decode_context(const pp_problem *p, const rot_tables *t, const double *keys,
                             int m, int32_t *log, int *log_n, pp_stats *st,
                             keyslot *ks, pp_bin *bins, int max_bins) {
    const int n = p->n_parts, H = p->plate_h[m], W = p->plate_w[m], P = p->n_types;
    int cnt = 0;
    for (int j = 0; j < n; j++) {
        if (machine_of(keys[n + j], p->thresholds, p->n_machines) == m) {
            ks[cnt].key = keys[j];
            ks[cnt].slot = j;
            cnt++;
        }
    }
    qsort(ks, (size_t)cnt, sizeof(keyslot), cmp_keyslot);
    int nb = 0;
    const double plate_area = (double)H * (double)W;
    for (int i = 0; i < cnt; i++) {
        const int ty = p->slot_type[ks[i].slot];
        const int r0 = p->rot_base[ty];
        const int h0 = p->rot_h[r0], w0 = p->rot_w[r0];
        if ((h0 > H || w0 > W) && (w0 > H || h0 > W)) return -1.0;   /* placement.py:47-53 */
        const double parea = p->area[ty];
        int placed = 0, pb = -1, py = 0, px = 0, prot = 0;
        for (int bi = 0; bi < nb && !placed; bi++) {
            pp_bin *b = &bins[bi];
            if (b->area + parea > plate_area) continue;              /* placement.py:57-59 */
            int best_len = 0, best_row = -1, best_x = 0, best_rot = -1;
            for (int r = 0; r < p->nrot[ty]; r++) {
                const int rg = r0 + r, h = p->rot_h[rg], w = p->rot_w[rg];
                if (h > H || w > W) continue;
                if (!vacancy_fit(b->vac, H, t->dens + t->dens_off[rg], h)) { st->vacancy_rejects++; continue; }
                st->tests++;
                int y, x;
                if (!bottom_left(b, H, W, t, rg, h, w, &y, &x, &st->positions)) continue;
                const int row = y + h - 1;
                const int new_len = b->enc_len > H - y ? b->enc_len : H - y;
                if (best_rot < 0 || new_len < best_len ||
                    (new_len == best_len && row > best_row) ||
                    (new_len == best_len && row == best_row && x < best_x)) {
                    best_len = new_len; best_row = row; best_x = x; best_rot = r;
                }
            }
            if (best_rot >= 0) {
                const int rg = r0 + best_rot;
                const int h = p->rot_h[rg], w = p->rot_w[rg];
                insert_part(b, H, W, best_row - h + 1, best_x, p->masks + p->mask_off[rg], h, w, parea);
                b->proc_time += p->proc_time[(int64_t)m * P + ty];
                const double pth = p->proc_time_height[(int64_t)m * P + ty];
                if (pth > b->proc_time_height) b->proc_time_height = pth;
                placed = 1; pb = bi; py = best_row; px = best_x; prot = best_rot;
            }
        }
        if (!placed) {
            if (nb >= max_bins) return -2.0;
            pp_bin *b = &bins[nb];
            memset(b->grid, 0, (size_t)H * W);
            for (int y = 0; y < H; y++) b->vac[y] = W;
            for (int y = 0; y < H; y++)
                for (int x = 0; x < W; x++) b->freer[(int64_t)y * W + x] = (uint16_t)(W - x);
            b->area = 0.0; b->proc_time = 0.0; b->proc_time_height = 0.0;
            b->min_row = H; b->max_row = -1; b->enc_len = 0;
            const int br = p->best_rot[ty], rg = r0 + br;
            const int h = p->rot_h[rg], w = p->rot_w[rg];
            insert_part(b, H, W, H - h, 0, p->masks + p->mask_off[rg], h, w, parea);
            b->proc_time += p->proc_time[(int64_t)m * P + ty];
            const double pth = p->proc_time_height[(int64_t)m * P + ty];
            if (pth > b->proc_time_height) b->proc_time_height = pth;
            st->bins_opened++;
            pb = nb; py = H - 1; px = 0; prot = br;
            nb++;
        }
        if (log) {
            int32_t *e = log + 6 * (*log_n);
            e[0] = m; e[1] = p->type_id[ty]; e[2] = pb; e[3] = px; e[4] = py; e[5] = prot;
            (*log_n)++;
        }
    }
    double total = 0.0;
    for (int bi = 0; bi < nb; bi++)
        total = total + ((bins[bi].proc_time + bins[bi].proc_time_height) + p->setup[m]);
    return total;
}

/* Decode a population. keys: [n_chrom][2*n_parts] float32 (key_is_f64=0) or
 * float64.  out: makespan per chromosome (1e16 when infeasible).
 * log (nullable): [n_chrom][n_parts][6] int32, unused rows set to -1.
 * Returns 0 on success, -1 on a bad argument, -2 when max_bins was exceeded. */
int oracle_decode(const pp_problem *p, const void *keys, int key_is_f64, int64_t n_chrom,
                  double *out, int32_t *log, int n_threads, pp_stats *stats_out) {
    if (!p || !keys || !out || n_chrom < 0 || p->n_parts <= 0 || p->n_machines <= 0) return -1;
    rot_tables t;
    build_tables(p, &t);
    int maxH = 0, maxW = 0;
    for (int m = 0; m < p->n_machines; m++) {
        if (p->plate_h[m] > maxH) maxH = p->plate_h[m];
        if (p->plate_w[m] > maxW) maxW = p->plate_w[m];
    }
    const int n = p->n_parts;
    const int max_bins = n;   /* one bin per part at most */
    int status = 0;
    pp_stats total = {0, 0, 0, 0};
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
#pragma omp parallel
    {
        pp_stats st = {0, 0, 0, 0};
        keyslot *ks = (keyslot *)malloc(sizeof(keyslot) * (size_t)n);
        double *kd = (double *)malloc(sizeof(double) * (size_t)(2 * n));
        pp_bin *bins = (pp_bin *)calloc((size_t)max_bins, sizeof(pp_bin));
        for (int b = 0; b < max_bins; b++) {
            bins[b].grid = (uint8_t *)malloc((size_t)maxH * maxW);
            bins[b].freer = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)maxH * maxW);
            bins[b].vac = (int32_t *)malloc(sizeof(int32_t) * (size_t)maxH);
        }
#pragma omp for schedule(dynamic, 1)
        for (int64_t c = 0; c < n_chrom; c++) {
            for (int j = 0; j < 2 * n; j++)
                kd[j] = key_is_f64 ? ((const double *)keys)[c * 2 * n + j]
                                   : (double)((const float *)keys)[c * 2 * n + j];
            int32_t *lg = log ? log + c * n * 6 : NULL;
            if (lg) for (int k = 0; k < n * 6; k++) lg[k] = -1;
            int ln = 0;
            double worst = 0.0;
            int feasible = 1;
            for (int m = 0; m < p->n_machines; m++) {
                double ms = decode_context(p, &t, kd, m, lg, &ln, &st, ks, bins, max_bins);
                if (ms == -2.0) {
#pragma omp atomic write
                    status = -2;
                }
                if (ms < 0.0) { feasible = 0; break; }
                if (ms > worst) worst = ms;
            }
            out[c] = feasible ? worst : 1e16;
        }
        for (int b = 0; b < max_bins; b++) { free(bins[b].grid); free(bins[b].freer); free(bins[b].vac); }
        free(bins); free(kd); free(ks);
#pragma omp critical
        {
            total.tests += st.tests; total.positions += st.positions;
            total.bins_opened += st.bins_opened; total.vacancy_rejects += st.vacancy_rejects;
        }
    }
    if (stats_out) *stats_out = total;
    free_tables(&t);
    return status;
}

/* Plain double sum of the overlap at bottom-row r, right-col c
 * (collision_backend.py:115 before rounding). */
int64_t oracle_overlap_count(const uint8_t *grid, int H, int W, const uint8_t *mask, int h, int w,
                             int r, int c) {
    int64_t s = 0;
    for (int a = 0; a < h; a++)
        for (int b = 0; b < w; b++)
            s += (int64_t)mask[a * w + b] * grid[(int64_t)(r - h + 1 + a) * W + (c - w + 1 + b)];
    return s;
}

/* Brute-force bottom-left search by full overlap counts (no pruning):
 * returns 1 and (col, row) as collision_backend.py:109-128 reports them. */
int oracle_bottom_left_bruteforce(const uint8_t *grid, int H, int W, const uint8_t *mask, int h,
                                  int w, int *out_col, int *out_row) {
    for (int r = H - 1; r >= h - 1; r--)
        for (int c = w - 1; c < W; c++)
            if (oracle_overlap_count(grid, H, W, mask, h, w, r, c) == 0) {
                *out_col = c - (w - 1);
                *out_row = r;
                return 1;
            }
    return 0;
}

/* ---------------------------------------------------------------------------
 * Greedy initial solution: BRKGA_alg3.py:445-603 with the batch class of
 * binClassInitialSol.py:41-174.
 *
 * Parts are taken in `order` (slots sorted by decreasing height, then
 * decreasing polygon area, ties in instance order).  For each part, every
 * machine whose plate fits rotation 0 in either orientation is tried: a new
 * batch (running machine makespan + setup + processing + height time, the
 * part at (0, L-1) in its best rotation), then every existing batch of that
 * machine in order with can_insert (vacancy filter, bottom-left collision
 * test and density -> row -> col choice over the rotations, no area check).
 * The candidate minimising max(current worst, machine makespan) wins, "<="
 * letting later candidates win ties, except that a second existing batch of
 * the same machine does not replace an equal one (placedInExist).
 *
 * The chromosome: the parts of machine m, in batch order then insertion
 * order, are mapped back to the first unused slot holding that part id; their
 * machine keys are machine_keys[m] (the reference draws one random.uniform
 * per machine inside the machine's threshold bucket) and their sequence keys
 * i * (1 / k) for the i-th of the machine's k parts (np.linspace(0, 1, k,
 * endpoint=False)), all stored as float32.
 *
 * Returns 0, or -1 when a part fits no machine (the reference then fails).
 * log (nullable): [n_parts][6] (machine, part id, batch, col, bottom row,
 * rotation) in greedy order.
 * ------------------------------------------------------------------------- */
typedef struct { pp_bin b; int32_t *parts; int np; } greedy_bin;

static int greedy_can_insert(const pp_problem *p, const rot_tables *t, const pp_bin *b, int m, int ty,
                             int *out_x, int *out_row, int *out_rot) {
    const int H = p->plate_h[m], W = p->plate_w[m], r0 = p->rot_base[ty];
    int best_len = 0, best_row = -1, best_x = 0, best_rot = -1;
    int64_t positions = 0;
    for (int r = 0; r < p->nrot[ty]; r++) {
        const int rg = r0 + r, h = p->rot_h[rg], w = p->rot_w[rg];
        if (h > H || w > W) continue;
        if (!vacancy_fit(b->vac, H, t->dens + t->dens_off[rg], h)) continue;
        int y, x;
        if (!bottom_left(b, H, W, t, rg, h, w, &y, &x, &positions)) continue;
        const int row = y + h - 1;
        const int new_len = b->enc_len > H - y ? b->enc_len : H - y;
        if (best_rot < 0 || new_len < best_len || (new_len == best_len && row > best_row) ||
            (new_len == best_len && row == best_row && x < best_x)) {
            best_len = new_len; best_row = row; best_x = x; best_rot = r;
        }
    }
    if (best_rot < 0) return 0;
    *out_x = best_x; *out_row = best_row; *out_rot = best_rot;
    return 1;
}

static void greedy_open(greedy_bin *g, int H, int W, int n) {
    pp_bin *b = &g->b;
    b->grid = (uint8_t *)calloc((size_t)H * W, 1);
    b->freer = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)H * W);
    b->vac = (int32_t *)malloc(sizeof(int32_t) * (size_t)H);
    for (int y = 0; y < H; y++) {
        b->vac[y] = W;
        for (int x = 0; x < W; x++) b->freer[(int64_t)y * W + x] = (uint16_t)(W - x);
    }
    b->area = 0.0; b->proc_time = 0.0; b->proc_time_height = 0.0;
    b->min_row = H; b->max_row = -1; b->enc_len = 0;
    g->parts = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    g->np = 0;
}

int oracle_initial_solution(const pp_problem *p, const int32_t *order, const double *machine_keys,
                            float *out_keys, double *out_makespan, int32_t *log) {
    const int n = p->n_parts, M = p->n_machines, P = p->n_types;
    rot_tables t;
    build_tables(p, &t);
    greedy_bin **bins = (greedy_bin **)calloc((size_t)M, sizeof(greedy_bin *));
    int *nbm = (int *)calloc((size_t)M, sizeof(int));
    double *mk = (double *)calloc((size_t)M, sizeof(double));
    for (int m = 0; m < M; m++) bins[m] = (greedy_bin *)calloc((size_t)n, sizeof(greedy_bin));
    double worst = 0.0, best_ms = 0.0;
    int rc = 0;
    for (int k = 0; k < n && rc == 0; k++) {
        const int ty = p->slot_type[order[k]];
        const int r0 = p->rot_base[ty];
        const int h0 = p->rot_h[r0], w0 = p->rot_w[r0];
        best_ms = 1.0 / 0.0;
        int kind = 0, cm = -1, cb = -1, cx = 0, crow = 0, crot = 0;   /* kind 1 new, 2 existing */
        double cmm = 0.0;
        for (int m = 0; m < M; m++) {
            const int H = p->plate_h[m], W = p->plate_w[m];
            int placed_exist = 0;
            if ((h0 > H || w0 > W) && (w0 > H || h0 > W)) continue;
            const int64_t mp = (int64_t)m * P + ty;
            const double pt = p->proc_time[mp], pth = p->proc_time_height[mp];
            double mm = ((mk[m] + p->setup[m]) + pt) + pth;
            double nm = worst > mm ? worst : mm;
            if (nm <= best_ms) { best_ms = nm; kind = 1; cm = m; cmm = mm; }
            for (int b = 0; b < nbm[m]; b++) {
                int x, row, rot;
                if (!greedy_can_insert(p, &t, &bins[m][b].b, m, ty, &x, &row, &rot)) continue;
                const double bpth = bins[m][b].b.proc_time_height;
                mm = bpth < pth ? ((mk[m] - bpth) + pth) + pt : mk[m] + pt;
                nm = worst > mm ? worst : mm;
                if (nm <= best_ms) {
                    if (placed_exist && nm == best_ms) continue;
                    placed_exist = 1; best_ms = nm; kind = 2; cm = m; cb = b; cx = x; crow = row; crot = rot;
                    cmm = mm;
                }
            }
        }
        if (kind == 0) { rc = -1; break; }
        const int H = p->plate_h[cm], W = p->plate_w[cm];
        if (kind == 1) {
            cb = nbm[cm]++;
            greedy_open(&bins[cm][cb], H, W, n);
            crot = p->best_rot[ty]; cx = 0; crow = H - 1;
        }
        greedy_bin *g = &bins[cm][cb];
        const int rg = r0 + crot, h = p->rot_h[rg], w = p->rot_w[rg];
        insert_part(&g->b, H, W, crow - h + 1, cx, p->masks + p->mask_off[rg], h, w, p->area[ty]);
        const int64_t mp = (int64_t)cm * P + ty;
        g->b.proc_time += p->proc_time[mp];
        if (p->proc_time_height[mp] > g->b.proc_time_height) g->b.proc_time_height = p->proc_time_height[mp];
        g->parts[g->np++] = p->type_id[ty];
        mk[cm] = cmm;
        if (cmm > worst) worst = cmm;
        if (log) {
            int32_t *e = log + 6 * k;
            e[0] = cm; e[1] = p->type_id[ty]; e[2] = cb; e[3] = cx; e[4] = crow; e[5] = crot;
        }
    }
    if (rc == 0) {
        char *used = (char *)calloc((size_t)n, 1);
        int32_t *pos = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
        for (int j = 0; j < 2 * n; j++) out_keys[j] = 0.0f;
        for (int m = 0; m < M; m++) {
            int cnt = 0;
            for (int b = 0; b < nbm[m]; b++)
                for (int q = 0; q < bins[m][b].np; q++) {
                    const int id = bins[m][b].parts[q];
                    for (int j = 0; j < n; j++)
                        if (!used[j] && p->type_id[p->slot_type[j]] == id) { used[j] = 1; pos[cnt++] = j; break; }
                }
            const double step = cnt ? 1.0 / (double)cnt : 0.0;
            for (int i = 0; i < cnt; i++) {
                out_keys[n + pos[i]] = (float)machine_keys[m];
                out_keys[pos[i]] = (float)((double)i * step);
            }
        }
        free(used); free(pos);
        *out_makespan = best_ms;
    }
    for (int m = 0; m < M; m++) {
        for (int b = 0; b < nbm[m]; b++) {
            free(bins[m][b].b.grid); free(bins[m][b].b.freer); free(bins[m][b].b.vac); free(bins[m][b].parts);
        }
        free(bins[m]);
    }
    free(bins); free(nbm); free(mk);
    free_tables(&t);
    return rc;
}
