/*
 * A C host of the decoder, through the C ABI only (no Python): builds a small
 * two-machine problem, decodes a seeded population with brkga_evaluate (host
 * buffers) and prints every chromosome's keys and makespan, which the GPU test
 * checks against the oracle.  Output lines:
 *   problem n_parts n_chrom
 *   keys i k_0 ... k_{2n-1}     (float32, %.9g round-trips exactly)
 *   makespan i value            (%.17g)
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "brkga_mi355x.h"

#define N_PARTS 8
#define N_TYPES 3
#define N_ROTS 7
#define N_CHROM 64

static uint64_t lcg = 0x2545f4914f6cdd1dull;
static float next_key(void) {
    lcg = lcg * 6364136223846793005ull + 1442695040888963407ull;
    return (float)(lcg >> 40) * (1.0f / 16777216.0f);
}

/* rectangle h x w */
static void rect(uint8_t *m, int h, int w) { memset(m, 1, (size_t)h * w); }

/* L: a 3-wide bar down the left and a 3-high bar along the bottom of a s x s box */
static void ell(uint8_t *m, int s) {
    for (int y = 0; y < s; y++)
        for (int x = 0; x < s; x++) m[y * s + x] = (uint8_t)(x < 3 || y >= s - 3);
}

/* quarter turn: out (w x h) from in (h x w) */
static void turn(const uint8_t *in, int h, int w, uint8_t *out) {
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) out[x * h + (h - 1 - y)] = in[y * w + x];
}

int main(void) {
    static uint8_t masks[4096];
    int32_t rot_h[N_ROTS], rot_w[N_ROTS];
    int64_t mask_off[N_ROTS];
    int64_t off = 0;
    int r = 0;
    /* type 0: 8 x 12 rectangle, two rotations */
    rot_h[r] = 8; rot_w[r] = 12; mask_off[r] = off; rect(masks + off, 8, 12); off += 96; r++;
    rot_h[r] = 12; rot_w[r] = 8; mask_off[r] = off; rect(masks + off, 12, 8); off += 96; r++;
    /* type 1: 10 x 10 L, four rotations */
    rot_h[r] = 10; rot_w[r] = 10; mask_off[r] = off; ell(masks + off, 10); off += 100; r++;
    for (int k = 0; k < 3; k++) {
        rot_h[r] = 10; rot_w[r] = 10; mask_off[r] = off;
        turn(masks + mask_off[r - 1], 10, 10, masks + off); off += 100; r++;
    }
    /* type 2: 6 x 6 square, one rotation */
    rot_h[r] = 6; rot_w[r] = 6; mask_off[r] = off; rect(masks + off, 6, 6); off += 36; r++;

    const double thresholds[1] = {0.5};
    const int32_t slot_type[N_PARTS] = {0, 1, 2, 0, 1, 2, 0, 2};
    const int32_t type_id[N_TYPES] = {10, 11, 12};
    const double area[N_TYPES] = {96.0, 51.0, 36.0};   /* ones in each mask */
    const int32_t nrot[N_TYPES] = {2, 4, 1};
    const int32_t best_rot[N_TYPES] = {0, 0, 0};
    const int32_t rot_base[N_TYPES] = {0, 2, 6};
    const int32_t plate_h[2] = {30, 25}, plate_w[2] = {30, 40};
    const double setup[2] = {1.0, 2.0};
    const double proc_time[2 * N_TYPES] = {1.5, 2.0, 0.5, 1.2, 2.5, 0.7};
    const double proc_time_height[2 * N_TYPES] = {0.1, 0.2, 0.05, 0.15, 0.1, 0.02};

    brkga_problem p;
    memset(&p, 0, sizeof p);
    p.n_parts = N_PARTS; p.n_machines = 2; p.n_types = N_TYPES; p.n_rots = N_ROTS;
    p.thresholds = thresholds; p.slot_type = slot_type; p.type_id = type_id; p.area = area;
    p.nrot = nrot; p.best_rot = best_rot; p.rot_base = rot_base; p.rot_h = rot_h; p.rot_w = rot_w;
    p.mask_off = mask_off; p.masks = masks; p.plate_h = plate_h; p.plate_w = plate_w;
    p.setup = setup; p.proc_time = proc_time; p.proc_time_height = proc_time_height;

    brkga_handle *h = NULL;
    if (brkga_create(&p, NULL, &h) != BRKGA_OK) {
        fprintf(stderr, "brkga_create: %s\n", brkga_last_error(h));
        return 1;
    }
    static float keys[N_CHROM * 2 * N_PARTS];
    for (int i = 0; i < N_CHROM * 2 * N_PARTS; i++) keys[i] = next_key();
    double ms[N_CHROM];
    const int rc = brkga_evaluate(h, keys, BRKGA_KEYS_F32, N_CHROM, 2 * N_PARTS, ms, NULL, NULL);
    if (rc != BRKGA_OK) {
        fprintf(stderr, "brkga_evaluate: %d %s\n", rc, brkga_last_error(h));
        brkga_destroy(h);
        return 1;
    }
    printf("problem %d %d\n", N_PARTS, N_CHROM);
    for (int i = 0; i < N_CHROM; i++) {
        printf("keys %d", i);
        for (int k = 0; k < 2 * N_PARTS; k++) printf(" %.9g", (double)keys[i * 2 * N_PARTS + k]);
        printf("\nmakespan %d %.17g\n", i, ms[i]);
    }
    brkga_destroy(h);
    return 0;
}
