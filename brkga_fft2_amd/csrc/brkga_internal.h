// brkga_internal.h -- entry points shared by the decoder and the device GA
// loop inside the library (not part of the public C ABI).
#ifndef BRKGA_INTERNAL_H
#define BRKGA_INTERNAL_H

#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/brkga_mi355x.h"

// Enqueue one population decode on stream s: keys [n_chrom][2n] on the device,
// makespans to d_out.  timed: bracket the decode kernel with the handle's HIP
// events (must be false while the stream is being captured into a graph).
int brkga_internal_launch(brkga_handle *h, const void *d_keys, int key_dtype, int64_t n_chrom,
                          double *d_out, int32_t *d_log, hipStream_t s, bool timed);
// Allocate the per-call buffers for populations of up to n_chrom (so that a
// captured launch allocates nothing).
int brkga_internal_reserve(brkga_handle *h, int64_t n_chrom);
int brkga_internal_n_parts(const brkga_handle *h);
// Changes whenever a device buffer used by the decode launch is reallocated
// (a graph captured before must be captured again).
unsigned brkga_internal_ws_gen(const brkga_handle *h);
// Contexts that have overflowed their bin workspace so far (synchronous read).
int brkga_internal_overflows(brkga_handle *h, unsigned long long *out);
int brkga_internal_device(const brkga_handle *h);
void brkga_internal_error(brkga_handle *h, const char *msg);

#endif
