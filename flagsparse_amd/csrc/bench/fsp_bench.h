/*
 * fsp_bench.h -- C ABI of libflagsparse_bench.so, the measurement kernels
 * bench.py times beside the operators.  Not part of the operator library
 * (libflagsparse_hip.so) and not a counterpart of anything in the reference.
 *
 * Return value: 0 or -1 (bad arguments) / -3 (launch failed);
 * fsp_bench_last_error() returns a thread-local message for the last failure.
 */
#ifndef FLAGSPARSE_BENCH_H
#define FLAGSPARSE_BENCH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* fsp_bench_last_error(void);

/* Gather-ceiling probe of the CSR SpMM: n_gathers reads of whole rows of
 * row_bytes (a power of two in [16, 1024]) from table_dev, which holds
 * table_bytes (>= n_rows * row_bytes) and starts on a 16-byte boundary, at
 * row ids in [0, n_rows) (1 <= n_rows < 2^32) from a hash (no index stream).
 * Nothing is written except, for a sum no ordinary table produces, one float
 * at sink_dev.  n_gathers * row_bytes / time is the practical rate of random
 * row reads of that size from a table of that footprint (the SpMM reads one
 * B row of N values per nonzero). */
int fsp_gather_rows_probe(const void* table_dev, int64_t table_bytes, int64_t n_rows, int32_t row_bytes,
                          int64_t n_gathers, void* sink_dev, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FLAGSPARSE_BENCH_H */
