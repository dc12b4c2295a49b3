/*
 * fsp_bench.h -- C ABI of libflagsparse_bench.so, the measurement kernels
 * bench.py times beside the operators (the SpMM's gather ceiling and the
 * column-slab schedule it was measured against).  Not part of the operator library
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

/* Column-slab schedule probe of the CSR SpMM (N = 64, float32): C = A B
 * computed slab by slab over n_slabs column ranges of A (row ranges of B),
 * each lane group keeping rows_per_group (4 or 8) virtual rows' accumulators
 * in registers across the slabs.  A virtual row v covers nonzeros
 * [vbeg[v], vbeg[v] + slab_end[v][n_slabs - 1]) of indices / data (sorted
 * columns, at most 65535 nonzeros); slab_end[v][k] (uint16, row-major
 * [n_virtual][n_slabs]) is the end offset of slab k inside it; vslot[v] is
 * its output row: < n_rows writes C[vslot] (row-major [n_rows, 64]), else
 * partials[vslot - n_rows] ([*, 64]).  B is row-major [K, 64].  spin_limit
 * >= 0: after every slab the workgroups with equal blockIdx % 8 wait for each
 * other (counters_dev: 8 int32, zeroed by the call; at most spin_limit polls,
 * then *flag_dev |= 1); spin_limit < 0: no waits.  n_workgroups 1024-thread
 * workgroups, launched cooperatively (every one resident, or -3). */
int fsp_slab_spmm_probe(int32_t rows_per_group, const int32_t* vbeg_dev, const uint16_t* slab_end_dev,
                        const int32_t* vslot_dev, const int32_t* indices_dev, const float* data_dev,
                        const float* B_dev, float* C_dev, float* partials_dev, int32_t n_rows,
                        int32_t n_virtual, int32_t n_slabs, int32_t n_workgroups, int32_t* counters_dev,
                        int32_t spin_limit, int32_t* flag_dev, void* stream);

/* Streaming-read probe: n_arrays (1 or 2) device arrays (16-byte aligned;
 * b_dev ignored for one) of n_chunks 16-byte chunks each, read in lock step
 * (chunk i of every array by one lane, nontemporal) by n_workgroups
 * 1024-thread workgroups, each a contiguous range of chunks (contiguous = 1)
 * or chunks b, b + grid, ... (0).  Writes nothing but, for a sum no
 * ordinary data produces, one float at sink_dev. */
int fsp_stream_read_probe(const void* a_dev, const void* b_dev, int64_t n_chunks, int32_t n_arrays,
                          int32_t n_workgroups, int32_t contiguous, void* sink_dev, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FLAGSPARSE_BENCH_H */
