/*
 * st_sort.h - stable device radix sort of (key, value) pairs (st_sort.hip).
 */
#ifndef ST_SORT_H
#define ST_SORT_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

/* Device scratch bytes for sorting n pairs. */
size_t st_radix_sort_scratch_bytes(uint32_t n);

/* Sorts (keys[i], vals[i]) by the low key_bits bits of the keys, ascending
   and stable, in place (keys_tmp / vals_tmp: n words each; scratch:
   st_radix_sort_scratch_bytes(n)). All work is queued on `stream`. Returns 0,
   or a negative value when n is beyond the sort's 67M limit or a launch
   failed. */
int st_radix_sort_pairs(hipStream_t stream, uint32_t *keys, uint32_t *vals, uint32_t *keys_tmp, uint32_t *vals_tmp,
                        uint32_t n, uint32_t key_bits, void *scratch);

/* Exclusive prefix sum of a[0, len) in place (sums: (len + 2047) / 2048
   words of scratch; len at most 2048 * 2048). Returns 0 on success. */
int st_scan_exclusive(hipStream_t stream, uint32_t *a, uint32_t len, uint32_t *sums);

#endif
