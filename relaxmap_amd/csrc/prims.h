// Device primitives: scans, stable LSD radix sort, fills.  All deterministic.
#pragma once

#include "common.h"

namespace rmx {

// Scratch arena owned by the caller; grows on demand (hipMallocAsync-free:
// the context allocates once and reuses).  A fixed arena wraps memory it does
// not own (a region of a static device array): it never allocates or frees,
// and a request beyond its size throws.
struct Scratch {
  void *ptr = nullptr;
  size_t bytes = 0;
  bool fixed = false;
  void *get(size_t need);
  void release();
};

// Bytes of scratch that a radix sort of n (K, V) pairs on `bits` key bits
// takes (the largest need of radix_sort_pairs_*), and an exclusive scan of n.
size_t radix_sort_scratch_bytes(int64_t n, int key_bytes, int val_bytes, int bits);
size_t exclusive_scan_scratch_bytes(int64_t n);

// out[i] = sum_{j<i} in[i]; *d_total (device, may be null) = sum of all.
// In-place (out == in) is allowed.  n < 2^32.
void exclusive_scan_u32(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *d_total,
                        Scratch &scratch, hipStream_t s);
// grows scratch to what exclusive_scan_u32 over n elements takes (so that
// the scan allocates nothing, e.g. under stream capture)
void exclusive_scan_reserve(int64_t n, Scratch &scratch);

// Stable LSD radix sort of (key, val) pairs on bits [begin_bit, end_bit).
// keys/vals are sorted in place; tmp buffers of n elements are taken from
// scratch.  Stability: equal keys keep their input order.
// With `prof`, every launch of the sort is timed and accounted under its
// kernel's name (bytes: the keys read by the digit count, the keys and
// values read and written by each pass).
struct Profiler;
void radix_sort_pairs_u32(uint32_t *keys, uint32_t *vals, int64_t n, int begin_bit, int end_bit,
                          Scratch &scratch, hipStream_t s, Profiler *prof = nullptr);
void radix_sort_pairs_u64(uint64_t *keys, uint32_t *vals, int64_t n, int begin_bit, int end_bit,
                          Scratch &scratch, hipStream_t s, Profiler *prof = nullptr);
// Leaves the result where the last pass wrote it (the inputs or the scratch,
// valid until the scratch is used again): *out_keys / *out_vals.
void radix_sort_pairs_u64_u64(uint64_t *keys, uint64_t *vals, int64_t n, int begin_bit, int end_bit,
                              Scratch &scratch, hipStream_t s, uint64_t **out_keys, uint64_t **out_vals,
                              Profiler *prof = nullptr);

void fill_u32(uint32_t *p, uint32_t v, int64_t n, hipStream_t s);
void fill_u8(uint8_t *p, uint8_t v, int64_t n, hipStream_t s);
void fill_i64(long long *p, long long v, int64_t n, hipStream_t s);
void iota_u32(uint32_t *p, int64_t n, hipStream_t s);

}  // namespace rmx
