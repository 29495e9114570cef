// On-device build of the synapse store (build_store.hip).
#pragma once
#include <stdint.h>

#include "ngpu_kernels.h"

namespace ngpu {

// Sorts S connections (insertion order) into the canonical store.  All
// pointers are device pointers; sg_in/sg_out may be null (no plasticity).
// Returns null on success or a static error message.
const char *BuildSynapseStore(const uint32_t *src, const uint32_t *tp,
                              const float *w, const uint16_t *d,
                              const uint8_t *sg_in, unsigned long long S,
                              uint32_t N, uint32_t D,
                              unsigned long long *rec, uint8_t *sg_out,
                              unsigned long long *base, uint32_t *seg,
                              void *stream);

// Plastic synapses grouped by target in canonical synapse order.
// rev_ptr (device, N+1) is written; *rev_idx_out is a new device array.
const char *BuildReverseIndex(const unsigned long long *rec,
                              const uint8_t *sg, unsigned long long S,
                              uint32_t N, unsigned long long *rev_ptr,
                              uint32_t **rev_idx_out,
                              unsigned long long *n_rev_out, void *stream);

// Tile-major copy of the canonical store for the tiled delivery: trec
// (S records) holds the records grouped by accumulator tile
// ((port * n_total + target) >> tile_shift), canonical order inside a tile;
// off ((N*D + 1) * T entries) gets off[(s*D + d)*T + t] = position in trec
// of the first record of source s, delay slot d in tile t (row N*D: the
// tile ends).  Requires S < 2^31 and T*N*D < 2^31.  Returns null on
// success or a static error message.
const char *BuildTiledStore(const unsigned long long *rec,
                            const unsigned long long *base,
                            const uint32_t *seg, unsigned long long S,
                            uint32_t N, uint32_t D, uint32_t T,
                            int tile_shift, long long n_total,
                            unsigned long long *trec, uint32_t *off,
                            void *stream);

// For the fixed-point tile sums of that store: bound >= the most records
// onto one accumulator entry, max_w = the largest |weight| (NaN if a
// weight is NaN).  Returns null on success or a static error message.
const char *TiledSumBound(const unsigned long long *trec, const uint32_t *off,
                          uint32_t N, uint32_t D, uint32_t T, int tile_shift,
                          long long n_total, unsigned long long *bound,
                          float *max_w, void *stream);

// Synapse counts of the canonical store (device pointers): for nodes lo ..
// hi-1, in_deg[t - lo] += synapses with target t, out_deg[s - lo] +=
// synapses with source s, delay_cnt[d] += synapses with source in the range
// and delay index d < n_delay; only synapses on receptor `port` (all ports
// if port < 0).  Outputs may be null and are added to (the caller zeroes
// them).  Enqueued on `stream`.
void CountSynapses(const unsigned long long *rec,
                   const unsigned long long *base, const uint32_t *seg,
                   uint32_t N, uint32_t D, int lo, int hi, int port,
                   int *in_deg, int *out_deg, unsigned long long *delay_cnt,
                   int n_delay, void *stream);

}  // namespace ngpu
