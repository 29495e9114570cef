// Connection rules generated on the device at calibration, straight into
// the coordinate arrays the synapse-store build sorts
// (src/connect_rules.h:92-340 are the reference's host loops).
//
// The connections are the ones the host generator in api.cpp makes for the
// same rule, seed and call index, in the same insertion order: the same
// counter-based Philox stream (key = seed ^ "RANDngpu", counter = (item,
// block, call, "conn")), the same bounded draws u*n >> 32, and the same
// sampler of k distinct integers per item (partial Fisher-Yates when
// n < 5k, otherwise rejection of repeats in draw order).
#include <hip/hip_runtime.h>

#include "ngpu_kernels.h"

namespace ngpu {

namespace {

constexpr int kGenBlock = 256;

__device__ __forceinline__ void philox4(uint32_t (&c)[4], uint32_t k0,
                                        uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint32_t lo0 = 0xD2511F53u * c[0];
    uint32_t hi0 = __umulhi(0xD2511F53u, c[0]);
    uint32_t lo1 = 0xCD9E8D57u * c[2];
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]);
    uint32_t n0 = hi1 ^ c[1] ^ k0;
    uint32_t n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// word `wi` of the stream of item `item`
__device__ __forceinline__ uint32_t stream_word(const ConnGenArgs &a,
                                                uint32_t item, uint32_t wi) {
  uint32_t c[4] = {item, wi >> 2, a.call, 0x636F6E6Eu};
  philox4(c, a.k0, a.k1);
  const uint32_t q = wi & 3u;
  return q == 0 ? c[0] : q == 1 ? c[1] : q == 2 ? c[2] : c[3];
}

__device__ __forceinline__ uint32_t bounded(uint32_t u, uint32_t n) {
  return (uint32_t)(((uint64_t)u * n) >> 32);
}

__device__ __forceinline__ int node_s(const ConnGenArgs &a, uint32_t k) {
  return a.s_arr ? a.s_arr[k] : a.s0 + (int)k;
}
__device__ __forceinline__ int node_t(const ConnGenArgs &a, uint32_t k) {
  return a.t_arr ? a.t_arr[k] : a.t0 + (int)k;
}

// connection k of the chunk (src/connect_rules.h:92-112 for the weight and
// delay of insertion index k)
__device__ __forceinline__ void put(const ConnGenArgs &a, unsigned long long k,
                                    int s, int t) {
  const float wv = a.w_arr ? a.w_arr[k]
                           : (a.wd.kind != kDistNone ? dist_sample(a.wd, k)
                                                     : a.w0);
  const float dv = a.d_arr ? a.d_arr[k]
                           : (a.dd.kind != kDistNone ? dist_sample(a.dd, k)
                                                     : a.d0);
  a.src[k] = (uint32_t)s;
  a.tp[k] = (uint32_t)t | a.port_word;
  a.w[k] = wv;
  a.d[k] = (uint16_t)((int)roundf(dv / a.dt) - 1);
  if (a.sg) a.sg[k] = a.sg_val;
}

// one_to_one, all_to_all, fixed_total_number: one thread per connection
__global__ __launch_bounds__(kGenBlock) void gen_simple_kernel(ConnGenArgs a) {
  const unsigned long long k =
      (unsigned long long)blockIdx.x * kGenBlock + threadIdx.x;
  if (k >= a.m) return;
  int s, t;
  if (a.rule == kGenOneToOne) {
    s = node_s(a, (uint32_t)k);
    t = node_t(a, (uint32_t)k);
  } else if (a.rule == kGenAllToAll) {
    s = node_s(a, (uint32_t)(k % (unsigned)a.ns));
    t = node_t(a, (uint32_t)(k / (unsigned)a.ns));
  } else {  // fixed_total_number: words 0 and 1 of block 0 of item k
    uint32_t c[4] = {(uint32_t)k, 0u, a.call, 0x636F6E6Eu};
    philox4(c, a.k0, a.k1);
    s = node_s(a, bounded(c[0], (uint32_t)a.ns));
    t = node_t(a, bounded(c[1], (uint32_t)a.nt));
  }
  put(a, k, s, t);
}

// fixed_indegree (items are targets, draws are sources) and
// fixed_outdegree (items are sources, draws are targets): one wave per
// item.  LDS: the sampler's state, then the k draws of the item.
__global__ __launch_bounds__(64) void gen_sample_kernel(ConnGenArgs a) {
  extern __shared__ uint32_t s_lds[];
  const uint32_t item = blockIdx.x;
  const int lane = threadIdx.x;
  const bool indeg = a.rule == kGenFixedIndegree;
  const uint32_t n = (uint32_t)(indeg ? a.ns : a.nt);
  const int K = a.K;
  uint32_t *out;
  if (n < 5u * (uint32_t)K) {
    // partial Fisher-Yates over [0, n): step i swaps i with
    // i + bounded(word i, n - i); later steps never touch entry i
    uint32_t *perm = s_lds;
    out = perm;
    for (uint32_t i = lane; i < n; i += 64) perm[i] = i;
    __syncthreads();
    if (lane == 0) {
      uint32_t c[4] = {0u, 0u, 0u, 0u};
      for (int i = 0; i < K; i++) {
        if ((i & 3) == 0) {
          c[0] = item;
          c[1] = (uint32_t)i >> 2;
          c[2] = a.call;
          c[3] = 0x636F6E6Eu;
          philox4(c, a.k0, a.k1);
        }
        const uint32_t u = c[i & 3];
        const uint32_t j = (uint32_t)i + bounded(u, n - (uint32_t)i);
        const uint32_t tmp = perm[i];
        perm[i] = perm[j];
        perm[j] = tmp;
      }
    }
    __syncthreads();
  } else {
    // rejection in draw order: 64 draws per round.  A draw is kept if its
    // value is neither in the table from earlier rounds nor drawn by a lower
    // lane in this round (the lowest lane of equal draws owns the slot).
    const uint32_t cap = (uint32_t)a.cap;
    uint32_t *table = s_lds;            // value + 1, 0 = empty
    uint32_t *owner = s_lds + cap;      // lowest lane that inserted
    out = s_lds + 2 * cap;
    for (uint32_t i = lane; i < cap; i += 64) {
      table[i] = 0u;
      owner[i] = 64u;
    }
    __syncthreads();
    int got = 0;
    for (uint32_t round = 0; got < K; round++) {
      const uint32_t u = stream_word(a, item, round * 64u + (uint32_t)lane);
      const uint32_t v = bounded(u, n);
      // earlier rounds
      uint32_t h = (v * 2654435761u) & (cap - 1u);
      bool dup = false;
      while (table[h] != 0u) {
        if (table[h] == v + 1u) {
          dup = true;
          break;
        }
        h = (h + 1u) & (cap - 1u);
      }
      __syncthreads();
      // insert (equal draws of this round meet in one slot)
      if (!dup) {
        for (;;) {
          const uint32_t old = atomicCAS(&table[h], 0u, v + 1u);
          if (old == 0u || old == v + 1u) break;
          h = (h + 1u) & (cap - 1u);
        }
        atomicMin(&owner[h], (uint32_t)lane);
      }
      __syncthreads();
      const bool keep = !dup && owner[h] == (uint32_t)lane;
      const unsigned long long m = __ballot(keep);
      const int before = __popcll(m & ((1ull << lane) - 1ull));
      if (keep && got + before < K) out[got + before] = v;
      got += __popcll(m);
      __syncthreads();
    }
  }
  // the item's k connections, insertion index item * k + i
  for (int i = lane; i < K; i += 64) {
    const unsigned long long kk = (unsigned long long)item * K + i;
    if (indeg)
      put(a, kk, node_s(a, out[i]), node_t(a, item));
    else
      put(a, kk, node_s(a, item), node_t(a, out[i]));
  }
}

}  // namespace

size_t ConnGenLdsBytes(int rule, int ns, int nt, int K, int *cap_out) {
  if (rule != kGenFixedIndegree && rule != kGenFixedOutdegree) return 0;
  const unsigned n = (unsigned)(rule == kGenFixedIndegree ? ns : nt);
  if (n < 5u * (unsigned)K) {
    *cap_out = 0;
    return (size_t)n * 4;
  }
  // at most k - 1 values before the last round plus 64 in it: the table
  // always keeps an empty slot, so every probe ends
  unsigned cap = 16;
  while (cap < 2u * (unsigned)K || cap < (unsigned)K + 65u) cap <<= 1;
  *cap_out = (int)cap;
  return ((size_t)2 * cap + (size_t)K) * 4;
}

const char *LaunchConnGen(const ConnGenArgs &a, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  if (a.m == 0) return nullptr;
  if (a.rule == kGenFixedIndegree || a.rule == kGenFixedOutdegree) {
    int cap = 0;
    const size_t lds = ConnGenLdsBytes(a.rule, a.ns, a.nt, a.K, &cap);
    if (lds > kConnGenMaxLds) return "connection rule state exceeds the LDS";
    ConnGenArgs b = a;
    b.cap = cap;
    const unsigned items = (unsigned)(a.rule == kGenFixedIndegree ? a.nt : a.ns);
    hipLaunchKernelGGL(gen_sample_kernel, dim3(items), dim3(64), lds, s, b);
  } else {
    const unsigned long long grid = (a.m + kGenBlock - 1) / kGenBlock;
    hipLaunchKernelGGL(gen_simple_kernel, dim3((unsigned)grid), dim3(kGenBlock),
                       0, s, a);
  }
  return nullptr;
}

}  // namespace ngpu
