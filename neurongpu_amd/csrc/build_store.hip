// On-device build of the synapse store at calibration.
//
// Input: connections in insertion order (coordinate form) already in HBM.
// Output: the canonical store the delivery kernel reads: synapses sorted by
// (source, delay index) with insertion order kept inside a delay group
// (the grouping of src/connect.cu:44-62 and the flattening of
// src/spike_buffer.cu:381-413), as 8-byte records {target|port<<26,
// weight}, plus per-source base offsets and per-(source, delay) segment
// offsets.  A stable LSD radix sort on key = source*D + delay with the
// insertion index as value gives exactly that order; a histogram + scan over
// the keys gives the segment offsets.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "build_store.h"

namespace ngpu {

namespace {

__global__ void make_keys(const uint32_t *src, const uint16_t *d, uint32_t D,
                          unsigned long long S, uint32_t *keys,
                          uint32_t *vals, uint32_t *count) {
  unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= S) return;
  uint32_t k = src[i] * D + d[i];
  keys[i] = k;
  vals[i] = (uint32_t)i;
  atomicAdd(count + k, 1u);
}

__global__ void gather_records(const uint32_t *vals, const uint32_t *tp,
                               const float *w, const uint8_t *sg_in,
                               unsigned long long S,
                               unsigned long long *rec, uint8_t *sg_out) {
  unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= S) return;
  uint32_t v = vals[i];
  rec[i] = (unsigned long long)tp[v] |
           ((unsigned long long)__float_as_uint(w[v]) << 32);
  if (sg_out) sg_out[i] = sg_in[v];
}

__global__ void segments(const unsigned long long *off, uint32_t N, uint32_t D,
                         unsigned long long *base, uint32_t *seg) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > N) return;
  unsigned long long b = off[(unsigned long long)s * D];
  base[s] = b;
  if (s == N) return;
  for (uint32_t d = 0; d <= D; d++)
    seg[(unsigned long long)s * (D + 1) + d] =
        (uint32_t)(off[(unsigned long long)s * D + d] - b);
}

// reverse index of plastic synapses: key = target, stable by synapse index
__global__ void rev_keys(const unsigned long long *rec, const uint8_t *sg,
                         unsigned long long S, uint32_t *keys, uint32_t *vals,
                         uint32_t *count, uint32_t *n_out) {
  unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= S) return;
  if (sg[i] == 0) return;
  uint32_t t = (uint32_t)rec[i] & kTargetMask;
  uint32_t pos = atomicAdd(n_out, 1u);
  keys[pos] = t;
  vals[pos] = (uint32_t)i;
  atomicAdd(count + t, 1u);
}

// Connection statistics: one wavefront per source walks its records in
// delay order.  Per-source and per-delay counts are wave sums (a source is
// owned by one wave); in-degrees are atomics on the target's counter.
__global__ __launch_bounds__(256) void count_synapses(
    const unsigned long long *rec, const unsigned long long *base,
    const uint32_t *seg, uint32_t s_begin, uint32_t s_end, uint32_t D,
    int lo, int hi, int port, int *in_deg, int *out_deg,
    unsigned long long *delay_cnt, int n_delay) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t n_wave = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t s = s_begin + ((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
       s < s_end; s += n_wave) {
    const bool own = (int)s >= lo && (int)s < hi;  // wave-uniform
    const unsigned long long *r = rec + base[s];
    const uint32_t *sg = seg + (unsigned long long)s * (D + 1);
    int n_own = 0;
    for (uint32_t d = 0; d < D; d++) {
      const uint32_t j1 = sg[d + 1];
      int n_d = 0;
      for (uint32_t j = sg[d] + lane; j < j1; j += 64u) {
        const uint32_t tp = (uint32_t)r[j];
        if (port >= 0 && (int)(tp >> kPortShift) != port) continue;
        n_d++;
        const uint32_t t = tp & kTargetMask;
        if (in_deg && (int)t >= lo && (int)t < hi)
          atomicAdd(in_deg + (t - (uint32_t)lo), 1);
      }
      if (own) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) n_d += __shfl_xor(n_d, off);
        n_own += n_d;
        if (delay_cnt && (int)d < n_delay && lane == 0 && n_d > 0)
          atomicAdd(delay_cnt + d, (unsigned long long)n_d);
      }
    }
    if (own && out_deg && lane == 0) out_deg[s - (uint32_t)lo] += n_own;
  }
}

// Tiled store: the accumulator tile of every canonical record, with its
// canonical index as the sort value.
__global__ void tile_keys(const unsigned long long *rec, unsigned long long S,
                          long long n_total, int tile_shift, uint16_t *keys,
                          uint32_t *vals) {
  unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= S) return;
  const uint32_t tp = (uint32_t)rec[i];
  const long long a = (long long)(tp >> kPortShift) * n_total + (tp & kTargetMask);
  keys[i] = (uint16_t)(a >> tile_shift);
  vals[i] = (uint32_t)i;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t m) {
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t y = (uint32_t)__shfl_xor((int)m, o);
    m = y > m ? y : m;
  }
  return m;
}

// Bound of the fixed-point tile sums: block (t, k) counts, in LDS, the
// records onto each entry of tile t in chunk k of gb of the tile's records
// and adds its largest count to bound[t] (at least the largest count of
// any entry of the tile, summed over the chunks); wmax gets the largest
// |weight| as float bits (NaN above all).
constexpr int kBoundBlock = 1024;
__global__ __launch_bounds__(kBoundBlock) void tile_bound(
    const unsigned long long *trec, const uint32_t *off, unsigned long long nd,
    uint32_t T, int tile_shift, long long n_total, uint32_t gb,
    uint32_t *bound, uint32_t *wmax) {
  extern __shared__ uint32_t s_cnt[];
  __shared__ uint32_t s_red[2];
  const uint32_t t = blockIdx.x / gb, k = blockIdx.x - t * gb;
  const int TL = 1 << tile_shift;
  for (int i = threadIdx.x; i < TL; i += kBoundBlock) s_cnt[i] = 0u;
  if (threadIdx.x == 0) s_red[0] = s_red[1] = 0u;
  __syncthreads();
  const uint32_t r0 = off[t], n = off[nd * T + t] - r0;
  const uint32_t c0 = r0 + (uint32_t)((unsigned long long)n * k / gb);
  const uint32_t c1 = r0 + (uint32_t)((unsigned long long)n * (k + 1) / gb);
  const long long a0 = (long long)t << tile_shift;
  uint32_t wb = 0;
  for (uint32_t j = c0 + threadIdx.x; j < c1; j += kBoundBlock) {
    const unsigned long long r = trec[j];
    const uint32_t tp = (uint32_t)r;
    const long long a =
        (long long)(tp >> kPortShift) * n_total + (tp & kTargetMask);
    atomicAdd(&s_cnt[a - a0], 1u);
    const uint32_t x = (uint32_t)(r >> 32) & 0x7FFFFFFFu;
    wb = x > wb ? x : wb;
  }
  __syncthreads();
  uint32_t m = 0;
  for (int i = threadIdx.x; i < TL; i += kBoundBlock)
    m = s_cnt[i] > m ? s_cnt[i] : m;
  m = wave_max_u32(m);
  wb = wave_max_u32(wb);
  if ((threadIdx.x & 63u) == 0) {
    atomicMax(&s_red[0], m);
    atomicMax(&s_red[1], wb);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(bound + t, s_red[0]);
    atomicMax(wmax, s_red[1]);
  }
}

__global__ void gather_u64(const uint32_t *idx, const unsigned long long *src,
                           unsigned long long S, unsigned long long *dst) {
  unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < S) dst[i] = src[idx[i]];
}

// Records per (tile, source, delay slot), tile-major: one wavefront per
// source walks its delay segments.
__global__ __launch_bounds__(256) void tile_counts(
    const unsigned long long *rec, const unsigned long long *base,
    const uint32_t *seg, uint32_t N, uint32_t D, long long n_total,
    int tile_shift, uint32_t *cnt) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t n_wave = (gridDim.x * blockDim.x) >> 6;
  const unsigned long long nd = (unsigned long long)N * D;
  for (uint32_t s = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; s < N;
       s += n_wave) {
    const unsigned long long *r = rec + base[s];
    const uint32_t *sg = seg + (unsigned long long)s * (D + 1);
    for (uint32_t d = 0; d < D; d++) {
      const uint32_t j1 = sg[d + 1];
      for (uint32_t j = sg[d] + lane; j < j1; j += 64u) {
        const uint32_t tp = (uint32_t)r[j];
        const long long a =
            (long long)(tp >> kPortShift) * n_total + (tp & kTargetMask);
        const unsigned long long t = (unsigned long long)(a >> tile_shift);
        atomicAdd(cnt + t * nd + (unsigned long long)s * D + d, 1u);
      }
    }
  }
}

// off[(sd)*T + t] = start[t*ND + sd] for sd in [0, ND]
__global__ void tile_transpose(const uint32_t *start, unsigned long long nd,
                               uint32_t T, uint32_t *off) {
  const unsigned long long n = (nd + 1) * T;
  for (unsigned long long k = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
       k < n; k += (unsigned long long)gridDim.x * blockDim.x) {
    const unsigned long long sd = k / T, t = k - sd * T;
    off[k] = start[t * nd + sd];
  }
}

template <class T>
struct DTmp {
  T *p = nullptr;
  explicit DTmp(size_t n) {
    if (hipMalloc((void **)&p, (n ? n : 1) * sizeof(T)) != hipSuccess) p = nullptr;
  }
  ~DTmp() {
    if (p) (void)hipFree(p);
  }
};

int bits_for(unsigned long long v) {
  int b = 0;
  while (b < 32 && (1ull << b) <= v) b++;
  return b < 1 ? 1 : b;
}

}  // namespace

const char *BuildSynapseStore(const uint32_t *src, const uint32_t *tp,
                              const float *w, const uint16_t *d,
                              const uint8_t *sg_in, unsigned long long S,
                              uint32_t N, uint32_t D,
                              unsigned long long *rec, uint8_t *sg_out,
                              unsigned long long *base, uint32_t *seg,
                              void *stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  unsigned long long nk = (unsigned long long)N * D;
  if (nk >= 0xFFFFFFFFull) return "source x delay key space exceeds 32 bits";
  DTmp<uint32_t> keys(S), keys2(S), vals(S), vals2(S), count(nk + 1);
  DTmp<unsigned long long> off(nk + 1);
  if (!keys.p || !keys2.p || !vals.p || !vals2.p || !count.p || !off.p)
    return "out of device memory building the synapse store";
  if (hipMemsetAsync(count.p, 0, (nk + 1) * sizeof(uint32_t), stream) != hipSuccess)
    return "hipMemsetAsync failed";
  unsigned int grid = (unsigned int)((S + 255) / 256);
  if (S > 0)
    hipLaunchKernelGGL(make_keys, dim3(grid), dim3(256), 0, stream, src, d, D,
                       S, keys.p, vals.p, count.p);
  // exclusive scan of the per-key counts -> first position of each key
  size_t tmp_bytes = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, count.p, off.p,
                                   (int)(nk + 1), stream);
  size_t sort_bytes = 0;
  hipcub::DoubleBuffer<uint32_t> kb(keys.p, keys2.p), vb(vals.p, vals2.p);
  int end_bit = bits_for(nk);
  if (S > 0)
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, kb, vb, (int)S, 0,
                                       end_bit, stream);
  DTmp<unsigned char> tmp(tmp_bytes > sort_bytes ? tmp_bytes : sort_bytes);
  if (!tmp.p) return "out of device memory building the synapse store";
  if (hipcub::DeviceScan::ExclusiveSum(tmp.p, tmp_bytes, count.p, off.p,
                                       (int)(nk + 1), stream) != hipSuccess)
    return "scan failed";
  if (S > 0) {
    size_t sb = tmp_bytes > sort_bytes ? tmp_bytes : sort_bytes;
    if (hipcub::DeviceRadixSort::SortPairs(tmp.p, sb, kb, vb, (int)S, 0,
                                           end_bit, stream) != hipSuccess)
      return "radix sort failed";
    hipLaunchKernelGGL(gather_records, dim3(grid), dim3(256), 0, stream,
                       vb.Current(), tp, w, sg_in, S, rec, sg_out);
  }
  hipLaunchKernelGGL(segments, dim3((N + 1 + 255) / 256), dim3(256), 0, stream,
                     off.p, N, D, base, seg);
  if (hipStreamSynchronize(stream) != hipSuccess) return "synapse store build failed";
  return nullptr;
}

const char *BuildReverseIndex(const unsigned long long *rec,
                              const uint8_t *sg, unsigned long long S,
                              uint32_t N, unsigned long long *rev_ptr,
                              uint32_t **rev_idx_out,
                              unsigned long long *n_rev_out, void *stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  DTmp<uint32_t> keys(S), keys2(S), vals(S), vals2(S), count(N + 1), n_out(1);
  if (!keys.p || !keys2.p || !vals.p || !vals2.p || !count.p || !n_out.p)
    return "out of device memory building the reverse index";
  (void)hipMemsetAsync(count.p, 0, (N + 1) * sizeof(uint32_t), stream);
  (void)hipMemsetAsync(n_out.p, 0, sizeof(uint32_t), stream);
  unsigned int grid = (unsigned int)((S + 255) / 256);
  if (S > 0)
    hipLaunchKernelGGL(rev_keys, dim3(grid), dim3(256), 0, stream, rec, sg, S,
                       keys.p, vals.p, count.p, n_out.p);
  uint32_t n_rev = 0;
  (void)hipMemcpyAsync(&n_rev, n_out.p, sizeof(uint32_t), hipMemcpyDeviceToHost,
                       stream);
  if (hipStreamSynchronize(stream) != hipSuccess) return "reverse index build failed";
  // the atomic append order is arbitrary: sort by (target, synapse index)
  // in two stable passes (index first, then target) for canonical order
  size_t tmp_bytes = 0, b2 = 0;
  hipcub::DoubleBuffer<uint32_t> kb(keys.p, keys2.p), vb(vals.p, vals2.p);
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, vb, kb, (int)n_rev, 0,
                                     32, stream);
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, b2, count.p, rev_ptr, (int)(N + 1),
                                   stream);
  size_t tb = tmp_bytes > b2 ? tmp_bytes : b2;
  DTmp<unsigned char> tmp(tb);
  if (!tmp.p) return "out of device memory building the reverse index";
  if (n_rev > 0) {
    if (hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, vb, kb, (int)n_rev, 0, 32,
                                           stream) != hipSuccess)
      return "radix sort failed";
    if (hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, kb, vb, (int)n_rev, 0,
                                           bits_for(N), stream) != hipSuccess)
      return "radix sort failed";
  }
  if (hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, count.p, rev_ptr,
                                       (int)(N + 1), stream) != hipSuccess)
    return "scan failed";
  uint32_t *idx = nullptr;
  if (hipMalloc((void **)&idx, ((size_t)n_rev + 1) * sizeof(uint32_t)) != hipSuccess)
    return "out of device memory building the reverse index";
  if (n_rev > 0)
    (void)hipMemcpyAsync(idx, vb.Current(), (size_t)n_rev * sizeof(uint32_t),
                         hipMemcpyDeviceToDevice, stream);
  if (hipStreamSynchronize(stream) != hipSuccess) return "reverse index build failed";
  *rev_idx_out = idx;
  *n_rev_out = n_rev;
  return nullptr;
}

const char *BuildTiledStore(const unsigned long long *rec,
                            const unsigned long long *base,
                            const uint32_t *seg, unsigned long long S,
                            uint32_t N, uint32_t D, uint32_t T,
                            int tile_shift, long long n_total,
                            unsigned long long *trec, uint32_t *off,
                            void *stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  const unsigned long long nd = (unsigned long long)N * D;
  if (S >= 0x7FFFFFFFull || nd * T >= 0x7FFFFFFFull || T > 0xFFFFu)
    return "synapse store too large for the tiled delivery";
  if (S > 0) {
    DTmp<uint16_t> k1(S), k2(S);
    DTmp<uint32_t> v1(S), v2(S);
    if (!k1.p || !k2.p || !v1.p || !v2.p)
      return "out of device memory building the tiled store";
    const unsigned int grid = (unsigned int)((S + 255) / 256);
    hipLaunchKernelGGL(tile_keys, dim3(grid), dim3(256), 0, stream, rec, S,
                       n_total, tile_shift, k1.p, v1.p);
    hipcub::DoubleBuffer<uint16_t> kb(k1.p, k2.p);
    hipcub::DoubleBuffer<uint32_t> vb(v1.p, v2.p);
    const int end_bit = bits_for(T > 1 ? T - 1 : 1);
    size_t sort_bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, kb, vb,
                                             (int)S, 0, end_bit, stream);
    DTmp<unsigned char> tmp(sort_bytes);
    if (!tmp.p) return "out of device memory building the tiled store";
    if (hipcub::DeviceRadixSort::SortPairs(tmp.p, sort_bytes, kb, vb, (int)S,
                                           0, end_bit, stream) != hipSuccess)
      return "radix sort failed";
    hipLaunchKernelGGL(gather_u64, dim3(grid), dim3(256), 0, stream,
                       vb.Current(), rec, S, trec);
    // the sort buffers are freed at the end of this scope
    if (hipStreamSynchronize(stream) != hipSuccess)
      return "tiled store build failed";
  }
  const unsigned long long nc = nd * T + 1;
  DTmp<uint32_t> cnt(nc), start(nc);
  if (!cnt.p || !start.p) return "out of device memory building the tiled store";
  if (hipMemsetAsync(cnt.p, 0, nc * sizeof(uint32_t), stream) != hipSuccess)
    return "hipMemsetAsync failed";
  if (N > 0 && D > 0) {
    unsigned int g = (N + 3) / 4;
    if (g > 65536u) g = 65536u;
    hipLaunchKernelGGL(tile_counts, dim3(g), dim3(256), 0, stream, rec, base,
                       seg, N, D, n_total, tile_shift, cnt.p);
  }
  size_t scan_bytes = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, cnt.p, start.p,
                                         (int)nc, stream);
  DTmp<unsigned char> stmp(scan_bytes);
  if (!stmp.p) return "out of device memory building the tiled store";
  if (hipcub::DeviceScan::ExclusiveSum(stmp.p, scan_bytes, cnt.p, start.p,
                                       (int)nc, stream) != hipSuccess)
    return "scan failed";
  hipLaunchKernelGGL(tile_transpose, dim3(4096), dim3(256), 0, stream,
                     start.p, nd, T, off);
  if (hipStreamSynchronize(stream) != hipSuccess)
    return "tiled store build failed";
  return nullptr;
}

const char *TiledSumBound(const unsigned long long *trec, const uint32_t *off,
                          uint32_t N, uint32_t D, uint32_t T, int tile_shift,
                          long long n_total, unsigned long long *bound,
                          float *max_w, void *stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  *bound = 0;
  *max_w = 0.0f;
  const unsigned long long nd = (unsigned long long)N * D;
  if (T == 0) return nullptr;
  DTmp<uint32_t> tb(T + 1);
  if (!tb.p) return "out of device memory";
  if (hipMemsetAsync(tb.p, 0, (T + 1) * sizeof(uint32_t), stream) !=
      hipSuccess)
    return "hipMemsetAsync failed";
  const size_t lds = sizeof(uint32_t) << tile_shift;
  static bool attr = false;
  if (!attr) {
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(tile_bound),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)(sizeof(uint32_t) << 15)) != hipSuccess)
      return "hipFuncSetAttribute failed";
    attr = true;
  }
  const uint32_t gb = 32;
  hipLaunchKernelGGL(tile_bound, dim3(T * gb), dim3(kBoundBlock), lds, stream,
                     trec, off, nd, T, tile_shift, n_total, gb, tb.p,
                     tb.p + T);
  std::vector<uint32_t> h(T + 1);
  if (hipMemcpyAsync(h.data(), tb.p, (T + 1) * sizeof(uint32_t),
                     hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess)
    return "tiled bound failed";
  for (uint32_t t = 0; t < T; t++) *bound = std::max<unsigned long long>(*bound, h[t]);
  std::memcpy(max_w, &h[T], sizeof(float));
  return nullptr;
}

void CountSynapses(const unsigned long long *rec,
                   const unsigned long long *base, const uint32_t *seg,
                   uint32_t N, uint32_t D, int lo, int hi, int port,
                   int *in_deg, int *out_deg, unsigned long long *delay_cnt,
                   int n_delay, void *stream) {
  // in-degrees need every source; the other counts only the range's own
  const uint32_t s_begin = in_deg ? 0u : (uint32_t)lo;
  const uint32_t s_end = in_deg ? N : (uint32_t)hi;
  if (s_end <= s_begin || D == 0) return;
  unsigned long long waves = s_end - s_begin;
  unsigned int grid = (unsigned int)((waves + 3) / 4);
  if (grid > 65536u) grid = 65536u;
  hipLaunchKernelGGL(count_synapses, dim3(grid), dim3(256), 0,
                     (hipStream_t)stream, rec, base, seg, s_begin, s_end, D,
                     lo, hi, port, in_deg, out_deg, delay_cnt, n_delay);
}

}  // namespace ngpu
