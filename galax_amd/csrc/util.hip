// util.hip — error plumbing, grow-only device buffers and the device prefix
// scans (sum / max over uint32) used by the tree build and list compaction.
#include <cstdarg>
#include <cstring>

#include "common.h"

namespace gx {

namespace {
thread_local char g_err[1024] = {0};
}

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

const char* last_error() { return g_err; }

int dbuf_reserve(DBuf& b, size_t bytes, hipStream_t s) {
  if (bytes <= b.bytes && b.p) return 0;
  if (b.p) {
    GX_HIP(hipStreamSynchronize(s));
    GX_HIP(hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
  }
  size_t want = bytes < 256 ? 256 : bytes;
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) {
    b.p = nullptr;
    return set_error(GALAX_ERR_ALLOC, "hipMalloc(%zu bytes) failed: %s", want, hipGetErrorString(e));
  }
  b.bytes = want;
  return 0;
}

void dbuf_free(DBuf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

// ------------------------------------------------------------------ scans ----

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_CHUNK = SCAN_THREADS * SCAN_ITEMS;

template <bool MAX>
__device__ __forceinline__ uint32_t sop(uint32_t a, uint32_t b) {
  return MAX ? (a > b ? a : b) : a + b;
}

// Exclusive scan of one value per thread over a 256-thread block; *tot gets
// the block reduction.
template <bool MAX>
__device__ uint32_t block_excl_scan(uint32_t v, uint32_t* tot) {
  __shared__ uint32_t wsum[SCAN_THREADS / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x = sop<MAX>(x, y);
  }
  uint32_t ex = __shfl_up(x, 1, 64);
  if (lane == 0) ex = 0;
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t woff = 0, t = 0;
#pragma unroll
  for (int i = 0; i < SCAN_THREADS / 64; ++i) {
    if (i < w) woff = sop<MAX>(woff, wsum[i]);
    t = sop<MAX>(t, wsum[i]);
  }
  __syncthreads();
  *tot = t;
  return sop<MAX>(woff, ex);
}

template <bool MAX>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_reduce(const uint32_t* __restrict__ in, size_t n,
                                                              uint32_t* __restrict__ part) {
  size_t base = (size_t)blockIdx.x * SCAN_CHUNK + (size_t)threadIdx.x * SCAN_ITEMS;
  uint32_t acc = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k)
    if (base + k < n) acc = sop<MAX>(acc, in[base + k]);
  uint32_t tot;
  (void)block_excl_scan<MAX>(acc, &tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

template <bool MAX>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_chunks(const uint32_t* in, uint32_t* out, size_t n,
                                                              const uint32_t* __restrict__ part_excl,
                                                              int inclusive, uint32_t* d_total) {
  size_t base = (size_t)blockIdx.x * SCAN_CHUNK + (size_t)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t acc = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0u;
    acc = sop<MAX>(acc, v[k]);
  }
  uint32_t tot;
  uint32_t ex = block_excl_scan<MAX>(acc, &tot);
  uint32_t pre = part_excl ? part_excl[blockIdx.x] : 0u;
  uint32_t run = sop<MAX>(pre, ex);
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    uint32_t nx = sop<MAX>(run, v[k]);
    if (base + k < n) out[base + k] = inclusive ? nx : run;
    run = nx;
  }
  if (d_total && blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_THREADS - 1) *d_total = sop<MAX>(pre, tot);
}

namespace {
int scan_rec(galax_fmm* h, const uint32_t* in, uint32_t* out, size_t n, bool inclusive, bool max_op,
             uint32_t* d_total, int depth) {
  if (n == 0) return 0;
  if (depth >= 6) return set_error(GALAX_ERR_INVALID, "scan too large");
  size_t nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
  hipStream_t s = h->stream;
  if (nb == 1) {
    if (max_op)
      k_scan_chunks<true><<<1, SCAN_THREADS, 0, s>>>(in, out, n, nullptr, inclusive, d_total);
    else
      k_scan_chunks<false><<<1, SCAN_THREADS, 0, s>>>(in, out, n, nullptr, inclusive, d_total);
    GX_HIP(hipGetLastError());
    return 0;
  }
  DBuf& pb = h->ws.scan[depth];
  GX_CHECK(dbuf_reserve(pb, nb * sizeof(uint32_t), s));
  uint32_t* part = pb.as<uint32_t>();
  if (max_op)
    k_scan_reduce<true><<<(unsigned)nb, SCAN_THREADS, 0, s>>>(in, n, part);
  else
    k_scan_reduce<false><<<(unsigned)nb, SCAN_THREADS, 0, s>>>(in, n, part);
  GX_HIP(hipGetLastError());
  GX_CHECK(scan_rec(h, part, part, nb, false, max_op, nullptr, depth + 1));
  if (max_op)
    k_scan_chunks<true><<<(unsigned)nb, SCAN_THREADS, 0, s>>>(in, out, n, part, inclusive, d_total);
  else
    k_scan_chunks<false><<<(unsigned)nb, SCAN_THREADS, 0, s>>>(in, out, n, part, inclusive, d_total);
  GX_HIP(hipGetLastError());
  return 0;
}
}  // namespace

int scan_u32(galax_fmm* h, const uint32_t* d_in, uint32_t* d_out, size_t n, bool inclusive, bool max_op,
             uint32_t* total) {
  if (n == 0) {
    if (total) *total = 0;
    return 0;
  }
  uint32_t* d_total = nullptr;
  if (total) {
    GX_CHECK(dbuf_reserve(h->flag, 64, h->stream));
    d_total = h->flag.as<uint32_t>() + 8;
  }
  GX_CHECK(scan_rec(h, d_in, d_out, n, inclusive, max_op, d_total, 0));
  if (total) {
    GX_HIP(hipMemcpyAsync(total, d_total, sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
    GX_HIP(hipStreamSynchronize(h->stream));
  }
  return 0;
}

// 64-bit sum of n uint32 values: a block sum per 2048 values, then one
// 64-bit atomic add per block (the order of the adds does not change an
// integer sum).
__global__ void k_sum_u32_u64(const uint32_t* __restrict__ in, size_t n, unsigned long long* __restrict__ out) {
  __shared__ unsigned long long part[SCAN_THREADS / 64];
  const size_t base = (size_t)blockIdx.x * SCAN_CHUNK;
  unsigned long long v = 0;
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const size_t i = base + (size_t)k * SCAN_THREADS + threadIdx.x;
    if (i < n) v += in[i];
  }
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_down(v, d, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < SCAN_THREADS / 64; ++w) t += part[w];
    atomicAdd(out, t);
  }
}

int sum_u32(galax_fmm* h, const uint32_t* d_in, size_t n, uint64_t* total) {
  *total = 0;
  if (n == 0) return 0;
  GX_CHECK(dbuf_reserve(h->flag, 64, h->stream));
  unsigned long long* d = reinterpret_cast<unsigned long long*>(h->flag.as<uint32_t>() + 14);
  GX_HIP(hipMemsetAsync(d, 0, 8, h->stream));
  k_sum_u32_u64<<<(unsigned)((n + SCAN_CHUNK - 1) / SCAN_CHUNK), SCAN_THREADS, 0, h->stream>>>(d_in, n, d);
  GX_HIP(hipGetLastError());
  unsigned long long v = 0;
  GX_HIP(hipMemcpyAsync(&v, d, 8, hipMemcpyDeviceToHost, h->stream));
  GX_HIP(hipStreamSynchronize(h->stream));
  *total = v;
  return 0;
}

int build_mark(galax_fmm* h, int i) {
  if (h->build_timed) GX_HIP(hipEventRecord(h->bev[i], h->stream));
  return 0;
}

}  // namespace gx
