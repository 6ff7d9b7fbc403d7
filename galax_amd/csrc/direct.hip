// direct.hip — the softened direct sum on the device (compute_n2_force,
// galax-core/src/lib.rs:1685-1721; direct_potential's sign, :1627-1635), FP64,
// for all bodies or a list of sampled targets.
//
// Work split. A target list of 10,000 samples is 40 blocks of 256 targets:
// far too few for 256 CUs. The sources are therefore cut into chunks; block
// (tile, chunk) sums its chunk for 256 targets and writes one partial per
// target and component to a scratch buffer, and a combine pass adds a
// target's partials in chunk order. The plan (chunk count and length) is a
// function of (n, n_targets) only, so every entry point gives the same sums
// in the same order: bit-identical results.
//
// Entry points:
//   galax_direct_sum          scratch (device copies of the inputs, partials)
//                             from hipMalloc, freed before return;
//   galax_direct_sum_noalloc  the same kernels on a file-scope device array of
//                             DIRECT_SCRATCH_BYTES; allocates nothing.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "common.h"

namespace gx {

namespace {

constexpr int DT = 256;             // targets per block, sources per LDS stage
constexpr size_t MIN_CHUNK = 2048;  // sources per chunk at least
constexpr size_t WANT_BLOCKS = 2048;

struct DirectPlan {
  size_t tiles, chunks, chunk_len;
};

DirectPlan direct_plan(size_t n, size_t nt) {
  DirectPlan p;
  p.tiles = (nt + DT - 1) / DT;
  size_t c = p.tiles ? (WANT_BLOCKS + p.tiles - 1) / p.tiles : 1;
  c = std::max<size_t>(1, std::min(c, n / MIN_CHUNK));
  size_t len = n ? (n + c - 1) / c : DT;
  len = (len + DT - 1) / DT * DT;
  p.chunk_len = len;
  p.chunks = n ? (n + len - 1) / len : 1;
  return p;
}

// Block (tile, chunk): targets tile*256 .. +256, sources of the chunk. With
// part == nullptr (a single chunk) the sums go straight to ax, ay, phi;
// otherwise to part[(chunk * 3 + q) * nt + k].
__global__ void __launch_bounds__(DT) k_direct_chunk(const double* __restrict__ x, const double* __restrict__ y,
                                                     const double* __restrict__ m, size_t n, double eps2,
                                                     const uint64_t* __restrict__ targets, size_t nt, size_t chunk_len,
                                                     double* __restrict__ part, double* __restrict__ ax,
                                                     double* __restrict__ ay, double* __restrict__ phi) {
  __shared__ double sx[DT], sy[DT], sm[DT];
  const size_t k = (size_t)blockIdx.x * DT + threadIdx.x;
  const bool valid = k < nt;
  const size_t i = valid ? (targets ? (size_t)targets[k] : k) : 0;
  const double xi = valid ? x[i] : 0.0, yi = valid ? y[i] : 0.0;
  const size_t c0 = (size_t)blockIdx.y * chunk_len;
  const size_t c1 = std::min(n, c0 + chunk_len);
  double accx = 0.0, accy = 0.0, accp = 0.0;
  for (size_t b = c0; b < c1; b += DT) {
    __syncthreads();
    const size_t j = b + threadIdx.x;
    if (j < c1) {
      sx[threadIdx.x] = x[j];
      sy[threadIdx.x] = y[j];
      sm[threadIdx.x] = m[j];
    }
    __syncthreads();
    const size_t cnt = std::min<size_t>(c1 - b, DT);
    if (valid) {
      for (size_t jj = 0; jj < cnt; ++jj) {
        if (b + jj == i) continue;  // no self term (a zero force, but -m/eps in phi)
        const double dx = sx[jj] - xi, dy = sy[jj] - yi;
        const double r2 = dx * dx + dy * dy + eps2;
        const double inv = 1.0 / sqrt(r2);
        const double f = sm[jj] * (inv * inv * inv);
        accx += f * dx;
        accy += f * dy;
        accp -= sm[jj] * inv;
      }
    }
  }
  if (!valid) return;
  if (part) {
    double* p = part + (size_t)blockIdx.y * 3 * nt;
    p[k] = accx;
    p[nt + k] = accy;
    p[2 * nt + k] = accp;
  } else {
    ax[k] = accx;
    ay[k] = accy;
    if (phi) phi[k] = accp;
  }
}

// Thread per target: its partials added in chunk order.
__global__ void k_direct_combine(const double* __restrict__ part, size_t nt, size_t chunks, double* __restrict__ ax,
                                 double* __restrict__ ay, double* __restrict__ phi) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nt) return;
  double sx = 0.0, sy = 0.0, sp = 0.0;
  for (size_t c = 0; c < chunks; ++c) {
    const double* p = part + c * 3 * nt;
    sx += p[k];
    sy += p[nt + k];
    sp += p[2 * nt + k];
  }
  ax[k] = sx;
  ay[k] = sy;
  if (phi) phi[k] = sp;
}

// Bytes of partials the plan needs (0 for one chunk).
size_t part_bytes(const DirectPlan& p, size_t nt) { return p.chunks > 1 ? p.chunks * 3 * nt * sizeof(double) : 0; }

int direct_launch(hipStream_t s, const DirectPlan& p, const double* x, const double* y, const double* m, size_t n,
                  double eps, const uint64_t* targets, size_t nt, double* part, double* ax, double* ay,
                  double* phi) {
  if (nt == 0) return 0;
  const bool split = p.chunks > 1;
  k_direct_chunk<<<dim3((unsigned)p.tiles, (unsigned)p.chunks), DT, 0, s>>>(
      x, y, m, n, eps * eps, targets, nt, p.chunk_len, split ? part : nullptr, ax, ay, phi);
  GX_HIP(hipGetLastError());
  if (split) {
    k_direct_combine<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(part, nt, p.chunks, ax, ay, phi);
    GX_HIP(hipGetLastError());
  }
  return 0;
}

int check_args(const double* x, const double* y, const double* m, size_t n, const uint64_t* targets, size_t nt,
               double* ax, double* ay) {
  if (n && (!x || !y || !m)) return set_error(GALAX_ERR_INVALID, "x, y, m are required");
  if (nt && (!ax || !ay)) return set_error(GALAX_ERR_INVALID, "ax, ay are required");
  if (n >= (1ull << 40) || nt >= (1ull << 40)) return set_error(GALAX_ERR_INVALID, "n too large");
  if (targets)
    for (size_t k = 0; k < nt; ++k)
      if (targets[k] >= n) return set_error(GALAX_ERR_INVALID, "target index %llu >= n", (unsigned long long)targets[k]);
  return 0;
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

// ------------------------------------------------------ no-alloc scratch ----

// The allocation-free entry point's device memory: staging for the inputs and
// outputs plus the partial sums. One copy per GPU, held until the process ends.
constexpr size_t DIRECT_SCRATCH_BYTES = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[DIRECT_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy at its own address). Host code uses
// this address, never the array's name, for kernel arguments and copies.
int scratch_address(void** out) {
  int dev = 0;
  GX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return set_error(GALAX_ERR_UNSUPPORTED, "device %d beyond %d", dev, MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(GALAX_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                       hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}
}  // namespace

}  // namespace gx

extern "C" {

int galax_direct_sum(const double* x, const double* y, const double* m, size_t n, double softening,
                     const uint64_t* targets, size_t n_targets, double* ax, double* ay, double* phi) {
  using namespace gx;
  const size_t nt = targets ? n_targets : n;
  GX_CHECK(check_args(x, y, m, n, targets, nt, ax, ay));
  const DirectPlan plan = direct_plan(n, nt);
  hipStream_t s;
  GX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  double *dx = nullptr, *dy = nullptr, *dm = nullptr, *dax = nullptr, *day = nullptr, *dphi = nullptr;
  double* dpart = nullptr;
  uint64_t* dt = nullptr;
  int rc = 0;
  auto fail = [&](hipError_t e) {
    if (e != hipSuccess && rc == 0) rc = set_error(GALAX_ERR_DEVICE, "direct sum: %s", hipGetErrorString(e));
  };
  const size_t nb = std::max<size_t>(n, 1) * 8, nbt = std::max<size_t>(nt, 1) * 8;
  fail(hipMalloc(&dx, nb));
  fail(hipMalloc(&dy, nb));
  fail(hipMalloc(&dm, nb));
  fail(hipMalloc(&dax, nbt));
  fail(hipMalloc(&day, nbt));
  fail(hipMalloc(&dphi, nbt));
  if (targets) fail(hipMalloc(&dt, nbt));
  if (part_bytes(plan, nt)) fail(hipMalloc(&dpart, part_bytes(plan, nt)));
  if (rc == 0 && n) {
    fail(hipMemcpyAsync(dx, x, n * 8, hipMemcpyHostToDevice, s));
    fail(hipMemcpyAsync(dy, y, n * 8, hipMemcpyHostToDevice, s));
    fail(hipMemcpyAsync(dm, m, n * 8, hipMemcpyHostToDevice, s));
    if (targets) fail(hipMemcpyAsync(dt, targets, nt * 8, hipMemcpyHostToDevice, s));
  }
  if (rc == 0) rc = direct_launch(s, plan, dx, dy, dm, n, softening, dt, nt, dpart, dax, day, dphi);
  if (rc == 0 && nt) {
    fail(hipMemcpyAsync(ax, dax, nt * 8, hipMemcpyDeviceToHost, s));
    fail(hipMemcpyAsync(ay, day, nt * 8, hipMemcpyDeviceToHost, s));
    if (phi) fail(hipMemcpyAsync(phi, dphi, nt * 8, hipMemcpyDeviceToHost, s));
  }
  fail(hipStreamSynchronize(s));
  for (void* p : {(void*)dx, (void*)dy, (void*)dm, (void*)dax, (void*)day, (void*)dphi, (void*)dt, (void*)dpart})
    if (p) (void)hipFree(p);
  (void)hipStreamDestroy(s);
  return rc;
}

size_t galax_direct_sum_noalloc_bytes(size_t n, size_t n_targets) {
  using namespace gx;
  const DirectPlan plan = direct_plan(n, n_targets);
  return 3 * align256(n * 8) + 4 * align256(n_targets * 8) + align256(part_bytes(plan, n_targets));
}

int galax_direct_sum_noalloc(const double* x, const double* y, const double* m, size_t n, double softening,
                             const uint64_t* targets, size_t n_targets, double* ax, double* ay, double* phi) {
  using namespace gx;
  const size_t nt = targets ? n_targets : n;
  GX_CHECK(check_args(x, y, m, n, targets, nt, ax, ay));
  const size_t need = galax_direct_sum_noalloc_bytes(n, nt);
  if (need > DIRECT_SCRATCH_BYTES)
    return set_error(GALAX_ERR_INVALID,
                     "direct sum of %zu bodies at %zu targets needs %zu bytes of scratch, above the %zu-byte maximum "
                     "of galax_direct_sum_noalloc",
                     n, nt, need, DIRECT_SCRATCH_BYTES);
  void* base = nullptr;
  GX_CHECK(scratch_address(&base));
  unsigned char* p = static_cast<unsigned char*>(base);
  auto take = [&](size_t bytes) {
    unsigned char* q = p;
    p += align256(bytes);
    return q;
  };
  double* dx = reinterpret_cast<double*>(take(n * 8));
  double* dy = reinterpret_cast<double*>(take(n * 8));
  double* dm = reinterpret_cast<double*>(take(n * 8));
  uint64_t* dt = reinterpret_cast<uint64_t*>(take(nt * 8));
  double* dax = reinterpret_cast<double*>(take(nt * 8));
  double* day = reinterpret_cast<double*>(take(nt * 8));
  double* dphi = reinterpret_cast<double*>(take(nt * 8));
  const DirectPlan plan = direct_plan(n, nt);
  double* dpart = reinterpret_cast<double*>(take(part_bytes(plan, nt)));
  hipStream_t s = hipStreamPerThread;  // no stream is created
  if (n) {
    GX_HIP(hipMemcpyAsync(dx, x, n * 8, hipMemcpyHostToDevice, s));
    GX_HIP(hipMemcpyAsync(dy, y, n * 8, hipMemcpyHostToDevice, s));
    GX_HIP(hipMemcpyAsync(dm, m, n * 8, hipMemcpyHostToDevice, s));
    if (targets && nt) GX_HIP(hipMemcpyAsync(dt, targets, nt * 8, hipMemcpyHostToDevice, s));
  }
  GX_CHECK(direct_launch(s, plan, dx, dy, dm, n, softening, targets ? dt : nullptr, nt, dpart, dax, day, dphi));
  if (nt) {
    GX_HIP(hipMemcpyAsync(ax, dax, nt * 8, hipMemcpyDeviceToHost, s));
    GX_HIP(hipMemcpyAsync(ay, day, nt * 8, hipMemcpyDeviceToHost, s));
    if (phi) GX_HIP(hipMemcpyAsync(phi, dphi, nt * 8, hipMemcpyDeviceToHost, s));
  }
  GX_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // extern "C"
