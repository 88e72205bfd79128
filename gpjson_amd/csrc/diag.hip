// diag.hip -- streaming copy and streaming read used to measure the achievable
// HBM bandwidth of the device, and the copy to calibrate rocprofv3's
// FETCH_SIZE / WRITE_SIZE counters against a kernel whose traffic is known
// exactly (n bytes read, n written).  No counterpart in the reference.
#include <hip/hip_runtime.h>

#include "runtime.h"

namespace gj {
namespace {

// Vectors of 16 bytes per thread per pass, and one pass over the buffer: the
// fastest shapes of a sweep over grids and vectors per thread on a 1 GiB
// buffer (copy: 4.6 TB/s with one vector per thread on 32 blocks per CU, 5.8
// TB/s with 4; read: 6.3 TB/s on 32 blocks per CU, 6.7 TB/s in one pass).
constexpr int kCopyUnroll = 4;
constexpr int kReadUnroll = 1;

// Copy of 16-byte vectors, kCopyUnroll per thread per pass (a block's passes
// cover consecutive vectors); n16 vectors, then a byte tail.
__global__ __launch_bounds__(256) void stream_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                          int64_t n16, const uint8_t* __restrict__ src_b,
                                                          uint8_t* __restrict__ dst_b, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x * kCopyUnroll;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x * kCopyUnroll + threadIdx.x; i < n16; i += stride) {
    uint4 v[kCopyUnroll];
#pragma unroll
    for (int u = 0; u < kCopyUnroll; u++) {
      const int64_t k = i + (int64_t)u * blockDim.x;
      if (k < n16) v[u] = src[k];
    }
#pragma unroll
    for (int u = 0; u < kCopyUnroll; u++) {
      const int64_t k = i + (int64_t)u * blockDim.x;
      if (k < n16) dst[k] = v[u];
    }
  }
  const int64_t t = n16 * 16 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst_b[t] = src_b[t];
}

// Read of 16-byte vectors (kReadUnroll per thread per pass), folded with XOR
// per thread; the fold is stored only if it matches `never` (the host passes a
// value the data cannot produce), so every load is kept but nothing is written.
__global__ __launch_bounds__(256) void stream_read_kernel(const uint4* __restrict__ src, int64_t n16, uint32_t never,
                                                          uint32_t* __restrict__ sink) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x * kReadUnroll;
  uint4 acc = make_uint4(0, 0, 0, 0);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x * kReadUnroll + threadIdx.x; i < n16; i += stride) {
#pragma unroll
    for (int u = 0; u < kReadUnroll; u++) {
      const int64_t k = i + (int64_t)u * blockDim.x;
      if (k < n16) {
        const uint4 v = src[k];
        acc.x ^= v.x;
        acc.y ^= v.y;
        acc.z ^= v.z;
        acc.w ^= v.w;
      }
    }
  }
  const uint32_t f = acc.x ^ acc.y ^ acc.z ^ acc.w;
  if (f == never) sink[0] = f;
}

// One pass over the whole buffer.
static int probe_grid(int64_t n16, int unroll) {
  const int64_t per_block = 256 * (int64_t)unroll;
  return (int)std::min<int64_t>(std::max<int64_t>((n16 + per_block - 1) / per_block, 1), 1 << 30);
}

// Times `iterations` launches of `launch` with HIP events on ctx's stream.
template <class Launch>
int time_launches(gpjson_ctx* ctx, int iterations, Launch launch, const char* what, double* ms_per_launch) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = GPJSON_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = fail(GPJSON_ERR_HIP, "hipEventCreate");
  if (rc == GPJSON_OK && launch() != hipSuccess) rc = fail(GPJSON_ERR_HIP, what);  // warm-up
  if (rc == GPJSON_OK) {
    (void)hipEventRecord(e0, ctx->stream);
    for (int i = 0; i < iterations && rc == GPJSON_OK; i++)
      if (launch() != hipSuccess) rc = fail(GPJSON_ERR_HIP, what);
    (void)hipEventRecord(e1, ctx->stream);
  }
  if (rc == GPJSON_OK) {
    hipError_t e = hipEventSynchronize(e1);
    if (e != hipSuccess) rc = hip_fail(e, "hipEventSynchronize");
  }
  if (rc == GPJSON_OK) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *ms_per_launch = ms / iterations;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(ctx->stream);
  return rc;
}

}  // namespace
}  // namespace gj

using namespace gj;

extern "C" {

int gpjson_time_copy(gpjson_ctx* ctx, int64_t n, int iterations, double* ms_per_copy) {
  if (!ctx || n <= 0 || iterations < 1 || !ms_per_copy) return fail(GPJSON_ERR_INVALID_ARG, "bad argument");
  GJ_HIP(hipSetDevice(ctx->device));
  DevBuf a, b;
  int rc = a.ensure((size_t)n);
  if (rc == GPJSON_OK) rc = b.ensure((size_t)n);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == GPJSON_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(GPJSON_ERR_HIP, "hipEventCreate");
  const int64_t n16 = n / 16;
  const int grid = probe_grid(n16, kCopyUnroll);
  auto launch = [&]() {
    hipLaunchKernelGGL(stream_copy_kernel, dim3(grid), dim3(256), 0, ctx->stream, (const uint4*)a.p, (uint4*)b.p,
                       n16, (const uint8_t*)a.p, (uint8_t*)b.p, n);
    return hipGetLastError();
  };
  if (rc == GPJSON_OK && hipMemsetAsync(a.p, 0x5a, (size_t)n, ctx->stream) != hipSuccess)
    rc = fail(GPJSON_ERR_HIP, "hipMemsetAsync");
  if (rc == GPJSON_OK && launch() != hipSuccess) rc = fail(GPJSON_ERR_HIP, "stream_copy_kernel");  // warm-up
  if (rc == GPJSON_OK) {
    (void)hipEventRecord(e0, ctx->stream);
    for (int i = 0; i < iterations && rc == GPJSON_OK; i++)
      if (launch() != hipSuccess) rc = fail(GPJSON_ERR_HIP, "stream_copy_kernel");
    (void)hipEventRecord(e1, ctx->stream);
  }
  if (rc == GPJSON_OK) {
    hipError_t e = hipEventSynchronize(e1);
    if (e != hipSuccess) rc = hip_fail(e, "hipEventSynchronize");
  }
  if (rc == GPJSON_OK) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *ms_per_copy = ms / iterations;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(ctx->stream);
  a.release();
  b.release();
  return rc;
}

int gpjson_time_read(gpjson_ctx* ctx, int64_t n, int iterations, double* ms_per_read) {
  if (!ctx || n < 16 || iterations < 1 || !ms_per_read) return fail(GPJSON_ERR_INVALID_ARG, "bad argument");
  GJ_HIP(hipSetDevice(ctx->device));
  DevBuf a, sink;
  int rc = a.ensure((size_t)n);
  if (rc == GPJSON_OK) rc = sink.ensure(256);
  // bytes of 0x5a: every 16-byte vector is the same, so each thread's XOR fold
  // is 0 or 0x5a5a5a5a per word and its final fold is 0, never 1
  if (rc == GPJSON_OK && hipMemsetAsync(a.p, 0x5a, (size_t)n, ctx->stream) != hipSuccess)
    rc = fail(GPJSON_ERR_HIP, "hipMemsetAsync");
  const int64_t n16 = n / 16;
  const int grid = probe_grid(n16, kReadUnroll);
  if (rc == GPJSON_OK)
    rc = time_launches(ctx, iterations, [&]() {
      hipLaunchKernelGGL(stream_read_kernel, dim3(grid), dim3(256), 0, ctx->stream, (const uint4*)a.p, n16, 1u,
                         (uint32_t*)sink.p);
      return hipGetLastError();
    }, "stream_read_kernel", ms_per_read);
  (void)hipStreamSynchronize(ctx->stream);
  a.release();
  sink.release();
  return rc;
}

}  // extern "C"
