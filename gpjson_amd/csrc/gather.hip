// gather.hip -- device-side gather of query results into one compact buffer.
//
// The reference materializes every result element on the host: it copies the
// span array back, maps the file a second time and decodes each element's
// bytes from the map on access (J/functions/QueryFunction.java:240,247-259,
// J/result/ResultItem.java:32-52).  Here the result bytes are gathered on the
// device right after the interpreter: element i = spans[i] = (start, end)
// gets bytes [off_i, off_i + end - start) of one contiguous buffer, where off
// is the exclusive prefix sum of the element lengths (0 for "no result"), and
// the caller copies that buffer and the rewritten spans back once.
//
// Kernels (all grid-stride, HBM/latency bound; nothing is compute heavy):
//   gather_sum_kernel    per-block sums of the element lengths
//   gather_base_kernel   exclusive scan of the block sums (one block)
//   gather_spans_kernel  per-element offsets: rewritten spans (-1 kept)
//   gather_copy_kernel   the bytes: a lane per short element, the whole
//                        wavefront for each element longer than 64 bytes
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gj {

namespace {

constexpr int kGatherBlock = 256;
constexpr int kGatherItems = 8;  // elements per thread in the scan kernels
constexpr int kGatherTile = kGatherBlock * kGatherItems;
constexpr int kShortBytes = 64;  // elements up to this long are copied by one lane

__device__ __forceinline__ int64_t elem_len(const int64_t* spans, int64_t i, int64_t n, int64_t N) {
  if (i >= n) return 0;
  const int64_t a = spans[2 * i], b = spans[2 * i + 1];
  if (a < 0) return 0;
  // spans lie inside the file; clamp as the host-side decoder does
  const int64_t lo = a < N ? a : N;
  const int64_t hi = b < lo ? lo : (b < N ? b : N);
  return hi - lo;
}

// Block-wide exclusive scan of one int64 per thread; returns the block total.
__device__ int64_t block_exclusive_scan(int64_t v, int64_t* excl) {
  __shared__ int64_t s_wave[kGatherBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned lo = __shfl_up((unsigned)(uint64_t)x, off), hi = __shfl_up((unsigned)((uint64_t)x >> 32), off);
    const int64_t y = (int64_t)(((uint64_t)hi << 32) | lo);
    if (lane >= off) x += y;
  }
  if (lane == 63) s_wave[wave] = x;
  __syncthreads();
  int64_t before = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kGatherBlock / 64; w++) {
    if (w < wave) before += s_wave[w];
    total += s_wave[w];
  }
  *excl = before + x - v;
  __syncthreads();
  return total;
}

}  // namespace

__global__ __launch_bounds__(kGatherBlock) void gather_sum_kernel(const int64_t* spans, int64_t n, int64_t N,
                                                                  int64_t* block_sums) {
  const int64_t base = (int64_t)blockIdx.x * kGatherTile + (int64_t)threadIdx.x * kGatherItems;
  int64_t t = 0;
#pragma unroll
  for (int k = 0; k < kGatherItems; k++) t += elem_len(spans, base + k, n, N);
  int64_t excl;
  const int64_t total = block_exclusive_scan(t, &excl);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// One block: exclusive scan of nb block sums in place; total_out = their sum.
__global__ __launch_bounds__(kGatherBlock) void gather_base_kernel(int64_t* block_sums, int64_t nb,
                                                                   int64_t* total_out) {
  int64_t carry = 0;
  for (int64_t c0 = 0; c0 < nb; c0 += kGatherBlock) {
    const int64_t i = c0 + threadIdx.x;
    const int64_t v = i < nb ? block_sums[i] : 0;
    int64_t excl;
    const int64_t total = block_exclusive_scan(v, &excl);
    if (i < nb) block_sums[i] = carry + excl;
    carry += total;
  }
  if (threadIdx.x == 0) *total_out = carry;
}

__global__ __launch_bounds__(kGatherBlock) void gather_spans_kernel(const int64_t* spans, int64_t n, int64_t N,
                                                                    const int64_t* block_base, int64_t* out_spans) {
  const int64_t base = (int64_t)blockIdx.x * kGatherTile + (int64_t)threadIdx.x * kGatherItems;
  int64_t len[kGatherItems];
  int64_t t = 0;
#pragma unroll
  for (int k = 0; k < kGatherItems; k++) {
    len[k] = elem_len(spans, base + k, n, N);
    t += len[k];
  }
  int64_t excl;
  block_exclusive_scan(t, &excl);
  int64_t off = block_base[blockIdx.x] + excl;
#pragma unroll
  for (int k = 0; k < kGatherItems; k++) {
    const int64_t i = base + k;
    if (i < n) {
      const bool null = spans[2 * i] < 0;
      out_spans[2 * i] = null ? -1 : off;
      out_spans[2 * i + 1] = null ? -1 : off + len[k];
    }
    off += len[k];
  }
}

__global__ __launch_bounds__(kGatherBlock) void gather_copy_kernel(const uint8_t* F, int64_t N, const int64_t* spans,
                                                                   const int64_t* out_spans, int64_t n,
                                                                   uint8_t* out) {
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * (kGatherBlock / 64);
  const int64_t wave_id = (int64_t)blockIdx.x * (kGatherBlock / 64) + (threadIdx.x >> 6);
  for (int64_t w0 = wave_id * 64; w0 < n; w0 += nwaves * 64) {
    const int64_t i = w0 + lane;
    int64_t src = 0, dst = 0, len = 0;
    if (i < n && out_spans[2 * i] >= 0) {
      dst = out_spans[2 * i];
      len = out_spans[2 * i + 1] - dst;
      src = spans[2 * i] < N ? spans[2 * i] : N;
    }
    if (len <= kShortBytes) {
      for (int64_t k = 0; k < len; k++) out[dst + k] = F[src + k];
    }
    // the long elements of these 64, one at a time with the whole wavefront
    unsigned long long longs = __ballot(len > kShortBytes);
    while (longs) {
      const int l = __builtin_ctzll(longs);
      longs &= longs - 1;
      const int64_t lsrc = __shfl(src, l), ldst = __shfl(dst, l), llen = __shfl(len, l);
      for (int64_t k = lane; k < llen; k += 64) out[ldst + k] = F[lsrc + k];
    }
  }
}

// Gathers the n elements of spans (int64[n][2] into the input F of N bytes)
// into out_bytes; out_spans receives the element offsets.  Phase 1 computes the
// total (written to *total_dev); the caller sizes out_bytes from it and calls
// phase 2.  block_sums must hold gather_blocks(n) int64.
int64_t gather_blocks(int64_t n) { return (n + kGatherTile - 1) / kGatherTile; }

hipError_t gather_phase1(const int64_t* spans, int64_t n, int64_t N, int64_t* block_sums, int64_t* total_dev,
                         hipStream_t stream) {
  const int64_t nb = gather_blocks(n);
  if (nb > 0)
    hipLaunchKernelGGL(gather_sum_kernel, dim3((unsigned)nb), dim3(kGatherBlock), 0, stream, spans, n, N,
                       block_sums);
  hipLaunchKernelGGL(gather_base_kernel, dim3(1), dim3(kGatherBlock), 0, stream, block_sums, nb, total_dev);
  return hipGetLastError();
}

hipError_t gather_phase2(const uint8_t* F, int64_t N, const int64_t* spans, int64_t n, const int64_t* block_sums,
                         int64_t* out_spans, uint8_t* out_bytes, int copy_grid, hipStream_t stream) {
  const int64_t nb = gather_blocks(n);
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(gather_spans_kernel, dim3((unsigned)nb), dim3(kGatherBlock), 0, stream, spans, n, N,
                     block_sums, out_spans);
  hipLaunchKernelGGL(gather_copy_kernel, dim3(copy_grid > 0 ? copy_grid : 1), dim3(kGatherBlock), 0, stream, F, N,
                     spans, out_spans, n, out_bytes);
  return hipGetLastError();
}

}  // namespace gj
