// Random-access HBM roofline for FM-index rank work on gfx950.
//
// The seeding kernels spend their time on dependent random reads of index
// lines (one occ line per backward-extension step, one line per LF step of
// the suffix-array walk).  The bound that matters is therefore not the
// streaming HBM bandwidth but the rate at which the memory system serves
// random, line-sized reads.  This file measures that rate two ways:
//   * independent gathers: every lane keeps `INFLIGHT` independent line reads
//     outstanding (throughput ceiling, the R_max used for roofline fractions);
//   * dependent chains: every lane follows a pointer chain, the next line
//     index depending on the data just read (the access pattern of one
//     extension stream), which measures loaded latency x concurrency.
// A plain streaming copy is provided to calibrate hardware byte counters
// (FETCH_SIZE / WRITE_SIZE) against a kernel of known traffic.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "gpuseed.h"

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  // splitmix64 finaliser (Steele, Lea, Flood 2014)
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t scale(uint64_t r, uint64_t n) {
  return __umul64hi(r, n);
}

template <int LINE, int INFLIGHT>
__global__ void __launch_bounds__(256) gather_independent(
    const uint4* __restrict__ buf, uint64_t nlines, int rounds, uint64_t seed,
    uint32_t* __restrict__ sink) {
  constexpr int W = LINE / 16;
  const uint64_t gid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t s = mix64(seed ^ (gid * 0x632be59bd9b4e019ull));
  uint32_t acc = 0;
  for (int r = 0; r < rounds; ++r) {
    uint4 v[INFLIGHT][W];
#pragma unroll
    for (int j = 0; j < INFLIGHT; ++j) {
      s = mix64(s);
      const uint4* p = buf + scale(s, nlines) * W;
#pragma unroll
      for (int w = 0; w < W; ++w) v[j][w] = p[w];
    }
#pragma unroll
    for (int j = 0; j < INFLIGHT; ++j)
#pragma unroll
      for (int w = 0; w < W; ++w) acc += v[j][w].x ^ v[j][w].y ^ v[j][w].z ^ v[j][w].w;
  }
  if (acc == 0x9e3779b9u) sink[gid & 1023] = acc;
}

template <int LINE>
__global__ void __launch_bounds__(256) gather_dependent(
    const uint4* __restrict__ buf, uint64_t nlines, int steps, uint64_t seed,
    uint32_t* __restrict__ sink) {
  constexpr int W = LINE / 16;
  const uint64_t gid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t s = mix64(seed ^ (gid * 0x632be59bd9b4e019ull));
  uint32_t acc = 0;
  for (int i = 0; i < steps; ++i) {
    const uint4* p = buf + scale(s, nlines) * W;
    uint32_t x = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      uint4 v = p[w];
      x ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    acc += x;
    s = mix64(s ^ x);  // next address depends on the data just read
  }
  if (acc == 0x9e3779b9u) sink[gid & 1023] = acc;
}

// One 16-byte chunk at a random offset inside a random LINE-byte line (mode 2):
// the lane requests 16 B but the line it lands in is LINE-aligned.
template <int LINE>
__global__ void __launch_bounds__(256) gather_subline(
    const uint4* __restrict__ buf, uint64_t nlines, int iters, uint64_t seed,
    uint32_t* __restrict__ sink) {
  constexpr int W = LINE / 16;
  const uint64_t gid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t s = mix64(seed ^ (gid * 0x632be59bd9b4e019ull));
  uint32_t acc = 0;
  for (int i = 0; i < iters; ++i) {
    s = mix64(s);
    uint4 v = buf[scale(s, nlines) * W + (s & (W - 1))];
    acc += v.x ^ v.y ^ v.z ^ v.w;
  }
  if (acc == 0x9e3779b9u) sink[gid & 1023] = acc;
}

// Cooperative line reads (mode 3): W = LINE/16 adjacent lanes read one line
// together, lane j taking bytes [16j, 16j+16), so each wave instruction issues
// 64/W coalesced line requests.
template <int LINE>
__global__ void __launch_bounds__(256) gather_coop(
    const uint4* __restrict__ buf, uint64_t nlines, int iters, uint64_t seed,
    uint32_t* __restrict__ sink) {
  constexpr int W = LINE / 16;
  const uint64_t gid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t grp = gid / W;
  const int j = (int)(gid % W);
  uint64_t s = mix64(seed ^ (grp * 0x632be59bd9b4e019ull));
  uint32_t acc = 0;
  for (int i = 0; i < iters; ++i) {
    s = mix64(s);
    uint4 v = buf[scale(s, nlines) * W + j];
    acc += v.x ^ v.y ^ v.z ^ v.w;
  }
  if (acc == 0x9e3779b9u) sink[gid & 1023] = acc;
}

// Random 4-byte words (mode 4): one dword per lane per read.
__global__ void __launch_bounds__(256) gather_dword(
    const uint32_t* __restrict__ buf, uint64_t nwords, int iters, uint64_t seed,
    uint32_t* __restrict__ sink) {
  const uint64_t gid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint64_t s = mix64(seed ^ (gid * 0x632be59bd9b4e019ull));
  uint32_t acc = 0;
  for (int i = 0; i < iters; ++i) {
    s = mix64(s);
    acc += buf[scale(s, nwords)];
  }
  if (acc == 0x9e3779b9u) sink[gid & 1023] = acc;
}

__global__ void fill_random(uint4* __restrict__ buf, uint64_t n16, uint64_t seed) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n16; i += stride) {
    uint64_t a = mix64(seed + 2 * i), b = mix64(seed + 2 * i + 1);
    buf[i] = make_uint4((uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32));
  }
}

__global__ void stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n16) {
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n16; i += stride) dst[i] = src[i];
}

template <int LINE>
void launch_gather(int mode, int inflight, int grid, const uint4* buf, uint64_t nlines, int iters,
                   uint64_t seed, uint32_t* sink, hipStream_t st) {
  if (mode == 1) {
    hipLaunchKernelGGL((gather_dependent<LINE>), dim3(grid), dim3(256), 0, st, buf, nlines, iters,
                       seed, sink);
    return;
  }
  if (mode == 2) {
    hipLaunchKernelGGL((gather_subline<LINE>), dim3(grid), dim3(256), 0, st, buf, nlines, iters,
                       seed, sink);
    return;
  }
  if (mode == 3) {
    hipLaunchKernelGGL((gather_coop<LINE>), dim3(grid), dim3(256), 0, st, buf, nlines, iters, seed,
                       sink);
    return;
  }
  if (mode == 4) {
    hipLaunchKernelGGL(gather_dword, dim3(grid), dim3(256), 0, st, (const uint32_t*)buf,
                       nlines * (LINE / 4), iters, seed, sink);
    return;
  }
  int rounds = iters / inflight;
  switch (inflight) {
    case 1:
      hipLaunchKernelGGL((gather_independent<LINE, 1>), dim3(grid), dim3(256), 0, st, buf, nlines,
                         rounds, seed, sink);
      break;
    case 2:
      hipLaunchKernelGGL((gather_independent<LINE, 2>), dim3(grid), dim3(256), 0, st, buf, nlines,
                         rounds, seed, sink);
      break;
    default:
      hipLaunchKernelGGL((gather_independent<LINE, 4>), dim3(grid), dim3(256), 0, st, buf, nlines,
                         rounds, seed, sink);
      break;
  }
}

}  // namespace

extern "C" int gs_roofline_random(uint64_t buffer_bytes, int line_bytes, int mode, int inflight,
                                  int grid, int iters, int reps, gs_roofline_result* out) {
  if (!out || buffer_bytes < (1u << 20) || iters <= 0 || reps <= 0 || grid <= 0) return GS_EINVAL;
  if (line_bytes != 16 && line_bytes != 32 && line_bytes != 64 && line_bytes != 128)
    return GS_EINVAL;
  if (mode < 0 || mode > 4) return GS_EINVAL;
  if (inflight != 1 && inflight != 2 && inflight != 4) return GS_EINVAL;
  memset(out, 0, sizeof(*out));
  uint4* buf = nullptr;
  uint32_t* sink = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = GS_OK;
  const uint64_t n16 = buffer_bytes / 16;
  const uint64_t nlines = buffer_bytes / (uint64_t)line_bytes;
  if (hipMalloc(&buf, n16 * 16) != hipSuccess) return GS_ENOMEM;
  if (hipMalloc(&sink, 4096) != hipSuccess) {
    (void)hipFree(buf);
    return GS_ENOMEM;
  }
  if (hipStreamCreate(&st) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = GS_EHIP;
    goto done;
  }
  hipLaunchKernelGGL(fill_random, dim3(4096), dim3(256), 0, st, buf, n16, 12345ull);
  {
    float best = 1e30f;
    for (int r = 0; r <= reps; ++r) {  // first round is an untimed warm-up
      (void)hipEventRecord(e0, st);
      switch (line_bytes) {
        case 16:
          launch_gather<16>(mode, inflight, grid, buf, nlines, iters, 77ull + r, sink, st);
          break;
        case 32:
          launch_gather<32>(mode, inflight, grid, buf, nlines, iters, 77ull + r, sink, st);
          break;
        case 64:
          launch_gather<64>(mode, inflight, grid, buf, nlines, iters, 77ull + r, sink, st);
          break;
        default:
          launch_gather<128>(mode, inflight, grid, buf, nlines, iters, 77ull + r, sink, st);
          break;
      }
      (void)hipEventRecord(e1, st);
      if (hipEventSynchronize(e1) != hipSuccess) {
        rc = GS_EHIP;
        goto done;
      }
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (r > 0 && ms < best) best = ms;
    }
    // lines (mode 0/1), 16-byte chunks (mode 2), lines (mode 3: one per W lanes),
    // dwords (mode 4) served per launch
    double per_lane = mode == 0 ? (double)((iters / inflight) * inflight) : (double)iters;
    double lines = (double)grid * 256.0 * per_lane;
    if (mode == 3) lines /= (double)(line_bytes / 16);
    out->seconds = best * 1e-3;
    out->lines_per_s = lines / out->seconds;
    out->bytes_per_s = out->lines_per_s * (mode == 2 ? 16 : mode == 4 ? 4 : line_bytes);
  }
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  (void)hipFree(sink);
  (void)hipFree(buf);
  return rc;
}

extern "C" int gs_stream_copy_bytes(uint64_t bytes, int reps, double* seconds_out) {
  if (!seconds_out || bytes < 4096 || reps <= 0) return GS_EINVAL;
  uint4 *a = nullptr, *b = nullptr;
  const uint64_t n16 = bytes / 16;
  if (hipMalloc(&a, n16 * 16) != hipSuccess) return GS_ENOMEM;
  if (hipMalloc(&b, n16 * 16) != hipSuccess) {
    (void)hipFree(a);
    return GS_ENOMEM;
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipLaunchKernelGGL(fill_random, dim3(4096), dim3(256), 0, nullptr, a, n16, 1ull);
  float best = 1e30f;
  for (int r = 0; r <= reps; ++r) {
    (void)hipEventRecord(e0, nullptr);
    hipLaunchKernelGGL(stream_copy, dim3(8192), dim3(256), 0, nullptr, a, b, n16);
    (void)hipEventRecord(e1, nullptr);
    (void)hipEventSynchronize(e1);
    float ms;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (r > 0 && ms < best) best = ms;
  }
  *seconds_out = best * 1e-3;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return GS_OK;
}
