// tk_micro.hip -- random-access ceilings of one gfx950 device, measured with
// the access shapes the hash-table kernels use, over a buffer far larger than
// the caches.  They bound the count pass (random 8-byte atomic increments)
// and the scoring pass (random 64-byte bucket reads) better than the
// streaming HBM peak does, because every random access moves a whole cache
// line between HBM and L2.
//
// Kinds:
//   TKM_GATHER8    one 8-byte read per lane at a random aligned address
//   TKM_GATHER64   eight lanes read one random 64-byte line (a table bucket)
//   TKM_GATHER128  sixteen lanes read one random 128-byte line
//   TKM_ATOMIC     one 8-byte atomicAdd per lane, result unused
//   TKM_ATOMIC_RET one 8-byte atomicAdd per lane, result consumed
//   TKM_CAS        one 8-byte atomicCAS per lane (compare value almost never
//                  matches), result consumed
//   TKM_ATOMIC_WG  one 8-byte atomicAdd per lane at workgroup scope, result
//                  consumed: performed in the XCD's L2 instead of at the
//                  memory side (the buffer's contents are meaningless after
//                  it; only the rate is measured)
//
// Benchmark infrastructure: not part of the k-mer path's C ABI.
#include <hip/hip_runtime.h>
#include <stdint.h>

extern "C" {
enum { TKM_GATHER8 = 0, TKM_GATHER64 = 1, TKM_GATHER128 = 2, TKM_ATOMIC = 3, TKM_ATOMIC_RET = 4, TKM_CAS = 5,
       TKM_ATOMIC_WG = 6 };
int tkm_run(int kind, void* buf, uint64_t nbytes, uint64_t n_ops, float* ms, uint64_t* ops_done);
// XCC (XCD) that each of `nblocks` workgroups ran on, into d_out (device, u32).
int tkm_xcc_map(uint32_t* d_out, uint32_t nblocks);
}

namespace {

constexpr int kUnroll = 8;

__device__ __forceinline__ uint64_t mix(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

// G lanes (8 bytes each) per lookup; `iters` lookups per lane group, a
// multiple of kUnroll.  Every index is < nlines, so every address is inside
// the nlines * G * 8 bytes of buf.
template <int G>
__global__ __launch_bounds__(256) void k_gather(const uint64_t* __restrict__ buf, uint64_t nlines,
                                                uint64_t iters, uint64_t* __restrict__ sink) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t grp = tid / G;
  const int sl = (int)(tid % G);
  uint64_t acc = 0;
  for (uint64_t i = 0; i < iters; i += kUnroll) {
    uint64_t v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const uint64_t line = __umul64hi(mix((grp * iters + i + u) * 0x9E3779B97F4A7C15ULL + 1), nlines);
      v[u] = buf[line * G + sl];
    }
#pragma unroll
    for (int u = 0; u < kUnroll; u++) acc ^= v[u];
  }
  sink[tid] = acc;
}

template <bool RET>
__global__ __launch_bounds__(256) void k_atomic(uint64_t* __restrict__ buf, uint64_t nslots, uint64_t iters,
                                                uint64_t* __restrict__ sink) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t acc = 0;
  for (uint64_t i = 0; i < iters; i += kUnroll) {
    uint64_t r[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const uint64_t s = __umul64hi(mix((tid * iters + i + u) * 0x9E3779B97F4A7C15ULL + 1), nslots);
      if (RET)
        r[u] = atomicAdd((unsigned long long*)&buf[s], 1ull);
      else
        atomicAdd((unsigned long long*)&buf[s], 1ull);
    }
    if (RET) {
#pragma unroll
      for (int u = 0; u < kUnroll; u++) acc ^= r[u];
    }
  }
  sink[tid] = acc;
}

template <int KIND>
__global__ __launch_bounds__(256) void k_atomic2(uint64_t* __restrict__ buf, uint64_t nslots, uint64_t iters,
                                                 uint64_t* __restrict__ sink) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t acc = 0;
  for (uint64_t i = 0; i < iters; i += kUnroll) {
    uint64_t r[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const uint64_t x = mix((tid * iters + i + u) * 0x9E3779B97F4A7C15ULL + 1);
      const uint64_t s = __umul64hi(x, nslots);
      if (KIND == TKM_CAS)
        r[u] = atomicCAS((unsigned long long*)&buf[s], (unsigned long long)(x | 1ull), 0ull);
      else
        r[u] = __hip_atomic_fetch_add((unsigned long long*)&buf[s], 1ull, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_WORKGROUP);
    }
#pragma unroll
    for (int u = 0; u < kUnroll; u++) acc ^= r[u];
  }
  sink[tid] = acc;
}

__global__ void k_xcc_map(uint32_t* __restrict__ out) {
  if (threadIdx.x == 0)
    out[blockIdx.x] = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 0xFu;  // XCC_ID[3:0]
}

}  // namespace

extern "C" int tkm_xcc_map(uint32_t* d_out, uint32_t nblocks) {
  if (!d_out || nblocks == 0) return 1;
  hipLaunchKernelGGL(k_xcc_map, dim3(nblocks), dim3(64), 0, 0, d_out);
  return hipDeviceSynchronize() == hipSuccess ? 0 : 4;
}

extern "C" int tkm_run(int kind, void* buf, uint64_t nbytes, uint64_t n_ops, float* ms, uint64_t* ops_done) {
  if (!buf || !ms || !ops_done || kind < 0 || kind > TKM_ATOMIC_WG || nbytes < 4096 || n_ops == 0) return 1;
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 2;
  const int block = 256;
  const uint64_t grid = (uint64_t)(ncu > 0 ? ncu : 256) * 8;
  const uint64_t threads = grid * block;
  const int G = kind == TKM_GATHER64 ? 8 : kind == TKM_GATHER128 ? 16 : 1;
  const uint64_t units = threads / G;  // lane groups, one lookup each per iteration
  uint64_t iters = (n_ops + units - 1) / units;
  iters = (iters + kUnroll - 1) / kUnroll * kUnroll;
  uint64_t* sink = nullptr;
  if (hipMalloc(&sink, threads * 8) != hipSuccess) return 3;
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  uint64_t* p = (uint64_t*)buf;
  (void)hipEventRecord(a, 0);
  switch (kind) {
    case TKM_GATHER8:
      hipLaunchKernelGGL(k_gather<1>, dim3(grid), dim3(block), 0, 0, p, nbytes / 8, iters, sink);
      break;
    case TKM_GATHER64:
      hipLaunchKernelGGL(k_gather<8>, dim3(grid), dim3(block), 0, 0, p, nbytes / 64, iters, sink);
      break;
    case TKM_GATHER128:
      hipLaunchKernelGGL(k_gather<16>, dim3(grid), dim3(block), 0, 0, p, nbytes / 128, iters, sink);
      break;
    case TKM_ATOMIC:
      hipLaunchKernelGGL(k_atomic<false>, dim3(grid), dim3(block), 0, 0, p, nbytes / 8, iters, sink);
      break;
    case TKM_ATOMIC_RET:
      hipLaunchKernelGGL(k_atomic<true>, dim3(grid), dim3(block), 0, 0, p, nbytes / 8, iters, sink);
      break;
    case TKM_CAS:
      hipLaunchKernelGGL(k_atomic2<TKM_CAS>, dim3(grid), dim3(block), 0, 0, p, nbytes / 8, iters, sink);
      break;
    default:
      hipLaunchKernelGGL(k_atomic2<TKM_ATOMIC_WG>, dim3(grid), dim3(block), 0, 0, p, nbytes / 8, iters, sink);
      break;
  }
  (void)hipEventRecord(b, 0);
  hipError_t e = hipEventSynchronize(b);
  if (e == hipSuccess) e = hipEventElapsedTime(ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipFree(sink);
  *ops_done = units * iters;
  return e == hipSuccess ? 0 : 4;
}
