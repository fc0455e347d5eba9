// amd_device.hpp -- device-side helpers shared by the HIP kernels.
//
// * persistent XCD-aware grids: MI355X dispatches workgroup b to XCD b % 8,
//   so the tile space is split into 8 contiguous ranges, one per XCD, and the
//   workgroups of an XCD walk their range together; the set of rows in flight
//   on one XCD (and the x rows they gather) then stays narrow and L2-resident.
// * 16-byte vector loads/stores of VPL consecutive scalars.
// * deterministic two-stage reductions: each workgroup (or wave) writes its
//   partial sums, a second kernel adds them in a fixed order.
#pragma once
#include "amd_common.hpp"

namespace phx {

constexpr int kXcds = 8;
constexpr int kBlock = 256;  // threads per workgroup (4 wave64)

// first and one-past-last tile of workgroup-walk for this block
struct TileWalk {
  int64_t t, end, stride;
};
__device__ inline TileWalk xcd_tiles(int64_t ntiles) {
  const int xcd = blockIdx.x % kXcds;
  const int k = blockIdx.x / kXcds;
  const int per = gridDim.x / kXcds;
  TileWalk w;
  w.t = (int64_t)xcd * ntiles / kXcds + k;
  w.end = (int64_t)(xcd + 1) * ntiles / kXcds;
  w.stride = per;
  return w;
}

// (at most 16-byte alignment: the widest load is 16 bytes, so wider blocks
// are several loads and need only that)
template <typename T, int N>
struct alignas(sizeof(T) * N < 16 ? sizeof(T) * N : 16) VecN {
  T v[N];
};

template <typename T, int N>
__device__ inline void vload(const T* p, T (&out)[N]) {
  if constexpr (N == 1) {
    out[0] = *p;
  } else {
    VecN<T, N> t = *reinterpret_cast<const VecN<T, N>*>(p);
#pragma unroll
    for (int e = 0; e < N; e++) out[e] = t.v[e];
  }
}

template <typename T, int N>
__device__ inline void vstore(T* p, const T (&in)[N]) {
  if constexpr (N == 1) {
    *p = in[0];
  } else {
    VecN<T, N> t;
#pragma unroll
    for (int e = 0; e < N; e++) t.v[e] = in[e];
    *reinterpret_cast<VecN<T, N>*>(p) = t;
  }
}

// vstore with non-temporal stores (the written lines are not kept in the
// caches): 16-byte vectors where N scalars make whole vectors, else scalars
template <typename T, int N>
__device__ inline void vstore_nt(T* p, const T (&in)[N]) {
  constexpr int V = 16 / sizeof(T);
  if constexpr (N % V == 0) {
    typedef T vec_t __attribute__((ext_vector_type(V)));
#pragma unroll
    for (int c = 0; c < N / V; c++) {
      vec_t v;
#pragma unroll
      for (int e = 0; e < V; e++) v[e] = in[c * V + e];
      __builtin_nontemporal_store(v, reinterpret_cast<vec_t*>(p) + c);
    }
  } else {
#pragma unroll
    for (int e = 0; e < N; e++) __builtin_nontemporal_store(in[e], p + e);
  }
}

// sum over the lanes of a wave64 that share (lane % STRIDE)
template <int STRIDE, typename T>
__device__ inline T wave_sum_strided(T v) {
#pragma unroll
  for (int off = STRIDE; off < 64; off <<= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Persistent-grid size for a kernel: resident workgroups (at most
// max_per_cu per CU when > 0), a multiple of 8.
template <typename K>
inline int persistent_blocks(K kernel, size_t lds_bytes, int64_t ntiles, int max_per_cu = 0) {
  Runtime& r = rt();
  int per_cu = occupancy_per_cu(reinterpret_cast<const void*>(kernel), kBlock, lds_bytes);
  if (max_per_cu > 0) per_cu = std::min(per_cu, max_per_cu);
  int64_t g = (int64_t)r.num_cu * per_cu;
  if (r.grid_blocks > 0) g = std::min<int64_t>(g, r.grid_blocks);
  int64_t need = ((ntiles + kXcds - 1) / kXcds) * kXcds;
  g = std::min(g, std::max<int64_t>(need, kXcds));
  g = std::max<int64_t>(kXcds, (g / kXcds) * kXcds);
  return (int)g;
}

// Adds the partial tiles of the four waves of a workgroup in LDS (wave 0,
// then 1, 2, 3) and stores the sum as the workgroup's tile out[0..nout).
// put(add) is called once per wave; each lane passes the elements (o, v) it
// owns to add(o, v), and every element of the tile is owned by exactly one
// lane of each wave.  sh holds nout elements (dynamic LDS).
template <typename T, typename F>
__device__ inline void block_tile_sum(T* sh, int nout, T* __restrict__ out, F&& put) {
  const int wv = threadIdx.x >> 6;
#pragma unroll 1
  for (int w = 0; w < kBlock / 64; w++) {
    if (wv == w) put([&](int o, T v) { sh[o] = w == 0 ? v : sh[o] + v; });
    __syncthreads();
  }
  for (int o = threadIdx.x; o < nout; o += kBlock) out[o] = sh[o];
}

// Sum over b of partials[b*width + o] for the 16 outputs o = 16*blockIdx.x +
// (tid & 15) of this workgroup: thread group p = tid / 16 adds b = p, p+16, ...
// in order, then the 16 group sums are added in order of p.  The result is
// returned on threads 0..15 (others get 0); every thread must call it.
template <typename T>
__device__ inline T block_sum16(const T* __restrict__ partials, int nblocks, int width, int group) {
  __shared__ T sh[16][17];
  const int o = group * 16 + (threadIdx.x & 15), grp = threadIdx.x >> 4;
  T s = T(0);
  if (o < width) {
#pragma unroll 8
    for (int b = grp; b < nblocks; b += kBlock / 16) s += partials[(int64_t)b * width + o];
  }
  sh[grp][threadIdx.x & 15] = s;
  __syncthreads();
  T t = T(0);
  if (threadIdx.x < 16)
#pragma unroll
    for (int q = 0; q < kBlock / 16; q++) t += sh[q][threadIdx.x];
  __syncthreads();  // sh may be reused by the next call
  return t;
}
template <typename T>
__device__ inline T block_sum16(const T* __restrict__ partials, int nblocks, int width) {
  return block_sum16(partials, nblocks, width, (int)blockIdx.x);
}

// Adds partials[b*width + i] over b in a fixed order; out[i] (device);
// launched with ceil(width / 16) workgroups.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_reduce_partials(const T* __restrict__ partials, int nblocks, int width,
                                                            T* __restrict__ out) {
  const T s = block_sum16(partials, nblocks, width);
  const int o = blockIdx.x * 16 + threadIdx.x;
  if (threadIdx.x < 16 && o < width) out[o] = s;
}

// The sums of k_reduce_partials written to the published-result buffer (host
// memory): with one workgroup the sums go there directly, else they are read
// from out; then the buffer's sequence word becomes seq.  The sums and the
// word are written by the first wave, whose system-scope release store of
// the word follows its stores of the sums.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_reduce_publish(const T* __restrict__ partials, int nblocks, int width,
                                                           const T* __restrict__ out, T* pub, uint32_t* word,
                                                           uint32_t seq) {
  if (width <= 16) {
    const T s = block_sum16(partials, nblocks, width);
    if (threadIdx.x < 16 && (int)threadIdx.x < width) pub[threadIdx.x] = s;
  } else if (threadIdx.x < 64) {
    for (int o = threadIdx.x; o < width; o += 64) pub[o] = out[o];
  }
  if (threadIdx.x == 0) __hip_atomic_store(word, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// reduce [nblocks][width] device partials and copy the width sums to host_out
template <typename T>
inline void reduce_partials_to_host(const T* d_partials, int nblocks, int width, T* host_out) {
  Runtime& r = rt();
  T* d_out = const_cast<T*>(d_partials) + (size_t)nblocks * width;
  void *hs = nullptr, *ds = nullptr;
  if ((size_t)width * sizeof(T) <= Runtime::kSignalBytes - Runtime::kSignalData - 2 * kSignalSlotBytes &&
      signal_buffer(&hs, &ds)) {
    // the device publishes the sums and a new sequence number; the host waits for that number
    if (width > 16) {
      hipLaunchKernelGGL(k_reduce_partials<T>, dim3((width + 15) / 16), dim3(kBlock), 0, r.stream, d_partials,
                         nblocks, width, d_out);
      PHX_HIP(hipGetLastError());
    }
    uint32_t seq = ++r.sig_seq;
    if (seq == 0) seq = ++r.sig_seq;
    hipLaunchKernelGGL(k_reduce_publish<T>, dim3(1), dim3(kBlock), 0, r.stream, d_partials, nblocks, width, d_out,
                       reinterpret_cast<T*>(static_cast<char*>(ds) + Runtime::kSignalData), static_cast<uint32_t*>(ds),
                       seq);
    PHX_HIP(hipGetLastError());
    wait_signal(seq);
    memcpy(host_out, static_cast<const char*>(hs) + Runtime::kSignalData, sizeof(T) * width);
    return;
  }
  hipLaunchKernelGGL(k_reduce_partials<T>, dim3((width + 15) / 16), dim3(kBlock), 0, r.stream, d_partials, nblocks,
                     width, d_out);
  PHX_HIP(hipGetLastError());
  T* h = (T*)host_work(sizeof(T) * width);
  PHX_HIP(hipMemcpyAsync(h, d_out, sizeof(T) * width, hipMemcpyDeviceToHost, r.stream));
  PHX_HIP(hipStreamSynchronize(r.stream));
  memcpy(host_out, h, sizeof(T) * width);
}

}  // namespace phx
