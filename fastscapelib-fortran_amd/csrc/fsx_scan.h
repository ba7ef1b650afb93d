// fsx_scan.h -- wave64 / workgroup scan helpers.
#pragma once
#include <hip/hip_runtime.h>

__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  return v;
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Inclusive scan over a workgroup of NT threads (NT a multiple of 64, <= 4096).
// lds must hold NT/64 ints.  Every thread receives the workgroup total.  Contains barriers:
// all threads of the workgroup must call it.
template <int NT>
__device__ __forceinline__ int block_incl_scan(int x, int* lds, int& total) {
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int v = wave_incl_scan(x);
  if (lane == 63) lds[wid] = v;
  __syncthreads();
  if (wid == 0) {
    int w = lane < NW ? lds[lane] : 0;
    w = wave_incl_scan(w);
    if (lane < NW) lds[lane] = w;
  }
  __syncthreads();
  int add = wid ? lds[wid - 1] : 0;
  total = lds[NW - 1];
  __syncthreads();
  return v + add;
}

template <int NT>
__device__ __forceinline__ int block_sum(int x, int* lds) {
  int t;
  block_incl_scan<NT>(x, lds, t);
  return t;
}
