// Block-wide status reduction shared by the HIP kernels: up to two maxima of
// non-negative doubles (as integer bit patterns, whose order matches the
// values' order, so the result does not depend on scheduling) and one count,
// combined across the block and then written with one atomic each from
// thread 0.  Every thread of the block must call it (it synchronises).
#pragma once

#include <hip/hip_runtime.h>

namespace vjx {

__device__ __forceinline__ void block_flags(unsigned long long m0, unsigned long long m1,
                                            unsigned int cnt, unsigned long long* o0,
                                            unsigned long long* o1, unsigned int* oc) {
  __shared__ unsigned long long s_m0[16], s_m1[16];
  __shared__ unsigned int s_cnt[16];
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long w0 = __shfl_xor(m0, o);
    const unsigned long long w1 = __shfl_xor(m1, o);
    m0 = w0 > m0 ? w0 : m0;
    m1 = w1 > m1 ? w1 : m1;
    cnt += __shfl_xor(cnt, o);
  }
  const int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_m0[wid] = m0;
    s_m1[wid] = m1;
    s_cnt[wid] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long a = 0, b = 0;
    unsigned int c = 0;
    const int nw = ((int)blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; ++w) {
      a = s_m0[w] > a ? s_m0[w] : a;
      b = s_m1[w] > b ? s_m1[w] : b;
      c += s_cnt[w];
    }
    if (a && o0) atomicMax(o0, a);
    if (b && o1) atomicMax(o1, b);
    if (c && oc) atomicAdd(oc, c);
  }
}

}  // namespace vjx
