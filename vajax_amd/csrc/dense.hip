// Dense trailing block of the LU refactorisation on the matrix cores.
//
// When the last pivots of the analysis form a block whose fill is dense (the
// top separators of the nested dissection, the supply rails), the scalar
// update program stops before them and the block is factorised as a dense
// row-major K x K FP64 matrix D:
//
//   1. D, Lp, Up cleared, then D gathered from the block's LU slots (as the
//      scalar program left them), Lp / Up from the L columns and U rows of the
//      earlier pivots with large contributions to the block;
//   2. D -= Lp Up (FP64 MFMA);
//   3. D = L U in place, blocked by panels of kDensePanel columns: the
//      panel's diagonal block in LDS, then L21 = A21 U11^-1 row by row and
//      U12 = L11^-1 A12 column by column (all rows and columns in parallel),
//      then A22 -= L21 U12 (FP64 MFMA);
//   4. the slots written back (FP64 or FP32 factors), folding max |L| and the
//      non-finite count into the pivot-growth status.
//
// The matrix product: C[M x N] -= A[M x Kd] B[Kd x N], row-major with leading
// dimensions, one 64 x 64 tile of C per workgroup of four wavefronts, each a
// 32 x 32 sub-tile as 2 x 2 v_mfma_f64_16x16x4f64 accumulators; A and B in
// 16-deep slices through LDS.  Operand layout of v_mfma_f64_16x16x4f64
// (wave64): lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16] and holds
// D[(l / 16) + 4 r][l % 16], r = 0..3: accumulator rows interleave across the
// four 16-lane groups.
#include "block_flags.h"
#include "internal.h"

namespace vjx {

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned long long dbits(double x) {
  return (unsigned long long)__double_as_longlong(x);
}

inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

__global__ void __launch_bounds__(256) k_dgemm_sub(int M, int N, int Kd,
                                                   const double* __restrict__ A, int lda,
                                                   const double* __restrict__ B, int ldb,
                                                   double* C, int ldc) {
  __shared__ double As[64][17];
  __shared__ double Bs[16][64];
  const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = (int)blockIdx.y * 64, n0 = (int)blockIdx.x * 64;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  dbl4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < Kd; k0 += 16) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = tid + u * 256;
      const int r = q >> 4, c = q & 15, gr = m0 + r, gc = k0 + c;
      As[r][c] = (gr < M && gc < Kd) ? A[(size_t)gr * lda + gc] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = tid + u * 256;
      const int r = q >> 6, c = q & 63, gr = k0 + r, gc = n0 + c;
      Bs[r][c] = (gr < Kd && gc < N) ? B[(size_t)gr * ldb + gc] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; kk += 4) {
      double a[2], b[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        a[t] = As[wr + t * 16 + (lane & 15)][kk + (lane >> 4)];
        b[t] = Bs[kk + (lane >> 4)][wc + t * 16 + (lane & 15)];
      }
#pragma unroll
      for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
          acc[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ti], b[tj], acc[ti][tj], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wr + ti * 16 + (lane >> 4) + 4 * r;
        const int col = n0 + wc + tj * 16 + (lane & 15);
        if (row < M && col < N) C[(size_t)row * ldc + col] -= acc[ti][tj][r];
      }
}

// The panel's w x w diagonal block (rows and columns o .. o+w-1) factorised in
// place in LDS by one workgroup: unit L below the diagonal, U on and above.
__global__ void __launch_bounds__(256) k_dense_diag(int K, int o, int w, double* D) {
  __shared__ double P[kDensePanel][kDensePanel + 1];
  const int tid = (int)threadIdx.x;
  for (int q = tid; q < w * w; q += 256) P[q / w][q % w] = D[(size_t)(o + q / w) * K + o + q % w];
  __syncthreads();
  for (int c = 0; c < w; ++c) {
    if (tid > c && tid < w) P[tid][c] /= P[c][c];
    __syncthreads();
    for (int q = tid; q < w * w; q += 256) {
      const int r = q / w, c2 = q % w;
      if (r > c && c2 > c) P[r][c2] -= P[r][c] * P[c][c2];
    }
    __syncthreads();
  }
  for (int q = tid; q < w * w; q += 256) D[(size_t)(o + q / w) * K + o + q % w] = P[q / w][q % w];
}

// L21 = A21 U11^-1 (blocks [0, nbr): one row per thread) and U12 = L11^-1 A12
// (the other blocks: one column per thread), reading the factorised diagonal
// block into LDS.
__global__ void __launch_bounds__(256) k_dense_panel(int K, int o, int w, int nbr, double* D) {
  __shared__ double P[kDensePanel][kDensePanel + 1];
  const int tid = (int)threadIdx.x;
  for (int q = tid; q < w * w; q += 256) P[q / w][q % w] = D[(size_t)(o + q / w) * K + o + q % w];
  __syncthreads();
  if ((int)blockIdx.x < nbr) {
    const int r = o + w + (int)blockIdx.x * 256 + tid;
    if (r >= K) return;
    double x[kDensePanel];
    double* row = D + (size_t)r * K + o;
    for (int c = 0; c < w; ++c) {
      double v = row[c];
      for (int c2 = 0; c2 < c; ++c2) v -= x[c2] * P[c2][c];
      x[c] = v / P[c][c];
    }
    for (int c = 0; c < w; ++c) row[c] = x[c];
  } else {
    const int j = o + w + ((int)blockIdx.x - nbr) * 256 + tid;
    if (j >= K) return;
    double y[kDensePanel];
    for (int r = 0; r < w; ++r) {
      double v = D[(size_t)(o + r) * K + j];
      for (int r2 = 0; r2 < r; ++r2) v -= P[r][r2] * y[r2];
      y[r] = v;
    }
    for (int r = 0; r < w; ++r) D[(size_t)(o + r) * K + j] = y[r];
  }
}

__global__ void k_dfill(size_t n, double* x) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    x[i] = 0.0;
}

template <typename T>
__global__ void k_dense_gather(int n, const int32_t* __restrict__ slot,
                               const int32_t* __restrict__ pos, const T* __restrict__ lu,
                               double* out) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) out[pos[i]] = (double)lu[slot[i]];
}

// lu[slot] = D[pos]; the block's L entries (below its diagonal) feed max |L|
// and the non-finite count of the pivot-growth monitor.
template <typename T>
__global__ void __launch_bounds__(256) k_dense_scatter(int n, int K,
                                                       const int32_t* __restrict__ slot,
                                                       const int32_t* __restrict__ pos,
                                                       const double* __restrict__ D, T* lu,
                                                       DevStatus* st) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  double mag = 0.0;
  unsigned int bad = 0;
  if (i < n) {
    const int p = pos[i];
    const T v = (T)D[p];
    lu[slot[i]] = v;
    if (p / K > p % K) {
      const double a = fabs((double)v);
      if (a <= 1.7976931348623157e308)
        mag = a;
      else
        bad = 1;
    }
  }
  if (st) block_flags(dbits(mag), 0ull, bad, &st->maxl_bits, nullptr, &st->nonfinite);
}

}  // namespace

void launch_dgemm_sub(int M, int N, int Kd, const double* A, int lda, const double* B, int ldb,
                      double* C, int ldc, hipStream_t s) {
  if (M <= 0 || N <= 0 || Kd <= 0) return;
  dim3 grid((unsigned)nblk(N, 64), (unsigned)nblk(M, 64));
  k_dgemm_sub<<<grid, 256, 0, s>>>(M, N, Kd, A, lda, B, ldb, C, ldc);
}

void launch_dense_lu(int K, double* D, hipStream_t s) {
  for (int o = 0; o < K; o += kDensePanel) {
    const int w = std::min(kDensePanel, K - o);
    k_dense_diag<<<1, 256, 0, s>>>(K, o, w, D);
    const int nt = K - o - w;
    if (nt <= 0) break;
    const int nbr = nblk(nt, 256);
    k_dense_panel<<<2 * nbr, 256, 0, s>>>(K, o, w, nbr, D);
    launch_dgemm_sub(nt, nt, w, D + (size_t)(o + w) * K + o, K, D + (size_t)o * K + o + w, K,
                     D + (size_t)(o + w) * K + o + w, K, s);
  }
}

template <typename T>
void launch_dense_block(DevLU& L, T* lu, DevStatus* st, hipStream_t s) {
  const LuSymbolic& S = L.sym;
  const int K = S.dense_k, m = S.dense_m;
  if (K <= 0) return;
  ktimer_begin(s);
  const size_t nd = (size_t)K * K, nl = (size_t)K * m;
  k_dfill<<<std::min(nblk((long long)nd, 256), 2048), 256, 0, s>>>(nd, L.dense_D.p);
  if (m > 0) {
    k_dfill<<<std::min(nblk((long long)nl, 256), 2048), 256, 0, s>>>(nl, L.dense_Lp.p);
    k_dfill<<<std::min(nblk((long long)nl, 256), 2048), 256, 0, s>>>(nl, L.dense_Up.p);
  }
  const int nds = (int)L.dense_slot.n, ndl = (int)L.dl_slot.n, ndu = (int)L.du_slot.n;
  if (nds) k_dense_gather<T><<<nblk(nds, 256), 256, 0, s>>>(nds, L.dense_slot.p, L.dense_pos.p, lu, L.dense_D.p);
  if (ndl) k_dense_gather<T><<<nblk(ndl, 256), 256, 0, s>>>(ndl, L.dl_slot.p, L.dl_pos.p, lu, L.dense_Lp.p);
  if (ndu) k_dense_gather<T><<<nblk(ndu, 256), 256, 0, s>>>(ndu, L.du_slot.p, L.du_pos.p, lu, L.dense_Up.p);
  if (m > 0) launch_dgemm_sub(K, K, m, L.dense_Lp.p, m, L.dense_Up.p, K, L.dense_D.p, K, s);
  launch_dense_lu(K, L.dense_D.p, s);
  if (nds) k_dense_scatter<T><<<nblk(nds, 256), 256, 0, s>>>(nds, K, L.dense_slot.p, L.dense_pos.p, L.dense_D.p, lu, st);
  ktimer_end(KC_LU_DENSE, s);
}

template void launch_dense_block<double>(DevLU&, double*, DevStatus*, hipStream_t);
template void launch_dense_block<float>(DevLU&, float*, DevStatus*, hipStream_t);

}  // namespace vjx

using namespace vjx;

extern "C" int vjx_dense_lu(int32_t K, double* D_dev) {
  if (K < 0 || (K > 0 && !D_dev)) {
    set_error("vjx_dense_lu: bad size or null matrix");
    return VJX_ERR_ARG;
  }
  launch_dense_lu(K, D_dev, 0);
  VJX_TRY_HIP(hipGetLastError());
  VJX_TRY_HIP(hipDeviceSynchronize());
  return VJX_OK;
}

extern "C" int vjx_dense_gemm_sub(int32_t M, int32_t N, int32_t Kd, const double* A_dev,
                                  int32_t lda, const double* B_dev, int32_t ldb, double* C_dev,
                                  int32_t ldc) {
  if (M < 0 || N < 0 || Kd < 0 || lda < Kd || ldb < N || ldc < N ||
      (M && N && Kd && (!A_dev || !B_dev || !C_dev))) {
    set_error("vjx_dense_gemm_sub: bad sizes or null arrays");
    return VJX_ERR_ARG;
  }
  launch_dgemm_sub(M, N, Kd, A_dev, lda, B_dev, ldb, C_dev, ldc, 0);
  VJX_TRY_HIP(hipGetLastError());
  VJX_TRY_HIP(hipDeviceSynchronize());
  return VJX_OK;
}
