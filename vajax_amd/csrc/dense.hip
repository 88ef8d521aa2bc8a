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
//   3. D = L U in place, blocked by panels of kDensePanel columns: one launch
//      per panel factorises its diagonal block in LDS (in every workgroup)
//      and solves the rows of L21 = A21 U11^-1 and the columns of
//      U12 = L11^-1 A12 one per thread, then A22 -= L21 U12 (FP64 MFMA);
//   4. the slots written back (FP64 or FP32 factors), folding max |L| and the
//      non-finite count into the pivot-growth status.
//
// The matrix product: C[M x N] -= A[M x Kd] B[Kd x N], row-major with leading
// dimensions, one 32 x 32 tile of C per workgroup of four wavefronts that
// split Kd and add their partial tiles in a fixed order.  Operand layout of
// v_mfma_f64_16x16x4f64 (wave64): lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] and holds D[(l / 16) + 4 r][l % 16], r = 0..3:
// accumulator rows interleave across the four 16-lane groups.
#include "block_flags.h"
#include "internal.h"

namespace vjx {

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned long long dbits(double x) {
  return (unsigned long long)__double_as_longlong(x);
}

inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

// Lane `lane`'s value of v (lane uniform across the wavefront).
__device__ __forceinline__ double lane_value(double v, int lane) {
  const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
  const unsigned int lo = (unsigned int)__builtin_amdgcn_readlane((int)(unsigned int)bits, lane);
  const unsigned int hi = (unsigned int)__builtin_amdgcn_readlane((int)(unsigned int)(bits >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// C[M x N] -= A[M x Kd] B[Kd x N]: one 32 x 32 tile of C per workgroup; each
// of its four wavefronts accumulates the whole tile (2 x 2 blocks of
// v_mfma_f64_16x16x4f64) over every fourth 4-deep step of Kd, reading its
// operands straight from memory in the instruction's lane layout, and the
// four partial tiles are added in wavefront order through LDS before C is
// updated (a fixed summation order: the result does not depend on timing).
__global__ void __launch_bounds__(256) k_dgemm_sub(int M, int N, int Kd,
                                                   const double* __restrict__ A, int lda,
                                                   const double* __restrict__ B, int ldb,
                                                   double* C, int ldc) {
  __shared__ double part[4][16][64];
  const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = (int)blockIdx.y * 32, n0 = (int)blockIdx.x * 32;
  dbl4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
  const int ar = lane & 15, ak = lane >> 4;  // A[row][k] / B[k][col] of this lane
  for (int k0 = wave * 4; k0 < Kd; k0 += 16) {
    const int k = k0 + ak;
    double a[2], b[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int r = m0 + t * 16 + ar, c = n0 + t * 16 + ar;
      a[t] = (r < M && k < Kd) ? A[(size_t)r * lda + k] : 0.0;
      b[t] = (k < Kd && c < N) ? B[(size_t)k * ldb + c] : 0.0;
    }
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
        acc[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ti], b[tj], acc[ti][tj], 0, 0, 0);
  }
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) part[wave][ti * 8 + tj * 4 + r][lane] = acc[ti][tj][r];
  __syncthreads();
  // thread -> (accumulator element, lane); lane l of element (ti, tj, r) is
  // C[m0 + 16 ti + l / 16 + 4 r][n0 + 16 tj + l % 16]
  for (int q = tid; q < 16 * 64; q += 256) {
    const int e = q >> 6, l = q & 63;
    const double v = ((part[0][e][l] + part[1][e][l]) + part[2][e][l]) + part[3][e][l];
    const int ti = e >> 3, tj = (e >> 2) & 1, r = e & 3;
    const int row = m0 + ti * 16 + (l >> 4) + 4 * r;
    const int col = n0 + tj * 16 + (l & 15);
    if (row < M && col < N) C[(size_t)row * ldc + col] -= v;
  }
}

// One panel of the blocked LU (columns o .. o+w-1, w <= kDensePanel): the
// first wavefront of every workgroup factorises the w x w diagonal block in
// registers (unit L below the diagonal, U on and above; workgroup 0 writes it
// back) and leaves it in LDS, then
// each thread solves one row of L21 = A21 U11^-1 (workgroups [0, nrb), 256
// rows each) or one column of U12 = L11^-1 A12 (the next ncb, 256 columns
// each) in registers, by substitution in the order of the host program (the
// loops are unrolled over kDensePanel, so the row stays in registers and the
// diagonal block's LDS reads are issued ahead of the dependent chain).
__global__ void __launch_bounds__(256) k_dense_panel(int K, int o, int w, int nrb, int ncb,
                                                     double* D) {
  constexpr int W = kDensePanel;
  __shared__ double P[W][W + 1];
  const int tid = (int)threadIdx.x;
  const int b = (int)blockIdx.x;
  double x[W];
  if (tid < 64) {
    // wavefront 0: lane r holds row r of the diagonal block; step c divides
    // the rows below c by the pivot and subtracts the multiple of row c,
    // whose entries every lane reads from lane c (no barrier between steps)
#pragma unroll
    for (int k = 0; k < W; ++k)
      x[k] = (tid < w && k < w) ? D[(size_t)(o + tid) * K + o + k] : 0.0;
    const bool below = tid < w;
#pragma unroll
    for (int c = 0; c < W; ++c) {
      if (c < w) {
        const double piv = lane_value(x[c], c);
        const bool act = below && tid > c;
        const double l = act ? x[c] / piv : x[c];
        x[c] = l;
#pragma unroll
        for (int k = c + 1; k < W; ++k) {
          const double u = lane_value(x[k], c);
          if (act) x[k] -= l * u;
        }
      }
    }
    if (below) {
#pragma unroll
      for (int k = 0; k < W; ++k) {
        P[tid][k] = x[k];
        if (b == 0 && k < w) D[(size_t)(o + tid) * K + o + k] = x[k];
      }
    }
  }
  __syncthreads();
  if (b < nrb) {
    const int r = o + w + b * 256 + tid;
    if (r >= K) return;
    double* row = D + (size_t)r * K + o;
#pragma unroll
    for (int c = 0; c < W; ++c) x[c] = c < w ? row[c] : 0.0;
#pragma unroll
    for (int c = 0; c < W; ++c)
      if (c < w) {
        double v = x[c];
#pragma unroll
        for (int c2 = 0; c2 < c; ++c2) v -= x[c2] * P[c2][c];
        x[c] = v / P[c][c];
      }
#pragma unroll
    for (int c = 0; c < W; ++c)
      if (c < w) row[c] = x[c];
  } else if (b < nrb + ncb) {
    const int j = o + w + (b - nrb) * 256 + tid;
    if (j >= K) return;
#pragma unroll
    for (int r = 0; r < W; ++r) x[r] = r < w ? D[(size_t)(o + r) * K + j] : 0.0;
#pragma unroll
    for (int r = 0; r < W; ++r)
      if (r < w) {
        double v = x[r];
#pragma unroll
        for (int r2 = 0; r2 < r; ++r2) v -= P[r][r2] * x[r2];
        x[r] = v;
      }
#pragma unroll
    for (int r = 0; r < W; ++r)
      if (r < w) D[(size_t)(o + r) * K + j] = x[r];
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
  dim3 grid((unsigned)nblk(N, 32), (unsigned)nblk(M, 32));
  k_dgemm_sub<<<grid, 256, 0, s>>>(M, N, Kd, A, lda, B, ldb, C, ldc);
}

void launch_dense_lu(int K, double* D, hipStream_t s) {
  for (int o = 0; o < K; o += kDensePanel) {
    const int w = std::min(kDensePanel, K - o);
    const int nt = K - o - w;
    const int nb = nblk(nt, 256);
    k_dense_panel<<<std::max(1, 2 * nb), 256, 0, s>>>(K, o, w, nb, nb, D);
    if (nt <= 0) break;
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
