// Dense contractions of block Lanczos on FP64 matrix cores (gfx950).
//
// Full reorthogonalisation of a block of p vectors against m stored ones is
// two skinny GEMMs per pass: C = V^H W (m x p, summed over the vector length)
// and W -= V C.  Both run on v_mfma_f64_16x16x4f64 (D += A B with A 16x4,
// B 4x16; lane l supplies A[l%16][l/16] and B[l/16][l%16] and holds
// D[l/16 + 4r][l%16] in its r-th accumulator).  Vectors are stored vector
// by vector (leading dimension ld doubles); complex vectors as interleaved
// (re, im) doubles, i.e. real vectors of twice the length:
//   Re <v, w> = sum_k v_k w_k,  Im <v, w> = sum_k v_k (J w)_k  with
//   (J w)_{2a} = w_{2a+1}, (J w)_{2a+1} = -w_{2a},
// so one real GEMM against the columns [W, J W] gives the complex product,
// and  w -= v c  is  w -= Re(c) v + Im(c) (J' v),  (J' v)_{2a} = -v_{2a+1},
// (J' v)_{2a+1} = v_{2a}: a second MFMA on lane-swapped operands.
// Sums run in a fixed order (no atomics): results are reproducible.
#include <algorithm>
#include <string>

#include "qes_internal.h"

namespace qes {

typedef double qes_d4 __attribute__((ext_vector_type(4)));

#define BLK_SLABS 4            // 16-vector slabs per wave in the projection
#define BLK_MAX_COL 16          // B columns of one MFMA

// 4 consecutive doubles at p[0..3] with p + 4 > end guarded (zero past len)
__device__ __forceinline__ void load4(const double* __restrict__ p, int64_t rho, int64_t len, double* out) {
  if (rho + 4 <= len) {
    const double2 a = *reinterpret_cast<const double2*>(p);
    const double2 b = *reinterpret_cast<const double2*>(p + 2);
    out[0] = a.x;
    out[1] = a.y;
    out[2] = b.x;
    out[3] = b.y;
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e) out[e] = (rho + e < len) ? p[e] : 0.0;
  }
}

// Partial projections: block b of slab group g sums the row chunks
// [c0, c1) (its four waves take every fourth chunk) of
// C[i][n] = sum_k V_i[k] B_n[k] for the vectors i of its BLK_SLABS slabs and
// the columns n < nb (B = W, then J W for complex); the waves' tiles are
// added in LDS and stored at part[(g * nblk + b) * BLK_SLABS * 256 + s * 256 +
// i_local * 16 + n].
__global__ void __launch_bounds__(256) block_dot_kernel(const double* __restrict__ V, int64_t ldv, int nv,
                                                        const double* __restrict__ W, int64_t ldw, int ncol,
                                                        int cplx, int64_t len, int nblk, double* __restrict__ part) {
  __shared__ double tile[4][BLK_SLABS * 256];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int g = blockIdx.x / nblk, b = blockIdx.x % nblk;
  const int64_t nchunk = (len + 15) / 16;
  const int64_t c0 = nchunk * b / nblk, c1 = nchunk * (b + 1) / nblk;
  const int m = lane & 15, kq = lane >> 4;
  const int nb = cplx ? 2 * ncol : ncol;
  const int n = m;   // B column of this lane
  const bool ncol_ok = n < nb;
  const int wq = cplx && n >= ncol ? n - ncol : n;
  const double* wrow = W + (int64_t)(ncol_ok ? wq : 0) * ldw;
  qes_d4 acc[BLK_SLABS];
#pragma unroll
  for (int s = 0; s < BLK_SLABS; ++s) acc[s] = qes_d4{0.0, 0.0, 0.0, 0.0};
  for (int64_t c = c0 + wid; c < c1; c += 4) {
    const int64_t rho = c * 16 + 4 * kq;   // this lane's 4 rows of the chunk
    double a[BLK_SLABS][4];
#pragma unroll
    for (int s = 0; s < BLK_SLABS; ++s) {
      const int i = (g * BLK_SLABS + s) * 16 + m;
#pragma unroll
      for (int e = 0; e < 4; ++e) a[s][e] = 0.0;
      if (i < nv) load4(V + (int64_t)i * ldv + rho, rho, len, a[s]);
    }
    double b4[4] = {0.0, 0.0, 0.0, 0.0};
    if (ncol_ok) {
      double t[4];
      load4(wrow + rho, rho, len, t);
      if (cplx && n >= ncol) {   // J w: (w1, -w0, w3, -w2)
        b4[0] = t[1];
        b4[1] = -t[0];
        b4[2] = t[3];
        b4[3] = -t[2];
      } else {
        b4[0] = t[0];
        b4[1] = t[1];
        b4[2] = t[2];
        b4[3] = t[3];
      }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int s = 0; s < BLK_SLABS; ++s) acc[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s][t], b4[t], acc[s], 0, 0, 0);
  }
#pragma unroll
  for (int s = 0; s < BLK_SLABS; ++s)
#pragma unroll
    for (int r = 0; r < 4; ++r) tile[wid][s * 256 + (kq + 4 * r) * 16 + m] = acc[s][r];
  __syncthreads();
  double* out = part + (int64_t)blockIdx.x * BLK_SLABS * 256;
  for (int e = threadIdx.x; e < BLK_SLABS * 256; e += 256)
    out[e] = (tile[0][e] + tile[1][e]) + (tile[2][e] + tile[3][e]);
}

// C[i][q] (real nv x ncol, or complex interleaved) from the partials, summed
// over the waves of each slab group in order.
__global__ void block_dot_finish_kernel(const double* __restrict__ part, int nv, int ncol, int cplx, int nblk,
                                        double* __restrict__ C) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nv * ncol) return;
  const int i = t / ncol, q = t % ncol;
  const int g = i / (16 * BLK_SLABS), s = (i / 16) % BLK_SLABS, il = i % 16;
  double re = 0.0, im = 0.0;
  for (int w = 0; w < nblk; ++w) {
    const double* p = part + ((int64_t)(g * nblk + w) * BLK_SLABS + s) * 256 + il * 16;
    re += p[q];
    if (cplx) im += p[ncol + q];
  }
  if (cplx) {
    C[2 * t] = re;
    C[2 * t + 1] = im;
  } else {
    C[t] = re;
  }
}

// Y_q = beta Y_q + sum_i V_i C[i][q] over rows [64 c, 64 c + 64) for every
// chunk c of this wave (grid-stride).  Lane (m, kq) reads rows 64 c + 4 m ..
// + 3 of vector i0 + kq as 32 bytes; accumulator t takes row 64 c + 4 m + t
// as MFMA row m, so the four accumulators are independent chains and hold
// D[kq + 4 r][m] <-> row 64 c + 4 (kq + 4 r) + t, column m.  Complex: the
// second MFMA adds Im(C) J'V, whose rows pair within a lane's four.
__global__ void __launch_bounds__(256) block_combine_kernel(const double* __restrict__ V, int64_t ldv, int nv,
                                                            const double* __restrict__ C, double beta,
                                                            double* __restrict__ Y, int64_t ldy, int ncol, int cplx,
                                                            int64_t len) {
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * 4;
  const int m = lane & 15, kq = lane >> 4;
  const int64_t nchunk = (len + 63) / 64;
  const bool q_ok = m < ncol;
  for (int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); c < nchunk; c += nwaves) {
    const int64_t rho = c * 64 + 4 * m;     // this lane's four A rows
    qes_d4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = qes_d4{0.0, 0.0, 0.0, 0.0};
    for (int i0 = 0; i0 < nv; i0 += 4) {
      const int i = i0 + kq;
      const bool ok = i < nv;
      double a[4] = {0.0, 0.0, 0.0, 0.0};
      if (ok) load4(V + (int64_t)i * ldv + rho, rho, len, a);
      if (cplx) {
        const double br = (ok && q_ok) ? C[2 * ((int64_t)i * ncol + m)] : 0.0;
        const double bi = (ok && q_ok) ? C[2 * ((int64_t)i * ncol + m) + 1] : 0.0;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          // (J'v)_{2a} = -v_{2a+1}, (J'v)_{2a+1} = v_{2a}; rho is a multiple of 4
          const double a2 = (t & 1) ? a[t - 1] : -a[t + 1];
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], br, acc[t], 0, 0, 0);
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, bi, acc[t], 0, 0, 0);
        }
      } else {
        const double b = (ok && q_ok) ? C[(int64_t)i * ncol + m] : 0.0;
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b, acc[t], 0, 0, 0);
      }
    }
    if (q_ok) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = c * 64 + 4 * (kq + 4 * r);
        double* y = Y + (int64_t)m * ldy + row;
#pragma unroll
        for (int t = 0; t < 4; ++t)
          if (row + t < len) y[t] = beta == 0.0 ? acc[t][r] : beta * y[t] + acc[t][r];
      }
    }
  }
}

// blocks (of four waves) per slab group: about 2048 blocks in all (eight
// per CU), at least one chunk per wave
static int block_dot_nblk(int nv, int64_t len) {
  const int ngroup = (((nv + 15) / 16) + BLK_SLABS - 1) / BLK_SLABS;
  const int64_t nchunk = (len + 15) / 16;
  int64_t b = std::max<int64_t>(1, 2048 / std::max(1, ngroup));
  b = std::min<int64_t>(b, std::max<int64_t>(1, (nchunk + 3) / 4));
  return (int)b;
}

// workspace of any call with at most nv vectors: ngroup * nblk <= max(2048, ngroup)
int64_t block_dot_ws_bytes(int nv, int64_t len) {
  (void)len;
  const int ngroup = (((nv + 15) / 16) + BLK_SLABS - 1) / BLK_SLABS;
  return (int64_t)std::max(2048, ngroup) * BLK_SLABS * 256 * 8;
}

static int64_t block_dot_ws_need(int nv, int64_t len) {
  const int ngroup = (((nv + 15) / 16) + BLK_SLABS - 1) / BLK_SLABS;
  return (int64_t)ngroup * block_dot_nblk(nv, len) * BLK_SLABS * 256 * 8;
}

static int check_block_args(const double* V, int64_t ldv, int nv, const double* W, int64_t ldw, int ncol, int64_t len,
                            int cplx) {
  if (!V || !W) return fail("null vector block");
  if (nv < 1 || ncol < 1) return fail("block sizes must be >= 1");
  if (ncol > (cplx ? BLK_MAX_COL / 2 : BLK_MAX_COL))
    return fail(std::string("at most ") + std::to_string(cplx ? BLK_MAX_COL / 2 : BLK_MAX_COL) +
                " columns per block (" + (cplx ? "complex" : "real") + ")");
  if (len < 1 || ldv < len || ldw < len) return fail("leading dimensions must be >= the vector length");
  if ((ldv & 1) || (ldw & 1) || (((uintptr_t)V | (uintptr_t)W) & 15))
    return fail("vector blocks must be 16-byte aligned with even leading dimensions");
  if (cplx && (len & 1)) return fail("complex vectors have an even number of doubles");
  return 0;
}

int block_dot(const double* V, int64_t ldv, int nv, const double* W, int64_t ldw, int ncol, int64_t len, int cplx,
              double* C, double* ws, int64_t ws_bytes, hipStream_t st) {
  if (check_block_args(V, ldv, nv, W, ldw, ncol, len, cplx)) return -1;
  if (!C || !ws) return fail("null output or workspace");
  const int64_t need = block_dot_ws_need(nv, len);
  if (ws_bytes < need) return fail("block dot workspace of " + std::to_string(ws_bytes) + " bytes; needs " +
                                   std::to_string(need));
  const int nblk = block_dot_nblk(nv, len);
  const int ngroup = (((nv + 15) / 16) + BLK_SLABS - 1) / BLK_SLABS;
  hipLaunchKernelGGL(block_dot_kernel, dim3(ngroup * nblk), dim3(256), 0, st, V, ldv, nv, W, ldw, ncol, cplx, len, nblk,
                     ws);
  QES_HIP(hipGetLastError());
  const int nt = nv * ncol;
  hipLaunchKernelGGL(block_dot_finish_kernel, dim3((nt + 255) / 256), dim3(256), 0, st, ws, nv, ncol, cplx, nblk, C);
  QES_HIP(hipGetLastError());
  return 0;
}

int block_combine(const double* V, int64_t ldv, int nv, const double* C, double beta, double* Y, int64_t ldy, int ncol,
                  int64_t len, int cplx, hipStream_t st) {
  if (check_block_args(V, ldv, nv, Y, ldy, ncol, len, cplx)) return -1;
  if (!C) return fail("null coefficient matrix");
  const int64_t nchunk = (len + 63) / 64;
  const int64_t nblk = std::min<int64_t>((nchunk + 3) / 4, 256 * 16);
  hipLaunchKernelGGL(block_combine_kernel, dim3((unsigned)nblk), dim3(256), 0, st, V, ldv, nv, C, beta, Y, ldy, ncol,
                     cplx, len);
  QES_HIP(hipGetLastError());
  return 0;
}

}  // namespace qes
