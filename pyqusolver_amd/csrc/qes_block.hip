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
#define BLK_DOT_WAVES_MAX 256   // waves along the vector length per slab group

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

// Partial projections: wave w of slab group g sums rows [r0, r1) of
// C[i][n] = sum_k V_i[k] B_n[k] for the vectors i of its BLK_SLABS slabs and
// the columns n < nb (B = W, then J W for complex), into
// part[(g * nwave + w) * BLK_SLABS * 256 + s * 256 + i_local * 16 + n].
__global__ void __launch_bounds__(256) block_dot_kernel(const double* __restrict__ V, int64_t ldv, int nv,
                                                        const double* __restrict__ W, int64_t ldw, int ncol,
                                                        int cplx, int64_t len, int nwave, double* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int ngroup = (((nv + 15) / 16) + BLK_SLABS - 1) / BLK_SLABS;
  if (wave >= ngroup * nwave) return;
  const int g = wave / nwave, w = wave % nwave;
  const int64_t nchunk = (len + 15) / 16;
  const int64_t c0 = nchunk * w / nwave, c1 = nchunk * (w + 1) / nwave;
  const int m = lane & 15, kq = lane >> 4;
  const int nb = cplx ? 2 * ncol : ncol;
  const int n = m;   // B column of this lane
  const bool ncol_ok = n < nb;
  const int wq = cplx && n >= ncol ? n - ncol : n;
  const double* wrow = W + (int64_t)(ncol_ok ? wq : 0) * ldw;
  qes_d4 acc[BLK_SLABS];
#pragma unroll
  for (int s = 0; s < BLK_SLABS; ++s) acc[s] = qes_d4{0.0, 0.0, 0.0, 0.0};
  for (int64_t c = c0; c < c1; ++c) {
    const int64_t rho = c * 16 + 4 * kq;   // this lane's 4 rows of the chunk
    double b[4] = {0.0, 0.0, 0.0, 0.0};
    if (ncol_ok) {
      double t[4];
      load4(wrow + rho, rho, len, t);
      if (cplx && n >= ncol) {   // J w: (w1, -w0, w3, -w2)
        b[0] = t[1];
        b[1] = -t[0];
        b[2] = t[3];
        b[3] = -t[2];
      } else {
        b[0] = t[0];
        b[1] = t[1];
        b[2] = t[2];
        b[3] = t[3];
      }
    }
#pragma unroll
    for (int s = 0; s < BLK_SLABS; ++s) {
      const int i = (g * BLK_SLABS + s) * 16 + m;
      double a[4] = {0.0, 0.0, 0.0, 0.0};
      if (i < nv) load4(V + (int64_t)i * ldv + rho, rho, len, a);
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b[t], acc[s], 0, 0, 0);
    }
  }
  double* out = part + (int64_t)wave * BLK_SLABS * 256;
#pragma unroll
  for (int s = 0; s < BLK_SLABS; ++s)
#pragma unroll
    for (int r = 0; r < 4; ++r) out[s * 256 + (kq + 4 * r) * 16 + m] = acc[s][r];
}

// C[i][q] (real nv x ncol, or complex interleaved) from the partials, summed
// over the waves of each slab group in order.
__global__ void block_dot_finish_kernel(const double* __restrict__ part, int nv, int ncol, int cplx, int nwave,
                                        double* __restrict__ C) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nv * ncol) return;
  const int i = t / ncol, q = t % ncol;
  const int g = i / (16 * BLK_SLABS), s = (i / 16) % BLK_SLABS, il = i % 16;
  double re = 0.0, im = 0.0;
  for (int w = 0; w < nwave; ++w) {
    const double* p = part + ((int64_t)(g * nwave + w) * BLK_SLABS + s) * 256 + il * 16;
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

// Y_q = beta Y_q + sum_i V_i C[i][q] over rows [16 c, 16 c + 16) for every
// chunk c of this wave (grid-stride); complex: the second MFMA adds Im(C) J'V.
__global__ void __launch_bounds__(256) block_combine_kernel(const double* __restrict__ V, int64_t ldv, int nv,
                                                            const double* __restrict__ C, double beta,
                                                            double* __restrict__ Y, int64_t ldy, int ncol, int cplx,
                                                            int64_t len) {
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * 4;
  const int m = lane & 15, kq = lane >> 4;
  const int64_t nchunk = (len + 15) / 16;
  const bool q_ok = m < ncol;
  for (int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); c < nchunk; c += nwaves) {
    const int64_t rho = c * 16 + m;     // A row of this lane
    qes_d4 acc = qes_d4{0.0, 0.0, 0.0, 0.0};
    for (int i0 = 0; i0 < nv; i0 += 4) {
      const int i = i0 + kq;
      const bool ok = i < nv;
      const double a = (ok && rho < len) ? V[(int64_t)i * ldv + rho] : 0.0;
      if (cplx) {
        const double br = (ok && q_ok) ? C[2 * ((int64_t)i * ncol + m)] : 0.0;
        const double bi = (ok && q_ok) ? C[2 * ((int64_t)i * ncol + m) + 1] : 0.0;
        const double partner = __shfl_xor(a, 1, 64);
        const double a2 = (m & 1) ? partner : -partner;   // (J'v)_{2a} = -v_{2a+1}, (J'v)_{2a+1} = v_{2a}
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, br, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, bi, acc, 0, 0, 0);
      } else {
        const double b = (ok && q_ok) ? C[(int64_t)i * ncol + m] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
      }
    }
    // lane holds D[kq + 4 r][m]: row c * 16 + kq + 4 r, column m
    if (q_ok) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = c * 16 + kq + 4 * r;
        if (row < len) {
          double* y = Y + (int64_t)m * ldy + row;
          *y = beta == 0.0 ? acc[r] : beta * (*y) + acc[r];
        }
      }
    }
  }
}

static int block_dot_nwave(int nv, int64_t len) {
  const int ngroup = (((nv + 15) / 16) + BLK_SLABS - 1) / BLK_SLABS;
  const int64_t nchunk = (len + 15) / 16;
  // about 2048 waves in all (8 per CU), at most BLK_DOT_WAVES_MAX per group
  int64_t w = std::max<int64_t>(1, 2048 / std::max(1, ngroup));
  w = std::min<int64_t>(w, BLK_DOT_WAVES_MAX);
  w = std::min<int64_t>(w, std::max<int64_t>(1, nchunk));
  return (int)w;
}

int64_t block_dot_ws_bytes(int nv, int64_t len) {
  const int ngroup = (((nv + 15) / 16) + BLK_SLABS - 1) / BLK_SLABS;
  return (int64_t)ngroup * block_dot_nwave(nv, len) * BLK_SLABS * 256 * 8;
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
  const int64_t need = block_dot_ws_bytes(nv, len);
  if (ws_bytes < need) return fail("block dot workspace of " + std::to_string(ws_bytes) + " bytes; needs " +
                                   std::to_string(need));
  const int nwave = block_dot_nwave(nv, len);
  const int ngroup = (((nv + 15) / 16) + BLK_SLABS - 1) / BLK_SLABS;
  const int nblk = (ngroup * nwave + 3) / 4;
  hipLaunchKernelGGL(block_dot_kernel, dim3(nblk), dim3(256), 0, st, V, ldv, nv, W, ldw, ncol, cplx, len, nwave, ws);
  QES_HIP(hipGetLastError());
  const int nt = nv * ncol;
  hipLaunchKernelGGL(block_dot_finish_kernel, dim3((nt + 255) / 256), dim3(256), 0, st, ws, nv, ncol, cplx, nwave, C);
  QES_HIP(hipGetLastError());
  return 0;
}

int block_combine(const double* V, int64_t ldv, int nv, const double* C, double beta, double* Y, int64_t ldy, int ncol,
                  int64_t len, int cplx, hipStream_t st) {
  if (check_block_args(V, ldv, nv, Y, ldy, ncol, len, cplx)) return -1;
  if (!C) return fail("null coefficient matrix");
  const int64_t nchunk = (len + 15) / 16;
  const int64_t nblk = std::min<int64_t>((nchunk + 3) / 4, 256 * 16);
  hipLaunchKernelGGL(block_combine_kernel, dim3((unsigned)nblk), dim3(256), 0, st, V, ldv, nv, C, beta, Y, ldy, ncol,
                     cplx, len);
  QES_HIP(hipGetLastError());
  return 0;
}

}  // namespace qes
