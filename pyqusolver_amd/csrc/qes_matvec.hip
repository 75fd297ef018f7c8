// Pull-form symmetric-sector matvec (float64 / complex128) on gfx950.
//
// The reference applies H column by column (push form,
// Algebra/Hilbert/matrix_builder.py:1347-1377):
//     y[idx] += val * conj(chi(g*)) * N_idx / N_k * x[k]
// with idx/g* read from dense 2^Ns maps.  Here each lane owns one output row
// a and gathers, using hermiticity H_ab = conj(H_ba):
//     y[a] = sum_{s' in H|r_a>} conj(val) * chi(g*(s')) * (N_b / N_a) * x[b]
// where b = index of the representative of s' (canonicalisation by the group
// + bucketed binary search over the sorted representative keys) and
// N_b = sqrt(|Stab(r_b)|) is the number of minimising group elements found
// during canonicalisation.  Every output is written by exactly one lane: no
// atomics, bitwise reproducible.
#include "qes_internal.h"

namespace qes {

template <bool CPLX> struct Acc;
template <> struct Acc<false> {
  double v = 0.0;
};
template <> struct Acc<true> {
  double r = 0.0, i = 0.0;
};

__device__ __forceinline__ double sel4(const double* t, int k) {
  double v0 = t[0], v1 = t[1], v2 = t[2], v3 = t[3];
  return k == 0 ? v0 : (k == 1 ? v1 : (k == 2 ? v2 : v3));
}

__device__ __forceinline__ int64_t xcd_block(int64_t bid, int64_t nblocks, int remap) {
  if (!remap) return bid;
  const int64_t nx = 8;
  int64_t per = nblocks / nx, rem = nblocks % nx;
  int64_t xcd = bid % nx, idx = bid / nx;
  return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

// NB = number of right-hand sides handled per lane (columns j0 .. j0+NB-1 of
// an (n, ldx) row-major block).
template <typename W, bool CPLX, bool LOADNORM, int NB>
__global__ void __launch_bounds__(256) matvec_kernel(QesGroup G, QesLookup L, const double* __restrict__ chi_re,
                                                     const double* __restrict__ chi_im,
                                                     const QesTermGroup* __restrict__ groups, int n_diag, int n_off,
                                                     const double* __restrict__ norms, const double* __restrict__ x,
                                                     double* __restrict__ y, int64_t n, int64_t ldx, int64_t j0,
                                                     int ncols, int accumulate, int remap) {
  __shared__ double s_chr[256];
  __shared__ double s_chi[256];
  const int gs = G.ntx * G.nty * G.nbase;
  for (int i = threadIdx.x; i < gs && i < 256; i += blockDim.x) {
    s_chr[i] = chi_re[i];
    s_chi[i] = CPLX ? chi_im[i] : 0.0;
  }
  __syncthreads();
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x, remap);
  const int64_t a = blk * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const W r = (W)L.keys64[a];
  const double na = norms[a];
  const double inv_na = 1.0 / na;
  const int cw = CPLX ? 2 : 1;

  double acc_r[NB], acc_i[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    acc_r[j] = 0.0;
    acc_i[j] = 0.0;
  }

  // diagonal (sum of all mask-0 groups), H_aa
  {
    double dr = 0.0, di = 0.0;
    for (int t = 0; t < n_diag; ++t) {
      const QesTermGroup& g = groups[t];
      int k = term_bits<W>(g, r);
      dr += sel4(g.re, k);
      if (CPLX) di += sel4(g.im, k);
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < ncols) {
        const double* xp = x + (a * ldx + j0 + j) * cw;
        if (CPLX) {
          double xr = xp[0], xi = xp[1];
          acc_r[j] += dr * xr - di * xi;
          acc_i[j] += dr * xi + di * xr;
        } else {
          acc_r[j] += dr * xp[0];
        }
      }
    }
  }

  for (int t = n_diag; t < n_diag + n_off; ++t) {
    const QesTermGroup& g = groups[t];
    const int k = term_bits<W>(g, r);
    const double vr = sel4(g.re, k);
    const double vi = CPLX ? sel4(g.im, k) : 0.0;
    if (vr == 0.0 && vi == 0.0) continue;
    const W sp = r ^ (W)g.mask;
    if (G.n_up >= 0 && popc<W>(sp) != G.n_up) continue;
    if (G.z2 != 0 && ((popc<W>(sp) & 1) ? -1 : 1) != G.z2) continue;
    const Canon<W> c = canonicalize<W>(G, sp);
    const int64_t b = lookup_rep<W>(L, c.rep);
    if (b < 0) continue;
    const double nb = LOADNORM ? norms[b] : sqrt((double)c.count);
    const double sc = nb * inv_na;
    // factor = conj(v) * chi[g*] * N_b/N_a
    const double cr = s_chr[c.g], ci = s_chi[c.g];
    double fr, fi;
    if (CPLX) {
      fr = (vr * cr + vi * ci) * sc;
      fi = (vr * ci - vi * cr) * sc;
    } else {
      fr = vr * cr * sc;
      fi = 0.0;
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j < ncols) {
        const double* xp = x + (b * ldx + j0 + j) * cw;
        if (CPLX) {
          double xr = xp[0], xi = xp[1];
          acc_r[j] += fr * xr - fi * xi;
          acc_i[j] += fr * xi + fi * xr;
        } else {
          acc_r[j] += fr * xp[0];
        }
      }
    }
  }

#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if (j < ncols) {
      double* yp = y + (a * ldx + j0 + j) * cw;
      if (CPLX) {
        if (accumulate) {
          yp[0] += acc_r[j];
          yp[1] += acc_i[j];
        } else {
          yp[0] = acc_r[j];
          yp[1] = acc_i[j];
        }
      } else {
        if (accumulate) yp[0] += acc_r[j]; else yp[0] = acc_r[j];
      }
    }
  }
}

template <typename W, bool CPLX, bool LOADNORM>
static int launch_matvec(const Op* op, const double* x, double* y, int64_t nbatch, int accumulate, hipStream_t st) {
  const Basis* B = op->basis;
  const int64_t n = B->n;
  if (n == 0) return 0;
  const int block = 256;
  const int64_t nblocks = (n + block - 1) / block;
  if (nblocks > 0x7fffffffll) return fail("matvec grid too large");
  for (int64_t j0 = 0; j0 < nbatch; j0 += 4) {
    int ncols = (int)std::min<int64_t>(4, nbatch - j0);
    if (ncols == 1)
      hipLaunchKernelGGL((matvec_kernel<W, CPLX, LOADNORM, 1>), dim3((unsigned)nblocks), dim3(block), 0, st, B->group,
                         B->lookup(), B->chi_re, B->chi_im, op->d_groups, op->n_diag, op->n_off, B->norms, x, y, n,
                         nbatch, j0, ncols, accumulate, op->xcd_remap);
    else
      hipLaunchKernelGGL((matvec_kernel<W, CPLX, LOADNORM, 4>), dim3((unsigned)nblocks), dim3(block), 0, st, B->group,
                         B->lookup(), B->chi_re, B->chi_im, op->d_groups, op->n_diag, op->n_off, B->norms, x, y, n,
                         nbatch, j0, ncols, accumulate, op->xcd_remap);
    QES_HIP(hipGetLastError());
  }
  return 0;
}

int matvec(const Op* op, const void* x, void* y, int64_t nbatch, int dtype, int accumulate, hipStream_t st) {
  const Basis* B = op->basis;
  const bool cplx = dtype == 1;
  if (!cplx && (op->is_complex || op->chi_complex))
    return fail("complex matrix elements (complex couplings or momentum sector) need complex128 vectors");
  if (nbatch < 1) return fail("nbatch must be >= 1");
  const bool w32 = B->group.ns <= 32;
  const bool ln = B->group.has_phase != 0;
  const double* xd = (const double*)x;
  double* yd = (double*)y;
#define QES_DISPATCH(W_, C_, L_) return launch_matvec<W_, C_, L_>(op, xd, yd, nbatch, accumulate, st)
  if (w32) {
    if (cplx) { if (ln) QES_DISPATCH(uint32_t, true, true); else QES_DISPATCH(uint32_t, true, false); }
    else { if (ln) QES_DISPATCH(uint32_t, false, true); else QES_DISPATCH(uint32_t, false, false); }
  } else {
    if (cplx) { if (ln) QES_DISPATCH(uint64_t, true, true); else QES_DISPATCH(uint64_t, true, false); }
    else { if (ln) QES_DISPATCH(uint64_t, false, true); else QES_DISPATCH(uint64_t, false, false); }
  }
#undef QES_DISPATCH
}

// ---------------------------------------------------------------------------
// Per-coupling outputs: y[c, a] = sum over terms of coupling c (single RHS).
// Used by the autograd backward (grad_J_c = Re <g, H_c v>).
// ---------------------------------------------------------------------------
#define QES_MAX_COUPLINGS 8
template <typename W, bool CPLX, bool LOADNORM>
__global__ void __launch_bounds__(256) matvec_couplings_kernel(QesGroup G, QesLookup L, const double* __restrict__ chi_re,
                                                               const double* __restrict__ chi_im,
                                                               const QesTermGroup* __restrict__ groups, int n_diag,
                                                               int n_off, int ncpl, const double* __restrict__ norms,
                                                               const double* __restrict__ x, double* __restrict__ y,
                                                               int64_t n) {
  __shared__ double s_chr[256];
  __shared__ double s_chi[256];
  const int gs = G.ntx * G.nty * G.nbase;
  for (int i = threadIdx.x; i < gs && i < 256; i += blockDim.x) {
    s_chr[i] = chi_re[i];
    s_chi[i] = CPLX ? chi_im[i] : 0.0;
  }
  __syncthreads();
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const W r = (W)L.keys64[a];
  const double inv_na = 1.0 / norms[a];
  const int cw = CPLX ? 2 : 1;
  double ar[QES_MAX_COUPLINGS], ai[QES_MAX_COUPLINGS];
#pragma unroll
  for (int c = 0; c < QES_MAX_COUPLINGS; ++c) {
    ar[c] = 0.0;
    ai[c] = 0.0;
  }
  const double xar = x[a * cw], xai = CPLX ? x[a * cw + 1] : 0.0;
  for (int t = 0; t < n_diag; ++t) {
    const QesTermGroup& g = groups[t];
    int k = term_bits<W>(g, r);
    double dr = sel4(g.re, k), di = CPLX ? sel4(g.im, k) : 0.0;
    double pr = dr * xar - di * xai, pi = dr * xai + di * xar;
#pragma unroll
    for (int c = 0; c < QES_MAX_COUPLINGS; ++c)
      if (c == g.coupling) {
        ar[c] += pr;
        ai[c] += pi;
      }
  }
  for (int t = n_diag; t < n_diag + n_off; ++t) {
    const QesTermGroup& g = groups[t];
    const int k = term_bits<W>(g, r);
    const double vr = sel4(g.re, k);
    const double vi = CPLX ? sel4(g.im, k) : 0.0;
    if (vr == 0.0 && vi == 0.0) continue;
    const W sp = r ^ (W)g.mask;
    if (G.n_up >= 0 && popc<W>(sp) != G.n_up) continue;
    if (G.z2 != 0 && ((popc<W>(sp) & 1) ? -1 : 1) != G.z2) continue;
    const Canon<W> cn = canonicalize<W>(G, sp);
    const int64_t b = lookup_rep<W>(L, cn.rep);
    if (b < 0) continue;
    const double nb = LOADNORM ? norms[b] : sqrt((double)cn.count);
    const double sc = nb * inv_na;
    const double cr = s_chr[cn.g], ci = s_chi[cn.g];
    double fr = (vr * cr + vi * ci) * sc;
    double fi = CPLX ? (vr * ci - vi * cr) * sc : 0.0;
    const double xr = x[b * cw], xi = CPLX ? x[b * cw + 1] : 0.0;
    double pr = fr * xr - fi * xi, pi = fr * xi + fi * xr;
#pragma unroll
    for (int c = 0; c < QES_MAX_COUPLINGS; ++c)
      if (c == g.coupling) {
        ar[c] += pr;
        ai[c] += pi;
      }
  }
  for (int c = 0; c < ncpl && c < QES_MAX_COUPLINGS; ++c) {
    double* yp = y + ((int64_t)c * n + a) * cw;
    yp[0] = ar[c];
    if (CPLX) yp[1] = ai[c];
  }
}

int matvec_couplings(const Op* op, const void* x, void* y, int dtype, hipStream_t st) {
  const Basis* B = op->basis;
  const bool cplx = dtype == 1;
  if (!cplx && (op->is_complex || op->chi_complex))
    return fail("complex matrix elements need complex128 vectors");
  if (op->n_coupling > QES_MAX_COUPLINGS) return fail("too many coupling groups (max 8)");
  const int64_t n = B->n;
  if (n == 0) return 0;
  dim3 grid((unsigned)((n + 255) / 256));
  const bool w32 = B->group.ns <= 32, ln = B->group.has_phase != 0;
  const double* xd = (const double*)x;
  double* yd = (double*)y;
#define QES_LC(W_, C_, L_)                                                                                       \
  hipLaunchKernelGGL((matvec_couplings_kernel<W_, C_, L_>), grid, dim3(256), 0, st, B->group, B->lookup(),        \
                     B->chi_re, B->chi_im, op->d_groups, op->n_diag, op->n_off, op->n_coupling, B->norms, xd, yd, n)
  if (w32) {
    if (cplx) { if (ln) QES_LC(uint32_t, true, true); else QES_LC(uint32_t, true, false); }
    else { if (ln) QES_LC(uint32_t, false, true); else QES_LC(uint32_t, false, false); }
  } else {
    if (cplx) { if (ln) QES_LC(uint64_t, true, true); else QES_LC(uint64_t, true, false); }
    else { if (ln) QES_LC(uint64_t, false, true); else QES_LC(uint64_t, false, false); }
  }
#undef QES_LC
  QES_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// E_off: number of off-diagonal gathers per matvec (stats pass)
// ---------------------------------------------------------------------------
template <typename W>
__global__ void count_offdiag_kernel(QesGroup G, QesLookup L, const QesTermGroup* __restrict__ groups, int n_diag,
                                     int n_off, int64_t n, unsigned long long* __restrict__ out) {
  __shared__ unsigned long long s_sum[256];
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long cnt = 0;
  if (a < n) {
    const W r = (W)L.keys64[a];
    for (int t = n_diag; t < n_diag + n_off; ++t) {
      const QesTermGroup& g = groups[t];
      const int k = term_bits<W>(g, r);
      if (sel4(g.re, k) == 0.0 && sel4(g.im, k) == 0.0) continue;
      const W sp = r ^ (W)g.mask;
      if (G.n_up >= 0 && popc<W>(sp) != G.n_up) continue;
      if (G.z2 != 0 && ((popc<W>(sp) & 1) ? -1 : 1) != G.z2) continue;
      const Canon<W> c = canonicalize<W>(G, sp);
      if (lookup_rep<W>(L, c.rep) >= 0) ++cnt;
    }
  }
  s_sum[threadIdx.x] = cnt;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) s_sum[threadIdx.x] += s_sum[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, s_sum[0]);
}

int count_offdiag(const Op* op, int64_t* e_off, hipStream_t st) {
  const Basis* B = op->basis;
  const int64_t n = B->n;
  *e_off = 0;
  if (n == 0) return 0;
  unsigned long long* d = nullptr;
  QES_HIP(hipMalloc(&d, sizeof(unsigned long long)));
  QES_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long), st));
  dim3 grid((unsigned)((n + 255) / 256));
  if (B->group.ns <= 32)
    hipLaunchKernelGGL(count_offdiag_kernel<uint32_t>, grid, dim3(256), 0, st, B->group, B->lookup(), op->d_groups,
                       op->n_diag, op->n_off, n, d);
  else
    hipLaunchKernelGGL(count_offdiag_kernel<uint64_t>, grid, dim3(256), 0, st, B->group, B->lookup(), op->d_groups,
                       op->n_diag, op->n_off, n, d);
  QES_HIP(hipGetLastError());
  unsigned long long h = 0;
  QES_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, st));
  QES_HIP(hipStreamSynchronize(st));
  QES_HIP(hipFree(d));
  *e_off = (int64_t)h;
  return 0;
}

}  // namespace qes
