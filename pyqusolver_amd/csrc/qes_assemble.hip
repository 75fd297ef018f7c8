// Assembled operator: the pull-form rows of H tabulated once in a sliced
// ELLPACK layout (64 rows per slice = one wavefront, column-major inside the
// slice), for callers that apply the same H many times (Lanczos).
//
// The reference offers the same trade: the explicit sparse build
// (build_operator_matrix / _build_sparse_same_sector_compact_jit,
// Algebra/Hilbert/matrix_builder.py:151-206, :911-1030) next to the
// matrix-free matvec_fun.  Here the table is built on the device by the same
// row visitor the matrix-free kernel uses, in the same order and with the
// same arithmetic (entry 0 is the diagonal, then the off-diagonal elements in
// visit order with their value already divided by N_a), so the assembled
// product is the matrix-free one up to the contraction of the final adds.
//
// Slice s covers rows 64s .. 64s+63; its len[s] = longest row in the slice;
// entry k of row 64s+l sits at off[s] + 64k + l.  Short rows are padded with
// (col = own row, value 0).  A matvec then streams the table with perfectly
// coalesced loads and gathers x[col]: no canonicalisation, no lookup.
#include <algorithm>
#include <string>
#include <vector>

#include "qes_internal.h"
#include "qes_rows.h"

namespace qes {

static constexpr int SELL_C = 64;

template <typename W, bool CPLX, int CK, bool FILL>
__global__ void __launch_bounds__(256) asm_rows_kernel(QesGroup G, QesLookup L, QesPlan P,
                                                       const double* __restrict__ chi_re,
                                                       const double* __restrict__ chi_im,
                                                       const QesTermGroup* __restrict__ groups, int n_diag,
                                                       const double* __restrict__ exval,
                                                       const uint32_t* __restrict__ lut,
                                                       const double* __restrict__ norms, int64_t n,
                                                       int32_t* __restrict__ rowlen, const int64_t* __restrict__ off,
                                                       const int32_t* __restrict__ slen, int32_t* __restrict__ col,
                                                       double* __restrict__ val) {
  extern __shared__ char smem[];
  const RowCtx C = load_ctx<CPLX>(G, L, P, chi_re, chi_im, groups, n_diag, exval, lut, smem);
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const W r = load_rep<W>(L, a);
  const double inv_na = 1.0 / norms[a];
  const int cw = CPLX ? 2 : 1;
  const int64_t s = a / SELL_C;
  const int l = (int)(a % SELL_C);
  int64_t base = FILL ? off[s] + l : 0;
  int k = 0;
  // entry 0: diagonal
  if (FILL) {
    double dr = 0.0, di = 0.0;
    for (int t = 0; t < n_diag; ++t) {
      const QesTermGroup& g = groups[t];
      const int kk = term_bits<W>(g, r);
      dr += sel4(g.re, kk);
      if (CPLX) di += sel4(g.im, kk);
    }
    col[base] = (int32_t)a;
    val[base * cw] = dr;
    if (CPLX) val[base * cw + 1] = di;
  }
  ++k;
  visit_row<W, CPLX, CK>(C, r, [&](int64_t b, double fr, double fi, int) {
    if (FILL) {
      const int64_t e = base + (int64_t)k * SELL_C;
      col[e] = (int32_t)b;
      val[e * cw] = fr * inv_na;
      if (CPLX) val[e * cw + 1] = fi * inv_na;
    }
    ++k;
  });
  if (FILL) {
    for (; k < slen[s]; ++k) {
      const int64_t e = base + (int64_t)k * SELL_C;
      col[e] = (int32_t)a;
      val[e * cw] = 0.0;
      if (CPLX) val[e * cw + 1] = 0.0;
    }
  } else {
    rowlen[a] = k;
  }
}

// y = A x with the sliced table; VC: complex values, XC: complex vectors
template <bool VC, bool XC>
__global__ void __launch_bounds__(256) sell_matvec_kernel(const int64_t* __restrict__ off,
                                                          const int32_t* __restrict__ slen,
                                                          const int32_t* __restrict__ col,
                                                          const double* __restrict__ val,
                                                          const double* __restrict__ x, double* __restrict__ y,
                                                          int64_t n, int accumulate) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const int64_t s = a / SELL_C;
  const int l = (int)(a % SELL_C);
  const int64_t base = off[s] + l;
  const int len = slen[s];
  const int vw = VC ? 2 : 1;
  double ar = 0.0, ai = 0.0;
  for (int k = 0; k < len; ++k) {
    const int64_t e = base + (int64_t)k * SELL_C;
    const int64_t b = col[e];
    const double vr = val[e * vw];
    const double vi = VC ? val[e * vw + 1] : 0.0;
    if (XC) {
      const double xr = x[2 * b], xi = x[2 * b + 1];
      ar += vr * xr - vi * xi;
      ai += vr * xi + vi * xr;
    } else {
      ar += vr * x[b];
    }
  }
  if (XC) {
    if (accumulate) {
      y[2 * a] += ar;
      y[2 * a + 1] += ai;
    } else {
      y[2 * a] = ar;
      y[2 * a + 1] = ai;
    }
  } else {
    if (accumulate) y[a] += ar; else y[a] = ar;
  }
}

void free_assembly(Op* op) {
  if (op->asm_off) hipFree(op->asm_off);
  if (op->asm_len) hipFree(op->asm_len);
  if (op->asm_col) hipFree(op->asm_col);
  if (op->asm_val) hipFree(op->asm_val);
  op->asm_off = nullptr;
  op->asm_len = nullptr;
  op->asm_col = nullptr;
  op->asm_val = nullptr;
  op->asm_entries = 0;
  op->asm_vc = 0;
}

#define QES_ASM_CK_SWITCH(CK_VAR, MACRO)                   \
  switch (CK_VAR) {                                        \
    case QES_CANON_CYCLIC: MACRO(QES_CANON_CYCLIC); break; \
    case QES_CANON_ROWS: MACRO(QES_CANON_ROWS); break;     \
    default: MACRO(QES_CANON_BRUTE); break;                \
  }

template <typename W, bool CPLX>
static int assemble_t(Op* op, hipStream_t st) {
  const Basis* B = op->basis;
  const int64_t n = B->n;
  const QesPlan& P = op->plan;
  const dim3 grid((unsigned)((n + 255) / 256));
  int32_t* d_rowlen = nullptr;
  QES_HIP(hipMalloc(&d_rowlen, sizeof(int32_t) * n));
#define QES_ASM(CK, FILL, ...)                                                                                      \
  hipLaunchKernelGGL((asm_rows_kernel<W, CPLX, CK, FILL>), grid, dim3(256), op->lds_bytes, st, B->group,          \
                     B->lookup(), P, B->chi_re, B->chi_im, op->d_gen, op->n_diag, op->d_exval, op->d_lut, B->norms, \
                     n, __VA_ARGS__)
#define QES_ASM_COUNT(CK) QES_ASM(CK, false, d_rowlen, nullptr, nullptr, nullptr, nullptr)
  QES_ASM_CK_SWITCH(P.canon, QES_ASM_COUNT)
#undef QES_ASM_COUNT
  QES_HIP(hipGetLastError());
  std::vector<int32_t> rowlen(n);
  QES_HIP(hipMemcpyAsync(rowlen.data(), d_rowlen, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
  QES_HIP(hipStreamSynchronize(st));
  hipFree(d_rowlen);
  const int64_t ns = (n + SELL_C - 1) / SELL_C;
  std::vector<int64_t> off(ns + 1);
  std::vector<int32_t> slen(ns);
  int64_t acc = 0;
  for (int64_t s = 0; s < ns; ++s) {
    int32_t m = 0;
    for (int64_t a = s * SELL_C; a < std::min<int64_t>(n, (s + 1) * SELL_C); ++a) m = std::max(m, rowlen[a]);
    slen[s] = m;
    off[s] = acc;
    acc += (int64_t)m * SELL_C;
  }
  off[ns] = acc;
  const int cw = CPLX ? 2 : 1;
  const double bytes = (double)acc * (4.0 + 8.0 * cw);
  if (bytes > op->asm_max_bytes)
    return fail("assembled table would take " + std::to_string(bytes / 1e9) + " GB (limit " +
                std::to_string(op->asm_max_bytes / 1e9) + " GB)");
  QES_HIP(hipMalloc(&op->asm_off, sizeof(int64_t) * (ns + 1)));
  QES_HIP(hipMalloc(&op->asm_len, sizeof(int32_t) * std::max<int64_t>(ns, 1)));
  QES_HIP(hipMalloc(&op->asm_col, sizeof(int32_t) * std::max<int64_t>(acc, 1)));
  QES_HIP(hipMalloc(&op->asm_val, sizeof(double) * cw * std::max<int64_t>(acc, 1)));
  QES_HIP(hipMemcpy(op->asm_off, off.data(), sizeof(int64_t) * (ns + 1), hipMemcpyHostToDevice));
  QES_HIP(hipMemcpy(op->asm_len, slen.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice));
  // rows past n in the last slice never exist: their entries stay unused
  // (the product kernel skips rows >= n), zero them for determinism
  QES_HIP(hipMemsetAsync(op->asm_col, 0, sizeof(int32_t) * acc, st));
  QES_HIP(hipMemsetAsync(op->asm_val, 0, sizeof(double) * cw * acc, st));
#define QES_ASM_FILL(CK) QES_ASM(CK, true, nullptr, op->asm_off, op->asm_len, op->asm_col, op->asm_val)
  QES_ASM_CK_SWITCH(P.canon, QES_ASM_FILL)
#undef QES_ASM_FILL
#undef QES_ASM
  QES_HIP(hipGetLastError());
  QES_HIP(hipStreamSynchronize(st));
  op->asm_entries = acc;
  op->asm_slices = ns;
  op->asm_vc = CPLX ? 1 : 0;
  return 0;
}

int op_assemble(Op* op, int enable, hipStream_t st) {
  free_assembly(op);
  if (!enable || op->basis->n == 0) return 0;
  if (op->basis->n >= (int64_t(1) << 31)) return fail("assembled operator needs fewer than 2^31 rows");
  const bool vc = op->is_complex || op->chi_complex;
  const bool w32 = op->basis->group.ns <= 32;
  int rc;
  if (w32) rc = vc ? assemble_t<uint32_t, true>(op, st) : assemble_t<uint32_t, false>(op, st);
  else rc = vc ? assemble_t<uint64_t, true>(op, st) : assemble_t<uint64_t, false>(op, st);
  if (rc) free_assembly(op);
  return rc;
}

int assembled_matvec(const Op* op, const double* x, double* y, int dtype, int accumulate, hipStream_t st) {
  const int64_t n = op->basis->n;
  if (n == 0) return 0;
  const dim3 grid((unsigned)((n + 255) / 256));
  const bool xc = dtype == 1;
  if (op->asm_vc && !xc) return fail("complex matrix elements need complex128 vectors");
  if (op->asm_vc)
    hipLaunchKernelGGL((sell_matvec_kernel<true, true>), grid, dim3(256), 0, st, op->asm_off, op->asm_len,
                       op->asm_col, op->asm_val, x, y, n, accumulate);
  else if (xc)
    hipLaunchKernelGGL((sell_matvec_kernel<false, true>), grid, dim3(256), 0, st, op->asm_off, op->asm_len,
                       op->asm_col, op->asm_val, x, y, n, accumulate);
  else
    hipLaunchKernelGGL((sell_matvec_kernel<false, false>), grid, dim3(256), 0, st, op->asm_off, op->asm_len,
                       op->asm_col, op->asm_val, x, y, n, accumulate);
  QES_HIP(hipGetLastError());
  return 0;
}

}  // namespace qes
