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
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "qes_internal.h"
#include "qes_rows.h"
#include "qes_sell.h"

namespace qes {


// Formats.  PLAIN: int32 column + f64 (or c128) value per entry, the
// diagonal as entry 0.  Real operators whose off-diagonal values take few
// distinct values (they are products of a coupling, a character and a norm
// ratio) store an index into a value dictionary instead: P4 packs column and
// index into one uint32 (column in the low bits), P6 keeps an int32 column
// and a uint16 index; the diagonal then lives in its own f64 array.
enum { PASS_COUNT = 0, PASS_DICT = 1, PASS_FILL = 2 };

static constexpr int DICT_HBITS = 16;                  // dictionary hash slots
static constexpr uint64_t DICT_EMPTY = ~0ull;          // a NaN bit pattern, never a value

__host__ __device__ __forceinline__ uint32_t dict_slot(uint64_t bits) {
  return (uint32_t)((bits * 0x9E3779B97F4A7C15ull) >> (64 - DICT_HBITS));
}

// insert (PASS_DICT) or find (PASS_FILL) a value's bit pattern
__device__ __forceinline__ uint32_t dict_insert(unsigned long long* tab, uint64_t bits) {
  uint32_t h = dict_slot(bits);
  const uint32_t mask = (1u << DICT_HBITS) - 1u;
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    const unsigned long long cur = tab[h];
    if (cur == bits) return h;
    if (cur == DICT_EMPTY) {
      const unsigned long long prev = atomicCAS(&tab[h], DICT_EMPTY, (unsigned long long)bits);
      if (prev == DICT_EMPTY || prev == bits) return h;
    }
    h = (h + 1) & mask;
  }
  return ~0u;   // table full: the host sees more than 2^16 distinct values and falls back
}

// every value met by the fill pass was inserted by the dictionary pass (same
// visitor, same arithmetic), so the probe ends at its slot; the bound only
// guards the loop
__device__ __forceinline__ uint32_t dict_find(const unsigned long long* __restrict__ tab, uint64_t bits) {
  uint32_t h = dict_slot(bits);
  const uint32_t mask = (1u << DICT_HBITS) - 1u;
  for (uint32_t probe = 0; probe <= mask && tab[h] != bits; ++probe) h = (h + 1) & mask;
  return h;
}

struct AsmOut {
  int fmt;
  int colbits;                          // P4: bits of the column field
  int32_t* rowlen;                      // PASS_COUNT
  unsigned long long* dict_tab;         // PASS_DICT / PASS_FILL: hash slots
  const uint16_t* slot_idx;             // PASS_FILL: dictionary index of each slot
  const int64_t* off;                   // slice offsets
  const int32_t* slen;                  // slice lengths (entries per row)
  int32_t* col;                         // PLAIN / P6
  double* val;                          // PLAIN
  uint32_t* packed;                     // P4
  uint16_t* vidx;                       // P6
  double* diag;                         // P4 / P6
};

template <typename W, bool CPLX, int CK, int PASS>
__global__ void __launch_bounds__(256) asm_rows_kernel(QesGroup G, QesLookup L, QesPlan P,
                                                       const double* __restrict__ chi_re,
                                                       const double* __restrict__ chi_im,
                                                       const QesTermGroup* __restrict__ groups, int n_diag,
                                                       const double* __restrict__ exval,
                                                       const uint32_t* __restrict__ lut,
                                                       const double* __restrict__ norms, int64_t n, AsmOut O) {
  extern __shared__ char smem[];
  const RowCtx C = load_ctx<CPLX>(G, L, P, chi_re, chi_im, groups, n_diag, exval, lut, smem);
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const W r = load_rep<W>(L, a);
  const double inv_na = 1.0 / norms[a];
  const int cw = CPLX ? 2 : 1;
  const int64_t s = a / SELL_C;
  const int l = (int)(a % SELL_C);
  const bool plain = O.fmt == ASM_PLAIN;
  const int64_t base = (PASS == PASS_FILL) ? O.off[s] + l : 0;
  int k = 0;
  if (PASS == PASS_FILL) {
    double dr = 0.0, di = 0.0;
    for (int t = 0; t < n_diag; ++t) {
      const QesTermGroup& g = groups[t];
      const int kk = term_bits<W>(g, r);
      dr += sel4(g.re, kk);
      if (CPLX) di += sel4(g.im, kk);
    }
    if (plain) {
      O.col[base] = (int32_t)a;
      O.val[base * cw] = dr;
      if (CPLX) O.val[base * cw + 1] = di;
    } else {
      O.diag[a] = dr;
    }
  }
  if (plain) ++k;   // entry 0 is the diagonal
  visit_row<W, CPLX, CK>(C, r, [&](int64_t b, double fr, double fi, int) {
    const double vr = fr * inv_na;
    if (PASS == PASS_DICT) {
      dict_insert(O.dict_tab, (uint64_t)__double_as_longlong(vr));
    } else if (PASS == PASS_FILL) {
      const int64_t e = base + (int64_t)k * SELL_C;
      if (plain) {
        O.col[e] = (int32_t)b;
        O.val[e * cw] = vr;
        if (CPLX) O.val[e * cw + 1] = fi * inv_na;
      } else {
        const uint32_t idx = O.slot_idx[dict_find(O.dict_tab, (uint64_t)__double_as_longlong(vr))];
        if (O.fmt == ASM_P4) {
          O.packed[e] = (uint32_t)b | (idx << O.colbits);
        } else {
          O.col[e] = (int32_t)b;
          O.vidx[e] = (uint16_t)idx;
        }
      }
    }
    ++k;
  });
  if (PASS == PASS_COUNT) {
    O.rowlen[a] = k;
  } else if (PASS == PASS_FILL) {
    // pad to the slice length with (own row, value 0)
    const uint32_t zidx = plain ? 0u : O.slot_idx[dict_find(O.dict_tab, 0ull)];
    for (; k < O.slen[s]; ++k) {
      const int64_t e = base + (int64_t)k * SELL_C;
      if (plain) {
        O.col[e] = (int32_t)a;
        O.val[e * cw] = 0.0;
        if (CPLX) O.val[e * cw + 1] = 0.0;
      } else if (O.fmt == ASM_P4) {
        O.packed[e] = (uint32_t)a | (zidx << O.colbits);
      } else {
        O.col[e] = (int32_t)a;
        O.vidx[e] = (uint16_t)zidx;
      }
    }
  }
}

// y = A x with the sliced table; XC: complex vectors
template <bool XC>
__global__ void __launch_bounds__(256) sell_matvec_kernel(AsmTable T, const double* __restrict__ x,
                                                          double* __restrict__ y, int64_t n, int accumulate) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  double ar, ai;
  sell_row<XC>(T, a, x, ar, ai);
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
  for (void* p : {(void*)op->asm_off, (void*)op->asm_len, (void*)op->asm_col, (void*)op->asm_val,
                  (void*)op->asm_packed, (void*)op->asm_vidx, (void*)op->asm_dict, (void*)op->asm_diag})
    if (p) hipFree(p);
  op->asm_off = nullptr;
  op->asm_len = nullptr;
  op->asm_col = nullptr;
  op->asm_val = nullptr;
  op->asm_packed = nullptr;
  op->asm_vidx = nullptr;
  op->asm_dict = nullptr;
  op->asm_diag = nullptr;
  op->asm_entries = 0;
  op->asm_vc = 0;
  op->asm_fmt = -1;
  op->asm_bytes = 0;
}

#define QES_ASM_CK_SWITCH(CK_VAR, MACRO)                   \
  switch (CK_VAR) {                                        \
    case QES_CANON_CYCLIC: MACRO(QES_CANON_CYCLIC); break; \
    case QES_CANON_ROWS: MACRO(QES_CANON_ROWS); break;     \
    default: MACRO(QES_CANON_BRUTE); break;                \
  }

static int bits_for(int64_t v) {   // bits needed to hold 0 .. v-1
  int b = 0;
  while ((int64_t(1) << b) < v) ++b;
  return b;
}

template <typename W, bool CPLX>
static int assemble_t(Op* op, int want_dict, hipStream_t st) {
  const Basis* B = op->basis;
  const int64_t n = B->n;
  const QesPlan& P = op->plan;
  const dim3 grid((unsigned)((n + 255) / 256));
  AsmOut O{};
#define QES_ASM(CK, PASS)                                                                                           \
  hipLaunchKernelGGL((asm_rows_kernel<W, CPLX, CK, PASS>), grid, dim3(256), op->lds_bytes, st, B->group,          \
                     B->lookup(), P, B->chi_re, B->chi_im, op->d_gen, op->n_diag, op->d_exval, op->d_lut, B->norms, \
                     n, O)
#define QES_ASM_COUNT(CK) QES_ASM(CK, PASS_COUNT)
#define QES_ASM_DICT(CK) QES_ASM(CK, PASS_DICT)
#define QES_ASM_FILL(CK) QES_ASM(CK, PASS_FILL)
  // pass A: row lengths (diagonal included)
  QES_HIP(hipMalloc(&O.rowlen, sizeof(int32_t) * n));
  QES_ASM_CK_SWITCH(P.canon, QES_ASM_COUNT)
  QES_HIP(hipGetLastError());
  std::vector<int32_t> rowlen(n);
  QES_HIP(hipMemcpyAsync(rowlen.data(), O.rowlen, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
  QES_HIP(hipStreamSynchronize(st));
  hipFree(O.rowlen);
  O.rowlen = nullptr;

  // pass B (real operators): value dictionary of the off-diagonal elements
  int fmt = ASM_PLAIN;
  std::vector<double> dict;
  std::vector<uint16_t> slot_idx;
  if (!CPLX && want_dict) {
    const size_t nslot = size_t(1) << DICT_HBITS;
    QES_HIP(hipMalloc(&O.dict_tab, sizeof(unsigned long long) * nslot));
    std::vector<unsigned long long> tab(nslot, DICT_EMPTY);
    tab[dict_slot(0ull)] = 0ull;   // padding value 0.0 is always present
    QES_HIP(hipMemcpy(O.dict_tab, tab.data(), sizeof(unsigned long long) * nslot, hipMemcpyHostToDevice));
    QES_ASM_CK_SWITCH(P.canon, QES_ASM_DICT)
    QES_HIP(hipGetLastError());
    QES_HIP(hipMemcpyAsync(tab.data(), O.dict_tab, sizeof(unsigned long long) * nslot, hipMemcpyDeviceToHost, st));
    QES_HIP(hipStreamSynchronize(st));
    std::vector<std::pair<double, uint32_t>> vals;
    for (uint32_t h = 0; h < nslot; ++h)
      if (tab[h] != DICT_EMPTY) {
        double v;
        memcpy(&v, &tab[h], sizeof(double));
        vals.push_back({v, h});
      }
    const int64_t nd = (int64_t)vals.size();
    if (nd < (int64_t)nslot * 3 / 4) {   // a full table means the insert gave up: no dictionary
      std::sort(vals.begin(), vals.end(), [](const std::pair<double, uint32_t>& x, const std::pair<double, uint32_t>& y) {
        return x.first < y.first || (x.first == y.first && x.second < y.second);
      });
      slot_idx.assign(nslot, 0);
      for (int64_t i = 0; i < nd; ++i) {
        dict.push_back(vals[i].first);
        slot_idx[vals[i].second] = (uint16_t)i;
      }
      const int ib = bits_for(nd), cb = bits_for(n);
      if (ib + cb <= 32) {
        fmt = ASM_P4;
        O.colbits = 32 - ib;
      } else if (nd <= 65536) {
        fmt = ASM_P6;
      }
    }
  }

  // slices: PLAIN holds the diagonal as entry 0, dictionary formats do not
  const int drop = fmt == ASM_PLAIN ? 0 : 1;
  const int64_t ns = (n + SELL_C - 1) / SELL_C;
  std::vector<int64_t> off(ns + 1);
  std::vector<int32_t> slen(ns);
  int64_t acc = 0;
  for (int64_t s = 0; s < ns; ++s) {
    int32_t m = 0;
    for (int64_t a = s * SELL_C; a < std::min<int64_t>(n, (s + 1) * SELL_C); ++a) m = std::max(m, rowlen[a] - drop);
    slen[s] = m;
    off[s] = acc;
    acc += (int64_t)m * SELL_C;
  }
  off[ns] = acc;
  const int cw = CPLX ? 2 : 1;
  double per_entry = fmt == ASM_PLAIN ? 4.0 + 8.0 * cw : (fmt == ASM_P4 ? 4.0 : 6.0);
  const double bytes = (double)acc * per_entry + (fmt == ASM_PLAIN ? 0.0 : 8.0 * n) + 12.0 * ns;
  if (bytes > op->asm_max_bytes) {
    if (O.dict_tab) hipFree(O.dict_tab);
    return fail("assembled table would take " + std::to_string(bytes / 1e9) + " GB (limit " +
                std::to_string(op->asm_max_bytes / 1e9) + " GB)");
  }
  O.fmt = fmt;
  QES_HIP(hipMalloc(&op->asm_off, sizeof(int64_t) * (ns + 1)));
  QES_HIP(hipMalloc(&op->asm_len, sizeof(int32_t) * std::max<int64_t>(ns, 1)));
  QES_HIP(hipMemcpy(op->asm_off, off.data(), sizeof(int64_t) * (ns + 1), hipMemcpyHostToDevice));
  QES_HIP(hipMemcpy(op->asm_len, slen.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice));
  const size_t ne = (size_t)std::max<int64_t>(acc, 1);
  if (fmt == ASM_PLAIN) {
    QES_HIP(hipMalloc(&op->asm_col, sizeof(int32_t) * ne));
    QES_HIP(hipMalloc(&op->asm_val, sizeof(double) * cw * ne));
    // entries of rows past n in the last slice are never read; zero them
    QES_HIP(hipMemsetAsync(op->asm_col, 0, sizeof(int32_t) * ne, st));
    QES_HIP(hipMemsetAsync(op->asm_val, 0, sizeof(double) * cw * ne, st));
  } else {
    QES_HIP(hipMalloc(&op->asm_diag, sizeof(double) * n));
    QES_HIP(hipMalloc(&op->asm_dict, sizeof(double) * dict.size()));
    QES_HIP(hipMemcpy(op->asm_dict, dict.data(), sizeof(double) * dict.size(), hipMemcpyHostToDevice));
    uint16_t* d_slot_idx = nullptr;
    QES_HIP(hipMalloc(&d_slot_idx, sizeof(uint16_t) * slot_idx.size()));
    QES_HIP(hipMemcpy(d_slot_idx, slot_idx.data(), sizeof(uint16_t) * slot_idx.size(), hipMemcpyHostToDevice));
    O.slot_idx = d_slot_idx;
    if (fmt == ASM_P4) {
      QES_HIP(hipMalloc(&op->asm_packed, sizeof(uint32_t) * ne));
      QES_HIP(hipMemsetAsync(op->asm_packed, 0, sizeof(uint32_t) * ne, st));
    } else {
      QES_HIP(hipMalloc(&op->asm_col, sizeof(int32_t) * ne));
      QES_HIP(hipMalloc(&op->asm_vidx, sizeof(uint16_t) * ne));
      QES_HIP(hipMemsetAsync(op->asm_col, 0, sizeof(int32_t) * ne, st));
      QES_HIP(hipMemsetAsync(op->asm_vidx, 0, sizeof(uint16_t) * ne, st));
    }
  }
  O.off = op->asm_off;
  O.slen = op->asm_len;
  O.col = op->asm_col;
  O.val = op->asm_val;
  O.packed = op->asm_packed;
  O.vidx = op->asm_vidx;
  O.diag = op->asm_diag;
  // pass C: the table
  QES_ASM_CK_SWITCH(P.canon, QES_ASM_FILL)
  QES_HIP(hipGetLastError());
  QES_HIP(hipStreamSynchronize(st));
#undef QES_ASM_COUNT
#undef QES_ASM_DICT
#undef QES_ASM_FILL
#undef QES_ASM
  if (O.dict_tab) hipFree(O.dict_tab);
  if (O.slot_idx) hipFree((void*)O.slot_idx);
  op->asm_entries = acc;
  op->asm_slices = ns;
  op->asm_vc = CPLX ? 1 : 0;
  op->asm_fmt = fmt;
  op->asm_colbits = O.colbits;
  op->asm_ndict = (int)dict.size();
  op->asm_bytes = (int64_t)bytes + (int64_t)dict.size() * 8;
  return 0;
}

int op_assemble(Op* op, int enable, int want_dict, hipStream_t st) {
  free_assembly(op);
  if (!enable || op->basis->n == 0) return 0;
  if (op->basis->n >= (int64_t(1) << 31)) return fail("assembled operator needs fewer than 2^31 rows");
  const bool vc = op->is_complex || op->chi_complex;
  const bool w32 = op->basis->group.ns <= 32;
  int rc;
  if (w32) rc = vc ? assemble_t<uint32_t, true>(op, want_dict, st) : assemble_t<uint32_t, false>(op, want_dict, st);
  else rc = vc ? assemble_t<uint64_t, true>(op, want_dict, st) : assemble_t<uint64_t, false>(op, want_dict, st);
  if (rc) free_assembly(op);
  return rc;
}

// Coordinate list of the assembled table: slot e of row a -> (a, col, value);
// dictionary formats append the diagonal at entries + a.  Slots of rows past
// n in the last slice are left untouched (callers pre-fill rows with -1).
__global__ void __launch_bounds__(256) asm_coo_kernel(AsmTable T, int64_t n, int64_t entries, int32_t* __restrict__ rows,
                                                      int32_t* __restrict__ cols, double* __restrict__ vals) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const int64_t s = a / SELL_C;
  const int l = (int)(a % SELL_C);
  const int64_t base = T.off[s] + l;
  const int len = T.len[s];
  const uint32_t cmask = (T.colbits >= 32) ? ~0u : ((1u << T.colbits) - 1u);
  const int vw = T.vc ? 2 : 1;
  for (int k = 0; k < len; ++k) {
    const int64_t e = base + (int64_t)k * SELL_C;
    rows[e] = (int32_t)a;
    if (T.fmt == ASM_PLAIN) {
      cols[e] = T.col[e];
      vals[e * vw] = T.val[e * vw];
      if (T.vc) vals[e * vw + 1] = T.val[e * vw + 1];
    } else if (T.fmt == ASM_P4) {
      const uint32_t p = T.packed[e];
      cols[e] = (int32_t)(p & cmask);
      vals[e] = T.dict[p >> T.colbits];
    } else {
      cols[e] = T.col[e];
      vals[e] = T.dict[T.vidx[e]];
    }
  }
  if (T.fmt != ASM_PLAIN) {
    rows[entries + a] = (int32_t)a;
    cols[entries + a] = (int32_t)a;
    vals[entries + a] = T.diag[a];
  }
}

int assembled_coo(const Op* op, int32_t* rows, int32_t* cols, double* vals, int64_t capacity, hipStream_t st) {
  if (op->asm_fmt < 0) return fail("operator is not assembled");
  const int64_t n = op->basis->n;
  const int64_t need = op->asm_entries + (op->asm_fmt == ASM_PLAIN ? 0 : n);
  if (capacity < need) return fail("coordinate buffers hold " + std::to_string(capacity) + " entries, need " +
                                   std::to_string(need));
  if (n == 0) return 0;
  const dim3 grid((unsigned)((n + 255) / 256));
  hipLaunchKernelGGL(asm_coo_kernel, grid, dim3(256), 0, st, asm_table_of(op), n, op->asm_entries, rows, cols, vals);
  QES_HIP(hipGetLastError());
  return 0;
}

int assembled_matvec(const Op* op, const double* x, double* y, int dtype, int accumulate, hipStream_t st) {
  const int64_t n = op->basis->n;
  if (n == 0) return 0;
  const dim3 grid((unsigned)((n + 255) / 256));
  const bool xc = dtype == 1;
  if (op->asm_vc && !xc) return fail("complex matrix elements need complex128 vectors");
  const AsmTable T = asm_table_of(op);
  if (xc) hipLaunchKernelGGL((sell_matvec_kernel<true>), grid, dim3(256), 0, st, T, x, y, n, accumulate);
  else hipLaunchKernelGGL((sell_matvec_kernel<false>), grid, dim3(256), 0, st, T, x, y, n, accumulate);
  QES_HIP(hipGetLastError());
  return 0;
}

}  // namespace qes
