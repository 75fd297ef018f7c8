// Row products of the assembled (sliced ELLPACK) operator, shared by the
// single-operator product kernels and the batched matvec.  A table is
// described by AsmTable; rows are grouped in slices of 64 (one wavefront),
// entry k of row 64s+l at off[s] + 64k + l.
#pragma once
#include "qes_internal.h"

namespace qes {

static constexpr int SELL_C = 64;
#ifndef QES_SELL_UNROLL
#define QES_SELL_UNROLL 4   // slots per lane in flight in sell_row_t
#endif
static constexpr int ASM_MAX_CPL = 8;   // coupling ids of a coupling-tagged table
#ifndef QES_SELL_NT
#define QES_SELL_NT 1
#endif

// load of a table word that is read once per product: with QES_SELL_NT the
// load carries the streaming (non-temporal) hint, so the table passing
// through leaves the caches to the gathered x lines (measured: 0.98 -> 0.95
// ms at L=32, 62.0 -> 60.7 ms at 6x6)
template <typename T>
__device__ __forceinline__ T stream_load(const T* p) {
#if QES_SELL_NT
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}
enum { ASM_PLAIN = 0, ASM_P4 = 1, ASM_P6 = 2 };

struct AsmTable {
  int32_t fmt;          // -1 none, ASM_PLAIN / ASM_P4 / ASM_P6
  int32_t vc;           // PLAIN: complex values
  int32_t dvc;          // P4 / P6: complex dictionary values and diagonal
  int32_t colbits;      // P4 column field width
  int32_t ncpl;         // > 0: diagonal = sum_c dscale[c] * diag_c[c] (per-coupling diagonals)
  int32_t nodiag;       // column panel after the first: no diagonal term, rows of empty slices untouched
  int64_t nrows;        // row count (stride of diag_c)
  const int64_t* off;
  const int32_t* len;
  const int32_t* col;       // PLAIN / P6
  const double* val;        // PLAIN (1 or 2 doubles per entry)
  const uint32_t* packed;   // P4
  const uint16_t* vidx;     // P6
  const double* dict;       // P4 / P6 (1 or 2 doubles per value)
  const double* diag;       // P4 / P6 total diagonal (1 or 2 doubles per row)
  const double* diag_c;     // per-coupling diagonals [c][row] (1 or 2 doubles per row)
  double dscale[ASM_MAX_CPL];
};

// entry e of a dictionary table -> (column, dictionary index)
__device__ __forceinline__ void sell_entry(const AsmTable& T, int64_t e, int64_t& b, uint32_t& idx) {
  if (T.fmt == ASM_P4) {
    const uint32_t p = stream_load(T.packed + e);
    const uint32_t cmask = (T.colbits >= 32) ? ~0u : ((1u << T.colbits) - 1u);
    b = p & cmask;
    idx = p >> T.colbits;
  } else {
    b = stream_load(T.col + e);
    idx = stream_load(T.vidx + e);
  }
}

// Row product specialised on the table format (FMT), complex vectors (XC)
// and complex values (CV: complex dictionary / plain values).  Slots are
// consumed four at a time: the four entry loads, then the four gathers and
// value loads, then the four multiply-adds in slot order, so each lane keeps
// several independent gathers in flight; the sums are formed in the same
// order as sell_row, so the results are identical.
template <int FMT, bool XC, bool CV>
__device__ __forceinline__ void sell_row_t(const AsmTable& T, int64_t a, const double* __restrict__ x, double& ar,
                                           double& ai) {
  constexpr int U = QES_SELL_UNROLL;
  constexpr int VW = CV ? 2 : 1;
  const int64_t s = a / SELL_C;
  const int l = (int)(a % SELL_C);
  const int64_t base = T.off[s] + l;
  const int len = T.len[s];
  ar = 0.0;
  ai = 0.0;
  if (FMT != ASM_PLAIN && !T.nodiag) {
    double dr, di = 0.0;
    if (T.ncpl > 0) {
      dr = 0.0;
#pragma unroll
      for (int c = 0; c < ASM_MAX_CPL; ++c) {
        if (c < T.ncpl) {
          const double* dc = T.diag_c + ((int64_t)c * T.nrows + a) * VW;
          dr += T.dscale[c] * stream_load(dc);
          if (CV) di += T.dscale[c] * stream_load(dc + 1);
        }
      }
    } else {
      dr = stream_load(T.diag + a * VW);
      if (CV) di = stream_load(T.diag + a * VW + 1);
    }
    if (XC) {
      const double xr = x[2 * a], xi = x[2 * a + 1];
      ar = dr * xr - di * xi;
      ai = dr * xi + di * xr;
    } else {
      ar = dr * x[a];
    }
  }
  const uint32_t cmask = (T.colbits >= 32) ? ~0u : ((1u << T.colbits) - 1u);
  auto entry = [&](int64_t e, int64_t& b, double& vr, double& vi) {
    if (FMT == ASM_P4) {
      const uint32_t p = stream_load(T.packed + e);
      b = p & cmask;
      const uint32_t idx = p >> T.colbits;
      vr = T.dict[idx * VW];
      vi = CV ? T.dict[idx * VW + 1] : 0.0;
    } else if (FMT == ASM_P6) {
      b = stream_load(T.col + e);
      const uint32_t idx = stream_load(T.vidx + e);
      vr = T.dict[idx * VW];
      vi = CV ? T.dict[idx * VW + 1] : 0.0;
    } else {
      b = stream_load(T.col + e);
      vr = stream_load(T.val + e * VW);
      vi = CV ? stream_load(T.val + e * VW + 1) : 0.0;
    }
  };
  auto accumulate = [&](int64_t b, double vr, double vi) {
    if (XC) {
      const double xr = x[2 * b], xi = x[2 * b + 1];
      ar += vr * xr - vi * xi;
      ai += vr * xi + vi * xr;
    } else {
      ar += vr * x[b];
    }
  };
  int k = 0;
  for (; k + U <= len; k += U) {
    int64_t b[U];
    double vr[U], vi[U], xr[U], xi[U];
#pragma unroll
    for (int u = 0; u < U; ++u) entry(base + (int64_t)(k + u) * SELL_C, b[u], vr[u], vi[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      xr[u] = XC ? x[2 * b[u]] : x[b[u]];
      xi[u] = XC ? x[2 * b[u] + 1] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (XC) {
        ar += vr[u] * xr[u] - vi[u] * xi[u];
        ai += vr[u] * xi[u] + vi[u] * xr[u];
      } else {
        ar += vr[u] * xr[u];
      }
    }
  }
  for (; k < len; ++k) {
    int64_t b;
    double vr, vi;
    entry(base + (int64_t)k * SELL_C, b, vr, vi);
    accumulate(b, vr, vi);
  }
}

inline AsmTable asm_table_of(const Op* op) {
  AsmTable T{};
  T.fmt = op->asm_fmt;
  T.vc = op->asm_vc;
  T.dvc = op->asm_dvc;
  T.colbits = op->asm_colbits;
  T.ncpl = 0;
  T.nodiag = 0;
  T.nrows = op->basis->n;
  T.off = op->asm_off;
  T.len = op->asm_len;
  T.col = op->asm_col;
  T.val = op->asm_val;
  T.packed = op->asm_packed;
  T.vidx = op->asm_vidx;
  T.dict = op->asm_dict;
  T.diag = op->asm_diag;
  T.diag_c = op->asm_diag_c;
  return T;
}

}  // namespace qes
