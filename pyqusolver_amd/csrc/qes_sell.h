// Row products of the assembled (sliced ELLPACK) operator, shared by the
// single-operator product kernels and the batched matvec.  A table is
// described by AsmTable; rows are grouped in slices of 64 (one wavefront),
// entry k of row 64s+l at off[s] + 64k + l.
#pragma once
#include "qes_internal.h"

namespace qes {

static constexpr int SELL_C = 64;
enum { ASM_PLAIN = 0, ASM_P4 = 1, ASM_P6 = 2 };

struct AsmTable {
  int32_t fmt;          // -1 none, ASM_PLAIN / ASM_P4 / ASM_P6
  int32_t vc;           // complex values (PLAIN only)
  int32_t colbits;      // P4 column field width
  int32_t pad;
  const int64_t* off;
  const int32_t* len;
  const int32_t* col;   // PLAIN / P6
  const double* val;    // PLAIN (1 or 2 doubles per entry)
  const uint32_t* packed;   // P4
  const uint16_t* vidx;     // P6
  const double* dict;       // P4 / P6
  const double* diag;       // P4 / P6
};

// (ar, ai) = row a of A times x; XC: complex x (interleaved)
template <bool XC>
__device__ __forceinline__ void sell_row(const AsmTable& T, int64_t a, const double* __restrict__ x, double& ar,
                                         double& ai) {
  const int64_t s = a / SELL_C;
  const int l = (int)(a % SELL_C);
  const int64_t base = T.off[s] + l;
  const int len = T.len[s];
  ar = 0.0;
  ai = 0.0;
  if (T.fmt == ASM_PLAIN) {
    const int vw = T.vc ? 2 : 1;
    for (int k = 0; k < len; ++k) {
      const int64_t e = base + (int64_t)k * SELL_C;
      const int64_t b = T.col[e];
      const double vr = T.val[e * vw];
      if (XC) {
        const double vi = T.vc ? T.val[e * vw + 1] : 0.0;
        const double xr = x[2 * b], xi = x[2 * b + 1];
        ar += vr * xr - vi * xi;
        ai += vr * xi + vi * xr;
      } else {
        ar += vr * x[b];
      }
    }
    return;
  }
  const double d = T.diag[a];
  if (XC) {
    ar = d * x[2 * a];
    ai = d * x[2 * a + 1];
  } else {
    ar = d * x[a];
  }
  const uint32_t cmask = (T.colbits >= 32) ? ~0u : ((1u << T.colbits) - 1u);
  for (int k = 0; k < len; ++k) {
    const int64_t e = base + (int64_t)k * SELL_C;
    int64_t b;
    double v;
    if (T.fmt == ASM_P4) {
      const uint32_t p = T.packed[e];
      b = p & cmask;
      v = T.dict[p >> T.colbits];
    } else {
      b = T.col[e];
      v = T.dict[T.vidx[e]];
    }
    if (XC) {
      ar += v * x[2 * b];
      ai += v * x[2 * b + 1];
    } else {
      ar += v * x[b];
    }
  }
}

inline AsmTable asm_table_of(const Op* op) {
  AsmTable T{};
  T.fmt = op->asm_fmt;
  T.vc = op->asm_vc;
  T.colbits = op->asm_colbits;
  T.off = op->asm_off;
  T.len = op->asm_len;
  T.col = op->asm_col;
  T.val = op->asm_val;
  T.packed = op->asm_packed;
  T.vidx = op->asm_vidx;
  T.dict = op->asm_dict;
  T.diag = op->asm_diag;
  return T;
}

}  // namespace qes
