// Assembled operator: the pull-form rows of H tabulated once in a sliced
// ELLPACK layout (64 rows per slice = one wavefront, column-major inside the
// slice), for callers that apply the same H many times (Lanczos).
//
// The reference offers the same trade: the explicit sparse build
// (build_operator_matrix / _build_sparse_same_sector_compact_jit,
// Algebra/Hilbert/matrix_builder.py:151-206, :911-1030) next to the
// matrix-free matvec_fun.  Here the table is built on the device by the same
// row visitor the matrix-free kernel uses, in the same order and with the
// same arithmetic (the diagonal first, then the off-diagonal elements in
// visit order with their value already divided by N_a), so the assembled
// product is the matrix-free one up to the contraction of the final adds.
//
// Slice s covers rows 64s .. 64s+63; its len[s] = longest row in the slice;
// entry k of row 64s+l sits at off[s] + 64k + l.  Short rows are padded with
// (col = own row, value 0).  A matvec then streams the table with perfectly
// coalesced loads and gathers x[col]: no canonicalisation, no lookup.
//
// Formats.  PLAIN: int32 column + f64 (or c128) value per entry, the
// diagonal as entry 0.  When the off-diagonal values take fewer than 2^16
// distinct values (products of a coupling, a character and a norm ratio) an
// entry stores an index into a value dictionary instead: P4 packs column and
// index into one uint32 (column in the low bits), P6 keeps an int32 column
// and a uint16 index; the diagonal lives in its own array.  A
// coupling-tagged table keys its dictionary by (value, coupling id) and keeps
// one diagonal per coupling, which gives H(J) x for any scale of each coupling
// and the per-coupling products Re<g, H_c v> from one pass over the table.
#include <algorithm>
#include <mutex>
#include <cmath>
#include <cstring>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "qes_internal.h"
#include "qes_rows.h"
#include "qes_sell.h"

namespace qes {

enum { PASS_COUNT = 0, PASS_DICT = 1, PASS_FILL = 2 };

static constexpr int DICT_HBITS = 16;                  // dictionary hash slots
static constexpr uint64_t DICT_EMPTY = ~0ull;

// fingerprint of a dictionary key (value re, value im, coupling id)
__host__ __device__ __forceinline__ uint64_t dict_key(double vr, double vi, int cpl) {
  uint64_t a, b;
  memcpy(&a, &vr, 8);
  memcpy(&b, &vi, 8);
  uint64_t h = a ^ ((b << 29) | (b >> 35)) ^ ((uint64_t)(cpl + 1) * 0xD6E8FEB86659FD93ull);
  h ^= h >> 31;
  h *= 0x7FB5D329728EA185ull;
  h ^= h >> 27;
  h *= 0x81DADEF4BC2DD44Dull;
  h ^= h >> 33;
  return h == DICT_EMPTY ? 0x9E3779B97F4A7C15ull : h;
}

__host__ __device__ __forceinline__ uint32_t dict_slot(uint64_t key) {
  return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> (64 - DICT_HBITS));
}

struct DictBuf {
  unsigned long long* key;   // slots
  double* vr;                // value of the slot (written by the inserting lane)
  double* vi;
  int32_t* cpl;
  const uint16_t* idx;       // PASS_FILL: dictionary index of each slot
  int32_t* bad;              // PASS_FILL: set when a probe finds another value under the key
};

__device__ __forceinline__ void dict_insert(const DictBuf& D, double vr, double vi, int cpl) {
  const uint64_t key = dict_key(vr, vi, cpl);
  uint32_t h = dict_slot(key);
  const uint32_t mask = (1u << DICT_HBITS) - 1u;
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    const unsigned long long cur = D.key[h];
    if (cur == key) return;
    if (cur == DICT_EMPTY) {
      const unsigned long long prev = atomicCAS(&D.key[h], DICT_EMPTY, (unsigned long long)key);
      if (prev == DICT_EMPTY) {
        D.vr[h] = vr;
        D.vi[h] = vi;
        D.cpl[h] = cpl;
        return;
      }
      if (prev == key) return;
    }
    h = (h + 1) & mask;
  }
  // table full: the host sees a full table and does not use a dictionary
}

// every key met by the fill pass was inserted by the dictionary pass (same
// visitor, same arithmetic); the bound only guards the loop, and the stored
// value is compared so that two values sharing a fingerprint are caught
__device__ __forceinline__ uint32_t dict_find(const DictBuf& D, double vr, double vi, int cpl) {
  const uint64_t key = dict_key(vr, vi, cpl);
  uint32_t h = dict_slot(key);
  const uint32_t mask = (1u << DICT_HBITS) - 1u;
  for (uint32_t probe = 0; probe <= mask && D.key[h] != key; ++probe) h = (h + 1) & mask;
  if (D.key[h] != key || D.vr[h] != vr || D.vi[h] != vi || D.cpl[h] != cpl) {
    atomicOr(D.bad, 1);
    return 0;
  }
  return D.idx[h];
}

struct AsmOut {
  int fmt;
  int colbits;                          // P4: bits of the column field
  int tagged;                           // dictionary keyed by coupling id, per-coupling diagonals
  int ncpl;                             // coupling ids of the operator
  int32_t* rowlen;                      // PASS_COUNT
  DictBuf D;                            // PASS_DICT / PASS_FILL
  const int32_t* excpl;                 // coupling id of exchange element (class, site)
  const int64_t* off;                   // slice offsets
  const int32_t* slen;                  // slice lengths (entries per row)
  int32_t* col;                         // PLAIN / P6
  double* val;                          // PLAIN
  uint32_t* packed;                     // P4
  uint16_t* vidx;                       // P6
  double* diag;                         // P4 / P6
  double* diag_c;                       // tagged: [c][row]
};

template <typename W, bool CPLX, int CK, int PASS, bool LEAN = false>
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
    double cr[ASM_MAX_CPL], ci[ASM_MAX_CPL];
#pragma unroll
    for (int c = 0; c < ASM_MAX_CPL; ++c) cr[c] = ci[c] = 0.0;
    for (int t = 0; t < n_diag; ++t) {
      const QesTermGroup& g = groups[t];
      const int kk = term_bits<W>(g, r);
      const double vr = sel4(g.re, kk), vi = CPLX ? sel4(g.im, kk) : 0.0;
      dr += vr;
      if (CPLX) di += vi;
      if (O.tagged) {
#pragma unroll
        for (int c = 0; c < ASM_MAX_CPL; ++c)
          if (c == g.coupling) {
            cr[c] += vr;
            ci[c] += vi;
          }
      }
    }
    if (plain) {
      O.col[base] = (int32_t)a;
      O.val[base * cw] = dr;
      if (CPLX) O.val[base * cw + 1] = di;
    } else {
      O.diag[a * cw] = dr;
      if (CPLX) O.diag[a * cw + 1] = di;
      if (O.tagged) {
#pragma unroll
        for (int c = 0; c < ASM_MAX_CPL; ++c) {
          if (c < O.ncpl) {
            O.diag_c[((int64_t)c * n + a) * cw] = cr[c];
            if (CPLX) O.diag_c[((int64_t)c * n + a) * cw + 1] = ci[c];
          }
        }
      }
    }
  }
  if (plain) ++k;   // entry 0 is the diagonal
  visit_row<W, CPLX, CK, LEAN>(C, r, [&](int64_t b, double fr, double fi, int tag) {
    const double vr = fr * inv_na;
    const double vi = CPLX ? fi * inv_na : 0.0;
    int cpl = 0;
    if (O.tagged) cpl = tag >= 0 ? O.excpl[tag] : groups[-1 - tag].coupling;
    if (PASS == PASS_DICT) {
      dict_insert(O.D, vr, vi, cpl);
    } else if (PASS == PASS_FILL) {
      const int64_t e = base + (int64_t)k * SELL_C;
      if (plain) {
        O.col[e] = (int32_t)b;
        O.val[e * cw] = vr;
        if (CPLX) O.val[e * cw + 1] = vi;
      } else {
        const uint32_t idx = dict_find(O.D, vr, vi, cpl);
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
    const uint32_t zidx = plain ? 0u : dict_find(O.D, 0.0, 0.0, 0);
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

// y = A x with the sliced table; XC: complex vectors, FMT / CV: table format
// and complex values
template <bool XC, int FMT, bool CV>
__global__ void __launch_bounds__(256) sell_matvec_kernel(AsmTable T, const double* __restrict__ x,
                                                          double* __restrict__ y, int64_t n, int accumulate) {
  // grid-stride over rows (one pass when the grid covers n)
  for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < n; a += (int64_t)gridDim.x * blockDim.x) {
    if (T.nodiag && T.len[a / SELL_C] == 0) continue;   // a later panel with nothing in this slice
    double ar, ai;
    sell_row_t<FMT, XC, CV>(T, a, x, ar, ai);
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
}

// Y[:, c0:c0+RC] = A X[:, c0:c0+RC] for (n, nb) row-major X, Y: each slot is
// read once for the RC columns, whose values at the neighbour are contiguous;
// per column the sums run in the order of the single-vector product, so each
// column equals it bit for bit.
template <int FMT, bool XC, bool CV, int RC>
__global__ void __launch_bounds__(256) sell_matvec_multi_kernel(AsmTable T, const double* __restrict__ x,
                                                                double* __restrict__ y, int64_t n, int nb, int c0,
                                                                int accumulate) {
  constexpr int VW = CV ? 2 : 1;
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const int cw = XC ? 2 : 1;
  double ar[RC], ai[RC];
  if (FMT != ASM_PLAIN) {
    const double dr = T.diag[a * VW], di = CV ? T.diag[a * VW + 1] : 0.0;
#pragma unroll
    for (int r = 0; r < RC; ++r) {
      const int64_t e = (a * nb + c0 + r) * cw;
      if (XC) {
        const double xr = x[e], xi = x[e + 1];
        ar[r] = dr * xr - di * xi;
        ai[r] = dr * xi + di * xr;
      } else {
        ar[r] = dr * x[e];
        ai[r] = 0.0;
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < RC; ++r) ar[r] = ai[r] = 0.0;
  }
  const int64_t s = a / SELL_C;
  const int l = (int)(a % SELL_C);
  const int64_t base = T.off[s] + l;
  const int len = T.len[s];
  const uint32_t cmask = (T.colbits >= 32) ? ~0u : ((1u << T.colbits) - 1u);
  for (int k = 0; k < len; ++k) {
    const int64_t e = base + (int64_t)k * SELL_C;
    int64_t b;
    double vr, vi = 0.0;
    if (FMT == ASM_P4) {
      const uint32_t p = T.packed[e];
      b = p & cmask;
      const uint32_t idx = p >> T.colbits;
      vr = T.dict[idx * VW];
      if (CV) vi = T.dict[idx * VW + 1];
    } else if (FMT == ASM_P6) {
      b = T.col[e];
      const uint32_t idx = T.vidx[e];
      vr = T.dict[idx * VW];
      if (CV) vi = T.dict[idx * VW + 1];
    } else {
      b = T.col[e];
      vr = T.val[e * VW];
      if (CV) vi = T.val[e * VW + 1];
    }
#pragma unroll
    for (int r = 0; r < RC; ++r) {
      const int64_t f = (b * nb + c0 + r) * cw;
      if (XC) {
        const double xr = x[f], xi = x[f + 1];
        ar[r] += vr * xr - vi * xi;
        ai[r] += vr * xi + vi * xr;
      } else {
        ar[r] += vr * x[f];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < RC; ++r) {
    const int64_t e = (a * nb + c0 + r) * cw;
    if (accumulate) {
      y[e] += ar[r];
      if (XC) y[e + 1] += ai[r];
    } else {
      y[e] = ar[r];
      if (XC) y[e + 1] = ai[r];
    }
  }
}

// d_c = sum_a conj(g_a) (H_c v)_a for every coupling c of a tagged table, in
// Row a of the coupling-resolved product: (H_c v)_a for every coupling id c
// (the table entries split by the coupling of their dictionary value, the
// per-coupling diagonals), and g_a.
template <bool XC>
__device__ __forceinline__ void dots_row(const AsmTable& T, const uint8_t* __restrict__ dict_cpl,
                                         const double* __restrict__ v, const double* __restrict__ g, int64_t a,
                                         double* hr, double* hi, double& gr, double& gi) {
  const int64_t s = a / SELL_C;
  const int l = (int)(a % SELL_C);
  const int64_t base = T.off[s] + l;
  const int len = T.len[s];
  const int dw = T.dvc ? 2 : 1;
  double vr_a, vi_a = 0.0;
  if (XC) {
    vr_a = v[2 * a];
    vi_a = v[2 * a + 1];
    gr = g[2 * a];
    gi = g[2 * a + 1];
  } else {
    vr_a = v[a];
    gr = g[a];
    gi = 0.0;
  }
#pragma unroll
  for (int c = 0; c < ASM_MAX_CPL; ++c) {
    hr[c] = hi[c] = 0.0;
    if (c < T.ncpl) {
      const double* dc = T.diag_c + ((int64_t)c * T.nrows + a) * dw;
      const double dr = dc[0], di = T.dvc ? dc[1] : 0.0;
      hr[c] = dr * vr_a - di * vi_a;
      hi[c] = dr * vi_a + di * vr_a;
    }
  }
  for (int k = 0; k < len; ++k) {
    const int64_t e = base + (int64_t)k * SELL_C;
    int64_t b;
    uint32_t idx;
    sell_entry(T, e, b, idx);
    const double wr = T.dict[idx * dw], wi = T.dvc ? T.dict[idx * dw + 1] : 0.0;
    const double xr = XC ? v[2 * b] : v[b], xi = XC ? v[2 * b + 1] : 0.0;
    const double pr = wr * xr - wi * xi, pi = wr * xi + wi * xr;
    const int cc = dict_cpl[idx];
#pragma unroll
    for (int c = 0; c < ASM_MAX_CPL; ++c)
      if (c == cc) {
        hr[c] += pr;
        hi[c] += pi;
      }
  }
}

// per-block sums of (pr[c], pi[c]) into part[block][c][re, im]
__device__ __forceinline__ void dots_block_reduce(double* pr_, double* pi_, double* __restrict__ part) {
  __shared__ double sr[4][ASM_MAX_CPL], si[4][ASM_MAX_CPL];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < ASM_MAX_CPL; ++c) {
    double pr = pr_[c], pi = pi_[c];
    for (int o = 32; o > 0; o >>= 1) {
      pr += __shfl_down(pr, o, 64);
      pi += __shfl_down(pi, o, 64);
    }
    if (lane == 0) {
      sr[wid][c] = pr;
      si[wid][c] = pi;
    }
  }
  __syncthreads();
  if (threadIdx.x < ASM_MAX_CPL) {
    const int c = threadIdx.x;
    double pr = 0.0, pi = 0.0;
    for (int w = 0; w < 4; ++w) {
      pr += sr[w][c];
      pi += si[w][c];
    }
    part[((int64_t)blockIdx.x * ASM_MAX_CPL + c) * 2] = pr;
    part[((int64_t)blockIdx.x * ASM_MAX_CPL + c) * 2 + 1] = pi;
  }
}

// one pass: conj(g_a) (H_c v)_a per row, per block partial sums
// part[block][c][re, im] (one row per thread)
template <bool XC>
__global__ void __launch_bounds__(256) sell_coupling_dots_kernel(AsmTable T, const uint8_t* __restrict__ dict_cpl,
                                                                 const double* __restrict__ v,
                                                                 const double* __restrict__ g, int64_t n,
                                                                 double* __restrict__ part) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double hr[ASM_MAX_CPL], hi[ASM_MAX_CPL], pr[ASM_MAX_CPL], pi[ASM_MAX_CPL];
  double gr = 0.0, gi = 0.0;
#pragma unroll
  for (int c = 0; c < ASM_MAX_CPL; ++c) hr[c] = hi[c] = 0.0;
  if (a < n) dots_row<XC>(T, dict_cpl, v, g, a, hr, hi, gr, gi);
#pragma unroll
  for (int c = 0; c < ASM_MAX_CPL; ++c) {
    pr[c] = gr * hr[c] + gi * hi[c];
    pi[c] = gr * hi[c] - gi * hr[c];
  }
  dots_block_reduce(pr, pi, part);
}

// Same sums with a fixed grid: each block strides over row blocks, so the
// partials need gridDim.x * ASM_MAX_CPL * 2 doubles whatever n is.
template <bool XC>
__global__ void __launch_bounds__(256) sell_coupling_dots_gs_kernel(AsmTable T, const uint8_t* __restrict__ dict_cpl,
                                                                    const double* __restrict__ v,
                                                                    const double* __restrict__ g, int64_t n,
                                                                    double* __restrict__ part) {
  double hr[ASM_MAX_CPL], hi[ASM_MAX_CPL], pr[ASM_MAX_CPL], pi[ASM_MAX_CPL];
#pragma unroll
  for (int c = 0; c < ASM_MAX_CPL; ++c) pr[c] = pi[c] = 0.0;
  for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < n; a += (int64_t)gridDim.x * blockDim.x) {
    double gr, gi;
    dots_row<XC>(T, dict_cpl, v, g, a, hr, hi, gr, gi);
#pragma unroll
    for (int c = 0; c < ASM_MAX_CPL; ++c) {
      pr[c] += gr * hr[c] + gi * hi[c];
      pi[c] += gr * hi[c] - gi * hr[c];
    }
  }
  dots_block_reduce(pr, pi, part);
}

// out[c][re, im] = sum over nblk partials (fixed order), one thread per value
__global__ void dots_finish_kernel(const double* __restrict__ part, int nblk, int ncpl, double* __restrict__ out) {
  const int t = threadIdx.x;
  if (t >= 2 * ncpl) return;
  const int c = t >> 1, q = t & 1;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += part[((int64_t)b * ASM_MAX_CPL + c) * 2 + q];
  out[t] = s;
}

void free_panels(Op* op);

void free_assembly(Op* op) {
  free_panels(op);
  for (void* p : {(void*)op->asm_off, (void*)op->asm_len, (void*)op->asm_col, (void*)op->asm_val,
                  (void*)op->asm_packed, (void*)op->asm_vidx, (void*)op->asm_dict, (void*)op->asm_diag,
                  (void*)op->asm_diag_c, (void*)op->asm_dict_cpl, (void*)op->asm_dict_s, (void*)op->asm_part})
    if (p) (void)hipFree(p);
  op->asm_off = nullptr;
  op->asm_len = nullptr;
  op->asm_col = nullptr;
  op->asm_val = nullptr;
  op->asm_packed = nullptr;
  op->asm_vidx = nullptr;
  op->asm_dict = nullptr;
  op->asm_diag = nullptr;
  op->asm_diag_c = nullptr;
  op->asm_dict_cpl = nullptr;
  op->asm_dict_s = nullptr;
  op->asm_part = nullptr;
  op->asm_entries = 0;
  op->asm_vc = 0;
  op->asm_dvc = 0;
  op->asm_fmt = -1;
  op->asm_bytes = 0;
  op->asm_tagged = 0;
  op->asm_ndict = 0;
  op->asm_h_dict.clear();
  op->asm_h_dict_cpl.clear();
  op->asm_h_scale.clear();
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

// want_dict: 0 plain values, 1 dictionary when possible, 2 coupling-tagged dictionary (required)
template <typename W, bool CPLX>
static int assemble_t(Op* op, int want_dict, hipStream_t st) {
  const Basis* B = op->basis;
  const int64_t n = B->n;
  const QesPlan& P = op->plan;
  const dim3 grid((unsigned)((n + 255) / 256));
  const bool tagged = want_dict == 2;
  if (tagged && op->n_coupling > ASM_MAX_CPL) return fail("a coupling-tagged table holds at most 8 coupling ids");
  AsmOut O{};
  O.tagged = tagged ? 1 : 0;
  O.ncpl = op->n_coupling;
  O.excpl = op->d_excpl;
  // lean table passes where the group and plan allow (lean_group)
  const bool lean = lean_group(op) && P.canon != QES_CANON_BRUTE;
#define QES_ASM(CK, PASS)                                                                                           \
  if (lean)                                                                                                         \
    hipLaunchKernelGGL((asm_rows_kernel<W, CPLX, CK, PASS, true>), grid, dim3(256), op->lds_bytes, st, B->group,  \
                       B->lookup(), P, B->chi_re, B->chi_im, op->d_gen, op->n_diag, op->d_exval, op->d_lut,       \
                       B->norms, n, O);                                                                             \
  else                                                                                                              \
    hipLaunchKernelGGL((asm_rows_kernel<W, CPLX, CK, PASS>), grid, dim3(256), op->lds_bytes, st, B->group,        \
                       B->lookup(), P, B->chi_re, B->chi_im, op->d_gen, op->n_diag, op->d_exval, op->d_lut,       \
                       B->norms, n, O)
#define QES_ASM_COUNT(CK) QES_ASM(CK, PASS_COUNT)
#define QES_ASM_DICT(CK) QES_ASM(CK, PASS_DICT)
#define QES_ASM_FILL(CK) QES_ASM(CK, PASS_FILL)
  // scratch buffers of the passes, freed on every return path (the table
  // itself lives in op and is freed by the caller when a pass fails)
  struct Scratch {
    void* p = nullptr;
    hipStream_t st = nullptr;
    ~Scratch() {
      if (p) {
        (void)hipStreamSynchronize(st);   // no pass may still be using it
        (void)hipFree(p);
      }
    }
  } s_rowlen, s_dict;
  s_rowlen.st = s_dict.st = st;
  // pass A: row lengths (diagonal included)
  QES_HIP(hipMalloc(&s_rowlen.p, sizeof(int32_t) * n));
  O.rowlen = (int32_t*)s_rowlen.p;
  QES_ASM_CK_SWITCH(P.canon, QES_ASM_COUNT)
  QES_HIP(hipGetLastError());
  std::vector<int32_t> rowlen(n);
  QES_HIP(hipMemcpyAsync(rowlen.data(), O.rowlen, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
  QES_HIP(hipStreamSynchronize(st));
  (void)hipFree(s_rowlen.p);   // before the table is allocated
  s_rowlen.p = nullptr;
  O.rowlen = nullptr;

  // pass B: value dictionary of the off-diagonal elements
  int fmt = ASM_PLAIN;
  std::vector<double> dict;          // 1 or 2 doubles per value
  std::vector<uint8_t> dict_cpl;
  std::vector<uint16_t> slot_idx;
  const size_t nslot = size_t(1) << DICT_HBITS;
  void* dbuf = nullptr;
  if (want_dict) {
    QES_HIP(hipMalloc(&s_dict.p, nslot * (8 + 8 + 8 + 4 + 2) + 64));
    dbuf = s_dict.p;
    char* p = (char*)dbuf;
    O.D.key = (unsigned long long*)p;
    O.D.vr = (double*)(p + nslot * 8);
    O.D.vi = (double*)(p + nslot * 16);
    O.D.cpl = (int32_t*)(p + nslot * 24);
    O.D.idx = (const uint16_t*)(p + nslot * 28);
    O.D.bad = (int32_t*)(p + nslot * 30);
    std::vector<unsigned long long> key(nslot, DICT_EMPTY);
    std::vector<double> vr(nslot, 0.0), vi(nslot, 0.0);
    std::vector<int32_t> vc(nslot, 0);
    const uint64_t zkey = dict_key(0.0, 0.0, 0);   // padding value 0.0 is always present
    key[dict_slot(zkey)] = zkey;
    QES_HIP(hipMemcpy(O.D.key, key.data(), nslot * 8, hipMemcpyHostToDevice));
    QES_HIP(hipMemcpy(O.D.vr, vr.data(), nslot * 8, hipMemcpyHostToDevice));
    QES_HIP(hipMemcpy(O.D.vi, vi.data(), nslot * 8, hipMemcpyHostToDevice));
    QES_HIP(hipMemcpy(O.D.cpl, vc.data(), nslot * 4, hipMemcpyHostToDevice));
    QES_HIP(hipMemsetAsync(O.D.bad, 0, sizeof(int32_t), st));
    QES_ASM_CK_SWITCH(P.canon, QES_ASM_DICT)
    QES_HIP(hipGetLastError());
    QES_HIP(hipMemcpyAsync(key.data(), O.D.key, nslot * 8, hipMemcpyDeviceToHost, st));
    QES_HIP(hipMemcpyAsync(vr.data(), O.D.vr, nslot * 8, hipMemcpyDeviceToHost, st));
    QES_HIP(hipMemcpyAsync(vi.data(), O.D.vi, nslot * 8, hipMemcpyDeviceToHost, st));
    QES_HIP(hipMemcpyAsync(vc.data(), O.D.cpl, nslot * 4, hipMemcpyDeviceToHost, st));
    QES_HIP(hipStreamSynchronize(st));
    std::vector<std::tuple<int, double, double, uint32_t>> vals;
    for (uint32_t h = 0; h < nslot; ++h)
      if (key[h] != DICT_EMPTY) vals.emplace_back(vc[h], vr[h], vi[h], h);
    const int64_t nd = (int64_t)vals.size();
    if (nd < (int64_t)nslot * 3 / 4) {   // a full table means the insert gave up: no dictionary
      std::sort(vals.begin(), vals.end());
      slot_idx.assign(nslot, 0);
      for (int64_t i = 0; i < nd; ++i) {
        dict.push_back(std::get<1>(vals[i]));
        if (CPLX) dict.push_back(std::get<2>(vals[i]));
        dict_cpl.push_back((uint8_t)std::get<0>(vals[i]));
        slot_idx[std::get<3>(vals[i])] = (uint16_t)i;
      }
      // at least one index bit, so that the column field is narrower than 32
      // bits and the shift that extracts the index stays defined
      const int ib = std::max(1, bits_for(nd)), cb = bits_for(n);
      if (ib + cb <= 32) {
        fmt = ASM_P4;
        O.colbits = 32 - ib;
      } else if (nd <= 65536) {
        fmt = ASM_P6;
      }
    }
    if (fmt == ASM_PLAIN && tagged) {
      return fail("the operator's values do not fit a dictionary: no coupling-tagged table");
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
  const double per_entry = fmt == ASM_PLAIN ? 4.0 + 8.0 * cw : (fmt == ASM_P4 ? 4.0 : 6.0);
  const double diag_bytes = fmt == ASM_PLAIN ? 0.0 : 8.0 * cw * n * (1 + (tagged ? op->n_coupling : 0));
  const double bytes = (double)acc * per_entry + diag_bytes + 12.0 * ns;
  if (bytes > op->asm_max_bytes) {
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
    QES_HIP(hipMalloc(&op->asm_diag, sizeof(double) * cw * n));
    QES_HIP(hipMalloc(&op->asm_dict, sizeof(double) * dict.size()));
    QES_HIP(hipMemcpy(op->asm_dict, dict.data(), sizeof(double) * dict.size(), hipMemcpyHostToDevice));
    QES_HIP(hipMalloc(&op->asm_dict_cpl, dict_cpl.size()));
    QES_HIP(hipMemcpy(op->asm_dict_cpl, dict_cpl.data(), dict_cpl.size(), hipMemcpyHostToDevice));
    QES_HIP(hipMemcpy((void*)O.D.idx, slot_idx.data(), sizeof(uint16_t) * nslot, hipMemcpyHostToDevice));
    if (tagged) QES_HIP(hipMalloc(&op->asm_diag_c, sizeof(double) * cw * n * op->n_coupling));
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
  O.diag_c = op->asm_diag_c;
  // pass C: the table
  QES_ASM_CK_SWITCH(P.canon, QES_ASM_FILL)
  QES_HIP(hipGetLastError());
  int32_t bad = 0;
  if (dbuf) QES_HIP(hipMemcpyAsync(&bad, O.D.bad, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  QES_HIP(hipStreamSynchronize(st));
#undef QES_ASM_COUNT
#undef QES_ASM_DICT
#undef QES_ASM_FILL
#undef QES_ASM
  if (bad) return fail("dictionary fingerprint collision");   // caller retries without a dictionary
  op->asm_entries = acc;
  op->asm_slices = ns;
  op->asm_vc = (fmt == ASM_PLAIN && CPLX) ? 1 : 0;
  op->asm_dvc = (fmt != ASM_PLAIN && CPLX) ? 1 : 0;
  op->asm_fmt = fmt;
  op->asm_colbits = O.colbits;
  op->asm_ndict = (int)dict_cpl.size();
  op->asm_tagged = tagged ? 1 : 0;
  op->asm_h_dict = dict;
  op->asm_h_dict_cpl = dict_cpl;
  op->asm_bytes = (int64_t)bytes + (int64_t)dict.size() * 8;
  return 0;
}

// ---- column panels ---------------------------------------------------------
// The 6x6 J1-J2 sector's x vector (1 GB) is four times the MALL, and every
// gathered x element of the one-pass product is a random HBM line.  The
// panelled table splits each row's entries by column range: pass p of the
// product streams panel p's slices and gathers only from x[p*pw, (p+1)*pw),
// a window that stays cache resident, and adds into y.  Entries keep their
// slot order inside a panel; zero entries (the padding of the one-pass
// table) are dropped and each panel pads its own slices with (column p*pw,
// value 0).

static constexpr int PANEL_MAX = 32;

struct PanelSrc {
  AsmTable T;          // the one-pass table
  int64_t n, ns, pw;
  int npanel;
  uint32_t zidx;       // dictionary index of the value 0 (dictionary formats)
};

// entry e of the one-pass table: column, and whether it carries a nonzero value
__device__ __forceinline__ bool panel_src_entry(const PanelSrc& S, int64_t e, int64_t& b, uint32_t& w0, uint32_t& w1,
                                                double& v0, double& v1) {
  const AsmTable& T = S.T;
  if (T.fmt == ASM_P4) {
    w0 = T.packed[e];
    const uint32_t cmask = (T.colbits >= 32) ? ~0u : ((1u << T.colbits) - 1u);
    b = w0 & cmask;
    return (w0 >> T.colbits) != S.zidx;
  }
  b = T.col[e];
  if (T.fmt == ASM_P6) {
    w1 = T.vidx[e];
    return w1 != S.zidx;
  }
  const int vw = T.vc ? 2 : 1;
  v0 = T.val[e * vw];
  v1 = T.vc ? T.val[e * vw + 1] : 0.0;
  return v0 != 0.0 || v1 != 0.0;
}

// slice lengths of every panel: entries per row counted in LDS, the maximum
// over the slice's 64 rows (one wavefront) written as plen[p][s]
__global__ void __launch_bounds__(256) panel_count_kernel(PanelSrc S, int32_t* __restrict__ plen) {
  __shared__ int32_t cnt[PANEL_MAX * 256];
  const int tid = threadIdx.x;
  for (int p = 0; p < S.npanel; ++p) cnt[p * 256 + tid] = 0;
  const int64_t a = (int64_t)blockIdx.x * 256 + tid;
  const int64_t s = a / SELL_C;
  if (a < S.n) {
    const int64_t base = S.T.off[s] + a % SELL_C;
    const int len = S.T.len[s];
    for (int k = 0; k < len; ++k) {
      int64_t b;
      uint32_t w0, w1;
      double v0, v1;
      if (panel_src_entry(S, base + (int64_t)k * SELL_C, b, w0, w1, v0, v1)) ++cnt[(int)(b / S.pw) * 256 + tid];
    }
  }
  for (int p = 0; p < S.npanel; ++p) {
    int v = cnt[p * 256 + tid];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    if ((tid & (SELL_C - 1)) == 0 && s < S.ns) plen[p * S.ns + s] = v;
  }
}

struct PanelDst {
  const int64_t* poff;   // [p][ns+1]
  const int32_t* plen;   // [p][ns]
  int32_t* col;          // PLAIN / P6
  double* val;           // PLAIN
  uint32_t* packed;      // P4
  uint16_t* vidx;        // P6
};

__global__ void __launch_bounds__(256) panel_fill_kernel(PanelSrc S, PanelDst D) {
  __shared__ int32_t cnt[PANEL_MAX * 256];
  const int tid = threadIdx.x;
  for (int p = 0; p < S.npanel; ++p) cnt[p * 256 + tid] = 0;
  const int64_t a = (int64_t)blockIdx.x * 256 + tid;
  if (a >= S.n) return;   // no wavefront-wide operation follows
  const int64_t s = a / SELL_C;
  const int l = (int)(a % SELL_C);
  const AsmTable& T = S.T;
  const int vw = T.vc ? 2 : 1;
  const int64_t base = T.off[s] + l;
  const int len = T.len[s];
  for (int k = 0; k < len; ++k) {
    int64_t b;
    uint32_t w0 = 0, w1 = 0;
    double v0 = 0.0, v1 = 0.0;
    if (!panel_src_entry(S, base + (int64_t)k * SELL_C, b, w0, w1, v0, v1)) continue;
    const int p = (int)(b / S.pw);
    const int kp = cnt[p * 256 + tid]++;
    const int64_t e = D.poff[p * (S.ns + 1) + s] + (int64_t)kp * SELL_C + l;
    if (T.fmt == ASM_P4) {
      D.packed[e] = w0;
    } else {
      D.col[e] = (int32_t)b;
      if (T.fmt == ASM_P6) {
        D.vidx[e] = (uint16_t)w1;
      } else {
        D.val[e * vw] = v0;
        if (T.vc) D.val[e * vw + 1] = v1;
      }
    }
  }
  for (int p = 0; p < S.npanel; ++p) {
    const int64_t b = (int64_t)p * S.pw;   // padding gathers x[p*pw], inside the panel's window
    const int64_t o = D.poff[p * (S.ns + 1) + s];
    for (int kp = cnt[p * 256 + tid]; kp < D.plen[p * S.ns + s]; ++kp) {
      const int64_t e = o + (int64_t)kp * SELL_C + l;
      if (T.fmt == ASM_P4) {
        D.packed[e] = (uint32_t)b | (S.zidx << T.colbits);
      } else {
        D.col[e] = (int32_t)b;
        if (T.fmt == ASM_P6) {
          D.vidx[e] = (uint16_t)S.zidx;
        } else {
          D.val[e * vw] = 0.0;
          if (T.vc) D.val[e * vw + 1] = 0.0;
        }
      }
    }
  }
}

void free_panels(Op* op) {
  for (void* p : {(void*)op->asm_poff, (void*)op->asm_plen, (void*)op->asm_pcol, (void*)op->asm_pval,
                  (void*)op->asm_ppacked, (void*)op->asm_pvidx})
    if (p) (void)hipFree(p);
  op->asm_poff = nullptr;
  op->asm_plen = nullptr;
  op->asm_pcol = nullptr;
  op->asm_pval = nullptr;
  op->asm_ppacked = nullptr;
  op->asm_pvidx = nullptr;
  op->asm_npanel = 0;
  op->asm_pw = 0;
  op->asm_pentries = 0;
  op->asm_pbytes = 0;
}

// columns per panel: the option, or automatically a 128 MB window of complex
// or real x where x itself exceeds 192 MB (panels cost one more read and
// write of y per pass, which smaller vectors do not win back).  Measured on
// MI355X, 6x6 J1-J2 k=(0,0) (x = 1 GB, one-pass product 61.1 ms): windows of
// 96 / 128 / 160 / 192 MB gave 44.7 / 44.7 / 46.5 / 47.8 ms; the panel
// tables hold 15.3 / 14.4 / 13.5 / 13.1 G slots against 10.1 G (more
// slices padded)
static int64_t panel_cols(const Op* op) {
  const int64_t n = op->basis->n;
  if (op->asm_panel_cols >= 0) return op->asm_panel_cols >= n ? 0 : op->asm_panel_cols;
  const int64_t xs = (op->asm_vc || op->asm_dvc) ? 16 : 8;
  if (n * xs <= (int64_t)192 << 20) return 0;
  return ((int64_t)128 << 20) / xs;
}

static int build_panels(Op* op, hipStream_t st) {
  free_panels(op);
  const int64_t pw = panel_cols(op);
  if (pw <= 0) return 0;
  const int64_t n = op->basis->n, ns = op->asm_slices;
  const int64_t np = (n + pw - 1) / pw;
  if (np < 2) return 0;
  if (np > PANEL_MAX) return fail("asm_panel_cols gives " + std::to_string(np) + " panels (at most " +
                                  std::to_string(PANEL_MAX) + ")");
  PanelSrc S{};
  S.T = asm_table_of(op);
  S.n = n;
  S.ns = ns;
  S.pw = pw;
  S.npanel = (int)np;
  S.zidx = 0;
  if (op->asm_fmt != ASM_PLAIN) {
    const int dw = op->asm_dvc ? 2 : 1;
    bool found = false;
    for (size_t i = 0; i < op->asm_h_dict_cpl.size() && !found; ++i) {
      if (op->asm_h_dict_cpl[i] == 0 && op->asm_h_dict[i * dw] == 0.0 && (dw == 1 || op->asm_h_dict[i * dw + 1] == 0.0)) {
        S.zidx = (uint32_t)i;
        found = true;
      }
    }
    if (!found) return fail("assembled dictionary holds no zero value");
  }
  const dim3 grid((unsigned)((n + 255) / 256));
  int32_t* d_plen = nullptr;
  QES_HIP(hipMalloc(&d_plen, sizeof(int32_t) * np * ns));
  op->asm_plen = d_plen;
  hipLaunchKernelGGL(panel_count_kernel, grid, dim3(256), 0, st, S, d_plen);
  QES_HIP(hipGetLastError());
  std::vector<int32_t> plen(np * ns);
  QES_HIP(hipMemcpyAsync(plen.data(), d_plen, sizeof(int32_t) * np * ns, hipMemcpyDeviceToHost, st));
  QES_HIP(hipStreamSynchronize(st));
  std::vector<int64_t> poff(np * (ns + 1));
  int64_t acc = 0;
  for (int64_t p = 0; p < np; ++p) {
    for (int64_t s = 0; s < ns; ++s) {
      poff[p * (ns + 1) + s] = acc;
      acc += (int64_t)plen[p * ns + s] * SELL_C;
    }
    poff[p * (ns + 1) + ns] = acc;
  }
  const int cw = op->asm_vc ? 2 : 1;
  const double per_entry = op->asm_fmt == ASM_PLAIN ? 4.0 + 8.0 * cw : (op->asm_fmt == ASM_P4 ? 4.0 : 6.0);
  const double bytes = (double)acc * per_entry + 12.0 * np * ns;
  if (op->asm_bytes + bytes > op->asm_max_bytes) {   // the one-pass table serves alone
    free_panels(op);
    return 0;
  }
  QES_HIP(hipMalloc(&op->asm_poff, sizeof(int64_t) * poff.size()));
  QES_HIP(hipMemcpy(op->asm_poff, poff.data(), sizeof(int64_t) * poff.size(), hipMemcpyHostToDevice));
  const size_t ne = (size_t)std::max<int64_t>(acc, 1);
  PanelDst D{};
  D.poff = op->asm_poff;
  D.plen = op->asm_plen;
  if (op->asm_fmt == ASM_P4) {
    QES_HIP(hipMalloc(&op->asm_ppacked, sizeof(uint32_t) * ne));
    QES_HIP(hipMemsetAsync(op->asm_ppacked, 0, sizeof(uint32_t) * ne, st));
  } else {
    QES_HIP(hipMalloc(&op->asm_pcol, sizeof(int32_t) * ne));
    QES_HIP(hipMemsetAsync(op->asm_pcol, 0, sizeof(int32_t) * ne, st));
    if (op->asm_fmt == ASM_P6) {
      QES_HIP(hipMalloc(&op->asm_pvidx, sizeof(uint16_t) * ne));
      QES_HIP(hipMemsetAsync(op->asm_pvidx, 0, sizeof(uint16_t) * ne, st));
    } else {
      QES_HIP(hipMalloc(&op->asm_pval, sizeof(double) * cw * ne));
      QES_HIP(hipMemsetAsync(op->asm_pval, 0, sizeof(double) * cw * ne, st));
    }
  }
  D.col = op->asm_pcol;
  D.val = op->asm_pval;
  D.packed = op->asm_ppacked;
  D.vidx = op->asm_pvidx;
  hipLaunchKernelGGL(panel_fill_kernel, grid, dim3(256), 0, st, S, D);
  QES_HIP(hipGetLastError());
  QES_HIP(hipStreamSynchronize(st));
  op->asm_npanel = (int)np;
  op->asm_pw = pw;
  op->asm_pentries = acc;
  op->asm_pbytes = (int64_t)bytes;
  return 0;
}

// table of panel p (arrays of the panels, the dictionary and diagonals of T0)
static AsmTable panel_table(const Op* op, const AsmTable& T0, int p) {
  AsmTable T = T0;
  const int64_t ns = op->asm_slices;
  T.off = op->asm_poff + (int64_t)p * (ns + 1);
  T.len = op->asm_plen + (int64_t)p * ns;
  T.col = op->asm_pcol;
  T.val = op->asm_pval;
  T.packed = op->asm_ppacked;
  T.vidx = op->asm_pvidx;
  T.nodiag = p > 0 ? 1 : 0;
  return T;
}

int op_assemble(Op* op, int enable, int want_dict, hipStream_t st) {
  free_assembly(op);
  if (!enable || op->basis->n == 0) return 0;
  if (op->basis->n >= (int64_t(1) << 31)) return fail("assembled operator needs fewer than 2^31 rows");
  const bool vc = op->is_complex || op->chi_complex;
  const bool w32 = op->basis->group.ns <= 32;
  auto run = [&](int wd) {
    if (w32) return vc ? assemble_t<uint32_t, true>(op, wd, st) : assemble_t<uint32_t, false>(op, wd, st);
    return vc ? assemble_t<uint64_t, true>(op, wd, st) : assemble_t<uint64_t, false>(op, wd, st);
  };
  int rc = run(want_dict);
  if (rc && want_dict == 1) {   // dictionary path failed (collision): plain values
    free_assembly(op);
    rc = run(0);
  }
  if (!rc) rc = build_panels(op, st);
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
  const bool cv = T.fmt == ASM_PLAIN ? T.vc : T.dvc;
  const int vw = cv ? 2 : 1;
  for (int k = 0; k < len; ++k) {
    const int64_t e = base + (int64_t)k * SELL_C;
    rows[e] = (int32_t)a;
    if (T.fmt == ASM_PLAIN) {
      cols[e] = T.col[e];
      vals[e * vw] = T.val[e * vw];
      if (cv) vals[e * vw + 1] = T.val[e * vw + 1];
    } else {
      int64_t b;
      uint32_t idx;
      sell_entry(T, e, b, idx);
      cols[e] = (int32_t)b;
      vals[e * vw] = T.dict[idx * vw];
      if (cv) vals[e * vw + 1] = T.dict[idx * vw + 1];
    }
  }
  if (T.fmt != ASM_PLAIN) {
    rows[entries + a] = (int32_t)a;
    cols[entries + a] = (int32_t)a;
    vals[(entries + a) * vw] = T.diag[a * vw];
    if (cv) vals[(entries + a) * vw + 1] = T.diag[a * vw + 1];
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

// blocks of the single-vector product: one per 256 rows, or (sell_grid > 0)
// sell_grid blocks per CU striding over the rows
static dim3 sell_grid_of(int sell_grid, int64_t n) {
  int64_t g = (n + 255) / 256;
  if (sell_grid > 0) {
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    g = std::min<int64_t>(g, (int64_t)ncu * sell_grid);
  }
  return dim3((unsigned)std::max<int64_t>(g, 1));
}

template <bool XC>
static void launch_sell_x(const AsmTable& T, const double* x, double* y, int64_t n, int accumulate, hipStream_t st,
                          int sell_grid, int sell_lds) {
  const dim3 grid = sell_grid_of(sell_grid, n);
  const bool cv = T.fmt == ASM_PLAIN ? T.vc != 0 : T.dvc != 0;
#define QES_SELL(F, C) \
  hipLaunchKernelGGL((sell_matvec_kernel<XC, F, C>), grid, dim3(256), sell_lds, st, T, x, y, n, accumulate)
  if (T.fmt == ASM_P4) {
    if (cv) QES_SELL(ASM_P4, true); else QES_SELL(ASM_P4, false);
  } else if (T.fmt == ASM_P6) {
    if (cv) QES_SELL(ASM_P6, true); else QES_SELL(ASM_P6, false);
  } else {
    if (cv) QES_SELL(ASM_PLAIN, true); else QES_SELL(ASM_PLAIN, false);
  }
#undef QES_SELL
}

static int launch_sell(const AsmTable& T, const double* x, double* y, int64_t n, bool xc, int accumulate,
                       hipStream_t st, int sell_grid = 0, int sell_lds = 0) {
  if (xc) launch_sell_x<true>(T, x, y, n, accumulate, st, sell_grid, sell_lds);
  else launch_sell_x<false>(T, x, y, n, accumulate, st, sell_grid, sell_lds);
  QES_HIP(hipGetLastError());
  return 0;
}

// single-vector product over the column panels when the operator has them
// (pass 0 adds the diagonal and honours accumulate, later passes add), else
// one pass over the table
static int launch_sell_panels(const Op* op, const AsmTable& T0, const double* x, double* y, int64_t n, bool xc,
                              int accumulate, hipStream_t st) {
  if (op->asm_npanel <= 0) return launch_sell(T0, x, y, n, xc, accumulate, st, op->sell_grid, op->sell_lds);
  for (int p = 0; p < op->asm_npanel; ++p)
    if (launch_sell(panel_table(op, T0, p), x, y, n, xc, p > 0 ? 1 : accumulate, st, op->sell_grid, op->sell_lds))
      return -1;
  return 0;
}

int assembled_matvec(const Op* op, const double* x, double* y, int dtype, int accumulate, hipStream_t st) {
  const int64_t n = op->basis->n;
  if (n == 0) return 0;
  const bool xc = dtype == 1;
  if ((op->asm_vc || op->asm_dvc) && !xc) return fail("complex matrix elements need complex128 vectors");
  return launch_sell_panels(op, asm_table_of(op), x, y, n, xc, accumulate, st);
}

template <bool XC, int FMT, bool CV>
static void launch_multi_fmt(const AsmTable& T, const double* x, double* y, int64_t n, int nb, int accumulate,
                             hipStream_t st) {
  const dim3 grid((unsigned)((n + 255) / 256));
  for (int c0 = 0; c0 < nb; c0 += 4) {
    const int rc = std::min(4, nb - c0);
#define QES_MULTI(R) \
  hipLaunchKernelGGL((sell_matvec_multi_kernel<FMT, XC, CV, R>), grid, dim3(256), 0, st, T, x, y, n, nb, c0, accumulate)
    if (rc == 4) QES_MULTI(4);
    else if (rc == 3) QES_MULTI(3);
    else if (rc == 2) QES_MULTI(2);
    else QES_MULTI(1);
#undef QES_MULTI
  }
}

template <bool XC>
static void launch_multi_x(const AsmTable& T, const double* x, double* y, int64_t n, int nb, int accumulate,
                           hipStream_t st) {
  const bool cv = T.fmt == ASM_PLAIN ? T.vc != 0 : T.dvc != 0;
  if (T.fmt == ASM_P4) {
    if (cv) launch_multi_fmt<XC, ASM_P4, true>(T, x, y, n, nb, accumulate, st);
    else launch_multi_fmt<XC, ASM_P4, false>(T, x, y, n, nb, accumulate, st);
  } else if (T.fmt == ASM_P6) {
    if (cv) launch_multi_fmt<XC, ASM_P6, true>(T, x, y, n, nb, accumulate, st);
    else launch_multi_fmt<XC, ASM_P6, false>(T, x, y, n, nb, accumulate, st);
  } else {
    if (cv) launch_multi_fmt<XC, ASM_PLAIN, true>(T, x, y, n, nb, accumulate, st);
    else launch_multi_fmt<XC, ASM_PLAIN, false>(T, x, y, n, nb, accumulate, st);
  }
}

int assembled_matvec_multi(const Op* op, const double* x, double* y, int64_t nbatch, int dtype, int accumulate,
                           hipStream_t st) {
  const int64_t n = op->basis->n;
  if (n == 0) return 0;
  if (nbatch > (int64_t(1) << 30)) return fail("nbatch too large");
  const bool xc = dtype == 1;
  if ((op->asm_vc || op->asm_dvc) && !xc) return fail("complex matrix elements need complex128 vectors");
  const AsmTable T = asm_table_of(op);
  if (xc) launch_multi_x<true>(T, x, y, n, (int)nbatch, accumulate, st);
  else launch_multi_x<false>(T, x, y, n, (int)nbatch, accumulate, st);
  QES_HIP(hipGetLastError());
  return 0;
}

int assembled_matvec_scaled(Op* op, const double* scale, int nscale, const double* x, double* y, int dtype,
                            hipStream_t st) {
  if (op->asm_fmt < 0 || !op->asm_tagged) return fail("operator has no coupling-tagged assembled table");
  if (nscale != op->n_coupling) return fail("one scale per coupling id needed");
  const int64_t n = op->basis->n;
  if (n == 0) return 0;
  const bool xc = dtype == 1;
  if (op->asm_dvc && !xc) return fail("complex matrix elements need complex128 vectors");
  // scaled dictionary value * scale[coupling of the value], uploaded when the
  // scales differ from the last call's (a recurrence at fixed J uploads once)
  const std::vector<double> sc(scale, scale + nscale);
  if (!op->asm_dict_s || sc != op->asm_h_scale) {
    const int dw = op->asm_dvc ? 2 : 1;
    std::vector<double> ds(op->asm_h_dict.size());
    for (size_t i = 0; i < op->asm_h_dict_cpl.size(); ++i)
      for (int w = 0; w < dw; ++w) ds[i * dw + w] = op->asm_h_dict[i * dw + w] * scale[op->asm_h_dict_cpl[i]];
    if (!op->asm_dict_s) QES_HIP(hipMalloc(&op->asm_dict_s, sizeof(double) * std::max<size_t>(ds.size(), 1)));
    QES_HIP(hipMemcpyAsync(op->asm_dict_s, ds.data(), sizeof(double) * ds.size(), hipMemcpyHostToDevice, st));
    // the host copy of the scaled values must outlive the asynchronous upload
    QES_HIP(hipStreamSynchronize(st));
    op->asm_h_scale = sc;
  }
  AsmTable T = asm_table_of(op);
  T.dict = op->asm_dict_s;
  T.ncpl = op->n_coupling;
  for (int c = 0; c < op->n_coupling; ++c) T.dscale[c] = scale[c];
  return launch_sell_panels(op, T, x, y, n, xc, 0, st);
}

int assembled_coupling_dots(Op* op, const double* v, const double* g, int dtype, double* out_host, hipStream_t st) {
  if (op->asm_fmt < 0 || !op->asm_tagged) return fail("operator has no coupling-tagged assembled table");
  const int64_t n = op->basis->n;
  const int nc = op->n_coupling;
  for (int c = 0; c < 2 * nc; ++c) out_host[c] = 0.0;
  if (n == 0) return 0;
  const bool xc = dtype == 1;
  if (op->asm_dvc && !xc) return fail("complex matrix elements need complex128 vectors");
  const int64_t nb = (n + 255) / 256;
  if (!op->asm_part) QES_HIP(hipMalloc(&op->asm_part, sizeof(double) * nb * ASM_MAX_CPL * 2));
  AsmTable T = asm_table_of(op);
  T.ncpl = nc;
  if (xc)
    hipLaunchKernelGGL((sell_coupling_dots_kernel<true>), dim3((unsigned)nb), dim3(256), 0, st, T, op->asm_dict_cpl, v,
                       g, n, op->asm_part);
  else
    hipLaunchKernelGGL((sell_coupling_dots_kernel<false>), dim3((unsigned)nb), dim3(256), 0, st, T, op->asm_dict_cpl,
                       v, g, n, op->asm_part);
  QES_HIP(hipGetLastError());
  std::vector<double> part((size_t)nb * ASM_MAX_CPL * 2);
  QES_HIP(hipMemcpyAsync(part.data(), op->asm_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, st));
  QES_HIP(hipStreamSynchronize(st));
  for (int64_t b = 0; b < nb; ++b)
    for (int c = 0; c < nc; ++c) {
      out_host[2 * c] += part[((size_t)b * ASM_MAX_CPL + c) * 2];
      out_host[2 * c + 1] += part[((size_t)b * ASM_MAX_CPL + c) * 2 + 1];
    }
  return 0;
}

// ---------------------------------------------------------------------------
// Allocation-free coupling dots.  Partial sums live in a file-scope device
// array, `scratch`: each GPU holds its own copy (found with
// hipGetSymbolAddress once per device, with that device current) until the
// process ends.  Layout: DOTS_NOALLOC_BLOCKS * ASM_MAX_CPL * 2 partials, then
// 2 * ASM_MAX_CPL results.  Every block writes all of its partial slots and
// the finishing kernel writes every result slot it returns, so nothing needs
// clearing between calls; calls that overlap on one GPU share the array and
// race.
// ---------------------------------------------------------------------------
#define DOTS_NOALLOC_BLOCKS 2048
#define DOTS_SCRATCH_DOUBLES (DOTS_NOALLOC_BLOCKS * ASM_MAX_CPL * 2 + 2 * ASM_MAX_CPL)
static __device__ __align__(256) double scratch[DOTS_SCRATCH_DOUBLES];

#define QES_MAX_DEVICES 64
static double* g_scratch_dev[QES_MAX_DEVICES];
static std::mutex g_scratch_mu;

// device address of this GPU's scratch array (hipGetSymbolAddress on the
// first call with that device current)
static int scratch_ptr(double** out) {
  int dev = 0;
  QES_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= QES_MAX_DEVICES) return fail("device index out of range for the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    double* p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return check_hip(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return 0;
}

int assembled_coupling_dots_noalloc(const Op* op, const double* v, const double* g, int dtype, double* out_host,
                                    hipStream_t st) {
  if (op->asm_fmt < 0 || !op->asm_tagged) return fail("operator has no coupling-tagged assembled table");
  const int64_t n = op->basis->n;
  const int nc = op->n_coupling;
  if (nc < 1 || nc > ASM_MAX_CPL)
    return fail("the allocation-free coupling dots hold at most " + std::to_string(ASM_MAX_CPL) +
                " coupling ids (operator has " + std::to_string(nc) + ")");
  for (int c = 0; c < 2 * nc; ++c) out_host[c] = 0.0;
  if (n == 0) return 0;
  const bool xc = dtype == 1;
  if (op->asm_dvc && !xc) return fail("complex matrix elements need complex128 vectors");
  double* sp = nullptr;
  if (scratch_ptr(&sp)) return -1;
  double* part = sp;
  double* res = sp + DOTS_NOALLOC_BLOCKS * ASM_MAX_CPL * 2;
  const int nblk = (int)std::min<int64_t>(DOTS_NOALLOC_BLOCKS, (n + 255) / 256);
  AsmTable T = asm_table_of(op);
  T.ncpl = nc;
  if (xc)
    hipLaunchKernelGGL((sell_coupling_dots_gs_kernel<true>), dim3((unsigned)nblk), dim3(256), 0, st, T,
                       op->asm_dict_cpl, v, g, n, part);
  else
    hipLaunchKernelGGL((sell_coupling_dots_gs_kernel<false>), dim3((unsigned)nblk), dim3(256), 0, st, T,
                       op->asm_dict_cpl, v, g, n, part);
  QES_HIP(hipGetLastError());
  hipLaunchKernelGGL(dots_finish_kernel, dim3(1), dim3(64), 0, st, part, nblk, nc, res);
  QES_HIP(hipGetLastError());
  QES_HIP(hipMemcpyAsync(out_host, res, sizeof(double) * 2 * nc, hipMemcpyDeviceToHost, st));
  QES_HIP(hipStreamSynchronize(st));
  return 0;
}

}  // namespace qes
