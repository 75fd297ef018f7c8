// Batched Lanczos over many independent sectors / disorder realisations.
//
// The reference sweeps sectors one HilbertSpace at a time
// (HilbertSpace.iter_momentum_sectors / iter_symmetry_sectors,
// Algebra/hilbert.py:412-652), each followed by its own diagonalize().
// Here all problems of a batch advance together: one persistent matvec
// launch walks a queue of row tiles drawn from every problem, and each
// vector step (dot, three-term update, norm, normalise) is one launch over
// chunks of every problem.  One Lanczos iteration of the whole batch is a
// fixed launch sequence, captured once into a hipGraph and replayed.
//
// Problems that share a basis and every off-diagonal term and differ only in
// their diagonal (disorder realisations of a random field) form a "block":
// R right-hand sides stored row-major (element (a, r) at a*R + r) that are
// advanced by one wavefront per row with lanes over the realisations, so a
// single canonicalisation + lookup serves all R columns and the gathers are
// contiguous.  Every column is its own Lanczos ("virtual problem") with its
// own alpha/beta; converged virtual problems are masked out.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/qes_mi.h"
#include "qes_internal.h"
#include "qes_lanczos.h"
#include "qes_rows.h"
#include "qes_sell.h"

namespace qes {

Op* qes_op_internal(qes_op_t h);   // qes_capi.hip

static constexpr int BRB = 256;        // threads per block
static constexpr int CHUNK = 4096;     // vector elements per chunk
static constexpr int MAX_RHS = 256;    // columns of a block problem (4 per lane)

struct QesProb {
  QesGroup G;
  QesLookup L;
  QesPlan P;
  const double* chi_re;
  const double* chi_im;
  const QesTermGroup* groups;
  int32_t n_diag;
  int32_t nrhs;         // 1: ordinary problem; R > 1: block of R realisations
  const double* exval;
  const uint32_t* lut;
  const double* norms;
  const double* D;      // block: per-realisation diagonal, (n, R) row-major
  int64_t n;            // rows
  int64_t off;          // element offset into the concatenated vectors
  int32_t vp0;          // first virtual problem
  int32_t blk;          // 1: block problem (diagonal from D)
  // block problems: off-diagonal neighbour table (pull form, value already
  // divided by N_a), built once at batch creation; null = matrix-free rows
  const int64_t* rowptr;
  const int32_t* col;
  const double* val;    // 1 (real) or 2 (complex) doubles per entry
  // ordinary problems: the operator's assembled table (fmt < 0: none).
  // Coupling-scaled block problems: its coupling-tagged table, streamed once
  // for all R columns; column r is H_r = sum_c cscale[r][c] H_c.
  AsmTable T;
  const double* cscale;   // [R][ASM_MAX_CPL] (null: not a coupling-scaled block)
  const double* dcols;    // [R][ndict] dictionary values scaled for each column (1 or 2 doubles each)
  int32_t ndict;
  int32_t pad2;
};

struct Tile {
  int32_t prob;
  int32_t rows;         // rows of this tile
  int64_t row0;
};

struct Chunk {
  int32_t prob;         // real problem
  int32_t rows;         // ordinary: elements; block: rows (x R columns)
  int64_t start;        // element offset
  int32_t nrhs;         // 1 or R
  int32_t vp0;          // virtual problem of column 0
  int32_t slot;         // partial-sum slot of column 0
  int32_t sstride;      // slot stride between columns (block)
};

// ---------------------------------------------------------------------------
// batched matvec: persistent grid over the tile queue
// ---------------------------------------------------------------------------
template <typename W, bool CPLX, int CK>
__device__ __forceinline__ void tile_rows(const RowCtx& C, const QesProb& P, int64_t a, const double* __restrict__ x,
                                          double* __restrict__ y) {
  const int cw = CPLX ? 2 : 1;
  const W r = load_rep<W>(C.L, a);
  const double inv_na = 1.0 / P.norms[a];
  const double* xp0 = x + P.off * cw;
  double* yp0 = y + P.off * cw;
  double dr = 0.0, di = 0.0;
  for (int t = 0; t < C.n_diag; ++t) {
    const QesTermGroup& g = C.groups[t];
    int k = term_bits<W>(g, r);
    dr += sel4(g.re, k);
    if (CPLX) di += sel4(g.im, k);
  }
  double ar, ai = 0.0;
  if (CPLX) {
    const double xr = xp0[a * 2], xi = xp0[a * 2 + 1];
    ar = dr * xr - di * xi;
    ai = dr * xi + di * xr;
  } else {
    ar = dr * xp0[a];
  }
  visit_row<W, CPLX, CK>(C, r, [&](int64_t b, double fr, double fi, int) {
    fr *= inv_na;
    fi *= inv_na;
    if (CPLX) {
      const double xr = xp0[b * 2], xi = xp0[b * 2 + 1];
      ar += fr * xr - fi * xi;
      ai += fr * xi + fi * xr;
    } else {
      ar += fr * xp0[b];
    }
  });
  if (CPLX) {
    yp0[a * 2] = ar;
    yp0[a * 2 + 1] = ai;
  } else {
    yp0[a] = ar;
  }
}

// Rp = lanes per row: R rounded up to a power of two, at most 64.  A
// wavefront holds 64/Rp rows; lane c0 of a row owns columns c0, c0+Rp, ...
__host__ __device__ __forceinline__ int block_lanes(int R) {
  int rp = 1;
  while (rp < R && rp < 64) rp <<= 1;
  return rp;
}

__device__ __forceinline__ unsigned block_colmask(const QesProb& P, int c0, int Rp, const int* __restrict__ active) {
  unsigned m = 0;
#pragma unroll
  for (int k = 0; k < MAX_RHS / 64; ++k) {
    const int c = c0 + Rp * k;
    if (c < P.nrhs && (!active || active[P.vp0 + c])) m |= 1u << k;
  }
  return m;
}

template <typename W, bool CPLX, int CK>
__device__ __forceinline__ void block_row(const RowCtx& C, const QesProb& P, int64_t a, int c0, int Rp,
                                          unsigned cm, const double* __restrict__ x, double* __restrict__ y) {
  const int cw = CPLX ? 2 : 1;
  const int R = P.nrhs;
  const W r = load_rep<W>(C.L, a);
  const double inv_na = 1.0 / P.norms[a];
  const double* xp0 = x + P.off * cw;
  double* yp0 = y + P.off * cw;
  double ar[MAX_RHS / 64], ai[MAX_RHS / 64];
#pragma unroll
  for (int k = 0; k < MAX_RHS / 64; ++k) {
    const int c = c0 + Rp * k;
    ar[k] = 0.0;
    ai[k] = 0.0;
    if ((cm >> k) & 1) {
      const int64_t e = a * R + c;
      const double d = P.D[e];
      ar[k] = d * xp0[e * cw];
      if (CPLX) ai[k] = d * xp0[e * cw + 1];
    }
  }
  visit_row<W, CPLX, CK>(C, r, [&](int64_t b, double fr, double fi, int) {
    fr *= inv_na;
    fi *= inv_na;
#pragma unroll
    for (int k = 0; k < MAX_RHS / 64; ++k) {
      const int c = c0 + Rp * k;
      if ((cm >> k) & 1) {
        const int64_t e = b * R + c;
        if (CPLX) {
          const double xr = xp0[e * 2], xi = xp0[e * 2 + 1];
          ar[k] += fr * xr - fi * xi;
          ai[k] += fr * xi + fi * xr;
        } else {
          ar[k] += fr * xp0[e];
        }
      }
    }
  });
#pragma unroll
  for (int k = 0; k < MAX_RHS / 64; ++k) {
    const int c = c0 + Rp * k;
    if ((cm >> k) & 1) {
      const int64_t e = a * R + c;
      yp0[e * cw] = ar[k];
      if (CPLX) yp0[e * cw + 1] = ai[k];
    }
  }
}

// Same row from the neighbour table: the lanes of a row read the same entry
// (one broadcast request), entries are consumed four at a time so that the
// table loads and the contiguous column gathers of four neighbours overlap.
template <bool CPLX>
__device__ __forceinline__ void block_row_table(const QesProb& P, int64_t a, int c0, int Rp, unsigned cm,
                                                const double* __restrict__ x, double* __restrict__ y) {
  const int cw = CPLX ? 2 : 1;
  const int R = P.nrhs;
  const double* xp0 = x + P.off * cw;
  double* yp0 = y + P.off * cw;
  double ar[MAX_RHS / 64], ai[MAX_RHS / 64];
#pragma unroll
  for (int k = 0; k < MAX_RHS / 64; ++k) {
    const int c = c0 + Rp * k;
    ar[k] = 0.0;
    ai[k] = 0.0;
    if ((cm >> k) & 1) {
      const int64_t e = a * R + c;
      const double d = P.D[e];
      ar[k] = d * xp0[e * cw];
      if (CPLX) ai[k] = d * xp0[e * cw + 1];
    }
  }
  const int64_t e0 = P.rowptr[a], e1 = P.rowptr[a + 1];
  auto acc = [&](int64_t b, double fr, double fi) {
#pragma unroll
    for (int k = 0; k < MAX_RHS / 64; ++k) {
      const int c = c0 + Rp * k;
      if ((cm >> k) & 1) {
        const int64_t e = b * R + c;
        if (CPLX) {
          const double xr = xp0[e * 2], xi = xp0[e * 2 + 1];
          ar[k] += fr * xr - fi * xi;
          ai[k] += fr * xi + fi * xr;
        } else {
          ar[k] += fr * xp0[e];
        }
      }
    }
  };
  int64_t e = e0;
  for (; e + 4 <= e1; e += 4) {
    int64_t b[4];
    double fr[4], fi[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      b[u] = P.col[e + u];
      fr[u] = P.val[(e + u) * cw];
      fi[u] = CPLX ? P.val[(e + u) * cw + 1] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) acc(b[u], fr[u], fi[u]);
  }
  for (; e < e1; ++e) acc(P.col[e], P.val[e * cw], CPLX ? P.val[e * cw + 1] : 0.0);
#pragma unroll
  for (int k = 0; k < MAX_RHS / 64; ++k) {
    const int c = c0 + Rp * k;
    if ((cm >> k) & 1) {
      const int64_t e = a * R + c;
      yp0[e * cw] = ar[k];
      if (CPLX) yp0[e * cw + 1] = ai[k];
    }
  }
}

// Row of a coupling-scaled block problem from the operator's coupling-tagged
// table, one lane per row (the wavefront reads 64 consecutive slots of a
// slice, as the single-operator product does) and the row's RC <= 4 columns
// in registers: every slot is read once for all columns, its neighbour's RC
// values are contiguous, and column r takes its value from its own scaled
// dictionary; the diagonal of column r is sum_c cscale[r][c] diag_c[a].
// Slots are taken four at a time as in sell_row_t.
template <int FMT, bool CPLX, bool CV, int RC>
__device__ __forceinline__ void block_row_sell(const QesProb& P, int64_t a, unsigned cm, const double* __restrict__ x,
                                               double* __restrict__ y) {
  constexpr int VW = CV ? 2 : 1;
  const AsmTable& T = P.T;
  const int cw = CPLX ? 2 : 1;
  const double* xp0 = x + P.off * cw;
  double* yp0 = y + P.off * cw;
  // diagonal of column r: sum over c of cscale[r][c] diag_c[a], summed in c
  // order (coupling outer, column inner keeps RC sums live instead of the
  // eight diagonals)
  double dr[RC], di[RC];
#pragma unroll
  for (int r = 0; r < RC; ++r) dr[r] = di[r] = 0.0;
#pragma unroll
  for (int c = 0; c < ASM_MAX_CPL; ++c) {
    if (c < T.ncpl) {
      const double* dc = T.diag_c + ((int64_t)c * T.nrows + a) * VW;
      const double dcr = stream_load(dc), dci = CV ? stream_load(dc + 1) : 0.0;
#pragma unroll
      for (int r = 0; r < RC; ++r) {
        const double s = P.cscale[r * ASM_MAX_CPL + c];
        dr[r] += s * dcr;
        if (CV) di[r] += s * dci;
      }
    }
  }
  double ar[RC], ai[RC];
#pragma unroll
  for (int r = 0; r < RC; ++r) {
    const int64_t e = a * RC + r;
    if (CPLX) {
      const double xr = xp0[e * 2], xi = xp0[e * 2 + 1];
      ar[r] = dr[r] * xr - di[r] * xi;
      ai[r] = dr[r] * xi + di[r] * xr;
    } else {
      ar[r] = dr[r] * xp0[e];
      ai[r] = 0.0;
    }
  }
  const int64_t s = a / SELL_C;
  const int l = (int)(a % SELL_C);
  const int64_t base = T.off[s] + l;
  const int len = T.len[s];
  const uint32_t cmask = (T.colbits >= 32) ? ~0u : ((1u << T.colbits) - 1u);
  auto entry = [&](int64_t e, int64_t& b, uint32_t& idx) {
    if (FMT == ASM_P4) {
      const uint32_t p = stream_load(T.packed + e);
      b = p & cmask;
      idx = p >> T.colbits;
    } else {
      b = stream_load(T.col + e);
      idx = stream_load(T.vidx + e);
    }
  };
  auto acc = [&](int64_t b, uint32_t idx) {
#pragma unroll
    for (int r = 0; r < RC; ++r) {
      const double* dv = P.dcols + ((int64_t)r * P.ndict + idx) * VW;
      const double vr = dv[0], vi = CV ? dv[1] : 0.0;
      const int64_t e = b * RC + r;
      if (CPLX) {
        const double xr = xp0[e * 2], xi = xp0[e * 2 + 1];
        ar[r] += vr * xr - vi * xi;
        ai[r] += vr * xi + vi * xr;
      } else {
        ar[r] += vr * xp0[e];
      }
    }
  };
  int k = 0;
  for (; k + 4 <= len; k += 4) {
    int64_t b[4];
    uint32_t idx[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) entry(base + (int64_t)(k + u) * SELL_C, b[u], idx[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) acc(b[u], idx[u]);
  }
  for (; k < len; ++k) {
    int64_t b;
    uint32_t idx;
    entry(base + (int64_t)k * SELL_C, b, idx);
    acc(b, idx);
  }
#pragma unroll
  for (int r = 0; r < RC; ++r) {
    if ((cm >> r) & 1) {
      const int64_t e = a * RC + r;
      yp0[e * cw] = ar[r];
      if (CPLX) yp0[e * cw + 1] = ai[r];
    }
  }
}

template <int FMT, bool CPLX, bool CV>
__device__ __forceinline__ void block_row_sell_r(const QesProb& P, int64_t a, unsigned cm, const double* __restrict__ x,
                                                 double* __restrict__ y) {
  switch (P.nrhs) {
    case 1: block_row_sell<FMT, CPLX, CV, 1>(P, a, cm, x, y); break;
    case 2: block_row_sell<FMT, CPLX, CV, 2>(P, a, cm, x, y); break;
    case 3: block_row_sell<FMT, CPLX, CV, 3>(P, a, cm, x, y); break;
    default: block_row_sell<FMT, CPLX, CV, 4>(P, a, cm, x, y); break;
  }
}

// next tile of a block: with a queue counter, taken by one lane with an
// atomic add and shared through LDS (blocks that finish early take more
// tiles); without, the static grid stride.  Every block leaves the loop once
// the counter passes ntiles.
__device__ __forceinline__ int next_tile(int* __restrict__ queue, int t, int* s_t) {
  if (!queue) return t < 0 ? (int)blockIdx.x : t + (int)gridDim.x;
  __syncthreads();   // every lane has read the previous tile index
  if (threadIdx.x == 0) *s_t = atomicAdd(queue, 1);
  __syncthreads();
  return *s_t;
}

template <typename W, bool CPLX>
__global__ void __launch_bounds__(BRB) batch_matvec_kernel(const QesProb* __restrict__ probs,
                                                           const Tile* __restrict__ tiles, int ntiles,
                                                           const int* __restrict__ prob_active,
                                                           const int* __restrict__ vp_active,
                                                           const double* __restrict__ x, double* __restrict__ y,
                                                           int* __restrict__ queue) {
  extern __shared__ char smem[];
  __shared__ int s_t;
  for (int t = next_tile(queue, -1, &s_t); t < ntiles; t = next_tile(queue, t, &s_t)) {
    const Tile T = tiles[t];
    if (prob_active && !prob_active[T.prob]) continue;   // uniform per block
    const QesProb& P = probs[T.prob];
    if (!P.blk && P.T.fmt >= 0) {   // assembled ordinary problem: stream its table
      const int64_t a = T.row0 + threadIdx.x;
      if (a < P.n) {
        const int cw = CPLX ? 2 : 1;
        double ar, ai;
        const double* xp = x + P.off * cw;
        const bool cv = P.T.fmt == ASM_PLAIN ? P.T.vc != 0 : P.T.dvc != 0;
        if (P.T.fmt == ASM_P4) {
          if (cv) sell_row_t<ASM_P4, CPLX, true>(P.T, a, xp, ar, ai);
          else sell_row_t<ASM_P4, CPLX, false>(P.T, a, xp, ar, ai);
        } else if (P.T.fmt == ASM_P6) {
          if (cv) sell_row_t<ASM_P6, CPLX, true>(P.T, a, xp, ar, ai);
          else sell_row_t<ASM_P6, CPLX, false>(P.T, a, xp, ar, ai);
        } else {
          if (cv) sell_row_t<ASM_PLAIN, CPLX, true>(P.T, a, xp, ar, ai);
          else sell_row_t<ASM_PLAIN, CPLX, false>(P.T, a, xp, ar, ai);
        }
        double* yp = y + (P.off + a) * cw;
        yp[0] = ar;
        if (CPLX) yp[1] = ai;
      }
      continue;
    }
    if (P.cscale) {   // coupling-scaled block: stream the tagged table once for all columns, lane per row
      unsigned cm = 0;
      for (int r = 0; r < P.nrhs; ++r)
        if (!vp_active || vp_active[P.vp0 + r]) cm |= 1u << r;
      const bool cv = P.T.dvc != 0;
      const int64_t a = T.row0 + threadIdx.x;
      if (a < P.n && a < T.row0 + T.rows) {
        if (P.T.fmt == ASM_P4) {
          if (cv) block_row_sell_r<ASM_P4, CPLX, true>(P, a, cm, x, y);
          else block_row_sell_r<ASM_P4, CPLX, false>(P, a, cm, x, y);
        } else {
          if (cv) block_row_sell_r<ASM_P6, CPLX, true>(P, a, cm, x, y);
          else block_row_sell_r<ASM_P6, CPLX, false>(P, a, cm, x, y);
        }
      }
      continue;
    }
    if (P.rowptr) {   // tabulated block problem: no operator context needed
      const int Rp = block_lanes(P.nrhs);
      const int c0 = threadIdx.x % Rp;
      const unsigned cm = block_colmask(P, c0, Rp, vp_active);
      for (int64_t a = T.row0 + threadIdx.x / Rp; a < (T.row0 + T.rows < P.n ? T.row0 + T.rows : P.n); a += BRB / Rp)
        block_row_table<CPLX>(P, a, c0, Rp, cm, x, y);
      continue;
    }
    __syncthreads();                                     // previous tile done with LDS
    const RowCtx C = load_ctx<CPLX>(P.G, P.L, P.P, P.chi_re, P.chi_im, P.groups, P.n_diag, P.exval, P.lut, smem);
    if (!P.blk) {
      const int64_t a = T.row0 + threadIdx.x;
      if (a < P.n) {
        switch (P.P.canon) {
          case QES_CANON_CYCLIC: tile_rows<W, CPLX, QES_CANON_CYCLIC>(C, P, a, x, y); break;
          case QES_CANON_ROWS: tile_rows<W, CPLX, QES_CANON_ROWS>(C, P, a, x, y); break;
          default: tile_rows<W, CPLX, QES_CANON_BRUTE>(C, P, a, x, y); break;
        }
      }
    } else {
      const int Rp = block_lanes(P.nrhs);
      const int c0 = threadIdx.x % Rp;
      const unsigned cm = block_colmask(P, c0, Rp, vp_active);
      for (int64_t a = T.row0 + threadIdx.x / Rp; a < (T.row0 + T.rows < P.n ? T.row0 + T.rows : P.n); a += BRB / Rp) {
        switch (P.P.canon) {
          case QES_CANON_CYCLIC: block_row<W, CPLX, QES_CANON_CYCLIC>(C, P, a, c0, Rp, cm, x, y); break;
          case QES_CANON_ROWS: block_row<W, CPLX, QES_CANON_ROWS>(C, P, a, c0, Rp, cm, x, y); break;
          default: block_row<W, CPLX, QES_CANON_BRUTE>(C, P, a, c0, Rp, cm, x, y); break;
        }
      }
    }
  }
}

// Batches whose problems are all coupling-scaled blocks (coupling sweeps):
// the same tile loop as batch_matvec_kernel with only the table-streaming
// branch, so the kernel is not sized by the matrix-free row visitor's
// registers and LDS (the combined kernel held 103-133 VGPRs and spilled
// SGPRs; this one keeps more waves resident for the gathers)
template <bool CPLX, int FMT, bool CV, int RC>
__global__ void __launch_bounds__(BRB) batch_scaled_kernel(const QesProb* __restrict__ probs,
                                                           const Tile* __restrict__ tiles, int ntiles,
                                                           const int* __restrict__ prob_active,
                                                           const int* __restrict__ vp_active,
                                                           const double* __restrict__ x, double* __restrict__ y,
                                                           int* __restrict__ queue) {
  __shared__ int s_t;
  for (int t = next_tile(queue, -1, &s_t); t < ntiles; t = next_tile(queue, t, &s_t)) {
    const Tile T = tiles[t];
    if (prob_active && !prob_active[T.prob]) continue;   // uniform per block
    const QesProb& P = probs[T.prob];
    unsigned cm = 0;
#pragma unroll
    for (int r = 0; r < RC; ++r)
      if (!vp_active || vp_active[P.vp0 + r]) cm |= 1u << r;
    const int64_t a = T.row0 + threadIdx.x;
    if (a < P.n && a < T.row0 + T.rows) block_row_sell<FMT, CPLX, CV, RC>(P, a, cm, x, y);
  }
}

// neighbour table of a block problem: count pass, then fill pass in the
// exact order (and arithmetic) of the matrix-free visitor
template <typename W, bool CPLX, bool FILL>
__global__ void __launch_bounds__(BRB) table_kernel(const QesProb* __restrict__ probs, int p,
                                                    int64_t* __restrict__ cnt, const int64_t* __restrict__ rowptr,
                                                    int32_t* __restrict__ col, double* __restrict__ val) {
  extern __shared__ char smem[];
  const QesProb& P = probs[p];
  const RowCtx C = load_ctx<CPLX>(P.G, P.L, P.P, P.chi_re, P.chi_im, P.groups, P.n_diag, P.exval, P.lut, smem);
  const int64_t a = (int64_t)blockIdx.x * BRB + threadIdx.x;
  if (a >= P.n) return;
  const W r = load_rep<W>(C.L, a);
  const double inv_na = 1.0 / P.norms[a];
  int64_t k = FILL ? rowptr[a] : 0;
  auto f = [&](int64_t b, double fr, double fi, int) {
    if (FILL) {
      col[k] = (int32_t)b;
      if (CPLX) {
        val[2 * k] = fr * inv_na;
        val[2 * k + 1] = fi * inv_na;
      } else {
        val[k] = fr * inv_na;
      }
    }
    ++k;
  };
  switch (P.P.canon) {
    case QES_CANON_CYCLIC: visit_row<W, CPLX, QES_CANON_CYCLIC>(C, r, f); break;
    case QES_CANON_ROWS: visit_row<W, CPLX, QES_CANON_ROWS>(C, r, f); break;
    default: visit_row<W, CPLX, QES_CANON_BRUTE>(C, r, f); break;
  }
  if (!FILL) cnt[a] = k;
}

// ---------------------------------------------------------------------------
// batched vector kernels.  Ordinary chunks: contiguous elements of one virtual
// problem, block-wide reduction.  Block chunks: rows x R columns; thread t
// works on column t % Rq (one virtual problem) of every (BRB/Rq)-th row, and
// the per-column partial sums are reduced across the BRB/Rq row groups.
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ int block_cols(int R) {
  int rq = 1;
  while (rq < R) rq <<= 1;
  return rq;   // R <= MAX_RHS == BRB
}

// mode 0: <a, b>, mode 1: <b, b>
template <bool CPLX>
__global__ void __launch_bounds__(BRB) bdot_kernel(const Chunk* __restrict__ chunks, int nchunks,
                                                   const int* __restrict__ active, const double* __restrict__ a,
                                                   const double* __restrict__ b, double* __restrict__ part) {
  __shared__ double sr[BRB], si[BRB];
  for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const Chunk K = chunks[c];
    if (K.nrhs > 1) {
      // Rq threads per row (R rounded up to a power of two), BRB/Rq rows at once
      const int Rq = block_cols(K.nrhs);
      const int col = threadIdx.x % Rq;
      double r = 0.0, im = 0.0;
      if (col < K.nrhs && (!active || active[K.vp0 + col])) {
        for (int i = threadIdx.x / Rq; i < K.rows; i += BRB / Rq) {
          const int64_t e = K.start + (int64_t)i * K.nrhs + col;
          if (CPLX) {
            const double ar = a[2 * e], ai = a[2 * e + 1], br = b[2 * e], bi = b[2 * e + 1];
            r += ar * br + ai * bi;
            im += ar * bi - ai * br;
          } else {
            r += a[e] * b[e];
          }
        }
      }
      __syncthreads();
      sr[threadIdx.x] = r;
      si[threadIdx.x] = im;
      __syncthreads();
      for (int off = BRB / 2; off >= Rq; off >>= 1) {
        if (threadIdx.x < (unsigned)off) {
          sr[threadIdx.x] += sr[threadIdx.x + off];
          si[threadIdx.x] += si[threadIdx.x + off];
        }
        __syncthreads();
      }
      if ((int)threadIdx.x < K.nrhs) {
        part[2 * (K.slot + threadIdx.x * K.sstride)] = sr[threadIdx.x];
        part[2 * (K.slot + threadIdx.x * K.sstride) + 1] = si[threadIdx.x];
      }
      continue;
    }
    double r = 0.0, im = 0.0;
    if (!active || active[K.vp0]) {
      for (int i = threadIdx.x; i < K.rows; i += BRB) {
        const int64_t e = K.start + i;
        if (CPLX) {
          const double ar = a[2 * e], ai = a[2 * e + 1], br = b[2 * e], bi = b[2 * e + 1];
          r += ar * br + ai * bi;
          im += ar * bi - ai * br;
        } else {
          r += a[e] * b[e];
        }
      }
    }
    __syncthreads();
    sr[threadIdx.x] = r;
    si[threadIdx.x] = im;
    __syncthreads();
    for (int off = BRB / 2; off > 0; off >>= 1) {
      if (threadIdx.x < (unsigned)off) {
        sr[threadIdx.x] += sr[threadIdx.x + off];
        si[threadIdx.x] += si[threadIdx.x + off];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      part[2 * K.slot] = sr[0];
      part[2 * K.slot + 1] = si[0];
    }
  }
}

// one block per virtual problem: scal[v] = sum of its partial slots (mode 1: sqrt)
__global__ void __launch_bounds__(BRB) bfinish_kernel(const int* __restrict__ vslot0, const int* __restrict__ vnslot,
                                                      int nvp, const int* __restrict__ active,
                                                      const double* __restrict__ part, double* __restrict__ scal,
                                                      int mode) {
  __shared__ double sr[BRB], si[BRB];
  for (int p = blockIdx.x; p < nvp; p += gridDim.x) {
    double r = 0.0, im = 0.0;
    const bool on = !active || active[p];
    if (on) {
      for (int c = threadIdx.x; c < vnslot[p]; c += BRB) {
        r += part[2 * (vslot0[p] + c)];
        im += part[2 * (vslot0[p] + c) + 1];
      }
    }
    __syncthreads();
    sr[threadIdx.x] = r;
    si[threadIdx.x] = im;
    __syncthreads();
    for (int off = BRB / 2; off > 0; off >>= 1) {
      if (threadIdx.x < (unsigned)off) {
        sr[threadIdx.x] += sr[threadIdx.x + off];
        si[threadIdx.x] += si[threadIdx.x + off];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0 && on) {
      scal[2 * p] = mode == 1 ? sqrt(sr[0]) : sr[0];
      scal[2 * p + 1] = mode == 1 ? 0.0 : si[0];
    }
  }
}

__global__ void bset_alpha_kernel(int nvp, const int* __restrict__ active, const double* __restrict__ scal,
                                  double* __restrict__ alpha, const int* __restrict__ jstep, int maxs) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nvp && (!active || active[p]) && jstep[p] < maxs) alpha[(int64_t)p * maxs + jstep[p]] = scal[2 * p];
}

__global__ void bset_beta_kernel(int nvp, const int* __restrict__ active, const double* __restrict__ scal,
                                 double* __restrict__ beta, int* __restrict__ jstep, int maxs) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nvp && (!active || active[p]) && jstep[p] < maxs) {
    beta[(int64_t)p * maxs + jstep[p]] = scal[2 * p];
    jstep[p] = jstep[p] + 1;
  }
}

// w <- w - alpha v - beta v_prev for every active problem, fused with the
// partial sums of |w|^2 (the slots bdot_kernel would write for <w, w>), so
// the norm needs no second pass over w.  The coefficients are loaded once
// per chunk (ordinary problems) or once per column (block problems).
template <bool CPLX>
__global__ void __launch_bounds__(BRB) brecur_norm_kernel(const Chunk* __restrict__ chunks, int nchunks,
                                                          const int* __restrict__ active, double* __restrict__ w,
                                                          const double* __restrict__ v, const double* __restrict__ vp,
                                                          const double* __restrict__ alpha,
                                                          const double* __restrict__ beta,
                                                          const int* __restrict__ jstep, int maxs,
                                                          double* __restrict__ part) {
  __shared__ double sr[BRB];
  const int cw = CPLX ? 2 : 1;
  for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const Chunk K = chunks[c];
    const bool blk = K.nrhs > 1;
    const int Rq = blk ? block_cols(K.nrhs) : 1;
    const int col = blk ? (int)(threadIdx.x % Rq) : 0;
    const int p = K.vp0 + col;
    double r = 0.0;
    if (col < K.nrhs && (!active || active[p])) {
      const int j = jstep[p];
      if (j < maxs) {
        const double al = alpha[(int64_t)p * maxs + j];
        const double be = j > 0 ? beta[(int64_t)p * maxs + j - 1] : 0.0;
        if (blk) {
          for (int i = threadIdx.x / Rq; i < K.rows; i += BRB / Rq) {
            const int64_t e = (K.start + (int64_t)i * K.nrhs + col) * cw;
#pragma unroll
            for (int q = 0; q < cw; ++q) {
              const double t = w[e + q] - al * v[e + q] - be * vp[e + q];
              w[e + q] = t;
              r += t * t;
            }
          }
        } else {
          for (int i = threadIdx.x; i < K.rows * cw; i += BRB) {
            const int64_t e = K.start * cw + i;
            const double t = w[e] - al * v[e] - be * vp[e];
            w[e] = t;
            r += t * t;
          }
        }
      }
    }
    __syncthreads();
    sr[threadIdx.x] = r;
    __syncthreads();
    const int stop = blk ? Rq : 1;
    for (int off = BRB / 2; off >= stop; off >>= 1) {
      if (threadIdx.x < (unsigned)off) sr[threadIdx.x] += sr[threadIdx.x + off];
      __syncthreads();
    }
    if (blk) {
      if ((int)threadIdx.x < K.nrhs) {
        part[2 * (K.slot + threadIdx.x * K.sstride)] = sr[threadIdx.x];
        part[2 * (K.slot + threadIdx.x * K.sstride) + 1] = 0.0;
      }
    } else if (threadIdx.x == 0) {
      part[2 * K.slot] = sr[0];
      part[2 * K.slot + 1] = 0.0;
    }
  }
}

// v_prev <- v ; v <- w / scal (the scale loaded once per chunk or column)
template <bool CPLX>
__global__ void __launch_bounds__(BRB) badvance_kernel(const Chunk* __restrict__ chunks, int nchunks,
                                                       const int* __restrict__ active, const double* __restrict__ w,
                                                       double* __restrict__ v, double* __restrict__ vp,
                                                       const double* __restrict__ scal, int first) {
  const int cw = CPLX ? 2 : 1;
  for (int c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const Chunk K = chunks[c];
    const bool blk = K.nrhs > 1;
    const int Rq = blk ? block_cols(K.nrhs) : 1;
    const int col = blk ? (int)(threadIdx.x % Rq) : 0;
    const int p = K.vp0 + col;
    if (col >= K.nrhs || (active && !active[p])) continue;
    const double be = scal[2 * p];
    const double inv = be > 0.0 ? 1.0 / be : 0.0;
    if (blk) {
      for (int i = threadIdx.x / Rq; i < K.rows; i += BRB / Rq) {
        const int64_t e = (K.start + (int64_t)i * K.nrhs + col) * cw;
#pragma unroll
        for (int q = 0; q < cw; ++q) {
          vp[e + q] = first ? 0.0 : v[e + q];
          v[e + q] = w[e + q] * inv;
        }
      }
    } else {
      for (int i = threadIdx.x; i < K.rows * cw; i += BRB) {
        const int64_t e = K.start * cw + i;
        vp[e] = first ? 0.0 : v[e];
        v[e] = w[e] * inv;
      }
    }
  }
}

__global__ void bzero_int_kernel(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
struct BlockSpec {
  Op* op;
  int nrhs;
  const double* D;
  const double* cscale;   // host [nrhs][ASM_MAX_CPL]: coupling-scaled block (D unused), or null
};

struct Batch {
  std::vector<Op*> ops;           // real problems (ordinary then blocks)
  bool acquired = false;          // ops counted in op->users
  std::vector<int> nrhs, vp0, blk;
  int nprob = 0, nvp = 0, maxs = 0, dtype = 0;
  bool w32 = true;
  int64_t total = 0;               // elements
  std::vector<int64_t> offs;
  QesProb* d_probs = nullptr;
  Tile* d_tiles = nullptr;
  Chunk* d_chunks = nullptr;
  int* d_vslot0 = nullptr;
  int* d_vnslot = nullptr;
  int ntiles = 0, nchunks = 0, nslots = 0;
  int* d_active = nullptr;         // per virtual problem
  int* d_pactive = nullptr;        // per real problem (any column active)
  int* d_j = nullptr;
  double *v = nullptr, *vp = nullptr, *w = nullptr;
  double *alpha = nullptr, *beta = nullptr, *scal = nullptr, *part = nullptr;
  size_t lds = 0;
  int grid_matvec = 0, grid_vec = 0;
  int tile_order = 1;     // 0: problem-contiguous tiles, 1: largest problems first, round-robin
  int grid_mult = 8;      // persistent matvec blocks per CU
  int check_min = -1;     // ground states: 0 = test every check_every steps; n >= 1 = adaptive, >= n apart;
                          // -1 = adaptive (n = 1) for at most 4 problems of >= 2^20 rows, else 0
  int block_rows = 16;    // rows per tile of a block problem (at least one per lane group)
  int block_table = 1;    // block problems use their neighbour table
  bool use_tables = true; // ordinary problems stream their operator's assembled table when it has one
  int scaled_sig = -1;     // every problem a coupling-scaled block of one table format, value type and
                           // width: fmt | cv << 2 | R << 3; else -1
  int scaled_kernel = 1;   // all_scaled batches: 1 their own kernel, 0 the combined one
  int scaled_lds = 0;      // dynamic LDS bytes per block of that kernel (caps resident blocks per CU)
  int tile_queue = 0;      // 1: blocks take tiles from an atomic counter (dynamic queue); 0: static grid stride
  int* d_queue = nullptr;  // the counter
  std::vector<QesProb> h_probs;
  std::vector<int64_t*> t_rowptr;   // per real problem (null for ordinary problems)
  std::vector<int32_t*> t_col;
  std::vector<double*> t_val;
  std::vector<double*> s_buf;       // per real problem: device cscale + scaled dictionaries (coupling-scaled blocks)
  int64_t table_entries = 0;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<int> h_active;
};

// Copies ordered on the batch's own stream (never the legacy stream), so a
// batch can run while another thread captures or replays its own graph.
static hipError_t batch_copy(Batch* B, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, B->stream);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(B->stream);
}

static int batch_build_tiles(Batch* B) {
  std::vector<std::vector<Tile>> per(B->nprob);
  for (int p = 0; p < B->nprob; ++p) {
    const bool lane_rows = !B->blk[p] || B->h_probs[p].cscale;   // one lane per row
    const int rows_per_tile = lane_rows ? BRB : std::max(BRB / block_lanes(B->nrhs[p]), B->block_rows);
    for (int64_t r = 0; r < B->ops[p]->basis->n; r += rows_per_tile) per[p].push_back(Tile{p, rows_per_tile, r});
  }
  std::vector<Tile> tiles;
  if (B->tile_order == 0) {
    for (auto& v : per) tiles.insert(tiles.end(), v.begin(), v.end());
  } else {
    // largest problems first, then round-robin across problems so that every
    // wave of the persistent grid mixes problems of all sizes
    std::vector<int> order(B->nprob);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return per[a].size() > per[b].size(); });
    size_t maxlen = 0;
    for (auto& v : per) maxlen = std::max(maxlen, v.size());
    for (size_t i = 0; i < maxlen; ++i)
      for (int p : order)
        if (i < per[p].size()) tiles.push_back(per[p][i]);
  }
  if (B->d_tiles) (void)hipFree(B->d_tiles);
  B->d_tiles = nullptr;
  B->ntiles = (int)tiles.size();
  QES_HIP(hipMalloc(&B->d_tiles, std::max<size_t>(tiles.size(), 1) * sizeof(Tile)));
  if (!tiles.empty()) QES_HIP(batch_copy(B, B->d_tiles, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice));
  int dev = 0, ncu = 256;
  QES_HIP(hipGetDevice(&dev));
  QES_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  B->grid_matvec = (int)std::max<int64_t>(1, std::min<int64_t>(B->ntiles, (int64_t)ncu * B->grid_mult));
  return 0;
}

// Build the neighbour tables of the block problems (h_probs must be on the
// device without table pointers), then upload the table pointers if enabled.
static int batch_build_tables(Batch* B) {
  B->t_rowptr.assign(B->nprob, nullptr);
  B->t_col.assign(B->nprob, nullptr);
  B->t_val.assign(B->nprob, nullptr);
  B->table_entries = 0;
  const int cw = B->dtype == 1 ? 2 : 1;
  for (int p = 0; p < B->nprob; ++p) {
    if (!B->blk[p] || B->h_probs[p].cscale) continue;
    const int64_t n = B->h_probs[p].n;
    if (n >= (int64_t(1) << 31)) return fail("block problem too large for 32-bit neighbour indices");
    int64_t* d_cnt = nullptr;
    QES_HIP(hipMalloc(&d_cnt, sizeof(int64_t) * (n + 1)));
    B->t_rowptr[p] = d_cnt;   // reused as rowptr after the scan
    const dim3 grid((unsigned)((n + BRB - 1) / BRB));
    const size_t lds = B->ops[p]->lds_bytes;
    const bool c = B->dtype == 1;
#define QES_TABLE_LAUNCH(FILL, ...)                                                                               \
    do {                                                                                                          \
      if (B->w32) {                                                                                               \
        if (c) hipLaunchKernelGGL((table_kernel<uint32_t, true, FILL>), grid, dim3(BRB), lds, B->stream, __VA_ARGS__); \
        else hipLaunchKernelGGL((table_kernel<uint32_t, false, FILL>), grid, dim3(BRB), lds, B->stream, __VA_ARGS__); \
      } else {                                                                                                    \
        if (c) hipLaunchKernelGGL((table_kernel<uint64_t, true, FILL>), grid, dim3(BRB), lds, B->stream, __VA_ARGS__); \
        else hipLaunchKernelGGL((table_kernel<uint64_t, false, FILL>), grid, dim3(BRB), lds, B->stream, __VA_ARGS__); \
      }                                                                                                           \
    } while (0)
    QES_TABLE_LAUNCH(false, B->d_probs, p, d_cnt, nullptr, nullptr, nullptr);
    QES_HIP(hipGetLastError());
    std::vector<int64_t> h(n + 1);
    QES_HIP(hipMemcpyAsync(h.data(), d_cnt, sizeof(int64_t) * n, hipMemcpyDeviceToHost, B->stream));
    QES_HIP(hipStreamSynchronize(B->stream));
    int64_t acc = 0;
    for (int64_t a = 0; a < n; ++a) {
      const int64_t k = h[a];
      h[a] = acc;
      acc += k;
    }
    h[n] = acc;
    QES_HIP(batch_copy(B, d_cnt, h.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice));
    QES_HIP(hipMalloc(&B->t_col[p], sizeof(int32_t) * std::max<int64_t>(acc, 1)));
    QES_HIP(hipMalloc(&B->t_val[p], sizeof(double) * cw * std::max<int64_t>(acc, 1)));
    QES_TABLE_LAUNCH(true, B->d_probs, p, nullptr, d_cnt, B->t_col[p], B->t_val[p]);
#undef QES_TABLE_LAUNCH
    QES_HIP(hipGetLastError());
    QES_HIP(hipStreamSynchronize(B->stream));
    B->table_entries += acc;
  }
  return 0;
}

static int batch_upload_probs(Batch* B) {
  std::vector<QesProb> h = B->h_probs;
  for (int p = 0; p < B->nprob; ++p) {
    const bool on = B->block_table && B->t_rowptr.size() == (size_t)B->nprob && B->t_rowptr[p];
    h[p].rowptr = on ? B->t_rowptr[p] : nullptr;
    h[p].col = on ? B->t_col[p] : nullptr;
    h[p].val = on ? B->t_val[p] : nullptr;
    if (!B->use_tables && !h[p].cscale) h[p].T.fmt = -1;   // coupling-scaled blocks always stream their table
  }
  QES_HIP(batch_copy(B, B->d_probs, h.data(), B->nprob * sizeof(QesProb), hipMemcpyHostToDevice));
  return 0;
}

static void batch_free(Batch* B) {
  for (auto p : B->t_rowptr) (void)hipFree(p);
  for (auto p : B->t_col) (void)hipFree(p);
  for (auto p : B->t_val) (void)hipFree(p);
  for (auto p : B->s_buf) (void)hipFree(p);
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  (void)hipFree(B->d_probs);
  (void)hipFree(B->d_tiles);
  (void)hipFree(B->d_chunks);
  (void)hipFree(B->d_vslot0);
  (void)hipFree(B->d_vnslot);
  (void)hipFree(B->v);
  (void)hipFree(B->vp);
  (void)hipFree(B->w);
  (void)hipFree(B->alpha);
  (void)hipFree(B->beta);
  (void)hipFree(B->scal);
  (void)hipFree(B->part);
  (void)hipFree(B->d_active);
  (void)hipFree(B->d_pactive);
  (void)hipFree(B->d_j);
  if (B->d_queue) (void)hipFree(B->d_queue);
  if (B->stream) (void)hipStreamDestroy(B->stream);
}

int batch_create(Op* const* ops, int nreg, const BlockSpec* blocks, int nblock, int max_steps, int dtype,
                 Batch** out) {
  if (nreg + nblock < 1) return fail("empty batch");
  if (max_steps < 1) return fail("max_steps must be >= 1");
  if (dtype != 0 && dtype != 1) return fail("dtype must be 0 or 1");
  Batch* B = new Batch();
  B->maxs = max_steps;
  B->dtype = dtype;
  for (int i = 0; i < nreg; ++i) {
    B->ops.push_back(ops[i]);
    B->nrhs.push_back(1);
    B->blk.push_back(0);
  }
  std::vector<const double*> D(nreg, nullptr);
  std::vector<const double*> CS(nreg, nullptr);
  for (int i = 0; i < nblock; ++i) {
    if (blocks[i].nrhs < 1 || blocks[i].nrhs > MAX_RHS) {
      delete B;
      return fail("block problems hold 1..256 right-hand sides");
    }
    if (blocks[i].cscale) {
      const Op* op = blocks[i].op;
      if (op->asm_fmt < 0 || !op->asm_tagged) {
        delete B;
        return fail("a coupling-scaled block needs its operator assembled with coupling tags (qes_op_assemble mode 3)");
      }
    } else if (!blocks[i].D) {
      delete B;
      return fail("block problem without diagonal array");
    }
    B->ops.push_back(blocks[i].op);
    B->nrhs.push_back(blocks[i].nrhs);
    B->blk.push_back(1);
    D.push_back(blocks[i].D);
    CS.push_back(blocks[i].cscale);
  }
  B->nprob = (int)B->ops.size();
  std::vector<QesProb> h(B->nprob);
  std::vector<Chunk> chunks;
  std::vector<int> vslot0, vnslot;
  int64_t off = 0;
  int slot = 0, vp = 0;
  for (int p = 0; p < B->nprob; ++p) {
    const Op* op = B->ops[p];
    if (dtype == 0 && (op->is_complex || op->chi_complex)) {
      delete B;
      return fail("problem " + std::to_string(p) + " needs complex128 vectors");
    }
    const Basis* Bs = op->basis;
    if (Bs->group.ns > 32) B->w32 = false;
    const int R = B->nrhs[p];
    QesProb& q = h[p];
    memset(&q, 0, sizeof(q));
    q.G = Bs->group;
    q.L = Bs->lookup();
    q.P = op->plan;
    q.chi_re = Bs->chi_re;
    q.chi_im = Bs->chi_im;
    q.groups = op->d_gen;
    q.n_diag = op->n_diag;   // blocks take their diagonal from D (visit_row skips these groups)
    q.nrhs = R;
    q.blk = B->blk[p];
    q.exval = op->d_exval;
    q.lut = op->d_lut;
    q.norms = Bs->norms;
    q.D = D[p];
    q.n = Bs->n;
    q.off = off;
    q.T = asm_table_of(op);
    if (R > 1 && !CS[p]) q.T.fmt = -1;
    if (CS[p]) {
      // device: cscale [R][8], then the dictionary scaled per column [R][ndict][vw]
      const int nd = op->asm_ndict, vw = op->asm_dvc ? 2 : 1;
      std::vector<double> buf((size_t)R * ASM_MAX_CPL + (size_t)R * nd * vw, 0.0);
      for (int r = 0; r < R; ++r)
        for (int c = 0; c < ASM_MAX_CPL; ++c) buf[(size_t)r * ASM_MAX_CPL + c] = c < op->n_coupling ? CS[p][r * ASM_MAX_CPL + c] : 0.0;
      double* dc = buf.data() + (size_t)R * ASM_MAX_CPL;
      for (int r = 0; r < R; ++r)
        for (int i = 0; i < nd; ++i)
          for (int w = 0; w < vw; ++w)
            dc[((size_t)r * nd + i) * vw + w] = op->asm_h_dict[(size_t)i * vw + w] * CS[p][r * ASM_MAX_CPL + op->asm_h_dict_cpl[i]];
      double* d = nullptr;
      if (hipMalloc(&d, buf.size() * sizeof(double)) != hipSuccess) {
        for (auto b : B->s_buf) (void)hipFree(b);
        delete B;
        return fail("hipMalloc failed in batch_create");
      }
      B->s_buf.push_back(d);
      const hipError_t ce = hipMemcpy(d, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice);
      if (ce != hipSuccess) {
        for (auto b : B->s_buf) (void)hipFree(b);
        delete B;
        return check_hip(ce, "batch_create: copy of the column scales");
      }
      q.cscale = d;
      q.dcols = d + (size_t)R * ASM_MAX_CPL;
      q.ndict = nd;
      q.T.ncpl = op->n_coupling;
    }
    q.vp0 = vp;
    B->vp0.push_back(vp);
    if (R == 1) {
      vslot0.push_back(slot);
      int cnt = 0;
      for (int64_t s = 0; s < Bs->n; s += CHUNK, ++cnt)
        chunks.push_back(Chunk{p, (int)std::min<int64_t>(CHUNK, Bs->n - s), off + s, 1, vp, slot + cnt, 0});
      vnslot.push_back(cnt);
      slot += cnt;
      vp += 1;
    } else {
      const int rows = std::max(1, CHUNK / R);
      const int nbc = (int)((Bs->n + rows - 1) / rows);
      for (int k = 0; k < nbc; ++k) {
        const int64_t a0 = (int64_t)k * rows;
        chunks.push_back(Chunk{p, (int)std::min<int64_t>(rows, Bs->n - a0), off + a0 * R, R, vp, slot + k, nbc});
      }
      for (int r = 0; r < R; ++r) {
        vslot0.push_back(slot + r * nbc);
        vnslot.push_back(nbc);
      }
      slot += R * nbc;
      vp += R;
    }
    B->offs.push_back(off);
    off += Bs->n * R;
    B->lds = std::max(B->lds, op->lds_bytes);
  }
  B->offs.push_back(off);
  B->total = off;
  B->nvp = vp;
  B->nslots = slot;
  B->nchunks = (int)chunks.size();
  const int cw = dtype == 1 ? 2 : 1;
  const size_t vb = std::max<int64_t>(off, 1) * cw * sizeof(double);
  int rc = 0;
#define QES_BALLOC(ptr, bytes) \
  if (!rc && hipMalloc(&(ptr), (bytes)) != hipSuccess) rc = fail("hipMalloc failed in batch_create")
  QES_BALLOC(B->d_probs, B->nprob * sizeof(QesProb));
  QES_BALLOC(B->d_chunks, std::max<size_t>(chunks.size(), 1) * sizeof(Chunk));
  QES_BALLOC(B->d_vslot0, std::max(B->nvp, 1) * sizeof(int));
  QES_BALLOC(B->d_vnslot, std::max(B->nvp, 1) * sizeof(int));
  QES_BALLOC(B->v, vb);
  QES_BALLOC(B->vp, vb);
  QES_BALLOC(B->w, vb);
  QES_BALLOC(B->alpha, sizeof(double) * B->nvp * max_steps);
  QES_BALLOC(B->beta, sizeof(double) * B->nvp * max_steps);
  QES_BALLOC(B->scal, sizeof(double) * 2 * B->nvp);
  QES_BALLOC(B->part, sizeof(double) * 2 * std::max(B->nslots, 1));
  QES_BALLOC(B->d_active, sizeof(int) * B->nvp);
  QES_BALLOC(B->d_pactive, sizeof(int) * B->nprob);
  QES_BALLOC(B->d_j, sizeof(int) * B->nvp);
#undef QES_BALLOC
  B->h_probs = h;
  for (size_t i = 0; i < h.size(); ++i) {
    const QesProb& q = h[i];
    const int sig = q.cscale && q.nrhs <= 4 ? (q.T.fmt | (q.T.dvc ? 4 : 0) | (q.nrhs << 3)) : -2;
    B->scaled_sig = (i == 0 || sig == B->scaled_sig) ? sig : -2;
  }
  if (B->scaled_sig < 0) B->scaled_sig = -1;
  if (B->scaled_sig >= 0) {
    // coupling sweeps, measured on MI355X (6x6 J1-J2, three columns, one
    // tagged table): a block per tile beats the persistent grid (real
    // 111 -> 96 ms, complex 173 -> 145 ms per product); real sweeps then run
    // their own kernel at two resident blocks per CU (87 ms: fewer gathers
    // in flight thrash the caches less), complex ones stay in the combined
    // kernel (145 ms against 172 ms in their own)
    B->grid_mult = 1 << 20;
    B->scaled_kernel = dtype == 1 ? 0 : 1;
    B->scaled_lds = dtype == 1 ? 0 : 54272;
  }
  if (!rc) {
    B->h_active.assign(B->nvp, 1);
    std::vector<int> pa(B->nprob, 1);
    hipError_t e = hipMemcpy(B->d_probs, h.data(), B->nprob * sizeof(QesProb), hipMemcpyHostToDevice);
    if (e == hipSuccess && !chunks.empty())
      e = hipMemcpy(B->d_chunks, chunks.data(), chunks.size() * sizeof(Chunk), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(B->d_vslot0, vslot0.data(), B->nvp * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(B->d_vnslot, vnslot.data(), B->nvp * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(B->d_active, B->h_active.data(), sizeof(int) * B->nvp, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(B->d_pactive, pa.data(), sizeof(int) * B->nprob, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking);
    if (e != hipSuccess) rc = check_hip(e, "batch_create: problem tables");
  }
  B->grid_vec = std::max(1, std::min(B->nchunks, 4096));
  if (!rc) rc = batch_build_tables(B);
  if (!rc) rc = batch_upload_probs(B);
  if (!rc) rc = batch_build_tiles(B);
  if (rc) {
    batch_free(B);
    delete B;
    return -1;
  }
  for (Op* op : B->ops) op_acquire(op);
  B->acquired = true;
  *out = B;
  return 0;
}

static void batch_drop_graph(Batch* B) {
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  B->exec = nullptr;
  B->graph = nullptr;
}

int batch_destroy(Batch* B) {
  if (!B) return 0;
  if (B->acquired)
    for (Op* op : B->ops) op_release(op);
  batch_free(B);
  delete B;
  return 0;
}

__global__ void queue_reset_kernel(int* q) { *q = 0; }

static int enqueue_bmatvec(Batch* B, const double* x, double* y, const int* pactive, const int* vactive,
                           hipStream_t st) {
  const bool c = B->dtype == 1;
  // dynamic tile queue: the counter starts at zero for every product (a
  // kernel, not a memset, so that captured graphs replay it)
  int* queue = nullptr;
  if (B->tile_queue && B->d_queue) {
    hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(1), 0, st, B->d_queue);
    queue = B->d_queue;
  }
  if (B->scaled_sig >= 0 && B->scaled_kernel) {
    const size_t lds = B->scaled_lds;
    // every problem: coupling-scaled block of the same format, value type and width
    const int fmt = B->scaled_sig & 3, cv = (B->scaled_sig >> 2) & 1, R = B->scaled_sig >> 3;
#define QES_SK(XC, F, CVV, RR)                                                                                     \
  hipLaunchKernelGGL((batch_scaled_kernel<XC, F, CVV, RR>), dim3(B->grid_matvec), dim3(BRB), lds, st, B->d_probs, \
                     B->d_tiles, B->ntiles, pactive, vactive, x, y, queue)
#define QES_SK_R(XC, F, CVV)              \
  switch (R) {                            \
    case 1: QES_SK(XC, F, CVV, 1); break; \
    case 2: QES_SK(XC, F, CVV, 2); break; \
    case 3: QES_SK(XC, F, CVV, 3); break; \
    default: QES_SK(XC, F, CVV, 4); break; \
  }
#define QES_SK_F(XC)                                                \
  if (fmt == ASM_P4) {                                              \
    if (cv) { QES_SK_R(XC, ASM_P4, true) } else { QES_SK_R(XC, ASM_P4, false) } \
  } else {                                                          \
    if (cv) { QES_SK_R(XC, ASM_P6, true) } else { QES_SK_R(XC, ASM_P6, false) } \
  }
    if (c) { QES_SK_F(true) } else { QES_SK_F(false) }
#undef QES_SK_F
#undef QES_SK_R
#undef QES_SK
  } else if (B->w32) {
    if (c) hipLaunchKernelGGL((batch_matvec_kernel<uint32_t, true>), dim3(B->grid_matvec), dim3(BRB), B->lds, st,
                               B->d_probs, B->d_tiles, B->ntiles, pactive, vactive, x, y, queue);
    else hipLaunchKernelGGL((batch_matvec_kernel<uint32_t, false>), dim3(B->grid_matvec), dim3(BRB), B->lds, st,
                            B->d_probs, B->d_tiles, B->ntiles, pactive, vactive, x, y, queue);
  } else {
    if (c) hipLaunchKernelGGL((batch_matvec_kernel<uint64_t, true>), dim3(B->grid_matvec), dim3(BRB), B->lds, st,
                               B->d_probs, B->d_tiles, B->ntiles, pactive, vactive, x, y, queue);
    else hipLaunchKernelGGL((batch_matvec_kernel<uint64_t, false>), dim3(B->grid_matvec), dim3(BRB), B->lds, st,
                            B->d_probs, B->d_tiles, B->ntiles, pactive, vactive, x, y, queue);
  }
  QES_HIP(hipGetLastError());
  return 0;
}

static void enqueue_bdot(Batch* B, const double* a, const double* b, int mode, const int* active, hipStream_t st) {
  if (B->dtype == 1)
    hipLaunchKernelGGL(bdot_kernel<true>, dim3(B->grid_vec), dim3(BRB), 0, st, B->d_chunks, B->nchunks, active, a, b,
                       B->part);
  else
    hipLaunchKernelGGL(bdot_kernel<false>, dim3(B->grid_vec), dim3(BRB), 0, st, B->d_chunks, B->nchunks, active, a, b,
                       B->part);
  hipLaunchKernelGGL(bfinish_kernel, dim3(std::min(B->nvp, 4096)), dim3(BRB), 0, st, B->d_vslot0, B->d_vnslot, B->nvp,
                     active, B->part, B->scal, mode);
}

static int enqueue_bstep(Batch* B, hipStream_t st) {
  const bool c = B->dtype == 1;
  const int pg = (B->nvp + 255) / 256;
  if (enqueue_bmatvec(B, B->v, B->w, B->d_pactive, B->d_active, st)) return -1;
  enqueue_bdot(B, B->v, B->w, 0, B->d_active, st);
  hipLaunchKernelGGL(bset_alpha_kernel, dim3(pg), dim3(256), 0, st, B->nvp, B->d_active, B->scal, B->alpha, B->d_j,
                     B->maxs);
  if (c)
    hipLaunchKernelGGL(brecur_norm_kernel<true>, dim3(B->grid_vec), dim3(BRB), 0, st, B->d_chunks, B->nchunks,
                       B->d_active, B->w, B->v, B->vp, B->alpha, B->beta, B->d_j, B->maxs, B->part);
  else
    hipLaunchKernelGGL(brecur_norm_kernel<false>, dim3(B->grid_vec), dim3(BRB), 0, st, B->d_chunks, B->nchunks,
                       B->d_active, B->w, B->v, B->vp, B->alpha, B->beta, B->d_j, B->maxs, B->part);
  hipLaunchKernelGGL(bfinish_kernel, dim3(std::min(B->nvp, 4096)), dim3(BRB), 0, st, B->d_vslot0, B->d_vnslot, B->nvp,
                     B->d_active, B->part, B->scal, 1);
  if (c)
    hipLaunchKernelGGL(badvance_kernel<true>, dim3(B->grid_vec), dim3(BRB), 0, st, B->d_chunks, B->nchunks,
                       B->d_active, B->w, B->v, B->vp, B->scal, 0);
  else
    hipLaunchKernelGGL(badvance_kernel<false>, dim3(B->grid_vec), dim3(BRB), 0, st, B->d_chunks, B->nchunks,
                       B->d_active, B->w, B->v, B->vp, B->scal, 0);
  hipLaunchKernelGGL(bset_beta_kernel, dim3(pg), dim3(256), 0, st, B->nvp, B->d_active, B->scal, B->beta, B->d_j,
                     B->maxs);
  QES_HIP(hipGetLastError());
  return 0;
}

int batch_set_active(Batch* B, const int* active) {
  B->h_active.assign(active, active + B->nvp);
  std::vector<int> pa(B->nprob, 0);
  for (int p = 0; p < B->nprob; ++p)
    for (int r = 0; r < B->nrhs[p]; ++r) pa[p] |= active[B->vp0[p] + r] ? 1 : 0;
  QES_HIP(batch_copy(B, B->d_active, active, sizeof(int) * B->nvp, hipMemcpyHostToDevice));
  QES_HIP(batch_copy(B, B->d_pactive, pa.data(), sizeof(int) * B->nprob, hipMemcpyHostToDevice));
  return 0;
}

int batch_start(Batch* B, const void* v0, hipStream_t user) {
  hipStream_t st = B->stream;
  hipEvent_t ev;
  QES_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  QES_HIP(hipEventRecord(ev, user));
  QES_HIP(hipStreamWaitEvent(st, ev, 0));
  const int cw = B->dtype == 1 ? 2 : 1;
  QES_HIP(hipMemcpyAsync(B->w, v0, B->total * cw * sizeof(double), hipMemcpyDeviceToDevice, st));
  QES_HIP(hipStreamSynchronize(st));
  std::vector<int> all(B->nvp, 1);
  if (batch_set_active(B, all.data())) return -1;
  hipLaunchKernelGGL(bzero_int_kernel, dim3((B->nvp + 255) / 256), dim3(256), 0, st, B->d_j, B->nvp);
  enqueue_bdot(B, B->w, B->w, 1, B->d_active, st);
  if (B->dtype == 1)
    hipLaunchKernelGGL(badvance_kernel<true>, dim3(B->grid_vec), dim3(BRB), 0, st, B->d_chunks, B->nchunks,
                       B->d_active, B->w, B->v, B->vp, B->scal, 1);
  else
    hipLaunchKernelGGL(badvance_kernel<false>, dim3(B->grid_vec), dim3(BRB), 0, st, B->d_chunks, B->nchunks,
                       B->d_active, B->w, B->v, B->vp, B->scal, 1);
  QES_HIP(hipGetLastError());
  QES_HIP(hipStreamSynchronize(st));
  QES_HIP(hipEventDestroy(ev));
  return 0;
}

int batch_run(Batch* B, int nsteps, int use_graph) {
  hipStream_t st = B->stream;
  if (nsteps < 0) return 0;
  if (nsteps == 0 && !use_graph) return 0;
  if (use_graph) {
    // nsteps == 0 with a graph: capture only (lets a caller capture every
    // batch before running several of them from different threads)
    if (!B->exec) {
      QES_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      int rc = enqueue_bstep(B, st);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(st, &g);
      if (rc) return rc;
      if (e != hipSuccess) return check_hip(e, "hipStreamEndCapture");
      B->graph = g;
      QES_HIP(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
    }
    for (int s = 0; s < nsteps; ++s) QES_HIP(hipGraphLaunch(B->exec, st));
  } else {
    for (int s = 0; s < nsteps; ++s)
      if (enqueue_bstep(B, st)) return -1;
  }
  QES_HIP(hipStreamSynchronize(st));
  return 0;
}

int batch_get(Batch* B, double* alpha, double* beta, int* steps) {
  if (alpha) QES_HIP(batch_copy(B, alpha, B->alpha, sizeof(double) * B->nvp * B->maxs, hipMemcpyDeviceToHost));
  if (beta) QES_HIP(batch_copy(B, beta, B->beta, sizeof(double) * B->nvp * B->maxs, hipMemcpyDeviceToHost));
  if (steps) QES_HIP(batch_copy(B, steps, B->d_j, sizeof(int) * B->nvp, hipMemcpyDeviceToHost));
  return 0;
}

int batch_matvec(Batch* B, const void* x, void* y, hipStream_t st) {
  return enqueue_bmatvec(B, (const double*)x, (double*)y, nullptr, nullptr, st);
}

int batch_set_option(Batch* B, const std::string& name, int64_t value) {
  if (name == "check_min") {   // no effect on the step graph or the tiles
    B->check_min = (int)std::max<int64_t>(-1, std::min<int64_t>(value, 1 << 20));
    return 0;
  }
  if (name == "tile_order") B->tile_order = (int)value;
  else if (name == "grid_mult") B->grid_mult = (int)std::max<int64_t>(1, std::min<int64_t>(value, 1 << 20));
  else if (name == "block_rows") B->block_rows = (int)std::max<int64_t>(1, std::min<int64_t>(value, 1 << 20));
  else if (name == "block_table") B->block_table = value ? 1 : 0;
  else if (name == "op_tables") B->use_tables = value != 0;
  else if (name == "scaled_kernel") B->scaled_kernel = value ? 1 : 0;
  else if (name == "scaled_lds") B->scaled_lds = (int)std::max<int64_t>(0, std::min<int64_t>(value, 65536));
  else if (name == "tile_queue") {
    if (value && !B->d_queue) QES_HIP(hipMalloc(&B->d_queue, sizeof(int)));
    B->tile_queue = value ? 1 : 0;
  }
  else return fail("unknown batch option " + name +
                   " (tile_order, grid_mult, block_rows, block_table, op_tables, scaled_kernel, scaled_lds, "
                   "tile_queue, check_min)");
  batch_drop_graph(B);
  if (batch_upload_probs(B)) return -1;
  return batch_build_tiles(B);
}

int batch_ground_states(Batch* B, double tol, int check_every, double* e0, int* steps_out) {
  const int V = B->nvp;
  std::vector<double> a((size_t)V * B->maxs), b((size_t)V * B->maxs);
  std::vector<int> steps(V), active(V, 1), k_eff(V, 0);
  std::vector<double> cur(V, 0.0), r_prev(V, 0.0);
  std::vector<int> m_prev(V, 0);
  std::vector<int64_t> dim(V);
  for (int p = 0; p < B->nprob; ++p)
    for (int r = 0; r < B->nrhs[p]; ++r) dim[B->vp0[p] + r] = B->ops[p]->basis->n;
  if (check_every < 1) check_every = 10;
  int check_min = B->check_min;
  if (check_min < 0) {   // a test costs a host round trip: worth placing only when steps are long and few problems set it
    int64_t nmax = 0;
    for (int p = 0; p < V; ++p) nmax = std::max(nmax, dim[p]);
    check_min = (V <= 4 && nmax >= ((int64_t)1 << 20)) ? 1 : 0;
  }
  int done_steps = 0, interval = check_every;
  while (done_steps < B->maxs) {
    const int ns = std::min(interval, B->maxs - done_steps);
    if (batch_run(B, ns, 1)) return -1;
    done_steps += ns;
    if (batch_get(B, a.data(), b.data(), steps.data())) return -1;
    bool any = false;
    int next_iv = check_every;
    for (int p = 0; p < V; ++p) {
      if (!active[p]) continue;
      const double* ap = a.data() + (size_t)p * B->maxs;
      const double* bp = b.data() + (size_t)p * B->maxs;
      // a problem that broke down since the last test ran on a zero vector:
      // keep only the rows that span its Krylov space
      const int k = effective_steps(ap, bp, steps[p]);
      k_eff[p] = k;
      cur[p] = tridiag_lowest(ap, bp, k);
      const bool brk = k < steps[p] || !(std::fabs(bp[k - 1]) > 1e-14 * std::max(1.0, std::fabs(cur[p])));
      const double target = tol * std::max(1.0, std::fabs(cur[p]));
      const double r = ritz_error(ap, bp, k, cur[p]);
      if (brk || k >= dim[p] || r <= target || k >= B->maxs) {
        active[p] = 0;
      } else {
        any = true;
        if (check_min > 0)
          next_iv = std::min(next_iv, next_check(m_prev[p], r_prev[p], k, r, target, check_min, check_every));
        m_prev[p] = k;
        r_prev[p] = r;
      }
    }
    if (batch_set_active(B, active.data())) return -1;
    if (!any) break;
    interval = next_iv;
  }
  for (int p = 0; p < V; ++p) {
    if (e0) e0[p] = cur[p];
    if (steps_out) steps_out[p] = k_eff[p];
  }
  return 0;
}

}  // namespace qes

using namespace qes;

struct qes_batch_s {
  Batch* B;
};

extern "C" {

int qes_batch_create(const qes_op_t* ops, int32_t nprob, int32_t max_steps, int32_t dtype, qes_batch_t* out) {
  return qes_batch_create_blocks(ops, nprob, nullptr, nullptr, nullptr, 0, max_steps, dtype, out);
}

int qes_batch_create_blocks(const qes_op_t* ops, int32_t nprob, const qes_op_t* block_ops, const int32_t* block_nrhs,
                            const void* const* block_diag_dev, int32_t nblock, int32_t max_steps, int32_t dtype,
                            qes_batch_t* out) {
  if (!out || (nprob > 0 && !ops) || (nblock > 0 && (!block_ops || !block_nrhs || !block_diag_dev)))
    return fail("null argument");
  std::vector<Op*> v(nprob);
  for (int i = 0; i < nprob; ++i) {
    if (!ops[i]) return fail("null operator in batch");
    v[i] = qes_op_internal(ops[i]);
  }
  std::vector<BlockSpec> bl(nblock);
  for (int i = 0; i < nblock; ++i) {
    if (!block_ops[i]) return fail("null block operator in batch");
    bl[i] = BlockSpec{qes_op_internal(block_ops[i]), block_nrhs[i], (const double*)block_diag_dev[i], nullptr};
  }
  Batch* B = nullptr;
  if (batch_create(v.data(), nprob, bl.data(), nblock, max_steps, dtype, &B)) return -1;
  auto* h = new qes_batch_s();
  h->B = B;
  *out = h;
  return 0;
}

int qes_batch_create_scaled(const qes_op_t* ops, int32_t nprob, int32_t ncol, const double* scales, int32_t max_steps,
                            int32_t dtype, qes_batch_t* out) {
  if (!out || (nprob > 0 && (!ops || !scales))) return fail("null argument");
  if (nprob < 1) return fail("empty batch");
  if (ncol < 1 || ncol > 4) return fail("ncol must be 1..4 (coupling sets per sector)");
  std::vector<BlockSpec> bl(nprob);
  for (int i = 0; i < nprob; ++i) {
    if (!ops[i]) return fail("null operator in batch");
    bl[i] = BlockSpec{qes_op_internal(ops[i]), ncol, nullptr, scales + (size_t)i * ncol * ASM_MAX_CPL};
  }
  Batch* B = nullptr;
  if (batch_create(nullptr, 0, bl.data(), nprob, max_steps, dtype, &B)) return -1;
  auto* h = new qes_batch_s();
  h->B = B;
  *out = h;
  return 0;
}

int qes_batch_offsets(qes_batch_t b, int64_t* offsets) {
  if (!b || !offsets) return fail("null argument");
  for (size_t i = 0; i < b->B->offs.size(); ++i) offsets[i] = b->B->offs[i];
  return 0;
}

int qes_batch_start(qes_batch_t b, const void* v0_dev, void* stream) {
  if (!b || !v0_dev) return fail("null argument");
  return batch_start(b->B, v0_dev, (hipStream_t)stream);
}

int qes_batch_run(qes_batch_t b, int32_t nsteps, int32_t use_graph) {
  if (!b) return fail("null argument");
  return batch_run(b->B, nsteps, use_graph);
}

int qes_batch_get(qes_batch_t b, double* alpha_host, double* beta_host, int32_t* steps_host) {
  if (!b) return fail("null argument");
  return batch_get(b->B, alpha_host, beta_host, steps_host);
}

int qes_batch_set_active(qes_batch_t b, const int32_t* active_host) {
  if (!b || !active_host) return fail("null argument");
  return batch_set_active(b->B, active_host);
}

int qes_batch_matvec(qes_batch_t b, const void* x_dev, void* y_dev, void* stream) {
  if (!b || !x_dev || !y_dev) return fail("null argument");
  return batch_matvec(b->B, x_dev, y_dev, (hipStream_t)stream);
}

int qes_batch_set_option(qes_batch_t b, const char* name, int64_t value) {
  if (!b || !name) return fail("null argument");
  return batch_set_option(b->B, name, value);
}

int qes_batch_ground_states(qes_batch_t b, double tol, int32_t check_every, double* e0_host, int32_t* steps_host) {
  if (!b) return fail("null argument");
  return batch_ground_states(b->B, tol, check_every, e0_host, steps_host);
}

int qes_batch_destroy(qes_batch_t b) {
  if (!b) return 0;
  batch_destroy(b->B);
  delete b;
  return 0;
}

}  // extern "C"
