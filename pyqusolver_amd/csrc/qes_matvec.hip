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
//
// Off-diagonal terms come in two forms (see QesPlan): exchange bonds, visited
// only where the two bits differ via (x ^ rot(x, d)) masks, and generic term
// groups evaluated one by one.
#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "qes_internal.h"
#include "qes_rows.h"

namespace qes {

// minimum resident waves per SIMD the register allocation must allow (0: the
// compiler's choice)
#ifndef QES_MV_WAVES
#define QES_MV_WAVES 8
#endif
#if QES_MV_WAVES > 0
#define QES_MV_ATTR __attribute__((amdgpu_waves_per_eu(QES_MV_WAVES)))
#else
#define QES_MV_ATTR
#endif


// NB = number of right-hand sides per lane (columns j0 .. j0+NB-1 of an
// (n, ldx) row-major block).  PRE (one column): the neighbours are gathered
// from xs = N .* x (prescale_kernel), so no element needs its N_b -- a
// direct rank-map hit loads one line of xs instead of a norm and an x line.
template <typename W, bool CPLX, int CK, int NB, bool PRE, int K, bool RG = false, bool LEAN = false>
__device__ __forceinline__ void matvec_rows(const QesGroup& G, const QesLookup& L, const QesPlan& P,
                                            const double* __restrict__ chi_re, const double* __restrict__ chi_im,
                                            const QesTermGroup* __restrict__ groups, int n_diag,
                                            const double* __restrict__ exval, const uint32_t* __restrict__ lut,
                                            const double* __restrict__ norms, const double* __restrict__ x,
                                            const double* __restrict__ xs, double* __restrict__ y, int64_t n,
                                            int64_t ldx, int64_t j0, int ncols, int accumulate, int remap) {
  extern __shared__ char smem[];
  const RowCtx C = load_ctx<CPLX>(G, L, P, chi_re, chi_im, groups, n_diag, exval, lut, smem);
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x, remap);
  const int64_t a = blk * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const W r = load_rep<W>(L, a);
  const double inv_na = 1.0 / norms[a];
  const int cw = CPLX ? 2 : 1;

  double acc_r[NB], acc_i[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    acc_r[j] = 0.0;
    acc_i[j] = 0.0;
  }
  // diagonal
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
  if (PRE) {
    // off-diagonal sum over N_b x_b, scaled by 1 / N_a once
    double sr = 0.0, si = 0.0;
    if (K > 1 && L.rmap) {
      pre_row_sum<W, CPLX, CK, K, RG, LEAN>(C, r, xs, sr, si);
    } else visit_row_t<W, CPLX, CK, false>(C, r, [&](int64_t b, double fr, double fi, int) {
      const double* xp = xs + b * cw;
      if (CPLX) {
        const double xr = xp[0], xi = xp[1];
        sr += fr * xr - fi * xi;
        si += fr * xi + fi * xr;
      } else {
        sr += fr * xp[0];
      }
    }, []() {});
    acc_r[0] += sr * inv_na;
    if (CPLX) acc_i[0] += si * inv_na;
  } else {
    visit_row<W, CPLX, CK>(C, r, [&](int64_t b, double fr, double fi, int) {
      fr *= inv_na;
      fi *= inv_na;
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
    });
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


#define QES_MV_PARAMS                                                                                            \
  QesGroup G, QesLookup L, QesPlan P, const double* __restrict__ chi_re, const double* __restrict__ chi_im,       \
      const QesTermGroup* __restrict__ groups, int n_diag, const double* __restrict__ exval,                      \
      const uint32_t* __restrict__ lut, const double* __restrict__ norms, const double* __restrict__ x,         \
      const double* __restrict__ xs, double* __restrict__ y, int64_t n, int64_t ldx, int64_t j0, int ncols,      \
      int accumulate, int remap
#define QES_MV_ARGS G, L, P, chi_re, chi_im, groups, n_diag, exval, lut, norms, x, xs, y, n, ldx, j0, ncols, accumulate, remap

// Up to 32 sites: registers capped so that 8 waves per SIMD stay resident
// (the lookups and gathers are latency bound; 7 -> 8 waves: 2.88 -> 2.68 ms
// at L=32, no spills).  Wider states keep the compiler's allocation: under
// the cap their cyclic canonicaliser spills to scratch.
template <bool CPLX, int CK, int NB, bool PRE, int K, bool RG = false, bool LEAN = false>
__global__ void __launch_bounds__(256) QES_MV_ATTR matvec_kernel(QES_MV_PARAMS) {
  matvec_rows<uint32_t, CPLX, CK, NB, PRE, K, RG, LEAN>(QES_MV_ARGS);
}
template <bool CPLX, int CK, int NB, bool PRE, int K, bool RG = false, bool LEAN = false>
__global__ void __launch_bounds__(256) matvec_kernel_wide(QES_MV_PARAMS) {
  matvec_rows<uint64_t, CPLX, CK, NB, PRE, K, RG, LEAN>(QES_MV_ARGS);
}
// The same capped at 8 resident waves per SIMD, for the row-LUT canonicaliser
// with four lookups in flight: measured on MI355X at 6x6 J1-J2 (|G| = 72),
// rank gather real 107.3 -> 102.7 ms and complex 132.7 -> 124.0 ms, dense
// gathers real 115.2 -> 108.8 ms; complex dense gathers spill 76 B per lane
// under the cap and slow down (146 -> 195 ms), as do the cyclic
// canonicaliser's variants (112-384 B), so those keep the uncapped kernel.
template <bool CPLX, int CK, int NB, bool PRE, int K, bool RG = false, bool LEAN = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
matvec_kernel_wide8(QES_MV_PARAMS) {
  matvec_rows<uint64_t, CPLX, CK, NB, PRE, K, RG, LEAN>(QES_MV_ARGS);
}

// rank-indexed gather vector: rxs[rank(rep_b) * cw ..] = N_b x_b (the ranks
// of the sorted representatives ascend, so the stores walk the vector once);
// an x whose bits equal the empty marker is stored as another NaN
__global__ void __launch_bounds__(256) rank_scatter_kernel(QesLookup L, const double* __restrict__ x,
                                                           const double* __restrict__ norms,
                                                           double* __restrict__ rxs, int64_t n, int cw) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t c = colex_rank(L, L.keys64[i]);
    const double nb = norms[i];
    for (int w = 0; w < cw; ++w) {
      double v = x[i * cw + w] * nb;
      if (__double_as_longlong(v) == (long long)QES_RXS_EMPTY) v = __longlong_as_double(0x7FF8000000000001ll);
      rxs[c * cw + w] = v;
    }
  }
}

// xs = N .* x (cw doubles per row): the gather vector of the PRE products
__global__ void __launch_bounds__(256) prescale_kernel(const double* __restrict__ x, const double* __restrict__ norms,
                                                       double* __restrict__ xs, int64_t n, int cw) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double nb = norms[i];
    if (cw == 2) {
      const double2 v = reinterpret_cast<const double2*>(x)[i];
      reinterpret_cast<double2*>(xs)[i] = make_double2(v.x * nb, v.y * nb);
    } else {
      xs[i] = x[i] * nb;
    }
  }
}
#undef QES_MV_PARAMS
#undef QES_MV_ARGS
// Per-coupling outputs y[c, a] (single RHS), c = coupling id of each term.
#define QES_MAX_COUPLINGS 8
template <typename W, bool CPLX, int CK>
__global__ void __launch_bounds__(256) matvec_couplings_kernel(QesGroup G, QesLookup L, QesPlan P,
                                                               const double* __restrict__ chi_re,
                                                               const double* __restrict__ chi_im,
                                                               const QesTermGroup* __restrict__ groups, int n_diag,
                                                               const double* __restrict__ exval,
                                                               const int32_t* __restrict__ excpl,
                                                               const uint32_t* __restrict__ lut, int ncpl,
                                                               const double* __restrict__ norms,
                                                               const double* __restrict__ x, double* __restrict__ y,
                                                               int64_t n) {
  extern __shared__ char smem[];
  const RowCtx C = load_ctx<CPLX>(G, L, P, chi_re, chi_im, groups, n_diag, exval, lut, smem);
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const W r = load_rep<W>(L, a);
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
  const int ns = G.ns;
  // tag >= 0: exchange entry (class * ns + bit position); tag < 0: generic group -1-tag
  visit_row<W, CPLX, CK>(C, r, [&](int64_t b, double fr, double fi, int tag) {
    const int cp = tag >= 0 ? excpl[tag] : groups[-1 - tag].coupling;
    fr *= inv_na;
    fi *= inv_na;
    const double xr = x[b * cw], xi = CPLX ? x[b * cw + 1] : 0.0;
    const double pr = fr * xr - fi * xi, pi = fr * xi + fi * xr;
#pragma unroll
    for (int c = 0; c < QES_MAX_COUPLINGS; ++c)
      if (c == cp) {
        ar[c] += pr;
        ai[c] += pi;
      }
  });
  (void)ns;
  for (int c = 0; c < ncpl && c < QES_MAX_COUPLINGS; ++c) {
    double* yp = y + ((int64_t)c * n + a) * cw;
    yp[0] = ar[c];
    if (CPLX) yp[1] = ai[c];
  }
}

template <typename W, int CK>
__global__ void __launch_bounds__(256) count_offdiag_kernel(QesGroup G, QesLookup L, QesPlan P,
                                                            const double* __restrict__ chi_re,
                                                            const double* __restrict__ chi_im,
                                                            const QesTermGroup* __restrict__ groups, int n_diag,
                                                            const double* __restrict__ exval,
                                                            const uint32_t* __restrict__ lut, int64_t n,
                                                            unsigned long long* __restrict__ out) {
  extern __shared__ char smem[];
  const RowCtx C = load_ctx<true>(G, L, P, chi_re, chi_im, groups, n_diag, exval, lut, smem);
  __shared__ unsigned long long s_sum[2][256];
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long cnt = 0, miss = 0;
  if (a < n) {
    const W r = load_rep<W>(L, a);
    visit_row_m<W, true, CK>(C, r, [&](int64_t, double, double, int) { ++cnt; }, [&]() { ++miss; });
  }
  s_sum[0][threadIdx.x] = cnt;
  s_sum[1][threadIdx.x] = miss;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) {
      s_sum[0][threadIdx.x] += s_sum[0][threadIdx.x + off];
      s_sum[1][threadIdx.x] += s_sum[1][threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(out, s_sum[0][0]);
    atomicAdd(out + 1, s_sum[1][0]);
  }
}

// Fast canonicalisers against the brute-force one on given states.
template <typename W, int CK>
__global__ void canon_check_kernel(QesGroup G, const uint32_t* __restrict__ lut, int lut_n,
                                   const double* __restrict__ chi_re, const double* __restrict__ chi_im,
                                   const uint64_t* __restrict__ states, int64_t n,
                                   unsigned long long* __restrict__ bad) {
  extern __shared__ char smem[];
  __shared__ uint64_t s_rowlo[QES_MAX_TX];
  uint32_t* s_lut = (uint32_t*)smem;
  for (int i = threadIdx.x; i < lut_n; i += blockDim.x) s_lut[i] = lut[i];
  const GV gv = make_gv(G, s_rowlo);
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const W s = (W)states[i];
  const Canon<W> a = canonicalize<W>(gv, s);
  Canon<W> b;
  if (CK == QES_CANON_CYCLIC) {
    b = (gv.ns == (int)(sizeof(W) * 8)) ? canonicalize_cyclic<W, true>(gv, s) : canonicalize_cyclic<W, false>(gv, s);
  } else {
    b = canonicalize_rows<W>(gv, s_lut, s);
  }
  // same representative and stabiliser size; same character of the reported
  // element whenever the state belongs to the sector (outside it the
  // character varies over the minimisers and the lookup rejects the state)
  double pr = 0.0, pim = 0.0;
  const bool in_sector = rep_test<W>(gv, chi_re, chi_im, a.rep, &pr, &pim) && hypot(pr, pim) > 1e-7;
  // (characters of distinct elements that are equal in exact arithmetic,
  // e.g. exp(i*pi*n) for even n, differ in the last bits of the FP64 table)
  const bool ok = a.rep == b.rep && a.count == b.count &&
                  (!in_sector || (fabs(chi_re[a.g] - chi_re[b.g]) <= 1e-12 && fabs(chi_im[a.g] - chi_im[b.g]) <= 1e-12));
  if (!ok) atomicAdd(bad, 1ull);
}

// ---------------------------------------------------------------------------
// host: plan construction and launches
// ---------------------------------------------------------------------------
static int pick_canon(const QesGroup& G);
int pick_canon_public(const QesGroup& G) { return pick_canon(G); }

static int pick_canon(const QesGroup& G) {
  const bool cyclic = G.nty == 1 && G.row_bits == G.ns && G.cell_bits == 1 && G.ntx == G.ns && G.ntx > 1;
  if (cyclic) return QES_CANON_CYCLIC;
  if ((G.ntx > 1 || G.nty > 1) && G.row_bits <= 12 && G.ns % G.row_bits == 0) return QES_CANON_ROWS;
  return QES_CANON_BRUTE;
}

static std::vector<uint32_t> row_lut(const QesGroup& G) {
  const int Wr = G.row_bits;
  const uint32_t nr = 1u << Wr;
  const uint32_t rm = nr - 1;
  std::vector<uint32_t> lut(nr);
  for (uint32_t v = 0; v < nr; ++v) {
    uint32_t best = 0xffffffffu, am = 0;
    for (int a = 0; a < G.ntx; ++a) {
      const int r = a * G.cell_bits;
      const uint32_t t = r ? (((v << r) | (v >> (Wr - r))) & rm) : v;
      if (t < best) {
        best = t;
        am = 1u << a;
      } else if (t == best) {
        am |= 1u << a;
      }
    }
    lut[v] = best | (am << 16);
  }
  return lut;
}

void free_plan(Op* op) {
  free_assembly(op);   // the table follows the plan's visit order
  if (op->d_exval) (void)hipFree(op->d_exval);
  if (op->d_excpl) (void)hipFree(op->d_excpl);
  if (op->d_lut) (void)hipFree(op->d_lut);
  if (op->d_gen) (void)hipFree(op->d_gen);
  op->d_exval = nullptr;
  op->d_excpl = nullptr;
  op->d_lut = nullptr;
  op->d_gen = nullptr;
}

int build_plan(Op* op, int canon_override) {
  free_plan(op);
  const Basis* B = op->basis;
  const QesGroup& G = B->group;
  const int ns = G.ns;
  QesPlan P{};
  P.canon = canon_override >= 0 ? canon_override : pick_canon(G);
  P.uniform = op->bond_order_uniform;
  P.direct = op->direct_lookup;
  P.top_test = op->top_row_test;
  if (P.canon == QES_CANON_CYCLIC &&
      !(G.nty == 1 && G.row_bits == ns && G.cell_bits == 1 && G.ntx == ns && G.ntx > 1))
    return fail("cyclic canonicaliser needs a chain translation group");
  if (P.canon == QES_CANON_ROWS && !(G.row_bits <= 12 && G.ns % G.row_bits == 0 && G.ntx <= 16))
    return fail("row canonicaliser needs rows of at most 12 bits");
  // classify off-diagonal groups
  std::vector<QesTermGroup> gen;
  std::map<int, int> cls_of;   // d -> class
  std::vector<std::vector<double>> exv;
  std::vector<std::vector<int32_t>> exc;
  std::vector<uint64_t> pm;
  std::vector<int> ds;
  for (int i = op->n_diag; i < (int)op->h_groups.size(); ++i) {
    const QesTermGroup& g = op->h_groups[i];
    const bool ex = g.p0 >= 0 && g.p1 >= 0 && g.p0 != g.p1 &&
                    g.mask == ((1ull << g.p0) | (1ull << g.p1)) && g.re[0] == 0.0 && g.im[0] == 0.0 &&
                    g.re[3] == 0.0 && g.im[3] == 0.0;
    int c = -1;
    if (ex) {
      const int d = ((g.p0 - g.p1) % ns + ns) % ns;
      auto it = cls_of.find(d);
      if (it != cls_of.end()) {
        c = it->second;
        if ((pm[c] >> g.p0) & 1ull) c = -2;   // same (d, p0) twice: keep generic
      } else if ((int)ds.size() < QES_MAX_CLASSES) {
        c = (int)ds.size();
        cls_of[d] = c;
        ds.push_back(d);
        pm.push_back(0);
        exv.emplace_back(ns * 4, 0.0);
        exc.emplace_back(ns, 0);
      }
    }
    if (c >= 0) {
      pm[c] |= 1ull << g.p0;
      // bit(p0) = 0 -> table entry 1 (b0=0,b1=1); bit(p0) = 1 -> entry 2
      exv[c][(g.p0 * 2 + 0) * 2 + 0] = g.re[1];
      exv[c][(g.p0 * 2 + 0) * 2 + 1] = g.im[1];
      exv[c][(g.p0 * 2 + 1) * 2 + 0] = g.re[2];
      exv[c][(g.p0 * 2 + 1) * 2 + 1] = g.im[2];
      exc[c][g.p0] = g.coupling;
    } else {
      gen.push_back(g);
    }
  }
  P.nclass = (int)ds.size();
  for (int c = 0; c < P.nclass; ++c) {
    P.cls[c].d = ds[c];
    P.cls[c].pmask = pm[c];
  }
  P.n_gen = (int)gen.size();
  std::vector<QesTermGroup> dg(op->h_groups.begin(), op->h_groups.begin() + op->n_diag);
  dg.insert(dg.end(), gen.begin(), gen.end());
  QES_HIP(hipMalloc(&op->d_gen, std::max<size_t>(dg.size(), 1) * sizeof(QesTermGroup)));
  if (!dg.empty()) QES_HIP(hipMemcpy(op->d_gen, dg.data(), dg.size() * sizeof(QesTermGroup), hipMemcpyHostToDevice));
  std::vector<double> exflat;
  std::vector<int32_t> cpflat;
  for (int c = 0; c < P.nclass; ++c) {
    exflat.insert(exflat.end(), exv[c].begin(), exv[c].end());
    cpflat.insert(cpflat.end(), exc[c].begin(), exc[c].end());
  }
  QES_HIP(hipMalloc(&op->d_exval, std::max<size_t>(exflat.size(), 1) * sizeof(double)));
  QES_HIP(hipMalloc(&op->d_excpl, std::max<size_t>(cpflat.size(), 1) * sizeof(int32_t)));
  if (!exflat.empty()) {
    QES_HIP(hipMemcpy(op->d_exval, exflat.data(), exflat.size() * sizeof(double), hipMemcpyHostToDevice));
    QES_HIP(hipMemcpy(op->d_excpl, cpflat.data(), cpflat.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  int nlut = 0;
  if (P.canon == QES_CANON_ROWS) {
    std::vector<uint32_t> lut = row_lut(G);
    P.lut_bits = G.row_bits;
    nlut = (int)lut.size();
    QES_HIP(hipMalloc(&op->d_lut, lut.size() * sizeof(uint32_t)));
    QES_HIP(hipMemcpy(op->d_lut, lut.data(), lut.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  } else {
    QES_HIP(hipMalloc(&op->d_lut, sizeof(uint32_t)));
  }
  op->plan = P;
  op->lds_bytes = QES_CTX_FIXED_LDS + (size_t)B->gsize * 2 * sizeof(double) + exflat.size() * sizeof(double) +
                  nlut * sizeof(uint32_t);
  return 0;
}

#define QES_CK_SWITCH(CK_VAR, MACRO)                       \
  switch (CK_VAR) {                                        \
    case QES_CANON_CYCLIC: MACRO(QES_CANON_CYCLIC); break; \
    case QES_CANON_ROWS: MACRO(QES_CANON_ROWS); break;     \
    default: MACRO(QES_CANON_BRUTE); break;                \
  }

// The lean kernels' conditions on the group and the plan: base elements the
// identity and at most one plain spin flip, no phases, no generic term
// groups, no top-row test, at most 32 rows for the row canonicaliser.
bool lean_group(const Op* op) {
  const QesGroup& G = op->basis->group;
  const QesPlan& P = op->plan;
  bool lean = op->lean != 0 && P.n_gen == 0 && !P.top_test && !G.has_phase && !G.has_bphase && G.nbase >= 1 &&
              G.nbase <= 2 && (P.canon != QES_CANON_ROWS || G.ns / G.row_bits <= 32);
  for (int b = 0; lean && b < G.nbase; ++b)
    for (int k = 0; k < QES_MAX_BASE_OPS; ++k)
      if (G.base_ops[b][k] != ((b == 1 && k == 0) ? QES_BOP_FLIP : QES_BOP_NONE)) lean = false;
  return lean;
}

template <typename W, bool CPLX>
static int launch_matvec(const Op* op, const double* x, double* y, int64_t nbatch, int accumulate, double* ws,
                         hipStream_t st) {
  const Basis* B = op->basis;
  const int64_t n = B->n;
  if (n == 0) return 0;
  const int block = 256;
  const int64_t nblocks = (n + block - 1) / block;
  if (nblocks > 0x7fffffffll) return fail("matvec grid too large");
  const QesPlan& P = op->plan;
  // the row sums take the colex tables' layout from the state word (32-bit
  // states exactly for ns <= 32)
  if ((sizeof(W) == 4) != (B->group.ns <= 32)) return fail("matvec: state word does not match the site count");
  const bool pre = ws != nullptr && nbatch == 1;
  // lookups in flight per lane in the direct and canonicalising passes
  // (32-bit states: two by default -- four spill in the general kernel --
  // and four in the real lean kernels, L=32 1.91 -> 1.85 ms with rank gather)
  const bool lean_ok = lean_group(op);
  const bool four32 = lean_ok && pre && !CPLX;
  const int kb = !B->rmap ? 1 : (op->direct_batch > 0 ? op->direct_batch : (sizeof(W) == 4 && !four32 ? 2 : 4));
  // rank gather: the basis's rank-indexed vector replaces the dense pre-scaled one
  // (the lookup must be the rank map: the row sum without it gathers by index)
  const bool rg = pre && op->rank_gather && B->lookup().rmap && B->rxs && B->rxs_cw == (CPLX ? 2 : 1) && kb > 1;
  // lean kernels (identity and at most one plain spin flip as
  // base elements, no generic term groups, no top-row test): no run-time
  // switch over base operations or term kinds, far fewer uniform values
  // held across the row loop
  const bool lean = pre && kb > 1 && B->lookup().rmap && lean_ok;
  if (rg) {
    const int64_t g = std::min<int64_t>(nblocks, 256 * 64);
    hipLaunchKernelGGL(rank_scatter_kernel, dim3((unsigned)g), dim3(block), 0, st, B->lookup(), x, B->norms, B->rxs,
                       n, CPLX ? 2 : 1);
    QES_HIP(hipGetLastError());
    ws = B->rxs;
  } else if (pre) {
    const int64_t g = std::min<int64_t>(nblocks, 256 * 64);
    hipLaunchKernelGGL(prescale_kernel, dim3((unsigned)g), dim3(block), 0, st, x, B->norms, ws, n, CPLX ? 2 : 1);
    QES_HIP(hipGetLastError());
  }
  // complex pre-scaled products with several lookups in flight (pre_row_sum)
  // keep each lane's misses per exchange class in LDS after the operator tables
  const size_t lds = (pre && kb > 1 && CPLX) ? ((op->lds_bytes + 7) & ~(size_t)7) + 8 + (size_t)P.nclass * block * 8
                                     : op->lds_bytes;
  if (lds > 65536) return fail("matvec: operator tables and miss slots exceed 64 KB of LDS");
  for (int64_t j0 = 0; j0 < nbatch; j0 += 4) {
    const int ncols = (int)std::min<int64_t>(4, nbatch - j0);
#define QES_LAUNCH_K(KER)                                                                                        \
  hipLaunchKernelGGL(KER, dim3((unsigned)nblocks), dim3(block), lds, st, B->group, B->lookup(), P,                 \
                     B->chi_re, B->chi_im, op->d_gen, op->n_diag, op->d_exval, op->d_lut, B->norms, x, ws, y, n,   \
                     nbatch, j0, ncols, accumulate, op->xcd_remap)
#define QES_LAUNCH(CK)                                                                                           \
  if (rg && sizeof(W) == 4) {                                                                                    \
    if (kb == 4 && lean && CK != QES_CANON_BRUTE) QES_LAUNCH_K((matvec_kernel<CPLX, CK, 1, true, 4, true, true>)); \
    else if (kb == 4) QES_LAUNCH_K((matvec_kernel<CPLX, CK, 1, true, 4, true>));                                 \
    else if (lean && CK != QES_CANON_BRUTE) QES_LAUNCH_K((matvec_kernel<CPLX, CK, 1, true, 2, true, true>));    \
    else QES_LAUNCH_K((matvec_kernel<CPLX, CK, 1, true, 2, true>));                                              \
  } else if (rg) {                                                                                               \
    if (kb == 8) QES_LAUNCH_K((matvec_kernel_wide<CPLX, CK, 1, true, 8, true>));                                 \
    else if (kb == 2) QES_LAUNCH_K((matvec_kernel_wide<CPLX, CK, 1, true, 2, true>));                            \
    else if (CK == QES_CANON_ROWS && lean) QES_LAUNCH_K((matvec_kernel_wide8<CPLX, CK, 1, true, 4, true, true>)); \
    else if (CK == QES_CANON_ROWS) QES_LAUNCH_K((matvec_kernel_wide8<CPLX, CK, 1, true, 4, true>));              \
    else QES_LAUNCH_K((matvec_kernel_wide<CPLX, CK, 1, true, 4, true>));                                         \
  } else if (sizeof(W) == 4) {                                                                                   \
    if (pre && kb >= 4 && lean && CK != QES_CANON_BRUTE) QES_LAUNCH_K((matvec_kernel<CPLX, CK, 1, true, 4, false, true>)); \
    else if (pre && kb >= 4) QES_LAUNCH_K((matvec_kernel<CPLX, CK, 1, true, 4>));                                \
    else if (pre && kb == 2 && lean && CK != QES_CANON_BRUTE)                                                    \
      QES_LAUNCH_K((matvec_kernel<CPLX, CK, 1, true, 2, false, true>));                                          \
    else if (pre && kb == 2) QES_LAUNCH_K((matvec_kernel<CPLX, CK, 1, true, 2>));                                \
    else if (pre) QES_LAUNCH_K((matvec_kernel<CPLX, CK, 1, true, 1>));                                           \
    else if (ncols == 1) QES_LAUNCH_K((matvec_kernel<CPLX, CK, 1, false, 1>));                                   \
    else QES_LAUNCH_K((matvec_kernel<CPLX, CK, 4, false, 1>));                                                   \
  } else {                                                                                                       \
    if (pre && kb >= 4 && lean && !CPLX && CK == QES_CANON_ROWS)                                                 \
      QES_LAUNCH_K((matvec_kernel_wide8<CPLX, CK, 1, true, 4, false, true>));                                    \
    else if (pre && kb >= 4 && lean && CK == QES_CANON_ROWS)                                                     \
      QES_LAUNCH_K((matvec_kernel_wide<CPLX, CK, 1, true, 4, false, true>));                                     \
    else if (pre && kb >= 4 && !CPLX && CK == QES_CANON_ROWS) QES_LAUNCH_K((matvec_kernel_wide8<CPLX, CK, 1, true, 4>)); \
    else if (pre && kb >= 4) QES_LAUNCH_K((matvec_kernel_wide<CPLX, CK, 1, true, 4>));                           \
    else if (pre && kb == 2) QES_LAUNCH_K((matvec_kernel_wide<CPLX, CK, 1, true, 2>));                           \
    else if (pre) QES_LAUNCH_K((matvec_kernel_wide<CPLX, CK, 1, true, 1>));                                      \
    else if (ncols == 1) QES_LAUNCH_K((matvec_kernel_wide<CPLX, CK, 1, false, 1>));                              \
    else QES_LAUNCH_K((matvec_kernel_wide<CPLX, CK, 4, false, 1>));                                              \
  }
    QES_CK_SWITCH(P.canon, QES_LAUNCH)
#undef QES_LAUNCH
#undef QES_LAUNCH_K
    QES_HIP(hipGetLastError());
  }
  return 0;
}

int matvec(const Op* op, const void* x, void* y, int64_t nbatch, int dtype, int accumulate, hipStream_t st,
           double* ws) {
  const Basis* B = op->basis;
  const bool cplx = dtype == 1;
  if (!cplx && (op->is_complex || op->chi_complex))
    return fail("complex matrix elements (complex couplings or momentum sector) need complex128 vectors");
  if (nbatch < 1) return fail("nbatch must be >= 1");
  if (op->asm_fmt >= 0 && nbatch == 1)
    return assembled_matvec(op, (const double*)x, (double*)y, dtype, accumulate, st);
  if (op->asm_fmt >= 0) return assembled_matvec_multi(op, (const double*)x, (double*)y, nbatch, dtype, accumulate, st);
  const bool w32 = B->group.ns <= 32;
  const double* xd = (const double*)x;
  double* yd = (double*)y;
  if (w32) return cplx ? launch_matvec<uint32_t, true>(op, xd, yd, nbatch, accumulate, ws, st)
                       : launch_matvec<uint32_t, false>(op, xd, yd, nbatch, accumulate, ws, st);
  return cplx ? launch_matvec<uint64_t, true>(op, xd, yd, nbatch, accumulate, ws, st)
              : launch_matvec<uint64_t, false>(op, xd, yd, nbatch, accumulate, ws, st);
}

int matvec_couplings(const Op* op, const void* x, void* y, int dtype, hipStream_t st) {
  const Basis* B = op->basis;
  const bool cplx = dtype == 1;
  if (!cplx && (op->is_complex || op->chi_complex)) return fail("complex matrix elements need complex128 vectors");
  if (op->n_coupling > QES_MAX_COUPLINGS) return fail("too many coupling groups (max 8)");
  const int64_t n = B->n;
  if (n == 0) return 0;
  dim3 grid((unsigned)((n + 255) / 256));
  const double* xd = (const double*)x;
  double* yd = (double*)y;
  const QesPlan& P = op->plan;
#define QES_LC(W_, C_, CK)                                                                                          \
  hipLaunchKernelGGL((matvec_couplings_kernel<W_, C_, CK>), grid, dim3(256), op->lds_bytes, st, B->group,          \
                     B->lookup(), P, B->chi_re, B->chi_im, op->d_gen, op->n_diag, op->d_exval, op->d_excpl,         \
                     op->d_lut, op->n_coupling, B->norms, xd, yd, n)
#define QES_LC32R(CK) QES_LC(uint32_t, false, CK)
#define QES_LC32C(CK) QES_LC(uint32_t, true, CK)
#define QES_LC64R(CK) QES_LC(uint64_t, false, CK)
#define QES_LC64C(CK) QES_LC(uint64_t, true, CK)
  if (B->group.ns <= 32) {
    if (cplx) { QES_CK_SWITCH(P.canon, QES_LC32C) } else { QES_CK_SWITCH(P.canon, QES_LC32R) }
  } else {
    if (cplx) { QES_CK_SWITCH(P.canon, QES_LC64C) } else { QES_CK_SWITCH(P.canon, QES_LC64R) }
  }
#undef QES_LC
  QES_HIP(hipGetLastError());
  return 0;
}

int count_offdiag(const Op* op, int64_t* e_off, hipStream_t st, int64_t* outside) {
  const Basis* B = op->basis;
  const int64_t n = B->n;
  *e_off = 0;
  if (outside) *outside = 0;
  if (n == 0) return 0;
  unsigned long long* d = nullptr;
  QES_HIP(hipMalloc(&d, 2 * sizeof(unsigned long long)));
  QES_HIP(hipMemsetAsync(d, 0, 2 * sizeof(unsigned long long), st));
  dim3 grid((unsigned)((n + 255) / 256));
  const QesPlan& P = op->plan;
#define QES_CO(W_, CK)                                                                                              \
  hipLaunchKernelGGL((count_offdiag_kernel<W_, CK>), grid, dim3(256), op->lds_bytes, st, B->group, B->lookup(), P, \
                     B->chi_re, B->chi_im, op->d_gen, op->n_diag, op->d_exval, op->d_lut, n, d)
#define QES_CO32(CK) QES_CO(uint32_t, CK)
#define QES_CO64(CK) QES_CO(uint64_t, CK)
  if (B->group.ns <= 32) {
    QES_CK_SWITCH(P.canon, QES_CO32)
  } else {
    QES_CK_SWITCH(P.canon, QES_CO64)
  }
#undef QES_CO
  QES_HIP(hipGetLastError());
  unsigned long long h[2] = {0, 0};
  QES_HIP(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, st));
  QES_HIP(hipStreamSynchronize(st));
  QES_HIP(hipFree(d));
  *e_off = (int64_t)h[0];
  if (outside) *outside = (int64_t)h[1];
  return 0;
}

// --------------------------------------------------------------------------
// Sector-changing (projected) product.  The reference pushes every input
// representative through every group element, the operator and every output
// element (_apply_op_batch_projected_compact_seq_jit,
// Algebra/Symmetries/jit/matrix_builder_jit.py:51-108), which is
//     y = U_out^dagger O U_in x
// with U the symmetric-basis isometries.  Pull form, one lane per output
// representative r: for every g, s' = g r, and every term t of O,
// u = s' ^ mask_t is an input state with <s'|O|u> = val_t(u), and
//     y[r] = sum_g chi_out(g) sum_t val_t(u) chi_in(g*_u) sqrt|Stab(u)| x[k(u)]
//            / (|G| N_r)
// (g*_u: first minimiser of u, k(u): index of its representative).
// No atomics; no intrinsic phases (parity-y sectors are refused).
// --------------------------------------------------------------------------
template <typename W>
__global__ void __launch_bounds__(256) projected_kernel(QesGroup G, QesLookup Lin, const uint64_t* __restrict__ reps_out,
                                                        const double* __restrict__ norms_out,
                                                        const double* __restrict__ chi_out_re,
                                                        const double* __restrict__ chi_out_im,
                                                        const double* __restrict__ chi_in_re,
                                                        const double* __restrict__ chi_in_im, int n_up_in, int z2_in,
                                                        const QesTermGroup* __restrict__ groups, int nterms, int gsize,
                                                        const double* __restrict__ x, double* __restrict__ y,
                                                        int64_t n_out) {
  __shared__ uint64_t s_rowlo[QES_MAX_TX];
  const GV gv = make_gv(G, s_rowlo);
  __syncthreads();
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n_out) return;
  const W r = (W)reps_out[a];
  double ar = 0.0, ai = 0.0;
  for (int g = 0; g < gsize; ++g) {
    const W sp = apply_element<W>(gv, g, r);
    const double cor = chi_out_re[g], coi = chi_out_im[g];
    for (int t = 0; t < nterms; ++t) {
      const QesTermGroup& tg = groups[t];
      const W u = sp ^ (W)tg.mask;
      const int kb = term_bits<W>(tg, u);
      const double vr = sel4(tg.re, kb), vi = sel4(tg.im, kb);
      if (vr * vr + vi * vi < 1e-30) continue;
      if (n_up_in >= 0 && popc<W>(u) != n_up_in) continue;
      if (z2_in != 0 && ((popc<W>(u) & 1) ? -1 : 1) != z2_in) continue;
      const Canon<W> cn = canonicalize<W>(gv, u);
      const int64_t k = lookup_rep<W>(Lin, cn.rep);
      if (k < 0) continue;
      const double nb = sqrt((double)cn.count);
      // w = chi_out(g) * val * chi_in(g*) * sqrt|Stab|
      const double cir = chi_in_re[cn.g], cii = chi_in_im[cn.g];
      const double pr = cor * vr - coi * vi, pi = cor * vi + coi * vr;
      const double wr = (pr * cir - pi * cii) * nb, wi = (pr * cii + pi * cir) * nb;
      const double xr = x[2 * k], xi = x[2 * k + 1];
      ar += wr * xr - wi * xi;
      ai += wr * xi + wi * xr;
    }
  }
  const double inv = 1.0 / ((double)gsize * norms_out[a]);
  y[2 * a] = ar * inv;
  y[2 * a + 1] = ai * inv;
}

int projected_matvec(const Op* op, const Basis* out, const double* x, double* y, hipStream_t st) {
  const Basis* in = op->basis;
  const QesGroup &Gi = in->group, &Go = out->group;
  if (Gi.ns != Go.ns || Gi.nbase != Go.nbase || Gi.ntx != Go.ntx || Gi.nty != Go.nty || Gi.row_bits != Go.row_bits ||
      Gi.cell_bits != Go.cell_bits || memcmp(Gi.base_ops, Go.base_ops, sizeof(Gi.base_ops)) != 0)
    return fail("input and output sectors must share the group elements (same generators); only the sector "
                "quantum numbers may differ");
  if (Gi.has_phase || Go.has_phase || Gi.has_bphase || Go.has_bphase)
    return fail("projected products with intrinsic phases (parity y, twisted translations) are not supported");
  if (out->n == 0) return 0;
  const int nterms = op->n_diag + op->n_off;
  const dim3 grid((unsigned)((out->n + 255) / 256));
  if (Gi.ns <= 32)
    hipLaunchKernelGGL(projected_kernel<uint32_t>, grid, dim3(256), 0, st, Go, in->lookup(), out->reps, out->norms,
                       out->chi_re, out->chi_im, in->chi_re, in->chi_im, Gi.n_up, Gi.z2, op->d_groups, nterms,
                       in->gsize, x, y, out->n);
  else
    hipLaunchKernelGGL(projected_kernel<uint64_t>, grid, dim3(256), 0, st, Go, in->lookup(), out->reps, out->norms,
                       out->chi_re, out->chi_im, in->chi_re, in->chi_im, Gi.n_up, Gi.z2, op->d_groups, nterms,
                       in->gsize, x, y, out->n);
  QES_HIP(hipGetLastError());
  return 0;
}

// H_aa for every row (real part; imaginary part into dim when non-null)
template <typename W>
__global__ void diagonal_kernel(const QesTermGroup* __restrict__ groups, int n_diag, const uint64_t* __restrict__ reps,
                                int64_t n, double* __restrict__ dre, double* __restrict__ dim) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const W r = (W)reps[a];
  double vr = 0.0, vi = 0.0;
  for (int t = 0; t < n_diag; ++t) {
    const QesTermGroup& g = groups[t];
    const int k = term_bits<W>(g, r);
    vr += sel4(g.re, k);
    vi += sel4(g.im, k);
  }
  dre[a] = vr;
  if (dim) dim[a] = vi;
}

int op_diagonal(const Op* op, double* dre, double* dim, hipStream_t st) {
  const Basis* B = op->basis;
  if (B->n == 0) return 0;
  dim3 grid((unsigned)((B->n + 255) / 256));
  if (B->group.ns <= 32)
    hipLaunchKernelGGL(diagonal_kernel<uint32_t>, grid, dim3(256), 0, st, op->d_gen, op->n_diag, B->reps, B->n, dre,
                       dim);
  else
    hipLaunchKernelGGL(diagonal_kernel<uint64_t>, grid, dim3(256), 0, st, op->d_gen, op->n_diag, B->reps, B->n, dre,
                       dim);
  QES_HIP(hipGetLastError());
  return 0;
}

int canon_check(const Basis* B, int kind, const uint32_t* d_lut_in, const uint64_t* d_states, int64_t n,
                int64_t* bad_out, hipStream_t st) {
  const QesGroup& G = B->group;
  *bad_out = 0;
  if (n <= 0) return 0;
  std::vector<uint32_t> lut;
  uint32_t* d_lut = nullptr;
  int nlut = 0;
  if (kind == QES_CANON_ROWS) {
    if (!(G.row_bits <= 12 && G.ns % G.row_bits == 0 && G.ntx <= 16)) return fail("rows canonicaliser not applicable");
    lut = row_lut(G);
    nlut = (int)lut.size();
    QES_HIP(hipMalloc(&d_lut, lut.size() * 4));
    QES_HIP(hipMemcpy(d_lut, lut.data(), lut.size() * 4, hipMemcpyHostToDevice));
  } else if (kind == QES_CANON_CYCLIC) {
    if (!(G.nty == 1 && G.row_bits == G.ns && G.cell_bits == 1 && G.ntx == G.ns && G.ntx > 1))
      return fail("cyclic canonicaliser not applicable");
    QES_HIP(hipMalloc(&d_lut, 4));
  } else {
    return fail("unknown canonicaliser");
  }
  (void)d_lut_in;
  unsigned long long* d_bad = nullptr;
  QES_HIP(hipMalloc(&d_bad, sizeof(unsigned long long)));
  QES_HIP(hipMemsetAsync(d_bad, 0, sizeof(unsigned long long), st));
  dim3 grid((unsigned)((n + 255) / 256));
  const size_t lds = nlut * 4;
  if (G.ns <= 32) {
    if (kind == QES_CANON_CYCLIC)
      hipLaunchKernelGGL((canon_check_kernel<uint32_t, QES_CANON_CYCLIC>), grid, dim3(256), lds, st, G, d_lut, nlut,
                         B->chi_re, B->chi_im, d_states, n, d_bad);
    else
      hipLaunchKernelGGL((canon_check_kernel<uint32_t, QES_CANON_ROWS>), grid, dim3(256), lds, st, G, d_lut, nlut,
                         B->chi_re, B->chi_im, d_states, n, d_bad);
  } else {
    if (kind == QES_CANON_CYCLIC)
      hipLaunchKernelGGL((canon_check_kernel<uint64_t, QES_CANON_CYCLIC>), grid, dim3(256), lds, st, G, d_lut, nlut,
                         B->chi_re, B->chi_im, d_states, n, d_bad);
    else
      hipLaunchKernelGGL((canon_check_kernel<uint64_t, QES_CANON_ROWS>), grid, dim3(256), lds, st, G, d_lut, nlut,
                         B->chi_re, B->chi_im, d_states, n, d_bad);
  }
  QES_HIP(hipGetLastError());
  unsigned long long h = 0;
  QES_HIP(hipMemcpyAsync(&h, d_bad, sizeof(h), hipMemcpyDeviceToHost, st));
  QES_HIP(hipStreamSynchronize(st));
  (void)hipFree(d_bad);
  (void)hipFree(d_lut);
  *bad_out = (int64_t)h;
  return 0;
}

}  // namespace qes
