// Row machinery shared by the single-operator and batched matvec kernels:
// LDS-staged operator context, the canonicaliser dispatch and the visitor
// that enumerates the off-diagonal elements of one representative row.
#pragma once
#include "qes_internal.h"

namespace qes {

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

// Everything a lane needs to enumerate the off-diagonal elements of a row.
// All run-time-indexed tables live in LDS (see make_gv).
struct RowCtx {
  GV G;
  QesLookup L;
  int nclass, n_gen;
  const int32_t* s_cls_d;       // LDS, nclass
  const uint64_t* s_cls_pm;     // LDS, nclass
  const QesTermGroup* groups;   // diag then generic off
  int n_diag;
  const double* s_chr;          // LDS
  const double* s_chi;
  const double* s_ex;           // LDS exchange values
  const uint32_t* s_lut;        // LDS row LUT
  uint64_t* s_fail;             // LDS miss slots of pre_row_sum (after the tables, 8-byte aligned)
  int uniform;                  // bond order of the direct-lookup pass
  int direct;                   // run the direct-lookup pass
  int top_test;                 // row lattices: top-row test before a direct lookup
};

#ifndef QES_LEAN_KC
#define QES_LEAN_KC 4
#endif

// bytes of dynamic LDS used by load_ctx before the per-op tables
#define QES_CTX_FIXED_LDS (QES_MAX_TX * 8 + QES_MAX_CLASSES * 8 + QES_MAX_CLASSES * 4 + 32)

// LEAN: the group's base elements are the identity and at most one spin
// flip (no phases), checked by the host before it launches a lean kernel
template <typename W, int CK, bool LEAN = false>
__device__ __forceinline__ Canon<W> canon_dispatch(const RowCtx& C, W s) {
  if (CK == QES_CANON_CYCLIC) {
    if (sizeof(W) * 8 == 32 || sizeof(W) * 8 == 64) {
      if (C.G.ns == (int)(sizeof(W) * 8)) return canonicalize_cyclic<W, true, LEAN>(C.G, s);
    }
    return canonicalize_cyclic<W, false, LEAN>(C.G, s);
  } else if (CK == QES_CANON_ROWS) {
    return canonicalize_rows<W, LEAN>(C.G, C.s_lut, s);
  }
  return canonicalize<W>(C.G, s);
}

// Visit every valid off-diagonal element of row r: f(b, fr, fi, coupling)
// with (fr, fi) = conj(val) * chi(g*) * N_b (the caller divides by N_a);
// NORM = false leaves out N_b (for products that gather N_b x_b from a
// pre-scaled vector).  miss() is called for every nonzero element whose
// target lies outside the sector (another U(1)/Z2 filling, or an orbit of
// zero norm): the targets the reference's kernel skips
// (Algebra/Hilbert/matrix_builder.py:1366-1371).
template <typename W, bool CPLX, int CK, bool NORM, bool LEAN = false, class F, class M>
__device__ __forceinline__ void visit_row_t(const RowCtx& C, W r, F&& f, M&& miss) {
  const int ns = C.G.ns;
  const W mask = (W)C.G.mask;
  auto emit = [&](W sp, double vr, double vi, int cpl) {
    const Canon<W> cn = canon_dispatch<W, CK, LEAN>(C, sp);
    int cnt_unused = 0;
    const int64_t b = (C.L.packed && !C.L.hash) ? lookup_packed(C.L, (uint64_t)cn.rep, &cnt_unused)
                                                : lookup_rep<W>(C.L, cn.rep);
    if (b < 0) {
      miss();
      return;
    }
    // with intrinsic (parity-y, boundary) phases N_b^2 is not |Stab|: read it
    const bool phased = C.G.has_phase || C.G.has_bphase;
    const double nb = NORM ? (phased ? C.L.norms[b] : C.L.sqrtn[cn.count]) : 1.0;
    const double cr = C.s_chr[cn.g];
    double fr, fi;
    if (CPLX) {
      const double ci = C.s_chi[cn.g];
      fr = (vr * cr + vi * ci) * nb;
      fi = (vr * ci - vi * cr) * nb;
      if (phased) {   // times conj(phase of the minimising element on s')
        double pr, pi;
        element_phase<W>(C.G, cn.g, sp, &pr, &pi);
        const double ur = fr * pr + fi * pi, ui = fi * pr - fr * pi;
        fr = ur;
        fi = ui;
      }
    } else {
      fr = vr * cr * nb;
      fi = 0.0;
    }
    f(b, fr, fi, cpl);
  };
  // Exchange classes.  Pass 1 looks a class's flipped states up directly
  // (optionally walking the bonds in wave-uniform order, so that the lanes'
  // targets and key probes cluster in a few cache lines): most flipped
  // states are representatives already, and a hit on the packed key array
  // gives the stabiliser size (character of the identity = 1).  Pass 2
  // canonicalises the flipped states that missed.  Order of the elements,
  // the order of pre_row_sum: real elements class by class (the class's
  // hits by bond position, then its misses); complex elements every class's
  // hits (class, then bond position), then every class's misses.
  W fails[QES_MAX_CLASSES];
#pragma unroll
  for (int cc = 0; cc < QES_MAX_CLASSES; ++cc) fails[cc] = 0;
  for (int c = 0; c < C.nclass; ++c) {
    const int d = C.s_cls_d[c];
    const W pm = (W)C.s_cls_pm[c];
    const W act = (r ^ rotl_ns<W>(r, d, ns, mask)) & pm;
    // direct hits: packed keys carry |Stab|; a rank-map hit takes N_b from
    // the norms (the first minimiser of a representative has character 1)
    // direct / uniform: 1 on, 0 off, -1 on exactly where a rank map serves
    const bool direct = C.direct > 0 ? (C.L.keyc || C.L.rmap) : (C.direct < 0 && C.L.rmap);
    const bool uniform = C.uniform > 0 || (C.uniform < 0 && C.L.rmap);
    W fail = direct ? (W)0 : act;
    if (direct) {
      W m = uniform ? pm : act;   // uniform order, or each lane's active bonds
      while (m) {
        const int p = (sizeof(W) == 4) ? __builtin_ctz((uint32_t)m) : __builtin_ctzll((uint64_t)m);
        m &= m - 1;
        if (!((act >> p) & 1)) continue;
        int p1 = p - d;
        if (p1 < 0) p1 += ns;
        const W sp = r ^ (((W)1 << p) | ((W)1 << p1));
        // row lattices: a state whose top row is not the smallest row
        // rotation is no representative -- skip its lookup
        if (!LEAN && CK == QES_CANON_ROWS && C.top_test && !rows_top_ok<W>(C.G, C.s_lut, sp)) {
          fail |= (W)1 << p;
          continue;
        }
        int cnt = 0;
        const int64_t b = C.L.rmap ? lookup_rank(C.L, (uint64_t)sp) : lookup_packed(C.L, (uint64_t)sp, &cnt);
        if (b < 0) {
          fail |= (W)1 << p;
          continue;
        }
        const int b0 = (int)((r >> p) & 1);
        const double* e = C.s_ex + ((c * ns + p) * 2 + b0) * 2;
        const double nb = NORM ? (C.L.rmap ? C.L.norms[b] : C.L.sqrtn[cnt]) : 1.0;
        f(b, e[0] * nb, CPLX ? -e[1] * nb : 0.0, c * ns + p);
      }
    }
    if (CPLX) {
#pragma unroll
      for (int cc = 0; cc < QES_MAX_CLASSES; ++cc)
        if (cc == c) fails[cc] = fail;
      continue;
    }
    while (fail) {
      const int p = (sizeof(W) == 4) ? __builtin_ctz((uint32_t)fail) : __builtin_ctzll((uint64_t)fail);
      fail &= fail - 1;
      int p1 = p - d;
      if (p1 < 0) p1 += ns;
      const W sp = r ^ (((W)1 << p) | ((W)1 << p1));
      const int b0 = (int)((r >> p) & 1);
      const double* e = C.s_ex + ((c * ns + p) * 2 + b0) * 2;
      emit(sp, e[0], CPLX ? e[1] : 0.0, c * ns + p);
    }
  }
  for (int c = 0; CPLX && c < C.nclass; ++c) {
    const int d = C.s_cls_d[c];
    W fail = 0;
#pragma unroll
    for (int cc = 0; cc < QES_MAX_CLASSES; ++cc)
      if (cc == c) fail = fails[cc];
    while (fail) {
      const int p = (sizeof(W) == 4) ? __builtin_ctz((uint32_t)fail) : __builtin_ctzll((uint64_t)fail);
      fail &= fail - 1;
      int p1 = p - d;
      if (p1 < 0) p1 += ns;
      const W sp = r ^ (((W)1 << p) | ((W)1 << p1));
      const int b0 = (int)((r >> p) & 1);
      const double* e = C.s_ex + ((c * ns + p) * 2 + b0) * 2;
      emit(sp, e[0], CPLX ? e[1] : 0.0, c * ns + p);
    }
  }
  // generic groups
  for (int t = C.n_diag; !LEAN && t < C.n_diag + C.n_gen; ++t) {
    const QesTermGroup& g = C.groups[t];
    const int k = term_bits<W>(g, r);
    const double vr = sel4(g.re, k);
    const double vi = CPLX ? sel4(g.im, k) : 0.0;
    if (vr == 0.0 && vi == 0.0) continue;
    const W sp = r ^ (W)g.mask;
    if ((C.G.n_up >= 0 && popc<W>(sp) != C.G.n_up) || (C.G.z2 != 0 && ((popc<W>(sp) & 1) ? -1 : 1) != C.G.z2)) {
      miss();
      continue;
    }
    emit(sp, vr, vi, -1 - t);
  }
}

// Off-diagonal sum of row r over a pre-scaled vector xs = N .* x (PRE
// products): sum_b conj(val) chi(g*) xs[b], the same elements in the same
// order as visit_row_t<NORM = false>.  With a rank map each pass takes K
// bonds at a time: K colex ranks, then K map words, then K gathers are in
// flight together, where visit_row_t waits for each lookup before the next.
// Complex elements need the LDS miss slots C.s_fail (nclass 8-byte words per
// lane of the block).
//
// RG (rank gather): xs is the basis's rank-indexed vector (N_b x_b at the
// colex rank of representative b, QES_RXS_EMPTY elsewhere), so a lookup is
// one load at the state's rank instead of a map word and then an x element;
// the same elements are summed in the same order.
//
// LEAN: the group is lean (canon_dispatch: no phases), the plan has no
// generic term groups and no top-row test -- the host checks.
template <typename W, bool CPLX, int CK, int K, bool RG = false, bool LEAN = false>
__device__ __forceinline__ void pre_row_sum(const RowCtx& C, W r, const double* __restrict__ xs, double& sr,
                                            double& si) {
  // states canonicalised together in the real canonicalising pass (each
  // canonicalisation holds its own candidate state; lean kernels may keep
  // fewer: with the lean row canonicaliser four measured fastest at 6x6,
  // 88.6 against 89.6 ms for two)
  constexpr int KC = (LEAN && K > QES_LEAN_KC) ? QES_LEAN_KC : K;
  const int ns = C.G.ns;
  const W mask = (W)C.G.mask;
  const int cw = CPLX ? 2 : 1;
  const bool direct = C.L.rmap != nullptr && C.direct != 0;
  auto add = [&](int64_t b, double fr, double fi) {
    const double* xp = xs + b * cw;
    if (CPLX) {
      const double xr = xp[0], xi = xp[1];
      sr += fr * xr - fi * xi;
      si += fr * xi + fi * xr;
    } else {
      sr += fr * xp[0];
    }
  };
  auto emit = [&](W sp, double vr, double vi) {
    const Canon<W> cn = canon_dispatch<W, CK, LEAN>(C, sp);
    int64_t b;
    if (RG) {   // the element of xs sits at the representative's rank (the marker: outside the sector)
      if (popc<W>(cn.rep) != C.L.n_up) return;
      b = (int64_t)colex_rank_w<W>(C.L, (uint64_t)cn.rep);
      if (__double_as_longlong(xs[b * cw]) == (long long)QES_RXS_EMPTY) return;
    } else {
      b = lookup_rep<W>(C.L, cn.rep);
    }
    if (b < 0) return;
    const double cr = C.s_chr[cn.g];
    if (CPLX) {
      const double ci = C.s_chi[cn.g];
      double fr = vr * cr + vi * ci, fi = vr * ci - vi * cr;
      if (C.G.has_phase || C.G.has_bphase) {   // times conj(phase of the minimising element on s')
        double pr, pi;
        element_phase<W>(C.G, cn.g, sp, &pr, &pi);
        const double ur = fr * pr + fi * pi, ui = fi * pr - fr * pi;
        fr = ur;
        fi = ui;
      }
      add(b, fr, fi);
    } else {
      add(b, vr * cr, 0.0);
    }
  };
  if (!CPLX) {
    // real elements: class by class, each class's direct hits, then its
    // misses (no LDS slots: the occupancy of the real kernels is worth more
    // than the steps the merged miss stream saves them)
    const bool uniform = C.uniform > 0 || (C.uniform < 0 && C.L.rmap);
    for (int c = 0; c < C.nclass; ++c) {
      const int d = C.s_cls_d[c];
      const W pm = (W)C.s_cls_pm[c];
      const W act = (r ^ rotl_ns<W>(r, d, ns, mask)) & pm;
      W fail = direct ? (W)0 : act;
      if (direct) {
        W m = uniform ? pm : act;
        while (m) {
          int pos[K];
          bool on[K];
          uint64_t rk[K];
  #pragma unroll
          for (int k = 0; k < K; ++k) {
            pos[k] = 0;
            on[k] = false;
            if (m) {
              pos[k] = (sizeof(W) == 4) ? __builtin_ctz((uint32_t)m) : __builtin_ctzll((uint64_t)m);
              m &= m - 1;
              on[k] = (act >> pos[k]) & 1;
            }
          }
  #pragma unroll
          for (int k = 0; k < K; ++k) {
            rk[k] = 0;
            if (on[k]) {
              int p1 = pos[k] - d;
              if (p1 < 0) p1 += ns;
              const W sp = r ^ (((W)1 << pos[k]) | ((W)1 << p1));
              // row lattices: a state whose top row exceeds the smallest
              // rotation of one of its rows (or of their images under the base
              // operations) is no representative -- no lookup; it goes to the
              // canonicalising pass as a failed lookup would, so the sums keep
              // their order
              if (!LEAN && CK == QES_CANON_ROWS && C.top_test && !rows_top_ok<W>(C.G, C.s_lut, sp)) {
                on[k] = false;
                fail |= (W)1 << pos[k];
                continue;
              }
              // an exchange keeps the filling, so the state has a colex rank
              rk[k] = colex_rank_w<W>(C.L, (uint64_t)sp);
            }
          }
          int64_t bk[K];
          double xr[K], xi[K];
          if (RG) {
            // one load per state at its rank; the marker: no representative
  #pragma unroll
            for (int k = 0; k < K; ++k) {
              xr[k] = 0.0;
              xi[k] = 0.0;
              bk[k] = -1;
              if (on[k]) {
                const double* xp = xs + rk[k] * cw;
                xr[k] = xp[0];
                if (CPLX) xi[k] = xp[1];
              }
            }
  #pragma unroll
            for (int k = 0; k < K; ++k) {
              if (on[k]) {
                bk[k] = __double_as_longlong(xr[k]) == (long long)QES_RXS_EMPTY ? -1 : 0;
                if (bk[k] < 0) fail |= (W)1 << pos[k];
              }
            }
          } else {
            ulonglong2 e[K];
  #pragma unroll
            for (int k = 0; k < K; ++k)
              e[k] = on[k] ? reinterpret_cast<const ulonglong2*>(C.L.rmap)[rk[k] >> 6] : make_ulonglong2(0ull, 0ull);
  #pragma unroll
            for (int k = 0; k < K; ++k) {
              bk[k] = on[k] ? rank_resolve(e[k], rk[k]) : -1;
              if (on[k] && bk[k] < 0) fail |= (W)1 << pos[k];
            }
  #pragma unroll
            for (int k = 0; k < K; ++k) {
              xr[k] = 0.0;
              xi[k] = 0.0;
              if (bk[k] >= 0) {
                const double* xp = xs + bk[k] * cw;
                xr[k] = xp[0];
                if (CPLX) xi[k] = xp[1];
              }
            }
          }
  #pragma unroll
          for (int k = 0; k < K; ++k) {
            if (bk[k] >= 0) {
              const int b0 = (int)((r >> pos[k]) & 1);
              const double* ev = C.s_ex + ((c * ns + pos[k]) * 2 + b0) * 2;
              const double fr = ev[0];
              if (CPLX) {
                const double fi = -ev[1];
                sr += fr * xr[k] - fi * xi[k];
                si += fr * xi[k] + fi * xr[k];
              } else {
                sr += fr * xr[k];
              }
            }
          }
        }
      }
      // states that are no representative: canonicalise KC at a time, then
      // their ranks, map words and gathers in flight together (rank map only)
      while (C.L.rmap && fail) {
        int pos[KC];
        bool on[KC];
        Canon<W> cn[KC];
        uint64_t rk[KC];
  #pragma unroll
        for (int k = 0; k < KC; ++k) {
          pos[k] = 0;
          on[k] = false;
          if (fail) {
            pos[k] = (sizeof(W) == 4) ? __builtin_ctz((uint32_t)fail) : __builtin_ctzll((uint64_t)fail);
            fail &= fail - 1;
            on[k] = true;
          }
        }
  #pragma unroll
        for (int k = 0; k < KC; ++k) {
          rk[k] = 0;
          if (on[k]) {
            int p1 = pos[k] - d;
            if (p1 < 0) p1 += ns;
            cn[k] = canon_dispatch<W, CK, LEAN>(C, r ^ (((W)1 << pos[k]) | ((W)1 << p1)));
            // outside the filling (not for a group of spin flips at half filling): no rank
            on[k] = popc<W>(cn[k].rep) == C.L.n_up;
            if (on[k]) rk[k] = colex_rank_w<W>(C.L, (uint64_t)cn[k].rep);
          }
        }
        int64_t bk[KC];
        double xr[KC], xi[KC];
        if (RG) {
  #pragma unroll
          for (int k = 0; k < KC; ++k) {
            xr[k] = 0.0;
            xi[k] = 0.0;
            bk[k] = -1;
            if (on[k]) {
              const double* xp = xs + rk[k] * cw;
              xr[k] = xp[0];
              if (CPLX) xi[k] = xp[1];
            }
          }
  #pragma unroll
          for (int k = 0; k < KC; ++k) {
            if (on[k]) bk[k] = __double_as_longlong(xr[k]) == (long long)QES_RXS_EMPTY ? -1 : 0;
            if (bk[k] < 0) xr[k] = xi[k] = 0.0;   // orbit outside the sector: skipped
          }
        } else {
          ulonglong2 e[KC];
  #pragma unroll
          for (int k = 0; k < KC; ++k)
            e[k] = on[k] ? reinterpret_cast<const ulonglong2*>(C.L.rmap)[rk[k] >> 6] : make_ulonglong2(0ull, 0ull);
  #pragma unroll
          for (int k = 0; k < KC; ++k) {
            bk[k] = on[k] ? rank_resolve(e[k], rk[k]) : -1;
            xr[k] = 0.0;
            xi[k] = 0.0;
            if (bk[k] >= 0) {
              const double* xp = xs + bk[k] * cw;
              xr[k] = xp[0];
              if (CPLX) xi[k] = xp[1];
            }
          }
        }
  #pragma unroll
        for (int k = 0; k < KC; ++k) {
          if (bk[k] >= 0) {
            const int b0 = (int)((r >> pos[k]) & 1);
            const double* ev = C.s_ex + ((c * ns + pos[k]) * 2 + b0) * 2;
            const double vr = ev[0], vi = CPLX ? ev[1] : 0.0;
            const double cr = C.s_chr[cn[k].g];
            if (CPLX) {
              const double ci = C.s_chi[cn[k].g];
              double fr = vr * cr + vi * ci, fi = vr * ci - vi * cr;
              if (C.G.has_phase || C.G.has_bphase) {   // times conj(phase of the minimising element on s')
                int p1 = pos[k] - d;
                if (p1 < 0) p1 += ns;
                double pr, pi;
                element_phase<W>(C.G, cn[k].g, r ^ (((W)1 << pos[k]) | ((W)1 << p1)), &pr, &pi);
                const double ur = fr * pr + fi * pi, ui = fi * pr - fr * pi;
                fr = ur;
                fi = ui;
              }
              sr += fr * xr[k] - fi * xi[k];
              si += fr * xi[k] + fi * xr[k];
            } else {
              sr += (vr * cr) * xr[k];
            }
          }
        }
      }
      while (fail) {
        const int p = (sizeof(W) == 4) ? __builtin_ctz((uint32_t)fail) : __builtin_ctzll((uint64_t)fail);
        fail &= fail - 1;
        int p1 = p - d;
        if (p1 < 0) p1 += ns;
        const W sp = r ^ (((W)1 << p) | ((W)1 << p1));
        const int b0 = (int)((r >> p) & 1);
        const double* ev = C.s_ex + ((c * ns + p) * 2 + b0) * 2;
        emit(sp, ev[0], CPLX ? ev[1] : 0.0);
      }
    }
  } else
  {
    if (!C.L.rmap) {   // (callers pass a rank map; without one every state is canonicalised, class by class)
      for (int c = 0; c < C.nclass; ++c) {
        const int d = C.s_cls_d[c];
        W fail = (r ^ rotl_ns<W>(r, d, ns, mask)) & (W)C.s_cls_pm[c];
        while (fail) {
          const int p = (sizeof(W) == 4) ? __builtin_ctz((uint32_t)fail) : __builtin_ctzll((uint64_t)fail);
          fail &= fail - 1;
          int p1 = p - d;
          if (p1 < 0) p1 += ns;
          const int b0 = (int)((r >> p) & 1);
          const double* ev = C.s_ex + ((c * ns + p) * 2 + b0) * 2;
          emit(r ^ (((W)1 << p) | ((W)1 << p1)), ev[0], CPLX ? ev[1] : 0.0);
        }
      }
    } else if (C.nclass > 0) {
      // Each lane walks one stream of (class, bond) items over all exchange
      // classes, K items per step, so a wavefront runs as many steps as its
      // longest row needs rather than the sum over classes of the longest
      // class.  Pass 1 looks the flipped states up directly and records the
      // misses in the lane's LDS slot of each class (C.s_fail, nclass words per
      // lane at a stride of blockDim.x); pass 2 walks the misses the same way.
      uint64_t* const sf = C.s_fail + threadIdx.x;
      const int bd = blockDim.x;
      auto act_of = [&](int cc, int dd) -> W { return (r ^ rotl_ns<W>(r, dd, ns, mask)) & (W)C.s_cls_pm[cc]; };
      if (direct) {
        int c = 0, dc = C.s_cls_d[0];
        W m = act_of(0, dc);
        sf[0] = 0;
        while (true) {
          int pos[K], cl[K], dk[K];
          bool on[K];
          uint64_t rk[K];
  #pragma unroll
          for (int k = 0; k < K; ++k) {
            while (!m && c + 1 < C.nclass) {   // next class with an active bond (every class's slot is cleared)
              ++c;
              dc = C.s_cls_d[c];
              m = act_of(c, dc);
              sf[c * bd] = 0;
            }
            pos[k] = 0;
            cl[k] = c;
            dk[k] = dc;
            on[k] = m != 0;
            if (on[k]) {
              pos[k] = (sizeof(W) == 4) ? __builtin_ctz((uint32_t)m) : __builtin_ctzll((uint64_t)m);
              m &= m - 1;
            }
          }
          if (!on[0]) break;   // slots fill in stream order: an empty first slot ends the stream
  #pragma unroll
          for (int k = 0; k < K; ++k) {
            rk[k] = 0;
            if (on[k]) {
              int p1 = pos[k] - dk[k];
              if (p1 < 0) p1 += ns;
              const W sp = r ^ (((W)1 << pos[k]) | ((W)1 << p1));
              // row lattices: a state whose top row exceeds the smallest
              // rotation of one of its rows (or of their images under the base
              // operations) is no representative -- no lookup; it goes to the
              // canonicalising pass as a failed lookup would
              if (!LEAN && CK == QES_CANON_ROWS && C.top_test && !rows_top_ok<W>(C.G, C.s_lut, sp)) {
                on[k] = false;
                __hip_atomic_fetch_or(sf + cl[k] * bd, 1ull << pos[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                continue;
              }
              // an exchange keeps the filling, so the state has a colex rank
              rk[k] = colex_rank_w<W>(C.L, (uint64_t)sp);
            }
          }
          bool hit[K];
          double xr[K], xi[K];
          if (RG) {
            // one load per state at its rank; the marker: no representative
  #pragma unroll
            for (int k = 0; k < K; ++k) {
              xr[k] = 0.0;
              xi[k] = 0.0;
              if (on[k]) {
                const double* xp = xs + rk[k] * cw;
                xr[k] = xp[0];
                if (CPLX) xi[k] = xp[1];
              }
            }
  #pragma unroll
            for (int k = 0; k < K; ++k) hit[k] = on[k] && __double_as_longlong(xr[k]) != (long long)QES_RXS_EMPTY;
          } else {
            ulonglong2 e[K];
            int64_t bk[K];
  #pragma unroll
            for (int k = 0; k < K; ++k)
              e[k] = on[k] ? reinterpret_cast<const ulonglong2*>(C.L.rmap)[rk[k] >> 6] : make_ulonglong2(0ull, 0ull);
  #pragma unroll
            for (int k = 0; k < K; ++k) {
              bk[k] = on[k] ? rank_resolve(e[k], rk[k]) : -1;
              hit[k] = bk[k] >= 0;
            }
  #pragma unroll
            for (int k = 0; k < K; ++k) {
              xr[k] = 0.0;
              xi[k] = 0.0;
              if (hit[k]) {
                const double* xp = xs + bk[k] * cw;
                xr[k] = xp[0];
                if (CPLX) xi[k] = xp[1];
              }
            }
          }
  #pragma unroll
          for (int k = 0; k < K; ++k) {
            if (on[k] && !hit[k])
              __hip_atomic_fetch_or(sf + cl[k] * bd, 1ull << pos[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          }
  #pragma unroll
          for (int k = 0; k < K; ++k) {
            if (hit[k]) {
              const int b0 = (int)((r >> pos[k]) & 1);
              const double* ev = C.s_ex + ((cl[k] * ns + pos[k]) * 2 + b0) * 2;
              const double fr = ev[0];
              if (CPLX) {
                const double fi = -ev[1];
                sr += fr * xr[k] - fi * xi[k];
                si += fr * xi[k] + fi * xr[k];
              } else {
                sr += fr * xr[k];
              }
            }
          }
        }
      }
      // states that are no representative (every active bond's state without
      // the direct pass): canonicalise K at a time, then their ranks, map words
      // and gathers in flight together
      int c = 0, dc = C.s_cls_d[0];
      W m = direct ? (W)sf[0] : act_of(0, dc);
      while (true) {
        int pos[K], cl[K], dk[K];
        bool on[K];
        Canon<W> cn[K];
        uint64_t rk[K];
  #pragma unroll
        for (int k = 0; k < K; ++k) {
          while (!m && c + 1 < C.nclass) {
            ++c;
            dc = C.s_cls_d[c];
            m = direct ? (W)sf[c * bd] : act_of(c, dc);
          }
          pos[k] = 0;
          cl[k] = c;
          dk[k] = dc;
          on[k] = m != 0;
          if (on[k]) {
            pos[k] = (sizeof(W) == 4) ? __builtin_ctz((uint32_t)m) : __builtin_ctzll((uint64_t)m);
            m &= m - 1;
          }
        }
        if (!on[0]) break;
  #pragma unroll
        for (int k = 0; k < K; ++k) {
          rk[k] = 0;
          if (on[k]) {
            int p1 = pos[k] - dk[k];
            if (p1 < 0) p1 += ns;
            cn[k] = canon_dispatch<W, CK, LEAN>(C, r ^ (((W)1 << pos[k]) | ((W)1 << p1)));
            // outside the filling (not for a group of spin flips at half filling): no rank
            on[k] = popc<W>(cn[k].rep) == C.L.n_up;
            if (on[k]) rk[k] = colex_rank_w<W>(C.L, (uint64_t)cn[k].rep);
          }
        }
        int64_t bk[K];
        double xr[K], xi[K];
        if (RG) {
  #pragma unroll
          for (int k = 0; k < K; ++k) {
            xr[k] = 0.0;
            xi[k] = 0.0;
            bk[k] = -1;
            if (on[k]) {
              const double* xp = xs + rk[k] * cw;
              xr[k] = xp[0];
              if (CPLX) xi[k] = xp[1];
            }
          }
  #pragma unroll
          for (int k = 0; k < K; ++k) {
            if (on[k]) bk[k] = __double_as_longlong(xr[k]) == (long long)QES_RXS_EMPTY ? -1 : 0;
            if (bk[k] < 0) xr[k] = xi[k] = 0.0;   // orbit outside the sector: skipped
          }
        } else {
          ulonglong2 e[K];
  #pragma unroll
          for (int k = 0; k < K; ++k)
            e[k] = on[k] ? reinterpret_cast<const ulonglong2*>(C.L.rmap)[rk[k] >> 6] : make_ulonglong2(0ull, 0ull);
  #pragma unroll
          for (int k = 0; k < K; ++k) {
            bk[k] = on[k] ? rank_resolve(e[k], rk[k]) : -1;
            xr[k] = 0.0;
            xi[k] = 0.0;
            if (bk[k] >= 0) {
              const double* xp = xs + bk[k] * cw;
              xr[k] = xp[0];
              if (CPLX) xi[k] = xp[1];
            }
          }
        }
  #pragma unroll
        for (int k = 0; k < K; ++k) {
          if (bk[k] >= 0) {
            const int b0 = (int)((r >> pos[k]) & 1);
            const double* ev = C.s_ex + ((cl[k] * ns + pos[k]) * 2 + b0) * 2;
            const double vr = ev[0], vi = CPLX ? ev[1] : 0.0;
            const double cr = C.s_chr[cn[k].g];
            if (CPLX) {
              const double ci = C.s_chi[cn[k].g];
              double fr = vr * cr + vi * ci, fi = vr * ci - vi * cr;
              if (C.G.has_phase || C.G.has_bphase) {   // times conj(phase of the minimising element on s')
                int p1 = pos[k] - dk[k];
                if (p1 < 0) p1 += ns;
                double pr, pi;
                element_phase<W>(C.G, cn[k].g, r ^ (((W)1 << pos[k]) | ((W)1 << p1)), &pr, &pi);
                const double ur = fr * pr + fi * pi, ui = fi * pr - fr * pi;
                fr = ur;
                fi = ui;
              }
              sr += fr * xr[k] - fi * xi[k];
              si += fr * xi[k] + fi * xr[k];
            } else {
              sr += (vr * cr) * xr[k];
            }
          }
        }
      }
    }
  }
  for (int t = C.n_diag; !LEAN && t < C.n_diag + C.n_gen; ++t) {
    const QesTermGroup& g = C.groups[t];
    const int k = term_bits<W>(g, r);
    const double vr = sel4(g.re, k);
    const double vi = CPLX ? sel4(g.im, k) : 0.0;
    if (vr == 0.0 && vi == 0.0) continue;
    const W sp = r ^ (W)g.mask;
    if ((C.G.n_up >= 0 && popc<W>(sp) != C.G.n_up) || (C.G.z2 != 0 && ((popc<W>(sp) & 1) ? -1 : 1) != C.G.z2))
      continue;
    emit(sp, vr, vi);
  }
}

template <typename W, bool CPLX, int CK, class F, class M>
__device__ __forceinline__ void visit_row_m(const RowCtx& C, W r, F&& f, M&& miss) {
  visit_row_t<W, CPLX, CK, true>(C, r, f, miss);
}

// LEAN: as in canon_dispatch, and the plan has no generic term groups and
// no top-row test (the host checks)
template <typename W, bool CPLX, int CK, bool LEAN = false, class F>
__device__ __forceinline__ void visit_row(const RowCtx& C, W r, F&& f) {
  visit_row_t<W, CPLX, CK, true, LEAN>(C, r, f, []() {});
}

// Load the per-block tables into LDS; returns the context.
template <bool CPLX>
__device__ __forceinline__ RowCtx load_ctx(const QesGroup& G, const QesLookup& L, const QesPlan& P,
                                           const double* __restrict__ chi_re, const double* __restrict__ chi_im,
                                           const QesTermGroup* __restrict__ groups, int n_diag,
                                           const double* __restrict__ exval, const uint32_t* __restrict__ lut,
                                           char* smem) {
  uint64_t* s_rowlo = (uint64_t*)smem;
  uint64_t* s_cls_pm = s_rowlo + QES_MAX_TX;
  int32_t* s_cls_d = (int32_t*)(s_cls_pm + QES_MAX_CLASSES);
  double* s_chr = (double*)(smem + QES_CTX_FIXED_LDS);
  const int gs = G.ntx * G.nty * G.nbase;
  double* s_chi = s_chr + gs;
  double* s_ex = s_chi + gs;
  const int nex = P.nclass * G.ns * 4;
  uint32_t* s_lut = (uint32_t*)(s_ex + nex);
  const int nlut = (P.canon == QES_CANON_ROWS) ? (1 << P.lut_bits) : 0;
  RowCtx C;
  C.G = make_gv(G, s_rowlo);
  if (threadIdx.x < QES_MAX_CLASSES) {
#pragma unroll
    for (int c = 0; c < QES_MAX_CLASSES; ++c)
      if ((int)threadIdx.x == c) {
        s_cls_pm[c] = P.cls[c].pmask;
        s_cls_d[c] = P.cls[c].d;
      }
  }
  for (int i = threadIdx.x; i < gs; i += blockDim.x) {
    s_chr[i] = chi_re[i];
    s_chi[i] = CPLX ? chi_im[i] : 0.0;
  }
  for (int i = threadIdx.x; i < nex; i += blockDim.x) s_ex[i] = exval[i];
  for (int i = threadIdx.x; i < nlut; i += blockDim.x) s_lut[i] = lut[i];
  __syncthreads();
  C.L = L;
  C.nclass = P.nclass;
  C.n_gen = P.n_gen;
  C.uniform = P.uniform;
  C.direct = P.direct;
  C.top_test = P.top_test;
  C.s_cls_d = s_cls_d;
  C.s_cls_pm = s_cls_pm;
  C.groups = groups;
  C.n_diag = n_diag;
  C.s_chr = s_chr;
  C.s_chi = s_chi;
  C.s_ex = s_ex;
  C.s_lut = s_lut;
  C.s_fail = (uint64_t*)(((uintptr_t)(s_lut + nlut) + 7) & ~(uintptr_t)7);
  return C;
}

template <typename W>
__device__ __forceinline__ W load_rep(const QesLookup& L, int64_t a) {
  if (sizeof(W) == 4 && L.keys32) return (W)L.keys32[a];
  return (W)L.keys64[a];
}


}  // namespace qes
