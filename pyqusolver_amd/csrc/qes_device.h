// Device-side building blocks for the symmetric-sector spin-1/2 path on gfx950.
//
// Group model.  Every symmetry group handled here is  G = T x B  where
//   B  = the "base" elements: all combinations of reflection (full bit
//        reversal) and spin flips (complement of all ns bits), applied first
//        (Algebra/Symmetries/symmetry_container.py:918-950, base + t order),
//   T  = translations T_x^a T_y^b on a lattice stored row-major with
//        row_bits = n_basis * Lx bits per row (translation.py:316-437):
//        T_x^a rotates every row left by a*cell_bits bits, T_y^b rotates the
//        whole ns-bit word left by b*row_bits bits.
// Element g has the reference index g = (a*nty + b)*nbase + B
// (symmetry_container.py:943-950 with translation powers from
// itertools.product over sorted directions, :1158-1172).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define QES_MAX_BASE 4
#define QES_MAX_BASE_OPS 4
#define QES_MAX_TX 64
#define QES_MAX_TERMS 1024

// base-op codes (one base element = up to QES_MAX_BASE_OPS ops in generator order)
#define QES_BOP_NONE 0
#define QES_BOP_REV 1      // reflection: bit i -> bit ns-1-i (reflection.py:143)
#define QES_BOP_FLIP 2     // parity x or z: complement, phase 1 (parity.py:42-54)
#define QES_BOP_FLIPY 3    // parity y: complement, phase +-i (parity.py:55-63)
#define QES_BOP_PERM 4     // site permutation (inversion.py:53-86): bit i -> bit perm[i]

struct QesGroup {
  int32_t ns;
  int32_t nbase;
  int32_t ntx, nty;
  int32_t row_bits;
  int32_t cell_bits;
  int32_t has_phase;     // some base element carries an intrinsic (+-i) phase
  int32_t n_up;          // U(1) popcount sector, -1 if none
  int32_t z2;            // popcount-parity sector (+1/-1), 0 if none
  int32_t has_bphase;    // translations carry boundary (flux) phases
  uint8_t base_ops[QES_MAX_BASE][QES_MAX_BASE_OPS];
  uint64_t mask;                      // (1<<ns)-1
  uint64_t rowlo[QES_MAX_TX];         // per x-power a: bits of each row below a*cell_bits
  // twisted translations (device arrays, set when has_bphase): bmask[a] =
  // bits that cross the boundary during a x steps, bmask[ntx + b] during b
  // y steps; bphase[(d * (ns + 1) + n) * 2 + {re, im}] = phase of n crossing
  // up spins along direction d (0: x, 1: y), multiplied per step
  // (symmetry_container_jit.py:91-98, translation.py:80-89)
  const uint64_t* bmask;
  const double* bphase;
  // QES_BOP_PERM: the permutation as npperm (mask, shift) pairs -- the bits
  // of mask move by shift (left if >= 0, right by -shift otherwise); device
  // array of 2 * npperm words, read with uniform indices
  const uint64_t* pperm;
  int32_t npperm;
};

// Register-sized view of a QesGroup used inside kernels.  Kernel arguments
// indexed with run-time indices would be copied to private (scratch) memory,
// so the per-x-power row masks live in LDS and the base ops are packed into
// one 64-bit word (3 bits per op).
struct GV {
  int ns, nbase, ntx, nty, row_bits, cell_bits, has_phase, n_up, z2, has_bphase, npperm;
  uint64_t bops;
  uint64_t mask;
  const uint64_t* rowlo;   // LDS copy, ntx entries
  const uint64_t* bmask;   // global, ntx + nty entries (has_bphase)
  const double* bphase;    // global, 2 (ns + 1) complex entries (has_bphase)
  const uint64_t* pperm;   // global, 2 npperm words (QES_BOP_PERM)
  __device__ __forceinline__ int bop(int B, int k) const { return (int)((bops >> ((B * 4 + k) * 3)) & 7u); }
};

// Stage the row masks in LDS (s_rowlo must hold QES_MAX_TX entries); every
// thread of the block must call this before __syncthreads().
__device__ __forceinline__ GV make_gv(const QesGroup& G, uint64_t* s_rowlo) {
  for (int i = threadIdx.x; i < G.ntx; i += blockDim.x) s_rowlo[i] = G.rowlo[i];
  GV v;
  v.ns = G.ns;
  v.nbase = G.nbase;
  v.ntx = G.ntx;
  v.nty = G.nty;
  v.row_bits = G.row_bits;
  v.cell_bits = G.cell_bits;
  v.has_phase = G.has_phase;
  v.n_up = G.n_up;
  v.z2 = G.z2;
  v.has_bphase = G.has_bphase;
  v.bmask = G.bmask;
  v.bphase = G.bphase;
  v.pperm = G.pperm;
  v.npperm = G.npperm;
  v.mask = G.mask;
  uint64_t b = 0;
#pragma unroll
  for (int B = 0; B < QES_MAX_BASE; ++B)
#pragma unroll
    for (int k = 0; k < QES_MAX_BASE_OPS; ++k) b |= ((uint64_t)G.base_ops[B][k] & 7u) << ((B * 4 + k) * 3);
  v.bops = b;
  v.rowlo = s_rowlo;
  return v;
}

// --------------------------------------------------------------------------
// word helpers
// --------------------------------------------------------------------------
template <typename W> __device__ __forceinline__ int popc(W v);
template <> __device__ __forceinline__ int popc<uint32_t>(uint32_t v) { return __popc(v); }
template <> __device__ __forceinline__ int popc<uint64_t>(uint64_t v) { return __popcll(v); }

template <typename W> __device__ __forceinline__ W brev(W v);
template <> __device__ __forceinline__ uint32_t brev<uint32_t>(uint32_t v) { return __builtin_bitreverse32(v); }
template <> __device__ __forceinline__ uint64_t brev<uint64_t>(uint64_t v) { return __builtin_bitreverse64(v); }

// rotate left by r within the low ns bits (0 <= r < ns)
template <typename W>
__device__ __forceinline__ W rotl_ns(W v, int r, int ns, W mask) {
  if (r == 0) return v;
  return ((v << r) | (v >> (ns - r))) & mask;
}

// rotate every row_bits-wide row left by r bits (0 <= r < row_bits);
// lo = bits of each row whose in-row position is < r
template <typename W>
__device__ __forceinline__ W rotl_rows(W v, int r, int row_bits, W lo, W mask) {
  if (r == 0) return v;
  W hi = mask ^ lo;
  return ((v << r) & hi) | ((v >> (row_bits - r)) & lo);
}

// apply base element B; returns new state and multiplies *ph by the intrinsic
// phase (only parity-y contributes, +i if the number of down spins is even)
template <typename W>
__device__ __forceinline__ W apply_base(const GV& G, int B, W s, int* ph_quarter) {
  W v = s;
  const int ns = G.ns;
  const W mask = (W)G.mask;
#pragma unroll
  for (int k = 0; k < QES_MAX_BASE_OPS; ++k) {
    int op = G.bop(B, k);
    if (op == QES_BOP_REV) {
      v = brev<W>(v) >> (int)(8 * sizeof(W) - ns);
    } else if (op == QES_BOP_FLIP) {
      v = (~v) & mask;
    } else if (op == QES_BOP_FLIPY) {
      int downs = ns - popc<W>(v);
      *ph_quarter += (downs & 1) ? 3 : 1;   // -i : +i  (quarter turns)
      v = (~v) & mask;
    } else if (op == QES_BOP_PERM) {
      W out = 0;
      for (int j = 0; j < G.npperm; ++j) {
        const W part = v & (W)G.pperm[2 * j];
        const int sh = (int)(int64_t)G.pperm[2 * j + 1];
        out |= sh >= 0 ? (W)(part << sh) : (W)(part >> -sh);
      }
      v = out;
    }
  }
  return v;
}

// base element B of a lean group (identity, then at most one spin flip; the
// host checks the group): no run-time switch over the operation codes
template <typename W, bool LEAN>
__device__ __forceinline__ W base_of(const GV& G, int B, W s) {
  if (LEAN) return B == 0 ? s : (W)((~s) & (W)G.mask);
  int dummy = 0;
  return apply_base<W>(G, B, s, &dummy);
}

// apply group element g = (a * nty + b) * nbase + B (reference order): base
// element B first, then a cell shifts along x and b row shifts along y
template <typename W>
__device__ __forceinline__ W apply_element(const GV& G, int g, W s) {
  const int B = g % G.nbase;
  const int ab = g / G.nbase;
  const int b = ab % G.nty, a = ab / G.nty;
  int dummy = 0;
  const W mask = (W)G.mask;
  W v = apply_base<W>(G, B, s, &dummy);
  v = rotl_rows<W>(v, a * G.cell_bits, G.row_bits, (W)G.rowlo[a], mask);
  return rotl_ns<W>(v, b * G.row_bits, G.ns, mask);
}

// Intrinsic phase of element g = (a * nty + b) * nbase + B acting on s: the
// parity-y quarter turns of its base element, then the boundary phase of
// every up spin that crosses during the a x steps and the b y steps
// (symmetry_container_jit.py:63-118; the base element acts first).
template <typename W>
__device__ __forceinline__ void element_phase(const GV& G, int g, W s, double* pre, double* pim) {
  const int B = g % G.nbase;
  const int ab = g / G.nbase;
  const int b = ab % G.nty, a = ab / G.nty;
  int q = 0;
  const W v = apply_base<W>(G, B, s, &q);
  double re = 1.0, im = 0.0;
  switch (q & 3) {
    case 1: re = 0.0; im = 1.0; break;
    case 2: re = -1.0; im = 0.0; break;
    case 3: re = 0.0; im = -1.0; break;
    default: break;
  }
  if (G.has_bphase) {
    const int nx = popc<W>(v & (W)G.bmask[a]);
    const int ny = popc<W>(v & (W)G.bmask[G.ntx + b]);
    const double* px = G.bphase + 2 * nx;
    const double* py = G.bphase + 2 * ((G.ns + 1) + ny);
    const double tr = px[0] * py[0] - px[1] * py[1], ti = px[0] * py[1] + px[1] * py[0];
    const double r2 = re * tr - im * ti, i2 = re * ti + im * tr;
    re = r2;
    im = i2;
  }
  *pre = re;
  *pim = im;
}

// --------------------------------------------------------------------------
// Canonicalisation: representative = min over G of g(s)
// (scan_chunk_find_representatives, jit/symmetry_container_jit.py:234-246).
// Returns the minimum, the number of minimising elements (|Stab(rep)|) and
// the reference index of the first minimiser.
// --------------------------------------------------------------------------
template <typename W>
struct Canon {
  W rep;
  int count;
  int g;
};

template <typename W>
__device__ __forceinline__ Canon<W> canonicalize(const GV& G, W s) {
  const int ns = G.ns;
  const W mask = (W)G.mask;
  W vb[QES_MAX_BASE];
#pragma unroll
  for (int B = 0; B < QES_MAX_BASE; ++B) {
    int dummy = 0;
    vb[B] = (B < G.nbase) ? apply_base<W>(G, B, s, &dummy) : (W)0;
  }
  Canon<W> c;
  c.rep = ~(W)0;
  c.count = 0;
  c.g = 0;
  for (int a = 0; a < G.ntx; ++a) {
    const int ra = a * G.cell_bits;
    const W lo = (W)G.rowlo[a];
    W ta[QES_MAX_BASE];
#pragma unroll
    for (int B = 0; B < QES_MAX_BASE; ++B) ta[B] = rotl_rows<W>(vb[B], ra, G.row_bits, lo, mask);
    for (int b = 0; b < G.nty; ++b) {
      const int rb = b * G.row_bits;
#pragma unroll
      for (int B = 0; B < QES_MAX_BASE; ++B) {
        if (B < G.nbase) {
          W v = rotl_ns<W>(ta[B], rb, ns, mask);
          int g = (a * G.nty + b) * G.nbase + B;
          if (v < c.rep) {
            c.rep = v;
            c.count = 1;
            c.g = g;
          } else if (v == c.rep) {
            c.count += 1;
          }
        }
      }
    }
  }
  return c;
}

// --------------------------------------------------------------------------
// Representative test used by the basis scan.  Returns false as soon as some
// g(s) < s.  Otherwise sums conj(chi_g)*phase_g over the stabiliser in the
// reference element order (compute_normalization, jit :154-174) and returns
// |sum| through *proj_abs2 components.
// --------------------------------------------------------------------------
template <typename W>
__device__ __forceinline__ bool rep_test(const GV& G, const double* __restrict__ chi_re,
                                         const double* __restrict__ chi_im, W s, double* pre,
                                         double* pim) {
  const int ns = G.ns;
  const W mask = (W)G.mask;
  W vb[QES_MAX_BASE];
  int phq[QES_MAX_BASE];
#pragma unroll
  for (int B = 0; B < QES_MAX_BASE; ++B) {
    phq[B] = 0;
    vb[B] = (B < G.nbase) ? apply_base<W>(G, B, s, &phq[B]) : (W)0;
  }
  double re = 0.0, im = 0.0;
  for (int a = 0; a < G.ntx; ++a) {
    const int ra = a * G.cell_bits;
    const W lo = (W)G.rowlo[a];
    W ta[QES_MAX_BASE];
#pragma unroll
    for (int B = 0; B < QES_MAX_BASE; ++B) ta[B] = rotl_rows<W>(vb[B], ra, G.row_bits, lo, mask);
    for (int b = 0; b < G.nty; ++b) {
      const int rb = b * G.row_bits;
#pragma unroll
      for (int B = 0; B < QES_MAX_BASE; ++B) {
        if (B < G.nbase) {
          W v = rotl_ns<W>(ta[B], rb, ns, mask);
          if (v < s) return false;
          if (v == s) {
            int g = (a * G.nty + b) * G.nbase + B;
            double cr = chi_re[g], ci = -chi_im[g];   // conj(chi)
            // times intrinsic phase i^q (exact: swaps/negations only)
            double tr, ti;
            switch (phq[B] & 3) {
              case 0: tr = cr; ti = ci; break;
              case 1: tr = -ci; ti = cr; break;
              case 2: tr = -cr; ti = -ci; break;
              default: tr = ci; ti = -cr; break;
            }
            if (G.has_bphase) {   // boundary phases of the crossing up spins
              const int nx = popc<W>(vb[B] & (W)G.bmask[a]);
              const int ny = popc<W>(vb[B] & (W)G.bmask[G.ntx + b]);
              const double* px = G.bphase + 2 * nx;
              const double* py = G.bphase + 2 * ((G.ns + 1) + ny);
              const double br = px[0] * py[0] - px[1] * py[1], bi = px[0] * py[1] + px[1] * py[0];
              const double ur = tr * br - ti * bi, ui = tr * bi + ti * br;
              tr = ur;
              ti = ui;
            }
            re += tr;
            im += ti;
          }
        }
      }
    }
  }
  *pre = re;
  *pim = im;
  return true;
}

// --------------------------------------------------------------------------
// Representative lookup: sorted keys + an exponent/mantissa bucket table.
// key(v): e = bit length of v; bucket = e<<mbits | (next mbits bits after the
// leading one), so bucket widths follow the roughly geometric density of
// representatives (which carry their longest zero run at the top).
// --------------------------------------------------------------------------
struct QesLookup {
  const uint64_t* keys64;
  const uint32_t* keys32;   // non-null when ns <= 32
  const uint32_t* bucket;   // ((e_max-e_min+1) << mbits) + 1 offsets
  int64_t n;
  int32_t mbits;
  int32_t e_min, e_max;     // bit-length range covered by the table
  int32_t ns;
  // open-addressing hash table (used when non-null): entry = key | index << ns,
  // empty slots are all ones; 2^hbits slots, linear probing
  const uint64_t* hash;
  int32_t hbits;
  int32_t packed;           // off-diagonal lookups probe keyc instead of keys32/keys64
  // packed keys (ns <= 56): key | (|Stab(rep)| - 1) << 56, sorted by key;
  // a hit on the packed array also yields the stabiliser size
  const uint64_t* keyc;
  // rank map (ns <= 48 with a U(1) filling n_up; used when non-null).  The
  // colex rank of v among the states of that filling is, for ns <= 32,
  // c = rbase[v >> 16] + rlo[v & 0xffff], and for 33..48 sites
  // c = rbase2[v >> 32] + rbase1[(k1 - k1min) * 65536 + (v >> 16 & 0xffff)]
  // + rlo[v & 0xffff] with k1 = popcount of the low 32 bits; rmap[c / 64]
  // holds the bit mask of the representatives among ranks 64*(c/64) ... +63
  // and the number of representatives of lower rank, so one 16-byte load
  // resolves a lookup.  The tables share one buffer, rtab, at fixed offsets
  // (one base pointer in the kernels): ns <= 32 rbase (u32[65536]) then rlo
  // (u16[65536]); 33..48 sites rbase2 (u64[65536]), rlo, then rbase1
  // (u32[rows * 65536]).
  const uint64_t* rmap;     // pairs {bits, count before}
  const uint64_t* rtab;     // the colex tables
  int32_t n_up;
  int32_t k1min;
  // sqrt(k) for k = 0..256 (stabiliser sizes; |G| <= 256): one cached load
  // per element instead of an FP64 square root
  const double* sqrtn;
  const double* norms;      // N_b of every representative (direct rank-map hits)
};

#define QES_KEY_BITS 56
#define QES_KEY_MASK ((1ull << QES_KEY_BITS) - 1ull)

// index of key v in the packed sorted array (bucket table narrows the range);
// returns -1 if absent, else the index and *count = |Stab|
__device__ __forceinline__ int64_t lookup_packed(const QesLookup& L, uint64_t v, int* count);

#define QES_HASH_EMPTY (~0ull)
__device__ __host__ __forceinline__ uint64_t qes_hash_slot(uint64_t key, int hbits) {
  return (key * 0x9E3779B97F4A7C15ull) >> (64 - hbits);
}

// first probe of a hash lookup (independent loads can be issued for several
// keys before any result is inspected)
__device__ __forceinline__ uint64_t hash_first(const QesLookup& L, uint64_t key, uint64_t* slot) {
  *slot = qes_hash_slot(key, L.hbits);
  return L.hash[*slot];
}

__device__ __forceinline__ int64_t hash_resolve(const QesLookup& L, uint64_t key, uint64_t slot, uint64_t e) {
  const uint64_t kmask = (L.ns >= 64) ? ~0ull : ((1ull << L.ns) - 1ull);
  const uint64_t hmask = (1ull << L.hbits) - 1ull;
  while (e != QES_HASH_EMPTY) {
    if ((e & kmask) == key) return (int64_t)(e >> L.ns);
    slot = (slot + 1) & hmask;
    e = L.hash[slot];
  }
  return -1;
}

__device__ __forceinline__ uint32_t bucket_of(uint64_t v, int mbits, int* e_out) {
  int e = 64 - __clzll(v);   // 0 for v == 0
  *e_out = e;
  uint64_t mant;
  if (e <= mbits) {
    mant = v;   // small values map one-to-one
  } else {
    mant = (v >> (e - 1 - mbits)) & ((1ull << mbits) - 1ull);
  }
  return ((uint32_t)e << mbits) | (uint32_t)mant;
}

__device__ __forceinline__ int64_t lookup_packed(const QesLookup& L, uint64_t v, int* count) {
  int e;
  uint32_t k = bucket_of(v, L.mbits, &e);
  if (e < L.e_min || e > L.e_max) return -1;
  k -= (uint32_t)L.e_min << L.mbits;
  uint32_t lo = L.bucket[k];
  uint32_t hi = L.bucket[k + 1];   // exclusive
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if ((L.keyc[mid] & QES_KEY_MASK) < v) lo = mid + 1; else hi = mid;
  }
  if (lo < (uint64_t)L.n) {
    const uint64_t e0 = L.keyc[lo];
    if ((e0 & QES_KEY_MASK) == v) {
      *count = (int)(e0 >> QES_KEY_BITS) + 1;
      return (int64_t)lo;
    }
  }
  return -1;
}

// offsets of the colex tables in rtab, in 8-byte words
#define QES_RTAB_NARROW_LO 32768ull            // after rbase u32[65536]
#define QES_RTAB_WIDE_LO 65536ull              // after rbase2 u64[65536]
#define QES_RTAB_WIDE_B1 (65536ull + 16384ull)  // after rlo u16[65536]

// colex rank of a state of the U(1) filling (the rank map's first phase:
// two or three loads from small cached tables, independent of each other);
// narrow: ns <= 32, wide: 33..48 sites
__device__ __forceinline__ uint64_t colex_rank_narrow(const QesLookup& L, uint64_t v) {
  const uint32_t u = (uint32_t)v;
  const uint32_t* rbase = reinterpret_cast<const uint32_t*>(L.rtab);
  const uint16_t* rlo = reinterpret_cast<const uint16_t*>(L.rtab + QES_RTAB_NARROW_LO);
  return rbase[u >> 16] + (uint32_t)rlo[u & 0xffffu];
}
__device__ __forceinline__ uint64_t colex_rank_wide(const QesLookup& L, uint64_t v) {
  const uint32_t lo32 = (uint32_t)v;
  const int row = __popc(lo32) - L.k1min;   // in 0..16 for a state of the filling
  const uint16_t* rlo = reinterpret_cast<const uint16_t*>(L.rtab + QES_RTAB_WIDE_LO);
  const uint32_t* rbase1 = reinterpret_cast<const uint32_t*>(L.rtab + QES_RTAB_WIDE_B1);
  return L.rtab[v >> 32] + (uint64_t)rbase1[(uint32_t)row * 65536u + (lo32 >> 16)] + (uint64_t)rlo[lo32 & 0xffffu];
}
__device__ __forceinline__ uint64_t colex_rank(const QesLookup& L, uint64_t v) {
  return L.ns <= 32 ? colex_rank_narrow(L, v) : colex_rank_wide(L, v);
}
// the same for kernels whose state word W follows the site count (32-bit
// states exactly for ns <= 32): no run-time branch
template <typename W>
__device__ __forceinline__ uint64_t colex_rank_w(const QesLookup& L, uint64_t v) {
  return sizeof(W) == 4 ? colex_rank_narrow(L, v) : colex_rank_wide(L, v);
}

// marker of the states of a rank-indexed gather vector that are no
// representative (a quiet NaN payload; a product writes any NaN of x with
// these bits as another NaN)
#define QES_RXS_EMPTY 0x7FFC5E17A11C0DE5ull

// second phase: the map word of rank c gives the index (-1: no representative)
__device__ __forceinline__ int64_t rank_resolve(ulonglong2 e, uint64_t c) {
  const uint32_t bit = (uint32_t)(c & 63u);
  if (!((e.x >> bit) & 1ull)) return -1;
  return (int64_t)e.y + (int64_t)__popcll(e.x & ((1ull << bit) - 1ull));
}

// rank-map lookup of a 32-bit state; -1 if it is not a representative
__device__ __forceinline__ int64_t lookup_rank(const QesLookup& L, uint64_t v) {
  if (L.ns < 64 && (v >> L.ns)) return -1;   // outside the ns-bit states (ns <= 48 here)
  if (__popcll(v) != L.n_up) return -1;
  const uint64_t c = colex_rank(L, v);
  const ulonglong2 e = reinterpret_cast<const ulonglong2*>(L.rmap)[c >> 6];
  const uint32_t bit = (uint32_t)(c & 63u);
  if (!((e.x >> bit) & 1ull)) return -1;
  return (int64_t)e.y + (int64_t)__popcll(e.x & ((1ull << bit) - 1ull));
}

template <typename W>
__device__ __forceinline__ int64_t lookup_rep(const QesLookup& L, W v) {
  if (L.rmap) return lookup_rank(L, (uint64_t)v);
  if (L.hash) {
    uint64_t slot;
    const uint64_t e0 = hash_first(L, (uint64_t)v, &slot);
    return hash_resolve(L, (uint64_t)v, slot, e0);
  }
  int e;
  uint32_t k = bucket_of((uint64_t)v, L.mbits, &e);
  if (e < L.e_min || e > L.e_max) return -1;
  k -= (uint32_t)L.e_min << L.mbits;
  int64_t lo = L.bucket[k];
  int64_t hi = L.bucket[k + 1];   // exclusive
  if (sizeof(W) == 4 && L.keys32 != nullptr) {
    const uint32_t* keys = L.keys32;
    while (lo < hi) {
      int64_t mid = (lo + hi) >> 1;
      uint32_t km = keys[mid];
      if (km < (uint32_t)v) lo = mid + 1; else hi = mid;
    }
    if (lo < L.n && keys[lo] == (uint32_t)v) return lo;
    return -1;
  } else {
    const uint64_t* keys = L.keys64;
    while (lo < hi) {
      int64_t mid = (lo + hi) >> 1;
      uint64_t km = keys[mid];
      if (km < (uint64_t)v) lo = mid + 1; else hi = mid;
    }
    if (lo < L.n && keys[lo] == (uint64_t)v) return lo;
    return -1;
  }
}

// --------------------------------------------------------------------------
// Compiled Hamiltonian terms.  One group = one flip mask and up to two read
// sites; its value for a state is val[(bit(p0)<<1) | bit(p1)].
// --------------------------------------------------------------------------
struct QesTermGroup {
  uint64_t mask;       // flip mask (0 for diagonal groups)
  int32_t p0, p1;      // bit positions read (-1: unused -> bit 0)
  int32_t coupling;    // coupling-group id (for per-coupling outputs)
  int32_t pad;
  double re[4];
  double im[4];
};

template <typename W>
__device__ __forceinline__ int term_bits(const QesTermGroup& t, W s) {
  int b0 = t.p0 >= 0 ? (int)((s >> t.p0) & 1) : 0;
  int b1 = t.p1 >= 0 ? (int)((s >> t.p1) & 1) : 0;
  return (b0 << 1) | b1;
}

// --------------------------------------------------------------------------
// Fast canonicalisers.  All return the same (rep, count) as canonicalize();
// the reported element g is *a* minimiser (not necessarily the first), which
// has the same character as the first one whenever the state belongs to the
// sector (characters of a one-dimensional irrep are constant on the coset
// of the stabiliser), i.e. whenever the lookup succeeds.
// --------------------------------------------------------------------------
#ifndef QES_BATCH
#define QES_BATCH 4
#endif
#define QES_CANON_BRUTE 0
#define QES_CANON_CYCLIC 1
#define QES_CANON_ROWS 2

// rotate left within ns bits; FULL: ns == word width
template <typename W, bool FULL>
__device__ __forceinline__ W rotc(W v, int r, int ns, W mask) {
  if (FULL) {
    if (sizeof(W) == 4) return (W)__builtin_rotateleft32((uint32_t)v, (uint32_t)r);
    return (W)__builtin_rotateleft64((uint64_t)v, (uint64_t)r);
  }
  return ((v << r) | (v >> (ns - r))) & mask;   // valid for 0 <= r < ns < width
}

// Cyclic group of rotations (one lattice row = the whole word, cell_bits 1)
// times the base elements.  The minimal rotation of v starts with the longest
// cyclic run of zeros of v, so only the starts of maximal runs are evaluated.
// run_masks: R_k[p] = 1 iff bits p, p-1, ..., p-k+1 (cyclic) of v are all 0.
template <typename W, bool FULL>
__device__ __forceinline__ int max_zero_run(W v, int ns, W mask, W* starts) {
  const W z = (~v) & mask;
  if (z == 0) {   // all ones: every rotation is equal, all are candidates
    *starts = mask;
    return 0;
  }
  if (z == mask) {
    *starts = mask;
    return ns;
  }
  W R[6];
  R[0] = z;
#pragma unroll
  for (int j = 1; j < 6; ++j) {
    const int sh = 1 << (j - 1);
    R[j] = (sh < ns) ? (R[j - 1] & rotc<W, FULL>(R[j - 1], sh, ns, mask)) : (W)0;
  }
  W cur = mask;
  int kk = 0;
#pragma unroll
  for (int j = 5; j >= 0; --j) {
    const int step = 1 << j;
    if (step < ns) {
      // FULL: the rotate takes its count modulo the word width itself
      const W cand = cur & rotc<W, FULL>(R[j], FULL ? kk : kk % ns, ns, mask);
      if (cand) {
        cur = cand;
        kk += step;
      }
    }
  }
  *starts = cur;
  return kk;
}

template <typename W, bool FULL, bool LEAN = false>
__device__ __forceinline__ Canon<W> canonicalize_cyclic(const GV& G, W s) {
  const int ns = G.ns;
  const W mask = (W)G.mask;
  constexpr int NB = LEAN ? 2 : QES_MAX_BASE;
  W vb[QES_MAX_BASE], st[QES_MAX_BASE];
  int run[QES_MAX_BASE];
  int best_run = -1;
#pragma unroll
  for (int B = 0; B < NB; ++B) {
    run[B] = -1;
    st[B] = 0;
    vb[B] = 0;
    if (B < G.nbase) {
      vb[B] = base_of<W, LEAN>(G, B, s);
      run[B] = max_zero_run<W, FULL>(vb[B], ns, mask, &st[B]);
      best_run = run[B] > best_run ? run[B] : best_run;
    }
  }
  Canon<W> c;
  c.rep = ~(W)0;
  c.count = 0;
  c.g = 0;
#pragma unroll
  for (int B = 0; B < NB; ++B) {
    if (B < G.nbase && run[B] == best_run) {
      W m = st[B];
      while (m) {
        const int p = (sizeof(W) == 4) ? __builtin_ctz((uint32_t)m) : __builtin_ctzll((uint64_t)m);
        m &= m - 1;
        const int r = (ns - 1 - p);   // rotation that brings bit p to the top
        const W v = rotc<W, FULL>(vb[B], r, ns, mask);
        const int g = r * G.nbase + B;
        if (v < c.rep) {
          c.rep = v;
          c.count = 1;
          c.g = g;
        } else if (v == c.rep) {
          c.count += 1;
        }
      }
    }
  }
  return c;
}

// Lattices with rows: the top row of the minimum must be the smallest
// row-rotation over all (base element, row that y-translations bring to the
// top).  lut[rv] = minval | (amask << 16) over x powers a (a*cell_bits bit
// rotation of the row).  Only candidates whose top row attains it are
// evaluated in full.
template <typename W, bool LEAN = false>
__device__ __forceinline__ Canon<W> canonicalize_rows(const GV& G, const uint32_t* __restrict__ lut, W s) {
  const int ns = G.ns;
  const W mask = (W)G.mask;
  const int Wr = G.row_bits;
  const int nrows = ns / Wr;
  const W rowmask = (W)((Wr == 64) ? ~0ull : ((1ull << Wr) - 1ull));
  const int q_lo = (G.nty > 1) ? 0 : nrows - 1;
  constexpr int NB = LEAN ? 2 : QES_MAX_BASE;
  if (LEAN) {
    // identity and at most one spin flip: a row of the flipped state is the
    // row of s complemented, so each row is extracted once and both LUT
    // entries read; the (base, row) pairs attaining the smallest top row are
    // collected while the minimum is found (bit 32 B + q, at most 32 rows:
    // the host checks) and then evaluated in the order of the general loop
    // (base element, then row), so the first minimiser is the same
    const uint32_t rm = (uint32_t)rowmask;
    const W flip = (~s) & mask;
    uint32_t mtop = 0xffffu;
    uint64_t cand = 0;
    for (int q = q_lo; q < nrows; ++q) {
      const uint32_t row = (uint32_t)(s >> (q * Wr)) & rm;
      const uint32_t e0 = lut[row] & 0xffffu;
      const uint64_t b0 = 1ull << q;
      cand = e0 < mtop ? b0 : (e0 == mtop ? (cand | b0) : cand);
      mtop = e0 < mtop ? e0 : mtop;
      if (G.nbase > 1) {
        const uint32_t e1 = lut[row ^ rm] & 0xffffu;
        const uint64_t b1 = 1ull << (32 + q);
        cand = e1 < mtop ? b1 : (e1 == mtop ? (cand | b1) : cand);
        mtop = e1 < mtop ? e1 : mtop;
      }
    }
    Canon<W> c;
    c.rep = ~(W)0;
    c.count = 0;
    c.g = 0;
    while (cand) {
      const int bi = __builtin_ctzll(cand);
      cand &= cand - 1;
      const int B = bi >> 5, q = bi & 31;
      const W vbB = B ? flip : s;
      uint32_t am = lut[(uint32_t)(vbB >> (q * Wr)) & rm] >> 16;
      const int b = nrows - 1 - q;
      while (am) {
        const int a = __builtin_ctz(am);
        am &= am - 1;
        const W t = rotl_rows<W>(vbB, a * G.cell_bits, Wr, (W)G.rowlo[a], mask);
        const W v = rotl_ns<W>(t, b * Wr, ns, mask);
        const int g = (a * G.nty + (G.nty > 1 ? b : 0)) * G.nbase + B;
        if (v < c.rep) {
          c.rep = v;
          c.count = 1;
          c.g = g;
        } else if (v == c.rep) {
          c.count += 1;
        }
      }
    }
    return c;
  }
  W vb[QES_MAX_BASE];
#pragma unroll
  for (int B = 0; B < NB; ++B) vb[B] = (B < G.nbase) ? base_of<W, LEAN>(G, B, s) : (W)0;
  uint32_t mtop = 0xffffu;
#pragma unroll
  for (int B = 0; B < NB; ++B) {
    if (B < G.nbase) {
      for (int q = q_lo; q < nrows; ++q) {
        const uint32_t e = lut[(uint32_t)((vb[B] >> (q * Wr)) & rowmask)] & 0xffffu;
        mtop = e < mtop ? e : mtop;
      }
    }
  }
  Canon<W> c;
  c.rep = ~(W)0;
  c.count = 0;
  c.g = 0;
#pragma unroll
  for (int B = 0; B < NB; ++B) {
    if (B < G.nbase) {
      for (int q = q_lo; q < nrows; ++q) {
        const uint32_t e = lut[(uint32_t)((vb[B] >> (q * Wr)) & rowmask)];
        if ((e & 0xffffu) != mtop) continue;
        uint32_t am = e >> 16;
        const int b = nrows - 1 - q;
        while (am) {
          const int a = __builtin_ctz(am);
          am &= am - 1;
          const W t = rotl_rows<W>(vb[B], a * G.cell_bits, Wr, (W)G.rowlo[a], mask);
          const W v = rotl_ns<W>(t, b * Wr, ns, mask);
          const int g = (a * G.nty + (G.nty > 1 ? b : 0)) * G.nbase + B;
          if (v < c.rep) {
            c.rep = v;
            c.count = 1;
            c.g = g;
          } else if (v == c.rep) {
            c.count += 1;
          }
        }
      }
    }
  }
  return c;
}

// Necessary condition of canonicalize_rows for s to be its own
// representative: the top row of s must be the smallest row rotation over
// all rows (that y-translations bring to the top) and base elements.  A
// state that fails it is not a representative; one that passes usually is.
#define QES_TOP_ROWS 8   // rows_top_ok reads at most this many rows (others: the test passes them)
template <typename W>
__device__ __forceinline__ bool rows_top_ok(const GV& G, const uint32_t* __restrict__ lut, W s) {
  const int ns = G.ns;
  const W mask = (W)G.mask;
  const int Wr = G.row_bits;
  const int nrows = ns / Wr;
  const W rowmask = (W)((Wr == 64) ? ~0ull : ((1ull << Wr) - 1ull));
  const int q_lo = (G.nty > 1) ? 0 : nrows - 1;
  const uint32_t top = (uint32_t)((s >> ((nrows - 1) * Wr)) & rowmask);
  // all lookups issued before any result is tested (no early exit: one wait
  // for the whole batch of LDS reads)
  uint32_t m = 0xffffu;
#pragma unroll
  for (int B = 0; B < QES_MAX_BASE; ++B) {
    if (B < G.nbase) {
      int dummy = 0;
      const W v = B == 0 ? s : apply_base<W>(G, B, s, &dummy);
#pragma unroll
      for (int q = 0; q < QES_TOP_ROWS; ++q)
        if (q >= q_lo && q < nrows) {
          const uint32_t e = lut[(uint32_t)((v >> (q * Wr)) & rowmask)] & 0xffffu;
          m = e < m ? e : m;
        }
    }
  }
  (void)mask;
  return m >= top;
}

// --------------------------------------------------------------------------
// Operator plan: exchange bonds grouped into "shift classes" plus generic
// term groups.  An exchange group flips bits p0 and p1 = (p0 - d) mod ns and
// is non-zero only when they differ; for a class d the active bonds of a
// state x are the set bits of (x ^ rot(x, d)) & pmask.
// Value table exval[((c*ns + p0)*2 + bit(p0))*2 + {re,im}].
// --------------------------------------------------------------------------
#define QES_MAX_CLASSES 8
struct QesExClass {
  uint64_t pmask;
  int32_t d;
  int32_t pad;
};
struct QesPlan {
  int32_t canon;     // QES_CANON_*
  int32_t nclass;
  int32_t n_gen;     // generic off-diagonal groups (after the diagonal ones)
  int32_t lut_bits;  // ROWS: row LUT has 1 << lut_bits entries
  int32_t uniform;   // 1: bonds in wavefront-uniform order, 0: each lane's active bonds
  int32_t direct;    // 1: look flipped states up directly before canonicalising
  int32_t top_test;  // ROWS: skip the direct lookup of states that fail rows_top_ok
  int32_t pad;
  QesExClass cls[QES_MAX_CLASSES];
};
