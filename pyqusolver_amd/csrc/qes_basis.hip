// Basis construction on the GPU: sorted representatives and normalisations.
//
// Replaces the 2^Ns scan of scan_chunk_find_representatives
// (Algebra/Symmetries/jit/symmetry_container_jit.py:204-255),
// fill_representatives (:193-199) and compute_normalization (:154-187).
// Instead of scanning every state into a dense 2^Ns map, only states that
// pass the global constraint are enumerated, in increasing order (unranking
// in the combinatorial number system + next-combination stepping for U(1)),
// each tested for "s == min_g g(s)" with an early exit, and the survivors
// are stream-compacted in order, so the list comes out sorted without a sort.
#include <algorithm>
#include <cstring>
#include <cmath>
#include <vector>

#include "qes_internal.h"

namespace qes {

__constant__ uint64_t c_binom[65][33];   // C(n, k) for k <= 32 (symmetric use)

static bool g_binom_ready = false;

static uint64_t host_binom(int n, int k) {
  if (k < 0 || k > n) return 0;
  if (k > n - k) k = n - k;
  unsigned __int128 r = 1;
  for (int i = 1; i <= k; ++i) r = r * (unsigned)(n - k + i) / (unsigned)i;
  return (uint64_t)r;
}

static int ensure_binom() {
  if (g_binom_ready) return 0;
  static uint64_t tab[65][33];
  for (int n = 0; n <= 64; ++n)
    for (int k = 0; k <= 32; ++k) tab[n][k] = host_binom(n, k);
  QES_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_binom), tab, sizeof(tab)));
  g_binom_ready = true;
  return 0;
}

__device__ __forceinline__ uint64_t dbinom(int n, int k) {
  if (k < 0 || k > n) return 0;
  if (k > n - k) k = n - k;
  return c_binom[n][k];
}

// k-th (0-based) ns-bit number with popcount m, in increasing numeric order
__device__ __forceinline__ uint64_t unrank_combination(uint64_t k, int ns, int m) {
  uint64_t s = 0;
  for (int p = ns - 1; p >= 0 && m > 0; --p) {
    uint64_t c = dbinom(p, m);   // numbers below 2^p with popcount m
    if (k >= c) {
      s |= 1ull << p;
      k -= c;
      --m;
    }
  }
  return s;
}

// next larger number with the same popcount (s != 0)
__device__ __forceinline__ uint64_t next_combination(uint64_t s) {
  uint64_t t = s | (s - 1);
  return (t + 1) | (((~t & (t + 1)) - 1) >> (__builtin_ctzll(s) + 1));
}

__device__ __forceinline__ bool passes_globals(const QesGroup& G, uint64_t s) {
  if (G.z2 != 0) {
    int par = (__popcll(s) & 1) ? -1 : 1;
    if (par != G.z2) return false;
  }
  return true;
}

// Thread t owns candidates [t*per, min((t+1)*per, total)).
template <typename W, bool WRITE>
__global__ void __launch_bounds__(256) basis_scan_kernel(QesGroup G, const double* __restrict__ chi_re,
                                                         const double* __restrict__ chi_im, uint64_t total,
                                                         uint32_t per, uint32_t* __restrict__ block_counts,
                                                         const int64_t* __restrict__ block_offsets,
                                                         uint64_t* __restrict__ reps, double* __restrict__ norms) {
  __shared__ uint32_t s_scan[256];
  __shared__ uint64_t s_rowlo[QES_MAX_TX];
  const GV gv = make_gv(G, s_rowlo);
  __syncthreads();
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t begin = t * per;
  uint64_t end = begin + per;
  if (end > total) end = total;
  uint32_t cnt = 0;
  const bool u1 = G.n_up >= 0;

  // pass 1 (count) and pass 2 (write) run the identical enumeration
  int64_t out = 0;
  if (WRITE) {
    // count first so the block scan can place this thread's outputs
    uint64_t s = 0;
    if (begin < end) s = u1 ? unrank_combination(begin, G.ns, G.n_up) : begin;
    for (uint64_t r = begin; r < end; ++r) {
      if (passes_globals(G, s)) {
        double pre, pim;
        if (rep_test<W>(gv, chi_re, chi_im, (W)s, &pre, &pim)) {
          if (hypot(pre, pim) > 1e-7) ++cnt;
        }
      }
      s = u1 ? (r + 1 < end ? next_combination(s) : s) : s + 1;
    }
    s_scan[threadIdx.x] = cnt;
    __syncthreads();
    // exclusive scan over the block (256 entries, simple Hillis-Steele)
    for (int off = 1; off < 256; off <<= 1) {
      uint32_t v = threadIdx.x >= (unsigned)off ? s_scan[threadIdx.x - off] : 0;
      __syncthreads();
      s_scan[threadIdx.x] += v;
      __syncthreads();
    }
    out = block_offsets[blockIdx.x] + (int64_t)(s_scan[threadIdx.x] - cnt);
    if (cnt == 0) return;
    s = u1 ? unrank_combination(begin, G.ns, G.n_up) : begin;
    for (uint64_t r = begin; r < end; ++r) {
      if (passes_globals(G, s)) {
        double pre, pim;
        if (rep_test<W>(gv, chi_re, chi_im, (W)s, &pre, &pim)) {
          double a = hypot(pre, pim);
          if (a > 1e-7) {
            reps[out] = s;
            norms[out] = sqrt(a);
            ++out;
          }
        }
      }
      s = u1 ? (r + 1 < end ? next_combination(s) : s) : s + 1;
    }
  } else {
    uint64_t s = 0;
    if (begin < end) s = u1 ? unrank_combination(begin, G.ns, G.n_up) : begin;
    for (uint64_t r = begin; r < end; ++r) {
      if (passes_globals(G, s)) {
        double pre, pim;
        if (rep_test<W>(gv, chi_re, chi_im, (W)s, &pre, &pim)) {
          if (hypot(pre, pim) > 1e-7) ++cnt;
        }
      }
      s = u1 ? (r + 1 < end ? next_combination(s) : s) : s + 1;
    }
    s_scan[threadIdx.x] = cnt;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
      if (threadIdx.x < (unsigned)off) s_scan[threadIdx.x] += s_scan[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) block_counts[blockIdx.x] = s_scan[0];
  }
}

__global__ void keys32_kernel(const uint64_t* __restrict__ reps, uint32_t* __restrict__ keys, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) keys[i] = (uint32_t)reps[i];
}

__device__ __forceinline__ int64_t rel_bucket(uint64_t v, int mbits, int e_min) {
  int e;
  uint32_t k = bucket_of(v, mbits, &e);
  return (int64_t)k - ((int64_t)e_min << mbits);
}

// bucket[k] = first index i with key(reps[i]) >= k (key is monotone in the
// representative, so a lower-bound search per bucket; k = nbucket gives n)
__global__ void bucket_kernel(const uint64_t* __restrict__ reps, int64_t n, int mbits, int e_min,
                              int64_t nbucket, uint32_t* __restrict__ bucket) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k > nbucket) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (rel_bucket(reps[mid], mbits, e_min) < k) lo = mid + 1; else hi = mid;
  }
  bucket[k] = (uint32_t)lo;
}

static int bitlen(uint64_t v) { return v ? 64 - __builtin_clzll(v) : 0; }

__global__ void keyc_kernel(const uint64_t* __restrict__ reps, const double* __restrict__ norms, int64_t n,
                            uint64_t* __restrict__ keyc);

int build_basis(Basis* B, hipStream_t st) {
  if (ensure_binom()) return -1;
  const QesGroup& G = B->group;
  uint64_t total;
  if (G.n_up >= 0) {
    if (G.n_up > G.ns) return fail("n_up > ns");
    total = host_binom(G.ns, G.n_up);
  } else {
    if (G.ns > 40) return fail("sector without U(1) limited to ns <= 40");
    total = 1ull << G.ns;
  }
  const uint32_t per = total > (1ull << 28) ? 256 : (total > (1ull << 20) ? 64 : 8);
  const uint64_t nthreads = (total + per - 1) / per;
  const uint64_t nblocks = (nthreads + 255) / 256;
  if (nblocks > 0x7fffffffull) return fail("basis enumeration too large");
  uint32_t* d_counts = nullptr;
  int64_t* d_offsets = nullptr;
  QES_HIP(hipMalloc(&d_counts, std::max<uint64_t>(nblocks, 1) * sizeof(uint32_t)));
  QES_HIP(hipMalloc(&d_offsets, std::max<uint64_t>(nblocks, 1) * sizeof(int64_t)));
  const bool w32 = G.ns <= 32;
  if (nblocks > 0) {
    if (w32)
      hipLaunchKernelGGL((basis_scan_kernel<uint32_t, false>), dim3((unsigned)nblocks), dim3(256), 0, st, G,
                         B->chi_re, B->chi_im, total, per, d_counts, nullptr, nullptr, nullptr);
    else
      hipLaunchKernelGGL((basis_scan_kernel<uint64_t, false>), dim3((unsigned)nblocks), dim3(256), 0, st, G,
                         B->chi_re, B->chi_im, total, per, d_counts, nullptr, nullptr, nullptr);
    QES_HIP(hipGetLastError());
  }
  std::vector<uint32_t> counts(nblocks);
  std::vector<int64_t> offsets(nblocks);
  if (nblocks) QES_HIP(hipMemcpyAsync(counts.data(), d_counts, nblocks * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  QES_HIP(hipStreamSynchronize(st));
  int64_t acc = 0;
  for (uint64_t b = 0; b < nblocks; ++b) {
    offsets[b] = acc;
    acc += counts[b];
  }
  B->n = acc;
  if (nblocks) QES_HIP(hipMemcpyAsync(d_offsets, offsets.data(), nblocks * sizeof(int64_t), hipMemcpyHostToDevice, st));
  const int64_t nalloc = std::max<int64_t>(acc, 1);
  QES_HIP(hipMalloc(&B->reps, nalloc * sizeof(uint64_t)));
  QES_HIP(hipMalloc(&B->norms, nalloc * sizeof(double)));
  if (nblocks > 0 && acc > 0) {
    if (w32)
      hipLaunchKernelGGL((basis_scan_kernel<uint32_t, true>), dim3((unsigned)nblocks), dim3(256), 0, st, G,
                         B->chi_re, B->chi_im, total, per, nullptr, d_offsets, B->reps, B->norms);
    else
      hipLaunchKernelGGL((basis_scan_kernel<uint64_t, true>), dim3((unsigned)nblocks), dim3(256), 0, st, G,
                         B->chi_re, B->chi_im, total, per, nullptr, d_offsets, B->reps, B->norms);
    QES_HIP(hipGetLastError());
  }
  QES_HIP(hipStreamSynchronize(st));
  QES_HIP(hipFree(d_counts));
  QES_HIP(hipFree(d_offsets));

  // lookup structures
  const int64_t n = B->n;
  {
    double tab[257];
    for (int k = 0; k <= 256; ++k) tab[k] = std::sqrt((double)k);
    QES_HIP(hipMalloc(&B->sqrtn, sizeof(tab)));
    QES_HIP(hipMemcpy(B->sqrtn, tab, sizeof(tab), hipMemcpyHostToDevice));
  }
  if (w32) {
    QES_HIP(hipMalloc(&B->keys32, nalloc * sizeof(uint32_t)));
    if (n) hipLaunchKernelGGL(keys32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, B->reps, B->keys32, n);
  }
  uint64_t lo_rep = 0, hi_rep = 0;
  if (n) {
    QES_HIP(hipMemcpyAsync(&lo_rep, B->reps, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    QES_HIP(hipMemcpyAsync(&hi_rep, B->reps + (n - 1), sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    QES_HIP(hipStreamSynchronize(st));
  }
  B->e_min = bitlen(lo_rep);
  B->e_max = bitlen(hi_rep);
  if (build_bucket(B, B->bucket_rows_log2, st)) return -1;
  if (G.ns <= QES_KEY_BITS && !G.has_phase && !G.has_bphase && n > 0) {
    QES_HIP(hipMalloc(&B->keyc, n * sizeof(uint64_t)));
    B->use_packed = G.ns > 32;   // 32-bit keys are narrower below 33 sites
    hipLaunchKernelGGL(keyc_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, B->reps, B->norms, n,
                       B->keyc);
    QES_HIP(hipGetLastError());
    QES_HIP(hipStreamSynchronize(st));
  }
  B->use_hash = false;
  // rank map by default where it applies (ns <= 48 with a U(1) filling and
  // C(ns, n_up)/4 bytes <= 4 GB: 150 MB at 32 sites, 2.3 GB at 36): one
  // 16-byte load per lookup instead of a bucket word and a key (L=32 matvec
  // 3.84 -> 2.88 ms); the sorted keys and bucket table stay for the other modes
  int ok = 0;
  if (n > 0 && build_rank_map(B, st, &ok) == 0) B->use_rank = ok != 0;
  else B->use_rank = false;   // no memory for it: the bucketed keys serve
  return 0;
}

// Bucket table over (bit length, next mbits bits) of the representatives,
// sized for about 2^rows_log2 representatives per bucket.
int build_bucket(Basis* B, int rows_log2, hipStream_t st) {
  const int64_t n = B->n;
  const int ne = B->e_max - B->e_min + 1;
  // rows_log2 < 0: more buckets than representatives (2^-rows_log2 per rep)
  const int64_t target = rows_log2 >= 0 ? std::max<int64_t>(n >> std::min(rows_log2, 40), 1)
                                        : std::max<int64_t>(n << std::min(-rows_log2, 4), 1);
  int mbits = 0;
  while (mbits < 24 && ((int64_t)ne << (mbits + 1)) <= target) ++mbits;   // bucket_of keys stay 32-bit
  if (B->bucket) (void)hipFree(B->bucket);
  B->bucket = nullptr;
  B->bucket_rows_log2 = rows_log2;
  B->mbits = mbits;
  B->nbucket = (int64_t)ne << mbits;
  QES_HIP(hipMalloc(&B->bucket, (B->nbucket + 1) * sizeof(uint32_t)));
  hipLaunchKernelGGL(bucket_kernel, dim3((unsigned)((B->nbucket + 1 + 255) / 256)), dim3(256), 0, st, B->reps, n,
                     mbits, B->e_min, B->nbucket, B->bucket);
  QES_HIP(hipGetLastError());
  QES_HIP(hipStreamSynchronize(st));
  return 0;
}

// packed key: representative | (|Stab| - 1) << 56 with |Stab| = N^2 (no
// intrinsic phases: the normalisation is sqrt of the stabiliser size)
__global__ void keyc_kernel(const uint64_t* __restrict__ reps, const double* __restrict__ norms, int64_t n,
                            uint64_t* __restrict__ keyc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double nn = norms[i];
  const uint64_t cnt = (uint64_t)llrint(nn * nn);
  keyc[i] = reps[i] | ((cnt - 1ull) << QES_KEY_BITS);
}

__global__ void hash_fill_kernel(uint64_t* __restrict__ t, uint64_t nslots) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nslots; i += (uint64_t)gridDim.x * blockDim.x)
    t[i] = QES_HASH_EMPTY;
}

// Insert every representative (linear probing, atomic claim of empty slots).
// Slot contents depend on insertion order only for colliding keys; every
// lookup returns the same index regardless.
__global__ void hash_insert_kernel(const uint64_t* __restrict__ reps, int64_t n, int ns, int hbits,
                                   unsigned long long* __restrict__ t) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t key = reps[i];
  const unsigned long long entry = key | ((unsigned long long)i << ns);
  const uint64_t hmask = (1ull << hbits) - 1ull;
  uint64_t slot = qes_hash_slot(key, hbits);
  for (uint64_t probe = 0; probe <= hmask; ++probe) {   // bounded: table is never full
    const unsigned long long prev = atomicCAS(&t[slot], (unsigned long long)QES_HASH_EMPTY, entry);
    if (prev == (unsigned long long)QES_HASH_EMPTY) return;
    slot = (slot + 1) & hmask;
  }
}

// Hash table representative -> index, load factor <= 0.6, when the index
// fits beside the key in 64 bits.
int build_hash(Basis* B, hipStream_t st) {
  const int64_t n = B->n;
  const int ns = B->group.ns;
  B->use_hash = false;
  if (n == 0 || ns >= 64 || (ns > 0 && (uint64_t)n > (1ull << (64 - ns)) - 1ull)) return 0;
  int hbits = 4;
  while ((double)n > 0.6 * (double)(1ull << hbits)) ++hbits;
  const uint64_t nslots = 1ull << hbits;
  QES_HIP(hipMalloc(&B->hash, nslots * sizeof(uint64_t)));
  B->hbits = hbits;
  hipLaunchKernelGGL(hash_fill_kernel, dim3((unsigned)std::min<uint64_t>((nslots + 255) / 256, 65536)), dim3(256), 0,
                     st, B->hash, nslots);
  hipLaunchKernelGGL(hash_insert_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, B->reps, n, ns, hbits,
                     (unsigned long long*)B->hash);
  QES_HIP(hipGetLastError());
  QES_HIP(hipStreamSynchronize(st));
  B->use_hash = true;
  return 0;
}

// ---------------------------------------------------------------------------
// Rank map: representatives of a U(1) filling k (ns <= 32) marked in a
// bitmap over the colex ranks of all states of that filling.  Integer order
// equals colex order for a fixed popcount, so the ranks of the sorted
// representatives are ascending and each bitmap word is filled from one
// contiguous range of them (no atomics).
// ---------------------------------------------------------------------------
__global__ void rank_of_reps_kernel(QesLookup L, const uint64_t* __restrict__ reps, int64_t n,
                                    uint64_t* __restrict__ rank) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  rank[i] = colex_rank(L, reps[i]);
}

// word w: bits of the ranks in [64w, 64w + 64) and the number of ranks below 64w
__global__ void rank_map_fill_kernel(const uint64_t* __restrict__ rank, int64_t n, int64_t nwords,
                                     uint64_t* __restrict__ rmap) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwords) return;
  const uint64_t base = (uint64_t)w * 64ull;
  int64_t lo = 0, hi = n;   // first index with rank >= base
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (rank[mid] < base) lo = mid + 1; else hi = mid;
  }
  uint64_t bits = 0;
  for (int64_t i = lo; i < n && rank[i] < base + 64ull; ++i) bits |= 1ull << (rank[i] - base);
  rmap[2 * w] = bits;
  rmap[2 * w + 1] = (uint64_t)lo;
}

int build_rank_map(Basis* B, hipStream_t st, int* eligible, double max_bytes) {
  const QesGroup& G = B->group;
  *eligible = 0;
  if (G.ns > 48 || G.n_up < 0 || G.n_up > G.ns) return 0;
  const int k = G.n_up;
  const uint64_t nstates = host_binom(G.ns, k);   // ranks 0 .. nstates-1
  const int64_t nwords = (int64_t)((nstates + 63) / 64);
  if ((double)nwords * 16.0 > max_bytes) return 0;
  *eligible = 1;
  if (B->rmap) return 0;
  // rlo[l] = #{l' < l : popc(l') = popc(l)}
  std::vector<uint16_t> rlo(65536);
  {
    uint32_t seen[17] = {0};
    for (uint32_t l = 0; l < 65536; ++l) rlo[l] = (uint16_t)seen[__builtin_popcount(l)]++;
  }
  // ns <= 32: rbase[h] = sum over h' < h of C(16, k - popc(h'))
  // 33..48:   rbase2[h] = sum over h' < h of C(32, k - popc(h'));
  //           rbase1[k1][h] = sum over h' < h of C(16, k1 - popc(h')), k1 = k1min..
  std::vector<uint32_t> rbase, rbase1;
  std::vector<uint64_t> rbase2;
  int k1min = 0, nrows = 0;
  if (G.ns <= 32) {
    rbase.resize(65536);
    uint64_t acc = 0;
    for (uint32_t h = 0; h < 65536; ++h) {
      rbase[h] = (uint32_t)acc;
      const int m = k - __builtin_popcount(h);
      acc += (m >= 0 && m <= 16) ? host_binom(16, m) : 0;
    }
  } else {
    rbase2.resize(65536);
    uint64_t acc = 0;
    for (uint32_t h = 0; h < 65536; ++h) {
      rbase2[h] = acc;
      const int m = k - __builtin_popcount(h);
      acc += (m >= 0 && m <= 32) ? host_binom(32, m) : 0;
    }
    k1min = std::max(0, k - (G.ns - 32));
    const int k1max = std::min(k, 32);
    nrows = k1max - k1min + 1;
    rbase1.resize((size_t)nrows * 65536);
    for (int r = 0; r < nrows; ++r) {
      const int k1 = k1min + r;
      uint64_t a = 0;
      for (uint32_t h = 0; h < 65536; ++h) {
        rbase1[(size_t)r * 65536 + h] = (uint32_t)a;
        const int m = k1 - __builtin_popcount(h);
        a += (m >= 0 && m <= 16) ? host_binom(16, m) : 0;
      }
    }
  }
  // one buffer, fixed offsets (QesLookup::rtab)
  std::vector<uint64_t> tab;
  if (G.ns <= 32) {
    tab.assign(QES_RTAB_NARROW_LO + 16384, 0);
    std::memcpy(tab.data(), rbase.data(), rbase.size() * sizeof(uint32_t));
    std::memcpy(tab.data() + QES_RTAB_NARROW_LO, rlo.data(), rlo.size() * sizeof(uint16_t));
  } else {
    tab.assign(QES_RTAB_WIDE_B1 + (rbase1.size() + 1) / 2, 0);
    std::memcpy(tab.data(), rbase2.data(), rbase2.size() * sizeof(uint64_t));
    std::memcpy(tab.data() + QES_RTAB_WIDE_LO, rlo.data(), rlo.size() * sizeof(uint16_t));
    std::memcpy(tab.data() + QES_RTAB_WIDE_B1, rbase1.data(), rbase1.size() * sizeof(uint32_t));
  }
  const int64_t n = B->n;
  uint64_t* d_rank = nullptr;
  hipError_t e = hipMalloc(&B->rtab, tab.size() * sizeof(uint64_t));
  if (e == hipSuccess) e = hipMemcpy(B->rtab, tab.data(), tab.size() * sizeof(uint64_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&B->rmap, (size_t)std::max<int64_t>(nwords, 1) * 16);
  if (e == hipSuccess) e = hipMalloc(&d_rank, (size_t)std::max<int64_t>(n, 1) * sizeof(uint64_t));
  if (e == hipSuccess) {
    B->k1min = k1min;
    QesLookup L = B->lookup();
    L.rtab = B->rtab;
    if (n) hipLaunchKernelGGL(rank_of_reps_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, L, B->reps, n,
                              d_rank);
    if (nwords) hipLaunchKernelGGL(rank_map_fill_kernel, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, st,
                                   d_rank, n, nwords, B->rmap);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (d_rank) (void)hipFree(d_rank);
  if (e != hipSuccess) {
    for (void* p : {(void*)B->rtab, (void*)B->rmap})
      if (p) (void)hipFree(p);
    B->rtab = nullptr;
    B->rmap = nullptr;
    return check_hip(e, "build_rank_map");
  }
  B->rmap_words = nwords;
  return 0;
}

// ---------------------------------------------------------------------------
// Rank-indexed gather vector: one slot of cw doubles per state of the U(1)
// filling, the empty marker everywhere until a product writes N_b x_b at the
// ranks of the representatives (rank_scatter_kernel).
// ---------------------------------------------------------------------------
__global__ void rxs_fill_kernel(uint64_t* __restrict__ p, int64_t words) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) p[i] = QES_RXS_EMPTY;
}

void free_rank_gather(Basis* B) {
  if (B->rxs) (void)hipFree(B->rxs);
  B->rxs = nullptr;
  B->rxs_states = 0;
  B->rxs_cw = 0;
}

int build_rank_gather(Basis* B, int cw, hipStream_t st) {
  if (!B->rmap) return fail("rank gather needs the basis's rank map (a U(1) filling and ns <= 48)");
  if (B->rxs && B->rxs_cw == cw) return 0;
  free_rank_gather(B);
  const int64_t states = B->rmap_words * 64;   // covers every colex rank of the filling
  const size_t bytes = (size_t)states * cw * sizeof(double);
  size_t free_b = 0, total_b = 0;
  QES_HIP(hipMemGetInfo(&free_b, &total_b));
  if (bytes > free_b / 2)
    return fail("rank gather vector would take " + std::to_string(bytes / 1e9) + " GB (" +
                std::to_string(free_b / 1e9) + " GB free; at most half is used)");
  QES_HIP(hipMalloc(&B->rxs, bytes));
  const int64_t words = states * cw;
  hipLaunchKernelGGL(rxs_fill_kernel, dim3((unsigned)std::min<int64_t>((words + 255) / 256, 256 * 64)), dim3(256), 0,
                     st, (uint64_t*)B->rxs, words);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    free_rank_gather(B);
    return check_hip(e, "build_rank_gather");
  }
  B->rxs_states = states;
  B->rxs_cw = cw;
  return 0;
}

template <typename W>
__global__ void map_states_kernel(QesGroup G, QesLookup L, const uint64_t* __restrict__ states, int64_t n,
                                  int64_t* __restrict__ idx, int32_t* __restrict__ gmin, int32_t* __restrict__ cnt) {
  __shared__ uint64_t s_rowlo[QES_MAX_TX];
  const GV gv = make_gv(G, s_rowlo);
  __syncthreads();
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t s = states[i];
  int64_t id = -1;
  int32_t g = -1, c = 0;
  bool ok = (s & ~G.mask) == 0 && passes_globals(G, s) && (G.n_up < 0 || __popcll(s) == G.n_up);
  if (ok) {
    Canon<W> cn = canonicalize<W>(gv, (W)s);
    id = lookup_rep<W>(L, cn.rep);
    g = cn.g;
    c = cn.count;
  }
  idx[i] = id;
  gmin[i] = id >= 0 ? g : -1;
  if (cnt) cnt[i] = id >= 0 ? c : 0;
}

// --------------------------------------------------------------------------
// Reduced <-> full space (no intrinsic phases).
//
// expand_to_full_state_jit (jit/symmetry_container_jit.py:261-315) pushes
//     psi[g r_k] += x_k conj(chi_g) / (N_k sqrt|G|)   for every g.
// The g with g r_k = s form a coset g_s Stab(r_k) on which chi is constant
// (chi is trivial on the stabiliser of a representative with N_k > 0), and
// |Stab| = N_k^2, so in pull form, one lane per full state s:
//     psi[s] = x_k chi(g*) N_k / sqrt|G|,  g* = first minimiser (g* s = r_k).
// transform_to_reduced_space (Algebra/Hilbert/hilbert_base.py:485-563):
//     red[k] = sum_{s -> k} conj(phase(s)) psi[s] / N_k,  phase(s) = chi(g*_s)
// in pull form over the orbit, one lane per representative: the orbit
// element g r_k has conj(chi(g*)) = chi(g), each distinct state is met |Stab|
// times, so red[k] = sum_g chi(g) psi[g r_k] / (N_k |Stab|).
// --------------------------------------------------------------------------
template <typename W, bool CPLX>
__global__ void expand_kernel(QesGroup G, QesLookup L, const double* __restrict__ chi_re,
                              const double* __restrict__ chi_im, const double* __restrict__ norms, double inv_sqrt_g,
                              const double* __restrict__ x, double* __restrict__ full, uint64_t nfull) {
  __shared__ uint64_t s_rowlo[QES_MAX_TX];
  const GV gv = make_gv(G, s_rowlo);
  __syncthreads();
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nfull) return;
  double vr = 0.0, vi = 0.0;
  const bool ok = passes_globals(G, s) && (G.n_up < 0 || __popcll(s) == G.n_up);
  if (ok) {
    const Canon<W> cn = canonicalize<W>(gv, (W)s);
    const int64_t k = lookup_rep<W>(L, cn.rep);
    if (k >= 0) {
      const double w = norms[k] * inv_sqrt_g;
      const double cr = chi_re[cn.g] * w, ci = chi_im[cn.g] * w;
      if (CPLX) {
        const double xr = x[2 * k], xi = x[2 * k + 1];
        vr = xr * cr - xi * ci;
        vi = xr * ci + xi * cr;
      } else {
        vr = x[k] * cr;
      }
    }
  }
  if (CPLX) {
    full[2 * s] = vr;
    full[2 * s + 1] = vi;
  } else {
    full[s] = vr;
  }
}

template <typename W, bool CPLX>
__global__ void reduce_kernel(QesGroup G, const double* __restrict__ chi_re, const double* __restrict__ chi_im,
                              const uint64_t* __restrict__ reps, const double* __restrict__ norms, int64_t n,
                              int gsize, const double* __restrict__ full, double* __restrict__ red) {
  __shared__ uint64_t s_rowlo[QES_MAX_TX];
  const GV gv = make_gv(G, s_rowlo);
  __syncthreads();
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const W r = (W)reps[k];
  double ar = 0.0, ai = 0.0;
  for (int g = 0; g < gsize; ++g) {
    const uint64_t s = (uint64_t)apply_element<W>(gv, g, r);
    const double cr = chi_re[g], ci = chi_im[g];
    if (CPLX) {
      const double pr = full[2 * s], pi = full[2 * s + 1];
      ar += cr * pr - ci * pi;
      ai += cr * pi + ci * pr;
    } else {
      ar += cr * full[s];
    }
  }
  const double nk = norms[k];
  const double inv = 1.0 / (nk * nk * nk);
  if (CPLX) {
    red[2 * k] = ar * inv;
    red[2 * k + 1] = ai * inv;
  } else {
    red[k] = ar * inv;
  }
}

int expand_reduce(const Basis* B, int reduce, const double* x_red, double* full, int dtype, hipStream_t st) {
  const QesGroup& G = B->group;
  if (G.has_phase || G.has_bphase)
    return fail("expansion with intrinsic phases (parity y, twisted translations) is not supported");
  if (G.ns > 36) return fail("full-space vectors are limited to 36 sites");
  bool chi_cplx = false;
  for (int g = 0; g < B->gsize; ++g) chi_cplx = chi_cplx || std::fabs(B->h_chi_im[g]) > 1e-12;
  if (dtype == 0 && chi_cplx) return fail("complex characters need complex128 vectors");
  const bool c = dtype == 1;
  const bool w32 = G.ns <= 32;
  if (!reduce) {
    const uint64_t nfull = 1ull << G.ns;
    const dim3 grid((unsigned)((nfull + 255) / 256));
    const double isg = 1.0 / std::sqrt((double)B->gsize);
#define QES_EX(W_, C_) \
  hipLaunchKernelGGL((expand_kernel<W_, C_>), grid, dim3(256), 0, st, G, B->lookup(), B->chi_re, B->chi_im, \
                     B->norms, isg, x_red, full, nfull)
    if (w32) { if (c) QES_EX(uint32_t, true); else QES_EX(uint32_t, false); }
    else { if (c) QES_EX(uint64_t, true); else QES_EX(uint64_t, false); }
#undef QES_EX
  } else {
    if (B->n == 0) return 0;
    const dim3 grid((unsigned)((B->n + 255) / 256));
#define QES_RD(W_, C_) \
  hipLaunchKernelGGL((reduce_kernel<W_, C_>), grid, dim3(256), 0, st, G, B->chi_re, B->chi_im, B->reps, B->norms, \
                     B->n, B->gsize, full, (double*)x_red)
    if (w32) { if (c) QES_RD(uint32_t, true); else QES_RD(uint32_t, false); }
    else { if (c) QES_RD(uint64_t, true); else QES_RD(uint64_t, false); }
#undef QES_RD
  }
  QES_HIP(hipGetLastError());
  return 0;
}

int map_states(const Basis* B, const uint64_t* d_states, int64_t n, int64_t* d_idx, int32_t* d_g, int32_t* d_cnt,
               hipStream_t st) {
  if (n <= 0) return 0;
  dim3 grid((unsigned)((n + 255) / 256));
  if (B->group.ns <= 32)
    hipLaunchKernelGGL(map_states_kernel<uint32_t>, grid, dim3(256), 0, st, B->group, B->lookup(), d_states, n,
                       d_idx, d_g, d_cnt);
  else
    hipLaunchKernelGGL(map_states_kernel<uint64_t>, grid, dim3(256), 0, st, B->group, B->lookup(), d_states, n,
                       d_idx, d_g, d_cnt);
  QES_HIP(hipGetLastError());
  return 0;
}

}  // namespace qes
