/*
 * qes_oracle.c -- C restatement of the reference push-form symmetric matvec.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ and by the cpu_baseline leg of
 * bench.py as the CPU port of the reference kernel.  Never linked into the
 * product library.
 *
 * Restates, with the reference's own data flow:
 *   - apply_group_element_compiled  (Algebra/Symmetries/jit/symmetry_container_jit.py:63-118)
 *     with _apply_translation_prim (Algebra/Symmetries/translation.py:50-89),
 *     _apply_perm_prim (reflection.py:35-44), _apply_parity_prim (parity.py:42-66);
 *   - the representative of a state = first minimiser over the group in
 *     element order (symmetry_container_jit.py:234-246), looked up by binary
 *     search in the sorted representative list (symmetry_container.py:72-102)
 *     where the reference reads its dense repr_map (that map is 2^ns * 5 bytes,
 *     21.5 GB at ns=32, so the lookup is restated as the search it was built from);
 *   - sigma_operator_composition_int (Algebra/Operator/impl/operators_spin.py:2318-2352);
 *   - _apply_op_batch_compact_seq_jit (Algebra/Hilbert/matrix_builder.py:1347-1377):
 *         y[idx] += val * conj(chi[g*]) * N_idx / N_k * x[k].
 *
 * Parity pinned by tests/test_oracle_c.py against the NumPy oracle (itself
 * pinned against brute force) on the same inputs.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OP_TRANSLATION 1
#define OP_REFLECTION 2
#define OP_PARITY 3

typedef struct {
  int ns;
  int n_group;
  int max_ops;
  const int32_t* n_ops;    /* [n_group] */
  const int32_t* op_code;  /* [n_group*max_ops] */
  const int64_t* arg0;     /* [n_group*max_ops] */
  const int64_t* arg1;     /* [n_group*max_ops] */
  const int64_t* trans_perm; /* [n_trans*ns] */
  const int64_t* refl_perm;  /* [n_refl*ns] */
  const int32_t* parity_axis;
  const double* chi_re;
  const double* chi_im;
} ogroup;

static int64_t apply_perm(int64_t s, int ns, const int64_t* perm) {
  int64_t out = 0;
  for (int src = 0; src < ns; ++src)
    if ((s >> src) & 1) out |= (int64_t)1 << perm[src];
  return out;
}

static int64_t apply_element(const ogroup* G, int64_t s, int g) {
  int64_t cur = s;
  const int m = G->n_ops[g];
  for (int j = 0; j < m; ++j) {
    const int code = G->op_code[g * G->max_ops + j];
    const int64_t a0 = G->arg0[g * G->max_ops + j];
    const int64_t a1 = G->arg1[g * G->max_ops + j];
    if (code == OP_TRANSLATION) {
      for (int64_t p = 0; p < a1; ++p) cur = apply_perm(cur, G->ns, G->trans_perm + a0 * G->ns);
    } else if (code == OP_REFLECTION) {
      cur = apply_perm(cur, G->ns, G->refl_perm + a0 * G->ns);
    } else if (code == OP_PARITY) {
      int64_t mask = G->ns == 64 ? -1 : (((int64_t)1 << G->ns) - 1);
      cur = (~cur) & mask;
    }
  }
  return cur;
}

static int64_t bsearch_rep(const int64_t* reps, int64_t n, int64_t s) {
  int64_t lo = 0, hi = n - 1;
  while (lo <= hi) {
    int64_t mid = (lo + hi) / 2;
    if (reps[mid] == s) return mid;
    if (reps[mid] < s) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

/* ---------------- VM (operators_spin.py) ---------------- */
static int elem(char kind, int64_t* s, int ns, int site, double complex* c) {
  const int pos = ns - 1 - site;
  const int b = (int)((*s >> pos) & 1);
  switch (kind) {
    case 'x': *s ^= (int64_t)1 << pos; *c *= 0.5; return 1;
    case 'y': *c *= I * (2.0 * b - 1.0) * 0.5; *s ^= (int64_t)1 << pos; return 1;
    case 'z': *c *= (2.0 * b - 1.0) * 0.5; return 1;
    case 'p': if (b) { *c = 0.0; return 0; } *s ^= (int64_t)1 << pos; *c *= 0.5; return 1;
    case 'm': if (!b) { *c = 0.0; return 0; } *s ^= (int64_t)1 << pos; *c *= 0.5; return 1;
  }
  return 0;
}

static const char* code_kinds(int code, int* diag) {
  *diag = 0;
  switch (code) {
    case 11: return "x";
    case 12: return "y";
    case 13: *diag = 1; return "z";
    case 14: return "p";
    case 15: return "m";
    case 20: return "xx";
    case 21: return "yy";
    case 22: *diag = 1; return "zz";
    case 23: return "pp";
    case 24: return "mm";
    case 25: return "xy";
    case 26: return "yx";
    case 27: return "xz";
    case 28: return "zx";
    case 29: return "yz";
    case 30: return "zy";
    case 31: return "pm";
    case 32: return "mp";
  }
  return NULL;
}

/* outputs: states/values; returns count (<= T+1) */
static int vm_apply(int64_t state, int ns, const int32_t* codes, const int32_t* sites, const double* cre,
                    const double* cim, int T, int is_complex, int64_t* out_s, double complex* out_v) {
  int ptr = 0;
  double complex diag = 0.0;
  for (int k = 0; k < T; ++k) {
    int is_diag;
    const char* kinds = code_kinds(codes[k], &is_diag);
    if (!kinds) continue;
    int64_t s = state;
    double complex c = 1.0;
    const int st[2] = {sites[2 * k], sites[2 * k + 1]};
    const int nk = (int)strlen(kinds);
    for (int i = 0; i < nk; ++i)
      if (!elem(kinds[i], &s, ns, st[i], &c)) break;
    double complex v;
    if (is_complex) v = c * (cre[k] + I * (cim ? cim[k] : 0.0));
    else v = creal(c) * cre[k];
    if (cabs(v) < 1e-15) continue;
    if (is_diag) diag += v;
    else {
      out_s[ptr] = s;
      out_v[ptr] = v;
      ++ptr;
    }
  }
  if (cabs(diag) > 1e-15) {
    out_s[ptr] = state;
    out_v[ptr] = diag;
    ++ptr;
  }
  return ptr;
}

/*
 * Push-form matvec over the columns k in [k_begin, k_end).  x and y are
 * complex (interleaved) when x_complex, else real (then the matrix elements
 * must be real: imaginary parts are dropped like a float64 destination).
 * y is accumulated into (caller zeroes it).  Returns the number of gathers.
 */
int64_t oracle_push_matvec(const ogroup* G, const int64_t* reps, const double* norms, int64_t n,
                           const int32_t* codes, const int32_t* sites, const double* cre, const double* cim, int T,
                           int is_complex, const double* x, double* y, int x_complex, int64_t k_begin,
                           int64_t k_end) {
  int64_t* os = (int64_t*)malloc(sizeof(int64_t) * (T + 1));
  double complex* ov = (double complex*)malloc(sizeof(double complex) * (T + 1));
  int64_t gathers = 0;
  for (int64_t k = k_begin; k < k_end; ++k) {
    const int64_t state = reps[k];
    const double norm_k = norms[k];
    const double inv_norm_k = norm_k > 0.0 ? 1.0 / norm_k : 0.0;
    const int nout = vm_apply(state, G->ns, codes, sites, cre, cim, T, is_complex, os, ov);
    for (int i = 0; i < nout; ++i) {
      /* representative: first minimiser in element order */
      int64_t best = INT64_MAX;
      int gbest = -1;
      for (int g = 0; g < G->n_group; ++g) {
        int64_t t = apply_element(G, os[i], g);
        if (t < best) {
          best = t;
          gbest = g;
        }
      }
      const int64_t idx = bsearch_rep(reps, n, best);
      if (idx < 0) continue;
      const double complex val = ov[i];
      if (cabs(val) < 1e-15) continue;
      const double complex phase = G->chi_re[gbest] + I * G->chi_im[gbest];
      const double complex factor = val * conj(phase) * norms[idx] * inv_norm_k;
      ++gathers;
      if (x_complex) {
        const double complex xv = x[2 * k] + I * x[2 * k + 1];
        const double complex p = factor * xv;
        y[2 * idx] += creal(p);
        y[2 * idx + 1] += cimag(p);
      } else {
        y[idx] += creal(factor) * x[k];
      }
    }
  }
  free(os);
  free(ov);
  return gathers;
}

/*
 * Representatives and normalisations of the states of popcount n_up whose
 * colex rank lies in [rank0, rank0 + count): the scan of
 * scan_chunk_find_representatives / compute_normalization
 * (Algebra/Symmetries/jit/symmetry_container_jit.py:154-255) -- s is a
 * representative iff no group element maps it below itself, and its
 * normalisation is sqrt|sum over g with g(s) = s of conj(chi_g)|, kept
 * when that sum exceeds 1e-7 (symmetry_container.py:55).  Colex order of
 * the states of one popcount is their integer order, so the output is
 * sorted.  Each element acts as the composite of its primitive ops
 * (apply_element), tabulated once per call as a site permutation plus a
 * complement flag (elements here are translations, reflections and spin
 * flips, which compose to exactly that).  No intrinsic (parity-y) phases.
 * Returns the number of representatives written.
 */
static uint64_t binom_u64(int n, int k) {
  if (k < 0 || k > n) return 0;
  if (k > n - k) k = n - k;
  unsigned __int128 r = 1;
  for (int i = 1; i <= k; ++i) r = r * (unsigned)(n - k + i) / (unsigned)i;
  return (uint64_t)r;
}

int64_t oracle_scan_window(const ogroup* G, int n_up, uint64_t rank0, int64_t count, int64_t* reps_out,
                           double* norms_out) {
  const int ns = G->ns;
  const int ng = G->n_group;
  const int64_t mask = ns == 64 ? -1 : (((int64_t)1 << ns) - 1);
  int32_t* perm = (int32_t*)malloc(sizeof(int32_t) * (size_t)ng * ns);
  int64_t* flip = (int64_t*)malloc(sizeof(int64_t) * (size_t)ng);
  for (int g = 0; g < ng; ++g) {
    flip[g] = apply_element(G, 0, g);   /* 0 or all ones */
    for (int src = 0; src < ns; ++src) {
      const int64_t img = apply_element(G, (int64_t)1 << src, g) ^ flip[g];
      int dst = 0;
      while (dst < ns && !((img >> dst) & 1)) ++dst;
      perm[g * ns + src] = dst;
    }
  }
  /* unrank: the state with colex rank rank0 among the n_up-subsets of ns bits */
  uint64_t r = rank0;
  int64_t s = 0;
  for (int i = n_up, c = ns - 1; i >= 1; --i) {
    while (c >= i - 1 && binom_u64(c, i) > r) --c;
    s |= (int64_t)1 << c;
    r -= binom_u64(c, i);
    --c;
  }
  int64_t found = 0;
  for (int64_t t = 0; t < count; ++t) {
    int rep = 1;
    double sre = 0.0, sim = 0.0;
    for (int g = 0; g < ng && rep; ++g) {
      int64_t img = 0;
      const int32_t* pg = perm + g * ns;
      for (int src = 0; src < ns; ++src)
        if ((s >> src) & 1) img |= (int64_t)1 << pg[src];
      img = (img ^ flip[g]) & mask;
      if (img < s) rep = 0;
      else if (img == s) {
        sre += G->chi_re[g];
        sim -= G->chi_im[g];
      }
    }
    if (rep) {
      const double a = sqrt(sre * sre + sim * sim);
      if (a > 1e-7) {
        reps_out[found] = s;
        norms_out[found] = sqrt(a);
        ++found;
      }
    }
    /* next state of the same popcount in increasing order (Gosper) */
    if (t + 1 < count) {
      const int64_t lowbit = s & -s;
      const int64_t ripple = s + lowbit;
      s = (((ripple ^ s) >> 2) / lowbit) | ripple;
    }
  }
  free(perm);
  free(flip);
  return found;
}
