// extern "C" boundary of libqes_mi (declarations in include/qes_mi.h).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/qes_mi.h"
#include "qes_internal.h"
#include "qes_lanczos.h"

namespace qes {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }
int fail(const std::string& msg) {
  g_err = msg;
  return -1;
}
int check_hip(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  // a failed allocation also sets the runtime's last-error state; clear it
  // so that the next launch check of an unrelated call does not report it
  (void)hipGetLastError();
  return -1;
}

// ---------------------------------------------------------------------------
// Spin-1/2 instruction semantics (Algebra/Operator/impl/operators_spin.py):
// site i addresses bit ns-1-i (:153); spin value 0.5; bit 1 = up.
// ---------------------------------------------------------------------------
using cplx = std::complex<double>;

enum {
  SX_L = 11, SY_L = 12, SZ_L = 13, SP_L = 14, SM_L = 15,
  SX_C = 20, SY_C = 21, SZ_C = 22, SP_C = 23, SM_C = 24,
  SXY_C = 25, SYX_C = 26, SXZ_C = 27, SZX_C = 28, SYZ_C = 29, SZY_C = 30, SPM_C = 31, SMP_C = 32
};

// one elementary spin operator on one site: returns false if it annihilates
static bool elem(char kind, uint64_t& s, int ns, int site, cplx& c) {
  const int pos = ns - 1 - site;
  const int b = (int)((s >> pos) & 1ull);
  switch (kind) {
    case 'x': s ^= 1ull << pos; c *= 0.5; return true;                               // :147-157
    case 'y': c *= cplx(0.0, (2.0 * b - 1.0) * 0.5); s ^= 1ull << pos; return true;   // :269-281
    case 'z': c *= (2.0 * b - 1.0) * 0.5; return true;                                 // :439-447
    case 'p': if (b) { c = 0.0; return false; } s ^= 1ull << pos; c *= 0.5; return true;   // :627-645
    case 'm': if (!b) { c = 0.0; return false; } s ^= 1ull << pos; c *= 0.5; return true;  // :698-737
  }
  return false;
}

// sigma_operator_composition_single_op (:2165-2313); returns is_diagonal
static bool apply_code(int code, uint64_t& s, int ns, int s1, int s2, cplx& c, bool& ok) {
  c = 1.0;
  ok = true;
  const char* k = nullptr;
  bool diag = false;
  switch (code) {
    case SX_L: k = "x"; break;
    case SY_L: k = "y"; break;
    case SZ_L: k = "z"; diag = true; break;
    case SP_L: k = "p"; break;
    case SM_L: k = "m"; break;
    case SX_C: k = "xx"; break;
    case SY_C: k = "yy"; break;
    case SZ_C: k = "zz"; diag = true; break;
    case SP_C: k = "pp"; break;
    case SM_C: k = "mm"; break;
    case SXY_C: k = "xy"; break;
    case SYX_C: k = "yx"; break;
    case SXZ_C: k = "xz"; break;
    case SZX_C: k = "zx"; break;
    case SYZ_C: k = "yz"; break;
    case SZY_C: k = "zy"; break;
    case SPM_C: k = "pm"; break;
    case SMP_C: k = "mp"; break;
    default: ok = false; return false;
  }
  const int sites[2] = {s1, s2};
  const int nk = (int)strlen(k);
  for (int i = 0; i < nk; ++i) {
    if (sites[i] < 0 || sites[i] >= ns) { ok = false; return false; }
    if (!elem(k[i], s, ns, sites[i], c)) break;
  }
  return diag;
}

struct GroupKey {
  uint64_t mask;
  int p0, p1, cpl;
  bool operator<(const GroupKey& o) const {
    return std::tie(mask, p0, p1, cpl) < std::tie(o.mask, o.p0, o.p1, o.cpl);
  }
};

static int compile_terms(int ns, const int32_t* codes, const int32_t* sites, const double* cre, const double* cim,
                         int T, bool is_complex, const int32_t* coupling, std::vector<QesTermGroup>& diag,
                         std::vector<QesTermGroup>& off, int* ncpl, bool* has_imag) {
  std::map<GroupKey, QesTermGroup> dmap, omap;
  std::vector<GroupKey> dorder, oorder;
  *ncpl = 1;
  *has_imag = false;
  for (int t = 0; t < T; ++t) {
    const int s1 = sites[2 * t], s2 = sites[2 * t + 1];
    const cplx coef(cre[t], cim ? cim[t] : 0.0);
    const int cp = coupling ? coupling[t] : 0;
    if (cp < 0 || cp > 7) return fail("coupling id must be in 0..7");
    if (cp + 1 > *ncpl) *ncpl = cp + 1;
    const int p0 = (s1 >= 0 && s1 < ns) ? ns - 1 - s1 : -1;
    const int p1 = (s2 >= 0 && s2 < ns) ? ns - 1 - s2 : -1;
    if (p0 < 0) return fail("term " + std::to_string(t) + ": first site out of range");
    // enumerate the bits at the read sites
    double vre[4] = {0, 0, 0, 0}, vim[4] = {0, 0, 0, 0};
    uint64_t mask = 0;
    bool mask_set = false, is_diag = false;
    for (int k = 0; k < 4; ++k) {
      const int b0 = (k >> 1) & 1, b1 = k & 1;
      if (p1 < 0 && b1) continue;
      uint64_t s = ((uint64_t)b0 << p0) | (p1 >= 0 ? ((uint64_t)b1 << p1) : 0ull);
      if (p1 >= 0 && p1 == p0 && b0 != b1) continue;   // same site: bits must agree
      const uint64_t s_in = s;
      cplx c;
      bool ok;
      is_diag = apply_code(codes[t], s, ns, s1, s2, c, ok);
      if (!ok) return fail("term " + std::to_string(t) + ": unsupported code " + std::to_string(codes[t]));
      cplx val = is_complex ? c * coef : cplx(c.real() * coef.real(), 0.0);
      if (std::abs(val) < 1e-15) val = 0.0;   // operators_spin.py:2335
      if (val != 0.0) {
        const uint64_t m = s ^ s_in;
        if (mask_set && m != mask) return fail("inconsistent flip mask");
        mask = m;
        mask_set = true;
      }
      vre[k] = val.real();
      vim[k] = val.imag();
      if (p1 < 0) {   // single read site: duplicate over the unused bit
        vre[k | 1] = vre[k];
        vim[k | 1] = vim[k];
      }
    }
    if (!mask_set) continue;   // identically zero term
    const bool as_diag = is_diag || mask == 0;
    GroupKey key{as_diag ? 0ull : mask, p0, p1, cp};
    auto& mp = as_diag ? dmap : omap;
    auto& order = as_diag ? dorder : oorder;
    auto it = mp.find(key);
    if (it == mp.end()) {
      QesTermGroup g{};
      g.mask = key.mask;
      g.p0 = p0;
      g.p1 = p1;
      g.coupling = cp;
      for (int k = 0; k < 4; ++k) {
        g.re[k] = 0.0;
        g.im[k] = 0.0;
      }
      it = mp.emplace(key, g).first;
      order.push_back(key);
    }
    for (int k = 0; k < 4; ++k) {
      it->second.re[k] += vre[k];
      it->second.im[k] += vim[k];
    }
  }
  for (auto& k : dorder) diag.push_back(dmap[k]);
  for (auto& k : oorder) {
    QesTermGroup g = omap[k];
    bool nz = false;
    for (int q = 0; q < 4; ++q) {
      if (std::fabs(g.re[q]) < 1e-15) g.re[q] = 0.0;
      if (std::fabs(g.im[q]) < 1e-15) g.im[q] = 0.0;
      nz = nz || g.re[q] != 0.0 || g.im[q] != 0.0;
    }
    if (nz) off.push_back(g);
  }
  for (auto& g : diag)
    for (int q = 0; q < 4; ++q) *has_imag = *has_imag || g.im[q] != 0.0;
  for (auto& g : off)
    for (int q = 0; q < 4; ++q) *has_imag = *has_imag || g.im[q] != 0.0;
  return 0;
}

}  // namespace qes

using namespace qes;

struct qes_basis_s {
  Basis b;
};
struct qes_op_s {
  Op op;
};
struct qes_lanczos_s {
  Lanczos* L;
};

namespace qes {
Op* qes_op_internal(qes_op_t h) { return &h->op; }

int op_check_unused(const Op* op, const char* what) {
  if (op->users > 0)
    return fail(std::string(what) + ": the operator is in use by " + std::to_string(op->users) +
                " Lanczos handle(s) or batch(es), which hold its plan and assembled table; destroy them first");
  return 0;
}

int basis_check_unused(const Basis* b, const char* what) {
  if (b->users > 0)
    return fail(std::string(what) + ": the basis is in use by " + std::to_string(b->users) +
                " Lanczos handle(s) or batch(es), which hold its lookup tables; destroy them first");
  return 0;
}
}  // namespace qes

extern "C" {

const char* qes_last_error(void) { return g_err.c_str(); }
const char* qes_version(void) { return "qes_mi 0.1 gfx950"; }

int qes_basis_create(const qes_group_spec* spec, qes_basis_t* out, int64_t* n_out) {
  if (!spec || !out || !spec->chi_re || !spec->chi_im) return fail("null argument");
  if (spec->ns < 1 || spec->ns > 64) return fail("ns must be in 1..64");
  if (spec->nbase < 1 || spec->nbase > QES_MAX_BASE) return fail("nbase must be in 1..4");
  if (spec->ntx < 1 || spec->nty < 1 || spec->ntx > QES_MAX_TX) return fail("bad translation extents");
  const int gs = spec->ntx * spec->nty * spec->nbase;
  if (gs > 256) return fail("group order above 256 is not supported");
  if (spec->row_bits < 1 || spec->ns % spec->row_bits != 0 || (spec->nty > 1 && spec->row_bits * spec->nty != spec->ns))
    return fail("row_bits must divide ns (and row_bits * nty == ns with y translations)");
  if (spec->cell_bits < 1 || spec->ntx * spec->cell_bits > spec->row_bits) return fail("bad cell_bits");
  if (spec->n_up > spec->ns) return fail("n_up > ns");
  bool uses_perm = false;
  for (int b = 0; b < spec->nbase; ++b)
    for (int k = 0; k < QES_MAX_BASE_OPS; ++k) {
      if (spec->base_ops[b][k] > QES_BOP_PERM) return fail("unknown base-op code");
      uses_perm = uses_perm || spec->base_ops[b][k] == QES_BOP_PERM;
    }
  // the permutation as (mask, shift) pairs: bits that move by the same
  // distance share one mask
  std::vector<uint64_t> pairs;
  if (uses_perm) {
    if (!spec->perm) return fail("a QES_BASE_OP_PERM base element needs spec->perm");
    std::vector<int> seen(spec->ns, 0);
    for (int i = 0; i < spec->ns; ++i) {
      const int j = spec->perm[i];
      if (j < 0 || j >= spec->ns || seen[j]++) return fail("perm is not a permutation of 0..ns-1");
    }
    for (int i = 0; i < spec->ns; ++i)
      if (spec->perm[spec->perm[i]] != i) return fail("perm must be an involution (inversion)");
    std::vector<std::pair<int, uint64_t>> by_shift;
    for (int i = 0; i < spec->ns; ++i) {
      const int d = spec->perm[i] - i;
      auto it = std::find_if(by_shift.begin(), by_shift.end(), [&](const std::pair<int, uint64_t>& e) {
        return e.first == d;
      });
      if (it == by_shift.end()) by_shift.emplace_back(d, 1ull << i);
      else it->second |= 1ull << i;
    }
    for (const auto& e : by_shift) {
      pairs.push_back(e.second);
      pairs.push_back((uint64_t)(int64_t)e.first);
    }
  }
  auto* h = new qes_basis_s();
  Basis& B = h->b;
  QesGroup& G = B.group;
  memset(&G, 0, sizeof(G));
  G.ns = spec->ns;
  G.nbase = spec->nbase;
  G.ntx = spec->ntx;
  G.nty = spec->nty;
  G.row_bits = spec->row_bits;
  G.cell_bits = spec->cell_bits;
  G.n_up = spec->n_up;
  G.z2 = spec->z2;
  G.mask = spec->ns == 64 ? ~0ull : ((1ull << spec->ns) - 1ull);
  G.has_phase = 0;
  for (int b = 0; b < QES_MAX_BASE; ++b)
    for (int k = 0; k < QES_MAX_BASE_OPS; ++k) {
      G.base_ops[b][k] = b < spec->nbase ? spec->base_ops[b][k] : 0;
      if (G.base_ops[b][k] == QES_BOP_FLIPY) G.has_phase = 1;
    }
  for (int a = 0; a < spec->ntx; ++a) {
    const int r = a * spec->cell_bits;
    uint64_t lo = 0;
    for (int row = 0; row < spec->ns / spec->row_bits; ++row)
      lo |= (r ? ((1ull << r) - 1ull) : 0ull) << (row * spec->row_bits);
    G.rowlo[a] = lo;
  }
  if ((spec->cross_masks == nullptr) != (spec->bphase == nullptr))
    return fail("twisted translations need both cross_masks and bphase");
  B.gsize = gs;
  B.h_chi_re.assign(spec->chi_re, spec->chi_re + gs);
  B.h_chi_im.assign(spec->chi_im, spec->chi_im + gs);
  hipStream_t st = nullptr;
  int rc = 0;
  if (hipMalloc(&B.chi_re, gs * sizeof(double)) != hipSuccess || hipMalloc(&B.chi_im, gs * sizeof(double)) != hipSuccess)
    rc = fail("hipMalloc failed for characters");
  if (!rc && hipMemcpy(B.chi_re, spec->chi_re, gs * 8, hipMemcpyHostToDevice) != hipSuccess) rc = fail("copy chi");
  if (!rc && hipMemcpy(B.chi_im, spec->chi_im, gs * 8, hipMemcpyHostToDevice) != hipSuccess) rc = fail("copy chi");
  if (!rc && spec->cross_masks) {
    const size_t nm = (size_t)spec->ntx + spec->nty, np = 2 * ((size_t)spec->ns + 1) * 2;
    uint64_t* dm = nullptr;
    double* dp = nullptr;
    if (hipMalloc(&dm, nm * 8) != hipSuccess || hipMalloc(&dp, np * 8) != hipSuccess ||
        hipMemcpy(dm, spec->cross_masks, nm * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dp, spec->bphase, np * 8, hipMemcpyHostToDevice) != hipSuccess) {
      if (dm) (void)hipFree(dm);
      if (dp) (void)hipFree(dp);
      rc = fail("device copy of the boundary phases failed");
    } else {
      G.bmask = dm;
      G.bphase = dp;
      G.has_bphase = 1;
    }
  }
  if (!rc && !pairs.empty()) {
    uint64_t* dp = nullptr;
    if (hipMalloc(&dp, pairs.size() * 8) != hipSuccess ||
        hipMemcpy(dp, pairs.data(), pairs.size() * 8, hipMemcpyHostToDevice) != hipSuccess) {
      if (dp) (void)hipFree(dp);
      rc = fail("device copy of the site permutation failed");
    } else {
      G.pperm = dp;
      G.npperm = (int32_t)(pairs.size() / 2);
    }
  }
  if (!rc) rc = build_basis(&B, st);
  if (rc) {
    qes_basis_destroy(h);
    return -1;
  }
  *out = h;
  if (n_out) *n_out = B.n;
  return 0;
}

int qes_basis_get(qes_basis_t h, uint64_t* reps_host, double* norms_host) {
  if (!h) return fail("null basis");
  const int64_t n = h->b.n;
  if (n == 0) return 0;
  if (reps_host) QES_HIP(hipMemcpy(reps_host, h->b.reps, n * 8, hipMemcpyDeviceToHost));
  if (norms_host) QES_HIP(hipMemcpy(norms_host, h->b.norms, n * 8, hipMemcpyDeviceToHost));
  return 0;
}

int qes_basis_device_ptrs(qes_basis_t h, const uint64_t** reps_dev, const double** norms_dev) {
  if (!h) return fail("null basis");
  if (reps_dev) *reps_dev = h->b.reps;
  if (norms_dev) *norms_dev = h->b.norms;
  return 0;
}

int qes_basis_map_states(qes_basis_t h, const uint64_t* states_host, int64_t n, int64_t* idx_out, int32_t* g_out) {
  if (!h) return fail("null basis");
  if (n <= 0) return 0;
  if (!states_host || !idx_out) return fail("null argument");
  uint64_t* ds = nullptr;
  int64_t* di = nullptr;
  int32_t* dg = nullptr;
  hipError_t e = hipMalloc(&ds, n * 8);
  if (e == hipSuccess) e = hipMalloc(&di, n * 8);
  if (e == hipSuccess) e = hipMalloc(&dg, n * 4);
  if (e == hipSuccess) e = hipMemcpy(ds, states_host, n * 8, hipMemcpyHostToDevice);
  int rc = e == hipSuccess ? map_states(&h->b, ds, n, di, dg, nullptr, nullptr) : check_hip(e, "qes_basis_map_states");
  if (!rc) {
    e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(idx_out, di, n * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess && g_out) e = hipMemcpy(g_out, dg, n * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = check_hip(e, "qes_basis_map_states");
  }
  (void)hipFree(ds);
  (void)hipFree(di);
  (void)hipFree(dg);
  return rc;
}

int qes_basis_destroy(qes_basis_t h) {
  if (!h) return 0;
  if (basis_check_unused(&h->b, "qes_basis_destroy")) return -1;
  Basis& B = h->b;
  if (B.group.bmask) (void)hipFree((void*)B.group.bmask);
  if (B.group.bphase) (void)hipFree((void*)B.group.bphase);
  if (B.group.pperm) (void)hipFree((void*)B.group.pperm);
  if (B.chi_re) (void)hipFree(B.chi_re);
  if (B.chi_im) (void)hipFree(B.chi_im);
  if (B.reps) (void)hipFree(B.reps);
  if (B.keys32) (void)hipFree(B.keys32);
  if (B.norms) (void)hipFree(B.norms);
  if (B.bucket) (void)hipFree(B.bucket);
  if (B.hash) (void)hipFree(B.hash);
  if (B.keyc) (void)hipFree(B.keyc);
  if (B.rmap) (void)hipFree(B.rmap);
  free_rank_gather(&B);
  if (B.rtab) (void)hipFree(B.rtab);
  if (B.sqrtn) (void)hipFree(B.sqrtn);
  delete h;
  return 0;
}

int qes_op_create(qes_basis_t bh, const int32_t* codes, const int32_t* sites, const double* coeff_re,
                  const double* coeff_im, int32_t n_terms, int32_t is_complex, const int32_t* coupling,
                  qes_op_t* out) {
  if (!bh || !out) return fail("null argument");
  if (n_terms < 0 || (n_terms > 0 && (!codes || !sites || !coeff_re))) return fail("bad term arrays");
  std::vector<QesTermGroup> diag, off;
  int ncpl = 1;
  bool has_imag = false;
  if (compile_terms(bh->b.group.ns, codes, sites, coeff_re, coeff_im, n_terms, is_complex != 0, coupling, diag, off,
                    &ncpl, &has_imag))
    return -1;
  if (diag.size() + off.size() > QES_MAX_TERMS) return fail("too many term groups");
  {
    // Reflection combined with a complex (generic-k) translation character is
    // not a symmetry sector (non-Hermitian block), nor is parity y.  The
    // boundary phases of twisted translations enter the matrix elements here
    // (phase of the minimising element on each target state; the reference's
    // kernel leaves them out of the elements, which makes those blocks
    // non-Hermitian).
    const QesGroup& G = bh->b.group;
    bool has_rev = false, chi_cplx = false;
    for (int b = 0; b < G.nbase; ++b)
      for (int k = 0; k < QES_MAX_BASE_OPS; ++k)
        has_rev = has_rev || G.base_ops[b][k] == QES_BOP_REV || G.base_ops[b][k] == QES_BOP_PERM;
    for (int g = 0; g < bh->b.gsize; ++g) chi_cplx = chi_cplx || std::fabs(bh->b.h_chi_im[g]) > 1e-12;
    if (G.has_phase)
      return fail("parity-y sectors are not supported for matvec: the reference's +-i phase does not make the "
                  "elements a representation (non-Hermitian block)");
    if (has_rev && chi_cplx)
      return fail("reflection or inversion at a generic momentum is not a symmetry sector (non-Hermitian reference "
                  "block)");
  }
  auto* h = new qes_op_s();
  Op& op = h->op;
  op.basis = &bh->b;
  op.n_diag = (int)diag.size();
  op.n_off = (int)off.size();
  op.n_coupling = ncpl;
  op.is_complex = has_imag;
  // complex vectors for complex characters and for intrinsic phases
  op.chi_complex = bh->b.group.has_phase || bh->b.group.has_bphase;
  for (int g = 0; g < bh->b.gsize; ++g)
    if (std::fabs(bh->b.h_chi_im[g]) > 1e-12) op.chi_complex = true;
  op.h_groups = diag;
  op.h_groups.insert(op.h_groups.end(), off.begin(), off.end());
  const size_t bytes = std::max<size_t>(op.h_groups.size(), 1) * sizeof(QesTermGroup);
  if (hipMalloc(&op.d_groups, bytes) != hipSuccess) {
    delete h;
    return fail("hipMalloc failed for term groups");
  }
  if (!op.h_groups.empty() &&
      hipMemcpy(op.d_groups, op.h_groups.data(), op.h_groups.size() * sizeof(QesTermGroup), hipMemcpyHostToDevice) !=
          hipSuccess) {
    (void)hipFree(op.d_groups);
    delete h;
    return fail("copy of term groups failed");
  }
  if (build_plan(&op, -1)) {
    (void)hipFree(op.d_groups);
    free_plan(&op);
    delete h;
    return -1;
  }
  *out = h;
  return 0;
}

int qes_op_set_canonicalizer(qes_op_t h, int32_t kind) {
  if (!h) return fail("null operator");
  if (kind < -1 || kind > 2) return fail("kind must be -1 (auto), 0 (brute), 1 (cyclic) or 2 (rows)");
  if (op_check_unused(&h->op, "qes_op_set_canonicalizer")) return -1;
  return build_plan(&h->op, kind);
}

int qes_op_set_option(qes_op_t h, const char* name, int64_t value) {
  if (!h || !name) return fail("null argument");
  Op& op = h->op;
  const std::string n(name);
  if (op_check_unused(&op, "qes_op_set_option")) return -1;
  if (n == "bond_order_uniform") op.bond_order_uniform = value < 0 ? -1 : (value ? 1 : 0);
  else if (n == "direct_lookup") op.direct_lookup = value < 0 ? -1 : (value ? 1 : 0);
  else if (n == "xcd_remap") op.xcd_remap = value ? 1 : 0;
  else if (n == "top_row_test") op.top_row_test = value ? 1 : 0;
  else if (n == "direct_batch") {
    if (value != -1 && value != 1 && value != 2 && value != 4 && value != 8)
      return fail("direct_batch must be -1 (auto), 1, 2, 4 or 8 (8: rank-gather products of 64-bit states)");
    op.direct_batch = (int)value;
  } else if (n == "sell_grid") {
    if (value < 0 || value > 64) return fail("sell_grid must be 0 (a block per 256 rows) or 1..64 blocks per CU");
    op.sell_grid = (int)value;
    return 0;   // launch shape only: the plan and the table stay
  } else if (n == "sell_lds") {
    if (value < 0 || value > 65536) return fail("sell_lds must be 0..65536 bytes");
    op.sell_lds = (int)value;
    return 0;
  } else if (n == "lean") {
    if (value != 0 && value != 1) return fail("lean must be 0 or 1");
    op.lean = (int)value;
    return 0;   // kernel choice only
  } else if (n == "rank_gather") {
    if (value != 0 && value != 1) return fail("rank_gather must be 0 or 1");
    if (basis_check_unused(op.basis, "qes_op_set_option(rank_gather)")) return -1;
    if (value) {
      if (build_rank_gather(op.basis, (op.is_complex || op.chi_complex) ? 2 : 1, nullptr)) return -1;
    } else {
      free_rank_gather(op.basis);   // other operators of the basis fall back to the dense gathers
    }
    op.rank_gather = (int)value;
    return 0;
  } else if (n == "asm_panel_cols") {
    if (value < -1) return fail("asm_panel_cols must be -1 (auto), 0 (off) or a column count");
    op.asm_panel_cols = value;
  } else
    return fail("unknown option " + n +
                " (bond_order_uniform, direct_lookup, xcd_remap, top_row_test, direct_batch, asm_panel_cols, sell_grid, "
                "sell_lds, rank_gather, lean)");
  return build_plan(&op, op.plan.canon == pick_canon_public(op.basis->group) ? -1 : op.plan.canon);
}

int qes_projected_matvec(qes_op_t op, qes_basis_t out, const void* x_dev, void* y_dev, void* stream) {
  if (!op || !out || !x_dev || !y_dev) return fail("null argument");
  return projected_matvec(&op->op, &out->b, (const double*)x_dev, (double*)y_dev, (hipStream_t)stream);
}

int qes_op_assemble(qes_op_t h, int32_t mode, double max_gb, void* stream) {
  if (!h) return fail("null operator");
  if (mode < 0 || mode > 3)
    return fail("mode must be 0 (free), 1 (auto format), 2 (plain values) or 3 (coupling-tagged dictionary)");
  if (mode == 0 && h->op.asm_fmt < 0) return 0;   // nothing to free
  if (op_check_unused(&h->op, "qes_op_assemble")) return -1;
  if (max_gb > 0) h->op.asm_max_bytes = max_gb * 1e9;
  const int want_dict = mode == 1 ? 1 : (mode == 3 ? 2 : 0);
  return op_assemble(&h->op, mode != 0, want_dict, (hipStream_t)stream);
}

int qes_op_matvec_scaled(qes_op_t h, const double* scale, int32_t nscale, const void* x_dev, void* y_dev,
                         int32_t dtype, void* stream) {
  if (!h || !scale || !x_dev || !y_dev) return fail("null argument");
  if (dtype != 0 && dtype != 1) return fail("dtype must be 0 (float64) or 1 (complex128)");
  return assembled_matvec_scaled(&h->op, scale, nscale, (const double*)x_dev, (double*)y_dev, dtype,
                                 (hipStream_t)stream);
}

int qes_op_coupling_dots(qes_op_t h, const void* v_dev, const void* g_dev, int32_t dtype, double* out,
                         void* stream) {
  if (!h || !v_dev || !g_dev || !out) return fail("null argument");
  if (dtype != 0 && dtype != 1) return fail("dtype must be 0 (float64) or 1 (complex128)");
  return assembled_coupling_dots(&h->op, (const double*)v_dev, (const double*)g_dev, dtype, out,
                                 (hipStream_t)stream);
}

int qes_op_coupling_dots_noalloc(qes_op_t h, const void* v_dev, const void* g_dev, int32_t dtype, double* out,
                                 void* stream) {
  if (!h || !v_dev || !g_dev || !out) return fail("null argument");
  if (dtype != 0 && dtype != 1) return fail("dtype must be 0 (float64) or 1 (complex128)");
  return assembled_coupling_dots_noalloc(&h->op, (const double*)v_dev, (const double*)g_dev, dtype, out,
                                         (hipStream_t)stream);
}

int qes_op_assembly_info(qes_op_t h, int64_t* entries, int64_t* bytes, int32_t* format, int32_t* n_values,
                         int32_t* tagged) {
  if (!h) return fail("null operator");
  const Op& op = h->op;
  const bool on = op.asm_fmt >= 0;
  if (entries) *entries = on ? op.asm_entries : 0;
  if (bytes) *bytes = on ? op.asm_bytes : 0;
  if (format) *format = op.asm_fmt;
  if (n_values) *n_values = op.asm_ndict;
  if (tagged) *tagged = on ? op.asm_tagged : 0;
  return 0;
}

int qes_op_panel_info(qes_op_t h, int32_t* npanel, int64_t* cols, int64_t* entries, int64_t* bytes) {
  if (!h) return fail("null operator");
  const Op& op = h->op;
  if (npanel) *npanel = op.asm_npanel;
  if (cols) *cols = op.asm_pw;
  if (entries) *entries = op.asm_pentries;
  if (bytes) *bytes = op.asm_pbytes;
  return 0;
}

int qes_op_table_coo(qes_op_t h, int32_t* rows_dev, int32_t* cols_dev, double* vals_dev, int64_t capacity,
                     int64_t* count, void* stream) {
  if (!h) return fail("null operator");
  const Op& op = h->op;
  const int64_t need = op.asm_fmt < 0 ? 0 : op.asm_entries + (op.asm_fmt == 0 ? 0 : op.basis->n);
  if (count) *count = need;
  if (!rows_dev) return op.asm_fmt < 0 ? fail("operator is not assembled") : 0;   // size query
  return assembled_coo(&op, rows_dev, cols_dev, vals_dev, capacity, (hipStream_t)stream);
}

int qes_op_get_plan(qes_op_t h, int32_t* canon, int32_t* n_classes, int32_t* n_generic) {
  if (!h) return fail("null operator");
  if (canon) *canon = h->op.plan.canon;
  if (n_classes) *n_classes = h->op.plan.nclass;
  if (n_generic) *n_generic = h->op.plan.n_gen;
  return 0;
}

int qes_basis_set_lookup(qes_basis_t b, int32_t mode) {
  if (!b) return fail("null basis");
  if (basis_check_unused(&b->b, "qes_basis_set_lookup")) return -1;
  if (mode < 0 || mode > 3)
    return fail("mode must be 0 (sorted keys + bucket table), 1 (hash table), 2 (packed keys + bucket table) "
                "or 3 (rank map)");
  if (mode == 1 && !b->b.hash) {
    if (build_hash(&b->b, nullptr)) return -1;
    if (!b->b.hash) return fail("no hash table for this basis (index does not fit beside the key)");
  }
  if (mode == 2 && !b->b.keyc) return fail("no packed keys for this basis (ns > 56 or intrinsic phases)");
  if (mode == 3) {
    int ok = 0;
    if (build_rank_map(&b->b, nullptr, &ok)) return -1;
    if (!ok) return fail("no rank map for this basis (needs ns <= 48, a U(1) filling and C(ns, n_up)/4 bytes <= 4 GB)");
  }
  b->b.use_hash = mode == 1;
  b->b.use_packed = mode == 2;
  b->b.use_rank = mode == 3;
  return 0;
}

int qes_basis_expand(qes_basis_t b, const void* x_red_dev, void* full_dev, int32_t dtype, void* stream) {
  if (!b || !x_red_dev || !full_dev) return fail("null argument");
  return expand_reduce(&b->b, 0, (const double*)x_red_dev, (double*)full_dev, dtype, (hipStream_t)stream);
}

int qes_basis_reduce(qes_basis_t b, const void* full_dev, void* x_red_dev, int32_t dtype, void* stream) {
  if (!b || !x_red_dev || !full_dev) return fail("null argument");
  return expand_reduce(&b->b, 1, (const double*)x_red_dev, (double*)full_dev, dtype, (hipStream_t)stream);
}

int qes_basis_set_bucket(qes_basis_t b, int32_t rows_log2) {
  if (!b) return fail("null basis");
  if (basis_check_unused(&b->b, "qes_basis_set_bucket")) return -1;
  if (rows_log2 < -4 || rows_log2 > 20) return fail("rows_log2 must be in -4..20");
  return build_bucket(&b->b, rows_log2, nullptr);
}

int qes_basis_canon_check(qes_basis_t b, int32_t kind, const uint64_t* states_host, int64_t n, int64_t* mismatches) {
  if (!b || !mismatches) return fail("null argument");
  if (n <= 0) {
    *mismatches = 0;
    return 0;
  }
  if (!states_host) return fail("null argument");
  uint64_t* ds = nullptr;
  QES_HIP(hipMalloc(&ds, n * 8));
  const hipError_t e = hipMemcpy(ds, states_host, n * 8, hipMemcpyHostToDevice);
  const int rc = e == hipSuccess ? canon_check(&b->b, kind, nullptr, ds, n, mismatches, nullptr)
                                 : check_hip(e, "qes_basis_canon_check");
  (void)hipFree(ds);
  return rc;
}

int qes_op_destroy(qes_op_t h) {
  if (!h) return 0;
  if (op_check_unused(&h->op, "qes_op_destroy")) return -1;
  free_plan(&h->op);
  if (h->op.d_groups) (void)hipFree(h->op.d_groups);
  delete h;
  return 0;
}

int qes_op_needs_complex(qes_op_t h, int32_t* needs) {
  if (!h || !needs) return fail("null argument");
  *needs = (h->op.is_complex || h->op.chi_complex) ? 1 : 0;
  return 0;
}

int qes_op_n_couplings(qes_op_t h, int32_t* n) {
  if (!h || !n) return fail("null argument");
  *n = h->op.n_coupling;
  return 0;
}

int qes_matvec(qes_op_t h, const void* x_dev, void* y_dev, int64_t nbatch, int32_t dtype, int32_t accumulate,
               void* stream) {
  if (!h) return fail("null operator");
  if (dtype != 0 && dtype != 1) return fail("dtype must be 0 (float64) or 1 (complex128)");
  if (h->op.basis->n > 0 && (!x_dev || !y_dev)) return fail("null vector");
  if (h->op.basis->n > 0 && nbatch > 0) {
    // the pull-form kernel gathers x from other rows while it writes y: the
    // two ranges must not overlap
    const uintptr_t bytes = (uintptr_t)h->op.basis->n * (uintptr_t)nbatch * (dtype == 1 ? 16u : 8u);
    const uintptr_t xa = (uintptr_t)x_dev, ya = (uintptr_t)y_dev;
    if (xa < ya + bytes && ya < xa + bytes) return fail("x and y must not overlap (in-place products are not supported)");
  }
  return matvec(&h->op, x_dev, y_dev, nbatch, dtype, accumulate, (hipStream_t)stream);
}

int qes_matvec_ws(qes_op_t h, const void* x_dev, void* y_dev, int64_t nbatch, int32_t dtype, int32_t accumulate,
                  void* ws_dev, int64_t ws_bytes, void* stream) {
  if (!h) return fail("null operator");
  const int64_t need = h->op.basis->n * (dtype == 1 ? 16 : 8);
  if (nbatch == 1 && h->op.basis->n > 0) {
    if (!ws_dev) return fail("null workspace");
    if (ws_bytes < need) return fail("workspace of " + std::to_string(ws_bytes) + " bytes; the product needs " +
                                     std::to_string(need) + " (n * 8, or n * 16 for complex128)");
    const uintptr_t wa = (uintptr_t)ws_dev, xa = (uintptr_t)x_dev, ya = (uintptr_t)y_dev, b = (uintptr_t)need;
    if ((wa < xa + b && xa < wa + b) || (wa < ya + b && ya < wa + b))
      return fail("the workspace must not overlap x or y");
  }
  if (dtype != 0 && dtype != 1) return fail("dtype must be 0 (float64) or 1 (complex128)");
  if (h->op.basis->n > 0 && (!x_dev || !y_dev)) return fail("null vector");
  if (h->op.basis->n > 0 && nbatch > 0) {
    const uintptr_t bytes = (uintptr_t)h->op.basis->n * (uintptr_t)nbatch * (dtype == 1 ? 16u : 8u);
    const uintptr_t xa = (uintptr_t)x_dev, ya = (uintptr_t)y_dev;
    if (xa < ya + bytes && ya < xa + bytes) return fail("x and y must not overlap (in-place products are not supported)");
  }
  return matvec(&h->op, x_dev, y_dev, nbatch, dtype, accumulate, (hipStream_t)stream,
                nbatch == 1 ? (double*)ws_dev : nullptr);
}

int qes_matvec_couplings(qes_op_t h, const void* x_dev, void* y_dev, int32_t dtype, void* stream) {
  if (!h) return fail("null operator");
  return matvec_couplings(&h->op, x_dev, y_dev, dtype, (hipStream_t)stream);
}

int qes_op_diagonal(qes_op_t h, void* re_dev, void* im_dev, void* stream) {
  if (!h || !re_dev) return fail("null argument");
  return op_diagonal(&h->op, (double*)re_dev, (double*)im_dev, (hipStream_t)stream);
}

int qes_op_count_offdiag(qes_op_t h, int64_t* e_off) {
  if (!h || !e_off) return fail("null argument");
  return count_offdiag(&h->op, e_off, nullptr);
}

int qes_op_count_outside(qes_op_t h, int64_t* outside) {
  if (!h || !outside) return fail("null argument");
  int64_t e = 0;
  return count_offdiag(&h->op, &e, nullptr, outside);
}

int qes_lanczos_create(qes_op_t op, int32_t max_steps, int32_t store_basis, int32_t dtype, qes_lanczos_t* out) {
  if (!op || !out) return fail("null argument");
  Lanczos* L = nullptr;
  if (lanczos_create(&op->op, max_steps, store_basis, dtype, &L)) return -1;
  auto* h = new qes_lanczos_s();
  h->L = L;
  *out = h;
  return 0;
}

int qes_lanczos_start(qes_lanczos_t h, const void* v0_dev, void* stream) {
  if (!h || !v0_dev) return fail("null argument");
  return lanczos_start(h->L, v0_dev, (hipStream_t)stream);
}

int qes_lanczos_run(qes_lanczos_t h, int32_t nsteps, int32_t reorth, int32_t use_graph) {
  if (!h) return fail("null argument");
  return lanczos_run(h->L, nsteps, reorth, use_graph);
}

int qes_lanczos_stats(qes_lanczos_t h, int32_t* second_passes) {
  if (!h || !second_passes) return fail("null argument");
  int c = 0;
  if (lanczos_second_passes(h->L, &c)) return -1;
  *second_passes = c;
  return 0;
}

int qes_lanczos_get(qes_lanczos_t h, double* alpha_host, double* beta_host, int32_t* steps_done) {
  if (!h) return fail("null argument");
  if (steps_done) *steps_done = lanczos_steps_done(h->L);
  return lanczos_get(h->L, alpha_host, beta_host);
}

int qes_lanczos_ritz(qes_lanczos_t h, const double* coeffs_host, int32_t k, void* out_dev, void* stream) {
  if (!h || !coeffs_host || !out_dev) return fail("null argument");
  return lanczos_combine(h->L, coeffs_host, k, out_dev, (hipStream_t)stream);
}

int qes_lanczos_accumulate(qes_lanczos_t h, const double* coeffs_host, int32_t nvec, void* out_dev, void* stream) {
  if (!h || !coeffs_host || !out_dev) return fail("null argument");
  (void)stream;
  return lanczos_accumulate(h->L, coeffs_host, nvec, out_dev, (hipStream_t)stream);
}

int qes_lanczos_destroy(qes_lanczos_t h) {
  if (!h) return 0;
  lanczos_destroy(h->L);
  delete h;
  return 0;
}

int qes_tridiag_ritz(const double* alpha_host, const double* beta_host, int32_t k, double* theta, double* resid,
                     double* quad) {
  if (!alpha_host || !beta_host || !theta || !resid) return fail("null argument");
  if (k < 1) return fail("k must be >= 1");
  *theta = tridiag_lowest(alpha_host, beta_host, k);
  *resid = ritz_error(alpha_host, beta_host, k, *theta, quad);
  return 0;
}

int qes_lanczos_ground_state(qes_op_t op, const void* v0_dev, int32_t dtype, int32_t max_steps, int32_t reorth,
                             double tol, int32_t check_every, double* alpha_host, double* beta_host,
                             int32_t* steps_done, double* e0, void* stream) {
  if (!op || !v0_dev || !alpha_host || !beta_host) return fail("null argument");
  if (check_every < 1) check_every = 10;
  Lanczos* L = nullptr;
  if (lanczos_create(&op->op, max_steps, reorth ? 1 : 0, dtype, &L)) return -1;
  int rc = lanczos_start(L, v0_dev, (hipStream_t)stream);
  double cur = 0.0;
  std::vector<double> a(max_steps), b(max_steps);
  // tests placed by the residual's decay when steps are long (N >= 2^20);
  // for small N a test's host round trip outweighs the steps it saves
  const bool adaptive = op->op.basis->n >= ((int64_t)1 << 20);
  int interval = check_every, m_prev = 0, k_eff = 0;
  double r_prev = 0.0;
  while (!rc && lanczos_steps_done(L) < max_steps) {
    rc = lanczos_run(L, std::min(interval, max_steps - lanczos_steps_done(L)), reorth, 1);
    if (rc) break;
    const int kall = lanczos_steps_done(L);
    rc = lanczos_get(L, a.data(), b.data());
    if (rc) break;
    // after a breakdown the recurrence runs on a zero vector until this test:
    // keep the rows that span the Krylov space
    const int k = effective_steps(a.data(), b.data(), kall);
    k_eff = k;
    cur = tridiag_lowest(a.data(), b.data(), k);
    const bool breakdown = k < kall || !(std::fabs(b[k - 1]) > 1e-14 * std::max(1.0, std::fabs(cur)));
    const double target = tol * std::max(1.0, std::fabs(cur));
    const double r = ritz_error(a.data(), b.data(), k, cur);
    if (breakdown || r <= target) break;
    if (adaptive) interval = next_check(m_prev, r_prev, k, r, target, 1, check_every);
    m_prev = k;
    r_prev = r;
  }
  const int k = k_eff;
  if (!rc) {
    memcpy(alpha_host, a.data(), sizeof(double) * k);
    memcpy(beta_host, b.data(), sizeof(double) * k);
    if (steps_done) *steps_done = k;
    if (e0) *e0 = cur;
  }
  lanczos_destroy(L);
  return rc;
}

int qes_block_dot_ws_bytes(int32_t nv, int64_t len, int64_t* bytes) {
  if (!bytes) return fail("null argument");
  if (nv < 1 || len < 1) return fail("block sizes must be >= 1");
  *bytes = block_dot_ws_bytes(nv, len);
  return 0;
}

int qes_block_dot(const double* V, int64_t ldv, int32_t nv, const double* W, int64_t ldw, int32_t ncol, int64_t len,
                  int32_t cplx, double* C, double* ws, int64_t ws_bytes, void* stream) {
  return block_dot(V, ldv, nv, W, ldw, ncol, len, cplx ? 1 : 0, C, ws, ws_bytes, (hipStream_t)stream);
}

int qes_block_combine(const double* V, int64_t ldv, int32_t nv, const double* C, double beta, double* Y, int64_t ldy,
                      int32_t ncol, int64_t len, int32_t cplx, void* stream) {
  return block_combine(V, ldv, nv, C, beta, Y, ldy, ncol, len, cplx ? 1 : 0, (hipStream_t)stream);
}

}  // extern "C"
