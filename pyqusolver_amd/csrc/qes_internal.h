// Host-side internals shared by the HIP translation units of libqes_mi.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "qes_device.h"

namespace qes {

void set_error(const std::string& msg);
int fail(const std::string& msg);           // set_error + return -1
int check_hip(hipError_t e, const char* what);

#define QES_HIP(call)                                   \
  do {                                                  \
    hipError_t _e = (call);                             \
    if (_e != hipSuccess) return qes::check_hip(_e, #call); \
  } while (0)

// Device basis of one symmetry sector.
struct Basis {
  QesGroup group{};
  int64_t n = 0;
  int gsize = 0;              // |G|
  double* chi_re = nullptr;   // device, |G|
  double* chi_im = nullptr;
  std::vector<double> h_chi_re, h_chi_im;
  uint64_t* reps = nullptr;   // device, n (sorted)
  uint32_t* keys32 = nullptr; // device, n (ns <= 32)
  double* norms = nullptr;    // device, n
  uint32_t* bucket = nullptr; // device lookup table
  int mbits = 0, e_min = 0, e_max = 0;
  int64_t nbucket = 0;
  bool owns = true;
  QesLookup lookup() const {
    QesLookup L;
    L.keys64 = reps;
    L.keys32 = keys32;
    L.bucket = bucket;
    L.n = n;
    L.mbits = mbits;
    L.e_min = e_min;
    L.e_max = e_max;
    L.ns = group.ns;
    L.hash = use_hash ? hash : nullptr;
    L.hbits = hbits;
    L.packed = (use_packed && keyc && !(use_rank && rmap)) ? 1 : 0;   // the rank map takes precedence
    L.keyc = keyc;
    L.rmap = (use_rank && rmap) ? rmap : nullptr;
    L.rtab = rtab;
    L.n_up = group.n_up;
    L.k1min = k1min;
    L.sqrtn = sqrtn;
    L.norms = norms;
    return L;
  }
  double* sqrtn = nullptr;    // device sqrt(0..256)
  // rank map (ns <= 48 with a U(1) filling): see QesLookup
  uint64_t* rmap = nullptr;
  uint64_t* rtab = nullptr;   // colex tables (layout: QesLookup)
  int k1min = 0;
  int64_t rmap_words = 0;
  bool use_rank = false;
  uint64_t* keyc = nullptr;   // device packed keys (key | (|Stab|-1) << 56), ns <= 56
  uint64_t* hash = nullptr;   // device hash table (representative -> index)
  int hbits = 0;
  bool use_hash = false;
  // Defaults from the MI355X lookup sweep (J1-J2 chain L=32, 6x6 square,
  // honeycomb 3x4): a bucket table with ~8 buckets per representative (most
  // lookups resolve with no search step) and the narrowest key array.
  bool use_packed = false;    // probe the packed keys (ns <= 56, no phases)
  int bucket_rows_log2 = -3;  // about 2^k representatives per bucket
  // rank-indexed gather vector (rank_gather products): rxs[c * rxs_cw ..] =
  // N_b x_b for the representative b of colex rank c, QES_RXS_EMPTY bits at
  // every other rank of the filling; filled once with the marker, the
  // representatives' slots rewritten by every product that uses it
  double* rxs = nullptr;
  int64_t rxs_states = 0;
  int rxs_cw = 0;
  // live Lanczos handles and batches over operators of this basis: they hold
  // its lookup tables (captured graphs keep the pointers), so the tables
  // must not be rebuilt while this is nonzero
  int users = 0;
};

// Compiled operator (Hamiltonian) bound to a basis.
struct Op {
  Basis* basis = nullptr;
  int n_diag = 0, n_off = 0;       // groups (n_off = all off-diagonal groups)
  QesPlan plan{};                  // exchange classes + generic groups + canon kind
  QesTermGroup* d_gen = nullptr;   // device: diag groups then generic off groups
  double* d_exval = nullptr;       // device exchange value table
  int32_t* d_excpl = nullptr;      // device exchange coupling ids [c][p]
  uint32_t* d_lut = nullptr;       // device row LUT (ROWS canonicaliser)
  size_t lds_bytes = 0;
  int n_coupling = 1;              // distinct coupling ids
  bool is_complex = false;         // some term value is complex
  bool chi_complex = false;        // sector characters complex
  QesTermGroup* d_groups = nullptr;   // device: diag groups then off groups
  std::vector<QesTermGroup> h_groups;
  int block = 256;
  // measured on MI355X (J1-J2 chain L=32 and 6x5 square): consecutive row
  // blocks spread over all XCDs keep one compact row window hot in the
  // shared MALL, faster than giving each XCD a contiguous slice
  int xcd_remap = 0;
  // direct_lookup -1 (default): on where the basis looks representatives up
  // in a rank map (direct hits skip the canonicalisation: L=32 2.70 -> 2.56
  // ms, 6x6 169 -> 146 ms); 0 off; 1 on.  bond_order_uniform: 0 (default)
  // each lane walks its own active bonds, which fills every slot of the
  // batched lookups (L=32 1.95 -> 1.89 ms, 6x6 116.9 -> 113.6 ms against the
  // wavefront-uniform order); 1 uniform order; -1 uniform where a rank map serves
  int bond_order_uniform = 0;
  int direct_lookup = -1;
  int top_row_test = 0;            // row lattices: rows_top_ok before each direct lookup (off: on the 6x6 J1-J2 product its VALU costs more than the skipped lookups save)
  // PRE products with a rank map: lookups in flight per lane (1, 2, 4; -1:
  // 2 for 32-bit states, whose kernel is capped at 8 waves per SIMD and
  // spills at 4, else 4; measured on MI355X: L=32 chain 2.37 / 1.94 / 1.96 ms,
  // 6x6 square 140.9 / 122.1 / 117.1 ms for 1 / 2 / 4)
  int direct_batch = -1;
  // matrix-free single-vector products with a workspace: 1 gathers from the
  // basis's rank-indexed vector (one load per looked-up state instead of a
  // map word and an x element); 0 (default) the pre-scaled dense vector
  int rank_gather = 0;
  int lean = 1;                    // lean row-sum kernels where the group and plan allow them (0: never)
  // assembled (sliced ELLPACK) table, see qes_assemble.hip
  int64_t* asm_off = nullptr;
  int32_t* asm_len = nullptr;
  int32_t* asm_col = nullptr;
  double* asm_val = nullptr;
  uint32_t* asm_packed = nullptr;
  uint16_t* asm_vidx = nullptr;
  double* asm_dict = nullptr;
  double* asm_diag = nullptr;
  double* asm_diag_c = nullptr;      // coupling-tagged: diagonal of each coupling [c][row]
  uint8_t* asm_dict_cpl = nullptr;   // coupling id of each dictionary value
  double* asm_dict_s = nullptr;      // scaled dictionary of the last assembled_matvec_scaled
  double* asm_part = nullptr;        // per-block partial sums of assembled_coupling_dots
  int64_t asm_entries = 0, asm_slices = 0, asm_bytes = 0;
  int asm_vc = 0, asm_dvc = 0, asm_fmt = -1, asm_colbits = 32, asm_ndict = 0, asm_tagged = 0;
  // column panels of the table (single-vector products): panel p holds the
  // entries whose column lies in [p*pw, (p+1)*pw) in its own sliced layout,
  // so one pass gathers from an x window that stays in the MALL.  Arrays of
  // all panels are concatenated: slice offsets [p][ns+1] (absolute), lengths
  // [p][ns].  asm_panel_cols: -1 automatic, 0 off, > 0 columns per panel
  int64_t asm_panel_cols = -1;
  int sell_grid = 0;   // single-vector table product: 0 one block per 256 rows, k > 0 k blocks per CU (grid stride)
  int sell_lds = 0;    // dynamic LDS bytes per block of that product (caps resident blocks per CU)
  int asm_npanel = 0;
  int64_t asm_pw = 0, asm_pentries = 0, asm_pbytes = 0;
  int64_t* asm_poff = nullptr;
  int32_t* asm_plen = nullptr;
  int32_t* asm_pcol = nullptr;
  double* asm_pval = nullptr;
  uint32_t* asm_ppacked = nullptr;
  uint16_t* asm_pvidx = nullptr;
  double asm_max_bytes = 64e9;
  std::vector<double> asm_h_dict;          // host copy of the dictionary
  std::vector<uint8_t> asm_h_dict_cpl;
  std::vector<double> asm_h_scale;         // scales of the dictionary in asm_dict_s
  // live Lanczos handles and batches over this operator: they keep pointers
  // to its plan and assembled table (and graphs captured with them), so
  // neither may change or be freed while this is nonzero
  int users = 0;
};

inline void op_acquire(Op* op) {
  ++op->users;
  ++op->basis->users;
}
inline void op_release(Op* op) {
  --op->users;
  --op->basis->users;
}
// fail (-1, message) when a live Lanczos handle or batch uses the operator
int op_check_unused(const Op* op, const char* what);
int basis_check_unused(const Basis* b, const char* what);

// launchers (qes_basis.hip)
int build_basis(Basis* B, hipStream_t st);
int build_hash(Basis* B, hipStream_t st);
// rank map of a sector with ns <= 48 and a U(1) filling whose map takes at
// most max_bytes; *eligible = 0 (and nothing built) otherwise
int build_rank_map(Basis* B, hipStream_t st, int* eligible, double max_bytes = 4e9);
int build_bucket(Basis* B, int rows_log2, hipStream_t st);
// rank-indexed gather vector of cw doubles per state of the filling (needs a
// rank map); allocated and filled with the empty marker once
int build_rank_gather(Basis* B, int cw, hipStream_t st);
void free_rank_gather(Basis* B);
// reduce == 0: full = expand(x_red) (2^ns entries); 1: x_red = reduce(full)
int expand_reduce(const Basis* B, int reduce, const double* x_red, double* full, int dtype, hipStream_t st);
int map_states(const Basis* B, const uint64_t* d_states, int64_t n, int64_t* d_idx, int32_t* d_g,
               int32_t* d_cnt, hipStream_t st);

// launchers (qes_matvec.hip)
// dtype 0: float64 vectors, 1: complex128 (interleaved re,im).  ws (may be
// null): n * (1 or 2) doubles of device workspace; with it a matrix-free
// single-vector product gathers from the pre-scaled vector N .* x
int matvec(const Op* op, const void* x, void* y, int64_t nbatch, int dtype, int accumulate,
           hipStream_t st, double* ws = nullptr);
int matvec_couplings(const Op* op, const void* x, void* y, int dtype, hipStream_t st);
// off-diagonal elements inside the sector; *outside (optional): nonzero
// elements whose target lies outside it (skipped by the product)
int count_offdiag(const Op* op, int64_t* e_off, hipStream_t st, int64_t* outside = nullptr);
int op_diagonal(const Op* op, double* dre, double* dim, hipStream_t st);
int projected_matvec(const Op* op, const Basis* out, const double* x, double* y, hipStream_t st);
int build_plan(Op* op, int canon_override);   // -1: automatic
bool lean_group(const Op* op);                 // the lean kernels' conditions on group and plan
// assembled operator (qes_assemble.hip)
int op_assemble(Op* op, int enable, int want_dict, hipStream_t st);
void free_assembly(Op* op);
int assembled_matvec(const Op* op, const double* x, double* y, int dtype, int accumulate, hipStream_t st);
// (n, nbatch) row-major vectors, columns in groups of four per table pass
int assembled_matvec_multi(const Op* op, const double* x, double* y, int64_t nbatch, int dtype, int accumulate,
                           hipStream_t st);
int assembled_coo(const Op* op, int32_t* rows, int32_t* cols, double* vals, int64_t capacity, hipStream_t st);
// coupling-tagged table: y = sum_c scale[c] H_c x; out_host[2c..2c+1] = <g, H_c v>
int assembled_matvec_scaled(Op* op, const double* scale, int nscale, const double* x, double* y, int dtype,
                            hipStream_t st);
int assembled_coupling_dots(Op* op, const double* v, const double* g, int dtype, double* out_host, hipStream_t st);
// same sums with no allocation: partials in a static device array
int assembled_coupling_dots_noalloc(const Op* op, const double* v, const double* g, int dtype, double* out_host,
                                    hipStream_t st);
int pick_canon_public(const QesGroup& G);
// block Lanczos contractions on FP64 matrix cores (qes_block.hip)
int64_t block_dot_ws_bytes(int nv, int64_t len);
int block_dot(const double* V, int64_t ldv, int nv, const double* W, int64_t ldw, int ncol, int64_t len, int cplx,
              double* C, double* ws, int64_t ws_bytes, hipStream_t st);
int block_combine(const double* V, int64_t ldv, int nv, const double* C, double beta, double* Y, int64_t ldy, int ncol,
                  int64_t len, int cplx, hipStream_t st);
void free_plan(Op* op);
int canon_check(const Basis* B, int kind, const uint32_t* d_lut, const uint64_t* d_states, int64_t n,
                int64_t* d_bad, hipStream_t st);

}  // namespace qes
