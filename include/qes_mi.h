/*
 * qes_mi.h -- C ABI of libqes_mi, the MI355X (gfx950) implementation of the
 * symmetric-sector spin-1/2 Hamiltonian matvec and the Lanczos ground-state
 * solve it drives.
 *
 * Plain C types only: opaque handles, host pointers for small tables,
 * device pointers (from hipMalloc or a torch tensor's data_ptr()) for
 * vectors, and hipStream_t passed as void*.  Every function returns 0 on
 * success and -1 on error; qes_last_error() then holds a message.
 *
 * Handles borrow what they are built from: an operator keeps a pointer to its
 * basis, a Lanczos handle or batch to its operators (and their assembled
 * tables), so destroy them in the reverse order of creation.  Device vectors
 * passed in are borrowed for the call (its work is ordered on the stream
 * argument where there is one) and are never freed by the library.
 *
 * Reference interfaces each entry point replaces are cited as
 * file:line relative to the reference package root Python/QES/.
 */
#ifndef QES_MI_H
#define QES_MI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qes_basis_s* qes_basis_t;
typedef struct qes_op_s* qes_op_t;
typedef struct qes_lanczos_s* qes_lanczos_t;

/* Base-op codes of one base group element (applied in generator order). */
#define QES_BASE_OP_NONE 0
#define QES_BASE_OP_REFLECT 1 /* reflection.py:143, full bit reversal        */
#define QES_BASE_OP_FLIP 2    /* parity x/z, parity.py:46-54, complement     */
#define QES_BASE_OP_FLIP_Y 3  /* parity y, parity.py:55-63, complement, +-i  */
#define QES_BASE_OP_PERM 4    /* inversion.py:53-86, bit i -> bit perm[i]     */

/*
 * Symmetry sector description (one irrep of G = translations x base).
 * Replaces SymmetryContainer.build_group/build_compiled_group and
 * get_character (Algebra/Symmetries/symmetry_container.py:866-1174,
 * :1399-1442).  chi[g] uses the reference element order
 * g = (a*nty + b)*nbase + B, with a the x power, b the y power and B the
 * base element (Algebra/Symmetries/symmetry_container.py:943-950).
 */
typedef struct {
  int32_t ns;             /* sites = bits (<= 64)                                */
  int32_t n_up;           /* U(1) popcount sector (globals.py:164-189), -1: none */
  int32_t z2;             /* popcount-parity sector +-1, 0: none (globals.py:336) */
  int32_t nbase;          /* number of base elements, 1..4                        */
  uint8_t base_ops[4][4]; /* QES_BASE_OP_* per base element                       */
  int32_t ntx;            /* x translation powers (Lx, or 1 if absent)            */
  int32_t nty;            /* y translation powers (Ly, or 1 if absent)            */
  int32_t row_bits;       /* bits per lattice row = n_basis * Lx                  */
  int32_t cell_bits;      /* bits per unit cell = n_basis                         */
  const double* chi_re;   /* host, ntx*nty*nbase                                  */
  const double* chi_im;   /* host, ntx*nty*nbase                                  */
  /* Twisted translations (boundary fluxes; NULL for none), the reference's
   * trans_cross masks and lattice.boundary_phases() table
   * (symmetry_container_jit.py:91-98, translation.py:80-89, base.py:246):
   * cross_masks[a] (a < ntx) = bits (bit = site) that cross the boundary
   * during a x translations, cross_masks[ntx + b] (b < nty) during b y
   * translations; bphase[(d * (ns + 1) + n) * 2 + {0, 1}] = re, im of the
   * phase for n crossing up spins along d (0 x, 1 y), per translation step. */
  const uint64_t* cross_masks;
  const double* bphase;
  /* Site permutation of QES_BASE_OP_PERM (an inversion whose permutation is
   * not the full site reversal: InversionSymmetry's op.perm,
   * inversion.py:135-230): bit i goes to bit perm[i]; ns entries, an
   * involution; NULL when no base element uses it. */
  const int32_t* perm;
} qes_group_spec;

/*
 * Build the sorted representative list and normalisations of a sector on the
 * GPU.  Replaces SymmetryContainer.generate_symmetric_basis
 * (Algebra/Symmetries/symmetry_container.py:1575-1765) with its kernels
 * scan_chunk_find_representatives / fill_representatives /
 * compute_normalization (Algebra/Symmetries/jit/symmetry_container_jit.py:154-255).
 * *n_out receives the reduced dimension (HilbertSpace.nh, hilbert.py:1202).
 */
int qes_basis_create(const qes_group_spec* spec, qes_basis_t* out, int64_t* n_out);

/* Host copies of the representative list and normalisations
 * (CompactSymmetryData.representative_list / normalization,
 *  symmetry_container.py:1746-1752). Either pointer may be NULL. */
int qes_basis_get(qes_basis_t b, uint64_t* reps_host, double* norms_host);

/* Device pointers of the same arrays (owned by the handle). */
int qes_basis_device_ptrs(qes_basis_t b, const uint64_t** reps_dev, const double** norms_dev);

/*
 * Entries of the reference's compact index map for arbitrary states:
 * idx_out[i] = repr_map[state] (rep index, -1 if not in the sector) and
 * g_out[i] = first minimising group element, from which
 * phase_idx = g_to_pidx[g] (symmetry_container_jit.py:248-255,
 * symmetry_container.py:184-195, :1767-1804).  Host arrays, any length.
 */
int qes_basis_map_states(qes_basis_t b, const uint64_t* states_host, int64_t n, int64_t* idx_out, int32_t* g_out);

/*
 * Reduced <-> full Hilbert space on the device (vectors of 2^ns entries on the
 * full side; dtype 0 float64, 1 complex128).  qes_basis_expand replaces
 * HilbertSpace.expand_state / expand_to_full_state_jit
 * (Algebra/hilbert.py:1659-1715, Algebra/Symmetries/jit/symmetry_container_jit.py:261-315):
 * psi[g r_k] += x_k conj(chi_g) / (N_k sqrt|G|), written in pull form (one
 * lane per full state, no atomics).  qes_basis_reduce replaces
 * transform_to_reduced_space (Algebra/Hilbert/hilbert_base.py:485-563):
 * red[k] = sum_{s in orbit k} conj(phase(s)) psi[s] / N_k (one lane per
 * representative).  Not for sectors with parity-y phases; ns <= 36.
 */
int qes_basis_expand(qes_basis_t b, const void* x_red_dev, void* full_dev, int32_t dtype, void* stream);
int qes_basis_reduce(qes_basis_t b, const void* full_dev, void* x_red_dev, int32_t dtype, void* stream);

int qes_basis_destroy(qes_basis_t b);

/*
 * Compile a spin-1/2 instruction list bound to a basis.  The arrays are the
 * ones SpecialOperator._finalize_instruction_arrays builds
 * (Algebra/Operator/special_operator.py:1392-1417): codes[T] (SpinLookupCodes,
 * operators_spin.py:1084-1118), sites[T][2] (-1 padded), coefficients
 * (re/im; im may be NULL), evaluated with the semantics of
 * sigma_composition_integer (operators_spin.py:2151-2354);
 * is_complex selects the complex build (real build keeps Re of each
 * operator value).  coupling[T] (may be NULL) tags each term with a coupling
 * id 0..7 for qes_matvec_couplings.
 */
int qes_op_create(qes_basis_t b, const int32_t* codes, const int32_t* sites, const double* coeff_re,
                  const double* coeff_im, int32_t n_terms, int32_t is_complex, const int32_t* coupling,
                  qes_op_t* out);
int qes_op_destroy(qes_op_t op);

/* 1 if the operator needs complex128 vectors (complex values or characters). */
int qes_op_needs_complex(qes_op_t op, int32_t* needs);

/*
 * y = H x on device vectors, x and y of shape (n, nbatch) row-major.
 * dtype 0: float64, 1: complex128 (interleaved).  accumulate != 0 adds into y.
 * Replaces SpecialOperator.matvec_fun (special_operator.py:593-709) and its
 * kernel _apply_op_batch_compact_seq_jit (Algebra/Hilbert/matrix_builder.py:1347-1377),
 * and Operator.matvec(..., hilbert_in=...) (Algebra/Operator/operator.py:1559-1602).
 */
int qes_matvec(qes_op_t op, const void* x_dev, void* y_dev, int64_t nbatch, int32_t dtype, int32_t accumulate,
               void* stream);
/*
 * Same product with caller workspace (device, at least n * 8 bytes for
 * float64 or n * 16 for complex128 when nbatch == 1; ignored otherwise).
 * A matrix-free single-vector product first writes N .* x there and
 * gathers the neighbours from it, so no element loads a neighbour's norm
 * (one random line fewer per neighbour that is a representative).  The
 * workspace must not overlap x or y; calls sharing one workspace must be
 * ordered on one stream.  Same reference interface as qes_matvec
 * (special_operator.py:593-709, matrix_builder.py:1347-1377).
 */
int qes_matvec_ws(qes_op_t op, const void* x_dev, void* y_dev, int64_t nbatch, int32_t dtype, int32_t accumulate,
                  void* ws_dev, int64_t ws_bytes, void* stream);

/*
 * Sector-changing product y_out = U_out^dagger O U_in x_in (complex128 in and
 * out): O compiled on the input basis (qes_op_create), the output basis a
 * sector of the same symmetry group with other quantum numbers (momentum,
 * parity, U(1) filling).  Replaces Operator.matvec(..., hilbert_out=...)
 * (Algebra/Operator/operator.py:1559-1602 -> adapters.py:193-206) and its
 * kernel _apply_op_batch_projected_compact_seq_jit
 * (Algebra/Symmetries/jit/matrix_builder_jit.py:51-108); pull form, one lane per
 * output representative, no atomics.
 */
int qes_projected_matvec(qes_op_t op, qes_basis_t out, const void* x_dev, void* y_dev, void* stream);

/* y[c] = H_c x for every coupling id c (y has n_couplings * n entries). */
int qes_matvec_couplings(qes_op_t op, const void* x_dev, void* y_dev, int32_t dtype, void* stream);
int qes_op_n_couplings(qes_op_t op, int32_t* n);

/* Diagonal H_aa of every row into device arrays (imaginary part optional). */
int qes_op_diagonal(qes_op_t op, void* re_dev, void* im_dev, void* stream);

/*
 * Assembled operator: tabulate every row of H once on the device (sliced
 * ELLPACK: 64 rows per slice, column-major inside a slice), after which
 * qes_matvec streams the table instead of canonicalising and looking up
 * neighbours (several right-hand sides: four columns per pass over the
 * table, each column bitwise equal to its single-vector product).  The device counterpart of the reference's explicit
 * sparse build (Algebra/Hilbert/matrix_builder.py:151-206, :911-1030), for
 * repeated application (Lanczos).  mode 0 frees it; 1 picks the format:
 * operators whose off-diagonal values fit a dictionary (< 2^16 distinct real
 * or complex values) store a value index beside the column (packed into 32
 * bits when index and column bits fit, format 1; else int32 + uint16, format
 * 2) and the diagonal separately; otherwise column + f64/c128 value with the
 * diagonal as entry 0 (format 0); mode 2 forces format 0; mode 3 builds a
 * coupling-tagged dictionary table (values keyed by coupling id, one
 * diagonal per coupling) for qes_op_matvec_scaled / qes_op_coupling_dots,
 * and fails when the values do not fit a dictionary or there are more than
 * 8 coupling ids.  The build fails if the table would exceed max_gb (<= 0:
 * 64 GB default).  Any change of the operator's plan (qes_op_set_option /
 * set_canonicalizer) drops the table.  qes_op_assembly_info: entries per
 * table, bytes, format (-1: none), dictionary size and 1 for a
 * coupling-tagged table (any output pointer may be NULL).
 */
int qes_op_assemble(qes_op_t op, int32_t mode, double max_gb, void* stream);
int qes_op_assembly_info(qes_op_t op, int64_t* entries, int64_t* bytes, int32_t* format, int32_t* n_values,
                         int32_t* tagged);
/* Column panels of the assembled table (option "asm_panel_cols"): number of
 * panels (0: single-vector products make one pass over the table), columns
 * per panel, entries and bytes of the panel tables (beside the one-pass
 * table, which the multi-vector, batched, coupling-dot and export paths use).
 * No reference counterpart: the reference's products have no blocked layout
 * (Algebra/Hilbert/matrix_builder.py:1347-1377 walks representatives once). */
int qes_op_panel_info(qes_op_t op, int32_t* npanel, int64_t* cols, int64_t* entries, int64_t* bytes);
/* With a coupling-tagged table (mode 3): y = sum_c scale[c] H_c x, one pass
 * over the table, for H(J) = sum_c (J_c / J0_c) H_c (the differentiable
 * forward; the reference builds a fresh operator per coupling set,
 * Algebra/hamil.py:1434-1557).  nscale = qes_op_n_couplings.  The scaled
 * value dictionary is kept on the device and re-uploaded (blocking the host
 * for the copy) only when the scales change; calls with different scales on
 * different streams race. */
int qes_op_matvec_scaled(qes_op_t op, const double* scale, int32_t nscale, const void* x_dev, void* y_dev,
                         int32_t dtype, void* stream);
/* With a coupling-tagged table: out[2c], out[2c+1] = Re, Im of
 * sum_a conj(g_a) (H_c v)_a for every coupling id c (host array of
 * 2 * n_couplings doubles), one pass over the table -- the coupling
 * gradient of <g, H(J) v>.  Blocks the host until done. */
int qes_op_coupling_dots(qes_op_t op, const void* v_dev, const void* g_dev, int32_t dtype, double* out,
                         void* stream);
/* No counterpart in the reference: the same sums as qes_op_coupling_dots
 * for callers that cannot allocate (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  It calls no hipMalloc / hipFree:
 * the per-block partial sums go to a static device array of 2048 * 8 * 2 + 16
 * doubles (256 KB, held by every GPU the library runs on until the process
 * ends); inputs of any length (a fixed grid of at most 2048 blocks strides
 * over the rows); at most 8 coupling ids, otherwise an error.  Calls that
 * overlap on one GPU share that array and race.  Blocks the host until done;
 * sums in another order than qes_op_coupling_dots (equal to rounding). */
int qes_op_coupling_dots_noalloc(qes_op_t op, const void* v_dev, const void* g_dev, int32_t dtype, double* out,
                                 void* stream);
/* Coordinate list (row, col, value) of the assembled table on the device --
 * the explicit sparse matrix of the reference's Hamiltonian.build
 * (Algebra/hamil.py:1434-1557 via matrix_builder.py:151-206).  Padding slots
 * carry value 0 at (row, row); slots of rows past n keep the caller's fill
 * (pre-fill rows with -1).  values: 1 (real) or 2 (complex) doubles per
 * entry.  rows_dev == NULL: only *count (entries needed) is returned. */
int qes_op_table_coo(qes_op_t op, int32_t* rows_dev, int32_t* cols_dev, double* vals_dev, int64_t capacity,
                     int64_t* count, void* stream);

/* Number of off-diagonal gathers of one matvec (E_off) -- stats pass. */
int qes_op_count_offdiag(qes_op_t op, int64_t* e_off);
/* Debug pass: nonzero off-diagonal elements whose target lies outside the
 * sector (another U(1) / Z2 filling, or a representative of zero norm) --
 * the targets the reference's kernel skips silently
 * (Algebra/Hilbert/matrix_builder.py:1366-1371); nonzero for an operator
 * that does not conserve the sector's quantum numbers. */
int qes_op_count_outside(qes_op_t op, int64_t* outside);

/*
 * Tuning and diagnostics (no reference counterpart; the reference has one
 * kernel).  Canonicaliser kinds: -1 automatic, 0 brute force over all group
 * elements, 1 cyclic chains (longest-zero-run candidates), 2 lattices with
 * rows (row-rotation lookup table).  qes_op_get_plan reports the chosen kind,
 * the number of exchange-bond classes and of generic term groups.
 * qes_basis_canon_check counts states on which canonicaliser `kind` differs
 * from the brute-force one (representative, stabiliser size, character).
 */
int qes_op_set_canonicalizer(qes_op_t op, int32_t kind);
/* Kernel options: "direct_lookup" (look each exchanged state up before
 * canonicalising it; hits skip the canonicalisation): 1 on, 0 off, -1
 * (default) on where the basis uses a rank map; "bond_order_uniform" (the
 * wavefront walks the bonds in one order in that pass): 1 on, 0 (default)
 * off, -1 on where a rank map serves; "xcd_remap" (0/1);
 * "top_row_test" (0 default / 1: row lattices test the top row before a
 * direct lookup; a state that fails it is no representative and goes
 * straight to canonicalisation); "direct_batch" (-1 default, 1, 2,
 * 4, 8: rank-map lookups a lane keeps in flight in a qes_matvec_ws product,
 * 8 for rank-gather products of more than 32 sites only; same sums in the
 * same order for every value); "asm_panel_cols" (-1
 * default: automatic; 0 off; > 0 columns per panel): the next qes_op_assemble
 * also splits the table into column panels for single-vector products (pass
 * p gathers only from x[p*cols, (p+1)*cols), a window that stays in the
 * cache; at most 32 panels; automatic: 128 MB windows where x exceeds 192 MB,
 * kept only while table and panels fit the max_gb limit); "rank_gather" (0
 * default, 1): qes_matvec_ws products of one column gather N_b x_b from a
 * vector indexed by the colex rank of each state of the U(1) filling (the
 * empty states hold a NaN marker), one load per looked-up state instead of
 * a rank-map word and then an x element; needs the rank map; the vector
 * belongs to the basis (8 or 16 bytes per state of the filling: 72.6 GB at
 * 6x6 J1-J2 Sz=0 real, built when the option is set, freed by setting 0),
 * so products that use it on one basis must not overlap in time; same sums
 * in the same order as without it (6x6: 113.6 -> 105.8 ms); products fall
 * back to the dense gathers while the basis looks representatives up by
 * another method.  Launch-shape knobs of the single-vector table product
 * (measured slower than the defaults, kept for measurements): "sell_grid"
 * (0 default: a block per 256 rows; 1..64: that many blocks per CU striding
 * over the rows) and "sell_lds" (bytes of LDS reserved per block, capping
 * the resident blocks).  "lean" (1 default, 0): pre-scaled products whose
 * group has no base elements but the identity and one plain spin flip and
 * whose operator has no generic term groups run kernels compiled for
 * exactly that (same sums in the same order).  The other options change the
 * order of the sums (results agree to rounding). */
int qes_op_set_option(qes_op_t op, const char* name, int64_t value);
int qes_op_get_plan(qes_op_t op, int32_t* canon, int32_t* n_classes, int32_t* n_generic);
int qes_basis_canon_check(qes_basis_t b, int32_t kind, const uint64_t* states_host, int64_t n, int64_t* mismatches);

/* Representative lookup used by the matvec: 0 sorted keys (32-bit when
 * ns <= 32; default there) + bucket table, 1 open-addressing hash table,
 * 2 packed 64-bit keys (key | stabiliser size) + bucket table (default for
 * 33..56 sites without intrinsic phases), 3 rank map (ns <= 48 with a U(1)
 * filling: a bitmap over the colex ranks of the filling's states with a
 * running count, C(ns, n_up)/4 bytes, at most 4 GB; one 16-byte load per
 * lookup; the default wherever it applies). */
int qes_basis_set_lookup(qes_basis_t b, int32_t mode);
/* Rebuild the bucket table for about 2^rows_log2 representatives per bucket
 * (-4..20; negative: several buckets per representative; default -3). */
int qes_basis_set_bucket(qes_basis_t b, int32_t rows_log2);

/*
 * Device Lanczos (FP64).  Replaces the eigensolver that
 * DiagonalizationEngine.diagonalize dispatches to
 * (Algebra/Hamil/hamil_diag_engine.py:176-300, called from
 * Algebra/hamil.py:2063-2081 with matvec=matvec_fun).
 * store_basis keeps all Lanczos vectors on the device (needed for full
 * reorthogonalisation and for Ritz vectors).  reorth 1: classical
 * Gram-Schmidt against all stored vectors every step, with a second pass
 * when the first removed more than half of |w|^2 (DGKS criterion); reorth 2:
 * always two passes; 0: plain three-term recurrence.
 */
int qes_lanczos_create(qes_op_t op, int32_t max_steps, int32_t store_basis, int32_t dtype, qes_lanczos_t* out);
int qes_lanczos_start(qes_lanczos_t L, const void* v0_dev, void* stream);
int qes_lanczos_run(qes_lanczos_t L, int32_t nsteps, int32_t reorth, int32_t use_graph);
int qes_lanczos_get(qes_lanczos_t L, double* alpha_host, double* beta_host, int32_t* steps_done);
int qes_lanczos_ritz(qes_lanczos_t L, const double* coeffs_host, int32_t k, void* out_dev, void* stream);
/* out[i] += coeffs[i] * v_j for i < nvec (1..16; out holds nvec consecutive
 * vectors), v_j = the current Lanczos vector (v_0 after start, v_{j+1} after
 * step j).  With a second run from the same start vector (the steps are
 * bitwise reproducible) this forms Ritz vectors without storing the basis. */
int qes_lanczos_accumulate(qes_lanczos_t L, const double* coeffs_host, int32_t nvec, void* out_dev, void* stream);
int qes_lanczos_destroy(qes_lanczos_t L);
/* Number of second Gram-Schmidt passes run since the last start. */
int qes_lanczos_stats(qes_lanczos_t L, int32_t* second_passes);

/*
 * All-in-one ground state: run up to max_steps Lanczos steps from v0 (device),
 * checking the lowest Ritz value theta every check_every steps (for N >=
 * 2^20 rows at most check_every steps apart: the next test goes where the
 * residual's geometric decay between the last two tests reaches tol, if that
 * is sooner); stop on
 * breakdown or when the residual norm of the Ritz pair r = |beta_k s_k|
 * (s: unit eigenvector of T_k) is <= tol*max(1,|theta|) -- the stopping test
 * of the reference's eigsh/ARPACK path (Algebra/Hamil/hamil_diag_engine.py:
 * 176-300); |theta - E0| <= r.  alpha/beta (host, max_steps) and *e0
 * receive the result.
 */
int qes_lanczos_ground_state(qes_op_t op, const void* v0_dev, int32_t dtype, int32_t max_steps, int32_t reorth,
                             double tol, int32_t check_every, double* alpha_host, double* beta_host,
                             int32_t* steps_done, double* e0, void* stream);
/* Host-only: lowest eigenvalue *theta of T_k (alpha[0..k-1], beta[0..k-2]),
 * the residual norm *resid of its Ritz pair used by the stopping test above
 * (beta[k-1] = beta_k) and, if quad is not NULL, the quadratic estimate
 * min(r, r^2/gap) of |theta - E0| (gap to the second Ritz value; not a bound
 * while a near-degenerate pair is unresolved).  Needs no device. */
int qes_tridiag_ritz(const double* alpha_host, const double* beta_host, int32_t k, double* theta, double* resid,
                     double* quad);

/*
 * Batched Lanczos over many independent operators (sectors or disorder
 * realisations).  Replaces the reference's sequential sector sweep
 * (HilbertSpace.iter_momentum_sectors / iter_symmetry_sectors,
 * Algebra/hilbert.py:412-652, one diagonalize() per sector) with one
 * persistent matvec launch over a queue of row tiles of all problems and one
 * launch per vector step, captured into a hipGraph.  Vectors of all problems
 * are concatenated (problem p occupies elements offsets[p]..offsets[p+1]).
 * Three-term recurrence without reorthogonalisation; converged problems are
 * masked out.
 */
typedef struct qes_batch_s* qes_batch_t;
int qes_batch_create(const qes_op_t* ops, int32_t nprob, int32_t max_steps, int32_t dtype, qes_batch_t* out);
/*
 * Same, plus nblock "block" problems: block i is R_i = block_nrhs[i] (1..256)
 * disorder realisations that share the basis and every off-diagonal term of
 * block_ops[i] and differ only in their diagonal, given on the device as
 * block_diag_dev[i] (float64, n x R_i row-major: H_aa of realisation r at
 * a*R_i + r, e.g. from qes_op_diagonal of each realisation).  Its vectors are
 * stored n x R_i row-major after the ordinary problems.  Each realisation is
 * its own Lanczos run (one virtual problem); virtual problems are numbered
 * ordinary problems first, then block 0 columns 0..R_0-1, and so on -- the
 * layout of alpha/beta/steps/active/e0.  Replaces one diagonalize() per
 * realisation of a disordered model (Algebra/hamil.py:2063-2081 called in a
 * loop over random fields).
 */
int qes_batch_create_blocks(const qes_op_t* ops, int32_t nprob, const qes_op_t* block_ops, const int32_t* block_nrhs,
                            const void* const* block_diag_dev, int32_t nblock, int32_t max_steps, int32_t dtype,
                            qes_batch_t* out);
/*
 * Coupling sweep: problem i is a block of ncol (1..4) columns over
 * ops[i], whose operator must be assembled with coupling tags
 * (qes_op_assemble mode 3); column r is its own Lanczos run on
 * H_r = sum_c scales[(i*ncol + r)*8 + c] H_c (c < qes_op_n_couplings; the
 * stride is 8 whatever the coupling count).  The tagged table is streamed
 * once per step for all columns (one lane per row, the row's columns in
 * registers) and every column has its own scaled value dictionary and
 * diagonal.  Replaces one diagonalize() per coupling set
 * (e.g. J2 in {0.4, 0.5, 0.6}) of the same sector.  The operators' tables
 * must outlive the batch.  Vectors are n x ncol row-major per problem.
 */
int qes_batch_create_scaled(const qes_op_t* ops, int32_t nprob, int32_t ncol, const double* scales, int32_t max_steps,
                            int32_t dtype, qes_batch_t* out);
/* element offsets of every real problem (ordinary, then blocks) + total */
int qes_batch_offsets(qes_batch_t b, int64_t* offsets /* nprob + nblock + 1 */);
int qes_batch_start(qes_batch_t b, const void* v0_dev, void* stream);
/* nsteps == 0 with use_graph != 0 captures the step graph without running it. */
int qes_batch_run(qes_batch_t b, int32_t nsteps, int32_t use_graph);
int qes_batch_get(qes_batch_t b, double* alpha_host, double* beta_host, int32_t* steps_host);
int qes_batch_set_active(qes_batch_t b, const int32_t* active_host);
int qes_batch_matvec(qes_batch_t b, const void* x_dev, void* y_dev, void* stream);
/* Schedule options: "tile_queue" (0, default: each persistent block takes
 * every gridDim-th tile; 1: blocks take the next tile from an atomic counter,
 * reset by a kernel at every product, so blocks that finish early take
 * more), "tile_order" (0 problem-contiguous, 1 largest-first
 * round-robin), "grid_mult" (persistent matvec blocks per CU), "block_rows"
 * (rows per tile of a block problem), "block_table" (1, default: block
 * problems read a neighbour table built once at creation -- column index and
 * pull-form value of every off-diagonal element, shared by all R columns and
 * all steps; 0: canonicalise and look up on the fly like qes_matvec),
 * "op_tables" (1, default: ordinary problems whose operator was assembled with
 * qes_op_assemble before qes_batch_create stream that table; the batch keeps
 * pointers to it, so the assembly must outlive the batch), "scaled_kernel"
 * (a batch of coupling-scaled blocks only -- qes_batch_create_scaled --: 1
 * runs its products in a kernel of their own, 0 in the combined kernel; the
 * same sums either way; default 1 for real, 0 for complex vectors),
 * "scaled_lds" (bytes of LDS reserved per block of that kernel, which caps
 * the resident blocks per CU; default 54272 for real vectors: two blocks),
 * such batches default to one block per tile (grid_mult 2^20), "check_min"
 * (qes_batch_ground_states: 0 tests every check_every steps; n >= 1 places
 * each test where the residual decay of the active problems predicts the
 * first crossing, at least n and at most check_every steps apart; -1,
 * default: 1 for batches of at most 4 problems with >= 2^20 rows, else 0). */
int qes_batch_set_option(qes_batch_t b, const char* name, int64_t value);
/* Lanczos on every problem of the batch; a problem stops on breakdown, at
 * its dimension, or under the residual test of qes_lanczos_ground_state. */
int qes_batch_ground_states(qes_batch_t b, double tol, int32_t check_every, double* e0_host, int32_t* steps_host);
int qes_batch_destroy(qes_batch_t b);

/*
 * Block Lanczos contractions on FP64 matrix cores (v_mfma_f64_16x16x4f64):
 * the dense parts of Hamiltonian.diagonalize(method='block_lanczos')
 * (Algebra/Hamil/hamil_diag_engine.py:59, :219 block_size; documented at
 * Algebra/hamil.py:1826-1828), whose eigensolver lives in the absent
 * general_python submodule.  Vectors are stored one after another (leading
 * dimension ld doubles, even, 16-byte aligned device memory); complex
 * vectors (cplx = 1) as interleaved (re, im) pairs, len counting doubles.
 * qes_block_dot: C = V^H W for nv vectors V and ncol vectors W (ncol <= 16
 * real, <= 8 complex); C is nv x ncol row-major on the device (complex:
 * interleaved); ws: device workspace for the partial sums (fixed
 * summation order) of qes_block_dot_ws_bytes(nv, len) bytes, enough for
 * every call with at most nv vectors.
 * qes_block_combine: Y_q = beta Y_q + sum_i V_i C[i][q] (beta = 0: Y is
 * not read; Y may be the same vectors as V).
 */
int qes_block_dot_ws_bytes(int32_t nv, int64_t len, int64_t* bytes);
int qes_block_dot(const double* V, int64_t ldv, int32_t nv, const double* W, int64_t ldw, int32_t ncol, int64_t len,
                  int32_t cplx, double* C, double* ws, int64_t ws_bytes, void* stream);
int qes_block_combine(const double* V, int64_t ldv, int32_t nv, const double* C, double beta, double* Y, int64_t ldy,
                      int32_t ncol, int64_t len, int32_t cplx, void* stream);

/* Message of the last failing call on this thread ("" if none). */
const char* qes_last_error(void);

/* Library version string. */
const char* qes_version(void);

#ifdef __cplusplus
}
#endif
#endif /* QES_MI_H */
