// Device Lanczos handle (internal).
#pragma once
#include "qes_internal.h"

namespace qes {
struct Lanczos;
int lanczos_create(Op* op, int max_steps, int store, int dtype, Lanczos** out);
int lanczos_destroy(Lanczos* L);
int lanczos_start(Lanczos* L, const void* v0, hipStream_t user);
int lanczos_run(Lanczos* L, int nsteps, int reorth, int use_graph);
#define LANCZOS_MAX_ACCUM 16
// out[i] += c[i] * (current Lanczos vector), i < nvec
int lanczos_accumulate(Lanczos* L, const double* c_host, int nvec, void* out, hipStream_t user);
int lanczos_get(Lanczos* L, double* alpha, double* beta);
int lanczos_second_passes(Lanczos* L, int* count);
int lanczos_combine(Lanczos* L, const double* c_host, int k, void* out, hipStream_t user);
int lanczos_steps_done(const Lanczos* L);
double tridiag_lowest(const double* a, const double* b, int k);
// Steps that span the Krylov space: i + 1 for the first |b[i]| <= 1e-14 *
// max(1, |theta_k|) (an invariant subspace; the rows after it are zero rows of
// a vector that is zero, whose eigenvalue 0 is spurious), else k.
int effective_steps(const double* a, const double* b, int k);
double tridiag_eig(const double* a, const double* b, int k, int idx);
// residual norm r of the lowest Ritz pair (b[k-1] = beta_k); *quad = min(r, r^2/gap)
double ritz_error(const double* a, const double* b, int k, double theta, double* quad = nullptr);
// steps to run before the next stopping test, given the residuals r0, r1 of
// the tests after m0 < m1 steps: where the geometric decay between them
// reaches target, clamped to [lo, hi]; hi when r is not decreasing
int next_check(int m0, double r0, int m1, double r1, double target, int lo, int hi);
}  // namespace qes
