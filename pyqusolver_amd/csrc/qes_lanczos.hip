// Device Lanczos (FP64 / complex128) driven by the pull-form matvec.
//
// The reference hands SpecialOperator.matvec_fun to the eigensolver of its
// (absent) general_python submodule (Algebra/Hamil/hamil_diag_engine.py:267-280,
// called from Algebra/hamil.py:2063-2081).  Here the three-term recurrence,
// the dot products, the axpys and the classical Gram-Schmidt
// reorthogonalisation (applied twice) run on the device.  The step index
// lives in device memory, so one Lanczos iteration is a fixed sequence of
// launches that is captured once into a hipGraph and replayed.  Only alpha /
// beta travel to the host, for the FP64 tridiagonal eigensolve.
#include <algorithm>
#include <cmath>
#include <vector>

#include "qes_internal.h"
#include "qes_lanczos.h"

namespace qes {

static constexpr int RB = 256;        // threads per block for vector kernels
static constexpr int VB = 8;          // Lanczos vectors per block in the reorth projection

// ---------------------------------------------------------------------------
// reductions: per-block partials then one deterministic finishing block
// ---------------------------------------------------------------------------
template <bool CPLX>
__global__ void __launch_bounds__(RB) dot_partial_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                                         int64_t n, double* __restrict__ part) {
  __shared__ double sr[RB], si[RB];
  double r = 0.0, im = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    if (CPLX) {
      double ar = a[2 * i], ai = a[2 * i + 1], br = b[2 * i], bi = b[2 * i + 1];
      r += ar * br + ai * bi;    // conj(a) * b
      im += ar * bi - ai * br;
    } else {
      r += a[i] * b[i];
    }
  }
  sr[threadIdx.x] = r;
  si[threadIdx.x] = im;
  __syncthreads();
  for (int off = RB / 2; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) {
      sr[threadIdx.x] += sr[threadIdx.x + off];
      si[threadIdx.x] += si[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sr[0];
    part[2 * blockIdx.x + 1] = si[0];
  }
}

// out[0..1] = sum of nb complex partials; mode 0: plain, 1: sqrt of real part
__global__ void __launch_bounds__(RB) finish_kernel(const double* __restrict__ part, int nb, double* __restrict__ out,
                                                    int mode) {
  __shared__ double sr[RB], si[RB];
  double r = 0.0, im = 0.0;
  for (int i = threadIdx.x; i < nb; i += RB) {
    r += part[2 * i];
    im += part[2 * i + 1];
  }
  sr[threadIdx.x] = r;
  si[threadIdx.x] = im;
  __syncthreads();
  for (int off = RB / 2; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) {
      sr[threadIdx.x] += sr[threadIdx.x + off];
      si[threadIdx.x] += si[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (mode == 1) {
      out[0] = sqrt(sr[0]);
      out[1] = 0.0;
    } else {
      out[0] = sr[0];
      out[1] = si[0];
    }
  }
}

// ---------------------------------------------------------------------------
// Lanczos step kernels.  Scalars: alpha[j], beta[j] on device; j = *d_j.
// ---------------------------------------------------------------------------

// alpha_j = Re <v, w>   (from finished dot in scal[0..1])
__global__ void set_alpha_kernel(const double* __restrict__ scal, double* __restrict__ alpha, const int* __restrict__ d_j) {
  alpha[*d_j] = scal[0];
}

// w -= alpha_j v + beta_{j-1} v_prev
template <bool CPLX>
__global__ void __launch_bounds__(RB) recur_kernel(double* __restrict__ w, const double* __restrict__ v,
                                                   const double* __restrict__ vp, const double* __restrict__ alpha,
                                                   const double* __restrict__ beta, const int* __restrict__ d_j, int64_t n) {
  const int j = *d_j;
  const double al = alpha[j];
  const double be = j > 0 ? beta[j - 1] : 0.0;
  const int cw = CPLX ? 2 : 1;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n * cw; i += (int64_t)gridDim.x * RB)
    w[i] = w[i] - al * v[i] - be * vp[i];
}

// partial projections c_i = <V_i, w> for i <= j; grid (nchunk, ceil(m/VB)).
// The vectors are read as 16-byte pairs of doubles (two real entries or one
// complex entry; V's leading dimension is even); an odd real length leaves
// one trailing entry, added by the first block column.
template <bool CPLX>
__global__ void __launch_bounds__(RB) proj_partial_kernel(const double* __restrict__ V, int64_t ldv,
                                                          const double* __restrict__ w, int64_t n,
                                                          const int* __restrict__ d_j, double* __restrict__ part,
                                                          int nchunk, const int* __restrict__ gate) {
  __shared__ double sr[VB][RB / 64], si[VB][RB / 64];
  if (gate && *gate == 0) return;
  const int j = *d_j;
  const int i0 = blockIdx.y * VB;
  if (i0 > j) return;
  double r[VB], im[VB];
#pragma unroll
  for (int k = 0; k < VB; ++k) {
    r[k] = 0.0;
    im[k] = 0.0;
  }
  const int64_t len = CPLX ? 2 * n : n;
  const int64_t m2 = len >> 1;
  const double2* __restrict__ w2 = reinterpret_cast<const double2*>(w);
  for (int64_t q = (int64_t)blockIdx.x * RB + threadIdx.x; q < m2; q += (int64_t)nchunk * RB) {
    const double2 wq = w2[q];
#pragma unroll
    for (int k = 0; k < VB; ++k) {
      if (i0 + k <= j) {
        const double2 vq = reinterpret_cast<const double2*>(V + (int64_t)(i0 + k) * ldv)[q];
        if (CPLX) {
          r[k] += vq.x * wq.x + vq.y * wq.y;
          im[k] += vq.x * wq.y - vq.y * wq.x;
        } else {
          r[k] += vq.x * wq.x;
          r[k] += vq.y * wq.y;
        }
      }
    }
  }
  if (!CPLX && (len & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < VB; ++k)
      if (i0 + k <= j) r[k] += V[(int64_t)(i0 + k) * ldv + len - 1] * w[len - 1];
  }
  // wavefront reduce then across the 4 waves
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < VB; ++k) {
    double a = r[k], b = im[k];
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_down(a, off, 64);
      b += __shfl_down(b, off, 64);
    }
    if (lane == 0) {
      sr[k][wid] = a;
      si[k][wid] = b;
    }
  }
  __syncthreads();
  if (threadIdx.x < VB) {
    const int k = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (int q = 0; q < RB / 64; ++q) {
      a += sr[k][q];
      b += si[k][q];
    }
    // part layout: [vector][chunk][2]
    part[((int64_t)(i0 + k) * nchunk + blockIdx.x) * 2] = a;
    part[((int64_t)(i0 + k) * nchunk + blockIdx.x) * 2 + 1] = b;
  }
}

// c_i = sum over chunks (one block per vector, i <= j)
__global__ void __launch_bounds__(RB) proj_finish_kernel(const double* __restrict__ part, int nchunk,
                                                         const int* __restrict__ d_j, double* __restrict__ coef,
                                                         const int* __restrict__ gate) {
  __shared__ double sr[RB], si[RB];
  if (gate && *gate == 0) return;
  const int i = blockIdx.x;
  if (i > *d_j) return;
  double r = 0.0, im = 0.0;
  for (int c = threadIdx.x; c < nchunk; c += RB) {
    r += part[((int64_t)i * nchunk + c) * 2];
    im += part[((int64_t)i * nchunk + c) * 2 + 1];
  }
  sr[threadIdx.x] = r;
  si[threadIdx.x] = im;
  __syncthreads();
  for (int off = RB / 2; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) {
      sr[threadIdx.x] += sr[threadIdx.x + off];
      si[threadIdx.x] += si[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    coef[2 * i] = sr[0];
    coef[2 * i + 1] = si[0];
  }
}

// w -= sum_{i<=j} c_i V_i   (c complex when CPLX).  Chunk-major: a block
// owns UP_PAIRS 16-byte pairs of rows (two real entries or one complex
// entry; V's leading dimension is even) and walks the stored vectors one
// after the other, so each wavefront streams one contiguous 16 KB piece of
// one vector at a time (5.5 -> 6.0 TB/s against interleaving four vectors
// per loop trip).  The sums keep the order i = 0, 1, ..., j.
static constexpr int UP_PT = 16;                 // pairs per thread
static constexpr int UP_PAIRS = UP_PT * RB;      // pairs per block
template <bool CPLX>
__global__ void __launch_bounds__(RB) proj_update_kernel(const double* __restrict__ V, int64_t ldv,
                                                         double* __restrict__ w, int64_t n, const int* __restrict__ d_j,
                                                         const double* __restrict__ coef, const int* __restrict__ gate) {
  __shared__ double sc[2 * 1024];
  if (gate && *gate == 0) return;
  const int j = *d_j;
  for (int i = threadIdx.x; i <= j && i < 1024; i += RB) {
    sc[2 * i] = coef[2 * i];
    sc[2 * i + 1] = coef[2 * i + 1];
  }
  __syncthreads();
  const int64_t len = CPLX ? 2 * n : n;
  const int64_t m2 = len >> 1;
  const int64_t p0 = (int64_t)blockIdx.x * UP_PAIRS;
  double ax[UP_PT], ay[UP_PT];   // real: entries 2q, 2q+1; complex: re, im of entry q
#pragma unroll
  for (int u = 0; u < UP_PT; ++u) ax[u] = ay[u] = 0.0;
  if (p0 < m2) {
    const int np = (int)((m2 - p0) < UP_PAIRS ? (m2 - p0) : UP_PAIRS);
#pragma unroll 1
    for (int i = 0; i <= j; ++i) {
      const double cr = sc[2 * i], ci = sc[2 * i + 1];
      const double2* __restrict__ v2 = reinterpret_cast<const double2*>(V + (int64_t)i * ldv) + p0;
#pragma unroll
      for (int u = 0; u < UP_PT; ++u) {
        const int p = threadIdx.x + RB * u;
        if (p < np) {
          const double2 vq = v2[p];
          if (CPLX) {
            ax[u] += cr * vq.x - ci * vq.y;
            ay[u] += cr * vq.y + ci * vq.x;
          } else {
            ax[u] += cr * vq.x;
            ay[u] += cr * vq.y;
          }
        }
      }
    }
    double2* __restrict__ w2 = reinterpret_cast<double2*>(w) + p0;
#pragma unroll
    for (int u = 0; u < UP_PT; ++u) {
      const int p = threadIdx.x + RB * u;
      if (p < np) {
        double2 wq = w2[p];
        wq.x -= ax[u];
        wq.y -= ay[u];
        w2[p] = wq;
      }
    }
  }
  if (!CPLX && (len & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    double a = 0.0;
    for (int i = 0; i <= j; ++i) a += sc[2 * i] * V[(int64_t)i * ldv + len - 1];
    w[len - 1] -= a;
  }
}

// beta_j = ||w|| (already in scal[0]); v_prev <- v; v <- w / beta_j; V[j+1] <- v
__global__ void __launch_bounds__(RB) advance_kernel(double* __restrict__ w, double* __restrict__ v,
                                                     double* __restrict__ vp, double* __restrict__ V, int64_t ldv,
                                                     int store, const double* __restrict__ scal,
                                                     const int* __restrict__ d_j, int64_t len, int maxv) {
  const int j = *d_j;
  const double be = scal[0];
  const double inv = be > 0.0 ? 1.0 / be : 0.0;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < len; i += (int64_t)gridDim.x * RB) {
    double nv = w[i] * inv;
    vp[i] = v[i];
    v[i] = nv;
    if (store && j + 1 < maxv) V[(int64_t)(j + 1) * ldv + i] = nv;
  }
}

__global__ void set_beta_step_kernel(const double* __restrict__ scal, double* __restrict__ beta, int* __restrict__ d_j) {
  beta[*d_j] = scal[0];
  *d_j = *d_j + 1;
}

__global__ void __launch_bounds__(RB) fill_kernel(double* __restrict__ p, int64_t len, double val) {
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < len; i += (int64_t)gridDim.x * RB) p[i] = val;
}

__global__ void __launch_bounds__(RB) scale_copy_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                                        double* __restrict__ dst2, int64_t len,
                                                        const double* __restrict__ scal) {
  const double inv = scal[0] > 0.0 ? 1.0 / scal[0] : 0.0;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < len; i += (int64_t)gridDim.x * RB) {
    double v = src[i] * inv;
    dst[i] = v;
    if (dst2) dst2[i] = v;
  }
}

__global__ void set_int_kernel(int* p, int v) { *p = v; }

// DGKS criterion: a second Gram-Schmidt pass only when the first removed
// more than half of |w|^2 (|w1| < |w0| / sqrt 2); s0, s1 = squared norms
__global__ void dgks_flag_kernel(const double* __restrict__ s0, const double* __restrict__ s1, int* __restrict__ flag,
                                 int* __restrict__ count) {
  const int f = s1[0] < 0.5 * s0[0] ? 1 : 0;
  *flag = f;
  *count += f;
}

// psi = sum_i c_i V_i  (real coefficients c)
template <bool CPLX>
__global__ void __launch_bounds__(RB) combine_kernel(const double* __restrict__ V, int64_t ldv, int k,
                                                     const double* __restrict__ c, double* __restrict__ out, int64_t n) {
  const int cw = CPLX ? 2 : 1;
  for (int64_t e = (int64_t)blockIdx.x * RB + threadIdx.x; e < n * cw; e += (int64_t)gridDim.x * RB) {
    double acc = 0.0;
    for (int i = 0; i < k; ++i) acc += c[i] * V[(int64_t)i * ldv + e];
    out[e] = acc;
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct Lanczos {
  Op* op = nullptr;
  bool acquired = false;   // counted in op->users
  int dtype = 0;
  int maxv = 0;          // stored vectors (max_steps + 1) when store
  int max_steps = 0;
  int store = 0;
  int64_t n = 0, len = 0;  // len = n * cw
  int64_t ldv = 0;         // leading dimension of V: len rounded up to even (16-byte rows)
  double *v = nullptr, *vp = nullptr, *w = nullptr, *V = nullptr, *ws = nullptr;
  double *alpha = nullptr, *beta = nullptr, *scal = nullptr, *part = nullptr, *coef = nullptr, *ppart = nullptr;
  int* d_j = nullptr;
  double* dg = nullptr;    // squared norms before / after the first Gram-Schmidt pass
  int* flag = nullptr;     // [0] second pass needed, [1] number of second passes run
  int nred = 0, nchunk = 0, nvec_blocks = 0;
  int steps_done = 0;
  int reorth_graph = -1;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static int grid_for(int64_t len) {
  int64_t g = (len + RB - 1) / RB;
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  return (int)g;
}

static int enqueue_dot(Lanczos* L, const double* a, const double* b, double* out, int mode, hipStream_t st) {
  if (L->dtype == 1)
    hipLaunchKernelGGL(dot_partial_kernel<true>, dim3(L->nred), dim3(RB), 0, st, a, b, L->n, L->part);
  else
    hipLaunchKernelGGL(dot_partial_kernel<false>, dim3(L->nred), dim3(RB), 0, st, a, b, L->n, L->part);
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(RB), 0, st, L->part, L->nred, out, mode);
  return 0;
}

static int enqueue_step(Lanczos* L, int reorth, hipStream_t st) {
  const bool c = L->dtype == 1;
  if (matvec(L->op, L->v, L->w, 1, L->dtype, 0, st, L->ws)) return -1;
  enqueue_dot(L, L->v, L->w, L->scal, 0, st);
  hipLaunchKernelGGL(set_alpha_kernel, dim3(1), dim3(1), 0, st, L->scal, L->alpha, L->d_j);
  const int g = grid_for(L->len);
  if (c)
    hipLaunchKernelGGL(recur_kernel<true>, dim3(g), dim3(RB), 0, st, L->w, L->v, L->vp, L->alpha, L->beta, L->d_j, L->n);
  else
    hipLaunchKernelGGL(recur_kernel<false>, dim3(g), dim3(RB), 0, st, L->w, L->v, L->vp, L->alpha, L->beta, L->d_j, L->n);
  if (reorth && L->store) {
    // reorth 1: second pass by the DGKS criterion; 2: always twice
    if (reorth == 1) enqueue_dot(L, L->w, L->w, L->dg, 0, st);
    for (int pass = 0; pass < 2; ++pass) {
      const int* gate = nullptr;
      if (pass == 1 && reorth == 1) {
        enqueue_dot(L, L->w, L->w, L->dg + 2, 0, st);
        hipLaunchKernelGGL(dgks_flag_kernel, dim3(1), dim3(1), 0, st, L->dg, L->dg + 2, L->flag, L->flag + 1);
        gate = L->flag;
      }
      dim3 pg(L->nchunk, L->nvec_blocks);
      if (c)
        hipLaunchKernelGGL(proj_partial_kernel<true>, pg, dim3(RB), 0, st, L->V, L->ldv, L->w, L->n, L->d_j, L->ppart,
                           L->nchunk, gate);
      else
        hipLaunchKernelGGL(proj_partial_kernel<false>, pg, dim3(RB), 0, st, L->V, L->ldv, L->w, L->n, L->d_j, L->ppart,
                           L->nchunk, gate);
      hipLaunchKernelGGL(proj_finish_kernel, dim3(L->maxv), dim3(RB), 0, st, L->ppart, L->nchunk, L->d_j, L->coef,
                         gate);
      const int gu = (int)std::max<int64_t>(1, ((L->len >> 1) + UP_PAIRS - 1) / UP_PAIRS);
      if (c)
        hipLaunchKernelGGL(proj_update_kernel<true>, dim3(gu), dim3(RB), 0, st, L->V, L->ldv, L->w, L->n, L->d_j,
                           L->coef, gate);
      else
        hipLaunchKernelGGL(proj_update_kernel<false>, dim3(gu), dim3(RB), 0, st, L->V, L->ldv, L->w, L->n, L->d_j,
                           L->coef, gate);
    }
  }
  enqueue_dot(L, L->w, L->w, L->scal, 1, st);
  hipLaunchKernelGGL(advance_kernel, dim3(g), dim3(RB), 0, st, L->w, L->v, L->vp, L->V, L->ldv, L->store, L->scal,
                     L->d_j, L->len, L->maxv);
  hipLaunchKernelGGL(set_beta_step_kernel, dim3(1), dim3(1), 0, st, L->scal, L->beta, L->d_j);
  QES_HIP(hipGetLastError());
  return 0;
}

int lanczos_create(Op* op, int max_steps, int store, int dtype, Lanczos** out) {
  if (max_steps < 1) return fail("max_steps must be >= 1");
  if (store && max_steps + 1 > 1024) return fail("stored Lanczos basis limited to 1023 steps");
  if (dtype != 0 && dtype != 1) return fail("dtype must be 0 (float64) or 1 (complex128)");
  if (dtype == 0 && (op->is_complex || op->chi_complex)) return fail("complex sector/model needs complex128");
  Lanczos* L = new Lanczos();
  L->op = op;
  L->dtype = dtype;
  L->max_steps = max_steps;
  L->store = store;
  L->n = op->basis->n;
  L->len = L->n * (dtype == 1 ? 2 : 1);
  L->ldv = (L->len + 1) & ~int64_t(1);
  L->maxv = max_steps + 1;
  const int64_t len = std::max<int64_t>(L->len, 1);
  L->nred = grid_for(L->n);
  L->nchunk = (int)std::min<int64_t>(std::max<int64_t>((L->n + RB - 1) / RB, 1), 512);
  L->nvec_blocks = (L->maxv + VB - 1) / VB;
  const size_t vbytes = store ? (size_t)std::max<int64_t>(L->ldv, 2) * 8 * L->maxv : 0;
  hipError_t e = hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking);
  const char* what = "stream";
  auto grab = [&](auto** p, size_t bytes, const char* name) {
    if (e == hipSuccess) {
      e = hipMalloc((void**)p, bytes);
      what = name;
    }
  };
  grab(&L->v, len * 8, "Lanczos vectors");
  grab(&L->vp, len * 8, "Lanczos vectors");
  grab(&L->w, len * 8, "Lanczos vectors");
  grab(&L->ws, len * 8, "Lanczos vectors");   // gather workspace of the matrix-free product
  if (store) grab(&L->V, vbytes, "stored Krylov basis");
  grab(&L->alpha, sizeof(double) * L->maxv, "coefficients");
  grab(&L->beta, sizeof(double) * L->maxv, "coefficients");
  grab(&L->scal, sizeof(double) * 4, "scalars");
  grab(&L->part, sizeof(double) * 2 * L->nred, "partial sums");
  grab(&L->coef, sizeof(double) * 2 * L->maxv, "coefficients");
  grab(&L->ppart, sizeof(double) * 2 * (size_t)L->nchunk * L->nvec_blocks * VB, "partial sums");
  grab(&L->d_j, sizeof(int), "scalars");
  grab(&L->dg, sizeof(double) * 4, "scalars");
  grab(&L->flag, sizeof(int) * 2, "scalars");
  if (e != hipSuccess) {   // nothing of the handle survives a failed create
    lanczos_destroy(L);
    std::string msg = std::string("lanczos_create (") + what;
    if (store) msg += "; the stored basis alone takes " + std::to_string(vbytes / 1e9) + " GB";
    return check_hip(e, (msg + ")").c_str());
  }
  op_acquire(op);
  L->acquired = true;
  *out = L;
  return 0;
}

int lanczos_destroy(Lanczos* L) {
  if (!L) return 0;
  if (L->acquired) op_release(L->op);
  if (L->exec) (void)hipGraphExecDestroy(L->exec);
  if (L->graph) (void)hipGraphDestroy(L->graph);
  (void)hipFree(L->v);
  (void)hipFree(L->vp);
  (void)hipFree(L->w);
  (void)hipFree(L->ws);
  if (L->V) (void)hipFree(L->V);
  (void)hipFree(L->alpha);
  (void)hipFree(L->beta);
  (void)hipFree(L->scal);
  (void)hipFree(L->part);
  (void)hipFree(L->coef);
  (void)hipFree(L->ppart);
  (void)hipFree(L->d_j);
  (void)hipFree(L->dg);
  (void)hipFree(L->flag);
  if (L->stream) (void)hipStreamDestroy(L->stream);
  delete L;
  return 0;
}

int lanczos_start(Lanczos* L, const void* v0, hipStream_t user) {
  hipStream_t st = L->stream;
  hipEvent_t ev;
  QES_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  QES_HIP(hipEventRecord(ev, user));
  QES_HIP(hipStreamWaitEvent(st, ev, 0));
  const int g = grid_for(L->len);
  // v = v0 / ||v0||; V[0] = v; v_prev = 0
  hipLaunchKernelGGL(fill_kernel, dim3(g), dim3(RB), 0, st, L->vp, L->len, 0.0);
  QES_HIP(hipMemcpyAsync(L->w, v0, L->len * 8, hipMemcpyDeviceToDevice, st));
  enqueue_dot(L, L->w, L->w, L->scal, 1, st);
  hipLaunchKernelGGL(scale_copy_kernel, dim3(g), dim3(RB), 0, st, L->w, L->v, L->store ? L->V : nullptr, L->len, L->scal);
  hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, st, L->d_j, 0);
  hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, st, L->flag + 1, 0);
  QES_HIP(hipGetLastError());
  L->steps_done = 0;
  QES_HIP(hipStreamSynchronize(st));
  QES_HIP(hipEventDestroy(ev));
  return 0;
}

int lanczos_run(Lanczos* L, int nsteps, int reorth, int use_graph) {
  hipStream_t st = L->stream;
  nsteps = std::min(nsteps, L->max_steps - L->steps_done);
  if (nsteps <= 0) return 0;
  if (use_graph) {
    if (!L->exec || L->reorth_graph != reorth) {
      if (L->exec) {
        (void)hipGraphExecDestroy(L->exec);
        (void)hipGraphDestroy(L->graph);
        L->exec = nullptr;
        L->graph = nullptr;
      }
      QES_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      int rc = enqueue_step(L, reorth, st);
      hipGraph_t gph = nullptr;
      hipError_t e = hipStreamEndCapture(st, &gph);
      if (rc) return rc;
      if (e != hipSuccess) return check_hip(e, "hipStreamEndCapture");
      L->graph = gph;
      QES_HIP(hipGraphInstantiate(&L->exec, L->graph, nullptr, nullptr, 0));
      L->reorth_graph = reorth;
    }
    for (int s = 0; s < nsteps; ++s) QES_HIP(hipGraphLaunch(L->exec, st));
  } else {
    for (int s = 0; s < nsteps; ++s)
      if (enqueue_step(L, reorth, st)) return -1;
  }
  QES_HIP(hipStreamSynchronize(st));
  L->steps_done += nsteps;
  return 0;
}

int lanczos_second_passes(Lanczos* L, int* count) {
  QES_HIP(hipMemcpy(count, L->flag + 1, sizeof(int), hipMemcpyDeviceToHost));
  return 0;
}

int lanczos_get(Lanczos* L, double* alpha, double* beta) {
  const int k = L->steps_done;
  if (k > 0) {
    QES_HIP(hipMemcpy(alpha, L->alpha, sizeof(double) * k, hipMemcpyDeviceToHost));
    QES_HIP(hipMemcpy(beta, L->beta, sizeof(double) * k, hipMemcpyDeviceToHost));
  }
  return 0;
}

// Make the handle's stream wait for the work already queued on the caller's
// stream (e.g. the zero fill of an output buffer): the handle's stream is a
// non-blocking one of its own, so nothing orders the two otherwise.
static int order_after(Lanczos* L, hipStream_t user) {
  if (user == L->stream) return 0;
  hipEvent_t ev;
  QES_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  hipError_t e = hipEventRecord(ev, user);
  if (e == hipSuccess) e = hipStreamWaitEvent(L->stream, ev, 0);
  (void)hipEventDestroy(ev);
  if (e != hipSuccess) return check_hip(e, "ordering after the caller's stream");
  return 0;
}

int lanczos_combine(Lanczos* L, const double* c_host, int k, void* out, hipStream_t user) {
  if (!L->store) return fail("Lanczos basis was not stored");
  if (k < 1 || k > L->steps_done + 1 || k > L->maxv) return fail("bad number of Ritz coefficients");
  double* dc = nullptr;
  QES_HIP(hipMalloc(&dc, sizeof(double) * k));
  QES_HIP(hipMemcpy(dc, c_host, sizeof(double) * k, hipMemcpyHostToDevice));
  if (order_after(L, user)) {
    (void)hipFree(dc);
    return -1;
  }
  const int g = grid_for(L->len);
  if (L->dtype == 1)
    hipLaunchKernelGGL(combine_kernel<true>, dim3(g), dim3(RB), 0, L->stream, L->V, L->ldv, k, dc, (double*)out, L->n);
  else
    hipLaunchKernelGGL(combine_kernel<false>, dim3(g), dim3(RB), 0, L->stream, L->V, L->ldv, k, dc, (double*)out, L->n);
  QES_HIP(hipGetLastError());
  QES_HIP(hipStreamSynchronize(L->stream));
  QES_HIP(hipFree(dc));
  return 0;
}

// out[i] += c[i] * v for i < nvec (out: nvec consecutive vectors of len doubles;
// complex vectors take real coefficients: the Ritz vectors of T_m are real)
struct AccumCoef {
  double c[LANCZOS_MAX_ACCUM];
};

__global__ void __launch_bounds__(RB) accumulate_kernel(const double* __restrict__ v, int64_t len, AccumCoef cf,
                                                        int nvec, double* __restrict__ out) {
  for (int64_t e = (int64_t)blockIdx.x * RB + threadIdx.x; e < len; e += (int64_t)gridDim.x * RB) {
    const double x = v[e];
    for (int i = 0; i < nvec; ++i) out[(int64_t)i * len + e] += cf.c[i] * x;
  }
}

int lanczos_accumulate(Lanczos* L, const double* c_host, int nvec, void* out, hipStream_t user) {
  if (nvec < 1 || nvec > LANCZOS_MAX_ACCUM) return fail("accumulate takes 1.." + std::to_string(LANCZOS_MAX_ACCUM) +
                                                        " vectors");
  if (order_after(L, user)) return -1;
  AccumCoef cf{};
  for (int i = 0; i < nvec; ++i) cf.c[i] = c_host[i];
  const int g = grid_for(L->len);
  hipLaunchKernelGGL(accumulate_kernel, dim3(g), dim3(RB), 0, L->stream, L->v, L->len, cf, nvec, (double*)out);
  QES_HIP(hipGetLastError());
  QES_HIP(hipStreamSynchronize(L->stream));
  return 0;
}

int lanczos_steps_done(const Lanczos* L) { return L->steps_done; }

// eigenvalue idx (0 = lowest) of the symmetric tridiagonal T_k with
// diagonal a[0..k-1] and off-diagonal b[0..k-2], by Sturm-count bisection
double tridiag_eig(const double* a, const double* b, int k, int idx) {
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < k; ++i) {
    double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i < k - 1 ? std::fabs(b[i]) : 0.0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  for (int it = 0; it < 200 && hi - lo > 1e-15 * std::max(1.0, std::fabs(lo)); ++it) {
    double mid = 0.5 * (lo + hi);
    int neg = 0;
    double d = 1.0;
    for (int i = 0; i < k; ++i) {
      double bb = i > 0 ? b[i - 1] * b[i - 1] : 0.0;
      d = a[i] - mid - (i > 0 ? bb / d : 0.0);
      if (d == 0.0) d = -1e-300;
      if (d < 0) ++neg;
    }
    if (neg >= idx + 1) hi = mid; else lo = mid;
  }
  return 0.5 * (lo + hi);
}

double tridiag_lowest(const double* a, const double* b, int k) { return tridiag_eig(a, b, k, 0); }

int effective_steps(const double* a, const double* b, int k) {
  if (k <= 1) return k;
  const double scale = std::max(1.0, std::fabs(tridiag_lowest(a, b, k)));
  for (int i = 0; i < k; ++i)
    if (!(std::fabs(b[i]) > 1e-14 * scale)) return i + 1;
  return k;
}

int next_check(int m0, double r0, int m1, double r1, double target, int lo, int hi) {
  if (lo < 1) lo = 1;
  if (hi < lo) hi = lo;
  if (!(m1 > m0 && r0 > 0.0 && r1 > 0.0 && r1 < r0 && target > 0.0 && r1 > target)) return hi;
  const double rate = (std::log(r1) - std::log(r0)) / (double)(m1 - m0);   // < 0
  const double need = std::ceil((std::log(target) - std::log(r1)) / rate);
  if (!(need < (double)hi)) return hi;
  return need < (double)lo ? lo : (int)need;
}

// Residual norm r = ||H y - theta y|| = |beta_k s_k| of the lowest Ritz pair
// after k Lanczos steps (b[k-1] = beta_k, the coefficient of the next basis
// vector; s = unit eigenvector of T_k for theta).  |theta - lambda| <= r for
// some eigenvalue lambda of H, and, once theta is separated from the rest of
// the spectrum, |theta - lambda| <~ r^2 / gap; *quad (optional) receives
// that quadratic estimate with gap = theta_2 - theta from the Ritz values.
// s comes from two steps of inverse iteration with T_k - theta I factorised
// by Gaussian elimination with row interchanges (stable when theta is an
// eigenvalue to working precision).
double ritz_error(const double* a, const double* b, int k, double theta, double* quad) {
  if (quad) *quad = 1e300;
  if (k <= 0) return 1e300;
  const double beta = std::fabs(b[k - 1]);
  if (k == 1) {
    if (quad) *quad = beta;
    return beta;
  }
  double tnorm = 0.0;
  for (int i = 0; i < k; ++i)
    tnorm = std::max(tnorm, std::fabs(a[i]) + (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i < k - 1 ? std::fabs(b[i]) : 0.0));
  const double small = std::max(tnorm, 1e-300) * 1e-16;
  std::vector<double> d(k), du(k, 0.0), du2(k, 0.0), dl(k, 0.0), x(k, 1.0), rhs(k);
  std::vector<char> sw(k, 0);
  for (int i = 0; i < k; ++i) d[i] = a[i] - theta;
  for (int i = 0; i < k - 1; ++i) du[i] = dl[i] = b[i];
  for (int i = 0; i < k - 1; ++i) {
    if (std::fabs(d[i]) >= std::fabs(dl[i])) {
      if (std::fabs(d[i]) < small) d[i] = d[i] < 0 ? -small : small;
      const double f = dl[i] / d[i];
      dl[i] = f;
      d[i + 1] -= f * du[i];
    } else {
      sw[i] = 1;
      const double f = d[i] / dl[i];
      d[i] = dl[i];
      dl[i] = f;
      const double t = du[i];
      du[i] = d[i + 1];
      d[i + 1] = t - f * d[i + 1];
      if (i < k - 2) {
        du2[i] = du[i + 1];
        du[i + 1] = -f * du[i + 1];
      }
    }
  }
  if (std::fabs(d[k - 1]) < small) d[k - 1] = d[k - 1] < 0 ? -small : small;
  for (int pass = 0; pass < 2; ++pass) {
    rhs = x;
    for (int i = 0; i < k - 1; ++i) {
      if (sw[i]) std::swap(rhs[i], rhs[i + 1]);
      rhs[i + 1] -= dl[i] * rhs[i];
    }
    for (int i = k - 1; i >= 0; --i) {
      double v = rhs[i];
      if (i + 1 < k) v -= du[i] * x[i + 1];
      if (i + 2 < k) v -= du2[i] * x[i + 2];
      x[i] = v / d[i];
    }
    double nrm = 0.0;
    for (int i = 0; i < k; ++i) nrm = std::max(nrm, std::fabs(x[i]));
    for (int i = 0; i < k; ++i) x[i] /= nrm;
  }
  double nrm2 = 0.0;
  for (int i = 0; i < k; ++i) nrm2 += x[i] * x[i];
  const double r = beta * std::fabs(x[k - 1]) / std::sqrt(nrm2);
  if (quad) {
    const double gap = tridiag_eig(a, b, k, 1) - theta;
    *quad = gap > 0.0 ? std::min(r, r * r / gap) : r;
  }
  return r;
}

}  // namespace qes
