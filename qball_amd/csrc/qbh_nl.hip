// Kleinman-Bylander projector contractions (NonLocalPotential::energy, norm-conserving
// branch, src/qball/NonLocalPotential.cc:1909-2171) on gfx950.
//
//   anl[p](G) = (-i)^l twnl_p(|k+G-A|) exp(-i (k+G-A).tau_a)        (comp_eigr/comp_anl)
//   F = anl^T psi  (Γ: real view over 2*ngw rows, G=0 counted once, x2)   (fnl_gemm)
//   Enl += occ * wt/Omega * |F|^2 ;  F *= wt/Omega                       (:2106-2148)
//   hpsi += anl F                                                         (enl_hpsi)
//
// The projector table anl is generated once per geometry / vector potential and kept in
// HBM (2*ngw*P doubles; 2.4 GB for 512 Si atoms), so an Hψ call streams it twice
// through the two GEMMs.  For the real Γ view they are this library's own matrix-core
// kernels (qbh_gemm.hip; F split over K slices); complex bases use rocBLAS
// (zgemm/cgemm, F split over K slices).  rocBLAS and hipBLASLt (a tuned algorithm over the
// full K) remain selectable for the Γ view as tuning choices.
#include <hip/hip_runtime.h>
#include <hipblaslt/hipblaslt.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/qbh.h"
#include "qbh_gemm.h"
#include "qbh_ham.h"

namespace qbh {

void* grow(void*& p, size_t& have, size_t need) {
  if (need <= have && p) return p;
  if (t_capturing) return nullptr;  // no free / malloc under graph capture
  if (p) hipFree(p);
  p = nullptr;
  have = 0;
  if (hipMalloc(&p, need) != hipSuccess) {
    p = nullptr;
    return nullptr;
  }
  have = need;
  return p;
}

int blas_handle(Ctx* c, void** out) {
  if (!c->blas) {
    rocblas_handle h;
    if (rocblas_create_handle(&h) != rocblas_status_success) return QBH_ERR_BLAS;
    c->blas = h;
  }
  rocblas_set_stream(static_cast<rocblas_handle>(c->blas), c->stream);
  *out = c->blas;
  return 0;
}

// hipBLASLt state of a context: the handle, one workspace, and per GEMM shape the
// descriptors and the algorithms the heuristic lists for it (in the order listed, so a
// tuned index names the same algorithm for the same shape and library)
struct LtPlan {
  hipblasLtMatmulDesc_t op = nullptr;
  hipblasLtMatrixLayout_t a = nullptr, b = nullptr, c = nullptr;
  std::vector<hipblasLtMatmulHeuristicResult_t> cand;
};
using LtKey = std::array<long, 8>;  // fp64, transA, m, n, k, lda, ldb, ldc
struct LtState {
  hipblasLtHandle_t h = nullptr;
  bool failed = false;  // handle creation failed: every shape uses rocBLAS
  int version = 0;      // hipblasLtGetVersion of the library loaded in this process
  void* ws = nullptr;
  size_t ws_bytes = 0;
  std::map<LtKey, LtPlan> plans;
};
constexpr size_t kLtWorkspaceCap = size_t(256) << 20;
constexpr int kLtCandidates = 16;

void blas_destroy(Ctx* c) {
  if (c->blas) rocblas_destroy_handle(static_cast<rocblas_handle>(c->blas));
  c->blas = nullptr;
  if (c->lt) {
    LtState* lt = static_cast<LtState*>(c->lt);
    for (auto& kv : lt->plans) {
      if (kv.second.a) hipblasLtMatrixLayoutDestroy(kv.second.a);
      if (kv.second.b) hipblasLtMatrixLayoutDestroy(kv.second.b);
      if (kv.second.c) hipblasLtMatrixLayoutDestroy(kv.second.c);
      if (kv.second.op) hipblasLtMatmulDescDestroy(kv.second.op);
    }
    if (lt->ws) hipFree(lt->ws);
    if (lt->h) hipblasLtDestroy(lt->h);
    delete lt;
    c->lt = nullptr;
  }
}

static bool lt_ready(Ctx* c) {
  if (!c->lt) c->lt = new LtState();
  LtState* lt = static_cast<LtState*>(c->lt);
  if (lt->failed) return false;
  if (!lt->h) {
    if (hipblasLtCreate(&lt->h) != HIPBLAS_STATUS_SUCCESS) {
      lt->h = nullptr;
      lt->failed = true;
      return false;
    }
    if (hipblasLtGetVersion(lt->h, &lt->version) != HIPBLAS_STATUS_SUCCESS) lt->version = 0;
  }
  return true;
}

// Built-in projector GEMM choice: hipBLASLt's second listed algorithm where the loaded
// hipBLASLt is 1.2 or newer (ROCm 7.2), rocBLAS otherwise.  On MI355X with hipBLASLt 1.2
// (512-atom Si, FP64 and FP32) every one of the 16 listed algorithms beat rocBLAS on both
// products and the second was within 1% of the fastest; an older hipBLASLt in the same
// process (as bundled with a PyTorch build for ROCm 7.0) listed 16x16-tile FP64 kernels
// about 5x slower than rocBLAS for the same shapes.
int nl_lt_builtin_algo(Ctx* c) {
  if (!lt_ready(c)) return 0;
  return static_cast<LtState*>(c->lt)->version >= 100200 ? 2 : 0;
}

// Built-in choice of both projector products: the matrix-core kernels of qbh_gemm.hip for
// the real Γ view, rocBLAS for complex bases.
int nl_builtin_algo(Ham* h, int which) {
  // complex bases: rocBLAS for Hψ += βF, which ran 0.145 against 0.216 ms per call (Si64
  // TD, beside the concurrent transforms) while F = βᴴψ ran 0.137 against 0.226 ms here
  return (which == 1 && !h->b->g.real) ? 0 : kOwnGemm;
}

Ctx* ctx_of(qbh_ctx c);

// D (m x n, ld ldc) = op(A) B + beta D over real FP64 or FP32 column-major operands, op(A) =
// A^T (A k x m, ld lda) or A (m x k); the plan is built on first use of the shape.  Returns
// nullptr when hipBLASLt lists no algorithm for it.
static LtPlan* lt_plan(Ctx* c, bool f64, bool ta, long m, long n, long k, long lda, long ldb,
                       long ldc) {
  if (!c->lt) c->lt = new LtState();
  LtState* lt = static_cast<LtState*>(c->lt);
  if (lt->failed) return nullptr;
  if (!lt_ready(c)) return nullptr;
  const LtKey key = {f64 ? 1L : 0L, ta ? 1L : 0L, m, n, k, lda, ldb, ldc};
  auto it = lt->plans.find(key);
  if (it != lt->plans.end()) return it->second.cand.empty() ? nullptr : &it->second;
  LtPlan& p = lt->plans[key];
  const hipDataType dt = f64 ? HIP_R_64F : HIP_R_32F;
  const hipblasOperation_t opa = ta ? HIPBLAS_OP_T : HIPBLAS_OP_N, opb = HIPBLAS_OP_N;
  bool ok = hipblasLtMatmulDescCreate(&p.op, f64 ? HIPBLAS_COMPUTE_64F : HIPBLAS_COMPUTE_32F,
                                      dt) == HIPBLAS_STATUS_SUCCESS;
  ok = ok && hipblasLtMatmulDescSetAttribute(p.op, HIPBLASLT_MATMUL_DESC_TRANSA, &opa,
                                             sizeof(opa)) == HIPBLAS_STATUS_SUCCESS;
  ok = ok && hipblasLtMatmulDescSetAttribute(p.op, HIPBLASLT_MATMUL_DESC_TRANSB, &opb,
                                             sizeof(opb)) == HIPBLAS_STATUS_SUCCESS;
  ok = ok && hipblasLtMatrixLayoutCreate(&p.a, dt, ta ? k : m, ta ? m : k, lda) ==
                 HIPBLAS_STATUS_SUCCESS;
  ok = ok && hipblasLtMatrixLayoutCreate(&p.b, dt, k, n, ldb) == HIPBLAS_STATUS_SUCCESS;
  ok = ok && hipblasLtMatrixLayoutCreate(&p.c, dt, m, n, ldc) == HIPBLAS_STATUS_SUCCESS;
  if (ok) {
    hipblasLtMatmulPreference_t pref;
    if (hipblasLtMatmulPreferenceCreate(&pref) == HIPBLAS_STATUS_SUCCESS) {
      uint64_t cap = kLtWorkspaceCap;
      hipblasLtMatmulHeuristicResult_t res[kLtCandidates];
      int nres = 0;
      if (hipblasLtMatmulPreferenceSetAttribute(pref, HIPBLASLT_MATMUL_PREF_MAX_WORKSPACE_BYTES,
                                                &cap, sizeof(cap)) == HIPBLAS_STATUS_SUCCESS &&
          hipblasLtMatmulAlgoGetHeuristic(lt->h, p.op, p.a, p.b, p.c, p.c, pref, kLtCandidates,
                                          res, &nres) == HIPBLAS_STATUS_SUCCESS)
        for (int i = 0; i < std::min(nres, kLtCandidates); ++i)
          if (res[i].workspaceSize <= kLtWorkspaceCap) p.cand.push_back(res[i]);
      hipblasLtMatmulPreferenceDestroy(pref);
    }
  }
  return p.cand.empty() ? nullptr : &p;
}

// runs candidate `algo` of plan p on the context's stream (D = C in place)
template <typename T>
static int lt_gemm(Ctx* c, LtPlan* p, int algo, const T* A, const T* B, T* C, T beta) {
  LtState* lt = static_cast<LtState*>(c->lt);
  const hipblasLtMatmulHeuristicResult_t& r = p->cand[algo];
  if (r.workspaceSize > lt->ws_bytes && !grow(lt->ws, lt->ws_bytes, r.workspaceSize)) {
    c->last_error = "hipBLASLt workspace allocation failed";
    return QBH_ERR_ALLOC;
  }
  const T one = T(1);
  const hipblasStatus_t st = hipblasLtMatmul(lt->h, p->op, &one, A, p->a, B, p->b, &beta, C, p->c,
                                             C, p->c, &r.algo, lt->ws, lt->ws_bytes, c->stream);
  if (st == HIPBLAS_STATUS_SUCCESS) return 0;
  c->last_error = "hipBLASLt matmul failed (status " + std::to_string((int)st) + ")";
  return QBH_ERR_BLAS;
}

// ------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
k_anl(int ngw, const int* __restrict__ col_l, const int* __restrict__ col_tw,
      const int* __restrict__ col_atom, const double* __restrict__ twnl,
      const double* __restrict__ tau, const double* __restrict__ kpgx, double ax, double ay,
      double az, T* __restrict__ anl) {
  const int ig = blockIdx.x * 256 + threadIdx.x;
  const int p = blockIdx.y;
  if (ig >= ngw) return;
  const int l = col_l[p];
  const double t = twnl[(long)col_tw[p] * ngw + ig];
  const int ia = col_atom[p];
  const double kx = kpgx[ig] - ax, ky = kpgx[ngw + ig] - ay, kz = kpgx[2 * ngw + ig] - az;
  const double arg = -(kx * tau[3 * ia] + ky * tau[3 * ia + 1] + kz * tau[3 * ia + 2]);
  double s, c;
  sincos(arg, &s, &c);
  double re, im;
  switch (l) {  // NonLocalPotential.cc:2000-2034
    case 0: re = t * c; im = t * s; break;
    case 1: re = t * s; im = -t * c; break;
    case 2: re = -t * c; im = -t * s; break;
    default: re = -t * s; im = t * c; break;
  }
  T* a = anl + (long)p * 2 * ngw + 2 * ig;
  a[0] = T(re);
  a[1] = T(im);
}

// Γ finish: F = 2 (F - 0.5 anl(G=0) psi(G=0)) (dger + x2, NonLocalPotential.cc:2071-2102),
// Enl partial sums, F *= wt/Omega.  anl / psi are the real views (stride 2ngw / bstride).
template <typename T>
__global__ void __launch_bounds__(256)
k_fnl_finish_real(int P, int nst, T* __restrict__ F, int nsplit, long pstride,
                  const T* __restrict__ anl, long astride, const T* __restrict__ psi, long bstride,
                  const double* __restrict__ fac, const double* __restrict__ occ,
                  double* __restrict__ partial) {
  __shared__ double red[256];
  const long total = (long)P * nst;
  double e = 0.0;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < total; i += (long)gridDim.x * 256) {
    const int p = (int)(i % P);
    const int n = (int)(i / P);
    const T a0 = anl[(long)p * astride];
    const T c0 = psi[(long)n * bstride];
    // split-K partial sums, in slice order; four slices' loads in flight at a time
    T f = F[i];
    int k = 1;
    for (; k + 3 < nsplit; k += 4) {
      const T a0 = F[k * pstride + i], a1 = F[(k + 1) * pstride + i];
      const T a2 = F[(k + 2) * pstride + i], a3 = F[(k + 3) * pstride + i];
      f += a0;
      f += a1;
      f += a2;
      f += a3;
    }
    for (; k < nsplit; ++k) f += F[k * pstride + i];
    f = f + a0 * (T(-0.5) * c0);
    f = T(2) * f;
    const double fp = fac[p];
    const double fd = double(f);
    e += (fp * occ[n]) * fd * fd;
    F[i] = T(fp) * f;
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// complex finish: Enl partial sums of occ*wt/Omega*|F|^2, F *= wt/Omega.
template <typename T>
__global__ void __launch_bounds__(256)
k_fnl_finish_cplx(int P, int nst, T* __restrict__ F, int nsplit, long pstride,
                  const double* __restrict__ fac, const double* __restrict__ occ,
                  double* __restrict__ partial) {
  __shared__ double red[256];
  const long total = (long)P * nst;
  double e = 0.0;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < total; i += (long)gridDim.x * 256) {
    const int p = (int)(i % P);
    const int n = (int)(i / P);
    const double fp = fac[p];
    // split-K partial sums, in slice order; four slices' loads in flight at a time
    T xr = F[2 * i], xi = F[2 * i + 1];
    int k = 1;
    for (; k + 3 < nsplit; k += 4) {
      T r[4], m[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        r[u] = F[(k + u) * pstride + 2 * i];
        m[u] = F[(k + u) * pstride + 2 * i + 1];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        xr += r[u];
        xi += m[u];
      }
    }
    for (; k < nsplit; ++k) {
      xr += F[k * pstride + 2 * i];
      xi += F[k * pstride + 2 * i + 1];
    }
    const double fr = double(xr), fi = double(xi);
    e += (fp * occ[n]) * (fr * fr + fi * fi);
    F[2 * i] = T(fp) * xr;
    F[2 * i + 1] = T(fp) * xi;
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_sum_partials(int n, double* __restrict__ partial) {
  __shared__ double red[256];
  double e = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) e += partial[i];
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[n] = red[0];
}

// psi (FP64, ld complex) -> compact FP32 copy (2*ngw floats per band)
// out (FP32 compact, 2*ngw floats per band) = psi (FP64, leading dimension ld complex);
// grid (ceil(2*ngw/1024), nst): each thread converts 4 consecutive doubles
__global__ void __launch_bounds__(256)
k_to_f32(int ngw, const double* __restrict__ psi, long ld, float* __restrict__ out) {
  const long n = blockIdx.y;
  const int r = (blockIdx.x * 256 + threadIdx.x) * 4;
  const int m = 2 * ngw;
  if (r >= m) return;
  const double* src = psi + n * 2 * ld + r;
  float* dst = out + n * (long)m + r;
  if (r + 4 <= m) {  // 2*ngw and 2*ld are even, so the pairs are 16-byte aligned
    const double2 a = *reinterpret_cast<const double2*>(src);
    const double2 b = *reinterpret_cast<const double2*>(src + 2);
    float2 lo = make_float2(float(a.x), float(a.y)), hi = make_float2(float(b.x), float(b.y));
    *reinterpret_cast<float2*>(dst) = lo;
    *reinterpret_cast<float2*>(dst + 2) = hi;
  } else {
    for (int k = r; k < m; ++k) out[n * (long)m + k] = float(psi[n * 2 * ld + k]);
  }
}

// hpsi (FP64, leading dimension ldh) += tmp (FP32 compact); grid as k_to_f32
__global__ void __launch_bounds__(256)
k_add_f32(int ngw, const float* __restrict__ tmp, double* __restrict__ h, long ldh) {
  const long n = blockIdx.y;
  const int r = (blockIdx.x * 256 + threadIdx.x) * 4;
  const int m = 2 * ngw;
  if (r >= m) return;
  const float* src = tmp + n * (long)m + r;
  double* dst = h + n * 2 * ldh + r;
  if (r + 4 <= m) {
    const float2 lo = *reinterpret_cast<const float2*>(src);
    const float2 hi = *reinterpret_cast<const float2*>(src + 2);
    double2 a = *reinterpret_cast<double2*>(dst);
    double2 b = *reinterpret_cast<double2*>(dst + 2);
    a.x += double(lo.x);
    a.y += double(lo.y);
    b.x += double(hi.x);
    b.y += double(hi.y);
    *reinterpret_cast<double2*>(dst) = a;
    *reinterpret_cast<double2*>(dst + 2) = b;
  } else {
    for (int k = r; k < m; ++k) h[n * 2 * ldh + k] += double(tmp[n * (long)m + k]);
  }
}

// ------------------------------------------------------------------------------------
static int blas_ok(Ctx* c, rocblas_status s, const char* what) {
  if (s == rocblas_status_success) return 0;
  c->last_error = std::string(what) + ": " + rocblas_status_to_string(s);
  return QBH_ERR_BLAS;
}

// the context's rocBLAS handle on its current stream, created on first use (only the
// rocBLAS choices and complex bases need one)
static int blas_rocblas(Ctx* c, rocblas_handle* out) {
  void* hb = nullptr;
  if (int r = blas_handle(c, &hb)) {
    c->last_error = "rocBLAS handle creation failed";
    return r;
  }
  *out = static_cast<rocblas_handle>(hb);
  return 0;
}

static int gemm_fail(Ctx* c, const char* what) {
  c->last_error = std::string(what) + " product: invalid sizes for the matrix-core kernel";
  return QBH_ERR_ARG;
}

int build_anl_c64(Ham* h, const double* rows, const double* A, double* out) {
  const int ngw = h->b->g.ngw;
  if (h->P == 0) return 0;
  dim3 grid((ngw + 255) / 256, h->P);
  k_anl<double><<<grid, 256, 0, h->ctx->stream>>>(ngw, h->col_l, h->col_tw, h->col_atom, rows,
                                                  h->tau, h->b->kpgx, A[0], A[1], A[2], out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    h->ctx->last_error = std::string("anl: ") + hipGetErrorString(e);
    return QBH_ERR_DEVICE;
  }
  return 0;
}

static int ensure_anl(Ham* h, bool fp32) {
  Ctx* c = h->ctx;
  const int ngw = h->b->g.ngw;
  const size_t n = 2 * (size_t)ngw * h->P;
  if (t_capturing && ((!h->anl64 && !h->anl64_ok) || (fp32 && !h->anl32 && !h->anl32_ok))) {
    c->last_error = "projector table would be allocated during graph capture";
    return QBH_ERR_ALLOC;
  }
  if (!h->anl64_ok) {
    if (!h->anl64 && hipMalloc(&h->anl64, n * sizeof(double)) != hipSuccess) {
      h->anl64 = nullptr;
      c->last_error = "anl allocation failed";
      return QBH_ERR_ALLOC;
    }
    TimedLaunch tl(c, "anl");
    dim3 grid((ngw + 255) / 256, h->P);
    k_anl<double><<<grid, 256, 0, c->stream>>>(ngw, h->col_l, h->col_tw, h->col_atom, h->twnl,
                                               h->tau, h->b->kpgx, h->A[0], h->A[1], h->A[2],
                                               static_cast<double*>(h->anl64));
    h->anl64_ok = true;
    h->anlc64_ok = false;
  }
  if (fp32 && !h->anl32_ok) {
    if (!h->anl32 && hipMalloc(&h->anl32, n * sizeof(float)) != hipSuccess) {
      h->anl32 = nullptr;
      c->last_error = "anl32 allocation failed";
      return QBH_ERR_ALLOC;
    }
    TimedLaunch tl(c, "anl");
    dim3 grid((ngw + 255) / 256, h->P);
    k_anl<float><<<grid, 256, 0, c->stream>>>(ngw, h->col_l, h->col_tw, h->col_atom, h->twnl,
                                              h->tau, h->b->kpgx, h->A[0], h->A[1], h->A[2],
                                              static_cast<float*>(h->anl32));
    h->anl32_ok = true;
    h->anlc32_ok = false;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    c->last_error = std::string("anl: ") + hipGetErrorString(e);
    return QBH_ERR_DEVICE;
  }
  return 0;
}

// Complex bases on this library's (real) matrix-core kernels: with rows 2p = beta_p and
// 2p + 1 = i beta_p of 2ngw interleaved reals, F = beta^H psi is one real product (row 2p:
// Re = br pr + bi pi, row 2p + 1: Im = br pi - bi pr, which lands in the interleaved
// complex layout of F), and psi += beta F is one real product over q = (p, re/im) with the
// same rows (Re = br Fr - bi Fi, Im = bi Fr + br Fi).
template <typename T>
__global__ void __launch_bounds__(256)
k_anl_pair(int ngw, const T* __restrict__ anl, T* __restrict__ out) {
  const int ig = blockIdx.x * 256 + threadIdx.x;
  if (ig >= ngw) return;
  const long p = blockIdx.y;
  const T re = anl[p * 2 * ngw + 2 * ig], im = anl[p * 2 * ngw + 2 * ig + 1];
  T* a = out + 2 * p * 2 * ngw + 2 * ig;
  a[0] = re;
  a[1] = im;
  a[2L * ngw] = -im;
  a[2L * ngw + 1] = re;
}

int anl_pair_c64(Ctx* c, int ngw, int rows, const double* anl, double* out) {
  dim3 grid((ngw + 255) / 256, rows);
  k_anl_pair<double><<<grid, 256, 0, c->stream>>>(ngw, anl, out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    c->last_error = std::string("anl pair: ") + hipGetErrorString(e);
    return QBH_ERR_DEVICE;
  }
  return 0;
}

// the combined table of a complex basis (after ensure_anl)
static int ensure_anlc(Ham* h, bool fp32) {
  Ctx* c = h->ctx;
  const int ngw = h->b->g.ngw;
  const size_t n = 4 * (size_t)ngw * h->P;
  void*& tab = fp32 ? h->anlc32 : h->anlc64;
  bool& ok = fp32 ? h->anlc32_ok : h->anlc64_ok;
  if (ok) return 0;
  if (t_capturing) {
    c->last_error = "projector table would be built during graph capture";
    return QBH_ERR_ALLOC;
  }
  if (!tab && hipMalloc(&tab, n * (fp32 ? sizeof(float) : sizeof(double))) != hipSuccess) {
    tab = nullptr;
    c->last_error = "combined projector table allocation failed";
    return QBH_ERR_ALLOC;
  }
  TimedLaunch tl(c, "anl");
  dim3 grid((ngw + 255) / 256, h->P);
  if (fp32)
    k_anl_pair<float><<<grid, 256, 0, c->stream>>>(ngw, static_cast<const float*>(h->anl32),
                                                   static_cast<float*>(tab));
  else
    k_anl_pair<double><<<grid, 256, 0, c->stream>>>(ngw, static_cast<const double*>(h->anl64),
                                                    static_cast<double*>(tab));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    c->last_error = std::string("anl pair: ") + hipGetErrorString(e);
    return QBH_ERR_DEVICE;
  }
  ok = true;
  return 0;
}

// rocBLAS dispatch by element type (real: T, complex: interleaved T pairs)
static rocblas_status gemm_sb(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m,
                              int n, int k, const double* A, long lda, long sa, const double* B,
                              long ldb, long sb, double* C, long ldc, long sc, int batch,
                              bool cplx, double beta) {
  if (!cplx) {
    const double one = 1.0;
    return rocblas_dgemm_strided_batched(h, ta, tb, m, n, k, &one, A, lda, sa, B, ldb, sb, &beta, C,
                                         ldc, sc, batch);
  }
  const rocblas_double_complex one(1.0, 0.0), b(beta, 0.0);
  return rocblas_zgemm_strided_batched(
      h, ta, tb, m, n, k, &one, reinterpret_cast<const rocblas_double_complex*>(A), lda, sa,
      reinterpret_cast<const rocblas_double_complex*>(B), ldb, sb, &b,
      reinterpret_cast<rocblas_double_complex*>(C), ldc, sc, batch);
}
static rocblas_status gemm_sb(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m,
                              int n, int k, const float* A, long lda, long sa, const float* B,
                              long ldb, long sb, float* C, long ldc, long sc, int batch, bool cplx,
                              double beta) {
  if (!cplx) {
    const float one = 1.f, bt = float(beta);
    return rocblas_sgemm_strided_batched(h, ta, tb, m, n, k, &one, A, lda, sa, B, ldb, sb, &bt, C,
                                         ldc, sc, batch);
  }
  const rocblas_float_complex one(1.f, 0.f), b(float(beta), 0.f);
  return rocblas_cgemm_strided_batched(
      h, ta, tb, m, n, k, &one, reinterpret_cast<const rocblas_float_complex*>(A), lda, sa,
      reinterpret_cast<const rocblas_float_complex*>(B), ldb, sb, &b,
      reinterpret_cast<rocblas_float_complex*>(C), ldc, sc, batch);
}

// Number of K-splits for F = anl^H psi: the P x nst output alone gives too few 128x128
// tiles to fill 256 CUs when P, nst ~ 1e3 and K ~ 1e5, so K is cut into slices whose
// partial products are summed in the finish kernel.
// Measured on MI355X for P=2048, nst=1024, K=145838 (FP64): 8 slices 51 TFLOP/s, 24-32
// slices 76 TFLOP/s; the target is ~4096 output tiles, at least 1024 rows per slice.
static int fnl_splits(int P, int nst, long K) {
  const long tiles = (long)((P + 127) / 128) * ((nst + 127) / 128);
  long s = (4096 + tiles - 1) / tiles;
  s = std::min<long>(s, 32);
  s = std::min<long>(s, std::max<long>(1, K / 1024));
  return (int)std::max<long>(1, s);
}

// F (nsplit partials, stride P*nst elements) = anl^T psi (real view) or anl^H psi.
// Kelem = contraction length in elements (2*ngw reals or ngw complex).
template <typename T>
static int fnl_gemm(Ctx* c, rocblas_handle bh, bool cplx, int P, int nst, long Kelem, const T* anl,
                    long lda, const T* psi, long ldb, T* F, long kc) {
  const rocblas_operation ta =
      cplx ? rocblas_operation_conjugate_transpose : rocblas_operation_transpose;
  const int e = cplx ? 2 : 1;  // T per element
  const long nfull = Kelem / kc;
  const long rem = Kelem - nfull * kc;
  const long pstride = (long)P * nst * e;
  rocblas_status st = rocblas_status_success;
  // batch strides are in BLAS elements (complex elements for zgemm/cgemm); pointer
  // offsets are in T
  if (nfull > 0)
    st = gemm_sb(bh, ta, rocblas_operation_none, P, nst, (int)kc, anl, lda, kc, psi, ldb, kc, F,
                 P, (long)P * nst, (int)nfull, cplx, 0.0);
  if (st == rocblas_status_success && rem > 0)
    st = gemm_sb(bh, ta, rocblas_operation_none, P, nst, (int)rem, anl + nfull * kc * e, lda, 0,
                 psi + nfull * kc * e, ldb, 0, F + nfull * pstride, P, (long)P * nst, 1, cplx,
                 0.0);
  return blas_ok(c, st, "fnl gemm");
}

// hipBLASLt plans of the two projector GEMMs of an Hψ call (real Γ view only): fnl is
// F (P x nst) = anl^T psi over K = 2 ngw rows, enl is hpsi (K x nst) += anl F (FP64, into
// hpsi) or tmp = anl F (FP32, into the compact product buffer).  nullptr: use rocBLAS.
static LtPlan* nl_plan(Ham* h, int which, int nst, long ld, long ldh, bool fp32) {
  if (!h->b->g.real || h->P == 0) return nullptr;
  const long K = 2L * h->b->g.ngw, P = h->P;
  if (which == 0) return lt_plan(h->ctx, !fp32, true, P, nst, K, K, fp32 ? K : 2 * ld, P);
  return lt_plan(h->ctx, !fp32, false, K, nst, P, K, P, fp32 ? K : 2 * ldh);
}

int nl_lt_candidates(Ham* h, int which, int nst, long ld, long ldh, int fp32) {
  LtPlan* p = nl_plan(h, which, nst, ld, ldh, fp32 != 0);
  return p ? (int)p->cand.size() : 0;
}

// Occupations on the device (h->red + kRedBlocks + 1, uploaded when they change) after
// the Enl reduction partials
constexpr int kRedBlocks = 512;
static int nl_occ(Ham* h, int nst, const double* occ_host, const double** occ) {
  Ctx* c = h->ctx;
  if (h->red_n < (size_t)(kRedBlocks + 1 + nst)) {
    if (t_capturing) {
      c->last_error = "reduction scratch would be allocated during graph capture";
      return QBH_ERR_ALLOC;
    }
    if (h->red) hipFree(h->red);
    h->red = nullptr;
    h->red_n = 0;
    h->occ_host.clear();
    if (hipMalloc(&h->red, (kRedBlocks + 1 + nst) * sizeof(double)) != hipSuccess) {
      c->last_error = "reduction scratch allocation failed";
      return QBH_ERR_ALLOC;
    }
    h->red_n = kRedBlocks + 1 + nst;
  }
  double* d = h->red + kRedBlocks + 1;
  // occupations change rarely: upload only when they differ from the device copy
  std::vector<double> o(nst, 0.0);
  if (occ_host)
    for (int n = 0; n < nst; ++n) o[n] = occ_host[n];
  if (o != h->occ_host) {
    if (t_capturing) {
      c->last_error = "occupations would be uploaded (a synchronising copy) during graph capture";
      return QBH_ERR_UNSUPPORTED;
    }
    hipMemcpyAsync(d, o.data(), nst * sizeof(double), hipMemcpyHostToDevice, c->stream);
    hipStreamSynchronize(c->stream);  // o is a stack buffer
    h->occ_host = o;
  }
  *occ = d;
  return 0;
}

// The projector coefficients of an Hψ call: F = anl^T psi over the real view (Γ) or
// anl^H psi, finished (K-slice partials summed; Γ: G=0 term and x2), with Enl partials in
// h->red and F scaled by wt/Omega (NonLocalPotential.cc:2040-2148).  FP32: psi is first
// converted into h->psi32 and F is FP32 (h->f32), else FP64 (h->f64).
struct FnlStage {
  void* F = nullptr;      // P x nst (complex: interleaved) in slot 0 of the partials
  int sel = 0;            // fnl_algo value in force (0: rocBLAS over K slices)
};

static int fnl_stage(Ham* h, rocblas_handle bh, const double* psi, long ld, int nst, long ldh,
                     bool fp32, const double* occ, FnlStage* o) {
  Ctx* c = h->ctx;
  const int ngw = h->b->g.ngw;
  const bool real = h->b->g.real;
  const bool cplx = !real;
  const int P = h->P;
  const long Kelem = real ? 2L * ngw : (long)ngw;  // contraction length in elements
  const int e = cplx ? 2 : 1;
  // hipBLASLt algorithm chosen for this shape (nullptr: rocBLAS)
  o->sel = h->fnl_algo >= 0 ? h->fnl_algo : nl_builtin_algo(h, 0);
  const bool own = o->sel == kOwnGemm;
  // own kernels: rows Ro of the real view (complex: rows (p, re/im) of the combined table)
  // over Qo = 2 ngw reals
  const int Ro = real ? P : 2 * P;
  const long Qo = 2L * ngw;
  if (own && cplx) {
    if (int r = ensure_anlc(h, fp32)) return r;
  }
  LtPlan* fplan = (o->sel > 0 && !own) ? nl_plan(h, 0, nst, ld, ldh, fp32) : nullptr;
  const int falgo = fplan ? std::min<int>(o->sel, (int)fplan->cand.size()) - 1 : -1;
  // K slices of kc elements; the slice count actually written (ceil(K / kc)) can be below
  // the requested split count, and the finish kernel sums exactly that many partials.
  // hipBLASLt contracts the full K in one product.
  // the own kernel cuts K into its own slices (multiples of its 16-row stage)
  const int want = fplan ? 1
                         : own ? mfma_slices(Ro, nst, Qo, h->nsplit)
                               : (h->nsplit > 0 ? h->nsplit : fnl_splits(P, nst, Kelem));
  const long kc = own ? mfma_slice_len(Qo, want) : (Kelem + want - 1) / want;
  const int nsplit = own ? (int)((Qo + kc - 1) / kc) : (int)((Kelem + kc - 1) / kc);
  const long pstride = (long)P * nst * e;
  const size_t fsz = (size_t)pstride * nsplit;
  if (!own && !bh) {
    if (int r = blas_rocblas(c, &bh)) return r;
  }
  if (!fp32) {
    double* F = static_cast<double*>(grow(h->f64, h->f64_bytes, fsz * sizeof(double)));
    const double* anl = static_cast<const double*>(h->anl64);
    if (!F) {
      c->last_error = "projector coefficient allocation failed";
      return QBH_ERR_ALLOC;
    }
    {
      TimedLaunch tl(c, "fnl_gemm");
      if (own) {
        const double* tab = real ? anl : static_cast<const double*>(h->anlc64);
        if (mfma_gemm_qq(c->stream, false, Ro, nst, Qo, tab, Qo, psi, 2 * ld, F, Ro, nsplit,
                         pstride))
          return gemm_fail(c, "fnl");
      } else if (int r = fplan ? lt_gemm<double>(c, fplan, falgo, anl, psi, F, 0.0)
                               : fnl_gemm<double>(c, bh, cplx, P, nst, Kelem, anl, Kelem, psi,
                                                  real ? 2 * ld : ld, F, kc)) {
        return r;
      }
    }
    TimedLaunch tl(c, "fnl_finish");
    if (real)
      k_fnl_finish_real<double><<<kRedBlocks, 256, 0, c->stream>>>(
          P, nst, F, nsplit, pstride, anl, 2L * ngw, psi, 2L * ld, h->col_fac, occ, h->red);
    else
      k_fnl_finish_cplx<double><<<kRedBlocks, 256, 0, c->stream>>>(P, nst, F, nsplit, pstride,
                                                                   h->col_fac, occ, h->red);
    o->F = F;
  } else {
    float* F = static_cast<float*>(grow(h->f32, h->f32_bytes, fsz * sizeof(float)));
    const size_t psz = 2 * (size_t)ngw * nst;
    float* p32 = static_cast<float*>(grow(h->psi32, h->psi32_bytes, psz * sizeof(float)));
    const float* anl = static_cast<const float*>(h->anl32);
    if (!F || !p32) {
      c->last_error = "FP32 projector scratch allocation failed";
      return QBH_ERR_ALLOC;
    }
    {
      TimedLaunch tl(c, "to_f32");
      k_to_f32<<<dim3((unsigned)((2 * ngw + 1023) / 1024), (unsigned)nst), 256, 0, c->stream>>>(
          ngw, psi, ld, p32);
    }
    {
      TimedLaunch tl(c, "fnl_gemm");
      if (own) {
        const float* tab = real ? anl : static_cast<const float*>(h->anlc32);
        if (mfma_gemm_qq(c->stream, true, Ro, nst, Qo, tab, Qo, p32, Qo, F, Ro, nsplit, pstride))
          return gemm_fail(c, "fnl");
      } else if (int r = fplan ? lt_gemm<float>(c, fplan, falgo, anl, p32, F, 0.0f)
                               : fnl_gemm<float>(c, bh, cplx, P, nst, Kelem, anl, Kelem, p32,
                                                 Kelem, F, kc)) {
        return r;
      }
    }
    TimedLaunch tl(c, "fnl_finish");
    if (real)
      k_fnl_finish_real<float><<<kRedBlocks, 256, 0, c->stream>>>(
          P, nst, F, nsplit, pstride, anl, 2L * ngw, p32, 2L * ngw, h->col_fac, occ, h->red);
    else
      k_fnl_finish_cplx<float><<<kRedBlocks, 256, 0, c->stream>>>(P, nst, F, nsplit, pstride,
                                                                  h->col_fac, occ, h->red);
    o->F = F;
  }
  return 0;
}

int nl_apply(Ham* h, const double* psi, long ld, int nst, const double* occ_host, double* hpsi,
             long ldh, int fp32, double* enl, const float** defer_add32, bool overwrite) {
  Ctx* c = h->ctx;
  if (h->P == 0) {
    if (enl) *enl = 0.0;
    return 0;
  }
  const int ngw = h->b->g.ngw;
  const bool real = h->b->g.real;
  const bool cplx = !real;
  const int P = h->P;
  if (int r = ensure_anl(h, fp32)) return r;
  rocblas_handle bh = nullptr;
  const double* occ;
  if (int r = nl_occ(h, nst, occ_host, &occ)) return r;
  FnlStage fs;
  if (int r = fnl_stage(h, nullptr, psi, ld, nst, ldh, fp32 != 0, occ, &fs)) return r;

  const long Kelem = real ? 2L * ngw : (long)ngw;
  int esel = h->enl_algo >= 0 ? h->enl_algo : nl_builtin_algo(h, 1);
  const bool own = esel == kOwnGemm;
  const int Ro = real ? P : 2 * P;  // own kernels: reduction rows (see fnl_stage)
  const long Qo = 2L * ngw;
  if (own && cplx) {
    if (int r = ensure_anlc(h, fp32 != 0)) return r;
  }
  LtPlan* eplan = (esel > 0 && !own) ? nl_plan(h, 1, nst, ld, ldh, fp32 != 0) : nullptr;
  const int ealgo = eplan ? std::min<int>(esel, (int)eplan->cand.size()) - 1 : -1;
  if (!own && !eplan) {
    if (int r = blas_rocblas(c, &bh)) return r;
  }
  if (!fp32) {
    const double* F = static_cast<const double*>(fs.F);
    const double* anl = static_cast<const double*>(h->anl64);
    TimedLaunch tl(c, "enl_gemm");
    const double beta = overwrite ? 0.0 : 1.0;
    if (own) {
      // hpsi[k, n] (+)= sum_p anl[k, p] F[p, n], taken transposed (r = band, c = row)
      const double* tab = real ? anl : static_cast<const double*>(h->anlc64);
      if (mfma_gemm_qc(c->stream, false, nst, (int)Qo, Ro, F, Ro, tab, Qo, hpsi, 2 * ldh,
                       !overwrite))
        return gemm_fail(c, "enl");
    } else if (eplan) {
      if (int r = lt_gemm<double>(c, eplan, ealgo, anl, F, hpsi, beta)) return r;
    } else if (int r = blas_ok(c,
                               gemm_sb(bh, rocblas_operation_none, rocblas_operation_none,
                                       (int)Kelem, nst, P, anl, Kelem, 0, F, P, 0, hpsi,
                                       real ? 2 * ldh : ldh, 0, 1, cplx, beta),
                               "enl gemm")) {
      return r;
    }
  } else {
    const float* F = static_cast<const float*>(fs.F);
    const size_t psz = 2 * (size_t)ngw * nst;
    float* t32 = static_cast<float*>(grow(h->tmp32, h->tmp32_bytes, psz * sizeof(float)));
    const float* anl = static_cast<const float*>(h->anl32);
    if (!t32) {
      c->last_error = "FP32 projector scratch allocation failed";
      return QBH_ERR_ALLOC;
    }
    {
      TimedLaunch tl(c, "enl_gemm");
      if (own) {
        const float* tab = real ? anl : static_cast<const float*>(h->anlc32);
        if (mfma_gemm_qc(c->stream, true, nst, (int)Qo, Ro, F, Ro, tab, Qo, t32, Qo, false))
          return gemm_fail(c, "enl");
      } else if (eplan) {
        if (int r = lt_gemm<float>(c, eplan, ealgo, anl, F, t32, 0.0f)) return r;
      } else if (int r = blas_ok(c,
                                 gemm_sb(bh, rocblas_operation_none, rocblas_operation_none,
                                         (int)Kelem, nst, P, anl, Kelem, 0, F, P, 0, t32, Kelem,
                                         0, 1, cplx, 0.0),
                                 "enl gemm")) {
        return r;
      }
    }
    if (defer_add32) {
      *defer_add32 = t32;  // the forward z pass adds it while accumulating into hpsi
    } else {
      TimedLaunch tl(c, "add_f32");
      k_add_f32<<<dim3((unsigned)((2 * ngw + 1023) / 1024), (unsigned)nst), 256, 0, c->stream>>>(
          ngw, t32, hpsi, ldh);
    }
  }
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    c->last_error = std::string("nonlocal: ") + hipGetErrorString(err);
    return QBH_ERR_DEVICE;
  }
  if (enl) {
    // Enl is reduced on the device; nl_read_enl fetches it once the stream has drained
    k_sum_partials<<<1, 256, 0, c->stream>>>(kRedBlocks, h->red);
    *enl = 0.0;
  }
  return 0;
}

// ------------------------------------------------------------------------------------
// Non-local ionic forces (NonLocalPotential.cc:2175-2323).  The reference tables, per
// direction j, anl_j = (k+G-A)_j twnl (-i)^(l+1) exp(-i (k+G-A).tau) = -i (k+G-A)_j anl
// and contracts dfnl_j = anl_j^T c (real view, x2 at Γ) or anl_j^H c.  The same numbers
// come from the resident anl table applied to psi'_j = i (k+G-A)_j psi, since
// Re(-i k a)^T Re(c) + Im(-i k a)^T Im(c) = Re(a)^T Re(i k c) + Im(a)^T Im(i k c) and
// (-i k a)^H c = a^H (i k c): one band-sized scaling pass per direction instead of three
// more P x ngw tables.

// psi'_j = i (k+G-A)_j psi, compact (2 ngw reals per band); grid (ceil(ngw/256), nst)
template <typename T>
__global__ void __launch_bounds__(256)
k_dpsi(int ngw, const double* __restrict__ psi, long ld, const double* __restrict__ kpgxj,
       double aj, T* __restrict__ out) {
  const int ig = blockIdx.x * 256 + threadIdx.x;
  if (ig >= ngw) return;
  const long n = blockIdx.y;
  const double2 v = *reinterpret_cast<const double2*>(psi + 2 * (n * ld + ig));
  const double k = kpgxj[ig] - aj;
  T* o = out + 2 * (n * ngw + ig);
  o[0] = T(-k * v.y);
  o[1] = T(k * v.x);
}

// One direction: part[s*P + p] = 2 scale sum_{n in chunk s} occ_n F[p,n] dF[p,n] (complex:
// Re(F conj(dF))), dF first summed over its nsplit K-slice partials.  F holds the wt/Omega
// scaled coefficients, 2 occ_n is facn of :2281-2282, scale is the x2 of dfnl at Γ
// (:2265-2267; there is no G=0 term to remove: the derivative factor vanishes at G=0
// without a vector potential).  Block: 64 columns x 4 band lanes; grid (ceil(P/64), S).
template <typename T, bool CPLX>
__global__ void __launch_bounds__(256)
k_nl_force(int P, int nst, const T* __restrict__ F, const T* __restrict__ dF, int nsplit,
           long pstride, const double* __restrict__ occ, double scale,
           double* __restrict__ part) {
  __shared__ double red[256];
  const int p = blockIdx.x * 64 + (threadIdx.x & 63);
  const int lane = threadIdx.x >> 6;
  double acc = 0.0;
  if (p < P) {
    for (int n = blockIdx.y * 4 + lane; n < nst; n += 4 * gridDim.y) {
      const long i = (long)n * P + p;
      if (!CPLX) {
        T d = dF[i];
        for (int k = 1; k < nsplit; ++k) d += dF[k * pstride + i];
        acc += occ[n] * (double(F[i]) * double(d));
      } else {
        T dr = dF[2 * i], di = dF[2 * i + 1];
        for (int k = 1; k < nsplit; ++k) {
          dr += dF[k * pstride + 2 * i];
          di += dF[k * pstride + 2 * i + 1];
        }
        acc += occ[n] * (double(F[2 * i]) * double(dr) + double(F[2 * i + 1]) * double(di));
      }
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  const int t = threadIdx.x;
  if (t < 64 && p < P)
    part[(long)blockIdx.y * P + p] = 2.0 * scale * ((red[t] + red[t + 64]) + (red[t + 128] + red[t + 192]));
}

// g[p] = sum_s part[s*P + p]
__global__ void __launch_bounds__(256)
k_col_sum(int P, int S, const double* __restrict__ part, double* __restrict__ g) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= P) return;
  double a = 0.0;
  for (int s = 0; s < S; ++s) a += part[(long)s * P + p];
  g[p] = a;
}

template <typename T>
static int dfnl_forces(Ham* h, rocblas_handle bh, const double* psi, long ld, int nst,
                       bool fp32, int sel, const T* anl, const T* F, const double* occ, T* dpsi,
                       double* g) {
  Ctx* c = h->ctx;
  const int ngw = h->b->g.ngw;
  const bool real = h->b->g.real;
  const bool cplx = !real;
  const int P = h->P;
  const long Kelem = real ? 2L * ngw : (long)ngw;
  const int e = cplx ? 2 : 1;
  // psi' is compact (leading dimension ngw); same library choice as the fnl product
  const bool own = sel == kOwnGemm;
  LtPlan* dplan = (sel > 0 && !own) ? nl_plan(h, 0, nst, ngw, ngw, fp32) : nullptr;
  const int dalgo = dplan ? std::min<int>(sel, (int)dplan->cand.size()) - 1 : -1;
  const int Ro = real ? P : 2 * P;  // own kernels: rows of the real view (see fnl_stage)
  const long Qo = 2L * ngw;
  const int want = dplan ? 1
                         : own ? mfma_slices(Ro, nst, Qo, h->nsplit)
                               : (h->nsplit > 0 ? h->nsplit : fnl_splits(P, nst, Kelem));
  const long kc = own ? mfma_slice_len(Qo, want) : (Kelem + want - 1) / want;
  const int nsplit = own ? (int)((Qo + kc - 1) / kc) : (int)((Kelem + kc - 1) / kc);
  if (!own && !dplan && !bh) {
    if (int r = blas_rocblas(c, &bh)) return r;
  }
  const long pstride = (long)P * nst * e;
  T* dF = static_cast<T*>(grow(h->df, h->df_bytes, (size_t)pstride * nsplit * sizeof(T)));
  const int S = std::max(1, std::min(64, (nst + 15) / 16));
  double* part = static_cast<double*>(
      grow(h->fpart, h->fpart_bytes, (size_t)S * P * sizeof(double)));
  if (!dF || !part) {
    c->last_error = "force scratch allocation failed";
    return QBH_ERR_ALLOC;
  }
  for (int j = 0; j < 3; ++j) {
    {
      TimedLaunch tl(c, "dpsi");
      k_dpsi<T><<<dim3((unsigned)((ngw + 255) / 256), (unsigned)nst), 256, 0, c->stream>>>(
          ngw, psi, ld, h->b->kpgx + (long)j * ngw, h->A[j], dpsi);
    }
    {
      TimedLaunch tl(c, "dfnl_gemm");
      if (own) {
        const T* tab = real ? anl
                            : static_cast<const T*>(fp32 ? h->anlc32 : h->anlc64);
        if (mfma_gemm_qq(c->stream, fp32, Ro, nst, Qo, tab, Qo, dpsi, Qo, dF, Ro, nsplit,
                         pstride))
          return gemm_fail(c, "dfnl");
      } else if (int r = dplan ? lt_gemm<T>(c, dplan, dalgo, anl, dpsi, dF, T(0))
                               : fnl_gemm<T>(c, bh, cplx, P, nst, Kelem, anl, Kelem, dpsi, Kelem,
                                             dF, kc)) {
        return r;
      }
    }
    TimedLaunch tl(c, "nl_force");
    const dim3 grid((unsigned)((P + 63) / 64), (unsigned)S);
    if (real)
      k_nl_force<T, false><<<grid, 256, 0, c->stream>>>(P, nst, F, dF, nsplit, pstride, occ, 2.0,
                                                        part);
    else
      k_nl_force<T, true><<<grid, 256, 0, c->stream>>>(P, nst, F, dF, nsplit, pstride, occ, 1.0,
                                                       part);
    k_col_sum<<<(P + 255) / 256, 256, 0, c->stream>>>(P, S, part, g + (long)j * P);
  }
  return 0;
}

int nl_forces(Ham* h, const double* psi, long ld, int nst, const double* occ_host, int fp32,
              double* fion, double* enl) {
  Ctx* c = h->ctx;
  long natoms = 0;
  for (int s = 0; s < h->nsp; ++s) natoms += h->na[s];
  std::fill(fion, fion + 3 * natoms, 0.0);
  if (enl) *enl = 0.0;
  if (h->P == 0) return 0;
  const int ngw = h->b->g.ngw;
  const int P = h->P;
  if (int r = ensure_anl(h, fp32)) return r;
  rocblas_handle bh = nullptr;
  const double* occ;
  if (int r = nl_occ(h, nst, occ_host, &occ)) return r;
  FnlStage fs;
  if (int r = fnl_stage(h, nullptr, psi, ld, nst, ngw, fp32 != 0, occ, &fs)) return r;
  double* g = static_cast<double*>(grow(h->fg, h->fg_bytes, 3 * (size_t)P * sizeof(double)));
  if (!g) {
    c->last_error = "force scratch allocation failed";
    return QBH_ERR_ALLOC;
  }
  const size_t psz = 2 * (size_t)ngw * nst;
  int r;
  if (!fp32) {
    double* dpsi = static_cast<double*>(grow(h->dpsi, h->dpsi_bytes, psz * sizeof(double)));
    if (!dpsi) {
      c->last_error = "force scratch allocation failed";
      return QBH_ERR_ALLOC;
    }
    r = dfnl_forces<double>(h, bh, psi, ld, nst, false, fs.sel,
                            static_cast<const double*>(h->anl64),
                            static_cast<const double*>(fs.F), occ, dpsi, g);
  } else {
    // the FP32 copy of psi is no longer read once F is finished; psi' reuses the product buffer
    float* dpsi = static_cast<float*>(grow(h->tmp32, h->tmp32_bytes, psz * sizeof(float)));
    if (!dpsi) {
      c->last_error = "force scratch allocation failed";
      return QBH_ERR_ALLOC;
    }
    r = dfnl_forces<float>(h, bh, psi, ld, nst, true, fs.sel, static_cast<const float*>(h->anl32),
                           static_cast<const float*>(fs.F), occ, dpsi, g);
  }
  if (r) return r;
  if (enl) k_sum_partials<<<1, 256, 0, c->stream>>>(kRedBlocks, h->red);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    c->last_error = std::string("nonlocal forces: ") + hipGetErrorString(err);
    return QBH_ERR_DEVICE;
  }
  std::vector<double> gh(3 * (size_t)P);
  hipMemcpyAsync(gh.data(), g, gh.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (enl) hipMemcpyAsync(enl, h->red + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, c->stream);
  err = hipStreamSynchronize(c->stream);
  if (err != hipSuccess) {
    c->last_error = std::string("nonlocal forces: ") + hipGetErrorString(err);
    return QBH_ERR_DEVICE;
  }
  // tmpfion[3*ia+j] -= ... summed over the columns of each atom (:2283-2288)
  for (int j = 0; j < 3; ++j)
    for (int p = 0; p < P; ++p) fion[3L * h->col_atom_host[p] + j] -= gh[(size_t)j * P + p];
  return 0;
}

}  // namespace qbh

extern "C" int qbh_ctx_library_info(qbh_ctx ctx, char* buf, int cap) {
  using namespace qbh;
  if (!ctx || !buf || cap <= 0) return QBH_ERR_ARG;
  Ctx* c = ctx_of(ctx);
  hipSetDevice(c->device);
  char rb[128] = {0};
  rocblas_get_version_string(rb, sizeof rb);
  int hv = 0;
  hipRuntimeGetVersion(&hv);
  const int lv = lt_ready(c) ? static_cast<LtState*>(c->lt)->version : 0;
  std::snprintf(buf, (size_t)cap,
                "hip %d rocblas %s hipblaslt %d projector-gemm-builtin mfma (complex-basis enl: "
                "rocblas)", hv, rb, lv);
  return QBH_OK;
}

namespace qbh {

double nl_read_enl(Ham* h, hipStream_t s) {
  double v = 0.0;
  if (!h->red || h->P == 0) return 0.0;
  hipMemcpyAsync(&v, h->red + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, s);
  hipStreamSynchronize(s);
  return v;
}

}  // namespace qbh
