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
// through the two MFMA GEMMs (rocBLAS dgemm/zgemm/sgemm/cgemm).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdio>
#include <string>

#include "../../include/qbh.h"
#include "qbh_ham.h"

namespace qbh {

void* grow(void*& p, size_t& have, size_t need) {
  if (need <= have && p) return p;
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

void blas_destroy(Ctx* c) {
  if (c->blas) rocblas_destroy_handle(static_cast<rocblas_handle>(c->blas));
  c->blas = nullptr;
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
k_fnl_finish_real(int P, int nst, T* __restrict__ F, const T* __restrict__ anl, long astride,
                  const T* __restrict__ psi, long bstride, const double* __restrict__ fac,
                  const double* __restrict__ occ, double* __restrict__ partial) {
  __shared__ double red[256];
  const long total = (long)P * nst;
  double e = 0.0;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < total; i += (long)gridDim.x * 256) {
    const int p = (int)(i % P);
    const int n = (int)(i / P);
    const T a0 = anl[(long)p * astride];
    const T c0 = psi[(long)n * bstride];
    T f = F[i] + a0 * (T(-0.5) * c0);
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
k_fnl_finish_cplx(int P, int nst, T* __restrict__ F, const double* __restrict__ fac,
                  const double* __restrict__ occ, double* __restrict__ partial) {
  __shared__ double red[256];
  const long total = (long)P * nst;
  double e = 0.0;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < total; i += (long)gridDim.x * 256) {
    const int p = (int)(i % P);
    const int n = (int)(i / P);
    const double fp = fac[p];
    const double fr = double(F[2 * i]), fi = double(F[2 * i + 1]);
    e += (fp * occ[n]) * (fr * fr + fi * fi);
    F[2 * i] = T(fp) * F[2 * i];
    F[2 * i + 1] = T(fp) * F[2 * i + 1];
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
__global__ void k_to_f32(int ngw, int nst, const double* __restrict__ psi, long ld,
                         float* __restrict__ out) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  const long total = 2L * ngw * nst;
  if (i >= total) return;
  const long n = i / (2L * ngw);
  const long r = i - n * 2L * ngw;
  out[i] = float(psi[n * 2 * ld + r]);
}

// hpsi (FP64) += tmp (FP32 compact)
__global__ void k_add_f32(int ngw, int nst, const float* __restrict__ tmp, double* __restrict__ h,
                          long ldh) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  const long total = 2L * ngw * nst;
  if (i >= total) return;
  const long n = i / (2L * ngw);
  const long r = i - n * 2L * ngw;
  h[n * 2 * ldh + r] += double(tmp[i]);
}

// ------------------------------------------------------------------------------------
static int blas_ok(Ctx* c, rocblas_status s, const char* what) {
  if (s == rocblas_status_success) return 0;
  c->last_error = std::string(what) + ": " + rocblas_status_to_string(s);
  return QBH_ERR_BLAS;
}

static int ensure_anl(Ham* h, bool fp32) {
  Ctx* c = h->ctx;
  const int ngw = h->b->g.ngw;
  const size_t n = 2 * (size_t)ngw * h->P;
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
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    c->last_error = std::string("anl: ") + hipGetErrorString(e);
    return QBH_ERR_DEVICE;
  }
  return 0;
}

int nl_apply(Ham* h, const double* psi, long ld, int nst, const double* occ_host, double* hpsi,
             long ldh, int fp32, double* enl) {
  Ctx* c = h->ctx;
  if (h->P == 0) {
    if (enl) *enl = 0.0;
    return 0;
  }
  const int ngw = h->b->g.ngw;
  const bool real = h->b->g.real;
  const int P = h->P;
  if (int r = ensure_anl(h, fp32)) return r;
  void* hb;
  if (int r = blas_handle(c, &hb)) return r;
  rocblas_handle bh = static_cast<rocblas_handle>(hb);

  // occupations and reduction scratch
  const int nblk = 512;
  if (h->red_n < (size_t)(nblk + 1 + nst)) {
    if (h->red) hipFree(h->red);
    h->red = nullptr;
    h->red_n = 0;
    h->occ_host.clear();
    if (hipMalloc(&h->red, (nblk + 1 + nst) * sizeof(double)) != hipSuccess) {
      c->last_error = "reduction scratch allocation failed";
      return QBH_ERR_ALLOC;
    }
    h->red_n = nblk + 1 + nst;
  }
  double* occ = h->red + nblk + 1;
  {
    // occupations change rarely: upload only when they differ from the device copy
    std::vector<double> o(nst, 0.0);
    if (occ_host)
      for (int n = 0; n < nst; ++n) o[n] = occ_host[n];
    if (o != h->occ_host) {
      hipMemcpyAsync(occ, o.data(), nst * sizeof(double), hipMemcpyHostToDevice, c->stream);
      hipStreamSynchronize(c->stream);  // o is a stack buffer
      h->occ_host = o;
    }
  }

  const size_t fsz = (size_t)P * nst * (real ? 1 : 2);
  if (!fp32) {
    double* F = static_cast<double*>(grow(h->f64, h->f64_bytes, fsz * sizeof(double)));
    const double* anl = static_cast<const double*>(h->anl64);
    if (!F) return QBH_ERR_ALLOC;
    if (real) {
      const double one = 1.0, zero = 0.0;
      {
        TimedLaunch tl(c, "fnl_gemm");
        if (int r = blas_ok(c,
                            rocblas_dgemm(bh, rocblas_operation_transpose, rocblas_operation_none,
                                          P, nst, 2 * ngw, &one, anl, 2 * ngw, psi, 2 * ld,
                                          &zero, F, P),
                            "fnl dgemm"))
          return r;
      }
      {
        TimedLaunch tl(c, "fnl_finish");
        k_fnl_finish_real<double><<<nblk, 256, 0, c->stream>>>(P, nst, F, anl, 2L * ngw, psi,
                                                               2L * ld, h->col_fac, occ, h->red);
      }
      TimedLaunch tl(c, "enl_gemm");
      if (int r = blas_ok(c,
                          rocblas_dgemm(bh, rocblas_operation_none, rocblas_operation_none,
                                        2 * ngw, nst, P, &one, anl, 2 * ngw, F, P, &one, hpsi,
                                        2 * ldh),
                          "enl dgemm"))
        return r;
    } else {
      const rocblas_double_complex one(1.0, 0.0), zero(0.0, 0.0);
      auto* A = reinterpret_cast<const rocblas_double_complex*>(anl);
      auto* B = reinterpret_cast<const rocblas_double_complex*>(psi);
      auto* Fz = reinterpret_cast<rocblas_double_complex*>(F);
      {
        TimedLaunch tl(c, "fnl_gemm");
        if (int r = blas_ok(c,
                            rocblas_zgemm(bh, rocblas_operation_conjugate_transpose,
                                          rocblas_operation_none, P, nst, ngw, &one, A, ngw, B,
                                          ld, &zero, Fz, P),
                            "fnl zgemm"))
          return r;
      }
      {
        TimedLaunch tl(c, "fnl_finish");
        k_fnl_finish_cplx<double><<<nblk, 256, 0, c->stream>>>(P, nst, F, h->col_fac, occ,
                                                               h->red);
      }
      TimedLaunch tl(c, "enl_gemm");
      if (int r = blas_ok(c,
                          rocblas_zgemm(bh, rocblas_operation_none, rocblas_operation_none, ngw,
                                        nst, P, &one, A, ngw, Fz, P, &one,
                                        reinterpret_cast<rocblas_double_complex*>(hpsi), ldh),
                          "enl zgemm"))
        return r;
    }
  } else {
    float* F = static_cast<float*>(grow(h->f32, h->f32_bytes, fsz * sizeof(float)));
    const size_t psz = 2 * (size_t)ngw * nst;
    float* p32 = static_cast<float*>(grow(h->psi32, h->psi32_bytes, psz * sizeof(float)));
    float* t32 = static_cast<float*>(grow(h->tmp32, h->tmp32_bytes, psz * sizeof(float)));
    const float* anl = static_cast<const float*>(h->anl32);
    if (!F || !p32 || !t32) {
      c->last_error = "FP32 projector scratch allocation failed";
      return QBH_ERR_ALLOC;
    }
    {
      TimedLaunch tl(c, "to_f32");
      k_to_f32<<<(unsigned)((psz + 255) / 256), 256, 0, c->stream>>>(ngw, nst, psi, ld, p32);
    }
    if (real) {
      const float one = 1.f, zero = 0.f;
      {
        TimedLaunch tl(c, "fnl_gemm");
        if (int r = blas_ok(c,
                            rocblas_sgemm(bh, rocblas_operation_transpose, rocblas_operation_none,
                                          P, nst, 2 * ngw, &one, anl, 2 * ngw, p32, 2 * ngw,
                                          &zero, F, P),
                            "fnl sgemm"))
          return r;
      }
      {
        TimedLaunch tl(c, "fnl_finish");
        k_fnl_finish_real<float><<<nblk, 256, 0, c->stream>>>(P, nst, F, anl, 2L * ngw, p32,
                                                              2L * ngw, h->col_fac, occ, h->red);
      }
      TimedLaunch tl(c, "enl_gemm");
      if (int r = blas_ok(c,
                          rocblas_sgemm(bh, rocblas_operation_none, rocblas_operation_none,
                                        2 * ngw, nst, P, &one, anl, 2 * ngw, F, P, &zero, t32,
                                        2 * ngw),
                          "enl sgemm"))
        return r;
    } else {
      const rocblas_float_complex one(1.f, 0.f), zero(0.f, 0.f);
      auto* A = reinterpret_cast<const rocblas_float_complex*>(anl);
      auto* B = reinterpret_cast<const rocblas_float_complex*>(p32);
      auto* Fz = reinterpret_cast<rocblas_float_complex*>(F);
      {
        TimedLaunch tl(c, "fnl_gemm");
        if (int r = blas_ok(c,
                            rocblas_cgemm(bh, rocblas_operation_conjugate_transpose,
                                          rocblas_operation_none, P, nst, ngw, &one, A, ngw, B,
                                          ngw, &zero, Fz, P),
                            "fnl cgemm"))
          return r;
      }
      {
        TimedLaunch tl(c, "fnl_finish");
        k_fnl_finish_cplx<float><<<nblk, 256, 0, c->stream>>>(P, nst, F, h->col_fac, occ,
                                                              h->red);
      }
      TimedLaunch tl(c, "enl_gemm");
      if (int r = blas_ok(c,
                          rocblas_cgemm(bh, rocblas_operation_none, rocblas_operation_none, ngw,
                                        nst, P, &one, A, ngw, Fz, P, &zero,
                                        reinterpret_cast<rocblas_float_complex*>(t32), ngw),
                          "enl cgemm"))
        return r;
    }
    {
      TimedLaunch tl(c, "add_f32");
      k_add_f32<<<(unsigned)((psz + 255) / 256), 256, 0, c->stream>>>(ngw, nst, t32, hpsi, ldh);
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    c->last_error = std::string("nonlocal: ") + hipGetErrorString(e);
    return QBH_ERR_DEVICE;
  }
  if (enl) {
    k_sum_partials<<<1, 256, 0, c->stream>>>(nblk, h->red);
    double v = 0.0;
    hipMemcpyAsync(&v, h->red + nblk, sizeof(double), hipMemcpyDeviceToHost, c->stream);
    hipStreamSynchronize(c->stream);
    *enl = v;
  }
  return 0;
}

}  // namespace qbh
