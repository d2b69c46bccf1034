// SCF driver on the device: the PSDA wavefunction stepper with its preconditioner, the
// Anderson charge mixer with Kerker weights, and the BOSampleStepper SCF loop for fixed ions.
//   Preconditioner::update/apply (src/qball/Preconditioner.cc:47-103): K_G = 0.5 /
//     max(|k+G|^2/2, ecutprec);
//   PSDAWavefunctionStepper::update (PSDAWavefunctionStepper.cc:38-375): A = C^T (HC) in the
//     basis metric, descent -K(HC - CA), Anderson extrapolation over one previous step;
//   AndersonMixer::update (AndersonMixer.cc): least squares over nmax previous pairs with
//     Tikhonov regularisation until sum theta^2 <= 1;
//   BOSampleStepper.cc:394-447 (Kerker and wls weights), :880-1130 (the loop):
//     update_density, charge mixing on rhog, update_rhor (ChargeDensity.cc:558-580),
//     update_vhxc, then nite x (energy(compute_hpsi) -> PSDA update -> gram).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/qbh.h"
#include "qbh_ham.h"

namespace qbh {
Basis* basis_of(qbh_basis b);
Ctx* ctx_of_basis(qbh_basis b);
Ctx* ctx_of(qbh_ctx c);
}  // namespace qbh

using namespace qbh;

namespace {

constexpr int NB = 256;

inline unsigned g256(long n) { return (unsigned)((n + 255) / 256); }

__device__ __forceinline__ double bsum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__global__ void k_prec_table(int ngw, const double* __restrict__ kpg2, double ecutprec,
                             double* __restrict__ K) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= ngw) return;
  const double e = 0.5 * kpg2[i];
  K[i] = (e < ecutprec) ? 0.5 / ecutprec : 0.5 / e;
}

// dc = -K * r in place (rows < ngw of each band)
__global__ void k_precondition(int ngw, int nst, const double* __restrict__ K, double* __restrict__ r,
                               long ldr) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i >= (long)ngw * nst) return;
  const long n = i / ngw, ig = i - n * ngw;
  double* p = r + 2 * (n * ldr + ig);
  const double k = -K[ig];
  p[0] *= k;
  p[1] *= k;
}

// partials of a = sum occ w_G f.df and b = sum occ w_G df.df, w_G = 2 except 1 at G = 0 (Γ)
__global__ void __launch_bounds__(256)
k_psda_ab(int ngw, int nst, int real, const double* __restrict__ occ,
          const double* __restrict__ dc, long ld, const double* __restrict__ dcl,
          double* __restrict__ part) {
  __shared__ double red[256];
  double a = 0.0, b = 0.0;
  const long tot = (long)ngw * nst;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < tot; i += (long)gridDim.x * 256) {
    const long n = i / ngw, ig = i - n * ngw;
    const double w = occ[n] * ((real && ig != 0) ? 2.0 : 1.0);
    const double f0 = dc[2 * (n * ld + ig)], f1 = dc[2 * (n * ld + ig) + 1];
    const double d0 = f0 - dcl[2 * i], d1 = f1 - dcl[2 * i + 1];
    a += w * (f0 * d0 + f1 * d1);
    b += w * (d0 * d0 + d1 * d1);
  }
  const double x = bsum(a, red), y = bsum(b, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = x;
    part[gridDim.x + blockIdx.x] = y;
  }
}

// c = x + th (x - xl) + f + th (f - fl); xl = x; fl = f.  th = 0 (the first step, or no
// extrapolation): c = x + f without reading xl, fl (the same value for finite history, and
// no 0 * NaN from history that was never written)
__global__ void k_psda_step(int ngw, int nst, double th, double* __restrict__ c, long ld,
                            const double* __restrict__ dc, long ldd, double* __restrict__ cl,
                            double* __restrict__ dcl) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i >= 2L * ngw * nst) return;
  const long n = i / (2L * ngw), r = i - n * 2L * ngw;
  const double x = c[n * 2 * ld + r], f = dc[n * 2 * ldd + r];
  c[n * 2 * ld + r] = th != 0.0 ? x + th * (x - cl[i]) + f + th * (f - dcl[i]) : x + f;
  cl[i] = x;
  dcl[i] = f;
}

// partials of sum occ_n w_G |k+G|^2/2 |c|^2
__global__ void __launch_bounds__(256)
k_ekin(int ngw, int nst, int real, const double* __restrict__ kpg2, const double* __restrict__ occ,
       const double* __restrict__ c, long ld, double* __restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  const long tot = (long)ngw * nst;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < tot; i += (long)gridDim.x * 256) {
    const long n = i / ngw, ig = i - n * ngw;
    const double w = occ[n] * ((real && ig != 0) ? 2.0 : 1.0);
    const double a = c[2 * (n * ld + ig)], b = c[2 * (n * ld + ig) + 1];
    s += w * 0.5 * kpg2[ig] * (a * a + b * b);
  }
  const double x = bsum(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = x;
}

// mixer dot products: D_i = f_k - f_{kmi}; out[p] partials for the n(n+1)/2 products
// D_i.D_j (j <= i, row-major) followed by the n products D_i.f_k
struct Hist {
  const double* f[8];
  int n;
};
__global__ void __launch_bounds__(256)
k_mix_dots(long m, const double* __restrict__ fk, Hist h, double* __restrict__ part) {
  __shared__ double red[256];
  double acc[44];
  const int np = h.n * (h.n + 1) / 2 + h.n;
  for (int p = 0; p < np; ++p) acc[p] = 0.0;
  for (long l = blockIdx.x * 256L + threadIdx.x; l < m; l += (long)gridDim.x * 256) {
    const double f = fk[l];
    double d[8];
    for (int i = 0; i < h.n; ++i) d[i] = f - h.f[i][l];
    int p = 0;
    for (int i = 0; i < h.n; ++i)
      for (int j = 0; j <= i; ++j) acc[p++] += d[i] * d[j];
    for (int i = 0; i < h.n; ++i) acc[p++] += d[i] * f;
  }
  for (int p = 0; p < np; ++p) {
    const double x = bsum(acc[p], red);
    if (threadIdx.x == 0) part[(long)p * gridDim.x + blockIdx.x] = x;
  }
}

// out = v_k - sum_i theta_i (v_k - v_{kmi})
struct Comb {
  const double* v[8];
  double th[8];
  int n;
};
__global__ void k_mix_combine(long m, const double* __restrict__ vk, Comb cb, double* __restrict__ out) {
  const long l = blockIdx.x * 256L + threadIdx.x;
  if (l >= m) return;
  const double x = vk[l];
  double s = x;
  for (int i = 0; i < cb.n; ++i) s -= cb.th[i] * (x - cb.v[i][l]);
  out[l] = s;
}

// Kerker / wls weights of the density basis (BOSampleStepper.cc:409-447)
__global__ void k_mix_weights(int n, const double* __restrict__ g2, double q1, double q0sq,
                              int kerker, double* __restrict__ wk, double* __restrict__ wls) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double g = g2[i];
  wls[i] = (g != 0.0) ? sqrt(g / (g + q1 * q1)) : 1.0;
  wk[i] = kerker ? g / (g + q0sq) : 1.0;
}

// drhog = (rhog - rhog_in) / wls
__global__ void k_mix_residual(int n, const double* __restrict__ rhog, const double* __restrict__ in,
                               const double* __restrict__ wls, double* __restrict__ d) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  d[2 * i] = (rhog[2 * i] - in[2 * i]) / wls[i];
  d[2 * i + 1] = (rhog[2 * i + 1] - in[2 * i + 1]) / wls[i];
}

// rhog_in = rhobar + alpha * (drhobar * wls) * wk; rhog = rhog_in
__global__ void k_mix_apply(int n, const double* __restrict__ rb, const double* __restrict__ drb,
                            const double* __restrict__ wls, const double* __restrict__ wk,
                            double alpha, double* __restrict__ in, double* __restrict__ rhog) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < 2; ++c) {
    const double d = drb[2 * i + c] * wls[i];
    const double v = rb[2 * i + c] + alpha * d * wk[i];
    in[2 * i + c] = v;
    rhog[2 * i + c] = v;
  }
}

__global__ void k_rhor_from_grid(long N, const double* __restrict__ f, double omega_inv,
                                 double* __restrict__ rhor) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i < N) rhor[i] = f[2 * i] * omega_inv;
}

__global__ void k_zero(double* __restrict__ p, long n) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

double host_sum(const std::vector<double>& v, int off, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += v[off + i];
  return s;
}

// solve the symmetric n x n system A x = b (A given in full) by Gaussian elimination with
// partial pivoting; false if singular
bool solve_sym(int n, std::vector<double> A, std::vector<double> b, std::vector<double>& x) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[i * n + k]) > std::fabs(A[p * n + k])) p = i;
    if (A[p * n + k] == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[p * n + j]);
      std::swap(b[k], b[p]);
    }
    for (int i = k + 1; i < n; ++i) {
      const double f = A[i * n + k] / A[k * n + k];
      for (int j = k; j < n; ++j) A[i * n + j] -= f * A[k * n + j];
      b[i] -= f * b[k];
    }
  }
  x.assign(n, 0.0);
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int j = i + 1; j < n; ++j) s -= A[i * n + j] * x[j];
    x[i] = s / A[i * n + i];
  }
  return true;
}

}  // namespace

static int sfail(Ctx* c, int code, const std::string& m) {
  c->last_error = m;
  return code;
}

static int scheck(Ctx* c, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return sfail(c, QBH_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  return QBH_OK;
}

// ---------------------------------------------------------------------------- PSDA
struct qbh_psda_s {
  Ctx* c = nullptr;
  Basis* b = nullptr;
  int nst = 0;
  double* K = nullptr;
  double* cl = nullptr;
  double* dcl = nullptr;
  double* A = nullptr;
  double* occ = nullptr;
  double* part = nullptr;
  double* scratch = nullptr;
  size_t scratch_bytes = 0;
  bool extrapolate = false;
  std::vector<double> occ_host;
};

extern "C" int qbh_psda_create(qbh_basis basis, int nst, double ecutprec, qbh_psda* out) {
  if (!basis || nst <= 0 || !(ecutprec > 0.0) || !out) return QBH_ERR_ARG;
  Ctx* c = ctx_of_basis(basis);
  hipSetDevice(c->device);
  auto* p = new qbh_psda_s();
  p->c = c;
  p->b = basis_of(basis);
  p->nst = nst;
  const int ngw = p->b->g.ngw;
  const size_t blk = 2 * (size_t)ngw * nst;
  bool ok = hipMalloc(&p->K, ngw * sizeof(double)) == hipSuccess &&
            hipMalloc(&p->cl, blk * sizeof(double)) == hipSuccess &&
            hipMalloc(&p->dcl, blk * sizeof(double)) == hipSuccess &&
            hipMalloc(&p->A, 2 * (size_t)nst * nst * sizeof(double)) == hipSuccess &&
            hipMalloc(&p->occ, nst * sizeof(double)) == hipSuccess &&
            hipMalloc(&p->part, 2 * NB * sizeof(double)) == hipSuccess;
  if (!ok) {
    qbh_psda_destroy(p);
    return sfail(c, QBH_ERR_ALLOC, "PSDA buffers allocation failed");
  }
  k_prec_table<<<g256(ngw), 256, 0, c->stream>>>(ngw, p->b->kpg2, ecutprec, p->K);
  // the extrapolation history starts cleared (written before it is first read)
  k_zero<<<g256((long)blk), 256, 0, c->stream>>>(p->cl, (long)blk);
  k_zero<<<g256((long)blk), 256, 0, c->stream>>>(p->dcl, (long)blk);
  *out = p;
  return scheck(c, "psda create");
}

extern "C" int qbh_psda_destroy(qbh_psda p) {
  if (!p) return QBH_ERR_ARG;
  hipSetDevice(p->c->device);
  hipStreamSynchronize(p->c->stream);
  for (void* q : {(void*)p->K, (void*)p->cl, (void*)p->dcl, (void*)p->A, (void*)p->occ,
                  (void*)p->part, (void*)p->scratch})
    if (q) hipFree(q);
  delete p;
  return QBH_OK;
}

extern "C" int qbh_psda_preprocess(qbh_psda p) {
  if (!p) return QBH_ERR_ARG;
  p->extrapolate = false;
  return QBH_OK;
}

extern "C" int qbh_psda_update(qbh_psda p, double* psi_dev, long ld, double* hpsi_dev, long ldh,
                               const double* occ, double* theta) {
  if (!p || !psi_dev || !hpsi_dev || !occ) return QBH_ERR_ARG;
  Ctx* c = p->c;
  hipSetDevice(c->device);
  const GridDev& g = p->b->g;
  const int ngw = g.ngw, nst = p->nst;
  const bool real = g.real;
  if (ld < ngw || ldh < ngw) return sfail(c, QBH_ERR_ARG, "ld < ngw");
  void* bh = nullptr;
  if (int r = blas_handle(c, &bh)) return r;
  rocblas_handle h = static_cast<rocblas_handle>(bh);
  std::vector<double> o(occ, occ + nst);
  if (o != p->occ_host) {
    hipMemcpyAsync(p->occ, o.data(), nst * sizeof(double), hipMemcpyHostToDevice, c->stream);
    hipStreamSynchronize(c->stream);
    p->occ_host = o;
  }
  // A = C^T (HC) in the basis metric (PSDAWavefunctionStepper.cc:57-64, :271)
  if (int r = band_product(c, bh, real, nst, nst, real ? 2L * ngw : (long)ngw, psi_dev,
                           real ? 2 * ld : ld, hpsi_dev, real ? 2 * ldh : ldh, real ? 2.0 : 1.0,
                           p->A, p->scratch, p->scratch_bytes))
    return r;
  // HC <- HC - C A
  rocblas_status st;
  if (real) {
    const double m1 = -1.0, one = 1.0;
    st = rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, 2 * ngw, nst, nst, &m1,
                       psi_dev, 2 * ld, p->A, nst, &one, hpsi_dev, 2 * ldh);
  } else {
    const rocblas_double_complex m1(-1.0, 0.0), one(1.0, 0.0);
    st = rocblas_zgemm(h, rocblas_operation_none, rocblas_operation_none, ngw, nst, nst, &m1,
                       reinterpret_cast<const rocblas_double_complex*>(psi_dev), ld,
                       reinterpret_cast<const rocblas_double_complex*>(p->A), nst, &one,
                       reinterpret_cast<rocblas_double_complex*>(hpsi_dev), ldh);
  }
  if (st != rocblas_status_success)
    return sfail(c, QBH_ERR_BLAS, std::string("psda residual: ") + rocblas_status_to_string(st));
  // dc = -K (HC - CA)
  k_precondition<<<g256((long)ngw * nst), 256, 0, c->stream>>>(ngw, nst, p->K, hpsi_dev, ldh);
  double th = 0.0;
  if (p->extrapolate) {
    k_psda_ab<<<NB, 256, 0, c->stream>>>(ngw, nst, real ? 1 : 0, p->occ, hpsi_dev, ldh, p->dcl,
                                         p->part);
    std::vector<double> hp(2 * NB);
    hipMemcpyAsync(hp.data(), p->part, 2 * NB * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    hipStreamSynchronize(c->stream);
    const double a = host_sum(hp, 0, NB), b = host_sum(hp, NB, NB);
    if (b != 0.0) th = -a / b;
    if (th < -1.0) th = 0.0;
    th = std::min(2.0, th);
  }
  k_psda_step<<<g256(2L * ngw * nst), 256, 0, c->stream>>>(ngw, nst, th, psi_dev, ld, hpsi_dev,
                                                          ldh, p->cl, p->dcl);
  p->extrapolate = true;
  if (theta) *theta = th;
  return scheck(c, "psda update");
}

extern "C" int qbh_kinetic_energy(qbh_basis basis, const double* psi_dev, long ld, int nst,
                                  const double* occ, double* ekin) {
  if (!basis || !psi_dev || !occ || nst <= 0 || !ekin) return QBH_ERR_ARG;
  Ctx* c = ctx_of_basis(basis);
  Basis* b = basis_of(basis);
  hipSetDevice(c->device);
  double* d = nullptr;
  if (hipMalloc(&d, (nst + NB) * sizeof(double)) != hipSuccess)
    return sfail(c, QBH_ERR_ALLOC, "kinetic energy scratch");
  hipMemcpyAsync(d, occ, nst * sizeof(double), hipMemcpyHostToDevice, c->stream);
  k_ekin<<<NB, 256, 0, c->stream>>>(b->g.ngw, nst, b->g.real ? 1 : 0, b->kpg2, d, psi_dev, ld,
                                    d + nst);
  std::vector<double> hp(NB);
  hipMemcpyAsync(hp.data(), d + nst, NB * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  hipStreamSynchronize(c->stream);
  hipFree(d);
  *ekin = host_sum(hp, 0, NB);
  return scheck(c, "kinetic energy");
}

// ---------------------------------------------------------------------------- mixer
struct qbh_mixer_s {
  Ctx* c = nullptr;
  long m = 0;
  int nmax = 0, n = -1, k = -1;
  std::vector<double*> x, f;
  double* part = nullptr;
};

static int mixer_new(Ctx* c, long m, int nmax, qbh_mixer* out) {
  auto* mx = new qbh_mixer_s();
  mx->c = c;
  mx->m = m;
  mx->nmax = nmax;
  bool ok = hipMalloc(&mx->part, 44 * NB * sizeof(double)) == hipSuccess;
  for (int i = 0; i <= nmax && ok; ++i) {
    double *a = nullptr, *b = nullptr;
    ok = hipMalloc(&a, m * sizeof(double)) == hipSuccess;
    if (ok) ok = hipMalloc(&b, m * sizeof(double)) == hipSuccess;
    mx->x.push_back(a);
    mx->f.push_back(b);
  }
  if (!ok) {
    qbh_mixer_destroy(mx);
    return sfail(c, QBH_ERR_ALLOC, "mixer history allocation failed");
  }
  *out = mx;
  return QBH_OK;
}

extern "C" int qbh_mixer_create(qbh_ctx ctx, long m, int nmax, qbh_mixer* out) {
  if (!ctx || m <= 0 || nmax < 0 || nmax > 8 || !out) return QBH_ERR_ARG;
  Ctx* c = ctx_of(ctx);
  hipSetDevice(c->device);
  return mixer_new(c, m, nmax, out);
}

extern "C" int qbh_mixer_destroy(qbh_mixer mx) {
  if (!mx) return QBH_ERR_ARG;
  hipSetDevice(mx->c->device);
  hipStreamSynchronize(mx->c->stream);
  for (double* p : mx->x) if (p) hipFree(p);
  for (double* p : mx->f) if (p) hipFree(p);
  if (mx->part) hipFree(mx->part);
  delete mx;
  return QBH_OK;
}

extern "C" int qbh_mixer_restart(qbh_mixer mx) {
  if (!mx) return QBH_ERR_ARG;
  mx->n = -1;
  mx->k = -1;
  return QBH_OK;
}

extern "C" int qbh_mixer_update(qbh_mixer mx, const double* x_dev, const double* f_dev,
                                double* xbar_dev, double* fbar_dev, double* theta) {
  if (!mx || !x_dev || !f_dev || !xbar_dev || !fbar_dev) return QBH_ERR_ARG;
  Ctx* c = mx->c;
  hipSetDevice(c->device);
  const int nm = mx->nmax;
  mx->k = (mx->k + 1) % (nm + 1);
  if (mx->n < nm) mx->n++;
  const int k = mx->k, n = mx->n;
  hipMemcpyAsync(mx->x[k], x_dev, mx->m * sizeof(double), hipMemcpyDeviceToDevice, c->stream);
  hipMemcpyAsync(mx->f[k], f_dev, mx->m * sizeof(double), hipMemcpyDeviceToDevice, c->stream);
  std::vector<double> th(std::max(n, 0), 0.0);
  if (n > 0) {
    Hist hs{};
    hs.n = n;
    for (int i = 0; i < n; ++i) hs.f[i] = mx->f[(k - i + nm) % (nm + 1)];
    const int np = n * (n + 1) / 2 + n;
    k_mix_dots<<<NB, 256, 0, c->stream>>>(mx->m, mx->f[k], hs, mx->part);
    std::vector<double> hp((size_t)np * NB);
    hipMemcpyAsync(hp.data(), mx->part, hp.size() * sizeof(double), hipMemcpyDeviceToHost,
                   c->stream);
    hipStreamSynchronize(c->stream);
    std::vector<double> A((size_t)n * n), b(n);
    int p = 0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j <= i; ++j) {
        const double s = host_sum(hp, p++ * NB, NB);
        A[i * n + j] = A[j * n + i] = s;
      }
    for (int i = 0; i < n; ++i) b[i] = host_sum(hp, p++ * NB, NB);
    // Tikhonov regularisation, doubled until sum theta^2 <= 1 (AndersonMixer.cc)
    double tp = 1e-12;
    for (int it = 0; it < 100; ++it) {
      std::vector<double> Ar = A;
      for (int i = 0; i < n; ++i) Ar[i * n + i] += tp;
      std::vector<double> t;
      if (!solve_sym(n, Ar, b, t)) return sfail(c, QBH_ERR_ARG, "Anderson mixer: singular system");
      th = t;
      double s2 = 0.0;
      for (double v : t) s2 += v * v;
      if (s2 <= 1.0) break;
      tp *= 2.0;
    }
  }
  Comb cf{}, cx{};
  cf.n = cx.n = n;
  for (int i = 0; i < n; ++i) {
    const int kmi = (k - i + nm) % (nm + 1);
    cf.v[i] = mx->f[kmi];
    cx.v[i] = mx->x[kmi];
    cf.th[i] = cx.th[i] = th[i];
  }
  k_mix_combine<<<g256(mx->m), 256, 0, c->stream>>>(mx->m, mx->f[k], cf, fbar_dev);
  k_mix_combine<<<g256(mx->m), 256, 0, c->stream>>>(mx->m, mx->x[k], cx, xbar_dev);
  if (theta)
    for (int i = 0; i < n; ++i) theta[i] = th[i];
  return scheck(c, "mixer update");
}

// ---------------------------------------------------------------------------- SCF loop
// SlaterDet::fermi (SlaterDet.cc:1897-1915): e, mu in Hartree, width in Rydberg
static double fermi(double e, double mu, double width) {
  if (width == 0.0) return e < mu ? 1.0 : (e == mu ? 0.5 : 0.0);
  const double kt = 0.5 * width;
  const double arg = (e - mu) / kt;
  if (arg < -30.0) return 1.0;
  if (arg > 30.0) return 0.0;
  return 1.0 / (1.0 + exp(arg));
}

// Wavefunction::update_occ, one spin and one k-point of weight 1, Fermi smearing
// (Wavefunction.cc:1214-1361 with SlaterDet::update_occ, SlaterDet.cc:1840-1881): mu by
// step-halving from 0 in steps of 10 eV until |sum occ - nel| <= 1e-8; occ = 2 f
static double update_occ(const std::vector<double>& eig, double nel, double width,
                         std::vector<double>& occ) {
  const double eVolt = 0.036749023;
  double mu = 0.0, dmu = 10.0 * eVolt;
  bool up = true;
  auto fill = [&]() {
    double s = 0.0;
    for (size_t n = 0; n < eig.size(); ++n) {
      occ[n] = 2.0 * fermi(eig[n], mu, width);
      s += occ[n];
    }
    return s;
  };
  double rhosum = fill();
  for (int niter = 0; niter < 10000 && fabs(rhosum - nel) > 1.e-8; ++niter) {
    if (rhosum < nel) {
      if (!up) dmu /= 2.0;
      mu += dmu;
      up = true;
    } else {
      if (up) dmu /= 2.0;
      mu -= dmu;
      up = false;
    }
    rhosum = fill();
  }
  return mu;
}

// SlaterDet::entropy, one spin (SlaterDet.cc:1948-1962)
static double entropy(const std::vector<double>& occ) {
  double sum = 0.0;
  for (double o : occ) {
    const double f = o / 2.0;
    if (f > 0.0 && f < 1.0) sum -= 2.0 * (f * log(f) + (1.0 - f) * log(1.0 - f));
  }
  return sum;
}

static int scf_impl(qbh_ham ham, qbh_basis vbasis, double* psi_dev, long ld, int nst,
                    std::vector<double>& occv, double nel, double width, bool diag, double omega,
                    const double* vion_g_dev, const double* rhops_g_dev, int nitscf, int nite,
                    double ecutprec, double mix_alpha, int mix_ndim, double rc_kerker,
                    int flags_inner, int flags_last, double* history, double* eig_out,
                    double* mu_out) {
  const double* occ = occv.data();
  Ham* h = &ham->h;
  Ctx* c = h->ctx;
  hipSetDevice(c->device);
  const GridDev& g = h->b->g;
  const GridDev& vg = basis_of(vbasis)->g;
  if (ld < g.ngw) return sfail(c, QBH_ERR_ARG, "ld < ngw");
  if (vg.np0 != g.np0 || vg.np1 != g.np1 || vg.np2 != g.np2)
    return sfail(c, QBH_ERR_ARG, "density basis grid differs from the wavefunction grid");
  if (ctx_of_basis(vbasis) != c) return sfail(c, QBH_ERR_ARG, "vbasis on another context");
  const long N = (long)g.np0 * g.np1 * g.np2;
  const int ngv = vg.ngw;
  const int nin = std::max(nite, 1);
  qbh_basis wb = reinterpret_cast<qbh_basis>(h->b_handle);
  if (!h->v && hipMalloc(&h->v, N * sizeof(double)) != hipSuccess) {
    h->v = nullptr;
    return sfail(c, QBH_ERR_ALLOC, "v_r allocation failed");
  }
  // device buffers
  double* mem = nullptr;
  const size_t nd = 2 * (size_t)g.ngw * nst + 3 * (size_t)N + 10 * (size_t)ngv + 2 * (size_t)ngv;
  if (hipMalloc(&mem, nd * sizeof(double)) != hipSuccess)
    return sfail(c, QBH_ERR_ALLOC, "SCF buffers allocation failed");
  double* hc = mem;
  double* rhor = hc + 2 * (size_t)g.ngw * nst;
  double* fgrid = rhor + N;  // 2N
  double* rhog = fgrid + 2 * N;
  double* rin = rhog + 2 * (size_t)ngv;
  double* drg = rin + 2 * (size_t)ngv;
  double* rbar = drg + 2 * (size_t)ngv;
  double* drbar = rbar + 2 * (size_t)ngv;
  double* wk = drbar + 2 * (size_t)ngv;
  double* wls = wk + ngv;
  qbh_psda psda = nullptr;
  qbh_mixer mixer = nullptr;
  int r = qbh_psda_create(wb, nst, ecutprec, &psda);
  if (r == QBH_OK) r = mixer_new(c, 2L * ngv, std::max(0, std::min(mix_ndim, 8)), &mixer);
  if (r == QBH_OK) {
    const double q1 = 2.0 * M_PI / 3.0;
    const double q0 = rc_kerker > 0.0 ? 2.0 * M_PI / rc_kerker : 0.0;
    k_mix_weights<<<g256(ngv), 256, 0, c->stream>>>(ngv, basis_of(vbasis)->kpg2, q1, q0 * q0,
                                                    rc_kerker > 0.0 ? 1 : 0, wk, wls);
  }
  for (int it = 0; it < nitscf && r == QBH_OK; ++it) {
    // ChargeDensity::update_density
    k_zero<<<g256(N), 256, 0, c->stream>>>(rhor, N);
    if ((r = qbh_density(wb, psi_dev, ld, nst, occ, 1.0, omega, rhor, 0))) break;
    if ((r = qbh_fft_forward_real(vbasis, rhor, omega, rhog, 0))) break;
    if (it == 0) {
      hipMemcpyAsync(rin, rhog, 2 * sizeof(double) * ngv, hipMemcpyDeviceToDevice, c->stream);
    } else {
      k_mix_residual<<<g256(ngv), 256, 0, c->stream>>>(ngv, rhog, rin, wls, drg);
      if ((r = qbh_mixer_update(mixer, rin, drg, rbar, drbar, nullptr))) break;
      k_mix_apply<<<g256(ngv), 256, 0, c->stream>>>(ngv, rbar, drbar, wls, wk, mix_alpha, rin,
                                                    rhog);
      // ChargeDensity::update_rhor
      if ((r = qbh_fft_backward(vbasis, rhog, ngv, 1, 0, fgrid, 0))) break;
      k_rhor_from_grid<<<g256(N), 256, 0, c->stream>>>(N, fgrid, 1.0 / omega, rhor);
    }
    double e3[3] = {0, 0, 0};
    if ((r = qbh_update_vhxc(vbasis, rhor, rhog, nullptr, vion_g_dev, rhops_g_dev, omega, h->v,
                             e3)))
      break;
    qbh_psda_preprocess(psda);
    for (int ite = 0; ite < nin && r == QBH_OK; ++ite) {
      const int fl = whole_h_flags((ite == nin - 1) ? flags_last : flags_inner);
      double enl = 0.0, ekin = 0.0, th = 0.0;
      if ((r = hpsi_impl(h, psi_dev, ld, nst, occ, hc, g.ngw, fl, &enl))) break;
      if ((r = qbh_kinetic_energy(wb, psi_dev, ld, nst, occ, &ekin))) break;
      if ((r = qbh_psda_update(psda, psi_dev, ld, hc, g.ngw, occ, &th))) break;
      if ((r = qbh_gram(wb, psi_dev, ld, nst))) break;
      // EnergyFunctional.cc:1602-1609: ets = -entropy * smearing_width / 2
      const double ets = width > 0.0 ? -entropy(occv) * width * 0.5 : 0.0;
      if (history) {
        double* hrow = history + 7 * ((size_t)it * nin + ite);
        hrow[0] = ekin + enl + e3[0] + e3[1] + e3[2] + ets;
        hrow[1] = ekin;
        hrow[2] = enl;
        hrow[3] = e3[0];
        hrow[4] = e3[1];
        hrow[5] = e3[2];
        hrow[6] = th;
      }
    }
    if (r == QBH_OK && (width > 0.0 || diag)) {
      // BOSampleStepper.cc:1150-1185: energy with Hψ, Wavefunction::diag with
      // eigenvectors, then (fractional occupations) update_occ
      std::vector<double> eig(nst);
      if ((r = hpsi_impl(h, psi_dev, ld, nst, occ, hc, g.ngw, whole_h_flags(flags_last),
                         nullptr)))
        break;
      if ((r = qbh_subspace_diag(wb, psi_dev, ld, hc, g.ngw, nst, eig.data(), 1))) break;
      const double mu = width > 0.0 ? update_occ(eig, nel, width, occv) : 0.0;
      if (eig_out) std::copy(eig.begin(), eig.end(), eig_out + (size_t)it * nst);
      if (mu_out) mu_out[it] = mu;
    }
  }
  hipStreamSynchronize(c->stream);
  if (psda) qbh_psda_destroy(psda);
  if (mixer) qbh_mixer_destroy(mixer);
  hipFree(mem);
  if (r) return r;
  return scheck(c, "scf");
}

extern "C" int qbh_scf_run(qbh_ham ham, qbh_basis vbasis, double* psi_dev, long ld, int nst,
                           const double* occ, double omega, const double* vion_g_dev,
                           const double* rhops_g_dev, int nitscf, int nite, double ecutprec,
                           double mix_alpha, int mix_ndim, double rc_kerker, int flags_inner,
                           int flags_last, double* history) {
  if (!ham || !vbasis || !psi_dev || !occ || nst <= 0 || nitscf < 1 || !(omega > 0.0))
    return QBH_ERR_ARG;
  std::vector<double> o(occ, occ + nst);
  return scf_impl(ham, vbasis, psi_dev, ld, nst, o, 0.0, 0.0, false, omega, vion_g_dev,
                  rhops_g_dev, nitscf, nite, ecutprec, mix_alpha, mix_ndim, rc_kerker,
                  flags_inner, flags_last, history, nullptr, nullptr);
}

extern "C" int qbh_scf_run_smeared(qbh_ham ham, qbh_basis vbasis, double* psi_dev, long ld,
                                   int nst, double* occ, double nel, double smearing,
                                   double omega, const double* vion_g_dev,
                                   const double* rhops_g_dev, int nitscf, int nite,
                                   double ecutprec, double mix_alpha, int mix_ndim,
                                   double rc_kerker, int flags_inner, int flags_last,
                                   double* history, double* eig, double* mu) {
  if (!ham || !vbasis || !psi_dev || !occ || nst <= 0 || nitscf < 1 || !(omega > 0.0) ||
      !(smearing >= 0.0) || (smearing > 0.0 && (!(nel > 0.0) || nel > 2.0 * nst)))
    return QBH_ERR_ARG;
  std::vector<double> o(occ, occ + nst);
  const int r = scf_impl(ham, vbasis, psi_dev, ld, nst, o, nel, smearing, true, omega,
                         vion_g_dev,
                         rhops_g_dev, nitscf, nite, ecutprec, mix_alpha, mix_ndim, rc_kerker,
                         flags_inner, flags_last, history, eig, mu);
  std::copy(o.begin(), o.end(), occ);
  return r;
}
