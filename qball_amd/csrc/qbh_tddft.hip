// Real-time TDDFT propagation over many steps and its discrete adjoint.
//
// Forward: the EhrenSampleStepper iteration (src/qball/EhrenSampleStepper.cc:280-1152) for
// fixed ions, one spin, one k-point, complex basis:
//   iteration start (:323-330): rho(psi_n) -> v_a = update_vhxc + v_ext; current j(psi_n,
//     A_n) (CurrentDensity::update_current, :359);
//   ETRS preupdate under H[v_a, A_n] (ExponentialWavefunctionStepper.cc:151-240);
//   A_n -> A_{n+1} (:480);
//   update: v_b = update_vhxc(rho(psi~)) + v_ext; psi_{n+1} = T(dt/2) psi(t+dt/2) under
//     H[v_b, A_{n+1}] (ExponentialWavefunctionStepper.cc:243-276).
// Objective J = sum_n w_n e.j(psi_n, A_n), j_d = -sum_b f_b sum_G (k+G)_d |c|^2 + A_d N_e.
// A enters H through 0.5 |k+G-A|^2 and, when the Hamiltonian's projector tables come from
// species (qbh_ham_set_species), through twnl re-tabled at |k+G-A| on every change of A
// (NonLocalPotential::update_twnl under a vector potential); the projector phases at k+G-A
// cancel in beta D beta^H.  So dH/dA_d = -(k+G-A)_d on the diagonal plus, with species,
// sum_p |d_d beta_p> D_p <beta_p| + h.c., d_d beta_p = -(-i)^l (d twnl_p / dq_d) e^{-i(k+G-A).tau}.
//
// Backward: exact derivative of that discrete map (Wirtinger gradients, dJ = 2 Re <g, dpsi>).
// psi is checkpointed every `checkpoint` steps; each segment is recomputed with its Krylov
// vectors H^k x stored, then walked in reverse:
//   T^H = Taylor with dt -> -dt; the local-potential gradient of one exponential is
//   mu(r) = (2/N) Re sum_b sum_k conj(z_k(r)) U_k(r), z_k = H^k lambda,
//   U_k = sum_m alpha_{k+m+1} u_m; dJ/dA_d = -2 Re sum conj(z_k) (k+G-A)_d U_k, plus with
//   species -2 Re sum_{p,b} D_p [conj(dF^z_d) F^U + conj(F^z) dF^U_d], F = anl^H x and
//   dF_d = danl_d^H x, danl_d built like anl from d twnl / dq_d;
//   through v = V_Hxc[rho]: nu = f_xc(rho) mu + K_H mu, lambda += (f_b N/Omega) V_loc[nu] psi.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/qbh.h"
#include "qbh_gemm.h"
#include "qbh_ham.h"

namespace qbh {
Basis* basis_of(qbh_basis b);
Ctx* ctx_of_basis(qbh_basis b);
}  // namespace qbh

using namespace qbh;

namespace {

constexpr int NBLK = 256;

__global__ void k_fill(double* __restrict__ p, long n) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

__global__ void k_axpy_real(long n, double a, const double* __restrict__ x, double* __restrict__ y) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}

__device__ __forceinline__ double block_sum(double v, double* red) {
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

// partials of S = sum_b f_b sum_G (e.K)|c|^2 and Ne = sum_b f_b sum_G |c|^2
__global__ void __launch_bounds__(256)
k_cur_part(int ngw, int nst, const double* __restrict__ kpgx, double e0, double e1, double e2,
           const double* __restrict__ f, const double* __restrict__ c, double* __restrict__ part) {
  __shared__ double red[256];
  double s = 0.0, ne = 0.0;
  const long tot = (long)ngw * nst;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < tot; i += (long)gridDim.x * 256) {
    const long b = i / ngw, ig = i - b * ngw;
    const double a2 = c[2 * i] * c[2 * i] + c[2 * i + 1] * c[2 * i + 1];
    const double ek = e0 * kpgx[ig] + e1 * kpgx[ngw + ig] + e2 * kpgx[2L * ngw + ig];
    s += f[b] * ek * a2;
    ne += f[b] * a2;
  }
  const double a = block_sum(s, red);
  const double bsum = block_sum(ne, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[gridDim.x + blockIdx.x] = bsum;
  }
}

// out[k] = scale * sum_i part[k * n + i] (+= when acc), k < m; one block
__global__ void __launch_bounds__(256)
k_final(int n, int m, const double* __restrict__ part, double scale, int acc,
        double* __restrict__ out) {
  __shared__ double red[256];
  for (int k = 0; k < m; ++k) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += part[k * n + i];
    const double t = block_sum(s, red);
    if (threadIdx.x == 0) out[k] = (acc ? out[k] : 0.0) + scale * t;
  }
}

// lam_b += w * (-f_b (e.(K - A))) c_b
__global__ void k_dj(int ngw, int nst, const double* __restrict__ kpgx, double e0, double e1,
                     double e2, double ea, const double* __restrict__ f, double w,
                     const double* __restrict__ c, double* __restrict__ lam) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i >= (long)ngw * nst) return;
  const long b = i / ngw, ig = i - b * ngw;
  const double ek = e0 * kpgx[ig] + e1 * kpgx[ngw + ig] + e2 * kpgx[2L * ngw + ig] - ea;
  const double s = -w * f[b] * ek;
  lam[2 * i] += s * c[2 * i];
  lam[2 * i + 1] += s * c[2 * i + 1];
}

// U = sum_{m < cnt} a_m u_m (complex coefficients), blocks of n complex
struct Combo {
  const double* u[8];
  double ar[8], ai[8];
  int cnt;
};
__global__ void k_combo(long n, Combo cb, double* __restrict__ U) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i >= n) return;
  double sr = 0.0, si = 0.0;
  for (int m = 0; m < cb.cnt; ++m) {
    const double xr = cb.u[m][2 * i], xi = cb.u[m][2 * i + 1];
    sr += cb.ar[m] * xr - cb.ai[m] * xi;
    si += cb.ar[m] * xi + cb.ai[m] * xr;
  }
  U[2 * i] = sr;
  U[2 * i + 1] = si;
}

// partials of Re sum conj(z) (K - A)_d U for d = 0..2
__global__ void __launch_bounds__(256)
k_crossk_part(int ngw, int nst, const double* __restrict__ kpgx, double A0, double A1, double A2,
              const double* __restrict__ z, const double* __restrict__ U, double* __restrict__ part) {
  __shared__ double red[256];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  const long tot = (long)ngw * nst;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < tot; i += (long)gridDim.x * 256) {
    const long ig = i % ngw;
    const double re = z[2 * i] * U[2 * i] + z[2 * i + 1] * U[2 * i + 1];
    s0 += (kpgx[ig] - A0) * re;
    s1 += (kpgx[ngw + ig] - A1) * re;
    s2 += (kpgx[2L * ngw + ig] - A2) * re;
  }
  const double a = block_sum(s0, red);
  const double b = block_sum(s1, red);
  const double c = block_sum(s2, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[gridDim.x + blockIdx.x] = b;
    part[2 * gridDim.x + blockIdx.x] = c;
  }
}

// partials of s_d = Re sum_{p,b} fac_p [conj(Fz[(1+d)P + p, b]) FU[p, b] + conj(Fz[p, b])
// FU[(1+d)P + p, b]] for d = 0..2; Fz, FU complex with leading dimension 4P
__global__ void __launch_bounds__(256)
k_nl_dA_part(int P, int nst, const double* __restrict__ fac, const double* __restrict__ Fz,
             const double* __restrict__ FU, double* __restrict__ part) {
  __shared__ double red[256];
  double s[3] = {0.0, 0.0, 0.0};
  const long tot = (long)P * nst;
  const long ld = 4L * P;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < tot; i += (long)gridDim.x * 256) {
    const long p = i % P, b = i / P;
    const long o = 2 * (b * ld + p);
    const double zr = Fz[o], zi = Fz[o + 1], ur = FU[o], ui = FU[o + 1];
    for (int d = 0; d < 3; ++d) {
      const long od = o + 2L * (1 + d) * P;
      s[d] += fac[p] * ((Fz[od] * ur + Fz[od + 1] * ui) + (zr * FU[od] + zi * FU[od + 1]));
    }
  }
  for (int d = 0; d < 3; ++d) {
    const double a = block_sum(s[d], red);
    if (threadIdx.x == 0) part[d * gridDim.x + blockIdx.x] = a;
  }
}

// mu[r] += scale * sum_{b < nb} Re(conj(fz_b(r)) fU_b(r))
__global__ void k_cross_acc(long N, int nb, double scale, const double* __restrict__ fz,
                            const double* __restrict__ fU, double* __restrict__ mu) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i >= N) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) {
    const long j = 2 * ((long)b * N + i);
    s += fz[j] * fU[j] + fz[j + 1] * fU[j + 1];
  }
  mu[i] += scale * s;
}

// d v_xc / d rho of the PZ81 LDA (same constants as update_vhxc), times mu
__global__ void k_fxc_mul(long N, const double* __restrict__ rho, const double* __restrict__ mu,
                          double* __restrict__ nu) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i >= N) return;
  const double c1 = 0.6203504908994001, c3 = -0.610887057711;
  const double A = 0.0311, B = -0.048, b1 = 1.0529, b2 = 0.3334, G = -0.1423;
  const double D = G / (1.0 + b1 + b2) - B;
  const double C = -A - D - G * ((b1 / 2.0 + b2) / ((1.0 + b1 + b2) * (1.0 + b1 + b2)));
  const double r = rho[i];
  double f = 0.0;
  if (r > 0.0) {
    const double rs = c1 / cbrt(r);
    const double dvx = (c3 / rs) / (3.0 * r);
    double dvc;
    if (rs < 1.0) {
      dvc = A / rs + (2.0 / 3.0) * C * (log(rs) + 1.0) + (2.0 * D - C) / 3.0;
    } else {
      const double s = sqrt(rs);
      const double den = 1.0 + b1 * s + b2 * rs;
      const double num = 1.0 + (7.0 / 6.0) * b1 * s + (4.0 / 3.0) * b2 * rs;
      const double dnum = (7.0 / 6.0) * b1 + (8.0 / 3.0) * b2 * s;
      const double dden = b1 + 2.0 * b2 * s;
      dvc = G * (dnum * den - 2.0 * num * dden) / (den * den * den) / (2.0 * s);
    }
    f = dvx + dvc * (-rs / (3.0 * r));
  }
  nu[i] = f * mu[i];
}

// x(G) *= 4 pi / G^2 (0 at G = 0), the Hartree kernel of update_vhxc
__global__ void k_hartree_scale(int n, const double* __restrict__ g2, double* __restrict__ x) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double gg = g2[i];
  const double gi = gg > 0.0 ? 1.0 / sqrt(gg) : 0.0;
  const double s = 4.0 * M_PI * gi * gi;
  x[2 * i] *= s;
  x[2 * i + 1] *= s;
}

__global__ void k_add_re(long N, const double* __restrict__ f, double* __restrict__ y) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i < N) y[i] += f[2 * i];
}

// lam_b += scale * f_b * x_b
__global__ void k_band_axpy(int ngw, int nst, const double* __restrict__ f, double scale,
                            const double* __restrict__ x, double* __restrict__ lam) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i >= (long)ngw * nst) return;
  const double s = scale * f[i / ngw];
  lam[2 * i] += s * x[2 * i];
  lam[2 * i + 1] += s * x[2 * i + 1];
}

inline unsigned nb256(long n) { return (unsigned)((n + 255) / 256); }

// out[i] = sum over the S slices of part (slice stride n), i < n
__global__ void __launch_bounds__(256)
k_sum_slices(long n, int S, const double* __restrict__ part, double* __restrict__ out) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    double a = part[i];
    for (int s = 1; s < S; ++s) a += part[(long)s * n + i];
    out[i] = a;
  }
}

}  // namespace

struct qbh_td_s {
  qbh_ham hh = nullptr;
  Ham* h = nullptr;
  qbh_basis vb = nullptr;
  Ctx* c = nullptr;
  int nst = 0, ngw = 0, ngv = 0;
  long N = 0;
  std::vector<double> occ, jfac;
  double weight = 1.0, omega = 0.0;
  const double* vion = nullptr;
  const double* rhops = nullptr;
  // last forward run
  bool have_run = false;
  int nsteps = 0, order = 4, ckpt = 20, flags = 0;
  double dt = 0.0;
  std::vector<double> A, w, jhost;
  double e[3] = {1.0, 0.0, 0.0};
  const double* vext = nullptr;
  // device memory, one allocation carved into pieces
  void* mem = nullptr;
  size_t mem_bytes = 0;
  double *fj = nullptr, *part = nullptr, *jacc = nullptr, *gA = nullptr;
  double *cur = nullptr, *nxt = nullptr, *psiM = nullptr, *tilde = nullptr, *half = nullptr;
  double *rhog = nullptr, *ckpts = nullptr, *seg = nullptr;
  double *lam = nullptr, *lam_h = nullptr, *lam_t = nullptr, *tmp = nullptr, *zk = nullptr,
         *ubuf = nullptr;
  double *mu = nullptr, *nu = nullptr, *gv = nullptr, *fgrid = nullptr, *gtmp = nullptr,
         *grids = nullptr;
  double *va = nullptr, *vb_buf = nullptr, *rho_a = nullptr, *rho_b = nullptr;
  int nb_cross = 1;
  // A-derivative of species-tabled projectors: rows at A and their q-gradient (3 blocks),
  // [anl | danl_x | danl_y | danl_z] (ngw x 4P complex), the products with z_k and U_k
  bool nl_adj = false;
  int nrows = 0;
  double *tw = nullptr, *dtw = nullptr, *anl4 = nullptr, *fz4 = nullptr, *fu4 = nullptr;
  // combined table of anl4 for the matrix-core products (8P rows of 2ngw reals) and their
  // K-slice partials (nsl slices of 8P x order*nst reals)
  double *anl8 = nullptr, *fsl = nullptr;
  int nsl = 1;
  bool tab_ok = false;  // anl4 holds the tables at tab_A
  double tab_A[3] = {0.0, 0.0, 0.0};
  size_t blk() const { return 2 * (size_t)ngw * nst; }
  size_t seg_step() const { return (3 * (size_t)order + 1) * blk() + 4 * (size_t)N; }
};

static int tfail(Ctx* c, int code, const std::string& m) {
  c->last_error = m;
  return code;
}

static int tcheck(Ctx* c, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return tfail(c, QBH_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  return QBH_OK;
}

static int set_A(qbh_td_s* t, const double* A) {
  Ham* h = t->h;
  if (h->A[0] == A[0] && h->A[1] == A[1] && h->A[2] == A[2] && h->kin) return QBH_OK;
  return qbh_ham_set_vector_potential(t->hh, A);
}

// v = update_vhxc(rho(c)) + v_ext; rho kept in `rho`
static int potential(qbh_td_s* t, const double* c, double* rho, double* v) {
  Ctx* cx = t->c;
  k_fill<<<nb256(t->N), 256, 0, cx->stream>>>(rho, t->N);
  if (int r = qbh_density(reinterpret_cast<qbh_basis>(t->h->b_handle), c, t->ngw, t->nst,
                          t->occ.data(), t->weight, t->omega, rho, t->flags & QBH_FP32))
    return r;
  if (int r = qbh_fft_forward_real(t->vb, rho, t->omega, t->rhog, 0)) return r;
  if (int r = qbh_update_vhxc(t->vb, rho, t->rhog, t->rhog, t->vion, t->rhops, t->omega, v,
                              nullptr))
    return r;
  if (t->vext) k_axpy_real<<<nb256(t->N), 256, 0, cx->stream>>>(t->N, 1.0, t->vext, v);
  return tcheck(cx, "td potential");
}

// current sums of step n into jacc[2n], jacc[2n+1]
static void current_sums(qbh_td_s* t, const double* c, int n) {
  Ctx* cx = t->c;
  k_cur_part<<<NBLK, 256, 0, cx->stream>>>(t->ngw, t->nst, t->h->b->kpgx, t->e[0], t->e[1],
                                           t->e[2], t->fj, c, t->part);
  k_final<<<1, 256, 0, cx->stream>>>(NBLK, 2, t->part, 1.0, 0, t->jacc + 2 * n);
}

// one TD iteration psi_n -> psi_{n+1}; st (nullable) receives the step's stored state:
// [u1 (order blocks) | u2 (order) | u3 (order) | tilde | rho_a | rho_b | v_a | v_b]
static int step_forward(qbh_td_s* t, int n, const double* in, double* out, double* st) {
  Ham* h = t->h;
  const size_t B = t->blk();
  const int o = t->order;
  double* u1 = st ? st : nullptr;
  double* u2 = st ? st + o * B : nullptr;
  double* u3 = st ? st + 2 * o * B : nullptr;
  double* tl = st ? st + 3 * o * B : t->tilde;
  double* ra = st ? tl + B : t->rho_a;
  double* rb = st ? ra + t->N : t->rho_b;
  double* va = st ? rb + t->N : t->va;
  double* vbf = st ? va + t->N : t->vb_buf;
  const int fl = whole_h_flags(t->flags);
  if (int r = potential(t, in, ra, va)) return r;
  h->v = va;
  if (int r = set_A(t, &t->A[3 * n])) return r;
  if (int r = expo_impl(h, in, t->ngw, t->half, t->ngw, t->nst, 0.5 * t->dt, o, fl, u1)) return r;
  if (int r = expo_impl(h, t->half, t->ngw, tl, t->ngw, t->nst, 0.5 * t->dt, o, fl, u2)) return r;
  if (int r = potential(t, tl, rb, vbf)) return r;
  h->v = vbf;
  if (int r = set_A(t, &t->A[3 * (n + 1)])) return r;
  return expo_impl(h, t->half, t->ngw, out, t->ngw, t->nst, 0.5 * t->dt, o, fl, u3);
}

// local-potential and A gradients of one exponential (forward step dt/2 on u, adjoint
// Krylov vectors zk): mu += (2/N) Re sum conj(z_k(r)) U_k(r); gA3 += -2 Re sum conj(z_k)
// (K - A)_d U_k
static int cross(qbh_td_s* t, const double* zk, const double* u, const double* A, double* mu,
                 double* gA3) {
  Ctx* cx = t->c;
  const size_t B = t->blk();
  const int o = t->order;
  const long nc = (long)t->ngw * t->nst;
  std::vector<std::complex<double>> a(o + 1);
  a[0] = 1.0;
  for (int N = 1; N <= o; ++N) a[N] = a[N - 1] * std::complex<double>(0.0, -0.5 * t->dt / N);
  qbh_basis hb = reinterpret_cast<qbh_basis>(t->h->b_handle);
  Ham* h = t->h;
  const int P = h->P;
  const bool nl = t->nl_adj && (whole_h_flags(t->flags) & QBH_NONLOCAL);
  if (nl && !(t->tab_ok && A[0] == t->tab_A[0] && A[1] == t->tab_A[1] && A[2] == t->tab_A[2])) {
    // tables at this exponential's A: twnl rows, their gradient, anl and danl_d (kept for
    // the next exponential under the same A)
    t->tab_ok = false;
    const long rs = (long)t->nrows * t->ngw;
    long row0 = 0;
    for (qbh_species sp : h->sps) {
      if (int r = species_table(sp, cx, h->b, A, 0, t->tw + row0 * t->ngw, 0)) return r;
      if (int r = species_table(sp, cx, h->b, A, 1, t->dtw + row0 * t->ngw, rs)) return r;
      row0 += species_nproj(sp);
    }
    const size_t cs = 2 * (size_t)t->ngw * P;
    if (int r = build_anl_c64(h, t->tw, A, t->anl4)) return r;
    for (int d = 0; d < 3; ++d)
      if (int r = build_anl_c64(h, t->dtw + d * rs, A, t->anl4 + (1 + d) * cs)) return r;
    if (int r = anl_pair_c64(cx, t->ngw, 4 * P, t->anl4, t->anl8)) return r;
    t->tab_ok = true;
    std::copy(A, A + 3, t->tab_A);
  }
  // F = [anl | danl_x | danl_y | danl_z]^H x for the order blocks of x at once: one real
  // matrix-core product over the combined table (row 2r: Re, row 2r + 1: Im of row r of
  // F, which is the interleaved complex layout), K slices summed into out
  auto products = [&](const double* x, double* out) -> bool {
    const long R8 = 8L * P, C = (long)o * t->nst, Q = 2L * t->ngw;
    if (mfma_gemm_qq(cx->stream, false, (int)R8, (int)C, Q, t->anl8, Q, x, Q, t->fsl, R8,
                     t->nsl, R8 * C))
      return false;
    k_sum_slices<<<nb256(R8 * C), 256, 0, cx->stream>>>(R8 * C, t->nsl, t->fsl, out);
    return true;
  };
  const int nb = t->nb_cross;
  double* fz = t->grids;
  double* fU = t->grids + 2 * (size_t)nb * t->N;
  // U_k = sum_m alpha_{k+m+1} u_m for every k (order blocks, like zk)
  for (int k = 0; k < o; ++k) {
    Combo cb{};
    cb.cnt = o - k;
    for (int m = 0; m < cb.cnt; ++m) {
      cb.u[m] = u + m * B;
      cb.ar[m] = a[k + m + 1].real();
      cb.ai[m] = a[k + m + 1].imag();
    }
    k_combo<<<nb256(nc), 256, 0, cx->stream>>>(nc, cb, t->ubuf + k * B);
  }
  if (nl) {
    if (!products(zk, t->fz4) || !products(t->ubuf, t->fu4))
      return tfail(cx, QBH_ERR_ARG, "td projector products: invalid shape");
    k_nl_dA_part<<<NBLK, 256, 0, cx->stream>>>(P, o * t->nst, h->col_fac, t->fz4, t->fu4,
                                               t->part);
    k_final<<<1, 256, 0, cx->stream>>>(NBLK, 3, t->part, -2.0, 1, gA3);
  }
  for (int k = 0; k < o; ++k) {
    const double* z = zk + k * B;
    const double* U = t->ubuf + k * B;
    k_crossk_part<<<NBLK, 256, 0, cx->stream>>>(t->ngw, t->nst, t->h->b->kpgx, A[0], A[1], A[2],
                                               z, U, t->part);
    k_final<<<1, 256, 0, cx->stream>>>(NBLK, 3, t->part, -2.0, 1, gA3);
    for (int b0 = 0; b0 < t->nst; b0 += nb) {
      const int m = std::min(nb, t->nst - b0);
      if (int r = qbh_fft_backward(hb, z + 2 * (size_t)b0 * t->ngw, t->ngw, m, 0, fz, 0)) return r;
      if (int r = qbh_fft_backward(hb, U + 2 * (size_t)b0 * t->ngw, t->ngw, m, 0, fU, 0))
        return r;
      k_cross_acc<<<nb256(t->N), 256, 0, cx->stream>>>(t->N, m, 2.0 / (double)t->N, fz, fU, mu);
    }
  }
  return tcheck(cx, "td cross");
}

// nu = f_xc(rho) mu + K_H mu
static int vhxc_adjoint(qbh_td_s* t, const double* rho, const double* mu, double* nu) {
  Ctx* cx = t->c;
  k_fxc_mul<<<nb256(t->N), 256, 0, cx->stream>>>(t->N, rho, mu, nu);
  if (int r = qbh_fft_forward_real(t->vb, mu, 1.0, t->gtmp, 0)) return r;
  k_hartree_scale<<<nb256(t->ngv), 256, 0, cx->stream>>>(t->ngv, basis_of(t->vb)->kpg2, t->gtmp);
  if (int r = qbh_fft_backward(t->vb, t->gtmp, t->ngv, 1, 0, t->fgrid, 0)) return r;
  k_add_re<<<nb256(t->N), 256, 0, cx->stream>>>(t->N, t->fgrid, nu);
  return tcheck(cx, "td vhxc adjoint");
}

// lam += (f_b N / Omega) V_loc[nu] c
static int density_adjoint(qbh_td_s* t, const double* c, double* nu, double* lam) {
  Ham* h = t->h;
  double* save = h->v;
  h->v = nu;
  std::vector<double> occ0(t->nst, 0.0);
  int r = hpsi_impl(h, c, t->ngw, t->nst, occ0.data(), t->tmp, t->ngw,
                    QBH_LOCAL | (t->flags & QBH_FP32), nullptr);
  h->v = save;
  if (r) return r;
  const long nc = (long)t->ngw * t->nst;
  k_band_axpy<<<nb256(nc), 256, 0, t->c->stream>>>(t->ngw, t->nst, t->fj,
                                                   (double)t->N / t->omega, t->tmp, lam);
  return tcheck(t->c, "td density adjoint");
}

static double* carve(char*& p, size_t doubles) {
  double* r = reinterpret_cast<double*>(p);
  p += ((doubles * sizeof(double) + 255) / 256) * 256;
  return r;
}

static int ensure_memory(qbh_td_s* t) {
  const size_t B = t->blk();
  const size_t N = t->N;
  const int nck = (t->nsteps + t->ckpt - 1) / t->ckpt;
  // bands per transform chunk of the gradient products: two complex grids per band
  t->nb_cross = (int)std::max<long>(1, std::min<long>(t->nst, (1L << 30) / (32L * (long)N)));
  size_t need = 0;
  auto add = [&](size_t d) { need += ((d * sizeof(double) + 255) / 256) * 256; };
  const size_t sizes[] = {(size_t)t->nst, 2 * (size_t)NBLK * 3, 2 * (size_t)(t->nsteps + 1),
                          3 * (size_t)(t->nsteps + 1), B, B, B, B, B, 2 * (size_t)t->ngv,
                          nck * B, t->ckpt * t->seg_step(), B, B, B, B, t->order * B,
                          t->order * B, N, N, N,
                          2 * N, 2 * (size_t)t->ngv, 4 * (size_t)t->nb_cross * N, N, N, N, N};
  for (size_t s : sizes) add(s);
  Ham* h = t->h;
  t->nl_adj = !h->sps.empty() && h->P > 0;
  t->nrows = 0;
  for (qbh_species sp : h->sps) t->nrows += species_nproj(sp);
  const size_t rows_d = (size_t)t->nrows * t->ngw;
  t->nsl = h->P > 0 ? mfma_slices(8 * h->P, t->order * t->nst, 2L * t->ngw, 16) : 1;
  const size_t tabs[] = {rows_d, 3 * rows_d, 8 * (size_t)t->ngw * h->P,
                         8 * (size_t)h->P * t->nst * t->order,
                         8 * (size_t)h->P * t->nst * t->order,
                         16 * (size_t)t->ngw * h->P,
                         (size_t)t->nsl * 8 * h->P * t->nst * t->order};
  t->tab_ok = false;
  if (t->nl_adj)
    for (size_t s : tabs) add(s);
  if (need > t->mem_bytes) {
    if (t->mem) hipFree(t->mem);
    t->mem = nullptr;
    t->mem_bytes = 0;
    if (hipMalloc(&t->mem, need) != hipSuccess) {
      t->mem = nullptr;
      return tfail(t->c, QBH_ERR_ALLOC, "TD buffers allocation failed (" +
                                            std::to_string(need >> 20) + " MiB)");
    }
    t->mem_bytes = need;
  }
  char* p = static_cast<char*>(t->mem);
  t->fj = carve(p, t->nst);
  t->part = carve(p, 2 * (size_t)NBLK * 3);
  t->jacc = carve(p, 2 * (size_t)(t->nsteps + 1));
  t->gA = carve(p, 3 * (size_t)(t->nsteps + 1));
  t->cur = carve(p, B);
  t->nxt = carve(p, B);
  t->psiM = carve(p, B);
  t->tilde = carve(p, B);
  t->half = carve(p, B);
  t->rhog = carve(p, 2 * (size_t)t->ngv);
  t->ckpts = carve(p, nck * B);
  t->seg = carve(p, t->ckpt * t->seg_step());
  t->lam = carve(p, B);
  t->lam_h = carve(p, B);
  t->lam_t = carve(p, B);
  t->tmp = carve(p, B);
  t->zk = carve(p, t->order * B);
  t->ubuf = carve(p, t->order * B);
  t->mu = carve(p, N);
  t->nu = carve(p, N);
  t->gv = carve(p, N);
  t->fgrid = carve(p, 2 * N);
  t->gtmp = carve(p, 2 * (size_t)t->ngv);
  t->grids = carve(p, 4 * (size_t)t->nb_cross * N);
  t->va = carve(p, N);
  t->vb_buf = carve(p, N);
  t->rho_a = carve(p, N);
  t->rho_b = carve(p, N);
  if (t->nl_adj) {
    t->tw = carve(p, tabs[0]);
    t->dtw = carve(p, tabs[1]);
    t->anl4 = carve(p, tabs[2]);
    t->fz4 = carve(p, tabs[3]);
    t->fu4 = carve(p, tabs[4]);
    t->anl8 = carve(p, tabs[5]);
    t->fsl = carve(p, tabs[6]);
  }
  hipMemcpyAsync(t->fj, t->jfac.data(), t->nst * sizeof(double), hipMemcpyHostToDevice,
                 t->c->stream);
  return QBH_OK;
}

extern "C" int qbh_td_create(qbh_ham ham, qbh_basis vbasis, int nst, const double* occ,
                             double weight, double omega, const double* vion_g_dev,
                             const double* rhops_g_dev, qbh_td* out) {
  if (!ham || !vbasis || !occ || nst <= 0 || !(omega > 0.0) || !out) return QBH_ERR_ARG;
  Ham* h = &ham->h;
  Ctx* c = h->ctx;
  const GridDev& g = h->b->g;
  const GridDev& vg = basis_of(vbasis)->g;
  if (g.real) return tfail(c, QBH_ERR_ARG, "TD propagation needs a complex basis");
  if (vg.real) return tfail(c, QBH_ERR_ARG, "TD density basis must be complex");
  if (ctx_of_basis(vbasis) != c) return tfail(c, QBH_ERR_ARG, "vbasis on another context");
  if (vg.np0 != g.np0 || vg.np1 != g.np1 || vg.np2 != g.np2)
    return tfail(c, QBH_ERR_ARG, "density basis grid differs from the wavefunction grid");
  auto* t = new qbh_td_s();
  t->hh = ham;
  t->h = h;
  t->vb = vbasis;
  t->c = c;
  t->nst = nst;
  t->ngw = g.ngw;
  t->ngv = vg.ngw;
  t->N = (long)g.np0 * g.np1 * g.np2;
  t->occ.assign(occ, occ + nst);
  t->weight = weight;
  t->omega = omega;
  t->vion = vion_g_dev;
  t->rhops = rhops_g_dev;
  t->jfac.resize(nst);
  for (int b = 0; b < nst; ++b) {
    const double f = weight * occ[b];
    t->jfac[b] = (f / omega > 0.0) ? f : 0.0;  // bands with positive density factor
  }
  *out = t;
  return QBH_OK;
}

extern "C" int qbh_td_destroy(qbh_td t) {
  if (!t) return QBH_ERR_ARG;
  hipSetDevice(t->c->device);
  hipStreamSynchronize(t->c->stream);
  if (t->mem) hipFree(t->mem);
  delete t;
  return QBH_OK;
}

extern "C" int qbh_td_forward(qbh_td t, const double* psi0_dev, long ld, int nsteps, double dt,
                              int order, const double* A, const double* w, const double* e,
                              const double* vext_dev, int checkpoint, int flags, double* J,
                              double* psi_out_dev, long ldo) {
  if (!t || !psi0_dev || nsteps < 0 || order < 1 || order > 8 || !A || !w || !e ||
      checkpoint < 1 || !J)
    return QBH_ERR_ARG;
  Ctx* c = t->c;
  hipSetDevice(c->device);
  if (ld < t->ngw || (psi_out_dev && ldo < t->ngw)) return tfail(c, QBH_ERR_ARG, "ld < ngw");
  t->have_run = false;
  t->nsteps = nsteps;
  t->dt = dt;
  t->order = order;
  t->ckpt = std::min(checkpoint, std::max(1, nsteps));
  t->flags = flags;
  t->A.assign(A, A + 3 * (size_t)(nsteps + 1));
  t->w.assign(w, w + (size_t)nsteps + 1);
  std::memcpy(t->e, e, sizeof(t->e));
  t->vext = vext_dev;
  if (int r = ensure_memory(t)) return r;
  Ham* h = t->h;
  double* v_own = h->v;
  double A_own[3] = {h->A[0], h->A[1], h->A[2]};
  const size_t B = t->blk();
  hipMemcpy2DAsync(t->cur, 2 * sizeof(double) * t->ngw, psi0_dev, 2 * sizeof(double) * ld,
                   2 * sizeof(double) * t->ngw, t->nst, hipMemcpyDeviceToDevice, c->stream);
  int r = QBH_OK;
  for (int n = 0; n < nsteps && r == QBH_OK; ++n) {
    if (n % t->ckpt == 0)
      hipMemcpyAsync(t->ckpts + (n / t->ckpt) * B, t->cur, B * sizeof(double),
                     hipMemcpyDeviceToDevice, c->stream);
    current_sums(t, t->cur, n);
    r = step_forward(t, n, t->cur, t->nxt, nullptr);
    std::swap(t->cur, t->nxt);
  }
  if (r == QBH_OK) {
    current_sums(t, t->cur, nsteps);
    hipMemcpyAsync(t->psiM, t->cur, B * sizeof(double), hipMemcpyDeviceToDevice, c->stream);
    if (psi_out_dev)
      hipMemcpy2DAsync(psi_out_dev, 2 * sizeof(double) * ldo, t->cur, 2 * sizeof(double) * t->ngw,
                       2 * sizeof(double) * t->ngw, t->nst, hipMemcpyDeviceToDevice, c->stream);
    t->jhost.assign(2 * (size_t)(nsteps + 1), 0.0);
    hipMemcpyAsync(t->jhost.data(), t->jacc, t->jhost.size() * sizeof(double),
                   hipMemcpyDeviceToHost, c->stream);
    hipStreamSynchronize(c->stream);
    double s = 0.0;
    for (int n = 0; n <= nsteps; ++n) {
      const double ea = e[0] * A[3 * n] + e[1] * A[3 * n + 1] + e[2] * A[3 * n + 2];
      s += w[n] * (-t->jhost[2 * n] + ea * t->jhost[2 * n + 1]);
    }
    *J = s;
    r = tcheck(c, "td forward");
  }
  // restore the caller's potential and vector potential on every exit path, keeping the
  // first error
  h->v = v_own;
  const int ra = qbh_ham_set_vector_potential(t->hh, A_own);
  if (r == QBH_OK) r = ra;
  if (r == QBH_OK) t->have_run = true;
  return r;
}

extern "C" int qbh_td_backward(qbh_td t, double* dvext_dev, double* dA) {
  if (!t) return QBH_ERR_ARG;
  Ctx* c = t->c;
  hipSetDevice(c->device);
  if (!t->have_run) return tfail(c, QBH_ERR_ARG, "qbh_td_backward needs a completed forward run");
  Ham* h = t->h;
  double* v_own = h->v;
  double A_own[3] = {h->A[0], h->A[1], h->A[2]};
  const size_t B = t->blk();
  const int M = t->nsteps, o = t->order;
  const long nc = (long)t->ngw * t->nst;
  const int fl = whole_h_flags(t->flags);
  const double hdt = 0.5 * t->dt;
  k_fill<<<nb256(3L * (M + 1)), 256, 0, c->stream>>>(t->gA, 3L * (M + 1));
  k_fill<<<nb256(t->N), 256, 0, c->stream>>>(t->gv, t->N);
  k_fill<<<nb256(2 * nc), 256, 0, c->stream>>>(t->lam, 2 * nc);
  const double* Ad = t->A.data();
  const double eaM = t->e[0] * Ad[3 * M] + t->e[1] * Ad[3 * M + 1] + t->e[2] * Ad[3 * M + 2];
  k_dj<<<nb256(nc), 256, 0, c->stream>>>(t->ngw, t->nst, h->b->kpgx, t->e[0], t->e[1], t->e[2],
                                         eaM, t->fj, t->w[M], t->psiM, t->lam);
  int r = tcheck(c, "td backward start");
  const int nseg = (M + t->ckpt - 1) / t->ckpt;
  for (int sgi = nseg - 1; sgi >= 0 && r == QBH_OK; --sgi) {
    const int s0 = sgi * t->ckpt, s1 = std::min(M, s0 + t->ckpt);
    // recompute the segment, storing each step's Krylov vectors and potentials
    hipMemcpyAsync(t->cur, t->ckpts + sgi * B, B * sizeof(double), hipMemcpyDeviceToDevice,
                   c->stream);
    for (int n = s0; n < s1 && r == QBH_OK; ++n) {
      r = step_forward(t, n, t->cur, t->nxt, t->seg + (n - s0) * t->seg_step());
      std::swap(t->cur, t->nxt);
    }
    for (int n = s1 - 1; n >= s0 && r == QBH_OK; --n) {
      double* st = t->seg + (n - s0) * t->seg_step();
      double* u1 = st;
      double* u2 = st + o * B;
      double* u3 = st + 2 * o * B;
      double* tl = st + 3 * o * B;
      double* ra = tl + B;
      double* rb = ra + t->N;
      double* va = rb + t->N;
      double* vbf = va + t->N;
      const double* An = Ad + 3 * n;
      const double* An1 = Ad + 3 * (n + 1);
      // psi_{n+1} = T_b half
      h->v = vbf;
      if ((r = set_A(t, An1))) break;
      if ((r = expo_impl(h, t->lam, t->ngw, t->lam_h, t->ngw, t->nst, -hdt, o, fl, t->zk))) break;
      k_fill<<<nb256(t->N), 256, 0, c->stream>>>(t->mu, t->N);
      if ((r = cross(t, t->zk, u3, An1, t->mu, t->gA + 3 * (n + 1)))) break;
      k_axpy_real<<<nb256(t->N), 256, 0, c->stream>>>(t->N, 1.0, t->mu, t->gv);
      // v_b = V[rho(tilde)]
      if ((r = vhxc_adjoint(t, rb, t->mu, t->nu))) break;
      k_fill<<<nb256(2 * nc), 256, 0, c->stream>>>(t->lam_t, 2 * nc);
      if ((r = density_adjoint(t, tl, t->nu, t->lam_t))) break;
      // tilde = T_a half
      h->v = va;
      if ((r = set_A(t, An))) break;
      if ((r = expo_impl(h, t->lam_t, t->ngw, t->tmp, t->ngw, t->nst, -hdt, o, fl, t->zk))) break;
      k_axpy_real<<<nb256(2 * nc), 256, 0, c->stream>>>(2 * nc, 1.0, t->tmp, t->lam_h);
      k_fill<<<nb256(t->N), 256, 0, c->stream>>>(t->mu, t->N);
      if ((r = cross(t, t->zk, u2, An, t->mu, t->gA + 3 * n))) break;
      // half = T_a psi_n
      if ((r = expo_impl(h, t->lam_h, t->ngw, t->lam, t->ngw, t->nst, -hdt, o, fl, t->zk))) break;
      if ((r = cross(t, t->zk, u1, An, t->mu, t->gA + 3 * n))) break;
      k_axpy_real<<<nb256(t->N), 256, 0, c->stream>>>(t->N, 1.0, t->mu, t->gv);
      // v_a = V[rho(psi_n)]
      if ((r = vhxc_adjoint(t, ra, t->mu, t->nu))) break;
      if ((r = density_adjoint(t, u1, t->nu, t->lam))) break;
      // current at step n
      const double ea = t->e[0] * An[0] + t->e[1] * An[1] + t->e[2] * An[2];
      k_dj<<<nb256(nc), 256, 0, c->stream>>>(t->ngw, t->nst, h->b->kpgx, t->e[0], t->e[1],
                                             t->e[2], ea, t->fj, t->w[n], u1, t->lam);
      r = tcheck(c, "td backward step");
    }
  }
  std::vector<double> ga(3 * (size_t)(M + 1), 0.0);
  if (r == QBH_OK) {
    hipMemcpyAsync(ga.data(), t->gA, ga.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (dvext_dev)
      hipMemcpyAsync(dvext_dev, t->gv, t->N * sizeof(double), hipMemcpyDeviceToDevice, c->stream);
  }
  hipStreamSynchronize(c->stream);
  h->v = v_own;
  if (r == QBH_OK) r = tcheck(c, "td backward");
  // restore the vector potential on every exit path, keeping the first error
  const int ra = qbh_ham_set_vector_potential(t->hh, A_own);
  if (r == QBH_OK) r = ra;
  if (r == QBH_OK && dA) {
    for (int n = 0; n <= M; ++n)
      for (int d = 0; d < 3; ++d) dA[3 * n + d] = ga[3 * n + d] + t->w[n] * t->e[d] * t->jhost[2 * n + 1];
  }
  return r;
}
