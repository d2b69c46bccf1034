// Self-consistent local potential on the density grid (EnergyFunctional::update_vhxc,
// src/qball/EnergyFunctional.cc:353-977, non-spin-polarised LDA, no ESM):
//   v_r = v_xc(rho(r)) + FFT^-1[ v_ion(G) + 4 pi (rho(G)/Omega + rho_ps(G)) / G^2 ]
// with the PZ81 LDA of LDAFunctional::xc_unpolarized (functionals/LDAFunctional.cc) and
// the energies exc, eps (= Omega <rho|v_ion>) and ehart computed on the way.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/qbh.h"
#include "qbh_internal.h"

namespace qbh {
Basis* basis_of(qbh_basis b);
Ctx* ctx_of_basis(qbh_basis b);
}  // namespace qbh

using namespace qbh;

// Perdew-Zunger 1981 parametrisation of Ceperley-Alder correlation plus Slater exchange
// (Phys. Rev. B 23, 5048), unpolarised, with C and D set by continuity of e_c and v_c at
// rs = 1 (the reference's choice, LDAFunctional.cc).  Returns e_xc per electron and v_xc.
__device__ __forceinline__ void pz81(double rh, double& ee, double& vv) {
  const double c1 = 0.6203504908994001;  // (3/(4 pi))^(1/3)
  const double c3 = -0.610887057711;     // -(4/3) (3/(4 pi)) / (4/(9 pi))^(1/3)
  const double A = 0.0311, B = -0.048, b1 = 1.0529, b2 = 0.3334, G = -0.1423;
  const double D = G / (1.0 + b1 + b2) - B;
  const double C = -A - D - G * ((b1 / 2.0 + b2) / ((1.0 + b1 + b2) * (1.0 + b1 + b2)));
  ee = 0.0;
  vv = 0.0;
  if (rh > 0.0) {
    const double ro13 = cbrt(rh);
    const double rs = c1 / ro13;
    const double vx = c3 / rs;
    const double ex = 0.75 * vx;
    double ec, vc;
    if (rs < 1.0) {
      const double logrs = log(rs);
      ec = A * logrs + B + C * rs * logrs + D * rs;
      vc = A * logrs + (B - A / 3.0) + (2.0 / 3.0) * C * rs * logrs + ((2.0 * D - C) / 3.0) * rs;
    } else {
      const double sqrtrs = sqrt(rs);
      const double den = 1.0 + b1 * sqrtrs + b2 * rs;
      ec = G / den;
      vc = ec * (1.0 + (7.0 / 6.0) * b1 * sqrtrs + (4.0 / 3.0) * b2 * rs) / den;
    }
    ee = ex + ec;
    vv = vx + vc;
  }
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

// v_r = v_xc(rho); partial[block] = sum rho * e_xc
__global__ void __launch_bounds__(256)
k_lda(long n, const double* __restrict__ rho, double* __restrict__ vr, double* __restrict__ partial) {
  __shared__ double red[256];
  double e = 0.0;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    double ee, vv;
    pz81(rho[i], ee, vv);
    vr[i] = vv;
    e += rho[i] * ee;
  }
  const double s = block_sum(e, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// vlocal_g = v_ion + 4 pi rhogt g2i, rhogt = rhog/Omega + rho_ps (Hamiltonian density for
// the Hartree term when hrhog is given: EnergyFunctional.cc:473-487, 512-514);
// partials: eps (ddot of rhoelg and v_ion) and ehart sums
__global__ void __launch_bounds__(256)
k_vlocal_g(int ngw, const double* __restrict__ g2, const double* __restrict__ rhog,
           const double* __restrict__ hrhog, const double* __restrict__ vion,
           const double* __restrict__ rhops, double omega_inv, double* __restrict__ vg,
           double* __restrict__ peps, double* __restrict__ peh) {
  __shared__ double red[256];
  const double fpi = 4.0 * M_PI;
  double se = 0.0, sh = 0.0;
  for (int ig = blockIdx.x * 256 + threadIdx.x; ig < ngw; ig += gridDim.x * 256) {
    const double gg = g2[ig];
    const double gi = gg > 0.0 ? 1.0 / sqrt(gg) : 0.0;  // Basis::update_g: gi = 1/|G|
    const double g2i = gi * gi;
    const double er = omega_inv * rhog[2 * ig], ei = omega_inv * rhog[2 * ig + 1];
    const double vr = vion ? vion[2 * ig] : 0.0, vi = vion ? vion[2 * ig + 1] : 0.0;
    se += er * vr + ei * vi;
    double tr = er, ti = ei;
    if (rhops) {
      tr += rhops[2 * ig];
      ti += rhops[2 * ig + 1];
    }
    double hr = tr, hi = ti;
    if (hrhog) {
      hr = omega_inv * hrhog[2 * ig] + (rhops ? rhops[2 * ig] : 0.0);
      hi = omega_inv * hrhog[2 * ig + 1] + (rhops ? rhops[2 * ig + 1] : 0.0);
      sh += (tr * hr + ti * hi) * g2i;  // Re(tmp conj(hamil_tmp)) g2i
    } else {
      sh += (tr * tr + ti * ti) * g2i;
    }
    vg[2 * ig] = vr + fpi * hr * g2i;
    vg[2 * ig + 1] = vi + fpi * hi * g2i;
  }
  const double a = block_sum(se, red);
  const double b = block_sum(sh, red);
  if (threadIdx.x == 0) {
    peps[blockIdx.x] = a;
    peh[blockIdx.x] = b;
  }
}

__global__ void k_add_real_part(long n, const double* __restrict__ f, double* __restrict__ vr) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i < n) vr[i] += f[2 * i];
}

static int vfail(Ctx* c, int code, const std::string& m) {
  c->last_error = m;
  return code;
}

extern "C" int qbh_update_vhxc(qbh_basis vbasis, const double* rhor_dev, const double* rhog_dev,
                               const double* hamil_rhog_dev, const double* vion_g_dev,
                               const double* rhops_g_dev, double omega, double* v_r_dev,
                               double* energies) {
  if (!vbasis || !rhor_dev || !rhog_dev || !v_r_dev || !(omega > 0.0)) return QBH_ERR_ARG;
  Basis* b = basis_of(vbasis);
  Ctx* c = ctx_of_basis(vbasis);
  hipSetDevice(c->device);
  const GridDev& g = b->g;
  const long N = (long)g.np0 * g.np1 * g.np2;
  const int ngw = g.ngw;
  const int nb = 256;
  double *part = nullptr, *vg = nullptr, *f = nullptr;
  if (hipMalloc(&part, 3 * nb * sizeof(double)) != hipSuccess ||
      hipMalloc(&vg, 2 * sizeof(double) * ngw) != hipSuccess ||
      hipMalloc(&f, 2 * sizeof(double) * N) != hipSuccess) {
    for (void* p : {(void*)part, (void*)vg, (void*)f})
      if (p) hipFree(p);
    return vfail(c, QBH_ERR_ALLOC, "update_vhxc scratch allocation failed");
  }
  {
    TimedLaunch tl(c, "vhxc_lda");
    k_lda<<<nb, 256, 0, c->stream>>>(N, rhor_dev, v_r_dev, part);
  }
  {
    TimedLaunch tl(c, "vhxc_g");
    k_vlocal_g<<<nb, 256, 0, c->stream>>>(ngw, b->kpg2, rhog_dev, hamil_rhog_dev, vion_g_dev,
                                          rhops_g_dev, 1.0 / omega, vg, part + nb, part + 2 * nb);
  }
  int r = qbh_fft_backward(vbasis, vg, ngw, 1, 0, f, 0);
  if (r == QBH_OK) {
    TimedLaunch tl(c, "vhxc_add");
    k_add_real_part<<<(unsigned)((N + 255) / 256), 256, 0, c->stream>>>(N, f, v_r_dev);
  }
  std::vector<double> hp(3 * nb, 0.0);
  hipMemcpyAsync(hp.data(), part, 3 * nb * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  hipStreamSynchronize(c->stream);
  for (void* p : {(void*)part, (void*)vg, (void*)f}) hipFree(p);
  if (r) return r;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return vfail(c, QBH_ERR_DEVICE, hipGetErrorString(e));
  if (energies) {
    double exc = 0.0, eps = 0.0, eh = 0.0;
    for (int i = 0; i < nb; ++i) {
      exc += hp[i];
      eps += hp[nb + i];
      eh += hp[2 * nb + i];
    }
    // Γ half-sphere sums count G and -G: 2 * sum - G=0 term (EnergyFunctional.cc:463-467)
    if (g.real) {
      std::vector<double> v0(2, 0.0), r0(2, 0.0);
      if (vion_g_dev) hipMemcpy(v0.data(), vion_g_dev, 2 * sizeof(double), hipMemcpyDeviceToHost);
      hipMemcpy(r0.data(), rhog_dev, 2 * sizeof(double), hipMemcpyDeviceToHost);
      eps = 2.0 * eps - (r0[0] / omega * v0[0] + r0[1] / omega * v0[1]);
    }
    energies[0] = eps * omega;                                   // eps
    energies[1] = (g.real ? 1.0 : 0.5) * omega * 4.0 * M_PI * eh;  // ehart
    energies[2] = exc * omega / (double)N;                       // exc
  }
  return QBH_OK;
}
