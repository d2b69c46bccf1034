// TDDFT observables: the electronic current (CurrentDensity::update_current,
// src/qball/CurrentDensity.cc:39-106) with the two-state density of
// SlaterDet::compute_density (SlaterDet.cc:935-966).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/qbh.h"
#include "qbh_internal.h"

namespace qbh {
Basis* basis_of(qbh_basis b);
Ctx* ctx_of_basis(qbh_basis b);
}  // namespace qbh

using namespace qbh;

// out[d][n][ig] = i (k+G)_d c[n][ig] for d = 0..2 (CurrentDensity.cc:70-73)
__global__ void k_ikc(int ngw, int nb, const double* __restrict__ kpgx,
                      const double* __restrict__ c, long ld, double* __restrict__ out) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i >= (long)ngw * nb) return;
  const long n = i / ngw, ig = i - n * ngw;
  const double cr = c[2 * (n * ld + ig)], ci = c[2 * (n * ld + ig) + 1];
  for (int d = 0; d < 3; ++d) {
    const double k = kpgx[(long)d * ngw + ig];
    double* o = out + 2 * ((long)d * ngw * nb + n * ngw + ig);
    o[0] = -k * ci;
    o[1] = k * cr;
  }
}

// j[d][r] -= sum_n fac_n Im(conj(f0_n(r)) fd_n(r)) over the nb bands of a chunk
__global__ void k_current_acc(long N, int nb, const double* __restrict__ fac,
                              const double* __restrict__ f0, const double* __restrict__ fd,
                              double* __restrict__ j) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i >= N) return;
  double a[3] = {0.0, 0.0, 0.0};
  for (int n = 0; n < nb; ++n) {
    const double w = fac[n];
    if (w <= 0.0) continue;
    const double ar = f0[2 * (n * N + i)], ai = f0[2 * (n * N + i) + 1];
    for (int d = 0; d < 3; ++d) {
      const double* p = fd + 2 * (((long)d * nb + n) * N + i);
      a[d] -= w * (ar * p[1] - ai * p[0]);
    }
  }
  for (int d = 0; d < 3; ++d) j[d * N + i] += a[d];
}

__global__ void k_fill0d(double* __restrict__ p, long n) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// per-block sums of the three components: part[d * nblk + block]
__global__ void __launch_bounds__(256)
k_sum3(long N, const double* __restrict__ j, double* __restrict__ part) {
  __shared__ double red[256];
  for (int d = 0; d < 3; ++d) {
    double s = 0.0;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < N; i += (long)gridDim.x * 256)
      s += j[d * N + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[d * gridDim.x + blockIdx.x] = red[0];
    __syncthreads();
  }
}

static int ofail(Ctx* c, int code, const std::string& m) {
  c->last_error = m;
  return code;
}

extern "C" int qbh_current(qbh_basis basis, const double* psi_dev, long ld, int nst,
                           const double* occ, double weight, double omega, const double* A,
                           double nelectrons, double* j_dev, double* total) {
  if (!basis || !psi_dev || !occ || nst <= 0 || !(omega > 0.0) || (!j_dev && !total))
    return QBH_ERR_ARG;
  Basis* b = basis_of(basis);
  Ctx* c = ctx_of_basis(basis);
  hipSetDevice(c->device);
  const GridDev& g = b->g;
  if (ld < g.ngw) return ofail(c, QBH_ERR_ARG, "ld < ngw");
  const long N = (long)g.np0 * g.np1 * g.np2;
  const int ngw = g.ngw;
  // bands per chunk: four complex grids per band, about 1 GiB of grids per chunk
  const int nb = (int)std::max(1L, std::min((long)nst, (1L << 30) / (64L * N)));
  const int nblk = 256;
  std::vector<double> fac(nst);
  for (int n = 0; n < nst; ++n) {
    const double f = weight / omega * occ[n];
    fac[n] = f > 0.0 ? f : 0.0;
  }
  const size_t bytes = sizeof(double) * (2 * 4 * (size_t)nb * N + 2 * 3 * (size_t)nb * ngw +
                                         (j_dev ? 0 : 3 * (size_t)N) + nst + 3 * nblk);
  double* buf = nullptr;
  if (hipMalloc(&buf, bytes) != hipSuccess)
    return ofail(c, QBH_ERR_ALLOC, "current scratch allocation failed");
  double* f0 = buf;
  double* fd = f0 + 2 * (size_t)nb * N;
  double* ikc = fd + 2 * 3 * (size_t)nb * N;
  double* jj = ikc + 2 * 3 * (size_t)nb * ngw;
  double* dfac = jj + (j_dev ? 0 : 3 * (size_t)N);
  double* part = dfac + nst;
  if (j_dev) jj = j_dev;
  hipMemcpyAsync(dfac, fac.data(), nst * sizeof(double), hipMemcpyHostToDevice, c->stream);
  k_fill0d<<<(unsigned)((3 * N + 255) / 256), 256, 0, c->stream>>>(jj, 3 * N);
  int r = QBH_OK;
  for (int n0 = 0; n0 < nst && r == QBH_OK; n0 += nb) {
    const int m = std::min(nb, nst - n0);
    if (std::all_of(fac.begin() + n0, fac.begin() + n0 + m, [](double f) { return f <= 0.0; }))
      continue;
    const double* cp = psi_dev + 2 * (size_t)ld * n0;
    k_ikc<<<(unsigned)(((long)ngw * m + 255) / 256), 256, 0, c->stream>>>(ngw, m, b->kpgx, cp, ld,
                                                                          ikc);
    r = qbh_fft_backward(basis, cp, ld, m, 0, f0, 0);
    for (int d = 0; d < 3 && r == QBH_OK; ++d)
      r = qbh_fft_backward(basis, ikc + 2 * (size_t)d * m * ngw, ngw, m, 0,
                           fd + 2 * (size_t)d * m * N, 0);
    if (r == QBH_OK)
      k_current_acc<<<(unsigned)((N + 255) / 256), 256, 0, c->stream>>>(N, m, dfac + n0, f0, fd,
                                                                        jj);
  }
  std::vector<double> hp(3 * nblk, 0.0);
  if (r == QBH_OK && total) {
    k_sum3<<<nblk, 256, 0, c->stream>>>(N, jj, part);
    hipMemcpyAsync(hp.data(), part, 3 * nblk * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  }
  hipStreamSynchronize(c->stream);
  hipFree(buf);
  if (r) return r;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return ofail(c, QBH_ERR_DEVICE, std::string("current: ") + hipGetErrorString(e));
  if (total) {
    for (int d = 0; d < 3; ++d) {
      double s = 0.0;
      for (int i = 0; i < nblk; ++i) s += hp[d * nblk + i];
      total[d] = omega / (double)N * s + (A ? A[d] * nelectrons : 0.0);
    }
  }
  return QBH_OK;
}
