// Dense band-space operations in FP64 on gfx950: overlap matrix + Cholesky
// orthonormalisation (SlaterDet::gram, src/qball/SlaterDet.cc:1043-1143) and subspace
// diagonalisation (Wavefunction::diag, src/qball/Wavefunction.cc:1510-1715).
//
// The band-space products S = C^H X (nst x nst, contraction over ~1e5 plane waves) run as
// K-split strided-batched rocBLAS GEMMs whose slices are summed by one kernel that also
// applies the Γ weights (2 Re sum over the half sphere, G = 0 counted once); the
// factorisations use rocSOLVER (potrf, syevd/heevd).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/qbh.h"
#include "qbh_ham.h"

namespace qbh {
Basis* basis_of(qbh_basis b);
Ctx* ctx_of_basis(qbh_basis b);
}  // namespace qbh

using namespace qbh;

static int dfail(Ctx* c, int code, const std::string& msg) {
  c->last_error = msg;
  return code;
}

// S = alpha * sum_k P[k] - (real ? a0 b0^T : 0), P slices of M*N (complex: 2*M*N doubles)
__global__ void k_sum_band_splits(int M, int N, int nsplit, int cplx, const double* __restrict__ P,
                                  double alpha, const double* __restrict__ A, long lda,
                                  const double* __restrict__ B, long ldb, int real,
                                  double* __restrict__ S) {
  const long e = cplx ? 2 : 1;
  const long total = (long)M * N * e;
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i >= total) return;
  double s = P[i];
  for (int k = 1; k < nsplit; ++k) s += P[(long)k * total + i];
  s = alpha * s;
  if (real) {
    const long m = i % M, n = i / M;
    s -= A[m * lda] * B[n * ldb];  // rank-1 G = 0 correction (ger -1.0 on row 0)
  }
  S[i] = s;
}

// S (M x N, ld M) = alpha * A^T B (real view, K rows) or A^H B (complex, K rows)
int qbh::band_product(Ctx* c, void* bhv, bool real, int M, int N, long K, const double* A,
                      long lda, const double* B, long ldb, double alpha, double* S,
                      double*& scratch, size_t& scratch_bytes) {
  rocblas_handle bh = static_cast<rocblas_handle>(bhv);
  const long tiles = (long)((M + 127) / 128) * ((N + 127) / 128);
  long ns = std::min<long>((4096 + tiles - 1) / tiles, 32);
  ns = std::max<long>(1, std::min<long>(ns, K / 1024));
  const long kc = (K + ns - 1) / ns;
  const long nfull = K / kc, rem = K - nfull * kc;
  const int nparts = (int)(nfull + (rem ? 1 : 0));
  const size_t e = real ? 1 : 2;
  const size_t need = sizeof(double) * e * (size_t)M * N * nparts;
  if (scratch_bytes < need) {
    if (scratch) hipFree(scratch);
    scratch = nullptr;
    scratch_bytes = 0;
    if (hipMalloc(&scratch, need) != hipSuccess) return dfail(c, QBH_ERR_ALLOC, "band product scratch");
    scratch_bytes = need;
  }
  rocblas_status st = rocblas_status_success;
  const long pst = (long)M * N;  // partial stride in BLAS elements
  if (real) {
    const double one = 1.0, zero = 0.0;
    if (nfull > 0)
      st = rocblas_dgemm_strided_batched(bh, rocblas_operation_transpose, rocblas_operation_none, M,
                                         N, (int)kc, &one, A, lda, kc, B, ldb, kc, &zero, scratch,
                                         M, pst, (int)nfull);
    if (st == rocblas_status_success && rem)
      st = rocblas_dgemm(bh, rocblas_operation_transpose, rocblas_operation_none, M, N, (int)rem,
                         &one, A + nfull * kc, lda, B + nfull * kc, ldb, &zero,
                         scratch + nfull * pst, M);
  } else {
    const rocblas_double_complex one(1.0, 0.0), zero(0.0, 0.0);
    auto* Az = reinterpret_cast<const rocblas_double_complex*>(A);
    auto* Bz = reinterpret_cast<const rocblas_double_complex*>(B);
    auto* Pz = reinterpret_cast<rocblas_double_complex*>(scratch);
    if (nfull > 0)
      st = rocblas_zgemm_strided_batched(bh, rocblas_operation_conjugate_transpose,
                                         rocblas_operation_none, M, N, (int)kc, &one, Az, lda, kc,
                                         Bz, ldb, kc, &zero, Pz, M, pst, (int)nfull);
    if (st == rocblas_status_success && rem)
      st = rocblas_zgemm(bh, rocblas_operation_conjugate_transpose, rocblas_operation_none, M, N,
                         (int)rem, &one, Az + nfull * kc, lda, Bz + nfull * kc, ldb, &zero,
                         Pz + nfull * pst, M);
  }
  if (st != rocblas_status_success)
    return dfail(c, QBH_ERR_BLAS, std::string("band product: ") + rocblas_status_to_string(st));
  const long total = (long)M * N * e;
  k_sum_band_splits<<<(unsigned)((total + 255) / 256), 256, 0, c->stream>>>(
      M, N, nparts, real ? 0 : 1, scratch, alpha, A, lda, B, ldb, real ? 1 : 0, S);
  return 0;
}

namespace {
struct Scratch {
  double* p = nullptr;
  size_t bytes = 0;
  ~Scratch() {
    if (p) hipFree(p);
  }
};
}  // namespace

extern "C" int qbh_gram(qbh_basis basis, double* psi_dev, long ld, int nst) {
  if (!basis || !psi_dev || nst <= 0) return QBH_ERR_ARG;
  Basis* b = basis_of(basis);
  Ctx* c = ctx_of_basis(basis);
  hipSetDevice(c->device);
  const int ngw = b->g.ngw;
  if (ld < ngw) return dfail(c, QBH_ERR_ARG, "ld < ngw");
  void* hb;
  if (int r = blas_handle(c, &hb)) return r;
  rocblas_handle bh = static_cast<rocblas_handle>(hb);
  const bool real = b->g.real;
  double* S = nullptr;
  rocblas_int* info = nullptr;
  const size_t ssz = (size_t)nst * nst * (real ? 1 : 2);
  if (hipMalloc(&S, ssz * sizeof(double)) != hipSuccess ||
      hipMalloc(&info, sizeof(rocblas_int)) != hipSuccess) {
    if (S) hipFree(S);
    return dfail(c, QBH_ERR_ALLOC, "gram scratch allocation failed");
  }
  Scratch scr;
  // S = 2 C^T C - c0 c0^T (SlaterDet.cc:1065-1066) or C^H C
  int r = band_product(c, bh, real, nst, nst, real ? 2L * ngw : (long)ngw, psi_dev,
                       real ? 2 * ld : ld, psi_dev, real ? 2 * ld : ld, real ? 2.0 : 1.0, S,
                       scr.p, scr.bytes);
  rocblas_status st = rocblas_status_success;
  if (!r) {
    if (real) {
      const double one = 1.0;
      st = rocsolver_dpotrf(bh, rocblas_fill_lower, nst, S, nst, info);
      if (st == rocblas_status_success)
        st = rocblas_dtrsm(bh, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                           rocblas_diagonal_non_unit, 2 * ngw, nst, &one, S, nst, psi_dev, 2 * ld);
    } else {
      const rocblas_double_complex zone(1.0, 0.0);
      auto* C = reinterpret_cast<rocblas_double_complex*>(psi_dev);
      auto* Sz = reinterpret_cast<rocblas_double_complex*>(S);
      st = rocsolver_zpotrf(bh, rocblas_fill_lower, nst, Sz, nst, info);
      if (st == rocblas_status_success)
        st = rocblas_ztrsm(bh, rocblas_side_right, rocblas_fill_lower,
                           rocblas_operation_conjugate_transpose, rocblas_diagonal_non_unit, ngw,
                           nst, &zone, Sz, nst, C, ld);
    }
  }
  rocblas_int hinfo = 0;
  hipMemcpyAsync(&hinfo, info, sizeof(rocblas_int), hipMemcpyDeviceToHost, c->stream);
  hipStreamSynchronize(c->stream);
  hipFree(S);
  hipFree(info);
  if (r) return r;
  if (st != rocblas_status_success)
    return dfail(c, QBH_ERR_BLAS, std::string("gram: ") + rocblas_status_to_string(st));
  if (hinfo != 0) return dfail(c, QBH_ERR_ARG, "gram: overlap matrix not positive definite");
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return dfail(c, QBH_ERR_DEVICE, hipGetErrorString(e));
  return QBH_OK;
}

extern "C" int qbh_subspace_diag(qbh_basis basis, double* psi_dev, long ld,
                                 const double* hpsi_dev, long ldh, int nst, double* eig,
                                 int eigvec) {
  if (!basis || !psi_dev || !hpsi_dev || nst <= 0 || !eig) return QBH_ERR_ARG;
  Basis* b = basis_of(basis);
  Ctx* c = ctx_of_basis(basis);
  hipSetDevice(c->device);
  const int ngw = b->g.ngw;
  if (ld < ngw || ldh < ngw) return dfail(c, QBH_ERR_ARG, "leading dimension < ngw");
  void* hb;
  if (int r = blas_handle(c, &hb)) return r;
  rocblas_handle bh = static_cast<rocblas_handle>(hb);
  const bool real = b->g.real;
  const size_t e = real ? 1 : 2;
  double *H = nullptr, *W = nullptr, *E = nullptr, *C0 = nullptr;
  rocblas_int* info = nullptr;
  const size_t csz = sizeof(double) * 2 * (size_t)ngw * nst;
  bool ok = hipMalloc(&H, sizeof(double) * e * nst * nst) == hipSuccess &&
            hipMalloc(&W, sizeof(double) * nst) == hipSuccess &&
            hipMalloc(&E, sizeof(double) * nst) == hipSuccess &&
            hipMalloc(&info, sizeof(rocblas_int)) == hipSuccess &&
            (!eigvec || hipMalloc(&C0, csz) == hipSuccess);
  auto release = [&]() {
    for (void* p : {(void*)H, (void*)W, (void*)E, (void*)C0, (void*)info})
      if (p) hipFree(p);
  };
  if (!ok) {
    release();
    return dfail(c, QBH_ERR_ALLOC, "subspace diag scratch allocation failed");
  }
  Scratch scr;
  // H = 2 C^T (HC) - c0 (hc)0^T (Wavefunction.cc:1546-1548) or C^H (HC) (:1634)
  int r = band_product(c, bh, real, nst, nst, real ? 2L * ngw : (long)ngw, psi_dev,
                       real ? 2 * ld : ld, hpsi_dev, real ? 2 * ldh : ldh, real ? 2.0 : 1.0, H,
                       scr.p, scr.bytes);
  rocblas_status st = rocblas_status_success;
  const rocblas_evect ev = eigvec ? rocblas_evect_original : rocblas_evect_none;
  if (!r) {
    if (real)
      st = rocsolver_dsyevd(bh, ev, rocblas_fill_lower, nst, H, nst, W, E, info);
    else
      st = rocsolver_zheevd(bh, ev, rocblas_fill_lower, nst,
                            reinterpret_cast<rocblas_double_complex*>(H), nst, W, E, info);
  }
  if (!r && st == rocblas_status_success && eigvec) {
    // C <- C Z (Wavefunction.cc:1605-1607)
    hipMemcpy2DAsync(C0, 2 * sizeof(double) * ngw, psi_dev, 2 * sizeof(double) * ld,
                     2 * sizeof(double) * ngw, nst, hipMemcpyDeviceToDevice, c->stream);
    if (real) {
      const double one = 1.0, zero = 0.0;
      st = rocblas_dgemm(bh, rocblas_operation_none, rocblas_operation_none, 2 * ngw, nst, nst,
                         &one, C0, 2 * ngw, H, nst, &zero, psi_dev, 2 * ld);
    } else {
      const rocblas_double_complex one(1.0, 0.0), zero(0.0, 0.0);
      st = rocblas_zgemm(bh, rocblas_operation_none, rocblas_operation_none, ngw, nst, nst, &one,
                         reinterpret_cast<rocblas_double_complex*>(C0), ngw,
                         reinterpret_cast<rocblas_double_complex*>(H), nst, &zero,
                         reinterpret_cast<rocblas_double_complex*>(psi_dev), ld);
    }
  }
  rocblas_int hinfo = 0;
  hipMemcpyAsync(&hinfo, info, sizeof(rocblas_int), hipMemcpyDeviceToHost, c->stream);
  hipMemcpyAsync(eig, W, sizeof(double) * nst, hipMemcpyDeviceToHost, c->stream);
  hipStreamSynchronize(c->stream);
  release();
  if (r) return r;
  if (st != rocblas_status_success)
    return dfail(c, QBH_ERR_BLAS, std::string("subspace diag: ") + rocblas_status_to_string(st));
  if (hinfo != 0) return dfail(c, QBH_ERR_ARG, "subspace diag: eigensolver did not converge");
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return dfail(c, QBH_ERR_DEVICE, hipGetErrorString(err));
  return QBH_OK;
}
