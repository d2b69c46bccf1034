// C-ABI entry points: transforms, Hamiltonian state, Hψ, density, Gram, exponential.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/qbh.h"
#include "qbh_ham.h"

using namespace qbh;

namespace qbh {
Basis* basis_of(qbh_basis b);
Ctx* ctx_of_basis(qbh_basis b);
Ctx* ctx_of(qbh_ctx c);
}  // namespace qbh

static int fail(Ctx* c, int code, const std::string& msg) {
  c->last_error = msg;
  return code;
}

static int dev_check(Ctx* c, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, QBH_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

__global__ void k_fill0(double* __restrict__ p, long n) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

static void fill0(Ctx* c, double* p, long n) {
  if (n <= 0) return;
  k_fill0<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(p, n);
}

// kinetic-only term when the local potential is off: h += kin * c
__global__ void k_kinetic(int ngw, int nst, const double* __restrict__ kin,
                          const double* __restrict__ psi, long ld, double* __restrict__ h,
                          long ldh) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i >= (long)ngw * nst) return;
  const long n = i / ngw, ig = i - n * ngw;
  const double k = kin[ig];
  const double* c = psi + 2 * (n * ld + ig);
  double* o = h + 2 * (n * ldh + ig);
  o[0] += k * c[0];
  o[1] += k * c[1];
}

// ------------------------------------------------------------------------------------
// chunked transform driver: calls body(args) for consecutive chunks of slots, with the
// transform workspace `ws` of `cap` bytes (at least one slot), or with the context's
// workspace (grown on demand up to its cap) when ws is null
template <typename F>
static int for_chunks(Ctx* c, Basis* b, int nslots_total, int prec, F&& body,
                      void* ws = nullptr, size_t cap = 0) {
  if (nslots_total <= 0) return 0;
  const size_t per = slot_bytes(b, prec);
  size_t S = std::max<size_t>(1, (ws ? cap : c->ws_cap) / per);
  S = std::min<size_t>(S, (size_t)nslots_total);
  S = std::min<size_t>(S, 65535);
  if (!ws) ws = ws_get(c, S * per);
  if (!ws) return fail(c, QBH_ERR_ALLOC, "transform workspace allocation failed");
  const size_t csz = prec == PREC_FP64 ? 16 : 8;
  char* t1 = static_cast<char*>(ws);
  char* t2 = t1 + S * csz * (size_t)b->g.ncol * b->g.np2;
  for (int s0 = 0; s0 < nslots_total; s0 += (int)S) {
    PipelineArgs a{};
    a.nslots = std::min<int>((int)S, nslots_total - s0);
    a.prec = prec;
    a.t1 = t1;
    a.t2 = t2;
    if (int r = body(a, s0)) return r;
  }
  return 0;
}

// ------------------------------------------------------------------------------------
extern "C" int qbh_fft_backward(qbh_basis basis, const double* c_dev, long ld, int nbands,
                                int pair, double* f_dev, int flags) {
  if (!basis || !c_dev || !f_dev || nbands <= 0) return QBH_ERR_ARG;
  Basis* b = basis_of(basis);
  Ctx* c = ctx_of_basis(basis);
  hipSetDevice(c->device);
  if (ld < b->g.ngw) return fail(c, QBH_ERR_ARG, "ld < ngw");
  if (pair && !b->g.real) return fail(c, QBH_ERR_ARG, "pair transforms need a real basis");
  const int prec = (flags & QBH_FP32) ? PREC_FP32 : PREC_FP64;
  const int mode = pair ? SLOT_PAIR : (b->g.real ? SLOT_SINGLE_REAL : SLOT_COMPLEX);
  const int nslots = pair ? (nbands + 1) / 2 : nbands;
  const long N = (long)b->g.np0 * b->g.np1 * b->g.np2;
  return for_chunks(c, b, nslots, prec, [&](PipelineArgs& a, int s0) {
    a.psi = c_dev;
    a.ld = ld;
    a.nst = nbands;
    a.band0 = pair ? 2 * s0 : s0;
    a.mode = mode;
    if (int r = pipeline_bwd_yz(c, b, a)) return r;
    return pipeline_x_store(c, b, a, f_dev + 2 * (long)s0 * N);
  });
}

// ------------------------------------------------------------------------------------
// Allocation-free backward transform.  Its transform workspace (the z-pencil and y-line
// tiles of each slot) is this file-scope array instead of memory the context allocates:
// each GPU the library runs on holds one copy from module load until the process ends.
// Host code never uses the array's name as a device address: qbh_ctx_create looks up this
// device's address once (hipGetSymbolAddress, with the device current) and every call
// passes that address.  Every call on one GPU shares the array and leaves its contents
// behind; the transform passes write every workspace element they read, so nothing has to
// start cleared.  Calls that overlap on one GPU (two streams, two threads) race.
constexpr size_t kNoallocScratchBytes = size_t(256) << 20;  // 256 MiB
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_addr;  // device -> device address of scratch
}  // namespace

void qbh::noalloc_scratch_setup(Ctx* c) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_addr.find(c->device);
  if (it == g_scratch_addr.end()) {
    // the caller (qbh_ctx_create) has made c->device current
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      (void)hipGetLastError();
      c->scratch = nullptr;
      c->scratch_err = std::string("hipGetSymbolAddress of the no-allocation scratch failed: ") +
                       hipGetErrorString(e);
      return;  // not cached: a later context on this device tries again
    }
    it = g_scratch_addr.emplace(c->device, ptr).first;
  }
  c->scratch = it->second;
}

extern "C" int qbh_fft_backward_noalloc(qbh_basis basis, const double* c_dev, long ld,
                                        int nbands, int pair, double* f_dev, int flags) {
  if (!basis || !c_dev || !f_dev || nbands <= 0) return QBH_ERR_ARG;
  Basis* b = basis_of(basis);
  Ctx* c = ctx_of_basis(basis);
  hipSetDevice(c->device);
  if (ld < b->g.ngw) return fail(c, QBH_ERR_ARG, "ld < ngw");
  if (pair && !b->g.real) return fail(c, QBH_ERR_ARG, "pair transforms need a real basis");
  if (!c->scratch) return fail(c, QBH_ERR_DEVICE, c->scratch_err);
  const int prec = (flags & QBH_FP32) ? PREC_FP32 : PREC_FP64;
  const size_t per = slot_bytes(b, prec);
  if (per > kNoallocScratchBytes)
    return fail(c, QBH_ERR_ARG,
                "one transform slot of this grid needs " + std::to_string(per) +
                    " bytes, more than the " + std::to_string(kNoallocScratchBytes) +
                    "-byte no-allocation scratch");
  const int mode = pair ? SLOT_PAIR : (b->g.real ? SLOT_SINGLE_REAL : SLOT_COMPLEX);
  const int nslots = pair ? (nbands + 1) / 2 : nbands;
  const long N = (long)b->g.np0 * b->g.np1 * b->g.np2;
  return for_chunks(
      c, b, nslots, prec,
      [&](PipelineArgs& a, int s0) {
        a.psi = c_dev;
        a.ld = ld;
        a.nst = nbands;
        a.band0 = pair ? 2 * s0 : s0;
        a.mode = mode;
        if (int r = pipeline_bwd_yz(c, b, a)) return r;
        return pipeline_x_store(c, b, a, f_dev + 2 * (long)s0 * N);
      },
      c->scratch, kNoallocScratchBytes);
}

extern "C" int qbh_fft_forward(qbh_basis basis, const double* f_dev, int nbands, int pair,
                               double* c_dev, long ld, int flags) {
  if (!basis || !c_dev || !f_dev || nbands <= 0) return QBH_ERR_ARG;
  Basis* b = basis_of(basis);
  Ctx* c = ctx_of_basis(basis);
  hipSetDevice(c->device);
  if (ld < b->g.ngw) return fail(c, QBH_ERR_ARG, "ld < ngw");
  if (pair && !b->g.real) return fail(c, QBH_ERR_ARG, "pair transforms need a real basis");
  const int prec = (flags & QBH_FP32) ? PREC_FP32 : PREC_FP64;
  const int mode = pair ? SLOT_PAIR : (b->g.real ? SLOT_SINGLE_REAL : SLOT_COMPLEX);
  const int nslots = pair ? (nbands + 1) / 2 : nbands;
  const long N = (long)b->g.np0 * b->g.np1 * b->g.np2;
  return for_chunks(c, b, nslots, prec, [&](PipelineArgs& a, int s0) {
    a.psi = nullptr;
    a.ld = ld;
    a.nst = nbands;
    a.band0 = pair ? 2 * s0 : s0;
    a.mode = mode;
    if (int r = pipeline_x_load(c, b, a, f_dev + 2 * (long)s0 * N, 0, 1.0)) return r;
    return pipeline_fwd_zy_accumulate(c, b, a, c_dev, ld, nullptr, 1);
  });
}

extern "C" int qbh_fft_forward_real(qbh_basis basis, const double* rho_dev, double scale,
                                    double* c_dev, int flags) {
  if (!basis || !rho_dev || !c_dev) return QBH_ERR_ARG;
  Basis* b = basis_of(basis);
  Ctx* c = ctx_of_basis(basis);
  hipSetDevice(c->device);
  const int prec = (flags & QBH_FP32) ? PREC_FP32 : PREC_FP64;
  const int mode = b->g.real ? SLOT_SINGLE_REAL : SLOT_COMPLEX;
  return for_chunks(c, b, 1, prec, [&](PipelineArgs& a, int) {
    a.psi = nullptr;
    a.ld = b->g.ngw;
    a.nst = 1;
    a.band0 = 0;
    a.mode = mode;
    if (int r = pipeline_x_load(c, b, a, rho_dev, 1, scale)) return r;
    return pipeline_fwd_zy_accumulate(c, b, a, c_dev, b->g.ngw, nullptr, 1);
  });
}

// ------------------------------------------------------------------------------------
static int upload_kin(Ham* h) {
  Basis* b = h->b;
  const int ngw = b->g.ngw;
  std::vector<double> kpg2(ngw), kx(3 * (size_t)ngw);
  hipMemcpy(kx.data(), b->kpgx, kx.size() * sizeof(double), hipMemcpyDeviceToHost);
  hipMemcpy(kpg2.data(), b->kpg2, kpg2.size() * sizeof(double), hipMemcpyDeviceToHost);
  const double* A = h->A;
  const bool zero = A[0] == 0.0 && A[1] == 0.0 && A[2] == 0.0;
  const double value2 = A[0] * A[0] + A[1] * A[1] + A[2] * A[2];
  std::vector<double> kin(ngw);
  for (int ig = 0; ig < ngw; ++ig) {
    double k2 = kpg2[ig];
    if (!zero) {  // VectorPotential::get_kpgpa2 (VectorPotential.h:84-93)
      k2 = kpg2[ig] + value2;
      k2 -= 2 * A[0] * kx[ig];
      k2 -= 2 * A[1] * kx[ngw + ig];
      k2 -= 2 * A[2] * kx[2 * (size_t)ngw + ig];
    }
    kin[ig] = 0.5 * k2;
  }
  if (!h->kin && hipMalloc(&h->kin, ngw * sizeof(double)) != hipSuccess) {
    h->kin = nullptr;
    return fail(h->ctx, QBH_ERR_ALLOC, "kinetic table allocation failed");
  }
  hipMemcpy(h->kin, kin.data(), ngw * sizeof(double), hipMemcpyHostToDevice);
  return 0;
}

extern "C" int qbh_ham_create(qbh_ctx ctx, qbh_basis basis, qbh_ham* out) {
  if (!ctx || !basis || !out) return QBH_ERR_ARG;
  Ctx* c = ctx_of(ctx);
  hipSetDevice(c->device);
  auto* hh = new qbh_ham_s();
  hh->h.ctx = c;
  hh->h.b = basis_of(basis);
  hh->h.b_handle = basis;
  if (int r = upload_kin(&hh->h)) {
    delete hh;
    return r;
  }
  *out = hh;
  return QBH_OK;
}

static void free_proj(Ham* h) {
  for (void* p : {(void*)h->col_l, (void*)h->col_tw, (void*)h->col_atom, (void*)h->col_fac,
                  (void*)h->twnl, (void*)h->tau, h->anl64, h->anl32, h->anlc64, h->anlc32})
    if (p) hipFree(p);
  h->col_l = h->col_tw = h->col_atom = nullptr;
  h->col_fac = h->twnl = h->tau = nullptr;
  h->anl64 = h->anl32 = h->anlc64 = h->anlc32 = nullptr;
  h->anl64_ok = h->anl32_ok = h->anlc64_ok = h->anlc32_ok = false;
  h->P = 0;
  h->col_atom_host.clear();
}

extern "C" int qbh_ham_destroy(qbh_ham ham) {
  if (!ham) return QBH_ERR_ARG;
  Ham* h = &ham->h;
  hipSetDevice(h->ctx->device);
  hipStreamSynchronize(h->ctx->stream);
  free_proj(h);
  for (void* p : {(void*)h->v, (void*)h->kin, h->f64, h->f32, h->psi32, h->tmp32, (void*)h->red,
                  h->etrs, h->dpsi, h->df, h->fpart, h->fg})
    if (p) hipFree(p);
  for (void* p : h->expo) hipFree(p);
  delete ham;
  return QBH_OK;
}

extern "C" int qbh_ham_set_vloc(qbh_ham ham, const double* v_dev) {
  if (!ham || !v_dev) return QBH_ERR_ARG;
  Ham* h = &ham->h;
  hipSetDevice(h->ctx->device);
  const long N = (long)h->b->g.np0 * h->b->g.np1 * h->b->g.np2;
  if (!h->v && hipMalloc(&h->v, N * sizeof(double)) != hipSuccess) {
    h->v = nullptr;
    return fail(h->ctx, QBH_ERR_ALLOC, "v_r allocation failed");
  }
  hipMemcpyAsync(h->v, v_dev, N * sizeof(double), hipMemcpyDeviceToDevice, h->ctx->stream);
  return dev_check(h->ctx, "set_vloc");
}

// projector columns and tables; twnl null: the table is allocated, not filled
static int set_proj(Ham* h, int nsp, const int* na, const int* npr, const int* lproj,
                    const double* wt, const double* twnl, const double* tau, double omega) {
  Ctx* c = h->ctx;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  free_proj(h);
  h->sps.clear();
  const int ngw = h->b->g.ngw;
  h->nsp = nsp;
  h->na.assign(na, na + nsp);
  h->npr.assign(npr, npr + nsp);
  h->omega = omega;
  std::vector<int> cl, ctw, cat;
  std::vector<double> cf;
  int tw_row0 = 0, atom0 = 0, ipr0 = 0;
  long ntw = 0, ntau = 0;
  for (int s = 0; s < nsp; ++s) {
    if (na[s] < 0 || npr[s] < 0) return fail(c, QBH_ERR_ARG, "negative na/npr");
    for (int ipr = 0; ipr < npr[s]; ++ipr) {
      const int l = lproj[ipr0 + ipr];
      if (l < 0 || l > 3) return fail(c, QBH_ERR_ARG, "lproj must be in 0..3");
      // column p = ia + ipr*na (NonLocalPotential.cc:1997 ipra = ia + ipr*na_block)
      for (int ia = 0; ia < na[s]; ++ia) {
        cl.push_back(l);
        ctw.push_back(tw_row0 + ipr);
        cat.push_back(atom0 + ia);
        cf.push_back(wt[ipr0 + ipr] * (1.0 / omega));
      }
    }
    tw_row0 += npr[s];
    atom0 += na[s];
    ipr0 += npr[s];
    ntw += (long)npr[s] * ngw;
    ntau += 3L * na[s];
  }
  h->P = (int)cl.size();
  if (h->P == 0) return QBH_OK;
  auto up = [&](auto*& dst, const auto* src, size_t n) {
    using X = std::remove_pointer_t<std::remove_reference_t<decltype(dst)>>;
    void* d = nullptr;
    if (hipMalloc(&d, std::max<size_t>(1, n) * sizeof(X)) != hipSuccess) return false;
    if (src) hipMemcpy(d, src, n * sizeof(X), hipMemcpyHostToDevice);
    dst = static_cast<X*>(d);
    return true;
  };
  if (!up(h->col_l, cl.data(), cl.size()) || !up(h->col_tw, ctw.data(), ctw.size()) ||
      !up(h->col_atom, cat.data(), cat.size()) || !up(h->col_fac, cf.data(), cf.size()) ||
      !up(h->twnl, twnl, (size_t)ntw) || !up(h->tau, tau, (size_t)ntau)) {
    free_proj(h);
    return fail(c, QBH_ERR_ALLOC, "projector table allocation failed");
  }
  h->col_atom_host = cat;
  return dev_check(c, "set_projectors");
}

extern "C" int qbh_ham_set_projectors(qbh_ham ham, int nsp, const int* na, const int* npr,
                                      const int* lproj, const double* wt, const double* twnl,
                                      const double* tau, double omega) {
  if (!ham || nsp < 0 || (nsp > 0 && (!na || !npr || !lproj || !wt || !twnl || !tau)) ||
      !(omega > 0.0))
    return QBH_ERR_ARG;
  return set_proj(&ham->h, nsp, na, npr, lproj, wt, twnl, tau, omega);
}

int qbh::retable_twnl(Ham* h) {
  if (!h->twnl || h->P == 0) return QBH_OK;
  long row0 = 0;
  for (qbh_species sp : h->sps) {
    if (int r = species_table(sp, h->ctx, h->b, h->A, 0, h->twnl + row0 * h->b->g.ngw, 0))
      return r;
    row0 += species_nproj(sp);
  }
  h->anl64_ok = h->anl32_ok = false;
  return QBH_OK;
}

extern "C" int qbh_ham_set_species(qbh_ham ham, int nsp, const qbh_species* sps, const int* na,
                                   const double* tau, double omega) {
  if (!ham || nsp < 0 || (nsp > 0 && (!sps || !na || !tau)) || !(omega > 0.0))
    return QBH_ERR_ARG;
  Ham* h = &ham->h;
  std::vector<int> npr(nsp), lproj;
  std::vector<double> wt;
  for (int s = 0; s < nsp; ++s) {
    if (!sps[s]) return QBH_ERR_ARG;
    int n = 0;
    qbh_species_info(sps[s], nullptr, nullptr, nullptr, &n, nullptr, nullptr);
    npr[s] = n;
    const size_t o = lproj.size();
    lproj.resize(o + n);
    wt.resize(o + n);
    qbh_species_info(sps[s], nullptr, nullptr, nullptr, nullptr, lproj.data() + o, wt.data() + o);
  }
  if (int r = set_proj(h, nsp, na, npr.data(), lproj.data(), wt.data(), nullptr, tau, omega))
    return r;
  h->sps.assign(sps, sps + nsp);
  if (int r = retable_twnl(h)) return r;
  return dev_check(h->ctx, "set_species");
}

extern "C" int qbh_ham_set_vector_potential(qbh_ham ham, const double* A) {
  if (!ham || !A) return QBH_ERR_ARG;
  Ham* h = &ham->h;
  hipSetDevice(h->ctx->device);
  hipStreamSynchronize(h->ctx->stream);
  std::memcpy(h->A, A, 3 * sizeof(double));
  h->anl64_ok = h->anl32_ok = false;
  // update_twnl at |k+G-A| (EnergyFunctional::vector_potential_changed) when the tables
  // come from species
  if (int r = retable_twnl(h)) return r;
  return upload_kin(h);
}

// ------------------------------------------------------------------------------------
static int ensure_side(Ctx* c) {
  if (c->side) return 0;
  if (hipStreamCreate(&c->side) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess)
    return fail(c, QBH_ERR_DEVICE, "side stream creation failed");
  return 0;
}

int qbh::hpsi_impl(Ham* h, const double* psi, long ld, int nst, const double* occ,
                   double* hpsi, long ldh, int flags, double* enl) {
  Ctx* c = h->ctx;
  Basis* b = h->b;
  const int ngw = b->g.ngw;
  if (ld < ngw || ldh < ngw) return fail(c, QBH_ERR_ARG, "leading dimension < ngw");
  if (psi == hpsi) return fail(c, QBH_ERR_ARG, "psi and hpsi must not alias");
  // every entry point that applies H comes through here (Hψ, exponential, ETRS step, TD,
  // SCF, autotuner, batch): complex elements are read as 16-byte vectors, and the
  // conversion and kinetic kernels put the band index in gridDim.y
  if ((reinterpret_cast<uintptr_t>(psi) | reinterpret_cast<uintptr_t>(hpsi)) & 15)
    return fail(c, QBH_ERR_ARG, "psi/hpsi must be 16-byte aligned");
  if (nst <= 0 || nst > 65535) return fail(c, QBH_ERR_ARG, "nst must be in 1..65535");
  const bool fp32 = (flags & QBH_FP32) != 0;
  const int prec = fp32 ? PREC_FP32 : PREC_FP64;
  if (enl) *enl = 0.0;
  const bool do_nl = (flags & QBH_NONLOCAL) && h->P > 0;
  const bool do_loc = (flags & QBH_LOCAL) != 0;
  const bool kin = (flags & QBH_KINETIC) != 0;
  if (do_loc && !h->v) return fail(c, QBH_ERR_ARG, "local potential not set");
  // dwf.c().clear(): with no padding rows (ldh == ngw) the first term written overwrites
  // instead -- FP64: the projector product (beta 0); FP32 with the local part: the forward
  // z pass (hpsi = projector + kinetic + local terms, with the projector term deferred)
  const bool fresh = !(flags & QBH_ACCUMULATE) && ldh == ngw;
  const bool nl_overwrites = fresh && !fp32 && do_nl;
  const bool z_overwrites = fresh && fp32 && do_loc;
  if (!(flags & QBH_ACCUMULATE) && !nl_overwrites && !z_overwrites)
    fill0(c, hpsi, 2 * ldh * (long)nst);
  hipStream_t main = c->stream;
  // Order of accumulation into hpsi: non-local (EnergyFunctional.cc:1500), then kinetic,
  // then local (:1654-1695).  The projector GEMMs run on the side stream, concurrently
  // with the backward/x/forward-y transform passes; the forward z pass, which adds the
  // kinetic and local terms to hpsi, waits for them.
  hipEvent_t join = nullptr;
  const float* add32 = nullptr;
  if (do_nl) {
    const bool overlap = do_loc && (h->overlap >= 0 ? h->overlap != 0 : c->overlap);
    if (overlap) {
      if (int r = ensure_side(c)) return r;
      hipEventRecord(c->ev_fork, main);
      hipStreamWaitEvent(c->side, c->ev_fork, 0);
      c->stream = c->side;
    }
    double dummy = 0.0;
    // FP32: the forward z pass adds the projector term while it accumulates into hpsi
    const int r = nl_apply(h, psi, ld, nst, occ, hpsi, ldh, fp32 ? 1 : 0, enl ? &dummy : nullptr,
                           (fp32 && do_loc) ? &add32 : nullptr, nl_overwrites);
    if (overlap) {
      hipEventRecord(c->ev_join, c->side);
      c->stream = main;
      join = c->ev_join;
    }
    if (r) {
      if (join) hipStreamWaitEvent(main, join, 0);
      return r;
    }
  }
  if (do_loc) {
    const int mode = b->g.real ? SLOT_PAIR : SLOT_COMPLEX;
    const int nslots = b->g.real ? (nst + 1) / 2 : nst;
    int r = for_chunks(c, b, nslots, prec, [&](PipelineArgs& a, int s0) {
      a.psi = psi;
      a.ld = ld;
      a.nst = nst;
      a.band0 = b->g.real ? 2 * s0 : s0;
      a.mode = mode;
      a.wait_z = join;
      a.add32 = add32;
      if (int r = pipeline_bwd_yz(c, b, a)) return r;
      if (int r = pipeline_x_vmul(c, b, a, h->v)) return r;
      return pipeline_fwd_zy_accumulate(c, b, a, hpsi, ldh, kin ? h->kin : nullptr,
                                        z_overwrites ? 2 : 0);
    });
    if (join) hipStreamWaitEvent(main, join, 0);  // also on early exit
    if (r) return r;
  } else if (kin) {
    TimedLaunch tl(c, "kinetic");
    const long n = (long)ngw * nst;
    k_kinetic<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(ngw, nst, h->kin, psi, ld,
                                                                  hpsi, ldh);
  }
  if (int r = dev_check(c, "hpsi")) return r;
  if (enl && do_nl) *enl = nl_read_enl(h, main);
  return 0;
}

extern "C" int qbh_hpsi(qbh_ham ham, const double* psi_dev, long ld, int nst, const double* occ,
                        double* hpsi_dev, long ldh, int flags, double* enl) {
  if (!ham || !psi_dev || !hpsi_dev || nst <= 0) return QBH_ERR_ARG;
  Ham* h = &ham->h;
  hipSetDevice(h->ctx->device);
  // complex elements: 16-byte aligned arrays (vector loads in the conversion kernels)
  if ((reinterpret_cast<uintptr_t>(psi_dev) | reinterpret_cast<uintptr_t>(hpsi_dev)) & 15)
    return fail(h->ctx, QBH_ERR_ARG, "psi/hpsi must be 16-byte aligned");
  if (nst > 65535) return fail(h->ctx, QBH_ERR_ARG, "nst > 65535");
  return hpsi_impl(h, psi_dev, ld, nst, occ, hpsi_dev, ldh, flags, enl);
}

extern "C" int qbh_nl_forces(qbh_ham ham, const double* psi_dev, long ld, int nst,
                             const double* occ, double* fion, int flags, double* enl) {
  if (!ham || !psi_dev || !fion || nst <= 0) return QBH_ERR_ARG;
  Ham* h = &ham->h;
  hipSetDevice(h->ctx->device);
  if (ld < h->b->g.ngw) return fail(h->ctx, QBH_ERR_ARG, "ld < ngw");
  if (reinterpret_cast<uintptr_t>(psi_dev) & 15)
    return fail(h->ctx, QBH_ERR_ARG, "psi must be 16-byte aligned");
  if (nst > 65535) return fail(h->ctx, QBH_ERR_ARG, "nst > 65535");
  if (flags & ~QBH_FP32) return fail(h->ctx, QBH_ERR_ARG, "flags: only QBH_FP32 applies");
  if (int r = nl_forces(h, psi_dev, ld, nst, occ, (flags & QBH_FP32) != 0, fion, enl)) return r;
  return dev_check(h->ctx, "nl_forces");
}

// ------------------------------------------------------------------------------------
// density: per-slot factors (fac1, fac2, beta1 = Im c1(G=0), beta2 = Im c2(G=0))
__global__ void k_dens_facs(int nslots, int band0, int nst, int pair, const double* __restrict__ fac,
                            const double* __restrict__ psi, long ld, double* __restrict__ out) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= nslots) return;
  if (pair) {
    const int n1 = band0 + 2 * s, n2 = n1 + 1;
    out[4 * s + 0] = fac[n1];
    out[4 * s + 1] = n2 < nst ? fac[n2] : 0.0;
    out[4 * s + 2] = psi[2 * (long)n1 * ld + 1];
    out[4 * s + 3] = n2 < nst ? psi[2 * (long)n2 * ld + 1] : 0.0;
  } else {
    out[s] = fac[band0 + s];
  }
}

extern "C" int qbh_density(qbh_basis basis, const double* psi_dev, long ld, int nst,
                           const double* occ, double weight, double omega, double* rho_dev,
                           int flags) {
  if (!basis || !psi_dev || !rho_dev || !occ || nst <= 0 || !(omega > 0.0)) return QBH_ERR_ARG;
  Basis* b = basis_of(basis);
  Ctx* c = ctx_of_basis(basis);
  hipSetDevice(c->device);
  if (ld < b->g.ngw) return fail(c, QBH_ERR_ARG, "ld < ngw");
  const int prec = (flags & QBH_FP32) ? PREC_FP32 : PREC_FP64;
  const double prefac = weight / omega;
  std::vector<double> fac(nst);
  for (int n = 0; n < nst; ++n) {
    const double f = prefac * occ[n];
    fac[n] = f > 0.0 ? f : 0.0;
  }
  double* dfac = nullptr;
  const bool pair = b->g.real;
  const int nslots = pair ? (nst + 1) / 2 : nst;
  if (hipMalloc(&dfac, (nst + 4 * (size_t)nslots) * sizeof(double)) != hipSuccess)
    return fail(c, QBH_ERR_ALLOC, "density factor allocation failed");
  hipMemcpyAsync(dfac, fac.data(), nst * sizeof(double), hipMemcpyHostToDevice, c->stream);
  double* slotf = dfac + nst;
  int r = for_chunks(c, b, nslots, prec, [&](PipelineArgs& a, int s0) {
    a.psi = psi_dev;
    a.ld = ld;
    a.nst = nst;
    a.band0 = pair ? 2 * s0 : s0;
    a.mode = pair ? SLOT_PAIR : SLOT_COMPLEX;
    k_dens_facs<<<(a.nslots + 255) / 256, 256, 0, c->stream>>>(a.nslots, a.band0, nst, pair ? 1 : 0,
                                                               dfac, psi_dev, ld,
                                                               slotf + 4 * (size_t)s0);
    if (int r = pipeline_bwd_yz(c, b, a)) return r;
    return pipeline_x_density(c, b, a, slotf + 4 * (size_t)s0, rho_dev);
  });
  hipStreamSynchronize(c->stream);
  hipFree(dfac);
  if (r) return r;
  return dev_check(c, "density");
}

// ------------------------------------------------------------------------------------
// ------------------------------------------------------------------------------------
__global__ void k_zaxpy(long n, double ar, double ai, const double* __restrict__ x, long ldx,
                        double* __restrict__ y, long ldy, int ngw) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i >= n) return;
  const long col = i / ngw, ig = i - col * ngw;
  const double* xp = x + 2 * (col * ldx + ig);
  double* yp = y + 2 * (col * ldy + ig);
  const double xr = xp[0], xi = xp[1];
  yp[0] += ar * xr - ai * xi;
  yp[1] += ar * xi + ai * xr;
}

// out = sum_{N<=order} (-i dt)^N/N! H^N in (ExponentialWavefunctionStepper::exponential,
// ExponentialWavefunctionStepper.cc:51-148, num_exp = 1); out may alias in.  krylov
// (nullable): receives H^k in for k < order, block k at krylov + k*2*ngw*nst (ld = ngw).
int qbh::expo_impl(Ham* h, const double* in, long ldin, double* outp, long ldout, int nst,
                   double dt, int order, int flags, double* krylov) {
  Ctx* c = h->ctx;
  const int ngw = h->b->g.ngw;
  const size_t bytes = 2 * sizeof(double) * (size_t)ngw * nst;
  const size_t blk = 2 * (size_t)ngw * nst;
  if (h->expo.size() != 3 || h->expo_bytes < bytes) {
    for (void* p : h->expo) hipFree(p);
    h->expo.assign(3, nullptr);
    h->expo_bytes = 0;
    for (auto& p : h->expo)
      if (hipMalloc(&p, bytes) != hipSuccess) {
        p = nullptr;
        return fail(c, QBH_ERR_ALLOC, "exponential buffers allocation failed");
      }
    h->expo_bytes = bytes;
  }
  double* expwf = static_cast<double*>(h->expo[0]);
  double* bufs[2] = {static_cast<double*>(h->expo[1]), static_cast<double*>(h->expo[2])};
  // expwf = in (compact, ld = ngw)
  hipMemcpy2DAsync(expwf, 2 * sizeof(double) * ngw, in, 2 * sizeof(double) * ldin,
                   2 * sizeof(double) * ngw, nst, hipMemcpyDeviceToDevice, c->stream);
  if (krylov)
    hipMemcpy2DAsync(krylov, 2 * sizeof(double) * ngw, in, 2 * sizeof(double) * ldin,
                     2 * sizeof(double) * ngw, nst, hipMemcpyDeviceToDevice, c->stream);
  std::vector<double> occ(nst, 0.0);
  double fr = 1.0, fi = 0.0;  // factor
  const double* x = in;
  long ldx = ldin;
  const int hflags = whole_h_flags(flags);
  for (int N = 1; N <= order; ++N) {
    // factor *= -i*dt/N (ExponentialWavefunctionStepper.cc:97-100)
    const double m = dt / double(N);
    const double nr = fi * m, ni = -fr * m;
    fr = nr;
    fi = ni;
    double* out = (krylov && N < order) ? krylov + N * blk : bufs[(N - 1) & 1];
    if (int r = hpsi_impl(h, x, ldx, nst, occ.data(), out, ngw, hflags, nullptr)) return r;
    const long n = (long)ngw * nst;
    k_zaxpy<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(n, fr, fi, out, ngw, expwf, ngw,
                                                                ngw);
    x = out;
    ldx = ngw;
  }
  hipMemcpy2DAsync(outp, 2 * sizeof(double) * ldout, expwf, 2 * sizeof(double) * ngw,
                   2 * sizeof(double) * ngw, nst, hipMemcpyDeviceToDevice, c->stream);
  return dev_check(c, "exponential");
}

extern "C" int qbh_exponential(qbh_ham ham, double* psi_dev, long ld, int nst, double dt,
                               int order, int flags) {
  if (!ham || !psi_dev || nst <= 0 || order < 1) return QBH_ERR_ARG;
  Ham* h = &ham->h;
  Ctx* c = h->ctx;
  hipSetDevice(c->device);
  if (ld < h->b->g.ngw) return fail(c, QBH_ERR_ARG, "ld < ngw");
  if (h->b->g.real) return fail(c, QBH_ERR_ARG, "exponential needs a complex basis");
  return expo_impl(h, psi_dev, ld, psi_dev, ld, nst, dt, order, flags, nullptr);
}

// One ETRS step (ExponentialWavefunctionStepper::preupdate + update,
// ExponentialWavefunctionStepper.cc:151-276, merge_exp_ false, not AETRS):
//   half  = T(dt/2, H(t)) psi            psi(t + dt/2)
//   tilde = T(dt/2, H(t)) half           approximation to psi(t + dt) (hamil_wf)
//   rho, rhog <- density(tilde)          hamil_cd_->update_density()
//   v_r <- update_vhxc(rho, rhog)        update_hamiltonian(); update_vhxc()
//   psi   = T(dt/2, H(t + dt)) half
extern "C" int qbh_etrs_step(qbh_ham ham, qbh_basis vbasis, double* psi_dev, long ld, int nst,
                             const double* occ, double weight, double dt, int order,
                             const double* vion_g_dev, const double* rhops_g_dev, double omega,
                             double* energies, int flags) {
  if (!ham || !vbasis || !psi_dev || !occ || nst <= 0 || order < 1 || !(omega > 0.0))
    return QBH_ERR_ARG;
  Ham* h = &ham->h;
  Ctx* c = h->ctx;
  hipSetDevice(c->device);
  const GridDev& g = h->b->g;
  const GridDev& vg = basis_of(vbasis)->g;
  if (ld < g.ngw) return fail(c, QBH_ERR_ARG, "ld < ngw");
  if (g.real) return fail(c, QBH_ERR_ARG, "etrs_step needs a complex basis");
  if (ctx_of_basis(vbasis) != c) return fail(c, QBH_ERR_ARG, "vbasis on another context");
  if (vg.np0 != g.np0 || vg.np1 != g.np1 || vg.np2 != g.np2)
    return fail(c, QBH_ERR_ARG, "density basis grid differs from the wavefunction grid");
  if (!h->v) return fail(c, QBH_ERR_ARG, "v_r not set");
  const long N = (long)g.np0 * g.np1 * g.np2;
  const size_t wbytes = 2 * sizeof(double) * (size_t)g.ngw * nst;
  const size_t need = 2 * wbytes + sizeof(double) * (size_t)(N + 2L * vg.ngw);
  if (!grow(h->etrs, h->etrs_bytes, need))
    return fail(c, QBH_ERR_ALLOC, "etrs buffers allocation failed");
  double* half = static_cast<double*>(h->etrs);
  double* tilde = half + 2 * (size_t)g.ngw * nst;
  double* rho = tilde + 2 * (size_t)g.ngw * nst;
  double* rhog = rho + N;
  const int fl = whole_h_flags(flags);
  if (int r = expo_impl(h, psi_dev, ld, half, g.ngw, nst, 0.5 * dt, order, fl, nullptr)) return r;
  if (int r = expo_impl(h, half, g.ngw, tilde, g.ngw, nst, 0.5 * dt, order, fl, nullptr)) return r;
  fill0(c, rho, N);
  if (int r = qbh_density(reinterpret_cast<qbh_basis>(h->b_handle), tilde, g.ngw, nst, occ,
                          weight, omega, rho, flags & QBH_FP32))
    return r;
  if (int r = qbh_fft_forward_real(vbasis, rho, omega, rhog, 0)) return r;
  if (int r = qbh_update_vhxc(vbasis, rho, rhog, rhog, vion_g_dev, rhops_g_dev, omega, h->v,
                              energies))
    return r;
  return expo_impl(h, half, g.ngw, psi_dev, ld, nst, 0.5 * dt, order, fl, nullptr);
}
