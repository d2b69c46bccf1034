// Launch tuning, the autotuner, and the multi-cell Hψ graph.
//
// Knobs (all choices compute the same Hψ; they differ in speed only):
//   basis: zy_reg (register two-factor z/y transforms where instantiated), x_reg (register x
//          pass), x_spw (slots per x-pass workgroup, 0 = occupancy heuristic), radix
//          (Stockham radix plan variant for lengths without a register transform);
//   ham:   nsplit (K slices of the projector contraction, 0 = heuristic), overlap (projector
//          GEMMs on a side stream concurrent with the transforms, -1 = context default),
//          fnl_algo / enl_algo (projector GEMMs by this library's matrix-core kernels (17),
//          by rocBLAS (0) or by the n-th algorithm hipBLASLt lists for the shape (1..16)).
// qbh_ham_autotune times candidate values with HIP events on the context's stream, one
// knob at a time (coordinate descent from the current choice), and caches the result per
// problem shape in the context.
//
// The batch graph captures Hψ of several cells into one hipGraph: cells are spread over
// branch streams (each with its own transform workspace and rocBLAS handle) forked from and
// joined to the capture stream, so small cells run concurrently; one hipGraphLaunch runs
// them all.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
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

static int ufail(Ctx* c, int code, const std::string& m) {
  c->last_error = m;
  return code;
}

static const char* const KNOBS[] = {"zy_reg", "x_reg",   "x_spw",    "radix",
                                    "nsplit", "overlap", "fnl_algo", "enl_algo"};
constexpr int NKNOB = 8;
constexpr int kMaxAlgo = 16;  // hipBLASLt algorithms listed per GEMM shape

static int knob_index(const char* name) {
  for (int i = 0; i < NKNOB; ++i)
    if (name && std::strcmp(name, KNOBS[i]) == 0) return i;
  return -1;
}

static int get_knob(Ham* h, int k) {
  Basis* b = h->b;
  switch (k) {
    case 0: return b->tune.zy_reg;
    case 1: return b->tune.x_reg;
    case 2: return b->tune.x_spw;
    case 3: return b->tune.radix;
    case 4: return h->nsplit;
    case 5: return h->overlap;
    case 6: return h->fnl_algo;
    default: return h->enl_algo;
  }
}

static int set_knob(Ham* h, int k, int v) {
  Basis* b = h->b;
  switch (k) {
    case 0: b->tune.zy_reg = v ? 1 : 0; return QBH_OK;
    case 1: b->tune.x_reg = v ? 1 : 0; return QBH_OK;
    case 2:
      if (v < 0 || v > 64) return QBH_ERR_ARG;
      b->tune.x_spw = v;
      return QBH_OK;
    case 3: {
      if (v < 0 || v > 2) return QBH_ERR_ARG;
      AxisPlan ax[3];
      const int n[3] = {b->g.np0, b->g.np1, b->g.np2};
      for (int a = 0; a < 3; ++a)
        if (!make_axis_variant(n[a], v, ax[a])) return QBH_ERR_UNSUPPORTED;
      for (int a = 0; a < 3; ++a) b->g.ax[a] = ax[a];
      b->tune.radix = v;
      return QBH_OK;
    }
    case 4:
      if (v < 0 || v > 64) return QBH_ERR_ARG;
      h->nsplit = v;
      return QBH_OK;
    case 5:
      if (v < -1 || v > 1) return QBH_ERR_ARG;
      h->overlap = v;
      return QBH_OK;
    default:
      if (v < -1 || (v > kMaxAlgo && v != kOwnGemm)) return QBH_ERR_ARG;
      (k == 6 ? h->fnl_algo : h->enl_algo) = v;
      return QBH_OK;
  }
}

extern "C" int qbh_ham_set_tuning(qbh_ham ham, const char* knob, int value) {
  if (!ham) return QBH_ERR_ARG;
  Ham* h = &ham->h;
  const int k = knob_index(knob);
  if (k < 0) return ufail(h->ctx, QBH_ERR_ARG, std::string("unknown tuning knob ") + (knob ? knob : "(null)"));
  hipSetDevice(h->ctx->device);
  hipStreamSynchronize(h->ctx->stream);
  const int r = set_knob(h, k, value);
  if (r) return ufail(h->ctx, r, std::string("value out of range for ") + knob);
  return QBH_OK;
}

extern "C" int qbh_ham_get_tuning(qbh_ham ham, const char* knob, int* value) {
  if (!ham || !value) return QBH_ERR_ARG;
  const int k = knob_index(knob);
  if (k < 0) return QBH_ERR_ARG;
  *value = get_knob(&ham->h, k);
  return QBH_OK;
}

// mean time (ms) of `reps` Hψ applications after one untimed one
static int time_hpsi(Ham* h, const double* psi, long ld, int nst, const double* occ, double* hpsi,
                     long ldh, int flags, int reps, float* ms) {
  Ctx* c = h->ctx;
  if (int r = hpsi_impl(h, psi, ld, nst, occ, hpsi, ldh, flags, nullptr)) return r;
  hipEvent_t a, b;
  hipEventCreate(&a);
  hipEventCreate(&b);
  hipEventRecord(a, c->stream);
  int r = QBH_OK;
  for (int i = 0; i < reps && r == QBH_OK; ++i)
    r = hpsi_impl(h, psi, ld, nst, occ, hpsi, ldh, flags, nullptr);
  hipEventRecord(b, c->stream);
  hipEventSynchronize(b);
  float t = 0.0f;
  hipEventElapsedTime(&t, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  *ms = t / std::max(1, reps);
  return r;
}

static std::string shape_key(Ham* h, int nst, int flags) {
  const GridDev& g = h->b->g;
  char buf[160];
  std::snprintf(buf, sizeof buf, "%d.%d.%d.%d.%d.%d.%d.%d", g.np0, g.np1, g.np2, g.ngw, g.real,
                nst, h->P, flags & (QBH_ALL | QBH_FP32));
  return buf;
}

extern "C" int qbh_ham_autotune(qbh_ham ham, const double* psi_dev, long ld, int nst,
                                const double* occ, double* hpsi_dev, long ldh, int flags,
                                int reps, int force, int knob_mask, char* report, int cap) {
  if (!ham || !psi_dev || !hpsi_dev || nst <= 0) return QBH_ERR_ARG;
  Ham* h = &ham->h;
  Ctx* c = h->ctx;
  hipSetDevice(c->device);
  const int fl = whole_h_flags(flags);
  const int mask = knob_mask == 0 ? (1 << NKNOB) - 1 : knob_mask;
  const std::string key = shape_key(h, nst, fl) + "." + std::to_string(mask);
  std::string rep;
  auto it = c->tune_cache.find(key);
  if (!force && it != c->tune_cache.end()) {
    for (int k = 0; k < NKNOB; ++k)
      if (int r = set_knob(h, k, it->second[k])) return ufail(c, r, "cached tuning invalid");
    rep = "cached " + key + "\n";
  } else {
    const GridDev& g = h->b->g;
    const bool nl = (fl & QBH_NONLOCAL) && h->P > 0, loc = (fl & QBH_LOCAL) != 0;
    const bool zy = register_edge(g.np1) || register_edge(g.np2);
    std::vector<std::pair<int, std::vector<int>>> plan;
    // transform knobs first, then the schedule, then the K split (whose best value depends
    // on whether the GEMMs share the GPU with the transforms)
    if (loc && zy) plan.push_back({0, {1, 0}});
    if (loc && register_edge(g.np0)) plan.push_back({1, {1, 0}});
    if (loc) plan.push_back({3, {0, 1, 2}});
    if (loc) plan.push_back({2, {0, 1, 2, 4, 8, 16}});
    if (nl && loc) plan.push_back({5, {1, 0}});
    if (nl) plan.push_back({4, {0, 1, 2, 4, 8, 16, 32}});
    // then the GEMM library and algorithm, against the tuned rocBLAS schedule
    for (int which = 0; nl && which < 2; ++which) {
      const int na = nl_lt_candidates(h, which, nst, ld, ldh, (fl & QBH_FP32) ? 1 : 0);
      std::vector<int> vals;
      vals.push_back(kOwnGemm);  // the matrix-core kernels of this library
      for (int v = 0; v <= std::min(na, kMaxAlgo); ++v) vals.push_back(v);
      if (vals.size() > 1) {
        // start from the value the built-in choice stands for, so it is not timed twice
        int& knob = which ? h->enl_algo : h->fnl_algo;
        if (knob < 0) knob = nl_builtin_algo(h, which);
        plan.push_back({6 + which, vals});
      }
    }
    plan.erase(std::remove_if(plan.begin(), plan.end(),
                              [&](const std::pair<int, std::vector<int>>& kv) {
                                return !(mask & (1 << kv.first));
                              }),
               plan.end());
    float best = 0.0f;
    if (int r = time_hpsi(h, psi_dev, ld, nst, occ, hpsi_dev, ldh, fl, reps, &best)) return r;
    char line[160];
    std::snprintf(line, sizeof line, "start %.4f ms\n", best);
    rep += line;
    for (auto& kv : plan) {
      const int k = kv.first;
      const int cur = get_knob(h, k);
      int bestv = cur;
      for (int v : kv.second) {
        if (v == cur) continue;
        if (set_knob(h, k, v) != QBH_OK) continue;
        float t = 0.0f;
        if (int r = time_hpsi(h, psi_dev, ld, nst, occ, hpsi_dev, ldh, fl, reps, &t)) {
          set_knob(h, k, cur);
          return r;
        }
        std::snprintf(line, sizeof line, "  %s=%d %.4f ms\n", KNOBS[k], v, t);
        rep += line;
        if (t < 0.99f * best) {
          best = t;
          bestv = v;
        }
      }
      set_knob(h, k, bestv);
      std::snprintf(line, sizeof line, "%s -> %d (%.4f ms)\n", KNOBS[k], bestv, best);
      rep += line;
    }
    std::vector<int> vals(NKNOB);
    for (int k = 0; k < NKNOB; ++k) vals[k] = get_knob(h, k);
    c->tune_cache[key] = vals;
  }
  if (report && cap > 0) {
    std::strncpy(report, rep.c_str(), (size_t)cap - 1);
    report[cap - 1] = '\0';
  }
  return QBH_OK;
}

// ------------------------------------------------------------------------------------
// multi-cell Hψ graph
constexpr int kBatchGraphEnabled = 0;

struct Cell {
  Ham* h;
  const double* psi;
  long ld;
  int nst;
  std::vector<double> occ;
  double* hpsi;
  long ldh;
};

struct qbh_batch_s {
  Ctx* user = nullptr;
  std::vector<Ctx*> branch;        // per-branch contexts (stream, workspace, BLAS handle)
  std::vector<Cell> cells;
  std::vector<int> cell_branch;
  int flags = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<hipEvent_t> events;
};

static void destroy_branch(Ctx* b) {
  blas_destroy(b);  // rocBLAS handle and hipBLASLt state
  if (b->ws) hipFree(b->ws);
  if (b->side) hipStreamDestroy(b->side);
  if (b->ev_fork) hipEventDestroy(b->ev_fork);
  if (b->ev_join) hipEventDestroy(b->ev_join);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}

extern "C" int qbh_batch_destroy(qbh_batch bt) {
  if (!bt) return QBH_ERR_ARG;
  hipSetDevice(bt->user->device);
  hipStreamSynchronize(bt->user->stream);
  if (bt->exec) hipGraphExecDestroy(bt->exec);
  if (bt->graph) hipGraphDestroy(bt->graph);
  for (hipEvent_t e : bt->events) hipEventDestroy(e);
  for (Ctx* b : bt->branch) destroy_branch(b);
  delete bt;
  return QBH_OK;
}

// runs every cell of branch `bi` on that branch's context
static int run_branch(qbh_batch_s* bt, int bi) {
  for (size_t i = 0; i < bt->cells.size(); ++i) {
    if (bt->cell_branch[i] != bi) continue;
    Cell& cl = bt->cells[i];
    Ctx* own = cl.h->ctx;
    cl.h->ctx = bt->branch[bi];
    const int r = hpsi_impl(cl.h, cl.psi, cl.ld, cl.nst, cl.occ.data(), cl.hpsi, cl.ldh,
                            bt->flags, nullptr);
    cl.h->ctx = own;
    if (r) {
      bt->user->last_error = bt->branch[bi]->last_error;
      return r;
    }
  }
  return QBH_OK;
}

extern "C" int qbh_batch_create(qbh_ctx ctx, int ncells, const qbh_ham* hams,
                                const double* const* psi_dev, const long* ld, const int* nst,
                                double* const* hpsi_dev, const long* ldh, int flags, int nbranch,
                                qbh_batch* out) {
  if (!ctx || ncells <= 0 || !hams || !psi_dev || !ld || !nst || !hpsi_dev || !ldh || !out)
    return QBH_ERR_ARG;
  Ctx* uc = ctx_of(ctx);
  // Switched off: the first GPU run of this entry point ended in a host segmentation fault
  // inside this call (eager warm-up or the stream capture of the cells' Hψ, which includes
  // the rocBLAS projector GEMMs); the cause is not identified, so the path stays disabled.
  if (kBatchGraphEnabled == 0)
    return ufail(uc, QBH_ERR_UNSUPPORTED,
                 "qbh_batch_create is disabled: an unresolved host fault during graph capture");
  hipSetDevice(uc->device);
  auto* bt = new qbh_batch_s();
  bt->user = uc;
  bt->flags = whole_h_flags(flags);
  // cells of one Hamiltonian share its scratch: keep them on one branch
  std::map<Ham*, int> ham_branch;
  nbranch = std::max(1, std::min(nbranch, ncells));
  for (int i = 0; i < ncells; ++i) {
    if (!hams[i] || !psi_dev[i] || !hpsi_dev[i] || nst[i] <= 0 ||
        ((reinterpret_cast<uintptr_t>(psi_dev[i]) | reinterpret_cast<uintptr_t>(hpsi_dev[i])) & 15) ||
        psi_dev[i] == hpsi_dev[i]) {
      qbh_batch_destroy(bt);
      return QBH_ERR_ARG;
    }
    Ham* h = &hams[i]->h;
    if (h->ctx != uc) {
      qbh_batch_destroy(bt);
      return ufail(uc, QBH_ERR_ARG, "batch cell on another context");
    }
    Cell cl{h, psi_dev[i], ld[i], nst[i], {}, hpsi_dev[i], ldh[i]};
    cl.occ.assign(nst[i], 0.0);  // occupations only feed Enl, which the graph does not return
    bt->cells.push_back(std::move(cl));
    auto f = ham_branch.find(h);
    const int b = f != ham_branch.end() ? f->second : (int)(ham_branch.size() % nbranch);
    ham_branch[h] = b;
    bt->cell_branch.push_back(b);
  }
  for (int b = 0; b < nbranch; ++b) {
    Ctx* c = new Ctx();
    c->device = uc->device;
    c->overlap = uc->overlap;
    c->ws_cap = uc->ws_cap;
    bt->branch.push_back(c);
    if (hipStreamCreate(&c->stream) != hipSuccess) {
      qbh_batch_destroy(bt);
      return ufail(uc, QBH_ERR_DEVICE, "branch stream creation failed");
    }
    c->own_stream = true;
  }
  // eager pass: allocates every workspace, table and scratch buffer the graph will use
  int r = QBH_OK;
  for (int b = 0; b < nbranch && r == QBH_OK; ++b) r = run_branch(bt, b);
  for (Ctx* c : bt->branch) hipStreamSynchronize(c->stream);
  if (r) {
    qbh_batch_destroy(bt);
    return r;
  }
  // capture: fork every branch from the user stream, join them back
  bt->events.resize(1 + nbranch, nullptr);
  for (auto& e : bt->events) hipEventCreateWithFlags(&e, hipEventDisableTiming);
  hipStream_t s0 = uc->stream;
  if (hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    qbh_batch_destroy(bt);
    return ufail(uc, QBH_ERR_DEVICE, "stream capture could not start (the stream must be a "
                                     "non-default stream)");
  }
  hipEventRecord(bt->events[0], s0);
  t_capturing = true;  // growth, uploads and read-backs fail instead of running
  for (int b = 0; b < nbranch && r == QBH_OK; ++b) {
    hipStreamWaitEvent(bt->branch[b]->stream, bt->events[0], 0);
    r = run_branch(bt, b);
    hipEventRecord(bt->events[1 + b], bt->branch[b]->stream);
    hipStreamWaitEvent(s0, bt->events[1 + b], 0);
  }
  t_capturing = false;
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(s0, &g);
  bt->graph = g;
  if (r == QBH_OK && ec != hipSuccess)
    r = ufail(uc, QBH_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(ec));
  if (r == QBH_OK && hipGraphInstantiate(&bt->exec, g, nullptr, nullptr, 0) != hipSuccess)
    r = ufail(uc, QBH_ERR_DEVICE, "graph instantiation failed");
  (void)hipGetLastError();
  if (r) {
    qbh_batch_destroy(bt);
    return r;
  }
  *out = bt;
  return QBH_OK;
}

extern "C" int qbh_batch_launch(qbh_batch bt) {
  if (!bt || !bt->exec) return QBH_ERR_ARG;
  hipSetDevice(bt->user->device);
  const hipError_t e = hipGraphLaunch(bt->exec, bt->user->stream);
  if (e != hipSuccess) return ufail(bt->user, QBH_ERR_DEVICE, std::string("graph launch: ") + hipGetErrorString(e));
  return QBH_OK;
}

// ------------------------------------------------------------------------------------
// Stream bandwidth of the device: read-only, write-only and copy kernels over two buffers
// of `bytes`, best over launch shapes of 1024, 4096 and 16384 workgroups of 256 threads
// with 1, 4 or 8 independent 16-byte accesses in flight per thread per iteration, and of
// `reps` launches each.  The copy rate counts read plus written bytes.  What a plain
// streaming pass reaches on this GPU is the practical ceiling against which the transform
// passes' HBM rates are read.
namespace qbh {
template <int U>
__global__ void __launch_bounds__(256)
k_bw_read(const double2* __restrict__ a, long n, double* __restrict__ sink) {
  double s = 0.0;
  const long stride = (long)gridDim.x * 256 * U;
  for (long i0 = blockIdx.x * 256L * U + threadIdx.x; i0 < n; i0 += stride) {
    double2 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long i = i0 + u * 256L;
      v[u] = i < n ? a[i] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) s += v[u].x + v[u].y;
  }
  if (s == -1.0) sink[0] = s;  // the buffer holds non-negative values: never stored
}
template <int U>
__global__ void __launch_bounds__(256) k_bw_write(double2* __restrict__ a, long n) {
  const long stride = (long)gridDim.x * 256 * U;
  for (long i0 = blockIdx.x * 256L * U + threadIdx.x; i0 < n; i0 += stride) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long i = i0 + u * 256L;
      if (i < n) a[i] = make_double2(1.0, 2.0);
    }
  }
}
template <int U>
__global__ void __launch_bounds__(256)
k_bw_copy(const double2* __restrict__ a, double2* __restrict__ b, long n) {
  const long stride = (long)gridDim.x * 256 * U;
  for (long i0 = blockIdx.x * 256L * U + threadIdx.x; i0 < n; i0 += stride) {
    double2 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long i = i0 + u * 256L;
      v[u] = i < n ? a[i] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long i = i0 + u * 256L;
      if (i < n) b[i] = v[u];
    }
  }
}

template <int U>
static void bw_launch(int k, unsigned grid, hipStream_t st, void* a, void* b, long n, void* sink) {
  if (k == 0) k_bw_read<U><<<grid, 256, 0, st>>>(static_cast<const double2*>(a), n,
                                                 static_cast<double*>(sink));
  if (k == 1) k_bw_write<U><<<grid, 256, 0, st>>>(static_cast<double2*>(b), n);
  if (k == 2) k_bw_copy<U><<<grid, 256, 0, st>>>(static_cast<const double2*>(a),
                                                 static_cast<double2*>(b), n);
}
}  // namespace qbh

extern "C" int qbh_ctx_stream_bandwidth(qbh_ctx ctx, double bytes, int reps, double* gbps) {
  using namespace qbh;
  if (!ctx || !gbps || reps < 1 || !(bytes >= 1048576.0) || bytes > 68719476736.0)
    return QBH_ERR_ARG;
  Ctx* c = ctx_of(ctx);
  hipSetDevice(c->device);
  const long n = (long)(bytes / sizeof(double2));
  const size_t nb = (size_t)n * sizeof(double2);
  void *a = nullptr, *b = nullptr, *sink = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int r = QBH_OK;
  if (hipMalloc(&a, nb) != hipSuccess || hipMalloc(&b, nb) != hipSuccess ||
      hipMalloc(&sink, 64) != hipSuccess) {
    r = ufail(c, QBH_ERR_ALLOC, "bandwidth buffers allocation failed");
  } else {
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    // a holds non-negative values (read and copied), b is written
    k_bw_write<1><<<2048, 256, 0, c->stream>>>(static_cast<double2*>(a), n);
    float best[3] = {1e30f, 1e30f, 1e30f};
    const unsigned grids[3] = {1024, 4096, 16384};
    const int unrolls[3] = {1, 4, 8};
    for (int k = 0; k < 3 && r == QBH_OK; ++k)
      for (unsigned grid : grids)
        for (int u : unrolls)
          for (int rep = 0; rep <= reps && r == QBH_OK; ++rep) {  // launch 0 is a warm-up
            hipEventRecord(e0, c->stream);
            if (u == 1) bw_launch<1>(k, grid, c->stream, a, b, n, sink);
            if (u == 4) bw_launch<4>(k, grid, c->stream, a, b, n, sink);
            if (u == 8) bw_launch<8>(k, grid, c->stream, a, b, n, sink);
            hipEventRecord(e1, c->stream);
            if (hipEventSynchronize(e1) != hipSuccess) {
              r = ufail(c, QBH_ERR_DEVICE, "bandwidth kernels failed");
              break;
            }
            float t = 0.f;
            hipEventElapsedTime(&t, e0, e1);
            if (rep > 0 && t < best[k]) best[k] = t;
          }
    if (r == QBH_OK) {
      gbps[0] = nb / (best[0] * 1e-3) / 1e9;
      gbps[1] = nb / (best[1] * 1e-3) / 1e9;
      gbps[2] = 2.0 * nb / (best[2] * 1e-3) / 1e9;
    }
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipStreamSynchronize(c->stream);
  for (void* p : {a, b, sink})
    if (p) hipFree(p);
  (void)hipGetLastError();
  return r;
}
