// Context and basis set-up: device tables for the transform pipeline.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "../../include/qbh.h"
#include "qbh_internal.h"

using namespace qbh;

struct qbh_ctx_s {
  Ctx c;
};
struct qbh_basis_s {
  qbh_ctx_s* ctx;
  Basis b;
};

namespace qbh {

TimedLaunch::TimedLaunch(Ctx* c, const char* name) : ctx(c), on(c->timer.enabled) {
  if (on) {
    rec.name = name;
    hipEventCreate(&rec.a);
    hipEventCreate(&rec.b);
    hipEventRecord(rec.a, ctx->stream);
  }
}
TimedLaunch::~TimedLaunch() {
  if (on) {
    hipEventRecord(rec.b, ctx->stream);
    ctx->timer.pending.push_back(rec);
  }
}

thread_local bool t_capturing = false;

void* ws_get(Ctx* ctx, size_t bytes) {
  if (bytes <= ctx->ws_bytes) return ctx->ws;
  if (t_capturing) return nullptr;  // no free / malloc / synchronisation under capture
  if (ctx->ws) {
    hipStreamSynchronize(ctx->stream);
    hipFree(ctx->ws);
    ctx->ws = nullptr;
    ctx->ws_bytes = 0;
  }
  if (hipMalloc(&ctx->ws, bytes) != hipSuccess) {
    ctx->ws = nullptr;
    return nullptr;
  }
  ctx->ws_bytes = bytes;
  return ctx->ws;
}

}  // namespace qbh

static int set_err(Ctx* c, int code, const std::string& msg) {
  c->last_error = msg;
  return code;
}

// ------------------------------------------------------------------------------------
extern "C" int qbh_ctx_create(int device, void* stream, qbh_ctx* out) {
  if (!out) return QBH_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return QBH_ERR_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return QBH_ERR_DEVICE;
  auto* h = new qbh_ctx_s();
  h->c.device = device;
  if (stream) {
    h->c.stream = static_cast<hipStream_t>(stream);
  } else {
    if (hipStreamCreate(&h->c.stream) != hipSuccess) {
      delete h;
      return QBH_ERR_DEVICE;
    }
    h->c.own_stream = true;
  }
  qbh::noalloc_scratch_setup(&h->c);
  *out = h;
  return QBH_OK;
}

namespace qbh {
void blas_destroy(Ctx* c);
}

extern "C" int qbh_ctx_destroy(qbh_ctx ctx) {
  if (!ctx) return QBH_ERR_ARG;
  hipSetDevice(ctx->c.device);
  hipStreamSynchronize(ctx->c.stream);
  for (auto& r : ctx->c.timer.pending) {
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  qbh::blas_destroy(&ctx->c);
  if (ctx->c.side) {
    hipStreamSynchronize(ctx->c.side);
    hipStreamDestroy(ctx->c.side);
  }
  if (ctx->c.ev_fork) hipEventDestroy(ctx->c.ev_fork);
  if (ctx->c.ev_join) hipEventDestroy(ctx->c.ev_join);
  if (ctx->c.ws) hipFree(ctx->c.ws);
  if (ctx->c.own_stream) hipStreamDestroy(ctx->c.stream);
  delete ctx;
  return QBH_OK;
}

extern "C" const char* qbh_last_error(qbh_ctx ctx) {
  return ctx ? ctx->c.last_error.c_str() : "null context";
}

extern "C" int qbh_sync(qbh_ctx ctx) {
  if (!ctx) return QBH_ERR_ARG;
  hipError_t e = hipStreamSynchronize(ctx->c.stream);
  if (e != hipSuccess) return set_err(&ctx->c, QBH_ERR_DEVICE, hipGetErrorString(e));
  return QBH_OK;
}

extern "C" int qbh_ctx_set_workspace_cap(qbh_ctx ctx, double bytes) {
  if (!ctx || bytes < 1.0) return QBH_ERR_ARG;
  ctx->c.ws_cap = (size_t)bytes;
  return QBH_OK;
}

extern "C" int qbh_ctx_set_overlap(qbh_ctx ctx, int on) {
  if (!ctx) return QBH_ERR_ARG;
  ctx->c.overlap = on != 0;
  return QBH_OK;
}

extern "C" int qbh_ctx_set_timing(qbh_ctx ctx, int on) {
  if (!ctx) return QBH_ERR_ARG;
  ctx->c.timer.enabled = on != 0;
  return QBH_OK;
}

extern "C" int qbh_ctx_timing_report(qbh_ctx ctx, char* buf, int cap) {
  if (!ctx) return QBH_ERR_ARG;
  hipStreamSynchronize(ctx->c.stream);
  auto& t = ctx->c.timer;
  for (auto& r : t.pending) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, r.a, r.b);
    auto& e = t.totals[r.name];
    e.first += ms;
    e.second += 1;
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  t.pending.clear();
  std::string s;
  for (auto& kv : t.totals) {
    char line[256];
    snprintf(line, sizeof(line), "%s %.6f %ld\n", kv.first.c_str(), kv.second.first,
             kv.second.second);
    s += line;
  }
  t.totals.clear();
  if (buf && cap > 0) {
    strncpy(buf, s.c_str(), cap - 1);
    buf[cap - 1] = 0;
  }
  return (int)s.size();
}

// ------------------------------------------------------------------------------------
// Stockham radix plan for one axis.  Variant 0: powers of two as 8s and 4s (a single 2
// only when the exponent is 1), then 3, 5, 7, 11; variant 1: powers of two as 4s (and one
// 2); variant 2: the odd radices first, then the powers of two of variant 0.  Every
// variant is the same transform; the choice only changes the pass count and strides.
bool qbh::make_axis_variant(int n, int variant, AxisPlan& p) {
  p.n = n;
  p.npass = 0;
  int m = n, e = 0;
  while (m % 2 == 0) {
    m /= 2;
    ++e;
  }
  auto push = [&](int r) {
    if (p.npass >= 8) return false;
    p.radix[p.npass++] = r;
    return true;
  };
  auto odd = [&]() {
    int q = m;
    for (int r : {3, 5, 7, 11})
      while (q % r == 0) {
        q /= r;
        if (!push(r)) return false;
      }
    return q == 1;
  };
  auto twos = [&]() {
    if (variant == 1) {
      for (int i = 0; i < e / 2; ++i) if (!push(4)) return false;
      if (e % 2 && !push(2)) return false;
      return true;
    }
    if (e == 1) return push(2);
    int e8 = e / 3, rem = e % 3;
    if (rem == 1 && e8 > 0) {  // 2^(3k+1) = 8^(k-1) * 4 * 4
      e8 -= 1;
      for (int i = 0; i < e8; ++i) if (!push(8)) return false;
      return push(4) && push(4);
    }
    for (int i = 0; i < e8; ++i) if (!push(8)) return false;
    if (rem == 2 && !push(4)) return false;
    if (rem == 1 && !push(2)) return false;
    return true;
  };
  if (variant == 2) return odd() && twos();
  return twos() && odd();
}

template <typename X>
static X* dev_upload(Basis& b, const std::vector<X>& h) {
  void* d = nullptr;
  const size_t bytes = std::max<size_t>(1, h.size()) * sizeof(X);
  if (hipMalloc(&d, bytes) != hipSuccess) return nullptr;
  b.allocations.push_back(d);
  if (!h.empty()) hipMemcpy(d, h.data(), h.size() * sizeof(X), hipMemcpyHostToDevice);
  return static_cast<X*>(d);
}

constexpr int kZL = 16;  // must match ZL in qbh_fft.hip

extern "C" int qbh_basis_create(qbh_ctx ctx, int real, int nrods, const int* rod_h,
                                const int* rod_k, const int* rod_lmin, const int* rod_size,
                                int np0, int np1, int np2, const double* bvec,
                                const double* kpoint, qbh_basis* out) {
  if (!ctx || !out || nrods <= 0 || !rod_h || !rod_k || !rod_lmin || !rod_size || !bvec)
    return QBH_ERR_ARG;
  Ctx* c = &ctx->c;
  *out = nullptr;
  hipSetDevice(c->device);
  const int nps[3] = {np0, np1, np2};
  for (int i = 0; i < 3; ++i)
    if (nps[i] < 2 || nps[i] > 256)
      return set_err(c, QBH_ERR_UNSUPPORTED, "grid sizes must be in [2, 256]");
  auto* hb = new qbh_basis_s();
  hb->ctx = ctx;
  Basis& b = hb->b;
  b.device = c->device;
  b.rod_h.assign(rod_h, rod_h + nrods);
  b.rod_k.assign(rod_k, rod_k + nrods);
  b.rod_lmin.assign(rod_lmin, rod_lmin + nrods);
  b.rod_size.assign(rod_size, rod_size + nrods);
  std::memcpy(b.b, bvec, 9 * sizeof(double));
  if (kpoint) std::memcpy(b.kpoint, kpoint, 3 * sizeof(double));
  GridDev& g = b.g;
  g.np0 = np0;
  g.np1 = np1;
  g.np2 = np2;
  g.real = real ? 1 : 0;
  g.nrods = nrods;
  auto fail = [&](const char* msg) {
    qbh_basis_destroy(hb);
    return set_err(c, QBH_ERR_ARG, msg);
  };
  if (real && (rod_h[0] != 0 || rod_k[0] != 0 || rod_lmin[0] != 0))
    return fail("real basis: first rod must be rod(0,0) starting at l=0");
  std::vector<int> first(nrods);
  long ngw = 0;
  for (int r = 0; r < nrods; ++r) {
    if (rod_size[r] <= 0) return fail("rod size must be positive");
    const int lo = rod_lmin[r], hi = rod_lmin[r] + rod_size[r] - 1;
    if (2 * std::max(std::abs(lo), std::abs(hi)) >= np2) return fail("rod exceeds np2/2");
    if (2 * std::abs(rod_h[r]) >= np0 || 2 * std::abs(rod_k[r]) >= np1)
      return fail("rod (h,k) exceeds np0/2 or np1/2");
    if (real && r > 0 && (rod_h[r] < 0 || (rod_h[r] == 0 && rod_k[r] <= 0)))
      return fail("real basis: rods must lie in the half plane h>0 or (h=0, k>0)");
    first[r] = (int)ngw;
    ngw += rod_size[r];
  }
  if (ngw > (1L << 30)) return fail("basis too large");
  g.ngw = (int)ngw;

  // columns
  struct Col {
    int x, y, rod, kind;
  };
  std::vector<Col> cols;
  auto wrapn = [](int v, int n) { return v < 0 ? v + n : v; };
  for (int r = 0; r < nrods; ++r) {
    const int h = rod_h[r], k = rod_k[r];
    if (real) {
      if (r == 0) {
        cols.push_back({0, 0, 0, 0});
      } else {
        cols.push_back({wrapn(h, np0), wrapn(k, np1), r, 1});
        cols.push_back({wrapn(-h, np0), wrapn(-k, np1), r, 2});
      }
    } else {
      cols.push_back({wrapn(h, np0), wrapn(k, np1), r, 3});
    }
  }
  const int ncol = (int)cols.size();
  g.ncol = ncol;
  std::vector<int> order(ncol);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int i, int j) {
    return cols[i].x != cols[j].x ? cols[i].x < cols[j].x : cols[i].y < cols[j].y;
  });
  std::vector<int> colpos(ncol), col_y(ncol);
  std::vector<int> line_start, line_x;
  for (int p = 0; p < ncol; ++p) {
    const Col& cc = cols[order[p]];
    if (p > 0) {
      const Col& pc = cols[order[p - 1]];
      if (pc.x == cc.x && pc.y == cc.y) return fail("two columns map to the same (x,y)");
    }
    if (p == 0 || cols[order[p - 1]].x != cc.x) {
      line_start.push_back(p);
      line_x.push_back(cc.x);
    }
    colpos[order[p]] = p;
    col_y[p] = cc.y;
  }
  line_start.push_back(ncol);
  g.nlx = (int)line_x.size();
  g.xpos_max = -1;
  g.xneg_min = np0;
  for (int x : line_x) {
    if (2 * x < np0) g.xpos_max = std::max(g.xpos_max, x);
    else g.xneg_min = std::min(g.xneg_min, x);
  }

  // z-pass groups: consecutive columns in creation order (the +G/-G images of a rod
  // are adjacent), at most kZL lines per group (4 per wavefront)
  const int maxl = kZL;
  std::vector<int> grp_start{0}, ent_col, ent_rod, ent_kind;
  int inl = 0;
  // Γ: rod(0,0) is followed by an empty line (kind 4) so that every +G/-G pair starts
  // at an even line and never straddles the 4-line share of a wavefront
  for (int r = 0, ci = 0; r < nrods; ++r) {
    const int nc = (real && r > 0) ? 2 : 1;
    const int nl = real ? 2 : 1;
    if (inl + nl > maxl) {
      grp_start.push_back((int)ent_col.size());
      inl = 0;
    }
    for (int j = 0; j < nc; ++j, ++ci) {
      ent_col.push_back(colpos[ci]);
      ent_rod.push_back(r);
      ent_kind.push_back(cols[ci].kind);
    }
    if (real && r == 0) {
      ent_col.push_back(-1);
      ent_rod.push_back(0);
      ent_kind.push_back(4);
    }
    inl += nl;
  }
  grp_start.push_back((int)ent_col.size());
  g.ngroups = (int)grp_start.size() - 1;

  // twiddles exp(+2 pi i m/n) in extended precision
  for (int ax = 0; ax < 3; ++ax) {
    const int n = nps[ax];
    if (!make_axis_variant(n, 0, g.ax[ax])) return fail("FFT length has unsupported factors");
    std::vector<cplx<double>> t64(n);
    std::vector<cplx<float>> t32(n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int m = 0; m < n; ++m) {
      const long double ang = 2.0L * pi * (long double)m / (long double)n;
      t64[m] = {(double)cosl(ang), (double)sinl(ang)};
      t32[m] = {(float)cosl(ang), (float)sinl(ang)};
    }
    g.tw64[ax] = dev_upload(b, t64);
    g.tw32[ax] = dev_upload(b, t32);
  }
  g.rod_first = dev_upload(b, first);
  g.rod_lmin = dev_upload(b, b.rod_lmin);
  g.rod_size = dev_upload(b, b.rod_size);
  g.grp_start = dev_upload(b, grp_start);
  g.ent_col = dev_upload(b, ent_col);
  g.ent_rod = dev_upload(b, ent_rod);
  g.ent_kind = dev_upload(b, ent_kind);
  g.col_y = dev_upload(b, col_y);
  g.line_start = dev_upload(b, line_start);
  g.line_x = dev_upload(b, line_x);

  // |k+G|^2 and k+G (Basis::update_g, Basis.cc:703-751, same evaluation order)
  std::vector<double> kpg2(ngw), kpgx(3 * ngw);
  {
    const double* B = bvec;
    long i = 0;
    for (int r = 0; r < nrods; ++r)
      for (int j = 0; j < rod_size[r]; ++j, ++i) {
        const double h = b.kpoint[0] + rod_h[r], k = b.kpoint[1] + rod_k[r],
                     l = b.kpoint[2] + (rod_lmin[r] + j);
        double v[3];
        for (int d = 0; d < 3; ++d) v[d] = (h * B[d] + k * B[3 + d]) + l * B[6 + d];
        kpgx[i] = v[0];
        kpgx[ngw + i] = v[1];
        kpgx[2 * ngw + i] = v[2];
        kpg2[i] = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
      }
  }
  b.kpg2 = dev_upload(b, kpg2);
  b.kpgx = dev_upload(b, kpgx);
  if (!g.rod_first || !g.line_x || !b.kpg2 || !b.kpgx) return fail("device allocation failed");
  *out = hb;
  return QBH_OK;
}

extern "C" int qbh_basis_destroy(qbh_basis basis) {
  if (!basis) return QBH_ERR_ARG;
  hipSetDevice(basis->b.device);
  hipDeviceSynchronize();
  for (void* p : basis->b.allocations) hipFree(p);
  delete basis;
  return QBH_OK;
}

extern "C" int qbh_basis_info(qbh_basis basis, int* ngw, int* ncol, int* nlx) {
  if (!basis) return QBH_ERR_ARG;
  if (ngw) *ngw = basis->b.g.ngw;
  if (ncol) *ncol = basis->b.g.ncol;
  if (nlx) *nlx = basis->b.g.nlx;
  return QBH_OK;
}

namespace qbh {
Basis* basis_of(qbh_basis b) { return &b->b; }
Ctx* ctx_of_basis(qbh_basis b) { return &b->ctx->c; }
Ctx* ctx_of(qbh_ctx c) { return &c->c; }
}  // namespace qbh
