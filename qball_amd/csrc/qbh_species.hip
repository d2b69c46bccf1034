// Setup of the Hψ path from a norm-conserving Kleinman-Bylander pseudopotential:
//
//   * the species radial tables of Species::initialize, KB branch (src/qball/Species.cc:
//     74-673): the radial mesh extended to ndft points with -zval/r, the local potential
//     minus the long-range part of the Gaussian ion charge (:1749-1759) and its transform
//     Vloc(G) (:280-330), the KB weights wsg (:340-360) and the projector transforms
//     (bessel_trans, :773-1000), each put on a cubic spline in G (pseudo/spline.cc);
//   * the projector tables twnl of NonLocalPotential::update_twnl (NonLocalPotential.cc:
//     261-1521, KB branch) at |k+G-A| of a wavefunction basis (GPU);
//   * v_ion(G) and rho_ps(G) of EnergyFunctional::cell_moved / atoms_moved
//     (EnergyFunctional.cc:1889-1915, 2001-2021) on the density basis (GPU);
//   * the pseudo-charge repulsion esr of atoms_moved (:1918-1988) and eself (Species.h:160).
//
// The radial tables are one-dimensional and built once per species on the host.  The
// radial sine and cosine transforms are the ones Numerical Recipes' sinft / cosft1 compute
// (sinft.cc), evaluated here as direct sums with exact angle indices (sin(pi j k / n) =
// table[(j k) mod 2n]), which equals the FFT form up to rounding.  The per-G tables are
// GPU kernels over the basis' k+G table.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/qbh.h"
#include "qbh_ham.h"
#include "qbh_internal.h"

namespace qbh {
Basis* basis_of(qbh_basis b);
Ctx* ctx_of_basis(qbh_basis b);
}  // namespace qbh

using namespace qbh;

namespace {

// Species.cc:45-62, extended Simpson rule on overlapping segments
double simpsn(const std::vector<double>& t) {
  const int n = (int)t.size();
  double s = (17.0 / 48.0) * (t[0] + t[n - 1]) + (59.0 / 48.0) * (t[1] + t[n - 2]) +
             (43.0 / 48.0) * (t[2] + t[n - 3]) + (49.0 / 48.0) * (t[3] + t[n - 4]);
  for (int i = 4; i < n - 4; ++i) s += t[i];
  return s;
}

// F_k = sum_{j=1}^{n-1} y_j sin(pi j k / n), k = 0..n-1 (what sinft computes)
std::vector<double> sine_transform(const std::vector<double>& y) {
  const long n = (long)y.size();
  std::vector<double> tab(2 * n), out(n, 0.0);
  for (long m = 0; m < 2 * n; ++m) tab[m] = std::sin(M_PI * (double)m / (double)n);
  for (long k = 1; k < n; ++k) {
    double s = 0.0;
    long idx = k;  // (j k) mod 2n for j = 1
    for (long j = 1; j < n; ++j) {
      s += y[j] * tab[idx];
      idx += k;
      if (idx >= 2 * n) idx -= 2 * n;
    }
    out[k] = s;
  }
  return out;
}

// y of n+1 points: F_k = (y_0 + (-1)^k y_n) / 2 + sum_{j=1}^{n-1} y_j cos(pi j k / n),
// k = 0..n-1 (what cosft1 computes)
std::vector<double> cosine_transform(const std::vector<double>& y) {
  const long n = (long)y.size() - 1;
  std::vector<double> tab(2 * n), out(n, 0.0);
  for (long m = 0; m < 2 * n; ++m) tab[m] = std::cos(M_PI * (double)m / (double)n);
  for (long k = 0; k < n; ++k) {
    double s = 0.5 * (y[0] + ((k & 1) ? -y[n] : y[n]));
    long idx = k;
    for (long j = 1; j < n; ++j) {
      s += y[j] * tab[idx];
      idx += k;
      if (idx >= 2 * n) idx -= 2 * n;
    }
    out[k] = s;
  }
  return out;
}

// pseudo/spline.cc spline(): second derivatives for first derivative yp1 at x[0]
// (natural when natural0) and a natural end at x[n-1]
std::vector<double> spline_fit(const std::vector<double>& x, const std::vector<double>& y,
                               bool natural0, double yp1) {
  const int n = (int)x.size();
  std::vector<double> y2(n), u(n);
  if (natural0) {
    y2[0] = 0.0;
    u[0] = 0.0;
  } else {
    y2[0] = -0.5;
    u[0] = (3.0 / (x[1] - x[0])) * ((y[1] - y[0]) / (x[1] - x[0]) - yp1);
  }
  for (int i = 1; i < n - 1; ++i) {
    const double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
    const double p = sig * y2[i - 1] + 2.0;
    y2[i] = (sig - 1.0) / p;
    u[i] = (6.0 * ((y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1])) /
                (x[i + 1] - x[i - 1]) -
            sig * u[i - 1]) /
           p;
  }
  const double qn = 0.0, un = 0.0;
  y2[n - 1] = (un - qn * u[n - 2]) / (qn * y2[n - 2] + 1.0);
  for (int k = n - 2; k >= 0; --k) y2[k] = y2[k] * y2[k + 1] + u[k];
  return y2;
}

// pseudo/spline.cc splintd(): value and derivative (bisection for the segment)
__host__ __device__ inline void splintd(const double* xa, const double* ya, const double* y2a,
                                        int n, double x, double* y, double* dy) {
  int klo = 0, khi = n - 1;
  while (khi - klo > 1) {
    const int k = (khi + klo) >> 1;
    if (xa[k] > x)
      khi = k;
    else
      klo = k;
  }
  const double h = xa[khi] - xa[klo];
  const double a = (xa[khi] - x) / h;
  const double b = (x - xa[klo]) / h;
  *y = a * ya[klo] + b * ya[khi] +
       h * h * (1.0 / 6.0) * ((a * a * a - a) * y2a[klo] + (b * b * b - b) * y2a[khi]);
  *dy = (ya[khi] - ya[klo]) / h +
        h * (((1.0 / 6.0) - 0.5 * a * a) * y2a[klo] + (0.5 * b * b - (1.0 / 6.0)) * y2a[khi]);
}

// the text between <tag...> and </tag> after position `from`; npos when absent
size_t find_element(const std::string& s, const std::string& tag, size_t from, size_t* body,
                    size_t* body_end, std::string* open_tag = nullptr) {
  size_t p = from;
  while (true) {
    p = s.find("<" + tag, p);
    if (p == std::string::npos) return p;
    const char c = s[p + 1 + tag.size()];
    if (c == '>' || c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '/') break;
    p += 1;
  }
  const size_t gt = s.find('>', p);
  if (gt == std::string::npos) return std::string::npos;
  if (open_tag) *open_tag = s.substr(p, gt - p + 1);
  const size_t close = s.find("</" + tag + ">", gt);
  if (close == std::string::npos) return std::string::npos;
  *body = gt + 1;
  *body_end = close;
  return close + tag.size() + 3;
}

bool attr_int(const std::string& open, const char* name, int* v) {
  const std::string key = std::string(name) + "=\"";
  const size_t p = open.find(key);
  if (p == std::string::npos) return false;
  *v = std::atoi(open.c_str() + p + key.size());
  return true;
}

constexpr int kMaxL = 3;

}  // namespace

// host-side species (radial tables) with per-device copies of the spline tables
struct qbh_species_s {
  double zval = 0, rcps = 0, deltar = 0;
  int lmax = 0, llocal = 0, ndft = 0;
  std::vector<double> gspl;                 // G mesh of the splines
  std::vector<double> vloc_y, vloc_y2;      // Vloc(G)
  std::vector<double> proj_y[kMaxL + 1], proj_y2[kMaxL + 1];  // KB projectors per l
  double wsg[kMaxL + 1] = {0, 0, 0, 0};
  std::mutex mu;
  std::map<int, double*> dev;  // device -> [gspl | vloc y | vloc y2 | per l: y | y2]
};

namespace {

int sp_fail(std::string* err, const std::string& m) {
  if (err) *err = m;
  return QBH_ERR_ARG;
}

// Species::initialize, KB branch (Species.cc:74-673)
int species_init(qbh_species_s* s, double zval, int lmax, int llocal, double deltar, int npts,
                 const double* vps, const double* phi, double rcps, std::string* err) {
  if (!(zval > 0.0)) return sp_fail(err, "zval must be positive");
  if (lmax < 0 || lmax > kMaxL) return sp_fail(err, "lmax must be in 0..3");
  if (llocal < 0 || llocal > lmax) return sp_fail(err, "llocal must be in 0..lmax");
  if (!(deltar > 0.0) || npts < 8) return sp_fail(err, "mesh spacing or size invalid");
  if (!(rcps > 0.0)) return sp_fail(err, "rcps must be positive");
  s->zval = zval;
  s->lmax = lmax;
  s->llocal = llocal;
  s->deltar = deltar;
  s->rcps = rcps;
  // ndft: the second power of two above npts, then doubled until it reaches 40 bohr
  int ndft = 1;
  while (ndft < npts) ndft *= 2;
  ndft *= 2;
  while (deltar * ndft < 40.0) ndft *= 2;
  s->ndft = ndft;
  const double dr = deltar;
  std::vector<double> r(ndft);
  for (int i = 0; i < ndft; ++i) r[i] = i * dr;
  std::vector<std::vector<double>> v(lmax + 1, std::vector<double>(ndft)),
      f(lmax + 1, std::vector<double>(ndft, 0.0));
  for (int l = 0; l <= lmax; ++l) {
    for (int i = 0; i < npts; ++i) v[l][i] = vps[(long)l * npts + i];
    for (int i = npts; i < ndft; ++i) v[l][i] = -zval / r[i];
    if (l != llocal && phi)
      for (int i = 0; i < npts; ++i) f[l][i] = phi[(long)l * npts + i];
  }
  const double fpi = 4.0 * M_PI;
  // local potential minus the long-range part of the Gaussian ion charge
  const std::vector<double>& vl = v[llocal];
  std::vector<double> vlocr(ndft);
  vlocr[0] = vl[0] + (zval / rcps) * M_2_SQRTPI;
  for (int i = 1; i < ndft; ++i) vlocr[i] = vl[i] + (zval / r[i]) * std::erf(r[i] / rcps);
  for (int i = 0; i < ndft; ++i) vlocr[i] *= fpi * r[i] * dr;
  std::vector<double> fint(ndft);
  for (int i = 0; i < ndft; ++i) fint[i] = vlocr[i] * r[i];
  const double v0 = simpsn(fint);
  std::vector<double> vlocg = sine_transform(vlocr);
  s->gspl.assign(ndft, 0.0);
  const double fac = M_PI / (ndft * dr);
  for (int i = 1; i < ndft; ++i) {
    s->gspl[i] = i * fac;
    vlocg[i] /= s->gspl[i];
  }
  vlocg[0] = v0;
  s->vloc_y = vlocg;
  s->vloc_y2 = spline_fit(s->gspl, vlocg, false, 0.0);
  const std::vector<double>& g = s->gspl;
  for (int l = 0; l <= lmax; ++l) {
    s->wsg[l] = 0.0;
    if (l == llocal) continue;
    if (!phi) return sp_fail(err, "radial functions are needed for the KB projectors");
    for (int i = 0; i < ndft; ++i) {
      const double tmp = f[l][i] * r[i];
      fint[i] = (v[l][i] - vl[i]) * tmp * tmp;
    }
    const double w = simpsn(fint);
    if (w == 0.0) return sp_fail(err, "<phi|dV|phi> vanishes for a projector");
    s->wsg[l] = 1.0 / (dr * w);
    std::vector<double> fnr(ndft);
    for (int i = 0; i < ndft; ++i) fnr[i] = fpi * dr * (v[l][i] - vl[i]) * f[l][i];
    // Species::bessel_trans: fng(G) = sum_r r^2 j_l(G r) fnr(r)
    std::vector<double> fng(ndft, 0.0), y(ndft), yc(ndft + 1, 0.0);
    if (l == 0) {
      for (int i = 0; i < ndft; ++i) fint[i] = fnr[i] * r[i] * r[i];
      const double f0 = simpsn(fint);
      for (int i = 0; i < ndft; ++i) y[i] = fnr[i] * r[i];
      fng = sine_transform(y);
      fng[0] = f0;
      for (int i = 1; i < ndft; ++i) fng[i] /= g[i];
    } else if (l == 1) {
      const std::vector<double> s1 = sine_transform(fnr);
      for (int i = 0; i < ndft; ++i) yc[i] = fnr[i] * r[i];
      yc[ndft] = 0.0;
      const std::vector<double> c1 = cosine_transform(yc);
      fng[0] = 0.0;
      for (int i = 1; i < ndft; ++i) fng[i] = s1[i] / (g[i] * g[i]) - c1[i] / g[i];
    } else if (l == 2) {
      y[0] = 0.0;
      for (int i = 1; i < ndft; ++i) y[i] = fnr[i] / r[i];
      const std::vector<double> s1 = sine_transform(y);
      for (int i = 0; i < ndft; ++i) y[i] = fnr[i] * r[i];
      const std::vector<double> s2 = sine_transform(y);
      for (int i = 0; i < ndft; ++i) yc[i] = fnr[i];
      yc[ndft] = 0.0;
      const std::vector<double> c1 = cosine_transform(yc);
      for (int i = 1; i < ndft; ++i) {
        double t = 3.0 * s1[i] / (g[i] * g[i] * g[i]);
        t += -s2[i] / g[i];
        t += -3.0 * c1[i] / (g[i] * g[i]);
        fng[i] = t;
      }
    } else {
      y[0] = 0.0;
      for (int i = 1; i < ndft; ++i) y[i] = fnr[i] / (r[i] * r[i]);
      const std::vector<double> s1 = sine_transform(y);
      const std::vector<double> s2 = sine_transform(fnr);
      yc[0] = 0.0;
      for (int i = 1; i < ndft; ++i) yc[i] = fnr[i] / r[i];
      yc[ndft] = 0.0;
      const std::vector<double> c1 = cosine_transform(yc);
      for (int i = 0; i < ndft; ++i) yc[i] = fnr[i] * r[i];
      yc[ndft] = 0.0;
      const std::vector<double> c2 = cosine_transform(yc);
      for (int i = 1; i < ndft; ++i) {
        double t = 15.0 * s1[i] / (g[i] * g[i] * g[i] * g[i]);
        t += -6.0 * s2[i] / (g[i] * g[i]);
        t += -15.0 * c1[i] / (g[i] * g[i] * g[i]);
        t += c2[i] / g[i];
        fng[i] = t;
      }
    }
    s->proj_y[l] = fng;
    // flat at G = 0, except p projectors (natural)
    s->proj_y2[l] = spline_fit(g, fng, l == 1, 0.0);
  }
  return QBH_OK;
}

int nproj_of(const qbh_species_s* s) {
  int n = 0;
  for (int l = 0; l <= s->lmax; ++l)
    if (l != s->llocal) n += 2 * l + 1;
  return n;
}

// packed device tables of a species on the current device
double* species_dev(qbh_species_s* s, int device) {
  std::lock_guard<std::mutex> lock(s->mu);
  auto it = s->dev.find(device);
  if (it != s->dev.end()) return it->second;
  const int n = s->ndft;
  std::vector<double> h;
  h.insert(h.end(), s->gspl.begin(), s->gspl.end());
  h.insert(h.end(), s->vloc_y.begin(), s->vloc_y.end());
  h.insert(h.end(), s->vloc_y2.begin(), s->vloc_y2.end());
  for (int l = 0; l <= kMaxL; ++l) {
    if (s->proj_y[l].empty()) {
      h.insert(h.end(), 2 * (size_t)n, 0.0);
    } else {
      h.insert(h.end(), s->proj_y[l].begin(), s->proj_y[l].end());
      h.insert(h.end(), s->proj_y2[l].begin(), s->proj_y2[l].end());
    }
  }
  double* d = nullptr;
  if (hipMalloc(&d, h.size() * sizeof(double)) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    hipFree(d);
    return nullptr;
  }
  s->dev[device] = d;
  return d;
}

// ---------------------------------------------------------------------------------------
// update_twnl, KB branch: rows (2l+1 per l) x ngw at |k+G-A| (NonLocalPotential.cc:290-1000,
// the same intermediate quantities: tg, tgi, tgx..., tgxx..., y1..y15)
template <int L>
__global__ void __launch_bounds__(256)
k_twnl(int ngw, const double* __restrict__ kpg2, const double* __restrict__ kpgx, double ax,
       double ay, double az, const double* __restrict__ gspl, const double* __restrict__ py,
       const double* __restrict__ py2, int n, double* __restrict__ out) {
  const int ig = blockIdx.x * 256 + threadIdx.x;
  if (ig >= ngw) return;
  // VectorPotential::get_kpgpa2 / get_kpgpa / get_kpgpai / get_kpgpax (A = 0: the basis'
  // own |k+G|, 1/|k+G| and k+G)
  const double value2 = ax * ax + ay * ay + az * az;
  double k2 = kpg2[ig] + value2;
  k2 -= 2 * ax * kpgx[ig];
  k2 -= 2 * ay * kpgx[ngw + ig];
  k2 -= 2 * az * kpgx[2 * ngw + ig];
  const double tg = sqrt(k2);
  const double tgi = k2 > 0.0 ? 1.0 / sqrt(k2) : 0.0;
  double v, dv;
  if (tg > gspl[n - 1]) {
    v = 0.0;
  } else {
    splintd(gspl, py, py2, n, tg, &v, &dv);
  }
  const double fpi = 4.0 * M_PI;
  if (L == 0) {
    out[ig] = sqrt(1.0 / fpi) * v;
    return;
  }
  const double tgx = kpgx[ig] - ax, tgy = kpgx[ngw + ig] - ay, tgz = kpgx[2 * ngw + ig] - az;
  if (L == 1) {
    const double s34pi = sqrt(3.0 / fpi);
    out[ig] = s34pi * tgx * tgi * v;
    out[ngw + ig] = s34pi * tgy * tgi * v;
    out[2L * ngw + ig] = s34pi * tgz * tgi * v;
    return;
  }
  const double tgi2 = tgi * tgi;
  const double tgxx = tgx * tgx * tgi2, tgyy = tgy * tgy * tgi2, tgzz = tgz * tgz * tgi2;
  if (L == 2) {
    const double s54pi = sqrt(5.0 / fpi), s3 = sqrt(3.0);
    const double tgxy = tgx * tgy * tgi2, tgyz = tgy * tgz * tgi2, tgxz = tgx * tgz * tgi2;
    out[ig] = s54pi * 0.5 * (3.0 * tgzz - 1.0) * v;
    out[ngw + ig] = s54pi * 0.5 * s3 * (tgxx - tgyy) * v;
    out[2L * ngw + ig] = s54pi * s3 * tgxy * v;
    out[3L * ngw + ig] = s54pi * s3 * tgyz * v;
    out[4L * ngw + ig] = s54pi * s3 * tgxz * v;
    return;
  }
  const double tgi3 = tgi2 * tgi;
  const double s74pi = sqrt(7.0 / fpi), s2132pi = sqrt(21.0 / (32. * M_PI)),
               s3532pi = sqrt(35.0 / (32. * M_PI)), s1054pi = sqrt(105.0 / fpi);
  out[ig] = s74pi * 0.5 * tgz * tgi * (5.0 * tgzz - 3.0) * v;
  out[ngw + ig] = s2132pi * tgx * tgi * (5.0 * tgzz - 1.0) * v;
  out[2L * ngw + ig] = s2132pi * tgy * tgi * (5.0 * tgzz - 1.0) * v;
  out[3L * ngw + ig] = s1054pi * tgx * tgy * tgz * tgi3 * v;
  out[4L * ngw + ig] = s1054pi * 0.5 * tgz * tgi * (tgxx - tgyy) * v;
  out[5L * ngw + ig] = s3532pi * tgx * tgi * (tgxx - 3.0 * tgyy) * v;
  out[6L * ngw + ig] = s3532pi * tgy * tgi * (3.0 * tgxx - tgyy) * v;
}

// Angular factor m of angular momentum L as a polynomial Y(u) of the direction components
// (the factors of k_twnl, in its row order) and its partial derivatives with the components
// taken independent
template <int L>
__device__ __forceinline__ void harmonic(int m, double ux, double uy, double uz, double& Y,
                                         double& gx, double& gy, double& gz) {
  const double fpi = 4.0 * M_PI;
  gx = gy = gz = 0.0;
  if (L == 0) {
    Y = sqrt(1.0 / fpi);
  } else if (L == 1) {
    const double c = sqrt(3.0 / fpi);
    Y = c * (m == 0 ? ux : m == 1 ? uy : uz);
    (m == 0 ? gx : m == 1 ? gy : gz) = c;
  } else if (L == 2) {
    const double c = sqrt(5.0 / fpi), s3 = sqrt(3.0);
    switch (m) {
      case 0: Y = c * 0.5 * (3.0 * uz * uz - 1.0); gz = c * 3.0 * uz; break;
      case 1: Y = c * 0.5 * s3 * (ux * ux - uy * uy); gx = c * s3 * ux; gy = -c * s3 * uy; break;
      case 2: Y = c * s3 * ux * uy; gx = c * s3 * uy; gy = c * s3 * ux; break;
      case 3: Y = c * s3 * uy * uz; gy = c * s3 * uz; gz = c * s3 * uy; break;
      default: Y = c * s3 * uz * ux; gx = c * s3 * uz; gz = c * s3 * ux; break;
    }
  } else {
    const double a = sqrt(7.0 / fpi), b = sqrt(21.0 / (32. * M_PI)),
                 c = sqrt(35.0 / (32. * M_PI)), d = sqrt(105.0 / fpi);
    const double z5 = 5.0 * uz * uz - 1.0;
    switch (m) {
      case 0: Y = a * 0.5 * uz * (5.0 * uz * uz - 3.0); gz = a * 0.5 * (15.0 * uz * uz - 3.0); break;
      case 1: Y = b * ux * z5; gx = b * z5; gz = b * 10.0 * ux * uz; break;
      case 2: Y = b * uy * z5; gy = b * z5; gz = b * 10.0 * uy * uz; break;
      case 3: Y = d * ux * uy * uz; gx = d * uy * uz; gy = d * ux * uz; gz = d * ux * uy; break;
      case 4:
        Y = d * 0.5 * uz * (ux * ux - uy * uy);
        gx = d * uz * ux; gy = -d * uz * uy; gz = d * 0.5 * (ux * ux - uy * uy);
        break;
      case 5:
        Y = c * ux * (ux * ux - 3.0 * uy * uy);
        gx = c * 3.0 * (ux * ux - uy * uy); gy = -c * 6.0 * ux * uy;
        break;
      default:
        Y = c * uy * (3.0 * ux * ux - uy * uy);
        gx = c * 6.0 * ux * uy; gy = c * 3.0 * (ux * ux - uy * uy);
        break;
    }
  }
}

// Gradient of the 2L+1 rows of k_twnl with respect to q = k+G-A: out[d * dstride + m * ngw
// + ig] = d/dq_d [v(|q|) Y_m(q/|q|)] = v' u_d Y + (v / |q|)(dY/du_d - u_d u.grad Y), zero at
// q = 0 (where k_twnl's direction factors vanish)
template <int L>
__global__ void __launch_bounds__(256)
k_dtwnl(int ngw, const double* __restrict__ kpg2, const double* __restrict__ kpgx, double ax,
        double ay, double az, const double* __restrict__ gspl, const double* __restrict__ py,
        const double* __restrict__ py2, int n, double* __restrict__ out, long dstride) {
  const int ig = blockIdx.x * 256 + threadIdx.x;
  if (ig >= ngw) return;
  const double value2 = ax * ax + ay * ay + az * az;
  double k2 = kpg2[ig] + value2;
  k2 -= 2 * ax * kpgx[ig];
  k2 -= 2 * ay * kpgx[ngw + ig];
  k2 -= 2 * az * kpgx[2 * ngw + ig];
  const double tg = sqrt(k2);
  const double tgi = k2 > 0.0 ? 1.0 / sqrt(k2) : 0.0;
  double v = 0.0, dv = 0.0;
  if (tg <= gspl[n - 1]) splintd(gspl, py, py2, n, tg, &v, &dv);
  const double u[3] = {(kpgx[ig] - ax) * tgi, (kpgx[ngw + ig] - ay) * tgi,
                       (kpgx[2 * ngw + ig] - az) * tgi};
  for (int m = 0; m < 2 * L + 1; ++m) {
    double Y, g[3];
    harmonic<L>(m, u[0], u[1], u[2], Y, g[0], g[1], g[2]);
    const double ug = u[0] * g[0] + u[1] * g[1] + u[2] * g[2];
    for (int d = 0; d < 3; ++d)
      out[d * dstride + (long)m * ngw + ig] =
          tgi > 0.0 ? dv * u[d] * Y + v * tgi * (g[d] - u[d] * ug) : 0.0;
  }
}

// v_ion(G) += S(G) vlocg(|G|) / omega, rho_ps(G) += S(G) rhopsg(|G|) / omega for one
// species, S(G) = sum_a exp(-i G.tau_a) (StructureFactor::update), complex interleaved
__global__ void __launch_bounds__(256)
k_local(int ngv, const double* __restrict__ kpg2, const double* __restrict__ kpgx, int na,
        const double* __restrict__ tau, const double* __restrict__ gspl,
        const double* __restrict__ vy, const double* __restrict__ vy2, int n, double zval,
        double rcps, double omega_inv, double* __restrict__ vion, double* __restrict__ rhops) {
  const int ig = blockIdx.x * 256 + threadIdx.x;
  if (ig >= ngv) return;
  const double gx = kpgx[ig], gy = kpgx[ngv + ig], gz = kpgx[2 * ngv + ig];
  double sr = 0.0, si = 0.0;
  for (int a = 0; a < na; ++a) {
    const double arg = gx * tau[3 * a] + gy * tau[3 * a + 1] + gz * tau[3 * a + 2];
    double s, c;
    sincos(arg, &s, &c);
    sr += c;
    si -= s;
  }
  const double g = sqrt(kpg2[ig]);
  double v = 0.0, dv;
  if (g <= gspl[n - 1]) splintd(gspl, vy, vy2, n, g, &v, &dv);
  const double vv = v * omega_inv;
  const double rr = (-zval * exp(-0.25 * rcps * rcps * g * g)) * omega_inv;
  vion[2 * ig] += sr * vv;
  vion[2 * ig + 1] += si * vv;
  rhops[2 * ig] += sr * rr;
  rhops[2 * ig + 1] += si * rr;
}

__global__ void k_zero2(double* p, long n) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// the 13 lattice translations whose bisecting planes bound the Wigner-Seitz cell, and
// half their squared lengths (UnitCell::set, UnitCell.cc:108-138); a holds the rows a0..a2
void ws_planes(const double* a, double an[13][3], double an2h[13]) {
  static const int c[13][3] = {{1, 0, 0},  {0, 1, 0},  {0, 0, 1},  {1, 1, 0},   {1, -1, 0},
                               {0, 1, 1},  {0, 1, -1}, {1, 0, 1},  {-1, 0, 1},  {1, 1, 1},
                               {1, -1, -1}, {1, 1, -1}, {1, -1, 1}};
  for (int i = 0; i < 13; ++i) {
    for (int d = 0; d < 3; ++d) an[i][d] = c[i][0] * a[d] + c[i][1] * a[3 + d] + c[i][2] * a[6 + d];
    an2h[i] = 0.5 * (an[i][0] * an[i][0] + an[i][1] * an[i][1] + an[i][2] * an[i][2]);
  }
}

// UnitCell::min_wsdist (UnitCell.cc:241-248): the shortest of the 13 translations
double min_wsdist(const double* a) {
  double an[13][3], an2h[13];
  ws_planes(a, an, an2h);
  double m = std::sqrt(2. * an2h[0]);
  for (int i = 1; i < 13; ++i) m = std::min(m, std::sqrt(2. * an2h[i]));
  return m;
}

// UnitCell::fold_in_ws (UnitCell.cc:251-281): shift v across the bisecting planes until it
// lies in the Wigner-Seitz cell (tolerance 1e-10, at most 10 sweeps)
void fold_in_ws(const double* a, double* v) {
  double an[13][3], an2h[13];
  ws_planes(a, an, an2h);
  const double epsilon = 1.e-10;
  auto dot = [&](int i) { return v[0] * an[i][0] + v[1] * an[i][1] + v[2] * an[i][2]; };
  bool done = false;
  for (int iter = 0; !done && iter < 10; ++iter) {
    done = true;
    for (int i = 0; i < 13 && done; ++i) {
      const double sp = dot(i);
      if (sp > an2h[i] + epsilon) {
        done = false;
        do {
          for (int d = 0; d < 3; ++d) v[d] -= an[i][d];
        } while (dot(i) > an2h[i] + epsilon);
      } else if (sp < -an2h[i] - epsilon) {
        done = false;
        do {
          for (int d = 0; d < 3; ++d) v[d] += an[i][d];
        } while (dot(i) < -an2h[i] - epsilon);
      }
    }
  }
}

}  // namespace

// =======================================================================================
extern "C" int qbh_species_create(double zval, int lmax, int llocal, double deltar, int npts,
                                  const double* vps, const double* phi, double rcps,
                                  qbh_species* out) {
  if (!out || !vps) return QBH_ERR_ARG;
  *out = nullptr;
  auto* s = new qbh_species_s();
  std::string err;
  if (int r = species_init(s, zval, lmax, llocal, deltar, npts, vps, phi, rcps, &err)) {
    delete s;
    return r;
  }
  *out = s;
  return QBH_OK;
}

// the <norm_conserving_pseudopotential> element of a species XML file (the fields the
// reference's SpeciesReader reads for a KB pseudopotential)
extern "C" int qbh_species_create_xml(const char* path, double rcps, qbh_species* out) {
  if (!path || !out) return QBH_ERR_ARG;
  *out = nullptr;
  std::ifstream f(path);
  if (!f) return QBH_ERR_ARG;
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string s = ss.str();
  size_t b, e;
  const size_t nc = find_element(s, "norm_conserving_pseudopotential", 0, &b, &e);
  if (nc == std::string::npos) return QBH_ERR_UNSUPPORTED;
  const std::string body = s.substr(b, e - b);
  auto scalar = [&](const char* tag, double* v) {
    size_t bb, ee;
    if (find_element(body, tag, 0, &bb, &ee) == std::string::npos) return false;
    char* endp = nullptr;
    const std::string t = body.substr(bb, ee - bb);
    *v = std::strtod(t.c_str(), &endp);
    return endp != t.c_str();
  };
  double zval, lmax_d, llocal_d, nquad_d, dr;
  if (!scalar("valence_charge", &zval) || !scalar("lmax", &lmax_d) ||
      !scalar("llocal", &llocal_d) || !scalar("mesh_spacing", &dr))
    return QBH_ERR_ARG;
  if (scalar("nquad", &nquad_d) && nquad_d != 0.0) return QBH_ERR_UNSUPPORTED;  // semilocal
  const int lmax = (int)lmax_d, llocal = (int)llocal_d;
  if (lmax < 0 || lmax > kMaxL) return QBH_ERR_UNSUPPORTED;
  int npts = -1;
  std::vector<std::vector<double>> vps(lmax + 1), phi(lmax + 1);
  size_t pos = 0;
  while (true) {
    std::string open;
    size_t pb, pe;
    const size_t next = find_element(body, "projector", pos, &pb, &pe, &open);
    if (next == std::string::npos) break;
    pos = next;
    int l = -1, size = -1;
    if (!attr_int(open, "l", &l) || !attr_int(open, "size", &size) || l < 0 || l > lmax ||
        size < 8)
      return QBH_ERR_ARG;
    if (npts < 0) npts = size;
    if (size != npts) return QBH_ERR_ARG;
    const std::string proj = body.substr(pb, pe - pb);
    auto numbers = [&](const char* tag, std::vector<double>* v) {
      size_t bb, ee;
      if (find_element(proj, tag, 0, &bb, &ee) == std::string::npos) return false;
      std::istringstream is(proj.substr(bb, ee - bb));
      double x;
      while (is >> x) v->push_back(x);
      return (int)v->size() == size;
    };
    if (!numbers("radial_potential", &vps[l])) return QBH_ERR_ARG;
    if (l != llocal && !numbers("radial_function", &phi[l])) return QBH_ERR_ARG;
  }
  if (npts < 0) return QBH_ERR_ARG;
  std::vector<double> V((size_t)(lmax + 1) * npts, 0.0), P((size_t)(lmax + 1) * npts, 0.0);
  for (int l = 0; l <= lmax; ++l) {
    if ((int)vps[l].size() != npts) return QBH_ERR_ARG;
    std::copy(vps[l].begin(), vps[l].end(), V.begin() + (size_t)l * npts);
    if (!phi[l].empty()) std::copy(phi[l].begin(), phi[l].end(), P.begin() + (size_t)l * npts);
  }
  return qbh_species_create(zval, lmax, llocal, dr, npts, V.data(), P.data(), rcps, out);
}

extern "C" int qbh_species_destroy(qbh_species sp) {
  if (!sp) return QBH_ERR_ARG;
  for (auto& kv : sp->dev) {
    hipSetDevice(kv.first);
    hipFree(kv.second);
  }
  delete sp;
  return QBH_OK;
}

extern "C" int qbh_species_info(qbh_species sp, double* zval, double* rcps, double* eself,
                                int* nproj, int* lproj, double* wt) {
  if (!sp) return QBH_ERR_ARG;
  if (zval) *zval = sp->zval;
  if (rcps) *rcps = sp->rcps;
  if (eself) *eself = sp->zval * sp->zval / (std::sqrt(2.0 * M_PI) * sp->rcps);
  if (nproj) *nproj = nproj_of(sp);
  int i = 0;
  for (int l = 0; l <= sp->lmax; ++l) {
    if (l == sp->llocal) continue;
    for (int m = 0; m < 2 * l + 1; ++m, ++i) {
      if (lproj) lproj[i] = l;
      if (wt) wt[i] = sp->wsg[l];
    }
  }
  return QBH_OK;
}

extern "C" int qbh_species_eval(qbh_species sp, int which, int n, const double* g, double* v,
                                double* dv) {
  if (!sp || !g || !v || n < 0 || which < -1 || which > sp->lmax) return QBH_ERR_ARG;
  const bool local = which < 0;
  if (!local && (which == sp->llocal)) {
    for (int i = 0; i < n; ++i) {
      v[i] = 0.0;
      if (dv) dv[i] = 0.0;
    }
    return QBH_OK;
  }
  const std::vector<double>& y = local ? sp->vloc_y : sp->proj_y[which];
  const std::vector<double>& y2 = local ? sp->vloc_y2 : sp->proj_y2[which];
  for (int i = 0; i < n; ++i) {
    double a = 0.0, b = 0.0;
    if (g[i] <= sp->gspl[sp->ndft - 1])
      splintd(sp->gspl.data(), y.data(), y2.data(), sp->ndft, g[i], &a, &b);
    v[i] = a;
    if (dv) dv[i] = b;
  }
  return QBH_OK;
}

// twnl rows of species sp at |k+G-A| on basis b (grad 0: rows x ngw into out), or their
// gradient with respect to k+G-A (grad 1: out + d * dstride, d = x, y, z), on c's stream
int qbh::species_table(qbh_species sp, Ctx* c, Basis* b, const double* A, int grad, double* out,
                       long dstride) {
  const double* tab = species_dev(sp, c->device);
  if (!tab) {
    c->last_error = "species table allocation failed";
    return QBH_ERR_ALLOC;
  }
  const int n = sp->ndft, ngw = b->g.ngw;
  const double ax = A ? A[0] : 0.0, ay = A ? A[1] : 0.0, az = A ? A[2] : 0.0;
  const dim3 grid((unsigned)((ngw + 255) / 256));
  long row = 0;
  for (int l = 0; l <= sp->lmax; ++l) {
    if (l == sp->llocal) continue;
    const double* py = tab + 3L * n + 2L * n * l;
    const double* py2 = py + n;
    double* o = out + row * ngw;
    hipStream_t st = c->stream;
    if (grad) {
      switch (l) {
        case 0: k_dtwnl<0><<<grid, 256, 0, st>>>(ngw, b->kpg2, b->kpgx, ax, ay, az, tab, py, py2, n, o, dstride); break;
        case 1: k_dtwnl<1><<<grid, 256, 0, st>>>(ngw, b->kpg2, b->kpgx, ax, ay, az, tab, py, py2, n, o, dstride); break;
        case 2: k_dtwnl<2><<<grid, 256, 0, st>>>(ngw, b->kpg2, b->kpgx, ax, ay, az, tab, py, py2, n, o, dstride); break;
        default: k_dtwnl<3><<<grid, 256, 0, st>>>(ngw, b->kpg2, b->kpgx, ax, ay, az, tab, py, py2, n, o, dstride); break;
      }
    } else {
      switch (l) {
        case 0: k_twnl<0><<<grid, 256, 0, st>>>(ngw, b->kpg2, b->kpgx, ax, ay, az, tab, py, py2, n, o); break;
        case 1: k_twnl<1><<<grid, 256, 0, st>>>(ngw, b->kpg2, b->kpgx, ax, ay, az, tab, py, py2, n, o); break;
        case 2: k_twnl<2><<<grid, 256, 0, st>>>(ngw, b->kpg2, b->kpgx, ax, ay, az, tab, py, py2, n, o); break;
        default: k_twnl<3><<<grid, 256, 0, st>>>(ngw, b->kpg2, b->kpgx, ax, ay, az, tab, py, py2, n, o); break;
      }
    }
    row += 2 * l + 1;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    c->last_error = std::string("species twnl: ") + hipGetErrorString(e);
    return QBH_ERR_DEVICE;
  }
  return QBH_OK;
}

int qbh::species_nproj(qbh_species sp) { return nproj_of(sp); }

extern "C" int qbh_species_twnl(qbh_species sp, qbh_basis basis, const double* A,
                                double* twnl_dev) {
  if (!sp || !basis || !twnl_dev) return QBH_ERR_ARG;
  Basis* b = basis_of(basis);
  Ctx* c = ctx_of_basis(basis);
  hipSetDevice(c->device);
  return species_table(sp, c, b, A, 0, twnl_dev, 0);
}

extern "C" int qbh_species_dtwnl(qbh_species sp, qbh_basis basis, const double* A,
                                 double* dtwnl_dev) {
  if (!sp || !basis || !dtwnl_dev) return QBH_ERR_ARG;
  Basis* b = basis_of(basis);
  Ctx* c = ctx_of_basis(basis);
  hipSetDevice(c->device);
  return species_table(sp, c, b, A, 1, dtwnl_dev, (long)nproj_of(sp) * b->g.ngw);
}

extern "C" int qbh_local_tables(qbh_basis vbasis, int nsp, const qbh_species* sps, const int* na,
                                const double* tau_dev, double omega, double* vion_dev,
                                double* rhops_dev) {
  if (!vbasis || nsp < 0 || (nsp > 0 && (!sps || !na || !tau_dev)) || !vion_dev || !rhops_dev ||
      !(omega > 0.0))
    return QBH_ERR_ARG;
  Basis* b = basis_of(vbasis);
  Ctx* c = ctx_of_basis(vbasis);
  hipSetDevice(c->device);
  const int ngv = b->g.ngw;
  const long n2 = 2L * ngv;
  k_zero2<<<(unsigned)((n2 + 255) / 256), 256, 0, c->stream>>>(vion_dev, n2);
  k_zero2<<<(unsigned)((n2 + 255) / 256), 256, 0, c->stream>>>(rhops_dev, n2);
  long atom0 = 0;
  for (int s = 0; s < nsp; ++s) {
    if (!sps[s] || na[s] < 0) return QBH_ERR_ARG;
    const double* tab = species_dev(sps[s], c->device);
    if (!tab) {
      c->last_error = "species table allocation failed";
      return QBH_ERR_ALLOC;
    }
    const int n = sps[s]->ndft;
    k_local<<<(unsigned)((ngv + 255) / 256), 256, 0, c->stream>>>(
        ngv, b->kpg2, b->kpgx, na[s], tau_dev + 3 * atom0, tab, tab + n, tab + 2L * n, n,
        sps[s]->zval, sps[s]->rcps, 1.0 / omega, vion_dev, rhops_dev);
    atom0 += na[s];
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    c->last_error = std::string("local tables: ") + hipGetErrorString(e);
    return QBH_ERR_DEVICE;
  }
  return QBH_OK;
}

// AtomSet::set_rcps (AtomSet.cc:1102-1163): from 0.3 in steps of 0.05 until the Gaussian
// charge's reciprocal tail (zval / sqrt(pi)) erfc(rcps sqrt(ecut)) is at most 1e-7 rcps /
// zval, capped at min_wsdist / (2 pi), then down in steps of 0.05 until the charge outside
// the box zval erfc(min_wsdist / (sqrt(2) rcps)) is at most 1e-8
extern "C" int qbh_species_auto_rcps(double zval, const double* a, double ecut, double* rcps) {
  if (!a || !rcps || !(zval > 0.0) || !(ecut > 0.0)) return QBH_ERR_ARG;
  const double ellmin = min_wsdist(a);
  if (!(ellmin > 0.0)) return QBH_ERR_ARG;
  double r = 0.3;
  const double recipmax = ellmin / (1.0 * 2. * M_PI);
  while ((zval / sqrt(M_PI)) * erfc(r * sqrt(ecut)) > 1.e-7 * r / zval) r += 0.05;
  if (r > recipmax) r = recipmax;
  while (zval * erfc(ellmin / (sqrt(2.) * r)) > 1.e-8) r -= 0.05;
  *rcps = r;
  return QBH_OK;
}

extern "C" int qbh_ewald_esr(const double* a, int nsp, const qbh_species* sps, const int* na,
                             const double* tau, double* esr) {
  if (!a || !esr || nsp < 0 || (nsp > 0 && (!sps || !na || !tau))) return QBH_ERR_ARG;
  const double wsd = min_wsdist(a);
  if (!(wsd > 0.0)) return QBH_ERR_ARG;
  std::vector<long> first(nsp + 1, 0);
  for (int s = 0; s < nsp; ++s) {
    if (!sps[s] || na[s] < 0) return QBH_ERR_ARG;
    first[s + 1] = first[s] + na[s];
  }
  double e = 0.0;
  for (int k = 0; k < nsp; ++k) {
    for (int j = k; j < nsp; ++j) {
      const double rk = sps[k]->rcps, rj = sps[j]->rcps;
      const double rckj = std::sqrt(rk * rk + rj * rj);
      const int newald = (int)((0.5 * 5 * rckj / wsd) + 0.5);
      const double zz = sps[k]->zval * sps[j]->zval;
      const int lax = na[k] - (k == j ? 1 : 0);
      for (int l = 0; l < lax; ++l) {
        const int inf = k == j ? l + 1 : 0;
        for (int m = inf; m < na[j]; ++m) {
          double v0[3];
          for (int d = 0; d < 3; ++d)
            v0[d] = tau[3 * (first[k] + l) + d] - tau[3 * (first[j] + m) + d];
          fold_in_ws(a, v0);
          for (int ux = -newald; ux <= newald; ++ux)
            for (int uy = -newald; uy <= newald; ++uy)
              for (int uz = -newald; uz <= newald; ++uz) {
                double v[3];
                for (int d = 0; d < 3; ++d) v[d] = v0[d] + ux * a[d] + uy * a[3 + d] + uz * a[6 + d];
                const double rlm = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                e += zz * std::erfc(rlm / rckj) / rlm;
              }
        }
      }
    }
  }
  *esr = e;
  return QBH_OK;
}
