// Host-side plane-wave sphere construction for one process row.
//
// Restates qball's Basis::resize (src/qball/Basis.cc:302-700) with nprow = 1 and
// UnitCell::set (src/qball/UnitCell.cc:36-90): rods (h, k, lmin, size) of the sphere
// |k+G|^2 < 2 ecut, the Γ half sphere (rod (0,0,l>=0), rods (0,k>0), rods (h>0, k)),
// rod order = ascending size with ties in insertion order (std::multiset<Rod> ordered
// by size, Basis.cc:190-193), rod (0,0) swapped to the front (:646-650), and the np
// rule 2*hmax+2 bumped by 2 to a factorizable length (:563-574).  Floating-point
// expressions keep the reference's evaluation order; this file is built without FP
// contraction so the sphere membership test rounds as the reference's does.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/qbh.h"

namespace {

struct V3 {
  double x, y, z;
};
inline V3 operator*(double a, const V3& v) { return {a * v.x, a * v.y, a * v.z}; }
inline V3 operator+(const V3& a, const V3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 cross(const V3& a, const V3& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm(const V3& a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline double length(const V3& a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

struct Rod {
  int h, k, lmin, size;
};

}  // namespace

extern "C" int qbh_factorizable(int n) {
  if (n <= 0) return 0;
  if (n % 11 == 0) n /= 11;
  if (n % 7 == 0) n /= 7;
  if (n % 5 == 0) n /= 5;
  if (n % 3 == 0) n /= 3;
  if (n % 3 == 0) n /= 3;
  while (n % 2 == 0) n /= 2;
  return n == 1;
}

extern "C" int qbh_rods_build(const double* a, double ecut, int real, const double* kpoint,
                              int max_rods, int* nrods_out, int* rod_h, int* rod_k,
                              int* rod_lmin, int* rod_size, int* np) {
  if (!a || !nrods_out || ecut <= 0.0) return QBH_ERR_ARG;
  const V3 a0{a[0], a[1], a[2]}, a1{a[3], a[4], a[5]}, a2{a[6], a[7], a[8]};
  const double volume = dot(a0, cross(a1, a2));
  if (!(volume > 0.0)) return QBH_ERR_ARG;
  const double fac_v = 1.0 / volume;
  const V3 amt0 = cross(fac_v * a1, a2);
  const V3 amt1 = cross(fac_v * a2, a0);
  const V3 amt2 = cross(fac_v * a0, a1);
  const V3 b0 = (2.0 * M_PI) * amt0, b1 = (2.0 * M_PI) * amt1, b2 = (2.0 * M_PI) * amt2;
  const double kpx = kpoint ? kpoint[0] : 0.0, kpy = kpoint ? kpoint[1] : 0.0,
               kpz = kpoint ? kpoint[2] : 0.0;
  const bool isreal = real && kpx == 0.0 && kpy == 0.0 && kpz == 0.0;

  const double two_ecut = 2.0 * ecut;
  const double twopi = 2.0 * M_PI;
  const double normb2 = norm(b2);
  const double b2inv2 = 1.0 / normb2;
  const double fac = std::sqrt(two_ecut) / twopi;
  const int hmax = (int)(0.5 + fac * length(a0));
  const int hmin = -hmax;
  const int kmax = (int)(0.5 + fac * length(a1));
  const int kmin = -kmax;
  const int lmax = (int)(0.5 + fac * length(a2));
  const int lmin = -lmax;

  std::vector<Rod> rods;
  auto scan = [&](double h, double k, double sl, int& ls, int& le) {
    ls = lmax;
    le = lmin;
    bool found = false;
    for (int l = lmin - 1; l <= lmax + 1; l++) {
      const V3 v = (h * b0 + k * b1) + (sl + l) * b2;
      if (norm(v) < two_ecut) {
        ls = std::min(l, ls);
        le = std::max(l, le);
        found = true;
      }
    }
    return found;
  };
  if (isreal) {
    const int lend = (int)(std::sqrt(two_ecut * b2inv2));
    rods.push_back({0, 0, 0, lend + 1});
    for (int k = 1; k <= kmax + 1; k++) {
      int ls, le;
      if (scan(0.0, (double)k, 0.0, ls, le)) rods.push_back({0, k, ls, le - ls + 1});
    }
    for (int h = 1; h <= hmax + 1; h++)
      for (int k = kmin - 1; k <= kmax + 1; k++) {
        int ls, le;
        if (scan((double)h, (double)k, 0.0, ls, le)) rods.push_back({h, k, ls, le - ls + 1});
      }
  } else {
    for (int h = hmin - 1; h <= hmax + 1; h++)
      for (int k = kmin - 1; k <= kmax + 1; k++) {
        int ls, le;
        if (scan(kpx + h, kpy + k, kpz, ls, le)) rods.push_back({h, k, ls, le - ls + 1});
      }
  }
  // multiset<Rod> by size: stable ordering of equal sizes
  std::stable_sort(rods.begin(), rods.end(),
                   [](const Rod& x, const Rod& y) { return x.size < y.size; });
  for (size_t i = 0; i < rods.size(); ++i)
    if (rods[i].h == 0 && rods[i].k == 0) {
      std::swap(rods[0], rods[i]);
      break;
    }
  if (np) {
    const int m[3] = {hmax, kmax, lmax};
    for (int i = 0; i < 3; ++i) {
      int n = 2 * m[i] + 2;
      while (!qbh_factorizable(n)) n += 2;
      np[i] = n;
    }
  }
  *nrods_out = (int)rods.size();
  if (max_rods > 0) {
    if (max_rods < (int)rods.size()) return QBH_ERR_ARG;
    for (size_t i = 0; i < rods.size(); ++i) {
      rod_h[i] = rods[i].h;
      rod_k[i] = rods[i].k;
      rod_lmin[i] = rods[i].lmin;
      rod_size[i] = rods[i].size;
    }
  }
  return QBH_OK;
}
