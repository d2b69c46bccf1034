// G-sphere <-> grid transform pipeline on gfx950.
//
// Behaviour follows qball's FourierTransform (src/qball/FourierTransform.cc): backward
// = unnormalised exp(+iG.r) sum of the sphere coefficients (Γ: c(G) at +G and conj(c(G))
// at -G, the -G value winning at G=0, :1624-1720); forward = exp(-iG.r)/N (:1338-1342)
// followed by the Γ pair unpacking c1 = (z+ + conj z-)/2, c2 = (z+ - conj z-)/(2i)
// (:1723-1752).  The decomposition is this library's own: a z pass over the nonzero
// columns only, a y pass over the x-lines that hold sphere points, and an x pass that
// owns whole real-space rows, so v(r)*psi(r) and |psi(r)|^2 are applied while a row is
// in LDS and the real-space grid never goes to HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

#include "qbh_internal.h"

namespace qbh {

constexpr int NTHR = 256;  // threads per workgroup, every pipeline kernel
constexpr int EPT = 16;    // lines * n <= EPT * NTHR in every LDS transform
constexpr int ZL = 16;     // max lines (columns) per z-pass group
constexpr int YB = 16;     // real-space rows per x-pass workgroup

template <typename T> struct ZBlk;
template <> struct ZBlk<double> { static constexpr int ZB = 8, LB = 2; };   // 128 B segments
template <> struct ZBlk<float> { static constexpr int ZB = 16, LB = 1; };   // 128 B segments

template <typename T>
__device__ __forceinline__ const cplx<T>* twtab(const GridDev& g, int axis);
template <>
__device__ __forceinline__ const cplx<double>* twtab<double>(const GridDev& g, int axis) {
  return g.tw64[axis];
}
template <>
__device__ __forceinline__ const cplx<float>* twtab<float>(const GridDev& g, int axis) {
  return g.tw32[axis];
}

__device__ __forceinline__ int wrap(int l, int n) { return l < 0 ? l + n : l; }

// ------------------------------------------------------------------------------------
// z pass, backward: gather sphere coefficients of one slot into its columns, transform
// along z (exp(+i)), store T1[slot][col][z].
template <typename T, int MODE>
__global__ void __launch_bounds__(NTHR)
k_zpass_bwd(GridDev g, const double* __restrict__ psi, long ld, int nst, int band0,
            cplx<T>* __restrict__ t1) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int np2 = g.np2;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + np2;
  const int grp = blockIdx.x;
  const int slot = blockIdx.y;
  const int e0 = g.grp_start[grp];
  const int nl = g.grp_start[grp + 1] - e0;
  const cplx<T>* twg = twtab<T>(g, 2);
  for (int i = threadIdx.x; i < np2; i += NTHR) tw[i] = twg[i];
  for (int i = threadIdx.x; i < nl * np2; i += NTHR) buf[i] = {T(0), T(0)};
  __syncthreads();

  int n1, n2;
  if (MODE == SLOT_PAIR) {
    n1 = band0 + 2 * slot;
    n2 = n1 + 1;
  } else {
    n1 = band0 + slot;
    n2 = nst;
  }
  const bool has2 = (MODE == SLOT_PAIR) && (n2 < nst);
  const double* p1 = psi + 2 * (long)n1 * ld;
  const double* p2 = psi + 2 * (long)(has2 ? n2 : n1) * ld;

  // 16 threads per line
  const int le = threadIdx.x >> 4;
  const int sub = threadIdx.x & 15;
  if (le < nl) {
    const int e = e0 + le;
    const int rod = g.ent_rod[e];
    const int kind = g.ent_kind[e];
    const int first = g.rod_first[rod];
    const int lmin = g.rod_lmin[rod];
    const int size = g.rod_size[rod];
    cplx<T>* line = buf + le * np2;
    for (int i = sub; i < size; i += 16) {
      const int l = lmin + i;
      const double a = p1[2 * (first + i)];
      const double b = p1[2 * (first + i) + 1];
      double c = 0.0, d = 0.0;
      if (has2) {
        c = p2[2 * (first + i)];
        d = p2[2 * (first + i) + 1];
      }
      double pr, pi, mr, mi;  // +G and -G values
      if (MODE == SLOT_PAIR) {
        pr = a - d; pi = b + c; mr = a + d; mi = c - b;
      } else {
        pr = a; pi = b; mr = a; mi = -b;
      }
      if (kind == 0) {
        if (i > 0) line[wrap(l, np2)] = {T(pr), T(pi)};
        line[wrap(-l, np2)] = {T(mr), T(mi)};
      } else if (kind == 1 || kind == 3) {
        line[wrap(l, np2)] = {T(pr), T(pi)};
      } else {
        line[wrap(-l, np2)] = {T(mr), T(mi)};
      }
    }
  }
  __syncthreads();
  fft_lines<+1, EPT>(buf, nl, np2, g.ax[2], tw);
  for (int i = threadIdx.x; i < nl * np2; i += NTHR) {
    const int ll = i / np2;
    const int z = i - ll * np2;
    const int col = g.ent_col[e0 + ll];
    t1[((long)slot * g.ncol + col) * np2 + z] = buf[i];
  }
}

// ------------------------------------------------------------------------------------
// y pass, backward: T1 columns of ZB planes x LB x-lines -> y transform -> T2 lines.
template <typename T>
__global__ void __launch_bounds__(NTHR)
k_ypass_bwd(GridDev g, const cplx<T>* __restrict__ t1, cplx<T>* __restrict__ t2) {
  constexpr int ZB = ZBlk<T>::ZB, LB = ZBlk<T>::LB;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int np1 = g.np1, np2 = g.np2, nlx = g.nlx;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + np1;
  const int nzb = (np2 + ZB - 1) / ZB;
  const int zb = blockIdx.x % nzb;
  const int lb = blockIdx.x / nzb;
  const int slot = blockIdx.y;
  const int z0 = zb * ZB, lx0 = lb * LB;
  const int lx1 = min(lx0 + LB, nlx);
  const cplx<T>* twg = twtab<T>(g, 1);
  for (int i = threadIdx.x; i < np1; i += NTHR) tw[i] = twg[i];
  for (int i = threadIdx.x; i < ZB * LB * np1; i += NTHR) buf[i] = {T(0), T(0)};
  __syncthreads();
  const int c0 = g.line_start[lx0], c1 = g.line_start[lx1];
  const int nwork = (c1 - c0) * ZB;
  for (int f = threadIdx.x; f < nwork; f += NTHR) {
    const int dz = f % ZB;
    const int c = c0 + f / ZB;
    const int z = z0 + dz;
    if (z < np2) {
      // line of column c: find by scanning the (<= LB) line starts
      int dl = 0;
      while (lx0 + dl + 1 < lx1 && c >= g.line_start[lx0 + dl + 1]) ++dl;
      buf[(dz + ZB * dl) * np1 + g.col_y[c]] = t1[((long)slot * g.ncol + c) * np2 + z];
    }
  }
  __syncthreads();
  fft_lines<+1, EPT>(buf, ZB * LB, np1, g.ax[1], tw);
  for (int i = threadIdx.x; i < ZB * LB * np1; i += NTHR) {
    const int li = i / np1;
    const int y = i - li * np1;
    const int dz = li % ZB, dl = li / ZB;
    const int z = z0 + dz, lx = lx0 + dl;
    if (z < np2 && lx < lx1) t2[(((long)slot * np2 + z) * nlx + lx) * np1 + y] = buf[i];
  }
}

// y pass, forward: T2 lines -> y transform (exp(-i)) -> sphere columns of T1.
template <typename T>
__global__ void __launch_bounds__(NTHR)
k_ypass_fwd(GridDev g, const cplx<T>* __restrict__ t2, cplx<T>* __restrict__ t1) {
  constexpr int ZB = ZBlk<T>::ZB, LB = ZBlk<T>::LB;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int np1 = g.np1, np2 = g.np2, nlx = g.nlx;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + np1;
  const int nzb = (np2 + ZB - 1) / ZB;
  const int zb = blockIdx.x % nzb;
  const int lb = blockIdx.x / nzb;
  const int slot = blockIdx.y;
  const int z0 = zb * ZB, lx0 = lb * LB;
  const int lx1 = min(lx0 + LB, nlx);
  const cplx<T>* twg = twtab<T>(g, 1);
  for (int i = threadIdx.x; i < np1; i += NTHR) tw[i] = twg[i];
  for (int i = threadIdx.x; i < ZB * LB * np1; i += NTHR) {
    const int li = i / np1;
    const int y = i - li * np1;
    const int dz = li % ZB, dl = li / ZB;
    const int z = z0 + dz, lx = lx0 + dl;
    cplx<T> v = {T(0), T(0)};
    if (z < np2 && lx < lx1) v = t2[(((long)slot * np2 + z) * nlx + lx) * np1 + y];
    buf[i] = v;
  }
  __syncthreads();
  fft_lines<-1, EPT>(buf, ZB * LB, np1, g.ax[1], tw);
  const int c0 = g.line_start[lx0], c1 = g.line_start[lx1];
  const int nwork = (c1 - c0) * ZB;
  for (int f = threadIdx.x; f < nwork; f += NTHR) {
    const int dz = f % ZB;
    const int c = c0 + f / ZB;
    const int z = z0 + dz;
    if (z < np2) {
      int dl = 0;
      while (lx0 + dl + 1 < lx1 && c >= g.line_start[lx0 + dl + 1]) ++dl;
      t1[((long)slot * g.ncol + c) * np2 + z] = buf[(dz + ZB * dl) * np1 + g.col_y[c]];
    }
  }
}

// ------------------------------------------------------------------------------------
// x pass.  One workgroup owns YB real-space rows (z fixed, y0..y0+YB) and loops over
// the slots [s0, s1).
enum XOp { X_VMUL = 0, X_STORE = 1, X_LOAD = 2, X_LOAD_REAL = 3, X_DENSITY = 4 };

template <typename T, int OP>
__global__ void __launch_bounds__(NTHR)
k_xpass(GridDev g, cplx<T>* __restrict__ t2, int nslots, int slots_per_wg,
        const double* __restrict__ v,      // VMUL: potential; LOAD*: input grid
        double* __restrict__ out,          // STORE: complex grid; DENSITY: rho
        const double* __restrict__ dens,   // DENSITY: per slot (fac1, fac2, beta1, beta2)
        double scale) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int np0 = g.np0, np1 = g.np1, np2 = g.np2, nlx = g.nlx;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + np0;
  const int nyb = (np1 + YB - 1) / YB;
  const int z = blockIdx.x / nyb;
  const int y0 = (blockIdx.x % nyb) * YB;
  const int nyl = min(YB, np1 - y0);  // rows in this block
  const int s0 = blockIdx.y * slots_per_wg;
  const int s1 = min(nslots, s0 + slots_per_wg);
  const cplx<T>* twg = twtab<T>(g, 0);
  for (int i = threadIdx.x; i < np0; i += NTHR) tw[i] = twg[i];
  const int nel = nyl * np0;
  const long grow0 = ((long)z * np1 + y0) * np0;  // grid offset of the block's first row

  T vr[EPT];
  double acc[EPT];
  if (OP == X_VMUL) {
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int e = threadIdx.x + k * NTHR;
      vr[k] = e < nel ? T(v[grow0 + e]) : T(0);
    }
  }
  if (OP == X_DENSITY) {
#pragma unroll
    for (int k = 0; k < EPT; ++k) acc[k] = 0.0;
  }

  for (int s = s0; s < s1; ++s) {
    __syncthreads();
    if (OP == X_LOAD || OP == X_LOAD_REAL) {
      for (int e = threadIdx.x; e < YB * np0; e += NTHR) {
        cplx<T> val = {T(0), T(0)};
        if (e < nel) {
          if (OP == X_LOAD) {
            const double* fp = v + 2 * ((long)s * np0 * np1 * np2 + grow0 + e);
            val = {T(fp[0]), T(fp[1])};
          } else {
            val = {T(scale * v[(long)s * np0 * np1 * np2 + grow0 + e]), T(0)};
          }
        }
        buf[e] = val;
      }
      __syncthreads();
    } else {
      for (int e = threadIdx.x; e < YB * np0; e += NTHR) buf[e] = {T(0), T(0)};
      __syncthreads();
      const int nwork = nlx * YB;
      for (int f = threadIdx.x; f < nwork; f += NTHR) {
        const int dy = f % YB;
        const int lx = f / YB;
        if (dy < nyl)
          buf[dy * np0 + g.line_x[lx]] =
              t2[(((long)s * np2 + z) * nlx + lx) * np1 + y0 + dy];
      }
      __syncthreads();
      fft_lines<+1, EPT>(buf, YB, np0, g.ax[0], tw);
    }

    if (OP == X_VMUL) {
#pragma unroll
      for (int k = 0; k < EPT; ++k) {
        const int e = threadIdx.x + k * NTHR;
        if (e < nel) {
          cplx<T> p = buf[e];
          buf[e] = {vr[k] * p.x, vr[k] * p.y};
        }
      }
      __syncthreads();
    } else if (OP == X_STORE) {
      for (int e = threadIdx.x; e < nel; e += NTHR) {
        double* fp = out + 2 * ((long)s * np0 * np1 * np2 + grow0 + e);
        fp[0] = double(buf[e].x);
        fp[1] = double(buf[e].y);
      }
      continue;
    } else if (OP == X_DENSITY) {
      const double f1 = dens[4 * s + 0], f2 = dens[4 * s + 1];
      const double b1 = dens[4 * s + 2], b2 = dens[4 * s + 3];
#pragma unroll
      for (int k = 0; k < EPT; ++k) {
        const int e = threadIdx.x + k * NTHR;
        if (e < nel) {
          const cplx<T> p = buf[e];
          // f = psi1 + i psi2 with psi_j = X_j + conj(c_j(G=0)): Re psi1 = Re f - beta2,
          // Re psi2 = Im f + beta1, Im psi_j = -beta_j
          if (f1 > 0.0) {
            const double r1 = double(p.x) - b2;
            acc[k] += f1 * (r1 * r1 + b1 * b1);
          }
          if (f2 > 0.0) {
            const double r2 = double(p.y) + b1;
            acc[k] += f2 * (r2 * r2 + b2 * b2);
          }
        }
      }
      continue;
    }

    if (OP == X_VMUL || OP == X_LOAD || OP == X_LOAD_REAL) {
      if (OP != X_VMUL) __syncthreads();
      fft_lines<-1, EPT>(buf, YB, np0, g.ax[0], tw);
      const int nwork = nlx * YB;
      for (int f = threadIdx.x; f < nwork; f += NTHR) {
        const int dy = f % YB;
        const int lx = f / YB;
        if (dy < nyl)
          t2[(((long)s * np2 + z) * nlx + lx) * np1 + y0 + dy] = buf[dy * np0 + g.line_x[lx]];
      }
    }
  }
  if (OP == X_DENSITY) {
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int e = threadIdx.x + k * NTHR;
      if (e < nel) out[grow0 + e] += acc[k];
    }
  }
}

// Complex-basis density: one band per slot, rho += fac * |f|^2.
template <typename T>
__global__ void __launch_bounds__(NTHR)
k_xpass_density_complex(GridDev g, const cplx<T>* __restrict__ t2, int nslots,
                        double* __restrict__ rho, const double* __restrict__ facs) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int np0 = g.np0, np1 = g.np1, np2 = g.np2, nlx = g.nlx;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + np0;
  const int nyb = (np1 + YB - 1) / YB;
  const int z = blockIdx.x / nyb;
  const int y0 = (blockIdx.x % nyb) * YB;
  const int nyl = min(YB, np1 - y0);
  const cplx<T>* twg = twtab<T>(g, 0);
  for (int i = threadIdx.x; i < np0; i += NTHR) tw[i] = twg[i];
  const int nel = nyl * np0;
  const long grow0 = ((long)z * np1 + y0) * np0;
  double acc[EPT];
#pragma unroll
  for (int k = 0; k < EPT; ++k) acc[k] = 0.0;
  for (int s = 0; s < nslots; ++s) {
    __syncthreads();
    for (int e = threadIdx.x; e < YB * np0; e += NTHR) buf[e] = {T(0), T(0)};
    __syncthreads();
    for (int f = threadIdx.x; f < nlx * YB; f += NTHR) {
      const int dy = f % YB;
      const int lx = f / YB;
      if (dy < nyl)
        buf[dy * np0 + g.line_x[lx]] = t2[(((long)s * np2 + z) * nlx + lx) * np1 + y0 + dy];
    }
    __syncthreads();
    fft_lines<+1, EPT>(buf, YB, np0, g.ax[0], tw);
    const double fac = facs[s];
    if (fac > 0.0) {
#pragma unroll
      for (int k = 0; k < EPT; ++k) {
        const int e = threadIdx.x + k * NTHR;
        if (e < nel) {
          const double re = double(buf[e].x), im = double(buf[e].y);
          acc[k] += fac * (re * re + im * im);
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int e = threadIdx.x + k * NTHR;
    if (e < nel) rho[grow0 + e] += acc[k];
  }
}

// ------------------------------------------------------------------------------------
// z pass, forward: T1 columns -> z transform (exp(-i)) -> x 1/N -> unpack -> output.
// out[n][G] = value (store) or out[n][G] = (out + kin[G]*psi[n][G]) + value.
template <typename T, int MODE>
__global__ void __launch_bounds__(NTHR)
k_zpass_fwd(GridDev g, const cplx<T>* __restrict__ t1, const double* __restrict__ psi, long ld,
            int nst, int band0, double* __restrict__ out, long ldo,
            const double* __restrict__ kin, int store) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int np2 = g.np2;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + np2;
  const int grp = blockIdx.x;
  const int slot = blockIdx.y;
  const int e0 = g.grp_start[grp];
  const int nl = g.grp_start[grp + 1] - e0;
  const cplx<T>* twg = twtab<T>(g, 2);
  for (int i = threadIdx.x; i < np2; i += NTHR) tw[i] = twg[i];
  for (int i = threadIdx.x; i < nl * np2; i += NTHR) {
    const int ll = i / np2;
    const int z = i - ll * np2;
    buf[i] = t1[((long)slot * g.ncol + g.ent_col[e0 + ll]) * np2 + z];
  }
  __syncthreads();
  fft_lines<-1, EPT>(buf, nl, np2, g.ax[2], tw);

  const double fac = 1.0 / ((double)g.np0 * (double)g.np1 * (double)g.np2);
  int n1, n2;
  if (MODE == SLOT_PAIR) {
    n1 = band0 + 2 * slot;
    n2 = n1 + 1;
  } else {
    n1 = band0 + slot;
    n2 = nst;
  }
  const bool has2 = (MODE == SLOT_PAIR) && (n2 < nst);
  const int le = threadIdx.x >> 4;
  const int sub = threadIdx.x & 15;
  if (le >= nl) return;
  const int e = e0 + le;
  const int kind = g.ent_kind[e];
  if (kind == 2) return;  // -G image: handled with its +G partner (next-lower line)
  const int rod = g.ent_rod[e];
  const int first = g.rod_first[rod];
  const int lmin = g.rod_lmin[rod];
  const int size = g.rod_size[rod];
  const cplx<T>* lp = buf + le * np2;
  const cplx<T>* lm = (kind == 1) ? buf + (le + 1) * np2 : lp;
  double* o1 = out + 2 * (long)n1 * ldo;
  double* o2 = out + 2 * (long)(has2 ? n2 : n1) * ldo;
  const double* q1 = psi ? psi + 2 * (long)n1 * ld : nullptr;
  const double* q2 = psi ? psi + 2 * (long)(has2 ? n2 : n1) * ld : nullptr;
  for (int i = sub; i < size; i += 16) {
    const int l = lmin + i;
    const int ig = first + i;
    const cplx<T> zp = lp[wrap(l, np2)];
    double r1, i1, r2 = 0.0, i2 = 0.0;
    if (MODE == SLOT_PAIR) {
      const cplx<T> zm = lm[wrap(-l, np2)];
      const double a = double(zp.x) * fac, b = double(zp.y) * fac;
      const double c = double(zm.x) * fac, d = double(zm.y) * fac;
      r1 = 0.5 * (a + c);
      i1 = 0.5 * (b - d);
      r2 = 0.5 * (b + d);
      i2 = 0.5 * (c - a);
    } else {
      r1 = double(zp.x) * fac;
      i1 = double(zp.y) * fac;
    }
    if (store) {
      o1[2 * ig] = r1;
      o1[2 * ig + 1] = i1;
      if (has2) {
        o2[2 * ig] = r2;
        o2[2 * ig + 1] = i2;
      }
    } else {
      double h0 = o1[2 * ig], h1 = o1[2 * ig + 1];
      if (kin) {
        const double k = kin[ig];
        h0 += k * q1[2 * ig];
        h1 += k * q1[2 * ig + 1];
      }
      o1[2 * ig] = h0 + r1;
      o1[2 * ig + 1] = h1 + i1;
      if (has2) {
        double g0 = o2[2 * ig], g1 = o2[2 * ig + 1];
        if (kin) {
          const double k = kin[ig];
          g0 += k * q2[2 * ig];
          g1 += k * q2[2 * ig + 1];
        }
        o2[2 * ig] = g0 + r2;
        o2[2 * ig + 1] = g1 + i2;
      }
    }
  }
}

// ------------------------------------------------------------------------------------
// host side
size_t slot_bytes(const Basis* b, int prec) {
  const size_t c = prec == PREC_FP64 ? 16 : 8;
  const GridDev& g = b->g;
  return c * ((size_t)g.ncol * g.np2 + (size_t)g.np2 * g.nlx * g.np1);
}

template <typename T>
static size_t lds_z(const GridDev& g) { return sizeof(cplx<T>) * (size_t)g.np2 * (1 + ZL); }
template <typename T>
static size_t lds_y(const GridDev& g) {
  return sizeof(cplx<T>) * (size_t)g.np1 * (1 + ZBlk<T>::ZB * ZBlk<T>::LB);
}
template <typename T>
static size_t lds_x(const GridDev& g) { return sizeof(cplx<T>) * (size_t)g.np0 * (1 + YB); }

static int check_launch(Ctx* ctx, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ctx->last_error = std::string(what) + ": " + hipGetErrorString(e);
    return -3;
  }
  return 0;
}

template <typename T>
static int bwd_yz_t(Ctx* ctx, Basis* b, const PipelineArgs& a) {
  const GridDev& g = b->g;
  auto* t1 = static_cast<cplx<T>*>(a.t1);
  auto* t2 = static_cast<cplx<T>*>(a.t2);
  {
    TimedLaunch tl(ctx, "zpass_bwd");
    dim3 grid(g.ngroups, a.nslots);
    if (a.mode == SLOT_PAIR)
      k_zpass_bwd<T, SLOT_PAIR><<<grid, NTHR, lds_z<T>(g), ctx->stream>>>(g, a.psi, a.ld, a.nst,
                                                                           a.band0, t1);
    else if (a.mode == SLOT_SINGLE_REAL)
      k_zpass_bwd<T, SLOT_SINGLE_REAL><<<grid, NTHR, lds_z<T>(g), ctx->stream>>>(
          g, a.psi, a.ld, a.nst, a.band0, t1);
    else
      k_zpass_bwd<T, SLOT_COMPLEX><<<grid, NTHR, lds_z<T>(g), ctx->stream>>>(g, a.psi, a.ld, a.nst,
                                                                              a.band0, t1);
  }
  if (int r = check_launch(ctx, "zpass_bwd")) return r;
  {
    TimedLaunch tl(ctx, "ypass_bwd");
    constexpr int ZB = ZBlk<T>::ZB, LB = ZBlk<T>::LB;
    dim3 grid(((g.np2 + ZB - 1) / ZB) * ((g.nlx + LB - 1) / LB), a.nslots);
    k_ypass_bwd<T><<<grid, NTHR, lds_y<T>(g), ctx->stream>>>(g, t1, t2);
  }
  return check_launch(ctx, "ypass_bwd");
}

int pipeline_bwd_yz(Ctx* ctx, Basis* b, const PipelineArgs& a) {
  return a.prec == PREC_FP64 ? bwd_yz_t<double>(ctx, b, a) : bwd_yz_t<float>(ctx, b, a);
}

template <typename T>
static int fwd_zy_t(Ctx* ctx, Basis* b, const PipelineArgs& a, double* out, long ldo,
                    const double* kin, int store) {
  const GridDev& g = b->g;
  auto* t1 = static_cast<cplx<T>*>(a.t1);
  auto* t2 = static_cast<cplx<T>*>(a.t2);
  {
    TimedLaunch tl(ctx, "ypass_fwd");
    constexpr int ZB = ZBlk<T>::ZB, LB = ZBlk<T>::LB;
    dim3 grid(((g.np2 + ZB - 1) / ZB) * ((g.nlx + LB - 1) / LB), a.nslots);
    k_ypass_fwd<T><<<grid, NTHR, lds_y<T>(g), ctx->stream>>>(g, t2, t1);
  }
  if (int r = check_launch(ctx, "ypass_fwd")) return r;
  {
    TimedLaunch tl(ctx, "zpass_fwd");
    dim3 grid(g.ngroups, a.nslots);
    if (a.mode == SLOT_PAIR)
      k_zpass_fwd<T, SLOT_PAIR><<<grid, NTHR, lds_z<T>(g), ctx->stream>>>(
          g, t1, a.psi, a.ld, a.nst, a.band0, out, ldo, kin, store);
    else if (a.mode == SLOT_SINGLE_REAL)
      k_zpass_fwd<T, SLOT_SINGLE_REAL><<<grid, NTHR, lds_z<T>(g), ctx->stream>>>(
          g, t1, a.psi, a.ld, a.nst, a.band0, out, ldo, kin, store);
    else
      k_zpass_fwd<T, SLOT_COMPLEX><<<grid, NTHR, lds_z<T>(g), ctx->stream>>>(
          g, t1, a.psi, a.ld, a.nst, a.band0, out, ldo, kin, store);
  }
  return check_launch(ctx, "zpass_fwd");
}

int pipeline_fwd_zy_accumulate(Ctx* ctx, Basis* b, const PipelineArgs& a, double* out, long ldo,
                               const double* kin, int store) {
  return a.prec == PREC_FP64 ? fwd_zy_t<double>(ctx, b, a, out, ldo, kin, store)
                             : fwd_zy_t<float>(ctx, b, a, out, ldo, kin, store);
}

// slots per x-pass workgroup: enough workgroups to fill the chip several times over
static int x_slots_per_wg(const GridDev& g, int nslots) {
  const long rows = (long)g.np2 * ((g.np1 + YB - 1) / YB);
  int spw = 1;
  while (spw < 8 && rows * ((nslots + 2 * spw - 1) / (2 * spw)) >= 4096) spw *= 2;
  return spw;
}

template <typename T, int OP>
static int xpass_t(Ctx* ctx, Basis* b, const PipelineArgs& a, const double* in, double* out,
                   const double* dens, double scale, const char* name) {
  const GridDev& g = b->g;
  auto* t2 = static_cast<cplx<T>*>(a.t2);
  const int nyb = (g.np1 + YB - 1) / YB;
  int spw = (OP == X_DENSITY) ? a.nslots : x_slots_per_wg(g, a.nslots);
  dim3 grid(g.np2 * nyb, (a.nslots + spw - 1) / spw);
  {
    TimedLaunch tl(ctx, name);
    k_xpass<T, OP><<<grid, NTHR, lds_x<T>(g), ctx->stream>>>(g, t2, a.nslots, spw, in, out, dens,
                                                             scale);
  }
  return check_launch(ctx, name);
}

int pipeline_x_vmul(Ctx* ctx, Basis* b, const PipelineArgs& a, const double* v) {
  return a.prec == PREC_FP64
             ? xpass_t<double, X_VMUL>(ctx, b, a, v, nullptr, nullptr, 1.0, "xpass_vmul")
             : xpass_t<float, X_VMUL>(ctx, b, a, v, nullptr, nullptr, 1.0, "xpass_vmul");
}
int pipeline_x_store(Ctx* ctx, Basis* b, const PipelineArgs& a, double* f) {
  return a.prec == PREC_FP64
             ? xpass_t<double, X_STORE>(ctx, b, a, nullptr, f, nullptr, 1.0, "xpass_store")
             : xpass_t<float, X_STORE>(ctx, b, a, nullptr, f, nullptr, 1.0, "xpass_store");
}
int pipeline_x_load(Ctx* ctx, Basis* b, const PipelineArgs& a, const double* f, int real_input,
                    double scale) {
  if (real_input)
    return a.prec == PREC_FP64
               ? xpass_t<double, X_LOAD_REAL>(ctx, b, a, f, nullptr, nullptr, scale, "xpass_load")
               : xpass_t<float, X_LOAD_REAL>(ctx, b, a, f, nullptr, nullptr, scale, "xpass_load");
  return a.prec == PREC_FP64
             ? xpass_t<double, X_LOAD>(ctx, b, a, f, nullptr, nullptr, 1.0, "xpass_load")
             : xpass_t<float, X_LOAD>(ctx, b, a, f, nullptr, nullptr, 1.0, "xpass_load");
}
int pipeline_x_density(Ctx* ctx, Basis* b, const PipelineArgs& a, const double* dens,
                       double* rho) {
  const GridDev& g = b->g;
  if (a.mode == SLOT_COMPLEX) {
    const int nyb = (g.np1 + YB - 1) / YB;
    dim3 grid(g.np2 * nyb, 1);
    TimedLaunch tl(ctx, "xpass_density");
    if (a.prec == PREC_FP64)
      k_xpass_density_complex<double><<<grid, NTHR, lds_x<double>(g), ctx->stream>>>(
          g, static_cast<cplx<double>*>(a.t2), a.nslots, rho, dens);
    else
      k_xpass_density_complex<float><<<grid, NTHR, lds_x<float>(g), ctx->stream>>>(
          g, static_cast<cplx<float>*>(a.t2), a.nslots, rho, dens);
    return check_launch(ctx, "xpass_density");
  }
  return a.prec == PREC_FP64
             ? xpass_t<double, X_DENSITY>(ctx, b, a, nullptr, rho, dens, 1.0, "xpass_density")
             : xpass_t<float, X_DENSITY>(ctx, b, a, nullptr, rho, dens, 1.0, "xpass_density");
}

}  // namespace qbh
