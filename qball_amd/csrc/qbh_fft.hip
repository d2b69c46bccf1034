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
//
// Work split: every workgroup has 4 wavefronts; each wavefront owns LPW lines of LDS and
// transforms them on its own (fft_lines_w: no workgroup barrier inside a transform).
// Global reads and writes are arranged so that a wavefront touches contiguous segments
// of >= 64 bytes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

#include "fft_reg.h"
#include "qbh_internal.h"

namespace qbh {

constexpr int NTHR = 256;  // threads per workgroup, every pipeline kernel
constexpr int NWAVE = 4;
constexpr int ZLPW = 4;    // z-pass lines per wavefront (16 per group)

// per precision: lines per wavefront in the y and x passes, y-pass plane block
template <typename T> struct Cfg;
// x-pass rows per wavefront and launch tuning (compile-time knobs)
#ifndef QBH_XR64
#define QBH_XR64 4
#endif
#ifndef QBH_XR32
#define QBH_XR32 4
#endif
// x pass: slots whose sphere lines are loaded ahead (register buffers of GPW elements
// each).  Measured on MI355X at Si512: two slots ahead slowed the FP32 x pass from 4.26
// to 4.60 ms per Hψ at three waves per SIMD (register spills) and to 5.57 ms at two; FP64
// has no registers to spare.
#ifndef QBH_XPF64
#define QBH_XPF64 1
#endif
#ifndef QBH_XPF32
#define QBH_XPF32 1
#endif
// x pass: minimum waves per SIMD the register allocation must allow (the counts the
// kernels reach without spilling: FP32 three, FP64 two)
#ifndef QBH_XWPE64
#define QBH_XWPE64 2
#endif
#ifndef QBH_XWPE32
#define QBH_XWPE32 3
#endif
// x pass: stage-B forward twiddles held in registers across slots (1) or formed per slot
#ifndef QBH_XTWB
#define QBH_XTWB 1
#endif
// x pass: padded LDS rows (element x at x + x/N2).  The x pass holds two workgroups per CU
// (VGPR-bound), so the larger rows cost it no occupancy, unlike the y/z passes.
#ifndef QBH_XLDS_PAD
#define QBH_XLDS_PAD 1
#endif
// y pass: ZB z-planes x LB x-lines per workgroup, LPW lines per wavefront (ZB*LB = 4*LPW).
// Measured on MI355X at Si512 (140^3): 16 x 1 beats 8 x 2, 4 x 4, 32 x 1 and the 8-line
// shapes in both precisions (FP64 forward y pass 4.58 -> 4.05 ms, FP32 y passes 5.78 ->
// 3.63 ms per Hψ): one x-line per workgroup and 256-byte (FP64) column segments.
#ifndef QBH_YZB64
#define QBH_YZB64 16
#endif
#ifndef QBH_YLB64
#define QBH_YLB64 1
#endif
#ifndef QBH_YZB32
#define QBH_YZB32 16
#endif
#ifndef QBH_YLB32
#define QBH_YLB32 1
#endif
template <> struct Cfg<double> {
  static constexpr int XR = QBH_XR64;  // x pass: rows per wavefront (64 B row segments)
  static constexpr int XPF = QBH_XPF64;  // x pass: slots of sphere lines in flight
  static constexpr int XWPE = QBH_XWPE64;  // x pass: minimum waves per SIMD asked of the compiler
  static constexpr int ZB = QBH_YZB64;  // y pass: z-planes per workgroup (column segments)
  static constexpr int LB = QBH_YLB64;  // y pass: x-lines per workgroup
  static constexpr int LPW = ZB * LB / 4;
};
template <> struct Cfg<float> {
  static constexpr int XR = QBH_XR32;
  static constexpr int XPF = QBH_XPF32;
  static constexpr int XWPE = QBH_XWPE32;
  static constexpr int ZB = QBH_YZB32;
  static constexpr int LB = QBH_YLB32;
  static constexpr int LPW = ZB * LB / 4;
};

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

// elements per lane of a wavefront owning `lpw` lines of length <= NMAX
template <int LPW, int NMAX> struct Epl { static constexpr int v = (LPW * NMAX + 63) / 64; };

// Wavefront-local transform of nl lines of length ld: the two-factor register transform
// when the length has an instantiated factorisation n = N1*N2, else the LDS Stockham passes.
template <int DIR, int EPL, int N1, int N2, int RPW, typename T>
__device__ __forceinline__ void line_fft(cplx<T>* buf, int nl, int ld, const AxisPlan& ax,
                                         const cplx<T>* tw, int lane) {
  if constexpr (N1 > 0)
    rows_fft<N1, N2, RPW, DIR>(buf, nl, ld, tw, lane);
  else
    fft_lines_w<DIR, EPL>(buf, nl, ld, ax, tw, lane);
}

// ------------------------------------------------------------------------------------
// z pass, backward: gather sphere coefficients of one slot into its columns, transform
// along z (exp(+i)), store T1[slot][col][z].  Line e of a group belongs to wavefront e/4.
template <typename T, int MODE, int NMAX, int N1, int N2>
__global__ void __launch_bounds__(NTHR)
k_zpass_bwd(GridDev g, const double* __restrict__ psi, long ld, int nst, int band0,
            cplx<T>* __restrict__ t1) {
  constexpr int EPL = Epl<ZLPW, NMAX>::v;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int np2 = g.np2;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int grp = blockIdx.x;
  const int slot = blockIdx.y;
  // this wavefront's lines and the rod of the line its 16-lane group gathers, read before
  // the workgroup barrier so the table loads overlap the twiddle load
  const int e0 = g.grp_start[grp] + wave * ZLPW;
  const int nl = min(ZLPW, g.grp_start[grp + 1] - e0);
  const int le = lane >> 4;
  const int sub = lane & 15;
  int kind = 4, first = 0, lmin = 0, size = 0;
  if (le < nl) {
    kind = g.ent_kind[e0 + le];
    if (kind != 4) {
      const int rod = g.ent_rod[e0 + le];
      first = g.rod_first[rod];
      lmin = g.rod_lmin[rod];
      size = g.rod_size[rod];
    }
  }
  const cplx<T>* twg = twtab<T>(g, 2);
  for (int i = threadIdx.x; i < np2; i += NTHR) tw[i] = twg[i];
  __syncthreads();
  if (nl <= 0) return;
  using PD = Pad<N1, N2>;
  const int LD = PD::row(np2);
  cplx<T>* buf = tw + np2 + wave * ZLPW * LD;
  for (int i = lane; i < nl * LD; i += 64) buf[i] = {T(0), T(0)};
  wave_sync();

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

  // 16 lanes per line; KZ coefficients per lane loaded together before they are stored
  constexpr int KZ = 4;
  if (le < nl && kind != 4) {
    cplx<T>* line = buf + le * LD;
    for (int i0 = sub; i0 < size; i0 += 16 * KZ) {
      double a[KZ], b[KZ], c[KZ], d[KZ];
      // branch-free: every lane loads an element of its rod (past the end: the last one,
      // never stored), so all KZ loads are in flight together
#pragma unroll
      for (int k = 0; k < KZ; ++k) {
        const long e = 2 * (long)(first + min(i0 + 16 * k, size - 1));
        a[k] = p1[e];
        b[k] = p1[e + 1];
        c[k] = d[k] = 0.0;
        if (MODE == SLOT_PAIR) {  // p2 is p1 for a lone last band; its values are dropped
          c[k] = p2[e];
          d[k] = p2[e + 1];
        }
      }
      if (!has2)
#pragma unroll
        for (int k = 0; k < KZ; ++k) c[k] = d[k] = 0.0;
#pragma unroll
      for (int k = 0; k < KZ; ++k) {
        const int i = i0 + 16 * k;
        if (i >= size) continue;
        const int l = lmin + i;
        double pr, pi, mr, mi;  // +G and -G values
        if (MODE == SLOT_PAIR) {
          pr = a[k] - d[k]; pi = b[k] + c[k]; mr = a[k] + d[k]; mi = c[k] - b[k];
        } else {
          pr = a[k]; pi = b[k]; mr = a[k]; mi = -b[k];
        }
        if (kind == 0) {
          if (i > 0) line[PD::idx(wrap(l, np2))] = {T(pr), T(pi)};
          line[PD::idx(wrap(-l, np2))] = {T(mr), T(mi)};
        } else if (kind == 1 || kind == 3) {
          line[PD::idx(wrap(l, np2))] = {T(pr), T(pi)};
        } else {
          line[PD::idx(wrap(-l, np2))] = {T(mr), T(mi)};
        }
      }
    }
  }
  wave_sync();
  line_fft<+1, EPL, N1, N2, ZLPW>(buf, nl, LD, g.ax[2], tw, lane);
  for (int ll = 0; ll < nl; ++ll) {
    const int col = g.ent_col[e0 + ll];
    if (col < 0) continue;
    cplx<T>* dst = t1 + ((long)slot * g.ncol + col) * np2;
    for (int z = lane; z < np2; z += 64) dst[z] = buf[ll * LD + PD::idx(z)];
  }
}

// ------------------------------------------------------------------------------------
// y pass, backward: T1 columns of ZB planes x LB x-lines -> y transform -> T2 lines.
// The workgroup gathers its columns cooperatively (ZB-element segments), then each
// wavefront transforms and stores its own LPW lines.
template <typename T, int NMAX, int N1, int N2>
__global__ void __launch_bounds__(NTHR)
k_ypass_bwd(GridDev g, const cplx<T>* __restrict__ t1, cplx<T>* __restrict__ t2) {
  constexpr int ZB = Cfg<T>::ZB, LB = Cfg<T>::LB, LPW = Cfg<T>::LPW;
  constexpr int EPL = Epl<LPW, NMAX>::v;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int np1 = g.np1, np2 = g.np2, nlx = g.nlx;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + np1;
  using PD = Pad<N1, N2>;
  const int LD = PD::row(np1);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nzb = (np2 + ZB - 1) / ZB;
  const int zb = blockIdx.x % nzb;
  const int lb = blockIdx.x / nzb;
  const int slot = blockIdx.y;
  const int z0 = zb * ZB, lx0 = lb * LB;
  const int lx1 = min(lx0 + LB, nlx);
  const cplx<T>* twg = twtab<T>(g, 1);
  for (int i = threadIdx.x; i < np1; i += NTHR) tw[i] = twg[i];
  for (int i = threadIdx.x; i < ZB * LB * LD; i += NTHR) buf[i] = {T(0), T(0)};
  __syncthreads();
  const int c0 = g.line_start[lx0], c1 = g.line_start[lx1];
  const int nwork = (c1 - c0) * ZB;
  if constexpr (sizeof(T) == 4) {
    // FP32: KG elements per thread and round, all loaded (with their column indices) before
    // any is stored; past the end a thread reloads the last element and stores nothing
    // (at Si512 1.64 -> 1.49 ms per Hψ; in FP64 the same form measured 3.33 -> 3.40-3.55 ms)
    constexpr int KG = 4;
    for (int f0 = threadIdx.x; f0 < nwork; f0 += NTHR * KG) {
      cplx<T> v[KG];
      int col[KG], cc[KG];
#pragma unroll
      for (int k = 0; k < KG; ++k) {
        const int f = min(f0 + k * NTHR, nwork - 1);
        const int c = c0 + f / ZB;
        const int z = min(z0 + f % ZB, np2 - 1);
        cc[k] = c;
        col[k] = g.col_y[c];
        v[k] = t1[((long)slot * g.ncol + c) * np2 + z];
      }
#pragma unroll
      for (int k = 0; k < KG; ++k) {
        const int f = f0 + k * NTHR;
        const int dz = f % ZB;
        if (f < nwork && z0 + dz < np2) {
          int dl = 0;
          while (lx0 + dl + 1 < lx1 && cc[k] >= g.line_start[lx0 + dl + 1]) ++dl;
          buf[(dz + ZB * dl) * LD + PD::idx(col[k])] = v[k];
        }
      }
    }
  } else {
    for (int f = threadIdx.x; f < nwork; f += NTHR) {
      const int dz = f % ZB;
      const int c = c0 + f / ZB;
      const int z = z0 + dz;
      if (z < np2) {
        int dl = 0;
        while (lx0 + dl + 1 < lx1 && c >= g.line_start[lx0 + dl + 1]) ++dl;
        buf[(dz + ZB * dl) * LD + PD::idx(g.col_y[c])] = t1[((long)slot * g.ncol + c) * np2 + z];
      }
    }
  }
  __syncthreads();
  // line li = dz + ZB*dl; wavefront w owns lines [w*LPW, (w+1)*LPW)
  cplx<T>* wbuf = buf + wave * LPW * LD;
  line_fft<+1, EPL, N1, N2, LPW>(wbuf, LPW, LD, g.ax[1], tw, lane);
  for (int ll = 0; ll < LPW; ++ll) {
    const int li = wave * LPW + ll;
    const int dz = li % ZB, dl = li / ZB;
    const int z = z0 + dz, lx = lx0 + dl;
    if (z >= np2 || lx >= lx1) continue;
    cplx<T>* dst = t2 + (((long)slot * np2 + z) * nlx + lx) * np1;
    for (int y = lane; y < np1; y += 64) dst[y] = wbuf[ll * LD + PD::idx(y)];
  }
}

// y pass, forward: T2 lines -> y transform (exp(-i)) -> sphere columns of T1.
template <typename T, int NMAX, int N1, int N2>
__global__ void __launch_bounds__(NTHR)
k_ypass_fwd(GridDev g, const cplx<T>* __restrict__ t2, cplx<T>* __restrict__ t1) {
  constexpr int ZB = Cfg<T>::ZB, LB = Cfg<T>::LB, LPW = Cfg<T>::LPW;
  constexpr int EPL = Epl<LPW, NMAX>::v;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int np1 = g.np1, np2 = g.np2, nlx = g.nlx;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + np1;
  using PD = Pad<N1, N2>;
  const int LD = PD::row(np1);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nzb = (np2 + ZB - 1) / ZB;
  const int zb = blockIdx.x % nzb;
  const int lb = blockIdx.x / nzb;
  const int slot = blockIdx.y;
  const int z0 = zb * ZB, lx0 = lb * LB;
  const int lx1 = min(lx0 + LB, nlx);
  const cplx<T>* twg = twtab<T>(g, 1);
  for (int i = threadIdx.x; i < np1; i += NTHR) tw[i] = twg[i];
  // LDS offset (without the z-plane term) of each column of this workgroup's lines, staged
  // once so the column scatter below waits on no index-table load
  int* cof = reinterpret_cast<int*>(buf + ZB * LB * LD);
  const int c0 = g.line_start[lx0], c1 = g.line_start[lx1];
  for (int i = threadIdx.x; i < c1 - c0; i += NTHR) {
    const int c = c0 + i;
    int dl = 0;
    while (lx0 + dl + 1 < lx1 && c >= g.line_start[lx0 + dl + 1]) ++dl;
    cof[i] = ZB * dl * LD + PD::idx(g.col_y[c]);
  }
  __syncthreads();
  cplx<T>* wbuf = buf + wave * LPW * LD;
  {
    // the wave's LPW lines (KY elements per lane each) are all loaded before any is stored
    // to LDS, so no load waits on the one before it (lines past the grid read a valid
    // line and store zeros)
    // elements per lane: a register transform's length is N1 N2 (which can exceed NMAX),
    // a generic one's at most NMAX
    constexpr int KY = ((N1 > 0 ? N1 * N2 : NMAX) + 63) / 64;
    cplx<T> v[LPW][KY];
    bool in[LPW][KY];
#pragma unroll
    for (int ll = 0; ll < LPW; ++ll) {
      const int li = wave * LPW + ll;
      const int dz = li % ZB, dl = li / ZB;
      const int z = z0 + dz, lx = lx0 + dl;
      const bool ok = z < np2 && lx < lx1;
      const cplx<T>* src = t2 + (((long)slot * np2 + (ok ? z : 0)) * nlx + (ok ? lx : 0)) * np1;
#pragma unroll
      for (int k = 0; k < KY; ++k) {
        const int y = lane + 64 * k;
        in[ll][k] = ok && y < np1;
        v[ll][k] = src[in[ll][k] ? y : 0];
      }
    }
#pragma unroll
    for (int ll = 0; ll < LPW; ++ll)
#pragma unroll
      for (int k = 0; k < KY; ++k) {
        const int y = lane + 64 * k;
        if (y < np1) wbuf[ll * LD + PD::idx(y)] = in[ll][k] ? v[ll][k] : cplx<T>{T(0), T(0)};
      }
  }
  wave_sync();
  line_fft<-1, EPL, N1, N2, LPW>(wbuf, LPW, LD, g.ax[1], tw, lane);
  __syncthreads();
  const int nwork = (c1 - c0) * ZB;
  for (int f = threadIdx.x; f < nwork; f += NTHR) {
    const int dz = f % ZB;
    const int c = c0 + f / ZB;
    const int z = z0 + dz;
    if (z < np2) t1[((long)slot * g.ncol + c) * np2 + z] = buf[dz * LD + cof[c - c0]];
  }
}

// ------------------------------------------------------------------------------------
// x pass.  Wavefront w of a workgroup owns LPW consecutive real-space rows (z fixed,
// y = y0 + w*LPW ...) and loops over the slots [s0, s1) with no workgroup barrier; the
// next slot's sphere lines are loaded into registers while the current slot transforms.
enum XOp { X_VMUL = 0, X_STORE = 1, X_LOAD = 2, X_LOAD_REAL = 3, X_DENSITY = 4 };

template <typename T, int OP, int NMAX>
__global__ void __launch_bounds__(NTHR)
k_xpass(GridDev g, cplx<T>* __restrict__ t2, int nslots, int slots_per_wg,
        const double* __restrict__ v,      // VMUL: potential; LOAD*: input grid
        double* __restrict__ out,          // STORE: complex grid; DENSITY: rho
        const double* __restrict__ dens,   // DENSITY: per slot (fac1, fac2, beta1, beta2)
        double scale, int cplx_density) {
  constexpr int LPW = Cfg<T>::LPW;
  constexpr int YB = LPW * NWAVE;
  constexpr int EPL = Epl<LPW, NMAX>::v;          // row elements per lane
  constexpr int GPL = EPL;  // sphere-line elements per lane (nlx <= np0)
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int np0 = g.np0, np1 = g.np1, np2 = g.np2, nlx = g.nlx;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nyb = (np1 + YB - 1) / YB;
  const int z = blockIdx.x / nyb;
  const int y0 = (blockIdx.x % nyb) * YB + wave * LPW;  // this wavefront's first row
  const int nyl = min(LPW, np1 - y0);                   // rows of this wavefront
  const int s0 = blockIdx.y * slots_per_wg;
  const int s1 = min(nslots, s0 + slots_per_wg);
  const cplx<T>* twg = twtab<T>(g, 0);
  for (int i = threadIdx.x; i < np0; i += NTHR) tw[i] = twg[i];
  __syncthreads();
  if (nyl <= 0) return;
  cplx<T>* buf = tw + np0 + wave * LPW * np0;
  const int nel = nyl * np0;
  const long grow0 = ((long)z * np1 + y0) * np0;  // grid offset of the wavefront's first row
  const long N = (long)np0 * np1 * np2;
  const int ng = nlx * LPW;                       // sphere-line elements of the rows

  T vr[EPL];
  double acc[EPL];
  if (OP == X_VMUL) {
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
      const int e = lane + k * 64;
      vr[k] = e < nel ? T(v[grow0 + e]) : T(0);
    }
  }
  if (OP == X_DENSITY) {
#pragma unroll
    for (int k = 0; k < EPL; ++k) acc[k] = 0.0;
  }

  // sphere-line element q = lane + k*64: row dy = q % LPW, line lx = q / LPW.  Its LDS
  // offset (-1 past the lines) and its T2 offset within a slot do not depend on the slot:
  // computed once, so the slot loop issues no index-table loads.
  const long sstride = (long)np2 * nlx * np1;
  int lofs[GPL], gofs[GPL];
#pragma unroll
  for (int k = 0; k < GPL; ++k) {
    const int q = lane + k * 64;
    const int dy = q % LPW, lx = q / LPW;
    lofs[k] = q < ng ? dy * np0 + g.line_x[lx] : -1;
    gofs[k] = (q < ng && dy < nyl) ? (z * nlx + lx) * np1 + y0 + dy : -1;
  }
  cplx<T> pf[GPL];
  auto prefetch = [&](int s) {
    const cplx<T>* src = t2 + s * sstride;
#pragma unroll
    for (int k = 0; k < GPL; ++k)
      pf[k] = gofs[k] >= 0 ? src[gofs[k]] : cplx<T>{T(0), T(0)};
  };
  const bool sparse_in = !(OP == X_LOAD || OP == X_LOAD_REAL);
  if (sparse_in && s0 < s1) prefetch(s0);

  for (int s = s0; s < s1; ++s) {
    if (sparse_in) {
      for (int i = lane; i < LPW * np0; i += 64) buf[i] = {T(0), T(0)};
      wave_sync();
#pragma unroll
      for (int k = 0; k < GPL; ++k)
        if (lofs[k] >= 0) buf[lofs[k]] = pf[k];
      wave_sync();
      if (s + 1 < s1) prefetch(s + 1);
      fft_lines_w<+1, EPL>(buf, LPW, np0, g.ax[0], tw, lane);
    } else {
      for (int e = lane; e < LPW * np0; e += 64) {
        cplx<T> val = {T(0), T(0)};
        if (e < nel) {
          if (OP == X_LOAD) {
            const double* fp = v + 2 * ((long)s * N + grow0 + e);
            val = {T(fp[0]), T(fp[1])};
          } else {
            val = {T(scale * v[(long)s * N + grow0 + e]), T(0)};
          }
        }
        buf[e] = val;
      }
      wave_sync();
    }

    if (OP == X_VMUL) {
#pragma unroll
      for (int k = 0; k < EPL; ++k) {
        const int e = lane + k * 64;
        if (e < nel) {
          const cplx<T> p = buf[e];
          buf[e] = {vr[k] * p.x, vr[k] * p.y};
        }
      }
      wave_sync();
    } else if (OP == X_STORE) {
      for (int e = lane; e < nel; e += 64) {
        double* fp = out + 2 * ((long)s * N + grow0 + e);
        fp[0] = double(buf[e].x);
        fp[1] = double(buf[e].y);
      }
      wave_sync();
      continue;
    } else if (OP == X_DENSITY) {
      if (cplx_density) {
        const double f = dens[s];
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
          const int e = lane + k * 64;
          if (e < nel && f > 0.0) {
            const double re = double(buf[e].x), im = double(buf[e].y);
            acc[k] += f * (re * re + im * im);
          }
        }
      } else {
        const double f1 = dens[4 * s + 0], f2 = dens[4 * s + 1];
        const double b1 = dens[4 * s + 2], b2 = dens[4 * s + 3];
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
          const int e = lane + k * 64;
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
      }
      wave_sync();
      continue;
    }

    // VMUL / LOAD*: forward x transform and store of the sphere lines
    fft_lines_w<-1, EPL>(buf, LPW, np0, g.ax[0], tw, lane);
    {
      cplx<T>* dst = t2 + s * sstride;
#pragma unroll
      for (int k = 0; k < GPL; ++k)
        if (gofs[k] >= 0) dst[gofs[k]] = buf[lofs[k]];
    }
    wave_sync();
  }
  if (OP == X_DENSITY) {
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
      const int e = lane + k * 64;
      if (e < nel) out[grow0 + e] += acc[k];
    }
  }
}

// ------------------------------------------------------------------------------------
// x pass with the two-factor register transform (fft_reg.h), rows of n = N1*N2:
//   stage A  (bwd): DFT_N1 over stride-N2 elements, twiddle W_n^{j2 k1}
//   stage B  (bwd): DFT_N2 over contiguous blocks -> psi(r) at x = k1 + N1*k2, in
//                   registers; VMUL multiplies by v(r) there and runs stage B of the
//                   forward transform (DFT_N2, twiddle W_n^{-k1 m2}) without leaving
//                   registers; DENSITY accumulates |psi|^2 there; STORE puts psi(r)
//                   back into the rows in natural order and writes the rows to the grid
//   stage A' (fwd): DFT_N1 over stride-N2 elements -> natural order
// Three LDS round trips per slot instead of one per radix pass, and the x-dependent
// multiply touches each element once, in registers.
template <typename T, int OP, int N1, int N2, int GPW>
__global__ void __launch_bounds__(NTHR)
__attribute__((amdgpu_waves_per_eu(Cfg<T>::XWPE)))
k_xpass2(GridDev g, cplx<T>* __restrict__ t2, int nslots, int slots_per_wg,
         const double* __restrict__ v, double* __restrict__ out, const double* __restrict__ dens,
         int cplx_density) {
  constexpr int LPW = Cfg<T>::XR;
  constexpr int YB = LPW * NWAVE;
  constexpr int n = N1 * N2;
  constexpr int TA = LPW * N2, TB = LPW * N1;
  constexpr int IA = (TA + 63) / 64, IB = (TB + 63) / 64;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int np1 = g.np1, np2 = g.np2, nlx = g.nlx;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nyb = (np1 + YB - 1) / YB;
  const int z = blockIdx.x / nyb;
  const int ywg = (blockIdx.x % nyb) * YB;  // the workgroup's first row
  const int y0 = ywg + wave * LPW;
  // rows of this wavefront; a wavefront with none still takes part in the workgroup
  // barriers below (its transforms run on zeros and every store it would make is masked)
  const int nyl = min(LPW, np1 - y0);
  const int s0 = blockIdx.y * slots_per_wg;
  const int s1 = min(nslots, s0 + slots_per_wg);
  const cplx<T>* twg = twtab<T>(g, 0);
  for (int i = threadIdx.x; i < n; i += NTHR) tw[i] = twg[i];
  __syncthreads();
  using PD = Pad<N1, N2, QBH_XLDS_PAD>;
  constexpr int LD = PD::row(n);
  constexpr int S = N2 + PD::P;
  cplx<T>* buf = tw + n + wave * LPW * LD;
  const long grow0 = ((long)z * np1 + y0) * n;

  T vr[IB][N2];
  double acc[IB][N2];
  // stage-B forward twiddles W_n^{-k1 m2}, m2 = 1..N2-1, of this lane's k1: the same
  // repeated products twiddle_row forms, computed once instead of per slot (measured at
  // 140 = 14*10: x pass 6.24 -> 6.01 ms per FP64 Hψ).  Not for the FP64 kernels with the
  // full line-element count, which have no registers to spare.
  constexpr int GS = (YB * (n / 2 + 8) + NTHR - 1) / NTHR;
  constexpr bool TWB = QBH_XTWB && (sizeof(T) == 4 || GPW <= GS);
  cplx<T> twb[IB][TWB ? N2 - 1 : 1];
  if constexpr (TWB && OP == X_VMUL) {
#pragma unroll
    for (int ib = 0; ib < IB; ++ib) {
      const int k1 = (lane + 64 * ib) % N1;
      cplx<T> w = tw[k1];
      w.y = -w.y;
      cplx<T> p = w;
#pragma unroll
      for (int m = 1; m < N2; ++m) {
        twb[ib][m - 1] = p;
        if (m + 1 < N2) p = cmul(p, w);
      }
    }
  }
#pragma unroll
  for (int ib = 0; ib < IB; ++ib) {
    const int t = lane + 64 * ib;
    const int r = t / N1, k1 = t % N1;
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) {
      if (OP == X_VMUL) vr[ib][k2] = (t < TB && r < nyl) ? T(v[grow0 + r * n + k1 + N1 * k2]) : T(0);
      if (OP == X_DENSITY) acc[ib][k2] = 0.0;
    }
  }

  // The workgroup moves the sphere lines of its YB rows together: element q = tid + NTHR*k
  // is row dy = q % YB (= tid % YB, as YB divides NTHR) of line lx = q / YB, so one load
  // instruction covers YB consecutive rows of NTHR/YB lines (YB*sizeof(cplx) contiguous
  // bytes per line) instead of each wavefront's LPW rows.  GPW (a launch parameter) covers
  // nlx*YB elements.  LDS offset (relative to the first row buffer) and T2 offset within a
  // slot are computed once, and the loads and LDS stores are unconditional: an element past
  // the lines loads element 0 of the slot and stores it to this thread's dump entry after
  // the row buffers; a row past np1 loads element 0 into its own (never stored) row.
  static_assert(NTHR % YB == 0, "rows per workgroup must divide the workgroup size");
  const bool rowok = ywg + (int)(threadIdx.x % YB) < np1;
  const long sstride = (long)np2 * nlx * np1;
  const int ngw = nlx * YB;
  cplx<T>* rows = tw + n;  // row buffers of all wavefronts, row dy at dy * LD
  constexpr int DUMP = NWAVE * LPW * LD;  // + threadIdx.x: one LDS entry per thread
  int lofs[GPW], gofs[GPW];
  unsigned gok = 0;  // bit k: element k is a line entry of a row below np1
#pragma unroll
  for (int k = 0; k < GPW; ++k) {
    const int q = threadIdx.x + k * NTHR;
    const int dy = q % YB, lx = q / YB;
    const bool in = q < ngw;
    lofs[k] = in ? dy * LD + PD::idx(g.line_x[lx]) : DUMP + (int)threadIdx.x;
    gofs[k] = (in && rowok) ? (z * nlx + lx) * np1 + ywg + dy : 0;
    if (in && rowok) gok |= 1u << k;
  }
  // the sphere lines of the next XPF slots are in flight in registers while a slot
  // transforms (buffer p of slot s is refilled with slot s + XPF)
  constexpr int XPF = Cfg<T>::XPF;
  cplx<T> pf[XPF][GPW];
  auto prefetch = [&](int s, cplx<T>(&p)[GPW]) {
    const cplx<T>* src = t2 + s * sstride;
#pragma unroll
    for (int k = 0; k < GPW; ++k) {
      const cplx<T> e = src[gofs[k]];
      p[k].x = e.x;
      p[k].y = e.y;
    }
  };
#pragma unroll
  for (int d = 0; d < XPF; ++d)
    if (s0 + d < s1) prefetch(s0 + d, pf[d]);

  // The sphere lines occupy x in [0, xpos_max] and [xneg_min, n).  Stage A input index
  // j1 in [ja, jb] is zero for every j2 and is not read; stage A' outputs there are not
  // needed (FP32 stores them anyway; they are never read).  Only the rest of the rows
  // outside the lines has to be cleared per slot:
  // nz0 <= N2-1 positions below N2*ja and nz1 <= N2-1 below xneg_min, so LPW*(nz0+nz1)
  // < LPW*2*N2 entries, ZK per lane, offsets again computed once.
  const int ja = g.xpos_max / N2 + 1;
  const int jb = g.xneg_min / N2 - 1;
  const int z0a = g.xpos_max + 1, z0b = min(N2 * ja, g.xneg_min);
  const int z1a = max(N2 * (jb + 1), z0b), z1b = g.xneg_min;
  const int nz0 = max(0, z0b - z0a), nz1 = max(0, z1b - z1a);
  constexpr int ZK = (LPW * 2 * N2 + 63) / 64;
  int zofs[ZK];
#pragma unroll
  for (int k = 0; k < ZK; ++k) {
    const int i = lane + 64 * k;
    const int nz = nz0 + nz1;
    zofs[k] = -1;
    if (i < LPW * nz) {
      const int r = i / nz, j = i % nz;
      zofs[k] = r * LD + PD::idx(j < nz0 ? z0a + j : z1a + j - nz0);
    }
  }

  // Per slot: clear (own rows) and scatter (all rows), barrier, transforms on own rows,
  // barrier, gather (all rows), barrier.  Cleared positions are never line positions, so
  // the clear needs no barrier against the scatter or the previous gather.
  auto run_slot = [&](int s, cplx<T>(&p)[GPW]) {
#pragma unroll
    for (int k = 0; k < ZK; ++k)
      if (zofs[k] >= 0) buf[zofs[k]] = {T(0), T(0)};
#pragma unroll
    for (int k = 0; k < GPW; ++k) rows[lofs[k]] = cplx<T>{p[k].x, p[k].y};
    __syncthreads();
    if (s + XPF < s1) prefetch(s + XPF, p);

    // stage A, backward
#pragma unroll
    for (int ia = 0; ia < IA; ++ia) {
      const int t = lane + 64 * ia;
      if (t < TA) {
        const int r = t / N2, j2 = t % N2;
        cplx<T>* row = buf + r * LD + j2;
        cplx<T> a[N1];
#pragma unroll
        for (int j1 = 0; j1 < N1; ++j1)
          a[j1] = (j1 < ja || j1 > jb) ? row[S * j1] : cplx<T>{T(0), T(0)};
        sdft<N1, +1>(a);
        twiddle_row<N1, +1>(a, tw, j2);
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) row[S * k1] = a[k1];
      }
    }
    wave_sync();
    if constexpr (OP == X_STORE) {
      // stage B backward for every block of the wavefront's rows first (in registers), then
      // psi(r) back into the rows in natural order and out to the slot's grid: each
      // wavefront's rows are one contiguous stretch of the grid (z fixed, y consecutive)
      cplx<T> bs[IB][N2];
#pragma unroll
      for (int ib = 0; ib < IB; ++ib) {
        const int t = lane + 64 * ib;
        if (t < TB) {
          const int r = t / N1, k1 = t % N1;
          const cplx<T>* blk = buf + r * LD + S * k1;
#pragma unroll
          for (int j2 = 0; j2 < N2; ++j2) bs[ib][j2] = blk[j2];
          sdft<N2, +1>(bs[ib]);
        }
      }
      wave_sync();
#pragma unroll
      for (int ib = 0; ib < IB; ++ib) {
        const int t = lane + 64 * ib;
        if (t < TB) {
          const int r = t / N1, k1 = t % N1;
#pragma unroll
          for (int k2 = 0; k2 < N2; ++k2) buf[r * LD + PD::idx(k1 + N1 * k2)] = bs[ib][k2];
        }
      }
      wave_sync();
      const long ngrid = (long)n * np1 * np2;
      for (int e = lane; e < nyl * n; e += 64) {
        const int r = e / n, x = e - r * n;
        const cplx<T> q = buf[r * LD + PD::idx(x)];
        double* fp = out + 2 * ((long)s * ngrid + grow0 + e);
        fp[0] = double(q.x);
        fp[1] = double(q.y);
      }
      __syncthreads();  // the next scatter overwrites the rows read above
      return;
    }
    // stage B: backward DFT_N2 -> psi(r); multiply / accumulate; forward DFT_N2
    double f1 = 0, f2 = 0, b1 = 0, b2 = 0;
    if (OP == X_DENSITY) {
      if (cplx_density) {
        f1 = dens[s];
      } else {
        f1 = dens[4 * s + 0];
        f2 = dens[4 * s + 1];
        b1 = dens[4 * s + 2];
        b2 = dens[4 * s + 3];
      }
    }
#pragma unroll
    for (int ib = 0; ib < IB; ++ib) {
      const int t = lane + 64 * ib;
      if (t < TB) {
        const int r = t / N1, k1 = t % N1;
        cplx<T>* blk = buf + r * LD + S * k1;
        cplx<T> b[N2];
#pragma unroll
        for (int j2 = 0; j2 < N2; ++j2) b[j2] = blk[j2];
        sdft<N2, +1>(b);
        if (OP == X_VMUL) {
#pragma unroll
          for (int k2 = 0; k2 < N2; ++k2) b[k2] = {vr[ib][k2] * b[k2].x, vr[ib][k2] * b[k2].y};
          sdft<N2, -1>(b);
          if constexpr (TWB) {
            if (k1 != 0) {
#pragma unroll
              for (int m = 1; m < N2; ++m) b[m] = cmul(b[m], twb[ib][m - 1]);
            }
          } else {
            twiddle_row<N2, -1>(b, tw, k1);
          }
#pragma unroll
          for (int m2 = 0; m2 < N2; ++m2) blk[m2] = b[m2];
        } else {
#pragma unroll
          for (int k2 = 0; k2 < N2; ++k2) {
            const double re = double(b[k2].x), im = double(b[k2].y);
            if (cplx_density) {
              if (f1 > 0.0) acc[ib][k2] += f1 * (re * re + im * im);
            } else {
              // f = psi1 + i psi2, psi_j = X_j + conj(c_j(G=0)) (see k_xpass)
              if (f1 > 0.0) {
                const double r1 = re - b2;
                acc[ib][k2] += f1 * (r1 * r1 + b1 * b1);
              }
              if (f2 > 0.0) {
                const double r2 = im + b1;
                acc[ib][k2] += f2 * (r2 * r2 + b2 * b2);
              }
            }
          }
        }
      }
    }
    if (OP == X_DENSITY) {
      __syncthreads();  // the next scatter overwrites line positions read above
      return;
    }
    wave_sync();
    // stage A', forward: natural order
#pragma unroll
    for (int ia = 0; ia < IA; ++ia) {
      const int t = lane + 64 * ia;
      if (t < TA) {
        const int r = t / N2, m2 = t % N2;
        cplx<T>* row = buf + r * LD + m2;
        cplx<T> a[N1];
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) a[k1] = row[S * k1];
        sdft<N1, -1>(a);
        // FP32: all N1 outputs, the unneeded ones in [ja, jb] too (stage A of the next slot
        // never reads them and the gather reads line positions only), with no branch per
        // store: FP32 x pass 3.62 -> 3.41 ms at Si512; FP64 3.00 -> 3.06 ms, so FP64 skips
        // them
#pragma unroll
        for (int m1 = 0; m1 < N1; ++m1)
          if (sizeof(T) == 4 || m1 < ja || m1 > jb) row[S * m1] = a[m1];
      }
    }
    __syncthreads();
    {
      cplx<T>* dst = t2 + s * sstride;
#pragma unroll
      for (int k = 0; k < GPW; ++k)
        if ((gok >> k) & 1u) dst[gofs[k]] = rows[lofs[k]];
    }
    __syncthreads();
  };
  for (int s = s0; s < s1; s += XPF) {
#pragma unroll
    for (int d = 0; d < XPF; ++d)
      if (s + d < s1) run_slot(s + d, pf[d]);
  }
  if (OP == X_DENSITY) {
#pragma unroll
    for (int ib = 0; ib < IB; ++ib) {
      const int t = lane + 64 * ib;
      const int r = t / N1, k1 = t % N1;
      if (t < TB && r < nyl) {
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) out[grow0 + r * n + k1 + N1 * k2] += acc[ib][k2];
      }
    }
  }
}

// ------------------------------------------------------------------------------------
// z pass, forward: T1 columns -> z transform (exp(-i)) -> x 1/N -> unpack -> output.
// store 0: out[n][G] = ((out + add32) + kin[G]*psi[n][G]) + value;
// store 1: out[n][G] = value;
// store 2: out[n][G] = (add32 + kin[G]*psi[n][G]) + value, out not read (a cleared output:
//          the same sums as store 0 on zeros).
template <typename T, int MODE, int NMAX, int N1, int N2>
__global__ void __launch_bounds__(NTHR)
k_zpass_fwd(GridDev g, const cplx<T>* __restrict__ t1, const double* __restrict__ psi, long ld,
            int nst, int band0, double* __restrict__ out, long ldo,
            const double* __restrict__ kin, int store, const float* __restrict__ add32) {
  constexpr int EPL = Epl<ZLPW, NMAX>::v;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int np2 = g.np2;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int grp = blockIdx.x;
  const int slot = blockIdx.y;
  // this wavefront's lines and the rod its 16-lane group writes out, read before the
  // workgroup barrier so the table loads overlap the twiddle load and the transform
  const int e0 = g.grp_start[grp] + wave * ZLPW;
  const int nl = min(ZLPW, g.grp_start[grp + 1] - e0);
  const int le = lane >> 4;
  const int sub = lane & 15;
  int kind = 4, first = 0, lmin = 0, size = 0;
  if (le < nl) {
    kind = g.ent_kind[e0 + le];
    if (kind != 2 && kind != 4) {
      const int rod = g.ent_rod[e0 + le];
      first = g.rod_first[rod];
      lmin = g.rod_lmin[rod];
      size = g.rod_size[rod];
    }
  }
  const cplx<T>* twg = twtab<T>(g, 2);
  for (int i = threadIdx.x; i < np2; i += NTHR) tw[i] = twg[i];
  __syncthreads();
  if (nl <= 0) return;
  using PD = Pad<N1, N2>;
  const int LD = PD::row(np2);
  cplx<T>* buf = tw + np2 + wave * ZLPW * LD;
  for (int ll = 0; ll < nl; ++ll) {
    const int col = g.ent_col[e0 + ll];
    const cplx<T>* src = t1 + ((long)slot * g.ncol + (col < 0 ? 0 : col)) * np2;
    for (int z = lane; z < np2; z += 64)
      buf[ll * LD + PD::idx(z)] = col < 0 ? cplx<T>{T(0), T(0)} : src[z];
  }
  wave_sync();
  line_fft<-1, EPL, N1, N2, ZLPW>(buf, nl, LD, g.ax[2], tw, lane);

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
  if (le >= nl) return;
  if (kind == 2 || kind == 4) return;  // -G image: handled with its +G partner (next line)
  const cplx<T>* lp = buf + le * LD;
  const cplx<T>* lm = (kind == 1) ? buf + (le + 1) * LD : lp;
  double* o1 = out + 2 * (long)n1 * ldo;
  double* o2 = out + 2 * (long)(has2 ? n2 : n1) * ldo;
  const double* q1 = psi ? psi + 2 * (long)n1 * ld : nullptr;
  const double* q2 = psi ? psi + 2 * (long)(has2 ? n2 : n1) * ld : nullptr;
  // KZ coefficients per lane: the accumulate mode loads hpsi, psi, kin and the FP32 term
  // of all KZ first, then combines and stores them
  constexpr int KZ = 2;
  for (int i0 = sub; i0 < size; i0 += 16 * KZ) {
    double h[KZ][4], q[KZ][4], kk[KZ];
#pragma unroll
    for (int k = 0; k < KZ; ++k) {
      const int i = i0 + 16 * k;
      const int ig = first + i;
      kk[k] = 0.0;
#pragma unroll
      for (int j = 0; j < 4; ++j) h[k][j] = q[k][j] = 0.0;
      if (store == 1 || i >= size) continue;
      if (store == 0) {
        h[k][0] = o1[2 * ig];
        h[k][1] = o1[2 * ig + 1];
      }
      if (add32) {  // FP32 projector term, in the order of a separate add before this pass
        h[k][0] += double(add32[2 * ((long)n1 * g.ngw + ig)]);
        h[k][1] += double(add32[2 * ((long)n1 * g.ngw + ig) + 1]);
      }
      if (kin) {
        kk[k] = kin[ig];
        q[k][0] = q1[2 * ig];
        q[k][1] = q1[2 * ig + 1];
      }
      if (has2) {
        if (store == 0) {
          h[k][2] = o2[2 * ig];
          h[k][3] = o2[2 * ig + 1];
        }
        if (add32) {
          h[k][2] += double(add32[2 * ((long)n2 * g.ngw + ig)]);
          h[k][3] += double(add32[2 * ((long)n2 * g.ngw + ig) + 1]);
        }
        if (kin) {
          q[k][2] = q2[2 * ig];
          q[k][3] = q2[2 * ig + 1];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < KZ; ++k) {
      const int i = i0 + 16 * k;
      if (i >= size) continue;
      const int l = lmin + i;
      const int ig = first + i;
      const cplx<T> zp = lp[PD::idx(wrap(l, np2))];
      double r1, i1, r2 = 0.0, i2 = 0.0;
      if (MODE == SLOT_PAIR) {
        const cplx<T> zm = lm[PD::idx(wrap(-l, np2))];
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
      if (store == 1) {
        o1[2 * ig] = r1;
        o1[2 * ig + 1] = i1;
        if (has2) {
          o2[2 * ig] = r2;
          o2[2 * ig + 1] = i2;
        }
      } else {
        double h0 = h[k][0], h1 = h[k][1];
        if (kin) {
          h0 += kk[k] * q[k][0];
          h1 += kk[k] * q[k][1];
        }
        o1[2 * ig] = h0 + r1;
        o1[2 * ig + 1] = h1 + i1;
        if (has2) {
          double g0 = h[k][2], g1 = h[k][3];
          if (kin) {
            g0 += kk[k] * q[k][2];
            g1 += kk[k] * q[k][3];
          }
          o2[2 * ig] = g0 + r2;
          o2[2 * ig + 1] = g1 + i2;
        }
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

// grid edges with a two-factor register transform (must match the dispatch tables)
static void reg_factors(int n, int& f1, int& f2) {
  f1 = f2 = 0;
  switch (n) {
    case 28: f1 = 7; f2 = 4; break;
    case 32: f1 = 8; f2 = 4; break;
    case 36: f1 = 6; f2 = 6; break;
    case 40: f1 = 8; f2 = 5; break;
    case 48: f1 = 8; f2 = 6; break;
    case 50: f1 = 10; f2 = 5; break;
    case 54: f1 = 9; f2 = 6; break;
    case 60: f1 = 10; f2 = 6; break;
    case 64: f1 = 8; f2 = 8; break;
    case 72: f1 = 9; f2 = 8; break;
    case 80: f1 = 10; f2 = 8; break;
    case 90: f1 = 10; f2 = 9; break;
    case 96: f1 = 12; f2 = 8; break;
    case 100: f1 = 10; f2 = 10; break;
    case 108: f1 = 12; f2 = 9; break;
    case 120: f1 = 12; f2 = 10; break;
    case 128: f1 = 16; f2 = 8; break;
    case 140: f1 = 14; f2 = 10; break;
    case 144: f1 = 12; f2 = 12; break;
    case 160: f1 = 16; f2 = 10; break;
    case 168: f1 = 14; f2 = 12; break;
    case 180: f1 = 15; f2 = 12; break;
    case 192: f1 = 16; f2 = 12; break;
    case 210: f1 = 15; f2 = 14; break;
    case 224: f1 = 16; f2 = 14; break;
    case 240: f1 = 16; f2 = 15; break;
    case 256: f1 = 16; f2 = 16; break;
    default: break;
  }
}
// LDS row length (Pad<N1,N2,pad>::row) of an edge of length n
static int lds_row(int n, int pad = QBH_LDS_PAD) {
  int f1, f2;
  reg_factors(n, f1, f2);
  return (pad && f1 > 0 && f2 % 2 == 0) ? n + n / f2 : n;
}
template <typename T>
static size_t lds_z(const GridDev& g) {
  return sizeof(cplx<T>) * ((size_t)g.np2 + (size_t)NWAVE * ZLPW * lds_row(g.np2));
}
template <typename T>
static size_t lds_y(const GridDev& g) {
  // twiddles, ZB*LB lines, and (forward pass) one int per column of LB lines
  return sizeof(cplx<T>) * ((size_t)g.np1 + (size_t)Cfg<T>::ZB * Cfg<T>::LB * lds_row(g.np1)) +
         sizeof(int) * (size_t)Cfg<T>::LB * g.np1;
}

static int check_launch(Ctx* ctx, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ctx->last_error = std::string(what) + ": " + hipGetErrorString(e);
    return -3;
  }
  return 0;
}

// NMAX dispatch: 144 covers every grid edge up to 144 (9 elements per lane for 4-line
// wavefronts), 256 the rest
#define QBH_NMAX_DISPATCH(n, CALL) \
  do {                             \
    if ((n) <= 144) {              \
      constexpr int NM = 144;      \
      CALL;                        \
    } else {                       \
      constexpr int NM = 256;      \
      CALL;                        \
    }                              \
  } while (0)

// per-length dispatch for the z and y passes: F1*F2 = n selects the two-factor register
// transform; F1 = 0 the LDS Stockham passes (also for every length when the basis' tuning
// turns the register transforms off)
#define QBH_SIZE_DISPATCH(n, CALL)                       \
  do {                                                   \
    switch (b->tune.zy_reg ? (n) : 0) {                                         \
      case 28: {                                         \
        constexpr int NM = 144, F1 = 7, F2 = 4;          \
        CALL;                                            \
      } break;                                           \
      case 32: {                                   \
        constexpr int NM = 144, F1 = 8, F2 = 4;    \
        CALL;                                      \
      } break;                                     \
      case 40: {                                   \
        constexpr int NM = 144, F1 = 8, F2 = 5;    \
        CALL;                                      \
      } break;                                     \
      case 48: {                                   \
        constexpr int NM = 144, F1 = 8, F2 = 6;    \
        CALL;                                      \
      } break;                                     \
      case 50: {                                   \
        constexpr int NM = 144, F1 = 10, F2 = 5;   \
        CALL;                                      \
      } break;                                     \
      case 54: {                                   \
        constexpr int NM = 144, F1 = 9, F2 = 6;    \
        CALL;                                      \
      } break;                                     \
      case 80: {                                   \
        constexpr int NM = 144, F1 = 10, F2 = 8;   \
        CALL;                                      \
      } break;                                     \
      case 100: {                                  \
        constexpr int NM = 144, F1 = 10, F2 = 10;  \
        CALL;                                      \
      } break;                                     \
      case 108: {                                  \
        constexpr int NM = 144, F1 = 12, F2 = 9;   \
        CALL;                                      \
      } break;                                     \
      case 128: {                                  \
        constexpr int NM = 144, F1 = 16, F2 = 8;   \
        CALL;                                      \
      } break;                                     \
      case 36: {                                         \
        constexpr int NM = 144, F1 = 6, F2 = 6;          \
        CALL;                                            \
      } break;                                           \
      case 60: {                                         \
        constexpr int NM = 144, F1 = 10, F2 = 6;         \
        CALL;                                            \
      } break;                                           \
      case 64: {                                         \
        constexpr int NM = 144, F1 = 8, F2 = 8;          \
        CALL;                                            \
      } break;                                           \
      case 72: {                                         \
        constexpr int NM = 144, F1 = 9, F2 = 8;          \
        CALL;                                            \
      } break;                                           \
      case 90: {                                         \
        constexpr int NM = 144, F1 = 10, F2 = 9;         \
        CALL;                                            \
      } break;                                           \
      case 96: {                                         \
        constexpr int NM = 144, F1 = 12, F2 = 8;         \
        CALL;                                            \
      } break;                                           \
      case 120: {                                        \
        constexpr int NM = 144, F1 = 12, F2 = 10;        \
        CALL;                                            \
      } break;                                           \
      case 140: {                                        \
        constexpr int NM = 144, F1 = 14, F2 = 10;        \
        CALL;                                            \
      } break;                                           \
      case 144: {                                        \
        constexpr int NM = 144, F1 = 12, F2 = 12;        \
        CALL;                                            \
      } break;                                           \
      case 168: {                                  \
        constexpr int NM = 144, F1 = 14, F2 = 12;  \
        CALL;                                      \
      } break;                                     \
      case 210: {                                  \
        constexpr int NM = 144, F1 = 15, F2 = 14;  \
        CALL;                                      \
      } break;                                     \
      case 224: {                                  \
        constexpr int NM = 144, F1 = 16, F2 = 14;  \
        CALL;                                      \
      } break;                                     \
      case 160: {                                  \
        constexpr int NM = 144, F1 = 16, F2 = 10;  \
        CALL;                                      \
      } break;                                     \
      case 180: {                                  \
        constexpr int NM = 144, F1 = 15, F2 = 12;  \
        CALL;                                      \
      } break;                                     \
      case 192: {                                  \
        constexpr int NM = 144, F1 = 16, F2 = 12;  \
        CALL;                                      \
      } break;                                     \
      case 240: {                                  \
        constexpr int NM = 144, F1 = 16, F2 = 15;  \
        CALL;                                      \
      } break;                                     \
      case 256: {                                  \
        constexpr int NM = 144, F1 = 16, F2 = 16;  \
        CALL;                                      \
      } break;                                     \
      default:                                           \
        if ((n) <= 144) {                                \
          constexpr int NM = 144, F1 = 0, F2 = 0;        \
          CALL;                                          \
        } else {                                         \
          constexpr int NM = 256, F1 = 0, F2 = 0;        \
          CALL;                                          \
        }                                                \
    }                                                    \
  } while (0)

template <typename T>
static int bwd_yz_t(Ctx* ctx, Basis* b, const PipelineArgs& a) {
  const GridDev& g = b->g;
  auto* t1 = static_cast<cplx<T>*>(a.t1);
  auto* t2 = static_cast<cplx<T>*>(a.t2);
  {
    TimedLaunch tl(ctx, "zpass_bwd");
    dim3 grid(g.ngroups, a.nslots);
    const size_t lds = lds_z<T>(g);
    if (a.mode == SLOT_PAIR)
      QBH_SIZE_DISPATCH(g.np2, (k_zpass_bwd<T, SLOT_PAIR, NM, F1, F2><<<grid, NTHR, lds, ctx->stream>>>(
                                   g, a.psi, a.ld, a.nst, a.band0, t1)));
    else if (a.mode == SLOT_SINGLE_REAL)
      QBH_SIZE_DISPATCH(g.np2,
                        (k_zpass_bwd<T, SLOT_SINGLE_REAL, NM, F1, F2><<<grid, NTHR, lds, ctx->stream>>>(
                            g, a.psi, a.ld, a.nst, a.band0, t1)));
    else
      QBH_SIZE_DISPATCH(g.np2, (k_zpass_bwd<T, SLOT_COMPLEX, NM, F1, F2><<<grid, NTHR, lds, ctx->stream>>>(
                                   g, a.psi, a.ld, a.nst, a.band0, t1)));
  }
  if (int r = check_launch(ctx, "zpass_bwd")) return r;
  {
    TimedLaunch tl(ctx, "ypass_bwd");
    constexpr int ZB = Cfg<T>::ZB, LB = Cfg<T>::LB;
    dim3 grid(((g.np2 + ZB - 1) / ZB) * ((g.nlx + LB - 1) / LB), a.nslots);
    QBH_SIZE_DISPATCH(g.np1, (k_ypass_bwd<T, NM, F1, F2><<<grid, NTHR, lds_y<T>(g), ctx->stream>>>(
                                 g, t1, t2)));
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
    constexpr int ZB = Cfg<T>::ZB, LB = Cfg<T>::LB;
    dim3 grid(((g.np2 + ZB - 1) / ZB) * ((g.nlx + LB - 1) / LB), a.nslots);
    QBH_SIZE_DISPATCH(g.np1, (k_ypass_fwd<T, NM, F1, F2><<<grid, NTHR, lds_y<T>(g), ctx->stream>>>(
                                 g, t2, t1)));
  }
  if (int r = check_launch(ctx, "ypass_fwd")) return r;
  {
    if (a.wait_z) hipStreamWaitEvent(ctx->stream, a.wait_z, 0);
    TimedLaunch tl(ctx, "zpass_fwd");
    dim3 grid(g.ngroups, a.nslots);
    const size_t lds = lds_z<T>(g);
    if (a.mode == SLOT_PAIR)
      QBH_SIZE_DISPATCH(g.np2, (k_zpass_fwd<T, SLOT_PAIR, NM, F1, F2><<<grid, NTHR, lds, ctx->stream>>>(
                                   g, t1, a.psi, a.ld, a.nst, a.band0, out, ldo, kin, store, a.add32)));
    else if (a.mode == SLOT_SINGLE_REAL)
      QBH_SIZE_DISPATCH(g.np2,
                        (k_zpass_fwd<T, SLOT_SINGLE_REAL, NM, F1, F2><<<grid, NTHR, lds, ctx->stream>>>(
                            g, t1, a.psi, a.ld, a.nst, a.band0, out, ldo, kin, store, a.add32)));
    else
      QBH_SIZE_DISPATCH(g.np2, (k_zpass_fwd<T, SLOT_COMPLEX, NM, F1, F2><<<grid, NTHR, lds, ctx->stream>>>(
                                   g, t1, a.psi, a.ld, a.nst, a.band0, out, ldo, kin, store, a.add32)));
  }
  return check_launch(ctx, "zpass_fwd");
}

int pipeline_fwd_zy_accumulate(Ctx* ctx, Basis* b, const PipelineArgs& a, double* out, long ldo,
                               const double* kin, int store) {
  return a.prec == PREC_FP64 ? fwd_zy_t<double>(ctx, b, a, out, ldo, kin, store)
                             : fwd_zy_t<float>(ctx, b, a, out, ldo, kin, store);
}

// slots per x-pass workgroup: keep >= 4096 workgroups in flight while amortising the
// potential rows and twiddles over several slots.  Measured on the Si64 complex basis (72^3,
// 128 slots): 4 slots per workgroup (11520 workgroups) 0.760/0.516 ms per FP64/FP32 Hψ, 8
// (5760) 0.747/0.495 ms, 16 (2880) 0.754/0.484 ms; Si512 and Al500 reach the cap of 16.
static int x_slots_per_wg(const GridDev& g, int nslots, int yb) {
  const long rows = (long)g.np2 * ((g.np1 + yb - 1) / yb);
  int spw = 1;
  while (spw < 16 && rows * ((nslots + 2 * spw - 1) / (2 * spw)) >= 4096) spw *= 2;
  return spw;
}

static bool has_x2(int n) {
  int f1, f2;
  reg_factors(n, f1, f2);
  return f1 > 0;
}

bool register_edge(int n) { return has_x2(n); }

// two-factor register x pass for the grid edges it is instantiated for
template <typename T, int OP>
static bool launch_x2(Ctx* ctx, const GridDev& g, dim3 grid, size_t lds, cplx<T>* t2, int nslots,
                      int spw, const double* in, double* out, const double* dens, int cd) {
  if constexpr (OP != X_VMUL && OP != X_DENSITY && OP != X_STORE) {
    return false;
  } else {
  // line elements per thread: a short count for spheres up to about half the edge (the
  // usual Hψ grid), the full edge otherwise
  constexpr int YB = Cfg<T>::XR * NWAVE;
  const int need = (g.nlx * YB + NTHR - 1) / NTHR;
#define QBH_X2(NN, A, B)                                                                      \
  case NN: {                                                                                    \
    constexpr int GS = (YB * (NN / 2 + 8) + NTHR - 1) / NTHR, GF = (YB * NN + NTHR - 1) / NTHR; \
    if (need <= GS)                                                                             \
      k_xpass2<T, OP, A, B, GS><<<grid, NTHR, lds, ctx->stream>>>(g, t2, nslots, spw, in, out, \
                                                                  dens, cd);                    \
    else                                                                                        \
      k_xpass2<T, OP, A, B, GF><<<grid, NTHR, lds, ctx->stream>>>(g, t2, nslots, spw, in, out, \
                                                                  dens, cd);                    \
    return true;                                                                                \
  }
  switch (g.np0) {
    QBH_X2(28, 7, 4)
    QBH_X2(32, 8, 4)
    QBH_X2(40, 8, 5)
    QBH_X2(48, 8, 6)
    QBH_X2(50, 10, 5)
    QBH_X2(54, 9, 6)
    QBH_X2(80, 10, 8)
    QBH_X2(100, 10, 10)
    QBH_X2(108, 12, 9)
    QBH_X2(128, 16, 8)
    QBH_X2(36, 6, 6)
    QBH_X2(60, 10, 6)
    QBH_X2(64, 8, 8)
    QBH_X2(72, 9, 8)
    QBH_X2(90, 10, 9)
    QBH_X2(96, 12, 8)
    QBH_X2(120, 12, 10)
    QBH_X2(140, 14, 10)
    QBH_X2(144, 12, 12)
    QBH_X2(160, 16, 10)
    QBH_X2(168, 14, 12)
    QBH_X2(210, 15, 14)
    QBH_X2(224, 16, 14)
    QBH_X2(180, 15, 12)
    QBH_X2(192, 16, 12)
    QBH_X2(240, 16, 15)
    QBH_X2(256, 16, 16)
    default:
      return false;
  }
#undef QBH_X2
  }
}

template <typename T, int OP>
static int xpass_t(Ctx* ctx, Basis* b, const PipelineArgs& a, const double* in, double* out,
                   const double* dens, double scale, int cplx_density, const char* name) {
  const GridDev& g = b->g;
  auto* t2 = static_cast<cplx<T>*>(a.t2);
  const bool reg = (OP == X_VMUL || OP == X_DENSITY || OP == X_STORE) && has_x2(g.np0) &&
                   b->tune.x_reg;
  const int rows = reg ? Cfg<T>::XR : Cfg<T>::LPW;  // rows per wavefront
  const int yb = rows * NWAVE;
  const int nyb = (g.np1 + yb - 1) / yb;
  // twiddles, row buffers and (register x pass) one dump entry per thread
  const size_t lds =
      sizeof(cplx<T>) * ((size_t)g.np0 + (size_t)NWAVE * rows *
                                             (reg ? lds_row(g.np0, QBH_XLDS_PAD) : g.np0) +
                         (reg ? NTHR : 0));
  int spw = (OP == X_DENSITY) ? a.nslots
            : b->tune.x_spw > 0 ? std::min(b->tune.x_spw, a.nslots)
                                : x_slots_per_wg(g, a.nslots, yb);
  dim3 grid(g.np2 * nyb, (a.nslots + spw - 1) / spw);
  {
    TimedLaunch tl(ctx, name);
    if (!(reg && launch_x2<T, OP>(ctx, g, grid, lds, t2, a.nslots, spw, in, out, dens,
                                  cplx_density)))
      QBH_NMAX_DISPATCH(g.np0, (k_xpass<T, OP, NM><<<grid, NTHR, lds, ctx->stream>>>(
                                   g, t2, a.nslots, spw, in, out, dens, scale, cplx_density)));
  }
  return check_launch(ctx, name);
}

int pipeline_x_vmul(Ctx* ctx, Basis* b, const PipelineArgs& a, const double* v) {
  return a.prec == PREC_FP64
             ? xpass_t<double, X_VMUL>(ctx, b, a, v, nullptr, nullptr, 1.0, 0, "xpass_vmul")
             : xpass_t<float, X_VMUL>(ctx, b, a, v, nullptr, nullptr, 1.0, 0, "xpass_vmul");
}
int pipeline_x_store(Ctx* ctx, Basis* b, const PipelineArgs& a, double* f) {
  return a.prec == PREC_FP64
             ? xpass_t<double, X_STORE>(ctx, b, a, nullptr, f, nullptr, 1.0, 0, "xpass_store")
             : xpass_t<float, X_STORE>(ctx, b, a, nullptr, f, nullptr, 1.0, 0, "xpass_store");
}
int pipeline_x_load(Ctx* ctx, Basis* b, const PipelineArgs& a, const double* f, int real_input,
                    double scale) {
  if (real_input)
    return a.prec == PREC_FP64
               ? xpass_t<double, X_LOAD_REAL>(ctx, b, a, f, nullptr, nullptr, scale, 0,
                                              "xpass_load")
               : xpass_t<float, X_LOAD_REAL>(ctx, b, a, f, nullptr, nullptr, scale, 0,
                                             "xpass_load");
  return a.prec == PREC_FP64
             ? xpass_t<double, X_LOAD>(ctx, b, a, f, nullptr, nullptr, 1.0, 0, "xpass_load")
             : xpass_t<float, X_LOAD>(ctx, b, a, f, nullptr, nullptr, 1.0, 0, "xpass_load");
}
int pipeline_x_density(Ctx* ctx, Basis* b, const PipelineArgs& a, const double* dens,
                       double* rho) {
  const int cd = a.mode == SLOT_COMPLEX ? 1 : 0;
  return a.prec == PREC_FP64
             ? xpass_t<double, X_DENSITY>(ctx, b, a, nullptr, rho, dens, 1.0, cd, "xpass_density")
             : xpass_t<float, X_DENSITY>(ctx, b, a, nullptr, rho, dens, 1.0, cd, "xpass_density");
}

}  // namespace qbh
