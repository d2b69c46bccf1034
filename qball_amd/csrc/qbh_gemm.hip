// Projector contractions on the matrix cores (gfx950): the two real GEMMs of the Γ
// Kleinman-Bylander path, NonLocalPotential::energy (src/qball/NonLocalPotential.cc):
//
//   fnl_gemm (:2053)  F[p, n] = sum_k anl[k, p] psi[k, n]          (reduction over K = 2 ngw)
//   enl_hpsi (:2158)  hpsi[k, n] (+)= sum_p anl[k, p] F[p, n]       (reduction over P)
//
// Both are written as C[r, c] = sum_q X[r, q] Y[q, c] with X read "q-contiguous"
// (X[r, q] at X[q + r ldx]): for fnl, r = p, q = k, c = n, X = anl and Y = psi, both
// q-contiguous (Y[q, c] at Y[q + c ldy]); for enl the output is taken transposed, r = n,
// c = k, q = p, X = F and Y = anl, which is c-contiguous (Y[q, c] at Y[c + q ldy]), so that
// 16 lanes of one matrix-core result row store 16 consecutive coefficients of a band.
//
// Workgroup: 256 threads = 4 waves in a 2 x 2 arrangement, a 128 x 128 tile of C (each
// wave 64 x 64 = 4 x 4 blocks of the 16 x 16 x 4 MFMA, 64 accumulators per lane), the
// reduction in stages 16 deep double-buffered through LDS: each stage's global loads are
// in flight while the previous stage's MFMAs run (64 per wave), one barrier per stage, two
// workgroups per CU in FP64 and three in FP32.  The operands are read through buffer
// descriptors (a scalar base per stage and a fixed 32-bit offset per lane, which frees the
// 64-bit address registers of pointer loads: FP64 9.7 / 9.9 ms against 10.1 / 10.0 ms,
// FP32 4.8 / 4.8 against 5.6 / 5.3 at the 512-atom shapes below).  Loads past the ends of
// the operands read a zero pair instead (no branches, no fix-up), so the compiler keeps all
// of a stage's loads in flight together.  The fnl product splits the reduction into slices
// (the 2048 x 1024 output of 512 Si atoms has only 128 tiles for 256 CUs), each writing its
// own partial, which the finish kernel sums.  Blocks are numbered so that the workgroups
// that run together on one XCD (blocks are dealt round-robin to the 8 XCDs) share a tile of
// the projector table in that XCD's L2.
//
// Lane layout of the 16 x 16 x 4 MFMA (measured on gfx950): lane i supplies A[i mod 16][i / 16]
// and B[i / 16][i mod 16]; lane i receives in accumulator j the result row 4 (i / 16) + j
// (FP32) or 4 j + i / 16 (FP64) of column i mod 16.  Any permutation of the reduction
// index that A and B share gives the same sum, so lane group g = i / 16 supplies
// reduction index (BQ / 4) g + j of a stage in MFMA step j and reads its values for several
// steps with one 16-byte LDS load.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "qbh_gemm.h"

namespace qbh {

namespace {

constexpr int kBR = 128, kBC = 128, kThreads = 256;

// stage depth (reduction elements per LDS stage) and waves per SIMD, per element type
// (measured at the 512-atom Si shapes, F = anl^T psi / hpsi += anl F, buffer loads: FP64
// 16-deep stages at two waves per SIMD 9.7 / 9.9 ms; FP32 16-deep 4.8 / 4.8 ms, 32-deep
// 4.8 / 5.1 ms.  Pointer loads: FP64 32-deep at one wave 12.4 / 12.6 ms, FP32 at four waves
// per SIMD (spilling) 8.2 / 8.7 ms)
#ifndef QBH_GEMM_BQ64
#define QBH_GEMM_BQ64 16
#endif
#ifndef QBH_GEMM_WPE64
#define QBH_GEMM_WPE64 2
#endif
#ifndef QBH_GEMM_BQ32
#define QBH_GEMM_BQ32 16
#endif
#ifndef QBH_GEMM_WPE32
#define QBH_GEMM_WPE32 2
#endif
#ifndef QBH_GEMM_BUF
#define QBH_GEMM_BUF 1
#endif
#ifndef QBH_GEMM_DIRECT
#define QBH_GEMM_DIRECT 1
#endif
constexpr int kBQ = QBH_GEMM_BQ64 > QBH_GEMM_BQ32 ? QBH_GEMM_BQ64 : QBH_GEMM_BQ32;  // slice unit


// Per element type: matrix-core operation, element pair, the 16-byte fragment vector (VN
// reduction steps per LDS load), the row of an accumulator, stage depth BQ (16 or 32),
// LDS row strides (elements) and waves per SIMD.  Strides: q-contiguous tiles hold BQ
// reduction elements per row plus a pad, c-contiguous tiles 128 columns plus a pad, chosen
// so that the fragment loads of a wave are free of bank conflicts with 32 or with 64
// four-byte banks.
template <typename T> struct Mf;
template <> struct Mf<double> {
  typedef double acc __attribute__((ext_vector_type(4)));
  typedef double2 v2;
  static constexpr int VN = 2;
  typedef double vn __attribute__((ext_vector_type(2)));
  static constexpr int BQ = QBH_GEMM_BQ64, XS = BQ + 2, CS = kBC + (BQ == 32 ? 2 : 4);
  static constexpr int WPE = QBH_GEMM_WPE64;
  __device__ static acc op(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // result row of accumulator j in lane group g
  __device__ static int row(int g, int j) { return 4 * j + g; }
};
template <> struct Mf<float> {
  typedef float acc __attribute__((ext_vector_type(4)));
  typedef float2 v2;
  static constexpr int VN = 4;
  typedef float vn __attribute__((ext_vector_type(4)));
  static constexpr int BQ = QBH_GEMM_BQ32, XS = BQ + 4, CS = kBC + (BQ == 32 ? 2 : 4);
  static constexpr int WPE = QBH_GEMM_WPE32;
  __device__ static acc op(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int g, int j) { return 4 * g + j; }
};

struct Args {
  int R, C;
  long Q;
  const void* X;
  long ldx;
  const void* Y;
  long ldy;
  void* O;
  long ldo;
  long qslice;   // reduction elements per slice (multiple of kBQ)
  long ostride;  // elements between the slices' outputs
  int tiles_r, tiles_c;
  int nblocks;   // tiles_r * tiles_c * slices (the grid is padded to a multiple of 8)
  int r_fastest; // consecutive blocks of one XCD step along r (else along c)
  int accumulate;
};

// zeros that out-of-range lanes load instead of operand elements
__device__ __attribute__((aligned(16))) double g_gemm_zero[2] = {0.0, 0.0};

// Element pair x[i], x[i + 1], each read only where ok && its index is below i + lim_rel
// and zero otherwise.  Branch-free and without a fix-up after the load: an out-of-range
// element is loaded from g_gemm_zero, so a stage's loads stay in flight while the previous
// stage's MFMAs run.  VEC: one 2-element vector load; the caller guarantees that lim_rel
// is never 1 and that i is even.
template <typename T, bool VEC>
__device__ inline typename Mf<T>::v2 ld2(const T* __restrict__ x, long i, long lim_rel, bool ok) {
  using v2 = typename Mf<T>::v2;
  const T* z = reinterpret_cast<const T*>(g_gemm_zero);
  if (VEC) {
    const bool in = ok && lim_rel >= 2;
    return *reinterpret_cast<const v2*>(in ? x + i : z);
  }
  const bool in0 = ok && lim_rel >= 1, in1 = ok && lim_rel >= 2;
  v2 v;
  v.x = *(in0 ? x + i : z);
  v.y = *(in1 ? x + i + 1 : z);
  return v;
}

// Buffer addressing: a descriptor per operand and stage (base, record count in bytes),
// one 32-bit byte offset per lane and element pair.  A lane whose pair lies outside the
// operand takes an offset past the record count, and the load returns zeros; the record
// count is the stage's byte extent, below kOutside.
constexpr unsigned kOutside = 0x80000000u;
constexpr int kRsrcWord3 = 0x00020000;  // raw dword access (gfx9 family)

template <typename T>
__device__ inline typename Mf<T>::v2 bload2(__amdgpu_buffer_rsrc_t r, unsigned off) {
  using v2 = typename Mf<T>::v2;
  if constexpr (sizeof(T) == 8) {
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    const u4 u = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2 d = __builtin_bit_cast(d2, u);
    v2 o;
    o.x = d.x;
    o.y = d.y;
    return o;
  } else {
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    const u2 u = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    typedef float f2 __attribute__((ext_vector_type(2)));
    const f2 f = __builtin_bit_cast(f2, u);
    v2 o;
    o.x = f.x;
    o.y = f.y;
    return o;
  }
}

__device__ inline __amdgpu_buffer_rsrc_t rsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                           kRsrcWord3);
}

// YC: Y is c-contiguous.  OUTC: the output is c-contiguous (O[c + r ldo]), else
// r-contiguous (O[r + c ldo]).  XV/YV: the operand's element pairs are aligned vectors that
// never straddle the end of the range.
//
// Reduction order: within a stage of BQ, lane group g supplies reduction index
// (BQ / 4) g + j in MFMA step j, for the X and the Y fragment alike, so each lane reads its
// X values of a stage as BQ / 4 consecutive elements of one row.
//
// BUF (with XV and YV): the operands are read through buffer descriptors based at the
// stage's first element, with per-lane offsets fixed for the whole reduction; the host
// selects it when one stage of either operand spans less than kOutside bytes.
template <typename T, bool YC, bool OUTC, bool XV, bool YV, bool BUF>
__global__ void __launch_bounds__(kThreads, Mf<T>::WPE) k_gemm(Args a) {
  using M = Mf<T>;
  using v2 = typename M::v2;
  using vn = typename M::vn;
  using acc_t = typename M::acc;
  constexpr int VN = M::VN;
  constexpr int BQ = M::BQ;                    // stage depth
  constexpr int QG = BQ / 4;                   // reduction indices per lane group and stage
  constexpr int XS = M::XS, CS = M::CS;
  constexpr int XTILE = kBR * XS;
  constexpr int YTILE = YC ? BQ * CS : kBC * XS;
  constexpr int NP = kBR * BQ / 2 / kThreads;  // element pairs per thread and operand
  constexpr int PPR = BQ / 2;                  // pairs per q-contiguous row
  constexpr int PPC = kBC / 2;                  // pairs per c-contiguous row (64)
  __shared__ __attribute__((aligned(16))) T lds[2 * (XTILE + YTILE)];

  const int nb8 = gridDim.x >> 3;
  const int bid = blockIdx.x;
  const int wid = (bid & 7) * nb8 + (bid >> 3);
  if (wid >= a.nblocks) return;
  const int tiles = a.tiles_r * a.tiles_c;
  const int s = wid / tiles;
  const int t = wid - s * tiles;
  int tr, tc;
  if (a.r_fastest) {
    tr = t % a.tiles_r;
    tc = t / a.tiles_r;
  } else {
    tc = t % a.tiles_c;
    tr = t / a.tiles_c;
  }
  const int r0 = tr * kBR, c0 = tc * kBC;
  const long q0 = (long)s * a.qslice;
  const long q1 = min(a.Q, q0 + a.qslice);

  const T* __restrict__ X = static_cast<const T*>(a.X);
  const T* __restrict__ Y = static_cast<const T*>(a.Y);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int l16 = lane & 15, lg = lane >> 4;

  // buffer addressing: byte offsets of this lane's pairs from the stage base, and where
  // each pair sits along the reduction (qx, qy) and across the tile (rows, columns)
  unsigned xo[NP], yo[NP];
  int xq[NP], yq[NP];
  bool xin[NP], yin[NP];
  unsigned xext = 0, yext = 0;
  if constexpr (BUF) {
#pragma unroll
    for (int it = 0; it < NP; ++it) {
      const int v = it * kThreads + tid;
      const int row = v / PPR, qv = (v % PPR) * 2;
      xo[it] = (unsigned)((qv + (long)row * a.ldx) * (long)sizeof(T));
      xq[it] = qv;
      xin[it] = r0 + row < a.R;
      if (YC) {
        const int qr = v / PPC, cv = (v % PPC) * 2;
        yo[it] = (unsigned)((cv + (long)qr * a.ldy) * (long)sizeof(T));
        yq[it] = qr;
        yin[it] = c0 + cv < a.C;  // C even: the pair is whole
      } else {
        const int col = v / PPR, qv2 = (v % PPR) * 2;
        yo[it] = (unsigned)((qv2 + (long)col * a.ldy) * (long)sizeof(T));
        yq[it] = qv2;
        yin[it] = c0 + col < a.C;
      }
    }
    const long rows = min(kBR, a.R - r0), cols = min(kBC, a.C - c0);
    xext = (unsigned)(((rows - 1) * a.ldx + BQ) * (long)sizeof(T));
    yext = YC ? (unsigned)(((BQ - 1) * a.ldy + cols) * (long)sizeof(T))
              : (unsigned)(((cols - 1) * a.ldy + BQ) * (long)sizeof(T));
  }

  // staging: NP element pairs of X and of Y per thread and stage
  auto gload = [&](v2 (&xs)[NP], v2 (&ys)[NP], long qb) {
    if constexpr (BUF) {
      const long left = q1 - qb;  // reduction elements left from the stage start
      const __amdgpu_buffer_rsrc_t rx = rsrc(X + qb + (long)r0 * a.ldx, xext);
      const __amdgpu_buffer_rsrc_t ry =
          rsrc(YC ? Y + c0 + qb * a.ldy : Y + qb + (long)c0 * a.ldy, yext);
#pragma unroll
      for (int it = 0; it < NP; ++it)
        xs[it] = bload2<T>(rx, xin[it] && xq[it] + 2 <= left ? xo[it] : kOutside);
#pragma unroll
      for (int it = 0; it < NP; ++it) {
        const bool in = YC ? yq[it] < left : yq[it] + 2 <= left;
        ys[it] = bload2<T>(ry, yin[it] && in ? yo[it] : kOutside);
      }
      return;
    }
#pragma unroll
    for (int it = 0; it < NP; ++it) {
      const int v = it * kThreads + tid;
      const int row = v / PPR, qv = (v % PPR) * 2;
      const int r = r0 + row;
      const long q = qb + qv;
      xs[it] = ld2<T, XV>(X, q + (long)r * a.ldx, q1 - q, r < a.R);
    }
#pragma unroll
    for (int it = 0; it < NP; ++it) {
      const int v = it * kThreads + tid;
      if (YC) {
        const int qr = v / PPC, cv = (v % PPC) * 2;
        const long q = qb + qr;
        const int c = c0 + cv;
        ys[it] = ld2<T, YV>(Y, (long)c + q * a.ldy, (long)a.C - c, q < q1);
      } else {
        const int col = v / PPR, qv = (v % PPR) * 2;
        const int c = c0 + col;
        const long q = qb + qv;
        ys[it] = ld2<T, YV>(Y, q + (long)c * a.ldy, q1 - q, c < a.C);
      }
    }
  };
  auto lstore = [&](const v2 (&xs)[NP], const v2 (&ys)[NP], int buf) {
    T* xl = lds + buf * (XTILE + YTILE);
    T* yl = xl + XTILE;
#pragma unroll
    for (int it = 0; it < NP; ++it) {
      const int v = it * kThreads + tid;
      *reinterpret_cast<v2*>(xl + (v / PPR) * XS + (v % PPR) * 2) = xs[it];
      if (YC)
        *reinterpret_cast<v2*>(yl + (v / PPC) * CS + (v % PPC) * 2) = ys[it];
      else
        *reinterpret_cast<v2*>(yl + (v / PPR) * XS + (v % PPR) * 2) = ys[it];
    }
  };

  acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  auto compute = [&](int buf) {
    const T* xl = lds + buf * (XTILE + YTILE);
    const T* yl = xl + XTILE;
    // VN MFMA steps per 16-byte fragment load
#pragma unroll
    for (int jv = 0; jv < QG / VN; ++jv) {
      vn xf[4], yf[4];
#pragma unroll
      for (int bi = 0; bi < 4; ++bi)
        xf[bi] = *reinterpret_cast<const vn*>(xl + (wr * 64 + bi * 16 + l16) * XS + QG * lg +
                                              VN * jv);
#pragma unroll
      for (int bj = 0; bj < 4; ++bj) {
        if (YC) {
          const T* p = yl + (QG * lg + VN * jv) * CS + wc * 64 + bj * 16 + l16;
#pragma unroll
          for (int e = 0; e < VN; ++e) yf[bj][e] = p[e * CS];
        } else {
          yf[bj] = *reinterpret_cast<const vn*>(yl + (wc * 64 + bj * 16 + l16) * XS + QG * lg +
                                                VN * jv);
        }
      }
#pragma unroll
      for (int e = 0; e < VN; ++e)
#pragma unroll
        for (int bi = 0; bi < 4; ++bi)
#pragma unroll
          for (int bj = 0; bj < 4; ++bj)
            acc[bi][bj] = M::op(xf[bi][e], yf[bj][e], acc[bi][bj]);
    }
  };

  v2 xa[NP], ya[NP];
  if (q0 < q1) {
    gload(xa, ya, q0);
    lstore(xa, ya, 0);
  }
  __syncthreads();
  int buf = 0;
  for (long qb = q0; qb < q1; qb += BQ) {
    const bool more = qb + BQ < q1;
    if (more) gload(xa, ya, qb + BQ);
    compute(buf);
    if (more) lstore(xa, ya, buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

  // epilogue: lane holds C[rb + M::row(lg, j)][cb + l16] of each block
  T* O = static_cast<T*>(a.O) + (long)s * a.ostride;
#pragma unroll
  for (int bi = 0; bi < 4; ++bi) {
    const int rb = r0 + wr * 64 + bi * 16;
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) {
      const int c = c0 + wc * 64 + bj * 16 + l16;
      if (c >= a.C) continue;
      const acc_t v = acc[bi][bj];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = rb + M::row(lg, j);
        if (r < a.R) {
          T* o = OUTC ? O + c + (long)r * a.ldo : O + r + (long)c * a.ldo;
          *o = a.accumulate ? *o + v[j] : v[j];
        }
      }
    }
  }
}

// Direct form (FP64 enl product): no LDS.  Each wave loads its own fragments from global
// memory through buffer descriptors into a ring of register stages (8 reduction elements
// each), two stages in flight while a third feeds the MFMAs, so there is no barrier and no
// LDS store.  The Y operand (the projector table, c-contiguous) is read one reduction row at
// a time, 16 lanes taking 16 consecutive elements of one 128-byte line.  At the 512-atom
// shape: 9.2 ms against 9.9 ms for the LDS form; F, whose q-contiguous rows it would read in
// 64-byte pieces, and FP32 are slower in this form (12.2 / 8.7 / 5.8 against 9.9 / 4.8 /
// 4.8 ms) and keep the LDS form.
template <typename T>
__device__ inline T bload1(__amdgpu_buffer_rsrc_t r, unsigned off) {
  if constexpr (sizeof(T) == 8) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  } else {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
  }
}

template <typename T, bool YC, bool OUTC>
__global__ void __launch_bounds__(kThreads, 2) k_gemm_direct(Args a) {
  using M = Mf<T>;
  using v2 = typename M::v2;
  using acc_t = typename M::acc;
  constexpr int QG = sizeof(T) == 8 ? 2 : 4;  // reduction values per lane, block and stage
  constexpr int BQ = 4 * QG;
  constexpr int NP = QG / 2;
  constexpr unsigned SZ = sizeof(T);
  struct Stage {
    T x[4][QG];
    T y[4][QG];
  };
  const int nb8 = gridDim.x >> 3;
  const int bid = blockIdx.x;
  const int wid = (bid & 7) * nb8 + (bid >> 3);
  if (wid >= a.nblocks) return;
  const int tiles = a.tiles_r * a.tiles_c;
  const int s = wid / tiles;
  const int t = wid - s * tiles;
  int tr, tc;
  if (a.r_fastest) {
    tr = t % a.tiles_r;
    tc = t / a.tiles_r;
  } else {
    tc = t % a.tiles_c;
    tr = t / a.tiles_c;
  }
  const int r0 = tr * kBR, c0 = tc * kBC;
  const long q0 = (long)s * a.qslice;
  const long q1 = min(a.Q, q0 + a.qslice);
  const T* __restrict__ X = static_cast<const T*>(a.X);
  const T* __restrict__ Y = static_cast<const T*>(a.Y);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wr = wave >> 1, wc = wave & 1;
  const int l16 = lane & 15, lg = lane >> 4;
  const int rw0 = r0 + wr * 64, cw0 = c0 + wc * 64;
  const int rows = a.R - rw0, cols = a.C - cw0;
  unsigned xo[4], yo[4];
  bool xin[4], yin[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    xo[b] = (unsigned)(((long)(b * 16 + l16) * a.ldx + QG * lg) * SZ);
    xin[b] = b * 16 + l16 < rows;
    yo[b] = YC ? (unsigned)(((long)(b * 16 + l16) + (long)QG * lg * a.ldy) * SZ)
               : (unsigned)(((long)(b * 16 + l16) * a.ldy + QG * lg) * SZ);
    yin[b] = b * 16 + l16 < cols;
  }
  const unsigned xext = rows > 0 ? (unsigned)(((long)(min(64, rows) - 1) * a.ldx + BQ) * SZ) : 0u;
  const unsigned yext =
      cols > 0 ? (YC ? (unsigned)(((long)(BQ - 1) * a.ldy + min(64, cols)) * SZ)
                     : (unsigned)(((long)(min(64, cols) - 1) * a.ldy + BQ) * SZ))
               : 0u;

  auto load = [&](Stage& st, long qb) {
    const long left = q1 - qb;
    const __amdgpu_buffer_rsrc_t rx = rsrc(X + qb + (long)rw0 * a.ldx, xext);
    const __amdgpu_buffer_rsrc_t ry =
        rsrc(YC ? Y + cw0 + qb * a.ldy : Y + qb + (long)cw0 * a.ldy, yext);
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        const bool ok = xin[b] && QG * lg + 2 * p + 2 <= left;
        const v2 v = bload2<T>(rx, ok ? xo[b] + 2 * p * SZ : kOutside);
        st.x[b][2 * p] = v.x;
        st.x[b][2 * p + 1] = v.y;
      }
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        if (YC) {
          const int k0 = QG * lg + 2 * p;
          const bool ok0 = yin[b] && k0 < left, ok1 = yin[b] && k0 + 1 < left;
          st.y[b][2 * p] = bload1<T>(ry, ok0 ? yo[b] + (unsigned)(2 * p * a.ldy * SZ) : kOutside);
          st.y[b][2 * p + 1] =
              bload1<T>(ry, ok1 ? yo[b] + (unsigned)((2 * p + 1) * a.ldy * SZ) : kOutside);
        } else {
          const bool ok = yin[b] && QG * lg + 2 * p + 2 <= left;
          const v2 v = bload2<T>(ry, ok ? yo[b] + 2 * p * SZ : kOutside);
          st.y[b][2 * p] = v.x;
          st.y[b][2 * p + 1] = v.y;
        }
      }
  };
  acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
  auto mma = [&](const Stage& st) {
#pragma unroll
    for (int j = 0; j < QG; ++j)
#pragma unroll
      for (int bi = 0; bi < 4; ++bi)
#pragma unroll
        for (int bj = 0; bj < 4; ++bj) {
          acc[bi][bj] = M::op(st.x[bi][j], st.y[bj][j], acc[bi][bj]);
        }
  };
  // loads are unconditional: past q1 every lane is outside and reads zeros
  Stage s0, s1, s2;
  load(s0, q0);
  load(s1, q0 + BQ);
  for (long qb = q0; qb < q1; qb += 3 * BQ) {
    load(s2, qb + 2 * BQ);
    mma(s0);
    load(s0, qb + 3 * BQ);
    mma(s1);
    load(s1, qb + 4 * BQ);
    mma(s2);
  }
  T* O = static_cast<T*>(a.O) + (long)s * a.ostride;
#pragma unroll
  for (int bi = 0; bi < 4; ++bi) {
    const int rb = rw0 + bi * 16;
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) {
      const int c = cw0 + bj * 16 + l16;
      if (c >= a.C) continue;
      const acc_t v = acc[bi][bj];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = rb + M::row(lg, j);
        if (r < a.R) {
          T* o = OUTC ? O + c + (long)r * a.ldo : O + r + (long)c * a.ldo;
          *o = a.accumulate ? *o + v[j] : v[j];
        }
      }
    }
  }
}

template <typename T, bool YC, bool OUTC>
void launch(const Args& a, hipStream_t st) {
  const unsigned grid = (unsigned)((a.nblocks + 7) / 8 * 8);
  // pairs are aligned vectors, and a pair never straddles the end of the range, when the
  // leading dimension and the range along the pair (Q for q-contiguous reads, C for
  // c-contiguous ones) are even (slice ends are multiples of 16)
  const bool xv = (a.ldx % 2) == 0 && (a.Q % 2) == 0;
  const bool yv = (a.ldy % 2) == 0 && ((YC ? (long)a.C : a.Q) % 2) == 0;
  // buffer addressing: a stage of each operand spans less than kOutside bytes
  constexpr long BQ = Mf<T>::BQ;
  const long xspan = ((kBR - 1) * a.ldx + BQ) * (long)sizeof(T);
  const long yspan = (YC ? (BQ - 1) * a.ldy + kBC : (kBC - 1) * a.ldy + BQ) * (long)sizeof(T);
  const bool buf = QBH_GEMM_BUF && xv && yv && xspan < (long)kOutside && yspan < (long)kOutside;
  if constexpr (YC && sizeof(T) == 8) {
    if (buf && QBH_GEMM_DIRECT) {
      k_gemm_direct<T, YC, OUTC><<<grid, kThreads, 0, st>>>(a);
      return;
    }
  }
  if (buf)
    k_gemm<T, YC, OUTC, true, true, true><<<grid, kThreads, 0, st>>>(a);
  else if (xv && yv)
    k_gemm<T, YC, OUTC, true, true, false><<<grid, kThreads, 0, st>>>(a);
  else if (xv)
    k_gemm<T, YC, OUTC, true, false, false><<<grid, kThreads, 0, st>>>(a);
  else if (yv)
    k_gemm<T, YC, OUTC, false, true, false><<<grid, kThreads, 0, st>>>(a);
  else
    k_gemm<T, YC, OUTC, false, false, false><<<grid, kThreads, 0, st>>>(a);
}

}  // namespace

int mfma_slices(int R, int C, long Q, int want) {
  const long tiles = (long)((R + kBR - 1) / kBR) * ((C + kBC - 1) / kBC);
  long s = want > 0 ? want : (2048 + tiles - 1) / tiles;
  s = std::min<long>(s, 64);
  s = std::min<long>(s, std::max<long>(1, Q / 512));  // at least 32 stages per slice
  s = std::max<long>(1, s);
  const long qs = ((Q + s - 1) / s + kBQ - 1) / kBQ * kBQ;
  return (int)((Q + qs - 1) / qs);
}

long mfma_slice_len(long Q, int slices) {
  return ((Q + slices - 1) / slices + kBQ - 1) / kBQ * kBQ;
}

int mfma_gemm_qq(hipStream_t st, bool fp32, int R, int C, long Q, const void* X, long ldx,
                 const void* Y, long ldy, void* O, long ldo, int slices, long ostride) {
  if (R <= 0 || C <= 0 || Q <= 0 || slices <= 0) return -1;
  Args a{};
  a.R = R;
  a.C = C;
  a.Q = Q;
  a.X = X;
  a.ldx = ldx;
  a.Y = Y;
  a.ldy = ldy;
  a.O = O;
  a.ldo = ldo;
  a.qslice = mfma_slice_len(Q, slices);
  if ((Q + a.qslice - 1) / a.qslice != slices) return -1;
  a.ostride = ostride;
  a.tiles_r = (R + kBR - 1) / kBR;
  a.tiles_c = (C + kBC - 1) / kBC;
  a.nblocks = a.tiles_r * a.tiles_c * slices;
  a.r_fastest = 0;
  a.accumulate = 0;
  if (fp32)
    launch<float, false, false>(a, st);
  else
    launch<double, false, false>(a, st);
  return 0;
}

int mfma_gemm_qc(hipStream_t st, bool fp32, int R, int C, long Q, const void* X, long ldx,
                 const void* Y, long ldy, void* O, long ldo, bool accumulate) {
  if (R <= 0 || C <= 0 || Q <= 0) return -1;
  Args a{};
  a.R = R;
  a.C = C;
  a.Q = Q;
  a.X = X;
  a.ldx = ldx;
  a.Y = Y;
  a.ldy = ldy;
  a.O = O;
  a.ldo = ldo;
  a.qslice = (Q + kBQ - 1) / kBQ * kBQ;
  a.ostride = 0;
  a.tiles_r = (R + kBR - 1) / kBR;
  a.tiles_c = (C + kBC - 1) / kBC;
  a.nblocks = a.tiles_r * a.tiles_c;
  a.r_fastest = 1;
  a.accumulate = accumulate ? 1 : 0;
  if (fp32)
    launch<float, true, true>(a, st);
  else
    launch<double, true, true>(a, st);
  return 0;
}

}  // namespace qbh
