// Register-resident small DFTs (lengths 2..16) and a two-factor row transform.
//
// A row of length n = N1*N2 is transformed in two register stages with one LDS
// exchange between them (Cooley-Tukey "four-step" with the transposition folded into
// the LDS addressing):
//   stage A : for each j2 < N2, DFT_N1 over x[j2 + N2*j1], twiddle W_n^{j2*k1}
//   stage B : for each k1 < N1, DFT_N2 over the N2 contiguous values -> X[k1 + N1*k2]
// Small DFTs of composite length use the prime-factor (Good-Thomas) index maps when the
// factors are coprime (6, 10, 12, 14, 15) and Cooley-Tukey with constant twiddles
// otherwise (9 = 3*3, 16 = 4*4).
#pragma once
#include "fft_device.h"

namespace qbh {

// cos/sin(2 pi m / L) tables for the Cooley-Tukey small lengths
template <int L> struct WTab;
template <> struct WTab<9> {
  static constexpr double c[9] = {1.0, 0.766044443118978, 0.17364817766693036, -0.5,
                                  -0.9396926207859084, -0.9396926207859084, -0.5,
                                  0.17364817766693036, 0.766044443118978};
  static constexpr double s[9] = {0.0, 0.6427876096865394, 0.984807753012208,
                                  0.8660254037844386, 0.3420201433256687, -0.3420201433256687,
                                  -0.8660254037844386, -0.984807753012208, -0.6427876096865394};
};
template <> struct WTab<16> {
  static constexpr double c[16] = {1.0, 0.9238795325112867, 0.7071067811865476, 0.3826834323650898,
                                   0.0, -0.3826834323650898, -0.7071067811865476,
                                   -0.9238795325112867, -1.0, -0.9238795325112867,
                                   -0.7071067811865476, -0.3826834323650898, 0.0,
                                   0.3826834323650898, 0.7071067811865476, 0.9238795325112867};
  static constexpr double s[16] = {0.0, 0.3826834323650898, 0.7071067811865476, 0.9238795325112867,
                                   1.0, 0.9238795325112867, 0.7071067811865476, 0.3826834323650898,
                                   0.0, -0.3826834323650898, -0.7071067811865476,
                                   -0.9238795325112867, -1.0, -0.9238795325112867,
                                   -0.7071067811865476, -0.3826834323650898};
};

constexpr int inv_mod(int a, int m) {
  for (int x = 1; x < m; ++x)
    if ((a * x) % m == 1) return x;
  return 1;
}

template <int L, int DIR, typename T> __device__ __forceinline__ void sdft(cplx<T>* a);

// Good-Thomas: L = A*B, gcd(A,B) = 1.  Input n = (a*B + b*A) mod L, output
// k = (ka*B*inv(B,A) + kb*A*inv(A,B)) mod L.
template <int A, int B, int DIR, typename T>
__device__ __forceinline__ void pfa(cplx<T>* x) {
  constexpr int L = A * B;
  cplx<T> m[B][A];
#pragma unroll
  for (int b = 0; b < B; ++b) {
#pragma unroll
    for (int a = 0; a < A; ++a) m[b][a] = x[(a * B + b * A) % L];
    sdft<A, DIR>(m[b]);
  }
  constexpr int cb = B * inv_mod(B % A, A);
  constexpr int ca = A * inv_mod(A % B, B);
#pragma unroll
  for (int ka = 0; ka < A; ++ka) {
    cplx<T> col[B];
#pragma unroll
    for (int b = 0; b < B; ++b) col[b] = m[b][ka];
    sdft<B, DIR>(col);
#pragma unroll
    for (int kb = 0; kb < B; ++kb) x[(ka * cb + kb * ca) % L] = col[kb];
  }
}

// Cooley-Tukey L = A*B with constant twiddles: X[k1 + A*k2] =
//   sum_j2 W_B^{j2 k2} W_L^{j2 k1} sum_j1 x[j2 + B*j1] W_A^{j1 k1}
template <int A, int B, int DIR, typename T>
__device__ __forceinline__ void ct(cplx<T>* x) {
  constexpr int L = A * B;
  cplx<T> m[B][A];
#pragma unroll
  for (int j2 = 0; j2 < B; ++j2) {
#pragma unroll
    for (int j1 = 0; j1 < A; ++j1) m[j2][j1] = x[j2 + B * j1];
    sdft<A, DIR>(m[j2]);
#pragma unroll
    for (int k1 = 1; k1 < A; ++k1) {
      const int e = (j2 * k1) % L;
      if (e != 0) {
        const cplx<T> w = {T(WTab<L>::c[e]), T(DIR * WTab<L>::s[e])};
        m[j2][k1] = cmul(m[j2][k1], w);
      }
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < A; ++k1) {
    cplx<T> col[B];
#pragma unroll
    for (int j2 = 0; j2 < B; ++j2) col[j2] = m[j2][k1];
    sdft<B, DIR>(col);
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) x[k1 + A * k2] = col[k2];
  }
}

template <int L, int DIR, typename T>
__device__ __forceinline__ void sdft(cplx<T>* a) {
  if constexpr (L == 1) {
  } else if constexpr (L == 2 || L == 3 || L == 4 || L == 5 || L == 7 || L == 8 || L == 11) {
    dft<L, DIR>(a);
  } else if constexpr (L == 6) {
    pfa<2, 3, DIR>(a);
  } else if constexpr (L == 10) {
    pfa<2, 5, DIR>(a);
  } else if constexpr (L == 12) {
    pfa<4, 3, DIR>(a);
  } else if constexpr (L == 14) {
    pfa<2, 7, DIR>(a);
  } else if constexpr (L == 15) {
    pfa<3, 5, DIR>(a);
  } else if constexpr (L == 9) {
    ct<3, 3, DIR>(a);
  } else if constexpr (L == 16) {
    ct<4, 4, DIR>(a);
  } else {
    static_assert(L == 2, "unsupported small DFT length");
  }
}

// powers w^k, k = 0..K-1, of w = tw[e] (conjugated for DIR < 0) by repeated products
template <int K, int DIR, typename T>
__device__ __forceinline__ void twiddle_row(cplx<T>* a, const cplx<T>* tw, int e) {
  if (e == 0) return;
  cplx<T> w = tw[e];
  if (DIR < 0) w.y = -w.y;
  cplx<T> p = w;
#pragma unroll
  for (int k = 1; k < K; ++k) {
    a[k] = cmul(a[k], p);
    if (k + 1 < K) p = cmul(p, w);
  }
}

// Padded LDS row layout for the two-factor transform: element x of a row lives at
// x + x / N2 when N2 is even, so the stage-B lanes (stride N2 + 1 elements, odd) hit
// distinct LDS banks; the row occupies N1 * (N2 + 1) elements.  N1 = 0 (no register
// factorisation) means an unpadded row.
// Off by default (QBH_LDS_PAD=0): measured on MI355X the padded rows raise the LDS
// footprint of the y/z passes from 38 to 42 KB per workgroup (4 -> 3 workgroups per CU)
// and cost more than the 2-way stage-B conflicts they remove.
#ifndef QBH_LDS_PAD
#define QBH_LDS_PAD 0
#endif
template <int N1, int N2, int ON = QBH_LDS_PAD> struct Pad {
  static constexpr int P = (ON && N1 > 0 && N2 % 2 == 0) ? 1 : 0;
  static __host__ __device__ constexpr int idx(int x) { return P ? x + x / N2 : x; }
  static __host__ __device__ constexpr int row(int n) { return P ? n + n / N2 : n; }
};

// Wavefront-local transform of `nrows` (<= RPW) LDS rows of length n = N1*N2 at
// buf + r*ld (padded layout, Pad<N1,N2>), natural order in and out, direction DIR:
//   stage A: task (r, j2): DFT_N1 over x[j2 + N2*j1], twiddle W_n^{DIR j2 k1}, in place
//   stage B: task (r, k1): DFT_N2 over the contiguous block N2*k1 + j2 -> X[k1 + N1*k2],
//            held in registers until every lane has read its block, then stored at the
//            natural positions.
template <int N1, int N2, int RPW, int DIR, typename T>
__device__ __forceinline__ void rows_fft(cplx<T>* buf, int nrows, int ld, const cplx<T>* tw,
                                         int lane) {
  constexpr int TA = RPW * N2, TB = RPW * N1;
  constexpr int IA = (TA + 63) / 64, IB = (TB + 63) / 64;
  constexpr int S = N2 + Pad<N1, N2>::P;  // padded stride of the N2-blocks
#pragma unroll
  for (int ia = 0; ia < IA; ++ia) {
    const int t = lane + 64 * ia;
    const int r = t / N2, j2 = t % N2;
    if (t < TA && r < nrows) {
      cplx<T>* row = buf + r * ld + j2;
      cplx<T> a[N1];
#pragma unroll
      for (int j1 = 0; j1 < N1; ++j1) a[j1] = row[S * j1];
      sdft<N1, DIR>(a);
      twiddle_row<N1, DIR>(a, tw, j2);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) row[S * k1] = a[k1];
    }
  }
  wave_sync();
  cplx<T> b[IB][N2];
#pragma unroll
  for (int ib = 0; ib < IB; ++ib) {
    const int t = lane + 64 * ib;
    const int r = t / N1, k1 = t % N1;
    if (t < TB && r < nrows) {
      const cplx<T>* blk = buf + r * ld + S * k1;
#pragma unroll
      for (int j2 = 0; j2 < N2; ++j2) b[ib][j2] = blk[j2];
      sdft<N2, DIR>(b[ib]);
    }
  }
  wave_sync();
#pragma unroll
  for (int ib = 0; ib < IB; ++ib) {
    const int t = lane + 64 * ib;
    const int r = t / N1, k1 = t % N1;
    if (t < TB && r < nrows) {
      cplx<T>* row = buf + r * ld;
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) row[Pad<N1, N2>::idx(k1 + N1 * k2)] = b[ib][k2];
    }
  }
  wave_sync();
}

}  // namespace qbh
