// Device-side 1-D FFT building blocks for gfx950: complex types, small-radix DFT
// butterflies and an LDS-resident Stockham pass.
//
// A line of length n (n = product of radices from {2,3,4,5,7,8,11}) lives in LDS.
// All threads of a workgroup cooperate on all lines of the workgroup: pass p with
// radix R has n/R butterflies per line; butterflies are dealt round-robin over the
// workgroup's threads, read into registers, synchronised, transformed and written
// back in self-sorting (Stockham) order, so the output is in natural order after the
// last pass (Stockham autosort formulation as in Govindaraju et al., SC'08).
//
// Direction DIR = +1 is the unnormalised exp(+i...) transform, DIR = -1 the
// exp(-i...) transform; scaling is applied by the callers.
#pragma once
#include <hip/hip_runtime.h>

namespace qbh {

template <typename T> struct cplx { T x, y; };

template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * (s*i) for s = +-1
template <int S, typename T>
__device__ __forceinline__ cplx<T> mul_si(cplx<T> a) {
  return S > 0 ? cplx<T>{-a.y, a.x} : cplx<T>{a.y, -a.x};
}

// cos/sin(2 pi m / R), m = 1..(R-1)/2, correctly rounded (computed in extended precision)
template <int R> struct OddConst;
template <> struct OddConst<3> {
  static constexpr double c[1] = {-0.5};
  static constexpr double s[1] = {0.8660254037844386};
};
template <> struct OddConst<5> {
  static constexpr double c[2] = {0.30901699437494745, -0.8090169943749475};
  static constexpr double s[2] = {0.9510565162951535, 0.5877852522924731};
};
template <> struct OddConst<7> {
  static constexpr double c[3] = {0.6234898018587335, -0.2225209339563144, -0.9009688679024191};
  static constexpr double s[3] = {0.7818314824680298, 0.9749279121818236, 0.4338837391175581};
};
template <> struct OddConst<11> {
  static constexpr double c[5] = {0.8412535328311812, 0.41541501300188644, -0.14231483827328514,
                                  -0.6548607339452851, -0.9594929736144974};
  static constexpr double s[5] = {0.5406408174555976, 0.9096319953545183, 0.9898214418809327,
                                  0.7557495743542583, 0.28173255684142967};
};

// Odd-radix DFT by conjugate-pair symmetry:
//   y_k = a_0 + sum_{j=1}^{h} (a_j + a_{R-j}) cos(2 pi jk/R) + DIR*i (a_j - a_{R-j}) sin(2 pi jk/R)
template <int R, int DIR, typename T>
__device__ __forceinline__ void dft_odd(cplx<T>* a) {
  constexpr int H = (R - 1) / 2;
  cplx<T> sp[H], sm[H];
  cplx<T> y0 = a[0];
#pragma unroll
  for (int j = 1; j <= H; ++j) {
    sp[j - 1] = cadd(a[j], a[R - j]);
    sm[j - 1] = csub(a[j], a[R - j]);
    y0 = cadd(y0, sp[j - 1]);
  }
  cplx<T> out[R];
  out[0] = y0;
#pragma unroll
  for (int k = 1; k <= H; ++k) {
    cplx<T> re = a[0];
    cplx<T> im = {T(0), T(0)};
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      const int m = (j * k) % R;
      const int mm = m <= H ? m : R - m;
      const T cs = T(OddConst<R>::c[mm - 1]);
      const T sn = m <= H ? T(OddConst<R>::s[mm - 1]) : -T(OddConst<R>::s[mm - 1]);
      re.x += sp[j - 1].x * cs;
      re.y += sp[j - 1].y * cs;
      im.x += sm[j - 1].x * sn;
      im.y += sm[j - 1].y * sn;
    }
    // DIR*i*im
    const cplx<T> t = mul_si<DIR>(im);
    out[k] = cadd(re, t);
    out[R - k] = csub(re, t);
  }
#pragma unroll
  for (int k = 0; k < R; ++k) a[k] = out[k];
}

template <int DIR, typename T>
__device__ __forceinline__ void dft2(cplx<T>* a) {
  const cplx<T> t = a[0];
  a[0] = cadd(t, a[1]);
  a[1] = csub(t, a[1]);
}

template <int DIR, typename T>
__device__ __forceinline__ void dft4(cplx<T>* a) {
  const cplx<T> s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
  const cplx<T> s13 = cadd(a[1], a[3]), d13 = mul_si<DIR>(csub(a[1], a[3]));
  a[0] = cadd(s02, s13);
  a[2] = csub(s02, s13);
  a[1] = cadd(d02, d13);
  a[3] = csub(d02, d13);
}

template <int DIR, typename T>
__device__ __forceinline__ void dft8(cplx<T>* a) {
  // decimation in time: E = DFT4(even), O = DFT4(odd); y_k = E_k + w^k O_k
  cplx<T> e[4] = {a[0], a[2], a[4], a[6]};
  cplx<T> o[4] = {a[1], a[3], a[5], a[7]};
  dft4<DIR>(e);
  dft4<DIR>(o);
  const T r = T(0.7071067811865476);
  // w^1 = (r, DIR r), w^2 = DIR i, w^3 = (-r, DIR r)
  const cplx<T> o1 = cmul(o[1], cplx<T>{r, DIR * r});
  const cplx<T> o2 = mul_si<DIR>(o[2]);
  const cplx<T> o3 = cmul(o[3], cplx<T>{-r, DIR * r});
  a[0] = cadd(e[0], o[0]);
  a[4] = csub(e[0], o[0]);
  a[1] = cadd(e[1], o1);
  a[5] = csub(e[1], o1);
  a[2] = cadd(e[2], o2);
  a[6] = csub(e[2], o2);
  a[3] = cadd(e[3], o3);
  a[7] = csub(e[3], o3);
}

template <int R, int DIR, typename T>
__device__ __forceinline__ void dft(cplx<T>* a) {
  if constexpr (R == 2) dft2<DIR>(a);
  else if constexpr (R == 4) dft4<DIR>(a);
  else if constexpr (R == 8) dft8<DIR>(a);
  else dft_odd<R, DIR>(a);
}

// Per-axis transform description (host fills it; passed by value in kernel args).
struct AxisPlan {
  int n;
  int npass;
  int radix[8];
};

// One Stockham pass with radix R over `nlines` lines of length n stored with leading
// dimension `ld` in LDS buffer `buf`.  `tw` holds exp(+2 pi i m / n), m = 0..n-1 (in LDS
// or global memory).  Ns = product of the radices of the previous passes.
// EPT bounds the elements per thread: nlines * n <= EPT * blockDim.x, so every thread
// holds at most ceil(EPT / R) butterflies of R points in registers between the read
// and the write half of the pass.
template <int R, int DIR, int EPT, typename T>
__device__ __forceinline__ void stockham_pass(cplx<T>* buf, int nlines, int ld, int n, int Ns,
                                              const cplx<T>* tw) {
  constexpr int MAXIT = (EPT + R - 1) / R;
  const int nb = n / R;
  const int total = nlines * nb;
  const int nthr = blockDim.x;
  const int step = n / (Ns * R);
  cplx<T> v[MAXIT][R];
  int dst[MAXIT];
#pragma unroll
  for (int it = 0; it < MAXIT; ++it) {
    const int w = threadIdx.x + it * nthr;
    dst[it] = -1;
    if (w < total) {
      const int line = w / nb;
      const int j = w - line * nb;
      const int k = j % Ns;
      cplx<T>* base = buf + line * ld;
#pragma unroll
      for (int r = 0; r < R; ++r) v[it][r] = base[j + r * nb];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cplx<T> t = tw[k * r * step];
        if (DIR < 0) t.y = -t.y;
        v[it][r] = cmul(v[it][r], t);
      }
      dft<R, DIR>(v[it]);
      dst[it] = line * ld + (j / Ns) * Ns * R + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < MAXIT; ++it) {
    if (dst[it] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[dst[it] + r * Ns] = v[it][r];
    }
  }
  __syncthreads();
}

// ------------------------------------------------------------------------------------
// Wavefront-local variants: the 64 lanes of one wavefront own `nlines` LDS lines and
// transform them with no workgroup barrier.  LDS operations of one wavefront complete
// in issue order, so a release/acquire fence at wavefront scope (which also keeps the
// compiler from moving LDS accesses across it) is the only synchronisation needed
// between the read and the write half of a pass.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// EPL bounds the elements per lane: nlines * n <= EPL * 64.
template <int R, int DIR, int EPL, typename T>
__device__ __forceinline__ void stockham_pass_w(cplx<T>* buf, int nlines, int ld, int n, int Ns,
                                                const cplx<T>* tw, int lane) {
  constexpr int MAXIT = (EPL + R - 1) / R;
  const int nb = n / R;
  const int total = nlines * nb;
  const int step = n / (Ns * R);
  cplx<T> v[MAXIT][R];
  int dst[MAXIT];
#pragma unroll
  for (int it = 0; it < MAXIT; ++it) {
    const int w = lane + it * 64;
    dst[it] = -1;
    if (w < total) {
      const int line = w / nb;
      const int j = w - line * nb;
      const int k = j % Ns;
      const cplx<T>* base = buf + line * ld;
#pragma unroll
      for (int r = 0; r < R; ++r) v[it][r] = base[j + r * nb];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cplx<T> t = tw[k * r * step];
        if (DIR < 0) t.y = -t.y;
        v[it][r] = cmul(v[it][r], t);
      }
      dft<R, DIR>(v[it]);
      dst[it] = line * ld + (j / Ns) * Ns * R + k;
    }
  }
  wave_sync();
#pragma unroll
  for (int it = 0; it < MAXIT; ++it) {
    if (dst[it] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[dst[it] + r * Ns] = v[it][r];
    }
  }
  wave_sync();
}

// Wavefront-local transform of `nlines` lines at `buf` (this wavefront's lines only).
// Caller: data written by this wavefront (or made visible by a barrier) before the call.
template <int DIR, int EPL, typename T>
__device__ __forceinline__ void fft_lines_w(cplx<T>* buf, int nlines, int ld, const AxisPlan& p,
                                            const cplx<T>* tw, int lane) {
  int Ns = 1;
  for (int ip = 0; ip < p.npass; ++ip) {
    const int R = p.radix[ip];
    switch (R) {
      case 2: stockham_pass_w<2, DIR, EPL>(buf, nlines, ld, p.n, Ns, tw, lane); break;
      case 3: stockham_pass_w<3, DIR, EPL>(buf, nlines, ld, p.n, Ns, tw, lane); break;
      case 4: stockham_pass_w<4, DIR, EPL>(buf, nlines, ld, p.n, Ns, tw, lane); break;
      case 5: stockham_pass_w<5, DIR, EPL>(buf, nlines, ld, p.n, Ns, tw, lane); break;
      case 7: stockham_pass_w<7, DIR, EPL>(buf, nlines, ld, p.n, Ns, tw, lane); break;
      case 8: stockham_pass_w<8, DIR, EPL>(buf, nlines, ld, p.n, Ns, tw, lane); break;
      case 11: stockham_pass_w<11, DIR, EPL>(buf, nlines, ld, p.n, Ns, tw, lane); break;
      default: break;
    }
    Ns *= R;
  }
}

// Full transform of `nlines` LDS lines.  Caller must __syncthreads() before (data in
// place); returns after a barrier (result visible to all threads).
template <int DIR, int EPT, typename T>
__device__ __forceinline__ void fft_lines(cplx<T>* buf, int nlines, int ld, const AxisPlan& p,
                                          const cplx<T>* tw) {
  int Ns = 1;
  for (int ip = 0; ip < p.npass; ++ip) {
    const int R = p.radix[ip];
    switch (R) {
      case 2: stockham_pass<2, DIR, EPT>(buf, nlines, ld, p.n, Ns, tw); break;
      case 3: stockham_pass<3, DIR, EPT>(buf, nlines, ld, p.n, Ns, tw); break;
      case 4: stockham_pass<4, DIR, EPT>(buf, nlines, ld, p.n, Ns, tw); break;
      case 5: stockham_pass<5, DIR, EPT>(buf, nlines, ld, p.n, Ns, tw); break;
      case 7: stockham_pass<7, DIR, EPT>(buf, nlines, ld, p.n, Ns, tw); break;
      case 8: stockham_pass<8, DIR, EPT>(buf, nlines, ld, p.n, Ns, tw); break;
      case 11: stockham_pass<11, DIR, EPT>(buf, nlines, ld, p.n, Ns, tw); break;
      default: break;
    }
    Ns *= R;
  }
}

}  // namespace qbh
