// Batched mixed-radix complex FFT of lines held in LDS (CDNA4 / gfx950).
//
// Self-sorting Stockham formulation: stage s with radix R combines R
// sub-transforms of size Ns into size Ns*R.  Butterfly j (0 <= j < N/R)
// reads x[j + r*N/R], applies twiddles W_{Ns R}^{(j mod Ns) r}, does a
// radix-R DFT and writes y[(j/Ns)*Ns*R + (j mod Ns) + r*Ns].  Output is in
// natural order after the last stage.  Each stage runs in place: every
// thread first pulls all of its butterfly inputs into registers, the block
// synchronises, then results are written back.  A thread therefore holds at
// most MAXC complex values, so a call may cover nlines*N <= MAXC*blockDim.
//
// Twiddles are a table tw[t] = exp(-2 pi i t / N) (forward sign, the sign of
// the reference's cufftExecR2C), computed in double on the host and staged
// in LDS by the caller.  Inverse transforms conjugate input and output.
#pragma once
#include <hip/hip_runtime.h>

namespace saxs {

constexpr int kMaxStages = 12;

struct FftDesc {
  int N;
  int nst;
  int radix[kMaxStages];
};

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// a * (-i)
__device__ __forceinline__ float2 cmul_mi(float2 a) { return make_float2(a.y, -a.x); }

template <int R>
struct Radix;

template <>
struct Radix<2> {
  static __device__ __forceinline__ void dft(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <>
struct Radix<4> {
  static __device__ __forceinline__ void dft(float2* v) {
    float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    float2 t2 = cadd(v[1], v[3]), t3 = cmul_mi(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};

template <>
struct Radix<8> {
  static __device__ __forceinline__ void dft(float2* v) {
    constexpr float c = 0.70710678118654752f;
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    Radix<4>::dft(e);
    Radix<4>::dft(o);
    // o_k *= exp(-i pi k / 4)
    float2 o1 = make_float2(c * (o[1].x + o[1].y), c * (o[1].y - o[1].x));
    float2 o2 = cmul_mi(o[2]);
    float2 o3 = make_float2(c * (o[3].y - o[3].x), -c * (o[3].x + o[3].y));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
};

// Twiddles of the odd radices: C[t] = cos(2 pi t/R), S[t] = -sin(2 pi t/R).
template <int R>
struct OddTables;
template <>
struct OddTables<3> {
  static constexpr float C[3] = {1.0f, -0.5f, -0.5f};
  static constexpr float S[3] = {0.0f, -0.86602540378443865f, 0.86602540378443865f};
};
template <>
struct OddTables<5> {
  static constexpr float C[5] = {1.0f, 0.30901699437494742f, -0.80901699437494742f,
                                 -0.80901699437494742f, 0.30901699437494742f};
  static constexpr float S[5] = {0.0f, -0.95105651629515357f, -0.58778525229247313f,
                                 0.58778525229247313f, 0.95105651629515357f};
};
template <>
struct OddTables<7> {
  static constexpr float C[7] = {1.0f, 0.62348980185873353f, -0.22252093395631440f,
                                 -0.90096886790241913f, -0.90096886790241913f,
                                 -0.22252093395631440f, 0.62348980185873353f};
  static constexpr float S[7] = {0.0f, -0.78183148246802981f, -0.97492791218182361f,
                                 -0.43388373911755812f, 0.43388373911755812f,
                                 0.97492791218182361f, 0.78183148246802981f};
};

// Odd radix R by conjugate-symmetric pairs: with a_j = v_j + v_{R-j} and
// b_j = v_j - v_{R-j} (j = 1..(R-1)/2) and t = 2 pi j k / R,
//   X_k     = v_0 + sum_j a_j cos t - i sum_j b_j sin t
//   X_{R-k} = v_0 + sum_j a_j cos t + i sum_j b_j sin t,
// (R-1)^2 + 5(R-1) real operations against 4R(R-1) for the direct sum.
template <int R>
struct RadixOdd {
  static __device__ __forceinline__ void dft(float2* v) {
    constexpr int M = (R - 1) / 2;
    float2 a[M], b[M];
#pragma unroll
    for (int j = 1; j <= M; ++j) {
      a[j - 1] = cadd(v[j], v[R - j]);
      b[j - 1] = csub(v[j], v[R - j]);
    }
    float2 out[R];
    float2 s0 = v[0];
#pragma unroll
    for (int j = 0; j < M; ++j) s0 = cadd(s0, a[j]);
    out[0] = s0;
#pragma unroll
    for (int k = 1; k <= M; ++k) {
      float2 A = v[0], B = make_float2(0.0f, 0.0f);
#pragma unroll
      for (int j = 1; j <= M; ++j) {
        const int t = (j * k) % R;
        const float c = OddTables<R>::C[t], sn = -OddTables<R>::S[t];   // cos t, sin t
        A.x += a[j - 1].x * c;
        A.y += a[j - 1].y * c;
        B.x += b[j - 1].x * sn;
        B.y += b[j - 1].y * sn;
      }
      out[k] = make_float2(A.x + B.y, A.y - B.x);
      out[R - k] = make_float2(A.x - B.y, A.y + B.x);
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = out[k];
  }
};
template <>
struct Radix<3> : RadixOdd<3> {};
template <>
struct Radix<5> : RadixOdd<5> {};
template <>
struct Radix<7> : RadixOdd<7> {};

// Composite radix A*B in registers (Cooley-Tukey): n = B n1 + n2, k = k1 + A k2,
// X[k] = sum_n2 W_B^{n2 k2} W_{AB}^{n2 k1} sum_n1 x[B n1 + n2] W_A^{n1 k1};
// natural order in and out.  Wider Stockham stages mean fewer LDS passes.
template <int A, int B>
struct RadixCT {
  static __device__ __forceinline__ void dft(float2* v) {
    constexpr int N = A * B;
    float2 y[B][A];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) {
#pragma unroll
      for (int n1 = 0; n1 < A; ++n1) y[n2][n1] = v[B * n1 + n2];
      Radix<A>::dft(y[n2]);
#pragma unroll
      for (int k1 = 1; k1 < A; ++k1) {
        if (n2 == 0) continue;
        const double ang = -2.0 * 3.14159265358979323846 * (double)(n2 * k1) / (double)N;
        y[n2][k1] = cmul(y[n2][k1], make_float2((float)__builtin_cos(ang), (float)__builtin_sin(ang)));
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
      float2 z[B];
#pragma unroll
      for (int n2 = 0; n2 < B; ++n2) z[n2] = y[n2][k1];
      Radix<B>::dft(z);
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = z[k2];
    }
  }
};
template <>
struct Radix<6> : RadixCT<2, 3> {};
template <>
struct Radix<9> : RadixCT<3, 3> {};
template <>
struct Radix<10> : RadixCT<2, 5> {};
template <>
struct Radix<12> : RadixCT<4, 3> {};

// One Stockham stage over nlines lines (line stride ls) of length N.
template <int R, int MAXC>
__device__ __forceinline__ void fft_stage(float2* buf, int ls, int nlines, int N, int Ns,
                                          const float2* __restrict__ tw) {
  constexpr int IT = MAXC / R;
  const unsigned nb = (unsigned)(N / R);
  const unsigned total = (unsigned)nlines * nb;
  const unsigned nsu = (unsigned)Ns;
  float2 v[IT][R];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const unsigned item = threadIdx.x + (unsigned)it * blockDim.x;
    if (item < total) {
      const unsigned line = item / nb, j = item - line * nb;
      const float2* src = buf + line * ls;
#pragma unroll
      for (int r = 0; r < R; ++r) v[it][r] = src[j + r * nb];
    }
  }
  __syncthreads();
  const unsigned tstride = (unsigned)N / (nsu * R);
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const unsigned item = threadIdx.x + (unsigned)it * blockDim.x;
    if (item < total) {
      const unsigned line = item / nb, j = item - line * nb;
      const unsigned q = j / nsu, k = j - q * nsu;
      if (k != 0) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[it][r] = cmul(v[it][r], tw[k * r * tstride]);
      }
      Radix<R>::dft(v[it]);
      float2* dst = buf + line * ls + q * nsu * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[r * nsu] = v[it][r];
    }
  }
  __syncthreads();
}

// First stage (Ns = 1, no twiddles) fed by a loader instead of LDS:
// ld(line, t) returns input element t of a line.  Results go to buf in
// place of the stage's output; the routine ends synchronised.  Reading from
// a different buffer than it writes, it needs no intermediate barrier.
template <int R, class Load>
__device__ __forceinline__ void fft_stage0(float2* buf, int ls, int nlines, int N, Load ld) {
  const unsigned nb = (unsigned)(N / R);
  const unsigned total = (unsigned)nlines * nb;
  for (unsigned item = threadIdx.x; item < total; item += blockDim.x) {
    const unsigned line = item / nb, j = item - line * nb;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = ld(line, j + r * nb);
    Radix<R>::dft(v);
    float2* dst = buf + line * ls + j * R;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[r] = v[r];
  }
  __syncthreads();
}

// Forward FFT of nlines lines of length d.N in LDS, from stage s0 on (the
// sub-transform size before stage s0 is Ns0).  The radix set RS (named POW2
// in the kernels) bounds the code, and so the registers, of one instance:
// 1 radices 8/4/2, 0 also 3/5/7, 2 also the composite radices 6, 9, 10 and
// 12 (fft_radix_set).  The caller must have synchronised after
// filling buf; the routine ends synchronised.
template <int MAXC, int POW2 = 0>
__device__ void fft_lines(float2* buf, int ls, int nlines, const FftDesc& d,
                          const float2* __restrict__ tw, int s0 = 0, int Ns0 = 1) {
  int Ns = Ns0;
  for (int s = s0; s < d.nst; ++s) {
    const int R = d.radix[s];
    if (R == 8) {
      fft_stage<8, MAXC>(buf, ls, nlines, d.N, Ns, tw);
    } else if (R == 4) {
      fft_stage<4, MAXC>(buf, ls, nlines, d.N, Ns, tw);
    } else if (R == 2) {
      fft_stage<2, MAXC>(buf, ls, nlines, d.N, Ns, tw);
    } else if constexpr (POW2 != 1) {
      if (R == 3) fft_stage<3, MAXC>(buf, ls, nlines, d.N, Ns, tw);
      else if (R == 5) fft_stage<5, MAXC>(buf, ls, nlines, d.N, Ns, tw);
      else if (R == 7) fft_stage<7, MAXC>(buf, ls, nlines, d.N, Ns, tw);
      else if constexpr (POW2 == 2) {
        if (R == 6) fft_stage<6, MAXC>(buf, ls, nlines, d.N, Ns, tw);
        else if (R == 9) fft_stage<9, MAXC>(buf, ls, nlines, d.N, Ns, tw);
        else if (R == 10) fft_stage<10, MAXC>(buf, ls, nlines, d.N, Ns, tw);
        else if (R == 12) fft_stage<12, MAXC>(buf, ls, nlines, d.N, Ns, tw);
      }
    }
    Ns *= R;
  }
}

// Whole transform with the first stage fed by a loader.
template <int MAXC, int POW2, class Load>
__device__ void fft_lines_loaded(float2* buf, int ls, int nlines, const FftDesc& d,
                                 const float2* __restrict__ tw, Load ld) {
  const int R = d.radix[0];
  if (R == 8) fft_stage0<8>(buf, ls, nlines, d.N, ld);
  else if (R == 4) fft_stage0<4>(buf, ls, nlines, d.N, ld);
  else if (R == 2) fft_stage0<2>(buf, ls, nlines, d.N, ld);
  else if constexpr (POW2 != 1) {
    if (R == 3) fft_stage0<3>(buf, ls, nlines, d.N, ld);
    else if (R == 5) fft_stage0<5>(buf, ls, nlines, d.N, ld);
    else if (R == 7) fft_stage0<7>(buf, ls, nlines, d.N, ld);
    else if constexpr (POW2 == 2) {
      if (R == 6) fft_stage0<6>(buf, ls, nlines, d.N, ld);
      else if (R == 9) fft_stage0<9>(buf, ls, nlines, d.N, ld);
      else if (R == 10) fft_stage0<10>(buf, ls, nlines, d.N, ld);
      else if (R == 12) fft_stage0<12>(buf, ls, nlines, d.N, ld);
    }
  }
  fft_lines<MAXC, POW2>(buf, ls, nlines, d, tw, 1, R);
}

// W16^m = exp(-2 pi i m / 16) for m = 0..15.
__device__ __forceinline__ float2 w16(int m) {
  constexpr float C[16] = {1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
                           0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f,
                           -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508977f,
                           0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f};
  return make_float2(C[m & 15], -C[(m + 12) & 15]);
}

// Radix-16 DFT in registers as 4 x 4: n = n1 + 4 n2, k = k2 + 4 k1.
// HALF: inputs 8..15 are zero.
template <bool HALF>
__device__ __forceinline__ void dft16(float2* v) {
  float2 u[4][4];
#pragma unroll
  for (int n1 = 0; n1 < 4; ++n1) {
    float2 t[4];
    if constexpr (HALF) {                   // DFT4 of (a, b, 0, 0), no adds of zeros
      const float2 a = v[n1], b = v[n1 + 4], mib = cmul_mi(b);
      t[0] = cadd(a, b);
      t[1] = cadd(a, mib);
      t[2] = csub(a, b);
      t[3] = csub(a, mib);
    } else {
      t[0] = v[n1];
      t[1] = v[n1 + 4];
      t[2] = v[n1 + 8];
      t[3] = v[n1 + 12];
      Radix<4>::dft(t);
    }
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) u[n1][k2] = (n1 * k2 == 0) ? t[k2] : cmul(t[k2], w16(n1 * k2));
  }
#pragma unroll
  for (int k2 = 0; k2 < 4; ++k2) {
    float2 t[4] = {u[0][k2], u[1][k2], u[2][k2], u[3][k2]};
    Radix<4>::dft(t);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) v[k2 + 4 * k1] = t[k1];
  }
}

// i * a.
__device__ __forceinline__ float2 cmul_i(float2 a) { return make_float2(-a.y, a.x); }

// dft16 restricted to the outputs k = k2 + 4 k1 with k1 = 0 (v[0..3]) and
// k1 = 3 (v[12..15]); the other outputs are left undefined.  Every
// first-stage DFT4 and twiddle is needed, but each second-stage DFT4 yields
// two sums: k1 = 0, sum_n1 u, and k1 = 3, sum_n1 W4^{3 n1} u = u0 + i u1 -
// u2 - i u3.  Outputs the caller never reads are removed by the compiler.
__device__ __forceinline__ void dft16_ends(float2* v) {
  float2 u[4][4];
#pragma unroll
  for (int n1 = 0; n1 < 4; ++n1) {
    float2 t[4] = {v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12]};
    Radix<4>::dft(t);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) u[n1][k2] = (n1 * k2 == 0) ? t[k2] : cmul(t[k2], w16(n1 * k2));
  }
#pragma unroll
  for (int k2 = 0; k2 < 4; ++k2) {
    const float2 s02 = cadd(u[0][k2], u[2][k2]), d02 = csub(u[0][k2], u[2][k2]);
    const float2 s13 = cadd(u[1][k2], u[3][k2]), d13 = csub(u[1][k2], u[3][k2]);
    v[k2] = cadd(s02, s13);
    v[k2 + 12] = cadd(d02, cmul_i(d13));
  }
}

// dft16 of an input that is zero except at n = 0, 1, 14, 15 (all 16 outputs):
// with n = n1 + 4 m the first-stage DFT4 over m has one nonzero input, at
// m = 0 for n1 = 0, 1 and m = 3 for n1 = 2, 3, so its outputs are that input
// times W4^{m k2} = 1 or i^k2.
__device__ __forceinline__ void dft16_in_ends(float2* v) {
  float2 u[4][4];
#pragma unroll
  for (int k2 = 0; k2 < 4; ++k2) {
    u[0][k2] = v[0];
    u[1][k2] = k2 == 0 ? v[1] : cmul(v[1], w16(k2));
    float2 a = v[14], c = v[15];                     // times i^k2
#pragma unroll
    for (int q = 0; q < k2; ++q) {
      a = cmul_i(a);
      c = cmul_i(c);
    }
    u[2][k2] = k2 == 0 ? a : cmul(a, w16(2 * k2));
    u[3][k2] = k2 == 0 ? c : cmul(c, w16(3 * k2));
  }
#pragma unroll
  for (int k2 = 0; k2 < 4; ++k2) {
    float2 t[4] = {u[0][k2], u[1][k2], u[2][k2], u[3][k2]};
    Radix<4>::dft(t);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) v[k2 + 4 * k1] = t[k1];
  }
}

// dft16 restricted to the outputs k < 8 (k = k2 + 4 k1, k1 = 0, 1): the
// second-stage DFT4s give sum_n1 u and u0 - i u1 - u2 + i u3 (W4 = -i).
__device__ __forceinline__ void dft16_lo(float2* v) {
  float2 u[4][4];
#pragma unroll
  for (int n1 = 0; n1 < 4; ++n1) {
    float2 t[4] = {v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12]};
    Radix<4>::dft(t);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) u[n1][k2] = (n1 * k2 == 0) ? t[k2] : cmul(t[k2], w16(n1 * k2));
  }
#pragma unroll
  for (int k2 = 0; k2 < 4; ++k2) {
    const float2 s02 = cadd(u[0][k2], u[2][k2]), d02 = csub(u[0][k2], u[2][k2]);
    const float2 s13 = cadd(u[1][k2], u[3][k2]), d13 = csub(u[1][k2], u[3][k2]);
    v[k2] = cadd(s02, s13);
    v[k2 + 4] = cadd(d02, cmul_mi(d13));
  }
}

// Radix-32 DFT in registers as 16 x 2: n = n1 + 2 n2 (n1 < 2, n2 < 16),
// k = k2 + 16 k1: X[k] = sum_n1 W32^{n1 k2} W2^{n1 k1} DFT16_{n2}(v[n1 + 2 n2])[k2].
__device__ __forceinline__ void dft32(float2* v) {
  float2 a[16], c[16];
#pragma unroll
  for (int n2 = 0; n2 < 16; ++n2) {
    a[n2] = v[2 * n2];
    c[n2] = v[2 * n2 + 1];
  }
  dft16<false>(a);
  dft16<false>(c);
#pragma unroll
  for (int k2 = 0; k2 < 16; ++k2) {
    const double ang = -2.0 * 3.14159265358979323846 * k2 / 32.0;
    const float2 w = make_float2((float)__builtin_cos(ang), (float)__builtin_sin(ang));
    const float2 t = (k2 == 0) ? c[k2] : cmul(c[k2], w);
    v[k2] = cadd(a[k2], t);
    v[k2 + 16] = csub(a[k2], t);
  }
}

// dft32 restricted to the outputs 0..3 and 28..31 (from the half-DFT16s'
// outputs 0..3 and 12..15).
__device__ __forceinline__ void dft32_ends(float2* v) {
  float2 a[16], c[16];
#pragma unroll
  for (int n2 = 0; n2 < 16; ++n2) {
    a[n2] = v[2 * n2];
    c[n2] = v[2 * n2 + 1];
  }
  dft16_ends(a);
  dft16_ends(c);
#pragma unroll
  for (int k2 = 0; k2 < 16; ++k2) {
    if (k2 >= 4 && k2 < 12) continue;
    const double ang = -2.0 * 3.14159265358979323846 * k2 / 32.0;
    const float2 w = make_float2((float)__builtin_cos(ang), (float)__builtin_sin(ang));
    const float2 t = (k2 == 0) ? c[k2] : cmul(c[k2], w);
    if (k2 < 4) v[k2] = cadd(a[k2], t);
    else v[k2 + 16] = csub(a[k2], t);
  }
}

// dft32 of an input that is zero except at n = 0..3 and 28..31 (all 32
// outputs): the even and odd halves are zero except at n2 = 0, 1, 14, 15.
__device__ __forceinline__ void dft32_in_ends(float2* v) {
  float2 a[16], c[16];
#pragma unroll
  for (int n2 = 0; n2 < 16; ++n2) {
    a[n2] = v[2 * n2];
    c[n2] = v[2 * n2 + 1];
  }
  dft16_in_ends(a);
  dft16_in_ends(c);
#pragma unroll
  for (int k2 = 0; k2 < 16; ++k2) {
    const double ang = -2.0 * 3.14159265358979323846 * k2 / 32.0;
    const float2 w = make_float2((float)__builtin_cos(ang), (float)__builtin_sin(ang));
    const float2 t = (k2 == 0) ? c[k2] : cmul(c[k2], w);
    v[k2] = cadd(a[k2], t);
    v[k2 + 16] = csub(a[k2], t);
  }
}

// Radix set of a factorisation: 1 only 8/4/2, 0 also 3/5/7, 2 composite radices.
__host__ __device__ inline int fft_radix_set(const FftDesc& d) {
  int rs = 1;
  for (int s = 0; s < d.nst; ++s) {
    const int r = d.radix[s];
    if (r == 6 || r == 9 || r == 10 || r == 12) return 2;
    if (r != 2 && r != 4 && r != 8) rs = 0;
  }
  return rs;
}

__host__ __device__ inline bool fft_is_pow2(const FftDesc& d) {
  for (int s = 0; s < d.nst; ++s)
    if (d.radix[s] != 2 && d.radix[s] != 4 && d.radix[s] != 8) return false;
  return true;
}

}  // namespace saxs
