// Per-frame q tables built on the device, for batches whose frames have
// different boxes (NPT trajectories; NVT with frame 0's binning).
//
// For every frame f and key (shell n = |m|^2, point (|m_x|, |m_y|, |m_z|), or
// signed point, as shell_key), k_frame_tables writes
//   bin[f][key]   = max{b : b^2 dq^2 <= q_b^2} if q^2 <= qcut^2 and b < num_bins,
//                   else -1   (q: the frame's own; q_b: the binning geometry)
//   ff[f][e][key] = f_e(q) = sum_l a_l exp(-b_l (q / 4 pi)^2) + c
// the reference's per-frame q of each point (saxsDeviceKernels.cu:84-91,
// 196-205 with the frame's oc, RunSaxs.cu:180-193) with dq, qcut and the bin
// count of the first frame (saxsKernel.cu:489-526), Cromer-Mann factors as
// opsfact.cu:29-35.  The bin and cut decisions are exact on the float32
// values of the geometry, dq and qcut (the rule the host tables use): each
// float is m 2^e with a 24-bit integer m, q^2 is formed as an unsigned
// 128-bit integer times a power of two, and the comparisons align the
// exponents without rounding.  The host sends geometries whose nonzero
// entries span at most kTabMaxShift binary orders (keeps every product below
// 2^128); others take the per-box path.
#pragma once
#include <hip/hip_runtime.h>

#include "saxs_kernels.hpp"

namespace saxs {

typedef unsigned __int128 u128;

constexpr int kTabMaxShift = 20;

// v = m 2^e with m < 2^24 (v >= 0; m = 0 for v = 0).
struct ExactF {
  unsigned long long m;
  int e;
};

__device__ __forceinline__ ExactF exact_of(float v) {
  ExactF r{0ull, 0};
  const float a = fabsf(v);
  if (a == 0.0f) return r;
  int ex = 0;
  const float fr = frexpf(a, &ex);                 // a = fr 2^ex, fr in [0.5, 1)
  r.m = (unsigned long long)(fr * 16777216.0f);    // exact: 24-bit significand
  r.e = ex - 24;
  return r;
}

__device__ __forceinline__ int bitlen128(u128 x) {
  const unsigned long long hi = (unsigned long long)(x >> 64), lo = (unsigned long long)x;
  if (hi) return 128 - __clzll(hi);
  if (lo) return 64 - __clzll(lo);
  return 0;
}

// T 2^E <= R 2^F, exactly.
__device__ __forceinline__ bool le_exact(u128 T, int E, u128 R, int F) {
  if (T == 0) return true;
  if (R == 0) return false;
  const int lt = bitlen128(T) + E, lr = bitlen128(R) + F;
  if (lt != lr) return lt < lr;
  if (E > F) T <<= (E - F);                         // same top bit: no overflow
  else R <<= (F - E);
  return T <= R;
}

// q^2 of key m in geometry g (key_type 0: g[0] = s; 1: g[0..2] = s_a; 2: rows
// b_aj) as T 2^E, plus its double value.
__device__ __forceinline__ void q2_exact(int key_type, const float* g, int mx, int my, int mz,
                                         u128& T, int& E, double& qd) {
  if (key_type == 0) {
    const ExactF s = exact_of(g[0]);
    const unsigned long long n = (unsigned long long)(mx * mx + my * my + mz * mz);
    T = (u128)n * (u128)(s.m * s.m);
    E = 2 * s.e;
    qd = (double)n * (double)g[0] * (double)g[0];
    return;
  }
  if (key_type == 1) {
    const int mm[3] = {mx, my, mz};
    ExactF s[3];
    int emin = 1 << 20;
    for (int a = 0; a < 3; ++a) {
      s[a] = exact_of(g[a]);
      if (s[a].m && mm[a]) emin = min(emin, s[a].e);
    }
    T = 0;
    qd = 0.0;
    for (int a = 0; a < 3; ++a) {
      if (!s[a].m || !mm[a]) continue;
      const unsigned long long m2 = (unsigned long long)(mm[a] * mm[a]);
      T += ((u128)m2 * (u128)(s[a].m * s[a].m)) << (2 * (s[a].e - emin));
      qd += (double)m2 * (double)g[a] * (double)g[a];
    }
    E = T ? 2 * emin : 0;
    return;
  }
  // signed point: q_j = sum_a m_a b_aj
  const int mm[3] = {mx, my, mz};
  int emin = 1 << 20;
  for (int a = 0; a < 3; ++a)
    for (int j = 0; j < 3; ++j) {
      const ExactF v = exact_of(g[3 * a + j]);
      if (v.m && mm[a]) emin = min(emin, v.e);
    }
  T = 0;
  qd = 0.0;
  for (int j = 0; j < 3; ++j) {
    __int128 qj = 0;
    double qjd = 0.0;
    for (int a = 0; a < 3; ++a) {
      const float bv = g[3 * a + j];
      const ExactF v = exact_of(bv);
      if (!v.m || !mm[a]) continue;
      const __int128 t = ((__int128)mm[a] * (__int128)v.m) << (v.e - emin);
      qj += bv < 0.0f ? -t : t;
      qjd += (double)mm[a] * (double)bv;
    }
    const u128 aq = (u128)(qj < 0 ? -qj : qj);
    T += aq * aq;
    qd += qjd * qjd;
  }
  E = T ? 2 * emin : 0;
}

// Frequencies of a table key (the inverse of shell_key's point layouts); for
// shells the key is n itself (returned as mx = 0, my = 0, mz^2 = n is not a
// lattice decomposition: q^2 = n s^2 is formed from n directly).
__device__ __forceinline__ void key_point(const Params& p, int key, int& mx, int& my, int& mz) {
  if (p.point_keys == 1) {
    const int kz1 = p.kmax[2] + 1, ky1 = p.kmax[1] + 1;
    mz = key % kz1;
    const int r = key / kz1;
    my = r % ky1;
    mx = r / ky1;
  } else {
    const int kz1 = 2 * p.kmax[2] + 1, ky1 = 2 * p.kmax[1] + 1;
    mz = key % kz1 - p.kmax[2];
    const int r = key / kz1;
    my = r % ky1 - p.kmax[1];
    mx = r / ky1 - p.kmax[0];
  }
}

// geom [F][9], bgeom [9] or NULL (bins by each frame's own q), coef [E][9].
__global__ void __launch_bounds__(kWG) k_frame_tables(Params p, const float* __restrict__ geom,
                                                      const float* __restrict__ bgeom,
                                                      const float* __restrict__ coef, float dq,
                                                      float qcut, int num_bins, int* bin_out,
                                                      float* ff_out) {
  const int key = blockIdx.x * blockDim.x + threadIdx.x, f = blockIdx.y;
  if (key >= p.lut_len) return;
  const int kt = p.point_keys;
  int mx = 0, my = 0, mz = 0;
  if (kt == 0) {
    // q^2 = n s^2: the shell enters as (0, 0, 0) with n carried separately
  } else {
    key_point(p, key, mx, my, mz);
  }
  const float* g = geom + (size_t)f * 9;
  u128 T;
  int E;
  double qd;
  if (kt == 0) {
    const ExactF s = exact_of(g[0]);
    T = (u128)(unsigned long long)key * (u128)(s.m * s.m);
    E = 2 * s.e;
    qd = (double)key * (double)g[0] * (double)g[0];
  } else {
    q2_exact(kt, g, mx, my, mz, T, E, qd);
  }
  const ExactF Q = exact_of(qcut);
  const bool in = le_exact(T, E, (u128)(Q.m * Q.m), 2 * Q.e);
  int bin = -1;
  if (in) {
    u128 Tb = T;
    int Eb = E;
    double qb = qd;
    if (bgeom) {
      if (kt == 0) {
        const ExactF s = exact_of(bgeom[0]);
        Tb = (u128)(unsigned long long)key * (u128)(s.m * s.m);
        Eb = 2 * s.e;
        qb = (double)key * (double)bgeom[0] * (double)bgeom[0];
      } else {
        q2_exact(kt, bgeom, mx, my, mz, Tb, Eb, qb);
      }
    }
    const bool inb = le_exact(Tb, Eb, (u128)(Q.m * Q.m), 2 * Q.e);
    if (inb) {
      const ExactF D = exact_of(dq);
      const u128 D2 = (u128)(D.m * D.m);
      int b = (int)floor(sqrt(qb) / (double)dq);
      b = max(0, min(b, num_bins));
      while (b > 0 && !le_exact((u128)(unsigned long long)(b * b) * D2, 2 * D.e, Tb, Eb)) --b;
      while (le_exact((u128)(unsigned long long)((b + 1) * (b + 1)) * D2, 2 * D.e, Tb, Eb) &&
             b < num_bins)
        ++b;
      bin = b < num_bins ? b : -1;
    }
  }
  bin_out[(size_t)f * p.lut_len + key] = bin;
  const double s2 = qd / (16.0 * M_PI * M_PI);
  for (int e = 0; e < p.E; ++e) {
    const float* c = coef + 9 * e;
    const double v = (double)c[0] * exp(-(double)c[4] * s2) + (double)c[1] * exp(-(double)c[5] * s2) +
                     (double)c[2] * exp(-(double)c[6] * s2) + (double)c[3] * exp(-(double)c[7] * s2) +
                     (double)c[8];
    ff_out[((size_t)f * p.E + e) * p.lut_len + key] = (float)v;
  }
}

}  // namespace saxs
