// Adjoint of the forward scattering path: dL/dx per atom for a loss whose
// gradient with respect to the per-frame histogram, gH = dL/dH_b, is given.
// The reference has no gradient; this is the exact transpose of the forward
// kernels in saxs_kernels.hpp:
//
//   forward  rho_e -(R2C z)-> Z_e -(FFT y)-> Y_e -(FFT x)-> F_e
//            A = sum_e f_e F_e, I = |A|^2 mod, H_b = sum_p w_p I_p (w = 2,
//            origin 1)
//   adjoint  Abar = 2 mod w gH_b A, Fbar_e = f_e Abar
//            Ybar_e = x-transpose (inverse DFT over kept m_x, crop x < n)
//            Zbar_e = y-transpose (same over m_y)
//            drho~_e(z) = Re sum_{k < kz} Zbar_e(k) e^{+2 pi i k z / nn}
//            (each half-spectrum point once: the transpose of R2C, not a
//            Hermitian C2R)
//            dL/dx_j = sum_stencil drho~ grad(theta theta theta)
//                      - (sum_cell drho~ / count) grad(border weight of j)
//
// The forward spectra are recomputed (spread_z, yfft) rather than stored;
// the x-pass adjoint recomputes F_e on the fly and overwrites Y with Ybar in
// place, the y-pass adjoint overwrites Z with Zbar.
#pragma once
#include "saxs_kernels.hpp"

namespace saxs {

struct AdjBuffers {
  const double* gH;     // [F][num_bins]
  float* grad;          // [F][N][3]
  double* padg;         // [F][E] sum over the cell of drho~
  double* padt;         // [F][E][ntx nty] the same sum over each column tile
  float* gpart;         // [F][nsx*nsy][N][3] per-tile partial gradients, entry
                        // (f, slot, r): record r of frame f's sorted run
  int nsx, nsy;         // tile offsets an atom's stencil reaches along x, y
};

// A sorted record as the z adjoints stage it: w carries the record's
// position in the sorted array (pos) instead of the atom index, so the
// gather writes each partial gradient at (frame, slot, pos - f N): the
// records of one source segment are consecutive, and so are the partials
// one tile writes for them.
__device__ __forceinline__ float4 staged_record(const Buffers& b, int pos) {
  float4 r = b.sorted[pos];
  r.w = __int_as_float(pos);
  return r;
}

// Slot of the partial gradient that tile (bx, by) writes for an atom in cell
// (mx, my): the tile's offset from the atom's home tile (the tile holding
// stencil point o = 0), ox + nsx oy.  Every (frame, slot, atom) has one
// writer, so the gather stores instead of adding atomically and the sum
// over slots (k_grad_finalize) is in a fixed order.  The tile divisions are
// exact float-reciprocal quotients (div_small: coordinates < 2^22).
template <int P>
__device__ __forceinline__ int partial_slot(int mx, int my, int bx, int by, const Params& p,
                                            const AdjBuffers& g) {
  int ox = bx - div_small(wrap_base(mx, P, p.n[0]), p.tx, 1.0f / (float)p.tx);
  int oy = by - div_small(wrap_base(my, P, p.n[1]), p.ty, 1.0f / (float)p.ty);
  if (ox < 0) ox += p.ntx;
  if (oy < 0) oy += p.nty;
  return ox + g.nsx * oy;
}

__device__ __forceinline__ int kept_to_index(int j, int k, int K, int N) {
  return (k == N) ? j : (j <= K ? j : N - (k - j));
}

// padg[fe] = sum over the column tiles of padt[fe][tile], in a fixed order
// (each lane sums tiles lane, lane + 64, ..., then a fixed shuffle tree), so
// the pad term of the gradient does not depend on the order in which the
// tiles ran.  One wavefront per (frame, element).
__global__ void __launch_bounds__(kWG) k_pad_reduce(const double* padt, double* padg, int ntiles,
                                                    int nfe) {
  const int fe = blockIdx.x * (kWG / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (fe >= nfe) return;
  const double* src = padt + (size_t)fe * ntiles;
  double s = 0.0;
  for (int t = lane; t < ntiles; t += 64) s += src[t];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) padg[fe] = s;
}

// x-pass adjoint: per (kept m_y, kz chunk, frame).  Recomputes A from Y,
// forms Abar, and for every element writes Ybar_e over ybuf in place.
template <int POW2>
__global__ void __launch_bounds__(kWG) k_adj_x(Params p, Buffers b, AdjBuffers g) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int jy = blockIdx.x, ch = blockIdx.y, f = blockIdx.z;
  const int k0 = ch * p.kcx;
  const int kc = min(p.kcx, p.kz - k0);
  const int nx = p.n[0], nnx = p.nn[0];
  const int nb = p.num_bins;
  float2* buf = reinterpret_cast<float2*>(smem);
  float2* tw = buf + (size_t)p.kcx * p.lsx;
  for (int i = threadIdx.x; i < nnx; i += blockDim.x) tw[i] = b.twx[i];
  const int iy = kept_to_index(jy, p.ky, p.Ky, p.nn[1]);
  const int my = iy < p.nf[1] ? iy : iy - p.nn[1];
  const int nout = p.kx * kc;
  float2 A[kOutMax];
#pragma unroll
  for (int u = 0; u < kOutMax; ++u) A[u] = make_float2(0.0f, 0.0f);

  for (int e = 0; e < p.E; ++e) {
    const float2* src = b.ybuf + ((((size_t)f * p.E + e) * p.ky + jy) * nx) * p.kz;
    for (int i = threadIdx.x; i < nx * kc; i += blockDim.x) {
      const int x = i / kc, k = i - x * kc;
      buf[k * p.lsx + x] = src[(size_t)x * p.kz + k0 + k];
    }
    for (int i = threadIdx.x; i < (nnx - nx) * kc; i += blockDim.x) {
      const int x = nx + i / kc, k = i % kc;
      buf[k * p.lsx + x] = make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    fft_lines<kMaxC, POW2>(buf, p.lsx, kc, p.fx, tw);
    const float* ff = b.lut_ff + (size_t)e * p.lut_len;
#pragma unroll
    for (int u = 0; u < kOutMax; ++u) {
      const int idx = threadIdx.x + u * blockDim.x;
      if (idx < nout) {
        const int jx = idx / kc, k = idx - jx * kc;
        const int ix = kept_to_index(jx, p.kx, p.Kx, nnx);
        const int mx = ix < p.nf[0] ? ix : ix - nnx;
        const int mz = k0 + k;
        const int nsq = shell_key(mx, my, mz, p);
        if (nsq < p.lut_len) {
          const float fv = ff[nsq];
          const float2 Fv = buf[k * p.lsx + ix];
          A[u].x += fv * Fv.x;
          A[u].y += fv * Fv.y;
        }
      }
    }
    __syncthreads();
  }
  // Abar = 2 * mod * w * gH_b * A at in-range points, 0 elsewhere
  int nsqs[kOutMax];
#pragma unroll
  for (int u = 0; u < kOutMax; ++u) {
    const int idx = threadIdx.x + u * blockDim.x;
    float2 ab = make_float2(0.0f, 0.0f);
    int nsq = 1 << 30;
    if (idx < nout) {
      const int jx = idx / kc, k = idx - jx * kc;
      const int ix = kept_to_index(jx, p.kx, p.Kx, nnx);
      const int mx = ix < p.nf[0] ? ix : ix - nnx;
      const int kz = k0 + k;
      nsq = shell_key(mx, my, kz, p);
      if (nsq < p.lut_len) {
        const int bin = b.lut_bin[nsq];
        if (bin >= 0 && bin < nb) {
          const float mod = b.modx[ix] * b.mody[iy] * b.modz[kz];
          const bool origin = (ix == 0) && (iy == 0) && (kz == 0);
          const float s = 2.0f * mod * (origin ? 1.0f : 2.0f) * (float)g.gH[(size_t)f * nb + bin];
          ab = make_float2(s * A[u].x, s * A[u].y);
        }
      }
    }
    A[u] = ab;
    nsqs[u] = nsq;
  }
  // per element: Fbar_e = f_e Abar at the kept m_x, inverse DFT along x,
  // crop to x < n_x, overwrite Y_e with Ybar_e
  for (int e = 0; e < p.E; ++e) {
    for (int i = threadIdx.x; i < kc * nnx; i += blockDim.x) {
      const int k = i / nnx, x = i - k * nnx;
      buf[k * p.lsx + x] = make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    const float* ff = b.lut_ff + (size_t)e * p.lut_len;
#pragma unroll
    for (int u = 0; u < kOutMax; ++u) {
      const int idx = threadIdx.x + u * blockDim.x;
      if (idx < nout && nsqs[u] < p.lut_len) {
        const int jx = idx / kc, k = idx - jx * kc;
        const int ix = kept_to_index(jx, p.kx, p.Kx, nnx);
        const float fv = ff[nsqs[u]];
        buf[k * p.lsx + ix] = make_float2(fv * A[u].x, -fv * A[u].y);   // conj
      }
    }
    __syncthreads();
    fft_lines<kMaxC, POW2>(buf, p.lsx, kc, p.fx, tw);
    float2* dst = b.ybuf + ((((size_t)f * p.E + e) * p.ky + jy) * nx) * p.kz;
    for (int i = threadIdx.x; i < nx * kc; i += blockDim.x) {
      const int x = i / kc, k = i - x * kc;
      const float2 v = buf[k * p.lsx + x];
      dst[(size_t)x * p.kz + k0 + k] = make_float2(v.x, -v.y);
    }
    __syncthreads();
  }
}

// Pruned x-pass adjoint for N = nn_x = 16 R2, same workgroups, lines and
// items as k_xfft_p (kept outputs at the ends, Kx < N / 8).  Phase 1
// recomputes A = sum_e f_e F_e at the kept m_x from Y; Abar = 2 mod w gH_b A;
// phase 2, per element: c = conj(f_e Abar) at the kept m = s + 16 r',
// U(x) = sum_m c(m) W_N^{m x} by the transposed stages (x = j + R2 r,
// W^{m x} = W_N^{s j} W_16^{s r} W_R2^{r' j}: radix-R2 DFT over r' times
// W_N^{s j}, LDS exchange, radix-16 DFT over s), Ybar(x) = conj(U(x)) for
// x < n_x written over Y_e.
template <int N>
__host__ __device__ inline size_t adjx_p_lds_bytes(const Params& p) {
  return (size_t)p.kcx * (17 * (N / 16) + 1) * sizeof(float2) + N * sizeof(float2);
}

// LOAD: phase 1 reads A from `amp`, saved by the forward k_xfft_p<N, true,
// HALF, true> of the same frames (same workgroups, threads and items, so
// the same values) instead of recomputing it from Y.
template <int N, bool HALF, bool LOAD = false>
__global__ void __launch_bounds__(512) k_adj_x_p(Params p, Buffers b, AdjBuffers g,
                                                 const float2* amp = nullptr) {
  constexpr int R2 = N / 16;
  constexpr int LY = 17 * R2 + 1;
  constexpr int KL = N / 128;
  constexpr int NO = 2 * KL;
  extern __shared__ __align__(16) unsigned char smem[];
  const int jy = blockIdx.x, ch = blockIdx.y, f = blockIdx.z;
  const int kz = p.kz, k0 = ch * p.kcx, kc = min(p.kcx, kz - k0);
  const int nx = p.n[0], nb = p.num_bins, Kx = p.Kx;
  float2* lines = reinterpret_cast<float2*>(smem);
  float2* tw = lines + (size_t)p.kcx * LY;            // W_N^{j s} stored [j][s]
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = b.twx[(i >> 4) * (i & 15)];
  const int iy = kept_to_index(jy, p.ky, p.Ky, p.nn[1]);
  const int my = iy < p.nf[1] ? iy : iy - p.nn[1];
  float2 A[kXItems][NO];
  int nsq_o[kXItems][NO];
#pragma unroll
  for (int u = 0; u < kXItems; ++u)
#pragma unroll
    for (int o = 0; o < NO; ++o) {
      A[u][o] = make_float2(0.0f, 0.0f);
      nsq_o[u][o] = 1 << 30;
      const int item = threadIdx.x + u * blockDim.x;
      if (item < kc * 16) {
        const int s = item / kc, k = item - s * kc;
        const int rr = o < KL ? o : R2 - NO + o;
        const int m = s + 16 * rr;
        if (m <= Kx || m >= N - Kx) {
          const int mx = m < p.nf[0] ? m : m - N;
          nsq_o[u][o] = shell_key(mx, my, k0 + k, p);
        }
      }
    }
  __syncthreads();
  // phase 1: A at the kept outputs
  if constexpr (LOAD) {
    const int blk = (f * p.ky + jy) * p.nchx + ch;
#pragma unroll
    for (int u = 0; u < kXItems; ++u)
      if (u * (int)blockDim.x < kc * 16) {           // the same for the whole workgroup
#pragma unroll
        for (int o = 0; o < NO; ++o) A[u][o] = amp[amp_index(blk, u, o, NO)];
      }
  }
  for (int e = 0; e < (LOAD ? 0 : p.E); ++e) {
    const float2* src = b.ybuf + ((((size_t)f * p.E + e) * p.ky + jy) * nx) * kz + k0;
    for (int item = threadIdx.x; item < kc * R2; item += blockDim.x) {
      const int j = item / kc, k = item - j * kc;
      float2 v[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) v[r] = make_float2(0.0f, 0.0f);
#pragma unroll
      for (int r = 0; r < (HALF ? 8 : 16); ++r) {
        const int x = j + R2 * r;
        if (x < nx) v[r] = src[(size_t)x * kz + k];
      }
      dft16<HALF>(v);
      float2* dst = lines + k * LY + 17 * j;
#pragma unroll
      for (int r = 0; r < 16; ++r) dst[r] = v[r];
    }
    __syncthreads();
    const float* ff = b.lut_ff + (size_t)e * p.lut_len;
#pragma unroll
    for (int u = 0; u < kXItems; ++u) {
      const int item = threadIdx.x + u * blockDim.x;
      if (item < kc * 16) {
        const int s = item / kc, k = item - s * kc;
        float2 t[R2];
#pragma unroll
        for (int r = 0; r < R2; ++r) {
          const float2 y = lines[k * LY + 17 * r + s];
          t[r] = (r == 0) ? y : cmul(y, tw[r * 16 + s]);
        }
        if constexpr (R2 == 8) Radix<8>::dft(t);
        else if constexpr (R2 == 16) dft16<false>(t);
        else dft32(t);
#pragma unroll
        for (int o = 0; o < NO; ++o) {
          const int rr = o < KL ? o : R2 - NO + o;
          if (nsq_o[u][o] < p.lut_len) {
            const float fv = ff[nsq_o[u][o]];
            A[u][o].x += fv * t[rr].x;
            A[u][o].y += fv * t[rr].y;
          }
        }
      }
    }
    __syncthreads();
  }
  // Abar = 2 * mod * w * gH_b * A at in-range points, 0 elsewhere
#pragma unroll
  for (int u = 0; u < kXItems; ++u) {
    const int item = threadIdx.x + u * blockDim.x;
    const int s = item / max(kc, 1), k = item - s * kc;
#pragma unroll
    for (int o = 0; o < NO; ++o) {
      float2 ab = make_float2(0.0f, 0.0f);
      const int nsq = nsq_o[u][o];
      if (item < kc * 16 && nsq < p.lut_len) {
        const int bin = b.lut_bin[nsq];
        if (bin >= 0 && bin < nb) {
          const int rr = o < KL ? o : R2 - NO + o;
          const int m = s + 16 * rr, mz = k0 + k;
          const float mod = b.modx[m] * b.mody[iy] * b.modz[mz];
          const bool origin = (m == 0) && (iy == 0) && (mz == 0);
          const float sc = 2.0f * mod * (origin ? 1.0f : 2.0f) * (float)g.gH[(size_t)f * nb + bin];
          ab = make_float2(sc * A[u][o].x, sc * A[u][o].y);
        }
      }
      A[u][o] = ab;
    }
  }
  // phase 2: per element, Ybar_e over Y_e
  for (int e = 0; e < p.E; ++e) {
    const float* ff = b.lut_ff + (size_t)e * p.lut_len;
#pragma unroll
    for (int u = 0; u < kXItems; ++u) {
      const int item = threadIdx.x + u * blockDim.x;
      if (item < kc * 16) {
        const int s = item / kc, k = item - s * kc;
        float2 t[R2];
#pragma unroll
        for (int r = 0; r < R2; ++r) t[r] = make_float2(0.0f, 0.0f);
#pragma unroll
        for (int o = 0; o < NO; ++o) {
          const int rr = o < KL ? o : R2 - NO + o;
          if (nsq_o[u][o] < p.lut_len) {
            const float fv = ff[nsq_o[u][o]];
            t[rr] = make_float2(fv * A[u][o].x, -fv * A[u][o].y);   // conj(f Abar)
          }
        }
        // nonzero inputs only at the KL = N/128 kept rows of each end
        if constexpr (R2 == 8) Radix<8>::dft(t);
        else if constexpr (R2 == 16) dft16_in_ends(t);
        else dft32_in_ends(t);
        float2* dst = lines + k * LY + s;
#pragma unroll
        for (int j = 0; j < R2; ++j) dst[17 * j] = (j == 0) ? t[0] : cmul(t[j], tw[j * 16 + s]);
      }
    }
    __syncthreads();
    float2* ydst = b.ybuf + ((((size_t)f * p.E + e) * p.ky + jy) * nx) * kz + k0;
    for (int item = threadIdx.x; item < kc * R2; item += blockDim.x) {
      const int j = item / kc, k = item - j * kc;
      float2 v[16];
      const float2* srcl = lines + k * LY + 17 * j;
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = srcl[q];
      dft16<false>(v);
#pragma unroll
      for (int r = 0; r < (HALF ? 8 : 16); ++r) {
        const int x = j + R2 * r;
        if (x < nx) ydst[(size_t)x * kz + k] = make_float2(v[r].x, -v[r].y);
      }
    }
    __syncthreads();
  }
}

__host__ __device__ inline size_t adjx_lds_bytes(const Params& p) {
  return (size_t)p.kcx * p.lsx * sizeof(float2) + (size_t)p.nn[0] * sizeof(float2);
}

// y-pass adjoint: per (x, kz chunk, frame*E + e): Zbar(x, y, kz) for y < n_y
// from Ybar over the kept m_y; writes zbuf.
template <int POW2>
__global__ void __launch_bounds__(kWG) k_adj_y(Params p, Buffers b) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int x = blockIdx.x, ch = blockIdx.y;
  const size_t fe = blockIdx.z;
  const int k0 = ch * p.kcy;
  const int kc = min(p.kcy, p.kz - k0);
  const int ny = p.n[1], nny = p.nn[1];
  float2* buf = reinterpret_cast<float2*>(smem);
  float2* tw = buf + (size_t)p.kcy * p.lsy;
  for (int i = threadIdx.x; i < nny; i += blockDim.x) tw[i] = b.twy[i];
  for (int i = threadIdx.x; i < kc * nny; i += blockDim.x) {
    const int k = i / nny, y = i - k * nny;
    buf[k * p.lsy + y] = make_float2(0.0f, 0.0f);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < p.ky * kc; i += blockDim.x) {
    const int j = i / kc, k = i - j * kc;
    const int iy = kept_to_index(j, p.ky, p.Ky, nny);
    const float2 v = b.ybuf[((fe * p.ky + j) * p.n[0] + x) * p.kz + k0 + k];
    buf[k * p.lsy + iy] = make_float2(v.x, -v.y);
  }
  __syncthreads();
  fft_lines<kMaxC, POW2>(buf, p.lsy, kc, p.fy, tw);
  float2* dst = b.zbuf + (fe * p.n[0] + x) * (size_t)ny * p.kz;
  for (int i = threadIdx.x; i < ny * kc; i += blockDim.x) {
    const int y = i / kc, k = i - y * kc;
    const float2 v = buf[k * p.lsy + y];
    dst[(size_t)y * p.kz + k0 + k] = make_float2(v.x, -v.y);
  }
}

// Pruned y-pass adjoint for N = nn_y = 16 R2, the transpose of k_yfft_p:
// U(y) = sum over kept m of conj(Ybar(m)) W_N^{m y}, Zbar(y) = conj(U(y)) for
// y < n_y.  With m = s + 16 r' and y = j + R2 r, W^{m y} = W_N^{s j}
// W_16^{s r} W_R2^{r' j}: stage A item (line, s) loads the kept r' and forms
// the radix-R2 DFT over r' times W_N^{s j}; one LDS exchange (T_s[j] at
// 17 j + s); stage B item (line, j) forms the radix-16 DFT over s and writes
// the y < n_y outputs.  Lines as in k_yfft_p; a workgroup takes
// kAdjYGroups consecutive line groups and issues the next group's loads
// before the current group's stage B, so they are in flight while it
// computes and stores (a thread has only its few kept-row loads otherwise).
constexpr int kAdjYGroups = 2;

template <int N, bool ENDS, bool HALF>
__global__ void __launch_bounds__(kWG) k_adj_y_p(Params p, Buffers b) {
  constexpr int R2 = N / 16;
  constexpr int LY = 17 * R2 + 1;
  constexpr int LW = yfft_p_lines<N>();
  constexpr int IPT = (LW * 16 + kWG - 1) / kWG;      // stage-A items per thread
  extern __shared__ __align__(16) unsigned char smem[];
  float2* tbuf = reinterpret_cast<float2*>(smem);
  float2* tw = tbuf + (size_t)LW * LY;                 // W_N^{j s} stored [j][s]
  // neighbouring groups share the 128-byte lines where the kz-long rows of
  // zbuf straddle group boundaries: each XCD (own L2) takes a contiguous range
  const unsigned grp = xcd_contiguous(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const size_t fe = grp / gridDim.x;
  const int nx = p.n[0], ny = p.n[1], kz = p.kz;
  const int nlines = nx * kz;
  const int gbase = (int)(grp % gridDim.x) * kAdjYGroups * LW;
  const int Ky = p.Ky, ky = p.ky;
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = b.twy[(i >> 4) * (i & 15)];
  float2 t[IPT][R2];
  // stage-A inputs of the group starting at line g0
  auto load = [&](int g0) {
#pragma unroll
    for (int u = 0; u < IPT; ++u) {
      const int item = threadIdx.x + u * kWG;
      const int l = item % LW, s = item / LW;
      const int g = g0 + l;
#pragma unroll
      for (int rr = 0; rr < R2; ++rr) t[u][rr] = make_float2(0.0f, 0.0f);
      if (item < LW * 16 && g < nlines) {
        const int x = g / kz, k = g - x * kz;
        const float2* in = b.ybuf + (fe * ky * nx + x) * kz + k;   // + jy nx kz
#pragma unroll
        for (int rr = 0; rr < R2; ++rr) {
          // ENDS: Ky < N / 8, kept m only at r' < KL or r' >= R2 - KL
          if (ENDS && rr >= N / 128 && rr < R2 - N / 128) continue;
          const int m = s + 16 * rr;
          if (m <= Ky || m >= N - Ky) {
            const int jy = m <= Ky ? m : ky - (N - m);
            const float2 v = in[(size_t)jy * nx * kz];
            t[u][rr] = make_float2(v.x, -v.y);
          }
        }
      }
    }
  };
  load(gbase);
  __syncthreads();
  for (int gi = 0; gi < kAdjYGroups; ++gi) {
    const int g0 = gbase + gi * LW;
    if (g0 >= nlines) break;                            // the same for the whole workgroup
#pragma unroll
    for (int u = 0; u < IPT; ++u) {
      const int item = threadIdx.x + u * kWG;
      if (item < LW * 16) {
        const int l = item % LW, s = item / LW;
        if constexpr (R2 == 8) Radix<8>::dft(t[u]);
        else if constexpr (ENDS && R2 == 16) dft16_in_ends(t[u]);    // rows rr < 2 or >= 14 only
        else if constexpr (ENDS && R2 == 32) dft32_in_ends(t[u]);    // rows rr < 4 or >= 28 only
        else if constexpr (R2 == 16) dft16<false>(t[u]);
        else dft32(t[u]);
        float2* dst = tbuf + l * LY + s;
#pragma unroll
        for (int j = 0; j < R2; ++j) dst[17 * j] = (j == 0) ? t[u][0] : cmul(t[u][j], tw[j * 16 + s]);
      }
    }
    __syncthreads();
    if (gi + 1 < kAdjYGroups) load(g0 + LW);           // in flight during stage B
    {
      const int l = threadIdx.x % LW, j = threadIdx.x / LW;     // j < R2
      const int g = g0 + l;
      float2 v[16];
      const float2* srcl = tbuf + l * LY + 17 * j;
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = srcl[q];
      dft16<false>(v);
      if (g < nlines) {
        const int x = g / kz, k = g - x * kz;
        float2* dst = b.zbuf + (fe * nx + x) * (size_t)ny * kz + k;
#pragma unroll
        for (int r = 0; r < (HALF ? 8 : 16); ++r) {      // HALF: 2 n_y <= N
          const int y = j + R2 * r;
          if (y < ny) dst[(size_t)y * kz] = make_float2(v[r].x, -v[r].y);
        }
      }
    }
    __syncthreads();                                    // tbuf is rewritten next group
  }
}

// LDS of k_adj_z: drho~ columns, FFT lines (the staged atom records alias
// them after the transform), twiddles.
__host__ __device__ inline size_t adjz_lds_bytes(const Params& p) {
  const size_t drho = (((size_t)p.tx * p.ty * p.n[2] * sizeof(float)) + 15) & ~size_t(15);
  size_t buf = (size_t)(p.tx * p.ty / 2) * p.lsz * sizeof(float2);
  const size_t recs = (size_t)kSpreadRecs * sizeof(float4);
  if (buf < recs) buf = recs;
  return drho + buf + (size_t)p.nn[2] * sizeof(float2);
}

// z-pass adjoint + B-spline gradient gather: per (column tile, e, f).
template <int P, int POW2>
__global__ void __launch_bounds__(kWG) k_adj_z(Params p, Buffers b, AdjBuffers g) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int tile = blockIdx.x, e = blockIdx.y, f = blockIdx.z;
  const int bx = tile % p.ntx, by = tile / p.ntx;
  const int X0 = bx * p.tx, Y0 = by * p.ty;
  const int wx = min(p.tx, p.n[0] - X0), wy = min(p.ty, p.n[1] - Y0);
  const int nz = p.n[2], nnz = p.nn[2];
  const int ncol = p.tx * p.ty;
  const int nl = ncol / 2;
  float* drho = reinterpret_cast<float*>(smem);
  float2* buf = reinterpret_cast<float2*>(smem + ((ncol * nz * sizeof(float) + 15) & ~size_t(15)));
  float4* recs = reinterpret_cast<float4*>(buf);
  const size_t buf_bytes = (size_t)nl * p.lsz * sizeof(float2);
  float2* tw = reinterpret_cast<float2*>(reinterpret_cast<unsigned char*>(buf) +
                                          (buf_bytes > kSpreadRecs * sizeof(float4)
                                               ? buf_bytes : kSpreadRecs * sizeof(float4)));
  __shared__ int s_start[9], s_cnt[9], s_pref[10];
  __shared__ double s_red[kWG / 64];
  if (threadIdx.x < 9) {
    const int2 r = tile_slot(p, b, f, e, bx, by, threadIdx.x);
    s_start[threadIdx.x] = r.x;
    s_cnt[threadIdx.x] = r.y;
  }
  for (int i = threadIdx.x; i < nnz; i += blockDim.x) tw[i] = b.twz[i];
  for (int i = threadIdx.x; i < nl * nnz; i += blockDim.x) {
    const int l = i / nnz, z = i - l * nnz;
    buf[l * p.lsz + z] = make_float2(0.0f, 0.0f);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int k = 0; k < 9; ++k) {
      s_pref[k] = acc;
      acc += s_cnt[k];
    }
    s_pref[9] = acc;
  }
  // S(k) = H_a(k) + i H_b(k), H(k) = (Zbar(k) + conj Zbar(-k)) / 2 with
  // Zbar = 0 for k >= kz; stored conjugated (inverse through the forward
  // transform).
  const size_t fe = (size_t)f * p.E + e;
  for (int i = threadIdx.x; i < nl * p.kz; i += blockDim.x) {
    const int l = i / p.kz, k = i - l * p.kz;
    const int c0 = 2 * l;
    const int lx = c0 / p.ty, ly = c0 - lx * p.ty;
    float2 a = make_float2(0.0f, 0.0f), c = make_float2(0.0f, 0.0f);
    if (lx < wx && ly < wy) {
      const size_t row = (fe * p.n[0] + (X0 + lx)) * p.n[1] + (Y0 + ly);
      a = b.zbuf[row * p.kz + k];
      if (ly + 1 < wy) c = b.zbuf[(row + 1) * p.kz + k];
    }
    float2* line = buf + l * p.lsz;
    if (k == 0 || 2 * k == nnz) {
      line[k] = make_float2(a.x, -c.x);
    } else {
      // S(k) = (a + i c)/2 ; S(N-k) = (conj a + i conj c)/2 ; store conj
      line[k] = make_float2(0.5f * (a.x - c.y), -0.5f * (a.y + c.x));
      line[nnz - k] = make_float2(0.5f * (a.x + c.y), -0.5f * (c.x - a.y));
    }
  }
  __syncthreads();
  fft_lines<kMaxC, POW2>(buf, p.lsz, nl, p.fz, tw);
  double part = 0.0;
  for (int i = threadIdx.x; i < ncol * nz; i += blockDim.x) {
    const int c = i / nz, z = i - c * nz;
    const int l = c >> 1;
    const int lx = c / p.ty, ly = c - lx * p.ty;
    const float2 v = buf[l * p.lsz + z];
    const float val = (lx < wx && ly < wy) ? ((c & 1) ? -v.y : v.x) : 0.0f;
    drho[i] = val;
    part += (double)val;
  }
  // block sum of drho~ over this tile -> padt[f][e][tile]
  for (int off = 32; off > 0; off >>= 1) part += __shfl_down(part, off, 64);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = part;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s_red[w];
    g.padt[fe * (p.ntx * p.nty) + tile] = t;
  }

  // gather: every atom whose stencil reaches this tile accumulates
  // sum drho~ grad(theta_x theta_y theta_z) over its in-tile stencil points
  const int total = s_pref[9];
  const float sx = p.tric ? (float)p.n[0] : (float)p.n[0] * p.oc[0],
              sy = p.tric ? (float)p.n[1] : (float)p.n[1] * p.oc[1],
              sz = p.tric ? (float)p.n[2] : (float)p.n[2] * p.oc[2];
  for (int base = 0; base < total; base += kSpreadRecs) {
    const int chunk = min(kSpreadRecs, total - base);
    __syncthreads();
    for (int i = threadIdx.x; i < chunk; i += blockDim.x) {
      const int a = base + i;
      int sgm = 0;
#pragma unroll
      for (int k = 1; k < 9; ++k) sgm += (a >= s_pref[k]) ? 1 : 0;
      recs[i] = staged_record(b, s_start[sgm] + (a - s_pref[sgm]));
    }
    __syncthreads();
    for (int i = threadIdx.x; i < chunk; i += blockDim.x) {
      const float4 rec = recs[i];
      const int mx = cell_index(rec.x, p.n[0]);
      const int my = cell_index(rec.y, p.n[1]);
      const int mz = cell_index(rec.z, p.n[2]);
      const float wxf = rec.x - (float)mx, wyf = rec.y - (float)my, wzf = rec.z - (float)mz;
      float thx[P], thy[P], thz[P], dtx[P], dty[P], dtz[P];
      bspline<P>(wxf, thx);
      bspline<P>(wyf, thy);
      bspline<P>(wzf, thz);
      bspline_d<P>(wxf, dtx);
      bspline_d<P>(wyf, dty);
      bspline_d<P>(wzf, dtz);
      int gz0 = mz - P;
      if (gz0 < 0) gz0 += nz;
      float gx = 0.0f, gy = 0.0f, gzs = 0.0f;
      bool hit = false;
#pragma unroll
      for (int o = 0; o < P; ++o) {
        int gxi = mx - P + o;
        if (gxi < 0) gxi += p.n[0];
        const int lx = gxi - X0;
        if (lx < 0 || lx >= wx) continue;
#pragma unroll
        for (int o2 = 0; o2 < P; ++o2) {
          int gyi = my - P + o2;
          if (gyi < 0) gyi += p.n[1];
          const int ly = gyi - Y0;
          if (ly < 0 || ly >= wy) continue;
          hit = true;
          const float* colp = drho + (lx * p.ty + ly) * nz;
          float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
          for (int q = 0; q < P; ++q) {
            int gz = gz0 + q;
            if (gz >= nz) gz -= nz;
            const float d = colp[gz];
            s0 += d * thz[q];
            s1 += d * dtz[q];
          }
          gx += dtx[o] * thy[o2] * s0;
          gy += thx[o] * dty[o2] * s0;
          gzs += thx[o] * thy[o2] * s1;
        }
      }
      if (hit) {
        const int r = __float_as_int(rec.w) - f * p.N;   // position in the frame's run
        const int slot = partial_slot<P>(mx, my, bx, by, p, g);
        float* gp = g.gpart + (((size_t)f * g.nsx * g.nsy + slot) * p.N + r) * 3;
        gp[0] = gx * sx;
        gp[1] = gy * sy;
        gp[2] = gzs * sz;
      }
    }
  }
}

// ---- pruned, all-element variant of the z adjoint (pow2 N = 16 * R2) ------
//
// x(z) = Re sum_{k < kz} Zbar(k) e^{+2 pi i k z / N} for two columns at once
// as conj(FFT(conj S)) with S(k) = (Zbar_a(k) + conj Zbar_a(-k))/2 +
// i (Zbar_b(k) + conj Zbar_b(-k))/2: stage 0 is a radix-16 DFT over the
// strided sparse input k = j + R2 r (loaded straight from zbuf), stage 1 a
// radix-R2 DFT giving z = j' + 16 r'; both in registers, one LDS transpose.

// Line stride (float2) of the pruned adjoint's stage-0 lines, equal to the
// column stride (float) of drho~: float columns 2l and 2l + 1 occupy exactly
// the bytes of line l.  Room for 17 R2 line entries and for a column with
// its z halo.
__host__ __device__ inline int adj_line_stride(int nz, int P, int N) {
  const int a = 17 * (N / 16), c = nz + P - 1;
  return a > c ? a : c;
}

__host__ __device__ inline size_t adj_p_lds_bytes(const Params& p) {
  const int N = p.nn[2];
  const int LY = adj_line_stride(p.n[2], p.order, N);
  const size_t A = ((size_t)(p.tx * p.ty / 2) * LY * sizeof(float2) + 15) & ~size_t(15);
  const size_t B = ((size_t)kZpRecs * sizeof(float4) + 15) & ~size_t(15);
  const size_t tw = (size_t)N * sizeof(float2);
  const size_t seg = (size_t)p.E * 28 * sizeof(int);
  return A + B + tw + seg;
}

// ---- output-pruned y adjoint for any N = 2H R2 with n_y <= N/2 -----------
//
// The same map as k_adj_y, out(y) = conj(sum_m conj(Ybar(m)) W_N^{my}) over
// the kept m_y (m <= Ky or N - m <= Ky) for y < n_y, as the transpose of
// k_yfft_g's stages: v(m) = conj(Ybar(m)) of each line staged in LDS, then
// one thread per (line, butterfly j) forms Tbar_j[s] = sum_{m = s mod 2H}
// v(m) W_N^{jm} and the half-output DFT of length 2H by two length-H DFTs
// (as k_adj_zg).  Lines of a workgroup have consecutive kz (coalesced reads
// of ybuf and writes of zbuf); groups map to XCDs contiguously.  The plan
// guarantees 2 Ky + 1 < N.

__host__ __device__ inline size_t adj_y_g_lds_bytes(const Params& p, int R1) {
  const int N = p.nn[1];
  return (size_t)(kWG / (N / R1)) * p.ky * sizeof(float2) + (size_t)N * sizeof(float2);
}

template <int H>
__global__ void __launch_bounds__(kWG) k_adj_y_g(Params p, Buffers b) {
  constexpr int R1 = 2 * H;
  const int N = p.nn[1], R2 = N / R1, LW = kWG / R2;
  const int Ky = p.Ky, ky = p.ky;
  extern __shared__ __align__(16) unsigned char smem[];
  float2* vbuf = reinterpret_cast<float2*>(smem);          // [LW][ky] in ybuf's jy order
  float2* tw = vbuf + (size_t)LW * ky;                     // W_N^m
  const int nx = p.n[0], ny = p.n[1], kz = p.kz;
  const int nlines = nx * kz;
  const unsigned grp = xcd_contiguous(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const size_t fe = grp / gridDim.x;
  const int g0 = (int)(grp % gridDim.x) * LW;
  const float inv_lw = 1.0f / (float)LW, inv_kz = 1.0f / (float)kz;
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = b.twy[i];
  for (int item = threadIdx.x; item < LW * ky; item += blockDim.x) {
    const int jy = div_small(item, LW, inv_lw), l = item - jy * LW;
    const int g = g0 + l;
    float2 v = make_float2(0.0f, 0.0f);
    if (g < nlines) {
      const int x = div_small(g, kz, inv_kz), k = g - x * kz;
      v = b.ybuf[((fe * ky + jy) * nx + x) * kz + k];
    }
    vbuf[l * ky + jy] = make_float2(v.x, -v.y);
  }
  __syncthreads();
  if ((int)threadIdx.x >= LW * R2) return;
  const int j = div_small(threadIdx.x, LW, inv_lw), l = threadIdx.x - j * LW;
  const int g = g0 + l;
  if (g >= nlines) return;
  const float2* vl = vbuf + l * ky;                        // v(m) = vl[m], v(N - m) = vl[ky - m]
  const int stepR = (R1 * j) % N;
  float2 Y[R1];
  int base = 0;                                            // (s j) mod N
#pragma unroll
  for (int s = 0; s < R1; ++s) {
    float2 acc = make_float2(0.0f, 0.0f);
    int idx = base;
    for (int m = s; m <= Ky; m += R1) {                    // v(m) W_N^{jm}
      const float2 w = tw[idx], v = vl[m];
      acc.x += v.x * w.x - v.y * w.y;
      acc.y += v.x * w.y + v.y * w.x;
      idx += stepR;
      if (idx >= N) idx -= N;
    }
    int m = s == 0 ? R1 : R1 - s;                          // N - m = s mod 2H, m >= 1
    int idh = stepR - base;                                // (m j) mod N
    if (idh < 0) idh += N;
    for (; m <= Ky; m += R1) {                             // v(N - m) conj(W_N^{jm})
      const float2 w = tw[idh], v = vl[ky - m];
      acc.x += v.x * w.x + v.y * w.y;
      acc.y += v.y * w.x - v.x * w.y;
      idh += stepR;
      if (idh >= N) idh -= N;
    }
    Y[s] = acc;
    base += j;
    if (base >= N) base -= N;
  }
  float2 ev[H], od[H];
#pragma unroll
  for (int t = 0; t < H; ++t) {
    ev[t] = Y[2 * t];
    od[t] = Y[2 * t + 1];
  }
  Radix<H>::dft(ev);
  Radix<H>::dft(od);
  const int x = div_small(g, kz, inv_kz), k = g - x * kz;
  float2* dst = b.zbuf + ((fe * nx + x) * (size_t)ny) * kz + k;
#pragma unroll
  for (int r = 0; r < H; ++r) {
    const int y = j + R2 * r;
    if (y >= ny) continue;
    const float2 o = r == 0 ? cadd(ev[0], od[0]) : cadd(ev[r], cmul(od[r], tw[r * R2]));
    dst[(size_t)y * kz] = make_float2(o.x, -o.y);
  }
}

// ---- output-pruned x adjoint for any N = 2H R2 with n_x <= N/2 -----------
//
// k_adj_x over the same (jy, z chunk, frame) workgroups with k_xfft_g's
// transforms: A = sum_e f_e F_e at the kept pairs m, N - m (k_xfft_g's two
// stages per element, each thread holding its pairs), Abar = 2 mod w gH_b A
// (the histogram's weights), then per element Ybar_e(x) =
// conj(sum_m conj(f_e Abar(m)) W_N^{mx}) for x < n_x by the transposed
// stages of k_adj_y_g (the kept values of each z line staged in LDS),
// written over Y_e.  The plan guarantees what k_xfft_g needs.

__host__ __device__ inline size_t adj_x_g_lds_bytes(const Params& p, int R1) {
  const int N = p.nn[0];
  return (size_t)p.kcx * yg_line_stride(R1, N / R1) * sizeof(float2) + (size_t)N * sizeof(float2);
}

// LOAD: A read from where the forward k_xfft_g<H, true> saved it (its
// pairs by the same threads and items), not recomputed from Y.
template <int H, bool LOAD = false>
__global__ void __launch_bounds__(kWG) k_adj_x_g(Params p, Buffers b, AdjBuffers g,
                                                 const float2* amp = nullptr) {
  constexpr int R1 = 2 * H, YS = R1 + 1;
  extern __shared__ __align__(16) unsigned char smem[];
  const int jy = blockIdx.x, ch = blockIdx.y, f = blockIdx.z;
  const int kz = p.kz, k0 = ch * p.kcx, kc = min(p.kcx, kz - k0);
  const int nx = p.n[0], N = p.nn[0], nb = p.num_bins, Kx = p.Kx, kx = p.kx;
  const int R2 = N / R1, LY = yg_line_stride(R1, R2);
  float2* lines = reinterpret_cast<float2*>(smem);         // stage-0 lines, then kept values
  float2* tw = lines + (size_t)p.kcx * LY;                  // W_N^m
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = b.twx[i];
  const int iy = kept_to_index(jy, p.ky, p.Ky, p.nn[1]);
  const int my = iy < p.nf[1] ? iy : iy - p.nn[1];
  const float inv_kc = 1.0f / (float)kc;
  const int nitems = kc * (Kx + 1);                         // (pair m, line k), k fastest
  float2 A[kXgItems][2];
#pragma unroll
  for (int u = 0; u < kXgItems; ++u) A[u][0] = A[u][1] = make_float2(0.0f, 0.0f);
  __syncthreads();                                          // twiddles
  if constexpr (LOAD) {
    const int blk = (f * p.ky + jy) * p.nchx + ch;
#pragma unroll
    for (int u = 0; u < kXgItems; ++u)
      if (u * (int)blockDim.x < nitems) {                   // the same for the whole workgroup
        A[u][0] = amp[amp_g_index(blk, u, 0)];
        A[u][1] = amp[amp_g_index(blk, u, 1)];
      }
  }
  // A = sum_e f_e F_e at the kept pairs (as k_xfft_g)
  for (int e = 0; e < (LOAD ? 0 : p.E); ++e) {
    const float2* src = b.ybuf + ((((size_t)f * p.E + e) * p.ky + jy) * nx) * kz + k0;
    for (int item = threadIdx.x; item < kc * R2; item += blockDim.x) {
      const int j = div_small(item, kc, inv_kc), k = item - j * kc;
      float2 xe[H], xo[H];
#pragma unroll
      for (int r = 0; r < H; ++r) {
        const int x = j + R2 * r;
        xe[r] = x < nx ? src[(size_t)x * kz + k] : make_float2(0.0f, 0.0f);
      }
      xo[0] = xe[0];
#pragma unroll
      for (int r = 1; r < H; ++r) xo[r] = cmul(xe[r], tw[r * R2]);
      Radix<H>::dft(xe);
      Radix<H>::dft(xo);
      float2* dst = lines + k * LY + YS * j;
#pragma unroll
      for (int t = 0; t < H; ++t) {
        dst[2 * t] = xe[t];
        dst[2 * t + 1] = xo[t];
      }
    }
    __syncthreads();
    const float* ff = b.lut_ff + (size_t)e * p.lut_len;
#pragma unroll
    for (int u = 0; u < kXgItems; ++u) {
      const int item = threadIdx.x + u * blockDim.x;
      if (item < nitems) {
        const int m = div_small(item, kc, inv_kc), k = item - m * kc;
        const int s = m % R1, sm = s == 0 ? 0 : R1 - s;
        const float2* yl = lines + k * LY;
        float2 Z = make_float2(0.0f, 0.0f), Zm = make_float2(0.0f, 0.0f);
        int idx = 0;
        for (int j = 0; j < R2; ++j) {
          const float2 w = tw[idx];
          idx += m;
          if (idx >= N) idx -= N;
          const float2 y1 = yl[YS * j + s], y2 = yl[YS * j + sm];
          Z.x += y1.x * w.x - y1.y * w.y;
          Z.y += y1.x * w.y + y1.y * w.x;
          Zm.x += y2.x * w.x + y2.y * w.y;
          Zm.y += y2.y * w.x - y2.x * w.y;
        }
        const int mz = k0 + k;
        const int n0 = shell_key(m, my, mz, p);
        if (n0 < p.lut_len) {
          const float fv = ff[n0];
          A[u][0].x += fv * Z.x;
          A[u][0].y += fv * Z.y;
        }
        if (m > 0) {
          const int n1 = shell_key(-m, my, mz, p);
          if (n1 < p.lut_len) {
            const float fv = ff[n1];
            A[u][1].x += fv * Zm.x;
            A[u][1].y += fv * Zm.y;
          }
        }
      }
    }
    __syncthreads();
  }
  // Abar = 2 mod w gH_b A at in-range points, 0 elsewhere (k_adj_x)
  int keys[kXgItems][2];
#pragma unroll
  for (int u = 0; u < kXgItems; ++u) {
    const int item = threadIdx.x + u * blockDim.x;
#pragma unroll
    for (int o = 0; o < 2; ++o) {
      float2 ab = make_float2(0.0f, 0.0f);
      int nsq = 1 << 30;
      if (item < nitems) {
        const int m = div_small(item, kc, inv_kc), k = item - m * kc;
        if (o == 0 || m > 0) {
          const int ix = o == 0 ? m : N - m, mx = o == 0 ? m : -m, mz = k0 + k;
          nsq = shell_key(mx, my, mz, p);
          if (nsq < p.lut_len) {
            const int bin = b.lut_bin[nsq];
            if (bin >= 0 && bin < nb) {
              const float mod = b.modx[ix] * b.mody[iy] * b.modz[mz];
              const bool origin = (ix == 0) && (iy == 0) && (mz == 0);
              const float sc = 2.0f * mod * (origin ? 1.0f : 2.0f) * (float)g.gH[(size_t)f * nb + bin];
              ab = make_float2(sc * A[u][o].x, sc * A[u][o].y);
            }
          }
        }
      }
      A[u][o] = ab;
      keys[u][o] = nsq;
    }
  }
  // per element: kept v(m) = conj(f_e Abar(m)) of every z line into LDS
  // (line k, kept index jx as in ybuf's x order), then Ybar_e(x), x < n_x
  float2* vbuf = lines;                                     // [kc][kx]
  for (int e = 0; e < p.E; ++e) {
    for (int i = threadIdx.x; i < kc * kx; i += blockDim.x) vbuf[i] = make_float2(0.0f, 0.0f);
    __syncthreads();
    const float* ff = b.lut_ff + (size_t)e * p.lut_len;
#pragma unroll
    for (int u = 0; u < kXgItems; ++u) {
      const int item = threadIdx.x + u * blockDim.x;
      if (item < nitems) {
        const int m = div_small(item, kc, inv_kc), k = item - m * kc;
#pragma unroll
        for (int o = 0; o < 2; ++o) {
          if (keys[u][o] >= p.lut_len) continue;
          const float fv = ff[keys[u][o]];
          const int jx = o == 0 ? m : kx - m;
          vbuf[k * kx + jx] = make_float2(fv * A[u][o].x, -fv * A[u][o].y);   // conj
        }
      }
    }
    __syncthreads();
    float2* dst = b.ybuf + ((((size_t)f * p.E + e) * p.ky + jy) * nx) * kz + k0;
    for (int item = threadIdx.x; item < kc * R2; item += blockDim.x) {
      const int j = div_small(item, kc, inv_kc), k = item - j * kc;
      const float2* vl = vbuf + k * kx;                     // v(m) = vl[m], v(N - m) = vl[kx - m]
      const int stepR = (R1 * j) % N;
      float2 Y[R1];
      int base = 0;
#pragma unroll
      for (int s = 0; s < R1; ++s) {
        float2 acc = make_float2(0.0f, 0.0f);
        int idx = base;
        for (int m = s; m <= Kx; m += R1) {
          const float2 w = tw[idx], v = vl[m];
          acc.x += v.x * w.x - v.y * w.y;
          acc.y += v.x * w.y + v.y * w.x;
          idx += stepR;
          if (idx >= N) idx -= N;
        }
        int m = s == 0 ? R1 : R1 - s;
        int idh = stepR - base;
        if (idh < 0) idh += N;
        for (; m <= Kx; m += R1) {
          const float2 w = tw[idh], v = vl[kx - m];
          acc.x += v.x * w.x + v.y * w.y;
          acc.y += v.y * w.x - v.x * w.y;
          idh += stepR;
          if (idh >= N) idh -= N;
        }
        Y[s] = acc;
        base += j;
        if (base >= N) base -= N;
      }
      float2 ev[H], od[H];
#pragma unroll
      for (int t = 0; t < H; ++t) {
        ev[t] = Y[2 * t];
        od[t] = Y[2 * t + 1];
      }
      Radix<H>::dft(ev);
      Radix<H>::dft(od);
#pragma unroll
      for (int r = 0; r < H; ++r) {
        const int x = j + R2 * r;
        if (x >= nx) continue;
        const float2 o = r == 0 ? cadd(ev[0], od[0]) : cadd(ev[r], cmul(od[r], tw[r * R2]));
        dst[(size_t)x * kz + k] = make_float2(o.x, -o.y);
      }
    }
    __syncthreads();                                        // vbuf rewritten next element
  }
}

// ---- y adjoint for N = 2H x 16, most outputs kept: transpose of k_yfft_gf ----
//
// yf_adj_lines consecutive (x, kz) lines a workgroup, as k_yfft_gf's.
// Item (residue s < 2H, line l) reads v(M) = conj(Ybar(M)) at its kept
// outputs M = s + 2H r' (M <= Ky or M >= N - Ky), takes the radix-16 DFT
// over r' and the factor W_N^{js}: Tbar_j[s] = sum_{M = s mod 2H} v(M)
// W_N^{jM} (into LDS); butterfly (j, l) then forms
// xbar(j + 16 r) = sum_s Tbar_j[s] W_2H^{rs}, r < H, by two length-H DFTs of
// the even and odd s, and writes Zbar(y) = conj(xbar) for y < n_y.
__host__ __device__ inline size_t adj_y_gf_lds_bytes(int N) {
  return (size_t)yf_adj_lines(N / 16, 16) * yg_line_stride(N / 16, 16) * sizeof(float2) +
         (size_t)N * sizeof(float2);
}

template <int H>
__global__ void __launch_bounds__(kWG) k_adj_y_gf(Params p, Buffers b) {
  constexpr int R2 = 16, R1 = 2 * H, YS = R1 + 1, N = R1 * R2, LW = yf_adj_lines(R1, R2);
  constexpr int LY = (R2 * YS) | 1;                  // yg_line_stride(R1, R2)
  extern __shared__ __align__(16) unsigned char smem[];
  float2* tbuf = reinterpret_cast<float2*>(smem);
  float2* tw = tbuf + LW * LY;                       // W_N^m
  const int nx = p.n[0], ny = p.n[1], kz = p.kz;
  const int nlines = nx * kz;
  const unsigned grp = xcd_contiguous(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const size_t fe = grp / gridDim.x;
  const int g0 = (int)(grp % gridDim.x) * LW;
  const float inv_kz = 1.0f / (float)kz;
  const int Ky = p.Ky, ky = p.ky;
  const size_t row = (size_t)nx * kz;
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = b.twy[i];
  __syncthreads();                                   // twiddles
  // transposed stage 1 (item = residue s, line l; l fastest)
  for (int it = threadIdx.x; it < LW * R1; it += blockDim.x) {
    const int s = it / LW, l = it % LW;
    const int g = g0 + l;
    float2 t[R2];
#pragma unroll
    for (int rr = 0; rr < R2; ++rr) t[rr] = make_float2(0.0f, 0.0f);
    if (g < nlines) {
      const int x = div_small(g, kz, inv_kz), k = g - x * kz;
      const float2* in = b.ybuf + (fe * ky * nx + x) * kz + k;   // + jy nx kz
#pragma unroll
      for (int rr = 0; rr < R2; ++rr) {
        const int m = s + R1 * rr;
        int jy = -1;
        if (m <= Ky) jy = m;
        else if (m >= N - Ky) jy = ky - (N - m);
        if (jy >= 0) {
          const float2 v = in[(size_t)jy * row];
          t[rr] = make_float2(v.x, -v.y);
        }
      }
    }
    dft16<false>(t);
    float2* dst = tbuf + l * LY + s;
    dst[0] = t[0];
#pragma unroll
    for (int j = 1; j < R2; ++j) dst[YS * j] = cmul(t[j], tw[j * s]);   // j s < N
  }
  __syncthreads();
  // transposed stage 0 (item = butterfly j, line l; l fastest): LW R2 <= kWG items
  {
    const int j = threadIdx.x / LW, l = threadIdx.x % LW;
    const int g = g0 + l;
    if (j < R2 && g < nlines) {
      const float2* yl = tbuf + l * LY + YS * j;
      float2 ev[H], od[H];
#pragma unroll
      for (int t = 0; t < H; ++t) {
        ev[t] = yl[2 * t];
        od[t] = yl[2 * t + 1];
      }
      Radix<H>::dft(ev);
      Radix<H>::dft(od);
      const int x = div_small(g, kz, inv_kz), k = g - x * kz;
      float2* dst = b.zbuf + ((fe * nx + x) * (size_t)ny) * kz + k;
#pragma unroll
      for (int r = 0; r < H; ++r) {
        const int y = j + R2 * r;
        if (y >= ny) continue;
        const float2 o = r == 0 ? cadd(ev[0], od[0]) : cadd(ev[r], cmul(od[r], tw[r * R2]));
        dst[(size_t)y * kz] = make_float2(o.x, -o.y);
      }
    }
  }
}

// ---- x adjoint for N = 2H x 16, most outputs kept: transpose of k_xfft_gf ----
//
// Same workgroups, threads and stage-1 items (residue s, line k) as
// k_xfft_gf.  Phase 1: A at the item's kept outputs, read from where the
// forward k_xfft_gf<..., SAVE> put it (LOAD: the same values) or recomputed
// by its stages; Abar = 2 mod w gH_b A (k_adj_x).  Phase 2 per element: item
// (s, k) puts v(r') = conj(f_e Abar) at its kept outputs, takes the radix-16
// DFT over r' and the factor W_N^{js}, Tbar_j[s] = sum_{M = s mod 2H} v(M)
// W_N^{jM} (into LDS); butterfly (j, k) then forms
// xbar(j + R2 r) = sum_s Tbar_j[s] W_2H^{rs}, r < H, by two length-H DFTs of
// the even and odd s, and writes Ybar(x) = conj(xbar) for x < n_x over Y_e.
__host__ __device__ inline size_t adj_x_gf_lds_bytes(const Params& p, int R2) {
  const int N = p.nn[0], R1 = N / R2;
  return (size_t)p.kcx * yg_line_stride(R1, R2) * sizeof(float2) + (size_t)N * sizeof(float2);
}

template <int H, int R2, int KLC, bool LOAD>
__global__ void __launch_bounds__(kWG, 3) k_adj_x_gf(Params p, Buffers b, AdjBuffers g,
                                                     const float2* amp) {
  static_assert(R2 == 16, "radix-16 stage 1");
  static_assert(2 * KLC <= R2, "kept outputs at the ends");
  constexpr int R1 = 2 * H, YS = R1 + 1, N = R1 * R2, NA = 2 * KLC;
  constexpr int LY = (R2 * YS) | 1;                  // yg_line_stride(R1, R2)
  extern __shared__ __align__(16) unsigned char smem[];
  const int jy = blockIdx.x, ch = blockIdx.y, f = blockIdx.z;
  const int kz = p.kz, k0 = ch * p.kcx, kc = min(p.kcx, kz - k0);
  const int nx = p.n[0], nb = p.num_bins, Kx = p.Kx;
  float2* lines = reinterpret_cast<float2*>(smem);
  float2* tw = lines + (size_t)p.kcx * LY;           // W_N^m
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = b.twx[i];
  const int iy = kept_to_index(jy, p.ky, p.Ky, p.nn[1]);
  const int my = iy < p.nf[1] ? iy : iy - p.nn[1];
  const float inv_kc = 1.0f / (float)kc;
  const bool act1 = (int)threadIdx.x < kc * R1;
  const int s1 = act1 ? div_small(threadIdx.x, kc, inv_kc) : 0;
  const int k1 = act1 ? (int)threadIdx.x - s1 * kc : 0;
  const int mz = k0 + k1;
  auto key = [&](int rr) -> int {
    const int m = s1 + R1 * rr;
    return (m <= Kx || m >= N - Kx) ? shell_key(m <= Kx ? m : m - N, my, mz, p) : p.lut_len;
  };
  auto out_of = [](int o) { return o < KLC ? o : R2 - NA + o; };   // accumulator -> r'
  float2 A[NA];
#pragma unroll
  for (int o = 0; o < NA; ++o) A[o] = make_float2(0.0f, 0.0f);
  __syncthreads();                                   // twiddles
  if constexpr (LOAD) {
    const int blk = (f * p.ky + jy) * p.nchx + ch;
#pragma unroll
    for (int o = 0; o < NA; ++o) A[o] = amp[amp_gf_index(blk, o, NA)];
  } else {
    for (int e = 0; e < p.E; ++e) {                  // k_xfft_gf's stages
      for (int item = threadIdx.x; item < kc * R2; item += blockDim.x) {
        const int j0 = div_small(item, kc, inv_kc), kk0 = item - j0 * kc;
        const float2* src = b.ybuf + ((((size_t)f * p.E + e) * p.ky + jy) * nx) * kz + k0 + kk0;
        float2 xe[H], xo[H];
#pragma unroll
        for (int r = 0; r < H; ++r) {
          const int x = j0 + R2 * r;
          xe[r] = x < nx ? src[(size_t)x * kz] : make_float2(0.0f, 0.0f);
        }
        xo[0] = xe[0];
#pragma unroll
        for (int r = 1; r < H; ++r) xo[r] = cmul(xe[r], tw[r * R2]);
        Radix<H>::dft(xe);
        Radix<H>::dft(xo);
        float2* dst = lines + kk0 * LY + YS * j0;
#pragma unroll
        for (int t = 0; t < H; ++t) {
          dst[2 * t] = xe[t];
          dst[2 * t + 1] = xo[t];
        }
      }
      __syncthreads();
      if (act1) {
        const float* ff = b.lut_ff + (size_t)e * p.lut_len;
        const float2* yl = lines + k1 * LY + s1;
        float2 t[R2];
        t[0] = yl[0];
#pragma unroll
        for (int j = 1; j < R2; ++j) t[j] = cmul(yl[YS * j], tw[j * s1]);
        dft16<false>(t);
#pragma unroll
        for (int o = 0; o < NA; ++o) {
          const int rr = out_of(o);
          const int nsq = key(rr);
          if (nsq < p.lut_len) {
            const float fv = ff[nsq];
            A[o].x += fv * t[rr].x;
            A[o].y += fv * t[rr].y;
          }
        }
      }
      __syncthreads();
    }
  }
  // Abar = 2 mod w gH_b A at in-range points, 0 elsewhere
#pragma unroll
  for (int o = 0; o < NA; ++o) {
    const int rr = out_of(o);
    const int nsq = act1 ? key(rr) : p.lut_len;
    float2 ab = make_float2(0.0f, 0.0f);
    if (nsq < p.lut_len) {
      const int bin = b.lut_bin[nsq];
      if (bin >= 0 && bin < nb) {
        const int ix = s1 + R1 * rr;
        const float mod = b.modx[ix] * b.mody[iy] * b.modz[mz];
        const bool origin = (ix == 0) && (iy == 0) && (mz == 0);
        const float sc = 2.0f * mod * (origin ? 1.0f : 2.0f) * (float)g.gH[(size_t)f * nb + bin];
        ab = make_float2(sc * A[o].x, sc * A[o].y);
      }
    }
    A[o] = ab;
  }
  // phase 2: per element, Ybar_e over Y_e
  for (int e = 0; e < p.E; ++e) {
    if (act1) {
      const float* ff = b.lut_ff + (size_t)e * p.lut_len;
      float2 t[R2];
#pragma unroll
      for (int r = 0; r < R2; ++r) t[r] = make_float2(0.0f, 0.0f);
#pragma unroll
      for (int o = 0; o < NA; ++o) {
        const int rr = out_of(o);
        const int nsq = key(rr);
        if (nsq < p.lut_len) {
          const float fv = ff[nsq];
          t[rr] = make_float2(fv * A[o].x, -fv * A[o].y);   // conj(f Abar)
        }
      }
      dft16<false>(t);
      float2* dst = lines + k1 * LY + s1;
      dst[0] = t[0];
#pragma unroll
      for (int j = 1; j < R2; ++j) dst[YS * j] = cmul(t[j], tw[j * s1]);   // j s < N
    }
    __syncthreads();
    float2* ydst = b.ybuf + ((((size_t)f * p.E + e) * p.ky + jy) * nx) * kz + k0;
    for (int item = threadIdx.x; item < kc * R2; item += blockDim.x) {
      const int j = div_small(item, kc, inv_kc), k = item - j * kc;
      const float2* yl = lines + k * LY + YS * j;
      float2 ev[H], od[H];
#pragma unroll
      for (int t = 0; t < H; ++t) {
        ev[t] = yl[2 * t];
        od[t] = yl[2 * t + 1];
      }
      Radix<H>::dft(ev);
      Radix<H>::dft(od);
#pragma unroll
      for (int r = 0; r < H; ++r) {
        const int x = j + R2 * r;
        if (x >= nx) continue;
        const float2 o = r == 0 ? cadd(ev[0], od[0]) : cadd(ev[r], cmul(od[r], tw[r * R2]));
        ydst[(size_t)x * kz + k] = make_float2(o.x, -o.y);
      }
    }
    __syncthreads();                                 // lines rewritten next element
  }
}

// Adjoint of spread_cols: for each staged atom, the in-tile part of
// sum_stencil drho~ grad(th_x th_y th_z), with drho~ in columns of stride cs
// whose z halo repeats z < P - 1 after nz.  The stencil is unrolled and
// predicated on the in-tile ranges; the partial gradient goes to the
// record's slot for this tile (partial_slot), at the record's position in
// the frame's sorted run (staged_record).
template <int P>
__device__ __forceinline__ void gather_cols(const float4* recs, int count, const float* drho,
                                            const Params& p, int X0, int Y0, int wx, int wy,
                                            int cs, const AdjBuffers& g, int f, int bx, int by) {
  const int nz = p.n[2];
  const int plane = p.ty * cs;
  const int nslot = g.nsx * g.nsy;
  float* gframe = g.gpart + (size_t)f * nslot * p.N * 3;   // this frame's slots
  const float sx = p.tric ? (float)p.n[0] : (float)p.n[0] * p.oc[0],
              sy = p.tric ? (float)p.n[1] : (float)p.n[1] * p.oc[1],
              sz = p.tric ? (float)p.n[2] : (float)p.n[2] * p.oc[2];
  for (int i = threadIdx.x; i < count; i += blockDim.x) {
    const float4 rec = recs[i];
    const int mx = cell_index(rec.x, p.n[0]);
    const int my = cell_index(rec.y, p.n[1]);
    const int mz = cell_index(rec.z, p.n[2]);
    int xlo, xhi, dx, ylo, yhi, dy;
    stencil_range<P>(mx, p.n[0], X0, wx, xlo, xhi, dx);
    stencil_range<P>(my, p.n[1], Y0, wy, ylo, yhi, dy);
    if (xlo >= xhi || ylo >= yhi) continue;
    const float wxf = rec.x - (float)mx, wyf = rec.y - (float)my, wzf = rec.z - (float)mz;
    float thx[P], thy[P], thz[P], dtx[P], dty[P], dtz[P];
    bspline<P>(wxf, thx);
    bspline<P>(wyf, thy);
    bspline<P>(wzf, thz);
    bspline_d<P>(wxf, dtx);
    bspline_d<P>(wyf, dty);
    bspline_d<P>(wzf, dtz);
    int gz0 = mz - P;
    if (gz0 < 0) gz0 += nz;
    int yo[P];
#pragma unroll
    for (int o2 = 0; o2 < P; ++o2) {
      int ly = dy + o2;
      if (ly >= p.n[1]) ly -= p.n[1];
      yo[o2] = (int)__umul24((unsigned)ly, (unsigned)cs) + gz0;   // LDS offsets < 2^24
    }
    float gx = 0.0f, gy = 0.0f, gzs = 0.0f;
#pragma unroll
    for (int o = 0; o < P; ++o) {
      if (o < xlo || o >= xhi) continue;
      int lx = dx + o;
      if (lx >= p.n[0]) lx -= p.n[0];
      const float* xp = drho + __umul24((unsigned)lx, (unsigned)plane);
      float ax = 0.0f, ay = 0.0f, az = 0.0f;
#pragma unroll
      for (int o2 = 0; o2 < P; ++o2) {
        if (o2 < ylo || o2 >= yhi) continue;
        const float* c = xp + yo[o2];
        float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
        for (int q = 0; q < P; ++q) {
          const float d = c[q];
          s0 += d * thz[q];
          s1 += d * dtz[q];
        }
        ax += thy[o2] * s0;
        ay += dty[o2] * s0;
        az += thy[o2] * s1;
      }
      gx += dtx[o] * ax;
      gy += thx[o] * ay;
      gzs += thx[o] * az;
    }
    const int r = __float_as_int(rec.w) - f * p.N;     // position in the frame's run
    const int slot = partial_slot<P>(mx, my, bx, by, p, g);
    // (slot N + r) 3 < 2^32 for every plan (N < 2^24, nslot <= 9)
    float* gp = gframe + ((__umul24((unsigned)slot, (unsigned)p.N) + (unsigned)r) * 3u);
    gp[0] = gx * sx;
    gp[1] = gy * sy;
    gp[2] = gzs * sz;
  }
}

// ---- output-pruned z adjoint for any N = 2H R2 with n_z <= N/2 ----------
//
// The transpose of k_spread_zg's two stages.  The kept spectrum gradient of
// a line (column pair) is first formed as in k_adj_zp into v(k), nonzero for
// k < kz and N - k < kz (staged in LDS); then one thread per butterfly j < R2
// of the line computes
//   Ybar_j[s] = sum_{k = s mod 2H} v(k) W_N^{jk}       (one or two terms a s
//                                                     when kz <= 2H)
//   out(j + R2 r) = sum_s Ybar_j[s] W_2H^{rs},  r < H  (two DFTs of length H:
//                   E_r + W_2H^r O_r over the even and odd s)
// which is out(z) = sum_k v(k) W_N^{kz} for every z < N/2, and writes the
// drho~ columns (Re, -Im as k_adj_zp) with their z halo straight to LDS: no
// exchange between the stages.  The pad-term sum and the gather follow
// k_adj_zp.  The plan guarantees kz <= kAdjZgMaxKz.

constexpr int kAdjZgMaxKz = 64;

__host__ __device__ inline int adjg_col_stride(int nz, int P) { return (nz + P - 1 + 1) & ~1; }

__host__ __device__ inline size_t adj_g_lds_bytes(const Params& p) {
  const int cs = adjg_col_stride(p.n[2], p.order);
  const size_t A = ((size_t)p.tx * p.ty * cs * sizeof(float) + 15) & ~size_t(15);
  const size_t V = ((size_t)(p.tx * p.ty / 2) * 2 * p.kz * sizeof(float2) + 15) & ~size_t(15);
  const size_t B = ((size_t)kZpRecs * sizeof(float4) + 15) & ~size_t(15);
  const size_t tw = (size_t)p.nn[2] * sizeof(float2);
  const size_t seg = (size_t)p.E * 28 * sizeof(int);
  return A + V + B + tw + seg;
}

template <int P, int H>
__global__ void __launch_bounds__(kWG) k_adj_zg(Params p, Buffers b, AdjBuffers g) {
  constexpr int R1 = 2 * H;
  extern __shared__ __align__(16) unsigned char smem[];
  const int2 tf = xcd_tile_frame();
  const int tile = tf.x, f = tf.y;
  const int bx = tile % p.ntx, by = tile / p.ntx;
  const int X0 = bx * p.tx, Y0 = by * p.ty;
  const int wx = min(p.tx, p.n[0] - X0), wy = min(p.ty, p.n[1] - Y0);
  const int nz = p.n[2], N = p.nn[2], R2 = N / R1;
  const int ncol = p.tx * p.ty;
  const int nl = ncol / 2;
  const int E = p.E, kz = p.kz;
  const int cs = adjg_col_stride(nz, P);
  const size_t A_b = ((size_t)ncol * cs * sizeof(float) + 15) & ~size_t(15);
  const size_t V_b = ((size_t)nl * 2 * kz * sizeof(float2) + 15) & ~size_t(15);
  const size_t B_b = ((size_t)kZpRecs * sizeof(float4) + 15) & ~size_t(15);
  float* drho = reinterpret_cast<float*>(smem);
  float2* vbuf = reinterpret_cast<float2*>(smem + A_b);   // [l][v_lo(0..kz) | v_hi(0..kz)]
  float4* recs = reinterpret_cast<float4*>(smem + A_b + V_b);
  float2* tw = reinterpret_cast<float2*>(smem + A_b + V_b + B_b);   // W_N^m
  int* s_start = reinterpret_cast<int*>(tw + N);
  int* s_cnt = s_start + 9 * E;
  int* s_pref = s_cnt + 9 * E;
  __shared__ double s_red[kWG / 64];

  for (int t = threadIdx.x; t < 9 * E; t += blockDim.x) {
    const int e = t / 9;
    const int2 r = tile_slot(p, b, f, e, bx, by, t - e * 9);
    s_start[t] = r.x;
    s_cnt[t] = r.y;
  }
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = b.twz[i];
  __syncthreads();
  for (int e = threadIdx.x; e < E; e += blockDim.x) {
    int acc = 0;
    for (int k = 0; k < 9; ++k) {
      s_pref[e * 10 + k] = acc;
      acc += s_cnt[e * 9 + k];
    }
    s_pref[e * 10 + 9] = acc;
  }
  __syncthreads();
  auto fetch = [&](int e, int a) -> float4 {
    const int* pref = s_pref + e * 10;
    int sgm = 0;
#pragma unroll
    for (int k = 1; k < 9; ++k) sgm += (a >= pref[k]) ? 1 : 0;
    return staged_record(b, s_start[e * 9 + sgm] + (a - pref[sgm]));
  };
  const float inv_ty = 1.0f / (float)p.ty;
  const float inv_r2 = 1.0f / (float)R2;
  const float inv_kz = 1.0f / (float)kz;

  for (int e = 0; e < E; ++e) {
    const size_t fe = (size_t)f * E + e;
    const int total = s_pref[e * 10 + 9];
    const int staged = min(total, kZpRecs);
    float4 pre[kZpPre];
#pragma unroll
    for (int u = 0; u < kZpPre; ++u) {
      const int a = threadIdx.x + u * blockDim.x;
      if (a < staged) pre[u] = fetch(e, a);
    }
    // v(k) (k < kz) and v(N - k) (0 < k < kz) of every line, from zbuf
    for (int item = threadIdx.x; item < nl * kz; item += blockDim.x) {
      const int l = div_small(item, kz, inv_kz), k = item - l * kz;
      const int c0 = 2 * l;
      const int lx = div_small(c0, p.ty, inv_ty), ly = c0 - lx * p.ty;
      float2 lo = make_float2(0.0f, 0.0f), hi = make_float2(0.0f, 0.0f);
      if (lx < wx && ly < wy) {
        const float2* za = b.zbuf + ((fe * p.n[0] + (X0 + lx)) * p.n[1] + (Y0 + ly)) * kz;
        const float2 a = za[k];
        const float2 c = (ly + 1 < wy) ? za[kz + k] : make_float2(0.0f, 0.0f);
        lo = k == 0 ? make_float2(a.x, -c.x) : make_float2(0.5f * (a.x - c.y), -0.5f * (a.y + c.x));
        hi = make_float2(0.5f * (a.x + c.y), -0.5f * (c.x - a.y));
      }
      vbuf[l * 2 * kz + k] = lo;
      vbuf[l * 2 * kz + kz + k] = hi;
    }
    __syncthreads();                                  // v staged
    double part = 0.0;
    for (int item = threadIdx.x; item < nl * R2; item += blockDim.x) {
      const int l = div_small(item, R2, inv_r2), j = item - l * R2;
      const int c0 = 2 * l;
      const int lx = div_small(c0, p.ty, inv_ty), ly = c0 - lx * p.ty;
      if (lx >= wx || ly >= wy) continue;            // never gathered
      const float2* vlo = vbuf + l * 2 * kz;
      const float2* vhi = vlo + kz;
      const int stepR = (R1 * j) % N;                 // (2H j) mod N
      float2 Y[R1];
      int base = 0;                                   // (s j) mod N
#pragma unroll
      for (int s = 0; s < R1; ++s) {
        float2 acc = make_float2(0.0f, 0.0f);
        int idx = base;
        for (int k = s; k < kz; k += R1) {            // v(k) W_N^{jk}
          const float2 w = tw[idx], v = vlo[k];
          acc.x += v.x * w.x - v.y * w.y;
          acc.y += v.x * w.y + v.y * w.x;
          idx += stepR;
          if (idx >= N) idx -= N;
        }
        int m = s == 0 ? R1 : R1 - s;                 // N - m = s mod 2H, m >= 1
        int idh = stepR - base;                       // (m j) mod N
        if (idh < 0) idh += N;
        for (; m < kz; m += R1) {                     // v(N - m) conj(W_N^{jm})
          const float2 w = tw[idh], v = vhi[m];
          acc.x += v.x * w.x + v.y * w.y;
          acc.y += v.y * w.x - v.x * w.y;
          idh += stepR;
          if (idh >= N) idh -= N;
        }
        Y[s] = acc;
        base += j;
        if (base >= N) base -= N;
      }
      float2 ev[H], od[H];
#pragma unroll
      for (int t = 0; t < H; ++t) {
        ev[t] = Y[2 * t];
        od[t] = Y[2 * t + 1];
      }
      Radix<H>::dft(ev);
      Radix<H>::dft(od);
      const bool ok1 = ly + 1 < wy;
      float* col0 = drho + c0 * cs;
      float* col1 = col0 + cs;
#pragma unroll
      for (int r = 0; r < H; ++r) {
        const int z = j + R2 * r;
        if (z >= nz) continue;
        const float2 x = r == 0 ? cadd(ev[0], od[0]) : cadd(ev[r], cmul(od[r], tw[r * R2]));
        const float a = x.x, c = ok1 ? -x.y : 0.0f;
        col0[z] = a;
        col1[z] = c;
        if (r == 0 && z < P - 1) {                    // z halo
          col0[nz + z] = a;
          col1[nz + z] = c;
        }
        part += (double)a + (double)c;
      }
    }
    for (int off = 32; off > 0; off >>= 1) part += __shfl_down(part, off, 64);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = part;
#pragma unroll
    for (int u = 0; u < kZpPre; ++u) {
      const int a = threadIdx.x + u * blockDim.x;
      if (a < staged) recs[a] = pre[u];
    }
    __syncthreads();                                  // drho~ and records ready
    if (threadIdx.x == 0) {
      double sum = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) sum += s_red[w];
      g.padt[fe * (p.ntx * p.nty) + tile] = sum;
    }
    gather_cols<P>(recs, staged, drho, p, X0, Y0, wx, wy, cs, g, f, bx, by);
    for (int base = kZpRecs; base < total; base += kZpRecs) {
      const int chunk = min(kZpRecs, total - base);
      __syncthreads();
      for (int i = threadIdx.x; i < chunk; i += blockDim.x) recs[i] = fetch(e, base + i);
      __syncthreads();
      gather_cols<P>(recs, chunk, drho, p, X0, Y0, wx, wy, cs, g, f, bx, by);
    }
    __syncthreads();                                  // gather done: columns reused
  }
}

// Pruned z-pass adjoint + gradient gather, all elements of one column tile
// per workgroup.  Stage 0: the kept spectrum of a column pair (kz <= N / 8,
// so only inputs r = 0, 1 of the low and r = 14, 15 of the mirrored
// frequencies can be nonzero) unpacked into conj(S), radix-16 DFT, line l
// in LDS.  Stage 1: radix-R2 DFT; its outputs overwrite the same line as
// the drho~ columns 2l and 2l + 1 (line stride = column stride, the threads
// of a line share one wavefront), z halo repeated after nz.  Then the
// gather (gather_cols) into the atoms' partial-gradient slots.
// EXACT: n_z == N/2, every stage-1 output z = j + 16 rr, rr < R2/2, is in the column.
template <int P, int N, bool HALF, bool EXACT = false>
__global__ void __launch_bounds__(kWG, 3) k_adj_zp(Params p, Buffers b, AdjBuffers g) {
  static_assert(!EXACT || HALF, "EXACT implies HALF");
  constexpr int R2 = N / 16;
  constexpr int ROUT = HALF ? R2 / 2 : R2;             // stage-1 outputs that can be z < nz
  extern __shared__ __align__(16) unsigned char smem[];
  const int2 tf = xcd_tile_frame();
  const int tile = tf.x, f = tf.y;
  const int bx = tile % p.ntx, by = tile / p.ntx;
  const int X0 = bx * p.tx, Y0 = by * p.ty;
  const int wx = min(p.tx, p.n[0] - X0), wy = min(p.ty, p.n[1] - Y0);
  const int nz = p.n[2];
  const int ncol = p.tx * p.ty;
  const int nl = ncol / 2;
  const int E = p.E, kz = p.kz;
  const float inv_ty = 1.0f / (float)p.ty;
  const int LY = adj_line_stride(nz, P, N);            // line stride = drho~ column stride
  const size_t A_b = ((size_t)nl * LY * sizeof(float2) + 15) & ~size_t(15);
  const size_t B_b = ((size_t)kZpRecs * sizeof(float4) + 15) & ~size_t(15);
  float* drho = reinterpret_cast<float*>(smem);
  float2* ybuf = reinterpret_cast<float2*>(smem);
  float4* recs = reinterpret_cast<float4*>(smem + A_b);
  float2* tw = reinterpret_cast<float2*>(smem + A_b + B_b);
  int* s_start = reinterpret_cast<int*>(tw + N);
  int* s_cnt = s_start + 9 * E;
  int* s_pref = s_cnt + 9 * E;
  __shared__ double s_red[kWG / 64];

  for (int t = threadIdx.x; t < 9 * E; t += blockDim.x) {
    const int e = t / 9;
    const int2 r = tile_slot(p, b, f, e, bx, by, t - e * 9);
    s_start[t] = r.x;
    s_cnt[t] = r.y;
  }
  // stage-1 twiddles W_N^{j r} stored [r][j]
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = b.twz[(i >> 4) * (i & 15)];
  __syncthreads();
  for (int e = threadIdx.x; e < E; e += blockDim.x) {
    int acc = 0;
    for (int k = 0; k < 9; ++k) {
      s_pref[e * 10 + k] = acc;
      acc += s_cnt[e * 9 + k];
    }
    s_pref[e * 10 + 9] = acc;
  }
  __syncthreads();
  auto fetch = [&](int e, int a) -> float4 {
    const int* pref = s_pref + e * 10;
    int sgm = 0;
#pragma unroll
    for (int k = 1; k < 9; ++k) sgm += (a >= pref[k]) ? 1 : 0;
    return staged_record(b, s_start[e * 9 + sgm] + (a - pref[sgm]));
  };

  // Stage-0 inputs of thread (line l, butterfly j): the spectrum gradients
  // of the line's two columns at k = j + R2 r (r = 0, 1, k < kz) and
  // N - k = N - (j + R2 r) (r = 14, 15, < kz), loaded one element ahead
  // (N <= 256) so that the loads are in flight during the previous
  // element's stage 1 and gather: 1.80 -> 1.66 ms per 100 frames at N = 256.
  // At N = 512, whose stage 1 holds 32 complex values a thread, the extra
  // registers cost more than the hidden latency (5.83 -> 7.15 ms per 16
  // frames of the 1M-atom system), and the element's own inputs are loaded
  // at its stage 0.
  constexpr bool kAhead = N <= 256;
  const bool act0 = (int)threadIdx.x < nl * R2;
  const int l0 = act0 ? (int)threadIdx.x / R2 : 0, j0 = (int)threadIdx.x - l0 * R2;
  const int lx0 = (2 * l0) / p.ty, ly0 = 2 * l0 - lx0 * p.ty;
  const bool ok0 = act0 && lx0 < wx && ly0 < wy, ok1 = ok0 && ly0 + 1 < wy;
  float2 za_v[4], zb_v[4];
  auto zload = [&](int e) {
#pragma unroll
    for (int q = 0; q < 4; ++q) za_v[q] = zb_v[q] = make_float2(0.0f, 0.0f);
    if (ok0) {
      const float2* za =
          b.zbuf + ((((size_t)f * E + e) * p.n[0] + (X0 + lx0)) * p.n[1] + (Y0 + ly0)) * kz;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int k = q < 2 ? j0 + R2 * q : N - (j0 + R2 * (12 + q));
        if (k < kz) {
          za_v[q] = za[k];
          if (ok1) zb_v[q] = za[kz + k];
        }
      }
    }
  };
  if constexpr (kAhead) zload(0);

  for (int e = 0; e < E; ++e) {
    const size_t fe = (size_t)f * E + e;
    const int total = s_pref[e * 10 + 9];
    const int staged = min(total, kZpRecs);
    // records of this element: loads in flight during the transform
    float4 pre[kZpPre];
#pragma unroll
    for (int u = 0; u < kZpPre; ++u) {
      const int a = threadIdx.x + u * blockDim.x;
      if (a < staged) pre[u] = fetch(e, a);
    }
    // stage 0 (item = line l, butterfly j): sparse conj(S), DFT16
    if (threadIdx.x < nl * R2) {
      const int l = threadIdx.x / R2, j = threadIdx.x - l * R2;
      float2 v[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) v[r] = make_float2(0.0f, 0.0f);
      if constexpr (kAhead) {
        if (ok0) {
#pragma unroll
          for (int r = 0; r < 2; ++r) {               // k = j + R2 r < kz
            const int k = j + R2 * r;
            if (k < kz) {
              const float2 a = za_v[r], c = zb_v[r];
              v[r] = k == 0 ? make_float2(a.x, -c.x)
                            : make_float2(0.5f * (a.x - c.y), -0.5f * (a.y + c.x));
            }
          }
#pragma unroll
          for (int r = 14; r < 16; ++r) {             // N - k < kz
            const int m = N - (j + R2 * r);
            if (m < kz) {
              const float2 a = za_v[r - 12], c = zb_v[r - 12];
              v[r] = make_float2(0.5f * (a.x + c.y), -0.5f * (c.x - a.y));
            }
          }
        }
        if (e + 1 < E) zload(e + 1);                   // the next element's inputs
      } else {
        const int c0 = 2 * l;
        const int lx = c0 / p.ty, ly = c0 - lx * p.ty;
        const bool okA = lx < wx && ly < wy, okB = lx < wx && ly + 1 < wy;
        if (okA) {
          const float2* za = b.zbuf + ((fe * p.n[0] + (X0 + lx)) * p.n[1] + (Y0 + ly)) * kz;
          const float2* zb = za + kz;
#pragma unroll
          for (int r = 0; r < 2; ++r) {               // k = j + R2 r < kz
            const int k = j + R2 * r;
            if (k < kz) {
              const float2 a = za[k];
              const float2 c = okB ? zb[k] : make_float2(0.0f, 0.0f);
              v[r] = k == 0 ? make_float2(a.x, -c.x)
                            : make_float2(0.5f * (a.x - c.y), -0.5f * (a.y + c.x));
            }
          }
#pragma unroll
          for (int r = 14; r < 16; ++r) {             // N - k < kz
            const int m = N - (j + R2 * r);
            if (m < kz) {
              const float2 a = za[m];
              const float2 c = okB ? zb[m] : make_float2(0.0f, 0.0f);
              v[r] = make_float2(0.5f * (a.x + c.y), -0.5f * (c.x - a.y));
            }
          }
        }
      }
      dft16_in_ends(v);                                // nonzero inputs r = 0, 1, 14, 15
      float2* dst = ybuf + l * LY + 17 * j;
#pragma unroll
      for (int r = 0; r < 16; ++r) dst[r] = v[r];
    }
    // lines complete: with R2 = 16 thread t takes line t / 16 in both stages
    // (a wavefront's own lines; its LDS operations execute in order)
    if constexpr (R2 == 16) __builtin_amdgcn_wave_barrier();
    else __syncthreads();
    // stage 1 (item = line l, butterfly s): radix-R2 DFT; x = conj(out):
    // column 2l = Re, column 2l+1 = -Im, written over line l
    double part = 0.0;
    for (int item = threadIdx.x; item < nl * 16; item += blockDim.x) {
      const int l1 = item >> 4, j1 = item & 15;
      float2 t[R2];
#pragma unroll
      for (int r = 0; r < R2; ++r) {
        const float2 y = ybuf[l1 * LY + 17 * r + j1];
        t[r] = (r == 0) ? y : cmul(y, tw[r * 16 + j1]);
      }
      if constexpr (R2 == 8) Radix<8>::dft(t);
      else if constexpr (R2 == 16) {
        if constexpr (HALF) dft16_lo(t);                // outputs rr < 8 only
        else dft16<false>(t);
      } else dft32(t);
      const int c0 = 2 * l1;
      const int lx = div_small(c0, p.ty, inv_ty), ly = c0 - lx * p.ty;
      float* col0 = drho + c0 * LY;
      float* col1 = col0 + LY;
      __builtin_amdgcn_wave_barrier();
      // columns outside the cell are never gathered and add nothing
      if (lx < wx && ly < wy) {
        const bool ok1 = ly + 1 < wy;
#pragma unroll
        for (int rr = 0; rr < ROUT; ++rr) {
          const int z = j1 + 16 * rr;
          if (EXACT || z < nz) {
            const float a = t[rr].x, c = ok1 ? -t[rr].y : 0.0f;
            col0[z] = a;
            col1[z] = c;
            if (rr == 0 && z < P - 1) {               // z halo
              col0[nz + z] = a;
              col1[nz + z] = c;
            }
            part += (double)a + (double)c;
          }
        }
      }
    }
    for (int off = 32; off > 0; off >>= 1) part += __shfl_down(part, off, 64);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = part;
#pragma unroll
    for (int u = 0; u < kZpPre; ++u) {
      const int a = threadIdx.x + u * blockDim.x;
      if (a < staged) recs[a] = pre[u];
    }
    __syncthreads();                                  // drho~ and records ready
    if (threadIdx.x == 0) {
      double sum = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) sum += s_red[w];
      g.padt[fe * (p.ntx * p.nty) + tile] = sum;
    }
    gather_cols<P>(recs, staged, drho, p, X0, Y0, wx, wy, LY, g, f, bx, by);
    for (int base = kZpRecs; base < total; base += kZpRecs) {
      const int chunk = min(kZpRecs, total - base);
      __syncthreads();
      for (int i = threadIdx.x; i < chunk; i += blockDim.x) recs[i] = fetch(e, base + i);
      __syncthreads();
      gather_cols<P>(recs, chunk, drho, p, X0, Y0, wx, wy, LY, g, f, bx, by);
    }
    __syncthreads();                                  // gather done: lines reused
  }
}

// Per sorted record r of frame blockIdx.y (one per atom with a valid
// element): the sum of the partial gradients of the tiles its stencil
// reaches (slots recomputed from the record's cell coordinates, as the
// gather computes them; read at (f, slot, r), consecutive along r), minus
// the bulk-solvent mean term when the pad is the border mean, written to the
// record's atom.  The mean term is (padg_e / count) grad(border weight),
// border weight = (sum th_x)(sum th_y)(sum over border z of th_z); the full
// sums over x and y are 1 and their derivatives 0 (partition of unity), so
// only its z component is nonzero: (padg_e / count) (sum over border z of
// dth_z).  Thread i also writes zeros for atom i when its element id is
// invalid (such atoms have no record).
template <int P>
__global__ void __launch_bounds__(kWG) k_grad_finalize(Params p, Buffers b, AdjBuffers g) {
  __shared__ int s_eb[kMaxElements + 1];             // element runs of the frame, from f N
  // a frame's blocks on one XCD: the scattered 12-byte gradient writes fill
  // whole lines in that XCD's L2 before they reach HBM
  const unsigned lg = xcd_contiguous(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const int f = (int)(lg / gridDim.x);
  const int blk = (int)(lg - (unsigned)f * gridDim.x);
  const int K = frame_keys(p), T = K / p.E;   // keys per frame, per element
  const int f0 = f * p.N;
  for (int e = threadIdx.x; e <= p.E; e += blockDim.x)
    s_eb[e] = e < p.E ? b.offsets[f * K + e * T] - f0
                      : b.offsets[f * K + K - 1] + b.counts[f * K + K - 1] - f0;
  __syncthreads();
  const int i = blk * blockDim.x + threadIdx.x;
  if (i >= p.N) return;
  {
    const int ei = b.elem[i];
    if (ei < 0 || ei >= p.E) {
      float* gz = g.grad + ((size_t)f0 + i) * 3;
      gz[0] = 0.0f;
      gz[1] = 0.0f;
      gz[2] = 0.0f;
    }
  }
  if (i >= s_eb[p.E]) return;
  const float4 rec = b.sorted[(size_t)f0 + i];
  int e = 0;
  for (int k = 1; k < p.E; ++k) e += (i >= s_eb[k]) ? 1 : 0;
  const float rr[3] = {rec.x, rec.y, rec.z};
  float w[3];
  int base[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int m = cell_index(rr[a], p.n[a]);
    w[a] = rr[a] - (float)m;
    base[a] = wrap_base(m, P, p.n[a]);
  }
  // tile offsets (from the home tile) that the stencil reaches along x, y
  const float inv_tx = 1.0f / (float)p.tx, inv_ty = 1.0f / (float)p.ty;
  unsigned vx = 0u, vy = 0u;
  const int hx = div_small(base[0], p.tx, inv_tx), hy = div_small(base[1], p.ty, inv_ty);
#pragma unroll
  for (int o = 0; o < P; ++o) {
    int cx = base[0] + o, cy = base[1] + o;
    if (cx >= p.n[0]) cx -= p.n[0];
    if (cy >= p.n[1]) cy -= p.n[1];
    int ox = div_small(cx, p.tx, inv_tx) - hx, oy = div_small(cy, p.ty, inv_ty) - hy;
    if (ox < 0) ox += p.ntx;
    if (oy < 0) oy += p.nty;
    vx |= 1u << ox;
    vy |= 1u << oy;
  }
  const int nslot = g.nsx * g.nsy;
  const float* part = g.gpart + ((size_t)f * nslot * p.N + i) * 3;
  float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
  for (int oy = 0; oy < g.nsy; ++oy) {
    if (!((vy >> oy) & 1u)) continue;
    for (int ox = 0; ox < g.nsx; ++ox) {
      if (!((vx >> ox) & 1u)) continue;
      const float* q = part + (size_t)(ox + g.nsx * oy) * p.N * 3;
      s0 += q[0];
      s1 += q[1];
      s2 += q[2];
    }
  }
  float t2 = s2;
  if (!b.pad_given) {                                // a given pad is a constant: no mean term
    float dth[P];
    bspline_d<P>(w[2], dth);
    double dSz = 0.0;
#pragma unroll
    for (int o = 0; o < P; ++o) {
      int z = base[2] + o;
      if (z >= p.n[2]) z -= p.n[2];
      if (z <= p.border || z >= p.n[2] - p.border) dSz += (double)dth[o];
    }
    if (dSz != 0.0) {
      const double c = g.padg[(size_t)f * p.E + e] * p.inv_count;
      t2 = s2 - (float)(c * dSz * (double)p.n[2] * (p.tric ? 1.0 : (double)p.oc[2]));
    }
  }
  const int idx = __float_as_int(rec.w);
  float* gp = g.grad + ((size_t)f0 + idx) * 3;
  if (!p.tric) {
    gp[0] = s0;
    gp[1] = s1;
    gp[2] = t2;
    return;
  }
  // triclinic: the sums are dL/dr_a n_a; dL/dx_j = sum_a (dL/dr_a n_a) M_aj
#pragma unroll
  for (int j = 0; j < 3; ++j) gp[j] = s0 * p.cm[j] + s1 * p.cm[3 + j] + t2 * p.cm[6 + j];
}

}  // namespace saxs
