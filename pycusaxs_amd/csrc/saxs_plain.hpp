// Reference-structured per-frame path on this engine's kernels: the stage
// order of saxsKernel::runPKernel (saxsKernel.cu:59-209) -- per element:
// zero the cell grid, spread with global atomics (rhoCartKernel), border
// mean (paddingKernel), embed into the FFT grid (superDensityKernel; here
// with the mean subtracted inside the cell and zeros outside, identical for
// every q != 0), full 3-D R2C FFT (cufftExecR2C), form-factor accumulate at
// q <= qcut (scatterKernel); then |A|^2 * moduli over the whole half
// spectrum (modulusKernel) and the histogram (calculate_histogram).
// One frame per call; it is the per-frame baseline of the batched path.
#pragma once
#include "saxs_kernels.hpp"

namespace saxs {

struct PlainBuffers {
  float* grid;     // [n_x][n_y][n_z]
  float* sup;      // [nn_x][nn_y][nn_z]
  float2* zspec;   // [nn_x][nn_y][h]
  float2* yspec;   // [nn_y][nn_x][h]
  float2* acc;     // [nn_x][nn_y][h]
  double* pad;     // [1]
};

__global__ void k_plain_fill(float* a, size_t n, float2* c, size_t nc, double* d) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0f;
  if (i < nc) c[i] = make_float2(0.0f, 0.0f);
  if (i == 0 && d) d[0] = 0.0;
}

template <int P>
__global__ void __launch_bounds__(kWG) k_plain_spread(Params p, const float* xyz, const int* elem,
                                                      int e, float* grid) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.N || elem[i] != e) return;
  int m[3];
  float th[3][P], rr[3];
  cell_coords(xyz + (size_t)i * 3, p, rr);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float r = rr[a];
    m[a] = cell_index(r, p.n[a]);
    bspline<P>(r - (float)m[a], th[a]);
  }
#pragma unroll
  for (int o = 0; o < P; ++o) {
    int gx = m[0] - P + o;
    if (gx < 0) gx += p.n[0];
#pragma unroll
    for (int o2 = 0; o2 < P; ++o2) {
      int gy = m[1] - P + o2;
      if (gy < 0) gy += p.n[1];
      const float fxy = th[0][o] * th[1][o2];
#pragma unroll
      for (int q = 0; q < P; ++q) {
        int gz = m[2] - P + q;
        if (gz < 0) gz += p.n[2];
        atomicAdd(grid + ((size_t)gx * p.n[1] + gy) * p.n[2] + gz, fxy * th[2][q]);
      }
    }
  }
}

// Sum over the border z-slabs (z <= d or z >= n_z - d).
__global__ void __launch_bounds__(kWG) k_plain_pad(Params p, const float* grid, double* pad) {
  __shared__ double red[kWG / 64];
  const int nb = 2 * p.border + 1;
  const size_t total = (size_t)p.n[0] * p.n[1] * nb;
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t xy = i / nb;
    const int t = (int)(i - xy * nb);
    const int z = t <= p.border ? t : p.n[2] - (nb - t);
    s += (double)grid[xy * p.n[2] + z];
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    atomicAdd(pad, t);
  }
}

// given >= 0: padding::given constant instead of the border mean.
__global__ void k_plain_embed(Params p, const float* grid, float* sup, const double* pad,
                              double given) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)p.nn[0] * p.nn[1] * p.nn[2];
  if (i >= total) return;
  const int z = (int)(i % p.nn[2]);
  const size_t xy = i / p.nn[2];
  const int y = (int)(xy % p.nn[1]);
  const int x = (int)(xy / p.nn[1]);
  float v = 0.0f;
  if (x < p.n[0] && y < p.n[1] && z < p.n[2]) {
    const double m = given >= 0.0 ? given : pad[0] * p.inv_count;
    const float hi = (float)m, lo = (float)(m - (double)hi);
    v = (grid[((size_t)x * p.n[1] + y) * p.n[2] + z] - hi) - lo;
  }
  sup[i] = v;
}

// R2C along z over all nn_x * nn_y lines of sup; line pairs (y, y+1).
template <int POW2>
__global__ void __launch_bounds__(kWG) k_plain_zfft(Params p, Buffers b, const float* sup, float2* zspec,
                                                    int pairs_per_wg) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nnz = p.nn[2], h = nnz / 2 + 1;
  float2* buf = reinterpret_cast<float2*>(smem);
  float2* tw = buf + (size_t)pairs_per_wg * p.lsz;
  const size_t npairs = (size_t)p.nn[0] * (p.nn[1] / 2);
  const size_t pair0 = (size_t)blockIdx.x * pairs_per_wg;
  const int np = (int)min((size_t)pairs_per_wg, npairs - pair0);
  for (int i = threadIdx.x; i < nnz; i += blockDim.x) tw[i] = b.twz[i];
  auto load = [&](unsigned l, unsigned z) -> float2 {
    if ((int)l >= np) return make_float2(0.0f, 0.0f);
    const size_t pr = pair0 + l;
    const size_t x = pr / (p.nn[1] / 2), y = 2 * (pr % (p.nn[1] / 2));
    const float* line = sup + (x * p.nn[1] + y) * nnz;
    return make_float2(line[z], line[nnz + z]);
  };
  __syncthreads();
  fft_lines_loaded<kMaxC, POW2>(buf, p.lsz, pairs_per_wg, p.fz, tw, load);
  for (int i = threadIdx.x; i < np * h; i += blockDim.x) {
    const int l = i / h, k = i - l * h;
    const size_t pr = pair0 + l;
    const size_t x = pr / (p.nn[1] / 2), y = 2 * (pr % (p.nn[1] / 2));
    const float2 Z = buf[l * p.lsz + k];
    const float2 Zm = buf[l * p.lsz + (k == 0 ? 0 : nnz - k)];
    const size_t row = x * p.nn[1] + y;
    zspec[row * h + k] = make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
    zspec[(row + 1) * h + k] = make_float2(0.5f * (Z.y + Zm.y), -0.5f * (Z.x - Zm.x));
  }
}

// FFT along x, then acc += f_e(q) F at q <= qcut (scatterKernel).  Input
// yspec layout [nn_y][nn_x][h] (from k_yfft on the full grid).
template <int POW2>
__global__ void __launch_bounds__(kWG) k_plain_xfft_scatter(Params p, Buffers b, const float2* yspec,
                                                            float2* acc, int e) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int iy = blockIdx.x, ch = blockIdx.y;
  const int h = p.kz;
  const int k0 = ch * p.kcx;
  const int kc = min(p.kcx, h - k0);
  const int nnx = p.nn[0];
  float2* buf = reinterpret_cast<float2*>(smem);
  float2* tw = buf + (size_t)p.kcx * p.lsx;
  for (int i = threadIdx.x; i < nnx; i += blockDim.x) tw[i] = b.twx[i];
  const float2* src = yspec + (size_t)iy * nnx * h;
  for (int i = threadIdx.x; i < nnx * kc; i += blockDim.x) {
    const int x = i / kc, k = i - x * kc;
    buf[k * p.lsx + x] = src[(size_t)x * h + k0 + k];
  }
  __syncthreads();
  fft_lines<kMaxC, POW2>(buf, p.lsx, kc, p.fx, tw);
  const int my = iy < p.nf[1] ? iy : iy - p.nn[1];
  for (int i = threadIdx.x; i < nnx * kc; i += blockDim.x) {
    const int ix = i / kc, k = i - ix * kc;
    const int kz = k0 + k;
    const int mx = ix < p.nf[0] ? ix : ix - nnx;
    const int mz = kz < p.nf[2] ? kz : kz - p.nn[2];
    const int nsq = shell_key(mx, my, mz, p);
    if (nsq < p.lut_len) {
      const float fv = b.lut_ff[(size_t)e * p.lut_len + nsq];
      const float2 F = buf[k * p.lsx + ix];
      float2* a = acc + ((size_t)ix * p.nn[1] + iy) * h + kz;
      const float2 old = *a;
      *a = make_float2(old.x + fv * F.x, old.y + fv * F.y);
    }
  }
}

// modulusKernel: acc = |acc|^2 * modX modY modZ over the whole half spectrum.
__global__ void k_plain_modulus(Params p, Buffers b, float2* acc) {
  const int h = p.nn[2] / 2 + 1;
  const size_t total = (size_t)p.nn[0] * p.nn[1] * h;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int k = (int)(i % h);
  const size_t xy = i / h;
  const int y = (int)(xy % p.nn[1]);
  const int x = (int)(xy / p.nn[1]);
  const float2 a = acc[i];
  const float mod = b.modx[x] * b.mody[y] * b.modz[k];
  acc[i] = make_float2((a.x * a.x + a.y * a.y) * mod, 0.0f);
}

// calculate_histogram: each in-range half-spectrum point adds I[i,j,k] +
// I[-i,-j,k] (count 2), the origin I once (count 1); per-block bins in LDS.
// Triclinic cells add 2 I[i,j,k] at the bin of (i,j,k) itself.
__global__ void __launch_bounds__(kWG) k_plain_hist(Params p, Buffers b, const float2* I, double* H, double* C) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nb = p.num_bins;
  double* Hs = reinterpret_cast<double*>(smem);
  double* Cs = Hs + nb;
  for (int i = threadIdx.x; i < 2 * nb; i += blockDim.x) Hs[i] = 0.0;
  __syncthreads();
  const int h = p.nn[2] / 2 + 1;
  const size_t total = (size_t)p.nn[0] * p.nn[1] * h;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(i % h);
    const size_t xy = i / h;
    const int y = (int)(xy % p.nn[1]);
    const int x = (int)(xy / p.nn[1]);
    const int mx = x < p.nf[0] ? x : x - p.nn[0];
    const int my = y < p.nf[1] ? y : y - p.nn[1];
    const int mz = k < p.nf[2] ? k : k - p.nn[2];
    const int nsq = shell_key(mx, my, mz, p);
    if (nsq >= p.lut_len) continue;
    const int bin = b.lut_bin[nsq];
    if (bin < 0 || bin >= nb) continue;
    if (x == 0 && y == 0 && k == 0) {
      atomicAdd(&Hs[bin], (double)I[i].x);
      atomicAdd(&Cs[bin], 1.0);
    } else if (p.point_keys == 2) {
      // triclinic: (-i,-j,k) is not at the same |q|; the point stands for
      // itself and its Hermitian mirror -m
      atomicAdd(&Hs[bin], 2.0 * (double)I[i].x);
      atomicAdd(&Cs[bin], 2.0);
    } else {
      const int xb = x == 0 ? 0 : p.nn[0] - x, yb = y == 0 ? 0 : p.nn[1] - y;
      const size_t j = ((size_t)xb * p.nn[1] + yb) * h + k;
      atomicAdd(&Hs[bin], (double)I[i].x + (double)I[j].x);
      atomicAdd(&Cs[bin], 2.0);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    if (Hs[i] != 0.0) atomicAdd(&H[i], Hs[i]);
    if (Cs[i] != 0.0) atomicAdd(&C[i], Cs[i]);
  }
}

}  // namespace saxs
