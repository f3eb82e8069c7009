// Device kernels of the forward scattering path (gfx950).
//
// Pipeline for a batch of F frames of one system (E element types):
//   k_clear         zero bin counters and pad sums
//   k_prep<P>       per atom: grid cell and fraction (saxsDeviceKernels.cu:352-368),
//                   column-tile key, slot in its bin, border-slab weight sum
//   k_seg_keys      per key: segment starts within its run and its total
//   k_scan          exclusive scan of the per-(frame, element, tile) counts
//   k_scatter<P>    counting-sort the atoms by (frame, element, tile)
//   k_spread_z<P>   per (frame, element, column tile): gather-spread the atoms
//                   into an LDS column block (saxsDeviceKernels.cu:338-398),
//                   subtract the bulk-solvent border mean (Q5 of the reference's
//                   pad/embed, :413-465), R2C along z, keep kz <= Kz
//   k_yfft          FFT along y, keep |m_y| <= K
//   k_xfft_epi      FFT along x, keep |m_x| <= K, then per point:
//                   A = sum_e f_e(q) F_e (:177-216), I = |A|^2 modX modY modZ
//                   (:142-162), weighted histogram into LDS (:67-125)
//   k_hist_reduce   deterministic reduction of the per-block histograms
//
// Layouts (complex64, x slowest):
//   zbuf [F][E][n_x][n_y][kz]     ybuf [F][E][ky][n_x][kz]
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_lds.hpp"

namespace saxs {

constexpr int kWG = 256;       // threads per workgroup (4 wave64)
constexpr int kMaxC = 16;      // complex values per thread in one FFT stage
constexpr int kOutMax = 8;     // x-pass output points per thread
constexpr int kMaxElements = 64;

struct Params {
  int n[3], nn[3];
  int order, E, border;
  int tx, ty, ntx, nty;
  int ncx = 1, ncy = 1;  // sort-key classes per axis (1, or 2: stencil leaves its home tile)
  int kz, ky, Ky, kx, Kx;
  int nf[3];
  int num_bins, lut_len;
  int F, N;
  float oc[3];
  int kcy, nchy, kcx, nchx;
  int lsz, lsy, lsx;     // LDS line strides (complex) of the three passes
  int point_keys;        // q-table keys: 0 shells |m|^2, 1 points |m_a|, 2 signed points
  int kmax[3];           // point-key box |m_a| <= kmax[a] (point_keys only)
  int tric;              // triclinic cell: u = cm x (row-major 3x3), else u_a = oc_a x_a
  float cm[9];
  double inv_count;      // 1 / number of border voxels
  FftDesc fz, fy, fx;
};

struct Buffers {
  const float* xyz;
  const int* elem;
  int* counts;           // [F * frame_keys + 1]
  int* offsets;          // [F * frame_keys + 1]
  int2* keyslot;         // [F*N]
  float4* sorted;        // [F*N]  (r_x, r_y, r_z, atom index bits)
  unsigned long long* pad_sum;   // [F*E] border-slab weight, fixed point (to_fixed_d)
  const unsigned long long* pad_given;  // [E] given pad in units of 2^-32, or NULL
  const int4* srcx;      // [ntx] {runs (1, 2), first x unit of run 0, of run 1, len0 | len1 << 16}
  const int4* srcy;      // [nty] {count (1..3), y unit 0, 1, 2}
  float2* zbuf;
  float2* ybuf;
  const float2* twz;
  const float2* twy;
  const float2* twx;
  const float* modx;
  const float* mody;
  const float* modz;
  const int* lut_bin;
  const float* lut_ff;   // [E][lut_len]
  double* partial;       // [F][ky][nchx][2*num_bins]
  int* segoff;           // [F][nseg][keys per frame] segmented sort (nseg > 0)
  const float* cells;    // [F][9] inverse cell of each frame (row-major), or NULL: the plan's
  int lut_fstride;       // 0: one q table for all frames; else frame f's tables start at
                         // lut_bin + f lut_fstride and lut_ff + f E lut_fstride
  int nseg, seglen;      // segments per frame and atoms per segment
  int segkr;             // keys per k_prep_seg pass
};

// Index of frequency m in the q tables (bin and form factor): the shell
// n = |m|^2 when the three reciprocal spacings are equal, else the point
// (|m_x|, |m_y|, |m_z|) of the box |m_a| <= kmax[a] (orthorhombic cells, where
// q^2 = sum_a (m_a s_a)^2 is not a function of |m|^2); lut_len when outside.
// Triclinic cells (q = sum_a m_a b_a is not even in each m_a) key the signed
// point, ((m_x + K_x)(2K_y + 1) + m_y + K_y)(2K_z + 1) + m_z + K_z.
__device__ __forceinline__ int shell_key(int mx, int my, int mz, const Params& p) {
  if (!p.point_keys) return mx * mx + my * my + mz * mz;
  const int ax = abs(mx), ay = abs(my), az = abs(mz);
  if (ax > p.kmax[0] || ay > p.kmax[1] || az > p.kmax[2]) return p.lut_len;
  if (p.point_keys == 1) return (ax * (p.kmax[1] + 1) + ay) * (p.kmax[2] + 1) + az;
  return ((mx + p.kmax[0]) * (2 * p.kmax[1] + 1) + my + p.kmax[1]) * (2 * p.kmax[2] + 1) + mz +
         p.kmax[2];
}

// Workgroups are dispatched round-robin over the chip's 8 XCDs, each with its
// own L2: linear block id `lin` runs on XCD lin % 8.  This bijection maps the
// blocks of XCD c onto the contiguous range of logical ids
// [c q + min(c, r), ...) (q = total / 8, r = total % 8), so logical
// neighbours share an L2.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_contiguous(unsigned lin, unsigned total) {
  const unsigned q = total / kXcds, r = total % kXcds;
  const unsigned c = lin % kXcds, i = lin / kXcds;
  return c * q + min(c, r) + i;
}

// (column tile, frame) of a workgroup of a (tiles, frames) grid, a frame's
// tiles kept on one XCD: neighbouring tiles fetch the same atom records
// (each record reaches about 2.4 tiles), which then come from that XCD's L2.
__device__ __forceinline__ int2 xcd_tile_frame() {
  const unsigned lg = xcd_contiguous(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const unsigned f = lg / gridDim.x;
  return make_int2((int)(lg - f * gridDim.x), (int)f);
}

// Bulk-solvent pad of (frame, element) fe in units of 2^-32 atoms per voxel:
// the border-slab mean (padding::avg, saxsDeviceKernels.cu:446-465) or the
// caller's constant (padding::given, Options.h:7-11).
__device__ __forceinline__ unsigned long long pad_units(const Buffers& b, const Params& p,
                                                        size_t fe, int e) {
  return b.pad_given ? b.pad_given[e]
                     : (unsigned long long)llrint((double)b.pad_sum[fe] * p.inv_count);
}

// ----------------------------------------------------------- atom helpers --

// Cell coordinate r_a = float(n_a (u - rint(u - 0.5))) of the fractional
// coordinate u, the float/double mix of saxsDeviceKernels.cu:355-361.
__device__ __forceinline__ float cell_coord_u(float u, int n) {
  const double ud = (double)u;
  const double r = (double)n * (ud - rint(ud - 0.5));
  return (float)r;
}

// Orthorhombic cell: u = float(oc_a x_a).
__device__ __forceinline__ float cell_coord(float x, float oc, int n) {
  return cell_coord_u(__fmul_rn(oc, x), n);
}

// All three cell coordinates of position x: orthorhombic, or triclinic with
// u_a = (M_a0 x + M_a1 y) + M_a2 z, each product and sum rounded to float
// (saxsDeviceKernels.cu:355-357 with the true inverse M of the cell).
__device__ __forceinline__ void cell_coords(const float* x, const Params& p, float r[3]) {
  if (!p.tric) {
#pragma unroll
    for (int a = 0; a < 3; ++a) r[a] = cell_coord(x[a], p.oc[a], p.n[a]);
    return;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float u = __fadd_rn(__fadd_rn(__fmul_rn(p.cm[3 * a], x[0]), __fmul_rn(p.cm[3 * a + 1], x[1])),
                              __fmul_rn(p.cm[3 * a + 2], x[2]));
    r[a] = cell_coord_u(u, p.n[a]);
  }
}

// Cell coordinates of position x in frame f: the frame's own inverse cell
// when the batch carries one per frame (u_a = (M_a0 x + M_a1 y) + M_a2 z,
// which equals M_aa x_a bit for bit when M is diagonal), else the plan's.
__device__ __forceinline__ void frame_cell_coords(const float* x, const Params& p,
                                                  const Buffers& b, int f, float r[3]) {
  if (!b.cells) {
    cell_coords(x, p, r);
    return;
  }
  const float* m = b.cells + (size_t)f * 9;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float u = __fadd_rn(__fadd_rn(__fmul_rn(m[3 * a], x[0]), __fmul_rn(m[3 * a + 1], x[1])),
                              __fmul_rn(m[3 * a + 2], x[2]));
    r[a] = cell_coord_u(u, p.n[a]);
  }
}

// Cardinal B-spline weights th[k] = M_P(w + P - 1 - k) by the recursion of
// Splines.cu:47-80 (Init then OnePass for k = 3..P).
// Cubic weights in closed form, times `scale`: with u = 1 - w,
//   th0 = u^3/6, th1 = (4 - 6w^2 + 3w^3)/6, th2 = (1 + 3w + 3w^2 - 3w^3)/6,
//   th3 = w^3/6
// (the order-4 recursion below expanded), in 12 operations with the scale
// folded into the constants.
__device__ __forceinline__ void bspline4(float w, float scale, float* th) {
  const float u = 1.0f - w, w2 = w * w, hw = w * (0.5f * scale);
  th[3] = w2 * (w * (scale / 6.0f));
  th[0] = (u * u) * (u * (scale / 6.0f));
  th[1] = fmaf(w2, fmaf(w, 0.5f * scale, -scale), scale * (2.0f / 3.0f));
  th[2] = fmaf(hw, fmaf(w, u, 1.0f), scale / 6.0f);
}

template <int P>
__device__ __forceinline__ void bspline(float w, float* th) {
  if constexpr (P == 4) {
    bspline4(w, 1.0f, th);
    return;
  }
  th[0] = 1.0f - w;
  th[1] = w;
#pragma unroll
  for (int i = 2; i < P; ++i) th[i] = 0.0f;
#pragma unroll
  for (int k = 3; k <= P; ++k) {
    const float div = 1.0f / (float)(k - 1);
    th[k - 1] = div * w * th[k - 2];
#pragma unroll
    for (int j = 1; j <= k - 2; ++j)
      th[k - j - 1] = div * ((w + (float)j) * th[k - j - 2] + ((float)(k - j) - w) * th[k - j - 1]);
    th[0] = div * (1.0f - w) * th[0];
  }
}

// dth/dw: order P-1 weights differenced (Splines.cu:84-92); for order 4 in
// closed form, dth0 = -u^2/2, dth1 = w (3w/2 - 2), dth2 = 1/2 + w - 3w^2/2,
// dth3 = w^2/2 with u = 1 - w (the quadratic weights differenced).
template <int P>
__device__ __forceinline__ void bspline_d(float w, float* dth) {
  if constexpr (P == 4) {
    const float u = 1.0f - w, w2 = w * w;
    dth[0] = -0.5f * (u * u);
    dth[1] = w * fmaf(w, 1.5f, -2.0f);
    dth[2] = fmaf(w2, -1.5f, w + 0.5f);
    dth[3] = 0.5f * w2;
    return;
  }
  float t[P];
  bspline<P - 1>(w, t);
  dth[0] = -t[0];
#pragma unroll
  for (int o = 1; o < P - 1; ++o) dth[o] = t[o - 1] - t[o];
  dth[P - 1] = t[P - 2];
}

// Cell index m = int(r), clamped to [0, n] so that non-finite coordinates
// cannot index outside the grid (finite inputs always give 0 <= r <= n).
__device__ __forceinline__ int cell_index(float r, int n) {
  return min(max((int)r, 0), n);
}

__device__ __forceinline__ int wrap_base(int m, int P, int n) {
  int b = m - P;
  return b < 0 ? b + n : b;
}

// Fixed-point accumulation: 2^32 units per unit weight in an unsigned
// 64-bit integer, used for the spread columns and the border-slab sums.
// LDS float atomics run at a small fraction of the integer rate on gfx950;
// integer sums are also exact and independent of the order of the atomics
// (and of the column tiling).  A spread contribution is a B-spline product
// v = th_x th_y th_z <= (2/3)^3 < 1 taken as (unsigned)(v * 2^32): it fits
// 32 bits, is an exact integer for v >= 2^-9 (24-bit mantissa) and is
// truncated below (error < 2^-32, under a tenth of a float32 ulp of a
// typical column value).  A sum overflows only past 2^32 units of weight.
constexpr double kFixInv = 1.0 / 4294967296.0;     // 2^-32

// The same scale for a weight in [0, 1] held in double (border-slab shares).
__device__ __forceinline__ unsigned long long to_fixed_d(double v) {
  return (unsigned long long)(v * 4294967296.0);
}

// q = a / d for 0 <= a < 2^22, 1 <= d < 2^22 via a float reciprocal and one
// correction each way (exact; cheaper than integer division by a runtime d).
__device__ __forceinline__ int div_small(int a, int d, float inv_d) {
  int q = (int)((float)a * inv_d);
  const int qd = (int)__umul24((unsigned)q, (unsigned)d);   // q d <= a + d < 2^23
  if (qd > a) --q;
  else if (qd + d <= a) ++q;
  return q;
}

// Column value minus pad, both in fixed-point units (2^-32 of a unit
// weight), as a float in the same units: the difference is exact in 64 bits
// and is rounded once, through double.
__device__ __forceinline__ float fixed_minus_units(unsigned long long v, unsigned long long pad) {
  const unsigned long long d = v - pad;
  return (float)((double)(int)(unsigned)(d >> 32) * 4294967296.0 + (double)(unsigned)d);
}

// ------------------------------------------------------------------ kernels --

__global__ void k_clear(int* counts, int nkeys, unsigned long long* pad_sum, int npad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nkeys) counts[i] = 0;
  if (i < npad) pad_sum[i] = 0ull;
}

// Sort keys of one frame.  A key is (element, y unit, x unit) with the unit
// of an axis S C + c: S the tile holding the stencil's first point (its home
// tile), c (with C = 2 classes) 1 when the stencil leaves that tile, else 0.
// A tile then fetches only atoms that reach it: its home units and the
// class-1 unit of the tile before it (C = 2), or every atom of each source
// tile (C = 1, axes whose stencils can reach three tiles).
__host__ __device__ inline int frame_keys(const Params& p) {
  return p.E * p.nty * p.ncy * p.ntx * p.ncx;
}

// Unit of stencil base `base` along an axis of n cells in tiles of t.
__device__ __forceinline__ int key_unit(int base, int t, int n, int C, int P) {
  const int S = base / t;
  if (C == 1) return S;
  const int X0 = S * t, w = min(t, n - X0);
  return 2 * S + (base - X0 + P - 1 >= w ? 1 : 0);
}

// Record range {first sorted position, count} of fetch slot sg < 9 of column
// tile (bx, by), element e, frame f: slot (i, j) = (x-unit run i < 2, y unit
// j < 3) as sg = i + 3 j.  A run of consecutive keys is one range of the
// sorted records (their offsets are an exclusive scan within the frame).
__device__ __forceinline__ int2 tile_slot(const Params& p, const Buffers& b, int f, int e, int bx,
                                          int by, int sg) {
  const int i = sg % 3, j = sg / 3;
  const int4 sx = b.srcx[bx], sy = b.srcy[by];
  if (i >= sx.x || j >= sy.x) return make_int2(0, 0);
  const int a = i == 0 ? sx.y : sx.z;
  const int len = i == 0 ? (sx.w & 0xffff) : (sx.w >> 16);
  const int yu = j == 0 ? sy.y : (j == 1 ? sy.z : sy.w);
  const int k0 = ((f * p.E + e) * (p.nty * p.ncy) + yu) * (p.ntx * p.ncx) + a;
  const int k1 = k0 + len - 1;
  const int st = b.offsets[k0];
  return make_int2(st, b.offsets[k1] + b.counts[k1] - st);
}

// Per atom of frame f: its sort key within the frame (frame_keys) and its
// border-slab share of unit weight in
// fixed point: the sum over the stencil of th_x th_y th_z restricted to
// z <= d or z >= n_z - d.
template <int P>
__device__ __forceinline__ void atom_key_border(const Params& p, const Buffers& b, int f, int i,
                                                int e, int& key, unsigned long long& border) {
  const float* x = b.xyz + ((size_t)f * p.N + i) * 3;
  int base[3];
  float w[3], rr[3];
  frame_cell_coords(x, p, b, f, rr);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float r = rr[a];
    const int m = cell_index(r, p.n[a]);
    w[a] = r - (float)m;
    base[a] = wrap_base(m, P, p.n[a]);
  }
  key = (e * (p.nty * p.ncy) + key_unit(base[1], p.ty, p.n[1], p.ncy, P)) * (p.ntx * p.ncx) +
        key_unit(base[0], p.tx, p.n[0], p.ncx, P);
  float th[P];
  double sx = 0.0, sy = 0.0, sz = 0.0;
  bspline<P>(w[0], th);
#pragma unroll
  for (int o = 0; o < P; ++o) sx += (double)th[o];
  bspline<P>(w[1], th);
#pragma unroll
  for (int o = 0; o < P; ++o) sy += (double)th[o];
  bspline<P>(w[2], th);
#pragma unroll
  for (int o = 0; o < P; ++o) {
    int z = base[2] + o;
    if (z >= p.n[2]) z -= p.n[2];
    if (z <= p.border || z >= p.n[2] - p.border) sz += (double)th[o];
  }
  border = to_fixed_d(sx * sy * sz);
}

// Counting sort with one global counter per key (any number of keys): the
// slot of an atom in its (frame, key) bin is an atomic increment.
template <int P>
__global__ void __launch_bounds__(kWG) k_prep(Params p, Buffers b) {
  __shared__ unsigned long long padacc[kMaxElements];
  const int f = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  for (int e = threadIdx.x; e < p.E; e += blockDim.x) padacc[e] = 0ull;
  __syncthreads();
  const int e = i < p.N ? b.elem[i] : -1;
  if (i < p.N && (e < 0 || e >= p.E)) b.keyslot[(size_t)f * p.N + i] = make_int2(-1, -1);
  if (i < p.N && e >= 0 && e < p.E) {
    int key;
    unsigned long long border;
    atom_key_border<P>(p, b, f, i, e, key, border);
    key += f * frame_keys(p);
    const int slot = atomicAdd(&b.counts[key], 1);
    b.keyslot[(size_t)f * p.N + i] = make_int2(key, slot);
    if (border != 0ull) atomicAdd(&padacc[e], border);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < p.E; e += blockDim.x) {
    const unsigned long long v = padacc[e];
    if (v != 0ull) atomicAdd(&b.pad_sum[f * p.E + e], v);
  }
}

// ---- segmented counting sort (one frame's keys fit in LDS) ---------------
//
// k_prep_seg: workgroup (segment s, frame f) takes atoms [s L, (s+1) L) and
// counts them per key in an LDS histogram (the slot within (key, segment)
// is the LDS atomic's old value); the histogram goes to segoff[f][s][k].
// k_seg_keys turns those counts into each (key, segment) run's start within
// the key and writes the per-key counts; k_scan then gives the per-key
// offsets the spread kernels read.  No global atomics on the counters.
constexpr int kSortWG = 1024;

// LDS of k_prep_seg for K keys a pass: the key histogram (32-bit counters,
// or with HW 16-bit counters two to a word: a (key, segment) count never
// exceeds the segment length, which the plan then keeps below 65536) and the
// per-element border sums.
__host__ __device__ inline size_t prep_seg_lds_bytes(int K, int E, bool hw = false) {
  const size_t hist = hw ? (size_t)((K + 1) / 2) * sizeof(unsigned) : (size_t)K * sizeof(int);
  return ((hist + 15) & ~size_t(15)) + (size_t)E * sizeof(unsigned long long);
}

constexpr int kSeg16MaxLen = 65535;   // atoms per segment with 16-bit counters

template <int P, bool HW>
__global__ void __launch_bounds__(kSortWG) k_prep_seg(Params p, Buffers b) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int K = frame_keys(p);
  const int KR = HW ? b.segkr : K;                   // keys per pass (LDS): 32-bit, one pass
  const int KRW = HW ? (KR + 1) / 2 : KR;
  unsigned* hist = reinterpret_cast<unsigned*>(smem);
  unsigned long long* padacc =
      reinterpret_cast<unsigned long long*>(smem + (((size_t)KRW * sizeof(unsigned) + 15) & ~size_t(15)));
  const int s = blockIdx.x, f = blockIdx.y;
  for (int e = threadIdx.x; e < p.E; e += blockDim.x) padacc[e] = 0ull;
  const int i0 = s * b.seglen, i1 = min(p.N, i0 + b.seglen);
  int* out = b.segoff + ((size_t)f * b.nseg + s) * K;
  // keys in passes of KR (one pass when a frame's keys fit the LDS); the
  // segment's atoms are read once per pass, the border sums added in the first
  for (int kb = 0; kb < K; kb += KR) {
    const int kn = min(KR, K - kb);
    const int knw = HW ? (kn + 1) / 2 : kn;
    for (int k = threadIdx.x; k < knw; k += blockDim.x) hist[k] = 0u;
    __syncthreads();
    for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) {
      const int e = b.elem[i];
      if (e < 0 || e >= p.E) {
        if (kb == 0) b.keyslot[(size_t)f * p.N + i] = make_int2(-1, -1);
        continue;
      }
      int key;
      unsigned long long border;
      atom_key_border<P>(p, b, f, i, e, key, border);
      const int kl = key - kb;
      if (kl >= 0 && kl < kn) {
        int slot;
        if (HW) {
          const unsigned sh = (unsigned)(kl & 1) * 16u;
          slot = (int)((atomicAdd(&hist[kl >> 1], 1u << sh) >> sh) & 0xffffu);
        } else {
          slot = (int)atomicAdd(&hist[kl], 1u);
        }
        b.keyslot[(size_t)f * p.N + i] = make_int2(f * K + key, slot);
      }
      // border-slab share: nonzero for the few atoms near the slab only (the
      // E accumulators are contended)
      if (kb == 0 && border != 0ull) atomicAdd(&padacc[e], border);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kn; k += blockDim.x)
      out[kb + k] = HW ? (int)((hist[k >> 1] >> ((k & 1) * 16)) & 0xffffu) : (int)hist[k];
    __syncthreads();                                 // hist is cleared for the next pass
  }
  for (int e = threadIdx.x; e < p.E; e += blockDim.x) {
    const unsigned long long v = padacc[e];
    if (v != 0ull) atomicAdd(&b.pad_sum[f * p.E + e], v);
  }
}

// Per key of frame blockIdx.y: the exclusive prefix of its per-segment
// counts (in place: segoff then holds each segment's start within the key's
// run) and the key's total.  One thread per key, loads and stores coalesced
// along the keys.
__global__ void __launch_bounds__(kWG) k_seg_keys(Params p, Buffers b) {
  const int K = frame_keys(p);
  const int k = blockIdx.x * blockDim.x + threadIdx.x, f = blockIdx.y;
  if (k >= K) return;
  int* seg = b.segoff + (size_t)f * b.nseg * K + k;
  int run = 0;
  for (int q = 0; q < b.nseg; ++q) {
    const int c = seg[(size_t)q * K];
    seg[(size_t)q * K] = run;
    run += c;
  }
  b.counts[(size_t)f * K + k] = run;
}

// Exclusive scan of the key counts of frame blockIdx.x from f N (the
// frame's first slot of the sorted array): one workgroup of 1024 threads,
// keys in tiles of 1024, one key per thread (coalesced); a round's loads are
// all issued before its tiles are scanned (wavefront shuffles, then the
// sixteen wavefront totals through LDS), with a running carry.
constexpr int kScanTiles = 16;

__global__ void __launch_bounds__(1024) k_scan(const int* __restrict__ counts, int* offsets, int K,
                                               int N) {
  __shared__ int wsum[16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const size_t off0 = (size_t)blockIdx.x * K;
  int carry = (int)blockIdx.x * N;
  for (int base = 0; base < K; base += 1024 * kScanTiles) {
    int v[kScanTiles];
#pragma unroll
    for (int i = 0; i < kScanTiles; ++i) {
      const int k = base + i * 1024 + t;
      v[i] = k < K ? counts[off0 + k] : 0;
    }
#pragma unroll
    for (int i = 0; i < kScanTiles; ++i) {
      if (base + i * 1024 >= K) break;               // uniform
      int x = v[i];
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
      }
      if (lane == 63) wsum[w] = x;
      __syncthreads();
      int before = 0, total = 0;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int sq = wsum[q];
        before += q < w ? sq : 0;
        total += sq;
      }
      const int k = base + i * 1024 + t;
      if (k < K) offsets[off0 + k] = carry + before + x - v[i];
      carry += total;
      __syncthreads();                               // wsum is rewritten
    }
  }
}

template <int P>
__global__ void __launch_bounds__(kWG) k_scatter(Params p, Buffers b) {
  // a frame's blocks on one XCD: its scattered 16-byte record writes then
  // fill whole lines in that XCD's L2 before they reach HBM
  const unsigned lg = xcd_contiguous(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const int f = (int)(lg / gridDim.x);
  const int i = (int)(lg - (unsigned)f * gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= p.N) return;
  const float* x = b.xyz + ((size_t)f * p.N + i) * 3;
  const int2 ks = b.keyslot[(size_t)f * p.N + i];
  if (ks.x < 0) return;
  float4 rec;
  float rr[3];
  frame_cell_coords(x, p, b, f, rr);
  rec.x = rr[0];
  rec.y = rr[1];
  rec.z = rr[2];
  rec.w = __int_as_float(i);
  const int K = frame_keys(p);
  int pos = b.offsets[ks.x] + ks.y;
  if (b.nseg > 0) pos += b.segoff[((size_t)f * b.nseg + i / b.seglen) * K + (ks.x - f * K)];
  b.sorted[pos] = rec;
}

// ---- all elements of a column tile in one workgroup ----------------------

// In-tile stencil range of one axis: stencil points base + o (mod n), o in
// [0, P), that fall in [T0, T0 + w); returns [lo, hi) in o and the tile
// offset of o = 0.
// A tile that spans the whole axis (w == n) takes every o, with local
// coordinate (base + o) mod n; otherwise the plan guarantees w + P - 1 <= n,
// so the in-tile points form one run of o.
template <int P>
__device__ __forceinline__ void stencil_range(int m, int n, int T0, int w, int& lo, int& hi,
                                              int& d) {
  int base = m - P;
  if (base < 0) base += n;
  if (w == n) {
    d = base;
    lo = 0;
    hi = P;
    return;
  }
  d = base - T0;                      // tile-local coordinate of o = 0
  if (d > n / 2) d -= n;
  if (d < -n / 2) d += n;
  lo = max(0, -d);
  hi = min(P, w - d);
}

// LDS column stride of the pruned spread: room for the z stencil to run
// P - 1 points past nz (folded back onto z < P - 1 when the column is read),
// rounded to 8 (mod 16) so that the two complex lines a half-wave reads in
// the first transform stage sit 16 banks apart.
template <int P>
__host__ __device__ constexpr int spread_col_stride(int nz) {
  return ((nz + P - 1 + 7) / 16) * 16 + 8;
}

// Adds the in-tile stencil points of `count` staged atoms (B-spline weights
// th_x th_y th_z of the atom's cell fraction, saxsDeviceKernels.cu:370-398)
// to LDS columns of stride cs with a z halo (no z wrap), in fixed point
// (kFixInv): the stencil is unrolled and predicated on the in-tile ranges,
// plane and column offsets are computed once per record, and LPR lanes share
// a record, lane s taking the x planes o = s, s + LPR, ...
template <int P, int LPR>
__device__ __forceinline__ void spread_cols(const float4* recs, int count,
                                            unsigned long long* rho, const Params& p, int X0,
                                            int Y0, int wx, int wy, int cs) {
  const int nz = p.n[2];
  const int plane = p.ty * cs;
  const int sub = threadIdx.x % LPR;
  for (int i = threadIdx.x / LPR; i < count; i += blockDim.x / LPR) {
    const float4 rec = recs[i];
    const int mx = cell_index(rec.x, p.n[0]);
    const int my = cell_index(rec.y, p.n[1]);
    const int mz = cell_index(rec.z, p.n[2]);
    int xlo, xhi, dx, ylo, yhi, dy;
    stencil_range<P>(mx, p.n[0], X0, wx, xlo, xhi, dx);
    stencil_range<P>(my, p.n[1], Y0, wy, ylo, yhi, dy);
    if (xlo >= xhi || ylo >= yhi) continue;
    float thx[P], thy[P], thz[P];
    bspline<P>(rec.x - (float)mx, thx);
    bspline<P>(rec.y - (float)my, thy);
    if constexpr (P == 4) {
      bspline4(rec.z - (float)mz, 4294967296.0f, thz);   // 2^32 (exact): to_fixed's scale
    } else {
      bspline<P>(rec.z - (float)mz, thz);
#pragma unroll
      for (int q = 0; q < P; ++q) thz[q] *= 4294967296.0f;
    }
    int gz0 = mz - P;
    if (gz0 < 0) gz0 += nz;
    int yo[P];
#pragma unroll
    for (int o2 = 0; o2 < P; ++o2) {
      int ly = dy + o2;
      if (ly >= p.n[1]) ly -= p.n[1];
      yo[o2] = (int)__umul24((unsigned)ly, (unsigned)cs) + gz0;   // LDS offsets < 2^24
    }
#pragma unroll
    for (int oo = 0; oo < P; oo += LPR) {
      const int o = oo + sub;
      if (o >= P || o < xlo || o >= xhi) continue;
      int lx = dx + o;
      if (lx >= p.n[0]) lx -= p.n[0];
      unsigned long long* xp = rho + __umul24((unsigned)lx, (unsigned)plane);
      const float fx = thx[o];
#pragma unroll
      for (int o2 = 0; o2 < P; ++o2) {
        if (o2 < ylo || o2 >= yhi) continue;
        unsigned long long* c = xp + yo[o2];
        const float f = fx * thy[o2];
#pragma unroll
        for (int q = 0; q < P; ++q) atomicAdd(c + q, (unsigned long long)(unsigned)(f * thz[q]));
      }
    }
  }
}

// Atom records staged per round in k_spread_z.
constexpr int kSpreadRecs = 512;

// LDS bytes of k_spread_z: density columns, FFT lines (the staged atom
// records alias the FFT lines; they are dead before the FFT starts) and
// twiddles.
__host__ __device__ inline size_t spread_lds_bytes(const Params& p) {
  const int cs = p.order == 6 ? spread_col_stride<6>(p.n[2]) : spread_col_stride<4>(p.n[2]);
  const size_t rho = (((size_t)p.tx * p.ty * cs * sizeof(unsigned long long)) + 15) & ~size_t(15);
  size_t buf = (size_t)(p.tx * p.ty / 2) * p.lsz * sizeof(float2);
  const size_t recs = (size_t)kSpreadRecs * sizeof(float4);
  if (buf < recs) buf = recs;
  const size_t tw = (size_t)p.nn[2] * sizeof(float2);
  return rho + buf + tw;
}

template <int P, int POW2>
__global__ void __launch_bounds__(kWG) k_spread_z(Params p, Buffers b) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int tile = blockIdx.x, e = blockIdx.y, f = blockIdx.z;
  const int bx = tile % p.ntx, by = tile / p.ntx;
  const int X0 = bx * p.tx, Y0 = by * p.ty;
  const int wx = min(p.tx, p.n[0] - X0), wy = min(p.ty, p.n[1] - Y0);
  const int nz = p.n[2], nnz = p.nn[2];
  const int ncol = p.tx * p.ty;
  const int nl = ncol / 2;
  const int cs = spread_col_stride<P>(nz);
  unsigned long long* rho = reinterpret_cast<unsigned long long*>(smem);
  float2* buf = reinterpret_cast<float2*>(
      smem + ((ncol * cs * sizeof(unsigned long long) + 15) & ~size_t(15)));
  float4* recs = reinterpret_cast<float4*>(buf);
  const size_t buf_bytes = (size_t)nl * p.lsz * sizeof(float2);
  float2* tw = reinterpret_cast<float2*>(reinterpret_cast<unsigned char*>(buf) +
                                          (buf_bytes > kSpreadRecs * sizeof(float4)
                                               ? buf_bytes : kSpreadRecs * sizeof(float4)));

  // source segments: up to 3x3 bins whose stencils reach this tile
  __shared__ int s_start[9], s_cnt[9], s_pref[10];
  if (threadIdx.x < 9) {
    const int2 r = tile_slot(p, b, f, e, bx, by, threadIdx.x);
    s_start[threadIdx.x] = r.x;
    s_cnt[threadIdx.x] = r.y;
  }
  for (int i = threadIdx.x; i < ncol * cs; i += blockDim.x) rho[i] = 0ull;
  for (int i = threadIdx.x; i < nnz; i += blockDim.x) tw[i] = b.twz[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int k = 0; k < 9; ++k) {
      s_pref[k] = acc;
      acc += s_cnt[k];
    }
    s_pref[9] = acc;
  }
  __syncthreads();
  const int total = s_pref[9];

  // Gather-spread in rounds of kSpreadRecs atoms: stage the records in LDS
  // with all threads, then one thread per atom computes its B-spline weights
  // once and adds its stencil columns that fall inside this tile.
  for (int base = 0; base < total; base += kSpreadRecs) {
    const int chunk = min(kSpreadRecs, total - base);
    for (int i = threadIdx.x; i < chunk; i += blockDim.x) {
      const int a = base + i;
      int sgm = 0;
#pragma unroll
      for (int k = 1; k < 9; ++k) sgm += (a >= s_pref[k]) ? 1 : 0;
      recs[i] = b.sorted[s_start[sgm] + (a - s_pref[sgm])];
    }
    __syncthreads();
    spread_cols<P, 1>(recs, chunk, rho, p, X0, Y0, wx, wy, cs);
    __syncthreads();
  }

  // z-transform: two real columns packed per complex line, the bulk-solvent
  // mean subtracted inside the cell in fixed point (one rounding to float,
  // in units of 2^-32), zero padding beyond n_z.  The first FFT stage reads
  // the columns directly.
  const unsigned long long pad = pad_units(b, p, (size_t)f * p.E + e, e);
  const int ty = p.ty;
  auto load = [&](unsigned l, unsigned z) -> float2 {
    const int c0 = 2 * (int)l;
    const int lx = c0 / ty, ly = c0 - lx * ty;
    float re = 0.0f, im = 0.0f;
    if ((int)z < nz && lx < wx) {
      const unsigned long long* c = rho + c0 * cs + z;
      const bool halo = (int)z < P - 1;               // fold the z halo back
      if (ly < wy) re = fixed_minus_units(c[0] + (halo ? c[nz] : 0ull), pad);
      if (ly + 1 < wy) im = fixed_minus_units(c[cs] + (halo ? c[cs + nz] : 0ull), pad);
    }
    return make_float2(re, im);
  };
  fft_lines_loaded<kMaxC, POW2>(buf, p.lsz, nl, p.fz, tw, load);

  // Unpack the two real transforms: X_a(k) = (Z(k) + conj Z(-k))/2,
  // X_b(k) = (Z(k) - conj Z(-k))/(2i).
  const size_t fe = (size_t)f * p.E + e;
  for (int i = threadIdx.x; i < nl * p.kz; i += blockDim.x) {
    const int l = i / p.kz, k = i - l * p.kz;
    const int c0 = 2 * l;
    const int lx = c0 / p.ty, ly = c0 - lx * p.ty;
    if (lx >= wx || ly >= wy) continue;
    const float2 Z = buf[l * p.lsz + k];
    const float2 Zm = buf[l * p.lsz + (k == 0 ? 0 : nnz - k)];
    constexpr float h = 0.5f / 4294967296.0f;       // 1/2 and the 2^-32 unit
    const float2 Xa = make_float2(h * (Z.x + Zm.x), h * (Z.y - Zm.y));
    const float2 Xb = make_float2(h * (Z.y + Zm.y), -h * (Z.x - Zm.x));
    const size_t row = (fe * p.n[0] + (X0 + lx)) * p.n[1] + (Y0 + ly);
    b.zbuf[row * p.kz + k] = Xa;
    if (ly + 1 < wy) b.zbuf[(row + 1) * p.kz + k] = Xb;
  }
}

// Lanes per record in the pruned kernels' spread_cols.  A column tile and
// element holds ~100 records at 0.05 atoms per cell, so one lane per record
// leaves most of a workgroup's 256 lanes idle in the spreading loop; two
// lanes (x planes o = s, s + 2) measured 1.89 ms against 1.98 per 100
// frames of the 128^3 trajectory, four 2.02 (every lane of a record
// computes its B-spline weights).
#ifndef SAXS_SPREAD_LANES
#define SAXS_SPREAD_LANES 2
#endif
constexpr int kSpreadLanes = SAXS_SPREAD_LANES;
constexpr int kPrefetch = 2;                       // records per thread prefetched
constexpr int kAllRecs = kPrefetch * kWG;          // staged records per round
// Records staged per round by the pruned z kernels: fewer, so that four
// workgroups fit the LDS of a CU (a column tile and element holds ~100
// records at 0.05 atoms per cell; denser tiles take more rounds).
constexpr int kZpRecs = 160;
constexpr int kZpPre = (kZpRecs + kWG - 1) / kWG;  // records prefetched per thread
// k_spread_zp at N = 512 (a 256-point z axis): its 512-entry twiddle table
// and 17 x 32 line entries per column leave room for four workgroups a CU
// only with 60 staged records, and its registers are capped to match (four
// waves a SIMD); the extra rounds cost less than the occupancy gains.
__host__ __device__ constexpr int zp_spread_recs(int N) { return N == 512 ? 60 : kZpRecs; }
__host__ __device__ constexpr int zp_spread_waves(int N) { return N == 512 ? 4 : 1; }

__host__ __device__ inline size_t spread_all_lds_bytes(const Params& p) {
  const int cs = p.order == 6 ? spread_col_stride<6>(p.n[2]) : spread_col_stride<4>(p.n[2]);
  const size_t rho = (((size_t)p.tx * p.ty * cs * sizeof(unsigned long long)) + 15) & ~size_t(15);
  size_t buf = (size_t)(p.tx * p.ty / 2) * p.lsz * sizeof(float2);
  const size_t recs = (size_t)kAllRecs * sizeof(float4);
  if (buf < recs) buf = recs;
  const size_t tw = (size_t)p.nn[2] * sizeof(float2);
  const size_t seg = (size_t)p.E * 28 * sizeof(int);
  return rho + buf + tw + ((seg + 15) & ~size_t(15));
}

// Same result as k_spread_z, one workgroup per (column tile, frame) looping
// over the elements: the bin segments of every element are fetched in one
// round trip, and while element e is transformed the first records of
// element e+1 are already in flight (loaded into registers, stored to LDS
// after the transform).
template <int P, int POW2>
__global__ void __launch_bounds__(kWG) k_spread_zall(Params p, Buffers b) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int tile = blockIdx.x, f = blockIdx.y;
  const int bx = tile % p.ntx, by = tile / p.ntx;
  const int X0 = bx * p.tx, Y0 = by * p.ty;
  const int wx = min(p.tx, p.n[0] - X0), wy = min(p.ty, p.n[1] - Y0);
  const int nz = p.n[2], nnz = p.nn[2];
  const int ncol = p.tx * p.ty;
  const int nl = ncol / 2;
  const int E = p.E;
  const int cs = spread_col_stride<P>(nz);
  unsigned long long* rho = reinterpret_cast<unsigned long long*>(smem);
  unsigned char* after_rho = smem + ((ncol * cs * sizeof(unsigned long long) + 15) & ~size_t(15));
  float2* buf = reinterpret_cast<float2*>(after_rho);
  float4* recs = reinterpret_cast<float4*>(after_rho);
  size_t buf_bytes = (size_t)nl * p.lsz * sizeof(float2);
  if (buf_bytes < kAllRecs * sizeof(float4)) buf_bytes = kAllRecs * sizeof(float4);
  float2* tw = reinterpret_cast<float2*>(after_rho + buf_bytes);
  int* s_start = reinterpret_cast<int*>(tw + nnz);   // [E][9]
  int* s_cnt = s_start + 9 * E;                       // [E][9]
  int* s_pref = s_cnt + 9 * E;                        // [E][10]

  for (int t = threadIdx.x; t < 9 * E; t += blockDim.x) {
    const int e = t / 9;
    const int2 r = tile_slot(p, b, f, e, bx, by, t - e * 9);
    s_start[t] = r.x;
    s_cnt[t] = r.y;
  }
  for (int i = threadIdx.x; i < nnz; i += blockDim.x) tw[i] = b.twz[i];
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

  // record a of element e (a < total of e)
  auto fetch = [&](int e, int a) -> float4 {
    const int* pref = s_pref + e * 10;
    int sgm = 0;
#pragma unroll
    for (int k = 1; k < 9; ++k) sgm += (a >= pref[k]) ? 1 : 0;
    return b.sorted[s_start[e * 9 + sgm] + (a - pref[sgm])];
  };

  // first round of element 0
  {
    const int tot0 = s_pref[9];
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) {
      const int a = threadIdx.x + u * blockDim.x;
      if (a < tot0) recs[a] = fetch(0, a);
    }
  }
  float4 pre[kPrefetch];
  for (int e = 0; e < E; ++e) {
    const int total = s_pref[e * 10 + 9];
    for (int i = threadIdx.x; i < ncol * cs; i += blockDim.x) rho[i] = 0ull;
    __syncthreads();
    // first round (already staged), then any further rounds synchronously
    spread_cols<P, 1>(recs, min(total, kAllRecs), rho, p, X0, Y0, wx, wy, cs);
    for (int base = kAllRecs; base < total; base += kAllRecs) {
      const int chunk = min(kAllRecs, total - base);
      __syncthreads();
      for (int i = threadIdx.x; i < chunk; i += blockDim.x) recs[i] = fetch(e, base + i);
      __syncthreads();
      spread_cols<P, 1>(recs, chunk, rho, p, X0, Y0, wx, wy, cs);
    }
    // prefetch the first round of element e+1 into registers
    const int tot_next = (e + 1 < E) ? s_pref[(e + 1) * 10 + 9] : 0;
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) {
      const int a = threadIdx.x + u * blockDim.x;
      if (a < tot_next) pre[u] = fetch(e + 1, a);
    }
    __syncthreads();

    const size_t fe = (size_t)f * E + e;
    const unsigned long long pad = pad_units(b, p, fe, e);
    const int ty = p.ty;
    auto load = [&](unsigned l, unsigned z) -> float2 {
      const int c0 = 2 * (int)l;
      const int lx = c0 / ty, ly = c0 - lx * ty;
      float re = 0.0f, im = 0.0f;
      if ((int)z < nz && lx < wx) {
        const unsigned long long* c = rho + c0 * cs + z;
        const bool halo = (int)z < P - 1;             // fold the z halo back
        if (ly < wy) re = fixed_minus_units(c[0] + (halo ? c[nz] : 0ull), pad);
        if (ly + 1 < wy) im = fixed_minus_units(c[cs] + (halo ? c[cs + nz] : 0ull), pad);
      }
      return make_float2(re, im);
    };
    fft_lines_loaded<kMaxC, POW2>(buf, p.lsz, nl, p.fz, tw, load);
    for (int i = threadIdx.x; i < nl * p.kz; i += blockDim.x) {
      const int l = i / p.kz, k = i - l * p.kz;
      const int c0 = 2 * l;
      const int lx = c0 / p.ty, ly = c0 - lx * p.ty;
      if (lx >= wx || ly >= wy) continue;
      const float2 Z = buf[l * p.lsz + k];
      const float2 Zm = buf[l * p.lsz + (k == 0 ? 0 : nnz - k)];
      constexpr float h = 0.5f / 4294967296.0f;     // 1/2 and the 2^-32 unit
      const float2 Xa = make_float2(h * (Z.x + Zm.x), h * (Z.y - Zm.y));
      const float2 Xb = make_float2(h * (Z.y + Zm.y), -h * (Z.x - Zm.x));
      const size_t row = (fe * p.n[0] + (X0 + lx)) * p.n[1] + (Y0 + ly);
      b.zbuf[row * p.kz + k] = Xa;
      if (ly + 1 < wy) b.zbuf[(row + 1) * p.kz + k] = Xb;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) {
      const int a = threadIdx.x + u * blockDim.x;
      if (a < tot_next) recs[a] = pre[u];
    }
  }
}

// ---- output-pruned z transform (N = 16 * R2, powers of two) --------------
//
// Stage 0 (radix 16): butterfly j < R2 takes x[j + R2 r], r < 16, two real
// columns packed per complex line, writes Y_j[s] at 17 j + s of the line.
// Stage 1 (radix R2): butterfly s < 16 takes Y_j[s] W_N^{j s} over j and
// keeps only the outputs X(s + 16 r') with k = s + 16 r' in [0, kz) or
// (N - kz, N), stored compactly: slot(k) = k for k < kz, kz - 1 + (N - k)
// for the mirrored ones.
//
// LDS: the spread columns (u64, stride cs) and the stage-0 lines (float2,
// stride 2 cs) share one region so that line l occupies exactly the bytes of
// columns 2l and 2l + 1.  The threads that read a column pair in stage 0 are
// the ones that write its line, and those that read the line in stage 1 are
// the ones that clear it for the next element -- all in one wavefront, whose
// LDS operations execute in order -- so these steps need no workgroup
// barrier.

__host__ __device__ inline int zp_col_stride(int nz, int P, int N) {
  const int a = nz + P - 1, b = 17 * (N / 16) / 2;   // z halo; 17 R2 line entries
  return ((a > b ? a : b) + 1) & ~1;
}

__host__ __device__ inline size_t spread_p_lds_bytes(const Params& p) {
  const int N = p.nn[2];
  const int cs = zp_col_stride(p.n[2], p.order, N);
  const size_t A = (size_t)p.tx * p.ty * cs * sizeof(unsigned long long);
  const size_t B = (size_t)zp_spread_recs(N) * sizeof(float4);
  const size_t tw = (size_t)N * sizeof(float2);          // W_N^{j s} as [j][s]
  const size_t seg = (size_t)p.E * (28 * sizeof(int) + sizeof(unsigned long long));
  return A + B + tw + seg;
}

// EXACT: n_z == N/2 (every sigma-2 power-of-two grid), so every stage-0
// input z = j + R2 r, r < 8, lies in the column and needs no bounds test.
template <int P, int N, bool HALF, bool EXACT = false>
__global__ void __launch_bounds__(kWG, zp_spread_waves(N)) k_spread_zp(Params p, Buffers b) {
  constexpr int RECS = zp_spread_recs(N);
  static_assert(!EXACT || HALF, "EXACT implies HALF");
  constexpr int R2 = N / 16;
  constexpr int RIN = HALF ? 8 : 16;                   // stage-0 inputs that can be nonzero
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
  const int cs = zp_col_stride(nz, P, N);
  const int LY = 2 * cs;                               // line stride (float2)
  const size_t A_b = (size_t)ncol * cs * sizeof(unsigned long long);
  const size_t B_b = (size_t)RECS * sizeof(float4);
  unsigned long long* rho = reinterpret_cast<unsigned long long*>(smem);
  float2* ybuf = reinterpret_cast<float2*>(smem);
  float4* recs = reinterpret_cast<float4*>(smem + A_b);
  float2* tw = reinterpret_cast<float2*>(smem + A_b + B_b);
  int* s_start = reinterpret_cast<int*>(tw + N);
  int* s_cnt = s_start + 9 * E;
  int* s_pref = s_cnt + 9 * E;
  // bulk-solvent mean of each element in the columns' fixed-point units (a
  // column outside the cell keeps its zeros), read once for the elements
  unsigned long long* s_pad = reinterpret_cast<unsigned long long*>(s_pref + 10 * E);

  for (int t = threadIdx.x; t < 9 * E; t += blockDim.x) {
    const int e = t / 9;
    const int2 r = tile_slot(p, b, f, e, bx, by, t - e * 9);
    s_start[t] = r.x;
    s_cnt[t] = r.y;
  }
  for (int e = threadIdx.x; e < E; e += blockDim.x) s_pad[e] = pad_units(b, p, (size_t)f * E + e, e);
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = b.twz[(i >> 4) * (i & 15)];
  for (int i = threadIdx.x; i < ncol * cs / 2; i += blockDim.x)
    reinterpret_cast<uint4*>(rho)[i] = make_uint4(0u, 0u, 0u, 0u);
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
    return b.sorted[s_start[e * 9 + sgm] + (a - pref[sgm])];
  };
  {
    const int tot0 = min(s_pref[9], RECS);
#pragma unroll
    for (int u = 0; u < kZpPre; ++u) {
      const int a = threadIdx.x + u * blockDim.x;
      if (a < tot0) recs[a] = fetch(0, a);
    }
  }
  const float inv_ty = 1.0f / (float)p.ty;
  float4 pre[kZpPre];
  for (int e = 0; e < E; ++e) {
    const int total = s_pref[e * 10 + 9];
    __syncthreads();                                 // records staged, columns clear
    spread_cols<P, kSpreadLanes>(recs, min(total, RECS), rho, p, X0, Y0, wx, wy, cs);
    for (int base = RECS; base < total; base += RECS) {
      const int chunk = min(RECS, total - base);
      __syncthreads();
      for (int i = threadIdx.x; i < chunk; i += blockDim.x) recs[i] = fetch(e, base + i);
      __syncthreads();
      spread_cols<P, kSpreadLanes>(recs, chunk, rho, p, X0, Y0, wx, wy, cs);
    }
    const int tot_next = min((e + 1 < E) ? s_pref[(e + 1) * 10 + 9] : 0, RECS);
#pragma unroll
    for (int u = 0; u < kZpPre; ++u) {
      const int a = threadIdx.x + u * blockDim.x;
      if (a < tot_next) pre[u] = fetch(e + 1, a);
    }
    __syncthreads();                                 // columns complete

    const size_t fe = (size_t)f * E + e;
    float2* zt = b.zbuf + ((fe * p.n[0] + X0) * p.n[1] + Y0) * kz;   // tile origin
    const unsigned long long pad = s_pad[e];
    // stage 0 (item = line l, butterfly j): columns 2l, 2l+1 -> line l in place
    if (threadIdx.x < nl * R2) {
      const int l = threadIdx.x / R2, j = threadIdx.x - l * R2;
      const int c0 = 2 * l;
      const int lx = div_small(c0, p.ty, inv_ty), ly = c0 - lx * p.ty;
      const unsigned long long pad0 = (lx < wx && ly < wy) ? pad : 0ull;
      const unsigned long long pad1 = (lx < wx && ly + 1 < wy) ? pad : 0ull;
      const unsigned long long* col0 = rho + c0 * cs;
      const unsigned long long* col1 = col0 + cs;
      // inputs in fixed-point units; the 2^-32 scale is applied to the kept
      // outputs (a power of two commutes exactly with the transform)
      unsigned long long a0[RIN], a1[RIN];
      unsigned big = (unsigned)(pad >> 31);
#pragma unroll
      for (int r = 0; r < RIN; ++r) {
        const int z = j + R2 * r;
        if (EXACT || z < nz) {
          a0[r] = col0[z];
          a1[r] = col1[z];
          if (r == 0 && z < P - 1) {                 // fold the z halo back
            a0[r] += col0[nz + z];
            a1[r] += col1[nz + z];
          }
        } else {
          a0[r] = pad0;
          a1[r] = pad1;
        }
        big |= (unsigned)((a0[r] | a1[r]) >> 31);
      }
      float2 v[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) v[r] = make_float2(0.0f, 0.0f);
      if (!__any(big != 0u)) {
        // every value and the pad below 2^31 units: exact 32-bit difference
#pragma unroll
        for (int r = 0; r < RIN; ++r)
          v[r] = make_float2((float)(int)((unsigned)a0[r] - (unsigned)pad0),
                             (float)(int)((unsigned)a1[r] - (unsigned)pad1));
      } else {
#pragma unroll
        for (int r = 0; r < RIN; ++r)
          v[r] = make_float2(fixed_minus_units(a0[r], pad0), fixed_minus_units(a1[r], pad1));
      }
      dft16<HALF>(v);
      __builtin_amdgcn_wave_barrier();
      float2* dst = ybuf + l * LY + 17 * j;
#pragma unroll
      for (int r = 0; r < 16; ++r) dst[r] = v[r];
    }
    // lines complete: with R2 = 16 thread t takes line t / 16 in both stages,
    // so a wavefront reads in stage 1 only the lines it wrote in stage 0 (its
    // LDS operations execute in order); otherwise lines cross wavefronts
    if constexpr (R2 == 16) __builtin_amdgcn_wave_barrier();
    else __syncthreads();
    // stage 1 (item = line l, butterfly s): kept outputs Z(k), k = s + 16 rr;
    // the two real columns of the line are separated with Z(N - k), which
    // lane (16 - s) mod 16 of the same line holds; then clear line l
    for (int item = threadIdx.x; item < nl * 16; item += blockDim.x) {
      const int l = item >> 4, j = item & 15;
      float2 t[R2];
#pragma unroll
      for (int r = 0; r < R2; ++r) {
        const float2 y = ybuf[l * LY + 17 * r + j];
        t[r] = (r == 0) ? y : cmul(y, tw[r * 16 + j]);
      }
      if constexpr (R2 == 8) Radix<8>::dft(t);
      else if constexpr (R2 == 16) dft16_ends(t);             // outputs 0, 1, 14, 15
      else dft32_ends(t);                                     // outputs 0..3, 28..31
      // kz <= N / 8 (plan check): at most KL = N / 128 kept outputs at each end
      constexpr int KL = N / 128;
      // Z(N - k) = t[R2 - 1 - rr] of lane (16 - j) mod 16; for j = 0 it is
      // t[0] (k = 0) and the lane's own t[R2 - rr]
      const int src = (threadIdx.x & ~15) | ((16 - j) & 15);
      float2 zm[KL];
#pragma unroll
      for (int rr = 0; rr < KL; ++rr) {
        const float2 m = t[R2 - 1 - rr];
        zm[rr] = make_float2(__shfl(m.x, src, 64), __shfl(m.y, src, 64));
      }
      if (j == 0) {
#pragma unroll
        for (int rr = KL - 1; rr >= 1; --rr) zm[rr] = zm[rr - 1];
        zm[0] = t[0];
      }
      const int c0 = 2 * l;
      const int lx = div_small(c0, p.ty, inv_ty), ly = c0 - lx * p.ty;
      if (lx < wx && ly < wy) {
        float2* dst = zt + (lx * p.n[1] + ly) * kz;
        constexpr float h = 0.5f / 4294967296.0f;   // 1/2 and the 2^-32 unit
#pragma unroll
        for (int rr = 0; rr < KL; ++rr) {
          const int k = j + 16 * rr;
          if (k < kz) {
            const float2 Z = t[rr], Zm = zm[rr];
            dst[k] = make_float2(h * (Z.x + Zm.x), h * (Z.y - Zm.y));
            if (ly + 1 < wy) dst[kz + k] = make_float2(h * (Z.y + Zm.y), -h * (Z.x - Zm.x));
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
      uint4* line = reinterpret_cast<uint4*>(ybuf + l * LY);   // LY / 2 uint4
      for (int i = j; i < LY / 2; i += 16) line[i] = make_uint4(0u, 0u, 0u, 0u);
    }
#pragma unroll
    for (int u = 0; u < kZpPre; ++u) {
      const int a = threadIdx.x + u * blockDim.x;
      if (a < tot_next) recs[a] = pre[u];
    }
  }
}

// ---- output-pruned z transform for any N = 2H R2 with n_z <= N/2 --------
//
// Stage 0 (radix R1 = 2H): butterfly j < R2 takes x[j + R2 r]; only r < H
// can be nonzero (z < n_z <= N/2), so its length-R1 DFT is two length-H
// DFTs, Y_j[2t] = DFT_H(x_r)[t] and Y_j[2t+1] = DFT_H(x_r W_R1^r)[t].
// Stage 1 forms only the kept outputs as sums over j,
//   Z(k)     = sum_j Y_j[k mod R1]  W_N^{jk}
//   Z(N - k) = sum_j Y_j[-k mod R1] conj(W_N^{jk}),   k < kz,
// evaluated by Horner's rule in W_N^k (2 kz R2 complex multiply-adds per
// line against the N log N of a full transform whose outputs are mostly
// discarded; kz is about N/30 here).
// Y_j[s] sits at (R1 + 1) j + s of its line, and line l occupies the bytes
// of columns 2l and 2l + 1 as in k_spread_zp.  Each wavefront owns a block
// of lines and takes them through both stages and the clear for the next
// element on its own: a stage-0 pass holds whole lines (64 / R2 of them),
// and a wavefront's LDS operations execute in order, so no workgroup
// barrier separates the stages.  Stage 1 splits each output's sum over up
// to 8 lanes (as many as fit kz pairs in a wavefront) and adds the parts
// with lane shuffles.

__host__ __device__ inline int zg_col_stride(int nz, int P, int N, int R1) {
  const int a = nz + P - 1, b = (N / R1 * (R1 + 1) + 1) / 2;
  return ((a > b ? a : b) + 1) & ~1;
}

__host__ __device__ inline size_t spread_g_lds_bytes(const Params& p, int R1) {
  const int N = p.nn[2];
  const int cs = zg_col_stride(p.n[2], p.order, N, R1);
  const size_t A = (size_t)p.tx * p.ty * cs * sizeof(unsigned long long);
  const size_t B = (size_t)kZpRecs * sizeof(float4);
  const size_t tw = (size_t)N * sizeof(float2);          // W_N^m
  const size_t seg = (size_t)p.E * (28 * sizeof(int) + sizeof(unsigned long long));
  return A + B + tw + seg;
}

// Horner accumulation for the pruned transforms' kept outputs: with y_t the
// entries p[s - t D] and p[sm - t D] (t = 0 .. T-1, read from the highest
// term down; D a compile-time stride, so the loads take immediate LDS
// offsets from two running pointers), returns Z = sum_t y1_t u^(T-1-t) and
// Zm = sum_t y2_t conj(u)^(T-1-t), i.e. Horner's rule over the terms in
// descending order.  Four terms' loads are issued before their multiply-
// adds: the chain is bound by LDS latency otherwise.
template <int D>
__device__ __forceinline__ void horner_pair_c(const float2* p, int s, int sm, int T, float2 u,
                                              float2& Z, float2& Zm) {
  float2 A = make_float2(0.0f, 0.0f), B = make_float2(0.0f, 0.0f);
  const float2* p1 = p + s;
  const float2* p2 = p + sm;
  int t = 0;
  for (; t + 4 <= T; t += 4) {
    float2 y1[4], y2[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      y1[v] = p1[-v * D];
      y2[v] = p2[-v * D];
    }
    p1 -= 4 * D;
    p2 -= 4 * D;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const float ax = fmaf(A.x, u.x, fmaf(-A.y, u.y, y1[v].x));
      const float ay = fmaf(A.x, u.y, fmaf(A.y, u.x, y1[v].y));
      const float bx = fmaf(B.x, u.x, fmaf(B.y, u.y, y2[v].x));
      const float by = fmaf(B.y, u.x, fmaf(-B.x, u.y, y2[v].y));
      A = make_float2(ax, ay);
      B = make_float2(bx, by);
    }
  }
  for (; t < T; ++t) {
    const float2 y1 = *p1, y2 = *p2;
    p1 -= D;
    p2 -= D;
    const float ax = fmaf(A.x, u.x, fmaf(-A.y, u.y, y1.x));
    const float ay = fmaf(A.x, u.y, fmaf(A.y, u.x, y1.y));
    const float bx = fmaf(B.x, u.x, fmaf(B.y, u.y, y2.x));
    const float by = fmaf(B.y, u.x, fmaf(-B.x, u.y, y2.y));
    A = make_float2(ax, ay);
    B = make_float2(bx, by);
  }
  Z = A;
  Zm = B;
}

// a mod n for 0 <= a < 2^22 (div_small's range).
__device__ __forceinline__ int mod_small(int a, int n, float inv_n) {
  return a - n * div_small(a, n, inv_n);
}

// The plan guarantees R2 <= 64 (a stage-0 pass holds a whole line) and
// R2 >= P (the z halo folds onto butterfly inputs with r = 0 only).
template <int P, int H>
__global__ void __launch_bounds__(kWG) k_spread_zg(Params p, Buffers b) {
  constexpr int R1 = 2 * H, YS = R1 + 1;
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
  const int cs = zg_col_stride(nz, P, N, R1);
  const int LY = 2 * cs;                               // line stride (float2)
  const size_t A_b = (size_t)ncol * cs * sizeof(unsigned long long);
  const size_t B_b = (size_t)kZpRecs * sizeof(float4);
  unsigned long long* rho = reinterpret_cast<unsigned long long*>(smem);
  float2* ybuf = reinterpret_cast<float2*>(smem);
  float4* recs = reinterpret_cast<float4*>(smem + A_b);
  float2* tw = reinterpret_cast<float2*>(smem + A_b + B_b);
  int* s_start = reinterpret_cast<int*>(tw + N);
  int* s_cnt = s_start + 9 * E;
  int* s_pref = s_cnt + 9 * E;
  // bulk-solvent mean of each element in the columns' fixed-point units (a
  // column outside the cell keeps its zeros), read once for the elements
  unsigned long long* s_pad = reinterpret_cast<unsigned long long*>(s_pref + 10 * E);

  for (int t = threadIdx.x; t < 9 * E; t += blockDim.x) {
    const int e = t / 9;
    const int2 r = tile_slot(p, b, f, e, bx, by, t - e * 9);
    s_start[t] = r.x;
    s_cnt[t] = r.y;
  }
  for (int e = threadIdx.x; e < E; e += blockDim.x) s_pad[e] = pad_units(b, p, (size_t)f * E + e, e);
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = b.twz[i];
  for (int i = threadIdx.x; i < ncol * cs / 2; i += blockDim.x)
    reinterpret_cast<uint4*>(rho)[i] = make_uint4(0u, 0u, 0u, 0u);
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
    return b.sorted[s_start[e * 9 + sgm] + (a - pref[sgm])];
  };
  {
    const int tot0 = min(s_pref[9], kZpRecs);
#pragma unroll
    for (int u = 0; u < kZpPre; ++u) {
      const int a = threadIdx.x + u * blockDim.x;
      if (a < tot0) recs[a] = fetch(0, a);
    }
  }
  const float inv_ty = 1.0f / (float)p.ty;
  const float inv_r2 = 1.0f / (float)R2;
  const int g0 = 64 / R2;                            // lines per stage-0 pass
  const int lpw = (nl + kWG / 64 - 1) / (kWG / 64);  // lines per wavefront
  const int lw0 = min(nl, (int)(threadIdx.x >> 6) * lpw), lw1 = min(nl, lw0 + lpw);
  // stage 1: nq output pairs of this wavefront, each summed by split lanes;
  // split depends on kz alone, so every output is summed in the same order
  // whatever the tile (and the spectra do not depend on it)
  const int nq = (lw1 - lw0) * kz;
  // (kz split <= 32: a wavefront's two or more lines then fill its lanes
  // without more iterations than an unsplit sum, which costs fewer
  // instructions than the shorter parts plus their shuffles)
  int lgs = 0;
  while (lgs < 3 && (kz << (lgs + 1)) <= 32) ++lgs;
  const int split = 1 << lgs;
  const int T0 = (R2 + split - 1) / split;           // terms of one part of a stage-1 sum
  const float inv_kz = 1.0f / (float)kz, inv_n = 1.0f / (float)N;
  float4 pre[kZpPre];
  for (int e = 0; e < E; ++e) {
    const int total = s_pref[e * 10 + 9];
    __syncthreads();                                 // records staged, columns clear
    spread_cols<P, kSpreadLanes>(recs, min(total, kZpRecs), rho, p, X0, Y0, wx, wy, cs);
    for (int base = kZpRecs; base < total; base += kZpRecs) {
      const int chunk = min(kZpRecs, total - base);
      __syncthreads();
      for (int i = threadIdx.x; i < chunk; i += blockDim.x) recs[i] = fetch(e, base + i);
      __syncthreads();
      spread_cols<P, kSpreadLanes>(recs, chunk, rho, p, X0, Y0, wx, wy, cs);
    }
    const int tot_next = min((e + 1 < E) ? s_pref[(e + 1) * 10 + 9] : 0, kZpRecs);
#pragma unroll
    for (int u = 0; u < kZpPre; ++u) {
      const int a = threadIdx.x + u * blockDim.x;
      if (a < tot_next) pre[u] = fetch(e + 1, a);
    }
    __syncthreads();                                 // columns complete

    const size_t fe = (size_t)f * E + e;
    float2* zt = b.zbuf + ((fe * p.n[0] + X0) * p.n[1] + Y0) * kz;   // tile origin
    const unsigned long long pad = s_pad[e];
    const int lane = threadIdx.x & 63;
    // stage 0, a pass of g0 whole lines: lane -> (line, butterfly j)
    {
      const int li = div_small(lane, R2, inv_r2), j = lane - li * R2;
      for (int lg = lw0; lg < lw1; lg += g0) {
        const int l = lg + li;
        if (li < g0 && l < lw1) {
          const int c0 = 2 * l;
          const int lx = div_small(c0, p.ty, inv_ty), ly = c0 - lx * p.ty;
          const unsigned long long pad0 = (lx < wx && ly < wy) ? pad : 0ull;
          const unsigned long long pad1 = (lx < wx && ly + 1 < wy) ? pad : 0ull;
          const unsigned long long* col0 = rho + c0 * cs;
          const unsigned long long* col1 = col0 + cs;
          unsigned long long a0[H], a1[H];
          unsigned big = (unsigned)(pad >> 31);
#pragma unroll
          for (int r = 0; r < H; ++r) {
            const int z = j + R2 * r;
            if (z < nz) {
              a0[r] = col0[z];
              a1[r] = col1[z];
              if (r == 0 && z < P - 1) {             // fold the z halo back
                a0[r] += col0[nz + z];
                a1[r] += col1[nz + z];
              }
            } else {
              a0[r] = pad0;
              a1[r] = pad1;
            }
            big |= (unsigned)((a0[r] | a1[r]) >> 31);
          }
          float2 xe[H], xo[H];
          if (big == 0u) {
            // every value and the pad below 2^31 units: exact 32-bit difference
#pragma unroll
            for (int r = 0; r < H; ++r)
              xe[r] = make_float2((float)(int)((unsigned)a0[r] - (unsigned)pad0),
                                  (float)(int)((unsigned)a1[r] - (unsigned)pad1));
          } else {
#pragma unroll
            for (int r = 0; r < H; ++r)
              xe[r] = make_float2(fixed_minus_units(a0[r], pad0), fixed_minus_units(a1[r], pad1));
          }
          xo[0] = xe[0];
#pragma unroll
          for (int r = 1; r < H; ++r) xo[r] = cmul(xe[r], tw[r * R2]);   // W_R1^r = W_N^{r R2}
          Radix<H>::dft(xe);
          Radix<H>::dft(xo);
          __builtin_amdgcn_wave_barrier();
          float2* dst = ybuf + l * LY + YS * j;
#pragma unroll
          for (int t = 0; t < H; ++t) {
            dst[2 * t] = xe[t];
            dst[2 * t + 1] = xo[t];
          }
        }
        __builtin_amdgcn_wave_barrier();
      }
    }
    // stage 1 over the wavefront's lines at once: lane -> (line, output pair
    // k, part of the sum over j), the parts added with lane shuffles; then
    // the two real columns of each line, X_a(k) = (Z(k) + conj Z(N-k))/2 and
    // X_b(k) = (Z(k) - conj Z(N-k))/(2i); then clear the lines
    for (int q0 = 0; q0 < nq; q0 += 64 >> lgs) {
      const int q = q0 + (lane >> lgs), part = lane & (split - 1);
      float2 Z = make_float2(0.0f, 0.0f), Zm = make_float2(0.0f, 0.0f);
      int l = 0, k = 0;
      if (q < nq) {
        l = div_small(q, kz, inv_kz);
        k = q - l * kz;
        l += lw0;
        const int s = k % R1, sm = s == 0 ? 0 : R1 - s;
        // the part's terms j = j0 .. j0 + T - 1 (j0 = part T0) by Horner's
        // rule in u = W_N^k from the last term down (conj u for Z(N - k)),
        // then the factor W_N^{j0 k}: sum_j y_j W_N^{jk} = W_N^{j0 k} sum_t
        // y_{j0 + t} u^t (k < kz <= N/2 + 1, so u = tw[k])
        const int j0 = part * T0, T = min(T0, R2 - j0);
        const float2 u = tw[k], wp = tw[mod_small(j0 * k, N, inv_n)];
        if (T > 0) horner_pair_c<YS>(ybuf + l * LY + YS * (j0 + T - 1), s, sm, T, u, Z, Zm);
        Z = cmul(Z, wp);
        Zm = cmul(Zm, make_float2(wp.x, -wp.y));
      }
      for (int m = 1; m < split; m <<= 1) {
        Z.x += __shfl_xor(Z.x, m, 64);
        Z.y += __shfl_xor(Z.y, m, 64);
        Zm.x += __shfl_xor(Zm.x, m, 64);
        Zm.y += __shfl_xor(Zm.y, m, 64);
      }
      if (part == 0 && q < nq) {
        const int c0 = 2 * l;
        const int lx = div_small(c0, p.ty, inv_ty), ly = c0 - lx * p.ty;
        if (lx < wx && ly < wy) {
          float2* dst = zt + (lx * p.n[1] + ly) * kz;
          constexpr float h = 0.5f / 4294967296.0f; // 1/2 and the 2^-32 unit
          dst[k] = make_float2(h * (Z.x + Zm.x), h * (Z.y - Zm.y));
          if (ly + 1 < wy) dst[kz + k] = make_float2(h * (Z.y + Zm.y), -h * (Z.x - Zm.x));
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    {
      uint4* lines = reinterpret_cast<uint4*>(ybuf + lw0 * LY);   // LY / 2 uint4 a line
      for (int i = lane; i < (lw1 - lw0) * (LY / 2); i += 64) lines[i] = make_uint4(0u, 0u, 0u, 0u);
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < kZpPre; ++u) {
      const int a = threadIdx.x + u * blockDim.x;
      if (a < tot_next) recs[a] = pre[u];
    }
  }
}

__host__ __device__ inline size_t yfft_lds_bytes(const Params& p) {
  return (size_t)p.kcy * p.lsy * sizeof(float2) + (size_t)p.nn[1] * sizeof(float2);
}

template <int POW2>
__global__ void __launch_bounds__(kWG) k_yfft(Params p, Buffers b) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int x = blockIdx.x, ch = blockIdx.y;
  const size_t fe = blockIdx.z;
  const int k0 = ch * p.kcy;
  const int kc = min(p.kcy, p.kz - k0);
  const int ny = p.n[1], nny = p.nn[1];
  float2* buf = reinterpret_cast<float2*>(smem);
  float2* tw = buf + (size_t)p.kcy * p.lsy;
  for (int i = threadIdx.x; i < nny; i += blockDim.x) tw[i] = b.twy[i];
  const float2* src = b.zbuf + (fe * p.n[0] + x) * (size_t)ny * p.kz;
  for (int i = threadIdx.x; i < ny * kc; i += blockDim.x) {
    const int y = i / kc, k = i - y * kc;
    buf[k * p.lsy + y] = src[(size_t)y * p.kz + k0 + k];
  }
  for (int i = threadIdx.x; i < (nny - ny) * kc; i += blockDim.x) {
    const int y = ny + i / kc, k = i % kc;
    buf[k * p.lsy + y] = make_float2(0.0f, 0.0f);
  }
  __syncthreads();
  fft_lines<kMaxC, POW2>(buf, p.lsy, kc, p.fy, tw);
  for (int i = threadIdx.x; i < p.ky * kc; i += blockDim.x) {
    const int j = i / kc, k = i - j * kc;
    const int iy = (p.ky == nny) ? j : (j <= p.Ky ? j : nny - (p.ky - j));
    b.ybuf[((fe * p.ky + j) * p.n[0] + x) * p.kz + k0 + k] = buf[k * p.lsy + iy];
  }
}

// Pruned y pass for N = nn_y = 16 R2 (R2 = 8, 16, 32): the n_x kz lines of
// one (frame, element) are taken 256 / R2 per workgroup; each line's
// y < n_y inputs (zero-padded to N) go straight from zbuf into registers,
// y = j + R2 r feeding the radix-16 DFT of item (line, j); one LDS exchange
// (line stride 17 R2 + 1, Y_j[s] at 17 j + s); stage 1 item (line, s) forms
// the radix-R2 DFT over j of Y_j[s] W_N^{j s} and writes only the kept
// outputs m = s + 16 r' with m <= Ky or m >= N - Ky.
template <int N>
__host__ __device__ constexpr int yfft_p_lines() { return kWG / (N / 16); }

template <int N>
__host__ __device__ inline size_t yfft_p_lds_bytes() {
  return (size_t)yfft_p_lines<N>() * (17 * (N / 16) + 1) * sizeof(float2) + N * sizeof(float2);
}

// ENDS: Ky < N / 8, so the kept outputs are r' < KL and r' >= R2 - KL with
// KL = N / 128 (the others are never formed).  HALF: 2 n_y <= N, so inputs
// r >= 8 are zero.
template <int N, bool ENDS, bool HALF>
__global__ void __launch_bounds__(kWG) k_yfft_p(Params p, Buffers b) {
  constexpr int R2 = N / 16;
  constexpr int LY = 17 * R2 + 1;
  constexpr int LW = yfft_p_lines<N>();
  extern __shared__ __align__(16) unsigned char smem[];
  float2* ybuf = reinterpret_cast<float2*>(smem);
  float2* tw = ybuf + (size_t)LW * LY;                 // W_N^{j s} stored [j][s]
  const int nx = p.n[0], ny = p.n[1], kz = p.kz;
  const int nlines = nx * kz;
  // Neighbouring groups share the 128-byte lines where the kz-long rows of
  // zbuf straddle group boundaries; give each XCD a contiguous range.
  const unsigned grp = xcd_contiguous(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const size_t fe = grp / gridDim.x;
  const int g0 = (int)(grp % gridDim.x) * LW;
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = b.twy[(i >> 4) * (i & 15)];
  const float2* src = b.zbuf + fe * (size_t)nx * ny * kz;
  {
    const int l = threadIdx.x % LW, j = threadIdx.x / LW;      // j < R2
    const int g = g0 + l;
    float2 v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = make_float2(0.0f, 0.0f);
    if (g < nlines) {
      const int x = g / kz, k = g - x * kz;
      const float2* col = src + (size_t)x * ny * kz + k;
#pragma unroll
      for (int r = 0; r < (HALF ? 8 : 16); ++r) {
        const int y = j + R2 * r;
        if (y < ny) v[r] = col[(size_t)y * kz];
      }
    }
    dft16<HALF>(v);
    float2* dst = ybuf + l * LY + 17 * j;
#pragma unroll
    for (int r = 0; r < 16; ++r) dst[r] = v[r];
  }
  __syncthreads();
  const int Ky = p.Ky, ky = p.ky;
  for (int item = threadIdx.x; item < LW * 16; item += blockDim.x) {
    const int l = item % LW, s = item / LW;
    const int g = g0 + l;
    if (g >= nlines) continue;
    float2 t[R2];
#pragma unroll
    for (int r = 0; r < R2; ++r) {
      const float2 y = ybuf[l * LY + 17 * r + s];
      t[r] = (r == 0) ? y : cmul(y, tw[r * 16 + s]);
    }
    if constexpr (R2 == 8) Radix<8>::dft(t);
    else if constexpr (R2 == 16) dft16<false>(t);
    else dft32(t);
    const int x = g / kz, k = g - x * kz;
    float2* out = b.ybuf + (fe * ky * nx + x) * kz + k;  // + jy nx kz
    if constexpr (ENDS) {
      constexpr int KL = N / 128;
#pragma unroll
      for (int rr = 0; rr < KL; ++rr) {
        const int m = s + 16 * rr;
        if (m <= Ky) out[(size_t)m * nx * kz] = t[rr];
      }
#pragma unroll
      for (int rr = R2 - KL; rr < R2; ++rr) {
        const int m = s + 16 * rr;
        if (m >= N - Ky) out[(size_t)(ky - (N - m)) * nx * kz] = t[rr];
      }
    } else {
#pragma unroll
      for (int rr = 0; rr < R2; ++rr) {
        const int m = s + 16 * rr;
        if (m <= Ky || m >= N - Ky) {
          const int jy = m <= Ky ? m : ky - (N - m);
          out[(size_t)jy * nx * kz] = t[rr];
        }
      }
    }
  }
}

__host__ __device__ inline size_t xfft_lds_bytes(const Params& p) {
  return (((size_t)p.num_bins * 12 + 15) & ~size_t(15)) + (size_t)p.kcx * p.lsx * sizeof(float2) +
         (size_t)p.nn[0] * sizeof(float2);
}

// ---- output-pruned y transform for any N = 2H R2 with n_y <= N/2 ---------
//
// The z-transformed rows of zbuf are transformed along y keeping only m_y in
// [0, Ky] and (N - Ky, N).  Stage 0 as in k_spread_zg: butterfly j < R2 of
// line (x, kz) takes the inputs y = j + R2 r, r < H (the rest is zero
// padding), as two length-H DFTs.  Stage 1 as sums over j, the pair m,
// N - m evaluated together by Horner's rule in W_N^m and its conjugate:
//   Y(m) = sum_j T_j[m mod R1] W_N^{jm},  Y(N-m) = sum_j T_j[-m mod R1] conj(W_N^{jm}).
// A workgroup takes LW = kWG / R2 lines with consecutive kz, so loads and
// stores are coalesced along kz.  The plan guarantees 2 Ky + 1 < N.

__host__ __device__ inline int yg_line_stride(int R1, int R2) { return (R2 * (R1 + 1)) | 1; }

__host__ __device__ inline size_t yfft_g_lds_bytes(int N, int R1) {
  const int R2 = N / R1;
  return (size_t)(kWG / R2) * yg_line_stride(R1, R2) * sizeof(float2) + (size_t)N * sizeof(float2);
}

template <int H>
__global__ void __launch_bounds__(kWG) k_yfft_g(Params p, Buffers b) {
  constexpr int R1 = 2 * H, YS = R1 + 1;
  const int N = p.nn[1], R2 = N / R1;
  const int LW = kWG / R2, LY = yg_line_stride(R1, R2);
  extern __shared__ __align__(16) unsigned char smem[];
  float2* tbuf = reinterpret_cast<float2*>(smem);
  float2* tw = tbuf + (size_t)LW * LY;
  const int nx = p.n[0], ny = p.n[1], kz = p.kz;
  const int nlines = nx * kz;
  const unsigned grp = xcd_contiguous(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const size_t fe = grp / gridDim.x;
  const int g0 = (int)(grp % gridDim.x) * LW;
  const float inv_lw = 1.0f / (float)LW, inv_kz = 1.0f / (float)kz;
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = b.twy[i];
  const float2* src = b.zbuf + fe * (size_t)nx * ny * kz;
  // stage 0 (item = butterfly j, line l; l fastest)
  const bool act0 = (int)threadIdx.x < LW * R2;
  const int j0 = div_small(threadIdx.x, LW, inv_lw), l0 = threadIdx.x - j0 * LW;
  float2 xe[H];
#pragma unroll
  for (int r = 0; r < H; ++r) xe[r] = make_float2(0.0f, 0.0f);
  if (act0 && g0 + l0 < nlines) {
    const int g = g0 + l0;
    const int x = div_small(g, kz, inv_kz), k = g - x * kz;
    const float2* col = src + (size_t)x * ny * kz + k;
#pragma unroll
    for (int r = 0; r < H; ++r) {
      const int y = j0 + R2 * r;
      if (y < ny) xe[r] = col[(size_t)y * kz];
    }
  }
  __syncthreads();                                   // twiddles
  if (act0) {
    float2 xo[H];
    xo[0] = xe[0];
#pragma unroll
    for (int r = 1; r < H; ++r) xo[r] = cmul(xe[r], tw[r * R2]);   // W_R1^r = W_N^{r R2}
    Radix<H>::dft(xe);
    Radix<H>::dft(xo);
    float2* dst = tbuf + l0 * LY + YS * j0;
#pragma unroll
    for (int t = 0; t < H; ++t) {
      dst[2 * t] = xe[t];
      dst[2 * t + 1] = xo[t];
    }
  }
  __syncthreads();                                   // lines complete
  // stage 1 (item = output pair m <= Ky, line l, part of the sum over j)
  const int Ky = p.Ky, ky = p.ky;
  const int nq = LW * (Ky + 1);
  int lgs = 0;
  while (lgs < 3 && (nq << (lgs + 1)) <= kWG) ++lgs;
  const int split = 1 << lgs;
  const int T0 = (R2 + split - 1) / split;           // terms of one part of a stage-1 sum
  const float inv_n = 1.0f / (float)N;
  for (int t0 = 0; t0 < (nq << lgs); t0 += kWG) {
    const int t = t0 + threadIdx.x;
    const int part = t & (split - 1), qq = t >> lgs;
    const int m = div_small(qq, LW, inv_lw), l = qq - m * LW;
    const bool act = qq < nq && g0 + l < nlines;
    float2 A = make_float2(0.0f, 0.0f), B = make_float2(0.0f, 0.0f);
    if (act) {
      const int s = m % R1, sm = s == 0 ? 0 : R1 - s;
      // sum_j y_j W_N^{jm} over the part's j = j0 .. j0 + T - 1 (j0 = part T0)
      // by Horner's rule in u = W_N^m (conj u for the mirrored output), times
      // W_N^{j0 m} (m <= Ky < N/2, so u = tw[m])
      const int j0 = part * T0, T = min(T0, R2 - j0);
      const float2 u = tw[m], wp = tw[mod_small(j0 * m, N, inv_n)];
      if (T > 0) horner_pair_c<YS>(tbuf + l * LY + YS * (j0 + T - 1), s, sm, T, u, A, B);
      A = cmul(A, wp);
      B = cmul(B, make_float2(wp.x, -wp.y));
    }
    for (int sh = 1; sh < split; sh <<= 1) {
      A.x += __shfl_xor(A.x, sh, 64);
      A.y += __shfl_xor(A.y, sh, 64);
      B.x += __shfl_xor(B.x, sh, 64);
      B.y += __shfl_xor(B.y, sh, 64);
    }
    if (act && part == 0) {
      const int g = g0 + l;
      const int x = div_small(g, kz, inv_kz), k = g - x * kz;
      float2* out = b.ybuf + (fe * ky * nx + x) * kz + k;   // + jy nx kz
      out[(size_t)m * nx * kz] = A;
      if (m > 0) out[(size_t)(ky - m) * nx * kz] = B;
    }
  }
}

// ---- y transform for N = 2H R2 (R2 = 16 or 32), most outputs kept -------
//
// Stage 0 as in k_yfft_g (butterfly j < R2 of a line: two length-H DFTs of
// the inputs y = j + R2 r, r < H).  Stage 1 forms every output: item
// (residue s < 2H, line l) takes the radix-R2 DFT over j of T_j[s] W_N^{js},
// which gives Y(s + 2H r'), r' < R2, and writes the kept ones (m <= Ky or
// m >= N - Ky).  A line then costs 2H radix-R2 DFTs against Horner's
// 2 (Ky + 1) R2 complex multiply-adds, fewer once Ky reaches about N / 8
// (the reference CLI's defaults keep half the axis).  A workgroup takes
// yf_lines consecutive lines.  The plan guarantees n_y <= N / 2 and
// 2 Ky + 1 < N.
// Lines a workgroup of the full-output y pass takes (one stage-0 item a
// thread; stage 1 then takes two rounds where 2H > R2), and of its adjoint
// (one item a thread in both stages: with the global reads in its first
// stage a second round would expose their latency again).
__host__ __device__ constexpr int yf_lines(int R1, int R2) { return kWG / R2; }
__host__ __device__ constexpr int yf_adj_lines(int R1, int R2) {
  return kWG / (R1 > R2 ? R1 : R2);
}

__host__ __device__ inline size_t yfft_gf_lds_bytes(int N, int R2) {
  const int R1 = N / R2;
  return (size_t)yf_lines(R1, R2) * yg_line_stride(R1, R2) * sizeof(float2) +
         (size_t)N * sizeof(float2);
}

template <int H, int R2>
__global__ void __launch_bounds__(kWG) k_yfft_gf(Params p, Buffers b) {
  static_assert(R2 == 16 || R2 == 32, "radix-R2 stage 1");
  constexpr int R1 = 2 * H, YS = R1 + 1, N = R1 * R2, LW = yf_lines(R1, R2);
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
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = b.twy[i];
  const float2* src = b.zbuf + fe * (size_t)nx * ny * kz;
  // stage 0 (item = butterfly j, line l; l fastest): LW R2 <= kWG items
  const int j0 = threadIdx.x / LW, l0 = threadIdx.x % LW;
  const bool act0 = j0 < R2;
  float2 xe[H];
#pragma unroll
  for (int r = 0; r < H; ++r) xe[r] = make_float2(0.0f, 0.0f);
  if (act0 && g0 + l0 < nlines) {
    const int g = g0 + l0;
    const int x = div_small(g, kz, inv_kz), k = g - x * kz;
    const float2* col = src + (size_t)x * ny * kz + k;
#pragma unroll
    for (int r = 0; r < H; ++r) {
      const int y = j0 + R2 * r;
      if (y < ny) xe[r] = col[(size_t)y * kz];
    }
  }
  __syncthreads();                                   // twiddles
  if (act0) {
    float2 xo[H];
    xo[0] = xe[0];
#pragma unroll
    for (int r = 1; r < H; ++r) xo[r] = cmul(xe[r], tw[r * R2]);   // W_R1^r = W_N^{r R2}
    Radix<H>::dft(xe);
    Radix<H>::dft(xo);
    float2* dst = tbuf + l0 * LY + YS * j0;
#pragma unroll
    for (int t = 0; t < H; ++t) {
      dst[2 * t] = xe[t];
      dst[2 * t + 1] = xo[t];
    }
  }
  __syncthreads();                                   // lines complete
  // stage 1 (item = residue s, line l; l fastest)
  const int Ky = p.Ky, ky = p.ky;
  const size_t row = (size_t)nx * kz;
  for (int it = threadIdx.x; it < LW * R1; it += blockDim.x) {
    const int s = it / LW, l = it % LW;
    const int g = g0 + l;
    if (g >= nlines) continue;
    const float2* yl = tbuf + l * LY + s;
    float2 t[R2];
    t[0] = yl[0];
#pragma unroll
    for (int j = 1; j < R2; ++j) t[j] = cmul(yl[YS * j], tw[j * s]);   // j s < N
    if constexpr (R2 == 16) dft16<false>(t);
    else dft32(t);
    const int x = div_small(g, kz, inv_kz), k = g - x * kz;
    float2* out = b.ybuf + (fe * ky * nx + x) * kz + k;     // + jy nx kz
#pragma unroll
    for (int rr = 0; rr < R2; ++rr) {
      const int m = s + R1 * rr;
      if (m <= Ky) out[(size_t)m * row] = t[rr];
      else if (m >= N - Ky) out[(size_t)(ky - (N - m)) * row] = t[rr];
    }
  }
}

template <int POW2>
__global__ void __launch_bounds__(kWG) k_xfft_epi(Params p, Buffers b) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int jy = blockIdx.x, ch = blockIdx.y, f = blockIdx.z;
  const int k0 = ch * p.kcx;
  const int kc = min(p.kcx, p.kz - k0);
  const int nx = p.n[0], nnx = p.nn[0];
  const int nb = p.num_bins;
  double* Hs = reinterpret_cast<double*>(smem);
  unsigned int* Cs = reinterpret_cast<unsigned int*>(Hs + nb);
  float2* buf = reinterpret_cast<float2*>(smem + (((size_t)nb * 12 + 15) & ~size_t(15)));
  float2* tw = buf + (size_t)p.kcx * p.lsx;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    Hs[i] = 0.0;
    Cs[i] = 0u;
  }
  for (int i = threadIdx.x; i < nnx; i += blockDim.x) tw[i] = b.twx[i];
  const int iy = (p.ky == p.nn[1]) ? jy : (jy <= p.Ky ? jy : p.nn[1] - (p.ky - jy));
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
    const float* ff = b.lut_ff + ((size_t)f * p.E * b.lut_fstride + (size_t)e * p.lut_len);
#pragma unroll
    for (int u = 0; u < kOutMax; ++u) {
      const int idx = threadIdx.x + u * blockDim.x;
      if (idx < nout) {
        const int jx = idx / kc, k = idx - jx * kc;
        const int ix = (p.kx == nnx) ? jx : (jx <= p.Kx ? jx : nnx - (p.kx - jx));
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
#pragma unroll
  for (int u = 0; u < kOutMax; ++u) {
    const int idx = threadIdx.x + u * blockDim.x;
    if (idx < nout) {
      const int jx = idx / kc, k = idx - jx * kc;
      const int ix = (p.kx == nnx) ? jx : (jx <= p.Kx ? jx : nnx - (p.kx - jx));
      const int mx = ix < p.nf[0] ? ix : ix - nnx;
      const int kz = k0 + k;
      const int nsq = shell_key(mx, my, kz, p);
      if (nsq < p.lut_len) {
        const int bin = b.lut_bin[(size_t)f * b.lut_fstride + nsq];
        if (bin >= 0 && bin < nb) {
          const float mod = b.modx[ix] * b.mody[iy] * b.modz[kz];
          const float I = (A[u].x * A[u].x + A[u].y * A[u].y) * mod;
          const bool origin = (ix == 0) && (iy == 0) && (kz == 0);
          atomicAdd(&Hs[bin], origin ? (double)I : 2.0 * (double)I);
          atomicAdd(&Cs[bin], origin ? 1u : 2u);
        }
      }
    }
  }
  __syncthreads();
  double* out = b.partial + (((size_t)f * p.ky + jy) * p.nchx + ch) * 2 * nb;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    out[i] = Hs[i];
    out[nb + i] = (double)Cs[i];
  }
}

// Pruned x pass + epilogue for N = nn_x = 16 R2 (R2 = 8, 16, 32): workgroup
// (kept m_y, kz chunk, frame) with kc = p.kcx lines (one per kz), looping
// over the elements.  Stage 0 item (line k, j): inputs x = j + R2 r straight
// from ybuf (k fastest: a wavefront reads contiguous rows), radix-16 DFT
// (HALF: 2 n_x <= N, inputs r >= 8 zero), one LDS exchange.  Stage 1 item
// (line k, s): radix-R2 DFT over j of Y_j[s] W_N^{j s}; only the kept outputs
// m = s + 16 r' are formed (ENDS: Kx < N / 8, r' < KL or >= R2 - KL with
// KL = N / 128) and each thread accumulates A = sum_e f_e(q) F_e for its
// outputs in registers.  Then |A|^2 x moduli into the LDS histogram.
template <int N>
__host__ __device__ inline size_t xfft_p_lds_bytes(const Params& p) {
  return (((size_t)p.num_bins * 12 + 15) & ~size_t(15)) +
         (size_t)p.kcx * (17 * (N / 16) + 1) * sizeof(float2) + N * sizeof(float2);
}

constexpr int kXItems = 2;       // stage-1 items per thread (kc * 16 <= 512 threads * 2)

// Entry (u, o) of thread t of x-pass workgroup blk in the saved amplitude
// array: [blk][u][o][t] (coalesced along t); blk = (f ky + jy) nchx + ch.
// Only the item rounds u that hold items (u blockDim < kc 16) are stored.
__device__ __forceinline__ size_t amp_index(int blk, int u, int o, int NO) {
  return (((size_t)blk * kXItems + u) * NO + o) * blockDim.x + threadIdx.x;
}

// SAVE: the amplitudes A at the kept outputs also go to `amp` (amp_index),
// for an x adjoint that follows to read instead of recomputing them.
template <int N, bool ENDS, bool HALF, bool SAVE = false>
__global__ void __launch_bounds__(512) k_xfft_p(Params p, Buffers b, float2* amp = nullptr) {
  constexpr int R2 = N / 16;
  constexpr int LY = 17 * R2 + 1;
  constexpr int NO = ENDS ? 2 * (N / 128) : R2;     // kept outputs per item, at most
  extern __shared__ __align__(16) unsigned char smem[];
  const int jy = blockIdx.x, ch = blockIdx.y, f = blockIdx.z;
  const int kz = p.kz, k0 = ch * p.kcx, kc = min(p.kcx, kz - k0);
  const int nx = p.n[0], nb = p.num_bins, Kx = p.Kx;
  double* Hs = reinterpret_cast<double*>(smem);
  unsigned int* Cs = reinterpret_cast<unsigned int*>(Hs + nb);
  float2* lines = reinterpret_cast<float2*>(smem + (((size_t)nb * 12 + 15) & ~size_t(15)));
  float2* tw = lines + (size_t)p.kcx * LY;            // W_N^{j s} stored [j][s]
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    Hs[i] = 0.0;
    Cs[i] = 0u;
  }
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = b.twx[(i >> 4) * (i & 15)];
  const int iy = (p.ky == p.nn[1]) ? jy : (jy <= p.Ky ? jy : p.nn[1] - (p.ky - jy));
  const int my = iy < p.nf[1] ? iy : iy - p.nn[1];
  // this thread's stage-1 items and their kept outputs (same for every element)
  float2 A[kXItems][NO];
#pragma unroll
  for (int u = 0; u < kXItems; ++u)
#pragma unroll
    for (int o = 0; o < NO; ++o) A[u][o] = make_float2(0.0f, 0.0f);
  __syncthreads();
  for (int e = 0; e < p.E; ++e) {
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
    const float* ff = b.lut_ff + ((size_t)f * p.E * b.lut_fstride + (size_t)e * p.lut_len);
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
        const int mz = k0 + k;
#pragma unroll
        for (int o = 0; o < NO; ++o) {
          const int rr = ENDS ? (o < NO / 2 ? o : R2 - NO + o) : o;
          const int m = s + 16 * rr;
          if (m <= Kx || m >= N - Kx) {
            const int mx = m < p.nf[0] ? m : m - N;
            const int nsq = shell_key(mx, my, mz, p);
            if (nsq < p.lut_len) {
              const float fv = ff[nsq];
              A[u][o].x += fv * t[rr].x;
              A[u][o].y += fv * t[rr].y;
            }
          }
        }
      }
    }
    __syncthreads();
  }
  if constexpr (SAVE) {
    const int blk = (f * p.ky + jy) * p.nchx + ch;
#pragma unroll
    for (int u = 0; u < kXItems; ++u)
      if (u * (int)blockDim.x < kc * 16) {           // the same for the whole workgroup
#pragma unroll
        for (int o = 0; o < NO; ++o) amp[amp_index(blk, u, o, NO)] = A[u][o];
      }
  }
#pragma unroll
  for (int u = 0; u < kXItems; ++u) {
    const int item = threadIdx.x + u * blockDim.x;
    if (item < kc * 16) {
      const int s = item / kc, k = item - s * kc;
      const int mz = k0 + k;
#pragma unroll
      for (int o = 0; o < NO; ++o) {
        const int rr = ENDS ? (o < NO / 2 ? o : R2 - NO + o) : o;
        const int m = s + 16 * rr;
        if (m <= Kx || m >= N - Kx) {
          const int mx = m < p.nf[0] ? m : m - N;
          const int nsq = shell_key(mx, my, mz, p);
          if (nsq < p.lut_len) {
            const int bin = b.lut_bin[(size_t)f * b.lut_fstride + nsq];
            if (bin >= 0 && bin < nb) {
              const float mod = b.modx[m] * b.mody[iy] * b.modz[mz];
              const float I = (A[u][o].x * A[u][o].x + A[u][o].y * A[u][o].y) * mod;
              const bool origin = (m == 0) && (iy == 0) && (mz == 0);
              atomicAdd(&Hs[bin], origin ? (double)I : 2.0 * (double)I);
              atomicAdd(&Cs[bin], origin ? 1u : 2u);
            }
          }
        }
      }
    }
  }
  __syncthreads();
  double* out = b.partial + (((size_t)f * p.ky + jy) * p.nchx + ch) * 2 * nb;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    out[i] = Hs[i];
    out[nb + i] = (double)Cs[i];
  }
}

// ---- output-pruned x transform for any N = 2H R2 with n_x <= N/2, fused
// with the form-factor sum, moduli and histogram ----------------------------
//
// Same result and accounting as k_xfft_epi over the same (jy, z chunk,
// frame) workgroups: stage 0 (butterfly j < R2 of line k: inputs x = j + R2 r,
// r < H, two length-H DFTs) per element, then stage 1 only for the kept
// pairs m, N - m (m <= Kx) as sums over j by Horner's rule in W_N^m, each
// thread accumulating
// A = sum_e f_e(q) F_e for its pairs in registers; then |A|^2 x moduli into
// the LDS histogram (2 I, or I at the origin).  The plan guarantees
// Kx < N/8, kx < N and kcx (Kx + 1) <= kXgItems kWG.

constexpr int kXgItems = 4;          // stage-1 output pairs per thread

__host__ __device__ inline size_t xfft_g_lds_bytes(const Params& p, int R1) {
  const int N = p.nn[0], R2 = N / R1;
  return (((size_t)p.num_bins * 12 + 15) & ~size_t(15)) +
         (size_t)p.kcx * yg_line_stride(R1, R2) * sizeof(float2) + (size_t)N * sizeof(float2);
}

// Entry (u, o) of thread t of generic x-pass workgroup blk in the saved
// amplitudes: [blk][u][o][t] (k_xfft_g's pairs, o = 0 for m, 1 for N - m).
__device__ __forceinline__ size_t amp_g_index(int blk, int u, int o) {
  return (((size_t)blk * kXgItems + u) * 2 + o) * blockDim.x + threadIdx.x;
}

// SAVE: the accumulated pairs also go to `amp` (amp_g_index) for the x
// adjoint of a backward that follows.
template <int H, bool SAVE = false>
__global__ void __launch_bounds__(kWG) k_xfft_g(Params p, Buffers b, float2* amp = nullptr) {
  constexpr int R1 = 2 * H, YS = R1 + 1;
  extern __shared__ __align__(16) unsigned char smem[];
  const int jy = blockIdx.x, ch = blockIdx.y, f = blockIdx.z;
  const int kz = p.kz, k0 = ch * p.kcx, kc = min(p.kcx, kz - k0);
  const int nx = p.n[0], N = p.nn[0], nb = p.num_bins, Kx = p.Kx;
  const int R2 = N / R1, LY = yg_line_stride(R1, R2);
  double* Hs = reinterpret_cast<double*>(smem);
  unsigned int* Cs = reinterpret_cast<unsigned int*>(Hs + nb);
  float2* lines = reinterpret_cast<float2*>(smem + (((size_t)nb * 12 + 15) & ~size_t(15)));
  float2* tw = lines + (size_t)p.kcx * LY;           // W_N^m
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    Hs[i] = 0.0;
    Cs[i] = 0u;
  }
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = b.twx[i];
  const int iy = (p.ky == p.nn[1]) ? jy : (jy <= p.Ky ? jy : p.nn[1] - (p.ky - jy));
  const int my = iy < p.nf[1] ? iy : iy - p.nn[1];
  const float inv_kc = 1.0f / (float)kc;
  const int nitems = kc * (Kx + 1);                  // (pair m, line k), k fastest
  float2 A[kXgItems][2];
#pragma unroll
  for (int u = 0; u < kXgItems; ++u) A[u][0] = A[u][1] = make_float2(0.0f, 0.0f);
  __syncthreads();                                   // twiddles
  for (int e = 0; e < p.E; ++e) {
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
      for (int r = 1; r < H; ++r) xo[r] = cmul(xe[r], tw[r * R2]);   // W_R1^r = W_N^{r R2}
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
    const float* ff = b.lut_ff + ((size_t)f * p.E * b.lut_fstride + (size_t)e * p.lut_len);
#pragma unroll
    for (int u = 0; u < kXgItems; ++u) {
      const int item = threadIdx.x + u * blockDim.x;
      if (item < nitems) {
        const int m = div_small(item, kc, inv_kc), k = item - m * kc;
        const int s = m % R1, sm = s == 0 ? 0 : R1 - s;
        const float2* yl = lines + k * LY;
        float2 Z = make_float2(0.0f, 0.0f), Zm = make_float2(0.0f, 0.0f);
        // sum_j y_j W_N^{jm} by Horner's rule in wm = W_N^m from j = R2 - 1
        // down (conj wm for the mirrored output)
        horner_pair_c<YS>(yl + YS * (R2 - 1), s, sm, R2, tw[m], Z, Zm);
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
  if constexpr (SAVE) {
    const int blk = (f * p.ky + jy) * p.nchx + ch;
#pragma unroll
    for (int u = 0; u < kXgItems; ++u)
      if (u * (int)blockDim.x < nitems) {            // the same for the whole workgroup
        amp[amp_g_index(blk, u, 0)] = A[u][0];
        amp[amp_g_index(blk, u, 1)] = A[u][1];
      }
  }
#pragma unroll
  for (int u = 0; u < kXgItems; ++u) {
    const int item = threadIdx.x + u * blockDim.x;
    if (item < nitems) {
      const int m = div_small(item, kc, inv_kc), k = item - m * kc;
      const int mz = k0 + k;
#pragma unroll
      for (int o = 0; o < 2; ++o) {
        if (o == 1 && m == 0) continue;
        const int ix = o == 0 ? m : N - m, mx = o == 0 ? m : -m;
        const int nsq = shell_key(mx, my, mz, p);
        if (nsq < p.lut_len) {
          const int bin = b.lut_bin[(size_t)f * b.lut_fstride + nsq];
          if (bin >= 0 && bin < nb) {
            const float mod = b.modx[ix] * b.mody[iy] * b.modz[mz];
            const float I = (A[u][o].x * A[u][o].x + A[u][o].y * A[u][o].y) * mod;
            const bool origin = (ix == 0) && (iy == 0) && (mz == 0);
            atomicAdd(&Hs[bin], origin ? (double)I : 2.0 * (double)I);
            atomicAdd(&Cs[bin], origin ? 1u : 2u);
          }
        }
      }
    }
  }
  __syncthreads();
  double* out = b.partial + (((size_t)f * p.ky + jy) * p.nchx + ch) * 2 * nb;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    out[i] = Hs[i];
    out[nb + i] = (double)Cs[i];
  }
}

// ---- x transform for N = 2H R2 (R2 = 16), most outputs kept, fused
// with the form-factor sum, moduli and histogram -----------------------------
//
// Same workgroups, result and accounting as k_xfft_g (line = kz value k of
// the chunk, for the workgroup's jy): stage 0 per element as there; stage 1
// forms every output, item (residue s < 2H, line k) taking the radix-R2 DFT
// over j of T_j[s] W_N^{js}, which gives X(s + 2H r'), r' < R2; each thread
// accumulates A = sum_e f_e(q) F_e at its item's kept outputs (m <= Kx or
// m >= N - Kx) in registers: with KL = floor(Kx / 2H) + 1 <= KLC those are
// among r' < KLC and r' >= R2 - KLC, the 2 KLC outputs it keeps (the
// transform's other outputs are then never formed).  The plan guarantees
// kcx 2H <= kWG (one stage-1 item a thread), n_x <= N / 2 and 2 Kx + 1 < N.
// SAVE: the accumulators also go to `amp` ([workgroup][accumulator][thread],
// amp_gf_index) for the x adjoint of a backward that follows.
__host__ __device__ inline size_t xfft_gf_lds_bytes(const Params& p, int R2) {
  const int N = p.nn[0], R1 = N / R2;
  return (((size_t)p.num_bins * 12 + 15) & ~size_t(15)) +
         (size_t)p.kcx * yg_line_stride(R1, R2) * sizeof(float2) + (size_t)N * sizeof(float2);
}

__device__ __forceinline__ size_t amp_gf_index(int blk, int o, int NA) {
  return ((size_t)blk * NA + o) * blockDim.x + threadIdx.x;
}

template <int H, int R2, int KLC, bool SAVE = false>
__global__ void __launch_bounds__(kWG, 3) k_xfft_gf(Params p, Buffers b, float2* amp = nullptr) {
  static_assert(R2 == 16, "radix-16 stage 1");
  static_assert(2 * KLC <= R2, "kept outputs at the ends");
  constexpr int R1 = 2 * H, YS = R1 + 1, N = R1 * R2, NA = 2 * KLC;
  constexpr int LY = (R2 * YS) | 1;                  // yg_line_stride(R1, R2)
  extern __shared__ __align__(16) unsigned char smem[];
  const int jy = blockIdx.x, ch = blockIdx.y, f = blockIdx.z;
  const int kz = p.kz, k0 = ch * p.kcx, kc = min(p.kcx, kz - k0);
  const int nx = p.n[0], nb = p.num_bins, Kx = p.Kx;
  double* Hs = reinterpret_cast<double*>(smem);
  unsigned int* Cs = reinterpret_cast<unsigned int*>(Hs + nb);
  float2* lines = reinterpret_cast<float2*>(smem + (((size_t)nb * 12 + 15) & ~size_t(15)));
  float2* tw = lines + (size_t)p.kcx * LY;           // W_N^m
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    Hs[i] = 0.0;
    Cs[i] = 0u;
  }
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = b.twx[i];
  const int iy = (p.ky == p.nn[1]) ? jy : (jy <= p.Ky ? jy : p.nn[1] - (p.ky - jy));
  const int my = iy < p.nf[1] ? iy : iy - p.nn[1];
  const float inv_kc = 1.0f / (float)kc;
  // this thread's stage-1 item (residue s, line k; k fastest); the shell key
  // of its output rr (lut_len where not kept) is formed where it is used
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
  for (int e = 0; e < p.E; ++e) {
    // stage 0 (item = butterfly j, line k; k fastest)
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
      for (int r = 1; r < H; ++r) xo[r] = cmul(xe[r], tw[r * R2]);   // W_R1^r = W_N^{r R2}
      Radix<H>::dft(xe);
      Radix<H>::dft(xo);
      float2* dst = lines + kk0 * LY + YS * j0;
#pragma unroll
      for (int t = 0; t < H; ++t) {
        dst[2 * t] = xe[t];
        dst[2 * t + 1] = xo[t];
      }
    }
    __syncthreads();                                 // lines complete
    if (act1) {
      const float* ff = b.lut_ff + ((size_t)f * p.E * b.lut_fstride + (size_t)e * p.lut_len);
      const float2* yl = lines + k1 * LY + s1;
      float2 t[R2];
      t[0] = yl[0];
#pragma unroll
      for (int j = 1; j < R2; ++j) t[j] = cmul(yl[YS * j], tw[j * s1]);   // j s < N
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
    __syncthreads();                                 // lines are rewritten
  }
  if constexpr (SAVE) {
    const int blk = (f * p.ky + jy) * p.nchx + ch;
#pragma unroll
    for (int o = 0; o < NA; ++o) amp[amp_gf_index(blk, o, NA)] = A[o];
  }
  if (act1) {
#pragma unroll
    for (int o = 0; o < NA; ++o) {
      const int rr = out_of(o);
      const int nsq = key(rr);
      if (nsq < p.lut_len) {
        const int ix = s1 + R1 * rr;
        const int bin = b.lut_bin[(size_t)f * b.lut_fstride + nsq];
        if (bin >= 0 && bin < nb) {
          const float mod = b.modx[ix] * b.mody[iy] * b.modz[mz];
          const float I = (A[o].x * A[o].x + A[o].y * A[o].y) * mod;
          const bool origin = (ix == 0) && (iy == 0) && (mz == 0);
          atomicAdd(&Hs[bin], origin ? (double)I : 2.0 * (double)I);
          atomicAdd(&Cs[bin], origin ? 1u : 2u);
        }
      }
    }
  }
  __syncthreads();
  double* out = b.partial + (((size_t)f * p.ky + jy) * p.nchx + ch) * 2 * nb;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    out[i] = Hs[i];
    out[nb + i] = (double)Cs[i];
  }
}

// Sums the per-block histograms of each frame (or of all frames) in a fixed
// order.  grid = (num_bins, frames_out); per_frame selects overwrite of
// Hf[f][b] vs accumulation into H[b].
__global__ void __launch_bounds__(kWG) k_hist_reduce(const double* __restrict__ partial, int nblk,
                                                     int frames_per_out, int nb, double* H,
                                                     double* C, int per_frame) {
  __shared__ double sh[kWG], sc[kWG];
  const int bin = blockIdx.x, fo = blockIdx.y;
  const size_t first = (size_t)fo * frames_per_out * nblk;
  const size_t count = (size_t)frames_per_out * nblk;
  double h = 0.0, c = 0.0;
  for (size_t i = threadIdx.x; i < count; i += blockDim.x) {
    const double* row = partial + (first + i) * 2 * nb;
    h += row[bin];
    c += row[nb + bin];
  }
  sh[threadIdx.x] = h;
  sc[threadIdx.x] = c;
  __syncthreads();
  for (int s = kWG / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sh[threadIdx.x] += sh[threadIdx.x + s];
      sc[threadIdx.x] += sc[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (per_frame) {
      H[(size_t)fo * nb + bin] = sh[0];
      C[(size_t)fo * nb + bin] = sc[0];
    } else {
      H[bin] += sh[0];
      C[bin] += sc[0];
    }
  }
}

}  // namespace saxs
