// C ABI of the MI355X SAXS engine: plan creation, workspace, launches.
// The declarations and the reference interfaces each call replaces are in
// include/saxs_engine.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "saxs_engine.h"
#include "saxs_kernels.hpp"
#include "saxs_adjoint.hpp"
#include "saxs_plain.hpp"
#include "saxs_tables.hpp"

// Device scratch of saxs_forward_static, the entry point that allocates
// nothing: every per-call buffer of the forward chain is carved from this
// array.  Each GPU holds its own copy from the library's load until the
// process ends.  In host code the name is only a symbol handle: the address
// of a device's copy comes from hipGetSymbolAddress (static_scratch below).
constexpr size_t kStaticScratchBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char msg[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(msg, sizeof(msg), fmt, ap);
  va_end(ap);
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(SAXS_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_));              \
  } while (0)

constexpr size_t kLdsLimit = 64 * 1024;

// Stockham stage radices of N.  Default: radix 8/4/2 for the power of two,
// then 3, 5, 7.  wide: greedy with the widest register radix first (12, 10,
// 9, 8, 7, 6, 5, 4, 3, 2), fewer stages and LDS passes for sizes with
// several odd factors, at fewer lines per call.
bool factor_radices(int N, saxs::FftDesc& d, bool wide = false) {
  d.N = N;
  d.nst = 0;
  if (wide) {
    int m = N, nst = 0;
    while (m > 1) {
      int r = 0;
      for (int c : {12, 10, 9, 8, 7, 6, 5, 4, 3, 2})
        if (m % c == 0) {
          r = c;
          break;
        }
      if (r == 0 || nst == saxs::kMaxStages) return false;
      d.radix[nst++] = r;
      m /= r;
    }
    if (nst == 0) d.radix[nst++] = 1;  // N == 1 is rejected by the caller
    d.nst = nst;
    return true;
  }
  int m = N, twos = 0;
  while (m % 2 == 0) {
    m /= 2;
    ++twos;
  }
  std::vector<int> rs;
  int eights = twos / 3, rem = twos % 3;
  if (rem == 1 && eights > 0) {
    --eights;
    rs.push_back(4);
    rs.push_back(4);
  } else if (rem == 1) {
    rs.push_back(2);
  } else if (rem == 2) {
    rs.push_back(4);
  }
  for (int i = 0; i < eights; ++i) rs.insert(rs.begin(), 8);
  for (int r : {3, 5, 7})
    while (m % r == 0) {
      m /= r;
      rs.push_back(r);
    }
  if (m != 1 || (int)rs.size() > saxs::kMaxStages) return false;
  if (rs.empty()) rs.push_back(1);  // N == 1 is rejected by the caller
  for (size_t i = 0; i < rs.size(); ++i) d.radix[i] = rs[i];
  d.nst = (int)rs.size();
  return true;
}

// Largest number of lines of length N one fft_lines call may hold.
int fft_line_cap(const saxs::FftDesc& d) {
  int per_thread = saxs::kMaxC;
  for (int s = 0; s < d.nst; ++s) per_thread = std::min(per_thread, (saxs::kMaxC / d.radix[s]) * d.radix[s]);
  return per_thread * saxs::kWG / d.N;
}

int lds_stride(int N) { return (N % 16 == 0) ? N + 1 : N; }

std::vector<float2> twiddles(int N) {
  std::vector<float2> t(N);
  for (int i = 0; i < N; ++i) {
    const double a = -2.0 * M_PI * (double)i / (double)N;
    t[i] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  return t;
}

// For each column tile of one axis of n cells in tiles of t: the key units
// (home tile S of the stencil's first point, times C, plus its class c; see
// saxs::frame_keys) whose atoms reach the tile, found by enumerating every
// stencil base.  x axis: the units as at most two runs of consecutive units
// {runs, first unit of run 0, of run 1, len0 | len1 << 16}; y axis: at most
// three units {count, u0, u1, u2}.  False when a tile needs more.
bool tile_units(int n, int t, int P, int C, bool x_axis, std::vector<int4>& out) {
  const int nt = (n + t - 1) / t;
  std::vector<std::vector<int>> units(nt);
  for (int base = 0; base < n; ++base) {
    const int S = base / t, X0 = S * t, w = std::min(t, n - X0);
    const int u = C == 1 ? S : 2 * S + (base - X0 + P - 1 >= w ? 1 : 0);
    for (int o = 0; o < P; ++o) {
      auto& v = units[((base + o) % n) / t];
      if (std::find(v.begin(), v.end(), u) == v.end()) v.push_back(u);
    }
  }
  out.assign(nt, make_int4(0, 0, 0, 0));
  for (int T = 0; T < nt; ++T) {
    std::vector<int>& v = units[T];
    std::sort(v.begin(), v.end());
    if (!x_axis) {
      if (v.size() > 3) return false;
      int4 r = make_int4((int)v.size(), 0, 0, 0);
      if (v.size() > 0) r.y = v[0];
      if (v.size() > 1) r.z = v[1];
      if (v.size() > 2) r.w = v[2];
      out[T] = r;
      continue;
    }
    std::vector<std::pair<int, int>> runs;            // (first unit, length)
    for (int u : v) {
      if (!runs.empty() && runs.back().first + runs.back().second == u)
        ++runs.back().second;
      else
        runs.push_back({u, 1});
    }
    if (runs.empty() || runs.size() > 2) return false;
    int4 r = make_int4((int)runs.size(), runs[0].first, 0, runs[0].second);
    if (runs.size() > 1) {
      r.z = runs[1].first;
      r.w |= runs[1].second << 16;
    }
    out[T] = r;
  }
  return true;
}

// Passes of the segmented counting sort over a frame's keys: 1 with 32-bit
// LDS counters, else the number of 16-bit passes over key ranges (at most
// 4), else a large number (the global-atomic sort).
constexpr size_t kSortLdsBudget = 48 * 1024;
int seg_kr16(int E) {
  int kr = 2 * (int)((kSortLdsBudget - 16 - E * sizeof(unsigned long long)) / sizeof(unsigned));
  while (kr > 2 && saxs::prep_seg_lds_bytes(kr, E, true) > kSortLdsBudget) kr -= 2;
  return kr;
}
int sort_passes(int kframe, int E) {
  if (saxs::prep_seg_lds_bytes(kframe, E) <= kSortLdsBudget) return 1;
  const int passes = (kframe + seg_kr16(E) - 1) / seg_kr16(E);
  return passes <= 4 ? passes : 1000;
}

// Largest kept |m| per axis: from the point-key box, or K with K^2 <= lut_len - 1
// for shell tables (shells n = |m|^2 < lut_len).
void kept_max(const saxs::Params& p, int lut_len, int K[3]) {
  if (p.point_keys) {
    for (int a = 0; a < 3; ++a) K[a] = p.kmax[a];
    return;
  }
  int k = 0;
  while ((long long)(k + 1) * (k + 1) < (long long)lut_len) ++k;
  K[0] = K[1] = K[2] = k;
}

// kz for tables of lut_len entries (as derive_ranges computes it).
int kz_of(const saxs::Params& p, int lut_len) {
  int K[3];
  kept_max(p, lut_len, K);
  return std::min(K[2] + 1, p.nn[2] / 2 + 1);
}

// Whether the radix-16 output-pruned spreading kernel (k_spread_zp) applies.
bool zp_applies(bool spread_all, bool pruned_z, const saxs::Params& p) {
  const int Nz = p.nn[2], nl = p.tx * p.ty / 2;
  return spread_all && pruned_z && (Nz == 128 || Nz == 256 || Nz == 512) &&
         2 * p.kz - 1 <= Nz / 4 && p.kz <= Nz / 8 && nl * (Nz / 16) <= saxs::kWG &&
         saxs::spread_p_lds_bytes(p) <= kLdsLimit;
}

// Half-length H of the first stage of the generic output-pruned spreading
// kernel (k_spread_zg, N = 2H R2), or 0 where it does not apply: it needs
// n_z <= N/2, an even R1 = 2H dividing N with order <= R2 <= 64 (a
// wavefront transforms whole lines), and the LDS.  H = 8, 7, 6, 5 come
// before 9 and 10: the length-9 and length-10 register DFTs need about 40
// more VGPRs, which costs a wave per SIMD and more time than the shorter
// stage-1 sums save.
int zg_half(bool spread_all, bool pruned_z, const saxs::Params& p) {
  const int Nz = p.nn[2];
  if (!spread_all || !pruned_z || 2 * p.n[2] > Nz) return 0;
  for (int H : {8, 7, 6, 5, 9, 10}) {
    const int R1 = 2 * H;
    if (Nz % R1 != 0 || Nz / R1 < p.order || Nz / R1 > 64) continue;
    if (saxs::spread_g_lds_bytes(p, R1) > kLdsLimit) continue;
    return H;
  }
  return 0;
}

// Half-length H of the generic output-pruned z adjoint (k_adj_zg, N = 2H R2),
// or 0 where it does not apply: n_z <= N/2, kz <= kAdjZgMaxKz, an even
// R1 = 2H dividing N with order <= R2, and the LDS (same preference as
// zg_half).
int adjg_half(bool spread_all, bool pruned_z, const saxs::Params& p) {
  const int Nz = p.nn[2];
  if (!spread_all || !pruned_z || 2 * p.n[2] > Nz || p.kz > saxs::kAdjZgMaxKz) return 0;
  if (saxs::adj_g_lds_bytes(p) > kLdsLimit) return 0;
  for (int H : {8, 7, 6, 5, 9, 10})
    if (Nz % (2 * H) == 0 && Nz / (2 * H) >= p.order) return H;
  return 0;
}

// Half-length H of the first stage of the generic output-pruned y transform
// (k_yfft_g, N = 2H R2), or 0 where it does not apply: it needs n_y <= N/2,
// not every output kept (ky < N), an even R1 = 2H dividing N with R2 <= 64,
// and the LDS (the same order of preference as zg_half).  Its stage-1
// direct sums cost about twice a full transform's arithmetic when half the
// outputs are kept (the reference CLI's q <= 2 on a 320-point axis), and in
// one LDS exchange it still beats the multi-stage Stockham kernel there
// (2.37 against 3.16 ms per 50 frames of 128^3 -> 320^3).
int yg_half(bool pruned_z, const saxs::Params& p) {
  const int Ny = p.nn[1];
  if (!pruned_z || 2 * p.n[1] > Ny || p.ky >= Ny) return 0;
  for (int H : {8, 7, 6, 5, 9, 10}) {
    const int R1 = 2 * H;
    if (Ny % R1 != 0 || Ny / R1 > 64) continue;
    if (saxs::yfft_g_lds_bytes(Ny, R1) > kLdsLimit) continue;
    return H;
  }
  return 0;
}

// Half-length H of the generic y transform whose stage 1 forms every output
// (k_yfft_gf<H, R2>, N = 2H R2 with R2 = 16 or 32), or 0 where it does not
// apply: what yg_half needs, and most outputs kept (Ky + 1 >= N / 8), where
// 2H radix-R2 DFTs a line cost less than the direct sums.
int yf_half(bool pruned_z, const saxs::Params& p, int& R2) {
  const int Ny = p.nn[1];
  if (!pruned_z || 2 * p.n[1] > Ny || p.ky >= Ny || 8 * (p.Ky + 1) < Ny) return 0;
  for (int r2 : {16, 32}) {
    if (Ny % (2 * r2) != 0) continue;
    const int H = Ny / (2 * r2);
    if (H < 5 || H > 10) continue;
    if (saxs::yfft_gf_lds_bytes(Ny, r2) > kLdsLimit) continue;
    R2 = r2;
    return H;
  }
  return 0;
}

// Whether the spreading kernel and the z adjoint that will run for column
// tile (p.tx, p.ty) fit: the pruned kernels where they apply, else the
// all-element or per-element Stockham ones (the launch picks in that order).
bool tile_fits(bool spread_all, bool pruned_z, saxs::Params p, int lut_len, bool fwd_only) {
  p.kz = kz_of(p, lut_len);
  const int Nz = p.nn[2], nl = p.tx * p.ty / 2;
  const bool line_ok = nl <= fft_line_cap(p.fz);
  const bool pruned = spread_all && pruned_z && (Nz == 128 || Nz == 256 || Nz == 512) &&
                      2 * p.kz - 1 <= Nz / 4 && p.kz <= Nz / 8 && nl * (Nz / 16) <= saxs::kWG;
  const bool fwd = (pruned && saxs::spread_p_lds_bytes(p) <= kLdsLimit) ||
                   zg_half(spread_all, pruned_z, p) > 0 ||
                   (line_ok && spread_all && saxs::spread_all_lds_bytes(p) <= kLdsLimit) ||
                   (line_ok && saxs::spread_lds_bytes(p) <= kLdsLimit);
  saxs::Params q = p;                      // room for one more kept z frequency
  q.kz = std::min(p.kz + 1, p.nn[2] / 2 + 1);
  const bool bwd = (pruned && saxs::adj_p_lds_bytes(p) <= kLdsLimit) ||
                   adjg_half(spread_all, pruned_z, q) > 0 ||
                   (line_ok && saxs::adjz_lds_bytes(p) <= kLdsLimit);
  return fwd && (bwd || fwd_only);
}

// Number of tile offsets (tile - home tile, mod the tile count) that one
// stencil base + [0, P) reaches along an axis of n cells in tiles of t; the
// home tile holds the stencil's first point.
int slot_count(int n, int t, int P) {
  const int nt = (n + t - 1) / t;
  int m = 1;
  for (int base = 0; base < n; ++base)
    for (int o = 0; o < P; ++o) {
      const int T = ((base + o) % n) / t;
      m = std::max(m, (T - base / t + nt) % nt + 1);
    }
  return m;
}

template <typename T>
int dev_alloc(T** p, size_t count, size_t& total) {
  *p = nullptr;
  if (count == 0) return SAXS_OK;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(SAXS_ENOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
  }
  total += count * sizeof(T);
  return SAXS_OK;
}

template <typename T>
int upload(T* dst, const T* src, size_t count) {
  if (count == 0) return SAXS_OK;
  HIP_TRY(hipMemcpy(dst, src, count * sizeof(T), hipMemcpyHostToDevice));
  return SAXS_OK;
}

}  // namespace

struct KernelTimer {
  bool on = false;
  std::vector<hipEvent_t> pool;                 // free events
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> rec;
  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
};

struct saxs_plan {
  KernelTimer timer;
  // state of the last saxs_forward_frames (for SAXS_BWD_REUSE_FORWARD)
  bool last_frames_valid = false;
  const float* last_xyz = nullptr;
  int last_N = 0, last_F = 0;
  bool spread_all = true;      // one workgroup per tile for all elements
  bool pruned_z = true;        // output-pruned radix-16 z transform where it applies
  saxs::Params p{};            // geometry; F/N/oc set per call
  saxs_desc desc{};
  int max_frames = 0, max_atoms = 0, nkeys_max = 0;
  int lut_cap = 0;             // q-table entries the tables were sized for
  int kz_cap = 0, ky_cap = 0, nchx_cap = 0;  // kept ranges the workspace was sized for
  size_t bytes = 0;
  // tables
  int4* srcx = nullptr;
  int4* srcy = nullptr;
  float2* twz = nullptr;
  float2* twy = nullptr;
  float2* twx = nullptr;
  float* modx = nullptr;
  float* mody = nullptr;
  float* modz = nullptr;
  int* lut_bin = nullptr;
  float* lut_ff = nullptr;
  // workspace
  int* counts = nullptr;
  int* offsets = nullptr;
  int2* keyslot = nullptr;
  float4* sorted = nullptr;
  unsigned long long* pad_sum = nullptr;
  unsigned long long* pad_given = nullptr;  // [E] padding::given pads, 2^-32 units
  std::vector<double> pad_given_host;       // empty: border-slab mean (padding::avg)
  float2* zbuf = nullptr;
  float2* ybuf = nullptr;
  double* partial = nullptr;
  double* hframe = nullptr;    // [max_frames][2*num_bins] scratch for gradients
  float* grid = nullptr;       // [E][n^3] adjoint density / plain-path grid
  float* sup = nullptr;        // [nn^3] plain-path embedded grid
  float2* spec = nullptr;      // [nn_x][nn_y][h] plain-path spectrum (z pass)
  float2* yspec = nullptr;     // [nn_y][nn_x][h] plain-path spectrum (y pass)
  float2* acc = nullptr;       // [nn_x][nn_y][h] plain-path accumulator
  double* padp = nullptr;      // [1] plain-path border sum
  saxs::Params pp{};           // plain-path geometry (the whole FFT grid as "cell")
  int plain_pairs = 0;
  double* padg = nullptr;      // [F*E] adjoint pad terms
  double* padt = nullptr;      // [F*E*ntx*nty] adjoint pad terms per column tile
  float* gpart = nullptr;      // [F][nsx*nsy][N][3] adjoint partials (first backward)
  int* segoff = nullptr;       // [(512 + max_frames) * keys per frame] segmented sort
  bool seg_sort = false;       // one frame's keys fit the LDS histogram of k_prep_seg
  bool seg16 = false;          // ... with 16-bit counters (segments <= kSeg16MaxLen atoms)
  int seg_kr = 0;              // keys per k_prep_seg pass (a frame's keys in <= 4 passes)
  bool x_pruned = false;       // x pass and its adjoint by k_xfft_p / k_adj_x_p
  bool fwd_only = false;       // created with SAXS_FLAG_FORWARD_ONLY
  float2* xamp = nullptr;      // x-pass amplitudes saved by saxs_forward_frames (k_xfft_p)
  size_t xamp_elems = 0;
  bool xamp_valid = false;     // xamp holds the last saxs_forward_frames' amplitudes
  int x_gen_h = 0;             // x pass by k_xfft_g<x_gen_h>, its adjoint by k_adj_x_g (0: Stockham)
  int x_full_h = 0, x_full_r2 = 0;  // forward x pass by k_xfft_gf<x_full_h, x_full_r2> (0: none)
  int nsx = 1, nsy = 1;        // tile offsets one stencil reaches along x, y
  // saxs_forward_boxes: per-frame cells, q geometries, coefficients and
  // tables (allocated on first use), and the pinned staging of the uploads
  float* fb_cells = nullptr;   // [max_frames][9]
  float* fb_geom = nullptr;    // [max_frames + 1][9] (the last row: binning geometry)
  float* fb_coef = nullptr;    // [E][9]
  int* fb_bin = nullptr;       // [max_frames][lut_cap]
  float* fb_ff = nullptr;      // [max_frames][E][lut_cap]
  float* fb_host = nullptr;    // pinned [max_frames * 18 + 9 + 9 E]
  hipEvent_t fb_copied = nullptr;
  bool fb_pending = false;
};

namespace {

// Kept frequencies and pass chunking from lut_len (shells n < lut_len).
int derive_ranges(saxs_plan* pl, int lut_len) {
  saxs::Params& p = pl->p;
  p.lut_len = lut_len;
  int K[3];
  kept_max(p, lut_len, K);
  auto kept = [&](int N, int Ka, int& k, int& Kout) {
    if (2 * Ka + 1 >= N) {
      k = N;
      Kout = N / 2;
    } else {
      k = 2 * Ka + 1;
      Kout = Ka;
    }
  };
  kept(p.nn[0], K[0], p.kx, p.Kx);
  kept(p.nn[1], K[1], p.ky, p.Ky);
  const int h = p.nn[2] / 2 + 1;
  p.kz = std::min(K[2] + 1, h);

  const int capy = fft_line_cap(p.fy), capx = fft_line_cap(p.fx);
  const size_t budget = 48 * 1024;
  int kcy = std::min({p.kz, capy, (int)(budget / (sizeof(float2) * p.lsy))});
  int kcx = std::min({p.kz, capx, (int)(budget / (sizeof(float2) * p.lsx)),
                      saxs::kOutMax * saxs::kWG / p.kx});
  if (kcy < 1 || kcx < 1) return fail(SAXS_EUNSUP, "FFT lines do not fit the LDS budget");
  p.nchy = (p.kz + kcy - 1) / kcy;
  p.kcy = (p.kz + p.nchy - 1) / p.nchy;
  p.nchx = (p.kz + kcx - 1) / kcx;
  p.kcx = (p.kz + p.nchx - 1) / p.nchx;
  // pruned x pass: kept m_x at the ends (Kx < N/8), one thread per stage-0
  // item (kc * N/16 <= 512), LDS for kc lines, twiddles and the histogram
  const int Nx = p.nn[0];
  pl->x_pruned = false;
  if (pl->pruned_z && (Nx == 128 || Nx == 256 || Nx == 512) && p.Kx < Nx / 8 && p.kx < Nx) {
    const int ly = 17 * (Nx / 16) + 1;
    const long long room = (long long)kLdsLimit - (((long long)p.num_bins * 12 + 15) & ~15LL) -
                           (long long)Nx * (long long)sizeof(float2);
    int kc = std::min({p.kz, 512 / (Nx / 16), (int)(room / (ly * (long long)sizeof(float2)))});
    if (kc < p.kz && kc >= 8) kc = kc / 8 * 8;
    if (kc >= 1) {
      pl->x_pruned = true;
      p.nchx = (p.kz + kc - 1) / kc;
      p.kcx = (p.kz + p.nchx - 1) / p.nchx;
    }
  }
  // generic pruned x pass and its adjoint (over the same z chunks): n_x <=
  // N/2, not every m_x kept, an even R1 = 2H dividing N, every kept pair of
  // a chunk in kXgItems per thread (at the reference CLI's q <= 2 on a
  // 320-point axis 3.04 against 4.31 ms per 50 frames for the Stockham pass)
  pl->x_gen_h = 0;
  if (pl->pruned_z && !pl->x_pruned && 2 * p.n[0] <= Nx && p.kx < Nx &&
      p.kcx * (p.Kx + 1) <= saxs::kXgItems * saxs::kWG) {
    for (int H : {8, 7, 6, 5, 9, 10}) {
      if (Nx % (2 * H) != 0 || Nx / (2 * H) > 64) continue;
      if (saxs::xfft_g_lds_bytes(p, 2 * H) > kLdsLimit) continue;
      pl->x_gen_h = H;
      break;
    }
  }
  // ... whose forward forms every stage-1 output by radix-16 DFTs where
  // most outputs are kept (N = 2H 16; one stage-1 item a thread over the
  // same z chunks)
  pl->x_full_h = pl->x_full_r2 = 0;
  if (pl->x_gen_h && 8 * (p.Kx + 1) >= Nx) {
    for (int r2 : {16}) {      // radix 32: 32 accumulated outputs a thread exceed the registers
      if (Nx % (2 * r2) != 0) continue;
      const int H = Nx / (2 * r2);
      if (H < 5 || H > 10 || p.kcx * 2 * H > saxs::kWG) continue;
      if (saxs::xfft_gf_lds_bytes(p, r2) > kLdsLimit) continue;
      pl->x_full_h = H;
      pl->x_full_r2 = r2;
      break;
    }
  }
  if (saxs::yfft_lds_bytes(p) > kLdsLimit || (!pl->x_pruned && saxs::xfft_lds_bytes(p) > kLdsLimit))
    return fail(SAXS_EUNSUP, "pass LDS exceeds 64 KiB (num_bins %d too large?)", p.num_bins);
  return SAXS_OK;
}

// A point-key box |m_a| <= kmax[a] must have lut_len = prod (kmax[a] + 1)
// entries, or prod (2 kmax[a] + 1) with signed keys.
int check_point_box(const int32_t* kmax, int32_t lut_len, bool signed_keys) {
  long long cells = 1;
  for (int a = 0; a < 3; ++a) {
    if (kmax[a] < 0 || kmax[a] > 4096) return fail(SAXS_EINVAL, "kmax[%d] = %d invalid", a, kmax[a]);
    cells *= signed_keys ? 2LL * kmax[a] + 1 : kmax[a] + 1;
  }
  if (cells != lut_len)
    return fail(SAXS_EINVAL, "lut_len %d != %lld entries of the kmax box", lut_len, cells);
  return SAXS_OK;
}

// Whether a row-major 3x3 inverse cell has off-diagonal terms.
bool is_triclinic(const float* cell) {
  return cell && (cell[1] != 0.0f || cell[2] != 0.0f || cell[3] != 0.0f || cell[5] != 0.0f ||
                  cell[6] != 0.0f || cell[7] != 0.0f);
}

void set_cell(saxs::Params& p, const float* cell) {
  p.tric = is_triclinic(cell) ? 1 : 0;
  for (int i = 0; i < 9; ++i) p.cm[i] = p.tric ? cell[i] : 0.0f;
}

// Replaces the q tables: shell keys (kmax NULL) or point keys, signed for a
// triclinic cell.
int set_tables(saxs_plan* pl, const float* cell, const int32_t* kmax, int32_t lut_len,
               const int32_t* bin_lut, const float* ff_lut) {
  if (!pl || !bin_lut || !ff_lut) return fail(SAXS_EINVAL, "NULL argument");
  if (lut_len < 1) return fail(SAXS_EINVAL, "lut_len %d < 1", lut_len);
  if (lut_len > pl->lut_cap)
    return fail(SAXS_ECAPACITY, "lut_len %d exceeds the plan's %d (create a plan for this box)",
                lut_len, pl->lut_cap);
  saxs::Params saved = pl->p;
  const bool tric = is_triclinic(cell);
  pl->p.point_keys = kmax ? (tric ? 2 : 1) : 0;
  for (int a = 0; a < 3; ++a) pl->p.kmax[a] = kmax ? kmax[a] : 0;
  set_cell(pl->p, cell);
  int rc = derive_ranges(pl, lut_len);
  if (!rc && (pl->p.kz > pl->kz_cap || pl->p.ky > pl->ky_cap || pl->p.nchx > pl->nchx_cap))
    rc = fail(SAXS_ECAPACITY, "new tables keep kz %d, ky %d, %d x chunks; the workspace holds "
              "%d, %d, %d (create a plan for this box)", pl->p.kz, pl->p.ky, pl->p.nchx,
              pl->kz_cap, pl->ky_cap, pl->nchx_cap);
  if (rc) {
    pl->p = saved;
    (void)derive_ranges(pl, saved.lut_len);          // restores the pass flags
    return rc;
  }
  pl->last_frames_valid = false;
  HIP_TRY(hipMemcpy(pl->lut_bin, bin_lut, sizeof(int32_t) * lut_len, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(pl->lut_ff, ff_lut, sizeof(float) * lut_len * pl->p.E, hipMemcpyHostToDevice));
  return SAXS_OK;
}

// Brackets one launch with events when the timing hook is on.
struct Probe {
  saxs_plan* pl;
  int id;
  hipStream_t st;
  hipEvent_t a = nullptr;
  Probe(saxs_plan* pl_, int id_, hipStream_t st_) : pl(pl_), id(id_), st(st_) {
    if (pl->timer.on && (a = pl->timer.get())) (void)hipEventRecord(a, st);
  }
  ~Probe() {
    if (!a) return;
    hipEvent_t b = pl->timer.get();
    if (!b) return;
    (void)hipEventRecord(b, st);
    pl->timer.rec.push_back({id, {a, b}});
  }
};

void free_plan(saxs_plan* pl) {
  for (auto& r : pl->timer.rec) {
    (void)hipEventDestroy(r.second.first);
    (void)hipEventDestroy(r.second.second);
  }
  for (hipEvent_t e : pl->timer.pool) (void)hipEventDestroy(e);
  if (pl->fb_copied) (void)hipEventDestroy(pl->fb_copied);
  if (pl->fb_host) (void)hipHostFree(pl->fb_host);
  void* fbp[] = {pl->fb_cells, pl->fb_geom, pl->fb_coef, pl->fb_bin, pl->fb_ff};
  for (void* q : fbp)
    if (q) (void)hipFree(q);
  void* ptrs[] = {pl->srcx, pl->srcy, pl->twz, pl->twy, pl->twx, pl->modx, pl->mody,
                  pl->modz, pl->lut_bin, pl->lut_ff, pl->counts, pl->offsets, pl->keyslot,
                  pl->sorted, pl->pad_sum, pl->zbuf, pl->ybuf, pl->partial, pl->hframe,
                  pl->grid, pl->sup, pl->spec, pl->acc, pl->padg, pl->padt, pl->yspec, pl->padp,
                  pl->gpart, pl->segoff, pl->pad_given, pl->xamp};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  delete pl;
}

// Per-call scratch of the forward chain: the plan's workspace, or carved
// from the static scratch array by saxs_forward_static.
struct Work {
  int* counts;
  int* offsets;
  int2* keyslot;
  float4* sorted;
  unsigned long long* pad_sum;
  int* segoff;
  float2* zbuf;
  float2* ybuf;
  double* partial;
  // per-frame geometry of a batch of boxes (saxs_forward_boxes), else NULL / 0
  const float* cells = nullptr;
  const int* lut_bin = nullptr;
  const float* lut_ff = nullptr;
  int lut_fstride = 0;
};

Work plan_work(const saxs_plan* pl) {
  return Work{pl->counts, pl->offsets, pl->keyslot, pl->sorted, pl->pad_sum,
              pl->segoff, pl->zbuf,    pl->ybuf,    pl->partial};
}

// Address of the current device's copy of `scratch`, looked up once per
// device (with that device current) and kept.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int static_scratch(void** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(SAXS_EUNSUP, "device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(SAXS_EHIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                  hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return SAXS_OK;
}

// Carves the forward chain's buffers for n_atoms x n_frames of the plan's
// current geometry out of `base` (256-byte aligned pieces); returns the
// bytes needed (base may be NULL to only measure).
size_t carve_work(const saxs_plan* pl, int n_atoms, int n_frames, unsigned char* base, Work& w) {
  const saxs::Params& p = pl->p;
  const size_t F = n_frames, N = n_atoms, E = p.E;
  const size_t kframe = saxs::frame_keys(p), nkeys = F * kframe;
  size_t off = 0;
  auto take = [&](size_t bytes) -> void* {
    void* q = base ? (void*)(base + off) : nullptr;
    off += (bytes + 255) & ~size_t(255);
    return q;
  };
  w.counts = (int*)take((nkeys + 1) * sizeof(int));
  w.offsets = (int*)take((nkeys + 1) * sizeof(int));
  w.keyslot = (int2*)take(F * N * sizeof(int2));
  w.sorted = (float4*)take(F * N * sizeof(float4));
  w.pad_sum = (unsigned long long*)take(F * E * sizeof(unsigned long long));
  w.segoff = pl->seg_sort ? (int*)take((512 + F) * kframe * sizeof(int)) : nullptr;
  w.zbuf = (float2*)take(F * E * p.n[0] * p.n[1] * p.kz * sizeof(float2));
  w.ybuf = (float2*)take(F * E * p.ky * p.n[0] * p.kz * sizeof(float2));
  w.partial = (double*)take(F * p.ky * p.nchx * 2 * p.num_bins * sizeof(double));
  return off;
}

saxs::Buffers make_buffers(saxs_plan* pl, const float* xyz, const int32_t* elem,
                           const Work* w = nullptr) {
  const Work W = w ? *w : plan_work(pl);
  saxs::Buffers b{};
  b.xyz = xyz;
  b.elem = elem;
  b.counts = W.counts;
  b.offsets = W.offsets;
  b.keyslot = W.keyslot;
  b.sorted = W.sorted;
  b.pad_sum = W.pad_sum;
  b.pad_given = pl->pad_given_host.empty() ? nullptr : pl->pad_given;
  b.srcx = pl->srcx;
  b.srcy = pl->srcy;
  b.zbuf = W.zbuf;
  b.ybuf = W.ybuf;
  b.twz = pl->twz;
  b.twy = pl->twy;
  b.twx = pl->twx;
  b.modx = pl->modx;
  b.mody = pl->mody;
  b.modz = pl->modz;
  b.lut_bin = W.lut_bin ? W.lut_bin : pl->lut_bin;
  b.lut_ff = W.lut_ff ? W.lut_ff : pl->lut_ff;
  b.cells = W.cells;
  b.lut_fstride = W.lut_fstride;
  b.partial = W.partial;
  b.segoff = W.segoff;
  b.nseg = 0;
  b.seglen = 1;
  b.segkr = 0;
  return b;
}

int check_call(saxs_plan* pl, const float* xyz, const int32_t* elem, int n_atoms, int n_frames,
               const float* oc_diag) {
  if (!pl) return fail(SAXS_EINVAL, "plan is NULL");
  if (n_atoms < 0 || n_atoms > pl->max_atoms)
    return fail(SAXS_EINVAL, "n_atoms %d outside [0, %d]", n_atoms, pl->max_atoms);
  if (n_frames < 1 || n_frames > pl->max_frames)
    return fail(SAXS_EINVAL, "n_frames %d outside [1, %d]", n_frames, pl->max_frames);
  if (n_atoms > 0 && (!xyz || !elem)) return fail(SAXS_EINVAL, "xyz/elem pointer is NULL");
  if (!oc_diag) return fail(SAXS_EINVAL, "oc_diag is NULL");
  for (int a = 0; a < 3; ++a) {
    if (!(oc_diag[a] > 0.0f) || !std::isfinite(oc_diag[a]))
      return fail(SAXS_EINVAL, "oc_diag[%d] = %g is not a positive finite inverse length", a,
                  (double)oc_diag[a]);
    if (pl->p.tric && oc_diag[a] != pl->p.cm[4 * a])
      return fail(SAXS_EINVAL, "oc_diag[%d] = %g differs from the plan's triclinic cell (%g)", a,
                  (double)oc_diag[a], (double)pl->p.cm[4 * a]);
  }
  return SAXS_OK;
}

saxs::Params call_params(saxs_plan* pl, int n_atoms, int n_frames, const float* oc_diag) {
  saxs::Params p = pl->p;
  p.F = n_frames;
  p.N = n_atoms;
  for (int a = 0; a < 3; ++a) p.oc[a] = oc_diag[a];
  return p;
}

// Store of the x pass's amplitudes at the kept outputs, which
// saxs_forward_frames saves for the x adjoint of a backward that reuses its
// spectra (k_xfft_p / k_adj_x_p, amp_index layout; k_xfft_gf / k_adj_x_gf,
// amp_gf_index; k_xfft_g / k_adj_x_g, amp_g_index): sized for the current
// kept ranges, grown when they grow.
int ensure_xamp(saxs_plan* pl) {
  const saxs::Params& p = pl->p;
  const int Nx = p.nn[0];
  const size_t bx = (size_t)(((p.kcx * (Nx / 16) + 63) / 64) * 64);
  const size_t need =
      pl->x_pruned ? (size_t)pl->max_frames * p.ky * p.nchx * saxs::kXItems * (2 * (Nx / 128)) * bx
      : pl->x_full_h ? (size_t)pl->max_frames * p.ky * p.nchx * 16 * saxs::kWG   // k_xfft_gf: <= 16 a thread
                     : (size_t)pl->max_frames * p.ky * p.nchx * saxs::kXgItems * 2 * saxs::kWG;
  if (need <= pl->xamp_elems) return SAXS_OK;
  if (pl->xamp) {
    HIP_TRY(hipFree(pl->xamp));
    pl->xamp = nullptr;
    pl->bytes -= pl->xamp_elems * sizeof(float2);
    pl->xamp_elems = 0;
  }
  const int rc = dev_alloc(&pl->xamp, need, pl->bytes);
  if (rc) return rc;
  pl->xamp_elems = need;
  return SAXS_OK;
}

// Enqueues the forward pipeline up to the per-block histograms (or, with
// with_x false, up to the y-pass output for the adjoint).
int launch_forward_core(saxs_plan* pl, const float* xyz, const int32_t* elem, int n_atoms,
                        int n_frames, const float* oc_diag, hipStream_t st, bool with_x = true,
                        const Work* work = nullptr, float2* amp = nullptr) {
  saxs::Params p = call_params(pl, n_atoms, n_frames, oc_diag);
  const Work W = work ? *work : plan_work(pl);
  saxs::Buffers b = make_buffers(pl, xyz, elem, &W);
  const int kframe = saxs::frame_keys(p);               // keys per frame
  const int nkeys = n_frames * kframe;
  const int nclear = std::max(nkeys + 1, n_frames * p.E);
  // 16-bit segment counters need segments of at most kSeg16MaxLen atoms
  // within the segment table's nseg * F <= 512 + F - 1
  const bool seg = pl->seg_sort && n_atoms > 0 &&
                   (!pl->seg16 || (n_atoms + saxs::kSeg16MaxLen - 1) / saxs::kSeg16MaxLen <=
                                      (512 + n_frames - 1) / n_frames);
  if (seg) {  // ~512 workgroups (nseg * F <= 512 + F - 1, the segoff allocation),
             // >= 4096 atoms per segment, and at least K/2 atoms per segment so
             // that the per-segment key histograms (K ints each, written by
             // k_prep_seg and rewritten by k_seg_keys) stay below the atom data
    b.nseg = std::max(1, std::min({(512 + n_frames - 1) / n_frames, (n_atoms + 4095) / 4096,
                                   (int)std::max(1LL, 2LL * n_atoms / kframe)}));
    if (pl->seg16)
      b.nseg = std::max(b.nseg, (n_atoms + saxs::kSeg16MaxLen - 1) / saxs::kSeg16MaxLen);
    b.seglen = (n_atoms + b.nseg - 1) / b.nseg;
    b.segkr = pl->seg_kr;
  }
  { Probe pr(pl, SAXS_K_CLEAR, st);
  hipLaunchKernelGGL(saxs::k_clear, dim3((nclear + 255) / 256), dim3(256), 0, st, W.counts,
                     nkeys + 1, W.pad_sum, n_frames * p.E); }
  if (n_atoms > 0) {
    Probe pr(pl, SAXS_K_PREP, st);
    if (seg) {
      dim3 gq(b.nseg, n_frames);
      const size_t lq = saxs::prep_seg_lds_bytes(pl->seg_kr, p.E, pl->seg16);
      if (pl->seg16) {
        if (p.order == 4)
          hipLaunchKernelGGL((saxs::k_prep_seg<4, true>), gq, dim3(saxs::kSortWG), lq, st, p, b);
        else
          hipLaunchKernelGGL((saxs::k_prep_seg<6, true>), gq, dim3(saxs::kSortWG), lq, st, p, b);
      } else {
        if (p.order == 4)
          hipLaunchKernelGGL((saxs::k_prep_seg<4, false>), gq, dim3(saxs::kSortWG), lq, st, p, b);
        else
          hipLaunchKernelGGL((saxs::k_prep_seg<6, false>), gq, dim3(saxs::kSortWG), lq, st, p, b);
      }
    } else {
      dim3 ga((n_atoms + saxs::kWG - 1) / saxs::kWG, n_frames);
      if (p.order == 4)
        hipLaunchKernelGGL(saxs::k_prep<4>, ga, dim3(saxs::kWG), 0, st, p, b);
      else
        hipLaunchKernelGGL(saxs::k_prep<6>, ga, dim3(saxs::kWG), 0, st, p, b);
    }
  }
  { Probe pr(pl, SAXS_K_SCAN, st);
  if (seg)
    hipLaunchKernelGGL(saxs::k_seg_keys, dim3((kframe + saxs::kWG - 1) / saxs::kWG, n_frames),
                       dim3(saxs::kWG), 0, st, p, b);
  hipLaunchKernelGGL(saxs::k_scan, dim3(n_frames), dim3(1024), 0, st, W.counts, W.offsets,
                     kframe, n_atoms); }
  if (n_atoms > 0) {
    Probe pr(pl, SAXS_K_SCATTER, st);
    dim3 ga((n_atoms + saxs::kWG - 1) / saxs::kWG, n_frames);
    if (p.order == 4)
      hipLaunchKernelGGL(saxs::k_scatter<4>, ga, dim3(saxs::kWG), 0, st, p, b);
    else
      hipLaunchKernelGGL(saxs::k_scatter<6>, ga, dim3(saxs::kWG), 0, st, p, b);
  }
  dim3 gs(p.ntx * p.nty, p.E, n_frames);
  const size_t lds_s = saxs::spread_lds_bytes(p);
  { Probe pr(pl, SAXS_K_SPREAD_Z, st);
  const int rsz = saxs::fft_radix_set(p.fz);
  const size_t lds_all = saxs::spread_all_lds_bytes(p);
  const size_t lds_p = saxs::spread_p_lds_bytes(p);
  const int Nz = p.nn[2];
  const bool pruned_ok = zp_applies(pl->spread_all, pl->pruned_z, p);
  const int zg_h = pruned_ok ? 0 : zg_half(pl->spread_all, pl->pruned_z, p);
  if (!pruned_ok && !zg_h && p.tx * p.ty / 2 > fft_line_cap(p.fz))
    return fail(SAXS_EUNSUP, "column tile %dx%d exceeds the z transform's line capacity", p.tx, p.ty);
  if (pruned_ok) {
    dim3 ga(p.ntx * p.nty, n_frames);
    const bool zhalf = 2 * p.n[2] <= Nz, zexact = 2 * p.n[2] == Nz;
#define SAXS_ZP(PP, NN)                                                                         \
  do {                                                                                          \
    if (zexact) hipLaunchKernelGGL((saxs::k_spread_zp<PP, NN, true, true>), ga, dim3(saxs::kWG), lds_p, st, p, b); \
    else if (zhalf) hipLaunchKernelGGL((saxs::k_spread_zp<PP, NN, true>), ga, dim3(saxs::kWG), lds_p, st, p, b); \
    else hipLaunchKernelGGL((saxs::k_spread_zp<PP, NN, false>), ga, dim3(saxs::kWG), lds_p, st, p, b);    \
  } while (0)
    if (p.order == 4) {
      if (Nz == 128) SAXS_ZP(4, 128); else if (Nz == 256) SAXS_ZP(4, 256); else SAXS_ZP(4, 512);
    } else {
      if (Nz == 128) SAXS_ZP(6, 128); else if (Nz == 256) SAXS_ZP(6, 256); else SAXS_ZP(6, 512);
    }
#undef SAXS_ZP
  } else if (zg_h) {
    dim3 ga(p.ntx * p.nty, n_frames);
    const size_t lds_g = saxs::spread_g_lds_bytes(p, 2 * zg_h);
#define SAXS_ZG(PP, HH) hipLaunchKernelGGL((saxs::k_spread_zg<PP, HH>), ga, dim3(saxs::kWG), lds_g, st, p, b)
#define SAXS_ZGH(PP)                                                                   \
  do {                                                                                 \
    switch (zg_h) {                                                                    \
      case 10: SAXS_ZG(PP, 10); break;                                                 \
      case 9: SAXS_ZG(PP, 9); break;                                                   \
      case 8: SAXS_ZG(PP, 8); break;                                                   \
      case 7: SAXS_ZG(PP, 7); break;                                                   \
      case 6: SAXS_ZG(PP, 6); break;                                                   \
      default: SAXS_ZG(PP, 5); break;                                                  \
    }                                                                                  \
  } while (0)
    if (p.order == 4) SAXS_ZGH(4); else SAXS_ZGH(6);
#undef SAXS_ZGH
#undef SAXS_ZG
  } else if (pl->spread_all && lds_all <= kLdsLimit) {
    dim3 ga(p.ntx * p.nty, n_frames);
#define SAXS_ZA(PP, RS) hipLaunchKernelGGL((saxs::k_spread_zall<PP, RS>), ga, dim3(saxs::kWG), lds_all, st, p, b)
    if (p.order == 4) {
      if (rsz == 1) SAXS_ZA(4, 1); else if (rsz == 0) SAXS_ZA(4, 0); else SAXS_ZA(4, 2);
    } else {
      if (rsz == 1) SAXS_ZA(6, 1); else if (rsz == 0) SAXS_ZA(6, 0); else SAXS_ZA(6, 2);
    }
#undef SAXS_ZA
  } else {
#define SAXS_ZS(PP, RS) hipLaunchKernelGGL((saxs::k_spread_z<PP, RS>), gs, dim3(saxs::kWG), lds_s, st, p, b)
    if (p.order == 4) {
      if (rsz == 1) SAXS_ZS(4, 1); else if (rsz == 0) SAXS_ZS(4, 0); else SAXS_ZS(4, 2);
    } else {
      if (rsz == 1) SAXS_ZS(6, 1); else if (rsz == 0) SAXS_ZS(6, 0); else SAXS_ZS(6, 2);
    }
#undef SAXS_ZS
  } }
  { Probe pr(pl, SAXS_K_YFFT, st);
  const int Ny = p.nn[1];
  const unsigned nfe = (unsigned)(n_frames * p.E);
  if (pl->pruned_z && (Ny == 128 || Ny == 256 || Ny == 512)) {
#define SAXS_YP(NN)                                                                          \
  do {                                                                                       \
    const dim3 gyp((p.n[0] * p.kz + saxs::yfft_p_lines<NN>() - 1) / saxs::yfft_p_lines<NN>(), nfe); \
    const bool ends = p.Ky < NN / 8 && p.ky < NN, yhalf = 2 * p.n[1] <= NN;                  \
    const size_t lq = saxs::yfft_p_lds_bytes<NN>();                                          \
    if (ends && yhalf)                                                                       \
      hipLaunchKernelGGL((saxs::k_yfft_p<NN, true, true>), gyp, dim3(saxs::kWG), lq, st, p, b); \
    else if (ends)                                                                           \
      hipLaunchKernelGGL((saxs::k_yfft_p<NN, true, false>), gyp, dim3(saxs::kWG), lq, st, p, b); \
    else if (yhalf)                                                                          \
      hipLaunchKernelGGL((saxs::k_yfft_p<NN, false, true>), gyp, dim3(saxs::kWG), lq, st, p, b); \
    else                                                                                     \
      hipLaunchKernelGGL((saxs::k_yfft_p<NN, false, false>), gyp, dim3(saxs::kWG), lq, st, p, b); \
  } while (0)
    if (Ny == 128) SAXS_YP(128); else if (Ny == 256) SAXS_YP(256); else SAXS_YP(512);
#undef SAXS_YP
  } else if (int yr2 = 0; const int yfh = yf_half(pl->pruned_z, p, yr2)) {
    const int LW = saxs::yf_lines(Ny / yr2, yr2);
    const dim3 gyf((p.n[0] * p.kz + LW - 1) / LW, nfe);
    const size_t lq = saxs::yfft_gf_lds_bytes(Ny, yr2);
#define SAXS_YF(HH, RR) hipLaunchKernelGGL((saxs::k_yfft_gf<HH, RR>), gyf, dim3(saxs::kWG), lq, st, p, b)
#define SAXS_YFH(RR)                                                                   \
  do {                                                                                 \
    switch (yfh) {                                                                     \
      case 10: SAXS_YF(10, RR); break;                                                 \
      case 9: SAXS_YF(9, RR); break;                                                   \
      case 8: SAXS_YF(8, RR); break;                                                   \
      case 7: SAXS_YF(7, RR); break;                                                   \
      case 6: SAXS_YF(6, RR); break;                                                   \
      default: SAXS_YF(5, RR); break;                                                  \
    }                                                                                  \
  } while (0)
    if (yr2 == 16) SAXS_YFH(16); else SAXS_YFH(32);
#undef SAXS_YFH
#undef SAXS_YF
  } else if (const int yh = yg_half(pl->pruned_z, p)) {
    const int R2 = Ny / (2 * yh), LW = saxs::kWG / R2;
    const dim3 gyg((p.n[0] * p.kz + LW - 1) / LW, nfe);
    const size_t lq = saxs::yfft_g_lds_bytes(Ny, 2 * yh);
#define SAXS_YG(HH) hipLaunchKernelGGL(saxs::k_yfft_g<HH>, gyg, dim3(saxs::kWG), lq, st, p, b)
    switch (yh) {
      case 10: SAXS_YG(10); break;
      case 9: SAXS_YG(9); break;
      case 8: SAXS_YG(8); break;
      case 7: SAXS_YG(7); break;
      case 6: SAXS_YG(6); break;
      default: SAXS_YG(5); break;
    }
#undef SAXS_YG
  } else {
    const int rsy = saxs::fft_radix_set(p.fy);
#define SAXS_Y(RS) hipLaunchKernelGGL(saxs::k_yfft<RS>, dim3(p.n[0], p.nchy, nfe), dim3(saxs::kWG), saxs::yfft_lds_bytes(p), st, p, b)
    if (rsy == 1) SAXS_Y(1); else if (rsy == 0) SAXS_Y(0); else SAXS_Y(2);
#undef SAXS_Y
  } }
  if (!with_x) {
    HIP_TRY(hipGetLastError());
    return SAXS_OK;
  }
  { Probe pr(pl, SAXS_K_XFFT, st);
  if (pl->x_pruned) {
    const int Nx = p.nn[0];
    const bool xhalf = 2 * p.n[0] <= Nx;
    const dim3 gx(p.ky, p.nchx, n_frames);
    const dim3 bx((unsigned)(((p.kcx * (Nx / 16) + 63) / 64) * 64));
#define SAXS_XP(NN)                                                                          \
  do {                                                                                       \
    const size_t lx = saxs::xfft_p_lds_bytes<NN>(p);                                         \
    if (amp && xhalf) hipLaunchKernelGGL((saxs::k_xfft_p<NN, true, true, true>), gx, bx, lx, st, p, b, amp); \
    else if (amp) hipLaunchKernelGGL((saxs::k_xfft_p<NN, true, false, true>), gx, bx, lx, st, p, b, amp); \
    else if (xhalf) hipLaunchKernelGGL((saxs::k_xfft_p<NN, true, true>), gx, bx, lx, st, p, b, nullptr); \
    else hipLaunchKernelGGL((saxs::k_xfft_p<NN, true, false>), gx, bx, lx, st, p, b, nullptr); \
  } while (0)
    if (Nx == 128) SAXS_XP(128); else if (Nx == 256) SAXS_XP(256); else SAXS_XP(512);
#undef SAXS_XP
  } else if (pl->x_full_h) {
    const dim3 gx(p.ky, p.nchx, n_frames);
    const size_t lq = saxs::xfft_gf_lds_bytes(p, pl->x_full_r2);
    // kept outputs of an item among the KLC first and last of its 16
    const int klc = p.Kx / (2 * pl->x_full_h) + 1 <= 4 ? 4 : 8;
#define SAXS_XF(HH, KK)                                                                \
  do {                                                                                 \
    if (amp) hipLaunchKernelGGL((saxs::k_xfft_gf<HH, 16, KK, true>), gx, dim3(saxs::kWG), lq, st, p, b, amp); \
    else hipLaunchKernelGGL((saxs::k_xfft_gf<HH, 16, KK>), gx, dim3(saxs::kWG), lq, st, p, b, nullptr); \
  } while (0)
#define SAXS_XFH(KK)                                                                   \
  do {                                                                                 \
    switch (pl->x_full_h) {                                                            \
      case 10: SAXS_XF(10, KK); break;                                                 \
      case 9: SAXS_XF(9, KK); break;                                                   \
      case 8: SAXS_XF(8, KK); break;                                                   \
      case 7: SAXS_XF(7, KK); break;                                                   \
      case 6: SAXS_XF(6, KK); break;                                                   \
      default: SAXS_XF(5, KK); break;                                                  \
    }                                                                                  \
  } while (0)
    if (klc == 4) SAXS_XFH(4); else SAXS_XFH(8);
#undef SAXS_XFH
#undef SAXS_XF
  } else if (pl->x_gen_h) {
    const dim3 gx(p.ky, p.nchx, n_frames);
    const size_t lq = saxs::xfft_g_lds_bytes(p, 2 * pl->x_gen_h);
#define SAXS_XG(HH)                                                                    \
  do {                                                                                 \
    if (amp) hipLaunchKernelGGL((saxs::k_xfft_g<HH, true>), gx, dim3(saxs::kWG), lq, st, p, b, amp); \
    else hipLaunchKernelGGL((saxs::k_xfft_g<HH>), gx, dim3(saxs::kWG), lq, st, p, b, nullptr); \
  } while (0)
    switch (pl->x_gen_h) {
      case 10: SAXS_XG(10); break;
      case 9: SAXS_XG(9); break;
      case 8: SAXS_XG(8); break;
      case 7: SAXS_XG(7); break;
      case 6: SAXS_XG(6); break;
      default: SAXS_XG(5); break;
    }
#undef SAXS_XG
  } else {
    const int rsx = saxs::fft_radix_set(p.fx);
#define SAXS_X(RS) hipLaunchKernelGGL(saxs::k_xfft_epi<RS>, dim3(p.ky, p.nchx, n_frames), dim3(saxs::kWG), saxs::xfft_lds_bytes(p), st, p, b)
    if (rsx == 1) SAXS_X(1); else if (rsx == 0) SAXS_X(0); else SAXS_X(2);
#undef SAXS_X
  } }
  HIP_TRY(hipGetLastError());
  return SAXS_OK;
}

}  // namespace

extern "C" {

int32_t saxs_abi_version(void) { return 6; }

const char* saxs_last_error(void) { return g_err.c_str(); }

int saxs_plan_create(const saxs_desc* d, saxs_plan** out) {
  g_err.clear();
  if (!d || !out) return fail(SAXS_EINVAL, "desc/out is NULL");
  *out = nullptr;
  for (int a = 0; a < 3; ++a) {
    if (d->n[a] < 1 || d->nn[a] < 1) return fail(SAXS_EINVAL, "grid sizes must be positive");
    if (d->nn[a] < d->n[a])
      return fail(SAXS_EUNSUP, "FFT grid %d smaller than cell grid %d on axis %d", d->nn[a],
                  d->n[a], a);
  }
  if (d->order != 4 && d->order != 6)
    return fail(SAXS_EUNSUP, "spline order %d unsupported (4 or 6)", d->order);
  for (int a = 0; a < 3; ++a)
    if (d->n[a] < 2 * d->order)
      return fail(SAXS_EUNSUP, "cell grid %d too small for order %d", d->n[a], d->order);
  if (d->num_elements < 1 || d->num_elements > saxs::kMaxElements)
    return fail(SAXS_EINVAL, "num_elements %d outside [1, %d]", d->num_elements, saxs::kMaxElements);
  if (d->num_bins < 1 || d->num_bins > 4096)
    return fail(SAXS_EUNSUP, "num_bins %d outside [1, 4096]", d->num_bins);
  if (d->border < 0 || 2 * d->border + 1 > d->n[2])
    return fail(SAXS_EINVAL, "border %d invalid for n_z %d", d->border, d->n[2]);
  if (d->lut_len < 1 || !d->bin_lut || !d->ff_lut || !d->mod_x || !d->mod_y || !d->mod_z)
    return fail(SAXS_EINVAL, "tables missing");
  if (d->max_atoms < 0 || d->max_frames < 1) return fail(SAXS_EINVAL, "capacity invalid");
  const bool tric = is_triclinic(d->cell);
  const bool point_keys = tric || d->kmax[0] > 0 || d->kmax[1] > 0 || d->kmax[2] > 0;
  if (point_keys) {
    const int rc = check_point_box(d->kmax, d->lut_len, tric);
    if (rc) return rc;
  }
  if ((long long)d->max_frames * d->num_elements > 65535)
    return fail(SAXS_EINVAL, "max_frames * num_elements must not exceed 65535");

  saxs_plan* pl = new (std::nothrow) saxs_plan();
  if (!pl) return fail(SAXS_ENOMEM, "host allocation failed");
  pl->desc = *d;
  saxs::Params& p = pl->p;
  for (int a = 0; a < 3; ++a) {
    p.n[a] = d->n[a];
    p.nn[a] = d->nn[a];
    p.nf[a] = (d->nn[a] % 2 == 0) ? d->nn[a] / 2 : d->nn[a] / 2 + 1;
  }
  p.order = d->order;
  p.E = d->num_elements;
  p.border = d->border;
  p.num_bins = d->num_bins;
  p.point_keys = point_keys ? (tric ? 2 : 1) : 0;
  for (int a = 0; a < 3; ++a) p.kmax[a] = point_keys ? d->kmax[a] : 0;
  set_cell(p, d->cell);
  p.inv_count = 1.0 / ((double)d->n[0] * d->n[1] * (2 * d->border + 1));
  const bool wide = (d->flags & SAXS_FLAG_WIDE_RADIX) != 0;
  if (!factor_radices(p.nn[2], p.fz, wide) || !factor_radices(p.nn[1], p.fy, wide) ||
      !factor_radices(p.nn[0], p.fx, wide) || p.nn[0] < 2 || p.nn[1] < 2 || p.nn[2] < 2) {
    free_plan(pl);
    return fail(SAXS_EUNSUP, "FFT sizes must be 2,3,5,7-smooth and >= 2");
  }
  p.lsz = lds_stride(p.nn[2]);
  p.lsy = lds_stride(p.nn[1]);
  p.lsx = lds_stride(p.nn[0]);

  // Column tile: even ty (two real columns per complex line), LDS <= 64 KiB.
  int tx = d->tile_x > 0 ? d->tile_x : 4, ty = d->tile_y > 0 ? d->tile_y : 8;
  tx = std::min(tx, p.n[0]);
  ty = std::min(ty, p.n[1]);
  if (ty % 2) ty = std::max(2, ty - 1);
  // tile + order - 1 <= n: the in-tile points of any stencil form one run
  if (tx + d->order - 1 > p.n[0]) tx = p.n[0] - d->order + 1;
  if (ty + d->order - 1 > p.n[1]) ty = (p.n[1] - d->order + 1) & ~1;
  pl->spread_all = (d->flags & SAXS_FLAG_PER_ELEMENT_SPREAD) == 0;
  pl->fwd_only = (d->flags & SAXS_FLAG_FORWARD_ONLY) != 0;
  pl->pruned_z = (d->flags & SAXS_FLAG_NO_PRUNED_Z) == 0;
  if (d->tile_x <= 0 && d->tile_y <= 0 && tx == 4 && ty == 8) {
    // default tile on long generic columns: 4 x 4 once the generic pruned
    // spreading kernel's LDS passes 48 KiB at 4 x 8 (two workgroups a CU;
    // measured faster from 160^3 cells up, slower below)
    p.tx = tx;
    p.ty = ty;
    const int h = zg_half(pl->spread_all, pl->pruned_z, p);
    if (h && saxs::spread_g_lds_bytes(p, 2 * h) > 48 * 1024) ty = 4;
  }
  for (;;) {
    p.tx = tx;
    p.ty = ty;
    if (tile_fits(pl->spread_all, pl->pruned_z, p, d->lut_len,
                  (d->flags & SAXS_FLAG_FORWARD_ONLY) != 0))
      break;
    if (tx >= ty && tx > 1)
      tx = (tx + 1) / 2;
    else if (ty > 2)
      ty = std::max(2, (ty / 2) & ~1);
    else {
      free_plan(pl);
      return fail(SAXS_EUNSUP, "cell column n_z=%d does not fit the LDS budget", p.n[2]);
    }
  }
  if (p.n[1] < 2) {
    free_plan(pl);
    return fail(SAXS_EUNSUP, "n_y must be >= 2");
  }
  p.ntx = (p.n[0] + p.tx - 1) / p.tx;
  p.nty = (p.n[1] + p.ty - 1) / p.ty;
  // Sort-key classes per axis: the most that keep every tile's fetch within
  // its nine record ranges and cost the counting sort no extra pass.
  std::vector<int4> sx, sy;
  {
    const int base_passes = sort_passes(p.E * p.nty * p.ntx, p.E);
    const bool classes = (d->flags & SAXS_FLAG_NO_KEY_CLASSES) == 0;
    const int cand[4][2] = {{2, 2}, {1, 2}, {2, 1}, {1, 1}};
    bool found = false;
    for (const auto& c : cand) {
      if (!classes && (c[0] != 1 || c[1] != 1)) continue;
      if ((c[0] != 1 || c[1] != 1) &&
          sort_passes(p.E * p.nty * c[1] * p.ntx * c[0], p.E) > base_passes)
        continue;
      if (!tile_units(p.n[0], p.tx, p.order, c[0], true, sx) ||
          !tile_units(p.n[1], p.ty, p.order, c[1], false, sy))
        continue;
      p.ncx = c[0];
      p.ncy = c[1];
      found = true;
      break;
    }
    if (!found) {
      free_plan(pl);
      return fail(SAXS_EUNSUP, "spreading tile %dx%d too narrow for order %d", p.tx, p.ty, p.order);
    }
  }
  pl->nsx = slot_count(p.n[0], p.tx, p.order);
  pl->nsy = slot_count(p.n[1], p.ty, p.order);
  int rc = derive_ranges(pl, d->lut_len);
  if (rc) {
    free_plan(pl);
    return rc;
  }
  // Room for a box change (NPT) that keeps one more frequency per axis, and
  // for switching between shell and point tables: q tables up to the point
  // box (K+2)^3, kept ranges one wider.
  {
    int K[3];
    kept_max(p, d->lut_len, K);
    const long long kb = 2LL * std::max({K[0], K[1], K[2]}) + 5;   // signed box, K + 2
    pl->lut_cap = (int)std::min<long long>(std::max<long long>(d->lut_len, kb * kb * kb), 1 << 26);
    pl->lut_cap = std::max(pl->lut_cap, d->lut_len);
    pl->kz_cap = std::min(p.kz + 1, p.nn[2] / 2 + 1);
    pl->ky_cap = std::min(p.ky + 2, p.nn[1]);
    pl->nchx_cap = p.nchx + 1;
  }
  pl->max_frames = d->max_frames;
  pl->max_atoms = d->max_atoms;
  pl->nkeys_max = d->max_frames * saxs::frame_keys(p);

  const size_t F = d->max_frames, N = d->max_atoms, E = p.E;
  const size_t n3 = (size_t)p.n[0] * p.n[1] * p.n[2];
  size_t& tot = pl->bytes;
#define ALLOC(ptr, cnt)                  \
  do {                                   \
    int r_ = dev_alloc(&(ptr), (cnt), tot); \
    if (r_) {                            \
      free_plan(pl);                     \
      return r_;                         \
    }                                    \
  } while (0)
  ALLOC(pl->srcx, sx.size());
  ALLOC(pl->srcy, sy.size());
  ALLOC(pl->twz, (size_t)p.nn[2]);
  ALLOC(pl->twy, (size_t)p.nn[1]);
  ALLOC(pl->twx, (size_t)p.nn[0]);
  ALLOC(pl->modx, (size_t)p.nn[0]);
  ALLOC(pl->mody, (size_t)p.nn[1]);
  ALLOC(pl->modz, (size_t)p.nn[2]);
  ALLOC(pl->lut_bin, (size_t)pl->lut_cap);
  ALLOC(pl->lut_ff, E * pl->lut_cap);
  ALLOC(pl->counts, (size_t)pl->nkeys_max + 1);
  ALLOC(pl->offsets, (size_t)pl->nkeys_max + 1);
  ALLOC(pl->keyslot, F * N);
  {
    const int kframe = saxs::frame_keys(p);
    // LDS histogram of a frame's keys: 32-bit counters in one pass, else
    // 16-bit counters in passes over key ranges (up to 4), else the global-
    // atomic sort (k_prep); k_prep_seg<P, false> always runs a single pass
    if (saxs::prep_seg_lds_bytes(kframe, E) <= kSortLdsBudget) {
      pl->seg_sort = true;
      pl->seg_kr = kframe;
    } else {
      const int kr = seg_kr16((int)E);
      if ((kframe + kr - 1) / kr <= 4) {
        pl->seg_sort = pl->seg16 = true;
        pl->seg_kr = std::min(kr, kframe);
      }
    }
    if (pl->seg_sort) ALLOC(pl->segoff, (size_t)(512 + F) * kframe);
  }
  ALLOC(pl->sorted, F * N);
  ALLOC(pl->pad_sum, F * E);
  ALLOC(pl->pad_given, E);
  ALLOC(pl->padg, F * E);
  ALLOC(pl->padt, F * E * p.ntx * p.nty);
  ALLOC(pl->zbuf, F * E * p.n[0] * p.n[1] * pl->kz_cap);
  ALLOC(pl->ybuf, F * E * pl->ky_cap * p.n[0] * pl->kz_cap);
  ALLOC(pl->partial, F * pl->ky_cap * pl->nchx_cap * 2 * p.num_bins);
  ALLOC(pl->hframe, F * 2 * p.num_bins);
#undef ALLOC
  const std::vector<float2> tz = twiddles(p.nn[2]), ty_ = twiddles(p.nn[1]), tx_ = twiddles(p.nn[0]);
  int r = 0;
  if ((r = upload(pl->srcx, sx.data(), sx.size())) || (r = upload(pl->srcy, sy.data(), sy.size())) ||
      (r = upload(pl->twz, tz.data(), tz.size())) || (r = upload(pl->twy, ty_.data(), ty_.size())) ||
      (r = upload(pl->twx, tx_.data(), tx_.size())) ||
      (r = upload(pl->modx, d->mod_x, (size_t)p.nn[0])) ||
      (r = upload(pl->mody, d->mod_y, (size_t)p.nn[1])) ||
      (r = upload(pl->modz, d->mod_z, (size_t)p.nn[2])) ||
      (r = upload(pl->lut_bin, d->bin_lut, (size_t)d->lut_len)) ||
      (r = upload(pl->lut_ff, d->ff_lut, E * d->lut_len))) {
    free_plan(pl);
    return r;
  }
  *out = pl;
  return SAXS_OK;
}

int saxs_plan_set_tables(saxs_plan* pl, int32_t lut_len, const int32_t* bin_lut, const float* ff_lut) {
  g_err.clear();
  return set_tables(pl, nullptr, nullptr, lut_len, bin_lut, ff_lut);
}

int saxs_plan_set_point_tables(saxs_plan* pl, const float* cell, const int32_t* kmax,
                               int32_t lut_len, const int32_t* bin_lut, const float* ff_lut) {
  g_err.clear();
  if (!kmax) return fail(SAXS_EINVAL, "NULL argument");
  const int rc = check_point_box(kmax, lut_len, is_triclinic(cell));
  if (rc) return rc;
  return set_tables(pl, cell, kmax, lut_len, bin_lut, ff_lut);
}

int saxs_plan_set_padding(saxs_plan* pl, const double* pad) {
  g_err.clear();
  if (!pl) return fail(SAXS_EINVAL, "plan is NULL");
  if (!pad) {
    pl->pad_given_host.clear();
    pl->last_frames_valid = false;
    return SAXS_OK;
  }
  const int E = pl->p.E;
  std::vector<unsigned long long> units(E);
  // pad * 2^32 fixed-point units, summed over at most n_z voxels of a column,
  // must stay below 2^63: pad <= 2^31 / n_z atoms per voxel
  const double pad_max = 2147483648.0 / (double)pl->p.n[2];
  for (int e = 0; e < E; ++e) {
    if (!(pad[e] >= 0.0 && pad[e] <= pad_max))
      return fail(SAXS_EINVAL, "pad[%d] = %g outside [0, 2^31 / n_z = %g] atoms per voxel", e,
                  pad[e], pad_max);
    units[e] = (unsigned long long)std::llrint(pad[e] * 4294967296.0);
  }
  HIP_TRY(hipMemcpy(pl->pad_given, units.data(), sizeof(unsigned long long) * E,
                    hipMemcpyHostToDevice));
  pl->pad_given_host.assign(pad, pad + E);
  pl->last_frames_valid = false;
  return SAXS_OK;
}

int saxs_forward(saxs_plan* pl, const float* xyz, const int32_t* elem, int32_t n_atoms,
                 int32_t n_frames, const float* oc_diag, double* H, double* C, void* stream) {
  g_err.clear();
  int rc = check_call(pl, xyz, elem, n_atoms, n_frames, oc_diag);
  if (rc) return rc;
  if (!H || !C) return fail(SAXS_EINVAL, "H/C pointer is NULL");
  hipStream_t st = (hipStream_t)stream;
  pl->last_frames_valid = false;
  rc = launch_forward_core(pl, xyz, elem, n_atoms, n_frames, oc_diag, st);
  if (rc) return rc;
  const int nblk = pl->p.ky * pl->p.nchx;
  {
    Probe pr(pl, SAXS_K_REDUCE, st);
    hipLaunchKernelGGL(saxs::k_hist_reduce, dim3(pl->p.num_bins, 1), dim3(saxs::kWG), 0, st,
                       pl->partial, nblk, n_frames, pl->p.num_bins, H, C, 0);
  }
  HIP_TRY(hipGetLastError());
  return SAXS_OK;
}

int saxs_forward_static(saxs_plan* pl, const float* xyz, const int32_t* elem, int32_t n_atoms,
                        int32_t n_frames, const float* oc_diag, double* H, double* C,
                        void* stream) {
  g_err.clear();
  if (!pl) return fail(SAXS_EINVAL, "plan is NULL");
  if (n_atoms < 0) return fail(SAXS_EINVAL, "n_atoms %d < 0", n_atoms);
  if (n_frames < 1 || (long long)n_frames * pl->p.E > 65535)
    return fail(SAXS_EINVAL, "n_frames %d outside [1, 65535 / num_elements]", n_frames);
  // the plan's capacity does not bound this call; the static scratch does
  int rc = check_call(pl, xyz, elem, 0, 1, oc_diag);
  if (rc) return rc;
  if (n_atoms > 0 && (!xyz || !elem)) return fail(SAXS_EINVAL, "xyz/elem pointer is NULL");
  if (!H || !C) return fail(SAXS_EINVAL, "H/C pointer is NULL");
  Work w{};
  const size_t need = carve_work(pl, n_atoms, n_frames, nullptr, w);
  if (need > kStaticScratchBytes)
    return fail(SAXS_EINVAL,
                "%d frames of %d atoms need %zu bytes of static scratch, more than its %zu "
                "(call with fewer frames)", n_frames, n_atoms, need, kStaticScratchBytes);
  void* base = nullptr;
  if ((rc = static_scratch(&base))) return rc;
  carve_work(pl, n_atoms, n_frames, (unsigned char*)base, w);
  hipStream_t st = (hipStream_t)stream;
  pl->last_frames_valid = false;
  const bool timing = pl->timer.on;      // no event creation on this path
  pl->timer.on = false;
  rc = launch_forward_core(pl, xyz, elem, n_atoms, n_frames, oc_diag, st, true, &w);
  if (!rc)
    hipLaunchKernelGGL(saxs::k_hist_reduce, dim3(pl->p.num_bins, 1), dim3(saxs::kWG), 0, st,
                       w.partial, pl->p.ky * pl->p.nchx, n_frames, pl->p.num_bins, H, C, 0);
  pl->timer.on = timing;
  if (rc) return rc;
  HIP_TRY(hipGetLastError());
  return SAXS_OK;
}

int saxs_forward_boxes(saxs_plan* pl, const float* xyz, const int32_t* elem, int32_t n_atoms,
                       int32_t n_frames, const saxs_frame_boxes* fb, double* H, double* C,
                       void* stream) {
  g_err.clear();
  if (!pl || !fb) return fail(SAXS_EINVAL, "plan/boxes is NULL");
  if (n_atoms < 0 || n_atoms > pl->max_atoms)
    return fail(SAXS_EINVAL, "n_atoms %d outside [0, %d]", n_atoms, pl->max_atoms);
  if (n_frames < 1 || n_frames > pl->max_frames)
    return fail(SAXS_EINVAL, "n_frames %d outside [1, %d]", n_frames, pl->max_frames);
  if (n_atoms > 0 && (!xyz || !elem)) return fail(SAXS_EINVAL, "xyz/elem pointer is NULL");
  if (!H || !C || !fb->cells || !fb->qgeom || !fb->form_coef)
    return fail(SAXS_EINVAL, "H/C/cells/qgeom/form_coef pointer is NULL");
  if (fb->key_type < 0 || fb->key_type > 2) return fail(SAXS_EINVAL, "key_type %d", fb->key_type);
  if (fb->lut_len < 1) return fail(SAXS_EINVAL, "lut_len %d < 1", fb->lut_len);
  if (fb->key_type > 0) {
    const int rc = check_point_box(fb->kmax, fb->lut_len, fb->key_type == 2);
    if (rc) return rc;
  }
  if (fb->lut_len > pl->lut_cap)
    return fail(SAXS_ECAPACITY, "lut_len %d exceeds the plan's %d (create a plan for the "
                "largest box)", fb->lut_len, pl->lut_cap);
  for (int f = 0; f < n_frames; ++f)
    for (int a = 0; a < 3; ++a)
      if (!(fb->cells[9 * f + 4 * a] > 0.0f) || !std::isfinite(fb->cells[9 * f + 4 * a]))
        return fail(SAXS_EINVAL, "frame %d: cell diagonal %d = %g is not a positive inverse "
                    "length", f, a, (double)fb->cells[9 * f + 4 * a]);
  // kept ranges of the batch's key box (restored when the launches are queued)
  const saxs::Params saved = pl->p;
  const bool saved_xp = pl->x_pruned;
  const int saved_xg = pl->x_gen_h, saved_xfh = pl->x_full_h, saved_xfr = pl->x_full_r2;
  auto restore = [&]() {
    pl->p = saved;
    pl->x_pruned = saved_xp;
    pl->x_gen_h = saved_xg;
    pl->x_full_h = saved_xfh;
    pl->x_full_r2 = saved_xfr;
  };
  pl->p.point_keys = fb->key_type;
  for (int a = 0; a < 3; ++a) pl->p.kmax[a] = fb->key_type ? fb->kmax[a] : 0;
  int rc = derive_ranges(pl, fb->lut_len);
  if (!rc && (pl->p.kz > pl->kz_cap || pl->p.ky > pl->ky_cap || pl->p.nchx > pl->nchx_cap))
    rc = fail(SAXS_ECAPACITY, "the batch keeps kz %d, ky %d, %d x chunks; the workspace holds "
              "%d, %d, %d (create a plan for the largest box)", pl->p.kz, pl->p.ky, pl->p.nchx,
              pl->kz_cap, pl->ky_cap, pl->nchx_cap);
  if (rc) {
    restore();
    return rc;
  }
  const size_t MF = pl->max_frames, E = pl->p.E;
  size_t tot = pl->bytes;
  int r = 0;
  if (!pl->fb_cells) {
    if ((r = dev_alloc(&pl->fb_cells, MF * 9, tot)) || (r = dev_alloc(&pl->fb_geom, (MF + 1) * 9, tot)) ||
        (r = dev_alloc(&pl->fb_coef, E * 9, tot)) ||
        (r = dev_alloc(&pl->fb_bin, MF * (size_t)pl->lut_cap, tot)) ||
        (r = dev_alloc(&pl->fb_ff, MF * E * (size_t)pl->lut_cap, tot))) {
      restore();
      return r;
    }
    if (hipHostMalloc((void**)&pl->fb_host, sizeof(float) * (MF * 18 + 9 + 9 * E)) != hipSuccess ||
        hipEventCreateWithFlags(&pl->fb_copied, hipEventDisableTiming) != hipSuccess) {
      restore();
      return fail(SAXS_ENOMEM, "pinned staging for per-frame boxes failed");
    }
    pl->bytes = tot;
  }
  hipStream_t st = (hipStream_t)stream;
  // the previous batch's upload must have left the staging buffer
  if (pl->fb_pending && hipEventSynchronize(pl->fb_copied) != hipSuccess) {
    restore();
    return fail(SAXS_EHIP, "waiting for the previous upload failed");
  }
  float* hc = pl->fb_host;
  float* hg = hc + MF * 9;
  float* hk = hg + (MF + 1) * 9;
  std::memcpy(hc, fb->cells, sizeof(float) * 9 * n_frames);
  std::memcpy(hg, fb->qgeom, sizeof(float) * 9 * n_frames);
  if (fb->bin_qgeom) std::memcpy(hg + MF * 9, fb->bin_qgeom, sizeof(float) * 9);
  std::memcpy(hk, fb->form_coef, sizeof(float) * 9 * E);
  if (hipMemcpyAsync(pl->fb_cells, hc, sizeof(float) * 9 * n_frames, hipMemcpyHostToDevice, st) ||
      hipMemcpyAsync(pl->fb_geom, hg, sizeof(float) * 9 * n_frames, hipMemcpyHostToDevice, st) ||
      (fb->bin_qgeom && hipMemcpyAsync(pl->fb_geom + MF * 9, hg + MF * 9, sizeof(float) * 9,
                                       hipMemcpyHostToDevice, st)) ||
      hipMemcpyAsync(pl->fb_coef, hk, sizeof(float) * 9 * E, hipMemcpyHostToDevice, st) ||
      hipEventRecord(pl->fb_copied, st)) {
    restore();
    return fail(SAXS_EHIP, "upload of the per-frame boxes failed");
  }
  pl->fb_pending = true;
  pl->last_frames_valid = false;
  saxs::Params tp = pl->p;
  tp.F = n_frames;
  tp.lut_len = fb->lut_len;
  hipLaunchKernelGGL(saxs::k_frame_tables, dim3((fb->lut_len + saxs::kWG - 1) / saxs::kWG, n_frames),
                     dim3(saxs::kWG), 0, st, tp, pl->fb_geom,
                     fb->bin_qgeom ? pl->fb_geom + MF * 9 : nullptr, pl->fb_coef, fb->dq, fb->qcut,
                     pl->p.num_bins, pl->fb_bin, pl->fb_ff);
  Work w = plan_work(pl);
  w.cells = pl->fb_cells;
  w.lut_bin = pl->fb_bin;
  w.lut_ff = pl->fb_ff;
  w.lut_fstride = fb->lut_len;
  const float oc1[3] = {fb->cells[0], fb->cells[4], fb->cells[8]};
  rc = launch_forward_core(pl, xyz, elem, n_atoms, n_frames, oc1, st, true, &w);
  if (!rc) {
    Probe pr(pl, SAXS_K_REDUCE, st);
    hipLaunchKernelGGL(saxs::k_hist_reduce, dim3(pl->p.num_bins, 1), dim3(saxs::kWG), 0, st,
                       pl->partial, pl->p.ky * pl->p.nchx, n_frames, pl->p.num_bins, H, C, 0);
  }
  restore();
  if (rc) return rc;
  HIP_TRY(hipGetLastError());
  return SAXS_OK;
}

int saxs_forward_frames(saxs_plan* pl, const float* xyz, const int32_t* elem, int32_t n_atoms,
                        int32_t n_frames, const float* oc_diag, double* Hf, double* Cf,
                        void* stream) {
  g_err.clear();
  int rc = check_call(pl, xyz, elem, n_atoms, n_frames, oc_diag);
  if (rc) return rc;
  if (!Hf || !Cf) return fail(SAXS_EINVAL, "Hf/Cf pointer is NULL");
  hipStream_t st = (hipStream_t)stream;
  pl->last_frames_valid = false;
  pl->xamp_valid = false;
  // a plan with a backward keeps the x pass's amplitudes for its x adjoint
  const bool save_amp =
      (pl->x_pruned || pl->x_full_h || pl->x_gen_h) && !pl->fwd_only && n_atoms > 0;
  if (save_amp) {
    rc = ensure_xamp(pl);
    if (rc) return rc;
  }
  rc = launch_forward_core(pl, xyz, elem, n_atoms, n_frames, oc_diag, st, true, nullptr,
                           save_amp ? pl->xamp : nullptr);
  if (rc) return rc;
  const int nblk = pl->p.ky * pl->p.nchx;
  {
    Probe pr(pl, SAXS_K_REDUCE, st);
    hipLaunchKernelGGL(saxs::k_hist_reduce, dim3(pl->p.num_bins, n_frames), dim3(saxs::kWG), 0,
                       st, pl->partial, nblk, 1, pl->p.num_bins, Hf, Cf, 1);
  }
  pl->xamp_valid = save_amp;
  pl->last_frames_valid = true;
  pl->last_xyz = xyz;
  pl->last_N = n_atoms;
  pl->last_F = n_frames;
  HIP_TRY(hipGetLastError());
  return SAXS_OK;
}

int saxs_backward(saxs_plan* pl, const float* xyz, const int32_t* elem, int32_t n_atoms,
                  int32_t n_frames, const float* oc_diag, const double* gH, float* grad,
                  void* stream) {
  return saxs_backward_ex(pl, xyz, elem, n_atoms, n_frames, oc_diag, gH, grad, 0, stream);
}

int saxs_backward_ex(saxs_plan* pl, const float* xyz, const int32_t* elem, int32_t n_atoms,
                     int32_t n_frames, const float* oc_diag, const double* gH, float* grad,
                     int32_t flags, void* stream) {
  g_err.clear();
  int rc = check_call(pl, xyz, elem, n_atoms, n_frames, oc_diag);
  if (rc) return rc;
  if (!gH || (n_atoms > 0 && !grad)) return fail(SAXS_EINVAL, "gH/grad pointer is NULL");
  {  // a forward-only plan's tile may not fit the z adjoint: refuse before any launch
    saxs::Params q = pl->p;
    const int Nz = q.nn[2];
    const bool pruned = pl->spread_all && pl->pruned_z && (Nz == 128 || Nz == 256 || Nz == 512) &&
                        2 * q.kz - 1 <= Nz / 4 && q.kz <= Nz / 8 &&
                        (q.tx * q.ty / 2) * (Nz / 16) <= saxs::kWG &&
                        saxs::adj_p_lds_bytes(q) <= kLdsLimit;
    if (!pruned && !adjg_half(pl->spread_all, pl->pruned_z, q) &&
        (q.tx * q.ty / 2 > fft_line_cap(q.fz) || saxs::adjz_lds_bytes(q) > kLdsLimit))
      return fail(SAXS_EUNSUP, "column tile %dx%d does not fit the z adjoint (plan created with "
                  "SAXS_FLAG_FORWARD_ONLY)", q.tx, q.ty);
  }
  hipStream_t st = (hipStream_t)stream;
  if (!pl->gpart && pl->max_atoms > 0) {   // partial-gradient slots, on first use
    rc = dev_alloc(&pl->gpart, (size_t)pl->max_frames * pl->nsx * pl->nsy * pl->max_atoms * 3,
                   pl->bytes);
    if (rc) return rc;
  }
  const bool reuse = (flags & SAXS_BWD_REUSE_FORWARD) != 0;
  if (reuse) {
    if (!pl->last_frames_valid || pl->last_xyz != xyz || pl->last_N != n_atoms ||
        pl->last_F != n_frames)
      return fail(SAXS_EINVAL, "no matching saxs_forward_frames state to reuse");
  } else {
    rc = launch_forward_core(pl, xyz, elem, n_atoms, n_frames, oc_diag, st, false);
    if (rc) return rc;
  }
  // the x adjoint reads the forward's saved amplitudes instead of recomputing them
  const float2* amp = (reuse && pl->xamp_valid) ? pl->xamp : nullptr;
  pl->last_frames_valid = false;           // the adjoint overwrites the spectra
  pl->xamp_valid = false;
  saxs::Params p = call_params(pl, n_atoms, n_frames, oc_diag);
  saxs::Buffers b = make_buffers(pl, xyz, elem);
  saxs::AdjBuffers g{gH, grad, pl->padg, pl->padt, pl->gpart, pl->nsx, pl->nsy};
  const int px = saxs::fft_radix_set(p.fx), py = saxs::fft_radix_set(p.fy),
            pz = saxs::fft_radix_set(p.fz);
  {
    Probe pr(pl, SAXS_K_ADJ_X, st);
    dim3 gx(p.ky, p.nchx, n_frames);
    if (pl->x_pruned) {
      const int Nx = p.nn[0];
      const bool xhalf = 2 * p.n[0] <= Nx;
      const dim3 bx((unsigned)(((p.kcx * (Nx / 16) + 63) / 64) * 64));
#define SAXS_AXP(NN)                                                                         \
  do {                                                                                       \
    const size_t lx = saxs::adjx_p_lds_bytes<NN>(p);                                         \
    if (amp && xhalf) hipLaunchKernelGGL((saxs::k_adj_x_p<NN, true, true>), gx, bx, lx, st, p, b, g, amp); \
    else if (amp) hipLaunchKernelGGL((saxs::k_adj_x_p<NN, false, true>), gx, bx, lx, st, p, b, g, amp); \
    else if (xhalf) hipLaunchKernelGGL((saxs::k_adj_x_p<NN, true>), gx, bx, lx, st, p, b, g, nullptr); \
    else hipLaunchKernelGGL((saxs::k_adj_x_p<NN, false>), gx, bx, lx, st, p, b, g, nullptr); \
  } while (0)
      if (Nx == 128) SAXS_AXP(128); else if (Nx == 256) SAXS_AXP(256); else SAXS_AXP(512);
#undef SAXS_AXP
    } else if (pl->x_full_h) {
      const size_t lq = saxs::adj_x_gf_lds_bytes(p, 16);
      const int klc = p.Kx / (2 * pl->x_full_h) + 1 <= 4 ? 4 : 8;
#define SAXS_AXF(HH, KK)                                                               \
  do {                                                                                 \
    if (amp) hipLaunchKernelGGL((saxs::k_adj_x_gf<HH, 16, KK, true>), gx, dim3(saxs::kWG), lq, st, p, b, g, amp); \
    else hipLaunchKernelGGL((saxs::k_adj_x_gf<HH, 16, KK, false>), gx, dim3(saxs::kWG), lq, st, p, b, g, nullptr); \
  } while (0)
#define SAXS_AXFH(KK)                                                                  \
  do {                                                                                 \
    switch (pl->x_full_h) {                                                            \
      case 10: SAXS_AXF(10, KK); break;                                                \
      case 9: SAXS_AXF(9, KK); break;                                                  \
      case 8: SAXS_AXF(8, KK); break;                                                  \
      case 7: SAXS_AXF(7, KK); break;                                                  \
      case 6: SAXS_AXF(6, KK); break;                                                  \
      default: SAXS_AXF(5, KK); break;                                                 \
    }                                                                                  \
  } while (0)
      if (klc == 4) SAXS_AXFH(4); else SAXS_AXFH(8);
#undef SAXS_AXFH
#undef SAXS_AXF
    } else if (pl->x_gen_h) {
      const size_t lq = saxs::adj_x_g_lds_bytes(p, 2 * pl->x_gen_h);
#define SAXS_AXG(HH)                                                                   \
  do {                                                                                 \
    if (amp) hipLaunchKernelGGL((saxs::k_adj_x_g<HH, true>), gx, dim3(saxs::kWG), lq, st, p, b, g, amp); \
    else hipLaunchKernelGGL((saxs::k_adj_x_g<HH>), gx, dim3(saxs::kWG), lq, st, p, b, g, nullptr); \
  } while (0)
      switch (pl->x_gen_h) {
        case 10: SAXS_AXG(10); break;
        case 9: SAXS_AXG(9); break;
        case 8: SAXS_AXG(8); break;
        case 7: SAXS_AXG(7); break;
        case 6: SAXS_AXG(6); break;
        default: SAXS_AXG(5); break;
      }
#undef SAXS_AXG
    } else {
#define SAXS_AX(RS) hipLaunchKernelGGL(saxs::k_adj_x<RS>, gx, dim3(saxs::kWG), saxs::adjx_lds_bytes(p), st, p, b, g)
      if (px == 1) SAXS_AX(1); else if (px == 0) SAXS_AX(0); else SAXS_AX(2);
#undef SAXS_AX
    }
  }
  const int Ny = p.nn[1];
  if (pl->pruned_z && (Ny == 128 || Ny == 256 || Ny == 512)) {
    Probe pr(pl, SAXS_K_ADJ_Y, st);
    const unsigned nfe = (unsigned)(n_frames * p.E);
#define SAXS_AYP(NN)                                                                         \
  do {                                                                                       \
    const int lw_ = saxs::yfft_p_lines<NN>() * saxs::kAdjYGroups;                           \
    const dim3 gay((p.n[0] * p.kz + lw_ - 1) / lw_, nfe);                                    \
    const bool ends = p.Ky < NN / 8 && p.ky < NN, yhalf = 2 * p.n[1] <= NN;                  \
    const size_t lq = saxs::yfft_p_lds_bytes<NN>();                                          \
    if (ends && yhalf)                                                                       \
      hipLaunchKernelGGL((saxs::k_adj_y_p<NN, true, true>), gay, dim3(saxs::kWG), lq, st, p, b); \
    else if (ends)                                                                           \
      hipLaunchKernelGGL((saxs::k_adj_y_p<NN, true, false>), gay, dim3(saxs::kWG), lq, st, p, b); \
    else if (yhalf)                                                                          \
      hipLaunchKernelGGL((saxs::k_adj_y_p<NN, false, true>), gay, dim3(saxs::kWG), lq, st, p, b); \
    else                                                                                     \
      hipLaunchKernelGGL((saxs::k_adj_y_p<NN, false, false>), gay, dim3(saxs::kWG), lq, st, p, b); \
  } while (0)
    if (Ny == 128) SAXS_AYP(128); else if (Ny == 256) SAXS_AYP(256); else SAXS_AYP(512);
#undef SAXS_AYP
  } else if (int r2f = 0, hf = yf_half(pl->pruned_z, p, r2f); hf && r2f == 16) {
    // the transpose of the full-output y pass (N = 2H x 16, most outputs kept)
    Probe pr(pl, SAXS_K_ADJ_Y, st);
    const int LW = saxs::yf_adj_lines(Ny / 16, 16);
    const dim3 gay((p.n[0] * p.kz + LW - 1) / LW, (unsigned)(n_frames * p.E));
    const size_t lq = saxs::adj_y_gf_lds_bytes(Ny);
#define SAXS_AYF(HH) hipLaunchKernelGGL(saxs::k_adj_y_gf<HH>, gay, dim3(saxs::kWG), lq, st, p, b)
    switch (hf) {
      case 10: SAXS_AYF(10); break;
      case 9: SAXS_AYF(9); break;
      case 8: SAXS_AYF(8); break;
      case 7: SAXS_AYF(7); break;
      case 6: SAXS_AYF(6); break;
      default: SAXS_AYF(5); break;
    }
#undef SAXS_AYF
  } else if (const int yh = yg_half(pl->pruned_z, p)) {
    Probe pr(pl, SAXS_K_ADJ_Y, st);
    const int R2 = Ny / (2 * yh), LW = saxs::kWG / R2;
    const dim3 gay((p.n[0] * p.kz + LW - 1) / LW, (unsigned)(n_frames * p.E));
    const size_t lq = saxs::adj_y_g_lds_bytes(p, 2 * yh);
#define SAXS_AYG(HH) hipLaunchKernelGGL(saxs::k_adj_y_g<HH>, gay, dim3(saxs::kWG), lq, st, p, b)
    switch (yh) {
      case 10: SAXS_AYG(10); break;
      case 9: SAXS_AYG(9); break;
      case 8: SAXS_AYG(8); break;
      case 7: SAXS_AYG(7); break;
      case 6: SAXS_AYG(6); break;
      default: SAXS_AYG(5); break;
    }
#undef SAXS_AYG
  } else {
    Probe pr(pl, SAXS_K_ADJ_Y, st);
    dim3 gy(p.n[0], p.nchy, n_frames * p.E);
#define SAXS_AY(RS) hipLaunchKernelGGL(saxs::k_adj_y<RS>, gy, dim3(saxs::kWG), saxs::yfft_lds_bytes(p), st, p, b)
    if (py == 1) SAXS_AY(1); else if (py == 0) SAXS_AY(0); else SAXS_AY(2);
#undef SAXS_AY
  }
  const int Nz = p.nn[2];
  const size_t lds_ap = saxs::adj_p_lds_bytes(p);
  const bool adj_pruned = pl->spread_all && pl->pruned_z && (Nz == 128 || Nz == 256 || Nz == 512) &&
                          2 * p.kz - 1 <= Nz / 4 && p.kz <= Nz / 8 &&
                          (p.tx * p.ty / 2) * (Nz / 16) <= saxs::kWG && lds_ap <= kLdsLimit;
  const int adj_h = adj_pruned ? 0 : adjg_half(pl->spread_all, pl->pruned_z, p);
  if (!adj_pruned && !adj_h &&
      (p.tx * p.ty / 2 > fft_line_cap(p.fz) || saxs::adjz_lds_bytes(p) > kLdsLimit))
    return fail(SAXS_EUNSUP, "column tile %dx%d does not fit the z adjoint", p.tx, p.ty);
  if (adj_h) {
    Probe pr(pl, SAXS_K_ADJ_Z, st);
    dim3 ga(p.ntx * p.nty, n_frames);
    const size_t lds = saxs::adj_g_lds_bytes(p);
#define SAXS_AZG(PP, HH) hipLaunchKernelGGL((saxs::k_adj_zg<PP, HH>), ga, dim3(saxs::kWG), lds, st, p, b, g)
#define SAXS_AZGH(PP)                                                                  \
  do {                                                                                 \
    switch (adj_h) {                                                                   \
      case 10: SAXS_AZG(PP, 10); break;                                                \
      case 9: SAXS_AZG(PP, 9); break;                                                  \
      case 8: SAXS_AZG(PP, 8); break;                                                  \
      case 7: SAXS_AZG(PP, 7); break;                                                  \
      case 6: SAXS_AZG(PP, 6); break;                                                  \
      default: SAXS_AZG(PP, 5); break;                                                 \
    }                                                                                  \
  } while (0)
    if (p.order == 4) SAXS_AZGH(4); else SAXS_AZGH(6);
#undef SAXS_AZGH
#undef SAXS_AZG
  } else if (adj_pruned) {
    Probe pr(pl, SAXS_K_ADJ_Z, st);
    dim3 ga(p.ntx * p.nty, n_frames);
    const bool zhalf = 2 * p.n[2] <= Nz, zexact = 2 * p.n[2] == Nz;
#define SAXS_AZP(PP, NN)                                                                       \
  do {                                                                                         \
    if (zexact) hipLaunchKernelGGL((saxs::k_adj_zp<PP, NN, true, true>), ga, dim3(saxs::kWG), lds_ap, st, p, b, g); \
    else if (zhalf) hipLaunchKernelGGL((saxs::k_adj_zp<PP, NN, true>), ga, dim3(saxs::kWG), lds_ap, st, p, b, g); \
    else hipLaunchKernelGGL((saxs::k_adj_zp<PP, NN, false>), ga, dim3(saxs::kWG), lds_ap, st, p, b, g);    \
  } while (0)
    if (p.order == 4) {
      if (Nz == 128) SAXS_AZP(4, 128); else if (Nz == 256) SAXS_AZP(4, 256); else SAXS_AZP(4, 512);
    } else {
      if (Nz == 128) SAXS_AZP(6, 128); else if (Nz == 256) SAXS_AZP(6, 256); else SAXS_AZP(6, 512);
    }
#undef SAXS_AZP
  } else {
    Probe pr(pl, SAXS_K_ADJ_Z, st);
    dim3 gz(p.ntx * p.nty, p.E, n_frames);
    const size_t lds = saxs::adjz_lds_bytes(p);
#define SAXS_AZ(PP, RS) hipLaunchKernelGGL((saxs::k_adj_z<PP, RS>), gz, dim3(saxs::kWG), lds, st, p, b, g)
    if (p.order == 4) {
      if (pz == 1) SAXS_AZ(4, 1); else if (pz == 0) SAXS_AZ(4, 0); else SAXS_AZ(4, 2);
    } else {
      if (pz == 1) SAXS_AZ(6, 1); else if (pz == 0) SAXS_AZ(6, 0); else SAXS_AZ(6, 2);
    }
#undef SAXS_AZ
  }
  if (n_atoms > 0) {
    Probe pr(pl, SAXS_K_GRAD_SUM, st);
    const int nfe = n_frames * p.E;
    constexpr int wpb = saxs::kWG / 64;    // k_pad_reduce: one wavefront per (frame, element)
    hipLaunchKernelGGL(saxs::k_pad_reduce, dim3((unsigned)((nfe + wpb - 1) / wpb)), dim3(saxs::kWG), 0, st,
                       pl->padt, pl->padg, p.ntx * p.nty, nfe);
    dim3 ga((n_atoms + saxs::kWG - 1) / saxs::kWG, n_frames);
    if (p.order == 4)
      hipLaunchKernelGGL(saxs::k_grad_finalize<4>, ga, dim3(saxs::kWG), 0, st, p, b, g);
    else
      hipLaunchKernelGGL(saxs::k_grad_finalize<6>, ga, dim3(saxs::kWG), 0, st, p, b, g);
  }
  HIP_TRY(hipGetLastError());
  return SAXS_OK;
}

}  // extern "C"

namespace {

// Full-spectrum geometry for the plain path and its buffers (allocated on
// first use).
int plain_setup(saxs_plan* pl) {
  if (pl->acc) return SAXS_OK;
  saxs::Params q = pl->p;
  for (int a = 0; a < 3; ++a) q.n[a] = q.nn[a];
  const int h = q.nn[2] / 2 + 1;
  q.kz = h;
  q.ky = q.nn[1];
  q.Ky = q.nn[1] / 2;
  q.kx = q.nn[0];
  q.Kx = q.nn[0] / 2;
  const size_t budget = 48 * 1024;
  int kcy = std::min({h, fft_line_cap(q.fy), (int)(budget / (sizeof(float2) * q.lsy))});
  int kcx = std::min({h, fft_line_cap(q.fx), (int)(budget / (sizeof(float2) * q.lsx))});
  if (kcy < 1 || kcx < 1) return fail(SAXS_EUNSUP, "plain path: FFT lines do not fit LDS");
  q.nchy = (h + kcy - 1) / kcy;
  q.kcy = (h + q.nchy - 1) / q.nchy;
  q.nchx = (h + kcx - 1) / kcx;
  q.kcx = (h + q.nchx - 1) / q.nchx;
  pl->plain_pairs = std::min(16, std::max(1, fft_line_cap(q.fz)));
  while (pl->plain_pairs > 1 &&
         (size_t)pl->plain_pairs * q.lsz * sizeof(float2) + q.nn[2] * sizeof(float2) > kLdsLimit)
    pl->plain_pairs /= 2;
  if (q.nn[1] % 2) return fail(SAXS_EUNSUP, "plain path needs an even nn_y");
  pl->pp = q;
  const size_t n3 = (size_t)pl->p.n[0] * pl->p.n[1] * pl->p.n[2];
  const size_t nn3 = (size_t)q.nn[0] * q.nn[1] * q.nn[2];
  const size_t nh = (size_t)q.nn[0] * q.nn[1] * h;
  size_t& tot = pl->bytes;
  int r;
  if ((r = dev_alloc(&pl->grid, n3, tot)) || (r = dev_alloc(&pl->sup, nn3, tot)) ||
      (r = dev_alloc(&pl->spec, nh, tot)) || (r = dev_alloc(&pl->yspec, nh, tot)) ||
      (r = dev_alloc(&pl->padp, 1, tot)) || (r = dev_alloc(&pl->acc, nh, tot)))
    return r;
  return SAXS_OK;
}

}  // namespace

extern "C" int saxs_forward_plain(saxs_plan* pl, const float* xyz, const int32_t* elem,
                                  int32_t n_atoms, const float* oc_diag, double* H, double* C,
                                  void* stream) {
  g_err.clear();
  int rc = check_call(pl, xyz, elem, n_atoms, 1, oc_diag);
  if (rc) return rc;
  if (!H || !C) return fail(SAXS_EINVAL, "H/C pointer is NULL");
  rc = plain_setup(pl);
  if (rc) return rc;
  pl->last_frames_valid = false;
  hipStream_t st = (hipStream_t)stream;
  saxs::Params p = call_params(pl, n_atoms, 1, oc_diag);     // cell geometry
  saxs::Params q = pl->pp;                                   // full grid geometry
  q.lut_len = p.lut_len;                                     // current q tables
  q.point_keys = p.point_keys;
  for (int a = 0; a < 3; ++a) q.kmax[a] = p.kmax[a];
  q.F = 1;
  q.N = n_atoms;
  for (int a = 0; a < 3; ++a) q.oc[a] = oc_diag[a];
  saxs::Buffers b = make_buffers(pl, xyz, elem);
  saxs::Buffers bq = b;
  bq.zbuf = pl->spec;
  bq.ybuf = pl->yspec;
  const int h = q.nn[2] / 2 + 1;
  const size_t n3 = (size_t)p.n[0] * p.n[1] * p.n[2];
  const size_t nn3 = (size_t)q.nn[0] * q.nn[1] * q.nn[2];
  const size_t nh = (size_t)q.nn[0] * q.nn[1] * h;
  const int pz = saxs::fft_radix_set(q.fz), py = saxs::fft_radix_set(q.fy),
            px = saxs::fft_radix_set(q.fx);
  Probe pr(pl, SAXS_K_PLAIN, st);
  hipLaunchKernelGGL(saxs::k_plain_fill, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, st,
                     (float*)nullptr, (size_t)0, pl->acc, nh, (double*)nullptr);
  for (int e = 0; e < p.E; ++e) {
    hipLaunchKernelGGL(saxs::k_plain_fill, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st,
                       pl->grid, n3, (float2*)nullptr, (size_t)0, pl->padp);
    if (n_atoms > 0) {
      dim3 ga((n_atoms + saxs::kWG - 1) / saxs::kWG);
      if (p.order == 4)
        hipLaunchKernelGGL(saxs::k_plain_spread<4>, ga, dim3(saxs::kWG), 0, st, p, xyz, elem, e,
                           pl->grid);
      else
        hipLaunchKernelGGL(saxs::k_plain_spread<6>, ga, dim3(saxs::kWG), 0, st, p, xyz, elem, e,
                           pl->grid);
    }
    hipLaunchKernelGGL(saxs::k_plain_pad, dim3(256), dim3(saxs::kWG), 0, st, p, pl->grid, pl->padp);
    hipLaunchKernelGGL(saxs::k_plain_embed, dim3((unsigned)((nn3 + 255) / 256)), dim3(256), 0, st,
                       p, pl->grid, pl->sup, pl->padp,
                       pl->pad_given_host.empty() ? -1.0 : pl->pad_given_host[e]);
    const int np = pl->plain_pairs;
    const size_t npairs = (size_t)q.nn[0] * (q.nn[1] / 2);
    const size_t lds_z = (size_t)np * q.lsz * sizeof(float2) + q.nn[2] * sizeof(float2);
    dim3 gz((unsigned)((npairs + np - 1) / np));
#define SAXS_PZ(RS) hipLaunchKernelGGL(saxs::k_plain_zfft<RS>, gz, dim3(saxs::kWG), lds_z, st, q, b, pl->sup, pl->spec, np)
    if (pz == 1) SAXS_PZ(1); else if (pz == 0) SAXS_PZ(0); else SAXS_PZ(2);
#undef SAXS_PZ
    dim3 gy(q.nn[0], q.nchy, 1);
#define SAXS_PY(RS) hipLaunchKernelGGL(saxs::k_yfft<RS>, gy, dim3(saxs::kWG), saxs::yfft_lds_bytes(q), st, q, bq)
    if (py == 1) SAXS_PY(1); else if (py == 0) SAXS_PY(0); else SAXS_PY(2);
#undef SAXS_PY
    dim3 gx(q.nn[1], q.nchx, 1);
    const size_t lds_x = (size_t)q.kcx * q.lsx * sizeof(float2) + q.nn[0] * sizeof(float2);
#define SAXS_PX(RS) hipLaunchKernelGGL(saxs::k_plain_xfft_scatter<RS>, gx, dim3(saxs::kWG), lds_x, st, q, b, pl->yspec, pl->acc, e)
    if (px == 1) SAXS_PX(1); else if (px == 0) SAXS_PX(0); else SAXS_PX(2);
#undef SAXS_PX
  }
  hipLaunchKernelGGL(saxs::k_plain_modulus, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, st, q,
                     b, pl->acc);
  hipLaunchKernelGGL(saxs::k_plain_hist, dim3(1024), dim3(saxs::kWG),
                     2 * q.num_bins * sizeof(double), st, q, b, pl->acc, H, C);
  HIP_TRY(hipGetLastError());
  return SAXS_OK;
}

extern "C" {

int saxs_plan_set_timing(saxs_plan* pl, int32_t enable) {
  if (!pl) return fail(SAXS_EINVAL, "plan is NULL");
  pl->timer.on = enable != 0;
  return SAXS_OK;
}

int saxs_plan_kernel_times(saxs_plan* pl, double* ms, int64_t* launches) {
  if (!pl || !ms || !launches) return fail(SAXS_EINVAL, "NULL argument");
  for (auto& r : pl->timer.rec) {
    HIP_TRY(hipEventSynchronize(r.second.second));
    float t = 0.0f;
    HIP_TRY(hipEventElapsedTime(&t, r.second.first, r.second.second));
    if (r.first >= 0 && r.first < SAXS_NKERNELS) {
      ms[r.first] += t;
      launches[r.first] += 1;
    }
    pl->timer.pool.push_back(r.second.first);
    pl->timer.pool.push_back(r.second.second);
  }
  pl->timer.rec.clear();
  return SAXS_OK;
}

int64_t saxs_plan_workspace_bytes(const saxs_plan* pl) { return pl ? (int64_t)pl->bytes : 0; }

int saxs_plan_kept(const saxs_plan* pl, int32_t* kx, int32_t* ky, int32_t* kz) {
  if (!pl || !kx || !ky || !kz) return fail(SAXS_EINVAL, "NULL argument");
  *kx = pl->p.kx;
  *ky = pl->p.ky;
  *kz = pl->p.kz;
  return SAXS_OK;
}

int saxs_plan_destroy(saxs_plan* pl) {
  if (pl) free_plan(pl);
  return SAXS_OK;
}

}  // extern "C"
