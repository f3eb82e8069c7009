// Frame reader for GROMACS XTC and TRR trajectories (C ABI in
// include/saxs_xdr.h).  Host code only: the file is memory-mapped, its
// frames are indexed once by walking the frame headers, and frames are
// decoded independently (each from its own offset), so a batch is split
// over host threads (kept for the handle's life) and written straight into
// the caller's buffers.  Reads on one handle are serialized.
//
// Formats (XDR: big-endian 4-byte words, opaque data padded to 4 bytes):
//
// TRR frame: int magic 1993; int slen (= strlen("GMX_trn_file") + 1);
//   XDR string (int length, bytes); ints ir_size, e_size, box_size,
//   vir_size, pres_size, top_size, sym_size, x_size, v_size, f_size,
//   natoms, step, nre; reals t, lambda; then box (9 reals), virial,
//   pressure, x, v, f, each present when its size is nonzero.  The real
//   width (4 or 8 bytes) is box_size / 9, else x_size / (3 natoms), else
//   v_size / (3 natoms), else f_size / (3 natoms).
//
// XTC frame: int magic 1995; int natoms; int step; float time; float
//   box[9]; int natoms again; then natoms <= 9: 3 natoms floats; else float
//   precision; int minint[3], maxint[3]; int smallidx; int nbytes; nbytes
//   of packed coordinates.  Coordinates are integers (nm times precision).
//   The first atom of each group is a full triplet (sizes maxint - minint
//   + 1 packed as one mixed-radix number, or three separate bit fields
//   when a size exceeds 2^24); a 1-bit flag then says whether a run length
//   follows (5 bits: run = code - code mod 3 atoms in the run, code mod 3 -
//   1 = change of the small-difference level).  The run's atoms are
//   triplets of differences from the previous atom, offset by half the
//   level's magic size and packed in `smallidx` bits; the first run atom is
//   emitted before the full atom it follows (water O-H-H order).  A run
//   length persists while the flag is 0.  The level table (kMagic) lists
//   the small-difference ranges; its entries are format constants.
//
// Positions and the box are returned in Angstrom as float32: (float)(int)
// times the float reciprocal of the precision (nm), then times 10.
#include "saxs_xdr.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

// Worker threads kept for the life of a handle (a batch of 16 frames would
// otherwise pay 16 thread creations per read).  run() hands out items
// 0..count-1 to the workers and the calling thread and returns when all are
// done; one run at a time per pool.
class Pool {
 public:
  ~Pool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    wake_.notify_all();
    for (auto& t : workers_) t.join();
  }

  void run(int nthreads, int count, const std::function<void(int)>& fn) {
    while ((int)workers_.size() < nthreads - 1) {
      const int index = (int)workers_.size();
      workers_.emplace_back([this, index] { loop(index); });
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      fn_ = &fn;
      count_ = count;
      next_ = 0;
      active_ = nthreads - 1;              // workers 0..active_-1 take items
      pending_ = (int)workers_.size();
      ++generation_;
    }
    wake_.notify_all();
    work();
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [this] { return pending_ == 0; });
    fn_ = nullptr;
  }

 private:
  void work() {
    for (int k = next_++; k < count_; k = next_++) (*fn_)(k);
  }
  void loop(int index) {
    unsigned long long seen = 0;
    for (;;) {
      bool take = false;
      {
        std::unique_lock<std::mutex> lk(mu_);
        wake_.wait(lk, [&] { return stop_ || generation_ != seen; });
        if (stop_) return;
        seen = generation_;
        take = index < active_;
      }
      if (take) work();
      std::lock_guard<std::mutex> lk(mu_);
      if (--pending_ == 0) done_.notify_all();
    }
  }

  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable wake_, done_;
  const std::function<void(int)>* fn_ = nullptr;
  int count_ = 0;
  std::atomic<int> next_{0};
  int pending_ = 0;
  int active_ = 0;
  unsigned long long generation_ = 0;
  bool stop_ = false;
};

}  // namespace

// Host threads that split one memory copy.  Between copies the workers spin
// on the job counter for about 0.2 ms before they block: the frame loop of
// a run hands over a frame every few tens of microseconds, and a blocking
// wake-up per frame costs about as much as the copy it would shorten.  The
// calling thread copies the first part itself and waits for the others.
struct saxs_copier {
  explicit saxs_copier(int threads) : nthreads(threads) {
    for (int i = 1; i < nthreads; ++i) workers.emplace_back([this, i] { loop(i); });
  }
  ~saxs_copier() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop.store(true);
    }
    cv.notify_all();
    for (auto& t : workers) t.join();
  }

  void part(int i) const {
    // parts of whole 64-byte lines, the last part takes the remainder
    const size_t lines = bytes / 64, per = lines / nthreads;
    const size_t b0 = (size_t)i * per * 64;
    const size_t b1 = i + 1 == nthreads ? bytes : b0 + per * 64;
    if (b1 > b0) std::memcpy(dst + b0, src + b0, b1 - b0);
  }

  void copy(char* d, const char* s_, size_t n) {
    if (nthreads == 1 || n < (size_t)nthreads * 4096) {
      if (n) std::memcpy(d, s_, n);
      return;
    }
    dst = d;
    src = s_;
    bytes = n;
    remaining.store(nthreads - 1, std::memory_order_relaxed);
    int sleeping;
    {
      std::lock_guard<std::mutex> lk(mu);
      job.fetch_add(1, std::memory_order_release);
      sleeping = sleepers;
    }
    if (sleeping) cv.notify_all();
    part(0);
    while (remaining.load(std::memory_order_acquire) > 0) std::this_thread::yield();
  }

  void loop(int index) {
    unsigned long long seen = 0;
    for (;;) {
      int spins = 0;
      while (job.load(std::memory_order_acquire) == seen && !stop.load()) {
        if (++spins < kSpin) {
          __builtin_ia32_pause();
        } else {
          std::unique_lock<std::mutex> lk(mu);
          ++sleepers;
          cv.wait(lk, [&] { return job.load() != seen || stop.load(); });
          --sleepers;
          spins = 0;
        }
      }
      if (stop.load()) return;
      seen = job.load(std::memory_order_acquire);
      part(index);
      remaining.fetch_sub(1, std::memory_order_release);
    }
  }

  static constexpr int kSpin = 5000;       // pause instructions, ~0.2 ms
  const int nthreads;
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv;
  std::atomic<unsigned long long> job{0};
  std::atomic<int> remaining{0};
  std::atomic<bool> stop{false};
  int sleepers = 0;                        // guarded by mu
  char* dst = nullptr;
  const char* src = nullptr;
  size_t bytes = 0;
};

struct saxs_xdr {
  int fd = -1;
  const unsigned char* data = nullptr;
  size_t size = 0;
  int kind = 0;
  int natoms = 0;
  std::vector<size_t> offsets;   // byte offset of each frame
  std::mutex read_mu;            // reads on one handle run one at a time
  Pool pool;                     // decoding threads, started on the first threaded read
};

namespace {

thread_local std::string g_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}

constexpr int kTrrMagic = 1993;
constexpr int kXtcMagic = 1995;

// Small-difference levels of the XTC packing (index = bits per triplet).
constexpr int kFirstIdx = 9;
constexpr unsigned kMagic[] = {
    0,       0,       0,       0,       0,        0,        0,        0,       0,
    8,       10,      12,      16,      20,       25,       32,       40,      50,
    64,      80,      101,     128,     161,      203,      256,      322,     406,
    512,     645,     812,     1024,    1290,     1625,     2048,     2580,    3250,
    4096,    5060,    6501,    8192,    10321,    13003,    16384,    20642,   26007,
    32768,   41285,   52015,   65536,   82570,    104031,   131072,   165140,  208063,
    262144,  330280,  416127,  524287,  660561,   832255,   1048576,  1321122, 1664510,
    2097152, 2642245, 3329021, 4194304, 5284491,  6658042,  8388607,  10568983, 13316085,
    16777216};
constexpr int kLastIdx = (int)(sizeof(kMagic) / sizeof(kMagic[0])) - 1;

// Bounds-checked big-endian cursor over one mapped frame.
struct Cursor {
  const unsigned char* p;
  size_t n;
  size_t o = 0;
  bool bad = false;

  bool has(size_t k) const { return !bad && o + k <= n; }
  unsigned u32() {
    if (!has(4)) {
      bad = true;
      return 0u;
    }
    const unsigned char* q = p + o;
    o += 4;
    return (unsigned)q[0] << 24 | (unsigned)q[1] << 16 | (unsigned)q[2] << 8 | (unsigned)q[3];
  }
  int i32() { return (int)u32(); }
  float f32() {
    const unsigned u = u32();
    float f;
    std::memcpy(&f, &u, 4);
    return f;
  }
  double f64() {
    const unsigned long long hi = u32(), lo = u32();
    const unsigned long long u = hi << 32 | lo;
    double d;
    std::memcpy(&d, &u, 8);
    return d;
  }
  double real(int width) { return width == 8 ? f64() : (double)f32(); }
  void skip(size_t k) {
    if (!has(k)) {
      bad = true;
      return;
    }
    o += k;
  }
};

size_t pad4(size_t k) { return (k + 3) & ~size_t(3); }

// ---- TRR --------------------------------------------------------------

struct TrrHeader {
  int ir, e, box, vir, pres, top, sym, x, v, f, natoms, step, nre;
  int width;       // bytes per real
  double t;
};

int trr_header(Cursor& c, TrrHeader& h) {
  if (c.i32() != kTrrMagic) return fail(SAXS_XDR_EFORMAT, "TRR frame at byte %zu: bad magic", c.o);
  const int slen = c.i32();
  const int len = c.i32();
  if (c.bad || slen < 1 || len < 0 || len > 1024)
    return fail(SAXS_XDR_EFORMAT, "TRR frame: bad version string");
  c.skip(pad4((size_t)len));
  int* fields[] = {&h.ir, &h.e, &h.box, &h.vir, &h.pres, &h.top, &h.sym,
                   &h.x,  &h.v, &h.f,   &h.natoms, &h.step, &h.nre};
  for (int* q : fields) *q = c.i32();
  if (c.bad) return fail(SAXS_XDR_EFORMAT, "TRR frame: truncated header");
  if (h.natoms < 0 || h.ir < 0 || h.e < 0 || h.box < 0 || h.vir < 0 || h.pres < 0 || h.top < 0 ||
      h.sym < 0 || h.x < 0 || h.v < 0 || h.f < 0)
    return fail(SAXS_XDR_EFORMAT, "TRR frame: negative size in header");
  const long long n3 = 3ll * h.natoms;
  int w = 0;
  if (h.box) w = h.box / 9;
  else if (h.x && n3) w = (int)(h.x / n3);
  else if (h.v && n3) w = (int)(h.v / n3);
  else if (h.f && n3) w = (int)(h.f / n3);
  if (w == 0 && !h.box && !h.x && !h.v && !h.f) w = 4;   // no data blocks
  if (w != 4 && w != 8) return fail(SAXS_XDR_EFORMAT, "TRR frame: real width %d", w);
  if ((h.box && h.box != 9 * w) || (h.vir && h.vir != 9 * w) || (h.pres && h.pres != 9 * w) ||
      (h.x && h.x != n3 * w) || (h.v && h.v != n3 * w) || (h.f && h.f != n3 * w))
    return fail(SAXS_XDR_EFORMAT, "TRR frame: block sizes disagree with %d atoms", h.natoms);
  h.width = w;
  h.t = c.real(w);
  (void)c.real(w);                                   // lambda
  if (c.bad) return fail(SAXS_XDR_EFORMAT, "TRR frame: truncated header");
  return SAXS_XDR_OK;
}

size_t trr_body_bytes(const TrrHeader& h) {
  return (size_t)h.ir + h.e + h.box + h.vir + h.pres + h.top + h.sym + h.x + h.v + h.f;
}

int trr_decode(const saxs_xdr* x, size_t off, size_t end, float* xyz, float* box, double* time,
               int64_t* step) {
  Cursor c{x->data + off, end - off};
  TrrHeader h{};
  int rc = trr_header(c, h);
  if (rc) return rc;
  c.skip((size_t)h.ir + h.e);
  if (box) {
    if (h.box) {
      for (int k = 0; k < 9; ++k) box[k] = (float)c.real(h.width) * 10.0f;
    } else {
      std::fill(box, box + 9, 0.0f);
    }
  } else {
    c.skip((size_t)h.box);
  }
  c.skip((size_t)h.vir + h.pres + h.top + h.sym);
  if (xyz) {
    if (!h.x) return fail(SAXS_XDR_EFORMAT, "TRR frame (step %d) holds no positions", h.step);
    const size_t n3 = 3 * (size_t)h.natoms;
    for (size_t k = 0; k < n3; ++k) xyz[k] = (float)c.real(h.width) * 10.0f;
  }
  if (c.bad) return fail(SAXS_XDR_EFORMAT, "TRR frame: truncated data");
  if (time) *time = h.t;
  if (step) *step = h.step;
  return SAXS_XDR_OK;
}

// ---- XTC --------------------------------------------------------------

struct XtcHeader {
  int natoms, step;
  float time;
  float box[9];
  size_t coord_off;   // offset (within the frame) of the coordinate block
  size_t bytes;       // frame length
};

int xtc_header(Cursor& c, XtcHeader& h) {
  if (c.i32() != kXtcMagic) return fail(SAXS_XDR_EFORMAT, "XTC frame at byte %zu: bad magic", c.o);
  h.natoms = c.i32();
  h.step = c.i32();
  h.time = c.f32();
  for (float& b : h.box) b = c.f32();
  h.coord_off = c.o;
  const int lsize = c.i32();
  if (c.bad) return fail(SAXS_XDR_EFORMAT, "XTC frame: truncated header");
  if (h.natoms < 0 || lsize != h.natoms)
    return fail(SAXS_XDR_EFORMAT, "XTC frame: atom counts %d and %d", h.natoms, lsize);
  if (h.natoms <= 9) {
    c.skip(12 * (size_t)h.natoms);
  } else {
    c.skip(4 + 12 + 12 + 4);                         // precision, minint, maxint, smallidx
    const int nbytes = c.i32();
    if (c.bad || nbytes < 0) return fail(SAXS_XDR_EFORMAT, "XTC frame: bad packed length");
    c.skip(pad4((size_t)nbytes));
  }
  if (c.bad) return fail(SAXS_XDR_EFORMAT, "XTC frame (step %d): truncated", h.step);
  h.bytes = c.o;
  return SAXS_XDR_OK;
}

// MSB-first bit reader over the packed bytes: a 64-bit window loaded at the
// current byte (big-endian) and shifted by the bit offset, so one call takes
// up to 57 bits; near the end of the data the window is assembled byte by
// byte with zeros past the end (reading past it sets `bad`).
struct Bits {
  const unsigned char* p;
  size_t n;
  size_t bit = 0;
  bool bad = false;

  unsigned long long window() const {
    const size_t byte = bit >> 3;
    unsigned long long w = 0;
    if (byte + 8 <= n) {
      std::memcpy(&w, p + byte, 8);
      w = __builtin_bswap64(w);
    } else {
      for (size_t k = 0; k < 8; ++k) w = (w << 8) | (byte + k < n ? p[byte + k] : 0u);
    }
    return w << (bit & 7);
  }
  unsigned long long get(int k) {   // 0 <= k <= 57
    if (k == 0) return 0ull;
    const unsigned long long v = window() >> (64 - k);
    bit += (size_t)k;
    if (bit > 8 * n) bad = true;
    return v;
  }
};

// Three integers packed as one mixed-radix number of `nbits` bits, stored
// as bytes least significant first (the last byte holds the top bits):
// value = (a0 sizes[1] + a1) sizes[2] + a2.  Up to 56 bits are read in one
// window and their bytes reordered; values below 2^32 or 2^64 take native
// divisions, wider ones 128-bit.
void get_triplet(Bits& br, int nbits, const unsigned sizes[3], int out[3]) {
  if (nbits <= 56) {
    const unsigned long long w = br.get(nbits);        // first byte in the top bits
    unsigned long long v = 0;
    int left = nbits, shift = 0;
    while (left > 8) {
      v |= ((w >> (left - 8)) & 0xffull) << shift;
      shift += 8;
      left -= 8;
    }
    v |= (w & ((1ull << left) - 1ull)) << shift;
    if (v <= 0xffffffffull) {
      unsigned u = (unsigned)v;
      out[2] = (int)(u % sizes[2]);
      u /= sizes[2];
      out[1] = (int)(u % sizes[1]);
      out[0] = (int)(u / sizes[1]);
      return;
    }
    out[2] = (int)(unsigned)(v % sizes[2]);
    v /= sizes[2];
    out[1] = (int)(unsigned)(v % sizes[1]);
    out[0] = (int)(unsigned long long)(v / sizes[1]);
    return;
  }
  unsigned __int128 v = 0;
  int shift = 0;
  while (nbits > 8) {
    v |= (unsigned __int128)br.get(8) << shift;
    shift += 8;
    nbits -= 8;
  }
  if (nbits > 0) v |= (unsigned __int128)br.get(nbits) << shift;
  out[2] = (int)(unsigned)(v % sizes[2]);
  v /= sizes[2];
  out[1] = (int)(unsigned)(v % sizes[1]);
  v /= sizes[1];
  out[0] = (int)(unsigned long long)v;
}

// Bits needed for values below `size` (one field).
int bits_of(unsigned size) {
  int k = 0;
  while (k < 32 && (1ull << k) <= (unsigned long long)size) ++k;
  return k;
}

// Bits of the packed triplet for sizes s0 s1 s2: the bit length of s0 s1 s2.
int bits_of3(const unsigned s[3]) {
  unsigned __int128 prod = (unsigned __int128)s[0] * s[1] * s[2];
  int k = 0;
  while (prod) {
    prod >>= 1;
    ++k;
  }
  return k;
}

int xtc_decode(const saxs_xdr* x, size_t off, size_t end, float* xyz, float* box, double* time,
               int64_t* step) {
  Cursor c{x->data + off, end - off};
  XtcHeader h{};
  int rc = xtc_header(c, h);
  if (rc) return rc;
  if (box)
    for (int k = 0; k < 9; ++k) box[k] = h.box[k] * 10.0f;
  if (time) *time = (double)h.time;
  if (step) *step = h.step;
  if (!xyz) return SAXS_XDR_OK;
  Cursor d{x->data + off + h.coord_off, h.bytes - h.coord_off};
  const int n = d.i32();
  if (n <= 9) {
    for (int k = 0; k < 3 * n; ++k) xyz[k] = d.f32() * 10.0f;
    return d.bad ? fail(SAXS_XDR_EFORMAT, "XTC frame: truncated coordinates") : SAXS_XDR_OK;
  }
  const float precision = d.f32();
  int minint[3], maxint[3];
  for (int& v : minint) v = d.i32();
  for (int& v : maxint) v = d.i32();
  int smallidx = d.i32();
  const int nbytes = d.i32();
  if (d.bad || !(precision > 0.0f))
    return fail(SAXS_XDR_EFORMAT, "XTC frame (step %d): bad precision", h.step);
  const float inv_precision = (float)(1.0 / (double)precision);
  unsigned sizeint[3];
  int bitsizeint[3] = {0, 0, 0};
  bool large = false;
  for (int k = 0; k < 3; ++k) {
    const long long s = (long long)maxint[k] - (long long)minint[k] + 1;
    if (s <= 0 || s > 0xffffffffll)
      return fail(SAXS_XDR_EFORMAT, "XTC frame (step %d): bad coordinate range", h.step);
    sizeint[k] = (unsigned)s;
    if (s > 0xffffff) large = true;
  }
  int bitsize = 0;
  if (large) {
    for (int k = 0; k < 3; ++k) bitsizeint[k] = bits_of(sizeint[k]);
  } else {
    bitsize = bits_of3(sizeint);
  }
  if (smallidx < kFirstIdx || smallidx > kLastIdx)
    return fail(SAXS_XDR_EFORMAT, "XTC frame (step %d): small level %d", h.step, smallidx);
  int smaller = (int)kMagic[std::max(kFirstIdx, smallidx - 1)] / 2;
  int smallnum = (int)kMagic[smallidx] / 2;
  unsigned sizesmall[3] = {kMagic[smallidx], kMagic[smallidx], kMagic[smallidx]};
  Bits br{x->data + off + h.coord_off + d.o, (size_t)nbytes};
  if (nbytes < 0 || d.o + (size_t)nbytes > d.n)
    return fail(SAXS_XDR_EFORMAT, "XTC frame (step %d): packed data past the frame", h.step);

  float* out = xyz;
  auto emit = [&](const int v[3]) {
    out[0] = ((float)v[0] * inv_precision) * 10.0f;
    out[1] = ((float)v[1] * inv_precision) * 10.0f;
    out[2] = ((float)v[2] * inv_precision) * 10.0f;
    out += 3;
  };
  int i = 0, run = 0;
  while (i < n) {
    int cur[3], prev[3];
    if (large) {
      for (int k = 0; k < 3; ++k) cur[k] = (int)br.get(bitsizeint[k]);
    } else {
      get_triplet(br, bitsize, sizeint, cur);
    }
    ++i;
    for (int k = 0; k < 3; ++k) {
      cur[k] += minint[k];
      prev[k] = cur[k];
    }
    int is_smaller = 0;
    if (br.get(1)) {
      run = (int)br.get(5);
      is_smaller = run % 3;
      run -= is_smaller;
      --is_smaller;
    }
    if (run > 0) {
      if (i + run / 3 > n)
        return fail(SAXS_XDR_EFORMAT, "XTC frame (step %d): run past the atom count", h.step);
      for (int k = 0; k < run; k += 3) {
        int s[3];
        get_triplet(br, smallidx, sizesmall, s);
        ++i;
        for (int a = 0; a < 3; ++a) s[a] += prev[a] - smallnum;
        if (k == 0) {
          // the run's first atom precedes the full atom in the output
          emit(s);
          emit(cur);
          for (int a = 0; a < 3; ++a) prev[a] = s[a];
        } else {
          emit(s);
          for (int a = 0; a < 3; ++a) prev[a] = s[a];
        }
      }
    } else {
      emit(cur);
    }
    smallidx += is_smaller;
    if (smallidx < kFirstIdx || smallidx > kLastIdx)
      return fail(SAXS_XDR_EFORMAT, "XTC frame (step %d): small level %d", h.step, smallidx);
    if (is_smaller < 0) {
      smallnum = smaller;
      smaller = smallidx > kFirstIdx ? (int)kMagic[smallidx - 1] / 2 : 0;
    } else if (is_smaller > 0) {
      smaller = smallnum;
      smallnum = (int)kMagic[smallidx] / 2;
    }
    sizesmall[0] = sizesmall[1] = sizesmall[2] = kMagic[smallidx];
    if (br.bad) return fail(SAXS_XDR_EFORMAT, "XTC frame (step %d): packed data ends early", h.step);
  }
  return SAXS_XDR_OK;
}

int decode_one(const saxs_xdr* x, int64_t frame, float* xyz, float* box, double* time,
               int64_t* step) {
  const size_t off = x->offsets[(size_t)frame];
  const size_t end = (size_t)frame + 1 < x->offsets.size() ? x->offsets[(size_t)frame + 1] : x->size;
  return x->kind == SAXS_XDR_XTC ? xtc_decode(x, off, end, xyz, box, time, step)
                                 : trr_decode(x, off, end, xyz, box, time, step);
}

}  // namespace

extern "C" {

int saxs_xdr_open(const char* path, saxs_xdr** out) {
  if (!path || !out) return fail(SAXS_XDR_EINVAL, "null argument");
  *out = nullptr;
  const int fd = ::open(path, O_RDONLY);
  if (fd < 0) return fail(SAXS_XDR_EIO, "cannot open %s", path);
  struct stat st;
  if (fstat(fd, &st) != 0 || st.st_size <= 0) {
    ::close(fd);
    return fail(SAXS_XDR_EFORMAT, "%s is empty or cannot be read", path);
  }
  const size_t size = (size_t)st.st_size;
  void* m = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
  if (m == MAP_FAILED) {
    ::close(fd);
    return fail(SAXS_XDR_EIO, "cannot map %s", path);
  }
  saxs_xdr* x = new saxs_xdr();
  x->fd = fd;
  x->data = static_cast<const unsigned char*>(m);
  x->size = size;
  Cursor first{x->data, size};
  const int magic = first.i32();
  if (magic == kXtcMagic) x->kind = SAXS_XDR_XTC;
  else if (magic == kTrrMagic) x->kind = SAXS_XDR_TRR;
  else {
    saxs_xdr_close(x);
    return fail(SAXS_XDR_EFORMAT, "%s: not an XTC or TRR file (magic %d)", path, magic);
  }
  size_t off = 0;
  while (off < size) {
    Cursor c{x->data + off, size - off};
    size_t len = 0;
    int natoms = 0, rc = 0;
    if (x->kind == SAXS_XDR_XTC) {
      XtcHeader h{};
      rc = xtc_header(c, h);
      len = h.bytes;
      natoms = h.natoms;
    } else {
      TrrHeader h{};
      rc = trr_header(c, h);
      natoms = h.natoms;
      if (!rc) {
        c.skip(trr_body_bytes(h));
        if (c.bad) rc = fail(SAXS_XDR_EFORMAT, "TRR frame %zu: truncated", x->offsets.size());
      }
      len = c.o;
    }
    if (rc) {
      const std::string msg = g_error;
      const size_t k = x->offsets.size();
      saxs_xdr_close(x);
      return fail(rc, "%s, frame %zu: %s", path, k, msg.c_str());
    }
    if (x->offsets.empty()) x->natoms = natoms;
    else if (natoms != x->natoms) {
      const size_t k = x->offsets.size();
      const int n0 = x->natoms;
      saxs_xdr_close(x);
      return fail(SAXS_XDR_EFORMAT, "%s: frame %zu has %d atoms, frame 0 has %d", path, k, natoms,
                  n0);
    }
    x->offsets.push_back(off);
    off += len;
  }
  *out = x;
  return SAXS_XDR_OK;
}

int saxs_xdr_info(const saxs_xdr* h, int* kind, int* n_atoms, int64_t* n_frames) {
  if (!h) return fail(SAXS_XDR_EINVAL, "null handle");
  if (kind) *kind = h->kind;
  if (n_atoms) *n_atoms = h->natoms;
  if (n_frames) *n_frames = (int64_t)h->offsets.size();
  return SAXS_XDR_OK;
}

int saxs_xdr_read(saxs_xdr* h, const int64_t* frames, int count, float* xyz, float* box,
                  double* time, int64_t* step, int threads) {
  if (!h || (count > 0 && !frames) || count < 0) return fail(SAXS_XDR_EINVAL, "bad argument");
  const int64_t nf = (int64_t)h->offsets.size();
  for (int k = 0; k < count; ++k)
    if (frames[k] < 0 || frames[k] >= nf)
      return fail(SAXS_XDR_EINVAL, "frame %lld out of range [0, %lld)", (long long)frames[k],
                  (long long)nf);
  std::lock_guard<std::mutex> serial(h->read_mu);
  const size_t n3 = 3 * (size_t)h->natoms;
  auto one = [&](int k) {
    return decode_one(h, frames[k], xyz ? xyz + k * n3 : nullptr, box ? box + 9 * (size_t)k : nullptr,
                      time ? time + k : nullptr, step ? step + k : nullptr);
  };
  const int nt = std::max(1, std::min(threads, count));
  if (nt == 1) {
    for (int k = 0; k < count; ++k) {
      const int rc = one(k);
      if (rc) return rc;
    }
    return SAXS_XDR_OK;
  }
  std::atomic<int> first_rc{0};
  std::mutex mu;
  std::string msg;
  const std::function<void(int)> item = [&](int k) {
    if (first_rc.load() != 0) return;
    const int rc = one(k);
    if (rc) {
      std::lock_guard<std::mutex> lock(mu);
      if (first_rc.load() == 0) {
        msg = g_error;
        first_rc = rc;
      }
    }
  };
  h->pool.run(nt, count, item);
  if (first_rc.load()) return fail(first_rc.load(), "%s", msg.c_str());
  return SAXS_XDR_OK;
}

const char* saxs_xdr_last_error(void) { return g_error.c_str(); }

int saxs_copier_create(int threads, saxs_copier** out) {
  if (!out) return fail(SAXS_XDR_EINVAL, "out is NULL");
  *out = nullptr;
  if (threads < 1 || threads > 64) return fail(SAXS_XDR_EINVAL, "threads %d outside [1, 64]", threads);
  *out = new saxs_copier(threads);
  return SAXS_XDR_OK;
}

int saxs_copier_copy(saxs_copier* c, void* dst, const void* src, size_t bytes) {
  if (!c || (bytes && (!dst || !src))) return fail(SAXS_XDR_EINVAL, "copier/dst/src is NULL");
  c->copy(static_cast<char*>(dst), static_cast<const char*>(src), bytes);
  return SAXS_XDR_OK;
}

void saxs_copier_destroy(saxs_copier* c) { delete c; }

void saxs_xdr_close(saxs_xdr* h) {
  if (!h) return;
  if (h->data) munmap(const_cast<unsigned char*>(h->data), h->size);
  if (h->fd >= 0) ::close(h->fd);
  delete h;
}

}  // extern "C"
