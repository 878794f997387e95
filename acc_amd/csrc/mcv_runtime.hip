// mcv_runtime.hip -- host runtime behind include/mcv_acc.h (libmcvacc.so).
//
// Owns device memory, the xoroshiro128+ stream seeding, module loading of the
// fused chain kernels the fuser compiled, launches and histogram read-back.
// Parameter packing keeps each component's reference parameter order and
// appends the few derived constants the device code wants (documented in
// mcv_components.h / mcv_samples.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mcv_acc.h"
#include "device/mcv_launch.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIPCHECK(expr)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) return fail(MCV_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// ------------------------------------------------------------------ stream seeding
// numba's create_xoroshiro128p_states (numba/cuda/random.py, numba 0.53.1):
// stream 0 = SplitMix64(seed) in both words, stream i = stream i-1 advanced by
// the 2^64-step jump.  The jump is linear over GF(2), so stream i = J^i s0; the
// device evaluates J^i as a product of precomputed J^(2^k) (one 128x128 bit
// matrix per set bit of i) instead of the reference's sequential host loop.
uint64_t splitmix64(uint64_t seed) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct V128 {
  uint64_t a, b;
};

inline V128 xnext(V128 s) {
  uint64_t s0 = s.a, s1 = s.b;
  s1 ^= s0;
  V128 r;
  r.a = ((s0 << 55) | (s0 >> 9)) ^ s1 ^ (s1 << 14);
  r.b = (s1 << 36) | (s1 >> 28);
  return r;
}

V128 xjump(V128 s) {
  static const uint64_t J[2] = {0xbeac0467eba5facbull, 0xd86b048b86aa9922ull};
  uint64_t a = 0, b = 0;
  for (int i = 0; i < 2; ++i)
    for (int bit = 0; bit < 64; ++bit) {
      if (J[i] & (1ull << bit)) {
        a ^= s.a;
        b ^= s.b;
      }
      s = xnext(s);
    }
  return V128{a, b};
}

// matrix = 128 columns (image of each basis bit: bits 0..63 of s0, then s1)
struct Mat {
  V128 col[128];
};

inline V128 apply(const Mat& M, V128 v) {
  V128 o{0, 0};
  for (int k = 0; k < 64; ++k)
    if ((v.a >> k) & 1) {
      o.a ^= M.col[k].a;
      o.b ^= M.col[k].b;
    }
  for (int k = 0; k < 64; ++k)
    if ((v.b >> k) & 1) {
      o.a ^= M.col[64 + k].a;
      o.b ^= M.col[64 + k].b;
    }
  return o;
}

constexpr int kJumpMats = 48;  // streams up to 2^48

const std::vector<uint64_t>& jump_matrices() {
  static std::vector<uint64_t> flat;
  static std::once_flag once;
  std::call_once(once, [] {
    Mat P;
    for (int k = 0; k < 128; ++k) {
      V128 e{k < 64 ? (1ull << k) : 0ull, k >= 64 ? (1ull << (k - 64)) : 0ull};
      P.col[k] = xjump(e);
    }
    flat.resize((size_t)kJumpMats * 256);
    for (int m = 0; m < kJumpMats; ++m) {
      for (int k = 0; k < 128; ++k) {
        flat[(size_t)m * 256 + 2 * k] = P.col[k].a;
        flat[(size_t)m * 256 + 2 * k + 1] = P.col[k].b;
      }
      Mat Q;
      for (int k = 0; k < 128; ++k) Q.col[k] = apply(P, P.col[k]);
      P = Q;
    }
  });
  return flat;
}

// Byte tables of the jump matrices for the device: for matrix m and byte j of
// the 128-bit state (j < 8: byte j of the first word, else byte j - 8 of the
// second), entry x (two words) = M applied to the state whose only set bits
// are x placed at byte j.  A product is then 16 table reads XORed together
// instead of 128 masked column XORs.  64 KB per matrix.
constexpr int kTabWords = 16 * 256 * 2;

const std::vector<uint64_t>& jump_byte_tables() {
  static std::vector<uint64_t> tabs;
  static std::once_flag once;
  std::call_once(once, [] {
    const auto& flat = jump_matrices();
    tabs.assign((size_t)kJumpMats * kTabWords, 0);
    for (int m = 0; m < kJumpMats; ++m) {
      const uint64_t* M = flat.data() + (size_t)m * 256;  // column c: M[2c], M[2c + 1]
      uint64_t* T = tabs.data() + (size_t)m * kTabWords;
      for (int j = 0; j < 16; ++j)
        for (int x = 1; x < 256; ++x) {
          const int low = __builtin_ctz((unsigned)x);
          const int c = 8 * j + low;  // column of the lowest set bit of x
          const size_t prev = ((size_t)j * 256 + (size_t)(x & (x - 1))) * 2, e = ((size_t)j * 256 + x) * 2;
          T[e] = T[prev] ^ M[2 * c];
          T[e + 1] = T[prev + 1] ^ M[2 * c + 1];
        }
    }
  });
  return tabs;
}

// state <- M state over GF(2), T the byte tables of M.
__device__ __forceinline__ void gf2_apply(const uint64_t* __restrict__ T, uint64_t& a, uint64_t& b) {
  const ulonglong2* __restrict__ E = reinterpret_cast<const ulonglong2*>(T);
  ulonglong2 r[16];
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = E[j * 256 + (int)((a >> (8 * j)) & 255u)];
#pragma unroll
  for (int j = 0; j < 8; ++j) r[8 + j] = E[(8 + j) * 256 + (int)((b >> (8 * j)) & 255u)];
  uint64_t oa = 0, ob = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    oa ^= r[j].x;
    ob ^= r[j].y;
  }
  a = oa;
  b = ob;
}

// Stream first + i's state is J^i applied to stream `first`'s state (a0, b0)
// (J: the 2^64-step jump; mats + k * kTabWords: the byte tables of J^(2^k)).  Thread t computes stream
// first + t from the set bits of t, then streams t + 2^k, t + 2 * 2^k, ...
// (kSeedChunk of them at most), each one product with J^(2^k) after the
// previous: most streams cost one matrix product instead of one per set bit
// of i, and a wavefront writes consecutive streams.  (Powers of J commute, so
// any order of the products gives J^i.)  Thread 0 also zeroes `nzero` launch
// counters, so the launch that follows on the same stream finds them cleared
// without a memset node.
constexpr int kSeedChunk = 8;

__global__ void __launch_bounds__(256) mcv_seed_streams(const uint64_t* __restrict__ mats, int k, uint64_t a0,
                                                        uint64_t b0, int64_t n, uint64_t* __restrict__ out,
                                                        unsigned long long* __restrict__ zero, int nzero) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0)
    for (int j = 0; j < nzero; ++j) zero[j] = 0;
  const int64_t step = int64_t(1) << k;
  if (t >= step || t >= n) return;
  uint64_t a = a0, b = b0;
  for (int bit = 0; bit < k; ++bit)
    if ((t >> bit) & 1) gf2_apply(mats + (size_t)bit * kTabWords, a, b);
  for (int64_t i = t; i < n; i += step) {
    out[2 * i] = a;
    out[2 * i + 1] = b;
    if (i + step < n) gf2_apply(mats + (size_t)k * kTabWords, a, b);
  }
}

// Scratch of mcv_instr_run_noalloc: the launch counters (processed neutrons,
// stream pool) in the first kScratchCounterWords words, then the states of up
// to kScratchStreams streams (two words each) -- 16 MiB + 64 B per device,
// held from the library's first use on a device until the process ends.  In
// host code `scratch` names a host-side stand-in; the device address is
// DeviceCtx::scratch, from hipGetSymbolAddress at the device's set-up.
constexpr int64_t kScratchCounterWords = 8;
constexpr int64_t kScratchStreams = int64_t(1) << 20;
static __device__ __align__(256) uint64_t scratch[kScratchCounterWords + 2 * kScratchStreams];

// Per-device state shared by every handle: the jump tables and one stream.
// All handles of a device queue their work on that one stream (calls are
// synchronous anyway): the runtime reserves a kernel's scratch (private
// memory) per hardware queue, and the multiple-scattering kernels need tens of
// KB of it per lane, so one queue per handle multiplied those reservations.
struct DeviceCtx {
  int device = -1;
  uint64_t* d_mats = nullptr;
  hipStream_t stream = nullptr;
  int cus = 1;                 // compute units
  int64_t scratch_budget = 0;  // bytes of lane scratch a launch may keep resident
  uint64_t* scratch = nullptr; // this device's copy of `scratch`
};

// Budget for the scratch (private segment) that the lanes of one launch keep
// resident: MCV_SCRATCH_BUDGET_MB from the environment, else an eighth of the
// device memory free when the device is first used.  The multiple-scattering
// kernels hold their neutron lists in lane scratch (25 KB per lane with the
// reference's 100-neutron lists); a launch whose resident lanes would need
// more is narrowed (persistent grids), split into several launches (the
// fused forward kernel), or refused with MCV_ERR_NOMEM, instead of leaving
// the runtime to fail the dispatch.
int64_t default_scratch_budget() {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    (void)hipGetLastError();
    return int64_t(4) << 30;
  }
  return (int64_t)(free_b / 8);
}

std::mutex g_ctx_mu;
std::deque<DeviceCtx> g_ctx;  // stable addresses across push_back

int device_ctx(DeviceCtx** out) {
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_ctx_mu);
  for (auto& c : g_ctx)
    if (c.device == dev) {
      *out = &c;
      return MCV_OK;
    }
  DeviceCtx c;
  c.device = dev;
  const auto& tabs = jump_byte_tables();
  HIPCHECK(hipMalloc(&c.d_mats, tabs.size() * sizeof(uint64_t)));
  HIPCHECK(hipMemcpy(c.d_mats, tabs.data(), tabs.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
  HIPCHECK(hipStreamCreate(&c.stream));
  HIPCHECK(hipDeviceGetAttribute(&c.cus, hipDeviceAttributeMultiprocessorCount, dev));
  c.scratch_budget = default_scratch_budget();
  // the no-allocation entry point's buffer: its address on this device (current here)
  HIPCHECK(hipGetSymbolAddress((void**)&c.scratch, HIP_SYMBOL(scratch)));
  if (!c.scratch) return fail(MCV_ERR_HIP, "no device address for the no-allocation scratch buffer");
  g_ctx.push_back(c);
  *out = &g_ctx.back();
  return MCV_OK;
}

int nbits_for(int64_t n) {
  int k = 0;
  while (k < 63 && (int64_t(1) << k) < n) ++k;
  return k;
}

// State of stream `first` of seed `seed`: J^first (z, z), z = SplitMix64(seed),
// from the host copy of the jump tables (one matrix-vector product per set bit).
V128 stream_state(uint64_t seed, int64_t first) {
  const uint64_t z = splitmix64(seed);
  V128 v{z, z};
  const auto& flat = jump_matrices();
  for (int m = 0; m < kJumpMats; ++m) {
    if (!((first >> m) & 1)) continue;
    const uint64_t* M = flat.data() + (size_t)m * 256;
    V128 o{0, 0};
    for (int bit = 0; bit < 64; ++bit)
      if ((v.a >> bit) & 1) {
        o.a ^= M[2 * bit];
        o.b ^= M[2 * bit + 1];
      }
    for (int bit = 0; bit < 64; ++bit)
      if ((v.b >> bit) & 1) {
        o.a ^= M[128 + 2 * bit];
        o.b ^= M[128 + 2 * bit + 1];
      }
    v = o;
  }
  return v;
}

// Seeds streams [first, first + n) into d_out (n x 2 words) on stream s, and
// zeroes nzero counters at `zero` first.  `ctx` must be the current device's.
int seed_streams(const DeviceCtx* ctx, hipStream_t s, uint64_t seed, int64_t first, int64_t n, uint64_t* d_out,
                 unsigned long long* zero = nullptr, int nzero = 0) {
  if (n <= 0 && nzero <= 0) return MCV_OK;
  if (first < 0 || nbits_for(first + std::max<int64_t>(n, 1)) > kJumpMats)
    return fail(MCV_ERR_ARG, "too many streams (%lld)", (long long)(first + n));
  const V128 s0 = stream_state(seed, first);
  // threads 0 .. 2^k - 1, each seeding up to kSeedChunk streams 2^k apart
  const int k = nbits_for((std::max<int64_t>(n, 1) + kSeedChunk - 1) / kSeedChunk);
  const int64_t threads = std::max<int64_t>(std::min<int64_t>(int64_t(1) << k, n), 1);
  const int bs = 256;
  const int64_t grid = (threads + bs - 1) / bs;
  hipLaunchKernelGGL(mcv_seed_streams, dim3((unsigned)grid), dim3(bs), 0, s, ctx->d_mats, k, s0.a, s0.b,
                     std::max<int64_t>(n, 0), d_out, zero, nzero);
  HIPCHECK(hipGetLastError());
  return MCV_OK;
}

// ------------------------------------------------------------------ packing
struct Packed {
  std::vector<double> params;
  std::vector<double> data;
  int64_t nbins = 0;
};

// Mirror-normal lengths the guides use at every reflection (guide.py:131-155,
// guide_tapering.py:236-259: nlen2 = l*l + w*w and sqrt(nlen2)), derived once
// here from the reference-order parameters.  Host IEEE sqrt is correctly
// rounded, so the values are the reference's.
void guide_norms(const double* P, double* out) {  // P = Guide params (ww, hh, ..., l at [4])
  const double l = P[4];
  const double l2 = l * l, w2 = P[0] * P[0], h2 = P[1] * P[1];
  out[0] = l2 + w2;
  out[1] = std::sqrt(out[0]);
  out[2] = l2 + h2;
  out[3] = std::sqrt(out[2]);
}

// Tapered guide: one 16-double record per segment from the component's
// 10 x nseg table (w1, w2, h1, h2, ww, hh, whalf, hhalf, lwhalf, lhhalf):
// [0] w1, [1] hhalf, [2] ww, [3] hh, [4] whalf * l_seg, [5] hhalf * l_seg
// (the products of cv1 and ch1, guide_tapering.py:199-201), [6] zr = seg *
// l_seg and [7] zr + l_seg (:168, :204), [8] nlen2 and [9] its sqrt for the
// vertical mirrors, [10] and [11] for the horizontal ones, [12] h1, [13] whalf,
// [14] 1 / [8] and [15] 1 / [10] (correctly rounded).
constexpr int64_t kTaperRec = 16;

void taper_record(double l_seg, int64_t seg, double w1, double h1, double ww, double hh, double whalf, double hhalf,
                  double* r) {
  const double l2 = l_seg * l_seg;
  const double w2 = ww * ww, h2 = hh * hh;
  r[0] = w1;
  r[1] = hhalf;
  r[2] = ww;
  r[3] = hh;
  r[4] = whalf * l_seg;
  r[5] = hhalf * l_seg;
  r[6] = (double)seg * l_seg;
  r[7] = r[6] + l_seg;
  r[8] = l2 + w2;
  r[9] = std::sqrt(r[8]);
  r[10] = l2 + h2;
  r[11] = std::sqrt(r[10]);
  r[12] = h1;
  r[13] = whalf;
  r[14] = 1.0 / r[8];  // for the bounce's 2 v.n / nlen2 (div_recip)
  r[15] = 1.0 / r[10];
}

void taper_records(double l_seg, int64_t nseg, const double* T, double* D) {  // T: 10 x nseg
  for (int64_t k = 0; k < nseg; ++k)
    taper_record(l_seg, k, T[k], T[2 * nseg + k], T[4 * nseg + k], T[5 * nseg + k], T[6 * nseg + k],
                 T[7 * nseg + k], D + kTaperRec * k);
}

// I(Q,E) monitor's angular window in the form the device test reads
// (mcv_components.h iqe_window), from P = Ei, Qmin, Qmax, Emin, Emax, max_in,
// min_in, max_out, min_out (degrees; acc/components/monitors/iqe_monitor.py:
// 93-101): out[0], out[1] = pseudo-angles of min_in and max_in, out[2..4] =
// kind (0 never, 1 always, 2 compare), cos, sin of the min_out test "el >
// min_out", out[5..7] the same for "el < max_out" (el in [0, 90]).
double iqe_pseudo_angle_host(double deg) {
  const double phi = deg * (M_PI / 180.0);
  const double vz = std::cos(phi), vx = std::sin(phi), s = std::fabs(vx) + std::fabs(vz);
  const double f = vx / s;
  return vz > 0 ? f : (vx > 0 ? 2.0 - f : -2.0 - f);
}

void iqe_window_params(const double* P, double* out) {
  const double max_in = P[5], min_in = P[6], max_out = P[7], min_out = P[8];
  // angle_in_plane lies in [-180, 180]: a > min_in always below -180, never from 180 on
  if (std::isnan(min_in)) out[0] = NAN;
  else if (min_in < -180.0) out[0] = -3.0;
  else if (min_in == -180.0) out[0] = -2.0;
  else if (min_in >= 180.0) out[0] = 3.0;
  else out[0] = iqe_pseudo_angle_host(min_in);
  if (std::isnan(max_in)) out[1] = NAN;
  else if (max_in > 180.0) out[1] = 3.0;
  else if (max_in == 180.0) out[1] = 2.0;
  else if (max_in <= -180.0) out[1] = -3.0;
  else out[1] = iqe_pseudo_angle_host(max_in);
  // el > min_out: always below 0, never from 90 on; el < max_out: never up to 0, always above 90
  const double lo = min_out * (M_PI / 180.0), hi = max_out * (M_PI / 180.0);
  out[2] = std::isnan(min_out) ? 0.0 : (min_out < 0.0 ? 1.0 : (min_out >= 90.0 ? 0.0 : 2.0));
  out[3] = std::cos(lo);
  out[4] = std::sin(lo);
  out[5] = std::isnan(max_out) ? 0.0 : (max_out <= 0.0 ? 0.0 : (max_out > 90.0 ? 1.0 : 2.0));
  out[6] = std::cos(hi);
  out[7] = std::sin(hi);
}

void cross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
void normalize3(double* v) {
  const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  const double s = 1.0 / l;
  v[0] *= s;
  v[1] *= s;
  v[2] *= s;
}

int expect(const mcv_comp_desc& d, int i, int np, int nip) {
  if (d.n_params != np || (np && !d.params))
    return fail(MCV_ERR_ARG, "component %d (type %d): expected %d params, got %d", i, d.type, np, d.n_params);
  if (d.n_iparams != nip || (nip && !d.iparams))
    return fail(MCV_ERR_ARG, "component %d (type %d): expected %d iparams, got %d", i, d.type, nip, d.n_iparams);
  return MCV_OK;
}

int pack(const mcv_comp_desc& d, int i, Packed& out) {
  auto copyp = [&](int np) { out.params.assign(d.params, d.params + np); };
  int rc;
  switch (d.type) {
    case MCV_ARM:
      return MCV_OK;
    case MCV_SOURCE_SIMPLE: {
      if ((rc = expect(d, i, 13, 0))) return rc;
      copyp(13);
      // randvec_target_rect's unit vectors (acc/components/sources/source_simple.py:25-34)
      const double target[3] = {0.0, 0.0, d.params[11]}, vertical[3] = {0.0, 1.0, 0.0};
      double p1[3], p2[3];
      cross3(target, vertical, p1);
      normalize3(p1);
      cross3(target, p1, p2);
      normalize3(p2);
      for (int k = 0; k < 3; ++k) out.params.push_back(p1[k]);
      for (int k = 0; k < 3; ++k) out.params.push_back(p2[k]);
      out.params.push_back(std::sqrt(target[0] * target[0] + target[1] * target[1] + target[2] * target[2]));
      return MCV_OK;
    }
    case MCV_SRC_IQE_NORMALIZER:  // Ei, Emin, Emax
      if ((rc = expect(d, i, 3, 0))) return rc;
      copyp(3);
      return MCV_OK;
    case MCV_SRC_STORAGE:  // P[0] = N; data = (N, 10)
      if ((rc = expect(d, i, 0, 1))) return rc;
      if (d.iparams[0] < 1 || d.n_data != 10 * d.iparams[0] || !d.data)
        return fail(MCV_ERR_ARG, "component %d: stored neutrons need (N >= 1, 10) data", i);
      out.params.push_back((double)d.iparams[0]);
      out.data.assign(d.data, d.data + d.n_data);
      return MCV_OK;
    case MCV_GUIDE:
      if ((rc = expect(d, i, 10, 0))) return rc;
      copyp(10);
      out.params.resize(14);
      guide_norms(out.params.data(), out.params.data() + 10);
      return MCV_OK;
    case MCV_SLIT:
    case MCV_BEAMSTOP:
      if ((rc = expect(d, i, 6, 0))) return rc;
      copyp(6);
      return MCV_OK;
    case MCV_MON_PSD_4PI:
    case MCV_MON_POSDIV: {
      if ((rc = expect(d, i, 5, 2))) return rc;
      if (d.iparams[0] < 1 || d.iparams[1] < 1 || d.n_bins != d.iparams[0] * d.iparams[1])
        return fail(MCV_ERR_ARG, "component %d: n_bins != n0*n1", i);
      copyp(5);
      out.params.push_back((double)d.iparams[0]);
      out.params.push_back((double)d.iparams[1]);
      out.nbins = d.n_bins;
      return MCV_OK;
    }
    case MCV_GUIDE_TAPERING: {
      if ((rc = expect(d, i, 9, 1))) return rc;
      const int64_t nseg = d.iparams[0];
      if (nseg < 1 || d.n_data != 10 * nseg || !d.data)
        return fail(MCV_ERR_ARG, "component %d: tapered guide needs a 10 x nseg table (nseg=%lld, n_data=%lld)", i,
                    (long long)nseg, (long long)d.n_data);
      copyp(9);
      out.params.push_back((double)nseg);
      out.data.assign(kTaperRec * nseg, 0.0);
      taper_records(d.params[0], nseg, d.data, out.data.data());
      return MCV_OK;
    }
    case MCV_SAMPLE_HSS:
    case MCV_SAMPLE_HMS: {
      if ((rc = expect(d, i, 16, 8))) return rc;
      copyp(16);
      for (int k = 0; k < 8; ++k) out.params.push_back((double)d.iparams[k]);
      const int64_t nn = d.iparams[0];
      if (nn < 1 || d.n_data < nn * MCV_CSG_NODE_SIZE || !d.data)
        return fail(MCV_ERR_ARG, "component %d: sample shape table too small", i);
      if (d.iparams[1] == MCV_KERNEL_SQE_GRID && d.iparams[5] + d.iparams[6] * d.iparams[7] > d.n_data)
        return fail(MCV_ERR_ARG, "component %d: S(Q,E) grid exceeds the data block", i);
      out.data.assign(d.data, d.data + d.n_data);
      return MCV_OK;
    }
    case MCV_MON_PSD: {
      if ((rc = expect(d, i, 4, 2))) return rc;
      if (d.n_bins != d.iparams[0] * d.iparams[1] || d.n_bins <= 0)
        return fail(MCV_ERR_ARG, "component %d: PSD n_bins != nx*ny", i);
      copyp(4);
      out.params.push_back((double)d.iparams[0]);
      out.params.push_back((double)d.iparams[1]);
      out.nbins = d.n_bins;
      return MCV_OK;
    }
    case MCV_MON_WAVELENGTH: {
      if ((rc = expect(d, i, 6, 1))) return rc;
      if (d.n_bins != d.iparams[0] || d.n_bins <= 0)
        return fail(MCV_ERR_ARG, "component %d: wavelength n_bins != nchan", i);
      copyp(6);
      out.params.push_back((double)d.iparams[0]);
      out.nbins = d.n_bins;
      return MCV_OK;
    }
    case MCV_MON_IQE: {
      if ((rc = expect(d, i, 9, 2))) return rc;
      if (d.n_bins != d.iparams[0] * d.iparams[1] || d.n_bins <= 0)
        return fail(MCV_ERR_ARG, "component %d: IQE n_bins != nQ*nE", i);
      copyp(9);
      out.params.push_back((double)d.iparams[0]);
      out.params.push_back((double)d.iparams[1]);
      out.params.resize(19);
      iqe_window_params(out.params.data(), out.params.data() + 11);
      out.nbins = d.n_bins;
      return MCV_OK;
    }
    case MCV_DET_TUBES: {
      if ((rc = expect(d, i, 8, 3))) return rc;
      if (d.n_bins != d.iparams[0] * d.iparams[1] * d.iparams[2] || d.n_bins <= 0)
        return fail(MCV_ERR_ARG, "component %d: detector n_bins != ntubes*npixels*ntof", i);
      copyp(8);
      for (int k = 0; k < 3; ++k) out.params.push_back((double)d.iparams[k]);
      out.nbins = d.n_bins;
      return MCV_OK;
    }
  }
  return fail(MCV_ERR_ARG, "component %d: unknown type %d", i, d.type);
}

int64_t align8(int64_t n) { return (n + 7) & ~int64_t(7); }

// Pack one instrument: blob = frame changes, then per-component parameters,
// then data tables (64-byte aligned blocks); offsets go into `args`.
// LDS budget (doubles) for privatised monitor histograms per workgroup:
// MCV_LDS_HIST_DOUBLES from the environment (0 turns privatisation off),
// default 2048 (16 KB).
int64_t lds_hist_budget() {
  const char* e = std::getenv("MCV_LDS_HIST_DOUBLES");
  if (!e || !*e) return 2048;
  const long long v = std::atoll(e);
  return v < 0 ? 0 : std::min<long long>(v, 8192);
}

// Replicas per global monitor histogram (MCV_HIST_REPLICAS from the
// environment, rounded down to a power of two; 0 or 1 turns replication off),
// default 128 (c1's direct-beam PSD, 1e7 neutrons: 6.94 ms unreplicated, 2.31
// with 64 copies, 1.92 with 128, 2.02 with 256), each monitor's replicas
// capped at 256 MB.
int hist_replicas() {
  const char* e = std::getenv("MCV_HIST_REPLICAS");
  long long v = (e && *e) ? std::atoll(e) : 128;
  if (v < 2) return 1;
  int r = 1;
  while ((long long)r * 2 <= std::min<long long>(v, 1024)) r *= 2;
  return r;
}
constexpr int64_t kReplicaBytes = int64_t(256) << 20;

// Slots of the LDS bin cache of plain kernels (MCV_LDS_CACHE_SLOTS from the
// environment, rounded down to a power of two, at least 64; 0 turns it off),
// default 4096 (113 KB of LDS per workgroup: one workgroup per CU).  Monitors
// of at most 2^20 bins that are not privatised whole use it.
int lds_cache_slots() {
  const char* e = std::getenv("MCV_LDS_CACHE_SLOTS");
  long long v = (e && *e) ? std::atoll(e) : 4096;
  if (v < 64) return 0;
  int s = 64;
  while ((long long)s * 2 <= std::min<long long>(v, 8192)) s *= 2;
  return s;
}
constexpr int64_t kCacheMaxBins = int64_t(1) << 20;
constexpr int64_t kLdsBytes = 160 * 1024;  // LDS of a gfx950 CU

// Dynamic LDS of a launch, in doubles: privatised histograms, then the bin
// cache (keys, three value rows, tile counters).
int64_t lds_doubles(const mcv::LaunchArgs& A) {
  if (A.cache_slots <= 0) return A.lds_total;
  return (int64_t)A.cache_off + A.cache_slots / 2 + 3 * (int64_t)A.cache_slots + (A.cache_ntiles + 2) / 2;
}

// Doubles of one workgroup's cache dump (keys, then N, sum p, sum p^2).
int64_t cache_dump_doubles(const mcv::LaunchArgs& A) {
  return A.cache_slots / 2 + 3 * (int64_t)A.cache_slots;
}

void no_cache(mcv::LaunchArgs& A) {
  A.cache_slots = 0;
  A.cache_off = 0;
  A.cmask = 0;
  A.cache_ntiles = 0;
  for (int i = 0; i < MCV_MAX_COMP; ++i) A.ckbase[i] = 0;
  A.cdump = nullptr;
  A.cdoff = nullptr;
}

int cache_tiles(int64_t nbins) { return (int)((nbins + MCV_CACHE_TILE - 1) >> MCV_CACHE_TILE_SHIFT); }

// Monitors of the bin cache, for the merge kernel: tile t of the dumps belongs
// to monitor comp[k] for base[k] <= t < base[k + 1], histogram at H + hoff[k].
struct CacheMergeArgs {
  double* H;
  const double* dump;
  const int32_t* doff;
  int32_t groups, slots, ntiles, nmon;
  int32_t split;  // workgroups per tile, each taking `per` dumps
  int32_t per;    // dumps per workgroup (<= 64)
  int32_t comp_base[MCV_MAX_COMP + 1];
  int64_t hoff[MCV_MAX_COMP], nbins[MCV_MAX_COMP];
};
constexpr int kMergeGroups = 8;   // dumps per merge workgroup of one launch's dumps
constexpr int kMergeBatch = 4;    // entries a thread loads before adding them

// `per` dumps per workgroup, `split` workgroups per tile of the cache's
// key space: a workgroup lays its dumps' runs for its tile end to end (a
// prefix sum over at most 64 runs), adds every entry into an LDS copy of the
// tile (LDS FP64 atomics, kMergeBatch independent loads per thread in
// flight), then the tile's non-empty bins into the histogram: with plain
// read-modify-writes when it is the tile's only workgroup, else with
// device-scope atomics (the fused kernel's direct adds are complete: the
// stream orders the launches).
__global__ __launch_bounds__(1024) void mcv_cache_merge(CacheMergeArgs M) {
  __shared__ double tile[3 * MCV_CACHE_TILE];
  __shared__ int run_lo[64], run_pre[65];
  const int t = (int)blockIdx.x / M.split, part = (int)blockIdx.x % M.split;
  const int g0 = part * M.per;
  int k = 0;
  while (k + 1 < M.nmon && M.comp_base[k + 1] <= t) ++k;
  for (int j = threadIdx.x; j < 3 * MCV_CACHE_TILE; j += blockDim.x) tile[j] = 0;
  if (threadIdx.x < 64) {
    const int lane = (int)threadIdx.x, g = g0 + lane;
    int lo = 0, n = 0;
    if (lane < M.per && g < M.groups) {
      const int32_t* off = M.doff + (int64_t)g * (M.ntiles + 1);
      lo = off[t];
      n = off[t + 1] - lo;
    }
    int incl = n;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(incl, o, 64);
      if (lane >= o) incl += u;
    }
    run_lo[lane] = lo;
    run_pre[lane + 1] = incl;
    if (lane == 0) run_pre[0] = 0;
  }
  __syncthreads();
  const int total = run_pre[64];
  const int64_t stride = M.slots / 2 + 3 * (int64_t)M.slots;
  for (int i0 = 0; i0 < total; i0 += kMergeBatch * (int)blockDim.x) {
    int bin[kMergeBatch];
    double v0[kMergeBatch], v1[kMergeBatch], v2[kMergeBatch];
#pragma unroll
    for (int u = 0; u < kMergeBatch; ++u) {
      const int i = i0 + u * (int)blockDim.x + (int)threadIdx.x;
      bin[u] = -1;
      if (i < total) {
        int lo = 0, hi = 64;  // run_pre[lo] <= i < run_pre[hi]: the last run starting at or before i
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (run_pre[mid] <= i) lo = mid; else hi = mid;
        }
        const int64_t e = (int64_t)(g0 + lo) * stride;
        const int j = run_lo[lo] + (i - run_pre[lo]);
        bin[u] = reinterpret_cast<const int32_t*>(M.dump + e)[j] & (MCV_CACHE_TILE - 1);
        const double* dval = M.dump + e + M.slots / 2;
        v0[u] = dval[j];
        v1[u] = dval[M.slots + j];
        v2[u] = dval[2 * M.slots + j];
      }
    }
#pragma unroll
    for (int u = 0; u < kMergeBatch; ++u) {
      if (bin[u] < 0) continue;
      atomicAdd(tile + bin[u], v0[u]);
      atomicAdd(tile + MCV_CACHE_TILE + bin[u], v1[u]);
      atomicAdd(tile + 2 * MCV_CACHE_TILE + bin[u], v2[u]);
    }
  }
  __syncthreads();
  const int64_t nb = M.nbins[k];
  const int64_t bin0 = (int64_t)(t - M.comp_base[k]) << MCV_CACHE_TILE_SHIFT;
  double* H = M.H + M.hoff[k];
  for (int j = threadIdx.x; j < MCV_CACHE_TILE; j += blockDim.x) {
    const int64_t b = bin0 + j;
    if (b >= nb || tile[j] == 0) continue;
    if (M.split == 1) {
      H[b] += tile[j];
      H[nb + b] += tile[MCV_CACHE_TILE + j];
      H[2 * nb + b] += tile[2 * MCV_CACHE_TILE + j];
    } else {
      unsafeAtomicAdd(H + b, tile[j]);
      unsafeAtomicAdd(H + nb + b, tile[MCV_CACHE_TILE + j]);
      unsafeAtomicAdd(H + 2 * nb + b, tile[2 * MCV_CACHE_TILE + j]);
    }
  }
}

// Monitors behind the bin cache take few global adds (the cache's misses and
// one flush per workgroup), so they are not replicated: drops their replicas
// and packs the others' offsets again.  Returns the replica doubles left.
int64_t unreplicate_cached(mcv::LaunchArgs& A, const std::vector<int64_t>& nbins) {
  int64_t rtot = 0;
  for (int i = 0; i < A.ncomp; ++i) {
    if (A.roff[i] < 0) continue;
    if ((A.cmask >> i) & 1u) {
      A.roff[i] = -1;
      A.rmask[i] = 0;
      continue;
    }
    A.roff[i] = rtot;
    rtot += (int64_t)(A.rmask[i] + 1) * 3 * nbins[i];
  }
  return rtot;
}

int build_layout(const mcv_comp_desc* comps, int32_t ncomp, const double* offsets, const double* rotmats,
                 mcv::LaunchArgs& args, std::vector<double>& B, std::vector<int64_t>& nbins, int64_t& hist_total,
                 std::vector<int>& types, std::vector<int>& nparams, bool replicate = false,
                 int64_t* rep_total = nullptr) {
  std::vector<Packed> packed(ncomp);
  for (int i = 0; i < ncomp; ++i) {
    int rc = pack(comps[i], i, packed[i]);
    if (rc) return rc;
    types.push_back(comps[i].type);
    nparams.push_back(comps[i].n_params);
  }
  B.clear();
  for (int i = 0; i + 1 < ncomp; ++i) {
    args.xoff[i] = (int64_t)B.size();
    for (int k = 0; k < 9; ++k) B.push_back(rotmats[9 * i + k]);
    for (int k = 0; k < 3; ++k) B.push_back(offsets[3 * i + k]);
    B.resize(align8((int64_t)B.size()), 0.0);
  }
  for (int i = 0; i < ncomp; ++i) {
    args.off[i] = (int64_t)B.size();
    B.insert(B.end(), packed[i].params.begin(), packed[i].params.end());
    B.resize(align8((int64_t)B.size()), 0.0);
  }
  for (int i = 0; i < ncomp; ++i) {
    args.doff[i] = (int64_t)B.size();
    B.insert(B.end(), packed[i].data.begin(), packed[i].data.end());
    B.resize(align8((int64_t)B.size()), 0.0);
  }
  if (B.empty()) B.resize(8, 0.0);
  hist_total = 0;
  for (int i = 0; i < ncomp; ++i) {
    args.hoff[i] = hist_total;
    nbins.push_back(packed[i].nbins);
    hist_total += 3 * packed[i].nbins;
  }
  // monitors kept in LDS by each workgroup: in chain order while they fit the budget
  const int64_t budget = lds_hist_budget();
  args.lds_nseg = 0;
  args.lds_total = 0;
  for (int i = 0; i < MCV_MAX_COMP; ++i) args.loff[i] = -1;
  for (int i = 0; i < ncomp; ++i) {
    const int64_t len = 3 * packed[i].nbins;
    if (len <= 0 || args.lds_nseg >= MCV_LDS_SEG || args.lds_total + len > budget) continue;
    args.loff[i] = args.lds_total;
    args.lds_loff[args.lds_nseg] = args.lds_total;
    args.lds_len[args.lds_nseg] = (int32_t)len;
    args.lds_hoff[args.lds_nseg] = args.hoff[i];
    args.lds_nseg += 1;
    args.lds_total += (int32_t)len;
  }
  // the other monitors: replicated in global memory while the copies fit
  args.R = nullptr;
  int64_t rtot = 0;
  const int reps = replicate ? hist_replicas() : 1;
  for (int i = 0; i < MCV_MAX_COMP; ++i) {
    args.roff[i] = -1;
    args.rmask[i] = 0;
  }
  for (int i = 0; i < ncomp && reps > 1; ++i) {
    const int64_t len = 3 * packed[i].nbins;
    if (len <= 0 || args.loff[i] >= 0) continue;
    int r = reps;
    while (r > 1 && len * (int64_t)sizeof(double) * r > kReplicaBytes) r /= 2;
    if (r < 2) continue;
    args.roff[i] = rtot;
    args.rmask[i] = r - 1;
    rtot += r * len;
  }
  if (rep_total) *rep_total = rtot;
  // the LDS bin cache: the monitors not kept whole in LDS, of at most 2^20
  // bins, in chain order while their tiles number at most MCV_CACHE_MAX_TILES
  no_cache(args);
  const int slots = lds_cache_slots();
  for (int i = 0; i < ncomp && slots > 0; ++i) {
    if (packed[i].nbins <= 0 || packed[i].nbins > kCacheMaxBins || args.loff[i] >= 0) continue;
    if (args.cache_ntiles + cache_tiles(packed[i].nbins) > MCV_CACHE_MAX_TILES) continue;
    args.cmask |= 1u << i;
    args.ckbase[i] = args.cache_ntiles;
    args.cache_ntiles += cache_tiles(packed[i].nbins);
  }
  if (args.cmask) {
    args.cache_slots = slots;
    args.cache_off = args.lds_total;
    while (args.cache_slots > 64 && lds_doubles(args) * (int64_t)sizeof(double) > kLdsBytes) args.cache_slots >>= 1;
  }
  args.ncomp = ncomp;
  return MCV_OK;
}

__global__ void mcv_zero_u64(unsigned long long* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// H[j] += sum over replicas r of R[r * len + j], replicas zeroed for the next launch
__global__ void mcv_fold_replicas(double* __restrict__ H, double* __restrict__ R, int64_t len, int reps) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < len; j += (int64_t)gridDim.x * blockDim.x) {
    double s = 0;
    for (int r = 0; r < reps; ++r) {
      s += R[r * len + j];
      R[r * len + j] = 0;
    }
    if (s != 0) H[j] += s;
  }
}

__global__ void mcv_zero_f64(double* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

}  // namespace

constexpr int kEvRing = 8;  // event triples of a run split into launches

struct mcv_instr {
  int ncomp = 0;
  int block = 256;  // workgroup size of the code object's kernels
  int spl = 1;      // random streams per lane (a workgroup serves block * spl streams)
  int pool = 0;     // the kernel takes streams from a pool (persistent grid)
  int slices = 0;   // the kernel runs slices of every stream, carrying the states (plain kernels)
  int64_t lane_scratch = 0;  // private segment bytes per lane of the fused kernel
  int per_cu = 0;       // its workgroups resident per CU
  DeviceCtx* ctx = nullptr;
  hipEvent_t rev[3 * kEvRing] = {};
  std::vector<int> types;
  std::vector<int> nparams;  // reference-order params per component (settable)
  mcv::LaunchArgs args{};
  std::vector<double> blob;
  std::vector<int64_t> nbins;
  int64_t hist_total = 0;
  double* d_blob = nullptr;
  double* d_hist = nullptr;
  double* d_rep = nullptr;  // monitor replicas (zero between launches)
  int64_t rep_total = 0;
  double* d_cdump = nullptr;  // bin-cache dumps of the resident workgroups
  int32_t* d_cdoff = nullptr;
  int64_t cdump_sets = 0;  // launches' worth of dumps allocated side by side
  uint64_t* d_rng = nullptr;
  int64_t rng_cap = 0;
  unsigned long long* d_counter = nullptr;
  double* d_dump = nullptr;
  int32_t* d_alive = nullptr;
  int64_t dump_first = 0, dump_count = 0;
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  // adjoint kernel (loaded on first use)
  hipModule_t adj_mod = nullptr;
  hipFunction_t adj_fn = nullptr;
  std::string adj_key;
  int adj_block = 256;
  int adj_spl = 1;
  int adj_pool = 0;
  int adj_ns = -1;  // score slots of the adjoint kernel
  int64_t adj_lane_scratch = 0;
  int adj_per_cu = 0;
  double* d_G = nullptr;
  double* d_grad = nullptr;
  int64_t grad_cap = 0;
  double* d_dense = nullptr;
  int64_t dense_cap = 0;
  // buffered-mode kernel (loaded on first use)
  hipModule_t proc_mod = nullptr;
  hipFunction_t proc_fn = nullptr;
  std::string proc_key;
  int proc_block = 256;
  int64_t proc_lane_scratch = 0;
  int proc_per_cu = 0;
  double* d_io = nullptr;
  int64_t io_cap = 0;
  double* d_io_out = nullptr;
  int64_t io_out_cap = 0;
};

namespace {

// Workgroup size a chain code object was built for (its `mcv_block_size`
// global; 256 for code objects without one).
int module_int(hipModule_t mod, const char* name, int dflt, int* out) {
  *out = dflt;
  hipDeviceptr_t p = nullptr;
  size_t sz = 0;
  if (hipModuleGetGlobal(&p, &sz, mod, name) != hipSuccess) {
    (void)hipGetLastError();
    return MCV_OK;
  }
  if (sz != sizeof(int)) return fail(MCV_ERR_MODULE, "%s has size %zu", name, sz);
  HIPCHECK(hipMemcpyDtoH(out, p, sizeof(int)));
  return MCV_OK;
}

// Workgroup size and streams per lane a chain code object was built for (its
// `mcv_block_size` / `mcv_streams_per_lane` globals; 256 and 1 when absent):
// one workgroup serves block * spl streams.
int module_block(hipModule_t mod, int* block, int* spl = nullptr) {
  int rc = module_int(mod, "mcv_block_size", 256, block);
  if (rc) return rc;
  if (*block < 64 || *block > 1024 || *block % 64)
    return fail(MCV_ERR_MODULE, "code object built for an unsupported workgroup size %d", *block);
  if (spl) {
    rc = module_int(mod, "mcv_streams_per_lane", 1, spl);
    if (rc) return rc;
    if (*spl < 1 || *spl > 16) return fail(MCV_ERR_MODULE, "code object built for %d streams per lane", *spl);
  }
  return MCV_OK;
}

// Scratch bytes per lane and resident workgroups per CU of a kernel at a
// dynamic LDS size (per_cu: the wave limit of a CU when the occupancy query
// fails).
void kernel_info(hipFunction_t fn, int block, size_t lds_bytes, int64_t* per_lane, int* per_cu) {
  int v = 0;
  if (hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, fn) != hipSuccess) {
    (void)hipGetLastError();
    v = 0;
  }
  *per_lane = v;
  int pc = 0;
  if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&pc, fn, block, lds_bytes) != hipSuccess || pc < 1) {
    (void)hipGetLastError();
    pc = std::max(1, 32 / std::max(1, block / 64));
  }
  *per_cu = pc;
}

// The scratch budget in force: MCV_SCRATCH_BUDGET_MB when set (read at every
// launch), else the device's default.
int64_t scratch_budget(const DeviceCtx* ctx) {
  const char* e = std::getenv("MCV_SCRATCH_BUDGET_MB");
  if (e && *e) return (int64_t)std::max<long long>(std::atoll(e), 0) << 20;
  return ctx->scratch_budget;
}

// Workgroups of `block` lanes needing `per_lane` bytes each that may be
// resident together within `budget` bytes, at most `resident`; 0 when not
// even one fits.
int64_t scratch_fit(int64_t per_lane, int64_t block, int64_t resident, int64_t budget) {
  if (resident <= 0) return 0;
  if (per_lane <= 0) return resident;
  return std::min<int64_t>(resident, budget / (per_lane * block));
}

// Grid of a launch of `needed` workgroups: persistent kernels (streams taken
// from a pool) run as many as the device holds at once, and any kernel as many
// as the scratch budget allows; 0 when not one fits (the caller fails).
// `persistent` = false: the grid is `needed` when its resident part fits, else 0.
int64_t launch_grid(const DeviceCtx* ctx, int64_t per_lane, int block, int per_cu, int64_t needed, bool persistent) {
  const int64_t resident = std::min<int64_t>(needed, (int64_t)ctx->cus * per_cu);
  const int64_t fit = scratch_fit(per_lane, block, resident, scratch_budget(ctx));
  if (persistent) return fit;
  return fit < resident ? 0 : needed;
}

int scratch_error(int64_t per_lane, int block, const DeviceCtx* ctx) {
  return fail(MCV_ERR_NOMEM,
              "kernel needs %lld bytes of scratch per lane; one workgroup of %d lanes exceeds the scratch budget "
              "of %lld bytes (MCV_SCRATCH_BUDGET_MB)",
              (long long)per_lane, block, (long long)scratch_budget(ctx));
}

void release(mcv_instr* h) {
  if (!h) return;
  if (h->d_blob) (void)hipFree(h->d_blob);
  if (h->d_hist) (void)hipFree(h->d_hist);
  if (h->d_rep) (void)hipFree(h->d_rep);
  if (h->d_cdump) (void)hipFree(h->d_cdump);
  if (h->d_cdoff) (void)hipFree(h->d_cdoff);
  if (h->d_rng) (void)hipFree(h->d_rng);
  if (h->d_counter) (void)hipFree(h->d_counter);
  if (h->d_dump) (void)hipFree(h->d_dump);
  if (h->d_alive) (void)hipFree(h->d_alive);
  for (auto& e : h->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : h->rev)
    if (e) (void)hipEventDestroy(e);
  if (h->mod) (void)hipModuleUnload(h->mod);
  if (h->adj_mod) (void)hipModuleUnload(h->adj_mod);
  if (h->proc_mod) (void)hipModuleUnload(h->proc_mod);
  if (h->d_io) (void)hipFree(h->d_io);
  if (h->d_io_out) (void)hipFree(h->d_io_out);
  if (h->d_G) (void)hipFree(h->d_G);
  if (h->d_grad) (void)hipFree(h->d_grad);
  if (h->d_dense) (void)hipFree(h->d_dense);
  delete h;
}

int create_impl(const mcv_comp_desc* comps, int32_t ncomp, const double* offsets, const double* rotmats,
                const char* code_object, const char* kernel_name, mcv_instr** out) {
  if (!comps || ncomp < 1) return fail(MCV_ERR_ARG, "empty component chain");
  if (ncomp > MCV_MAX_COMP) return fail(MCV_ERR_ARG, "chain longer than %d components", MCV_MAX_COMP);
  if (ncomp > 1 && (!offsets || !rotmats)) return fail(MCV_ERR_ARG, "missing frame transformations");
  if (!code_object || !kernel_name) return fail(MCV_ERR_ARG, "missing code object / kernel name");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(MCV_ERR_HIP, "no HIP device available");

  mcv_instr* h = new mcv_instr();
  *out = h;
  h->ncomp = ncomp;
  int rc = build_layout(comps, ncomp, offsets, rotmats, h->args, h->blob, h->nbins, h->hist_total, h->types,
                        h->nparams, true, &h->rep_total);
  if (rc) return rc;
  std::vector<double>& B = h->blob;

  {
    DeviceCtx* ctx = nullptr;
    int rc = device_ctx(&ctx);
    if (rc) return rc;
    h->stream = ctx->stream;  // shared, not owned
    h->ctx = ctx;
  }
  for (auto& e : h->ev) HIPCHECK(hipEventCreate(&e));
  for (auto& e : h->rev) HIPCHECK(hipEventCreate(&e));
  HIPCHECK(hipMalloc(&h->d_blob, B.size() * sizeof(double)));
  HIPCHECK(hipMemcpy(h->d_blob, B.data(), B.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHECK(hipMalloc(&h->d_hist, std::max<int64_t>(h->hist_total, 1) * sizeof(double)));
  HIPCHECK(hipMemset(h->d_hist, 0, std::max<int64_t>(h->hist_total, 1) * sizeof(double)));
  HIPCHECK(hipMalloc(&h->d_counter, 2 * sizeof(unsigned long long)));  // processed, stream pool

  hipError_t e = hipModuleLoad(&h->mod, code_object);
  if (e != hipSuccess) return fail(MCV_ERR_MODULE, "hipModuleLoad(%s): %s", code_object, hipGetErrorString(e));
  e = hipModuleGetFunction(&h->fn, h->mod, kernel_name);
  if (e != hipSuccess)
    return fail(MCV_ERR_MODULE, "kernel %s not in %s: %s", kernel_name, code_object, hipGetErrorString(e));
  rc = module_block(h->mod, &h->block, &h->spl);
  if (rc) return rc;
  rc = module_int(h->mod, "mcv_stream_pool", 0, &h->pool);
  if (rc) return rc;
  rc = module_int(h->mod, "mcv_slices", 0, &h->slices);
  if (rc) return rc;
  int cache = 0;  // plain kernels run the LDS bin cache (and a grid of resident workgroups)
  rc = module_int(h->mod, "mcv_lds_cache", 0, &cache);
  if (rc) return rc;
  if (!cache) no_cache(h->args);
  h->rep_total = unreplicate_cached(h->args, h->nbins);
  if (h->rep_total > 0) {
    HIPCHECK(hipMalloc(&h->d_rep, h->rep_total * sizeof(double)));
    HIPCHECK(hipMemset(h->d_rep, 0, h->rep_total * sizeof(double)));
  }
  kernel_info(h->fn, h->block, lds_doubles(h->args) * sizeof(double), &h->lane_scratch, &h->per_cu);
  if (h->args.cache_slots > 0) {  // one dump per resident workgroup (the cache's launch grid)
    const int64_t groups = (int64_t)h->ctx->cus * std::max(h->per_cu, 1);
    HIPCHECK(hipMalloc(&h->d_cdump, (size_t)(groups * cache_dump_doubles(h->args)) * sizeof(double)));
    HIPCHECK(hipMalloc(&h->d_cdoff, (size_t)(groups * (h->args.cache_ntiles + 1)) * sizeof(int32_t)));
    h->args.cdump = h->d_cdump;
    h->args.cdoff = h->d_cdoff;
    h->cdump_sets = 1;
  }
  return MCV_OK;
}

// A run in slices keeps the bin-cache dumps of up to kDumpSets launches side
// by side (set s of a launch grid of g workgroups at groups [s g, (s + 1) g))
// and merges them together: one merge per kDumpSets launches instead of one
// per launch, each of which pays its own ramp over a few MB of dumps.
constexpr int64_t kDumpSets = 16;
constexpr int64_t kDumpBytes = int64_t(1) << 30;  // all sets of a handle

// Sets for a run of `want` launches, within kDumpSets and kDumpBytes; grown
// when `may_grow` (a failed allocation keeps the sets there are), else the
// sets the handle has.
int dump_sets(mcv_instr* h, int64_t want, bool may_grow, int64_t& sets) {
  const int64_t groups = (int64_t)h->ctx->cus * std::max(h->per_cu, 1);
  const int64_t dd = groups * cache_dump_doubles(h->args), dn = groups * (h->args.cache_ntiles + 1);
  const int64_t set_bytes = dd * (int64_t)sizeof(double) + dn * (int64_t)sizeof(int32_t);
  want = std::max<int64_t>(1, std::min({want, kDumpSets, kDumpBytes / std::max<int64_t>(set_bytes, 1)}));
  if (may_grow && want > h->cdump_sets) {
    double* d = nullptr;
    int32_t* o = nullptr;
    if (hipMalloc(&d, (size_t)(want * dd) * sizeof(double)) == hipSuccess &&
        hipMalloc(&o, (size_t)(want * dn) * sizeof(int32_t)) == hipSuccess) {
      HIPCHECK(hipFree(h->d_cdump));
      HIPCHECK(hipFree(h->d_cdoff));
      h->d_cdump = d;
      h->d_cdoff = o;
      h->args.cdump = d;
      h->args.cdoff = o;
      h->cdump_sets = want;
    } else {
      if (d) (void)hipFree(d);
      (void)hipGetLastError();  // the run goes on with the sets there are
    }
  }
  sets = std::max<int64_t>(1, std::min(want, h->cdump_sets));
  return MCV_OK;
}

// Adds the bin-cache dumps of `groups` workgroups (one launch's, or several
// launches' side by side) into the histograms.  Dumps of several launches go
// up to 64 to a workgroup, so that the merge's workgroups stay a few per CU:
// each zeroes and scans a whole tile, whatever its dumps hold.
int merge_cache(mcv_instr* h, int64_t groups, bool sets = false) {
  const mcv::LaunchArgs& A = h->args;
  CacheMergeArgs M{};
  M.H = h->d_hist;
  M.dump = h->d_cdump;
  M.doff = h->d_cdoff;
  M.groups = (int32_t)groups;
  M.slots = A.cache_slots;
  M.ntiles = A.cache_ntiles;
  for (int i = 0; i < h->ncomp; ++i) {
    if (!((A.cmask >> i) & 1u)) continue;
    M.comp_base[M.nmon] = A.ckbase[i];
    M.hoff[M.nmon] = A.hoff[i];
    M.nbins[M.nmon] = h->nbins[i];
    M.nmon += 1;
  }
  M.comp_base[M.nmon] = A.cache_ntiles;
  M.per = kMergeGroups;
  if (sets)
    M.per = (int32_t)std::min<int64_t>(64, std::max<int64_t>(kMergeGroups, (groups * A.cache_ntiles + 4 * h->ctx->cus - 1) /
                                                                             (4 * (int64_t)h->ctx->cus)));
  M.split = (int32_t)std::max<int64_t>(1, (groups + M.per - 1) / M.per);
  hipLaunchKernelGGL(mcv_cache_merge, dim3((unsigned)(A.cache_ntiles * M.split)), dim3(1024), 0, h->stream, M);
  HIPCHECK(hipGetLastError());
  return MCV_OK;
}

void partition(uint64_t N, int64_t ntotalthreads, int32_t tpb, int64_t& nblocks, int64_t& nthreads, int64_t& npt) {
  if (ntotalthreads <= 0) ntotalthreads = 1000000;  // acc/config.py:1
  if (tpb <= 0) tpb = 512;                          // acc/config.py:2
  const int64_t ntot = (int64_t)std::min<uint64_t>(N, (uint64_t)ntotalthreads);
  nblocks = (ntot + tpb - 1) / tpb;
  nthreads = nblocks * tpb;
  npt = nthreads > 0 ? (int64_t)((N + (uint64_t)nthreads - 1) / (uint64_t)nthreads) : 0;
}

// Adds the replicated monitors' copies into their histograms (and zeroes them).
int fold_replicas(mcv_instr* h) {
  for (int i = 0; i < h->ncomp; ++i) {
    if (h->args.roff[i] < 0) continue;
    const int64_t len = 3 * h->nbins[i];
    const int64_t grid = std::min<int64_t>((len + 255) / 256, 4096);
    hipLaunchKernelGGL(mcv_fold_replicas, dim3((unsigned)grid), dim3(256), 0, h->stream, h->d_hist + h->args.hoff[i],
                       h->d_rep + h->args.roff[i], len, h->args.rmask[i] + 1);
    HIPCHECK(hipGetLastError());
  }
  return MCV_OK;
}

// Streams per launch of a run: all the streams that hold neutrons, or whole
// streams of about `buffer_size` neutrons (at least one stream) when given,
// then as few as the scratch budget allows resident at once.  A stream is
// never split: chunk c runs streams [c*S, (c+1)*S) of the one partition of
// ncount, each seeded as in a single launch, so a run in chunks computes
// every neutron exactly as one launch does (the reference's buffered engine
// instead reseeds identical streams for every buffer).  0 on a scratch error.
int64_t streams_per_launch(const mcv_instr* h, int64_t s_used, int64_t npt, uint64_t buffer_size) {
  int64_t S = s_used;
  if (buffer_size > 0) S = std::min<int64_t>(S, std::max<int64_t>(1, (int64_t)(buffer_size / (uint64_t)npt)));
  if (!h->pool) {
    const int64_t per_group = (int64_t)h->block * h->spl;
    const int64_t resident = std::min<int64_t>((S + per_group - 1) / per_group, (int64_t)h->ctx->cus * h->per_cu);
    const int64_t fit = scratch_fit(h->lane_scratch, h->block, resident, scratch_budget(h->ctx));
    if (fit <= 0) return 0;
    if (fit < resident) S = std::min<int64_t>(S, fit * per_group);
  }
  return S;
}

// Neutrons of every stream per launch when a run of a plain kernel is split
// by buffer_size: slices of about buffer_size / streams neutrons of every
// stream, the states carried in place from launch to launch, so that each
// launch is as wide as the whole run (ranges of whole streams would leave a
// launch of 1e6 neutrons at 10 per stream with 1e5 lanes).  0: split by
// stream ranges instead (other kernels, one slice would be the whole stream,
// or the states of all streams do not fit `max_states` or one launch).
int64_t slice_per_launch(const mcv_instr* h, int64_t s_used, int64_t npt, uint64_t buffer_size, uint64_t ncount,
                         int64_t max_states) {
  if (!h->slices || buffer_size == 0 || buffer_size >= ncount || npt < 2 || s_used > max_states) return 0;
  if (streams_per_launch(h, s_used, npt, 0) != s_used) return 0;
  const int64_t k = std::max<int64_t>(1, (int64_t)(buffer_size / (uint64_t)s_used));
  return k < npt ? k : 0;
}

// One run of ncount neutrons in launches over stream ranges or stream slices.
// `states(S, &p)` provides a device buffer for the states of S streams (at
// most `max_states`); `counter` holds the processed count and the stream-pool
// counter (zeroed by the first seeding).
template <class States>
int run_impl(mcv_instr* h, uint64_t ncount, uint64_t seed, int64_t ntotalthreads, int32_t tpb, uint32_t flags,
             uint64_t buffer_size, mcv_run_stats* stats, States states, unsigned long long* counter,
             int64_t max_states = INT64_MAX, bool may_alloc = true) {
  const auto t0 = std::chrono::steady_clock::now();
  int64_t nblocks, nthreads, npt;
  partition(ncount, ntotalthreads, tpb, nblocks, nthreads, npt);
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->nblocks = nblocks;
    stats->nthreads = nthreads;
    stats->npt = npt;
    stats->scratch_per_lane = h->lane_scratch;
  }
  if (ncount == 0) return MCV_OK;
  // streams past the last one holding neutrons run nothing
  const int64_t s_used = std::min<int64_t>(nthreads, (int64_t)((ncount + (uint64_t)npt - 1) / (uint64_t)npt));
  const int64_t slice = slice_per_launch(h, s_used, npt, buffer_size, ncount, max_states);
  const int64_t S = slice > 0 ? s_used : streams_per_launch(h, s_used, npt, buffer_size);
  if (S <= 0) return scratch_error(h->lane_scratch, h->block, h->ctx);
  uint64_t* d_states = nullptr;
  int rc = states(S, &d_states);
  if (rc) return rc;
  const int64_t nchunks = slice > 0 ? (npt + slice - 1) / slice : (s_used + S - 1) / S;
  const int64_t per_group = (int64_t)h->block * h->spl;
  // a persistent grid: as many workgroups as are resident at once within the scratch budget
  const int64_t pool_grid =
      h->pool ? launch_grid(h->ctx, h->lane_scratch, h->block, h->per_cu, (S + per_group - 1) / per_group, true) : 0;
  if (h->pool && pool_grid <= 0) return scratch_error(h->lane_scratch, h->block, h->ctx);
  const unsigned lds = (unsigned)(lds_doubles(h->args) * sizeof(double));
  // with the bin cache, a grid of the resident workgroups, each running many stream blocks
  const int64_t cache_grid = h->args.cache_slots > 0 ? (int64_t)h->ctx->cus * std::max(h->per_cu, 1) : 0;
  // slices: every launch has the same grid, so launches' dumps can share a merge
  int64_t nsets = 1;
  if (cache_grid > 0 && slice > 0 && nchunks > 1 && (rc = dump_sets(h, nchunks, may_alloc, nsets))) return rc;
  double ms_rng = 0, ms_k = 0;
  auto harvest = [&](int64_t c) -> int {
    hipEvent_t* e = h->rev + 3 * (c % kEvRing);
    HIPCHECK(hipEventSynchronize(e[2]));
    float a = 0, b = 0;
    HIPCHECK(hipEventElapsedTime(&a, e[0], e[1]));
    HIPCHECK(hipEventElapsedTime(&b, e[1], e[2]));
    ms_rng += a;
    ms_k += b;
    return MCV_OK;
  };
  for (int64_t c = 0; c < nchunks; ++c) {
    if (c >= kEvRing && (rc = harvest(c - kEvRing))) return rc;
    hipEvent_t* e = h->rev + 3 * (c % kEvRing);
    const int64_t s0 = slice > 0 ? 0 : c * S, s1 = std::min<int64_t>(s0 + S, s_used);
    HIPCHECK(hipEventRecord(e[0], h->stream));
    // the first launch also zeroes the processed count; every launch of
    // stream ranges seeds its streams and zeroes its stream pool; slices
    // continue the states the previous launch left
    if (slice == 0 || c == 0) {
      rc = seed_streams(h->ctx, h->stream, seed, s0, s1 - s0, d_states, c == 0 ? counter : counter + 1,
                        c == 0 ? 2 : 1);
      if (rc) return rc;
    }
    HIPCHECK(hipEventRecord(e[1], h->stream));
    mcv::LaunchArgs A = h->args;
    A.P = h->d_blob;
    A.H = h->d_hist;
    A.R = h->d_rep;
    A.rng = d_states;
    A.rng_first = s0;
    A.counter = counter;
    A.pool = h->pool ? counter + 1 : nullptr;
    A.dump = h->d_dump;
    A.dump_alive = h->d_alive;
    A.dump_first = h->dump_first;
    A.dump_count = h->d_dump ? h->dump_count : 0;
    A.N = (int64_t)ncount;
    A.nthreads = nthreads;
    A.npt = npt;
    A.tid_base = s0;
    A.tid_end = s1;
    A.slice0 = slice > 0 ? c * slice : 0;
    A.slice_len = slice;
    A.rng_out = slice > 0 ? d_states : nullptr;
    A.gate = (flags & MCV_FLAG_GATE_DEAD) ? 1 : 0;
    int64_t grid = (s1 - s0 + per_group - 1) / per_group;
    if (h->pool) grid = std::min(grid, pool_grid);
    if (cache_grid > 0) grid = std::min(grid, cache_grid);
    const int64_t set = c % nsets;
    if (cache_grid > 0) {
      A.cdump = h->d_cdump + set * grid * cache_dump_doubles(h->args);
      A.cdoff = h->d_cdoff + set * grid * (h->args.cache_ntiles + 1);
    }
    void* kargs[] = {&A};
    HIPCHECK(hipModuleLaunchKernel(h->fn, (unsigned)grid, 1, 1, (unsigned)h->block, 1, 1, lds, h->stream, kargs,
                                   nullptr));
    // the sets filled so far, before the next launch reuses the first
    if (cache_grid > 0 && (set + 1 == nsets || c + 1 == nchunks) &&
        (rc = merge_cache(h, grid * (set + 1), nsets > 1)))
      return rc;
    if (c + 1 == nchunks && (rc = fold_replicas(h))) return rc;
    HIPCHECK(hipEventRecord(e[2], h->stream));
  }
  HIPCHECK(hipStreamSynchronize(h->stream));
  for (int64_t c = std::max<int64_t>(0, nchunks - kEvRing); c < nchunks; ++c)
    if ((rc = harvest(c))) return rc;
  unsigned long long processed = 0;
  HIPCHECK(hipMemcpy(&processed, counter, sizeof(processed), hipMemcpyDeviceToHost));
  if (stats) {
    stats->processed = processed;
    stats->ms_rng_init = ms_rng;
    stats->ms_kernel = ms_k;
    stats->nchunks = nchunks;
    stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  if (processed != ncount)
    return fail(MCV_ERR_COUNT, "Processed neutron count %llu does not match requested neutron count %llu",
                (unsigned long long)processed, (unsigned long long)ncount);
  return MCV_OK;
}

}  // namespace

extern "C" {

int mcv_abi_version(void) { return MCV_ABI_VERSION; }

const char* mcv_last_error(void) { return g_last_error.c_str(); }

int mcv_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

mcv_instr* mcv_instr_create(const mcv_comp_desc* comps, int32_t ncomp, const double* offsets, const double* rotmats,
                            const char* code_object, const char* kernel_name) {
  mcv_instr* h = nullptr;
  int rc = create_impl(comps, ncomp, offsets, rotmats, code_object, kernel_name, &h);
  if (rc) {
    std::string msg = g_last_error;
    release(h);
    g_last_error = msg;
    return nullptr;
  }
  return h;
}

int mcv_instr_set_params(mcv_instr* h, int32_t comp, const double* params, int32_t n_params) {
  if (!h || comp < 0 || comp >= h->ncomp) return fail(MCV_ERR_ARG, "bad handle or component index");
  if (n_params != h->nparams[comp] || (n_params && !params))
    return fail(MCV_ERR_ARG, "component %d expects %d params", comp, h->nparams[comp]);
  if (h->types[comp] == MCV_SOURCE_SIMPLE)
    return fail(MCV_ERR_ARG, "Source_simple parameters are fixed at create time");
  const int64_t off = h->args.off[comp];
  std::memcpy(h->blob.data() + off, params, sizeof(double) * n_params);
  HIPCHECK(hipMemcpy(h->d_blob + off, params, sizeof(double) * n_params, hipMemcpyHostToDevice));
  // values derived from the parameters follow them
  if (h->types[comp] == MCV_GUIDE) {
    double* P = h->blob.data() + off;
    guide_norms(P, P + 10);
    HIPCHECK(hipMemcpy(h->d_blob + off + 10, P + 10, sizeof(double) * 4, hipMemcpyHostToDevice));
  } else if (h->types[comp] == MCV_GUIDE_TAPERING) {
    const int64_t nseg = (int64_t)h->blob[off + 9];
    double* D = h->blob.data() + h->args.doff[comp];
    for (int64_t k = 0; k < nseg; ++k) {
      double* r = D + kTaperRec * k;
      taper_record(params[0], k, r[0], r[12], r[2], r[3], r[13], r[1], r);
    }
    HIPCHECK(hipMemcpy(h->d_blob + h->args.doff[comp], D, sizeof(double) * kTaperRec * nseg, hipMemcpyHostToDevice));
  } else if (h->types[comp] == MCV_MON_IQE) {
    double* P = h->blob.data() + off;
    iqe_window_params(P, P + 11);
    HIPCHECK(hipMemcpy(h->d_blob + off + 11, P + 11, sizeof(double) * 8, hipMemcpyHostToDevice));
  }
  return MCV_OK;
}

int mcv_instr_run(mcv_instr* h, uint64_t ncount, uint64_t seed, int64_t ntotalthreads, int32_t tpb, uint32_t flags,
                  uint64_t buffer_size, mcv_run_stats* stats) {
  if (!h) return fail(MCV_ERR_ARG, "null handle");
  // stream states of one launch in the handle's buffer, grown on demand
  auto states = [h](int64_t streams, uint64_t** d_states) -> int {
    if (streams > h->rng_cap) {
      if (h->d_rng) HIPCHECK(hipFree(h->d_rng));
      h->d_rng = nullptr;
      h->rng_cap = 0;
      HIPCHECK(hipMalloc(&h->d_rng, (size_t)streams * 2 * sizeof(uint64_t)));
      h->rng_cap = streams;
    }
    *d_states = h->d_rng;
    return MCV_OK;
  };
  return run_impl(h, ncount, seed, ntotalthreads, tpb, flags, buffer_size, stats, states, h->d_counter);
}

int mcv_instr_run_noalloc(mcv_instr* h, uint64_t ncount, uint64_t seed, int64_t ntotalthreads, int32_t tpb,
                          uint32_t flags, uint64_t buffer_size, mcv_run_stats* stats) {
  if (!h) return fail(MCV_ERR_ARG, "null handle");
  // stream states and counters in the device's static scratch: nothing is allocated
  uint64_t* base = h->ctx->scratch;
  auto states = [base](int64_t streams, uint64_t** d_states) -> int {
    if (streams > kScratchStreams)
      return fail(MCV_ERR_ARG,
                  "mcv_instr_run_noalloc: a launch of %lld streams exceeds the scratch buffer's %lld streams; "
                  "split the run with a buffer_size of at most %lld streams' worth of neutrons",
                  (long long)streams, (long long)kScratchStreams, (long long)kScratchStreams);
    *d_states = base + kScratchCounterWords;
    return MCV_OK;
  };
  return run_impl(h, ncount, seed, ntotalthreads, tpb, flags, buffer_size, stats, states,
                  reinterpret_cast<unsigned long long*>(base), kScratchStreams, false);
}

int mcv_monitor_read(mcv_instr* h, int32_t comp, double* out3) {
  if (!h || comp < 0 || comp >= h->ncomp) return fail(MCV_ERR_ARG, "bad handle or component index");
  if (h->nbins[comp] <= 0) return fail(MCV_ERR_ARG, "component %d is not a monitor", comp);
  HIPCHECK(hipMemcpy(out3, h->d_hist + h->args.hoff[comp], 3 * h->nbins[comp] * sizeof(double),
                     hipMemcpyDeviceToHost));
  return MCV_OK;
}

int mcv_monitor_clear(mcv_instr* h) {
  if (!h) return fail(MCV_ERR_ARG, "null handle");
  if (h->hist_total > 0) HIPCHECK(hipMemset(h->d_hist, 0, h->hist_total * sizeof(double)));
  return MCV_OK;
}

int mcv_instr_set_dump(mcv_instr* h, int64_t first, int64_t count) {
  if (!h || first < 0 || count < 0) return fail(MCV_ERR_ARG, "bad dump range");
  if (h->d_dump) HIPCHECK(hipFree(h->d_dump));
  if (h->d_alive) HIPCHECK(hipFree(h->d_alive));
  h->d_dump = nullptr;
  h->d_alive = nullptr;
  h->dump_first = first;
  h->dump_count = count;
  if (count == 0) return MCV_OK;
  HIPCHECK(hipMalloc(&h->d_dump, (size_t)count * 10 * sizeof(double)));
  HIPCHECK(hipMalloc(&h->d_alive, (size_t)count * sizeof(int32_t)));
  HIPCHECK(hipMemset(h->d_dump, 0, (size_t)count * 10 * sizeof(double)));
  HIPCHECK(hipMemset(h->d_alive, 0xff, (size_t)count * sizeof(int32_t)));
  return MCV_OK;
}

int mcv_instr_adjoint(mcv_instr* h, const char* code_object, const char* kernel_name, uint64_t ncount,
                      uint64_t seed, int64_t ntotalthreads, int32_t tpb, uint32_t flags, const double* g_hist,
                      int32_t nscore, double* grad, int64_t ndense, double* grad_dense, mcv_run_stats* stats) {
  if (!h || !code_object || !kernel_name || !g_hist || nscore < 1 || !grad)
    return fail(MCV_ERR_ARG, "bad adjoint arguments");
  if (ndense < 0 || (ndense > 0 && !grad_dense)) return fail(MCV_ERR_ARG, "bad dense gradient buffer");
  // the dense gradient is the S(Q,E) grid of the chain's grid-kernel sample: its size must match
  int64_t grid_nodes = 0;
  for (int c = 0; c < h->ncomp; ++c)
    if (h->types[c] == MCV_SAMPLE_HSS && (int64_t)h->blob[h->args.off[c] + 17] == MCV_KERNEL_SQE_GRID)
      grid_nodes = (int64_t)h->blob[h->args.off[c] + 22] * (int64_t)h->blob[h->args.off[c] + 23];
  if (ndense != 0 && ndense != grid_nodes)
    return fail(MCV_ERR_ARG, "dense gradient of %lld values for a grid of %lld nodes", (long long)ndense,
                (long long)grid_nodes);
  const auto t0 = std::chrono::steady_clock::now();
  const std::string key = std::string(code_object) + "\n" + kernel_name;
  if (key != h->adj_key) {
    if (h->adj_mod) (void)hipModuleUnload(h->adj_mod);
    h->adj_mod = nullptr;
    h->adj_fn = nullptr;
    h->adj_key.clear();
    hipError_t e = hipModuleLoad(&h->adj_mod, code_object);
    if (e != hipSuccess) {
      h->adj_mod = nullptr;
      return fail(MCV_ERR_MODULE, "hipModuleLoad(%s): %s", code_object, hipGetErrorString(e));
    }
    if (hipModuleGetFunction(&h->adj_fn, h->adj_mod, kernel_name) != hipSuccess)
      return fail(MCV_ERR_MODULE, "kernel %s not in %s", kernel_name, code_object);
    int rc = module_block(h->adj_mod, &h->adj_block, &h->adj_spl);
    if (!rc) rc = module_int(h->adj_mod, "mcv_stream_pool", 0, &h->adj_pool);
    if (!rc) rc = module_int(h->adj_mod, "mcv_adjoint_nscore", -1, &h->adj_ns);
    if (rc) return rc;
    kernel_info(h->adj_fn, h->adj_block, 0, &h->adj_lane_scratch, &h->adj_per_cu);
    h->adj_key = key;
  }
  // the kernel adds into grad[0 .. NS): the caller's buffer must be exactly that long
  if (h->adj_ns < 0) return fail(MCV_ERR_MODULE, "adjoint code object %s has no mcv_adjoint_nscore", code_object);
  if (nscore != h->adj_ns)
    return fail(MCV_ERR_ARG, "adjoint kernel has %d score slots, caller passed nscore=%d", h->adj_ns, nscore);
  int64_t nblocks, nthreads, npt;
  partition(ncount, ntotalthreads, tpb, nblocks, nthreads, npt);
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->nblocks = nblocks;
    stats->nthreads = nthreads;
    stats->npt = npt;
  }
  if (!h->d_G) HIPCHECK(hipMalloc(&h->d_G, std::max<int64_t>(h->hist_total, 1) * sizeof(double)));
  if (nscore > h->grad_cap) {
    if (h->d_grad) HIPCHECK(hipFree(h->d_grad));
    h->d_grad = nullptr;
    HIPCHECK(hipMalloc(&h->d_grad, (size_t)nscore * sizeof(double)));
    h->grad_cap = nscore;
  }
  if (ndense > h->dense_cap) {
    if (h->d_dense) HIPCHECK(hipFree(h->d_dense));
    h->d_dense = nullptr;
    HIPCHECK(hipMalloc(&h->d_dense, (size_t)ndense * sizeof(double)));
    h->dense_cap = ndense;
  }
  if (nthreads > h->rng_cap) {
    if (h->d_rng) HIPCHECK(hipFree(h->d_rng));
    h->d_rng = nullptr;
    HIPCHECK(hipMalloc(&h->d_rng, (size_t)nthreads * 2 * sizeof(uint64_t)));
    h->rng_cap = nthreads;
  }
  if (h->hist_total > 0)
    HIPCHECK(hipMemcpyAsync(h->d_G, g_hist, h->hist_total * sizeof(double), hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(mcv_zero_f64, dim3(1), dim3(256), 0, h->stream, h->d_grad, (int64_t)nscore);
  if (ndense > 0)
    hipLaunchKernelGGL(mcv_zero_f64, dim3((unsigned)std::min<int64_t>((ndense + 255) / 256, 4096)), dim3(256), 0,
                       h->stream, h->d_dense, ndense);
  hipLaunchKernelGGL(mcv_zero_u64, dim3(1), dim3(64), 0, h->stream, h->d_counter, (int64_t)2);
  HIPCHECK(hipGetLastError());
  if (ncount > 0) {
    HIPCHECK(hipEventRecord(h->ev[0], h->stream));
    int rc = seed_streams(h->ctx, h->stream, seed, 0, nthreads, h->d_rng);
    if (rc) return rc;
    HIPCHECK(hipEventRecord(h->ev[1], h->stream));
    mcv::LaunchArgs A = h->args;
    A.P = h->d_blob;
    A.H = nullptr;
    A.rng_first = 0;
    A.tid_end = nthreads;
    A.rng = h->d_rng;
    A.counter = h->d_counter;
    A.pool = h->adj_pool ? h->d_counter + 1 : nullptr;
    A.dump = nullptr;
    A.dump_alive = nullptr;
    A.dump_count = 0;
    A.N = (int64_t)ncount;
    A.nthreads = nthreads;
    A.npt = npt;
    A.tid_base = 0;
    A.gate = (flags & MCV_FLAG_GATE_DEAD) ? 1 : 0;
    A.lds_nseg = 0;  // the adjoint kernel fills no histograms
    A.lds_total = 0;
    no_cache(A);
    mcv::AdjointArgs J{h->d_G, h->d_grad, ndense > 0 ? h->d_dense : nullptr};
    const unsigned bs = (unsigned)h->adj_block;
    const int64_t per_group = (int64_t)h->adj_block * h->adj_spl;
    int64_t grid = (nthreads + per_group - 1) / per_group;
    grid = launch_grid(h->ctx, h->adj_lane_scratch, h->adj_block, h->adj_per_cu, grid, h->adj_pool != 0);
    if (grid <= 0) return scratch_error(h->adj_lane_scratch, h->adj_block, h->ctx);
    void* kargs[] = {&A, &J};
    HIPCHECK(hipModuleLaunchKernel(h->adj_fn, (unsigned)grid, 1, 1, bs, 1, 1, 0, h->stream, kargs, nullptr));
    HIPCHECK(hipEventRecord(h->ev[2], h->stream));
  }
  HIPCHECK(hipStreamSynchronize(h->stream));
  unsigned long long processed = 0;
  HIPCHECK(hipMemcpy(&processed, h->d_counter, sizeof(processed), hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(grad, h->d_grad, (size_t)nscore * sizeof(double), hipMemcpyDeviceToHost));
  if (ndense > 0) HIPCHECK(hipMemcpy(grad_dense, h->d_dense, (size_t)ndense * sizeof(double), hipMemcpyDeviceToHost));
  if (stats) {
    stats->processed = processed;
    if (ncount > 0) {
      float ms_rng = 0, ms_k = 0;
      HIPCHECK(hipEventElapsedTime(&ms_rng, h->ev[0], h->ev[1]));
      HIPCHECK(hipEventElapsedTime(&ms_k, h->ev[1], h->ev[2]));
      stats->ms_rng_init = ms_rng;
      stats->ms_kernel = ms_k;
    }
    stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  if (processed != ncount)
    return fail(MCV_ERR_COUNT, "Processed neutron count %llu does not match requested neutron count %llu",
                (unsigned long long)processed, (unsigned long long)ncount);
  return MCV_OK;
}

int mcv_instr_process(mcv_instr* h, const char* code_object, const char* kernel_name, double* neutrons, int64_t n,
                      uint64_t seed, int64_t ntotalthreads, int32_t tpb, uint32_t flags, double* out_ms,
                      int32_t m, mcv_run_stats* stats) {
  if (!h || !code_object || !kernel_name || n < 0 || (n > 0 && !neutrons) || m < 0 || (m > 0 && !out_ms))
    return fail(MCV_ERR_ARG, "bad buffered-mode arguments");
  const auto t0 = std::chrono::steady_clock::now();
  const std::string key = std::string(code_object) + "\n" + kernel_name;
  if (key != h->proc_key) {
    if (h->proc_mod) (void)hipModuleUnload(h->proc_mod);
    h->proc_mod = nullptr;
    h->proc_fn = nullptr;
    h->proc_key.clear();
    hipError_t e = hipModuleLoad(&h->proc_mod, code_object);
    if (e != hipSuccess) {
      h->proc_mod = nullptr;
      return fail(MCV_ERR_MODULE, "hipModuleLoad(%s): %s", code_object, hipGetErrorString(e));
    }
    if (hipModuleGetFunction(&h->proc_fn, h->proc_mod, kernel_name) != hipSuccess)
      return fail(MCV_ERR_MODULE, "kernel %s not in %s", kernel_name, code_object);
    int rc = module_block(h->proc_mod, &h->proc_block);
    if (rc) return rc;
    kernel_info(h->proc_fn, h->proc_block, m > 0 ? 0 : h->args.lds_total * sizeof(double), &h->proc_lane_scratch,
                &h->proc_per_cu);
    h->proc_key = key;
  }
  int64_t nblocks, nthreads, npt;
  partition((uint64_t)n, ntotalthreads, tpb, nblocks, nthreads, npt);
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->nblocks = nblocks;
    stats->nthreads = nthreads;
    stats->npt = npt;
  }
  if (n == 0) return MCV_OK;
  if (n > h->io_cap) {
    if (h->d_io) HIPCHECK(hipFree(h->d_io));
    h->d_io = nullptr;
    HIPCHECK(hipMalloc(&h->d_io, (size_t)n * 10 * sizeof(double)));
    h->io_cap = n;
  }
  if (m > 0 && n * m > h->io_out_cap) {
    if (h->d_io_out) HIPCHECK(hipFree(h->d_io_out));
    h->d_io_out = nullptr;
    HIPCHECK(hipMalloc(&h->d_io_out, (size_t)n * m * 10 * sizeof(double)));
    h->io_out_cap = n * m;
  }
  if (nthreads > h->rng_cap) {
    if (h->d_rng) HIPCHECK(hipFree(h->d_rng));
    h->d_rng = nullptr;
    HIPCHECK(hipMalloc(&h->d_rng, (size_t)nthreads * 2 * sizeof(uint64_t)));
    h->rng_cap = nthreads;
  }
  HIPCHECK(hipMemcpyAsync(h->d_io, neutrons, (size_t)n * 10 * sizeof(double), hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(mcv_zero_u64, dim3(1), dim3(64), 0, h->stream, h->d_counter, (int64_t)1);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipEventRecord(h->ev[0], h->stream));
  int rc = seed_streams(h->ctx, h->stream, seed, 0, nthreads, h->d_rng);
  if (rc) return rc;
  HIPCHECK(hipEventRecord(h->ev[1], h->stream));
  mcv::LaunchArgs A = h->args;
  A.P = h->d_blob;
  A.H = h->d_hist;
  A.R = h->d_rep;
  A.rng = h->d_rng;
  A.rng_first = 0;
  A.tid_end = nthreads;
  A.counter = h->d_counter;
  A.dump = nullptr;
  A.dump_alive = nullptr;
  A.dump_count = 0;
  A.N = n;
  A.nthreads = nthreads;
  A.npt = npt;
  A.tid_base = 0;
  A.gate = (flags & MCV_FLAG_GATE_DEAD) ? 1 : 0;
  A.io = h->d_io;
  A.io_out = m > 0 ? h->d_io_out : nullptr;
  A.io_m = m;
  no_cache(A);
  if (m > 0) {  // the multiple-scattering process kernel keeps no histograms
    A.lds_nseg = 0;
    A.lds_total = 0;
  }
  const unsigned bs = (unsigned)h->proc_block;
  if (launch_grid(h->ctx, h->proc_lane_scratch, h->proc_block, h->proc_per_cu, (nthreads + bs - 1) / bs, false) <= 0)
    return scratch_error(h->proc_lane_scratch, h->proc_block, h->ctx);
  void* kargs[] = {&A};
  HIPCHECK(hipModuleLaunchKernel(h->proc_fn, (unsigned)((nthreads + bs - 1) / bs), 1, 1, bs, 1, 1,
                                 (unsigned)(A.lds_total * sizeof(double)), h->stream, kargs, nullptr));
  if (m == 0) {
    rc = fold_replicas(h);
    if (rc) return rc;
  }
  HIPCHECK(hipEventRecord(h->ev[2], h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  unsigned long long processed = 0;
  HIPCHECK(hipMemcpy(&processed, h->d_counter, sizeof(processed), hipMemcpyDeviceToHost));
  if (m > 0)
    HIPCHECK(hipMemcpy(out_ms, h->d_io_out, (size_t)n * m * 10 * sizeof(double), hipMemcpyDeviceToHost));
  else
    HIPCHECK(hipMemcpy(neutrons, h->d_io, (size_t)n * 10 * sizeof(double), hipMemcpyDeviceToHost));
  if (stats) {
    float ms_rng = 0, ms_k = 0;
    HIPCHECK(hipEventElapsedTime(&ms_rng, h->ev[0], h->ev[1]));
    HIPCHECK(hipEventElapsedTime(&ms_k, h->ev[1], h->ev[2]));
    stats->processed = processed;
    stats->ms_rng_init = ms_rng;
    stats->ms_kernel = ms_k;
    stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  if ((int64_t)processed != n)
    return fail(MCV_ERR_COUNT, "Processed neutron count %llu does not match requested neutron count %lld",
                (unsigned long long)processed, (long long)n);
  return MCV_OK;
}

int mcv_instr_read_dump(mcv_instr* h, double* states, int32_t* alive) {
  if (!h || !h->d_dump) return fail(MCV_ERR_ARG, "no dump configured");
  HIPCHECK(hipMemcpy(states, h->d_dump, (size_t)h->dump_count * 10 * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(alive, h->d_alive, (size_t)h->dump_count * sizeof(int32_t), hipMemcpyDeviceToHost));
  return MCV_OK;
}

int mcv_rng_states(uint64_t seed, int64_t first, int64_t n, uint64_t* out) {
  if (first < 0 || n < 0 || (n > 0 && !out)) return fail(MCV_ERR_ARG, "bad stream range");
  if (n == 0) return MCV_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(MCV_ERR_HIP, "no HIP device available");
  DeviceCtx* ctx = nullptr;
  int rc = device_ctx(&ctx);
  if (rc) return rc;
  uint64_t* d = nullptr;
  HIPCHECK(hipMalloc(&d, (size_t)n * 2 * sizeof(uint64_t)));
  rc = seed_streams(ctx, ctx->stream, seed, first, n, d);
  if (rc == MCV_OK) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(out, d, (size_t)n * 2 * sizeof(uint64_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = fail(MCV_ERR_HIP, "stream seeding: %s", hipGetErrorString(e));
  }
  (void)hipFree(d);
  return rc;
}

void mcv_instr_destroy(mcv_instr* h) { release(h); }

int64_t mcv_scratch_fit(int64_t scratch_per_lane, int32_t block, int64_t resident_groups, int64_t budget_bytes) {
  if (block < 1 || budget_bytes < 0) return 0;
  return scratch_fit(scratch_per_lane, block, resident_groups, budget_bytes);
}

}  // extern "C"

// ------------------------------------------------------------------ batched scans
struct mcv_batch {
  int ninst = 0, ncomp = 0, block = 256, spl = 1;
  int pool = 0;          // the single-instance kernel takes streams from a pool
  int item_streams = 0;  // streams per work item of the batch kernel
  mcv::LaunchArgs args{};
  int64_t pstride = 0, hstride = 0;
  std::vector<int64_t> nbins;
  double* d_blob = nullptr;
  double* d_hist = nullptr;
  uint64_t* d_rng = nullptr;
  int64_t rng_cap = 0;
  int64_t* d_meta = nullptr;                 // item0 (ninst+1), n, nthreads, npt, instance (ninst each)
  unsigned long long* d_counters = nullptr;  // processed (ninst), work item counter, stream pool
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr, fn_batch = nullptr;
  int64_t lane_scratch = 0, lane_scratch_batch = 0;  // private segment bytes per lane
  int per_cu = 0, per_cu_batch = 0;        // resident workgroups per CU
  hipStream_t stream = nullptr;
  DeviceCtx* ctx = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<int64_t> graph_key;
  double* h_stage = nullptr;  // pinned host copy of d_hist (mcv_batch_histograms_deliver)
  std::vector<hipEvent_t> stage_ev;  // one per staged piece of d_hist
};

namespace {

void release_batch(mcv_batch* b) {
  if (!b) return;
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->d_blob) (void)hipFree(b->d_blob);
  if (b->d_hist) (void)hipFree(b->d_hist);
  if (b->d_rng) (void)hipFree(b->d_rng);
  if (b->d_meta) (void)hipFree(b->d_meta);
  if (b->d_counters) (void)hipFree(b->d_counters);
  for (auto& e : b->ev)
    if (e) (void)hipEventDestroy(e);
  if (b->mod) (void)hipModuleUnload(b->mod);
  if (b->h_stage) (void)hipHostFree(b->h_stage);
  for (auto& e : b->stage_ev) (void)hipEventDestroy(e);
  delete b;
}

int batch_create_impl(const mcv_comp_desc* comps, int32_t ninst, int32_t ncomp, const double* offsets,
                      const double* rotmats, const char* code_object, const char* kernel_name, mcv_batch** out) {
  if (!comps || ninst < 1 || ncomp < 1) return fail(MCV_ERR_ARG, "empty batch");
  if (ncomp > MCV_MAX_COMP) return fail(MCV_ERR_ARG, "chain longer than %d components", MCV_MAX_COMP);
  if (ncomp > 1 && (!offsets || !rotmats)) return fail(MCV_ERR_ARG, "missing frame transformations");
  if (!code_object || !kernel_name) return fail(MCV_ERR_ARG, "missing code object / kernel name");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(MCV_ERR_HIP, "no HIP device available");
  mcv_batch* b = new mcv_batch();
  *out = b;
  b->ninst = ninst;
  b->ncomp = ncomp;
  std::vector<double> all;
  int64_t hist_total = 0;
  for (int k = 0; k < ninst; ++k) {
    mcv::LaunchArgs a{};
    std::vector<double> blob;
    std::vector<int64_t> nb;
    std::vector<int> types, np;
    int64_t ht = 0;
    int rc = build_layout(comps + (size_t)k * ncomp, ncomp, offsets + (size_t)k * (ncomp - 1) * 3,
                          rotmats + (size_t)k * (ncomp - 1) * 9, a, blob, nb, ht, types, np);
    if (rc) return rc;
    if (k == 0) {
      b->args = a;
      b->nbins = nb;
      b->pstride = (int64_t)blob.size();
      hist_total = ht;
    } else if ((int64_t)blob.size() != b->pstride || ht != hist_total ||
               std::memcmp(a.off, b->args.off, sizeof(a.off)) || std::memcmp(a.doff, b->args.doff, sizeof(a.doff)) ||
               std::memcmp(a.hoff, b->args.hoff, sizeof(a.hoff))) {
      return fail(MCV_ERR_ARG, "instance %d differs in layout from instance 0 (tables and histograms must match)", k);
    }
    all.insert(all.end(), blob.begin(), blob.end());
  }
  b->hstride = hist_total;
  {
    DeviceCtx* ctx = nullptr;
    int rc = device_ctx(&ctx);
    if (rc) return rc;
    b->stream = ctx->stream;  // shared, not owned
    b->ctx = ctx;
  }
  for (auto& e : b->ev) HIPCHECK(hipEventCreate(&e));
  HIPCHECK(hipMalloc(&b->d_blob, all.size() * sizeof(double)));
  HIPCHECK(hipMemcpy(b->d_blob, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice));
  const int64_t hn = std::max<int64_t>(hist_total * ninst, 1);
  HIPCHECK(hipMalloc(&b->d_hist, hn * sizeof(double)));
  HIPCHECK(hipMemset(b->d_hist, 0, hn * sizeof(double)));
  HIPCHECK(hipMalloc(&b->d_meta, (size_t)(5 * ninst + 1) * sizeof(int64_t)));
  HIPCHECK(hipMalloc(&b->d_counters, (size_t)(ninst + 2) * sizeof(unsigned long long)));
  hipError_t e = hipModuleLoad(&b->mod, code_object);
  if (e != hipSuccess) return fail(MCV_ERR_MODULE, "hipModuleLoad(%s): %s", code_object, hipGetErrorString(e));
  if (hipModuleGetFunction(&b->fn, b->mod, kernel_name) != hipSuccess)
    return fail(MCV_ERR_MODULE, "kernel %s not in %s", kernel_name, code_object);
  std::string bn = std::string(kernel_name) + "_batch";
  if (hipModuleGetFunction(&b->fn_batch, b->mod, bn.c_str()) != hipSuccess)
    return fail(MCV_ERR_MODULE, "kernel %s not in %s", bn.c_str(), code_object);
  int rc = module_block(b->mod, &b->block, &b->spl);
  if (rc) return rc;
  rc = module_int(b->mod, "mcv_item_streams", b->block * b->spl, &b->item_streams);
  if (rc) return rc;
  if (b->item_streams < b->block * b->spl || b->item_streams % (b->block * b->spl))
    return fail(MCV_ERR_MODULE, "code object built for %d streams per work item", b->item_streams);
  rc = module_int(b->mod, "mcv_stream_pool", 0, &b->pool);
  if (rc) return rc;
  const size_t lds = b->args.lds_total * sizeof(double);
  kernel_info(b->fn, b->block, lds, &b->lane_scratch, &b->per_cu);
  kernel_info(b->fn_batch, b->block, lds, &b->lane_scratch_batch, &b->per_cu_batch);
  return MCV_OK;
}

int ensure_rng(mcv_batch* b, int64_t n) {
  if (n <= b->rng_cap) return MCV_OK;
  if (b->d_rng) HIPCHECK(hipFree(b->d_rng));
  b->d_rng = nullptr;
  HIPCHECK(hipMalloc(&b->d_rng, (size_t)n * 2 * sizeof(uint64_t)));
  b->rng_cap = n;
  if (b->gexec) {  // the graph captured the old buffer
    (void)hipGraphExecDestroy(b->gexec);
    (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
  }
  return MCV_OK;
}

// stream seeding + counter reset + persistent batch kernel, on b->stream
int enqueue_batch(mcv_batch* b, uint64_t seed, int64_t nthreads_max, int64_t total_items, uint32_t flags,
                  int32_t grid) {
  int rc = seed_streams(b->ctx, b->stream, seed, 0, nthreads_max, b->d_rng);
  if (rc) return rc;
  hipLaunchKernelGGL(mcv_zero_u64, dim3((unsigned)((b->ninst + 1 + 255) / 256)), dim3(256), 0, b->stream,
                     b->d_counters, (int64_t)(b->ninst + 1));
  HIPCHECK(hipGetLastError());
  mcv::BatchArgs B{};
  B.base = b->args;
  B.base.P = b->d_blob;
  B.base.H = b->d_hist;
  B.base.rng = b->d_rng;
  B.base.dump = nullptr;
  B.base.dump_alive = nullptr;
  B.base.dump_count = 0;
  B.base.tid_base = 0;
  B.base.tid_end = INT64_MAX;  // each instance's own stream count bounds its streams
  B.base.rng_first = 0;
  no_cache(B.base);
  B.base.gate = (flags & MCV_FLAG_GATE_DEAD) ? 1 : 0;
  B.pstride = b->pstride;
  B.hstride = b->hstride;
  B.item0 = b->d_meta;
  B.inst_n = b->d_meta + b->ninst + 1;
  B.inst_nthreads = b->d_meta + 2 * b->ninst + 1;
  B.inst_npt = b->d_meta + 3 * b->ninst + 1;
  B.inst_index = b->d_meta + 4 * b->ninst + 1;
  B.inst_processed = b->d_counters;
  B.work = b->d_counters + b->ninst;
  B.ninst = b->ninst;
  int64_t g = std::min<int64_t>(grid > 0 ? grid : 1024, std::max<int64_t>(total_items, 1));
  g = launch_grid(b->ctx, b->lane_scratch_batch, b->block, b->per_cu_batch, g, true);
  if (g <= 0) return scratch_error(b->lane_scratch_batch, b->block, b->ctx);
  void* kargs[] = {&B};
  HIPCHECK(hipModuleLaunchKernel(b->fn_batch, (unsigned)g, 1, 1, (unsigned)b->block, 1, 1,
                                 (unsigned)(B.base.lds_total * sizeof(double)), b->stream, kargs, nullptr));
  return MCV_OK;
}

}  // namespace

extern "C" {

mcv_batch* mcv_batch_create(const mcv_comp_desc* comps, int32_t ninst, int32_t ncomp, const double* offsets,
                            const double* rotmats, const char* code_object, const char* kernel_name) {
  mcv_batch* b = nullptr;
  int rc = batch_create_impl(comps, ninst, ncomp, offsets, rotmats, code_object, kernel_name, &b);
  if (rc) {
    std::string msg = g_last_error;
    release_batch(b);
    g_last_error = msg;
    return nullptr;
  }
  return b;
}

int mcv_batch_run(mcv_batch* b, const uint64_t* ncounts, uint64_t seed, int64_t ntotalthreads, int32_t tpb,
                  uint32_t flags, int32_t grid, int32_t use_graph, mcv_run_stats* stats) {
  if (!b || !ncounts) return fail(MCV_ERR_ARG, "bad batch arguments");
  const auto t0 = std::chrono::steady_clock::now();
  const int n = b->ninst;
  // work order: instances with more neutrons per stream (longer work items)
  // first, so that the launch ends on short items and no workgroup is left
  // with a long one while the others run dry
  std::vector<int64_t> nt_of(n), npt_of(n);
  std::vector<int> order(n);
  for (int k = 0; k < n; ++k) {
    int64_t nb;
    partition(ncounts[k], ntotalthreads, tpb, nb, nt_of[k], npt_of[k]);
    order[k] = k;
  }
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return npt_of[x] > npt_of[y]; });
  std::vector<int64_t> meta(5 * n + 1, 0);
  int64_t nthreads_max = 0;
  uint64_t total_n = 0;
  for (int j = 0; j < n; ++j) {
    const int k = order[j];
    const int64_t nt = nt_of[k], npt = npt_of[k];
    const int64_t per_item = b->item_streams;
    meta[j + 1] = meta[j] + (nt + per_item - 1) / per_item;
    meta[n + 1 + j] = (int64_t)ncounts[k];
    meta[2 * n + 1 + j] = nt;
    meta[3 * n + 1 + j] = npt;
    meta[4 * n + 1 + j] = k;
    nthreads_max = std::max(nthreads_max, nt);
    total_n += ncounts[k];
  }
  const int64_t total_items = meta[n];
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->nthreads = nthreads_max;
  }
  int rc = ensure_rng(b, std::max<int64_t>(nthreads_max, 1));
  if (rc) return rc;
  HIPCHECK(hipMemcpyAsync(b->d_meta, meta.data(), meta.size() * sizeof(int64_t), hipMemcpyHostToDevice, b->stream));
  HIPCHECK(hipEventRecord(b->ev[0], b->stream));
  if (total_items > 0) {
    if (use_graph) {
      std::vector<int64_t> key = {(int64_t)seed, nthreads_max, total_items, (int64_t)flags, grid};
      if (!b->gexec || key != b->graph_key) {
        if (b->gexec) {
          (void)hipGraphExecDestroy(b->gexec);
          (void)hipGraphDestroy(b->graph);
          b->gexec = nullptr;
          b->graph = nullptr;
        }
        HIPCHECK(hipStreamSynchronize(b->stream));
        DeviceCtx* ctx = nullptr;  // allocate the jump tables before capture, not inside it
        rc = device_ctx(&ctx);
        if (rc) return rc;
        HIPCHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_batch(b, seed, nthreads_max, total_items, flags, grid);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc) {
          if (g) (void)hipGraphDestroy(g);
          return rc;
        }
        if (e != hipSuccess) return fail(MCV_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
        b->graph = g;
        HIPCHECK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
        b->graph_key = key;
        HIPCHECK(hipEventRecord(b->ev[0], b->stream));
      }
      HIPCHECK(hipGraphLaunch(b->gexec, b->stream));
    } else {
      rc = enqueue_batch(b, seed, nthreads_max, total_items, flags, grid);
      if (rc) return rc;
    }
  }
  HIPCHECK(hipEventRecord(b->ev[1], b->stream));
  HIPCHECK(hipStreamSynchronize(b->stream));
  std::vector<unsigned long long> processed(n + 1, 0);
  HIPCHECK(hipMemcpy(processed.data(), b->d_counters, (n + 1) * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  float ms = 0;
  HIPCHECK(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
  uint64_t done = 0;
  for (int k = 0; k < n; ++k) done += total_items > 0 ? processed[k] : 0;
  if (stats) {
    stats->processed = done;
    stats->ms_kernel = ms;
    stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  for (int k = 0; k < n; ++k)
    if (total_items > 0 && processed[k] != ncounts[k])
      return fail(MCV_ERR_COUNT, "instance %d: processed %llu of %llu neutrons", k,
                  (unsigned long long)processed[k], (unsigned long long)ncounts[k]);
  return MCV_OK;
}

int mcv_batch_run_sequential(mcv_batch* b, const uint64_t* ncounts, uint64_t seed, int64_t ntotalthreads,
                             int32_t tpb, uint32_t flags, mcv_run_stats* stats) {
  if (!b || !ncounts) return fail(MCV_ERR_ARG, "bad batch arguments");
  const auto t0 = std::chrono::steady_clock::now();
  double ms_sum = 0;
  uint64_t done = 0;
  for (int k = 0; k < b->ninst; ++k) {
    if (ncounts[k] == 0) continue;
    int64_t nb, nt, npt;
    partition(ncounts[k], ntotalthreads, tpb, nb, nt, npt);
    int rc = ensure_rng(b, nt);
    if (rc) return rc;
    HIPCHECK(hipEventRecord(b->ev[0], b->stream));
    rc = seed_streams(b->ctx, b->stream, seed, 0, nt, b->d_rng);
    if (rc) return rc;
    hipLaunchKernelGGL(mcv_zero_u64, dim3(1), dim3(64), 0, b->stream, b->d_counters, (int64_t)1);
    hipLaunchKernelGGL(mcv_zero_u64, dim3(1), dim3(64), 0, b->stream, b->d_counters + b->ninst + 1, (int64_t)1);
    mcv::LaunchArgs A = b->args;
    A.P = b->d_blob + k * b->pstride;
    A.H = b->d_hist + k * b->hstride;
    A.rng = b->d_rng;
    A.counter = b->d_counters;
    A.pool = b->pool ? b->d_counters + b->ninst + 1 : nullptr;
    A.dump = nullptr;
    A.dump_alive = nullptr;
    A.dump_count = 0;
    A.N = (int64_t)ncounts[k];
    A.nthreads = nt;
    A.npt = npt;
    A.tid_base = 0;
    A.tid_end = nt;
    A.rng_first = 0;
    no_cache(A);
    A.gate = (flags & MCV_FLAG_GATE_DEAD) ? 1 : 0;
    void* kargs[] = {&A};
    const int64_t per_group = (int64_t)b->block * b->spl;
    int64_t grid = (nt + per_group - 1) / per_group;
    grid = launch_grid(b->ctx, b->lane_scratch, b->block, b->per_cu, grid, b->pool != 0);
    if (grid <= 0) return scratch_error(b->lane_scratch, b->block, b->ctx);
    HIPCHECK(hipModuleLaunchKernel(b->fn, (unsigned)grid, 1, 1, (unsigned)b->block, 1,
                                   1, (unsigned)(A.lds_total * sizeof(double)), b->stream, kargs,
                                   nullptr));
    HIPCHECK(hipEventRecord(b->ev[1], b->stream));
    HIPCHECK(hipStreamSynchronize(b->stream));
    unsigned long long processed = 0;
    HIPCHECK(hipMemcpy(&processed, b->d_counters, sizeof(processed), hipMemcpyDeviceToHost));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    ms_sum += ms;
    done += processed;
    if (processed != ncounts[k])
      return fail(MCV_ERR_COUNT, "instance %d: processed %llu of %llu neutrons", k, processed,
                  (unsigned long long)ncounts[k]);
  }
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->processed = done;
    stats->ms_kernel = ms_sum;
    stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return MCV_OK;
}

int mcv_batch_monitor_read(mcv_batch* b, int32_t inst, int32_t comp, double* out3) {
  if (!b || inst < 0 || inst >= b->ninst || comp < 0 || comp >= b->ncomp)
    return fail(MCV_ERR_ARG, "bad batch, instance or component index");
  if (b->nbins[comp] <= 0) return fail(MCV_ERR_ARG, "component %d is not a monitor", comp);
  HIPCHECK(hipMemcpy(out3, b->d_hist + inst * b->hstride + b->args.hoff[comp], 3 * b->nbins[comp] * sizeof(double),
                     hipMemcpyDeviceToHost));
  return MCV_OK;
}

int mcv_batch_histograms_read(mcv_batch* b, double* out, int64_t n) {
  if (!b || (!out && n)) return fail(MCV_ERR_ARG, "bad batch or output buffer");
  if (n != b->hstride * b->ninst)
    return fail(MCV_ERR_ARG, "histograms of the batch hold %lld doubles, not %lld",
                (long long)(b->hstride * b->ninst), (long long)n);
  if (n > 0) HIPCHECK(hipMemcpy(out, b->d_hist, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
  return MCV_OK;
}

int mcv_batch_histograms_deliver(mcv_batch* b, int64_t nblocks, const int64_t* src, const int64_t* len,
                                 double* const* dst, int32_t accumulate) {
  if (!b || nblocks < 0 || (nblocks > 0 && (!src || !len || !dst)))
    return fail(MCV_ERR_ARG, "bad batch or block lists");
  const int64_t n = b->hstride * b->ninst;
  int64_t bytes = 0;
  for (int64_t j = 0; j < nblocks; ++j) {
    if (src[j] < 0 || len[j] < 0 || src[j] > n - len[j] || (len[j] > 0 && !dst[j]))
      return fail(MCV_ERR_ARG, "block %lld (offset %lld, %lld doubles) is outside the %lld histogram doubles",
                  (long long)j, (long long)src[j], (long long)len[j], (long long)n);
    bytes += len[j] * (int64_t)sizeof(double);
  }
  if (n == 0 || bytes == 0) return MCV_OK;
  if (!b->h_stage) HIPCHECK(hipHostMalloc((void**)&b->h_stage, (size_t)n * sizeof(double), hipHostMallocDefault));
  // the copy goes in pieces, each followed by an event, so that blocks are
  // written out while the later pieces are still in flight
  constexpr int64_t kPiece = int64_t(1) << 20;  // doubles (8 MB)
  const int64_t npiece = (n + kPiece - 1) / kPiece;
  while ((int64_t)b->stage_ev.size() < npiece) {
    hipEvent_t e;
    HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b->stage_ev.push_back(e);
  }
  for (int64_t c = 0; c < npiece; ++c) {
    const int64_t c0 = c * kPiece, cn = std::min(kPiece, n - c0);
    HIPCHECK(hipMemcpyAsync(b->h_stage + c0, b->d_hist + c0, (size_t)cn * sizeof(double), hipMemcpyDeviceToHost,
                            b->stream));
    HIPCHECK(hipEventRecord(b->stage_ev[c], b->stream));
  }
  // blocks dealt to threads in turn; a few MB per thread at least
  const int64_t nt = std::max<int64_t>(1, std::min<int64_t>({8, nblocks, bytes >> 22}));
  std::atomic<int> err{(int)hipSuccess};
  auto work = [&](int64_t t) {
    for (int64_t j = t; j < nblocks; j += nt) {
      if (len[j] == 0) continue;
      // the pieces arrive in order: the one holding the block's last value is enough
      const hipError_t e = hipEventSynchronize(b->stage_ev[(src[j] + len[j] - 1) / kPiece]);
      if (e != hipSuccess) {
        int ok = (int)hipSuccess;
        err.compare_exchange_strong(ok, (int)e);
        return;
      }
      const double* s = b->h_stage + src[j];
      double* d = dst[j];
      if (accumulate) {
        for (int64_t i = 0; i < len[j]; ++i) d[i] += s[i];
      } else {
        std::memcpy(d, s, (size_t)len[j] * sizeof(double));
      }
    }
  };
  std::vector<std::thread> pool;
  std::vector<int64_t> mine{0};  // thread indices this thread works through
  for (int64_t t = 1; t < nt; ++t) {
    try {
      pool.emplace_back(work, t);
    } catch (...) {  // no thread to be had: this one does that share too
      mine.push_back(t);
    }
  }
  for (const int64_t t : mine) work(t);
  for (auto& th : pool) th.join();
  HIPCHECK((hipError_t)err.load());
  HIPCHECK(hipStreamSynchronize(b->stream));  // every piece is in before the staging buffer is reused
  return MCV_OK;
}

int mcv_batch_clear(mcv_batch* b) {
  if (!b) return fail(MCV_ERR_ARG, "null batch");
  const int64_t n = b->hstride * b->ninst;
  if (n > 0) {
    hipLaunchKernelGGL(mcv_zero_f64, dim3(1024), dim3(256), 0, b->stream, b->d_hist, n);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(b->stream));
  }
  return MCV_OK;
}

void mcv_batch_destroy(mcv_batch* b) { release_batch(b); }

}  // extern "C"
