// oracle.cpp -- CPU restatement of mcvine.acc's per-neutron propagate path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP engine is
// compared against; only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline may load it.  It is never part of the product path.
//
// Every function restates the reference's algorithm statement by statement
// (same operation order, same float32/float64 mix, same quirks), citing the
// reference file:line it follows.  Paths are relative to the mcvine/acc
// checkout.  The code is deliberately plain: AoS neutron[10] arrays
// (x,y,z,vx,vy,vz,s1,s2,t,p as in acc/neutron.py), recursion for CSG, local
// arrays for the multiple-scattering buffers, and the numba xoroshiro128+
// generator with its sequential jump-ahead seeding.  Compile with
// -ffp-contract=off so no FMA contraction changes the reference's rounding.
//
// Parity status of each part (see tests/test_oracle_*.py):
//  * RNG: xoroshiro128+ transition pinned by full-period + jump-polynomial
//    checks; SplitMix64 by published outputs; numba's "same z in both words"
//    seeding and float32 conversion follow numba/cuda/random.py (numba 0.53.1,
//    not vendored) -- parity unpinned at that boundary.
//  * Conversion constants V2K/SE2V/VS2E are McStas' published values
//    (mcni.utils.conversion is not vendored); K2V = 1/V2K -- parity unpinned.
//  * Geometry, HSS time bookkeeping, HMS interactions, frame transforms: pinned
//    by the reference tests' literal vectors (tests/golden/*.json).
//  * Source_simple (draw order, float32 radius, mirrored target point), the
//    S(Q,E)-grid lookup, the phonon-DOS weight, the tube detector, wavelength
//    binning, the I(Q,E) normalizer's neutrons and the segmented guide's
//    sloped-wall reflection: pinned by closed forms computed from the
//    neutron's own state (tests/test_oracle_pins.py); the grid and DOS kernels
//    and the detector have no reference counterpart, so those pins state what
//    this engine defines them to be.

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#include <omp.h>

#include "../include/mcv_acc.h"

// Flop-counting hooks (oracle_flops.cpp defines them; no-ops here).
#ifndef ORC_PAUSE
#define ORC_PAUSE ((void)0)
#define ORC_RESUME ((void)0)
#define ORC_DRAW ((void)0)
#endif

namespace orc {

// Conditioning probe (test hook): when g_nudge != 0 every transcendental
// result is moved by one ulp: up (1), down (-1), or in a direction drawn per
// call from a hash of the value (2).  A neutron whose final state
// moves by more than the parity tolerance under this nudge is ill-conditioned:
// a 1-ulp difference between two correct libm implementations (the host's and
// the GPU's) legitimately shows up in its final state.
static int g_nudge = 0;
// Reproduction hook (test hook): for the one neutron g_pin_idx, the k-th
// transcendental result of its history (counted from 0 at its start) is moved
// by g_pin_dir[k] ulps (-1, 0, +1); every other neutron is left alone.  A
// device state that this reproduces bit for bit differs from the oracle's by
// exactly those libm ulps.
static int64_t g_pin_idx = -1;
static std::vector<int> g_pin_dir;
#ifdef ORC_FLOP_BUILD
template <class T>
static inline T tr(T v) { return v; }
#else
static thread_local bool t_pin = false;
static thread_local size_t t_call = 0;
static inline double tr(double v) {
  if (t_pin) {
    const size_t k = t_call++;
    if (k < g_pin_dir.size() && g_pin_dir[k]) v = std::nextafter(v, g_pin_dir[k] > 0 ? INFINITY : -INFINITY);
    return v;
  }
  if (!g_nudge) return v;
  int dir = g_nudge;
  if (g_nudge == 2) {  // sign per call, from a hash of the value's bits
    uint64_t b;
    std::memcpy(&b, &v, sizeof b);
    b ^= b >> 33;
    b *= 0xff51afd7ed558ccdull;
    b ^= b >> 33;
    dir = (b & 1) ? 1 : -1;
  }
  return std::nextafter(v, dir > 0 ? INFINITY : -INFINITY);
}
#endif

#define TR_SIN(x) tr(std::sin(x))
#define TR_COS(x) tr(std::cos(x))
#define TR_EXP(x) tr(std::exp(x))
#define TR_TANH(x) tr(std::tanh(x))
#define TR_ATAN(x) tr(std::atan(x))
#define TR_ATAN2(a, b) tr(std::atan2(a, b))
#define TR_ASIN(x) tr(std::asin(x))

// mcni.utils.conversion (McStas constants, not vendored)
static const double V2K = 1.58825361e-3;
static const double SE2V = 437.393377;
static const double VS2E = 5.22703725e-6;
static const double K2V = 1.0 / 1.58825361e-3;
static const double PI = 3.141592653589793;  // math.pi

// ---------------------------------------------------------------- RNG (R1)
// numba/cuda/random.py (numba 0.53.1): init_xoroshiro128p_state,
// xoroshiro128p_next, xoroshiro128p_jump, uint64_to_unit_float32.
struct Rng {
  uint64_t s0, s1;
};

static inline uint64_t rotl(uint64_t x, unsigned k) { return (x << k) | (x >> (64u - k)); }

static inline uint64_t splitmix64(uint64_t seed) {
  uint64_t z = seed + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static inline uint64_t next(Rng& s) {
  uint64_t s0 = s.s0, s1 = s.s1;
  uint64_t result = s0 + s1;
  s1 ^= s0;
  s.s0 = rotl(s0, 55) ^ s1 ^ (s1 << 14);
  s.s1 = rotl(s1, 36);
  return result;
}

static inline void jump(Rng& s) {
  static const uint64_t J[2] = {0xbeac0467eba5facbull, 0xd86b048b86aa9922ull};
  uint64_t a = 0, b = 0;
  for (int i = 0; i < 2; ++i)
    for (int bit = 0; bit < 64; ++bit) {
      if (J[i] & (1ull << bit)) {
        a ^= s.s0;
        b ^= s.s1;
      }
      next(s);
    }
  s.s0 = a;
  s.s1 = b;
}

static inline float uniform_f32(Rng& s) {
  uint64_t x = next(s);
  ORC_DRAW;
  ORC_PAUSE;  // the draw's integer-to-float conversion is RNG work, counted apart
  double d = (double)(x >> 11) * (1.0 / 9007199254740992.0);
  ORC_RESUME;
  return (float)d;
}

// create_xoroshiro128p_states(n, seed): state0 from SplitMix64, then jump chain.
static void init_states(uint64_t seed, int64_t n, Rng* out) {
  if (n < 1) return;
  uint64_t z = splitmix64(seed);
  out[0].s0 = z;
  out[0].s1 = z;
  for (int64_t i = 1; i < n; ++i) {
    out[i] = out[i - 1];
    jump(out[i]);
  }
}

// ---------------------------------------------------------------- vec3 (acc/vec3.py)
static inline double dot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void cross(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double length(const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
static inline void scale(double* v, double s) {
  v[0] *= s;
  v[1] *= s;
  v[2] *= s;
}
static inline void normalize(double* v) {
  double l = length(v);
  scale(v, 1.0 / l);
}
static inline void add(const double* a, const double* b, double* o) {
  o[0] = a[0] + b[0];
  o[1] = a[1] + b[1];
  o[2] = a[2] + b[2];
}
static inline void copy3(const double* a, double* b) {
  b[0] = a[0];
  b[1] = a[1];
  b[2] = a[2];
}
static inline void mXv(const double* m, const double* v, double* o) {
  for (int i = 0; i < 3; ++i) o[i] = m[3 * i + 0] * v[0] + m[3 * i + 1] * v[1] + m[3 * i + 2] * v[2];
}

// acc/neutron.py:18-22 abs2rel; acc/vec3.py:71-75 _abs2rel: r' = R (r - o); v' = R v
static inline void apply_transformation(double* n, const double* rot, const double* off) {
  double rt[3], vt[3], tmp[3];
  copy3(n, rt);
  copy3(n + 3, vt);
  tmp[0] = rt[0] - off[0];
  tmp[1] = rt[1] - off[1];
  tmp[2] = rt[2] - off[2];
  mXv(rot, tmp, n);
  mXv(rot, vt, n + 3);
}

// acc/neutron.py:44-51 prop_dt_inplace
static inline void prop_dt_inplace(double* n, double dt) {
  double x = n[0], y = n[1], z = n[2], vx = n[3], vy = n[4], vz = n[5], t = n[8];
  n[0] = x + vx * dt;
  n[1] = y + vy * dt;
  n[2] = z + vz * dt;
  n[8] = t + dt;
}

// acc/neutron.py:60-66 prop_z0
static inline void prop_z0(const double* n, double& x, double& y, double& t) {
  double dt = -n[2] / n[5];
  x = n[0] + n[3] * dt;
  y = n[1] + n[4] * dt;
  t = n[8] + dt;
}

static inline bool is_moving(const double* n) {  // acc/neutron.py:38-42
  double v = std::sqrt(n[3] * n[3] + n[4] * n[4] + n[5] * n[5]);
  return v > 1.e-6;
}

// ---------------------------------------------------------------- Source_simple (R3)
// acc/components/sources/source_simple.py:17-42
static double randvec_target_rect(const double* target, double width, double height, float rand1, float rand2,
                                  double* vecout) {
  double dx = width * ((double)rand1 * 2 - 1) / 2.0;
  double dy = height * ((double)rand2 * 2 - 1) / 2.0;
  ORC_PAUSE;  // loop-invariant: depends on the target only
  double dist = length(target);
  double p1[3], vertical[3] = {0.0, 1.0, 0.0};
  cross(target, vertical, p1);
  normalize(p1);
  double p2[3];
  cross(target, p1, p2);
  normalize(p2);
  ORC_RESUME;
  scale(p1, dx);
  scale(p2, dy);
  double tmp[3];
  add(p1, target, tmp);
  add(p2, tmp, vecout);
  double dist2 = std::sqrt(dx * dx + dy * dy + dist * dist);
  return (width * height * dist) / (dist2 * dist2 * dist2);
}

// acc/components/sources/source_simple.py:114-161
static void source_simple(Rng& rng, double* n, const double* P) {
  bool square = P[0] != 0;
  double width = P[1], height = P[2], radius = P[3];
  bool wl_distr = P[4] != 0;
  double Lambda0 = P[5], dLambda = P[6], E0 = P[7], dE = P[8], xw = P[9], yh = P[10], dist = P[11], pmul = P[12];
  float r1 = uniform_f32(rng), r2 = uniform_f32(rng), r3 = uniform_f32(rng), r4 = uniform_f32(rng),
        r5 = uniform_f32(rng);
  double x, y;
  if (square) {
    x = width * ((double)r1 - 0.5);
    y = height * ((double)r2 - 0.5);
  } else {
    double chi = 2 * PI * (double)r1;
    double r = (double)std::sqrt(r2) * radius;  // float32 sqrt (numba types sqrt(float32) as float32)
    x = r * TR_COS(chi);
    y = r * TR_SIN(chi);
  }
  n[0] = x;
  n[1] = y;
  n[2] = 0.;
  double target[3] = {0.0, 0.0, dist};
  double vec_f[3];
  double solidangle = randvec_target_rect(target, xw, yh, r3, r4, vec_f);
  double dx = vec_f[0] - x, dy = vec_f[1] - y;
  double dist1 = std::sqrt(dx * dx + dy * dy + dist * dist);
  double v;
  if (wl_distr) {
    double L = Lambda0 + dLambda * ((double)r5 * 2 - 1);
    v = K2V * (2 * PI / L);
  } else {
    double E = E0 + dE * ((double)r5 * 2 - 1);
    v = SE2V * std::sqrt(E);
  }
  n[3] = v * dx / dist1;
  n[4] = v * dy / dist1;
  n[5] = v * dist / dist1;
  n[8] = 0;
  n[9] = pmul * solidangle;
}

// I(Q,E)-monitor normalizer neutrons (acc/components/monitors/iqe_monitor.py:146-174,
// makeNeutrons), one neutron per call: E, cos(theta), phi drawn in that order.
static void iqe_normalizer_source(Rng& rng, double* n, const double* P) {
  double Ei = P[0], Emin = P[1], Emax = P[2];
  double E = Emin + (double)uniform_f32(rng) * (Emax - Emin);
  double Ef = Ei - E;
  double vi = std::sqrt(Ei) * SE2V, vf = std::sqrt(Ef) * SE2V;
  double cos_t = (double)uniform_f32(rng) * 2 - 1;
  double theta = std::acos(cos_t);
  double sin_t = TR_SIN(theta);
  double phi = (double)uniform_f32(rng) * 2 * PI;
  n[0] = 0; n[1] = 0; n[2] = 0;
  n[3] = vf * sin_t * TR_COS(phi);
  n[4] = vf * sin_t * TR_SIN(phi);
  n[5] = vf * cos_t;
  n[8] = 0;
  n[9] = vf / vi;
}

// ---------------------------------------------------------------- reflectivity (R6)
// acc/components/optics/_guide_utils.py:6-23
static inline double calc_reflectivity(double Q, double R0, double Qc, double alpha, double m, double W) {
  double R = R0;
  if (Q > Qc) {
    double tmp = (Q - m * Qc) / W;
    if (tmp < 10)
      R *= (1 - TR_TANH(tmp)) * (1 - alpha * (Q - Qc)) / 2;
    else
      R = 0;
  }
  return R;
}

// ---------------------------------------------------------------- Guide (R5)
// acc/components/optics/guide.py:59-162
static void guide(double* n, const double* P) {
  double ww = P[0], hh = P[1], hw1 = P[2], hh1 = P[3], l = P[4], R0 = P[5], Qc = P[6], alpha = P[7], m = P[8],
         W = P[9];
  double x = n[0], y = n[1], z = n[2], vx = n[3], vy = n[4], vz = n[5];
  double t = n[8], prob = n[9];
  double dt = -z / vz;
  x += vx * dt;
  y += vy * dt;
  z = 0.;
  t += dt;
  if (x <= -hw1 || x >= hw1 || y <= -hh1 || y >= hh1) {
    n[9] = -1;  // absorb(): state is not written back
    return;
  }
  for (int nb = 0; nb < 100000; ++nb) {
    double av = l * vx, bv = ww * vz, ah = l * vy, bh = hh * vz;
    double vdotn_v1 = bv + av, vdotn_v2 = bv - av, vdotn_h1 = bh + ah, vdotn_h2 = bh - ah;
    double cv1 = -hw1 * l - z * ww, cv2 = x * l, ch1 = -hh1 * l - z * hh, ch2 = y * l;
    double t1 = (l - z) / vz, t2;
    int i = 0;
    if (vdotn_v1 < 0) {
      t2 = (cv1 - cv2) / vdotn_v1;
      if (t2 < t1) { t1 = t2; i = 1; }
    }
    if (vdotn_v2 < 0) {
      t2 = (cv1 + cv2) / vdotn_v2;
      if (t2 < t1) { t1 = t2; i = 2; }
    }
    if (vdotn_h1 < 0) {
      t2 = (ch1 - ch2) / vdotn_h1;
      if (t2 < t1) { t1 = t2; i = 3; }
    }
    if (vdotn_h2 < 0) {
      t2 = (ch1 + ch2) / vdotn_h2;
      if (t2 < t1) { t1 = t2; i = 4; }
    }
    if (i == 0) break;
    x += vx * t1;
    y += vy * t1;
    z += vz * t1;
    t += t1;
    double q = 0, d, nlen2;
    if (i == 1) {
      nlen2 = l * l + ww * ww;
      q = V2K * (-2) * vdotn_v1 / std::sqrt(nlen2);
      d = 2 * vdotn_v1 / nlen2;
      vx = vx - d * l;
      vz = vz - d * ww;
    } else if (i == 2) {
      nlen2 = l * l + ww * ww;
      q = V2K * (-2) * vdotn_v2 / std::sqrt(nlen2);
      d = 2 * vdotn_v2 / nlen2;
      vx = vx + d * l;
      vz = vz - d * ww;
    } else if (i == 3) {
      nlen2 = l * l + hh * hh;
      q = V2K * (-2) * vdotn_h1 / std::sqrt(nlen2);
      d = 2 * vdotn_h1 / nlen2;
      vy = vy - d * l;
      vz = vz - d * hh;
    } else {
      nlen2 = l * l + hh * hh;
      q = V2K * (-2) * vdotn_h2 / std::sqrt(nlen2);
      d = 2 * vdotn_h2 / nlen2;
      vy = vy + d * l;
      vz = vz - d * hh;
    }
    double R = calc_reflectivity(q, R0, Qc, alpha, m, W);
    prob *= R;
    if (prob <= 0) break;
  }
  n[0] = x; n[1] = y; n[2] = z; n[3] = vx; n[4] = vy; n[5] = vz;
  n[8] = t;
  n[9] = prob;
}

// ---------------------------------------------------------------- Guide_tapering (R7)
// acc/components/optics/guide_tapering.py:141-291; table rows per prep_inputs_kernel (:117-139)
static void guide_tapering(double* n, const double* P, int64_t nseg, const double* data) {
  const double *w1 = data, *ww = data + 4 * nseg, *hh = data + 5 * nseg, *whalf = data + 6 * nseg,
               *hhalf = data + 7 * nseg, *h1 = data + 2 * nseg;
  double l_seg = P[0], R0 = P[1], Qcx = P[2], Qcy = P[3], alphax = P[4], alphay = P[5], mx = P[6], my = P[7],
         W = P[8];
  double x = n[0], y = n[1], z = n[2], vx = n[3], vy = n[4], vz = n[5];
  double t = n[8], prob = n[9];
  double dt = -z / vz;
  x += vx * dt;
  y += vy * dt;
  z = 0.0;
  t += dt;
  for (int64_t seg = 0; seg < nseg; ++seg) {
    double zr = (double)seg * l_seg;
    dt = (zr - z) / vz;
    x += vx * dt;
    y += vy * dt;
    z += vz * dt;
    if (x <= -0.5 * w1[seg] || x >= 0.5 * w1[seg] || y <= -hhalf[seg] || y >= hhalf[seg]) {
      n[9] = 0;
      return;
    }
    x += 0.5 * w1[seg];
    double edge = std::floor(x / w1[seg]) * w1[seg];
    if (x - edge > w1[seg]) {
      n[9] = 0;
      return;
    }
    x -= (edge + 0.5 * w1[seg]);
    double hadj = edge + (0.5 * w1[seg]) - 0.5 * w1[seg];
    for (int nb = 0; nb < 100000; ++nb) {
      double av = l_seg * vx, bv = ww[seg] * vz, ah = l_seg * vy, bh = hh[seg] * vz;
      double vdotn_v1 = bv + av, vdotn_v2 = bv - av, vdotn_h1 = bh + ah, vdotn_h2 = bh - ah;
      double cv1 = -whalf[seg] * l_seg - (z - zr) * ww[seg], cv2 = x * l_seg;
      double ch1 = -hhalf[seg] * l_seg - (z - zr) * hh[seg], ch2 = y * l_seg;
      double t1 = (zr + l_seg - z) / vz, t2;
      int i = 0;
      if (vdotn_v1 < 0) {
        t2 = (cv1 - cv2) / vdotn_v1;
        if (t2 < t1) { t1 = t2; i = 1; }
      }
      if (vdotn_v2 < 0) {
        t2 = (cv1 + cv2) / vdotn_v2;
        if (t2 < t1) { t1 = t2; i = 2; }
      }
      if (vdotn_h1 < 0) {
        t2 = (ch1 - ch2) / vdotn_h1;
        if (t2 < t1) { t1 = t2; i = 3; }
      }
      if (vdotn_h2 < 0) {
        t2 = (ch1 + ch2) / vdotn_h2;
        if (t2 < t1) { t1 = t2; i = 4; }
      }
      if (i == 0) break;
      x += vx * t1;
      y += vy * t1;
      z += vz * t1;
      t += t1;
      double q, d, nlen2;
      if (i == 1) {
        nlen2 = l_seg * l_seg + ww[seg] * ww[seg];
        q = V2K * (-2.0) * vdotn_v1 / std::sqrt(nlen2);
        d = 2.0 * vdotn_v1 / nlen2;
        vx = vx - d * l_seg;
        vz = vz - d * ww[seg];
      } else if (i == 2) {
        nlen2 = l_seg * l_seg + ww[seg] * ww[seg];
        q = V2K * (-2.0) * vdotn_v2 / std::sqrt(nlen2);
        d = 2.0 * vdotn_v2 / nlen2;
        vx = vx + d * l_seg;
        vz = vz - d * ww[seg];
      } else if (i == 3) {
        nlen2 = l_seg * l_seg + hh[seg] * hh[seg];
        q = V2K * (-2.0) * vdotn_h1 / std::sqrt(nlen2);
        d = 2.0 * vdotn_h1 / nlen2;
        vy = vy - d * l_seg;
        vz = vz - d * hh[seg];
      } else {
        nlen2 = l_seg * l_seg + hh[seg] * hh[seg];
        q = V2K * (-2.0) * vdotn_h2 / std::sqrt(nlen2);
        d = 2.0 * vdotn_h2 / nlen2;
        vy = vy + d * l_seg;
        vz = vz - d * hh[seg];
      }
      if ((i <= 2 && mx == 0) || (i > 2 && my == 0)) {
        x += hadj;
        break;
      }
      if (i <= 2) {
        double m = mx;
        if (m <= 0.0) m = std::fabs(mx * w1[0] / w1[seg]);
        double R = calc_reflectivity(q, R0, Qcx, alphax, m, W);
        if (R > 1e-10)
          prob *= R;
        else {
          n[9] = 0;
          return;
        }
      } else {
        double m = my;
        if (m <= 0.0) m = std::fabs(my * h1[0] / h1[seg]);
        double R = calc_reflectivity(q, R0, Qcy, alphay, m, W);
        if (R > 1e-10)
          prob *= R;
        else {
          n[9] = 0;
          return;
        }
      }
    }
    x += hadj;
  }
  n[0] = x; n[1] = y; n[2] = z; n[3] = vx; n[4] = vy; n[5] = vz;
  n[8] = t;
  n[9] = prob;
}

// ---------------------------------------------------------------- CSG (R8)
// acc/geometry/__init__.py:2, location.py:1-3
static const double EPS = 1e-10;
enum { OUTSIDE = 1, ONBORDER = 0, INSIDE = -1 };
static const int MAXI = 10;  // acc/geometry/arrow_intersect.py:16 max_intersections

struct Shape {
  const double* nodes;  // MCV_CSG_NODE_SIZE doubles per node
  int n;
};

// acc/geometry/locate.py:85-122 (sphere, cylinder, block)
static int locate_sphere(double x, double y, double z, double R) {
  double dist2 = x * x + y * y + z * z;
  if (dist2 > (R + EPS) * (R + EPS)) return OUTSIDE;
  if (dist2 < (R - EPS) * (R - EPS)) return INSIDE;
  return ONBORDER;
}
static int locate_cylinder(double x, double y, double z, double R, double H) {
  if (std::fabs(z) - H / 2. > EPS) return OUTSIDE;
  double r2 = x * x + y * y;
  if (r2 > (R + EPS) * (R + EPS)) return OUTSIDE;
  if (H / 2. - std::fabs(z) > EPS && r2 < (R - EPS) * (R - EPS)) return INSIDE;
  return ONBORDER;
}
static int locate_box(double x, double y, double z, double W, double H, double D) {
  if (std::fabs(x) - 0.5 * W > EPS) return OUTSIDE;
  if (std::fabs(y) - 0.5 * H > EPS) return OUTSIDE;
  if (std::fabs(z) - 0.5 * D > EPS) return OUTSIDE;
  if ((-0.5 * W < x && x < 0.5 * W) && (-0.5 * H < y && y < 0.5 * H) && (-0.5 * D < z && z < 0.5 * D))
    return INSIDE;
  return ONBORDER;
}

// acc/geometry/locate.py:68-82,125-151 (union, difference, intersection)
static int locate(const Shape& s, int idx, double x, double y, double z) {
  const double* nd = s.nodes + MCV_CSG_NODE_SIZE * idx;
  int type = (int)nd[0];
  switch (type) {
    case MCV_CSG_SPHERE: return locate_sphere(x, y, z, nd[1]);
    case MCV_CSG_CYLINDER: return locate_cylinder(x, y, z, nd[1], nd[2]);
    case MCV_CSG_BLOCK: return locate_box(x, y, z, nd[1], nd[2], nd[3]);
    case MCV_CSG_UNION: {
      bool isoutside = true;
      int l = locate(s, (int)nd[4], x, y, z);
      if (l == INSIDE) return INSIDE;
      isoutside = isoutside && l == OUTSIDE;
      l = locate(s, (int)nd[5], x, y, z);
      if (l == INSIDE) return INSIDE;
      isoutside = isoutside && l == OUTSIDE;
      if (isoutside) return OUTSIDE;
      return ONBORDER;
    }
    case MCV_CSG_DIFFERENCE: {
      int p1 = locate(s, (int)nd[4], x, y, z);
      int p2 = locate(s, (int)nd[5], x, y, z);
      if (p1 == OUTSIDE || p2 == INSIDE) return OUTSIDE;
      if (p1 == ONBORDER || p2 == ONBORDER) return ONBORDER;
      return INSIDE;
    }
    case MCV_CSG_INTERSECTION: {
      bool isinside = true;
      int l = locate(s, (int)nd[4], x, y, z);
      if (l == OUTSIDE) return OUTSIDE;
      isinside = isinside && l == INSIDE;
      l = locate(s, (int)nd[5], x, y, z);
      if (l == OUTSIDE) return OUTSIDE;
      isinside = isinside && l == INSIDE;
      if (isinside) return INSIDE;
      return ONBORDER;
    }
  }
  return OUTSIDE;
}

// acc/geometry/_utils.py:4-32
static int insert_into_sorted_list(double d, double* l, int N, int len) {
  if (N >= len) return N;
  if (N == 0) {
    l[0] = d;
    return 1;
  }
  int minidx = 0, maxidx = N - 1, idx = minidx;
  while (minidx < maxidx - 1) {
    int mididx = (minidx + maxidx) / 2;
    double mid = l[mididx];
    if (d > mid)
      minidx = mididx;
    else
      maxidx = mididx;
  }
  if (minidx == maxidx) {
    idx = minidx;
    if (d > l[idx]) idx = idx + 1;
  } else {
    if (d >= l[maxidx])
      idx = maxidx + 1;
    else if (d <= l[minidx])
      idx = minidx;
    else
      idx = minidx + 1;
  }
  for (int i = N; i > idx; --i) l[i] = l[i - 1];
  l[idx] = d;
  return N + 1;
}

// acc/geometry/arrow_intersect.py:154-168
static void intersect_sphere(double x, double y, double z, double vx, double vy, double vz, double R, double& t1,
                             double& t2) {
  double r0dotv = x * vx + y * vy + z * vz;
  double r0Xv_x = y * vz - z * vy, r0Xv_y = z * vx - x * vz, r0Xv_z = x * vy - y * vx;
  double v2 = vx * vx + vy * vy + vz * vz;
  double r0Xv2 = r0Xv_x * r0Xv_x + r0Xv_y * r0Xv_y + r0Xv_z * r0Xv_z;
  double b2m4ac = v2 * R * R - r0Xv2;
  t1 = NAN;
  t2 = NAN;
  if (b2m4ac < 0) return;
  double sq = std::sqrt(b2m4ac);
  t1 = -(r0dotv + sq) / v2;
  t2 = -(r0dotv - sq) / v2;
}

// acc/geometry/arrow_intersect.py:178-204
static int cyl_side(double x, double y, double z, double vx, double vy, double vz, double R, double H, double& ta,
                    double& tb) {
  double a = vx * vx + vy * vy, b = x * vx + y * vy, c = x * x + y * y - R * R;
  double k = b * b - a * c, hh = H / 2.;
  ta = tb = NAN;
  if (k < 0 || a == 0) return 0;
  if (k == 0) {
    double t = -b / a;
    if (std::fabs(z + vz * t) < hh) {
      ta = t;
      return 1;
    }
  }
  k = std::sqrt(k);
  double t1 = (-b + k) / a, t2 = (-b - k) / a;
  bool g1 = std::fabs(z + vz * t1) < hh, g2 = std::fabs(z + vz * t2) < hh;
  if (g1) {
    if (g2) {
      ta = t2;
      tb = t1;
      return 2;
    }
    ta = t1;
    return 1;
  }
  if (g2) {
    ta = t2;
    return 1;
  }
  return 0;
}

// acc/geometry/arrow_intersect.py:213-239
static int cyl_topbottom(double x, double y, double z, double vx, double vy, double vz, double R, double H,
                         double& ta, double& tb) {
  ta = tb = NAN;
  if (vz == 0) return 0;
  double hh = H / 2, r2 = R * R;
  double t1 = (hh - z) / vz;
  double x1 = x + vx * t1, y1 = y + vy * t1;
  bool g1 = x1 * x1 + y1 * y1 <= r2;
  double t2 = (-hh - z) / vz;
  x1 = x + vx * t2;
  y1 = y + vy * t2;
  bool g2 = x1 * x1 + y1 * y1 <= r2;
  if (g1) {
    if (g2) {
      if (t1 < t2) {
        ta = t1;
        tb = t2;
      } else {
        ta = t2;
        tb = t1;
      }
      return 2;
    }
    ta = t1;
    return 1;
  }
  if (g2) {
    ta = t2;
    return 1;
  }
  return 0;
}

static inline double pymin(double a, double b) { return b < a ? b : a; }  // Python min(a, b)
static inline double pymax(double a, double b) { return b > a ? b : a; }  // Python max(a, b)

// acc/geometry/arrow_intersect.py:241-267 (n==0 falls through to the n==4 branch: all NaN)
static void intersect_cylinder(double x, double y, double z, double vx, double vy, double vz, double R, double H,
                               double& o1, double& o2) {
  double t1, t2, t3, t4;
  int n1 = cyl_side(x, y, z, vx, vy, vz, R, H, t1, t2);
  int n2 = cyl_topbottom(x, y, z, vx, vy, vz, R, H, t3, t4);
  int n = n1 + n2;
  if (n == 1) {
    if (n1 == 1) { o1 = t1; o2 = t1; }
    else { o1 = t3; o2 = t3; }
  } else if (n == 2) {
    if (n1 == 1) {
      if (t1 > t3) { o1 = t3; o2 = t1; }
      else { o1 = t1; o2 = t3; }
    } else if (n1 == 0) { o1 = t3; o2 = t4; }
    else { o1 = t1; o2 = t2; }
  } else if (n == 3) {
    if (n1 == 1) { o1 = pymin(t1, t3); o2 = pymax(t1, t4); }
    else { o1 = pymin(t3, t1); o2 = pymax(t3, t2); }
  } else {
    o1 = pymin(t1, t3);
    o2 = pymax(t2, t4);
  }
}

// acc/geometry/onbox.py:8-56 (the reference computes this in float32; restated in float64)
static double intersect_rectangle(double rx, double ry, double rz, double vx, double vy, double vz, double X, double Y) {
  double t = -rz / vz;
  double r1x = rx + vx * t, r1y = ry + vy * t;
  if (std::fabs(r1x) > X / 2 || std::fabs(r1y) > Y / 2) t = NAN;
  return t;
}
static void update_intersections(double& t1, double& t2, double t) {
  if (!std::isfinite(t)) return;
  if (!std::isfinite(t1)) { t1 = t; return; }
  if (!std::isfinite(t2)) {
    if (t > t1) { t2 = t; }
    else { t2 = t1; t1 = t; }
    return;
  }
  if (t < t1) { t1 = t; return; }
  if (t < t2) return;
  t2 = t;
}
static void intersect_box(double x, double y, double z, double vx, double vy, double vz, double X, double Y, double Z,
                          double& t1, double& t2) {
  t1 = NAN;
  t2 = NAN;
  if (vz != 0) {
    update_intersections(t1, t2, intersect_rectangle(x, y, z - Z / 2, vx, vy, vz, X, Y));
    update_intersections(t1, t2, intersect_rectangle(x, y, z + Z / 2, vx, vy, vz, X, Y));
  }
  if (vx != 0) {
    update_intersections(t1, t2, intersect_rectangle(y, z, x - X / 2, vy, vz, vx, Y, Z));
    update_intersections(t1, t2, intersect_rectangle(y, z, x + X / 2, vy, vz, vx, Y, Z));
  }
  if (vy != 0) {
    update_intersections(t1, t2, intersect_rectangle(z, x, y - Y / 2, vz, vx, vy, Z, X));
    update_intersections(t1, t2, intersect_rectangle(z, x, y + Y / 2, vz, vx, vy, Z, X));
  }
}

// ArrowIntersectFuncFactory (acc/geometry/arrow_intersect.py:18-123)
static int intersect(const Shape& s, int idx, double x, double y, double z, double vx, double vy, double vz,
                     double* ts) {
  const double* nd = s.nodes + MCV_CSG_NODE_SIZE * idx;
  int type = (int)nd[0];
  double t1, t2;
  switch (type) {
    case MCV_CSG_SPHERE:
      intersect_sphere(x, y, z, vx, vy, vz, nd[1], t1, t2);
      if (std::isnan(t1)) return 0;
      ts[0] = t1;
      ts[1] = t2;
      return 2;
    case MCV_CSG_CYLINDER:
      intersect_cylinder(x, y, z, vx, vy, vz, nd[1], nd[2], t1, t2);
      if (std::isnan(t1)) return 0;
      ts[0] = t1;
      ts[1] = t2;
      return 2;
    case MCV_CSG_BLOCK:
      intersect_box(x, y, z, vx, vy, vz, nd[1], nd[2], nd[3], t1, t2);
      if (std::isnan(t1)) return 0;
      ts[0] = t1;
      ts[1] = t2;
      return 2;
    default: {
      // intersectComposite (acc/geometry/arrow_intersect.py:126-142)
      double ts1[MAXI], ts2[MAXI];
      int N1 = intersect(s, (int)nd[4], x, y, z, vx, vy, vz, ts1);
      int N2 = intersect(s, (int)nd[5], x, y, z, vx, vy, vz, ts2);
      int N = 0;
      for (int i = 0; i < N1; ++i) {
        double t = ts1[i];
        if (locate(s, idx, x + vx * t, y + vy * t, z + vz * t) == ONBORDER) {
          if (N < MAXI) ts[N] = t;
          N += 1;
        }
      }
      for (int i = 0; i < N2; ++i) {
        double t = ts2[i];
        if (locate(s, idx, x + vx * t, y + vy * t, z + vz * t) == ONBORDER) {
          if (N < MAXI) ts[N] = t;
          N = insert_into_sorted_list(t, ts, N, MAXI);
        }
      }
      return N;
    }
  }
}

// ---------------------------------------------------------------- propagation helpers
// acc/geometry/propagation.py:9-191
static int forward_intersect(const Shape& s, double x, double y, double z, double vx, double vy, double vz,
                             double* ts) {
  int N = intersect(s, 0, x, y, z, vx, vy, vz, ts);
  int Nm = N;
  for (int i = 0; i < N; ++i)
    if (ts[i] >= 0) {
      Nm = i;
      break;
    }
  int Np = N - Nm;
  for (int i = 0; i < Np; ++i) ts[i] = ts[i + Nm];
  return Np;
}

static bool is_exiting(const Shape& s, const double* n) {
  double ts[MAXI];
  double x = n[0], y = n[1], z = n[2], vx = n[3], vy = n[4], vz = n[5];
  int loc = locate(s, 0, x, y, z);
  if (loc == INSIDE) return false;
  int N = forward_intersect(s, x, y, z, vx, vy, vz, ts);
  if (N == 0) return true;
  if (loc == OUTSIDE) return false;
  double previous = 0.;
  for (int it = 0; it < N; ++it) {
    double now = ts[it];
    double middle = (previous + now) / 2.;
    if (locate(s, 0, x + vx * middle, y + vy * middle, z + vz * middle) != ONBORDER) return false;
    previous = now;
  }
  return true;
}

static void propagate_out(const Shape& s, double* n) {
  double ts[MAXI];
  int N = forward_intersect(s, n[0], n[1], n[2], n[3], n[4], n[5], ts);
  if (N == 0) return;
  prop_dt_inplace(n, ts[N - 1]);
}

static void propagate_to_next_incident_surface(const Shape& s, double* n) {
  double ts[MAXI];
  double x = n[0], y = n[1], z = n[2], vx = n[3], vy = n[4], vz = n[5];
  int loc = locate(s, 0, x, y, z);
  if (loc == INSIDE) return;
  int N = forward_intersect(s, x, y, z, vx, vy, vz, ts);
  if (N == 0) return;
  double t = 0;
  if (loc == OUTSIDE)
    t = ts[0];
  else {
    bool found = false;
    for (int i = 0; i < N; ++i) {
      int loc1 = locate(s, 0, x + ts[i] / 2 * vx, y + ts[i] / 2 * vy, z + ts[i] / 2 * vz);
      if (loc1 == INSIDE) return;
      if (loc1 == OUTSIDE) {
        t = ts[i];
        found = true;
        break;
      }
    }
    if (!found) return;
  }
  prop_dt_inplace(n, t);
}

static void propagate_to_next_exiting_surface(const Shape& s, double* n) {
  double ts[MAXI];
  double x = n[0], y = n[1], z = n[2], vx = n[3], vy = n[4], vz = n[5];
  int loc = locate(s, 0, x, y, z);
  double t;
  if (loc == INSIDE) {
    forward_intersect(s, x, y, z, vx, vy, vz, ts);
    t = ts[0];
  } else {
    if (loc == OUTSIDE) {
      propagate_to_next_incident_surface(s, n);
      x = n[0]; y = n[1]; z = n[2]; vx = n[3]; vy = n[4]; vz = n[5];
    }
    int N = forward_intersect(s, x, y, z, vx, vy, vz, ts);
    bool found = false;
    t = 0;
    for (int i = 0; i < N; ++i) {
      int loc1 = locate(s, 0, x + ts[i] / 2 * vx, y + ts[i] / 2 * vy, z + ts[i] / 2 * vz);
      if (loc1 == OUTSIDE) return;
      if (loc1 == INSIDE) {
        t = ts[i];
        found = true;
        break;
      }
    }
    if (!found) return;
  }
  prop_dt_inplace(n, t);
}

static double tof_before_first_exit(const Shape& s, const double* n) {
  double ts[MAXI];
  double x = n[0], y = n[1], z = n[2], vx = n[3], vy = n[4], vz = n[5];
  int loc = locate(s, 0, x, y, z);
  if (loc == OUTSIDE) return 0.;
  int N = forward_intersect(s, x, y, z, vx, vy, vz, ts);
  if (N == 0) return 0.;
  if (loc == INSIDE) return ts[0];
  double t = 0.;
  for (int i = 0; i < N; ++i) {
    int loc1 = locate(s, 0, x + ts[i] / 2 * vx, y + ts[i] / 2 * vy, z + ts[i] / 2 * vz);
    if (loc1 == OUTSIDE) return 0.;
    if (loc1 == INSIDE) {
      t = ts[i];
      break;
    }
  }
  return t;
}

// ---------------------------------------------------------------- HSS time bookkeeping
// acc/components/samples/homogeneous_single_scatterer.py:121-195
static void calc_time_to_point_of_scattering(const double* ts, int ninter, double rand, double& total_time_in_shape,
                                             double& ttt_sp, double& ttis_sp) {
  double tps[MAXI] = {0}, tpsg[MAXI] = {0};
  double T = 0.0, time_to_enter = -1.0;
  int sign = 0, ip = 0;
  for (int i = 0; i < ninter; ++i) {
    double t = ts[i];
    if (t < 0) continue;
    if (i % 2 == 0 && time_to_enter < 0.0) time_to_enter = t;
    if (sign == 0)
      sign = (i % 2 == 0) ? -1 : 1;
    else
      sign *= -1;
    T += sign * t;
    if (sign > 0) {
      tps[ip] = T;
      tpsg[ip] = t;
      ip += 1;
    }
  }
  total_time_in_shape = T;
  double ttistsp = rand * total_time_in_shape;
  int i = ninter - 1;  // numba reuses the first loop's variable if the second loop does not run
  for (int k = 0; k < ip; ++k) {
    i = k;
    if (ttistsp < tps[k]) break;
  }
  ttt_sp = ttistsp - tps[i] + tpsg[i];
  ttis_sp = ttistsp;
}

static double total_time_in_shape(const double* ts, int ninter) {
  double T = 0;
  int sign = 0;
  for (int i = 0; i < ninter; ++i) {
    double t = ts[i];
    if (t < 0) continue;
    if (sign == 0)
      sign = (i % 2 == 0) ? -1 : 1;
    else
      sign *= -1;
    T += sign * t;
  }
  return T;
}

static double time_to_enter(const double* ts, int ninter) {
  for (int i = 0; i < ninter; ++i) {
    double t = ts[i];
    if (t < 0) continue;
    return (i % 2 == 0) ? t : 0;
  }
  return 0;
}

// ---------------------------------------------------------------- kernels (R11-R13)
struct Sample {
  Shape shape;
  int kernel;
  double mu, sigma, min_prob;
  const double* kp;  // kernel params
  const double* kdata;
  int64_t kn0, kn1;
  int max_scattered, max_ms_loops, max_ms_loops_path1;
};

// acc/kernels/__init__.py:149-161: sigma constant; mu/v*2200
static inline double calc_scattering_coeff(const Sample& S, const double* n) { (void)n; return S.sigma; }
static inline double calc_absorption_coeff(const Sample& S, const double* n) {
  double v = length(n + 3);
  return S.mu / v * 2200;
}

// acc/kernels/isotropic.py:5-21 (+ acc/kernels/__init__.py:36-44: p *= sigma)
static void isotropic_scatter(Rng& rng, double* n, double sigma) {
  n[9] *= sigma;
  double cos_t = (double)uniform_f32(rng) * 2 - 1;
  double phi = (double)uniform_f32(rng) * (2 * PI);
  if (cos_t > 1) cos_t = 1;
  double sin_t = std::sqrt(1 - cos_t * cos_t);
  double sin_p = TR_SIN(phi), cos_p = TR_COS(phi);
  double vx = n[3], vy = n[4], vz = n[5];
  double vi = std::sqrt(vx * vx + vy * vy + vz * vz);
  n[3] = vi * sin_t * cos_p;
  n[4] = vi * sin_t * sin_p;
  n[5] = vi * cos_t;
}

// bilinear interpolation on a grid of node values, index clamping as
// acc/kernels/SANS2D_ongrid.py:16-56
static double grid_interp(const double* g, int64_t nx, int64_t ny, double x, double xmin, double xmax, double y,
                          double ymin, double ymax) {
  double dx = (xmax - xmin) / nx, dy = (ymax - ymin) / ny;
  int64_t ix = std::max<int64_t>(0, (int64_t)((x - xmin) / dx));
  int64_t iy = std::max<int64_t>(0, (int64_t)((y - ymin) / dy));
  ix = std::min<int64_t>(ix, nx - 2);
  iy = std::min<int64_t>(iy, ny - 2);
  double f00 = g[ix * ny + iy], f10 = g[(ix + 1) * ny + iy], f01 = g[ix * ny + iy + 1],
         f11 = g[(ix + 1) * ny + iy + 1];
  double fx = (x - xmin) / dx - ix, fy = (y - ymin) / dy - iy;
  double a00 = f00, a10 = f10 - f00, a01 = f01 - f00, a11 = f11 - f10 - f01 + f00;
  return a00 + a10 * fx + a01 * fy + a11 * fx * fy;
}

// Phonon-DOS kernel (no reference): S(Q,E) = Q^2 * sum_i A_i exp(-(E-c_i)^2/(2 w_i^2))
static double dos_S(const double* kp, double Q, double E) {
  double g = 0;
  for (int k = 0; k < 3; ++k) {
    double A = kp[4 + 3 * k], c = kp[5 + 3 * k], w = kp[6 + 3 * k];
    double u = (E - c) / w;
    g += A * TR_EXP(-0.5 * u * u);
  }
  return Q * Q * g;
}

// S(Q,E) kernels (no reference).  Sampling and kinematic closure follow the
// E_Q kernel (acc/kernels/E_Q.py:37-99) with E drawn uniformly as well; a
// kinematically forbidden draw sets p = 0 (one draw, no retry, so the estimate
// stays unbiased).  Weight: sigma * S(Q,E) (vf/vi) Q dQ dE / (2 ki kf).
static void sqe_scatter(Rng& rng, double* n, const Sample& S) {
  n[9] *= S.sigma;
  const double* kp = S.kp;
  double Qmin = kp[0], Qmax = kp[1], Emin = kp[2], Emax = kp[3];
  double v[3] = {n[3], n[4], n[5]};
  double vi = length(v);
  double Ei = vi * vi * VS2E;
  double Q = (double)uniform_f32(rng) * (Qmax - Qmin);
  Q += Qmin;
  double E = (double)uniform_f32(rng) * (Emax - Emin);
  E += Emin;
  double Ef = Ei - E;
  if (Ef <= 0) {
    n[9] = 0;
    return;
  }
  double vf = std::sqrt(Ef) * SE2V;
  double ki = V2K * vi, kf = V2K * vf;
  double cost = (ki * ki + kf * kf - Q * Q) / (2 * ki * kf);
  double cost2 = cost * cost;
  if (cost2 > 1) {
    n[9] = 0;
    return;
  }
  double sint = std::sqrt(1.0 - cost2);
  double phi = (double)uniform_f32(rng) * 2.0 * PI;
  double v2 = vf * sint * TR_COS(phi), v3 = vf * sint * TR_SIN(phi), v1 = vf * cost;
  double e1[3], e2[3], e3[3], norm[3];
  copy3(v, e1);
  normalize(e1);
  if (std::fabs(e1[0]) > 1e-1 || std::fabs(e1[1]) > 1e-1) {
    norm[0] = 0; norm[1] = 0; norm[2] = 1;
  } else {
    norm[0] = 1; norm[1] = 0; norm[2] = 0;
  }
  cross(norm, e1, e2);
  normalize(e2);
  cross(e1, e2, e3);
  scale(e1, v1);
  scale(e2, v2);
  scale(e3, v3);
  n[3] = e1[0] + e2[0] + e3[0];
  n[4] = e1[1] + e2[1] + e3[1];
  n[5] = e1[2] + e2[2] + e3[2];
  double Sv = (S.kernel == MCV_KERNEL_SQE_GRID) ? grid_interp(S.kdata, S.kn0, S.kn1, Q, Qmin, Qmax, E, Emin, Emax)
                                                : dos_S(kp, Q, E);
  double prob = Sv * (vf / vi) * Q * (Qmax - Qmin) * (Emax - Emin) / (kf * ki) / 2;
  n[9] *= prob;
}

static void kernel_scatter(Rng& rng, double* n, const Sample& S) {
  if (S.kernel == MCV_KERNEL_ISOTROPIC)
    isotropic_scatter(rng, n, S.sigma);
  else
    sqe_scatter(rng, n, S);
}

// ---------------------------------------------------------------- HSS (R9)
// acc/components/samples/homogeneous_single_scatterer.py:59-100
static void hss_propagate(Rng& rng, double* n, const Sample& S) {
  double x = n[0], y = n[1], z = n[2], vx = n[3], vy = n[4], vz = n[5];
  double ts[MAXI];
  int ninter = intersect(S.shape, 0, x, y, z, vx, vy, vz, ts);
  if (ninter < 2) return;
  if (ts[ninter - 1] <= 0) return;
  float rand = uniform_f32(rng);
  double tis, dt, ttis;
  calc_time_to_point_of_scattering(ts, ninter, (double)rand, tis, dt, ttis);
  if (dt <= 0) return;
  prop_dt_inplace(n, dt);
  double v = std::sqrt(vx * vx + vy * vy + vz * vz);
  double dist = v * ttis;
  double fulllen = v * tis;
  double sigma = calc_scattering_coeff(S, n);
  double mu = calc_absorption_coeff(S, n);
  double atten = TR_EXP(-(mu + sigma) * dist);
  double prob = fulllen * atten;
  n[9] *= prob;
  kernel_scatter(rng, n, S);
  if (n[9] <= 0) return;  // kernel absorb() is a no-op for these kernels
  x = n[0]; y = n[1]; z = n[2]; vx = n[3]; vy = n[4]; vz = n[5];
  ninter = intersect(S.shape, 0, x, y, z, vx, vy, vz, ts);
  double dt3 = total_time_in_shape(ts, ninter);
  prop_dt_inplace(n, dt3);
  sigma = calc_scattering_coeff(S, n);
  mu = calc_absorption_coeff(S, n);
  double atten2 = TR_EXP(-(mu + sigma) * v * dt3);
  n[9] *= atten2;
}

// ---------------------------------------------------------------- HMS (R10)
// acc/components/samples/homogeneous_multiple_scatterer.py:63-303
static inline void clone(const double* a, double* b) { std::memcpy(b, a, 10 * sizeof(double)); }

static double atten_coeff(const Sample& S, const double* n) {  // :63-67
  double sigma = calc_scattering_coeff(S, n);
  double mu = calc_absorption_coeff(S, n);
  return sigma + mu;
}

static void propagate_to_next_exiting_surface_with_attenuation(const Sample& S, double* n) {  // :69-79
  double vx = n[3], vy = n[4], vz = n[5];
  double tof = tof_before_first_exit(S.shape, n);
  double v = std::sqrt(vx * vx + vy * vy + vz * vz);
  double distance = tof * v;
  double atten = atten_coeff(S, n);
  double transmission_prob = TR_EXP(-atten * distance);
  propagate_to_next_exiting_surface(S.shape, n);
  n[9] *= transmission_prob;
}

// :81-121; out has >= 2 rows
static int interactM1(Rng& rng, const Sample& S, double* out, double* n) {
  double x = n[0], y = n[1], z = n[2], vx = n[3], vy = n[4], vz = n[5];
  int loc = locate(S.shape, 0, x, y, z);
  if (loc != INSIDE) propagate_to_next_incident_surface(S.shape, n);
  double tof = tof_before_first_exit(S.shape, n);
  if (tof == 0.) return 0;
  double v = std::sqrt(vx * vx + vy * vy + vz * vz);
  double distance = tof * v;
  double sigma = calc_scattering_coeff(S, n);
  double mu = calc_absorption_coeff(S, n);
  double transmission_prob = TR_EXP(-(mu + sigma) * distance);
  double* transmitted = out;
  clone(n, transmitted);
  propagate_to_next_exiting_surface(S.shape, transmitted);
  transmitted[9] *= transmission_prob;
  // absorption branch: temp = out[1] is overwritten by the scattered neutron
  // below and the kernels' absorb() is a no-op (acc/kernels/__init__.py:163-165)
  float rand = uniform_f32(rng);
  double xs = distance * (double)rand;
  double prob = distance * TR_EXP(-(mu + sigma) * xs);
  double* scattered = out + 10;
  clone(n, scattered);
  scattered[9] *= prob;
  prop_dt_inplace(scattered, xs / v);
  kernel_scatter(rng, scattered, S);
  return 2;
}

// :128-178
static int interactM_path1(Rng& rng, const Sample& S, double* out, int out_len, double* n) {
  if (n[9] < S.min_prob) return 0;
  const int M = S.max_scattered;
  std::vector<double> buf(2 * M * 10 + 2 * 10);
  double* tbs = buf.data();
  double* tbs2 = tbs + M * 10;
  double* scattered = tbs2 + M * 10;
  int N_tbs = 1;
  clone(n, tbs);
  int out_index = 0;
  int lim = std::min(M, out_len);
  for (int iloop = 0; iloop < S.max_ms_loops_path1; ++iloop) {
    int s2 = 0;
    for (int ineutron = 0; ineutron < N_tbs; ++ineutron) {
      double* n1 = tbs + 10 * ineutron;
      int ns1 = interactM1(rng, S, scattered, n1);
      for (int is = 0; is < ns1; ++is) {
        double* sc1 = scattered + 10 * is;
        if (sc1[9] < S.min_prob) continue;
        if (locate(S.shape, 0, sc1[0], sc1[1], sc1[2]) != INSIDE) {
          clone(sc1, out + 10 * out_index);
          out_index += 1;
          if (out_index >= lim) return out_index;
          continue;
        }
        clone(sc1, tbs2 + 10 * s2);
        s2 += 1;
        if (s2 >= M) break;
      }
      if (s2 >= M) break;
    }
    std::swap(tbs, tbs2);
    N_tbs = s2;
    if (!N_tbs) break;
  }
  for (int i = 0; i < N_tbs; ++i) {
    if (out_index >= lim) break;
    clone(tbs + 10 * i, out + 10 * out_index);
    out_index += 1;
  }
  return out_index;
}

// :214-273
static int scatterM(Rng& rng, const Sample& S, double* out, int out_len, double* n) {
  if (is_exiting(S.shape, n)) {
    clone(n, out);
    return 1;
  }
  const int M = S.max_scattered;
  std::vector<double> buf(3 * M * 10);
  double* tbs = buf.data();
  double* tbs2 = tbs + M * 10;
  double* scattered = tbs2 + M * 10;
  int N_tbs = 1;
  clone(n, tbs);
  int out_index = 0;
  int lim = std::min(M, out_len);
  for (int iloop = 0; iloop < S.max_ms_loops; ++iloop) {
    int s2 = 0;
    for (int ineutron = 0; ineutron < N_tbs; ++ineutron) {
      double* n1 = tbs + 10 * ineutron;
      int ns1 = interactM_path1(rng, S, scattered, M, n1);
      for (int is = 0; is < ns1; ++is) {
        double* sc1 = scattered + 10 * is;
        if (sc1[9] < S.min_prob) continue;
        if (!is_moving(sc1)) continue;
        if (locate(S.shape, 0, sc1[0], sc1[1], sc1[2]) == INSIDE)
          propagate_to_next_exiting_surface_with_attenuation(S, sc1);
        if (is_exiting(S.shape, sc1)) {
          clone(sc1, out + 10 * out_index);
          out_index += 1;
          if (out_index >= lim) return out_index;
          continue;
        }
        clone(sc1, tbs2 + 10 * s2);
        s2 += 1;
        if (s2 >= M) break;
      }
      if (s2 >= M) break;
    }
    std::swap(tbs, tbs2);
    N_tbs = s2;
    if (!N_tbs) break;
  }
  for (int i = 0; i < N_tbs; ++i) {
    if (out_index >= lim) break;
    double* nn = tbs + 10 * i;
    double x = nn[0], y = nn[1], z = nn[2];
    propagate_out(S.shape, nn);
    double xo = nn[0], yo = nn[1], zo = nn[2];
    double dist = std::sqrt((xo - x) * (xo - x) + (yo - y) * (yo - y) + (zo - z) * (zo - z));
    double atten = atten_coeff(S, nn);
    double transmission_prob = TR_EXP(-atten * dist);
    nn[9] *= transmission_prob;
    clone(nn, out + 10 * out_index);
    out_index += 1;
  }
  return out_index;
}

// ---------------------------------------------------------------- slit, beamstop
// acc/components/optics/slit.py:71-108 and beamstop.py:75-112 (absorb = p -> -1)
static void aperture(double* n, const double* P, bool beamstop) {
  double xmin = P[0], xmax = P[1], ymin = P[2], ymax = P[3], radius_squared = P[4], cut = P[5];
  double x = n[0], y = n[1], z = n[2], vx = n[3], vy = n[4], vz = n[5];
  if ((z < 0 && vz <= 0) || (z > 0 && vz >= 0)) return;
  double t = n[8], prob = n[9];
  if (prob < cut) {
    n[9] = -1;
    return;
  }
  if (z != 0) {
    double dt = -z / vz;
    x += vx * dt;
    y += vy * dt;
    t += dt;
  }
  if (radius_squared != 0) {
    if (beamstop ? (x * x + y * y <= radius_squared) : (x * x + y * y > radius_squared)) {
      n[9] = -1;
      return;
    }
  } else {
    bool within = xmin <= x && x <= xmax && ymin <= y && y <= ymax;
    if (beamstop ? within : !within) {
      n[9] = -1;
      return;
    }
  }
  n[0] = x;
  n[1] = y;
  n[2] = 0.;
  n[8] = t;
}

// One monitor event: (count, sum p, sum p^2) into bin b of an nb-bin block.
// Runs whose per-thread histogram copies would be too large share one
// histogram between threads; their adds are atomic.
static thread_local bool g_shared_hist = false;
static inline void hist_event(double* out, int64_t nb, int64_t b, double p) {
  const double p2 = p * p;
  if (g_shared_hist) {
#ifdef ORC_FLOP_BUILD  // the counting type is not an atomic operand
#pragma omp critical(orc_hist)
    {
      out[b] += 1;
      out[nb + b] += p;
      out[2 * nb + b] += p2;
    }
#else
#pragma omp atomic
    out[b] += 1;
#pragma omp atomic
    out[nb + b] += p;
#pragma omp atomic
    out[2 * nb + b] += p2;
#endif
  } else {
    out[b] += 1;
    out[nb + b] += p;
    out[2 * nb + b] += p2;
  }
}

// ---------------------------------------------------------------- monitors (R14-R17)
// acc/components/monitors/psd_monitor.py:66-83
static void psd_monitor(const double* n, const double* P, int64_t nx, int64_t ny, double* out) {
  double xmin = P[0], xmax = P[1], ymin = P[2], ymax = P[3];
  double t0 = n[8], x, y, t;
  prop_z0(n, x, y, t);
  if (t0 > t) return;
  double p = n[9];
  // the reference's window test admits NaN (then indexes out of bounds); NaN is rejected here
  if (!(x > xmin && x < xmax && y > ymin && y < ymax)) return;
  int64_t ix = (int64_t)std::floor((x - xmin) / (xmax - xmin) * nx);
  int64_t iy = (int64_t)std::floor((y - ymin) / (ymax - ymin) * ny);
  if (ix >= nx) ix = nx - 1;  // the reference writes past the row here (no bounds check)
  if (iy >= ny) iy = ny - 1;
  int64_t nb = nx * ny, b = ix * ny + iy;
  hist_event(out, nb, b, p);
}

// acc/components/monitors/posdiv_monitor.py:67-89 (NaN rejected and indices
// clamped, as for the PSD monitor)
static void posdiv_monitor(const double* n, const double* P, int64_t npos, int64_t ndiv, double* out) {
  double xmin = P[0], xmax = P[1], ymin = P[2], ymax = P[3], maxdiv = P[4];
  double t0 = n[8], x, y, t;
  prop_z0(n, x, y, t);
  if (t0 > t) return;
  double p = n[9];
  double vx = n[3], vz = n[5];
  if (!(x > xmin && x < xmax && y > ymin && y < ymax)) return;
  double div = TR_ATAN(vx / vz) * (180. / PI);
  if (!(div < maxdiv && div > -maxdiv)) return;
  int64_t ix = (int64_t)std::floor((x - xmin) / (xmax - xmin) * npos);
  int64_t idiv = (int64_t)std::floor((div + maxdiv) / (2 * maxdiv) * ndiv);
  ix = std::min<int64_t>(std::max<int64_t>(ix, 0), npos - 1);
  idiv = std::min<int64_t>(std::max<int64_t>(idiv, 0), ndiv - 1);
  int64_t nb = npos * ndiv, b = ix * ndiv + idiv;
  hist_event(out, nb, b, p);
}

// acc/components/monitors/psd_monitor_4pi.py:56-79 with
// acc/geometry/arrow_intersect.py:155-168 (moves the neutron onto the sphere;
// a line that misses the sphere is not counted here)
static void psd4pi_monitor(double* n, const double* P, int64_t nphi, int64_t ntheta, double* out) {
  double radius = P[4];
  double x = n[0], y = n[1], z = n[2], vx = n[3], vy = n[4], vz = n[5];
  double r0dotv = x * vx + y * vy + z * vz;
  double r0Xv_x = y * vz - z * vy, r0Xv_y = z * vx - x * vz, r0Xv_z = x * vy - y * vx;
  double v2 = vx * vx + vy * vy + vz * vz;
  double r0Xv2 = r0Xv_x * r0Xv_x + r0Xv_y * r0Xv_y + r0Xv_z * r0Xv_z;
  double b2m4ac = v2 * radius * radius - r0Xv2;
  if (!(b2m4ac >= 0)) return;
  double sq = std::sqrt(b2m4ac);
  double t1 = -(r0dotv + sq) / v2, t2 = -(r0dotv - sq) / v2;
  if (t2 < 0) return;
  if (t1 < 0) t1 = t2;
  prop_dt_inplace(n, t1);
  double phi = TR_ATAN2(n[0], n[2]);
  double theta = TR_ASIN(n[1] / radius);
  int64_t iphi = (int64_t)(nphi * (phi / 2 / PI + 0.5));
  int64_t itheta = (int64_t)(ntheta * (theta + PI / 2) / PI + 0.5);
  iphi = std::min<int64_t>(std::max<int64_t>(iphi, 0), nphi - 1);
  itheta = std::min<int64_t>(std::max<int64_t>(itheta, 0), ntheta - 1);
  double p = n[9];
  int64_t nb = nphi * ntheta, b = iphi * ntheta + itheta;
  hist_event(out, nb, b, p);
}

// acc/components/monitors/wavelength_monitor.py:52-74
static void wavelength_monitor(const double* n, const double* P, int64_t nchan, double* out) {
  double xmin = P[0], xmax = P[1], ymin = P[2], ymax = P[3], Lmin = P[4], Lmax = P[5];
  double t0 = n[8], x, y, t;
  prop_z0(n, x, y, t);
  if (t0 > t) return;
  double p = n[9];
  double vx = n[3], vy = n[4], vz = n[5];
  if (!(x > xmin && x < xmax && y > ymin && y < ymax)) return;
  double v = std::sqrt(vx * vx + vy * vy + vz * vz);
  double L = 2 * PI / (v * V2K);
  if (!(L > Lmin && L < Lmax)) return;
  int64_t iL = (int64_t)std::floor((L - Lmin) / (Lmax - Lmin) * nchan);
  if (iL >= nchan) iL = nchan - 1;
  hist_event(out, nchan, iL, p);
}

// acc/components/monitors/iqe_monitor.py:80-113
static void iqe_monitor(const double* n, const double* P, int64_t nQ, int64_t nE, double* out) {
  double Ei = P[0], Qmin = P[1], Qmax = P[2], Emin = P[3], Emax = P[4], max_in = P[5], min_in = P[6],
         max_out = P[7], min_out = P[8];
  double vx = n[3], vy = n[4], vz = n[5];
  double angle_in_plane = TR_ATAN2(vx, vz) / PI * 180.;
  double angle_out_of_plane;
  if (vy != 0) {
    double theta = TR_ATAN(std::sqrt(vx * vx + vz * vz) / vy) / PI * 180.;
    angle_out_of_plane = vy > 0 ? 90. - theta : 90. + theta;
  } else
    angle_out_of_plane = 0.;
  if (min_in < angle_in_plane && angle_in_plane < max_in && min_out < angle_out_of_plane &&
      angle_out_of_plane < max_out) {
    double vi = std::sqrt(Ei) * SE2V;
    double Q = std::sqrt(vx * vx + vy * vy + (vz - vi) * (vz - vi)) * V2K;
    double E = Ei - (vx * vx + vy * vy + vz * vz) * VS2E;
    if ((Q >= Qmin && Q < Qmax) && (E >= Emin && E < Emax)) {
      int64_t iQ = (int64_t)std::floor((Q - Qmin) / (Qmax - Qmin) * nQ);
      int64_t iE = (int64_t)std::floor((E - Emin) / (Emax - Emin) * nE);
      if (iQ >= nQ) iQ = nQ - 1;
      if (iE >= nE) iE = nE - 1;
      double p = n[9];
      int64_t nb = nQ * nE, b = iQ * nE + iE;
      hist_event(out, nb, b, p);
    }
  }
}

// Cylindrical tube-detector array (no reference; monitor conventions of
// psd_monitor.py: non-destructive, N/sum p/sum p^2).  Tubes are vertical (y)
// on a cylinder of `radius` around the component origin; tube j covers
// azimuth phi = atan2(x, z) in [phi_min + j dphi, phi_min + (j+1) dphi).
static void tube_detector(const double* n, const double* P, int64_t ntubes, int64_t npix, int64_t ntof,
                          double* out) {
  double R = P[0], H = P[1], phimin = P[2], phimax = P[3], tmin = P[4], tmax = P[5];
  double x = n[0], y = n[1], z = n[2], vx = n[3], vy = n[4], vz = n[5];
  double a = vx * vx + vz * vz;
  if (a == 0) return;
  double b = x * vx + z * vz, c = x * x + z * z - R * R;
  double k = b * b - a * c;
  if (k < 0) return;
  double dt = (-b + std::sqrt(k)) / a;
  if (dt < 0) return;
  double xh = x + vx * dt, yh = y + vy * dt, zh = z + vz * dt, tof = n[8] + dt;
  double phi = TR_ATAN2(xh, zh) / PI * 180.;
  if (!(phi >= phimin && phi < phimax)) return;
  if (!(yh > -H / 2 && yh < H / 2)) return;
  if (!(tof >= tmin && tof < tmax)) return;
  int64_t it = (int64_t)std::floor((phi - phimin) / (phimax - phimin) * ntubes);
  int64_t ip = (int64_t)std::floor((yh + H / 2) / H * npix);
  int64_t ib = (int64_t)std::floor((tof - tmin) / (tmax - tmin) * ntof);
  if (it >= ntubes) it = ntubes - 1;
  if (ip >= npix) ip = npix - 1;
  if (ib >= ntof) ib = ntof - 1;
  double p = n[9];
  int64_t nb = ntubes * npix * ntof, bin = (it * npix + ip) * ntof + ib;
  hist_event(out, nb, bin, p);
}

// ---------------------------------------------------------------- chain
struct Comp {
  int type;
  const double* P;
  const int64_t* IP;
  const double* data;
  int64_t nbins;
  int64_t hist_off;  // offset of this monitor's (3, nbins) block
  Sample sample;
};

static Sample make_sample(const mcv_comp_desc& d) {
  Sample S;
  S.shape.nodes = (const double*)d.data;
  S.shape.n = (int)d.iparams[0];
  S.kernel = (int)d.iparams[1];
  S.max_scattered = (int)d.iparams[2];
  S.max_ms_loops = (int)d.iparams[3];
  S.max_ms_loops_path1 = (int)d.iparams[4];
  S.kdata = (const double*)d.data + d.iparams[5];
  S.kn0 = d.iparams[6];
  S.kn1 = d.iparams[7];
  S.mu = d.params[0];
  S.sigma = d.params[1];
  S.min_prob = d.params[2];
  S.kp = (const double*)d.params + 3;
  return S;
}

struct Chain {
  std::vector<Comp> comps;
  const double* offsets;
  const double* rotmats;
  int64_t hist_total;
  int ms_index;  // index of the first multiple-scattering component, or -1
  bool gate;
};

// acc/components/sources/neutronfromstorage.py:70-82
static void storage_source(double* n, const int64_t* IP, const double* data, int64_t nidx) {
  const int64_t N = IP[0];
  std::memcpy(n, data + 10 * (nidx % N), 10 * sizeof(double));
}

static void run_component(const Chain& C, int i, Rng& rng, double* n, double* hist, int64_t nidx) {
  const Comp& c = C.comps[i];
  switch (c.type) {
    case MCV_ARM: break;
    case MCV_SRC_STORAGE: storage_source(n, c.IP, c.data, nidx); break;
    case MCV_SOURCE_SIMPLE: source_simple(rng, n, c.P); break;
    case MCV_SRC_IQE_NORMALIZER: iqe_normalizer_source(rng, n, c.P); break;
    case MCV_GUIDE: guide(n, c.P); break;
    case MCV_SLIT: aperture(n, c.P, false); break;
    case MCV_BEAMSTOP: aperture(n, c.P, true); break;
    case MCV_GUIDE_TAPERING: guide_tapering(n, c.P, c.IP[0], c.data); break;
    case MCV_MON_POSDIV: posdiv_monitor(n, c.P, c.IP[0], c.IP[1], hist + c.hist_off); break;
    case MCV_MON_PSD_4PI: psd4pi_monitor(n, c.P, c.IP[0], c.IP[1], hist + c.hist_off); break;
    case MCV_SAMPLE_HSS: hss_propagate(rng, n, c.sample); break;
    case MCV_MON_PSD: psd_monitor(n, c.P, c.IP[0], c.IP[1], hist + c.hist_off); break;
    case MCV_MON_WAVELENGTH: wavelength_monitor(n, c.P, c.IP[0], hist + c.hist_off); break;
    case MCV_MON_IQE: iqe_monitor(n, c.P, c.IP[0], c.IP[1], hist + c.hist_off); break;
    case MCV_DET_TUBES: tube_detector(n, c.P, c.IP[0], c.IP[1], c.IP[2], hist + c.hist_off); break;
    default: break;
  }
}

static inline bool dead(const Chain& C, const double* n) { return C.gate && !(n[9] > 0); }

// Components [from, end) on one neutron (the tail of the fused kernel body,
// generated by acc/run_script.py:56-93).  Returns true if it reached the end alive.
static bool run_tail(const Chain& C, int from, Rng& rng, double* n, double* hist, int64_t nidx) {
  int nc = (int)C.comps.size();
  for (int i = from; i < nc; ++i) {
    if (i > 0) apply_transformation(n, C.rotmats + 9 * (i - 1), C.offsets + 3 * (i - 1));
    run_component(C, i, rng, n, hist, nidx);
    if (dead(C, n)) return false;
  }
  return true;
}

// One neutron through the whole chain; final state left in n.  For a chain
// with a multiple-scattering sample, the sample's outputs each run the rest of
// the chain (the loop acc/run_script.py:87-92 generates); n then holds output 0's final state and
// the return value is the number of outputs.
static int run_neutron(const Chain& C, Rng& rng, double* n, double* hist, std::vector<double>& msbuf,
                       bool& alive, int64_t nidx) {
  int nc = (int)C.comps.size();
  if (C.ms_index < 0) {
    alive = run_tail(C, 0, rng, n, hist, nidx);
    return 1;
  }
  int ms = C.ms_index;
  for (int i = 0; i < ms; ++i) {
    if (i > 0) apply_transformation(n, C.rotmats + 9 * (i - 1), C.offsets + 3 * (i - 1));
    run_component(C, i, rng, n, hist, nidx);
    if (dead(C, n)) {
      alive = false;
      return 0;
    }
  }
  if (ms > 0) apply_transformation(n, C.rotmats + 9 * (ms - 1), C.offsets + 3 * (ms - 1));
  const Sample& S = C.comps[ms].sample;
  int M = S.max_scattered;
  double* out = msbuf.data();
  int num = scatterM(rng, S, out, M, n);
  alive = false;
  for (int k = 0; k < num; ++k) {
    double* o = out + 10 * k;
    bool a = true;
    if (dead(C, o))
      a = false;
    else if (ms + 1 < nc)
      a = run_tail(C, ms + 1, rng, o, hist, nidx);
    if (k == 0) {
      clone(o, n);
      alive = a;
    }
  }
  return num;
}

}  // namespace orc

using namespace orc;

static bool build_chain(const mcv_comp_desc* comps, int ncomp, const double* offsets, const double* rotmats,
                        uint32_t flags, Chain& C) {
  C.offsets = offsets;
  C.rotmats = rotmats;
  C.hist_total = 0;
  C.ms_index = -1;
  C.gate = (flags & MCV_FLAG_GATE_DEAD) != 0;
  for (int i = 0; i < ncomp; ++i) {
    Comp c;
    c.type = comps[i].type;
    c.P = (const double*)comps[i].params;
    c.IP = comps[i].iparams;
    c.data = (const double*)comps[i].data;
    c.nbins = comps[i].n_bins;
    c.hist_off = C.hist_total;
    C.hist_total += 3 * c.nbins;
    if (c.type == MCV_SAMPLE_HSS || c.type == MCV_SAMPLE_HMS) c.sample = make_sample(comps[i]);
    if (c.type == MCV_SAMPLE_HMS && C.ms_index < 0) C.ms_index = i;
    C.comps.push_back(c);
  }
  return true;
}

extern "C" {

void orc_set_trans_nudge(int ulps) { orc::g_nudge = ulps; }
void orc_set_pinned_nudge(int64_t nidx, const int32_t* dirs, int64_t ndirs) {
  orc::g_pin_idx = nidx;
  orc::g_pin_dir.assign(dirs, dirs + (ndirs > 0 ? ndirs : 0));
}

// ---- unit-level entry points (tests pin each against the reference's KATs)
void orc_rng_states(uint64_t seed, int64_t n, uint64_t* out) {
  std::vector<Rng> s(n > 0 ? n : 1);
  init_states(seed, n, s.data());
  for (int64_t i = 0; i < n; ++i) {
    out[2 * i] = s[i].s0;
    out[2 * i + 1] = s[i].s1;
  }
}

uint64_t orc_splitmix64(uint64_t seed) { return splitmix64(seed); }

// advance a state through k next() calls, writing raw outputs and float32 draws
void orc_rng_draws(uint64_t* state2, int64_t k, uint64_t* raw, float* f32) {
  Rng s{state2[0], state2[1]};
  for (int64_t i = 0; i < k; ++i) {
    Rng c = s;
    if (raw) raw[i] = next(c);
    f32[i] = uniform_f32(s);
  }
  state2[0] = s.s0;
  state2[1] = s.s1;
}

void orc_rng_jump(uint64_t* state2) {
  Rng s{state2[0], state2[1]};
  jump(s);
  state2[0] = s.s0;
  state2[1] = s.s1;
}

int orc_locate(const double* nodes, int nnodes, double x, double y, double z) {
  Shape s{nodes, nnodes};
  return locate(s, 0, x, y, z);
}

int orc_intersect(const double* nodes, int nnodes, double x, double y, double z, double vx, double vy, double vz,
                  double* ts) {
  Shape s{nodes, nnodes};
  return intersect(s, 0, x, y, z, vx, vy, vz, ts);
}

int orc_insert_into_sorted_list(double d, double* l, int N, int len) { return insert_into_sorted_list(d, l, N, len); }

void orc_hss_times(const double* ts, int ninter, double rand, double* out5) {
  double a, b, c;
  calc_time_to_point_of_scattering(ts, ninter, rand, a, b, c);
  out5[0] = a;
  out5[1] = b;
  out5[2] = c;
  out5[3] = total_time_in_shape(ts, ninter);
  out5[4] = time_to_enter(ts, ninter);
}

// op: 0 propagate_out, 1 tof_before_first_exit (returned), 2 to_next_incident,
//     3 to_next_exiting, 4 is_exiting (returned as 0/1)
double orc_propagation(int op, const double* nodes, int nnodes, double* n) {
  Shape s{nodes, nnodes};
  switch (op) {
    case 0: propagate_out(s, n); return 0;
    case 1: return tof_before_first_exit(s, n);
    case 2: propagate_to_next_incident_surface(s, n); return 0;
    case 3: propagate_to_next_exiting_surface(s, n); return 0;
    case 4: return is_exiting(s, n) ? 1.0 : 0.0;
  }
  return 0;
}

void orc_transform(double* n, const double* rot, const double* off) { apply_transformation(n, rot, off); }

// Run one component on one neutron with a given stream state (state2 updated).
// hist: (3, n_bins) for monitors.  For MCV_SAMPLE_HMS the return value is the
// output count and `out` (max_scattered x 10) receives the outputs; `which`
// selects 0 scatterM, 1 interactM_path1, 2 interactM1.
int orc_component(const mcv_comp_desc* d, double* n, uint64_t* state2, double* hist, double* out, int which) {
  Rng rng{state2[0], state2[1]};
  int ret = 0;
  switch (d->type) {
    case MCV_SOURCE_SIMPLE: source_simple(rng, n, (const double*)d->params); break;
    case MCV_SRC_IQE_NORMALIZER: iqe_normalizer_source(rng, n, (const double*)d->params); break;
    case MCV_GUIDE: guide(n, (const double*)d->params); break;
    case MCV_SLIT: aperture(n, (const double*)d->params, false); break;
    case MCV_BEAMSTOP: aperture(n, (const double*)d->params, true); break;
    case MCV_GUIDE_TAPERING: guide_tapering(n, (const double*)d->params, d->iparams[0], (const double*)d->data); break;
    case MCV_MON_POSDIV: posdiv_monitor(n, (const double*)d->params, d->iparams[0], d->iparams[1], hist); break;
    case MCV_MON_PSD_4PI: psd4pi_monitor(n, (const double*)d->params, d->iparams[0], d->iparams[1], hist); break;
    case MCV_SAMPLE_HSS: {
      Sample S = make_sample(*d);
      hss_propagate(rng, n, S);
      break;
    }
    case MCV_SAMPLE_HMS: {
      Sample S = make_sample(*d);
      if (which == 0)
        ret = scatterM(rng, S, out, S.max_scattered, n);
      else if (which == 1)
        ret = interactM_path1(rng, S, out, S.max_scattered, n);
      else
        ret = interactM1(rng, S, out, n);
      break;
    }
    case MCV_MON_PSD: psd_monitor(n, (const double*)d->params, d->iparams[0], d->iparams[1], hist); break;
    case MCV_MON_WAVELENGTH: wavelength_monitor(n, (const double*)d->params, d->iparams[0], hist); break;
    case MCV_MON_IQE: iqe_monitor(n, (const double*)d->params, d->iparams[0], d->iparams[1], hist); break;
    case MCV_DET_TUBES: tube_detector(n, (const double*)d->params, d->iparams[0], d->iparams[1], d->iparams[2], hist); break;
    default: return -1;  // not a component this entry point runs
  }
  state2[0] = rng.s0;
  state2[1] = rng.s1;
  return ret;
}

// Reference partition (acc/components/sources/SourceBase.py:36-52).
void orc_partition(uint64_t N, int64_t ntotalthreads, int32_t tpb, int64_t* out3) {
  if (ntotalthreads <= 0) ntotalthreads = 1000000;
  if (tpb <= 0) tpb = 512;
  int64_t ntot = (int64_t)std::min<uint64_t>(N, (uint64_t)ntotalthreads);
  int64_t nblocks = (ntot + tpb - 1) / tpb;
  int64_t nthreads = nblocks * tpb;
  int64_t npt = nthreads > 0 ? (int64_t)((N + nthreads - 1) / nthreads) : 0;
  out3[0] = nblocks;
  out3[1] = nthreads;
  out3[2] = npt;
}

// Full fused chain over streams [tid0, tid1) of the reference partition for N
// neutrons (acc/run_script.py:182-196 body; acc/components/sources/SourceBase.py:36-52
// partition).  hist: concatenated (3, n_bins) blocks in component order.
// dump: final states of neutron indices [dump_first, dump_first+dump_count).
// nthreads_omp <= 0 uses every core.  Returns the number of neutrons processed.
int64_t orc_run(const mcv_comp_desc* comps, int ncomp, const double* offsets, const double* rotmats, uint64_t N,
                uint64_t seed, int64_t ntotalthreads, int32_t tpb, uint32_t flags, int64_t tid0, int64_t tid1,
                int nthreads_omp, double* hist, int64_t dump_first, int64_t dump_count, double* dump_states,
                int32_t* dump_alive) {
  Chain C;
  build_chain(comps, ncomp, offsets, rotmats, flags, C);
  int64_t part[3];
  orc_partition(N, ntotalthreads, tpb, part);
  int64_t nthreads = part[1], npt = part[2];
  if (tid1 < 0 || tid1 > nthreads) tid1 = nthreads;
  if (tid0 < 0) tid0 = 0;
  if (tid1 <= tid0) return 0;
  std::vector<Rng> states(tid1);
  init_states(seed, tid1, states.data());
  int nt = nthreads_omp > 0 ? nthreads_omp : omp_get_max_threads();
  int64_t processed = 0;
  // per-thread histograms, or one shared (atomic adds) when the copies would
  // exceed ORC_SHARED_HIST_MB (default 512)
  const char* lim = std::getenv("ORC_SHARED_HIST_MB");
  const double limit_mb = lim && *lim ? std::atof(lim) : 512.0;
  const bool shared = nt > 1 && (double)C.hist_total * 8.0 * nt > limit_mb * (1 << 20);
  std::vector<std::vector<double>> hists(shared ? 0 : nt, std::vector<double>(C.hist_total, 0.0));
  int maxms = 1;
  for (auto& c : C.comps)
    if (c.type == MCV_SAMPLE_HMS) maxms = std::max(maxms, c.sample.max_scattered);
#pragma omp parallel num_threads(nt) reduction(+ : processed)
  {
    int me = omp_get_thread_num();
    g_shared_hist = shared;
    double* h = shared ? hist : hists[me].data();
    std::vector<double> msbuf(10 * maxms);
#pragma omp for schedule(dynamic, 64)
    for (int64_t tid = tid0; tid < tid1; ++tid) {
      Rng rng = states[tid];
      double n[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
      int64_t start = tid * npt;
      int64_t end = std::min<int64_t>(start + npt, (int64_t)N);
      for (int64_t idx = start; idx < end; ++idx) {
        processed += 1;
        bool alive = true;
#ifndef ORC_FLOP_BUILD
        t_pin = idx == g_pin_idx;
        t_call = 0;
#endif
        int num = run_neutron(C, rng, n, h, msbuf, alive, idx);
#ifndef ORC_FLOP_BUILD
        t_pin = false;
#endif
        if (dump_states && idx >= dump_first && idx < dump_first + dump_count) {
          int64_t k = idx - dump_first;
          std::memcpy(dump_states + 10 * k, n, 10 * sizeof(double));
          dump_alive[k] = C.ms_index >= 0 ? num : (alive ? 1 : 0);
        }
      }
    }
    g_shared_hist = false;
  }
  for (size_t t = 0; t < hists.size(); ++t)
    for (int64_t k = 0; k < C.hist_total; ++k) hist[k] += hists[t][k];
  return processed;
}

// Buffered mode (acc/components/ComponentBase.py:79-107 with the launch of
// StochasticComponentBase.call_process, :21-55): the chain runs in place over
// an (n, 10) array, stream tid taking rows [tid*npt, min((tid+1)*npt, n)) of
// the partition of n, streams seeded from `seed`.  A chain that is a single
// multiple-scattering sample writes neutron i's outputs to rows
// [i*M, i*M + num) of out_ms (n*M, 10) and leaves the rest zero.
int64_t orc_process(const mcv_comp_desc* comps, int ncomp, const double* offsets, const double* rotmats,
                    double* neutrons, int64_t n, uint64_t seed, int64_t ntotalthreads, int32_t tpb, uint32_t flags,
                    double* hist, double* out_ms, int nthreads_omp) {
  Chain C;
  build_chain(comps, ncomp, offsets, rotmats, flags, C);
  if (C.ms_index >= 0 && (ncomp != 1 || !out_ms)) return -1;
  int64_t part[3];
  orc_partition((uint64_t)n, ntotalthreads, tpb, part);
  int64_t nthreads = part[1], npt = part[2];
  std::vector<Rng> states(nthreads > 0 ? nthreads : 1);
  init_states(seed, nthreads, states.data());
  int nt = nthreads_omp > 0 ? nthreads_omp : omp_get_max_threads();
  std::vector<std::vector<double>> hists(nt, std::vector<double>(C.hist_total, 0.0));
  int64_t processed = 0;
#pragma omp parallel num_threads(nt) reduction(+ : processed)
  {
    double* h = hists[omp_get_thread_num()].data();
#pragma omp for schedule(dynamic, 64)
    for (int64_t tid = 0; tid < nthreads; ++tid) {
      Rng rng = states[tid];
      int64_t start = tid * npt;
      int64_t end = std::min<int64_t>(start + npt, n);
      for (int64_t idx = start; idx < end; ++idx) {
        processed += 1;
        double* nn = neutrons + 10 * idx;
        if (C.ms_index < 0) {
          run_tail(C, 0, rng, nn, h, idx);
        } else {
          const Sample& S = C.comps[0].sample;
          const int M = S.max_scattered;
          double* o = out_ms + 10 * (int64_t)M * idx;
          std::memset(o, 0, 10 * (size_t)M * sizeof(double));
          scatterM(rng, S, o, M, nn);
        }
      }
    }
    g_shared_hist = false;
  }
  for (size_t t = 0; t < hists.size(); ++t)
    for (int64_t k = 0; k < C.hist_total; ++k) hist[k] += hists[t][k];
  return processed;
}

}  // extern "C"
