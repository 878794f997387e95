// mcv_components.h -- per-component device code of the fused kernels.
//
// Each function is the gfx950 form of a reference component's `propagate`
// (cited per function).  Parameters arrive as a pointer into the packed
// parameter blob of the instrument (layouts in include/mcv_acc.h); all lanes
// read the same addresses, so they compile to scalar loads.  Arithmetic keeps
// the reference's operation order and the kernels are compiled without FMA
// contraction, so each operation rounds as the reference's does.
#pragma once
#include "mcv_core.h"

namespace mcv {

// ------------------------------------------------------------------ Source_simple
// acc/components/sources/source_simple.py:114-161 (+ randvec_target_rect :17-42).
// P[0..12] = reference params; P[13..15] = p1, P[16..18] = p2 (the two unit
// vectors randvec_target_rect derives from the target, computed on the host
// with the reference's formulas); P[19] = length(target).
__device__ __forceinline__ void source_simple(Rng& rng, Neutron& n, const double* __restrict__ P) {
  const double r1 = rng.uniform(), r2 = rng.uniform(), r3 = rng.uniform(), r4 = rng.uniform(), r5 = rng.uniform();
  double x, y;
  if (P[0] != 0.0) {
    x = P[1] * (r1 - 0.5);
    y = P[2] * (r2 - 0.5);
  } else {
    const double chi = (2 * PI) * r1;
    // numba types sqrt(float32) as float32: the correctly rounded float sqrt,
    // obtained exactly as float(sqrt_double) (double rounding is innocuous for
    // sqrt since 53 >= 2*24 + 2)
    const double r = (double)(float)sqrt(r2) * P[3];
    double s, c;
    sincos(chi, &s, &c);
    x = r * c;
    y = r * s;
  }
  const double xw = P[9], yh = P[10], dist = P[11];
  // randvec_target_rect(target=(0,0,dist), xw, yh, r3, r4)
  const double dx0 = xw * (r3 * 2 - 1) / 2.0;
  const double dy0 = yh * (r4 * 2 - 1) / 2.0;
  const double L = P[19];
  const double vf0 = P[16] * dy0 + (P[13] * dx0 + 0.0);
  const double vf1 = P[17] * dy0 + (P[14] * dx0 + 0.0);
  const double vf2 = P[18] * dy0 + (P[15] * dx0 + dist);
  const double dist2 = sqrt(dx0 * dx0 + dy0 * dy0 + L * L);
  const double solidangle = (xw * yh * L) / (dist2 * dist2 * dist2);
  (void)vf2;
  const double dx = vf0 - x, dy = vf1 - y;
  const double dist1 = sqrt(dx * dx + dy * dy + dist * dist);
  double v;
  if (P[4] != 0.0) {
    const double Lw = P[5] + P[6] * (r5 * 2 - 1);
    v = K2V * ((2 * PI) / Lw);
  } else {
    const double E = P[7] + P[8] * (r5 * 2 - 1);
    v = SE2V * sqrt(E);
  }
  n.x = x;
  n.y = y;
  n.z = 0.0;
  n.vx = v * dx / dist1;
  n.vy = v * dy / dist1;
  n.vz = v * dist / dist1;
  n.t = 0.0;
  n.p = P[12] * solidangle;
}

// ------------------------------------------------------------------ supermirror
// acc/components/optics/_guide_utils.py:6-23
__device__ __forceinline__ double calc_reflectivity(double Q, double R0, double Qc, double alpha, double m,
                                                    double W) {
  double R = R0;
  if (Q > Qc) {
    const double tmp = (Q - m * Qc) / W;
    if (tmp < 10)
      R *= (1 - tanh(tmp)) * (1 - alpha * (Q - Qc)) / 2;
    else
      R = 0;
  }
  return R;
}

// ------------------------------------------------------------------ exit-time selection
// The guides pick the earliest of several exit times t_k = n_k / d_k with a
// chain of strict "t2 < t1" tests over the walls the neutron moves towards
// (d_k < 0; guide.py:100-123, guide_tapering.py:204-227).  Here the chain
// compares quotients by cross-multiplication, without branches: a losing
// candidate costs two products and a few selects instead of an FP64 division
// (the costliest instruction sequence of the loop, which a wavefront would
// execute whenever any of its lanes takes the branch).  The winning wall's
// time is divided afterwards from the same operands as in the reference, so
// a path is bit for bit the reference's unless two times agree to the last
// bits (a corner hit).  A neutron not moving forward (vz <= 0 or NaN) with a
// wall ahead sets `odd`, and earliest_exit then runs the reference's chain
// with divisions.
struct ExitChoice {
  double n, d;  // the earliest exit time so far is n / d
  int i;        // 0: the segment (guide) end; 1-4: a wall
  bool odd;
};

// e.d < 0 (see earliest_exit) and d < 0, so d * e.d > 0 and n/d < e.n/e.d
// is n * e.d < e.n * d.
__device__ __forceinline__ void exit_candidate(ExitChoice& e, double n, double d, int k) {
  const double l = n * e.d, r = e.n * d;
  const bool take = (d < 0) && (l < r);
  e.n = take ? n : e.n;
  e.d = take ? d : e.d;
  e.i = take ? k : e.i;
}

// The reference's chain itself, for the rare `odd` case.
__device__ __forceinline__ void exit_divided(ExitChoice& e, double& t1, double n, double d, int k) {
  if (d < 0) {
    const double t2 = n / d;
    if (t2 < t1) {
      t1 = t2;
      e.n = n;
      e.d = d;
      e.i = k;
    }
  }
}

__device__ __forceinline__ ExitChoice earliest_exit(double n0, double d0, double n1, double d1, double n2,
                                                    double d2, double n3, double d3, double n4, double d4) {
  // the end's time n0 / d0 enters as (-n0) / (-d0), so that e.d < 0
  // throughout (negation is exact, and the products and their comparison only
  // change sign); a taken wall has d < 0.  A neutron not moving forward (d0
  // <= 0 or NaN) with a wall ahead takes the reference's divided chain.
  ExitChoice e{-n0, -d0, 0, false};
  exit_candidate(e, n1, d1, 1);
  exit_candidate(e, n2, d2, 2);
  exit_candidate(e, n3, d3, 3);
  exit_candidate(e, n4, d4, 4);
  e.odd = !(d0 > 0) && (d1 < 0 || d2 < 0 || d3 < 0 || d4 < 0);
  if (e.odd) {
    double t1 = n0 / d0;
    e.n = n0;
    e.d = d0;
    e.i = 0;
    exit_divided(e, t1, n1, d1, 1);
    exit_divided(e, t1, n2, d2, 2);
    exit_divided(e, t1, n3, d3, 3);
    exit_divided(e, t1, n4, d4, 4);
  }
  return e;
}

// ------------------------------------------------------------------ Guide
// acc/components/optics/guide.py:59-162.  P = ww, hh, hw1, hh1, l, R0, Qc, alpha, m, W,
// then the mirror normals' squared lengths and lengths the host derives from
// them (l*l + ww*ww, its sqrt, l*l + hh*hh, its sqrt: the reference's nlen2
// and sqrt(nlen2), correctly rounded on either side).
// The guide runs in two parts: the entrance (move to z = 0, test the opening)
// and the bounces.  A compacted kernel can end its per-stream phase at the
// entrance, where three in four neutrons of a beam wider than the opening
// die, and run the bounces with whole wavefronts of neutrons that entered.
// guide_enter returns false when it absorbs (p = -1, the rest of the state
// left as it was); otherwise the neutron is at z = 0.
__device__ __forceinline__ bool guide_enter(Neutron& n, const double* __restrict__ P) {
  const double hw1 = P[2], hh1 = P[3];
  const double dt = -n.z / n.vz;
  const double x = n.x + n.vx * dt, y = n.y + n.vy * dt;
  if (x <= -hw1 || x >= hw1 || y <= -hh1 || y >= hh1) {
    n.p = -1;  // absorb()
    return false;
  }
  n.x = x;
  n.y = y;
  n.z = 0.;
  n.t += dt;
  return true;
}

template <class T>
__device__ __forceinline__ void guide_bounces_t(Neutron& n, const double* __restrict__ P, T tr) {
  const double ww = P[0], hh = P[1], hw1 = P[2], hh1 = P[3], l = P[4];
  double x = n.x, y = n.y, z = n.z, vx = n.vx, vy = n.vy, vz = n.vz, t = n.t, prob = n.p;
  for (int nb = 0; nb < 100000; ++nb) {
    const double av = l * vx, bv = ww * vz, ah = l * vy, bh = hh * vz;
    const double vdotn_v1 = bv + av, vdotn_v2 = bv - av, vdotn_h1 = bh + ah, vdotn_h2 = bh - ah;
    const double cv1 = -hw1 * l - z * ww, cv2 = x * l, ch1 = -hh1 * l - z * hh, ch2 = y * l;
    const ExitChoice e = earliest_exit(l - z, vz, cv1 - cv2, vdotn_v1, cv1 + cv2, vdotn_v2, ch1 - ch2, vdotn_h1,
                                       ch1 + ch2, vdotn_h2);
    const int i = e.i;
    if (i == 0) break;  // the guide-end time itself is never used
    const double t1 = e.n / e.d;
    x += vx * t1;
    y += vy * t1;
    z += vz * t1;
    t += t1;
    // reflection off mirror i: the four cases share one form with (axis, sign)
    const bool vert = i <= 2;
    const double wh = vert ? ww : hh;
    const double vdn = e.d;
    const double nlen2 = vert ? P[10] : P[12];
    const double q = V2K * (-2) * vdn / (vert ? P[11] : P[13]);
    const double d = 2 * vdn / nlen2;
    const double dl = d * l;
    vx = i == 1 ? vx - dl : i == 2 ? vx + dl : vx;
    vy = i == 3 ? vy - dl : i == 4 ? vy + dl : vy;
    vz = vz - d * wh;
    const double R = calc_reflectivity(q, P[5], P[6], P[7], P[8], P[9]);
    if (R > 0) tr.mirror(vert ? 0 : 1, q, P[5], P[6], P[7], P[8], P[9], 1.0);
    prob *= R;
    if (prob <= 0) break;
  }
  n.x = x; n.y = y; n.z = z; n.vx = vx; n.vy = vy; n.vz = vz;
  n.t = t;
  n.p = prob;
}

template <class T>
__device__ __forceinline__ void guide_t(Neutron& n, const double* __restrict__ P, T tr) {
  if (guide_enter(n, P)) guide_bounces_t(n, P, tr);
}

__device__ __forceinline__ void guide(Neutron& n, const double* __restrict__ P) { guide_t(n, P, NoTrack{}); }
__device__ __forceinline__ void guide_bounces(Neutron& n, const double* __restrict__ P) {
  guide_bounces_t(n, P, NoTrack{});
}

// ------------------------------------------------------------------ Slit / Beamstop
// acc/components/optics/slit.py:71-108, beamstop.py:75-112.
// P = xmin, xmax, ymin, ymax, radius^2, cut.  A neutron that never reaches
// z = 0 is left as it is; below the weight cut or stopped it is absorbed
// (p = -1, the rest of the state unchanged); otherwise it moves to z = 0.
template <bool BEAMSTOP>
__device__ __forceinline__ void aperture(Neutron& n, const double* __restrict__ P) {
  const double z = n.z, vx = n.vx, vy = n.vy, vz = n.vz;
  if ((z < 0 && vz <= 0) || (z > 0 && vz >= 0)) return;
  double x = n.x, y = n.y, t = n.t;
  if (n.p < P[5]) {
    n.p = -1;
    return;
  }
  if (z != 0) {
    const double dt = -z / vz;
    x += vx * dt;
    y += vy * dt;
    t += dt;
  }
  bool stop;
  if (P[4] != 0) {
    stop = BEAMSTOP ? (x * x + y * y <= P[4]) : (x * x + y * y > P[4]);
  } else {
    const bool within = P[0] <= x && x <= P[1] && P[2] <= y && y <= P[3];
    stop = BEAMSTOP ? within : !within;
  }
  if (stop) {
    n.p = -1;
    return;
  }
  n.x = x;
  n.y = y;
  n.z = 0.;
  n.t = t;
}

__device__ __forceinline__ void slit(Neutron& n, const double* __restrict__ P) { aperture<false>(n, P); }
__device__ __forceinline__ void beamstop(Neutron& n, const double* __restrict__ P) { aperture<true>(n, P); }

// ------------------------------------------------------------------ Guide_tapering
// acc/components/optics/guide_tapering.py:141-291.
// P = l_seg, R0, Qcx, Qcy, alphax, alphay, mx, my, W, nseg (as double);
// D = one 16-double record per segment, packed by the runtime from the
// component's 10 x nseg table (w1, w2, h1, h2, ww, hh, whalf, hhalf, lwhalf,
// lhhalf): [0] w1, [1] hhalf, [2] ww, [3] hh, [4] whalf * l_seg, [5] hhalf *
// l_seg (the reference's products in cv1 and ch1), [6] zr = seg * l_seg, [7]
// zr + l_seg, [8] l_seg^2 + ww^2 and [9] its sqrt, [10] l_seg^2 + hh^2 and
// [11] its sqrt (the reference's nlen2 and sqrt(nlen2)), [12] h1, [13] whalf,
// [14] 1 / [8] and [15] 1 / [10] (correctly rounded, for div_recip).
// A segment entrance reads its eight values from one 128-byte record instead
// of rows nseg apart.
// The reference's two nested loops (segments, then bounces inside a segment)
// run here as one event loop: each pass computes the next wall or segment-end
// time of the current segment, then either reflects or moves on to the next
// segment.  Every neutron still performs exactly the reference's operations in
// the reference's order, but lanes of a wavefront no longer wait for each
// other at every segment: a wave runs max(segments + bounces) passes instead
// of the sum over segments of the per-segment maximum (init / step / finish
// below).
// Reflectivity is the costliest part of a bounce (a division for q, another in
// R and a tanh), and a wavefront would evaluate it on every pass in which any
// of its lanes bounces.  A bounce therefore only records its wall's v.n, its
// segment and its axis as pending.  A lane's
// pending reflectivity falls due before its next reflection or its exit from
// the guide; whenever one falls due, every pending lane of the wavefront
// evaluates its own in the same pass.  Each neutron's weight is still
// multiplied by its R factors in the reference's order.  A neutron that a
// pending R kills (R <= 1e-10) may have moved on meanwhile; it ends dead all
// the same, which is all the reference's early return leaves behind (p = 0,
// state as it entered).  (Also evaluating when half the wavefront held a
// pending R measured slower on the 100-segment guide: 26.3 vs 25.8 ms per
// 1e8 neutrons.)

struct TaperState {
  double x, y, z, vx, vy, vz, t, prob;
  double zr, zend, wws, hhh, lwh, lhh, hadj;  // current segment (entrance and end z, slopes, l_seg x half sizes, offset)
  double pvdn;                          // pending reflection: v.n of its wall
  int seg, nb;
  int pseg, pax;  // pending reflection's segment and axis (0 vertical, 1 horizontal; -1 none)
  bool enter, died;
};

__device__ __forceinline__ void taper_init(TaperState& s, const Neutron& n, const double* __restrict__ P) {
  const int nseg = (int)P[9];
  s.x = n.x; s.y = n.y; s.z = n.z; s.vx = n.vx; s.vy = n.vy; s.vz = n.vz; s.t = n.t; s.prob = n.p;
  const double dt = -s.z / s.vz;
  s.x += s.vx * dt;
  s.y += s.vy * dt;
  s.z = 0.0;
  s.t += dt;
  s.seg = 0;
  s.nb = 0;
  s.zr = s.zend = s.wws = s.hhh = s.lwh = s.lhh = s.hadj = 0;
  s.pvdn = 0;
  s.pseg = 0;
  s.pax = -1;
  s.enter = nseg > 0;
  s.died = false;
}

// Applies the pending reflection (guide_tapering.py:236-284); returns true
// when it kills the neutron.
template <class T>
__device__ __forceinline__ bool taper_reflect(TaperState& s, const double* __restrict__ P,
                                              const double* __restrict__ D, T& tr) {
  const bool vert = s.pax == 0;
  const double* __restrict__ rec = D + 16 * s.pseg;
  const double q = V2K * (-2.0) * s.pvdn / rec[vert ? 9 : 11];
  const double mxy = vert ? P[6] : P[7];
  double m = mxy, dm = 1.0;
  if (m <= 0.0) {  // m = |mxy * w1[0] / w1[seg]| (or h1)
    const double a0 = vert ? D[0] : D[12], as = vert ? rec[0] : rec[12];
    const double ratio = a0 / as;
    m = fabs(mxy * a0 / as);
    dm = mxy * ratio < 0 ? -ratio : ratio;  // d|mxy * ratio| / d mxy
  }
  const double Qc = vert ? P[2] : P[3], alpha = vert ? P[4] : P[5];
  const double R = calc_reflectivity(q, P[1], Qc, alpha, m, P[8]);
  s.pax = -1;
  if (R > 1e-10) {
    tr.mirror(vert ? 0 : 1, q, P[1], Qc, alpha, m, P[8], dm);
    s.prob *= R;
    return false;
  }
  s.died = true;
  return true;
}

// A segment entrance (guide_tapering.py:170-186): moves (x, y, z) to the
// segment's entrance plane (t is left alone, as the reference leaves it) and
// returns whether the neutron dies there: outside the opening (p = 0), or
// beyond the channel after the shift.  x1 and edge are the shift's values.
// Both tests are evaluated without short-circuit branches.  After the shift x
// is in [0, w1s] (rounding keeps x > -w1s/2 from going negative), so floor(x /
// w1s) is 0 unless the quotient rounds up to 1: divide only then.
__device__ __forceinline__ bool taper_entrance(const double* __restrict__ rec, double& x, double& y, double& z,
                                               double vx, double vy, double vz, double& x1, double& edge) {
  const double w1s = rec[0], hhs = rec[1];
  const double dt = (rec[6] - z) / vz;
  x += vx * dt;
  y += vy * dt;
  z += vz * dt;
  const bool outside = (x <= -0.5 * w1s) | (x >= 0.5 * w1s) | (y <= -hhs) | (y >= hhs);
  x1 = x + 0.5 * w1s;
  edge = x1 < 0.75 * w1s ? 0.0 : floor(x1 / w1s) * w1s;
  return outside | (x1 - edge > w1s);
}

// One pass; returns true when the neutron has left the guide or died.
template <class T>
__device__ __forceinline__ bool taper_step(TaperState& s, const double* __restrict__ P,
                                           const double* __restrict__ D, T& tr) {
  const double l_seg = P[0];
  const int nseg = (int)P[9];
  if (s.enter) {  // segment entrance
    const double* __restrict__ rec = D + 16 * s.seg;
    const double w1s = rec[0];
    s.wws = rec[2];
    s.hhh = rec[3];
    s.lwh = rec[4];
    s.lhh = rec[5];
    s.zr = rec[6];
    s.zend = rec[7];
    double x1, edge;
    if (taper_entrance(rec, s.x, s.y, s.z, s.vx, s.vy, s.vz, x1, edge)) {
      s.died = true;
      return true;
    }
    s.x = x1 - (edge + 0.5 * w1s);
    s.hadj = edge + (0.5 * w1s) - 0.5 * w1s;
    s.nb = 0;
    s.enter = false;
  } else if (s.seg >= nseg) {
    return true;
  }
  // one pass of the bounce loop (guide_tapering.py:189-284); nb counts the
  // reference's max_bounces passes per segment
  bool seg_end = s.nb >= 100000;
  const double zr = s.zr, wws = s.wws, hhh = s.hhh, z = s.z, x = s.x, y = s.y, vx = s.vx, vy = s.vy, vz = s.vz;
  const double av = l_seg * vx, bv = wws * vz, ah = l_seg * vy, bh = hhh * vz;
  const double vdotn_v1 = bv + av, vdotn_v2 = bv - av, vdotn_h1 = bh + ah, vdotn_h2 = bh - ah;
  const double cv1 = -s.lwh - (z - zr) * wws, cv2 = x * l_seg;
  const double ch1 = -s.lhh - (z - zr) * hhh, ch2 = y * l_seg;
  ExitChoice e{0.0, 0.0, 0, false};
  if (!seg_end)
    e = earliest_exit(s.zend - z, vz, cv1 - cv2, vdotn_v1, cv1 + cv2, vdotn_v2, ch1 - ch2, vdotn_h1, ch1 + ch2,
                      vdotn_h2);
  const int i = e.i;
  // pending reflectivities: due before this lane reflects again or leaves the guide
  {
    const bool pend = s.pax >= 0;
    const bool due = pend && (i != 0 || s.seg + 1 >= nseg);
    if (__ballot(due) != 0ull && pend && taper_reflect(s, P, D, tr)) return true;
  }
  if (!seg_end) {
    if (i == 0) {
      seg_end = true;  // the neutron leaves the segment (the segment-end time is never used)
    } else {
      const double t1 = e.n / e.d;
      s.x = x + vx * t1;
      s.y = y + vy * t1;
      s.z = z + vz * t1;
      s.t += t1;
      const bool vert = i <= 2;
      const double wh = vert ? wws : hhh;
      const double vdn = e.d;
      const double* __restrict__ rec = D + 16 * s.seg;
      const double d = div_recip(2.0 * vdn, rec[vert ? 8 : 10], rec[vert ? 14 : 15]);
      const double dl = d * l_seg;
      s.vx = i == 1 ? vx - dl : i == 2 ? vx + dl : vx;
      s.vy = i == 3 ? vy - dl : i == 4 ? vy + dl : vy;
      s.vz = vz - d * wh;
      const double mxy = vert ? P[6] : P[7];
      if (mxy == 0) {
        s.x += s.hadj;  // the reference adds hadj here and again at the segment end
        seg_end = true;
      } else {
        s.pvdn = vdn;  // q, m and R are evaluated by taper_reflect
        s.pseg = s.seg;
        s.pax = vert ? 0 : 1;
        ++s.nb;
      }
    }
  }
  if (seg_end) {
    s.x += s.hadj;
    ++s.seg;
    s.enter = s.seg < nseg;
    if (!s.enter) return true;
  }
  return false;
}

// Whether the guide lets the neutron through its first segment's opening:
// false exactly when guide_tapering would leave it as it came with p = 0 at
// that opening (the same arithmetic on a copy).  Unchanged neutrons make this
// a test a kernel can run before deciding where the guide runs.
__device__ __forceinline__ bool taper_admits(const Neutron& n, const double* __restrict__ P,
                                             const double* __restrict__ D) {
  if ((int)P[9] <= 0) return true;
  TaperState s;
  taper_init(s, n, P);
  double x1, edge;
  return !taper_entrance(D, s.x, s.y, s.z, s.vx, s.vy, s.vz, x1, edge);
}

// A dead neutron keeps the state it entered with (p = 0); a transmitted one
// takes the guide's final state (guide_tapering.py:285-289).
__device__ __forceinline__ void taper_finish(const TaperState& s, Neutron& n) {
  if (s.died) {
    n.p = 0;
    return;
  }
  n.x = s.x; n.y = s.y; n.z = s.z; n.vx = s.vx; n.vy = s.vy; n.vz = s.vz;
  n.t = s.t;
  n.p = s.prob;
}

template <class T>
__device__ __forceinline__ void guide_tapering_t(Neutron& n, const double* __restrict__ P,
                                                 const double* __restrict__ D, T tr) {
  TaperState s;
  taper_init(s, n, P);
  while (!taper_step(s, P, D, tr)) {
  }
  taper_finish(s, n);
}

__device__ __forceinline__ void guide_tapering(Neutron& n, const double* __restrict__ P,
                                               const double* __restrict__ D) {
  guide_tapering_t(n, P, D, NoTrack{});
}

// ------------------------------------------------------------------ monitors
// Each monitor is its binning (`*_bin`: the bin index, or -1 when the neutron
// is not counted, and the number of bins); the fused kernel adds (1, p, p^2)
// to the bin (mcv::hist_count) and the adjoint kernel weighs the same bin.
// acc/neutron.py:60-66 prop_z0 (monitors do not write it back)
__device__ __forceinline__ void prop_z0(const Neutron& n, double& x, double& y, double& t) {
  const double dt = -n.z / n.vz;
  x = n.x + n.vx * dt;
  y = n.y + n.vy * dt;
  t = n.t + dt;
}

// acc/components/monitors/psd_monitor.py:66-83.  P = xmin, xmax, ymin, ymax, nx, ny
__device__ __forceinline__ int64_t psd_bin(const Neutron& n, const double* __restrict__ P, int64_t& nbins) {
  const int64_t nx = (int64_t)P[4], ny = (int64_t)P[5];
  nbins = nx * ny;
  double x, y, t;
  prop_z0(n, x, y, t);
  if (n.t > t) return -1;
  const double xmin = P[0], xmax = P[1], ymin = P[2], ymax = P[3];
  // strict window as the reference; written so that NaN is rejected too
  if (!(x > xmin && x < xmax && y > ymin && y < ymax)) return -1;
  int64_t ix = (int64_t)floor((x - xmin) / (xmax - xmin) * (double)nx);
  int64_t iy = (int64_t)floor((y - ymin) / (ymax - ymin) * (double)ny);
  ix = ix < 0 ? 0 : (ix < nx ? ix : nx - 1);
  iy = iy < 0 ? 0 : (iy < ny ? iy : ny - 1);
  return ix * ny + iy;
}

// acc/components/monitors/wavelength_monitor.py:52-74.  P = xmin..ymax, Lmin, Lmax, nchan
__device__ __forceinline__ int64_t wavelength_bin(const Neutron& n, const double* __restrict__ P,
                                                  int64_t& nbins) {
  const int64_t nchan = (int64_t)P[6];
  nbins = nchan;
  double x, y, t;
  prop_z0(n, x, y, t);
  if (n.t > t) return -1;
  if (!(x > P[0] && x < P[1] && y > P[2] && y < P[3])) return -1;
  const double v = sqrt(n.vx * n.vx + n.vy * n.vy + n.vz * n.vz);
  const double L = 2 * PI / (v * V2K);
  const double Lmin = P[4], Lmax = P[5];
  if (!(L > Lmin && L < Lmax)) return -1;
  int64_t iL = (int64_t)floor((L - Lmin) / (Lmax - Lmin) * (double)nchan);
  iL = iL < 0 ? 0 : (iL < nchan ? iL : nchan - 1);
  return iL;
}

// acc/components/monitors/posdiv_monitor.py:67-89.  P = xmin, xmax, ymin, ymax, maxdiv, npos, ndiv
__device__ __forceinline__ int64_t posdiv_bin(const Neutron& n, const double* __restrict__ P, int64_t& nbins) {
  const int64_t npos = (int64_t)P[5], ndiv = (int64_t)P[6];
  nbins = npos * ndiv;
  double x, y, t;
  prop_z0(n, x, y, t);
  if (n.t > t) return -1;
  const double xmin = P[0], xmax = P[1], ymin = P[2], ymax = P[3], maxdiv = P[4];
  if (!(x > xmin && x < xmax && y > ymin && y < ymax)) return -1;
  const double div = atan(n.vx / n.vz) * (180. / PI);
  if (!(div < maxdiv && div > -maxdiv)) return -1;
  int64_t ix = (int64_t)floor((x - xmin) / (xmax - xmin) * (double)npos);
  int64_t idiv = (int64_t)floor((div + maxdiv) / (2 * maxdiv) * (double)ndiv);
  ix = ix < 0 ? 0 : (ix < npos ? ix : npos - 1);
  idiv = idiv < 0 ? 0 : (idiv < ndiv ? idiv : ndiv - 1);
  return ix * ndiv + idiv;
}

// acc/components/monitors/psd_monitor_4pi.py:56-79.  P = phimin, phimax,
// thetamin, thetamax, radius, nphi, ntheta.  Moves the neutron onto the sphere
// (the reference's monitor is destructive) and bins (phi, theta) with the
// reference's index formulas (theta's rounds half up).  A neutron whose line
// misses the sphere is not counted and left as it is.
__device__ __forceinline__ int64_t psd4pi_bin(Neutron& n, const double* __restrict__ P, int64_t& nbins) {
  const int64_t nphi = (int64_t)P[5], ntheta = (int64_t)P[6];
  nbins = nphi * ntheta;
  const double R = P[4];
  const double x = n.x, y = n.y, z = n.z, vx = n.vx, vy = n.vy, vz = n.vz;
  // acc/geometry/arrow_intersect.py:155-168
  const double r0dotv = x * vx + y * vy + z * vz;
  const double cx = y * vz - z * vy, cy = z * vx - x * vz, cz = x * vy - y * vx;
  const double v2 = vx * vx + vy * vy + vz * vz;
  const double r0Xv2 = cx * cx + cy * cy + cz * cz;
  const double b2m4ac = v2 * R * R - r0Xv2;
  if (!(b2m4ac >= 0)) return -1;
  const double sq = sqrt(b2m4ac);
  double t1 = -(r0dotv + sq) / v2;
  const double t2 = -(r0dotv - sq) / v2;
  if (t2 < 0) return -1;
  if (t1 < 0) t1 = t2;
  prop_dt_inplace(n, t1);
  const double phi = atan2(n.x, n.z);
  const double theta = asin(n.y / R);
  int64_t iphi = (int64_t)((double)nphi * (phi / 2 / PI + 0.5));
  int64_t itheta = (int64_t)((double)ntheta * (theta + PI / 2) / PI + 0.5);
  iphi = iphi < 0 ? 0 : (iphi < nphi ? iphi : nphi - 1);
  itheta = itheta < 0 ? 0 : (itheta < ntheta ? itheta : ntheta - 1);
  return iphi * ntheta + itheta;
}

// acc/components/monitors/iqe_monitor.py:80-113.
// P = Ei, Qmin, Qmax, Emin, Emax, max_in, min_in, max_out, min_out, nQ, nE,
// then the angular window in the form iqe_window reads (derived by the
// runtime): [11], [12] the pseudo-angles of min_in and max_in (-3 / 3 for
// bounds no direction reaches), [13] kind of the min_out test (0 never, 1
// always, 2 compare), [14], [15] cos and sin of min_out, [16]..[18] the same
// for max_out.

// The reference's window test: in-plane angle atan2(vx, vz) and out-of-plane
// angle 90 -+ atan(sqrt(vx^2 + vz^2) / vy) in degrees, strict bounds.  Kept
// out of line: iqe_window calls it only for directions within 1e-12 of a
// window edge, and its atan2 / atan would otherwise raise the register
// pressure of every fused kernel holding an I(Q,E) monitor (marked cold, so
// the register allocator keeps the saves around the call on that rare path:
// c2s 24.6 ms with the angles inline, 24.2 as a plain call, 23.8 cold).
__device__ __noinline__ __attribute__((cold)) bool iqe_window_exact(double vx, double vy, double vz, const double* __restrict__ P) {
  const double angle_in_plane = atan2(vx, vz) / PI * 180.;
  double angle_out_of_plane = 0.;
  if (vy != 0) {
    const double theta = atan(sqrt(vx * vx + vz * vz) / vy) / PI * 180.;
    angle_out_of_plane = vy > 0 ? 90. - theta : 90. + theta;
  }
  return P[6] < angle_in_plane && angle_in_plane < P[5] && P[8] < angle_out_of_plane && angle_out_of_plane < P[7];
}

// Pseudo-angle of the in-plane direction (vz, vx): vx / (|vx| + |vz|) for vz >
// 0, 2 - that for vx > 0 >= vz, -2 - that for vx < 0 >= vz; strictly
// increasing in atan2(vx, vz) over (-pi, pi), from -2 to 2, with slope at least
// 1/2 per radian.
__device__ __forceinline__ double iqe_pseudo_angle(double vx, double vz, double s) {
  const double f = vx / s;
  return vz > 0 ? f : (vx > 0 ? 2.0 - f : -2.0 - f);
}

// The same decision without atan2 / atan.  The in-plane bounds compare
// pseudo-angles; the out-of-plane angle is the elevation el = atan(|vy| / r),
// r = sqrt(vx^2 + vz^2), in [0, 90] degrees (the reference's 90 -+ theta for
// either sign of vy), and el > b is the sign of |vy| cos b - r sin b.  Every
// quantity is within ~1e-15 (relative to its scale) of its exact value, and
// the reference's angles within ~1e-15 rad of the true ones, so a direction
// more than 1e-12 from every bound gets the reference's answer from the signs;
// one closer (or with vx = 0, non-finite components or bounds) takes
// iqe_window_exact.
__device__ __forceinline__ bool iqe_window(double vx, double vy, double vz, const double* __restrict__ P) {
  constexpr double kMargin = 1e-12;
  const double s = fabs(vx) + fabs(vz), ay = fabs(vy);
  const double p = iqe_pseudo_angle(vx, vz, s);
  const double r = sqrt(vx * vx + vz * vz), w = ay + r;
  const double qlo = ay * P[14] - r * P[15], qhi = ay * P[17] - r * P[18];
  const bool tlo = P[13] == 2.0, thi = P[16] == 2.0;
  const bool sure = vx != 0 && s < INFINITY && ay < INFINITY && fabs(p - P[11]) > kMargin &&
                    fabs(p - P[12]) > kMargin && (!tlo || fabs(qlo) > kMargin * w) &&
                    (!thi || fabs(qhi) > kMargin * w);
  if (__builtin_expect(!sure, 0)) return iqe_window_exact(vx, vy, vz, P);
  const bool above = tlo ? qlo > 0 : P[13] == 1.0;  // el > min_out
  const bool below = thi ? qhi < 0 : P[16] == 1.0;  // el < max_out
  return p > P[11] && p < P[12] && above && below;
}

__device__ __forceinline__ int64_t iqe_bin(const Neutron& n, const double* __restrict__ P, int64_t& nbins) {
  const int64_t nQ = (int64_t)P[9], nE = (int64_t)P[10];
  nbins = nQ * nE;
  const double vx = n.vx, vy = n.vy, vz = n.vz;
  if (!iqe_window(vx, vy, vz, P)) return -1;
  const double Ei = P[0], Qmin = P[1], Qmax = P[2], Emin = P[3], Emax = P[4];
  const double vi = sqrt(Ei) * SE2V;
  const double Q = sqrt(vx * vx + vy * vy + (vz - vi) * (vz - vi)) * V2K;
  const double E = Ei - (vx * vx + vy * vy + vz * vz) * VS2E;
  if (!((Q >= Qmin && Q < Qmax) && (E >= Emin && E < Emax))) return -1;
  int64_t iQ = (int64_t)floor((Q - Qmin) / (Qmax - Qmin) * (double)nQ);
  int64_t iE = (int64_t)floor((E - Emin) / (Emax - Emin) * (double)nE);
  iQ = iQ < 0 ? 0 : (iQ < nQ ? iQ : nQ - 1);
  iE = iE < 0 ? 0 : (iE < nE ? iE : nE - 1);
  return iQ * nE + iE;
}

// Cylindrical array of vertical tubes (no reference counterpart; monitor
// conventions of psd_monitor.py).  P = radius, height, phi_min, phi_max (deg),
// tof_min, tof_max (s), tube_diameter, unused, ntubes, npixels, ntof.
__device__ __forceinline__ int64_t tube_bin(const Neutron& n, const double* __restrict__ P, int64_t& nbins) {
  const int64_t ntubes = (int64_t)P[8], npix = (int64_t)P[9], ntof = (int64_t)P[10];
  nbins = ntubes * npix * ntof;
  const double R = P[0], Ht = P[1], phimin = P[2], phimax = P[3], tmin = P[4], tmax = P[5];
  const double x = n.x, y = n.y, z = n.z, vx = n.vx, vy = n.vy, vz = n.vz;
  const double a = vx * vx + vz * vz;
  if (a == 0) return -1;
  const double b = x * vx + z * vz, c = x * x + z * z - R * R;
  const double k = b * b - a * c;
  if (k < 0) return -1;
  const double dt = (-b + sqrt(k)) / a;
  if (dt < 0) return -1;
  const double xh = x + vx * dt, yh = y + vy * dt, zh = z + vz * dt, tof = n.t + dt;
  const double phi = atan2(xh, zh) / PI * 180.;
  if (!(phi >= phimin && phi < phimax)) return -1;
  if (!(yh > -Ht / 2 && yh < Ht / 2)) return -1;
  if (!(tof >= tmin && tof < tmax)) return -1;
  int64_t it = (int64_t)floor((phi - phimin) / (phimax - phimin) * (double)ntubes);
  int64_t ip = (int64_t)floor((yh + Ht / 2) / Ht * (double)npix);
  int64_t ib = (int64_t)floor((tof - tmin) / (tmax - tmin) * (double)ntof);
  it = it < 0 ? 0 : (it < ntubes ? it : ntubes - 1);
  ip = ip < 0 ? 0 : (ip < npix ? ip : npix - 1);
  ib = ib < 0 ? 0 : (ib < ntof ? ib : ntof - 1);
  return (it * npix + ip) * ntof + ib;
}

}  // namespace mcv

namespace mcv {

// ------------------------------------------------------------------ stored neutrons
// acc/components/sources/neutronfromstorage.py:70-82: neutron i of the run is
// row i % N of the stored (N, 10) array.  P[0] = N, D = the rows.
__device__ __forceinline__ void storage_source(Neutron& n, const double* __restrict__ P,
                                               const double* __restrict__ D, int64_t nidx) {
  const int64_t N = (int64_t)P[0];
  const double* r = D + 10 * (nidx % N);
  n.x = r[0]; n.y = r[1]; n.z = r[2]; n.vx = r[3]; n.vy = r[4]; n.vz = r[5];
  n.s1 = r[6]; n.s2 = r[7]; n.t = r[8]; n.p = r[9];
}

// ------------------------------------------------------------------ IQE normalizer source
// Neutrons of the I(Q,E) monitor's Monte-Carlo normalizer
// (acc/components/monitors/iqe_monitor.py:146-174, makeNeutrons): energy
// transfer E uniform in [Emin, Emax), final speed e2v(Ei - E), direction
// uniform on 4 pi, weight vf/vi, at the origin.  P = Ei, Emin, Emax.
__device__ __forceinline__ void iqe_normalizer_source(Rng& rng, Neutron& n, const double* __restrict__ P) {
  const double Ei = P[0], Emin = P[1], Emax = P[2];
  const double E = Emin + rng.uniform() * (Emax - Emin);
  const double Ef = Ei - E;
  const double vi = sqrt(Ei) * SE2V, vf = sqrt(Ef) * SE2V;
  const double cos_t = rng.uniform() * 2 - 1;
  const double sin_t = sin(acos(cos_t));
  const double phi = rng.uniform() * 2 * PI;
  double sp, cp;
  sincos(phi, &sp, &cp);
  n.x = 0; n.y = 0; n.z = 0;
  n.vx = vf * sin_t * cp;
  n.vy = vf * sin_t * sp;
  n.vz = vf * cos_t;
  n.t = 0;
  n.p = vf / vi;
}

}  // namespace mcv
