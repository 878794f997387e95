// mcv_samples.h -- CSG geometry, scattering kernels and the single-scatterer
// sample on the device.
//
// The reference renders a shape tree into nested Numba closures that fill a
// local ts[10] array (acc/geometry/arrow_intersect.py:18-142).  Here the tree is
// a C++ type (CsgSphere<0>, CsgDifference<0, CsgCylinder<1>, CsgCylinder<2>>,
// ...) emitted by the fuser, so every root list has a compile-time bound K and
// stays in registers: the sorted insertion of intersectComposite becomes a
// branch-free merge of the children's on-border roots, and index shifts become
// unrolled selects.  Node parameters are read from the sample's data block
// (MCV_CSG_NODE_SIZE doubles per node).
#pragma once
#include "mcv_core.h"

namespace mcv {

constexpr double CSG_EPS = 1e-10;  // acc/geometry/__init__.py:2
constexpr int OUTSIDE = 1, ONBORDER = 0, INSIDE = -1;
constexpr int MAXI = 10;  // acc/geometry/arrow_intersect.py:16

constexpr int cmin(int a, int b) { return a < b ? a : b; }

// acc/geometry/locate.py:85-122
__device__ __forceinline__ int locate_sphere(double x, double y, double z, double R) {
  const double d2 = x * x + y * y + z * z;
  if (d2 > (R + CSG_EPS) * (R + CSG_EPS)) return OUTSIDE;
  if (d2 < (R - CSG_EPS) * (R - CSG_EPS)) return INSIDE;
  return ONBORDER;
}
__device__ __forceinline__ int locate_cylinder(double x, double y, double z, double R, double H) {
  if (fabs(z) - H / 2. > CSG_EPS) return OUTSIDE;
  const double r2 = x * x + y * y;
  if (r2 > (R + CSG_EPS) * (R + CSG_EPS)) return OUTSIDE;
  if (H / 2. - fabs(z) > CSG_EPS && r2 < (R - CSG_EPS) * (R - CSG_EPS)) return INSIDE;
  return ONBORDER;
}
__device__ __forceinline__ int locate_box(double x, double y, double z, double W, double H, double D) {
  if (fabs(x) - 0.5 * W > CSG_EPS) return OUTSIDE;
  if (fabs(y) - 0.5 * H > CSG_EPS) return OUTSIDE;
  if (fabs(z) - 0.5 * D > CSG_EPS) return OUTSIDE;
  if ((-0.5 * W < x && x < 0.5 * W) && (-0.5 * H < y && y < 0.5 * H) && (-0.5 * D < z && z < 0.5 * D))
    return INSIDE;
  return ONBORDER;
}

// acc/geometry/arrow_intersect.py:154-168
__device__ __forceinline__ int intersect_sphere(double x, double y, double z, double vx, double vy, double vz,
                                                double R, double* ts) {
  const double r0dotv = x * vx + y * vy + z * vz;
  const double cx = y * vz - z * vy, cy = z * vx - x * vz, cz = x * vy - y * vx;
  const double v2 = vx * vx + vy * vy + vz * vz;
  const double r0Xv2 = cx * cx + cy * cy + cz * cz;
  const double b2m4ac = v2 * R * R - r0Xv2;
  if (!(b2m4ac >= 0)) return 0;
  const double sq = sqrt(b2m4ac);
  ts[0] = -(r0dotv + sq) / v2;
  ts[1] = -(r0dotv - sq) / v2;
  return 2;
}

// acc/geometry/arrow_intersect.py:178-267 (side + top/bottom, combined as the
// reference combines them; n == 0 -> no roots)
__device__ __forceinline__ int intersect_cylinder(double x, double y, double z, double vx, double vy, double vz,
                                                  double R, double H, double* ts) {
  // side
  int n1 = 0;
  double t1 = 0, t2 = 0;
  {
    const double a = vx * vx + vy * vy, b = x * vx + y * vy, c = x * x + y * y - R * R;
    double k = b * b - a * c;
    const double hh = H / 2.;
    bool done = false;
    if (k < 0 || a == 0) done = true;
    if (!done && k == 0) {
      const double t = -b / a;
      if (fabs(z + vz * t) < hh) {
        n1 = 1;
        t1 = t;
        done = true;
      }
    }
    if (!done) {
      k = sqrt(k);
      const double ta = (-b + k) / a, tb = (-b - k) / a;
      const bool g1 = fabs(z + vz * ta) < hh, g2 = fabs(z + vz * tb) < hh;
      if (g1 && g2) { n1 = 2; t1 = tb; t2 = ta; }
      else if (g1) { n1 = 1; t1 = ta; }
      else if (g2) { n1 = 1; t1 = tb; }
    }
  }
  // top / bottom
  int n2 = 0;
  double t3 = 0, t4 = 0;
  if (vz != 0) {
    const double hh = H / 2, r2 = R * R;
    const double ta = (hh - z) / vz;
    double x1 = x + vx * ta, y1 = y + vy * ta;
    const bool g1 = x1 * x1 + y1 * y1 <= r2;
    const double tb = (-hh - z) / vz;
    x1 = x + vx * tb;
    y1 = y + vy * tb;
    const bool g2 = x1 * x1 + y1 * y1 <= r2;
    if (g1 && g2) {
      n2 = 2;
      if (ta < tb) { t3 = ta; t4 = tb; }
      else { t3 = tb; t4 = ta; }
    } else if (g1) { n2 = 1; t3 = ta; }
    else if (g2) { n2 = 1; t3 = tb; }
  }
  const int n = n1 + n2;
  if (n == 0) return 0;
  double o1, o2;
  if (n == 1) {
    o1 = o2 = (n1 == 1) ? t1 : t3;
  } else if (n == 2) {
    if (n1 == 1) {
      if (t1 > t3) { o1 = t3; o2 = t1; }
      else { o1 = t1; o2 = t3; }
    } else if (n1 == 0) { o1 = t3; o2 = t4; }
    else { o1 = t1; o2 = t2; }
  } else if (n == 3) {
    if (n1 == 1) { o1 = t3 < t1 ? t3 : t1; o2 = t4 > t1 ? t4 : t1; }
    else { o1 = t1 < t3 ? t1 : t3; o2 = t2 > t3 ? t2 : t3; }
  } else {
    o1 = t3 < t1 ? t3 : t1;
    o2 = t4 > t2 ? t4 : t2;
  }
  ts[0] = o1;
  ts[1] = o2;
  return 2;
}

// acc/geometry/onbox.py:8-56, evaluated in float64 (the reference uses float32)
__device__ __forceinline__ double intersect_rectangle(double rx, double ry, double rz, double vx, double vy, double vz,
                                                      double X, double Y) {
  double t = -rz / vz;
  const double r1x = rx + vx * t, r1y = ry + vy * t;
  if (fabs(r1x) > X / 2 || fabs(r1y) > Y / 2) t = __builtin_nan("");
  return t;
}
__device__ __forceinline__ void update_intersections(double& t1, double& t2, double t) {
  if (!isfinite(t)) return;
  if (!isfinite(t1)) { t1 = t; return; }
  if (!isfinite(t2)) {
    if (t > t1) t2 = t;
    else { t2 = t1; t1 = t; }
    return;
  }
  if (t < t1) { t1 = t; return; }
  if (t < t2) return;
  t2 = t;
}
__device__ __forceinline__ int intersect_box(double x, double y, double z, double vx, double vy, double vz, double X,
                                             double Y, double Z, double* ts) {
  double t1 = __builtin_nan(""), t2 = __builtin_nan("");
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
  if (isnan(t1)) return 0;
  ts[0] = t1;
  ts[1] = t2;
  return 2;
}

// ---------------------------------------------------------------- CSG node types
#define MCV_NODE(S, I, k) (S[6 * (I) + (k)])

template <int I>
struct CsgSphere {
  static constexpr int K = 2;
  __device__ static __forceinline__ int locate(const double* S, double x, double y, double z) {
    return locate_sphere(x, y, z, MCV_NODE(S, I, 1));
  }
  __device__ static __forceinline__ int intersect(const double* S, double x, double y, double z, double vx, double vy,
                                                  double vz, double* ts) {
    return intersect_sphere(x, y, z, vx, vy, vz, MCV_NODE(S, I, 1), ts);
  }
};

template <int I>
struct CsgCylinder {
  static constexpr int K = 2;
  __device__ static __forceinline__ int locate(const double* S, double x, double y, double z) {
    return locate_cylinder(x, y, z, MCV_NODE(S, I, 1), MCV_NODE(S, I, 2));
  }
  __device__ static __forceinline__ int intersect(const double* S, double x, double y, double z, double vx, double vy,
                                                  double vz, double* ts) {
    return intersect_cylinder(x, y, z, vx, vy, vz, MCV_NODE(S, I, 1), MCV_NODE(S, I, 2), ts);
  }
};

template <int I>
struct CsgBlock {
  static constexpr int K = 2;
  __device__ static __forceinline__ int locate(const double* S, double x, double y, double z) {
    return locate_box(x, y, z, MCV_NODE(S, I, 1), MCV_NODE(S, I, 2), MCV_NODE(S, I, 3));
  }
  __device__ static __forceinline__ int intersect(const double* S, double x, double y, double z, double vx, double vy,
                                                  double vz, double* ts) {
    return intersect_box(x, y, z, vx, vy, vz, MCV_NODE(S, I, 1), MCV_NODE(S, I, 2), MCV_NODE(S, I, 3), ts);
  }
};

// Sorted on-border roots of two children (intersectComposite,
// acc/geometry/arrow_intersect.py:126-142).  Both child lists are sorted, so
// the reference's append + sorted-insert yields the sorted multiset of the
// kept roots; a compare-exchange network over KA+KB slots (dropped roots set
// to +inf) produces the same list without dynamic indexing.
template <class Self, class A, class B>
__device__ __forceinline__ int composite_intersect(const double* S, double x, double y, double z, double vx,
                                                   double vy, double vz, double* ts) {
  constexpr int KA = A::K, KB = B::K, KC = KA + KB;
  double ta[KA], tb[KB], c[KC];
  const int na = A::intersect(S, x, y, z, vx, vy, vz, ta);
  const int nb = B::intersect(S, x, y, z, vx, vy, vz, tb);
  int cnt = 0;
#pragma unroll
  for (int i = 0; i < KA; ++i) {
    const double t = ta[i];
    const bool ok = i < na && Self::locate(S, x + vx * t, y + vy * t, z + vz * t) == ONBORDER;
    c[i] = ok ? t : __builtin_inf();
    cnt += ok;
  }
#pragma unroll
  for (int i = 0; i < KB; ++i) {
    const double t = tb[i];
    const bool ok = i < nb && Self::locate(S, x + vx * t, y + vy * t, z + vz * t) == ONBORDER;
    c[KA + i] = ok ? t : __builtin_inf();
    cnt += ok;
  }
  // odd-even transposition sort (KC <= 10)
#pragma unroll
  for (int r = 0; r < KC; ++r) {
#pragma unroll
    for (int i = r & 1; i + 1 < KC; i += 2) {
      const double lo = c[i] < c[i + 1] ? c[i] : c[i + 1];
      const double hi = c[i] < c[i + 1] ? c[i + 1] : c[i];
      c[i] = lo;
      c[i + 1] = hi;
    }
  }
#pragma unroll
  for (int i = 0; i < cmin(KC, MAXI); ++i) ts[i] = c[i];
  return cnt < MAXI ? cnt : MAXI;
}

template <int I, class A, class B>
struct CsgUnion {
  static constexpr int K = cmin(A::K + B::K, MAXI);
  __device__ static __forceinline__ int locate(const double* S, double x, double y, double z) {
    const int l1 = A::locate(S, x, y, z);
    if (l1 == INSIDE) return INSIDE;
    const int l2 = B::locate(S, x, y, z);
    if (l2 == INSIDE) return INSIDE;
    if (l1 == OUTSIDE && l2 == OUTSIDE) return OUTSIDE;
    return ONBORDER;
  }
  __device__ static __forceinline__ int intersect(const double* S, double x, double y, double z, double vx, double vy,
                                                  double vz, double* ts) {
    return composite_intersect<CsgUnion, A, B>(S, x, y, z, vx, vy, vz, ts);
  }
};

template <int I, class A, class B>
struct CsgDifference {
  static constexpr int K = cmin(A::K + B::K, MAXI);
  __device__ static __forceinline__ int locate(const double* S, double x, double y, double z) {
    const int p1 = A::locate(S, x, y, z), p2 = B::locate(S, x, y, z);
    if (p1 == OUTSIDE || p2 == INSIDE) return OUTSIDE;
    if (p1 == ONBORDER || p2 == ONBORDER) return ONBORDER;
    return INSIDE;
  }
  __device__ static __forceinline__ int intersect(const double* S, double x, double y, double z, double vx, double vy,
                                                  double vz, double* ts) {
    return composite_intersect<CsgDifference, A, B>(S, x, y, z, vx, vy, vz, ts);
  }
};

template <int I, class A, class B>
struct CsgIntersection {
  static constexpr int K = cmin(A::K + B::K, MAXI);
  __device__ static __forceinline__ int locate(const double* S, double x, double y, double z) {
    const int l1 = A::locate(S, x, y, z);
    if (l1 == OUTSIDE) return OUTSIDE;
    const int l2 = B::locate(S, x, y, z);
    if (l2 == OUTSIDE) return OUTSIDE;
    if (l1 == INSIDE && l2 == INSIDE) return INSIDE;
    return ONBORDER;
  }
  __device__ static __forceinline__ int intersect(const double* S, double x, double y, double z, double vx, double vy,
                                                  double vz, double* ts) {
    return composite_intersect<CsgIntersection, A, B>(S, x, y, z, vx, vy, vz, ts);
  }
};

// ---------------------------------------------------------------- time bookkeeping
// acc/components/samples/homogeneous_single_scatterer.py:121-178, over a
// register array of compile-time bound K.
template <int K>
__device__ __forceinline__ double total_time_in_shape(const double* ts, int ninter) {
  double T = 0;
  int sign = 0;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    if (i < ninter) {
      const double t = ts[i];
      if (!(t < 0)) {
        sign = (sign == 0) ? ((i % 2 == 0) ? -1 : 1) : -sign;
        T += sign * t;
      }
    }
  }
  return T;
}

template <int K>
__device__ __forceinline__ void calc_time_to_point_of_scattering(const double* ts, int ninter, double rand,
                                                                 double& total, double& ttt_sp, double& ttis_sp) {
  constexpr int KP = (K + 1) / 2 + 1;
  double tps[KP], tpsg[KP];
#pragma unroll
  for (int j = 0; j < KP; ++j) tps[j] = tpsg[j] = 0;
  double T = 0.0;
  int sign = 0, ip = 0;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    if (i < ninter) {
      const double t = ts[i];
      if (!(t < 0)) {
        sign = (sign == 0) ? ((i % 2 == 0) ? -1 : 1) : -sign;
        T += sign * t;
        if (sign > 0) {
#pragma unroll
          for (int j = 0; j < KP; ++j)
            if (j == ip) { tps[j] = T; tpsg[j] = t; }
          ip += 1;
        }
      }
    }
  }
  total = T;
  const double ttistsp = rand * T;
  // first partial sum exceeding ttistsp, else the last one
  double a = 0, b = 0;
  bool found = false;
#pragma unroll
  for (int j = 0; j < KP; ++j) {
    if (j < ip && !found) {
      a = tps[j];
      b = tpsg[j];
      if (ttistsp < tps[j]) found = true;
    }
  }
  ttt_sp = ttistsp - a + b;
  ttis_sp = ttistsp;
}

// ---------------------------------------------------------------- kernels
// Sample parameter block (packed): P[0] mu, P[1] sigma, P[2] min_prob,
// P[3..15] kernel params, P[16..23] iparams (n_nodes, kernel_type,
// max_scattered, max_ms_loops, max_ms_loops_path1, kernel_data_offset, n0, n1).
// D = sample data block: CSG nodes, then kernel tables at P[21].

// acc/kernels/__init__.py:156-161 (mu / v * 2200)
__device__ __forceinline__ double absorption_coeff(const double* P, const Neutron& n) {
  const double v = sqrt(n.vx * n.vx + n.vy * n.vy + n.vz * n.vz);
  return P[0] / v * 2200;
}

// acc/kernels/isotropic.py:5-21 with p *= sigma (acc/kernels/__init__.py:36-44)
template <class T>
__device__ __forceinline__ void isotropic_scatter(Rng& rng, Neutron& n, double sigma, T& tr) {
  tr.sigma(sigma);
  n.p *= sigma;
  double cos_t = rng.uniform() * 2 - 1;
  const double phi = rng.uniform() * (2 * PI);
  if (cos_t > 1) cos_t = 1;
  const double sin_t = sqrt(1 - cos_t * cos_t);
  double sin_p, cos_p;
  sincos(phi, &sin_p, &cos_p);
  const double vi = sqrt(n.vx * n.vx + n.vy * n.vy + n.vz * n.vz);
  n.vx = vi * sin_t * cos_p;
  n.vy = vi * sin_t * sin_p;
  n.vz = vi * cos_t;
}

// bilinear grid lookup with the index clamping of acc/kernels/SANS2D_ongrid.py:16-56;
// (ix, iy) is the lower node of the cell used and (fx, fy) the position in it.
__device__ __forceinline__ double grid_interp(const double* __restrict__ g, int64_t nx, int64_t ny, double x,
                                              double xmin, double xmax, double y, double ymin, double ymax,
                                              int64_t& ix, int64_t& iy, double& fx, double& fy) {
  const double dx = (xmax - xmin) / (double)nx, dy = (ymax - ymin) / (double)ny;
  ix = (int64_t)((x - xmin) / dx);
  iy = (int64_t)((y - ymin) / dy);
  ix = ix < 0 ? 0 : ix;
  iy = iy < 0 ? 0 : iy;
  ix = ix < nx - 2 ? ix : nx - 2;
  iy = iy < ny - 2 ? iy : ny - 2;
  const double f00 = g[ix * ny + iy], f10 = g[(ix + 1) * ny + iy], f01 = g[ix * ny + iy + 1],
               f11 = g[(ix + 1) * ny + iy + 1];
  fx = (x - xmin) / dx - (double)ix;
  fy = (y - ymin) / dy - (double)iy;
  const double a00 = f00, a10 = f10 - f00, a01 = f01 - f00, a11 = f11 - f10 - f01 + f00;
  return a00 + a10 * fx + a01 * fy + a11 * fx * fy;
}

// phonon-DOS kernel value (no reference): Q^2 sum_k A_k exp(-(E-c_k)^2 / (2 w_k^2))
__device__ __forceinline__ double dos_S(const double* __restrict__ kp, double Q, double E) {
  double g = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double u = (E - kp[5 + 3 * k]) / kp[6 + 3 * k];
    g += kp[4 + 3 * k] * exp(-0.5 * u * u);
  }
  return Q * Q * g;
}

// S(Q,E) kernels (no reference): E_Q-style kinematic closure
// (acc/kernels/E_Q.py:37-99) with Q and E both drawn uniformly, one draw, p = 0
// when kinematically forbidden.  KIND = MCV_KERNEL_SQE_GRID (2) or PHONON_DOS (3).
template <int KIND, class T>
__device__ __forceinline__ void sqe_scatter(Rng& rng, Neutron& n, const double* __restrict__ P,
                                            const double* __restrict__ D, T& tr) {
  tr.sigma(P[1]);
  n.p *= P[1];
  const double* kp = P + 3;
  const double Qmin = kp[0], Qmax = kp[1], Emin = kp[2], Emax = kp[3];
  const double vx0 = n.vx, vy0 = n.vy, vz0 = n.vz;
  const double vi = sqrt(vx0 * vx0 + vy0 * vy0 + vz0 * vz0);
  const double Ei = vi * vi * VS2E;
  double Q = rng.uniform() * (Qmax - Qmin);
  Q += Qmin;
  double E = rng.uniform() * (Emax - Emin);
  E += Emin;
  const double Ef = Ei - E;
  if (Ef <= 0) {
    n.p = 0;
    return;
  }
  const double vf = sqrt(Ef) * SE2V;
  const double ki = V2K * vi, kf = V2K * vf;
  const double cost = (ki * ki + kf * kf - Q * Q) / (2 * ki * kf);
  const double cost2 = cost * cost;
  if (cost2 > 1) {
    n.p = 0;
    return;
  }
  const double sint = sqrt(1.0 - cost2);
  const double phi = rng.uniform() * 2.0 * PI;
  double sp, cp;
  sincos(phi, &sp, &cp);
  const double v2 = vf * sint * cp, v3 = vf * sint * sp, v1 = vf * cost;
  const double il = 1.0 / vi;
  const double e1x = vx0 * il, e1y = vy0 * il, e1z = vz0 * il;
  double nx = 1, nz = 0;
  if (fabs(e1x) > 1e-1 || fabs(e1y) > 1e-1) { nx = 0; nz = 1; }
  // e2 = normalize(norm x e1); e3 = e1 x e2
  double e2x = 0 * e1z - nz * e1y, e2y = nz * e1x - nx * e1z, e2z = nx * e1y - 0 * e1x;
  const double l2 = 1.0 / sqrt(e2x * e2x + e2y * e2y + e2z * e2z);
  e2x *= l2;
  e2y *= l2;
  e2z *= l2;
  const double e3x = e1y * e2z - e1z * e2y, e3y = e1z * e2x - e1x * e2z, e3z = e1x * e2y - e1y * e2x;
  n.vx = e1x * v1 + e2x * v2 + e3x * v3;
  n.vy = e1y * v1 + e2y * v2 + e3y * v3;
  n.vz = e1z * v1 + e2z * v2 + e3z * v3;
  double Sv;
  if constexpr (KIND == 2) {
    const int64_t n0 = (int64_t)P[22], n1 = (int64_t)P[23];
    int64_t ix, iy;
    double fx, fy;
    Sv = grid_interp(D + (int64_t)P[21], n0, n1, Q, Qmin, Qmax, E, Emin, Emax, ix, iy, fx, fy);
    tr.grid(ix * n1 + iy, (ix + 1) * n1 + iy, ix * n1 + iy + 1, (ix + 1) * n1 + iy + 1, fx, fy, Sv);
  } else {
    Sv = dos_S(kp, Q, E);
    tr.dos(kp, E);
  }
  const double prob = Sv * (vf / vi) * Q * (Qmax - Qmin) * (Emax - Emin) / (kf * ki) / 2;
  n.p *= prob;
}

template <int KIND, class T>
__device__ __forceinline__ void kernel_scatter(Rng& rng, Neutron& n, const double* __restrict__ P,
                                               const double* __restrict__ D, T& tr) {
  if constexpr (KIND == 1)
    isotropic_scatter(rng, n, P[1], tr);
  else
    sqe_scatter<KIND>(rng, n, P, D, tr);
}

template <int KIND>
__device__ __forceinline__ void kernel_scatter(Rng& rng, Neutron& n, const double* __restrict__ P,
                                               const double* __restrict__ D) {
  NoTrack tr;
  kernel_scatter<KIND>(rng, n, P, D, tr);
}

// ---------------------------------------------------------------- HSS
// acc/components/samples/homogeneous_single_scatterer.py:59-100
template <class Shape, int KIND, class T>
__device__ __forceinline__ void hss_propagate_t(Rng& rng, Neutron& n, const double* __restrict__ P,
                                                const double* __restrict__ D, T tr) {
  constexpr int K = Shape::K;
  double ts[K];
  const double vx = n.vx, vy = n.vy, vz = n.vz;
  int ninter = Shape::intersect(D, n.x, n.y, n.z, vx, vy, vz, ts);
  if (ninter < 2) return;
  double last = ts[0];
#pragma unroll
  for (int i = 1; i < K; ++i)
    if (i == ninter - 1) last = ts[i];
  if (last <= 0) return;
  const double rand = rng.uniform();
  double tis, dt, ttis;
  calc_time_to_point_of_scattering<K>(ts, ninter, rand, tis, dt, ttis);
  if (dt <= 0) return;
  prop_dt_inplace(n, dt);
  const double v = sqrt(vx * vx + vy * vy + vz * vz);
  const double dist = v * ttis;
  const double fulllen = v * tis;
  const double sigma = P[1];
  double mu = absorption_coeff(P, n);
  const double atten = exp(-(mu + sigma) * dist);
  tr.attenuate(dist, 1.0 / v);
  n.p *= fulllen * atten;
  kernel_scatter<KIND>(rng, n, P, D, tr);
  if (n.p <= 0) return;
  ninter = Shape::intersect(D, n.x, n.y, n.z, n.vx, n.vy, n.vz, ts);
  const double dt3 = total_time_in_shape<K>(ts, ninter);
  prop_dt_inplace(n, dt3);
  mu = absorption_coeff(P, n);
  tr.attenuate(v * dt3, 1.0 / sqrt(n.vx * n.vx + n.vy * n.vy + n.vz * n.vz));
  n.p *= exp(-(mu + sigma) * v * dt3);
}

template <class Shape, int KIND>
__device__ __forceinline__ void hss_propagate(Rng& rng, Neutron& n, const double* __restrict__ P,
                                              const double* __restrict__ D) {
  hss_propagate_t<Shape, KIND>(rng, n, P, D, NoTrack{});
}

}  // namespace mcv
