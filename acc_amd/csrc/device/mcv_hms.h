// mcv_hms.h -- propagation helpers and the homogeneous multiple scatterer on
// the device.
//
// acc/geometry/propagation.py:9-191 and
// acc/components/samples/homogeneous_multiple_scatterer.py:63-303, restated
// over compile-time bounded root lists (Shape::K).  The reference's lists of
// neutrons still to scatter (to_be_scattered, to_be_scattered2) are per-lane
// arrays indexed at run time, so they live in the lane's private scratch; its
// lists of an encounter's outputs (scattered) and of the scatterer's outputs
// are not kept: each output is handed on as it is produced (hms_scatter).
// The scatterer and its first-encounter loop are inlined into the chain
// kernel: as separate functions their frames and by-reference neutrons went
// through scratch at every call (c3, 1e8 neutrons: 75 ms with both out of
// line, 70 with the first-encounter loop inlined, 60 with both; 48.2 -> 43.4
// ms later with the output lists gone, 36 ms with each root list and location
// computed once where the reference computes it twice).  Keeping each list's
// first entry in registers measured slower (52 ms: the register file is full,
// and spills replace the scratch traffic).
#pragma once
#include "mcv_samples.h"

namespace mcv {

// forward_intersect: roots >= 0 shifted to the front (propagation.py:11-22)
template <class Shape>
__device__ __forceinline__ int forward_intersect(const double* S, double x, double y, double z, double vx, double vy,
                                                 double vz, double* out) {
  constexpr int K = Shape::K;
  double ts[K];
  const int N = Shape::intersect(S, x, y, z, vx, vy, vz, ts);
  int Nm = N;
#pragma unroll
  for (int i = K - 1; i >= 0; --i)
    if (i < N && ts[i] >= 0) Nm = i;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    double v = 0;
#pragma unroll
    for (int j = 0; j < K; ++j)
      if (j == i + Nm) v = ts[j];
    out[i] = v;
  }
  return N - Nm;
}

template <class Shape>
__device__ __forceinline__ int locate0(const double* S, double x, double y, double z) {
  return Shape::locate(S, x, y, z);
}

// A location not computed yet (the reference's OUTSIDE, ONBORDER and INSIDE
// are 1, 0 and -1).  The helpers below that take `loc` skip the reference's
// first locate() of the neutron when the caller already knows its result for
// the same position: locate() is a pure function of the position.
constexpr int LOC_UNKNOWN = 2;

// propagation.py:24-50
template <class Shape>
__device__ __forceinline__ bool is_exiting(const double* S, const Neutron& n, int loc = LOC_UNKNOWN) {
  constexpr int K = Shape::K;
  if (loc == LOC_UNKNOWN) loc = locate0<Shape>(S, n.x, n.y, n.z);
  if (loc == INSIDE) return false;
  double ts[K];
  const int N = forward_intersect<Shape>(S, n.x, n.y, n.z, n.vx, n.vy, n.vz, ts);
  if (N == 0) return true;
  if (loc == OUTSIDE) return false;
  double previous = 0.;
#pragma unroll
  for (int it = 0; it < K; ++it) {
    if (it < N) {
      const double now = ts[it];
      const double middle = (previous + now) / 2.;
      if (locate0<Shape>(S, n.x + n.vx * middle, n.y + n.vy * middle, n.z + n.vz * middle) != ONBORDER) return false;
      previous = now;
    }
  }
  return true;
}

// propagation.py:52-59
template <class Shape>
__device__ __forceinline__ void propagate_out(const double* S, Neutron& n) {
  constexpr int K = Shape::K;
  double ts[K];
  const int N = forward_intersect<Shape>(S, n.x, n.y, n.z, n.vx, n.vy, n.vz, ts);
  if (N == 0) return;
  double last = ts[0];
#pragma unroll
  for (int i = 1; i < K; ++i)
    if (i == N - 1) last = ts[i];
  prop_dt_inplace(n, last);
}

// propagation.py:61-105
template <class Shape>
__device__ __forceinline__ void propagate_to_next_incident_surface(const double* S, Neutron& n,
                                                                   int loc = LOC_UNKNOWN) {
  constexpr int K = Shape::K;
  const double x = n.x, y = n.y, z = n.z, vx = n.vx, vy = n.vy, vz = n.vz;
  if (loc == LOC_UNKNOWN) loc = locate0<Shape>(S, x, y, z);
  if (loc == INSIDE) return;
  double ts[K];
  const int N = forward_intersect<Shape>(S, x, y, z, vx, vy, vz, ts);
  if (N == 0) return;
  double t = 0;
  if (loc == OUTSIDE) {
    t = ts[0];
  } else {
    bool found = false, done = false;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      if (i < N && !done) {
        const int loc1 = locate0<Shape>(S, x + ts[i] / 2 * vx, y + ts[i] / 2 * vy, z + ts[i] / 2 * vz);
        if (loc1 == INSIDE) return;
        if (loc1 == OUTSIDE) {
          t = ts[i];
          found = true;
          done = true;
        }
      }
    }
    if (!found) return;
  }
  prop_dt_inplace(n, t);
}

// propagation.py:107-152
template <class Shape>
__device__ __forceinline__ void propagate_to_next_exiting_surface(const double* S, Neutron& n) {
  constexpr int K = Shape::K;
  double ts[K];
  const int loc = locate0<Shape>(S, n.x, n.y, n.z);
  double t;
  if (loc == INSIDE) {
    forward_intersect<Shape>(S, n.x, n.y, n.z, n.vx, n.vy, n.vz, ts);
    t = ts[0];
  } else {
    if (loc == OUTSIDE) propagate_to_next_incident_surface<Shape>(S, n);
    const double x = n.x, y = n.y, z = n.z, vx = n.vx, vy = n.vy, vz = n.vz;
    const int N = forward_intersect<Shape>(S, x, y, z, vx, vy, vz, ts);
    bool found = false, done = false;
    t = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      if (i < N && !done) {
        const int loc1 = locate0<Shape>(S, x + ts[i] / 2 * vx, y + ts[i] / 2 * vy, z + ts[i] / 2 * vz);
        if (loc1 == OUTSIDE) return;
        if (loc1 == INSIDE) {
          t = ts[i];
          found = true;
          done = true;
        }
      }
    }
    if (!found) return;
  }
  prop_dt_inplace(n, t);
}

// propagation.py:154-191
template <class Shape>
__device__ __forceinline__ double tof_before_first_exit(const double* S, const Neutron& n) {
  constexpr int K = Shape::K;
  const double x = n.x, y = n.y, z = n.z, vx = n.vx, vy = n.vy, vz = n.vz;
  const int loc = locate0<Shape>(S, x, y, z);
  if (loc == OUTSIDE) return 0.;
  double ts[K];
  const int N = forward_intersect<Shape>(S, x, y, z, vx, vy, vz, ts);
  if (N == 0) return 0.;
  if (loc == INSIDE) return ts[0];
  double t = 0.;
  bool done = false;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    if (i < N && !done) {
      const int loc1 = locate0<Shape>(S, x + ts[i] / 2 * vx, y + ts[i] / 2 * vy, z + ts[i] / 2 * vz);
      if (loc1 == OUTSIDE) return 0.;
      if (loc1 == INSIDE) {
        t = ts[i];
        done = true;
      }
    }
  }
  return t;
}

__device__ __forceinline__ bool is_moving(const Neutron& n) {  // acc/neutron.py:38-42
  return sqrt(n.vx * n.vx + n.vy * n.vy + n.vz * n.vz) > 1.e-6;
}

__device__ __forceinline__ double atten_coeff(const double* P, const Neutron& n) {  // :63-67
  return P[1] + absorption_coeff(P, n);
}

// :69-79, for a neutron inside the shape (its only use).  Inside, the
// reference's tof_before_first_exit is the first forward root (0 without
// one) and propagate_to_next_exiting_surface moves by that same first root:
// one intersection serves both.
template <class Shape>
__device__ __forceinline__ void propagate_inside_to_exiting_surface_with_attenuation(const double* P,
                                                                                    const double* D, Neutron& n) {
  const double vx = n.vx, vy = n.vy, vz = n.vz;
  double ts[Shape::K];
  const int N = forward_intersect<Shape>(D, n.x, n.y, n.z, vx, vy, vz, ts);
  const double tof = N == 0 ? 0. : ts[0];
  const double v = sqrt(vx * vx + vy * vy + vz * vz);
  const double distance = tof * v;
  const double atten = atten_coeff(P, n);
  const double transmission_prob = exp(-atten * distance);
  prop_dt_inplace(n, ts[0]);
  n.p *= transmission_prob;
}

// :81-121; out[0] transmitted, out[1] scattered
// (`loc`: the neutron's location if the caller knows it.)
template <class Shape, int KIND>
__device__ __forceinline__ int interactM1(Rng& rng, const double* P, const double* D, Neutron* out, Neutron& n,
                                          int loc = LOC_UNKNOWN) {
  const double vx = n.vx, vy = n.vy, vz = n.vz;
  if (loc == LOC_UNKNOWN) loc = locate0<Shape>(D, n.x, n.y, n.z);
  double tof;
  if (loc == INSIDE) {  // the neutron stays put: tof_before_first_exit is its first root
    double ts[Shape::K];
    const int N = forward_intersect<Shape>(D, n.x, n.y, n.z, vx, vy, vz, ts);
    tof = N == 0 ? 0. : ts[0];
  } else {
    propagate_to_next_incident_surface<Shape>(D, n, loc);
    tof = tof_before_first_exit<Shape>(D, n);
  }
  if (tof == 0.) return 0;
  const double v = sqrt(vx * vx + vy * vy + vz * vz);
  const double distance = tof * v;
  const double sigma = P[1];
  const double mu = absorption_coeff(P, n);
  const double transmission_prob = exp(-(mu + sigma) * distance);
  // propagate_to_next_exiting_surface(out[0]) moves it by exactly tof here:
  // the neutron is inside or on the border (else tof is 0), and for both the
  // reference computes the same roots and the same midpoint tests as
  // tof_before_first_exit did
  out[0] = n;
  prop_dt_inplace(out[0], tof);
  out[0].p *= transmission_prob;
  // the absorption branch writes out[1] and is overwritten below; the kernels'
  // absorb() is a no-op (acc/kernels/__init__.py:163-165)
  const double rand = rng.uniform();
  const double xs = distance * rand;
  const double prob = distance * exp(-(mu + sigma) * xs);
  out[1] = n;
  out[1].p *= prob;
  prop_dt_inplace(out[1], xs / v);
  kernel_scatter<KIND>(rng, out[1], P, D);
  return 2;
}

// The breadth-first lists hold neutrons without their spins (nothing here
// changes them, so every neutron of the lists carries the incident one's):
// 64 instead of 80 bytes of lane scratch per entry.  A list entry is copied
// into registers before it is worked on, and written back only where the
// reference stores it into a list.
struct NeutronS {
  double x, y, z, vx, vy, vz, t, p;
};

__device__ __forceinline__ void store_s(NeutronS& d, const Neutron& n) {
  d.x = n.x; d.y = n.y; d.z = n.z; d.vx = n.vx; d.vy = n.vy; d.vz = n.vz; d.t = n.t; d.p = n.p;
}

__device__ __forceinline__ void store_s(Neutron& d, const Neutron& n) { d = n; }

__device__ __forceinline__ Neutron load_s(const NeutronS& d, double s1, double s2) {
  Neutron n;
  n.x = d.x; n.y = d.y; n.z = d.z; n.vx = d.vx; n.vy = d.vy; n.vz = d.vz; n.t = d.t; n.p = d.p;
  n.s1 = s1;
  n.s2 = s2;
  return n;
}

// :128-178.  Each output is handed to emit(const Neutron&, int loc) in the
// order in which the reference stores it into its output list (no list is
// kept), with its location when known (LOC_UNKNOWN otherwise); the return
// value is the reference's output count.  Entries of the list of neutrons to
// scatter are inside the shape (that is the test that put them there), except
// the incident neutron.
template <class Shape, int KIND, int M, class Emit>
__device__ __forceinline__ int interactM_path1(Rng& rng, const double* P, const double* D, int out_len,
                                               const Neutron& n, Emit&& emit) {
  const double min_prob = P[2];
  const int loops = (int)P[20];
  if (n.p < min_prob) return 0;
  NeutronS bufA[M], bufB[M];
  Neutron scattered[2];
  NeutronS* tbs = bufA;
  NeutronS* tbs2 = bufB;
  int N_tbs = 1;
  store_s(tbs[0], n);
  int out_index = 0;
  const int lim = M < out_len ? M : out_len;
  for (int iloop = 0; iloop < loops; ++iloop) {
    int s2 = 0;
    for (int ineutron = 0; ineutron < N_tbs; ++ineutron) {
      Neutron cur = load_s(tbs[ineutron], n.s1, n.s2);
      const int ns1 = interactM1<Shape, KIND>(rng, P, D, scattered, cur, iloop > 0 ? INSIDE : LOC_UNKNOWN);
      // at most two outputs: constant indices keep them in registers
#pragma unroll
      for (int is = 0; is < 2; ++is) {
        if (is >= ns1) break;
        const Neutron& sc1 = scattered[is];
        if (sc1.p < min_prob) continue;
        const int loc = locate0<Shape>(D, sc1.x, sc1.y, sc1.z);
        if (loc != INSIDE) {
          emit(sc1, loc);
          out_index += 1;
          if (out_index >= lim) return out_index;
          continue;
        }
        store_s(tbs2[s2], sc1);
        s2 += 1;
        if (s2 >= M) break;
      }
      if (s2 >= M) break;
    }
    NeutronS* tmp = tbs;
    tbs = tbs2;
    tbs2 = tmp;
    N_tbs = s2;
    if (!N_tbs) break;
  }
  for (int i = 0; i < N_tbs; ++i) {
    if (out_index >= lim) break;
    emit(load_s(tbs[i], n.s1, n.s2), loops > 0 ? INSIDE : LOC_UNKNOWN);
    out_index += 1;
  }
  return out_index;
}

// :214-273.  P[19] = max_ms_loops, P[20] = max_ms_loops_path1 (packed iparams).
// Outputs go to emit(const Neutron&) in the reference's output-list order.
// The reference lets the first encounter (interactM_path1) finish, with all
// its draws, before it reads that encounter's outputs, and none of that
// reading draws; so reading each output as the encounter hands it over, and
// ignoring the rest once the reference would stop reading (its list of
// neutrons to scatter is full, or M outputs are out), keeps every draw and
// every output.  The incident neutron is taken by value: an emit that
// overwrites the caller's neutron does not change the spins read here.
// CHECKED: the caller has already found that is_exiting(n) is false (the
// compacted kernels test it to decide which neutrons interact).
template <class Shape, int KIND, int M, bool CHECKED = false, class Emit>
__device__ __forceinline__ int hms_scatter(Rng& rng, const Neutron n, const double* __restrict__ P,
                                           const double* __restrict__ D, Emit&& emit) {
  if (!CHECKED && is_exiting<Shape>(D, n)) {
    emit(n);
    return 1;
  }
  const double min_prob = P[2];
  const int loops = (int)P[19];
  NeutronS bufA[M], bufB[M];
  NeutronS* tbs = bufA;
  NeutronS* tbs2 = bufB;
  int N_tbs = 1;
  store_s(tbs[0], n);
  int out_index = 0;
  for (int iloop = 0; iloop < loops; ++iloop) {
    int s2 = 0;
    bool full = false, done = false;  // the reference's break (s2 reached M) and return (M outputs)
    for (int ineutron = 0; ineutron < N_tbs; ++ineutron) {
      const Neutron cur = load_s(tbs[ineutron], n.s1, n.s2);
      interactM_path1<Shape, KIND, M>(rng, P, D, M, cur, [&](const Neutron& o, int oloc) {
        if (full || done) return;
        Neutron sc1 = o;
        if (sc1.p < min_prob) return;
        if (!is_moving(sc1)) return;
        const int loc = oloc == LOC_UNKNOWN ? locate0<Shape>(D, sc1.x, sc1.y, sc1.z) : oloc;
        bool exiting;
        if (loc == INSIDE) {
          propagate_inside_to_exiting_surface_with_attenuation<Shape>(P, D, sc1);
          exiting = is_exiting<Shape>(D, sc1);
        } else {
          exiting = is_exiting<Shape>(D, sc1, loc);
        }
        if (exiting) {
          emit(sc1);
          out_index += 1;
          done = out_index >= M;
          return;
        }
        store_s(tbs2[s2], sc1);
        s2 += 1;
        full = s2 >= M;
      });
      if (done) return out_index;
      if (full) break;
    }
    NeutronS* tmp = tbs;
    tbs = tbs2;
    tbs2 = tmp;
    N_tbs = s2;
    if (!N_tbs) break;
  }
  for (int i = 0; i < N_tbs; ++i) {
    if (out_index >= M) break;
    Neutron nn = load_s(tbs[i], n.s1, n.s2);
    const double x = nn.x, y = nn.y, z = nn.z;
    propagate_out<Shape>(D, nn);
    const double dist = sqrt((nn.x - x) * (nn.x - x) + (nn.y - y) * (nn.y - y) + (nn.z - z) * (nn.z - z));
    const double atten = atten_coeff(P, nn);
    nn.p *= exp(-atten * dist);
    emit(nn);
    out_index += 1;
  }
  return out_index;
}

// The scatterer with an output list: `out` holds Neutron rows (buffered mode)
// or spin-less NeutronS entries (fused chains whose downstream components
// draw, which restore the incident neutron's spins).
template <class Shape, int KIND, int M, class Out>
__device__ __forceinline__ int hms_scatterM(Rng& rng, const Neutron& n, const double* __restrict__ P,
                                            const double* __restrict__ D, Out* out) {
  int k = 0;
  return hms_scatter<Shape, KIND, M>(rng, n, P, D, [&](const Neutron& o) {
    store_s(out[k], o);
    ++k;
  });
}

}  // namespace mcv
