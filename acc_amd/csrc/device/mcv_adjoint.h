// mcv_adjoint.h -- vector-Jacobian product of the monitor histograms with
// respect to the weight-carrying parameters of a chain.
//
// No reference counterpart: the reference only runs the forward simulation.
// For fixed random streams a neutron's path (positions, velocities, the bins
// it lands in) does not depend on the supermirror coating (R0, Qc, alpha, m,
// W), the sample's absorption/scattering coefficients or its S(Q,E) values:
// those only scale its weight p.  So for the histogram sums
//   H[m][b] = sum_i p_i [bin_m(i) = b],   H2[m][b] = sum_i p_i^2 [bin_m(i) = b]
// and an upstream gradient (G1, G2) of the same shape,
//   dL/dtheta = sum_i sum_m (G1[m][bin] p_i + 2 G2[m][bin] p_i^2) dlog p_i/dtheta,
// with dlog p_i/dtheta accumulated along the path as each factor is applied
// (the "score").  The adjoint kernel re-runs the forward chain with the same
// streams (path segments are recomputed, nothing is stored), carries the
// score in registers, and at every monitor crossing adds the weighted score
// to per-lane sums that are reduced across the wave into `grad` at the end.
#pragma once
#include "mcv_kernel.h"

namespace mcv {

// d log R / d (R0, Qc, alpha, m, W) of calc_reflectivity, for R > 0.
__device__ __forceinline__ void reflectivity_score(double q, double R0, double Qc, double alpha, double m, double W,
                                                   double& dR0, double& dQc, double& dalpha, double& dm,
                                                   double& dW) {
  dR0 = 1.0 / R0;
  dQc = dalpha = dm = dW = 0.0;
  if (q > Qc) {
    const double a = (q - m * Qc) / W;
    const double k = 1.0 + tanh(a);  // -d log(1 - tanh a) / da
    const double lin = 1.0 - alpha * (q - Qc);
    dm = k * Qc / W;
    dW = k * a / W;
    dQc = k * m / W + alpha / lin;
    dalpha = -(q - Qc) / lin;
  }
}

// Straight guide: score slots R0, Qc, alpha, m, W.
struct GuideScore : NoTrack {
  double* s;
  __device__ explicit GuideScore(double* s_) : s(s_) {}
  __device__ __forceinline__ void mirror(int, double q, double R0, double Qc, double alpha, double m, double W,
                                         double dmdp) {
    double a, b, c, d, e;
    reflectivity_score(q, R0, Qc, alpha, m, W, a, b, c, d, e);
    s[0] += a;
    s[1] += b;
    s[2] += c;
    s[3] += d * dmdp;
    s[4] += e;
  }
};

// Tapered guide: score slots R0, Qcx, Qcy, alphax, alphay, mx, my, W.
struct TaperScore : NoTrack {
  double* s;
  __device__ explicit TaperScore(double* s_) : s(s_) {}
  __device__ __forceinline__ void mirror(int axis, double q, double R0, double Qc, double alpha, double m, double W,
                                         double dmdp) {
    double a, b, c, d, e;
    reflectivity_score(q, R0, Qc, alpha, m, W, a, b, c, d, e);
    s[0] += a;
    if (axis == 0) {
      s[1] += b;
      s[3] += c;
      s[5] += d * dmdp;
    } else {
      s[2] += b;
      s[4] += c;
      s[6] += d * dmdp;
    }
    s[7] += e;
  }
};

// Grid-kernel cell of the one scattering event (dense gradient), weights dS/dnode / S.
struct DenseTrack {
  int64_t idx[4];
  double w[4];
  bool set;
};

// Sample: score slots mu_a(2200 m/s), sigma, then the phonon-DOS peaks
// (A_k, c_k, w_k for k = 1..3, the packed order P[7..15]) when present.
struct SampleScore : NoTrack {
  double* s;
  DenseTrack* dz;
  __device__ SampleScore(double* s_, DenseTrack* dz_) : s(s_), dz(dz_) {}
  __device__ __forceinline__ void attenuate(double path, double inv_v) {
    s[0] -= 2200.0 * inv_v * path;
    s[1] -= path;
  }
  __device__ __forceinline__ void sigma(double sig) { s[1] += 1.0 / sig; }
  __device__ __forceinline__ void dos(const double* kp, double E) {
    double e[3], g = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double u = (E - kp[5 + 3 * k]) / kp[6 + 3 * k];
      e[k] = exp(-0.5 * u * u);
      g += kp[4 + 3 * k] * e[k];
    }
    if (!(g > 0)) return;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double A = kp[4 + 3 * k], c = kp[5 + 3 * k], w = kp[6 + 3 * k];
      const double u = (E - c) / w;
      s[2 + 3 * k] += e[k] / g;
      s[3 + 3 * k] += A * e[k] * u / (w * g);
      s[4 + 3 * k] += A * e[k] * u * u / (w * g);
    }
  }
  __device__ __forceinline__ void grid(int64_t i00, int64_t i10, int64_t i01, int64_t i11, double fx, double fy,
                                       double S) {
    if (!(S > 0)) return;
    dz->idx[0] = i00;
    dz->idx[1] = i10;
    dz->idx[2] = i01;
    dz->idx[3] = i11;
    dz->w[0] = (1 - fx) * (1 - fy) / S;
    dz->w[1] = fx * (1 - fy) / S;
    dz->w[2] = (1 - fx) * fy / S;
    dz->w[3] = fx * fy / S;
    dz->set = true;
  }
};

// Weight of one monitor crossing in dL/dtheta.
__device__ __forceinline__ double adjoint_weight(const double* __restrict__ G, int64_t nbins, int64_t b, double p) {
  return G[nbins + b] * p + 2.0 * G[2 * nbins + b] * p * p;
}

__device__ __forceinline__ void dense_add(const AdjointArgs& J, const DenseTrack& dz, double c) {
  if (dz.set && J.dense != nullptr && c != 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) unsafeAtomicAdd(J.dense + dz.idx[k], c * dz.w[k]);
  }
}

// Chain::body(rng, n, A, J, acc, nidx) runs neutron nidx and adds its monitor
// contributions to acc[0 .. Chain::NS).
template <class Chain>
__device__ __forceinline__ void run_lane_adjoint(const LaunchArgs& A, const AdjointArgs& J, int64_t tid) {
  Rng rng;
  int64_t idx, end;
  stream_range(A, tid, rng, idx, end);
  double acc[Chain::NS];
#pragma unroll
  for (int k = 0; k < Chain::NS; ++k) acc[k] = 0;
  Neutron n;
  n.s1 = 0;
  n.s2 = 0;
  for (int64_t i = idx; i < end; ++i) Chain::body(rng, n, A, J, acc, i);
  count_processed(A.counter, end - idx);
#pragma unroll
  for (int k = 0; k < Chain::NS; ++k) {
    double v = acc[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0 && v != 0) unsafeAtomicAdd(J.grad + k, v);
  }
}

// Compacted adjoint: the forward kernel's stream-carrying two-phase scheme
// (run_block_compacted_rng) with the score slots of the phase-A components
// (Chain::NX doubles) travelling with the queued neutron.  Chain::phase_a(rng,
// n, A, J, acc, xa, nidx) returns whether the neutron reaches phase B with
// its phase-A scores in xa; Chain::phase_b(rng, n, A, J, acc, xa) finishes it.
template <int NX>
struct QueueAdjLDS : QueueRngLDS {
  double xt[NX][MCV_QCAP];
};

template <class Chain>
__device__ __forceinline__ void run_block_adjoint_rng(const LaunchArgs& A, const AdjointArgs& J, int64_t tid,
                                                      QueueAdjLDS<Chain::NX>& q, unsigned long long* pool = nullptr,
                                                      int64_t pool_base = 0) {
  const int lt = threadIdx.x;
  constexpr unsigned ALL = (1u << MCV_SPL) - 1u;
  double acc[Chain::NS];
#pragma unroll
  for (int k = 0; k < Chain::NS; ++k) acc[k] = 0;
  if (lt == 0) q.count = 0;
  unsigned busy = 0, done = 0;
  bool dry = pool == nullptr;
  int64_t mine = 0;
  slots_init(A, q, tid, done, mine);
  int par = 0, cnt = 0, rr = 0;
  __syncthreads();
  while (true) {
    if (!dry) slots_refill(A, q, pool, pool_base, INT64_MAX, busy, done, dry, mine);
    const unsigned avail = ~(busy | done) & ALL;
    bool has = false;
    Neutron n;
    n.s1 = 0;
    n.s2 = 0;
    double xa[Chain::NX];
#pragma unroll
    for (int k = 0; k < Chain::NX; ++k) xa[k] = 0;
    int sel = -1;
    if (avail) {
      const unsigned rot = ((avail >> rr) | (avail << (MCV_SPL - rr))) & ALL;
      sel = (rr + __builtin_ctz(rot)) % MCV_SPL;
      const int o = lt * MCV_SPL + sel;
      Rng w;
      w.s0 = q.st0[o];
      w.s1 = q.st1[o];
      const int64_t my = q.next[o];
      has = Chain::phase_a(w, n, A, J, acc, xa, my) != 0;
      q.st0[o] = w.s0;
      q.st1[o] = w.s1;
      q.next[o] = my + 1;
      if (my + 1 >= q.end[o]) done |= 1u << sel;
      if (has) busy |= 1u << sel;
      rr = (sel + 1) % MCV_SPL;
    }
    const int s = q_append(q, has);
    if (has) {
      q_store_state(q, s, n);
      q.owner[s] = lt * MCV_SPL + sel;
#pragma unroll
      for (int k = 0; k < Chain::NX; ++k) q.xt[k][s] = xa[k];
    }
    int n_new;
    bool can_a;
    round_totals(q, par, has, (~(busy | done) & ALL) != 0 || (!dry && (done & ~busy & ALL) != 0), n_new, can_a);
    cnt += n_new;
    if (cnt >= MCV_BLOCK || (!can_a && cnt > 0)) {
      const int take = cnt >= MCV_BLOCK ? MCV_BLOCK : cnt;
      if (lt < take) {
        Neutron b;
        const int slot = cnt - take + lt;
        q_load_state(q, slot, b);
        double xb[Chain::NX];
#pragma unroll
        for (int k = 0; k < Chain::NX; ++k) xb[k] = q.xt[k][slot];
        const int o = q.owner[slot];
        Rng r;
        r.s0 = q.st0[o];
        r.s1 = q.st1[o];
        Chain::phase_b(r, b, A, J, acc, xb);
        q.st0[o] = r.s0;
        q.st1[o] = r.s1;
        q.returned[o] = 1;
      }
      cnt -= take;
      __syncthreads();
#pragma unroll
      for (int k = 0; k < MCV_SPL; ++k) {
        const int o = lt * MCV_SPL + k;
        if (q.returned[o]) {
          q.returned[o] = 0;
          busy &= ~(1u << k);
        }
      }
      if (lt == 0) q.count = cnt;
      __syncthreads();
    } else if (!can_a) {
      break;
    }
  }
  count_processed(A.counter, mine);
#pragma unroll
  for (int k = 0; k < Chain::NS; ++k) {
    double v = acc[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0 && v != 0) unsafeAtomicAdd(J.grad + k, v);
  }
}

}  // namespace mcv

#define MCV_DEFINE_ADJOINT_KERNEL_COMPACTED(NAME, CHAIN)                                                \
  extern "C" __device__ int mcv_stream_pool = 1;                                                       \
  extern "C" __device__ int mcv_adjoint_nscore = CHAIN::NS;                                            \
  extern "C" __global__ void MCV_KERNEL_ATTR NAME(const mcv::LaunchArgs A, const mcv::AdjointArgs J) { \
    __shared__ mcv::QueueAdjLDS<CHAIN::NX> q;                                                          \
    const int64_t tid = A.tid_base + (int64_t)blockIdx.x * (blockDim.x * MCV_SPL) + threadIdx.x;      \
    mcv::run_block_adjoint_rng<CHAIN>(A, J, tid, q, A.pool,                                            \
                                      A.tid_base + (int64_t)gridDim.x * (MCV_BLOCK * MCV_SPL));        \
  }

#define MCV_DEFINE_ADJOINT_KERNEL(NAME, CHAIN)                                                          \
  extern "C" __device__ int mcv_adjoint_nscore = CHAIN::NS;                                            \
  extern "C" __global__ void MCV_KERNEL_ATTR NAME(const mcv::LaunchArgs A, const mcv::AdjointArgs J) { \
    const int64_t tid = A.tid_base + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;                   \
    mcv::run_lane_adjoint<CHAIN>(A, J, tid);                                                           \
  }
