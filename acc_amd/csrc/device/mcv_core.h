// mcv_core.h -- device primitives of the fused neutron ray-tracing kernels (gfx950).
//
// Neutron state lives in registers as ten doubles (the layout of
// acc/neutron.py: x, y, z, vx, vy, vz, s1, s2, t, p).  The random stream is
// numba's xoroshiro128+ (numba/cuda/random.py, numba 0.53.1) with the state
// held in four VGPRs instead of being re-read from global memory per draw.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcv {

// mcni.utils.conversion (McStas values); K2V = 1/V2K.
constexpr double V2K = 1.58825361e-3;
constexpr double SE2V = 437.393377;
constexpr double VS2E = 5.22703725e-6;
constexpr double K2V = 1.0 / 1.58825361e-3;
constexpr double PI = 3.141592653589793;

struct Neutron {
  double x, y, z, vx, vy, vz, s1, s2, t, p;
};
static_assert(sizeof(Neutron) == 10 * sizeof(double), "a neutron is one 10-double row");

// ------------------------------------------------------------------ RNG
struct Rng {
  uint64_t s0, s1;

  __device__ __forceinline__ uint64_t next() {
    const uint64_t a = s0, b0 = s1;
    const uint64_t result = a + b0;
    const uint64_t b = b0 ^ a;
    s0 = ((a << 55) | (a >> 9)) ^ b ^ (b << 14);
    s1 = (b << 36) | (b >> 28);
    return result;
  }
  // uint64_to_unit_float32: float32(float64(x >> 11) * 2^-53); may round to 1.0f
  __device__ __forceinline__ double uniform() {
    const uint64_t x = next() >> 11;
    const double d = (double)x * 0x1p-53;
    return (double)(float)d;
  }
};

// ------------------------------------------------------------------ frames
// acc/neutron.py:18-22 (abs2rel): r' = R (r - o); v' = R v.  T = {R[9], o[3]}.
__device__ __forceinline__ void apply_transformation(Neutron& n, const double* __restrict__ T) {
  const double rx = n.x - T[9], ry = n.y - T[10], rz = n.z - T[11];
  const double vx = n.vx, vy = n.vy, vz = n.vz;
  n.x = T[0] * rx + T[1] * ry + T[2] * rz;
  n.y = T[3] * rx + T[4] * ry + T[5] * rz;
  n.z = T[6] * rx + T[7] * ry + T[8] * rz;
  n.vx = T[0] * vx + T[1] * vy + T[2] * vz;
  n.vy = T[3] * vx + T[4] * vy + T[5] * vz;
  n.vz = T[6] * vx + T[7] * vy + T[8] * vz;
}

__device__ __forceinline__ void prop_dt_inplace(Neutron& n, double dt) {
  n.x = n.x + n.vx * dt;
  n.y = n.y + n.vy * dt;
  n.z = n.z + n.vz * dt;
  n.t = n.t + dt;
}

// ------------------------------------------------------------------ division by a known reciprocal
// a / b rounded to nearest, given y = 1 / b rounded to nearest (computed once,
// e.g. on the host, for a divisor that many divisions share).  Markstein's
// theorem (P. Markstein, IBM J. Res. Dev. 34, 1990; Muller et al., Handbook of
// Floating-Point Arithmetic, "Markstein's theorem"): if y is 1 / b rounded to
// nearest and q is within one ulp of a / b, then r = a - b q is exact (one
// FMA) and q + r y rounded to nearest is a / b rounded to nearest.  q0 = a y
// is within 1.5 ulp, so one correction brings it within ~0.5 ulp and the
// second applies the theorem.  Five FP64 instructions instead of the ~11 of a
// division; valid while a, b, a / b and the remainders stay in the normal
// range (the guides' operands are of order 1e-3 to 1e4).
__device__ __forceinline__ double div_recip(double a, double b, double y) {
  const double q0 = a * y;
  const double q1 = fma(fma(-b, q0, a), y, q0);
  return fma(fma(-b, q1, a), y, q1);
}

// ------------------------------------------------------------------ histograms
// Monitors own (3, nbins) float64 blocks: N, sum p, sum p^2, updated with
// device-scope FP64 atomic adds that return nothing (the wave does not wait
// for them).  With eight XCDs each keeping its own L2, a device-scope atomic
// is performed past L2: rocprofv3 counts one 32-byte memory write per add.
__device__ __forceinline__ void hist_add3(double* __restrict__ H, int64_t nbins, int64_t b, double p) {
  unsafeAtomicAdd(H + b, 1.0);
  unsafeAtomicAdd(H + nbins + b, p);
  unsafeAtomicAdd(H + 2 * nbins + b, p * p);
}

// ------------------------------------------------------------------ score tracking
// Components that carry differentiable weight factors take a tracker; the
// forward kernels pass NoTrack, whose hooks compile to nothing, so forward
// and adjoint kernels share one copy of every component's arithmetic (and the
// adjoint retraces exactly the forward paths).  The hooks:
//   mirror(axis, q, R0, Qc, alpha, m, W, dm)  weight *= R(q; ...) > 0, with
//                                             dm = d m / d(m parameter)
//   attenuate(path, inv_v)                    weight *= exp(-(mu_a(v) + sigma) path)
//   sigma()                                   weight *= sigma
//   dos(kp, E)                                weight *= Q^2 sum_k A_k exp(-u_k^2 / 2)
//   grid(i00, i10, i01, i11, fx, fy, S)       weight *= S (bilinear in the 4 nodes)
struct NoTrack {
  __device__ __forceinline__ void mirror(int, double, double, double, double, double, double, double) {}
  __device__ __forceinline__ void attenuate(double, double) {}
  __device__ __forceinline__ void sigma(double) {}
  __device__ __forceinline__ void dos(const double*, double) {}
  __device__ __forceinline__ void grid(int64_t, int64_t, int64_t, int64_t, double, double, double) {}
};

}  // namespace mcv
