// Device-side building blocks shared by the spc2 kernels (gfx950).
//
// Every translation unit that includes this header is compiled with
// -ffp-contract=off: the FP64 forms must round exactly like the reference
// (no fused a*b+c), so each FP64 expression keeps the reference's left-to-right
// evaluation order.  The FP32 forms use explicit fmaf where they want fusion.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spc2 {

// ---- geometry (PeriodicBoundaryConditions.cpp) -----------------------------

struct Box {
  double Lx, Ly;    // x_size_, y_size_
  double rLx, rLy;  // x_rsize_ = 1.0 / x_size (PeriodicBoundaryConditions.cpp:13-14)
};

// ClassAEffectiveParticleDistance, PeriodicBoundaryConditions.cpp:22-34:
// dx = x_j - x_i; dx -= int(dx * 2.0 / L) * L (truncation toward zero).
__device__ __forceinline__ double min_image(double d, double L, double rL) {
  return d - (double)(int)(d * 2.0 * rL) * L;
}

// Scalar wrap, PeriodicBoundaryConditions.cpp:117-121: x - floor(x / L) * L.
__device__ __forceinline__ double wrap(double x, double L, double rL) { return x - floor(x * rL) * L; }

// The conditional wrap applied before binning and before each read in the
// linked-list loops (ChimeraSystem.cpp:299-302, :402-405).
__device__ __forceinline__ void cond_wrap(const Box &b, double &x, double &y) {
  if (x < 0.0 || x >= b.Lx || y < 0.0 || y >= b.Ly) {
    x = wrap(x, b.Lx, b.rLx);
    y = wrap(y, b.Ly, b.rLy);
  }
}

// AdjustNeighboringCellToPeriodicBoundaries, ChimeraSystem.cpp:905-922.
__device__ __forceinline__ int adjust_cell(int c, int n) { return c < 0 ? n - 1 : (c >= n ? 0 : c); }

// Cell index of a (conditionally wrapped) position: int(x / L * n), divide
// then multiply (ChimeraSystem.cpp:304-306).  An index that rounds up to n
// (x a hair below L) is out of range in the reference; it is clamped here, and
// so is any non-finite coordinate, so that no cell index leaves [0, n).
__device__ __forceinline__ int cell_coord(double x, double L, int n) {
  double f = x / L * n;
  if (!(f >= 0.0)) return 0;        // negative zero-crossing artefacts and NaN
  if (f >= (double)n) return n - 1;
  return (int)f;
}

// ---- stable rank inside a cell ----------------------------------------------
// The counting sort places a workgroup's slots in cells of increasing index, so
// their cells span [start of the first slot's cell, end of the last slot's
// cell).  stage_cell_ids copies the placed ids of that span to LDS when it holds
// at most kRankStage of them; every thread of the workgroup calls it (lo, hi:
// the thread's cell; unused past n).  Returns the span's first slot, or -1 when
// the ids stay in global memory.
constexpr int kRankStage = 2048;
__device__ __forceinline__ int stage_cell_ids(int *ids, int *span, const int *tmpid, int n, int lo, int hi) {
  const int k0 = blockIdx.x * blockDim.x, k1 = min(n, k0 + (int)blockDim.x) - 1;
  const int k = k0 + (int)threadIdx.x;
  if (k == k0) span[0] = lo;
  if (k == k1) span[1] = hi;
  __syncthreads();
  const int s0 = span[0], s1 = span[1];
  const bool staged = s1 - s0 <= kRankStage;
  if (staged)
    for (int m = s0 + (int)threadIdx.x; m < s1; m += (int)blockDim.x) ids[m - s0] = tmpid[m];
  __syncthreads();
  return staged ? s0 : -1;
}

// The rank of id among the cnt ids at a: how many are smaller.
__device__ __forceinline__ int rank_in_cell(const int *a, int cnt, int id) {
  int r = 0;
  for (int m = 0; m < cnt; ++m) r += a[m] < id;
  return r;
}

// ---- Kahan-compensated FP32 sum ----------------------------------------------

struct Kahan {
  float s, c;  // running sum and (negated) low-order part
  __device__ __forceinline__ void init() { s = 0.0f; c = 0.0f; }
  __device__ __forceinline__ void add(float v) {
    float y = v - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  // add(a * b) with the product's rounding folded into the compensated term
  __device__ __forceinline__ void add_prod(float a, float b) {
    float y = __builtin_fmaf(a, b, -c);
    float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};

// ---- Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11) ---------------------

struct U4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
    U4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// ---- normal deviates from IEEE basic operations only -------------------------
//
// The transform uses +, -, *, /, sqrt, floor and frexp only (all correctly
// rounded or exact), in a fixed order, so the same bits come out of any IEEE
// double implementation: the device stream can be regenerated on a host.
//
// log(x), x in (0, 1]: x = m 2^e with m in [sqrt(1/2), sqrt(2));
// log m = 2 atanh(s), s = (m - 1) / (m + 1), odd series to s^23;
// log x = e ln2_hi + (e ln2_lo + log m)  (Cody-Waite split of ln 2).
__device__ __forceinline__ double log_basic(double x) {
  int e;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e = e - 1;
  }
  double s = (m - 1.0) / (m + 1.0);
  double z = s * s;
  double p = 0.043478260869565216;
  p = p * z + 0.047619047619047616;
  p = p * z + 0.052631578947368418;
  p = p * z + 0.058823529411764705;
  p = p * z + 0.066666666666666666;
  p = p * z + 0.076923076923076927;
  p = p * z + 0.090909090909090912;
  p = p * z + 0.11111111111111110;
  p = p * z + 0.14285714285714285;
  p = p * z + 0.20000000000000001;
  p = p * z + 0.33333333333333331;
  double logm = 2.0 * s + 2.0 * s * (z * p);
  double de = (double)e;
  return de * 6.93147180369123816490e-01 + (de * 1.90821492927058770002e-10 + logm);
}

// sin and cos of (pi/2) g, |g| <= 1/2, Taylor series to x^17 / x^18.
__device__ __forceinline__ void sincos_quarter(double g, double &sn, double &cs) {
  double x = g * 1.5707963267948966;
  double z = x * x;
  double ps = 2.8114572543455206e-15;
  ps = ps * z - 7.6471637318198164e-13;
  ps = ps * z + 1.6059043836821613e-10;
  ps = ps * z - 2.5052108385441720e-08;
  ps = ps * z + 2.7557319223985893e-06;
  ps = ps * z - 1.9841269841269841e-04;
  ps = ps * z + 8.3333333333333332e-03;
  ps = ps * z - 1.6666666666666666e-01;
  sn = x + x * (z * ps);
  double pc = 1.5619206968586225e-16;
  pc = pc * z - 4.7794773323873853e-14;
  pc = pc * z + 1.1470745597729725e-11;
  pc = pc * z - 2.0876756987868100e-09;
  pc = pc * z + 2.7557319223985888e-07;
  pc = pc * z - 2.4801587301587302e-05;
  pc = pc * z + 1.3888888888888889e-03;
  pc = pc * z - 4.1666666666666664e-02;
  pc = pc * z + 0.5;
  cs = 1.0 - z * pc;
}

// Box-Muller pair from one Philox block: u1 in (0, 1] from (w1:w0), u2 in
// [0, 1) from (w3:w2), r = sqrt(-2 log u1), angle 2 pi u2 reduced exactly
// by quadrant q = round(4 u2), g = 4 u2 - q.
__device__ __forceinline__ void box_muller(U4 w, double &z0, double &z1) {
  uint64_t a = ((uint64_t)w.y << 32) | w.x;
  uint64_t b = ((uint64_t)w.w << 32) | w.z;
  double u1 = (double)((a >> 11) + 1) * 0x1.0p-53;
  double u2 = (double)(b >> 11) * 0x1.0p-53;
  double r = sqrt(-2.0 * log_basic(u1));
  double f = u2 * 4.0;
  double q = floor(f + 0.5);
  double g = f - q;
  double sn, cs;
  sincos_quarter(g, sn, cs);
  int iq = ((int)q) & 3;
  double c, s;
  if (iq == 0) {
    c = cs;
    s = sn;
  } else if (iq == 1) {
    c = -sn;
    s = cs;
  } else if (iq == 2) {
    c = -cs;
    s = -sn;
  } else {
    c = sn;
    s = -cs;
  }
  z0 = r * c;
  z1 = r * s;
}

// Noise streams: counter (particle, stream, step_lo, step_hi), key (seed).
enum : uint32_t { kStreamStep = 0, kStreamInitPos = 1, kStreamInitPhase = 2, kStreamOmega = 3 };

__device__ __forceinline__ void normal_pair(uint64_t seed, uint64_t step, uint32_t stream, uint32_t particle,
                                            double &z0, double &z1) {
  U4 c = {particle, stream, (uint32_t)step, (uint32_t)(step >> 32)};
  U4 w = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  box_muller(w, z0, z1);
}

__device__ __forceinline__ void uniform_pair(uint64_t seed, uint64_t step, uint32_t stream, uint32_t particle,
                                             double &u0, double &u1) {
  U4 c = {particle, stream, (uint32_t)step, (uint32_t)(step >> 32)};
  U4 w = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  uint64_t a = ((uint64_t)w.y << 32) | w.x;
  uint64_t b = ((uint64_t)w.w << 32) | w.z;
  u0 = (double)(a >> 11) * 0x1.0p-53;
  u1 = (double)(b >> 11) * 0x1.0p-53;
}

// Histogram increment with one atomic per distinct cell per wavefront: lanes
// holding the same cell take consecutive arrival ranks from a single
// atomicAdd issued by the lowest such lane.  Every lane of the wavefront takes
// part (valid == false for lanes past the end).
__device__ __forceinline__ int wave_aggregated_rank(int *count, int c, bool valid) {
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  bool done = !valid;
  int rank = 0;
  while (true) {
    unsigned long long pending = __ballot(!done);
    if (pending == 0ull) break;
    const int leader = __builtin_ctzll(pending);
    const int lc = __shfl(c, leader, 64);
    const bool mine = !done && c == lc;
    const unsigned long long same = __ballot(mine);
    int base = 0;
    if (lane == leader) base = atomicAdd(&count[lc], __popcll(same));
    base = __shfl(base, leader, 64);
    if (mine) {
      rank = base + __popcll(same & below);
      done = true;
    }
  }
  return rank;
}

}  // namespace spc2
