// Order-parameter reduction, Philox-keyed initial state, materialised noise
// and a fill kernel (graph-safe replacement for hipMemsetAsync).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>

#include "spc2_device.h"
#include "spc2_internal.h"

namespace spc2 {

namespace {

constexpr int kBlock = 256;
constexpr int kRedBlocks = 1024;
constexpr double kTwoPi = 6.283185307179586;  // 2.0 * M_PI

inline int blocks_for(long long n, int b = kBlock) { return (int)((n + b - 1) / b); }

// Kuramoto order parameter, two-level FP64 reduction (deterministic: fixed
// partition and fixed tree); replaces the float sum of
// BinaryObserverPtr::SaveSummaryStatistics (BinaryObserverPtr.cpp:237-247).
// Partial sums over particle ids (the phase read through the id-to-home map),
// so that the result does not depend on where re-homing has put the particles.
__global__ __launch_bounds__(kBlock) void k_order_partial(int n, const double *sp, const int *hpos, double *red) {
  __shared__ double sc[kBlock], ss[kBlock];
  double c = 0.0, s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double sn, cs;
    sincos(sp[hpos[i]], &sn, &cs);
    c += cs;
    s += sn;
  }
  sc[threadIdx.x] = c;
  ss[threadIdx.x] = s;
  __syncthreads();
  for (int d = kBlock / 2; d > 0; d >>= 1) {
    if ((int)threadIdx.x < d) {
      sc[threadIdx.x] += sc[threadIdx.x + d];
      ss[threadIdx.x] += ss[threadIdx.x + d];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    red[2 * blockIdx.x] = sc[0];
    red[2 * blockIdx.x + 1] = ss[0];
  }
}

__global__ __launch_bounds__(kBlock) void k_order_final(int nb, const double *red, double *out2) {
  __shared__ double sc[kBlock], ss[kBlock];
  double c = 0.0, s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    c += red[2 * b];
    s += red[2 * b + 1];
  }
  sc[threadIdx.x] = c;
  ss[threadIdx.x] = s;
  __syncthreads();
  for (int d = kBlock / 2; d > 0; d >>= 1) {
    if ((int)threadIdx.x < d) {
      sc[threadIdx.x] += sc[threadIdx.x + d];
      ss[threadIdx.x] += ss[threadIdx.x + d];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out2[0] = sc[0];
    out2[1] = ss[0];
  }
}

// x, y ~ U[0, L), phi ~ U[0, 2 pi), omega ~ N(0, omega_std) from Philox
// streams 1, 2, 3 at step 0 (replaces SimulationEnginePtr.cpp:31-50).
__global__ __launch_bounds__(kBlock) void k_init_state(int n, double Lx, double Ly, uint64_t seed, double omega_std,
                                                       const int *hid, double *sx, double *sy, double *sp,
                                                       double *omega) {
  int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= n) return;
  const uint32_t i = (uint32_t)hid[h];
  double ux, uy, up, unused;
  uniform_pair(seed, 0, kStreamInitPos, i, ux, uy);
  uniform_pair(seed, 0, kStreamInitPhase, i, up, unused);
  sx[h] = ux * Lx;
  sy[h] = uy * Ly;
  sp[h] = up * kTwoPi;
  double z0, z1;
  normal_pair(seed, 0, kStreamOmega, i, z0, z1);
  omega[h] = z0 * omega_std;
}

// id-ordered host arrays <-> home-ordered device arrays
__global__ __launch_bounds__(kBlock) void k_to_home(int n, const int *hpos, const double *src_id, double *dst_home) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id < n) dst_home[hpos[id]] = src_id[id];
}

__global__ __launch_bounds__(kBlock) void k_from_home(int n, const int *hpos, const double *src_home, double *dst_id) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id < n) dst_id[id] = src_home[hpos[id]];
}

// Per-slot values to particle-id order through a slot -> id map.
__global__ __launch_bounds__(kBlock) void k_slot_to_id(int n, const int *id_of_slot, const double *src_slot,
                                                       double *dst_id) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst_id[id_of_slot[k]] = src_slot[k];
}

__global__ __launch_bounds__(kBlock) void k_from_home_int(int n, const int *hpos, const int *src_home, int *dst_id) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id < n) dst_id[id] = src_home[hpos[id]];
}

// Trajectory frame payload: float AoS [x, y, phi] by particle id
// (BinaryObserverPtr.cpp:163-168 writes the float copy of the double state).
__global__ __launch_bounds__(kBlock) void k_snapshot(int n, const int *hpos, const double *sx, const double *sy,
                                                     const double *sp, float *out) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  const int h = hpos[id];
  out[3 * (size_t)id] = (float)sx[h];
  out[3 * (size_t)id + 1] = (float)sy[h];
  out[3 * (size_t)id + 2] = (float)sp[h];
}

__global__ __launch_bounds__(kBlock) void k_iota(int n, int *a, int *b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    a[i] = i;
    b[i] = i;
  }
}

// Re-homing: the particle in cell-sorted slot k moves to home k.
struct PermuteSet {
  const double *src[kPermuteArrays];
  double *dst[kPermuteArrays];
};

// The first `count` arrays of the set move (the rest are dead at the call).
__global__ __launch_bounds__(kBlock) void k_permute(int n, const int *slot_h, PermuteSet ps, int count,
                                                    const int *hid_in, int *hid_out) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int h = slot_h[k];
#pragma unroll
  for (int a = 0; a < kPermuteArrays; ++a)
    if (a < count) ps.dst[a][k] = ps.src[a][h];
  hid_out[k] = hid_in[h];
}

// hpos[hid[h]] = h: the home of every particle id.
__global__ __launch_bounds__(kBlock) void k_hpos(int n, const int *hid, int *hpos) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h < n) hpos[hid[h]] = h;
}

__global__ __launch_bounds__(kBlock) void k_noise(int n, int per_particle, uint64_t seed, uint64_t step,
                                                  double *out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double z0, z1;
  normal_pair(seed, step, kStreamStep, (uint32_t)i, z0, z1);
  if (per_particle == 1) {
    out[i] = z0;
  } else {
    out[2 * i] = z0;
    out[2 * i + 1] = z1;
  }
}

__global__ __launch_bounds__(kBlock) void k_fill(double *p, double v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace


void launch_init_state(Sim &s, double omega_std) {
  hipLaunchKernelGGL(k_init_state, dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream, s.n, s.P.Lx, s.P.Ly, s.P.seed,
                     omega_std, s.hid, s.sx, s.sy, s.sp, s.omega);
}

namespace {
void ensure_hpos(Sim &s) {
  if (!s.hpos_stale) return;
  hipLaunchKernelGGL(k_hpos, dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream, s.n, (const int *)s.hid, s.hpos);
  s.hpos_stale = false;
}
}  // namespace

void launch_order_param(Sim &s, double *out2) {
  ensure_hpos(s);
  int nb = std::min(kRedBlocks, std::max(1, blocks_for(s.n)));
  hipLaunchKernelGGL(k_order_partial, dim3(nb), dim3(kBlock), 0, s.stream, s.n, (const double *)s.sp,
                     (const int *)s.hpos, s.red);
  hipLaunchKernelGGL(k_order_final, dim3(1), dim3(kBlock), 0, s.stream, nb, (const double *)s.red, out2);
}

void launch_to_home(Sim &s, const double *src_id, double *dst_home) {
  ensure_hpos(s);
  hipLaunchKernelGGL(k_to_home, dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream, s.n, s.hpos, src_id, dst_home);
}

void launch_from_home(Sim &s, const double *src_home, double *dst_id) {
  ensure_hpos(s);
  hipLaunchKernelGGL(k_from_home, dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream, s.n, s.hpos, src_home, dst_id);
}

void launch_slot_to_id(Sim &s, const int *id_of_slot, const double *src_slot, double *dst_id) {
  hipLaunchKernelGGL(k_slot_to_id, dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream, s.n, id_of_slot, src_slot, dst_id);
}

void launch_from_home_int(Sim &s, const int *src_home, int *dst_id) {
  ensure_hpos(s);
  hipLaunchKernelGGL(k_from_home_int, dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream, s.n, s.hpos, src_home,
                     dst_id);
}

void launch_snapshot(Sim &s, float *out) {
  ensure_hpos(s);
  hipLaunchKernelGGL(k_snapshot, dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream, s.n, s.hpos, s.sx, s.sy, s.sp,
                     out);
}

void launch_identity_homes(Sim &s) {
  s.hpos_stale = false;
  hipLaunchKernelGGL(k_iota, dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream, s.n, s.hid, s.hpos);
}

// Moves every persistent per-particle array to the cell order of the last
// build (slot k becomes home k) and swaps in the permuted buffers.
void launch_rehome(Sim &s, bool k_xy_live) {
  // state, omega and the phi components of k_1..k_4 (in/out across steps: a
  // particle without partners keeps its previous phi derivative), then the
  // x / y components, which a step always writes before reading
  double **arrs[kPermuteArrays] = {&s.sx, &s.sy, &s.sp, &s.omega, &s.k[0][2], &s.k[1][2], &s.k[2][2], &s.k[3][2],
                                   &s.k[0][0], &s.k[0][1], &s.k[1][0], &s.k[1][1], &s.k[2][0], &s.k[2][1],
                                   &s.k[3][0], &s.k[3][1]};
  const int count = k_xy_live ? kPermuteArrays : 8;
  PermuteSet ps;
  for (int a = 0; a < kPermuteArrays; ++a) {
    ps.src[a] = *arrs[a];
    ps.dst[a] = s.spare[a];
  }
  hipLaunchKernelGGL(k_permute, dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream, s.n, s.slot_h, ps, count, s.hid,
                     s.spare_i);
  for (int a = 0; a < count; ++a) std::swap(*arrs[a], s.spare[a]);
  std::swap(s.hid, s.spare_i);
  s.hpos_stale = true;  // a scatter by id, deferred to the next transfer by id
}

void launch_noise(Sim &s, int per_particle, uint64_t step, double *out) {
  hipLaunchKernelGGL(k_noise, dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream, s.n, per_particle, s.P.seed, step,
                     out);
}

void launch_fill(Sim &s, double *p, double v, size_t n) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s.stream, p, v, n);
}

}  // namespace spc2
