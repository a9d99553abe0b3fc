// Linked-cell build on gfx950: stage + binning + histogram, a three-phase
// wavefront scan of the histogram, and a stable counting sort that writes the
// cell-sorted tiles the pair kernels read.  Compiled with -ffp-contract=off.
//
// Particle data live in "home" order h (a permutation of the particle ids kept
// close to cell order, see spc2_api.hip); hid[h] is the particle id at home h.
// The cell list itself is the reference's: cells in row-major order, members
// in ascending particle id.
#include <hip/hip_runtime.h>

#include "spc2_device.h"
#include "spc2_internal.h"

namespace spc2 {

namespace {

constexpr int kBlock = 256;
constexpr int kScanItems = 4;                   // counts per thread
constexpr int kScanTile = kBlock * kScanItems;  // counts per workgroup

inline int blocks_for(long long n, int b = kBlock) { return (int)((n + b - 1) / b); }

// K1: stage s + (k_coef * dt) * k (ChimeraSystem.cpp:274-276, evaluated as
// (k_coef * dt) * k_prev like the reference); on the linked-cell path the
// position is wrapped if outside the box and binned (ChimeraSystem.cpp:291-309).
// `local` is the arrival rank inside the cell, used only to place the particle
// before the stable re-ranking by particle id.  apf4 (the FP32 form on the
// all-pairs path): each particle's tile, whose cos and sin of the stage phase
// the pair sums read instead of an FP64 sincos per pair (the same values).
__global__ __launch_bounds__(kBlock) void k_stage_bin(int n, int linked, Box b, int nx, int ny, const double *sx,
                                                      const double *sy, const double *sp, const double *kx,
                                                      const double *ky, const double *kp, double kcdt, double *gx,
                                                      double *gy, double *gp, int *cell_of, int *count, int *local,
                                                      float4 *apf4) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = h < n;
  int c = 0;
  if (valid) {
    double x = sx[h], y = sy[h], p = sp[h];
    if (kx) {
      x = x + kcdt * kx[h];
      y = y + kcdt * ky[h];
      p = p + kcdt * kp[h];
    }
    if (linked) {
      cond_wrap(b, x, y);
      c = cell_coord(y, b.Ly, ny) * nx + cell_coord(x, b.Lx, nx);
      cell_of[h] = c;
    }
    gx[h] = x;
    gy[h] = y;
    gp[h] = p;
    if (apf4) {
      double sn, cs;
      sincos(p, &sn, &cs);
      apf4[h] = make_float4((float)x, (float)y, (float)cs, (float)sn);
    }
  }
  if (linked) {
    int r = wave_aggregated_rank(count, c, valid);
    if (valid) local[h] = r;
  }
}

// Inclusive scan of one int per thread across a 256-thread workgroup:
// shuffles inside each 64-lane wavefront, wavefront totals through LDS.
__device__ __forceinline__ int block_inclusive_scan(int v, int *wtot) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  if (lane == 63) wtot[w] = v;
  __syncthreads();
  int add = 0;
#pragma unroll
  for (int q = 0; q < kBlock / 64; ++q) add += (q < w) ? wtot[q] : 0;
  __syncthreads();
  return v + add;
}

// K2a: per-workgroup totals of kScanTile consecutive counts.
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const int *count, int nc, int *bsum) {
  __shared__ int wtot[kBlock / 64];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int s = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) s += (base + q < nc) ? count[base + q] : 0;
  int incl = block_inclusive_scan(s, wtot);
  if (threadIdx.x == kBlock - 1) bsum[blockIdx.x] = incl;
}

// K2b: exclusive scan of the workgroup totals (one workgroup, carried tiles);
// writes the grand total to start[nc].
__global__ __launch_bounds__(kBlock) void k_scan_bsum(int *bsum, int nb, int *start, int nc) {
  __shared__ int wtot[kBlock / 64];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += kBlock) {
    int i = base + threadIdx.x;
    int v = i < nb ? bsum[i] : 0;
    int incl = block_inclusive_scan(v, wtot);
    int c0 = carry;
    if (i < nb) bsum[i] = c0 + incl - v;
    __syncthreads();
    if (threadIdx.x == kBlock - 1) carry = c0 + incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) start[nc] = carry;
}

// K2c: exclusive scan inside each tile plus the tile offset -> start[];
// clears the histogram for the next build.
__global__ __launch_bounds__(kBlock) void k_scan_apply(int *count, int nc, const int *bsum, int *start) {
  __shared__ int wtot[kBlock / 64];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    v[q] = (base + q < nc) ? count[base + q] : 0;
    s += v[q];
  }
  int run = block_inclusive_scan(s, wtot) - s + bsum[blockIdx.x];
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    if (base + q < nc) {
      start[base + q] = run;
      count[base + q] = 0;
    }
    run += v[q];
  }
}

// K2 for at most kScanSingle counts: the whole exclusive scan in one
// workgroup, tile by tile with a carried total (the three-launch scan's values),
// writing start[nc] and clearing the histogram.
constexpr int kScanSingle = 4 * kScanTile;
__global__ __launch_bounds__(kBlock) void k_scan_single(int *count, int nc, int *start) {
  __shared__ int wtot[kBlock / 64];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int tile = 0; tile < nc; tile += kScanTile) {
    const int base = tile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
      v[q] = (base + q < nc) ? count[base + q] : 0;
      s += v[q];
    }
    int run = carry + block_inclusive_scan(s, wtot) - s;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
      if (base + q < nc) {
        start[base + q] = run;
        count[base + q] = 0;
      }
      run += v[q];
    }
    __syncthreads();  // every lane has read carry
    if (threadIdx.x == kBlock - 1) carry = run;
    __syncthreads();
  }
  if (threadIdx.x == 0) start[nc] = carry;
}

// K3a: place each particle at cell start + arrival rank (home and id).
__global__ __launch_bounds__(kBlock) void k_place(int n, const int *cell_of, const int *local, const int *start,
                                                  const int *hid, int *tmp, int *tmpid) {
  int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= n) return;
  int pos = start[cell_of[h]] + local[h];
  tmp[pos] = h;
  tmpid[pos] = hid[h];
}

// K3b: stable counting sort.  The final position of a particle inside its cell
// is the number of cell members with a smaller particle id, which reproduces
// the reference's ascending push_back order (ChimeraSystem.cpp:335-338).  The
// same thread gathers the particle's stage data into the cell-sorted tiles.
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_rank_tiles(int n, int nx, double hx, double hy, const int *tmp,
                                                       const int *tmpid, const int *cell_of, const int *start,
                                                       const double *gx, const double *gy, const double *gp,
                                                       int *slot_h, int *slot_id, Tiles t) {
  __shared__ int ids[kRankStage];
  __shared__ int span[2];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  int h = 0, id = 0, c = 0, lo = 0, hi = 0;
  if (k < n) {
    h = tmp[k];
    id = tmpid[k];
    c = cell_of[h];
    lo = start[c];
    hi = start[c + 1];
  }
  const int s0 = stage_cell_ids(ids, span, tmpid, n, lo, hi);
  if (k >= n) return;
  const int r = s0 >= 0 ? rank_in_cell(ids + (lo - s0), hi - lo, id) : rank_in_cell(tmpid + lo, hi - lo, id);
  const int dst = lo + r;
  slot_h[dst] = h;
  slot_id[dst] = id;
  t.cell[dst] = c;
  double x = gx[h], y = gy[h], p = gp[h];
  if (PREC == P32K) {
    int cx = c % nx, cy = c / nx;
    double sn, cs;
    sincos(p, &sn, &cs);
    t.f4[dst] = make_float4((float)(x - (double)cx * hx), (float)(y - (double)cy * hy), (float)cs, (float)sn);
    t.x[dst] = x;
    t.y[dst] = y;
  } else {
    t.x[dst] = x;
    t.y[dst] = y;
    t.p[dst] = p;
    if (PREC == P64F) {
      double sn, cs;
      sincos(p, &sn, &cs);
      t.c[dst] = cs;
      t.s[dst] = sn;
    }
  }
}

}  // namespace

void launch_stage_bin(Sim &s, const double *kx, const double *ky, const double *kp, double kcdt) {
  Box b{s.kp.Lx, s.kp.Ly, s.kp.rLx, s.kp.rLy};
  hipLaunchKernelGGL(k_stage_bin, dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream, s.n, (int)s.linked, b, s.nx,
                     s.ny, s.sx, s.sy, s.sp, kx, ky, kp, kcdt, s.gx, s.gy, s.gp, s.cell_of, s.count, s.local,
                     !s.linked && s.prec == P32K ? s.t.f4 : nullptr);
}

int scan_blocks(int nc) { return (nc + kScanTile - 1) / kScanTile; }

// Exclusive scan of count[0..m) into out[0..m], out[m] = total; clears count.
// bsum holds scan_blocks(m) + 1 ints.
void launch_scan_raw(hipStream_t st, int *bsum, int *count, int m, int *out) {
  if (m <= kScanSingle) {  // small grids: one launch instead of three
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kBlock), 0, st, count, m, out);
    return;
  }
  const int nb = scan_blocks(m);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kBlock), 0, st, count, m, bsum);
  hipLaunchKernelGGL(k_scan_bsum, dim3(1), dim3(kBlock), 0, st, bsum, nb, out, m);
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kBlock), 0, st, count, m, bsum, out);
}

void launch_scan_array(Sim &s, int *count, int m, int *out) { launch_scan_raw(s.stream, s.bsum, count, m, out); }

void launch_scan(Sim &s) { launch_scan_array(s, s.count, s.nc, s.start); }

void launch_place_raw(hipStream_t st, int n, const int *cell_of, const int *local, const int *start, const int *hid,
                      int *tmp, int *tmpid) {
  hipLaunchKernelGGL(k_place, dim3(blocks_for(n)), dim3(kBlock), 0, st, n, cell_of, local, start, hid, tmp, tmpid);
}

void launch_place(Sim &s) { launch_place_raw(s.stream, s.n, s.cell_of, s.local, s.start, s.hid, s.tmp, s.tmpid); }

void launch_rank_tiles(Sim &s) {
  dim3 g(blocks_for(s.n)), b(kBlock);
  switch (s.prec) {
    case P64:
      hipLaunchKernelGGL(k_rank_tiles<P64>, g, b, 0, s.stream, s.n, s.nx, s.kp.hx, s.kp.hy, s.tmp, s.tmpid,
                         s.cell_of, s.start, s.gx, s.gy, s.gp, s.slot_h, s.slot_id, s.t);
      break;
    case P64F:
      hipLaunchKernelGGL(k_rank_tiles<P64F>, g, b, 0, s.stream, s.n, s.nx, s.kp.hx, s.kp.hy, s.tmp, s.tmpid,
                         s.cell_of, s.start, s.gx, s.gy, s.gp, s.slot_h, s.slot_id, s.t);
      break;
    default:
      hipLaunchKernelGGL(k_rank_tiles<P32K>, g, b, 0, s.stream, s.n, s.nx, s.kp.hx, s.kp.hy, s.tmp, s.tmpid,
                         s.cell_of, s.start, s.gx, s.gy, s.gp, s.slot_h, s.slot_id, s.t);
  }
}

}  // namespace spc2
