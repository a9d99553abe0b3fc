// Single-system pair kernels on gfx950 (chimera and vortex arrays): the
// linked-cell kernels over the row-segment LDS staging of spc2_lds_stencil.h,
// the all-pairs path, and their dispatch.  Compiled with -ffp-contract=off.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "spc2_device.h"
#include "spc2_internal.h"
#include "spc2_lds_stencil.h"

namespace spc2 {

namespace {

constexpr int kBlock = 256;

inline int blocks_for(long long n, int b = kBlock) { return (int)((n + b - 1) / b); }

// =============================================================================
// Pass-1 and pass-2 kernels over linked cells
// =============================================================================

template <class Acc, int PREC, int EPI>
__global__ __launch_bounds__(kBlock) void k_pass1_lds(KParams P, const int *__restrict__ start, Tiles t,
                                                      const int *__restrict__ slot_h, const int *__restrict__ slot_id,
                                                      const double *gp, EpiCtx e, unsigned long long *counts) {
  extern __shared__ float4 dyn_lds[];
  pass1_segment<Acc, PREC, EPI>(P, start, t, slot_h, slot_id, gp, e, counts, xcd_segment(blockIdx.x, gridDim.x),
                                dyn_lds);
}

template <class Acc, int PREC, int EPI>
__global__ __launch_bounds__(kBlock) void k_pass2_lds(KParams P, const int *__restrict__ start, Tiles t,
                                                      const int *__restrict__ slot_h, const int *__restrict__ slot_id,
                                                      const double *gp, EpiCtx e) {
  extern __shared__ float4 dyn_lds[];
  pass2_segment<Acc, PREC, EPI>(P, start, t, slot_h, slot_id, gp, e, xcd_segment(blockIdx.x, gridDim.x), dyn_lds);
}

// =============================================================================
// All pairs (rho > L/4): thread per particle (the FP32 chimera pass 1: a group
// of allpairs_lanes lanes per particle), j = 0..N-1, stage positions not
// wrapped (ChimeraSystem.cpp:131-167, :184-254)
// =============================================================================

template <int PREC, int SYS, int EPI>
__global__ __launch_bounds__(kBlock) void k_pass1_allpairs(KParams P, const double *gx, const double *gy,
                                                           const double *gp, const float4 *tiles, EpiCtx e,
                                                           unsigned long long *counts) {
  constexpr int G = allpairs_lanes<PREC, SYS>();  // lanes per particle
  const int t = blockIdx.x * blockDim.x + threadIdx.x, i = t / G, g = t % G;
  const bool active = i < P.n, lead = active && g == 0;  // whole lane groups
  Pass1Out o{};
  if (active) {
    if constexpr (G > 1) allpairs_pass1_group(P, i, g, gx, gy, gp, tiles, o);
    else allpairs_pass1<PREC, SYS>(P, i, gx, gy, gp, o, tiles);
    if (lead) epilogue1<EPI>(P, e, i, i, i, o);  // all pairs: home order is the id order
  }
  if (counts) count_add(counts, lead ? P.n - 1 : 0, lead ? o.n : 0, lead ? o.n2 : 0);
}

template <int PREC, int SYS, int EPI>
__global__ __launch_bounds__(kBlock) void k_pass2_allpairs(KParams P, const double *gx, const double *gy,
                                                           const double *gp, EpiCtx e) {
  constexpr int G = allpairs_lanes<PREC, SYS>();  // lanes per particle
  const int t = blockIdx.x * blockDim.x + threadIdx.x, i = t / G, g = t % G;
  if (i >= P.n) return;  // whole lane groups
  Pass2Out o;
  const double in0 = EPI == EPI_ITO2 ? e.a0[i] : 0.0, in1 = EPI == EPI_ITO2 ? e.a1[i] : 0.0;
  if constexpr (G > 1)
    allpairs_pass2_group(P, i, g, gx, gy, gp, e.aphi_slot, e.n_slot, in0, in1, o);
  else
    allpairs_pass2<SYS>(P, i, gx, gy, gp, e.aphi_slot, e.n_slot, e.n2_slot, in0, in1, o);
  if (g == 0) epilogue2<EPI>(P, e, i, i, o);
}

// =============================================================================
// dispatch
// =============================================================================

template <int PREC, int PASS>
size_t lds_bytes(const Sim &s) {
  return (size_t)s.kp.lds_cap * LdsTile<PREC, PASS>::bytes;
}

template <int PREC, int SYS, int EPI>
void pass1_typed(Sim &s, const EpiCtx &e) {
  unsigned long long *cnt = s.counting ? s.counts : nullptr;
  if (s.linked) {
    const dim3 grid(s.ny * s.kp.segs), blk(kBlock);
    const size_t sh = lds_bytes<PREC, 1>(s);
    if (SYS == SPC2_CHIMERA)
      launch_pair(s, k_pass1_lds<ChimAcc<PREC>, PREC, EPI>, grid, blk, sh, s.kp, (const int *)s.start, s.t,
                  (const int *)s.slot_h, (const int *)s.slot_id, (const double *)s.gp, e, cnt);
    else
      launch_pair(s, k_pass1_lds<VortAcc<PREC>, PREC, EPI>, grid, blk, sh, s.kp, (const int *)s.start, s.t,
                  (const int *)s.slot_h, (const int *)s.slot_id, (const double *)s.gp, e, cnt);
  } else {
    // the FP32 form reads cos and sin of phi_j from the stage's tiles
    launch_pair(s, k_pass1_allpairs<PREC, SYS, EPI>, dim3(blocks_for((long long)s.n * allpairs_lanes<PREC, SYS>())),
                dim3(kBlock), 0, s.kp,
                (const double *)s.gx, (const double *)s.gy, (const double *)s.gp,
                PREC == P32K ? (const float4 *)s.t.f4 : nullptr, e, cnt);
  }
}

template <int PREC, int SYS>
void pass1_epi(Sim &s, int epi, const EpiCtx &e) {
  switch (epi) {
    case EPI_K: pass1_typed<PREC, SYS, EPI_K>(s, e); break;
    case EPI_ITO1: pass1_typed<PREC, SYS, EPI_ITO1>(s, e); break;
    case EPI_SE: pass1_typed<PREC, SYS, EPI_SE>(s, e); break;
    case EPI_RK4_FINAL: pass1_typed<PREC, SYS, EPI_RK4_FINAL>(s, e); break;
  }
}

template <int PREC>
void pass1_sys(Sim &s, int epi, const EpiCtx &e) {
  if (s.P.system == SPC2_CHIMERA)
    pass1_epi<PREC, SPC2_CHIMERA>(s, epi, e);
  else
    pass1_epi<PREC, SPC2_VORTEX>(s, epi, e);
}

template <int PREC, int SYS, int EPI>
void pass2_typed(Sim &s, const EpiCtx &e) {
  if (s.linked) {
    const size_t sh = lds_bytes<PREC, 2>(s);
    const dim3 grid(s.ny * s.kp.segs), blk(kBlock);
    if (SYS == SPC2_CHIMERA)
      launch_pair(s, k_pass2_lds<Chim2Acc<PREC>, PREC, EPI>, grid, blk, sh, s.kp, (const int *)s.start, s.t,
                  (const int *)s.slot_h, (const int *)s.slot_id, (const double *)s.gp, e);
    else
      launch_pair(s, k_pass2_lds<Vort2Acc<PREC>, PREC, EPI>, grid, blk, sh, s.kp, (const int *)s.start, s.t,
                  (const int *)s.slot_h, (const int *)s.slot_id, (const double *)s.gp, e);
  } else {
    launch_pair(s, k_pass2_allpairs<PREC, SYS, EPI>, dim3(blocks_for((long long)s.n * allpairs_lanes<PREC, SYS>())),
                dim3(kBlock), 0, s.kp, (const double *)s.gx,
                (const double *)s.gy, (const double *)s.gp, e);
  }
}

template <int PREC, int SYS>
void pass2_epi(Sim &s, int epi, const EpiCtx &e) {
  if (epi == EPI_ITO2)
    pass2_typed<PREC, SYS, EPI_ITO2>(s, e);
  else
    pass2_typed<PREC, SYS, EPI_SRK>(s, e);
}

template <int PREC>
void pass2_sys(Sim &s, int epi, const EpiCtx &e) {
  if (s.P.system == SPC2_CHIMERA)
    pass2_epi<PREC, SPC2_CHIMERA>(s, epi, e);
  else
    pass2_epi<PREC, SPC2_VORTEX>(s, epi, e);
}

}  // namespace

// The pair workgroups of a linked system: about `target` particles each and an
// LDS capacity for their halo.
// * Rows of at least kChunkPerCell particles per cell are cut into chunks of
//   `target` slots rounded to whole wavefronts, so that every wavefront but a
//   row's last is full.  A row gets workgroups for its mean count plus four
//   standard deviations; its last workgroup takes whatever remains.
// * Sparser rows are cut into segments of seg_w cells (at most kMaxSegW), whose
//   particle counts follow the cells.
// The LDS capacity is the mean halo with a margin, within 60 KB; a workgroup
// whose halo exceeds it reads the same tiles from global memory.
void configure_segments_kp(KParams &k, int prec, double target) {
  k.chunk = 0;
  if (!k.linked) {
    k.seg_w = k.segs = 1;
    k.lds_cap = 0;
    return;
  }
  const double per_cell = (double)k.n / ((double)k.nx * (double)k.ny);
  double halo;
  if (per_cell >= kChunkPerCell) {
    int chunk = 64 * std::max(1, (int)std::lround(target / 64.0));
    // at most kMaxSegW - 3 cells' worth, so a chunk's columns fit the halo tables
    chunk = std::min(chunk, 64 * std::max(1, (int)std::floor((kMaxSegW - 3) * per_cell / 64.0)));
    const double row = (double)k.n / (double)k.ny;
    k.chunk = chunk;
    k.segs = std::max(1, (int)std::ceil((row + 4.0 * std::sqrt(row) + 8.0) / chunk));
    k.seg_w = std::min(k.nx, (int)std::ceil(chunk / per_cell) + 1);
    halo = 1.25 * 3.0 * (chunk + 3.0 * per_cell);  // chunk's cells, a partial cell at each end, and the sides
  } else {
    int w = (int)std::floor(target / std::max(per_cell, 1e-9));
    w = std::max(1, std::min({w, kMaxSegW, k.nx}));
    k.seg_w = w;
    k.segs = (k.nx + w - 1) / w;
    halo = 1.6 * 3.0 * (w + 2) * per_cell;
  }
  int cap = (int)std::ceil((halo + 160.0) / 64.0) * 64;
  cap = std::min(cap, kLdsBytes / lds_elem_bytes(prec));
  if (const char *env = std::getenv("SPC2_LDS_CAP")) cap = std::max(0, std::min(cap, std::atoi(env)));
  k.lds_cap = cap;
}

int lds_elem_bytes(int prec) {
  switch (prec) {
    case P32K: return LdsTile<P32K, 2>::bytes;
    case P64: return LdsTile<P64, 2>::bytes;
    default: return LdsTile<P64F, 2>::bytes;
  }
}

void configure_segments(Sim &s) { configure_segments_kp(s.kp, s.prec, kSegTarget); }

void launch_pass1(Sim &s, int epi, double *ox, double *oy, double *op, const StepArgs &sa, double *const k1[3],
                  double *const k2[3], double *const k3[3]) {
  EpiCtx e = make_ctx(s, ox, oy, op, sa, k1, k2, k3);
  switch (s.prec) {
    case P64: pass1_sys<P64>(s, epi, e); break;
    case P64F: pass1_sys<P64F>(s, epi, e); break;
    default: pass1_sys<P32K>(s, epi, e);
  }
}

void launch_pass2(Sim &s, int epi, double *a0, double *a1, const StepArgs &sa) {
  EpiCtx e = make_ctx(s, nullptr, nullptr, nullptr, sa, nullptr, nullptr, nullptr);
  e.a0 = a0;
  e.a1 = a1;
  switch (s.prec) {
    case P64: pass2_sys<P64>(s, epi, e); break;
    case P64F: pass2_sys<P64F>(s, epi, e); break;
    default: pass2_sys<P32K>(s, epi, e);
  }
}

}  // namespace spc2
