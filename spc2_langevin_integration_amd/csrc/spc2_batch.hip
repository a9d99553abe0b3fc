// Batched replicas: many independent systems of uneven N (parameter sweeps,
// serving requests) stepped together, one launch per phase per step for the
// whole batch.
//
// The replicas' particles are concatenated: replica r owns homes (and global
// ids) poff[r] .. poff[r] + n[r] - 1, its cells coff[r] .. in one global cell
// array, and its linked-cell row segments soff[r] .. in one global segment
// numbering.  A step bins every particle into its replica's cells (one
// histogram), scans all cells at once, counting-sorts all particles, then one
// pair launch per pass covers the all-pairs replicas' blocks, first, and every
// row segment of every linked replica (a block looks its replica up in a
// table).  The
// per-replica arithmetic, summation order and noise keys (replica seed, the
// replica's own step, particle id within the replica) are those of a single
// handle, so a replica's trajectory is bit-identical alone or in any batch.
//
// Methods: the Stochastic Euler step (StochasticEulerStepperPtr.hpp:32-76) and
// the stochastic Runge-Kutta step (StochasticRungeKuttaStepperPtr.hpp:32-97:
// pass 1 with the Ito intermediates, pass 2 with the update) of chimera and
// vortex-arrays replicas.  Compiled with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <vector>

#include "spc2_lds_stencil.h"

using namespace spc2;

namespace spc2 {
namespace {

constexpr int kB = 256;
constexpr int kRehomeEvery = 16;

inline int bfor(long long n) { return (int)((n + kB - 1) / kB); }

// One replica's entry in the device table.
struct Rep {
  KParams kp;           // geometry and model; kp.linked = 0 on the all-pairs path
  int system;           // spc2_system
  int poff, n;          // first home / global id, particles
  int coff;             // first cell (all-pairs replicas own one cell)
  int soff, nseg;       // first row segment, segments (linked replicas)
  int apoff, apblocks;  // first block and blocks of the all-pairs launch
  uint64_t seed;
  int64_t step;         // the replica's own step counter at the start of the call
  int64_t steps;        // steps the replica takes in the current call
  double dt, sqdt, sq2D, D, h1;  // h1 = dt^1.5 / 2 (the SRK step's delta_Z)
};

// K1: stage (the state itself for SE) and binning into the replica's cells.
__global__ __launch_bounds__(kB) void k_b_stage_bin(int n, const Rep *reps, const int *rep_of, const double *sx,
                                                    const double *sy, const double *sp, double *gx, double *gy,
                                                    double *gp, int *cell_of, int *count, int *local) {
  const int h = blockIdx.x * kB + threadIdx.x;
  const bool valid = h < n;
  int c = 0;
  if (valid) {
    const Rep &R = reps[rep_of[h]];
    double x = sx[h], y = sy[h];
    int lc = 0;
    if (R.kp.linked) {
      Box b{R.kp.Lx, R.kp.Ly, R.kp.rLx, R.kp.rLy};
      cond_wrap(b, x, y);
      lc = cell_coord(y, b.Ly, R.kp.ny) * R.kp.nx + cell_coord(x, b.Lx, R.kp.nx);
    }
    c = R.coff + lc;
    cell_of[h] = c;
    gx[h] = x;
    gy[h] = y;
    gp[h] = sp[h];
  }
  const int r = wave_aggregated_rank(count, c, valid);
  if (valid) local[h] = r;
}

// K3b: stable counting sort (rank by particle id inside the cell) and the
// tiles; slot_id holds the id within the replica (the noise key) and t.cell
// the cell within the replica.
template <int PREC>
__global__ __launch_bounds__(kB) void k_b_rank_tiles(int n, const Rep *reps, const int *rep_of, const int *tmp,
                                                     const int *tmpid, const int *cell_of, const int *start,
                                                     const double *gx, const double *gy, const double *gp,
                                                     int *slot_h, int *slot_id, Tiles t) {
  __shared__ int ids[kRankStage];
  __shared__ int span[2];
  const int k = blockIdx.x * kB + threadIdx.x;
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
  const Rep &R = reps[rep_of[h]];
  const int lc = c - R.coff;
  slot_h[dst] = h;
  slot_id[dst] = id - R.poff;
  t.cell[dst] = lc;
  const double x = gx[h], y = gy[h], p = gp[h];
  t.x[dst] = x;
  t.y[dst] = y;
  if (PREC == P32K) {
    const int cx = lc % R.kp.nx, cy = lc / R.kp.nx;
    double sn, cs;
    sincos(p, &sn, &cs);
    t.f4[dst] = make_float4((float)(x - (double)cx * R.kp.hx), (float)(y - (double)cy * R.kp.hy), (float)cs,
                            (float)sn);
  } else {
    t.p[dst] = p;
    if (PREC == P64F) {
      double sn, cs;
      sincos(p, &sn, &cs);
      t.c[dst] = cs;
      t.s[dst] = sn;
    }
  }
}

__device__ __forceinline__ EpiCtx rep_ctx(const EpiCtx &e, const Rep &R, int k) {
  EpiCtx c = e;
  c.sa.dt = R.dt;
  c.sa.sqdt = R.sqdt;
  c.sa.sq2D = R.sq2D;
  c.sa.seed = R.seed;
  c.sa.step = (uint64_t)(R.step + k);
  c.sa.omega = R.system == SPC2_CHIMERA ? e.sa.omega : nullptr;  // no natural frequency in the vortex drift
  c.sa.D = R.D;
  c.sa.h1 = R.h1;
  c.bcoff = R.coff;
  return c;
}

// All-pairs replicas (rho > L/4): a thread per particle (the FP32 chimera
// pass 1: a group of allpairs_lanes lanes per particle); their homes are in id
// order (one cell, sorted by id), so a particle's slot is its home.  Block ab
// of the all-pairs blocks.
template <int PREC, int EPI>
__device__ __forceinline__ void b_allpairs1(const Rep *reps, const int *ap_rep, int ab, const double *gx,
                                            const double *gy, const double *gp, const float4 *f4, const EpiCtx &e,
                                            int k) {
  const Rep &R = reps[ap_rep[ab]];
  if (k >= R.steps) return;
  const int G = allpairs_lanes(PREC, R.system);  // lanes per particle, as a single handle's
  const int t = (ab - R.apoff) * kB + threadIdx.x, i = t / G, g = t % G;
  if (i >= R.n) return;  // whole lane groups
  const KParams P = R.kp;
  const EpiCtx c = rep_ctx(e, R, k);
  Pass1Out o{};
  if (R.system == SPC2_CHIMERA) {  // the replica's one cell holds its particles in id order: tile j is id j
    if constexpr (PREC == P32K)
      allpairs_pass1_group(P, i, g, gx + R.poff, gy + R.poff, gp + R.poff, f4 + R.poff, o);
    else
      allpairs_pass1<PREC, SPC2_CHIMERA>(P, i, gx + R.poff, gy + R.poff, gp + R.poff, o);
  } else {
    allpairs_pass1<PREC, SPC2_VORTEX>(P, i, gx + R.poff, gy + R.poff, gp + R.poff, o);
  }
  if (g == 0) epilogue1<EPI>(P, c, R.poff + i, i, R.poff + i, o);
}

template <int PREC>
__device__ __forceinline__ void b_allpairs2_srk(const Rep *reps, const int *ap_rep, int ab, const double *gx,
                                                const double *gy, const double *gp, const EpiCtx &e, int k) {
  const Rep &R = reps[ap_rep[ab]];
  if (k >= R.steps) return;
  const int G = allpairs_lanes(PREC, R.system);  // lanes per particle, as a single handle's
  const int t = (ab - R.apoff) * kB + threadIdx.x, i = t / G, g = t % G;
  if (i >= R.n) return;  // whole lane groups
  const KParams P = R.kp;
  const EpiCtx c = rep_ctx(e, R, k);
  const int o0 = R.poff;
  Pass2Out o;
  if (R.system == SPC2_CHIMERA) {
    if constexpr (PREC == P32K)
      allpairs_pass2_group(P, i, g, gx + o0, gy + o0, gp + o0, e.aphi_slot + o0, e.n_slot + o0, 0.0, 0.0, o);
    else
      allpairs_pass2<SPC2_CHIMERA>(P, i, gx + o0, gy + o0, gp + o0, e.aphi_slot + o0, e.n_slot + o0,
                                   e.n2_slot + o0, 0.0, 0.0, o);
  } else {
    allpairs_pass2<SPC2_VORTEX>(P, i, gx + o0, gy + o0, gp + o0, e.aphi_slot + o0, e.n_slot + o0, e.n2_slot + o0,
                                0.0, 0.0, o);
  }
  if (g == 0) epilogue2<EPI_SRK>(P, c, o0 + i, i, o);
}

// The all-pairs blocks lead the pass launches: the first ap.pad blocks (ap.n
// of them used, ap.pad a multiple of 8 so that the row segments keep their XCD
// mapping) take all-pairs replicas.  They are few and long, and starting first
// they run beside the row segments instead of after them.
struct ApBlocks {
  const int *rep;
  int n, pad;
};

// Pass 1 over every row segment of every linked replica: pair sums and the SE
// update (EPI_SE), or the SRK step's Ito intermediates by slot (EPI_ITO1).
template <int PREC, int EPI>
__global__ __launch_bounds__(kB) void k_b_pass1(const Rep *reps, const int *seg_rep, const int *start, Tiles t,
                                                const int *slot_h, const int *slot_id, const double *gx,
                                                const double *gy, const double *gp, EpiCtx e, int k, ApBlocks ap) {
  extern __shared__ float4 dyn_lds[];
  if ((int)blockIdx.x < ap.pad) {
    if ((int)blockIdx.x < ap.n) b_allpairs1<PREC, EPI>(reps, ap.rep, blockIdx.x, gx, gy, gp, t.f4, e, k);
    return;
  }
  const int sg = xcd_segment(blockIdx.x - ap.pad, gridDim.x - ap.pad);
  const Rep &R = reps[seg_rep[sg]];
  if (k >= R.steps) return;  // the whole workgroup: the replica is done in this call
  const KParams P = R.kp;
  const EpiCtx c = rep_ctx(e, R, k);
  if (R.system == SPC2_CHIMERA)
    pass1_segment<ChimAcc<PREC>, PREC, EPI>(P, start + R.coff, t, slot_h, slot_id, gp, c, nullptr, sg - R.soff,
                                            dyn_lds);
  else
    pass1_segment<VortAcc<PREC>, PREC, EPI>(P, start + R.coff, t, slot_h, slot_id, gp, c, nullptr, sg - R.soff,
                                            dyn_lds);
}

// Pass 2 of the SRK step over every row segment of every linked replica.
template <int PREC>
__global__ __launch_bounds__(kB) void k_b_pass2_srk(const Rep *reps, const int *seg_rep, const int *start, Tiles t,
                                                    const int *slot_h, const int *slot_id, const double *gx,
                                                    const double *gy, const double *gp, EpiCtx e, int k,
                                                    ApBlocks ap) {
  extern __shared__ float4 dyn_lds[];
  if ((int)blockIdx.x < ap.pad) {
    if ((int)blockIdx.x < ap.n) b_allpairs2_srk<PREC>(reps, ap.rep, blockIdx.x, gx, gy, gp, e, k);
    return;
  }
  const int sg = xcd_segment(blockIdx.x - ap.pad, gridDim.x - ap.pad);
  const Rep &R = reps[seg_rep[sg]];
  if (k >= R.steps) return;
  const KParams P = R.kp;
  const EpiCtx c = rep_ctx(e, R, k);
  if (R.system == SPC2_CHIMERA)
    pass2_segment<Chim2Acc<PREC>, PREC, EPI_SRK>(P, start + R.coff, t, slot_h, slot_id, gp, c, sg - R.soff, dyn_lds);
  else
    pass2_segment<Vort2Acc<PREC>, PREC, EPI_SRK>(P, start + R.coff, t, slot_h, slot_id, gp, c, sg - R.soff, dyn_lds);
}

// Re-homing of the SE arrays (x, y, phi, omega): slot k becomes home k.
__global__ __launch_bounds__(kB) void k_b_permute(int n, const int *slot_h, const double *x, const double *y,
                                                  const double *p, const double *w, double *x2, double *y2,
                                                  double *p2, double *w2, const int *hid_in, int *hid_out,
                                                  int *hpos) {
  const int k = blockIdx.x * kB + threadIdx.x;
  if (k >= n) return;
  const int h = slot_h[k];
  x2[k] = x[h];
  y2[k] = y[h];
  p2[k] = p[h];
  w2[k] = w[h];
  const int g = hid_in[h];
  hid_out[k] = g;
  hpos[g] = k;
}

// Philox initial state of every replica, keyed by its seed and the particle's
// id within it (the single-handle keys).
__global__ __launch_bounds__(kB) void k_b_init_state(int n, const Rep *reps, const int *rep_of, const int *hid,
                                                     double omega_std, const int *only, double *sx, double *sy,
                                                     double *sp, double *omega) {
  const int h = blockIdx.x * kB + threadIdx.x;
  if (h >= n) return;
  const int r = rep_of[h];
  if (only && !only[r]) return;
  const Rep &R = reps[r];
  const uint32_t i = (uint32_t)(hid[h] - R.poff);
  double ux, uy, up, unused, z0, z1;
  uniform_pair(R.seed, 0, kStreamInitPos, i, ux, uy);
  uniform_pair(R.seed, 0, kStreamInitPhase, i, up, unused);
  normal_pair(R.seed, 0, kStreamOmega, i, z0, z1);
  sx[h] = ux * R.kp.Lx;
  sy[h] = uy * R.kp.Ly;
  sp[h] = up * 6.283185307179586;
  omega[h] = z0 * omega_std;
}

// id order <-> home order for one replica
__global__ __launch_bounds__(kB) void k_b_to_home(int n, int off, const int *hpos, const double *src, double *dst) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i < n) dst[hpos[off + i]] = src[i];
}

__global__ __launch_bounds__(kB) void k_b_from_home(int n, int off, const int *hpos, const double *src, double *dst) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i < n) dst[i] = src[hpos[off + i]];
}

// Carries n values of one replica from another batch (home order there, id
// order = home order here in a fresh batch).
__global__ __launch_bounds__(kB) void k_b_gather(int n, const int *src_hpos, int src_off, const double *src,
                                                 int dst_off, double *dst) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i < n) dst[dst_off + i] = src[src_hpos[src_off + i]];
}

// Order parameter sums per replica (one workgroup per replica, FP64).
__global__ __launch_bounds__(kB) void k_b_order(const Rep *reps, const double *sp, double *out) {
  __shared__ double sc[kB], ss[kB];
  const Rep &R = reps[blockIdx.x];
  double c = 0.0, s = 0.0;
  for (int i = threadIdx.x; i < R.n; i += kB) {
    double sn, cs;
    sincos(sp[R.poff + i], &sn, &cs);
    c += cs;
    s += sn;
  }
  sc[threadIdx.x] = c;
  ss[threadIdx.x] = s;
  __syncthreads();
  for (int d = kB / 2; d > 0; d >>= 1) {
    if ((int)threadIdx.x < d) {
      sc[threadIdx.x] += sc[threadIdx.x + d];
      ss[threadIdx.x] += ss[threadIdx.x + d];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = sc[0];
    out[2 * blockIdx.x + 1] = ss[0];
  }
}

// Float frames of every replica, concatenated in replica order, AoS by id.
__global__ __launch_bounds__(kB) void k_b_snapshot(int n, const int *hpos, const double *sx, const double *sy,
                                                   const double *sp, float *out) {
  const int g = blockIdx.x * kB + threadIdx.x;
  if (g >= n) return;
  const int h = hpos[g];
  out[3 * (size_t)g] = (float)sx[h];
  out[3 * (size_t)g + 1] = (float)sy[h];
  out[3 * (size_t)g + 2] = (float)sp[h];
}

__global__ __launch_bounds__(kB) void k_b_iota(int n, int *a, int *b) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i < n) {
    a[i] = i;
    b[i] = i;
  }
}

}  // namespace
}  // namespace spc2

// =============================================================================
// Host side
// =============================================================================

struct spc2_batch {
  int count = 0, prec = P32K;
  int nt = 0, nct = 0, nseg = 0, nap = 0;  // particles, cells, row segments, all-pairs blocks
  hipStream_t stream = nullptr;
  std::vector<spc2_params> params;
  std::vector<Rep> reps;
  std::vector<double> target;  // particles per row segment, per replica
  std::vector<int64_t> steps_done;
  Rep *d_reps = nullptr;
  int *rep_of = nullptr, *seg_rep = nullptr, *ap_rep = nullptr;
  double *sx = nullptr, *sy = nullptr, *sp = nullptr, *omega = nullptr, *gx = nullptr, *gy = nullptr, *gp = nullptr;
  double *spare[4] = {};
  int *cell_of = nullptr, *count_ = nullptr, *local = nullptr, *start = nullptr, *tmp = nullptr, *tmpid = nullptr,
      *slot_h = nullptr, *slot_id = nullptr, *hid = nullptr, *hpos = nullptr, *spare_i = nullptr, *bsum = nullptr;
  Tiles t{};
  double *io = nullptr, *red = nullptr;
  size_t io_cap = 0;
  // the SRK step's Ito intermediates by slot (allocated at the first SRK step)
  double *aphi_slot = nullptr, *n_slot = nullptr, *n2_slot = nullptr;
  float *aphi_slot_f = nullptr;
  float *snap = nullptr;
  hipEvent_t snap_ev = nullptr;
  bool snap_pending = false;
  size_t lds = 0;
  int since_rehome = 0;
};

namespace {

#define B_TRY(expr)                                                                               \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) return set_error(SPC2_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
hipError_t balloc(T **p, size_t count) {
  return hipMalloc((void **)p, sizeof(T) * std::max<size_t>(count, 1));
}

void bfree(spc2_batch &b) {
  auto f = [](void *p) {
    if (p) (void)hipFree(p);
  };
  f(b.d_reps); f(b.rep_of); f(b.seg_rep); f(b.ap_rep);
  f(b.sx); f(b.sy); f(b.sp); f(b.omega); f(b.gx); f(b.gy); f(b.gp);
  for (auto &p : b.spare) f(p);
  f(b.cell_of); f(b.count_); f(b.local); f(b.start); f(b.tmp); f(b.tmpid); f(b.slot_h); f(b.slot_id);
  f(b.hid); f(b.hpos); f(b.spare_i); f(b.bsum);
  f(b.t.x); f(b.t.y); f(b.t.p); f(b.t.c); f(b.t.s); f(b.t.f4); f(b.t.cell);
  f(b.io); f(b.red); f(b.snap);
  f(b.aphi_slot); f(b.n_slot); f(b.n2_slot); f(b.aphi_slot_f);
  if (b.snap_ev) (void)hipEventDestroy(b.snap_ev);
  if (b.stream) (void)hipStreamDestroy(b.stream);
}

spc2_status check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(SPC2_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return SPC2_OK;
}

// Replica layout: KParams, cells and segments of each replica (tile targets
// from b.target), the segment -> replica and all-pairs block -> replica maps.
spc2_status layout(spc2_batch &b, std::vector<int> &seg_rep, std::vector<int> &ap_rep) {
  int poff = 0, coff = 0, soff = 0, apoff = 0;
  size_t lds_cap = 0;
  seg_rep.clear();
  ap_rep.clear();
  for (int r = 0; r < b.count; ++r) {
    const spc2_params &p = b.params[r];
    Rep &R = b.reps[r];
    bool linked;
    int nx, ny;
    if (spc2_status st = cell_grid(p, linked, nx, ny)) return st;
    std::memset((void *)&R.kp, 0, sizeof(R.kp));
    fill_kparams_from(R.kp, p, (int)p.n, nx, ny, linked);
    configure_segments_kp(R.kp, b.prec, b.target[r]);
    R.system = p.system;
    R.poff = poff;
    R.n = (int)p.n;
    R.coff = coff;
    R.seed = p.seed;
    if (linked) {
      R.soff = soff;
      R.nseg = ny * R.kp.segs;
      R.apoff = R.apblocks = 0;
      for (int q = 0; q < R.nseg; ++q) seg_rep.push_back(r);
      soff += R.nseg;
      coff += nx * ny;
      lds_cap = std::max(lds_cap, (size_t)R.kp.lds_cap);
    } else {
      R.soff = R.nseg = 0;
      R.apoff = apoff;
      R.apblocks = bfor((long long)R.n * allpairs_lanes(b.prec, p.system));
      for (int q = 0; q < R.apblocks; ++q) ap_rep.push_back(r);
      apoff += R.apblocks;
      coff += 1;
    }
    poff += R.n;
  }
  b.nt = poff;
  b.nct = coff;
  b.nseg = soff;
  b.nap = apoff;
  b.lds = lds_cap * (size_t)lds_elem_bytes(b.prec);
  return SPC2_OK;
}

spc2_status upload_layout(spc2_batch &b) {
  std::vector<int> seg_rep, ap_rep;
  if (spc2_status st = layout(b, seg_rep, ap_rep)) return st;
  if (b.seg_rep) (void)hipFree(b.seg_rep);
  if (b.ap_rep) (void)hipFree(b.ap_rep);
  b.seg_rep = b.ap_rep = nullptr;
  B_TRY(balloc(&b.seg_rep, seg_rep.size()));
  B_TRY(balloc(&b.ap_rep, ap_rep.size()));
  if (!seg_rep.empty())
    B_TRY(hipMemcpyAsync(b.seg_rep, seg_rep.data(), sizeof(int) * seg_rep.size(), hipMemcpyHostToDevice, b.stream));
  if (!ap_rep.empty())
    B_TRY(hipMemcpyAsync(b.ap_rep, ap_rep.data(), sizeof(int) * ap_rep.size(), hipMemcpyHostToDevice, b.stream));
  B_TRY(hipMemcpyAsync(b.d_reps, b.reps.data(), sizeof(Rep) * b.reps.size(), hipMemcpyHostToDevice, b.stream));
  B_TRY(hipStreamSynchronize(b.stream));
  return SPC2_OK;
}

spc2_status alloc_batch(spc2_batch &b) {
  const size_t n = (size_t)b.nt, nc = (size_t)b.nct;
#define BA(ptr, cnt) B_TRY(balloc(&(ptr), (cnt)))
  BA(b.d_reps, (size_t)b.count);
  BA(b.rep_of, n);
  BA(b.sx, n); BA(b.sy, n); BA(b.sp, n); BA(b.omega, n); BA(b.gx, n); BA(b.gy, n); BA(b.gp, n);
  for (auto &p : b.spare) BA(p, n);
  BA(b.cell_of, n); BA(b.count_, nc); BA(b.local, n); BA(b.start, nc + 1); BA(b.tmp, n); BA(b.tmpid, n);
  BA(b.slot_h, n); BA(b.slot_id, n); BA(b.hid, n); BA(b.hpos, n); BA(b.spare_i, n);
  BA(b.bsum, (size_t)scan_blocks((int)std::max<size_t>(nc, 1)) + 1);
  BA(b.t.cell, n); BA(b.t.x, n); BA(b.t.y, n);
  if (b.prec == P32K) {
    BA(b.t.f4, n);
  } else {
    BA(b.t.p, n);
    if (b.prec == P64F) {
      BA(b.t.c, n); BA(b.t.s, n);
    }
  }
  BA(b.red, 2 * (size_t)b.count);
#undef BA
  B_TRY(hipMemsetAsync(b.count_, 0, sizeof(int) * nc, b.stream));
  B_TRY(hipMemsetAsync(b.omega, 0, sizeof(double) * n, b.stream));
  B_TRY(hipMemsetAsync(b.sx, 0, sizeof(double) * n, b.stream));
  B_TRY(hipMemsetAsync(b.sy, 0, sizeof(double) * n, b.stream));
  B_TRY(hipMemsetAsync(b.sp, 0, sizeof(double) * n, b.stream));
  std::vector<int> rep_of(n);
  for (int r = 0; r < b.count; ++r)
    for (int i = 0; i < b.reps[r].n; ++i) rep_of[(size_t)b.reps[r].poff + i] = r;
  B_TRY(hipMemcpyAsync(b.rep_of, rep_of.data(), sizeof(int) * n, hipMemcpyHostToDevice, b.stream));
  hipLaunchKernelGGL(k_b_iota, dim3(bfor(b.nt)), dim3(kB), 0, b.stream, b.nt, b.hid, b.hpos);
  B_TRY(hipStreamSynchronize(b.stream));
  return check_launch("batch init");
}

spc2_status ensure_io(spc2_batch &b, size_t count) {
  if (count <= b.io_cap) return SPC2_OK;
  if (b.io) (void)hipFree(b.io);
  b.io = nullptr;
  b.io_cap = 0;
  B_TRY(balloc(&b.io, count));
  b.io_cap = count;
  return SPC2_OK;
}

// One step of every replica.  staged: the previous step's state-writing pass
// already staged and binned this step's state (gx, gy, gp, cell_of, local and
// the cell counts), so the stage-and-bin launch is skipped.  fuse: this step's
// state-writing pass stages and bins the next step's state into spare[0..2]
// (the caller then swaps them with gx, gy, gp); every replica must take the
// next step and no re-homing may come between.
template <int PREC>
void launch_step_kernels(spc2_batch &b, const EpiCtx &e, int k, bool srk, bool staged, bool fuse) {
  const dim3 g(bfor(b.nt)), blk(kB);
  if (!staged)
    hipLaunchKernelGGL(k_b_stage_bin, g, blk, 0, b.stream, b.nt, (const Rep *)b.d_reps, (const int *)b.rep_of,
                       (const double *)b.sx, (const double *)b.sy, (const double *)b.sp, b.gx, b.gy, b.gp,
                       b.cell_of, b.count_, b.local);
  launch_scan_raw(b.stream, b.bsum, b.count_, b.nct, b.start);
  launch_place_raw(b.stream, b.nt, b.cell_of, b.local, b.start, b.hid, b.tmp, b.tmpid);
  hipLaunchKernelGGL(k_b_rank_tiles<PREC>, g, blk, 0, b.stream, b.nt, (const Rep *)b.d_reps, (const int *)b.rep_of,
                     (const int *)b.tmp, (const int *)b.tmpid, (const int *)b.cell_of, (const int *)b.start,
                     (const double *)b.gx, (const double *)b.gy, (const double *)b.gp, b.slot_h, b.slot_id, b.t);
  const ApBlocks ap{b.ap_rep, b.nap, (b.nap + 7) / 8 * 8};
  const dim3 gpass(ap.pad + b.nseg);
  if (gpass.x == 0) return;
  EpiCtx es = e;  // the state-writing pass's context
  if (fuse) {
    es.bgx = b.spare[0];
    es.bgy = b.spare[1];
    es.bgp = b.spare[2];
    es.bcell = b.cell_of;
    es.bcount = b.count_;
    es.blocal = b.local;
  }
  if (!srk) {
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_b_pass1<PREC, EPI_SE>), gpass, blk, b.lds, b.stream, (const Rep *)b.d_reps,
                       (const int *)b.seg_rep, (const int *)b.start, b.t, (const int *)b.slot_h,
                       (const int *)b.slot_id, (const double *)b.gx, (const double *)b.gy, (const double *)b.gp, es,
                       k, ap);
    return;
  }
  // SRK (StochasticRungeKuttaStepperPtr.hpp:32-97): both passes read the same
  // stage tiles; pass 2 reads every replica's pass-1 values
  EpiCtx e1 = e;
  e1.sa.zero_in = 1;  // derivative zero-filled each step
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_b_pass1<PREC, EPI_ITO1>), gpass, blk, b.lds, b.stream, (const Rep *)b.d_reps,
                     (const int *)b.seg_rep, (const int *)b.start, b.t, (const int *)b.slot_h,
                     (const int *)b.slot_id, (const double *)b.gx, (const double *)b.gy, (const double *)b.gp, e1, k,
                     ap);
  hipLaunchKernelGGL(k_b_pass2_srk<PREC>, gpass, blk, b.lds, b.stream, (const Rep *)b.d_reps, (const int *)b.seg_rep,
                     (const int *)b.start, b.t, (const int *)b.slot_h, (const int *)b.slot_id, (const double *)b.gx,
                     (const double *)b.gy, (const double *)b.gp, es, k, ap);
}

void launch_rehome(spc2_batch &b) {
  hipLaunchKernelGGL(k_b_permute, dim3(bfor(b.nt)), dim3(kB), 0, b.stream, b.nt, (const int *)b.slot_h,
                     (const double *)b.sx, (const double *)b.sy, (const double *)b.sp, (const double *)b.omega,
                     b.spare[0], b.spare[1], b.spare[2], b.spare[3], (const int *)b.hid, b.spare_i, b.hpos);
  std::swap(b.sx, b.spare[0]);
  std::swap(b.sy, b.spare[1]);
  std::swap(b.sp, b.spare[2]);
  std::swap(b.omega, b.spare[3]);
  std::swap(b.hid, b.spare_i);
}

spc2_status run_steps(spc2_batch &b, int method, double dt, int64_t nsteps, const int64_t *limit) {
  const bool srk = method == SPC2_SRK;
  if (srk && !b.aphi_slot) {
    const size_t n = (size_t)b.nt;
    B_TRY(balloc(&b.aphi_slot, n));
    B_TRY(balloc(&b.n_slot, n));
    B_TRY(balloc(&b.n2_slot, n));
    B_TRY(balloc(&b.aphi_slot_f, n));
  }
  int64_t kmax = 0, kmin = std::numeric_limits<int64_t>::max();
  for (int r = 0; r < b.count; ++r) {
    Rep &R = b.reps[r];
    R.step = b.steps_done[r];
    R.steps = limit ? std::min<int64_t>(nsteps, std::max<int64_t>(limit[r], 0)) : nsteps;
    R.dt = dt;
    R.sqdt = std::sqrt(dt);
    R.sq2D = std::sqrt(2.0 * b.params[r].D);
    R.D = b.params[r].D;
    R.h1 = 0.5 * dt * std::sqrt(dt);
    kmax = std::max(kmax, R.steps);
    kmin = std::min(kmin, R.steps);
  }
  B_TRY(hipMemcpyAsync(b.d_reps, b.reps.data(), sizeof(Rep) * b.reps.size(), hipMemcpyHostToDevice, b.stream));
  EpiCtx e{};
  e.sx = b.sx;
  e.sy = b.sy;
  e.sp = b.sp;
  e.sa.wrap = 1;
  e.sa.sqrt3 = std::sqrt(3.0);
  e.aphi_slot = b.aphi_slot;
  e.aphi_slot_f = b.aphi_slot_f;
  e.n_slot = b.n_slot;
  e.n2_slot = b.n2_slot;
  bool staged = false;
  for (int64_t k = 0; k < kmax; ++k) {
    e.sx = b.sx;
    e.sy = b.sy;
    e.sp = b.sp;
    e.sa.omega = b.omega;
    // the next step's stage and bins come from this step's state-writing pass
    // when every replica takes that step and no re-homing comes between
    const bool fuse = k + 1 < kmin && b.since_rehome + 1 < kRehomeEvery && b.nap + b.nseg > 0;
    switch (b.prec) {
      case P64: launch_step_kernels<P64>(b, e, (int)k, srk, staged, fuse); break;
      case P64F: launch_step_kernels<P64F>(b, e, (int)k, srk, staged, fuse); break;
      default: launch_step_kernels<P32K>(b, e, (int)k, srk, staged, fuse);
    }
    if (fuse) {
      std::swap(b.gx, b.spare[0]);
      std::swap(b.gy, b.spare[1]);
      std::swap(b.gp, b.spare[2]);
    }
    staged = fuse;
    if (++b.since_rehome >= kRehomeEvery) {
      launch_rehome(b);
      b.since_rehome = 0;
    }
    if (spc2_status st = check_launch("batch step")) return st;
  }
  B_TRY(hipStreamSynchronize(b.stream));
  for (int r = 0; r < b.count; ++r) b.steps_done[r] += b.reps[r].steps;
  return SPC2_OK;
}

spc2_status check_batch(spc2_batch_t b) {
  if (!b) return set_error(SPC2_ERR_INVALID, "null batch");
  return SPC2_OK;
}

spc2_status check_replica(spc2_batch_t b, int32_t r) {
  if (check_batch(b)) return SPC2_ERR_INVALID;
  if (r < 0 || r >= b->count) return set_error(SPC2_ERR_INVALID, "replica index out of range");
  return SPC2_OK;
}

}  // namespace

extern "C" {

uint64_t spc2_batch_footprint(const spc2_params *p) {
  if (!p || p->n <= 0) return 0;
  bool linked;
  int nx, ny;
  const size_t n = (size_t)p->n;
  const size_t cells = cell_grid(*p, linked, nx, ny) == SPC2_OK ? (size_t)nx * (size_t)ny : 1;
  const size_t tile = p->precision == SPC2_FP32_KAHAN ? 16 + 16 : (p->precision == SPC2_FP64 ? 24 : 40);
  // 11 doubles (state, omega, stage, spare), 11 ints, tiles; 2 ints per cell
  // (SE stepping; the first SRK step adds 28 B per particle)
  return n * (11 * 8 + 11 * 4 + tile) + cells * 8 + sizeof(Rep) + 64;
}

spc2_status spc2_batch_create_from(const spc2_params *replicas, int32_t count, spc2_batch_t prev,
                                   const int32_t *prev_index, double omega_std, spc2_batch_t *out) {
  if (!replicas || !out || count <= 0) return set_error(SPC2_ERR_INVALID, "null argument or empty batch");
  *out = nullptr;
  if (prev && !prev_index) return set_error(SPC2_ERR_INVALID, "prev_index is needed with a previous batch");
  if (!(omega_std >= 0.0)) return set_error(SPC2_ERR_INVALID, "omega_std must be >= 0");
  long long total = 0;
  for (int r = 0; r < count; ++r) {
    const spc2_params &p = replicas[r];
    if (p.n <= 0 || !(p.Lx > 0.0) || !(p.Ly > 0.0) || !(p.rho > 0.0) || !(p.D >= 0.0))
      return set_error(SPC2_ERR_INVALID, "replica " + std::to_string(r) + ": n, Lx, Ly, rho must be > 0, D >= 0");
    if (p.system != SPC2_CHIMERA && p.system != SPC2_VORTEX)
      return set_error(SPC2_ERR_INVALID, "replica " + std::to_string(r) + ": unknown system");
    if (p.precision != replicas[0].precision)
      return set_error(SPC2_ERR_INVALID, "all replicas of a batch share one precision");
    if (p.neighbour != SPC2_NEIGH_AUTO)
      return set_error(SPC2_ERR_UNSUPPORTED, "batched replicas use the reference's neighbour dispatch (AUTO)");
    if (prev && prev_index[r] >= 0) {
      if (prev_index[r] >= prev->count || prev->params[prev_index[r]].n != p.n)
        return set_error(SPC2_ERR_INVALID, "replica " + std::to_string(r) + ": bad carried replica");
    }
    total += p.n;
  }
  if (replicas[0].precision < SPC2_FP64 || replicas[0].precision > SPC2_FP64_FACTORED)
    return set_error(SPC2_ERR_INVALID, "unknown precision");
  if (total >= (1LL << 31) - 1) return set_error(SPC2_ERR_INVALID, "a batch holds fewer than 2^31 particles");
  spc2_batch *b = new (std::nothrow) spc2_batch();
  if (!b) return set_error(SPC2_ERR_NOMEM, "host allocation failed");
  b->count = count;
  b->prec = precision_code(replicas[0].precision);
  b->params.assign(replicas, replicas + count);
  b->reps.assign(count, Rep{});
  b->target.assign(count, kSegTarget);
  b->steps_done.assign(count, 0);
  std::vector<int> seg_rep, ap_rep;
  spc2_status st = layout(*b, seg_rep, ap_rep);
  hipError_t he = hipSuccess;
  if (st == SPC2_OK) {
    he = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (he != hipSuccess) st = set_error(SPC2_ERR_HIP, std::string("HIP init: ") + hipGetErrorString(he));
  }
  if (st == SPC2_OK) st = alloc_batch(*b);
  if (st == SPC2_OK) st = upload_layout(*b);
  if (st == SPC2_OK) {
    // carried replicas take state, frequencies and step counters from prev;
    // new ones get the Philox initial state
    std::vector<int> fresh(count, 1);
    bool any_fresh = false;
    if (prev) {
      (void)hipStreamSynchronize(prev->stream);
      for (int r = 0; r < count && st == SPC2_OK; ++r) {
        const int q = prev_index[r];
        if (q < 0) {
          any_fresh = true;
          continue;
        }
        fresh[r] = 0;
        b->steps_done[r] = prev->steps_done[q];
        const Rep &S = prev->reps[q];
        const Rep &D = b->reps[r];
        const double *src[4] = {prev->sx, prev->sy, prev->sp, prev->omega};
        double *dst[4] = {b->sx, b->sy, b->sp, b->omega};
        for (int a = 0; a < 4; ++a)
          hipLaunchKernelGGL(k_b_gather, dim3(bfor(D.n)), dim3(kB), 0, b->stream, D.n, (const int *)prev->hpos,
                             S.poff, src[a], D.poff, dst[a]);
        st = check_launch("carry replicas");
      }
    } else {
      any_fresh = true;
    }
    if (st == SPC2_OK && any_fresh) {
      int *d_only = nullptr;
      he = balloc(&d_only, (size_t)count);
      if (he == hipSuccess)
        he = hipMemcpyAsync(d_only, fresh.data(), sizeof(int) * count, hipMemcpyHostToDevice, b->stream);
      if (he == hipSuccess) {
        hipLaunchKernelGGL(k_b_init_state, dim3(bfor(b->nt)), dim3(kB), 0, b->stream, b->nt,
                           (const Rep *)b->d_reps, (const int *)b->rep_of, (const int *)b->hid, omega_std,
                           (const int *)d_only, b->sx, b->sy, b->sp, b->omega);
        he = hipStreamSynchronize(b->stream);
      }
      if (d_only) (void)hipFree(d_only);
      if (he != hipSuccess) st = set_error(SPC2_ERR_HIP, std::string("batch init state: ") + hipGetErrorString(he));
    }
    if (st == SPC2_OK) {
      he = hipStreamSynchronize(b->stream);
      if (he != hipSuccess) st = set_error(SPC2_ERR_HIP, hipGetErrorString(he));
    }
  }
  if (st != SPC2_OK) {
    bfree(*b);
    delete b;
    return st;
  }
  *out = b;
  return SPC2_OK;
}

spc2_status spc2_batch_create(const spc2_params *replicas, int32_t count, spc2_batch_t *out) {
  return spc2_batch_create_from(replicas, count, nullptr, nullptr, 0.0, out);
}

spc2_status spc2_batch_destroy(spc2_batch_t b) {
  if (!b) return SPC2_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  bfree(*b);
  delete b;
  return SPC2_OK;
}

spc2_status spc2_batch_init_random_state(spc2_batch_t b, double omega_std) {
  if (check_batch(b)) return SPC2_ERR_INVALID;
  if (!(omega_std >= 0.0)) return set_error(SPC2_ERR_INVALID, "omega_std must be >= 0");
  hipLaunchKernelGGL(k_b_init_state, dim3(bfor(b->nt)), dim3(kB), 0, b->stream, b->nt, (const Rep *)b->d_reps,
                     (const int *)b->rep_of, (const int *)b->hid, omega_std, (const int *)nullptr, b->sx, b->sy,
                     b->sp, b->omega);
  if (spc2_status st = check_launch("batch init state")) return st;
  B_TRY(hipStreamSynchronize(b->stream));
  return SPC2_OK;
}

spc2_status spc2_batch_set_state(spc2_batch_t b, int32_t r, const double *state) {
  if (check_replica(b, r) || !state) return set_error(SPC2_ERR_INVALID, "bad replica or null state");
  const Rep &R = b->reps[r];
  if (spc2_status st = ensure_io(*b, (size_t)R.n)) return st;
  std::vector<double> buf((size_t)R.n);
  double *dst[3] = {b->sx, b->sy, b->sp};
  for (int c = 0; c < 3; ++c) {
    for (int i = 0; i < R.n; ++i) buf[i] = state[3 * (size_t)i + c];
    B_TRY(hipMemcpyAsync(b->io, buf.data(), sizeof(double) * R.n, hipMemcpyHostToDevice, b->stream));
    hipLaunchKernelGGL(k_b_to_home, dim3(bfor(R.n)), dim3(kB), 0, b->stream, R.n, R.poff, (const int *)b->hpos,
                       (const double *)b->io, dst[c]);
    B_TRY(hipStreamSynchronize(b->stream));
  }
  return check_launch("batch set state");
}

spc2_status spc2_batch_get_state(spc2_batch_t b, int32_t r, double *state) {
  if (check_replica(b, r) || !state) return set_error(SPC2_ERR_INVALID, "bad replica or null state");
  const Rep &R = b->reps[r];
  if (spc2_status st = ensure_io(*b, (size_t)R.n)) return st;
  std::vector<double> buf((size_t)R.n);
  const double *src[3] = {b->sx, b->sy, b->sp};
  for (int c = 0; c < 3; ++c) {
    hipLaunchKernelGGL(k_b_from_home, dim3(bfor(R.n)), dim3(kB), 0, b->stream, R.n, R.poff, (const int *)b->hpos,
                       src[c], b->io);
    B_TRY(hipMemcpyAsync(buf.data(), b->io, sizeof(double) * R.n, hipMemcpyDeviceToHost, b->stream));
    B_TRY(hipStreamSynchronize(b->stream));
    for (int i = 0; i < R.n; ++i) state[3 * (size_t)i + c] = buf[i];
  }
  return check_launch("batch get state");
}

spc2_status spc2_batch_set_omega(spc2_batch_t b, int32_t r, const double *omega) {
  if (check_replica(b, r) || !omega) return set_error(SPC2_ERR_INVALID, "bad replica or null omega");
  const Rep &R = b->reps[r];
  if (spc2_status st = ensure_io(*b, (size_t)R.n)) return st;
  B_TRY(hipMemcpyAsync(b->io, omega, sizeof(double) * R.n, hipMemcpyHostToDevice, b->stream));
  hipLaunchKernelGGL(k_b_to_home, dim3(bfor(R.n)), dim3(kB), 0, b->stream, R.n, R.poff, (const int *)b->hpos,
                     (const double *)b->io, b->omega);
  B_TRY(hipStreamSynchronize(b->stream));
  return check_launch("batch set omega");
}

spc2_status spc2_batch_step(spc2_batch_t b, int32_t method, double dt, int64_t nsteps, const int64_t *limit) {
  if (check_batch(b)) return SPC2_ERR_INVALID;
  if (method != SPC2_SE && method != SPC2_SRK)
    return set_error(SPC2_ERR_UNSUPPORTED, "batched replicas step with Stochastic Euler or stochastic Runge-Kutta");
  if (!(dt > 0.0) || nsteps < 0) return set_error(SPC2_ERR_INVALID, "dt must be > 0, steps >= 0");
  return run_steps(*b, method, dt, nsteps, limit);
}

spc2_status spc2_batch_steps_done(spc2_batch_t b, int64_t *steps) {
  if (check_batch(b) || !steps) return set_error(SPC2_ERR_INVALID, "null argument");
  for (int r = 0; r < b->count; ++r) steps[r] = b->steps_done[r];
  return SPC2_OK;
}

spc2_status spc2_batch_order_parameters(spc2_batch_t b, double *R, double *psi) {
  if (check_batch(b) || !R || !psi) return set_error(SPC2_ERR_INVALID, "null argument");
  hipLaunchKernelGGL(k_b_order, dim3(b->count), dim3(kB), 0, b->stream, (const Rep *)b->d_reps, (const double *)b->sp,
                     b->red);
  if (spc2_status st = check_launch("batch order parameter")) return st;
  std::vector<double> s(2 * (size_t)b->count);
  B_TRY(hipMemcpyAsync(s.data(), b->red, sizeof(double) * s.size(), hipMemcpyDeviceToHost, b->stream));
  B_TRY(hipStreamSynchronize(b->stream));
  for (int r = 0; r < b->count; ++r) {
    const double c = s[2 * r] / b->reps[r].n, sn = s[2 * r + 1] / b->reps[r].n;
    R[r] = std::sqrt(c * c + sn * sn);
    psi[r] = std::atan2(sn, c);
  }
  return SPC2_OK;
}

spc2_status spc2_batch_snapshot_async(spc2_batch_t b, float *dst) {
  if (check_batch(b) || !dst) return set_error(SPC2_ERR_INVALID, "null argument");
  if (!b->snap) B_TRY(balloc(&b->snap, 3 * (size_t)b->nt));
  if (!b->snap_ev) B_TRY(hipEventCreateWithFlags(&b->snap_ev, hipEventDisableTiming));
  hipLaunchKernelGGL(k_b_snapshot, dim3(bfor(b->nt)), dim3(kB), 0, b->stream, b->nt, (const int *)b->hpos,
                     (const double *)b->sx, (const double *)b->sy, (const double *)b->sp, b->snap);
  if (spc2_status st = check_launch("batch snapshot")) return st;
  B_TRY(hipMemcpyAsync(dst, b->snap, sizeof(float) * 3 * (size_t)b->nt, hipMemcpyDeviceToHost, b->stream));
  B_TRY(hipEventRecord(b->snap_ev, b->stream));
  b->snap_pending = true;
  return SPC2_OK;
}

spc2_status spc2_batch_snapshot_wait(spc2_batch_t b) {
  if (check_batch(b)) return SPC2_ERR_INVALID;
  if (b->snap_pending) {
    B_TRY(hipEventSynchronize(b->snap_ev));
    b->snap_pending = false;
  }
  return SPC2_OK;
}

spc2_status spc2_batch_layout(spc2_batch_t b, int32_t *first_particle, int32_t *particles) {
  if (check_batch(b) || !first_particle || !particles) return set_error(SPC2_ERR_INVALID, "null argument");
  for (int r = 0; r < b->count; ++r) {
    first_particle[r] = b->reps[r].poff;
    particles[r] = b->reps[r].n;
  }
  return SPC2_OK;
}

spc2_status spc2_batch_set_tile_targets(spc2_batch_t b, const double *particles_per_segment) {
  if (check_batch(b) || !particles_per_segment) return set_error(SPC2_ERR_INVALID, "null argument");
  for (int r = 0; r < b->count; ++r) {
    if (!(particles_per_segment[r] >= 16.0 && particles_per_segment[r] <= 4096.0))
      return set_error(SPC2_ERR_INVALID, "tile targets must lie in [16, 4096] particles per workgroup");
  }
  b->target.assign(particles_per_segment, particles_per_segment + b->count);
  return upload_layout(*b);
}

// Autotuner: per bucket of linked replicas with the same floor(log2 N), the
// pair-kernel tile (particles per workgroup) that minimises the step
// time of the whole batch, by coordinate descent over the buckets.  Trials run
// on a copy of the batch, so the caller's replicas do not move.
spc2_status spc2_batch_autotune(spc2_batch_t b, double dt, int32_t trial_steps, double *ms_per_step) {
  if (check_batch(b)) return SPC2_ERR_INVALID;
  if (!(dt > 0.0) || trial_steps < 1 || trial_steps > 1000)
    return set_error(SPC2_ERR_INVALID, "dt must be > 0 and trial_steps in [1, 1000]");
  static const double kCandidates[] = {128.0, 192.0, 256.0, 320.0, 448.0, 704.0};
  std::vector<int32_t> same(b->count);
  for (int r = 0; r < b->count; ++r) same[r] = r;
  spc2_batch_t trial = nullptr;
  if (spc2_status st = spc2_batch_create_from(b->params.data(), b->count, b, same.data(), 0.0, &trial)) return st;
  std::vector<int> bucket(b->count, -1);
  int nb = 0;
  for (int r = 0; r < b->count; ++r) {
    if (!trial->reps[r].kp.linked) continue;
    bucket[r] = (int)std::floor(std::log2((double)b->params[r].n));
    nb = std::max(nb, bucket[r] + 1);
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  spc2_status st = SPC2_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    st = set_error(SPC2_ERR_HIP, "event creation failed");
  auto time_trial = [&](double &ms) -> spc2_status {
    if (spc2_status q = upload_layout(*trial)) return q;
    if (spc2_status q = run_steps(*trial, SPC2_SE, dt, 1, nullptr)) return q;  // warm-up
    B_TRY(hipEventRecord(e0, trial->stream));
    if (spc2_status q = run_steps(*trial, SPC2_SE, dt, trial_steps, nullptr)) return q;
    B_TRY(hipEventRecord(e1, trial->stream));
    B_TRY(hipEventSynchronize(e1));
    float f = 0.0f;
    B_TRY(hipEventElapsedTime(&f, e0, e1));
    ms = f / trial_steps;
    return SPC2_OK;
  };
  double best_all = 0.0;
  if (st == SPC2_OK) st = time_trial(best_all);
  for (int k = 0; k < nb && st == SPC2_OK; ++k) {
    bool present = false;
    for (int r = 0; r < b->count; ++r) present = present || bucket[r] == k;
    if (!present) continue;
    double best_t = 0.0;
    for (int r = 0; r < b->count; ++r)
      if (bucket[r] == k) best_t = trial->target[r];
    for (double c : kCandidates) {
      if (c == best_t) continue;
      std::vector<double> keep = trial->target;
      for (int r = 0; r < b->count; ++r)
        if (bucket[r] == k) trial->target[r] = c;
      double ms = 0.0;
      st = time_trial(ms);
      if (st) break;
      if (ms < best_all) {
        best_all = ms;
        best_t = c;
      } else {
        trial->target = keep;
      }
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st == SPC2_OK) {
    b->target = trial->target;
    st = upload_layout(*b);
  }
  spc2_batch_destroy(trial);
  if (st == SPC2_OK && ms_per_step) *ms_per_step = best_all;
  return st;
}

spc2_status spc2_batch_get_tile_targets(spc2_batch_t b, double *particles_per_segment) {
  if (check_batch(b) || !particles_per_segment) return set_error(SPC2_ERR_INVALID, "null argument");
  std::copy(b->target.begin(), b->target.end(), particles_per_segment);
  return SPC2_OK;
}

}  // extern "C"
