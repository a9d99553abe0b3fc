// Verlet neighbour lists on gfx950 (the list the reference keeps in
// ChimeraSystemPtr::EvaluateInteractionsWithVerletNeighborList,
// ChimeraSystemPtr.cpp:650-927): every pair within rho_out + skin of the
// positions at the build, found over a 3 x 3 stencil of cells of side
// >= rho_out + skin.  The list stays valid while no particle has moved more
// than skin / 2 since the build; the RHS kernels then visit ~ (1 + skin/rho)^2
// pi / 9 of the cell-stencil candidates and skip the per-RHS cell build.
//
// Layout.  At a build the particle arrays are re-homed into the build's cell
// order, so home h is also the tile index and homes of one cell are
// contiguous.  The list is wavefront-interleaved ("wave-ELL"): the entries of
// the 64 homes of wavefront w occupy rows wrow[w] .. wrow[w+1]-1, entry e of
// lane l at (wrow[w] + e) * 64 + l, so a wavefront reads its e-th entries as
// one 256-byte line.  Each entry packs the neighbour's home (bits 4..31) and
// its stencil cell offset (dcx + 1) | (dcy + 1) << 2, which the FP32 form needs
// to turn cell-local coordinates into a separation.
//
// Per RHS a stage kernel writes the tiles by home (FP64 wrapped stage position
// for the cutoff decisions, and the precision's own tile), then one pass kernel
// per RHS pass runs the same accumulators and epilogues as the cell-stencil
// kernels, a thread per particle.  The passes stage each workgroup's halo of
// tiles in LDS and walk lists of 16-bit halo indices (all three precision forms,
// below); a build whose halo does not fit keeps the gather layout, whose passes
// read the neighbours' tiles from memory (spatially close in home order, so L1 /
// L2 hits).  Compiled with -ffp-contract=off.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <climits>
#include <type_traits>

#include "spc2_device.h"
#include "spc2_internal.h"
#include "spc2_lds_stencil.h"
#include "spc2_pair_accumulators.h"

namespace spc2 {

namespace {

constexpr int kBlock = 256;

inline int blocks_for(long long n, int b = kBlock) { return (int)((n + b - 1) / b); }
constexpr int kRunMax = kVerletRunMax;

// List layout: entries of the 64 homes of wavefront w are stored one int4 per
// lane and group, group g of lane p at int4 (wrow[w] + g) * 64 + p, so that a
// wavefront reads 1 KiB per group load: four 32-bit entries per group in the
// gather layout, eight 16-bit ones in the LDS layout (EntLds).  Gather entry =
// (j << 4) | code, code = (dcx + 1) | (dcy + 1) << 2 for the stencil cell
// (dcx, dcy) of partner j.  Every
// lane's column runs to the wavefront's group count, padded with a valid entry
// (the home itself, never summed: the passes stop at the count; in the LDS
// layout the halo's sentinel tile), and kVerletSpareGroups valid groups follow
// the list, so the passes may read one group ahead without bounds.

// Visits the candidates of home i in a 3 x 3 stencil of build cells (cells of
// side >= rlist), ascending home within each cell, in the reference's stencil
// order; calls f(j, code, in) for every candidate, in = (j != i and within
// rlist).  FP64 forms: min image of the wrapped build positions.  FP32 form:
// the cell-local FP32 tiles with the pass kernels' offset arithmetic, against
// rlist^2 (1 + 2^-13), a superset of the FP64 test (the passes test every entry
// against the cutoffs themselves, and a pair outside rlist at the build stays
// outside rho while the list is used).
// FP32 tiles of the candidates: TilesGlobal reads tile j from memory; the
// count pass reads them from its workgroup's staged halo (HaloTiles), where the
// tiles of build cell jc sit at base(jc) + j.
struct TilesGlobal {
  const float4 *f4;
  __device__ int base(int) const { return 0; }
  __device__ float4 at(int k) const { return f4[k]; }
};

template <int PREC, class F, class TS = TilesGlobal>
__device__ __forceinline__ void for_each_build_candidate(const KParams &P, const int *start, const Tiles &t, int i,
                                                         double rlist2, F &&f, const TS &ts = TS{nullptr}) {
  const int c = t.cell[i];
  const int cx = c % P.nx, cy = c / P.nx;
  if constexpr (PREC == P32K) {
    const float4 vi = t.f4[i];
    const float kx = -P.hxf - vi.x, ky = -P.hyf - vi.y;
    const float rl2 = (float)(rlist2 * (1.0 + 0x1p-13));
    for (int dyo = -1; dyo <= 1; ++dyo) {
      const int row = adjust_cell(cy + dyo, P.ny) * P.nx;
      const float oy = __builtin_fmaf((float)(dyo + 1), P.hyf, ky);
      for (int dxo = -1; dxo <= 1; ++dxo) {
        const int jc = row + adjust_cell(cx + dxo, P.nx);
        const int code = (dxo + 1) | ((dyo + 1) << 2);
        const float ox = __builtin_fmaf((float)(dxo + 1), P.hxf, kx);
        const int e = start[jc + 1];
        const int tb = ts.base(jc);
#pragma unroll 4
        for (int j = start[jc]; j < e; ++j) {
          const float4 v = std::is_same<TS, TilesGlobal>::value ? t.f4[j] : ts.at(tb + j);
          const float dx = v.x + ox, dy = v.y + oy;
          f(j, code, (__builtin_fmaf(dx, dx, dy * dy) <= rl2) & (j != i));
        }
      }
    }
  } else {
    const double xi = t.x[i], yi = t.y[i];
    for (int dyo = -1; dyo <= 1; ++dyo) {
      const int row = adjust_cell(cy + dyo, P.ny) * P.nx;
      for (int dxo = -1; dxo <= 1; ++dxo) {
        const int jc = row + adjust_cell(cx + dxo, P.nx);
        const int code = (dxo + 1) | ((dyo + 1) << 2);
        const int e = start[jc + 1];
#pragma unroll 2
        for (int j = start[jc]; j < e; ++j) {
          double dx = min_image(t.x[j] - xi, P.Lx, P.rLx);
          double dy = min_image(t.y[j] - yi, P.Ly, P.rLy);
          f(j, code, (dx * dx + dy * dy <= rlist2) & (j != i));
        }
      }
    }
  }
}

// The count pass records each candidate's test as one bit (bit q of word q / 32
// at mask[(q / 32) * n + i], candidate q in visiting order), up to mwords words;
// a home with more candidates is tested again by the fill.
// Visits the entries of home i: calls f(j, code) for every candidate within
// rlist, from the mask when it holds all candidates, else by testing again.
// The mask's bits are read as one stream in visiting order; qend / jbase (the
// caller's LDS, a column per lane) hold, per stencil cell, the end of its
// candidates in that order and the home of candidate q less q.
template <int PREC, class F>
__device__ __forceinline__ void for_each_build_neighbour(const KParams &P, const int *start, const Tiles &t, int i,
                                                         double rlist2, const uint32_t *mask, int mwords,
                                                         int (*qend)[kBlock], int (*jbase)[kBlock], F &&f) {
  const int c = t.cell[i];
  const int cx = c % P.nx, cy = c / P.nx;
  const int me = threadIdx.x;
  int total = 0;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int jc = adjust_cell(cy + k / 3 - 1, P.ny) * P.nx + adjust_cell(cx + k % 3 - 1, P.nx);
    const int s0 = start[jc];
    jbase[k][me] = s0 - total;
    total += start[jc + 1] - s0;
    qend[k][me] = total;
  }
  if (total > 32 * mwords) {
    for_each_build_candidate<PREC>(P, start, t, i, rlist2, [&](int j, int code, bool in) {
      if (in) f(j, code);
    });
    return;
  }
  int k = 0, qe = qend[0][me], jb = jbase[0][me], code = 0;  // cell k's stencil code (k % 3) | (k / 3) << 2
  for (int q0 = 0; q0 < total; q0 += 32) {
    uint32_t m = mask[(size_t)(q0 >> 5) * P.n + i];
    if (total - q0 < 32) m &= (1u << (total - q0)) - 1u;
    while (m) {
      const int q = q0 + __ffs(m) - 1;
      m &= m - 1u;
      while (q >= qe) {
        ++k;
        qe = qend[k][me];
        jb = jbase[k][me];
        code = (k % 3) | ((k / 3) << 2);
      }
      f(jb + q, code);
    }
  }
}

// LDS index base of build cell c in workgroup b's halo: index = base + home.
__device__ __forceinline__ int halo_base(const int *run, const int *off, int nrun, const int *start, int c) {
  int k = 0;
  while (k + 1 < nrun && run[2 * k + 2] <= c) ++k;
  return off[k] - start[run[2 * k]];
}

// The tiles of a workgroup's staged halo, for the candidate loop.
struct HaloTiles {
  const float4 *tiles;
  const int *run, *off, *start;
  int nrun;
  __device__ int base(int c) const { return halo_base(run, off, nrun, start, c); }
  __device__ float4 at(int k) const { return tiles[k]; }
};

// Build pass 1: entries per home, their candidate bits and each wavefront's
// group count (its longest list in groups of four; wgroups8, when set, in
// groups of eight, for the LDS layout).  The FP32 form reads the candidates'
// tiles from the workgroup's halo staged in LDS when the halo runs are given
// (bh_nrun) and the halo fits kCountHalo tiles, else from memory: the same
// values either way.
constexpr int kCountHalo = 1536;  // 24 KiB of LDS per count workgroup

template <int PREC>
__global__ __launch_bounds__(kBlock) void k_verlet_count(KParams P, const int *start, Tiles t, double rlist2,
                                                         uint32_t *mask, int mwords, int *vcnt, int *wgroups,
                                                         int *wgroups8, const int *bh_nrun, const int *bh_run,
                                                         const int *bh_off) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int cnt = 0;
  uint32_t word = 0;
  int q = 0;
  const auto bit = [&](int, int, bool in) {
    word |= (uint32_t)in << (q & 31);
    cnt += in;
    if ((q & 31) == 31) {
      if ((q >> 5) < mwords) mask[(size_t)(q >> 5) * P.n + i] = word;
      word = 0;
    }
    ++q;
  };
  bool staged = false;
  if constexpr (PREC == P32K) {
    __shared__ float4 halo[kCountHalo];
    __shared__ int run[2 * kRunMax], off[kRunMax + 1];
    const int b = blockIdx.x, nrun = bh_nrun ? bh_nrun[b] : -1;
    if (nrun >= 0) {  // uniform over the workgroup
      if ((int)threadIdx.x < 2 * kRunMax) run[threadIdx.x] = bh_run[(size_t)b * 2 * kRunMax + threadIdx.x];
      if ((int)threadIdx.x <= kRunMax) off[threadIdx.x] = bh_off[(size_t)b * (kRunMax + 1) + threadIdx.x];
      __syncthreads();
      const int total = off[nrun];
      staged = total <= kCountHalo;
      if (staged) {
        for (int el = threadIdx.x; el < total; el += kBlock) {
          int k = 0;
          while (k + 1 < nrun && off[k + 1] <= el) ++k;
          halo[el] = t.f4[start[run[2 * k]] + (el - off[k])];
        }
        __syncthreads();
        if (i < P.n) for_each_build_candidate<PREC>(P, start, t, i, rlist2, bit, HaloTiles{halo, run, off, start, nrun});
      }
    }
  }
  if (!staged && i < P.n) for_each_build_candidate<PREC>(P, start, t, i, rlist2, bit);
  if (i < P.n) {
    if ((q & 31) && (q >> 5) < mwords) mask[(size_t)(q >> 5) * P.n + i] = word;
    vcnt[i] = cnt;
  }
  int m = cnt;
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) m = max(m, __shfl_xor(m, d, 64));
  const int w = i >> 6;
  if ((threadIdx.x & 63) == 0 && w * 64 < P.n) {
    wgroups[w] = (m + 3) >> 2;
    if (wgroups8) wgroups8[w] = (m + 7) >> 3;
  }
}

// Build pass 2: the entries, in the order of the count, four per store, then
// the column padded to the wavefront's group count.
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_verlet_fill(KParams P, const int *start, Tiles t, double rlist2,
                                                        const uint32_t *mask, int mwords, const int *wrow,
                                                        int *vlist) {
  __shared__ int qend[9][kBlock], jbase[9][kBlock];  // per lane and stencil cell
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const int w = i >> 6, groups = wrow[w + 1] - wrow[w];
  int4 *dst = (int4 *)vlist + (size_t)wrow[w] * 64 + (i & 63);
  const int self = (i << 4) | 5;
  // the group being formed: entries enter at the top and move down a word
  // (register moves instead of a switch on the entry's position)
  int r0 = self, r1 = self, r2 = self, r3 = self;
  int e = 0;
  for_each_build_neighbour<PREC>(P, start, t, i, rlist2, mask, mwords, qend, jbase, [&](int j, int code) {
    r0 = r1;
    r1 = r2;
    r2 = r3;
    r3 = (j << 4) | code;
    if ((e & 3) == 3) dst[(size_t)(e >> 2) * 64] = make_int4(r0, r1, r2, r3);
    ++e;
  });
  if (e & 3) {  // the last group, completed with the home itself (never summed)
    for (int k = e & 3; k < 4; ++k) {
      r0 = r1;
      r1 = r2;
      r2 = r3;
      r3 = self;
    }
    dst[(size_t)(e >> 2) * 64] = make_int4(r0, r1, r2, r3);
  }
  for (int g = (e + 3) >> 2; g < groups; ++g) dst[(size_t)g * 64] = make_int4(self, self, self, self);
}

// ---- LDS-staged passes (FP32 form) -----------------------------------------------
// A pass workgroup covers kBlock consecutive homes: a run of consecutive build
// cells in one grid row, or the end of one row and the start of the next.  Its
// halo is the union of those cells' 3 x 3 stencils.  Since the homes of
// consecutive cells are contiguous, the halo is a few runs of consecutive cells
// (one per halo row and column interval), each one contiguous run of tiles.  At a
// build every workgroup's runs and their LDS offsets are recorded, and list
// entries hold the partner's index in its workgroup's staged halo instead of its
// home.  The passes stage the halo tiles in LDS with coalesced reads and gather
// from LDS.  A build where a workgroup spans more than two cell rows, or whose
// largest halo does not fit the LDS capacity, keeps the gather layout.

// Runs of halo cells of workgroup b in increasing cell order: cells
// [run0[r], run1[r]], LDS offset roff[r] (roff[nrun] = halo size).  nrun = -1:
// the workgroup spans more than two rows (or needs more than kRunMax runs).
__global__ __launch_bounds__(64) void k_block_halo(KParams P, const int *start, const int *cell, int nblk,
                                                   int *bh_nrun, int *bh_run, int *bh_off, int *bh_max) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  const int nx = P.nx, ny = P.ny;
  const int c0 = cell[b * kBlock], c1 = cell[min(P.n, (b + 1) * kBlock) - 1];
  const int y0 = c0 / nx, y1 = c1 / nx;
  int *run = bh_run + (size_t)b * 2 * kRunMax;  // run[2r], run[2r + 1]: first and last cell
  int nrun = 0;
  // at least four build rows, so that the halo rows y0 - 1 .. y1 + 1 are
  // distinct and each has one row offset from y0 (the staged tiles' y frame)
  bool ok = y1 - y0 <= 1 && ny >= 4;
  // column intervals of halo row r: the workgroup's intervals in rows r - 1 .. r + 1
  // widened by one column (and wrapped), emitted as runs in increasing column order
  for (int r = y0 - 1; ok && r <= y1 + 1; ++r) {
    int a[2], e[2], m = 0;
    for (int yb = y0; yb <= y1; ++yb) {
      if (yb < r - 1 || yb > r + 1) continue;
      a[m] = (yb == y0 ? c0 % nx : 0) - 1;
      e[m] = (yb == y1 ? c1 % nx : nx - 1) + 1;
      ++m;
    }
    if (m == 2 && a[1] <= e[0] + 1 && a[0] <= e[1] + 1) {
      a[0] = min(a[0], a[1]);
      e[0] = max(e[0], e[1]);
      m = 1;
    }
    const int row = adjust_cell(r, ny) * nx;
    for (int q = 0; q < m; ++q) {
      int lo = a[q], hi = e[q];
      if (hi - lo + 1 >= nx) {
        lo = 0;
        hi = nx - 1;
      }
      // [lo, hi] within [-1, nx]: at most three pieces after wrapping
      const int plo[3] = {lo < 0 ? nx - 1 : -1, max(lo, 0), hi >= nx ? 0 : -1};
      const int phi[3] = {lo < 0 ? nx - 1 : -2, min(hi, nx - 1), hi >= nx ? 0 : -2};
      for (int k = 0; k < 3; ++k) {
        if (phi[k] < plo[k]) continue;
        if (nrun == kRunMax) {
          ok = false;
          break;
        }
        run[2 * nrun] = row + plo[k];
        run[2 * nrun + 1] = row + phi[k];
        ++nrun;
      }
    }
  }
  if (!ok) {
    bh_nrun[b] = -1;
    atomicMax(bh_max, INT_MAX);
    return;
  }
  // sort the runs by first cell, then merge overlapping or adjacent ones of one
  // row (wrapped columns may repeat cells); a run stays within one row
  for (int k = 1; k < nrun; ++k) {
    const int f = run[2 * k], l = run[2 * k + 1];
    int j = k - 1;
    while (j >= 0 && run[2 * j] > f) {
      run[2 * j + 2] = run[2 * j];
      run[2 * j + 3] = run[2 * j + 1];
      --j;
    }
    run[2 * j + 2] = f;
    run[2 * j + 3] = l;
  }
  int u = 0;
  for (int k = 0; k < nrun; ++k) {
    if (u > 0 && run[2 * k] <= run[2 * u - 1] + 1 && run[2 * k] / nx == run[2 * u - 1] / nx) {
      run[2 * u - 1] = max(run[2 * u - 1], run[2 * k + 1]);
    } else {
      run[2 * u] = run[2 * k];
      run[2 * u + 1] = run[2 * k + 1];
      ++u;
    }
  }
  int *off = bh_off + (size_t)b * (kRunMax + 1);
  int tot = 0;
  for (int k = 0; k < u; ++k) {
    off[k] = tot;
    tot += start[run[2 * k + 1] + 1] - start[run[2 * k]];
  }
  off[u] = tot;
  bh_nrun[b] = u;
  atomicMax(bh_max, tot);
}

// Lane slots of the LDS layout: within each pass workgroup (kBlock homes) the
// homes by decreasing list length (ties by home), so that a wavefront's lists
// are of similar length and it walks about their mean rather than the longest
// of 64 independent lists.  The workgroup's halo is the same whichever lane
// takes a home, but the per-particle reads and writes of the epilogue scatter
// over the workgroup's homes; the build uses the slots only where lists are
// long (see verlet_build).  vslot[slot] = home (null: slot = home);
// wmax_s[w] = wavefront w's group count (groups of eight), from the length at
// its first slot.
// The order is a bitonic sort of the workgroup's 256 keys (0xffff - length) << 8
// | lane in LDS, ascending: decreasing length, ties by home (lengths clamped to
// 0xffff for the key; missing homes of the last workgroup sort last).  A
// wavefront's group count is the maximum of its lanes' lengths, whatever the
// order.
__global__ __launch_bounds__(kBlock) void k_verlet_order(int n, const int *vcnt, int *vslot, int *wmax_s) {
  static_assert(kBlock == 256, "the key holds the lane in 8 bits");
  __shared__ uint32_t key[kBlock];
  const int base = blockIdx.x * kBlock, m = min(kBlock, n - base);
  const int me = threadIdx.x;
  key[me] = me < m ? ((uint32_t)(0xffff - min(vcnt[base + me], 0xffff)) << 8) | (uint32_t)me : 0xffffffffu;
  __syncthreads();
  for (int k = 2; k <= kBlock; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int p = me ^ j;
      if (p > me) {
        const uint32_t a = key[me], b = key[p];
        if (((me & k) == 0) == (a > b)) {
          key[me] = b;
          key[p] = a;
        }
      }
      __syncthreads();
    }
  }
  const uint32_t kv = key[me];  // position me of the order
  const int home = base + (int)(kv & 0xffu);
  int len = me < m ? vcnt[home] : 0;
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) len = max(len, __shfl_xor(len, d, 64));
  if (me >= m) return;
  const int slot = base + me;
  vslot[slot] = home;
  if ((slot & 63) == 0) wmax_s[slot >> 6] = (len + 7) >> 3;
}

// FP32 list entry formats.
//   Gather lists: (j << 4) | (dcx + 1) | (dcy + 1) << 2, j a home, so that j << 4,
//   the tile's byte offset, is one mask from the entry; four entries per lane
//   and group.  Entries past the count are skipped.
//   LDS lists: 16-bit entries (h << 4) | (dcx + 1) | (dcy + 1) << 2, h an index
//   into the workgroup's staged halo (below 2^12), so that h << 4, the staged
//   tile's LDS address, is one mask from the entry; eight entries per lane and
//   group, two to a 32-bit word (the first in the low half).  Padding entries
//   name the sentinel tile staged after the halo, far from every particle, so
//   that every entry of a group is summed unguarded.
struct EntGather {
  __device__ static uint32_t off(int e) { return (uint32_t)e & ~15u; }
  __device__ static float4 tile(const char *f4b, uint32_t off) { return *(const float4 *)(f4b + off); }
  __device__ static int index(int e) { return e >> 4; }
  __device__ static float fx(int e) { return (float)(e & 3); }
  __device__ static float fy(int e) { return (float)((e >> 2) & 3); }
};
struct EntLds {
  __device__ static uint32_t off(uint32_t e) { return e & 0xfff0u; }
  // the staged tiles start at LDS address 0 (the passes have no static LDS), so
  // the byte offset is the LDS address: no base is added per entry
  __device__ static float4 tile(uint32_t off) {
    typedef float v4f __attribute__((ext_vector_type(4)));
    const v4f x = *(const __attribute__((address_space(3))) v4f *)(uintptr_t)off;
    return make_float4(x.x, x.y, x.z, x.w);
  }
  __device__ static int index(uint32_t e) { return (int)((e >> 4) & 0xfffu); }
  __device__ static float fx(uint32_t e) { return (float)(e & 3u); }
  __device__ static float fy(uint32_t e) { return (float)((e >> 2) & 3u); }
  // entry for halo index h and stencil code (dcx + 1) | (dcy + 1) << 2
  __host__ __device__ static uint32_t make(int h, int code) { return ((uint32_t)h << 4) | (uint32_t)code; }
};
constexpr float kSentinelCoord = 1.0e6f;  // sentinel tile x and y: r^2 ~ 2e12, beyond every cutoff
static_assert(kLdsBytes / 20 < (1 << 12), "staged halo indices must fit EntLds's 12 bits");

// An LDS pass workgroup's dynamic LDS: the staged tiles (16 B each, from
// address 0), for pass 2 the pass-1 derivatives (4 B each), then this tail with
// the halo's runs (which also give a staged tile's home) and the pass-1 count
// partials.
struct LdsTail {
  int gst[kRunMax];      // first home of each run
  int off[kRunMax + 1];  // LDS index of each run's first tile; off[n] = halo size
  int n;                 // run count
  long long part[3][kBlock / 64];
  float ryf[kRunMax];  // each run's y frame offset (row - r0) * h_y (exact)
};
static_assert(sizeof(LdsTail) % 8 == 0, "LdsTail keeps the 8-byte alignment of its offset");
constexpr size_t lds_pass_bytes(int cap, int pass) { return (size_t)cap * (pass == 2 ? 20 : 16) + sizeof(LdsTail); }
// pass 1 of the vortex system: the cohesion queue (walk_lds) follows the tail
constexpr int kCohQueue = 4;
constexpr size_t kCohQueueBytes = sizeof(uint32_t) * kBlock * kCohQueue;

// Build pass 2 for the LDS layout (FP32 form), per lane slot: as k_verlet_fill,
// with EntLds entries (the partner's index in the workgroup's staged halo in
// place of its home), eight per group, and the halo's sentinel tile as padding.
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_verlet_fill_lds(KParams P, const int *start, Tiles t, double rlist2,
                                                            const uint32_t *mask, int mwords, const int *vslot,
                                                            const int *wrow, const int *bh_nrun, const int *bh_run,
                                                            const int *bh_off, int *vlist) {
  // the workgroup is one pass workgroup: its halo runs go to LDS once
  __shared__ int run[2 * kRunMax], off[kRunMax + 1];
  __shared__ int qend[9][kBlock];  // per lane and stencil cell, below
  __shared__ uint32_t ebase[9][kBlock];
  const int b = blockIdx.x, nrun = bh_nrun[b];
  if ((int)threadIdx.x < 2 * kRunMax) run[threadIdx.x] = bh_run[(size_t)b * 2 * kRunMax + threadIdx.x];
  if ((int)threadIdx.x <= kRunMax) off[threadIdx.x] = bh_off[(size_t)b * (kRunMax + 1) + threadIdx.x];
  __syncthreads();
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= P.n) return;
  const int i = vslot ? vslot[slot] : slot;
  const int w = slot >> 6, groups = wrow[w + 1] - wrow[w];
  int4 *dst = (int4 *)vlist + (size_t)wrow[w] * 64 + (slot & 63);
  const int c = t.cell[i], cx = c % P.nx, cy = c / P.nx;
  const uint32_t pad = EntLds::make(off[nrun], 5);  // the sentinel tile after the halo
  const int pad2 = (int)(pad | pad << 16);
  // the group being formed, a 128-bit shift register of 16-bit entries: each
  // entry enters at the top and the others move down 16 bits (one funnel shift
  // per word), so that after eight entries word k holds entries 2k, 2k + 1
  uint32_t r0 = 0, r1 = 0, r2 = 0, r3 = 0;
  const auto push = [&](uint32_t ent) {
    r0 = __builtin_amdgcn_alignbit(r1, r0, 16);
    r1 = __builtin_amdgcn_alignbit(r2, r1, 16);
    r2 = __builtin_amdgcn_alignbit(r3, r2, 16);
    r3 = __builtin_amdgcn_alignbit(ent, r3, 16);
  };
  int e = 0;
  const auto emit = [&](uint32_t ent) {
    push(ent);
    if ((e & 7) == 7) dst[(size_t)(e >> 3) * 64] = make_int4((int)r0, (int)r1, (int)r2, (int)r3);
    ++e;
  };
  // Per build cell k of the stencil, in visiting order: the end of its
  // candidates in that order (qend) and the entry of candidate q less q << 4
  // (ebase: the halo index base and the stencil code; the sum is exact modulo
  // 2^32 and its low four bits stay the code)
  int total = 0;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int dyo = k / 3 - 1, dxo = k % 3 - 1;
    const int jc = adjust_cell(cy + dyo, P.ny) * P.nx + adjust_cell(cx + dxo, P.nx);
    const int s0 = start[jc];
    ebase[k][threadIdx.x] = EntLds::make(halo_base(run, off, nrun, start, jc) + s0 - total,
                                         (dxo + 1) | ((dyo + 1) << 2));
    total += start[jc + 1] - s0;
    qend[k][threadIdx.x] = total;
  }
  if (total > 32 * mwords) {  // the count pass kept no bits for this home: test again
    int last = -1, base = 0;
    for_each_build_candidate<PREC>(P, start, t, i, rlist2, [&](int j, int code, bool in) {
      if (!in) return;
      if (code != last) {
        const int jc = adjust_cell(cy + (code >> 2) - 1, P.ny) * P.nx + adjust_cell(cx + (code & 3) - 1, P.nx);
        base = halo_base(run, off, nrun, start, jc);
        last = code;
      }
      emit(EntLds::make(base + j, code));
    });
  } else {
    // the candidate bits as one stream, word by word; candidate q lies in the
    // first cell whose qend exceeds it (cells only advance)
    int k = 0, qe = qend[0][threadIdx.x];
    uint32_t eb = ebase[0][threadIdx.x];
    for (int q0 = 0; q0 < total; q0 += 32) {
      uint32_t m = mask[(size_t)(q0 >> 5) * P.n + i];
      if (total - q0 < 32) m &= (1u << (total - q0)) - 1u;
      while (m) {
        const int q = q0 + __ffs(m) - 1;
        m &= m - 1u;
        while (q >= qe) {
          ++k;
          qe = qend[k][threadIdx.x];
          eb = ebase[k][threadIdx.x];
        }
        emit(eb + ((uint32_t)q << 4));
      }
    }
  }
  if (e & 7) {  // the last group, completed with sentinel entries
    for (int k = e & 7; k < 8; ++k) push(pad);
    dst[(size_t)(e >> 3) * 64] = make_int4((int)r0, (int)r1, (int)r2, (int)r3);
  }
  for (int g = (e + 7) >> 3; g < groups; ++g) dst[(size_t)g * 64] = make_int4(pad2, pad2, pad2, pad2);
}

// Per-RHS stage: s + (k_coef dt) k by home, the FP64 tile (conditionally
// wrapped, ChimeraSystem.cpp:291-309) and the precision's own tile.  FP32
// coordinates are relative to the particle's build cell, taken at the nearest
// periodic image (a particle is within skin / 2 of its build position).
template <int PREC>
__device__ __forceinline__ void write_stage(const KParams &P, const Tiles &t, double *gp, int h, double x, double y,
                                            double p) {
  Box b{P.Lx, P.Ly, P.rLx, P.rLy};
  cond_wrap(b, x, y);
  t.x[h] = x;
  t.y[h] = y;
  gp[h] = p;
  if (PREC == P32K) {
    const int c = t.cell[h];
    double xl = x - (double)(c % P.nx) * P.hx, yl = y - (double)(c / P.nx) * P.hy;
    if (xl > 0.5 * P.Lx) xl -= P.Lx;
    else if (xl < -0.5 * P.Lx) xl += P.Lx;
    if (yl > 0.5 * P.Ly) yl -= P.Ly;
    else if (yl < -0.5 * P.Ly) yl += P.Ly;
    double sn, cs;
    sincos(p, &sn, &cs);
    t.f4[h] = make_float4((float)xl, (float)yl, (float)cs, (float)sn);
  } else {
    t.p[h] = p;
    if (PREC == P64F) {
      double sn, cs;
      sincos(p, &sn, &cs);
      t.c[h] = cs;
      t.s[h] = sn;
    }
  }
}

template <int PREC>
__global__ __launch_bounds__(kBlock) void k_verlet_stage(KParams P, const double *sx, const double *sy,
                                                         const double *sp, const double *kx, const double *ky,
                                                         const double *kp, double kcdt, Tiles t, double *gp) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= P.n) return;
  double x = sx[h], y = sy[h], p = sp[h];
  if (kx) {
    x = x + kcdt * kx[h];
    y = y + kcdt * ky[h];
    p = p + kcdt * kp[h];
  }
  write_stage<PREC>(P, t, gp, h, x, y, p);
}

// One list entry into the accumulator: the FP32 form learns the stencil cell
// offset through row() and the column bounds (sel = (m >= b2) - (m < b1)).
template <class Acc, class S>
__device__ __forceinline__ void visit_entry(Acc &acc, const S &src, int ent, const KParams &P) {
  const int j = ent >> 4, dcx = (ent & 3) - 1, dcy = ((ent >> 2) & 3) - 1;
  const int b1 = dcx < 0 ? j + 1 : (dcx == 0 ? j : INT_MIN);
  const int b2 = dcx < 0 ? INT_MAX : (dcx == 0 ? j + 1 : j);
  acc.row(dcy, b1, b2, P);
  acc.pair(src, j, P);
}

// The FP32 passes' y frame.  The homes of a pass workgroup (kBlock consecutive
// homes) lie in build rows r0 and r0 + 1, r0 the row of its first home.  A
// tile's y enters the walk as its cell-local y plus (row - r0) h_y, exact for
// row - r0 in {-1, 0, 1, 2} (staged once per halo run in the LDS layout, formed
// per entry from the stencil row in the gather layout, the same value), and the
// home's y as yl + (row_i - r0) h_y, so that a separation's y is one
// difference.  Returns row_i - r0; a home whose workgroup spans more rows
// (gather layout only) takes its own row as r0.
__device__ __forceinline__ int yframe_rows(const KParams &P, const int *cell, int i) {
  const int ri = cell[i] / P.nx - cell[i & ~(kBlock - 1)] / P.nx;
  return ri > 1 ? 0 : ri;
}

// FP32 list walk over a gather list, a group of four entries in flight: the
// next group's entries and the current group's partner tiles (and pass-2
// derivatives) are requested before the group is summed, in list order, so one
// memory latency covers four entries.  Every eight entries are summed into the
// accumulator's partials and flushed into its compensated sums, as walk_lds
// does, so that both layouts give the same sums.
template <int PASS, class Acc>
__device__ __forceinline__ void walk32(Acc &acc, const Src32 &src, const int4 *l, int cnt, const KParams &P,
                                       float ysh) {
  using F = EntGather;
  if (cnt <= 0) return;
  const char *f4b = (const char *)src.f4;
  const char *ajb = (const char *)src.aj;
  int4 g = *l;
  for (int q = 0; q < cnt; q += 8) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (h == 1 && q + 4 >= cnt) break;
      const int e[4] = {g.x, g.y, g.z, g.w};
      float4 v[4];
      float a[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t off = F::off(e[u]);
        v[u] = F::tile(f4b, off);
        a[u] = PASS == 2 ? *(const float *)(ajb + (off >> 2)) : 0.0f;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)  // into the frame: + (row_i - r0 + dcy) h_y, ysh = (row_i - r0 - 1) h_y
        v[u].y = v[u].y + __builtin_fmaf(F::fy(e[u]), P.hyf, ysh);
      l += 64;
      g = *l;  // the next group (a spare group follows the list)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (h == 0 && u == 0)
          acc.template pair_c<SUM_GROUP_FIRST>(src, F::index(e[u]), F::fx(e[u]), F::fy(e[u]), v[u], a[u], P);
        else if (q + 4 * h + u < cnt)
          acc.template pair_c<SUM_GROUP>(src, F::index(e[u]), F::fx(e[u]), F::fy(e[u]), v[u], a[u], P);
      }
    }
    acc.flush();
  }
}

// FP32 list walk over an LDS list: groups of eight 16-bit entries, taken as two
// halves of four whose staged tiles (and pass-2 derivatives) are requested
// before the half is summed; the lane's next group of entries is requested
// before the current group is summed (walk_lds_sets, for the walks without a
// cohesion queue: in two sets of word registers taken in turn).  Sentinel
// padding is summed like any entry (it adds zero), and each group's partials
// are flushed once.  RHO_BITS (chimera pass 1): the group's eight within-rho
// decisions are stored as one byte, bit k for entry k, at mb (the list's
// interleaved layout, a byte per lane and group).  COH_QUEUE (vortex pass 1):
// entries within xi_r's band, a few per particle, are queued in the lane's LDS
// column cq (kCohQueue slots, stride kBlock) and their cohesion terms summed
// after the walk, in list order (a full queue is emptied first), so that the
// divergent cohesion branch runs once per wavefront rather than at every entry
// where one lane needs it.  The vortex walk keeps one set of word registers: a
// second costs it more in registers than it saves in copies.

template <class Acc>
__device__ __forceinline__ void coh_drain(Acc &acc, const Src32 &src, const uint32_t *cq, int qn,
                                          const KParams &P) {
  using F = EntLds;
  for (int j = 0; j < qn; ++j) {
    const uint32_t e = cq[j * kBlock];
    const float4 v = F::tile(F::off(e));
    // the separation and r^2 exactly as pair_c forms them
    const float dx = v.x + acc.co.cdx(F::fx(e), P.hxf), dy = v.y - acc.co.yf;
    acc.cohesion(src, F::index(e), dx, dy, __builtin_fmaf(dx, dx, dy * dy), P);
  }
}

// One group of walk_lds's list without a cohesion queue (chimera pass 1 with or
// without rho bits, pass 2): the eight entries as two halves of four, each
// half's staged tiles (and pass-2 derivatives) requested before it is summed.
template <int PASS, bool RHO_BITS, class Acc>
__device__ __forceinline__ void walk_group(Acc &acc, const Src32 &src, const uint32_t (&w)[4], const KParams &P,
                                           uint8_t *&mb) {
  using F = EntLds;
  const char *ajb = (const char *)src.aj;
  uint32_t bits = 0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint32_t e[4] = {w[2 * h] & 0xffffu, w[2 * h] >> 16, w[2 * h + 1] & 0xffffu, w[2 * h + 1] >> 16};
    float4 v[4];
    float a[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t off = F::off(e[u]);
      v[u] = F::tile(off);
      a[u] = PASS == 2 ? *(const float *)(ajb + (off >> 2)) : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if constexpr (RHO_BITS) {
        const bool in =
            h == 0 && u == 0
                ? acc.template pair_c<SUM_GROUP_FIRST>(src, F::index(e[u]), F::fx(e[u]), F::fy(e[u]), v[u], a[u], P)
                : acc.template pair_c<SUM_GROUP>(src, F::index(e[u]), F::fx(e[u]), F::fy(e[u]), v[u], a[u], P);
        bits |= (uint32_t)in << (4 * h + u);
      } else if (h == 0 && u == 0) {
        acc.template pair_c<SUM_GROUP_FIRST>(src, F::index(e[u]), F::fx(e[u]), F::fy(e[u]), v[u], a[u], P);
      } else {
        acc.template pair_c<SUM_GROUP>(src, F::index(e[u]), F::fx(e[u]), F::fy(e[u]), v[u], a[u], P);
      }
    }
  }
  acc.flush();
  if constexpr (RHO_BITS) {
    *mb = (uint8_t)bits;
    mb += 64;
  }
}

// walk_lds without a cohesion queue: the groups two at a time in two sets of
// word registers taken in turn, so that the lane's next group is requested
// before the current one is summed with no copies between the sets.  The
// terms, groups and rho bits are those of the one-set walk, in the same order.
template <int PASS, bool RHO_BITS, class Acc>
__device__ __forceinline__ void walk_lds_sets(Acc &acc, const Src32 &src, const int4 *l, int cnt, const KParams &P,
                                              uint8_t *mb) {
  if (cnt <= 0) return;
  int4 g0 = *l, g1;
  for (int q = 0;; q += 16) {
    l += 64;
    if (q + 8 < cnt) g1 = *l;  // the lane's next group, if it has one
    const uint32_t w0[4] = {(uint32_t)g0.x, (uint32_t)g0.y, (uint32_t)g0.z, (uint32_t)g0.w};
    walk_group<PASS, RHO_BITS>(acc, src, w0, P, mb);
    if (q + 8 >= cnt) break;
    l += 64;
    if (q + 16 < cnt) g0 = *l;
    const uint32_t w1[4] = {(uint32_t)g1.x, (uint32_t)g1.y, (uint32_t)g1.z, (uint32_t)g1.w};
    walk_group<PASS, RHO_BITS>(acc, src, w1, P, mb);
    if (q + 16 >= cnt) break;
  }
}

template <int PASS, bool RHO_BITS = false, bool COH_QUEUE = false, class Acc>
__device__ __forceinline__ void walk_lds(Acc &acc, const Src32 &src, const int4 *l, int cnt, const KParams &P,
                                         uint8_t *mb = nullptr, uint32_t *cq = nullptr) {
  using F = EntLds;
  if constexpr (!COH_QUEUE) {
    walk_lds_sets<PASS, RHO_BITS>(acc, src, l, cnt, P, mb);
    return;
  }
  if (cnt <= 0) return;
  int qn = 0;
  int4 g = *l;
  for (int q = 0; q < cnt; q += 8) {
    const uint32_t w[4] = {(uint32_t)g.x, (uint32_t)g.y, (uint32_t)g.z, (uint32_t)g.w};
    l += 64;
    if (q + 8 < cnt) g = *l;  // the lane's next group, if it has one
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t e[4] = {w[2 * h] & 0xffffu, w[2 * h] >> 16, w[2 * h + 1] & 0xffffu, w[2 * h + 1] >> 16};
      float4 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = F::tile(F::off(e[u]));
      if constexpr (COH_QUEUE) {
        // two entries at a time: one branch for both rho band re-checks and one
        // for both queue tests
#pragma unroll
        for (int u = 0; u < 4; u += 2) {
          const float dxa = v[u].x + acc.co.cdx(F::fx(e[u]), P.hxf), dya = v[u].y - acc.co.yf;
          const float dxb = v[u + 1].x + acc.co.cdx(F::fx(e[u + 1]), P.hxf), dyb = v[u + 1].y - acc.co.yf;
          const float r2a = __builtin_fmaf(dxa, dxa, dya * dya), r2b = __builtin_fmaf(dxb, dxb, dyb * dyb);
          bool ina, inb;
          const bool ba = acc.rho_test(r2a, ina, P), bb = acc.rho_test(r2b, inb, P);
          if (ba | bb) {
            if (ba) ina = acc.rho_exact(src, F::index(e[u]), P);
            if (bb) inb = acc.rho_exact(src, F::index(e[u + 1]), P);
          }
          if (h == 0 && u == 0) acc.template term<SUM_GROUP_FIRST>(v[u], r2a, ina, P);
          else acc.template term<SUM_GROUP>(v[u], r2a, ina, P);
          acc.template term<SUM_GROUP>(v[u + 1], r2b, inb, P);
          const bool na = r2a <= P.xr2f_hi, nb = r2b <= P.xr2f_hi;
          if (na | nb) {
            // the entries from an opaque copy of their word, so that the common
            // path keeps its one-instruction decode of the high half
            uint32_t wq = w[2 * h + (u >> 1)];
            asm volatile("" : "+v"(wq));
            if (na) {
              if (qn == kCohQueue) {
                coh_drain(acc, src, cq, qn, P);
                qn = 0;
              }
              cq[qn * kBlock] = wq & 0xffffu;
              ++qn;
            }
            if (nb) {
              if (qn == kCohQueue) {
                coh_drain(acc, src, cq, qn, P);
                qn = 0;
              }
              cq[qn * kBlock] = wq >> 16;
              ++qn;
            }
          }
        }
        continue;
      }
    }
    acc.flush();
  }
  if constexpr (COH_QUEUE) coh_drain(acc, src, cq, qn, P);
}

// Chimera pass-2 walk with pass 1's within-rho decisions (one byte per group at
// mb): the staged tiles hold (cos, sin, aj, 0) of each halo particle, and no
// separation is formed.  Same terms, order and groups as walk_lds<2>.
template <class Acc>
__device__ __forceinline__ void walk_lds_rho_bits(Acc &acc, const int4 *l, const uint8_t *mb, int cnt) {
  using F = EntLds;
  if (cnt <= 0) return;
  int4 g = *l;
  uint32_t bits = *mb;
  for (int q = 0; q < cnt; q += 8) {
    const uint32_t w[4] = {(uint32_t)g.x, (uint32_t)g.y, (uint32_t)g.z, (uint32_t)g.w};
    const uint32_t b = bits;
    l += 64;
    mb += 64;
    if (q + 8 < cnt) {  // the lane's next group, if it has one
      g = *l;
      bits = *mb;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t e[4] = {w[2 * h] & 0xffffu, w[2 * h] >> 16, w[2 * h + 1] & 0xffffu, w[2 * h + 1] >> 16};
      float4 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = F::tile(F::off(e[u]));
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const bool in = (b >> (4 * h + u)) & 1u;
        if (h == 0 && u == 0)
          acc.template pair_m<SUM_GROUP_FIRST>(v[u], in);
        else
          acc.template pair_m<SUM_GROUP>(v[u], in);
      }
    }
    acc.flush();
  }
}

// FP64 forms: the entries of each group in turn.
template <class Acc, class S>
__device__ __forceinline__ void walk64(Acc &acc, const S &src, const int4 *l, int cnt, const KParams &P) {
  for (int q = 0; q < cnt; q += 4) {
    const int4 g = l[(size_t)(q >> 2) * 64];
    visit_entry(acc, src, g.x, P);
    if (q + 1 < cnt) visit_entry(acc, src, g.y, P);
    if (q + 2 < cnt) visit_entry(acc, src, g.z, P);
    if (q + 3 < cnt) visit_entry(acc, src, g.w, P);
  }
}

template <int PREC>
__device__ __forceinline__ typename SrcOf<PREC>::type verlet_src(const Tiles &t, const EpiCtx &e, int pass) {
  if constexpr (PREC == P32K) {
    return Src32{t.f4, nullptr, pass == 2 ? e.aphi_slot_f : nullptr};
  } else {
    return Src64{t.x, t.y, t.p, t.c, t.s, pass == 2 ? e.aphi_slot : nullptr};
  }
}

// Home of the thread: workgroups are dealt to the 8 XCDs round robin; each XCD
// takes a contiguous run of homes so that its L2 holds neighbouring tiles.
__device__ __forceinline__ int verlet_home() {
  return xcd_segment(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
}

template <class Acc, int PREC, int EPI>
__global__ __launch_bounds__(kBlock) void k_pass1_verlet(KParams P, const int *__restrict__ vcnt,
                                                         const int *__restrict__ wrow, const int *__restrict__ vlist,
                                                         Tiles t, const int *__restrict__ hid, const double *gp,
                                                         EpiCtx e, unsigned long long *counts) {
  const int i = verlet_home();
  const bool active = i < P.n;
  Pass1Out o{};
  int cnt = 0;
  if (active) {
    const auto src = verlet_src<PREC>(t, e, 1);
    Acc acc;
    acc.init(src, i, i, t, P);
    cnt = vcnt[i];
    const int4 *l = (const int4 *)vlist + (size_t)wrow[i >> 6] * 64 + (i & 63);
    if constexpr (PREC == P32K) {
      const int ri = yframe_rows(P, t.cell, i);
      acc.co.frame((float)ri * P.hyf);
      walk32<1>(acc, src, l, cnt, P, (float)(ri - 1) * P.hyf);
    } else {
      walk64(acc, src, l, cnt, P);
    }
    acc.finish(P, PREC == P32K ? gp[i] : 0.0, o);
    const EpiState r = epilogue1<EPI>(P, e, i, hid[i], i, o);
    if (EPI == EPI_K && e.next == 1) {
      // the next stage s + kcdt_n k, as k_verlet_stage forms it from k in memory
      write_stage<PREC>(P, e.tn, e.gpn, i, e.sx[i] + e.kcdt_n * r.ax, e.sy[i] + e.kcdt_n * r.ay,
                        e.sp[i] + e.kcdt_n * r.ap);
    } else if ((EPI == EPI_RK4_FINAL || EPI == EPI_SE) && e.next == 2) {
      write_stage<PREC>(P, e.tn, e.gpn, i, r.x, r.y, r.p);  // the next step's first stage
    }
  }
  if (counts) count_add(counts, cnt, active ? o.n : 0, active ? o.n2 : 0);
}

template <class Acc, int PREC, int EPI>
__global__ __launch_bounds__(kBlock) void k_pass2_verlet(KParams P, const int *__restrict__ vcnt,
                                                         const int *__restrict__ wrow, const int *__restrict__ vlist,
                                                         Tiles t, const int *__restrict__ hid, const double *gp,
                                                         EpiCtx e) {
  const int i = verlet_home();
  if (i >= P.n) return;
  const auto src = verlet_src<PREC>(t, e, 2);
  Pass2In in{e.aphi_slot[i], e.n_slot[i], e.n2_slot[i], 0.0, 0.0};
  if (EPI == EPI_ITO2) {
    in.in0 = e.a0[i];
    in.in1 = e.a1[i];
  }
  Acc acc;
  acc.init(src, i, i, t, P, in);
  const int cnt = vcnt[i];
  const int4 *l = (const int4 *)vlist + (size_t)wrow[i >> 6] * 64 + (i & 63);
  if constexpr (PREC == P32K) {
    const int ri = yframe_rows(P, t.cell, i);
    acc.co.frame((float)ri * P.hyf);
    walk32<2>(acc, src, l, cnt, P, (float)(ri - 1) * P.hyf);
  } else {
    walk64(acc, src, l, cnt, P);
  }
  Pass2Out o;
  acc.finish(P, PREC == P32K ? gp[i] : 0.0, o);
  const EpiState r = epilogue2<EPI>(P, e, i, hid[i], o);
  if (EPI == EPI_SRK && e.next == 2) write_stage<PREC>(P, e.tn, e.gpn, i, r.x, r.y, r.p);  // the next step's stage
}

// Stages workgroup seg's halo: FP32 tiles, for pass 2 the pass-1 phi
// derivatives, and the sentinel tile after the halo.  Returns the LDS sources
// (a tile's home, for the FP64 re-checks, comes from the runs); *tail is the
// dynamic area's LdsTail.
template <int PASS>
__device__ __forceinline__ Src32 stage_halo(const KParams &P, const int *start, const int *bh_nrun,
                                            const int *bh_run, const int *bh_off, int seg, const Tiles &t,
                                            const EpiCtx &e, void *dyn, int cap, LdsTail **tail) {
  float4 *f4l = (float4 *)dyn;
  float *ajl = (float *)(f4l + cap);
  LdsTail *tl = (LdsTail *)(ajl + (PASS == 2 ? cap : 0));
  *tail = tl;
  if (threadIdx.x == 0) tl->n = bh_nrun[seg];
  if (threadIdx.x <= kRunMax) {
    const int k = threadIdx.x, nrun = bh_nrun[seg];
    if (k <= nrun) tl->off[k] = bh_off[(size_t)seg * (kRunMax + 1) + k];
    if (k < nrun) {
      const int c = bh_run[(size_t)seg * 2 * kRunMax + 2 * k];  // the run's first cell
      tl->gst[k] = start[c];
      int d = c / P.nx - t.cell[seg * kBlock] / P.nx;  // its row less r0, unwrapped to -1 .. 2
      if (d > 2) d -= P.ny;
      else if (d < -1) d += P.ny;
      tl->ryf[k] = (float)d * P.hyf;
    }
  }
  __syncthreads();
  const int nrun = tl->n, total = tl->off[nrun];
  for (int el = threadIdx.x; el < total; el += kBlock) {
    int k = 0;
    while (k + 1 < nrun && tl->off[k + 1] <= el) ++k;
    const int g = tl->gst[k] + (el - tl->off[k]);
    float4 v = t.f4[g];
    v.y = v.y + tl->ryf[k];  // into the workgroup's y frame
    f4l[el] = v;
    if (PASS == 2) ajl[el] = e.aphi_slot_f[g];
  }
  if (threadIdx.x == 0) {  // the sentinel tile that padding entries name (total < cap)
    f4l[total] = make_float4(kSentinelCoord, kSentinelCoord, 0.0f, 0.0f);
    if (PASS == 2) ajl[total] = 0.0f;
  }
  __syncthreads();
  return Src32{f4l, nullptr, PASS == 2 ? ajl : nullptr, tl->gst, tl->off, nrun};
}

template <class Acc, int EPI, bool RHO_BITS = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_pass1_verlet_lds(KParams P, const int *__restrict__ vcnt,
                                                             const int *__restrict__ vslot,
                                                             const int *__restrict__ wrow,
                                                             const int *__restrict__ vlist, const int *start,
                                                             const int *bh_nrun, const int *bh_run,
                                                             const int *bh_off, Tiles t, const int *__restrict__ hid,
                                                             const double *gp, EpiCtx e,
                                                             unsigned long long *counts, uint8_t *rho_bits) {
  extern __shared__ float4 dyn_lds[];
  const int seg = xcd_segment(blockIdx.x, gridDim.x);
  LdsTail *tl;
  const Src32 src = stage_halo<1>(P, start, bh_nrun, bh_run, bh_off, seg, t, e, dyn_lds, P.lds_cap, &tl);
  const int slot = seg * kBlock + threadIdx.x;
  const bool active = slot < P.n;
  Pass1Out o{};
  int cnt = 0;
  if (active) {
    const int i = vslot ? vslot[slot] : slot;
    Acc acc;
    acc.init(Src32{t.f4, nullptr, nullptr}, i, i, t, P);
    acc.co.frame((float)yframe_rows(P, t.cell, i) * P.hyf);
    cnt = vcnt[i];
    const size_t at = (size_t)wrow[slot >> 6] * 64 + (slot & 63);
    constexpr bool coh = std::is_same<Acc, VortAcc<P32K>>::value;
    uint32_t *cq = (uint32_t *)(tl + 1) + threadIdx.x;  // the lane's cohesion queue column (vortex)
    walk_lds<1, RHO_BITS, coh>(acc, src, (const int4 *)vlist + at, cnt, P, RHO_BITS ? rho_bits + at : nullptr,
                               cq);
    acc.finish(P, gp[i], o);
    const EpiState r = epilogue1<EPI>(P, e, i, hid[i], i, o);
    if (EPI == EPI_K && e.next == 1) {
      write_stage<P32K>(P, e.tn, e.gpn, i, e.sx[i] + e.kcdt_n * r.ax, e.sy[i] + e.kcdt_n * r.ay,
                        e.sp[i] + e.kcdt_n * r.ap);
    } else if ((EPI == EPI_RK4_FINAL || EPI == EPI_SE) && e.next == 2) {
      write_stage<P32K>(P, e.tn, e.gpn, i, r.x, r.y, r.p);
    }
  }
  if (counts) count_add_lds(tl->part, counts, cnt, active ? o.n : 0, active ? o.n2 : 0);
}

template <class Acc, int EPI>
__global__ __launch_bounds__(kBlock) void k_pass2_verlet_lds(KParams P, const int *__restrict__ vcnt,
                                                             const int *__restrict__ vslot,
                                                             const int *__restrict__ wrow,
                                                             const int *__restrict__ vlist, const int *start,
                                                             const int *bh_nrun, const int *bh_run,
                                                             const int *bh_off, Tiles t, const int *__restrict__ hid,
                                                             const double *gp, EpiCtx e) {
  extern __shared__ float4 dyn_lds[];
  const int seg = xcd_segment(blockIdx.x, gridDim.x);
  LdsTail *tl;
  const Src32 src = stage_halo<2>(P, start, bh_nrun, bh_run, bh_off, seg, t, e, dyn_lds, P.lds_cap, &tl);
  const int slot = seg * kBlock + threadIdx.x;
  if (slot >= P.n) return;
  const int i = vslot ? vslot[slot] : slot;
  Pass2In in{e.aphi_slot[i], e.n_slot[i], e.n2_slot[i], 0.0, 0.0};
  if (EPI == EPI_ITO2) {
    in.in0 = e.a0[i];
    in.in1 = e.a1[i];
  }
  Acc acc;
  acc.init(Src32{t.f4, nullptr, nullptr}, i, i, t, P, in);
  acc.co.frame((float)yframe_rows(P, t.cell, i) * P.hyf);
  walk_lds<2>(acc, src, (const int4 *)vlist + (size_t)wrow[slot >> 6] * 64 + (slot & 63), vcnt[i], P);
  Pass2Out o;
  acc.finish(P, gp[i], o);
  const EpiState r = epilogue2<EPI>(P, e, i, hid[i], o);
  if (EPI == EPI_SRK && e.next == 2) write_stage<P32K>(P, e.tn, e.gpn, i, r.x, r.y, r.p);  // the next step's stage
}

// Stages workgroup seg's halo for the chimera pass 2 with pass 1's rho bits:
// (cos, sin, pass-1 phi derivative, 0) per halo particle, and a zero tile for
// the padding entries.
__device__ __forceinline__ void stage_halo_cs(const int *start, const int *bh_nrun, const int *bh_run,
                                              const int *bh_off, int seg, const Tiles &t, const EpiCtx &e,
                                              void *dyn, int cap, LdsTail **tail) {
  float4 *f4l = (float4 *)dyn;
  LdsTail *tl = (LdsTail *)(f4l + cap);
  *tail = tl;
  if (threadIdx.x == 0) tl->n = bh_nrun[seg];
  if (threadIdx.x <= kRunMax) {
    const int k = threadIdx.x, nrun = bh_nrun[seg];
    if (k <= nrun) tl->off[k] = bh_off[(size_t)seg * (kRunMax + 1) + k];
    if (k < nrun) tl->gst[k] = start[bh_run[(size_t)seg * 2 * kRunMax + 2 * k]];
  }
  __syncthreads();
  const int nrun = tl->n, total = tl->off[nrun];
  for (int el = threadIdx.x; el < total; el += kBlock) {
    int k = 0;
    while (k + 1 < nrun && tl->off[k + 1] <= el) ++k;
    const int g = tl->gst[k] + (el - tl->off[k]);
    const float4 v = t.f4[g];
    f4l[el] = make_float4(v.z, v.w, e.aphi_slot_f[g], 0.0f);
  }
  if (threadIdx.x == 0) f4l[total] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  __syncthreads();
}

template <int EPI>
__global__ __launch_bounds__(kBlock) void k_pass2_verlet_lds_rho_bits(
    KParams P, const int *__restrict__ vcnt, const int *__restrict__ vslot, const int *__restrict__ wrow,
    const int *__restrict__ vlist, const uint8_t *__restrict__ rho_bits, const int *start, const int *bh_nrun,
    const int *bh_run, const int *bh_off, Tiles t, const int *__restrict__ hid, const double *gp, EpiCtx e) {
  extern __shared__ float4 dyn_lds[];
  const int seg = xcd_segment(blockIdx.x, gridDim.x);
  LdsTail *tl;
  stage_halo_cs(start, bh_nrun, bh_run, bh_off, seg, t, e, dyn_lds, P.lds_cap, &tl);
  const int slot = seg * kBlock + threadIdx.x;
  if (slot >= P.n) return;
  const int i = vslot ? vslot[slot] : slot;
  Pass2In in{e.aphi_slot[i], e.n_slot[i], e.n2_slot[i], 0.0, 0.0};
  if (EPI == EPI_ITO2) {
    in.in0 = e.a0[i];
    in.in1 = e.a1[i];
  }
  Chim2Acc<P32K> acc;
  acc.init(Src32{t.f4, nullptr, nullptr}, i, i, t, P, in);
  const size_t at = (size_t)wrow[slot >> 6] * 64 + (slot & 63);
  walk_lds_rho_bits(acc, (const int4 *)vlist + at, rho_bits + at, vcnt[i]);
  Pass2Out o;
  acc.finish(P, gp[i], o);
  const EpiState r = epilogue2<EPI>(P, e, i, hid[i], o);
  if (EPI == EPI_SRK && e.next == 2) write_stage<P32K>(P, e.tn, e.gpn, i, r.x, r.y, r.p);  // the next step's stage
}

// ---- LDS-staged passes (FP64 forms) ---------------------------------------------
// The FP64 forms walk the LDS layout too: the same 16-bit halo indices, with the
// workgroup's halo staged as FP64 arrays (LdsTile: wrapped x, y and the phase,
// or cos and sin for the factored form; for pass 2 the pass-1 derivatives).
// Their pair arithmetic and order are those of the gather passes, so the sums
// are the same bit for bit; the list's padding entries are never read (the walk
// stops at the count).

template <int PREC, int PASS>
constexpr size_t lds_pass64_bytes(int cap) {
  return (size_t)cap * LdsTile<PREC, PASS>::bytes + sizeof(LdsTail);
}

template <int PREC, int PASS>
__device__ __forceinline__ Src64 stage_halo64(const int *start, const int *bh_nrun, const int *bh_run,
                                              const int *bh_off, int seg, const Tiles &t, const EpiCtx &e,
                                              void *dyn, int cap, LdsTail **tail) {
  LdsTile<PREC, PASS> lt(dyn, cap);
  LdsTail *tl = (LdsTail *)((char *)dyn + (size_t)cap * LdsTile<PREC, PASS>::bytes);
  *tail = tl;
  if (threadIdx.x == 0) tl->n = bh_nrun[seg];
  if (threadIdx.x <= kRunMax) {
    const int k = threadIdx.x, nrun = bh_nrun[seg];
    if (k <= nrun) tl->off[k] = bh_off[(size_t)seg * (kRunMax + 1) + k];
    if (k < nrun) tl->gst[k] = start[bh_run[(size_t)seg * 2 * kRunMax + 2 * k]];
  }
  __syncthreads();
  const int nrun = tl->n, total = tl->off[nrun];
  for (int el = threadIdx.x; el < total; el += kBlock) {
    int k = 0;
    while (k + 1 < nrun && tl->off[k + 1] <= el) ++k;
    lt.load(el, tl->gst[k] + (el - tl->off[k]), t, e);
  }
  __syncthreads();
  return lt.src();
}

// FP64 walk over an LDS list: the entries below the count, in list order.
template <class Acc>
__device__ __forceinline__ void walk64_lds(Acc &acc, const Src64 &src, const int4 *l, int cnt, const KParams &P) {
  for (int q = 0; q < cnt; q += 8) {
    const int4 g = l[(size_t)(q >> 3) * 64];
    const uint32_t w[4] = {(uint32_t)g.x, (uint32_t)g.y, (uint32_t)g.z, (uint32_t)g.w};
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (q + u < cnt) acc.pair(src, EntLds::index((u & 1) ? w[u >> 1] >> 16 : w[u >> 1] & 0xffffu), P);
  }
}

template <class Acc, int PREC, int EPI>
__global__ __launch_bounds__(kBlock) void k_pass1_verlet_lds64(KParams P, const int *__restrict__ vcnt,
                                                               const int *__restrict__ vslot,
                                                               const int *__restrict__ wrow,
                                                               const int *__restrict__ vlist, const int *start,
                                                               const int *bh_nrun, const int *bh_run,
                                                               const int *bh_off, Tiles t,
                                                               const int *__restrict__ hid, EpiCtx e,
                                                               unsigned long long *counts) {
  static_assert(PREC != P32K, "the FP32 form has its own LDS passes");
  extern __shared__ double dyn_lds64[];
  const int seg = xcd_segment(blockIdx.x, gridDim.x);
  LdsTail *tl;
  const Src64 src = stage_halo64<PREC, 1>(start, bh_nrun, bh_run, bh_off, seg, t, e, dyn_lds64, P.lds_cap, &tl);
  const int slot = seg * kBlock + threadIdx.x;
  const bool active = slot < P.n;
  Pass1Out o{};
  int cnt = 0;
  if (active) {
    const int i = vslot ? vslot[slot] : slot;
    Acc acc;
    acc.init(verlet_src<PREC>(t, e, 1), i, i, t, P);
    cnt = vcnt[i];
    walk64_lds(acc, src, (const int4 *)vlist + (size_t)wrow[slot >> 6] * 64 + (slot & 63), cnt, P);
    acc.finish(P, 0.0, o);
    const EpiState r = epilogue1<EPI>(P, e, i, hid[i], i, o);
    if (EPI == EPI_K && e.next == 1) {
      write_stage<PREC>(P, e.tn, e.gpn, i, e.sx[i] + e.kcdt_n * r.ax, e.sy[i] + e.kcdt_n * r.ay,
                        e.sp[i] + e.kcdt_n * r.ap);
    } else if ((EPI == EPI_RK4_FINAL || EPI == EPI_SE) && e.next == 2) {
      write_stage<PREC>(P, e.tn, e.gpn, i, r.x, r.y, r.p);
    }
  }
  if (counts) count_add_lds(tl->part, counts, cnt, active ? o.n : 0, active ? o.n2 : 0);
}

template <class Acc, int PREC, int EPI>
__global__ __launch_bounds__(kBlock) void k_pass2_verlet_lds64(KParams P, const int *__restrict__ vcnt,
                                                               const int *__restrict__ vslot,
                                                               const int *__restrict__ wrow,
                                                               const int *__restrict__ vlist, const int *start,
                                                               const int *bh_nrun, const int *bh_run,
                                                               const int *bh_off, Tiles t,
                                                               const int *__restrict__ hid, EpiCtx e) {
  static_assert(PREC != P32K, "the FP32 form has its own LDS passes");
  extern __shared__ double dyn_lds64[];
  const int seg = xcd_segment(blockIdx.x, gridDim.x);
  LdsTail *tl;
  const Src64 src = stage_halo64<PREC, 2>(start, bh_nrun, bh_run, bh_off, seg, t, e, dyn_lds64, P.lds_cap, &tl);
  const int slot = seg * kBlock + threadIdx.x;
  if (slot >= P.n) return;
  const int i = vslot ? vslot[slot] : slot;
  Pass2In in{e.aphi_slot[i], e.n_slot[i], e.n2_slot[i], 0.0, 0.0};
  if (EPI == EPI_ITO2) {
    in.in0 = e.a0[i];
    in.in1 = e.a1[i];
  }
  Acc acc;
  acc.init(verlet_src<PREC>(t, e, 2), i, i, t, P, in);
  walk64_lds(acc, src, (const int4 *)vlist + (size_t)wrow[slot >> 6] * 64 + (slot & 63), vcnt[i], P);
  Pass2Out o;
  acc.finish(P, 0.0, o);
  const EpiState r = epilogue2<EPI>(P, e, i, hid[i], o);
  if (EPI == EPI_SRK && e.next == 2) write_stage<PREC>(P, e.tn, e.gpn, i, r.x, r.y, r.p);  // the next step's stage
}

// Largest squared displacement since the build (min image), two-level and
// without atomics: per-workgroup maxima, then one workgroup.
__global__ __launch_bounds__(kBlock) void k_drift_partial(KParams P, const double *sx, const double *sy,
                                                          const double *xb, const double *yb, double *red) {
  __shared__ double sm[kBlock];
  double m = 0.0;
  for (int h = blockIdx.x * blockDim.x + threadIdx.x; h < P.n; h += gridDim.x * blockDim.x) {
    double dx = min_image(sx[h] - xb[h], P.Lx, P.rLx), dy = min_image(sy[h] - yb[h], P.Ly, P.rLy);
    m = fmax(m, dx * dx + dy * dy);
  }
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int d = kBlock / 2; d > 0; d >>= 1) {
    if ((int)threadIdx.x < d) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + d]);
    __syncthreads();
  }
  if (threadIdx.x == 0) red[blockIdx.x] = sm[0];
}

__global__ __launch_bounds__(kBlock) void k_drift_final(int nb, const double *red, double *out) {
  __shared__ double sm[kBlock];
  double m = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) m = fmax(m, red[b]);
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int d = kBlock / 2; d > 0; d >>= 1) {
    if ((int)threadIdx.x < d) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + d]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sm[0];
}

template <int PREC, int SYS, int EPI>
void vpass1_typed(Sim &s, const EpiCtx &e) {
  unsigned long long *cnt = s.counting ? s.counts : nullptr;
  const dim3 grid(blocks_for(s.n)), blk(kBlock);
  if constexpr (PREC == P32K) {
    if (s.vlds) {
      KParams kp = s.kp;
      kp.lds_cap = s.vlds_cap;
      const size_t sh = lds_pass_bytes(s.vlds_cap, 1);
      if constexpr (SYS == SPC2_CHIMERA && EPI == EPI_ITO1) {
        if (s.vrho) {  // the pass 2 that follows reads the rho decisions
          launch_pair(s, k_pass1_verlet_lds<ChimAcc<P32K>, EPI, true>, grid, blk, sh, kp, (const int *)s.vcnt,
                      (const int *)(s.vsorted ? s.vslot : nullptr), (const int *)s.wrow, (const int *)s.vlist,
                      (const int *)s.start, (const int *)s.bh_nrun, (const int *)s.bh_run, (const int *)s.bh_off,
                      s.t, (const int *)s.hid, (const double *)s.gp, e, cnt, s.vrho);
          s.vrho_fresh = true;
          return;
        }
      }
      if (SYS == SPC2_CHIMERA)
        launch_pair(s, k_pass1_verlet_lds<ChimAcc<P32K>, EPI>, grid, blk, sh, kp, (const int *)s.vcnt,
                    (const int *)(s.vsorted ? s.vslot : nullptr), (const int *)s.wrow, (const int *)s.vlist, (const int *)s.start,
                    (const int *)s.bh_nrun,
                    (const int *)s.bh_run, (const int *)s.bh_off, s.t, (const int *)s.hid, (const double *)s.gp,
                    e, cnt, (uint8_t *)nullptr);
      else
        launch_pair(s, k_pass1_verlet_lds<VortAcc<P32K>, EPI>, grid, blk, sh + kCohQueueBytes, kp, (const int *)s.vcnt,
                    (const int *)(s.vsorted ? s.vslot : nullptr), (const int *)s.wrow, (const int *)s.vlist, (const int *)s.start,
                    (const int *)s.bh_nrun,
                    (const int *)s.bh_run, (const int *)s.bh_off, s.t, (const int *)s.hid, (const double *)s.gp,
                    e, cnt, (uint8_t *)nullptr);
      return;
    }
  } else if (s.vlds) {
    KParams kp = s.kp;
    kp.lds_cap = s.vlds_cap;
    const size_t sh = lds_pass64_bytes<PREC, 1>(s.vlds_cap);
    if (SYS == SPC2_CHIMERA)
      launch_pair(s, k_pass1_verlet_lds64<ChimAcc<PREC>, PREC, EPI>, grid, blk, sh, kp, (const int *)s.vcnt,
                  (const int *)(s.vsorted ? s.vslot : nullptr), (const int *)s.wrow, (const int *)s.vlist,
                  (const int *)s.start, (const int *)s.bh_nrun, (const int *)s.bh_run, (const int *)s.bh_off, s.t,
                  (const int *)s.hid, e, cnt);
    else
      launch_pair(s, k_pass1_verlet_lds64<VortAcc<PREC>, PREC, EPI>, grid, blk, sh, kp, (const int *)s.vcnt,
                  (const int *)(s.vsorted ? s.vslot : nullptr), (const int *)s.wrow, (const int *)s.vlist,
                  (const int *)s.start, (const int *)s.bh_nrun, (const int *)s.bh_run, (const int *)s.bh_off, s.t,
                  (const int *)s.hid, e, cnt);
    return;
  }
  if (SYS == SPC2_CHIMERA)
    launch_pair(s, k_pass1_verlet<ChimAcc<PREC>, PREC, EPI>, grid, blk, 0, s.kp, (const int *)s.vcnt,
                (const int *)s.wrow, (const int *)s.vlist, s.t, (const int *)s.hid,
                (const double *)s.gp, e, cnt);
  else
    launch_pair(s, k_pass1_verlet<VortAcc<PREC>, PREC, EPI>, grid, blk, 0, s.kp, (const int *)s.vcnt,
                (const int *)s.wrow, (const int *)s.vlist, s.t, (const int *)s.hid,
                (const double *)s.gp, e, cnt);
}

template <int PREC, int SYS>
void vpass1_epi(Sim &s, int epi, const EpiCtx &e) {
  switch (epi) {
    case EPI_K: vpass1_typed<PREC, SYS, EPI_K>(s, e); break;
    case EPI_ITO1: vpass1_typed<PREC, SYS, EPI_ITO1>(s, e); break;
    case EPI_SE: vpass1_typed<PREC, SYS, EPI_SE>(s, e); break;
    case EPI_RK4_FINAL: vpass1_typed<PREC, SYS, EPI_RK4_FINAL>(s, e); break;
  }
}

template <int PREC>
void vpass1_sys(Sim &s, int epi, const EpiCtx &e) {
  if (s.P.system == SPC2_CHIMERA)
    vpass1_epi<PREC, SPC2_CHIMERA>(s, epi, e);
  else
    vpass1_epi<PREC, SPC2_VORTEX>(s, epi, e);
}

template <int PREC, int SYS, int EPI>
void vpass2_typed(Sim &s, const EpiCtx &e) {
  const dim3 grid(blocks_for(s.n)), blk(kBlock);
  if constexpr (PREC == P32K) {
    if (s.vlds) {
      KParams kp = s.kp;
      kp.lds_cap = s.vlds_cap;
      if (SYS == SPC2_CHIMERA && s.vrho_fresh) {  // pass 1 left its rho decisions
        launch_pair(s, k_pass2_verlet_lds_rho_bits<EPI>, grid, blk, lds_pass_bytes(s.vlds_cap, 1), kp,
                    (const int *)s.vcnt, (const int *)(s.vsorted ? s.vslot : nullptr), (const int *)s.wrow,
                    (const int *)s.vlist, (const uint8_t *)s.vrho, (const int *)s.start, (const int *)s.bh_nrun,
                    (const int *)s.bh_run, (const int *)s.bh_off, s.t, (const int *)s.hid, (const double *)s.gp, e);
        return;
      }
      const size_t sh = lds_pass_bytes(s.vlds_cap, 2);
      if (SYS == SPC2_CHIMERA)
        launch_pair(s, k_pass2_verlet_lds<Chim2Acc<P32K>, EPI>, grid, blk, sh, kp, (const int *)s.vcnt,
                    (const int *)(s.vsorted ? s.vslot : nullptr), (const int *)s.wrow, (const int *)s.vlist, (const int *)s.start,
                    (const int *)s.bh_nrun,
                    (const int *)s.bh_run, (const int *)s.bh_off, s.t, (const int *)s.hid, (const double *)s.gp,
                    e);
      else
        launch_pair(s, k_pass2_verlet_lds<Vort2Acc<P32K>, EPI>, grid, blk, sh, kp, (const int *)s.vcnt,
                    (const int *)(s.vsorted ? s.vslot : nullptr), (const int *)s.wrow, (const int *)s.vlist, (const int *)s.start,
                    (const int *)s.bh_nrun,
                    (const int *)s.bh_run, (const int *)s.bh_off, s.t, (const int *)s.hid, (const double *)s.gp,
                    e);
      return;
    }
  } else if (s.vlds) {
    KParams kp = s.kp;
    kp.lds_cap = s.vlds_cap;
    const size_t sh = lds_pass64_bytes<PREC, 2>(s.vlds_cap);
    if (SYS == SPC2_CHIMERA)
      launch_pair(s, k_pass2_verlet_lds64<Chim2Acc<PREC>, PREC, EPI>, grid, blk, sh, kp, (const int *)s.vcnt,
                  (const int *)(s.vsorted ? s.vslot : nullptr), (const int *)s.wrow, (const int *)s.vlist,
                  (const int *)s.start, (const int *)s.bh_nrun, (const int *)s.bh_run, (const int *)s.bh_off, s.t,
                  (const int *)s.hid, e);
    else
      launch_pair(s, k_pass2_verlet_lds64<Vort2Acc<PREC>, PREC, EPI>, grid, blk, sh, kp, (const int *)s.vcnt,
                  (const int *)(s.vsorted ? s.vslot : nullptr), (const int *)s.wrow, (const int *)s.vlist,
                  (const int *)s.start, (const int *)s.bh_nrun, (const int *)s.bh_run, (const int *)s.bh_off, s.t,
                  (const int *)s.hid, e);
    return;
  }
  if (SYS == SPC2_CHIMERA)
    launch_pair(s, k_pass2_verlet<Chim2Acc<PREC>, PREC, EPI>, grid, blk, 0, s.kp, (const int *)s.vcnt,
                (const int *)s.wrow, (const int *)s.vlist, s.t, (const int *)s.hid,
                (const double *)s.gp, e);
  else
    launch_pair(s, k_pass2_verlet<Vort2Acc<PREC>, PREC, EPI>, grid, blk, 0, s.kp, (const int *)s.vcnt,
                (const int *)s.wrow, (const int *)s.vlist, s.t, (const int *)s.hid,
                (const double *)s.gp, e);
}

template <int PREC>
void vpass2_sys(Sim &s, int epi, const EpiCtx &e) {
  const bool ito = epi == EPI_ITO2;
  if (s.P.system == SPC2_CHIMERA) {
    if (ito) vpass2_typed<PREC, SPC2_CHIMERA, EPI_ITO2>(s, e);
    else vpass2_typed<PREC, SPC2_CHIMERA, EPI_SRK>(s, e);
  } else {
    if (ito) vpass2_typed<PREC, SPC2_VORTEX, EPI_ITO2>(s, e);
    else vpass2_typed<PREC, SPC2_VORTEX, EPI_SRK>(s, e);
  }
}

// True when every LDS pass kernel of one system has no static LDS, so that its
// dynamic area starts at LDS address 0, where EntLds::tile reads the staged
// tiles.
template <class A1, class A2>
bool lds_passes_unbased() {
  const void *kern[] = {
      reinterpret_cast<const void *>(&k_pass1_verlet_lds<A1, EPI_K>),
      reinterpret_cast<const void *>(&k_pass1_verlet_lds<A1, EPI_ITO1>),
      reinterpret_cast<const void *>(&k_pass1_verlet_lds<A1, EPI_SE>),
      reinterpret_cast<const void *>(&k_pass1_verlet_lds<A1, EPI_RK4_FINAL>),
      reinterpret_cast<const void *>(&k_pass2_verlet_lds<A2, EPI_ITO2>),
      reinterpret_cast<const void *>(&k_pass2_verlet_lds<A2, EPI_SRK>),
      // the chimera passes with rho decisions (harmless duplicates for vortex)
      reinterpret_cast<const void *>(&k_pass1_verlet_lds<ChimAcc<P32K>, EPI_ITO1, true>),
      reinterpret_cast<const void *>(&k_pass2_verlet_lds_rho_bits<EPI_ITO2>),
      reinterpret_cast<const void *>(&k_pass2_verlet_lds_rho_bits<EPI_SRK>),
  };
  for (const void *k : kern) {
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, k) != hipSuccess || a.sharedSizeBytes != 0) return false;
  }
  return true;
}

}  // namespace

bool verlet_lds_unbased(int system) {
  if (system == SPC2_CHIMERA) {
    static const bool ok = lds_passes_unbased<ChimAcc<P32K>, Chim2Acc<P32K>>();
    return ok;
  }
  static const bool ok = lds_passes_unbased<VortAcc<P32K>, Vort2Acc<P32K>>();
  return ok;
}

void launch_verlet_count(Sim &s, bool halo) {
  const int *hn = halo ? (const int *)s.bh_nrun : nullptr;
  const double r2 = s.rlist * s.rlist * (1.0 + 1e-12);
  const dim3 g(blocks_for(s.n)), b(kBlock);
  switch (s.prec) {
    case P64:
      hipLaunchKernelGGL(k_verlet_count<P64>, g, b, 0, s.stream, s.kp, (const int *)s.start, s.t, r2, s.vmask,
                         s.vmask_words, s.vcnt, s.wmax, s.wmax8, hn,
                         (const int *)s.bh_run, (const int *)s.bh_off);
      break;
    case P64F:
      hipLaunchKernelGGL(k_verlet_count<P64F>, g, b, 0, s.stream, s.kp, (const int *)s.start, s.t, r2, s.vmask,
                         s.vmask_words, s.vcnt, s.wmax, s.wmax8, hn,
                         (const int *)s.bh_run, (const int *)s.bh_off);
      break;
    default:
      hipLaunchKernelGGL(k_verlet_count<P32K>, g, b, 0, s.stream, s.kp, (const int *)s.start, s.t, r2, s.vmask,
                         s.vmask_words, s.vcnt, s.wmax, s.wmax8, hn,
                         (const int *)s.bh_run, (const int *)s.bh_off);
  }
}

void launch_verlet_fill(Sim &s) {
  const double r2 = s.rlist * s.rlist * (1.0 + 1e-12);
  const dim3 g(blocks_for(s.n)), b(kBlock);
  switch (s.prec) {
    case P64:
      hipLaunchKernelGGL(k_verlet_fill<P64>, g, b, 0, s.stream, s.kp, (const int *)s.start, s.t, r2,
                         (const uint32_t *)s.vmask, s.vmask_words, (const int *)s.wrow, s.vlist);
      break;
    case P64F:
      hipLaunchKernelGGL(k_verlet_fill<P64F>, g, b, 0, s.stream, s.kp, (const int *)s.start, s.t, r2,
                         (const uint32_t *)s.vmask, s.vmask_words, (const int *)s.wrow, s.vlist);
      break;
    default:
      hipLaunchKernelGGL(k_verlet_fill<P32K>, g, b, 0, s.stream, s.kp, (const int *)s.start, s.t, r2,
                         (const uint32_t *)s.vmask, s.vmask_words, (const int *)s.wrow, s.vlist);
  }
}

void launch_block_halo(Sim &s) {
  hipLaunchKernelGGL(k_block_halo, dim3(blocks_for(s.nblk, 64)), dim3(64), 0, s.stream, s.kp, (const int *)s.start,
                     (const int *)s.t.cell, s.nblk, s.bh_nrun, s.bh_run, s.bh_off, s.bh_max);
}

void launch_verlet_fill_lds(Sim &s) {
  const double r2 = s.rlist * s.rlist * (1.0 + 1e-12);
  const auto k = s.prec == P64 ? k_verlet_fill_lds<P64> : s.prec == P64F ? k_verlet_fill_lds<P64F> : k_verlet_fill_lds<P32K>;
  hipLaunchKernelGGL(k, dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream, s.kp, (const int *)s.start,
                     s.t, r2, (const uint32_t *)s.vmask, s.vmask_words,
                     (const int *)(s.vsorted ? s.vslot : nullptr), (const int *)s.wrow, (const int *)s.bh_nrun,
                     (const int *)s.bh_run, (const int *)s.bh_off, s.vlist);
}

int verlet_lds_tile_bytes(int prec) {
  // a staged halo particle's bytes in the larger (pass-2) form
  return prec == P64 ? LdsTile<P64, 2>::bytes : prec == P64F ? LdsTile<P64F, 2>::bytes : 20;
}

void launch_verlet_order(Sim &s) {
  hipLaunchKernelGGL(k_verlet_order, dim3(blocks_for(s.n)), dim3(kBlock), 0, s.stream, s.n, (const int *)s.vcnt,
                     s.vslot, s.wmax_s);
}

void launch_verlet_stage(Sim &s, const double *kx, const double *ky, const double *kp, double kcdt) {
  const dim3 g(blocks_for(s.n)), b(kBlock);
  s.vrho_fresh = false;  // new tiles: the rho decisions no longer hold
  switch (s.prec) {
    case P64:
      hipLaunchKernelGGL(k_verlet_stage<P64>, g, b, 0, s.stream, s.kp, s.sx, s.sy, s.sp, kx, ky, kp, kcdt, s.t, s.gp);
      break;
    case P64F:
      hipLaunchKernelGGL(k_verlet_stage<P64F>, g, b, 0, s.stream, s.kp, s.sx, s.sy, s.sp, kx, ky, kp, kcdt, s.t,
                         s.gp);
      break;
    default:
      hipLaunchKernelGGL(k_verlet_stage<P32K>, g, b, 0, s.stream, s.kp, s.sx, s.sy, s.sp, kx, ky, kp, kcdt, s.t,
                         s.gp);
  }
}

namespace {
// A pass launch that writes the next stage's tiles (s.vnext, set by
// next_stage): the launch context names the second tile set, and afterwards the
// written tiles become the current ones (cells are shared).
int take_next(Sim &s, EpiCtx &e) {
  const int next = s.vnext;
  s.vnext = 0;
  if (next) {
    e.tn = s.t2;
    e.gpn = s.gp2;
    e.kcdt_n = s.vnext_kcdt;
    e.next = next;
  }
  return next;
}
void swap_tiles(Sim &s) {
  std::swap(s.t.x, s.t2.x);
  std::swap(s.t.y, s.t2.y);
  std::swap(s.t.p, s.t2.p);
  std::swap(s.t.c, s.t2.c);
  std::swap(s.t.s, s.t2.s);
  std::swap(s.t.f4, s.t2.f4);
  std::swap(s.gp, s.gp2);
}
}  // namespace

void launch_verlet_pass1(Sim &s, int epi, double *ox, double *oy, double *op, const StepArgs &sa, double *const k1[3],
                         double *const k2[3], double *const k3[3]) {
  EpiCtx e = make_ctx(s, ox, oy, op, sa, k1, k2, k3);
  s.vrho_fresh = false;  // set again by a launch that writes the rho decisions
  const int next = take_next(s, e);
  switch (s.prec) {
    case P64: vpass1_sys<P64>(s, epi, e); break;
    case P64F: vpass1_sys<P64F>(s, epi, e); break;
    default: vpass1_sys<P32K>(s, epi, e);
  }
  if (next) swap_tiles(s);
}

void launch_verlet_pass2(Sim &s, int epi, double *a0, double *a1, const StepArgs &sa) {
  EpiCtx e = make_ctx(s, nullptr, nullptr, nullptr, sa, nullptr, nullptr, nullptr);
  e.a0 = a0;
  e.a1 = a1;
  const int next = take_next(s, e);
  switch (s.prec) {
    case P64: vpass2_sys<P64>(s, epi, e); break;
    case P64F: vpass2_sys<P64F>(s, epi, e); break;
    default: vpass2_sys<P32K>(s, epi, e);
  }
  if (next) swap_tiles(s);
}

void launch_drift(Sim &s, double *out) {
  const int nb = std::min(1024, std::max(1, blocks_for(s.n)));
  hipLaunchKernelGGL(k_drift_partial, dim3(nb), dim3(kBlock), 0, s.stream, s.kp, s.sx, s.sy, s.xb, s.yb, s.red);
  hipLaunchKernelGGL(k_drift_final, dim3(1), dim3(kBlock), 0, s.stream, nb, s.red, out);
}

}  // namespace spc2
