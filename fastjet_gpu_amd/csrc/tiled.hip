/*
 * tiled.hip — the (rapidity, phi)-tiled clustering engine for gfx950.
 *
 * Replaces the reference's grid engine (grid.cu): jets are binned into tiles of
 * side >= R in rapidity and phi (phi wraps around), so every pair closer than R
 * sits in the same or adjacent tiles (grid.cu:55-106), and after each step only
 * the 3x3 neighbourhoods of the jets involved are revisited (grid.cu:403-549).
 * Unlike the reference the rapidity range comes from the event itself (edge
 * tiles absorb anything beyond it), the tile lists have no fixed capacity, and
 * the decisions are exactly FastJet's N2Plain ones, ties included:
 *   - every jet carries its N2Plain slot number as a *label* (the slot it would
 *     occupy in N2Plain's compacted array); the physical storage never moves;
 *   - the argmin is the lexicographic (diJ, label) minimum, nearest-neighbour
 *     ties go to the lowest label, and N2Plain's compaction is a relabelling of
 *     the last jet (label n-1 -> label of the removed slot);
 *   - nearest neighbours are physical indices, so the relabelling needs no
 *     pointer fix-up.
 *
 * One wave (64 lanes) clusters one event at a time; persistent workgroups pull
 * events from the queue of LaunchArgs.  Per-jet records live in global memory,
 * laid out in tile order when the event starts (a tile's members, and a row of
 * adjacent tiles, are contiguous); per tile, the lexicographic (diJ, label)
 * minimum and the member segment descriptor live in LDS.
 *
 * Per step (n live jets):
 *   1. argmin over the tile minima (LDS loads, a per-lane tree, a DPP wave
 *      reduction); the winner's tie-break key also names its partner;
 *   2. the merge (E-scheme, kinematics of the new jet) and the structural
 *      updates (tile membership, relabelling) by lane 0;
 *   3. one gather of the 3x3 tiles around the two old positions and the new one:
 *      orphans (jets whose nearest neighbour was removed or merged), jets now
 *      closer to the new jet, and the new jet's own nearest neighbour;
 *   4. a 3x3 rescan per orphan (from the lanes' registers when the gather
 *      covers, or provably bounds, its neighbourhood);
 *   5. the minima of the gathered tiles recomputed with LDS atomic minima (diJ
 *      bits, then key, then the winner's index) from the members' entries in
 *      registers, or, for gathers of several chunks, incrementally from the
 *      entries that changed.
 */
#include <hip/hip_runtime.h>

#include "fjmi_internal.h"
#include "fjmi_kernels.h"
#include "fjmi_math.h"

namespace fjmi {

namespace tiled {

constexpr int W = 64;       /* one wave per event */
constexpr int MAXG = 32;    /* tiles gathered per step (3 x 3x3 + 1 <= 28) */
constexpr int MAXO = 128;   /* orphans listed per step (more: rescan in rounds) */
constexpr int BIG = 0x7fffffff;
/* The gather counter restarts (every tile's stamp cleared) when it reaches this
 * value, so a long event cannot wrap it onto stamps still in the table; a
 * restart within a step only makes the orphan rescans' coverage test (stamp ==
 * the step's gather) fail, i.e. take the fresh-gather path. */
#ifndef FJMI_TILED_STAMP_LIMIT
#define FJMI_TILED_STAMP_LIMIT 0xFFFFFF00u
#endif
constexpr unsigned kStampLimit = FJMI_TILED_STAMP_LIMIT;

/* A tile minimum's tie-break key: the minimum jet's label in the high half and
 * its nearest neighbour's index (0xFFFF: none) in the low half, offset so that
 * signed order is label order (labels are distinct; events hold at most
 * kTiledMaxN particles).  The merge reads the partner straight from the key. */
__device__ __forceinline__ int tkey(int label, int nn) {
  return (int)((((unsigned)label << 16) | ((unsigned)nn & 0xFFFFu)) ^ 0x80000000u);
}
__device__ __forceinline__ int tkey_nn(int key) {
  const int v = key & 0xFFFF;
  return v == 0xFFFF ? -1 : v;
}

/* (diJ, label) minimum of a tile: diJ as its bit pattern (diJ >= 0, so the bits
 * order like the values), the key tkey(label, nearest neighbour), and the jet's
 * physical index.  Empty tile: d = ~0, l = BIG, p = -1. */
struct TMin {
  unsigned long long d;
  int l;
  int p;
};

/* A list of gathered tiles and the prefix sums of their member counts. */
struct Gather {
  int tiles[MAXG];
  int pref[MAXG + 1];
  int gpos[W]; /* pool position of member j < W (further members: search pref) */
};

/* Per-workgroup state of the event in flight; MAXT = tiles per event (the
 * tile table size: fewer tiles need less LDS, so more workgroups fit a CU). */
template <int MAXT>
struct Shared {
  TMin tmin[MAXT];
  /* Changed-entry table of multi-chunk gathers: the nchg jets whose entry
   * changed in the step, with their diJ bits, key tkey(label, nearest
   * neighbour), physical index and tile. */
  unsigned long long mt_d[W];
  int mt_l[W];
  int mt_p[W];
  int mt_t[W];
  int nchg;
  int seg_start[MAXT], seg_count[MAXT];
  /* segment capacities, at most 0xFFFF: a tile never holds more than an
   * event's kTiledMaxN = 65535 jets (16 bits, so 320 tiles fit 12 workgroups
   * per CU) */
  unsigned short seg_cap[MAXT];
  Gather g[2]; /* [0]: the step's gather, [1]: orphan rescans */
  int orph[MAXO];
  int norph;
  /* the last steps' history entries, written to global memory W at a time */
  int hb_p1[W], hb_p2[W];
  double hb_dij[W];
  unsigned stamp[MAXT]; /* last gather that listed the tile (dedup) */
  int pool_top;
  int rebuild;
  int ev;
  ServeDesc sd; /* serving: the claimed event's descriptor */
  int ny, nphi;
  double y0, sy, sphi;
};

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ double uni_d(double x) {
  return __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(x)),
                          __builtin_amdgcn_readfirstlane(__double2loint(x)));
}

__device__ __forceinline__ double readlane_d(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l), __builtin_amdgcn_readlane(__double2loint(v), l));
}

template <int CTRL, int RM>
__device__ __forceinline__ double dpp_d(double v) {
  int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), CTRL, RM, 0xF, false);
  int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), CTRL, RM, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int CTRL, int RM>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(v, v, CTRL, RM, 0xF, false);
}
__device__ __forceinline__ double wave_min_d(double v) {
  v = __builtin_fmin(v, dpp_d<0xB1, 0xF>(v));
  v = __builtin_fmin(v, dpp_d<0x4E, 0xF>(v));
  v = __builtin_fmin(v, dpp_d<0x141, 0xF>(v));
  v = __builtin_fmin(v, dpp_d<0x140, 0xF>(v));
  v = __builtin_fmin(v, dpp_d<0x142, 0xA>(v));
  v = __builtin_fmin(v, dpp_d<0x143, 0xC>(v));
  return readlane_d(v, 63);
}
__device__ __forceinline__ int wave_min_i(int v) {
  v = min(v, dpp_i<0xB1, 0xF>(v));
  v = min(v, dpp_i<0x4E, 0xF>(v));
  v = min(v, dpp_i<0x141, 0xF>(v));
  v = min(v, dpp_i<0x140, 0xF>(v));
  v = min(v, dpp_i<0x142, 0xA>(v));
  v = min(v, dpp_i<0x143, 0xC>(v));
  return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int wave_max_i(int v) {
  v = max(v, dpp_i<0xB1, 0xF>(v));
  v = max(v, dpp_i<0x4E, 0xF>(v));
  v = max(v, dpp_i<0x141, 0xF>(v));
  v = max(v, dpp_i<0x140, 0xF>(v));
  v = max(v, dpp_i<0x142, 0xA>(v));
  v = max(v, dpp_i<0x143, 0xC>(v));
  return __builtin_amdgcn_readlane(v, 63);
}

/* Lexicographic (d, lab) minimum over the wave; returns the winner's lane or -1. */
__device__ __forceinline__ int wave_argmin(double d, int lab, double *dmin, int *lmin) {
  const double m = wave_min_d(d);
  unsigned long long bal = __ballot(d == m);
  int ls;
  if (__popcll(bal) == 1) {
    const int ln = __ffsll(bal) - 1;
    ls = __builtin_amdgcn_readlane(lab, ln);
    *dmin = m;
    *lmin = ls;
    return ls == BIG ? -1 : ln;
  }
  const int key = (d == m) ? lab : BIG;
  ls = wave_min_i(key);
  *dmin = m;
  *lmin = ls;
  if (ls == BIG) return -1;
  bal = __ballot(key == ls);
  return __ffsll(bal) - 1;
}

/* Inclusive prefix sum over the wave: row_shr 1,2,4,8 (zero beyond the row
 * start) gives the scan within each row of 16; row_bcast15 (rows 1,3) and
 * row_bcast31 (rows 2,3) carry the row totals. */
__device__ __forceinline__ int wave_incl_scan(int v, int) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, true);
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, true);
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, true);
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, true);
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);
  return v;
}

/* The workgroup is exactly one wavefront (64 threads), and a wavefront's
 * memory operations, global and LDS, take effect in program order for all its
 * lanes, so lanes hand data to each other through memory with no wait on the
 * memory system: a wavefront-scope fence (a compiler barrier) replaces the
 * workgroup barrier and fence, which would wait for every outstanding store to
 * complete.  (FJMI_TILED_WG_FENCES=1 builds the workgroup-scope form.) */
#ifndef FJMI_TILED_WG_FENCES
#define FJMI_TILED_WG_FENCES 0
#endif
__device__ __forceinline__ void wave_sync() {
#if FJMI_TILED_WG_FENCES
  __syncthreads();
#else
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#endif
}

struct Geo {
  int ny, nphi;
  double y0, sy, sphi;
  double rsy, rsphi; /* 1 / sy, 1 / sphi */
  float rnphi;       /* 1 / nphi */
  /* A tile boundary may fall either way within rounding of the reciprocals; the
   * tiles' side exceeds R by far more, so every 3x3 neighbourhood still holds
   * every jet within R of its centre tile's members. */
  __device__ int tile(double y, double phi) const {
    const double fy = (y - y0) * rsy;
    const int ty = fy < 0.0 ? 0 : (fy >= (double)ny ? ny - 1 : (int)fy);
    int tp = (int)(phi * rsphi);
    tp = tp >= nphi ? nphi - 1 : (tp < 0 ? 0 : tp);
    return ty * nphi + tp;
  }
  /* row of tile t: an FP32 quotient estimate (off by less than one for
   * t < 2^16), corrected once */
  __device__ int row(int t) const {
    const int ty = (int)(((float)t + 0.5f) * rnphi);
    const int tp = t - ty * nphi;
    return tp < 0 ? ty - 1 : (tp >= nphi ? ty + 1 : ty);
  }
  /* column tp + dp, |dp| <= 1, wrapped in phi */
  __device__ int wrap(int tp) const { return tp < 0 ? tp + nphi : (tp >= nphi ? tp - nphi : tp); }
};

/* The per-event views of the global scratch. */
struct Ev {
  JetRec *rec;  /* per-jet record: everything a gather or a rescan reads */
  double4 *p4;
  int *hist;
  unsigned short *lab2; /* label -> physical */
  unsigned short *pool; /* tile segments */
  int pool_size;
};

/* (rapidity, azimuth, weight) of a momentum, bit-identical to fjmi_rap,
 * fjmi_phi and fjmi_weight.  In the common case (finite nonzero px and py, a
 * normal positive log argument) both go through the branch-free cores in one
 * straight-line block, so that the two dependent chains interleave. */
__device__ __forceinline__ double4 kin_coords_t(double px, double py, double pz, double E, int p) {
  const double kt2 = px * px + py * py;
  const double w = fjmi_weight(kt2, p);
  const double apz = __builtin_fabs(pz);
  const double m2 = (E + pz) * (E - pz) - kt2;
  const double em2 = m2 > 0.0 ? m2 : 0.0;
  const double epz = E + apz;
  const double arg = (kt2 + em2) / (epz * epz);
  const unsigned long long ua = (unsigned long long)__double_as_longlong(arg);
  const double ax = __builtin_fabs(px), ay = __builtin_fabs(py);
  const bool fast = kt2 != 0.0 && ua >= 0x0010000000000000ull && ua < 0x7FF0000000000000ull &&
                    ax != 0.0 && ay != 0.0 && ax < __builtin_inf() && ay < __builtin_inf();
  if (fast) {
    double rap = 0.5 * fjmi_log_core(arg, 0);
    rap = pz > 0 ? -rap : rap;
    double phi = fjmi_atan2_core(py, px);
    phi = phi < 0.0 ? phi + FJMI_TWOPI : phi;
    phi = phi >= FJMI_TWOPI ? phi - FJMI_TWOPI : phi;
    return make_double4(rap, phi, w, 0.0);
  }
  return make_double4(fjmi_rap(pz, E, kt2), fjmi_phi(px, py, kt2), w, 0.0);
}

/* particle s of the event starting at base, from the caller's input rows */
__device__ __forceinline__ double4 load_in(const LaunchArgs &a, int64_t base, int s) {
  const double *q = a.in + (size_t)(base + s) * a.in_stride + a.in_off;
  return make_double4(q[0], q[1], q[2], q[3]);
}

__device__ __forceinline__ double diJ_of(double nnd, double w, int nn, double wnn) {
  return nnd * ((nn >= 0 && wnn < w) ? wnn : w);
}

/* Segment capacity = count + (count >> FJMI_TILED_SLACK_SHIFT) +
 * FJMI_TILED_SLACK_ADD: 50 % slack plus two.  Tests build with less slack (and
 * a small pool) so that segments relocate and the pool runs out;
 * FJMI_TILED_ON_REBUILD() lets them count the rebuilds that follow. */
#ifndef FJMI_TILED_SLACK_SHIFT
#define FJMI_TILED_SLACK_SHIFT 1
#endif
#ifndef FJMI_TILED_SLACK_ADD
#define FJMI_TILED_SLACK_ADD 2
#endif
#ifndef FJMI_TILED_ON_REBUILD
#define FJMI_TILED_ON_REBUILD() ((void)0)
#endif

/* Rebuilds every tile segment from the live jets (labels 0..n-1): counting
 * sort into the pool with slack.  Used when a relocation finds the pool
 * exhausted. */
template <class SH>
__device__ void rebuild_segments(SH &sh, const Ev &ev, int n, int ntiles, int lane) {
  if (lane == 0) FJMI_TILED_ON_REBUILD();
  for (int t = lane; t < ntiles; t += W) sh.seg_count[t] = 0;
  wave_sync();
  for (int l = lane; l < n; l += W) {
    const int p = ev.lab2[l];
    atomicAdd(&sh.seg_count[ev.rec[p].meta.z], 1);
  }
  wave_sync();
  int carry = 0;
  for (int t0 = 0; t0 < ntiles; t0 += W) {
    const int t = t0 + lane;
    const int c = t < ntiles ? sh.seg_count[t] : 0;
    const int cap = c + (c >> FJMI_TILED_SLACK_SHIFT) + FJMI_TILED_SLACK_ADD;
    const int incl = wave_incl_scan(cap, lane);
    if (t < ntiles) {
      sh.seg_start[t] = carry + incl - cap;
      sh.seg_cap[t] = (unsigned short)min(cap, 0xFFFF);
      sh.seg_count[t] = 0;
    }
    carry += __builtin_amdgcn_readlane(incl, 63);
  }
  if (lane == 0) {
    sh.pool_top = carry;
    sh.rebuild = 0;
  }
  wave_sync();
  for (int l = lane; l < n; l += W) {
    const int p = ev.lab2[l];
    const int t = ev.rec[p].meta.z;
    const int pos = sh.seg_start[t] + atomicAdd(&sh.seg_count[t], 1);
    ev.pool[pos] = p;
    ev.rec[p].meta.w = pos;
  }
  wave_sync();
}

template <class SH>
__device__ __forceinline__ int gathered(const SH &sh, const Gather &G, const Ev &ev, int ng, int j);

/* Recomputes the (diJ, label) minimum of tiles listed in G.tiles[0..ng) from
 * all their members. */
template <class SH>
__device__ void refresh_minima(SH &sh, const Gather &G, const Ev &ev, int ng, int lane) {
  if (lane < ng) {
    const int t = G.tiles[lane];
    sh.tmin[t].d = ~0ull;
    sh.tmin[t].l = BIG;
    sh.tmin[t].p = -1;
  }
  wave_sync();
  const int total = G.pref[ng];
  /* three passes over all members: the diJ minimum, then the lowest key among
   * members at that minimum, then its index (each pass must see the final result
   * of the previous one across every chunk of 64 members).  Up to two chunks
   * are loaded once and kept in registers; larger gathers reload per pass. */
  if (total <= 2 * W) {
    unsigned long long bits[2];
    int key[2], t[2], p[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) p[c] = gathered(sh, G, ev, ng, c * W + lane);
#pragma unroll
    for (int c = 0; c < 2; ++c)
      if (p[c] >= 0) {
        const int4 m = ev.rec[p[c]].meta;
        t[c] = m.z;
        key[c] = tkey(m.y, m.x);
        bits[c] = (unsigned long long)__double_as_longlong(ev.rec[p[c]].dij);
      }
#pragma unroll
    for (int c = 0; c < 2; ++c)
      if (p[c] >= 0) atomicMin(&sh.tmin[t[c]].d, bits[c]);
    wave_sync();
#pragma unroll
    for (int c = 0; c < 2; ++c)
      if (p[c] >= 0 && sh.tmin[t[c]].d == bits[c]) atomicMin(&sh.tmin[t[c]].l, key[c]);
    wave_sync();
#pragma unroll
    for (int c = 0; c < 2; ++c)
      if (p[c] >= 0 && sh.tmin[t[c]].d == bits[c] && sh.tmin[t[c]].l == key[c]) sh.tmin[t[c]].p = p[c];
    wave_sync();
    return;
  }
  for (int phase = 0; phase < 3; ++phase) {
    for (int j0 = 0; j0 < total; j0 += W) {
      const int p = gathered(sh, G, ev, ng, j0 + lane);
      if (p >= 0) {
        const int4 m = ev.rec[p].meta;
        const int t = m.z;
        const unsigned long long bits = (unsigned long long)__double_as_longlong(ev.rec[p].dij);
        if (phase == 0) {
          atomicMin(&sh.tmin[t].d, bits);
        } else if (sh.tmin[t].d == bits) {
          const int key = tkey(m.y, m.x);
          if (phase == 1)
            atomicMin(&sh.tmin[t].l, key);
          else if (sh.tmin[t].l == key)
            sh.tmin[t].p = p;
        }
      }
    }
    wave_sync();
  }
}

/* Tile minima of the gathered tiles (gathers of at most 64 members) from the
 * members' entries in their lanes' registers: lane j < total holds member j
 * (physical myq in tile mytile).  The lexicographic (diJ, key) minimum per tile
 * in three passes of LDS atomics, as refresh_minima does with loaded entries. */
template <class SH>
__device__ void refresh_from_regs(SH &sh, const Gather &G, int ng, int total, int lane, int myq, int mytile,
                                  int mykey, unsigned long long mybits) {
  if (lane < ng) {
    const int t = G.tiles[lane];
    sh.tmin[t].d = ~0ull;
    sh.tmin[t].l = BIG;
    sh.tmin[t].p = -1;
  }
  wave_sync();
  const bool mem = lane < total;
  if (mem) atomicMin(&sh.tmin[mytile].d, mybits);
  wave_sync();
  const bool atd = mem && sh.tmin[mytile].d == mybits;
  if (atd) atomicMin(&sh.tmin[mytile].l, mykey);
  wave_sync();
  if (atd && sh.tmin[mytile].l == mykey) sh.tmin[mytile].p = myq;
  wave_sync();
}

template <class SH>
__device__ int gather_tiles(SH &sh, Gather &G, int cand, int lane, unsigned &stampv, int *total = nullptr);

/* Tile minima of a several-chunk gather, incrementally.  The step changed the
 * entries of the jets listed in the member table (nchg of them: jets that took
 * the merged jet as their neighbour, the relabelled jet, the rescanned orphans,
 * the merged jet in its new tile) and removed phys_hi.  A tile whose minimum
 * jet is one of these is recomputed from all its members (GF gathers them):
 * its entry may have risen (an orphan's neighbour is farther; for kt the
 * merged jet's weight is larger) or left the tile.  In every other tile the
 * old minimum still holds against the unchanged members, so the listed
 * entries enter the tiles' minima by comparison. */
template <class SH>
__device__ void refresh_changed(SH &sh, const Gather &G, Gather &GF, const Ev &ev, int ng, int phys_hi,
                                int lane, unsigned &stampv) {
  int cand = -1;
  if (lane < ng) {
    const int t = G.tiles[lane];
    const int pm = sh.tmin[t].p;
    bool full = pm >= 0 && pm == phys_hi;
    const int nc = sh.nchg;
    for (int i = 0; i < nc && !full; ++i) full = sh.mt_p[i] == pm;
    if (full) cand = t;
  }
  const int nf = gather_tiles(sh, GF, cand, lane, stampv);
  if (nf > 0) refresh_minima(sh, GF, ev, nf, lane);
  /* lane i holds listed entry i; the best entry of each tile (every lane
   * compares its entry with all listed ones) then meets the tile's minimum */
  const int nc = uni(sh.nchg);
  const bool mine = lane < nc;
  unsigned long long d = ~0ull;
  int k = BIG, t = -1, pe = -1;
  if (mine) {
    d = sh.mt_d[lane];
    k = sh.mt_l[lane];
    t = sh.mt_t[lane];
    pe = sh.mt_p[lane];
  }
  bool best = mine;
  for (int i = 0; i < nc; ++i) {
    const int ti = __builtin_amdgcn_readlane(t, i);
    const unsigned long long di = ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(d >> 32), i) << 32) |
                                  (unsigned)__builtin_amdgcn_readlane((int)d, i);
    const int ki = __builtin_amdgcn_readlane(k, i);
    /* another entry of the same tile that is smaller, or equal and listed first */
    if (ti == t && i != lane && (di < d || (di == d && (ki < k || (ki == k && i < lane))))) best = false;
  }
  if (best && (d < sh.tmin[t].d || (d == sh.tmin[t].d && k < sh.tmin[t].l))) {
    sh.tmin[t].d = d;
    sh.tmin[t].l = k;
    sh.tmin[t].p = pe;
  }
  wave_sync();
}

/* Sets G.tiles to the distinct union of the 3x3 neighbourhoods of the given
 * tiles (centres < 0 are skipped) plus `extra` (if >= 0), and G.pref to the
 * prefix sums of their member counts.  Lane i < 27 proposes neighbour i%9 of
 * centre i/9 (c0, c1, c2); duplicates are dropped through a per-tile stamp:
 * stampv, the wave's gather counter (kept in a register), is advanced and
 * written into every listed tile's stamp. */
template <class SH>
__device__ int build_gather(SH &sh, Gather &G, const Geo &geo, int c0, int c1, int c2, int extra, int lane,
                            unsigned &stampv, int *total = nullptr) {
  int cand = -1;
  if (lane < 27) {
    const int c = lane < 9 ? c0 : (lane < 18 ? c1 : c2);
    if (c >= 0) {
      const int ty = geo.row(c), tp = c - ty * geo.nphi;
      const int y = ty + (lane % 9) / 3 - 1;
      if (y >= 0 && y < geo.ny) cand = y * geo.nphi + geo.wrap(tp + (lane % 9) % 3 - 1);
    }
  } else if (lane == 27) {
    cand = extra;
  }
  return gather_tiles(sh, G, cand, lane, stampv, total);
}

/* The gather of the tiles the lanes propose (cand >= 0; duplicates dropped);
 * returns the number of tiles and, through `total`, of members. */
template <class SH>
__device__ int gather_tiles(SH &sh, Gather &G, int cand, int lane, unsigned &stampv, int *total) {
  if (stampv >= kStampLimit) { /* the counter starts again from zero before it could wrap */
    wave_sync();
    for (int t = lane; t < (int)(sizeof sh.stamp / sizeof sh.stamp[0]); t += W) sh.stamp[t] = 0u;
    stampv = 0;
  }
  stampv += 1;
  wave_sync(); /* the callers' LDS writes (segment counts) come first */
  bool keep = false;
  if (cand >= 0) keep = atomicExch(&sh.stamp[cand], stampv) != stampv;
  const unsigned long long kb = __ballot(keep);
  const int idx = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(kb >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)kb, 0u));
  const int cnt = keep ? sh.seg_count[cand] : 0;
  const int st = keep ? sh.seg_start[cand] : 0;
  const int incl = wave_incl_scan(cnt, lane);
  if (keep) {
    G.tiles[idx] = cand;
    G.pref[idx + 1] = incl;
    for (int u = 0, j = incl - cnt; u < cnt && j < W; ++u, ++j) G.gpos[j] = st + u;
  }
  if (lane == 0) G.pref[0] = 0;
  wave_sync();
  if (total) *total = __builtin_amdgcn_readlane(incl, 63);
  return (int)__popcll(kb);
}

/* Member j of the gathered tiles: physical index (or -1 past the end). */
template <class SH>
__device__ __forceinline__ int gathered(const SH &sh, const Gather &G, const Ev &ev, int ng, int j) {
  if (j >= G.pref[ng]) return -1;
  if (j < W) return ev.pool[G.gpos[j]];
  int lo = 0, hi = ng - 1; /* largest g with g_pref[g] <= j */
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (G.pref[mid] <= j)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int t = G.tiles[lo];
  return ev.pool[sh.seg_start[t] + (j - G.pref[lo])];
}

/* Appends physical p to tile t's segment; relocates the segment to the end of
 * the pool with twice the capacity when full, or returns true (a rebuild is
 * needed) when the pool is exhausted (p is then recorded only in its metadata).
 * The capacity stays at most 0xFFFF: an appended jet is the merged one, whose
 * parent lo is in another tile, so the segment then holds at most n - 1 <=
 * 65534 jets and a full segment has a capacity below 0xFFFF. */
template <class SH>
__device__ bool seg_insert(SH &sh, const Ev &ev, int t, int p) {
  if (sh.seg_count[t] == (int)sh.seg_cap[t]) {
    const int ncap = min(2 * (int)sh.seg_cap[t] + 2, 0xFFFF);
    if (sh.pool_top + ncap > ev.pool_size) {
      sh.rebuild = 1;
      return true;
    }
    const int ns = sh.pool_top;
    for (int i = 0; i < sh.seg_count[t]; ++i) {
      const int q = ev.pool[sh.seg_start[t] + i];
      ev.pool[ns + i] = q;
      ev.rec[q].meta.w = ns + i;
    }
    sh.seg_start[t] = ns;
    sh.seg_cap[t] = ncap;
    sh.pool_top += ncap;
  }
  const int pos = sh.seg_start[t] + sh.seg_count[t];
  ev.pool[pos] = p;
  ev.rec[p].meta.w = pos;
  sh.seg_count[t] += 1;
  return false;
}

/* Lexicographic (distance, label) nearest neighbour of physical o among the
 * gathered tiles (all lanes). Returns the winner's physical index or -1. */
template <int NCH, class SH>
__device__ int nn_scan(const SH &sh, const Gather &G, const Ev &ev, int ng, int o, double orap, double ophi, double R2,
                       int lane,
                       double *best_d, double *best_w, int *best_l = nullptr) {
  const int total = G.pref[ng];
  double bd = R2, bw = 0.0;
  int bl = BIG, bp = -1;
  /* NC chunks of 64 members at a time: their pool entries load together,
   * then their records (one chunk at a time where registers are short) */
  constexpr int NC = NCH;
  for (int j0 = 0; j0 < total; j0 += NC * W) {
    int q[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) q[c] = j0 + c * W < total ? gathered(sh, G, ev, ng, j0 + c * W + lane) : -1;
    double2 rp[NC];
    double wq[NC];
    int lab[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c)
      if (q[c] >= 0) {
        rp[c] = ev.rec[q[c]].rp;
        wq[c] = ev.rec[q[c]].wn.x;
        lab[c] = ev.rec[q[c]].meta.y;
      }
#pragma unroll
    for (int c = 0; c < NC; ++c)
      if (q[c] >= 0 && q[c] != o) {
        const double d = fjmi_dr2(orap, ophi, rp[c].x, rp[c].y);
        if (d < R2 && (d < bd || (d == bd && lab[c] < bl))) {
          bd = d;
          bl = lab[c];
          bp = q[c];
          bw = wq[c];
        }
      }
  }
  double md;
  int ml;
  const int ln = wave_argmin(bd, bl, &md, &ml);
  *best_d = md;
  if (best_l) *best_l = ml;
  if (ln < 0) {
    *best_w = 0.0;
    return -1;
  }
  *best_w = readlane_d(bw, ln);
  return __builtin_amdgcn_readlane(bp, ln);
}

/* Nearest neighbour (lexicographic (distance, label), within R) of jet o at
 * (oy, ophi) in tile otile over its 3x3 neighbourhood: its own tile first, then
 * only the neighbour tiles that could hold a jet at most as far as the best so
 * far (lane i < 9 bounds the distance to neighbour tile i from below by the
 * jet's distance to its own tile's edges on that side, less a margin far above
 * the rounding of the tile boundaries).  For dense events, whose 3x3
 * neighbourhoods span several chunks of 64; a sparse neighbourhood is one
 * chunk, scanned faster in one pass. */
template <int NCH, class SH>
__device__ int nn_rescan(SH &sh, Gather &GO, const Geo &geo, const Ev &ev, int o, int otile, double oy, double ophi,
                         double R2, int lane, unsigned &stampv, double *best_d, double *best_w) {
  int ngo = gather_tiles(sh, GO, lane == 0 ? otile : -1, lane, stampv);
  double d1, w1;
  int l1;
  const int n1 = nn_scan<NCH>(sh, GO, ev, ngo, o, oy, ophi, R2, lane, &d1, &w1, &l1);
  int cand = -1;
  if (lane < 9 && lane != 4) {
    const int ty = geo.row(otile), tp = otile - ty * geo.nphi, dy = lane / 3 - 1, dp = lane % 3 - 1;
    const int y = ty + dy;
    if (y >= 0 && y < geo.ny) {
      double by = dy < 0 ? oy - (geo.y0 + ty * geo.sy) : (dy > 0 ? geo.y0 + (ty + 1) * geo.sy - oy : 0.0);
      double bp = dp < 0 ? ophi - tp * geo.sphi : (dp > 0 ? (tp + 1) * geo.sphi - ophi : 0.0);
      by = dy != 0 ? (by - 1e-10 > 0.0 ? by - 1e-10 : 0.0) : 0.0;
      bp = dp != 0 ? (bp - 1e-10 > 0.0 ? bp - 1e-10 : 0.0) : 0.0;
      if (!(by * by + bp * bp > d1)) cand = y * geo.nphi + geo.wrap(tp + dp);
    }
  }
  if (__ballot(cand >= 0) == 0) {
    *best_d = d1;
    *best_w = w1;
    return n1;
  }
  ngo = gather_tiles(sh, GO, cand, lane, stampv);
  double d2, w2;
  int l2;
  const int n2 = nn_scan<NCH>(sh, GO, ev, ngo, o, oy, ophi, R2, lane, &d2, &w2, &l2);
  const bool take2 = n2 >= 0 && (n1 < 0 || d2 < d1 || (d2 == d1 && l2 < l1));
  *best_d = take2 ? d2 : d1;
  *best_w = take2 ? w2 : w1;
  return take2 ? n2 : n1;
}

}  // namespace tiled

using namespace tiled;

/* Serving: claims the next published event descriptor (lane 0 only).  Waves
 * claim through a device-memory head and a device-memory mirror of the host's
 * published count.  At most one wave per poll interval reads the host words over
 * PCIe (the poll lock, tested before it is taken) and refreshes the mirrors, and
 * idle waves back off to ~55 us between looks, so the idle waves' traffic on the
 * control words (each in its own 128-byte line) stays far below what would queue
 * behind it the busy waves' memory accesses.  Returns the descriptor index, or
 * -1 when the launch ends: the host asked to stop, or no wave has claimed
 * anything for idle_ticks while no event is being clustered.  The first wave to
 * see that latches the launch's epoch into the exit word, after which no wave of
 * the launch claims again, so the launch ends as a whole and the host's next
 * submit or poll relaunches it with every workgroup; a launch never shrinks to
 * the few waves that happened to be busy. */
__device__ __forceinline__ long long serve_claim(const ServeArgs &s) {
  unsigned long long *head = s.dev + kServeHead, *tail = s.dev + kServeTail, *lock = s.dev + kServeLock,
                     *stop = s.dev + kServeStop, *last_claim = s.dev + kServeLastClaim,
                     *last_poll = s.dev + kServeLastPoll, *exit_epoch = s.dev + kServeExitEpoch;
  constexpr unsigned long long kPollTicks = 100; /* 1 us of the 100 MHz wall clock */
  const unsigned long long t0 = wall_clock64();
  int nap = 1;
  for (;;) {
    if (__hip_atomic_load(exit_epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == s.epoch) return -1;
    const unsigned long long h = __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long t = __hip_atomic_load(tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    if (h < t) {
      if (atomicCAS(head, h, h + 1) == h) {
        __hip_atomic_fetch_max(last_claim, (unsigned long long)wall_clock64(), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long busy = atomicAdd(s.dev + kServeBusy, 1ull) + 1;
        __hip_atomic_fetch_max(s.dev + kServePeakBusy, busy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return (long long)h;
      }
      continue;
    }
    if (__hip_atomic_load(stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return -1;
    const unsigned long long now = wall_clock64();
    if (now - __hip_atomic_load(last_poll, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > kPollTicks &&
        __hip_atomic_load(lock, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0ull &&
        atomicCAS(lock, 0ull, 1ull) == 0ull) {
      __hip_atomic_store(last_poll, now, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned long long ht = __hip_atomic_load(&s.host->tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
      const unsigned int st = __hip_atomic_load(&s.host->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_fetch_max(tail, ht, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      if (st) __hip_atomic_store(stop, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(lock, 0ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      if (ht > h || st) {
        nap = 1;
        continue;
      }
    }
    const unsigned long long lc = __hip_atomic_load(last_claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (now - (lc > t0 ? lc : t0) > s.idle_ticks &&
        __hip_atomic_load(s.dev + kServeBusy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0ull) {
      __hip_atomic_store(exit_epoch, s.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return -1;
    }
    for (int k = 0; k < nap; ++k) __builtin_amdgcn_s_sleep(64);
    nap = nap < 32 ? 2 * nap : 32;
  }
}

/* Serving: copies an event's history and jets from the device arenas to the
 * host result arrays with wide coalesced stores, then announces the slot on the
 * completion ring (after a system-scope fence, so the host sees the results
 * first).  njets = -1 marks an event whose clustering failed. */
__device__ __forceinline__ void serve_finish(const LaunchArgs &a, int slot, int64_t base, int N, int nj, int lane) {
  const ServeArgs &s = a.serve;
  wave_sync();
  for (int i = lane; i < N; i += W) {
    s.res_p1[base + i] = a.hist_p1[base + i];
    s.res_p2[base + i] = a.hist_p2[base + i];
    s.res_dij[base + i] = a.hist_dij[base + i];
  }
  const int nd = nj > 0 ? 4 * nj : 0;
  for (int i = lane; i < nd; i += W) s.res_jets[4 * base + i] = a.jets[4 * base + i];
  if (lane == 0) s.res_nj[slot] = nj;
  __threadfence_system();
  wave_sync();
  if (lane == 0) {
    atomicSub(s.dev + kServeBusy, 1ull);
    const unsigned long long c = atomicAdd(s.dev + kServeCompleted, 1ull);
    __hip_atomic_store(s.comp + (unsigned)(c % s.ring_cap), slot + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

/* MAXT tiles per event; at least WPE waves per SIMD (the register budget:
 * 512 / WPE VGPRs per lane).  SERVE: events come from the host-published
 * descriptor ring of LaunchArgs::serve instead of a fixed queue, and each
 * event's results are copied to host memory and announced on completion. */
template <int MAXT, int WPE, bool SERVE>
__global__ __launch_bounds__(64, WPE) void tiled_kernel(LaunchArgs a, TiledScratch ts) {
  static_assert(MAXT % W == 0, "the argmin loads MAXT / W tile minima per lane");
  __shared__ Shared<MAXT> sh;
  const int lane = threadIdx.x;
  unsigned stampv = 0; /* gather counter; the stamps are cleared below it at each event */
  int pass = 0;        /* a launch without a queue head (one event) takes queue entry 0 and leaves */
  for (;;) {
    wave_sync();
    if (lane == 0) {
      if constexpr (SERVE) {
        const long long h = serve_claim(a.serve);
        sh.ev = h < 0 ? -1 : 0;
        if (h >= 0) {
          ServeDesc *d = const_cast<ServeDesc *>(a.serve.ring) + (unsigned long long)h % a.serve.ring_cap;
          const unsigned want = serve_seq((unsigned long long)h);
          const unsigned got = __hip_atomic_load(&d->seq, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
          sh.sd.base = d->base;
          sh.sd.pool_off = d->pool_off;
          sh.sd.n = d->n;
          sh.sd.p = d->p;
          sh.sd.R = d->R;
          sh.sd.slot = d->slot;
          if (got != want) { /* the host broke the ring protocol: fail the launch's status, skip nothing */
            atomicOr(a.status, (unsigned)STATUS_INTERNAL);
          }
          /* the cell may be written again once the copy above is complete */
          __hip_atomic_store(&d->seq, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
      } else {
        sh.ev = a.head ? (int)(atomicAdd(a.head, 1u) - a.head_base) : pass;
      }
    }
    ++pass;
    wave_sync();
    const int qi = uni(sh.ev);
    if (SERVE ? qi < 0 : (qi < 0 || qi >= a.n_queue)) {
      if (a.t_end && threadIdx.x == 0) atomicMax(a.t_end, (unsigned long long)wall_clock64());
      break;
    }
    int e, N, p;
    int64_t base, pool_off;
    double R;
    if constexpr (SERVE) {
      e = uni(sh.sd.slot);
      N = uni(sh.sd.n);
      p = uni(sh.sd.p);
      R = uni_d(sh.sd.R);
      base = ((int64_t)(unsigned)uni((int)(sh.sd.base >> 32)) << 32) | (unsigned)uni((int)sh.sd.base);
      pool_off = ((int64_t)(unsigned)uni((int)(sh.sd.pool_off >> 32)) << 32) | (unsigned)uni((int)sh.sd.pool_off);
    } else {
      e = a.queue ? uni(a.queue[qi]) : qi;
      base = a.offsets ? a.offsets[e] : 0;
      N = a.offsets ? uni((int)(a.offsets[e + 1] - base)) : a.n_single;
      p = a.p_arr ? a.p_arr[e] : a.p_uni;
      R = a.R_arr ? a.R_arr[e] : a.R_uni;
      pool_off = 3 * base + (int64_t)e * ts.pool_per_event;
    }
    if (N > kTiledMaxN) { /* the host rejects such events; never clustered here */
      if (lane == 0) atomicOr(a.status, (unsigned)STATUS_CAPACITY);
      if constexpr (SERVE) {
        serve_finish(a, e, base, 0, -1, lane);
      } else if (lane == 0) {
        a.njets[e] = -1; /* failed */
      }
      continue;
    }
    if (N <= 0) {
      if constexpr (SERVE) {
        serve_finish(a, e, base, 0, 0, lane);
      } else if (lane == 0) {
        a.njets[e] = 0;
      }
      continue;
    }
    const double R2 = R * R;
    const double invR2 = 1.0 / R2;
    for (int t = lane; t < MAXT; t += W) sh.stamp[t] = 0;
    stampv = 0;
    Ev ev;
    ev.rec = ts.rec + base;
    ev.p4 = reinterpret_cast<double4 *>(ts.p4) + base;
    ev.hist = ts.hist + base;
    ev.lab2 = ts.lab2 + base;
    ev.pool = ts.pool + pool_off;
    ev.pool_size = 3 * N + ts.pool_per_event;

    /* ---- prologue.  Jets are laid out in tile order: particle s (label s)
     * gets the physical index i = (members of lower tiles) + (its rank in
     * its tile), so that the members of a tile, and of a row of adjacent
     * tiles, are contiguous records.  Labels, not physical indices, decide
     * ties, so the layout does not change any result. ----
     * 1. kinematics into a staging copy indexed by s (the p4 array, which is
     *    written in tile order only in step 4), and the rapidity range */
    double4 *stage = ev.p4;
    /* loads in flight per lane in the prologue's passes (PB particles) and
     * scans (NB candidates); fewer where more waves share the registers */
    constexpr int PB = WPE <= 2 ? 4 : (WPE == 3 ? 2 : 1);
    constexpr int kScanChunks = WPE <= 2 ? 4 : 1; /* chunks per batch in the orphan rescans */
    constexpr int NB = WPE <= 2 ? 8 : (WPE == 3 ? 4 : 2);
    double ylo = 1e300, yhi = -1e300;
    /* the next particle's momentum is loaded before this one's kinematics */
    double4 qn = make_double4(0.0, 0.0, 0.0, 0.0);
    if (lane < N) qn = load_in(a, base, lane);
    for (int s = lane; s < N; s += W) {
      const double4 q4 = qn;
      if (s + W < N) qn = load_in(a, base, s + W);
      const double4 c = kin_coords_t(q4.x, q4.y, q4.z, q4.w, p);
      stage[s] = make_double4(c.x, c.y, c.z, 0.0);
      if (c.x > -1e4 && c.x < 1e4) {
        ylo = c.x < ylo ? c.x : ylo;
        yhi = c.x > yhi ? c.x : yhi;
      }
    }
    {
      const double mn = wave_min_d(ylo);
      const double mx = -wave_min_d(-yhi);
      if (lane == 0) {
        /* side >= R: the 3x3 neighbourhood of a tile holds every jet within R of its members */
        const double side = R * ((a.tile_q > 4 ? a.tile_q : 4) * 0.25) * (1.0 + 0x1p-20);
        double lo = mn, hi = mx;
        if (!(lo <= hi)) {
          lo = 0.0;
          hi = 0.0;
        }
        int np = (int)(FJMI_TWOPI / side);
        np = np < 1 ? 1 : (np > MAXT ? MAXT : np);
        int nyc = (int)((hi - lo) / side) + 1;
        const int nymax = MAXT / np;
        nyc = nyc > nymax ? nymax : (nyc < 1 ? 1 : nyc);
        sh.ny = nyc;
        sh.nphi = np;
        sh.y0 = lo;
        sh.sy = (hi - lo) / nyc > side ? (hi - lo) / nyc : side;
        sh.sphi = FJMI_TWOPI / np;
      }
    }
    wave_sync();
    Geo geo;
    geo.ny = uni(sh.ny);
    geo.nphi = uni(sh.nphi);
    geo.y0 = uni_d(sh.y0);
    geo.sy = uni_d(sh.sy);
    geo.sphi = uni_d(sh.sphi);
    geo.rsy = 1.0 / geo.sy;
    geo.rsphi = 1.0 / geo.sphi;
    geo.rnphi = 1.0f / (float)geo.nphi;
    const int ntiles = geo.ny * geo.nphi;
    for (int t = lane; t < ntiles; t += W) sh.seg_count[t] = 0;
    wave_sync();
    /* 2. each particle's tile and its rank among the tile's members (the
     *    order of equal-tile particles is immaterial: ties go by label) */
    for (int s0 = lane; s0 < N; s0 += PB * W) { /* PB particles' loads at once */
      double2 c[PB];
#pragma unroll
      for (int k = 0; k < PB; ++k)
        if (s0 + k * W < N) c[k] = make_double2(stage[s0 + k * W].x, stage[s0 + k * W].y);
#pragma unroll
      for (int k = 0; k < PB; ++k)
        if (s0 + k * W < N) {
          const int t = geo.tile(c[k].x, c[k].y);
          const int rank = atomicAdd(&sh.seg_count[t], 1);
          stage[s0 + k * W].w = __hiloint2double(t, rank);
        }
    }
    wave_sync();
    /* 3. segments with 50% slack in the pool, and each tile's first physical
     *    index (kept in tmin[t].p until the tile minima are set) */
    {
      int carry = 0, pcarry = 0;
      for (int t0 = 0; t0 < ntiles; t0 += W) {
        const int t = t0 + lane;
        const int c = t < ntiles ? sh.seg_count[t] : 0;
        const int cap = c + (c >> FJMI_TILED_SLACK_SHIFT) + FJMI_TILED_SLACK_ADD;
        const int incl = wave_incl_scan(cap, lane);
        const int pincl = wave_incl_scan(c, lane);
        if (t < ntiles) {
          sh.seg_start[t] = carry + incl - cap;
          sh.seg_cap[t] = (unsigned short)min(cap, 0xFFFF);
          sh.tmin[t].p = pcarry + pincl - c;
        }
        carry += __builtin_amdgcn_readlane(incl, 63);
        pcarry += __builtin_amdgcn_readlane(pincl, 63);
      }
      if (lane == 0) {
        sh.pool_top = carry;
        sh.rebuild = 0;
      }
    }
    wave_sync();
    /* 4. the records in tile order */
    for (int s0 = lane; s0 < N; s0 += PB * W) {
      double4 c[PB];
#pragma unroll
      for (int k = 0; k < PB; ++k)
        if (s0 + k * W < N) c[k] = stage[s0 + k * W];
#pragma unroll
      for (int k = 0; k < PB; ++k)
        if (s0 + k * W < N) {
          const int s = s0 + k * W;
          const int t = __double2hiint(c[k].w), rank = __double2loint(c[k].w);
          const int i = sh.tmin[t].p + rank;
          const int pos = sh.seg_start[t] + rank;
          ev.rec[i].rp = make_double2(c[k].x, c[k].y);
          ev.rec[i].wn = make_double2(c[k].z, R2);
          ev.rec[i].meta = make_int4(-1, s, t, pos);
          ev.pool[pos] = i;
          ev.lab2[s] = i;
          ev.hist[i] = s;
        }
    }
    wave_sync(); /* the staging copy is read in full before p4 is overwritten */
    for (int s = lane; s < N; s += W) ev.p4[ev.lab2[s]] = load_in(a, base, s);

    /* 5. initial nearest neighbours, one jet per lane in tile order.  The 3x3
     *    neighbourhood of a tile is, per tile row, one contiguous range of
     *    physical indices (two when the row wraps around in phi, the whole
     *    row when it has at most 3 columns). */
    for (int i = lane; i < N; i += W) {
      const double2 rp = ev.rec[i].rp;
      const double wi = ev.rec[i].wn.x;
      const int ti = ev.rec[i].meta.z;
      const int ty0 = geo.row(ti), tp0 = ti - ty0 * geo.nphi;
      double bd = R2;
      int bl = BIG, bp = -1;
      for (int r = 0; r < 6; ++r) {
        const int y = ty0 + r / 2 - 1;
        if (y < 0 || y >= geo.ny) continue;
        const int row = y * geo.nphi;
        int c0, c1; /* columns [c0, c1] */
        if (geo.nphi <= 3) {
          if (r & 1) continue;
          c0 = 0;
          c1 = geo.nphi - 1;
        } else if (tp0 == 0) {
          c0 = (r & 1) ? geo.nphi - 1 : 0;
          c1 = (r & 1) ? geo.nphi - 1 : 1;
        } else if (tp0 == geo.nphi - 1) {
          c0 = (r & 1) ? 0 : tp0 - 1;
          c1 = (r & 1) ? 0 : tp0;
        } else {
          if (r & 1) continue;
          c0 = tp0 - 1;
          c1 = tp0 + 1;
        }
        const int j0 = sh.tmin[row + c0].p, j1 = sh.tmin[row + c1].p + sh.seg_count[row + c1];
        /* candidates in blocks of NB, all of a block's loads issued together
         * (indices past the range repeat its last member and are masked) */
        for (int j = j0; j < j1; j += NB) {
          double2 rq[NB];
          int lq[NB];
#pragma unroll
          for (int k = 0; k < NB; ++k) {
            const int jk = min(j + k, j1 - 1);
            rq[k] = ev.rec[jk].rp;
            lq[k] = ev.rec[jk].meta.y;
          }
#pragma unroll
          for (int k = 0; k < NB; ++k) {
            const double d = fjmi_dr2(rp.x, rp.y, rq[k].x, rq[k].y);
            const bool take = (j + k < j1) & (j + k != i) & ((d < bd) | ((d == bd) & (d < R2) & (lq[k] < bl)));
            bd = take ? d : bd;
            bl = take ? lq[k] : bl;
            bp = take ? j + k : bp;
          }
        }
      }
      const double bw = bp >= 0 ? ev.rec[bp].wn.x : 0.0;
      ev.rec[i].meta.x = bp;
      ev.rec[i].wn = make_double2(wi, bd);
      ev.rec[i].dij = diJ_of(bd, wi, bp, bw);
    }
    wave_sync();
    /* 6. tile minima, one tile per lane over its contiguous members */
    for (int t = lane; t < ntiles; t += W) {
      const int j0 = sh.tmin[t].p, j1 = j0 + sh.seg_count[t];
      unsigned long long bd = ~0ull;
      int bl = BIG, bp = -1;
      for (int j = j0; j < j1; j += NB / 2) {
        unsigned long long dq[NB / 2];
        int lq[NB];
#pragma unroll
        for (int k = 0; k < NB / 2; ++k) {
          const int jk = min(j + k, j1 - 1);
          dq[k] = (unsigned long long)__double_as_longlong(ev.rec[jk].dij);
          lq[k] = tkey(ev.rec[jk].meta.y, ev.rec[jk].meta.x);
        }
#pragma unroll
        for (int k = 0; k < NB / 2; ++k) {
          const bool take = (j + k < j1) & ((dq[k] < bd) | ((dq[k] == bd) & (lq[k] < bl)));
          bd = take ? dq[k] : bd;
          bl = take ? lq[k] : bl;
          bp = take ? j + k : bp;
        }
      }
      sh.tmin[t].d = bd;
      sh.tmin[t].l = bl;
      sh.tmin[t].p = bp;
    }
    wave_sync();
    Gather &G = sh.g[0];
    Gather &GO = sh.g[1];

    /* ---- the N recombination steps ---- */
    int nj = 0;
    bool bad = false;
    int step = 0;
    for (; step < N; ++step) {
      if ((step & (W - 1)) == 0 && step > 0) { /* the previous W history entries, one coalesced store each */
        wave_sync();
        a.hist_p1[base + step - W + lane] = sh.hb_p1[lane];
        a.hist_p2[base + step - W + lane] = sh.hb_p2[lane];
        a.hist_dij[base + step - W + lane] = sh.hb_dij[lane];
      }
      const int n_old = N - step;
      /* the jet with the last label moves in this step (N2Plain compaction); its
       * index is loaded first so the load overlaps the argmin */
      const int tail = n_old - 1;
      const int pt_raw = ev.lab2[tail];
      /* 1. argmin over the tile minima: every lane loads all its MAXT/W entries
       * at once (always in bounds; entries past ntiles are masked), then a
       * pairwise tree of lexicographic (diJ, label) compares */
      constexpr int KT = MAXT / W;
      double kd[KT];
      int kl[KT], kt[KT], kp[KT];
      uint4 raw[KT];
#pragma unroll
      for (int k = 0; k < KT; ++k) raw[k] = *reinterpret_cast<const uint4 *>(&sh.tmin[k * W + lane]);
#pragma unroll
      for (int k = 0; k < KT; ++k) {
        const int l = (int)raw[k].z;
        const bool ok = (k * W + lane < ntiles) & (l != BIG);
        kd[k] = ok ? __hiloint2double((int)raw[k].y, (int)raw[k].x) : __builtin_inf();
        kl[k] = ok ? l : BIG;
        kt[k] = k * W + lane;
        kp[k] = (int)raw[k].w;
      }
#pragma unroll
      for (int w = 1; w < KT; w *= 2) {
#pragma unroll
        for (int k = 0; k + w < KT; k += 2 * w) {
          const bool c = (kd[k + w] < kd[k]) | ((kd[k + w] == kd[k]) & (kl[k + w] < kl[k]));
          kd[k] = c ? kd[k + w] : kd[k];
          kl[k] = c ? kl[k + w] : kl[k];
          kt[k] = c ? kt[k + w] : kt[k];
          kp[k] = c ? kp[k + w] : kp[k];
        }
      }
      const double bd = kd[0];
      const int bl = kl[0], bt = kt[0];
      double dmin;
      int lA;
      const int wl = wave_argmin(bd, bl, &dmin, &lA);
      const int tA = wl < 0 ? -1 : uni(__builtin_amdgcn_readlane(bt, wl));
      const int A = tA < 0 ? -1 : uni(__builtin_amdgcn_readlane(kp[0], wl));
      const int B = tA < 0 ? -1 : uni(tkey_nn(lA));
      if (A < 0) { /* no live jet found: inconsistent state, stop this event */
        if (lane == 0) atomicOr(a.status, 2u);
        bad = true;
        break;
      }
      /* 2. the merge.  The partner B comes with the tile minimum, so both
       * jets' records load at once.  Every value loaded for the merge is taken
       * into registers before the merge's first store: a wait for a load also
       * waits for every store issued before it. */
      const int4 mA = ev.rec[A].meta;
      const double4 qA = ev.p4[A];
      const int hA = ev.hist[A];
      int4 mB = make_int4(-1, -1, -1, -1);
      double4 qB = make_double4(0.0, 0.0, 0.0, 0.0);
      int hB = -1;
      if (B >= 0) {
        mB = ev.rec[B].meta;
        qB = ev.p4[B];
        hB = ev.hist[B];
      }
      const int pt = uni(pt_raw);
      const int tile_pt_raw = ev.rec[pt].meta.z;
      double4 qt_raw = make_double4(0.0, 0.0, 0.0, 0.0);
      if (a.pj_out) qt_raw = ev.p4[pt];
      const bool a_hi = B < 0 || uni(mA.y) > uni(mB.y);
      const int phys_hi = a_hi ? A : B;
      const int phys_lo = B < 0 ? -1 : (a_hi ? B : A);
      const int lab_hi = uni(a_hi ? mA.y : mB.y);
      const int tile_hi = uni(a_hi ? mA.z : mB.z);
      const int tile_lo_old = phys_lo < 0 ? -1 : uni(a_hi ? mB.z : mA.z);
      /* Segment removals swap the removed jet with its segment's last entry.  The
       * pool entries that may be needed are loaded now, by every lane, so that
       * their latency overlaps the new jet's kinematics:
       *   removal 1: phys_hi leaves tile_hi (last entry at last_hi);
       *   removal 2: phys_lo leaves tile_lo_old if the merged jet changes tile;
       *   when both are in one tile, removal 2 sees the segment as removal 1
       *   left it (one shorter, phys_lo possibly moved into phys_hi's place). */
      const int pos_hi = uni(a_hi ? mA.w : mB.w);
      const int pos_lo = phys_lo < 0 ? -1 : uni(a_hi ? mB.w : mA.w);
      const int cnt_hi = uni(sh.seg_count[tile_hi]);
      const int last_hi = uni(sh.seg_start[tile_hi]) + cnt_hi - 1;
      const int q_last_hi_raw = ev.pool[last_hi];
      const bool same_tile = phys_lo >= 0 && tile_lo_old == tile_hi;
      int last2 = -1, q_last2_raw = -1;
      if (phys_lo >= 0) {
        last2 = uni(sh.seg_start[tile_lo_old]) + uni(sh.seg_count[tile_lo_old]) - 1 - (same_tile ? 1 : 0);
        q_last2_raw = ev.pool[last2];
      }

      double nrap = 0.0, nphi = 0.0, nw = 0.0;
      int tile_new = -1;
      const double4 qhi = a_hi ? qA : qB;
      const int hhi = a_hi ? hA : hB;
      const double4 qlo = a_hi ? qB : qA;
      const int hlo = a_hi ? hB : hA;
      const double4 qn = make_double4(qhi.x + qlo.x, qhi.y + qlo.y, qhi.z + qlo.z, qhi.w + qlo.w);
      if (phys_lo >= 0) {
        const double4 c = kin_coords_t(qn.x, qn.y, qn.z, qn.w, p);
        nrap = c.x;
        nphi = c.y;
        nw = c.z;
        tile_new = uni(geo.tile(nrap, nphi));
      }
      const int tile_pt = uni(tile_pt_raw);
      const int q_last_hi = uni(q_last_hi_raw);
      const int q_last2_v = uni(q_last2_raw);
      const double4 qt = qt_raw;

      /* the stores: the history entry (to the LDS block) and the new jet's
       * record, then removal 1 and the N2Plain compaction (the last label
       * moves), then, if the merged jet changes tile, removal 2 (from the
       * segment as removal 1 left it) and its insertion.  The stores nothing
       * later in this step reads (the new jet's four-momentum and history
       * index, the beam jet, the removed jet's record, the label map, the
       * drop-in record) wait until the members' loads have been issued, so
       * that those loads do not wait for them (loads and stores share one
       * counter). */
      bool need_rebuild = false; /* lane 0's segment insertion exhausted the pool */
      const int jet_slot = nj;
      if (lane == 0) {
        if (phys_lo >= 0) {
          sh.hb_p1[step & (W - 1)] = min(hhi, hlo);
          sh.hb_p2[step & (W - 1)] = max(hhi, hlo);
          sh.hb_dij[step & (W - 1)] = dmin * invR2;
          ev.rec[phys_lo].rp = make_double2(nrap, nphi);
          ev.rec[phys_lo].wn = make_double2(nw, R2);
          ev.rec[phys_lo].meta.x = -1;
        } else {
          sh.hb_p1[step & (W - 1)] = hhi;
          sh.hb_p2[step & (W - 1)] = -1;
          sh.hb_dij[step & (W - 1)] = dmin * invR2;
        }
        if (pos_hi != last_hi) {
          ev.pool[pos_hi] = q_last_hi;
          ev.rec[q_last_hi].meta.w = pos_hi;
        }
        sh.seg_count[tile_hi] = cnt_hi - 1;
        if (pt != phys_hi) ev.rec[pt].meta.y = lab_hi;
        if (phys_lo >= 0 && tile_new != tile_lo_old) {
          const bool moved = same_tile && pos_lo == last_hi && pos_hi != last_hi;
          const int pos2 = moved ? pos_hi : pos_lo;
          const int q_last2 = (same_tile && last2 == pos_hi) ? q_last_hi : q_last2_v;
          if (pos2 != last2) {
            ev.pool[pos2] = q_last2;
            ev.rec[q_last2].meta.w = pos2;
          }
          sh.seg_count[tile_lo_old] -= 1;
          ev.rec[phys_lo].meta.z = tile_new;
          need_rebuild = seg_insert(sh, ev, tile_new, phys_lo);
        }
      }
      auto deferred_stores = [&]() {
        if (lane != 0) return;
        if (phys_lo >= 0) {
          ev.p4[phys_lo] = qn;
          ev.hist[phys_lo] = N + step;
        } else {
          double *jo = a.jets + 4 * (base + jet_slot);
          jo[0] = qhi.x;
          jo[1] = qhi.y;
          jo[2] = qhi.z;
          jo[3] = qhi.w;
        }
        ev.rec[phys_hi].meta = make_int4(-1, -1, -1, -1);
        if (pt != phys_hi) ev.lab2[lab_hi] = pt;
        if (a.pj_out) {
          unsigned char *rec = a.pj_out + 40 * (size_t)(base + tail);
          reinterpret_cast<int *>(rec)[0] = tail;
          reinterpret_cast<int *>(rec)[1] = phys_lo < 0 ? 1 : 0;
          double *dp = reinterpret_cast<double *>(rec + 8);
          const double4 o = phys_lo < 0 ? qhi : qt;
          dp[0] = o.x;
          dp[1] = o.y;
          dp[2] = o.z;
          dp[3] = o.w;
        }
      };
      if (phys_lo < 0) ++nj;
      wave_sync();
      const int n = n_old - 1;
      if (n == 0) {
        deferred_stores();
        ++step; /* this step's entry is recorded */
        break;
      }
      bool deferred_done = false;
      if (__builtin_amdgcn_readlane((int)need_rebuild, 0)) { /* pool exhausted by a relocation: lay all segments out again */
        deferred_stores(); /* the rebuild reads the label map */
        deferred_done = true;
        wave_sync();
        rebuild_segments(sh, ev, n, ntiles, lane);
      }

      /* 3. gather the 3x3 tiles around the old and new positions and the relabelled jet's tile */
      int total;
      const int ng = build_gather(sh, G, geo, tile_hi, tile_lo_old, tile_new, tile_pt, lane, stampv, &total);
      const unsigned main_stamp = stampv;
      if (lane == 0) {
        sh.norph = 0;
        sh.nchg = 0;
      }
      wave_sync();
      double cbd = R2, cbw = 0.0;
      int cbl = BIG, cbp = -1;
      /* a gather of at most 64 members keeps each member's (index, label, diJ)
       * in its lane for the tile-minimum refresh */
      const bool single = total <= W;
      int myq = -1, mylab = BIG, mytile = -1, mynn = -1;
      unsigned long long mybits = ~0ull;
      double2 myrp = make_double2(0.0, 0.0);
      double myw = 0.0;
      unsigned long long orph_lanes = 0; /* single-chunk gathers: lanes holding orphans */
      int q_next = gathered(sh, G, ev, ng, lane);
      for (int j0 = 0; j0 < total; j0 += W) {
        const int q = q_next;
        if (j0 + W < total) q_next = gathered(sh, G, ev, ng, j0 + W + lane); /* the next chunk's pool entries load meanwhile */
        bool orph = false;
        if (q >= 0) { /* every member's record, kept in its lane for the orphan rescans */
          const JetRec &r = ev.rec[q];
          const int4 mq0 = r.meta;
          myq = q;
          mylab = mq0.y;
          mytile = mq0.z;
          mynn = mq0.x;
          myrp = r.rp;
          myw = r.wn.x;
        }
        bool chg = false; /* multi-chunk gathers: the member's tile-minimum entry changed */
        int chg_key = 0, chg_tile = -1;
        if (q >= 0 && q != phys_lo) {
          int4 mq = ev.rec[q].meta;
          double dq = ev.rec[q].dij;
          orph = (mq.x == phys_hi) || (phys_lo >= 0 && mq.x == phys_lo);
          chg = q == pt && !orph; /* relabelled */
          if (phys_lo >= 0) {
            const double2 rq = ev.rec[q].rp;
            const double2 wq = ev.rec[q].wn;
            const double d = fjmi_dr2(rq.x, rq.y, nrap, nphi);
            if (!orph && d < wq.y) {
              dq = d * (nw < wq.x ? nw : wq.x);
              ev.rec[q].meta.x = phys_lo;
              ev.rec[q].wn = make_double2(wq.x, d);
              ev.rec[q].dij = dq;
              mynn = phys_lo;
              chg = true;
            }
            if (d < R2 && (d < cbd || (d == cbd && mq.y < cbl))) {
              cbd = d;
              cbl = mq.y;
              cbp = q;
              cbw = wq.x;
            }
          }
          mybits = (unsigned long long)__double_as_longlong(dq);
          chg_key = tkey(mq.y, mynn);
          chg_tile = mq.z;
        }
        if (!single) { /* list the changed entries for the incremental refresh */
          const unsigned long long cb = __ballot(chg);
          if (cb) {
            int cbase = 0;
            if (lane == 0) cbase = atomicAdd(&sh.nchg, (int)__popcll(cb));
            cbase = uni(cbase);
            const int idx = cbase + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(cb >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((unsigned)cb, 0u));
            if (chg && idx < W) {
              sh.mt_d[idx] = mybits;
              sh.mt_l[idx] = chg_key;
              sh.mt_p[idx] = q;
              sh.mt_t[idx] = chg_tile;
            }
          }
        }
        const unsigned long long ob = __ballot(orph);
        orph_lanes = ob;
        if (!single && ob) { /* single-chunk gathers keep their orphans in orph_lanes */
          int obase = 0;
          if (lane == 0) obase = atomicAdd(&sh.norph, (int)__popcll(ob));
          obase = uni(obase);
          const int idx = obase + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(ob >> 32),
                                                                 __builtin_amdgcn_mbcnt_lo((unsigned)ob, 0u));
          if (orph && idx < MAXO) sh.orph[idx] = q;
        }
      }
      if (!deferred_done) deferred_stores();
      if (phys_lo >= 0) { /* the new jet's own nearest neighbour */
        double md;
        int ml;
        const int ln = wave_argmin(cbd, cbl, &md, &ml);
        const int nnB = ln < 0 ? -1 : __builtin_amdgcn_readlane(cbp, ln);
        const double wB = ln < 0 ? 0.0 : readlane_d(cbw, ln);
        const double dB = diJ_of(md, nw, nnB, wB);
        if (lane == 0) {
          ev.rec[phys_lo].meta.x = nnB;
          ev.rec[phys_lo].wn = make_double2(nw, md);
          ev.rec[phys_lo].dij = dB;
        }
        if (myq == phys_lo) {
          mybits = (unsigned long long)__double_as_longlong(dB);
          mynn = nnB;
        }
        const int lab_lo = phys_lo == pt ? lab_hi : uni(a_hi ? mB.y : mA.y); /* after the relabelling */
        if (!single && lane == 0) {
          const int i = sh.nchg;
          sh.nchg = i + 1;
          if (i < W) {
            sh.mt_d[i] = (unsigned long long)__double_as_longlong(dB);
            sh.mt_l[i] = tkey(lab_lo, nnB);
            sh.mt_p[i] = phys_lo;
            sh.mt_t[i] = tile_new;
          }
        }
      }
      wave_sync();

      /* 4. orphan rescans over their own 3x3 tiles.  With a single-chunk gather
       * whose tiles hold the orphan's whole 3x3 neighbourhood, the candidates
       * are already in the lanes' registers: one distance per lane and one
       * wave argmin (gathered jets outside the neighbourhood are farther than
       * R and drop out).  Otherwise the neighbourhood is gathered afresh. */
      const int norph = single ? (int)__popcll(orph_lanes) : uni(sh.norph);
      if (norph > 0 && single) {
        unsigned long long pend = orph_lanes;
        while (pend) {
          const int lo = __ffsll(pend) - 1;
          pend &= pend - 1;
          const int o = __builtin_amdgcn_readlane(myq, lo);
          const int otile = __builtin_amdgcn_readlane(mytile, lo);
          const double oy = readlane_d(myrp.x, lo), ophi = readlane_d(myrp.y, lo), ow = readlane_d(myw, lo);
          /* the nearest gathered candidate, from the lanes' registers */
          const bool cand = myq >= 0 && myq != o;
          const double dc = cand ? fjmi_dr2(oy, ophi, myrp.x, myrp.y) : R2;
          const bool in = cand && dc < R2;
          double md;
          int ml;
          const int ln = wave_argmin(in ? dc : R2, in ? mylab : BIG, &md, &ml);
          /* It is the orphan's nearest neighbour unless a tile of its 3x3
           * neighbourhood that the gather missed could hold a jet at distance
           * <= md: lane i < 9 bounds from below the distance to neighbour tile
           * i (the orphan's distance to its own tile's edges on that side, less
           * a margin far above the rounding of the tile boundaries). */
          bool open = false;
          if (lane < 9) {
            const int ty = geo.row(otile), tp = otile - ty * geo.nphi, dy = lane / 3 - 1, dp = lane % 3 - 1;
            const int y = ty + dy;
            if (y >= 0 && y < geo.ny) {
              const int t = y * geo.nphi + geo.wrap(tp + dp);
              if (sh.stamp[t] != main_stamp) {
                double by = dy < 0 ? oy - (geo.y0 + ty * geo.sy) : (dy > 0 ? geo.y0 + (ty + 1) * geo.sy - oy : 0.0);
                double bp = dp < 0 ? ophi - tp * geo.sphi : (dp > 0 ? (tp + 1) * geo.sphi - ophi : 0.0);
                by = dy != 0 ? (by - 1e-10 > 0.0 ? by - 1e-10 : 0.0) : 0.0;
                bp = dp != 0 ? (bp - 1e-10 > 0.0 ? bp - 1e-10 : 0.0) : 0.0;
                open = !(by * by + bp * bp > md);
              }
            }
          }
          double od, owb;
          int nno;
          if (__ballot(open) == 0) {
            od = md;
            nno = ln < 0 ? -1 : __builtin_amdgcn_readlane(myq, ln);
            owb = ln < 0 ? 0.0 : readlane_d(myw, ln);
          } else {
            const int ngo = build_gather(sh, GO, geo, otile, -1, -1, -1, lane, stampv);
            nno = nn_scan<kScanChunks>(sh, GO, ev, ngo, o, oy, ophi, R2, lane, &od, &owb);
          }
          const double dO = diJ_of(od, ow, nno, owb);
          if (lane == 0) {
            ev.rec[o].meta.x = nno;
            ev.rec[o].wn = make_double2(ow, od);
            ev.rec[o].dij = dO;
          }
          if (lane == lo) {
            mybits = (unsigned long long)__double_as_longlong(dO);
            mynn = nno;
          }
        }
        wave_sync();
      } else if (norph > 0) {
        const int nlist = min(norph, MAXO);
        for (int i = 0; i < nlist; ++i) {
          const int o = uni(sh.orph[i]);
          const int4 mo = ev.rec[o].meta;
          const double2 ro = ev.rec[o].rp;
          const double wo = ev.rec[o].wn.x;
          double od, owb;
          const int nno = nn_rescan<kScanChunks>(sh, GO, geo, ev, o, mo.z, ro.x, ro.y, R2, lane, stampv, &od, &owb);
          const double dO = diJ_of(od, wo, nno, owb);
          if (lane == 0) {
            ev.rec[o].meta.x = nno;
            ev.rec[o].wn = make_double2(wo, od);
            ev.rec[o].dij = dO;
            const int c = sh.nchg;
            sh.nchg = c + 1;
            if (c < W) {
              sh.mt_d[c] = (unsigned long long)__double_as_longlong(dO);
              sh.mt_l[c] = tkey(mo.y, nno);
              sh.mt_p[c] = o;
              sh.mt_t[c] = mo.z;
            }
          }
          if (myq == o) {
            mybits = (unsigned long long)__double_as_longlong(dO);
            mynn = nno;
          }
          wave_sync();
        }
        if (norph > MAXO) { /* more orphans than the list holds: rescan the rest from the gather */
          const int ng2 = ng;
          const int tot2 = G.pref[ng2];
          for (int j0 = 0; j0 < tot2; j0 += W) {
            const int q = gathered(sh, G, ev, ng2, j0 + lane);
            const bool orph = q >= 0 && q != phys_lo &&
                              (ev.rec[q].meta.x == phys_hi || (phys_lo >= 0 && ev.rec[q].meta.x == phys_lo));
            unsigned long long ob = __ballot(orph);
            while (ob) {
              const int ln = __ffsll(ob) - 1;
              ob &= ob - 1;
              const int o = __builtin_amdgcn_readlane(q, ln);
              const int4 mo = ev.rec[o].meta;
              const double2 ro = ev.rec[o].rp;
              const double wo = ev.rec[o].wn.x;
              const int ngo = build_gather(sh, GO, geo, mo.z, -1, -1, -1, lane, stampv);
              double od, owb;
              const int nno = nn_scan<kScanChunks>(sh, GO, ev, ngo, o, ro.x, ro.y, R2, lane, &od, &owb);
              if (lane == 0) {
                ev.rec[o].meta.x = nno;
                ev.rec[o].wn = make_double2(wo, od);
                ev.rec[o].dij = diJ_of(od, wo, nno, owb);
              }
              wave_sync();
            }
          }
        }
      }

      /* 5. tile minima of the gathered tiles: from the member table (single
       * chunk), incrementally from the changed entries (several chunks), or
       * from all members */
      if (single && norph <= MAXO)
        refresh_from_regs(sh, G, ng, total, lane, myq, mytile, tkey(mylab, mynn), mybits);
      else if (norph <= MAXO && uni(sh.nchg) <= W)
        refresh_changed(sh, G, GO, ev, ng, phys_hi, lane, stampv);
      else
        refresh_minima(sh, G, ev, ng, lane);
    }
    { /* the history entries recorded since the last full block */
      wave_sync();
      const int done = step;
      const int first = (done - 1) & ~(W - 1);
      if (done > 0 && first + lane < done) {
        a.hist_p1[base + first + lane] = sh.hb_p1[lane];
        a.hist_p2[base + first + lane] = sh.hb_p2[lane];
        a.hist_dij[base + first + lane] = sh.hb_dij[lane];
      }
    }
    if constexpr (SERVE) {
      serve_finish(a, e, base, N, bad ? -1 : nj, lane);
    } else if (lane == 0) {
      a.njets[e] = bad ? -1 : nj; /* -1: the clustering failed */
    }
  }
}

/* Workgroups per CU, from 160 KiB of LDS and the register budget of the
 * waves per SIMD requested: 512 tiles, 2 waves (the full register budget):
 * 8 per CU; 384 tiles (25 rows of 15 at R = 0.4, |y| < 5), 3 waves: 10 per CU
 * (LDS-bound); 320 tiles, 3 waves: 12 per CU; 256 tiles, 3 waves: 12 per CU
 * (register-bound; more tiles per event than 192 at the same occupancy, and
 * fewer spills than at 4 waves). */
const void *const kTiledKernels[kNumTiledKernels] = {
    reinterpret_cast<const void *>(&tiled_kernel<512, 2, false>), reinterpret_cast<const void *>(&tiled_kernel<384, 3, false>),
    reinterpret_cast<const void *>(&tiled_kernel<320, 3, false>), reinterpret_cast<const void *>(&tiled_kernel<256, 3, false>)};
/* serving launches: the same variants, events from the host-published ring */
const void *const kTiledServeKernels[kNumTiledKernels] = {
    reinterpret_cast<const void *>(&tiled_kernel<512, 2, true>), reinterpret_cast<const void *>(&tiled_kernel<384, 3, true>),
    reinterpret_cast<const void *>(&tiled_kernel<320, 3, true>), reinterpret_cast<const void *>(&tiled_kernel<256, 3, true>)};
const int kTiledMaxTiles[kNumTiledKernels] = {512, 384, 320, 256};

}  // namespace fjmi
