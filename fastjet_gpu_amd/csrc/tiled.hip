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

/* Inclusive prefix maximum over the wave of values >= 0 (as wave_incl_scan,
 * with max for +). */
__device__ __forceinline__ int wave_incl_max(int v) {
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, true));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, true));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, true));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, true));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false));
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
template <bool GEN = false>
__device__ __forceinline__ double4 kin_coords_t(double px, double py, double pz, double E, int p, double pd = 0.0) {
  const double kt2 = px * px + py * py;
  /* GEN: the genkt weight for a real exponent pd (fjmi_weight_gen) */
  const double w = GEN ? fjmi_weight_gen(kt2, pd) : fjmi_weight(kt2, p);
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

/* ---- general jet definitions (the tiled_gen_kernel variants): any real
 * genkt exponent and FastJet's recombination schemes ---- */

/* an input particle as the recombination scheme takes it (fjmi_preprocess) */
__device__ __forceinline__ double4 preprocess4(int scheme, double4 q) {
  fjmi_preprocess(scheme, &q.x, &q.y, &q.z, &q.w);
  return q;
}

/* The merged jet of the parents with the higher label (qh, its rapidity and
 * azimuth rh) and the lower label (ql, rl): FastJet's N2Plain recombines jetA,
 * the higher slot, with jetB.  Its (rapidity, azimuth, weight) go to *kin, as
 * the scheme fixes them or from the merged momentum. */
__device__ __forceinline__ double4 merge_gen(int scheme, double pd, double4 qh, double2 rh, double4 ql, double2 rl,
                                             double4 *kin) {
  const double qa[4] = {qh.x, qh.y, qh.z, qh.w}, qb[4] = {ql.x, ql.y, ql.z, ql.w};
  double qo[4], y = 0.0, phi = 0.0;
  const int fixed = fjmi_recombine(scheme, qa, rh.x, rh.y, qb, rl.x, rl.y, qo, &y, &phi);
  const double4 q = make_double4(qo[0], qo[1], qo[2], qo[3]);
  if (fixed)
    *kin = make_double4(y, phi, fjmi_weight_gen(q.x * q.x + q.y * q.y, pd), 0.0);
  else
    *kin = kin_coords_t<true>(q.x, q.y, q.z, q.w, 0, pd);
  return q;
}

/* a jet definition the kernel can run: finite exponent, known scheme */
__device__ __forceinline__ bool jetdef_ok(double pd, int scheme) {
  return pd == pd && pd > -1e300 && pd < 1e300 && scheme >= 0 && scheme < FJMI_NUM_SCHEMES;
}

/* Segment capacity = count + (count >> FJMI_TILED_SLACK_SHIFT) +
 * FJMI_TILED_SLACK_ADD: 50 % slack plus two.  The CPU emulation of this source
 * defines these (and FJMI_TILED_STAMP_LIMIT) as its own run-time settings, and
 * runs with less slack and a small pool so that segments relocate and the pool
 * runs out; FJMI_TILED_ON_REBUILD() lets it count the rebuilds that follow. */
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
 * all their members; gathers of up to NK chunks are loaded once. */
template <int NK = 2, class SH>
__device__ void refresh_minima(SH &sh, const Gather &G, const Ev &ev, int ng, int lane) {
  if (ng == 1 && G.pref[1] > W) { /* one large tile: its minimum by a wave reduction, each member loaded once */
    const int total = G.pref[1];
    double bd = __builtin_inf();
    int bk = BIG, bp = -1;
    for (int j0 = 0; j0 < total; j0 += NK * W) {
      int p[NK];
#pragma unroll
      for (int c = 0; c < NK; ++c) p[c] = gathered(sh, G, ev, 1, j0 + c * W + lane);
      int4 m[NK];
      double dq[NK];
#pragma unroll
      for (int c = 0; c < NK; ++c)
        if (p[c] >= 0) {
          m[c] = ev.rec[p[c]].meta;
          dq[c] = ev.rec[p[c]].dij;
        }
#pragma unroll
      for (int c = 0; c < NK; ++c)
        if (p[c] >= 0) {
          const int key = tkey(m[c].y, m[c].x);
          if (dq[c] < bd || (dq[c] == bd && key < bk)) {
            bd = dq[c];
            bk = key;
            bp = p[c];
          }
        }
    }
    double md;
    int mk;
    const int ln = wave_argmin(bd, bk, &md, &mk);
    const int pw = ln < 0 ? -1 : __builtin_amdgcn_readlane(bp, ln);
    if (lane == 0) {
      const int t = G.tiles[0];
      sh.tmin[t].d = ln < 0 ? ~0ull : (unsigned long long)__double_as_longlong(md);
      sh.tmin[t].l = ln < 0 ? BIG : mk;
      sh.tmin[t].p = pw;
    }
    wave_sync();
    return;
  }
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
   * of the previous one across every chunk of 64 members).  Up to NK chunks
   * are loaded once and kept in registers; larger gathers reload per pass. */
  if (total <= NK * W) {
    unsigned long long bits[NK];
    int key[NK], t[NK], p[NK];
#pragma unroll
    for (int c = 0; c < NK; ++c) p[c] = gathered(sh, G, ev, ng, c * W + lane);
#pragma unroll
    for (int c = 0; c < NK; ++c)
      if (p[c] >= 0) {
        const int4 m = ev.rec[p[c]].meta;
        t[c] = m.z;
        key[c] = tkey(m.y, m.x);
        bits[c] = (unsigned long long)__double_as_longlong(ev.rec[p[c]].dij);
      }
#pragma unroll
    for (int c = 0; c < NK; ++c)
      if (p[c] >= 0) atomicMin(&sh.tmin[t[c]].d, bits[c]);
    wave_sync();
#pragma unroll
    for (int c = 0; c < NK; ++c)
      if (p[c] >= 0 && sh.tmin[t[c]].d == bits[c]) atomicMin(&sh.tmin[t[c]].l, key[c]);
    wave_sync();
#pragma unroll
    for (int c = 0; c < NK; ++c)
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
__device__ int gather_tiles(SH &sh, Gather &G, int cand, int lane, unsigned &stampv, int *total = nullptr, int *lpos = nullptr);

/* Tile minima of a several-chunk gather, incrementally.  The step changed the
 * entries of the jets listed in the member table (nchg of them: jets that took
 * the merged jet as their neighbour, the relabelled jet, the rescanned orphans,
 * the merged jet in its new tile) and removed phys_hi.  A tile whose minimum
 * jet is one of these is recomputed from all its members (GF gathers them):
 * its entry may have risen (an orphan's neighbour is farther; for kt the
 * merged jet's weight is larger) or left the tile.  In every other tile the
 * old minimum still holds against the unchanged members, so the listed
 * entries enter the tiles' minima by comparison. */
template <int NK = 2, class SH>
__device__ void refresh_changed(SH &sh, const Gather &G, Gather &GF, const Ev &ev, int ng, int phys_hi,
                                int lane, unsigned &stampv) {
  /* lane i holds listed entry i */
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
  int cand = -1;
  {
    const int tg = lane < ng ? G.tiles[lane] : 0;
    const int pm = lane < ng ? sh.tmin[tg].p : -1;
    bool full = pm >= 0 && pm == phys_hi;
    for (int i = 0; i < nc; ++i) {
      const int pi = __builtin_amdgcn_readlane(pe, i);
      full |= pm >= 0 && pi == pm;
    }
    if (lane < ng && full) cand = tg;
  }
  const int nf = gather_tiles(sh, GF, cand, lane, stampv);
  if (nf > 0) refresh_minima<NK>(sh, GF, ev, nf, lane);
  /* the best listed entry of each tile (every lane compares its entry with
   * all listed ones) then meets the tile's minimum */
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
                            unsigned &stampv, int *total = nullptr, int *lpos = nullptr) {
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
  return gather_tiles(sh, G, cand, lane, stampv, total, lpos);
}

/* The gather of the tiles the lanes propose (cand >= 0; duplicates dropped);
 * returns the number of tiles and, through `total`, of members, and through
 * `lpos` each lane's member's pool position (members 0..63, as G.gpos). */
template <class SH>
__device__ int gather_tiles(SH &sh, Gather &G, int cand, int lane, unsigned &stampv, int *total, int *lpos) {
  if (stampv >= FJMI_TILED_STAMP_LIMIT) { /* the counter starts again from zero before it could wrap */
    wave_sync();
    for (int t = lane; t < (int)(sizeof sh.stamp / sizeof sh.stamp[0]); t += W) sh.stamp[t] = 0u;
    stampv = 0;
  }
  stampv += 1;
  wave_sync(); /* the callers' LDS writes (segment counts) come first */
  const unsigned long long cb = __ballot(cand >= 0);
  if (cb == 0) { /* no tile */
    if (lane == 0) G.pref[0] = 0;
    wave_sync();
    if (total) *total = 0;
    return 0;
  }
  if (__popcll(cb) == 1) { /* one tile (an orphan's own tile, a refreshed tile): nothing to deduplicate */
    const int t = __builtin_amdgcn_readlane(cand, (int)__builtin_ctzll(cb));
    const int cnt = uni(sh.seg_count[t]), st = uni(sh.seg_start[t]);
    if (lane == 0) {
      sh.stamp[t] = stampv;
      G.tiles[0] = t;
      G.pref[0] = 0;
      G.pref[1] = cnt;
    }
    G.gpos[lane] = st + lane;
    if (lpos) *lpos = st + lane;
    wave_sync();
    if (total) *total = cnt;
    return 1;
  }
  bool keep = false;
  if (cand >= 0) keep = atomicExch(&sh.stamp[cand], stampv) != stampv;
  const unsigned long long kb = __ballot(keep);
  const int idx = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(kb >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)kb, 0u));
  const int cnt = keep ? sh.seg_count[cand] : 0;
  const int st = keep ? sh.seg_start[cand] : 0;
  const int incl = wave_incl_scan(cnt, lane);
  const int tot = __builtin_amdgcn_readlane(incl, 63);
  const int ngt = (int)__popcll(kb);
  /* The first 64 members' pool positions: lane j finds the tile of member j with
   * a prefix-maximum scan over the tiles' first members (each marked with its
   * proposing lane, plus one), rather than each tile's lane listing its members
   * one after another (a loop as long as the largest tile, and divergent); the
   * proposing lane's segment start then comes by a cross-lane read. */
  G.gpos[lane] = 0;
  if (keep) {
    G.tiles[idx] = cand;
    G.pref[idx + 1] = incl;
  }
  if (lane == 0) G.pref[0] = 0;
  wave_sync();
  if (keep && cnt > 0 && incl - cnt < W) G.gpos[incl - cnt] = lane + 1; /* the first member of each tile */
  wave_sync();
  const int src = wave_incl_max(G.gpos[lane]) - 1; /* the lane of the last tile starting at or before member `lane` */
  const int pos = __builtin_amdgcn_ds_bpermute(4 * (src < 0 ? 0 : src), st - (incl - cnt)) + lane;
  wave_sync();
  G.gpos[lane] = pos;
  if (lpos) *lpos = pos;
  wave_sync();
  if (total) *total = tot;
  return ngt;
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

/* ---- cooperative steps (tiled_coop_kernel): several waves per event ----
 * For large events a step's gathers span many chunks of 64 members, and one
 * wave issuing every chunk's instructions bounds the step.  In the cooperative
 * kernels wave 0 runs the event exactly as the one-wave kernels do, but at the
 * member pass over a gather of several chunks it posts a command in LDS, and
 * every wave of the workgroup takes chunks wid, wid + NW, ... of it.  Waves
 * meet at two workgroup barriers per command (posted, done).  The results are
 * the one-wave kernel's: per-member updates do not depend on each other, the
 * lists of changed entries and orphans are sets (their order decides nothing),
 * and the winner is the lexicographic (distance, label) minimum over all
 * candidates, whichever wave saw it.  The orphans' scans stay with wave 0: a
 * scan does less per member than the member pass, and sharing one cost more in
 * the barriers' waits than it saved (one event of N = 19996, R = 0.4: 190 ms
 * with wave 0 alone, 210 ms shared). */
constexpr int kCoopMaxWaves = 16;
enum : int { COOP_EXIT = 0, COOP_MEMBERS = 1 };
struct CoopCmd {
  int op;
  int ng, gsel;               /* the gather: tiles, and which (0: the step's) */
  int phys_hi, phys_lo, pt;   /* the step's parents, the relabelled jet */
  double nrap, nphi, nw, R2;  /* the merged jet */
  JetRec *rec;                /* the event's views */
  unsigned short *pool;
  /* each wave's lexicographic nearest candidate */
  double bd[kCoopMaxWaves], bw[kCoopMaxWaves];
  int bl[kCoopMaxWaves], bp[kCoopMaxWaves];
};

/* Wave wid's share of a member pass over a gather of several chunks (the
 * one-wave pass's multi-chunk path, chunk by chunk); its nearest candidate for
 * the merged jet goes to slot wid.  Member j's pool position: lane g < ng holds
 * gathered tile g's first member index and segment start; the tile of member j
 * is the last one starting at or before j, from the tiles starting before the
 * chunk (a ballot) and a prefix maximum over the starts inside it, marked in
 * the wave's 64 words of `mark`; the position then comes from that tile's lane
 * (a binary search over the tiles for every member took five dependent LDS
 * reads). */
template <class SH>
__device__ void coop_members(SH &sh, const Gather &G, const Ev &ev, const CoopCmd &c, int wid, int nwaves, int lane,
                             CoopCmd *out, int *mark) {
  const int total = G.pref[c.ng];
  const int phys_hi = c.phys_hi, phys_lo = c.phys_lo, pt = c.pt;
  double cbd = c.R2, cbw = 0.0;
  int cbl = BIG, cbp = -1;
  const int gs = lane < c.ng ? G.pref[lane] : BIG;                       /* tile `lane`'s first member */
  const int gofs = lane < c.ng ? sh.seg_start[G.tiles[lane]] - gs : 0; /* its pool position less that index */
  int *const wmark = mark + wid * W;
  for (int j0 = wid * W; j0 < total; j0 += nwaves * W) {
    const int g0 = (int)__popcll(__ballot(gs <= j0)) - 1; /* the tile holding member j0 */
    wmark[lane] = 0;
    wave_sync();
    if (gs > j0 && gs < j0 + W) atomicMax(&wmark[gs - j0], lane + 1);
    wave_sync();
    const int gm = wave_incl_max(wmark[lane]) - 1;
    const int own = gm > g0 ? gm : g0; /* the tile holding member j0 + lane */
    const int pos = __builtin_amdgcn_ds_bpermute(4 * own, gofs) + j0 + lane;
    const int q = j0 + lane < total ? (int)ev.pool[pos] : -1;
    bool orph = false, chg = false;
    int chg_key = 0, chg_tile = -1;
    unsigned long long bits = ~0ull;
    if (q >= 0 && q != phys_lo) {
      const int4 mq = ev.rec[q].meta;
      double dq = ev.rec[q].dij;
      int nn = mq.x;
      orph = (mq.x == phys_hi) || (phys_lo >= 0 && mq.x == phys_lo);
      chg = q == pt && !orph; /* relabelled */
      if (phys_lo >= 0) {
        const double2 rq = ev.rec[q].rp;
        const double2 wq = ev.rec[q].wn;
        const double d = fjmi_dr2(rq.x, rq.y, c.nrap, c.nphi);
        if (!orph && d < wq.y) {
          dq = d * (c.nw < wq.x ? c.nw : wq.x);
          ev.rec[q].meta.x = phys_lo;
          ev.rec[q].wn = make_double2(wq.x, d);
          ev.rec[q].dij = dq;
          nn = phys_lo;
          chg = true;
        }
        if (d < c.R2 && (d < cbd || (d == cbd && mq.y < cbl))) {
          cbd = d;
          cbl = mq.y;
          cbp = q;
          cbw = wq.x;
        }
      }
      bits = (unsigned long long)__double_as_longlong(dq);
      chg_key = tkey(mq.y, nn);
      chg_tile = mq.z;
    }
    const unsigned long long cb = __ballot(chg);
    if (cb) {
      int cbase = 0;
      if (lane == 0) cbase = atomicAdd(&sh.nchg, (int)__popcll(cb));
      cbase = uni(cbase);
      const int idx = cbase + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(cb >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((unsigned)cb, 0u));
      if (chg && idx < W) {
        sh.mt_d[idx] = bits;
        sh.mt_l[idx] = chg_key;
        sh.mt_p[idx] = q;
        sh.mt_t[idx] = chg_tile;
      }
    }
    const unsigned long long ob = __ballot(orph);
    if (ob) {
      int obase = 0;
      if (lane == 0) obase = atomicAdd(&sh.norph, (int)__popcll(ob));
      obase = uni(obase);
      const int idx = obase + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(ob >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((unsigned)ob, 0u));
      if (orph && idx < MAXO) sh.orph[idx] = q;
    }
  }
  double md;
  int ml;
  const int ln = wave_argmin(cbd, cbl, &md, &ml);
  const int bp = ln < 0 ? -1 : __builtin_amdgcn_readlane(cbp, ln);
  const double bw = ln < 0 ? 0.0 : readlane_d(cbw, ln);
  if (lane == 0) {
    out->bd[wid] = md;
    out->bl[wid] = ln < 0 ? BIG : ml;
    out->bp[wid] = bp;
    out->bw[wid] = bw;
  }
}

/* The lexicographic (distance, label) best of the waves' slots (wave 0, after
 * the command's second barrier): its slot index, or -1 if no wave found one. */
__device__ __forceinline__ int coop_best(const CoopCmd *cc, int nwaves) {
  int best = -1;
  for (int k = 0; k < nwaves; ++k) {
    const double d = cc->bd[k];
    const int l = cc->bl[k];
    if (l != BIG && (best < 0 || d < cc->bd[best] || (d == cc->bd[best] && l < cc->bl[best]))) best = k;
  }
  return best;
}

/* Waves 1 .. NW-1 of a cooperative kernel: run wave 0's commands until it posts
 * COOP_EXIT. */
template <class SH>
__device__ void coop_helper(SH &sh, CoopCmd *cc, int *mark, int wid, int nwaves, int lane) {
  for (;;) {
    __syncthreads(); /* posted */
    const CoopCmd c = *cc;
    if (c.op == COOP_EXIT) return;
    Ev ev;
    ev.rec = c.rec;
    ev.pool = c.pool;
    ev.p4 = nullptr;
    ev.hist = nullptr;
    ev.lab2 = nullptr;
    ev.pool_size = 0;
    coop_members(sh, sh.g[c.gsel], ev, c, wid, nwaves, lane, cc, mark);
    __syncthreads(); /* done */
  }
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

/* The tiled kernels, whose body is tiled_body.inc: MAXT tiles per event; at least WPE waves per
 * SIMD (the register budget: 512 / WPE VGPRs per lane).  SERVE: events come
 * from the host-published descriptor ring of LaunchArgs::serve instead of a
 * fixed queue, and each event's results are copied to host memory and
 * announced on completion.  GEN: the jet definition is LaunchArgs::def_arr[e]
 * (or def_uni), any real genkt exponent with a recombination scheme, instead
 * of p in {1, 0, -1} with the E-scheme.
 * The body is included into each kernel rather than called as a forced-inline
 * function: inlined from a function, the same code came out with more spilled
 * registers (18 -> 23 VGPRs on the 256-tile kernel). */
template <int MAXT, int WPE, bool SERVE>
__global__ __launch_bounds__(64, WPE) void tiled_kernel(LaunchArgs a, TiledScratch ts) {
  constexpr bool GEN = false;
  constexpr int NW = 1;
  CoopCmd *const coop = nullptr;
  int *const coop_mark = nullptr;
#include "tiled_body.inc"
}

/* general jet definitions: the same engine with a real genkt exponent and a
 * recombination scheme per event (LaunchArgs::def_arr / def_uni) */
template <int MAXT, int WPE>
__global__ __launch_bounds__(64, WPE) void tiled_gen_kernel(LaunchArgs a, TiledScratch ts) {
  constexpr bool SERVE = false, GEN = true;
  constexpr int NW = 1;
  CoopCmd *const coop = nullptr;
  int *const coop_mark = nullptr;
#include "tiled_body.inc"
}

/* large events: NW waves per event, 2 waves per SIMD (the full register
 * budget), one event per CU at NW = 8; wave 0 runs the event and the others
 * share its gathers of several chunks */
template <int MAXT, int NW>
__global__ __launch_bounds__(64 * NW, 2) void tiled_coop_kernel(LaunchArgs a, TiledScratch ts) {
  static_assert(NW >= 2 && NW <= kCoopMaxWaves, "waves per event");
  constexpr bool SERVE = false, GEN = false;
  constexpr int WPE = 2;
  __shared__ CoopCmd coop_cmd;
  __shared__ int coop_marks[NW * W]; /* each wave's member-position marks */
  CoopCmd *const coop = &coop_cmd;
  int *const coop_mark = coop_marks;
#include "tiled_body.inc"
}

/* The library builds this file twice: FJMI_TILED_PART 1 instantiates the
 * one-wave kernels (batch, serving, general definitions) and 2 the cooperative
 * ones, so that each part gets the machine-scheduler strategy that measured
 * fastest for it; 0, the default, instantiates both. */
#ifndef FJMI_TILED_PART
#define FJMI_TILED_PART 0
#endif
#if FJMI_TILED_PART != 2
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
/* general jet definitions: the same variants */
const void *const kTiledGenKernels[kNumTiledKernels] = {
    reinterpret_cast<const void *>(&tiled_gen_kernel<512, 2>), reinterpret_cast<const void *>(&tiled_gen_kernel<384, 3>),
    reinterpret_cast<const void *>(&tiled_gen_kernel<320, 3>), reinterpret_cast<const void *>(&tiled_gen_kernel<256, 3>)};
const int kTiledMaxTiles[kNumTiledKernels] = {512, 384, 320, 256};
#endif
#if FJMI_TILED_PART != 1
/* large events, 8 waves per event: variants 0 and 1 (the other variants'
 * rows take these too) */
const void *const kTiledCoopKernels[kNumTiledKernels] = {
    reinterpret_cast<const void *>(&tiled_coop_kernel<512, 8>), reinterpret_cast<const void *>(&tiled_coop_kernel<384, 8>),
    reinterpret_cast<const void *>(&tiled_coop_kernel<384, 8>), reinterpret_cast<const void *>(&tiled_coop_kernel<384, 8>)};
#endif

}  // namespace fjmi
