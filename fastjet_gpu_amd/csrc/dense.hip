/*
 * dense.hip — the dense nearest-neighbour clustering engine for gfx950.
 *
 * One workgroup clusters one event at a time; persistent workgroups pull event
 * indices from a queue (LaunchArgs::queue/head), so the same kernel serves the
 * one-event-per-launch path (grid of 1, queue of 1) and the batched path (a
 * persistent grid over a multiplicity bucket).
 *
 * Semantics: FastJet's N2Plain (the method the reference approximates with its
 * all-pairs triangle, tri_matrix.cu:138-159 and 286-336, and its single-block recompute loop,
 * n_array.cu:129-200), restated so that every decision, including exact ties, is
 * the same as the CPU oracle's:
 *   - jets live in slots 0..n-1; each has its nearest neighbour (NN) within R in
 *     (y, phi) (lowest slot on equal distance), NN distance and
 *     diJ = NN_dist * min(w_i, w_NN) (R^2 w_i without a NN);
 *   - each step takes the lowest slot with the smallest diJ (the reference's
 *     reduction_min/reduction_blocks argmin, tri_matrix.cu:161-257, with an
 *     index tie-break instead of CUB's tree order);
 *   - a pair merges (E-scheme, tri_matrix.cu:247-250) into the lower slot and the
 *     last slot moves into the higher one; a beam step removes the jet and the
 *     last slot moves into its place (tri_matrix.cu:240-255);
 *   - after a step only the merged jet's row is evaluated against every jet, and
 *     jets that had the removed or merged jet as NN rescan (incremental update,
 *     instead of the reference's full-triangle rescan, tri_matrix.cu:161-202).
 *
 * Layout (MI355X-first):
 *   - slot s is owned by thread s % T and lives in that thread's registers at
 *     index s / T: y, phi, w, NN distance, diJ, NN slot (11 VGPRs per slot);
 *   - LDS holds only what other threads need: the four-momenta and history
 *     indices of the slots (or, for the largest events, the four-momenta live in
 *     global memory and LDS only stages coordinates for the initial NN pass);
 *   - every argmin is a 64-lane DPP reduction (value, then lowest slot among the
 *     minima), combined across waves through LDS;
 *   - the merge (including atan2/log of the new jet) is computed by wave 0 only
 *     and broadcast through LDS; per-slot updates are branch-free selects, so the
 *     register arrays are never copied at control-flow joins.
 * Per step: one barrier after the merge (multi-wave groups), one after the update pass, and
 * one per round of up to MO orphan rescans (only when jets lost their NN).
 */
#include <hip/hip_runtime.h>

#include "fjmi_internal.h"
#include "fjmi_math.h"

namespace fjmi {

/* (y, phi, w) of a four-momentum.  Not inlined: the double-double atan2/log are
 * long, and one out-of-line copy keeps them from inflating the register
 * allocation of the clustering loop. */
__device__ __noinline__ double4 kin_coords(double px, double py, double pz, double E, int p) {
  const double kt2 = px * px + py * py;
  return make_double4(fjmi_rap(pz, E, kt2), fjmi_phi(px, py, kt2), fjmi_weight(kt2, p), 0.0);
}

constexpr int BIG = 0x7fffffff;
constexpr int OCH = 64; /* orphan list entries per chunk */
constexpr int MO = 2;   /* orphans rescanned per reduction round */

struct Red {
  double d;
  int s;
  int pi;
  double pd;
};

struct Orph {
  double rap, phi, w;
  int label;
  int pad;
};

struct Mail {
  double rap, phi, w, nnd, dij;
  int nn;
  int pend;
};

template <int W>
struct Shared {
  Red ru[W];
  Red rb[W];
  Red ro[2][MO][W];
  Orph orph[2][OCH];
  int norph[2];
  Mail mail;
  double4 newc; /* (y, phi, w) of the jet made by the current step */
  int ev;
};

__device__ __forceinline__ double readlane_d(double v, int l) {
  int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
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

/* Minimum over the 64 lanes, returned to every lane.  quad_perm xor1, xor2,
 * row_half_mirror, row_mirror give the row-of-16 minimum in every lane;
 * row_bcast15 (rows 1,3) and row_bcast31 (rows 2,3) fold the rows into lane 63. */
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

/* Lexicographic (d, s) minimum over the wave with the winner's payload. */
__device__ __forceinline__ Red wave_argmin(double d, int s, int pi, double pd) {
  Red r;
  r.d = wave_min_d(d);
  unsigned long long bal = __ballot(d == r.d);
  if (__popcll(bal) != 1) { /* ties across lanes (rare): lowest slot among the minima */
    const int key = (d == r.d) ? s : BIG;
    r.s = wave_min_i(key);
    bal = __ballot(key == r.s && r.s != BIG);
  } else {
    r.s = __builtin_amdgcn_readlane(s, __ffsll(bal) - 1);
    if (r.s == BIG) bal = 0;
  }
  if (bal) {
    const int ln = __ffsll(bal) - 1;
    r.pi = __builtin_amdgcn_readlane(pi, ln);
    r.pd = readlane_d(pd, ln);
  } else {
    r.pi = -1;
    r.pd = 0.0;
  }
  return r;
}

__device__ __forceinline__ bool lex_less(double d1, int s1, double d2, int s2) {
  return d1 < d2 || (d1 == d2 && s1 < s2);
}

template <int W>
__device__ __forceinline__ Red combine(const Red *r) {
  Red b = r[0];
#pragma unroll
  for (int i = 1; i < W; ++i) {
    Red x = r[i];
    if (lex_less(x.d, x.s, b.d, b.s)) b = x;
  }
  return b;
}

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

/* Writes one PseudoJet record (int index; bool isJet; double px,py,pz,E). */
__device__ __forceinline__ void write_pseudojet(unsigned char *rec, int index, bool is_jet, double4 q) {
  int *ip = reinterpret_cast<int *>(rec);
  ip[0] = index;
  ip[1] = is_jet ? 1 : 0;
  double *dp = reinterpret_cast<double *>(rec + 8);
  dp[0] = q.x;
  dp[1] = q.y;
  dp[2] = q.z;
  dp[3] = q.w;
}

template <int T, int K, bool P4G, int OCC>
__global__ __launch_bounds__(T, OCC) void dense_kernel(LaunchArgs a) {
  constexpr int W = T / 64;
  constexpr int CAP = T * K;
  static_assert(T % 64 == 0 && T <= 1024, "workgroup must be whole waves");
  static_assert(K >= 1 && K <= 32, "slot mask is 32 bits");
  __shared__ Shared<W> sh;
  extern __shared__ __align__(16) unsigned char smem[];

  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wv = t >> 6;
  const int CH = P4G ? a.stage_chunk : CAP; /* staging entries */
  double4 *stage = reinterpret_cast<double4 *>(smem);
  double4 *lp4 = reinterpret_cast<double4 *>(smem); /* aliases the stage when !P4G */
  int *lhist = reinterpret_cast<int *>(smem + sizeof(double4) * (size_t)CH);

  double rap[K], phi[K], w[K], nnd[K], dij[K];
  int nn[K];

  /* a launch without a queue head (one event) takes queue entry 0 and leaves */
  int pass = 0;
  for (;;) {
    __syncthreads(); /* previous event (or queue fetch) fully consumed */
    if (t == 0) sh.ev = a.head ? (int)(atomicAdd(a.head, 1u) - a.head_base) : pass;
    ++pass;
    __syncthreads();
    const int qi = uni(sh.ev);
    if (qi < 0 || qi >= a.n_queue) {
      if (a.t_end && threadIdx.x == 0) atomicMax(a.t_end, (unsigned long long)wall_clock64());
      break;
    }
    const int e = a.queue ? uni(a.queue[qi]) : qi;
    const int64_t base = a.offsets ? a.offsets[e] : 0;
    const int N = a.offsets ? uni((int)(a.offsets[e + 1] - base)) : a.n_single;
    if (N > CAP) {
      if (t == 0) {
        atomicOr(a.status, (unsigned)STATUS_CAPACITY);
        a.njets[e] = -1; /* failed */
      }
      continue;
    }
    if (N <= 0) {
      if (t == 0) a.njets[e] = 0;
      continue;
    }
    const int p = a.p_arr ? a.p_arr[e] : a.p_uni;
    const double R = a.R_arr ? a.R_arr[e] : a.R_uni;
    const double R2 = R * R;
    const double invR2 = 1.0 / R2;
    double4 *gp4 = P4G ? reinterpret_cast<double4 *>(a.p4_scratch) + base : nullptr;

    /* ---- prologue: coordinates of every particle (one loop, not unrolled per
     * slot, so the atan2/log code exists once) into the LDS stage, or into global
     * scratch when the event is larger than the stage ---- */
    double4 *gcoord = P4G ? reinterpret_cast<double4 *>(a.coord_scratch) + base : nullptr;
#pragma unroll 1
    for (int s = t; s < N; s += T) {
      const double *q = a.in + (size_t)(base + s) * a.in_stride + a.in_off;
      const double px = q[0], py = q[1], pz = q[2], E = q[3];
      const double4 c = kin_coords(px, py, pz, E, p);
      if (P4G) {
        gcoord[s] = c;
        gp4[s] = make_double4(px, py, pz, E);
      } else {
        stage[s] = c;
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int s = k * T + t;
      nnd[k] = R2;
      dij[k] = 0.0;
      nn[k] = -1;
      double4 c = make_double4(0.0, 0.0, 0.0, 0.0);
      if (s < N) c = P4G ? gcoord[s] : stage[s];
      rap[k] = c.x;
      phi[k] = c.y;
      w[k] = c.z;
    }

    /* ---- initial nearest neighbours: all pairs, read from the LDS stage (in
     * chunks when the coordinates live in global memory) ---- */
    double wnn[K];
#pragma unroll
    for (int k = 0; k < K; ++k) wnn[k] = 0.0;
    for (int c0 = 0; c0 < N; c0 += CH) {
      const int cn = min(CH, N - c0);
      if (P4G) {
        __syncthreads();
        for (int s = t; s < cn; s += T) stage[s] = gcoord[c0 + s];
        __syncthreads();
      }
      for (int j = 0; j < cn; ++j) {
        const double4 q = stage[j];
        const int sj = c0 + j;
#pragma unroll
        for (int k = 0; k < K; ++k) {
          if (k * T >= N) continue;
          const int s = k * T + t;
          const double d = fjmi_dr2(rap[k], phi[k], q.x, q.y);
          const bool c = s < N && s != sj && d < nnd[k];
          nnd[k] = c ? d : nnd[k];
          nn[k] = c ? sj : nn[k];
          wnn[k] = c ? q.z : wnn[k];
        }
      }
    }
    double uD = __builtin_inf();
    int uS = BIG, uNN = -1;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int s = k * T + t;
      if (s < N) {
        const double kt2 = (nn[k] >= 0 && wnn[k] < w[k]) ? wnn[k] : w[k];
        dij[k] = nnd[k] * kt2;
        if (lex_less(dij[k], s, uD, uS)) {
          uD = dij[k];
          uS = s;
          uNN = nn[k];
        }
      }
    }
    __syncthreads(); /* staging consumed */
    for (int s = t; s < N; s += T) {
      if (!P4G) {
        const double *q = a.in + (size_t)(base + s) * a.in_stride + a.in_off;
        lp4[s] = make_double4(q[0], q[1], q[2], q[3]);
      }
      lhist[s] = s;
    }
    if (t == 0) {
      sh.norph[0] = 0;
      sh.norph[1] = 0;
    }
    Red r0 = wave_argmin(uD, uS, uNN, 0.0);
    if (W > 1 && lane == 0) sh.ru[wv] = r0;
    __syncthreads();
    if (W > 1) r0 = combine<W>(sh.ru);

    double dmin = r0.d;
    int A = uni(r0.s), nnA = uni(r0.pi);
    int nj = 0;

    /* ---- the N recombination steps ---- */
    for (int step = 0; step < N; ++step) {
      const int n_old = N - step;
      const int hi = nnA >= 0 ? max(A, nnA) : A;
      const int lo = nnA >= 0 ? min(A, nnA) : -1;
      const int tail = n_old - 1;
      const int par = step & 1;

      /* The merge: wave 0 reads the two four-momenta, records the step and
       * computes the new jet's (y, phi, w) once; the other waves pick the three
       * numbers up from LDS after the barrier.  Only wave 0 reads the four-momenta
       * and history indices in LDS, so lane 0 updates them in program order. */
      double nrap = 0.0, nphi = 0.0, nw = 0.0;
      if (wv == 0) {
        const double4 qhi = P4G ? gp4[hi] : lp4[hi];
        const double4 qtail = P4G ? gp4[tail] : lp4[tail];
        const int hhi = lhist[hi];
        const int htail = lhist[tail];
        double4 qn = make_double4(0.0, 0.0, 0.0, 0.0);
        if (lo >= 0) {
          const double4 qlo = P4G ? gp4[lo] : lp4[lo];
          const int hlo = lhist[lo];
          qn = make_double4(qhi.x + qlo.x, qhi.y + qlo.y, qhi.z + qlo.z, qhi.w + qlo.w);
          const double4 c = kin_coords(qn.x, qn.y, qn.z, qn.w, p);
          nrap = c.x;
          nphi = c.y;
          nw = c.z;
          if (t == 0) {
            a.hist_p1[base + step] = min(hhi, hlo);
            a.hist_p2[base + step] = max(hhi, hlo);
            a.hist_dij[base + step] = dmin * invR2;
            if (W > 1) sh.newc = c;
          }
        } else if (t == 0) {
          a.hist_p1[base + step] = hhi;
          a.hist_p2[base + step] = -1;
          a.hist_dij[base + step] = dmin * invR2;
          double *jo = a.jets + 4 * (base + nj);
          jo[0] = qhi.x;
          jo[1] = qhi.y;
          jo[2] = qhi.z;
          jo[3] = qhi.w;
        }
        if (t == 0) {
          if (a.pj_out) write_pseudojet(a.pj_out + 40 * (size_t)(base + tail), tail, lo < 0, lo < 0 ? qhi : qtail);
          if (lo >= 0) {
            if (P4G)
              gp4[lo] = qn;
            else
              lp4[lo] = qn;
            lhist[lo] = N + step;
          }
          if (hi != tail) {
            if (P4G)
              gp4[hi] = qtail;
            else
              lp4[hi] = qtail;
            lhist[hi] = htail;
          }
        }
      }
      if (lo < 0) ++nj;
      if (W > 1) {
        __syncthreads(); /* the new jet's coordinates are in LDS */
        if (wv != 0 && lo >= 0) {
          const double4 c = sh.newc;
          nrap = c.x;
          nphi = c.y;
          nw = c.z;
        }
      }
      const int n = n_old - 1;
      if (n == 0) break;

      /* ---- update pass over the own slots ---- */
      double bD = R2, bW = 0.0;
      int bS = BIG;
      uD = __builtin_inf();
      uS = BIG;
      uNN = -1;
      unsigned pend = 0;
      /* Per slot: selects and bitwise predicates only (divergent branches
       * around updates of the register arrays make the compiler copy whole
       * arrays at the joins).  Slots are visited in increasing order, so a
       * strict < keeps the lowest slot on ties; only the moved jet (visited
       * last, relabelled hi) needs the explicit slot comparison. */
      const bool hasB = lo >= 0;
      unsigned omask = 0;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (k * T >= n_old) continue; /* uniform */
        const int s = k * T + t;
        const bool live = (s < n_old) & (s != hi);
        const bool isT = s == tail;
        const int label = isT ? hi : s;
        const bool isB = live & (label == lo);
        const bool normal = live & (label != lo);
        rap[k] = isB ? nrap : rap[k];
        phi[k] = isB ? nphi : phi[k];
        w[k] = isB ? nw : w[k];
        int m = nn[k];
        const bool orph = normal & ((m == hi) | (hasB & (m == lo)));
        m = (normal & !orph & (m == tail)) ? hi : m;
        const double d = fjmi_dr2(rap[k], phi[k], nrap, nphi);
        const bool closer = normal & hasB & !orph & (d < nnd[k]);
        const double wmin = nw < w[k] ? nw : w[k];
        dij[k] = closer ? d * wmin : dij[k];
        nnd[k] = closer ? d : nnd[k];
        m = closer ? lo : m;
        nn[k] = m;
        const bool bc = normal & hasB & ((d < bD) | (isT & (d == bD) & (d < R2) & (hi < bS)));
        bD = bc ? d : bD;
        bS = bc ? label : bS;
        bW = bc ? w[k] : bW;
        const bool uc = normal & !orph & ((dij[k] < uD) | (isT & (dij[k] == uD) & (hi < uS)));
        uD = uc ? dij[k] : uD;
        uS = uc ? label : uS;
        uNN = uc ? m : uNN;
        omask |= orph ? (1u << k) : 0u;
      }
      /* orphans (about one per step): one LDS atomic per wave and slot index */
      if (__ballot(omask != 0)) { /* uniform */
#pragma unroll
        for (int k = 0; k < K; ++k) {
          if (k * T >= n_old) continue;
          const bool orph = (omask >> k) & 1u;
          const unsigned long long obal = __ballot(orph);
          if (obal) { /* uniform */
            int obase = 0;
            if (lane == 0) obase = atomicAdd(&sh.norph[par], (int)__popcll(obal));
            obase = __builtin_amdgcn_readfirstlane(obase);
            const int idx = obase + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(obal >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((unsigned)obal, 0u));
            if (orph & (idx < OCH)) {
              Orph o;
              o.rap = rap[k];
              o.phi = phi[k];
              o.w = w[k];
              o.label = (k * T + t == tail) ? hi : k * T + t;
              o.pad = 0;
              sh.orph[par][idx] = o;
            }
            pend |= (orph & (idx >= OCH)) ? (1u << k) : 0u;
          }
        }
      }
      if (hi != tail) { /* uniform: hand the moved jet to slot hi's owner */
        const int kt = tail / T;
        if ((tail % T) == t) {
#pragma unroll
          for (int k = 0; k < K; ++k)
            if (k == kt) {
              Mail ml;
              ml.rap = rap[k];
              ml.phi = phi[k];
              ml.w = w[k];
              ml.nnd = nnd[k];
              ml.dij = dij[k];
              ml.nn = nn[k];
              ml.pend = (pend >> k) & 1u;
              sh.mail = ml;
            }
        }
        pend &= ((tail % T) == t) ? ~(1u << kt) : ~0u;
      }
      Red ru = wave_argmin(uD, uS, uNN, 0.0);
      Red rb = wave_argmin(bD, bS, 0, bW);
      if (W > 1 && lane == 0) {
        sh.ru[wv] = ru;
        sh.rb[wv] = rb;
      }
      __syncthreads();
      if (W > 1) {
        ru = combine<W>(sh.ru);
        rb = combine<W>(sh.rb);
      }
      if (t == 0) sh.norph[par ^ 1] = 0; /* the other parity's list was last read before this barrier */
      int norph = uni(sh.norph[par]);

      /* the moved jet changes owner */
      if (hi != tail) { /* uniform */
        const int kh = hi / T;
        const bool own = (hi % T) == t;
        const Mail ml = sh.mail;
#pragma unroll
        for (int k = 0; k < K; ++k)
          if (k == kh) { /* uniform */
            rap[k] = own ? ml.rap : rap[k];
            phi[k] = own ? ml.phi : phi[k];
            w[k] = own ? ml.w : w[k];
            nnd[k] = own ? ml.nnd : nnd[k];
            dij[k] = own ? ml.dij : dij[k];
            nn[k] = own ? ml.nn : nn[k];
            pend |= (own && ml.pend) ? (1u << k) : 0u;
          }
      }

      /* candidates of the next argmin: non-orphans (ru), the merged jet, the orphans */
      double bd = ru.d;
      int bs = ru.s, bnn = ru.pi;
      if (lo >= 0) {
        const int nnB = rb.s == BIG ? -1 : rb.s;
        const double nndB = rb.d;
        const double dijB = nndB * ((nnB >= 0 && rb.pd < nw) ? rb.pd : nw);
        const int kl = lo / T;
        const bool own = (lo % T) == t;
#pragma unroll
        for (int k = 0; k < K; ++k)
          if (k == kl) { /* uniform */
            nn[k] = own ? nnB : nn[k];
            nnd[k] = own ? nndB : nnd[k];
            dij[k] = own ? dijB : dij[k];
          }
        if (lex_less(dijB, lo, bd, bs)) {
          bd = dijB;
          bs = lo;
          bnn = nnB;
        }
      }

      /* ---- orphan rescans ---- */
      int oround = 0;
      while (norph > 0) {
        const int chunk = min(norph, OCH);
        for (int o0 = 0; o0 < chunk; o0 += MO) {
          const int mo = min(MO, chunk - o0);
          double orap[MO], ophi[MO], od[MO], ow[MO];
          int olab[MO], os[MO];
#pragma unroll
          for (int j = 0; j < MO; ++j) {
            const Orph o = sh.orph[par][o0 + (j < mo ? j : 0)];
            orap[j] = o.rap;
            ophi[j] = o.phi;
            olab[j] = j < mo ? o.label : -1;
            od[j] = R2;
            os[j] = BIG;
            ow[j] = 0.0;
          }
#pragma unroll
          for (int k = 0; k < K; ++k) {
            if (k * T >= n) continue;
            const int s = k * T + t;
#pragma unroll
            for (int j = 0; j < MO; ++j) {
              const double d = fjmi_dr2(rap[k], phi[k], orap[j], ophi[j]);
              const bool c = s < n && s != olab[j] && d < R2 && lex_less(d, s, od[j], os[j]);
              od[j] = c ? d : od[j];
              os[j] = c ? s : os[j];
              ow[j] = c ? w[k] : ow[j];
            }
          }
          Red ro[MO];
#pragma unroll
          for (int j = 0; j < MO; ++j) ro[j] = wave_argmin(od[j], os[j], 0, ow[j]);
          if (W > 1) {
            const int rp = oround & 1;
            if (lane == 0) {
#pragma unroll
              for (int j = 0; j < MO; ++j) sh.ro[rp][j][wv] = ro[j];
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < MO; ++j) ro[j] = combine<W>(sh.ro[rp][j]);
          }
          ++oround;
#pragma unroll
          for (int j = 0; j < MO; ++j) {
            if (j < mo) {
              const int lab = uni(olab[j]);
              const int nno = ro[j].s == BIG ? -1 : ro[j].s;
              const double nndo = ro[j].d;
              const double wo = sh.orph[par][o0 + j].w;
              const double dijo = nndo * ((nno >= 0 && ro[j].pd < wo) ? ro[j].pd : wo);
              const int kk = lab / T;
              const bool own = (lab % T) == t;
#pragma unroll
              for (int k = 0; k < K; ++k)
                if (k == kk) { /* uniform */
                  nn[k] = own ? nno : nn[k];
                  nnd[k] = own ? nndo : nnd[k];
                  dij[k] = own ? dijo : dij[k];
                }
              if (lex_less(dijo, lab, bd, bs)) {
                bd = dijo;
                bs = lab;
                bnn = nno;
              }
            }
          }
        }
        norph -= chunk;
        if (norph <= 0) break;
        /* more orphans than one list holds: refill the list from the pending ones */
        __syncthreads();
        if (t == 0) sh.norph[par] = 0;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; ++k) {
          if (pend & (1u << k)) {
            const int idx = atomicAdd(&sh.norph[par], 1);
            if (idx < OCH) {
              Orph o;
              o.rap = rap[k];
              o.phi = phi[k];
              o.w = w[k];
              o.label = k * T + t;
              o.pad = 0;
              sh.orph[par][idx] = o;
              pend &= ~(1u << k);
            }
          }
        }
        __syncthreads();
        norph = uni(sh.norph[par]); /* every pending orphan counted once */
      }
      dmin = bd;
      A = uni(bs);
      nnA = uni(bnn);
    }
    if (t == 0) a.njets[e] = nj;
  }
}

/* Per-step launch floor (the analogue of the reference's tri_matrix_empty.cu):
 * an empty kernel launched as often as a per-step-launch design would. */
__global__ void empty_kernel(int) {}

}  // namespace fjmi

/* ---- registry of the compiled configurations (see fjmi_kernels.h) ---- */
#include "fjmi_kernels.h"

namespace fjmi {

#define FJMI_DENSE(T, K, G, O) {T, K, G, O, reinterpret_cast<const void *>(&dense_kernel<T, K, G, O>)}
const KernelDesc kDenseKernels[] = {
#ifdef FJMI_KERNEL_LIST
    FJMI_KERNEL_LIST
#else
#include "dense_configs.inc"
#endif
};
#undef FJMI_DENSE
const int kNumDenseKernels = sizeof(kDenseKernels) / sizeof(kDenseKernels[0]);
const void *const kEmptyKernel = reinterpret_cast<const void *>(&empty_kernel);

/* Start of a batch graph: zeroes the queue heads and status word, and the
 * per-bucket end times; t[0] receives the start time. */
__global__ void zero_u32_kernel(unsigned int *p, int n, unsigned long long *t, int nt) {
  for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < n; i += blockDim.x * gridDim.x) p[i] = 0u;
  for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < nt; i += blockDim.x * gridDim.x)
    t[i] = i == 0 ? (unsigned long long)wall_clock64() : 0ull;
}
const void *const kZeroKernel = reinterpret_cast<const void *>(&zero_u32_kernel);

}  // namespace fjmi
