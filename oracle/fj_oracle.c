/*
 * fj_oracle.c — CPU oracle (TEST INFRASTRUCTURE ONLY; see fj_oracle.h).
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off).  The shared kinematics
 * header is the product's fjmi_math.h: both sides must use the same bits for
 * rapidity and phi; that header is pinned separately against high-precision
 * decimal arithmetic and glibc in tests/test_math.py.
 */
#include "fj_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../fastjet_gpu_amd/csrc/fjmi_math.h"

double fjo_log(double x) { return fjmi_log(x); }
double fjo_atan2(double y, double x) { return fjmi_atan2(y, x); }
double fjo_exp(double x) { return fjmi_exp(x); }
double fjo_pow(double x, double p) { return fjmi_pow_pos(x, p); }
double fjo_sqrt(double x) { return fjmi_sqrt(x); }
double fjo_sin(double x) {
  double s, c;
  fjmi_sincos(x, &s, &c);
  return s;
}
double fjo_cos(double x) {
  double s, c;
  fjmi_sincos(x, &s, &c);
  return c;
}
double fjo_weight_gen(double kt2, double p) { return fjmi_weight_gen(kt2, p); }

/* PseudoJet rap/phi (tri_matrix.cu:61-100) with the shared or glibc log/atan2. */
static void rap_phi(const double *q, int math_mode, double *rap, double *phi) {
  const double px = q[0], py = q[1], pz = q[2], E = q[3];
  const double kt2 = px * px + py * py;
  if (math_mode == FJO_MATH_LIBM) {
    double ph = (kt2 == 0.0) ? 0.0 : atan2(py, px);
    if (ph < 0.0) ph += FJMI_TWOPI;
    if (ph >= FJMI_TWOPI) ph -= FJMI_TWOPI;
    *phi = ph;
    if (E == fabs(pz) && kt2 == 0.0) {
      double mr = FJMI_MAXRAP + fabs(pz);
      *rap = (pz >= 0.0) ? mr : -mr;
    } else {
      double m2 = (E + pz) * (E - pz) - kt2;
      double em2 = m2 > 0.0 ? m2 : 0.0;
      double epz = E + fabs(pz);
      double r = 0.5 * log((kt2 + em2) / (epz * epz));
      *rap = (pz > 0) ? -r : r;
    }
  } else {
    *phi = fjmi_phi(px, py, kt2);
    *rap = fjmi_rap(pz, E, kt2);
  }
}

/* PseudoJet rap/phi and algorithm weight (tri_matrix.cu:61-100, grid.cu:153-166). */
static void set_jet(const double *q, int p, int math_mode, double *rap, double *phi, double *w) {
  rap_phi(q, math_mode, rap, phi);
  *w = fjmi_weight(q[0] * q[0] + q[1] * q[1], p);
}

/* A jet definition: the generalised-kt exponent (1 kt, 0 C/A, -1 anti-kt, any
 * other real genkt; FastJet's JetDefinition with genkt_algorithm) and the
 * recombination scheme (FJMI_SCHEME_*, FastJet's DefaultRecombiner). */
typedef struct {
  double p;
  int scheme;
} jdef_t;

static int jdef_ok(const jdef_t *d) {
  return d->p == d->p && d->p > -1e300 && d->p < 1e300 && d->scheme >= 0 && d->scheme < FJMI_NUM_SCHEMES;
}

void fjo_set_jet(const double *p4, int p, int math_mode, double *rap, double *phi, double *w) {
  set_jet(p4, p, math_mode, rap, phi, w);
}

/* ---------------------------------------------------------------------------
 * Shared bookkeeping: the FastJet jets/history arrays.
 * jets p4 storage holds the N inputs followed by one entry per merge.
 * ------------------------------------------------------------------------- */
typedef struct {
  double *p4;      /* 4 * (2N) */
  double *rap, *phi; /* per stored jet: from its momentum, or as the scheme fixes them */
  int32_t *hist;   /* history index of each jet */
  int njets_store; /* number of jets stored */
  int N;
  int step;
  int32_t *hp1, *hp2;
  double *hdij;
  double *out_jets;
  int32_t nout;
  jdef_t def;
  int mm;
  double *hist_jets; /* optional: step s's new entry's momentum at 4 s */
} book_t;

/* the hist_jets output of the *_defh entry points for the calling thread's clustering */
static _Thread_local double *t_hist_jets;

/* the inputs, as the recombination scheme takes them, with their rap/phi */
static int book_init(book_t *b, const double *p4, int n, const jdef_t *def, int mm, int32_t *hp1, int32_t *hp2,
                     double *hdij, double *jets) {
  b->p4 = (double *)malloc(sizeof(double) * 4 * (size_t)(2 * n + 1));
  b->rap = (double *)malloc(sizeof(double) * (size_t)(2 * n + 1));
  b->phi = (double *)malloc(sizeof(double) * (size_t)(2 * n + 1));
  b->hist = (int32_t *)malloc(sizeof(int32_t) * (size_t)(2 * n + 1));
  if (!b->p4 || !b->hist || !b->rap || !b->phi) {
    free(b->p4);
    free(b->rap);
    free(b->phi);
    free(b->hist);
    return -1;
  }
  memcpy(b->p4, p4, sizeof(double) * 4 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    double *q = b->p4 + 4 * (size_t)i;
    fjmi_preprocess(def->scheme, &q[0], &q[1], &q[2], &q[3]);
    rap_phi(q, mm, &b->rap[i], &b->phi[i]);
    b->hist[i] = i;
  }
  b->njets_store = n;
  b->N = n;
  b->step = 0;
  b->hp1 = hp1;
  b->hp2 = hp2;
  b->hdij = hdij;
  b->out_jets = jets;
  b->nout = 0;
  b->def = *def;
  b->mm = mm;
  b->hist_jets = t_hist_jets;
  return 0;
}

static void book_free(book_t *b) {
  free(b->p4);
  free(b->rap);
  free(b->phi);
  free(b->hist);
}

/* the algorithm weight of stored jet `jet` */
static double book_w(const book_t *b, int jet) {
  const double *q = b->p4 + 4 * (size_t)jet;
  return fjmi_weight_gen(q[0] * q[0] + q[1] * q[1], b->def.p);
}

/* Recombination (E-scheme: tri_matrix.cu:247-250; other schemes: FastJet's
 * DefaultRecombiner, fjmi_recombine): returns the new jet index. */
static int book_merge(book_t *b, int ja, int jb, double dij) {
  int k = b->njets_store++;
  double y, ph;
  if (fjmi_recombine(b->def.scheme, b->p4 + 4 * (size_t)ja, b->rap[ja], b->phi[ja], b->p4 + 4 * (size_t)jb,
                     b->rap[jb], b->phi[jb], b->p4 + 4 * (size_t)k, &y, &ph)) {
    b->rap[k] = y;
    b->phi[k] = ph;
  } else {
    rap_phi(b->p4 + 4 * (size_t)k, b->mm, &b->rap[k], &b->phi[k]);
  }
  if (b->hist_jets) memcpy(b->hist_jets + 4 * (size_t)b->step, b->p4 + 4 * (size_t)k, sizeof(double) * 4);
  int ha = b->hist[ja], hb = b->hist[jb];
  b->hist[k] = b->N + b->step;
  b->hp1[b->step] = ha < hb ? ha : hb;
  b->hp2[b->step] = ha < hb ? hb : ha;
  b->hdij[b->step] = dij;
  b->step++;
  return k;
}

static void book_beam(book_t *b, int ja, double dij) {
  b->hp1[b->step] = b->hist[ja];
  b->hp2[b->step] = -1;
  b->hdij[b->step] = dij;
  b->step++;
  memcpy(b->out_jets + 4 * (size_t)b->nout, b->p4 + 4 * (size_t)ja, sizeof(double) * 4);
  b->nout++;
}

/* ---------------------------------------------------------------------------
 * N2Plain restatement.
 * ------------------------------------------------------------------------- */
typedef struct {
  double eta, phi, kt2; /* kt2 holds the algorithm weight w = kt^(2p) */
  double nn_dist;
  int nn; /* slot of the nearest neighbour within R, or -1 */
  int jet;   /* index into book_t jets */
} bj_t;

static inline double bj_dist(const bj_t *a, const bj_t *b) { return fjmi_dr2(a->eta, a->phi, b->eta, b->phi); }

static inline double bj_diJ(const bj_t *B, int i) {
  double kt2 = B[i].kt2;
  if (B[i].nn >= 0 && B[B[i].nn].kt2 < kt2) kt2 = B[B[i].nn].kt2;
  return B[i].nn_dist * kt2;
}

static void bj_set(bj_t *x, const book_t *b, int jet, double R2) {
  x->eta = b->rap[jet];
  x->phi = b->phi[jet];
  x->kt2 = book_w(b, jet);
  x->jet = jet;
  x->nn_dist = R2;
  x->nn = -1;
}

/* nearest neighbour among all other slots in [0,n), lowest slot on ties */
static void bj_set_nn_nocross(bj_t *B, int i, int n, double R2) {
  double best = R2;
  int nn = -1;
  for (int j = 0; j < n; ++j) {
    if (j == i) continue;
    double d = bj_dist(&B[i], &B[j]);
    if (d < best) {
      best = d;
      nn = j;
    }
  }
  B[i].nn = nn;
  B[i].nn_dist = best;
}

static int cluster_n2plain(const double *p4, int n, const jdef_t *def, double R, int mm, int32_t *hp1, int32_t *hp2,
                           double *hdij,
                        double *jets, int32_t *njets) {
  if (n < 0 || !(R > 0.0) || !jdef_ok(def)) return -1;
  *njets = 0;
  if (n == 0) return 0;
  const double R2 = R * R, invR2 = 1.0 / R2;
  book_t bk;
  if (book_init(&bk, p4, n, def, mm, hp1, hp2, hdij, jets)) return -2;
  bj_t *B = (bj_t *)malloc(sizeof(bj_t) * (size_t)n);
  double *diJ = (double *)malloc(sizeof(double) * (size_t)n);
  if (!B || !diJ) {
    free(B);
    free(diJ);
    book_free(&bk);
    return -2;
  }
  for (int i = 0; i < n; ++i) bj_set(&B[i], &bk, i, R2);
  /* initial NN by cross-checking each slot against the lower slots */
  for (int a = 1; a < n; ++a) {
    double best = R2;
    int nn = -1;
    for (int b = 0; b < a; ++b) {
      double d = bj_dist(&B[a], &B[b]);
      if (d < best) {
        best = d;
        nn = b;
      }
      if (d < B[b].nn_dist) {
        B[b].nn_dist = d;
        B[b].nn = a;
      }
    }
    B[a].nn = nn;
    B[a].nn_dist = best;
  }
  for (int i = 0; i < n; ++i) diJ[i] = bj_diJ(B, i);

  while (n > 0) {
    double dmin = diJ[0];
    int amin = 0;
    for (int i = 0; i < n; ++i)
      if (diJ[i] < dmin) {
        amin = i;
        dmin = diJ[i];
      }
    int ja = amin;
    int jb = B[ja].nn;
    double dij = dmin * invR2;
    if (jb >= 0) {
      if (ja < jb) {
        int t = ja;
        ja = jb;
        jb = t;
      }
      int k = book_merge(&bk, B[ja].jet, B[jb].jet, dij);
      bj_set(&B[jb], &bk, k, R2);
    } else {
      book_beam(&bk, B[ja].jet, dij);
    }
    const int tail = n - 1;
    n--;
    B[ja] = B[tail];
    diJ[ja] = diJ[tail];
    for (int i = 0; i < n; ++i) {
      if (B[i].nn == ja || (jb >= 0 && B[i].nn == jb)) {
        bj_set_nn_nocross(B, i, n, R2);
        diJ[i] = bj_diJ(B, i);
      }
      if (jb >= 0) {
        double d = bj_dist(&B[i], &B[jb]);
        if (d < B[i].nn_dist && i != jb) {
          B[i].nn_dist = d;
          B[i].nn = jb;
          diJ[i] = bj_diJ(B, i);
        }
        if (d < B[jb].nn_dist && i != jb) {
          B[jb].nn_dist = d;
          B[jb].nn = i;
        }
      }
      if (B[i].nn == tail) B[i].nn = ja;
    }
    if (jb >= 0) diJ[jb] = bj_diJ(B, jb);
  }
  *njets = bk.nout;
  free(B);
  free(diJ);
  book_free(&bk);
  return 0;
}

/* ---------------------------------------------------------------------------
 * Brute force: same slot semantics, NN recomputed from scratch every step.
 * ------------------------------------------------------------------------- */
static int cluster_brute(const double *p4, int n, const jdef_t *def, double R, int mm, int32_t *hp1, int32_t *hp2,
                           double *hdij,
                      double *jets, int32_t *njets) {
  if (n < 0 || !(R > 0.0) || !jdef_ok(def)) return -1;
  *njets = 0;
  if (n == 0) return 0;
  const double R2 = R * R, invR2 = 1.0 / R2;
  book_t bk;
  if (book_init(&bk, p4, n, def, mm, hp1, hp2, hdij, jets)) return -2;
  bj_t *B = (bj_t *)malloc(sizeof(bj_t) * (size_t)n);
  if (!B) {
    book_free(&bk);
    return -2;
  }
  for (int i = 0; i < n; ++i) bj_set(&B[i], &bk, i, R2);
  while (n > 0) {
    for (int i = 0; i < n; ++i) bj_set_nn_nocross(B, i, n, R2);
    double dmin = bj_diJ(B, 0);
    int amin = 0;
    for (int i = 1; i < n; ++i) {
      double d = bj_diJ(B, i);
      if (d < dmin) {
        dmin = d;
        amin = i;
      }
    }
    int ja = amin, jb = B[ja].nn;
    double dij = dmin * invR2;
    if (jb >= 0) {
      if (ja < jb) {
        int t = ja;
        ja = jb;
        jb = t;
      }
      int k = book_merge(&bk, B[ja].jet, B[jb].jet, dij);
      bj_set(&B[jb], &bk, k, R2);
    } else {
      book_beam(&bk, B[ja].jet, dij);
    }
    B[ja] = B[n - 1];
    n--;
  }
  *njets = bk.nout;
  free(B);
  book_free(&bk);
  return 0;
}

/* ---------------------------------------------------------------------------
 * Textbook pairwise generalised-kt (no NN bookkeeping).
 * ------------------------------------------------------------------------- */
static int cluster_pairwise(const double *p4, int n, const jdef_t *def, double R, int mm, int32_t *hp1, int32_t *hp2,
                           double *hdij,
                         double *jets, int32_t *njets) {
  if (n < 0 || !(R > 0.0) || !jdef_ok(def)) return -1;
  *njets = 0;
  if (n == 0) return 0;
  const double R2 = R * R, invR2 = 1.0 / R2;
  book_t bk;
  if (book_init(&bk, p4, n, def, mm, hp1, hp2, hdij, jets)) return -2;
  int *live = (int *)malloc(sizeof(int) * (size_t)n);
  double *rap = (double *)malloc(sizeof(double) * (size_t)(2 * n));
  double *phi = (double *)malloc(sizeof(double) * (size_t)(2 * n));
  double *w = (double *)malloc(sizeof(double) * (size_t)(2 * n));
  if (!live || !rap || !phi || !w) {
    free(live);
    free(rap);
    free(phi);
    free(w);
    book_free(&bk);
    return -2;
  }
  for (int i = 0; i < n; ++i) {
    live[i] = i;
    rap[i] = bk.rap[i];
    phi[i] = bk.phi[i];
    w[i] = book_w(&bk, i);
  }
  int nl = n;
  while (nl > 0) {
    double best = 0.0;
    int bi = -1, bj = -1;
    for (int a = 0; a < nl; ++a) {
      int i = live[a];
      double d = w[i] * R2;
      if (bi < 0 || d < best) {
        best = d;
        bi = a;
        bj = -1;
      }
      for (int c = a + 1; c < nl; ++c) {
        int j = live[c];
        double dr2 = fjmi_dr2(rap[i], phi[i], rap[j], phi[j]);
        if (!(dr2 < R2)) continue;
        double wm = w[i] < w[j] ? w[i] : w[j];
        double dd = dr2 * wm;
        if (dd < best) {
          best = dd;
          bi = a;
          bj = c;
        }
      }
    }
    double dij = best * invR2;
    if (bj >= 0) {
      int k = book_merge(&bk, live[bi], live[bj], dij);
      rap[k] = bk.rap[k];
      phi[k] = bk.phi[k];
      w[k] = book_w(&bk, k);
      live[bi] = k;
      for (int c = bj; c < nl - 1; ++c) live[c] = live[c + 1];
    } else {
      book_beam(&bk, live[bi], dij);
      for (int c = bi; c < nl - 1; ++c) live[c] = live[c + 1];
    }
    nl--;
  }
  *njets = bk.nout;
  free(live);
  free(rap);
  free(phi);
  free(w);
  book_free(&bk);
  return 0;
}

/* ---------------------------------------------------------------------------
 * Tiled N2Plain: the same decisions as fjo_cluster_n2plain (every tie included),
 * with nearest neighbours maintained through a (rapidity, phi) tiling whose tile
 * side is at least R (phi wraps around), in the spirit of FastJet's N2Tiled and of
 * the reference's grid engine (grid.cu:55-106 geometry, 337-549 neighbour
 * repair).  Jets keep their N2Plain slot number as a *label*: the argmin breaks
 * ties by label, the nearest-neighbour search breaks distance ties by label, and
 * the compaction of N2Plain only relabels the last jet.  Any two jets closer than
 * R sit in the same or adjacent tiles, so every candidate is found.
 * ------------------------------------------------------------------------- */
typedef struct {
  double eta, phi, w, nnd;
  int nn;    /* physical index of the nearest neighbour, or -1 */
  int label; /* N2Plain slot */
  int tile;
  int prev, next; /* tile list */
  int jet;        /* index into book_t */
} tj_t;

typedef struct {
  int ny, nphi;
  double y0, sy, sphi;
} tgeom_t;

static void tgeom_init(tgeom_t *g, const tj_t *J, int n, double R) {
  const double side = R * (1.0 + 0x1p-20); /* margin against rounding at tile edges */
  double lo = 0.0, hi = 0.0;
  int any = 0;
  for (int i = 0; i < n; ++i) {
    double y = J[i].eta;
    if (!(y > -1e4 && y < 1e4)) continue; /* zero-pt momenta at +-(1e5+|pz|) go to the edge tiles */
    if (!any || y < lo) lo = y;
    if (!any || y > hi) hi = y;
    any = 1;
  }
  int ny = (int)((hi - lo) / side) + 1;
  if (ny > 256) ny = 256;
  g->ny = ny;
  g->y0 = lo;
  g->sy = (hi - lo) / ny > side ? (hi - lo) / ny : side;
  int np = (int)(FJMI_TWOPI / side);
  if (np < 1) np = 1;
  g->nphi = np;
  g->sphi = FJMI_TWOPI / np;
}

static int tgeom_tile(const tgeom_t *g, double y, double phi) {
  double fy = (y - g->y0) / g->sy;
  int ty = fy >= 0.0 ? (fy < (double)g->ny ? (int)fy : g->ny - 1) : 0; /* NaN: tile row 0 */
  const double fp = phi / g->sphi;
  int tp = fp >= 0.0 ? (fp < (double)g->nphi ? (int)fp : g->nphi - 1) : 0;
  return ty * g->nphi + tp;
}

/* the distinct tiles of the 3x3 neighbourhood of `tile` (up to 9) */
static int tgeom_neigh(const tgeom_t *g, int tile, int *out) {
  int ty = tile / g->nphi, tp = tile % g->nphi, k = 0;
  for (int dy = -1; dy <= 1; ++dy) {
    int y = ty + dy;
    if (y < 0 || y >= g->ny) continue;
    for (int dp = -1; dp <= 1; ++dp) {
      int q = ((tp + dp) % g->nphi + g->nphi) % g->nphi;
      int t = y * g->nphi + q, dup = 0;
      for (int m = 0; m < k; ++m) dup |= out[m] == t;
      if (!dup) out[k++] = t;
    }
  }
  return k;
}

static void tl_insert(tj_t *J, int *head, int p) {
  J[p].prev = -1;
  J[p].next = head[J[p].tile];
  if (J[p].next >= 0) J[J[p].next].prev = p;
  head[J[p].tile] = p;
}

static void tl_remove(tj_t *J, int *head, int p) {
  if (J[p].prev >= 0)
    J[J[p].prev].next = J[p].next;
  else
    head[J[p].tile] = J[p].next;
  if (J[p].next >= 0) J[J[p].next].prev = J[p].prev;
}

static inline double tj_dij(const tj_t *J, int p) {
  double kt2 = J[p].w;
  if (J[p].nn >= 0 && J[J[p].nn].w < kt2) kt2 = J[J[p].nn].w;
  return J[p].nnd * kt2;
}

/* pair distances evaluated by the tiled restatement on this thread (the
 * algorithmic work of the tiled method; read by fjo_tiled_pair_evals) */
static _Thread_local int64_t t_pair_evals;

/* lexicographic (distance, label) nearest neighbour of p among the 3x3 tiles */
static void tj_rescan(tj_t *J, const int *head, const tgeom_t *g, int p, double R2) {
  int nb[9];
  int k = tgeom_neigh(g, J[p].tile, nb);
  double best = R2;
  int bn = -1, bl = 0;
  for (int m = 0; m < k; ++m)
    for (int q = head[nb[m]]; q >= 0; q = J[q].next) {
      if (q == p) continue;
      ++t_pair_evals;
      double d = fjmi_dr2(J[p].eta, J[p].phi, J[q].eta, J[q].phi);
      if (d < best || (d == best && bn >= 0 && J[q].label < bl)) {
        best = d;
        bn = q;
        bl = J[q].label;
      }
    }
  J[p].nn = bn;
  J[p].nnd = best;
}

static int cluster_tiled(const double *p4, int n, const jdef_t *def, double R, int mm, int32_t *hp1, int32_t *hp2,
                           double *hdij,
                      double *jets, int32_t *njets) {
  if (n < 0 || !(R > 0.0) || !jdef_ok(def)) return -1;
  *njets = 0;
  if (n == 0) return 0;
  const double R2 = R * R, invR2 = 1.0 / R2;
  book_t bk;
  if (book_init(&bk, p4, n, def, mm, hp1, hp2, hdij, jets)) return -2;
  tj_t *J = (tj_t *)malloc(sizeof(tj_t) * (size_t)n);
  int *lab2 = (int *)malloc(sizeof(int) * (size_t)n);
  double *dij = (double *)malloc(sizeof(double) * (size_t)n);
  if (!J || !lab2 || !dij) {
    free(J);
    free(lab2);
    free(dij);
    book_free(&bk);
    return -2;
  }
  for (int i = 0; i < n; ++i) {
    J[i].eta = bk.rap[i];
    J[i].phi = bk.phi[i];
    J[i].w = book_w(&bk, i);
    J[i].jet = i;
    J[i].label = i;
    lab2[i] = i;
  }
  tgeom_t g;
  tgeom_init(&g, J, n, R);
  const int ntiles = g.ny * g.nphi;
  int *head = (int *)malloc(sizeof(int) * (size_t)ntiles);
  if (!head) {
    free(J);
    free(lab2);
    free(dij);
    book_free(&bk);
    return -2;
  }
  for (int t = 0; t < ntiles; ++t) head[t] = -1;
  for (int i = n - 1; i >= 0; --i) {
    J[i].tile = tgeom_tile(&g, J[i].eta, J[i].phi);
    tl_insert(J, head, i);
  }
  for (int i = 0; i < n; ++i) tj_rescan(J, head, &g, i, R2);
  for (int i = 0; i < n; ++i) dij[i] = tj_dij(J, i);

  int nb[9], nb2[9];
  while (n > 0) {
    /* argmin in label order: strict < keeps the lowest label */
    int A = lab2[0];
    double dmin = dij[A];
    for (int l = 1; l < n; ++l) {
      int q = lab2[l];
      if (dij[q] < dmin) {
        dmin = dij[q];
        A = q;
      }
    }
    int B = J[A].nn;
    double dd = dmin * invR2;
    int phi_, plo; /* physical slots of the higher and lower label */
    if (B >= 0) {
      phi_ = J[A].label > J[B].label ? A : B;
      plo = phi_ == A ? B : A;
    } else {
      phi_ = A;
      plo = -1;
    }
    const int hl = J[phi_].label;
    const int tail = n - 1;
    const int pt = lab2[tail];
    const int oldtile_hi = J[phi_].tile;
    const int oldtile_lo = plo >= 0 ? J[plo].tile : -1;
    if (plo >= 0) {
      int k = book_merge(&bk, J[phi_].jet, J[plo].jet, dd);
      tl_remove(J, head, plo);
      J[plo].eta = bk.rap[k];
      J[plo].phi = bk.phi[k];
      J[plo].w = book_w(&bk, k);
      J[plo].jet = k;
      J[plo].nn = -1;
      J[plo].nnd = R2;
      J[plo].tile = tgeom_tile(&g, J[plo].eta, J[plo].phi);
      tl_insert(J, head, plo);
    } else {
      book_beam(&bk, J[phi_].jet, dd);
    }
    tl_remove(J, head, phi_);
    J[phi_].tile = -1;
    J[phi_].label = -1;
    if (pt != phi_) { /* N2Plain compaction: the last jet takes the removed slot */
      J[pt].label = hl;
      lab2[hl] = pt;
    }
    n--;
    if (n == 0) break;
    /* orphans: jets whose NN was the removed or the merged jet; they lie in the
     * 3x3 tiles of the two old positions */
    int norph_t = tgeom_neigh(&g, oldtile_hi, nb);
    if (oldtile_lo >= 0) {
      int k2 = tgeom_neigh(&g, oldtile_lo, nb2);
      for (int m = 0; m < k2; ++m) {
        int dup = 0;
        for (int u = 0; u < norph_t; ++u) dup |= nb[u] == nb2[m];
        if (!dup) nb[norph_t++] = nb2[m];
      }
    }
    /* nb may now hold up to 18 tiles */
    int orph[64], no = 0, overflow = 0;
    for (int m = 0; m < norph_t; ++m)
      for (int q = head[nb[m]]; q >= 0; q = J[q].next)
        if (q != plo && (J[q].nn == phi_ || (plo >= 0 && J[q].nn == plo))) {
          if (no < 64)
            orph[no++] = q;
          else
            overflow = 1;
        }
    if (overflow) { /* pathological: rescan every jet of those tiles that lost its NN */
      for (int m = 0; m < norph_t; ++m)
        for (int q = head[nb[m]]; q >= 0; q = J[q].next)
          if (q != plo && (J[q].nn == phi_ || (plo >= 0 && J[q].nn == plo))) J[q].nn = -2;
      for (int m = 0; m < norph_t; ++m)
        for (int q = head[nb[m]]; q >= 0; q = J[q].next)
          if (J[q].nn == -2) {
            tj_rescan(J, head, &g, q, R2);
            dij[q] = tj_dij(J, q);
          }
    } else {
      for (int o = 0; o < no; ++o) {
        tj_rescan(J, head, &g, orph[o], R2);
        dij[orph[o]] = tj_dij(J, orph[o]);
      }
    }
    if (plo >= 0) {
      /* the merged jet: update the jets around it and find its own NN */
      int k = tgeom_neigh(&g, J[plo].tile, nb);
      double best = R2;
      int bn = -1, bl = 0;
      for (int m = 0; m < k; ++m)
        for (int q = head[nb[m]]; q >= 0; q = J[q].next) {
          if (q == plo) continue;
          ++t_pair_evals;
          double d = fjmi_dr2(J[q].eta, J[q].phi, J[plo].eta, J[plo].phi);
          if (d < J[q].nnd) { /* orphans were rescanned against the new jet already */
            J[q].nnd = d;
            J[q].nn = plo;
            dij[q] = tj_dij(J, q);
          }
          if (d < best || (d == best && bn >= 0 && J[q].label < bl)) {
            best = d;
            bn = q;
            bl = J[q].label;
          }
        }
      J[plo].nn = bn;
      J[plo].nnd = best;
      dij[plo] = tj_dij(J, plo);
    }
  }
  *njets = bk.nout;
  free(J);
  free(lab2);
  free(dij);
  free(head);
  book_free(&bk);
  return 0;
}

/* ---------------------------------------------------------------------------
 * Results derived from a clustering history (the same outputs as the GPU's
 * post-processing kernels).
 * ------------------------------------------------------------------------- */

/* main.cc:309-327: keep the jets with px^2 + py^2 >= ptmin^2 (the filter tests
 * "px*px + py*py < ptmin*ptmin" for removal) and sort them by decreasing
 * px^2 + py^2.  The reference's std::sort leaves equal keys in an unspecified
 * order; here they keep their promotion order. */
int fjo_inclusive_jets(const double *jets, int njets, double ptmin, double *out, int32_t *n_out) {
  if (njets < 0 || !n_out) return -1;
  const double pt2min = ptmin * ptmin;
  int k = 0;
  for (int i = 0; i < njets; ++i) {
    const double *q = jets + 4 * (size_t)i;
    const double key = q[0] * q[0] + q[1] * q[1];
    if (key < pt2min) continue;
    int rank = 0; /* kept jets before it in the sorted order */
    for (int j = 0; j < njets; ++j) {
      const double *r = jets + 4 * (size_t)j;
      const double kj = r[0] * r[0] + r[1] * r[1];
      if (kj < pt2min) continue;
      if (kj > key || (kj == key && j < i)) ++rank;
    }
    memcpy(out + 4 * (size_t)rank, q, sizeof(double) * 4);
    ++k;
  }
  *n_out = k;
  return 0;
}

/* FastJet's ClusterSequence history view of our N steps: history entries
 * 0..N-1 are the particles and N+s the result of step s.  child[k] is the entry
 * whose step consumed entry k (-3, FastJet's Invalid, for none: the results of
 * beam steps and nothing else once the clustering is complete);
 * max_dij[s] = max(dij[0..s]) (FastJet's max_dij_so_far, whose particle
 * entries hold 0). */
int fjo_history_children(const int32_t *hp1, const int32_t *hp2, const double *hdij, int n, int32_t *child,
                         double *max_dij) {
  if (n < 0) return -1;
  for (int k = 0; k < 2 * n; ++k) child[k] = -3;
  double m = 0.0;
  for (int s = 0; s < n; ++s) {
    if (hp1[s] < 0 || hp1[s] >= n + s || hp2[s] >= n + s) return -2;
    child[hp1[s]] = n + s;
    if (hp2[s] >= 0) child[hp2[s]] = n + s;
    m = hdij[s] > m ? hdij[s] : m;
    max_dij[s] = m;
  }
  return 0;
}

/* FastJet's exclusive jets from the history: ClusterSequence::exclusive_jets(njets)
 * (mode 0, param = njets <= N) or exclusive_jets(dcut) (mode 1, param = dcut, via
 * n_exclusive_jets: the last entry whose max_dij_so_far <= dcut is the stop
 * point).  The jets are the entries before the stop point consumed at or after
 * it, in history order of the consuming step (parent1 then parent2).  Their
 * momenta are the E-scheme sums of the history, replayed in step order. */
int fjo_exclusive_jets(const double *p4, const int32_t *hp1, const int32_t *hp2, const double *hdij, int n, int mode,
                       double param, double *out, int32_t *n_out) {
  if (n < 0 || !n_out) return -1;
  int njets;
  if (mode == 0) {
    njets = (int)param;
    if (!(param >= 0.0) || njets > n || (double)njets != param) return -1;
  } else {
    int i = 2 * n - 1;
    double m = 0.0;
    double *mx = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    if (!mx) return -2;
    for (int s = 0; s < n; ++s) {
      m = hdij[s] > m ? hdij[s] : m;
      mx[s] = m;
    }
    while (i >= 0) {
      const double ms = i >= n ? mx[i - n] : 0.0;
      if (ms <= param) break;
      --i;
    }
    free(mx);
    njets = 2 * n - (i + 1);
  }
  const int stop = 2 * n - njets;
  double *h = (double *)malloc(sizeof(double) * 4 * (size_t)(2 * n + 1));
  if (!h) return -2;
  memcpy(h, p4, sizeof(double) * 4 * (size_t)n);
  for (int s = 0; s + n < stop; ++s)
    if (hp2[s] >= 0)
      for (int c = 0; c < 4; ++c) h[4 * (size_t)(n + s) + c] = h[4 * (size_t)hp1[s] + c] + h[4 * (size_t)hp2[s] + c];
  int k = 0;
  for (int e = stop; e < 2 * n; ++e) {
    const int s = e - n;
    if (hp1[s] < stop) memcpy(out + 4 * (size_t)k++, h + 4 * (size_t)hp1[s], sizeof(double) * 4);
    if (hp2[s] < stop && hp2[s] > 0) memcpy(out + 4 * (size_t)k++, h + 4 * (size_t)hp2[s], sizeof(double) * 4);
  }
  free(h);
  *n_out = k;
  return 0;
}

/* Exclusive jets of a clustering under recombination scheme `scheme`
 * (ClusterSequence::exclusive_jets reads its stored jets): the stop point as in
 * fjo_exclusive_jets, the particles as the scheme preprocesses them, merged
 * entries N + s from hist_jets[4 s]. */
int fjo_exclusive_jets_def(const double *p4, const int32_t *hp1, const int32_t *hp2, const double *hdij,
                           const double *hist_jets, int n, int scheme, int mode, double param, double *out,
                           int32_t *n_out) {
  if (n < 0 || !n_out || scheme < 0 || scheme >= FJMI_NUM_SCHEMES) return -1;
  int njets;
  if (mode == 0) {
    njets = (int)param;
    if (!(param >= 0.0) || njets > n || (double)njets != param) return -1;
  } else {
    int i = 2 * n - 1;
    double m = 0.0;
    double *mx = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    if (!mx) return -2;
    for (int s = 0; s < n; ++s) {
      m = hdij[s] > m ? hdij[s] : m;
      mx[s] = m;
    }
    while (i >= 0) {
      const double ms = i >= n ? mx[i - n] : 0.0;
      if (ms <= param) break;
      --i;
    }
    free(mx);
    njets = 2 * n - (i + 1);
  }
  const int stop = 2 * n - njets;
  int k = 0;
  for (int e = stop; e < 2 * n; ++e) {
    const int s = e - n;
    const int sel[2] = {hp1[s], hp2[s]};
    for (int c = 0; c < 2; ++c) {
      const int x = sel[c];
      if (x < 0 || x >= stop || (c == 1 && x == 0)) continue;
      double *o = out + 4 * (size_t)k++;
      if (x >= n) {
        memcpy(o, hist_jets + 4 * (size_t)(x - n), sizeof(double) * 4);
      } else {
        memcpy(o, p4 + 4 * (size_t)x, sizeof(double) * 4);
        fjmi_preprocess(scheme, &o[0], &o[1], &o[2], &o[3]);
      }
    }
  }
  *n_out = k;
  return 0;
}

/* Public entry points: kt / C/A / anti-kt with the E-scheme, and any jet definition. */
#define FJO_WRAP(name)                                                                                              \
  int fjo_cluster_##name(const double *p4, int n, int p, double R, int mm, int32_t *hp1, int32_t *hp2, double *hdij, \
                         double *jets, int32_t *njets) {                                                             \
    if (p < -1 || p > 1) return -1;                                                                                  \
    const jdef_t d = {(double)p, FJMI_SCHEME_E};                                                                     \
    return cluster_##name(p4, n, &d, R, mm, hp1, hp2, hdij, jets, njets);                                            \
  }                                                                                                                  \
  int fjo_cluster_##name##_def(const double *p4, int n, double p, int scheme, double R, int mm, int32_t *hp1,      \
                               int32_t *hp2, double *hdij, double *jets, int32_t *njets) {                          \
    const jdef_t d = {p, scheme};                                                                                    \
    return cluster_##name(p4, n, &d, R, mm, hp1, hp2, hdij, jets, njets);                                            \
  }                                                                                                                  \
  int fjo_cluster_##name##_defh(const double *p4, int n, double p, int scheme, double R, int mm, int32_t *hp1,     \
                                int32_t *hp2, double *hdij, double *jets, int32_t *njets, double *hist_jets) {      \
    const jdef_t d = {p, scheme};                                                                                    \
    t_hist_jets = hist_jets;                                                                                         \
    const int rc = cluster_##name(p4, n, &d, R, mm, hp1, hp2, hdij, jets, njets);                                    \
    t_hist_jets = NULL;                                                                                              \
    return rc;                                                                                                       \
  }
FJO_WRAP(n2plain)
FJO_WRAP(brute)
FJO_WRAP(pairwise)
FJO_WRAP(tiled)

int64_t fjo_tiled_pair_evals(const double *p4, int n, int p, double R) {
  if (n <= 0) return 0;
  int32_t *h1 = (int32_t *)malloc(sizeof(int32_t) * (size_t)n), *h2 = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  double *hd = (double *)malloc(sizeof(double) * (size_t)n), *jt = (double *)malloc(sizeof(double) * 4 * (size_t)n);
  int32_t nj = 0;
  int64_t r = -1;
  if (h1 && h2 && hd && jt) {
    t_pair_evals = 0;
    if (fjo_cluster_tiled(p4, n, p, R, FJO_MATH_SHARED, h1, h2, hd, jt, &nj) == 0) r = t_pair_evals;
  }
  free(h1);
  free(h2);
  free(hd);
  free(jt);
  return r;
}

/* ---------------------------------------------------------------------------
 * Batches on host threads: the CPU baseline (timed) and the parity reference
 * for whole GPU batches (outputs kept).
 * ------------------------------------------------------------------------- */
typedef struct {
  const double *p4;
  const int64_t *off;
  const int32_t *p_arr;
  const double *R_arr;
  int n_events, nthreads, tid, method, mm;
  jdef_t def;
  double R;
  int32_t *o1, *o2, *onj;
  double *od, *oj;
  int64_t *opairs;
  int err;
} tb_arg_t;

static void *tb_worker(void *vp) {
  tb_arg_t *a = (tb_arg_t *)vp;
  int64_t maxn = 0;
  for (int e = 0; e < a->n_events; ++e) {
    int64_t m = a->off[e + 1] - a->off[e];
    if (m > maxn) maxn = m;
  }
  int32_t *h1 = (int32_t *)malloc(sizeof(int32_t) * (size_t)(maxn + 1));
  int32_t *h2 = (int32_t *)malloc(sizeof(int32_t) * (size_t)(maxn + 1));
  double *hd = (double *)malloc(sizeof(double) * (size_t)(maxn + 1));
  double *jt = (double *)malloc(sizeof(double) * 4 * (size_t)(maxn + 1));
  if (!h1 || !h2 || !hd || !jt) a->err = 1;
  for (int e = a->tid; e < a->n_events && !a->err; e += a->nthreads) {
    int32_t nj;
    const int64_t b = a->off[e];
    const double *q = a->p4 + 4 * b;
    const int n = (int)(a->off[e + 1] - b);
    jdef_t de = a->def;
    if (a->p_arr) de.p = (double)a->p_arr[e];
    const double Re = a->R_arr ? a->R_arr[e] : a->R;
    t_pair_evals = 0;
    int32_t *w1 = a->o1 ? a->o1 + b : h1, *w2 = a->o2 ? a->o2 + b : h2;
    double *wd = a->od ? a->od + b : hd, *wj = a->oj ? a->oj + 4 * b : jt;
    int rc = a->method ? cluster_tiled(q, n, &de, Re, a->mm, w1, w2, wd, wj, &nj)
                       : cluster_n2plain(q, n, &de, Re, a->mm, w1, w2, wd, wj, &nj);
    if (rc) a->err = 1;
    if (a->onj) a->onj[e] = nj;
    if (a->opairs) a->opairs[e] = a->method ? t_pair_evals : -1;
  }
  free(h1);
  free(h2);
  free(hd);
  free(jt);
  return NULL;
}

static double cluster_batch(const double *p4, const int64_t *offsets, int n_events, jdef_t def, double R,
                            const int32_t *p_arr, const double *R_arr, int nthreads, int method, int math_mode,
                            int32_t *hist_p1, int32_t *hist_p2, double *hist_dij, double *jets, int32_t *njets,
                            int64_t *pair_evals) {
  if (nthreads < 1 || nthreads > 1024 || n_events < 0) return -1.0;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
  tb_arg_t *args = (tb_arg_t *)malloc(sizeof(tb_arg_t) * (size_t)nthreads);
  if (!th || !args) {
    free(th);
    free(args);
    return -1.0;
  }
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int t = 0; t < nthreads; ++t) {
    args[t] = (tb_arg_t){p4, offsets, p_arr, R_arr, n_events, nthreads, t, method, math_mode, def, R,
                         hist_p1, hist_p2, njets, hist_dij, jets, pair_evals, 0};
    pthread_create(&th[t], NULL, tb_worker, &args[t]);
  }
  int err = 0;
  for (int t = 0; t < nthreads; ++t) {
    pthread_join(th[t], NULL);
    err |= args[t].err;
  }
  clock_gettime(CLOCK_MONOTONIC, &t1);
  free(th);
  free(args);
  if (err) return -1.0;
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

double fjo_cluster_batch(const double *p4, const int64_t *offsets, int n_events, int p, double R,
                         const int32_t *p_arr, const double *R_arr, int nthreads, int method, int math_mode,
                         int32_t *hist_p1, int32_t *hist_p2, double *hist_dij, double *jets, int32_t *njets,
                         int64_t *pair_evals) {
  if (p < -1 || p > 1) return -1.0;
  if (p_arr)
    for (int e = 0; e < n_events; ++e)
      if (p_arr[e] < -1 || p_arr[e] > 1) return -1.0;
  const jdef_t d = {(double)p, FJMI_SCHEME_E};
  return cluster_batch(p4, offsets, n_events, d, R, p_arr, R_arr, nthreads, method, math_mode, hist_p1, hist_p2,
                       hist_dij, jets, njets, pair_evals);
}

double fjo_cluster_batch_def(const double *p4, const int64_t *offsets, int n_events, double p, int scheme, double R,
                             const double *R_arr, int nthreads, int method, int32_t *hist_p1, int32_t *hist_p2,
                             double *hist_dij, double *jets, int32_t *njets) {
  const jdef_t d = {p, scheme};
  if (!jdef_ok(&d)) return -1.0;
  return cluster_batch(p4, offsets, n_events, d, R, NULL, R_arr, nthreads, method, FJO_MATH_SHARED, hist_p1, hist_p2,
                       hist_dij, jets, njets, NULL);
}

double fjo_time_batch(const double *p4, const int64_t *offsets, int n_events, int p, double R, int nthreads,
                      int method) {
  return fjo_cluster_batch(p4, offsets, n_events, p, R, NULL, NULL, nthreads, method, FJO_MATH_SHARED, NULL, NULL,
                           NULL, NULL, NULL, NULL);
}
