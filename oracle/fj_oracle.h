/*
 * fj_oracle.h — CPU oracle for sequential-recombination jet clustering.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the CLI,
 * the drop-in cluster()) may include, link or call this code.  It is used by
 * tests/ as the parity checker, by __graft_entry__.smoke(), and by bench.py's
 * CPU-baseline leg.
 *
 * What it restates:
 *  - FastJet 3.3.2's N2Plain strategy (ClusterSequence::_simple_N2_cluster, the
 *    published nearest-neighbour method of Cacciari & Salam, hep-ph/0512210),
 *    which is the algorithm the reference engines approximate on the GPU
 *    (README.md:1-4; tri_matrix.cu:286-336 all-pairs form; n_array.cu:129-200).
 *    Kinematics and distances follow tri_matrix.cu:56-113 and grid.cu:110-201.
 *    Recombination / compaction follow tri_matrix.cu:240-255 (beam: the jet is
 *    removed and the last slot moves into its place; pair: the merged jet takes
 *    the lower slot and the last slot moves into the higher one).
 *  - The FastJet history numbering: particles are 0..N-1, the jet made at step s
 *    has history index N+s; a step records (min parent, max parent) or
 *    (parent, -1) for a beam step, and dij = min diJ / R^2.
 *
 * Output convention shared with the GPU library (include/fjmi.h):
 *   hist_p1[s], hist_p2[s], hist_dij[s] for s = 0..N-1 (there are exactly N steps);
 *   jets[4*j .. 4*j+3] = (px,py,pz,E) of the j-th jet promoted to the beam;
 *   *njets = number of jets.
 */
#ifndef FJ_ORACLE_H
#define FJ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* math_mode: 0 = the shared correctly-rounded fjmi_math.h functions (what the GPU
 * uses), 1 = glibc atan2/log (what FastJet uses). */
enum { FJO_MATH_SHARED = 0, FJO_MATH_LIBM = 1 };

/* FastJet N2Plain restatement. Returns 0 on success, negative on bad arguments. */
int fjo_cluster_n2plain(const double *p4, int n, int p, double R, int math_mode, int32_t *hist_p1,
                        int32_t *hist_p2, double *hist_dij, double *jets, int32_t *njets);

/* Same semantics, but every step recomputes every nearest neighbour from scratch
 * (O(N^3)).  Equal to n2plain except on exact nearest-neighbour distance ties. */
int fjo_cluster_brute(const double *p4, int n, int p, double R, int math_mode, int32_t *hist_p1, int32_t *hist_p2,
                      double *hist_dij, double *jets, int32_t *njets);

/* Textbook generalised-kt: each step takes the smallest of all pair distances
 * min(w_i,w_j) dR^2 (dR^2 < R^2) and beam distances w_i R^2, with no nearest-
 * neighbour bookkeeping and no slot compaction (live jets kept in creation order,
 * ties -> first found).  Equal to n2plain except on exact ties. */
int fjo_cluster_pairwise(const double *p4, int n, int p, double R, int math_mode, int32_t *hist_p1,
                         int32_t *hist_p2, double *hist_dij, double *jets, int32_t *njets);

/* N2Plain decisions with (rapidity, phi) tiling of side >= R for the nearest-
 * neighbour bookkeeping (N2Plain slot numbers kept as labels for every tie). */
int fjo_cluster_tiled(const double *p4, int n, int p, double R, int math_mode, int32_t *hist_p1, int32_t *hist_p2,
                      double *hist_dij, double *jets, int32_t *njets);

/* The same four methods for any jet definition: the generalised-kt exponent p
 * (1 kt, 0 C/A, -1 anti-kt, any other finite real genkt, with FastJet's
 * kt^2 >= 1e-300 guard for p <= 0) and a recombination scheme FJMI_SCHEME_*
 * (FastJet's DefaultRecombiner: inputs preprocessed as the scheme requires,
 * merged jets built by fjmi_recombine).  Neither is in the reference, which has
 * p in {1, 0, -1} and the E-scheme (cluster.h:4, tri_matrix.cu:247-250). */
int fjo_cluster_n2plain_def(const double *p4, int n, double p, int scheme, double R, int math_mode, int32_t *hist_p1,
                            int32_t *hist_p2, double *hist_dij, double *jets, int32_t *njets);
int fjo_cluster_brute_def(const double *p4, int n, double p, int scheme, double R, int math_mode, int32_t *hist_p1,
                          int32_t *hist_p2, double *hist_dij, double *jets, int32_t *njets);
int fjo_cluster_pairwise_def(const double *p4, int n, double p, int scheme, double R, int math_mode, int32_t *hist_p1,
                             int32_t *hist_p2, double *hist_dij, double *jets, int32_t *njets);
int fjo_cluster_tiled_def(const double *p4, int n, double p, int scheme, double R, int math_mode, int32_t *hist_p1,
                          int32_t *hist_p2, double *hist_dij, double *jets, int32_t *njets);
/* The same, and the momentum of the history entry each merge step creates into
 * hist_jets[4 s .. 4 s + 3] (beam steps leave theirs untouched): FastJet's
 * ClusterSequence jets beyond the inputs. */
int fjo_cluster_n2plain_defh(const double *p4, int n, double p, int scheme, double R, int math_mode, int32_t *hist_p1,
                             int32_t *hist_p2, double *hist_dij, double *jets, int32_t *njets, double *hist_jets);
int fjo_cluster_brute_defh(const double *p4, int n, double p, int scheme, double R, int math_mode, int32_t *hist_p1,
                           int32_t *hist_p2, double *hist_dij, double *jets, int32_t *njets, double *hist_jets);
int fjo_cluster_pairwise_defh(const double *p4, int n, double p, int scheme, double R, int math_mode,
                              int32_t *hist_p1, int32_t *hist_p2, double *hist_dij, double *jets, int32_t *njets,
                              double *hist_jets);
int fjo_cluster_tiled_defh(const double *p4, int n, double p, int scheme, double R, int math_mode, int32_t *hist_p1,
                           int32_t *hist_p2, double *hist_dij, double *jets, int32_t *njets, double *hist_jets);
/* A batch with one jet definition (and optional per-event R) on nthreads host threads. */
double fjo_cluster_batch_def(const double *p4, const int64_t *offsets, int n_events, double p, int scheme, double R,
                             const double *R_arr, int nthreads, int method, int32_t *hist_p1, int32_t *hist_p2,
                             double *hist_dij, double *jets, int32_t *njets);

/* Kinematics exported for tests. */
double fjo_log(double x);
double fjo_atan2(double y, double x);
void fjo_set_jet(const double *p4, int p, int math_mode, double *rap, double *phi, double *w);
/* The functions of general jet definitions (fjmi_math.h), exported for tests. */
double fjo_exp(double x);
double fjo_pow(double x, double p); /* x > 0 finite, p finite */
double fjo_sqrt(double x);
double fjo_sin(double x);
double fjo_cos(double x);
double fjo_weight_gen(double kt2, double p);

/* Clusters n_events events (p4 packed, offsets[n_events+1] in particles) on
 * nthreads threads with n2plain (method 0) or tiled (method 1) in math mode
 * math_mode; per-event p_arr / R_arr, or NULL for p / R.  Outputs are indexed
 * like the particles (jets of event e from 4*offsets[e]), njets per event, and
 * for the tiled method the pair distances each event evaluated (-1 for
 * n2plain); any output may be NULL (discarded).  Returns wall seconds, or -1. */
double fjo_cluster_batch(const double *p4, const int64_t *offsets, int n_events, int p, double R,
                         const int32_t *p_arr, const double *R_arr, int nthreads, int method, int math_mode,
                         int32_t *hist_p1, int32_t *hist_p2, double *hist_dij, double *jets, int32_t *njets,
                         int64_t *pair_evals);

/* CPU baseline: fjo_cluster_batch with shared math and discarded outputs;
 * returns wall seconds, or -1. */
double fjo_time_batch(const double *p4, const int64_t *offsets, int n_events, int p, double R, int nthreads,
                      int method);

/* Inclusive jets above ptmin sorted by decreasing pt^2 (main.cc:309-327; equal
 * keys keep promotion order): writes n_out jets to out. */
int fjo_inclusive_jets(const double *jets, int njets, double ptmin, double *out, int32_t *n_out);

/* FastJet's ClusterSequence history view: child[2n] (-3 = none) and max_dij[n]
 * (max_dij_so_far of the step entries). */
int fjo_history_children(const int32_t *hp1, const int32_t *hp2, const double *hdij, int n, int32_t *child,
                         double *max_dij);

/* FastJet's exclusive jets from a complete history: mode 0 njets = param,
 * mode 1 dcut = param.  out holds up to n jets. */
int fjo_exclusive_jets(const double *p4, const int32_t *hp1, const int32_t *hp2, const double *hdij, int n, int mode,
                       double param, double *out, int32_t *n_out);

/* FastJet's exclusive jets of a clustering under a general jet definition
 * (recombination scheme `scheme`), from its stored history momenta hist_jets
 * (fjo_cluster_*_defh): mode and param as fjo_exclusive_jets. */
int fjo_exclusive_jets_def(const double *p4, const int32_t *hp1, const int32_t *hp2, const double *hdij,
                           const double *hist_jets, int n, int scheme, int mode, double param, double *out,
                           int32_t *n_out);

/* Pair distances the tiled restatement evaluates clustering one event (its
 * initial nearest neighbours, orphan rescans and merged-jet updates): the
 * algorithmic work of the tiled method.  -1 on error. */
int64_t fjo_tiled_pair_evals(const double *p4, int n, int p, double R);

#ifdef __cplusplus
}
#endif

#endif
