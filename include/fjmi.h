/*
 * fjmi.h — C ABI of libfjmi, the MI355X (gfx950) jet-clustering engine.
 *
 * Plain C: pointers, sizes and status codes; no HIP or C++ types in the
 * signatures (a stream is passed as `void *`, a hipStream_t; NULL is the null
 * stream).  Every function returns FJMI_OK (0) or a negative fjmi_status and
 * sets a thread-local message readable with fjmi_last_error().
 *
 * Output convention (all entry points): an event of N particles has exactly N
 * clustering steps.  Step s writes hist_p1[s], hist_p2[s], hist_dij[s]:
 *   pair merge:  (min, max) of the parents' history indices, dij = d_ij / R^2;
 *   beam step:   (parent, -1), dij = d_iB / R^2;
 * history indices are FastJet's: particles 0..N-1, the jet made at step s is N+s.
 * jets[4j..4j+3] = (px, py, pz, E) of the j-th jet promoted to the beam, and
 * njets = number of jets (-1 if a kernel failed on the event: then the calls
 * that report errors also return one).  In the batched calls every per-particle array is
 * indexed like the input particles (event e starts at offsets[e]), and njets has
 * one entry per event.
 *
 * The reference has one entry point, the C++ function cluster() (cluster.h:6 of
 * the reference), bound at link time to one engine (Makefile:31-50).  The
 * functions below replace it and the launch/allocation code inside each of its
 * engines; each cites what it replaces.
 */
#ifndef FJMI_H
#define FJMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fjmi_context fjmi_context;
typedef struct fjmi_batch fjmi_batch;

enum fjmi_status {
  FJMI_OK = 0,
  FJMI_EINVAL = -1,      /* bad argument */
  FJMI_EHIP = -2,        /* a HIP runtime call failed */
  FJMI_ECAPACITY = -3,   /* an event is larger than any configured kernel */
  FJMI_ENOMEM = -4,      /* allocation failed / over the memory budget */
  FJMI_EUNSUPPORTED = -5, /* engine not available */
  FJMI_EAGAIN = -6,       /* serving: the memory budget or the event slots are full; poll, release, retry */
  FJMI_EINTERNAL = -7     /* a kernel found an inconsistent clustering state (its event's output is invalid) */
};

enum fjmi_engine {
  FJMI_ENGINE_AUTO = 0,  /* batches: the engine of the event's multiplicity bucket.  One-event calls:
                          * the engine that finishes one event alone soonest (dense for R >= 0.5 up to
                          * its 8192-particle capacity; below R = 0.5 dense up to N = 500 at R <= 0.2,
                          * rising linearly to 1200 at R = 0.3 and above; tiled otherwise, with 384
                          * tiles, or 512 for R < 0.3 and N > 2000), until fjmi_set_buckets installs
                          * a table, which then decides for them too */
  FJMI_ENGINE_DENSE = 1, /* replaces tri_matrix.cu / n_array.cu */
  FJMI_ENGINE_TILED = 2, /* replaces grid.cu; a single event takes tile side R with 384 tiles (512 for
                          * R < 0.3 and N > 2000) under the default table, else the caller's table's
                          * tiled row for its size, or tile side R and variant 0 when it has none */
  FJMI_ENGINE_EMPTY = 3  /* replaces tri_matrix_empty.cu: per-step launch floor, no result */
};

/* One row of the multiplicity-bucket table: events with N <= max_n (and above
 * the previous row's max_n) run on `engine`:
 *   FJMI_ENGINE_DENSE: `threads` threads per workgroup, `slots` slots per thread
 *     (capacity threads*slots >= max_n), four-momenta in LDS (p4_global = 0) or
 *     global memory (p4_global = 1);
 *   FJMI_ENGINE_TILED: one 64-lane wave per event (threads = 64); `slots` is the
 *     tile side in quarters of R, 4..16 (0 means 4: side R); `p4_global` selects
 *     the kernel variant: 0 = 512 tiles (8 events per CU, the full register
 *     budget), 1 = 384 tiles (10 per CU), 2 = 320 tiles (12 per CU), 3 = 256
 *     tiles (12 per CU).  A smaller tile table takes less LDS, so more events
 *     run per CU; an event whose tiles would not fit gets taller tiles in
 *     rapidity.
 * Replaces the reference's occupancy-derived launch shapes (launch.h:11-60). */
typedef struct {
  int32_t max_n;
  int32_t engine;
  int32_t threads;
  int32_t slots;
  int32_t p4_global;
} fjmi_bucket;

/* Last error message of the calling thread ("" if none). */
const char *fjmi_last_error(void);

/* Library version, major*10000 + minor*100 + patch. */
int fjmi_version(void);

/* Context on HIP device `device`: device buffers, kernels, tuning table.
 * Replaces the per-call cudaMalloc/cudaFree of every engine
 * (tri_matrix.cu:288-299, :331-335; grid.cu:586-593, :636-639). */
int fjmi_create(int device, fjmi_context **out);
int fjmi_destroy(fjmi_context *ctx);

/* Kernel configurations compiled into the library: writes up to `cap` rows
 * (max_n = capacity of the configuration; the tiled engine reports INT32_MAX,
 * one row per tile table size) and returns how many exist. */
int fjmi_kernel_configs(fjmi_bucket *out, int cap);

/* Replace / read the multiplicity-bucket table (rows sorted by max_n).  After
 * fjmi_set_buckets the table also picks the engine of one-event calls with
 * FJMI_ENGINE_AUTO, in place of the latency rule (see fjmi_engine). */
int fjmi_set_buckets(fjmi_context *ctx, const fjmi_bucket *rows, int n_rows);
int fjmi_get_buckets(fjmi_context *ctx, fjmi_bucket *rows, int cap, int *n_rows);

/* Single event, host memory in and out, synchronous.  Replaces one call of the
 * reference's cluster() with the CLI's H2D/D2H around it (main.cc:291-298,
 * tri_matrix.cu:286-336).  hist_* and jets hold n and 4n entries. */
int fjmi_cluster_event(fjmi_context *ctx, const double *p4, int n, int p, double R, int engine, int32_t *hist_p1,
                       int32_t *hist_p2, double *hist_dij, double *jets, int32_t *njets);

/* Single event, device memory in and out, asynchronous on `stream`: one kernel
 * launch (the "plain" one-event-per-launch path).  d_p4 holds 4n doubles.
 * The one-event device entry points (this one and fjmi_cluster_pseudojets) keep
 * their scratch and a status word per stream: calls on different streams may
 * overlap, calls on one stream run in order.  A failure inside the kernel
 * (an event beyond the engine's capacity, an inconsistent state) is reported by
 * the next fjmi_stream_check on that stream, not by the call itself. */
int fjmi_cluster_event_device(fjmi_context *ctx, const double *d_p4, int n, int p, double R, int engine,
                              int32_t *d_hist_p1, int32_t *d_hist_p2, double *d_hist_dij, double *d_jets,
                              int32_t *d_njets, void *stream);

/* Drop-in: in-place clustering of `n` PseudoJet records (40 bytes each) in device
 * memory, the contract of the reference's cluster() (cluster.h:6; output layout
 * of tri_matrix.cu:240-245, 272-284): afterwards records with isJet hold the jets.
 * Asynchronous on `stream`. */
int fjmi_cluster_pseudojets(fjmi_context *ctx, void *d_particles, int n, int p, double R, int engine, void *stream);

/* Same as fjmi_cluster_event_device on the tiled engine, but allocates nothing,
 * ever: no device or host allocation on any call (the context's creation set it
 * up).  No counterpart in the reference; for callers that cannot allocate, such
 * as code under graph capture or tight serving loops.  Its scratch is a static
 * array of the library, one per device, sized for events of up to 20000
 * particles (about 2.2 MB); larger events return FJMI_ECAPACITY.  Every call on
 * a device shares that device's array, so calls that overlap on one device (on
 * different streams, or from several threads or contexts) race and must not be
 * issued; calls in order on one stream are safe.  A kernel failure shows as
 * *d_njets == -1.  Asynchronous on `stream`. */
int fjmi_cluster_event_device_noalloc(fjmi_context *ctx, const double *d_p4, int n, int p, double R,
                                      int32_t *d_hist_p1, int32_t *d_hist_p2, double *d_hist_dij, double *d_jets,
                                      int32_t *d_njets, void *stream);

/* Synchronises `stream` and reports (then clears) any failure raised by the
 * kernels of the one-event device entry points launched on it since the last
 * check: FJMI_OK, FJMI_ECAPACITY or FJMI_EINTERNAL.  The reference's cluster()
 * has no error return; its engines abort on CUDA errors (cudaCheck.h:11-47) and
 * the drop-in cluster() calls this to do the same. */
int fjmi_stream_check(fjmi_context *ctx, void *stream);

/* ---- Results derived from a clustering on the device --------------------------
 * All asynchronous on `stream`, device memory in and out.  A batch passes
 * d_offsets (n_events + 1 entries, as for the batch calls); a single event passes
 * n_events = 1 and d_offsets = NULL (with n_single particles where needed). */

/* Inclusive jets above ptmin, sorted by decreasing pt^2 = px^2 + py^2: the jets
 * with px^2 + py^2 >= ptmin^2 of each event (d_njets[e] jets at 4*offsets[e] of
 * d_jets, as every entry point writes them) go to 4*offsets[e] of d_out_jets,
 * their count to d_out_njets[e] (-1 for an event whose njets is -1).  Equal pt^2
 * keep promotion order.  Replaces the host-side filter and sort of the
 * reference's driver after each cluster() call (main.cc:309-327). */
int fjmi_inclusive_jets(fjmi_context *ctx, int n_events, const int64_t *d_offsets, const double *d_jets,
                        const int32_t *d_njets, double ptmin, double *d_out_jets, int32_t *d_out_njets,
                        void *stream);

/* The same on n PseudoJet records after the drop-in cluster(): the records with
 * isJet and pt^2 >= ptmin^2, whole and sorted, to d_out (n records of room), and
 * their count to *d_count (main.cc:309-327). */
int fjmi_inclusive_pseudojets(fjmi_context *ctx, const void *d_particles, int n, double ptmin, void *d_out,
                              int32_t *d_count, void *stream);

/* FastJet's ClusterSequence history fields implied by the step arrays.  History
 * entries 0..N-1 are the particles, N+s the result of step s.  d_child gets 2
 * entries per particle (event e's 2N entries from 2*offsets[e]): the entry whose
 * step consumed entry k, or -3 (FastJet's Invalid) if none.  d_max_dij gets one
 * per particle: max(dij[0..s]) (FastJet's max_dij_so_far; its particle entries
 * hold 0).  Not in the reference, which exposes no history. */
int fjmi_history_view(fjmi_context *ctx, int n_events, const int64_t *d_offsets, int n_single,
                      const int32_t *d_hist_p1, const int32_t *d_hist_p2, const double *d_hist_dij, int32_t *d_child,
                      double *d_max_dij, void *stream);

/* FastJet's exclusive jets from a complete history (ClusterSequence::
 * exclusive_jets): mode FJMI_EXCLUSIVE_NJETS with param = njets (events with
 * fewer particles get d_out_njets[e] = -1), or FJMI_EXCLUSIVE_DCUT with param =
 * dcut >= 0 (compared with the normalised dij of the history, as FastJet's
 * n_exclusive_jets does).  The jets are the history entries alive at the stop
 * point, in the order of the steps that consume them; their momenta are the
 * E-scheme sums of the history replayed in step order, bit-identical to the
 * clustering's.  d_p4 is the event's input (4 doubles per particle), d_work
 * holds 8 doubles per particle, d_out_jets 4 per particle.  Not in the
 * reference (it has inclusive jets only). */
enum { FJMI_EXCLUSIVE_NJETS = 0, FJMI_EXCLUSIVE_DCUT = 1 };
int fjmi_exclusive_jets(fjmi_context *ctx, int n_events, const int64_t *d_offsets, int n_single, const double *d_p4,
                        const int32_t *d_hist_p1, const int32_t *d_hist_p2, const double *d_hist_dij, int mode,
                        double param, double *d_work, double *d_out_jets, int32_t *d_out_njets, void *stream);

/* Batched clustering of many events in ONE graph launch.  Not in the reference
 * (it clusters one event per call, main.cc:269-368).  The plan sorts events into
 * multiplicity buckets (fjmi_set_buckets), orders each bucket's queue by cost,
 * and captures one hipGraph whose kernels are persistent workgroups pulling
 * events from those queues.  All pointers are device memory and stay bound to
 * the plan: h_offsets (host) and d_offsets (device) both hold n_events+1 entries.
 * d_p / d_R give per-event exponent and radius, or NULL for p_uni / R_uni. */
int fjmi_batch_create(fjmi_context *ctx, int n_events, const int64_t *h_offsets, const double *d_p4,
                      const int64_t *d_offsets, const int32_t *d_p, const double *d_R, int p_uni, double R_uni,
                      int32_t *d_hist_p1, int32_t *d_hist_p2, double *d_hist_dij, double *d_jets, int32_t *d_njets,
                      fjmi_batch **out);
/* Enqueue the captured graph on `stream` (asynchronous). */
int fjmi_batch_launch(fjmi_batch *b, void *stream);
/* Synchronise `stream` and check the kernels' status word. */
int fjmi_batch_sync(fjmi_batch *b, void *stream);
int fjmi_batch_destroy(fjmi_batch *b);

/* Per-bucket statistics of a batch plan, with each bucket kernel's completion
 * time (ms after the graph's queue reset) in the last synchronised launch.
 * Not in the reference (its CLI times a whole cluster() call, main.cc:281-306). */
typedef struct {
  fjmi_bucket bucket; /* the table row (max_n) and the kernel configuration */
  int32_t n_events;
  int32_t grid; /* persistent workgroups launched */
  int64_t particles;
  double sum_n2; /* sum over the bucket's events of N^2 */
  float ms;
} fjmi_bucket_stats;
int fjmi_batch_stats(fjmi_batch *b, fjmi_bucket_stats *out, int cap, int *n_buckets);

/* ---- General jet definitions --------------------------------------------------
 * FastJet's genkt algorithm for any finite real exponent p (weight kt^(2p), with
 * FastJet's kt^2 >= 1e-300 guard for p <= 0; p = 1, 0, -1 give kt, C/A and
 * anti-kt bit for bit) under one of FastJet's recombination schemes (its
 * JetDefinition::DefaultRecombiner: inputs preprocessed as the scheme requires,
 * merged jets built by the scheme).  Not in the reference, which has p in
 * {1, 0, -1} (cluster.h:4) and the E-scheme (tri_matrix.cu:247-250).  These
 * calls run on the tiled engine (a dense row of a caller's table becomes the
 * 384-tile row with side R); outputs as above, with the jets' momenta as the
 * scheme builds them.  The scheme numbers are FastJet's RecombinationScheme
 * values. */
#define FJMI_SCHEME_E 0        /* four-momentum sum */
#define FJMI_SCHEME_PT 1       /* massless (pt, y, phi), pt-weighted y and phi; inputs made massless, E = |p| */
#define FJMI_SCHEME_PT2 2      /* as pt, pt^2-weighted */
#define FJMI_SCHEME_ET 3       /* as pt; inputs made massless by rescaling |p| to E */
#define FJMI_SCHEME_ET2 4      /* as Et, pt^2-weighted */
#define FJMI_SCHEME_BIPT 5     /* as pt, inputs unchanged */
#define FJMI_SCHEME_BIPT2 6    /* as pt2, inputs unchanged */
#define FJMI_SCHEME_WTA_PT 7   /* winner takes all: the harder jet's y, phi and mass, pt = pta + ptb */
#define FJMI_SCHEME_WTA_MODP 8 /* winner takes all: the harder jet's direction, |p| = |pa| + |pb| */
#define FJMI_NUM_SCHEMES 9

typedef struct {
  double p;       /* genkt exponent, finite */
  int32_t scheme; /* FJMI_SCHEME_* */
  int32_t pad;    /* 0 */
} fjmi_jetdef;

/* The calls below take an optional hist_jets output (NULL: none), indexed like
 * the particles, 4 doubles each: step s of a merge writes the momentum of the
 * history entry N + s it creates (beam steps leave theirs unwritten).  These
 * are FastJet's ClusterSequence jets beyond the inputs; fjmi_exclusive_jets_def
 * reads them. */

/* One event, host memory in and out, synchronous (as fjmi_cluster_event). */
int fjmi_cluster_event_def(fjmi_context *ctx, const double *p4, int n, const fjmi_jetdef *def, double R,
                           int32_t *hist_p1, int32_t *hist_p2, double *hist_dij, double *jets, int32_t *njets,
                           double *hist_jets);
/* One event, device memory, asynchronous on `stream` (as fjmi_cluster_event_device;
 * failures are reported by fjmi_stream_check). */
int fjmi_cluster_event_device_def(fjmi_context *ctx, const double *d_p4, int n, const fjmi_jetdef *def, double R,
                                  int32_t *d_hist_p1, int32_t *d_hist_p2, double *d_hist_dij, double *d_jets,
                                  int32_t *d_njets, double *d_hist_jets, void *stream);
/* A batch plan (as fjmi_batch_create) with one definition per event (d_defs,
 * device memory, n_events entries) or one for all (def, host memory; d_defs
 * NULL), and per-event radii d_R or R_uni.  An event whose device-side
 * definition is invalid gets njets = -1 and fails fjmi_batch_sync with
 * FJMI_EINVAL. */
int fjmi_batch_create_def(fjmi_context *ctx, int n_events, const int64_t *h_offsets, const double *d_p4,
                          const int64_t *d_offsets, const fjmi_jetdef *d_defs, const fjmi_jetdef *def,
                          const double *d_R, double R_uni, int32_t *d_hist_p1, int32_t *d_hist_p2,
                          double *d_hist_dij, double *d_jets, int32_t *d_njets, double *d_hist_jets,
                          fjmi_batch **out);
/* A host-memory batch (as fjmi_cluster_batch: pinned chunks, copies overlapped
 * with the kernels) with one definition per event (defs, host memory) or one
 * for all (def; defs NULL), per-event radii R or R_uni, and optionally the
 * history momenta into hist_jets (4 doubles per particle). */
int fjmi_cluster_batch_def(fjmi_context *ctx, int n_events, const int64_t *offsets, const double *p4,
                           const fjmi_jetdef *defs, const fjmi_jetdef *def, const double *R, double R_uni,
                           int32_t *hist_p1, int32_t *hist_p2, double *hist_dij, double *jets, int32_t *njets,
                           double *hist_jets);
/* Exclusive jets (as fjmi_exclusive_jets) of a clustering under a general jet
 * definition with recombination scheme `scheme`: the entries alive at the stop
 * point, particles as the scheme preprocesses them and merged entries from
 * d_hist_jets.  A replay of the history cannot rebuild those for schemes other
 * than E: their merge depends on the order of its operands, which the history
 * (parents as min, max, as FastJet stores them) does not keep.  Not in the
 * reference. */
int fjmi_exclusive_jets_def(fjmi_context *ctx, int n_events, const int64_t *d_offsets, int n_single,
                            const double *d_p4, const int32_t *d_hist_p1, const int32_t *d_hist_p2,
                            const double *d_hist_dij, const double *d_hist_jets, int scheme, int mode, double param,
                            double *d_out_jets, int32_t *d_out_njets, void *stream);

/* Host-memory batch: copies in, clusters, copies out, synchronises.
 * Per-particle outputs are sized like p4/4 (total particles).  The events are
 * taken longest first in up to four chunks, each on its own stream with pinned
 * staging that the context keeps: a chunk's copies overlap the others' kernels,
 * and only the history and each event's njets jets come back.  For a stream of
 * batches already in device memory, alternate two fjmi_batch plans (own
 * outputs) on two streams, so that one batch's tail overlaps the next one's
 * start. */
int fjmi_cluster_batch(fjmi_context *ctx, int n_events, const int64_t *offsets, const double *p4, const int32_t *p,
                       const double *R, int p_uni, double R_uni, int32_t *hist_p1, int32_t *hist_p2, double *hist_dij,
                       double *jets, int32_t *njets);

/* ---- Binary event-batch files ------------------------------------------------
 * Many events' four-momenta (4 doubles per particle) and int64 offsets, with
 * optional per-event exponent p and radius R, for throughput runs and serving
 * replays.  Layout: "FJMIEVT1", uint64 events, uint64 particles, uint32 flags
 * (bit 0: p present, bit 1: R present), uint32 0, int64 offsets[E+1],
 * double p4[4P], int32 p[E], double R[E]; little-endian, no padding.  The
 * reference reads text events only (main.cc:34-85); its text files convert to
 * this form losslessly.  Host memory; no GPU needed. */
int fjmi_events_write(const char *path, int n_events, const int64_t *offsets, const double *p4, const int32_t *p,
                      const double *R);
/* Sizes of a file's contents; has_p / has_R may be NULL. */
int fjmi_events_info(const char *path, int *n_events, int64_t *n_particles, int *has_p, int *has_R);
/* Reads a file into caller buffers sized from fjmi_events_info (p / R may be
 * NULL to skip them); checks magic, sizes, offsets, exponents and radii. */
int fjmi_events_read(const char *path, int64_t *offsets, double *p4, int32_t *p, double *R);

/* ---- Serving: continuous batching of arriving events ------------------------
 * Not in the reference (its CLI clusters one event per call, main.cc:269-368).
 * A server keeps a fixed device-memory arena for the state of admitted events
 * (per particle: tiled scratch, history and jets; per event: a tile-segment
 * pool) and pinned host memory for their inputs and results.  submit() copies
 * an event into pinned memory and publishes it; persistent tiled-engine
 * workgroups on the GPU claim published events as they free up, cluster them,
 * copy the results to pinned memory and announce each completion, so events of
 * any size flow through without batch boundaries.  An event is admitted only
 * if its state fits the budget; otherwise submit() returns FJMI_EAGAIN and the
 * caller keeps it queued.  The GPU workgroups leave together once no event has
 * been claimed for idle_us and none is being clustered, and are relaunched by
 * the next submit() or poll().
 * A server is driven by one host thread; distinct servers are independent. */
typedef struct fjmi_server fjmi_server;

typedef struct {
  int64_t device_budget; /* bytes of device memory for event state (0: 16 GiB) */
  int64_t host_budget;   /* bytes of pinned host memory for inputs and results (0: 8 GiB) */
  int32_t max_events;    /* event slots, i.e. events admitted at once (0: 65536) */
  int32_t variant;       /* tiled kernel variant, as the p4_global field of a tiled bucket row
                          * (1 serves the most events per second; 0 has a slightly shorter
                          * per-event latency at light load) */
  int32_t tile_q;        /* tile side in quarters of R (0: R) */
  int32_t idle_us;       /* microseconds without any claimed event before the GPU workgroups leave (0: 20000) */
} fjmi_server_config;

typedef struct {
  uint64_t tag;    /* as given to fjmi_server_submit */
  int32_t handle;  /* pass to fjmi_server_release once the result has been read */
  int32_t n;       /* particles */
  int32_t njets;   /* jets, or -1 if the clustering failed */
  int32_t pad;
  /* the event's results in pinned host memory, valid until fjmi_server_release:
   * n history steps and njets jets, laid out as in the other entry points */
  const int32_t *hist_p1;
  const int32_t *hist_p2;
  const double *hist_dij;
  const double *jets;
} fjmi_server_result;

typedef struct {
  int64_t particle_capacity; /* particles the arenas hold at once */
  int32_t event_capacity;
  int32_t grid;              /* workgroups of a serving launch */
  int64_t device_bytes;      /* device memory held by the server */
  int64_t host_bytes;        /* pinned host memory held by the server */
  int64_t submitted, completed;
  int64_t in_flight;         /* admitted and not yet released */
  int64_t particles_in_flight, peak_particles_in_flight;
  int64_t launches;          /* serving kernel launches so far */
  int64_t gpu_peak_busy;     /* the most events the GPU has clustered at once */
} fjmi_server_info;

/* cfg may be NULL (all defaults). */
int fjmi_server_create(fjmi_context *ctx, const fjmi_server_config *cfg, fjmi_server **out);
/* Admits one event (p4: 4n doubles of host memory, copied before returning):
 * FJMI_OK, FJMI_EAGAIN (no room now), or an error. */
int fjmi_server_submit(fjmi_server *s, const double *p4, int n, int p, double R, uint64_t tag);
/* Completed events since the last poll, in completion order (non-blocking);
 * writes up to cap results and their count to *n_out. */
int fjmi_server_poll(fjmi_server *s, fjmi_server_result *out, int cap, int *n_out);
/* Frees a completed event's memory and slot. */
int fjmi_server_release(fjmi_server *s, int32_t handle);
int fjmi_server_info_get(fjmi_server *s, fjmi_server_info *out);
/* Stops the GPU workgroups (waiting for events already claimed) and frees everything. */
int fjmi_server_destroy(fjmi_server *s);

#ifdef __cplusplus
}
#endif

#endif /* FJMI_H */
