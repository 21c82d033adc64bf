/*
 * fjmi_internal.h — structures shared between the HIP kernels and the host
 * runtime of libfjmi (not part of the public C ABI).
 */
#ifndef FJMI_INTERNAL_H
#define FJMI_INTERNAL_H

#include <stddef.h>
#include <stdint.h>

namespace fjmi {

/* Arguments of one clustering launch: a queue of events (indices into the
 * packed particle arrays) that persistent workgroups pull from. */
/* Serving: one admitted event, published by the host in fine-grained host memory. */
struct ServeDesc {
  int64_t base;     /* first particle slot in the arenas */
  int64_t pool_off; /* first tile-segment pool entry (3n + pool_per_event entries) */
  int32_t n;
  int32_t p;
  double R;
  int32_t slot; /* event slot: result njets index, completion token - 1 */
  /* written by the host as (descriptor number mod 2^31-1) + 1, and set to 0 by the
   * GPU once it has copied the descriptor: the host writes a cell again only
   * after that, so a claimant never reads a cell that has been reused */
  uint32_t seq;
};

/* Serving: the sequence tag of descriptor number h (never 0, which marks a
 * consumed cell). */
#if defined(__HIPCC__)
__host__ __device__
#endif
inline unsigned serve_seq(unsigned long long h) { return (unsigned)(h % 0x7FFFFFFFull) + 1u; }

/* Serving: words the host writes and the kernels read (fine-grained host memory). */
struct ServeHost {
  unsigned long long tail; /* descriptors published (host: release store) */
  unsigned int stop;       /* the host asks the kernels to leave */
  unsigned int pad;
};

/* Serving: indices of the device-memory control words in ServeArgs::dev, 16
 * words (128 bytes) apart so that idle waves polling one of them do not queue
 * behind atomics on another. */
enum : int {
  kServeHead = 0,        /* descriptors claimed */
  kServeTail = 16,       /* mirror of ServeHost::tail */
  kServeLock = 32,       /* poll lock: one wave at a time reads the host words */
  kServeStop = 48,       /* mirror of ServeHost::stop */
  kServeCompleted = 64,  /* completions pushed */
  kServeLastClaim = 80,  /* wall clock of the last claim */
  kServeBusy = 96,       /* events being clustered now */
  kServePeakBusy = 112,  /* the most ever clustered at once */
  kServeLastPoll = 128,  /* wall clock of the last read of the host words */
  kServeExitEpoch = 144, /* launch epoch whose waves are leaving */
  kServeCtlWords = 160
};

/* Serving: what a serving launch of the tiled kernel needs beside LaunchArgs. */
struct ServeArgs {
  ServeHost *host;        /* fine-grained host memory */
  const ServeDesc *ring;  /* fine-grained host memory, ring_cap entries */
  int32_t *comp;          /* fine-grained host memory, ring_cap entries: slot + 1 per completion, in order */
  int32_t *res_nj;        /* fine-grained host memory, per slot (-1: the event failed) */
  int32_t *res_p1, *res_p2;
  double *res_dij, *res_jets; /* fine-grained host memory, indexed like the particles */
  /* device-memory control words, each in its own 128-byte line (indices kServe*) */
  unsigned long long *dev;
  unsigned int ring_cap;
  unsigned long long idle_ticks; /* wall_clock64 ticks without any claim after which waves leave */
  unsigned long long epoch;      /* this launch's number (1, 2, ...) */
};

/* A general jet definition (the tiled_gen_kernel variants): genkt exponent p
 * (any finite real) and recombination scheme FJMI_SCHEME_*; the layout of the
 * public fjmi_jetdef. */
struct JetDef {
  double p;
  int32_t scheme;
  int32_t pad;
};

struct LaunchArgs {
  /* input particles: event e occupies particles [offsets[e], offsets[e+1]);
   * particle i's (px,py,pz,E) are in[i*in_stride + in_off + 0..3] */
  const double *in;
  int in_stride; /* in doubles: 4 for packed double4, 5 for the PseudoJet record */
  int in_off;    /* in doubles: 0 for packed double4, 1 for the PseudoJet record */
  const int64_t *offsets; /* nullptr: a single event of n_single particles */
  int n_single;
  /* event queue */
  const int32_t *queue; /* event indices, processed in order; nullptr: identity */
  int n_queue;
  unsigned int *head;      /* queue head counter; nullptr: each workgroup takes entry 0 once (one event) */
  unsigned int head_base;  /* value of *head when the launch starts */
  /* per-event parameters (nullptr -> uniform values) */
  const int32_t *p_arr;
  const double *R_arr;
  int p_uni;
  double R_uni;
  /* outputs, indexed like the particles (an event with N particles has N steps) */
  int32_t *hist_p1;
  int32_t *hist_p2;
  double *hist_dij;
  double *jets;   /* 4 doubles per jet, jets of event e start at 4*offsets[e] */
  int32_t *njets; /* per event */
  /* optional in-place PseudoJet output (cluster.h contract), stride 40 bytes */
  unsigned char *pj_out;
  /* global scratch for the four-momenta when they are not kept in LDS: 4 doubles per particle */
  double *p4_scratch;
  /* global scratch for (y, phi, w) when the event does not fit the LDS stage: 4 doubles per particle */
  double *coord_scratch;
  /* error word: bit 0 = event larger than the kernel's capacity, bit 1 = internal,
   * bit 2 = a per-event jet definition the kernel cannot run */
  unsigned int *status;
  /* optional: each workgroup leaving the queue raises this to its wall-clock
   * time (the constant-rate counter of wall_clock64(), 100 MHz on gfx950) */
  unsigned long long *t_end;
  int stage_chunk; /* number of staging entries in LDS when four-momenta live in global memory */
  int tile_q;      /* tiled engine: tile side = max(tile_q, 4)/4 * R (rounded up) */
  ServeArgs serve; /* serving launches of the tiled engine only */
  /* general jet definitions (tiled_gen_kernel only): per event, or nullptr -> def_uni */
  const JetDef *def_arr;
  JetDef def_uni;
  /* tiled_gen_kernel, optional: 4 doubles per particle, the momentum of the
   * history entry each merge step creates (beam steps leave theirs unwritten) */
  double *hist_jets;
};

enum { STATUS_CAPACITY = 1, STATUS_INTERNAL = 2, STATUS_DEFINITION = 4 };

/* The tiled engine's per-jet record: one 64-byte block (half a cache line), so
 * that visiting a jet in a gather or a rescan costs one line, not four. */
struct alignas(64) JetRec {
  double2 rp;  /* rapidity, phi */
  double2 wn;  /* weight, nearest-neighbour distance */
  int4 meta;   /* nearest neighbour (particle index or -1), label, tile, position in the tile segment */
  double dij;  /* diJ = NN distance * min(w, w_NN), or R^2 w without a NN */
  double pad;
};

/* Global per-particle scratch of the tiled engine (indexed by particle like the
 * input), plus the tile-segment pool: kTiledPoolPerParticle entries per particle
 * and pool_per_event per event (event e's pool starts at 3*offsets[e] +
 * e*pool_per_event).  The label map and the pool hold particle indices of one
 * event, below kTiledMaxN = 65535, in 16 bits. */
struct TiledScratch {
  JetRec *rec;           /* 1 per particle */
  double *p4;            /* 4 per particle */
  int *hist;             /* 1 per particle */
  unsigned short *lab2;  /* 1 per particle: label -> particle */
  unsigned short *pool;  /* tile segments */
  int pool_per_event;
};

constexpr size_t kTiledPoolPerParticle = 3;
constexpr size_t kTiledPoolEntryBytes = sizeof(unsigned short);

/* Bytes of tiled scratch per particle: record, four-momentum, history index,
 * label map and its pool entries. */
constexpr size_t kTiledBytesPerParticle =
    sizeof(JetRec) + 32 + 4 + sizeof(unsigned short) + kTiledPoolPerParticle * kTiledPoolEntryBytes;

/* Bytes of tiled scratch for `particles` particles in `events` events. */
constexpr size_t tiled_scratch_bytes(int64_t particles, int64_t events, int pool_per_event) {
  return (size_t)particles * kTiledBytesPerParticle + (size_t)events * pool_per_event * kTiledPoolEntryBytes + 1024;
}

/* Carves a TiledScratch out of one allocation of tiled_scratch_bytes() (the
 * records first, so they keep the allocation's 256-byte alignment). */
inline TiledScratch tiled_scratch_carve(void *mem, int64_t particles, int pool_per_event) {
  char *c = static_cast<char *>(mem);
  TiledScratch t;
  const size_t P = (size_t)particles;
  t.rec = reinterpret_cast<JetRec *>(c);
  c += sizeof(JetRec) * P;
  t.p4 = reinterpret_cast<double *>(c);
  c += 32 * P;
  t.hist = reinterpret_cast<int *>(c);
  c += 4 * P;
  t.lab2 = reinterpret_cast<unsigned short *>(c);
  c += sizeof(unsigned short) * P;
  t.pool = reinterpret_cast<unsigned short *>(c);
  t.pool_per_event = pool_per_event;
  return t;
}

}  // namespace fjmi

#endif
