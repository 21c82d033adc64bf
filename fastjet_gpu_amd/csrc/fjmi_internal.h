/*
 * fjmi_internal.h — structures shared between the HIP kernels and the host
 * runtime of libfjmi (not part of the public C ABI).
 */
#ifndef FJMI_INTERNAL_H
#define FJMI_INTERNAL_H

#include <stdint.h>

namespace fjmi {

/* Arguments of one clustering launch: a queue of events (indices into the
 * packed particle arrays) that persistent workgroups pull from. */
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
  unsigned int *head;      /* queue head counter */
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
  /* error word: bit 0 = event larger than the kernel's capacity */
  unsigned int *status;
  int stage_chunk; /* number of staging entries in LDS when four-momenta live in global memory */
};

enum { STATUS_CAPACITY = 1 };

}  // namespace fjmi

#endif
