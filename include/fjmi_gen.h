/*
 * fjmi_gen.h — seeded synthetic event generator (host only).
 *
 * Not part of the reference's interface: the reference reads its events from
 * text files (main.cc:34-85).  This generator produces the benchmark and test
 * workloads deterministically so that the GPU paths and the CPU baseline
 * consume identical bytes.  Workload numbers 1-5 are described in
 * fastjet_gpu_amd/csrc/fjmi_gen.c.
 */
#ifndef FJMI_GEN_H
#define FJMI_GEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Multiplicity of event `event` of workload `workload` (1..5), or -1. */
int fjmi_gen_event_size(int workload, int64_t event);

/* Writes (px,py,pz,E) of each particle to p4[4*i..4*i+3]; returns N, -1 for an
 * unknown workload, -2 if N > capacity. */
int fjmi_gen_event(int workload, int64_t event, double *p4, int capacity);

/* Serving attributes of event `event`: generalised-kt exponent p, radius R and a
 * unit-mean exponential inter-arrival gap. */
int fjmi_gen_serving_attrs(int64_t event, int *p, double *R, double *gap);

#ifdef __cplusplus
}
#endif

#endif
