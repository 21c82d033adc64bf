/*
 * fjmi_kernels.h — the kernels compiled into libfjmi, as seen by the host runtime.
 */
#ifndef FJMI_KERNELS_H
#define FJMI_KERNELS_H

namespace fjmi {

/* One compiled configuration of the dense kernel: T threads per workgroup, K slots
 * per thread (capacity T*K particles), four-momenta in global memory (p4g) or LDS. */
struct KernelDesc {
  int T;
  int K;
  bool p4g;
  int occ; /* minimum waves per SIMD the register allocation was built for */
  const void *fn; /* __global__ void(LaunchArgs) */
};

extern const KernelDesc kDenseKernels[];
extern const int kNumDenseKernels;
extern const void *const kEmptyKernel; /* __global__ void(int) */
extern const void *const kZeroKernel;  /* __global__ void(unsigned *, int, unsigned long long *, int) */

/* tiled engine, __global__ void(LaunchArgs, TiledScratch), 64 threads, by tile
 * table size (index = a tiled bucket row's p4_global field) */
constexpr int kNumTiledKernels = 4;
/* largest event of the tiled engine: labels and particle indices share one
 * 32-bit word (16 bits each) in its tile minima */
constexpr int kTiledMaxN = 65535;
extern const void *const kTiledKernels[kNumTiledKernels];
extern const void *const kTiledServeKernels[kNumTiledKernels]; /* same, serving mode (LaunchArgs::serve) */
extern const void *const kTiledGenKernels[kNumTiledKernels];   /* same, general jet definitions (LaunchArgs::def_*) */
/* large events: 8 waves per event (512 threads; the table's `threads` field of a
 * tiled row), one event per CU */
constexpr int kTiledCoopThreads = 512;
extern const void *const kTiledCoopKernels[kNumTiledKernels];
extern const int kTiledMaxTiles[kNumTiledKernels];

/* post-processing (postproc.hip), 256 threads, one workgroup per event */
extern const void *const kInclusiveKernel;   /* (jets, njets, offsets, ptmin, out, n_out) */
extern const void *const kInclusivePjKernel; /* (records, n, ptmin, out, n_out) */
extern const void *const kHistoryKernel;     /* (offsets, n_single, p1, p2, dij, child, max_dij) */
extern const void *const kExclusiveKernel;
extern const void *const kJetsScanKernel; /* (njets, n, joff[n+1]), one workgroup of 256 */
extern const void *const kJetsPackKernel; /* (jets, offsets, njets, joff, out), one workgroup of 64 per event */   /* (offsets, n_single, p4, p1, p2, dij, mode, param, work, out, n_out) */

/* Pool entries per event of the tiled engine beyond 3 per particle (>= 2 per tile). */
constexpr int kTiledPoolPerEvent = 2048;

/* LDS staging entries used by the p4g kernels for the initial nearest-neighbour pass. */
constexpr int kStageChunk = 2048;

}  // namespace fjmi

#endif
