/*
 * postproc.hip — results derived on the GPU from a clustering's output, one
 * workgroup of 256 threads per event:
 *
 *   inclusive jets   the jets with pt^2 >= ptmin^2 in order of decreasing pt^2:
 *                    the filter and sort the reference's driver runs on the host
 *                    after cluster() (main.cc:309-327), for packed jets or in
 *                    place of the PseudoJet records' isJet flags;
 *   history view     FastJet's ClusterSequence fields that the step arrays imply:
 *                    each history entry's child and max_dij_so_far;
 *   exclusive jets   FastJet's exclusive_jets(njets) / exclusive_jets(dcut): the
 *                    entries alive at the history's stop point, their momenta the
 *                    E-scheme sums of the history replayed in step order, or, for
 *                    a general jet definition, the momenta its clustering stored
 *                    per merge step (FastJet's jets of the ClusterSequence; a
 *                    replay cannot rebuild them: a scheme's merge depends on the
 *                    order of its two operands, which the history does not keep).
 *
 * Every comparison and sum follows the oracle's restatements
 * (oracle/fj_oracle.c: fjo_inclusive_jets, fjo_history_children,
 * fjo_exclusive_jets) operation for operation; ties in the sort keep the lower
 * index first.
 */
#include <hip/hip_runtime.h>

#include "fjmi_internal.h"
#include "fjmi_kernels.h"
#include "fjmi_math.h"

namespace fjmi {

namespace post {

constexpr int T = 256;

__device__ __forceinline__ int64_t ev_base(const int64_t *offsets, int e) { return offsets ? offsets[e] : 0; }
__device__ __forceinline__ int ev_size(const int64_t *offsets, int e, int n_single) {
  return offsets ? (int)(offsets[e + 1] - offsets[e]) : n_single;
}

/* Sum over the workgroup (all threads get the total). */
__device__ int block_sum(int v, int *scratch) {
  __syncthreads();
  scratch[threadIdx.x] = v;
  __syncthreads();
  for (int s = T / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) scratch[threadIdx.x] += scratch[threadIdx.x + s];
    __syncthreads();
  }
  const int r = scratch[0];
  __syncthreads();
  return r;
}

/* Rank of key (at position i) among the kept keys, in decreasing order with the
 * lower position first on ties; keys of jets not kept are staged as -1 (every
 * kept key is >= 0).  All threads of the workgroup call it with the same count. */
__device__ int rank_of(double key, int i, bool kept, int count, double *skey, const void *src, int stride_doubles,
                       int off_doubles, const unsigned char *flags, int flag_stride, double pt2min) {
  int rank = 0;
  for (int j0 = 0; j0 < count; j0 += T) {
    __syncthreads();
    const int j = j0 + (int)threadIdx.x;
    double kj = -1.0;
    if (j < count) {
      const double *q = static_cast<const double *>(src) + (size_t)j * stride_doubles + off_doubles;
      const double k = q[0] * q[0] + q[1] * q[1];
      const bool f = flags ? flags[(size_t)j * flag_stride] != 0 : true;
      kj = (f && !(k < pt2min)) ? k : -1.0;
    }
    skey[threadIdx.x] = kj;
    __syncthreads();
    const int m = count - j0 < T ? count - j0 : T;
    if (kept)
      for (int jj = 0; jj < m; ++jj) {
        const double k2 = skey[jj];
        rank += (k2 > key) | ((k2 == key) & (j0 + jj < i));
      }
  }
  return rank;
}

}  // namespace post

using namespace post;

/* Inclusive jets of packed results: event e's njets[e] jets at 4*offsets[e] of
 * `jets`; the kept ones, sorted, to 4*offsets[e] of `out`, their count to
 * n_out[e] (-1 when njets[e] < 0, a failed clustering). */
__global__ __launch_bounds__(256) void inclusive_kernel(const double *jets, const int32_t *njets,
                                                        const int64_t *offsets, double ptmin, double *out,
                                                        int32_t *n_out) {
  __shared__ double skey[T];
  __shared__ int sred[T];
  const int e = blockIdx.x;
  const int64_t base = ev_base(offsets, e);
  const int J = njets[e];
  if (J < 0) {
    if (threadIdx.x == 0) n_out[e] = -1;
    return;
  }
  const double pt2min = ptmin * ptmin;
  const double *src = jets + 4 * base;
  int kept_total = 0;
  for (int i0 = 0; i0 < J; i0 += T) {
    const int i = i0 + (int)threadIdx.x;
    double key = 0.0;
    double4 q = make_double4(0.0, 0.0, 0.0, 0.0);
    bool kept = false;
    if (i < J) {
      q = *reinterpret_cast<const double4 *>(src + 4 * (size_t)i);
      key = q.x * q.x + q.y * q.y;
      kept = !(key < pt2min);
    }
    const int rank = rank_of(key, i, kept, J, skey, src, 4, 0, nullptr, 0, pt2min);
    if (kept) *reinterpret_cast<double4 *>(out + 4 * (base + rank)) = q;
    kept_total += block_sum(kept ? 1 : 0, sred);
  }
  if (threadIdx.x == 0) n_out[e] = kept_total;
}

/* Inclusive jets of PseudoJet records (40 bytes: int index, bool isJet, then
 * px, py, pz, E at byte 8) after the drop-in cluster(): the records with isJet
 * and pt^2 >= ptmin^2, sorted, copied whole to `out`; their count to *n_out. */
__global__ __launch_bounds__(256) void inclusive_pj_kernel(const unsigned char *recs, int n, double ptmin,
                                                           unsigned char *out, int32_t *n_out) {
  __shared__ double skey[T];
  __shared__ int sred[T];
  const double pt2min = ptmin * ptmin;
  const double *src = reinterpret_cast<const double *>(recs); /* record j: doubles 5j+1 .. 5j+4 */
  int kept_total = 0;
  for (int i0 = 0; i0 < n; i0 += T) {
    const int i = i0 + (int)threadIdx.x;
    double key = 0.0;
    bool kept = false;
    if (i < n) {
      const double *q = src + 5 * (size_t)i + 1;
      key = q[0] * q[0] + q[1] * q[1];
      kept = recs[40 * (size_t)i + 4] != 0 && !(key < pt2min);
    }
    const int rank = rank_of(key, i, kept, n, skey, src, 5, 1, recs + 4, 40, pt2min);
    if (kept) {
      const uint2 *a = reinterpret_cast<const uint2 *>(recs + 40 * (size_t)i);
      uint2 *b = reinterpret_cast<uint2 *>(out + 40 * (size_t)rank);
      for (int w = 0; w < 5; ++w) b[w] = a[w];
    }
    kept_total += block_sum(kept ? 1 : 0, sred);
  }
  if (threadIdx.x == 0) *n_out = kept_total;
}

/* History view: child[2*base + k] for the 2N entries of event e (-3: none) and
 * max_dij[base + s], the running maximum of dij over steps 0..s. */
__global__ __launch_bounds__(256) void history_kernel(const int64_t *offsets, int n_single, const int32_t *p1,
                                                      const int32_t *p2, const double *dij, int32_t *child,
                                                      double *max_dij) {
  __shared__ double sm[T];
  const int e = blockIdx.x;
  const int64_t base = ev_base(offsets, e);
  const int N = ev_size(offsets, e, n_single);
  int32_t *c = child + 2 * base;
  for (int k = threadIdx.x; k < 2 * N; k += T) c[k] = -3;
  __syncthreads();
  for (int s = threadIdx.x; s < N; s += T) {
    const int a = p1[base + s], b = p2[base + s];
    if (a >= 0 && a < N + s) c[a] = N + s;
    if (b >= 0 && b < N + s) c[b] = N + s;
  }
  /* running maximum in chunks of T: Hillis-Steele within the chunk, then the carry */
  double carry = 0.0;
  for (int s0 = 0; s0 < N; s0 += T) {
    const int s = s0 + (int)threadIdx.x;
    double v = s < N ? dij[base + s] : 0.0;
    v = v > carry ? v : carry;
    __syncthreads();
    sm[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < T; d <<= 1) {
      const double o = (int)threadIdx.x >= d ? sm[threadIdx.x - d] : 0.0;
      __syncthreads();
      v = o > v ? o : v;
      sm[threadIdx.x] = v;
      __syncthreads();
    }
    if (s < N) max_dij[base + s] = v;
    carry = sm[T - 1];
  }
}

/* Exclusive jets of event e (mode 0: njets = param; mode 1: dcut = param >= 0),
 * written to 4*base of `out`, their count to n_out[e] (-1 if njets > N).
 * E-scheme history (hist_jets null): work holds 8 doubles per particle (the 2N
 * history momenta), replayed in step order.  General jet definition: the
 * particles as `scheme` preprocesses them and the merge steps' momenta from
 * hist_jets (4 doubles per particle, as the clustering wrote them).  The replay
 * and the selection are sequential in history order, in one thread. */
__global__ __launch_bounds__(256) void exclusive_kernel(const int64_t *offsets, int n_single, const double *p4,
                                                        const int32_t *p1, const int32_t *p2, const double *dij,
                                                        int mode, double param, double *work, double *out,
                                                        int32_t *n_out, const double *hist_jets, int scheme) {
  const int e = blockIdx.x;
  const int64_t base = ev_base(offsets, e);
  const int N = ev_size(offsets, e, n_single);
  const double4 *in = reinterpret_cast<const double4 *>(p4 + 4 * base);
  /* the stop point */
  int stop;
  if (mode == 0) {
    const int nj = (int)param;
    if (nj > N) {
      if (threadIdx.x == 0) n_out[e] = -1;
      return;
    }
    stop = 2 * N - nj;
  } else {
    /* the first step whose running maximum exceeds dcut; the entries before it
     * (particles hold 0 <= dcut) end at the stop point */
    stop = 2 * N;
    double m = 0.0;
    if (threadIdx.x == 0)
      for (int s = 0; s < N; ++s) {
        const double d = dij[base + s];
        m = d > m ? d : m;
        if (m > param) {
          stop = N + s;
          break;
        }
      }
  }
  if (hist_jets) { /* a general jet definition: the stored momenta, no replay */
    if (threadIdx.x != 0) return;
    const double4 *hj = reinterpret_cast<const double4 *>(hist_jets + 4 * base);
    auto entry = [&](int x) {
      if (x >= N) return hj[x - N];
      double4 q = in[x];
      fjmi_preprocess(scheme, &q.x, &q.y, &q.z, &q.w);
      return q;
    };
    int k = 0;
    double4 *o = reinterpret_cast<double4 *>(out + 4 * base);
    for (int en = stop; en < 2 * N; ++en) {
      const int s = en - N;
      const int a = p1[base + s], b = p2[base + s];
      if (a >= 0 && a < stop && k < N) o[k++] = entry(a);
      if (b < stop && b > 0 && k < N) o[k++] = entry(b);
    }
    n_out[e] = k;
    return;
  }
  double4 *h = reinterpret_cast<double4 *>(work + 8 * base);
  for (int k = threadIdx.x; k < N; k += T) h[k] = in[k];
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int s = 0; s + N < stop; ++s) {
    const int a = p1[base + s], b = p2[base + s];
    if (b >= 0 && a >= 0 && a < N + s && b < N + s) { /* a valid history always passes these bounds */
      const double4 x = h[a], y = h[b];
      h[N + s] = make_double4(x.x + y.x, x.y + y.y, x.z + y.z, x.w + y.w);
    }
  }
  int k = 0; /* at most N: each entry before the stop point is consumed once */
  double4 *o = reinterpret_cast<double4 *>(out + 4 * base);
  for (int en = stop; en < 2 * N; ++en) {
    const int s = en - N;
    const int a = p1[base + s], b = p2[base + s];
    if (a >= 0 && a < stop && k < N) o[k++] = h[a];
    if (b < stop && b > 0 && k < N) o[k++] = h[b];
  }
  n_out[e] = k;
}

/* Jets of a batch packed back to back (for copies to the host): joff[e] = sum
 * of max(njets, 0) over the events before e, joff[n] the total.  One workgroup. */
__global__ __launch_bounds__(256) void jets_scan_kernel(const int32_t *njets, int n, int32_t *joff) {
  __shared__ int sm[T];
  int carry = 0;
  for (int e0 = 0; e0 < n; e0 += T) {
    const int e = e0 + (int)threadIdx.x;
    const int v = e < n && njets[e] > 0 ? njets[e] : 0;
    __syncthreads();
    sm[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < T; d <<= 1) { /* inclusive scan of the chunk */
      const int o = (int)threadIdx.x >= d ? sm[threadIdx.x - d] : 0;
      __syncthreads();
      sm[threadIdx.x] += o;
      __syncthreads();
    }
    if (e < n) joff[e] = carry + sm[threadIdx.x] - v;
    carry += sm[T - 1];
  }
  if (threadIdx.x == 0) joff[n] = carry;
}

/* Copies event e's njets[e] jets from 4*offsets[e] of `jets` to 4*joff[e] of `out`. */
__global__ __launch_bounds__(64) void jets_pack_kernel(const double *jets, const int64_t *offsets,
                                                       const int32_t *njets, const int32_t *joff, double *out) {
  const int e = blockIdx.x;
  const int nd = 4 * (njets[e] > 0 ? njets[e] : 0);
  const double *src = jets + 4 * offsets[e];
  double *dst = out + 4 * (int64_t)joff[e];
  for (int i = threadIdx.x; i < nd; i += 64) dst[i] = src[i];
}

const void *const kJetsScanKernel = reinterpret_cast<const void *>(&jets_scan_kernel);
const void *const kJetsPackKernel = reinterpret_cast<const void *>(&jets_pack_kernel);
const void *const kInclusiveKernel = reinterpret_cast<const void *>(&inclusive_kernel);
const void *const kInclusivePjKernel = reinterpret_cast<const void *>(&inclusive_pj_kernel);
const void *const kHistoryKernel = reinterpret_cast<const void *>(&history_kernel);
const void *const kExclusiveKernel = reinterpret_cast<const void *>(&exclusive_kernel);

}  // namespace fjmi
