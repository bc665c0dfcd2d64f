// cf_sched.hip -- cost-ordered work lists of the tuned pair kernel (gfx950).
//
// The persistent pair kernel pulls work items from a queue in list order, so
// the order and the size of the items set its load balance. The scheduler
//   1. computes kNumFeatures features per cell pair              (k_cp_features)
//   2. predicts each cell pair's cost, in log2 units, either with the plain
//      N1*N2 heuristic (k_cost_heuristic) or with the cost model's network
//      (k_mlp_predict, cf_model.hip)
//   3. moves the heavy cell pairs (>= kHeavyFactor x the mean cost) to the front,
//      heaviest first, with a counting sort over quarter-octave buckets (block-local
//      LDS histograms, a scan, block-local placement); the rest keep list order
//                                   (k_cost_total, k_cost_hist, scan, k_cost_place)
//   4. splits heavy cell pairs into parts of their secondary window
//      (k_count_chunks / k_emit_chunks with the cost array).
// The counts do not depend on the order or the split: every (i, j) pair is
// still tested exactly once. Every kernel takes an optional device count
// (`ncp_dev`) that bounds `ncp`: a graph-captured batch launches them over its
// capacity and the count its earlier steps wrote.
#include "cf_device.h"

namespace cfamd {

__global__ void k_cp_features(int ncp, const CellPairRec* __restrict__ cps, const DevLattice* __restrict__ lats,
                              int batch, PairParams P, float* __restrict__ feats,
                              const int* __restrict__ ncp_dev = nullptr)
{
    if (ncp_dev) ncp = min(ncp, *ncp_dev);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncp) return;
    const CellPairRec cp = cps[t];
    const DevLattice& A = lats[batch ? cp.pad : 0];
    const DevLattice& B = lats[batch ? cp.pad : 1];
    const double n1 = (double)cell_count(A.bstart, cp.c1), n2 = (double)cell_count(B.bstart, cp.c2);
    const double* b1 = A.bounds + (int64_t)cp.c1 * 6;
    const double* b2 = B.bounds + (int64_t)cp.c2 * 6;
    double maxsep2;
    const int dense = cp_dense(P, cp, b1, b2, &maxsep2);
    const int is2d = P.mode == MODE_RPPI || P.mode == MODE_WP;
    const double minsep2 = cp.mdx * cp.mdx + cp.mdy * cp.mdy + (is2d ? 0.0 : cp.mdz * cp.mdz);
    const double oz = cp.wz * P.zwrap;
    const double zlo = b1[4] + oz - P.zreach, zhi = b1[5] + oz + P.zreach;
    const double ext2 = b2[5] - b2[4];
    double zfrac = 1.0;
    if (ext2 > 0) {
        const double ov = fmin(zhi, b2[5]) - fmax(zlo, b2[4]);
        zfrac = fmin(1.0, fmax(0.0, ov / ext2));
    }
    const double inv = P.sqr_rmax > 0 ? 1.0 / P.sqr_rmax : 0.0;
    float* f = feats + (int64_t)t * kNumFeatures;
    f[0] = (float)log2(1.0 + n1);
    f[1] = (float)log2(1.0 + n2);
    f[2] = (float)log2(fmax(1.0, ceil(n1 / 64.0)));
    f[3] = (float)cp.same;
    f[4] = (float)dense;
    f[5] = (float)fmin(8.0, minsep2 * inv);
    f[6] = (float)fmin(8.0, maxsep2 * inv);
    f[7] = (float)zfrac;
    f[8] = (float)is2d;
    f[9] = (float)log2(1.0 + n2 * zfrac);
}

// The plain heuristic: log2 of the number of pairs tested, N1*N2 (N1(N1-1)/2
// for a cell with itself).
__global__ void k_cost_heuristic(int ncp, const CellPairRec* __restrict__ cps, const DevLattice* __restrict__ lats,
                                 int batch, float* __restrict__ cost, const int* __restrict__ ncp_dev = nullptr)
{
    if (ncp_dev) ncp = min(ncp, *ncp_dev);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncp) return;
    const CellPairRec cp = cps[t];
    const double n1 = (double)cell_count(lats[batch ? cp.pad : 0].bstart, cp.c1);
    const double n2 = (double)cell_count(lats[batch ? cp.pad : 1].bstart, cp.c2);
    const double pairs = cp.same ? 0.5 * n1 * (n1 - 1.0) : n1 * n2;
    cost[t] = (float)log2(1.0 + pairs);
}

// Sort keys: cell pairs at least `heavy` times the mean cost go first,
// heaviest first (4 buckets per octave); all others keep their list order in
// one last bucket, which keeps neighbouring cells' pairs together (L2 reuse).
__device__ __forceinline__ int sort_bucket(float c, float light_below)
{
    c = sane_cost(c);
    if (c < light_below) return kSortBuckets - 1;
    int b = (int)((kCostTop - c) * (float)kSortPerOctave);
    return b < 0 ? 0 : (b > kSortBuckets - 2 ? kSortBuckets - 2 : b);
}

// total = sum over cell pairs of 2^cost.
__global__ void k_cost_total(int ncp, const float* __restrict__ cost, double* __restrict__ total,
                             const int* __restrict__ ncp_dev = nullptr)
{
    if (ncp_dev) ncp = min(ncp, *ncp_dev);
    double v = 0.0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < ncp; t += gridDim.x * blockDim.x)
        v += exp2((double)sane_cost(cost[t]));
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0 && v != 0.0) atomicAdd(total, v);
}

__device__ __forceinline__ float light_threshold(int ncp, const double* total, float heavy)
{
    return (float)log2(fmax(*total / (double)(ncp > 0 ? ncp : 1) * (double)heavy, 1.0));
}

// Per-block bucket histogram (LDS atomics), written bucket-major:
// counts[b * nblocks + block]. kSortBlock elements per block.
__global__ __launch_bounds__(256) void k_cost_hist(int ncp, const float* __restrict__ cost,
                                                   const double* __restrict__ total, float heavy, int nblocks,
                                                   int* __restrict__ counts, const int* __restrict__ ncp_dev = nullptr)
{
    if (ncp_dev) ncp = min(ncp, *ncp_dev);
    __shared__ int h[kSortBuckets];
    for (int b = threadIdx.x; b < kSortBuckets; b += blockDim.x) h[b] = 0;
    __syncthreads();
    const float lt = light_threshold(ncp, total, heavy);
    for (int k = 0; k < kSortBlock / 256; k++) {
        const int t = blockIdx.x * kSortBlock + k * 256 + threadIdx.x;
        if (t < ncp) atomicAdd(&h[sort_bucket(cost[t], lt)], 1);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kSortBuckets; b += blockDim.x) counts[(int64_t)b * nblocks + blockIdx.x] = h[b];
}

// order[pos] = cell pair, pos = offs[b * nblocks + block] + rank within the block.
__global__ __launch_bounds__(256) void k_cost_place(int ncp, const float* __restrict__ cost,
                                                    const double* __restrict__ total, float heavy, int nblocks,
                                                    const int* __restrict__ offs, int* __restrict__ order,
                                                    const int* __restrict__ ncp_dev = nullptr)
{
    if (ncp_dev) ncp = min(ncp, *ncp_dev);
    __shared__ int cursor[kSortBuckets];
    for (int b = threadIdx.x; b < kSortBuckets; b += blockDim.x)
        cursor[b] = offs[(int64_t)b * nblocks + blockIdx.x];
    __syncthreads();
    const float lt = light_threshold(ncp, total, heavy);
    for (int k = 0; k < kSortBlock / 256; k++) {
        const int t = blockIdx.x * kSortBlock + k * 256 + threadIdx.x;
        if (t < ncp) order[atomicAdd(&cursor[sort_bucket(cost[t], lt)], 1)] = t;
    }
}

// Integer and FP64 fills (graph-capturable, unlike memset nodes).
__global__ void k_fill_i32(int64_t n, int* __restrict__ p, int v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_u64(int64_t n, unsigned long long* __restrict__ p, unsigned long long v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Training samples of the cost model: the features and log2 of the recorded
// ticks of every cell pair that took any time, appended to a bounded store.
__global__ void k_append_samples(int ncp, const float* __restrict__ feats,
                                 const unsigned long long* __restrict__ ticks, int cap, int* __restrict__ count,
                                 float* __restrict__ sx, float* __restrict__ sy)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncp || ticks[t] == 0) return;
    const int k = atomicAdd(count, 1);
    if (k >= cap) return;
    for (int f = 0; f < kNumFeatures; f++) sx[(int64_t)k * kNumFeatures + f] = feats[(int64_t)t * kNumFeatures + f];
    sy[k] = (float)log2((double)ticks[t]);
}

// One row of the per-cell-pair timer, laid out as the reference's
// struct api_cell_timings (utils/defs.h:37-45).
struct TimingRow {
    int64_t N1, N2, time_in_ns;
    int first_cellindex, second_cellindex, tid;
    int pad;
};
static_assert(sizeof(TimingRow) == 40, "TimingRow is 40 bytes");

__global__ void k_timing_rows(int ncp, const CellPairRec* __restrict__ cps, const DevLattice* __restrict__ lats,
                              int batch, const unsigned long long* __restrict__ ticks, const int* __restrict__ cu,
                              double ns_per_tick, TimingRow* __restrict__ rows)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncp) return;
    const CellPairRec cp = cps[t];
    TimingRow r;
    r.N1 = cell_count(lats[batch ? cp.pad : 0].bstart, cp.c1);
    r.N2 = cell_count(lats[batch ? cp.pad : 1].bstart, cp.c2);
    r.time_in_ns = (int64_t)((double)ticks[t] * ns_per_tick);
    r.first_cellindex = cp.c1;
    r.second_cellindex = cp.c2;
    r.tid = cu[t];
    r.pad = 0;
    rows[t] = r;
}

}  // namespace cfamd
