// cf_kernels.hip -- HIP kernels of the MI355X theory pair-counting engine (gfx950).
//
// Device pipeline of one countpairs* call:
//   k_minmax / k_minmax_final     data extent              (reference gridlink_utils.c.src:52-70)
//   k_bucket_count                cell index + z-slab per particle, bucket histogram
//                                                          (gridlink_impl.c.src:162-201)
//   k_scan_*                      exclusive prefix sum of the bucket histogram
//   k_scatter                     counting-sort scatter into SoA arrays in HBM (FP64 + quantized
//                                 cell-local positions for the integer pair filter)
//   k_cell_bounds                 per-cell bounding boxes  (gridlink_impl.c.src:292-316)
//   k_cellpair_flag / _emit       neighbour-cell work list with periodic wraps and
//                                 min-separation culling   (gridlink_impl.c.src:439-625)
//   k_pairs_plain                 FP64 pair kernel, one workgroup per cell pair
//   k_pairs_tuned                 integer filter + FP64 replay, persistent waves, LDS histograms
//
// All FP64 arithmetic that decides a pair's bin is written in the reference's
// operation order; this file is compiled with -ffp-contract=off so the compiler
// never fuses it, and fma() appears only where the reference's AVX512F kernels
// use an FMA.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>
#include "cf_device.h"

namespace cfamd {

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ double wave_min_d(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_max_d(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// ------------------------------------------------------------ extent (a2)
template <typename T>
__global__ void k_minmax(int64_t n, const T* __restrict__ X, const T* __restrict__ Y, const T* __restrict__ Z,
                         double* __restrict__ partial)
{
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double v[3] = {(double)X[i], (double)Y[i], (double)Z[i]};
        for (int d = 0; d < 3; d++) { mn[d] = fmin(mn[d], v[d]); mx[d] = fmax(mx[d], v[d]); }
    }
    __shared__ double s[6][256];
    for (int d = 0; d < 3; d++) { s[d][threadIdx.x] = mn[d]; s[3 + d][threadIdx.x] = mx[d]; }
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int d = 0; d < 3; d++) {
                s[d][threadIdx.x] = fmin(s[d][threadIdx.x], s[d][threadIdx.x + st]);
                s[3 + d][threadIdx.x] = fmax(s[3 + d][threadIdx.x], s[3 + d][threadIdx.x + st]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < 6; d++) partial[blockIdx.x * 6 + d] = s[d][0];
}

// One wave: the lanes reduce strided subsets of the block partials, then the
// wave (min/max are exact, so the order does not matter).
__global__ void k_minmax_final(int nblocks, const double* __restrict__ partial, double* __restrict__ out)
{
    const int lane = (int)threadIdx.x;
    double r[6] = {DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int b = lane; b < nblocks; b += 64)
        for (int d = 0; d < 3; d++) {
            r[d] = fmin(r[d], partial[b * 6 + d]);
            r[3 + d] = fmax(r[3 + d], partial[b * 6 + 3 + d]);
        }
    for (int d = 0; d < 3; d++) {
        r[d] = wave_min_d(r[d]);
        r[3 + d] = wave_max_d(r[3 + d]);
    }
    if (lane == 0)
        for (int d = 0; d < 6; d++) out[d] = r[d];
}

// --------------------------------------------------------- gridding (a5)
struct GridArgs {
    double xmin, ymin, zmin;
    double xinv, yinv, zinv;
    double xbs, ybs, zbs;
    int nx, ny, nz;
    double qinv;  // quanta per unit length of the quantized positions
};

// Quantized cell-local coordinate: rint(v * qinv) clamped to [-kQMax, kQMax].
__device__ __forceinline__ int quantize(double v, double qinv)
{
    const double t = fmin(fmax(v * qinv, -(double)kQMax), (double)kQMax);
    return (int)rint(t);
}
__device__ __forceinline__ uint2 quantize_pos(double dx, double dy, double dz, double qinv)
{
    const int qx = quantize(dx, qinv), qy = quantize(dy, qinv), qz = quantize(dz, qinv);
    return make_uint2(((unsigned)qx & 0xffffu) | ((unsigned)qy << 16), (unsigned)qz);
}

// Cell index exactly as the reference computes it (gridlink_impl.c.src:167-178),
// followed by the z-slab inside the cell.
__device__ __forceinline__ int cell_and_slab(const GridArgs& g, double X, double Y, double Z, int* cell, int* oob)
{
    int ix = (int)((X - g.xmin) * g.xinv);
    int iy = (int)((Y - g.ymin) * g.yinv);
    const double zr = (Z - g.zmin) * g.zinv;
    int iz = (int)zr;
    if (ix > g.nx - 1) ix--;
    if (iy > g.ny - 1) iy--;
    if (iz > g.nz - 1) iz--;
    *oob = ix < 0 || ix >= g.nx || iy < 0 || iy >= g.ny || iz < 0 || iz >= g.nz;
    *cell = (ix * g.ny + iy) * g.nz + iz;
    int s = (int)((zr - (double)iz) * kSlabs);
    s = s < 0 ? 0 : (s > kSlabs - 1 ? kSlabs - 1 : s);
    return s;
}

template <typename T>
__device__ __forceinline__ void bucket_count_one(int64_t i, const T* __restrict__ X, const T* __restrict__ Y,
                                                 const T* __restrict__ Z, const GridArgs& g, int* __restrict__ bucket,
                                                 int* __restrict__ counts, unsigned long long* __restrict__ oob_count)
{
    int c, oob;
    const int s = cell_and_slab(g, (double)X[i], (double)Y[i], (double)Z[i], &c, &oob);
    if (oob) {
        atomicAdd(oob_count, 1ull);
        bucket[i] = -1;
        return;
    }
    const int b = c * kSlabs + s;
    bucket[i] = b;
    atomicAdd(&counts[b], 1);
}

template <typename T>
__global__ void k_bucket_count(int64_t n, const T* __restrict__ X, const T* __restrict__ Y, const T* __restrict__ Z,
                               GridArgs g, int* __restrict__ bucket, int* __restrict__ counts,
                               unsigned long long* __restrict__ oob_count)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) bucket_count_one(i, X, Y, Z, g, bucket, counts, oob_count);
}

// Exclusive scan over int32 counts, three phases, 1024 items per block.
constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;

__device__ int block_exclusive_scan(int v, int* total)
{
    __shared__ int ws[kScanBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int incl = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) ws[w] = incl;
    __syncthreads();
    int off = 0, tot = 0;
    for (int k = 0; k < kScanBlock / 64; k++) {
        if (k < w) off += ws[k];
        tot += ws[k];
    }
    __syncthreads();
    *total = tot;
    return off + incl - v;
}

__global__ void k_scan_blocks(int64_t n, const int* __restrict__ in, int* __restrict__ out, int* __restrict__ block_sums)
{
    const int64_t base = (int64_t)blockIdx.x * kScanBlock * kScanItems + (int64_t)threadIdx.x * kScanItems;
    int v[kScanItems], s = 0;
    for (int k = 0; k < kScanItems; k++) {
        v[k] = base + k < n ? in[base + k] : 0;
        s += v[k];
    }
    int tot;
    int ex = block_exclusive_scan(s, &tot);
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
    if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

__global__ void k_scan_add(int64_t n, int* __restrict__ out, const int* __restrict__ block_offsets)
{
    const int64_t base = (int64_t)blockIdx.x * kScanBlock * kScanItems + (int64_t)threadIdx.x * kScanItems;
    const int add = block_offsets[blockIdx.x];
    for (int k = 0; k < kScanItems; k++)
        if (base + k < n) out[base + k] += add;
}

template <typename T>
__device__ __forceinline__ void scatter_one(int64_t i, const T* __restrict__ X, const T* __restrict__ Y,
                                            const T* __restrict__ Z, const T* __restrict__ W, const GridArgs& g,
                                            const int* __restrict__ bucket, const int* __restrict__ bstart,
                                            int* __restrict__ fill, double* __restrict__ x, double* __restrict__ y,
                                            double* __restrict__ z, double* __restrict__ w, double4* __restrict__ d4,
                                            uint2* __restrict__ q)
{
    const int b = bucket[i];
    if (b < 0) return;  // outside the grid: counted by k_bucket_count, reported by the host
    const int pos = bstart[b] + atomicAdd(&fill[b], 1);
    const double xv = (double)X[i], yv = (double)Y[i], zv = (double)Z[i];
    x[pos] = xv;
    y[pos] = yv;
    z[pos] = zv;
    const double wv = W ? (double)W[i] : 0.0;
    if (W) w[pos] = wv;
    d4[pos] = make_double4(xv, yv, zv, wv);
    const int c = b / kSlabs;
    const int iz = c % g.nz, ixy = c / g.nz, iy = ixy % g.ny, ix = ixy / g.ny;
    const double ox = g.xmin + ix * g.xbs, oy = g.ymin + iy * g.ybs, oz = g.zmin + iz * g.zbs;
    q[pos] = quantize_pos(xv - ox, yv - oy, zv - oz, g.qinv);
}

template <typename T>
__global__ void k_scatter(int64_t n, const T* __restrict__ X, const T* __restrict__ Y, const T* __restrict__ Z,
                          const T* __restrict__ W, GridArgs g, const int* __restrict__ bucket,
                          const int* __restrict__ bstart, int* __restrict__ fill,
                          double* __restrict__ x, double* __restrict__ y, double* __restrict__ z,
                          double* __restrict__ w, double4* __restrict__ d4, uint2* __restrict__ q)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) scatter_one(i, X, Y, Z, W, g, bucket, bstart, fill, x, y, z, w, d4, q);
}

// Per-cell bounding boxes, one wave per cell (lanes stride over the cell's
// particles, then a shuffle reduction; min/max are exact, so the order of the
// reduction does not matter). Empty cells get (+DBL_MAX, -DBL_MAX).
__device__ __forceinline__ void cell_bounds_one(int64_t c, const int* __restrict__ bstart,
                                                const double* __restrict__ x, const double* __restrict__ y,
                                                const double* __restrict__ z, double* __restrict__ bounds)
{
    const int lane = threadIdx.x & 63;
    double b[6] = {DBL_MAX, -DBL_MAX, DBL_MAX, -DBL_MAX, DBL_MAX, -DBL_MAX};
    const int e = bstart[(c + 1) * kSlabs];
    for (int k = bstart[c * kSlabs] + lane; k < e; k += 64) {
        b[0] = x[k] < b[0] ? x[k] : b[0];
        b[1] = x[k] > b[1] ? x[k] : b[1];
        b[2] = y[k] < b[2] ? y[k] : b[2];
        b[3] = y[k] > b[3] ? y[k] : b[3];
        b[4] = z[k] < b[4] ? z[k] : b[4];
        b[5] = z[k] > b[5] ? z[k] : b[5];
    }
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int d = 0; d < 6; d += 2) {
            const double lo = __shfl_xor(b[d], o, 64), hi = __shfl_xor(b[d + 1], o, 64);
            b[d] = lo < b[d] ? lo : b[d];
            b[d + 1] = hi > b[d + 1] ? hi : b[d + 1];
        }
    }
    if (lane == 0)
        for (int d = 0; d < 6; d++) bounds[c * 6 + d] = b[d];
}

__global__ void k_cell_bounds(int64_t ncells, const int* __restrict__ bstart, const double* __restrict__ x,
                              const double* __restrict__ y, const double* __restrict__ z, double* __restrict__ bounds)
{
    const int64_t c = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    if (c < ncells) cell_bounds_one(c, bstart, x, y, z, bounds);  // uniform per wave
}

__global__ void k_weight_sums(int64_t n, const double* __restrict__ w, double* __restrict__ out)
{
    double s = 0, s2 = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        s += w[i];
        s2 += w[i] * w[i];
    }
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        s2 += __shfl_xor(s2, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&out[0], s);
        atomicAdd(&out[1], s2);
    }
}

// Device-to-device copy: n16 16-byte words through the vector path (grid-stride),
// then the byte tail. Reads and writes exactly the bytes asked for, which also
// makes it the reference kernel for scaling the HBM byte counters.
__global__ void k_copy16(int64_t n16, const uint4* __restrict__ src, uint4* __restrict__ dst, int tail,
                         const unsigned char* __restrict__ src_tail, unsigned char* __restrict__ dst_tail)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
    if (blockIdx.x == 0 && (int)threadIdx.x < tail) dst_tail[threadIdx.x] = src_tail[threadIdx.x];
}

// ------------------------------------------------------- cell pairs (a6)
struct NgbArgs {
    int nx, ny, nz;
    int rx, ry, rz;
    int px, py, pz;
    int autocorr, min_sep_opt, dupcheck;
    double xwrap, ywrap, zwrap;
    double max3, max2, max1;
};

__device__ __forceinline__ int64_t cell_count(const int* bstart, int64_t c)
{
    return (int64_t)bstart[(c + 1) * kSlabs] - bstart[c * kSlabs];
}

// Evaluates neighbour k of primary cell icell in the reference's loop order
// (iix outer, iiy, iiz inner). Returns 1 and fills rec when the reference would
// append this cell pair (before its duplicate check).
__device__ int eval_neighbour(const NgbArgs& a, int64_t icell, int k, const int* __restrict__ bs1,
                              const int* __restrict__ bs2, const double* __restrict__ bd1,
                              const double* __restrict__ bd2, CellPairRec* rec)
{
    const int ny_n = 2 * a.ry + 1, nz_n = 2 * a.rz + 1;
    const int iix = k / (ny_n * nz_n) - a.rx;
    const int iiy = (k / nz_n) % ny_n - a.ry;
    const int iiz = k % nz_n - a.rz;
    const int iz = (int)(icell % a.nz);
    const int ix = (int)(icell / ((int64_t)a.ny * a.nz));
    const int iy = (int)((icell - iz - (int64_t)ix * a.nz * a.ny) / a.nz);
    const int iiix = a.px ? (ix + iix + a.nx) % a.nx : ix + iix;
    if (iiix < 0 || iiix >= a.nx) return 0;
    const int iiiy = a.py ? (iy + iiy + a.ny) % a.ny : iy + iiy;
    if (iiiy < 0 || iiiy >= a.ny) return 0;
    const int iiiz = a.pz ? (iz + iiz + a.nz) % a.nz : iz + iiz;
    if (iiiz < 0 || iiiz >= a.nz) return 0;
    const int wx = (ix + iix >= 0 && ix + iix < a.nx) ? 0 : (ix + iix < 0 ? 1 : -1);
    const int wy = (iy + iiy >= 0 && iy + iiy < a.ny) ? 0 : (iy + iiy < 0 ? 1 : -1);
    const int wz = (iz + iiz >= 0 && iz + iiz < a.nz) ? 0 : (iz + iiz < 0 ? 1 : -1);
    const int64_t icell2 = iiiz + (int64_t)a.nz * iiiy + (int64_t)a.nz * a.ny * iiix;
    if ((a.autocorr == 1 && icell2 > icell) || cell_count(bs2, icell2) == 0) return 0;
    const double offx = wx * a.xwrap, offy = wy * a.ywrap, offz = wz * a.zwrap;
    double cx = 0, cy = 0, cz = 0, mdx = 0, mdy = 0, mdz = 0;
    if (a.min_sep_opt) {
        const double* f = bd1 + icell * 6;
        const double* s = bd2 + icell2 * 6;
        const double xl = f[0] + offx, xh = f[1] + offx;
        const double yl = f[2] + offy, yh = f[3] + offy;
        const double zl = f[4] + offz, zh = f[5] + offz;
        cx = iix < 0 ? xl : (iix > 0 ? xh : 0.0);
        cy = iiy < 0 ? yl : (iiy > 0 ? yh : 0.0);
        cz = iiz < 0 ? zl : (iiz > 0 ? zh : 0.0);
        mdx = iix != 0 ? ((iix < 0 ? xl : xh) - (iix < 0 ? s[1] : s[0])) : 0.0;
        mdy = iiy != 0 ? ((iiy < 0 ? yl : yh) - (iiy < 0 ? s[3] : s[2])) : 0.0;
        mdz = iiz != 0 ? ((iiz < 0 ? zl : zh) - (iiz < 0 ? s[5] : s[4])) : 0.0;
        if (a.max3 > 0) {
            const double t = mdx * mdx + mdy * mdy + mdz * mdz;
            if (t >= a.max3 * a.max3) return 0;
        }
        if (a.max2 > 0) {
            const double t = mdx * mdx + mdy * mdy;
            if (t >= a.max2 * a.max2) return 0;
        }
        if (a.max1 > 0 && iiz != 0) {
            if (mdz * mdz >= a.max1 * a.max1) return 0;
        }
    }
    rec->c1 = (int)icell;
    rec->c2 = (int)icell2;
    rec->wx = (signed char)wx;
    rec->wy = (signed char)wy;
    rec->wz = (signed char)wz;
    rec->same = (a.autocorr == 1 && icell2 == icell) ? 1 : 0;
    rec->pad = 0;
    rec->mdx = mdx; rec->mdy = mdy; rec->mdz = mdz;
    rec->cx = cx; rec->cy = cy; rec->cz = cz;
    return 1;
}

__device__ int neighbour_added(const NgbArgs& a, int64_t icell, int k, const int* bs1, const int* bs2,
                               const double* bd1, const double* bd2, CellPairRec* rec)
{
    if (cell_count(bs1, icell) == 0) return 0;
    if (!eval_neighbour(a, icell, k, bs1, bs2, bd1, bd2, rec)) return 0;
    if (a.dupcheck) {
        // the reference skips a neighbour already appended for this primary cell with
        // the same secondary cell and the same wrap (gridlink_utils.h.src:46-72).
        // prev is initialized: eval_neighbour writes it only when it returns 1, and a
        // compiler may evaluate the comparisons below on an unwritten record and
        // combine them with the result (observed: neighbours dropped depending on
        // leftover register contents)
        CellPairRec prev = {};
        for (int kk = 0; kk < k; kk++) {
            const int added = eval_neighbour(a, icell, kk, bs1, bs2, bd1, bd2, &prev);
            if (added && prev.c2 == rec->c2 && prev.wx == rec->wx && prev.wy == rec->wy && prev.wz == rec->wz)
                return 0;
        }
    }
    return 1;
}

__global__ void k_cellpair_flag(NgbArgs a, int64_t ncells, int nngb, const int* __restrict__ bs1,
                                const int* __restrict__ bs2, const double* __restrict__ bd1,
                                const double* __restrict__ bd2, int* __restrict__ flags)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= ncells * nngb) return;
    CellPairRec r = {};
    flags[t] = neighbour_added(a, t / nngb, (int)(t % nngb), bs1, bs2, bd1, bd2, &r);
}

// Grid-stride over the candidates; the pairs tested are summed per workgroup
// (integers below 2^53, exact in any order) and added with one atomic per
// workgroup, as every atomic on the single counter is serialized.
__global__ void k_cellpair_emit(NgbArgs a, int64_t ncells, int nngb, const int* __restrict__ bs1,
                                const int* __restrict__ bs2, const double* __restrict__ bd1,
                                const double* __restrict__ bd2, const int* __restrict__ pos,
                                CellPairRec* __restrict__ out, double* __restrict__ tested, int cat)
{
    __shared__ double part[4];
    const int64_t total = ncells * nngb;
    double my = 0;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        CellPairRec r = {};
        if (neighbour_added(a, t / nngb, (int)(t % nngb), bs1, bs2, bd1, bd2, &r)) {
            r.pad = cat;
            out[pos[t]] = r;
            const double n1 = (double)cell_count(bs1, r.c1), n2 = (double)cell_count(bs2, r.c2);
            my += r.same ? n1 * (n1 - 1) * 0.5 : n1 * n2;
        }
    }
    for (int o = 32; o > 0; o >>= 1) my += __shfl_xor(my, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = my;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double s = part[0] + part[1] + part[2] + part[3];
        if (s != 0) atomicAdd(tested, s);
    }
}

// ------------------------------------------------- exact pair replay (a8-a12)
__device__ __forceinline__ double r2_3d(int f, double dx, double dy, double dz)
{
    switch (f) {
    case R2_FMA_DZ_DY_DX: return fma(dz, dz, fma(dy, dy, dx * dx));
    case R2_X_SUM_YZ: return dx * dx + (dy * dy + dz * dz);
    case R2_FMA_DX_DY_DZ: return fma(dx, dx, fma(dy, dy, dz * dz));
    default: return (dx * dx + dy * dy) + dz * dz;
    }
}

__device__ __forceinline__ double rp2_f(int f, double dx, double dy)
{
    return f == RP_FMA ? fma(dy, dy, dx * dx) : dx * dx + dy * dy;
}

// first k from the top with r2 >= rupp_sqr[k-1] (countpairs_kernels.c.src:982-993)
__device__ __forceinline__ int find_kbin(const double* __restrict__ e, int nbin, int sorted, double r2)
{
    if (sorted) {
        int lo = 0, hi = nbin - 1;  // count of e[0..nbin-2] <= r2
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (e[mid] <= r2) lo = mid + 1; else hi = mid;
        }
        return lo;
    }
    for (int k = nbin - 1; k >= 1; k--)
        if (r2 >= e[k - 1]) return k;
    return 0;
}

// On-sky pair (mocks/DDrppi_mocks/countpairs_rp_pi_mocks_kernels.c.src:1240-1290,
// mocks/DDsmu_mocks/countpairs_s_mu_mocks_kernels.c.src fallback): s = x_j - x_i,
// l = x_i + x_j; rp-pi: pi^2 = (s.l)^2 / |l|^2, rp^2 = s^2 - pi^2; s-mu:
// mu^2 = (s.l)^2 / (|l|^2 s^2). Returns 0 for a rejected pair, else the squared
// separation that selects the bin (*key: rp^2 or s^2) and the pi / mu column
// (*j2; n2 where pi^2 >= pimax^2 or mu^2 >= mu_max^2, as the reference).
__device__ __forceinline__ int los_split(const PairParams& P, int mode, double xi, double yi, double zi, double xj,
                                         double yj, double zj, double* key, int* j2)
{
    const double parx = xi + xj, pary = yi + yj, parz = zi + zj;
    const double perpx = xj - xi, perpy = yj - yi, perpz = zj - zi;
    const double sqr_s = perpx * perpx + perpy * perpy + perpz * perpz;
    const double s_dot_l = parx * perpx + pary * perpy + parz * perpz;
    const double sqr_l = parx * parx + pary * pary + parz * parz;
    const double sqr_s_dot_l = s_dot_l * s_dot_l;
    const double sqr_lim = P.sqr_mu_max;
    if (mode == MODE_RPPI_MOCKS) {
        if (sqr_s >= P.sqr_rmax) return 0;
        if (sqr_s_dot_l >= sqr_l * sqr_lim) return 0;
        const double sqr_par = sqr_s_dot_l / sqr_l;
        *j2 = sqr_par >= sqr_lim ? P.n2 : (int)(sqrt(sqr_par) * P.inv_d2);
        const double sqr_perp = sqr_s - sqr_par;
        if (sqr_perp >= P.sqr_rpmax || sqr_perp < P.sqr_rmin) return 0;
        *key = sqr_perp;
        return 1;
    }
    if (sqr_s >= P.sqr_rmax || sqr_s < P.sqr_rmin) return 0;
    const double sqr_mu = sqr_s_dot_l / (sqr_l * sqr_s);
    *j2 = sqr_mu >= sqr_lim ? P.n2 : (int)(sqrt(sqr_mu) * P.inv_d2);
    *key = sqr_s;
    return 1;
}

struct ITerms {
    double xpos, ypos, zpos;
    double max_dz, target;
    int skip;
};

// The reference kernels' per-particle prologue (countpairs_kernels.c.src:900-950 and
// the rp_pi / s_mu / wp / xi counterparts): position + wrap, minimum separation of
// this particle to the secondary cell, z window.
__device__ __forceinline__ void i_terms(const PairParams& P, int mode, const CellPairRec& cp, double xi, double yi,
                                        double zi, ITerms* t)
{
    t->xpos = xi + cp.wx * P.xwrap;
    t->ypos = yi + cp.wy * P.ywrap;
    t->zpos = zi + cp.wz * P.zwrap;
    if (cp.same) {
        t->skip = 0;
        t->max_dz = (mode == MODE_RPPI || mode == MODE_WP) ? P.pimax : P.max_all_dz_same;
        t->target = t->zpos - t->max_dz;
        return;
    }
    const double mdx = cp.mdx > 0 ? cp.mdx + fabs(t->xpos - cp.cx) : cp.mdx;
    const double mdy = cp.mdy > 0 ? cp.mdy + fabs(t->ypos - cp.cy) : cp.mdy;
    const double mdz = cp.mdz > 0 ? cp.mdz + fabs(t->zpos - cp.cz) : cp.mdz;
    if (mode == MODE_RPPI || mode == MODE_WP) {
        t->skip = (mdx * mdx + mdy * mdy) >= P.sqr_rmax || mdz >= P.pimax;
        t->max_dz = P.pimax;
        t->target = t->zpos - P.pimax;
    } else if (mode == MODE_SMU) {
        t->skip = ((mdx * mdx + mdy * mdy) + mdz * mdz) >= P.sqr_rmax || mdz >= P.pimax;
        const double sq = (P.sqr_rmax - mdx * mdx) - mdy * mdy;
        t->max_dz = sq < P.pimax * P.pimax ? sqrt(sq) : P.pimax;
        t->target = t->zpos - t->max_dz;
    } else {
        t->skip = ((mdx * mdx + mdy * mdy) + mdz * mdz) >= P.sqr_rmax;
        t->max_dz = sqrt((P.sqr_rmax - mdx * mdx) - mdy * mdy);
        t->target = t->zpos - t->max_dz;
    }
}

// Decides one (i, j) pair exactly as the reference does. Returns the histogram
// slot or -1; *sep receives sqrt(r^2) (or sqrt(s^2)).
// `window` = 0 skips the z-window / per-particle tests, which the caller may do
// only when it has shown the pair lies outside the edge windows (see k_pairs_tuned).
__device__ __forceinline__ int replay_pair(const PairParams& P, int mode, const ITerms& t, int same, int window,
                                           double xj, double yj, double zj, const double* __restrict__ e,
                                           double* sep)
{
    const double dx = xj - t.xpos, dy = yj - t.ypos, dz = zj - t.zpos;
    if (mode == MODE_RPPI || mode == MODE_WP) {
        if (window && ((!same && (t.skip || zj <= t.target)))) return -1;
        if (dz <= -P.pimax || dz >= P.pimax) return -1;
        const double adz = fabs(dz);
        const double r2 = rp2_f(P.r2form, dx, dy);
        if (r2 >= P.sqr_rmax || r2 < P.sqr_rmin) return -1;
        const int k = find_kbin(e, P.nbin, P.sorted_edges, r2);
        if (k == 0) return -1;
        *sep = sqrt(r2);
        if (mode == MODE_WP) return k;
        if (P.float_index) return (int)((double)k * (double)(P.n2 + 1) + adz * P.inv_d2);
        int pb = (int)(adz * P.inv_d2);
        pb = pb > P.n2 ? P.n2 : pb;
        return k * (P.n2 + 1) + pb;
    }
    if (window) {
        if (t.skip) return -1;
        if (zj <= t.target) return -1;
        if (dz >= t.max_dz) return -1;
    }
    if (mode == MODE_RPPI_MOCKS || mode == MODE_SMU_MOCKS) {
        double key;
        int j2;
        if (!los_split(P, mode, t.xpos, t.ypos, t.zpos, xj, yj, zj, &key, &j2)) return -1;
        const int k = find_kbin(e, P.nbin, P.sorted_edges, key);
        if (k == 0) return -1;
        *sep = sqrt(key);
        return k * (P.n2 + 1) + j2;
    }
    if (mode == MODE_SMU) {
        const double sqr_dz = dz * dz;
        const double s2 = P.r2form == R2_FMA_DX_DY_DZ ? fma(dx, dx, fma(dy, dy, sqr_dz)) : (dx * dx + dy * dy) + sqr_dz;
        if (s2 >= P.sqr_rmax || s2 < P.sqr_rmin) return -1;
        if (sqr_dz >= s2 * P.sqr_mu_max) return -1;
        const double mu = sqrt(sqr_dz / s2);
        const int k = find_kbin(e, P.nbin, P.sorted_edges, s2);
        if (k == 0) return -1;
        *sep = sqrt(s2);
        if (P.float_index) return (int)((double)k * (double)(P.n2 + 1) + mu * P.inv_d2);
        int mb = (int)(mu * P.inv_d2);
        mb = mb > P.n2 ? P.n2 : mb;
        return k * (P.n2 + 1) + mb;
    }
    const double r2 = r2_3d(P.r2form, dx, dy, dz);
    if (r2 >= P.sqr_rmax || r2 < P.sqr_rmin) return -1;
    const int k = find_kbin(e, P.nbin, P.sorted_edges, r2);
    if (k == 0) return -1;
    *sep = sqrt(r2);
    return k;
}

// Histogram of one workgroup in LDS: counts (u64), sum of separations, sum of weights.
struct LdsHist {
    unsigned long long* np;
    double* avg;
    double* w;
};

__device__ __forceinline__ void hist_add(const PairParams& P, LdsHist& h, int slot, double sep, double wprod)
{
    atomicAdd(&h.np[slot], 1ull);
    if (P.need_avg) atomicAdd(&h.avg[slot], sep);
    if (P.need_w) atomicAdd(&h.w[slot], wprod);
}

__device__ void hist_flush(const PairParams& P, LdsHist& h, unsigned long long* g_np, double* g_avg, double* g_w)
{
    for (int s = threadIdx.x; s < P.nslots; s += blockDim.x) {
        const unsigned long long c = h.np[s];
        if (c) {
            atomicAdd(&g_np[s], c);
            if (P.need_avg) atomicAdd(&g_avg[s], h.avg[s]);
            if (P.need_w) atomicAdd(&g_w[s], h.w[s]);
        }
    }
}

__device__ void hist_init(const PairParams& P, LdsHist& h)
{
    for (int s = threadIdx.x; s < P.nslots; s += blockDim.x) {
        h.np[s] = 0;
        h.avg[s] = 0;
        h.w[s] = 0;
    }
}

// -------------------------------------------------- plain FP64 pair kernel
// One workgroup per cell pair; every (i, j) of the pair is decided in FP64 with
// the full reference predicate. This is the "plain GPU form". Its histogram is
// FP64 in LDS (GH = false), or the global histogram itself (GH = true) when
// the slots do not fit the workgroup's LDS.
template <bool GH>
__global__ __launch_bounds__(128) void k_pairs_plain(PairParams P, DevLattice L1, DevLattice L2,
                                                     const CellPairRec* __restrict__ cps,
                                                     const double* __restrict__ rupp_sqr,
                                                     unsigned long long* __restrict__ g_np,
                                                     double* __restrict__ g_avg, double* __restrict__ g_w)
{
    extern __shared__ unsigned long long smem_u64[];
    LdsHist h;
    const double* e;
    if (GH) {
        // no LDS: histogram and edges in global memory
        h.np = g_np;
        h.avg = g_avg;
        h.w = g_w;
        e = rupp_sqr;
    } else {
        h.np = smem_u64;
        h.avg = reinterpret_cast<double*>(smem_u64 + P.nslots);
        h.w = h.avg + P.nslots;
        double* el = h.w + P.nslots;
        hist_init(P, h);
        for (int s = threadIdx.x; s < P.nbin; s += blockDim.x) el[s] = rupp_sqr[s];
        e = el;
        __syncthreads();
    }
    const CellPairRec cp = cps[blockIdx.x];
    const int mode = P.mode;
    const int i0 = L1.bstart[(int64_t)cp.c1 * kSlabs], i1 = L1.bstart[(int64_t)(cp.c1 + 1) * kSlabs];
    const int j0 = L2.bstart[(int64_t)cp.c2 * kSlabs], j1 = L2.bstart[(int64_t)(cp.c2 + 1) * kSlabs];
    for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        ITerms t = {};
        i_terms(P, mode, cp, L1.x[i], L1.y[i], L1.z[i], &t);
        if (t.skip) continue;
        const double wi = P.need_w ? L1.w[i] : 0.0;
        const int jstart = cp.same ? i + 1 : j0;
        for (int j = jstart; j < j1; j++) {
            double sep = 0;
            const int slot = replay_pair(P, mode, t, cp.same, 1, L2.x[j], L2.y[j], L2.z[j], e, &sep);
            if (slot >= 0) hist_add(P, h, slot, sep, P.need_w ? wi * L2.w[j] : 0.0);
        }
    }
    if (GH) return;
    __syncthreads();
    hist_flush(P, h, g_np, g_avg, g_w);
}
template __global__ void k_pairs_plain<false>(PairParams, DevLattice, DevLattice, const CellPairRec*, const double*,
                                              unsigned long long*, double*, double*);
template __global__ void k_pairs_plain<true>(PairParams, DevLattice, DevLattice, const CellPairRec*, const double*,
                                             unsigned long long*, double*, double*);

// ------------------------------------------------- tuned pair kernel
// Work item: (cell pair, chunk of <= 64 consecutive primary particles).
// Each wave pulls work items from a device counter until the list is exhausted.
// Lane l holds primary particle i0+l. The secondary particles of the cell pair,
// restricted to the z-slabs the chunk can reach, are read with scalar loads: the
// index is wave-uniform, so a secondary lands in scalar registers that every
// lane's vector instructions read directly (no per-lane load, no broadcast
// instruction). Two paths per work item:
//  * dense (every pair of the cell pair is inside the separation range): each
//    (i, j) is decided in FP64 in lockstep from a 32-byte (x, y, z, w) record;
//  * sparse: an integer test on quantized cell-local positions (packed 16-bit
//    differences, one dot-product instruction per pair), with thresholds that
//    provably pass every pair the FP64 test can accept, rejects the pairs that
//    cannot be in range; the survivors are compacted into a per-wave LDS queue
//    and decided 64 at a time in FP64.
// Either way every counted pair is decided by the reference's FP64 arithmetic.
constexpr int kQueues = 8;          // work sub-queues (one counter each)
constexpr int kCounterStride = 32;  // ints between sub-queue counters (128 B)
constexpr int kBlockJ = 4;          // secondaries per scalar load of the sparse path
// Survivor buffer per wave: entries (j << 6) | lane at [head, tail). Drains take
// 64 * DP entries from the head (DP survivors per lane: 1, or 2 with both pairs'
// records loaded before either is decided); a scan runs while fewer than that
// wait, a block adds <= 64 kBlockJ, and lanes without a survivor write up to 127
// slots past the tail (see enqueue); once head >= kCompactAt the remaining
// entries move to the front. At a block's start head <= 64 (DP = 1; head
// advances by 64 and is reset at 128) or head = 0 (DP = 2: every drain round
// ends with the move) and tail - head < 64 DP, so tail <= 127 and the
// furthest write is 127 + 64 (kBlockJ - 1) + 127 < kQueueAlloc.
constexpr int kCompactAt = 128;
constexpr int kQueueAlloc = kCompactAt + 64 + 64 * (kBlockJ - 1) + 127;
// Queue entries hold secondary indices relative to the start of the window
// being scanned, ((j - jbase) << 6) | lane in 32 bits: a secondary window longer
// than P.max_window (2^25 by default, a multiple of kBlockJ) is scanned in
// consecutive sub-windows of that length, with the queue drained at the end of
// each.

// Read-only particle data seen through the constant address space: loads with a
// wave-uniform index become scalar loads.
typedef const __attribute__((address_space(4))) double KDouble;
typedef const __attribute__((address_space(4))) unsigned KUint;
__device__ __forceinline__ double4 kload_d4(KDouble* p, int j)
{
    return make_double4(p[4 * (int64_t)j], p[4 * (int64_t)j + 1], p[4 * (int64_t)j + 2], p[4 * (int64_t)j + 3]);
}
__device__ __forceinline__ uint2 kload_u2(KUint* p, int j)
{
    return make_uint2(p[2 * (int64_t)j], p[2 * (int64_t)j + 1]);
}
typedef short short2v __attribute__((ext_vector_type(2)));

// sqrt(r2) to ~1e-14 relative: the hardware FP32 square root (about 1 ulp)
// refined by one Newton step in FP64.
__device__ __forceinline__ double fast_sqrt(double r2)
{
    const float f = (float)r2;
    if (!(f > 1e-30f)) return sqrt(r2);
    const float s = __builtin_amdgcn_sqrtf(f);
    const double sd = (double)s;
    const double res = fma(-sd, sd, r2);
    const double half_inv = 0.5 * (double)__builtin_amdgcn_rcpf(s);
    return fma(res, half_inv, sd);
}

// Histogram of one workgroup of the tuned kernel, replicated R ways with layout
// [slot][R]: lanes of one wave that land in the same slot add to different
// copies, which sit in different banks. Counts u64, sums FP64 (native
// ds_add_u64 / ds_add_f64: the LDS atomics issue on the LDS pipe, not on the
// VALU, which bounds this kernel). GH kernels (histograms larger than the
// workgroup's LDS) add to the global histogram instead.
struct KernelLds {
    unsigned long long* np;
    double* avg;
    double* w;
    int rmask;                 // R - 1
    int lane_copy;             // lane & (R - 1)
    const double* e;           // rupp_sqr[0..nbin-1]
    const unsigned short* lut; // bin lookup
    unsigned long long* gnp;   // the global histogram (GH)
    double* gavg;
    double* gw;
};

// Per-launch flags of the tuned kernel as compile-time constants: FL >= 0 is the
// mask of the launch's separation sums (kFlAvg), weight sums (kFlW) and
// single-step bin lookup (kFlLut1: sorted edges, at most one edge per lookup
// bucket); FL = kFlRuntime reads them from P (the global-histogram kernels).
// Read from P, every such test in the per-pair code is a uniform branch on a
// scalar lane mask held across the persistent loop, and the masks crowd the
// scalar registers into spills.
constexpr int kFlAvg = 1, kFlW = 2, kFlLut1 = 4, kFlRuntime = -1;
template <int FL>
__device__ __forceinline__ bool fl_avg(const PairParams& P)
{
    return FL < 0 ? P.need_avg != 0 : (FL & kFlAvg) != 0;
}
template <int FL>
__device__ __forceinline__ bool fl_w(const PairParams& P)
{
    return FL < 0 ? P.need_w != 0 : (FL & kFlW) != 0;
}
template <int FL>
__device__ __forceinline__ bool fl_lut1(const PairParams& P)
{
    return FL < 0 ? (P.sorted_edges != 0 && P.lut_single != 0) : (FL & kFlLut1) != 0;
}
// The reference's AVX512F flattened (rp, pi) / (s, mu) index rule: used exactly
// when the rounding order replayed is the AVX512F kernel's (the host sets
// P.float_index by the same rule).
template <int MODE, int R2F>
__device__ __forceinline__ constexpr bool float_index_of()
{
    return (MODE == MODE_RPPI && R2F == RP_FMA) || (MODE == MODE_SMU && R2F == R2_FMA_DX_DY_DZ);
}

// Adds a counted pair (slot, squared separation r2, weight product) to the histogram.
template <bool GH, int FL>
__device__ __forceinline__ void hist_pair(const PairParams& P, KernelLds& S, int slot, double r2, double wprod)
{
    if (GH) {
        atomicAdd(&S.gnp[slot], 1ull);
        if (fl_avg<FL>(P)) atomicAdd(&S.gavg[slot], fast_sqrt(r2));
        if (fl_w<FL>(P)) atomicAdd(&S.gw[slot], wprod);
        return;
    }
    const int a = slot * (S.rmask + 1) + S.lane_copy;
    atomicAdd(&S.np[a], 1ull);
    if (fl_avg<FL>(P)) atomicAdd(&S.avg[a], fast_sqrt(r2));
    if (fl_w<FL>(P)) atomicAdd(&S.w[a], wprod);
}

// Number of edges e[0..nbin-2] <= r2 (the reference's top-down bin search,
// countpairs_kernels.c.src:982-993, for non-decreasing edges): start from the
// lookup table's lower bound for the 1/16-octave bucket of (float)r2 and step up.
template <int FL>
__device__ __forceinline__ int lut_kbin(const PairParams& P, const KernelLds& S, double r2)
{
    if (!fl_lut1<FL>(P) && !P.sorted_edges) return find_kbin(S.e, P.nbin, 0, r2);
    int idx = (int)(__float_as_uint((float)r2) >> 19) - P.lut_base;
    idx = idx < 0 ? 0 : (idx >= P.lut_n ? P.lut_n - 1 : idx);
    int k = S.lut[idx];
    if (fl_lut1<FL>(P)) return k + (S.e[k] <= r2 ? 1 : 0);  // e[nbin-1] = rmax^2 > r2
    while (k < P.nbin - 1 && S.e[k] <= r2) k++;
    return k;
}

// FP64 separation in the rounding order R2F of the reference kernel being replayed.
template <int R2F>
__device__ __forceinline__ double sep2(double dx, double dy, double dz)
{
    if (R2F == R2_FMA_DZ_DY_DX) return fma(dz, dz, fma(dy, dy, dx * dx));
    if (R2F == R2_X_SUM_YZ) return dx * dx + (dy * dy + dz * dz);
    if (R2F == R2_FMA_DX_DY_DZ) return fma(dx, dx, fma(dy, dy, dz * dz));
    if (R2F == RP_FMA) return fma(dy, dy, dx * dx);
    if (R2F == RP_SUM) return dx * dx + dy * dy;
    return (dx * dx + dy * dy) + dz * dz;
}

// The same pointer, opaque to the compiler: loads through it are issued where
// they are written (not hoisted, not merged with earlier ones), which keeps
// values that only rare paths need out of the registers of the hot loops.
template <typename T>
__device__ __forceinline__ const T* fresh(const T* p)
{
    asm volatile("" : "+s"(p));
    return p;
}
// The same for read-only data seen through the constant address space, so the
// loads are scalar.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* fresh_k(const T* p)
{
    asm volatile("" : "+s"(p));
    return (const __attribute__((address_space(4))) T*)(const void*)p;
}

// Decides one (i, j) pair in FP64 exactly as the reference does and adds it to
// the LDS histogram. (xpos, ypos, zpos) = primary position plus the cell pair's
// wrap; (xj, yj, zj) = secondary position. `ok` masks lanes without a pair.
template <int MODE, int R2F, bool GH, int FL>
__device__ __forceinline__ void exact_pair(const PairParams& P, const CellPairRec* cpp, const DevLattice* l1p, int i,
                                           int ok, double xpos, double ypos, double zpos, double wi, double xj,
                                           double yj, double zj, double wj, KernelLds& S)
{
    constexpr bool is2d = MODE == MODE_RPPI || MODE == MODE_WP;
    constexpr bool onsky = MODE == MODE_RPPI_MOCKS || MODE == MODE_SMU_MOCKS;
    const double dx = xj - xpos, dy = yj - ypos, dz = zj - zpos;
    if (onsky) {
        double key;
        int j2;
        if (!ok || !los_split(P, MODE, xpos, ypos, zpos, xj, yj, zj, &key, &j2)) return;
        // pruning replay for pairs near the 3-D limit (the 3-D kernels' rules)
        if (sep2<R2F>(dx, dy, dz) >= P.sqr_rmax - P.near_r2) {
            const CellPairRec* c = fresh(cpp);
            const DevLattice* l1 = fresh(l1p);
            ITerms t = {};
            i_terms(P, MODE, *c, l1->x[i], l1->y[i], l1->z[i], &t);
            if (t.skip || zj <= t.target || dz >= t.max_dz) return;
        }
        const int k = lut_kbin<FL>(P, S, key);
        if (k == 0) return;
        hist_pair<GH, FL>(P, S, k * (P.n2 + 1) + j2, key, wi * wj);
        return;
    }
    const double sqr_dz = dz * dz;
    const double r2 = sep2<R2F>(dx, dy, dz);
    const double adz = fabs(dz);
    bool in = ok && r2 < P.sqr_rmax && r2 >= P.sqr_rmin;
    if (is2d) in = in && dz > -P.pimax && dz < P.pimax;
    if (MODE == MODE_SMU) in = in && sqr_dz < r2 * P.sqr_mu_max;
    if (!in) return;
    // The reference's pruning tests (per-particle minimum separation, z window)
    // can only reject pairs within a few ulps of the coordinate scale of rmax or
    // pimax; only such pairs replay them.
    const bool near = r2 >= P.sqr_rmax - P.near_r2 || (MODE != MODE_DD && MODE != MODE_XI && adz >= P.pimax - P.near_dz);
    if (near) {
        // the cell pair record and the primary set, loaded here
        const CellPairRec* c = fresh(cpp);
        const DevLattice* l1 = fresh(l1p);
        ITerms t = {};
        i_terms(P, MODE, *c, l1->x[i], l1->y[i], l1->z[i], &t);
        if (is2d) {
            if (!c->same && (t.skip || zj <= t.target)) return;
        } else {
            if (t.skip || zj <= t.target || dz >= t.max_dz) return;
        }
    }
    const bool findex = FL < 0 ? P.float_index != 0 : float_index_of<MODE, R2F>();
    int k;
    if (fl_lut1<FL>(P) || P.sorted_edges) {
        int idx = (int)(__float_as_uint((float)r2) >> 19) - P.lut_base;
        idx = idx < 0 ? 0 : (idx >= P.lut_n ? P.lut_n - 1 : idx);
        k = S.lut[idx];
        if (fl_lut1<FL>(P)) {
            // at most one edge above lut[idx] can lie below r2 (e[nbin-1] = rmax^2 > r2)
            k += S.e[k] <= r2 ? 1 : 0;
        } else if (k < P.nbin - 1 && S.e[k] <= r2) {
            k++;
            while (k < P.nbin - 1 && S.e[k] <= r2) k++;
        }
    } else {
        k = find_kbin(S.e, P.nbin, 0, r2);
    }
    if (k == 0) return;
    int slot = k;
    if (MODE == MODE_RPPI) {
        if (findex) {
            slot = (int)((double)k * (double)(P.n2 + 1) + adz * P.inv_d2);
        } else {
            int pb = (int)(adz * P.inv_d2);
            pb = pb > P.n2 ? P.n2 : pb;
            slot = k * (P.n2 + 1) + pb;
        }
    } else if (MODE == MODE_SMU) {
        // mu bin from an FP32 estimate t ~ mu * inv_dmu: |dz|, s^2 and inv_dmu
        // rounded to FP32 (2^-24 each), the hardware reciprocal square root (1
        // ulp) and two products put t within 4e-7 * t of the reference's FP64
        // value, so its integer part is the reference's unless t lies within
        // 1e-6 * (t + 1) of an integer; only those pairs (and non-finite
        // estimates, clamped to 1e9) compute mu = sqrt(dz^2 / s^2) in FP64 as the
        // reference does (countpairs_s_mu_kernels.c.src:1154-1171).
        const float t = fminf((float)adz * __builtin_amdgcn_rsqf((float)r2) * (float)P.inv_d2, 1e9f);
        const int mbf = (int)t;
        const float fr = t - (float)mbf;
        const float tol = 1e-6f * (t + 1.0f);
        if (fr >= tol && fr <= 1.0f - tol) {
            slot = k * (P.n2 + 1) + (mbf > P.n2 ? P.n2 : mbf);
        } else {
            const double mu = sqrt(sqr_dz / r2);
            if (findex) {
                slot = (int)((double)k * (double)(P.n2 + 1) + mu * P.inv_d2);
            } else {
                int mb = (int)(mu * P.inv_d2);
                mb = mb > P.n2 ? P.n2 : mb;
                slot = k * (P.n2 + 1) + mb;
            }
        }
    }
    hist_pair<GH, FL>(P, S, slot, r2, wi * wj);
}

// Decides up to 64 queued survivors (entries ((j - jbase) << 6) | lane) in FP64.
template <int MODE, int R2F, bool GH, int FL>
__device__ __forceinline__ void drain_queue(const PairParams& P, const CellPairRec& cp, const CellPairRec* cpp,
                                            const DevLattice* l1p, const double4* d4_1, const double4* d4_2,
                                            const double3& wrap, int i0, int istride, const unsigned* qe,
                                            const double4* prim, unsigned head, int ndrain, int jbase,
                                            KernelLds& S)
{
    const int lane = lane_id();
    if (lane < ndrain) {
        const unsigned e = qe[head + lane];
        const int j = jbase + (int)(e >> 6), l = (int)(e & 63);
        const int i = i0 + l * istride;
        double4 pi;
        if (P.prim_cache) {
            pi = prim[l];  // the primary's wrapped position and weight
        } else {
            pi = d4_1[i];
            pi.x += wrap.x;
            pi.y += wrap.y;
            pi.z += wrap.z;
        }
        const double4 pj = d4_2[j];
        exact_pair<MODE, R2F, GH, FL>(P, cpp, l1p, i, 1, pi.x, pi.y, pi.z, pi.w, pj.x, pj.y, pj.z, pj.w, S);
    }
}

// Decides up to 128 queued survivors, two per lane: both pairs' records are
// loaded before either is decided, so two gathers are in flight per lane (the
// DP = 2 kernels, for launches whose LDS already limits the waves per SIMD to
// five or fewer: the extra registers then cost no residency).
template <int MODE, int R2F, bool GH, int FL>
__device__ __forceinline__ void drain_queue2(const PairParams& P, const CellPairRec& cp, const CellPairRec* cpp,
                                             const DevLattice* l1p, const double4* d4_1, const double4* d4_2,
                                             const double3& wrap, int i0, int istride, const unsigned* qe,
                                             const double4* prim, unsigned head, int ndrain, int jbase,
                                             KernelLds& S)
{
    const int lane = lane_id();
    const int oka = lane < ndrain, okb = lane + 64 < ndrain;
    // entries past the tail read as 0: j = jbase, primary i0 (valid indices)
    const unsigned ea0 = qe[head + lane], eb0 = qe[head + 64 + lane];
    const unsigned ea = oka ? ea0 : 0u, eb = okb ? eb0 : 0u;
    const int ja = jbase + (int)(ea >> 6), la = (int)(ea & 63);
    const int jb = jbase + (int)(eb >> 6), lb = (int)(eb & 63);
    const int ia = i0 + la * istride, ib = i0 + lb * istride;
    double4 pia, pib;
    if (P.prim_cache) {
        pia = prim[la];
        pib = prim[lb];
    } else {
        pia = d4_1[ia];
        pib = d4_1[ib];
        pia.x += wrap.x;
        pia.y += wrap.y;
        pia.z += wrap.z;
        pib.x += wrap.x;
        pib.y += wrap.y;
        pib.z += wrap.z;
    }
    const double4 pja = d4_2[ja], pjb = d4_2[jb];
    exact_pair<MODE, R2F, GH, FL>(P, cpp, l1p, ia, oka, pia.x, pia.y, pia.z, pia.w, pja.x, pja.y, pja.z, pja.w, S);
    exact_pair<MODE, R2F, GH, FL>(P, cpp, l1p, ib, okb, pib.x, pib.y, pib.z, pib.w, pjb.x, pjb.y, pjb.z, pjb.w, S);
}

// Integer pair filter on quantized cell-local positions (secondary s, primary
// qxy / qz). 3-D: (dqx^2 + dqy^2 + dqz^2 + qlo) mod 2^32 < qrange; projected:
// the same on (dqx, dqy) and |dqz| < zq, as (dqz + zq - 1) unsigned < 2 zq - 1
// with zoff = zq - 1 - qz. Differences fit 16 bits and sums of squares 32 bits
// (|q| <= kQMax), so the only rounding is that of the positions themselves.
struct QFilter {
    unsigned qxy;
    int qz, qlo, zoff;
    unsigned qrange, zspan;
    int ilim;                  // same cell: the lane takes j > ilim
};
// The wave's mask of lanes whose pair with secondary s (index j) survives.
template <int MODE, bool SAME>
__device__ __forceinline__ unsigned long long qpass(const QFilter& f, uint2 s, int j)
{
    const short2v d = __builtin_bit_cast(short2v, s.x) - __builtin_bit_cast(short2v, f.qxy);
    unsigned long long m;
    if (MODE == MODE_RPPI || MODE == MODE_WP) {
        const int r = __builtin_amdgcn_sdot2(d, d, f.qlo, false);
        m = __builtin_amdgcn_ballot_w64((unsigned)r < f.qrange) &
            __builtin_amdgcn_ballot_w64((unsigned)((int)s.y + f.zoff) < f.zspan);
    } else {
        const int dz = (int)s.y - f.qz;
        const int r = __builtin_amdgcn_sdot2(d, d, __mul24(dz, dz) + f.qlo, false);
        m = __builtin_amdgcn_ballot_w64((unsigned)r < f.qrange);
    }
    // same cell: the lane's own j > i (lanes without a primary have quantized
    // positions no secondary passes)
    return SAME ? m & __builtin_amdgcn_ballot_w64(j > f.ilim) : m;
}

// Per-lane select by a lane mask: the lane's bit of m set ? a : b (one
// v_cndmask; a C select here tends to become a branch around the computation of a).
__device__ __forceinline__ unsigned lane_select(unsigned long long m, unsigned a, unsigned b)
{
    unsigned r;
    asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(b), "v"(a), "s"(m));
    return r;
}

// Appends the lanes of m (nonzero) to the survivor buffer at tail with entry
// (j << 6) | lane. Branch-free: the other lanes write slot tail + 64 + lane, past
// this secondary's survivors, where a later survivor overwrites it or it stays
// beyond the tail.
__device__ __forceinline__ void enqueue(unsigned* qe, unsigned long long m, unsigned& tail, unsigned entry)
{
    const unsigned lane = (unsigned)lane_id();
    const unsigned b =
        (unsigned)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0));
    unsigned* qt = qe + __builtin_amdgcn_readfirstlane(tail);
    qt[lane_select(m, b, 64u + lane)] = entry;
    tail += (unsigned)__popcll(m);
}

// A block of kBlockJ quantized secondaries in scalar registers.
struct QBlock {
    uint2 s[kBlockJ];
};
__device__ __forceinline__ QBlock qload(KUint* qs, int jb)
{
    KUint* p = qs + 2 * (int64_t)jb;
    QBlock b;
#pragma unroll
    for (int t = 0; t < kBlockJ; t++) b.s[t] = make_uint2(p[2 * t], p[2 * t + 1]);
    return b;
}

// Filters block b (secondaries jb .. jb + kBlockJ - 1, of which the first
// `nlive` exist) and enqueues the survivors relative to jbase.
template <int MODE, bool SAME, bool PARTIAL>
__device__ __forceinline__ void qblock(const QFilter& f, const QBlock& b, int jb, int nlive, unsigned* qe,
                                       unsigned& tail, int jbase)
{
    unsigned long long m[kBlockJ];
#pragma unroll
    for (int t = 0; t < kBlockJ; t++) {
        m[t] = qpass<MODE, SAME>(f, b.s[t], jb + t);
        if (PARTIAL) m[t] = t < nlive ? m[t] : 0ull;
    }
    const unsigned e0 = ((unsigned)(jb - jbase) << 6) | (unsigned)lane_id();
#pragma unroll
    for (int t = 0; t < kBlockJ; t++)
        if (m[t]) enqueue(qe, m[t], tail, e0 + 64u * t);
}

// Filters the full blocks from jb on and enqueues their survivors, until fewer
// than kBlockJ secondaries are left or 64 survivors are waiting; returns the next
// jb. On entry and exit `a` holds the block at jb: the loop alternates between
// two register blocks so that each block is loaded one block ahead (q is padded
// by kQPad entries, so the look-ahead stays inside the array).
template <int MODE, bool SAME, int DRAIN>
__device__ __forceinline__ int qscan(const QFilter& f, KUint* qs, int jb, int jhi, QBlock& a, unsigned* qe,
                                     unsigned head, unsigned& tail, int jbase)
{
    int nfull = (jhi - jb) / kBlockJ;
    for (;;) {
        if (nfull <= 0 || tail - head >= DRAIN) return jb;
        const QBlock b = qload(qs, jb + kBlockJ);
        qblock<MODE, SAME, false>(f, a, jb, kBlockJ, qe, tail, jbase);
        jb += kBlockJ;
        if (--nfull <= 0 || tail - head >= DRAIN) {
            a = b;
            return jb;
        }
        a = qload(qs, jb + kBlockJ);
        qblock<MODE, SAME, false>(f, b, jb, kBlockJ, qe, tail, jbase);
        jb += kBlockJ;
        --nfull;
    }
}

// One work item of the tuned kernel (see the comment above kQueues).
template <int MODE, int R2F, bool GH, int DP, int FL>
__device__ __forceinline__ void process_item(const PairParams& P, const DevLattice* l1p, const DevLattice* l2p,
                                             const CellPairRec& cp, const CellPairRec* cpp, const WorkItem& it,
                                             KernelLds& S, unsigned* qe, double4* prim)
{
    // the lattices and the set-up parameters are loaded per item (fresh), so
    // that they do not occupy registers across the persistent loop
    const __attribute__((address_space(4))) DevLattice& L1 = *fresh_k(l1p);
    const __attribute__((address_space(4))) DevLattice& L2 = *fresh_k(l2p);
    const PairParams& Pi = P;
    const int lane = lane_id();
    const int i0 = __builtin_amdgcn_readfirstlane(it.i0), i1 = __builtin_amdgcn_readfirstlane(it.i1);
    const int flags = __builtin_amdgcn_readfirstlane(it.flags);
    const int dense = Pi.dense_mode == 1 ? 1 : (Pi.dense_mode == 2 ? 0 : (flags & kItemDense));
    // with P.dense_stride, dense cross-cell items take strided primaries (a
    // whole-cell sample per wave, spreading the lanes' pairs over more
    // histogram slots); else 64 consecutive ones
    const int stride = flags >> kStrideShift;
    const int i = i0 + lane * (stride > 0 ? stride : 1);
    const int ivalid = i < i1;
    const int c2 = cp.c2;
    const int same = cp.same;
    // a lane takes secondary j iff j > ilim (its primary exists; same cell: j > i)
    const int ilim = !ivalid ? INT_MAX : (same ? i : INT_MIN);
    const double3 wrap = make_double3(cp.wx * Pi.xwrap, cp.wy * Pi.ywrap, cp.wz * Pi.zwrap);
    const double4* d4_1 = L1.d4;
    const double4* d4_2 = L2.d4;
    double xpos = 0, ypos = 0, zpos = 0, wi = 0;
    if (ivalid) {
        const double4 pi = d4_1[i];
        xpos = pi.x + wrap.x;
        ypos = pi.y + wrap.y;
        zpos = pi.z + wrap.z;
        wi = pi.w;
    }
    // secondary z-slabs this chunk can reach
    const int iz2 = c2 % L2.nz;
    const double zlo = wave_min_d(ivalid ? zpos : DBL_MAX) - Pi.zreach - Pi.slab_margin;
    const double zhi = wave_max_d(ivalid ? zpos : -DBL_MAX) + Pi.zreach + Pi.slab_margin;
    int jlo, jhi;
    {
        const int64_t cb = (int64_t)c2 * kSlabs;
        double flo = floor(((zlo - L2.zmin) * L2.zinv - (double)iz2) * kSlabs);
        double fhi = floor(((zhi - L2.zmin) * L2.zinv - (double)iz2) * kSlabs);
        if (L2.zinv == 0.0) { flo = 0.0; fhi = kSlabs - 1; }
        if (fhi < 0.0 || flo > (double)(kSlabs - 1)) {
            jlo = jhi = 0;
        } else {
            const int slo = flo < 0.0 ? 0 : (int)flo;
            const int shi = fhi > (double)(kSlabs - 1) ? kSlabs - 1 : (int)fhi;
            jlo = L2.bstart[cb + slo];
            jhi = L2.bstart[cb + shi + 1];
        }
    }
    if (same && jlo < i0 + 1) jlo = i0 + 1;
    jlo = __builtin_amdgcn_readfirstlane(jlo);
    jhi = __builtin_amdgcn_readfirstlane(jhi);
    {
        const int nparts = __builtin_amdgcn_readfirstlane(it.nparts);
        if (nparts > 1 && jhi > jlo) {
            const int part = __builtin_amdgcn_readfirstlane(it.part);
            const int64_t len = jhi - jlo;
            const int a = jlo + (int)(len * part / nparts), b = jlo + (int)(len * (part + 1) / nparts);
            jlo = a;
            jhi = b;
        }
    }

    if (jhi <= jlo) return;

    if (dense) {
        // FP64 lockstep; the secondary record is a scalar load (d4 is padded, so
        // the look-ahead load of j + 1 stays inside the array)
        KDouble* ds = (KDouble*)(const void*)d4_2;
        double4 cur = kload_d4(ds, jlo);
        for (int j = jlo; j < jhi; j++) {
            const double4 nxt = kload_d4(ds, j + 1);
            exact_pair<MODE, R2F, GH, FL>(P, cpp, l1p, i, j > ilim, xpos, ypos, zpos, wi, cur.x, cur.y, cur.z,
                                          fl_w<FL>(P) ? cur.w : 0.0, S);
            cur = nxt;
        }
        return;
    }

    // sparse: quantize the wave's primaries relative to the secondary cell
    const int ixy2 = c2 / L2.nz, iy2 = ixy2 % L2.ny, ix2 = ixy2 / L2.ny;
    const double ox = L2.xmin + ix2 * L2.xbs, oy = L2.ymin + iy2 * L2.ybs, oz = L2.zmin + iz2 * L2.zbs;
    // a lane without a primary gets (qx, qy) = (-32768, -32768): every difference
    // to a secondary (0 <= q <= kQMax) wraps to |dq| >= 32768 - kQMax per axis, so
    // its squared distance exceeds any threshold (which is < (kQMax + 2)^2 + 1)
    // without overflowing 32 bits
    uint2 qi = quantize_pos(xpos - ox, ypos - oy, zpos - oz, L2.qinv);
    if (!ivalid) qi = make_uint2(0x80008000u, 0u);
    QFilter f;
    f.qxy = qi.x;
    f.qz = (int)qi.y;
    f.qlo = L2.qlo_neg;
    f.qrange = L2.qrange;
    f.zoff = L2.qz - 1 - f.qz;
    f.zspan = 2u * (unsigned)L2.qz - 1u;
    f.ilim = ilim;
    const int istride = stride > 0 ? stride : 1;
    unsigned head = 0, tail = 0;
    // the wave's primaries (wrapped, FP64) for the drains
    if (P.prim_cache) {
        prim[lane] = make_double4(xpos, ypos, zpos, wi);
        __builtin_amdgcn_wave_barrier();
    }
    KUint* qs = (KUint*)(const void*)L2.q;
    const int max_window = Pi.max_window;
    constexpr int kDrain = 64 * DP;
    for (int wlo = jlo; wlo < jhi; wlo += max_window) {
        const int whi = jhi - wlo > max_window ? wlo + max_window : jhi;
        QBlock a = qload(qs, wlo);
        int jb = wlo;
        while (jb < whi) {
            if (same)
                jb = qscan<MODE, true, kDrain>(f, qs, jb, whi, a, qe, head, tail, wlo);
            else
                jb = qscan<MODE, false, kDrain>(f, qs, jb, whi, a, qe, head, tail, wlo);
            if (jb < whi && jb + kBlockJ > whi && tail - head < kDrain) {
                // the last, partial block
                if (same)
                    qblock<MODE, true, true>(f, a, jb, whi - jb, qe, tail, wlo);
                else
                    qblock<MODE, false, true>(f, a, jb, whi - jb, qe, tail, wlo);
                jb = whi;
            }
            while (tail - head >= kDrain) {
                __builtin_amdgcn_wave_barrier();
                if (DP == 2)
                    drain_queue2<MODE, R2F, GH, FL>(P, cp, cpp, l1p, d4_1, d4_2, wrap, i0, istride, qe, prim, head,
                                                kDrain, wlo, S);
                else
                    drain_queue<MODE, R2F, GH, FL>(P, cp, cpp, l1p, d4_1, d4_2, wrap, i0, istride, qe, prim, head,
                                               kDrain, wlo, S);
                __builtin_amdgcn_wave_barrier();
                head += kDrain;
            }
            if (head >= kCompactAt) {
                // move the < kDrain waiting entries to the front
                const unsigned n = tail - head;
                unsigned v[DP];
#pragma unroll
                for (int p = 0; p < DP; p++) {
                    v[p] = 0;
                    if ((unsigned)(lane + 64 * p) < n) v[p] = qe[head + lane + 64 * p];
                }
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int p = 0; p < DP; p++)
                    if ((unsigned)(lane + 64 * p) < n) qe[lane + 64 * p] = v[p];
                __builtin_amdgcn_wave_barrier();
                head = 0;
                tail = n;
            }
        }
        if (tail != head) {
            __builtin_amdgcn_wave_barrier();
            if (DP == 2)
                drain_queue2<MODE, R2F, GH, FL>(P, cp, cpp, l1p, d4_1, d4_2, wrap, i0, istride, qe, prim, head,
                                            (int)(tail - head), wlo, S);
            else
                drain_queue<MODE, R2F, GH, FL>(P, cp, cpp, l1p, d4_1, d4_2, wrap, i0, istride, qe, prim, head,
                                           (int)(tail - head), wlo, S);
            __builtin_amdgcn_wave_barrier();
        }
        head = tail = 0;
    }
}

// Adds a (replicated) LDS histogram region to the global histogram at `base`
// and clears it; `t0`/`tstep` split the slots among the calling threads.
__device__ __forceinline__ void flush_region(const PairParams& P, KernelLds& S, int R, int64_t base, int t0,
                                             int tstep, unsigned long long* g_np, double* g_avg, double* g_w)
{
    for (int sl = t0; sl < P.nslots; sl += tstep) {
        unsigned long long c = 0;
        double sa = 0, sw = 0;
        for (int r = 0; r < R; r++) {
            c += S.np[sl * R + r];
            S.np[sl * R + r] = 0;
            if (P.need_avg) { sa += S.avg[sl * R + r]; S.avg[sl * R + r] = 0; }
            if (P.need_w) { sw += S.w[sl * R + r]; S.w[sl * R + r] = 0; }
        }
        if (c) {
            atomicAdd(&g_np[base + sl], c);
            if (P.need_avg) atomicAdd(&g_avg[base + sl], sa);
            if (P.need_w) atomicAdd(&g_w[base + sl], sw);
        }
    }
}

// The tuned pair kernel. lats[0], lats[1]: primary / secondary particle sets;
// in a batch (BATCH = true) every cell pair carries its catalog index k in
// `pad`, both sets are lats[k], each wave keeps its own LDS histogram for the
// catalog it is working on, and the histogram of catalog k lives at
// g_*[k*nslots ...]. GH = true: no LDS histogram; every counted pair is added
// to the global histogram (histograms too large for the workgroup's LDS).
// DP = survivors per lane per drain (1: compiled for six waves per SIMD, <= 80
// VGPRs; 2: for five, <= 96 VGPRs). FL = the launch's flags (fl_avg etc.):
// a mask for the LDS-histogram kernels, kFlRuntime for the GH ones.
template <int MODE, int R2F, bool BATCH, bool GH, int DP, int FL>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(DP == 2 ? 5 : 6))) void k_pairs_tuned(PairParams P, const DevLattice* __restrict__ lats,
                                                     const CellPairRec* __restrict__ cps,
                                                     const WorkItem* __restrict__ items,
                                                     const int* __restrict__ nitems_p, int* __restrict__ counter,
                                                     const double* __restrict__ rupp_sqr,
                                                     const unsigned short* __restrict__ lut_g,
                                                     unsigned long long* __restrict__ g_np,
                                                     double* __restrict__ g_avg, double* __restrict__ g_w,
                                                     unsigned long long* __restrict__ cp_ticks,
                                                     int* __restrict__ cp_cu, const int* __restrict__ single_p)
{
    extern __shared__ unsigned long long smem_u64[];
    KernelLds S;
    const int R = GH ? 1 : P.hist_copies;
    const int nwaves = blockDim.x >> 6;
    const int nrep = GH ? 0 : P.nslots * R;         // one histogram region
    const int nreg = BATCH ? nrep * nwaves : nrep;  // all regions of the workgroup
    const int wave = threadIdx.x >> 6, lane = lane_id();
    unsigned long long* np0 = smem_u64;
    double* avg0 = reinterpret_cast<double*>(smem_u64 + nreg);
    double* w0 = avg0 + (P.need_avg ? nreg : 0);
    double* e = w0 + (P.need_w ? nreg : 0);
    const int roff = BATCH ? wave * nrep : 0;
    S.np = np0 + roff;
    S.avg = avg0 + roff;
    S.w = w0 + roff;
    S.rmask = R - 1;
    S.lane_copy = lane & (R - 1);
    S.gnp = g_np;
    S.gavg = g_avg;
    S.gw = g_w;
    unsigned short* lut = reinterpret_cast<unsigned short*>(e + P.nbin);
    int* qbase = reinterpret_cast<int*>(lut + ((P.lut_n + 1) & ~1));
    S.e = e;
    S.lut = lut;
    unsigned* qe = reinterpret_cast<unsigned*>(qbase) + __builtin_amdgcn_readfirstlane(wave) * kQueueAlloc;
    // primaries cache: 16-byte aligned after the queues
    // (offsets from the LDS base, so the pointer stays an LDS pointer)
    const size_t poff = ((size_t)(reinterpret_cast<char*>(qbase + nwaves * kQueueAlloc) -
                                  reinterpret_cast<char*>(smem_u64)) + 15) & ~(size_t)15;
    double4* prim = reinterpret_cast<double4*>(reinterpret_cast<char*>(smem_u64) + poff) + wave * 64;
    for (int s = threadIdx.x; s < P.nbin; s += blockDim.x) e[s] = rupp_sqr[s];
    for (int s = threadIdx.x; s < P.lut_n; s += blockDim.x) lut[s] = lut_g[s];
    for (int s = threadIdx.x; s < nreg; s += blockDim.x) {
        np0[s] = 0;
        if (P.need_avg) avg0[s] = 0;
        if (P.need_w) w0[s] = 0;
    }
    __syncthreads();
    const int nitems = *nitems_p;
    const DevLattice* l1p = lats;
    const DevLattice* l2p = lats + (BATCH ? 0 : 1);
    int cur = -1;  // catalog whose histogram the wave's LDS region holds (BATCH)
    // Work queue: the list is cut into blocks of B = 2^qblock_log2 consecutive
    // items, and block b belongs to sub-queue b % kQueues (the cost order is kept
    // inside every sub-queue): the k-th item of sub-queue q is
    // g = ((k / B) * kQueues + q) * B + k % B. Sub-queue q has its own counter at
    // counter[q * kCounterStride], so the device-wide atomics spread over kQueues
    // addresses. A wave starts at sub-queue blockIdx.x % kQueues -- workgroups are
    // dispatched to the XCDs round-robin, so that is mostly the wave's own XCD --
    // and moves on to the next when that one is exhausted, until all are. Blocks
    // keep the items of a cell pair, which scan the same secondary window, on one
    // XCD and its L2. Items below *single_p (the heavy cell pairs the scheduler
    // puts first) are taken one at a time; the others P.fetch per atomic while
    // more than 2 P.fetch items per wave of the grid remain (as last seen), then
    // one at a time.
    // The loop has no lane-divergent branch around its wave-wide steps: every
    // lane executes the queue atomic (lane 0 adds the batch, the others 0) and
    // the broadcast reads lane 0's result. A divergent `if (lane == 0)` here lets
    // the compiler start the next iteration's broadcast without lane 0
    // (observed in the generated code).
    const int tail_zone = (int)min((long)nitems, (long)gridDim.x * nwaves * P.fetch * 2);
    const int n_single = *single_p;
    const int home = blockIdx.x & (kQueues - 1);
    const int lb = P.qblock_log2, bmask = (1 << lb) - 1;
    const long round = (long)kQueues << lb;  // items of one block per sub-queue
    int qi = 0, q = 0, next = 0, end = 0, last = 0;
    for (;;) {
        if (next >= end) {
            bool got = false;
            while (qi < kQueues) {
                q = (home + qi) & (kQueues - 1);
                const long full = (long)nitems / round;
                const long rem = (long)nitems - full * round - ((long)q << lb);
                const int cnt = (int)((full << lb) + (rem < 0 ? 0 : (rem > bmask ? bmask + 1 : rem)));
                const int want = (last >= n_single && last < nitems - tail_zone) ? P.fetch : 1;
                const int fetched = atomicAdd(&counter[q * kCounterStride], lane == 0 ? want : 0);
                const int k = __builtin_amdgcn_readlane(fetched, 0);
                if (k < cnt) {
                    next = k;
                    end = min(k + want, cnt);
                    got = true;
                    break;
                }
                qi++;
            }
            if (!got) break;
        }
        const int item = (((next >> lb) * kQueues + q) << lb) + (next & bmask);
        next++;
        last = item;
        const WorkItem it = items[item];
        const CellPairRec* cpp = cps + __builtin_amdgcn_readfirstlane(it.cp);
        const CellPairRec cp = *cpp;
        if (BATCH) {
            const int k = __builtin_amdgcn_readfirstlane(cp.pad);
            if (k != cur) {
                if (cur >= 0) {
                    __builtin_amdgcn_wave_barrier();
                    flush_region(P, S, R, (int64_t)cur * P.nslots, lane, 64, g_np, g_avg, g_w);
                    __builtin_amdgcn_wave_barrier();
                }
                cur = k;
                l1p = lats + k;
                l2p = l1p;
            }
        }
        // per-cell-pair cost record: wall-clock ticks of the wave on this item
        // (branch-free in the lanes: lane 0 adds dt, the others 0; every lane
        // stores the same compute-unit id)
        const unsigned long long t0 = cp_ticks ? wall_clock64() : 0ull;
        process_item<MODE, R2F, GH, DP, FL>(P, l1p, l2p, cp, cpp, it, S, qe, prim);
        if (cp_ticks) {
            const unsigned long long dt = wall_clock64() - t0;
            atomicAdd(&cp_ticks[it.cp], lane == 0 ? dt : 0ull);
            cp_cu[it.cp] = (int)__smid();
        }
    }
    if (GH) return;
    if (BATCH) {
        if (cur >= 0) {
            __builtin_amdgcn_wave_barrier();
            flush_region(P, S, R, (int64_t)cur * P.nslots, lane, 64, g_np, g_avg, g_w);
        }
        return;
    }
    __syncthreads();
    flush_region(P, S, R, 0, threadIdx.x, blockDim.x, g_np, g_avg, g_w);
}

#define CF_TUNED_INST0(M, F, B, G, D, FL)                                                                      \
    template __global__ void k_pairs_tuned<M, F, B, G, D, FL>(PairParams, const DevLattice*, const CellPairRec*, \
                                                              const WorkItem*, const int*, int*, const double*,  \
                                                              const unsigned short*, unsigned long long*,         \
                                                              double*, double*, unsigned long long*, int*,        \
                                                              const int*);
// LDS-histogram kernels: one per flag mask; global-histogram kernels: runtime flags
#define CF_TUNED_INST_FL(M, F, B, D)                                                                           \
    CF_TUNED_INST0(M, F, B, false, D, 0) CF_TUNED_INST0(M, F, B, false, D, 1)                                   \
    CF_TUNED_INST0(M, F, B, false, D, 2) CF_TUNED_INST0(M, F, B, false, D, 3)                                   \
    CF_TUNED_INST0(M, F, B, false, D, 4) CF_TUNED_INST0(M, F, B, false, D, 5)                                   \
    CF_TUNED_INST0(M, F, B, false, D, 6) CF_TUNED_INST0(M, F, B, false, D, 7)
#define CF_TUNED_INST(M, F, B) CF_TUNED_INST_FL(M, F, B, 1) CF_TUNED_INST_FL(M, F, B, 2)
#define CF_TUNED_INST2(M, F)                                                                                   \
    CF_TUNED_INST(M, F, false) CF_TUNED_INST0(M, F, false, true, 1, kFlRuntime)                                 \
    CF_TUNED_INST0(M, F, false, true, 2, kFlRuntime)
CF_TUNED_INST2(MODE_DD, R2_SUM_XY_Z)
CF_TUNED_INST2(MODE_DD, R2_FMA_DZ_DY_DX)
CF_TUNED_INST2(MODE_XI, R2_SUM_XY_Z)
CF_TUNED_INST2(MODE_XI, R2_X_SUM_YZ)
CF_TUNED_INST2(MODE_XI, R2_FMA_DZ_DY_DX)
CF_TUNED_INST2(MODE_RPPI, RP_SUM)
CF_TUNED_INST2(MODE_RPPI, RP_FMA)
CF_TUNED_INST2(MODE_WP, RP_SUM)
CF_TUNED_INST2(MODE_WP, RP_FMA)
CF_TUNED_INST2(MODE_SMU, R2_SUM_XY_Z)
CF_TUNED_INST2(MODE_SMU, R2_FMA_DX_DY_DZ)
CF_TUNED_INST2(MODE_RPPI_MOCKS, R2_SUM_XY_Z)
CF_TUNED_INST2(MODE_SMU_MOCKS, R2_SUM_XY_Z)
CF_TUNED_INST(MODE_DD, R2_SUM_XY_Z, true)
CF_TUNED_INST(MODE_DD, R2_FMA_DZ_DY_DX, true)
#undef CF_TUNED_INST2
#undef CF_TUNED_INST
#undef CF_TUNED_INST_FL
#undef CF_TUNED_INST0

// Largest squared separation between the bounding boxes of a cell pair
// (primary box shifted by the wrap), per the geometry of the mode, and whether
// the pair takes the dense path: every pair of its particles is inside the range,
// or the estimated in-range fraction -- the share of the squared-separation span
// [min_sep^2, max_sep^2] below rmax^2 (times the |dz| share below pimax in the
// projected modes) -- reaches P.dense_frac.
__device__ __forceinline__ int cp_dense(const PairParams& P, const CellPairRec& cp, const double* b1,
                                        const double* b2, double* maxsep2)
{
    const double ox = cp.wx * P.xwrap, oy = cp.wy * P.ywrap, oz = cp.wz * P.zwrap;
    const double mx = fmax(fabs(b2[1] - (b1[0] + ox)), fabs((b1[1] + ox) - b2[0]));
    const double my = fmax(fabs(b2[3] - (b1[2] + oy)), fabs((b1[3] + oy) - b2[2]));
    const double mz = fmax(fabs(b2[5] - (b1[4] + oz)), fabs((b1[5] + oz) - b2[4]));
    const bool is2d = P.mode == MODE_RPPI || P.mode == MODE_WP;
    const double hi = is2d ? mx * mx + my * my : mx * mx + my * my + mz * mz;
    *maxsep2 = hi;
    if (is2d ? (hi < 0.98 * P.sqr_rmax && mz < 0.98 * P.pimax) : hi < 0.98 * P.sqr_rmax) return 1;
    const double lo = cp.mdx * cp.mdx + cp.mdy * cp.mdy + (is2d ? 0.0 : cp.mdz * cp.mdz);
    double f = hi > lo ? (P.sqr_rmax - lo) / (hi - lo) : 0.0;
    f = fmin(1.0, fmax(0.0, f));
    if (is2d) f *= mz > cp.mdz ? fmin(1.0, fmax(0.0, (P.pimax - cp.mdz) / (mz - cp.mdz))) : 1.0;
    return f >= (double)P.dense_frac;
}

// A predicted log2 cost clamped to [0, kCostTop] (NaN -> 0), so that any model
// output gives a valid bucket and a finite total.
__device__ __forceinline__ float sane_cost(float c)
{
    return c >= 0.f ? (c <= kCostTop ? c : kCostTop) : 0.f;
}

// Work-item construction: chunks of 64 primary particles per cell pair, each
// split into `parts` slices of its secondary window. Position t of the list
// holds cell pair order[t] (order == nullptr: t). With a cost array (log2
// units), a chunk whose predicted cost exceeds total / cap_div is split into
// ceil(cost / cap) parts (at most kMaxParts).
// ncp_dev (optional): a device count that bounds ncp; entries of the count
// array past it are 0 (a graph-captured batch scans its whole capacity).
__global__ void k_count_chunks(int ncp, const int* __restrict__ order, const CellPairRec* __restrict__ cps,
                               const DevLattice* __restrict__ lats, int batch, const float* __restrict__ cost,
                               const double* __restrict__ total, double cap_div, int* __restrict__ nparts,
                               int* __restrict__ nitems, const int* __restrict__ ncp_dev = nullptr)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncp) return;
    if (ncp_dev && t >= *ncp_dev) {
        nitems[t] = 0;
        return;
    }
    const int c = order ? order[t] : t;
    const int* bs1 = lats[batch ? cps[c].pad : 0].bstart;
    const int n = (int)cell_count(bs1, cps[c].c1);
    const int nch = (n + 63) / 64;
    int parts = 1;
    if (cost && nch > 0) {
        const double cap = *total / cap_div;
        const double per_chunk = exp2((double)sane_cost(cost[c])) / nch;
        if (cap > 0 && per_chunk > cap) parts = (int)fmin((double)kMaxParts, ceil(per_chunk / cap));
    }
    if (nparts) nparts[t] = parts;
    nitems[t] = nch * parts;
}

// Items are written only below items_cap (a graph-captured batch's capacity).
__global__ void k_emit_chunks(int ncp, const int* __restrict__ order, const CellPairRec* __restrict__ cps,
                              const DevLattice* __restrict__ lats, int batch, PairParams P,
                              const int* __restrict__ off, const int* __restrict__ nparts,
                              WorkItem* __restrict__ items, const int* __restrict__ ncp_dev = nullptr,
                              int items_cap = INT_MAX)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncp || (ncp_dev && t >= *ncp_dev)) return;
    const int c = order ? order[t] : t;
    const CellPairRec cp = cps[c];
    const int* bs1 = lats[batch ? cp.pad : 0].bstart;
    const double* bd1 = lats[batch ? cp.pad : 0].bounds;
    const double* bd2 = lats[batch ? cp.pad : 1].bounds;
    double maxsep2;
    const int dense = cp_dense(P, cp, bd1 + (int64_t)cp.c1 * 6, bd2 + (int64_t)cp.c2 * 6, &maxsep2);
    const int s = bs1[(int64_t)cp.c1 * kSlabs], e = bs1[(int64_t)(cp.c1 + 1) * kSlabs];
    const int parts = nparts ? nparts[t] : 1;
    int o = off[t];
    const int nch = (e - s + 63) / 64;
    for (int ch = 0; ch < nch; ch++) {
        WorkItem w;
        w.cp = c;
        if (dense && P.dense_stride && !cp.same) {
            w.i0 = s + ch;
            w.i1 = e;
            w.flags = kItemDense | (nch << kStrideShift);
        } else {
            w.i0 = s + 64 * ch;
            w.i1 = w.i0 + 64 < e ? w.i0 + 64 : e;
            w.flags = dense ? kItemDense : 0;
        }
        w.nparts = parts;
        for (int p = 0; p < parts; p++) {
            w.part = p;
            if (o < items_cap) items[o] = w;
            o++;
        }
    }
}

// dst[0] = src[0] (a device-side gather of a count computed by a scan).
__global__ void k_copy_int(const int* __restrict__ src, int* __restrict__ dst)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = *src;
}

// *single = item index where the light cell pairs start: the work-list offset
// of ordered position *nheavy.
__global__ void k_single_boundary(const int* __restrict__ nheavy, const int* __restrict__ off,
                                  int* __restrict__ single)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *single = off[*nheavy];
}

}  // namespace cfamd
