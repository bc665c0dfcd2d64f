// cf_device.h -- shared host/device declarations of the MI355X pair-counting engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cfamd {

// MODE_RPPI_MOCKS / MODE_SMU_MOCKS: on-sky catalogs (Cartesian positions from
// RA, DEC, distance), (rp, pi) / (s, mu) split along the pair's mean line of sight.
enum Mode { MODE_DD = 0, MODE_RPPI = 1, MODE_SMU = 2, MODE_XI = 3, MODE_WP = 4, MODE_RPPI_MOCKS = 5, MODE_SMU_MOCKS = 6 };

// FP64 rounding orders of the reference per-ISA kernels (which mode and ISA uses which: r2form_for).
enum R2Form {
    R2_SUM_XY_Z = 0,      // (dx*dx + dy*dy) + dz*dz          fallback / SSE / AVX (DD, DDsmu)
    R2_FMA_DZ_DY_DX = 1,  // fma(dz,dz,fma(dy,dy,dx*dx))      AVX512F DD, xi
    R2_X_SUM_YZ = 2,      // dx*dx + (dy*dy + dz*dz)          AVX/AVX2 xi
    R2_FMA_DX_DY_DZ = 3,  // fma(dx,dx,fma(dy,dy,dz*dz))      AVX512F DDsmu
    RP_SUM = 10,          // dx*dx + dy*dy                    fallback rp
    RP_FMA = 11           // fma(dy,dy,dx*dx)                 AVX512F rp_pi / wp, AVX2 wp
};

// Number of z-slabs each cell is bucketed into by the gridding sort.
constexpr int kSlabs = 16;

// Quantized positions: |q| <= kQMax, so differences fit int16 and a sum of three
// squared differences fits uint32; kQPad entries pad the array for block loads.
constexpr int kQMax = 16000;
constexpr int kQPad = 8;

// One neighbour-cell pair of the work list (64 bytes). Mirrors the fields of
// the reference's struct cell_pair (utils/cell_pair.h.src:20-37); the wrap
// offsets are stored as signs so off = sign * wrap is exact.
struct CellPairRec {
    int c1, c2;
    signed char wx, wy, wz, same;
    int pad;
    double mdx, mdy, mdz;   // min_dx, min_dy, min_dz
    double cx, cy, cz;      // closest_x1, closest_y1, closest_z1
};
static_assert(sizeof(CellPairRec) == 64, "CellPairRec must be 64 bytes");

// A gridded particle set resident in HBM, sorted by (cell, z-slab).
struct DevLattice {
    int nx, ny, nz;
    int64_t ncells;
    int64_t n;
    double xmin, ymin, zmin;
    double xbs, ybs, zbs;      // cell sizes
    double xinv, yinv, zinv;   // 1/cell size (0 if the size is 0)
    double *x, *y, *z, *w;     // FP64 positions / weight, cell order
    double4 *d4;               // the same as one 32-byte record (x, y, z, w) per particle
    // Cell-local quantized positions for the integer pair filter: per particle
    // (qx & 0xffff) | (qy << 16) and qz, q = rint((x - ox) * qinv), ox = xmin + ix*xbs,
    // followed by kQPad padding entries.
    uint2 *q;
    double qinv;               // quanta per unit length: kQMax quanta span a cell plus the pair reach
    unsigned qrange;           // filter: pass iff (dq^2 + qlo_neg) mod 2^32 < qrange
    int qlo_neg;
    int qz;                    // projected modes: pass iff |dqz| < qz
    int qpad_;
    int *bstart;               // [ncells*kSlabs + 1] bucket starts
    double *bounds;            // [ncells*6] x lo,hi, y lo,hi, z lo,hi
};

// Parameters of one pair-counting launch.
struct PairParams {
    int mode;          // Mode
    int r2form;        // R2Form
    int float_index;   // AVX512F flattened (rp,pi)/(s,mu) index rule
    int nbin;          // edges
    int n2;            // npibin / nmu_bins (0 for 1-D modes)
    int nslots;        // histogram slots
    int need_avg, need_w;
    int autocorr;
    int sorted_edges;  // rupp_sqr non-decreasing -> binary search
    double sqr_rmax, sqr_rmin; // on-sky modes: sqr_rmax is the 3-D limit (rpmax^2 + pimax^2, or smax^2)
    double pimax;      // DD/xi: rmax; rp_pi/wp: pimax; s_mu: smax*mu_max; on-sky: pimax or mu_max
    double sqr_rpmax;  // on-sky (rp, pi): rpmax^2
    double inv_d2, sqr_mu_max;
    double max_all_dz_same;  // same-cell z window (DD/xi/s_mu)
    double xwrap, ywrap, zwrap;
    double near_r2, near_dz; // edge windows below which pruning replay is skipped
    double zreach;           // z reach used for the slab window
    double slab_margin;
    // bin lookup: slot count of edges below the 1/16-octave bucket of (float)r^2
    int lut_base;            // bucket index of lut[0]
    int lut_n;               // entries
    int lut_single;          // 1: no bucket holds more than one edge, so one compare finishes the lookup
    int dense_mode;          // 0: by work-item flag, 1: always dense, 2: never dense
    int hist_copies;         // LDS histogram replication R (power of 2): lane l adds to copy l % R
    int dense_stride;        // 1: dense work items take strided primaries
    float dense_frac;        // a cell pair whose estimated in-range fraction reaches this takes the dense path
    int fetch;               // work items per queue atomic outside the single-fetch ranges
    int prim_cache;          // 1: the sparse path keeps the wave's primaries in LDS for its drains
    int max_window;          // longest secondary window scanned with one queue base (2^25, a multiple of 4)
    int qblock_log2;         // work items go to sub-queues in blocks of 2^qblock_log2 consecutive items
};

// Work item of the tuned kernel: (cell pair, <= 64 primary particles, one part
// of their secondary window).
struct WorkItem {
    int cp;
    int i0, i1;      // contiguous items: primaries [i0, i1); strided: i0 + lane*stride < i1
    int flags;       // bit 0: dense (every pair of the cell pair inside the range); bits 8..: stride
    int part, nparts;  // the item scans part `part` of `nparts` equal slices of the secondary window
};
constexpr int kItemDense = 1;
constexpr int kStrideShift = 8;
constexpr int kMaxParts = 32;

// Per-cell-pair features of the cost model (computed on the device).
constexpr int kNumFeatures = 10;
// f0 log2(1+N1)       f1 log2(1+N2)        f2 log2(chunks of 64 primaries)
// f3 same cell        f4 dense (bounding boxes entirely in range)
// f5 min separation^2 / rmax^2              f6 max separation^2 / rmax^2 (<= 8)
// f7 fraction of the secondary cell's z extent within z reach of the primaries
// f8 projected (rp, pi) mode                f9 log2(1 + N2 * f7)

// Cost-ordered work lists: heavy cell pairs are bucketed by log2 cost, descending.
constexpr float kCostTop = 48.0f;        // log2 cost mapped to bucket 0
constexpr int kSortPerOctave = 4;
constexpr int kSortBuckets = 48 * 4 + 1; // + the light bucket (list order)
constexpr int kSortBlock = 1024;         // cell pairs per block of the sort
constexpr float kHeavyFactor = 4.0f;     // default heavy threshold: cost >= 4 x the mean cell pair

}  // namespace cfamd
