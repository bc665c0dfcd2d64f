// cf_bgraph.hip -- device side of the graph-captured batch (cf_amd_batch_graph_*).
//
// cf_amd_countpairs_batch takes its grid, cell-pair and work-item decisions on
// the host from the data (one synchronisation per phase). Here every decision
// that depends on the positions is taken on the device, and every later step
// reads its counts from device memory and is launched over capacities fixed
// when the graph is captured, so one hipGraph replays the whole batch --
// extents, grids, counting sorts, cell pairs, work list and the persistent
// pair kernel -- for new positions of the same catalog sizes (jackknife and
// bootstrap resamples). A capacity that the new positions would exceed is
// never written past: the step clamps and sets a bit of the fail word, which
// the launch reports.
#include "cf_device.h"

namespace cfamd {

// Options of a graph batch (host values at capture; DD(r) autocorrelations).
struct BgOpts {
    double rmax;
    double sqr_rmax, sqr_rmin;  // squared outer and inner edges (quantized filter thresholds)
    double wrap[3];
    int per[3];
    int refine[3];      // the options' bin refine factors
    int custom;         // custom refine factors: no 0.05-wrap rule and no boost
    int max_cells;      // options->max_cells_per_dim
    int min_sep_opt;
    int64_t ncells_cap; // cells (and 16 z-slabs each) the catalog's buffers hold
};

// Per-catalog plan written by k_bg_grid and read by the later steps.
struct BgCat {
    GridArgs g;
    NgbArgs a;
    int64_t ncells;
    int nngb;
    int refine[3];
};

constexpr int kBgFailGrid = 1;   // rmax >= wrap / 2 on a periodic axis
constexpr int kBgFailCells = 2;  // the grid needs more cells than captured
constexpr int kBgFailOob = 4;    // positions outside their grid (NaN)
constexpr int kBgFailPairs = 8;  // more cell pairs than captured
constexpr int kBgFailItems = 16; // more work items than captured

// get_binsize (gridlink_utils.c.src:31-49), as the host's version.
__device__ int bg_binsize(double xdiff, double xwrap, double rmax, int refine, int max_ncells, double* binsize,
                          int* nl)
{
    int nmesh = (int)(refine * xdiff / rmax);
    nmesh = nmesh < 1 ? 1 : nmesh;
    if (xwrap > 0 && rmax >= xwrap / 2) return 1;
    if (nmesh > max_ncells) nmesh = max_ncells;
    if (nmesh < 2) nmesh = 2;
    *binsize = xdiff / nmesh;
    *nl = nmesh;
    return 0;
}

__device__ int bg_make_grid(const BgOpts& o, const double* ext, const int refine[3], int n[3], double bs[3])
{
    int bad = 0;
    for (int d = 0; d < 3; d++) bad |= bg_binsize(ext[3 + d] - ext[d], o.wrap[d], o.rmax, refine[d], o.max_cells, &bs[d],
                                                  &n[d]);
    return bad;
}

// One thread: the grid of catalog k from its extent, with the host's rules for
// DD(r) (countpairs_impl.c.src:272-332: refine 1 on an axis with rmax below
// 0.05 of the wrap, boost x and y by one when the grid is small or crowded),
// then the lattice record the pair kernel reads and the quantized filter's
// thresholds (as set_quant_filter). A grid beyond the captured cells, or an
// invalid one, is replaced by a 2 x 2 x 2 grid (which every buffer holds) and
// flagged.
__global__ void k_bg_grid(int k, BgOpts o, const double* __restrict__ ext, int64_t n, DevLattice tmpl,
                          BgCat* __restrict__ cats, DevLattice* __restrict__ lats, int* __restrict__ fail)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double* e = ext + 6 * (int64_t)k;
    int refine[3] = {o.refine[0], o.refine[1], o.refine[2]};
    if (!o.custom)
        for (int d = 0; d < 3; d++)
            if (o.rmax < 0.05 * o.wrap[d]) refine[d] = 1;
    int nm[3];
    double bs[3];
    int bad = bg_make_grid(o, e, refine, nm, bs);
    if (!bad) {
        const double avg_np = ((double)n) / ((double)nm[0] * nm[1] * nm[2]);
        const int max_nmesh = nm[0] > nm[1] ? (nm[0] > nm[2] ? nm[0] : nm[2]) : (nm[1] > nm[2] ? nm[1] : nm[2]);
        if ((max_nmesh <= 10 || avg_np >= 250) && max_nmesh < o.max_cells && !o.custom) {
            refine[0] += 1;
            refine[1] += 1;
        }
        bad = bg_make_grid(o, e, refine, nm, bs);
    }
    int f = bad ? kBgFailGrid : 0;
    if (!bad && (int64_t)nm[0] * nm[1] * nm[2] > o.ncells_cap) f |= kBgFailCells;
    if (f) {
        for (int d = 0; d < 3; d++) {
            nm[d] = 2;
            bs[d] = (e[3 + d] - e[d]) / 2;
        }
        atomicOr(fail, f);
    }
    BgCat c;
    c.g.xmin = e[0]; c.g.ymin = e[1]; c.g.zmin = e[2];
    c.g.xbs = bs[0]; c.g.ybs = bs[1]; c.g.zbs = bs[2];
    c.g.xinv = bs[0] > 0 ? 1.0 / bs[0] : 0.;
    c.g.yinv = bs[1] > 0 ? 1.0 / bs[1] : 0.;
    c.g.zinv = bs[2] > 0 ? 1.0 / bs[2] : 0.;
    c.g.nx = nm[0]; c.g.ny = nm[1]; c.g.nz = nm[2];
    const double qreach = o.rmax * (1.0 + 1e-6);
    c.g.qinv = (double)kQMax / (fmax(bs[0], fmax(bs[1], bs[2])) + qreach);
    c.ncells = (int64_t)nm[0] * nm[1] * nm[2];
    c.refine[0] = refine[0]; c.refine[1] = refine[1]; c.refine[2] = refine[2];
    NgbArgs& a = c.a;
    a.nx = nm[0]; a.ny = nm[1]; a.nz = nm[2];
    a.rx = refine[0]; a.ry = refine[1]; a.rz = refine[2];
    a.px = o.per[0]; a.py = o.per[1]; a.pz = o.per[2];
    a.autocorr = 1;
    a.min_sep_opt = o.min_sep_opt;
    a.dupcheck = (o.per[0] || o.per[1] || o.per[2]) &&
                 (nm[0] < 2 * refine[0] + 1 || nm[1] < 2 * refine[1] + 1 || nm[2] < 2 * refine[2] + 1);
    a.xwrap = o.wrap[0]; a.ywrap = o.wrap[1]; a.zwrap = o.wrap[2];
    a.max3 = o.rmax; a.max2 = -1; a.max1 = -1;
    c.nngb = (2 * refine[0] + 1) * (2 * refine[1] + 1) * (2 * refine[2] + 1);
    cats[k] = c;
    DevLattice L = tmpl;
    L.nx = nm[0]; L.ny = nm[1]; L.nz = nm[2];
    L.ncells = c.ncells;
    L.n = n;
    L.xmin = c.g.xmin; L.ymin = c.g.ymin; L.zmin = c.g.zmin;
    L.xbs = c.g.xbs; L.ybs = c.g.ybs; L.zbs = c.g.zbs;
    L.xinv = c.g.xinv; L.yinv = c.g.yinv; L.zinv = c.g.zinv;
    L.qinv = c.g.qinv;
    // set_quant_filter (3-D)
    const double rhi = sqrt(o.sqr_rmax) * (1.0 + 1e-9) * L.qinv + 2.0;
    const double rlo = sqrt(o.sqr_rmin) * (1.0 - 1e-9) * L.qinv - 2.0;
    const double hi = floor(rhi * rhi) + 1.0;
    const double lo = rlo > 0 ? floor(rlo * rlo) : 0.0;
    L.qrange = (unsigned)(hi - lo);
    L.qlo_neg = -(int)lo;
    L.qz = 0;
    lats[k] = L;
}

template <typename T>
__global__ void k_bg_bucket_count(int64_t n, const T* __restrict__ X, const T* __restrict__ Y,
                                  const T* __restrict__ Z, const BgCat* __restrict__ cat, int* __restrict__ bucket,
                                  int* __restrict__ counts, unsigned long long* __restrict__ oob)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) bucket_count_one(i, X, Y, Z, cat->g, bucket, counts, oob);
}

template <typename T>
__global__ void k_bg_scatter(int64_t n, const T* __restrict__ X, const T* __restrict__ Y, const T* __restrict__ Z,
                             const T* __restrict__ W, const BgCat* __restrict__ cat, const int* __restrict__ bucket,
                             const int* __restrict__ bstart, int* __restrict__ fill, double* __restrict__ x,
                             double* __restrict__ y, double* __restrict__ z, double* __restrict__ w,
                             double4* __restrict__ d4, uint2* __restrict__ q)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) scatter_one(i, X, Y, Z, W, cat->g, bucket, bstart, fill, x, y, z, w, d4, q);
}

__global__ void k_bg_cell_bounds(const BgCat* __restrict__ cat, const int* __restrict__ bstart,
                                 const double* __restrict__ x, const double* __restrict__ y,
                                 const double* __restrict__ z, double* __restrict__ bounds)
{
    const int64_t c = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    if (c < cat->ncells) cell_bounds_one(c, bstart, x, y, z, bounds);  // uniform per wave
}

// Candidates (cell, neighbour) over the captured capacity `cap`; those beyond
// the catalog's own ncells * nngb are not candidates.
__global__ void k_bg_cellpair_flag(const BgCat* __restrict__ cat, int64_t cap, const int* __restrict__ bs,
                                   const double* __restrict__ bd, int* __restrict__ flags)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= cap) return;
    const int nngb = cat->nngb;
    int f = 0;
    if (t < cat->ncells * nngb) {
        CellPairRec r = {};
        f = neighbour_added(cat->a, t / nngb, (int)(t % nngb), bs, bs, bd, bd, &r);
    }
    flags[t] = f;
}

// off[k + 1] = off[k] + cell pairs of catalog k (the scan's total at pos[cap]);
// past the list's capacity the total is clamped and flagged.
__global__ void k_bg_cp_offset(int k, const int* __restrict__ pos, int64_t cap, int* __restrict__ off, int pairs_cap,
                               int* __restrict__ fail)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int64_t v = (int64_t)off[k] + pos[cap];
    if (v > pairs_cap) {
        v = pairs_cap;
        atomicOr(fail, kBgFailPairs);
    }
    off[k + 1] = (int)v;
}

// The cell pairs of catalog k at off[k] + their scan position (within the list's
// capacity), and the pairs tested summed into tested[k].
__global__ void k_bg_cellpair_emit(const BgCat* __restrict__ cat, int64_t cap, const int* __restrict__ bs,
                                   const double* __restrict__ bd, const int* __restrict__ pos,
                                   const int* __restrict__ off, int k, int pairs_cap, CellPairRec* __restrict__ out,
                                   double* __restrict__ tested)
{
    __shared__ double part[4];
    const int nngb = cat->nngb;
    const int64_t total = min(cap, cat->ncells * nngb);
    const int base = off[k];
    double my = 0;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        CellPairRec r = {};
        if (neighbour_added(cat->a, t / nngb, (int)(t % nngb), bs, bs, bd, bd, &r)) {
            r.pad = k;
            const int64_t o = (int64_t)base + pos[t];
            if (o < pairs_cap) out[o] = r;
            const double n1 = (double)cell_count(bs, r.c1), n2 = (double)cell_count(bs, r.c2);
            my += r.same ? n1 * (n1 - 1) * 0.5 : n1 * n2;
        }
    }
    for (int o = 32; o > 0; o >>= 1) my += __shfl_xor(my, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = my;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double s = part[0] + part[1] + part[2] + part[3];
        if (s != 0) atomicAdd(&tested[k], s);
    }
}

// *nitems_eff = the work items, or 0 -- the pair kernel then does nothing --
// when they exceed the capacity (flagged) or an earlier step has failed.
__global__ void k_bg_items_clamp(const int* __restrict__ total, int items_cap, int* __restrict__ nitems_eff,
                                 int* __restrict__ fail)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int v = *total;
    if (v > items_cap) {
        v = 0;
        atomicOr(fail, kBgFailItems);
    }
    if (*fail) v = 0;
    *nitems_eff = v;
}

// fail |= bit when any particle fell outside its grid (counted by the bucket kernels)
__global__ void k_oob_flag(const unsigned long long* __restrict__ oob, int* __restrict__ fail, int bit)
{
    if (threadIdx.x == 0 && blockIdx.x == 0 && *oob) atomicOr(fail, bit);
}

}  // namespace cfamd
