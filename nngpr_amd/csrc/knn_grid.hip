// Exact k-nearest-neighbour search on uniform grids (d <= 3, k <= 64), gfx950.
//
// Same answer as the brute-force scans: for every query the k refs with the
// smallest (key, ref index), key = the FP64 squared distance of the scaled
// coordinates with the reference's FMA order (cuda_kernel_accelerators.py:711-742),
// restricted to refs below the Vecchia limit, padded with -1.  Refs with a
// non-finite coordinate are never neighbours (their key is never below +inf), and
// a query with a non-finite coordinate gets -1 only, as in the scans.
//
// Grid build: finite bounding box -> cell size h for about kPointsPerCell refs per
// cell -> counting sort of the refs by cell (count, exclusive scan, atomic
// scatter; the order inside a cell is arbitrary and does not matter because the
// top-k is ordered by (key, index)).  A few far refs would stretch the box and put
// most refs into a few cells: when at most kMaxOutside refs fall outside an inner
// box sampled from the refs (second smallest / largest of 128 sampled refs per
// feature, widened by half their distance) and that box holds under 1/8 of the
// volume, the grid spans the refs inside it and the others go to an "outside"
// bucket that every search keys before its first ring.  The refs in the cells then
// still lie inside their cells, so every distance bound of the search holds for any
// query, inside or outside the box; queries outside it (which would walk many
// rings) are searched one per wave.
//
// Search, one lane per query, 64 queries per wave: the queries are sorted by tiles
// of cells, so a wave's lanes search neighbouring space.  Each lane walks rings of
// cells around its own cell (R = 0, 1, ...), loading the refs of each cell run
// clipped to its current search ball.  A lane appends the refs below its
// admission threshold to its column of an LDS buffer; full buffers are merged
// into the lanes' sorted register lists in lock-step.  After a ring, a lane is
// done when its k-th key is below the squared distance from its query to the
// nearest unvisited cell (minus a slack covering every rounding of the cell
// assignment and of the key); the wave stops when all its lanes are done.  (A
// wave-shared walk, every lane testing every ref of the union of the lanes'
// regions through scalar loads, was measured 3x slower on clustered data: a
// group of 64 queries grown by the neighbour radius holds ~20x the pairs the
// lanes need, at about the same VALU cost per pair.)
//
// Searches of at most kWaveSearchMaxQueries queries run one query per wave
// instead: the same rings, stop test and run clipping, with the wave's lanes
// keying 64 refs of the ring's runs at a time and the 64 smallest pairs kept
// sorted across the lanes.  A lane-per-query wave's latency (its lanes' ring walks
// one after the other, ~0.5 ms on the road-mixture data) would otherwise bound
// every search below ~10^5 queries; on the full 434,874-row Vecchia search a wave
// per query costs 2.7x the lane kernel.
//
// Vecchia cut (row q searches refs [0, q + cut)): rows are split into bands of
// doubling limit [L, 2L).  A band's grid holds refs [0, 2L) only, so at least half
// of the refs a band streams are admissible for each of its rows.  Rows with
// limits below the first band are scanned brute force by the caller.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>

#include "common.h"
#include "knn.h"
#include "knn_device.h"

namespace nngp {

namespace {

constexpr int kPointsPerCell = 2;
constexpr int kScanBlock = 1024;   // items per block of the exclusive scan (256 x 4)
constexpr int kCoopWaves = 4;      // waves per block of the cost pass
// Waves per block of the lane-per-query search: one, so that a finished wave frees
// its slot at once (with 4, a block's slots stayed held until its slowest wave
// ended: 66 % of the wave slots busy over the 434,874-row Vecchia search)
constexpr int kLaneWaves = 1;
constexpr int kCoopSlots = 16;     // LDS append slots per lane (lists of 16)
constexpr int kCoopSlotsWide = 24; // LDS append slots per lane (lists of 32 and 64)
// Searches whose full grid (the top band's) is at most this many cells wide in x
// take the instantiation that walks only the boundary rows of rings wider than the
// x extent.
constexpr int kNarrowCells = 16;
// Searches of at most this many queries run one query per wave.  A lane-per-query
// wave walks its 64 queries' rings one after the other: 0.61-0.71 ms on the
// road-mixture data for any query search of 1,024 to 87,000 queries against its
// 434,874 refs (whole call: query binning and launches included).  A wave per query
// keys 64 refs per step and ends in tens of microseconds, so its time grows with the
// count: 0.19 ms (1,024 queries), 0.26 ms (16,384), 0.38 ms (43,487), 0.49 ms
// (65,536), 0.58 ms (86,974); the two cross near 110,000.  On the full Vecchia
// self-search a wave per query costs 2.7x the lane kernel.
constexpr int64_t kWaveSearchMaxQueries = 100000;

__device__ __forceinline__ bool point_finite(const double* u, int d) {
    bool fin = true;
    for (int f = 0; f < d; ++f) fin = fin && __builtin_isfinite(u[f]);
    return fin;
}

constexpr int kBBoxBlocks = 256;  // blocks of the bounding-box pass (one partial each)
// Box memory (long long words): [0, 6) the bounding box of the finite refs as
// order-preserving keys, [8, 14) the candidate inner box, 14 the count of finite
// refs outside it, [16, 22) the bounding box of the refs inside it, block partials
// of the bounding-box pass from 24.
constexpr int kBoxInner = 8;
constexpr int kBoxOutCount = 14;
constexpr int kBoxTight = 16;
constexpr int kBoxPartials = 24;
constexpr int kBoxSample = 128;   // refs sampled for the inner box
constexpr int kMaxOutside = 256;  // most refs a narrowed grid keeps out of its cells

__device__ __forceinline__ void bbox_wave_reduce(long long (&lo)[3], long long (&hi)[3]) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const long long ol = __shfl_xor(lo[f], off, 64);
            const long long oh = __shfl_xor(hi[f], off, 64);
            lo[f] = ol < lo[f] ? ol : lo[f];
            hi[f] = oh > hi[f] ? oh : hi[f];
        }
}

// Reduces the block's 4 waves (after bbox_wave_reduce) into thread 0's lo / hi.
__device__ __forceinline__ void bbox_block_reduce(long long (&lo)[3], long long (&hi)[3]) {
    __shared__ long long wl[3][4], wh[3][4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            wl[f][w] = lo[f];
            wh[f][w] = hi[f];
        }
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int f = 0; f < 3; ++f)
            for (int v = 1; v < 4; ++v) {
                lo[f] = wl[f][v] < lo[f] ? wl[f][v] : lo[f];
                hi[f] = wh[f][v] > hi[f] ? wh[f][v] : hi[f];
            }
}

// Bounding box of the points whose coordinates are all finite, as order-preserving
// integer keys: one partial (lo, hi per feature) per block, reduced by
// grid_bbox_final_kernel (no same-address atomics).
__global__ __launch_bounds__(256) void grid_bbox_kernel(const double* u64, int64_t n, int d,
                                                        long long* part) {
    long long lo[3], hi[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        lo[f] = 0x7fffffffffffffffLL;
        hi[f] = (long long)0x8000000000000000ULL;
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double* u = u64 + i * d;
        if (!point_finite(u, d)) continue;
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            if (f < d) {
                const long long k = dbl_key(u[f]);
                lo[f] = k < lo[f] ? k : lo[f];
                hi[f] = k > hi[f] ? k : hi[f];
            }
        }
    }
    bbox_wave_reduce(lo, hi);
    bbox_block_reduce(lo, hi);
    if (threadIdx.x == 0)
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            part[blockIdx.x * 6 + 2 * f] = lo[f];
            part[blockIdx.x * 6 + 2 * f + 1] = hi[f];
        }
}

// bb = the reduction of nb block partials (empty: lo = max, hi = min).
__global__ __launch_bounds__(256) void grid_bbox_final_kernel(const long long* part, int nb,
                                                              long long* bb) {
    long long lo[3], hi[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        lo[f] = 0x7fffffffffffffffLL;
        hi[f] = (long long)0x8000000000000000ULL;
    }
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const long long l = part[b * 6 + 2 * f], h = part[b * 6 + 2 * f + 1];
            lo[f] = l < lo[f] ? l : lo[f];
            hi[f] = h > hi[f] ? h : hi[f];
        }
    bbox_wave_reduce(lo, hi);
    bbox_block_reduce(lo, hi);
    if (threadIdx.x == 0)
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            bb[2 * f] = lo[f];
            bb[2 * f + 1] = hi[f];
        }
}

// Candidate inner box: per feature, the second smallest and second largest of
// kBoxSample finite refs sampled at even strides of [0, n) (rank by (value, sample)),
// widened on each side by half their distance; unbounded with fewer than 64 finite
// samples.  Clears the outside count and the inside box.  One block; n > 0.
__global__ __launch_bounds__(kBoxSample) void grid_box_sample_kernel(const double* u64, int64_t n,
                                                                     int d, long long* bb) {
    __shared__ double v[3][kBoxSample];
    __shared__ double qv[3][2];
    const int t = threadIdx.x;
    int64_t i = (int64_t)(((double)t + 0.5) * (double)n / (double)kBoxSample);
    i = i < n ? i : n - 1;
    const double* u = u64 + i * d;
    const bool fin = point_finite(u, d);
    double x[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        x[f] = (f < d && fin) ? u[f] : __builtin_inf();
        v[f][t] = x[f];
    }
    const int nf = __syncthreads_count(fin);
    if (t == 0) {
        bb[kBoxOutCount] = 0;
        for (int f = 0; f < 3; ++f) {
            bb[kBoxTight + 2 * f] = 0x7fffffffffffffffLL;
            bb[kBoxTight + 2 * f + 1] = (long long)0x8000000000000000ULL;
        }
    }
    // the three features' ranks together, 8 samples per step (independent loads)
    int rank[3] = {0, 0, 0};
#pragma unroll 8
    for (int j = 0; j < kBoxSample; ++j) {
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const double y = v[f][j];
            rank[f] += (int)((y < x[f]) | ((y == x[f]) & (j < t)));
        }
    }
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        if (fin && rank[f] == 1) qv[f][0] = x[f];
        if (fin && rank[f] == nf - 2) qv[f][1] = x[f];
    }
    __syncthreads();
    if (t == 0) {
        for (int f = 0; f < 3; ++f) {
            double lo = -__builtin_inf(), hi = __builtin_inf();
            if (f < d && nf >= 64) {
                const double pad = 0.5 * (qv[f][1] - qv[f][0]);
                lo = qv[f][0] - pad;
                hi = qv[f][1] + pad;
            }
            bb[kBoxInner + 2 * f] = dbl_key(lo);
            bb[kBoxInner + 2 * f + 1] = dbl_key(hi);
        }
    }
}

// Volume of the candidate inner box over the bounding box's (features of zero
// extent left out); the inner box clipped to the bounding box in lo / ext.
__device__ __forceinline__ double inner_box_ratio(const long long* bb, int d, double (&lo)[3],
                                                  double (&ext)[3]) {
    double ratio = 1.0;
    for (int f = 0; f < d; ++f) {
        const double l = key_dbl(bb[2 * f]), h = key_dbl(bb[2 * f + 1]);
        const double il = fmax(key_dbl(bb[kBoxInner + 2 * f]), l);
        const double ih = fmin(key_dbl(bb[kBoxInner + 2 * f + 1]), h);
        lo[f] = il;
        ext[f] = fmax(ih - il, 0.0);
        const double e = h - l;
        if (e > 0.0) ratio *= ext[f] / e;
    }
    return ratio;
}

// A narrowed grid must shrink the box's volume at least this much.
constexpr double kNarrowVolume = 0.125;

// Finite refs outside the candidate inner box, and the bounding box of those inside
// it (only when that box is small enough to be used): block partials added to
// bb[kBoxOutCount], bb[kBoxTight ...].
__global__ __launch_bounds__(256) void grid_outside_count_kernel(const double* u64, int64_t n,
                                                                 int d, long long* bb) {
    double lo[3], ext[3];
    if (!(inner_box_ratio(bb, d, lo, ext) < kNarrowVolume)) return;
    double il[3], ih[3];
    long long tl[3], th[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        il[f] = key_dbl(bb[kBoxInner + 2 * f]);
        ih[f] = key_dbl(bb[kBoxInner + 2 * f + 1]);
        tl[f] = 0x7fffffffffffffffLL;
        th[f] = (long long)0x8000000000000000ULL;
    }
    unsigned long long c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double* u = u64 + i * d;
        if (!point_finite(u, d)) continue;
        bool out = false;
        for (int f = 0; f < d; ++f) out = out || u[f] < il[f] || u[f] > ih[f];
        if (out) {
            ++c;
            continue;
        }
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            if (f < d) {
                const long long k = dbl_key(u[f]);
                tl[f] = k < tl[f] ? k : tl[f];
                th[f] = k > th[f] ? k : th[f];
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    __shared__ unsigned long long wc[4];
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = c;
    bbox_wave_reduce(tl, th);
    bbox_block_reduce(tl, th);  // (its barriers also publish wc)
    if (threadIdx.x == 0) {
        const unsigned long long tot = wc[0] + wc[1] + wc[2] + wc[3];
        if (tot) atomicAdd(reinterpret_cast<unsigned long long*>(bb + kBoxOutCount), tot);
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            atomicMin(bb + kBoxTight + 2 * f, tl[f]);
            atomicMax(bb + kBoxTight + 2 * f + 1, th[f]);
        }
    }
}

// Cell size and grid dimensions from the bounding box, for a grid over n refs.  The cell size targets n / kPointsPerCell cells over the box; it
// grows until the grid fits in `cap` cells.  Query tiles are 4 cells wide along
// features with at least 8 cells; the tiled bucket count must fit in capq.
__device__ void setup_grid(const long long* bb, int d, int64_t n, int cap, int capq,
                           GridParams* gp) {
    double lo[3] = {0.0, 0.0, 0.0}, ext[3] = {0.0, 0.0, 0.0};
    bool any = true;
    for (int f = 0; f < d; ++f) {
        const double l = key_dbl(bb[2 * f]), h = key_dbl(bb[2 * f + 1]);
        if (!(l <= h)) any = false;
        lo[f] = l;
        ext[f] = h - l;
    }
    if (!any) {  // no finite ref: one empty cell
        for (int f = 0; f < 3; ++f) {
            lo[f] = 0.0;
            ext[f] = 0.0;
        }
    }
    // A few far refs stretch the box and coarsen every cell: when at most
    // kMaxOutside finite refs fall outside the sampled inner box and it holds under
    // kNarrowVolume of the box's volume, the grid covers the inner box only and
    // those refs are kept out of the cells (keyed by every search instead).
    for (int f = 0; f < 3; ++f) {
        gp->in_lo[f] = -__builtin_inf();
        gp->in_hi[f] = __builtin_inf();
    }
    if (any) {
        const long long n_out = bb[kBoxOutCount];
        double nlo[3] = {0.0, 0.0, 0.0}, next[3] = {0.0, 0.0, 0.0};
        bool tight = true;  // some ref lies inside the inner box
        for (int f = 0; f < d; ++f) tight = tight && bb[kBoxTight + 2 * f] <= bb[kBoxTight + 2 * f + 1];
        if (n_out > 0 && n_out <= kMaxOutside && tight &&
            inner_box_ratio(bb, d, nlo, next) < kNarrowVolume) {
            // the grid spans the refs inside the inner box; the inner box decides
            // which points stay out
            for (int f = 0; f < d; ++f) {
                lo[f] = key_dbl(bb[kBoxTight + 2 * f]);
                ext[f] = key_dbl(bb[kBoxTight + 2 * f + 1]) - lo[f];
                gp->in_lo[f] = key_dbl(bb[kBoxInner + 2 * f]);
                gp->in_hi[f] = key_dbl(bb[kBoxInner + 2 * f + 1]);
            }
        }
    }
    for (int f = 0; f < 3; ++f)
        if (!__builtin_isfinite(ext[f])) ext[f] = 1.0e300;
    const double target = fmax(1.0, (double)n / kPointsPerCell);
    int deff = 0;
    double lv = 0.0;
    for (int f = 0; f < d; ++f)
        if (ext[f] > 0.0) {
            ++deff;
            lv += log(ext[f]);
        }
    double h = deff ? exp((lv - log(target)) / deff) : 1.0;
    if (!(h > 1.0e-290) || !__builtin_isfinite(h)) h = deff ? 1.0e-290 : 1.0;
    int g[3] = {1, 1, 1};
    for (int it = 0; it < 4000; ++it) {
        double prod = 1.0;
        for (int f = 0; f < 3; ++f) {
            g[f] = 1;
            if (f < d && ext[f] > 0.0) {
                const double c = ceil(ext[f] / h);
                g[f] = c > (double)cap ? cap + 1 : (c < 1.0 ? 1 : (int)c);
            }
            prod *= (double)g[f];
        }
        if (prod <= (double)cap) break;
        h *= 1.25;
    }
    double mag = 0.0;
    int64_t nqb = 1;
    for (int f = 0; f < 3; ++f) {
        gp->lo[f] = lo[f];
        gp->g[f] = g[f];
        mag = fmax(mag, fabs(lo[f]) + (double)g[f] * h);
        const int t = g[f] >= 8 ? 4 : 1;
        gp->t[f] = t;
        gp->tg[f] = (g[f] + t - 1) / t;
        nqb *= (int64_t)gp->tg[f] * t;
    }
    if (nqb > capq) {  // untiled (cell order)
        nqb = 1;
        for (int f = 0; f < 3; ++f) {
            gp->t[f] = 1;
            gp->tg[f] = g[f];
            nqb *= g[f];
        }
    }
    gp->n_qb = (int)nqb;
    gp->h = h;
    gp->inv_h = 1.0 / h;
    gp->tol = 1.0e-9 * mag + 1.0e-300;
    gp->n_cells = g[0] * g[1] * g[2];
}

__device__ __forceinline__ int cell_coord(double u, double lo, double inv_h, int g) {
    double c = floor((u - lo) * inv_h);
    c = c < 0.0 ? 0.0 : c;
    c = c > (double)(g - 1) ? (double)(g - 1) : c;
    return (int)c;
}

// ---------------------------------------------------------------------------
// Counting sorts of points into grid buckets, several grids per launch
// ---------------------------------------------------------------------------
// One counting sort: points = rows [row0, row0 + n) of u64, into nb buckets of
// grid gp (the last bucket for points with a non-finite coordinate).  Refs
// (tiled == 0): the bucket is the cell in row-major order, and the occupancy
// words of the grid are filled.  Queries (tiled != 0): the cell in tile order.
// t_off / f_off / s_off / w_off: first block of the job in the point, bucket,
// scan and occupancy-word launches.
struct SortJob {
    GridParams* gp;
    int32_t* cnt;
    int32_t* start;
    int32_t* fill;
    int32_t* sums;
    int32_t* bucket_of;
    GridPt* pts;
    uint64_t* rowmask;
    int64_t n;
    int64_t row0;
    int32_t t_off, f_off, s_off, w_off;
    int32_t nb;
    int32_t tiled;
    int32_t cap;
    int32_t capq;
};

constexpr int kJobsPerLaunch = 15;  // keeps the argument block small

struct SortArgs {
    const double* u64;
    const long long* bbox;
    int d;
    int n_jobs;
    SortJob job[kJobsPerLaunch];
};

// Job of block b given each job's first block (field `off` of the jobs).
template <int32_t SortJob::*OFF>
__device__ __forceinline__ int job_of_block(const SortArgs& a, int64_t b) {
    int j = 0;
    for (int t = 1; t < a.n_jobs; ++t)
        if ((int64_t)(a.job[t].*OFF) <= b) j = t;
    return j;
}

__host__ __device__ inline int64_t scan_blocks(int64_t n) { return (n + kScanBlock - 1) / kScanBlock; }

// Grid parameters of every ref job (one thread each).
__global__ void sort_setup_kernel(SortArgs a) {
    const int j = threadIdx.x;
    if (j < a.n_jobs) {
        const SortJob& J = a.job[j];
        setup_grid(a.bbox, a.d, J.n, J.cap, J.capq, J.gp);
    }
}

// Zeroes the counts and fill cursors (nb + 1 entries per job).
__global__ __launch_bounds__(256) void sort_clear_kernel(SortArgs a) {
    const SortJob& J = a.job[job_of_block<&SortJob::f_off>(a, blockIdx.x)];
    const int64_t i = ((int64_t)blockIdx.x - J.f_off) * 256 + threadIdx.x;
    if (i <= J.nb) {
        J.cnt[i] = 0;
        J.fill[i] = 0;
    }
}

__global__ __launch_bounds__(256) void sort_count_kernel(SortArgs a) {
    const SortJob& J = a.job[job_of_block<&SortJob::t_off>(a, blockIdx.x)];
    const int64_t i = ((int64_t)blockIdx.x - J.t_off) * 256 + threadIdx.x;
    if (i >= J.n) return;
    const int d = a.d;
    const double* u = a.u64 + (J.row0 + i) * d;
    const GridParams* gp = J.gp;
    int b = J.nb - 1;  // non-finite
    bool out = false;
    if (point_finite(u, d))
        for (int f = 0; f < d; ++f) out = out || u[f] < gp->in_lo[f] || u[f] > gp->in_hi[f];
    if (out) {
        b = J.nb - 2;  // outside the grid's box
    } else if (point_finite(u, d)) {
        const double ih = gp->inv_h;
        int c[3] = {0, 0, 0};
#pragma unroll
        for (int f = 0; f < 3; ++f)
            if (f < d) c[f] = cell_coord(u[f], gp->lo[f], ih, gp->g[f]);
        if (J.tiled) {
            const int t0 = gp->t[0], t1 = gp->t[1], t2 = gp->t[2];
            const int tile = c[0] / t0 + gp->tg[0] * (c[1] / t1 + gp->tg[1] * (c[2] / t2));
            const int local = c[0] % t0 + t0 * (c[1] % t1 + t1 * (c[2] % t2));
            b = tile * (t0 * t1 * t2) + local;
        } else {
            b = c[0] + gp->g[0] * (c[1] + gp->g[1] * c[2]);
        }
    }
    J.bucket_of[i] = b;
    atomicAdd(&J.cnt[b], 1);
}

// Point i (row row0 + i) to its place in bucket order.
__global__ __launch_bounds__(256) void sort_scatter_kernel(SortArgs a) {
    const SortJob& J = a.job[job_of_block<&SortJob::t_off>(a, blockIdx.x)];
    const int64_t i = ((int64_t)blockIdx.x - J.t_off) * 256 + threadIdx.x;
    if (i >= J.n) return;
    const int d = a.d;
    const int c = J.bucket_of[i];
    const int pos = J.start[c] + atomicAdd(&J.fill[c], 1);
    GridPt p;
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        p.u[f] = 0.0;
        if (f < d) p.u[f] = a.u64[(J.row0 + i) * d + f];
    }
    p.idx = J.row0 + i;
    J.pts[pos] = p;
}

// Exclusive prefix sum of the 256 threads' values; *total gets the block sum.
__device__ __forceinline__ int32_t block_excl_scan_256(int32_t v, int32_t* total) {
    __shared__ int32_t wsum[4];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    int32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int32_t woff = 0;
    for (int i = 0; i < w; ++i) woff += wsum[i];
    *total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    return woff + x - v;
}

// Exclusive scan of cnt into start (nb + 1 entries per job): per-block scans,
// then each job's block sums (one block per job), then the block offsets.
__global__ __launch_bounds__(256) void sort_scan_local_kernel(SortArgs a) {
    const SortJob& J = a.job[job_of_block<&SortJob::s_off>(a, blockIdx.x)];
    const int64_t blk = (int64_t)blockIdx.x - J.s_off;
    const int64_t n = (int64_t)J.nb + 1;
    const int64_t base = blk * kScanBlock + threadIdx.x * 4;
    int32_t v[4];
    int32_t s = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        v[i] = base + i < n ? J.cnt[base + i] : 0;
        s += v[i];
    }
    int32_t tot;
    int32_t run = block_excl_scan_256(s, &tot);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (base + i < n) J.start[base + i] = run;
        run += v[i];
    }
    if (threadIdx.x == 0) J.sums[blk] = tot;
}

__global__ __launch_bounds__(256) void sort_scan_sums_kernel(SortArgs a) {
    const SortJob& J = a.job[blockIdx.x];
    const int nb = (int)scan_blocks((int64_t)J.nb + 1);
    int32_t carry = 0;
    for (int c0 = 0; c0 < nb; c0 += 256) {
        const int i = c0 + threadIdx.x;
        const int32_t v = i < nb ? J.sums[i] : 0;
        int32_t tot;
        const int32_t ex = block_excl_scan_256(v, &tot);
        if (i < nb) J.sums[i] = carry + ex;
        carry += tot;
    }
}

__global__ __launch_bounds__(256) void sort_scan_add_kernel(SortArgs a) {
    const SortJob& J = a.job[job_of_block<&SortJob::s_off>(a, blockIdx.x)];
    const int64_t blk = (int64_t)blockIdx.x - J.s_off;
    const int64_t n = (int64_t)J.nb + 1;
    const int64_t base = blk * kScanBlock;
    const int32_t off = J.sums[blk];
    for (int i = threadIdx.x; i < kScanBlock; i += 256)
        if (base + i < n) J.start[base + i] += off;
}

// Occupancy words of the ref grids: one thread per word (grid dimensions read
// on the device; threads past the grid's last word return).
__global__ __launch_bounds__(256) void sort_rowmask_kernel(SortArgs a) {
    const SortJob& J = a.job[job_of_block<&SortJob::w_off>(a, blockIdx.x)];
    const int64_t w = ((int64_t)blockIdx.x - J.w_off) * 256 + threadIdx.x;
    const GridParams* gp = J.gp;
    const int g0 = gp->g[0];
    const int wpr = (g0 + 63) >> 6;
    const int64_t rows = (int64_t)gp->g[1] * gp->g[2];
    if (w >= rows * wpr) return;
    const int64_t row = w / wpr;
    const int x0 = (int)(w % wpr) * 64;
    const int xn = min(64, g0 - x0);
    const int32_t* st = J.start + row * g0 + x0;
    uint64_t m = 0;
    for (int b = 0; b < xn; ++b)
        if (st[b + 1] > st[b]) m |= 1ull << b;
    J.rowmask[w] = m;
}

// Occupancy words of a grid of at most `cap` cells: every row holds
// ceil(g_0 / 64) <= g_0 / 64 + 1 words, over at most `cap` rows.
size_t rowmask_words_max(int cap) { return (size_t)cap + (size_t)cap / 64 + 2; }

// Runs the counting sorts of `jobs` (refs: grid setup first and occupancy words
// after), kJobsPerLaunch jobs per launch.
void run_sorts(const double* u64, const long long* bbox, int d, const SortJob* jobs, int n_jobs,
               bool refs, hipStream_t st) {
    for (int j0 = 0; j0 < n_jobs; j0 += kJobsPerLaunch) {
        SortArgs a{};
        a.u64 = u64;
        a.bbox = bbox;
        a.d = d;
        a.n_jobs = std::min(kJobsPerLaunch, n_jobs - j0);
        int64_t tb = 0, fb = 0, sb = 0, wb = 0;
        for (int t = 0; t < a.n_jobs; ++t) {
            SortJob J = jobs[j0 + t];
            J.t_off = (int32_t)tb;
            J.f_off = (int32_t)fb;
            J.s_off = (int32_t)sb;
            J.w_off = (int32_t)wb;
            tb += (J.n + 255) / 256;
            fb += ((int64_t)J.nb + 1 + 255) / 256;
            sb += scan_blocks((int64_t)J.nb + 1);
            if (refs) wb += (int64_t)(rowmask_words_max(J.cap) + 255) / 256;
            a.job[t] = J;
        }
        if (refs) hipLaunchKernelGGL(sort_setup_kernel, dim3(1), dim3(64), 0, st, a);
        hipLaunchKernelGGL(sort_clear_kernel, dim3((unsigned)fb), dim3(256), 0, st, a);
        if (tb > 0) hipLaunchKernelGGL(sort_count_kernel, dim3((unsigned)tb), dim3(256), 0, st, a);
        hipLaunchKernelGGL(sort_scan_local_kernel, dim3((unsigned)sb), dim3(256), 0, st, a);
        hipLaunchKernelGGL(sort_scan_sums_kernel, dim3((unsigned)a.n_jobs), dim3(256), 0, st, a);
        hipLaunchKernelGGL(sort_scan_add_kernel, dim3((unsigned)sb), dim3(256), 0, st, a);
        if (tb > 0)
            hipLaunchKernelGGL(sort_scatter_kernel, dim3((unsigned)tb), dim3(256), 0, st, a);
        if (refs) hipLaunchKernelGGL(sort_rowmask_kernel, dim3((unsigned)wb), dim3(256), 0, st, a);
    }
}

// ---------------------------------------------------------------------------
// Search kernel
// ---------------------------------------------------------------------------
// One band of a search: its grid and its queries (tile order); its waves are
// [wave_off, wave_off + ceil(n_q / 64)).
struct BandDesc {
    const GridParams* gp;
    const int32_t* cell_start;  // ref cells, row-major
    const uint64_t* rowmask;    // occupancy bits of the ref cells
    const GridPt* rpts;         // refs in cell order
    const GridPt* qpts;         // the band's queries in tile order
    const int32_t* qstart;      // the band's query bucket starts
    int32_t qfar;               // bucket of its queries outside the grid's box
    int32_t ref_out;            // bucket of the grid's refs outside its box
    int64_t n_q;
    int64_t wave_off;
};

struct SearchArgs {
    int n_bands;
    BandDesc band[kMaxGridBands + 1];
    int64_t n_waves;          // waves over all bands
    const int32_t* perm;      // launch slot -> wave (most expensive first)
    int64_t n_ref;
    int has_cut;
    int64_t cut;
    int k;
    int32_t* out_idx;
    // when non-null, every wave of lane_knn_kernel adds the number of (query, ref)
    // pairs its lanes examined (the measured work of the search).  The wave-per-query
    // kernels do not count: one same-address atomic per query (43,487 of them for the
    // road-mixture test set) cost ~0.2 ms of contention whenever timing was on.
    unsigned long long* pair_count;
    // queries lane_knn_kernel could not order exactly (equal keys at the k-th place
    // or out of index order): (band, position) pairs and their count, searched again
    // by wave_far_kernel; the count is cleared by wave_sort_kernel
    int32_t* redo;
    int32_t* redo_n;
    // when non-null, wave_far_kernel adds the number of queued queries to it
    int64_t* rows_recomputed;
};

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x3 __attribute__((ext_vector_type(3)));

// Sorted list of K entries, the first K - k of them fixed at (-inf, -2): the
// last entry is then the k-th smallest admitted pair.
template <int K>
__device__ __forceinline__ void init_list(double (&keys)[K], int (&ids)[K], int pad) {
#pragma unroll
    for (int t = 0; t < K; ++t) {
        keys[t] = t < pad ? -__builtin_inf() : __builtin_inf();
        ids[t] = t < pad ? -2 : -1;
    }
}

template <int K>
__device__ __forceinline__ void store_list(const int (&ids)[K], int pad, int32_t* o) {
#pragma unroll
    for (int t = 0; t < K; ++t)
        if (t >= pad) o[t - pad] = ids[t];
}

// Merges a lane's buffered candidates in slots [off, min(off + S, cnt)) (LDS
// column `lane`) into its sorted list.
template <int K, int S>
__device__ __forceinline__ void merge_slots(double (&keys)[K], int (&ids)[K],
                                            const double (*bk)[64], const int (*bi)[64], int lane,
                                            int off, int cnt) {
    double ck[S];
    int ci[S];
#pragma unroll
    for (int b = 0; b < S; ++b) {
        const bool v = off + b < cnt;
        const double x = bk[off + b][lane];
        const int xi = bi[off + b][lane];
        ck[b] = v ? x : __builtin_inf();
        ci[b] = v ? xi : -1;
    }
    pair_merge<K, S>(keys, ids, ck, ci);
}

// As merge_slots, with the key-only network (key_merge); dropped_min collects the
// smallest key dropped.
template <int K, int S>
__device__ __forceinline__ void merge_slots_keys(double (&keys)[K], int (&ids)[K],
                                                 const double (*bk)[64], const int (*bi)[64],
                                                 int lane, int off, int cnt, double& dropped_min) {
    double ck[S];
    int ci[S];
#pragma unroll
    for (int b = 0; b < S; ++b) {
        const bool v = off + b < cnt;
        const double x = bk[off + b][lane];
        const int xi = bi[off + b][lane];
        ck[b] = v ? x : __builtin_inf();
        ci[b] = v ? xi : -1;
    }
    key_merge<K, S>(keys, ids, ck, ci, dropped_min);
}

// One lane per query, 64 neighbouring queries per wave.  A lane walks the rings
// of cells around its own cell (R = 0, 1, ...; a ring's rows (y, z) one after the
// other: a new row's whole x run, an old row's two new end cells) and appends the
// refs below its admission threshold to its LDS column; when any lane's column is
// full, or a lane finished a ring, the wave merges every column into the lanes'
// sorted register lists in lock-step (the single merge site of the loop).  After a
// ring the lane stops once its k-th key is below the squared distance from its
// query to the nearest cell outside the visited block.
// K = 32 and 64 order their lists by key alone (key_merge, admission below the k-th
// key): a lane that meets a key equal to its k-th (admitted or dropped) or ends
// with equal keys out of index order is queued for the exact wave-per-query search
// (wave_far_kernel), which rewrites its row.
// Two instantiations share each launch set: NARROW = false serves searches whose
// full grid is wider than kNarrowCells in x, NARROW = true the others (every wave of
// the instantiation that does not serve the search returns at once).
template <int D, int K, bool NARROW>
__global__ __launch_bounds__(64 * kLaneWaves)
// (K = 64: one wave per SIMD, so that the lists live in VGPRs and AGPRs instead of
// spilling to scratch; 16 % faster on the Vecchia search, 40 % on a query search.)
__attribute__((amdgpu_waves_per_eu(K <= 16 ? 3 : (K <= 32 ? 2 : 1))))
void lane_knn_kernel(SearchArgs a) {
    constexpr int B = K == 16 ? kCoopSlots : kCoopSlotsWide;
    constexpr bool KEYS = K != 16;  // key-only list order (see above)
    // refs per step of a lane, their loads in flight together (with (key, index)
    // merges 4: 2.14 ms, 8: 1.97 ms, 12: 2.06 ms, 16: 2.62 ms on the 434,874-row
    // Vecchia search; with the key-only merges of lists of 32 and 64, whole kNN call
    // 6: 2.08-2.10 ms, 8: 2.02-2.04, 12: 1.97-2.00); B - CH is the merge trigger
    constexpr int CH = K == 16 ? 8 : 12;
    __shared__ double sk[kLaneWaves][B][64];
    __shared__ int si[kLaneWaves][B][64];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const int lane = threadIdx.x & 63;
    const int64_t slot = (int64_t)blockIdx.x * kLaneWaves + w;
    // slots past the last wave map past every band (no query)
    const int64_t wg = slot < a.n_waves ? (int64_t)__builtin_amdgcn_readfirstlane(a.perm[slot])
                                        : a.n_waves;
    int bj = 0;
    for (int t = 1; t < a.n_bands; ++t)
        if (a.band[t].wave_off <= wg) bj = t;
    const BandDesc& bd0 = a.band[bj];
    if ((__builtin_amdgcn_readfirstlane(a.band[a.n_bands - 1].gp->g[0]) <= kNarrowCells) !=
        NARROW)
        return;
    const int64_t j = (wg - bd0.wave_off) * 64 + lane;
    // queries outside the grid's box are searched one per wave (wave_far_kernel)
    const int32_t far0 = bd0.qstart[bd0.qfar], far1 = bd0.qstart[bd0.qfar + 1];
    const bool jv = j < bd0.n_q && !(j >= far0 && j < far1);
    double qu[3] = {0.0, 0.0, 0.0};
    int64_t q = 0;
    if (jv) {
        const GridPt qp = bd0.qpts[j];
#pragma unroll
        for (int f = 0; f < 3; ++f) qu[f] = qp.u[f];
        q = qp.idx;
    }
    int64_t lim = 0;
    if (jv) {
        lim = a.n_ref;
        if (a.has_cut) {
            lim = q + a.cut;
            lim = lim < 0 ? 0 : (lim > a.n_ref ? a.n_ref : lim);
        }
    }
    bool fin = jv;
#pragma unroll
    for (int f = 0; f < D; ++f) fin = fin && __builtin_isfinite(qu[f]);
    const bool active = fin && lim > 0;
    const GridParams& P = *bd0.gp;
    int g[3], c[3];
    double lo[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        g[f] = P.g[f];
        lo[f] = P.lo[f];
        c[f] = (f < D && active) ? cell_coord(qu[f], lo[f], P.inv_h, g[f]) : 0;
    }
    // K - k leading placeholder entries; laundered through a register so that
    // the per-entry masks are rebuilt here and at the store, not kept live
    int pad = K - a.k;
    asm volatile("" : "+v"(pad));
    double keys[K];
    int ids[K];
    init_list<K>(keys, ids, pad);

    uint32_t examined = 0;  // refs this lane keyed
    bool tie = false;       // (KEYS: an equal key met at the k-th place)
    if (__builtin_amdgcn_ballot_w64(active)) {
        const double h = P.h, tol = P.tol;
        const int32_t* cs = bd0.cell_start;
        const GridPt* rpts = bd0.rpts;
        const int32_t lim32 = (int32_t)lim;
        double (*bk)[64] = sk[w];
        int (*bi)[64] = si[w];
        // admission threshold pair; inactive lanes admit nothing
        double tk = active ? __builtin_inf() : -__builtin_inf();
        int ti = -1;  // (K = 16: index of the threshold pair)
        // squared search radius with slack, (sqrt(tk) + tol)^2; +inf until k refs
        double rr = __builtin_inf();
        const double inv_h = P.inv_h;
        int cnt = 0;
        // per-lane walk: ring R, segment sg of the ring (two per row), refs [p, e)
        int R = 0, sg = 0, nseg = 0, ny = 1;
        int x0 = 0, x1 = 0, y0 = 0, z0 = 0;
        // The rows of ring R are every row of its (y, z) block, or - when both end
        // cells of an interior row lie off the grid (the ring is wider than the x
        // extent, as in a grid one cell wide in x) - only the block's boundary rows,
        // whose whole runs are then the ring's only cells: a bottom and a top row
        // (z = c_z -+ R, all y) and the side rows y = c_y -+ R of the z in between.
        bool edge = false;
        auto ring_bounds = [&]() {
            x0 = max(c[0] - R, 0);
            x1 = min(c[0] + R, g[0] - 1);
            y0 = max(c[1] - R, 0);
            const int y1 = min(c[1] + R, g[1] - 1);
            z0 = max(c[2] - R, 0);
            const int z1 = min(c[2] + R, g[2] - 1);
            ny = y1 - y0 + 1;
            edge = NARROW && R > 0 && c[0] - R < 0 && c[0] + R > g[0] - 1;
            if (edge) {
                const int nbt = (c[2] - R >= 0 ? ny : 0) + (c[2] + R <= g[2] - 1 ? ny : 0);
                const int nzm = max(min(c[2] + R - 1, g[2] - 1) - max(c[2] - R + 1, 0) + 1, 0);
                const int ne = (int)(c[1] - R >= 0) + (int)(c[1] + R <= g[1] - 1);
                nseg = 2 * (nbt + nzm * ne);
            } else {
                nseg = 2 * ny * (z1 - z0 + 1);
            }
        };
        // (y, z) of row r of the current ring in the boundary-row enumeration
        auto edge_row = [&](int r, int& y, int& z) {
            const int nbot = c[2] - R >= 0 ? ny : 0;
            if (r < nbot) {
                y = y0 + r;
                z = c[2] - R;
                return;
            }
            r -= nbot;
            const int ntop = c[2] + R <= g[2] - 1 ? ny : 0;
            if (r < ntop) {
                y = y0 + r;
                z = c[2] + R;
                return;
            }
            r -= ntop;
            const bool ylo = c[1] - R >= 0, yhi = c[1] + R <= g[1] - 1;
            const bool two = ylo && yhi;
            z = max(c[2] - R + 1, 0) + (two ? (r >> 1) : r);
            y = (ylo && !(two && (r & 1))) ? c[1] - R : c[1] + R;
        };
        ring_bounds();
        const uint64_t* rmask = bd0.rowmask;
        const int wpr = (g[0] + 63) >> 6;
        // Rows of the ring run y fastest: row r is (y0 + r % ny, z0 + r / ny); a row
        // has two segments: part 0 (a full row's whole run, or the interior row's
        // left end cell) and part 1 (the interior row's right end cell).  Runs are
        // clipped to the ball of squared radius rr around the query: a ref outside
        // it has key > tk (rr includes the slack tol), so it would not be
        // admitted.  The clip is taken in cell units, q_x / h -+ rx / h, with rx / h
        // from an FP32 square root enlarged by a relative 1e-6 and 1e-6 of a cell:
        // that covers every rounding of the FP64 floor map of the binning, so the
        // clip only ever widens.
        const double qxc = (qu[0] - lo[0]) * inv_h;
        auto row_runs = [&](int y, int z, int& a0, int& b0, int& a1, int& b1) {
            const bool full = R == 0 || z == c[2] - R || z == c[2] + R || y == c[1] - R ||
                              y == c[1] + R;
            a0 = full ? x0 : c[0] - R;
            b0 = full ? x1 : c[0] - R;
            a1 = full ? 1 : c[0] + R;
            b1 = full ? 0 : c[0] + R;
            if (a0 < 0) {  // interior row whose left end is off the grid
                a0 = 1;
                b0 = 0;
            }
            if (a1 >= g[0]) {
                a1 = 1;
                b1 = 0;
            }
            if (rr < __builtin_inf()) {
                double s2 = 0.0;
                if (D >= 2) {
                    const double yl = lo[1] + (double)y * h;
                    const double dy = fmax(fmax(yl - qu[1], qu[1] - (yl + h)), 0.0);
                    s2 = dy * dy;
                }
                if (D >= 3) {
                    const double zl = lo[2] + (double)z * h;
                    const double dz = fmax(fmax(zl - qu[2], qu[2] - (zl + h)), 0.0);
                    s2 = fma(dz, dz, s2);
                }
                const double rem = rr - s2;
                int xl = 0, xh = -1;
                if (rem >= 0.0) {
                    const double rc = (double)(__builtin_sqrtf((float)(rem * inv_h * inv_h)) *
                                                   1.000001f +
                                               1.0e-6f);
                    const double gx = (double)g[0];
                    xl = (int)fmin(fmax(floor(qxc - rc), -1.0), gx);
                    xh = (int)fmin(fmax(floor(qxc + rc), -1.0), gx);
                }
                a0 = max(a0, xl);
                b0 = min(b0, xh);
                a1 = max(a1, xl);
                b1 = min(b1, xh);
            }
        };
        // does run [a, b] of row `row` hold a ref?  (runs across two occupancy
        // words are not checked)
        auto run_has_ref = [&](int row, int a, int b) -> bool {
            if (a > b) return false;
            if ((a >> 6) != (b >> 6)) return true;
            const uint64_t m = rmask[(int64_t)row * wpr + (a >> 6)];
            const int wdt = b - a + 1;
            const uint64_t sel = wdt >= 64 ? ~0ull : ((1ull << wdt) - 1ull);
            return ((m >> (a & 63)) & sel) != 0ull;
        };
        // the refs outside the grid's box first (every lane keys them all)
        int32_t p = 0, e = 0;
        if (active) {
            p = cs[bd0.ref_out];
            e = cs[bd0.ref_out + 1];
        }
        bool ring_end = false;   // ring walked: wait for the merge, then the stop test
        bool done = !active;
#pragma unroll 1
        for (;;) {
            if (__builtin_amdgcn_ballot_w64((cnt > B - CH) | (ring_end & (cnt > 0)))) {
                const int nb = wave_max32(cnt);
                // K = 16: insertion ladders (the networks would not fit its 3-wave
                // register budget); K = 32 and 64: slices of 8 slots through one
                // network instance (24 slots, merged when a lane holds more than
                // B - CH = 12; with (key, index) merges and CH = 8, 24 slots merged
                // above 16 took 1.91 ms at K = 32 and 5.83 ms at K = 64 on the
                // 434,874-row Vecchia search against 1.99 and 6.14 ms for 16 slots
                // merged above 8; a 16-slot network beside the 8-slot one spills)
                if (K == 16) {
#pragma unroll 1
                    for (int b = 0; b < nb; ++b) {
                        const bool v = b < cnt;
                        const double x = bk[b][lane];
                        const int xi = bi[b][lane];
                        topk_insert<double, K>(keys, ids, v ? x : __builtin_inf(), v ? xi : -1);
                    }
                } else {
                    double dmin = __builtin_inf();
#pragma unroll 1
                    for (int o = 0; o < nb; o += 8)
                        merge_slots_keys<K, 8>(keys, ids, bk, bi, lane, o, cnt, dmin);
                    // dropped keys are >= the kept ones: a tie at the k-th place
                    tie |= (dmin == keys[K - 1]) & (keys[K - 1] < __builtin_inf());
                }
                cnt = 0;
                if (active) {
                    tk = keys[K - 1];
                    if (!KEYS) ti = ids[K - 1];
                    if (tk < __builtin_inf()) {
                        const double r = sqrt(tk) + tol;
                        rr = r * r;
                    }
                }
            }
            if (ring_end) {
                // Lower bound of the distance from the query to any ref in a cell
                // outside the block walked so far: the nearest block face with
                // cells beyond it.
                double bd = __builtin_inf();
#pragma unroll
                for (int f = 0; f < D; ++f) {
                    if (c[f] - R >= 1) bd = fmin(bd, qu[f] - (lo[f] + (double)(c[f] - R) * h));
                    if (c[f] + R + 1 <= g[f] - 1)
                        bd = fmin(bd, (lo[f] + (double)(c[f] + R + 1) * h) - qu[f]);
                }
                done = bd == __builtin_inf();  // every cell visited
                if (!done) {
                    const double bb = bd - tol;
                    done = bb > 0.0 && keys[K - 1] < bb * bb;
                }
                ring_end = false;
                if (!done) {
                    ++R;
                    ring_bounds();
                    sg = 0;
                }
            }
            if (!__builtin_amdgcn_ballot_w64(!done)) break;
            if (done) continue;
            if (p < e) {
                // the next CH refs of the segment, loaded together
                // Only the words used are loaded (a loaded but unused register
                // would be reused while its load is in flight, serialising the
                // loads): u_0, u_1 as 16 bytes, u_2 and the low word of idx as 12.
                double ru[CH][3];
                int rix[CH];
#pragma unroll
                for (int u = 0; u < CH; ++u) {
                    const GridPt* rq = rpts + min(p + u, e - 1);
                    const u32x4 w01 = *reinterpret_cast<const u32x4*>(rq);
                    const u32x3 w2 = *reinterpret_cast<const u32x3*>(&rq->u[2]);
                    ru[u][0] = __hiloint2double((int)w01.y, (int)w01.x);
                    ru[u][1] = __hiloint2double((int)w01.w, (int)w01.z);
                    ru[u][2] = __hiloint2double((int)w2.y, (int)w2.x);
                    rix[u] = (int)w2.z;
                }
#pragma unroll
                for (int u = 0; u < CH; ++u) {
                    const double t0 = ru[u][0] - qu[0];
                    double key = t0 * t0;
#pragma unroll
                    for (int f = 1; f < D; ++f) {
                        const double tt = ru[u][f] - qu[f];
                        key = fma(tt, tt, key);
                    }
                    const int ri = rix[u];
                    const bool in = (p + u < e) & (ri < lim32);
                    if (KEYS) tie |= in & (key == tk);
                    if (in & ((key < tk) | (!KEYS & (key == tk) & (ri < ti)))) {
                        bk[cnt][lane] = key;
                        bi[cnt][lane] = ri;
                        ++cnt;
                    }
                }
                examined += (uint32_t)min(CH, e - p);
                p += CH;
            }
            // a lane whose run is finished finds its next run in the same step
            if (p >= e) {
              if (sg < nseg) {
                // The segments of rows r0 and r0 + 1 from sg on: their runs and
                // occupancy; the first run with a ref is loaded, the empty ones
                // before it skipped.
                const int r0 = sg >> 1;
                const int rz0 = edge ? 0 : r0 / ny;
                const int ry0 = r0 - rz0 * ny;
                int ra[4], rb[4], rrow[4];
                bool has[4];
#pragma unroll
                for (int v = 0; v < 2; ++v) {
                    const bool rv = 2 * (r0 + v) < nseg;
                    int yy = ry0 + v, zz = rz0;
                    if (yy >= ny) {
                        yy -= ny;
                        ++zz;
                    }
                    int y = y0 + yy, z = z0 + zz;
                    if constexpr (NARROW)
                        if (edge && rv) edge_row(r0 + v, y, z);
                    int a0 = 1, b0 = 0, a1 = 1, b1 = 0;
                    if (rv) row_runs(y, z, a0, b0, a1, b1);
                    const int row = y + g[1] * z;
                    if (v == 0 && (sg & 1)) {  // part 0 of row r0 already done
                        a0 = 1;
                        b0 = 0;
                    }
                    ra[2 * v] = a0;
                    rb[2 * v] = b0;
                    ra[2 * v + 1] = a1;
                    rb[2 * v + 1] = b1;
                    rrow[2 * v] = rrow[2 * v + 1] = row;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) has[u] = run_has_ref(rrow[u], ra[u], rb[u]);
                int pick = 4;
#pragma unroll
                for (int u = 3; u >= 0; --u)
                    if (has[u]) pick = u;
                int xa = 1, xb = 0, xrow = 0;
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (pick == u) {
                        xa = ra[u];
                        xb = rb[u];
                        xrow = rrow[u];
                    }
                if (pick < 4) {
                    const int xbase = g[0] * xrow;
                    p = cs[xbase + xa];
                    e = cs[xbase + xb + 1];
                }
                // next segment: after the picked one, else after row r0 + 1
                sg = min(pick < 4 ? 2 * r0 + pick + 1 : 2 * r0 + 4, nseg);
              } else {
                ring_end = true;
              }
            }
        }
    }
    int pad_out = K - a.k;
    asm volatile("" : "+v"(pad_out));
    if (jv) store_list<K>(ids, pad_out, a.out_idx + q * a.k);
    if (KEYS) {
        // equal keys out of index order (placeholders share one index)
#pragma unroll
        for (int t = 0; t + 1 < K; ++t)
            tie |= (keys[t] == keys[t + 1]) & ((unsigned)ids[t] > (unsigned)ids[t + 1]);
        if (jv & tie) {
            const int r = atomicAdd(a.redo_n, 1);
            a.redo[2 * r] = bj;
            a.redo[2 * r + 1] = (int32_t)j;
        }
    }
    if (a.pair_count) {
        unsigned long long s = examined;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (lane == 0 && s) atomicAdd(a.pair_count, s);
    }
}

// ---------------------------------------------------------------------------
// Wave-per-query search
// ---------------------------------------------------------------------------
// Value of lane `src` (wave-uniform) of a double / int.
__device__ __forceinline__ double lane_d(double v, int src) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Candidates of one chunk merged one by one at most this many, else by a sort.
constexpr int kWaveInsertMax = 6;

// One query per wave, every control decision wave-uniform.  The wave keeps the 64
// smallest admitted (key, index) pairs sorted across its lanes (lane t: the t-th),
// so the admission threshold is the pair in lane k - 1.  Ring R of cells around
// the query's cell is enumerated one row (y, z) per lane: a row's whole x run when
// it is a boundary row of the ring, else its two end cells, each run clipped to
// the ball of the current admission radius.  The refs of the rows' runs are then
// examined 64 at a time (lane l takes position l of the rows' concatenation, found
// by a binary search over the exclusive prefix of the run lengths): the FP64 key,
// the Vecchia limit and the threshold.  A few admitted refs are inserted into the
// list one by one; more are sorted by a bitonic network and merged (elementwise
// minimum against the reversed chunk, then half-cleaners).  After ring R the
// query is done when its k-th key is below the squared distance to the nearest
// face of the visited block with cells beyond it (minus tol), or when the block
// covers the grid.  Rings whose interior rows have both end cells off the grid
// (wider than the grid in x) enumerate their boundary rows only.
template <int D>
__device__ __forceinline__ void wave_search_query(const SearchArgs& a, const BandDesc& bd,
                                                  int64_t jq, int lane) {
    const GridPt qp = bd.qpts[jq];
    double qu[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) qu[f] = qp.u[f];
    const int64_t q = qp.idx;
    int64_t lim = a.n_ref;
    if (a.has_cut) {
        lim = q + a.cut;
        lim = lim < 0 ? 0 : (lim > a.n_ref ? a.n_ref : lim);
    }
    bool fin = true;
#pragma unroll
    for (int f = 0; f < D; ++f) fin = fin && __builtin_isfinite(qu[f]);
    const int k = a.k;
    // +inf through a register: held as a constant in an SGPR pair it was emitted as
    // s_mov_b64 with a 64-bit literal, which this toolchain encoded as 0 for gfx950
    double bk = __builtin_inf();
    asm volatile("" : "+v"(bk));
    int bi = -1;
    if (fin && lim > 0) {
        const GridParams& P = *bd.gp;
        int g[3], c[3];
        double lo[3];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            g[f] = P.g[f];
            lo[f] = P.lo[f];
            c[f] = f < D ? cell_coord(qu[f], lo[f], P.inv_h, g[f]) : 0;
        }
        const double h = P.h, inv_h = P.inv_h, tol = P.tol;
        const double qxc = (qu[0] - lo[0]) * inv_h;
        const int32_t* cs = bd.cell_start;
        const GridPt* rpts = bd.rpts;
        const int32_t lim32 = (int32_t)lim;
        double tk, rr = bk;  // admission threshold pair (the k-th); squared clip radius
        int ti;
        auto set_threshold = [&]() {
            tk = lane_d(bk, k - 1);
            ti = __builtin_amdgcn_readlane(bi, k - 1);
            if (tk < __builtin_inf()) {
                const double r = sqrt(tk) + tol;
                rr = r * r;
            }
        };
        set_threshold();
        // key the ref at position pos of the grid's point array (pv: the lane holds
        // one) and admit it
        auto consume = [&](int pos, bool pv) {
            const GridPt* rq = rpts + pos;
            const u32x4 w01 = *reinterpret_cast<const u32x4*>(rq);
            const u32x3 w2 = *reinterpret_cast<const u32x3*>(&rq->u[2]);
            double ru[3];
            ru[0] = __hiloint2double((int)w01.y, (int)w01.x);
            ru[1] = __hiloint2double((int)w01.w, (int)w01.z);
            ru[2] = __hiloint2double((int)w2.y, (int)w2.x);
            const int ri = (int)w2.z;
            const double t0 = ru[0] - qu[0];
            double key = t0 * t0;
#pragma unroll
            for (int f = 1; f < D; ++f) {
                const double tt = ru[f] - qu[f];
                key = fma(tt, tt, key);
            }
            const bool adm = pv & (ri < lim32) & pair_less(key, ri, tk, ti);
            uint64_t bal = __builtin_amdgcn_ballot_w64(adm);
            if (!bal) return;
            if (__builtin_popcountll(bal) <= kWaveInsertMax) {
                do {
                    const int s = (int)__builtin_ctzll(bal);
                    bal &= bal - 1;
                    const double ck = lane_d(key, s);
                    const int ci = __builtin_amdgcn_readlane(ri, s);
                    if (!pair_less(ck, ci, tk, ti)) continue;
                    const uint64_t m = __builtin_amdgcn_ballot_w64(pair_less(ck, ci, bk, bi));
                    const int at = (int)__builtin_ctzll(m);
                    const double pk = __shfl_up(bk, 1, 64);
                    const int pi = __shfl_up(bi, 1, 64);
                    if (lane > at) {
                        bk = pk;
                        bi = pi;
                    } else if (lane == at) {
                        bk = ck;
                        bi = ci;
                    }
                    set_threshold();
                } while (bal);
            } else {
                double sk = adm ? key : __builtin_inf();
                int si = adm ? ri : -1;
#pragma unroll
                for (int size = 2; size <= 64; size <<= 1)
#pragma unroll
                    for (int stride = size >> 1; stride > 0; stride >>= 1)
                        wave_pair_step_x(sk, si, lane, size, stride);
                const double rk = __shfl(sk, 63 - lane, 64);
                const int rix = __shfl(si, 63 - lane, 64);
                if (pair_less(rk, rix, bk, bi)) {
                    bk = rk;
                    bi = rix;
                }
#pragma unroll
                for (int stride = 32; stride > 0; stride >>= 1)
                    wave_pair_step_x(bk, bi, lane, 64, stride);
                set_threshold();
            }
        };
        // the refs outside the grid's box first
        {
            const int32_t os = cs[bd.ref_out], oe = cs[bd.ref_out + 1];
#pragma unroll 1
            for (int32_t cb = os; cb < oe; cb += 64) {
                const int32_t pos = cb + lane;
                consume(pos < oe ? pos : os, pos < oe);
            }
        }
#pragma unroll 1
        for (int R = 0;; ++R) {
            const int y0 = max(c[1] - R, 0), y1 = min(c[1] + R, g[1] - 1);
            const int z0 = max(c[2] - R, 0), z1 = min(c[2] + R, g[2] - 1);
            const int ny = y1 - y0 + 1, nz = z1 - z0 + 1;
            const bool edge = R > 0 && c[0] - R < 0 && c[0] + R > g[0] - 1;
            const int nbot = c[2] - R >= 0 ? ny : 0;
            const int ntop = (R > 0 && c[2] + R <= g[2] - 1) ? ny : 0;
            const int zm0 = max(c[2] - R + 1, 0), zm1 = min(c[2] + R - 1, g[2] - 1);
            const int nzm = max(zm1 - zm0 + 1, 0);
            const bool ylo = c[1] - R >= 0, yhi = R > 0 && c[1] + R <= g[1] - 1;
            const int nside = (int)ylo + (int)yhi;
            const int nrows = edge ? nbot + ntop + nzm * nside : ny * nz;
#pragma unroll 1
            for (int rb = 0; rb < nrows; rb += 64) {
                const int r = rb + lane;
                int y = 0, z = 0;
                if (!edge) {
                    y = y0 + r % ny;
                    z = z0 + r / ny;
                } else {
                    int rr2 = r;
                    if (rr2 < nbot) {
                        y = y0 + rr2;
                        z = c[2] - R;
                    } else if ((rr2 -= nbot) < ntop) {
                        y = y0 + rr2;
                        z = c[2] + R;
                    } else {
                        rr2 -= ntop;
                        z = zm0 + (nside == 2 ? (rr2 >> 1) : rr2);
                        y = (ylo && !(nside == 2 && (rr2 & 1))) ? c[1] - R : c[1] + R;
                    }
                }
                int a0 = 1, b0 = 0, a1 = 1, b1 = 0;
                if (r < nrows) {
                    const bool full = R == 0 || z == c[2] - R || z == c[2] + R ||
                                      y == c[1] - R || y == c[1] + R;
                    if (full) {
                        a0 = max(c[0] - R, 0);
                        b0 = min(c[0] + R, g[0] - 1);
                    } else {
                        if (c[0] - R >= 0) a0 = b0 = c[0] - R;
                        if (c[0] + R <= g[0] - 1) a1 = b1 = c[0] + R;
                    }
                    if (rr < __builtin_inf()) {
                        // clip to the ball (cell units, widened as in lane_knn_kernel)
                        double s2 = 0.0;
                        if (D >= 2) {
                            const double yl = lo[1] + (double)y * h;
                            const double dy = fmax(fmax(yl - qu[1], qu[1] - (yl + h)), 0.0);
                            s2 = dy * dy;
                        }
                        if (D >= 3) {
                            const double zl = lo[2] + (double)z * h;
                            const double dz = fmax(fmax(zl - qu[2], qu[2] - (zl + h)), 0.0);
                            s2 = fma(dz, dz, s2);
                        }
                        const double rem = rr - s2;
                        int xl = 0, xh = -1;
                        if (rem >= 0.0) {
                            const double rc =
                                (double)(__builtin_sqrtf((float)(rem * inv_h * inv_h)) *
                                             1.000001f +
                                         1.0e-6f);
                            const double gx = (double)g[0];
                            xl = (int)fmin(fmax(floor(qxc - rc), -1.0), gx);
                            xh = (int)fmin(fmax(floor(qxc + rc), -1.0), gx);
                        }
                        a0 = max(a0, xl);
                        b0 = min(b0, xh);
                        a1 = max(a1, xl);
                        b1 = min(b1, xh);
                    }
                }
                const int64_t rowb = (int64_t)g[0] * (y + (int64_t)g[1] * z);
                int p0 = 0, n0 = 0, p1 = 0, n1 = 0;
                if (a0 <= b0) {
                    p0 = cs[rowb + a0];
                    n0 = cs[rowb + b0 + 1] - p0;
                }
                if (a1 <= b1) {
                    p1 = cs[rowb + a1];
                    n1 = cs[rowb + b1 + 1] - p1;
                }
                // exclusive prefix of the lanes' ref counts
                const int cntl = n0 + n1;
                int incl = cntl;
#pragma unroll
                for (int off = 1; off < 64; off <<= 1) {
                    const int o = __shfl_up(incl, off, 64);
                    if (lane >= off) incl += o;
                }
                const int ex = incl - cntl;
                const int T = __builtin_amdgcn_readlane(incl, 63);
#pragma unroll 1
                for (int cb = 0; cb < T; cb += 64) {
                    const int Pn = cb + lane;
                    // the lane whose run holds position Pn: the last with ex <= Pn
                    int src = 0, es = 0;
#pragma unroll
                    for (int s = 32; s > 0; s >>= 1) {
                        const int e = __shfl(ex, src + s, 64);
                        if (e <= Pn) {
                            src += s;
                            es = e;
                        }
                    }
                    const int o = Pn - es;
                    const int sp0 = __shfl(p0, src, 64);
                    const int sn0 = __shfl(n0, src, 64);
                    const int sp1 = __shfl(p1, src, 64);
                    const bool pv = Pn < T;
                    const int pos = pv ? (o < sn0 ? sp0 + o : sp1 + (o - sn0)) : 0;
                    consume(pos, pv);
                }
            }
            // stop test: the nearest face of the visited block with cells beyond it
            double bdist = __builtin_inf();
#pragma unroll
            for (int f = 0; f < D; ++f) {
                if (c[f] - R >= 1) bdist = fmin(bdist, qu[f] - (lo[f] + (double)(c[f] - R) * h));
                if (c[f] + R + 1 <= g[f] - 1)
                    bdist = fmin(bdist, (lo[f] + (double)(c[f] + R + 1) * h) - qu[f]);
            }
            if (bdist == __builtin_inf()) break;  // every cell visited
            const double bb = bdist - tol;
            if (bb > 0.0 && tk < bb * bb) break;
        }
    }
    if (lane < k) a.out_idx[q * k + lane] = bi;
}

// Every query of every band, one per wave (band table: wave_off = first query slot
// of the band).
template <int D>
__global__ __launch_bounds__(256) void wave_knn_kernel(SearchArgs a) {
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int64_t gq = (int64_t)blockIdx.x * 4 + w;  // query slot over all bands
    if (gq >= a.n_waves) return;
    int bj = 0;
    for (int t = 1; t < a.n_bands; ++t)
        if (a.band[t].wave_off <= gq) bj = t;
    wave_search_query<D>(a, a.band[bj], gq - a.band[bj].wave_off, lane);
}

// The queries outside the grid's box (each band's far bucket, counts on the
// device), then the queries lane_knn_kernel queued for an exact search (ties at
// the k-th key), one per wave, the waves striding over them; after
// lane_knn_kernel, which leaves the far queries out.
template <int D>
__global__ __launch_bounds__(256) void wave_far_kernel(SearchArgs a) {
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    // lane t: band t's far queries [f0, f0 + cnt) and the inclusive prefix of the
    // counts (bands <= kMaxGridBands + 1 <= 64)
    int32_t f0 = 0, cnt = 0;
    if (lane < a.n_bands) {
        const BandDesc& b = a.band[lane];
        f0 = b.qstart[b.qfar];
        cnt = b.qstart[b.qfar + 1] - f0;
    }
    int32_t incl = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int32_t o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
    }
    const int64_t total = __builtin_amdgcn_readlane(incl, 63);
    const int64_t redo = a.redo_n ? __builtin_amdgcn_readfirstlane(*a.redo_n) : 0;
    if (a.rows_recomputed && blockIdx.x == 0 && threadIdx.x == 0) *a.rows_recomputed += redo;
    const int64_t stride = (int64_t)gridDim.x * 4;
#pragma unroll 1
    for (int64_t s = (int64_t)blockIdx.x * 4 + w; s < total + redo; s += stride) {
        if (s >= total) {
            const int32_t* e = a.redo + 2 * (s - total);
            const int bj = __builtin_amdgcn_readfirstlane(e[0]);
            const int64_t jq = __builtin_amdgcn_readfirstlane(e[1]);
            wave_search_query<D>(a, a.band[bj], jq, lane);
            continue;
        }
        // the first band whose inclusive prefix exceeds s
        const uint64_t m = __builtin_amdgcn_ballot_w64((int64_t)incl > s);
        const int bj = (int)__builtin_ctzll(m);
        const int64_t before = __builtin_amdgcn_readlane(incl - cnt, bj);
        const int64_t jq = __builtin_amdgcn_readlane(f0, bj) + (s - before);
        wave_search_query<D>(a, a.band[bj], jq, lane);
    }
}

// One launch over every band's queries, one wave each.
void launch_wave_search(SearchArgs& c, int d, hipStream_t st) {
    int64_t nq = 0;
    for (int t = 0; t < c.n_bands; ++t) {
        c.band[t].wave_off = nq;
        nq += c.band[t].n_q;
    }
    if (nq == 0) return;
    c.n_waves = nq;
    const unsigned blocks = (unsigned)((nq + 3) / 4);
    if (d == 1) hipLaunchKernelGGL(wave_knn_kernel<1>, dim3(blocks), dim3(256), 0, st, c);
    else if (d == 2) hipLaunchKernelGGL(wave_knn_kernel<2>, dim3(blocks), dim3(256), 0, st, c);
    else hipLaunchKernelGGL(wave_knn_kernel<3>, dim3(blocks), dim3(256), 0, st, c);
}


// Estimated cost of each search wave: the largest number of refs in the 3 x 3 x 3
// cells around one of its queries.
__global__ __launch_bounds__(64 * kCoopWaves) void wave_cost_kernel(SearchArgs a,
                                                                    int32_t* cost) {
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int64_t wg = (int64_t)blockIdx.x * kCoopWaves + w;
    if (wg >= a.n_waves) return;
    int bj = 0;
    for (int t = 1; t < a.n_bands; ++t)
        if (a.band[t].wave_off <= wg) bj = t;
    const BandDesc& bd = a.band[bj];
    const int64_t j = (wg - bd.wave_off) * 64 + lane;
    const int32_t far0 = bd.qstart[bd.qfar], far1 = bd.qstart[bd.qfar + 1];
    int c = 0;
    if (j < bd.n_q && !(j >= far0 && j < far1)) {
        const GridPt qp = bd.qpts[j];
        const GridParams& P = *bd.gp;
        bool fin = true;
        int cc[3];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            fin = fin && __builtin_isfinite(qp.u[f]);
            cc[f] = fin ? cell_coord(qp.u[f], P.lo[f], P.inv_h, P.g[f]) : 0;
        }
        if (fin) {
            const int x0 = max(cc[0] - 1, 0), x1 = min(cc[0] + 1, P.g[0] - 1);
            for (int z = max(cc[2] - 1, 0); z <= min(cc[2] + 1, P.g[2] - 1); ++z)
                for (int y = max(cc[1] - 1, 0); y <= min(cc[1] + 1, P.g[1] - 1); ++y) {
                    const int base = P.g[0] * (y + P.g[1] * z);
                    c += bd.cell_start[base + x1 + 1] - bd.cell_start[base + x0];
                }
        }
    }
    c = wave_max32(c);
    if (lane == 0) cost[wg] = c;
}

// One block: counting sort of the waves by cost bucket (4 per doubling),
// most expensive first (order inside a bucket is arbitrary).
constexpr int kCostBuckets = 64;
__global__ __launch_bounds__(1024) void wave_sort_kernel(const int32_t* cost, int64_t n,
                                                        int32_t* perm, int32_t* redo_n) {
    __shared__ int hist[kCostBuckets];
    if (threadIdx.x < kCostBuckets) hist[threadIdx.x] = 0;
    if (threadIdx.x == 0) *redo_n = 0;  // the search's redo queue starts empty
    __syncthreads();
    auto bucket = [](int c) {
        const int b = (int)(4.0f * __log2f(1.0f + (float)c));
        return kCostBuckets - 1 - min(max(b, 0), kCostBuckets - 1);
    };
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) atomicAdd(&hist[bucket(cost[i])], 1);
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int b = 0; b < kCostBuckets; ++b) {
            const int v = hist[b];
            hist[b] = run;
            run += v;
        }
    }
    __syncthreads();
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x)
        perm[atomicAdd(&hist[bucket(cost[i])], 1)] = (int32_t)i;
}

template <int D, int K>
void launch_search_dk(const SearchArgs& c, hipStream_t st) {
    const unsigned blocks = (unsigned)((c.n_waves + kLaneWaves - 1) / kLaneWaves);
    hipLaunchKernelGGL((lane_knn_kernel<D, K, false>), dim3(blocks), dim3(64 * kLaneWaves), 0, st,
                       c);
    hipLaunchKernelGGL((lane_knn_kernel<D, K, true>), dim3(blocks), dim3(64 * kLaneWaves), 0, st,
                       c);
}

// One launch over every band's waves.
// Waves are launched most expensive first, so that long waves do not start last.
void launch_search(SearchArgs& c, int d, const GridSearchMem& M, hipStream_t st) {
    int64_t waves = 0;
    for (int t = 0; t < c.n_bands; ++t) {
        c.band[t].wave_off = waves;
        waves += (c.band[t].n_q + 63) / 64;
    }
    if (waves == 0) return;
    c.n_waves = waves;
    c.perm = M.wave_perm;
    const unsigned blocks = (unsigned)((waves + kCoopWaves - 1) / kCoopWaves);
    hipLaunchKernelGGL(wave_cost_kernel, dim3(blocks), dim3(64 * kCoopWaves), 0, st, c,
                       M.wave_cost);
    c.redo = M.redo;
    c.redo_n = M.redo_n;
    hipLaunchKernelGGL(wave_sort_kernel, dim3(1), dim3(1024), 0, st, M.wave_cost, waves,
                       M.wave_perm, M.redo_n);
    const int K = c.k <= 16 ? 16 : (c.k <= 32 ? 32 : 64);
    // the queries outside the grid's box, which the lane kernel leaves out: their
    // count is on the device, so the waves stride over them; lists of 32 and 64
    // also queue the queries with ties at the k-th key (on data with many equal
    // distances, most of them): one round of the kernel's occupancy (5 waves per
    // SIMD, 1,280 blocks of 4 waves) for those, 512 waves for 16
    int64_t nq = 0;
    for (int t = 0; t < c.n_bands; ++t) nq += c.band[t].n_q;
    const unsigned fb = (unsigned)std::min<int64_t>((nq + 3) / 4, K == 16 ? 128 : 1280);
    switch (d * 100 + K) {
        case 116: launch_search_dk<1, 16>(c, st); break;
        case 132: launch_search_dk<1, 32>(c, st); break;
        case 164: launch_search_dk<1, 64>(c, st); break;
        case 216: launch_search_dk<2, 16>(c, st); break;
        case 232: launch_search_dk<2, 32>(c, st); break;
        case 264: launch_search_dk<2, 64>(c, st); break;
        case 316: launch_search_dk<3, 16>(c, st); break;
        case 332: launch_search_dk<3, 32>(c, st); break;
        default: launch_search_dk<3, 64>(c, st); break;
    }
    if (d == 1) hipLaunchKernelGGL(wave_far_kernel<1>, dim3(fb), dim3(256), 0, st, c);
    else if (d == 2) hipLaunchKernelGGL(wave_far_kernel<2>, dim3(fb), dim3(256), 0, st, c);
    else hipLaunchKernelGGL(wave_far_kernel<3>, dim3(fb), dim3(256), 0, st, c);
}

int grid_cap(int64_t n_pts) {
    return (int)(2 * std::max<int64_t>(1, n_pts / kPointsPerCell) + 64);
}

// Query buckets of a grid with `cap` cells: tiling pads each tiled feature by at
// most 3 cells, which for the features it applies to (>= 8 cells) is <= 1.375x.
int grid_capq(int cap) { return 3 * cap + 64; }

void carve_grid(Carver& c, int64_t n_pts, GridMem* M) {
    const int cap = grid_cap(n_pts);
    const int64_t ns = (int64_t)cap + 3;
    M->cap = cap;
    M->params = c.take<GridParams>(1);
    M->cnt = c.take<int32_t>((size_t)ns);
    M->start = c.take<int32_t>((size_t)ns);
    M->fill = c.take<int32_t>((size_t)ns);
    M->sums = c.take<int32_t>((size_t)scan_blocks(ns) + 1);
    M->cell_of = c.take<int32_t>((size_t)n_pts + 1);
    M->pts = c.take<GridPt>((size_t)n_pts + 1);
    M->rowmask = c.take<uint64_t>(rowmask_words_max(cap));
}

// Counting-sort job of a grid over the first n_pts refs.
SortJob ref_job(const GridMem& M, int64_t n_pts) {
    SortJob J{};
    J.gp = M.params;
    J.cnt = M.cnt;
    J.start = M.start;
    J.fill = M.fill;
    J.sums = M.sums;
    J.bucket_of = M.cell_of;
    J.pts = M.pts;
    J.rowmask = M.rowmask;
    J.n = n_pts;
    J.row0 = 0;
    J.nb = M.cap + 2;
    J.tiled = 0;
    J.cap = M.cap;
    J.capq = grid_capq(M.cap);
    return J;
}

// Counting-sort job of query rows [qa, qb) into the tiles of grid M, written to
// qpts[pos ...], and the band it describes.
SortJob query_job(const GridMem& M, int64_t qa, int64_t qb, int64_t pos,
                  const GridQueryBins& bins, const GridSearchMem& S, BandDesc* band) {
    SortJob J{};
    J.gp = M.params;
    J.cnt = bins.cnt;
    J.start = bins.start;
    J.fill = bins.fill;
    J.sums = bins.sums;
    J.bucket_of = S.qcell_of + pos;
    J.pts = S.qpts + pos;
    J.rowmask = nullptr;
    J.n = qb - qa;
    J.row0 = qa;
    J.nb = bins.capq + 2;
    J.tiled = 1;
    J.cap = M.cap;
    J.capq = bins.capq;
    *band = BandDesc{};
    band->gp = M.params;
    band->cell_start = M.start;
    band->rowmask = M.rowmask;
    band->rpts = M.pts;
    band->qpts = S.qpts + pos;
    band->qstart = bins.start;
    band->qfar = bins.capq;
    band->ref_out = M.cap;
    band->n_q = qb - qa;
    return J;
}

void carve_bins(Carver& c, int capq, GridQueryBins* Q) {
    const int64_t ns = (int64_t)capq + 3;
    Q->capq = capq;
    Q->cnt = c.take<int32_t>((size_t)ns);
    Q->start = c.take<int32_t>((size_t)ns);
    Q->fill = c.take<int32_t>((size_t)ns);
    Q->sums = c.take<int32_t>((size_t)scan_blocks(ns) + 1);
}

}  // namespace

bool grid_supported(int d, int k) { return d >= 1 && d <= 3 && k >= 1 && k <= 64; }

int64_t grid_prefix_limit(int64_t n_ref, int k) {
    (void)n_ref;
    return std::max<int64_t>(64, 2 * (int64_t)k);
}

void carve_grid_refs(Carver& c, int64_t n_ref, GridRefMem* M) {
    M->bbox = c.take<long long>(kBoxPartials + 6 * (size_t)kBBoxBlocks);  // boxes, then block partials
    carve_grid(c, n_ref, &M->full);
}

void carve_grid_search(Carver& c, int64_t n_ref, int64_t n_query, int k, bool with_cut,
                       GridSearchMem* M) {
    *M = GridSearchMem{};
    int nl = 0;
    if (with_cut) {
        // lower bands: refs [0, 2L) with 2L < n_ref
        for (int64_t L = grid_prefix_limit(n_ref, k); 2 * L < n_ref && nl < kMaxGridBands; L *= 2) {
            carve_grid(c, 2 * L, &M->band[nl]);
            carve_bins(c, grid_capq(M->band[nl].cap), &M->qb[nl]);
            ++nl;
        }
    }
    M->n_lower = nl;
    carve_bins(c, grid_capq(grid_cap(n_ref)), &M->qb[nl]);
    M->qcell_of = c.take<int32_t>((size_t)n_query + 1);
    M->qpts = c.take<GridPt>((size_t)n_query + 1);
    // every band adds at most one partial wave
    const size_t max_waves = (size_t)(n_query / 64) + (size_t)nl + 2;
    M->wave_cost = c.take<int32_t>(max_waves);
    M->wave_perm = c.take<int32_t>(max_waves);
    M->redo = c.take<int32_t>(2 * ((size_t)n_query + 1));
    M->redo_n = c.take<int32_t>(1);
}

int grid_build(const double* u64, int64_t n, int d, const GridRefMem& M, hipStream_t st) {
    if (d < 1 || d > 3) return NNGP_ERR_INVALID_ARGUMENT;
    timing_begin(TK_KNN_GRID_BUILD, st);
    long long* part = M.bbox + kBoxPartials;
    const int gb = (int)std::min<int64_t>(kBBoxBlocks, (n + 255) / 256);
    if (gb > 0)
        hipLaunchKernelGGL(grid_bbox_kernel, dim3((unsigned)gb), dim3(256), 0, st, u64, n, d, part);
    hipLaunchKernelGGL(grid_bbox_final_kernel, dim3(1), dim3(256), 0, st, part, gb, M.bbox);
    if (n > 0) {
        hipLaunchKernelGGL(grid_box_sample_kernel, dim3(1), dim3(kBoxSample), 0, st, u64, n, d,
                           M.bbox);
        hipLaunchKernelGGL(grid_outside_count_kernel, dim3((unsigned)gb), dim3(256), 0, st, u64, n,
                           d, M.bbox);
    }
    const SortJob J = ref_job(M.full, n);
    run_sorts(u64, M.bbox, d, &J, 1, true, st);
    timing_end(TK_KNN_GRID_BUILD, st);
    NNGP_HIP_CHECK(hipGetLastError());
    return NNGP_OK;
}

int grid_search(const GridRefMem& G, const double* ru64, int64_t n_ref, int d,
                const double* qu64, int64_t n_query, int k, int has_cut, int64_t cut,
                int64_t row0, const GridSearchMem& M, int32_t* out_idx,
                int64_t* rows_recomputed, hipStream_t st) {
    if (!grid_supported(d, k)) return NNGP_ERR_UNSUPPORTED;
    SearchArgs sa{};
    sa.rows_recomputed = rows_recomputed;
    sa.n_ref = n_ref;
    sa.has_cut = has_cut;
    sa.cut = cut;
    sa.k = k;
    sa.out_idx = out_idx;
    timing_begin(TK_KNN_GRID_BUILD, st);
    SortJob rjobs[kMaxGridBands], qjobs[kMaxGridBands + 1];
    int n_r = 0, n_q = 0;
    if (!has_cut) {
        qjobs[n_q++] = query_job(G.full, 0, n_query, 0, M.qb[M.n_lower], M, &sa.band[0]);
        sa.n_bands = 1;
    } else {
        // Band j: rows whose limit q + cut lies in [L, 2L), L = L0 * 2^j; the top
        // band (2L >= n_ref) takes every row above and searches the full grid.
        const int64_t L0 = grid_prefix_limit(n_ref, k);
        int j = 0;
        for (int64_t L = L0;; L *= 2, ++j) {
            const bool top = 2 * L >= n_ref;
            if (!top && j >= M.n_lower) return NNGP_ERR_WORKSPACE_TOO_SMALL;
            int64_t qa = std::max<int64_t>(L - cut, row0);
            int64_t qb = top ? n_query : std::min<int64_t>(2 * L - cut, n_query);
            qa = std::min<int64_t>(qa, n_query);
            if (qb > qa) {
                const GridMem* gm = &G.full;
                if (!top) {
                    rjobs[n_r++] = ref_job(M.band[j], 2 * L);
                    gm = &M.band[j];
                }
                qjobs[n_q++] = query_job(*gm, qa, qb, qa - row0, M.qb[top ? M.n_lower : j], M,
                                         &sa.band[sa.n_bands]);
                ++sa.n_bands;
            }
            if (top) break;
        }
    }
    // grids over the lower bands' ref prefixes (bounding box of all refs), then
    // every band's queries into tile order
    if (n_r > 0) run_sorts(ru64, G.bbox, d, rjobs, n_r, true, st);
    run_sorts(qu64, G.bbox, d, qjobs, n_q, false, st);
    timing_end(TK_KNN_GRID_BUILD, st);
    sa.pair_count = timing_work_counter(TK_KNN_GRID);
    timing_begin(TK_KNN_GRID, st);
    const int meth = knn_method();
    int64_t n_searched = 0;
    for (int t = 0; t < sa.n_bands; ++t) n_searched += sa.band[t].n_q;
    if (meth == NNGP_KNN_GRID_WAVES ||
        (meth != NNGP_KNN_GRID_LANES && n_searched <= kWaveSearchMaxQueries))
        launch_wave_search(sa, d, st);
    else
        launch_search(sa, d, M, st);
    timing_end(TK_KNN_GRID, st);
    NNGP_HIP_CHECK(hipGetLastError());
    return NNGP_OK;
}

}  // namespace nngp
