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
// top-k is ordered by (key, index)).
//
// Search, one wave per 64 queries: the queries are sorted by tiles of cells, so
// a wave's queries sit in a compact box of cells.  The wave visits that box and
// then rings of cells around it (R = 1, 2, ...), streaming each cell run's refs
// once for all 64 lanes through wave-uniform (scalar) loads.  A lane appends the
// refs below its admission threshold to its column of an LDS buffer; full buffers
// are merged into the lanes' sorted register lists in lock-step.  After a ring, a
// lane is done when its k-th key is below the squared distance from its query to
// the nearest unvisited cell (minus a slack covering every rounding of the cell
// assignment and of the key); the wave stops when all its lanes are done.
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
constexpr int kCoopWaves = 4;      // waves per search block
constexpr int kCoopSlots = 16;     // LDS append slots per lane

__global__ void grid_fill_i32_kernel(int32_t* p, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__device__ __forceinline__ bool point_finite(const double* u, int d) {
    bool fin = true;
    for (int f = 0; f < d; ++f) fin = fin && __builtin_isfinite(u[f]);
    return fin;
}

__global__ void grid_bbox_init_kernel(long long* bb) {
    const int f = threadIdx.x;
    if (f < 3) {
        bb[2 * f] = 0x7fffffffffffffffLL;
        bb[2 * f + 1] = (long long)0x8000000000000000ULL;
    }
}

// Bounding box of the points whose coordinates are all finite.
__global__ __launch_bounds__(256) void grid_bbox_kernel(const double* u64, int64_t n, int d,
                                                        long long* bb) {
    __shared__ long long slo[3][256], shi[3][256];
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
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        slo[f][threadIdx.x] = lo[f];
        shi[f][threadIdx.x] = hi[f];
    }
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (int f = 0; f < 3; ++f) {
                slo[f][threadIdx.x] = min(slo[f][threadIdx.x], slo[f][threadIdx.x + s]);
                shi[f][threadIdx.x] = max(shi[f][threadIdx.x], shi[f][threadIdx.x + s]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int f = 0; f < d; ++f) {
            atomicMin(&bb[2 * f], slo[f][0]);
            atomicMax(&bb[2 * f + 1], shi[f][0]);
        }
    }
}

// One thread: cell size and grid dimensions from the bounding box, for a grid
// over n refs.  The cell size targets n / kPointsPerCell cells over the box; it
// grows until the grid fits in `cap` cells.  Query tiles are 4 cells wide along
// features with at least 8 cells; the tiled bucket count must fit in capq.
__global__ void grid_setup_kernel(const long long* bb, int d, int64_t n, int cap, int capq,
                                  GridParams* gp) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
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

// Bucket of every point and the per-bucket counts.  Refs (tiled == 0): the cell in
// row-major order.  Queries (tiled != 0): the cell in tile order.  A point with a
// non-finite coordinate goes to bucket `nonfin`.
__global__ void grid_count_kernel(const double* u64, int64_t n, int d, const GridParams* gp,
                                  int tiled, int nonfin, int32_t* cnt, int32_t* bucket_of) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* u = u64 + i * d;
    int b = nonfin;
    if (point_finite(u, d)) {
        const double ih = gp->inv_h;
        int c[3] = {0, 0, 0};
#pragma unroll
        for (int f = 0; f < 3; ++f)
            if (f < d) c[f] = cell_coord(u[f], gp->lo[f], ih, gp->g[f]);
        if (tiled) {
            const int t0 = gp->t[0], t1 = gp->t[1], t2 = gp->t[2];
            const int tile = c[0] / t0 + gp->tg[0] * (c[1] / t1 + gp->tg[1] * (c[2] / t2));
            const int local = c[0] % t0 + t0 * (c[1] % t1 + t1 * (c[2] % t2));
            b = tile * (t0 * t1 * t2) + local;
        } else {
            b = c[0] + gp->g[0] * (c[1] + gp->g[1] * c[2]);
        }
    }
    bucket_of[i] = b;
    atomicAdd(&cnt[b], 1);
}

// Point i (row idx_off + i) to its place in bucket order.
__global__ void grid_scatter_kernel(const double* u64, int64_t n, int d, int64_t idx_off,
                                    const int32_t* bucket_of, const int32_t* start, int32_t* fill,
                                    GridPt* pts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = bucket_of[i];
    const int pos = start[c] + atomicAdd(&fill[c], 1);
    GridPt p;
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        p.u[f] = 0.0;
        if (f < d) p.u[f] = u64[i * d + f];
    }
    p.idx = idx_off + i;
    pts[pos] = p;
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

__global__ __launch_bounds__(256) void scan_local_kernel(const int32_t* in, int64_t n, int32_t* out,
                                                         int32_t* sums) {
    const int64_t base = (int64_t)blockIdx.x * kScanBlock + threadIdx.x * 4;
    int32_t v[4];
    int32_t s = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        v[i] = base + i < n ? in[base + i] : 0;
        s += v[i];
    }
    int32_t tot;
    int32_t run = block_excl_scan_256(s, &tot);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void scan_sums_kernel(int32_t* sums, int nb) {
    int32_t carry = 0;
    for (int c0 = 0; c0 < nb; c0 += 256) {
        const int i = c0 + threadIdx.x;
        const int32_t v = i < nb ? sums[i] : 0;
        int32_t tot;
        const int32_t ex = block_excl_scan_256(v, &tot);
        if (i < nb) sums[i] = carry + ex;
        carry += tot;
    }
}

__global__ __launch_bounds__(256) void scan_add_kernel(int32_t* out, int64_t n, const int32_t* sums) {
    const int64_t base = (int64_t)blockIdx.x * kScanBlock;
    const int32_t off = sums[blockIdx.x];
    for (int i = threadIdx.x; i < kScanBlock; i += 256)
        if (base + i < n) out[base + i] += off;
}

int64_t scan_blocks(int64_t n) { return (n + kScanBlock - 1) / kScanBlock; }

// Counting sort of n points (rows idx_off ...) into nb buckets (the last, nb - 1,
// for non-finite points); start gets nb + 1 entries.
void bin_points(const double* u64, int64_t n, int d, int64_t idx_off, const GridParams* gp,
                int tiled, int nb, int32_t* cnt, int32_t* start, int32_t* fill, int32_t* sums,
                int32_t* bucket_of, GridPt* pts, hipStream_t st) {
    const int64_t ns = (int64_t)nb + 1;
    const unsigned gz = (unsigned)((ns + 255) / 256);
    hipLaunchKernelGGL(grid_fill_i32_kernel, dim3(gz), dim3(256), 0, st, cnt, ns, 0);
    hipLaunchKernelGGL(grid_fill_i32_kernel, dim3(gz), dim3(256), 0, st, fill, ns, 0);
    if (n > 0) {
        const unsigned gn = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(grid_count_kernel, dim3(gn), dim3(256), 0, st, u64, n, d, gp, tiled,
                           nb - 1, cnt, bucket_of);
    }
    const int64_t nsb = scan_blocks(ns);
    hipLaunchKernelGGL(scan_local_kernel, dim3((unsigned)nsb), dim3(256), 0, st, cnt, ns, start,
                       sums);
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(256), 0, st, sums, (int)nsb);
    hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)nsb), dim3(256), 0, st, start, ns, sums);
    if (n > 0) {
        const unsigned gn = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(grid_scatter_kernel, dim3(gn), dim3(256), 0, st, u64, n, d, idx_off,
                           bucket_of, start, fill, pts);
    }
}

// ---------------------------------------------------------------------------
// Search kernel
// ---------------------------------------------------------------------------
struct CoopArgs {
    const GridParams* gp;
    const int32_t* cell_start;  // ref cells, row-major
    const GridPt* rpts;         // refs in cell order
    const GridPt* qpts;         // the band's queries in tile order
    const int32_t* qstart;      // start of each query bucket (capq + 2 entries)
    const int32_t* wp;          // first work item of each tile; wp[capq + 1] = item count
    const int32_t* item_tile;   // tile of each work item
    int capq;
    int64_t n_ref;
    int has_cut;
    int64_t cut;
    int k;
    int32_t* out_idx;
};

// Queries [b, e) of tile t in the sorted query array; tile n_tiles stands for the
// bucket of queries with a non-finite coordinate.
__device__ __forceinline__ void tile_range(const GridParams& P, const int32_t* qstart, int capq,
                                           int t, int32_t& b, int32_t& e) {
    const int tv = P.t[0] * P.t[1] * P.t[2];
    const int nt = P.n_qb / tv;
    if (t < nt) {
        b = qstart[t * tv];
        e = qstart[(t + 1) * tv];
    } else if (t == nt) {
        b = qstart[capq];
        e = qstart[capq + 1];
    } else {
        b = e = 0;
    }
}

// Work items of each tile: its queries in groups of 64 (one wave each).
__global__ void items_kernel(const GridParams* gp, const int32_t* qstart, int capq,
                             int32_t* items) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= capq + 2) return;
    int32_t b, e;
    tile_range(*gp, qstart, capq, t, b, e);
    items[t] = (e - b + 63) / 64;
}

__global__ void expand_items_kernel(const int32_t* items, const int32_t* wp, int n_tiles,
                                    int32_t* item_tile) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tiles) return;
    const int32_t c = items[t], o = wp[t];
    for (int32_t i = 0; i < c; ++i) item_tile[o + i] = t;
}

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

// Value of v in lane src (wave-uniform result).
__device__ __forceinline__ double read_lane_f64(double v, int src) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int D, int K>
__global__ __launch_bounds__(64 * kCoopWaves)
__attribute__((amdgpu_waves_per_eu(K <= 16 ? 3 : 2)))
void coop_knn_kernel(CoopArgs a) {
    constexpr int B = kCoopSlots;
    constexpr int CH = 4;
    __shared__ double sk[kCoopWaves][B][64];
    __shared__ int si[kCoopWaves][B][64];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const int lane = threadIdx.x & 63;
    const GridParams& P = *a.gp;
    const int64_t n_items = a.wp[a.capq + 1];
    const int64_t n_waves = (int64_t)gridDim.x * kCoopWaves;
    double (*bk)[64] = sk[w];
    int (*bi)[64] = si[w];
#pragma unroll 1
    for (int64_t item = (int64_t)blockIdx.x * kCoopWaves + w; item < n_items; item += n_waves) {
        const int t = a.item_tile[item];
        int32_t tb, te;
        tile_range(P, a.qstart, a.capq, t, tb, te);
        const int64_t j = (int64_t)tb + 64 * (int64_t)(item - a.wp[t]) + lane;
        const bool jv = j < te;
        double qu[3] = {0.0, 0.0, 0.0};
        int64_t q = 0;
        if (jv) {
            const GridPt qp = a.qpts[j];
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
        int g[3], c[3];
        double lo[3];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            g[f] = P.g[f];
            lo[f] = P.lo[f];
            c[f] = (f < D && active) ? cell_coord(qu[f], lo[f], P.inv_h, g[f]) : 0;
        }
        // the wave's box of query cells
        int bmin[3], bmax[3];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            bmin[f] = wave_min32(active ? c[f] : INT_MAX);
            bmax[f] = wave_max32(active ? c[f] : INT_MIN);
        }
        // K - k leading placeholder entries; laundered through a register so that
        // the per-entry masks are rebuilt here and at the store, not kept live
        int pad = K - a.k;
        asm volatile("" : "+v"(pad));
        double keys[K];
        int ids[K];
        init_list<K>(keys, ids, pad);

        if (__builtin_amdgcn_ballot_w64(active)) {
            const double h = P.h, tol = P.tol;
            const int32_t* cs = a.cell_start;
            const int32_t lim32 = (int32_t)lim;
            // admission threshold pair; inactive lanes admit nothing
            double tk = active ? __builtin_inf() : -__builtin_inf();
            int ti = -1;
            int cnt = 0;
            auto merge = [&]() {
                const int nb = wave_max32(cnt);
#pragma unroll 1
                for (int b = 0; b < nb; ++b) {
                    const bool v = b < cnt;
                    const double x = bk[b][lane];
                    const int xi = bi[b][lane];
                    topk_insert<double, K>(keys, ids, v ? x : __builtin_inf(), v ? xi : -1);
                }
                cnt = 0;
                if (active) {
                    tk = keys[K - 1];
                    ti = ids[K - 1];
                }
            };
            // The traversal is one flat loop (rings -> batches of 64 rows -> rounds
            // -> groups of CH refs) so that the list merge has a single call site.
            int R = 0, rb = 0, nrows = 0, ny = 1;
            int x0 = 0, x1 = 0, y0 = 0, y1 = 0, z0 = 0, z1 = 0;
            auto ring_bounds = [&]() {
                x0 = max(bmin[0] - R, 0);
                x1 = min(bmax[0] + R, g[0] - 1);
                y0 = max(bmin[1] - R, 0);
                y1 = min(bmax[1] + R, g[1] - 1);
                z0 = max(bmin[2] - R, 0);
                z1 = min(bmax[2] + R, g[2] - 1);
                ny = y1 - y0 + 1;
                nrows = ny * (z1 - z0 + 1);
            };
            ring_bounds();
            bool need_batch = true;  // set up the next 64 rows of the ring
            bool flush = false;      // ring finished: merge, then stop test
            // the batch's rows, one per lane: refs [sA, sA + lA) then [sB, ...), and
            // the exclusive prefix `off` of the row lengths (the batch's ref stream)
            int32_t sA = 0, lA = 0, sB = 0, off = 0, tot = 0, t0 = 0;
            GridPt nxt, cp;
            uint64_t m = 0;          // lanes whose round ref is still to be scored
            // Ref at stream position t of the current batch: the row holding it by a
            // binary search over the lanes' prefixes.  Called with every lane active
            // (the shuffles read the other lanes' registers).
            auto ref_at = [&](int32_t t) -> int32_t {
                int j = 0;
#pragma unroll
                for (int step = 32; step > 0; step >>= 1) {
                    const int c2 = j + step;
                    if (__shfl(off, c2, 64) <= t) j = c2;
                }
                // all shuffles unconditionally: a shuffle under a per-lane branch
                // would read lanes that did not take that branch
                const int32_t loc = t - __shfl(off, j, 64);
                const int32_t la = __shfl(lA, j, 64);
                const int32_t sa = __shfl(sA, j, 64);
                const int32_t sb = __shfl(sB, j, 64);
                return loc < la ? sa + loc : sb + (loc - la);
            };
#pragma unroll 1
            for (;;) {
                if (__builtin_amdgcn_ballot_w64((cnt > B - CH) | (flush & (cnt > 0)))) merge();
                if (flush) {
                    // Lower bound of the distance from the query to any ref in a cell
                    // outside the visited box: the nearest box face with cells beyond.
                    double bd = __builtin_inf();
#pragma unroll
                    for (int f = 0; f < D; ++f) {
                        if (bmin[f] - R >= 1)
                            bd = fmin(bd, qu[f] - (lo[f] + (double)(bmin[f] - R) * h));
                        if (bmax[f] + R + 1 <= g[f] - 1)
                            bd = fmin(bd, (lo[f] + (double)(bmax[f] + R + 1) * h) - qu[f]);
                    }
                    bool done = !active || bd == __builtin_inf();  // infinite: all visited
                    if (!done) {
                        const double bb = bd - tol;
                        done = bb > 0.0 && keys[K - 1] < bb * bb;
                    }
                    if (!__builtin_amdgcn_ballot_w64(!done)) break;
                    ++R;
                    ring_bounds();
                    rb = 0;
                    flush = false;
                    need_batch = true;
                    continue;
                }
                if (need_batch) {
                    if (rb >= nrows) {
                        flush = true;
                        continue;
                    }
                    // Rows (y, z) of the ring, one per lane: a new row contributes
                    // its whole x run, an old row its two new end cells.
                    const int r = rb + lane;
                    rb += 64;
                    sA = 0;
                    lA = 0;
                    sB = 0;
                    int32_t lB = 0;
                    if (r < nrows) {
                        const int y = y0 + r % ny, z = z0 + r / ny;
                        const int base = g[0] * (y + g[1] * z);
                        const bool full = R == 0 || z == bmin[2] - R || z == bmax[2] + R ||
                                          y == bmin[1] - R || y == bmax[1] + R;
                        if (full) {
                            sA = cs[base + x0];
                            lA = cs[base + x1 + 1] - sA;
                        } else {
                            const int xa = bmin[0] - R, xb = bmax[0] + R;
                            if (xa >= 0) {
                                sA = cs[base + xa];
                                lA = cs[base + xa + 1] - sA;
                            }
                            if (xb < g[0]) {
                                sB = cs[base + xb];
                                lB = cs[base + xb + 1] - sB;
                            }
                        }
                    }
                    // exclusive prefix of the row lengths over the wave
                    const int32_t len = lA + lB;
                    int32_t inc = len;
#pragma unroll
                    for (int o = 1; o < 64; o <<= 1) {
                        const int32_t y2 = __shfl_up(inc, o, 64);
                        if (lane >= o) inc += y2;
                    }
                    off = inc - len;
                    tot = __builtin_amdgcn_readlane(inc, 63);
                    t0 = 0;
                    need_batch = false;
                    m = 0;
                    {
                        // every lane runs the search (the shuffles read all lanes)
                        const int32_t pi = ref_at(min(lane, max(tot - 1, 0)));
                        if (lane < tot) nxt = a.rpts[pi];
                    }
                }
                if (m == 0) {
                    // next round: 64 consecutive refs of the batch's stream, one per
                    // lane (the following round is prefetched)
                    if (t0 >= tot) {
                        need_batch = true;
                        continue;
                    }
                    m = __builtin_amdgcn_ballot_w64(t0 + lane < tot);
                    cp = nxt;
                    t0 += 64;
                    if (t0 < tot) {
                        const int32_t pi = ref_at(min(t0 + lane, tot - 1));
                        if (t0 + lane < tot) nxt = a.rpts[pi];
                    }
                }
                // score up to CH refs of the round, read from the lanes holding them
                double key[CH];
                int rid[CH];
                bool hit[CH];
                bool any = false;
#pragma unroll
                for (int u = 0; u < CH; ++u) {
                    const bool ok = m != 0;
                    const int src = ok ? (int)__builtin_ctzll(m) : 0;
                    m &= m - 1;
                    const double t0 = read_lane_f64(cp.u[0], src) - qu[0];
                    key[u] = t0 * t0;
#pragma unroll
                    for (int f = 1; f < D; ++f) {
                        const double tt = read_lane_f64(cp.u[f], src) - qu[f];
                        key[u] = fma(tt, tt, key[u]);
                    }
                    rid[u] = __builtin_amdgcn_readlane((int)cp.idx, src);
                    hit[u] = ok & (rid[u] < lim32) &
                             ((key[u] < tk) | ((key[u] == tk) & (rid[u] < ti)));
                    any = any | hit[u];
                }
                if (any) {
#pragma unroll
                    for (int u = 0; u < CH; ++u) {
                        if (hit[u]) {
                            bk[cnt][lane] = key[u];
                            bi[cnt][lane] = rid[u];
                            ++cnt;
                        }
                    }
                }
            }
        }
        int pad_out = K - a.k;
        asm volatile("" : "+v"(pad_out));
        if (jv) store_list<K>(ids, pad_out, a.out_idx + q * a.k);
    }
}

template <int D, int K>
void launch_coop_dk(const CoopArgs& c, unsigned blocks, hipStream_t st) {
    hipLaunchKernelGGL((coop_knn_kernel<D, K>), dim3(blocks), dim3(64 * kCoopWaves), 0, st, c);
}

void launch_coop(const CoopArgs& c, int d, unsigned blocks, hipStream_t st) {
    const int K = c.k <= 16 ? 16 : (c.k <= 32 ? 32 : 64);
    switch (d * 100 + K) {
        case 116: launch_coop_dk<1, 16>(c, blocks, st); break;
        case 132: launch_coop_dk<1, 32>(c, blocks, st); break;
        case 164: launch_coop_dk<1, 64>(c, blocks, st); break;
        case 216: launch_coop_dk<2, 16>(c, blocks, st); break;
        case 232: launch_coop_dk<2, 32>(c, blocks, st); break;
        case 264: launch_coop_dk<2, 64>(c, blocks, st); break;
        case 316: launch_coop_dk<3, 16>(c, blocks, st); break;
        case 332: launch_coop_dk<3, 32>(c, blocks, st); break;
        default: launch_coop_dk<3, 64>(c, blocks, st); break;
    }
}

int grid_cap(int64_t n_pts) {
    return (int)(2 * std::max<int64_t>(1, n_pts / kPointsPerCell) + 64);
}

// Query buckets of a grid with `cap` cells: tiling pads each tiled feature by at
// most 3 cells, which for the features it applies to (>= 8 cells) is <= 1.375x.
int grid_capq(int cap) { return 3 * cap + 64; }

void carve_grid(Carver& c, int64_t n_pts, GridMem* M) {
    const int cap = grid_cap(n_pts);
    const int64_t ns = (int64_t)cap + 2;
    M->cap = cap;
    M->params = c.take<GridParams>(1);
    M->cnt = c.take<int32_t>((size_t)ns);
    M->start = c.take<int32_t>((size_t)ns);
    M->fill = c.take<int32_t>((size_t)ns);
    M->sums = c.take<int32_t>((size_t)scan_blocks(ns) + 1);
    M->cell_of = c.take<int32_t>((size_t)n_pts + 1);
    M->pts = c.take<GridPt>((size_t)n_pts + 1);
}

// Grid over the first n_pts scaled refs of u64, with the bounding box of all refs.
void build_grid(const double* u64, int64_t n_pts, int d, const long long* bbox, const GridMem& M,
                hipStream_t st) {
    hipLaunchKernelGGL(grid_setup_kernel, dim3(1), dim3(64), 0, st, bbox, d, n_pts, M.cap,
                       grid_capq(M.cap), M.params);
    bin_points(u64, n_pts, d, 0, M.params, 0, M.cap + 1, M.cnt, M.start, M.fill, M.sums,
               M.cell_of, M.pts, st);
}

// Searches query rows [qa, qb) against grid M, which holds every ref the rows
// may reach.
void search_band(const GridMem& M, const double* qu64, int64_t qa, int64_t qb, int d, int k,
                 int has_cut, int64_t cut, int64_t n_ref, const GridSearchMem& S, int32_t* out_idx,
                 hipStream_t st) {
    const int64_t nq = qb - qa;
    if (nq <= 0) return;
    const int64_t nt = (int64_t)S.capq + 2;  // tiles (bound) + non-finite bucket + total
    timing_begin(TK_KNN_GRID_BUILD, st);
    bin_points(qu64 + qa * d, nq, d, qa, M.params, 1, S.capq + 1, S.qcnt, S.qstart, S.qfill,
               S.qsums, S.qcell_of, S.qpts, st);
    // work items: qcnt <- items per tile, qfill <- their exclusive prefix
    const unsigned gt = (unsigned)((nt + 255) / 256);
    hipLaunchKernelGGL(items_kernel, dim3(gt), dim3(256), 0, st, M.params, S.qstart, S.capq,
                       S.qcnt);
    const int64_t nsb = scan_blocks(nt);
    hipLaunchKernelGGL(scan_local_kernel, dim3((unsigned)nsb), dim3(256), 0, st, S.qcnt, nt,
                       S.qfill, S.qsums);
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(256), 0, st, S.qsums, (int)nsb);
    hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)nsb), dim3(256), 0, st, S.qfill, nt, S.qsums);
    hipLaunchKernelGGL(expand_items_kernel, dim3(gt), dim3(256), 0, st, S.qcnt, S.qfill, (int)nt,
                       S.qitem);
    timing_end(TK_KNN_GRID_BUILD, st);
    CoopArgs c{};
    c.gp = M.params;
    c.cell_start = M.start;
    c.rpts = M.pts;
    c.qpts = S.qpts;
    c.qstart = S.qstart;
    c.wp = S.qfill;
    c.item_tile = S.qitem;
    c.capq = S.capq;
    c.n_ref = n_ref;
    c.has_cut = has_cut;
    c.cut = cut;
    c.k = k;
    c.out_idx = out_idx;
    // waves stride over the work items (their count is known on the device only)
    const int64_t waves = std::min<int64_t>(8192, std::max<int64_t>(4, (nq + 31) / 32));
    timing_begin(TK_KNN_GRID, st);
    launch_coop(c, d, (unsigned)((waves + kCoopWaves - 1) / kCoopWaves), st);
    timing_end(TK_KNN_GRID, st);
}

}  // namespace

bool grid_supported(int d, int k) { return d >= 1 && d <= 3 && k >= 1 && k <= 64; }

int64_t grid_prefix_limit(int64_t n_ref, int k) {
    return std::max<int64_t>(std::max<int64_t>(64, 2 * (int64_t)k),
                             std::min<int64_t>(16384, n_ref / 16));
}

void carve_grid_refs(Carver& c, int64_t n_ref, GridRefMem* M) {
    M->bbox = c.take<long long>(8);
    carve_grid(c, n_ref, &M->full);
}

void carve_grid_search(Carver& c, int64_t n_ref, int64_t n_query, bool with_cut,
                       GridSearchMem* M) {
    *M = GridSearchMem{};
    // lower bands hold fewer than n_ref refs
    if (with_cut) carve_grid(c, n_ref, &M->band);
    const int capq = grid_capq(grid_cap(n_ref));
    const int64_t ns = (int64_t)capq + 2;
    M->capq = capq;
    M->qcnt = c.take<int32_t>((size_t)ns);
    M->qstart = c.take<int32_t>((size_t)ns);
    M->qfill = c.take<int32_t>((size_t)ns);
    M->qsums = c.take<int32_t>((size_t)scan_blocks(ns) + 1);
    M->qcell_of = c.take<int32_t>((size_t)n_query + 1);
    M->qpts = c.take<GridPt>((size_t)n_query + 1);
    // work items: at most one per 64 queries plus one per non-empty tile
    M->qitem = c.take<int32_t>((size_t)(n_query + n_query / 64 + 2));
}

int grid_build(const double* u64, int64_t n, int d, const GridRefMem& M, hipStream_t st) {
    if (d < 1 || d > 3) return NNGP_ERR_INVALID_ARGUMENT;
    timing_begin(TK_KNN_GRID_BUILD, st);
    hipLaunchKernelGGL(grid_bbox_init_kernel, dim3(1), dim3(64), 0, st, M.bbox);
    if (n > 0) {
        const unsigned gb = (unsigned)std::min<int64_t>(512, (n + 255) / 256);
        hipLaunchKernelGGL(grid_bbox_kernel, dim3(gb), dim3(256), 0, st, u64, n, d, M.bbox);
    }
    build_grid(u64, n, d, M.bbox, M.full, st);
    timing_end(TK_KNN_GRID_BUILD, st);
    NNGP_HIP_CHECK(hipGetLastError());
    return NNGP_OK;
}

int grid_search(const GridRefMem& G, const double* ru64, int64_t n_ref, int d,
                const double* qu64, int64_t n_query, int k, int has_cut, int64_t cut,
                int64_t row0, const GridSearchMem& M, int32_t* out_idx, hipStream_t st) {
    if (!grid_supported(d, k)) return NNGP_ERR_UNSUPPORTED;
    if (!has_cut) {
        search_band(G.full, qu64, 0, n_query, d, k, 0, 0, n_ref, M, out_idx, st);
        NNGP_HIP_CHECK(hipGetLastError());
        return NNGP_OK;
    }
    if (!M.band.params) return NNGP_ERR_WORKSPACE_TOO_SMALL;
    // Band j: rows whose limit q + cut lies in [L, 2L), L = L0 * 2^j; the top band
    // (2L >= n_ref) takes every row above and searches the full grid.
    const int64_t L0 = grid_prefix_limit(n_ref, k);
    for (int64_t L = L0;; L *= 2) {
        const bool top = 2 * L >= n_ref;
        int64_t qa = std::max<int64_t>(L - cut, row0);
        int64_t qb = top ? n_query : std::min<int64_t>(2 * L - cut, n_query);
        qa = std::min<int64_t>(qa, n_query);
        if (qb > qa) {
            const GridMem* gm = &G.full;
            if (!top) {
                timing_begin(TK_KNN_GRID_BUILD, st);
                build_grid(ru64, 2 * L, d, G.bbox, M.band, st);
                timing_end(TK_KNN_GRID_BUILD, st);
                gm = &M.band;
            }
            search_band(*gm, qu64, qa, qb, d, k, 1, cut, n_ref, M, out_idx, st);
        }
        if (top) break;
    }
    NNGP_HIP_CHECK(hipGetLastError());
    return NNGP_OK;
}

}  // namespace nngp
