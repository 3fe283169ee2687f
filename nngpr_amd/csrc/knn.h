// Internal interface of the kNN module (knn.hip, knn_grid.hip).
#pragma once
#include "common.h"

namespace nngp {

// ---------------------------------------------------------------------------
// Uniform-grid exact search (d <= 3, k <= 64), knn_grid.hip
// ---------------------------------------------------------------------------
// One point of a cell-sorted array: scaled coordinates (unused features 0) and
// the point's row index.
struct alignas(32) GridPt {
    double u[3];
    long long idx;
};

// Grid geometry, computed on the device from the bounding box of the finite refs.
// Cell of a point: c_f = clamp(floor((u_f - lo_f) * inv_h), 0, g_f - 1); linear id
// c_0 + g_0 * (c_1 + g_1 * c_2) (refs are stored in this order).  Features >= d
// have g_f = 1.  Queries are ordered by tiles of t_0 x t_1 x t_2 cells (tg_f tiles
// per feature, n_qb buckets), so that consecutive queries are spatially compact.
struct GridParams {
    double lo[3];
    double h;
    double inv_h;
    double tol;       // slack subtracted from every distance lower bound
    // A finite point with a coordinate outside [in_lo, in_hi] is kept out of the
    // cells: the "outside" bucket of refs, keyed by every search, or the far
    // bucket of queries.  The box is the grid's own extent when a few far points
    // would otherwise stretch the grid (see setup_grid), else unbounded.
    double in_lo[3];
    double in_hi[3];
    int g[3];
    int n_cells;
    int t[3];
    int tg[3];
    int n_qb;
};

// A grid over a set of refs and the refs sorted by cell.  Bucket `cap` collects the
// finite points outside the grid's box, bucket `cap + 1` the points with a
// non-finite coordinate (never neighbours); start has cap + 3 entries.
struct GridMem {
    GridParams* params;
    int32_t* cnt;
    int32_t* start;
    int32_t* fill;
    int32_t* sums;
    int32_t* cell_of;
    GridPt* pts;
    // occupancy bits: row (c_1, c_2) holds ceil(g_0 / 64) words, bit x of word
    // x / 64 set when cell x of the row has a ref
    uint64_t* rowmask;
    int cap;
};

// Ref side, kept with the refs: bounding box and the grid over all refs.
struct GridRefMem {
    long long* bbox;
    GridMem full;
};

// Query buckets of one band (tile order over the band's grid; bucket capq holds the
// finite queries outside the grid's box, capq + 1 the non-finite ones).
struct GridQueryBins {
    int32_t* cnt;
    int32_t* start;
    int32_t* fill;
    int32_t* sums;
    int capq;
};

constexpr int kMaxGridBands = 28;

// Per search: grids over prefixes of the refs (the lower Vecchia bands, band j
// holding refs [0, 2 L0 2^j)), the query buckets of every band (band n_lower is
// the top band, searched on the full grid; without a cut it is the only band),
// and the queries sorted band by band.
struct GridSearchMem {
    int n_lower;
    GridMem band[kMaxGridBands];
    GridQueryBins qb[kMaxGridBands + 1];
    int32_t* qcell_of;
    GridPt* qpts;
    // search waves (64 queries each) in launch order, by estimated cost, and
    // their costs
    int32_t* wave_cost;
    int32_t* wave_perm;
    // queries the lane-per-query search queues for an exact search: (band,
    // position) pairs, at most one per query, and their count
    int32_t* redo;
    int32_t* redo_n;
};

bool grid_supported(int d, int k);
// Vecchia rows whose ref limit is below this are scanned brute force.
int64_t grid_prefix_limit(int64_t n_ref, int k);
void carve_grid_refs(Carver& c, int64_t n_ref, GridRefMem* M);
void carve_grid_search(Carver& c, int64_t n_ref, int64_t n_query, int k, bool with_cut,
                       GridSearchMem* M);
// Builds the grid over n scaled points (u64, n x d AoS).
int grid_build(const double* u64, int64_t n, int d, const GridRefMem& M, hipStream_t st);
// Exact top-k of query rows [row0, n_query) (rows below have limits under
// grid_prefix_limit and are scanned elsewhere; row0 = 0 without a cut).
int grid_search(const GridRefMem& G, const double* ru64, int64_t n_ref, int d,
                const double* qu64, int64_t n_query, int k, int has_cut, int64_t cut,
                int64_t row0, const GridSearchMem& M, int32_t* out_idx,
                int64_t* rows_recomputed, hipStream_t st);

// ---------------------------------------------------------------------------
// Ref-side state of the kNN search
// ---------------------------------------------------------------------------
// Search layout of a ref set, chosen from d, the precision and the process-wide
// method when the refs are prepared:
//   kRefGrid: the uniform grid (AUTO, d <= 3);
//   kRefMfma: the bf16x3 matrix-core screen's operand fragments (FP32_REFINED,
//             d >= 4, unless the method is BRUTE_FORCE_VALU);
//   kRefScan: the blocked layouts of the VALU scans (FP32 screen or FP64 exact).
enum RefLayout { kRefScan = 0, kRefGrid = 1, kRefMfma = 2 };
int refs_layout(int d, int precision);

// Scaled coordinates (AoS), and the layout the search reads, with the screen
// centre and max radius.
struct KnnRefs {
    const double* u64;
    const double* s64;
    const float* s32;
    const uint16_t* frag;        // kRefMfma: ref operand fragments
    const double* centre;        // screen centre (mean of the finite scaled refs)
    const unsigned int* smax;
    GridRefMem grid;
    int layout;                  // RefLayout
    int use_grid;                // layout == kRefGrid
    int64_t n;
    int d;
    int precision;
    int n_scale;
    double scale[NNGP_MAX_DIM];
};

// Process-wide search method: NNGP_KNN_AUTO or NNGP_KNN_BRUTE_FORCE.  Read when
// refs are prepared.
int knn_method();
int knn_set_method(int method);

size_t knn_refs_bytes(int64_t n_ref, int d, int precision);
int knn_prepare_refs(const double* ref, int64_t n_ref, int d, const double* scale, int n_scale,
                     int precision, void* mem, size_t bytes, KnnRefs* out, hipStream_t st);
// with_cut: budget for searches with a Vecchia cut (nngp_knn); the serving
// model's query searches have none.  layout: the refs' RefLayout (only the
// buffers of the search that layout runs are budgeted).
size_t knn_search_bytes(int64_t n_ref, int64_t n_query, int d, int k, int precision, bool same,
                        bool with_cut, int layout);
// query == NULL searches the ref set itself (Vecchia self-search).
int knn_search(const KnnRefs& R, const double* query, int64_t n_query, int k, int has_cut,
               int64_t ref_cut, int32_t* out_idx, void* ws, size_t ws_bytes, WsHeader* hdr,
               hipStream_t st);

size_t knn_workspace_bytes(int64_t n_ref, int64_t n_query, int d, int k, int precision,
                           bool same);

// Exact recomputation of listed query rows: rows[0 .. *count), each with an upper
// bound pair (bound_key, bound_idx) of its true k-th (key, index) pair (a pair of
// an actual ref).  One wave per row keeps every pair at or below the bound and
// ranks them; rows keeping more than 1,024 pairs go to the serial per-row kernel
// (over_rows, scratch lkey / lidx of exact_rows_scratch_count(k) entries).
struct ExactRows {
    const int64_t* rows;
    const unsigned long long* count;
    const double* bound_key;
    const int32_t* bound_idx;
    int64_t* over_rows;              // n_query entries
    unsigned long long* over_count;
    double* lkey;
    int32_t* lidx;
};
size_t exact_rows_scratch_count(int k);
void launch_exact_rows(const double* ru64, const double* qu64, int d, int64_t n_ref,
                       int64_t n_query, int has_cut, int64_t cut, int k, const ExactRows& er,
                       int32_t* out_idx, hipStream_t st);

// ---------------------------------------------------------------------------
// bf16x3 matrix-core screen of the brute-force search, d >= 4 (knn_mfma.hip)
// ---------------------------------------------------------------------------
bool mfma_screen_supported(int d);
// Candidates per row for k neighbours (k + 8 to k + 23, one of 24, 40, 56, 72).
int mfma_screen_K(int k);
// bf16 entries of the operand fragments of n points (two padding tiles included).
size_t mfma_frag_elems(int64_t n, int d);
// Relative error bound of a screen score for the re-rank certificate.
double mfma_screen_eps(int d);
// Operand fragments of points [0, n) from their scaled coordinates u64: role 0 =
// refs (also raises *smax, the largest centred norm), role 1 = queries (also
// writes p2 = |u - centre|^2).
void mfma_prep(const double* u64, int64_t n, int d, const double* centre, int role,
               uint16_t* frag, double* p2, unsigned int* smax, hipStream_t st);
// Screen of n_query rows (K candidates each): writes cand / thr and returns 0, or,
// with ref stripes (spk given and the rows too few to fill the chip), writes
// sorted per-stripe lists to spk and returns their count (a stripe merge
// follows); -1 if K or d is not supported.  spk holds at least
// min(8 n_query, 131072) x K entries.
int launch_mfma_screen(const uint16_t* rfrag, int64_t n_ref, const uint16_t* qfrag,
                       int64_t n_query, int d, int K, int has_cut, int64_t cut, int32_t* cand,
                       float* thr, uint64_t* spk, hipStream_t st);

// ---------------------------------------------------------------------------
// Exact lists at a new scale from exact lists at an old one (knn_rescale.hip)
// ---------------------------------------------------------------------------
int rescale_max_candidates();
size_t rescale_state_bytes(int64_t n_ref, int64_t n_query, int d, int k, int K);
int rescale_init(const double* ref, int64_t n_ref, const double* query, int64_t n_query, int d,
                 const double* scale, int n_scale, const int32_t* cand, int K, int k, void* state,
                 size_t state_bytes, hipStream_t st);
int rescale_run(const double* ref, int64_t n_ref, const double* query, int64_t n_query, int d,
                const double* scale_old, const double* scale_new, int n_scale, int has_cut,
                int64_t ref_cut, const int32_t* cand, int K, int k, int32_t* out_idx, void* state,
                size_t state_bytes, hipStream_t st);
int knn_run(const double* ref, int64_t n_ref, const double* query, int64_t n_query, int d,
            const double* scale, int n_scale, int k, int has_cut, int64_t ref_cut, int precision,
            int32_t* out_idx, void* ws, size_t ws_bytes, hipStream_t st, bool reset_stats);

}  // namespace nngp
