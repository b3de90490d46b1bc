// fg_pixelwise_hbm.h -- launchers of the per-plane grain generation and the
// splat-from-HBM pixel-wise kernel (fg_pixelwise_hbm.hip).
#ifndef FG_PIXELWISE_HBM_H
#define FG_PIXELWISE_HBM_H

#include "fg_kernels.h"

namespace fg {

constexpr int kBinSide = 16;  // a bin is 16 x 16 cells (one wave; a lane holds four cells)
// The draw pass (bin_grains) gives each wave kDrawChunk consecutive bins of
// a plane: the wave's set-up (its arguments, and in a batch the search for
// its plane) is spread over the chunk's bins (on a batch of 64 uneven
// images, 4 or 16 bins a wave are 0.3% and 0.5% slower than 8).
constexpr int kDrawChunk = 8;

// The plane's extended cell grid: every cell some point's window can reach.
struct PixGeom {
    int Ix0, Jy0;  // first cell
    int W, H;      // cells per row / column (each <= 65535)
    int nbx, nby;  // bins per row / column
};

// The N samples ordered by the fractional part of their scaled x offset
// X = ox * zoom, with a bucket index over that fraction.  For a grain, the
// samples whose candidate box has a column in it are those whose fraction
// lies in one cyclic interval, so they form one cyclic run of this order.
struct SampleTable {
    const float2 *xy;      // (ox * zoom, oy * zoom), sorted by frac(ox * zoom), ties by index
    const float2 *o;       // (ox, oy) in input pixels, same order
    const uint16_t *k;     // sample index, same order
    const uint16_t *qidx;  // qbuckets + 1 entries: first position whose fraction >= q / qbuckets
    const float4 *ext;     // (min X, max X, min Y, max Y) of the scaled offsets
    const float *key;      // the fractions, same order
    const uint16_t *yord;  // the sorted positions in (X.y, position) order (filled for n <= 2048)
    const uint16_t *ybk;   // kRowBuckets + 1 entries: first index of yord whose row bucket is >= b
    int n;                 // samples in the table
    int qbuckets;          // power of two
    float eps;             // widening of the fraction interval (covers f32 rounding of the box)
    float safe_r;          // a grain of radius <= safe_r that passes the disk test of a point
                           // lies in that point's cell window
    // the splat's candidate lists over this order (launch_splat_lists), or
    // nullptr: (start, length) per (radius class, cell), the entries (sorted
    // position * 16) of each list in order of the scaled y offset, and per
    // cell kRowBuckets counts: entry b - 1 is the list's entries in row
    // buckets below b (row_bucket; entry kRowBuckets - 1 is its length)
    const uint2 *cl_cell;
    const uint16_t *cl_list;
    const uint16_t *cl_yidx;
};

// Row buckets of a table's (X.y, position) order and of its candidate
// lists: 16 intervals of the scaled y offsets X.y, from their extent.
constexpr int kRowBuckets = 16;

// Bytes of the device buffer behind a SampleTable for n samples.
size_t sample_table_bytes(int n, int qbuckets);
// Builds the table in `buf` (sample_table_bytes bytes) and points *t at it.
hipError_t launch_sample_table(const float2 *offs, int n, float zoom, int qbuckets, void *buf, SampleTable *t,
                               hipStream_t s);
// The table of the first n samples of `full` (full.n <= 2048), built by
// filtering full's order instead of sorting again.
hipError_t launch_sample_prefix(const SampleTable &full, int n, int qbuckets, void *buf, SampleTable *t,
                                hipStream_t s);

// Candidate lists of the HBM splat (PixelArgs::cl_*): q * q cells per class,
// class k of radius (k + 1) / (2 q).  e is an integer >= (2 st.eps + 2^-11)
// q; list needs splat_lists_entries(q, classes, e, st.n) entries, cell
// classes * q * q (start, length) pairs, yidx classes * q * q * kRowBuckets
// counts, fill classes counters (cleared here).
// A class beyond its room sets kStatusCapacity in *status.
size_t splat_lists_entries(int q, int classes, int e, int n);
// Writes t to device memory at dst (the single-plane splat reads its table
// from PixelArgs::level_tables).
hipError_t launch_store_table(const SampleTable &t, SampleTable *dst, hipStream_t s);
hipError_t launch_splat_lists(const SampleTable &st, int q, int classes, int e, uint32_t *fill, uint2 *cell,
                              uint16_t *yidx, uint16_t *list, int *status, hipStream_t s);

// Also sets kStatusLambda in *a.status for lambda values outside the accepted range.
hipError_t launch_threshold(const PixelArgs &a, double *T, hipStream_t s);
// bin_mask: four 64-bit words per bin, the nonempty cells (bin_grains reads them).
hipError_t launch_bin_count(const PixelArgs &a, const PixGeom &g, const double *T, uint32_t *bin_count,
                            unsigned long long *bin_mask, hipStream_t s);
// hipCUB exclusive sum; with temp == nullptr only sets *temp_bytes.
hipError_t scan_bins(void *temp, size_t *temp_bytes, const uint32_t *in, uint32_t *out, int n, hipStream_t s);
// The planes of one frame (same seed and geometry): their thresholds, and
// where their bin counts and nonempty masks go.
constexpr int kFramePlanes = 3;
struct FramePlanes {
    const double *T[kFramePlanes];
    uint32_t *cnt[kFramePlanes];
    unsigned long long *mask[kFramePlanes];
};
// bin_count for n_planes (1..3) planes in one pass: each cell hashed once
// (the planes' cell streams are the same), counted against every plane's
// threshold.
hipError_t launch_bin_count_planes(const PixelArgs &a, const PixGeom &g, const FramePlanes &fp, int n_planes,
                                   hipStream_t s);
hipError_t launch_bin_grains(const PixelArgs &a, const PixGeom &g, int dist_kind, const double *T,
                             const uint32_t *bin_start, const unsigned long long *bin_mask, float4 *grains,
                             hipStream_t s);
// Grain-wise integrator over the sorted sample order (fg_grainwise.hip); the
// table is built from the output-pixel offsets with zoom 1.  Needs N <= 2048
// and tiles of at most 128 x 256.
size_t grainwise_sorted_lds(int tile_w, int tile_h, int n_samples, int qbuckets);
// Static LDS of the sorted grain-wise kernel per workgroup of `block` threads.
size_t grainwise_sorted_static_lds(int block);
// Per input pixel: grain count and stream state after the count draws
// (a.pix_q, a.pix_s01, a.pix_s23), read by the sorted grain-wise kernel.
hipError_t launch_pixel_state(const GrainArgs &a, hipStream_t s);
hipError_t launch_grainwise_sorted(const GrainArgs &a, const SampleTable &st, int dist_kind, int block,
                                   hipStream_t s);
// Candidate lists of the sorted grain-wise kernel for one constant output
// radius r_out (GrainArgs::cell_*): q x q cells, start (q * q + 1 entries),
// list (room for q * q * n positions); c_base = max(out_w, out_h) + the
// largest output radius + 8 (the kernel's widening bound without the offsets).
hipError_t launch_cell_lists(const SampleTable &st, int n, int q, float r_out, float c_base, uint32_t *start,
                             uint16_t *list, hipStream_t s);

// Heaviest-first tile order: w[t] = the work of tile t (grains of the bins
// covering its region), ids[t] = t; sort_tiles orders ids by w, descending
// (hipCUB radix sort; with temp == nullptr only sets *temp_bytes).
hipError_t launch_pixel_tile_weights(const PixelArgs &a, const PixGeom &g, const uint32_t *bin_start, uint32_t *w,
                                     int *ids, hipStream_t s);
hipError_t sort_tiles(void *temp, size_t *temp_bytes, const uint32_t *w_in, uint32_t *w_out, const int *ids_in,
                      int *ids_out, int n, hipStream_t s);
// Grain-wise: w[t] = grains of the input pixels whose grains can reach tile t.
hipError_t launch_grain_tile_weights(const GrainArgs &a, const SampleTable &st, uint32_t *w, int *ids,
                                     hipStream_t s);

// One plane of a batch: everything its splat tiles read.
struct BatchPlane {
    PixelArgs a;
    PixGeom geo;
    SampleTable st;
    const uint32_t *bin_start;  // this plane's bins inside the batch's bin_start
    const float4 *grains;       // the batch's grain array (bin_start holds batch offsets)
    long long item_base;        // first (plane, tile) item of this plane
};

// The generation of a batch's planes: the thresholds, the draws and the
// sample tables each one launch over all planes, in which a workgroup (or
// wave) finds its plane by a binary search over first-block offsets (one
// entry more than the planes).  In the draw pass a wave takes kDrawChunk
// consecutive bins of one plane (the offsets count chunks).  Every pointer
// is device memory.
struct BatchGen {
    const BatchPlane *planes;     // n planes: their a, geo, st and bin_start
    int n;
    const long long *thr_off;     // first threshold workgroup of each plane (kBlock pixels each)
    long long thr_blocks;
    const long long *bin_off;     // first bin of each plane in cnt and mask
    long long total_bins;
    double *const *T;             // each plane's thresholds (in_w x in_h)
    // the planes of each radius distribution (0 constant, 1 lognormal) and
    // their first chunks in that sequence (n_kind[k] + 1 entries)
    const int *kind_plane[2];
    const long long *kind_off[2];
    int n_kind[2];
    long long kind_chunks[2];     // chunks of the planes of each distribution
    uint32_t *cnt;
    unsigned long long *mask;
    float4 *grains;
    // sample tables (each plane's st): the candidate lists' plan per plane,
    // (q, classes, e, 0) with classes == 0 for a plane without lists, their
    // counters, and the first splat-lists workgroup of each plane
    const int4 *list_plan;
    uint32_t *const *list_fill;
    const long long *list_off;
    long long list_blocks;
    int sort_blocks;              // workgroups per plane of the sample sort: ceil(max n_samples / kBlock)
};
// Thresholds of every plane (sets kStatusLambda as launch_threshold does).
hipError_t launch_batch_threshold(const BatchGen &bg, hipStream_t s);
// bin_grains of every plane (bin_start of each plane is the batch scan's).
hipError_t launch_batch_bin_grains(const BatchGen &bg, hipStream_t s);
// Every plane's sample table and candidate lists.
hipError_t launch_batch_sample_tables(const BatchGen &bg, hipStream_t s);

// Adaptive sampling: tile_level[t] = first level (level_n ascending) covering
// the largest budget in tile t (a.budget).
hipError_t launch_tile_levels(const PixelArgs &a, const int *level_n, int n_levels, uint8_t *tile_level,
                              hipStream_t s);
// Pointers of a sample table laid out in buf (as launch_sample_table does).
void sample_table_layout(void *buf, int n, int qbuckets, SampleTable *t);
// The plane records of one frame, written to device memory by a kernel
// (they travel as its argument: no host buffer has to outlive a copy).
struct FramePlaneRecords {
    BatchPlane p[kFramePlanes];
};
hipError_t launch_store_frame_planes(const FramePlaneRecords &r, int n, BatchPlane *dst, hipStream_t s);
// Resident workgroups of the persistent splat kernel per CU.
int splat_batch_blocks_per_cu(int block, int lds_bytes);
// Persistent splat over (plane, tile) items in launch order (items[i] =
// (plane, tile)); the counter must be 0 at launch.
hipError_t launch_splat_batch(const BatchPlane *planes, const int2 *items, int n_items, unsigned int *counter,
                              int grid, int block, int lds_bytes, hipStream_t s);

// Graph-safe batch helpers (kernels only, no library calls; fg_batch.hip).
// Exclusive scan of n u32 values (n + 1 outputs, out[n] = total) with a
// caller-provided partials buffer of scan_partials(n) entries.
size_t scan_partials(long long n);
hipError_t launch_scan_u32(const uint32_t *in, uint32_t *out, long long n, uint32_t *partials, hipStream_t s);
// Fill n u32 words with v.
hipError_t launch_fill_u32(uint32_t *p, long long n, uint32_t v, hipStream_t s);
// Per item: the work of (plane, tile) as in launch_pixel_tile_weights.
hipError_t launch_batch_weights(const BatchPlane *planes, int n_planes, uint32_t *w, int n_items, hipStream_t s);
// items in heaviest-first order (buckets of floor(log2(w + 1)), descending);
// scratch holds 2 * 64 u32.
hipError_t launch_batch_order(const BatchPlane *planes, int n_planes, const uint32_t *w, int n_items, int2 *items,
                              uint32_t *scratch, hipStream_t s);
// *out += the sum of n u32 values, in 64 bits (the caller clears *out): a
// plane's or batch's grain total from its bin counts.
hipError_t launch_sum_u64(const uint32_t *in, long long n, unsigned long long *out, hipStream_t s);
// Sets kStatusCapacity in *status when the batch's grain total (the 64-bit
// sum of its bin counts) exceeds cap.
hipError_t launch_capacity_check(const unsigned long long *total, uint32_t cap, int *status, hipStream_t s);
// Sets kStatusLambda in *status when a lambda value is NaN, +inf or above
// kMaxLambda (for the formulations whose kernels read lambda directly).
hipError_t launch_lambda_check(const float *lambda, long long n, int *status, hipStream_t s);

// LDS bytes the splat kernel needs beyond its static arrays for one tile.
size_t splat_hbm_lds(int tile_w, int tile_h, int n_samples, int qbuckets);
// Static LDS of the splat kernels per workgroup of `block` threads: per wave a
// candidate queue and 64 grain records.
inline size_t splat_hbm_static_lds(int block) { return (size_t)block / 64 * (kQueueCap * 4 + 64 * 16) + 64; }
hipError_t launch_splat_hbm(const PixelArgs &a, const PixGeom &g, const SampleTable &t, const uint32_t *bin_start,
                            const float4 *grains, int block, hipStream_t s);

}  // namespace fg

#endif
