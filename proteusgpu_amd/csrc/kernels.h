// Host-side launchers of the libpgpu kernels (internal interface).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pg {

// ---------------------------------------------------------------- scan.hip
// Exclusive prefix sum of n int32 values; writes total (device int32) when
// total != nullptr.  tmp must hold scan_tmp_elems(n) int32.
int64_t scan_tmp_elems(int64_t n);
void scan_exclusive_i32(const int32_t* in, int32_t* out, int64_t n, int32_t* tmp, int32_t* total,
                        hipStream_t s);

// ---------------------------------------------------------------- grid.hip
// periodic_mesh.cpp:63-112: number of ghost regions each point falls in.
void launch_ghost_count(int dim, const double* pts, int64_t n, double buff, int32_t* cnt,
                        hipStream_t s);
// periodic_mesh.cpp:50-65,119-126: write scaled points, hydro first then the
// ghosts in (point, region) order; ghost_src[g] = source hydro index.
void launch_ghost_write(int dim, const double* pts, int64_t n, double buff, const int32_t* off,
                        double* out, int64_t* ghost_src, hipStream_t s);
// knn.cpp:119-222: grid cell of every point and per-cell counts.
void launch_cell_count(int dim, const double* pts, int64_t n, int N, int32_t* cell,
                       int32_t* counts, hipStream_t s);
// Scatter point indices into their cells' ranges (any order within a cell).
void launch_cell_scatter(int64_t n, const int32_t* cell, const int32_t* starts, int32_t* fill,
                         uint32_t* idx, hipStream_t s);
// Sort every cell's range ascending: the order of the reference's serial
// store loop (knn.cpp:187-204), i.e. stable by input index.  big (ncell
// int32) and nbig (one int32) are scratch for ranges longer than 32.
void launch_cell_sort(int64_t ncell, const int32_t* starts, const int32_t* counts, uint32_t* idx,
                      int32_t* big, int32_t* nbig, hipStream_t s);
// out = in displaced by at most amp per axis, deterministically from each
// point's index and salt (the degeneracy fallback of a mesh build); periodic:
// coordinates wrap into [0, 1), otherwise they are clamped into it.
void launch_displace(int dim, const double* in, int64_t n, double amp, uint64_t salt, int periodic, double* out,
                     hipStream_t s);
// sorted[s] = pts[perm[s]]
void launch_gather_points(int dim, int64_t n, const uint32_t* perm, const double* pts,
                          double* sorted, hipStream_t s);

// Sub-grids of the crowded grid cells (clustered inputs): a cell holding
// more than SUB_MIN points is split m ways per axis (m^dim sub-cells of about
// 3 points), and its points (sorted ids starts[c] .. starts[c] + counts[c])
// are listed again by sub-cell in idx[starts[c] ..].  m[c] = 0: no
// sub-grid; else the cell's m^dim sub-cells are start/cnt[ptr[c] + j]
// (first position in idx, count), j = a + b m (+ c m^2) for sub-cell
// (a, b, c).  Searches use it only to skip parts of a cell their ball
// misses; the sorted order, and with it every row, is unchanged.
constexpr int SUB_MIN = 32;
struct SubGrid {
    const int32_t* m = nullptr;
    const int32_t* ptr = nullptr;
    const int32_t* start = nullptr;
    const int32_t* cnt = nullptr;
    const uint32_t* idx = nullptr;
};
// Build it after launch_cell_sort (whose list of cells longer than 32 it
// takes) and launch_gather_points.  m, ptr: ncell int32; start, cnt: cap
// int32 (cap >= n / 3 + 64 always suffices); idx: n uint32; alloc: one
// int32 of scratch.
void launch_subgrid(int dim, const double* sorted, int64_t n, int N, int64_t ncell, const int32_t* starts,
                    const int32_t* counts, const int32_t* big, const int32_t* nbig, int32_t* m, int32_t* ptr,
                    int32_t* start, int32_t* cnt, int64_t cap, uint32_t* idx, int32_t* alloc, hipStream_t s);

// ----------------------------------------------------------------- knn.hip
// A KNN launch for seeds whose clip will resume: their rows need only the
// neighbours after the saved row length (resume[s] & 0xffff, the last of
// them lastnb[s]); others get full rows.  tag: the shape tag the clip
// kernel resumes under.  resume == nullptr: full rows for every seed.
struct KnnTail {
    const int32_t* resume = nullptr;
    const uint32_t* lastnb = nullptr;
    int32_t tag = 0;
};

// k nearest neighbours of queries (sorted ids; nullptr = all n points).
// Rows ascending by (FP64 dist^2, original index); out row stride k.
void launch_knn(int dim, const double* sorted, const uint32_t* perm, int64_t n, int N,
                const int32_t* starts, const int32_t* counts, const uint32_t* queries, int64_t nq,
                int k, uint32_t* out, hipStream_t s, KnnTail tail = {}, SubGrid sg = {});

// ------------------------------------------------------------- voronoi.hip
struct FaceArrays {
    // face-slot-major padded storage: entry (f, cell s) at f * stride + s
    int64_t stride;
    int32_t* nbr;      // sorted neighbour id, -1 = box plane
    double* area;
    double* nrm;       // 3 planes of stride*maxf doubles: x | y | z
    double* cen;       // face centroid, same layout
    uint8_t* count;    // faces per cell
    double* vol;       // cell volume
    double* ccen;      // cell centroid, 3 planes of stride doubles
    int32_t* status;   // PG status per cell
    int32_t* kcert;    // 1 + index of the certifying neighbour, 0 if none
    // Certificate restarts resume where the previous pass stopped: the fast
    // kernel saves an uncertified cell (planes, vertex planes, adjacency;
    // CLIP_STATE_BYTES per sorted seed) and sets resume[s] = shape tag | row
    // length, else 0.  nullptr: nothing is saved or resumed.
    uint8_t* cstate = nullptr;
    int32_t* resume = nullptr;
    uint32_t* lastnb = nullptr;  // saved seeds: the last neighbour of the row (sorted id)
    // exact-predicate runs: tie-break key per sorted id (the input index;
    // nullptr: the sorted id itself)
    const uint32_t* key = nullptr;
    // optional cell-major copy of nbr (FACE_SLOTS per cell), for the face
    // ownership of the hydro update
    int32_t* nbr_cm = nullptr;
};


// Face slots per cell in the padded face arrays (the reference's _MAX_P_).
constexpr int FACE_SLOTS = 64;

// Saved clip state of one seed (64 planes, up to 64 vertex slots).
constexpr int64_t CLIP_STATE_BYTES = 3328;

// Clip + certify + extract the seeds listed in `seeds` (sorted ids; nullptr
// = all n) with their neighbour rows knn[row q] (row q belongs to seeds[q]).
// big != 0 selects the 128-slot kernel; restart launches the fast kernel
// under the name k_voronoi_restart (profiles tell the passes apart).
// lanes x slots: the fast kernel's shape (valid_clip_config), taken once
// per build by the caller so that every launch and the KNN tail tag of a
// build agree.  exact: the exact-predicate kernel (exact.h; 128 slots, no
// saved cells), for seeds whose filtered predicate was degenerate.
void launch_voronoi(int dim, const double* sorted, int64_t n, const uint32_t* seeds, int64_t nq,
                    const uint32_t* knn, int k, int max_p, int max_t, int big, FaceArrays fa, int lanes, int slots,
                    hipStream_t s, bool restart = false, bool exact = false);

// Vertex-ball completion (k_voronoi_ball): the seeds listed, uncertified at
// the longest row, clipped by their rows of length k and then by the points
// their vertex balls hold until every ball is empty (kcert = kmark); grid
// tables of the sorted points (N per axis, starts, counts) and the input
// index per sorted id.
void launch_voronoi_ball(int dim, const double* sorted, int64_t n, const uint32_t* seeds, int64_t nq,
                         const uint32_t* knn, int k, int max_p, int max_t, FaceArrays fa, int N,
                         const int32_t* starts, const int32_t* counts, const uint32_t* perm, int kmark,
                         hipStream_t s, SubGrid sg = {});

// Process-wide default shape of the fast clip kernel for the following
// builds: lanes per seed (8 or 16) and vertex slots per seed (48 or 64).
bool valid_clip_config(int lanes, int slots);
void set_clip_config(int lanes, int slots);
int clip_lanes();
int clip_slots();
// Certificate restarts resume saved cells (true) or clip from scratch.
bool clip_resume();
void set_clip_resume(bool on);
// Collect the sorted ids whose status is not SUCCESS (order: ascending id).
// nfail (device int32) receives the count; tmp as for scan.
void launch_collect_failed(int64_t n, const int32_t* status, uint32_t* list, int32_t* flags,
                           int32_t* offs, int32_t* tmp, int32_t* nfail, hipStream_t s);
void launch_collect_status(int64_t nq, const uint32_t* seeds, const int32_t* status, int code,
                           uint32_t* list, int32_t* flags, int32_t* offs, int32_t* tmp,
                           int32_t* nout, hipStream_t s);

// ---------------------------------------------------------- learned_k.hip
// Classifier of the neighbour count: 16 features -> 64 -> 64 -> 6 buckets.
constexpr int LK_NFEAT = 16, LK_NCLASS = 6;
constexpr int LK_NPARAM = 16 * 64 + 64 + 64 * 64 + 64 + 64 * 6 + 6;
int lk_bucket(int c);  // 16, 24, 32, 48, 64, 96
// Features of every seed (sorted space) into out[16 * (out_index ? out_index[s] : s)]:
//   f0..f3  log2(1 + x), x = mean count per in-grid cell of the own cell /
//           face / edge / corner neighbours of the 3^dim cube, over the mean
//           occupancy (the logarithm keeps clustered inputs, x up to ~1e3,
//           in the range the bfloat16 network trains on)
//   f4      in-grid fraction of the cube
//   f5..f7  |2u - 1|, u = position inside the own cell per axis (0 in 2D z)
//   f8,f9   distance to the 4th / 8th nearest point of the cube, in cell
//           widths (2 if absent);  f10 the nearest;  f15 = f9 / f8
//   f11     log2(1 + cube density / mean density)
//   f12     min(distance to the unit box faces in cell widths, 3) / 3
//   f13     log2(1 + (max - min neighbour count) / mean occupancy)
//   f14     log2(1 + |sum of count * offset| / (3^dim * mean occupancy))
// sg: the grid's sub-grids (crowded cells are scanned through them; the
// features are the same as a full scan's)
void launch_kfeatures(int dim, const double* sorted, int64_t n, int N, const int32_t* counts,
                      const int32_t* starts, const uint32_t* out_index, float* out, hipStream_t s,
                      SubGrid sg = {});
// cls[s] = bucket with the least expected cost of its restart chain (a pass
// at row length k costs k + r m, m = clip steps, r = a0); logits
// (n x 6, optional) receives the network outputs.
// n_dev (optional): seed count on the device, n then is the capacity.
void launch_kpredict(const float* X, int64_t n, const float* params, float a0, const uint32_t* out_index,
                     int32_t* cls, float* logits, hipStream_t s, const int64_t* n_dev = nullptr);
// Features of batched boxes (sorted seeds of every box, box-major).
void launch_bkfeatures(int dim, const double* sorted, const int64_t* n_dev, int64_t cap, const int32_t* sbox,
                       const int64_t* soff, const int32_t* N, const int64_t* coff, const int32_t* counts,
                       const int32_t* starts, float* out, hipStream_t s);
int64_t lk_grad_blocks(int64_t B);
// Forward + backward of a batch of B rows (seed != 0: rows drawn from n by a
// hash of (seed, step, i)); part: lk_grad_blocks(B) x LK_NPARAM partial sums.
void launch_kgrad(const float* X, const int32_t* y, int64_t n, int64_t B, uint64_t seed, int64_t step,
                  const float* params, float* part, float* part_loss, hipStream_t s);
// Reduce the partials in block order; Adam step when lr > 0.
void launch_kadam(const float* part, const float* part_loss, int nblk, float* params, float* m, float* v,
                  float* grad_out, float* loss_out, int64_t step, float lr, hipStream_t s);
// Stable split of seeds 0..n-1 (n = *n_dev when n_dev != nullptr, else
// n_cap) by class: class c into lists + c * list_cap (ascending seed id, or
// in the sequence order[0..n) when order != nullptr), its size into
// counts[c] (device int64).  cnt, off: NC * lk_bucket_blocks int32; tmp as
// for scan of that size.
int64_t lk_bucket_blocks(int64_t n_cap);
void launch_kbucket(const int32_t* cls, const int64_t* n_dev, int64_t n_cap, int32_t* cnt, int32_t* off,
                    int32_t* tmp, int32_t* total, int64_t list_cap, uint32_t* lists, int64_t* counts,
                    hipStream_t s, const uint32_t* order = nullptr);
// Clip order of the seeds of a learned build: list[0..n) (list == nullptr:
// the seeds 0..n-1; n = *n_dev when n_dev != nullptr, else n_cap) into out,
// sorted stably by the previous certified k of each seed's point (o =
// perm[s], through id_map when given; prev[o] for o in [0, prev_n), else
// key 0) within tiles of 4096 consecutive positions, so that the seeds a
// wave clips together certify after similar numbers of steps while a
// launch still sweeps space tile by tile.  Results never depend on the
// order.  out must not alias list.  Graph capturable.
void launch_order_list(const uint32_t* list, const int64_t* n_dev, int64_t n_cap, const uint32_t* perm,
                       const int32_t* id_map, const int32_t* prev, int64_t prev_n, uint32_t* out, hipStream_t s);
// prev[o] = kcert[s] for the sorted seeds s (as above) whose point o lies in
// [0, prev_n).
void launch_keep_kcert(const int64_t* n_dev, int64_t n_cap, const uint32_t* perm, const int32_t* id_map,
                       const int32_t* kcert, int64_t prev_n, int32_t* prev, hipStream_t s);
// Training label (bucket index) of every seed from its certified k.
void launch_klabels(const int32_t* kcert, int64_t n, int32_t* y, hipStream_t s);
// -------------------------------------------------------------- batch.hip
// Batched boxes, box-major arrays; counts that change per step are device
// scalars (n_dev, nq_dev) and launches depend only on capacities (graph
// capturable).  See batch.hip.
void launch_box_setup(int nbox, int dim, const int64_t* hoff, const int32_t* goff, const int32_t* gtot,
                      int64_t seed_cap, int64_t cell_cap, int64_t* soff, int32_t* N, int64_t* coff, int32_t* err,
                      hipStream_t s);
void launch_bseeds(int dim, const double* pts, int64_t n_h, double buff, int periodic, const int32_t* hbox,
                   const int64_t* hoff, const int32_t* goff, const int64_t* soff, const int32_t* err, double* out,
                   int32_t* sbox, int32_t* orig_h, hipStream_t s);
void launch_bgrid(int dim, const double* pts, const int64_t* n_dev, int64_t cap, int64_t cell_cap,
                  const int32_t* sbox, const int32_t* N, const int64_t* coff, int32_t* cell, int32_t* counts,
                  int32_t* starts, int32_t* fill, int32_t* tmp, int32_t* total, uint32_t* perm, double* sorted,
                  int32_t* sorted_box, int32_t* big, hipStream_t s);
void launch_bknn(int dim, const double* sorted, const uint32_t* perm, const int32_t* sbox, const int32_t* N,
                 const int64_t* coff, const int32_t* starts, const int32_t* counts, const uint32_t* queries,
                 const int64_t* nq_dev, int64_t cap, int k, uint32_t* out, int wave_min_k, int blocks_per_cu,
                 unsigned long long* ticket, hipStream_t s, KnnTail tail = {});
// ticket: a device counter, the work queue of the persistent kernels
// (cleared by the launcher, in stream order)
// loop: persistent blocks over the list (short lists); else one block per
// group of seeds over the capacity (lists that fill most of it)
void launch_bvoronoi(int dim, const double* sorted, const uint32_t* seeds, const int64_t* nq_dev, int64_t cap,
                     const uint32_t* knn, int k, int max_p, int max_t, int big, int lanes, int slots, FaceArrays fa,
                     int blocks_per_cu, unsigned long long* ticket, bool loop, hipStream_t s);
void launch_bcollect(const int64_t* nq_dev, int64_t cap, const uint32_t* seeds, const int32_t* status, int code,
                     uint32_t* list, int32_t* flags, int32_t* offs, int32_t* tmp, int32_t* total, int64_t* nout,
                     hipStream_t s);
void launch_bcount_failed(const int64_t* n_dev, int64_t cap, const int32_t* status, int32_t* nfail, hipStream_t s);
void launch_fill_i32(int32_t* a, int64_t n, int32_t v, hipStream_t s);
// ------------------------------------------------------------- noalloc.hip
// Size of the static device array behind the allocation-free entry points
// (one copy per GPU, held until the process ends).
constexpr int64_t NOALLOC_SCRATCH_BYTES = (int64_t)32 << 20;
// Device address of this GPU's array (the device must be current); looked
// up once per device, then cached.
hipError_t noalloc_setup(int device, void** ptr);
// Scratch bytes the allocation-free KNN build needs for n points on an N^dim grid.
int64_t noalloc_knn_bytes(int dim, int64_t n, int N);
// knn::init + knn::solve on device points: sorted points, permutation and
// rows of length k, scratch carved from `scratch` (noalloc_setup's address),
// no allocation, no memset, no synchronisation.
void noalloc_knn(void* scratch, int dim, const double* pts, int64_t n, int N, int k, double* sorted, uint32_t* perm,
                 uint32_t* rows, hipStream_t s);

// ---------------------------------------------------------------- probe.hip
// Streaming device copy of `bytes` (multiple of 32).
void launch_stream_copy(const void* src, void* dst, int64_t bytes, hipStream_t s);

}  // namespace pg
