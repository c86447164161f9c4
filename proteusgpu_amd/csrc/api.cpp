// C ABI of libpgpu (include/pgpu.h): device memory, the rebuild sequence and
// the restart loop.  Host code only; every kernel lives in the .hip files.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pgpu.h"
#include "common.h"
#include "hydro.h"
#include "kernels.h"

namespace {

thread_local std::string g_err;

struct Error {
    int code;
    std::string msg;
};

void hipchk(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error{PG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e)};
}
#define HIPCHK(x) hipchk((x), #x)

void argchk(bool ok, const char* what) {
    if (!ok) throw Error{PG_ERR_ARG, what};
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return PG_OK;
    } catch (const Error& e) {
        g_err = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return PG_ERR_NOMEM;
    }
}

// Grow-only device buffer.
struct DBuf {
    void* p = nullptr;
    size_t cap = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
    template <class T>
    T* get(size_t count) {
        size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        if (bytes > cap) {
            if (p) HIPCHK(hipFree(p));
            p = nullptr;
            cap = 0;
            hipError_t e = hipMalloc(&p, bytes);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                throw Error{PG_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed"};
            }
            cap = bytes;
        }
        return (T*)p;
    }
    template <class T>
    T* as() const {
        return (T*)p;
    }
};

int ngrid(int dim, int64_t n) { return pg::ngrid_of(dim, n); }

// Row lengths of the first pass and of certificate restarts.  Rows beyond
// 256 serve seeds beside dense clusters (their security radius holds the
// whole cluster); the wave KNN keeps up to KNN_MAX_K entries.
const int KSCHED[] = {16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024};
constexpr int KNN_MAX_K = 1024;
constexpr int MAXP = 64;
// offset pattern of the degeneracy fallback's t-th retry: salt (t + 1) x this
constexpr uint64_t DISPLACE_SALT = 0xD1B54A32D192ED03ull;

}  // namespace

struct pg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    void* scratch = nullptr;  // this GPU's static scratch array (allocation-free entry points)
    DBuf hll_io[5];  // pg_hll_flux staging (left, right, normal, wall, flux), grow-only
};

namespace {

void use(pg_ctx* c) {
    argchk(c != nullptr, "null context");
    HIPCHK(hipSetDevice(c->device));
}

// Uniform grid + counting sort (knn::init).
struct Grid {
    int dim = 3;
    int64_t n = 0;
    int N = 0;
    int64_t ncell = 0;
    DBuf pts, sorted, perm, cell, counts, starts, fill, tmp, total, big;
    // sub-grids of the crowded cells (pg::SubGrid), built with the grid
    DBuf sub_m, sub_ptr, sub_start, sub_cnt, sub_idx, sub_alloc;
    bool has_sub = false;
    int32_t crowded = -1;  // cells over 32 points, read back by note_crowded (-1: not read)

    void build(hipStream_t s) {
        ncell = (int64_t)N * N * (dim == 3 ? N : 1);
        int32_t* cnt = counts.get<int32_t>(ncell);
        int32_t* st = starts.get<int32_t>(ncell);
        int32_t* fl = fill.get<int32_t>(ncell);
        int32_t* tp = tmp.get<int32_t>(pg::scan_tmp_elems(std::max(ncell, n)));
        HIPCHK(hipMemsetAsync(cnt, 0, sizeof(int32_t) * ncell, s));
        HIPCHK(hipMemsetAsync(fl, 0, sizeof(int32_t) * ncell, s));
        pg::launch_cell_count(dim, pts.as<double>(), n, N, cell.get<int32_t>(n), cnt, s);
        pg::scan_exclusive_i32(cnt, st, ncell, tp, total.get<int32_t>(1), s);
        uint32_t* pm = perm.get<uint32_t>(n);
        pg::launch_cell_scatter(n, cell.as<int32_t>(), st, fl, pm, s);
        int32_t* bg = big.get<int32_t>(ncell + 1);
        pg::launch_cell_sort(ncell, st, cnt, pm, bg + 1, bg, s);
        pg::launch_gather_points(dim, n, pm, pts.as<double>(), sorted.get<double>((size_t)n * dim), s);
        // a sub-cell table holds at most n / 3 + 64 entries (about 3 points
        // per sub-cell, at least 8 (2D: 9) sub-cells per cell of over 32)
        const int64_t cap = n / 2 + 1024;
        pg::launch_subgrid(dim, sorted.as<double>(), n, N, ncell, st, cnt, bg + 1, bg, sub_m.get<int32_t>(ncell),
                           sub_ptr.get<int32_t>(ncell), sub_start.get<int32_t>(cap), sub_cnt.get<int32_t>(cap), cap,
                           sub_idx.get<uint32_t>(n), sub_alloc.get<int32_t>(1), s);
        has_sub = true;
        crowded = -1;
        HIPCHK(hipGetLastError());
    }
    // Queue the copy of the crowded-cell count to crowded (read after the
    // stream's next synchronisation).
    void note_crowded(hipStream_t s) {
        if (has_sub) HIPCHK(hipMemcpyAsync(&crowded, big.as<int32_t>(), sizeof(int32_t), hipMemcpyDeviceToHost, s));
    }
    // The sub-grid for kernels that look at it cell by cell (whatever the
    // crowded count): empty only when it was not built.
    pg::SubGrid sub_any() const {
        pg::SubGrid g;
        if (!has_sub) return g;
        g.m = sub_m.as<int32_t>();
        g.ptr = sub_ptr.as<int32_t>();
        g.start = sub_start.as<int32_t>();
        g.cnt = sub_cnt.as<int32_t>();
        g.idx = sub_idx.as<uint32_t>();
        return g;
    }
    // The sub-grid for the searches: empty without crowded cells (the
    // kernels without sub-grid code run), or when the count was not read.
    pg::SubGrid sub() const {
        pg::SubGrid g;
        if (!has_sub || crowded <= 0) return g;
        g.m = sub_m.as<int32_t>();
        g.ptr = sub_ptr.as<int32_t>();
        g.start = sub_start.as<int32_t>();
        g.cnt = sub_cnt.as<int32_t>();
        g.idx = sub_idx.as<uint32_t>();
        return g;
    }
};

}  // namespace

struct pg_knn {
    Grid g;
    DBuf knn;
    int k = 0;
};

struct pg_mesh {
    pg_mesh_params prm{};
    int64_t n_in = 0, n_seeds = 0, n_hydro = 0;
    Grid g;
    DBuf in_pts, disp_pts, gcnt, goff, gsrc, gtot;
    DBuf knn, rows, list, list2, list3, flags, offs, cnt1;
    DBuf feat, kcls, klist, kcnt;  // learned k: features, bucket per seed, seeds by bucket, bucket sizes
    // faces (padded, face-slot major) and cells, sorted order
    DBuf nbr, area, nrm, cen, count, vol, ccen, status, kcert;
    DBuf cstate, resume, lastnb;  // saved cells of uncertified seeds (restarts resume them)
    bool use_resume = false;
    int lanes = 16, slots = 64;  // fast clip kernel shape of the current build
    bool learned_rows = false;  // first pass with per-seed row lengths (knn rows not kept)
    // cells of a KNN problem (pg_voronoi_compute): outputs stay in the
    // problem's sorted order (iperm = identity), while g.perm keeps the
    // problem's permutation for the KNN tie order
    bool sorted_space = false;
    DBuf iperm;
    // learned builds: previous certified k per input id orders the bucket
    // lists (nullptr: ascending seed id)
    const int32_t* order_prev = nullptr;
    int64_t order_prev_n = 0;
    DBuf ovals;  // seeds in clip order
    uint32_t* out_order() { return sorted_space ? iperm.as<uint32_t>() : g.perm.as<uint32_t>(); }
    // hydro tables
    DBuf hid, area_u, vol_u, o2s, bad;
    // owner-computes-once update: hydro code per sorted seed, twin code per
    // face slot, and scratch for the sorted cell records and the per-slot
    // face fluxes
    DBuf hcode, nbr_cm, twin, solve, frec, fbuf;
    bool want_cm = false;  // the clip keeps nbr_cm (meshes of simulations)
    bool finalized = false;
    pg_mesh_info_t info{};

    pg::FaceArrays faces() {
        pg::FaceArrays fa;
        fa.stride = n_seeds;
        fa.nbr = nbr.as<int32_t>();
        fa.area = area.as<double>();
        fa.nrm = nrm.as<double>();
        fa.cen = cen.as<double>();
        fa.count = count.as<uint8_t>();
        fa.vol = vol.as<double>();
        fa.ccen = ccen.as<double>();
        fa.status = status.as<int32_t>();
        fa.kcert = kcert.as<int32_t>();
        if (use_resume) {
            fa.cstate = cstate.as<uint8_t>();
            fa.resume = resume.as<int32_t>();
            fa.lastnb = lastnb.as<uint32_t>();
        }
        // exact-predicate builds: every pass flags the tests that rounding
        // could decide wrongly (the tie-break keys switch that mode on)
        if (prm.exact) fa.key = g.perm.as<uint32_t>();
        if (want_cm) fa.nbr_cm = nbr_cm.as<int32_t>();
        return fa;
    }
    double scale() const { return prm.boundary == PG_BOUNDARY_PERIODIC ? 1. + (2 * prm.buff) : 1.0; }
};

struct pg_kmodel {
    int device = 0;
    float a0 = 1.0f;  // cost model: clip cost per step over KNN cost per row entry
    int64_t step = 0;  // Adam steps taken
    DBuf P, m, v, part, ploss, loss, grad, X, y, cls, logits;
};

struct pg_state {
    int dim = 3;
    int64_t n = 0;
    DBuf rho[2], vel[2], E[2];
    int cur = 0;
};

namespace {

int32_t read_i32(const int32_t* d, hipStream_t s) {
    int32_t h = 0;
    HIPCHK(hipMemcpyAsync(&h, d, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return h;
}

// Seeds of `seeds` (nullptr: all) whose status equals code, into out.
int64_t collect_list(pg_mesh* m, const uint32_t* seeds, int64_t nq, int code, uint32_t* out,
                     hipStream_t s) {
    int32_t* fl = m->flags.get<int32_t>(m->n_seeds);
    int32_t* of = m->offs.get<int32_t>(m->n_seeds);
    int32_t* tp = m->g.tmp.get<int32_t>(pg::scan_tmp_elems(std::max(m->n_seeds, m->g.ncell)));
    int32_t* c = m->cnt1.get<int32_t>(1);
    pg::launch_collect_status(nq, seeds, m->status.as<int32_t>(), code, out, fl, of, tp, c, s);
    HIPCHK(hipGetLastError());
    return read_i32(c, s);
}

// The shape tag under which the fast clip kernel saves and resumes cells.
int32_t clip_shape_tag(int lanes, int slots) { return (lanes << 24) | ((slots / lanes) << 16); }

void run_list(pg_mesh* m, const uint32_t* list, int64_t nq, int k, int big, hipStream_t s, bool exact = false) {
    uint32_t* rows = m->rows.get<uint32_t>((size_t)nq * k);
    pg::KnnTail tail;  // seeds that resume need only the rest of their row
    if (m->use_resume && !big && !exact)
        tail = {m->resume.as<int32_t>(), m->lastnb.as<uint32_t>(), clip_shape_tag(m->lanes, m->slots)};
    pg::launch_knn(m->prm.dim, m->g.sorted.as<double>(), m->g.perm.as<uint32_t>(), m->n_seeds, m->g.N,
                   m->g.starts.as<int32_t>(), m->g.counts.as<int32_t>(), list, nq, k, rows, s, tail, m->g.sub());
    pg::FaceArrays fa = m->faces();
    if (exact) {
        // the exact kernel neither saves nor resumes cells; ties are broken
        // by the input index
        fa.cstate = nullptr;
        fa.resume = nullptr;
        fa.lastnb = nullptr;
        fa.key = m->g.perm.as<uint32_t>();
    }
    pg::launch_voronoi(m->prm.dim, m->g.sorted.as<double>(), m->n_seeds, list, nq, rows, k,
                       m->prm.max_p, m->prm.max_t, big, fa, m->lanes, m->slots, s, true, exact);
    HIPCHK(hipGetLastError());
    m->info.row_entries += nq * k;
}

// Next row length of a restart: the first bucket at least twice k (a
// restarted seed repeats its clips from the start, so k grows
// geometrically: few rounds, and a longer KNN row costs far less than a
// repeated failed clip), else the cap.  -1 when k is at the cap.  The
// result does not depend on the schedule: a longer row extends a shorter
// one and certification happens at the same neighbour.
int next_k(int k, int cap) {
    for (int v : KSCHED) {
        if (v <= k || v > cap) continue;
        if (v >= 2 * k || v == cap || v == KSCHED[sizeof(KSCHED) / sizeof(int) - 1]) return v;
    }
    return k < cap ? cap : -1;
}

// Row length of the 128-slot kernel's pass over overflowed seeds at chain
// step k: k without restarts, else at least BIG_ROW (capped).  A longer row
// extends a shorter one and the cell certifies at the same neighbour, so
// the pass equals one at k, while a seed that would need the next steps of
// the chain (the fast kernel overflows again at the same clip, then the
// 128-slot kernel runs again) is done at once: the few overflowed seeds run
// as one or two waves, latency rather than throughput.  Rows up to 256 stay
// cheap in the wave KNN kernel.
constexpr int BIG_ROW = 256;
int big_row(int k, int cap, bool restart) { return restart ? std::max(k, std::min(cap, BIG_ROW)) : k; }

// Seeds cur[0..nq) (nullptr: all) were just clipped with rows of length k;
// uncertified seeds go on at the next k, until all pass or k reaches its
// cap.  Seeds that outgrow the fast kernel's vertex slots keep that status
// through the chain, and the 128-slot kernel clips all of them once, at
// big_row(k), when the chain's k would pass that row (or the chain ends);
// after every longer step it runs again, and the next step's list is then
// taken over all of cur, so that its uncertified seeds continue too.
// Returns the largest k used.
int restart_chain(pg_mesh* m, const uint32_t* cur, int64_t nq, int k, hipStream_t s) {
    const pg_mesh_params& p = m->prm;
    const size_t ns = (size_t)m->n_seeds;
    uint32_t* buf[2] = {m->list.get<uint32_t>(ns), m->list2.get<uint32_t>(ns)};
    uint32_t* ov = m->list3.get<uint32_t>(ns);
    const int cap = (int)std::min<int64_t>(std::min(std::max(p.k_max, k), KNN_MAX_K), m->n_seeds - 1);
    const bool restart = p.restart != 0;
    const int kb = big_row(k, cap, restart);
    const uint32_t* L = cur;
    int64_t nL = nq;
    int kmax = k;
    bool big_done = false;
    auto big_pass = [&](int kk) -> int64_t {
        const int64_t no = collect_list(m, cur, nq, pg::SLOT_OVERFLOW, ov, s);
        if (no > 0) {
            run_list(m, ov, no, kk, 1, s);
            m->info.n_restarted += no;
            kmax = std::max(kmax, kk);
            L = cur;
            nL = nq;
        }
        return no;
    };
    while (restart && nL > 0) {
        const int kn = next_k(k, cap);
        if (kn < 0) break;
        if (!big_done && kn > kb) {
            big_pass(kb);
            big_done = true;
        }
        uint32_t* dst = (L == buf[0]) ? buf[1] : buf[0];
        const int64_t nr = collect_list(m, L, nL, PG_SECURITY_RADIUS_NOT_REACHED, dst, s);
        if (nr == 0) break;
        k = kn;
        run_list(m, dst, nr, k, 0, s);
        m->info.n_restarted += nr;
        kmax = std::max(kmax, k);
        L = dst;
        nL = nr;
        if (big_done) big_pass(k);
    }
    if (!big_done) {
        const int64_t no = big_pass(kb);
        // a 128-slot pass below the cap may leave seeds that longer rows
        // certify: they continue in a chain from kb (the only uncertified
        // seeds of cur at this point)
        if (restart && no > 0 && kb < cap &&
            collect_list(m, ov, no, PG_SECURITY_RADIUS_NOT_REACHED, buf[0], s) > 0)
            kmax = std::max(kmax, restart_chain(m, cur, nq, kb, s));
    }
    return kmax;
}

// Vertex-ball completion: seeds still uncertified after the restart chain
// (its rows stop at the cap) are clipped by their first min(cap, 64)
// neighbours and then by the points their vertex balls hold, until every
// ball is empty (kcert = cap + 1).
void ball_pass(pg_mesh* m, hipStream_t s) {
    const pg_mesh_params& p = m->prm;
    const size_t ns = (size_t)m->n_seeds;
    uint32_t* L = m->list.get<uint32_t>(ns);
    const int64_t nL = collect_list(m, nullptr, m->n_seeds, PG_SECURITY_RADIUS_NOT_REACHED, L, s);
    m->info.n_ball = nL;
    if (nL == 0) return;
    const int cap = (int)std::min<int64_t>(std::min(std::max(p.k_max, p.k), KNN_MAX_K), m->n_seeds - 1);
    const int kb = std::min(cap, 64);
    uint32_t* rows = m->rows.get<uint32_t>((size_t)nL * kb);
    pg::launch_knn(p.dim, m->g.sorted.as<double>(), m->g.perm.as<uint32_t>(), m->n_seeds, m->g.N,
                   m->g.starts.as<int32_t>(), m->g.counts.as<int32_t>(), L, nL, kb, rows, s, {}, m->g.sub());
    pg::launch_voronoi_ball(p.dim, m->g.sorted.as<double>(), m->n_seeds, L, nL, rows, kb, p.max_p, p.max_t,
                            m->faces(), m->g.N, m->g.starts.as<int32_t>(), m->g.counts.as<int32_t>(),
                            m->g.perm.as<uint32_t>(), cap + 1, s, m->g.sub());
    HIPCHK(hipGetLastError());
    m->info.row_entries += nL * kb;
}

// Exact-predicate fallback (pg_mesh_params.exact): seeds left with
// needs_exact_predicates are clipped again by the exact kernel from the
// first neighbour at the configured k, then at growing k while uncertified
// (the schedule of restart_chain).  Returns the largest k used.
int exact_chain(pg_mesh* m, hipStream_t s) {
    const pg_mesh_params& p = m->prm;
    const size_t ns = (size_t)m->n_seeds;
    uint32_t* buf[2] = {m->list.get<uint32_t>(ns), m->list2.get<uint32_t>(ns)};
    int64_t nL = collect_list(m, nullptr, m->n_seeds, PG_NEEDS_EXACT_PREDICATES, buf[0], s);
    m->info.n_exact = nL;
    if (nL == 0) return 0;
    const int cap = (int)std::min<int64_t>(std::min(std::max(p.k_max, p.k), KNN_MAX_K), m->n_seeds - 1);
    int k = (int)std::min<int64_t>(p.k, m->n_seeds - 1);
    uint32_t* L = buf[0];
    run_list(m, L, nL, k, 1, s, true);
    int kmax = k;
    while (p.restart) {
        const int kn = next_k(k, cap);
        if (kn < 0) break;
        uint32_t* dst = (L == buf[0]) ? buf[1] : buf[0];
        const int64_t nr = collect_list(m, L, nL, PG_SECURITY_RADIUS_NOT_REACHED, dst, s);
        if (nr == 0) break;
        k = kn;
        run_list(m, dst, nr, k, 1, s, true);
        m->info.n_restarted += nr;
        kmax = k;
        L = dst;
        nL = nr;
    }
    return kmax;
}

void count_faces(pg_mesh* m, hipStream_t s);

// Stage markers of one rebuild (HIP events on the context stream).
struct Marks {
    hipEvent_t ev[12] = {};
    bool on = false;
    void mark(int i, hipStream_t s) {
        if (on) HIPCHK(hipEventRecord(ev[i], s));
    }
};
enum { MK_START, MK_GHOSTS, MK_GRID, MK_PREDICT, MK_KNN, MK_CLIP, MK_RESTART, MK_FINAL, MK_HYDRO, MK_END, MK_N };

// Learned builds order each bucket's seeds by their previous certified k
// (PGPU_LK_ORDER=0 keeps ascending seed ids).
bool lk_order_on() {
    static const bool on = [] {
        const char* e = std::getenv("PGPU_LK_ORDER");
        return !(e && std::atoi(e) == 0);
    }();
    return on;
}

// Face and cell outputs of ns seeds.
void alloc_cells(pg_mesh* m, size_t ns) {
    m->nbr.get<int32_t>(ns * MAXP);
    m->area.get<double>(ns * MAXP);
    m->nrm.get<double>(3 * ns * MAXP);
    m->cen.get<double>(3 * ns * MAXP);
    m->count.get<uint8_t>(ns);
    m->vol.get<double>(ns);
    m->ccen.get<double>(3 * ns);
    m->status.get<int32_t>(ns);
    m->kcert.get<int32_t>(ns);
    if (m->want_cm) m->nbr_cm.get<int32_t>(ns * MAXP);
}

// One build of the mesh of pts (device memory when on_device).  count:
// also learn VMesh.num_faces (builds whose mesh is finalized learn it there).
void mesh_build_once(pg_ctx* c, pg_mesh* m, const double* pts, int64_t n, bool on_device, bool count, Marks* mk,
                     const pg_kmodel* km) {
    hipStream_t s = c->stream;
    const pg_mesh_params& p = m->prm;
    const int dim = p.dim;
    m->n_in = n;
    double* din;
    if (on_device) {
        din = const_cast<double*>(pts);
    } else {
        din = m->in_pts.get<double>((size_t)n * dim);
        HIPCHK(hipMemcpyAsync(din, pts, sizeof(double) * n * dim, hipMemcpyHostToDevice, s));
    }
    int64_t nseeds = n;
    if (p.boundary == PG_BOUNDARY_PERIODIC) {
        int32_t* gc = m->gcnt.get<int32_t>(n);
        int32_t* go = m->goff.get<int32_t>(n);
        int32_t* tp = m->g.tmp.get<int32_t>(pg::scan_tmp_elems(n));
        int32_t* gt = m->gtot.get<int32_t>(1);
        pg::launch_ghost_count(dim, din, n, p.buff, gc, s);
        pg::scan_exclusive_i32(gc, go, n, tp, gt, s);
        int64_t ng = read_i32(gt, s);
        nseeds = n + ng;
        double* gp = m->g.pts.get<double>((size_t)nseeds * dim);
        pg::launch_ghost_write(dim, din, n, p.buff, go, gp, m->gsrc.get<int64_t>(std::max<int64_t>(ng, 1)), s);
        m->n_hydro = n;
    } else {
        double* gp = m->g.pts.get<double>((size_t)n * dim);
        HIPCHK(hipMemcpyAsync(gp, din, sizeof(double) * n * dim, hipMemcpyDeviceToDevice, s));
        m->n_hydro = p.boundary == PG_BOUNDARY_REFLECTIVE ? n : 0;
    }
    HIPCHK(hipGetLastError());
    m->n_seeds = nseeds;
    argchk(nseeds - 1 >= p.k, "k must be below the number of seeds");
    m->g.dim = dim;
    m->g.n = nseeds;
    m->g.N = ngrid(dim, nseeds);
    if (p.boundary != PG_BOUNDARY_REFLECTIVE && m->g.N < 16)
        throw Error{PG_ERR_GRID, "KNN: N_grid < 16 (the reference's minimum mesh size)"};
    if (mk) mk->mark(MK_GHOSTS, s);
    m->g.build(s);
    if (mk) mk->mark(MK_GRID, s);

    const size_t ns = (size_t)nseeds;
    alloc_cells(m, ns);
    // the clip shape and resume switch are read once per build
    m->use_resume = p.restart && pg::clip_resume();
    m->lanes = pg::clip_lanes();
    m->slots = pg::clip_slots();
    if (m->use_resume) {
        m->cstate.get<uint8_t>(ns * pg::CLIP_STATE_BYTES);
        m->lastnb.get<uint32_t>(ns);
        // saved cells belong to this build only: no pass may resume a cell
        // of an earlier build (or of recycled memory)
        pg::launch_fill_i32(m->resume.get<int32_t>(ns), (int64_t)ns, 0, s);
    }
    m->finalized = false;

    m->info = pg_mesh_info_t{};
    m->learned_rows = km != nullptr;
    int k_used = p.k;
    if (!km) {
        // first pass: every seed at the configured k
        const int k = p.k;
        m->g.note_crowded(s);  // the searches take the sub-grid only if some cell is crowded
        HIPCHK(hipStreamSynchronize(s));
        if (mk) mk->mark(MK_PREDICT, s);
        uint32_t* rows = m->knn.get<uint32_t>(ns * k);
        pg::launch_knn(dim, m->g.sorted.as<double>(), m->g.perm.as<uint32_t>(), nseeds, m->g.N,
                       m->g.starts.as<int32_t>(), m->g.counts.as<int32_t>(), nullptr, nseeds, k, rows, s, {},
                       m->g.sub());
        if (mk) mk->mark(MK_KNN, s);
        pg::launch_voronoi(dim, m->g.sorted.as<double>(), nseeds, nullptr, nseeds, rows, k, p.max_p, p.max_t,
                           0, m->faces(), m->lanes, m->slots, s);
        HIPCHK(hipGetLastError());
        m->info.row_entries += nseeds * k;
        if (mk) mk->mark(MK_CLIP, s);
        k_used = restart_chain(m, nullptr, nseeds, k, s);
    } else {
        // learned first pass: features -> bucket per seed -> seeds grouped by
        // bucket (ascending id inside a bucket) -> KNN + clip per bucket
        float* feat = m->feat.get<float>(ns * pg::LK_NFEAT);
        int32_t* cls = m->kcls.get<int32_t>(ns);
        pg::launch_kfeatures(dim, m->g.sorted.as<double>(), nseeds, m->g.N, m->g.counts.as<int32_t>(),
                             m->g.starts.as<int32_t>(), nullptr, feat, s, m->g.sub_any());
        pg::launch_kpredict(feat, nseeds, km->P.as<float>(), km->a0, nullptr, cls, nullptr, s);
        HIPCHK(hipGetLastError());
        int64_t hc[pg::LK_NCLASS];
        {
            const int64_t nbk = pg::lk_bucket_blocks(nseeds) * pg::LK_NCLASS;
            int32_t* tp = m->g.tmp.get<int32_t>(pg::scan_tmp_elems(std::max({nseeds, m->g.ncell, nbk})));
            int64_t* dcnt = m->kcnt.get<int64_t>(pg::LK_NCLASS);
            // seeds in the order of their previous certified k, so that the
            // seeds a wave clips together certify after similar numbers of
            // steps (each bucket keeps this order)
            const uint32_t* order = nullptr;
            if (m->order_prev && lk_order_on()) {
                uint32_t* ord = m->ovals.get<uint32_t>(ns);
                pg::launch_order_list(nullptr, nullptr, nseeds, m->g.perm.as<uint32_t>(), nullptr, m->order_prev,
                                      m->order_prev_n, ord, s);
                order = ord;
            }
            pg::launch_kbucket(cls, nullptr, nseeds, m->flags.get<int32_t>(std::max<size_t>(ns, nbk)),
                               m->offs.get<int32_t>(std::max<size_t>(ns, nbk)), tp, m->cnt1.get<int32_t>(1), ns,
                               m->klist.get<uint32_t>(ns * pg::LK_NCLASS), dcnt, s, order);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(hc, dcnt, sizeof(hc), hipMemcpyDeviceToHost, s));
            m->g.note_crowded(s);
            HIPCHK(hipStreamSynchronize(s));
        }
        uint32_t* kl = m->klist.as<uint32_t>();
        int bk[pg::LK_NCLASS];
        int64_t roff[pg::LK_NCLASS + 1] = {0};
        for (int c = 0; c < pg::LK_NCLASS; c++) {
            m->info.n_class[c] = hc[c];
            bk[c] = (int)std::min<int64_t>(pg::lk_bucket(c), nseeds - 1);
            roff[c + 1] = roff[c] + hc[c] * bk[c];
        }
        if (mk) mk->mark(MK_PREDICT, s);
        uint32_t* rows = m->knn.get<uint32_t>((size_t)std::max<int64_t>(roff[pg::LK_NCLASS], 1));
        for (int c = 0; c < pg::LK_NCLASS; c++)
            if (hc[c] > 0)
                pg::launch_knn(dim, m->g.sorted.as<double>(), m->g.perm.as<uint32_t>(), nseeds, m->g.N,
                               m->g.starts.as<int32_t>(), m->g.counts.as<int32_t>(), kl + (size_t)c * ns, hc[c],
                               bk[c], rows + roff[c], s, {}, m->g.sub());
        if (mk) mk->mark(MK_KNN, s);
        for (int c = 0; c < pg::LK_NCLASS; c++)
            if (hc[c] > 0) {
                pg::launch_voronoi(dim, m->g.sorted.as<double>(), nseeds, kl + (size_t)c * ns, hc[c], rows + roff[c],
                                   bk[c], p.max_p, p.max_t, 0, m->faces(), m->lanes, m->slots, s);
                m->info.row_entries += hc[c] * bk[c];
            }
        HIPCHK(hipGetLastError());
        if (mk) mk->mark(MK_CLIP, s);
        // every rejected seed (any bucket) continues in one restart chain:
        // a seed recomputed at a row length at least its certified k gives
        // the same cell whatever its first bucket, so the chain may start at
        // 64 (overflowed seeds at 64, uncertified ones at 128, then 256)
        k_used = std::max(restart_chain(m, nullptr, nseeds, std::min<int64_t>(64, nseeds - 1), s),
                          *std::max_element(bk, bk + pg::LK_NCLASS));
    }
    if (p.restart) ball_pass(m, s);
    if (p.exact) k_used = std::max(k_used, exact_chain(m, s));
    if (mk) mk->mark(MK_RESTART, s);
    m->info.n_seeds = nseeds;
    m->info.n_hydro = m->n_hydro;
    m->info.N_grid = m->g.N;
    m->info.k_used = k_used;
    m->info.n_failed = collect_list(m, nullptr, nseeds, -1, m->list.get<uint32_t>(ns), s);
    if (count) count_faces(m, s);
}

// Cells of the last build whose predicates were degenerate.
int64_t degenerate_cells(pg_mesh* m, hipStream_t s) {
    uint32_t* tmp = m->list.get<uint32_t>((size_t)m->n_seeds);
    return collect_list(m, nullptr, m->n_seeds, PG_NEEDS_EXACT_PREDICATES, tmp, s) +
           collect_list(m, nullptr, m->n_seeds, PG_INCONSISTENT_BOUNDARY, tmp, s);
}

// A mesh build with the degeneracy fallback of pg_mesh_params.perturb:
// builds that leave degenerate cells are repeated on displaced generators
// (amplitude perturb x n^(-1/dim), times ten per retry, three builds at most).
void mesh_build(pg_ctx* c, pg_mesh* m, const double* pts, int64_t n, bool on_device = false,
                Marks* mk = nullptr, const pg_kmodel* km = nullptr) {
    mesh_build_once(c, m, pts, n, on_device, !on_device, mk, km);
    if (m->prm.perturb <= 0 || m->info.n_failed == 0) return;
    hipStream_t s = c->stream;
    const int64_t ndeg = degenerate_cells(m, s);
    if (ndeg == 0) return;
    const int dim = m->prm.dim;
    // the generators as given, on the device (a host input was copied there)
    const double* src = on_device ? pts : m->in_pts.as<double>();
    double* dsp = m->disp_pts.get<double>((size_t)n * dim);
    double amp = m->prm.perturb * std::pow((double)n, -1.0 / dim);
    for (int t = 0; t < 2; t++, amp *= 10.0) {
        pg::launch_displace(dim, src, n, amp, DISPLACE_SALT * (uint64_t)(t + 1),
                            m->prm.boundary == PG_BOUNDARY_PERIODIC ? 1 : 0, dsp, s);
        HIPCHK(hipGetLastError());
        mesh_build_once(c, m, dsp, n, true, !on_device, mk, km);
        m->info.displacement = amp;
        if (m->info.n_failed == 0 || degenerate_cells(m, s) == 0) break;
    }
    m->info.n_degenerate = ndeg;
}

// Total faces over all cells (VMesh.num_faces).
void count_faces(pg_mesh* m, hipStream_t s) {
    const size_t ns = (size_t)m->n_seeds;
    const int64_t nseeds = m->n_seeds;
    int32_t* co = m->flags.get<int32_t>(ns);
    int32_t* po = m->offs.get<int32_t>(ns);
    int32_t* tp = m->g.tmp.get<int32_t>(pg::scan_tmp_elems(std::max<int64_t>(nseeds, m->g.ncell)));
    pg::launch_counts_orig(nseeds, m->out_order(), m->count.as<uint8_t>(), co, s);
    pg::scan_exclusive_i32(co, po, nseeds, tp, m->cnt1.get<int32_t>(1), s);
    m->info.num_faces = read_i32(m->cnt1.as<int32_t>(), s);
}

void mesh_finalize(pg_ctx* c, pg_mesh* m) {
    if (m->finalized) return;
    hipStream_t s = c->stream;
    const size_t ns = (size_t)m->n_seeds;
    const double sc = m->scale();
    const double as = m->prm.dim == 2 ? sc : sc * sc;
    const double vs = m->prm.dim == 2 ? sc * sc : sc * sc * sc;
    int32_t* bad = m->bad.get<int32_t>(2);
    HIPCHK(hipMemsetAsync(bad, 0, 2 * sizeof(int32_t), s));
    const int64_t* gsrc = m->prm.boundary == PG_BOUNDARY_PERIODIC ? m->gsrc.as<int64_t>() : nullptr;
    pg::launch_finalize(m->n_seeds, m->n_hydro, m->n_seeds, m->out_order(), gsrc,
                        m->count.as<uint8_t>(), m->nbr.as<int32_t>(), m->area.as<double>(),
                        m->vol.as<double>(), as, vs, m->hid.get<int32_t>(ns * MAXP),
                        m->area_u.get<double>(ns * MAXP), m->vol_u.get<double>(ns),
                        m->o2s.get<int32_t>(std::max<int64_t>(m->n_hydro, 1)), m->hcode.get<int32_t>(ns), bad, s);
    HIPCHK(hipGetLastError());
    // faces over all cells (VMesh.num_faces), read with the error counters:
    // meshes built on the device (simulations) learn their face count here
    int32_t* co = m->flags.get<int32_t>(ns);
    int32_t* po = m->offs.get<int32_t>(ns);
    int32_t* tp = m->g.tmp.get<int32_t>(pg::scan_tmp_elems(std::max<int64_t>(m->n_seeds, m->g.ncell)));
    pg::launch_counts_orig(m->n_seeds, m->out_order(), m->count.as<uint8_t>(), co, s);
    pg::scan_exclusive_i32(co, po, m->n_seeds, tp, m->cnt1.get<int32_t>(1), s);
    HIPCHK(hipGetLastError());
    int32_t hb[2] = {0, 0}, nf = 0;
    HIPCHK(hipMemcpyAsync(hb, bad, sizeof(hb), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&nf, m->cnt1.as<int32_t>(), sizeof(nf), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    m->info.num_faces = nf;
    if (hb[0] != 0) throw Error{PG_ERR_CELLS, "hydro cells without faces (uncertified cells)"};
    if (hb[1] != 0)
        throw Error{PG_ERR_CELLS, std::to_string(hb[1]) +
                                      " hydro faces reach the box plane: the periodic ghost layer (buff) is "
                                      "thinner than the cells; increase buff"};
    m->finalized = true;
}

}  // namespace

// =================================================================== C ABI
extern "C" {

const char* pg_version(void) { return "pgpu 0.1 (gfx950)"; }
const char* pg_last_error(void) { return g_err.c_str(); }

int pg_set_clip_config(int lanes, int slots) {
    return guarded([&] {
        argchk(pg::valid_clip_config(lanes, slots), "clip lanes must be 8 or 16, slots 48 or 64");
        pg::set_clip_config(lanes, slots);
    });
}

int pg_get_clip_config(int* lanes, int* slots) {
    return guarded([&] {
        argchk(lanes && slots, "null argument");
        *lanes = pg::clip_lanes();
        *slots = pg::clip_slots();
    });
}

int pg_set_clip_resume(int on) {
    return guarded([&] { pg::set_clip_resume(on != 0); });
}

int pg_get_clip_resume(int* on) {
    return guarded([&] {
        argchk(on, "null argument");
        *on = pg::clip_resume() ? 1 : 0;
    });
}

int pg_device_count(int* count) {
    return guarded([&] {
        argchk(count != nullptr, "null count");
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            n = 0;
        }
        *count = n;
    });
}

int pg_ctx_create(int device, pg_ctx** out) {
    return guarded([&] {
        argchk(out != nullptr, "null out");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
            (void)hipGetLastError();
            throw Error{PG_ERR_NODEV, "no GPU device visible"};
        }
        argchk(device >= 0 && device < n, "device index out of range");
        HIPCHK(hipSetDevice(device));
        pg_ctx* c = new pg_ctx();
        c->device = device;
        hipError_t e = pg::noalloc_setup(device, &c->scratch);
        if (e != hipSuccess) {
            delete c;
            hipchk(e, "hipGetSymbolAddress of the scratch array");
        }
        e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            hipchk(e, "hipStreamCreate");
        }
        *out = c;
    });
}

int pg_ctx_destroy(pg_ctx* c) {
    return guarded([&] {
        if (!c) return;
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamDestroy(c->stream);
        delete c;
    });
}

int pg_ctx_sync(pg_ctx* c) {
    return guarded([&] {
        use(c);
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int pg_stream_copy_probe(pg_ctx* c, int64_t bytes, int reps, double* ms) {
    return guarded([&] {
        use(c);
        argchk(ms != nullptr && bytes >= 32 && reps >= 1, "bad argument");
        bytes &= ~(int64_t)31;
        DBuf a, b;
        void* pa = a.get<char>((size_t)bytes);
        void* pb = b.get<char>((size_t)bytes);
        HIPCHK(hipMemsetAsync(pa, 1, (size_t)bytes, c->stream));
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        pg::launch_stream_copy(pa, pb, bytes, c->stream);  // warm-up
        HIPCHK(hipEventRecord(e0, c->stream));
        for (int r = 0; r < reps; r++) pg::launch_stream_copy(pa, pb, bytes, c->stream);
        HIPCHK(hipEventRecord(e1, c->stream));
        HIPCHK(hipEventSynchronize(e1));
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms = (double)t / reps;
    });
}

// ------------------------------------------------------------------ KNN
int pg_knn_init(pg_ctx* c, int dim, const double* pts, int64_t n, pg_knn** out) {
    return guarded([&] {
        use(c);
        argchk(out && pts, "null argument");
        argchk(dim == 2 || dim == 3, "dim must be 2 or 3");
        argchk(n >= 2 && n < (int64_t)1 << 31, "n out of range");
        int N = ngrid(dim, n);
        if (N < 16) throw Error{PG_ERR_GRID, "KNN: N_grid < 16 (the reference's minimum mesh size)"};
        pg_knn* k = new pg_knn();
        try {
            k->g.dim = dim;
            k->g.n = n;
            k->g.N = N;
            double* d = k->g.pts.get<double>((size_t)n * dim);
            HIPCHK(hipMemcpyAsync(d, pts, sizeof(double) * n * dim, hipMemcpyHostToDevice, c->stream));
            k->g.build(c->stream);
            k->g.note_crowded(c->stream);
            HIPCHK(hipStreamSynchronize(c->stream));
        } catch (...) {
            delete k;
            throw;
        }
        *out = k;
    });
}

int pg_knn_solve(pg_ctx* c, pg_knn* k, int kk) {
    return guarded([&] {
        use(c);
        argchk(k != nullptr, "null knn");
        argchk(kk >= 1 && kk <= 256, "k must be in [1, 256]");
        uint32_t* o = k->knn.get<uint32_t>((size_t)k->g.n * kk);
        pg::launch_knn(k->g.dim, k->g.sorted.as<double>(), k->g.perm.as<uint32_t>(), k->g.n, k->g.N,
                       k->g.starts.as<int32_t>(), k->g.counts.as<int32_t>(), nullptr, k->g.n, kk, o,
                       c->stream, {}, k->g.sub());
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        k->k = kk;
    });
}

int pg_knn_solve_noalloc(pg_ctx* c, int dim, const double* pts, int64_t n, int kk, double* sorted, uint32_t* perm,
                         uint32_t* rows, int* N_grid) {
    static_assert(PG_NOALLOC_SCRATCH_BYTES == pg::NOALLOC_SCRATCH_BYTES, "scratch size");
    return guarded([&] {
        use(c);
        argchk(pts && sorted && perm && rows, "null argument");
        argchk(dim == 2 || dim == 3, "dim must be 2 or 3");
        argchk(n >= 2 && n < (int64_t)1 << 31, "n out of range");
        argchk(kk >= 1 && kk <= 256, "k must be in [1, 256]");
        const int N = ngrid(dim, n);
        if (N < 16) throw Error{PG_ERR_GRID, "KNN: N_grid < 16 (the reference's minimum mesh size)"};
        const int64_t need = pg::noalloc_knn_bytes(dim, n, N);
        if (need > pg::NOALLOC_SCRATCH_BYTES)
            throw Error{PG_ERR_ARG, "pg_knn_solve_noalloc: " + std::to_string(n) + " points need " +
                                        std::to_string(need) + " bytes of scratch, more than the " +
                                        std::to_string(pg::NOALLOC_SCRATCH_BYTES) + " reserved per GPU"};
        argchk(c->scratch != nullptr, "context without scratch array");
        pg::noalloc_knn(c->scratch, dim, pts, n, N, kk, sorted, perm, rows, c->stream);
        HIPCHK(hipGetLastError());
        if (N_grid) *N_grid = N;
    });
}

int pg_knn_info(const pg_knn* k, int64_t* n, int* N, int* kk) {
    return guarded([&] {
        argchk(k != nullptr, "null knn");
        if (n) *n = k->g.n;
        if (N) *N = k->g.N;
        if (kk) *kk = k->k;
    });
}

int pg_knn_get_points(pg_ctx* c, const pg_knn* k, double* out) {
    return guarded([&] {
        use(c);
        argchk(k && out, "null argument");
        HIPCHK(hipMemcpyAsync(out, k->g.sorted.p, sizeof(double) * k->g.n * k->g.dim, hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int pg_knn_get_knearest(pg_ctx* c, const pg_knn* k, uint32_t* out) {
    return guarded([&] {
        use(c);
        argchk(k && out, "null argument");
        argchk(k->k > 0, "knn not solved");
        HIPCHK(hipMemcpyAsync(out, k->knn.p, sizeof(uint32_t) * k->g.n * k->k, hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int pg_knn_get_permutation(pg_ctx* c, const pg_knn* k, uint32_t* out) {
    return guarded([&] {
        use(c);
        argchk(k && out, "null argument");
        HIPCHK(hipMemcpyAsync(out, k->g.perm.p, sizeof(uint32_t) * k->g.n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int pg_knn_free(pg_knn** k) {
    return guarded([&] {
        if (!k || !*k) return;
        delete *k;
        *k = nullptr;
    });
}

// -------------------------------------------------------------- Voronoi
void pg_mesh_params_default(pg_mesh_params* p) {
    if (!p) return;
    p->dim = 3;
    p->boundary = PG_BOUNDARY_PERIODIC;
    p->buff = 0.04;
    p->k = 32;
    p->max_p = 64;
    p->max_t = 96;
    p->restart = 1;
    p->k_max = KNN_MAX_K;
    p->perturb = 0.0;
    p->exact = 0;
}

int pg_mesh_compute_k(pg_ctx* c, const pg_mesh_params* prm, const pg_kmodel* km, const double* pts, int64_t n,
                      pg_mesh** out) {
    return guarded([&] {
        use(c);
        argchk(!km || km->device == c->device, "model lives on another device");
        argchk(prm && pts && out, "null argument");
        argchk(prm->dim == 2 || prm->dim == 3, "dim must be 2 or 3");
        argchk(prm->boundary >= 0 && prm->boundary <= 2, "unknown boundary");
        argchk(prm->k >= 1 && prm->k <= 256, "k must be in [1, 256]");
        argchk(prm->k_max <= KNN_MAX_K, "k_max must be at most 1024");
        argchk(prm->max_p >= 2 * prm->dim + 1 && prm->max_p <= MAXP, "max_p must be in [2*dim+1, 64]");
        argchk(prm->max_t >= 8 && prm->max_t <= 128, "max_t must be in [8, 128]");
        argchk(prm->boundary != PG_BOUNDARY_PERIODIC || (prm->buff > 0 && prm->buff < 0.5),
               "buff must be in (0, 0.5)");
        argchk(prm->perturb >= 0.0 && prm->perturb <= 1e-3, "perturb must be in [0, 1e-3]");
        argchk(prm->exact == 0 || prm->exact == 1, "exact must be 0 or 1");
        argchk(n >= 2 && n < (int64_t)1 << 29, "n out of range");
        pg_mesh* m = new pg_mesh();
        try {
            m->prm = *prm;
            if (m->prm.k_max < m->prm.k) m->prm.k_max = m->prm.k;
            mesh_build(c, m, pts, n, false, nullptr, km);
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}

int pg_mesh_compute(pg_ctx* c, const pg_mesh_params* prm, const double* pts, int64_t n, pg_mesh** out) {
    return pg_mesh_compute_k(c, prm, nullptr, pts, n, out);
}

int pg_voronoi_compute(pg_ctx* c, const pg_knn* kc, int max_p, int max_t, pg_mesh** out, int32_t* status) {
    return guarded([&] {
        use(c);
        argchk(kc && out, "null argument");
        pg_knn* kp = const_cast<pg_knn*>(kc);
        argchk(kp->k >= 1, "the KNN problem is not solved (pg_knn_solve)");
        // a row longer than the other points would hold missing entries
        // (UINT_MAX); unreachable while pg_knn_init needs N_grid >= 16
        argchk(kp->g.n - 1 >= kp->k, "k must be below the number of points");
        const int dim = kp->g.dim;
        argchk(max_p >= 2 * dim + 1 && max_p <= MAXP, "max_p must be in [2*dim+1, 64]");
        argchk(max_t >= 8 && max_t <= 128, "max_t must be in [8, 128]");
        hipStream_t s = c->stream;
        pg_mesh* m = new pg_mesh();
        try {
            pg_mesh_params_default(&m->prm);
            m->prm.dim = dim;
            m->prm.boundary = PG_BOUNDARY_NONE;
            m->prm.k = m->prm.k_max = kp->k;
            m->prm.max_p = max_p;
            m->prm.max_t = max_t;
            m->prm.restart = 0;  // compute_cells has no restarts (voronoi.cpp:104-110)
            const int64_t n = kp->g.n;
            const size_t nn = (size_t)n;
            m->n_in = m->n_seeds = n;
            m->n_hydro = 0;
            // the KNN problem's grid and rows; the mesh's outputs are indexed
            // by the problem's sorted ids, as compute_cells leaves them
            // before unpermute_vmesh
            Grid& g = m->g;
            g.dim = dim;
            g.n = n;
            g.N = kp->g.N;
            g.ncell = kp->g.ncell;
            HIPCHK(hipMemcpyAsync(g.sorted.get<double>(nn * dim), kp->g.sorted.p, sizeof(double) * nn * dim,
                                  hipMemcpyDeviceToDevice, s));
            HIPCHK(hipMemcpyAsync(g.counts.get<int32_t>(g.ncell), kp->g.counts.p, sizeof(int32_t) * g.ncell,
                                  hipMemcpyDeviceToDevice, s));
            HIPCHK(hipMemcpyAsync(g.starts.get<int32_t>(g.ncell), kp->g.starts.p, sizeof(int32_t) * g.ncell,
                                  hipMemcpyDeviceToDevice, s));
            HIPCHK(hipMemcpyAsync(g.perm.get<uint32_t>(nn), kp->g.perm.p, sizeof(uint32_t) * nn,
                                  hipMemcpyDeviceToDevice, s));
            std::vector<uint32_t> iota(nn);
            for (size_t i = 0; i < nn; i++) iota[i] = (uint32_t)i;
            HIPCHK(hipMemcpyAsync(m->iperm.get<uint32_t>(nn), iota.data(), sizeof(uint32_t) * nn,
                                  hipMemcpyHostToDevice, s));
            m->sorted_space = true;
            const int k = kp->k;
            uint32_t* rows = m->knn.get<uint32_t>(nn * k);
            HIPCHK(hipMemcpyAsync(rows, kp->knn.p, sizeof(uint32_t) * nn * k, hipMemcpyDeviceToDevice, s));
            alloc_cells(m, nn);
            m->use_resume = false;
            m->lanes = pg::clip_lanes();
            m->slots = pg::clip_slots();
            m->finalized = false;
            m->info = pg_mesh_info_t{};
            pg::launch_voronoi(dim, g.sorted.as<double>(), n, nullptr, n, rows, k, max_p, max_t, 0, m->faces(),
                               m->lanes, m->slots, s);
            HIPCHK(hipGetLastError());
            m->info.row_entries = n * k;
            // seeds that outgrew the fast kernel's slots: the 128-slot kernel
            m->info.k_used = restart_chain(m, nullptr, n, k, s);
            m->info.n_seeds = n;
            m->info.N_grid = g.N;
            m->info.n_failed = collect_list(m, nullptr, n, -1, m->list.get<uint32_t>(nn), s);
            count_faces(m, s);
            if (status) {
                HIPCHK(hipMemcpyAsync(status, m->status.p, sizeof(int32_t) * nn, hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
            }
            // iota must outlive the copy above
            HIPCHK(hipStreamSynchronize(s));
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}

int pg_mesh_unpermute(pg_ctx* c, pg_mesh* m) {
    return guarded([&] {
        use(c);
        argchk(m != nullptr, "null mesh");
        argchk(m->sorted_space, "the mesh is already in input order");
        m->sorted_space = false;
    });
}

int pg_mesh_info(const pg_mesh* m, pg_mesh_info_t* info) {
    return guarded([&] {
        argchk(m && info, "null argument");
        *info = m->info;
    });
}

int pg_mesh_get_knn(pg_ctx* c, const pg_mesh* m, double* points, uint32_t* nearest, uint32_t* permutation) {
    return guarded([&] {
        use(c);
        argchk(m != nullptr, "null mesh");
        argchk(!m->learned_rows, "the mesh was built with learned k: its rows have per-seed lengths");
        hipStream_t s = c->stream;
        const size_t ns = (size_t)m->n_seeds;
        if (points)
            HIPCHK(hipMemcpyAsync(points, m->g.sorted.p, sizeof(double) * ns * m->prm.dim, hipMemcpyDeviceToHost, s));
        if (nearest)
            HIPCHK(hipMemcpyAsync(nearest, m->knn.p, sizeof(uint32_t) * ns * m->prm.k, hipMemcpyDeviceToHost, s));
        if (permutation) HIPCHK(hipMemcpyAsync(permutation, m->g.perm.p, sizeof(uint32_t) * ns, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    });
}

int pg_mesh_copy_out(pg_ctx* c, const pg_mesh* mc, pg_vmesh_view* o) {
    return guarded([&] {
        use(c);
        argchk(mc && o, "null argument");
        pg_mesh* m = const_cast<pg_mesh*>(mc);
        hipStream_t s = c->stream;
        const int64_t n = m->n_seeds, F = m->info.num_faces;
        const int dim = m->prm.dim;
        const double sc = m->scale();
        const double as = dim == 2 ? sc : sc * sc;
        const double vs = dim == 2 ? sc * sc : sc * sc * sc;
        // face pointers in original order
        int32_t* co = m->flags.get<int32_t>(n);
        int32_t* po = m->offs.get<int32_t>(n);
        int32_t* tp = m->g.tmp.get<int32_t>(pg::scan_tmp_elems(std::max<int64_t>(n, m->g.ncell)));
        pg::launch_counts_orig(n, m->out_order(), m->count.as<uint8_t>(), co, s);
        pg::scan_exclusive_i32(co, po, n, tp, m->cnt1.get<int32_t>(1), s);
        // the face arrays below are sized by the caller from pg_mesh_info:
        // refuse a mesh whose face count differs from the one reported
        int32_t nf = 0;
        HIPCHK(hipMemcpyAsync(&nf, m->cnt1.as<int32_t>(), sizeof(nf), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if ((int64_t)nf != F)
            throw Error{PG_ERR_ARG, "mesh has " + std::to_string(nf) + " faces, pg_mesh_info reported " +
                                        std::to_string(F) + "; query pg_mesh_info again"};
        DBuf onbr, onrm, oarea, ocen, oseeds, ovol, ost, okc, occ;
        int32_t* dnbr = onbr.get<int32_t>(F);
        double* dnrm = onrm.get<double>(3 * F);
        double* dar = oarea.get<double>(F);
        double* dcen = ocen.get<double>(3 * F);
        pg::launch_pack_faces(n, n, n * MAXP, dim, m->out_order(), m->count.as<uint8_t>(), po,
                              m->nbr.as<int32_t>(), m->nrm.as<double>(), m->area.as<double>(),
                              m->cen.as<double>(), as, sc, dnbr, dnrm, dar, dcen, s);
        double* dse = oseeds.get<double>(3 * n);
        double* dvo = ovol.get<double>(n);
        int32_t* dst = ost.get<int32_t>(n);
        int32_t* dkc = okc.get<int32_t>(n);
        double* dcc = occ.get<double>(3 * n);
        pg::launch_cells_orig(n, n, dim, m->out_order(), m->g.sorted.as<double>(),
                              m->status.as<int32_t>(), m->kcert.as<int32_t>(), m->vol.as<double>(),
                              m->ccen.as<double>(), sc, vs, dse, dvo, dst, dkc, dcc, s);
        HIPCHK(hipGetLastError());
        std::vector<int32_t> hcnt(n), hptr(n);
        HIPCHK(hipMemcpyAsync(hcnt.data(), co, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(hptr.data(), po, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
        if (o->neighbor_cell) HIPCHK(hipMemcpyAsync(o->neighbor_cell, dnbr, sizeof(int32_t) * F, hipMemcpyDeviceToHost, s));
        if (o->face_normal) HIPCHK(hipMemcpyAsync(o->face_normal, dnrm, sizeof(double) * 3 * F, hipMemcpyDeviceToHost, s));
        if (o->face_area) HIPCHK(hipMemcpyAsync(o->face_area, dar, sizeof(double) * F, hipMemcpyDeviceToHost, s));
        if (o->face_centroid) HIPCHK(hipMemcpyAsync(o->face_centroid, dcen, sizeof(double) * 3 * F, hipMemcpyDeviceToHost, s));
        if (o->seeds) HIPCHK(hipMemcpyAsync(o->seeds, dse, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
        if (o->volumes) HIPCHK(hipMemcpyAsync(o->volumes, dvo, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        if (o->status) HIPCHK(hipMemcpyAsync(o->status, dst, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
        if (o->kcert) HIPCHK(hipMemcpyAsync(o->kcert, dkc, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
        if (o->cell_centroid) HIPCHK(hipMemcpyAsync(o->cell_centroid, dcc, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
        const int64_t ng = n - m->n_in;
        if (o->ghost_ids && m->prm.boundary == PG_BOUNDARY_PERIODIC && ng > 0)
            HIPCHK(hipMemcpyAsync(o->ghost_ids, m->gsrc.p, sizeof(int64_t) * ng, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (int64_t i = 0; i < n; i++) {
            if (o->cell_ids) o->cell_ids[i] = i;
            if (o->face_counts) o->face_counts[i] = hcnt[i];
            if (o->face_ptr) o->face_ptr[i] = hptr[i];
        }
    });
}

int pg_mesh_free(pg_mesh** m) {
    return guarded([&] {
        if (!m || !*m) return;
        delete *m;
        *m = nullptr;
    });
}

// ---------------------------------------------------------------- hydro
int pg_state_create(pg_ctx* c, const pg_mesh* m, const double* rho, const double* vel, const double* E,
                    pg_state** out) {
    return guarded([&] {
        use(c);
        argchk(m && rho && vel && E && out, "null argument");
        argchk(m->n_hydro > 0, "mesh has no hydro cells");
        pg_state* st = new pg_state();
        try {
            st->dim = m->prm.dim;
            st->n = m->n_hydro;
            const size_t n = (size_t)st->n, d = (size_t)st->dim;
            for (int b = 0; b < 2; b++) {
                st->rho[b].get<double>(n);
                st->vel[b].get<double>(n * d);
                st->E[b].get<double>(n);
            }
            HIPCHK(hipMemcpyAsync(st->rho[0].p, rho, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(st->vel[0].p, vel, sizeof(double) * n * d, hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(st->E[0].p, E, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        } catch (...) {
            delete st;
            throw;
        }
        *out = st;
    });
}

int pg_dt_cfl(pg_ctx* c, const pg_mesh* mc, const pg_state* st, double gamma, double cfl, double* dt) {
    return guarded([&] {
        use(c);
        argchk(mc && st && dt, "null argument");
        pg_mesh* m = const_cast<pg_mesh*>(mc);
        argchk(st->n == m->n_hydro && st->dim == m->prm.dim, "state does not match mesh");
        mesh_finalize(c, m);
        DBuf out;
        unsigned long long* d = out.get<unsigned long long>(1);
        double init = 1e100;  // finite_volume_solver.cpp:153
        HIPCHK(hipMemcpyAsync(d, &init, sizeof(double), hipMemcpyHostToDevice, c->stream));
        const int b = st->cur;
        pg::launch_dt_cfl(st->dim, st->n, m->o2s.as<int32_t>(), m->vol_u.as<double>(), gamma, cfl,
                          st->rho[b].as<double>(), st->vel[b].as<double>(), st->E[b].as<double>(), d,
                          c->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(dt, d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int pg_hydro_step(pg_ctx* c, const pg_mesh* mc, pg_state* st, double gamma, double dt, int flux_mode) {
    return guarded([&] {
        use(c);
        argchk(mc && st, "null argument");
        argchk(flux_mode == PG_FLUX_REFERENCE || flux_mode == PG_FLUX_PHYSICAL, "unknown flux mode");
        pg_mesh* m = const_cast<pg_mesh*>(mc);
        argchk(st->n == m->n_hydro && st->dim == m->prm.dim, "state does not match mesh");
        mesh_finalize(c, m);
        const int b = st->cur, nb = 1 - b;
        pg::launch_hll_static(st->dim, m->n_seeds, m->n_hydro, m->n_seeds, m->n_seeds * MAXP,
                              m->out_order(), m->count.as<uint8_t>(), m->hid.as<int32_t>(),
                              m->nrm.as<double>(), m->area_u.as<double>(), m->vol_u.as<double>(), gamma, dt,
                              flux_mode, st->rho[b].as<double>(), st->vel[b].as<double>(),
                              st->E[b].as<double>(), st->rho[nb].as<double>(), st->vel[nb].as<double>(),
                              st->E[nb].as<double>(), c->stream);
        HIPCHK(hipGetLastError());
        st->cur = nb;
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int pg_state_copy_out(pg_ctx* c, const pg_state* st, double* rho, double* vel, double* E) {
    return guarded([&] {
        use(c);
        argchk(st != nullptr, "null state");
        const int b = st->cur;
        const size_t n = (size_t)st->n;
        if (rho) HIPCHK(hipMemcpyAsync(rho, st->rho[b].p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        if (vel) HIPCHK(hipMemcpyAsync(vel, st->vel[b].p, sizeof(double) * n * st->dim, hipMemcpyDeviceToHost, c->stream));
        if (E) HIPCHK(hipMemcpyAsync(E, st->E[b].p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int pg_hll_flux(pg_ctx* c, int dim, double gamma, int64_t n, const double* left, const double* right,
                const double* normal, const uint8_t* wall, int flux_mode, double* flux) {
    return guarded([&] {
        use(c);
        argchk(dim == 2 || dim == 3, "dim must be 2 or 3");
        argchk(gamma > 1.0, "gamma > 1 required");
        argchk(n >= 0 && n < (int64_t)1 << 28, "n out of range");
        argchk(flux_mode == PG_FLUX_REFERENCE || flux_mode == PG_FLUX_PHYSICAL, "unknown flux mode");
        if (n == 0) return;
        argchk(left && right && normal && flux, "null argument");
        hipStream_t s = c->stream;
        const size_t w = (size_t)dim + 2, nn = (size_t)n;
        DBuf &dl = c->hll_io[0], &dr = c->hll_io[1], &dn = c->hll_io[2], &dw = c->hll_io[3], &df = c->hll_io[4];
        HIPCHK(hipMemcpyAsync(dl.get<double>(nn * w), left, sizeof(double) * nn * w, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(dr.get<double>(nn * w), right, sizeof(double) * nn * w, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(dn.get<double>(nn * 3), normal, sizeof(double) * nn * 3, hipMemcpyHostToDevice, s));
        const uint8_t* dwp = nullptr;
        if (wall) {
            HIPCHK(hipMemcpyAsync(dw.get<uint8_t>(nn), wall, nn, hipMemcpyHostToDevice, s));
            dwp = dw.as<uint8_t>();
        }
        pg::launch_hll_faces(dim, n, gamma, dl.as<double>(), dr.as<double>(), dn.as<double>(), dwp, flux_mode,
                             df.get<double>(nn * w), s);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(flux, df.as<double>(), sizeof(double) * nn * w, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    });
}

int pg_state_free(pg_state** st) {
    return guarded([&] {
        if (!st || !*st) return;
        delete *st;
        *st = nullptr;
    });
}

}  // extern "C"

// ============================================== objects that fail mid-step
// A simulation or batch step that throws leaves the state advanced and the
// mesh tables possibly from two different builds: the object records the
// error and refuses every later call that would launch on it.
namespace {

template <class T>
void usable(const T* o, const char* what) {
    argchk(o != nullptr, (std::string("null ") + what).c_str());
    if (!o->failed.empty())
        throw Error{PG_ERR_FAILED, std::string("the ") + what + " failed in an earlier step (" + o->failed +
                                       ") and is unusable"};
}

// Run one advancing call; any error marks the object failed.
template <class T, class F>
void fail_guard(T* o, F&& f) {
    try {
        f();
    } catch (const Error& e) {
        o->failed = e.msg;
        throw;
    } catch (...) {
        o->failed = "unknown error";
        throw;
    }
}

}  // namespace

// ================================================================ simulation
struct pg_sim {
    pg_sim_params prm{};
    const pg_kmodel* km = nullptr;
    pg_mesh mesh;
    int64_t n = 0;
    DBuf x, mass, mom, en, rho, vel, E, dt, part, wgen;
    DBuf kc_prev;  // certified k of the last rebuild, by point id
    bool have_prev = false;
    // set when a step throws: the state was already advanced and the mesh
    // tables may belong to two different builds, so nothing may run on it
    std::string failed;
    Marks mk;
    pg_sim_times last{};
    ~pg_sim() {
        for (hipEvent_t& e : mk.ev)
            if (e) (void)hipEventDestroy(e);
    }
    pg::SimView view() {
        return pg::SimView{n, mass.as<double>(), mom.as<double>(), en.as<double>(),
                           rho.as<double>(), vel.as<double>(), E.as<double>()};
    }
};

namespace {

void sim_rebuild(pg_ctx* c, pg_sim* S, Marks* mk) {
    // a learned build clips seeds in the order of their last certified k
    S->mesh.order_prev = (S->km && S->have_prev) ? S->kc_prev.as<int32_t>() : nullptr;
    S->mesh.order_prev_n = S->n;
    mesh_build(c, &S->mesh, S->x.as<double>(), S->n, true, mk, S->km);
    if (S->mesh.info.n_failed)
        throw Error{PG_ERR_CELLS, std::to_string(S->mesh.info.n_failed) + " cells left uncertified"};
    S->mesh.finalized = false;
    mesh_finalize(c, &S->mesh);
    pg_mesh& m = S->mesh;
    pg::launch_face_twins(nullptr, m.n_seeds, m.n_seeds, m.hcode.as<int32_t>(), m.count.as<uint8_t>(),
                          m.nbr_cm.as<int32_t>(), m.twin.get<uint8_t>((size_t)m.n_seeds * MAXP),
                          m.solve.get<uint64_t>((size_t)m.n_seeds), c->stream);
    pg::launch_keep_kcert(nullptr, S->mesh.n_seeds, S->mesh.g.perm.as<uint32_t>(), nullptr,
                          S->mesh.kcert.as<int32_t>(), S->n, S->kc_prev.get<int32_t>((size_t)S->n), c->stream);
    HIPCHK(hipGetLastError());
    S->have_prev = true;
    if (mk) mk->mark(MK_FINAL, c->stream);
}

// primitives, CFL dt (device scalar), HLL update; drift when the mesh moves
void sim_hydro(pg_ctx* c, pg_sim* S, double dt_max, bool moving) {
    hipStream_t s = c->stream;
    pg_mesh& m = S->mesh;
    const int dim = S->prm.mesh.dim;
    const double sc = m.scale();
    pg::launch_prim_from_q(dim, S->view(), m.o2s.as<int32_t>(), m.vol_u.as<double>(), s);
    unsigned long long* d = S->dt.get<unsigned long long>(1);
    const double init = 1e100;
    HIPCHK(hipMemcpyAsync(d, &init, sizeof(double), hipMemcpyHostToDevice, s));
    const double* w = nullptr;  // steered generator velocities (regularised motion)
    if (moving && S->prm.move_mesh == PG_MOVE_REGULARIZED) {
        double* wg = S->wgen.get<double>((size_t)S->n * dim);
        pg::launch_gen_velocity(dim, S->n, S->x.as<double>(), m.o2s.as<int32_t>(), m.vol_u.as<double>(),
                                m.ccen.as<double>(), m.n_seeds, sc, S->prm.gamma, S->rho.as<double>(),
                                S->vel.as<double>(), S->E.as<double>(), wg, s);
        w = wg;
    }
    pg::CflMesh cm;
    if (w) cm = {m.count.as<uint8_t>(), m.nbr.as<int32_t>(), m.g.sorted.as<double>(), m.n_seeds, sc};
    pg::launch_dt_cfl(dim, S->n, m.o2s.as<int32_t>(), m.vol_u.as<double>(), S->prm.gamma, S->prm.cfl,
                      S->rho.as<double>(), S->vel.as<double>(), S->E.as<double>(), d, s, w, cm);
    const size_t nplane = (size_t)m.n_seeds * MAXP;
    pg::OwnedUpdate u;
    u.n = m.n_seeds;
    u.stride = m.n_seeds;
    u.nplane = (int64_t)nplane;
    u.hcode = m.hcode.as<int32_t>();
    u.count = m.count.as<uint8_t>();
    u.twin = m.twin.as<uint8_t>();
    u.solve = m.solve.as<uint64_t>();
    u.nbr = m.nbr.as<int32_t>();
    u.nrm = m.nrm.as<double>();
    u.area_u = m.area_u.as<double>();
    u.cen = m.cen.as<double>();
    u.sorted = m.g.sorted.as<double>();
    u.scale = sc;
    u.gam = S->prm.gamma;
    u.flux_mode = S->prm.flux_mode;
    u.move = moving ? 1 : 0;
    u.dt_dev = (const double*)d;
    u.dt_max = dt_max;
    u.sv = S->view();
    u.wgen = w;
    u.rec = m.frec.get<pg::CellRec>((size_t)m.n_seeds);
    u.fbuf = m.fbuf.get<double>((size_t)(dim + 2) * nplane);
    pg::launch_hll_owned(dim, u, s);
    if (moving)
        pg::launch_drift(dim, S->n, S->x.as<double>(), w ? w : S->vel.as<double>(), (const double*)d, dt_max,
                         S->prm.mesh.boundary == PG_BOUNDARY_PERIODIC ? 1 : 0, s);
    HIPCHK(hipGetLastError());
}

double sim_read_dt(pg_ctx* c, pg_sim* S, double dt_max) {
    double dt = 0;
    HIPCHK(hipMemcpyAsync(&dt, S->dt.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (dt_max > 0 && dt > dt_max) dt = dt_max;
    return dt;
}

}  // namespace

extern "C" {

int pg_sim_create(pg_ctx* c, const pg_sim_params* prm, const double* pts, const double* rho,
                  const double* vel, const double* E, int64_t n, pg_sim** out) {
    return guarded([&] {
        use(c);
        argchk(prm && pts && rho && vel && E && out, "null argument");
        const pg_mesh_params& mp = prm->mesh;
        argchk(mp.dim == 2 || mp.dim == 3, "dim must be 2 or 3");
        argchk(mp.boundary == PG_BOUNDARY_PERIODIC || mp.boundary == PG_BOUNDARY_REFLECTIVE,
               "a simulation needs a periodic or reflective box");
        argchk(mp.k >= 1 && mp.k <= 256, "k must be in [1, 256]");
        argchk(mp.k_max <= KNN_MAX_K, "k_max must be at most 1024");
        argchk(mp.max_p >= 2 * mp.dim + 1 && mp.max_p <= MAXP, "max_p must be in [2*dim+1, 64]");
        argchk(mp.max_t >= 8 && mp.max_t <= 128, "max_t must be in [8, 128]");
        argchk(mp.boundary != PG_BOUNDARY_PERIODIC || (mp.buff > 0 && mp.buff < 0.5), "buff must be in (0, 0.5)");
        argchk(prm->gamma > 1.0 && prm->cfl > 0.0, "gamma > 1 and cfl > 0 required");
        argchk(prm->flux_mode == PG_FLUX_REFERENCE || prm->flux_mode == PG_FLUX_PHYSICAL, "unknown flux mode");
        argchk(prm->move_mesh >= PG_MOVE_STATIC && prm->move_mesh <= PG_MOVE_REGULARIZED, "unknown move_mesh mode");
        argchk(mp.perturb >= 0.0 && mp.perturb <= 1e-3, "perturb must be in [0, 1e-3]");
        argchk(mp.exact == 0 || mp.exact == 1, "exact must be 0 or 1");
        argchk(n >= 2 && n < (int64_t)1 << 29, "n out of range");
        pg_sim* S = new pg_sim();
        try {
            S->prm = *prm;
            S->mesh.prm = prm->mesh;
            if (S->mesh.prm.k_max < S->mesh.prm.k) S->mesh.prm.k_max = S->mesh.prm.k;
            S->mesh.want_cm = true;  // face ownership of the moving-mesh update
            S->n = n;
            const int dim = mp.dim;
            const size_t nn = (size_t)n;
            for (hipEvent_t& e : S->mk.ev) HIPCHK(hipEventCreate(&e));
            S->mk.on = true;
            HIPCHK(hipMemcpyAsync(S->x.get<double>(nn * dim), pts, sizeof(double) * nn * dim,
                                  hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(S->rho.get<double>(nn), rho, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(S->vel.get<double>(nn * dim), vel, sizeof(double) * nn * dim,
                                  hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(S->E.get<double>(nn), E, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
            S->mass.get<double>(nn);
            S->mom.get<double>(nn * dim);
            S->en.get<double>(nn);
            sim_rebuild(c, S, nullptr);
            pg::launch_q_from_prim(dim, S->view(), S->mesh.o2s.as<int32_t>(), S->mesh.vol_u.as<double>(), c->stream);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(c->stream));
        } catch (...) {
            delete S;
            throw;
        }
        *out = S;
    });
}

int pg_sim_step(pg_ctx* c, pg_sim* S, double dt_max, double* dt_used) {
    return guarded([&] {
        use(c);
        usable(S, "simulation");
        fail_guard(S, [&] {
        hipStream_t s = c->stream;
        Marks& mk = S->mk;
        mk.mark(MK_START, s);
        sim_hydro(c, S, dt_max, S->prm.move_mesh != 0);
        mk.mark(MK_HYDRO, s);
        if (S->prm.move_mesh) {
            sim_rebuild(c, S, &mk);
        } else {
            // static mesh: rebuild anyway (same generators) so every step has
            // the same cost model; the result equals the current mesh
            sim_rebuild(c, S, &mk);
        }
        mk.mark(MK_END, s);
        HIPCHK(hipStreamSynchronize(s));
        auto el = [&](int a, int b) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, mk.ev[a], mk.ev[b]));
            return (double)ms;
        };
        pg_sim_times& t = S->last;
        t.hydro_ms = el(MK_START, MK_HYDRO);
        t.ghosts_ms = el(MK_HYDRO, MK_GHOSTS);
        t.grid_ms = el(MK_GHOSTS, MK_GRID);
        t.predict_ms = el(MK_GRID, MK_PREDICT);
        t.knn_ms = el(MK_PREDICT, MK_KNN);
        t.row_entries = S->mesh.info.row_entries;
        t.clip_ms = el(MK_KNN, MK_CLIP);
        t.restart_ms = el(MK_CLIP, MK_RESTART);
        t.final_ms = el(MK_RESTART, MK_FINAL);
        t.total_ms = el(MK_START, MK_END);
        t.n_restarted = S->mesh.info.n_restarted;
        t.k_used = S->mesh.info.k_used;
        double dt = sim_read_dt(c, S, dt_max);
        if (dt_used) *dt_used = dt;
        });
    });
}

int pg_sim_hydro_steps(pg_ctx* c, pg_sim* S, int nsteps, double dt_max, double* t_adv) {
    return guarded([&] {
        use(c);
        usable(S, "simulation");
        argchk(nsteps >= 0, "bad argument");
        fail_guard(S, [&] {
            double t = 0;
            for (int i = 0; i < nsteps; i++) {
                sim_hydro(c, S, dt_max, false);
                t += sim_read_dt(c, S, dt_max);
            }
            if (t_adv) *t_adv = t;
        });
    });
}

int pg_sim_times_last(const pg_sim* S, pg_sim_times* t) {
    return guarded([&] {
        argchk(S && t, "null argument");
        *t = S->last;
    });
}

int pg_sim_mesh(const pg_sim* S, const pg_mesh** m) {
    return guarded([&] {
        usable(S, "simulation");
        argchk(m != nullptr, "null argument");
        *m = &S->mesh;
    });
}

int pg_sim_get(pg_ctx* c, const pg_sim* Sc, double* pts, double* rho, double* vel, double* E) {
    return guarded([&] {
        use(c);
        usable(Sc, "simulation");
        pg_sim* S = const_cast<pg_sim*>(Sc);
        hipStream_t s = c->stream;
        const int dim = S->prm.mesh.dim;
        const size_t n = (size_t)S->n;
        pg::launch_prim_from_q(dim, S->view(), S->mesh.o2s.as<int32_t>(), S->mesh.vol_u.as<double>(), s);
        HIPCHK(hipGetLastError());
        if (pts) HIPCHK(hipMemcpyAsync(pts, S->x.p, sizeof(double) * n * dim, hipMemcpyDeviceToHost, s));
        if (rho) HIPCHK(hipMemcpyAsync(rho, S->rho.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        if (vel) HIPCHK(hipMemcpyAsync(vel, S->vel.p, sizeof(double) * n * dim, hipMemcpyDeviceToHost, s));
        if (E) HIPCHK(hipMemcpyAsync(E, S->E.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    });
}

int pg_sim_totals(pg_ctx* c, pg_sim* S, double* tot) {
    return guarded([&] {
        use(c);
        usable(S, "simulation");
        argchk(tot != nullptr, "null argument");
        const int dim = S->prm.mesh.dim;
        const int64_t nb = pg::totals_blocks(S->n);
        double* part = S->part.get<double>((size_t)5 * nb);
        pg::launch_totals(dim, S->view(), part, c->stream);
        HIPCHK(hipGetLastError());
        std::vector<double> h((size_t)5 * nb);
        HIPCHK(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int q = 0; q < 5; q++) tot[q] = 0;
        for (int64_t b = 0; b < nb; b++)
            for (int q = 0; q < 5; q++) tot[q] += h[(size_t)5 * b + q];
    });
}

int pg_sim_free(pg_sim** S) {
    return guarded([&] {
        if (!S || !*S) return;
        delete *S;
        *S = nullptr;
    });
}

int pg_sim_set_kmodel(pg_sim* S, pg_kmodel* km) {
    return guarded([&] {
        argchk(S != nullptr, "null sim");
        S->km = km;
    });
}

// ---------------------------------------------------------------- learned k
namespace {

uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

void km_use(pg_ctx* c, const pg_kmodel* km) {
    use(c);
    argchk(km != nullptr, "null model");
    argchk(km->device == c->device, "model lives on another device");
}

// Forward + backward of rows X[0..B) (device) on the stream; the summed
// gradient lands in km->grad and the loss in km->loss.  lr > 0 also takes
// one Adam step.
void km_grad(pg_ctx* c, pg_kmodel* km, const float* X, const int32_t* y, int64_t n, int64_t B,
             uint64_t seed, float lr) {
    const int64_t nblk = pg::lk_grad_blocks(B);
    float* part = km->part.get<float>((size_t)nblk * pg::LK_NPARAM);
    float* pl = km->ploss.get<float>((size_t)nblk);
    pg::launch_kgrad(X, y, n, B, seed, km->step, km->P.as<float>(), part, pl, c->stream);
    pg::launch_kadam(part, pl, (int)nblk, km->P.as<float>(), km->m.as<float>(), km->v.as<float>(),
                     km->grad.get<float>(pg::LK_NPARAM), km->loss.get<float>(1), km->step, lr, c->stream);
    HIPCHK(hipGetLastError());
    if (lr > 0.f) km->step++;
}

void km_upload_batch(pg_ctx* c, pg_kmodel* km, const float* feat, const int32_t* cls, int64_t n) {
    HIPCHK(hipMemcpyAsync(km->X.get<float>((size_t)n * pg::LK_NFEAT), feat, sizeof(float) * n * pg::LK_NFEAT,
                          hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(km->y.get<int32_t>((size_t)n), cls, sizeof(int32_t) * n, hipMemcpyHostToDevice,
                          c->stream));
}

double km_read_loss(pg_ctx* c, pg_kmodel* km) {
    float l = 0.f;
    HIPCHK(hipMemcpyAsync(&l, km->loss.p, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return (double)l;
}

void check_classes(const int32_t* cls, int64_t n) {
    for (int64_t i = 0; i < n; i++) argchk(cls[i] >= 0 && cls[i] < pg::LK_NCLASS, "class out of range");
}

}  // namespace

int pg_kmodel_create(pg_ctx* c, uint64_t seed, pg_kmodel** out) {
    return guarded([&] {
        use(c);
        argchk(out != nullptr, "null out");
        static_assert(PG_KMODEL_NPARAM == pg::LK_NPARAM, "parameter count");
        static_assert(PG_KMODEL_NFEAT == pg::LK_NFEAT && PG_KMODEL_NCLASS == pg::LK_NCLASS, "model shape");
        pg_kmodel* km = new pg_kmodel();
        try {
            km->device = c->device;
            std::vector<float> h(pg::LK_NPARAM, 0.f);
            const int fan[3][2] = {{16, 64}, {64, 64}, {64, 6}};
            const int off[3] = {0, 16 * 64 + 64, 16 * 64 + 64 + 64 * 64 + 64};
            uint64_t st = seed;
            for (int L = 0; L < 3; L++) {
                const double a = std::sqrt(6.0 / (fan[L][0] + fan[L][1]));
                for (int i = 0; i < fan[L][0] * fan[L][1]; i++) {
                    const double u = (double)(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
                    h[off[L] + i] = (float)((2.0 * u - 1.0) * a);
                }
            }
            HIPCHK(hipMemcpyAsync(km->P.get<float>(pg::LK_NPARAM), h.data(), sizeof(float) * pg::LK_NPARAM,
                                  hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipMemsetAsync(km->m.get<float>(pg::LK_NPARAM), 0, sizeof(float) * pg::LK_NPARAM, c->stream));
            HIPCHK(hipMemsetAsync(km->v.get<float>(pg::LK_NPARAM), 0, sizeof(float) * pg::LK_NPARAM, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        } catch (...) {
            delete km;
            throw;
        }
        *out = km;
    });
}

int pg_kmodel_set_params(pg_ctx* c, pg_kmodel* km, const float* params) {
    return guarded([&] {
        km_use(c, km);
        argchk(params != nullptr, "null params");
        HIPCHK(hipMemcpyAsync(km->P.p, params, sizeof(float) * pg::LK_NPARAM, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemsetAsync(km->m.p, 0, sizeof(float) * pg::LK_NPARAM, c->stream));
        HIPCHK(hipMemsetAsync(km->v.p, 0, sizeof(float) * pg::LK_NPARAM, c->stream));
        km->step = 0;
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int pg_kmodel_get_params(pg_ctx* c, const pg_kmodel* km, float* params) {
    return guarded([&] {
        km_use(c, km);
        argchk(params != nullptr, "null params");
        HIPCHK(hipMemcpyAsync(params, km->P.p, sizeof(float) * pg::LK_NPARAM, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int pg_kmodel_set_cost(pg_kmodel* km, double a0) {
    return guarded([&] {
        argchk(km != nullptr, "null model");
        argchk(a0 >= 0.0 && a0 < 1e6, "a0 out of range");
        km->a0 = (float)a0;
    });
}

int pg_kmodel_predict(pg_ctx* c, const pg_kmodel* kmc, const float* feat, int64_t n, int32_t* cls,
                      float* logits) {
    return guarded([&] {
        km_use(c, kmc);
        argchk(feat && cls && n >= 1 && n < (int64_t)1 << 31, "bad argument");
        pg_kmodel* km = const_cast<pg_kmodel*>(kmc);
        float* X = km->X.get<float>((size_t)n * pg::LK_NFEAT);
        HIPCHK(hipMemcpyAsync(X, feat, sizeof(float) * n * pg::LK_NFEAT, hipMemcpyHostToDevice, c->stream));
        int32_t* dc = km->cls.get<int32_t>((size_t)n);
        float* dl = logits ? km->logits.get<float>((size_t)n * pg::LK_NCLASS) : nullptr;
        pg::launch_kpredict(X, n, km->P.as<float>(), km->a0, nullptr, dc, dl, c->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(cls, dc, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
        if (logits)
            HIPCHK(hipMemcpyAsync(logits, dl, sizeof(float) * n * pg::LK_NCLASS, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int pg_kmodel_grad(pg_ctx* c, const pg_kmodel* kmc, const float* feat, const int32_t* cls, int64_t n,
                   float* grad, double* loss) {
    return guarded([&] {
        km_use(c, kmc);
        argchk(feat && cls && grad && n >= 1 && n < (int64_t)1 << 31, "bad argument");
        check_classes(cls, n);
        pg_kmodel* km = const_cast<pg_kmodel*>(kmc);
        km_upload_batch(c, km, feat, cls, n);
        km_grad(c, km, km->X.as<float>(), km->y.as<int32_t>(), n, n, 0, 0.f);
        HIPCHK(hipMemcpyAsync(grad, km->grad.p, sizeof(float) * pg::LK_NPARAM, hipMemcpyDeviceToHost, c->stream));
        const double l = km_read_loss(c, km);
        if (loss) *loss = l;
    });
}

int pg_kmodel_train_step(pg_ctx* c, pg_kmodel* km, const float* feat, const int32_t* cls, int64_t n, double lr,
                         double* loss) {
    return guarded([&] {
        km_use(c, km);
        argchk(feat && cls && n >= 1 && n < (int64_t)1 << 31 && lr > 0.0, "bad argument");
        check_classes(cls, n);
        km_upload_batch(c, km, feat, cls, n);
        km_grad(c, km, km->X.as<float>(), km->y.as<int32_t>(), n, n, 0, (float)lr);
        const double l = km_read_loss(c, km);
        if (loss) *loss = l;
    });
}

int pg_kmodel_train_mesh(pg_ctx* c, pg_kmodel* km, const pg_mesh* mc, int steps, int64_t batch, double lr,
                         uint64_t seed, double* loss) {
    return guarded([&] {
        km_use(c, km);
        argchk(mc != nullptr, "null mesh");
        argchk(steps >= 1 && batch >= 16 && batch <= (1 << 22) && lr > 0.0, "bad training argument");
        pg_mesh* m = const_cast<pg_mesh*>(mc);
        const int64_t n = m->n_seeds;
        argchk(n >= 1 && m->status.p, "mesh is not built");
        float* X = km->X.get<float>((size_t)n * pg::LK_NFEAT);
        int32_t* y = km->y.get<int32_t>((size_t)n);
        pg::launch_kfeatures(m->prm.dim, m->g.sorted.as<double>(), n, m->g.N, m->g.counts.as<int32_t>(),
                             m->g.starts.as<int32_t>(), nullptr, X, c->stream, m->g.sub_any());
        pg::launch_klabels(m->kcert.as<int32_t>(), n, y, c->stream);
        HIPCHK(hipGetLastError());
        const uint64_t sd = seed ? seed : 1;
        for (int t = 0; t < steps; t++) km_grad(c, km, X, y, n, batch, sd, (float)lr);
        const double l = km_read_loss(c, km);
        if (loss) *loss = l;
    });
}

int pg_kmodel_free(pg_kmodel** km) {
    return guarded([&] {
        if (!km || !*km) return;
        delete *km;
        *km = nullptr;
    });
}

int pg_mesh_features(pg_ctx* c, const pg_mesh* mc, float* out) {
    return guarded([&] {
        use(c);
        argchk(mc && out, "null argument");
        pg_mesh* m = const_cast<pg_mesh*>(mc);
        const int64_t n = m->n_seeds;
        argchk(n >= 1 && m->g.sorted.p, "mesh is not built");
        DBuf tmp;
        float* d = tmp.get<float>((size_t)n * pg::LK_NFEAT);
        pg::launch_kfeatures(m->prm.dim, m->g.sorted.as<double>(), n, m->g.N, m->g.counts.as<int32_t>(),
                             m->g.starts.as<int32_t>(), m->g.perm.as<uint32_t>(), d, c->stream, m->g.sub_any());
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, d, sizeof(float) * n * pg::LK_NFEAT, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

}  // extern "C"

// ================================================================== batches
struct pg_batch {
    pg_sim_params prm{};
    std::string failed;  // error of a failed step (the batch is then unusable)
    int nbox = 0, dim = 3;
    bool periodic = true;
    int64_t n_h = 0, cap = 0, cell_cap = 0;
    std::vector<int64_t> hoff_h;
    std::vector<int> ksched, ksched_learned;
    // tuning (timing only; results do not depend on it)
    int lanes = 16, slots = 64, wave_min_k = 20, bpc = 64;
    bool use_graph = true;
    const pg_kmodel* km = nullptr;  // learned k (nullptr: fixed k)
    // static
    DBuf hoff, hbox, x, xs;
    DBuf feat, kcls, klists, kcnt;  // learned k
    DBuf kc_prev, ovals;  // clip order: last certified k per hydro cell, seeds in that order
    // conserved state and primitives (hydro order)
    DBuf mass, mom, en, rho, vel, E, dt_box, wgen;
    DBuf mass_s, mom_s, en_s;  // autotuner save area
    // per step
    DBuf gcnt, goff, gtot, soff, N, coff, stat, spts, sbox, orig_h;
    DBuf cell, counts, starts, fill, tmp, total, perm, sorted, sorted_box, big;
    DBuf knn, lists, ncount, flags, offs;
    DBuf nbr, area, nrm, cen, count, vol, ccen, status, kcert;
    DBuf cstate, resume, lastnb;  // saved cells of uncertified seeds (restarts resume them)
    DBuf ticket;          // work-queue counter of the persistent kernels
    DBuf hid, area_u, vol_u, o2s, hsorted, tot;
    DBuf hcode, nbr_cm, twin, solve, frec, fbuf;  // owner-computes-once update
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float last_ms = 0.f;
    ~pg_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
    pg::SimView view() {
        return pg::SimView{n_h, mass.as<double>(), mom.as<double>(), en.as<double>(), rho.as<double>(),
                           vel.as<double>(), E.as<double>()};
    }
    pg::FaceArrays faces() {
        pg::FaceArrays fa;
        fa.stride = cap;
        fa.nbr = nbr.as<int32_t>();
        fa.area = area.as<double>();
        fa.nrm = nrm.as<double>();
        fa.cen = cen.as<double>();
        fa.count = count.as<uint8_t>();
        fa.vol = vol.as<double>();
        fa.ccen = ccen.as<double>();
        fa.status = status.as<int32_t>();
        fa.kcert = kcert.as<int32_t>();
        if (resume.p) {
            fa.cstate = cstate.as<uint8_t>();
            fa.resume = resume.as<int32_t>();
            fa.lastnb = lastnb.as<uint32_t>();
        }
        fa.nbr_cm = nbr_cm.as<int32_t>();
        return fa;
    }
    double scale() const { return periodic ? 1. + (2 * prm.mesh.buff) : 1.0; }
    int64_t* n_dev() { return soff.as<int64_t>() + nbox; }
    // status block: [0] capacity error bits, [1] hydro cells without faces,
    // [2] periodic hydro faces on the box plane, [3] uncertified seeds
    int32_t* st() { return stat.as<int32_t>(); }
};

namespace {

// The rebuild of every box (ghosts, grids, KNN, clip, restarts, hydro
// tables), enqueued on the stream with device-side counts only.
void batch_rebuild(pg_batch* B, hipStream_t s) {
    const int dim = B->dim;
    const pg_mesh_params& p = B->prm.mesh;
    const int64_t cap = B->cap;
    pg::launch_fill_i32(B->st(), 4, 0, s);
    // saved cells belong to this rebuild only
    if (B->resume.p) pg::launch_fill_i32(B->resume.as<int32_t>(), cap, 0, s);
    const int32_t* goff = nullptr;
    if (B->periodic) {
        pg::launch_ghost_count(dim, B->x.as<double>(), B->n_h, p.buff, B->gcnt.as<int32_t>(), s);
        pg::scan_exclusive_i32(B->gcnt.as<int32_t>(), B->goff.as<int32_t>(), B->n_h, B->tmp.as<int32_t>(),
                               B->gtot.as<int32_t>(), s);
        goff = B->goff.as<int32_t>();
    }
    pg::launch_box_setup(B->nbox, dim, B->hoff.as<int64_t>(), goff, B->gtot.as<int32_t>(), cap, B->cell_cap,
                         B->soff.as<int64_t>(), B->N.as<int32_t>(), B->coff.as<int64_t>(), B->st(), s);
    pg::launch_bseeds(dim, B->x.as<double>(), B->n_h, p.buff, B->periodic ? 1 : 0, B->hbox.as<int32_t>(),
                      B->hoff.as<int64_t>(), goff, B->soff.as<int64_t>(), B->st(), B->spts.as<double>(),
                      B->sbox.as<int32_t>(), B->orig_h.as<int32_t>(), s);
    const int64_t* n_dev = B->n_dev();
    pg::launch_bgrid(dim, B->spts.as<double>(), n_dev, cap, B->cell_cap, B->sbox.as<int32_t>(), B->N.as<int32_t>(),
                     B->coff.as<int64_t>(), B->cell.as<int32_t>(), B->counts.as<int32_t>(), B->starts.as<int32_t>(),
                     B->fill.as<int32_t>(), B->tmp.as<int32_t>(), B->total.as<int32_t>(), B->perm.as<uint32_t>(),
                     B->sorted.as<double>(), B->sorted_box.as<int32_t>(), B->big.as<int32_t>(), s);
    auto knn = [&](const uint32_t* q, const int64_t* nq, int k, bool big) {
        pg::KnnTail tail;  // seeds that resume need only the rest of their row
        if (B->resume.p && !big)
            tail = {B->resume.as<int32_t>(), B->lastnb.as<uint32_t>(), clip_shape_tag(B->lanes, B->slots)};
        pg::launch_bknn(dim, B->sorted.as<double>(), B->perm.as<uint32_t>(), B->sorted_box.as<int32_t>(),
                        B->N.as<int32_t>(), B->coff.as<int64_t>(), B->starts.as<int32_t>(), B->counts.as<int32_t>(),
                        q, nq, cap, k, B->knn.as<uint32_t>(), B->wave_min_k, B->bpc,
                        B->ticket.as<unsigned long long>(), s, tail);
    };
    // 3D first pass and learned buckets: one block per seed group (the 3D
    // clip spills in a persistent loop); the short restart lists, and every
    // 2D list (whose clip does not spill): persistent blocks
    auto clip = [&](const uint32_t* q, const int64_t* nq, int k, int big, bool restart_list) {
        pg::launch_bvoronoi(dim, B->sorted.as<double>(), q, nq, cap, B->knn.as<uint32_t>(), k, p.max_p, p.max_t, big,
                            B->lanes, B->slots, B->faces(), B->bpc, B->ticket.as<unsigned long long>(),
                            restart_list || dim == 2, s);
    };
    auto collect = [&](const uint32_t* L, const int64_t* nL, int code, uint32_t* out, int64_t* nout) {
        pg::launch_bcollect(nL, cap, L, B->status.as<int32_t>(), code, out, B->flags.as<int32_t>(),
                            B->offs.as<int32_t>(), B->tmp.as<int32_t>(), B->total.as<int32_t>(), nout, s);
    };
    // first pass: every seed at the fixed k, or every bucket of the learned
    // k at its own k; then the restart chain of the single-box build
    const std::vector<int>& sched = B->km ? B->ksched_learned : B->ksched;
    int k = sched[0];
    if (!B->km) {
        knn(nullptr, n_dev, k, false);
        clip(nullptr, n_dev, k, 0, false);
    } else {
        float* feat = B->feat.as<float>();
        int32_t* cls = B->kcls.as<int32_t>();
        pg::launch_bkfeatures(dim, B->sorted.as<double>(), n_dev, cap, B->sorted_box.as<int32_t>(),
                              B->soff.as<int64_t>(), B->N.as<int32_t>(), B->coff.as<int64_t>(), B->counts.as<int32_t>(),
                              B->starts.as<int32_t>(), feat, s);
        pg::launch_kpredict(feat, cap, B->km->P.as<float>(), B->km->a0, nullptr, cls, nullptr, s, n_dev);
        // seeds in the order of their certified k of the last rebuild (all
        // zero before the first: ascending ids), kept inside every bucket
        const uint32_t* order = nullptr;
        if (lk_order_on()) {
            pg::launch_order_list(nullptr, n_dev, cap, B->perm.as<uint32_t>(), B->orig_h.as<int32_t>(),
                                  B->kc_prev.as<int32_t>(), B->n_h, B->ovals.as<uint32_t>(), s);
            order = B->ovals.as<uint32_t>();
        }
        pg::launch_kbucket(cls, n_dev, cap, B->flags.as<int32_t>(), B->offs.as<int32_t>(), B->tmp.as<int32_t>(),
                           B->total.as<int32_t>(), cap, B->klists.as<uint32_t>(), B->kcnt.as<int64_t>(), s, order);
        for (int c = 0; c < pg::LK_NCLASS; c++) {
            const uint32_t* lc = B->klists.as<uint32_t>() + (size_t)c * cap;
            knn(lc, B->kcnt.as<int64_t>() + c, pg::lk_bucket(c), false);
            clip(lc, B->kcnt.as<int64_t>() + c, pg::lk_bucket(c), 0, false);
        }
    }
    uint32_t* lst[3] = {B->lists.as<uint32_t>(), B->lists.as<uint32_t>() + cap, B->lists.as<uint32_t>() + 2 * cap};
    int64_t* cnt[3] = {B->ncount.as<int64_t>(), B->ncount.as<int64_t>() + 1, B->ncount.as<int64_t>() + 2};
    const uint32_t* L = nullptr;
    const int64_t* nL = n_dev;
    // overflowed seeds keep that status through the chain; the 128-slot
    // kernel clips all of them once the chain's k would pass big_row, and
    // after every longer step, whose list is then taken over the whole batch
    // (see restart_chain)
    const int kb = big_row(k, sched.back(), p.restart != 0);
    bool big_done = false;
    auto big_pass = [&](int kk) {
        collect(nullptr, n_dev, pg::SLOT_OVERFLOW, lst[2], cnt[2]);
        knn(lst[2], cnt[2], kk, true);
        clip(lst[2], cnt[2], kk, 1, true);
        L = nullptr;
        nL = n_dev;
    };
    for (size_t r = 1; p.restart && r < sched.size(); r++) {
        if (!big_done && sched[r] > kb) {
            big_pass(kb);
            big_done = true;
        }
        const int d = (L == lst[0]) ? 1 : 0;
        collect(L, nL, PG_SECURITY_RADIUS_NOT_REACHED, lst[d], cnt[d]);
        k = sched[r];
        knn(lst[d], cnt[d], k, false);
        clip(lst[d], cnt[d], k, 0, true);
        L = lst[d];
        nL = cnt[d];
        if (big_done) big_pass(k);
    }
    if (!big_done) big_pass(kb);
    const double sc = B->scale();
    const double as = dim == 2 ? sc : sc * sc, vs = dim == 2 ? sc * sc : sc * sc * sc;
    pg::launch_bfinalize(n_dev, cap, cap, B->perm.as<uint32_t>(), B->orig_h.as<int32_t>(), B->count.as<uint8_t>(),
                         B->nbr.as<int32_t>(), B->area.as<double>(), B->vol.as<double>(), as, vs, B->periodic ? 1 : 0,
                         B->hid.as<int32_t>(), B->area_u.as<double>(), B->vol_u.as<double>(), B->o2s.as<int32_t>(),
                         B->hsorted.as<int32_t>(), B->hcode.as<int32_t>(), B->st() + 1, s);
    pg::launch_face_twins(n_dev, cap, cap, B->hcode.as<int32_t>(), B->count.as<uint8_t>(), B->nbr_cm.as<int32_t>(),
                          B->twin.as<uint8_t>(), B->solve.as<uint64_t>(), s);
    pg::launch_bcount_failed(n_dev, cap, B->status.as<int32_t>(), B->st() + 3, s);
    if (B->km && lk_order_on())
        pg::launch_keep_kcert(n_dev, cap, B->perm.as<uint32_t>(), B->orig_h.as<int32_t>(), B->kcert.as<int32_t>(),
                              B->n_h, B->kc_prev.as<int32_t>(), s);
}

// primitives, per-box CFL step, ALE HLL update, drift
void batch_hydro(pg_batch* B, hipStream_t s) {
    const int dim = B->dim;
    const bool move = B->prm.move_mesh != 0;
    pg::launch_prim_from_q(dim, B->view(), B->o2s.as<int32_t>(), B->vol_u.as<double>(), s);
    const double* w = nullptr;  // steered generator velocities (regularised motion)
    if (move && B->prm.move_mesh == PG_MOVE_REGULARIZED) {
        double* wg = B->wgen.as<double>();
        pg::launch_gen_velocity(dim, B->n_h, B->x.as<double>(), B->o2s.as<int32_t>(), B->vol_u.as<double>(),
                                B->ccen.as<double>(), B->cap, B->scale(), B->prm.gamma, B->rho.as<double>(),
                                B->vel.as<double>(), B->E.as<double>(), wg, s);
        w = wg;
    }
    pg::launch_bdt(dim, B->n_h, B->nbox, B->hbox.as<int32_t>(), B->o2s.as<int32_t>(), B->vol_u.as<double>(),
                   B->prm.gamma, B->prm.cfl, B->rho.as<double>(), B->vel.as<double>(), B->E.as<double>(),
                   B->dt_box.as<double>(), s, w,
                   w ? pg::CflMesh{B->count.as<uint8_t>(), B->nbr.as<int32_t>(), B->sorted.as<double>(), B->cap,
                                   B->scale()}
                     : pg::CflMesh{});
    pg::OwnedUpdate u;
    u.n_dev = B->n_dev();
    u.n = B->cap;
    u.stride = B->cap;
    u.nplane = B->cap * MAXP;
    u.hcode = B->hcode.as<int32_t>();
    u.count = B->count.as<uint8_t>();
    u.twin = B->twin.as<uint8_t>();
    u.solve = B->solve.as<uint64_t>();
    u.nbr = B->nbr.as<int32_t>();
    u.nrm = B->nrm.as<double>();
    u.area_u = B->area_u.as<double>();
    u.cen = B->cen.as<double>();
    u.sorted = B->sorted.as<double>();
    u.scale = B->scale();
    u.gam = B->prm.gamma;
    u.flux_mode = B->prm.flux_mode;
    u.move = move ? 1 : 0;
    u.dt_box = B->dt_box.as<double>();
    u.hbox = B->hbox.as<int32_t>();
    u.sv = B->view();
    u.wgen = w;
    u.rec = B->frec.as<pg::CellRec>();
    u.fbuf = B->fbuf.as<double>();
    pg::launch_hll_owned(dim, u, s);
    if (move)
        pg::launch_bdrift(dim, B->n_h, B->hbox.as<int32_t>(), B->x.as<double>(), w ? w : B->vel.as<double>(),
                          B->dt_box.as<double>(), B->periodic ? 1 : 0, s);
}

void batch_check(pg_ctx* c, pg_batch* B) {
    int32_t h[4] = {0, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(h, B->st(), sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (h[0]) throw Error{PG_ERR_CELLS, "batch: seeds or grid cells exceed the capacity reserved at creation"};
    if (h[3]) throw Error{PG_ERR_CELLS, std::to_string(h[3]) + " cells left uncertified"};
    if (h[1]) throw Error{PG_ERR_CELLS, "hydro cells without faces (uncertified cells)"};
    if (h[2])
        throw Error{PG_ERR_CELLS, std::to_string(h[2]) +
                                      " hydro faces reach the box plane: the periodic ghost layer (buff) is "
                                      "thinner than the cells; increase buff"};
}

void batch_step_eager(pg_ctx* c, pg_batch* B) {
    batch_hydro(B, c->stream);
    batch_rebuild(B, c->stream);
    HIPCHK(hipGetLastError());
}

void batch_capture(pg_ctx* c, pg_batch* B) {
    if (B->exec) {
        HIPCHK(hipGraphExecDestroy(B->exec));
        B->exec = nullptr;
    }
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    try {
        batch_hydro(B, c->stream);
        batch_rebuild(B, c->stream);
    } catch (...) {
        // end the capture so that the stream is usable again
        if (hipStreamEndCapture(c->stream, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        throw;
    }
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        hipchk(e, "hipStreamEndCapture");
    }
    e = hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    hipchk(e, "hipGraphInstantiate");
}

void batch_step(pg_ctx* c, pg_batch* B) {
    HIPCHK(hipEventRecord(B->e0, c->stream));
    if (B->use_graph) {
        if (!B->exec) batch_capture(c, B);
        HIPCHK(hipGraphLaunch(B->exec, c->stream));
    } else {
        batch_step_eager(c, B);
    }
    HIPCHK(hipEventRecord(B->e1, c->stream));
    batch_check(c, B);
    HIPCHK(hipEventElapsedTime(&B->last_ms, B->e0, B->e1));
}

void batch_copy_state(pg_batch* B, bool save, hipStream_t s) {
    const size_t n = (size_t)B->n_h, d = (size_t)B->dim;
    auto cp = [&](DBuf& a, DBuf& b, size_t bytes) {
        if (save) HIPCHK(hipMemcpyAsync(b.p, a.p, bytes, hipMemcpyDeviceToDevice, s));
        else HIPCHK(hipMemcpyAsync(a.p, b.p, bytes, hipMemcpyDeviceToDevice, s));
    };
    cp(B->x, B->xs, sizeof(double) * n * d);
    cp(B->mass, B->mass_s, sizeof(double) * n);
    cp(B->mom, B->mom_s, sizeof(double) * n * d);
    cp(B->en, B->en_s, sizeof(double) * n);
}

}  // namespace

extern "C" {

int pg_batch_create(pg_ctx* c, const pg_sim_params* prm, int nbox, const int64_t* n_per_box, const double* pts,
                    const double* rho, const double* vel, const double* E, pg_batch** out) {
    return guarded([&] {
        use(c);
        argchk(prm && n_per_box && pts && rho && vel && E && out, "null argument");
        const pg_mesh_params& mp = prm->mesh;
        argchk(mp.dim == 2 || mp.dim == 3, "dim must be 2 or 3");
        argchk(mp.boundary == PG_BOUNDARY_PERIODIC || mp.boundary == PG_BOUNDARY_REFLECTIVE,
               "a batch needs periodic or reflective boxes");
        argchk(mp.perturb == 0.0, "batches do not apply the degeneracy fallback (perturb must be 0)");
        argchk(mp.exact == 0, "batches do not apply the exact-predicate fallback (exact must be 0)");
        argchk(mp.k >= 1 && mp.k <= 256 && mp.k_max <= 256, "k and k_max must be in [1, 256]");
        argchk(mp.max_p >= 2 * mp.dim + 1 && mp.max_p <= MAXP, "max_p must be in [2*dim+1, 64]");
        argchk(mp.max_t >= 8 && mp.max_t <= 128, "max_t must be in [8, 128]");
        argchk(mp.boundary != PG_BOUNDARY_PERIODIC || (mp.buff > 0 && mp.buff < 0.5), "buff must be in (0, 0.5)");
        argchk(prm->gamma > 1.0 && prm->cfl > 0.0, "gamma > 1 and cfl > 0 required");
        argchk(prm->flux_mode == PG_FLUX_REFERENCE || prm->flux_mode == PG_FLUX_PHYSICAL, "unknown flux mode");
        argchk(prm->move_mesh >= PG_MOVE_STATIC && prm->move_mesh <= PG_MOVE_REGULARIZED, "unknown move_mesh mode");
        argchk(nbox >= 1 && nbox <= 4096, "nbox must be in [1, 4096]");
        pg_batch* B = new pg_batch();
        try {
            B->prm = *prm;
            if (B->prm.mesh.k_max < B->prm.mesh.k) B->prm.mesh.k_max = B->prm.mesh.k;
            B->nbox = nbox;
            B->dim = mp.dim;
            B->periodic = mp.boundary == PG_BOUNDARY_PERIODIC;
            const int kneed = std::max(B->prm.mesh.k, B->prm.mesh.k_max);
            B->hoff_h.assign(nbox + 1, 0);
            std::vector<int32_t> hb;
            int64_t cap = 0, ccap = 0;
            const double gfrac = B->periodic ? std::pow(1.0 + 2.0 * mp.buff, mp.dim) - 1.0 : 0.0;
            for (int b = 0; b < nbox; b++) {
                const int64_t nb = n_per_box[b];
                argchk(nb >= kneed + 2 && nb < (int64_t)1 << 26, "every box needs more than k_max + 1 points");
                B->hoff_h[b + 1] = B->hoff_h[b] + nb;
                const int64_t cb = nb + (int64_t)std::ceil(2.0 * gfrac * (double)nb) + 64;
                cap += cb;
                const int64_t Nb = ngrid(mp.dim, cb) + 1;
                ccap += mp.dim == 2 ? Nb * Nb : Nb * Nb * Nb;
                if (!B->periodic) argchk(ngrid(mp.dim, nb) >= 1, "box too small");
            }
            argchk(cap < (int64_t)1 << 30, "batch too large");
            B->n_h = B->hoff_h[nbox];
            B->cap = cap;
            B->cell_cap = ccap;
            hb.resize(B->n_h);
            for (int b = 0; b < nbox; b++)
                for (int64_t i = B->hoff_h[b]; i < B->hoff_h[b + 1]; i++) hb[i] = b;
            // restart schedule of the single-box build (first k, then next_k)
            B->ksched.push_back(B->prm.mesh.k);
            if (B->prm.mesh.restart)
                for (int kk = next_k(B->prm.mesh.k, B->prm.mesh.k_max); kk > 0; kk = next_k(kk, B->prm.mesh.k_max))
                    B->ksched.push_back(kk);
            // learned k: after the bucket pass every rejected seed continues
            // from 64 (a recomputation at any row length at least its
            // certified k gives the same cell)
            B->ksched_learned.push_back(std::min(64, B->prm.mesh.k_max));
            if (B->prm.mesh.restart)
                for (int kk = next_k(B->ksched_learned[0], B->prm.mesh.k_max); kk > 0;
                     kk = next_k(kk, B->prm.mesh.k_max))
                    B->ksched_learned.push_back(kk);
            const int kmax = std::max(96, *std::max_element(B->ksched.begin(), B->ksched.end()));
            const size_t n = (size_t)B->n_h, d = (size_t)mp.dim, cs = (size_t)cap;
            hipStream_t s = c->stream;
            HIPCHK(hipMemcpyAsync(B->hoff.get<int64_t>(nbox + 1), B->hoff_h.data(), sizeof(int64_t) * (nbox + 1),
                                  hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(B->hbox.get<int32_t>(n), hb.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(B->x.get<double>(n * d), pts, sizeof(double) * n * d, hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(B->rho.get<double>(n), rho, sizeof(double) * n, hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(B->vel.get<double>(n * d), vel, sizeof(double) * n * d, hipMemcpyHostToDevice, s));
            if (B->prm.move_mesh == PG_MOVE_REGULARIZED) B->wgen.get<double>(n * d);
            HIPCHK(hipMemcpyAsync(B->E.get<double>(n), E, sizeof(double) * n, hipMemcpyHostToDevice, s));
            B->xs.get<double>(n * d);
            B->mass.get<double>(n);
            B->mom.get<double>(n * d);
            B->en.get<double>(n);
            B->mass_s.get<double>(n);
            B->mom_s.get<double>(n * d);
            B->en_s.get<double>(n);
            B->dt_box.get<double>(nbox);
            B->gcnt.get<int32_t>(n);
            B->goff.get<int32_t>(n);
            B->gtot.get<int32_t>(1);
            B->soff.get<int64_t>(nbox + 1);
            B->N.get<int32_t>(nbox);
            B->coff.get<int64_t>(nbox + 1);
            B->stat.get<int32_t>(4);
            B->spts.get<double>(cs * d);
            B->sbox.get<int32_t>(cs);
            B->orig_h.get<int32_t>(cs);
            B->cell.get<int32_t>(cs);
            B->counts.get<int32_t>(ccap);
            B->starts.get<int32_t>(ccap);
            B->fill.get<int32_t>(ccap);
            B->big.get<int32_t>(ccap + 1);
            B->tmp.get<int32_t>(pg::scan_tmp_elems(
                std::max<int64_t>({ccap, cap, B->n_h, pg::lk_bucket_blocks(cap) * pg::LK_NCLASS})));
            B->total.get<int32_t>(1);
            B->perm.get<uint32_t>(cs);
            B->sorted.get<double>(cs * d);
            B->sorted_box.get<int32_t>(cs);
            B->knn.get<uint32_t>(cs * kmax);
            B->lists.get<uint32_t>(3 * cs);
            B->ncount.get<int64_t>(3);
            B->flags.get<int32_t>(cs);
            B->offs.get<int32_t>(cs);
            B->nbr.get<int32_t>(cs * MAXP);
            B->area.get<double>(cs * MAXP);
            B->nrm.get<double>(3 * cs * MAXP);
            B->cen.get<double>(3 * cs * MAXP);
            B->count.get<uint8_t>(cs);
            B->vol.get<double>(cs);
            B->ccen.get<double>(3 * cs);
            B->status.get<int32_t>(cs);
            B->kcert.get<int32_t>(cs);
            B->ticket.get<unsigned long long>(1);
            if (B->prm.mesh.restart && pg::clip_resume()) {
                B->cstate.get<uint8_t>(cs * pg::CLIP_STATE_BYTES);
                B->lastnb.get<uint32_t>(cs);
                B->resume.get<int32_t>(cs);
            }
            B->hid.get<int32_t>(cs * MAXP);
            B->area_u.get<double>(cs * MAXP);
            B->vol_u.get<double>(cs);
            B->o2s.get<int32_t>(n);
            B->hsorted.get<int32_t>(cs);
            B->hcode.get<int32_t>(cs);
            B->nbr_cm.get<int32_t>(cs * MAXP);
            B->twin.get<uint8_t>(cs * MAXP);
            B->solve.get<uint64_t>(cs);
            B->frec.get<pg::CellRec>(cs);
            B->fbuf.get<double>((size_t)(mp.dim + 2) * cs * MAXP);
            B->tot.get<double>(5 * (size_t)nbox);
            B->feat.get<float>(cs * pg::LK_NFEAT);
            B->kcls.get<int32_t>(cs);
            B->klists.get<uint32_t>(cs * pg::LK_NCLASS);
            B->kcnt.get<int64_t>(pg::LK_NCLASS);
            // clip order of learned steps: last certified k per hydro cell
            B->ovals.get<uint32_t>(cs);
            pg::launch_fill_i32(B->kc_prev.get<int32_t>(n), n, 0, s);
            HIPCHK(hipEventCreate(&B->e0));
            HIPCHK(hipEventCreate(&B->e1));
            B->lanes = pg::clip_lanes();
            B->slots = pg::clip_slots();
            batch_rebuild(B, s);
            HIPCHK(hipGetLastError());
            batch_check(c, B);
            pg::launch_q_from_prim(mp.dim, B->view(), B->o2s.as<int32_t>(), B->vol_u.as<double>(), s);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(s));
        } catch (...) {
            delete B;
            throw;
        }
        *out = B;
    });
}

int pg_batch_step(pg_ctx* c, pg_batch* B, double* ms) {
    return guarded([&] {
        use(c);
        usable(B, "batch");
        fail_guard(B, [&] { batch_step(c, B); });
        if (ms) *ms = B->last_ms;
    });
}

int pg_batch_set_kmodel(pg_batch* B, const pg_kmodel* km) {
    return guarded([&] {
        argchk(B != nullptr, "null batch");
        if (B->km != km && B->exec) {
            (void)hipGraphExecDestroy(B->exec);
            B->exec = nullptr;
        }
        B->km = km;
    });
}

int pg_batch_set_tuning(pg_batch* B, const pg_batch_tuning* t) {
    return guarded([&] {
        argchk(B && t, "null argument");
        argchk(pg::valid_clip_config(t->clip_lanes, t->clip_slots), "clip_lanes must be 8 or 16, clip_slots 48 or 64");
        argchk(t->knn_wave_min_k >= 1 && t->blocks_per_cu >= 1 && t->blocks_per_cu <= 64, "bad tuning");
        const bool changed = B->lanes != t->clip_lanes || B->slots != t->clip_slots || B->wave_min_k != t->knn_wave_min_k ||
                             B->bpc != t->blocks_per_cu || B->use_graph != (t->use_graph != 0);
        B->lanes = t->clip_lanes;
        B->slots = t->clip_slots;
        B->wave_min_k = t->knn_wave_min_k;
        B->bpc = t->blocks_per_cu;
        B->use_graph = t->use_graph != 0;
        if (changed && B->exec) {
            (void)hipGraphExecDestroy(B->exec);
            B->exec = nullptr;
        }
    });
}

int pg_batch_get_tuning(const pg_batch* B, pg_batch_tuning* t) {
    return guarded([&] {
        argchk(B && t, "null argument");
        t->clip_lanes = B->lanes;
        t->clip_slots = B->slots;
        t->knn_wave_min_k = B->wave_min_k;
        t->blocks_per_cu = B->bpc;
        t->use_graph = B->use_graph ? 1 : 0;
    });
}

int pg_batch_autotune(pg_ctx* c, pg_batch* B, int reps, pg_batch_tuning* best, double* best_ms) {
    return guarded([&] {
        use(c);
        usable(B, "batch");
        argchk(reps >= 1 && reps <= 16, "bad argument");
        fail_guard(B, [&] {
        hipStream_t s = c->stream;
        batch_copy_state(B, true, s);
        const pg_batch_tuning start{B->lanes, B->slots, B->wave_min_k, B->bpc, B->use_graph ? 1 : 0};
        pg_batch_tuning winner = start;
        double wms = 1e300;
        const int shape_c[4][2] = {{16, 64}, {16, 48}, {8, 64}, {8, 48}};
        const int wave_c[2] = {33, 1};
        const int bpc_c[2] = {8, 64};
        for (int a = 0; a < 4; a++)
            for (int w = 0; w < 2; w++)
                for (int q = 0; q < 2; q++) {
                    const pg_batch_tuning t{shape_c[a][0], shape_c[a][1], wave_c[w], bpc_c[q], 1};
                    if (pg_batch_set_tuning(B, &t) != PG_OK) throw Error{PG_ERR_ARG, g_err};
                    double tot = 0;
                    for (int r = 0; r < reps; r++) {
                        batch_copy_state(B, false, s);
                        batch_rebuild(B, s);  // the mesh of the saved state
                        batch_step(c, B);
                        tot += B->last_ms;
                    }
                    if (tot / reps < wms) {
                        wms = tot / reps;
                        winner = t;
                    }
                }
        if (pg_batch_set_tuning(B, &winner) != PG_OK) throw Error{PG_ERR_ARG, g_err};
        batch_copy_state(B, false, s);
        batch_rebuild(B, s);
        HIPCHK(hipGetLastError());
        batch_check(c, B);
        if (best) *best = winner;
        if (best_ms) *best_ms = wms;
        });
    });
}

int pg_batch_get(pg_ctx* c, pg_batch* B, double* pts, double* rho, double* vel, double* E) {
    return guarded([&] {
        use(c);
        usable(B, "batch");
        hipStream_t s = c->stream;
        const size_t n = (size_t)B->n_h, d = (size_t)B->dim;
        pg::launch_prim_from_q(B->dim, B->view(), B->o2s.as<int32_t>(), B->vol_u.as<double>(), s);
        HIPCHK(hipGetLastError());
        if (pts) HIPCHK(hipMemcpyAsync(pts, B->x.p, sizeof(double) * n * d, hipMemcpyDeviceToHost, s));
        if (rho) HIPCHK(hipMemcpyAsync(rho, B->rho.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        if (vel) HIPCHK(hipMemcpyAsync(vel, B->vel.p, sizeof(double) * n * d, hipMemcpyDeviceToHost, s));
        if (E) HIPCHK(hipMemcpyAsync(E, B->E.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    });
}

int pg_batch_dt(pg_ctx* c, const pg_batch* B, double* dt) {
    return guarded([&] {
        use(c);
        usable(B, "batch");
        argchk(dt != nullptr, "null argument");
        HIPCHK(hipMemcpyAsync(dt, B->dt_box.p, sizeof(double) * B->nbox, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int pg_batch_totals(pg_ctx* c, pg_batch* B, double* tot) {
    return guarded([&] {
        use(c);
        usable(B, "batch");
        argchk(tot != nullptr, "null argument");
        pg::launch_btotals(B->dim, B->nbox, B->hoff.as<int64_t>(), B->view(), B->tot.as<double>(), c->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(tot, B->tot.p, sizeof(double) * 5 * B->nbox, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    });
}

int pg_batch_info(pg_ctx* c, const pg_batch* B, pg_batch_info_t* info) {
    return guarded([&] {
        use(c);
        argchk(B && info, "null argument");
        int64_t ns = 0;
        HIPCHK(hipMemcpyAsync(&ns, B->soff.as<int64_t>() + B->nbox, sizeof(int64_t), hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        info->nbox = B->nbox;
        info->n_hydro = B->n_h;
        info->n_seeds = ns;
        info->seed_capacity = B->cap;
        info->last_step_ms = B->last_ms;
    });
}

int pg_batch_free(pg_batch** B) {
    return guarded([&] {
        if (!B || !*B) return;
        delete *B;
        *B = nullptr;
    });
}

}  // extern "C"
