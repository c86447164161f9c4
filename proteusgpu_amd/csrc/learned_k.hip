// Learned neighbour count: per-seed features, a 16-64-64-6 classifier on
// MFMA (bfloat16 operands, float32 accumulation), its training step and the
// per-seed choice of k.  The reference has no counterpart (its neighbour
// count is the compile-time _K_, knn.h:20); the certificate still judges
// every cell, so a wrong prediction costs a restart, never a wrong cell.
//
// Matrix layout (v_mfma_f32_16x16x16_bf16, lane l = 16 q + r):
//   A (16x16, row x k): lane holds A[r][4q + j], j = 0..3
//   B (16x16, k x col): lane holds B[4q + j][r]
//   D (16x16, row x col): lane holds D[4q + j][r]
// Each layer is computed transposed, Z^T = W^T X^T, with the 16 seeds of a
// group as the columns: the D tile of one layer (rows = 16 outputs, column =
// seed) is exactly the B operand of the next layer's k-block, so activations
// never leave registers in the forward pass.  The weight gradients contract
// over seeds and go through a 16x16 transpose in LDS.
#include <algorithm>

#include "common.h"
#include "kernels.h"
#include "subgrid.h"

namespace pg {
namespace {

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int NF = LK_NFEAT, NH = 64, NC = LK_NCLASS;
constexpr int OW1 = 0, OB1 = OW1 + NF * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NH * NC;
static_assert(OB3 + NC == LK_NPARAM, "parameter layout");

__constant__ int c_buckets[NC] = {16, 24, 32, 48, 64, 96};
// representative certified k of each bucket's interval (midpoints; 80 for
// the open last bucket)
__constant__ float c_krep[NC] = {12.f, 20.f, 28.f, 40.f, 56.f, 80.f};

// Cost, in KNN row entries, of a seed with certified k = K started at row
// length k0: a pass costs k (row) + r m (clip steps; m = K once certified,
// else k); a rejected pass repeats at 2k (at most 256).
PG_DEV float chain_cost(float k0, float K, float r) {
    float k = k0, cost = 0.f;
    for (int stage = 0; stage < 6; stage++) {
        if (K <= k) return cost + k + r * K;
        cost += k + r * k;
        k = fminf(2.f * k, 256.f);
    }
    return cost;
}

PG_DEV short f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (short)0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

PG_DEV f4 mfma(bf4 a, bf4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

PG_DEV f4 zero4() { return f4{0.f, 0.f, 0.f, 0.f}; }

// ------------------------------------------------------------ features
// Features of sorted seed s of a box of n seeds with grid N whose cells are
// cbase + i + j N + k N^2; see the feature list in kernels.h.
template <int DIM>
PG_DEV void kfeatures_one(const double* __restrict__ sorted, int64_t s, int64_t n, int N, int64_t cbase,
                          const int32_t* __restrict__ cnt, const int32_t* __restrict__ start, float (&f)[16],
                          SubGrid sg = {}) {
    double p[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; d++) p[d] = sorted[(size_t)DIM * s + d];
    const int c = grid_cell<DIM>(p, N);
    const int ci = c % N, cj = (c / N) % N, ck = DIM == 3 ? c / (N * N) : 0;
    const double ncell = (double)N * N * (DIM == 3 ? N : 1);
    const double lam = (double)n / ncell;
    double best[8];  // the 8 smallest squared distances, ascending (+inf: none)
#pragma unroll
    for (int j = 0; j < 8; j++) best[j] = __builtin_inf();
    // keep d2 if it is among the 8 smallest so far: insertion without
    // branches, new best[j] = max(best[j-1], min(best[j], d2)), top down
    auto keep = [&](double d2) {
#pragma unroll
        for (int j = 7; j > 0; j--) best[j] = fmax(best[j - 1], fmin(best[j], d2));
        best[0] = fmin(best[0], d2);
    };
    auto keep_point = [&](int64_t q) {
        double pq[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < DIM; d++) pq[d] = sorted[(size_t)DIM * q + d];
        keep(dist2<DIM>(p, pq));
    };
    uint32_t crowded_cells = 0;  // bit: cube cell (di, dj, dk) has a sub-grid
    double csum[4] = {0, 0, 0, 0};
    int cin[4] = {0, 0, 0, 0};
    int inside = 0;
    double total = 0, gx = 0, gy = 0, gz = 0;
    int nmin = 0x7fffffff, nmax = -1;
    const int kr = DIM == 3 ? 1 : 0;
    for (int dk = -kr; dk <= kr; dk++)
        for (int dj = -1; dj <= 1; dj++)
            for (int di = -1; di <= 1; di++) {
                const int ii = ci + di, jj = cj + dj, kk = ck + dk;
                if (ii < 0 || jj < 0 || kk < 0 || ii >= N || jj >= N || kk >= N) continue;
                const int64_t cc = cbase + ii + jj * N + (int64_t)kk * N * N;
                const int m = (di != 0) + (dj != 0) + (dk != 0);
                const int ct = cnt[cc];
                inside++;
                csum[m] += ct;
                cin[m]++;
                total += ct;
                gx += (double)ct * di;
                gy += (double)ct * dj;
                gz += (double)ct * dk;
                if (m > 0) {
                    nmin = min(nmin, ct);
                    nmax = max(nmax, ct);
                }
                if (sg.m != nullptr && sg.m[cc] > 0) {
                    // a crowded cell: its sub-cells below, pruned by the 8th
                    // distance (the same 8 smallest values as a full scan)
                    crowded_cells |= 1u << ((dk + 1) * 9 + (dj + 1) * 3 + (di + 1));
                    continue;
                }
                const int32_t b = start[cc];
                for (int32_t q = b; q < b + ct; q++)
                    if (q != s) keep_point(q);
            }
    // crowded cells, the own one first: the sub-cells around p, then those
    // of each cell whose lower bound is below the current 8th distance (a
    // sub-cell whose bound is at least it holds no smaller distance)
    if (crowded_cells != 0u) {
        double P[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < DIM; d++) P[d] = p[d] * (double)N;
        const int own = 13;  // (0, 0, 0) in the cube's bit numbering (2D: dk = 0)
        for (int pass = 0; pass < 2; pass++) {
            uint32_t todo = crowded_cells;
            while (todo != 0u) {
                const int bit = (pass == 0 && ((todo >> own) & 1u)) ? own : __builtin_ctz(todo);
                todo &= ~(1u << bit);
                if (pass == 0 && bit != own) continue;
                const int dk = DIM == 3 ? bit / 9 - 1 : 0, dj = (bit / 3) % 3 - 1, di = bit % 3 - 1;
                const int ix[3] = {ci + di, cj + dj, ck + dk};
                const int64_t cc = cbase + ix[0] + ix[1] * N + (int64_t)ix[2] * N * N;
                const int m = sg.m[cc];
                const int64_t base = sg.ptr[cc];
                // pass 0: the own cell's sub-cells next to p (a first 8th
                // distance); pass 1: every crowded cell, pruned
                int lo[3], hi[3];
                sub_box<DIM>(P, ix, m, N, pass == 0 ? 0.0 : best[7], lo, hi);
                for (int c3 = lo[2]; c3 <= hi[2]; c3++)
                    for (int b3 = lo[1]; b3 <= hi[1]; b3++)
                        for (int a3 = lo[0]; a3 <= hi[0]; a3++) {
                            const int sx[3] = {a3, b3, c3};
                            const bool near = pass == 0;
                            if (!near && sub_lower_d2<DIM>(P, ix, sx, m, N) >= best[7]) continue;
                            if (pass == 1 && bit == own) {
                                // already taken in pass 0: the sub-cells of
                                // the pass-0 box
                                int l0[3], h0[3];
                                sub_box<DIM>(P, ix, m, N, 0.0, l0, h0);
                                if (a3 >= l0[0] && a3 <= h0[0] && b3 >= l0[1] && b3 <= h0[1] && c3 >= l0[2] &&
                                    c3 <= h0[2])
                                    continue;
                            }
                            const int64_t j = base + a3 + (int64_t)m * (b3 + (int64_t)m * c3);
                            const int32_t b = sg.start[j], e = b + sg.cnt[j];
                            for (int32_t t = b; t < e; t++) {
                                const uint32_t q = sg.idx[t];
                                if ((int64_t)q != s) keep_point(q);
                            }
                        }
            }
        }
    }
    // points of the cube other than the seed (a full scan's count)
    const int nb = (int)total - 1;
    const int cube = DIM == 3 ? 27 : 9;
#pragma unroll
    for (int m = 0; m < 4; m++) f[m] = cin[m] ? (float)log2(1.0 + csum[m] / cin[m] / lam) : 0.0f;
    f[4] = (float)((double)inside / cube);
    const int cidx[3] = {ci, cj, ck};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (d < DIM) {
            const double u = p[d] * (double)N - (double)cidx[d];
            f[5 + d] = (float)fabs(2.0 * u - 1.0);
        } else {
            f[5 + d] = 0.0f;
        }
    }
    const double r4 = nb >= 4 ? sqrt(best[3]) * N : 2.0;
    const double r8 = nb >= 8 ? sqrt(best[7]) * N : 2.0;
    const double r1 = nb >= 1 ? sqrt(best[0]) * N : 2.0;
    f[8] = (float)r4;
    f[9] = (float)r8;
    f[10] = (float)r1;
    f[11] = (float)log2(1.0 + total / (inside * lam));
    double wall = 1e300;
#pragma unroll
    for (int d = 0; d < DIM; d++) wall = fmin(wall, fmin(p[d], 1.0 - p[d]));
    f[12] = (float)(fmin(fmax(wall, 0.0) * N, 3.0) / 3.0);
    f[13] = nmax >= 0 ? (float)log2(1.0 + (double)(nmax - nmin) / lam) : 0.0f;
    f[14] = (float)log2(1.0 + sqrt(gx * gx + gy * gy + gz * gz) / (cube * lam));
    f[15] = (float)(r8 / r4);
}

PG_DEV void store_features(float* out, int64_t o, const float (&f)[16]) {
    float4* dst = (float4*)(out + 16 * o);
#pragma unroll
    for (int v = 0; v < 4; v++) dst[v] = make_float4(f[4 * v], f[4 * v + 1], f[4 * v + 2], f[4 * v + 3]);
}

// One thread per seed (sorted space).
template <int DIM>
__global__ void __launch_bounds__(256)
    k_kfeatures(const double* __restrict__ sorted, int64_t n, int N, const int32_t* __restrict__ cnt,
                const int32_t* __restrict__ start, const uint32_t* __restrict__ out_index,
                float* __restrict__ out, SubGrid sg) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float f[16];
    kfeatures_one<DIM>(sorted, s, n, N, 0, cnt, start, f, sg);
    store_features(out, out_index ? (int64_t)out_index[s] : s, f);
}

// Batched boxes: the box of each sorted seed gives its grid and seed count.
template <int DIM>
__global__ void __launch_bounds__(256)
    k_bkfeatures(const double* __restrict__ sorted, const int64_t* __restrict__ n_dev, const int32_t* __restrict__ sbox,
                 const int64_t* __restrict__ soff, const int32_t* __restrict__ N, const int64_t* __restrict__ coff,
                 const int32_t* __restrict__ cnt, const int32_t* __restrict__ start, float* __restrict__ out) {
    const int64_t n = *n_dev;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.x * blockDim.x) {
        const int b = sbox[s];
        float f[16];
        kfeatures_one<DIM>(sorted, s, soff[b + 1] - soff[b], N[b], coff[b], cnt, start, f);
        store_features(out, s, f);
    }
}

// ------------------------------------------------------------ weights
struct FwdWeights {
    bf4 a1[4], a2[4][4], a3[4];
    float b1[4][4], b2[4][4], b3[4];
};

PG_DEV void load_fwd(const float* __restrict__ P, int q, int r, FwdWeights& w) {
#pragma unroll
    for (int t = 0; t < 4; t++) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            w.a1[t][j] = f2bf(P[OW1 + (4 * q + j) * NH + 16 * t + r]);
            w.b1[t][j] = P[OB1 + 16 * t + 4 * q + j];
            w.b2[t][j] = P[OB2 + 16 * t + 4 * q + j];
        }
#pragma unroll
        for (int kb = 0; kb < 4; kb++)
#pragma unroll
            for (int j = 0; j < 4; j++) w.a2[t][kb][j] = f2bf(P[OW2 + (16 * kb + 4 * q + j) * NH + 16 * t + r]);
    }
#pragma unroll
    for (int kb = 0; kb < 4; kb++)
#pragma unroll
        for (int j = 0; j < 4; j++) w.a3[kb][j] = r < NC ? f2bf(P[OW3 + (16 * kb + 4 * q + j) * NC + r]) : (short)0;
#pragma unroll
    for (int j = 0; j < 4; j++) w.b3[j] = 4 * q + j < NC ? P[OB3 + 4 * q + j] : 0.f;
}

struct FwdOut {
    f4 z1[4], z2[4], z3;
    bf4 x, h1[4], h2[4];
};

// Forward pass of one group of 16 seeds; xb holds features 4q..4q+3 of the
// group's seed r.
PG_DEV void forward16(const FwdWeights& w, bf4 xb, FwdOut& o) {
    o.x = xb;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        f4 z = mfma(w.a1[t], xb, zero4());
#pragma unroll
        for (int j = 0; j < 4; j++) {
            z[j] += w.b1[t][j];
            o.h1[t][j] = f2bf(fmaxf(z[j], 0.f));
        }
        o.z1[t] = z;
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        f4 z = zero4();
#pragma unroll
        for (int kb = 0; kb < 4; kb++) z = mfma(w.a2[t][kb], o.h1[kb], z);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            z[j] += w.b2[t][j];
            o.h2[t][j] = f2bf(fmaxf(z[j], 0.f));
        }
        o.z2[t] = z;
    }
    f4 z = zero4();
#pragma unroll
    for (int kb = 0; kb < 4; kb++) z = mfma(w.a3[kb], o.h2[kb], z);
#pragma unroll
    for (int j = 0; j < 4; j++) z[j] += w.b3[j];
    o.z3 = z;
}

PG_DEV bf4 load_x(const float* __restrict__ X, int64_t row, bool valid, int q) {
    bf4 xb;
    if (valid) {
        const float4 v = *(const float4*)(X + 16 * row + 4 * q);
        xb[0] = f2bf(v.x);
        xb[1] = f2bf(v.y);
        xb[2] = f2bf(v.z);
        xb[3] = f2bf(v.w);
    } else {
        xb = bf4{0, 0, 0, 0};
    }
    return xb;
}

// Softmax over the 6 logits of seed r: classes 0-3 sit in lane r (q = 0),
// classes 4-5 in lane r + 16 (q = 1, j < 2).  Returns p in the same layout
// (0 on padding entries).
PG_DEV f4 softmax16(f4 z, int q) {
    float m = -INFINITY;
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (4 * q + j < NC) m = fmaxf(m, z[j]);
    m = fmaxf(m, __shfl_xor(m, 16, 64));
    f4 e;
    float sum = 0.f;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        e[j] = 4 * q + j < NC ? expf(z[j] - m) : 0.f;
        sum += e[j];
    }
    sum += __shfl_xor(sum, 16, 64);
#pragma unroll
    for (int j = 0; j < 4; j++) e[j] = 4 * q + j < NC ? e[j] / sum : 0.f;
    return e;
}

// ------------------------------------------------------------ inference
// grid-stride over groups of 16 seeds, one wave per group.  cls[s] = the
// bucket with the least expected restart-chain cost (chain_cost, r = a0)
// under the softmax of the logits.
__global__ void __launch_bounds__(256)
    k_kpredict(const float* __restrict__ X, int64_t n_host, const int64_t* __restrict__ n_dev,
               const float* __restrict__ P, float a0, const uint32_t* __restrict__ out_index,
               int32_t* __restrict__ cls, float* __restrict__ logits) {
    const int64_t n = n_dev ? *n_dev : n_host;
    const int l = threadIdx.x & 63, q = l >> 4, r = l & 15;
    FwdWeights w;
    load_fwd(P, q, r, w);
    const int64_t groups = (n + 15) / 16;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwave = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t g = wave; g < groups; g += nwave) {
        const int64_t row = g * 16 + r;
        const bool valid = row < n;
        FwdOut o;
        forward16(w, load_x(X, row, valid, q), o);
        const f4 p = softmax16(o.z3, q);
        const float p4 = __shfl(p[0], (l + 16) & 63, 64), p5 = __shfl(p[1], (l + 16) & 63, 64);
        if (valid && logits && q < 2) {
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (4 * q + j < NC) logits[NC * row + 4 * q + j] = o.z3[j];
        }
        if (valid && q == 0) {
            const float pc[NC] = {p[0], p[1], p[2], p[3], p4, p5};
            float best = INFINITY;
            int bi = 0;
#pragma unroll
            for (int b = 0; b < NC; b++) {
                float cost = 0.f;
#pragma unroll
                for (int c = 0; c < NC; c++) cost += pc[c] * chain_cost((float)c_buckets[b], c_krep[c], a0);
                if (cost < best) {
                    best = cost;
                    bi = b;
                }
            }
            const int64_t dst = out_index ? (int64_t)out_index[row] : row;
            cls[dst] = bi;
        }
    }
}

// ------------------------------------------------------------ training
struct BwdWeights {
    bf4 a3b[4], a2b[4][4];
};

PG_DEV void load_bwd(const float* __restrict__ P, int q, int r, BwdWeights& w) {
#pragma unroll
    for (int t = 0; t < 4; t++) {
#pragma unroll
        for (int j = 0; j < 4; j++) w.a3b[t][j] = 4 * q + j < NC ? f2bf(P[OW3 + (16 * t + r) * NC + 4 * q + j]) : (short)0;
#pragma unroll
        for (int kb = 0; kb < 4; kb++)
#pragma unroll
            for (int j = 0; j < 4; j++) w.a2b[t][kb][j] = f2bf(P[OW2 + (16 * t + r) * NH + 16 * kb + 4 * q + j]);
    }
}

PG_DEV uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// LDS tile [16 seeds][16 features] of bf16, written from the D layout and
// read transposed (lane holds entries of seeds 4q..4q+3 for feature r).
PG_DEV void tile_put(short* T, int q, int r, bf4 v) {
    *(bf4*)(T + r * 16 + 4 * q) = v;
}
PG_DEV bf4 tile_getT(const short* T, int q, int r) {
    bf4 v;
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] = T[(4 * q + j) * 16 + r];
    return v;
}

constexpr int GROUPS_PER_BLOCK = 4;  // 64 seeds per training block

// One wave per block.  Rows: batch entry i of this step is row
// mix64(seed + step * B + i) mod n when seed != 0, else row i (i < n).
// part[blk][LK_NPARAM] and part_loss[blk] receive this block's sums.
__global__ void __launch_bounds__(64)
    k_kgrad(const float* __restrict__ X, const int32_t* __restrict__ y, int64_t n, int64_t B,
            uint64_t seed, int64_t step, const float* __restrict__ P, float* __restrict__ part,
            float* __restrict__ part_loss) {
    __shared__ short T[18][256];
    const int l = threadIdx.x, q = l >> 4, r = l & 15;
    FwdWeights w;
    load_fwd(P, q, r, w);
    BwdWeights wb;
    load_bwd(P, q, r, wb);
    f4 acc1[4], acc2[4][4], acc3[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        acc1[t] = zero4();
        acc3[t] = zero4();
#pragma unroll
        for (int u = 0; u < 4; u++) acc2[t][u] = zero4();
    }
    f4 db1[4], db2[4], db3 = zero4();
#pragma unroll
    for (int t = 0; t < 4; t++) {
        db1[t] = zero4();
        db2[t] = zero4();
    }
    float loss = 0.f;
    const float invB = 1.0f / (float)B;
    for (int gi = 0; gi < GROUPS_PER_BLOCK; gi++) {
        const int64_t i = ((int64_t)blockIdx.x * GROUPS_PER_BLOCK + gi) * 16 + r;
        const bool valid = i < B;
        int64_t row = 0;
        if (valid) row = seed ? (int64_t)(mix64(seed + (uint64_t)(step * B + i)) % (uint64_t)n) : i;
        FwdOut o;
        forward16(w, load_x(X, row, valid, q), o);
        const f4 p = softmax16(o.z3, q);
        const int yr = valid ? y[row] : -1;
        f4 g3;
        bf4 g3b;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int c = 4 * q + j;
            float g = 0.f;
            if (valid && c < NC) {
                g = (p[j] - (c == yr ? 1.f : 0.f)) * invB;
                if (c == yr) loss -= logf(p[j]);
            }
            g3[j] = g;
            g3b[j] = f2bf(g);
        }
        db3 += g3;
        // hidden gradients
        f4 g2[4];
        bf4 g2b[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const f4 d = mfma(wb.a3b[t], g3b, zero4());
#pragma unroll
            for (int j = 0; j < 4; j++) {
                g2[t][j] = o.z2[t][j] > 0.f ? d[j] : 0.f;
                g2b[t][j] = f2bf(g2[t][j]);
            }
            db2[t] += g2[t];
        }
        bf4 g1b[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            f4 d = zero4();
#pragma unroll
            for (int kb = 0; kb < 4; kb++) d = mfma(wb.a2b[t][kb], g2b[kb], d);
            f4 g1;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                g1[j] = o.z1[t][j] > 0.f ? d[j] : 0.f;
                g1b[t][j] = f2bf(g1[j]);
            }
            db1[t] += g1;
        }
        // weight gradients: contract over the 16 seeds through LDS
        __syncthreads();
        tile_put(T[0], q, r, o.x);
        tile_put(T[1], q, r, g3b);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            tile_put(T[2 + t], q, r, o.h1[t]);
            tile_put(T[6 + t], q, r, o.h2[t]);
            tile_put(T[10 + t], q, r, g2b[t]);
            tile_put(T[14 + t], q, r, g1b[t]);
        }
        __syncthreads();
        const bf4 xT = tile_getT(T[0], q, r), g3T = tile_getT(T[1], q, r);
        bf4 g1T[4], g2T[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            g1T[t] = tile_getT(T[14 + t], q, r);
            g2T[t] = tile_getT(T[10 + t], q, r);
        }
#pragma unroll
        for (int to = 0; to < 4; to++) acc1[to] = mfma(xT, g1T[to], acc1[to]);
#pragma unroll
        for (int ti = 0; ti < 4; ti++) {
            const bf4 h1T = tile_getT(T[2 + ti], q, r), h2T = tile_getT(T[6 + ti], q, r);
#pragma unroll
            for (int to = 0; to < 4; to++) acc2[ti][to] = mfma(h1T, g2T[to], acc2[ti][to]);
            acc3[ti] = mfma(h2T, g3T, acc3[ti]);
        }
    }
    float* dst = part + (size_t)blockIdx.x * LK_NPARAM;
    // D layout: dW[i = 16 ti + 4q + j][o = 16 to + r]
#pragma unroll
    for (int to = 0; to < 4; to++)
#pragma unroll
        for (int j = 0; j < 4; j++) dst[OW1 + (4 * q + j) * NH + 16 * to + r] = acc1[to][j];
#pragma unroll
    for (int ti = 0; ti < 4; ti++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
#pragma unroll
            for (int to = 0; to < 4; to++) dst[OW2 + (16 * ti + 4 * q + j) * NH + 16 * to + r] = acc2[ti][to][j];
            if (r < NC) dst[OW3 + (16 * ti + 4 * q + j) * NC + r] = acc3[ti][j];
        }
    // bias gradients: sum the 16 seed lanes of each q
#pragma unroll
    for (int j = 0; j < 4; j++) {
#pragma unroll
        for (int t = 0; t < 4; t++) {
            float a = db1[t][j], b = db2[t][j];
#pragma unroll
            for (int m = 1; m < 16; m <<= 1) {
                a += __shfl_xor(a, m, 64);
                b += __shfl_xor(b, m, 64);
            }
            if (r == 0) {
                dst[OB1 + 16 * t + 4 * q + j] = a;
                dst[OB2 + 16 * t + 4 * q + j] = b;
            }
        }
        float c = db3[j];
#pragma unroll
        for (int m = 1; m < 16; m <<= 1) c += __shfl_xor(c, m, 64);
        if (r == 0 && 4 * q + j < NC) dst[OB3 + 4 * q + j] = c;
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) loss += __shfl_xor(loss, m, 64);
    if (l == 0) part_loss[blockIdx.x] = loss * invB;
}

// Sum the block partials in block order (deterministic) and, when lr > 0,
// take one Adam step (t = step + 1) on the float32 parameters.
__global__ void __launch_bounds__(256)
    k_kadam(const float* __restrict__ part, const float* __restrict__ part_loss, int nblk, float* P,
            float* m, float* v, float* grad_out, float* loss_out, int64_t step, float lr) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0 && loss_out) {
        float L = 0.f;
        for (int b = 0; b < nblk; b++) L += part_loss[b];
        *loss_out = L;
    }
    if (p >= LK_NPARAM) return;
    float g = 0.f;
    for (int b = 0; b < nblk; b++) g += part[(size_t)b * LK_NPARAM + p];
    if (grad_out) grad_out[p] = g;
    if (lr <= 0.f) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mm = b1 * m[p] + (1.f - b1) * g;
    const float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float t = (float)(step + 1);
    const float mh = mm / (1.f - powf(b1, t)), vh = vv / (1.f - powf(b2, t));
    P[p] -= lr * mh / (sqrtf(vh) + eps);
}

// class of a seed from its minimal certified k (1 + certifying neighbour)
__global__ void k_klabels(const int32_t* __restrict__ kcert, int64_t n, int32_t* __restrict__ y) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int kc = kcert[s];
    int c = NC - 1;
    if (kc > 0) {
        for (int b = 0; b < NC; b++)
            if (c_buckets[b] >= kc) {
                c = b;
                break;
            }
    }
    y[s] = c;
}

// ------------------------------------------------------------ bucketing
// Stable split of seeds 0..n-1 by class into per-class lists (class c at
// lists + c * list_cap, ascending seed id).  Pass 1: per block of 256 seeds
// the count of each class, stored class-major (cnt[c * nblk + blk]); an
// exclusive scan of that array; pass 2: each seed's rank among the seeds of
// its class before it.  n from n_dev when given (capacity-shaped launch).
constexpr int BTB = 256;

PG_DEV int64_t seed_count(const int64_t* n_dev, int64_t n) { return n_dev ? *n_dev : n; }

__global__ void __launch_bounds__(BTB)
    k_bucket_hist(const int32_t* __restrict__ cls, const uint32_t* __restrict__ order, const int64_t* __restrict__ n_dev,
                  int64_t n_host, int nblk, int32_t* __restrict__ cnt) {
    __shared__ int32_t wc[BTB / 64][NC];
    const int64_t n = seed_count(n_dev, n_host);
    const int64_t i = (int64_t)blockIdx.x * BTB + threadIdx.x;
    const int c = i < n ? cls[order ? order[i] : (uint32_t)i] : -1;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < NC; k++) {
        const uint64_t m = (uint64_t)__ballot(c == k);
        if (l == 0) wc[w][k] = __popcll(m);
    }
    __syncthreads();
    if (threadIdx.x < NC) {
        int t = 0;
#pragma unroll
        for (int ww = 0; ww < BTB / 64; ww++) t += wc[ww][threadIdx.x];
        cnt[threadIdx.x * nblk + blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(BTB)
    k_bucket_scatter(const int32_t* __restrict__ cls, const uint32_t* __restrict__ order,
                     const int64_t* __restrict__ n_dev, int64_t n_host, int nblk, const int32_t* __restrict__ off,
                     int64_t list_cap, uint32_t* __restrict__ lists) {
    __shared__ int32_t wc[BTB / 64][NC];
    const int64_t n = seed_count(n_dev, n_host);
    const int64_t i = (int64_t)blockIdx.x * BTB + threadIdx.x;
    const uint32_t seed = order && i < n ? order[i] : (uint32_t)i;
    const int c = i < n ? cls[seed] : -1;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    int rank = 0;
#pragma unroll
    for (int k = 0; k < NC; k++) {
        const uint64_t m = (uint64_t)__ballot(c == k);
        if (l == 0) wc[w][k] = __popcll(m);
        if (c == k) rank = __popcll(m & ((1ull << l) - 1ull));
    }
    __syncthreads();
    if (c < 0) return;
    for (int ww = 0; ww < w; ww++) rank += wc[ww][c];
    const int64_t pos = (int64_t)off[c * nblk + blockIdx.x] - off[c * nblk] + rank;
    lists[(int64_t)c * list_cap + pos] = seed;
}

// counts[c] = seeds of class c (int64, device)
__global__ void k_bucket_counts(const int32_t* __restrict__ off, const int32_t* __restrict__ total, int nblk,
                                int64_t* __restrict__ counts) {
    const int c = threadIdx.x;
    if (c >= NC) return;
    const int64_t a = off[c * nblk];
    const int64_t b = c + 1 < NC ? (int64_t)off[(c + 1) * nblk] : (int64_t)*total;
    counts[c] = b - a;
}

}  // namespace

int64_t lk_bucket_blocks(int64_t n_cap) { return (n_cap + BTB - 1) / BTB; }

void launch_kbucket(const int32_t* cls, const int64_t* n_dev, int64_t n_cap, int32_t* cnt, int32_t* off,
                    int32_t* tmp, int32_t* total, int64_t list_cap, uint32_t* lists, int64_t* counts,
                    hipStream_t s, const uint32_t* order) {
    const int nblk = (int)lk_bucket_blocks(n_cap);
    if (nblk <= 0) return;
    k_bucket_hist<<<nblk, BTB, 0, s>>>(cls, order, n_dev, n_cap, nblk, cnt);
    scan_exclusive_i32(cnt, off, (int64_t)NC * nblk, tmp, total, s);
    k_bucket_scatter<<<nblk, BTB, 0, s>>>(cls, order, n_dev, n_cap, nblk, off, list_cap, lists);
    k_bucket_counts<<<1, 64, 0, s>>>(off, total, nblk, counts);
}

namespace {
constexpr int NKEY = 256;  // order key: min(previous certified k, 255)

// key of sorted seed s: the previous certified k of its point (o = perm[s],
// through id_map when given; o outside [0, prev_n) keys 0)
PG_DEV int order_key(uint32_t s, const uint32_t* __restrict__ perm, const int32_t* __restrict__ id_map,
                     const int32_t* __restrict__ prev, int64_t prev_n) {
    int64_t o = perm[s];
    if (id_map) o = id_map[o];
    if (o < 0 || o >= prev_n) return 0;
    return min(max(prev[o], 0), 255);
}

constexpr int OTILE = 4096;  // list positions per tile of the clip order, at most (LDS)

// One tile of `tile` (at most OTILE) consecutive list positions per block, reordered in
// place-of-tile (into out) by key, stably: key histogram in LDS, its
// exclusive scan, then chunks of BTB entries ranked by ballots (lane order
// within a wave, wave order within the chunk, chunk order within the tile).
__global__ void __launch_bounds__(BTB)
    k_order_tiles(const uint32_t* __restrict__ list, const int64_t* __restrict__ n_dev, int64_t n_host,
                  const uint32_t* __restrict__ perm, const int32_t* __restrict__ id_map,
                  const int32_t* __restrict__ prev, int64_t prev_n, uint32_t* __restrict__ out, int tile) {
    static_assert(BTB == NKEY, "one key counter per thread");
    __shared__ uint8_t key_s[OTILE];
    __shared__ int32_t base[NKEY];
    __shared__ int32_t wc[BTB / 64][NKEY];
    __shared__ int32_t wt[BTB / 64];
    const int64_t n = seed_count(n_dev, n_host);
    const int64_t t0 = (int64_t)blockIdx.x * tile;
    if (t0 >= n) return;  // whole block
    const int len = (int)min<int64_t>(tile, n - t0);
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    base[tid] = 0;
    __syncthreads();
    for (int j = tid; j < len; j += BTB) {
        const uint32_t e = list ? list[t0 + j] : (uint32_t)(t0 + j);
        const int key = order_key(e, perm, id_map, prev, prev_n);
        key_s[j] = (uint8_t)key;
        atomicAdd(&base[key], 1);
    }
    __syncthreads();
    {  // exclusive scan of the NKEY counts, one per thread
        const int v = base[tid];
        int incl = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int u = __shfl_up(incl, d, 64);
            if (l >= d) incl += u;
        }
        if (l == 63) wt[w] = incl;
        __syncthreads();
        for (int ww = 0; ww < w; ww++) incl += wt[ww];
        base[tid] = incl - v;
        __syncthreads();
    }
    for (int c0 = 0; c0 < len; c0 += BTB) {
        const int j = c0 + tid;
        const int key = j < len ? (int)key_s[j] : -1;
        for (int t = tid; t < (BTB / 64) * NKEY; t += BTB) (&wc[0][0])[t] = 0;
        __syncthreads();
        int rank = 0;
        // peel the wave's distinct keys one at a time (todo is wave-uniform)
        uint64_t todo = (uint64_t)__ballot(key >= 0);
        while (todo) {
            const int leader = __ffsll((unsigned long long)todo) - 1;
            const int lk = __shfl(key, leader, 64);
            const uint64_t m = (uint64_t)__ballot(key == lk);
            if (key == lk) rank = __popcll(m & ((1ull << l) - 1ull));
            if (l == leader) wc[w][lk] = __popcll(m);
            todo &= ~m;
        }
        __syncthreads();
        if (key >= 0) {
            for (int ww = 0; ww < w; ww++) rank += wc[ww][key];
            out[t0 + base[key] + rank] = list ? list[t0 + j] : (uint32_t)(t0 + j);
        }
        __syncthreads();
        int add = 0;
#pragma unroll
        for (int ww = 0; ww < BTB / 64; ww++) add += wc[ww][tid];
        base[tid] += add;
        __syncthreads();
    }
}

// prev[o] = kcert[s] for the sorted seeds s whose point o (perm[s], through
// id_map when given) lies in [0, prev_n)
__global__ void k_keep_kcert(const int64_t* __restrict__ n_dev, int64_t n_host, const uint32_t* __restrict__ perm,
                             const int32_t* __restrict__ id_map, const int32_t* __restrict__ kcert, int64_t prev_n,
                             int32_t* __restrict__ prev) {
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= seed_count(n_dev, n_host)) return;
    int64_t o = perm[s];
    if (id_map) o = id_map[o];
    if (o >= 0 && o < prev_n) prev[o] = kcert[s];
}
}  // namespace

void launch_order_list(const uint32_t* list, const int64_t* n_dev, int64_t n_cap, const uint32_t* perm,
                       const int32_t* id_map, const int32_t* prev, int64_t prev_n, uint32_t* out, hipStream_t s) {
    if (n_cap <= 0) return;
    // the tile can be shortened with PGPU_LK_OTILE for tuning runs
    // (256 positions: the face records the clip writes then stay within a
    // few hundred consecutive ids per wave, which halves its write traffic
    // against 4096 at the same step time)
    static const int tile = [] {
        const char* e = getenv("PGPU_LK_OTILE");
        const int t = e ? atoi(e) : 256;
        return t < 64 ? 64 : (t > OTILE ? OTILE : t);
    }();
    const unsigned nb = (unsigned)((n_cap + tile - 1) / tile);
    k_order_tiles<<<nb, BTB, 0, s>>>(list, n_dev, n_cap, perm, id_map, prev, prev_n, out, tile);
}

void launch_keep_kcert(const int64_t* n_dev, int64_t n_cap, const uint32_t* perm, const int32_t* id_map,
                       const int32_t* kcert, int64_t prev_n, int32_t* prev, hipStream_t s) {
    if (n_cap <= 0) return;
    k_keep_kcert<<<(unsigned)((n_cap + 255) / 256), 256, 0, s>>>(n_dev, n_cap, perm, id_map, kcert, prev_n, prev);
}

int lk_bucket(int c) {
    static const int b[LK_NCLASS] = {16, 24, 32, 48, 64, 96};
    return b[c];
}

void launch_kfeatures(int dim, const double* sorted, int64_t n, int N, const int32_t* counts,
                      const int32_t* starts, const uint32_t* out_index, float* out, hipStream_t s, SubGrid sg) {
    if (n <= 0) return;
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (dim == 2) k_kfeatures<2><<<nb, 256, 0, s>>>(sorted, n, N, counts, starts, out_index, out, sg);
    else k_kfeatures<3><<<nb, 256, 0, s>>>(sorted, n, N, counts, starts, out_index, out, sg);
}

void launch_kpredict(const float* X, int64_t n, const float* params, float a0, const uint32_t* out_index,
                     int32_t* cls, float* logits, hipStream_t s, const int64_t* n_dev) {
    if (n <= 0) return;
    const int64_t groups = (n + 15) / 16;
    const int64_t blocks = std::min<int64_t>((groups + 3) / 4, 4096);
    k_kpredict<<<(unsigned)blocks, 256, 0, s>>>(X, n, n_dev, params, a0, out_index, cls, logits);
}

void launch_bkfeatures(int dim, const double* sorted, const int64_t* n_dev, int64_t cap, const int32_t* sbox,
                       const int64_t* soff, const int32_t* N, const int64_t* coff, const int32_t* counts,
                       const int32_t* starts, float* out, hipStream_t s) {
    const unsigned nb = (unsigned)std::min<int64_t>(std::max<int64_t>((cap + 255) / 256, 1), 4096);
    if (dim == 2) k_bkfeatures<2><<<nb, 256, 0, s>>>(sorted, n_dev, sbox, soff, N, coff, counts, starts, out);
    else k_bkfeatures<3><<<nb, 256, 0, s>>>(sorted, n_dev, sbox, soff, N, coff, counts, starts, out);
}

int64_t lk_grad_blocks(int64_t B) { return (B + 16 * GROUPS_PER_BLOCK - 1) / (16 * GROUPS_PER_BLOCK); }

void launch_kgrad(const float* X, const int32_t* y, int64_t n, int64_t B, uint64_t seed, int64_t step,
                  const float* params, float* part, float* part_loss, hipStream_t s) {
    k_kgrad<<<(unsigned)lk_grad_blocks(B), 64, 0, s>>>(X, y, n, B, seed, step, params, part, part_loss);
}

void launch_kadam(const float* part, const float* part_loss, int nblk, float* params, float* m, float* v,
                  float* grad_out, float* loss_out, int64_t step, float lr, hipStream_t s) {
    k_kadam<<<(LK_NPARAM + 255) / 256, 256, 0, s>>>(part, part_loss, nblk, params, m, v, grad_out, loss_out,
                                                    step, lr);
}

void launch_klabels(const int32_t* kcert, int64_t n, int32_t* y, hipStream_t s) {
    if (n <= 0) return;
    k_klabels<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(kcert, n, y);
}

}  // namespace pg
