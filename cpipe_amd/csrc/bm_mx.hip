// Mixed-precision block matching for gfx950: the patch-distance cross term on
// MFMA, an exact FP32 re-rank of the shortlisted candidates.
//
// Same workgroup tiling as the exact kernel (16 ref columns x 14 ref rows),
// 16 waves: wave w takes the 4 x 4-ref sub-block at ref column 4 (w mod 4),
// ref row 4 (w / 4) (the last sub-block row has 2 real rows). Phases of a
// workgroup:
//  1. stage the search region in LDS, in FP32 (`tf`) and centred on its mean
//     mu and rounded to FP16 (or BF16) (`th`), and the squared norm
//     ||c - mu||^2 of every candidate patch of the rounded tile (`nrm`);
//  2. scan: for each sub-block, the union of its 16 refs' windows (48 x 48
//     positions) in column groups of 4 (x = xa .. xa+3, xa a multiple of 4 in
//     tile coordinates) by 16 rows. With A[i][k] the rounded tile over the
//     12 x 8 pixels (columns xa .. xa+11, rows cy_i .. cy_i+7) and B_s[k][j]
//     ref j placed at column offset s (zero elsewhere), D_s = A B_s holds the
//     cross terms <c - mu, r_j - mu> of the 16 candidates (xa + s, cy_i) with
//     the 16 refs: 3 MFMAs v_mfma_f32_16x16x32 (K = 96 = 8 rows x 12 columns;
//     2 for s = 0, whose last 4 columns are zero), A read with aligned 8-byte
//     LDS loads and shared by the 4 offsets. Operand slots (lane group g,
//     element e): K-block 0 = patch row g, column e; K-block 1 = row 4 + g;
//     K-block 2 = rows 2g (e < 4) and 2g + 1 (e >= 4), column 8 + (e mod 4);
//  3. each lane (ref j, quarter q) holds D rows 4q .. 4q+3, i.e. candidates
//     at rows cyb + 4t + ((q + xa/4) mod 4): quarter q collects one row class
//     mod 4, rotated by the column group, so candidates along a row, a column
//     or a diagonal spread over the 4 quarters. It ranks them by
//     ||c - mu||^2 - 2 <c - mu, r - mu> (= SSD - ||r - mu||^2 of the rounded
//     tile) and keeps its best 12 in a sorted register list; the position
//     (row, column within the window union) rides in the low 12 bits of the
//     FP32 score, so an insertion is one v_med3 per entry. Candidates below
//     the lane's admission bound wait in a short LDS queue;
//  4. the 4 x 12 shortlisted positions of each ref are re-ranked with the
//     exact FP32 SSD in the exact kernel's operation order (chunks of 3, 3
//     and 2 columns per row and rows per patch); the 15 smallest by (SSD, scan order) with
//     SSD < tau follow the ref. Whenever the exact top 15 are among the
//     shortlist (the rounded distances perturb ranks by far less than the gaps
//     a quarter-window top-12 leaves), the lists equal the exact kernel's.
#include <hip/hip_runtime.h>

#include "bm3d_shared.h"
#include "cpipe_bm3d.h"

namespace cpipe_internal {
namespace {

constexpr int TX = 16, TY = 14;                         // refs per workgroup
constexpr int LW = 104;                                 // FP32 tile row (px)
constexpr int LWH = 108;                                // rounded tile row (px): 54 dwords, so the
                                                        // 16 rows one A load touches hit distinct banks
constexpr int LH = (TY + 1) * STRIDE + 2 * R + 3 + 1;   // 87 rows
constexpr int NW = 88;                                  // norm columns (candidate x < 84)
constexpr int NH = (TY - 1) * STRIDE + 2 * R + 1;       // 78 candidate rows
constexpr int KL = 12;                                  // shortlist per lane
constexpr int QP = 16;                                  // queue slots per lane
constexpr int NT = 1024;                                // threads per workgroup (4 waves per SIMD)
// A row block of 16 candidate rows may run past the union's last row (up to
// tile row 83, patch rows to 90): those rows of the rounded tile and of the
// norms are padding (zero), read and never kept.
constexpr int LHP = LH + 4;
constexpr int NHP = NH + 6;

// Members in this order keep every vector access naturally aligned (16-byte
// norm rows, 8-byte rounded-tile groups); an unaligned LDS access stalls.
struct BmLds {
    float nrm[NHP * NW];       // ||c - mu||^2 of the rounded tile per candidate position
    float tf[LH * LW];         // the FP32 tile for the re-rank (0 outside the image)
    uint16_t th[LHP * LWH];    // rounded, centred tile
    uint32_t pend[QP * NT];    // admission queue (packed scores)
};
static_assert(sizeof(float) * NHP * NW % 16 == 0 && sizeof(float) * LH * LW % 16 == 0 &&
                  NW % 4 == 0 && LWH % 4 == 0,
              "aligned norm rows and rounded-tile groups");
static_assert(sizeof(BmLds) + 64 <= 160 * 1024, "one workgroup per CU");

typedef float f4 __attribute__((ext_vector_type(4)));

template <int PREC>
struct Bx;

template <>
struct Bx<CPIPE_BM3D_PRECISION_FP16> {
    typedef _Float16 vec8 __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ uint16_t enc(float x) {
        return __builtin_bit_cast(uint16_t, (_Float16)x);
    }
    static __device__ __forceinline__ float dec(uint16_t b) {
        return (float)__builtin_bit_cast(_Float16, b);
    }
    static __device__ __forceinline__ f4 mma(uint4 a, uint4 b, f4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(vec8, a), __builtin_bit_cast(vec8, b),
                                                      c, 0, 0, 0);
    }
};

template <>
struct Bx<CPIPE_BM3D_PRECISION_BF16> {
    typedef __bf16 vec8 __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ uint16_t enc(float x) {  // round to nearest even (v_cvt_pk_bf16_f32)
        return __builtin_bit_cast(uint16_t, (__bf16)x);
    }
    static __device__ __forceinline__ float dec(uint16_t b) {
        return __uint_as_float((uint32_t)b << 16);
    }
    static __device__ __forceinline__ f4 mma(uint4 a, uint4 b, f4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(vec8, a), __builtin_bit_cast(vec8, b),
                                                       c, 0, 0, 0);
    }
};

__host__ __device__ inline int rpos(int i, int n) {
    const int last = n - K;
    const int p = i * STRIDE;
    return p < last ? p : last;
}

__device__ __forceinline__ uint32_t pk(uint32_t lo, uint32_t hi) { return lo | (hi << 16); }

// B operand, K-blocks 0 / 1: element e = ref pixel (row, e - S), 0 for e < S.
template <int S>
__device__ __forceinline__ uint4 b_head(const uint16_t (&r)[8]) {
    uint32_t v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = e >= S ? r[e - S] : 0u;
    return make_uint4(pk(v[0], v[1]), pk(v[2], v[3]), pk(v[4], v[5]), pk(v[6], v[7]));
}

// B operand, K-block 2: element e < 4 = ref pixel (row a, 8 + e - S), element
// 4 + e = (row b, 8 + e - S); 0 where that column is past the patch.
template <int S>
__device__ __forceinline__ uint4 b_tail(const uint16_t (&ra)[8], const uint16_t (&rb)[8]) {
    uint32_t v[8];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        v[e] = e < S ? ra[8 + e - S] : 0u;
        v[4 + e] = e < S ? rb[8 + e - S] : 0u;
    }
    return make_uint4(pk(v[0], v[1]), pk(v[2], v[3]), pk(v[4], v[5]), pk(v[6], v[7]));
}

// Sorted ascending list; inserting x keeps the KL smallest (x = +inf: no change).
__device__ __forceinline__ void insert_k(float (&bd)[KL], float x) {
#pragma unroll
    for (int k = KL - 1; k >= 1; --k) bd[k] = __builtin_amdgcn_fmed3f(bd[k - 1], x, bd[k]);
    bd[0] = __builtin_amdgcn_fmed3f(-__builtin_inff(), x, bd[0]);
}

// Geometry of the sub-block at ref column gx, ref row gy of the workgroup, for ref j.
struct SubBlock {
    bool real;
    int bx, by, rx, ry, lrx, lry;
    int cy0, cy1;     // union rows (absolute, inclusive)
    int xa0, xend;    // first column group / last union column (tile coordinates)
};

__device__ __forceinline__ SubBlock sub_block(int gx, int gy, int j, int bx0, int by0, int X0, int Y0,
                                              int w, int h, int nx, int ny, bool regular) {
    SubBlock b;
    const int a = gy + (j >> 2);
    b.bx = bx0 + gx + (j & 3);
    b.by = by0 + a;
    b.real = b.bx < nx && b.by < ny && a < TY;
    b.rx = rpos(min(b.bx, nx - 1), w);
    b.ry = regular ? rpos(by0, h) + STRIDE * min(a, TY - 1) : rpos(min(b.by, ny - 1), h);
    b.lrx = b.rx - X0;
    b.lry = b.ry - Y0;
    const int ry_first = regular ? rpos(by0, h) + STRIDE * gy : rpos(by0 + gy, h);
    const int a_last = min(gy + 3, min(TY - 1, ny - 1 - by0));
    const int ry_last = regular ? rpos(by0, h) + STRIDE * a_last : rpos(by0 + a_last, h);
    b.cy0 = max(ry_first - R, 0);
    b.cy1 = min(ry_last + R, h - K);
    b.xa0 = (rpos(min(bx0 + gx, nx - 1), w) - R - X0) & ~3;
    b.xend = min(rpos(min(bx0 + gx + 3, nx - 1), w) + R, w - K) - X0;
    return b;
}

template <int PREC>
__global__ __launch_bounds__(NT) void k_bm_mx(const FrameTab* __restrict__ tab,
                                               const int4* __restrict__ items, int step, float tau,
                                               int use_gate) {
    typedef Bx<PREC> X;
    __shared__ __align__(16) BmLds L;
    __shared__ float red[2 * (NT / 64)];

    const int4 it = items[blockIdx.x];
    const FrameTab& F = tab[it.x];
    if (use_gate && *F.gate == 0) return;
    const float* __restrict__ plane = step == 1 ? F.z[0] : F.bz[0];
    const int w = F.w, h = F.h, nx = F.nx, ny = F.ny;
    uint32_t* __restrict__ pos = F.pos[step - 1];
    uint8_t* __restrict__ cnt = F.cnt[step - 1];
    const int bx0 = it.y * TX, by0 = it.z * TY;
    const int X0 = rpos(bx0, w) - R;
    const int Y0 = rpos(by0, h) - R;
    const int tid = threadIdx.x;
    constexpr int NST = LH * LW, IT = (NST + NT - 1) / NT;

    // ---- 1. stage the region; its mean centres the rounded values
    {
        float sv[IT];
        float sum = 0.0f;
        int nin = 0;
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const int idx = tid + NT * i;
            const int yy = idx / LW, xx = idx - yy * LW;
            const int gxx = X0 + xx, gyy = Y0 + yy;
            sv[i] = 0.0f;
            if (idx < NST && gxx >= 0 && gxx < w && gyy >= 0 && gyy < h) sv[i] = plane[(size_t)gyy * w + gxx];
        }
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const int idx = tid + NT * i;
            const int yy = idx / LW, xx = idx - yy * LW;
            const int gxx = X0 + xx, gyy = Y0 + yy;
            if (idx < NST && gxx >= 0 && gxx < w && gyy >= 0 && gyy < h) {
                sum += sv[i];
                ++nin;
            }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            sum += __shfl_xor(sum, off);
            nin += __shfl_xor(nin, off);
        }
        if ((tid & 63) == 0) {
            red[tid >> 6] = sum;
            red[NT / 64 + (tid >> 6)] = __int_as_float(nin);
        }
        __syncthreads();
        float tot = 0.0f;
        int ntot = 0;
#pragma unroll
        for (int i = 0; i < NT / 64; ++i) {
            tot += red[i];
            ntot += __float_as_int(red[NT / 64 + i]);
        }
        const float mu = ntot > 0 ? tot / (float)ntot : 0.0f;
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const int idx = tid + NT * i;
            const int yy = idx / LW, xx = idx - yy * LW;
            if (idx < NST) {
                L.tf[idx] = sv[i];
                L.th[yy * LWH + xx] = X::enc(sv[i] - mu);
            }
        }
        for (int i = tid; i < (LHP - LH) * LWH; i += NT) L.th[LH * LWH + i] = 0;
        for (int i = tid; i < (NHP - NH) * NW; i += NT) L.nrm[NH * NW + i] = 0.0f;
    }
    __syncthreads();

    // ---- candidate norms: 4 threads per column, sliding over 8 rows
    if (tid < 4 * NW) {
        const int lx = tid % NW, part = tid / NW;
        const int ya = part * (NH / 4 + 1), yb = min(ya + NH / 4 + 1, NH);
        float ring[8];
        auto rowsq = [&](int y) {
            float s = 0.0f;
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) {
                const float v = X::dec(L.th[y * LWH + lx + jj]);
                s = __builtin_fmaf(v, v, s);
            }
            return s;
        };
#pragma unroll
        for (int i = 0; i < 7; ++i) ring[i] = rowsq(ya + i);
        for (int y = ya; y < yb; ++y) {
            ring[7] = rowsq(y + 7);
            float s = 0.0f;
#pragma unroll
            for (int i = 0; i < 8; ++i) s += ring[i];
            L.nrm[y * NW + lx] = s;
#pragma unroll
            for (int i = 0; i < 7; ++i) ring[i] = ring[i + 1];
        }
    }
    __syncthreads();

    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int j = lane & 15, q = lane >> 4;  // D: ref j, rows 4q .. 4q+3; A/B: matrix row (ref) j, K group q
    const int last_in = ny - 1 - by0;
    const bool regular = !(last_in >= 0 && last_in < TY && rpos(ny - 1, h) != rpos(by0, h) + STRIDE * last_in);
    const float inf = __builtin_inff();
    const int gx = 4 * (wave & 3);

    // ---- 2./3. scan a sub-block, then 4. re-rank its shortlists
    const int gy = 4 * (wave >> 2);
    if (by0 + gy < ny) {  // uniform: row groups past the image have no refs
        float bd[KL];  // this lane's shortlist, ascending packed scores
        {  // scan
            const SubBlock b = sub_block(gx, gy, j, bx0, by0, X0, Y0, w, h, nx, ny, regular);
            // ref j as B operands (K group q), for column offsets 0..3
            uint4 B0[4], B1[4], B2[4];
            {
                uint16_t r0[8], r1[8], r2[8], r3[8];
                const uint16_t* rp = L.th + b.lry * LWH + b.lrx;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    r0[e] = rp[q * LWH + e];
                    r1[e] = rp[(4 + q) * LWH + e];
                    r2[e] = rp[(2 * q) * LWH + e];
                    r3[e] = rp[(2 * q + 1) * LWH + e];
                }
                B0[0] = b_head<0>(r0);
                B0[1] = b_head<1>(r0);
                B0[2] = b_head<2>(r0);
                B0[3] = b_head<3>(r0);
                B1[0] = b_head<0>(r1);
                B1[1] = b_head<1>(r1);
                B1[2] = b_head<2>(r1);
                B1[3] = b_head<3>(r1);
                B2[0] = make_uint4(0u, 0u, 0u, 0u);
                B2[1] = b_tail<1>(r2, r3);
                B2[2] = b_tail<2>(r2, r3);
                B2[3] = b_tail<3>(r2, r3);
            }
            // this lane's window (tile x, absolute y); empty if not real. The ref
            // itself is scanned too (it is the best match, and is dropped at the re-rank).
            const int wlo = b.real ? max(b.lrx - R, -X0) : 1, whi = b.real ? min(b.lrx + R, w - K - X0) : 0;
            const int ylo = b.real ? max(b.ry - R, 0) : 1, yhi = b.real ? min(b.ry + R, h - K) : 0;
#pragma unroll
            for (int k = 0; k < KL; ++k) bd[k] = inf;
            // admission bound: this lane's 12th best, and the largest 6th best of the
            // ref's four lanes (24 candidates are at least that good, so a worse one
            // cannot reach the ref's best 15 with any margin)
            float lim = inf;
            uint32_t* const pq = L.pend + tid;  // this lane's queue, slot k at pq[k * NT]
            uint32_t* qp = pq;                  // next free slot
            uint32_t* const qfull = pq + (QP - 4) * NT;
            auto flush = [&]() {
                const int np = (int)(qp - pq) / NT;
#pragma unroll
                for (int k0 = 0; k0 < QP; k0 += 8) {
                    if (!__any(k0 < np)) break;
                    float e[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) e[k] = __uint_as_float(pq[(k0 + k) * NT]);  // in flight together
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        if (!__any(k0 + k < np)) break;
                        insert_k(bd, k0 + k < np ? e[k] : inf);
                    }
                }
                qp = pq;
                float t6 = bd[5];
                t6 = fmaxf(t6, __shfl_xor(t6, 16));
                t6 = fmaxf(t6, __shfl_xor(t6, 32));
                lim = fminf(bd[KL - 1], t6);
            };
            // row blocks: the one holding the refs' own rows first (near candidates
            // tighten the admission bound early), then the ones above and below
            for (int bi = 0; bi < 3; ++bi) {
                const int cyb = b.cy0 + 16 * (bi == 0 ? 1 : (bi == 1 ? 0 : 2));
                if (cyb > b.cy1) continue;
                // column groups in 4 passes of equal rotation (xa / 4 mod 4), so every
                // row offset below is fixed within a pass
#pragma unroll
                for (int rot = 0; rot < 4; ++rot) {
                    const int xa_first = b.xa0 + 4 * ((rot - (b.xa0 >> 2)) & 3);
                    // A: matrix row j = 4 q' + t' is candidate row cyb + 4 t' + ((q' + rot) & 3)
                    const int rowa = cyb + 4 * (j & 3) + (((j >> 2) + rot) & 3) - Y0;
                    const uint16_t* ta0 = L.th + (rowa + q) * LWH;
                    const uint16_t* ta1 = L.th + (rowa + 4 + q) * LWH;
                    const uint16_t* ta2 = L.th + (rowa + 2 * q) * LWH + 8;
                    // this lane's D rows: element t is candidate row cyb + 4 t + rq
                    const int rq = (q + rot) & 3;
                    const float* tn = L.nrm + (cyb + rq - Y0) * NW;
                    bool yok[4];
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        const int cy = cyb + 4 * t + rq;
                        yok[t] = cy >= ylo && cy <= yhi;
                    }
                    // position code: (row - cy0) << 6 | (column - xa0), < 48 x 64
                    const uint32_t prow = (uint32_t)((cyb - b.cy0 + rq) << 6) - (uint32_t)b.xa0;
                    for (int xa = xa_first; xa <= b.xend; xa += 16) {
                        const uint2 a00 = *reinterpret_cast<const uint2*>(ta0 + xa);
                        const uint2 a01 = *reinterpret_cast<const uint2*>(ta0 + xa + 4);
                        const uint2 a10 = *reinterpret_cast<const uint2*>(ta1 + xa);
                        const uint2 a11 = *reinterpret_cast<const uint2*>(ta1 + xa + 4);
                        const uint2 a20 = *reinterpret_cast<const uint2*>(ta2 + xa);
                        const uint2 a21 = *reinterpret_cast<const uint2*>(ta2 + LWH + xa);
                        const uint4 A0 = make_uint4(a00.x, a00.y, a01.x, a01.y);
                        const uint4 A1 = make_uint4(a10.x, a10.y, a11.x, a11.y);
                        const uint4 A2 = make_uint4(a20.x, a20.y, a21.x, a21.y);
                        f4 nv[4];
#pragma unroll
                        for (int t = 0; t < 4; ++t) nv[t] = *reinterpret_cast<const f4*>(tn + 4 * t * NW + xa);
                        f4 acc[4];
#pragma unroll
                        for (int s = 0; s < 4; ++s) {
                            const f4 z = {0.0f, 0.0f, 0.0f, 0.0f};
                            acc[s] = X::mma(A0, B0[s], z);
                            acc[s] = X::mma(A1, B1[s], acc[s]);
                            if (s > 0) acc[s] = X::mma(A2, B2[s], acc[s]);
                        }
                        const uint32_t pb = prow + (uint32_t)xa;
#pragma unroll
                        for (int s = 0; s < 4; ++s) {
                            if (__any(qp > qfull)) flush();
                            const bool xok = xa + s >= wlo && xa + s <= whi;
#pragma unroll
                            for (int t = 0; t < 4; ++t) {
                                const float nvv = (xok && yok[t]) ? nv[t][s] : inf;
                                const float sc = __builtin_fmaf(-2.0f, acc[s][t], nvv);
                                if (sc < lim) {
                                    *qp = (__float_as_uint(sc) & ~0xfffu) | (pb + (uint32_t)(((4 * t) << 6) + s));
                                    qp += NT;
                                }
                            }
                        }
                    }
                }
            }
            flush();
        }

        // ---- 4. exact FP32 re-rank (chunks of 3, 3 and 2 columns and rows).
        // The sub-block geometry is recomputed here rather than kept live through
        // the scan (it would be spilled to scratch).
        int jr = j;
        asm volatile("" : "+v"(jr));
        const SubBlock b = sub_block(gx, gy, jr, bx0, by0, X0, Y0, w, h, nx, ny, regular);
        uint32_t bp[KL];  // (y << 16) | x
        float es[KL], gs[KL];
        int cb[KL];       // FP32-tile offset of a live entry; 0 (one shared address) otherwise
        uint32_t lm = 0;  // live entries, bit k
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            const uint32_t rel = __float_as_uint(bd[k]) & 0xfffu;
            const int qy = b.cy0 + (int)(rel >> 6), qx = X0 + b.xa0 + (int)(rel & 63u);
            bp[k] = ((uint32_t)qy << 16) | (uint32_t)qx;
            es[k] = 0.0f;
            const bool live = b.real && bd[k] < inf && !(qx == b.rx && qy == b.ry);
            lm |= live ? 1u << k : 0u;
            cb[k] = live ? (qy - Y0) * LW + (qx - X0) : 0;
        }
        // rows in groups of 3, 3 and 2: gs = the current group's sum, es = the
        // finished groups' (one select per row and candidate; the group
        // hand-over is one uniform branch per row)
#pragma unroll 1
        for (int i = 0; i < 8; ++i) {
            if (i == 3) {
#pragma unroll
                for (int k = 0; k < KL; ++k) es[k] = gs[k];
            } else if (i == 6) {
#pragma unroll
                for (int k = 0; k < KL; ++k) es[k] = es[k] + gs[k];
            }
            const bool first = i == 0 || i == 3 || i == 6;
            const float* ri = L.tf + (b.lry + i) * LW + b.lrx;
            float rr[8];
#pragma unroll
            for (int jj = 0; jj < 8; ++jj) rr[jj] = ri[jj];
            // three candidate rows in flight at a time
#pragma unroll
            for (int k0 = 0; k0 < KL; k0 += 3) {
                float c[3][8];
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) {
                    const float* ci = L.tf + cb[k0 + kk] + i * LW;
#pragma unroll
                    for (int jj = 0; jj < 8; ++jj) c[kk][jj] = ci[jj];
                }
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) {
                    float e[8];
#pragma unroll
                    for (int jj = 0; jj < 8; ++jj) e[jj] = rr[jj] - c[kk][jj];
                    const float r = ssd_row(e);
                    gs[k0 + kk] = first ? r : gs[k0 + kk] + r;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < KL; ++k) es[k] = es[k] + gs[k];
        // order key (ssd bits, scan order (dy + R) * 64 + (dx + R)); the ref
        // itself, empty slots and ssd >= tau drop out
        uint64_t key[KL];
        const uint64_t none = ~0ull;
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            const int qx = (int)(bp[k] & 0xffffu), qy = (int)(bp[k] >> 16);
            const bool ok = ((lm >> k) & 1u) && es[k] < tau;
            key[k] = ok ? ((uint64_t)__float_as_uint(es[k]) << 32) | (uint64_t)((qy - b.ry + R) * 64 + (qx - b.rx + R))
                        : none;
        }
        // rank each entry among the ref's 4 x 12 (lanes j, j+16, j+32, j+48)
        int rank[KL];
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            rank[k] = 0;
#pragma unroll
            for (int m = 0; m < KL; ++m)
                if (m != k) rank[k] += key[m] < key[k] ? 1 : 0;
        }
#pragma unroll 1
        for (int o = 1; o < 4; ++o) {
            const int src = (lane + 16 * o) & 63;
            uint32_t lo[KL], hi[KL];  // the other lane's keys, all fetches in flight
#pragma unroll
            for (int m = 0; m < KL; ++m) {
                lo[m] = (uint32_t)__shfl((int)(uint32_t)key[m], src);
                hi[m] = (uint32_t)__shfl((int)(uint32_t)(key[m] >> 32), src);
            }
#pragma unroll
            for (int m = 0; m < KL; ++m) {
                const uint64_t ok = ((uint64_t)hi[m] << 32) | lo[m];
#pragma unroll
                for (int k = 0; k < KL; ++k) rank[k] += ok < key[k] ? 1 : 0;
            }
        }
        int nvalid = 0;
#pragma unroll
        for (int k = 0; k < KL; ++k) nvalid += key[k] != none ? 1 : 0;
        nvalid += __shfl_xor(nvalid, 16);
        nvalid += __shfl_xor(nvalid, 32);
        const int n = 1 + min(NMAX - 1, nvalid);
        if (b.real) {
            const size_t r = (size_t)b.by * nx + b.bx;
            uint32_t* o = pos + r * NMAX;
#pragma unroll
            for (int k = 0; k < KL; ++k)
                if (key[k] != none && rank[k] < NMAX - 1) o[1 + rank[k]] = bp[k];
            if (q == 0) {
                o[0] = ((uint32_t)b.ry << 16) | (uint32_t)b.rx;
                for (int k = n; k < NMAX; ++k) o[k] = 0u;
                cnt[r] = (uint8_t)n;
            }
        }
    }
}

}  // namespace

int bm_mx_launch(int precision, const FrameTab* tab, const int4* items, int n_items, int step,
                 float tau, int use_gate, hipStream_t st) {
    if (precision == CPIPE_BM3D_PRECISION_BF16)
        k_bm_mx<CPIPE_BM3D_PRECISION_BF16><<<n_items, NT, 0, st>>>(tab, items, step, tau, use_gate);
    else
        k_bm_mx<CPIPE_BM3D_PRECISION_FP16><<<n_items, NT, 0, st>>>(tab, items, step, tau, use_gate);
    return hipGetLastError();
}

}  // namespace cpipe_internal
