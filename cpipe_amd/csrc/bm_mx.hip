// Mixed-precision block matching for gfx950: the patch-distance cross term on
// MFMA, an exact FP32 re-rank of the shortlisted candidates.
//
// Same workgroup tiling as the exact kernel (16 ref columns x 14 ref rows,
// search region staged in LDS); each wave takes 4 x 4-ref sub-blocks. For the
// 16 refs of a sub-block the union of their windows (48 x 48 positions) is
// scanned as blocks of 16 candidate positions (one candidate row, 16 x). With the tile centred on
// its mean mu and rounded to FP16 (or BF16), one 16x16 block of cross terms
// <c - mu, r - mu> is 4 MFMAs (v_mfma_f32_16x16x16, K = 64 pixels):
//   A[i = candidate][k = pixel]  lane i + 16 (k/4): 4 consecutive pixels of a
//                                candidate row, read from the FP32 tile in LDS,
//                                centred and rounded on the fly
//   B[k = pixel][j = ref]        lane j + 16 (k/4): the refs, loaded once per row
//   D[i][j]                      lane j + 16 (i/4), element i % 4
// so lane (ref j, quarter q) receives matrix rows 4q .. 4q+3 of every block
// (the candidates whose x + y = q mod 4, see below) and ranks them by ||c - mu||^2 - 2 <c - mu, r - mu> (= SSD - ||r - mu||^2 of the
// rounded tile; the norms come from the same rounded values, summed in FP32).
// Each lane keeps its best 12 (a quarter of the window), candidates below its
// admission bound waiting in a short LDS queue as in the exact kernel. The
// 4 x 12 shortlisted positions of a ref are then re-ranked with the exact FP32
// SSD (from the FP32 tile) in the exact kernel's operation order (row sums as fma chains, rows
// added in order); the 15 smallest by (SSD, scan order) with SSD < tau follow
// the ref. Whenever the exact top 15 are among the shortlist (the rounded
// distances perturb ranks by far less than the gaps a quarter-window top-12
// leaves), the lists equal the exact kernel's.
#include <hip/hip_runtime.h>

#include "bm3d_shared.h"
#include "cpipe_bm3d.h"

namespace cpipe_internal {
namespace {

constexpr int TX = 16, TY = 14;                         // refs per workgroup
constexpr int LW = 104;                                 // f16 tile width (6 blocks of 16 + 8)
constexpr int LH = (TY + 1) * STRIDE + 2 * R + 3 + 1;   // 86 rows
constexpr int NB = 6;                                   // candidate blocks of 16 per row (96 positions)
constexpr int NW = 88;                                  // stored norms per row (positions >= 84 never match)
constexpr int NH = (TY - 1) * STRIDE + 2 * R + 1;       // 78 candidate rows
constexpr int KL = 12;                                  // shortlist per lane
constexpr int QP = 8;                                   // queue slots per lane

typedef float f4 __attribute__((ext_vector_type(4)));

template <int PREC>
struct Bx;

template <>
struct Bx<CPIPE_BM3D_PRECISION_FP16> {
    typedef _Float16 vec __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ uint16_t enc(float x) {
        return __builtin_bit_cast(uint16_t, (_Float16)x);
    }
    static __device__ __forceinline__ float dec(uint16_t b) {
        return (float)__builtin_bit_cast(_Float16, b);
    }
    static __device__ __forceinline__ f4 mma(vec a, vec b, f4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
    }
};

template <>
struct Bx<CPIPE_BM3D_PRECISION_BF16> {
    typedef short vec __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ uint16_t enc(float x) {  // round to nearest even
        uint32_t u = __float_as_uint(x);
        u += 0x7fffu + ((u >> 16) & 1u);
        return (uint16_t)(u >> 16);
    }
    static __device__ __forceinline__ float dec(uint16_t b) {
        return __uint_as_float((uint32_t)b << 16);
    }
    static __device__ __forceinline__ f4 mma(vec a, vec b, f4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    }
};

__host__ __device__ inline int rpos(int i, int n) {
    const int last = n - K;
    const int p = i * STRIDE;
    return p < last ? p : last;
}

// 4 consecutive tile values starting at column c of row y, centred and
// rounded, as an operand.
template <int PREC>
__device__ __forceinline__ typename Bx<PREC>::vec tile4(const float* tf, int y, int c, float mu) {
    typedef typename Bx<PREC>::vec V;
    const float* p = tf + y * LW + c;
    const uint64_t u = (uint64_t)Bx<PREC>::enc(p[0] - mu) | ((uint64_t)Bx<PREC>::enc(p[1] - mu) << 16) |
                       ((uint64_t)Bx<PREC>::enc(p[2] - mu) << 32) | ((uint64_t)Bx<PREC>::enc(p[3] - mu) << 48);
    return __builtin_bit_cast(V, u);
}

__device__ __forceinline__ void insert_k(float (&bd)[KL], uint32_t (&bp)[KL], float s, uint32_t p) {
    // bd sorted ascending; s < bd[KL - 1]
#pragma unroll
    for (int k = KL - 1; k >= 1; --k) {
        const bool sh = s < bd[k - 1];
        const bool here = s < bd[k];
        bd[k] = sh ? bd[k - 1] : (here ? s : bd[k]);
        bp[k] = sh ? bp[k - 1] : (here ? p : bp[k]);
    }
    const bool first = s < bd[0];
    bd[0] = first ? s : bd[0];
    bp[0] = first ? p : bp[0];
}

template <int PREC>
__global__ __launch_bounds__(256) void k_bm_mx(const FrameTab* __restrict__ tab,
                                               const int4* __restrict__ items, int step, float tau,
                                               int use_gate) {
    typedef Bx<PREC> X;
    typedef typename X::vec V;
    __shared__ float tf[LH * LW];                 // the search region, FP32 (0 outside the image)
    __shared__ __align__(16) float nrm[NH * NW];  // ||c - mu||^2 of the rounded values per position
    __shared__ float2 pend[QP * 256];
    __shared__ float red[8];

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

    // ---- stage the region; its mean centres the rounded values
    float sum = 0.0f;
    int nin = 0;
    {
        // every load of the thread in flight before the first LDS store
        constexpr int NST = LH * LW, IT = (NST + 255) / 256;
        float sv[IT];
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const int idx = tid + 256 * i;
            const int yy = idx / LW, xx = idx - yy * LW;
            const int gx = X0 + xx, gy = Y0 + yy;
            sv[i] = 0.0f;
            if (idx < NST && gx >= 0 && gx < w && gy >= 0 && gy < h) sv[i] = plane[(size_t)gy * w + gx];
        }
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const int idx = tid + 256 * i;
            const int yy = idx / LW, xx = idx - yy * LW;
            const int gx = X0 + xx, gy = Y0 + yy;
            if (idx < NST && gx >= 0 && gx < w && gy >= 0 && gy < h) {
                sum += sv[i];
                ++nin;
            }
            if (idx < NST) tf[idx] = sv[i];
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        sum += __shfl_xor(sum, off);
        nin += __shfl_xor(nin, off);
    }
    if ((tid & 63) == 0) {
        red[tid >> 6] = sum;
        red[4 + (tid >> 6)] = __int_as_float(nin);
    }
    __syncthreads();
    const int ntot = __float_as_int(red[4]) + __float_as_int(red[5]) + __float_as_int(red[6]) +
                     __float_as_int(red[7]);
    const float mu = ntot > 0 ? (red[0] + red[1] + red[2] + red[3]) / (float)ntot : 0.0f;

    // ---- candidate norms: 2 threads per column, sliding over 8 rows
    if (tid < 2 * NW) {
        const int lx = tid % NW, half = tid / NW;
        const int ya = half * (NH / 2), yb = half ? NH : NH / 2;
        float ring[8];
        auto rowsq = [&](int y) {
            float s = 0.0f;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const float v = X::dec(X::enc(tf[y * LW + lx + j] - mu));
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
            nrm[y * NW + lx] = s;
#pragma unroll
            for (int i = 0; i < 7; ++i) ring[i] = ring[i + 1];
        }
    }
    __syncthreads();

    const int wave = tid >> 6, lane = tid & 63;
    const int j = lane & 15, q = lane >> 4;
    const int last_in = ny - 1 - by0;
    const bool regular = !(last_in >= 0 && last_in < TY && rpos(ny - 1, h) != rpos(by0, h) + STRIDE * last_in);
    float2* qq = pend + tid;
    const float inf = __builtin_inff();

    // 16 sub-blocks of 4 x 4 refs (the last row group has 2 real rows), four per wave;
    // lane j of a sub-block is ref (column j % 4, row j / 4)
    for (int sb = wave; sb < 16; sb += 4) {
        const int gx = (sb & 3) * 4, gy = (sb >> 2) * 4;
        if (by0 + gy >= ny) break;  // uniform: row groups past the image
        const int a = gy + (j >> 2);
        const int bx = bx0 + gx + (j & 3), by = by0 + a;
        const bool real = bx < nx && by < ny && a < TY;
        const int rx = rpos(min(bx, nx - 1), w);
        const int ry = regular ? rpos(by0, h) + STRIDE * min(a, TY - 1) : rpos(min(by, ny - 1), h);
        const int lry = ry - Y0, lrx = rx - X0;
        // refs as the B operand: lane (j, q) holds pixels 16 kb + 4 q + t of ref j
        V bref[4];
#pragma unroll
        for (int kb = 0; kb < 4; ++kb)
            bref[kb] = tile4<PREC>(tf, lry + 2 * kb + (q >> 1), lrx + 4 * (q & 1), mu);

        float bd[KL];
        uint32_t bp[KL];
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            bd[k] = inf;
            bp[k] = 0u;
        }
        int np = 0;
        // admission bound: this lane's 12th best, and the largest 6th best of the
        // ref's four lanes (24 candidates are at least that good, so a worse one
        // cannot reach the ref's best 15 with any margin)
        float lim = inf;
        auto flush = [&]() {
#pragma unroll
            for (int k = 0; k < QP; ++k) {
                if (k < np) {
                    const float2 e = qq[k * 256];
                    if (e.x < bd[KL - 1]) insert_k(bd, bp, e.x, __float_as_uint(e.y));
                }
            }
            np = 0;
            float t6 = bd[5];
            t6 = fmaxf(t6, __shfl_xor(t6, 16));
            t6 = fmaxf(t6, __shfl_xor(t6, 32));
            lim = fminf(bd[KL - 1], t6);
        };
        // union of the sub-block's windows: rows of its first / last ref row, columns
        // from its first ref column (3 blocks of 16 cover 4 columns 3 px apart)
        const int ry_first = regular ? rpos(by0, h) + STRIDE * gy : rpos(by0 + gy, h);
        const int a_last = min(gy + 3, min(TY - 1, ny - 1 - by0));
        const int ry_last = regular ? rpos(by0, h) + STRIDE * a_last : rpos(by0 + a_last, h);
        const int cy0 = max(ry_first - R, 0), cy1 = min(ry_last + R, h - K);
        const int lxb = rpos(min(bx0 + gx, nx - 1), w) - R - X0;
        // this lane's window (clipped to the image); empty if not real
        const int wlo = real ? max(lrx - R, -X0) : 1, whi = real ? min(lrx + R, w - K - X0) : 0;
        const int ylo = real ? max(ry - R, 0) : 1, yhi = real ? min(ry + R, h - K) : 0;
        for (int cy = cy0; cy <= cy1; ++cy) {
            const int ly = cy - Y0;
            const bool yin = cy >= ylo && cy <= yhi;
            const int self = cy == ry ? lrx : -1;  // the ref itself is pinned, not a candidate
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const int lx0 = lxb + 16 * b;
                // A: matrix row i = 4 q' + t' holds candidate lx0 + 4 t' + ((q' - cy) & 3),
                // so lane quarter q receives the positions with (lx + cy) = q (mod 4):
                // good matches cluster along edges (same x, or same y), and the
                // diagonal classes spread them over the four shortlists
                f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                const int ca = lx0 + 4 * (j & 3) + (((j >> 2) - cy) & 3);
#pragma unroll
                for (int kb = 0; kb < 4; ++kb) {
                    const V av = tile4<PREC>(tf, ly + 2 * kb + (q >> 1), ca + 4 * (q & 1), mu);
                    acc = X::mma(av, bref[kb], acc);
                }
                if (__any(np > QP - 4)) flush();
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int lx = lx0 + 4 * t + ((q - cy) & 3);  // element t of quarter q: row 4q + t
                    const float nv = nrm[ly * NW + min(lx, NW - 1)];
                    const float sc = __builtin_fmaf(-2.0f, acc[t], nv);
                    if (yin && lx >= wlo && lx <= whi && lx != self && sc < lim) {
                        qq[np * 256] = make_float2(
                            sc, __uint_as_float(((uint32_t)cy << 16) | (uint32_t)(X0 + lx)));
                        ++np;
                    }
                }
            }
        }
        flush();

        // ---- exact FP32 re-rank of this lane's shortlist (same order as the exact kernel)
        float es[KL];
        uint32_t key[KL];  // scan order: (dy + R) * 64 + (dx + R)
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            es[k] = inf;
            key[k] = 0xffffffffu;
        }
        if (real) {
            const float* rp = tf + lry * LW + lrx;
#pragma unroll
            for (int k = 0; k < KL; ++k) {
                if (bd[k] < inf) {
                    const int qx = (int)(bp[k] & 0xffffu), qy = (int)(bp[k] >> 16);
                    const float* c = tf + (qy - Y0) * LW + (qx - X0);
                    float ssd = 0.0f;
#pragma unroll 1
                    for (int i = 0; i < 8; ++i) {
                        const float* ri = rp + i * LW;
                        const float* ci = c + i * LW;
                        float e = ri[0] - ci[0];
                        float r = e * e;
#pragma unroll
                        for (int jj = 1; jj < 8; ++jj) {
                            e = ri[jj] - ci[jj];
                            r = __builtin_fmaf(e, e, r);
                        }
                        ssd = i == 0 ? r : ssd + r;
                    }
                    if (ssd < tau) {
                        es[k] = ssd;
                        key[k] = (uint32_t)((qy - ry + R) * 64 + (qx - rx + R));
                    }
                }
            }
        }
        // rank each entry among the ref's 4 x 12 (lanes j, j+16, j+32, j+48) by
        // (ssd, scan order) ascending
        int rank[KL];
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            rank[k] = 0;
#pragma unroll
            for (int m = 0; m < KL; ++m)
                rank[k] += (m != k && (es[m] < es[k] || (es[m] == es[k] && key[m] < key[k]))) ? 1 : 0;
        }
#pragma unroll 1
        for (int o = 1; o < 4; ++o) {
            const int src = (lane + 16 * o) & 63;
#pragma unroll 1
            for (int m = 0; m < KL; ++m) {
                const float os = __shfl(es[m], src);
                const uint32_t ok = (uint32_t)__shfl((int)key[m], src);
#pragma unroll
                for (int k = 0; k < KL; ++k)
                    rank[k] += (os < es[k] || (os == es[k] && ok < key[k])) ? 1 : 0;
            }
        }
        int nvalid = 0;
#pragma unroll
        for (int k = 0; k < KL; ++k) nvalid += es[k] < inf ? 1 : 0;
        nvalid += __shfl_xor(nvalid, 16);
        nvalid += __shfl_xor(nvalid, 32);
        const int n = 1 + min(NMAX - 1, nvalid);
        if (real) {
            const size_t r = (size_t)by * nx + bx;
            uint32_t* o = pos + r * NMAX;
#pragma unroll
            for (int k = 0; k < KL; ++k)
                if (es[k] < inf && rank[k] < NMAX - 1) o[1 + rank[k]] = bp[k];
            if (q == 0) {
                o[0] = ((uint32_t)ry << 16) | (uint32_t)rx;
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
        k_bm_mx<CPIPE_BM3D_PRECISION_BF16><<<n_items, 256, 0, st>>>(tab, items, step, tau, use_gate);
    else
        k_bm_mx<CPIPE_BM3D_PRECISION_FP16><<<n_items, 256, 0, st>>>(tab, items, step, tau, use_gate);
    return hipGetLastError();
}

}  // namespace cpipe_internal
