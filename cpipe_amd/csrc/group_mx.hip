// Mixed-precision BM3D group filtering for gfx950.
//
// The 3D transform of a group (2D DCT-II of each 8x8 member, 1D orthonormal
// Haar along the stack) and its inverse run as 16x16x16 MFMA tiles with FP16
// or BF16 operands and FP32 accumulation. Shrinkage (hard threshold, Wiener),
// group weights and the tile aggregation stay FP32, and the aggregation uses
// the same fixed order as the exact kernels (groups in ref raster order via
// the LDS turn token, members in match order), so a run is reproducible.
//
// Operand layouts of v_mfma_f32_16x16x16 (f16 and bf16_1k), lane l:
//   A[i][k]: lane i + 16 (k / 4), element k % 4
//   B[k][j]: lane j + 16 (k / 4), element k % 4
//   D[i][j]: lane j + 16 (i / 4), element i % 4
// A group of up to 16 members is the 16 x 64 matrix X (member x pixel,
// pixel = 8 row + col); rows past the group size are zero.
//   forward DCT   Y  = X Kt    A = X (lane r = member, 4 consecutive pixels), B = blocks of Kt
//   forward Haar  Z  = H Y     A = H_n, B = Y (the D layout of the previous product)
//   inverse Haar  Vt = Zt H    A = Zt (Z's D layout read as an A operand), B = H_n
//   inverse DCT   Xt = Kt Vt   A = blocks of Kt, B = Vt (D layout of the previous product)
// with K = C (x) C, K[u][p] = C[u / 8][p / 8] C[u % 8][p % 8], and H_n the
// n-point multi-level orthonormal Haar matrix (the exact kernels' butterfly,
// applied to unit vectors) zero-padded to 16 x 16. Xt ends in D layout (lane
// r = member, 4 consecutive pixels of one patch row); a per-wave LDS scratch
// turns it into lane = pixel for the ordered aggregation.
#include <hip/hip_runtime.h>

#include <cmath>

#include "bm3d_shared.h"
#include "cpipe_bm3d.h"

namespace cpipe_internal {
namespace {

__constant__ float c_mx_dct[64];        // C[u][i]
__constant__ float c_mx_kaiser[64];     // Kaiser(8, beta=2) outer product
__constant__ float c_mx_haar[5][256];   // H_n for n = 1, 2, 4, 8, 16, [row][col], zero-padded

typedef float f4 __attribute__((ext_vector_type(4)));

template <int PREC>
struct Mx;

template <>
struct Mx<CPIPE_BM3D_PRECISION_FP16> {
    typedef _Float16 vec __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ vec pack(float a, float b, float c, float d) {
        vec r;
        r[0] = (_Float16)a;
        r[1] = (_Float16)b;
        r[2] = (_Float16)c;
        r[3] = (_Float16)d;
        return r;
    }
    static __device__ __forceinline__ f4 mma(vec a, vec b, f4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
    }
    typedef _Float16 vec8 __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ f4 mma32(vec8 a, vec8 b, f4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

template <>
struct Mx<CPIPE_BM3D_PRECISION_BF16> {
    typedef short vec __attribute__((ext_vector_type(4)));
    // round to nearest even: gfx950's v_cvt_pk_bf16_f32, two values per instruction
    static __device__ __forceinline__ vec pack(float a, float b, float c, float d) {
        typedef __bf16 bvec __attribute__((ext_vector_type(4)));
        const bvec r = {(__bf16)a, (__bf16)b, (__bf16)c, (__bf16)d};
        return __builtin_bit_cast(vec, r);
    }
    static __device__ __forceinline__ f4 mma(vec a, vec b, f4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    }
    typedef __bf16 vec8 __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ f4 mma32(vec8 a, vec8 b, f4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

template <int PREC>
__device__ __forceinline__ typename Mx<PREC>::vec pack4(const f4& x) {
    return Mx<PREC>::pack(x[0], x[1], x[2], x[3]);
}

// Two K = 16 operands of one lane as one K = 32 operand. The 64-long DCT
// contractions run on 16x16x32 MFMAs with their K index permuted the same way
// in A and B (lane group g, element e < 4: k = 4 g + e of the first 16-block;
// e >= 4: k = 4 g + e - 4 of the second), so each D tile of a 16x16 product
// still feeds the next product in place.
template <int PREC>
__device__ __forceinline__ typename Mx<PREC>::vec8 cat8(typename Mx<PREC>::vec a, typename Mx<PREC>::vec b) {
    return __builtin_bit_cast(typename Mx<PREC>::vec8, __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7));
}

__device__ __forceinline__ float kmat(int u, int p) {
    return c_mx_dct[(u >> 3) * 8 + (p >> 3)] * c_mx_dct[(u & 7) * 8 + (p & 7)];
}

// Orders this wave's LDS accesses before/after the point (lanes exchange data).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Issues the loads of one stack's rows for lane (member r, quarter q): rows
// 2 kb + q/2, columns 4 (q & 1) .. + 3 of the member patch at pm (zeros for
// r >= n); forward3d consumes them.
__device__ __forceinline__ void load_stack(float (&a)[16], const float* __restrict__ src, int w,
                                           uint32_t pm, bool live, int q) {
    const int qx = (int)(pm & 0xffffu), qy = (int)(pm >> 16);
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
        if (live) {
            const float* s = src + (size_t)(qy + 2 * kb + (q >> 1)) * w + qx + 4 * (q & 1);
#pragma unroll
            for (int t = 0; t < 4; ++t) a[4 * kb + t] = s[t];
        } else {
#pragma unroll
            for (int t = 0; t < 4; ++t) a[4 * kb + t] = 0.0f;
        }
    }
}

// BF16 keeps 8 mantissa bits, so rounding a stack at its own level (pixels
// near 0.5) costs more than the residual noise of a quiet frame. BF16 groups
// are therefore centred: a level mu is taken out of each stack before
// rounding, the 3D DC coefficient (the centred stack's plus 8 mu sqrt(n):
// orthonormal DCT and Haar) is carried in FP32 through the shrinkage, and the
// shrunk DC comes back after the inverse as DC / (8 sqrt(n)) on every member
// pixel. The transforms are linear, so any mu is exact; mu is the mean of the
// members' quarter-0 pixels (the first 16-lane row: four DPP adds, no LDS).
__device__ __forceinline__ float stack_level(const float (&a)[16], int n) {
    float s = 0.0f;  // rows past the group hold zeros
#pragma unroll
    for (int t = 0; t < 16; ++t) s = s + a[t];
#define CPIPE_ROW_SHR(v, ctrl) \
    __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), ctrl, 0xf, 0xf, true))
    s = s + CPIPE_ROW_SHR(s, 0x111);  // row_shr:1
    s = s + CPIPE_ROW_SHR(s, 0x112);  // row_shr:2
    s = s + CPIPE_ROW_SHR(s, 0x114);  // row_shr:4
    s = s + CPIPE_ROW_SHR(s, 0x118);  // row_shr:8: lane 15 holds the row's sum
#undef CPIPE_ROW_SHR
    const float row0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, s), 15));
    return row0 * (1.0f / 16.0f) / (float)n;  // 16 n pixels; powers of two
}

__device__ __forceinline__ void centre_stack(float (&c)[16], const float (&a)[16], float mu, bool live) {
    const float m = live ? mu : 0.0f;  // rows past the group stay zero
#pragma unroll
    for (int t = 0; t < 16; ++t) c[t] = a[t] - m;
}

// sqrt(n) and 1 / (8 sqrt(n)) for the group sizes 1, 2, 4, 8, 16 (n is wave-uniform)
__device__ __forceinline__ float root_n(int n) {
    return n == 1 ? 1.0f : n == 2 ? 1.41421356f : n == 4 ? 2.0f : n == 8 ? 2.82842712f : 4.0f;
}
__device__ __forceinline__ float inv8root_n(int n) {
    return n == 1 ? 0.125f : n == 2 ? 0.0883883476f : n == 4 ? 0.0625f : n == 8 ? 0.0441941738f : 0.03125f;
}

// Forward 3D transform of one loaded stack: z[nb] = (H_n X Kt)[:, 16 nb .. 16 nb + 15] in D layout.
template <int PREC>
__device__ __forceinline__ void forward3d(f4 (&z)[4], const float (&a)[16],
                                          const typename Mx<PREC>::vec8 (&kf)[2][4],
                                          const typename Mx<PREC>::vec& ha) {
    typedef Mx<PREC> M;
    typename M::vec xa[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) xa[kb] = M::pack(a[4 * kb], a[4 * kb + 1], a[4 * kb + 2], a[4 * kb + 3]);
    const f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    const typename M::vec8 x0 = cat8<PREC>(xa[0], xa[1]), x1 = cat8<PREC>(xa[2], xa[3]);
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        f4 y = M::mma32(x0, kf[0][nb], zero);
        y = M::mma32(x1, kf[1][nb], y);
        z[nb] = M::mma(ha, pack4<PREC>(y), zero);
    }
}

constexpr int MX_SCR = 4;   // members per LDS transposition chunk
constexpr int MX_SPP = 68;  // scratch floats per member: the chunk's four 16-byte writes hit distinct banks

// Filters one group from its loaded stacks (load_stack: xs noisy, xb basic):
// 3D transform, shrinkage, inverse; leaves lane p = pixel p of member m in v[m]
// (m < n) and returns the group's aggregation weight.
template <int STEP, int PREC>
__device__ __forceinline__ float filter_mx(float (&v)[NMAX], int n, int r, int q, int lane,
                                           const float (&xs)[16], const float (&xb)[16], float* sb,
                                           const typename Mx<PREC>::vec8 (&kf)[2][4],
                                           const typename Mx<PREC>::vec8 (&ki)[4][2], float thr,
                                           float s2) {
    typedef Mx<PREC> M;
    const float* hm = c_mx_haar[31 - __builtin_clz((unsigned)n)];
    const typename M::vec ha =
        M::pack(hm[r * 16 + 4 * q], hm[r * 16 + 4 * q + 1], hm[r * 16 + 4 * q + 2], hm[r * 16 + 4 * q + 3]);
    const typename M::vec hb = M::pack(hm[(4 * q) * 16 + r], hm[(4 * q + 1) * 16 + r],
                                       hm[(4 * q + 2) * 16 + r], hm[(4 * q + 3) * 16 + r]);
    constexpr bool CENTRE = PREC == CPIPE_BM3D_PRECISION_BF16;
    const float rootn = CENTRE ? root_n(n) : 0.0f;
    f4 z[4];
    if constexpr (CENTRE) {
        const float mu = stack_level(xs, n);
        float c[16];
        centre_stack(c, xs, mu, r < n);
        forward3d<PREC>(z, c, kf, ha);
        if (lane == 0) z[0][0] = z[0][0] + 8.0f * mu * rootn;  // the 3D DC: Haar row 0, frequency 0
    } else {
        forward3d<PREC>(z, xs, kf, ha);
    }
    float wgt;
    if (STEP == 1) {
        // coefficient (member 4q + t, frequency 16 nb + r); rows >= n are exactly zero
        int nnz = 0;
#pragma unroll
        for (int nb = 0; nb < 4; ++nb)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const bool keep = fabsf(z[nb][t]) > thr;
                z[nb][t] = keep ? z[nb][t] : 0.0f;
                nnz += __popcll(__ballot(keep));
            }
        wgt = nnz > 0 ? 1.0f / (float)nnz : 1.0f;
    } else {
        f4 zb[4];
        if constexpr (CENTRE) {
            const float mu = stack_level(xb, n);
            float c[16];
            centre_stack(c, xb, mu, r < n);
            forward3d<PREC>(zb, c, kf, ha);
            if (lane == 0) zb[0][0] = zb[0][0] + 8.0f * mu * rootn;
        } else {
            forward3d<PREC>(zb, xb, kf, ha);
        }
        float aw = 0.0f;
#pragma unroll
        for (int nb = 0; nb < 4; ++nb)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float bb = zb[nb][t] * zb[nb][t];
                const float d = bb + s2;
                float W = d > 0.0f ? bb * __builtin_amdgcn_rcpf(d) : 1.0f;  // 1-ulp reciprocal
                W = (4 * q + t) < n ? W : 0.0f;
                z[nb][t] = W * z[nb][t];
                aw = __builtin_fmaf(W, W, aw);
            }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) aw = aw + __shfl_xor(aw, off);
        wgt = aw > 0.0f ? __builtin_amdgcn_rcpf(aw) : 1.0f;
    }
    // BF16: the shrunk DC leaves the MFMA path and is added back per pixel below
    float dc_px = 0.0f;
    if constexpr (CENTRE) {
        dc_px = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(z[0][0]))) * inv8root_n(n);
        if (lane == 0) z[0][0] = 0.0f;
    }
    const f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    typename M::vec vt[4];
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) vt[nb] = pack4<PREC>(M::mma(pack4<PREC>(z[nb]), hb, zero));
    const typename M::vec8 v0 = cat8<PREC>(vt[0], vt[1]), v1 = cat8<PREC>(vt[2], vt[3]);
    f4 xt[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) xt[mb] = M::mma32(ki[mb][1], v1, M::mma32(ki[mb][0], v0, zero));
    // lane (member r, quarter q) holds pixels 16 mb + 4 q + t: transpose through
    // LDS, MX_SCR members at a time
#pragma unroll
    for (int hf = 0; hf < NMAX / MX_SCR; ++hf) {
        if (hf * MX_SCR < n) {
            if (r / MX_SCR == hf) {
#pragma unroll
                for (int mb = 0; mb < 4; ++mb)
                    *reinterpret_cast<f4*>(sb + (r % MX_SCR) * MX_SPP + 16 * mb + 4 * q) = xt[mb];
            }
            wave_lds_sync();
#pragma unroll
            for (int m = 0; m < MX_SCR; ++m)
                if (hf * MX_SCR + m < n) {
                    if constexpr (CENTRE)
                        v[hf * MX_SCR + m] = sb[m * MX_SPP + lane] + dc_px;
                    else
                        v[hf * MX_SCR + m] = sb[m * MX_SPP + lane];
                }
            wave_lds_sync();
        }
    }
    return wgt;
}

// One workgroup per (tile of 16 x TY refs, channel); 16 x 8 tiles keep the LDS
// footprint small enough for three workgroups per CU.
template <int STEP, int PREC, int TY, bool MAP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
void k_group_mx(GroupLaunch L) {
    typedef Mx<PREC> M;
    constexpr int DOFS = FOOT * foot_edge(TY);
    __shared__ float2 acc[DOFS];  // (num, den) per footprint pixel
    __shared__ __align__(16) float scratch[4][MX_SCR * MX_SPP];
    __shared__ int turn;  // number of groups of this tile already accumulated
    const int2 item = L.items[blockIdx.x];
    const FrameTab& F = L.tab[item.x];
    if (L.use_gate && *F.gate == 0) return;
    const GroupView p = group_view(F, STEP);
    const int tile = item.y;
    const int tx = tile % p.ntx, ty = tile / p.ntx;
    const int c = blockIdx.y;
    if (threadIdx.x == 0) turn = 0;
    bool timed_out = false;
    // footprint bounds as scalars (member offsets y * fw + x stay scalar products)
    const int fx0 = __builtin_amdgcn_readfirstlane(p.fx0[tx]), fy0 = __builtin_amdgcn_readfirstlane(p.fy0[ty]);
    const int fw = __builtin_amdgcn_readfirstlane(p.fx1[tx]) - fx0;
    const int fh = __builtin_amdgcn_readfirstlane(p.fy1[ty]) - fy0;
    for (int idx = threadIdx.x; idx < DOFS; idx += blockDim.x) acc[idx] = make_float2(0.0f, 0.0f);
    __syncthreads();

    // the wave index as a scalar: group indices, the round loop and the turn
    // wait below stay wave-uniform (scalar branches, no exec-mask bookkeeping)
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int r = lane & 15, q = lane >> 4;
    const int li = lane >> 3, lj = lane & 7;
    float* sb = scratch[wave];
    typename M::vec8 kf[2][4], ki[4][2];
    {
        typename M::vec f16[4][4], i16[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                // B[k = pixel][j = freq] = K[freq][pixel]
                f16[a][nb] = M::pack(kmat(16 * nb + r, 16 * a + 4 * q), kmat(16 * nb + r, 16 * a + 4 * q + 1),
                                     kmat(16 * nb + r, 16 * a + 4 * q + 2), kmat(16 * nb + r, 16 * a + 4 * q + 3));
                // A[i = pixel][k = freq] = K[freq][pixel]
                i16[a][nb] = M::pack(kmat(16 * nb + 4 * q, 16 * a + r), kmat(16 * nb + 4 * q + 1, 16 * a + r),
                                     kmat(16 * nb + 4 * q + 2, 16 * a + r), kmat(16 * nb + 4 * q + 3, 16 * a + r));
            }
        // K = 32 operands: pixel (forward) or frequency (inverse) blocks 2h and 2h + 1
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                kf[h][nb] = cat8<PREC>(f16[2 * h][nb], f16[2 * h + 1][nb]);
                ki[nb][h] = cat8<PREC>(i16[nb][2 * h], i16[nb][2 * h + 1]);
            }
    }
    const float kw = c_mx_kaiser[lane];
    const float* __restrict__ src = F.z[c];
    const float* __restrict__ bsrc = F.bz[c];
    const float sc = F.sig[c];
    const float thr = 2.7f * sc;
    const float s2 = sc * sc;
    // noise-profile map (MAP): each group shrinks with its reference patch's sigma
    // and weighs its aggregation by 1/sigma^2 (which a single sigma cancels)
    const float* __restrict__ smap = MAP ? F.smap + (size_t)c * p.nx * p.ny : nullptr;
    const int w = p.w;
    const int nbx = min(TILE, p.nx - tx * TILE), nby = min(TY, p.ny - ty * TY);
    const int G = nbx * nby;

    const int lo = (li - fy0) * fw + (lj - fx0);

    // Group g of this wave: its match list through the scalar cache
    // (wave-uniform address), size and sigma; this lane's member r = mp[r].
    // The stacks are loaded after the previous group's turn: with the DCT
    // operand fragments held in registers, stacks in flight during the turn
    // would not fit in the three-waves-per-SIMD register budget.
    for (int rnd = 0; rnd < (G + 3) / 4; ++rnd) {
        const int g = rnd * 4 + wave;
        const bool has = g < G;
        int n = 0;
        float wgt = 0.0f;
        float v[NMAX];
        uint32_t mp[NMAX], pm = 0;
        if (has) {
            const int bx = tx * TILE + g % nbx, by = ty * TY + g / nbx;
            const size_t ref = (size_t)by * p.nx + bx;
            n = floor_pow2(load_count(p.cnt, ref));
            const uint32_t CONST_AS* lp = (const uint32_t CONST_AS*)(p.pos + ref * NMAX);
#pragma unroll
            for (int m = 0; m < NMAX; ++m) mp[m] = lp[m];
            pm = mp[0];
#pragma unroll
            for (int m = 1; m < NMAX; ++m) pm = r == m ? mp[m] : pm;
            float gthr = thr, gs2 = s2;
            if (MAP) {
                const float sg = smap[ref];
                gthr = 2.7f * sg;
                gs2 = sg * sg;
            }
            float xs[16], xb[16];
            const bool live = r < n;
            load_stack(xs, src, w, pm, live, q);
            if (STEP == 2) load_stack(xb, bsrc, w, pm, live, q);
            wgt = filter_mx<STEP, PREC>(v, n, r, q, lane, xs, xb, sb, kf, ki, gthr, gs2);
            if (MAP) wgt = wgt * (1.0f / (gs2 > 1e-30f ? gs2 : 1.0f));
        }
        timed_out |= add_group_in_turn(acc, &turn, g, has, n, mp, pm, v, wgt * kw, lo, fw);
    }
    if (timed_out && lane == 0) atomicOr(L.err, 1);
    __syncthreads();
    float* dst = p.partial + ((size_t)c * p.ntx * p.nty + tile) * p.tile_stride;
    for (int idx = threadIdx.x; idx < fw * fh; idx += blockDim.x) {
        const float2 a = acc[idx];
        dst[idx] = a.x;
        dst[p.dofs + idx] = a.y;
    }
}

}  // namespace

int group_mx_tables(int device, const float* dct64, const float* kaiser64) {
    (void)device;  // the caller has made `device` current
    float haar[5][256] = {};
    for (int lvl = 0; lvl < 5; ++lvl) {
        const int n = 1 << lvl;
        for (int j = 0; j < n; ++j) {
            double v[16] = {};
            v[j] = 1.0;
            const double s = std::sqrt(0.5);
            for (int len = n; len > 1; len >>= 1) {
                double tmp[16];
                const int hh = len >> 1;
                for (int k = 0; k < hh; ++k) {
                    tmp[k] = (v[2 * k] + v[2 * k + 1]) * s;
                    tmp[hh + k] = (v[2 * k] - v[2 * k + 1]) * s;
                }
                for (int k = 0; k < len; ++k) v[k] = tmp[k];
            }
            for (int i = 0; i < n; ++i) haar[lvl][i * 16 + j] = (float)v[i];
        }
    }
    hipError_t e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_mx_dct), dct64, 64 * sizeof(float))) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_mx_kaiser), kaiser64, 64 * sizeof(float))) != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(c_mx_haar), haar, sizeof(haar));
}

template <int TY, bool MAP>
void launch_ty(int step, bool bf, const GroupLaunch& p, dim3 grid, hipStream_t st) {
    if (step == 1) {
        if (bf)
            k_group_mx<1, CPIPE_BM3D_PRECISION_BF16, TY, MAP><<<grid, 256, 0, st>>>(p);
        else
            k_group_mx<1, CPIPE_BM3D_PRECISION_FP16, TY, MAP><<<grid, 256, 0, st>>>(p);
    } else {
        if (bf)
            k_group_mx<2, CPIPE_BM3D_PRECISION_BF16, TY, MAP><<<grid, 256, 0, st>>>(p);
        else
            k_group_mx<2, CPIPE_BM3D_PRECISION_FP16, TY, MAP><<<grid, 256, 0, st>>>(p);
    }
}

int group_mx_launch(int step, int precision, int tile_y, bool map, const GroupLaunch& p, dim3 grid,
                    hipStream_t st) {
    const bool bf = precision == CPIPE_BM3D_PRECISION_BF16;
    if (tile_y == TILE / 2 && map)
        launch_ty<TILE / 2, true>(step, bf, p, grid, st);
    else if (tile_y == TILE / 2)
        launch_ty<TILE / 2, false>(step, bf, p, grid, st);
    else if (tile_y == TILE && map)
        launch_ty<TILE, true>(step, bf, p, grid, st);
    else if (tile_y == TILE)
        launch_ty<TILE, false>(step, bf, p, grid, st);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace cpipe_internal
