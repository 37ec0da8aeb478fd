// BM3D denoise engine for MI355X (gfx950): kernels and the C ABI of include/cpipe_bm3d.h.
//
// Data layout in HBM (per plan, dense FP32 planes, row pitch = width):
//   z[c]      noisy planes in the opponent colour space (c < channels)
//   bz[c]     basic estimate (step-1 output), opponent space
//   pos/cnt   match lists: per reference position 16 packed (y<<16 | x) + raw count
//   partial   per (channel, tile) num/den planes of the tile footprint (<= 91x91)
// Reference positions: x in {0, 3, 6, ...} < W-8, plus W-8 (same in y).
// A tile is 16x4 reference positions (x by y) in exact FP32 plans, 16x8 (or,
// retuned, 16x16) in mixed-precision plans; its footprint is every pixel a member
// patch of one of its groups can touch (refs +- 19 px search, + 8 px patch).
//
// Arithmetic is FP32 with a fixed operation order (explicit fmaf, -ffp-contract=off):
//   SSD      = (((r0 + r1) + r2) + ((r3 + r4) + r5)) + (r6 + r7), row sum
//              r_i = (sq3(e0..e2) + sq3(e3..e5)) + sq2(e6, e7), e = ref - cand,
//              sq2/sq3 = product + fma chain
//   DCT      = row pass then column pass, each an fma chain in index order
//   Haar     = orthonormal multi-level, (a+b)*s / (a-b)*s
//   Wiener   = W = b^2/(b^2+s^2); weight 1/sum W^2 (lane tree 32,16,..,1)
//   HT       = keep |c| > 2.7 sigma_c; weight 1/N_nonzero
//   tile sum = groups in ref raster order, members in match order, num += wk*v
//   merge    = tiles in (ty, tx) ascending order, out = num / den
// so the result is reproducible run to run and equal to a CPU FP32 restatement
// that follows the same order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "bm3d_shared.h"
#include "host_staging.h"
#include "cpipe_bm3d.h"
#include "pilot.h"

// Learned-pilot weights embedded at build time (cpipe_amd/weights/pilot_v1.bin).
extern "C" const unsigned char CPIPE_BM3D_PILOT_BLOB[];
extern "C" const size_t CPIPE_BM3D_PILOT_BLOB_SIZE;

using namespace cpipe_internal;

namespace {
// Block matching workgroup: 16 ref columns x 14 ref rows, two waves. A wave
// owns 4 column pairs; a pair is 16 consecutive lanes (one DPP row): 14 refs
// of each of its two columns + 2 helper lanes that only compute patch-row sums
// for the refs above them. Launches too small to fill the GPU run one column
// per DPP row instead (four waves per workgroup).
constexpr int BM_TX = 16;
constexpr int BM_TY = 14;
constexpr int BM_NT = 128;                                      // threads per workgroup (pairs)
constexpr int BM_COLS = (BM_TX - 1) * STRIDE + 2 * R + K + 1;  // 92: +1 for the dx=20 overrun
constexpr int BM_LH = (BM_TY + 1) * STRIDE + 2 * R + 3;        // 86: helper rows included
constexpr int BM_LW = BM_COLS;  // LDS row pitch: a wave's 16 rows x 4 pairs (6 px apart) meet
                                // at most 2-way bank conflicts with any pitch
constexpr float BM_OUTSIDE = 1.0e18f;  // staged for pixels outside the image (SSD >= 1e36 > tau)
constexpr int PEND = 8;     // per-ref queue of candidate keys awaiting top-16 insertion
// four workgroups per CU: 31.6 KB of search region + 8 KB of queues each
static_assert(4 * (sizeof(float) * BM_LW * BM_LH + sizeof(float) * PEND * 2 * BM_NT) <= 160 * 1024,
              "block matching LDS");

enum Status { OK = 0, FAILED = 1, OOM = 3, BAD_INDEX = 5, INTERNAL = 7, UNSUPPORTED = 8 };

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_CHECK(expr)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            return fail(e_ == hipErrorOutOfMemory ? OOM : INTERNAL,                      \
                        std::string(#expr) + ": " + hipGetErrorString(e_));              \
        }                                                                                \
    } while (0)

__constant__ float c_kaiser2[64]; // Kaiser(8, beta=2) outer product

__host__ __device__ inline int ref_pos(int i, int n) {
    const int last = n - K;
    const int p = i * STRIDE;
    return p < last ? p : last;
}

inline int ref_count(int n) { return n < K ? 0 : (n - K + STRIDE - 1) / STRIDE + 1; }

// Tau on the un-normalised 64-pixel SSD (tau * k^2), Lebrun's 2500 and 400 on 0..255.
inline float tau_ssd(int step) {
    const double t = step == 1 ? 2500.0 / (255.0 * 255.0) : 400.0 / (255.0 * 255.0);
    return (float)(64.0 * t);
}

// ---------------------------------------------------------------------------
// Colour
// ---------------------------------------------------------------------------

__global__ void k_opp_fwd(const float* __restrict__ r, const float* __restrict__ g,
                          const float* __restrict__ b, size_t pitch, float* __restrict__ z0,
                          float* __restrict__ z1, float* __restrict__ z2, int w, int h) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const size_t i = (size_t)y * pitch + x;
    const size_t o = (size_t)y * w + x;
    const float rv = r[i], gv = g[i], bv = b[i];
    const float c13 = (float)(1.0 / 3.0);
    z0[o] = ((rv + gv) + bv) * c13;
    z1[o] = (rv - bv) * 0.5f;
    z2[o] = ((rv + bv) - (gv + gv)) * 0.25f;
}

// Input colour transform for every frame of a plan: one block row per pixel
// row of a frame rectangle (blockIdx.z = rectangle, blockIdx.y = row in it).
__global__ void k_colour_in(const FrameTab* __restrict__ tab, const FrameRect* __restrict__ rects) {
    const FrameRect rc = rects[blockIdx.z];
    const FrameTab& F = tab[rc.f];
    const int x = rc.x0 + blockIdx.x * blockDim.x + threadIdx.x;
    const int y = rc.y0 + blockIdx.y;
    if (x >= rc.x1 || y >= rc.y1) return;
    const size_t i = (size_t)y * F.in_pitch + x;
    const size_t o = (size_t)y * F.w + x;
    if (F.c == 3) {
        const float rv = F.in[0][i], gv = F.in[1][i], bv = F.in[2][i];
        const float c13 = (float)(1.0 / 3.0);
        F.z[0][o] = ((rv + gv) + bv) * c13;
        F.z[1][o] = (rv - bv) * 0.5f;
        F.z[2][o] = ((rv + bv) - (gv + gv)) * 0.25f;
    } else {
        F.z[0][o] = F.in[0][i];
    }
}

// Noise-profile sigma map of frame blockIdx.z (one thread per reference position):
// x = mean of the reference's 8x8 input patch per channel (row-major FP32 sum
// times 1/64), v = max(0, beta + lambda x); opponent channel k gets
// sqrt(sum_c A_kc^2 v_c) for the opponent rows A = (1/3, 1/3, 1/3),
// (1/2, 0, -1/2), (1/4, -1/2, 1/4); grey planes sqrt(v).
__global__ void k_sigma_map(const FrameTab* __restrict__ tab) {
    const FrameTab& F = tab[blockIdx.z];
    const int bx = blockIdx.x * blockDim.x + threadIdx.x, by = blockIdx.y;
    if (F.smap == nullptr || bx >= F.nx || by >= F.ny) return;
    float* __restrict__ smap = F.smap;
    const int x0 = ref_pos(bx, F.w), y0 = ref_pos(by, F.h);
    float v[3] = {0.0f, 0.0f, 0.0f};
    for (int c = 0; c < F.c; ++c) {
        const float* __restrict__ src = F.in[c] + (size_t)y0 * F.in_pitch + x0;
        float sum = 0.0f;
        for (int i = 0; i < K; ++i)
            for (int j = 0; j < K; ++j) sum = sum + src[(size_t)i * F.in_pitch + j];
        const float t = F.nbeta[c] + F.nlambda[c] * (sum * 0.015625f);
        v[c] = t > 0.0f ? t : 0.0f;
    }
    const size_t nref = (size_t)F.nx * F.ny, r = (size_t)by * F.nx + bx;
    if (F.c == 1) {
        smap[r] = sqrtf(v[0]);
    } else {
        smap[r] = sqrtf(((v[0] + v[1]) + v[2]) * (float)(1.0 / 9.0));
        smap[nref + r] = sqrtf((v[0] + v[2]) * 0.25f);
        smap[2 * nref + r] = sqrtf(((v[0] + v[2]) * 0.0625f) + (v[1] * 0.25f));
    }
}

// Frames smaller than one patch: output = input (blockIdx.z = frame rectangle).
__global__ void k_copy_through(const FrameTab* __restrict__ tab, const FrameRect* __restrict__ rects) {
    const FrameRect rc = rects[blockIdx.z];
    const FrameTab& F = tab[rc.f];
    const int x = rc.x0 + blockIdx.x * blockDim.x + threadIdx.x;
    const int y = rc.y0 + blockIdx.y;
    if (x >= rc.x1 || y >= rc.y1) return;
    for (int c = 0; c < F.c; ++c) F.out[c][(size_t)y * F.out_pitch + x] = F.in[c][(size_t)y * F.in_pitch + x];
}

// Sets one int to a value (a kernel rather than hipMemsetAsync, so that runs
// captured into a graph clear the plan's error flag on every replay).
__global__ void k_set_flag(int* __restrict__ flag, int v) { *flag = v; }

__global__ void k_copy_plane(const float* __restrict__ src, size_t pitch, float* __restrict__ dst,
                             size_t dpitch, int w) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    dst[(size_t)y * dpitch + x] = src[(size_t)y * pitch + x];
}

__global__ void k_opp_inv(const float* __restrict__ z0, const float* __restrict__ z1,
                          const float* __restrict__ z2, float* __restrict__ r,
                          float* __restrict__ g, float* __restrict__ b, size_t pitch, int w) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const size_t i = (size_t)y * w + x;
    const size_t o = (size_t)y * pitch + x;
    const float c23 = (float)(2.0 / 3.0), c43 = (float)(4.0 / 3.0);
    const float yv = z0[i], u = z1[i], v = z2[i];
    r[o] = (yv + u) + v * c23;
    g[o] = yv - v * c43;
    b[o] = (yv - u) + v * c23;
}

__device__ inline float h2f(uint16_t bits) { return (float)__builtin_bit_cast(_Float16, bits); }
__device__ inline uint16_t f2h(float v) { return __builtin_bit_cast(uint16_t, (_Float16)v); }

__global__ void k_unpack_rgba16f(const uint16_t* __restrict__ in, size_t isp, float* __restrict__ r,
                                 float* __restrict__ g, float* __restrict__ b, int w) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const uint16_t* p = in + ((size_t)y * isp + x) * 4;
    const size_t o = (size_t)y * w + x;
    r[o] = h2f(p[0]);
    g[o] = h2f(p[1]);
    b[o] = h2f(p[2]);
}

__global__ void k_pack_rgba16f(const float* __restrict__ r, const float* __restrict__ g,
                               const float* __restrict__ b, const uint16_t* __restrict__ in,
                               size_t isp, uint16_t* __restrict__ out, size_t osp, int w) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const size_t i = (size_t)y * w + x;
    uint16_t* q = out + ((size_t)y * osp + x) * 4;
    q[0] = f2h(r[i]);
    q[1] = f2h(g[i]);
    q[2] = f2h(b[i]);
    q[3] = in[((size_t)y * isp + x) * 4 + 3];
}

// ---------------------------------------------------------------------------
// Reference-compatible filter: bm3d_step1_generator.cpp:9-16,42-56 (2x2 box
// mean blended by s = clamp(sigma/0.03, 0, 1), RGB clamped >= 0, alpha kept;
// repeat_edge at the borders) and bm3d_step2_generator.cpp:33-42 (identity).
// ---------------------------------------------------------------------------

__global__ void k_compat_rgba16f(const uint16_t* __restrict__ in, size_t isp,
                                 uint16_t* __restrict__ out, size_t osp, int w, int h, float sigma,
                                 int stage) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const float strength = fminf(1.0f, fmaxf(0.0f, sigma / 0.03f));
    const int bx = (x / 2) * 2, by = (y / 2) * 2;
    const int x0 = bx, x1 = min(bx + 1, w - 1);
    const int y0 = by, y1 = min(by + 1, h - 1);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float v = h2f(in[((size_t)y * isp + x) * 4 + c]);
        float r;
        if (stage == 2) {
            r = v + sigma * 0.0f;
        } else if (c == 3) {
            r = v;
        } else {
            const float a00 = h2f(in[((size_t)y0 * isp + x0) * 4 + c]);
            const float a10 = h2f(in[((size_t)y0 * isp + x1) * 4 + c]);
            const float a01 = h2f(in[((size_t)y1 * isp + x0) * 4 + c]);
            const float a11 = h2f(in[((size_t)y1 * isp + x1) * 4 + c]);
            const float basic = (((a00 + a10) + a01) + a11) * 0.25f;
            const float pa = v * (1.0f - strength);
            const float pb = basic * strength;
            r = fmaxf(0.0f, pa + pb);
        }
        out[((size_t)y * osp + x) * 4 + c] = f2h(r);
    }
}

// ---------------------------------------------------------------------------
// Block matching, exact FP32.
// One workgroup (two waves) = 16 x 14 reference positions. Lane (row a, pair
// cg) takes the refs at tile columns 2 cg and 2 cg + 1, 3 pixels apart: for one
// displacement both refs use the differences of the same 11 pixel columns, and
// with patch rows summed in column chunks of 3, 3 and 2 (ssd_row), the second
// ref's first chunk is the first ref's middle chunk and its second chunk
// extends the first ref's last one, so a pair costs 13 operations per patch
// row and ref instead of 16.
// The tile's search region (92 x 86 px) is staged in LDS once; every lane
// scans the clipped 39x39 window in (dy, dx) raster order, 4 displacements at
// a time (14 LDS reads per patch row serve 8 SSDs), and keeps a sorted top-16
// of rank keys (bm_key: SSD to 12 mantissa bits, then scan position) per ref in
// registers, reference pinned first. Only SSD < tau enters. Qualifying
// candidates wait as keys in a short LDS queue per ref and are inserted in
// batches, so the wave runs the insertion network once per batch instead of
// once per candidate. Refs of a
// column are 3 rows apart and the 8 row sums add up in groups of 3, 3 and 2,
// so each lane sums its own 3 patch rows and takes the next group and the
// pair after it from the two lanes below (DPP row shifts): same sums, same
// order.
// ---------------------------------------------------------------------------

// Keyed slots per ref: the reference and the 19 candidates with the smallest
// rank keys. The list keeps the 15 of those 19 that come first by exact (SSD,
// scan position): a candidate of the exact top 15 is among the 19 smallest
// keys unless five candidates after it in scan order with larger exact SSDs
// share its 2^-12 bucket, so the lists are the exact (SSD, position) top 15 in
// practice and by definition equal the CPU FP32 restatement's, which keeps the
// same two stages.
constexpr int BM_NK = 20;

__device__ __forceinline__ void insert_key(float (&bd)[BM_NK], float s) {
    // bd: sorted keys, bd[0] the pinned reference (below every key); the new
    // key of slot k is the median of (bd[k-1], s, bd[k]) (one v_med3). A key
    // at or above bd[BM_NK - 1] (or +inf: no candidate) changes nothing.
#pragma unroll
    for (int k = BM_NK - 1; k >= 1; --k) bd[k] = __builtin_amdgcn_fmed3f(bd[k - 1], s, bd[k]);
}

// Exact SSDs, in the chunked order, of the ref patch at tile offset
// (lry, lrx) against M candidate patches at tile offsets cb[m].
template <int M>
__device__ __forceinline__ void tile_ssds(const float* __restrict__ tile, int lry, int lrx, const int (&cb)[M],
                                          float (&es)[M]) {
    // row sums r0..r7 added as ((r0 + r1) + r2) + ((r3 + r4) + r5) + (r6 + r7)
    float gs[M];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        float rr[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) rr[j] = tile[(lry + i) * BM_LW + lrx + j];
#pragma unroll
        for (int m = 0; m < M; ++m) {
            float e[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) e[j] = rr[j] - tile[cb[m] + i * BM_LW + j];
            const float rs = ssd_row(e);
            gs[m] = (i == 0 || i == 3 || i == 6) ? rs : gs[m] + rs;
            if (i == 2) es[m] = gs[m];
            if (i == 5 || i == 7) es[m] = es[m] + gs[m];
        }
    }
}

__device__ __forceinline__ float dpp_row_shl(float v, int n) {
    // lane l receives lane l+n of its 16-lane row (0 past the row end)
    const int x = __builtin_bit_cast(int, v);
    const int r = n == 1 ? __builtin_amdgcn_update_dpp(0, x, 0x101, 0xf, 0xf, true)
                         : __builtin_amdgcn_update_dpp(0, x, 0x102, 0xf, 0xf, true);
    return __builtin_bit_cast(float, r);
}

// One workgroup per item (frame, tile column, tile row) of `items`; RPL refs
// per lane: 2 (pairs of columns, two waves) or 1 (four waves, for launches too
// small to fill the GPU with pairs). Both sum every SSD in the same order.
template <int RPL>
__global__ __launch_bounds__(BM_NT * 2 / RPL) void k_block_match(const FrameTab* __restrict__ tab,
                                                                 const int4* __restrict__ items, int step,
                                                                 float tau, int use_gate) {
    constexpr int NT = BM_NT * 2 / RPL;
    __shared__ float tile[BM_LW * BM_LH];
    __shared__ float qs[PEND * RPL * NT];  // [slot][ref of the lane][thread]: rank key
    const int4 it = items[blockIdx.x];
    const FrameTab& F = tab[it.x];
    if (use_gate && *F.gate == 0) return;  // step skipped (decided on the device)
    const float* __restrict__ plane = step == 1 ? F.z[0] : F.bz[0];
    const int w = F.w, h = F.h, nx = F.nx, ny = F.ny;
    uint32_t* __restrict__ pos = F.pos[step - 1];
    uint8_t* __restrict__ cnt = F.cnt[step - 1];
    const int bx0 = it.y * BM_TX, by0 = it.z * BM_TY;
    const int X0 = ref_pos(bx0, w) - R;  // LDS origin in image coordinates (may be < 0)
    const int Y0 = ref_pos(by0, h) - R;
    const int tid = threadIdx.x;
    {
        // every load of the thread in flight before the first LDS store
        constexpr int NST = BM_COLS * BM_LH, IT = (NST + NT - 1) / NT;
        float sv[IT];
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const int idx = tid + NT * i;
            const int yy = idx / BM_COLS, xx = idx - yy * BM_COLS;
            const int gx = X0 + xx, gy = Y0 + yy;
            sv[i] = BM_OUTSIDE;
            if (idx < NST && gx >= 0 && gx < w && gy >= 0 && gy < h) sv[i] = plane[(size_t)gy * w + gx];
        }
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const int idx = tid + NT * i;
            if (idx < NST) tile[idx] = sv[i];
        }
    }
    __syncthreads();

    const int l = tid & 63;
    const int a = l & 15;                            // ref row within the tile (14, 15: helpers)
    const int bxp = bx0 + RPL * ((tid >> 6) * 4 + (l >> 4));  // the lane's (first) ref column
    const int by = by0 + a;
    bool real[RPL];
    int rx[RPL];
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
        real[k] = a < BM_TY && bxp + k < nx && by < ny;
        rx[k] = ref_pos(min(bxp + k, nx - 1), w);
    }
    // Regular tile: refs of a column sit exactly 3 rows apart, so a ref's
    // patch rows 3..7 are the rows 0..2 / 0..1 of the refs one and two below.
    // Only a tile holding the image's last ref row at a shorter stride is not.
    const int last_in = ny - 1 - by0;
    const bool regular = !(last_in >= 0 && last_in < BM_TY &&
                           ref_pos(ny - 1, h) != ref_pos(by0, h) + STRIDE * last_in);
    // Likewise the two refs of a pair sit 3 columns apart, unless the image's
    // last ref column, at a shorter stride, is the second of its pair.
    const int lastx_in = nx - 1 - bx0;
    const bool paired = RPL == 1 || !(lastx_in >= 0 && lastx_in < BM_TX && (lastx_in & 1) &&
                                      ref_pos(nx - 1, w) != ref_pos(nx - 2, w) + STRIDE);
    const int ry = regular ? ref_pos(by0, h) + STRIDE * a : ref_pos(min(by, ny - 1), h);
    const int lry = ry - Y0;

    const float inf = __builtin_inff();
    float bd[RPL][BM_NK];  // rank keys, ascending; slot 0: the reference, pinned below every key
    // queue filter: below the ref's threshold (tau; -1 for a ref outside the
    // frame) and below the end of the last slot's bucket (stale between merges)
    float thr0[RPL], lim[RPL];
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
#pragma unroll
        for (int m = 0; m < BM_NK; ++m) bd[k][m] = inf;
        bd[k][0] = -1.0f;
        thr0[k] = real[k] ? tau : -1.0f;
        lim[k] = thr0[k];
    }

    // Candidates that may precede the last kept key are queued (in scan
    // order) and merged in batches; the queue filter is stale and looser, so
    // the result equals inserting each candidate immediately. A ref's queue is
    // merged before a block of 4 displacements when some lane could overflow
    // it, so the wave runs the insertion network on batches of 5 to 8; lanes
    // with fewer queued keys insert +inf, which changes nothing.
    float* const qsl = qs + tid;  // slot s of ref k at [(RPL s + k) * NT]
    int np[RPL];
#pragma unroll
    for (int k = 0; k < RPL; ++k) np[k] = 0;
    auto flush_ref = [&](int k) __attribute__((always_inline)) {
        float e[PEND];
#pragma unroll
        for (int s = 0; s < PEND; ++s) e[s] = qsl[(RPL * s + k) * NT];  // reads in flight together
#pragma unroll
        for (int s = 0; s < PEND; ++s) {
            if (!__any(s < np[k])) break;
            insert_key(bd[k], s < np[k] ? e[s] : inf);
        }
        np[k] = 0;
        lim[k] = fminf(thr0[k], bm_key_bucket_end(bd[k][BM_NK - 1]));
    };
    auto flush = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < RPL; ++k) flush_ref(k);
    };
    auto push = [&](int k, float s, int code) {
        qsl[(RPL * np[k] + k) * NT] = bm_key(s, (uint32_t)code);
        ++np[k];
    };

    if (regular && paired) {
        // Candidates whose patch leaves the image read BM_OUTSIDE pixels, so
        // their SSD (>= 1e36) fails s < tau by itself; what is left to reject
        // is uniform: dx = R + 1 (the last 4-wide block's overrun), the ref
        // itself, and helper lanes (threshold -1). The row sums then stay
        // outside any branch and the DPP row shifts fold into the adds.
        const int lrx = rx[0] - X0;
        constexpr int RW = 3 * RPL + 5;  // columns of the lane's refs: 11 (pair) or 8
        float ref[3][RW];  // the lane's 3 patch rows
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < RW; ++j) ref[i][j] = tile[(lry + i) * BM_LW + lrx + j];
        for (int dy = -R; dy <= R; ++dy) {
            const float* rowp = tile + (lry + dy) * BM_LW + lrx;
            const int cy = (dy + R) * BM_WIN + R;
            for (int dx0 = -R; dx0 <= R; dx0 += 4) {
                // room for this block's 4 candidates in every lane's queues
#pragma unroll
                for (int k = 0; k < RPL; ++k)
                    if (__any(np[k] > PEND - 4)) flush_ref(k);
                float rs[RPL][3][4];
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    float c[RW + 3];
#pragma unroll
                    for (int t = 0; t < RW + 3; ++t) c[t] = rowp[i * BM_LW + dx0 + t];
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) {
                        float e[RW];
#pragma unroll
                        for (int j = 0; j < RW; ++j) e[j] = ref[i][j] - c[kk + j];
                        const float a3 = sq3(e[3], e[4], e[5]);
                        const float b6 = sq2(e[6], e[7]);
                        rs[0][i][kk] = (sq3(e[0], e[1], e[2]) + a3) + b6;
                        if constexpr (RPL == 2)
                            rs[RPL - 1][i][kk] = (a3 + __builtin_fmaf(e[8], e[8], b6)) + sq2(e[9], e[10]);
                    }
                }
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const int dx = dx0 + kk;
                    const bool skip = dx > R || (dy == 0 && dx == 0);
#pragma unroll
                    for (int k = 0; k < RPL; ++k) {
                        // rows 0..2 (this lane), 3..5 (the lane below), 6..7 (two below)
                        const float q = rs[k][0][kk] + rs[k][1][kk];
                        const float pr = q + rs[k][2][kk];
                        float sum = pr + dpp_row_shl(pr, 1);
                        sum = sum + dpp_row_shl(q, 2);
                        if (!skip && sum < lim[k]) push(k, sum, cy + dx);
                    }
                }
            }
        }
    } else {
        // A tile at the image's last ref row or column: every ref sums all 8
        // of its patch rows itself, with the same row sums and order.
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
            const int lrx = rx[k] - X0;
            float ref[64];
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j) ref[i * 8 + j] = tile[(lry + i) * BM_LW + lrx + j];
            for (int dy = -R; dy <= R; ++dy) {
                const int qy = ry + dy;
                if (!real[k] || qy < 0 || qy > h - K) continue;
                const float* rowp = tile + (lry + dy) * BM_LW + lrx;
                const int cy = (dy + R) * BM_WIN + R;
                for (int dx0 = -R; dx0 <= R; dx0 += 4) {
                    if (__any(np[k] > PEND - 4)) flush();
                    float s[4], g[4];  // patch sum, sum of the current group of rows
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        float c[11];
#pragma unroll
                        for (int t = 0; t < 11; ++t) c[t] = rowp[i * BM_LW + dx0 + t];
#pragma unroll
                        for (int kk = 0; kk < 4; ++kk) {
                            float e[8];
#pragma unroll
                            for (int j = 0; j < 8; ++j) e[j] = ref[i * 8 + j] - c[kk + j];
                            const float r = ssd_row(e);
                            g[kk] = (i == 0 || i == 3 || i == 6) ? r : g[kk] + r;
                            if (i == 2) s[kk] = g[kk];
                            if (i == 5 || i == 7) s[kk] = s[kk] + g[kk];
                        }
                    }
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) {
                        const int dx = dx0 + kk, qx = rx[k] + dx;
                        const bool valid = dx <= R && qx >= 0 && qx <= w - K && (dx != 0 || dy != 0);
                        if (valid && s[kk] < lim[k]) push(k, s[kk], cy + dx);
                    }
                }
            }
        }
    }
    flush();
    // the 19 keyed candidates in exact (SSD, scan position) order; the first 15
    // follow the reference
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
        if (!real[k]) continue;
        constexpr int M = BM_NK - 1;
        const int lrx = rx[k] - X0;
        int cb[M], nv = 0;
        uint32_t code[M];
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const bool ok = bd[k][1 + m] < inf;
            const uint32_t c = __float_as_uint(bd[k][1 + m]) & BM_CODE_MASK;
            const int cdy = (int)(c / BM_WIN), cdx = (int)c - cdy * BM_WIN;
            code[m] = ok ? c : 0xffffu;
            cb[m] = ok ? (lry + cdy - R) * BM_LW + lrx + cdx - R : lry * BM_LW + lrx;
            nv += ok ? 1 : 0;
        }
        // Order key (SSD bits << 11) | position code. Kept keys in different
        // 2^-12 buckets are already ordered by their truncated SSD, so only the
        // slots whose bucket a neighbouring slot shares need their exact SSD:
        // each lane takes its next such slot per pass (same code path for all
        // lanes, different slots), the others keep the truncated bits.
        uint64_t ex[M];
        uint32_t need = 0;  // slots sharing a bucket with a neighbour
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const uint32_t t = __float_as_uint(bd[k][1 + m]) & ~BM_CODE_MASK;
            ex[m] = code[m] != 0xffffu ? ((uint64_t)t << 11) | code[m] : ~0ull;
            if (m > 0 && code[m] != 0xffffu &&
                (__float_as_uint(bd[k][m]) & ~BM_CODE_MASK) == t)
                need |= 3u << (m - 1);
        }
        while (__any(need != 0)) {
            const bool has = need != 0;
            const int j = has ? __builtin_ctz(need) : 0;
            need &= need - 1u;
            int cbj[1] = {cb[0]};
#pragma unroll
            for (int m = 1; m < M; ++m) cbj[0] = j == m ? cb[m] : cbj[0];
            float sj[1];
            tile_ssds(tile, lry, lrx, cbj, sj);
#pragma unroll
            for (int m = 0; m < M; ++m)
                if (has && j == m) ex[m] = ((uint64_t)__float_as_uint(sj[0]) << 11) | code[m];
        }
        const size_t r = (size_t)by * nx + bxp + k;
        uint32_t* o = pos + r * NMAX;
        const int n = 1 + min(NMAX - 1, nv);
        o[0] = ((uint32_t)ry << 16) | (uint32_t)rx[k];
#pragma unroll
        for (int m = 0; m < M; ++m) {
            int rank = 0;
#pragma unroll
            for (int j = 0; j < M; ++j) rank += ex[j] < ex[m] ? 1 : 0;
            const uint32_t c = code[m];
            const uint32_t cdy = c / BM_WIN;
            const uint32_t qy = (uint32_t)ry + cdy - R, qx = (uint32_t)rx[k] + (c - cdy * BM_WIN) - R;
            if (c != 0xffffu && rank < NMAX - 1) o[1 + rank] = (qy << 16) | qx;
        }
#pragma unroll
        for (int t = 1; t < NMAX; ++t)
            if (t >= n) o[t] = 0u;
        cnt[r] = (uint8_t)n;
    }
}

// Exact block matching over n items: pairs of refs per lane (two waves per
// item) once a launch has items for about four workgroups per CU; one ref per
// lane (four waves per item) below that, where the pairs' fewer waves would
// leave most of the GPU idle.
constexpr int BM_PAIR_ITEMS = 1024;

void launch_block_match(int n_items, const FrameTab* tab, const int4* items, int step, int use_gate,
                        hipStream_t st) {
    if (n_items >= BM_PAIR_ITEMS)
        k_block_match<2><<<n_items, BM_NT, 0, st>>>(tab, items, step, tau_ssd(step), use_gate);
    else
        k_block_match<1><<<n_items, 2 * BM_NT, 0, st>>>(tab, items, step, tau_ssd(step), use_gate);
}

// ---------------------------------------------------------------------------
// Group transform + shrinkage + tile aggregation (fused), exact FP32.
// One workgroup (4 waves) = one (tile, channel). Lane p = (i, j) holds pixel
// (i, j) of every member patch of its wave's group in registers; the 2D DCT
// passes go through a per-wave LDS scratch (16-byte row reads); the Haar transform
// along the stack is in-register. Groups are taken in ref raster order, four
// at a time (one per wave), and added to the LDS num/den footprint strictly in
// group order (an LDS turn token passed from wave to wave), so the accumulation
// order is fixed while the other waves keep filtering their next groups
// (plain read-add-write: LDS float atomics cost ~180 cycles per instruction here).
// ---------------------------------------------------------------------------


// 8-point orthonormal DCT-II and inverse, even/odd butterfly (36 operations
// each; the CPU FP32 restatement uses the same operations in the same order).
// Q0 = 1/sqrt(8), Qk = cos(k pi/16)/2, each rounded once to FP32.
constexpr float DQ0 = 0x1.6a09e6p-2f, DQ1 = 0x1.f6297cp-2f, DQ2 = 0x1.d906bcp-2f,
                DQ3 = 0x1.a9b662p-2f, DQ5 = 0x1.1c73b4p-2f, DQ6 = 0x1.87de2ap-3f,
                DQ7 = 0x1.8f8b84p-4f;

__device__ __forceinline__ void fdct8(const float (&x)[8], float (&X)[8]) {
    const float a0 = x[0] + x[7], a1 = x[1] + x[6], a2 = x[2] + x[5], a3 = x[3] + x[4];
    const float b0 = x[0] - x[7], b1 = x[1] - x[6], b2 = x[2] - x[5], b3 = x[3] - x[4];
    const float e0 = a0 + a3, e1 = a1 + a2, f0 = a0 - a3, f1 = a1 - a2;
    X[0] = (e0 + e1) * DQ0;
    X[4] = (e0 - e1) * DQ0;
    X[2] = __builtin_fmaf(DQ2, f0, DQ6 * f1);
    X[6] = __builtin_fmaf(DQ6, f0, -(DQ2 * f1));
    X[1] = __builtin_fmaf(DQ7, b3, __builtin_fmaf(DQ5, b2, __builtin_fmaf(DQ3, b1, DQ1 * b0)));
    X[3] = __builtin_fmaf(-DQ5, b3, __builtin_fmaf(-DQ1, b2, __builtin_fmaf(-DQ7, b1, DQ3 * b0)));
    X[5] = __builtin_fmaf(DQ3, b3, __builtin_fmaf(DQ7, b2, __builtin_fmaf(-DQ1, b1, DQ5 * b0)));
    X[7] = __builtin_fmaf(-DQ1, b3, __builtin_fmaf(DQ3, b2, __builtin_fmaf(-DQ5, b1, DQ7 * b0)));
}

__device__ __forceinline__ void idct8(const float (&X)[8], float (&x)[8]) {
    const float e0 = (X[0] + X[4]) * DQ0, e1 = (X[0] - X[4]) * DQ0;
    const float f0 = __builtin_fmaf(DQ2, X[2], DQ6 * X[6]);
    const float f1 = __builtin_fmaf(DQ6, X[2], -(DQ2 * X[6]));
    const float a0 = e0 + f0, a3 = e0 - f0, a1 = e1 + f1, a2 = e1 - f1;
    const float b0 = __builtin_fmaf(DQ7, X[7], __builtin_fmaf(DQ5, X[5], __builtin_fmaf(DQ3, X[3], DQ1 * X[1])));
    const float b1 = __builtin_fmaf(-DQ5, X[7], __builtin_fmaf(-DQ1, X[5], __builtin_fmaf(-DQ7, X[3], DQ3 * X[1])));
    const float b2 = __builtin_fmaf(DQ3, X[7], __builtin_fmaf(DQ7, X[5], __builtin_fmaf(-DQ1, X[3], DQ5 * X[1])));
    const float b3 = __builtin_fmaf(-DQ1, X[7], __builtin_fmaf(DQ3, X[5], __builtin_fmaf(-DQ5, X[3], DQ7 * X[1])));
    x[0] = a0 + b0;
    x[7] = a0 - b0;
    x[1] = a1 + b1;
    x[6] = a1 - b1;
    x[2] = a2 + b2;
    x[5] = a2 - b2;
    x[3] = a3 + b3;
    x[4] = a3 - b3;
}

// Separable 2D DCT of the N member patches a wave holds (lane p = (i, j) holds
// pixel p of every member), eight members at a time through a per-wave LDS
// scratch of 8 patches (patch pitch 72 floats). A pass gives lane l = 8 r + q the whole line q (a row, or a
// column stored transposed) of patch r as two 16-byte reads and transforms it
// in registers; the results go back as one scalar write per coefficient.
// Forward: rows, then columns. Inverse: columns, then rows.
constexpr int SCR = 8;   // members per LDS scratch chunk
constexpr int SPP = 72;  // scratch floats per patch
// Offset of line q (a row, or a column stored transposed) inside a patch:
// lines 4..7 sit 4 floats further on, so the eight 16-byte line reads of a
// patch spread over more banks (measured: group kernels -1.4 %).
__device__ __forceinline__ int lo8(int q) { return q * 8 + (q >= 4 ? 4 : 0); }

__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void load_line(const float* p, float (&x)[8]) {
    const float4 a = *reinterpret_cast<const float4*>(p);
    const float4 b = *reinterpret_cast<const float4*>(p + 4);
    x[0] = a.x, x[1] = a.y, x[2] = a.z, x[3] = a.w;
    x[4] = b.x, x[5] = b.y, x[6] = b.z, x[7] = b.w;
}

// Forward, from the rows a lane loaded (load_rows): v[8 c + k] is element k
// of row q of member 8 c + r for lane l = 8 r + q, so the row pass runs in
// registers and only the column pass goes through the scratch. Returns the
// coefficients in the pixel layout (lane (u, v) holds t[u][v] of every
// member) in v[0 .. N).
template <int N>
__device__ __forceinline__ void dct_fwd_rows(float (&v)[NMAX], float* sb, int lane) {
    const int r = lane >> 3, q = lane & 7;
    const int pix = lo8(r) + q;  // this lane's coefficient (u, v) = (r, q) at [u][v]
    float* line = sb + r * SPP + lo8(q);
#pragma unroll
    for (int c0 = 0; c0 < N; c0 += SCR) {
        constexpr int CH = N < SCR ? N : SCR;
        float x[8], X[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) x[k] = v[c0 + k];  // row q of member c0 + r
        fdct8(x, X);
#pragma unroll
        for (int k = 0; k < 8; ++k) sb[r * SPP + lo8(k) + q] = X[k];  // stored [v][i]
        wave_lds_fence();
        load_line(line, x);  // column q
        fdct8(x, X);
#pragma unroll
        for (int k = 0; k < 8; ++k) sb[r * SPP + lo8(k) + q] = X[k];  // t[u][v = q]
        wave_lds_fence();
#pragma unroll
        for (int m = 0; m < CH; ++m) v[c0 + m] = sb[m * SPP + pix];  // lane = (u, v)
        wave_lds_fence();
    }
}

template <int N>
__device__ __forceinline__ void dct_inv_n(float (&v)[NMAX], float* sb, int lane, int li, int lj) {
    const int r = lane >> 3, q = lane & 7;
    const int pix = lo8(li) + lj, tpix = lo8(lj) + li;
    float* line = sb + r * SPP + lo8(q);
#pragma unroll
    for (int c0 = 0; c0 < N; c0 += SCR) {
        constexpr int CH = N < SCR ? N : SCR;
#pragma unroll
        for (int m = 0; m < CH; ++m) sb[m * SPP + tpix] = v[c0 + m];  // t stored [v][u]
        wave_lds_fence();
        float x[8], X[8];
        load_line(line, X);  // column q of patch r
        idct8(X, x);
#pragma unroll
        for (int k = 0; k < 8; ++k) sb[r * SPP + lo8(k) + q] = x[k];  // y[i][v = q]
        wave_lds_fence();
        load_line(line, X);  // row q
        idct8(X, x);
        *reinterpret_cast<float4*>(line) = make_float4(x[0], x[1], x[2], x[3]);
        *reinterpret_cast<float4*>(line + 4) = make_float4(x[4], x[5], x[6], x[7]);
        wave_lds_fence();
#pragma unroll
        for (int m = 0; m < CH; ++m) v[c0 + m] = sb[m * SPP + pix];  // x[i][j]
        wave_lds_fence();
    }
}

// Multi-level Haar along the stack without its 1/sqrt(2) factors; output
// [approx, d_L, .., d_1]. Coefficient i of an N = 2^L stack finished at level
// haar_level<N>(i) (L for i = 0, 1; L - floor(log2 i) otherwise) and is the
// orthonormal coefficient divided by 2^(-level/2); the factors are carried by
// powers of two in the shrinkage instead (thresholds fl(thr fl(2^(level/2))),
// Wiener noise levels sigma^2 2^level, a scale 2^-level on kept coefficients),
// as in the CPU FP32 restatement.
template <int N>
constexpr int haar_level(int i) {
    int L = 0;
    while ((1 << L) < N) ++L;
    if (i == 0) return L;
    int f = 0;
    while ((2 << f) <= i) ++f;
    return L - f;
}

constexpr float HAAR_ROOT2[5] = {1.0f, 0x1.6a09e6p+0f, 2.0f, 0x1.6a09e6p+1f, 4.0f};  // fl(2^(l/2))
constexpr float HAAR_INV_POW2[5] = {1.0f, 0.5f, 0.25f, 0.125f, 0.0625f};             // 2^-l

template <int N>
__device__ __forceinline__ void haar_fwd_n(float (&v)[NMAX]) {
#pragma unroll
    for (int len = N; len > 1; len >>= 1) {
        float tmp[NMAX];
        const int hh = len >> 1;
#pragma unroll
        for (int k = 0; k < hh; ++k) {
            const float a = v[2 * k], b = v[2 * k + 1];
            tmp[k] = a + b;
            tmp[hh + k] = a - b;
        }
#pragma unroll
        for (int k = 0; k < len; ++k) v[k] = tmp[k];
    }
}

template <int N>
__device__ __forceinline__ void haar_inv_n(float (&v)[NMAX]) {
#pragma unroll
    for (int len = 2; len <= N; len <<= 1) {
        float tmp[NMAX];
        const int hh = len >> 1;
#pragma unroll
        for (int k = 0; k < hh; ++k) {
            const float a = v[k], d = v[hh + k];
            tmp[2 * k] = a + d;
            tmp[2 * k + 1] = a - d;
        }
#pragma unroll
        for (int k = 0; k < len; ++k) v[k] = tmp[k];
    }
}

// Reciprocal of d > 0 (normal, below 2^126): integer seed (relative error
// < 0.051) and three Newton steps, within 0.5 ulp of 1/d over 1e-20..1e19.
// Replaces the IEEE division in the Wiener gain (same operations as the CPU
// FP32 restatement).
__device__ __forceinline__ float rcp_nr(float d) {
    float r = __uint_as_float(0x7EF311C3u - __float_as_uint(d));
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float e = __builtin_fmaf(-d, r, 1.0f);
        r = __builtin_fmaf(r, e, r);
    }
    return r;
}

// Issues the loads of a group's members as rows: lane l = 8 r + q loads row q
// of member r into v[0 .. 8) and, for n = 16, of member 8 + r into v[8 .. 16)
// (two 16-byte buffer loads per row; and the basic estimate's into b for
// step 2). Lanes past n load the patch their list slot names (a valid
// position, or 0 past the list's count); their rows are never read back.
// mr0 / mr1: the list entries of members r and 8 + r.
template <int STEP>
__device__ __forceinline__ void load_rows(float (&v)[NMAX], float (&b)[NMAX], uint32_t mr0, uint32_t mr1, int n,
                                          __amdgpu_buffer_rsrc_t src, __amdgpu_buffer_rsrc_t bsrc, int w,
                                          int q) {
    auto row = [&](float(&d)[NMAX], int o, __amdgpu_buffer_rsrc_t rs, uint32_t mpos) {
        const uint32_t off = (((mpos >> 16) + (uint32_t)q) * (uint32_t)w + (mpos & 0xffffu)) * (uint32_t)sizeof(float);
        const uint4 a = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
        const uint4 c = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, off + 16, 0, 0));
        d[o + 0] = __uint_as_float(a.x), d[o + 1] = __uint_as_float(a.y);
        d[o + 2] = __uint_as_float(a.z), d[o + 3] = __uint_as_float(a.w);
        d[o + 4] = __uint_as_float(c.x), d[o + 5] = __uint_as_float(c.y);
        d[o + 6] = __uint_as_float(c.z), d[o + 7] = __uint_as_float(c.w);
    };
    row(v, 0, src, mr0);
    if (STEP == 2) row(b, 0, bsrc, mr0);
    if (n == NMAX) {
        row(v, 8, src, mr1);
        if (STEP == 2) row(b, 8, bsrc, mr1);
    }
}

// Wiener shrinkage of a transformed stack v with the basic estimate's b:
// W = B^2 / (B^2 + sigma^2) on the orthonormal coefficients: Wl = W 2^-l
// = bb / fma(bb, 2^l, sigma^2 4^l), so Wl v is the kept coefficient scaled
// for the inverse. Returns this lane's sum W^2 / 4^l over the stack, kept at
// the current level's scale (times 4 when the level drops; the caller scales
// it back after the lane tree). GUARD: a zero denominator (sigma = 0) keeps
// the coefficient; without it the caller has sigma^2 > 2^-100, so every
// denominator is at least sigma^2 4^l > 2^-100 and the test always passes.
template <int N, bool GUARD>
__device__ __forceinline__ float wiener_shrink(float (&v)[NMAX], const float (&b)[NMAX], float s2) {
    float aw = 0.0f;
#pragma unroll
    for (int m = 0; m < N; ++m) {
        const int l = haar_level<N>(m);
        if (m > 0 && l != haar_level<N>(m - 1)) aw = aw * 4.0f;
        const float bb = b[m] * b[m];
        const float d = __builtin_fmaf(bb, (float)(1 << l), s2 * (float)(1 << (2 * l)));
        const float W = (!GUARD || d > 0x1p-100f) ? bb * rcp_nr(d) : HAAR_INV_POW2[l];
        v[m] = W * v[m];
        aw = (m == 0) ? W * W : __builtin_fmaf(W, W, aw);
    }
    return aw;
}

// Filters one group in place from its loaded members (load_rows): 3D
// transform, shrinkage, inverse; returns the group's aggregation weight
// (1/N_nonzero for the hard threshold, 1/sum W^2 for Wiener).
template <int STEP, int N>
__device__ __forceinline__ float filter_group(float (&v)[NMAX], float (&b)[NMAX], float* sb, int lane,
                                              int li, int lj, float thr, float s2) {
    dct_fwd_rows<N>(v, sb, lane);
    haar_fwd_n<N>(v);
    float wgt;
    if (STEP == 1) {
        int nnz = 0;  // kept coefficients of the wave, counted per member by ballot
#pragma unroll
        for (int m = 0; m < N; ++m) {
            const int l = haar_level<N>(m);
            const float t = l == 0 ? thr : thr * HAAR_ROOT2[l];
            const bool keep = fabsf(v[m]) > t;
            nnz += __popcll(__ballot(keep));
            v[m] = keep ? v[m] * HAAR_INV_POW2[l] : 0.0f;
        }
        wgt = nnz > 0 ? 1.0f / (float)nnz : 1.0f;
    } else {
        dct_fwd_rows<N>(b, sb, lane);
        haar_fwd_n<N>(b);
        float aw = s2 > 0x1p-100f ? wiener_shrink<N, false>(v, b, s2) : wiener_shrink<N, true>(v, b, s2);
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) aw = aw + __shfl_xor(aw, off);
        aw = aw * (float)(1 << (2 * haar_level<N>(N - 1)));
        wgt = aw > 0.0f ? 1.0f / aw : 1.0f;
    }
    haar_inv_n<N>(v);
    dct_inv_n<N>(v, sb, lane, li, lj);
    return wgt;
}

// Fetches group g of a tile of 16 x TY refs (rows of nbx refs): its match list
// through the scalar cache (wave-uniform address) and, member m in lane m mod
// 16, as a vector (pv), size n, the sigma of the noise-profile map (MAP), and
// issues its member row loads (load_rows).
template <int STEP, int TY, bool MAP>
__device__ __forceinline__ void fetch_group(const GroupView& p, const float* __restrict__ smap, int tx,
                                            int ty, int nbx, int g, __amdgpu_buffer_rsrc_t src,
                                            __amdgpu_buffer_rsrc_t bsrc, int w, int lane,
                                            int& n, uint32_t (&mp)[NMAX], uint32_t& pv, float (&v)[NMAX],
                                            float (&b)[NMAX], float& gthr, float& gs2) {
    const int bx = tx * TILE + g % nbx, by = ty * TY + g / nbx;
    const size_t ref = (size_t)by * p.nx + bx;
    n = floor_pow2(load_count(p.cnt, ref));
    const uint32_t CONST_AS* lp = (const uint32_t CONST_AS*)(p.pos + ref * NMAX);
#pragma unroll
    for (int m = 0; m < NMAX; ++m) mp[m] = lp[m];
    pv = p.pos[ref * NMAX + (lane & 15)];  // lane m (mod 16): member m, for the run plan
    const uint32_t mr0 = p.pos[ref * NMAX + (lane >> 3)], mr1 = p.pos[ref * NMAX + 8 + (lane >> 3)];
    if (MAP) {
        const float sg = smap[ref];
        gthr = 2.7f * sg;
        gs2 = sg * sg;
    }
    load_rows<STEP>(v, b, mr0, mr1, n, src, bsrc, w, lane & 7);
}

// Tiles of 16 x TY refs, NW waves per workgroup (one group per wave per round).
// Exact FP32 plans aggregate over 16 x 4-ref tiles with 4-wave workgroups:
// each workgroup's ordered aggregation is one chain, and three workgroups
// (40 KB footprint + 9 KB scratch each) keep three chains per CU.
constexpr int EXACT_TY = 4, EXACT_NW = 4;

template <int STEP, int TY, int NW, bool MAP>
__global__ __launch_bounds__(64 * NW) void k_group(GroupLaunch L) {
    constexpr int DOFS = FOOT * foot_edge(TY);
    __shared__ float2 acc[DOFS];  // (num, den) per footprint pixel
    __shared__ __align__(16) float scratch[NW][SCR * SPP];
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
    const int li = lane >> 3, lj = lane & 7;
    float* sb = scratch[wave];
    const float kw = c_kaiser2[lane];
    const int w = p.w;
    const uint32_t plane_bytes = (uint32_t)((size_t)w * p.h * sizeof(float));
    const __amdgpu_buffer_rsrc_t src = plane_rsrc(F.z[c], plane_bytes);
    const __amdgpu_buffer_rsrc_t bsrc = plane_rsrc(F.bz[c], plane_bytes);
    const float sc = F.sig[c];
    const float thr = 2.7f * sc;
    const float s2 = sc * sc;
    // noise-profile map (MAP): each group shrinks with its reference patch's sigma
    // and weighs its aggregation by 1/sigma^2 (which a single sigma cancels)
    const float* __restrict__ smap = MAP ? F.smap + (size_t)c * p.nx * p.ny : nullptr;
    const int nbx = min(TILE, p.nx - tx * TILE), nby = min(TY, p.ny - ty * TY);
    const int G = nbx * nby;
    const int lo = (li - fy0) * fw + (lj - fx0);

    // Group g of this wave: its match list through the scalar cache
    // (wave-uniform address), size, sigma, and its member loads issued
    // (fetch_group). The next group is fetched before the current group's
    // turn, so its loads are in flight while the wave waits for the token and
    // adds.
    int n = 0;
    uint32_t mp[NMAX], pv = 0;
    float v[NMAX], b[NMAX];
    float gthr = thr, gs2 = s2;
    if (wave < G) fetch_group<STEP, TY, MAP>(p, smap, tx, ty, nbx, wave, src, bsrc, w, lane, n, mp, pv, v, b, gthr, gs2);
    for (int rnd = 0; rnd < (G + NW - 1) / NW; ++rnd) {
        const int g = rnd * NW + wave;
        const bool has = g < G;
        float wgt = 0.0f;
        if (has) {
            switch (n) {
                case 16: wgt = filter_group<STEP, 16>(v, b, sb, lane, li, lj, gthr, gs2); break;
                case 8: wgt = filter_group<STEP, 8>(v, b, sb, lane, li, lj, gthr, gs2); break;
                case 4: wgt = filter_group<STEP, 4>(v, b, sb, lane, li, lj, gthr, gs2); break;
                case 2: wgt = filter_group<STEP, 2>(v, b, sb, lane, li, lj, gthr, gs2); break;
                default: wgt = filter_group<STEP, 1>(v, b, sb, lane, li, lj, gthr, gs2); break;
            }
            if (MAP) wgt = wgt * (1.0f / (gs2 > 1e-30f ? gs2 : 1.0f));
        }
        PreparedGroup P;
        prepare_group(P, has, n, mp, pv, v, wgt * kw, lo, fw);
        if (g + NW < G) fetch_group<STEP, TY, MAP>(p, smap, tx, ty, nbx, g + NW, src, bsrc, w, lane, n, mp, pv, v, b, gthr, gs2);
        timed_out |= take_turn(acc, &turn, g, P);
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

// ---------------------------------------------------------------------------
// Tile merge + normalisation. Every pixel sums the partial num/den of the
// tiles whose footprint contains it, in (ty, tx) ascending order.
// Step 1 writes the basic estimate (opponent space); step 2 also applies the
// inverse colour transform and writes the caller's output planes.
// ---------------------------------------------------------------------------

// Destinations of a merge: the step-1 basic estimate (opponent planes), the
// plain step-1 planes on spot-checked rectangles, or the caller's output
// (inverse colour transform for 3 channels).
enum MergeDest { TO_BASIC = 0, TO_CHECK = 1, TO_OUTPUT = 2 };

// One block row per pixel row of a frame rectangle (blockIdx.z = rectangle).
__global__ void k_merge(const FrameTab* __restrict__ tab, const FrameRect* __restrict__ rects,
                        int dest, int use_gate) {
    const FrameRect rc = rects[blockIdx.z];
    const FrameTab& F = tab[rc.f];
    if (use_gate && *F.gate == 0) return;
    const int x = rc.x0 + blockIdx.x * blockDim.x + threadIdx.x;
    const int y = rc.y0 + blockIdx.y;
    if (x >= rc.x1 || y >= rc.y1) return;
    const int txl = F.txlo[x], txh = F.txhi[x], tyl = F.tylo[y], tyh = F.tyhi[y];
    const size_t ntiles = (size_t)F.ntx * F.nty;
    // a pixel lies in at most MTY x MTX tile footprints (16-ref tiles along x:
    // 3; 4-ref tiles along y: 5); the loads of a channel are all issued before
    // the sums, which still run in (ty, tx) ascending order
    constexpr int MTX = 3, MTY = 5;
    size_t offs[MTY * MTX];
    bool live[MTY * MTX];
#pragma unroll
    for (int i = 0; i < MTY; ++i)
#pragma unroll
        for (int k = 0; k < MTX; ++k) {
            const int ty = tyl + i, tx = txl + k;
            live[i * MTX + k] = ty <= tyh && tx <= txh;
            const int tyc = min(ty, tyh), txc = min(tx, txh);
            const int fw = F.fx1[txc] - F.fx0[txc];
            offs[i * MTX + k] = ((size_t)tyc * F.ntx + txc) * F.tile_stride +
                                (size_t)((y - F.fy0[tyc]) * fw + (x - F.fx0[txc]));
        }
    float o[3] = {0.0f, 0.0f, 0.0f};
    for (int c = 0; c < F.c; ++c) {
        const float* pc = F.partial + (size_t)c * ntiles * F.tile_stride;
        float vn[MTY * MTX], vd[MTY * MTX];
#pragma unroll
        for (int i = 0; i < MTY * MTX; ++i) {
            vn[i] = live[i] ? pc[offs[i]] : 0.0f;
            vd[i] = live[i] ? pc[offs[i] + F.dofs] : 0.0f;
        }
        float num = 0.0f, den = 0.0f;
#pragma unroll
        for (int i = 0; i < MTY * MTX; ++i) {
            if (live[i]) {
                num = num + vn[i];
                den = den + vd[i];
            }
        }
        o[c] = num / den;
    }
    if (dest == TO_OUTPUT) {
        const size_t oi = (size_t)y * F.out_pitch + x;
        if (F.c == 3) {
            const float c23 = (float)(2.0 / 3.0), c43 = (float)(4.0 / 3.0);
            F.out[0][oi] = (o[0] + o[1]) + o[2] * c23;
            F.out[1][oi] = o[0] - o[2] * c43;
            F.out[2][oi] = (o[0] - o[1]) + o[2] * c23;
        } else {
            F.out[0][oi] = o[0];
        }
    } else {
        float* const* d = dest == TO_BASIC ? F.bz : F.bzp;
        const size_t oi = (size_t)y * F.w + x;
        for (int c = 0; c < F.c; ++c) d[c][oi] = o[c];
    }
}

// ---------------------------------------------------------------------------
// Learned-pilot spot checks. For each checked rectangle the plain step-1
// estimate (bzp) is computed exactly; the method-noise ratio
// delta = |10 log10(|z - pilot|^2 / |z - plain|^2)| (three channels) decides
// whether the rectangle falls back to the plain estimate. If more than a set
// fraction of the checked rectangles fall back, the gate enables a full plain
// step 1.
// ---------------------------------------------------------------------------

// The network's output on a checked rectangle comes from its own run on a
// region of P x P pixels around it, P = block edge + 2 PILOT_HALO: the six 3x3
// layers reach 6 pixels, so inside the rectangle it is bit-identical to the
// whole-frame run, and the checks can decide while that run is still going.
// The region keeps PILOT_HALO pixels on each side of the rectangle, shifted
// inwards at the frame's edges (there the frame's own padding applies), so all
// checks of a frame share one region size and run as one stack of images.
// Each check's slot of `io` holds the region's 3 input and 3 output planes of
// pitch P: [check][6][P * P].
constexpr int PILOT_HALO = 6;

// Region origin along one axis: rectangle start v0, frame extent n.
__host__ __device__ inline int pilot_region_lo(int v0, int n, int P) {
    const int lo = v0 - PILOT_HALO < n - P ? v0 - PILOT_HALO : n - P;
    return lo > 0 ? lo : 0;
}

// The regions of all checks into their input slots, and each check's noise
// level from its frame's (block (row, check)).
__global__ __launch_bounds__(256) void k_pilot_gather(const FrameTab* __restrict__ tab,
                                                      const FrameRect* __restrict__ rects,
                                                      const float* __restrict__ frame_levels,
                                                      float* __restrict__ io, float* __restrict__ levels,
                                                      int P) {
    const FrameRect r = rects[blockIdx.y];
    const FrameTab& F = tab[r.f];
    if (blockIdx.x == 0 && threadIdx.x == 0) levels[blockIdx.y] = frame_levels[r.f];
    const int hx0 = pilot_region_lo(r.x0, F.w, P), hy0 = pilot_region_lo(r.y0, F.h, P);
    const int hw = min(P, F.w), hh = min(P, F.h);
    if ((int)blockIdx.x >= hh) return;
    const int y = hy0 + blockIdx.x;
    float* slot = io + (size_t)blockIdx.y * 6 * P * P + (size_t)blockIdx.x * P;
    for (int x = threadIdx.x; x < hw; x += blockDim.x) {
#pragma unroll
        for (int c = 0; c < 3; ++c) slot[(size_t)c * P * P + x] = F.z[c][(size_t)y * F.w + hx0 + x];
    }
}

// Clears every frame's gate before the network and the checks start.
__global__ void k_clear_gates(const FrameTab* __restrict__ tab, int n_frames) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < n_frames) __hip_atomic_store(tab[f].gate, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Three launches over all checked rectangles of all frames: per-row squared
// residual sums (block (row, rectangle)), per-rectangle totals summed over the
// rows in a fixed order (one block per rectangle; report entry blockIdx.x),
// and the fallback copy of the rectangles that fail.
__global__ __launch_bounds__(256) void k_pilot_check_rows(const FrameTab* __restrict__ tab,
                                                          const FrameRect* __restrict__ rects,
                                                          const float* __restrict__ io, int P,
                                                          float2* __restrict__ part, int rows) {
    __shared__ float ra[256], rb[256];
    const FrameRect r = rects[blockIdx.y];
    const int y = r.y0 + blockIdx.x;
    if (y >= r.y1) return;  // uniform per block
    const FrameTab& F = tab[r.f];
    // the block network's output planes (slot planes 3..5)
    const float* pil = io + ((size_t)blockIdx.y * 6 + 3) * P * P +
                       (size_t)(y - pilot_region_lo(r.y0, F.h, P)) * P - pilot_region_lo(r.x0, F.w, P);
    float a = 0.0f, b = 0.0f;
    for (int x = r.x0 + threadIdx.x; x < r.x1; x += blockDim.x) {
        const size_t o = (size_t)y * F.w + x;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float zs = F.z[c][o];
            const float dl = zs - pil[(size_t)c * P * P + x], dp = zs - F.bzp[c][o];
            a = __builtin_fmaf(dl, dl, a);
            b = __builtin_fmaf(dp, dp, b);
        }
    }
    ra[threadIdx.x] = a;
    rb[threadIdx.x] = b;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            ra[threadIdx.x] += ra[threadIdx.x + off];
            rb[threadIdx.x] += rb[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.y * rows + blockIdx.x] = make_float2(ra[0], rb[0]);
}

__global__ __launch_bounds__(256) void k_pilot_check_sum(const FrameRect* __restrict__ rects,
                                                         const float2* __restrict__ part, int rows,
                                                         float thr_db, float under_db,
                                                         float* __restrict__ report) {
    __shared__ float ra[256], rb[256];
    const FrameRect r = rects[blockIdx.x];
    const int nrow = r.y1 - r.y0;
    float a = 0.0f, b = 0.0f;
    for (int i = threadIdx.x; i < nrow; i += blockDim.x) {
        const float2 v = part[(size_t)blockIdx.x * rows + i];
        a += v.x;
        b += v.y;
    }
    ra[threadIdx.x] = a;
    rb[threadIdx.x] = b;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            ra[threadIdx.x] += ra[threadIdx.x + off];
            rb[threadIdx.x] += rb[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float sa = ra[0], sb = rb[0];
        const float r = (sa > 0.0f && sb > 0.0f) ? 10.0f * log10f(sa / sb)
                                                  : (sa == sb ? 0.0f : (sa > sb ? 1e30f : -1e30f));
        report[2 * blockIdx.x] = r;
        report[2 * blockIdx.x + 1] = (r > thr_db || r < -under_db) ? 1.0f : 0.0f;
    }
}

__global__ __launch_bounds__(256) void k_pilot_check_apply(const FrameTab* __restrict__ tab,
                                                           const FrameRect* __restrict__ rects,
                                                           const float* __restrict__ report) {
    if (report[2 * blockIdx.y + 1] < 0.5f) return;  // this rectangle keeps the pilot
    const FrameRect r = rects[blockIdx.y];
    const int y = r.y0 + blockIdx.x;
    if (y >= r.y1) return;
    const FrameTab& F = tab[r.f];
    for (int x = r.x0 + threadIdx.x; x < r.x1; x += blockDim.x) {
        const size_t o = (size_t)y * F.w + x;
#pragma unroll
        for (int c = 0; c < 3; ++c) F.bz[c][o] = F.bzp[c][o];
    }
}

// Per frame: the gate runs the plain step 1 on the whole frame when more than
// `frac` of its checked rectangles fell back.
__global__ void k_pilot_decide(const FrameTab* __restrict__ tab, int n_frames,
                               const float* __restrict__ report, float frac) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_frames) return;
    const FrameTab& F = tab[f];
    int falls = 0;
    for (int i = F.chk0; i < F.chk0 + F.nchk; ++i) falls += report[2 * i + 1] > 0.5f ? 1 : 0;
    // agent scope: the frame's network, running on another stream, polls it
    __hip_atomic_store(F.gate, ((float)falls > frac * (float)F.nchk) ? 1 : 0, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

std::mutex g_tables_mu;
bool g_tables_ready[64] = {};

int ensure_tables(int device) {
    std::lock_guard<std::mutex> lk(g_tables_mu);
    if (device < 0 || device >= 64) return fail(BAD_INDEX, "device ordinal out of range");
    if (g_tables_ready[device]) return OK;
    double c[64], k1[8];
    for (int u = 0; u < 8; ++u) {
        const double a = (u == 0) ? std::sqrt(1.0 / 8.0) : std::sqrt(2.0 / 8.0);
        for (int i = 0; i < 8; ++i) c[u * 8 + i] = a * std::cos((2.0 * i + 1.0) * u * M_PI / 16.0);
    }
    auto i0 = [](double x) {
        double sum = 1.0, term = 1.0;
        for (int k = 1; k < 200; ++k) {
            term *= (x / (2.0 * k));
            const double t2 = term * term;
            sum += t2;
            if (t2 < 1e-18 * sum) break;
        }
        return sum;
    };
    for (int n = 0; n < 8; ++n) {
        const double r = 2.0 * n / 7.0 - 1.0;
        k1[n] = i0(2.0 * std::sqrt(1.0 - r * r)) / i0(2.0);
    }
    float cf[64], kf[64];
    for (int i = 0; i < 64; ++i) cf[i] = (float)c[i];
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) kf[i * 8 + j] = (float)(k1[i] * k1[j]);
    HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_kaiser2), kf, sizeof(kf)));
    HIP_CHECK((hipError_t)group_mx_tables(device, cf, kf));
    g_tables_ready[device] = true;
    return OK;
}

class DeviceGuard {
public:
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev_) != hipSuccess) prev_ = -1;
        if (prev_ != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        if (prev_ >= 0) (void)hipSetDevice(prev_);
    }

private:
    int prev_ = -1;
};

}  // namespace

// A plan owns one or more frames (a batch). Every stage is one launch over a
// work list that spans all frames: block-matching tiles (frame, tx, ty),
// aggregation tiles (frame, tile) and pixel-row bands (frame rectangles); the
// kernels read each frame's geometry, buffers, sigma and I/O planes from the
// device frame table. A single-image plan is a batch of one frame.
struct cpipe_bm3d_plan_s {
    struct Frame {
        int w = 0, h = 0, nx = 0, ny = 0, ntx = 0, nty = 0;
        size_t npx = 0;
        bool small = false;         // smaller than one patch: copied through
        std::vector<int> geo;       // txlo[w] txhi[w] tylo[h] tyhi[h] fx0[ntx] fx1[ntx] fy0[nty] fy1[nty]
        float* z = nullptr;         // c planes each
        float* bz = nullptr;
        float* bzp = nullptr;       // learned pilot: plain step-1 planes on checked rectangles
        uint32_t* pos[2] = {nullptr, nullptr};
        uint8_t* cnt[2] = {nullptr, nullptr};
        float* partial = nullptr;
        int* geo_dev = nullptr;
    };
    struct Check {
        int f;
        int btx0, bty0, btx1, bty1;  // block-matching tiles
        int tx0, ty0, tx1, ty1;      // aggregation tiles
        int x0, y0, x1, y1;          // pixels
        int hx0, hy0, hw, hh;        // the block network's region: the pixels + the receptive half-width
    };
    struct Lists {  // one stage's work over a set of frames / rectangles
        int4* bm = nullptr;
        int2* grp = nullptr;
        FrameRect* rects = nullptr;
        int n_bm = 0, n_grp = 0, n_rects = 0;
        int rect_w = 0, rect_h = 0;  // largest rectangle
    };

    int c = 0, precision = 0, device = 0;
    int tile_y = TILE / 2;           // refs per tile column (4 exact FP32, 8 or 16 mixed precision)
    int dofs = FOOT2;                // den offset in a partial block
    size_t tile_stride = 2 * (size_t)FOOT2;
    int order = 0;                   // work-list order: 0 frame raster, 1 XCD-contiguous chunks
    int band_rows = 32;              // rows per colour / merge rectangle
    int bm_mfma = 0;                 // MFMA cross-term block matching (mixed-precision default; else exact FP32)
    std::vector<Frame> fr;
    std::vector<FrameTab> tab;       // host image of the device frame table
    FrameTab* tab_dev = nullptr;     // followed by one float noise level per frame
    void* staging = nullptr;         // pinned host copy of the table, uploaded every run
    size_t staging_bytes = 0;
    hipEvent_t staged = nullptr;     // the last upload (or graph replay) has read `staging`
    bool staged_pending = false;
    int* gates = nullptr;            // per frame
    int* err = nullptr;              // device flag: an ordering wait timed out (cleared by every run)
    int* err_host = nullptr;         // pinned: the flag as the last run left it (copied at its end)
    hipEvent_t done = nullptr;       // recorded after the last run's flag copy
    Lists full;                      // every frame, whole image
    FrameRect* small_rects = nullptr;  // frames smaller than a patch (copied through)
    int n_small = 0, small_w = 0, small_h = 0;
    float* rgb = nullptr;            // 6 planes for the RGBA16F path (single-frame plans)
    // noise-profile sigma map: per frame, per (channel, ref) sigma
    bool profile = false;
    std::vector<float> pbeta, plambda;  // [frame][3]
    std::vector<size_t> smap_off;       // per frame (floats into smap)
    float* smap = nullptr;
    int smap_nx = 0, smap_ny = 0;       // largest reference grid (launch shape)
    bool ran = false;
    // timing (eager runs)
    bool timing = false;
    hipEvent_t ev[2 * CPIPE_BM3D_T_COUNT] = {};
    bool stage_ran[CPIPE_BM3D_T_COUNT] = {};
    float ms[CPIPE_BM3D_T_COUNT] = {};
    int runs = 0;
    // learned pilot (replaces step 1) and its spot checks
    cpipe_bm3d_pilot* pilot = nullptr;
    void* pilot_ws = nullptr;        // the network's activations for this plan's largest frame
    float pilot_thr_db = 0.5f;
    float pilot_under_db = 0.25f;
    float pilot_full_frac = 0.5f;
    cpipe_bm3d_pilot_opts pilot_opts{};
    std::vector<Check> checks;
    Lists chk;                       // the checked rectangles' work, all checks in one launch per kernel
    float* report = nullptr;         // [n][2]: delta dB, fell back
    float2* chk_part = nullptr;      // [n][rect rows]: per-row squared-residual sums of each check
    float* chk_io = nullptr;         // per check: 3 input + 3 output planes of its block network
    float* chk_levels = nullptr;     // per check: its frame's noise level (filled every run)
    void* chk_ws = nullptr;          // the block networks' activations (one block at a time)
    int chk_pitch = 0;               // row pitch (and edge) of a block network's planes
    hipStream_t side = nullptr;      // the spot checks' plain step 1 runs here, beside the network
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    // one graph per plan: the whole batch as a single submission
    bool use_graph = false;
    hipGraphExec_t exec = nullptr;
    bool dirty = true;               // work lists or pilot changed since the graph was captured
};

namespace {

void tile_box_host(int t, int tile, int nref, int n, int* f0, int* f1) {
    const int first = ref_pos(t * tile, n);
    const int lastref = std::min(t * tile + tile - 1, nref - 1);
    const int last = ref_pos(lastref, n);
    *f0 = std::max(0, first - R);
    *f1 = std::min(n, last + R + K);
}

void free_lists(cpipe_bm3d_plan_s::Lists& l) {
    (void)hipFree(l.bm);
    (void)hipFree(l.grp);
    (void)hipFree(l.rects);
    l = cpipe_bm3d_plan_s::Lists{};
}

void drop_graph(cpipe_bm3d_plan* p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    p->exec = nullptr;
    p->dirty = true;
}

void free_frame_buffers(cpipe_bm3d_plan_s::Frame& f) {
    (void)hipFree(f.z);
    (void)hipFree(f.bz);
    (void)hipFree(f.bzp);
    for (int s = 0; s < 2; ++s) {
        (void)hipFree(f.pos[s]);
        (void)hipFree(f.cnt[s]);
    }
    (void)hipFree(f.partial);
    (void)hipFree(f.geo_dev);
    f.z = f.bz = f.bzp = f.partial = nullptr;
    f.pos[0] = f.pos[1] = nullptr;
    f.cnt[0] = f.cnt[1] = nullptr;
    f.geo_dev = nullptr;
}

void free_plan(cpipe_bm3d_plan* p) {
    if (!p) return;
    DeviceGuard g(p->device);
    (void)hipDeviceSynchronize();  // runs may still read the scratch being freed
    drop_graph(p);
    for (auto& f : p->fr) free_frame_buffers(f);
    free_lists(p->full);
    free_lists(p->chk);
    (void)hipFree(p->small_rects);
    (void)hipFree(p->tab_dev);
    if (p->staging) (void)hipHostFree(p->staging);
    if (p->staged) (void)hipEventDestroy(p->staged);
    (void)hipFree(p->gates);
    (void)hipFree(p->err);
    if (p->err_host) (void)hipHostFree(p->err_host);
    if (p->done) (void)hipEventDestroy(p->done);
    (void)hipFree(p->rgb);
    (void)hipFree(p->smap);
    (void)hipFree(p->report);
    (void)hipFree(p->chk_part);
    (void)hipFree(p->chk_io);
    (void)hipFree(p->chk_levels);
    (void)hipFree(p->chk_ws);
    (void)hipFree(p->pilot_ws);
    if (p->side) (void)hipStreamDestroy(p->side);
    if (p->fork_ev) (void)hipEventDestroy(p->fork_ev);
    if (p->join_ev) (void)hipEventDestroy(p->join_ev);
    for (auto& e : p->ev)
        if (e) (void)hipEventDestroy(e);
    delete p;
}

int dev_alloc(void** ptr, size_t bytes) {
    hipError_t e = hipMalloc(ptr, std::max<size_t>(bytes, 4));
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? OOM : INTERNAL, std::string("hipMalloc: ") + hipGetErrorString(e));
    return OK;
}

template <typename T>
int upload(T** dst, const std::vector<T>& v) {
    int rc = dev_alloc((void**)dst, v.size() * sizeof(T));
    if (rc) return rc;
    if (!v.empty()) HIP_CHECK(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return OK;
}

// Work-list order. Workgroups are handed to the 8 XCDs round robin; order 1
// permutes a list so that each XCD receives one contiguous run of it
// (neighbouring tiles, whose footprints overlap, share that XCD's L2).
template <typename T>
std::vector<T> apply_order(const std::vector<T>& v, int order) {
    if (order != 1 || v.size() < 16) return v;
    const size_t n = v.size();
    constexpr size_t X = 8;
    size_t start[X], size[X], acc = 0;
    for (size_t x = 0; x < X; ++x) {
        size[x] = (n > x) ? (n - x + X - 1) / X : 0;
        start[x] = acc;
        acc += size[x];
    }
    std::vector<T> out(n);
    for (size_t i = 0; i < n; ++i) out[i] = v[start[i % X] + i / X];
    return out;
}

// Geometry, buffers and the device table entry of frame f for the plan's tile shape.
int build_frame(cpipe_bm3d_plan* p, int f) {
    auto& F = p->fr[f];
    free_frame_buffers(F);
    F.npx = (size_t)F.w * F.h;
    F.small = F.w < K || F.h < K;
    FrameTab& T = p->tab[f];
    T = FrameTab{};
    T.w = F.w;
    T.h = F.h;
    T.c = p->c;
    T.gate = p->gates + f;
    if (F.small) return OK;
    F.nx = ref_count(F.w);
    F.ny = ref_count(F.h);
    F.ntx = (F.nx + TILE - 1) / TILE;
    F.nty = (F.ny + p->tile_y - 1) / p->tile_y;
    const size_t nref = (size_t)F.nx * F.ny;
    const size_t ntiles = (size_t)F.ntx * F.nty;
    int rc;
    if ((rc = dev_alloc((void**)&F.z, sizeof(float) * F.npx * p->c)) ||
        (rc = dev_alloc((void**)&F.bz, sizeof(float) * F.npx * p->c)) ||
        (rc = dev_alloc((void**)&F.pos[0], sizeof(uint32_t) * nref * NMAX)) ||
        (rc = dev_alloc((void**)&F.pos[1], sizeof(uint32_t) * nref * NMAX)) ||
        // counts: whole dwords (the group kernels read a count through its aligned dword)
        (rc = dev_alloc((void**)&F.cnt[0], (nref + 3) & ~(size_t)3)) ||
        (rc = dev_alloc((void**)&F.cnt[1], (nref + 3) & ~(size_t)3)) ||
        (rc = dev_alloc((void**)&F.partial, sizeof(float) * p->tile_stride * ntiles * p->c)))
        return rc;
    const size_t ngeo = 2 * (size_t)F.w + 2 * (size_t)F.h + 2 * (size_t)F.ntx + 2 * (size_t)F.nty;
    F.geo.assign(ngeo, 0);
    int* txlo = F.geo.data();
    int* txhi = txlo + F.w;
    int* tylo = txhi + F.w;
    int* tyhi = tylo + F.h;
    int* fx0 = tyhi + F.h;
    int* fx1 = fx0 + F.ntx;
    int* fy0 = fx1 + F.ntx;
    int* fy1 = fy0 + F.nty;
    for (int t = 0; t < F.ntx; ++t) tile_box_host(t, TILE, F.nx, F.w, &fx0[t], &fx1[t]);
    for (int t = 0; t < F.nty; ++t) tile_box_host(t, p->tile_y, F.ny, F.h, &fy0[t], &fy1[t]);
    for (int x = 0; x < F.w; ++x) {
        txlo[x] = -1;
        txhi[x] = -2;
    }
    for (int t = 0; t < F.ntx; ++t)
        for (int x = fx0[t]; x < fx1[t]; ++x) {
            if (txlo[x] < 0) txlo[x] = t;
            txhi[x] = t;
        }
    for (int y = 0; y < F.h; ++y) {
        tylo[y] = -1;
        tyhi[y] = -2;
    }
    for (int t = 0; t < F.nty; ++t)
        for (int y = fy0[t]; y < fy1[t]; ++y) {
            if (tylo[y] < 0) tylo[y] = t;
            tyhi[y] = t;
        }
    // every pixel lies in 1..3 tile footprints along x and 1..5 along y (the
    // merge kernel reads at most 5 x 3)
    for (int x = 0; x < F.w; ++x)
        if (txlo[x] < 0 || txhi[x] - txlo[x] >= 3) return fail(INTERNAL, "tile cover in x out of range");
    for (int y = 0; y < F.h; ++y)
        if (tylo[y] < 0 || tyhi[y] - tylo[y] >= 5) return fail(INTERNAL, "tile cover in y out of range");
    if ((rc = upload(&F.geo_dev, F.geo))) return rc;
    T.nx = F.nx;
    T.ny = F.ny;
    T.ntx = F.ntx;
    T.nty = F.nty;
    T.tile_y = p->tile_y;
    T.dofs = p->dofs;
    T.tile_stride = p->tile_stride;
    for (int ch = 0; ch < 3; ++ch) {
        const int cc = ch < p->c ? ch : 0;
        T.z[ch] = F.z + cc * F.npx;
        T.bz[ch] = F.bz + cc * F.npx;
    }
    T.pos[0] = F.pos[0];
    T.pos[1] = F.pos[1];
    T.cnt[0] = F.cnt[0];
    T.cnt[1] = F.cnt[1];
    T.partial = F.partial;
    const int* g = F.geo_dev;
    T.txlo = g;
    T.txhi = g + F.w;
    T.tylo = T.txhi + F.w;
    T.tyhi = T.tylo + F.h;
    T.fx0 = T.tyhi + F.h;
    T.fx1 = T.fx0 + F.ntx;
    T.fy0 = T.fx1 + F.ntx;
    T.fy1 = T.fy0 + F.nty;
    return OK;
}

// Whole-frame work lists of every frame, in the plan's order.
int build_full_lists(cpipe_bm3d_plan* p) {
    free_lists(p->full);
    (void)hipFree(p->small_rects);
    p->small_rects = nullptr;
    std::vector<int4> bm;
    std::vector<int2> grp;
    std::vector<FrameRect> bands, small;
    int maxw = 0, sw = 0, sh = 0;
    for (int f = 0; f < (int)p->fr.size(); ++f) {
        const auto& F = p->fr[f];
        if (F.small) {
            small.push_back({f, 0, 0, F.w, F.h, 0, 0, 0});
            sw = std::max(sw, F.w);
            sh = std::max(sh, F.h);
            continue;
        }
        const int nbtx = (F.nx + BM_TX - 1) / BM_TX, nbty = (F.ny + BM_TY - 1) / BM_TY;
        for (int ty = 0; ty < nbty; ++ty)
            for (int tx = 0; tx < nbtx; ++tx) bm.push_back(make_int4(f, tx, ty, 0));
        for (int t = 0; t < F.ntx * F.nty; ++t) grp.push_back(make_int2(f, t));
        for (int y0 = 0; y0 < F.h; y0 += p->band_rows)
            bands.push_back({f, 0, y0, F.w, std::min(F.h, y0 + p->band_rows), 0, 0, 0});
        maxw = std::max(maxw, F.w);
    }
    int rc;
    auto& L = p->full;
    if ((rc = upload(&L.bm, apply_order(bm, p->order))) || (rc = upload(&L.grp, apply_order(grp, p->order))) ||
        (rc = upload(&L.rects, bands)) || (rc = upload(&p->small_rects, small)))
        return rc;
    L.n_bm = (int)bm.size();
    L.n_grp = (int)grp.size();
    L.n_rects = (int)bands.size();
    L.rect_w = maxw;
    L.rect_h = p->band_rows;
    p->n_small = (int)small.size();
    p->small_w = sw;
    p->small_h = sh;
    p->dirty = true;
    return OK;
}

// Tile shape: exact FP32 plans use 16 x 4-ref tiles (the CPU FP32 restatement
// aggregates over the same tiling): a 91 x 55 px (num, den) footprint, three
// 4-wave workgroups and so three ordered aggregation chains per CU. Mixed-
// precision plans start with 16 x 8-ref tiles (91 x 67 px, three 4-wave
// workgroups per CU) and may be retuned to 16 x 16.
int set_tile_shape(cpipe_bm3d_plan* p, int tile_y) {
    p->tile_y = tile_y;
    p->dofs = FOOT * foot_edge(tile_y);
    p->tile_stride = 2 * (size_t)p->dofs;
    for (int f = 0; f < (int)p->fr.size(); ++f) {
        const int rc = build_frame(p, f);
        if (rc) return rc;
    }
    return build_full_lists(p);
}

int create_core(cpipe_bm3d_plan** out, const int* widths, const int* heights, int n, int channels,
                int precision, int device) {
    *out = nullptr;
    if (n <= 0 || n > 4096) return fail(BAD_INDEX, "frame count out of range (1..4096)");
    for (int f = 0; f < n; ++f) {
        if (widths[f] <= 0 || heights[f] <= 0 || widths[f] > 65535 || heights[f] > 65535)
            return fail(BAD_INDEX, "width/height out of range (1..65535)");
        if ((size_t)widths[f] * heights[f] * sizeof(float) > 0xffffffffull)
            return fail(BAD_INDEX, "frame planes must stay below 4 GiB (width x height <= 2^30 pixels)");
    }
    if (channels != 1 && channels != 3) return fail(BAD_INDEX, "channels must be 1 or 3");
    if (precision != CPIPE_BM3D_PRECISION_FP32 && precision != CPIPE_BM3D_PRECISION_FP16 &&
        precision != CPIPE_BM3D_PRECISION_BF16)
        return fail(BAD_INDEX, "unknown precision");
    const int ndev = cpipe_bm3d_device_count();
    if (ndev <= 0) return fail(UNSUPPORTED, "no HIP device visible");
    if (device < 0 || device >= ndev) return fail(BAD_INDEX, "device ordinal out of range");
    DeviceGuard guard(device);
    int rc = ensure_tables(device);
    if (rc) return rc;
    auto* p = new cpipe_bm3d_plan_s();
    p->c = channels;
    p->precision = precision;
    p->bm_mfma = precision != CPIPE_BM3D_PRECISION_FP32;  // the faster kernel (lists as the exact kernel's for >= 99.5 % of refs)
    p->device = device;
    p->fr.resize(n);
    p->tab.resize(n);
    for (int f = 0; f < n; ++f) {
        p->fr[f].w = widths[f];
        p->fr[f].h = heights[f];
    }
    const size_t tab_bytes = sizeof(FrameTab) * n + sizeof(float) * n;
    p->staging_bytes = tab_bytes;
    hipError_t e = hipHostMalloc(&p->staging, tab_bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        free_plan(p);
        return fail(OOM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    }
    if ((rc = dev_alloc((void**)&p->tab_dev, tab_bytes)) || (rc = dev_alloc((void**)&p->gates, sizeof(int) * n)) ||
        (rc = dev_alloc((void**)&p->err, sizeof(int)))) {
        free_plan(p);
        return rc;
    }
    e = hipMemset(p->err, 0, sizeof(int));
    if (e == hipSuccess) e = hipMemset(p->gates, 0, sizeof(int) * n);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->staged, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipHostMalloc((void**)&p->err_host, sizeof(int), hipHostMallocDefault);
    if (e == hipSuccess) *p->err_host = 0;
    if (e != hipSuccess) {
        free_plan(p);
        return fail(INTERNAL, std::string("init: ") + hipGetErrorString(e));
    }
    rc = set_tile_shape(p, precision == CPIPE_BM3D_PRECISION_FP32 ? EXACT_TY : TILE / 2);
    if (rc) {
        free_plan(p);
        return rc;
    }
    *out = p;
    return OK;
}

float* levels_dev(cpipe_bm3d_plan* p) {
    return reinterpret_cast<float*>(p->tab_dev + p->fr.size());
}

// Per-run fields of frame f (sigma, noise level, I/O planes).
void set_run(cpipe_bm3d_plan* p, int f, const float* const* in, size_t in_pitch, float* const* out,
             size_t out_pitch, float sigma) {
    FrameTab& T = p->tab[f];
    T.sigma = sigma;
    if (p->c == 1) {
        T.sig[0] = T.sig[1] = T.sig[2] = sigma;
    } else {
        T.sig[0] = sigma * (float)std::sqrt(1.0 / 3.0);
        T.sig[1] = sigma * (float)std::sqrt(0.5);
        T.sig[2] = sigma * (float)std::sqrt(0.375);
    }
    for (int ch = 0; ch < 3; ++ch) {
        const int cc = ch < p->c ? ch : 0;
        T.in[ch] = in ? in[cc] : nullptr;
        T.out[ch] = out ? out[cc] : nullptr;
    }
    T.in_pitch = in_pitch;
    T.out_pitch = out_pitch;
    const bool map = p->profile && p->smap && !p->fr[f].small;
    T.smap = map ? p->smap + p->smap_off[f] : nullptr;
    for (int ch = 0; ch < 3; ++ch) {
        T.nbeta[ch] = map ? p->pbeta[3 * f + (ch < p->c ? ch : 0)] : 0.0f;
        T.nlambda[ch] = map ? p->plambda[3 * f + (ch < p->c ? ch : 0)] : 0.0f;
    }
}

// Copies the frame table into the pinned staging buffer once the previous
// upload has consumed it.
int fill_staging(cpipe_bm3d_plan* p) {
    if (p->staged_pending) {
        HIP_CHECK(hipEventSynchronize(p->staged));
        p->staged_pending = false;
    }
    const size_t n = p->fr.size();
    std::memcpy(p->staging, p->tab.data(), sizeof(FrameTab) * n);
    float* lv = reinterpret_cast<float*>(static_cast<char*>(p->staging) + sizeof(FrameTab) * n);
    for (size_t f = 0; f < n; ++f) lv[f] = 10.0f * p->tab[f].sigma;
    return OK;
}

struct RunCtx {
    cpipe_bm3d_plan* p;
    hipStream_t st;
    bool timed;
    int mark(int stage, bool end) {
        if (!timed) return OK;
        if (!end) p->stage_ran[stage] = true;
        HIP_CHECK(hipEventRecord(p->ev[2 * stage + (end ? 1 : 0)], st));
        return OK;
    }
};

// One BM3D step over a work list: block matching, group filtering and
// aggregation, merge into `dest`. With use_gate a frame runs only if its
// gate is set (decided on the device).
int run_step(RunCtx& r, int step, const cpipe_bm3d_plan_s::Lists& L, int dest, int use_gate, bool timed) {
    cpipe_bm3d_plan* p = r.p;
    const int sbm = step == 1 ? CPIPE_BM3D_T_BM1 : CPIPE_BM3D_T_BM2;
    const int sgr = step == 1 ? CPIPE_BM3D_T_GROUP1 : CPIPE_BM3D_T_GROUP2;
    const int smg = step == 1 ? CPIPE_BM3D_T_MERGE1 : CPIPE_BM3D_T_MERGE2;
    timed = timed && r.timed;
    RunCtx rt{p, r.st, timed};
    int rc;
    if ((rc = rt.mark(sbm, false))) return rc;
    if (L.n_bm > 0) {
        if (p->bm_mfma) {
            HIP_CHECK((hipError_t)bm_mx_launch(p->precision, p->tab_dev, L.bm, L.n_bm, step,
                                               tau_ssd(step), use_gate, r.st));
        } else {
            launch_block_match(L.n_bm, p->tab_dev, L.bm, step, use_gate, r.st);
        }
        HIP_CHECK(hipGetLastError());
    }
    if ((rc = rt.mark(sbm, true))) return rc;
    if ((rc = rt.mark(sgr, false))) return rc;
    if (L.n_grp > 0) {
        const GroupLaunch gl{p->tab_dev, L.grp, use_gate, p->err};
        const dim3 ggrid(L.n_grp, p->c);
        const bool map = p->profile && p->smap;  // kernels built with the per-ref sigma read
        if (p->precision != CPIPE_BM3D_PRECISION_FP32) {
            HIP_CHECK((hipError_t)group_mx_launch(step, p->precision, p->tile_y, map, gl, ggrid, r.st));
        } else if (p->tile_y != EXACT_TY) {
            return fail(INTERNAL, "exact FP32 plans run 16 x 4-ref tiles only");
        } else if (step == 1) {
            if (map)
                k_group<1, EXACT_TY, EXACT_NW, true><<<ggrid, 64 * EXACT_NW, 0, r.st>>>(gl);
            else
                k_group<1, EXACT_TY, EXACT_NW, false><<<ggrid, 64 * EXACT_NW, 0, r.st>>>(gl);
        } else {
            if (map)
                k_group<2, EXACT_TY, EXACT_NW, true><<<ggrid, 64 * EXACT_NW, 0, r.st>>>(gl);
            else
                k_group<2, EXACT_TY, EXACT_NW, false><<<ggrid, 64 * EXACT_NW, 0, r.st>>>(gl);
        }
        HIP_CHECK(hipGetLastError());
    }
    if ((rc = rt.mark(sgr, true))) return rc;
    if ((rc = rt.mark(smg, false))) return rc;
    if (L.n_rects > 0) {
        const dim3 mgrid((L.rect_w + 255) / 256, L.rect_h, L.n_rects);
        k_merge<<<mgrid, 256, 0, r.st>>>(p->tab_dev, L.rects, dest, use_gate);
        HIP_CHECK(hipGetLastError());
    }
    if ((rc = rt.mark(smg, true))) return rc;
    return OK;
}

// ---------------------------------------------------------------------------
// Process-global cache behind the host-memory conveniences (the node path):
// the reference host creates and destroys a node instance on every run, so
// plans, their device buffers and the embedded learned pilot live here and are
// reused while the image shape stays the same.
// ---------------------------------------------------------------------------
enum HostKind { HOST_PLANES = 0, HOST_RGBA = 1, HOST_COMPAT = 2 };

struct HostSlot {
    int kind = 0, w = 0, h = 0, c = 0, precision = 0, device = -1, flags = 0;
    cpipe_bm3d_plan* plan = nullptr;  // none for the compat filter
    void* dev = nullptr;              // device in/out staging
    void* pin = nullptr;              // pinned host in/out staging, same size
    size_t dev_bytes = 0;
    hipEvent_t ev[3] = {};            // per-plane transfer completion
    uint64_t used = 0;
};

struct HostCache {
    std::mutex mu;
    std::vector<HostSlot> slots;
    std::map<int, cpipe_bm3d_pilot*> nets;  // embedded pilot weights, per device
    std::map<int, OwnStream> streams;       // host-buffer runs, per device (non-blocking)
    uint64_t clock = 0;
};

// Never destroyed: HIP must not be called from static destructors at exit.
HostCache& host_cache() {
    static HostCache* c = new HostCache();
    return *c;
}

void drop_slot(HostSlot& sl) {
    DeviceGuard guard(sl.device);
    if (sl.plan) free_plan(sl.plan);
    (void)hipFree(sl.dev);
    if (sl.pin) (void)hipHostFree(sl.pin);
    for (hipEvent_t& e : sl.ev) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    sl.plan = nullptr;
    sl.dev = nullptr;
    sl.pin = nullptr;
}

int host_net(HostCache& hc, int device, cpipe_bm3d_pilot** out) {
    auto it = hc.nets.find(device);
    if (it != hc.nets.end()) {
        *out = it->second;
        return OK;
    }
    cpipe_bm3d_pilot* net = nullptr;
    int rc = cpipe_bm3d_pilot_create(device, &net);
    if (rc) return fail(rc, std::string("pilot: ") + cpipe_bm3d_pilot_last_error());
    rc = cpipe_bm3d_pilot_load(net, CPIPE_BM3D_PILOT_BLOB, CPIPE_BM3D_PILOT_BLOB_SIZE);
    if (rc) {
        cpipe_bm3d_pilot_destroy(net);
        return fail(rc, std::string("embedded pilot weights: ") + cpipe_bm3d_pilot_last_error());
    }
    hc.nets[device] = net;
    *out = net;
    return OK;
}

// Grows a slot's device and pinned staging to `bytes` each (none for the
// device-resident path, which passes 0).
int slot_staging(HostSlot& sl, size_t bytes) {
    if (bytes <= sl.dev_bytes) return OK;
    DeviceGuard guard(sl.device);
    (void)hipFree(sl.dev);
    if (sl.pin) (void)hipHostFree(sl.pin);
    sl.dev = sl.pin = nullptr;
    sl.dev_bytes = 0;
    hipError_t e = hipMalloc(&sl.dev, bytes);
    if (e == hipSuccess) e = hipHostMalloc(&sl.pin, bytes, hipHostMallocDefault);
    if (e != hipSuccess) return fail(OOM, std::string("staging buffers: ") + hipGetErrorString(e));
    sl.dev_bytes = bytes;
    return OK;
}

// The calling thread's noise profile for CPIPE_BM3D_HOST_PROFILE calls.
thread_local float g_prof_beta[3] = {0.0f, 0.0f, 0.0f}, g_prof_lambda[3] = {0.0f, 0.0f, 0.0f};
thread_local int g_prof_n = 0;

// Puts the thread's noise profile on the slot's plan (one pair per channel, or
// one pair for every channel), or leaves the plan on the scalar sigma.
int apply_profile(HostSlot* sl, int flags) {
    if (!(flags & CPIPE_BM3D_HOST_PROFILE)) return OK;
    if (g_prof_n != 1 && g_prof_n != sl->c)
        return fail(BAD_INDEX, "CPIPE_BM3D_HOST_PROFILE: set one (beta, lambda) pair or one per channel first");
    float b[3], l[3];
    for (int c = 0; c < sl->c; ++c) {
        b[c] = g_prof_beta[g_prof_n == 1 ? 0 : c];
        l[c] = g_prof_lambda[g_prof_n == 1 ? 0 : c];
    }
    return cpipe_bm3d_plan_set_noise_profile(sl->plan, b, l, sl->c);
}

// Finds or creates the slot for this call's shape (caller holds hc.mu).
int host_slot(HostCache& hc, int kind, int w, int h, int c, int precision, int device, int flags,
              size_t dev_bytes, HostSlot** out) {
    const int ndev = cpipe_bm3d_device_count();
    if (ndev <= 0) return fail(UNSUPPORTED, "no HIP device visible");
    if (device < 0 || device >= ndev) return fail(BAD_INDEX, "device ordinal out of range");
    if ((flags & CPIPE_BM3D_HOST_PILOT) && c != 3)
        return fail(UNSUPPORTED, "the learned pilot is trained for RGB (3-plane) images");
    if ((flags & CPIPE_BM3D_HOST_PILOT) && (flags & CPIPE_BM3D_HOST_PROFILE))
        return fail(UNSUPPORTED, "the learned pilot takes a scalar noise level");
    if (kind == HOST_COMPAT) precision = 0, flags = 0;
    for (HostSlot& sl : hc.slots) {
        if (sl.kind == kind && sl.w == w && sl.h == h && sl.c == c && sl.precision == precision &&
            sl.device == device && sl.flags == flags) {
            int rc = slot_staging(sl, dev_bytes);
            if (rc) return rc;
            sl.used = ++hc.clock;
            *out = &sl;
            return OK;
        }
    }
    if (hc.slots.size() >= CPIPE_BM3D_HOST_SLOTS) {  // evict the least recently used shape
        size_t lru = 0;
        for (size_t i = 1; i < hc.slots.size(); ++i)
            if (hc.slots[i].used < hc.slots[lru].used) lru = i;
        drop_slot(hc.slots[lru]);
        hc.slots.erase(hc.slots.begin() + (long)lru);
    }
    HostSlot sl;
    sl.kind = kind, sl.w = w, sl.h = h, sl.c = c, sl.precision = precision, sl.device = device;
    sl.flags = flags;
    int rc = OK;
    if (kind != HOST_COMPAT) {
        cpipe_bm3d_plan_desc d{sizeof(cpipe_bm3d_plan_desc), w, h, c, precision, device};
        if ((rc = cpipe_bm3d_plan_create(&d, &sl.plan))) return rc;
        if (flags & CPIPE_BM3D_HOST_PILOT) {
            cpipe_bm3d_pilot* net = nullptr;
            if (!(rc = host_net(hc, device, &net))) rc = cpipe_bm3d_plan_set_pilot(sl.plan, net, nullptr);
        }
    }
    if (!rc) {
        DeviceGuard guard(device);
        hipError_t e = hipSuccess;
        for (hipEvent_t& ev : sl.ev)
            if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) rc = fail(INTERNAL, std::string("slot events: ") + hipGetErrorString(e));
    }
    if (!rc) rc = slot_staging(sl, dev_bytes);
    if (rc) {
        drop_slot(sl);
        return rc;
    }
    sl.used = ++hc.clock;
    hc.slots.push_back(sl);
    *out = &hc.slots.back();
    return OK;
}

// The whole run: table upload, colour, step 1 (or the learned pilot with its
// spot checks and device-side fallback), step 2 into the callers' planes.
int enqueue_run(cpipe_bm3d_plan* p, hipStream_t st, bool timed, bool capturing);

// Brackets a run with the error flag: cleared first, copied to the pinned host
// word after the last kernel (cpipe_bm3d_plan_check reads it once `done` has
// passed; outside graph capture the event is recorded here).
int flag_begin(cpipe_bm3d_plan* p, hipStream_t st) {
    k_set_flag<<<1, 1, 0, st>>>(p->err, 0);
    HIP_CHECK(hipGetLastError());
    return OK;
}

int flag_end(cpipe_bm3d_plan* p, hipStream_t st, bool capturing) {
    HIP_CHECK(hipMemcpyAsync(p->err_host, p->err, sizeof(int), hipMemcpyDeviceToHost, st));
    if (!capturing) HIP_CHECK(hipEventRecord(p->done, st));
    return OK;
}

int enqueue(cpipe_bm3d_plan* p, hipStream_t st, bool timed, bool capturing) {
    int rc = flag_begin(p, st);
    if (!rc) rc = enqueue_run(p, st, timed, capturing);
    if (!rc) rc = flag_end(p, st, capturing);
    return rc;
}

int enqueue_run(cpipe_bm3d_plan* p, hipStream_t st, bool timed, bool capturing) {
    RunCtx r{p, st, timed};
    for (bool& b : p->stage_ran) b = false;
    const size_t n = p->fr.size();
    HIP_CHECK(hipMemcpyAsync(p->tab_dev, p->staging, p->staging_bytes, hipMemcpyHostToDevice, st));
    if (!capturing) {
        HIP_CHECK(hipEventRecord(p->staged, st));
        p->staged_pending = true;
    }
    if (p->n_small > 0) {  // frames smaller than one patch pass through
        const dim3 g((p->small_w + 255) / 256, p->small_h, p->n_small);
        k_copy_through<<<g, 256, 0, st>>>(p->tab_dev, p->small_rects);
        HIP_CHECK(hipGetLastError());
    }
    if (p->full.n_rects == 0) return OK;
    int rc;
    if ((rc = r.mark(CPIPE_BM3D_T_COLOR, false))) return rc;
    {
        const dim3 g((p->full.rect_w + 255) / 256, p->full.rect_h, p->full.n_rects);
        k_colour_in<<<g, 256, 0, st>>>(p->tab_dev, p->full.rects);
        HIP_CHECK(hipGetLastError());
    }
    if (p->profile && p->smap) {
        const dim3 g((unsigned)((p->smap_nx + 63) / 64), (unsigned)p->smap_ny, (unsigned)n);
        k_sigma_map<<<g, 64, 0, st>>>(p->tab_dev);
        HIP_CHECK(hipGetLastError());
    }
    if ((rc = r.mark(CPIPE_BM3D_T_COLOR, true))) return rc;

    if (p->pilot) {
        // Learned pilot in place of step 1. The spot checks run on the side stream
        // beside the whole-frame network: the plain step 1 and the network on each
        // checked rectangle (+ its halo), the residual ratios and each frame's gate.
        // A frame's network layers skip their work once its gate says it falls
        // back (its output is then replaced by the plain step 1).
        if ((rc = r.mark(CPIPE_BM3D_T_PILOT, false))) return rc;
        const int nchk = p->chk.n_rects;
        if (nchk > 0) {
            k_clear_gates<<<((int)n + 63) / 64, 64, 0, st>>>(p->tab_dev, (int)n);
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipEventRecord(p->fork_ev, st));
            HIP_CHECK(hipStreamWaitEvent(p->side, p->fork_ev, 0));
            RunCtx rs{p, p->side, false};
            if ((rc = run_step(rs, 1, p->chk, TO_CHECK, 0, false))) return rc;
            const int P = p->chk_pitch;
            const size_t plane = (size_t)P * P;
            k_pilot_gather<<<dim3(P, nchk), 256, 0, p->side>>>(p->tab_dev, p->chk.rects, levels_dev(p), p->chk_io,
                                                               p->chk_levels, P);
            HIP_CHECK(hipGetLastError());
            for (int f = 0; f < (int)n; ++f) {  // a frame's checks as one stack (slot stride 6 P^2)
                const FrameTab& T = p->tab[f];
                if (T.nchk == 0) continue;
                const auto& c = p->checks[T.chk0];
                float* slot = p->chk_io + (size_t)T.chk0 * 6 * plane;
                const float* bin[3] = {slot, slot + plane, slot + 2 * plane};
                float* bout[3] = {slot + 3 * plane, slot + 4 * plane, slot + 5 * plane};
                rc = cpipe_internal::pilot_forward(p->pilot, bin, (size_t)P, 6 * plane, c.hw, c.hh, T.nchk, 0.0f,
                                                   p->chk_levels + T.chk0, bout, p->chk_ws, p->side);
                if (rc) return fail(rc, std::string("pilot: ") + cpipe_bm3d_pilot_last_error());
            }
            const int rows = p->chk.rect_h;
            k_pilot_check_rows<<<dim3(rows, nchk), 256, 0, p->side>>>(p->tab_dev, p->chk.rects, p->chk_io, P,
                                                                       p->chk_part, rows);
            HIP_CHECK(hipGetLastError());
            k_pilot_check_sum<<<nchk, 256, 0, p->side>>>(p->chk.rects, p->chk_part, rows, p->pilot_thr_db,
                                                          p->pilot_under_db, p->report);
            HIP_CHECK(hipGetLastError());
            k_pilot_decide<<<((int)n + 63) / 64, 64, 0, p->side>>>(p->tab_dev, (int)n, p->report,
                                                                   p->pilot_full_frac);
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipEventRecord(p->join_ev, p->side));
        }
        for (int f = 0; f < (int)n; ++f) {
            const auto& F = p->fr[f];
            if (F.small) continue;
            float* zin[3] = {p->tab[f].z[0], p->tab[f].z[1], p->tab[f].z[2]};
            float* bout[3] = {p->tab[f].bz[0], p->tab[f].bz[1], p->tab[f].bz[2]};
            rc = cpipe_internal::pilot_forward(p->pilot, zin, (size_t)F.w, 0, F.w, F.h, 1, 0.0f,
                                               levels_dev(p) + f, bout, p->pilot_ws, st,
                                               nchk > 0 ? p->tab[f].gate : nullptr);
            if (rc) return fail(rc, std::string("pilot: ") + cpipe_bm3d_pilot_last_error());
        }
        if (nchk > 0) {
            HIP_CHECK(hipStreamWaitEvent(st, p->join_ev, 0));
            k_pilot_check_apply<<<dim3(p->chk.rect_h, nchk), 256, 0, st>>>(p->tab_dev, p->chk.rects, p->report);
            HIP_CHECK(hipGetLastError());
        }
        if ((rc = r.mark(CPIPE_BM3D_T_PILOT, true))) return rc;
        if (nchk > 0 && (rc = run_step(r, 1, p->full, TO_BASIC, 1, true))) return rc;
    } else {
        if ((rc = run_step(r, 1, p->full, TO_BASIC, 0, true))) return rc;
    }
    return run_step(r, 2, p->full, TO_OUTPUT, 0, true);
}

int collect_timing(cpipe_bm3d_plan* p) {
    HIP_CHECK(hipEventSynchronize(p->ev[2 * CPIPE_BM3D_T_MERGE2 + 1]));
    for (int i = 0; i < CPIPE_BM3D_T_COUNT; ++i) {
        if (!p->stage_ran[i]) continue;
        float ms = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&ms, p->ev[2 * i], p->ev[2 * i + 1]));
        p->ms[i] += ms;
    }
    p->runs++;
    return OK;
}

// Runs the plan with the frame table already set: eagerly (timed if asked),
// or as one replay of the captured graph (captured on first use after an
// eager run has sized every lazily allocated buffer).
int run_plan(cpipe_bm3d_plan* p, hipStream_t st) {
    int rc = fill_staging(p);
    if (rc) return rc;
    const bool timed = p->timing && p->full.n_rects > 0;
    if (!p->use_graph || timed || p->dirty) {
        if ((rc = enqueue(p, st, timed, false))) return rc;
        p->ran = true;
        if (timed && (rc = collect_timing(p))) return rc;
        if (p->use_graph && p->dirty) {
            // capture the next runs' graph now (buffers exist after this eager run)
            HIP_CHECK(hipEventSynchronize(p->staged));
            p->staged_pending = false;
            drop_graph(p);
            hipStream_t cs;
            HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                rc = enqueue(p, cs, false, true);
                hipError_t e2 = hipStreamEndCapture(cs, &g);
                if (!rc && e2 != hipSuccess) rc = fail(INTERNAL, std::string("graph capture: ") + hipGetErrorString(e2));
                if (!rc) {
                    e2 = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
                    if (e2 != hipSuccess) rc = fail(INTERNAL, std::string("graph instantiate: ") + hipGetErrorString(e2));
                }
                if (g) (void)hipGraphDestroy(g);
            } else {
                rc = fail(INTERNAL, std::string("begin capture: ") + hipGetErrorString(e));
            }
            (void)hipStreamDestroy(cs);
            if (rc) return rc;
            p->dirty = false;
        }
        return OK;
    }
    HIP_CHECK(hipGraphLaunch(p->exec, st));
    HIP_CHECK(hipEventRecord(p->staged, st));
    HIP_CHECK(hipEventRecord(p->done, st));
    p->staged_pending = true;
    p->ran = true;
    return OK;
}

// Spot-check rectangles of every frame and their work lists.
int build_check_lists(cpipe_bm3d_plan* p, const cpipe_bm3d_pilot_opts& o) {
    free_lists(p->chk);
    (void)hipFree(p->report);
    (void)hipFree(p->chk_part);
    (void)hipFree(p->chk_io);
    (void)hipFree(p->chk_levels);
    (void)hipFree(p->chk_ws);
    p->report = nullptr;
    p->chk_part = nullptr;
    p->chk_io = nullptr;
    p->chk_levels = nullptr;
    p->chk_ws = nullptr;
    p->checks.clear();
    std::vector<FrameRect> rects;
    std::vector<int4> bm;
    std::vector<int2> grp;
    int cw = 0, chh = 0;
    const int B = o.check_block;
    for (int f = 0; f < (int)p->fr.size(); ++f) {
        const auto& F = p->fr[f];
        FrameTab& T = p->tab[f];
        T.chk0 = (int)p->checks.size();
        T.nchk = 0;
        if (F.small || o.check_every <= 0) continue;
        const int* txlo = F.geo.data();
        const int* txhi = txlo + F.w;
        const int* tylo = txhi + F.w;
        const int* tyhi = tylo + F.h;
        const int nbx = (F.w + B - 1) / B, nby = (F.h + B - 1) / B;
        const int nbtx = (F.nx + BM_TX - 1) / BM_TX, nbty = (F.ny + BM_TY - 1) / BM_TY;
        std::vector<char> bm_used((size_t)nbtx * nbty, 0), tile_used((size_t)F.ntx * F.nty, 0);
        auto add_check = [&](int b) {
            cpipe_bm3d_plan_s::Check c{};
            c.f = f;
            c.x0 = (b % nbx) * B;
            c.y0 = (b / nbx) * B;
            c.x1 = std::min(F.w, c.x0 + B);
            c.y1 = std::min(F.h, c.y0 + B);
            c.hx0 = pilot_region_lo(c.x0, F.w, B + 2 * PILOT_HALO);
            c.hy0 = pilot_region_lo(c.y0, F.h, B + 2 * PILOT_HALO);
            c.hw = std::min(B + 2 * PILOT_HALO, F.w);
            c.hh = std::min(B + 2 * PILOT_HALO, F.h);
            c.tx0 = txlo[c.x0];
            c.tx1 = txhi[c.x1 - 1] + 1;
            c.ty0 = tylo[c.y0];
            c.ty1 = tyhi[c.y1 - 1] + 1;
            const int rx0 = c.tx0 * TILE, rx1 = std::min(c.tx1 * TILE, F.nx);
            const int ry0 = c.ty0 * p->tile_y, ry1 = std::min(c.ty1 * p->tile_y, F.ny);
            c.btx0 = rx0 / BM_TX;
            c.btx1 = (rx1 + BM_TX - 1) / BM_TX;
            c.bty0 = ry0 / BM_TY;
            c.bty1 = (ry1 + BM_TY - 1) / BM_TY;
            p->checks.push_back(c);
            rects.push_back({f, c.x0, c.y0, c.x1, c.y1, 0, 0, 0});
            T.nchk++;
            for (int ty = c.bty0; ty < c.bty1; ++ty)
                for (int tx = c.btx0; tx < c.btx1; ++tx) bm_used[(size_t)ty * nbtx + tx] = 1;
            for (int ty = c.ty0; ty < c.ty1; ++ty)
                for (int tx = c.tx0; tx < c.tx1; ++tx) tile_used[(size_t)ty * F.ntx + tx] = 1;
            cw = std::max(cw, c.x1 - c.x0);
            chh = std::max(chh, c.y1 - c.y0);
        };
        // every check_every-th block in raster order, the phase shifted by one per block row;
        // a frame the pattern misses (fewer blocks than it needs) is checked on its central
        // block, so no frame runs the network unchecked
        for (int b = 0; b < nbx * nby; ++b)
            if ((b + b / nbx) % o.check_every == o.check_every / 2) add_check(b);
        if (T.nchk == 0) add_check((nby / 2) * nbx + nbx / 2);
        // each tile once: overlapping checks must not run one tile twice in a launch
        for (int i = 0; i < nbtx * nbty; ++i)
            if (bm_used[i]) bm.push_back(make_int4(f, i % nbtx, i / nbtx, 0));
        for (int i = 0; i < F.ntx * F.nty; ++i)
            if (tile_used[i]) grp.push_back(make_int2(f, i));
    }
    int rc;
    auto& L = p->chk;
    if ((rc = upload(&L.bm, bm)) || (rc = upload(&L.grp, grp)) || (rc = upload(&L.rects, rects)) ||
        (rc = dev_alloc((void**)&p->report, std::max<size_t>(1, rects.size()) * 2 * sizeof(float))))
        return rc;
    HIP_CHECK(hipMemset(p->report, 0, std::max<size_t>(1, rects.size()) * 2 * sizeof(float)));
    if ((rc = dev_alloc((void**)&p->chk_part, std::max<size_t>(1, rects.size() * (size_t)std::max(1, chh)) * sizeof(float2))))
        return rc;
    p->chk_pitch = B + 2 * PILOT_HALO;
    const size_t plane = (size_t)p->chk_pitch * p->chk_pitch;
    size_t max_stack = 1;  // a frame's checks run as one stack of images
    for (const auto& T : p->tab) max_stack = std::max(max_stack, (size_t)T.nchk);
    if ((rc = dev_alloc((void**)&p->chk_io, std::max<size_t>(1, rects.size()) * 6 * plane * sizeof(float))) ||
        (rc = dev_alloc((void**)&p->chk_levels, std::max<size_t>(1, rects.size()) * sizeof(float))) ||
        (rc = dev_alloc(&p->chk_ws, cpipe_internal::pilot_workspace_bytes(plane * max_stack))))
        return rc;
    L.n_bm = (int)bm.size();
    L.n_grp = (int)grp.size();
    L.n_rects = (int)rects.size();
    L.rect_w = cw;
    L.rect_h = chh;
    return OK;
}

// Autotuner cache: best (tile height, work-list order, band rows) per plan shape.
struct TuneKey {
    int precision, c;
    std::vector<int> dims;
    bool operator<(const TuneKey& o) const {
        return std::tie(precision, c, dims) < std::tie(o.precision, o.c, o.dims);
    }
};
struct TuneChoice {
    int tile_y, order, band_rows, bm_mfma;
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneChoice> g_tune_cache;

TuneKey tune_key(const cpipe_bm3d_plan* p) {
    TuneKey k{p->precision, p->c, {}};
    for (const auto& f : p->fr) {
        k.dims.push_back(f.w);
        k.dims.push_back(f.h);
    }
    return k;
}

int apply_choice(cpipe_bm3d_plan* p, const TuneChoice& ch) {
    const bool reshape = ch.tile_y != p->tile_y;
    p->order = ch.order;
    p->band_rows = ch.band_rows;
    p->bm_mfma = ch.bm_mfma;
    drop_graph(p);
    int rc = reshape ? set_tile_shape(p, ch.tile_y) : build_full_lists(p);
    if (rc) return rc;
    // a new tile shape rebuilds the frames: re-create the pilot's buffers and checks
    if (reshape && p->pilot) rc = cpipe_bm3d_plan_set_pilot(p, p->pilot, &p->pilot_opts);
    return rc;
}

}  // namespace

// ---------------------------------------------------------------------------
// The entry point that allocates nothing (cpipe_bm3d_noalloc_planes). Every
// buffer of an exact FP32 run -- frame table, work lists, tile geometry,
// opponent and basic-estimate planes, match lists, tile partials and the
// consistency flag -- is carved from `scratch`, one file-scope device array
// that each GPU the library runs on holds until the process ends. A setup
// kernel writes the table, the geometry and the work lists, so a call only
// enqueues kernels: no allocation and no host-to-device copy.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t NOALLOC_BYTES = (size_t)CPIPE_BM3D_NOALLOC_SCRATCH_MIB << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

std::mutex g_noalloc_mu;
unsigned char* g_noalloc_base[64] = {};  // device address of `scratch` on each device

// The device address of `scratch` on the current device `dev`, looked up on the
// first call there (the array's name is only a host-side stand-in).
int noalloc_base(int dev, unsigned char** out) {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_base[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(INTERNAL, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_noalloc_base[dev] = static_cast<unsigned char*>(ptr);
    }
    *out = g_noalloc_base[dev];
    return OK;
}

constexpr int NOALLOC_BAND = 32;  // rows per colour / merge rectangle (the plans' default)

// Byte offsets into `scratch` (256-aligned) and the work counts of one frame,
// laid out as a single-frame exact FP32 plan lays out its buffers.
struct NoallocLayout {
    size_t tab, rects, flags, bm, grp, geo, z, bz, pos0, pos1, cnt0, cnt1, partial, total;
    int small, nx, ny, ntx, nty, nbtx, nbty, n_bm, n_grp, n_rects;
};

NoallocLayout noalloc_layout(int w, int h, int c) {
    NoallocLayout L{};
    size_t off = 0;
    auto take = [&off](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t npx = (size_t)w * h;
    L.small = w < K || h < K;
    L.n_rects = L.small ? 1 : (h + NOALLOC_BAND - 1) / NOALLOC_BAND;
    L.tab = take(sizeof(FrameTab));
    L.rects = take(sizeof(FrameRect) * L.n_rects);
    L.flags = take(2 * sizeof(int));  // [0] consistency flag, [1] gate (unused: no pilot)
    if (!L.small) {
        L.nx = ref_count(w);
        L.ny = ref_count(h);
        L.ntx = (L.nx + TILE - 1) / TILE;
        L.nty = (L.ny + EXACT_TY - 1) / EXACT_TY;
        L.nbtx = (L.nx + BM_TX - 1) / BM_TX;
        L.nbty = (L.ny + BM_TY - 1) / BM_TY;
        L.n_bm = L.nbtx * L.nbty;
        L.n_grp = L.ntx * L.nty;
        const size_t nref = (size_t)L.nx * L.ny;
        const size_t tile_stride = 2 * (size_t)FOOT * foot_edge(EXACT_TY);
        L.bm = take(sizeof(int4) * L.n_bm);
        L.grp = take(sizeof(int2) * L.n_grp);
        L.geo = take(sizeof(int) * (2 * (size_t)w + 2 * (size_t)h + 2 * (size_t)L.ntx + 2 * (size_t)L.nty));
        L.z = take(sizeof(float) * npx * c);
        L.bz = take(sizeof(float) * npx * c);
        L.pos0 = take(sizeof(uint32_t) * nref * NMAX);
        L.pos1 = take(sizeof(uint32_t) * nref * NMAX);
        L.cnt0 = take(nref);
        L.cnt1 = take(nref);
        L.partial = take(sizeof(float) * tile_stride * L.n_grp * c);
    }
    L.total = off;
    return L;
}

struct NoallocSetup {
    unsigned char* base;
    NoallocLayout L;
    int w, h, c;
    float sigma, sig[3];
    const float* in[3];
    float* out[3];
    size_t in_pitch, out_pitch;
};

__device__ inline void tile_box_dev(int t, int tile, int nref, int n, int* f0, int* f1) {
    const int first = ref_pos(t * tile, n);
    const int lastref = min(t * tile + tile - 1, nref - 1);
    const int last = ref_pos(lastref, n);
    *f0 = max(0, first - R);
    *f1 = min(n, last + R + K);
}

// One workgroup: the frame table, the merge / colour rectangles, the work lists
// and the tile geometry of one frame, as the plans build them on the host.
__global__ __launch_bounds__(256) void k_noalloc_setup(NoallocSetup S) {
    const NoallocLayout& L = S.L;
    unsigned char* b = S.base;
    FrameTab* T = reinterpret_cast<FrameTab*>(b + L.tab);
    FrameRect* rects = reinterpret_cast<FrameRect*>(b + L.rects);
    int* flags = reinterpret_cast<int*>(b + L.flags);
    const int tid = threadIdx.x, nt = blockDim.x, w = S.w, h = S.h;
    int* txlo = reinterpret_cast<int*>(b + L.geo);
    int* txhi = txlo + w;
    int* tylo = txhi + w;
    int* tyhi = tylo + h;
    int* fx0 = tyhi + h;
    int* fx1 = fx0 + L.ntx;
    int* fy0 = fx1 + L.ntx;
    int* fy1 = fy0 + L.nty;
    if (tid == 0) {
        FrameTab t = {};
        t.w = w;
        t.h = h;
        t.c = S.c;
        t.gate = flags + 1;
        for (int ch = 0; ch < 3; ++ch) {
            const int cc = ch < S.c ? ch : 0;
            t.sig[ch] = S.sig[ch];
            t.in[ch] = S.in[cc];
            t.out[ch] = S.out[cc];
        }
        t.sigma = S.sigma;
        t.in_pitch = S.in_pitch;
        t.out_pitch = S.out_pitch;
        if (!L.small) {
            const size_t npx = (size_t)w * h;
            t.nx = L.nx;
            t.ny = L.ny;
            t.ntx = L.ntx;
            t.nty = L.nty;
            t.tile_y = EXACT_TY;
            t.dofs = FOOT * foot_edge(EXACT_TY);
            t.tile_stride = 2 * (size_t)t.dofs;
            for (int ch = 0; ch < 3; ++ch) {
                const int cc = ch < S.c ? ch : 0;
                t.z[ch] = reinterpret_cast<float*>(b + L.z) + cc * npx;
                t.bz[ch] = reinterpret_cast<float*>(b + L.bz) + cc * npx;
            }
            t.pos[0] = reinterpret_cast<uint32_t*>(b + L.pos0);
            t.pos[1] = reinterpret_cast<uint32_t*>(b + L.pos1);
            t.cnt[0] = b + L.cnt0;
            t.cnt[1] = b + L.cnt1;
            t.partial = reinterpret_cast<float*>(b + L.partial);
            t.txlo = txlo;
            t.txhi = txhi;
            t.tylo = tylo;
            t.tyhi = tyhi;
            t.fx0 = fx0;
            t.fx1 = fx1;
            t.fy0 = fy0;
            t.fy1 = fy1;
        }
        *T = t;
        flags[0] = 0;
        flags[1] = 0;
    }
    if (L.small) {
        if (tid == 0) rects[0] = FrameRect{0, 0, 0, w, h, 0, 0, 0};
        return;
    }
    for (int i = tid; i < L.n_rects; i += nt)
        rects[i] = FrameRect{0, 0, i * NOALLOC_BAND, w, min(h, (i + 1) * NOALLOC_BAND), 0, 0, 0};
    int4* bm = reinterpret_cast<int4*>(b + L.bm);
    for (int i = tid; i < L.n_bm; i += nt) bm[i] = make_int4(0, i % L.nbtx, i / L.nbtx, 0);
    int2* grp = reinterpret_cast<int2*>(b + L.grp);
    for (int i = tid; i < L.n_grp; i += nt) grp[i] = make_int2(0, i);
    for (int t = tid; t < L.ntx; t += nt) tile_box_dev(t, TILE, L.nx, w, &fx0[t], &fx1[t]);
    for (int t = tid; t < L.nty; t += nt) tile_box_dev(t, EXACT_TY, L.ny, h, &fy0[t], &fy1[t]);
    // first and last tile whose footprint covers each column / row
    for (int x = tid; x < w; x += nt) {
        int lo = -1, hi = -2;
        for (int t = 0; t < L.ntx; ++t) {
            int a0, a1;
            tile_box_dev(t, TILE, L.nx, w, &a0, &a1);
            if (x >= a0 && x < a1) {
                if (lo < 0) lo = t;
                hi = t;
            }
        }
        txlo[x] = lo;
        txhi[x] = hi;
    }
    for (int y = tid; y < h; y += nt) {
        int lo = -1, hi = -2;
        for (int t = 0; t < L.nty; ++t) {
            int a0, a1;
            tile_box_dev(t, EXACT_TY, L.ny, h, &a0, &a1);
            if (y >= a0 && y < a1) {
                if (lo < 0) lo = t;
                hi = t;
            }
        }
        tylo[y] = lo;
        tyhi[y] = hi;
    }
}

}  // namespace

extern "C" {

const char* cpipe_bm3d_last_error(void) { return g_err.c_str(); }

int cpipe_bm3d_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int cpipe_bm3d_plan_create(const cpipe_bm3d_plan_desc* d, cpipe_bm3d_plan** out_plan) {
    if (!d || !out_plan) return fail(BAD_INDEX, "null argument");
    *out_plan = nullptr;
    if (d->struct_size < sizeof(cpipe_bm3d_plan_desc))
        return fail(BAD_INDEX, "cpipe_bm3d_plan_desc.struct_size too small");
    int rc = create_core(out_plan, &d->width, &d->height, 1, d->channels, d->precision, d->device);
    if (rc) return rc;
    cpipe_bm3d_plan* p = *out_plan;
    if (p->c == 3) {
        DeviceGuard guard(p->device);
        if ((rc = dev_alloc((void**)&p->rgb, sizeof(float) * 6 * (size_t)d->width * d->height))) {
            free_plan(p);
            *out_plan = nullptr;
            return rc;
        }
    }
    return OK;
}

void cpipe_bm3d_plan_destroy(cpipe_bm3d_plan* plan) { free_plan(plan); }

int cpipe_bm3d_plan_run(cpipe_bm3d_plan* p, const float* const* in, size_t in_pitch,
                        float* const* out, size_t out_pitch, float sigma, void* stream) {
    if (!p || !in || !out) return fail(BAD_INDEX, "null argument");
    if (p->fr.size() != 1) return fail(BAD_INDEX, "a batch runs with cpipe_bm3d_batch_run");
    for (int c = 0; c < p->c; ++c)
        if (!in[c] || !out[c]) return fail(BAD_INDEX, "null plane pointer");
    if (in_pitch < (size_t)p->fr[0].w || out_pitch < (size_t)p->fr[0].w)
        return fail(BAD_INDEX, "pitch smaller than width");
    if (!(sigma >= 0.0f) || !std::isfinite(sigma)) return fail(BAD_INDEX, "sigma must be finite and >= 0");
    DeviceGuard guard(p->device);
    set_run(p, 0, in, in_pitch, out, out_pitch, sigma);
    return run_plan(p, (hipStream_t)stream);
}

int cpipe_bm3d_batch_create(const cpipe_bm3d_batch_desc* d, cpipe_bm3d_batch** out) {
    if (!d || !out) return fail(BAD_INDEX, "null argument");
    *out = nullptr;
    if (d->struct_size < sizeof(cpipe_bm3d_batch_desc))
        return fail(BAD_INDEX, "cpipe_bm3d_batch_desc.struct_size too small");
    if (!d->widths || !d->heights) return fail(BAD_INDEX, "null frame sizes");
    int rc = create_core(out, d->widths, d->heights, d->n_frames, d->channels, d->precision, d->device);
    if (rc) return rc;
    cpipe_bm3d_plan* p = *out;
    p->use_graph = d->use_graph != 0;
    TuneChoice ch{};
    bool cached = false;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(tune_key(p));
        if (it != g_tune_cache.end()) {
            ch = it->second;
            cached = true;
        }
    }
    if (cached) {
        DeviceGuard guard(p->device);
        if ((rc = apply_choice(p, ch))) {
            free_plan(p);
            *out = nullptr;
            return rc;
        }
    }
    return OK;
}

void cpipe_bm3d_batch_destroy(cpipe_bm3d_batch* b) { free_plan(b); }

int cpipe_bm3d_batch_run(cpipe_bm3d_batch* b, const float* const* in, const size_t* in_pitch,
                         float* const* out, const size_t* out_pitch, const float* sigmas,
                         void* stream) {
    if (!b || !in || !out || !in_pitch || !out_pitch || !sigmas) return fail(BAD_INDEX, "null argument");
    const int n = (int)b->fr.size(), C = b->c;
    for (int f = 0; f < n; ++f) {
        for (int c = 0; c < C; ++c)
            if (!in[f * C + c] || !out[f * C + c]) return fail(BAD_INDEX, "null plane pointer");
        if (in_pitch[f] < (size_t)b->fr[f].w || out_pitch[f] < (size_t)b->fr[f].w)
            return fail(BAD_INDEX, "pitch smaller than width");
        if (!(sigmas[f] >= 0.0f) || !std::isfinite(sigmas[f]))
            return fail(BAD_INDEX, "sigma must be finite and >= 0");
    }
    DeviceGuard guard(b->device);
    for (int f = 0; f < n; ++f) set_run(b, f, in + f * C, in_pitch[f], out + f * C, out_pitch[f], sigmas[f]);
    return run_plan(b, (hipStream_t)stream);
}

int cpipe_bm3d_batch_tune(cpipe_bm3d_batch* b, const float* const* in, const size_t* in_pitch,
                          float* const* out, const size_t* out_pitch, const float* sigmas, int iters,
                          void* stream, cpipe_bm3d_tune_report* report) {
    if (!b || !in || !out || !in_pitch || !out_pitch || !sigmas) return fail(BAD_INDEX, "null argument");
    if (iters < 1) iters = 1;
    DeviceGuard guard(b->device);
    hipStream_t st = (hipStream_t)stream;
    const int n = (int)b->fr.size(), C = b->c;
    auto set_all = [&]() {
        for (int f = 0; f < n; ++f) set_run(b, f, in + f * C, in_pitch[f], out + f * C, out_pitch[f], sigmas[f]);
    };
    const bool graph = b->use_graph, timing = b->timing;
    b->use_graph = false;
    b->timing = false;
    std::vector<TuneChoice> cands;
    std::vector<int> tys = {b->tile_y}, bms = {0};
    if (b->precision != CPIPE_BM3D_PRECISION_FP32) {
        tys = {TILE / 2, TILE};
        bms = {0, 1};
    }
    for (int bm : bms)
        for (int ty : tys)
            for (int order : {0, 1})
                for (int rows : {16, 64}) cands.push_back({ty, order, rows, bm});
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    float best_ms = 1e30f;
    TuneChoice best = cands[0];
    int rc = OK, k = 0;
    for (const auto& ch : cands) {
        if ((rc = apply_choice(b, ch))) break;
        set_all();
        if ((rc = fill_staging(b)) || (rc = enqueue(b, st, false, false))) break;  // warm-up
        HIP_CHECK(hipEventRecord(e0, st));
        for (int i = 0; i < iters && !rc; ++i) {
            if ((rc = fill_staging(b)) || (rc = enqueue(b, st, false, false))) break;
        }
        if (rc) break;
        HIP_CHECK(hipEventRecord(e1, st));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        ms /= (float)iters;
        if (report && k < CPIPE_BM3D_TUNE_MAX) {
            report->tile_y[k] = ch.tile_y;
            report->order[k] = ch.order;
            report->band_rows[k] = ch.band_rows;
            report->bm_mfma[k] = ch.bm_mfma;
            report->ms[k] = ms;
        }
        ++k;
        if (ms < best_ms) {
            best_ms = ms;
            best = ch;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->use_graph = graph;
    b->timing = timing;
    if (rc) return rc;
    if ((rc = apply_choice(b, best))) return rc;
    set_all();
    if (report) {
        report->n = std::min(k, (int)CPIPE_BM3D_TUNE_MAX);
        report->best = -1;
        for (int i = 0; i < report->n; ++i)
            if (report->tile_y[i] == best.tile_y && report->order[i] == best.order &&
                report->band_rows[i] == best.band_rows && report->bm_mfma[i] == best.bm_mfma)
                report->best = i;
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[tune_key(b)] = best;
    return OK;
}

int cpipe_bm3d_batch_config(const cpipe_bm3d_batch* b, int* tile_y, int* order, int* band_rows,
                            int* bm_mfma) {
    if (!b) return fail(BAD_INDEX, "null batch");
    if (tile_y) *tile_y = b->tile_y;
    if (order) *order = b->order;
    if (band_rows) *band_rows = b->band_rows;
    if (bm_mfma) *bm_mfma = b->bm_mfma;
    return OK;
}

// Per-frame noise profiles (beta/lambda: [frame][channel], `stride` floats per
// frame; stride 0 = one set of `p->c` pairs for every frame), or none.
static int set_profiles(cpipe_bm3d_plan* p, const float* beta, const float* lambda, size_t stride, bool on) {
    const int nf = (int)p->fr.size();
    if (on && (!beta || !lambda)) return fail(BAD_INDEX, "null noise profile");
    if (on) {
        for (int f = 0; f < nf; ++f)
            for (int c = 0; c < p->c; ++c)
                if (!std::isfinite(beta[f * stride + c]) || !std::isfinite(lambda[f * stride + c]))
                    return fail(BAD_INDEX, "noise profile values must be finite");
        if (p->pilot) return fail(UNSUPPORTED, "the learned pilot takes a scalar noise level");
    }
    DeviceGuard guard(p->device);
    if (on != p->profile) {
        HIP_CHECK(hipDeviceSynchronize());
        drop_graph(p);
    }
    if (on && !p->smap) {
        std::vector<size_t> off(nf, 0);
        size_t total = 0;
        int mx = 0, my = 0;
        for (int f = 0; f < nf; ++f) {
            const auto& F = p->fr[f];
            off[f] = total;
            if (F.small) continue;
            total += (size_t)p->c * F.nx * F.ny;
            mx = std::max(mx, F.nx);
            my = std::max(my, F.ny);
        }
        if (total > 0 && hipMalloc(&p->smap, sizeof(float) * total) != hipSuccess) {
            p->smap = nullptr;
            return fail(OOM, "noise-profile map");
        }
        p->smap_off = off;
        p->smap_nx = mx;
        p->smap_ny = my;
    }
    p->profile = on;
    p->pbeta.assign(3 * (size_t)nf, 0.0f);
    p->plambda.assign(3 * (size_t)nf, 0.0f);
    for (int f = 0; on && f < nf; ++f)
        for (int c = 0; c < p->c; ++c) {
            p->pbeta[3 * f + c] = beta[f * stride + c];
            p->plambda[3 * f + c] = lambda[f * stride + c];
        }
    return OK;
}

int cpipe_bm3d_plan_set_noise_profile(cpipe_bm3d_plan* p, const float* beta, const float* lambda, int n) {
    if (!p) return fail(BAD_INDEX, "null plan");
    if (n != 0 && n != p->c) return fail(BAD_INDEX, "a noise profile needs one (beta, lambda) pair per channel");
    return set_profiles(p, beta, lambda, 0, n != 0);
}

int cpipe_bm3d_batch_set_noise_profiles(cpipe_bm3d_batch* b, const float* beta, const float* lambda,
                                        int pairs_per_frame) {
    if (!b) return fail(BAD_INDEX, "null batch");
    if (pairs_per_frame != 0 && pairs_per_frame != b->c)
        return fail(BAD_INDEX, "a noise profile needs one (beta, lambda) pair per channel");
    return set_profiles(b, beta, lambda, (size_t)b->c, pairs_per_frame != 0);
}

int cpipe_bm3d_plan_set_block_matching(cpipe_bm3d_plan* p, int mfma) {
    if (!p) return fail(BAD_INDEX, "null plan");
    if (mfma && p->precision == CPIPE_BM3D_PRECISION_FP32)
        return fail(UNSUPPORTED, "MFMA block matching needs a mixed-precision plan");
    if ((p->bm_mfma != 0) != (mfma != 0)) {
        DeviceGuard guard(p->device);
        HIP_CHECK(hipDeviceSynchronize());
        drop_graph(p);
    }
    p->bm_mfma = mfma != 0;
    return OK;
}

int cpipe_bm3d_plan_run_rgba16f(cpipe_bm3d_plan* p, const uint16_t* in, size_t isp, uint16_t* out,
                                size_t osp, float sigma, void* stream) {
    if (!p || !in || !out) return fail(BAD_INDEX, "null argument");
    if (p->fr.size() != 1 || p->c != 3) return fail(BAD_INDEX, "RGBA16F needs a 3-channel single-frame plan");
    const int w = p->fr[0].w, h = p->fr[0].h;
    if (isp < (size_t)w || osp < (size_t)w) return fail(BAD_INDEX, "stride smaller than width");
    DeviceGuard guard(p->device);
    hipStream_t st = (hipStream_t)stream;
    const dim3 rowgrid((w + 255) / 256, h);
    const size_t npx = p->fr[0].npx;
    float* rin[3] = {p->rgb, p->rgb + npx, p->rgb + 2 * npx};
    float* rout[3] = {p->rgb + 3 * npx, p->rgb + 4 * npx, p->rgb + 5 * npx};
    k_unpack_rgba16f<<<rowgrid, 256, 0, st>>>(in, isp, rin[0], rin[1], rin[2], w);
    HIP_CHECK(hipGetLastError());
    const int rc = cpipe_bm3d_plan_run(p, rin, (size_t)w, rout, (size_t)w, sigma, stream);
    if (rc) return rc;
    k_pack_rgba16f<<<rowgrid, 256, 0, st>>>(rout[0], rout[1], rout[2], in, isp, out, osp, w);
    HIP_CHECK(hipGetLastError());
    return OK;
}

int cpipe_bm3d_plan_basic(cpipe_bm3d_plan* p, float* const* out, size_t out_pitch, void* stream) {
    if (!p || !out) return fail(BAD_INDEX, "null argument");
    if (!p->ran) return fail(BAD_INDEX, "plan has not run yet");
    const auto& F = p->fr[0];
    if (F.small) return fail(UNSUPPORTED, "image smaller than one patch has no basic estimate");
    DeviceGuard guard(p->device);
    hipStream_t st = (hipStream_t)stream;
    const size_t npx = F.npx;
    const dim3 rowgrid((F.w + 255) / 256, F.h);
    if (p->c == 3) {
        k_opp_inv<<<rowgrid, 256, 0, st>>>(F.bz, F.bz + npx, F.bz + 2 * npx, out[0], out[1], out[2],
                                           out_pitch, F.w);
    } else {
        k_copy_plane<<<rowgrid, 256, 0, st>>>(F.bz, (size_t)F.w, out[0], out_pitch, F.w);
    }
    HIP_CHECK(hipGetLastError());
    return OK;
}

int cpipe_bm3d_plan_matches(cpipe_bm3d_plan* p, int step, uint32_t* pos, uint8_t* count,
                            void* stream) {
    if (!p || !pos || !count || (step != 1 && step != 2)) return fail(BAD_INDEX, "bad argument");
    const auto& F = p->fr[0];
    if (!p->ran || F.nx == 0) return fail(BAD_INDEX, "no match lists");
    DeviceGuard guard(p->device);
    hipStream_t st = (hipStream_t)stream;
    const size_t nref = (size_t)F.nx * F.ny;
    HIP_CHECK(hipMemcpyAsync(pos, F.pos[step - 1], nref * NMAX * sizeof(uint32_t),
                             hipMemcpyDeviceToDevice, st));
    HIP_CHECK(hipMemcpyAsync(count, F.cnt[step - 1], nref, hipMemcpyDeviceToDevice, st));
    return OK;
}

int cpipe_bm3d_plan_step1_opponent(cpipe_bm3d_plan* p, const float* const* in, size_t in_pitch,
                                   float sigma, float* const* noisy_out, float* const* basic_out,
                                   size_t out_pitch, void* stream) {
    if (!p || !in || !noisy_out || !basic_out) return fail(BAD_INDEX, "null argument");
    if (p->fr.size() != 1 || p->c != 3) return fail(UNSUPPORTED, "opponent planes exist for 3-channel single-frame plans");
    const auto& F = p->fr[0];
    if (F.small) return fail(UNSUPPORTED, "image smaller than one patch");
    if (in_pitch < (size_t)F.w || out_pitch < (size_t)F.w) return fail(BAD_INDEX, "pitch smaller than width");
    DeviceGuard guard(p->device);
    hipStream_t st = (hipStream_t)stream;
    set_run(p, 0, in, in_pitch, nullptr, 0, sigma);
    int rc = fill_staging(p);
    if (rc) return rc;
    HIP_CHECK(hipMemcpyAsync(p->tab_dev, p->staging, p->staging_bytes, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipEventRecord(p->staged, st));
    p->staged_pending = true;
    p->ran = true;
    if ((rc = flag_begin(p, st))) return rc;
    const dim3 g((p->full.rect_w + 255) / 256, p->full.rect_h, p->full.n_rects);
    k_colour_in<<<g, 256, 0, st>>>(p->tab_dev, p->full.rects);
    HIP_CHECK(hipGetLastError());
    RunCtx r{p, st, false};
    if ((rc = run_step(r, 1, p->full, TO_BASIC, 0, false))) return rc;
    if ((rc = flag_end(p, st, false))) return rc;
    for (int c = 0; c < 3; ++c) {
        HIP_CHECK(hipMemcpy2DAsync(noisy_out[c], out_pitch * sizeof(float), p->tab[0].z[c], (size_t)F.w * sizeof(float),
                                   (size_t)F.w * sizeof(float), F.h, hipMemcpyDeviceToDevice, st));
        HIP_CHECK(hipMemcpy2DAsync(basic_out[c], out_pitch * sizeof(float), p->tab[0].bz[c], (size_t)F.w * sizeof(float),
                                   (size_t)F.w * sizeof(float), F.h, hipMemcpyDeviceToDevice, st));
    }
    return OK;
}

int cpipe_bm3d_plan_step2_opponent(cpipe_bm3d_plan* p, const float* const* in, size_t in_pitch, float sigma,
                                   const float* const* basic_in, size_t basic_pitch, float* const* out,
                                   size_t out_pitch, void* stream) {
    if (!p || !in || !basic_in || !out) return fail(BAD_INDEX, "null argument");
    if (p->fr.size() != 1 || p->c != 3) return fail(UNSUPPORTED, "opponent planes exist for 3-channel single-frame plans");
    const auto& F = p->fr[0];
    if (F.small) return fail(UNSUPPORTED, "image smaller than one patch");
    for (int c = 0; c < 3; ++c)
        if (!in[c] || !basic_in[c] || !out[c]) return fail(BAD_INDEX, "null plane pointer");
    if (in_pitch < (size_t)F.w || basic_pitch < (size_t)F.w || out_pitch < (size_t)F.w)
        return fail(BAD_INDEX, "pitch smaller than width");
    if (!(sigma >= 0.0f) || !std::isfinite(sigma)) return fail(BAD_INDEX, "sigma must be finite and >= 0");
    DeviceGuard guard(p->device);
    hipStream_t st = (hipStream_t)stream;
    set_run(p, 0, in, in_pitch, out, out_pitch, sigma);
    int rc = fill_staging(p);
    if (rc) return rc;
    HIP_CHECK(hipMemcpyAsync(p->tab_dev, p->staging, p->staging_bytes, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipEventRecord(p->staged, st));
    p->staged_pending = true;
    p->ran = true;
    if ((rc = flag_begin(p, st))) return rc;
    const dim3 g((p->full.rect_w + 255) / 256, p->full.rect_h, p->full.n_rects);
    k_colour_in<<<g, 256, 0, st>>>(p->tab_dev, p->full.rects);
    HIP_CHECK(hipGetLastError());
    for (int c = 0; c < 3; ++c)
        HIP_CHECK(hipMemcpy2DAsync(p->tab[0].bz[c], (size_t)F.w * sizeof(float), basic_in[c],
                                   basic_pitch * sizeof(float), (size_t)F.w * sizeof(float), F.h,
                                   hipMemcpyDeviceToDevice, st));
    RunCtx r{p, st, false};
    if ((rc = run_step(r, 2, p->full, TO_OUTPUT, 0, false))) return rc;
    return flag_end(p, st, false);
}

int cpipe_bm3d_plan_set_pilot(cpipe_bm3d_plan* p, cpipe_bm3d_pilot* net,
                              const cpipe_bm3d_pilot_opts* opts) {
    if (!p) return fail(BAD_INDEX, "null plan");
    if (net && p->profile) return fail(UNSUPPORTED, "the learned pilot takes a scalar noise level");
    DeviceGuard guard(p->device);
    HIP_CHECK(hipDeviceSynchronize());
    drop_graph(p);
    free_lists(p->chk);
    (void)hipFree(p->report);
    (void)hipFree(p->chk_part);
    (void)hipFree(p->chk_io);
    (void)hipFree(p->chk_levels);
    (void)hipFree(p->chk_ws);
    p->report = nullptr;
    p->chk_part = nullptr;
    p->chk_io = nullptr;
    p->chk_levels = nullptr;
    p->chk_ws = nullptr;
    for (auto& f : p->fr) {
        (void)hipFree(f.bzp);
        f.bzp = nullptr;
    }
    for (auto& t : p->tab) {
        t.bzp[0] = t.bzp[1] = t.bzp[2] = nullptr;
        t.chk0 = t.nchk = 0;
    }
    p->checks.clear();
    p->pilot = nullptr;
    (void)hipFree(p->pilot_ws);
    p->pilot_ws = nullptr;
    if (!net) return OK;
    if (p->c != 3) return fail(UNSUPPORTED, "the learned pilot is trained for RGB (3-plane) images");
    bool any = false;
    for (const auto& f : p->fr) any = any || !f.small;
    if (!any) return fail(UNSUPPORTED, "image smaller than one patch");
    cpipe_bm3d_pilot_opts o{sizeof(cpipe_bm3d_pilot_opts), 0.5f, 256, 16, 0.5f, 0.25f};
    if (opts) {
        constexpr size_t v1 = offsetof(cpipe_bm3d_pilot_opts, under_db);  // layout before under_db
        if (opts->struct_size < v1) return fail(BAD_INDEX, "opts.struct_size too small");
        std::memcpy(&o, opts, v1);
        o.under_db = opts->struct_size >= sizeof(cpipe_bm3d_pilot_opts) ? opts->under_db : 0.5f * o.fallback_db;
        o.struct_size = sizeof(cpipe_bm3d_pilot_opts);
    }
    if (!(o.fallback_db >= 0.0f) || !(o.under_db >= 0.0f) || o.check_block < 16 || o.check_every < 0)
        return fail(BAD_INDEX, "bad pilot options");
    int rc;
    size_t max_px = 0;
    for (int f = 0; f < (int)p->fr.size(); ++f) {
        auto& F = p->fr[f];
        if (F.small) continue;
        if ((rc = dev_alloc((void**)&F.bzp, sizeof(float) * 3 * F.npx))) return rc;
        for (int c = 0; c < 3; ++c) p->tab[f].bzp[c] = F.bzp + c * F.npx;
        max_px = std::max(max_px, (size_t)F.w * F.h);
    }
    // frames run the network one after another on the plan's stream, so one
    // workspace for the largest serves them all
    if ((rc = dev_alloc(&p->pilot_ws, cpipe_internal::pilot_workspace_bytes(max_px)))) return rc;
    if ((rc = build_check_lists(p, o))) return rc;
    if (!p->side) {
        HIP_CHECK(hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreateWithFlags(&p->fork_ev, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&p->join_ev, hipEventDisableTiming));
    }
    p->pilot = net;
    p->pilot_opts = o;
    p->pilot_thr_db = o.fallback_db;
    p->pilot_under_db = o.under_db;
    p->pilot_full_frac = o.full_fraction;
    return OK;
}

int cpipe_bm3d_plan_pilot_report(cpipe_bm3d_plan* p, int* n_checked, int* n_fallback,
                                 float* max_delta_db, int* full_fallback) {
    if (!p) return fail(BAD_INDEX, "null plan");
    int nc = 0, nf = 0, full = 0;
    float md = 0.0f;
    if (p->pilot) {
        DeviceGuard guard(p->device);
        if (p->ran) HIP_CHECK(hipEventSynchronize(p->done));
        nc = (int)p->checks.size();
        std::vector<float> rep(2 * std::max(1, nc));
        if (nc) HIP_CHECK(hipMemcpy(rep.data(), p->report, 2 * nc * sizeof(float), hipMemcpyDeviceToHost));
        for (int i = 0; i < nc; ++i) {
            md = std::max(md, std::fabs(rep[2 * i]));
            nf += rep[2 * i + 1] > 0.5f ? 1 : 0;
        }
        if (nc) {
            std::vector<int> g(p->fr.size());
            HIP_CHECK(hipMemcpy(g.data(), p->gates, sizeof(int) * g.size(), hipMemcpyDeviceToHost));
            for (size_t f = 0; f < g.size(); ++f)
                if (p->tab[f].nchk > 0) full += g[f] ? 1 : 0;
        }
    }
    if (n_checked) *n_checked = nc;
    if (n_fallback) *n_fallback = nf;
    if (max_delta_db) *max_delta_db = md;
    if (full_fallback) *full_fallback = full;
    return OK;
}

int cpipe_bm3d_plan_check(cpipe_bm3d_plan* p) {
    if (!p) return fail(BAD_INDEX, "null plan");
    if (!p->ran) return OK;
    DeviceGuard guard(p->device);
    HIP_CHECK(hipEventSynchronize(p->done));  // the last run only, not the device
    if (*p->err_host) return fail(INTERNAL, "an aggregation ordering wait timed out; results are invalid");
    return OK;
}

int cpipe_bm3d_plan_debug_flag(cpipe_bm3d_plan* p, void* stream) {
    if (!p) return fail(BAD_INDEX, "null plan");
    DeviceGuard guard(p->device);
    k_set_flag<<<1, 1, 0, (hipStream_t)stream>>>(p->err, 1);
    HIP_CHECK(hipGetLastError());
    return flag_end(p, (hipStream_t)stream, false);
}

int cpipe_bm3d_plan_ref_grid(const cpipe_bm3d_plan* p, int* nx, int* ny) {
    if (!p || !nx || !ny) return fail(BAD_INDEX, "null argument");
    *nx = p->fr[0].nx;
    *ny = p->fr[0].ny;
    return OK;
}

int cpipe_bm3d_plan_set_timing(cpipe_bm3d_plan* p, int enable) {
    if (!p) return fail(BAD_INDEX, "null plan");
    DeviceGuard guard(p->device);
    if (enable && !p->ev[0]) {
        for (auto& e : p->ev) HIP_CHECK(hipEventCreate(&e));
    }
    p->timing = enable != 0;
    for (auto& m : p->ms) m = 0.0f;
    p->runs = 0;
    return OK;
}

int cpipe_bm3d_plan_timing(cpipe_bm3d_plan* p, float* ms, int n, int* runs) {
    if (!p || !ms) return fail(BAD_INDEX, "null argument");
    for (int i = 0; i < n && i < CPIPE_BM3D_T_COUNT; ++i) ms[i] = p->ms[i];
    if (runs) *runs = p->runs;
    return OK;
}

int cpipe_bm3d_compat_rgba16f(const uint16_t* in, size_t isp, uint16_t* out, size_t osp, int w,
                              int h, float sigma, int stage, void* stream) {
    if (!in || !out) return fail(BAD_INDEX, "null image pointer");
    if (w <= 0 || h <= 0) return fail(BAD_INDEX, "empty image");
    if (isp < (size_t)w || osp < (size_t)w) return fail(BAD_INDEX, "stride smaller than width");
    if (stage != 1 && stage != 2) return fail(BAD_INDEX, "stage must be 1 or 2");
    const dim3 rowgrid((w + 255) / 256, h);
    k_compat_rgba16f<<<rowgrid, 256, 0, (hipStream_t)stream>>>(in, isp, out, osp, w, h, sigma, stage);
    HIP_CHECK(hipGetLastError());
    return OK;
}

int cpipe_bm3d_host_planes(const float* const* in, size_t in_pitch, float* const* out,
                           size_t out_pitch, int w, int h, int channels, float sigma, int precision,
                           int device, int flags) {
    if (!in || !out) return fail(BAD_INDEX, "null argument");
    if (channels != 1 && channels != 3) return fail(BAD_INDEX, "channels must be 1 or 3");
    if (w <= 0 || h <= 0) return fail(BAD_INDEX, "empty image");
    HostCache& hc = host_cache();
    std::lock_guard<std::mutex> lock(hc.mu);
    const size_t npx = (size_t)w * h, pb = npx * sizeof(float), row = (size_t)w * sizeof(float);
    HostSlot* sl = nullptr;
    int rc = host_slot(hc, HOST_PLANES, w, h, channels, precision, device, flags, pb * 2 * channels, &sl);
    if (rc || (rc = apply_profile(sl, flags))) return rc;
    DeviceGuard guard(device);
    hipStream_t st = nullptr;
    HIP_CHECK(hc.streams[device].get(&st));
    float* dev = static_cast<float*>(sl->dev);
    float* pin = static_cast<float*>(sl->pin);
    float* din[3];
    float* dout[3];
    for (int c = 0; c < channels; ++c) {
        din[c] = dev + c * npx;
        dout[c] = dev + (channels + c) * npx;
    }
    // plane c's DMA runs while plane c+1 is copied into the pinned staging
    for (int c = 0; c < channels; ++c) {
        par_rows(pin + c * npx, row, in[c], in_pitch * sizeof(float), row, h);
        HIP_CHECK(hipMemcpyAsync(din[c], pin + c * npx, pb, hipMemcpyHostToDevice, st));
    }
    if ((rc = cpipe_bm3d_plan_run(sl->plan, din, (size_t)w, dout, (size_t)w, sigma, st))) return rc;
    float* pout = pin + channels * npx;
    for (int c = 0; c < channels; ++c) {
        HIP_CHECK(hipMemcpyAsync(pout + c * npx, dout[c], pb, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipEventRecord(sl->ev[c], st));
    }
    for (int c = 0; c < channels; ++c) {
        HIP_CHECK(hipEventSynchronize(sl->ev[c]));
        par_rows(out[c], out_pitch * sizeof(float), pout + c * npx, row, row, h);
    }
    return cpipe_bm3d_plan_check(sl->plan);
}

int cpipe_bm3d_host_rgba16f(const uint16_t* in, size_t isp, uint16_t* out, size_t osp, int w, int h,
                            float sigma, int mode, int precision, int device, int flags) {
    if (!in || !out) return fail(BAD_INDEX, "null argument");
    if (w <= 0 || h <= 0) return fail(BAD_INDEX, "empty image");
    if (mode == 0 && (flags & (CPIPE_BM3D_HOST_PILOT | CPIPE_BM3D_HOST_PROFILE)))
        return fail(BAD_INDEX, "the learned pilot and the noise-profile map need mode bm3d");
    HostCache& hc = host_cache();
    std::lock_guard<std::mutex> lock(hc.mu);
    const size_t bytes = (size_t)w * h * 8;
    HostSlot* sl = nullptr;
    int rc = host_slot(hc, mode == 0 ? HOST_COMPAT : HOST_RGBA, w, h, 3, precision, device, flags,
                       2 * bytes, &sl);
    if (rc || (rc = apply_profile(sl, flags))) return rc;
    DeviceGuard guard(device);
    hipStream_t st = nullptr;
    HIP_CHECK(hc.streams[device].get(&st));
    uint16_t* din = static_cast<uint16_t*>(sl->dev);
    uint16_t* dout = din + (size_t)w * h * 4;
    uint16_t* pin = static_cast<uint16_t*>(sl->pin);
    uint16_t* pout = pin + (size_t)w * h * 4;
    const size_t row = (size_t)w * 8;
    par_rows(pin, row, in, isp * 8, row, h);
    HIP_CHECK(hipMemcpyAsync(din, pin, bytes, hipMemcpyHostToDevice, st));
    if (mode == 0) {
        rc = cpipe_bm3d_compat_rgba16f(din, (size_t)w, dout, (size_t)w, w, h, sigma, 1, st);
    } else {
        rc = cpipe_bm3d_plan_run_rgba16f(sl->plan, din, (size_t)w, dout, (size_t)w, sigma, st);
    }
    if (rc) return rc;
    HIP_CHECK(hipMemcpyAsync(pout, dout, bytes, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    par_rows(out, osp * 8, pout, row, row, h);
    return mode == 0 ? OK : cpipe_bm3d_plan_check(sl->plan);
}

int cpipe_bm3d_device_planes(const float* const* in, size_t in_pitch, float* const* out,
                             size_t out_pitch, int w, int h, int channels, float sigma, int precision,
                             int device, int flags, void* stream) {
    if (!in || !out) return fail(BAD_INDEX, "null argument");
    if (channels != 1 && channels != 3) return fail(BAD_INDEX, "channels must be 1 or 3");
    if (w <= 0 || h <= 0) return fail(BAD_INDEX, "empty image");
    HostCache& hc = host_cache();
    std::lock_guard<std::mutex> lock(hc.mu);
    HostSlot* sl = nullptr;
    int rc = host_slot(hc, HOST_PLANES, w, h, channels, precision, device, flags, 0, &sl);
    if (rc || (rc = apply_profile(sl, flags))) return rc;
    DeviceGuard guard(device);
    if ((rc = cpipe_bm3d_plan_run(sl->plan, in, in_pitch, out, out_pitch, sigma, stream))) return rc;
    HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
    return cpipe_bm3d_plan_check(sl->plan);
}

int cpipe_bm3d_device_rgba16f(const uint16_t* in, size_t isp, uint16_t* out, size_t osp, int w, int h,
                              float sigma, int mode, int precision, int device, int flags, void* stream) {
    if (!in || !out) return fail(BAD_INDEX, "null argument");
    if (w <= 0 || h <= 0) return fail(BAD_INDEX, "empty image");
    if (mode == 0 && (flags & (CPIPE_BM3D_HOST_PILOT | CPIPE_BM3D_HOST_PROFILE)))
        return fail(BAD_INDEX, "the learned pilot and the noise-profile map need mode bm3d");
    HostCache& hc = host_cache();
    std::lock_guard<std::mutex> lock(hc.mu);
    HostSlot* sl = nullptr;
    int rc = host_slot(hc, mode == 0 ? HOST_COMPAT : HOST_RGBA, w, h, 3, precision, device, flags, 0, &sl);
    if (rc || (rc = apply_profile(sl, flags))) return rc;
    DeviceGuard guard(device);
    rc = mode == 0 ? cpipe_bm3d_compat_rgba16f(in, isp, out, osp, w, h, sigma, 1, stream)
                   : cpipe_bm3d_plan_run_rgba16f(sl->plan, in, isp, out, osp, sigma, stream);
    if (rc) return rc;
    HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
    return mode == 0 ? OK : cpipe_bm3d_plan_check(sl->plan);
}

int cpipe_bm3d_host_noise_profile(const float* beta, const float* lambda, int n) {
    if (n < 0 || n > 3) return fail(BAD_INDEX, "noise profile: 0 to 3 (beta, lambda) pairs");
    if (n > 0 && (!beta || !lambda)) return fail(BAD_INDEX, "null noise profile");
    for (int c = 0; c < n; ++c) {
        if (!std::isfinite(beta[c]) || !std::isfinite(lambda[c]))
            return fail(BAD_INDEX, "noise profile values must be finite");
        g_prof_beta[c] = beta[c];
        g_prof_lambda[c] = lambda[c];
    }
    g_prof_n = n;
    return OK;
}

void cpipe_bm3d_host_release(void) {
    HostCache& hc = host_cache();
    std::lock_guard<std::mutex> lock(hc.mu);
    for (HostSlot& sl : hc.slots) drop_slot(sl);
    hc.slots.clear();
    for (auto& kv : hc.nets) {
        DeviceGuard guard(kv.first);
        cpipe_bm3d_pilot_destroy(kv.second);
    }
    hc.nets.clear();
}

int cpipe_bm3d_noalloc_planes(const float* const* in, size_t in_pitch, float* const* out,
                              size_t out_pitch, int w, int h, int channels, float sigma, void* stream) {
    if (!in || !out) return fail(BAD_INDEX, "null argument");
    if (channels != 1 && channels != 3) return fail(BAD_INDEX, "channels must be 1 or 3");
    for (int c = 0; c < channels; ++c)
        if (!in[c] || !out[c]) return fail(BAD_INDEX, "null plane");
    if (w <= 0 || h <= 0) return fail(BAD_INDEX, "empty image");
    if (w > CPIPE_BM3D_NOALLOC_MAX_SIDE || h > CPIPE_BM3D_NOALLOC_MAX_SIDE ||
        (size_t)w * h > (size_t)CPIPE_BM3D_NOALLOC_MAX_PIXELS)
        return fail(BAD_INDEX, "cpipe_bm3d_noalloc_planes: frame " + std::to_string(w) + "x" + std::to_string(h) +
                                   " exceeds its maximum (width x height <= 1048576 pixels, sides <= 8192)");
    if (in_pitch < (size_t)w || out_pitch < (size_t)w) return fail(BAD_INDEX, "pitch smaller than width");
    if (!(sigma >= 0.0f) || !std::isfinite(sigma)) return fail(BAD_INDEX, "sigma must be finite and >= 0");
    const NoallocLayout L = noalloc_layout(w, h, channels);
    if (L.total > NOALLOC_BYTES) return fail(INTERNAL, "no-allocation scratch too small for an admitted frame");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return fail(UNSUPPORTED, "no HIP device visible");
    if (dev >= 64) return fail(BAD_INDEX, "device ordinal out of range");
    int rc = ensure_tables(dev);
    if (rc) return rc;
    unsigned char* base = nullptr;
    if ((rc = noalloc_base(dev, &base))) return rc;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    NoallocSetup S{};
    S.base = base;
    S.L = L;
    S.w = w;
    S.h = h;
    S.c = channels;
    S.sigma = sigma;
    if (channels == 1) {
        S.sig[0] = S.sig[1] = S.sig[2] = sigma;
    } else {  // the opponent channels' sigmas, as the plans set them
        S.sig[0] = sigma * (float)std::sqrt(1.0 / 3.0);
        S.sig[1] = sigma * (float)std::sqrt(0.5);
        S.sig[2] = sigma * (float)std::sqrt(0.375);
    }
    for (int c = 0; c < 3; ++c) {
        S.in[c] = c < channels ? in[c] : nullptr;
        S.out[c] = c < channels ? out[c] : nullptr;
    }
    S.in_pitch = in_pitch;
    S.out_pitch = out_pitch;
    k_noalloc_setup<<<1, 256, 0, st>>>(S);
    HIP_CHECK(hipGetLastError());
    const FrameTab* tab = reinterpret_cast<const FrameTab*>(base + L.tab);
    const FrameRect* rects = reinterpret_cast<const FrameRect*>(base + L.rects);
    if (L.small) {  // smaller than one patch: copied through
        k_copy_through<<<dim3((w + 255) / 256, h, 1), 256, 0, st>>>(tab, rects);
        HIP_CHECK(hipGetLastError());
        return OK;
    }
    const int4* bm = reinterpret_cast<const int4*>(base + L.bm);
    const int2* grp = reinterpret_cast<const int2*>(base + L.grp);
    int* err = reinterpret_cast<int*>(base + L.flags);
    const dim3 pg((w + 255) / 256, NOALLOC_BAND, L.n_rects);
    k_colour_in<<<pg, 256, 0, st>>>(tab, rects);
    for (int step = 1; step <= 2; ++step) {
        launch_block_match(L.n_bm, tab, bm, step, 0, st);
        const GroupLaunch gl{tab, grp, 0, err};
        const dim3 gg(L.n_grp, channels);
        if (step == 1)
            k_group<1, EXACT_TY, EXACT_NW, false><<<gg, 64 * EXACT_NW, 0, st>>>(gl);
        else
            k_group<2, EXACT_TY, EXACT_NW, false><<<gg, 64 * EXACT_NW, 0, st>>>(gl);
        k_merge<<<pg, 256, 0, st>>>(tab, rects, step == 1 ? TO_BASIC : TO_OUTPUT, 0);
    }
    HIP_CHECK(hipGetLastError());
    return OK;
}

const void* cpipe_bm3d_pilot_embedded(size_t* size) {
    if (size) *size = CPIPE_BM3D_PILOT_BLOB_SIZE;
    return CPIPE_BM3D_PILOT_BLOB;
}

}  // extern "C"
