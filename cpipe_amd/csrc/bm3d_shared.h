// Definitions shared by the BM3D engine's translation units (plan + exact FP32
// kernels in bm3d_device.hip, mixed-precision group kernels in group_mx.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace cpipe_internal {

constexpr int K = 8;          // patch edge
constexpr int R = 19;         // search radius (39x39 window)
constexpr int STRIDE = 3;     // reference stride
constexpr int NMAX = 16;      // max group size
constexpr int TILE = 16;      // refs per tile row (tile width); tile height 4 (exact FP32), 8 or 16 (mixed)
// Footprint edge of a tile of t refs: every pixel a member patch of one of its
// groups can touch (t refs at stride 3, +- 19 px search, + 8 px patch).
constexpr int foot_edge(int t) { return (t - 1) * STRIDE + 2 * R + K; }
constexpr int FOOT = foot_edge(TILE);  // 91
constexpr int FOOT2 = FOOT * FOOT;     // largest footprint (16 x 16 refs)

// Patch SSD in chunks of 3, 3 and 2, along both axes. A row sum is
// (sq3(e0..e2) + sq3(e3..e5)) + sq2(e6, e7), with sq3 = e0^2 + e1^2 + e2^2 as one
// product and two fmas (its first two terms are sq2); the 8 row sums r0..r7 add
// up as (((r0 + r1) + r2) + ((r3 + r4) + r5)) + (r6 + r7). Refs 3 columns or 3
// rows apart share chunks (the block matching kernels rely on it), and the CPU
// FP32 restatement sums in the same order.
__device__ __forceinline__ float sq2(float e0, float e1) { return __builtin_fmaf(e1, e1, e0 * e0); }
__device__ __forceinline__ float sq3(float e0, float e1, float e2) { return __builtin_fmaf(e2, e2, sq2(e0, e1)); }
__device__ __forceinline__ float ssd_row(const float (&e)[8]) {
    return (sq3(e[0], e[1], e[2]) + sq3(e[3], e[4], e[5])) + sq2(e[6], e[7]);
}

// Rank key of block matching's first stage: the SSD's bit pattern with its
// low 11 bits replaced by the position code (dy + R) * 39 + (dx + R) (< 1521),
// i.e. the SSD rounded down to 12 explicit mantissa bits, then the window
// position in scan order. For SSD >= 0 keys order like the floats they are,
// so a sorted list of keys is kept with one v_med3 per slot and the position
// rides along; the kept candidates are then re-ranked by exact (SSD,
// position), as in the CPU FP32 restatement.
constexpr int BM_WIN = 2 * R + 1;            // 39
constexpr uint32_t BM_CODE_MASK = 0x7ffu;    // codes < 39 * 39 fit 11 bits
__device__ __forceinline__ float bm_key(float ssd, uint32_t code) {
    return __uint_as_float((__float_as_uint(ssd) & ~BM_CODE_MASK) | code);
}
// The smallest value above every SSD whose key can precede `key` (the end of
// its 2^-12 bucket): a loose filter for candidates before insertion. +inf
// (an empty slot) stays +inf.
__device__ __forceinline__ float bm_key_bucket_end(float key) {
    return __uint_as_float(min(__float_as_uint(key) | BM_CODE_MASK, 0x7f7fffffu) + 1u);
}

// Per-frame parameters in device memory (one entry per frame of a plan or
// batch). Geometry and buffers are fixed at creation; sigma, noise level and
// the caller's input/output planes are rewritten before every run (one small
// host-to-device copy), so a captured graph replays with new frames and sigmas.
struct FrameTab {
    int w, h, c, nx, ny, ntx, nty, tile_y;
    int dofs;               // den offset in a tile's partial block (FOOT x foot height)
    int chk0, nchk;         // this frame's spot checks: report entries [chk0, chk0 + nchk)
    int pad0;
    size_t tile_stride;     // floats per (channel, tile) partial block
    float* z[3];            // noisy opponent planes (pitch w); unused channels alias z[0]
    float* bz[3];           // basic estimate planes
    float* bzp[3];          // plain step-1 planes on the checked rectangles (learned pilot)
    uint32_t* pos[2];       // match lists of steps 1 and 2
    uint8_t* cnt[2];
    float* partial;         // per (channel, tile) num/den footprints
    const int* txlo;        // per column: first / last tile whose footprint covers it
    const int* txhi;
    const int* tylo;        // per row
    const int* tyhi;
    const int* fx0;         // per tile column / row: footprint bounds
    const int* fx1;
    const int* fy0;
    const int* fy1;
    int* gate;              // learned pilot: 1 = run the plain step 1 on the whole frame
    // per run
    float sig[3];           // per-channel sigma (opponent: sigma * |row|)
    float sigma;
    float* smap;            // noise-profile map: per (channel, ref) sigma, or null (scalar sig[])
    float nbeta[3];         // noise profile of the input channels: variance nbeta + nlambda * x
    float nlambda[3];
    const float* in[3];     // caller's input planes
    float* out[3];          // caller's output planes
    size_t in_pitch, out_pitch;
};

// A rectangle of one frame (merge bands, spot-check blocks).
struct FrameRect {
    int f, x0, y0, x1, y1, pad0, pad1, pad2;
};

// Arguments of the group kernels: one workgroup per (item, channel); item =
// (frame, tile index ty * ntx + tx).
struct GroupLaunch {
    const FrameTab* tab;
    const int2* items;
    int use_gate;           // run a frame's groups only if its gate is set
    int* err;               // set to 1 if an ordering wait timed out (never expected)
};

// The per-frame scalars a group kernel needs, read once per workgroup (the
// per-channel planes and sigmas are read from the table by channel index).
struct GroupView {
    int w, h, nx, ny, ntx, nty, dofs;
    size_t tile_stride;
    const uint32_t* pos;
    const uint8_t* cnt;
    const int* fx0;
    const int* fx1;
    const int* fy0;
    const int* fy1;
    float* partial;
};

__device__ __forceinline__ GroupView group_view(const FrameTab& F, int step) {
    GroupView v;
    v.w = F.w;
    v.h = F.h;
    v.nx = F.nx;
    v.ny = F.ny;
    v.ntx = F.ntx;
    v.nty = F.nty;
    v.dofs = F.dofs;
    v.tile_stride = F.tile_stride;
    v.pos = F.pos[step - 1];
    v.cnt = F.cnt[step - 1];
    v.fx0 = F.fx0;
    v.fx1 = F.fx1;
    v.fy0 = F.fy0;
    v.fy1 = F.fy1;
    v.partial = F.partial;
    return v;
}

// Loads through the scalar cache: a pointer in the constant address space is
// read with scalar loads when the address is wave-uniform (the data must not
// change during the kernel).
#define CONST_AS __attribute__((address_space(4)))

// The match count of ref r (one byte per ref) through the scalar cache: the
// aligned dword that holds it. Count buffers start 256-byte aligned and span
// whole dwords (plans round their size up; the static no-allocation layout
// rounds every buffer to 256 bytes), so that dword lies inside the buffer.
__device__ __forceinline__ int load_count(const uint8_t* cnt, size_t r) {
    const uintptr_t a = (uintptr_t)(cnt + r);
    const uint32_t CONST_AS* wp = (const uint32_t CONST_AS*)(a & ~(uintptr_t)3);
    return (int)((*wp >> (8u * (uint32_t)(a & 3u))) & 0xffu);
}

// A buffer resource over one plane of `bytes` (< 4 GiB; plans refuse larger
// planes): a load takes the lane's 32-bit byte offset in a VGPR and a
// wave-uniform byte offset in an SGPR, so it needs no address arithmetic.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const float* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p), (short)0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ float plane_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// One read/add/write round over members [S, S + L) of a group (pairwise
// disjoint patches): all reads, then all writes.
// o[m]: byte offset of the lane's (num, den) pair for member m.
template <int S, int L>
__device__ __forceinline__ void rmw_run(float2* acc, const int (&o)[NMAX], const float (&c)[NMAX],
                                        float wk) {
    char* const base = reinterpret_cast<char*>(acc);
    float2 a[L];
#pragma unroll
    for (int m = 0; m < L; ++m) a[m] = *reinterpret_cast<const float2*>(base + o[S + m]);
#pragma unroll
    for (int m = 0; m < L; ++m)
        *reinterpret_cast<float2*>(base + o[S + m]) = make_float2(a[m].x + c[S + m], a[m].y + wk);
}

// Members m0 .. m0+3 in rounds of consecutive pairwise-disjoint runs; bit k of
// `brk` (k = 1..3) starts a new run at member m0 + k.
template <int M0>
__device__ __forceinline__ void rmw_block4(float2* acc, const int (&o)[NMAX], const float (&c)[NMAX],
                                           float wk, uint32_t brk) {
    switch (brk & 0xeu) {
        case 0x0: rmw_run<M0, 4>(acc, o, c, wk); break;
        case 0x8: rmw_run<M0, 3>(acc, o, c, wk); rmw_run<M0 + 3, 1>(acc, o, c, wk); break;
        case 0x4: rmw_run<M0, 2>(acc, o, c, wk); rmw_run<M0 + 2, 2>(acc, o, c, wk); break;
        case 0xc:
            rmw_run<M0, 2>(acc, o, c, wk);
            rmw_run<M0 + 2, 1>(acc, o, c, wk);
            rmw_run<M0 + 3, 1>(acc, o, c, wk);
            break;
        case 0x2: rmw_run<M0, 1>(acc, o, c, wk); rmw_run<M0 + 1, 3>(acc, o, c, wk); break;
        case 0xa:
            rmw_run<M0, 1>(acc, o, c, wk);
            rmw_run<M0 + 1, 2>(acc, o, c, wk);
            rmw_run<M0 + 3, 1>(acc, o, c, wk);
            break;
        case 0x6:
            rmw_run<M0, 1>(acc, o, c, wk);
            rmw_run<M0 + 1, 1>(acc, o, c, wk);
            rmw_run<M0 + 2, 2>(acc, o, c, wk);
            break;
        default:
            rmw_run<M0, 1>(acc, o, c, wk);
            rmw_run<M0 + 1, 1>(acc, o, c, wk);
            rmw_run<M0 + 2, 1>(acc, o, c, wk);
            rmw_run<M0 + 3, 1>(acc, o, c, wk);
            break;
    }
}

// Ordered aggregation of filtered groups into a tile's LDS footprint of
// (num, den) pairs (one 8-byte read and write per member pixel). Lane p = (i, j)
// holds pixel (i, j) of member m in v[m]; mp[m] (wave-uniform) is the member's
// footprint position (y << 16 | x); lo = (i - fy0) * fw + (j - fx0). Group g is
// accumulated right after group g-1: the wave waits for the LDS turn token to
// reach g, adds num += wk * v and den += wk, then passes the token on. Every
// wave takes its turns in increasing g and turn g-1 never waits on turn g, so
// the chain always completes; the spin is bounded anyway and reports a timeout
// instead of hanging.
//
// The turns form the kernel's critical path, so everything that does not
// touch the footprint (offsets, products, the batching plan) is computed
// before the wait (prepare_group). Members are added in read/add/write rounds:
// each block of four consecutive members is cut greedily into runs of
// consecutive members whose patches are pairwise disjoint, one round per run. A
// pixel's members are added in member order either way (members covering one
// pixel overlap, so they never share a run).
struct PreparedGroup {
    int o[NMAX];    // byte offsets of the lane's (num, den) pair per member
    float c[NMAX];
    float wk;
    uint32_t plan;  // run breaks: bit 4b + k set when member 4b + k starts a new run (k = 1..3)
    int n;
    bool has;       // false: nothing to add (no group this round)
};

// Bit m (m < 16) set when member m's patch overlaps member m - D's. Lane m
// (mod 16) holds member m's position pv = (y << 16) | x; x7 = x + 7, y7 = y + 7
// of the same lane. One DPP row shift brings member m - D's position.
template <int D>
__device__ __forceinline__ uint32_t overlap_prev(uint32_t pv, uint32_t x7, uint32_t y7) {
    const uint32_t pk = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)pv, 0x110 + D, 0xf, 0xf, false);  // row_shr:D
    const uint32_t dx = x7 - (pk & 0xffffu), dy = y7 - (pk >> 16);
    return (uint32_t)__ballot(dx < 15u && dy < 15u) & 0xffffu;
}

// mp: the members' positions (wave-uniform, footprint-relative offsets are
// y * fw + x with fw wave-uniform); pv: lane m (mod 16) holds mp[m]. The run
// plan of each block of four members comes from three lane-parallel overlap
// masks (member m against m - 1, m - 2, m - 3).
__device__ __forceinline__ void prepare_group(PreparedGroup& P, bool has, int n, const uint32_t (&mp)[NMAX],
                                              uint32_t pv, const float (&v)[NMAX], float wk, int lo, int fw) {
    P.has = has;
    P.n = n;
    P.wk = wk;
    const int lo8 = lo * (int)sizeof(float2);
#pragma unroll
    for (int m = 0; m < NMAX; ++m) {
        // the member's corner as a scalar byte offset, plus the lane's pixel
        P.o[m] = ((int)(mp[m] >> 16) * fw + (int)(mp[m] & 0xffffu)) * (int)sizeof(float2) + lo8;
        P.c[m] = wk * v[m];
    }
    const uint32_t x7 = (pv & 0xffffu) + 7u, y7 = (pv >> 16) + 7u;
    const uint32_t A = overlap_prev<1>(pv, x7, y7), B = overlap_prev<2>(pv, x7, y7),
                   C = overlap_prev<3>(pv, x7, y7);
    // per block m0 = 4b: b1 = A(m0+1); b2 = A(m0+2) | (!b1 & B(m0+2));
    // b3 = A(m0+3) | (!b2 & B(m0+3)) | (!b1 & !b2 & C(m0+3))
    const uint32_t B1 = A & 0x2222u;
    const uint32_t B2 = (A | (B & ~(B1 << 1))) & 0x4444u;
    const uint32_t B3 = (A | (B & ~(B2 << 1)) | (C & ~(B1 << 2) & ~(B2 << 1))) & 0x8888u;
    uint32_t plan = B1 | B2 | B3;
    if (n < 4) plan = n == 2 ? (plan & 2u) | 0xcu : 0xeu;  // n = 1: one member; n = 2: at most 2
    P.plan = __builtin_amdgcn_readfirstlane(plan);
    // keep the precomputation ahead of the wait
#pragma unroll
    for (int m = 0; m < NMAX; ++m) asm volatile("" : "+v"(P.o[m]), "+v"(P.c[m]));
    asm volatile("" : "+s"(P.plan));
}

// Waits (bounded) until the token reaches g; true if the wait timed out.
__device__ __forceinline__ bool wait_turn(int* turn, int g) {
    int spins = 0;
    // every lane reads the same token: compare it as a scalar
    while (__builtin_amdgcn_readfirstlane(
               __hip_atomic_load(turn, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)) != g) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > (1 << 24)) return true;
    }
    return false;
}

__device__ __forceinline__ void pass_turn(int* turn, int g) {
    __hip_atomic_store(turn, g, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void rmw_group(float2* acc, const PreparedGroup& P) {
    if (!P.has) return;
    const int n = P.n;
    if (n >= 4) {
        rmw_block4<0>(acc, P.o, P.c, P.wk, P.plan);
        if (n >= 8) {
            rmw_block4<4>(acc, P.o, P.c, P.wk, P.plan >> 4);
            if (n >= 16) {
                rmw_block4<8>(acc, P.o, P.c, P.wk, P.plan >> 8);
                rmw_block4<12>(acc, P.o, P.c, P.wk, P.plan >> 12);
            }
        }
    } else if (n == 2) {
        if (P.plan & 2u) {
            rmw_run<0, 1>(acc, P.o, P.c, P.wk);
            rmw_run<1, 1>(acc, P.o, P.c, P.wk);
        } else {
            rmw_run<0, 2>(acc, P.o, P.c, P.wk);
        }
    } else {
        rmw_run<0, 1>(acc, P.o, P.c, P.wk);
    }
}

// The turn of group g: wait for the token, add, pass it on. The wave raises
// its issue priority for the turn, so the chain's LDS reads and writes are not
// queued behind the filtering waves' instructions (3 % off the exact
// hard-threshold kernel, results unchanged).
__device__ __forceinline__ bool take_turn(float2* acc, int* turn, int g, const PreparedGroup& P) {
    __builtin_amdgcn_s_setprio(2);
    const bool timed_out = wait_turn(turn, g);
    rmw_group(acc, P);
    pass_turn(turn, g + 1);
    __builtin_amdgcn_s_setprio(0);
    return timed_out;
}

// One group per turn (token g -> g + 1). `has` false: only pass the token.
__device__ __forceinline__ bool add_group_in_turn(float2* acc, int* turn, int g, bool has, int n,
                                                  const uint32_t (&mp)[NMAX], uint32_t pv,
                                                  const float (&v)[NMAX], float wk, int lo,
                                                  int fw) {
    PreparedGroup P;
    prepare_group(P, has, n, mp, pv, v, wk, lo, fw);
    return take_turn(acc, turn, g, P);
}

__device__ __forceinline__ int floor_pow2(int c) {
    return c >= 16 ? 16 : c >= 8 ? 8 : c >= 4 ? 4 : c >= 2 ? 2 : 1;
}

// Mixed-precision group kernels (group_mx.hip). group_mx_tables uploads the
// 8x8 DCT-II matrix C[u][i] and the 8x8 Kaiser window (row-major, FP32) to the
// current device, once per device; group_mx_launch enqueues the step-1 or
// step-2 group kernel with FP16 or BF16 MFMA transforms on the launch grid
// (items, channels) for tiles of 16 x tile_y refs (tile_y = 8 or 16), with the
// per-ref sigma of the frame's noise-profile map when `map` is set.
// Both return 0 or a hipError_t.
int group_mx_tables(int device, const float* dct64, const float* kaiser64);
int group_mx_launch(int step, int precision, int tile_y, bool map, const GroupLaunch& L, dim3 grid,
                    hipStream_t st);

// Mixed-precision block matching (bm_mx.hip): FP16 or BF16 MFMA cross terms,
// exact FP32 re-rank of a 48-candidate shortlist per ref. Same work items as
// the exact kernel (frame, 16-ref column block, 14-ref row block). Returns 0
// or a hipError_t.
int bm_mx_launch(int precision, const FrameTab* tab, const int4* items, int n_items, int step,
                 float tau, int use_gate, hipStream_t st);

}  // namespace cpipe_internal
