// Learned pilot for BM3D step 2: a 6-layer 3x3 convolutional network that
// maps the noisy opponent planes (+ a noise-level channel) to the basic
// estimate, replacing BM3D step 1. Forward, backward and Adam are HIP kernels
// for gfx950; convolutions run on MFMA (FP16 operands, FP32 accumulation).
//
// Layers (channels padded to multiples of 16 for the 16x16x16 MFMA):
//   L1 conv 4(->16) -> 32, ReLU     input: Y, U, V, 10*sigma
//   L2..L5 conv 32 -> 32, ReLU
//   L6 conv 32 -> 3(->16), output = input[0..2] + L6 (residual)
// Activations: NHWC FP16, channel stride = padded channel count.
// Weights: W[tap 0..8][cin][cout] FP32 master copy + FP16 copy, bias FP32.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "cpipe_bm3d.h"
#include "pilot.h"

namespace pilot {

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));

constexpr int TW = 64;  // output tile width (pixels) per workgroup
constexpr int TH = 4;   // weight-gradient tile height
constexpr int CTH = 8;  // convolution output tile height: two rows per wave

// Layer geometry (padded channel counts).
constexpr int NL = 6;
constexpr int CIN[NL] = {16, 32, 32, 32, 32, 32};
constexpr int COUT[NL] = {32, 32, 32, 32, 32, 16};
constexpr int CIN_REAL[NL] = {4, 32, 32, 32, 32, 32};
constexpr int COUT_REAL[NL] = {32, 32, 32, 32, 32, 3};

inline size_t layer_w(int l) { return (size_t)9 * CIN[l] * COUT[l]; }
inline size_t layer_params(int l) { return layer_w(l) + COUT[l]; }
inline size_t total_params() {
    size_t n = 0;
    for (int l = 0; l < NL; ++l) n += layer_params(l);
    return n;
}

// ---------------------------------------------------------------------------
// Packing: opponent planes (FP32) + noise level -> NHWC FP16 (16 channels)
// ---------------------------------------------------------------------------

__global__ void k_pack_input(const float* __restrict__ y, const float* __restrict__ u,
                             const float* __restrict__ v, size_t pitch, float level,
                             const float* __restrict__ levels, int w, int h,
                             size_t img_stride_planes, _Float16* __restrict__ out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int yy = blockIdx.y;
    const int img = blockIdx.z;
    if (x >= w) return;
    const size_t i = img * img_stride_planes + (size_t)yy * pitch + x;
    _Float16* o = out + ((size_t)img * w * h + (size_t)yy * w + x) * 16;
    typedef _Float16 half8 __attribute__((ext_vector_type(8)));
    const float lv = levels ? levels[img] : level;
    half8 a = {(_Float16)y[i], (_Float16)u[i], (_Float16)v[i], (_Float16)lv, 0, 0, 0, 0};
    half8 b = {0, 0, 0, 0, 0, 0, 0, 0};
    *reinterpret_cast<half8*>(o) = a;
    *reinterpret_cast<half8*>(o + 8) = b;
}

// ---------------------------------------------------------------------------
// 3x3 convolution, stride 1, zero padding, on MFMA 16x16x16 (FP16 -> FP32).
// Workgroup: 64x8 output pixels, 4 waves (two adjacent output rows each, both
// accumulated at once so that each A fragment read from LDS feeds every row
// whose window covers it), 4 segments of 16 pixels per row. The input tile
// (66x10 pixels) is staged in LDS with a padded pixel pitch (CIN + 4 halves)
// to spread banks, every global load of a thread in flight before its first
// LDS store; weights are loaded as pre-arranged B fragments; outputs leave
// through a per-wave LDS buffer as 16-byte stores.
// MFMA operand layouts (wave64, 16x16x16): A[i][k] at lane i + 16*(k/4),
// element k%4; B[k][j] at lane j + 16*(k/4), element k%4; D[i][j] at lane
// j + 16*(i/4), element i%4.
// EPI 0: out = relu(acc + bias) (FP16 NHWC)
// EPI 1: out planes (FP32) = residual planes + acc + bias, channels < 3
// EPI 2: out = acc * (mask > 0) (FP16 NHWC) -- backward through a ReLU
// ---------------------------------------------------------------------------

struct ConvArgs {
    const _Float16* in;    // NHWC, channel stride CIN, images stacked
    const _Float16* frag;  // weights as B fragments [tap][kb][nb][lane][e] (see k_frag)
    const float* bias;     // [cout] (EPI 0/1)
    _Float16* out;         // NHWC, channel stride COUT (EPI 0/2)
    const _Float16* mask;  // NHWC, channel stride COUT (EPI 2)
    const float* res[3];   // residual planes (EPI 1)
    float* outp[3];        // output planes (EPI 1)
    size_t pitch;          // plane row pitch (EPI 1)
    size_t img_stride_planes;
    const int* gate;       // inference: skip the launch's work once *gate is set (null: never)
    int w, h;
};

// B fragments of a layer's weights, in the order the convolution reads them.
// A layer operator with K = KI input channels uses MFMA 16x16x32 when KI is a
// multiple of 32 (fragments of 8 halves: k = 32 kb + 8 (lane / 16) + e), else
// 16x16x16 (4 halves: k = 16 kb + 4 (lane / 16) + e); j = 16 nb + lane % 16.
// frag[((t * KB + kb) * NB + nb) * 64 + lane] holds one lane's halves, so each
// fragment is one coalesced load per wave. Forward operator (CI -> CO):
// B[k = cin][j = cout] = W[t][k][j]; backward operator (CO -> CI, flipped):
// B[k = cout][j = cin] = W[8 - t][j][k].
constexpr int kchunk(int ki) { return ki % 32 == 0 ? 32 : 16; }

template <int KI, int NJ, bool BWD, int CI, int CO>
__device__ __forceinline__ void frag_one(const _Float16* __restrict__ w16, _Float16* __restrict__ dst, int idx) {
    constexpr int KC = kchunk(KI), E = KC / 4, KB = KI / KC, NB = NJ / 16;
    if (idx >= 9 * KB * NB * 64) return;
    const int lane = idx & 63, lr = lane & 15, lq = lane >> 4;
    const int f = idx >> 6, nb = f % NB, kb = (f / NB) % KB, t = f / (NB * KB);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int k = kb * KC + lq * E + e, j = nb * 16 + lr;
        dst[(size_t)idx * E + e] = BWD ? w16[((size_t)(8 - t) * CI + j) * CO + k]
                                       : w16[((size_t)t * CI + k) * CO + j];
    }
}

template <int CI, int CO>
__global__ void k_frag(const _Float16* __restrict__ w16, _Float16* __restrict__ fwd,
                       _Float16* __restrict__ bwd) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    frag_one<CI, CO, false, CI, CO>(w16, fwd, idx);
    frag_one<CO, CI, true, CI, CO>(w16, bwd, idx);
}

// A frame whose spot checks fell back discards the network's output, so its
// layers skip their work once the frame's gate is set (the checks decide on
// another stream while the network runs; the gate only goes from 0 to 1
// during a run). One read per workgroup, shared through LDS, so that every
// wave of the workgroup takes the same branch around its barriers.
__device__ __forceinline__ bool gate_set(const int* gate) {
    __shared__ int skip;
    if (threadIdx.x == 0) skip = __hip_atomic_load(gate, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return skip != 0;
}

template <int CI, int CO, int EPI>
__global__ __launch_bounds__(256) void k_conv3x3(ConvArgs a) {
    constexpr int PS = CI + 4;  // LDS pixel pitch in halves (8-byte aligned pixels)
    constexpr int RPW = CTH / 4;  // adjacent output rows per wave
    __shared__ __align__(16) _Float16 tile[(CTH + 2) * (TW + 2) * PS];
    __shared__ __align__(16) _Float16 ob[4][16 * CO];  // per wave: 16 output pixels x CO
    // EPI 1: per wave, 16 pixels x 3 output channels (acc + bias) before the residual add
    __shared__ float ofp[EPI == 1 ? 4 * 3 * 16 : 1];
    if (a.gate && gate_set(a.gate)) return;  // uniform: the kernel argument and the shared read
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * CTH, img = blockIdx.z;
    const int w = a.w, h = a.h;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    const _Float16* in = a.in + (size_t)img * w * h * CI;
    typedef _Float16 half8 __attribute__((ext_vector_type(8)));
    // weights as B fragments (forward or backward operator, see k_frag); their
    // loads are issued first so that they overlap the tile's
    constexpr int KC = kchunk(CI), E = KC / 4, KB = CI / KC, NB = CO / 16;
    typedef _Float16 hvec __attribute__((ext_vector_type(E)));
    hvec bw[9][KB][NB];
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int kb = 0; kb < KB; ++kb)
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
                bw[t][kb][nb] = *reinterpret_cast<const hvec*>(a.frag + (size_t)(((t * KB + kb) * NB + nb) * 64 + lane) * E);
    // stage (TW+2) x (CTH+2) pixels, CI channels: every 16-byte load of the
    // thread is in flight before the first LDS store
    constexpr int OPP = CI / 8;  // 16-byte chunks per pixel
    constexpr int NCH = (CTH + 2) * (TW + 2) * OPP;
    constexpr int IT = (NCH + 255) / 256;
    half8 sv[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int idx = threadIdx.x + 256 * i;
        const int q = idx % OPP, pix = idx / OPP;
        const int ty = pix / (TW + 2), tx = pix - ty * (TW + 2);
        const int gx = x0 + tx - 1, gy = y0 + ty - 1;
        sv[i] = (half8){0, 0, 0, 0, 0, 0, 0, 0};
        if (idx < NCH && gx >= 0 && gx < w && gy >= 0 && gy < h)
            sv[i] = *reinterpret_cast<const half8*>(in + ((size_t)gy * w + gx) * CI + q * 8);
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int idx = threadIdx.x + 256 * i;
        if (idx < NCH) {
            const int q = idx % OPP, pix = idx / OPP;
            _Float16* d = &tile[pix * PS + q * 8];
            *reinterpret_cast<half4*>(d) = __builtin_shufflevector(sv[i], sv[i], 0, 1, 2, 3);
            *reinterpret_cast<half4*>(d + 4) = __builtin_shufflevector(sv[i], sv[i], 4, 5, 6, 7);
        }
    }
    __syncthreads();
    _Float16* o16 = ob[wave];
    const int ly0 = wave * RPW;  // this wave's first output row in the tile
    // biases read once: the output stores could alias them as far as the
    // compiler knows, so a read inside the loop is repeated for every row
    float bias_l[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
        bias_l[nb] = (EPI == 0 || (EPI == 1 && nb == 0 && lr < 3)) ? a.bias[nb * 16 + lr] : 0.0f;

#pragma unroll 1
    for (int seg = 0; seg < TW / 16; ++seg) {
        // rows ly0 .. ly0+RPW-1 together: each A fragment (an input row at one
        // horizontal tap) feeds every output row whose 3x3 window covers it
        float4v acc[RPW][NB];
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) acc[r][nb] = (float4v){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int iy = 0; iy < RPW + 2; ++iy) {
#pragma unroll
            for (int kx = 0; kx < 3; ++kx) {
                const _Float16* row = &tile[((ly0 + iy) * (TW + 2) + seg * 16 + lr + kx) * PS];
#pragma unroll
                for (int kb = 0; kb < KB; ++kb) {
                    hvec av;
                    if constexpr (E == 8) {  // two 8-byte reads (pixels are 8-byte aligned)
                        const half4 lo = *reinterpret_cast<const half4*>(row + kb * KC + lq * E);
                        const half4 hi = *reinterpret_cast<const half4*>(row + kb * KC + lq * E + 4);
                        av = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
                    } else {
                        av = *reinterpret_cast<const hvec*>(row + kb * KC + lq * E);
                    }
#pragma unroll
                    for (int r = 0; r < RPW; ++r) {
                        const int ky = iy - r;
                        if (ky < 0 || ky > 2) continue;
#pragma unroll
                        for (int nb = 0; nb < NB; ++nb) {
                            if constexpr (KC == 32)
                                acc[r][nb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(av, bw[ky * 3 + kx][kb][nb], acc[r][nb], 0, 0, 0);
                            else
                                acc[r][nb] = __builtin_amdgcn_mfma_f32_16x16x16f16(av, bw[ky * 3 + kx][kb][nb], acc[r][nb], 0, 0, 0);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const int oy = y0 + ly0 + r;
            if (oy >= h) continue;  // uniform per wave
            if (EPI == 1) {
                // out = residual + (acc + bias): staged through LDS so that lane
                // 16 c + p stores pixel p of plane c (coalesced per plane)
                float* of = ofp + wave * 48;
                if (lr < 3) {
#pragma unroll
                    for (int e = 0; e < 4; ++e) of[lr * 16 + lq * 4 + e] = acc[r][0][e] + bias_l[0];
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const int c = lane >> 4, ox = x0 + seg * 16 + (lane & 15);
                if (lane < 48 && ox < w) {
                    const size_t q = img * a.img_stride_planes + (size_t)oy * a.pitch + ox;
                    a.outp[c][q] = a.res[c][q] + of[lane];
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                continue;
            }
            // D[i = pixel][j = channel] (lane lr + 16 lq, element e: pixel 4 lq + e):
            // through the wave's LDS buffer, then one 16-byte store per lane
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const int co = nb * 16 + lr;
                const float b = bias_l[nb];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    float v = acc[r][nb][e] + b;
                    if (EPI == 0) v = v > 0.0f ? v : 0.0f;
                    o16[(lq * 4 + e) * CO + co] = (_Float16)v;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            constexpr int CH = CO / 8;  // 16-byte chunks per pixel
#pragma unroll
            for (int k = lane; k < 16 * CH; k += 64) {
                const int px = k / CH, ch = k % CH;
                const int ox = x0 + seg * 16 + px;
                if (ox < w) {
                    const size_t p = (size_t)img * w * h + (size_t)oy * w + ox;
                    half8 v = *reinterpret_cast<const half8*>(o16 + px * CO + ch * 8);
                    if (EPI == 2) {
                        const half8 m = *reinterpret_cast<const half8*>(a.mask + p * CO + ch * 8);
#pragma unroll
                        for (int e = 0; e < 8; ++e) v[e] = (float)m[e] > 0.0f ? v[e] : (_Float16)0.0f;
                    }
                    *reinterpret_cast<half8*>(a.out + p * CO + ch * 8) = v;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
}

// ---------------------------------------------------------------------------
// First layer for inference, fused with the input packing: reads the three
// FP32 opponent planes (+ the noise level as channel 3) straight into an LDS
// tile of 4 FP16 channels per pixel, and packs four taps x four channels into
// the K = 16 of one v_mfma_f32_16x16x16f16: K index k = 4 t + c for tap
// t = 3 ky + kx (t >= 9: zero). Three MFMAs per 16 pixels x 16 output
// channels instead of nine with twelve zero channels. Output as EPI 0
// (ReLU(acc + bias), FP16 NHWC, 32 channels). Same tile shape as k_conv3x3.
// ---------------------------------------------------------------------------

constexpr int FIRST_KB = 3;  // ceil(9 taps * 4 channels / 16)

// B fragments of the packed first layer: frag[(kb * NB + nb) * 64 + lane][e]
// = W[t][c][j] with k = 16 kb + 4 (lane / 16) + e = 4 t + c, j = 16 nb + lane % 16.
__global__ void k_frag_first(const _Float16* __restrict__ w16, _Float16* __restrict__ dst) {
    constexpr int CO = COUT[0], CI = CIN[0], NB = CO / 16;
    const int idx = threadIdx.x;  // (kb * NB + nb) * 64 + lane
    if (idx >= FIRST_KB * NB * 64) return;
    const int lane = idx & 63, f = idx >> 6, nb = f % NB, kb = f / NB;
    const int j = nb * 16 + (lane & 15);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int k = kb * 16 + (lane >> 4) * 4 + e, t = k >> 2, c = k & 3;
        dst[(size_t)idx * 4 + e] = t < 9 ? w16[((size_t)t * CI + c) * CO + j] : (_Float16)0.0f;
    }
}

struct FirstArgs {
    const float* in[3];       // opponent planes
    size_t pitch, img_stride_planes;
    float level;              // 10 sigma (when levels is null)
    const float* levels;      // per image
    const _Float16* frag;     // k_frag_first layout
    const float* bias;        // [32]
    _Float16* out;            // NHWC, 32 channels
    const int* gate;          // skip once *gate is set (null: never)
    int w, h;
};

__global__ __launch_bounds__(256) void k_conv_first(FirstArgs a) {
    constexpr int CO = COUT[0], NB = CO / 16, RPW = CTH / 4;
    constexpr int TP = (CTH + 2) * (TW + 2);
    __shared__ __align__(16) half4 tile[TP];
    __shared__ __align__(16) _Float16 ob[4][16 * CO];
    if (a.gate && gate_set(a.gate)) return;
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * CTH, img = blockIdx.z;
    const int w = a.w, h = a.h;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    half4 bw[FIRST_KB][NB];
#pragma unroll
    for (int kb = 0; kb < FIRST_KB; ++kb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
            bw[kb][nb] = *reinterpret_cast<const half4*>(a.frag + (size_t)((kb * NB + nb) * 64 + lane) * 4);
    const float lv = a.levels ? a.levels[img] : a.level;
    constexpr int IT = (TP + 255) / 256;
    float sv[IT][3];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int pix = threadIdx.x + 256 * i;
        const int ty = pix / (TW + 2), tx = pix - ty * (TW + 2);
        const int gx = x0 + tx - 1, gy = y0 + ty - 1;
        sv[i][0] = sv[i][1] = sv[i][2] = 0.0f;
        if (pix < TP && gx >= 0 && gx < w && gy >= 0 && gy < h) {
            const size_t q = img * a.img_stride_planes + (size_t)gy * a.pitch + gx;
            sv[i][0] = a.in[0][q];
            sv[i][1] = a.in[1][q];
            sv[i][2] = a.in[2][q];
        }
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int pix = threadIdx.x + 256 * i;
        const int ty = pix / (TW + 2), tx = pix - ty * (TW + 2);
        const int gx = x0 + tx - 1, gy = y0 + ty - 1;
        const bool in = gx >= 0 && gx < w && gy >= 0 && gy < h;
        if (pix < TP)
            tile[pix] = (half4){(_Float16)sv[i][0], (_Float16)sv[i][1], (_Float16)sv[i][2],
                                in ? (_Float16)lv : (_Float16)0.0f};
    }
    __syncthreads();
    _Float16* o16 = ob[wave];
    typedef _Float16 half8 __attribute__((ext_vector_type(8)));
    float bias_l[NB];  // read once (the output stores could alias them)
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) bias_l[nb] = a.bias[nb * 16 + lr];
#pragma unroll 1
    for (int r = 0; r < RPW; ++r) {
        const int ly = wave * RPW + r, oy = y0 + ly;
#pragma unroll 1
        for (int seg = 0; seg < TW / 16; ++seg) {
            float4v acc[NB];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) acc[nb] = (float4v){0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < FIRST_KB; ++kb) {
                const int t = kb * 4 + lq, ky = t / 3, kx = t - 3 * (t / 3);
                half4 av = {0, 0, 0, 0};
                if (t < 9) av = tile[(ly + ky) * (TW + 2) + seg * 16 + lr + kx];
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
                    acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16f16(av, bw[kb][nb], acc[nb], 0, 0, 0);
            }
            if (oy >= h) continue;  // uniform per wave
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const int co = nb * 16 + lr;
                const float b = bias_l[nb];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float v = acc[nb][e] + b;
                    o16[(lq * 4 + e) * CO + co] = (_Float16)(v > 0.0f ? v : 0.0f);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            constexpr int CH = CO / 8;
#pragma unroll
            for (int k = lane; k < 16 * CH; k += 64) {
                const int px = k / CH, ch = k % CH;
                const int ox = x0 + seg * 16 + px;
                if (ox < w) {
                    const size_t p = (size_t)img * w * h + (size_t)oy * w + ox;
                    *reinterpret_cast<half8*>(a.out + p * CO + ch * 8) = *reinterpret_cast<const half8*>(o16 + px * CO + ch * 8);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
}

// ---------------------------------------------------------------------------
// Weight gradients: dW[t][ci][co] = sum_p X[p + t][ci] * dY[p][co], db[co] =
// sum_p dY[p][co], as MFMA with K = pixels. Each workgroup reduces its own
// pixel tile (TW x TH) and writes FP32 partials; k_reduce sums them in a
// fixed order.
// ---------------------------------------------------------------------------

struct WgradArgs {
    const _Float16* x;   // NHWC, stride CI (layer input)
    const _Float16* dy;  // NHWC, stride CO (gradient at the layer output, pre-activation)
    float* partial;      // [wg][9*CI*CO + CO]
    int w, h;
};

template <int CI, int CO>
__global__ __launch_bounds__(256) void k_conv3x3_wgrad(WgradArgs a) {
    // channel-major staging so that 4 consecutive pixels are contiguous
    constexpr int XW = TW + 2, XH = TH + 2;
    __shared__ _Float16 xs[CI * XH * XW];
    __shared__ _Float16 gs[CO * TH * TW];
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH, img = blockIdx.z;
    const int w = a.w, h = a.h;
    const _Float16* xin = a.x + (size_t)img * w * h * CI;
    const _Float16* gin = a.dy + (size_t)img * w * h * CO;
    for (int idx = threadIdx.x; idx < XH * XW * CI; idx += blockDim.x) {
        const int c = idx % CI, pix = idx / CI;
        const int ty = pix / XW, tx = pix - ty * XW;
        const int gx = x0 + tx - 1, gy = y0 + ty - 1;
        _Float16 v = 0;
        if (gx >= 0 && gx < w && gy >= 0 && gy < h) v = xin[((size_t)gy * w + gx) * CI + c];
        xs[c * XH * XW + pix] = v;
    }
    for (int idx = threadIdx.x; idx < TH * TW * CO; idx += blockDim.x) {
        const int c = idx % CO, pix = idx / CO;
        const int ty = pix / TW, tx = pix - ty * TW;
        const int gx = x0 + tx, gy = y0 + ty;
        _Float16 v = 0;
        if (gx < w && gy < h) v = gin[((size_t)gy * w + gx) * CO + c];
        gs[c * TH * TW + pix] = v;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    constexpr int MB = CI / 16, NB = CO / 16;
    const size_t wg = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    float* part = a.partial + wg * (9 * CI * CO + CO);
    // 9 taps x MB x NB output blocks, spread over the 4 waves
    for (int job = wave; job < 9 * MB * NB; job += 4) {
        const int t = job / (MB * NB), mb = (job / NB) % MB, nb = job % NB;
        const int ky = t / 3, kx = t % 3;
        float4v acc = {0.f, 0.f, 0.f, 0.f};
        // K = the TH*TW pixels of the tile, 16 at a time (4 per lane)
#pragma unroll 4
        for (int k0 = 0; k0 < TH * TW; k0 += 16) {
            const int py = k0 / TW, px = k0 % TW + lq * 4;
            half4 av, bv;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                av[e] = xs[(mb * 16 + lr) * XH * XW + (py + ky) * XW + px + e + kx];
                bv[e] = gs[(nb * 16 + lr) * TH * TW + py * TW + px + e];
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16f16(av, bv, acc, 0, 0, 0);
        }
        // D[i = ci][j = co]: lane holds rows lq*4 + e, column lr
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int ci = mb * 16 + lq * 4 + e, co = nb * 16 + lr;
            part[((size_t)t * CI + ci) * CO + co] = acc[e];
        }
    }
    if (threadIdx.x < CO) {
        float s = 0.0f;
        for (int pix = 0; pix < TH * TW; ++pix) s += (float)gs[threadIdx.x * TH * TW + pix];
        part[9 * CI * CO + threadIdx.x] = s;
    }
}

// Sum partials in workgroup order (deterministic), scale, write FP32 grads.
__global__ void k_reduce(const float* __restrict__ partial, int nwg, int n, float scale,
                         float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int k = 0; k < nwg; ++k) s += partial[(size_t)k * n + i];
    out[i] = s * scale;
}

// Loss gradient at the output planes (layout [img][c][plane] for both the
// output and the target): dY (FP16 NHWC, 16 ch) = out - target for channels
// < 3, and per-block squared-error sums for the loss value.
__global__ void k_loss_grad(const float* __restrict__ out, const float* __restrict__ tgt,
                            size_t plane, const float* __restrict__ wimg, _Float16* __restrict__ dy,
                            float* __restrict__ sq_partial) {
    __shared__ float red[256];
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t img = blockIdx.y;
    const float wi = wimg ? wimg[img] : 1.0f;
    float s = 0.0f;
    if (p < plane) {
        typedef _Float16 half8 __attribute__((ext_vector_type(8)));
        half8 g = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int c = 0; c < 3; ++c) {
            const size_t q = (img * 3 + c) * plane + p;
            const float d = out[q] - tgt[q];
            g[c] = (_Float16)(wi * d);
            s += wi * (d * d);
        }
        _Float16* o = dy + (img * plane + p) * 16;
        *reinterpret_cast<half8*>(o) = g;
        *reinterpret_cast<half8*>(o + 8) = (half8){0, 0, 0, 0, 0, 0, 0, 0};
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) sq_partial[img * gridDim.x + blockIdx.x] = red[0];
}

// Adam on the FP32 master weights; refreshes the FP16 copy; padded entries
// (mask == 0) stay zero.
__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, const uint8_t* __restrict__ live,
                       _Float16* __restrict__ w16, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!live[i]) {
        w[i] = 0.0f;
        w16[i] = 0;
        return;
    }
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    w16[i] = (_Float16)wi;
}

__global__ void k_to_half(const float* __restrict__ w, _Float16* __restrict__ w16, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w16[i] = (_Float16)w[i];
}

// ===========================================================================
// Host side
// ===========================================================================

namespace {

thread_local std::string g_perr;

int perr(int code, const std::string& msg) {
    g_perr = msg;
    return code;
}

#define PHIP(expr)                                                                      \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return perr(e_ == hipErrorOutOfMemory ? 3 : 7,                              \
                        std::string(#expr) + ": " + hipGetErrorString(e_));             \
    } while (0)

constexpr uint32_t kBlobMagic = 0x4C504350u;  // "CPPL"
constexpr uint32_t kBlobVersion = 1;

struct BlobHeader {
    uint32_t magic;
    uint32_t version;
    uint32_t nparam;
    uint32_t flags;  // bit 0: optimizer state follows the weights
    uint64_t step;
    uint64_t seed;
};

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double uniform01(uint64_t& s) { return ((splitmix64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

}  // namespace
}  // namespace pilot

using namespace pilot;

struct cpipe_bm3d_pilot_s {
    int device = 0;
    size_t nparam = 0;
    size_t off[NL] = {};
    float* w32 = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    float* g = nullptr;
    _Float16* w16 = nullptr;
    _Float16* frag = nullptr;  // per layer: forward then backward B fragments (k_frag)
    size_t frag_off[NL][2] = {};
    size_t frag_first = 0;  // first layer, taps packed into K (k_frag_first)
    uint8_t* live = nullptr;
    uint64_t step = 0;
    uint64_t seed = 0;
    std::vector<uint8_t> live_host;
    // workspace (grown on demand)
    size_t cap_px = 0;
    bool cap_train = false;
    _Float16* act[NL] = {};
    _Float16* dyb[2] = {};
    float* outp = nullptr;  // 3 planes (training)
    float* partial = nullptr;
    size_t partial_cap = 0;
    float* sq = nullptr;
    size_t sq_cap = 0;
    float* levels = nullptr;
    size_t levels_cap = 0;
    // loss: 0 = 1/2 MSE; 1 = 1/2 MSE with image i weighted by (ref_sigma / sigma_i)^2
    int loss_mode = 0;
    float loss_ref_sigma = 0.05f;
    float* wimg = nullptr;
    size_t wimg_cap = 0;
};

namespace {

class PDevice {
public:
    explicit PDevice(int dev) {
        if (hipGetDevice(&prev_) != hipSuccess) prev_ = -1;
        if (prev_ != dev) (void)hipSetDevice(dev);
    }
    ~PDevice() {
        if (prev_ >= 0) (void)hipSetDevice(prev_);
    }

private:
    int prev_ = -1;
};

void free_ws(cpipe_bm3d_pilot* n) {
    for (auto& a : n->act) {
        (void)hipFree(a);
        a = nullptr;
    }
    for (auto& d : n->dyb) {
        (void)hipFree(d);
        d = nullptr;
    }
    (void)hipFree(n->outp);
    n->outp = nullptr;
    n->cap_px = 0;
}

// Workspace for `px` pixels: inference needs the input and two ping-pong
// activations; training keeps every activation plus two gradient buffers.
int ensure_ws(cpipe_bm3d_pilot* n, size_t px, bool train) {
    if (px <= n->cap_px && (!train || n->cap_train)) return 0;
    (void)hipDeviceSynchronize();
    free_ws(n);
    const size_t want = px;
    if (train) PHIP(hipMalloc((void**)&n->act[0], want * 16 * sizeof(_Float16)));  // packed input
    const int nact = train ? NL : 3;
    for (int l = 1; l < nact; ++l) PHIP(hipMalloc((void**)&n->act[l], want * 32 * sizeof(_Float16)));
    if (train) {
        for (auto& d : n->dyb) PHIP(hipMalloc((void**)&d, want * 32 * sizeof(_Float16)));
        PHIP(hipMalloc((void**)&n->outp, want * 3 * sizeof(float)));
    }
    n->cap_px = want;
    n->cap_train = train;
    return 0;
}

int ensure_buf(float** p, size_t* cap, size_t n) {
    if (n <= *cap) return 0;
    (void)hipDeviceSynchronize();
    (void)hipFree(*p);
    *p = nullptr;
    PHIP(hipMalloc((void**)p, n * sizeof(float)));
    *cap = n;
    return 0;
}

// Rebuilds every layer's B fragments from the FP16 weights (after each change).
int refresh_frags(cpipe_bm3d_pilot* n, hipStream_t st) {
    // one thread per fragment lane: at most 9 * CI * CO / 4 per operator
    const int blocks[NL] = {(int)((layer_w(0) / 4 + 255) / 256), (int)((layer_w(1) / 4 + 255) / 256),
                            (int)((layer_w(2) / 4 + 255) / 256), (int)((layer_w(3) / 4 + 255) / 256),
                            (int)((layer_w(4) / 4 + 255) / 256), (int)((layer_w(5) / 4 + 255) / 256)};
    auto f = [&](int l) { return n->frag + n->frag_off[l][0]; };
    auto b = [&](int l) { return n->frag + n->frag_off[l][1]; };
    k_frag<16, 32><<<blocks[0], 256, 0, st>>>(n->w16 + n->off[0], f(0), b(0));
    for (int l = 1; l < NL - 1; ++l) k_frag<32, 32><<<blocks[l], 256, 0, st>>>(n->w16 + n->off[l], f(l), b(l));
    k_frag<32, 16><<<blocks[NL - 1], 256, 0, st>>>(n->w16 + n->off[NL - 1], f(NL - 1), b(NL - 1));
    k_frag_first<<<1, FIRST_KB * (COUT[0] / 16) * 64, 0, st>>>(n->w16 + n->off[0], n->frag + n->frag_first);
    PHIP(hipGetLastError());
    return 0;
}

template <int CI, int CO, int EPI>
int launch_conv(const ConvArgs& a, int n_img, hipStream_t st) {
    const dim3 grid((a.w + TW - 1) / TW, (a.h + CTH - 1) / CTH, n_img);
    k_conv3x3<CI, CO, EPI><<<grid, 256, 0, st>>>(a);
    PHIP(hipGetLastError());
    return 0;
}

int forward(cpipe_bm3d_pilot* n, const float* const* in, size_t pitch, size_t img_stride, int w,
            int h, int n_img, float level, const float* levels, float* const* out, bool train,
            hipStream_t st, void* ws = nullptr, const int* gate = nullptr) {
    const size_t px = (size_t)w * h * n_img;
    _Float16* act[NL] = {};
    if (ws && !train) {  // inference on the caller's workspace: layers ping-pong between two buffers
        act[1] = static_cast<_Float16*>(ws);
        act[2] = act[1] + px * 32;
    } else {
        int rc = ensure_ws(n, px, train);
        if (rc) return rc;
        for (int l = 0; l < NL; ++l) act[l] = n->act[l];
    }
    int rc = 0;
    ConvArgs a{};
    a.w = w;
    a.h = h;
    a.pitch = pitch;
    a.img_stride_planes = img_stride;
    a.gate = train ? nullptr : gate;
    if (train) {
        // L1 on the packed input, which the weight gradient reads later
        k_pack_input<<<dim3((w + 255) / 256, h, n_img), 256, 0, st>>>(in[0], in[1], in[2], pitch, level,
                                                                       levels, w, h, img_stride, act[0]);
        PHIP(hipGetLastError());
        a.in = act[0];
        a.frag = n->frag + n->frag_off[0][0];
        a.bias = n->w32 + n->off[0] + layer_w(0);
        a.out = act[1];
        if ((rc = launch_conv<16, 32, 0>(a, n_img, st))) return rc;
    } else {
        // L1 fused with the packing, taps packed into K
        FirstArgs f{};
        for (int c = 0; c < 3; ++c) f.in[c] = in[c];
        f.pitch = pitch;
        f.img_stride_planes = img_stride;
        f.level = level;
        f.levels = levels;
        f.frag = n->frag + n->frag_first;
        f.bias = n->w32 + n->off[0] + layer_w(0);
        f.out = act[1];
        f.gate = gate;
        f.w = w;
        f.h = h;
        k_conv_first<<<dim3((w + TW - 1) / TW, (h + CTH - 1) / CTH, n_img), 256, 0, st>>>(f);
        PHIP(hipGetLastError());
    }
    int cur = 1;
    for (int l = 1; l < NL - 1; ++l) {
        const int nxt = train ? l + 1 : (cur == 1 ? 2 : 1);
        a.in = act[cur];
        a.frag = n->frag + n->frag_off[l][0];
        a.bias = n->w32 + n->off[l] + layer_w(l);
        a.out = act[nxt];
        if ((rc = launch_conv<32, 32, 0>(a, n_img, st))) return rc;
        cur = nxt;
    }
    a.in = act[cur];
    a.frag = n->frag + n->frag_off[NL - 1][0];
    a.bias = n->w32 + n->off[NL - 1] + layer_w(NL - 1);
    for (int c = 0; c < 3; ++c) {
        a.res[c] = in[c];
        a.outp[c] = out[c];
    }
    return launch_conv<32, 16, 1>(a, n_img, st);
}

template <int CI, int CO>
int wgrad(cpipe_bm3d_pilot* n, int l, const _Float16* x, const _Float16* dy, int w, int h,
          int n_img, float scale, hipStream_t st) {
    const dim3 grid((w + TW - 1) / TW, (h + TH - 1) / TH, n_img);
    const size_t nwg = (size_t)grid.x * grid.y * grid.z;
    const size_t per = (size_t)9 * CI * CO + CO;
    int rc = ensure_buf(&n->partial, &n->partial_cap, nwg * per);
    if (rc) return rc;
    WgradArgs a{x, dy, n->partial, w, h};
    k_conv3x3_wgrad<CI, CO><<<grid, 256, 0, st>>>(a);
    PHIP(hipGetLastError());
    k_reduce<<<(unsigned)((per + 255) / 256), 256, 0, st>>>(n->partial, (int)nwg, (int)per, scale,
                                                             n->g + n->off[l]);
    PHIP(hipGetLastError());
    return 0;
}

// forward + loss + backward into n->g; returns the loss (host).
// Inputs and targets are laid out [img][c][plane] (in[c] = base + c * plane).
int backward(cpipe_bm3d_pilot* n, const float* const* in, const float* const* tgt, int w, int h,
             int n_img, const float* levels_dev, float* loss, hipStream_t st) {
    const size_t plane = (size_t)w * h;
    const size_t px = plane * n_img;
    int rc = ensure_ws(n, px, true);
    if (rc) return rc;
    float* outp[3] = {n->outp, n->outp + plane, n->outp + 2 * plane};
    if ((rc = forward(n, in, (size_t)w, 3 * plane, w, h, n_img, 0.0f, levels_dev, outp, true, st)))
        return rc;
    const size_t nb_img = (plane + 255) / 256;
    const size_t nblk = nb_img * n_img;
    if ((rc = ensure_buf(&n->sq, &n->sq_cap, nblk))) return rc;
    k_loss_grad<<<dim3((unsigned)nb_img, n_img), 256, 0, st>>>(n->outp, tgt[0], plane,
                                                                n->loss_mode ? n->wimg : nullptr, n->dyb[0], n->sq);
    PHIP(hipGetLastError());
    std::vector<float> sqh(nblk);
    const float scale = 1.0f / (float)(px * 3);
    // layer 6 (index 5): x = act[5], dy = dyb[0] (16 ch)
    if ((rc = wgrad<32, 16>(n, NL - 1, n->act[NL - 1], n->dyb[0], w, h, n_img, scale, st))) return rc;
    ConvArgs a{};
    a.w = w;
    a.h = h;
    a.in = n->dyb[0];
    a.frag = n->frag + n->frag_off[NL - 1][1];
    a.mask = n->act[NL - 1];
    a.out = n->dyb[1];
    if ((rc = launch_conv<16, 32, 2>(a, n_img, st))) return rc;
    int cur = 1;
    for (int l = NL - 2; l >= 1; --l) {
        if ((rc = wgrad<32, 32>(n, l, n->act[l], n->dyb[cur], w, h, n_img, scale, st))) return rc;
        a.in = n->dyb[cur];
        a.frag = n->frag + n->frag_off[l][1];
        a.mask = n->act[l];
        a.out = n->dyb[1 - cur];
        if ((rc = launch_conv<32, 32, 2>(a, n_img, st))) return rc;
        cur = 1 - cur;
    }
    if ((rc = wgrad<16, 32>(n, 0, n->act[0], n->dyb[cur], w, h, n_img, scale, st))) return rc;
    PHIP(hipMemcpyAsync(sqh.data(), n->sq, nblk * sizeof(float), hipMemcpyDeviceToHost, st));
    PHIP(hipStreamSynchronize(st));
    double acc = 0.0;
    for (size_t i = 0; i < nblk; ++i) acc += sqh[i];
    if (loss) *loss = (float)(0.5 * acc / (double)(px * 3));
    return 0;
}

int upload_levels(cpipe_bm3d_pilot* n, const float* sigma, int n_img, hipStream_t st) {
    int rc = ensure_buf(&n->levels, &n->levels_cap, (size_t)n_img);
    if (rc) return rc;
    if ((rc = ensure_buf(&n->wimg, &n->wimg_cap, (size_t)n_img))) return rc;
    std::vector<float> lv(2 * (size_t)n_img);
    for (int i = 0; i < n_img; ++i) {
        lv[i] = 10.0f * sigma[i];
        const float r = sigma[i] > 0.0f ? n->loss_ref_sigma / sigma[i] : 1.0f;
        lv[n_img + i] = r * r;
    }
    PHIP(hipMemcpyAsync(n->levels, lv.data(), n_img * sizeof(float), hipMemcpyHostToDevice, st));
    PHIP(hipMemcpyAsync(n->wimg, lv.data() + n_img, n_img * sizeof(float), hipMemcpyHostToDevice, st));
    PHIP(hipStreamSynchronize(st));
    return 0;
}

}  // namespace

namespace cpipe_internal {
size_t pilot_workspace_bytes(size_t px) { return 2 * px * 32 * sizeof(_Float16); }

int pilot_forward(cpipe_bm3d_pilot* net, const float* const* in, size_t pitch, size_t img_stride,
                  int w, int h, int n_img, float level, const float* levels, float* const* out,
                  void* ws, hipStream_t st, const int* gate) {
    return forward(net, in, pitch, img_stride, w, h, n_img, level, levels, out, false, st, ws, gate);
}
}  // namespace cpipe_internal

extern "C" {

const char* cpipe_bm3d_pilot_last_error(void) { return g_perr.c_str(); }

size_t cpipe_bm3d_pilot_param_count(void) { return total_params(); }

int cpipe_bm3d_pilot_create(int device, cpipe_bm3d_pilot** out) {
    if (!out) return perr(5, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return perr(8, "no HIP device visible");
    if (device < 0 || device >= ndev) return perr(5, "device ordinal out of range");
    PDevice g(device);
    auto* n = new cpipe_bm3d_pilot_s();
    n->device = device;
    size_t o = 0;
    for (int l = 0; l < NL; ++l) {
        n->off[l] = o;
        o += layer_params(l);
    }
    n->nparam = o;
    n->live_host.assign(o, 1);
    for (int l = 0; l < NL; ++l) {
        for (int t = 0; t < 9; ++t)
            for (int ci = 0; ci < CIN[l]; ++ci)
                for (int co = 0; co < COUT[l]; ++co)
                    if (ci >= CIN_REAL[l] || co >= COUT_REAL[l])
                        n->live_host[n->off[l] + ((size_t)t * CIN[l] + ci) * COUT[l] + co] = 0;
        for (int co = COUT_REAL[l]; co < COUT[l]; ++co) n->live_host[n->off[l] + layer_w(l) + co] = 0;
    }
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = hipMalloc((void**)&n->w32, o * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&n->m, o * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&n->v, o * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&n->g, o * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&n->w16, o * sizeof(_Float16));
    if (e == hipSuccess) e = hipMalloc((void**)&n->live, o);
    size_t fo = 0;  // fragments: 9 * CI * CO halves per operator
    for (int l = 0; l < NL; ++l) {
        n->frag_off[l][0] = fo;
        fo += layer_w(l);
        n->frag_off[l][1] = fo;
        fo += layer_w(l);
    }
    n->frag_first = fo;
    fo += (size_t)FIRST_KB * (COUT[0] / 16) * 64 * 4;
    if (e == hipSuccess) e = hipMalloc((void**)&n->frag, fo * sizeof(_Float16));
    if (e == hipSuccess) e = hipMemset(n->frag, 0, fo * sizeof(_Float16));
    if (e == hipSuccess) e = hipMemcpy(n->live, n->live_host.data(), o, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(n->w32, 0, o * sizeof(float));
    if (e == hipSuccess) e = hipMemset(n->m, 0, o * sizeof(float));
    if (e == hipSuccess) e = hipMemset(n->v, 0, o * sizeof(float));
    if (e == hipSuccess) e = hipMemset(n->w16, 0, o * sizeof(_Float16));
    if (e != hipSuccess) {
        cpipe_bm3d_pilot_destroy(n);
        return perr(e == hipErrorOutOfMemory ? 3 : 7, std::string("pilot alloc: ") + hipGetErrorString(e));
    }
    *out = n;
    return 0;
}

void cpipe_bm3d_pilot_destroy(cpipe_bm3d_pilot* n) {
    if (!n) return;
    PDevice g(n->device);
    (void)hipDeviceSynchronize();
    free_ws(n);
    for (void* p : {(void*)n->w32, (void*)n->m, (void*)n->v, (void*)n->g, (void*)n->w16,
                    (void*)n->frag, (void*)n->live, (void*)n->partial, (void*)n->sq, (void*)n->levels,
                    (void*)n->wimg})
        (void)hipFree(p);
    delete n;
}

int cpipe_bm3d_pilot_init(cpipe_bm3d_pilot* n, uint64_t seed) {
    if (!n) return perr(5, "null pilot");
    PDevice g(n->device);
    std::vector<float> w(n->nparam, 0.0f);
    uint64_t s = seed;
    for (int l = 0; l < NL; ++l) {
        const double fan_in = 9.0 * CIN_REAL[l];
        double std = std::sqrt(2.0 / fan_in);  // He init for ReLU layers
        if (l == NL - 1) std *= 0.1;           // residual output starts near zero
        for (size_t i = 0; i < layer_w(l); ++i) {
            const double u1 = uniform01(s), u2 = uniform01(s);
            const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
            if (n->live_host[n->off[l] + i]) w[n->off[l] + i] = (float)(z * std);
        }
    }
    n->step = 0;
    n->seed = seed;
    PHIP(hipMemcpy(n->w32, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
    PHIP(hipMemset(n->m, 0, n->nparam * sizeof(float)));
    PHIP(hipMemset(n->v, 0, n->nparam * sizeof(float)));
    k_to_half<<<(unsigned)((n->nparam + 255) / 256), 256>>>(n->w32, n->w16, (int)n->nparam);
    PHIP(hipGetLastError());
    if (int rc = refresh_frags(n, nullptr)) return rc;
    PHIP(hipDeviceSynchronize());
    return 0;
}

size_t cpipe_bm3d_pilot_blob_size(int with_optimizer) {
    return sizeof(BlobHeader) + total_params() * sizeof(float) * (with_optimizer ? 3 : 1);
}

int cpipe_bm3d_pilot_save(const cpipe_bm3d_pilot* n, void* blob, size_t cap, int with_optimizer) {
    if (!n || !blob) return perr(5, "null argument");
    const size_t need = cpipe_bm3d_pilot_blob_size(with_optimizer);
    if (cap < need) return perr(5, "blob buffer too small");
    PDevice g(n->device);
    BlobHeader h{kBlobMagic, kBlobVersion, (uint32_t)n->nparam, with_optimizer ? 1u : 0u, n->step, n->seed};
    std::memcpy(blob, &h, sizeof(h));
    auto* f = reinterpret_cast<float*>(static_cast<char*>(blob) + sizeof(h));
    PHIP(hipDeviceSynchronize());
    PHIP(hipMemcpy(f, n->w32, n->nparam * sizeof(float), hipMemcpyDeviceToHost));
    if (with_optimizer) {
        PHIP(hipMemcpy(f + n->nparam, n->m, n->nparam * sizeof(float), hipMemcpyDeviceToHost));
        PHIP(hipMemcpy(f + 2 * n->nparam, n->v, n->nparam * sizeof(float), hipMemcpyDeviceToHost));
    }
    return 0;
}

int cpipe_bm3d_pilot_load(cpipe_bm3d_pilot* n, const void* blob, size_t size) {
    if (!n || !blob) return perr(5, "null argument");
    if (size < sizeof(BlobHeader)) return perr(5, "blob too small");
    BlobHeader h;
    std::memcpy(&h, blob, sizeof(h));
    if (h.magic != kBlobMagic || h.version != kBlobVersion) return perr(5, "not a pilot weight blob (magic/version)");
    if (h.nparam != n->nparam) return perr(5, "blob was written for another network shape");
    const bool opt = (h.flags & 1u) != 0;
    if (size < cpipe_bm3d_pilot_blob_size(opt ? 1 : 0)) return perr(5, "blob truncated");
    PDevice g(n->device);
    const float* f = reinterpret_cast<const float*>(static_cast<const char*>(blob) + sizeof(h));
    PHIP(hipDeviceSynchronize());
    PHIP(hipMemcpy(n->w32, f, n->nparam * sizeof(float), hipMemcpyHostToDevice));
    if (opt) {
        PHIP(hipMemcpy(n->m, f + n->nparam, n->nparam * sizeof(float), hipMemcpyHostToDevice));
        PHIP(hipMemcpy(n->v, f + 2 * n->nparam, n->nparam * sizeof(float), hipMemcpyHostToDevice));
    } else {
        PHIP(hipMemset(n->m, 0, n->nparam * sizeof(float)));
        PHIP(hipMemset(n->v, 0, n->nparam * sizeof(float)));
    }
    n->step = opt ? h.step : 0;
    n->seed = h.seed;
    k_to_half<<<(unsigned)((n->nparam + 255) / 256), 256>>>(n->w32, n->w16, (int)n->nparam);
    PHIP(hipGetLastError());
    if (int rc = refresh_frags(n, nullptr)) return rc;
    PHIP(hipDeviceSynchronize());
    return 0;
}

int cpipe_bm3d_pilot_forward(cpipe_bm3d_pilot* n, const float* const* in, size_t pitch,
                             size_t img_stride, int w, int h, int n_img, float sigma,
                             float* const* out, void* stream) {
    if (!n || !in || !out) return perr(5, "null argument");
    if (w <= 0 || h <= 0 || n_img <= 0 || pitch < (size_t)w) return perr(5, "bad geometry");
    PDevice g(n->device);
    return forward(n, in, pitch, img_stride, w, h, n_img, 10.0f * sigma, nullptr, out, false,
                   (hipStream_t)stream);
}

int cpipe_bm3d_pilot_gradients(cpipe_bm3d_pilot* n, const float* const* in, const float* const* tgt,
                               int w, int h, int n_img, const float* sigma, float* grads_host,
                               float* loss) {
    if (!n || !in || !tgt || !sigma) return perr(5, "null argument");
    PDevice g(n->device);
    hipStream_t st = nullptr;
    int rc = upload_levels(n, sigma, n_img, st);
    if (rc) return rc;
    if ((rc = backward(n, in, tgt, w, h, n_img, n->levels, loss, st))) return rc;
    if (grads_host) {
        PHIP(hipStreamSynchronize(st));
        PHIP(hipMemcpy(grads_host, n->g, n->nparam * sizeof(float), hipMemcpyDeviceToHost));
    }
    return 0;
}

int cpipe_bm3d_pilot_set_loss(cpipe_bm3d_pilot* n, int mode, float ref_sigma) {
    if (!n) return perr(5, "null argument");
    if ((mode != 0 && mode != 1) || !(ref_sigma > 0.0f)) return perr(5, "bad loss mode or reference sigma");
    n->loss_mode = mode;
    n->loss_ref_sigma = ref_sigma;
    return 0;
}

int cpipe_bm3d_pilot_train_step(cpipe_bm3d_pilot* n, const float* const* in, const float* const* tgt,
                                int w, int h, int n_img, const float* sigma, float lr, float* loss) {
    if (!n || !in || !tgt || !sigma) return perr(5, "null argument");
    PDevice g(n->device);
    hipStream_t st = nullptr;
    int rc = upload_levels(n, sigma, n_img, st);
    if (rc) return rc;
    if ((rc = backward(n, in, tgt, w, h, n_img, n->levels, loss, st))) return rc;
    n->step++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - std::pow(b1, (float)n->step), c2 = 1.0f - std::pow(b2, (float)n->step);
    k_adam<<<(unsigned)((n->nparam + 255) / 256), 256, 0, st>>>(n->w32, n->m, n->v, n->g, n->live,
                                                                 n->w16, (int)n->nparam, lr, b1, b2,
                                                                 eps, c1, c2);
    PHIP(hipGetLastError());
    if (int rc2 = refresh_frags(n, st)) return rc2;
    return 0;
}

}  // extern "C"
