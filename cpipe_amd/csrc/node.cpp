// com.cpipe.denoise.bm3d -- the cpipe node (plugin ABI 0.4) driving the MI355X engine,
// and, further down, com.cpipe.denoise.wavelet_bayes_shrink, the next node of the
// reference's classic pipeline (examples/pipelines/full-classic-pipeline.cpipe.json:127-129),
// and com.cpipe.precision_convert, the conversion node its precision planner inserts.
//
// Replaces src/cpipe/nodes/denoise/bm3d.cpp (DenoiseBm3d) and its registration
// CPIPE_REGISTER_NODE (bm3d.cpp:111, include/cpipe/sdk/registry.hpp:14-32).
// Behaviour kept from the reference:
//   * action dispatch and the exception barrier of sdk::detail::dispatch
//     (include/cpipe/sdk/sdk.hpp:615-702): describe/prepare/unknown -> REPLY_DEFAULT,
//     create writes the instance to *(void**)out_ctx, destroy takes it from in_ctx,
//     process gets the instance as `node` and a cpipe_process_ctx as in_ctx;
//     anything thrown -> CPIPE_INTERNAL_ERROR + host->log(host, 4, ...);
//   * arity check -> CPIPE_BAD_INDEX (bm3d.cpp:84-87);
//   * format gate -> CPIPE_BAD_PRECISION (bm3d.cpp:57-68);
//   * sigma precedence sigma -> sigma_override -> metadata noise profile -> 0.03,
//     params clamped to [0, 0.2] (bm3d.cpp:21-55);
//   * add_applied_step("denoise.bm3d") on success (bm3d.cpp:103).
// Added: `mode` (compat | bm3d) and `precision` params, FP32 plane I/O
// (1 x R32_SFLOAT grey or 3 x R32_SFLOAT RGB planes), a defined status
// when no HIP device is visible (CPIPE_UNSUPPORTED, logged), and, for hosts
// that offer the "hip_buffer" suite, a device-resident path: when every input
// and output is in HIP device memory on one device, the three nodes run on the
// device addresses on the host's stream, with no lock_cpu and no host copy.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "cpipe_bm3d.h"
#include "cpipe_convert.h"
#include "cpipe_node_abi.h"
#include "cpipe_wavelet.h"

extern "C" const char CPIPE_BM3D_MANIFEST_JSON[];
extern "C" const char CPIPE_WAVELET_MANIFEST_JSON[];
extern "C" const char CPIPE_CONVERT_MANIFEST_JSON[];

namespace {

const char kNodeId[] = "com.cpipe.denoise.bm3d";
const char kNodeVersion[] = "2.0.0";

struct Suites {
    const cpipe_buffer_suite_v1* buffer = nullptr;
    const cpipe_param_suite_v1* param = nullptr;
    const cpipe_metadata_suite_v1* metadata = nullptr;
    const cpipe_metadata_builder_suite_v1* builder = nullptr;
    const cpipe_hip_buffer_suite_v1* hip = nullptr;  // optional: device-resident buffers
};

Suites fetch_suites(cpipe_host_t* host) {
    Suites s;
    if (host == nullptr || host->get_suite == nullptr) return s;
    s.buffer = static_cast<const cpipe_buffer_suite_v1*>(host->get_suite(host, "buffer", 1));
    s.param = static_cast<const cpipe_param_suite_v1*>(host->get_suite(host, "param", 1));
    s.metadata = static_cast<const cpipe_metadata_suite_v1*>(host->get_suite(host, "metadata", 1));
    s.builder = static_cast<const cpipe_metadata_builder_suite_v1*>(
        host->get_suite(host, "metadata_builder", 1));
    s.hip = static_cast<const cpipe_hip_buffer_suite_v1*>(host->get_suite(host, "hip_buffer", 1));
    return s;
}

// Device addresses of every input and output when the host keeps them all in
// HIP device memory on one device; false when the host has no "hip_buffer"
// suite or any buffer is host memory (the node then goes through lock_cpu).
bool device_resident(const Suites& s, const cpipe_process_ctx* ctx, std::vector<void*>* in,
                     std::vector<void*>* out, int* device) {
    if (s.hip == nullptr || s.hip->get_device_ptr == nullptr) return false;
    int dev = -1;
    auto one = [&](const cpipe_buffer_t* b, void** p) {
        int d = -1;
        if (s.hip->get_device_ptr(b, p, &d) != CPIPE_OK || *p == nullptr || d < 0) return false;
        if (dev < 0) dev = d;
        return d == dev;
    };
    in->assign(ctx->n_in, nullptr);
    out->assign(ctx->n_out, nullptr);
    for (size_t i = 0; i < ctx->n_in; ++i)
        if (!one(ctx->inputs[i], &(*in)[i])) return false;
    for (size_t i = 0; i < ctx->n_out; ++i)
        if (!one(ctx->outputs[i], &(*out)[i])) return false;
    *device = dev;
    return true;
}

// The stream the host orders this node's device work on (NULL: the null stream).
int device_stream(const Suites& s, const cpipe_process_ctx* ctx, void** stream) {
    *stream = nullptr;
    if (s.hip == nullptr || s.hip->get_stream == nullptr) return CPIPE_OK;
    return s.hip->get_stream(ctx->compute, stream);
}

// Launches an entry point's kernels on `device`, then waits for `stream`.
template <class F>
int on_device(int device, void* stream, F&& launch) {
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(device) != hipSuccess) return CPIPE_BAD_INDEX;
    int rc = launch();
    if (rc == CPIPE_OK && hipStreamSynchronize(static_cast<hipStream_t>(stream)) != hipSuccess)
        rc = CPIPE_INTERNAL_ERROR;
    (void)hipSetDevice(prev);
    return rc;
}

void host_log(cpipe_host_t* host, int level, const std::string& msg) {
    if (host != nullptr && host->log != nullptr) host->log(host, level, msg.c_str());
}

struct Bm3dNode {
    int unused = 0;
};

// The input's noise-profile pairs (variance a + b x per channel), false when
// it carries none (the metadata path of bm3d.cpp:29-52).
bool noise_profile(const Suites& s, const cpipe_buffer_t* input, std::vector<float>* a, std::vector<float>* b) {
    if (s.buffer == nullptr || s.buffer->get_metadata == nullptr || s.metadata == nullptr ||
        s.metadata->get_calibration == nullptr)
        return false;
    const cpipe_metadata_t* meta = nullptr;
    if (s.buffer->get_metadata(input, &meta) != CPIPE_OK || meta == nullptr) return false;
    cpipe_calibration_view view{};
    if (s.metadata->get_calibration(meta, &view) != CPIPE_OK || view.get_noise_profile == nullptr) return false;
    size_t total = 0;
    if (view.get_noise_profile(meta, 0, &total, nullptr, nullptr) != CPIPE_OK || total == 0) return false;
    a->assign(total, 0.0F);
    b->assign(total, 0.0F);
    return view.get_noise_profile(meta, total, &total, a->data(), b->data()) == CPIPE_OK && total > 0;
}

// Whether the params name sigma or sigma_override (which take precedence over metadata).
bool explicit_sigma(const Suites& s, const cpipe_props_t* params) {
    double v = 0.0;
    return s.param != nullptr && s.param->get_double != nullptr && params != nullptr &&
           (s.param->get_double(params, "sigma", &v) == CPIPE_OK ||
            s.param->get_double(params, "sigma_override", &v) == CPIPE_OK);
}

// bm3d.cpp:21-55
float resolve_sigma(const Suites& s, const cpipe_props_t* params, const cpipe_buffer_t* input) {
    if (s.param != nullptr && s.param->get_double != nullptr && params != nullptr) {
        double v = 0.0;
        if (s.param->get_double(params, "sigma", &v) == CPIPE_OK)
            return std::clamp(static_cast<float>(v), 0.0F, 0.2F);
        double o = 0.0;
        if (s.param->get_double(params, "sigma_override", &o) == CPIPE_OK)
            return std::clamp(static_cast<float>(o), 0.0F, 0.2F);
    }
    if (s.buffer == nullptr || s.buffer->get_metadata == nullptr || s.metadata == nullptr ||
        s.metadata->get_calibration == nullptr)
        return 0.03F;
    const cpipe_metadata_t* meta = nullptr;
    if (s.buffer->get_metadata(input, &meta) != CPIPE_OK || meta == nullptr) return 0.03F;
    cpipe_calibration_view view{};
    if (s.metadata->get_calibration(meta, &view) != CPIPE_OK || view.get_noise_profile == nullptr)
        return 0.03F;
    size_t total = 0;
    if (view.get_noise_profile(meta, 0, &total, nullptr, nullptr) != CPIPE_OK || total == 0)
        return 0.03F;
    std::vector<float> a(total), b(total);
    if (view.get_noise_profile(meta, total, &total, a.data(), b.data()) != CPIPE_OK || total == 0)
        return 0.03F;
    double variance = 0.0;
    for (size_t i = 0; i < total; ++i) variance += std::max(0.0F, a[i] + (b[i] * 0.25F));
    variance /= static_cast<double>(total);
    return static_cast<float>(std::sqrt(variance));
}

std::string enum_param(const Suites& s, const cpipe_props_t* params, const char* key,
                       const char* env, const char* dflt) {
    if (const char* e = std::getenv(env); e != nullptr && *e != '\0') return e;
    if (s.param != nullptr && s.param->get_enum != nullptr && params != nullptr) {
        const char* v = nullptr;
        if (s.param->get_enum(params, key, &v) == CPIPE_OK && v != nullptr) return v;
    }
    return dflt;
}

struct Image {
    int format = 0;
    uint32_t w = 0, h = 0;
    size_t row_bytes = 0;
};

int describe_buffer(const Suites& s, const cpipe_buffer_t* b, Image* out) {
    if (s.buffer == nullptr || s.buffer->get_format == nullptr || s.buffer->get_dims == nullptr)
        return CPIPE_BAD_INDEX;
    int rc = s.buffer->get_format(b, &out->format);
    if (rc != CPIPE_OK) return rc;
    uint8_t ndim = 0;
    uint32_t dims[8] = {};
    rc = s.buffer->get_dims(b, &ndim, dims);
    if (rc != CPIPE_OK) return rc;
    if (ndim < 2) return CPIPE_BAD_INDEX;
    out->w = dims[0];
    out->h = dims[1];
    const size_t bpp = out->format == CPIPE_FORMAT_R16G16B16A16_SFLOAT ? 8 : 4;
    out->row_bytes = (size_t)out->w * bpp;
    if (s.buffer->get_stride != nullptr) {
        uint64_t st[8] = {};
        if (s.buffer->get_stride(b, st) == CPIPE_OK) {
            if (st[0] != 0 && st[0] != bpp) return CPIPE_UNSUPPORTED;
            if (st[1] != 0) out->row_bytes = st[1];
        }
    }
    if (out->row_bytes % bpp != 0 || out->row_bytes < (size_t)out->w * bpp) return CPIPE_BAD_INDEX;
    return CPIPE_OK;
}

int map_engine_status(int rc, cpipe_host_t* host) {
    if (rc != CPIPE_OK) host_log(host, 3, std::string("denoise.bm3d: ") + cpipe_bm3d_last_error());
    return rc;
}

int process(cpipe_host_t* host, const cpipe_props_t* params, const cpipe_process_ctx* ctx) {
    const Suites s = fetch_suites(host);
    if (ctx == nullptr || ctx->inputs == nullptr || ctx->outputs == nullptr ||
        ctx->out_metadata == nullptr || ctx->n_in == 0 || ctx->n_in != ctx->n_out ||
        (ctx->n_in != 1 && ctx->n_in != 3) || ctx->out_metadata[0] == nullptr)
        return CPIPE_BAD_INDEX;
    for (size_t i = 0; i < ctx->n_in; ++i)
        if (ctx->inputs[i] == nullptr || ctx->outputs[i] == nullptr) return CPIPE_BAD_INDEX;
    if (s.buffer == nullptr || s.buffer->lock_cpu == nullptr || s.buffer->unlock_cpu == nullptr)
        return CPIPE_BAD_INDEX;

    std::vector<Image> in(ctx->n_in), out(ctx->n_out);
    for (size_t i = 0; i < ctx->n_in; ++i) {
        int rc = describe_buffer(s, ctx->inputs[i], &in[i]);
        if (rc != CPIPE_OK) return rc;
        rc = describe_buffer(s, ctx->outputs[i], &out[i]);
        if (rc != CPIPE_OK) return rc;
    }
    const std::string mode = enum_param(s, params, "mode", "CPIPE_BM3D_MODE", "compat");
    const std::string prec = enum_param(s, params, "precision", "CPIPE_BM3D_PRECISION", "fp32");
    const std::string pilot = enum_param(s, params, "pilot", "CPIPE_BM3D_PILOT", "off");
    const std::string nmap = enum_param(s, params, "noise_map", "CPIPE_BM3D_NOISE_MAP", "off");
    if (mode != "compat" && mode != "bm3d") return CPIPE_BAD_INDEX;
    if (pilot != "off" && pilot != "learned") return CPIPE_BAD_INDEX;
    if (pilot == "learned" && mode != "bm3d") return CPIPE_BAD_INDEX;
    if (nmap != "off" && nmap != "profile") return CPIPE_BAD_INDEX;
    if (nmap == "profile" && (mode != "bm3d" || pilot == "learned")) return CPIPE_BAD_INDEX;
    int flags = pilot == "learned" ? CPIPE_BM3D_HOST_PILOT : 0;
    int precision = CPIPE_BM3D_PRECISION_FP32;
    if (prec == "bf16") {
        precision = CPIPE_BM3D_PRECISION_BF16;
    } else if (prec == "fp16") {
        precision = CPIPE_BM3D_PRECISION_FP16;
    } else if (prec != "fp32") {
        return CPIPE_BAD_INDEX;
    }

    const bool rgba = ctx->n_in == 1 && in[0].format == CPIPE_FORMAT_R16G16B16A16_SFLOAT &&
                      out[0].format == CPIPE_FORMAT_R16G16B16A16_SFLOAT;
    bool planes = true;
    for (size_t i = 0; i < ctx->n_in; ++i)
        planes = planes && in[i].format == CPIPE_FORMAT_R32_SFLOAT &&
                 out[i].format == CPIPE_FORMAT_R32_SFLOAT;
    if (!rgba && !(planes && mode == "bm3d")) return CPIPE_BAD_PRECISION;
    for (size_t i = 0; i < ctx->n_in; ++i)
        if (in[i].w != in[0].w || in[i].h != in[0].h || out[i].w != in[0].w || out[i].h != in[0].h)
            return CPIPE_BAD_INDEX;

    const float sigma = resolve_sigma(s, params, ctx->inputs[0]);
    if (cpipe_bm3d_device_count() <= 0) {
        host_log(host, 3, "denoise.bm3d: no HIP device visible; the node runs on the GPU only");
        return CPIPE_UNSUPPORTED;
    }
    // noise_map=profile: a sigma per reference patch from the noise profile
    // (one pair per channel, else the pairs' mean for every channel); an explicit
    // sigma param still wins, and without a profile the scalar sigma is used
    if (nmap == "profile" && !explicit_sigma(s, params)) {
        std::vector<float> a, b;
        if (noise_profile(s, ctx->inputs[0], &a, &b)) {
            const size_t nc = rgba ? 3 : ctx->n_in;
            float pb[3] = {0.0F, 0.0F, 0.0F}, pl[3] = {0.0F, 0.0F, 0.0F};
            for (size_t c = 0; c < nc; ++c) {
                if (a.size() == nc) {
                    pb[c] = a[c];
                    pl[c] = b[c];
                } else {
                    for (size_t i = 0; i < a.size(); ++i) {
                        pb[c] += a[i];
                        pl[c] += b[i];
                    }
                    pb[c] /= static_cast<float>(a.size());
                    pl[c] /= static_cast<float>(a.size());
                }
            }
            if (cpipe_bm3d_host_noise_profile(pb, pl, (int)nc) != CPIPE_OK) return CPIPE_BAD_INDEX;
            flags |= CPIPE_BM3D_HOST_PROFILE;
        } else {
            host_log(host, 2, "denoise.bm3d: noise_map=profile without a noise profile; using one sigma");
        }
    }

    // One engine call on host (locked) or device addresses.
    auto run = [&](const std::vector<void*>& ip, const std::vector<void*>& op, bool dev_path, int device,
                   void* stream) -> int {
        const int w = (int)in[0].w, h = (int)in[0].h;
        if (rgba) {
            const auto* i16 = static_cast<const uint16_t*>(ip[0]);
            auto* o16 = static_cast<uint16_t*>(op[0]);
            const int m = mode == "bm3d" ? 1 : 0;
            return dev_path ? cpipe_bm3d_device_rgba16f(i16, in[0].row_bytes / 8, o16, out[0].row_bytes / 8, w,
                                                        h, sigma, m, precision, device, flags, stream)
                            : cpipe_bm3d_host_rgba16f(i16, in[0].row_bytes / 8, o16, out[0].row_bytes / 8, w, h,
                                                      sigma, m, precision, 0, flags);
        }
        const float* fi[3] = {};
        float* fo[3] = {};
        for (size_t c = 0; c < ctx->n_in; ++c) {
            fi[c] = static_cast<const float*>(ip[c]);
            fo[c] = static_cast<float*>(op[c]);
        }
        for (size_t c = 1; c < ctx->n_in; ++c)
            if (in[c].row_bytes != in[0].row_bytes || out[c].row_bytes != out[0].row_bytes)
                return CPIPE_UNSUPPORTED;
        const size_t ipitch = in[0].row_bytes / 4, opitch = out[0].row_bytes / 4;
        return dev_path ? cpipe_bm3d_device_planes(fi, ipitch, fo, opitch, w, h, (int)ctx->n_in, sigma,
                                                   precision, device, flags, stream)
                        : cpipe_bm3d_host_planes(fi, ipitch, fo, opitch, w, h, (int)ctx->n_in, sigma,
                                                 precision, 0, flags);
    };

    std::vector<void*> inp, outp;
    int device = 0;
    int rc = CPIPE_OK;
    if (device_resident(s, ctx, &inp, &outp, &device)) {
        void* stream = nullptr;
        rc = device_stream(s, ctx, &stream);
        if (rc == CPIPE_OK) rc = map_engine_status(run(inp, outp, true, device, stream), host);
        if (rc != CPIPE_OK) return rc;
        if (s.builder == nullptr || s.builder->add_applied_step == nullptr) return CPIPE_BAD_INDEX;
        return s.builder->add_applied_step(ctx->out_metadata[0], "denoise.bm3d");
    }
    inp.assign(ctx->n_in, nullptr);
    outp.assign(ctx->n_out, nullptr);
    size_t locked_in = 0, locked_out = 0;
    for (size_t i = 0; i < ctx->n_in && rc == CPIPE_OK; ++i) {
        rc = s.buffer->lock_cpu(const_cast<cpipe_buffer_t*>(ctx->inputs[i]), CPIPE_CPU_ACCESS_READ,
                                &inp[i]);
        if (rc == CPIPE_OK) ++locked_in;
    }
    for (size_t i = 0; i < ctx->n_out && rc == CPIPE_OK; ++i) {
        rc = s.buffer->lock_cpu(ctx->outputs[i], CPIPE_CPU_ACCESS_WRITE, &outp[i]);
        if (rc == CPIPE_OK) ++locked_out;
    }
    if (rc == CPIPE_OK) rc = map_engine_status(run(inp, outp, false, 0, nullptr), host);
    for (size_t i = 0; i < locked_out; ++i) {
        s.buffer->unlock_cpu(ctx->outputs[i]);
        if (rc == CPIPE_OK && s.buffer->flush_cpu_writes != nullptr)
            s.buffer->flush_cpu_writes(ctx->outputs[i]);
    }
    for (size_t i = 0; i < locked_in; ++i)
        s.buffer->unlock_cpu(const_cast<cpipe_buffer_t*>(ctx->inputs[i]));
    if (rc != CPIPE_OK) return rc;
    if (s.builder == nullptr || s.builder->add_applied_step == nullptr) return CPIPE_BAD_INDEX;
    return s.builder->add_applied_step(ctx->out_metadata[0], "denoise.bm3d");
}

int node_main_entry(const char* action, cpipe_host_t* host, cpipe_node_t* node,
                    cpipe_props_t* params, void* in_ctx, void* out_ctx) {
    try {
        if (action == nullptr) return CPIPE_REPLY_DEFAULT;
        if (std::strcmp(action, CPIPE_ACTION_DESCRIBE) == 0) return CPIPE_REPLY_DEFAULT;
        if (std::strcmp(action, CPIPE_ACTION_CREATE) == 0) {
            if (out_ctx == nullptr) return CPIPE_BAD_INDEX;
            *static_cast<void**>(out_ctx) = new Bm3dNode();
            return CPIPE_OK;
        }
        if (std::strcmp(action, CPIPE_ACTION_DESTROY) == 0) {
            delete static_cast<Bm3dNode*>(in_ctx);
            return CPIPE_OK;
        }
        if (std::strcmp(action, CPIPE_ACTION_PROCESS) == 0) {
            if (node == nullptr) return CPIPE_BAD_INDEX;
            return process(host, params, static_cast<const cpipe_process_ctx*>(in_ctx));
        }
    } catch (...) {
        host_log(host, 4, "plugin dispatch threw an exception");
        return CPIPE_INTERNAL_ERROR;
    }
    return CPIPE_REPLY_DEFAULT;
}

__attribute__((used, section("cpipe_registry"), aligned(8))) const cpipe_plugin_desc_t kBm3dDescriptor = {
    CPIPE_ABI_MAJOR, CPIPE_ABI_MINOR, kNodeId, kNodeVersion, CPIPE_BM3D_MANIFEST_JSON,
    &node_main_entry};

// ---------------------------------------------------------------------------
// com.cpipe.denoise.wavelet_bayes_shrink: replaces DenoiseWaveletBayesShrink
// (src/cpipe/nodes/denoise/wavelet_bayes_shrink.cpp:27-47) and its
// CPIPE_REGISTER_NODE (:54-55). Behaviour kept: one input, one output and an
// output metadata builder, else CPIPE_BAD_INDEX (:31-35); chroma_strength read
// as a number, 1 when absent, clamped to [0, 2] (:37-38 with
// src/cpipe/nodes/ParamUtils.hpp:13-29); add_applied_step
// ("denoise.wavelet_bayes_shrink") on success (:46). The Halide submit becomes
// cpipe_wavelet_host_rgba16f (host buffers) or cpipe_wavelet_rgba16f
// (device-resident buffers). Added: a format/size gate (RGBA16F in and out of
// one size, else CPIPE_BAD_PRECISION / CPIPE_BAD_INDEX) and CPIPE_UNSUPPORTED
// (logged) when no HIP device is visible.
// ---------------------------------------------------------------------------

const char kWaveletId[] = "com.cpipe.denoise.wavelet_bayes_shrink";
const char kWaveletVersion[] = "2.0.0";

struct WaveletNode {
    int unused = 0;
};

float clamped_param_float_or(const Suites& s, const cpipe_props_t* params, const char* key,
                             float fallback, float lo, float hi) {
    double v = fallback;
    if (s.param != nullptr && s.param->get_double != nullptr && params != nullptr) {
        double d = 0.0;
        if (s.param->get_double(params, key, &d) == CPIPE_OK) v = d;
    }
    return std::clamp(static_cast<float>(v), lo, hi);
}

int process_wavelet(cpipe_host_t* host, const cpipe_props_t* params, const cpipe_process_ctx* ctx) {
    const Suites s = fetch_suites(host);
    if (ctx == nullptr || ctx->n_in != 1 || ctx->n_out != 1 || ctx->inputs == nullptr ||
        ctx->outputs == nullptr || ctx->inputs[0] == nullptr || ctx->outputs[0] == nullptr ||
        ctx->out_metadata == nullptr || ctx->out_metadata[0] == nullptr) {
        host_log(host, 4, "denoise.wavelet_bayes_shrink missing buffers");
        return CPIPE_BAD_INDEX;
    }
    if (s.buffer == nullptr || s.buffer->lock_cpu == nullptr || s.buffer->unlock_cpu == nullptr)
        return CPIPE_BAD_INDEX;
    Image in, out;
    int rc = describe_buffer(s, ctx->inputs[0], &in);
    if (rc != CPIPE_OK) return rc;
    rc = describe_buffer(s, ctx->outputs[0], &out);
    if (rc != CPIPE_OK) return rc;
    if (in.format != CPIPE_FORMAT_R16G16B16A16_SFLOAT || out.format != CPIPE_FORMAT_R16G16B16A16_SFLOAT)
        return CPIPE_BAD_PRECISION;
    if (in.w != out.w || in.h != out.h || in.w == 0 || in.h == 0) return CPIPE_BAD_INDEX;
    const float strength = clamped_param_float_or(s, params, "chroma_strength", 1.0F, 0.0F, 2.0F);
    if (cpipe_bm3d_device_count() <= 0) {
        host_log(host, 3, "denoise.wavelet_bayes_shrink: no HIP device visible; the node runs on the GPU only");
        return CPIPE_UNSUPPORTED;
    }
    std::vector<void*> dip, dop;
    int device = 0;
    if (device_resident(s, ctx, &dip, &dop, &device)) {
        void* stream = nullptr;
        rc = device_stream(s, ctx, &stream);
        if (rc == CPIPE_OK)
            rc = on_device(device, stream, [&] {
                return cpipe_wavelet_rgba16f(static_cast<const uint16_t*>(dip[0]), in.row_bytes / 8,
                                             static_cast<uint16_t*>(dop[0]), out.row_bytes / 8, (int)in.w,
                                             (int)in.h, strength, stream);
            });
        if (rc != CPIPE_OK) {
            host_log(host, 3, std::string("denoise.wavelet_bayes_shrink: ") + cpipe_wavelet_last_error());
            return rc;
        }
        if (s.builder == nullptr || s.builder->add_applied_step == nullptr) return CPIPE_BAD_INDEX;
        return s.builder->add_applied_step(ctx->out_metadata[0], "denoise.wavelet_bayes_shrink");
    }
    void* ip = nullptr;
    void* op = nullptr;
    rc = s.buffer->lock_cpu(const_cast<cpipe_buffer_t*>(ctx->inputs[0]), CPIPE_CPU_ACCESS_READ, &ip);
    if (rc != CPIPE_OK) return rc;
    rc = s.buffer->lock_cpu(ctx->outputs[0], CPIPE_CPU_ACCESS_WRITE, &op);
    if (rc != CPIPE_OK) {
        s.buffer->unlock_cpu(const_cast<cpipe_buffer_t*>(ctx->inputs[0]));
        return rc;
    }
    rc = cpipe_wavelet_host_rgba16f(static_cast<const uint16_t*>(ip), in.row_bytes / 8,
                                    static_cast<uint16_t*>(op), out.row_bytes / 8, (int)in.w, (int)in.h,
                                    strength);
    if (rc != CPIPE_OK)
        host_log(host, 3, std::string("denoise.wavelet_bayes_shrink: ") + cpipe_wavelet_last_error());
    s.buffer->unlock_cpu(ctx->outputs[0]);
    if (rc == CPIPE_OK && s.buffer->flush_cpu_writes != nullptr) s.buffer->flush_cpu_writes(ctx->outputs[0]);
    s.buffer->unlock_cpu(const_cast<cpipe_buffer_t*>(ctx->inputs[0]));
    if (rc != CPIPE_OK) return rc;
    if (s.builder == nullptr || s.builder->add_applied_step == nullptr) return CPIPE_BAD_INDEX;
    return s.builder->add_applied_step(ctx->out_metadata[0], "denoise.wavelet_bayes_shrink");
}

int wavelet_main_entry(const char* action, cpipe_host_t* host, cpipe_node_t* node,
                       cpipe_props_t* params, void* in_ctx, void* out_ctx) {
    try {
        if (action == nullptr) return CPIPE_REPLY_DEFAULT;
        if (std::strcmp(action, CPIPE_ACTION_DESCRIBE) == 0) return CPIPE_REPLY_DEFAULT;
        if (std::strcmp(action, CPIPE_ACTION_CREATE) == 0) {
            if (out_ctx == nullptr) return CPIPE_BAD_INDEX;
            *static_cast<void**>(out_ctx) = new WaveletNode();
            return CPIPE_OK;
        }
        if (std::strcmp(action, CPIPE_ACTION_DESTROY) == 0) {
            delete static_cast<WaveletNode*>(in_ctx);
            return CPIPE_OK;
        }
        if (std::strcmp(action, CPIPE_ACTION_PROCESS) == 0) {
            if (node == nullptr) return CPIPE_BAD_INDEX;
            return process_wavelet(host, params, static_cast<const cpipe_process_ctx*>(in_ctx));
        }
    } catch (...) {
        host_log(host, 4, "plugin dispatch threw an exception");
        return CPIPE_INTERNAL_ERROR;
    }
    return CPIPE_REPLY_DEFAULT;
}

__attribute__((used, section("cpipe_registry"), aligned(8))) const cpipe_plugin_desc_t kWaveletDescriptor = {
    CPIPE_ABI_MAJOR, CPIPE_ABI_MINOR, kWaveletId, kWaveletVersion, CPIPE_WAVELET_MANIFEST_JSON,
    &wavelet_main_entry};

// ---------------------------------------------------------------------------
// com.cpipe.precision_convert: replaces PrecisionConvert
// (src/cpipe/nodes/precision/precision_convert.cpp:40-80) and its
// CPIPE_REGISTER_NODE (:86). Behaviour kept: one input, one output and an output
// metadata builder, else CPIPE_BAD_INDEX (:49-52); the string param
// target_format, its status when absent (:54-57, ParamView::string in
// include/cpipe/sdk/sdk.hpp:557-569); the output format must be the target
// (:30-36,63-66) and the pair one of the four conversions (:13-28,68-72), else
// CPIPE_BAD_PRECISION; add_applied_step("precision_convert") (:78). The Halide
// submit becomes cpipe_convert_host (host buffers) or cpipe_convert
// (device-resident buffers). Added: CPIPE_UNSUPPORTED (logged) when no
// HIP device is visible, CPIPE_BAD_INDEX on a size mismatch.
// ---------------------------------------------------------------------------

const char kConvertId[] = "com.cpipe.precision_convert";
const char kConvertVersion[] = "2.0.0";

struct ConvertNode {
    int unused = 0;
};

int format_of_target(const std::string& t) {
    if (t == "R32_SFLOAT") return CPIPE_CONVERT_R32_SFLOAT;
    if (t == "R16G16B16A16_SFLOAT") return CPIPE_CONVERT_R16G16B16A16_SFLOAT;
    if (t == "R8G8B8A8_UNORM") return CPIPE_CONVERT_R8G8B8A8_UNORM;
    if (t == "R16G16B16A16_UNORM") return CPIPE_CONVERT_R16G16B16A16_UNORM;
    return -1;
}

size_t pixel_bytes(int format) {
    switch (format) {
        case CPIPE_CONVERT_R16_UINT: return 2;
        case CPIPE_CONVERT_R8G8B8A8_UNORM: return 4;
        case CPIPE_CONVERT_R32_SFLOAT: return 4;
        case CPIPE_CONVERT_R16G16B16A16_SFLOAT:
        case CPIPE_CONVERT_R16G16B16A16_UNORM: return 8;
        default: return 0;
    }
}

// Format, size and row bytes of a buffer of any of the node's formats.
int describe_any(const Suites& s, const cpipe_buffer_t* b, Image* out) {
    if (s.buffer == nullptr || s.buffer->get_format == nullptr || s.buffer->get_dims == nullptr)
        return CPIPE_BAD_INDEX;
    int rc = s.buffer->get_format(b, &out->format);
    if (rc != CPIPE_OK) return rc;
    uint8_t ndim = 0;
    uint32_t dims[8] = {};
    rc = s.buffer->get_dims(b, &ndim, dims);
    if (rc != CPIPE_OK) return rc;
    if (ndim < 2) return CPIPE_BAD_INDEX;
    out->w = dims[0];
    out->h = dims[1];
    const size_t bpp = pixel_bytes(out->format);
    out->row_bytes = (size_t)out->w * (bpp ? bpp : 1);
    if (bpp != 0 && s.buffer->get_stride != nullptr) {
        uint64_t st[8] = {};
        if (s.buffer->get_stride(b, st) == CPIPE_OK) {
            if (st[0] != 0 && st[0] != bpp) return CPIPE_UNSUPPORTED;
            if (st[1] != 0) out->row_bytes = st[1];
        }
    }
    return CPIPE_OK;
}

int process_convert(cpipe_host_t* host, const cpipe_props_t* params, const cpipe_process_ctx* ctx) {
    const Suites s = fetch_suites(host);
    if (ctx == nullptr || ctx->n_in != 1 || ctx->n_out != 1 || ctx->inputs == nullptr ||
        ctx->outputs == nullptr || ctx->inputs[0] == nullptr || ctx->outputs[0] == nullptr ||
        ctx->out_metadata == nullptr || ctx->out_metadata[0] == nullptr) {
        host_log(host, 4, "precision_convert missing buffers");
        return CPIPE_BAD_INDEX;
    }
    if (s.param == nullptr || s.param->get_enum == nullptr || params == nullptr) return CPIPE_NEED_PARAM;
    const char* target = nullptr;
    int rc = s.param->get_enum(params, "target_format", &target);
    if (rc != CPIPE_OK) return rc;
    if (target == nullptr) return CPIPE_NEED_PARAM;
    if (s.buffer == nullptr || s.buffer->lock_cpu == nullptr || s.buffer->unlock_cpu == nullptr)
        return CPIPE_BAD_INDEX;
    Image in, out;
    rc = describe_any(s, ctx->inputs[0], &in);
    if (rc != CPIPE_OK) return rc;
    rc = describe_any(s, ctx->outputs[0], &out);
    if (rc != CPIPE_OK) return rc;
    const int tf = format_of_target(target);
    if (tf < 0 || out.format != tf) {
        host_log(host, 4, "precision_convert target/output mismatch");
        return CPIPE_BAD_PRECISION;
    }
    if (!cpipe_convert_supported(in.format, tf)) {
        host_log(host, 4, "precision_convert unsupported conversion");
        return CPIPE_BAD_PRECISION;
    }
    if (in.w != out.w || in.h != out.h || in.w == 0 || in.h == 0) return CPIPE_BAD_INDEX;
    if (cpipe_bm3d_device_count() <= 0) {
        host_log(host, 3, "precision_convert: no HIP device visible; the node runs on the GPU only");
        return CPIPE_UNSUPPORTED;
    }
    std::vector<void*> dip, dop;
    int device = 0;
    if (device_resident(s, ctx, &dip, &dop, &device)) {
        void* stream = nullptr;
        rc = device_stream(s, ctx, &stream);
        if (rc == CPIPE_OK)
            rc = on_device(device, stream, [&] {
                return cpipe_convert(in.format, dip[0], in.row_bytes, tf, dop[0], out.row_bytes, (int)in.w,
                                     (int)in.h, stream);
            });
        if (rc != CPIPE_OK) {
            host_log(host, 3, std::string("precision_convert: ") + cpipe_convert_last_error());
            return rc;
        }
        if (s.builder == nullptr || s.builder->add_applied_step == nullptr) return CPIPE_BAD_INDEX;
        return s.builder->add_applied_step(ctx->out_metadata[0], "precision_convert");
    }
    void* ip = nullptr;
    void* op = nullptr;
    rc = s.buffer->lock_cpu(const_cast<cpipe_buffer_t*>(ctx->inputs[0]), CPIPE_CPU_ACCESS_READ, &ip);
    if (rc != CPIPE_OK) return rc;
    rc = s.buffer->lock_cpu(ctx->outputs[0], CPIPE_CPU_ACCESS_WRITE, &op);
    if (rc != CPIPE_OK) {
        s.buffer->unlock_cpu(const_cast<cpipe_buffer_t*>(ctx->inputs[0]));
        return rc;
    }
    rc = cpipe_convert_host(in.format, ip, in.row_bytes, tf, op, out.row_bytes, (int)in.w, (int)in.h);
    if (rc != CPIPE_OK) host_log(host, 3, std::string("precision_convert: ") + cpipe_convert_last_error());
    s.buffer->unlock_cpu(ctx->outputs[0]);
    if (rc == CPIPE_OK && s.buffer->flush_cpu_writes != nullptr) s.buffer->flush_cpu_writes(ctx->outputs[0]);
    s.buffer->unlock_cpu(const_cast<cpipe_buffer_t*>(ctx->inputs[0]));
    if (rc != CPIPE_OK) return rc;
    if (s.builder == nullptr || s.builder->add_applied_step == nullptr) return CPIPE_BAD_INDEX;
    return s.builder->add_applied_step(ctx->out_metadata[0], "precision_convert");
}

int convert_main_entry(const char* action, cpipe_host_t* host, cpipe_node_t* node,
                       cpipe_props_t* params, void* in_ctx, void* out_ctx) {
    try {
        if (action == nullptr) return CPIPE_REPLY_DEFAULT;
        if (std::strcmp(action, CPIPE_ACTION_DESCRIBE) == 0) return CPIPE_REPLY_DEFAULT;
        if (std::strcmp(action, CPIPE_ACTION_CREATE) == 0) {
            if (out_ctx == nullptr) return CPIPE_BAD_INDEX;
            *static_cast<void**>(out_ctx) = new ConvertNode();
            return CPIPE_OK;
        }
        if (std::strcmp(action, CPIPE_ACTION_DESTROY) == 0) {
            delete static_cast<ConvertNode*>(in_ctx);
            return CPIPE_OK;
        }
        if (std::strcmp(action, CPIPE_ACTION_PROCESS) == 0) {
            if (node == nullptr) return CPIPE_BAD_INDEX;
            return process_convert(host, params, static_cast<const cpipe_process_ctx*>(in_ctx));
        }
    } catch (...) {
        host_log(host, 4, "plugin dispatch threw an exception");
        return CPIPE_INTERNAL_ERROR;
    }
    return CPIPE_REPLY_DEFAULT;
}

__attribute__((used, section("cpipe_registry"), aligned(8))) const cpipe_plugin_desc_t kConvertDescriptor = {
    CPIPE_ABI_MAJOR, CPIPE_ABI_MINOR, kConvertId, kConvertVersion, CPIPE_CONVERT_MANIFEST_JSON,
    &convert_main_entry};

}  // namespace

// Section bounds, defined by the linker for the `cpipe_registry` section.
extern "C" const cpipe_plugin_desc_t __start_cpipe_registry[];
extern "C" const cpipe_plugin_desc_t __stop_cpipe_registry[];

extern "C" const cpipe_plugin_desc_t* cpipe_bm3d_node_descriptor(void) { return &kBm3dDescriptor; }
extern "C" const cpipe_plugin_desc_t* cpipe_wavelet_node_descriptor(void) { return &kWaveletDescriptor; }
extern "C" const cpipe_plugin_desc_t* cpipe_convert_node_descriptor(void) { return &kConvertDescriptor; }

// Descriptors of this library's `cpipe_registry` section, for hosts that dlopen
// the node instead of linking it (the reference host walks the section of its
// own executable, src/cpipe/runtime/Registry.cpp:6-26).
extern "C" size_t cpipe_registry_entries(const cpipe_plugin_desc_t** first) {
    if (first != nullptr) *first = __start_cpipe_registry;
    return (size_t)(__stop_cpipe_registry - __start_cpipe_registry);
}

// Link anchors with the reference nodes' names (src/cpipe/nodes/denoise/bm3d.cpp:115-117,
// wavelet_bayes_shrink.cpp:59-61, src/cpipe/nodes/precision/precision_convert.cpp:90-92),
// so a host's builtin_link.cpp pulls this object in from the static archive.
extern "C" void cpipe_link_builtin_denoise_bm3d(void) {}
extern "C" void cpipe_link_builtin_denoise_wavelet_bayes_shrink(void) {}
extern "C" void cpipe_link_builtin_precision_convert(void) {}
