// Minimal cpipe host: the suites a node sees through the plugin ABI 0.4, a
// registry walker for the `cpipe_registry` ELF section, and a one-node runner.
//
// It mirrors what the reference host provides to a single node
// (src/cpipe/runtime/HostContext.cpp:257-283 get_suite, :297-390 buffer suite,
// :388-400 and :625-644 calibration / noise profile, :756-763 add_applied_step,
// :843-851 get_double; Registry.cpp:6-26 section walk and ABI filter) and the
// unit-test driver process_bm3d (tests/unit/test_node_denoise_bm3d.cpp:89-125):
// create -> process -> freeze metadata -> destroy. It is not the reference's
// Pipeline/Scheduler; it drives one node on host buffers.
//
// Buffers are host memory (CpuBuffer analogue) or HIP device memory (the
// device-resident kind, offered to nodes through the "hip_buffer" suite; its
// lock_cpu stages the pixels through a host copy, as a mapped VulkanBuffer
// gives the host a view of device memory, VulkanBuffer.hpp:38-40).
//
// Exposed as a small C API (mh_*) so Python tests and tools can drive nodes.
#include <dlfcn.h>
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "cpipe_node_abi.h"

struct cpipe_metadata_s {
    std::string cs_role;
    std::vector<std::string> applied_steps;
    bool has_calibration = false;
    std::vector<std::pair<float, float>> noise_profile;
};

struct cpipe_buffer_s {
    int format = 0;
    uint32_t w = 0, h = 0;
    uint64_t row_bytes = 0;
    uint64_t pixel_bytes = 0;
    std::vector<uint8_t> storage;
    uint8_t* data = nullptr;  // 64-byte aligned view into storage (the host copy of a device buffer)
    void* dev = nullptr;      // device-resident pixels (row_bytes * h), or null for a host buffer
    int device = -1;
    int locks = 0;
    int lock_access = 0;
    cpipe_metadata_s meta;
};

struct cpipe_props_s {
    std::map<std::string, double> numbers;
    std::map<std::string, std::string> enums;
};

struct cpipe_metadata_builder_s {
    cpipe_metadata_s meta;
};

namespace {

thread_local std::string g_log;

int bpp_of(int format) {
    switch (format) {
        case CPIPE_FORMAT_R16G16B16A16_SFLOAT: return 8;
        case CPIPE_FORMAT_R32_SFLOAT: return 4;
        case 8: return 16;  // R32G32B32A32_SFLOAT
        case 1: return 2;   // R16_UINT
        case 3: return 4;   // R8G8B8A8_UNORM
        case 16: return 8;  // R16G16B16A16_UNORM
        default: return 0;
    }
}

// ---- buffer suite --------------------------------------------------------
int b_get_dims(const cpipe_buffer_t* b, uint8_t* ndim, uint32_t d[8]) {
    if (!b || !ndim || !d) return CPIPE_BAD_INDEX;
    *ndim = 2;
    d[0] = b->w;
    d[1] = b->h;
    return CPIPE_OK;
}
int b_get_format(const cpipe_buffer_t* b, int* f) {
    if (!b || !f) return CPIPE_BAD_INDEX;
    *f = b->format;
    return CPIPE_OK;
}
int b_get_kind(const cpipe_buffer_t* b, int* k) {
    if (!b || !k) return CPIPE_BAD_INDEX;
    *k = 0;  // Image2D
    return CPIPE_OK;
}
int b_get_stride(const cpipe_buffer_t* b, uint64_t s[8]) {
    if (!b || !s) return CPIPE_BAD_INDEX;
    s[0] = b->pixel_bytes;
    s[1] = b->row_bytes;
    return CPIPE_OK;
}
int b_get_color_role(const cpipe_buffer_t* b, const char** r) {
    if (!b || !r) return CPIPE_BAD_INDEX;
    *r = b->meta.cs_role.c_str();
    return CPIPE_OK;
}
// Copies between a device buffer and its host copy on the buffer's device.
bool dev_copy(cpipe_buffer_t* b, bool to_device) {
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(b->device) != hipSuccess) return false;
    const size_t n = (size_t)b->row_bytes * b->h;
    const hipError_t e = to_device ? hipMemcpy(b->dev, b->data, n, hipMemcpyHostToDevice)
                                   : hipMemcpy(b->data, b->dev, n, hipMemcpyDeviceToHost);
    (void)hipSetDevice(prev);
    return e == hipSuccess;
}
int b_lock(cpipe_buffer_t* b, int access, void** p) {
    if (!b || !p) return CPIPE_BAD_INDEX;
    if (b->locks != 0) return CPIPE_FAILED;  // single outstanding lock, like CpuBuffer
    if (b->dev && access != CPIPE_CPU_ACCESS_WRITE && !dev_copy(b, false)) return CPIPE_FAILED;
    b->locks = 1;
    b->lock_access = access;
    *p = b->data;
    return CPIPE_OK;
}
int b_unlock(cpipe_buffer_t* b) {
    if (!b) return CPIPE_BAD_INDEX;
    const bool wrote = b->locks != 0 && b->lock_access != CPIPE_CPU_ACCESS_READ;
    b->locks = 0;
    if (b->dev && wrote && !dev_copy(b, true)) return CPIPE_FAILED;
    return CPIPE_OK;
}
int b_flush(cpipe_buffer_t* b) { return b ? CPIPE_OK : CPIPE_BAD_INDEX; }
int b_get_metadata(const cpipe_buffer_t* b, const cpipe_metadata_t** out) {
    if (!b || !out) return CPIPE_BAD_INDEX;
    *out = &b->meta;
    return CPIPE_OK;
}

// ---- hip_buffer suite ----------------------------------------------------------
int hb_get_device_ptr(const cpipe_buffer_t* b, void** p, int* device) {
    if (!b || !p || !device) return CPIPE_BAD_INDEX;
    if (!b->dev) return CPIPE_UNSUPPORTED;
    *p = b->dev;
    *device = b->device;
    return CPIPE_OK;
}
int hb_get_stream(cpipe_compute_t*, void** stream) {
    if (!stream) return CPIPE_BAD_INDEX;
    *stream = nullptr;  // the null stream: ordered with the host copies of lock_cpu
    return CPIPE_OK;
}

// ---- metadata suite --------------------------------------------------------
int m_noise_profile(const cpipe_metadata_t* m, size_t max, size_t* n, float* a, float* b) {
    if (!m || !n || (max > 0 && (!a || !b))) return CPIPE_BAD_INDEX;
    if (!m->has_calibration) {
        *n = 0;
        return CPIPE_OK;
    }
    *n = m->noise_profile.size();
    for (size_t i = 0; i < max && i < m->noise_profile.size(); ++i) {
        a[i] = m->noise_profile[i].first;
        b[i] = m->noise_profile[i].second;
    }
    return CPIPE_OK;
}
int m_get_calibration(const cpipe_metadata_t* m, cpipe_calibration_view* out) {
    if (!m || !out) return CPIPE_BAD_INDEX;
    *out = cpipe_calibration_view{};
    out->get_noise_profile = &m_noise_profile;
    return CPIPE_OK;
}
int m_get_cs_role(const cpipe_metadata_t* m, const char** out) {
    if (!m || !out) return CPIPE_BAD_INDEX;
    *out = m->cs_role.c_str();
    return CPIPE_OK;
}
int m_has_step(const cpipe_metadata_t* m, const char* step, int* out) {
    if (!m || !step || !out) return CPIPE_BAD_INDEX;
    *out = 0;
    for (const auto& s : m->applied_steps)
        if (s == step) *out = 1;
    return CPIPE_OK;
}
int m_list_steps(const cpipe_metadata_t* m, size_t max, size_t* n, const char** out) {
    if (!m || !n) return CPIPE_BAD_INDEX;
    *n = m->applied_steps.size();
    for (size_t i = 0; i < max && i < m->applied_steps.size(); ++i) out[i] = m->applied_steps[i].c_str();
    return CPIPE_OK;
}

// ---- metadata builder suite -------------------------------------------------
int mb_add_step(cpipe_metadata_builder_t* b, const char* step) {
    if (!b || !step) return CPIPE_BAD_INDEX;
    for (const auto& s : b->meta.applied_steps)
        if (s == step) return CPIPE_OK;
    b->meta.applied_steps.emplace_back(step);
    return CPIPE_OK;
}
int mb_set_cs_role(cpipe_metadata_builder_t* b, const char* role) {
    if (!b || !role) return CPIPE_BAD_INDEX;
    b->meta.cs_role = role;
    return CPIPE_OK;
}

// ---- param suite -------------------------------------------------------------
int p_get_double(const cpipe_props_t* p, const char* key, double* out) {
    if (!p || !key || !out) return CPIPE_NEED_PARAM;
    auto it = p->numbers.find(key);
    if (it == p->numbers.end()) return CPIPE_NEED_PARAM;
    *out = it->second;
    return CPIPE_OK;
}
int p_get_enum(const cpipe_props_t* p, const char* key, const char** out) {
    if (!p || !key || !out) return CPIPE_NEED_PARAM;
    auto it = p->enums.find(key);
    if (it == p->enums.end()) return CPIPE_NEED_PARAM;
    *out = it->second.c_str();
    return CPIPE_OK;
}

struct Host {
    cpipe_host_t host{};
    cpipe_buffer_suite_v1 buffer{};
    cpipe_metadata_suite_v1 metadata{};
    cpipe_metadata_builder_suite_v1 builder{};
    cpipe_param_suite_v1 param{};
    cpipe_compute_suite_v1 compute{};
    cpipe_inference_suite_v1 inference{};
    cpipe_hip_buffer_suite_v1 hip{};
    bool offer_hip = true;  // false: behave as a host without the suite
};

const void* h_get_suite(cpipe_host_t* self, const char* name, int version);
void h_log(cpipe_host_t*, int level, const char* msg) {
    g_log += "[" + std::to_string(level) + "] " + (msg ? msg : "") + "\n";
}

Host& host_singleton() {
    static Host h = [] {
        Host x;
        x.host.abi_major = CPIPE_ABI_MAJOR;
        x.host.abi_minor = CPIPE_ABI_MINOR;
        x.host.get_suite = &h_get_suite;
        x.host.log = &h_log;
        x.host.alloc = [](cpipe_host_t*, size_t n) { return std::malloc(n); };
        x.host.free = [](cpipe_host_t*, void* p) { std::free(p); };
        x.buffer = {&b_get_dims, &b_get_format, &b_get_kind, &b_get_stride, &b_get_color_role,
                    &b_lock, &b_unlock, &b_flush, &b_get_metadata};
        x.metadata.get_calibration = &m_get_calibration;
        x.metadata.get_cs_role = &m_get_cs_role;
        x.metadata.has_applied_step = &m_has_step;
        x.metadata.list_applied_steps = &m_list_steps;
        x.builder.add_applied_step = &mb_add_step;
        x.builder.set_cs_role = &mb_set_cs_role;
        x.param.get_double = &p_get_double;
        x.param.get_enum = &p_get_enum;
        x.hip.get_device_ptr = &hb_get_device_ptr;
        x.hip.get_stream = &hb_get_stream;
        return x;
    }();
    return h;
}

const void* h_get_suite(cpipe_host_t*, const char* name, int version) {
    if (version != 1 || !name) return nullptr;
    Host& h = host_singleton();
    if (!std::strcmp(name, "buffer")) return &h.buffer;
    if (!std::strcmp(name, "metadata")) return &h.metadata;
    if (!std::strcmp(name, "metadata_builder")) return &h.builder;
    if (!std::strcmp(name, "param")) return &h.param;
    if (!std::strcmp(name, "compute")) return &h.compute;
    if (!std::strcmp(name, "inference")) return &h.inference;
    if (!std::strcmp(name, "hip_buffer") && h.offer_hip) return &h.hip;
    return nullptr;
}

}  // namespace

namespace {
// Section bounds of a loaded library: its cpipe_registry_entries() accessor, or
// the linker-defined __start/__stop symbols when they are exported.
bool registry_bounds(void* h, const cpipe_plugin_desc_t** start, const cpipe_plugin_desc_t** stop) {
    using entries_fn = size_t (*)(const cpipe_plugin_desc_t**);
    if (auto fn = reinterpret_cast<entries_fn>(dlsym(h, "cpipe_registry_entries"))) {
        const size_t n = fn(start);
        *stop = *start + n;
        return *start != nullptr;
    }
    *start = static_cast<const cpipe_plugin_desc_t*>(dlsym(h, "__start_cpipe_registry"));
    *stop = static_cast<const cpipe_plugin_desc_t*>(dlsym(h, "__stop_cpipe_registry"));
    return *start != nullptr && *stop != nullptr;
}
}  // namespace

extern "C" {

const char* mh_log(void) { return g_log.c_str(); }
void mh_clear_log(void) { g_log.clear(); }

// Registry walk over a loaded library's `cpipe_registry` section (Registry.cpp:6-26):
// descriptors with abi_major != 0 or abi_minor > 4 are skipped.
const cpipe_plugin_desc_t* mh_find_node(const char* lib_path, const char* node_id) {
    void* h = dlopen(lib_path, RTLD_NOW | RTLD_LOCAL);
    if (!h) {
        g_log += std::string("dlopen: ") + dlerror() + "\n";
        return nullptr;
    }
    const cpipe_plugin_desc_t* start = nullptr;
    const cpipe_plugin_desc_t* stop = nullptr;
    if (!registry_bounds(h, &start, &stop)) {
        g_log += "no cpipe_registry section\n";
        return nullptr;
    }
    for (auto* d = start; d < stop; ++d) {
        if (d->abi_major != 0 || d->abi_minor > CPIPE_ABI_MINOR) continue;
        if (d->node_id && node_id && !std::strcmp(d->node_id, node_id)) return d;
    }
    return nullptr;
}

int mh_registry_count(const char* lib_path) {
    void* h = dlopen(lib_path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1;
    const cpipe_plugin_desc_t* start = nullptr;
    const cpipe_plugin_desc_t* stop = nullptr;
    if (!registry_bounds(h, &start, &stop)) return -1;
    return (int)(stop - start);
}

cpipe_buffer_t* mh_buffer_new(int format, uint32_t w, uint32_t h, uint64_t row_bytes) {
    const int bpp = bpp_of(format);
    if (bpp == 0 || w == 0 || h == 0) return nullptr;
    if (row_bytes == 0) row_bytes = (uint64_t)w * bpp;
    if (row_bytes < (uint64_t)w * bpp) return nullptr;
    auto* b = new cpipe_buffer_s();
    b->format = format;
    b->w = w;
    b->h = h;
    b->pixel_bytes = (uint64_t)bpp;
    b->row_bytes = row_bytes;
    b->storage.assign(row_bytes * h + 64, 0);
    const uintptr_t base = reinterpret_cast<uintptr_t>(b->storage.data());
    b->data = reinterpret_cast<uint8_t*>((base + 63) & ~uintptr_t(63));
    return b;
}
// A device-resident buffer on HIP device `device` (zero-filled), or null when
// the layout is bad or the device allocation fails (no device, out of memory).
cpipe_buffer_t* mh_buffer_new_device(int format, uint32_t w, uint32_t h, uint64_t row_bytes, int device) {
    cpipe_buffer_t* b = mh_buffer_new(format, w, h, row_bytes);
    if (!b) return nullptr;
    int prev = 0;
    if (device < 0 || hipGetDevice(&prev) != hipSuccess || hipSetDevice(device) != hipSuccess) {
        g_log += "device buffer: no such HIP device\n";
        delete b;
        return nullptr;
    }
    const size_t n = (size_t)b->row_bytes * h;
    hipError_t e = hipMalloc(&b->dev, n);
    if (e == hipSuccess) e = hipMemset(b->dev, 0, n);
    (void)hipSetDevice(prev);
    if (e != hipSuccess) {
        g_log += std::string("device buffer: ") + hipGetErrorString(e) + "\n";
        if (b->dev) (void)hipFree(b->dev);
        delete b;
        return nullptr;
    }
    b->device = device;
    return b;
}
void mh_buffer_free(cpipe_buffer_t* b) {
    if (b && b->dev) {
        int prev = 0;
        if (hipGetDevice(&prev) == hipSuccess && hipSetDevice(b->device) == hipSuccess) {
            (void)hipFree(b->dev);
            (void)hipSetDevice(prev);
        }
    }
    delete b;
}
int mh_buffer_is_device(const cpipe_buffer_t* b) { return b && b->dev ? 1 : 0; }
// lock_cpu / unlock_cpu of the buffer suite, for tools filling and reading buffers.
void* mh_buffer_lock(cpipe_buffer_t* b, int access) {
    void* p = nullptr;
    return b_lock(b, access, &p) == CPIPE_OK ? p : nullptr;
}
int mh_buffer_unlock(cpipe_buffer_t* b) { return b_unlock(b); }
// Whether get_suite offers "hip_buffer" (1, the default) or not (0).
void mh_offer_hip_buffer(int on) { host_singleton().offer_hip = on != 0; }
void* mh_buffer_data(cpipe_buffer_t* b) { return b ? b->data : nullptr; }
uint64_t mh_buffer_row_bytes(const cpipe_buffer_t* b) { return b ? b->row_bytes : 0; }
int mh_buffer_set_noise_profile(cpipe_buffer_t* b, const float* beta, const float* lambda, size_t n) {
    if (!b) return CPIPE_BAD_INDEX;
    b->meta.has_calibration = true;
    b->meta.noise_profile.clear();
    for (size_t i = 0; i < n; ++i) b->meta.noise_profile.emplace_back(beta[i], lambda[i]);
    return CPIPE_OK;
}
int mh_buffer_set_cs_role(cpipe_buffer_t* b, const char* role) {
    if (!b || !role) return CPIPE_BAD_INDEX;
    b->meta.cs_role = role;
    return CPIPE_OK;
}
int mh_buffer_add_step(cpipe_buffer_t* b, const char* step) {
    if (!b || !step) return CPIPE_BAD_INDEX;
    b->meta.applied_steps.emplace_back(step);
    return CPIPE_OK;
}
const char* mh_buffer_cs_role(const cpipe_buffer_t* b) { return b ? b->meta.cs_role.c_str() : ""; }
int mh_buffer_step_count(const cpipe_buffer_t* b) { return b ? (int)b->meta.applied_steps.size() : 0; }
const char* mh_buffer_step(const cpipe_buffer_t* b, int i) {
    if (!b || i < 0 || i >= (int)b->meta.applied_steps.size()) return nullptr;
    return b->meta.applied_steps[i].c_str();
}

cpipe_props_t* mh_params_new(void) { return new cpipe_props_s(); }
void mh_params_free(cpipe_props_t* p) { delete p; }
void mh_params_set_double(cpipe_props_t* p, const char* k, double v) { p->numbers[k] = v; }
void mh_params_set_enum(cpipe_props_t* p, const char* k, const char* v) { p->enums[k] = v; }

// One node run on host buffers: create -> process -> freeze -> destroy, like
// process_bm3d (test_node_denoise_bm3d.cpp:89-125). The output metadata starts
// as a copy of input 0's (inherits_from) and gets the builder's steps.
// Returns the process status; create/destroy failures return their status + 1000.
int mh_run_node(const cpipe_plugin_desc_t* desc, cpipe_props_t* params, cpipe_buffer_t** inputs,
                size_t n_in, cpipe_buffer_t** outputs, size_t n_out) {
    if (!desc || !desc->main_entry) return CPIPE_BAD_INDEX;
    Host& h = host_singleton();
    void* instance = nullptr;
    int rc = desc->main_entry(CPIPE_ACTION_CREATE, &h.host, nullptr, params, nullptr, &instance);
    if (rc != CPIPE_OK) return 1000 + rc;
    std::vector<cpipe_metadata_builder_s> builders(n_out);
    std::vector<cpipe_metadata_builder_t*> bptr(n_out);
    for (size_t i = 0; i < n_out; ++i) {
        if (n_in > 0) builders[i].meta = inputs[0]->meta;
        bptr[i] = &builders[i];
    }
    std::vector<const cpipe_buffer_t*> ins(inputs, inputs + n_in);
    cpipe_process_ctx ctx{nullptr, nullptr, ins.data(), n_in, outputs, n_out, bptr.data()};
    rc = desc->main_entry(CPIPE_ACTION_PROCESS, &h.host, reinterpret_cast<cpipe_node_t*>(instance),
                          params, &ctx, nullptr);
    for (size_t i = 0; i < n_out; ++i) outputs[i]->meta = builders[i].meta;
    const int drc = desc->main_entry(CPIPE_ACTION_DESTROY, &h.host, nullptr, nullptr, instance, nullptr);
    if (drc != CPIPE_OK) return 1000 + drc;
    return rc;
}

int mh_call_action(const cpipe_plugin_desc_t* desc, const char* action) {
    if (!desc || !desc->main_entry) return CPIPE_BAD_INDEX;
    return desc->main_entry(action, &host_singleton().host, nullptr, nullptr, nullptr, nullptr);
}

}  // extern "C"
