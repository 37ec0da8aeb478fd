// DNG NoiseProfile ingest (include/cpipe_dng.h): a bounds-checked walk over the
// TIFF IFDs of a DNG held in memory, reading tag 51041. The reference reads its
// TIFF tags in src/cpipe/ingest/dng_opcode/OpcodeListParser.cpp but has no
// NoiseProfile among them; this fills the gap for
// CalibrationBlock::noise_profile (include/cpipe/core/CalibrationBlock.hpp:54).
#include "cpipe_dng.h"

#include <cmath>
#include <cstdint>
#include <cstring>

namespace {

constexpr int OK = 0, FAILED = 1, BAD_INDEX = 5;
constexpr uint16_t TAG_NEW_SUBFILE_TYPE = 254;
constexpr uint16_t TAG_SUB_IFDS = 330;
constexpr uint16_t TAG_NOISE_PROFILE = 51041;
constexpr uint16_t TYPE_SHORT = 3, TYPE_LONG = 4, TYPE_DOUBLE = 12, TYPE_IFD = 13;
constexpr int MAX_IFDS = 64;       // IFDs visited at most (guards against offset loops)
constexpr uint32_t MAX_PAIRS = 16;  // a NoiseProfile holds 1 or ColorPlanes pairs

class Bytes {
public:
    Bytes(const uint8_t* p, size_t n, bool big) : p_(p), n_(n), big_(big) {}
    bool has(uint64_t off, uint64_t len) const { return off <= n_ && len <= n_ - off; }
    bool u16(uint64_t off, uint16_t* v) const {
        if (!has(off, 2)) return false;
        *v = big_ ? (uint16_t)(p_[off] << 8 | p_[off + 1]) : (uint16_t)(p_[off + 1] << 8 | p_[off]);
        return true;
    }
    bool u32(uint64_t off, uint32_t* v) const {
        if (!has(off, 4)) return false;
        uint32_t r = 0;
        for (int i = 0; i < 4; ++i) r |= (uint32_t)p_[off + i] << (big_ ? 8 * (3 - i) : 8 * i);
        *v = r;
        return true;
    }
    bool f64(uint64_t off, double* v) const {
        if (!has(off, 8)) return false;
        uint64_t r = 0;
        for (int i = 0; i < 8; ++i) r |= (uint64_t)p_[off + i] << (big_ ? 8 * (7 - i) : 8 * i);
        std::memcpy(v, &r, 8);
        return true;
    }

private:
    const uint8_t* p_;
    size_t n_;
    bool big_;
};

struct Found {
    bool any = false;
    bool main_image = false;  // from an IFD with NewSubFileType 0
    uint32_t count = 0;       // doubles
    uint32_t offset = 0;
};

}  // namespace

extern "C" int cpipe_dng_noise_profile(const void* bytes, size_t size, float* beta, float* lambda,
                                       int max_pairs, int* n_pairs) {
    if (!bytes || !n_pairs || max_pairs < 0 || (max_pairs > 0 && (!beta || !lambda))) return BAD_INDEX;
    *n_pairs = 0;
    const uint8_t* p = static_cast<const uint8_t*>(bytes);
    if (size < 8) return FAILED;
    bool big;
    if (p[0] == 'I' && p[1] == 'I')
        big = false;
    else if (p[0] == 'M' && p[1] == 'M')
        big = true;
    else
        return FAILED;
    const Bytes b(p, size, big);
    uint16_t magic = 0;
    uint32_t ifd0 = 0;
    if (!b.u16(2, &magic) || magic != 42 || !b.u32(4, &ifd0)) return FAILED;

    uint32_t queue[MAX_IFDS];
    int head = 0, tail = 0;
    auto push = [&](uint32_t off) {
        if (off == 0 || tail >= MAX_IFDS) return;
        for (int i = 0; i < tail; ++i)
            if (queue[i] == off) return;  // visited or queued already
        queue[tail++] = off;
    };
    push(ifd0);
    Found best;
    while (head < tail) {
        const uint32_t off = queue[head++];
        uint16_t n = 0;
        if (!b.u16(off, &n) || !b.has((uint64_t)off + 2, (uint64_t)n * 12 + 4)) return FAILED;
        uint32_t subfile = 0;  // TIFF default: a main image
        Found here;
        for (uint32_t e = 0; e < n; ++e) {
            const uint64_t ent = (uint64_t)off + 2 + 12 * (uint64_t)e;
            uint16_t tag = 0, type = 0;
            uint32_t count = 0, value = 0;
            b.u16(ent, &tag);
            b.u16(ent + 2, &type);
            b.u32(ent + 4, &count);
            b.u32(ent + 8, &value);
            if (tag == TAG_NEW_SUBFILE_TYPE && count == 1) {
                if (type == TYPE_LONG) subfile = value;
                if (type == TYPE_SHORT) {
                    uint16_t s = 0;
                    b.u16(ent + 8, &s);
                    subfile = s;
                }
            } else if (tag == TAG_SUB_IFDS && (type == TYPE_LONG || type == TYPE_IFD) && count > 0) {
                if (count == 1) {
                    push(value);
                } else {
                    for (uint32_t k = 0; k < count && k < (uint32_t)MAX_IFDS; ++k) {
                        uint32_t sub = 0;
                        if (!b.u32((uint64_t)value + 4 * (uint64_t)k, &sub)) return FAILED;
                        push(sub);
                    }
                }
            } else if (tag == TAG_NOISE_PROFILE) {
                if (type != TYPE_DOUBLE || count == 0 || (count & 1u) || count > 2 * MAX_PAIRS) return FAILED;
                if (!b.has(value, (uint64_t)count * 8)) return FAILED;
                here.any = true;
                here.count = count;
                here.offset = value;
            }
        }
        if (here.any && (!best.any || (subfile == 0 && !best.main_image))) {
            best = here;
            best.main_image = subfile == 0;
        }
        uint32_t next = 0;
        if (!b.u32((uint64_t)off + 2 + 12 * (uint64_t)n, &next)) return FAILED;
        push(next);
    }
    if (!best.any) return OK;  // no NoiseProfile: *n_pairs = 0
    const int np = (int)(best.count / 2);
    for (int i = 0; i < np; ++i) {
        double s = 0.0, o = 0.0;
        b.f64((uint64_t)best.offset + 16 * (uint64_t)i, &s);
        b.f64((uint64_t)best.offset + 16 * (uint64_t)i + 8, &o);
        if (!std::isfinite(s) || !std::isfinite(o)) return FAILED;
        if (i < max_pairs) {
            beta[i] = (float)o;
            lambda[i] = (float)s;
        }
    }
    *n_pairs = np;
    return OK;
}
