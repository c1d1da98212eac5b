#include "sfmt.h"

#include <cstring>

namespace trw {
namespace {

constexpr int kLanes = 156;   // 128-bit lanes
constexpr int kPos1 = 122;
constexpr int kShiftLeft = 18;
constexpr int kShiftRight = 11;
constexpr uint32_t kMask[4] = {0xdfffffefU, 0xddfecb7fU, 0xbffaffffU, 0xbffffff6U};
constexpr uint32_t kParity[4] = {0x00000001U, 0x00000000U, 0x00000000U, 0x13c9e684U};

struct Lane {
    uint64_t lo, hi;  // little-endian 128-bit value: word0..1 in lo, word2..3 in hi
};

inline Lane load(const uint32_t *p) {
    Lane v;
    v.lo = static_cast<uint64_t>(p[0]) | (static_cast<uint64_t>(p[1]) << 32);
    v.hi = static_cast<uint64_t>(p[2]) | (static_cast<uint64_t>(p[3]) << 32);
    return v;
}

inline void store(uint32_t *p, Lane v) {
    p[0] = static_cast<uint32_t>(v.lo);
    p[1] = static_cast<uint32_t>(v.lo >> 32);
    p[2] = static_cast<uint32_t>(v.hi);
    p[3] = static_cast<uint32_t>(v.hi >> 32);
}

// One lane of the recursion: a ^ (a << 8 bits as 128-bit) ^ ((b >> 11) & mask per word)
// ^ (c >> 8 bits as 128-bit) ^ (d << 18 per word).
inline Lane mix(Lane a, const uint32_t *b, Lane c, const uint32_t *d) {
    Lane x{a.lo << 8, (a.hi << 8) | (a.lo >> 56)};
    Lane y{(c.lo >> 8) | (c.hi << 56), c.hi >> 8};
    uint32_t t[4];
    for (int i = 0; i < 4; ++i) t[i] = ((b[i] >> kShiftRight) & kMask[i]) ^ (d[i] << kShiftLeft);
    Lane z = load(t);
    return Lane{a.lo ^ x.lo ^ y.lo ^ z.lo, a.hi ^ x.hi ^ y.hi ^ z.hi};
}

}  // namespace

void Sfmt19937::seed_with(uint32_t seed) {
    state_[0] = seed;
    for (uint32_t i = 1; i < static_cast<uint32_t>(kWords); ++i)
        state_[i] = 1812433253U * (state_[i - 1] ^ (state_[i - 1] >> 30)) + i;
    pos_ = kWords;
    certify();
}

void Sfmt19937::certify() {
    uint32_t inner = 0;
    for (int i = 0; i < 4; ++i) inner ^= state_[i] & kParity[i];
    for (int s = 16; s > 0; s >>= 1) inner ^= inner >> s;
    if (inner & 1U) return;
    for (int i = 0; i < 4; ++i)
        for (uint32_t bit = 1U; bit != 0; bit <<= 1)
            if (bit & kParity[i]) {
                state_[i] ^= bit;
                return;
            }
}

void Sfmt19937::regenerate() {
    const uint32_t *r1 = state_ + 4 * (kLanes - 2);
    const uint32_t *r2 = state_ + 4 * (kLanes - 1);
    for (int i = 0; i < kLanes; ++i) {
        uint32_t *cur = state_ + 4 * i;
        const uint32_t *b = state_ + 4 * ((i + kPos1) % kLanes);
        store(cur, mix(load(cur), b, load(r1), r2));
        r1 = r2;
        r2 = cur;
    }
}

void Sfmt19937::fill(uint32_t *out, int64_t n) {
    while (n > 0) {
        if (pos_ >= kWords) {
            regenerate();
            pos_ = 0;
        }
        int64_t take = kWords - pos_;
        if (take > n) take = n;
        std::memcpy(out, state_ + pos_, static_cast<size_t>(take) * sizeof(uint32_t));
        pos_ += static_cast<int>(take);
        out += take;
        n -= take;
    }
}

void Sfmt19937::skip(uint64_t n) {
    while (n > 0) {
        if (pos_ >= kWords) {
            regenerate();
            pos_ = 0;
        }
        uint64_t take = static_cast<uint64_t>(kWords - pos_);
        if (take > n) take = n;
        pos_ += static_cast<int>(take);
        n -= take;
    }
}

}  // namespace trw
