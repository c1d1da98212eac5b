// SFMT-19937 on the host: the replay stream.  Parameters and recursion as
// published by Saito & Matsumoto (SFMT 1.5.1, dependency/SFMT-src-1.5.1:
// SFMT-params19937.h:5-17, SFMT-common.h:142-158, SFMT.c:170-200, 349-362).
#pragma once

#include <cstdint>

namespace trw {

class Sfmt19937 {
public:
    static constexpr int kWords = 624;

    explicit Sfmt19937(uint32_t seed) { seed_with(seed); }

    void seed_with(uint32_t seed);

    uint32_t next() {
        if (pos_ >= kWords) {
            regenerate();
            pos_ = 0;
        }
        return state_[pos_++];
    }

    // Writes the next n words.
    void fill(uint32_t *out, int64_t n);
    void skip(uint64_t n);

private:
    void regenerate();
    void certify();

    alignas(16) uint32_t state_[kWords];
    int pos_ = kWords;
};

}  // namespace trw
