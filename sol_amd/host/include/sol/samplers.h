// Random numbers on the host side: the PCG32 generator of sol (include/sol/samplers.h:20-46 of
// the reference: XSH-RR output, multiplier 6364136223846793005, increment 1, seeded by
// state = 0; next(); state += seed; next()) and its float conversion, which in sol comes from
// libstdc++'s std::uniform_real_distribution<float>(0, 1): one 32-bit draw x gives
// min(float(x) * 2^-32, largest float below 1).
// The procedural scene generators and ray generators draw from it.
#ifndef SOL_SAMPLERS_H
#define SOL_SAMPLERS_H

#include <cstdint>
#include <cmath>

namespace sol {

struct PcgGenerator {
    uint64_t state = 0;

    explicit PcgGenerator(uint64_t init_state) { seed(init_state); }

    void seed(uint64_t init_state) {
        state = 0;
        next();
        state += init_state;
        next();
    }

    uint32_t next() {
        uint64_t old = state;
        state = old * UINT64_C(6364136223846793005) + UINT64_C(1);
        uint32_t xorshifted = static_cast<uint32_t>(((old >> 18) ^ old) >> 27);
        uint32_t rot = static_cast<uint32_t>(old >> 59);
        return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
    }
};

// Uniform float in [0, 1) from one 32-bit draw.
inline float uint_to_unit_float(uint32_t x) {
    float f = static_cast<float>(x) * 0x1p-32f;
    return f < 0x1.fffffep-1f ? f : 0x1.fffffep-1f;
}

class PcgSampler {
public:
    explicit PcgSampler(uint64_t seed) : gen_(seed) {}
    float operator()() { return uint_to_unit_float(gen_.next()); }
    // Uniform in [a, b).
    float uniform(float a, float b) { float u = (*this)(); return a + (b - a) * u; }
    float log_uniform(float a, float b) {
        float u = (*this)();
        return std::exp(std::log(a) + (std::log(b) - std::log(a)) * u);
    }
private:
    PcgGenerator gen_;
};

} // namespace sol

#endif
