// Counter-based random stream that replaces the reference RandomGenerator
// (include/tools/randomGenerator.h:31-94: per-OMP-thread mt19937_64 on the
// CPU, cuRAND XORWOW states indexed by thread id on the GPU).  Philox4x32-10
// as published by Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as
// easy as 1, 2, 3" (SC'11): ten rounds of two 32x32->64 multiplies with a Weyl
// key schedule.  Because it has no state, every draw is addressable by its
// counter and the order in which agents are processed cannot change results.
//
// Stream layout (shared by every kernel, the host population builder and any
// CPU re-statement that wants identical draws):
//   key = {seed low 32 bits, seed high 32 bits}
//   ctr = {entity id (agent or location), time step, (phase << 20) | draw, replica}
// "draw" counts the draws an entity makes inside one phase of one step.
#pragma once
#include <stdint.h>

#include "ps_hd.h"

namespace ps {

struct u32x4 {
    uint32_t v[4];
};

PS_INLINE void mulhilo32(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

PS_INLINE u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#if defined(__HIPCC__)
#pragma unroll
#endif
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(M0, c.v[0], hi0, lo0);
        mulhilo32(M1, c.v[2], hi1, lo1);
        u32x4 n;
        n.v[0] = hi1 ^ c.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ c.v[3] ^ k1;
        n.v[3] = lo0;
        c = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// Phases of the stream (the "phase" field of the counter).
enum Phase : uint32_t {
    PH_INIT_DAYS = 1,      // initial daysBeforeNextState of every agent
    PH_IMM_CATEGORY = 2,   // vaccination-willingness draws (init)
    PH_MOVE = 3,           // per-step movement
    PH_PLAN_SCHOOLWORK = 4,  // midnight: school/work quarantine propagation
    PH_TESTING = 5,        // midnight testing
    PH_PROGRESSION = 6,    // midnight disease progression
    PH_OTHER_DISEASE = 7,  // midnight non-COVID hospitalisation
    PH_INFECT = 8,         // 8 + variant: per-step infection draw
    PH_MUTATION = 16,      // ExposeToMutation closure rule
    PH_VACCINATION = 17,   // vaccination / booster rollout
    PH_TEST_FLAG = 18,     // midnight testing: flagging locations
    PH_PLAN_UNDERAGE = 19, // midnight: children kept at home
    PH_DUMMY_MOVE = 20,    // DummyMovement: random relocation
    PH_POP_LOCATION = 32,  // dummy population: locations
    PH_POP_AGENT = 33,     // dummy population: agents
    PH_SYNTH = 40          // synthetic household population builder
};

// One draw == one Philox block.  The helpers below take what they need from it.
struct Stream {
    uint32_t k0, k1, id, step, phase, replica;
    uint32_t draw;

    PS_INLINE Stream(uint64_t seed, uint32_t id_, uint32_t step_, uint32_t phase_, uint32_t replica_ = 0)
        : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), id(id_), step(step_), phase(phase_), replica(replica_), draw(0) {}

    PS_INLINE u32x4 next() {
        u32x4 c;
        c.v[0] = id;
        c.v[1] = step;
        c.v[2] = (phase << 20) | (draw & 0xFFFFFu);
        c.v[3] = replica;
        ++draw;
        return philox4x32_10(c, k0, k1);
    }
    // uniform double in [0, 1): the top 53 bits of words 0..1
    PS_INLINE double unit() {
        u32x4 r = next();
        uint64_t x = ((uint64_t)r.v[0] << 32) | (uint64_t)r.v[1];
        return (double)(x >> 11) * (1.0 / 9007199254740992.0);
    }
    // uniform double in [0, max)
    PS_INLINE double real(double max) { return max * unit(); }
    // a draw whose value is not used: the counter moves on, no block is computed
    PS_INLINE void skip() { ++draw; }
    // uniform integer in [0, max): multiply-shift on word 2.  max == 0 returns
    // 0 without consuming a draw, as randomUnsigned(0) does.
    PS_INLINE uint32_t uint_below(uint32_t max) {
        if (max == 0) return 0u;
        u32x4 r = next();
        return (uint32_t)(((uint64_t)r.v[2] * (uint64_t)max) >> 32);
    }
};

}  // namespace ps
