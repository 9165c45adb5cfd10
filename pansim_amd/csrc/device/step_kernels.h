// Per-step kernels: movement (mover scan + dense mover processing) and
// infection.  Included once, by engine.hip.
#pragma once
#include "../common/philox.h"
#include "../common/psmath.h"
#include "engine_types.h"

namespace ps {
namespace dev {

// ---------------------------------------------------------------- TimeDay
// (hours, minutes) as unsigned chars, src/tools/timeDay.cpp
struct TD {
    uint8_t h, m;
};
__device__ __forceinline__ unsigned td_min(TD t) { return (unsigned)t.h * 60u + (unsigned)t.m; }
__device__ __forceinline__ TD td_from_min(unsigned mins) {
    TD t;
    t.h = (uint8_t)(mins / 60u);
    t.m = (uint8_t)(mins % 60u);
    return t;
}
__device__ __forceinline__ bool td_lt(TD a, TD b) { return a.h < b.h || (a.h == b.h && a.m < b.m); }
__device__ __forceinline__ TD td_plus(TD a, TD d) {
    unsigned mins = td_min(d);
    TD r = a;
    r.m = (uint8_t)(r.m + mins % 60u);
    unsigned of = r.m / 60u;
    r.m = (uint8_t)(r.m % 60u);
    r.h = (uint8_t)(r.h + mins / 60u + of);
    return r;
}
__device__ __forceinline__ TD td_minus(TD a, TD b) { return td_from_min(td_min(a) - td_min(b)); }
__device__ __forceinline__ TD ev_start(const pansim_event& e) { return TD{e.start_h, e.start_m}; }
__device__ __forceinline__ TD ev_end(const pansim_event& e) { return TD{e.end_h, e.end_m}; }
__device__ __forceinline__ TD ev_dur(const pansim_event& e) { return TD{e.dur_h, e.dur_m}; }

__device__ __forceinline__ long long fixed32(float x) { return (long long)((double)x * 4294967296.0); }

// wake step after a path that sets the countdown to v and returns (no decrement)
__device__ __forceinline__ uint32_t wake_keep(uint32_t t, uint32_t v) {
    unsigned long long w = (unsigned long long)t + 1ull + (unsigned long long)v;
    return w >= 0xFFFFFFFFull ? NEVER : (uint32_t)w;
}
// wake step after the normal path: countdown v, then decremented (wraps at 0)
__device__ __forceinline__ uint32_t wake_dec(uint32_t t, uint32_t v) {
    if (v == 0) return NEVER;
    unsigned long long w = (unsigned long long)t + (unsigned long long)v;
    return w >= 0xFFFFFFFFull ? NEVER : (uint32_t)w;
}

// Location record of l; a missing location (NOLOC, the result of a lookup
// for a type the agent has no entry for) reads as closed, never quarantined
// and of no type.
__device__ __forceinline__ LocRec lrec(const Locs& L, uint32_t l) {
    if (l < L.n) return L.rec[l];
    LocRec z;
    z.quarUntil = 0;
    z.closedUntil = 0;
    z.noWork = 0;
    z.type = LOC_TYPE_NONE;
    z.open = 0;
    z.essential = 0;
    z.flag = 0;
    return z;
}

// Active window of every schedule slot at every step of the day: the first
// event containing the time (ab) and the first event starting after it (ae),
// -1 when none; ae is the slot end when the last event contains the time
// (the scan of doMovement, movementPolicies.h:470-492).  With them, event
// ab's location type and duration (hours | minutes << 8), so that a mover
// with a single event in its window needs no read of the event itself.
__global__ void k_event_windows(const uint32_t* __restrict__ evOff, const pansim_event* __restrict__ ev, uint32_t nOff,
                                uint32_t stepsPerDay, uint32_t dt, int4* __restrict__ win) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nOff * stepsPerDay) return;
    const uint32_t step = g / nOff, off = g % nOff;
    const TD simTime = td_from_min(step * dt);
    const uint32_t eb = evOff[off], ee = evOff[off + 1];
    int ab = -1, ae = -1;
    for (uint32_t j = eb; j < ee; ++j) {
        const pansim_event& e = ev[j];
        bool in = !td_lt(simTime, ev_start(e)) && td_lt(simTime, ev_end(e));
        if (in && ab == -1) ab = (int)j;
        if (td_lt(simTime, ev_start(e))) {
            ae = (int)j;
            break;
        }
        if (j == ee - 1 && in) ae = (int)ee;
    }
    const pansim_event first = ab >= 0 ? ev[ab] : pansim_event{};
    win[g] = make_int4(ab, ae, (int)first.loc_type, (int)((uint32_t)first.dur_h | ((uint32_t)first.dur_m << 8)));
}

// ------------------------------------------------ findActualLocationForType
// The possible-location list of one agent: [b, e) in possLoc/possType, and
// its first home-type entry.
struct Csr {
    uint32_t b, e, home;
};
__device__ __forceinline__ Csr csr_of(const AgentRec& r) { return Csr{r.csrBegin, r.csrBegin + r.csrCount, r.home}; }

// movementPolicies.h:69-122; `rs` may be null when the type needs no draw.
// The reference's lookup takes the first entry, without a draw, for three
// types passed by the caller: home, school and work everywhere except in the
// planning helpers (checkUnderageAtHome, checkSchoolWorkQuarantine,
// movementPolicies.h:1073-1090, 1208-1213), which pass home, school and
// classroom: there a classroom is the first entry of its type and a
// workplace is drawn among the agent's workplaces.  `plan` selects the
// second set.
__device__ __forceinline__ bool first_match_type(const pansim_location_types& lt, uint32_t ty, bool plan) {
    return ty == lt.school || ty == (plan ? lt.classroom : lt.work);
}

__device__ uint32_t find_loc(const Agents& A, const Locs& L, const pansim_location_types& lt, Csr c, uint32_t ty,
                             int retry, Stream* rs, bool plan = false) {
    if (ty == lt.home) return c.home;
    const uint32_t b = c.b, e = c.e;
    if (first_match_type(lt, ty, plan)) {
        for (uint32_t i = b; i < e; ++i)
            if (A.possType[i] == ty) return A.possLoc[i];
        return NOLOC;
    }
    uint32_t counter = 0, pos = 0;
    for (uint32_t i = b; i < e; ++i)
        if (A.possType[i] == ty) {
            if (counter == 0) pos = i;
            ++counter;
        }
    if (counter == 1) return A.possLoc[pos];
    if (counter > 1) {
        uint32_t counter2 = counter, randIdx = counter2;
        int redraws = 0;
        for (;;) {
            while (randIdx == counter2) {
                if (redraws == 65) {
                    randIdx = (counter2 + 1) % counter;
                    break;
                }
                randIdx = rs->uint_below(counter);
                ++redraws;
            }
            counter2 = 0;
            for (uint32_t i = pos; i < e; ++i) {
                if (A.possType[i] == ty) {
                    if (counter2 == randIdx) {
                        if (retry && loc_special(L, A.possLoc[i]) && L.rec[A.possLoc[i]].open == 0) {
                            retry = 0;
                            break;
                        }
                        return A.possLoc[i];
                    }
                    ++counter2;
                }
            }
        }
    }
    return NOLOC;
}

// The location-type bytes of one agent's list, held in registers when the
// list has at most CSR_REG_MAX entries: entry k is byte (k & 3) of t[k >> 2].
// The type array is padded so that the aligned word loads stay inside it.
constexpr uint32_t CSR_REG_MAX = 16;
struct CsrReg {
    uint32_t b, n, home;
    uint32_t t[4];
};

// Agent i's list (record r): the types of a short list come from its
// 16-byte record in A.types16 (zero for a list of more than CSR_REG_MAX
// entries), one aligned load that does not wait for r.
__device__ __forceinline__ CsrReg load_csr(const Agents& A, uint32_t i, const AgentRec& r) {
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    const v4u t = *reinterpret_cast<const v4u*>(A.types16 + 4 * (size_t)i);
    CsrReg c;
    c.b = r.csrBegin;
    c.n = r.csrCount;
    c.home = r.home;
    c.t[0] = t.x;
    c.t[1] = t.y;
    c.t[2] = t.z;
    c.t[3] = t.w;
    return c;
}

// Entries of the register list whose type is ty, one bit per entry (bit k =
// entry k).  Four type bytes per word are compared at once: byte b of x is
// zero exactly when entry b matches, and t below has bit 7 of that byte set
// exactly then (the 7-bit sums cannot carry between bytes); the four bit-7
// flags are gathered into bits 28..31 by one multiplication.
__device__ __forceinline__ uint32_t csr_match(const CsrReg& c, uint32_t ty) {
    const uint32_t rep = ty * 0x01010101u;
    uint32_t mask = 0u;
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
        const uint32_t x = c.t[j] ^ rep;
        const uint32_t t = ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x | 0x7F7F7F7Fu);
        mask |= (((t >> 7) * 0x10204080u) >> 28) << (4u * j);
    }
    return mask & ((1u << c.n) - 1u);  // entries past the list (n <= 16 here)
}

// Entry index of the r-th (0-based) set bit of mask; r < popcount(mask).
__device__ __forceinline__ uint32_t nth_bit(uint32_t mask, uint32_t r) {
    for (uint32_t k = 0; k < r; ++k) mask &= mask - 1u;
    return (uint32_t)__builtin_ctz(mask);
}

// findActualLocationForType over the register copy of the types; lists longer
// than CSR_REG_MAX take the memory version.
__device__ __forceinline__ uint32_t find_loc(const Agents& A, const Locs& L, const pansim_location_types& lt,
                                             const CsrReg& c, uint32_t ty, int retry, Stream* rs, bool plan = false) {
    if (ty == lt.home) return c.home;
    if (c.n > CSR_REG_MAX) return find_loc(A, L, lt, Csr{c.b, c.b + c.n, c.home}, ty, retry, rs, plan);
    const uint32_t mask = ty <= 0xFFu ? csr_match(c, ty) : 0u;  // list types are bytes
    if (mask == 0u) return NOLOC;
    // the caller's first-match types, or a single entry: no draw
    if (first_match_type(lt, ty, plan) || (mask & (mask - 1u)) == 0u)
        return A.possLoc[c.b + (uint32_t)__builtin_ctz(mask)];
    const uint32_t counter = (uint32_t)__builtin_popcount(mask);
    uint32_t counter2 = counter, randIdx = counter;
    int redraws = 0;
    for (;;) {
        while (randIdx == counter2) {
            if (redraws == 65) {
                randIdx = (counter2 + 1) % counter;
                break;
            }
            randIdx = rs->uint_below(counter);
            ++redraws;
        }
        const uint32_t l = A.possLoc[c.b + nth_bit(mask, randIdx)];
        if (retry && loc_special(L, l) && L.rec[l].open == 0) {  // closed: draw another one, once
            retry = 0;
            counter2 = randIdx;
            continue;
        }
        return l;
    }
}

// The same for agent a, reading its record.
__device__ __forceinline__ uint32_t find_loc(const Agents& A, const Locs& L, const pansim_location_types& lt, uint32_t a,
                                             uint32_t ty, int retry, Stream* rs) {
    return find_loc(A, L, lt, csr_of(A.rec[a]), ty, retry, rs);
}

struct MoveArgs {
    Agents A;
    Locs L;
    Tables T;
    StepScratch S;
    pansim_step_params p;
    uint64_t seed;
    uint32_t replica;
    uint32_t curfewStep;  // all agents take the full path this step
    uint32_t stepOfDay;   // timestamp % steps per day
    uint32_t nEvents;
    uint32_t pendPar;  // parity of the pending-write counter used this step
    unsigned long long* counters;  // MOVE_CTR_SLOTS pairs (evaluated, relocated)
    // step constants (move_set_step): minutes into the day, steps left until
    // midnight, first step of the day
    uint32_t minutes, stepsUntilMidnight, dayStart;
    // run constants (set by the host): steps per day, quarantine length in
    // steps, steps per hour and per half hour, ceil(2^32 / time step)
    uint32_t spd, qLen, hourSteps, halfHourSteps;
    uint64_t dtMagic;
};

// The step-dependent members of m for timestamp ts (stepOfDay = ts % steps
// per day).
__host__ __device__ inline void move_set_step(MoveArgs& m, uint32_t ts, uint32_t stepOfDay) {
    m.p.timestamp = ts;
    m.stepOfDay = stepOfDay;
    m.curfewStep = (m.p.enable_curfew && m.p.curfew_begin == stepOfDay) ? 1u : 0u;
    m.minutes = stepOfDay * m.p.time_step;
    m.stepsUntilMidnight = (1440u - m.minutes) / m.p.time_step;
    m.dayStart = ts - stepOfDay;
}

// x / time step for x < 2^16 (minute counts of the day's clock): with
// magic = ceil(2^32 / d), floor(x * magic / 2^32) is exact for d <= 1440 and
// x < 2^32 / d (checked exhaustively for all such d and x < 2^16).
__device__ __forceinline__ uint32_t div_dt(const MoveArgs& m, uint32_t x) {
    return (uint32_t)(((uint64_t)x * m.dtMagic) >> 32);
}

constexpr uint32_t MOVE_BLOCK = 256;
// 32 chunks of 256 agents per block: 1221 blocks for 10M agents, one round
// of the 256 CUs x 5 resident blocks, so that a step with few movers
// (latency bound: one mover's dependent loads) costs one round, not two
constexpr uint32_t MOVE_ITEMS = 32;
constexpr uint32_t MOVE_TILE = MOVE_BLOCK * MOVE_ITEMS;  // agents per block
constexpr uint32_t MOVE_CTR_SLOTS = 64;

// A pending quarantine write to loc by the agent of input index `agent`; of
// the writes to one location in one step, the highest input index wins.
__device__ __forceinline__ void pend_write(const MoveArgs& m, uint32_t agent, uint32_t loc, uint32_t until) {
    unsigned long long key = ((unsigned long long)agent << 32) | (unsigned long long)until;
    unsigned long long old = atomicMax(&m.L.pendKey[loc], key);
    if (old == 0ull) {
        uint32_t slot = atomicAdd(&m.S.pendCount[m.pendPar], 1u);
        if (slot < m.S.pendCap) m.S.pendList[slot] = loc;
    }
}

// The location part of quarantineAgent (movementPolicies.h:202-241) for a
// diagnosed agent: the home, and with quarantinePolicy > 2 the classroom and
// workplace, get pending quarantine writes (applied at the end of the step).
// It runs once, after the rest of the agent's movement routine, so its
// classroom draw is the routine's last draw.
__device__ __forceinline__ void quarantine_locations(const MoveArgs& m, uint32_t key, const CsrReg& c, uint32_t until,
                                                     uint32_t prev, Stream& rs) {
    const pansim_location_types& lt = m.T.lt;
    const uint32_t ts = m.p.timestamp;
    if (m.p.quarantine_policy > 1 && c.home != NOLOC) pend_write(m, key, c.home, until);
    if (m.p.quarantine_policy > 2 && prev < ts) {
        uint32_t toClose[2] = {find_loc(m.A, m.L, lt, c, lt.classroom, 0, &rs), find_loc(m.A, m.L, lt, c, lt.work, 0, nullptr)};
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const uint32_t loc = toClose[k];
            if (loc == NOLOC) continue;
            const LocRec lr = m.L.rec[loc];
            if ((lr.type == lt.work || lr.type == lt.classroom) && lr.quarUntil < ts && lr.open && lr.closedUntil < ts)
                pend_write(m, key, loc, until);
        }
    }
}

// RealMovementOps::doMovement (movementPolicies.h:258-1041) for agent i.
// Writes location, wake step, flags and stats of agent i; keeps the
// per-location occupancy and infectiousness sums up to date.  Returns
// whether the agent changed location.
__device__ __forceinline__ bool do_movement(const MoveArgs& m, uint32_t i) {
    // one 32-byte load: every agent field the routine reads
    const AgentRec r0 = m.A.rec[i];
    AgentRec r = r0;
    const CsrReg csr = load_csr(m.A, i, r0);
    const Agents& A = m.A;
    const Locs& L = m.L;
    const Tables& T = m.T;
    const pansim_location_types& lt = T.lt;
    const uint32_t ts = m.p.timestamp;
    const unsigned minutes = m.minutes;
    const TD simTime = td_from_min(minutes);
    const uint32_t stepsUntilMidnight = m.stepsUntilMidnight;
    const uint32_t qLen = m.qLen;
    const uint32_t agentType = r.type;
    const uint32_t oldLoc = A.location[i];
    uint32_t loc = oldLoc;
    const uint32_t dayStart = m.dayStart;
    // a mover's wake step is this step, except on the curfew step (everyone)
    const uint16_t wake0 = m.curfewStep ? A.wake[i] : (uint16_t)m.stepOfDay;
    uint32_t wake = wake0 == WAKE_NONE ? NEVER : dayStart + wake0;
    const uint32_t key = A.id[i];  // the agent's input index: random streams and tie-breaks
    Stream rs(m.seed, key, ts, PH_MOVE, m.replica);
    const uint8_t st = r.state;
    const uint8_t w = r.wb;  // T.stWB[st], kept in the record
    // The loads below depend only on the agent's record and location: they
    // are issued together, ahead of the branches that use them, and their
    // values are combined only where used, so that a wave waits for one
    // round trip where a branch per load would wait for one each.
    // The schedule window of this step (agentType.cpp:13-15, precomputed per
    // step of the day; a dead agent, w = 4, has no window: entry 0 stands
    // in, unused).
    const int4 win = T.evWin[(size_t)m.stepOfDay * T.nOff +
                             (w <= 3u ? 7u * w + m.p.day_of_week + agentType * 28u : 0u)];
    // state and variant are fixed during the routine
    const float stInfV = T.stInf[st], varMulV = T.varMul[r.variant];
    const uint8_t stInfectedV = T.stInfected[st];
    // the special-bitmap word of the agent's present location (oldLoc is a
    // location of this population, below L.n)
    const uint32_t oldWord = L.special[oldLoc >> 5];
    // the window's first and following events, read from valid indices
    // (entry 0 otherwise) and replaced by the empty event when absent, so
    // that neither load waits for a branch
    const bool hasFirst = win.x >= 0, hasNext = win.y >= 0 && (uint32_t)win.y < m.nEvents;
    const pansim_event e0 = T.ev[hasFirst ? (uint32_t)win.x : 0u];
    const pansim_event e1 = T.ev[hasNext ? (uint32_t)win.y : 0u];
    const pansim_event none = {};
    const pansim_event evFirst = hasFirst ? e0 : none, evNext = hasNext ? e1 : none;
    const float xInf = stInfV * varMulV;
    const bool infectedState = stInfectedV != 0;
    const bool oldSpecial = ((oldWord >> (oldLoc & 31u)) & 1u) != 0u;
    // the last location record read (movement writes no location record);
    // it starts as the record of a missing location (NOLOC), which a lookup
    // of a type the agent has no entry for returns
    uint32_t lastL = NOLOC;
    LocRec lastR = lrec(L, NOLOC);
    auto lrec_c = [&](uint32_t l) {
        if (l != lastL) {
            lastR = lrec(L, l);
            lastL = l;
        }
        return lastR;
    };
    // the last special bit read (loc_special), starting with the agent's
    // present location: the checks of one destination share one read
    uint32_t lastSL = oldLoc;
    bool lastS = oldSpecial;
    auto special_c = [&](uint32_t l) {
        if (l != lastSL) {
            lastS = loc_special(L, l);
            lastSL = l;
        }
        return lastS;
    };

    // quarantineAgent (movementPolicies.h:171-245) without its location
    // writes: an agent is quarantined at most once per routine, and only a
    // diagnosed one leaves location writes, done after the routine
    bool qLoc = false;
    uint32_t qLocUntil = 0, qLocPrev = 0;
    auto quarantine = [&](uint32_t until) {
        if (m.p.quarantine_policy == 0) return;
        if (r.diagTs > 0 && !infectedState && m.p.quarantine_immune_active == 0) return;
        if (!(r.flags & F_DIAGNOSED) && A.immTs[i] > 0 && m.p.quarantine_immune_active == 0) return;
        r.flags |= F_QUARANTINED;
        A.quarTs[i] = ts;
        const uint32_t prev = r.quarUntil;
        r.quarUntil = until;
        A.daysInQ[i] += (until - ts) / m.spd;
        if (r.flags & F_DIAGNOSED) {
            qLoc = true;
            qLocUntil = until;
            qLocPrev = prev;
        }
    };

    bool done = false;
    // curfew (movementPolicies.h:267-326)
    if (m.p.enable_curfew && m.p.curfew_begin == m.stepOfDay) {
        bool deadOrHosp = (w == 4 || w == 3);
        bool hospNonCovid = r.hospTs <= ts && r.hospUntil > ts;
        if (!deadOrHosp && !hospNonCovid) {
            if (agentType + 1 == 7) {
                uint32_t workplace = find_loc(A, L, lt, csr, lt.work, 0, nullptr);
                const bool workOpen = !loc_special(L, workplace) || lrec_c(workplace).open != 0;
                if (loc != workplace || !workOpen) {
                    loc = csr.home;
                    if (!workOpen) {
                        wake = wake_keep(ts, stepsUntilMidnight);
                        done = true;
                    }
                }
            } else {
                wake = wake_keep(ts, stepsUntilMidnight);
                loc = csr.home;
                done = true;
            }
        }
    }
    if (!done && wake != ts) done = true;  // countdown > 0: nothing else happens this step

    if (!done) do {
        if (r.quarUntil <= ts) r.flags &= (uint8_t)~F_QUARANTINED;
        if (w == 4) {  // dead
            wake = NEVER;
            loc = L.cemetery;
            break;
        }
        const uint32_t agentHome = csr.home;
        uint32_t sum = 0;
        auto diagnose = [&]() {
            if (!(r.flags & F_DIAGNOSED) && xInf > 0.0f) {
                r.flags |= F_DIAGNOSED;
                uint32_t d = ts;
                if (stepsUntilMidnight == m.spd) d++;
                r.diagTs = d;
                quarantine(ts + qLen);
            }
        };
        if (r.hospTs <= ts && r.hospUntil > ts && w != 3 && w != 4) {
            uint32_t v = r.hospUntil - ts - 1;
            sum = v < stepsUntilMidnight ? v : stepsUntilMidnight;
            loc = find_loc(A, L, lt, csr, lt.hospital, 0, &rs);
            diagnose();
            wake = wake_keep(ts, sum);
            break;
        }
        if (w == 3) {
            sum = stepsUntilMidnight;
            loc = find_loc(A, L, lt, csr, lt.hospital, 0, &rs);
            diagnose();
            wake = wake_keep(ts, sum);
            break;
        }
        auto quarantinable = [&](const LocRec& lr) {
            uint32_t t = lr.type;
            return t == lt.home || t == lt.school || t == lt.classroom || t == lt.work;
        };
        auto stayedHomeCheck = [&]() {
            if (agentHome != NOLOC && loc != agentHome && sum > m.hourSteps) r.flags &= (uint8_t)~F_STAYED_HOME;
        };
        // (a location that is not special is not quarantined today)
        if (m.p.quarantine_policy > 1 && special_c(loc)) {
            const LocRec lr = lrec_c(loc);
            if (ts < lr.quarUntil && quarantinable(lr)) {
                if (!(r.flags & F_QUARANTINED)) quarantine(lr.quarUntil);
                if (r.flags & F_QUARANTINED) {
                    uint32_t v = lr.quarUntil - ts - 1;
                    sum = v < stepsUntilMidnight ? v : stepsUntilMidnight;
                    if (csr.home != loc) loc = csr.home;
                    stayedHomeCheck();
                    wake = wake_keep(ts, sum);
                    break;
                }
            }
        }
        {
            bool qp = m.p.quarantine_policy > 0;
            uint32_t dts = r.diagTs;
            if ((qp && ((r.flags & F_DIAGNOSED) || (dts > 0 && (ts - dts) < qLen))) || (qp && (r.flags & F_QUARANTINED))) {
                if (qp && !(r.flags & F_QUARANTINED)) quarantine(ts + qLen);
                if (r.flags & F_QUARANTINED) {
                    loc = csr.home;
                    sum = stepsUntilMidnight;
                    stayedHomeCheck();
                    wake = wake_keep(ts, sum);
                    break;
                }
            }
        }
        const int ab = win.x, ae = win.y;
        uint32_t newType = NOLOC;
        if (ab == -1 && ae == -1) {  // Case 1
            newType = (w == 3) ? lt.hospital : lt.home;
            loc = find_loc(A, L, lt, csr, newType, 0, &rs);
            sum = stepsUntilMidnight;
        }
        if (ab != -1) {  // Cases 2 and 4
            uint32_t nPot = (uint32_t)((ae == -1 ? ab + 1 : ae) - ab);
            uint32_t picked = 0;
            TD dur;
            if (nPot == 1) {  // event ab, from the window record
                newType = (uint32_t)win.z;
                dur = TD{(uint8_t)(win.w & 0xFF), (uint8_t)((win.w >> 8) & 0xFF)};
            } else {
                double r = rs.real(1.0);
                double thr = evFirst.chance;
                uint32_t idx = 0;
                // the last event read: the picked one when the loop ends (an
                // exhausted loop steps back to the last one it read)
                pansim_event cur = evFirst;
                while (r > thr && idx < nPot) {
                    idx++;
                    if (idx < nPot) {
                        cur = T.ev[ab + idx];
                        thr += cur.chance;
                    }
                }
                if (idx == nPot) idx--;
                picked = idx;
                newType = cur.loc_type;
                dur = ev_dur(cur);
            }
            if ((r.flags & F_NO_WORK) && newType == lt.work) {
                newType = lt.home;
                nPot = 1;
                if (!(r.flags & F_QUARANTINED)) A.daysInQ[i] += 1;
            }
            uint32_t newLoc = find_loc(A, L, lt, csr, newType, nPot == 1, &rs);
            bool tooOld = (newType == lt.school || newType == lt.classroom) && r.age >= m.p.school_age_restriction;
            bool lockdown = m.p.lockdown_nonvacc_active && A.immTs[i] == 0;
            auto closedHere = [&](uint32_t l, uint32_t t) {
                if (t == lt.work || !special_c(l)) return false;
                const LocRec lr = lrec_c(l);
                return lr.open == 0 || lr.closedUntil > ts;
            };
            if (tooOld || closedHere(newLoc, newType) || lockdown) {
                if (nPot > 1) {
                    double r = rs.real(1.0);
                    double denom = 1.0 - (double)T.ev[ab + picked].chance;
                    double thr = (picked == 0) ? 0.0 : (double)T.ev[ab].chance / denom;
                    uint32_t idx = 0;
                    while (r > thr && idx < nPot) {
                        idx++;
                        if (idx < nPot) thr += (picked == idx) ? 0.0 : (double)T.ev[ab + idx].chance / denom;
                    }
                    if (idx == nPot) idx--;
                    newType = T.ev[ab + idx].loc_type;
                    dur = ev_dur(T.ev[ab + idx]);
                    newLoc = find_loc(A, L, lt, csr, newType, 0, &rs);
                    if (tooOld || closedHere(newLoc, newType) || lockdown) newLoc = csr.home;
                } else {
                    newLoc = csr.home;
                }
            } else if (newType == lt.school) {
                uint32_t cls = find_loc(A, L, lt, csr, lt.classroom, 0, &rs);
                if (cls != NOLOC && loc_special(L, cls) && L.rec[cls].closedUntil > ts) newLoc = csr.home;
            }
            loc = newLoc;
            if (dur.h > 24) {
                sum = stepsUntilMidnight;
            } else if (ae == -1) {
                TD arrive = td_plus(simTime, dur);
                sum = arrive.h >= 24 ? stepsUntilMidnight : div_dt(m, td_min(dur));
            } else {
                // ae can equal the slot end: the reference reads the next slot's first
                // event, or a default event past the end of the array
                pansim_event ne;
                if ((uint32_t)ae < m.nEvents) ne = evNext;
                else {
                    ne.loc_type = 0;
                    ne.chance = -1.0f;
                    ne.start_h = ne.start_m = ne.end_h = ne.end_m = ne.dur_h = ne.dur_m = 0;
                }
                TD arrive = td_plus(simTime, dur);
                if (td_lt(arrive, ev_start(ne))) sum = div_dt(m, td_min(dur));
                else if (td_lt(ev_end(ne), arrive)) sum = div_dt(m, td_min(td_minus(ev_end(ne), simTime))) - 1u;
                else {
                    uint32_t stp = div_dt(m, td_min(td_minus(ev_end(ne), arrive)));
                    uint32_t extra = rs.uint_below(stp);
                    sum = div_dt(m, td_min(dur)) + extra;
                }
            }
            if (agentType + 1 == 7 && m.p.enable_curfew &&
                (m.p.curfew_begin <= m.stepOfDay || m.p.curfew_end > m.stepOfDay)) {
                bool workButClosed = false;
                if (newType == lt.work && loc_special(L, newLoc)) {
                    const LocRec lr = lrec_c(newLoc);
                    workButClosed = lr.open == 0 || lr.closedUntil > ts;
                }
                if (workButClosed || (newType != lt.work && newType != lt.home)) loc = agentHome;
            }
        }
        if (ab == -1 && ae != -1) {  // Case 3 (ae is an event of the slot: ae < nEvents)
            const pansim_event& ne = evNext;
            uint32_t lenSteps = div_dt(m, td_min(td_minus(ev_end(ne), ev_start(ne))));
            uint32_t delay = rs.uint_below(lenSteps);
            sum = div_dt(m, td_min(td_minus(ev_start(ne), simTime))) + delay;
            if (sum < m.halfHourSteps) {
            } else if (sum < m.hourSteps) {
                newType = lt.public_space;
                loc = find_loc(A, L, lt, csr, lt.public_space, 0, &rs);
            } else {
                newType = lt.home;
                loc = csr.home;
            }
        }
        // (a location that is not special is not quarantined today)
        if (m.p.quarantine_policy > 1 && special_c(loc)) {
            const LocRec lr = lrec_c(loc);
            if (ts < lr.quarUntil && quarantinable(lr)) {
                if (!(r.flags & F_QUARANTINED)) quarantine(lr.quarUntil);
                if (r.flags & F_QUARANTINED) {
                    uint32_t v = lr.quarUntil - ts - 1;
                    sum = v < stepsUntilMidnight ? v : stepsUntilMidnight;
                    if (csr.home != loc) loc = csr.home;
                    stayedHomeCheck();
                    wake = wake_keep(ts, sum);
                    break;
                }
            }
        }
        if (newType == lt.hospital || (newType == lt.doctor && rs.real(1.0) < 0.7)) diagnose();
        stayedHomeCheck();
        wake = wake_dec(ts, sum);
    } while (false);

    if (qLoc) quarantine_locations(m, key, csr, qLocUntil, qLocPrev, rs);
    if (r.flags != r0.flags) A.rec[i].flags = r.flags;
    if (r.diagTs != r0.diagTs) A.rec[i].diagTs = r.diagTs;
    if (r.quarUntil != r0.quarUntil) A.rec[i].quarUntil = r.quarUntil;
    // a wake step beyond today is never reached: the midnight planning resets it
    const uint16_t wake16 = (wake == NEVER || wake - dayStart >= m.spd) ? WAKE_NONE : (uint16_t)(wake - dayStart);
    if (wake16 != wake0) A.wake[i] = wake16;
    // a location type the agent has no entry for leaves it where it is
    if (loc == NOLOC) loc = oldLoc;
    if (loc == oldLoc) return false;
    A.location[i] = loc;
    const uint8_t var = r.variant;
    if (xInf != 0.0f) {
        long long q = fixed32(xInf);
        atomicAdd((unsigned long long*)loc_sum(L, var, oldLoc), (unsigned long long)(-q));
        atomicAdd((unsigned long long*)loc_sum(L, var, loc), (unsigned long long)q);
        mark_visited(L, loc);
    }
    atomicSub(&L.dyn[oldLoc].count, 1u);
    atomicAdd(&L.dyn[loc].count, 1u);
    return true;
}

// One movement step for block blk of the nBlk blocks working on this
// population.  The block takes the 256-agent chunks blk, blk + nBlk,
// blk + 2 nBlk, ... (ITEMS of them): agents that move at the same steps (one
// schedule's agents, which a population may store next to each other) are
// spread over every block instead of loading a few.  The agents whose wake
// step is this step (everyone on the curfew step) are compacted into LDS,
// then processed by the same block with do_movement.  The order in which
// movers are processed does not change any result: the per-agent writes are
// private, the per-location sums are integer atomics and the location
// quarantine writes are deferred to the end of the step.  Evaluated /
// relocated counts go to one of MOVE_CTR_SLOTS counter pairs.
template <uint32_t ITEMS>
__device__ __forceinline__ void move_tile(const MoveArgs& m, uint32_t blk, uint32_t nBlk) {
    static_assert(MOVE_BLOCK == 256 && ITEMS <= 256, "list entries are (chunk << 8) | agent of the chunk");
    __shared__ uint16_t list[MOVE_BLOCK * ITEMS];  // movers, as (chunk of this block << 8) | agent of the chunk
    __shared__ uint32_t nList, nMoved;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t n = m.A.n;
    if (tid == 0) {
        nList = 0;
        nMoved = 0;
    }
    // wave v scans the chunks v, v + 4, ... whole: one 8-byte wake load per
    // chunk and lane (lane l holds agents 4l .. 4l + 3 of the chunk; the wake
    // array is padded to whole movement tiles), then the chunk's movers are
    // appended to the LDS list in agent order with one LDS atomic
    static_assert(ITEMS % 4 == 0, "four waves take whole chunks");
    constexpr uint32_t PER_WAVE = ITEMS / 4;
    typedef unsigned int v2u __attribute__((ext_vector_type(2)));
    const uint32_t wave = tid >> 6;
    v2u w[PER_WAVE];
#pragma unroll
    for (uint32_t c = 0; c < PER_WAVE; ++c) {
        const uint32_t base = (blk + (wave + 4u * c) * nBlk) * MOVE_BLOCK;
        w[c] = __builtin_nontemporal_load(reinterpret_cast<const v2u*>(m.A.wake + base) + lane);
    }
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
    const uint32_t today = m.stepOfDay;
#pragma unroll
    for (uint32_t c = 0; c < PER_WAVE; ++c) {
        const uint32_t k = wave + 4u * c;
        const uint32_t first = (blk + k * nBlk) * MOVE_BLOCK + 4u * lane;
        bool mv[4];
#pragma unroll
        for (uint32_t j = 0; j < 4u; ++j) {
            const uint32_t wv = ((j < 2u ? w[c].x : w[c].y) >> (16u * (j & 1u))) & 0xFFFFu;
            mv[j] = first + j < n && (m.curfewStep || wv == today);
        }
        // the chunk's movers, and this lane's movers' position among them
        // (agents before its own)
        uint32_t total = 0, before = 0;
#pragma unroll
        for (uint32_t j = 0; j < 4u; ++j) {
            const unsigned long long bj = __ballot(mv[j]);
            total += (uint32_t)__popcll(bj);
            before += (uint32_t)__popcll(bj & below);
        }
        if (total != 0u) {
            uint32_t pos = 0;
            if (lane == 0) pos = atomicAdd(&nList, total);
            pos = __shfl(pos, 0, 64) + before;
#pragma unroll
            for (uint32_t j = 0; j < 4u; ++j)
                if (mv[j]) list[pos++] = (uint16_t)((k << 8) | (4u * lane + j));
        }
    }
    __syncthreads();
    const uint32_t cnt = nList;
    uint32_t moved = 0;
    for (uint32_t j = tid; j < cnt; j += MOVE_BLOCK) {
        const uint32_t e = list[j];
        moved += do_movement(m, (blk + (e >> 8) * nBlk) * MOVE_BLOCK + (e & 0xFFu)) ? 1u : 0u;
    }
    if (moved) atomicAdd(&nMoved, moved);
    __syncthreads();
    if (tid == 0 && cnt) {
        unsigned long long* c = m.counters + 2u * (blk % MOVE_CTR_SLOTS);
        atomicAdd(&c[0], (unsigned long long)cnt);
        if (nMoved) atomicAdd(&c[1], (unsigned long long)nMoved);
    }
}

__global__ void __launch_bounds__(MOVE_BLOCK) k_move_tile(MoveArgs m) { move_tile<MOVE_ITEMS>(m, blockIdx.x, gridDim.x); }

// End of the step's movement: the deferred quarantine writes; the highest
// writing agent index of each location wins.  The step's pending list count
// is pendCount[par]; the other parity's count is cleared for the next step.
// Block `blk` of `nBlk` blocks working on this population.
__device__ __forceinline__ void resolve_pending(const Locs& L, const StepScratch& S, uint32_t par, uint32_t blk,
                                                uint32_t nBlk) {
    if (blk == 0 && threadIdx.x == 0) S.pendCount[par ^ 1u] = 0u;
    uint32_t pc = S.pendCount[par];
    if (pc > S.pendCap) pc = S.pendCap;
    for (uint32_t j = blk * blockDim.x + threadIdx.x; j < pc; j += nBlk * blockDim.x) {
        uint32_t loc = S.pendList[j];
        unsigned long long key = atomicExch(&L.pendKey[loc], 0ull);
        if (key != 0ull) {
            L.rec[loc].quarUntil = (uint32_t)key;
            mark_special(L, loc);
        }
    }
}

__global__ void __launch_bounds__(256) k_resolve_pending(Locs L, StepScratch S, uint32_t par) {
    resolve_pending(L, S, par, blockIdx.x, gridDim.x);
}

constexpr uint32_t HAZARD_TYPES = 256;  // location-type multipliers, by type id
// J, dJ/dk, dJ/dm[type], then the four finite-difference probes: the change
// of J at k (1 + h) and k (1 - h), and at m_p (1 + h), m_p (1 - h) for one
// probe type p, along the same trajectory
constexpr uint32_t HAZARD_PROBE0 = 2 + HAZARD_TYPES;
constexpr uint32_t HAZARD_SLOTS = HAZARD_PROBE0 + 4;

struct InfectArgs {
    Agents A;
    Locs L;
    Tables T;
    StepScratch S;
    uint64_t seed;
    uint32_t replica;
    uint32_t timestamp;
    uint32_t mask;
    float k;
    float season;
    float dt;
    const float* typeMul;  // [HAZARD_TYPES] infectiousness multiplier of each location type (1 = reference)
    double* hazard;        // [HAZARD_SLOTS] accumulated expected infections and gradient, or null
    float probeH;          // relative step of the finite-difference probes (0: none)
    uint32_t probeType;    // location type of the multiplier probe
    uint32_t* dumpList;    // agents infected by this launch are appended here (infectious-list dump), or null
    uint32_t* dumpCount;
    uint32_t resolve;      // also apply the preceding movement step's quarantine writes
    uint32_t resolvePar;   // their pending-counter parity
    uint32_t nTiles;       // infection blocks of this population (for the resolve stride)
};

// Infection of all active variants.  LocationsList::infectAgents
// (locationList.h:251-305) with the ratio of infectionPolicies.h:523-546
// evaluated from the location's exact sums.  Each thread takes INFECT_ITEMS
// agents 64 apart and issues their independent loads together, so that a
// wave keeps several memory round trips in flight.
constexpr uint32_t INFECT_BLOCK = 256;
constexpr uint32_t INFECT_ITEMS = 8;  // agents per thread
constexpr uint32_t INFECT_TILE = INFECT_BLOCK * INFECT_ITEMS;
static_assert(INFECT_BLOCK % 64 == 0, "the infection pass maps 64-agent runs to waves");

// The expected-infection hazard of one susceptible agent at a location with
// infectious visitors, s * (1 - exp(-x)) with x = k * density * inf * m_type *
// season * dt, and its derivatives with respect to k and m_type, with the
// trajectory held fixed (FP32 terms, FP64 sums).  The terms every candidate
// adds to (J, dJ/dk and the four probe sums) are
// kept in registers, acc[0..5], and reduced once per wave at the end of the
// tile; the term of its location's type goes to the block's LDS slot.
constexpr uint32_t HAZARD_ACC = 6;
__device__ __forceinline__ void hazard_terms(double* acc, double* sh, float k, float density, float inf, float m,
                                             uint32_t type, float season, float dt, float x, float s, float probeH,
                                             uint32_t probeType) {
    // each term in FP32 (1 - e^-x through expm1: no cancellation for small
    // x), accumulated in FP64
    const float em = ps_expm1f(-x);  // e^-x - 1
    const float w = s * (1.0f + em);  // s e^-x
    acc[0] += (double)(s * -em);
    acc[1] += (double)(w * (density * inf * m * season * dt));
    atomicAdd(&sh[2 + type], (double)(w * (k * density * inf * season * dt)));
    if (probeH > 0.0f) {
        // x is linear in k and in the multiplier of its location's type: J at
        // a parameter scaled by (1 +- h) changes by s e^-x (1 - e^-+xh)
        const float xh = x * probeH;
        const float dp = -w * ps_expm1f(-xh);
        const float dm = -w * ps_expm1f(xh);
        acc[2] += (double)dp;
        acc[3] += (double)dm;
        if (type == probeType) {
            acc[4] += (double)dp;
            acc[5] += (double)dm;
        }
    }
}

// infm: the location's infectiousness times its type's multiplier (LocDyn).
template <bool HAZARD>
__device__ __forceinline__ void infect_agent(const InfectArgs& a, uint32_t i, uint32_t v, uint32_t loc, long long e,
                                             uint32_t count, float infm, float s, double* acc, double* sh) {
    const uint32_t n = a.A.n;
    float num = (float)((double)e * (1.0 / 4294967296.0));
    float density = num / (float)count;
    const float x = a.k * density * infm * a.season * a.dt;
    if (HAZARD) {
        const uint32_t type = a.L.type8[loc];
        const float m = a.typeMul[type];
        // infm = inf * m (the LocDyn value loaded with the sum): the raw
        // infectiousness without a second gather, unless m is 0
        const float inf = m != 0.0f ? infm / m : a.L.infRaw[loc];
        hazard_terms(acc, sh, a.k, density, inf, m, type, a.season, a.dt, x, s, a.probeH, a.probeType);
    }
    float ratio = 1.0f - ps_expf(-x);
    if (ratio == 0.0f) return;
    Stream rs(a.seed, a.A.id[i], a.timestamp, PH_INFECT + v, a.replica);
    if (rs.unit() < (double)(ratio * s)) {
        // gotInfected (dynamicPPState.cpp:235-244) + stats (locationList.h:286-291)
        uint8_t fi = (uint8_t)a.T.firstInfected;
        a.A.rec[i].state = fi;
        a.A.rec[i].wb = a.T.stWB[fi];
        a.A.rec[i].variant = (uint8_t)v;
        a.A.days[i] = -2;
        float sf = a.T.stSusc[fi] ? 1.0f : 0.0f;
        for (uint32_t u = 0; u < a.T.nVar; ++u) a.A.susc[(size_t)u * a.A.suscStride + i] = sf;
        a.A.infTs[i] = a.timestamp;
        a.A.infCount[i] = (uint8_t)(a.A.infCount[i] + 1);
        a.A.infLoc[i] = loc;
        a.A.worstState[i] = (int8_t)fi;
        a.A.worstTs[i] = a.timestamp;
        a.A.statVar[i] = (uint8_t)(a.A.statVar[i] | (1u << v));
        if (a.dumpList) a.dumpList[atomicAdd(a.dumpCount, 1u)] = i;
    }
}

// HAZARD: accumulate the expected-infection hazard and its gradient
// (a separate instantiation, so that the plain pass carries no FP64 state).
template <bool HAZARD, uint32_t ITEMS = INFECT_ITEMS>
__device__ __forceinline__ void infect_tile(const InfectArgs& a, uint32_t tile) {
    constexpr uint32_t INFECT_ITEMS = ITEMS;  // agents per thread of this instance
    constexpr uint32_t INFECT_TILE = INFECT_BLOCK * ITEMS;
    if (a.mask == 0u) {
        if (a.resolve) resolve_pending(a.L, a.S, a.resolvePar, tile, a.nTiles);
        return;
    }
    // each wave takes 64 * INFECT_ITEMS consecutive agents, item k of lane
    // l being agent base + 64k: every load instruction covers 64 consecutive
    // agents, so household members (consecutive agents at one location, and
    // neighbouring households' homes) share cache lines of the bitmap below
    // (the location and susceptibility arrays are padded to whole tiles)
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t base = tile * INFECT_TILE + (threadIdx.x >> 6) * (64u * INFECT_ITEMS) + lane;
    uint32_t loc[INFECT_ITEMS];
#pragma unroll
    for (uint32_t k = 0; k < INFECT_ITEMS; ++k) loc[k] = __builtin_nontemporal_load(&a.A.location[base + 64u * k]);
    // the visited bit of each agent's location (an L2-resident bitmap): an
    // agent at a location without an infectious visitor today is done, for
    // every variant, without reading its susceptibility
    uint32_t cand = 0;
#pragma unroll
    for (uint32_t k = 0; k < INFECT_ITEMS; ++k)
        if ((a.L.visited[loc[k] >> 5] >> (loc[k] & 31u)) & 1u) cand |= 1u << k;
    // the movement step's deferred location writes: the infection pass never
    // reads quarantine state, so they share its launch, issued after the
    // loads above so that the pending count's round trip overlaps them
    if (a.resolve) resolve_pending(a.L, a.S, a.resolvePar, tile, a.nTiles);
    __shared__ double shHazard[HAZARD ? HAZARD_SLOTS : 1];
    double* sh = shHazard;
    double acc[HAZARD ? HAZARD_ACC : 1];
#pragma unroll
    for (uint32_t j = 0; j < (HAZARD ? HAZARD_ACC : 1); ++j) acc[j] = 0.0;
    if (HAZARD) {
        for (uint32_t j = threadIdx.x; j < HAZARD_SLOTS; j += blockDim.x) shHazard[j] = 0.0;
        __syncthreads();
    }
    for (uint32_t v = 0; cand != 0u && v < a.T.nVar; ++v) {
        if (!((a.mask >> v) & 1u)) continue;
        // the location's sum of this variant, then the susceptibility, which
        // is re-read per variant: an infection by an earlier variant in this
        // launch has already rewritten it
        const long long* sv = v == 0 ? &a.L.dyn[0].sum0 : a.L.sumX + (size_t)(v - 1) * a.L.n;
        const uint32_t pitch = v == 0 ? (uint32_t)(sizeof(LocDyn) / sizeof(long long)) : 1u;
        const float* su = a.A.susc + (size_t)v * a.A.suscStride + base;
        // one round of loads for all items, issued before any result is used
        // (a wave waits for one memory round trip, not two per item): the
        // sums at the candidates' locations and the candidates'
        // susceptibilities
        long long e[INFECT_ITEMS];
        float s[INFECT_ITEMS];
#pragma unroll
        for (uint32_t k = 0; k < INFECT_ITEMS; ++k) {
            e[k] = 0;
            s[k] = 0.0f;
            if ((cand >> k) & 1u) {
                e[k] = sv[(size_t)loc[k] * pitch];
                s[k] = su[64u * k];
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < INFECT_ITEMS; ++k)
            if (e[k] == 0) s[k] = 0.0f;
        uint32_t hits = 0;
#pragma unroll
        for (uint32_t k = 0; k < INFECT_ITEMS; ++k)
            if (s[k] > 0.0f) hits |= 1u << k;
        // the rare agents at a location with infectious visitors: one at a time
        while (hits) {
            const uint32_t k = (uint32_t)__builtin_ctz(hits);
            hits &= hits - 1u;
            const uint32_t i = base + 64u * k;
            const uint32_t l = loc[k];
            const LocDyn d = a.L.dyn[l];
            infect_agent<HAZARD>(a, i, v, l, e[k], d.count, d.inf, s[k], acc, sh);
        }
    }
    if (HAZARD) {
        // the register sums: one LDS add per wave
        const uint32_t slot[HAZARD_ACC] = {0u, 1u, HAZARD_PROBE0, HAZARD_PROBE0 + 1u, HAZARD_PROBE0 + 2u,
                                           HAZARD_PROBE0 + 3u};
#pragma unroll
        for (uint32_t j = 0; j < (HAZARD ? HAZARD_ACC : 1); ++j) {
            double v = acc[j];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if ((threadIdx.x & 63u) == 0u && v != 0.0) atomicAdd(&sh[slot[j]], v);
        }
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < HAZARD_SLOTS; j += blockDim.x)
            if (shHazard[j] != 0.0) atomicAdd(&a.hazard[j], shHazard[j]);
    }
}

// amdgpu_waves_per_eu(8): the plain pass is bound by the latency of its
// dependent loads, so it asks for the most resident waves (the compiler then
// keeps it within 96 SGPRs, moving rarely used pointers into VGPR lanes);
// the hazard pass, with its FP64 terms, is left to the compiler's choice
template <bool HAZARD>
__global__ void __launch_bounds__(INFECT_BLOCK) __attribute__((amdgpu_waves_per_eu(HAZARD ? 1 : 8)))
k_infect(InfectArgs a) {
    infect_tile<HAZARD>(a, blockIdx.x);
}

// ------------------------------------------------------------ DummyMovement
// DummyMovement::movement (movementPolicies.h:44-64): an agent whose
// countdown is zero moves to a uniformly drawn location (any of the L,
// cemetery included) and draws a new countdown below 36; every countdown
// then decrements, wrapping from 0 to 2^32 - 1 (the reference's unsigned
// underflow: an agent that draws 0 never moves again).  Occupancy and sums
// are kept up to date as in the real movement.
__global__ void k_dummy_move(Agents A, Locs L, Tables T, uint32_t* wait, uint64_t seed, uint32_t replica, uint32_t ts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    uint32_t w = wait[i];
    if (w == 0u) {
        Stream rs(seed, A.id[i], ts, PH_DUMMY_MOVE, replica);
        const uint32_t loc = rs.uint_below(L.n);
        w = rs.uint_below(144u / 4u);
        const uint32_t old = A.location[i];
        if (loc != old) {
            A.location[i] = loc;
            const AgentRec r = A.rec[i];
            const float x = T.stInf[r.state] * T.varMul[r.variant];
            if (x != 0.0f) {
                const long long q = fixed32(x);
                atomicAdd((unsigned long long*)loc_sum(L, r.variant, old), (unsigned long long)(-q));
                atomicAdd((unsigned long long*)loc_sum(L, r.variant, loc), (unsigned long long)q);
                mark_visited(L, loc);
            }
            atomicSub(&L.dyn[old].count, 1u);
            atomicAdd(&L.dyn[loc].count, 1u);
        }
    }
    wait[i] = w - 1u;
}

// ------------------------------------------------------------ replica batches
// Many populations (replicas) stepped by one launch per kernel: every block
// finds its replica in a per-block table and runs the same tile code on that
// replica's arrays.  The per-day parameters of every replica live in device
// arrays; the step's timestamp is a kernel argument, shared by all replicas.
struct BatchArgs {
    const MoveArgs* move;       // [replicas]
    const InfectArgs* infect;   // [replicas]
    const uint32_t* moveRep;    // [movement blocks]: replica of each block
    const uint32_t* moveFirst;  // [replicas]: its first movement block
    const uint32_t* infRep;     // [infection blocks]
    const uint32_t* infFirst;   // [replicas]
    const uint32_t* dayStart;   // device word: timestamp of the day's first step (set with the day's arguments)
    uint32_t stepOfDay;
    uint32_t pendPar;
    uint32_t nRep;        // replicas
    uint32_t moveBlocks;  // movement blocks of all replicas
    uint32_t infBlocks;   // infection blocks of all replicas
};

template <uint32_t ITEMS>
__global__ void __launch_bounds__(MOVE_BLOCK) k_move_batch(BatchArgs b) {
    const uint32_t r = b.moveRep[blockIdx.x];
    MoveArgs m = b.move[r];
    move_set_step(m, *b.dayStart + b.stepOfDay, b.stepOfDay);
    m.pendPar = b.pendPar;
    // this replica's blocks at this tile size
    const uint32_t nBlk = (r + 1 < b.nRep ? b.moveFirst[r + 1] : b.moveBlocks) - b.moveFirst[r];
    move_tile<ITEMS>(m, blockIdx.x - b.moveFirst[r], nBlk);
}

template <bool HAZARD, uint32_t ITEMS>
__global__ void __launch_bounds__(INFECT_BLOCK) __attribute__((amdgpu_waves_per_eu(HAZARD ? 1 : 8)))
k_infect_batch(BatchArgs b) {
    const uint32_t r = b.infRep[blockIdx.x];
    InfectArgs a = b.infect[r];
    a.timestamp = *b.dayStart + b.stepOfDay;
    a.resolve = 1u;  // every batch step is a movement step followed by this launch
    a.resolvePar = b.pendPar;
    // this replica's blocks at this tile size (the resolve stride)
    a.nTiles = (r + 1 < b.nRep ? b.infFirst[r + 1] : b.infBlocks) - b.infFirst[r];
    // a replica without hazard accumulation in a batch that has some: its
    // blocks take the plain path
    if (HAZARD && a.hazard) infect_tile<true, ITEMS>(a, blockIdx.x - b.infFirst[r]);
    else infect_tile<false, ITEMS>(a, blockIdx.x - b.infFirst[r]);
}

}  // namespace dev
}  // namespace ps
