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

// ------------------------------------------------ findActualLocationForType
// movementPolicies.h:69-122; `rs` may be null when the type needs no draw.
__device__ uint32_t find_loc(const Agents& A, const Locs& L, const pansim_location_types& lt, uint32_t a, uint32_t ty,
                             int retry, Stream* rs) {
    uint32_t b = A.locOffset[a], e = A.locOffset[a + 1];
    if (ty == lt.home || ty == lt.school || ty == lt.work) {
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
                        if (retry && L.open[A.possLoc[i]] == 0) {
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

struct MoveArgs {
    Agents A;
    Locs L;
    Tables T;
    StepScratch S;
    pansim_step_params p;
    uint64_t seed;
    uint32_t replica;
    uint32_t curfewStep;  // all agents take the full path this step
    uint32_t nEvents;
    uint32_t pendPar;  // parity of the pending-write counter used this step
    unsigned long long* counters;  // MOVE_CTR_SLOTS pairs (evaluated, relocated)
};

constexpr uint32_t MOVE_BLOCK = 256;
constexpr uint32_t MOVE_ITEMS = 16;
constexpr uint32_t MOVE_TILE = MOVE_BLOCK * MOVE_ITEMS;  // agents per block
constexpr uint32_t MOVE_CTR_SLOTS = 64;

__device__ __forceinline__ void pend_write(const MoveArgs& m, uint32_t agent, uint32_t loc, uint32_t until) {
    unsigned long long key = ((unsigned long long)agent << 32) | (unsigned long long)until;
    unsigned long long old = atomicMax(&m.L.pendKey[loc], key);
    if (old == 0ull) {
        uint32_t slot = atomicAdd(&m.S.pendCount[m.pendPar], 1u);
        if (slot < m.S.pendCap) m.S.pendList[slot] = loc;
    }
}

// movementPolicies.h:171-245 with the location writes deferred to the end of the step
__device__ void quarantine_agent(const MoveArgs& m, uint32_t i, uint32_t until, uint8_t& fl, Stream& rs) {
    const Agents& A = m.A;
    const Locs& L = m.L;
    const pansim_location_types& lt = m.T.lt;
    uint32_t ts = m.p.timestamp;
    if (m.p.quarantine_policy == 0) return;
    bool infected = m.T.stInfected[A.state[i]] != 0;
    if (A.diagTs[i] > 0 && !infected && m.p.quarantine_immune_active == 0) return;
    if (!(fl & F_DIAGNOSED) && A.immTs[i] > 0 && m.p.quarantine_immune_active == 0) return;
    fl |= F_QUARANTINED;
    A.quarTs[i] = ts;
    uint32_t prev = A.quarUntil[i];
    A.quarUntil[i] = until;
    A.daysInQ[i] += (until - ts) / (24u * 60u / m.p.time_step);
    if (fl & F_DIAGNOSED) {
        uint32_t home = find_loc(A, L, lt, i, lt.home, 0, nullptr);
        if (m.p.quarantine_policy > 1 && home != NOLOC) pend_write(m, i, home, until);
        if (m.p.quarantine_policy > 2 && prev < ts) {
            uint32_t cls = find_loc(A, L, lt, i, lt.classroom, 0, &rs);
            uint32_t work = find_loc(A, L, lt, i, lt.work, 0, nullptr);
            uint32_t toClose[2] = {cls, work};
            for (int k = 0; k < 2; ++k) {
                uint32_t loc = toClose[k];
                if (loc != NOLOC && (L.type[loc] == lt.work || L.type[loc] == lt.classroom) && L.quarUntil[loc] < ts &&
                    L.open[loc] && L.closedUntil[loc] < ts)
                    pend_write(m, i, loc, until);
            }
        }
    }
}

// RealMovementOps::doMovement (movementPolicies.h:258-1041) for agent i.
// Writes location, wake step, flags and stats of agent i; keeps the
// per-location occupancy and infectiousness sums up to date.  Returns
// whether the agent changed location.
__device__ bool do_movement(const MoveArgs& m, uint32_t i) {
    const Agents& A = m.A;
    const Locs& L = m.L;
    const Tables& T = m.T;
    const pansim_location_types& lt = T.lt;
    const uint32_t ts = m.p.timestamp;
    const uint32_t dt = m.p.time_step;
    const unsigned minutes = (ts % (1440u / dt)) * dt;
    const TD simTime = td_from_min(minutes);
    const uint32_t stepsUntilMidnight = (1440u - minutes) / dt;
    const uint32_t qLen = m.p.quarantine_length * 24u * 60u / dt;
    const uint32_t agentType = A.type[i];
    const uint32_t oldLoc = A.location[i];
    uint32_t loc = oldLoc;
    uint8_t fl = A.flags[i];
    const uint8_t fl0 = fl;
    uint32_t wake = A.wake[i];
    Stream rs(m.seed, i, ts, PH_MOVE, m.replica);
    const uint8_t st = A.state[i];
    const uint8_t w = T.stWB[st];

    bool done = false;
    // curfew (movementPolicies.h:267-326)
    if (m.p.enable_curfew && m.p.curfew_begin == minutes / dt) {
        bool deadOrHosp = (w == 4 || w == 3);
        bool hospNonCovid = A.hospTs[i] <= ts && A.hospUntil[i] > ts;
        if (!deadOrHosp && !hospNonCovid) {
            if (agentType + 1 == 7) {
                uint32_t workplace = find_loc(A, L, lt, i, lt.work, 0, nullptr);
                if (loc != workplace || L.open[workplace] == 0) {
                    loc = find_loc(A, L, lt, i, lt.home, 0, nullptr);
                    if (L.open[workplace] == 0) {
                        wake = wake_keep(ts, stepsUntilMidnight);
                        done = true;
                    }
                }
            } else {
                wake = wake_keep(ts, stepsUntilMidnight);
                loc = find_loc(A, L, lt, i, lt.home, 0, nullptr);
                done = true;
            }
        }
    }
    if (!done && wake != ts) done = true;  // countdown > 0: nothing else happens this step

    if (!done) do {
        if (A.quarUntil[i] <= ts) fl &= (uint8_t)~F_QUARANTINED;
        if (w == 4) {  // dead
            wake = NEVER;
            loc = L.cemetery;
            break;
        }
        const uint32_t agentHome = find_loc(A, L, lt, i, lt.home, 0, nullptr);
        uint32_t sum = 0;
        auto diagnose = [&]() {
            if (!(fl & F_DIAGNOSED) && T.stInf[st] * T.varMul[A.variant[i]] > 0.0f) {
                fl |= F_DIAGNOSED;
                uint32_t d = ts;
                if (stepsUntilMidnight == 24u * 60u / dt) d++;
                A.diagTs[i] = d;
                quarantine_agent(m, i, ts + qLen, fl, rs);
            }
        };
        if (A.hospTs[i] <= ts && A.hospUntil[i] > ts && w != 3 && w != 4) {
            uint32_t v = A.hospUntil[i] - ts - 1;
            sum = v < stepsUntilMidnight ? v : stepsUntilMidnight;
            loc = find_loc(A, L, lt, i, lt.hospital, 0, &rs);
            diagnose();
            wake = wake_keep(ts, sum);
            break;
        }
        if (w == 3) {
            sum = stepsUntilMidnight;
            loc = find_loc(A, L, lt, i, lt.hospital, 0, &rs);
            diagnose();
            wake = wake_keep(ts, sum);
            break;
        }
        auto quarantinable = [&](uint32_t l) {
            uint32_t t = L.type[l];
            return t == lt.home || t == lt.school || t == lt.classroom || t == lt.work;
        };
        auto stayedHomeCheck = [&]() {
            if (agentHome != NOLOC && loc != agentHome && sum > 60u / dt) fl &= (uint8_t)~F_STAYED_HOME;
        };
        if (m.p.quarantine_policy > 1 && ts < L.quarUntil[loc] && quarantinable(loc)) {
            if (!(fl & F_QUARANTINED)) quarantine_agent(m, i, L.quarUntil[loc], fl, rs);
            if (fl & F_QUARANTINED) {
                uint32_t v = L.quarUntil[loc] - ts - 1;
                sum = v < stepsUntilMidnight ? v : stepsUntilMidnight;
                uint32_t homeLocation = find_loc(A, L, lt, i, lt.home, 0, nullptr);
                if (homeLocation != loc) loc = homeLocation;
                stayedHomeCheck();
                wake = wake_keep(ts, sum);
                break;
            }
        }
        {
            bool qp = m.p.quarantine_policy > 0;
            uint32_t dts = A.diagTs[i];
            if ((qp && ((fl & F_DIAGNOSED) || (dts > 0 && (ts - dts) < qLen))) || (qp && (fl & F_QUARANTINED))) {
                if (qp && !(fl & F_QUARANTINED)) quarantine_agent(m, i, ts + qLen, fl, rs);
                if (fl & F_QUARANTINED) {
                    loc = find_loc(A, L, lt, i, lt.home, 0, nullptr);
                    sum = stepsUntilMidnight;
                    stayedHomeCheck();
                    wake = wake_keep(ts, sum);
                    break;
                }
            }
        }
        // schedule window (agentType.cpp:13-15)
        uint32_t off = 7u * w + m.p.day_of_week + agentType * 28u;
        uint32_t eb = T.evOff[off], ee = T.evOff[off + 1];
        int ab = -1, ae = -1;
        for (uint32_t j = eb; j < ee; ++j) {
            const pansim_event& e = T.ev[j];
            bool in = !td_lt(simTime, ev_start(e)) && td_lt(simTime, ev_end(e));
            if (in && ab == -1) ab = (int)j;
            if (td_lt(simTime, ev_start(e))) {
                ae = (int)j;
                break;
            }
            if (j == ee - 1 && in) ae = (int)ee;
        }
        uint32_t newType = NOLOC;
        if (ab == -1 && ae == -1) {  // Case 1
            newType = (w == 3) ? lt.hospital : lt.home;
            loc = find_loc(A, L, lt, i, newType, 0, &rs);
            sum = stepsUntilMidnight;
        }
        if (ab != -1) {  // Cases 2 and 4
            uint32_t nPot = (uint32_t)((ae == -1 ? ab + 1 : ae) - ab);
            uint32_t picked = 0;
            TD dur;
            if (nPot == 1) {
                newType = T.ev[ab].loc_type;
                dur = ev_dur(T.ev[ab]);
            } else {
                double r = rs.real(1.0);
                double thr = T.ev[ab].chance;
                uint32_t idx = 0;
                while (r > thr && idx < nPot) {
                    idx++;
                    if (idx < nPot) thr += T.ev[ab + idx].chance;
                }
                if (idx == nPot) idx--;
                picked = idx;
                newType = T.ev[ab + idx].loc_type;
                dur = ev_dur(T.ev[ab + idx]);
            }
            if ((fl & F_NO_WORK) && newType == lt.work) {
                newType = lt.home;
                nPot = 1;
                if (!(fl & F_QUARANTINED)) A.daysInQ[i] += 1;
            }
            uint32_t newLoc = find_loc(A, L, lt, i, newType, nPot == 1, &rs);
            bool tooOld = (newType == lt.school || newType == lt.classroom) && A.age[i] >= m.p.school_age_restriction;
            bool lockdown = m.p.lockdown_nonvacc_active && A.immTs[i] == 0;
            auto closedHere = [&](uint32_t l, uint32_t t) {
                return (L.open[l] == 0 || L.closedUntil[l] > ts) && t != lt.work;
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
                    newLoc = find_loc(A, L, lt, i, newType, 0, &rs);
                    if (tooOld || closedHere(newLoc, newType) || lockdown) newLoc = find_loc(A, L, lt, i, lt.home, 0, nullptr);
                } else {
                    newLoc = find_loc(A, L, lt, i, lt.home, 0, nullptr);
                }
            } else if (newType == lt.school) {
                uint32_t cls = find_loc(A, L, lt, i, lt.classroom, 0, &rs);
                if (cls != NOLOC && L.closedUntil[cls] > ts) newLoc = find_loc(A, L, lt, i, lt.home, 0, nullptr);
            }
            loc = newLoc;
            if (dur.h > 24) {
                sum = stepsUntilMidnight;
            } else if (ae == -1) {
                TD arrive = td_plus(simTime, dur);
                sum = arrive.h >= 24 ? stepsUntilMidnight : td_min(dur) / dt;
            } else {
                // ae can equal the slot end: the reference reads the next slot's first
                // event, or a default event past the end of the array
                pansim_event ne;
                if ((uint32_t)ae < m.nEvents) ne = T.ev[ae];
                else {
                    ne.loc_type = 0;
                    ne.chance = -1.0f;
                    ne.start_h = ne.start_m = ne.end_h = ne.end_m = ne.dur_h = ne.dur_m = 0;
                }
                TD arrive = td_plus(simTime, dur);
                if (td_lt(arrive, ev_start(ne))) sum = td_min(dur) / dt;
                else if (td_lt(ev_end(ne), arrive)) sum = td_min(td_minus(ev_end(ne), simTime)) / dt - 1u;
                else {
                    uint32_t stp = td_min(td_minus(ev_end(ne), arrive)) / dt;
                    uint32_t extra = rs.uint_below(stp);
                    sum = td_min(dur) / dt + extra;
                }
            }
            if (agentType + 1 == 7 && m.p.enable_curfew &&
                (m.p.curfew_begin <= minutes / dt || m.p.curfew_end > minutes / dt)) {
                bool workButClosed = newType == lt.work && (L.open[newLoc] == 0 || L.closedUntil[newLoc] > ts);
                if (workButClosed || (newType != lt.work && newType != lt.home)) loc = agentHome;
            }
        }
        if (ab == -1 && ae != -1) {  // Case 3
            const pansim_event& ne = T.ev[ae];
            uint32_t lenSteps = td_min(td_minus(ev_end(ne), ev_start(ne))) / dt;
            uint32_t delay = rs.uint_below(lenSteps);
            sum = td_min(td_minus(ev_start(ne), simTime)) / dt + delay;
            if (sum < 30u / dt) {
            } else if (sum < 60u / dt) {
                newType = lt.public_space;
                loc = find_loc(A, L, lt, i, lt.public_space, 0, &rs);
            } else {
                newType = lt.home;
                loc = find_loc(A, L, lt, i, lt.home, 0, nullptr);
            }
        }
        if (m.p.quarantine_policy > 1 && ts < L.quarUntil[loc] && quarantinable(loc)) {
            if (!(fl & F_QUARANTINED)) quarantine_agent(m, i, L.quarUntil[loc], fl, rs);
            if (fl & F_QUARANTINED) {
                uint32_t v = L.quarUntil[loc] - ts - 1;
                sum = v < stepsUntilMidnight ? v : stepsUntilMidnight;
                uint32_t homeLocation = find_loc(A, L, lt, i, lt.home, 0, nullptr);
                if (homeLocation != loc) loc = homeLocation;
                stayedHomeCheck();
                wake = wake_keep(ts, sum);
                break;
            }
        }
        if (newType == lt.hospital || (newType == lt.doctor && rs.real(1.0) < 0.7)) diagnose();
        stayedHomeCheck();
        wake = wake_dec(ts, sum);
    } while (false);

    if (fl != fl0) A.flags[i] = fl;
    if (wake != A.wake[i]) A.wake[i] = wake;
    if (loc == oldLoc) return false;
    A.location[i] = loc;
    atomicSub(&L.count[oldLoc], 1u);
    atomicAdd(&L.count[loc], 1u);
    float x = T.stInf[st] * T.varMul[A.variant[i]];
    if (x != 0.0f) {
        long long q = fixed32(x);
        long long* s = L.sum + (size_t)A.variant[i] * L.n;
        atomicAdd((unsigned long long*)&s[oldLoc], (unsigned long long)(-q));
        atomicAdd((unsigned long long*)&s[loc], (unsigned long long)q);
    }
    return true;
}

// One movement step over a tile of MOVE_TILE consecutive agents per block:
// the agents whose countdown reaches zero at this step (everyone on the
// curfew step) are compacted into LDS, then processed by the same block.
// The order in which movers are processed does not change any result: the
// per-agent writes are private, the per-location sums are integer atomics and
// the location quarantine writes are deferred to k_resolve_pending.
// Evaluated / relocated counts go to one of MOVE_CTR_SLOTS counter pairs.
__global__ void __launch_bounds__(MOVE_BLOCK) k_move_tile(MoveArgs m) {
    __shared__ uint32_t list[MOVE_TILE];
    __shared__ uint32_t nList, nMoved;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t base = blockIdx.x * MOVE_TILE;
    const uint32_t n = m.A.n;
    const uint32_t ts = m.p.timestamp;
    if (tid == 0) {
        nList = 0;
        nMoved = 0;
    }
    uint32_t w[MOVE_ITEMS];
#pragma unroll
    for (uint32_t k = 0; k < MOVE_ITEMS; ++k) {
        uint32_t i = base + k * MOVE_BLOCK + tid;
        w[k] = i < n ? __builtin_nontemporal_load(&m.A.wake[i]) : ts + 1u;
    }
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (uint32_t k = 0; k < MOVE_ITEMS; ++k) {
        uint32_t i = base + k * MOVE_BLOCK + tid;
        bool mover = i < n && (m.curfewStep || w[k] == ts);
        unsigned long long mask = __ballot(mover);
        if (mask != 0ull) {
            uint32_t pos = 0;
            if (lane == 0) pos = atomicAdd(&nList, (uint32_t)__popcll(mask));
            pos = __shfl(pos, 0, 64);
            if (mover) list[pos + (uint32_t)__popcll(mask & below)] = i;
        }
    }
    __syncthreads();
    const uint32_t cnt = nList;
    uint32_t moved = 0;
    for (uint32_t j = tid; j < cnt; j += MOVE_BLOCK) moved += do_movement(m, list[j]) ? 1u : 0u;
    if (moved) atomicAdd(&nMoved, moved);
    __syncthreads();
    if (tid == 0 && cnt) {
        unsigned long long* c = m.counters + 2u * (blockIdx.x % MOVE_CTR_SLOTS);
        atomicAdd(&c[0], (unsigned long long)cnt);
        if (nMoved) atomicAdd(&c[1], (unsigned long long)nMoved);
    }
}

// End of the step's movement: the deferred quarantine writes; the highest
// writing agent index of each location wins.  The step's pending list count
// is pendCount[par]; the other parity's count is cleared for the next step.
__global__ void __launch_bounds__(256) k_resolve_pending(Locs L, StepScratch S, uint32_t par) {
    if (blockIdx.x == 0 && threadIdx.x == 0) S.pendCount[par ^ 1u] = 0u;
    uint32_t pc = S.pendCount[par];
    if (pc > S.pendCap) pc = S.pendCap;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < pc; j += gridDim.x * blockDim.x) {
        uint32_t loc = S.pendList[j];
        unsigned long long key = atomicExch(&L.pendKey[loc], 0ull);
        if (key != 0ull) L.quarUntil[loc] = (uint32_t)key;
    }
}

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
};

// Infection of all active variants.  LocationsList::infectAgents
// (locationList.h:251-305) with the ratio of infectionPolicies.h:523-546
// evaluated from the location's exact sums.  Each thread takes INFECT_ITEMS
// agents (stride INFECT_BLOCK) and issues their independent loads together,
// so that a wave keeps several memory round trips in flight.
constexpr uint32_t INFECT_BLOCK = 256;
constexpr uint32_t INFECT_ITEMS = 4;
constexpr uint32_t INFECT_TILE = INFECT_BLOCK * INFECT_ITEMS;

__device__ __forceinline__ void infect_agent(const InfectArgs& a, uint32_t i, uint32_t v, uint32_t loc, long long e,
                                             float s) {
    const uint32_t n = a.A.n;
    float num = (float)((double)e * (1.0 / 4294967296.0));
    float density = num / (float)a.L.count[loc];
    float ratio = 1.0f - ps_expf(-a.k * density * a.L.inf[loc] * a.season * a.dt);
    if (ratio == 0.0f) return;
    Stream rs(a.seed, i, a.timestamp, PH_INFECT + v, a.replica);
    if (rs.unit() < (double)(ratio * s)) {
        // gotInfected (dynamicPPState.cpp:235-244) + stats (locationList.h:286-291)
        uint8_t fi = (uint8_t)a.T.firstInfected;
        a.A.state[i] = fi;
        a.A.variant[i] = (uint8_t)v;
        a.A.days[i] = -2;
        float sf = a.T.stSusc[fi] ? 1.0f : 0.0f;
        for (uint32_t u = 0; u < a.T.nVar; ++u) a.A.susc[(size_t)u * n + i] = sf;
        a.A.infTs[i] = a.timestamp;
        a.A.infCount[i] = (uint8_t)(a.A.infCount[i] + 1);
        a.A.infLoc[i] = loc;
        a.A.worstState[i] = (int8_t)fi;
        a.A.worstTs[i] = a.timestamp;
        a.A.statVar[i] = (uint8_t)(a.A.statVar[i] | (1u << v));
    }
}

__global__ void __launch_bounds__(INFECT_BLOCK) k_infect(InfectArgs a) {
    if (a.mask == 0u) return;
    const uint32_t n = a.A.n;
    const uint32_t base = blockIdx.x * INFECT_TILE + threadIdx.x;
    uint32_t loc[INFECT_ITEMS];
#pragma unroll
    for (uint32_t k = 0; k < INFECT_ITEMS; ++k) {
        uint32_t i = base + k * INFECT_BLOCK;
        loc[k] = i < n ? __builtin_nontemporal_load(&a.A.location[i]) : NOLOC;
    }
    for (uint32_t v = 0; v < a.T.nVar; ++v) {
        if (!((a.mask >> v) & 1u)) continue;
        // susceptibility is re-read per variant: an infection by an earlier
        // variant in this launch has already rewritten it
        float s[INFECT_ITEMS];
#pragma unroll
        for (uint32_t k = 0; k < INFECT_ITEMS; ++k) {
            uint32_t i = base + k * INFECT_BLOCK;
            s[k] = i < n ? a.A.susc[(size_t)v * n + i] : 0.0f;
        }
        long long e[INFECT_ITEMS];
#pragma unroll
        for (uint32_t k = 0; k < INFECT_ITEMS; ++k) e[k] = s[k] > 0.0f ? a.L.sum[(size_t)v * a.L.n + loc[k]] : 0ll;
#pragma unroll
        for (uint32_t k = 0; k < INFECT_ITEMS; ++k)
            if (e[k] != 0) infect_agent(a, base + k * INFECT_BLOCK, v, loc[k], e[k], s[k]);
    }
}

}  // namespace dev
}  // namespace ps
