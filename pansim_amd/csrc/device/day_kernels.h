// Midnight kernels: initial state draws, occupancy rebuild, testing,
// progression, immunity waning, non-COVID hospitalisation, statistics,
// location planning, closure toggles.  Included once, by engine.hip.
#pragma once
#include "step_kernels.h"

namespace ps {
namespace dev {

// ------------------------------------------------------------ common helpers
__device__ __forceinline__ int geometric(double p, Stream& rs) {
    double log1mp = ps_log(1.0 - p);
    const double naf = (1.0 - 2.220446049250313e-16) / 2.0;
    const double thr = 2147483647.0 + naf;
    double cand;
    int guard = 0;
    do {
        cand = floor(ps_log(1.0 - rs.unit()) / log1mp);
    } while (cand >= thr && ++guard < 1000);
    if (cand >= thr) cand = 2147483647.0;
    return (int)(cand + naf);
}

// BasicLengthAbstract::LengthOfState::calculateDays (basicLengthAbstract.cpp:48-53)
__device__ __forceinline__ int calc_days(const Tables& T, uint32_t prog, uint32_t state, uint32_t variant, Stream& rs) {
    size_t k = ((size_t)prog * T.nStates + state) * 7 + variant;
    if (T.lenAvg[k] == -1.0f) return -1;
    double p = T.lenP[k];
    float mx = T.lenMax[k];
    int days = geometric(p, rs);
    int guard = 0;
    while (mx < (float)days && ++guard < 1000) days = geometric(p, rs);
    if (mx < (float)days) days = (int)mx;
    return days;
}

__device__ __forceinline__ float scaling_symptoms(const Agents& A, uint32_t i, uint32_t v, uint32_t s) {
    // scalingSymptoms[v*7 + MAX(6, s)] (agentMeta.cpp:128-129): only slot 6 of
    // each variant is ever written; other slots hold scalingAgeSex
    return s <= 6 ? A.slot6[(size_t)v * A.n + i] : A.scal[i];
}

__device__ __forceinline__ void update_meta(const Agents& A, const Tables& T, uint32_t i, uint8_t st) {
    float sf = T.stSusc[st] ? 1.0f : 0.0f;
    for (uint32_t v = 0; v < T.nVar; ++v) A.susc[(size_t)v * A.suscStride + i] = sf;
}

// ------------------------------------------------------------ init
__global__ void k_init_agents(Agents A, Tables T, uint64_t seed, uint32_t replica, uint32_t qp, uint32_t qlenSteps,
                              uint32_t qlenDays) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    uint8_t st = A.rec[i].state;
    A.rec[i].variant = 0;
    update_meta(A, T, i, st);
    for (uint32_t v = 0; v < T.nVar; ++v) A.slot6[(size_t)v * A.n + i] = A.scal[i];
    Stream rs(seed, A.id[i], 0, PH_INIT_DAYS, replica);
    A.days[i] = (int16_t)calc_days(T, A.prog[i], st, 0, rs);
    bool inf = T.stInfected[st] != 0;
    A.infTs[i] = inf ? 0u : NEVER;
    A.infCount[i] = inf ? 1 : 0;
    A.worstState[i] = inf ? (int8_t)st : 0;
    A.worstTs[i] = 0;
    A.worstEndTs[i] = 0;
    A.infLoc[i] = 0;
    A.recTs[i] = NEVER;
    A.rec[i].hospTs = 0;
    A.rec[i].hospUntil = 0;
    A.immTs[i] = 0;
    A.immCount[i] = 0;
    A.statVar[i] = 0;
    A.lastTest[i] = NEVER;
    A.wake[i] = 0;
    uint8_t fl = A.rec[i].flags;  // diagnosed / hcw preset by the host upload
    bool diag = fl & F_DIAGNOSED;
    A.rec[i].diagTs = diag ? 1u : 0u;
    if (diag && qp > 0) {
        A.quarTs[i] = 0;
        A.rec[i].quarUntil = qlenSteps;
        A.daysInQ[i] = qlenDays;
        fl |= F_QUARANTINED;
    } else {
        A.quarTs[i] = 0;
        A.rec[i].quarUntil = 0;
        A.daysInQ[i] = 0;
    }
    fl |= F_STAYED_HOME;
    A.rec[i].flags = fl;
}

// ------------------------------------------------------------ occupancy
__global__ void k_zero_locations(Locs L, uint32_t nVar) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < L.n; l += stride) {
        L.dyn[l].count = 0u;
        L.dyn[l].sum0 = 0;
        for (uint32_t v = 1; v < nVar; ++v) L.sumX[(size_t)(v - 1) * L.n + l] = 0;
        if ((l & 31u) == 0) L.visited[l >> 5] = 0u;
    }
}

// Runs of equal values among the consecutive lanes of a wave: `head` is set
// on the first lane of each run, the return value is the run's length there.
// Valid lanes form a prefix of the wave.
__device__ __forceinline__ uint32_t lane_run(uint32_t v, bool valid, bool& head) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t prev = __shfl_up(v, 1, 64);
    head = valid && (lane == 0 || prev != v);
    const unsigned long long heads = __ballot(head);
    const uint32_t nValid = (uint32_t)__popcll(__ballot(valid));
    const unsigned long long after = heads & ~((2ull << lane) - 1ull);
    uint32_t next = after ? (uint32_t)(__ffsll((long long)after) - 1) : 64u;
    if (next > nValid) next = nValid;
    return next - lane;
}

__global__ void k_scatter_locations(Agents A, Locs L, Tables T) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < A.n;
    const uint32_t loc = valid ? A.location[i] : NOLOC;
    // members of a household are consecutive agents: one occupancy atomic per run
    bool head;
    const uint32_t run = lane_run(loc, valid, head);
    if (!valid) return;
    if (head) atomicAdd(&L.dyn[loc].count, run);
    const AgentRec r = A.rec[i];
    float x = T.stInf[r.state] * T.varMul[r.variant];
    if (x != 0.0f) {
        atomicAdd((unsigned long long*)loc_sum(L, r.variant, loc), (unsigned long long)fixed32(x));
        mark_visited(L, loc);
    }
}

// The sums-only rebuild of the midnight planning: no agent changes location
// between the day's last movement step and the planning, so the occupancy
// counts stand; the infectiousness sums (which the midnight passes change
// through state and variant) are cleared with the visited bits, then
// refilled by the infectious agents alone.
__global__ void k_zero_sums(Locs L, uint32_t nVar) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < L.n; l += stride) {
        L.dyn[l].sum0 = 0;
        for (uint32_t v = 1; v < nVar; ++v) L.sumX[(size_t)(v - 1) * L.n + l] = 0;
        if ((l & 31u) == 0) L.visited[l >> 5] = 0u;
    }
}

__global__ void k_scatter_sums(Agents A, Locs L, Tables T) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const AgentRec r = A.rec[i];
    const float x = T.stInf[r.state] * T.varMul[r.variant];
    if (x == 0.0f) return;
    const uint32_t loc = A.location[i];
    atomicAdd((unsigned long long*)loc_sum(L, r.variant, loc), (unsigned long long)fixed32(x));
    mark_visited(L, loc);
}

// ------------------------------------------------------------ infectious-list dump
// (BasicInfection::dumpLocationInfectiousList, infectionPolicies.h:262-441)
// The locations of the agents infected since the last collection: flagged
// in a bitmap, and recorded next to each agent.
__global__ void k_dump_mark(Agents A, const uint32_t* agents, const uint32_t* count, uint32_t* locOut, uint32_t* flags) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= *count) return;
    const uint32_t l = A.location[agents[j]];
    locOut[j] = l;
    atomicOr(&flags[l >> 5], 1u << (l & 31u));
}

// Every agent with non-zero infectiousness at a flagged location.
__global__ void k_dump_infectious(Agents A, Tables T, const uint32_t* flags, uint32_t* outAgent, uint32_t* outLoc,
                                  float* outValue, uint32_t* count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const uint32_t l = A.location[i];
    if (!((flags[l >> 5] >> (l & 31u)) & 1u)) return;
    const AgentRec r = A.rec[i];
    const float x = T.stInf[r.state] * T.varMul[r.variant];
    if (x == 0.0f) return;
    const uint32_t slot = atomicAdd(count, 1u);
    outAgent[slot] = i;
    outLoc[slot] = l;
    outValue[slot] = x;
}

__global__ void k_count_variants(Agents A, Tables T, unsigned long long* out) {
    __shared__ unsigned int sh[MAXV];
    if (threadIdx.x < MAXV) sh[threadIdx.x] = 0;
    __syncthreads();
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < A.n && T.stInfected[A.rec[i].state]) atomicAdd(&sh[A.rec[i].variant], 1u);
    __syncthreads();
    if (threadIdx.x < T.nVar && sh[threadIdx.x]) atomicAdd(&out[threadIdx.x], (unsigned long long)sh[threadIdx.x]);
}

// ------------------------------------------------------------ testing
struct TestArgs {
    Agents A;
    Locs L;
    Tables T;
    uint64_t seed;
    uint32_t replica, ts, dt, delay, qlen, usePCR;
    double pRandom, pHome, pWork, pSchool, pHospital, pNursery;
    unsigned long long* counters;  // tests, positive now, positive in the last day
};

// The testing flags live in a bitmap (1 bit per location, L2 resident): the
// testing pass reads the flags of every agent's home, workplace and school.
__device__ __forceinline__ bool test_flag(const Locs& L, uint32_t l) {
    return ((L.flagBits[l >> 5] >> (l & 31u)) & 1u) != 0u;
}
__device__ __forceinline__ void set_test_flag(const Locs& L, uint32_t l) {
    atomicOr(&L.flagBits[l >> 5], 1u << (l & 31u));
}

__global__ void k_clear_flags(Locs L) {
    uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t words = (L.n + 31u) / 32u;
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < words; w += stride) L.flagBits[w] = 0u;
}

// DetailedTestingOps::flagLocations (testingPolicies.h:59-148)
__global__ void k_flag_locations(TestArgs a) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.A.n) return;
    const uint32_t day = 24u * 60u / a.dt;
    if (!(a.A.rec[i].diagTs > a.ts - day)) return;
    const pansim_location_types& lt = a.T.lt;
    const uint32_t ql = a.qlen * 24u * 60u / a.dt;
    auto ok = [&](uint32_t l) {
        uint32_t q = a.L.rec[l].quarUntil;
        return q == 0 || (q != 0 && (q - ql >= a.ts - day || q < a.ts));
    };
    const CsrReg csr = load_csr(a.A, i, a.A.rec[i]);
    uint32_t home = csr.home;
    if (home != NOLOC) set_test_flag(a.L, home);
    uint32_t work = find_loc(a.A, a.L, lt, csr, lt.work, 0, nullptr);
    if (work != NOLOC && ok(work)) set_test_flag(a.L, work);
    uint32_t school = find_loc(a.A, a.L, lt, csr, lt.school, 0, nullptr);
    if (school != NOLOC && ok(school)) {
        set_test_flag(a.L, school);
        Stream rs(a.seed, a.A.id[i], a.ts, PH_TEST_FLAG, a.replica);
        uint32_t cls = find_loc(a.A, a.L, lt, csr, lt.classroom, 0, &rs);
        if (cls != NOLOC && ok(cls)) set_test_flag(a.L, cls);
    }
}

// DetailedTestingOps::doTesting (testingPolicies.h:157-281) + the counts of performTests.
// The agent's record is read once, and the location records of its home,
// workplace and school are loaded together, so that a thread waits for a
// few memory round trips instead of one per field.
__global__ void k_do_testing(TestArgs a) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ unsigned int sh[3];
    if (threadIdx.x < 3) sh[threadIdx.x] = 0;
    __syncthreads();
    if (i < a.A.n) {
        const Agents& A = a.A;
        const uint32_t day = 24u * 60u / a.dt;
        const uint32_t ts = a.ts;
        const pansim_location_types& lt = a.T.lt;
        const AgentRec r = A.rec[i];
        uint32_t lastTest = A.lastTest[i];
        uint32_t diagTs = r.diagTs;
        const uint8_t fl = r.flags;
        const bool skip = ts > a.delay * day && lastTest != NEVER && lastTest > ts - a.delay * day;
        const uint8_t st = r.state;
        if (!skip && !(r.wb == 4 || (fl & F_DIAGNOSED))) {  // wb: T.stWB[st], kept in the record
            Stream rs(a.seed, A.id[i], ts, PH_TESTING, a.replica);
            const CsrReg csr = load_csr(A, i, r);
            const uint32_t home = csr.home;
            const uint32_t work = find_loc(A, a.L, lt, csr, lt.work, 0, nullptr);
            const uint32_t school = find_loc(A, a.L, lt, csr, lt.school, 0, nullptr);
            // flags from the bitmap, the workplace's type from the type bytes
            const bool homeFlag = home != NOLOC && test_flag(a.L, home);
            const bool workFlag = work != NOLOC && test_flag(a.L, work);
            const uint32_t workType = work != NOLOC ? (uint32_t)a.L.type8[work] : (uint32_t)LOC_TYPE_NONE;
            bool schoolFlag = false, classFlag = false;
            if (school != NOLOC) {
                schoolFlag = test_flag(a.L, school);
                uint32_t cls = find_loc(A, a.L, lt, csr, lt.classroom, 0, &rs);
                if (cls != NOLOC) classFlag = true;
            }
            double prob = a.pRandom;
            prob += (double)homeFlag * a.pHome;
            prob += (double)workFlag * a.pWork;
            prob += (double)schoolFlag * a.pSchool;
            prob += (double)classFlag * 3.0 * a.pSchool;
            if (work != NOLOC && (workType == lt.doctor || workType == lt.hospital)) prob += a.pHospital;
            if (work != NOLOC && workType == lt.nursery_home) prob += a.pNursery;
            if (r.hospTs <= ts && r.hospUntil > ts) prob += a.pHospital;
            if (prob > 1.0 || rs.real(1.0) < prob) {
                A.lastTest[i] = ts;
                lastTest = ts;
                if (a.T.stInfected[st]) {
                    float pr = a.usePCR ? a.T.accPCR[st] : a.T.accAnt[st];
                    if ((double)pr > rs.real(1.0)) {
                        A.rec[i].flags = (uint8_t)(fl | F_DIAGNOSED);
                        A.rec[i].diagTs = ts;
                        diagTs = ts;
                    }
                } else {
                    if (r.quarUntil > ts && (home != NOLOC && a.L.rec[home].quarUntil < ts)) {
                        if (A.daysInQ[i] > 0) A.daysInQ[i] -= (r.quarUntil - ts) / day;
                        A.rec[i].quarUntil = ts;
                    }
                }
            }
        }
        if (lastTest == ts) atomicAdd(&sh[0], 1u);
        if (diagTs == ts) atomicAdd(&sh[1], 1u);
        if (diagTs < ts && diagTs > ts - day) atomicAdd(&sh[2], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 3 && sh[threadIdx.x]) atomicAdd(&a.counters[threadIdx.x], (unsigned long long)sh[threadIdx.x]);
}

// ------------------------------------------------------------ progression
// Simulation::updateAgents (simulation.h:425-471), DynamicPPState::update
// (dynamicPPState.cpp:246-303), NextStates::selectNext (multiBadMatrix.cpp:9-31)
__global__ void k_progression(Agents A, Tables T, uint64_t seed, uint32_t replica, uint32_t ts, uint32_t dt) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    int16_t days = A.days[i];
    if (days > 0) {
        A.days[i] = (int16_t)(days - 1);
        return;
    }
    if (days != -2 && days != 0) return;
    uint8_t st = A.rec[i].state;
    uint8_t v = A.rec[i].variant;
    uint32_t prog = A.prog[i];
    float progS;
    if (st == 7) progS = T.deathScaling[v] / scaling_symptoms(A, i, v, st);
    else progS = T.progScaling[v];
    float scaling = scaling_symptoms(A, i, v, st) * progS;
    Stream rs(seed, A.id[i], ts, PH_PROGRESSION, replica);
    if (days == -2) {
        days = (int16_t)calc_days(T, prog, st, v, rs);
        if (v >= 3) {
            double d = days * 0.7;
            days = (int16_t)(d == 0 ? 1.0 : d);
        }
    }
    if (days == 0) {
        uint8_t old = st;
        bool wasInf = T.stInfected[st] != 0;
        // selectNext
        size_t ps_ = (size_t)prog * T.nStates + st;
        uint32_t eb = T.edgeOff[ps_], nb = T.badCount[ps_], ee = T.edgeOff[ps_ + 1];
        uint32_t nNeu = ee - eb - nb;
        float sc = nNeu == 0 ? 1.0f : scaling;
        double random = rs.unit();
        double preSum = 0.0, badSum = 0.0;
        uint32_t next = 0xFFFFFFFFu;
        bool bad = false;
        for (uint32_t j = 0; j < nb; ++j) {
            float term = T.edgeChance[eb + j] * sc;
            preSum += (double)term;
            if (random < preSum) {
                next = T.edgeTo[eb + j];
                bad = true;
                break;
            }
            badSum += (double)T.edgeChance[eb + j];
        }
        if (!bad && nNeu == 0) {  // bad chances summed below the draw
            next = T.edgeTo[eb + nb - 1];
            bad = true;
        }
        if (!bad) {
            double badScaling = 1.0 + ((badSum - preSum) / (1.0 - badSum));
            if (nb == 0) badScaling = 1.0;
            uint32_t idx = 0;
            do {
                preSum += (double)T.edgeChance[eb + nb + idx] * badScaling;
                ++idx;
            } while (preSum < random && idx < nNeu);
            idx--;
            next = T.edgeTo[eb + nb + idx];
        }
        int nd = calc_days(T, prog, next, v, rs);
        st = (uint8_t)next;
        A.rec[i].state = st;
        A.rec[i].wb = T.stWB[st];
        update_meta(A, T, i, st);
        days = (int16_t)nd;
        if (v >= 3 && st < 8) {
            float mul = st < 6 ? 0.7f : 0.5f;
            float d = (float)days * mul;
            days = (int16_t)(d == 0 ? 1.0f : d);
        }
        if (bad) {
            A.worstState[i] = (int8_t)st;
            A.worstTs[i] = ts;
            A.worstEndTs[i] = ts + (uint32_t)((int)days * 60 * 24) / dt;
        } else if (old != st) {
            if (wasInf && !T.stInfected[st]) {
                A.recTs[i] = ts;
                A.rec[i].flags = (uint8_t)(A.rec[i].flags & ~F_DIAGNOSED);
            }
        }
    }
    A.days[i] = days;
}

// ------------------------------------------------------------ immunity
struct ImmArgs {
    Agents A;
    Tables T;
    uint32_t ts, dt;
    float eff[5][6][2];  // inf, 2 dose, 3 dose, inf+2, inf+3; rows (comp + 3*different); (waning, level)
    int sim[MAXV];
    float simMul[MAXV];
};

__device__ __forceinline__ float fmin_ref(float a, float b) { return a < b ? a : b; }

// Immunization::update, susceptibility of state-0 agents (immunization.h:547-695)
__global__ void k_immunization(ImmArgs a) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const Agents& A = a.A;
    if (i >= A.n) return;
    if (A.rec[i].state != 0) return;
    const uint8_t infC = A.infCount[i], immC = A.immCount[i];
    // never infected nor immunized: susceptibility 1 for every variant, which
    // it already is when state 0 is a susceptible state (set with that state
    // when the agent was created, and written since only by an infection or
    // by the branches below)
    if (infC == 0 && immC == 0 && a.T.stSusc[0]) return;
    const uint32_t nVar = a.T.nVar;
    const uint32_t ts = a.ts == 0 ? 1u : a.ts;
    const uint32_t week = 7u * 24u * 60u / a.dt;
    float susc[MAXV];
    for (uint32_t v = 0; v < nVar; ++v) susc[v] = 1.0f;
    const uint32_t wImm = (ts - A.immTs[i]) / week;
    const uint32_t wInf = (ts - A.infTs[i]) / week;
    const uint32_t prior = A.statVar[i];
    float scal = A.scal[i];
    auto lin = [&](int e, int row, uint32_t weeks) -> float {
        return -1.0f * a.eff[e][row][0] * (float)weeks + a.eff[e][row][1];
    };
    auto mix = [&](int e, int comp, int similar, uint32_t v, uint32_t weeks) -> float {
        int d = (1 - similar) * 3 + comp, s = similar * 3 + comp;
        return -1.0f * (a.eff[e][d][0] + (a.eff[e][s][0] - a.eff[e][d][0]) * a.simMul[v]) * (float)weeks + a.eff[e][d][1] +
               (a.eff[e][s][1] - a.eff[e][d][1]) * a.simMul[v];
    };
    auto setSlot = [&](uint32_t v, float x) { A.slot6[(size_t)v * A.n + i] = scal * x; };
    if (infC > 0 && (immC == 0 || (immC == 1 && wImm < 4))) {
        for (uint32_t v = 0; v < nVar; ++v) {
            int similar = 0;
            for (uint32_t j = 0; j < nVar; ++j) similar = similar || ((prior & (1u << j)) && a.sim[v] == a.sim[j]);
            int d = (1 - similar) * 3;
            float vsInf = (float)((double)(-1.0f * a.eff[0][d][0] * (float)wInf) +
                                  (((1u << v) & prior) ? 0.95 : (double)a.eff[0][d][1]));
            bool mixIt = similar && !((1u << v) & prior) && a.simMul[v] > 0.0f;
            if (mixIt) vsInf = mix(0, 0, similar, v, wInf);
            susc[v] = 1.0f - vsInf;
            float vsSympt = mixIt ? mix(0, 1, similar, v, wInf) : lin(0, d + 1, wInf);
            float vsHosp = mixIt ? mix(0, 2, similar, v, wInf) : lin(0, d + 2, wInf);
            setSlot(v, fmin_ref(1.0f, (1.0f - vsHosp) / (1.0f - vsSympt)));  // slots 2, 4, 5 share index 6; the last write wins
        }
    } else if (infC == 0 && immC > 0) {
        for (uint32_t v = 0; v < nVar; ++v) {
            int similar = a.sim[v] == 0;
            int d = (1 - similar) * 3;
            if (immC == 1 && wImm == 0) continue;
            if (immC == 1 && wImm < 4) {
                float vsInf = (-1.0f * a.eff[1][d][0] * (float)wImm + a.eff[1][d][1]) / 2.0f;
                susc[v] = 1.0f - vsInf;
                float vsSympt = (-1.0f * a.eff[1][d + 1][0] * (float)wImm + a.eff[1][d + 1][1]) / 2.0f;
                float vsHosp = (-1.0f * a.eff[1][d + 2][0] * (float)wImm + a.eff[1][d + 2][1]) / 2.0f;
                setSlot(v, (1.0f - vsHosp) / (1.0f - vsSympt));
            } else {
                int e = immC == 1 ? 1 : 2;
                bool mixIt = similar && a.simMul[v] > 0.0f;
                float vsInf = mixIt ? mix(e, 0, similar, v, wImm) : lin(e, d + 0, wImm);
                susc[v] = 1.0f - vsInf;
                float vsSympt = mixIt ? mix(e, 1, similar, v, wImm) : lin(e, d + 1, wImm);
                float vsHosp = mixIt ? mix(e, 2, similar, v, wImm) : lin(e, d + 2, wImm);
                setSlot(v, (1.0f - vsHosp) / (1.0f - vsSympt));
            }
        }
    } else if (infC > 0 && immC > 0) {
        uint32_t weeks = wInf < wImm ? wInf : wImm;
        int e = immC == 1 ? 3 : 4;
        for (uint32_t v = 0; v < nVar; ++v) {
            int similar = 0;
            for (uint32_t j = 0; j < nVar; ++j) similar = similar || ((prior & (1u << j)) && a.sim[v] == a.sim[j]);
            int d = (1 - similar) * 3;
            bool mixIt = similar && !((1u << v) & prior) && a.simMul[v] > 0.0f;
            float vsInf = (float)((double)(-1.0f * a.eff[e][d][0] * (float)weeks) +
                                  (((1u << v) & prior) ? 0.95 : (double)a.eff[e][d][1]));
            if (mixIt) vsInf = mix(e, 0, similar, v, weeks);
            susc[v] = 1.0f - vsInf;
            float vsSympt = mixIt ? mix(e, 1, similar, v, weeks) : lin(e, d + 1, weeks);
            float vsHosp = mixIt ? mix(e, 2, similar, v, weeks) : lin(e, d + 2, weeks);
            setSlot(v, fmin_ref(1.0f, (1.0f - vsHosp) / (1.0f - vsSympt)));
        }
    }
    for (uint32_t v = 0; v < nVar; ++v) A.susc[(size_t)v * A.suscStride + i] = fmin_ref(1.0f, susc[v]);
}

// ------------------------------------------------------------ vaccination rollout
// Membership of the ten categories of initCategories (immunization.h:229-339),
// indexed as the reference's pass loop indexes them; location-type ids 2, 3,
// 4, 12, 14 and 22 are literals there.
__device__ __forceinline__ bool in_vacc_category(const Agents& A, const Locs& L, uint32_t i, const AgentRec& r, uint32_t g,
                                                 uint8_t precond) {
    const uint32_t age = r.age;
    auto anyEntry = [&](uint32_t kind) {
        for (uint32_t e = r.csrBegin; e < r.csrBegin + r.csrCount; ++e) {
            const uint32_t t = A.possType[e];
            const LocRec lr = L.rec[A.possLoc[e]];
            bool hit = false;
            if (kind == 0) hit = t == 4 && (lr.type == 12 || lr.type == 14);
            else if (kind == 1) hit = (t == 4 || t == 2) && lr.type == 22;
            else if (kind == 4) hit = t == 4 && lr.essential == 1;
            else hit = t == 4 && lr.type == 3;
            if (hit) return true;
        }
        return false;
    };
    switch (g) {
        case 0: return anyEntry(0);
        case 1: return anyEntry(1);
        case 2: return age >= 60;
        case 3: return precond > 0 && age >= 18 && age < 60;
        case 4: return anyEntry(4);
        case 5: return age > 17 && age < 60;
        case 6: return precond > 0 && age >= 60;
        case 7: return anyEntry(7);
        case 8: return age >= 12 && age < 18;
        case 9: return age >= 5 && age < 12;
    }
    return false;
}

struct VaccCategoryArgs {
    Agents A;
    Locs L;
    uint64_t seed;
    uint32_t replica;
    uint32_t nPasses;
    uint8_t passes[10];   // group indices in visiting order
    uint8_t round[10];    // round of each group (its immunizationOrder value)
    float level[10];      // willingness of each group
};

// initCategories: an agent is decided by the first visited group it belongs
// to, with one draw.
__global__ void k_vacc_categories(VaccCategoryArgs a) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.A.n) return;
    const AgentRec r = a.A.rec[i];
    const uint8_t pre = a.A.precond[i];
    uint8_t round = 0;
    for (uint32_t p = 0; p < a.nPasses; ++p) {
        const uint32_t g = a.passes[p];
        if (!in_vacc_category(a.A, a.L, i, r, g, pre)) continue;
        Stream rs(a.seed, a.A.id[i], 0, PH_IMM_CATEGORY, a.replica);
        round = rs.unit() < (double)a.level[g] ? a.round[g] : (uint8_t)255;
        break;
    }
    a.A.immRound[i] = round;
}

__device__ __forceinline__ int age_group_of(uint32_t age) { return age < 90 ? (int)(age / 10) : (age < 200 ? 9 : -1); }

// ages per 10-year group (initAgeGroups, immunization.h:421-479)
__global__ void k_age_groups(Agents A, unsigned long long* out) {
    __shared__ unsigned int sh[10];
    if (threadIdx.x < 10) sh[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < A.n; i += gridDim.x * blockDim.x) {
        int g = age_group_of(A.rec[i].age);
        if (g >= 0) atomicAdd(&sh[g], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 10 && sh[threadIdx.x]) atomicAdd(&out[threadIdx.x], (unsigned long long)sh[threadIdx.x]);
}

struct VaccDayArgs {
    Agents A;
    Tables T;
    uint64_t seed;
    uint32_t replica;
    uint32_t ts;     // timestamp, 1 at timestamp 0 (immunization.h:485)
    uint32_t tsKey;  // timestamp of the draws
    uint32_t spd;    // steps per day
    uint32_t round;  // booster round / category round
    uint32_t draw;   // draw index of this pass
    uint32_t anyRound;
    float prob[10];
};

// booster eligibility (immunization.h:707-717)
__device__ __forceinline__ bool booster_eligible(const VaccDayArgs& a, uint32_t i, const AgentRec& r) {
    const uint32_t it = a.A.immTs[i];
    return it > 0 && a.A.immCount[i] >= 1 && (a.ts - it) / a.spd > 4u * 30u && a.T.stWB[r.state] == 0 &&
           (a.ts >= a.spd * 3u * 30u && r.diagTs < a.ts - a.spd * 4u * 30u);
}

// first-dose eligibility (immunization.h:849-854)
__device__ __forceinline__ bool vacc_eligible(const VaccDayArgs& a, uint32_t i, const AgentRec& r) {
    return a.A.immCount[i] == 0 && a.T.stWB[r.state] == 0 &&
           ((a.ts < a.spd * 3u * 30u && r.diagTs == 0) || (a.ts >= a.spd * 3u * 30u && r.diagTs < a.ts - a.spd * 3u * 30u));
}

__global__ void k_booster_count(VaccDayArgs a, unsigned long long* out) {
    __shared__ unsigned int sh[10];
    if (threadIdx.x < 10) sh[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.A.n; i += gridDim.x * blockDim.x) {
        const AgentRec r = a.A.rec[i];
        int g = age_group_of(r.age);
        if (g >= 0 && booster_eligible(a, i, r)) atomicAdd(&sh[g], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 10 && sh[threadIdx.x]) atomicAdd(&out[threadIdx.x], (unsigned long long)sh[threadIdx.x]);
}

// the age groups are disjoint, so one pass applies all ten in the
// reference's group-by-group order
__global__ void k_booster_apply(VaccDayArgs a) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.A.n) return;
    const AgentRec r = a.A.rec[i];
    int g = age_group_of(r.age);
    if (g < 0 || !booster_eligible(a, i, r)) return;
    const float prob = a.prob[g];
    Stream rs(a.seed, a.A.id[i], a.tsKey, PH_VACCINATION, a.replica);
    rs.draw = a.draw;
    if (prob == 1.0f || rs.unit() < (double)prob) {
        a.A.immTs[i] = a.ts;
        a.A.immCount[i] = (uint8_t)(a.A.immCount[i] + 1);
    }
}

// per-thread counts, one LDS atomic per thread and one global atomic per block
__device__ __forceinline__ void block_count_add(unsigned int mine, unsigned long long* out) {
    __shared__ unsigned int sh;
    if (threadIdx.x == 0) sh = 0;
    __syncthreads();
    if (mine) atomicAdd(&sh, mine);
    __syncthreads();
    if (threadIdx.x == 0 && sh) atomicAdd(out, (unsigned long long)sh);
}

__global__ void k_vacc_count(VaccDayArgs a, unsigned long long* out) {
    unsigned int mine = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.A.n; i += gridDim.x * blockDim.x) {
        const AgentRec r = a.A.rec[i];
        mine += ((a.A.immRound[i] == a.round || a.anyRound) && vacc_eligible(a, i, r)) ? 1u : 0u;
    }
    block_count_add(mine, out);
}

__global__ void k_vacc_apply(VaccDayArgs a) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.A.n) return;
    if (a.A.immRound[i] != a.round) return;
    const AgentRec r = a.A.rec[i];
    if (!vacc_eligible(a, i, r)) return;
    Stream rs(a.seed, a.A.id[i], a.tsKey, PH_VACCINATION, a.replica);
    rs.draw = a.draw;
    if (a.prob[0] == 1.0f || rs.unit() < (double)a.prob[0]) {
        a.A.immTs[i] = a.ts;
        a.A.immCount[i] = 1;
    }
}

__global__ void k_count_immunized(Agents A, uint32_t ts, unsigned long long* out) {
    unsigned int mine = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < A.n; i += gridDim.x * blockDim.x)
        mine += A.immTs[i] == ts ? 1u : 0u;
    block_count_add(mine, out);
}

// ------------------------------------------------------------ non-COVID hospitalisation
__constant__ double c_randomHospProbs[84] = {
    0.00017785,  0.000118567, 0.000614453, 0.003174864, 0.001016573, 0.0011019,   0.000926381, 9.52062E-05,
    6.34708E-05, 0.000328928, 0.000986754, 0.00044981,  0.000623826, 0.000855563, 0.006333678, 0.004222452,
    0.008444904, 0.030227275, 0.036202732, 0.03924143,  0.032990764, 0.003390532, 0.002260355, 0.00452071,
    0.009394699, 0.016018856, 0.022216016, 0.030468759, 0.000364207, 0.000242805, 0.001240249, 0.006390329,
    0.002081773, 0.002256508, 0.001897075, 9.97603E-05, 6.65068E-05, 0.000335,    0.000999373, 0.000471326,
    0.000653666, 0.000896488, 0.002751133, 0.001834089, 0.003668178, 0.013129695, 0.015725228, 0.017045135,
    0.014330059, 0.001472731, 0.000981821, 0.001963642, 0.004080736, 0.006958043, 0.009649877, 0.013234587,
    0.00036181,  0.000241207, 0.001237053, 0.006378891, 0.002068074, 0.002241659, 0.001884591, 9.84773E-05,
    6.56515E-05, 0.000333289, 0.000995818, 0.000465264, 0.000645259, 0.000884959, 0.012211473, 0.008140982,
    0.016281963, 0.058278861, 0.069799674, 0.075658351, 0.063606928, 0.006537022, 0.004358015, 0.008716029,
    0.018113189, 0.030884712, 0.042832973, 0.058744446};

// Simulation::otherDisease per-agent part (simulation.h:206-422)
__global__ void k_other_disease(Agents A, Tables T, uint64_t seed, uint32_t replica, uint32_t ts, uint32_t dt,
                                float probabilityMul) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    uint8_t st = A.rec[i].state;
    if (T.stWB[st] == 4) return;
    Stream rs(seed, A.id[i], ts, PH_OTHER_DISEASE, replica);
    uint8_t age = A.rec[i].age;
    uint32_t ag = age < 5 ? 0 : age < 15 ? 1 : age < 30 ? 2 : age < 60 ? 3 : age < 70 ? 4 : age < 80 ? 5 : 6;
    uint32_t notSex = A.sex[i] ? 0u : 1u;
    uint8_t pc = A.precond[i];
    uint32_t type = pc == 2 ? 1 : (pc == 4 ? 2 : 0);
    if (ts > 0 && A.rec[i].hospUntil == ts) {
        if (rs.real(1.0) < c_randomHospProbs[type * 4 * 7 + 2 * 7 + notSex * 7 + ag]) {
            // die(false) (dynamicPPState.cpp:313-325)
            uint8_t d = (uint8_t)T.nonCOVIDDead;
            A.rec[i].state = d;
            A.rec[i].wb = T.stWB[d];
            A.days[i] = -1;
            update_meta(A, T, i, d);
            A.worstState[i] = (int8_t)d;
            A.worstTs[i] = ts;
            return;
        }
    }
    uint8_t w = T.stWB[st];
    if (w == 4 || w == 3 || ts < A.rec[i].hospUntil) return;
    rs.skip();  // sudden death: disabled in the reference, its draw still made (and unused)
    const int multDay[4] = {99, 99 + 365, 99 + 2 * 365, 99 + 3 * 365};
    const double mult[4] = {0.82, 0.67, 0.9, 0.9};
    int day = (int)(ts / (24u * 60u / dt));
    int d = 0;
    while (d < 3 && day > multDay[d]) d++;
    double probability = c_randomHospProbs[type * 4 * 7 + notSex * 7 + ag] *
                         (probabilityMul > 0.0f ? (double)probabilityMul : mult[d]);
    if (rs.real(1.0) < probability) {
        const double avgLengths[3] = {5.55, 2.78, 5.24};
        unsigned avg = (unsigned)avgLengths[type];
        double p = 1.0 / (double)avg;
        unsigned length = (unsigned)geometric(p, rs);
        if (length == 0) length = 1;
        A.rec[i].hospTs = ts;
        A.rec[i].hospUntil = ts + length * 24u * 60u / dt;
    }
}

__global__ void k_count_covid_hospital(Agents A, Locs L, Tables T, uint32_t ts, uint32_t dueWith,
                                       unsigned long long* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool c = false;
    if (i < A.n) {
        uint8_t st = A.rec[i].state;
        if (dueWith == 0) c = st >= 6 && st <= 8;
        else
            c = T.stInfected[st] && L.rec[A.location[i]].type == T.lt.hospital &&
                !(T.stWB[st] < 3 && ts > 0 && A.rec[i].hospUntil == ts);
    }
    unsigned long long m = __ballot(c);
    if ((threadIdx.x & 63u) == 0 && m) atomicAdd(out, (unsigned long long)__popcll(m));
}

// ------------------------------------------------------------ statistics
enum StatSlot {
    S_HOSP = 0, S_STAYED, S_Q, S_QT, S_NQ, S_ALLINF, S_NI, S_INF, S_REINF, S_BSTR, S_HCI, S_HCE, S_INFH, S_VNI,
    S_VARINF = 14,            // + v
    S_INFV = S_VARINF + MAXV, // + v
    S_IMM = S_INFV + MAXV,    // + v (2^-24 fixed point)
    S_STATE = S_IMM + MAXV,   // + state
    S_COUNT = S_STATE + 128
};

__global__ void __launch_bounds__(256) k_stats(Agents A, Locs L, Tables T, uint32_t ts, uint32_t dt,
                                               unsigned long long* out) {
    __shared__ unsigned long long sh[S_COUNT];
    for (uint32_t k = threadIdx.x; k < S_COUNT; k += blockDim.x) sh[k] = 0;
    __syncthreads();
    const uint32_t day = 24u * 60u / dt;
    const uint32_t nVar = T.nVar;
    uint32_t stride = gridDim.x * blockDim.x;
    unsigned int c[14] = {0};
    unsigned int varInf[MAXV] = {0}, infV[MAXV] = {0};
    unsigned long long imm[MAXV] = {0};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < A.n; i += stride) {
        const AgentRec r = A.rec[i];
        uint8_t st = r.state;
        uint8_t w = T.stWB[st];
        uint8_t fl = r.flags;
        bool inf = T.stInfected[st] != 0;
        {  // state histogram: one LDS atomic per distinct state of the wave
            const uint32_t lane = threadIdx.x & 63u;
            unsigned long long todo = __ballot(true);
            bool mine = true;
            while (todo) {
                const uint32_t first = (uint32_t)(__ffsll((long long)todo) - 1);
                const uint32_t s0 = (uint32_t)__shfl((int)st, (int)first, 64);
                const unsigned long long m = __ballot(mine && st == s0);
                if (lane == first) atomicAdd(&sh[S_STATE + s0], (unsigned long long)__popcll(m));
                if (st == s0) mine = false;
                todo &= ~m;
            }
        }
        c[S_HOSP] += (w != 4 && w != 3 && ts < r.hospUntil);
        c[S_STAYED] += (fl & F_STAYED_HOME) ? 1 : 0;
        bool isQ = A.quarTs[i] <= ts && r.quarUntil > ts;
        c[S_Q] += isQ;
        c[S_QT] += isQ && inf;
        c[S_NQ] += !isQ && inf;
        c[S_ALLINF] += inf;
        uint32_t it = A.infTs[i];
        bool newInf = it > ts - day && it <= ts;
        c[S_NI] += newInf;
        uint8_t ic = A.infCount[i], mc = A.immCount[i];
        c[S_INF] += ic > 0;
        c[S_REINF] += ic > 1 ? ic - 1u : 0u;
        c[S_BSTR] += mc > 1 ? mc - 1u : 0u;
        if (fl & F_HCW) {
            c[S_HCI] += inf || (fl & F_QUARANTINED);
            c[S_HCE] += newInf;
        }
        c[S_INFH] += inf && L.rec[A.location[i]].type == T.lt.hospital;
        c[S_VNI] += ic == 0 && mc > 0;
        uint8_t sv = A.statVar[i];
#pragma unroll
        for (uint32_t v = 0; v < (uint32_t)MAXV; ++v) {
            if (v >= nVar) break;
            varInf[v] += (inf && r.variant == v) ? 1u : 0u;
            infV[v] += (sv >> v) & 1u;
            float s = A.susc[(size_t)v * A.suscStride + i];
            imm[v] += (unsigned long long)(long long)((double)(1.0f - s) * 16777216.0);
        }
    }
    for (int k = 0; k < 14; ++k)
        if (c[k]) atomicAdd(&sh[k], (unsigned long long)c[k]);
#pragma unroll
    for (uint32_t v = 0; v < (uint32_t)MAXV; ++v) {
        if (varInf[v]) atomicAdd(&sh[S_VARINF + v], (unsigned long long)varInf[v]);
        if (infV[v]) atomicAdd(&sh[S_INFV + v], (unsigned long long)infV[v]);
        if (imm[v]) atomicAdd(&sh[S_IMM + v], imm[v]);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < S_COUNT; k += blockDim.x)
        if (sh[k]) atomicAdd(&out[k], sh[k]);
}

// ------------------------------------------------------------ planning
struct PlanArgs {
    Agents A;
    Locs L;
    Tables T;
    uint64_t seed;
    uint32_t replica, ts, dt, curfewEnd, schoolAge;
};

// checkSchoolQuarantine (movementPolicies.h:1285-1310), one thread per school
__global__ void k_plan_school(PlanArgs a) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.L.nSchools) return;
    LocRec* q = a.L.rec;
    uint32_t school = a.L.schools[s];
    if (q[school].quarUntil > a.ts) return;
    uint32_t counter = 0, mx = 0;
    for (uint32_t c = a.L.classOff[s]; c < a.L.classOff[s + 1]; ++c) {
        uint32_t cu = q[a.L.classrooms[c]].quarUntil;
        if (cu > a.ts) {
            counter++;
            mx = mx > cu ? mx : cu;
        }
    }
    if (counter > 1) {
        q[school].quarUntil = mx;
        mark_special(a.L, school);
        for (uint32_t c = a.L.classOff[s]; c < a.L.classOff[s + 1]; ++c) {
            q[a.L.classrooms[c]].quarUntil = mx;
            mark_special(a.L, a.L.classrooms[c]);
        }
    }
}

// Per-agent part of planLocations in one pass: the resets of
// movementPolicies.h:1446-1451 (stepsUntilMove := curfewEnd, stayedHome,
// noWorkAgent) and, with quarantinePolicy > 2, checkSchoolWorkQuarantine
// (movementPolicies.h:1188-1240).
__global__ void k_plan_agents(PlanArgs a, uint32_t schoolwork) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const Agents& A = a.A;
    if (i >= A.n) return;
    A.wake[i] = a.curfewEnd < 1440u / a.dt ? (uint16_t)a.curfewEnd : WAKE_NONE;
    const AgentRec r = A.rec[i];
    uint8_t fl = (uint8_t)((r.flags | F_STAYED_HOME) & ~F_NO_WORK);
    if (schoolwork && !(fl & F_QUARANTINED) && !(r.diagTs > 0 && !a.T.stInfected[r.state])) {
        const pansim_location_types& lt = a.T.lt;
        Stream rs(a.seed, a.A.id[i], a.ts, PH_PLAN_SCHOOLWORK, a.replica);
        const CsrReg csr = load_csr(A, i, r);
        // the planning lookup (home, school, classroom first-match): a
        // workplace is drawn when the agent has several, in this order
        uint32_t wl = find_loc(A, a.L, lt, csr, lt.work, 0, &rs, true);
        uint32_t sl = find_loc(A, a.L, lt, csr, lt.school, 0, nullptr, true);
        uint32_t cl = find_loc(A, a.L, lt, csr, lt.classroom, 0, nullptr, true);
        // the three quarantine times loaded together, then the first that
        // applies; a location whose special bit is clear was last quarantined
        // until at most the first step of the day that ends now (every
        // quarantine write since, tonight's school quarantines included, sets
        // the bit), so its record is not read
        const LocRec* q = a.L.rec;
        const uint32_t qs = sl != NOLOC && loc_special(a.L, sl) ? q[sl].quarUntil : 0u;
        const uint32_t qc = cl != NOLOC && loc_special(a.L, cl) ? q[cl].quarUntil : 0u;
        const uint32_t qw = wl != NOLOC && loc_special(a.L, wl) ? q[wl].quarUntil : 0u;
        uint32_t until = 0;
        if (qs > a.ts) until = qs;
        else if (qc > a.ts) until = qc;
        else if (qw > a.ts) until = qw;
        if (until > 0) {
            fl |= F_QUARANTINED;
            A.quarTs[i] = a.ts;
            A.rec[i].quarUntil = until;
            A.daysInQ[i] += (until - a.ts) / (24u * 60u / a.dt);
        }
    }
    if (fl != r.flags) A.rec[i].flags = fl;
}

// Per-location part: noWork reset (movementPolicies.h:1449) and, with
// quarantinePolicy > 2, the quarantine -> closedUntil conversion
// (movementPolicies.h:1544-1557).
// Also recomputes the special-location bits for the day (loc_special): each
// wave covers 64 consecutive locations (block size and grid stride are
// multiples of 64) and writes its two bitmap words from a ballot.
__global__ void k_plan_locs(PlanArgs a, uint32_t close) {
    uint32_t stride = gridDim.x * blockDim.x;
    const pansim_location_types& lt = a.T.lt;
    const uint32_t dayStart = a.ts - a.ts % (1440u / a.dt);
    for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < a.L.n; l += stride) {
        LocRec* lr = &a.L.rec[l];
        lr->noWork = 0u;
        uint32_t q = lr->quarUntil;
        if (close) {
            uint32_t t = lr->type;
            if (q > a.ts && (t == lt.school || t == lt.classroom)) {
                lr->closedUntil = q;
                lr->quarUntil = q = 0;
            } else if (q > a.ts && t == lt.work) {
                lr->quarUntil = q = 0;
            }
        }
        const bool sp = q > dayStart || lr->closedUntil > dayStart || lr->open == 0;
        const unsigned long long bits = __ballot(sp);
        const uint32_t lane = threadIdx.x & 63u;
        if (lane == 0u) a.L.special[l >> 5] = (uint32_t)bits;
        if (lane == 32u) a.L.special[l >> 5] = (uint32_t)(bits >> 32);
        if ((lane & 31u) == 0u) a.L.noWorkBits[l >> 5] = 0u;  // with noWork above
    }
}

constexpr uint32_t NOWORK_FLAGGED = 0xFFFFFFFFu;

// checkUnderageAtHome (movementPolicies.h:1056-1094), over the list of the
// agents aged 14 or less (`under`, n entries)
__global__ void k_plan_underage(PlanArgs a, const uint32_t* under, uint32_t n) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const Agents& A = a.A;
    if (j >= n) return;
    const uint32_t i = under[j];
    const pansim_location_types& lt = a.T.lt;
    bool flag = false;
    if (A.rec[i].flags & F_QUARANTINED) flag = true;
    else {
        const CsrReg csr = load_csr(A, i, A.rec[i]);
        uint32_t sl = find_loc(A, a.L, lt, csr, lt.school, 0, nullptr, true);
        uint32_t cl = find_loc(A, a.L, lt, csr, lt.classroom, 0, nullptr, true);
        bool tooOld = A.rec[i].age >= a.schoolAge;
        flag = tooOld || (sl != NOLOC && (a.L.rec[sl].open == 0 || a.L.rec[sl].closedUntil > a.ts)) ||
               (cl != NOLOC && (a.L.rec[cl].open == 0 || a.L.rec[cl].closedUntil > a.ts));
    }
    if (flag) {
        uint32_t h = A.rec[i].home;
        if (h != NOLOC) {
            a.L.rec[h].noWork = NOWORK_FLAGGED;
            atomicOr(&a.L.noWorkBits[h >> 5], 1u << (h & 31u));
        }
    }
}

// A home flagged by tonight's planning (an L2-resident bitmap read before the
// home's location record)
__device__ __forceinline__ bool nowork_flagged(const Locs& L, uint32_t h) {
    return h != NOLOC && ((L.noWorkBits[h >> 5] >> (h & 31u)) & 1u) != 0u;
}

// setNoWorkToday (movementPolicies.h:1137-1158): the lowest eligible agent of
// a flagged home claims it (claim pass, then resolve pass), over the list of
// the agents aged 27-64 (`adults`, n entries)
__global__ void k_plan_nowork_claim(PlanArgs a, const uint32_t* adults, uint32_t n) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const Agents& A = a.A;
    if (j >= n) return;
    const uint32_t i = adults[j];
    uint32_t h = A.rec[i].home;
    // 0 = not flagged, NOWORK_FLAGGED = flagged, else (claiming agent's
    // input index + 1)
    if (nowork_flagged(a.L, h)) atomicMin(&a.L.rec[h].noWork, A.id[i] + 1u);
}

__global__ void k_plan_nowork_resolve(PlanArgs a, const uint32_t* adults, uint32_t n) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const Agents& A = a.A;
    if (j >= n) return;
    const uint32_t i = adults[j];
    uint32_t h = A.rec[i].home;
    if (nowork_flagged(a.L, h) && a.L.rec[h].noWork == A.id[i] + 1u) A.rec[i].flags = (uint8_t)(A.rec[i].flags | F_NO_WORK);
}

// ------------------------------------------------------------ closures
__global__ void k_set_open(Locs L, const uint32_t* types, uint32_t n, uint8_t open) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < L.n; l += stride) {
        if (L.rec[l].essential == 1) continue;
        for (uint32_t k = 0; k < n; ++k)
            if (L.rec[l].type == types[k]) {
                L.rec[l].open = open;
                if (!open) mark_special(L, l);
            }
    }
}

__global__ void k_scale_inf(Locs L, uint32_t homeType, double factor, uint32_t divide, const float* typeMul) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < L.n; l += stride) {
        const uint32_t t = L.rec[l].type;
        if (t == homeType) continue;
        const float v = divide ? (float)((double)L.infRaw[l] / factor) : (float)((double)L.infRaw[l] * factor);
        L.infRaw[l] = v;
        L.dyn[l].inf = v * typeMul[t];
    }
}

// The infection factor of every location after a change of the type
// multipliers.
__global__ void k_loc_infm(Locs L, const float* typeMul) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < L.n; l += stride)
        L.dyn[l].inf = L.infRaw[l] * typeMul[L.rec[l].type];
}

// ExposeToMutation (closurePolicies.h:356-390)
__global__ void k_expose_mutation(Agents A, Tables T, uint64_t seed, uint32_t replica, uint32_t ts, uint32_t v,
                                  double fraction) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    float s = A.susc[(size_t)v * A.suscStride + i];
    if (!(s > 0.0f)) return;
    Stream rs(seed, A.id[i], ts, PH_MUTATION, replica);
    if (rs.unit() < fraction * (double)s) {
        uint8_t fi = (uint8_t)T.firstInfected;
        A.rec[i].state = fi;
        A.rec[i].wb = T.stWB[fi];
        A.rec[i].variant = (uint8_t)v;
        A.days[i] = -2;
        update_meta(A, T, i, fi);
        A.infTs[i] = ts;
        A.infLoc[i] = A.location[i];
        A.infCount[i] = (uint8_t)(A.infCount[i] + 1);
        A.worstState[i] = (int8_t)fi;
        A.worstTs[i] = ts;
        A.statVar[i] = (uint8_t)(A.statVar[i] | (1u << v));
    }
}

// ------------------------------------------------------------ test helpers
__global__ void k_expf(const float* in, float* out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = ps_expf(in[i]);
}
__global__ void k_log(const double* in, double* out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = ps_log(in[i]);
}
__global__ void k_philox(const uint32_t* ctr, uint32_t k0, uint32_t k1, uint32_t* out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u32x4 c;
    for (int j = 0; j < 4; ++j) c.v[j] = ctr[4 * i + j];
    u32x4 r = philox4x32_10(c, k0, k1);
    for (int j = 0; j < 4; ++j) out[4 * i + j] = r.v[j];
}

}  // namespace dev
}  // namespace ps
