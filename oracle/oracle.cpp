// ORACLE -- test infrastructure only.  Nothing in the product links or calls
// this file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load it, as the checker.
//
// A sequential CPU re-statement of the reference PanSim per-timestep agent
// pipeline (khbence/pansim), kept close to the reference's own structure: an
// array of per-agent records (PPState, AgentMeta, AgentStats), the
// stepsUntilMove countdown, a per-step full rebuild of the per-location agent
// counts and infectiousness sums, and one function per reference routine.
// Its random numbers come from the Philox stream of ref_rng.h keyed on
// (seed, agent, step, phase, draw) instead of the reference's unseeded
// per-thread generators, so a run is reproducible and the GPU path can be
// compared with it bit for bit.
//
// Semantics that the reference leaves undefined and that this re-statement
// fixes (the GPU path implements the same rules):
//  * location quarantine writes made during movement (quarantineAgent,
//    movementPolicies.h:202-241) become visible at the end of the step; reads
//    during the step see the values at its start; among several writes to one
//    location in one step the highest agent index wins (= the last writer of a
//    sequential ascending pass).  The reference races here.
//  * setNoWorkToday (movementPolicies.h:1137-1158): the lowest eligible agent
//    index of a flagged home claims it (= the sequential CPU order).
//  * per-location infectiousness sums (util.h:47-103) are exact: every agent's
//    infectiousness is a float >= 2^-8 or 0, so the sum in 2^-32 fixed point
//    is exact and its conversion to float is the correctly rounded sum.  The
//    reference adds in ascending agent order (CPU) or with float atomics (GPU).
//  * IMM statistic (simulation.h:594-602): exact sum of the float terms,
//    truncated; the reference uses an order-dependent float reduction.
//  * weekday = (startDay + day) % 7 and run length = weeks*7 days, instead of
//    the wall-clock localtime arithmetic of timeHandler.cpp:90-120.
//  * bounded retry loops (a cap no realistic run reaches): the closed-location
//    redraw of findActualLocationForType (64 redraws), the geometric
//    resampling of calculateDays (1000 draws), the neutral scan of selectNext
//    (stops at the last neutral entry).
//
// What pins each part against something other than itself (tests/): the
// Philox known-answer vectors, exp/log within one ulp of correctly rounded
// values, the truncated-geometric mean, the geometric distribution, the
// next-state shares of selectNext against the progression files' chances,
// seasonality by date properties and exact cosine points, hand-computed
// movement trajectories and planning midnights on schedules written in the
// tests (with the event and re-pick shares), the input files read
// independently in plain Python, the testing pass one probability term at a
// time, every case of the immunization pass (first dose, after an infection,
// infection plus doses, boosters), the three closure conditions, the Masks,
// ExposeToMutation, Curfew, HolidayMode and ReduceMovement actions, the
// non-COVID hospital stays against otherDisease's probability table, the
// statistics vector recounted from the fields, and the brute-force recount
// of the location sums.  The movement and planning pins and the full-size
// property checks are also run against a planted mistake (the codes below,
// and wrong reported fields) and seen to fail.  Parity unpinned: the spread
// of outcomes over replicas (the ensemble bands), which needs a runnable
// reference.
#include <stdint.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../pansim_amd/csrc/host/scenario.h"
#include "ref_rng.h"
#ifdef ORACLE_OMP
#include <omp.h>

#include <parallel/algorithm>
#endif

namespace oracle {

// Planted mistakes, for the tests that show a pin fails on a wrong oracle
// (0 = none; set through oracle_plant).  Each code breaks one rule:
//  1 Case 3 waits up to 60 min instead of 30 before moving (:883)
//  2 a single event's stay is one step longer than its duration (:796)
//  3 the cumulative event pick compares with >= (:666-672)
//  4 a closed destination is not noticed (:706-709)
//  5 a location-quarantined agent is not sent home (:964-975)
//  6 a diagnosed agent's work/classroom is quarantined only above policy 3 (:209)
//  7 the last eligible adult of a home takes the child care (:1149-1155)
//  8 one quarantined classroom already quarantines the school (:1297)
int plant = 0;

// quarantineLocations pending for the agent being moved (per thread: the
// OpenMP build moves agents in parallel)
thread_local bool qLoc = false;
thread_local unsigned qLocUntil = 0, qLocPrev = 0;

using ps::Scenario;
constexpr unsigned UMAX = 0xFFFFFFFFu;
constexpr int MAXS = 7;

// ---------------------------------------------------------------- TimeDay
// src/tools/timeDay.cpp: (hours, minutes) in unsigned chars
struct TD {
    uint8_t h = 0, m = 0;
    static TD fromMinutes(unsigned mins) {
        TD t;
        t.h = (uint8_t)(mins / 60);
        t.m = (uint8_t)(mins % 60);
        return t;
    }
    unsigned minutes() const { return (unsigned)h * 60u + (unsigned)m; }
    unsigned steps(unsigned dt) const { return minutes() / dt; }
    unsigned stepsUntilMidnight(unsigned dt) const { return (1440u - minutes()) / dt; }
    bool overMidnight() const { return h >= 24; }
    TD plus(TD dur) const {  // TimeDay::operator+=
        TD r = *this;
        unsigned mins = dur.minutes();
        r.m = (uint8_t)(r.m + mins % 60);
        unsigned overflow = r.m / 60;
        r.m = (uint8_t)(r.m % 60);
        r.h = (uint8_t)(r.h + mins / 60 + overflow);
        return r;
    }
    TD minus(TD o) const { return fromMinutes(minutes() - o.minutes()); }  // TimeDay - TimeDay
};
inline bool operator<(TD a, TD b) { return a.h < b.h || (a.h == b.h && a.m < b.m); }
inline bool operator>(TD a, TD b) { return b < a; }
inline bool operator>=(TD a, TD b) { return !(a < b); }

struct Event {
    unsigned locationType;
    float chance;
    TD start, end, duration;
};

// ---------------------------------------------------------------- records
struct PPState {  // dynamicPPState.h:14-52
    float infectious = 0;
    unsigned progressionID = 0;
    int state = 0;
    int16_t days = -1;
    float susceptible[MAXS] = {0};
    uint8_t variant = 0;
};

struct Meta {  // agentMeta.h:10-46
    float scalingSymptoms[7 * MAXS];
    float scalingAgeSex = 1.0f;
    uint8_t age = 0, preCondIdx = 0;
    bool sex = false;
    float get(unsigned v, unsigned s) const {
        unsigned i = v * 7 + (s > 6 ? s : 6);
        if (i < 7 * MAXS) return scalingSymptoms[i];
        return scalingAgeSex;  // index 49 is the next member in the reference layout
    }
    void set(float scaling, unsigned s, unsigned v) {
        unsigned i = v * 7 + (s > 6 ? s : 6);
        if (i < 7 * MAXS) scalingSymptoms[i] = scalingAgeSex * scaling;
    }
};

struct Stats {  // agentStats.h:5-24
    unsigned infectedTimestamp = UMAX;
    unsigned infectedLocation = 0;
    unsigned worstStateTimestamp = 0;
    unsigned worstStateEndTimestamp = 0;
    unsigned diagnosedTimestamp = 0;
    unsigned quarantinedTimestamp = 0;
    unsigned quarantinedUntilTimestamp = 0;
    unsigned daysInQuarantine = 0;
    unsigned hospitalizedTimestamp = 0;
    unsigned hospitalizedUntilTimestamp = 0;
    unsigned immunizationTimestamp = 0;
    unsigned recoveredTimestamp = UMAX;
    uint8_t immunizationCount = 0;
    unsigned variant = 0;
    uint8_t infectedCount = 0;
    int8_t worstState = 0;
};

// ---------------------------------------------------------------- progression
struct LengthOfState {
    float avg[MAXS], max[MAXS];
    double p[MAXS];
};

struct Matrix {
    std::vector<LengthOfState> len;
    std::vector<std::vector<std::pair<unsigned, float>>> bad, neutral;
};

// (unsigned)round(max) as compiled for x86-64: a 64-bit truncating
// conversion whose low 32 bits are kept; out-of-range values give 0.
static unsigned x86_round_to_unsigned(float max) {
    double r = std::round((double)max);
    if (!(r < 9223372036854775808.0) || !(r >= -9223372036854775808.0)) return 0u;
    return (unsigned)(uint64_t)(int64_t)r;
}

// basicLengthAbstract.cpp:4-14
static double expectedLength(double pInput, float max) {
    if (pInput < 0.0) return 0.0;
    double sumExpected = 0.0, sumChance = 0.0;
    unsigned n = x86_round_to_unsigned(max);
    for (unsigned i = 0; i < n; ++i) {
        double chance = std::pow(1 - pInput, (double)i) * pInput;
        sumExpected += (i + 1) * chance;
        sumChance += chance;
    }
    return sumExpected * (1.0 / sumChance);
}

// basicLengthAbstract.cpp:16-29 with SecantMethod of smallTools.h:13-29
static double calculateModifiedP(float avgLength, float maxLength) {
    if (avgLength == -1.0f) return 1.0;
    double n = (double)avgLength;
    double pNew = 1.0 / n;
    float maxLocal = maxLength;
    auto f = [maxLocal, n](double p) { return expectedLength(p, maxLocal) - n; };
    bool stuck = false;
    double x_k = pNew - (pNew * 0.1), x_prev = pNew;
    unsigned it = 0;
    while (0.0001 < std::abs(f(x_k))) {
        double temp = x_k;
        double diff = f(x_k) - f(x_prev);
        if (diff != 0.0) x_k -= f(x_k) * (x_k - x_prev) / diff;
        else {
            stuck = true;
            break;
        }
        x_prev = temp;
        if (++it > 10000) {
            stuck = true;
            break;
        }
    }
    if (!stuck) pNew = x_k;
    if ((pNew * 10.0 < (1.0 / n)) || pNew > 1.0) pNew = 1.0 / n;
    return pNew;
}

// ---------------------------------------------------------------- simulator
class Sim {
public:
    Scenario sc;
    unsigned N = 0, L = 0, dt = 10, stepsPerDay = 144, nStates = 0, nVar = 1;
    uint64_t seed = 0;
    // state tables
    std::vector<float> infTable, accPCR, accAntigen;
    std::vector<uint8_t> wb, suscFlag, infectedFlag;
    unsigned firstInfected = 0, nonCOVIDDead = 0, deadState = 0;
    std::vector<float> variantMul, progScaling, deathScaling;
    std::vector<Matrix> matrices;
    // schedules
    std::vector<unsigned> eventOffset;
    std::vector<Event> events;
    // agents
    std::vector<PPState> pp;
    std::vector<Meta> meta;
    std::vector<Stats> stats;
    std::vector<uint8_t> diagnosed, quarantined, stayedHome, noWorkAgent, hcw;
    std::vector<unsigned> location, types, stepsUntilMove, lastTest;
    std::vector<uint32_t> locOffset, possibleLocations;
    std::vector<uint8_t> possibleTypes;
    // locations
    std::vector<unsigned> locType, closedUntil, quarantineUntil, noWorkLoc, schools, classrooms, classroomOffsets;
    std::vector<float> locInf;
    std::vector<uint8_t> locOpen, essential, locFlags;
    unsigned cemetery = 0;
    unsigned homeT, workT, schoolT, classroomT, hospitalT, doctorT, publicT, nurseryT;
    // per-step scratch
    std::vector<unsigned> locCount;
    std::vector<int64_t> locSum;  // [variant][location], 2^-32 fixed point
    std::vector<float> ratio;
    // policy state
    unsigned timestamp = 0;
    unsigned quarantinePolicy = 3, quarantineLength = 10;
    bool enableCurfew = false;
    unsigned curfewBegin = 0, curfewEnd = 0;
    bool curfewTimeConverted = false;
    unsigned schoolAgeRestriction = 99;
    bool holidayModeActive = false, quarantineImmuneActive = false, lockdownNonvaccActive = false;
    double testingRandom = 0.005, testingHome = 0.2, testingWork = 0.1, testingSchool = 0.1, testingRandomHospital = 0.2,
           testingNurseryHome = 0.3;
    unsigned testingDelay = 5;
    bool usePCR = true;
    unsigned testStats[3] = {0, 0, 0};
    bool testingInitialized = false;
    double k = 0.000347;
    double currentMaskValue = 1.0;
    std::vector<unsigned> variantCounts;
    unsigned healthcareWorkerCount = 0;
    unsigned immunizedToday = 0;
    // with snapOn, a copy of the whole state at the moment the day's
    // statistics are computed (for recounting them in the tests)
    std::vector<unsigned> sortedAgents, sortKey, locOffsets;  // OpenMP build: the per-step location index
    bool snapOn = false;
    std::shared_ptr<Sim> statSnap;
    // calibration interface: location-type multipliers, expected-infection
    // hazard and its gradient (FP64), hazard sums at perturbed parameters
    std::vector<float> typeMul = std::vector<float>(256, 1.0f);
    bool hazardOn = false;
    std::vector<double> hazard = std::vector<double>(258, 0.0);
    struct Probe {
        double dk;
        std::vector<double> dm;
    };
    std::vector<Probe> probes;
    std::vector<double> probeJ;
    // vaccination rollout (immunization.h:24-61)
    std::vector<uint8_t> immRound;  // category round of every agent, 255 = not willing
    unsigned currentCategory = 0;
    bool ageGroupsReady = false;
    std::vector<std::array<unsigned, 10>> ageGroupSize;
    std::vector<std::array<float, 10>> ageGroupFrac;
    std::vector<std::vector<unsigned>> dayStats;  // every statistics vector produced
    std::vector<unsigned> lastStats;
    // closures
    struct Cond {
        std::vector<unsigned> headerPos;
        std::vector<double> history;
        int pos = 0;
        bool active = false;
        std::function<bool(Cond*, std::vector<unsigned>&)> condition;
    };
    struct Rule {
        std::string name;
        std::vector<size_t> conds;
        std::function<void(Rule*)> rule;
        bool previousOpenState = true;
    };
    std::vector<Cond> conds;
    std::vector<Rule> rules;
    std::vector<std::string> header;

    explicit Sim(const Scenario& s);
    void initClosures();
    void initDays();

    // reference routines
    unsigned findLoc(unsigned agent, unsigned locType, int retry, bool useStates, Rng* rng, bool plan = false);
    void quarantineAgent(unsigned i, unsigned until);
    void quarantineLocations(unsigned i, unsigned until, unsigned previousQuarantineUntil,
                             std::vector<std::pair<unsigned, unsigned>>& writes, Rng& rng);
    void doMovementRoutine(unsigned i, unsigned day, TD simTime, Rng& rng);

    void doMovement(unsigned i, unsigned day, TD simTime, std::vector<std::pair<unsigned, unsigned>>& writes);
    void movement();
    void rebuildLocationSums();
    double seasonality();
    void infectionsAtLocations(unsigned variant);
    void planLocations();
    std::vector<unsigned> dummyStepsUntilMove;  // DummyMovement::stepsUntilMove
    void dummyMovement();
    void performTests();
    void updateAgents();
    void immunizationUpdate();
    bool inVaccinationCategory(unsigned i, unsigned g) const;
    void initCategories();
    void initAgeGroups();
    void vaccinationRollout();
    void otherDisease();
    std::vector<unsigned> refreshStatistics();
    void closureMidnight(std::vector<unsigned>& st);
    void countVariantCases();
    void midnight();
    void step();  // one iteration of runSimulation's loop
    void processFlags(const std::vector<std::string>& flags);
    // output files (agentsList.h:280-327, infectionPolicies.h:27-114)
    void finalize();
    void writeAgentStatJSON(const std::string& path) const;
    std::string agentStatDiags() const;
    std::string diagsOut;  // what --diags printed at finalize
    // DumpStore (infectionPolicies.h:27-85)
    std::vector<unsigned> dTs, dVar, dLocIds, dLocOff{0}, dNew, dNewOff{0}, dInf, dInfOff{0};
    std::vector<float> dInfness;
    void dumpInsert(unsigned variant, const std::vector<unsigned>& newlyInfected);
    void writeDump(const std::string& path) const;
    std::vector<unsigned> runForDay(const std::vector<std::string>& flags);

    // helpers
    void updateMeta(PPState& p) {
        p.infectious = infTable[p.state] * variantMul[p.variant];
        for (int i = 0; i < MAXS; ++i) p.susceptible[i] = suscFlag[p.state] ? 1.0f : 0.0f;
    }
    void gotInfected(PPState& p, uint8_t v) {
        p.state = (int)firstInfected;
        p.variant = v;
        p.days = -2;
        updateMeta(p);
    }
    int die(PPState& p, bool covid) {
        p.days = -1;
        p.state = covid ? (int)deadState : (int)nonCOVIDDead;
        updateMeta(p);
        return p.state;
    }
    int geometric(double p, Rng& rng) {
        double log1p_ = log_f64(1.0 - p);
        const double naf = (1 - DBL_EPSILON) / 2;
        const double thr = 2147483647.0 + naf;
        double cand;
        int guard = 0;
        do {
            cand = std::floor(log_f64(1.0 - rng.unit()) / log1p_);
        } while (cand >= thr && ++guard < 1000);
        if (cand >= thr) cand = 2147483647.0;
        return (int)(cand + naf);
    }
    int calculateDays(unsigned prog, unsigned state, int variant, Rng& rng) {
        const LengthOfState& l = matrices[prog].len[state];
        if (l.avg[variant] == -1.0) return -1;
        int days = geometric(l.p[variant], rng);
        int guard = 0;
        while (l.max[variant] < (float)days && ++guard < 1000) days = geometric(l.p[variant], rng);
        if (l.max[variant] < (float)days) days = (int)l.max[variant];
        return days;
    }
    std::pair<unsigned, bool> selectNext(unsigned prog, unsigned state, float scaling, Rng& rng) {
        return selectNextFor(prog, state, scaling, rng.unit());
    }
    // NextStates::selectNext for the uniform draw `random`
    std::pair<unsigned, bool> selectNextFor(unsigned prog, unsigned state, float scaling, double random) const {
        const auto& bad = matrices[prog].bad[state];
        const auto& neu = matrices[prog].neutral[state];
        if (neu.empty()) scaling = 1.0f;
        double preSum = 0.0, badSum = 0.0;
        for (size_t i = 0; i < bad.size(); ++i) {
            float term = bad[i].second * scaling;
            preSum += term;
            if (random < preSum) return {bad[i].first, true};
            badSum += bad[i].second;
        }
        if (neu.empty()) return {bad.back().first, true};  // bad chances summed below the draw
        double badScaling = 1.0 + ((badSum - preSum) / (1.0 - badSum));
        if (bad.empty()) badScaling = 1.0;
        size_t idx = 0;
        do {
            preSum += neu[idx].second * badScaling;
            ++idx;
        } while (preSum < random && idx < neu.size());
        idx--;
        return {neu[idx].first, false};
    }
    bool isInfected(unsigned i) const { return infectedFlag[pp[i].state] != 0; }
    uint8_t wbOf(unsigned i) const { return wb[pp[i].state]; }
    unsigned dayOfWeek() const { return (sc.opt.startDay + timestamp / stepsPerDay) % 7; }
};

Sim::Sim(const Scenario& s) : sc(s) {
    N = sc.agents.n;
    L = sc.locs.n;
    dt = sc.opt.deltat;
    stepsPerDay = sc.stepsPerDay;
    seed = sc.opt.seed;
    nStates = (unsigned)sc.st.names.size();
    nVar = (unsigned)sc.variantMultiplier.size();
    infTable = sc.st.infectious;
    accPCR = sc.st.accuracyPCR;
    accAntigen = sc.st.accuracyAntigen;
    wb = sc.st.wb;
    suscFlag = sc.st.susceptible;
    infectedFlag = sc.st.infected;
    firstInfected = sc.st.firstInfected;
    nonCOVIDDead = sc.st.nonCOVIDDead;
    deadState = sc.st.deadState;
    variantMul = sc.variantMultiplier;
    progScaling = sc.progressionScaling;
    deathScaling = sc.deathScaling;
    // matrices with derived geometric parameters
    for (auto& raw : sc.prog) {
        Matrix m;
        for (auto& row : raw.rows) {
            LengthOfState l;
            for (int v = 0; v < MAXS; ++v) {
                l.avg[v] = row.avgLength[v];
                l.max[v] = row.maxLength[v];
                l.p[v] = calculateModifiedP(l.avg[v], l.max[v]);
            }
            m.len.push_back(l);
            std::vector<std::pair<unsigned, float>> b, n;
            for (auto& x : row.bad) b.emplace_back(x.state, x.chance);
            for (auto& x : row.neutral) n.emplace_back(x.state, x.chance);
            m.bad.push_back(b);
            m.neutral.push_back(n);
        }
        matrices.push_back(m);
    }
    eventOffset.assign(sc.sched.eventOffset.begin(), sc.sched.eventOffset.end());
    for (auto& e : sc.sched.events) {
        Event x;
        x.locationType = e.locType;
        x.chance = e.chance;
        x.start.h = e.start_h;
        x.start.m = e.start_m;
        x.end.h = e.end_h;
        x.end.m = e.end_m;
        x.duration.h = e.dur_h;
        x.duration.m = e.dur_m;
        events.push_back(x);
    }
    homeT = sc.lt.home;
    workT = sc.lt.work;
    schoolT = sc.lt.school;
    classroomT = sc.lt.classroom;
    hospitalT = sc.lt.hospital;
    doctorT = sc.lt.doctor;
    publicT = sc.lt.publicSpace;
    nurseryT = sc.lt.nurseryhome;
    // locations
    locType.assign(sc.locs.type.begin(), sc.locs.type.end());
    locInf = sc.locs.infectiousness;
    locOpen = sc.locs.open;
    essential = sc.locs.essential;
    closedUntil.assign(L, 0);
    quarantineUntil.assign(L, 0);
    noWorkLoc.assign(L, 0);
    locFlags.assign(L, 0);
    schools = sc.locs.schools;
    classrooms = sc.locs.classrooms;
    classroomOffsets = sc.locs.classroomOffsets;
    cemetery = sc.locs.cemetery;
    // agents
    quarantinePolicy = sc.opt.quarantinePolicy;
    quarantineLength = sc.opt.quarantineLength;
    k = sc.opt.k;
    testingDelay = sc.opt.testingRepeatDelay;
    {
        auto pr = ps::split_double(sc.opt.testingProbabilities, ',');
        double* dst[6] = {&testingRandom, &testingHome, &testingWork, &testingSchool, &testingRandomHospital,
                          &testingNurseryHome};
        for (size_t i = 0; i < pr.size() && i < 6; ++i) *dst[i] = pr[i];
    }
    usePCR = sc.opt.testingMethod == "PCR";
    const auto& A = sc.agents;
    pp.resize(N);
    meta.resize(N);
    stats.resize(N);
    diagnosed.resize(N);
    quarantined.resize(N);
    stayedHome.assign(N, 1);
    noWorkAgent.assign(N, 0);
    location.assign(A.location.begin(), A.location.end());
    types.resize(N);
    stepsUntilMove.assign(N, 0);
    lastTest.assign(N, UMAX);
    hcw = A.healthcareWorker;
    locOffset = A.locOffset;
    possibleLocations = A.possibleLocations;
    possibleTypes = A.possibleTypes;
    for (unsigned i = 0; i < N; ++i) {
        PPState& p = pp[i];
        p.progressionID = A.progID[i];
        p.state = A.state[i];
        for (int v = 0; v < MAXS; ++v) p.susceptible[v] = 1.0f;
        updateMeta(p);
        Meta& m = meta[i];
        m.scalingAgeSex = A.scalingAgeSex[i];
        for (int j = 0; j < 7 * MAXS; ++j) m.scalingSymptoms[j] = m.scalingAgeSex;
        m.age = A.age[i];
        m.preCondIdx = A.preCond[i];
        m.sex = A.sex[i] != 0;
        Stats& st = stats[i];
        if (infectedFlag[p.state]) {
            st.infectedTimestamp = 0;
            st.infectedCount = 1;
            st.worstState = (int8_t)p.state;
            st.worstStateTimestamp = 0;
        }
        if (A.diagnosed[i]) {
            st.diagnosedTimestamp = 1;
            if (quarantinePolicy > 0) {
                st.quarantinedTimestamp = 0;
                st.quarantinedUntilTimestamp = quarantineLength * 24 * 60 / dt;
                st.daysInQuarantine = quarantineLength;
            }
        }
        diagnosed[i] = A.diagnosed[i];
        quarantined[i] = (A.diagnosed[i] && quarantinePolicy > 0) ? 1 : 0;
        types[i] = A.typeIdx[i];
    }
    healthcareWorkerCount = 0;
    for (unsigned i = 0; i < N; ++i) healthcareWorkerCount += hcw[i];
    initDays();
    locCount.assign(L, 0);
    locSum.assign((size_t)nVar * L, 0);
    ratio.assign(L, 0.f);
    variantCounts.assign(nVar, 0);
    header.clear();
    {
        std::string h = sc.statesHeader, cur;
        for (char c : h) {
            if (c == '\t') {
                if (!cur.empty()) header.push_back(cur);
                cur.clear();
            } else cur.push_back(c);
        }
        if (!cur.empty()) header.push_back(cur);
    }
    initClosures();
    initCategories();
}

// DynamicPPState(name, progressionID): daysBeforeNextState = calculateJustDays(state, 0)
void Sim::initDays() {
    for (unsigned i = 0; i < N; ++i) {
        Rng rng(seed, i, 0, PHASE_INIT_DAYS);
        pp[i].days = (int16_t)calculateDays(pp[i].progressionID, (unsigned)pp[i].state, 0, rng);
    }
}

// movementPolicies.h:69-122
// movementPolicies.h:69-122.  The reference's callers name three types whose
// first entry is taken without a draw: home, school and work, except the
// planning helpers checkUnderageAtHome and checkSchoolWorkQuarantine
// (:1073-1090, :1208-1213), which name home, school and classroom (`plan`).
unsigned Sim::findLoc(unsigned agent, unsigned lt, int retry, bool useStates, Rng* rng, bool plan) {
    uint32_t b = locOffset[agent], e = locOffset[agent + 1];
    if (lt == homeT || lt == schoolT || lt == (plan ? classroomT : workT)) {
        for (uint32_t i = b; i < e; ++i)
            if (lt == possibleTypes[i]) return possibleLocations[i];
        return UMAX;
    }
    unsigned counter = 0, pos = 0;
    for (uint32_t i = b; i < e; ++i)
        if (lt == possibleTypes[i]) {
            if (counter == 0) pos = i;
            counter++;
        }
    if (counter == 1) return possibleLocations[pos];
    if (counter > 1) {
        unsigned counter2 = counter, randIdx = counter2;
        int redraws = 0;
        while (true) {
            while (randIdx == counter2) {
                if (redraws == 65) {
                    randIdx = (counter2 + 1) % counter;
                    break;
                }
                randIdx = rng->below(counter);
                ++redraws;
            }
            counter2 = 0;
            for (uint32_t i = pos; i < e; ++i) {
                if (lt == possibleTypes[i]) {
                    if (counter2 == randIdx) {
                        if (retry && useStates && locOpen[possibleLocations[i]] == 0) {
                            retry = 0;
                            break;
                        }
                        return possibleLocations[i];
                    }
                    counter2++;
                }
            }
        }
    }
    return UMAX;
}

// movementPolicies.h:171-245 without its location writes.  An agent is
// quarantined at most once per movement routine; for a diagnosed one the
// location writes of quarantineLocations follow after the routine (so its
// classroom draw is the routine's last draw).
void Sim::quarantineAgent(unsigned i, unsigned until) {
    if (quarantinePolicy == 0) return;
    if (stats[i].diagnosedTimestamp > 0 && !isInfected(i) && quarantineImmuneActive == false) return;
    if (!diagnosed[i] && stats[i].immunizationTimestamp > 0 && quarantineImmuneActive == false) return;
    quarantined[i] = 1;
    stats[i].quarantinedTimestamp = timestamp;
    unsigned previousQuarantineUntil = stats[i].quarantinedUntilTimestamp;
    stats[i].quarantinedUntilTimestamp = until;
    stats[i].daysInQuarantine += (until - timestamp) / (24 * 60 / dt);
    if (diagnosed[i]) {
        qLoc = true;
        qLocUntil = until;
        qLocPrev = previousQuarantineUntil;
    }
}

// movementPolicies.h:202-241: the home, and with quarantinePolicy > 2 the
// classroom and workplace, of a quarantined diagnosed agent.  Location writes
// are collected in `writes`.
void Sim::quarantineLocations(unsigned i, unsigned until, unsigned previousQuarantineUntil,
                              std::vector<std::pair<unsigned, unsigned>>& writes, Rng& rng) {
    unsigned myHome = findLoc(i, homeT, 0, false, nullptr);
    if (quarantinePolicy > 1 && myHome != UMAX) writes.emplace_back(myHome, until);  // the reference writes out of bounds
    if (quarantinePolicy > (plant == 6 ? 3u : 2u) && previousQuarantineUntil < timestamp) {
        unsigned classroom = findLoc(i, classroomT, 0, false, &rng);
        unsigned work = findLoc(i, workT, 0, false, nullptr);
        unsigned toClose[2] = {classroom, work};
        for (unsigned loc : toClose) {
            if (loc != UMAX && (locType[loc] == workT || locType[loc] == classroomT) && quarantineUntil[loc] < timestamp &&
                locOpen[loc] && closedUntil[loc] < timestamp)
                writes.emplace_back(loc, until);
        }
    }
}

static inline unsigned umin(unsigned a, unsigned b) { return a < b ? a : b; }
// MIN(a, b) of datatypes.h:211 (NaN in b propagates, unlike std::min)
static inline float fmin_ref(float a, float b) { return a < b ? a : b; }

// movementPolicies.h:258-1041
void Sim::doMovement(unsigned i, unsigned day, TD simTime, std::vector<std::pair<unsigned, unsigned>>& writes) {
    Rng rng(seed, i, timestamp, PHASE_MOVE);
    qLoc = false;
    doMovementRoutine(i, day, simTime, rng);
    if (qLoc) quarantineLocations(i, qLocUntil, qLocPrev, writes, rng);
}

void Sim::doMovementRoutine(unsigned i, unsigned day, TD simTime, Rng& rng) {
    unsigned agentType = types[i];
    unsigned& sum = stepsUntilMove[i];
    Stats& st = stats[i];
    const unsigned stepsUntilMidnight = simTime.stepsUntilMidnight(dt);
    const unsigned qLen = quarantineLength * 24 * 60 / dt;
    // a missing location (UMAX: no entry of the requested type) reads as
    // closed, never quarantined and of no type; the GPU path does the same
    auto locQ = [&](unsigned loc) { return loc < L ? quarantineUntil[loc] : 0u; };
    auto openAt = [&](unsigned loc) { return loc < L ? (unsigned)locOpen[loc] : 0u; };
    auto closedAt = [&](unsigned loc) { return loc < L ? closedUntil[loc] : 0u; };

    if (enableCurfew && curfewBegin == simTime.minutes() / dt) {
        uint8_t w = wbOf(i);
        bool deadOrHosp = (w == ps::WB_D || w == ps::WB_S);
        bool hospNonCovid = st.hospitalizedTimestamp <= timestamp && st.hospitalizedUntilTimestamp > timestamp;
        if (!deadOrHosp && !hospNonCovid) {
            if (agentType + 1 == 7) {
                unsigned workplace = findLoc(i, workT, 0, false, nullptr);
                if (location[i] != workplace || openAt(workplace) == 0) {
                    location[i] = findLoc(i, homeT, 0, false, nullptr);
                    if (openAt(workplace) == 0) {
                        sum = stepsUntilMidnight;
                        return;
                    }
                }
            } else {
                sum = stepsUntilMidnight;
                location[i] = findLoc(i, homeT, 0, false, nullptr);
                return;
            }
        }
    }

    if (sum > 0) {
        sum--;
        return;
    }
    if (st.quarantinedUntilTimestamp <= timestamp) quarantined[i] = 0;
    uint8_t w = wbOf(i);
    if (w == ps::WB_D) {
        sum = UMAX;
        location[i] = cemetery;
        return;
    }
    unsigned agentHome = findLoc(i, homeT, 0, false, nullptr);

    auto diagnoseHere = [&]() {
        if (!diagnosed[i] && pp[i].infectious > 0.0f) {
            diagnosed[i] = 1;
            st.diagnosedTimestamp = timestamp;
            if (stepsUntilMidnight == 24 * 60 / dt) st.diagnosedTimestamp++;
            quarantineAgent(i, timestamp + qLen);
        }
    };

    if (st.hospitalizedTimestamp <= timestamp && st.hospitalizedUntilTimestamp > timestamp && w != ps::WB_S && w != ps::WB_D) {
        sum = umin(st.hospitalizedUntilTimestamp - timestamp - 1, stepsUntilMidnight);
        location[i] = findLoc(i, hospitalT, 0, false, &rng);
        diagnoseHere();
        return;
    }
    if (w == ps::WB_S) {
        sum = stepsUntilMidnight;
        location[i] = findLoc(i, hospitalT, 0, false, &rng);
        diagnoseHere();
        return;
    }
    auto quarantinableHere = [&](unsigned loc) {
        unsigned t = loc < L ? locType[loc] : 255u;
        return t == homeT || t == schoolT || t == classroomT || t == workT;
    };
    if (quarantinePolicy > 1 && timestamp < locQ(location[i]) && quarantinableHere(location[i])) {
        if (quarantined[i] == 0) quarantineAgent(i, locQ(location[i]));
        if (quarantined[i]) {
            sum = umin(locQ(location[i]) - timestamp - 1, stepsUntilMidnight);
            unsigned homeLocation = findLoc(i, homeT, 0, false, nullptr);
            if (homeLocation != location[i] && plant != 5) location[i] = homeLocation;
            if (agentHome != UMAX && location[i] != agentHome && sum > 60 / dt) stayedHome[i] = 0;
            return;
        }
    }
    if ((quarantinePolicy > 0 &&
         (diagnosed[i] || (st.diagnosedTimestamp > 0 && (timestamp - st.diagnosedTimestamp) < qLen))) ||
        (quarantinePolicy > 0 && quarantined[i])) {
        if (quarantinePolicy > 0 && !quarantined[i]) quarantineAgent(i, timestamp + qLen);
        if (quarantined[i]) {
            location[i] = findLoc(i, homeT, 0, false, nullptr);
            sum = stepsUntilMidnight;
            if (agentHome != UMAX && location[i] != agentHome && sum > 60 / dt) stayedHome[i] = 0;
            return;
        }
    }

    unsigned off = 7 * (unsigned)w + day + agentType * 28;
    unsigned eventsBegin = eventOffset[off], eventsEnd = eventOffset[off + 1];
    int activeBegin = -1, activeEnd = -1;
    for (unsigned j = eventsBegin; j < eventsEnd; j++) {
        if (simTime >= events[j].start && simTime < events[j].end && activeBegin == -1) activeBegin = (int)j;
        if (simTime < events[j].start) {
            activeEnd = (int)j;
            break;
        }
        if (j == eventsEnd - 1)
            if (simTime >= events[j].start && simTime < events[j].end) activeEnd = (int)eventsEnd;
    }

    unsigned newLocationType = UMAX;
    // Case 1
    if (activeBegin == -1 && activeEnd == -1) {
        newLocationType = (w == ps::WB_S) ? hospitalT : homeT;
        location[i] = findLoc(i, newLocationType, 0, false, &rng);
        sum = stepsUntilMidnight;
    }
    // Cases 2 and 4
    if (activeBegin != -1) {
        unsigned numPotential = (unsigned)((activeEnd == -1 ? activeBegin + 1 : activeEnd) - activeBegin);
        unsigned pickedIdx = 0;
        TD basicDuration;
        if (numPotential == 1) {
            newLocationType = events[activeBegin].locationType;
            basicDuration = events[activeBegin].duration;
        } else {
            double rand = rng.real(1.0);
            double thr = events[activeBegin].chance;
            unsigned idx = 0;
            while ((plant == 3 ? rand >= thr - 0.1 : rand > thr) && idx < numPotential) {
                idx++;
                if (idx < numPotential) thr += events[activeBegin + idx].chance;  // the reference also reads one past
            }
            if (idx == numPotential) idx--;
            pickedIdx = idx;
            newLocationType = events[activeBegin + idx].locationType;
            basicDuration = events[activeBegin + idx].duration;
        }
        if (noWorkAgent[i] != 0 && newLocationType == workT) {
            newLocationType = homeT;
            numPotential = 1;
            if (quarantined[i] == 0) st.daysInQuarantine++;
        }
        unsigned newLocation = findLoc(i, newLocationType, numPotential == 1, true, &rng);
        bool schoolAndTooOld =
            (newLocationType == schoolT || newLocationType == classroomT) && meta[i].age >= schoolAgeRestriction;
        auto closedHere = [&](unsigned loc, unsigned lt) {
            if (plant == 4) return false;
            return (openAt(loc) == 0 || closedAt(loc) > timestamp) && lt != workT;
        };
        if (schoolAndTooOld || closedHere(newLocation, newLocationType) ||
            (lockdownNonvaccActive && st.immunizationTimestamp == 0)) {
            if (numPotential > 1) {
                double rand = rng.real(1.0);
                double denom = 1.0 - events[activeBegin + pickedIdx].chance;
                double thr = (pickedIdx == 0) ? 0.0 : events[activeBegin].chance / denom;
                unsigned idx = 0;
                while (rand > thr && idx < numPotential) {
                    idx++;
                    if (idx < numPotential) thr += (pickedIdx == idx) ? 0.0 : events[activeBegin + idx].chance / denom;
                }
                if (idx == numPotential) idx--;
                newLocationType = events[activeBegin + idx].locationType;
                basicDuration = events[activeBegin + idx].duration;
                newLocation = findLoc(i, newLocationType, 0, false, &rng);
                if (schoolAndTooOld || closedHere(newLocation, newLocationType) ||
                    (lockdownNonvaccActive && st.immunizationTimestamp == 0)) {
                    newLocation = findLoc(i, homeT, 0, false, nullptr);
                }
            } else {
                newLocation = findLoc(i, homeT, 0, false, nullptr);
            }
        } else if (newLocationType == schoolT) {
            unsigned myClassroom = findLoc(i, classroomT, 0, false, &rng);
            if (myClassroom != UMAX && closedUntil[myClassroom] > timestamp) newLocation = findLoc(i, homeT, 0, false, nullptr);
        }
        location[i] = newLocation;
        if (basicDuration.h > 24) {
            sum = stepsUntilMidnight;
        } else if (activeEnd == -1) {
            if (simTime.plus(basicDuration).overMidnight()) sum = stepsUntilMidnight;
            else sum = basicDuration.steps(dt);
        } else {
            // activeEnd can equal eventsEnd: the reference then reads the first
            // event of the next schedule slot (a default event past the array end)
            const Event ne = (unsigned)activeEnd < events.size() ? events[activeEnd] : Event{0, -1.0f, TD(), TD(), TD()};
            TD arrive = simTime.plus(basicDuration);
            if (arrive < ne.start) sum = basicDuration.steps(dt) + (plant == 2 ? 1u : 0u);
            else if (arrive > ne.end) sum = ne.end.minus(simTime).steps(dt) - 1;
            else {
                TD window = ne.end.minus(arrive);
                unsigned stp = window.steps(dt);
                unsigned randExtra = rng.below(stp);
                sum = basicDuration.steps(dt) + randExtra;
            }
        }
        if (agentType + 1 == 7 && enableCurfew &&
            (curfewBegin <= simTime.minutes() / dt || curfewEnd > simTime.minutes() / dt)) {
            bool workButClosed =
                newLocationType == workT && (openAt(newLocation) == 0 || closedAt(newLocation) > timestamp);
            if (workButClosed || (newLocationType != workT && newLocationType != homeT)) location[i] = agentHome;
        }
    }
    // Case 3
    if (activeBegin == -1 && activeEnd != -1) {
        const Event& ne = events[activeEnd];
        TD length = ne.end.minus(ne.start);
        unsigned lengthSteps = length.steps(dt);
        unsigned randDelay = rng.below(lengthSteps);
        sum = ne.start.minus(simTime).steps(dt) + randDelay;
        unsigned timeLeft = sum;
        if (timeLeft < (plant == 1 ? 60u : 30u) / dt) {
            // stay
        } else if (timeLeft < 60 / dt) {
            newLocationType = publicT;
            location[i] = findLoc(i, publicT, 0, false, &rng);
        } else {
            newLocationType = homeT;
            location[i] = findLoc(i, homeT, 0, false, nullptr);
        }
    }
    // arrived somewhere under quarantine
    if (quarantinePolicy > 1 && timestamp < locQ(location[i]) && quarantinableHere(location[i])) {
        if (!quarantined[i]) quarantineAgent(i, locQ(location[i]));
        if (quarantined[i]) {
            sum = umin(locQ(location[i]) - timestamp - 1, stepsUntilMidnight);
            unsigned homeLocation = findLoc(i, homeT, 0, false, nullptr);
            if (homeLocation != location[i] && plant != 5) location[i] = homeLocation;
            if (agentHome != UMAX && location[i] != agentHome && sum > 60 / dt) stayedHome[i] = 0;
            return;
        }
    }
    if (newLocationType == hospitalT || (newLocationType == doctorT && rng.real(1.0) < 0.7)) diagnoseHere();
    if (agentHome != UMAX && location[i] != agentHome && sum > 60 / dt) stayedHome[i] = 0;
    sum--;
}

// DummyMovement::movement (movementPolicies.h:44-64): thrust::for_each over
// the agents; randomUnsigned(numberOfLocations), then randomUnsigned(144/4),
// then the unsigned decrement (an agent that draws 0 wraps and never moves
// again: quirk Q6).
void Sim::dummyMovement() {
    const unsigned numberOfLocations = L;  // locationListOffsets.size() - 1
    for (unsigned i = 0; i < N; ++i) {
        unsigned& s = dummyStepsUntilMove[i];
        if (s == 0) {
            Rng rng(seed, i, timestamp, PHASE_DUMMY_MOVE);
            location[i] = rng.below(numberOfLocations);
            s = rng.below(144 / 4);
        }
        s--;
    }
}

void Sim::movement() {
    if (sc.opt.movementPolicy == "none") return;  // NoMovement::movement
    if (sc.opt.movementPolicy == "dummy") {
        dummyMovement();
        return;
    }
    TD simTime = TD::fromMinutes((timestamp % stepsPerDay) * dt);
    unsigned day = holidayModeActive ? 6u : dayOfWeek();
#ifdef ORACLE_OMP
    // the reference's OpenMP path: one parallel for over the agents
    // (movementPolicies.h:1717-1719); each thread keeps its location writes
    // in agent order and the static schedule hands out ascending agent
    // ranges, so concatenating them in thread order keeps the agent order
    std::vector<std::vector<std::pair<unsigned, unsigned>>> perThread(omp_get_max_threads());
#pragma omp parallel
    {
        auto& mine = perThread[omp_get_thread_num()];
        std::vector<std::pair<unsigned, unsigned>> w;
#pragma omp for schedule(static)
        for (unsigned i = 0; i < N; ++i) {
            w.clear();
            unsigned old = location[i];
            doMovement(i, day, simTime, w);
            if (location[i] == UMAX) location[i] = old;
            mine.insert(mine.end(), w.begin(), w.end());
        }
    }
    for (auto& v : perThread)
        for (auto& x : v) quarantineUntil[x.first] = x.second;
#else
    std::vector<std::pair<unsigned, unsigned>> all, w;
    for (unsigned i = 0; i < N; ++i) {
        w.clear();
        unsigned old = location[i];
        doMovement(i, day, simTime, w);
        if (location[i] == UMAX) location[i] = old;  // no entry of the requested type: stay
        all.insert(all.end(), w.begin(), w.end());
    }
    for (auto& x : all) quarantineUntil[x.first] = x.second;  // ascending agent order: last writer wins
#endif
}

// Per-location occupancy and infectiousness sums (updatePerLocationAgentLists
// + reduce_by_location), rebuilt from scratch.
void Sim::rebuildLocationSums() {
#ifdef ORACLE_OMP
    // the reference's OpenMP structure (util.cpp:55-110, util.h:47-103):
    // agents stably sorted by location every step, offsets from the sorted
    // keys, then each location's sum taken sequentially over its agents in
    // ascending agent order (the fixed-point sum of the sequential build).
    // With one thread the reference sorts nothing (simulation.h:720-723) and
    // scatters in agent order, as below.
    if (omp_get_max_threads() > 1) {
    sortedAgents.resize(N);
    sortKey.resize(N);
#pragma omp parallel for schedule(static)
    for (unsigned i = 0; i < N; ++i) sortedAgents[i] = i;
    __gnu_parallel::stable_sort(sortedAgents.begin(), sortedAgents.end(),
                                [this](unsigned a, unsigned b) { return location[a] < location[b]; });
#pragma omp parallel for schedule(static)
    for (unsigned i = 0; i < N; ++i) sortKey[i] = location[sortedAgents[i]];
    locOffsets.assign(L + 1, 0u);
#pragma omp parallel for schedule(static)
    for (unsigned l = 0; l <= L; ++l)
        locOffsets[l] = (unsigned)(std::lower_bound(sortKey.begin(), sortKey.end(), l) - sortKey.begin());
#pragma omp parallel for schedule(dynamic, 256)
    for (unsigned l = 0; l < L; ++l) {
        locCount[l] = locOffsets[l + 1] - locOffsets[l];
        for (unsigned v = 0; v < nVar; ++v) locSum[(size_t)v * L + l] = 0;
        for (unsigned k = locOffsets[l]; k < locOffsets[l + 1]; ++k) {
            const unsigned i = sortedAgents[k];
            float x = pp[i].infectious;
            if (x != 0.0f) locSum[(size_t)pp[i].variant * L + l] += (int64_t)((double)x * 4294967296.0);
        }
    }
    return;
    }
#endif
    std::fill(locCount.begin(), locCount.end(), 0u);
    std::fill(locSum.begin(), locSum.end(), (int64_t)0);
    for (unsigned i = 0; i < N; ++i) {
        locCount[location[i]]++;
        float x = pp[i].infectious;
        if (x != 0.0f) locSum[(size_t)pp[i].variant * L + location[i]] += (int64_t)((double)x * 4294967296.0);
    }
}

// infectionPolicies.h:443-478
double Sim::seasonality() {
    unsigned simDay = timestamp / stepsPerDay;
    int d = (int)sc.opt.startDate + (int)simDay;
    int d_peak = 59 + 22;
    d += 6;
    double c0 = 3.08;
    double trunc_val = 0.5;
    if (simDay > 250 && simDay < 450) {
        d += sc.dOffset.empty() ? 0 : sc.dOffset[0];
        d_peak += sc.opt.d_peak_offset;
        c0 = sc.opt.c0;
    } else if (simDay > 600) {
        trunc_val = sc.opt.trunc;
        if (sc.dOffset.size() > 1) d += sc.dOffset[1];
    }
    int d_mod = d % 366;
    if (d_mod > 366 / 2) d_mod = 366 - d_mod;
    double normed = (0.5 * c0 * std::cos(2.0 * M_PI * (double)(d_mod - d_peak) / 366.0) + (1.0 - 0.5 * c0)) /
                    (0.5 * c0 * std::cos(2.0 * M_PI * (double)(d_peak - d_peak) / 366.0) + (1.0 - 0.5 * c0));
    double calc = std::min(std::max(normed, trunc_val), 1.0);
    if (d_mod < d_peak) return 1.0;
    return calc;
}

// infectionPolicies.h:481-558 + locationList.h:233-310
void Sim::infectionsAtLocations(unsigned variant) {
    rebuildLocationSums();
    float tmpK = (float)k;
    float season = (float)seasonality();
    const int64_t* sums = &locSum[(size_t)variant * L];
#ifdef ORACLE_OMP
#pragma omp parallel for schedule(static)
#endif
    for (unsigned l = 0; l < L; ++l) {
        float num = (float)((double)sums[l] * (1.0 / 4294967296.0));
        if (num == 0.0f) {
            ratio[l] = 0.0f;
            continue;
        }
        float density = num / (float)locCount[l];
        // location-type multiplier of the calibration interface (1 = reference)
        ratio[l] = 1.0f - exp_f32(-tmpK * density * (locInf[l] * typeMul[locType[l] & 255u]) * season * (float)dt);
    }
    if (hazardOn) {
        // expected infections and their gradient along the trajectory, all in
        // FP64 (x recomputed in double), plus the hazard sums of the probes
        for (unsigned i = 0; i < N; ++i) {
            const double s = pp[i].susceptible[variant];
            const unsigned l = location[i];
            if (!(s > 0.0) || sums[l] == 0) continue;
            const double density = (double)(float)((double)sums[l] * (1.0 / 4294967296.0)) / (double)locCount[l];
            const unsigned t = locType[l] & 255u;
            const double base = density * (double)locInf[l] * (double)season * (double)dt;
            const double m = typeMul[t];
            const double x = (double)tmpK * base * m;
            const double ex = std::exp(-x);
            hazard[0] += s * (1.0 - ex);
            hazard[1] += s * ex * base * m;
            hazard[2 + t] += s * ex * (double)tmpK * base;
            for (size_t q = 0; q < probes.size(); ++q) {
                const Probe& pr = probes[q];
                const double xp = ((double)tmpK + pr.dk) * base * (m + pr.dm[t]);
                probeJ[q] += s * (1.0 - std::exp(-xp));
            }
        }
    }
    std::vector<unsigned> newlyInfected;  // ascending agent index
#ifdef ORACLE_OMP
    const bool dump = !sc.opt.dumpLocationInfectiousList.empty();
#pragma omp parallel for schedule(static) if (!dump)
#endif
    for (unsigned i = 0; i < N; ++i) {
        PPState& p = pp[i];
        if (p.susceptible[variant] > 0.0f) {
            float rl = ratio[location[i]];
            if (rl == 0.0f) continue;  // u < 0 never holds; the draw is counter-addressed
            Rng rng(seed, i, timestamp, PHASE_INFECT + variant);
            if (rng.unit() < (double)(rl * p.susceptible[variant])) {
#ifdef ORACLE_OMP
                if (dump)
#endif
                    newlyInfected.push_back(i);
                gotInfected(p, (uint8_t)variant);
                Stats& st = stats[i];
                st.infectedTimestamp = timestamp;
                st.infectedCount++;
                st.infectedLocation = location[i];
                st.worstState = (int8_t)p.state;
                st.worstStateTimestamp = timestamp;
                st.variant |= 1u << variant;
            }
        }
    }
    if (!sc.opt.dumpLocationInfectiousList.empty()) dumpInsert(variant, newlyInfected);
}

// movementPolicies.h:1428-1622
void Sim::planLocations() {
    if (sc.opt.movementPolicy == "none") return;  // NoMovement::planLocations
    if (sc.opt.movementPolicy == "dummy") {       // DummyMovement::planLocations (:35-42)
        dummyStepsUntilMove.assign(N, 0u);
        return;
    }
    if (!curfewTimeConverted) {
        curfewBegin = curfewBegin / dt;
        curfewEnd = curfewEnd / dt;
        curfewTimeConverted = true;
    }
    std::fill(stepsUntilMove.begin(), stepsUntilMove.end(), curfewEnd);
    std::fill(stayedHome.begin(), stayedHome.end(), (uint8_t)1);
    std::fill(noWorkLoc.begin(), noWorkLoc.end(), 0u);
    std::fill(noWorkAgent.begin(), noWorkAgent.end(), (uint8_t)0);
    if (quarantinePolicy > 3) {
        unsigned numSchools = (unsigned)schools.size();
        for (unsigned s = 0; s < numSchools; ++s) {
            unsigned schoolIdx = schools[s];
            if (quarantineUntil[schoolIdx] > timestamp) continue;
            unsigned counter = 0, mx = 0;
            for (unsigned c = classroomOffsets[s]; c < classroomOffsets[s + 1]; ++c) {
                unsigned cl = classrooms[c];
                if (quarantineUntil[cl] > timestamp) {
                    counter++;
                    mx = mx > quarantineUntil[cl] ? mx : quarantineUntil[cl];
                }
            }
            if (counter > (plant == 8 ? 0u : 1u)) {
                quarantineUntil[schoolIdx] = mx;
                for (unsigned c = classroomOffsets[s]; c < classroomOffsets[s + 1]; ++c) quarantineUntil[classrooms[c]] = mx;
            }
        }
    }
    if (quarantinePolicy > 2) {
        for (unsigned i = 0; i < N; ++i) {  // checkSchoolWorkQuarantine
            if (quarantined[i]) continue;
            if (stats[i].diagnosedTimestamp > 0 && !isInfected(i)) continue;
            Rng rng(seed, i, timestamp, PHASE_PLAN_SCHOOLWORK);
            unsigned wl = findLoc(i, workT, 0, false, &rng, true);
            unsigned sl = findLoc(i, schoolT, 0, false, nullptr, true);
            unsigned cl = findLoc(i, classroomT, 0, false, nullptr, true);
            unsigned until = 0;
            if (sl != UMAX && quarantineUntil[sl] > timestamp) until = quarantineUntil[sl];
            else if (cl != UMAX && quarantineUntil[cl] > timestamp) until = quarantineUntil[cl];
            else if (wl != UMAX && quarantineUntil[wl] > timestamp) until = quarantineUntil[wl];
            if (until > 0) {
                quarantined[i] = 1;
                stats[i].quarantinedTimestamp = timestamp;
                stats[i].quarantinedUntilTimestamp = until;
                stats[i].daysInQuarantine += (until - timestamp) / (24 * 60 / dt);
            }
        }
        for (unsigned l = 0; l < L; ++l) {
            if (quarantineUntil[l] > timestamp && (locType[l] == schoolT || locType[l] == classroomT)) {
                closedUntil[l] = quarantineUntil[l];
                quarantineUntil[l] = 0;
            } else if (quarantineUntil[l] > timestamp && locType[l] == workT) {
                quarantineUntil[l] = 0;
            }
        }
    }
    for (unsigned i = 0; i < N; ++i) {  // checkUnderageAtHome
        if (meta[i].age > 14) continue;
        if (quarantined[i]) {
            unsigned h = findLoc(i, homeT, 0, false, nullptr);
            if (h != UMAX) noWorkLoc[h] = 1;
        } else {
            unsigned sl = findLoc(i, schoolT, 0, false, nullptr, true);
            unsigned cl = findLoc(i, classroomT, 0, false, nullptr, true);
            bool tooOld = meta[i].age >= schoolAgeRestriction;
            if (tooOld || (sl != UMAX && (locOpen[sl] == 0 || closedUntil[sl] > timestamp)) ||
                (cl != UMAX && (locOpen[cl] == 0 || closedUntil[cl] > timestamp))) {
                unsigned h = findLoc(i, homeT, 0, false, nullptr);
                if (h != UMAX) noWorkLoc[h] = 1;
            }
        }
    }
    for (unsigned k = 0; k < N; ++k) {  // setNoWorkToday, ascending agent order
        const unsigned i = plant == 7 ? N - 1 - k : k;
        if (meta[i].age > 26 && meta[i].age < 65) {
            uint32_t b = locOffset[i], e = locOffset[i + 1];
            unsigned h = UMAX;
            for (uint32_t j = b; j < e; ++j)
                if (possibleTypes[j] == homeT) {
                    h = possibleLocations[j];
                    break;
                }
            if (h != UMAX && noWorkLoc[h] == 1) {
                noWorkLoc[h] = 2;
                noWorkAgent[i] = 1;
            }
        }
    }
}

// testingPolicies.h:59-281, 370-475
void Sim::performTests() {
    if (sc.opt.testingPolicy == "none") return;  // NoTesting: no tests, statistics stay 0 (testingPolicies.h:11-20)
    std::fill(locFlags.begin(), locFlags.end(), (uint8_t)0);
    const unsigned day = 24 * 60 / dt;
    const unsigned qlen = quarantineLength * 24 * 60 / dt;
    auto testedOK = [&](unsigned loc) {
        unsigned q = quarantineUntil[loc];
        return q == 0 || (q != 0 && (q - qlen >= timestamp - day || q < timestamp));
    };
    for (unsigned i = 0; i < N; ++i) {  // flagLocations
        if (stats[i].diagnosedTimestamp > timestamp - day) {
            unsigned home = findLoc(i, homeT, 0, false, nullptr);
            if (home != UMAX) locFlags[home] = 1;
            unsigned work = findLoc(i, workT, 0, false, nullptr);
            if (work != UMAX && testedOK(work)) locFlags[work] = 1;
            unsigned school = findLoc(i, schoolT, 0, false, nullptr);
            if (school != UMAX && testedOK(school)) {
                locFlags[school] = 1;
                Rng rng(seed, i, timestamp, PHASE_TEST_FLAG);
                unsigned classroom = findLoc(i, classroomT, 0, false, &rng);
                if (classroom != UMAX && testedOK(classroom)) locFlags[classroom] = 1;
            }
        }
    }
    for (unsigned i = 0; i < N; ++i) {  // doTesting
        Rng rng(seed, i, timestamp, PHASE_TESTING);
        if (timestamp > testingDelay * day && lastTest[i] != UMAX && lastTest[i] > timestamp - testingDelay * day) continue;
        if (wbOf(i) == ps::WB_D || diagnosed[i]) continue;
        unsigned home = findLoc(i, homeT, 0, false, nullptr);
        bool homeFlag = home != UMAX ? locFlags[home] != 0 : false;
        unsigned work = findLoc(i, workT, 0, false, nullptr);
        bool workFlag = work != UMAX ? locFlags[work] != 0 : false;
        unsigned school = findLoc(i, schoolT, 0, false, nullptr);
        bool schoolFlag = false, classroomFlag = false;
        if (school != UMAX) {
            schoolFlag = locFlags[school] != 0;
            unsigned classroom = findLoc(i, classroomT, 0, false, &rng);
            if (classroom != UMAX) classroomFlag = true;
        }
        double prob = testingRandom;
        prob += homeFlag * testingHome;
        prob += workFlag * testingWork;
        prob += schoolFlag * testingSchool;
        prob += classroomFlag * 3.0 * testingSchool;
        if (work != UMAX && (locType[work] == doctorT || locType[work] == hospitalT)) prob += testingRandomHospital;
        if (work != UMAX && locType[work] == nurseryT) prob += testingNurseryHome;
        if (stats[i].hospitalizedTimestamp <= timestamp && stats[i].hospitalizedUntilTimestamp > timestamp)
            prob += testingRandomHospital;
        if (prob > 1.0 || rng.real(1.0) < prob) {
            lastTest[i] = timestamp;
            if (isInfected(i)) {
                float probability = usePCR ? accPCR[pp[i].state] : accAntigen[pp[i].state];
                if (probability > rng.real(1.0)) {
                    diagnosed[i] = 1;
                    stats[i].diagnosedTimestamp = timestamp;
                }
            } else {
                if (stats[i].quarantinedUntilTimestamp > timestamp && (home != UMAX && quarantineUntil[home] < timestamp)) {
                    if (stats[i].daysInQuarantine > 0)
                        stats[i].daysInQuarantine -= (stats[i].quarantinedUntilTimestamp - timestamp) / day;
                    stats[i].quarantinedUntilTimestamp = timestamp;
                }
            }
        }
    }
    unsigned tests = 0, pos1 = 0, pos2 = 0;
    for (unsigned i = 0; i < N; ++i) {
        tests += lastTest[i] == timestamp;
        pos1 += stats[i].diagnosedTimestamp == timestamp;
        pos2 += stats[i].diagnosedTimestamp < timestamp && stats[i].diagnosedTimestamp > timestamp - day;
    }
    testStats[0] = tests;
    testStats[1] = pos1;
    testStats[2] = pos2;
}

// simulation.h:425-471 + dynamicPPState.cpp:246-303 + multiBadMatrix.cpp:120-125
void Sim::updateAgents() {
#ifdef ORACLE_OMP
#pragma omp parallel for schedule(static)
#endif
    for (unsigned i = 0; i < N; ++i) {
        PPState& p = pp[i];
        Meta& m = meta[i];
        Stats& st = stats[i];
        float progS;
        if (p.state == 7) progS = deathScaling[p.variant] / m.get(p.variant, (unsigned)p.state);
        else progS = progScaling[p.variant];
        float scaling = m.get(p.variant, (unsigned)p.state) * progS;
        Rng rng(seed, i, timestamp, PHASE_PROGRESSION);
        bool recovered = false;
        if (p.days == -2) {
            p.days = (int16_t)calculateDays(p.progressionID, (unsigned)p.state, p.variant, rng);
            if (p.variant >= 3) {
                double d = p.days * 0.7;
                p.days = (int16_t)(d == 0 ? 1.0 : d);
            }
        } else if (p.days > 0) {
            --p.days;
            continue;
        }
        if (p.days == 0) {
            int oldState = p.state;
            bool wasInfected = infectedFlag[p.state] != 0;
            auto nx = selectNext(p.progressionID, (unsigned)p.state, scaling, rng);
            int nd = calculateDays(p.progressionID, nx.first, p.variant, rng);
            p.state = (int)nx.first;
            updateMeta(p);
            p.days = (int16_t)nd;
            if (p.variant >= 3 && p.state < 8) {
                float mul = p.state < 6 ? 0.7f : 0.5f;
                float d = p.days * mul;
                p.days = (int16_t)(d == 0 ? 1.0f : d);
            }
            if (nx.second) {
                st.worstState = (int8_t)p.state;
                st.worstStateTimestamp = timestamp;
                st.worstStateEndTimestamp = timestamp + (unsigned)(p.days * 60 * 24) / dt;
            } else if (oldState != p.state) {
                if (wasInfected && !infectedFlag[p.state]) {
                    st.recoveredTimestamp = timestamp;
                    recovered = true;
                }
            }
        }
        if (recovered) diagnosed[i] = 0;
    }
}

// immunization.h:481-695 (susceptibility of state-0 agents).  Vaccination and
// booster rollout (immunization.h:697-900) start on day immunizationStart /
// boosterStart; the oracle refuses to run into them.
void Sim::immunizationUpdate() {
    unsigned ts = timestamp == 0 ? 1u : timestamp;
    auto pi = ps::split_float(sc.opt.protectionInfection, ',');
    auto piw = ps::split_float(sc.opt.protectionInfectionWaning, ',');
    auto psy = ps::split_float(sc.opt.protectionSymptomatic, ',');
    auto psyw = ps::split_float(sc.opt.protectionSymptomaticWaning, ',');
    auto ph = ps::split_float(sc.opt.protectionHospitalization, ',');
    auto phw = ps::split_float(sc.opt.protectionHospitalizationWaning, ',');
    auto vsim = ps::split_int(sc.opt.variantSimilarity, ',');
    auto vsm = ps::split_float(sc.opt.variantSimilarMultiplier, ',');
    if (pi.size() < 10 || piw.size() < 10 || psy.size() < 10 || psyw.size() < 10 || ph.size() < 10 || phw.size() < 10)
        throw ps::InitError(ps::ERR_BAD_ARG, "protection parameters need 10 values");
    // efficiencyInf etc.: pairs (waning, level) for {inf, sympt, hosp} x {similar, different}
    float effInf[6][2], eff2[6][2], eff3[6][2], effI2[6][2], effI3[6][2];
    auto fill = [&](float e[6][2], int base) {
        for (int s = 0; s < 2; ++s) {
            e[s * 3 + 0][0] = piw[base + s];
            e[s * 3 + 0][1] = pi[base + s];
            e[s * 3 + 1][0] = psyw[base + s];
            e[s * 3 + 1][1] = psy[base + s];
            e[s * 3 + 2][0] = phw[base + s];
            e[s * 3 + 2][1] = ph[base + s];
        }
    };
    fill(effInf, 0);
    fill(eff2, 2);
    fill(eff3, 4);
    fill(effI2, 6);
    fill(effI3, 8);
    int simLocal[MAXS] = {0};
    float simMul[MAXS] = {0};
    for (size_t i = 0; i < vsim.size() && i < (size_t)MAXS; ++i) simLocal[i] = vsim[i];
    for (size_t i = 0; i < vsm.size() && i < (size_t)MAXS; ++i) simMul[i] = vsm[i];
    const unsigned week = 7 * 24 * 60 / dt;
    for (unsigned a = 0; a < N; ++a) {
        PPState& p = pp[a];
        if (p.state != 0) continue;
        Stats& st = stats[a];
        Meta& m = meta[a];
        float susc[MAXS];
        for (unsigned v = 0; v < nVar; ++v) susc[v] = 1.0f;
        unsigned wImm = (ts - st.immunizationTimestamp) / week;
        unsigned wInf = (ts - st.infectedTimestamp) / week;
        auto lin = [&](float e[6][2], int row, unsigned weeks) -> float {  // -w*weeks + level, all float
            return -1.0f * e[row][0] * (float)weeks + e[row][1];
        };
        auto mix = [&](float e[6][2], int comp, int similar, unsigned v, unsigned weeks) -> float {
            // -(w_d + (w_s - w_d)*m)*weeks + l_d + (l_s - l_d)*m
            int d = (1 - similar) * 3 + comp, s = similar * 3 + comp;
            return -1.0f * (e[d][0] + (e[s][0] - e[d][0]) * simMul[v]) * (float)weeks + e[d][1] +
                   (e[s][1] - e[d][1]) * simMul[v];
        };
        if (st.infectedCount > 0 && (st.immunizationCount == 0 || (st.immunizationCount == 1 && wImm < 4))) {
            unsigned prior = st.variant;
            for (unsigned v = 0; v < nVar; ++v) {
                int similar = 0;
                for (unsigned j = 0; j < nVar; ++j) similar = similar || ((prior & (1u << j)) && simLocal[v] == simLocal[j]);
                int d = (1 - similar) * 3;
                float vsInf = (float)((double)(-1.0f * effInf[d + 0][0] * (float)wInf) +
                                      (((1u << v) & prior) ? 0.95 : (double)effInf[d + 0][1]));
                bool mixIt = similar && !((1u << v) & prior) && simMul[v] > 0.0f;
                if (mixIt) vsInf = mix(effInf, 0, similar, v, wInf);
                susc[v] = 1.0f - vsInf;
                float vsSympt = lin(effInf, d + 1, wInf);
                if (mixIt) vsSympt = mix(effInf, 1, similar, v, wInf);
                float vsHosp = lin(effInf, d + 2, wInf);
                if (mixIt) vsHosp = mix(effInf, 2, similar, v, wInf);
                m.set(fmin_ref(1.0f, (1.0f - vsSympt) / (1.0f - vsInf)), 2, v);
                m.set(fmin_ref(1.0f, (1.0f - vsHosp) / (1.0f - vsSympt)), 4, v);
                m.set(fmin_ref(1.0f, (1.0f - vsHosp) / (1.0f - vsSympt)), 5, v);
            }
        } else if (st.infectedCount == 0 && st.immunizationCount > 0) {
            for (unsigned v = 0; v < nVar; ++v) {
                int similar = simLocal[v] == 0;
                int d = (1 - similar) * 3;
                if (st.immunizationCount == 1 && wImm == 0) continue;
                if (st.immunizationCount == 1 && wImm < 4) {
                    float vsInf = (-1.0f * eff2[d + 0][0] * (float)wImm + eff2[d + 0][1]) / 2.0f;
                    susc[v] = 1.0f - vsInf;
                    float vsSympt = (-1.0f * eff2[d + 1][0] * (float)wImm + eff2[d + 1][1]) / 2.0f;
                    float vsHosp = (-1.0f * eff2[d + 2][0] * (float)wImm + eff2[d + 2][1]) / 2.0f;
                    m.set((1.0f - vsSympt) / (1.0f - vsInf), 2, v);
                    m.set((1.0f - vsHosp) / (1.0f - vsSympt), 4, v);
                    m.set((1.0f - vsHosp) / (1.0f - vsSympt), 5, v);
                } else {
                    float(*e)[2] = st.immunizationCount == 1 ? eff2 : eff3;
                    bool mixIt = similar && simMul[v] > 0.0f;
                    float vsInf = mixIt ? mix(e, 0, similar, v, wImm) : lin(e, d + 0, wImm);
                    susc[v] = 1.0f - vsInf;
                    float vsSympt = mixIt ? mix(e, 1, similar, v, wImm) : lin(e, d + 1, wImm);
                    float vsHosp = mixIt ? mix(e, 2, similar, v, wImm) : lin(e, d + 2, wImm);
                    m.set((1.0f - vsSympt) / (1.0f - vsInf), 2, v);
                    m.set((1.0f - vsHosp) / (1.0f - vsSympt), 4, v);
                    m.set((1.0f - vsHosp) / (1.0f - vsSympt), 5, v);
                }
            }
        } else if (st.infectedCount > 0 && st.immunizationCount > 0) {
            unsigned prior = st.variant;
            unsigned weeks = std::min(wInf, wImm);
            for (unsigned v = 0; v < nVar; ++v) {
                int similar = 0;
                for (unsigned j = 0; j < nVar; ++j) similar = similar || ((prior & (1u << j)) && simLocal[v] == simLocal[j]);
                int d = (1 - similar) * 3;
                float(*e)[2] = st.immunizationCount == 1 ? effI2 : effI3;
                if (st.immunizationCount < 1) continue;
                bool mixIt = similar && !((1u << v) & prior) && simMul[v] > 0.0f;
                float vsInf = (float)((double)(-1.0f * e[d + 0][0] * (float)weeks) +
                                      (((1u << v) & prior) ? 0.95 : (double)e[d + 0][1]));
                if (mixIt) vsInf = mix(e, 0, similar, v, weeks);
                susc[v] = 1.0f - vsInf;
                float vsSympt = mixIt ? mix(e, 1, similar, v, weeks) : lin(e, d + 1, weeks);
                float vsHosp = mixIt ? mix(e, 2, similar, v, weeks) : lin(e, d + 2, weeks);
                m.set(fmin_ref(1.0f, (1.0f - vsSympt) / (1.0f - vsInf)), 2, v);
                m.set(fmin_ref(1.0f, (1.0f - vsHosp) / (1.0f - vsSympt)), 4, v);
                m.set(fmin_ref(1.0f, (1.0f - vsHosp) / (1.0f - vsSympt)), 5, v);
            }
        }
        for (unsigned v = 0; v < nVar; ++v)
            if (p.state == 0) p.susceptible[v] = fmin_ref(1.0f, susc[v]);
    }
    vaccinationRollout();
}

// The ten vaccination categories of initCategories (immunization.h:229-339),
// indexed as the reference's pass loop indexes them (groupIdx).  Location-type
// ids 2, 3, 4, 12, 14, 22 are literals in the reference.
bool Sim::inVaccinationCategory(unsigned i, unsigned g) const {
    const unsigned age = meta[i].age, pre = meta[i].preCondIdx;
    auto anyEntry = [&](auto pred) {
        for (uint32_t e = locOffset[i]; e < locOffset[i + 1]; ++e)
            if (pred((unsigned)possibleTypes[e], possibleLocations[e])) return true;
        return false;
    };
    switch (g) {
        case 0: return anyEntry([&](unsigned t, unsigned l) { return t == 4 && (locType[l] == 12 || locType[l] == 14); });
        case 1: return anyEntry([&](unsigned t, unsigned l) { return (t == 4 || t == 2) && locType[l] == 22; });
        case 2: return age >= 60;
        case 3: return pre > 0 && age >= 18 && age < 60;
        case 4: return anyEntry([&](unsigned t, unsigned l) { return t == 4 && essential[l] == 1; });
        case 5: return age > 17 && age < 60;
        case 6: return pre > 0 && age >= 60;
        case 7: return anyEntry([&](unsigned t, unsigned l) { return t == 4 && locType[l] == 3; });
        case 8: return age >= 12 && age < 18;
        case 9: return age >= 5 && age < 12;
    }
    return false;
}

// immunization.h:229-420: every agent's category round (its category's
// position in immunizationOrder) or 255 when it declines.  Groups are visited
// in ascending order value, ties by ascending index; an agent is decided by
// the first visited group it belongs to (one draw).  The level of group g is
// vaccinationGroupLevel[{0,1,3,4,5,7,2,6,8,9}[g]], as the reference's lambdas
// capture it.
void Sim::initCategories() {
    static const int levelOf[10] = {0, 1, 3, 4, 5, 7, 2, 6, 8, 9};
    const auto& v = sc.vacc;
    immRound.assign(N, 0);
    std::vector<unsigned> passes;
    for (int o = 1; o <= 10; ++o)
        for (unsigned g = 0; g < 10; ++g)
            if (v.order[g] == o) passes.push_back(g);
    for (unsigned i = 0; i < N; ++i) {
        Rng rng(seed, i, 0, PHASE_IMM_CATEGORY);
        for (unsigned g : passes) {
            if (!inVaccinationCategory(i, g)) continue;
            immRound[i] = rng.unit() < (double)v.groupLevel[levelOf[g]] ? (uint8_t)v.order[g] : (uint8_t)255;
            break;
        }
    }
}

static inline int ageGroupOf(unsigned age) { return age < 90 ? (int)(age / 10) : (age < 200 ? 9 : -1); }

// immunization.h:421-479
void Sim::initAgeGroups() {
    ageGroupsReady = true;
    const auto& v = sc.vacc;
    const size_t rounds = v.boosterStart.size();
    ageGroupSize.assign(rounds, {});
    ageGroupFrac.assign(rounds, {});
    std::array<unsigned, 10> ages{};
    for (unsigned i = 0; i < N; ++i) {
        int g = ageGroupOf(meta[i].age);
        if (g >= 0) ages[g]++;
    }
    for (size_t r = 0; r < rounds; ++r) {
        unsigned sum = 0;
        for (int g = 0; g < 10; ++g) {
            ageGroupSize[r][g] = (unsigned)((float)ages[g] * v.boosterAgePercent[r][g]);
            sum += ageGroupSize[r][g];
        }
        for (int g = 0; g < 10; ++g) ageGroupFrac[r][g] = (float)ageGroupSize[r][g] / (float)sum;
    }
}

// immunization.h:696-900: booster rounds, then first doses category by
// category, then immunizedToday = agents immunized at this timestamp.
void Sim::vaccinationRollout() {
    const auto& v = sc.vacc;
    const unsigned spd = stepsPerDay;
    const unsigned ts = timestamp == 0 ? 1u : timestamp;
    const unsigned day = timestamp / spd;
    if (!ageGroupsReady) initAgeGroups();
    auto isW = [&](unsigned i) { return wbOf(i) == ps::WB_W; };
    auto boosterEligible = [&](unsigned i) {
        const Stats& s = stats[i];
        return s.immunizationTimestamp > 0 && s.immunizationCount >= 1 && (ts - s.immunizationTimestamp) / spd > 4 * 30 &&
               isW(i) && (ts >= spd * 3 * 30 && s.diagnosedTimestamp < ts - spd * 4 * 30);
    };
    for (size_t r = 0; r < v.boosterStart.size(); ++r) {
        unsigned available = 0;
        if ((int)day >= v.boosterStart[r]) {
            if (v.boosterPerWeek[r].empty()) available = v.dailyBoosters;
            else {
                unsigned week = (day - (unsigned)v.boosterStart[r]) / 7;
                const auto& w = v.boosterPerWeek[r];
                available = (unsigned)(w[week >= w.size() ? w.size() - 1 : week] / 7.0);
            }
        }
        if (!available) continue;
        std::array<unsigned, 10> counts{};
        for (unsigned i = 0; i < N; ++i) {
            int g = ageGroupOf(meta[i].age);
            if (g >= 0 && boosterEligible(i)) counts[g]++;
        }
        for (int g = 0; g < 10; ++g) {
            unsigned availG = (unsigned)((float)available * ageGroupFrac[r][g]);
            float prob = counts[g] < availG ? 1.0f : (float)availG / (float)counts[g];
            for (unsigned i = 0; i < N; ++i) {
                if (ageGroupOf(meta[i].age) != g || !boosterEligible(i)) continue;
                Rng rng(seed, i, timestamp, PHASE_VACCINATION);
                rng.seek((uint32_t)r);
                if (prob == 1.0f || rng.unit() < (double)prob) {
                    stats[i].immunizationTimestamp = ts;
                    stats[i].immunizationCount++;
                }
            }
        }
    }
    unsigned available = 0;
    if (day >= v.startDay) {
        if (v.vaccPerWeek.empty()) available = v.dailyDoses;
        else {
            unsigned week = (day - v.startDay) / 7;
            available = (unsigned)(v.vaccPerWeek[week >= v.vaccPerWeek.size() ? v.vaccPerWeek.size() - 1 : week] / 7.0);
        }
    }
    auto vaccEligible = [&](unsigned i) {
        const Stats& s = stats[i];
        return s.immunizationCount == 0 && isW(i) &&
               ((ts < spd * 3 * 30 && s.diagnosedTimestamp == 0) ||
                (ts >= spd * 3 * 30 && s.diagnosedTimestamp < ts - spd * 3 * 30));
    };
    while (available > 0 && currentCategory < 10) {
        unsigned count = 0;
        while (count == 0 && currentCategory <= 10) {
            const unsigned cc = currentCategory + 1;  // rounds start at 1; 11 counts every round
            count = 0;
            for (unsigned i = 0; i < N; ++i) count += (immRound[i] == cc || cc == 11) && vaccEligible(i);
            if (count == 0) currentCategory++;
        }
        float prob = count < available ? 1.0f : (float)available / (float)count;
        const unsigned cc = currentCategory + 1;
        for (unsigned i = 0; i < N; ++i) {
            if (immRound[i] != cc || !vaccEligible(i)) continue;
            Rng rng(seed, i, timestamp, PHASE_VACCINATION);
            rng.seek(32u + currentCategory);
            if (prob == 1.0f || rng.unit() < (double)prob) {
                stats[i].immunizationTimestamp = ts;
                stats[i].immunizationCount = 1;
            }
        }
        available = count < available ? available - count : 0;
    }
    immunizedToday = 0;
    for (unsigned i = 0; i < N; ++i) immunizedToday += stats[i].immunizationTimestamp == ts;
}

// simulation.h:169-423
void Sim::otherDisease() {
    static const double randomHospProbs[] = {
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
    static const double avgLengths[] = {5.55, 2.78, 5.24};
    unsigned inHospitalWithCovid = 0;
    if (sc.opt.dueWithCOVID == 0) {
        for (unsigned i = 0; i < N; ++i) inHospitalWithCovid += (pp[i].state >= 6 && pp[i].state <= 8);
    } else {
        for (unsigned i = 0; i < N; ++i)
            inHospitalWithCovid += isInfected(i) && locType[location[i]] == hospitalT &&
                                   !(wbOf(i) < ps::WB_S && timestamp > 0 && stats[i].hospitalizedUntilTimestamp == timestamp);
    }
    float probabilityMul = 0.0f;
    unsigned simDay = timestamp / stepsPerDay;
    if (simDay < sc.totalHospitalizations.size()) {
        unsigned target = (unsigned)std::max(0.0, double(sc.totalHospitalizations[simDay]) -
                                                      double(inHospitalWithCovid) * (9730000.0 / 179500.0));
        probabilityMul = float(target) / 48000.0;
    }
    for (unsigned i = 0; i < N; ++i) {
        PPState& p = pp[i];
        Stats& st = stats[i];
        if (wbOf(i) == ps::WB_D) continue;
        Rng rng(seed, i, timestamp, PHASE_OTHER_DISEASE);
        uint8_t age = meta[i].age;
        unsigned ageGroup = age < 5 ? 0 : age < 15 ? 1 : age < 30 ? 2 : age < 60 ? 3 : age < 70 ? 4 : age < 80 ? 5 : 6;
        bool sex = meta[i].sex;
        unsigned type = meta[i].preCondIdx == 2 ? 1 : (meta[i].preCondIdx == 4 ? 2 : 0);
        if (timestamp > 0 && st.hospitalizedUntilTimestamp == timestamp) {
            if (rng.real(1.0) < randomHospProbs[type * 4 * 7 + 2 * 7 + !sex * 7 + ageGroup]) {
                st.worstState = (int8_t)die(p, false);
                st.worstStateTimestamp = timestamp;
                continue;
            }
        }
        if (wbOf(i) == ps::WB_D || wbOf(i) == ps::WB_S || timestamp < st.hospitalizedUntilTimestamp) continue;
        (void)rng.real(1.0);  // sudden death: disabled in the reference, but its draw is made
        static const int multDay[] = {99, 99 + 365, 99 + 2 * 365, 99 + 3 * 365};
        static const double mult[] = {0.82, 0.67, 0.9, 0.9};
        int day = (int)(timestamp / (24 * 60 / dt));
        int d = 0;
        while (d < 3 && day > multDay[d]) d++;
        double probability =
            randomHospProbs[type * 4 * 7 + !sex * 7 + ageGroup] * (probabilityMul > 0.0f ? probabilityMul : mult[d]);
        if (rng.real(1.0) < probability) {
            unsigned avgLength = (unsigned)avgLengths[type];
            double pgeo = 1.0 / (double)avgLength;
            unsigned length = (unsigned)geometric(pgeo, rng);
            if (length == 0) length = 1;
            st.hospitalizedTimestamp = timestamp;
            st.hospitalizedUntilTimestamp = timestamp + length * 24 * 60 / dt;
        }
    }
}

// simulation.h:473-665 (+ statistics.h:28-57, agentsList.h:55-78)
std::vector<unsigned> Sim::refreshStatistics() {
    std::vector<unsigned> out(nStates, 0);
    for (unsigned i = 0; i < N; ++i) out[(unsigned)pp[i].state]++;
    unsigned hospitalized = 0, stayed = 0;
    for (unsigned i = 0; i < N; ++i) {
        uint8_t w = wbOf(i);
        hospitalized += (w != ps::WB_D && w != ps::WB_S && timestamp < stats[i].hospitalizedUntilTimestamp);
        stayed += stayedHome[i];
    }
    out.push_back(hospitalized);
    out.push_back(testStats[0]);
    out.push_back(testStats[1]);
    out.push_back(testStats[2]);
    unsigned q = 0, qt = 0, nq = 0;
    for (unsigned i = 0; i < N; ++i) {
        bool isQ = stats[i].quarantinedTimestamp <= timestamp && stats[i].quarantinedUntilTimestamp > timestamp;
        bool inf = isInfected(i);
        q += isQ;
        qt += isQ && inf;
        nq += !isQ && inf;
    }
    out.push_back(q);
    out.push_back(qt);
    out.push_back(nq);
    if (nVar > 1) {
        unsigned all = 0;
        for (unsigned i = 0; i < N; ++i) all += isInfected(i);
        for (unsigned v = 0; v + 1 < nVar; ++v) {
            unsigned c = 0;
            for (unsigned i = 0; i < N; ++i) c += isInfected(i) && pp[i].variant == v + 1;
            double pct = double(c) / double(all) * 100.0;
            out.push_back(pct != pct ? 0u : (unsigned)(int64_t)pct);
        }
    } else {
        out.push_back(0);
    }
    out.push_back(stayed - out[10] - out[11]);
    out.push_back(immunizedToday);
    const unsigned day = 24 * 60 / dt;
    unsigned ni = 0, inf = 0, reinf = 0, bstr = 0;
    for (unsigned i = 0; i < N; ++i) {
        ni += stats[i].infectedTimestamp > timestamp - day && stats[i].infectedTimestamp <= timestamp;
        inf += stats[i].infectedCount > 0;
        reinf += stats[i].infectedCount > 1 ? stats[i].infectedCount - 1u : 0u;
        bstr += stats[i].immunizationCount > 1 ? stats[i].immunizationCount - 1u : 0u;
    }
    out.push_back(ni);
    out.push_back(inf);
    out.push_back(reinf);
    out.push_back(bstr);
    for (unsigned v = 0; v < nVar; ++v) {
        int64_t acc = 0;  // 2^-24 fixed point: every term 1 - s is a multiple of 2^-24
        for (unsigned i = 0; i < N; ++i) acc += (int64_t)((double)(1.0f - pp[i].susceptible[v]) * 16777216.0);
        out.push_back((unsigned)(acc >> 24));
    }
    unsigned hci = 0, hce = 0;
    for (unsigned i = 0; i < N; ++i) {
        if (!hcw[i]) continue;
        hci += isInfected(i) || quarantined[i];
        hce += stats[i].infectedTimestamp > timestamp - day && stats[i].infectedTimestamp <= timestamp;
    }
    out.push_back((hci * 100) / std::max(1u, healthcareWorkerCount));
    out.push_back(hce);
    for (unsigned v = 0; v < nVar; ++v) {
        unsigned c = 0;
        for (unsigned i = 0; i < N; ++i) c += ((1u << v) & stats[i].variant) ? 1 : 0;
        out.push_back(c);
    }
    unsigned infh = 0, vni = 0;
    for (unsigned i = 0; i < N; ++i) {
        infh += isInfected(i) && locType[location[i]] == hospitalT;
        vni += stats[i].infectedCount == 0 && stats[i].immunizationCount > 0;
    }
    out.push_back(infh);
    out.push_back(vni);
    return out;
}

void Sim::countVariantCases() {
    for (unsigned v = 0; v < nVar; ++v) {
        unsigned c = 0;
        for (unsigned i = 0; i < N; ++i) c += isInfected(i) && pp[i].variant == v;
        variantCounts[v] = c;
    }
}

// closurePolicies.h:84-479
void Sim::initClosures() {
    if (sc.opt.closurePolicy == "none") return;  // NoClosure::init reads no rules (closurePolicies.h:41)
    unsigned enable = sc.opt.enableClosures;
    std::vector<unsigned> hospitalStates;
    for (unsigned i = 0; i < header.size(); ++i)
        if (header[i].find("_h") != std::string::npos) hospitalStates.push_back(i);
    const unsigned nAgents = N;
    static const char* alwaysOn[] = {"Masks",           "Curfew",           "HolidayMode", "SchoolAgeRestriction",
                                     "ExposeToMutation", "ReduceMovement",   "QuarantinePolicy",
                                     "QuarantineImmune", "LockdownNonvacc",  "TestingProb"};
    conds.reserve(sc.rules.size());
    rules.reserve(sc.rules.size());
    for (const auto& r : sc.rules) {
        bool keep = enable != 0;
        for (auto* a : alwaysOn) keep = keep || r.name == a;
        if (!keep) continue;
        Cond c;
        int closeAfter = r.closeAfter, openAfter = r.openAfter;
        double threshold = r.threshold, threshold2 = r.threshold2;
        auto hysteresis = [threshold, threshold2, openAfter, closeAfter](Cond* c, double value) {
            c->history[c->pos] = value;
            c->pos = (c->pos + 1) % (int)c->history.size();
            int hs = (int)c->history.size();
            auto mod = [](int a, int b) {
                int x = a % b;
                return x < 0 ? x + b : x;
            };
            if (c->active == false && value > threshold) {
                bool wasBelow = false;
                for (unsigned i = (unsigned)mod(c->pos - closeAfter, hs); i != (unsigned)c->pos; i = (i + 1) % hs)
                    wasBelow |= c->history[i] < threshold;
                return !wasBelow ? true : c->active;
            } else if (c->active == true && value <= threshold2) {
                bool wasAbove = false;
                for (unsigned i = (unsigned)mod(c->pos - openAfter, hs); i != (unsigned)c->pos; i = (i + 1) % hs)
                    wasAbove |= c->history[i] >= threshold2;
                return !wasAbove ? false : c->active;
            }
            return c->active;
        };
        if (r.conditionType == "afterDays") {
            if (closeAfter >= 0 || openAfter <= 0)
                throw ps::InitError(ps::ERR_BAD_CLOSURE, "For closure rule 'afterDays', closeAfter must be -1, openAfter must be >0");
            c.condition = [openAfter, threshold](Cond* c, std::vector<unsigned>&) {
                int day = (int)c->history[0]++;
                return day >= threshold && day < threshold + openAfter;
            };
            c.history.assign(1, 0.0);
        } else if (r.conditionType == "hospitalizedFraction") {
            if (closeAfter <= 0 || openAfter <= 0)
                throw ps::InitError(ps::ERR_BAD_CLOSURE, "For closure rule 'hospitalizedFraction', closeAfter and openAfter must be >0");
            c.headerPos = hospitalStates;
            c.condition = [hysteresis, nAgents](Cond* c, std::vector<unsigned>& s) {
                double accum = 0.0;
                for (unsigned i = 0; i < c->headerPos.size(); i++) accum += s[c->headerPos[i]];
                return hysteresis(c, accum / (double)nAgents);
            };
            c.history.assign(std::max(closeAfter, openAfter) + 2, 0.0);
        } else if (r.conditionType == "newDeadFraction") {
            if (closeAfter <= 0 || openAfter <= 0)
                throw ps::InitError(ps::ERR_BAD_CLOSURE, "For closure rule 'newDeadFraction', closeAfter and openAfter must be >0");
            for (unsigned i = 0; i < header.size(); i++)
                if (header[i] == "D1") c.headerPos.push_back(i);
            c.headerPos.push_back(0);
            c.condition = [hysteresis, nAgents](Cond* c, std::vector<unsigned>& s) {
                double value = (s[c->headerPos[0]] - c->headerPos[1]) / (double)nAgents;
                c->headerPos[1] = s[c->headerPos[0]];
                return hysteresis(c, value);
            };
            c.history.assign(std::max(closeAfter, openAfter) + 2, 0.0);
        } else {
            throw ps::InitError(ps::ERR_BAD_CLOSURE, "Unknown closure type " + r.conditionType);
        }
        conds.push_back(c);
        size_t ci = conds.size() - 1;
        Rule rule;
        rule.name = r.name;
        rule.conds = {ci};
        auto closeNow = [this](Rule* rr) {
            bool close = true;
            for (size_t c : rr->conds) close = close && conds[c].active;
            return close;
        };
        if (r.name == "Masks") {
            double coef = std::atof(r.parameter.c_str());
            rule.rule = [this, closeNow, coef](Rule* rr) {
                bool shouldBeOpen = !closeNow(rr);
                if (rr->previousOpenState != shouldBeOpen) {
                    currentMaskValue = shouldBeOpen ? 1.0 : coef;
                    for (unsigned l = 0; l < L; ++l)
                        if (locType[l] != homeT)
                            locInf[l] = shouldBeOpen ? (float)(locInf[l] / coef) : (float)(locInf[l] * coef);
                    rr->previousOpenState = shouldBeOpen;
                }
            };
        } else if (r.name == "Curfew") {
            unsigned cb = 0, ce = 0;
            if (r.parameter.size() == 9) {
                cb = (unsigned)std::atoi(r.parameter.substr(0, 2).c_str()) * 60 + (unsigned)std::atoi(r.parameter.substr(2, 2).c_str());
                ce = (unsigned)std::atoi(r.parameter.substr(5, 2).c_str()) * 60 + (unsigned)std::atoi(r.parameter.substr(7, 2).c_str());
            } else if (!r.parameter.empty())
                throw ps::InitError(ps::ERR_BAD_CLOSURE, "curfew parameter string needs to be exactly 9 characters long");
            rule.rule = [this, closeNow, cb, ce](Rule* rr) {
                bool close = closeNow(rr);
                if (rr->previousOpenState != !close) {
                    enableCurfew = close;
                    curfewBegin = cb;
                    curfewEnd = ce;
                    curfewTimeConverted = false;
                    rr->previousOpenState = !close;
                }
            };
        } else if (r.name == "HolidayMode" || r.name == "QuarantineImmune" || r.name == "LockdownNonvacc") {
            bool* target = r.name == "HolidayMode" ? &holidayModeActive
                                                   : (r.name == "QuarantineImmune" ? &quarantineImmuneActive : &lockdownNonvaccActive);
            rule.rule = [closeNow, target](Rule* rr) {
                bool close = closeNow(rr);
                if (rr->previousOpenState != !close) {
                    *target = close;
                    rr->previousOpenState = !close;
                }
            };
        } else if (r.name == "QuarantinePolicy") {
            int qp = std::atoi(r.parameter.c_str());
            rule.rule = [this, closeNow, qp](Rule* rr) {
                bool close = closeNow(rr);
                if (close) {
                    quarantinePolicy = (unsigned)qp;
                    rr->previousOpenState = close;
                }
            };
        } else if (r.name == "SchoolAgeRestriction") {
            unsigned limit = (unsigned)std::atoi(r.parameter.c_str());
            rule.rule = [this, closeNow, limit](Rule* rr) {
                bool close = closeNow(rr);
                if (rr->previousOpenState != !close) {
                    schoolAgeRestriction = close ? limit : 99;
                    rr->previousOpenState = !close;
                }
            };
        } else if (r.name == "ExposeToMutation") {
            double fraction = std::atof(r.parameter.c_str());
            unsigned variantIdx = (unsigned)(int)threshold2;
            if (variantIdx >= nVar)
                throw ps::InitError(ps::ERR_BAD_CLOSURE, "More ExposeToMutation events defined than infectiousnessMultiplier specified on the command line");
            rule.rule = [this, closeNow, fraction, variantIdx](Rule* rr) {
                if (!closeNow(rr)) return;
                for (unsigned i = 0; i < N; ++i) {
                    PPState& p = pp[i];
                    if (p.susceptible[variantIdx] > 0.0f) {
                        Rng rng(seed, i, timestamp, PHASE_MUTATION);
                        if (rng.unit() < fraction * p.susceptible[variantIdx]) {
                            gotInfected(p, (uint8_t)variantIdx);
                            Stats& st = stats[i];
                            st.infectedTimestamp = timestamp;
                            st.infectedLocation = location[i];
                            st.infectedCount++;
                            st.worstState = (int8_t)p.state;
                            st.worstStateTimestamp = timestamp;
                            st.variant |= 1u << variantIdx;
                        }
                    }
                }
            };
        } else if (r.name == "ReduceMovement") {
            double fraction = std::atof(r.parameter.c_str());
            rule.rule = [this, closeNow, fraction](Rule* rr) {
                bool shouldBeOpen = !closeNow(rr);
                if (rr->previousOpenState != shouldBeOpen) {
                    // agentType.cpp:42-109 on the host copy of the schedules
                    std::vector<unsigned>& off = eventOffset;
                    std::vector<Event>& ev = events;
                    for (unsigned s = 0; s + 1 < off.size(); s++) {
                        unsigned ws = off[s], we = ws;
                        while (ws < off[s + 1]) {
                            while (we < ev.size() && !(ev[we].start < ev[ws].start) && !(ev[ws].start < ev[we].start)) we++;
                            if (shouldBeOpen) {
                                for (unsigned i = ws; i < we; i++)
                                    if (ev[i].locationType != homeT) ev[i].chance = (float)(ev[i].chance / (1.0 - fraction));
                                for (unsigned i = ws; i < we; i++)
                                    if (ev[i].locationType == homeT) {
                                        ev[i].chance = (float)((ev[i].chance - fraction) / (1.0 - fraction));
                                        if (std::abs(ev[i].chance) < 1e-10) {
                                            ev.erase(ev.begin() + i);
                                            for (size_t j = s + 1; j < off.size(); j++) off[j]--;
                                        }
                                        break;
                                    }
                            } else {
                                bool homePresent = false;
                                for (unsigned i = ws; i < we; i++)
                                    if (ev[i].locationType == homeT) {
                                        homePresent = true;
                                        ev[i].chance = (float)(ev[i].chance + (1.0 - ev[i].chance) * fraction);
                                        break;
                                    }
                                if (homePresent && we - ws == 1) {
                                    ws = we;
                                    continue;
                                }
                                if (!homePresent) {
                                    Event he;
                                    he.start = ev[ws].start;
                                    he.end = ev[ws].end;
                                    he.locationType = homeT;
                                    he.chance = (float)fraction;
                                    he.duration = ev[ws].duration;
                                    ev.insert(ev.begin() + ws, he);
                                    we++;
                                    for (size_t j = s + 1; j < off.size(); j++) off[j]++;
                                }
                                for (unsigned i = ws; i < we; i++)
                                    if (ev[i].locationType != homeT) ev[i].chance = (float)(ev[i].chance * (1.0 - fraction));
                            }
                            ws = we;
                        }
                    }
                    rr->previousOpenState = shouldBeOpen;
                }
            };
        } else if (r.name == "TestingProb") {
            std::string probs = r.parameter;
            rule.rule = [this, closeNow, probs](Rule* rr) {
                if (!closeNow(rr)) return;
                auto pr = ps::split_double(probs, ',');
                double* dst[6] = {&testingRandom, &testingHome, &testingWork, &testingSchool, &testingRandomHospital,
                                  &testingNurseryHome};
                for (size_t i = 0; i < pr.size() && i < 6; ++i) *dst[i] = pr[i];
            };
        } else {
            std::vector<int> toClose = r.locationTypesToClose;
            if (toClose.size() >= 10)
                throw ps::InitError(ps::ERR_BAD_CLOSURE, "Error, Closure Rule " + r.name + " has over 10 location types to close");
            rule.rule = [this, closeNow, toClose](Rule* rr) {
                bool shouldBeOpen = !closeNow(rr);
                if (rr->previousOpenState != shouldBeOpen) {
                    for (unsigned l = 0; l < L; ++l) {
                        if (essential[l] == 1) continue;
                        for (int t : toClose)
                            if (locType[l] == (unsigned)t) locOpen[l] = shouldBeOpen ? 1 : 0;
                    }
                    rr->previousOpenState = shouldBeOpen;
                }
            };
        }
        rules.push_back(rule);
    }
}

void Sim::closureMidnight(std::vector<unsigned>& st) {
    if (sc.opt.closurePolicy == "none") return;  // NoClosure::midnight (closurePolicies.h:42)
    for (auto& c : conds) c.active = c.condition(&c, st);
    for (auto& r : rules) r.rule(&r);
}

void Sim::midnight() {
    if (timestamp > 0) performTests();
    if (timestamp > 0) updateAgents();
    immunizationUpdate();
    if (sc.opt.otherDisease) otherDisease();
    auto st = refreshStatistics();
    if (snapOn) statSnap = std::make_shared<Sim>(*this);
    lastStats = st;
    dayStats.push_back(st);
    closureMidnight(st);
    planLocations();
    countVariantCases();
}

void Sim::step() {
    if (timestamp % stepsPerDay == 0) midnight();
    movement();
    for (unsigned v = 0; v < nVar; ++v)
        if (variantCounts[v] > 0) infectionsAtLocations(v);
    ++timestamp;
}

// simulation.h:779-897
void Sim::processFlags(const std::vector<std::string>& flags) {
    for (auto& flag : flags) {
        std::string prefix = flag.substr(0, 2);
        auto setProbs = [&](const char* s) {
            auto pr = ps::split_double(s, ',');
            double* dst[6] = {&testingRandom, &testingHome, &testingWork, &testingSchool, &testingRandomHospital,
                              &testingNurseryHome};
            for (size_t i = 0; i < pr.size() && i < 6; ++i) *dst[i] = pr[i];
        };
        if (prefix == "TP") {
            if (flag == "TPdef") setProbs("0.00005,0.01,0.0005,0.0005,0.005,0.05");
            else if (flag == "TP015") setProbs("0.00005,0.2,0.04,0.04,0.005,0.05");
            else if (flag == "TP035") setProbs("0.00005,1.0,0.2,0.2,0.005,0.05");
        } else if (prefix == "PL") {
            static const unsigned fix[4] = {5, 6, 22, 44};
            if (flag == "PLNONE" || flag == "PL0") {
                for (unsigned l = 0; l < L; ++l) {
                    if (essential[l] == 1) continue;
                    for (unsigned f : fix)
                        if (locType[l] == f) locOpen[l] = flag == "PLNONE" ? 1 : 0;
                }
            }
        } else if (prefix == "CF") {
            if (flag == "CFNONE") {
                enableCurfew = false;
                curfewBegin = 0;
                curfewEnd = 0;
                curfewTimeConverted = false;
            } else if (flag == "CF2000-0500") {
                enableCurfew = true;
                curfewBegin = 20 * 60;
                curfewEnd = 5 * 60;
                curfewTimeConverted = false;
            }
        } else if (prefix == "SO") {
            if (flag == "SONONE") schoolAgeRestriction = 99;
            else if (flag == "SO12") schoolAgeRestriction = 12;
            else if (flag == "SO3") schoolAgeRestriction = 3;
        } else if (prefix == "QU") {
            if (flag == "QU0") quarantinePolicy = 0;
            else if (flag == "QU1") quarantinePolicy = 1;
            else if (flag == "QU2") quarantinePolicy = 2;
            else if (flag == "QU3") quarantinePolicy = 3;
        } else if (prefix == "MA") {
            if (flag == "MA1.0" && currentMaskValue != 1.0) {
                double cur = currentMaskValue;
                for (unsigned l = 0; l < L; ++l)
                    if (locType[l] != homeT) locInf[l] = (float)(locInf[l] / cur);
                currentMaskValue = 1.0;
            } else if (flag == "MA0.8" && currentMaskValue != 0.8) {
                currentMaskValue = 0.8;
                for (unsigned l = 0; l < L; ++l)
                    if (locType[l] != homeT) locInf[l] = (float)(locInf[l] * 0.8);
            }
        }
    }
}

// simulation.h:928-956
std::vector<unsigned> Sim::runForDay(const std::vector<std::string>& flags) {
    processFlags(flags);
    auto st = refreshStatistics();
    closureMidnight(st);
    planLocations();
    countVariantCases();
    for (unsigned s = 0; s < stepsPerDay; ++s) {
        movement();
        for (unsigned v = 0; v < nVar; ++v)
            if (variantCounts[v] > 0) infectionsAtLocations(v);
        ++timestamp;
    }
    if (timestamp > 0) performTests();
    if (timestamp > 0) updateAgents();
    immunizationUpdate();
    if (sc.opt.otherDisease) otherDisease();
    st = refreshStatistics();
    lastStats = st;
    dayStats.push_back(st);
    return st;
}

// dumpLocationInfectiousList (infectionPolicies.h:262-441) and
// DumpStore::insert (:46-59): the locations where this pass infected
// someone (ascending), the infected agents by location, and at each of those
// locations the agents with non-zero infectiousness (ascending index, the
// order of the stable sort by location) with their infectiousness.
void Sim::dumpInsert(unsigned variant, const std::vector<unsigned>& newlyInfected) {
    if (newlyInfected.empty()) return;
    std::vector<unsigned> locs;
    for (unsigned i : newlyInfected) locs.push_back(location[i]);
    std::sort(locs.begin(), locs.end());
    locs.erase(std::unique(locs.begin(), locs.end()), locs.end());
    std::vector<int> slot(L, -1);
    for (size_t q = 0; q < locs.size(); ++q) slot[locs[q]] = (int)q;
    std::vector<std::vector<unsigned>> newBy(locs.size()), infBy(locs.size());
    for (unsigned i : newlyInfected) newBy[slot[location[i]]].push_back(i);
    for (unsigned i = 0; i < N; ++i)
        if (slot[location[i]] >= 0 && pp[i].infectious != 0.0f) infBy[slot[location[i]]].push_back(i);
    dTs.push_back(timestamp);
    dVar.push_back(variant);
    for (unsigned l : locs) dLocIds.push_back(l);
    dLocOff.push_back((unsigned)dLocIds.size());
    for (auto& v : newBy)
        for (unsigned i : v) dNew.push_back(i);
    dNewOff.push_back((unsigned)dNew.size());
    for (auto& v : infBy) {
        for (unsigned i : v) {
            dInf.push_back(i);
            dInfness.push_back(pp[i].infectious);
        }
        dInfOff.push_back((unsigned)dInf.size());
    }
}

// DumpStore::write (infectionPolicies.h:60-84): the count of inserts, then
// one line per array, every value followed by a space.
void Sim::writeDump(const std::string& path) const {
    std::ofstream f(path.c_str());
    if (!f.is_open()) return;
    f << dTs.size() << "\n";
    auto put = [&f](const auto& v) {
        for (const auto& x : v) f << x << " ";
        f << "\n";
    };
    put(dTs);
    put(dVar);
    put(dLocOff);
    put(dLocIds);
    put(dNewOff);
    put(dNew);
    put(dInfOff);
    put(dInf);
    put(dInfness);
}

// AgentStatOutput (agentStatOutput.cpp:5-44) through rapidjson's
// PrettyWriter (4-space indentation, JSONWriter.cpp:15-25).
void Sim::writeAgentStatJSON(const std::string& path) const {
    if (path.empty()) return;
    std::ofstream f(path.c_str());
    if (!f.is_open()) return;
    auto quoted = [](const std::string& s) {
        std::string o = "\"";
        for (char ch : s) {
            if (ch == '"' || ch == '\\') o += '\\';
            o += ch;
        }
        return o + "\"";
    };
    std::ostringstream out;
    out << "{\n    \"Statistics\": [";
    for (unsigned i = 0; i < N; ++i) {
        const Stats& e = stats[i];
        out << (i == 0 ? "\n" : ",\n") << "        {\n";
        const std::pair<const char*, unsigned long long> kv[] = {
            {"ID", i},
            {"infectionTime", e.infectedTimestamp},
            {"infectedCount", e.infectedCount},
            {"variant", e.variant},
            {"InfectionLoc", e.infectedLocation},
            {"diagnosisTime", e.diagnosedTimestamp},
            {"quarantinedTime", e.quarantinedTimestamp},
            {"quarantinedUntilTime", e.quarantinedUntilTimestamp},
            {"daysInQuarantine", e.daysInQuarantine},
            {"hospitalizedTime", e.hospitalizedTimestamp},
            {"hospitalizedUntilTime", e.hospitalizedUntilTimestamp},
            {"immunizationTimestamp", e.immunizationTimestamp},
            {"immunizationCount", e.immunizationCount},
            {"recoveredTimestamp", e.recoveredTimestamp},
        };
        for (const auto& p : kv) out << "            \"" << p.first << "\": " << p.second << ",\n";
        out << "            \"worstState\": {\n";
        out << "                \"name\": " << quoted(sc.st.names[(unsigned)(int)e.worstState]) << ",\n";
        out << "                \"begin\": " << e.worstStateTimestamp << ",\n";
        out << "                \"end\": ";
        if (e.worstStateEndTimestamp == 0) out << "-1";
        else out << e.worstStateEndTimestamp;
        out << "\n            }\n        }";
    }
    out << (N ? "\n    ]\n}" : "]\n}");
    f << out.str();
}

// The --diags block of printAgentStatJSON (agentsList.h:283-326).
std::string Sim::agentStatDiags() const {
    std::ostringstream os;
    unsigned days = 0;
    for (unsigned i = 0; i < N; ++i) days += stats[i].daysInQuarantine * (meta[i].age > 20 && meta[i].age <= 65);
    os << "Total number of days in quarantine for agents 20<age<=65\n" << days << "\n";
    const unsigned cemeteryLocation = L - 1;
    std::vector<unsigned> types(N);
    for (unsigned i = 0; i < N; ++i)
        types[i] = locType[stats[i].infectedTimestamp == UMAX ? cemeteryLocation : stats[i].infectedLocation];
    std::sort(types.begin(), types.end());
    const size_t width = sc.lt.types.size() + 1 + 2;  // generalLocationTypes (with the cemetery) + 2
    std::vector<unsigned> ids(width, 0u), counts(width, 0u);
    size_t groups = 0;
    for (unsigned i = 0; i < N; ++i) {
        if (i == 0 || types[i] != types[i - 1]) {
            if (groups < width) ids[groups] = types[i];
            ++groups;
        }
        if (groups - 1 < width) counts[groups - 1]++;
    }
    for (unsigned x : ids) os << x << " ";
    os << "\n";
    for (unsigned x : counts) os << x << " ";
    os << "\n";
    return os.str();
}

// Simulation::finalize (simulation.h:958-961): printAgentStatJSON, then
// BasicInfection::finalize (infectionPolicies.h:110-114).
void Sim::finalize() {
    writeAgentStatJSON(sc.opt.outAgentStat);
    diagsOut = sc.opt.diags > 0 ? agentStatDiags() : std::string();
    if (!sc.opt.dumpLocationInfectiousList.empty()) writeDump(sc.opt.dumpLocationInfectiousList);
}

}  // namespace oracle

// ---------------------------------------------------------------------------
// C interface for the Python tests (ctypes)
// ---------------------------------------------------------------------------
namespace {
std::string g_err;
}

extern "C" {

void* oracle_create(int argc, const char** argv) {
    try {
        ps::Options opt = ps::parse_options(argc, argv, false);
        ps::Scenario sc = ps::build_scenario(opt);
        return new oracle::Sim(sc);
    } catch (const ps::InitError& e) {
        g_err = e.what();
    } catch (const std::exception& e) {
        g_err = e.what();
    }
    return nullptr;
}

const char* oracle_last_error() { return g_err.c_str(); }

void oracle_plant(int code) { oracle::plant = code; }

void oracle_snapshot_stats(void* h, int on) { ((oracle::Sim*)h)->snapOn = on != 0; }

void oracle_destroy(void* h) { delete (oracle::Sim*)h; }

// Advances n iterations of runSimulation's loop; returns 0 or -1 on error.
int oracle_steps(void* h, int n) {
    auto* s = (oracle::Sim*)h;
    try {
        for (int i = 0; i < n; ++i) s->step();
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
    return 0;
}

int oracle_run_for_day(void* h, int nflags, const char** flags, unsigned* out, int cap) {
    auto* s = (oracle::Sim*)h;
    try {
        std::vector<std::string> f;
        for (int i = 0; i < nflags; ++i) f.push_back(flags[i]);
        auto st = s->runForDay(f);
        for (int i = 0; i < cap && i < (int)st.size(); ++i) out[i] = st[i];
        return (int)st.size();
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Writes the output files; copies what --diags printed into out (cap
// bytes, NUL-terminated) and returns its length, or -1 on error.
int oracle_finalize(void* h, char* out, int cap) {
    auto* s = (oracle::Sim*)h;
    try {
        s->finalize();
        const std::string& d = s->diagsOut;
        if (out && cap > 0) {
            const size_t n = std::min(d.size(), (size_t)cap - 1);
            memcpy(out, d.data(), n);
            out[n] = 0;
        }
        return (int)d.size();
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

unsigned oracle_timestamp(void* h) { return ((oracle::Sim*)h)->timestamp; }
unsigned oracle_num_agents(void* h) { return ((oracle::Sim*)h)->N; }
unsigned oracle_num_locations(void* h) { return ((oracle::Sim*)h)->L; }
int oracle_num_days(void* h) { return (int)((oracle::Sim*)h)->dayStats.size(); }

// Copies the statistics vector of day d; returns its length.
int oracle_day_stats(void* h, int d, unsigned* out, int cap) {
    auto* s = (oracle::Sim*)h;
    if (d < 0 || d >= (int)s->dayStats.size()) return -1;
    auto& v = s->dayStats[d];
    for (int i = 0; i < cap && i < (int)v.size(); ++i) out[i] = v[i];
    return (int)v.size();
}

// Named per-agent / per-location fields as uint32 (floats as raw bits).
int oracle_field(void* h, const char* name, unsigned* out) {
    auto* s = (oracle::Sim*)h;
    std::string n = name;
    if (n.rfind("snap:", 0) == 0) {  // the copy taken when the last statistics were computed
        if (!s->statSnap) return -1;
        s = s->statSnap.get();
        n = n.substr(5);
    }
    unsigned N = s->N, L = s->L;
    auto fbits = [](float f) {
        unsigned u;
        memcpy(&u, &f, 4);
        return u;
    };
    if (n == "state") for (unsigned i = 0; i < N; ++i) out[i] = (unsigned)s->pp[i].state;
    else if (n == "agentType") for (unsigned i = 0; i < N; ++i) out[i] = s->types[i];
    else if (n == "age") for (unsigned i = 0; i < N; ++i) out[i] = s->meta[i].age;
    else if (n == "sex") for (unsigned i = 0; i < N; ++i) out[i] = s->meta[i].sex ? 1u : 0u;
    else if (n == "preCond") for (unsigned i = 0; i < N; ++i) out[i] = s->meta[i].preCondIdx;
    else if (n == "variant") for (unsigned i = 0; i < N; ++i) out[i] = s->pp[i].variant;
    else if (n == "days") for (unsigned i = 0; i < N; ++i) out[i] = (unsigned)(int)s->pp[i].days;
    else if (n == "location") for (unsigned i = 0; i < N; ++i) out[i] = s->location[i];
    else if (n == "stepsUntilMove") for (unsigned i = 0; i < N; ++i) out[i] = s->stepsUntilMove[i];
    else if (n == "diagnosed") for (unsigned i = 0; i < N; ++i) out[i] = s->diagnosed[i];
    else if (n == "quarantined") for (unsigned i = 0; i < N; ++i) out[i] = s->quarantined[i];
    else if (n == "stayedHome") for (unsigned i = 0; i < N; ++i) out[i] = s->stayedHome[i];
    else if (n == "noWorkAgent") for (unsigned i = 0; i < N; ++i) out[i] = s->noWorkAgent[i];
    else if (n == "lastTest") for (unsigned i = 0; i < N; ++i) out[i] = s->lastTest[i];
    else if (n.rfind("susceptible", 0) == 0 && n.size() == 12 && n[11] >= '0' && n[11] < '0' + oracle::MAXS) {
        unsigned v = (unsigned)(n[11] - '0');
        for (unsigned i = 0; i < N; ++i) out[i] = fbits(s->pp[i].susceptible[v]);
    }
    else if (n == "healthcareWorker") for (unsigned i = 0; i < N; ++i) out[i] = s->hcw[i];
    else if (n == "timestamp") out[0] = s->timestamp;
    else if (n == "infectedTimestamp") for (unsigned i = 0; i < N; ++i) out[i] = s->stats[i].infectedTimestamp;
    else if (n == "infectedLocation") for (unsigned i = 0; i < N; ++i) out[i] = s->stats[i].infectedLocation;
    else if (n == "worstStateTimestamp") for (unsigned i = 0; i < N; ++i) out[i] = s->stats[i].worstStateTimestamp;
    else if (n == "worstStateEndTimestamp") for (unsigned i = 0; i < N; ++i) out[i] = s->stats[i].worstStateEndTimestamp;
    else if (n == "diagnosedTimestamp") for (unsigned i = 0; i < N; ++i) out[i] = s->stats[i].diagnosedTimestamp;
    else if (n == "quarantinedTimestamp") for (unsigned i = 0; i < N; ++i) out[i] = s->stats[i].quarantinedTimestamp;
    else if (n == "quarantinedUntilTimestamp") for (unsigned i = 0; i < N; ++i) out[i] = s->stats[i].quarantinedUntilTimestamp;
    else if (n == "daysInQuarantine") for (unsigned i = 0; i < N; ++i) out[i] = s->stats[i].daysInQuarantine;
    else if (n == "hospitalizedTimestamp") for (unsigned i = 0; i < N; ++i) out[i] = s->stats[i].hospitalizedTimestamp;
    else if (n == "hospitalizedUntilTimestamp") for (unsigned i = 0; i < N; ++i) out[i] = s->stats[i].hospitalizedUntilTimestamp;
    else if (n == "recoveredTimestamp") for (unsigned i = 0; i < N; ++i) out[i] = s->stats[i].recoveredTimestamp;
    else if (n == "immunizationTimestamp") for (unsigned i = 0; i < N; ++i) out[i] = s->stats[i].immunizationTimestamp;
    else if (n == "immunizationCount") for (unsigned i = 0; i < N; ++i) out[i] = s->stats[i].immunizationCount;
    else if (n == "vaccinationRound") for (unsigned i = 0; i < N; ++i) out[i] = s->immRound[i];
    else if (n == "infectedCount") for (unsigned i = 0; i < N; ++i) out[i] = s->stats[i].infectedCount;
    else if (n == "statsVariant") for (unsigned i = 0; i < N; ++i) out[i] = s->stats[i].variant;
    else if (n == "worstState") for (unsigned i = 0; i < N; ++i) out[i] = (unsigned)(int)s->stats[i].worstState;
    else if (n == "locQuarantineUntil") for (unsigned l = 0; l < L; ++l) out[l] = s->quarantineUntil[l];
    else if (n == "locClosedUntil") for (unsigned l = 0; l < L; ++l) out[l] = s->closedUntil[l];
    else if (n == "locOpen") for (unsigned l = 0; l < L; ++l) out[l] = s->locOpen[l];
    else if (n == "locEssential") for (unsigned l = 0; l < L; ++l) out[l] = s->essential[l];
    else if (n == "locInfectiousness") for (unsigned l = 0; l < L; ++l) out[l] = fbits(s->locInf[l]);
    else if (n == "locCount") {
        s->rebuildLocationSums();
        for (unsigned l = 0; l < L; ++l) out[l] = s->locCount[l];
    } else {
        g_err = "unknown field " + n;
        return -1;
    }
    return 0;
}

// Exact per-location infectiousness sum of variant v (2^-32 fixed point, as int64 pairs).
int oracle_location_sums(void* h, unsigned v, long long* out) {
    auto* s = (oracle::Sim*)h;
    if (v >= s->nVar) return -1;
    s->rebuildLocationSums();
    for (unsigned l = 0; l < s->L; ++l) out[l] = s->locSum[(size_t)v * s->L + l];
    return 0;
}

// Structure accessors for the invariant tests.
unsigned oracle_csr_size(void* h) { return (unsigned)((oracle::Sim*)h)->possibleLocations.size(); }
void oracle_csr(void* h, unsigned* offsets, unsigned* locs, unsigned char* types) {
    auto* s = (oracle::Sim*)h;
    for (size_t i = 0; i < s->locOffset.size(); ++i) offsets[i] = s->locOffset[i];
    for (size_t i = 0; i < s->possibleLocations.size(); ++i) {
        locs[i] = s->possibleLocations[i];
        types[i] = s->possibleTypes[i];
    }
}
unsigned oracle_cemetery(void* h) { return ((oracle::Sim*)h)->cemetery; }
// The schedule tables: eventOffset (28 * types + 1 entries) and, per event,
// its location type, chance and (start, end, duration) as (hours, minutes).
unsigned oracle_num_events(void* h) { return (unsigned)((oracle::Sim*)h)->events.size(); }
unsigned oracle_num_offsets(void* h) { return (unsigned)((oracle::Sim*)h)->eventOffset.size(); }
void oracle_schedule(void* h, unsigned* offsets, unsigned* types, float* chances, unsigned char* times) {
    auto* s = (oracle::Sim*)h;
    for (size_t i = 0; i < s->eventOffset.size(); ++i) offsets[i] = s->eventOffset[i];
    for (size_t i = 0; i < s->events.size(); ++i) {
        const auto& e = s->events[i];
        types[i] = e.locationType;
        chances[i] = e.chance;
        unsigned char t[6] = {e.start.h, e.start.m, e.end.h, e.end.m, e.duration.h, e.duration.m};
        for (int k = 0; k < 6; ++k) times[6 * i + k] = t[k];
    }
}
void oracle_location_types(void* h, unsigned* out) {
    auto* s = (oracle::Sim*)h;
    for (unsigned l = 0; l < s->L; ++l) out[l] = s->locType[l];
}
unsigned oracle_num_states(void* h) { return ((oracle::Sim*)h)->nStates; }
const char* oracle_state_name(void* h, unsigned k) {
    auto* s = (oracle::Sim*)h;
    return k < s->sc.st.names.size() ? s->sc.st.names[k].c_str() : "";
}
void oracle_state_table(void* h, unsigned char* wb, float* infectious) {
    auto* s = (oracle::Sim*)h;
    for (unsigned k = 0; k < s->nStates; ++k) {
        wb[k] = s->wb[k];
        infectious[k] = s->infTable[k];
    }
}
// The next state NextStates::selectNext picks for progression `prog`, state
// `state`, symptom scaling `scaling` and uniform draw `u`; bit 16 set for a
// bad progression.  -1 for indices out of range.
int oracle_select_next(void* h, unsigned prog, unsigned state, float scaling, double u) {
    auto* s = (oracle::Sim*)h;
    if (prog >= s->matrices.size() || state >= s->matrices[prog].bad.size()) return -1;
    auto r = s->selectNextFor(prog, state, scaling, u);
    return (int)(r.first | (r.second ? 0x10000u : 0u));
}
unsigned oracle_num_progressions(void* h) { return (unsigned)((oracle::Sim*)h)->matrices.size(); }
unsigned oracle_variant_multipliers(void* h, float* out, unsigned cap) {
    auto* s = (oracle::Sim*)h;
    for (unsigned v = 0; v < s->nVar && v < cap; ++v) out[v] = s->variantMul[v];
    return s->nVar;
}
void oracle_agent_infectious(void* h, float* out) {
    auto* s = (oracle::Sim*)h;
    for (unsigned i = 0; i < s->N; ++i) out[i] = s->pp[i].infectious;
}

// Calibration interface (hazard of expected infections and its gradient).
void oracle_set_type_multipliers(void* h, const float* m, unsigned n) {
    auto* s = (oracle::Sim*)h;
    for (unsigned t = 0; t < 256; ++t) s->typeMul[t] = t < n ? m[t] : 1.0f;
}
void oracle_hazard_enable(void* h, int on) {
    auto* s = (oracle::Sim*)h;
    s->hazardOn = on != 0;
    std::fill(s->hazard.begin(), s->hazard.end(), 0.0);
}
// a hazard probe at k + dk, m + dm[256] (perturbations, FP64)
int oracle_hazard_probe(void* h, double dk, const double* dm) {
    auto* s = (oracle::Sim*)h;
    oracle::Sim::Probe p{dk, std::vector<double>(dm, dm + 256)};
    s->probes.push_back(p);
    s->probeJ.push_back(0.0);
    return (int)s->probes.size() - 1;
}
void oracle_hazard_read(void* h, double* out, double* probeJ) {
    auto* s = (oracle::Sim*)h;
    for (int j = 0; j < 258; ++j) out[j] = s->hazard[j];
    for (size_t q = 0; probeJ && q < s->probeJ.size(); ++q) probeJ[q] = s->probeJ[q];
}

// Known-answer helpers for the tests.
void oracle_philox(const unsigned* ctr, unsigned k0, unsigned k1, unsigned* out) { oracle::philox10(ctr, k0, k1, out); }
float oracle_expf(float x) { return oracle::exp_f32(x); }
double oracle_log(double x) { return oracle::log_f64(x); }
double oracle_seasonality(void* h, unsigned ts) {
    auto* s = (oracle::Sim*)h;
    unsigned keep = s->timestamp;
    s->timestamp = ts;
    double r = s->seasonality();
    s->timestamp = keep;
    return r;
}
double oracle_length_p(float avg, float max) { return oracle::calculateModifiedP(avg, max); }
int oracle_geometric(unsigned long long seed, unsigned id, double p) {
    oracle::Rng rng(seed, id, 0, 0x7f0);
    oracle::Sim* none = nullptr;
    (void)none;
    double log1p_ = oracle::log_f64(1.0 - p);
    double cand = std::floor(oracle::log_f64(1.0 - rng.unit()) / log1p_);
    return (int)(cand + (1 - DBL_EPSILON) / 2);
}
}
