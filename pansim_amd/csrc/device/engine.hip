// C ABI of the gfx950 agent pipeline (include/pansim_hip.h): context with
// the device arrays of one population, its stream, and the launch sequences
// of every per-step and per-midnight routine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "day_kernels.h"
#include "engine_types.h"
#include "step_kernels.h"

using namespace ps::dev;

struct pansim_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    pansim_dims dims{};
    uint64_t seed = 0;
    uint32_t replica = 0;
    Agents A{};
    Locs L{};
    Tables T{};
    StepScratch S{};
    std::vector<void*> allocations;
    unsigned long long* counters = nullptr;  // device scratch for reductions (S_COUNT + 16)
    unsigned long long* hostCounters = nullptr;
    uint32_t hcwCount = 0;
    int numCUs = 256;
    uint32_t timing = 0;  // PANSIM_TIMING_* mode
    // per-call HIP event pairs on the context stream (kernel mode), or one
    // pair per movement + infection step (pipeline mode, in moveEv)
    std::vector<hipEvent_t> moveEv, infectEv;
    size_t moveN = 0, infectN = 0;
    bool pipeOpen = false;  // pipeline mode: a pair is open (a run of movement + infection launches)
    size_t pipeSteps = 0;   // pipeline mode: movement launches inside the pairs
    unsigned long long* moveCounters = nullptr;  // MOVE_CTR_SLOTS pairs (evaluated movers, relocated agents)
    uint32_t pendPar = 0;                         // parity of the pending-write counter of the next step
    bool resolvePending = false;                  // a movement step's quarantine writes are not applied yet
    uint32_t resolvePar = 0;
    // schedule storage (re-uploadable)
    uint32_t* evOffDev = nullptr;
    pansim_event* evDev = nullptr;
    uint32_t evOffCap = 0, evCap = 0;
    uint32_t nEvents = 0;
    uint32_t nCsr = 0;  // entries of the possible-location lists
    int4* evWinDev = nullptr;
    float* typeMul = nullptr;  // [HAZARD_TYPES]
    double* hazard = nullptr;  // [HAZARD_SLOTS] while hazard accumulation is on
    bool hazardOn = false;
    float probeH = 0.0f;      // finite-difference probes of the hazard (0: none)
    uint32_t probeType = 0;
    // per-variant infected counts of the last statistics pass, valid while no
    // state-changing launch has run since (stateEpoch counts those)
    uint64_t stateEpoch = 1, varEpoch = 0;
    uint32_t varCache[MAXV] = {0};
    uint32_t* dummyWait = nullptr;  // DummyMovement countdowns (allocated on first use)
    // agents by age band for the planning passes (ages do not change): under
    // 15 (checkUnderageAtHome) and 27-64 (setNoWorkToday)
    uint32_t* underageList = nullptr;
    uint32_t* workAgeList = nullptr;
    uint32_t nUnderage = 0, nWorkAge = 0;
    // infectious-list dump: the agents infected since the last collection,
    // and the collection's buffers (allocated when the dump is enabled)
    bool dumpOn = false;
    uint32_t *dumpNew = nullptr, *dumpNewLoc = nullptr, *dumpCnt = nullptr, *dumpFlags = nullptr;
    uint32_t *dumpInfAgent = nullptr, *dumpInfLoc = nullptr;
    float* dumpInfVal = nullptr;
    uint32_t dt = 10;
    uint32_t wakeDay = 0;  // first step of the day the stored wake steps count from
    std::vector<uint32_t> perm;  // device agent index -> input agent index
};

// A batch of replicas stepped together (include/pansim_hip.h): the member
// contexts keep their arrays; the batch owns the per-replica argument tables
// and the stream every member works on while it is batched.
// tile-size candidates of the replica batches' autotuner: agents per thread
// of the movement tile (its LDS mover list holds 256 x that many offsets)
// and of the infection tile
static constexpr uint32_t kMoveItems[3] = {8, 16, 32};
static constexpr uint32_t kInfectItems[2] = {4, 8};

struct pansim_batch {
    std::vector<pansim_ctx*> reps;
    std::vector<hipStream_t> ownStreams;  // members' streams, restored on destroy
    hipStream_t stream = nullptr;
    std::string err;
    uint32_t dt = 10;
    uint32_t pendPar = 0;
    ps::dev::MoveArgs* dMove = nullptr;
    ps::dev::InfectArgs* dInf = nullptr;
    // block -> replica tables for every tile size the autotuner may pick:
    // per movement size moveRep | moveFirst, per infection size infRep | infFirst
    uint32_t* dTables = nullptr;
    uint32_t moveOff[3] = {0, 0, 0}, moveBlocks[3] = {0, 0, 0};
    uint32_t infOff[2] = {0, 0}, infBlocks[2] = {0, 0};
    // tile sizes in use (indices into the candidate lists); -1 = still tuning
    int moveCfg = -1, infCfg = -1;
    bool autotune = true;
    struct TuneSample {
        int move, inf;
        hipEvent_t e0, e1;
    };
    std::vector<TuneSample> samples;
    ps::dev::MoveArgs* hMove = nullptr;  // pinned staging of the per-day arguments
    ps::dev::InfectArgs* hInf = nullptr;
    uint32_t* dDayStart = nullptr;  // the day's first timestamp, read by the batch kernels
    uint32_t* hDayStart = nullptr;  // its pinned staging
    bool haveDay = false;
    // one simulated day of batch steps captured as a graph (pansim_batch_run_day);
    // it depends on the hazard choice and the pending-write parity at the day start
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graphHazard = -1;
    uint32_t graphPar = 0;
};

namespace {

std::string g_create_error;

int fail(pansim_ctx* c, int code, const std::string& m) {
    if (c) c->err = m;
    return code;
}

#define HIP_OR_FAIL(ctx, call)                                                                        \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            return fail(ctx, PANSIM_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));     \
    } while (0)

template <class T>
int dalloc(pansim_ctx* c, T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) return fail(c, PANSIM_E_DEVICE, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    c->allocations.push_back(q);
    *p = (T*)q;
    return 0;
}

template <class T>
int upload(pansim_ctx* c, T** p, const T* src, size_t count) {
    int r = dalloc(c, p, count);
    if (r) return r;
    if (src && count) HIP_OR_FAIL(c, hipMemcpyAsync(*p, src, count * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return 0;
}

inline unsigned blocks_for(size_t n, unsigned bs = 256) { return (unsigned)((n + bs - 1) / bs); }

int check_launch(pansim_ctx* c, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, PANSIM_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    return 0;
}

// Applies a movement step's deferred quarantine writes when no infection
// launch has taken them (every entry point but pansim_infection_step calls
// this first, so the location state it sees is complete), and ends an open
// pipeline timing pair.
// Timing events: timestamps only, so no system-scope fence (cache
// writeback) at each record.
static int record_event(pansim_ctx* c, std::vector<hipEvent_t>& v, size_t idx) {
    while (v.size() <= idx) {
        hipEvent_t e;
        HIP_OR_FAIL(c, hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        v.push_back(e);
    }
    HIP_OR_FAIL(c, hipEventRecord(v[idx], c->stream));
    return 0;
}

// Ends an open pipeline timing pair after the launches enqueued so far.
static int pipe_close(pansim_ctx* c) {
    if (!c->pipeOpen) return 0;
    c->pipeOpen = false;
    if (record_event(c, c->moveEv, 2 * c->moveN + 1)) return PANSIM_E_DEVICE;
    c->moveN++;
    return 0;
}


// keepPipe: called by pansim_movement_step, which continues an open
// pipeline timing pair; every other entry point closes it.
static int flush_resolve(pansim_ctx* c, bool keepPipe = false) {
    if (!keepPipe)
        if (int r = pipe_close(c)) return r;
    if (!c->resolvePending) return 0;
    c->resolvePending = false;
    hipLaunchKernelGGL(k_resolve_pending, dim3(64), dim3(256), 0, c->stream, c->L, c->S, c->resolvePar);
    return check_launch(c, "resolve_pending");
}

unsigned persistent_blocks(pansim_ctx* c) { return (unsigned)c->numCUs * 8u; }

// Schedule offsets must be non-decreasing and stay within the events.
bool offsets_ok(const uint32_t* off, uint32_t nOff, uint32_t nEvents) {
    for (uint32_t k = 0; k + 1 < nOff; ++k)
        if (off[k] > off[k + 1]) return false;
    return nOff > 0 && off[nOff - 1] <= nEvents;
}

// Recomputes the per-step schedule windows after the schedules changed.
int compute_windows(pansim_ctx* c) {
    const uint32_t spd = 1440u / c->dt;
    const uint32_t total = c->T.nOff * spd;
    if (total == 0) return 0;
    hipLaunchKernelGGL(k_event_windows, dim3(blocks_for(total)), dim3(256), 0, c->stream, c->T.evOff, c->T.ev, c->T.nOff, spd,
                       c->dt, c->evWinDev);
    return check_launch(c, "event windows");
}

}  // namespace

extern "C" {

const char* pansim_create_error(void) { return g_create_error.c_str(); }
const char* pansim_last_error(const pansim_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int pansim_create(pansim_ctx** out, const pansim_dims* d, const pansim_tables* t, const pansim_locations* l,
                  const pansim_agents* a, uint64_t seed, uint32_t replica, uint32_t time_step, uint32_t quarantine_policy,
                  uint32_t quarantine_length) {
    if (!out || !d || !t || !l || !a) {
        g_create_error = "pansim_create: null argument";
        return PANSIM_E_ARG;
    }
    *out = nullptr;
    if (d->n_agents == 0 || d->n_locations == 0 || d->n_variants == 0 || d->n_variants > 7 || d->n_states == 0 ||
        d->n_states > 127 || time_step == 0 || 1440 % time_step != 0) {
        g_create_error = "pansim_create: invalid dimensions";
        return PANSIM_E_ARG;
    }
    // the fused infection pass assumes susceptible states carry no infectiousness
    for (uint32_t s = 0; s < d->n_states; ++s)
        if (t->state_susceptible[s] && t->state_infectious[s] != 0.0f) {
            g_create_error = "pansim_create: a susceptible state with non-zero infectiousness is not supported";
            return PANSIM_E_UNSUPPORTED;
        }
    if (t->first_infected >= d->n_states || t->non_covid_dead >= d->n_states || t->dead_state >= d->n_states) {
        g_create_error = "pansim_create: first infected, non-COVID dead or dead state out of range";
        return PANSIM_E_ARG;
    }
    // The per-location infectiousness sums are kept exactly in 2^-32 fixed
    // point.  An agent's infectiousness is the float product of its state's
    // value and its variant's multiplier; a float of at least 2^-9 is a
    // multiple of 2^-32, and below 2^31 / n_agents no sum can leave the
    // 64-bit range.  Other values would be truncated: refuse them.
    for (uint32_t s = 0; s < d->n_states; ++s)
        for (uint32_t v = 0; v < d->n_variants; ++v) {
            const float x = t->state_infectious[s] * t->variant_multiplier[v];
            if (x == 0.0f) continue;
            if (!(x >= 0x1p-9f) || !((double)x * (double)d->n_agents < 2147483648.0)) {
                g_create_error = "pansim_create: an agent infectiousness (state value times variant multiplier) "
                                 "outside [2^-9, 2^31 / n_agents) is not supported (exact fixed-point sums)";
                return PANSIM_E_UNSUPPORTED;
            }
        }
    if (t->state_infectious[t->first_infected] != 0.0f) {
        g_create_error = "pansim_create: non-zero infectiousness of the first infected state is not supported";
        return PANSIM_E_UNSUPPORTED;
    }
    pansim_ctx* c = new pansim_ctx();
    c->dims = *d;
    c->seed = seed;
    c->replica = replica;
    auto bail = [&](int code) {
        g_create_error = c->err;
        pansim_destroy(c);
        return code;
    };
    if (hipGetDevice(&c->device) != hipSuccess) {
        c->err = "no HIP device";
        return bail(PANSIM_E_DEVICE);
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) c->numCUs = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        c->err = "hipStreamCreate failed";
        return bail(PANSIM_E_DEVICE);
    }
    const uint32_t N = d->n_agents, NL = d->n_locations, V = d->n_variants, NS = d->n_states;
    const uint32_t NP = d->n_progressions;
    int r = 0;
#define TRY(x)                  \
    do {                        \
        if ((r = (x)) != 0) return bail(r); \
    } while (0)
    // ---- tables
    Tables& T = c->T;
    T.nStates = NS;
    T.nVar = V;
    T.nProg = NP;
    T.firstInfected = t->first_infected;
    T.nonCOVIDDead = t->non_covid_dead;
    T.deadState = t->dead_state;
    T.lt = t->lt;
    TRY(upload(c, (float**)&T.stInf, t->state_infectious, NS));
    TRY(upload(c, (uint8_t**)&T.stWB, t->state_wb, NS));
    TRY(upload(c, (uint8_t**)&T.stSusc, t->state_susceptible, NS));
    TRY(upload(c, (uint8_t**)&T.stInfected, t->state_infected, NS));
    TRY(upload(c, (float**)&T.accPCR, t->state_accuracy_pcr, NS));
    TRY(upload(c, (float**)&T.accAnt, t->state_accuracy_antigen, NS));
    TRY(upload(c, (float**)&T.varMul, t->variant_multiplier, V));
    TRY(upload(c, (float**)&T.progScaling, t->progression_scaling, V));
    TRY(upload(c, (float**)&T.deathScaling, t->death_scaling, V));
    TRY(upload(c, (uint32_t**)&T.edgeOff, t->edge_off, (size_t)NP * NS + 1));
    TRY(upload(c, (uint32_t**)&T.badCount, t->bad_count, (size_t)NP * NS));
    TRY(upload(c, (uint32_t**)&T.edgeTo, t->edge_to, d->n_matrix_edges));
    TRY(upload(c, (float**)&T.edgeChance, t->edge_chance, d->n_matrix_edges));
    TRY(upload(c, (float**)&T.lenAvg, t->length_avg, (size_t)NP * NS * 7));
    TRY(upload(c, (float**)&T.lenMax, t->length_max, (size_t)NP * NS * 7));
    TRY(upload(c, (double**)&T.lenP, t->length_p, (size_t)NP * NS * 7));
    c->evOffCap = d->n_agent_types * 28 + 1;
    if (!offsets_ok(t->event_offset, c->evOffCap, d->n_events)) {
        c->err = "pansim_create: schedule offsets out of range";
        return bail(PANSIM_E_ARG);
    }
    c->evCap = d->n_events + 4096;
    TRY(dalloc(c, &c->evOffDev, c->evOffCap));
    TRY(dalloc(c, &c->evDev, c->evCap));
    HIP_OR_FAIL(c, hipMemcpyAsync(c->evOffDev, t->event_offset, c->evOffCap * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    if (d->n_events)
        HIP_OR_FAIL(c, hipMemcpyAsync(c->evDev, t->events, d->n_events * sizeof(pansim_event), hipMemcpyHostToDevice, c->stream));
    T.evOff = c->evOffDev;
    T.ev = c->evDev;
    c->nEvents = d->n_events;
    c->dt = time_step;
    T.nOff = c->evOffCap - 1;
    TRY(dalloc(c, &c->evWinDev, (size_t)T.nOff * (1440u / time_step)));
    T.evWin = c->evWinDev;
    if ((r = compute_windows(c)) != 0) return bail(r);
    // ---- input checks (indexes the kernels rely on)
    for (uint32_t k = 0; k < NL; ++k)
        if (l->type[k] >= LOC_TYPE_NONE) {
            c->err = "pansim_create: location type ids must be below 255";
            return bail(PANSIM_E_UNSUPPORTED);
        }
    if (l->cemetery >= NL) {
        c->err = "pansim_create: cemetery index out of range";
        return bail(PANSIM_E_ARG);
    }
    if (a->loc_offset[0] != 0 || a->loc_offset[N] != d->n_csr) {
        c->err = "pansim_create: inconsistent possible-location offsets";
        return bail(PANSIM_E_ARG);
    }
    for (uint32_t k = 0; k < N; ++k)
        if (a->location[k] >= NL || a->loc_offset[k + 1] < a->loc_offset[k] || a->loc_offset[k + 1] - a->loc_offset[k] > 0xFFFFu) {
            c->err = "pansim_create: agent location or location list out of range";
            return bail(PANSIM_E_ARG);
        }
    for (uint32_t k = 0; k < d->n_csr; ++k)
        if (a->possible_locations[k] >= NL) {
            c->err = "pansim_create: possible location out of range";
            return bail(PANSIM_E_ARG);
        }
    // ---- storage order: agents grouped by schedule (a counting sort by agent
    // type, input order within a type), so that the agents that move at one
    // step sit next to each other; every agent array below is uploaded in
    // this order and c->perm (device index -> input index) maps back
    std::vector<uint8_t> pAge(N), pSex(N), pPre(N), pType(N), pProg(N), pState(N), pDiag(N), pHcw(N), pPType(d->n_csr);
    std::vector<uint32_t> pLoc(N), pOff((size_t)N + 1), pPLoc(d->n_csr);
    std::vector<float> pScal(N);
    pansim_agents ap = *a;
    {
        uint32_t first[257] = {0};
        for (uint32_t k = 0; k < N; ++k) first[a->agent_type[k] + 1]++;
        for (int b = 0; b < 256; ++b) first[b + 1] += first[b];
        c->perm.resize(N);
        for (uint32_t k = 0; k < N; ++k) c->perm[first[a->agent_type[k]]++] = k;
        pOff[0] = 0;
        for (uint32_t i = 0; i < N; ++i) {
            const uint32_t k = c->perm[i];
            pAge[i] = a->age[k];
            pSex[i] = a->sex[k];
            pPre[i] = a->precond[k];
            pType[i] = a->agent_type[k];
            pProg[i] = a->progression[k];
            pState[i] = a->state[k];
            pDiag[i] = a->diagnosed[k];
            pHcw[i] = a->healthcare_worker[k];
            pLoc[i] = a->location[k];
            pScal[i] = a->scaling_age_sex[k];
            const uint32_t b = a->loc_offset[k], e = a->loc_offset[k + 1];
            std::copy(a->possible_locations + b, a->possible_locations + e, pPLoc.begin() + pOff[i]);
            std::copy(a->possible_types + b, a->possible_types + e, pPType.begin() + pOff[i]);
            pOff[i + 1] = pOff[i] + (e - b);
        }
        ap.age = pAge.data();
        ap.sex = pSex.data();
        ap.precond = pPre.data();
        ap.agent_type = pType.data();
        ap.progression = pProg.data();
        ap.state = pState.data();
        ap.diagnosed = pDiag.data();
        ap.healthcare_worker = pHcw.data();
        ap.location = pLoc.data();
        ap.scaling_age_sex = pScal.data();
        ap.loc_offset = pOff.data();
        ap.possible_locations = pPLoc.data();
        ap.possible_types = pPType.data();
        a = &ap;
    }
    // ---- locations
    Locs& L = c->L;
    L.n = NL;
    L.cemetery = l->cemetery;
    L.nSchools = d->n_schools;
    {
        std::vector<LocRec> lr(NL);
        std::vector<LocDyn> ld(NL);
        std::vector<uint8_t> t8(NL);
        for (uint32_t k = 0; k < NL; ++k) {
            lr[k].quarUntil = 0;
            lr[k].closedUntil = 0;
            lr[k].noWork = 0;
            lr[k].type = t8[k] = (uint8_t)l->type[k];
            lr[k].open = l->open[k];
            lr[k].essential = l->essential[k];
            lr[k].flag = 0;
            ld[k].sum0 = 0;
            ld[k].count = 0;
            ld[k].inf = l->infectiousness[k];  // type multipliers start at 1
        }
        TRY(upload(c, &L.rec, lr.data(), NL));
        TRY(upload(c, &L.dyn, ld.data(), NL));
        TRY(upload(c, &L.infRaw, l->infectiousness, NL));
        TRY(upload(c, (uint8_t**)&L.type8, t8.data(), NL));
        HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    }
    TRY(dalloc(c, &L.sumX, (size_t)(V - 1) * NL));
    TRY(dalloc(c, &L.visited, ((size_t)NL + 31) / 32));
    HIP_OR_FAIL(c, hipMemsetAsync(L.visited, 0, ((size_t)NL + 31) / 32 * sizeof(uint32_t), c->stream));
    // every location special until the first midnight planning computes the bits
    TRY(dalloc(c, &L.special, ((size_t)NL + 31) / 32));
    HIP_OR_FAIL(c, hipMemsetAsync(L.special, 0xFF, ((size_t)NL + 31) / 32 * sizeof(uint32_t), c->stream));
    TRY(dalloc(c, &L.noWorkBits, ((size_t)NL + 31) / 32));
    HIP_OR_FAIL(c, hipMemsetAsync(L.noWorkBits, 0, ((size_t)NL + 31) / 32 * sizeof(uint32_t), c->stream));
    TRY(dalloc(c, &L.flagBits, ((size_t)NL + 31) / 32));
    HIP_OR_FAIL(c, hipMemsetAsync(L.flagBits, 0, ((size_t)NL + 31) / 32 * sizeof(uint32_t), c->stream));
    TRY(dalloc(c, &L.pendKey, NL));
    {
        std::vector<float> ones(HAZARD_TYPES, 1.0f);
        TRY(upload(c, &c->typeMul, ones.data(), HAZARD_TYPES));
        TRY(dalloc(c, &c->hazard, HAZARD_SLOTS));
        HIP_OR_FAIL(c, hipMemsetAsync(c->hazard, 0, HAZARD_SLOTS * sizeof(double), c->stream));
        HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    }
    TRY(upload(c, (uint32_t**)&L.schools, l->schools, d->n_schools));
    TRY(upload(c, (uint32_t**)&L.classrooms, l->classrooms, d->n_classrooms));
    TRY(upload(c, (uint32_t**)&L.classOff, l->classroom_offsets, d->n_schools + 1));
    HIP_OR_FAIL(c, hipMemsetAsync(L.pendKey, 0, NL * sizeof(unsigned long long), c->stream));
    // ---- agents
    Agents& A = c->A;
    A.n = N;
    TRY(upload(c, (uint32_t**)&A.id, c->perm.data(), N));
    {
        std::vector<AgentRec> ar(N);
        uint32_t hcw = 0;
        const uint32_t homeT = t->lt.home;
        for (uint32_t k = 0; k < N; ++k) {
            AgentRec& r = ar[k];
            r.quarUntil = r.diagTs = r.hospTs = r.hospUntil = 0;  // set by k_init_agents
            r.csrBegin = a->loc_offset[k];
            r.csrCount = (uint16_t)(a->loc_offset[k + 1] - a->loc_offset[k]);
            r.home = NOLOC;
            for (uint32_t e = a->loc_offset[k]; e < a->loc_offset[k + 1]; ++e)
                if (a->possible_types[e] == homeT) {
                    r.home = a->possible_locations[e];
                    break;
                }
            r.type = a->agent_type[k];
            r.state = a->state[k];
            r.flags = (uint8_t)((a->diagnosed[k] ? F_DIAGNOSED : 0) | (a->healthcare_worker[k] ? F_HCW : 0));
            r.variant = 0;
            r.age = a->age[k];
            r.wb = t->state_wb[r.state];
            hcw += a->healthcare_worker[k] ? 1 : 0;
        }
        c->hcwCount = hcw;
        TRY(upload(c, &A.rec, ar.data(), N));
        std::vector<uint32_t> under, work;
        for (uint32_t k = 0; k < N; ++k) {
            if (a->age[k] <= 14) under.push_back(k);
            if (a->age[k] > 26 && a->age[k] < 65) work.push_back(k);
        }
        c->nUnderage = (uint32_t)under.size();
        c->nWorkAge = (uint32_t)work.size();
        if (!under.empty()) TRY(upload(c, &c->underageList, under.data(), under.size()));
        if (!work.empty()) TRY(upload(c, &c->workAgeList, work.data(), work.size()));
        HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    }
    TRY(upload(c, &A.prog, a->progression, N));
    TRY(upload(c, &A.sex, a->sex, N));
    TRY(upload(c, &A.precond, a->precond, N));
    TRY(dalloc(c, &A.days, N));
    // location and susceptibility arrays padded to whole infection tiles
    // (16-byte vector loads); the padding holds location 0 and susceptibility 0
    const size_t nPad = ((size_t)N + INFECT_TILE - 1) / INFECT_TILE * INFECT_TILE;
    A.suscStride = (uint32_t)nPad;
    TRY(dalloc(c, &A.location, nPad));
    HIP_OR_FAIL(c, hipMemsetAsync(A.location, 0, nPad * sizeof(uint32_t), c->stream));
    HIP_OR_FAIL(c, hipMemcpyAsync(A.location, a->location, (size_t)N * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    {
        // padded to whole movement tiles with WAKE_NONE: the scan reads 16-byte vectors
        const size_t nw = ((size_t)N + MOVE_TILE - 1) / MOVE_TILE * MOVE_TILE;
        TRY(dalloc(c, &A.wake, nw));
        HIP_OR_FAIL(c, hipMemsetAsync(A.wake, 0xFF, nw * sizeof(uint16_t), c->stream));
    }
    TRY(dalloc(c, &A.lastTest, N));
    TRY(upload(c, &A.scal, a->scaling_age_sex, N));
    TRY(dalloc(c, &A.susc, (size_t)V * nPad));
    HIP_OR_FAIL(c, hipMemsetAsync(A.susc, 0, (size_t)V * nPad * sizeof(float), c->stream));
    TRY(dalloc(c, &A.slot6, (size_t)V * N));
    uint32_t** u32s[] = {&A.infTs, &A.infLoc, &A.worstTs, &A.worstEndTs, &A.quarTs, &A.daysInQ, &A.immTs, &A.recTs};
    for (auto p : u32s) TRY(dalloc(c, p, N));
    TRY(dalloc(c, &A.immCount, N));
    TRY(dalloc(c, &A.immRound, N));
    HIP_OR_FAIL(c, hipMemsetAsync(A.immRound, 0, N, c->stream));
    TRY(dalloc(c, &A.statVar, N));
    TRY(dalloc(c, &A.infCount, N));
    TRY(dalloc(c, &A.worstState, N));
    TRY(upload(c, (uint32_t**)&A.possLoc, a->possible_locations, d->n_csr));
    c->nCsr = d->n_csr;
    {
        TRY(upload(c, (uint8_t**)&A.possType, a->possible_types, d->n_csr));
        // the types of every list of at most 16 entries as one 16-byte record per agent
        std::vector<uint32_t> t16((size_t)N * 4, 0u);
        for (uint32_t k = 0; k < N; ++k) {
            const uint32_t b = a->loc_offset[k], e = a->loc_offset[k + 1];
            if (e - b > 16u) continue;
            for (uint32_t j = b; j < e; ++j)
                t16[(size_t)k * 4 + (j - b) / 4] |= (uint32_t)a->possible_types[j] << (8u * ((j - b) % 4));
        }
        TRY(upload(c, (uint32_t**)&A.types16, t16.data(), (size_t)N * 4));
        HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    }
    // ---- scratch
    c->S.pendCap = d->n_locations + 16u;
    TRY(dalloc(c, &c->S.pendList, c->S.pendCap));
    TRY(dalloc(c, &c->S.pendCount, 2));
    HIP_OR_FAIL(c, hipMemsetAsync(c->S.pendCount, 0, 2 * sizeof(uint32_t), c->stream));
    TRY(dalloc(c, &c->counters, S_COUNT + 16));
    if (hipHostMalloc((void**)&c->hostCounters, (S_COUNT + 16) * sizeof(unsigned long long), hipHostMallocDefault) !=
        hipSuccess) {
        c->err = "hipHostMalloc failed";
        return bail(PANSIM_E_DEVICE);
    }
    TRY(dalloc(c, &c->moveCounters, 2 * MOVE_CTR_SLOTS));
    HIP_OR_FAIL(c, hipMemsetAsync(c->moveCounters, 0, 2 * MOVE_CTR_SLOTS * sizeof(unsigned long long), c->stream));
    // ---- initial agent state
    uint32_t qlenSteps = quarantine_length * 24u * 60u / time_step;
    hipLaunchKernelGGL(k_init_agents, dim3(blocks_for(N)), dim3(256), 0, c->stream, c->A, c->T, seed, replica,
                       quarantine_policy, qlenSteps, quarantine_length);
    if ((r = check_launch(c, "k_init_agents")) != 0) return bail(r);
    hipLaunchKernelGGL(k_zero_locations, dim3(persistent_blocks(c)), dim3(256), 0, c->stream, c->L, c->T.nVar);
    hipLaunchKernelGGL(k_scatter_locations, dim3(blocks_for(N)), dim3(256), 0, c->stream, c->A, c->L, c->T);
    if ((r = check_launch(c, "initial occupancy")) != 0) return bail(r);
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
#undef TRY
    *out = c;
    return 0;
}

void pansim_destroy(pansim_ctx* c) {
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    for (void* p : c->allocations) hipFree(p);
    if (c->hostCounters) hipHostFree(c->hostCounters);
    for (auto e : c->moveEv) hipEventDestroy(e);
    for (auto e : c->infectEv) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int pansim_sync(pansim_ctx* c) {
    if (!c) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    return 0;
}

int pansim_enable_timing(pansim_ctx* c, uint32_t on) {
    if (!c) return PANSIM_E_ARG;
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    if (on > PANSIM_TIMING_MOVEMENT) return fail(c, PANSIM_E_ARG, "enable_timing: unknown mode");
    c->timing = on;
    c->moveN = c->infectN = 0;
    c->pipeOpen = false;
    c->pipeSteps = 0;
    return 0;
}



int pansim_timing_summary(pansim_ctx* c, uint32_t* n_move, double* move_ms, uint32_t* n_infect, double* infect_ms) {
    if (!c) return PANSIM_E_ARG;
    if (int r = pipe_close(c)) return r;
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    double m = 0.0, f = 0.0;
    for (size_t k = 0; k + 1 < 2 * c->moveN + 1 && 2 * k + 1 < c->moveEv.size() && k < c->moveN; ++k) {
        float x = 0.f;
        HIP_OR_FAIL(c, hipEventElapsedTime(&x, c->moveEv[2 * k], c->moveEv[2 * k + 1]));
        m += x;
    }
    for (size_t k = 0; k < c->infectN && 2 * k + 1 < c->infectEv.size(); ++k) {
        float x = 0.f;
        HIP_OR_FAIL(c, hipEventElapsedTime(&x, c->infectEv[2 * k], c->infectEv[2 * k + 1]));
        f += x;
    }
    // pipeline mode: the steps (movement launches) the pairs cover
    if (n_move) *n_move = (uint32_t)(c->timing == PANSIM_TIMING_PIPELINE ? c->pipeSteps : c->moveN);
    if (move_ms) *move_ms = m;
    if (n_infect) *n_infect = (uint32_t)c->infectN;
    if (infect_ms) *infect_ms = f;
    return 0;
}

int pansim_last_kernel_ms(pansim_ctx* c, float* move_ms, float* infect_ms) {
    if (!c) return PANSIM_E_ARG;
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    float m = 0.f, f = 0.f;
    if (c->moveN) HIP_OR_FAIL(c, hipEventElapsedTime(&m, c->moveEv[2 * (c->moveN - 1)], c->moveEv[2 * (c->moveN - 1) + 1]));
    if (c->infectN)
        HIP_OR_FAIL(c, hipEventElapsedTime(&f, c->infectEv[2 * (c->infectN - 1)], c->infectEv[2 * (c->infectN - 1) + 1]));
    if (move_ms) *move_ms = m;
    if (infect_ms) *infect_ms = f;
    return 0;
}

int pansim_movement_counters(pansim_ctx* c, uint64_t* evaluated, uint64_t* relocated, uint32_t reset) {
    if (!c) return PANSIM_E_ARG;
    unsigned long long h[2 * MOVE_CTR_SLOTS];
    HIP_OR_FAIL(c, hipMemcpyAsync(h, c->moveCounters, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    uint64_t ev = 0, re = 0;
    for (uint32_t s = 0; s < MOVE_CTR_SLOTS; ++s) {
        ev += h[2 * s];
        re += h[2 * s + 1];
    }
    if (evaluated) *evaluated = ev;
    if (relocated) *relocated = re;
    if (reset) HIP_OR_FAIL(c, hipMemsetAsync(c->moveCounters, 0, sizeof(h), c->stream));
    return 0;
}

static MoveArgs move_args(pansim_ctx* c, const pansim_step_params* p) {
    MoveArgs m;
    m.A = c->A;
    m.L = c->L;
    m.T = c->T;
    m.S = c->S;
    m.p = *p;
    m.seed = c->seed;
    m.replica = c->replica;
    m.nEvents = c->nEvents;
    m.counters = c->moveCounters;
    m.pendPar = c->pendPar;
    const uint32_t dt = p->time_step;
    m.spd = 1440u / dt;
    m.qLen = p->quarantine_length * 24u * 60u / dt;
    m.hourSteps = 60u / dt;
    m.halfHourSteps = 30u / dt;
    m.dtMagic = ((1ull << 32) + dt - 1u) / dt;
    move_set_step(m, p->timestamp, p->timestamp % m.spd);
    return m;
}

static InfectArgs infect_args(pansim_ctx* c, uint32_t timestamp, uint32_t mask, float k, float season, uint32_t dt) {
    InfectArgs a;
    a.A = c->A;
    a.L = c->L;
    a.T = c->T;
    a.S = c->S;
    a.seed = c->seed;
    a.replica = c->replica;
    a.timestamp = timestamp;
    a.mask = mask & ((1u << c->T.nVar) - 1u);
    a.k = k;
    a.season = season;
    a.dt = (float)dt;
    a.typeMul = c->typeMul;
    a.hazard = c->hazardOn ? c->hazard : nullptr;
    a.probeH = c->probeH;
    a.probeType = c->probeType;
    a.dumpList = c->dumpOn ? c->dumpNew : nullptr;
    a.dumpCount = c->dumpCnt;
    a.resolve = 0;
    a.resolvePar = 0;
    a.nTiles = blocks_for(c->A.n, INFECT_TILE);
    return a;
}


int pansim_movement_step(pansim_ctx* c, const pansim_step_params* p) {
    if (!c || !p) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c, true)) return fr;
    if (p->time_step != c->dt) return fail(c, PANSIM_E_ARG, "movement_step: time step differs from the one given at creation");
    MoveArgs m = move_args(c, p);
    if (p->timestamp - m.stepOfDay != c->wakeDay)
        return fail(c, PANSIM_E_ARG, "movement_step: pansim_plan_locations must run at the start of every day");
    c->pendPar ^= 1u;
    const bool perLaunch = c->timing == PANSIM_TIMING_KERNELS || c->timing == PANSIM_TIMING_MOVEMENT;
    if (perLaunch && record_event(c, c->moveEv, 2 * c->moveN)) return PANSIM_E_DEVICE;
    if (c->timing == PANSIM_TIMING_PIPELINE && !c->pipeOpen) {
        if (record_event(c, c->moveEv, 2 * c->moveN)) return PANSIM_E_DEVICE;
        c->pipeOpen = true;
    }
    hipLaunchKernelGGL(k_move_tile, dim3(blocks_for(c->A.n, MOVE_TILE)), dim3(MOVE_BLOCK), 0, c->stream, m);
    c->resolvePending = true;
    c->resolvePar = m.pendPar;
    if (perLaunch) {
        if (record_event(c, c->moveEv, 2 * c->moveN + 1)) return PANSIM_E_DEVICE;
        c->moveN++;
    }
    if (c->timing == PANSIM_TIMING_PIPELINE) c->pipeSteps++;
    return check_launch(c, "movement");
}

int pansim_infection_step(pansim_ctx* c, uint32_t timestamp, uint32_t mask, float k, float season, uint32_t dt) {
    if (!c) return PANSIM_E_ARG;
    InfectArgs a = infect_args(c, timestamp, mask, k, season, dt);
    if (a.mask) c->stateEpoch++;
    a.resolve = c->resolvePending ? 1u : 0u;
    a.resolvePar = c->resolvePar;
    c->resolvePending = false;
    const unsigned grid = blocks_for(c->A.n, INFECT_TILE);
    if (c->timing == PANSIM_TIMING_KERNELS && record_event(c, c->infectEv, 2 * c->infectN)) return PANSIM_E_DEVICE;
    if (a.hazard) hipLaunchKernelGGL(k_infect<true>, dim3(grid), dim3(INFECT_BLOCK), 0, c->stream, a);
    else hipLaunchKernelGGL(k_infect<false>, dim3(grid), dim3(INFECT_BLOCK), 0, c->stream, a);
    if (c->timing == PANSIM_TIMING_KERNELS) {
        if (record_event(c, c->infectEv, 2 * c->infectN + 1)) return PANSIM_E_DEVICE;
        c->infectN++;
    }
    return check_launch(c, "infection");
}

static int rebuild_locations(pansim_ctx* c) {
    hipLaunchKernelGGL(k_zero_locations, dim3(persistent_blocks(c)), dim3(256), 0, c->stream, c->L, c->T.nVar);
    hipLaunchKernelGGL(k_scatter_locations, dim3(blocks_for(c->A.n)), dim3(256), 0, c->stream, c->A, c->L, c->T);
    return check_launch(c, "rebuild_locations");
}

int pansim_rebuild_locations(pansim_ctx* c) {
    if (!c) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    return rebuild_locations(c);
}

int pansim_dummy_plan_locations(pansim_ctx* c) {
    if (!c) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    if (!c->dummyWait)
        if (int r = dalloc(c, &c->dummyWait, c->A.n)) return r;
    HIP_OR_FAIL(c, hipMemsetAsync(c->dummyWait, 0, (size_t)c->A.n * sizeof(uint32_t), c->stream));
    return rebuild_locations(c);
}

int pansim_dummy_movement_step(pansim_ctx* c, uint32_t timestamp) {
    if (!c) return PANSIM_E_ARG;
    if (!c->dummyWait) return fail(c, PANSIM_E_ARG, "dummy_movement_step: pansim_dummy_plan_locations was never called");
    if (int fr = flush_resolve(c)) return fr;
    hipLaunchKernelGGL(k_dummy_move, dim3(blocks_for(c->A.n)), dim3(256), 0, c->stream, c->A, c->L, c->T, c->dummyWait,
                       c->seed, c->replica, timestamp);
    return check_launch(c, "dummy_movement_step");
}

int pansim_plan_locations(pansim_ctx* c, const pansim_step_params* p) {
    if (!c || !p) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    PlanArgs a;
    a.A = c->A;
    a.L = c->L;
    a.T = c->T;
    a.seed = c->seed;
    a.replica = c->replica;
    a.ts = p->timestamp;
    a.dt = p->time_step;
    a.curfewEnd = p->curfew_end;
    a.schoolAge = p->school_age_restriction;
    const unsigned gN = blocks_for(c->A.n);
    if (p->time_step != c->dt) return fail(c, PANSIM_E_ARG, "plan_locations: time step differs from the one given at creation");
    c->wakeDay = p->timestamp - p->timestamp % (1440u / c->dt);
    if (p->quarantine_policy > 3 && c->L.nSchools)
        hipLaunchKernelGGL(k_plan_school, dim3(blocks_for(c->L.nSchools)), dim3(256), 0, c->stream, a);
    const uint32_t qp2 = p->quarantine_policy > 2 ? 1u : 0u;
    hipLaunchKernelGGL(k_plan_agents, dim3(gN), dim3(256), 0, c->stream, a, qp2);
    hipLaunchKernelGGL(k_plan_locs, dim3(persistent_blocks(c)), dim3(256), 0, c->stream, a, qp2);
    if (c->nUnderage)
        hipLaunchKernelGGL(k_plan_underage, dim3(blocks_for(c->nUnderage)), dim3(256), 0, c->stream, a,
                           (const uint32_t*)c->underageList, c->nUnderage);
    if (c->nWorkAge) {
        hipLaunchKernelGGL(k_plan_nowork_claim, dim3(blocks_for(c->nWorkAge)), dim3(256), 0, c->stream, a,
                           (const uint32_t*)c->workAgeList, c->nWorkAge);
        hipLaunchKernelGGL(k_plan_nowork_resolve, dim3(blocks_for(c->nWorkAge)), dim3(256), 0, c->stream, a,
                           (const uint32_t*)c->workAgeList, c->nWorkAge);
    }
    int r = check_launch(c, "plan_locations");
    if (r) return r;
    // the day's last movement step left every agent where it is now: the
    // counts stand, the sums are rebuilt
    hipLaunchKernelGGL(k_zero_sums, dim3(persistent_blocks(c)), dim3(256), 0, c->stream, c->L, c->T.nVar);
    hipLaunchKernelGGL(k_scatter_sums, dim3(blocks_for(c->A.n)), dim3(256), 0, c->stream, c->A, c->L, c->T);
    return check_launch(c, "rebuild_sums");
}

int pansim_testing_day(pansim_ctx* c, uint32_t ts, uint32_t dt, const double probs[6], uint32_t delay, uint32_t qlen,
                       uint32_t usePCR, uint32_t out[3]) {
    if (!c || !probs) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    TestArgs a;
    a.A = c->A;
    a.L = c->L;
    a.T = c->T;
    a.seed = c->seed;
    a.replica = c->replica;
    a.ts = ts;
    a.dt = dt;
    a.delay = delay;
    a.qlen = qlen;
    a.usePCR = usePCR;
    a.pRandom = probs[0];
    a.pHome = probs[1];
    a.pWork = probs[2];
    a.pSchool = probs[3];
    a.pHospital = probs[4];
    a.pNursery = probs[5];
    a.counters = c->counters;
    HIP_OR_FAIL(c, hipMemsetAsync(c->counters, 0, 3 * sizeof(unsigned long long), c->stream));
    hipLaunchKernelGGL(k_clear_flags, dim3(persistent_blocks(c)), dim3(256), 0, c->stream, c->L);
    hipLaunchKernelGGL(k_flag_locations, dim3(blocks_for(c->A.n)), dim3(256), 0, c->stream, a);
    hipLaunchKernelGGL(k_do_testing, dim3(blocks_for(c->A.n)), dim3(256), 0, c->stream, a);
    int r = check_launch(c, "testing");
    if (r) return r;
    HIP_OR_FAIL(c, hipMemcpyAsync(c->hostCounters, c->counters, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                  c->stream));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    if (out)
        for (int i = 0; i < 3; ++i) out[i] = (uint32_t)c->hostCounters[i];
    return 0;
}

int pansim_progression_day(pansim_ctx* c, uint32_t ts, uint32_t dt) {
    if (!c) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    c->stateEpoch++;
    hipLaunchKernelGGL(k_progression, dim3(blocks_for(c->A.n)), dim3(256), 0, c->stream, c->A, c->T, c->seed, c->replica,
                       ts, dt);
    return check_launch(c, "progression");
}

int pansim_immunization_day(pansim_ctx* c, uint32_t ts, uint32_t dt, const float prot[60], const int32_t sim[7],
                            const float simMul[7]) {
    if (!c || !prot || !sim || !simMul) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    ImmArgs a;
    a.A = c->A;
    a.T = c->T;
    a.ts = ts;
    a.dt = dt;
    // prot = level[10] (inf), waning[10] (inf), level[10] (sympt), waning[10], level[10] (hosp), waning[10]
    for (int e = 0; e < 5; ++e)
        for (int s = 0; s < 2; ++s)
            for (int comp = 0; comp < 3; ++comp) {
                a.eff[e][s * 3 + comp][0] = prot[comp * 20 + 10 + 2 * e + s];
                a.eff[e][s * 3 + comp][1] = prot[comp * 20 + 2 * e + s];
            }
    for (int v = 0; v < MAXV; ++v) {
        a.sim[v] = sim[v];
        a.simMul[v] = simMul[v];
    }
    hipLaunchKernelGGL(k_immunization, dim3(blocks_for(c->A.n)), dim3(256), 0, c->stream, a);
    return check_launch(c, "immunization");
}

int pansim_vaccination_init(pansim_ctx* c, const int32_t order[10], const float group_level[10]) {
    if (!c || !order || !group_level) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    // the willingness of group g is vaccinationGroupLevel[levelOf[g]], as the
    // reference's category lambdas capture it (immunization.h:238-322)
    static const int levelOf[10] = {0, 1, 3, 4, 5, 7, 2, 6, 8, 9};
    VaccCategoryArgs a;
    a.A = c->A;
    a.L = c->L;
    a.seed = c->seed;
    a.replica = c->replica;
    a.nPasses = 0;
    for (int o = 1; o <= 10; ++o)
        for (int g = 0; g < 10; ++g)
            if (order[g] == o) a.passes[a.nPasses++] = (uint8_t)g;
    for (int g = 0; g < 10; ++g) {
        if (order[g] > 10) return fail(c, PANSIM_E_ARG, "immunizationOrder values have to be at most 10");
        a.round[g] = (uint8_t)(order[g] > 0 ? order[g] : 0);
        a.level[g] = group_level[levelOf[g]];
    }
    hipLaunchKernelGGL(k_vacc_categories, dim3(blocks_for(c->A.n)), dim3(256), 0, c->stream, a);
    return check_launch(c, "vaccination_init");
}

int pansim_age_group_counts(pansim_ctx* c, uint32_t out[10]) {
    if (!c || !out) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    HIP_OR_FAIL(c, hipMemsetAsync(c->counters, 0, 10 * sizeof(unsigned long long), c->stream));
    hipLaunchKernelGGL(k_age_groups, dim3(persistent_blocks(c)), dim3(256), 0, c->stream, c->A, c->counters);
    int r = check_launch(c, "age_group_counts");
    if (r) return r;
    HIP_OR_FAIL(c, hipMemcpyAsync(c->hostCounters, c->counters, 10 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                  c->stream));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    for (int g = 0; g < 10; ++g) out[g] = (uint32_t)c->hostCounters[g];
    return 0;
}

static VaccDayArgs vacc_day_args(pansim_ctx* c, uint32_t ts, uint32_t dt) {
    VaccDayArgs a;
    a.A = c->A;
    a.T = c->T;
    a.seed = c->seed;
    a.replica = c->replica;
    a.tsKey = ts;
    a.ts = ts == 0 ? 1u : ts;
    a.spd = 24u * 60u / dt;
    a.round = a.draw = a.anyRound = 0;
    for (int g = 0; g < 10; ++g) a.prob[g] = 0.0f;
    return a;
}

int pansim_booster_round(pansim_ctx* c, uint32_t ts, uint32_t dt, uint32_t round, uint32_t available,
                         const float age_group_frac[10]) {
    if (!c || !age_group_frac || dt == 0) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    if (available == 0) return 0;
    VaccDayArgs a = vacc_day_args(c, ts, dt);
    a.round = round;
    a.draw = round;
    HIP_OR_FAIL(c, hipMemsetAsync(c->counters, 0, 10 * sizeof(unsigned long long), c->stream));
    hipLaunchKernelGGL(k_booster_count, dim3(persistent_blocks(c)), dim3(256), 0, c->stream, a, c->counters);
    int r = check_launch(c, "booster_count");
    if (r) return r;
    HIP_OR_FAIL(c, hipMemcpyAsync(c->hostCounters, c->counters, 10 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                  c->stream));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    for (int g = 0; g < 10; ++g) {
        const uint32_t count = (uint32_t)c->hostCounters[g];
        const uint32_t availG = (uint32_t)((float)available * age_group_frac[g]);
        a.prob[g] = count < availG ? 1.0f : (float)availG / (float)count;
    }
    hipLaunchKernelGGL(k_booster_apply, dim3(blocks_for(c->A.n)), dim3(256), 0, c->stream, a);
    return check_launch(c, "booster_apply");
}

int pansim_vaccinate(pansim_ctx* c, uint32_t ts, uint32_t dt, uint32_t available, uint32_t* current_category,
                     uint32_t* immunized_today) {
    if (!c || !current_category || dt == 0) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    VaccDayArgs a = vacc_day_args(c, ts, dt);
    uint32_t& cat = *current_category;
    auto count_round = [&](uint32_t cc, uint32_t any, uint32_t* out) -> int {
        a.round = cc;
        a.anyRound = any;
        HIP_OR_FAIL(c, hipMemsetAsync(c->counters, 0, sizeof(unsigned long long), c->stream));
        hipLaunchKernelGGL(k_vacc_count, dim3(persistent_blocks(c)), dim3(256), 0, c->stream, a, c->counters);
        int r = check_launch(c, "vaccination_count");
        if (r) return r;
        HIP_OR_FAIL(c, hipMemcpyAsync(c->hostCounters, c->counters, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                      c->stream));
        HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
        *out = (uint32_t)c->hostCounters[0];
        return 0;
    };
    while (available > 0 && cat < 10) {
        uint32_t count = 0;
        while (count == 0 && cat <= 10) {
            const uint32_t cc = cat + 1;  // rounds start at 1; round 11 counts every agent
            int r = count_round(cc, cc == 11 ? 1u : 0u, &count);
            if (r) return r;
            if (count == 0) cat++;
        }
        a.round = cat + 1;
        a.anyRound = 0;
        a.draw = 32u + cat;
        a.prob[0] = count < available ? 1.0f : (float)available / (float)count;
        hipLaunchKernelGGL(k_vacc_apply, dim3(blocks_for(c->A.n)), dim3(256), 0, c->stream, a);
        int r = check_launch(c, "vaccination_apply");
        if (r) return r;
        available = count < available ? available - count : 0;
    }
    if (immunized_today) {
        HIP_OR_FAIL(c, hipMemsetAsync(c->counters, 0, sizeof(unsigned long long), c->stream));
        hipLaunchKernelGGL(k_count_immunized, dim3(persistent_blocks(c)), dim3(256), 0, c->stream, c->A, a.ts, c->counters);
        int r = check_launch(c, "count_immunized");
        if (r) return r;
        HIP_OR_FAIL(c, hipMemcpyAsync(c->hostCounters, c->counters, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                      c->stream));
        HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
        *immunized_today = (uint32_t)c->hostCounters[0];
    }
    return 0;
}

int pansim_other_disease_day(pansim_ctx* c, uint32_t ts, uint32_t dt, float probabilityMul) {
    if (!c) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    c->stateEpoch++;
    hipLaunchKernelGGL(k_other_disease, dim3(blocks_for(c->A.n)), dim3(256), 0, c->stream, c->A, c->T, c->seed,
                       c->replica, ts, dt, probabilityMul);
    return check_launch(c, "other_disease");
}

int pansim_count_hospitalized_covid(pansim_ctx* c, uint32_t ts, uint32_t dueWith, uint32_t* out) {
    if (!c || !out) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    HIP_OR_FAIL(c, hipMemsetAsync(c->counters, 0, sizeof(unsigned long long), c->stream));
    hipLaunchKernelGGL(k_count_covid_hospital, dim3(blocks_for(c->A.n)), dim3(256), 0, c->stream, c->A, c->L, c->T, ts,
                       dueWith, c->counters);
    int r = check_launch(c, "count_hospitalized");
    if (r) return r;
    HIP_OR_FAIL(c, hipMemcpyAsync(c->hostCounters, c->counters, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    *out = (uint32_t)c->hostCounters[0];
    return 0;
}

int pansim_count_variants(pansim_ctx* c, uint32_t* out) {
    if (!c || !out) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    if (c->varEpoch == c->stateEpoch) {  // no state change since the statistics pass counted them
        for (uint32_t v = 0; v < c->T.nVar; ++v) out[v] = c->varCache[v];
        return 0;
    }
    HIP_OR_FAIL(c, hipMemsetAsync(c->counters, 0, MAXV * sizeof(unsigned long long), c->stream));
    hipLaunchKernelGGL(k_count_variants, dim3(blocks_for(c->A.n)), dim3(256), 0, c->stream, c->A, c->T, c->counters);
    int r = check_launch(c, "count_variants");
    if (r) return r;
    HIP_OR_FAIL(c, hipMemcpyAsync(c->hostCounters, c->counters, MAXV * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                  c->stream));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    for (uint32_t v = 0; v < c->T.nVar; ++v) out[v] = c->varCache[v] = (uint32_t)c->hostCounters[v];
    c->varEpoch = c->stateEpoch;
    return 0;
}

int pansim_stats_day(pansim_ctx* c, uint32_t ts, uint32_t dt, const uint32_t test[3], uint32_t immunizedToday,
                     uint32_t* out, size_t cap, size_t* len) {
    if (!c || !test) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    HIP_OR_FAIL(c, hipMemsetAsync(c->counters, 0, S_COUNT * sizeof(unsigned long long), c->stream));
    hipLaunchKernelGGL(k_stats, dim3(persistent_blocks(c)), dim3(256), 0, c->stream, c->A, c->L, c->T, ts, dt, c->counters);
    int r = check_launch(c, "stats");
    if (r) return r;
    HIP_OR_FAIL(c, hipMemcpyAsync(c->hostCounters, c->counters, S_COUNT * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                  c->stream));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    const unsigned long long* h = c->hostCounters;
    const uint32_t NS = c->T.nStates, V = c->T.nVar;
    for (uint32_t k = 0; k < V; ++k) c->varCache[k] = (uint32_t)h[S_VARINF + k];
    c->varEpoch = c->stateEpoch;
    std::vector<uint32_t> v;
    for (uint32_t s = 0; s < NS; ++s) v.push_back((uint32_t)h[S_STATE + s]);
    v.push_back((uint32_t)h[S_HOSP]);
    v.push_back(test[0]);
    v.push_back(test[1]);
    v.push_back(test[2]);
    v.push_back((uint32_t)h[S_Q]);
    v.push_back((uint32_t)h[S_QT]);
    v.push_back((uint32_t)h[S_NQ]);
    if (V > 1) {
        for (uint32_t k = 0; k + 1 < V; ++k) {
            double pct = double(h[S_VARINF + k + 1]) / double(h[S_ALLINF]) * 100.0;
            v.push_back(pct != pct ? 0u : (uint32_t)(int64_t)pct);
        }
    } else {
        v.push_back(0);
    }
    uint32_t d10 = NS > 10 ? v[10] : 0, d11 = NS > 11 ? v[11] : 0;
    v.push_back((uint32_t)h[S_STAYED] - d10 - d11);
    v.push_back(immunizedToday);
    v.push_back((uint32_t)h[S_NI]);
    v.push_back((uint32_t)h[S_INF]);
    v.push_back((uint32_t)h[S_REINF]);
    v.push_back((uint32_t)h[S_BSTR]);
    for (uint32_t k = 0; k < V; ++k) v.push_back((uint32_t)(h[S_IMM + k] >> 24));
    uint32_t hcw = c->hcwCount > 0 ? c->hcwCount : 1u;
    v.push_back(((uint32_t)h[S_HCI] * 100u) / hcw);
    v.push_back((uint32_t)h[S_HCE]);
    for (uint32_t k = 0; k < V; ++k) v.push_back((uint32_t)h[S_INFV + k]);
    v.push_back((uint32_t)h[S_INFH]);
    v.push_back((uint32_t)h[S_VNI]);
    if (len) *len = v.size();
    if (out)
        for (size_t i = 0; i < cap && i < v.size(); ++i) out[i] = v[i];
    return 0;
}

int pansim_set_open_by_type(pansim_ctx* c, const uint32_t* types, uint32_t n, uint32_t open) {
    if (!c || (n && !types) || n > 16) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    uint32_t* dtypes = (uint32_t*)(c->counters + S_COUNT);  // 16 u64 words of scratch
    HIP_OR_FAIL(c, hipMemcpyAsync(dtypes, types, n * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_set_open, dim3(persistent_blocks(c)), dim3(256), 0, c->stream, c->L, dtypes, n, (uint8_t)(open ? 1 : 0));
    int r = check_launch(c, "set_open_by_type");
    if (r) return r;
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    return 0;
}

int pansim_scale_infectiousness(pansim_ctx* c, double factor, uint32_t divide) {
    if (!c) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    hipLaunchKernelGGL(k_scale_inf, dim3(persistent_blocks(c)), dim3(256), 0, c->stream, c->L, c->T.lt.home, factor, divide,
                       c->typeMul);
    return check_launch(c, "scale_infectiousness");
}

int pansim_expose_mutation(pansim_ctx* c, uint32_t ts, uint32_t v, double fraction) {
    if (!c || v >= c->T.nVar) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    c->stateEpoch++;
    hipLaunchKernelGGL(k_expose_mutation, dim3(blocks_for(c->A.n)), dim3(256), 0, c->stream, c->A, c->T, c->seed, c->replica,
                       ts, v, fraction);
    return check_launch(c, "expose_mutation");
}

int pansim_upload_schedules(pansim_ctx* c, const uint32_t* off, uint32_t nOff, const pansim_event* ev, uint32_t nEv) {
    if (!c || !off || nOff != c->evOffCap) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    if (nEv > c->evCap) return fail(c, PANSIM_E_ARG, "schedule grew beyond its reserved capacity");
    if (!offsets_ok(off, nOff, nEv)) return fail(c, PANSIM_E_ARG, "schedule offsets out of range");
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    HIP_OR_FAIL(c, hipMemcpy(c->evOffDev, off, nOff * sizeof(uint32_t), hipMemcpyHostToDevice));
    if (nEv) HIP_OR_FAIL(c, hipMemcpy(c->evDev, ev, nEv * sizeof(pansim_event), hipMemcpyHostToDevice));
    c->nEvents = nEv;
    return compute_windows(c);
}

int pansim_download_location_sums(pansim_ctx* c, uint32_t v, int64_t* out, size_t n) {
    if (!c || !out || v >= c->T.nVar || n < c->L.n) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    if (v == 0) {
        // the LocDyn records copied whole in chunks, the sums picked out on the host
        const uint32_t chunk = 1u << 20;
        std::vector<LocDyn> tmp(std::min(c->L.n, chunk));
        for (uint32_t first = 0; first < c->L.n; first += chunk) {
            const uint32_t m = std::min(chunk, c->L.n - first);
            HIP_OR_FAIL(c, hipMemcpy(tmp.data(), c->L.dyn + first, (size_t)m * sizeof(LocDyn), hipMemcpyDeviceToHost));
            for (uint32_t l = 0; l < m; ++l) out[first + l] = tmp[l].sum0;
        }
    } else
        HIP_OR_FAIL(c, hipMemcpy(out, c->L.sumX + (size_t)(v - 1) * c->L.n, c->L.n * sizeof(int64_t), hipMemcpyDeviceToHost));
    return 0;
}

// A field in device (storage) order.
static int download_field_stored(pansim_ctx* c, const char* name, uint32_t ts, uint32_t* out, size_t n) {
    if (!c || !name || !out) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    const uint32_t N = c->A.n, NL = c->L.n;
    std::string f = name;
    auto get32 = [&](const uint32_t* src, uint32_t cnt) -> int {
        if (n < cnt) return fail(c, PANSIM_E_ARG, "output buffer too small");
        HIP_OR_FAIL(c, hipMemcpy(out, src, cnt * sizeof(uint32_t), hipMemcpyDeviceToHost));
        return 0;
    };
    auto get8 = [&](const void* src, uint32_t cnt, bool isSigned) -> int {
        if (n < cnt) return fail(c, PANSIM_E_ARG, "output buffer too small");
        std::vector<uint8_t> tmp(cnt);
        HIP_OR_FAIL(c, hipMemcpy(tmp.data(), src, cnt, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < cnt; ++i) out[i] = isSigned ? (uint32_t)(int32_t)(int8_t)tmp[i] : tmp[i];
        return 0;
    };
    // one member of an array of records: the records are copied whole, in
    // chunks of at most 2^20 (contiguous transfers; a strided 2D copy of
    // millions of rows is slow), and the member is picked out on the host
    auto getMember = [&](const void* base, size_t pitch, size_t offset, size_t width, uint32_t cnt) -> int {
        if (n < cnt) return fail(c, PANSIM_E_ARG, "output buffer too small");
        const uint32_t chunk = 1u << 20;
        std::vector<uint8_t> tmp((size_t)std::min(cnt, chunk) * pitch);
        for (uint32_t first = 0; first < cnt; first += chunk) {
            const uint32_t m = std::min(chunk, cnt - first);
            HIP_OR_FAIL(c, hipMemcpy(tmp.data(), (const uint8_t*)base + (size_t)first * pitch, (size_t)m * pitch,
                                     hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < m; ++i) {
                const uint8_t* p = &tmp[(size_t)i * pitch + offset];
                uint32_t& o = out[first + i];
                if (width == 4) memcpy(&o, p, 4);
                else if (width == 2) {
                    uint16_t v;
                    memcpy(&v, p, 2);
                    o = v;
                } else o = *p;
            }
        }
        return 0;
    };
#define AGENT_MEMBER(m) getMember(c->A.rec, sizeof(AgentRec), offsetof(AgentRec, m), sizeof(AgentRec::m), N)
#define LOC_MEMBER(arr, T, m) getMember(c->L.arr, sizeof(T), offsetof(T, m), sizeof(T::m), NL)
    auto getFlag = [&](uint8_t bit) -> int {
        int r = AGENT_MEMBER(flags);
        if (r) return r;
        for (uint32_t i = 0; i < N; ++i) out[i] = (out[i] & bit) ? 1u : 0u;
        return 0;
    };
    const Agents& A = c->A;
    if (f == "state") return AGENT_MEMBER(state);
    if (f == "variant") return AGENT_MEMBER(variant);
    if (f == "days") {
        if (n < N) return PANSIM_E_ARG;
        std::vector<int16_t> tmp(N);
        HIP_OR_FAIL(c, hipMemcpy(tmp.data(), A.days, N * sizeof(int16_t), hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < N; ++i) out[i] = (uint32_t)(int32_t)tmp[i];
        return 0;
    }
    if (f == "location") return get32(A.location, N);
    if (f == "csrBegin") return AGENT_MEMBER(csrBegin);
    if (f == "csrCount") return AGENT_MEMBER(csrCount);
    if (f == "possibleLocations") return get32(A.possLoc, c->nCsr);
    if (f == "stepsUntilMove") {
        if (n < N) return fail(c, PANSIM_E_ARG, "output buffer too small");
        std::vector<uint16_t> tmp(N);
        HIP_OR_FAIL(c, hipMemcpy(tmp.data(), A.wake, N * sizeof(uint16_t), hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < N; ++i) out[i] = tmp[i] == WAKE_NONE ? NEVER : c->wakeDay + tmp[i] - ts;
        return 0;
    }
    if (f == "diagnosed") return getFlag(F_DIAGNOSED);
    if (f == "quarantined") return getFlag(F_QUARANTINED);
    if (f == "stayedHome") return getFlag(F_STAYED_HOME);
    if (f == "noWorkAgent") return getFlag(F_NO_WORK);
    if (f == "lastTest") return get32(A.lastTest, N);
    if (f == "susceptible0") return get32((const uint32_t*)A.susc, N);
    if (f == "infectedTimestamp") return get32(A.infTs, N);
    if (f == "infectedLocation") return get32(A.infLoc, N);
    if (f == "worstStateTimestamp") return get32(A.worstTs, N);
    if (f == "worstStateEndTimestamp") return get32(A.worstEndTs, N);
    if (f == "diagnosedTimestamp") return AGENT_MEMBER(diagTs);
    if (f == "quarantinedTimestamp") return get32(A.quarTs, N);
    if (f == "quarantinedUntilTimestamp") return AGENT_MEMBER(quarUntil);
    if (f == "daysInQuarantine") return get32(A.daysInQ, N);
    if (f == "hospitalizedTimestamp") return AGENT_MEMBER(hospTs);
    if (f == "hospitalizedUntilTimestamp") return AGENT_MEMBER(hospUntil);
    if (f == "recoveredTimestamp") return get32(A.recTs, N);
    if (f == "immunizationTimestamp") return get32(A.immTs, N);
    if (f == "immunizationCount") return get8(A.immCount, N, false);
    if (f == "vaccinationRound") return get8(A.immRound, N, false);
    if (f == "infectedCount") return get8(A.infCount, N, false);
    if (f == "statsVariant") return get8(A.statVar, N, false);
    if (f == "worstState") return get8(A.worstState, N, true);
    if (f == "locQuarantineUntil") return LOC_MEMBER(rec, LocRec, quarUntil);
    if (f == "locClosedUntil") return LOC_MEMBER(rec, LocRec, closedUntil);
    if (f == "locOpen") return LOC_MEMBER(rec, LocRec, open);
    if (f == "locInfectiousness") return get32((const uint32_t*)c->L.infRaw, NL);
    if (f == "locCount") return LOC_MEMBER(dyn, LocDyn, count);
#undef AGENT_MEMBER
#undef LOC_MEMBER
    return fail(c, PANSIM_E_ARG, "unknown field " + f);
}

// A field in input order: agent fields are moved from the storage order
// back to the input's; the possible-location lists are rebuilt in the
// input's layout.
int pansim_download_field(pansim_ctx* c, const char* name, uint32_t ts, uint32_t* out, size_t n) {
    if (!c || !name || !out) return PANSIM_E_ARG;
    const std::string f = name;
    const uint32_t N = c->A.n;
    const std::vector<uint32_t>& perm = c->perm;
    if (f == "possibleLocations" || f == "csrBegin") {
        if (n < (f == "csrBegin" ? (size_t)N : (size_t)c->nCsr)) return fail(c, PANSIM_E_ARG, "output buffer too small");
        std::vector<uint32_t> cnt(N), beg(N), inBeg(N), pl(c->nCsr);
        int r = download_field_stored(c, "csrCount", ts, cnt.data(), N);
        if (!r) r = download_field_stored(c, "csrBegin", ts, beg.data(), N);
        if (!r) r = download_field_stored(c, "possibleLocations", ts, pl.data(), pl.size());
        if (r) return r;
        std::vector<uint32_t> cntIn(N);
        for (uint32_t i = 0; i < N; ++i) cntIn[perm[i]] = cnt[i];
        uint32_t acc = 0;
        for (uint32_t k = 0; k < N; ++k) {
            inBeg[k] = acc;
            acc += cntIn[k];
        }
        if (f == "csrBegin") {
            std::copy(inBeg.begin(), inBeg.end(), out);
            return 0;
        }
        for (uint32_t i = 0; i < N; ++i) std::copy(pl.begin() + beg[i], pl.begin() + beg[i] + cnt[i], out + inBeg[perm[i]]);
        return 0;
    }
    const int r = download_field_stored(c, name, ts, out, n);
    if (r) return r;
    static const char* const locFields[] = {"locQuarantineUntil", "locClosedUntil", "locOpen", "locInfectiousness",
                                            "locCount"};
    for (const char* lf : locFields)
        if (f == lf) return 0;
    std::vector<uint32_t> stored(out, out + N);
    for (uint32_t i = 0; i < N; ++i) out[perm[i]] = stored[i];
    return 0;
}

int pansim_device_expf(const float* in, float* out, size_t n) {
    if (n == 0) return 0;
    if (!in || !out) return PANSIM_E_ARG;
    float *din = nullptr, *dout = nullptr;
    if (hipMalloc(&din, n * sizeof(float)) != hipSuccess) return PANSIM_E_DEVICE;
    if (hipMalloc(&dout, n * sizeof(float)) != hipSuccess) {
        hipFree(din);
        return PANSIM_E_DEVICE;
    }
    hipMemcpy(din, in, n * sizeof(float), hipMemcpyHostToDevice);
    hipLaunchKernelGGL(k_expf, dim3(blocks_for(n)), dim3(256), 0, 0, din, dout, n);
    hipError_t e = hipMemcpy(out, dout, n * sizeof(float), hipMemcpyDeviceToHost);
    hipFree(din);
    hipFree(dout);
    return e == hipSuccess ? 0 : PANSIM_E_DEVICE;
}

int pansim_device_log(const double* in, double* out, size_t n) {
    if (n == 0) return 0;
    if (!in || !out) return PANSIM_E_ARG;
    double *din = nullptr, *dout = nullptr;
    if (hipMalloc(&din, n * sizeof(double)) != hipSuccess) return PANSIM_E_DEVICE;
    if (hipMalloc(&dout, n * sizeof(double)) != hipSuccess) {
        hipFree(din);
        return PANSIM_E_DEVICE;
    }
    hipMemcpy(din, in, n * sizeof(double), hipMemcpyHostToDevice);
    hipLaunchKernelGGL(k_log, dim3(blocks_for(n)), dim3(256), 0, 0, din, dout, n);
    hipError_t e = hipMemcpy(out, dout, n * sizeof(double), hipMemcpyDeviceToHost);
    hipFree(din);
    hipFree(dout);
    return e == hipSuccess ? 0 : PANSIM_E_DEVICE;
}

int pansim_device_philox(const uint32_t* ctr, uint32_t k0, uint32_t k1, uint32_t* out, size_t n) {
    if (n == 0) return 0;
    if (!ctr || !out) return PANSIM_E_ARG;
    uint32_t *din = nullptr, *dout = nullptr;
    if (hipMalloc(&din, n * 16) != hipSuccess) return PANSIM_E_DEVICE;
    if (hipMalloc(&dout, n * 16) != hipSuccess) {
        hipFree(din);
        return PANSIM_E_DEVICE;
    }
    hipMemcpy(din, ctr, n * 16, hipMemcpyHostToDevice);
    hipLaunchKernelGGL(k_philox, dim3(blocks_for(n)), dim3(256), 0, 0, din, k0, k1, dout, n);
    hipError_t e = hipMemcpy(out, dout, n * 16, hipMemcpyDeviceToHost);
    hipFree(din);
    hipFree(dout);
    return e == hipSuccess ? 0 : PANSIM_E_DEVICE;
}

// Philox blocks without any allocation: the counters and the blocks are
// staged in `scratch`, a file-scope device array (each GPU holds its own copy
// until the process ends).  Its address on the current device comes from
// hipGetSymbolAddress at the first call on that device and is kept per
// device; host code uses only that address, never the array's name.  Calls
// that overlap on one GPU share the array and race.
static __device__ __align__(256) uint32_t scratch[2 * 4 * PANSIM_PHILOX_NOALLOC_MAX];
static std::mutex g_scratch_mu;
static uint32_t* g_scratch_ptr[64];

int pansim_device_philox_noalloc(const uint32_t* ctr, uint32_t k0, uint32_t k1, uint32_t* out, size_t n) {
    if (n == 0) return 0;
    if (!ctr || !out) return PANSIM_E_ARG;
    if (n > PANSIM_PHILOX_NOALLOC_MAX) {
        g_create_error = "pansim_device_philox_noalloc: " + std::to_string(n) + " counters exceed the scratch of " +
                         std::to_string(PANSIM_PHILOX_NOALLOC_MAX);
        return PANSIM_E_ARG;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return PANSIM_E_DEVICE;
    uint32_t* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_ptr[dev]) {
            void* p = nullptr;
            hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                g_create_error = std::string("pansim_device_philox_noalloc: hipGetSymbolAddress failed: ") +
                                 hipGetErrorString(e);
                return PANSIM_E_DEVICE;
            }
            g_scratch_ptr[dev] = (uint32_t*)p;
        }
        base = g_scratch_ptr[dev];
    }
    uint32_t* din = base;
    uint32_t* dout = base + 4 * (size_t)PANSIM_PHILOX_NOALLOC_MAX;
    if (hipMemcpy(din, ctr, n * 16, hipMemcpyHostToDevice) != hipSuccess) return PANSIM_E_DEVICE;
    hipLaunchKernelGGL(k_philox, dim3(blocks_for(n)), dim3(256), 0, 0, din, k0, k1, dout, n);
    if (hipGetLastError() != hipSuccess) return PANSIM_E_DEVICE;
    return hipMemcpy(out, dout, n * 16, hipMemcpyDeviceToHost) == hipSuccess ? 0 : PANSIM_E_DEVICE;
}

// ------------------------------------------------------------ hazard gradient
int pansim_infection_dump_enable(pansim_ctx* c, uint32_t on) {
    if (!c) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    if (on && !c->dumpNew) {
        const size_t N = c->A.n, words = ((size_t)c->L.n + 31) / 32;
        int r = 0;
        if ((r = dalloc(c, &c->dumpNew, N)) || (r = dalloc(c, &c->dumpNewLoc, N)) || (r = dalloc(c, &c->dumpCnt, 4)) ||
            (r = dalloc(c, &c->dumpFlags, words)) || (r = dalloc(c, &c->dumpInfAgent, N)) ||
            (r = dalloc(c, &c->dumpInfLoc, N)) || (r = dalloc(c, &c->dumpInfVal, N))) {
            c->dumpNew = nullptr;  // stays "never enabled" (the partial buffers are freed with the context)
            return r;
        }
        HIP_OR_FAIL(c, hipMemsetAsync(c->dumpCnt, 0, 4 * sizeof(uint32_t), c->stream));
        HIP_OR_FAIL(c, hipMemsetAsync(c->dumpFlags, 0, words * sizeof(uint32_t), c->stream));
    }
    c->dumpOn = on != 0;
    return 0;
}

int pansim_infection_dump_collect(pansim_ctx* c, uint32_t* new_agent, uint32_t* new_loc, size_t cap_new, size_t* n_new,
                                  uint32_t* inf_agent, uint32_t* inf_loc, float* inf_value, size_t cap_inf,
                                  size_t* n_inf) {
    if (!c || !n_new || !n_inf) return PANSIM_E_ARG;
    if (!c->dumpNew) return fail(c, PANSIM_E_ARG, "infection_dump_collect: the dump was never enabled");
    if (int fr = flush_resolve(c)) return fr;
    const unsigned gN = blocks_for(c->A.n);
    HIP_OR_FAIL(c, hipMemsetAsync(c->dumpCnt + 1, 0, sizeof(uint32_t), c->stream));
    hipLaunchKernelGGL(k_dump_mark, dim3(gN), dim3(256), 0, c->stream, c->A, c->dumpNew, c->dumpCnt, c->dumpNewLoc,
                       c->dumpFlags);
    hipLaunchKernelGGL(k_dump_infectious, dim3(gN), dim3(256), 0, c->stream, c->A, c->T, c->dumpFlags, c->dumpInfAgent,
                       c->dumpInfLoc, c->dumpInfVal, c->dumpCnt + 1);
    if (int r = check_launch(c, "infection_dump_collect")) return r;
    uint32_t cnt[2];
    HIP_OR_FAIL(c, hipMemcpyAsync(cnt, c->dumpCnt, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    HIP_OR_FAIL(c, hipMemsetAsync(c->dumpFlags, 0, ((size_t)c->L.n + 31) / 32 * sizeof(uint32_t), c->stream));
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    *n_new = cnt[0];
    *n_inf = cnt[1];
    const bool query = !new_agent && !new_loc && !inf_agent && !inf_loc && !inf_value;
    if (query) return 0;  // counts only; the record is kept
    if (!new_agent || !new_loc || !inf_agent || !inf_loc || !inf_value || cap_new < cnt[0] || cap_inf < cnt[1])
        return fail(c, PANSIM_E_ARG, "infection_dump_collect: output arrays missing or too small");
    if (cnt[0]) {
        HIP_OR_FAIL(c, hipMemcpy(new_agent, c->dumpNew, cnt[0] * sizeof(uint32_t), hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < cnt[0]; ++k) new_agent[k] = c->perm[new_agent[k]];
        HIP_OR_FAIL(c, hipMemcpy(new_loc, c->dumpNewLoc, cnt[0] * sizeof(uint32_t), hipMemcpyDeviceToHost));
    }
    if (cnt[1]) {
        HIP_OR_FAIL(c, hipMemcpy(inf_agent, c->dumpInfAgent, cnt[1] * sizeof(uint32_t), hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < cnt[1]; ++k) inf_agent[k] = c->perm[inf_agent[k]];
        HIP_OR_FAIL(c, hipMemcpy(inf_loc, c->dumpInfLoc, cnt[1] * sizeof(uint32_t), hipMemcpyDeviceToHost));
        HIP_OR_FAIL(c, hipMemcpy(inf_value, c->dumpInfVal, cnt[1] * sizeof(float), hipMemcpyDeviceToHost));
    }
    HIP_OR_FAIL(c, hipMemsetAsync(c->dumpCnt, 0, sizeof(uint32_t), c->stream));
    return 0;
}

int pansim_set_location_type_multipliers(pansim_ctx* c, const float* m, uint32_t n) {
    if (!c || (n && !m) || n > HAZARD_TYPES) return PANSIM_E_ARG;
    std::vector<float> v(HAZARD_TYPES, 1.0f);
    for (uint32_t t = 0; t < n; ++t) v[t] = m[t];
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    HIP_OR_FAIL(c, hipMemcpy(c->typeMul, v.data(), HAZARD_TYPES * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_loc_infm, dim3(persistent_blocks(c)), dim3(256), 0, c->stream, c->L, c->typeMul);
    return check_launch(c, "set_location_type_multipliers");
}

int pansim_hazard_enable(pansim_ctx* c, uint32_t on) {
    if (!c) return PANSIM_E_ARG;
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    HIP_OR_FAIL(c, hipMemset(c->hazard, 0, HAZARD_SLOTS * sizeof(double)));
    c->hazardOn = on != 0;
    return 0;
}

int pansim_hazard_set_probe(pansim_ctx* c, double h, uint32_t type) {
    if (!c || !(h >= 0.0 && h < 1.0) || type >= HAZARD_TYPES) return PANSIM_E_ARG;
    c->probeH = (float)h;
    c->probeType = type;
    return 0;
}

int pansim_hazard_read(pansim_ctx* c, double* out, uint32_t n, uint32_t reset) {
    if (!c || !out || n < HAZARD_PROBE0) return PANSIM_E_ARG;
    if (int fr = flush_resolve(c)) return fr;
    HIP_OR_FAIL(c, hipStreamSynchronize(c->stream));
    const uint32_t m = n < HAZARD_SLOTS ? HAZARD_PROBE0 : HAZARD_SLOTS;
    HIP_OR_FAIL(c, hipMemcpy(out, c->hazard, m * sizeof(double), hipMemcpyDeviceToHost));
    if (reset) HIP_OR_FAIL(c, hipMemset(c->hazard, 0, HAZARD_SLOTS * sizeof(double)));
    return 0;
}

// ------------------------------------------------------------ replica batches
static int batch_fail(pansim_batch* b, int code, const std::string& m) {
    if (b) b->err = m;
    return code;
}

void pansim_batch_destroy(pansim_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (size_t r = 0; r < b->reps.size() && r < b->ownStreams.size(); ++r) b->reps[r]->stream = b->ownStreams[r];
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (auto& x : b->samples) {
        (void)hipEventDestroy(x.e0);
        (void)hipEventDestroy(x.e1);
    }
    if (b->dDayStart) (void)hipFree(b->dDayStart);
    if (b->hDayStart) (void)hipHostFree(b->hDayStart);
    if (b->dMove) (void)hipFree(b->dMove);
    if (b->dInf) (void)hipFree(b->dInf);
    if (b->dTables) (void)hipFree(b->dTables);
    if (b->hMove) (void)hipHostFree(b->hMove);
    if (b->hInf) (void)hipHostFree(b->hInf);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char* pansim_batch_error(const pansim_batch* b) { return b ? b->err.c_str() : g_create_error.c_str(); }

int pansim_batch_create(pansim_batch** out, pansim_ctx* const* replicas, uint32_t n) {
    if (!out || !replicas || n == 0) {
        g_create_error = "pansim_batch_create: null argument or no replicas";
        return PANSIM_E_ARG;
    }
    *out = nullptr;
    auto* b = new pansim_batch();
    auto bail = [&](int code, const std::string& m) {
        g_create_error = m;
        pansim_batch_destroy(b);
        return code;
    };
    for (uint32_t r = 0; r < n; ++r) {
        pansim_ctx* c = replicas[r];
        if (!c) return bail(PANSIM_E_ARG, "pansim_batch_create: null replica");
        if (c->dt != replicas[0]->dt || c->device != replicas[0]->device)
            return bail(PANSIM_E_ARG, "pansim_batch_create: replicas differ in time step or device");
        for (uint32_t q = 0; q < r; ++q)
            if (replicas[q] == c) return bail(PANSIM_E_ARG, "pansim_batch_create: a replica appears twice");
    }
    b->dt = replicas[0]->dt;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(PANSIM_E_DEVICE, "pansim_batch_create: hipStreamCreate failed");
    for (uint32_t r = 0; r < n; ++r) {
        pansim_ctx* c = replicas[r];
        if (flush_resolve(c) != 0 || hipStreamSynchronize(c->stream) != hipSuccess)
            return bail(PANSIM_E_DEVICE, "pansim_batch_create: sync failed");
        b->reps.push_back(c);
        b->ownStreams.push_back(c->stream);
        c->stream = b->stream;
        // at rest neither pending-write counter holds anything: restart all at parity 0
        if (hipMemsetAsync(c->S.pendCount, 0, 2 * sizeof(uint32_t), b->stream) != hipSuccess)
            return bail(PANSIM_E_DEVICE, "pansim_batch_create: memset failed");
        c->pendPar = 0;
    }
    // block -> replica tables of every candidate tile size
    std::vector<uint32_t> tables;
    auto addTables = [&](uint32_t tile, uint32_t& off, uint32_t& blocks) {
        std::vector<uint32_t> rep, first(n);
        for (uint32_t r = 0; r < n; ++r) {
            first[r] = (uint32_t)rep.size();
            for (uint32_t t = 0; t < blocks_for(replicas[r]->A.n, tile); ++t) rep.push_back(r);
        }
        off = (uint32_t)tables.size();
        blocks = (uint32_t)rep.size();
        tables.insert(tables.end(), rep.begin(), rep.end());
        tables.insert(tables.end(), first.begin(), first.end());
    };
    for (int k = 0; k < 3; ++k) addTables(MOVE_BLOCK * kMoveItems[k], b->moveOff[k], b->moveBlocks[k]);
    for (int k = 0; k < 2; ++k) addTables(INFECT_BLOCK * kInfectItems[k], b->infOff[k], b->infBlocks[k]);
    const size_t nt = tables.size();
    if (hipMalloc(&b->dMove, n * sizeof(MoveArgs)) != hipSuccess || hipMalloc(&b->dInf, n * sizeof(InfectArgs)) != hipSuccess ||
        hipMalloc(&b->dTables, nt * sizeof(uint32_t)) != hipSuccess)
        return bail(PANSIM_E_DEVICE, "pansim_batch_create: hipMalloc failed");
    if (hipMalloc(&b->dDayStart, sizeof(uint32_t)) != hipSuccess)
        return bail(PANSIM_E_DEVICE, "pansim_batch_create: hipMalloc failed");
    if (hipHostMalloc((void**)&b->hMove, n * sizeof(MoveArgs), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&b->hInf, n * sizeof(InfectArgs), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&b->hDayStart, sizeof(uint32_t), hipHostMallocDefault) != hipSuccess)
        return bail(PANSIM_E_DEVICE, "pansim_batch_create: hipHostMalloc failed");
    if (hipMemcpy(b->dTables, tables.data(), nt * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return bail(PANSIM_E_DEVICE, "pansim_batch_create: table upload failed");
    *out = b;
    return 0;
}

int pansim_batch_set_day(pansim_batch* b, const pansim_step_params* params, const uint32_t* variant_masks, const float* k,
                         const float* seasonality) {
    if (!b || !params || !variant_masks || !k || !seasonality) return PANSIM_E_ARG;
    const uint32_t n = (uint32_t)b->reps.size();
    // the staging buffers may still feed the previous day's copy
    if (hipStreamSynchronize(b->stream) != hipSuccess) return batch_fail(b, PANSIM_E_DEVICE, "pansim_batch_set_day: sync failed");
    for (uint32_t r = 0; r < n; ++r) {
        pansim_ctx* c = b->reps[r];
        if (params[r].time_step != b->dt) return batch_fail(b, PANSIM_E_ARG, "pansim_batch_set_day: time step differs");
        b->hMove[r] = move_args(c, &params[r]);
        b->hInf[r] = infect_args(c, params[r].timestamp, variant_masks[r], k[r], seasonality[r], b->dt);
    }
    const uint32_t spd = 1440u / b->dt;
    *b->hDayStart = params[0].timestamp - params[0].timestamp % spd;
    if (hipMemcpyAsync(b->dMove, b->hMove, n * sizeof(MoveArgs), hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipMemcpyAsync(b->dInf, b->hInf, n * sizeof(InfectArgs), hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipMemcpyAsync(b->dDayStart, b->hDayStart, sizeof(uint32_t), hipMemcpyHostToDevice, b->stream) != hipSuccess)
        return batch_fail(b, PANSIM_E_DEVICE, "pansim_batch_set_day: upload failed");
    b->haveDay = true;
    return 0;
}

static bool batch_hazard(const pansim_batch* b) {
    for (const pansim_ctx* c : b->reps)
        if (c->hazardOn) return true;
    return false;
}

// One batch step: the movement launch of every replica, then the infection
// launch (which also applies the movement's deferred location writes), with
// the tile sizes of candidates mc (movement) and ic (infection).
extern "C++" {
template <uint32_t MI, uint32_t II>
static void launch_batch_sized(pansim_batch* b, const BatchArgs& a, uint32_t mc, uint32_t ic, bool hazard) {
    hipLaunchKernelGGL(k_move_batch<MI>, dim3(b->moveBlocks[mc]), dim3(MOVE_BLOCK), 0, b->stream, a);
    if (hazard) hipLaunchKernelGGL((k_infect_batch<true, II>), dim3(b->infBlocks[ic]), dim3(INFECT_BLOCK), 0, b->stream, a);
    else hipLaunchKernelGGL((k_infect_batch<false, II>), dim3(b->infBlocks[ic]), dim3(INFECT_BLOCK), 0, b->stream, a);
}
}

static void launch_batch_step(pansim_batch* b, uint32_t stepOfDay, uint32_t pendPar, bool hazard, int mc, int ic) {
    const uint32_t n = (uint32_t)b->reps.size();
    BatchArgs a;
    a.move = b->dMove;
    a.infect = b->dInf;
    a.moveRep = b->dTables + b->moveOff[mc];
    a.moveFirst = a.moveRep + b->moveBlocks[mc];
    a.infRep = b->dTables + b->infOff[ic];
    a.infFirst = a.infRep + b->infBlocks[ic];
    a.dayStart = b->dDayStart;
    a.stepOfDay = stepOfDay;
    a.pendPar = pendPar;
    a.nRep = n;
    a.moveBlocks = b->moveBlocks[mc];
    a.infBlocks = b->infBlocks[ic];
    switch (mc * 2 + ic) {
    case 0: launch_batch_sized<kMoveItems[0], kInfectItems[0]>(b, a, 0, 0, hazard); break;
    case 1: launch_batch_sized<kMoveItems[0], kInfectItems[1]>(b, a, 0, 1, hazard); break;
    case 2: launch_batch_sized<kMoveItems[1], kInfectItems[0]>(b, a, 1, 0, hazard); break;
    case 3: launch_batch_sized<kMoveItems[1], kInfectItems[1]>(b, a, 1, 1, hazard); break;
    case 4: launch_batch_sized<kMoveItems[2], kInfectItems[0]>(b, a, 2, 0, hazard); break;
    default: launch_batch_sized<kMoveItems[2], kInfectItems[1]>(b, a, 2, 1, hazard); break;
    }
}

// The autotuner: while no tile sizes are chosen, eager batch steps cycle
// through the six candidate pairs with an event pair around each step (the
// results do not depend on the tile sizes: every agent's work is its own and
// the per-location updates are integer atomics); pansim_batch_run_day then
// picks the pair with the least mean step time before it captures the day.
static int batch_pick_tiles(pansim_batch* b) {
    if (b->moveCfg >= 0) return 0;
    double sum[6] = {0, 0, 0, 0, 0, 0};
    uint32_t cnt[6] = {0, 0, 0, 0, 0, 0};
    if (!b->samples.empty() && hipEventSynchronize(b->samples.back().e1) != hipSuccess)
        return batch_fail(b, PANSIM_E_DEVICE, "autotuner: event sync failed");
    for (auto& x : b->samples) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, x.e0, x.e1) == hipSuccess) {
            sum[x.move * 2 + x.inf] += ms;
            cnt[x.move * 2 + x.inf]++;
        }
        (void)hipEventDestroy(x.e0);
        (void)hipEventDestroy(x.e1);
    }
    b->samples.clear();
    int best = 5;  // without samples: 32 movement and 8 infection agents per thread
    double bestT = 1e300;
    for (int k = 0; k < 6; ++k)
        if (cnt[k] && sum[k] / cnt[k] < bestT) {
            bestT = sum[k] / cnt[k];
            best = k;
        }
    b->moveCfg = best / 2;
    b->infCfg = best % 2;
    return 0;
}

int pansim_batch_step(pansim_batch* b, uint32_t timestamp) {
    if (!b) return PANSIM_E_ARG;
    if (!b->haveDay) return batch_fail(b, PANSIM_E_ARG, "pansim_batch_step: pansim_batch_set_day was not called");
    const uint32_t spd = 1440u / b->dt;
    const uint32_t stepOfDay = timestamp % spd;
    for (pansim_ctx* c : b->reps)
        if (timestamp - stepOfDay != c->wakeDay)
            return batch_fail(b, PANSIM_E_ARG, "pansim_batch_step: pansim_plan_locations must run at the start of every day");
    if (timestamp - stepOfDay != *b->hDayStart)
        return batch_fail(b, PANSIM_E_ARG, "pansim_batch_step: the day's arguments are not this day's");
    for (pansim_ctx* c : b->reps) c->stateEpoch++;
    if (b->moveCfg >= 0) {
        launch_batch_step(b, stepOfDay, b->pendPar, batch_hazard(b), b->moveCfg, b->infCfg);
    } else {
        const int k = (int)(b->samples.size() % 6);
        pansim_batch::TuneSample x{k / 2, k % 2, nullptr, nullptr};
        if (hipEventCreate(&x.e0) != hipSuccess || hipEventCreate(&x.e1) != hipSuccess)
            return batch_fail(b, PANSIM_E_DEVICE, "autotuner: hipEventCreate failed");
        (void)hipEventRecord(x.e0, b->stream);
        launch_batch_step(b, stepOfDay, b->pendPar, batch_hazard(b), x.move, x.inf);
        (void)hipEventRecord(x.e1, b->stream);
        b->samples.push_back(x);
    }
    b->pendPar ^= 1u;
    for (pansim_ctx* c : b->reps) c->pendPar = b->pendPar;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return batch_fail(b, PANSIM_E_DEVICE, std::string("pansim_batch_step: ") + hipGetErrorString(e));
    return 0;
}

int pansim_batch_set_tiles(pansim_batch* b, uint32_t move_items, uint32_t infect_items) {
    if (!b) return PANSIM_E_ARG;
    if (hipStreamSynchronize(b->stream) != hipSuccess) return batch_fail(b, PANSIM_E_DEVICE, "set_tiles: sync failed");
    if (move_items == 0 && infect_items == 0) {
        b->moveCfg = b->infCfg = -1;  // tune again
    } else {
        int mc = -1, ic = -1;
        for (int k = 0; k < 3; ++k)
            if (kMoveItems[k] == move_items) mc = k;
        for (int k = 0; k < 2; ++k)
            if (kInfectItems[k] == infect_items) ic = k;
        if (mc < 0 || ic < 0) return batch_fail(b, PANSIM_E_ARG, "set_tiles: not a candidate tile size");
        b->moveCfg = mc;
        b->infCfg = ic;
    }
    for (auto& x : b->samples) {
        (void)hipEventDestroy(x.e0);
        (void)hipEventDestroy(x.e1);
    }
    b->samples.clear();
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    return 0;
}

int pansim_batch_tiles(pansim_batch* b, uint32_t* move_items, uint32_t* infect_items) {
    if (!b || !move_items || !infect_items) return PANSIM_E_ARG;
    *move_items = b->moveCfg >= 0 ? kMoveItems[b->moveCfg] : 0u;
    *infect_items = b->infCfg >= 0 ? kInfectItems[b->infCfg] : 0u;
    return 0;
}

int pansim_batch_run_day(pansim_batch* b, uint32_t timestamp) {
    if (!b) return PANSIM_E_ARG;
    if (!b->haveDay) return batch_fail(b, PANSIM_E_ARG, "pansim_batch_run_day: pansim_batch_set_day was not called");
    const uint32_t spd = 1440u / b->dt;
    if (timestamp % spd != 0 || timestamp != *b->hDayStart)
        return batch_fail(b, PANSIM_E_ARG, "pansim_batch_run_day: not the first step of the day set by pansim_batch_set_day");
    for (pansim_ctx* c : b->reps)
        if (timestamp != c->wakeDay)
            return batch_fail(b, PANSIM_E_ARG, "pansim_batch_run_day: pansim_plan_locations must run at the start of every day");
    const int hazard = batch_hazard(b) ? 1 : 0;
    if (int rc = batch_pick_tiles(b)) return rc;
    if (!b->exec || b->graphHazard != hazard || b->graphPar != b->pendPar) {
        // capture the day: spd (movement, infection) launch pairs and nothing
        // else (no memset nodes: the counters the steps clear are cleared by
        // the kernels themselves)
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
        if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return batch_fail(b, PANSIM_E_DEVICE, "pansim_batch_run_day: hipStreamBeginCapture failed");
        for (uint32_t s = 0; s < spd; ++s) launch_batch_step(b, s, b->pendPar ^ (s & 1u), hazard != 0, b->moveCfg, b->infCfg);
        hipError_t e = hipStreamEndCapture(b->stream, &b->graph);
        if (e != hipSuccess) return batch_fail(b, PANSIM_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) return batch_fail(b, PANSIM_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        b->graphHazard = hazard;
        b->graphPar = b->pendPar;
    }
    hipError_t e = hipGraphLaunch(b->exec, b->stream);
    if (e != hipSuccess) return batch_fail(b, PANSIM_E_DEVICE, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
    b->pendPar ^= spd & 1u;
    for (pansim_ctx* c : b->reps) {
        c->stateEpoch += spd;
        c->pendPar = b->pendPar;
    }
    return 0;
}

int pansim_batch_sync(pansim_batch* b) {
    if (!b) return PANSIM_E_ARG;
    if (hipStreamSynchronize(b->stream) != hipSuccess) return batch_fail(b, PANSIM_E_DEVICE, "pansim_batch_sync failed");
    return 0;
}

}  // extern "C"
