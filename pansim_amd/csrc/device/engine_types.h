// Device-side layout of one population and the launch parameter blocks.
// Everything is passed to kernels by value (plain pointers), so a kernel
// never chases a host-side object.
//
// HBM layout, all arrays 256-byte aligned:
//   agents are stored grouped by schedule (agent type), so that the agents
//     that move at one step share cache lines; id[] maps back to the input;
//   agents, per-agent record AgentRec (32 B): the fields that the movement
//     of one agent reads, so that a mover touches one cache line instead of
//     one line per field;
//   agents, structure of arrays: location u32 and wake u16 (streamed by
//     every step; the wake step counts from the first step of the current
//     day, since the midnight planning resets every agent's), susceptibility f32 x V (streamed by the infection pass,
//     variant-major), and the fields only the midnight passes or rare paths
//     touch: progression, sex, precond u8, days i16, lastTest u32,
//     scalingAgeSex f32, symptom scaling slot f32 x V, AgentStats
//     timestamps u32 and u8 counters, the possible-location lists (CSR);
//   locations, LocRec (16 B): what movement reads of a location;
//   locations, LocDyn (16 B): occupancy, exact infectiousness sum of
//     variant 0 and infectiousness, the line that a relocation updates and
//     the infection pass gathers; sums of variants 1.. in a separate array;
//     pending quarantine key u64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/pansim_hip.h"

namespace ps {
namespace dev {

constexpr uint32_t NEVER = 0xFFFFFFFFu;
constexpr uint32_t NOLOC = 0xFFFFFFFFu;
constexpr int MAXV = 7;
constexpr uint8_t LOC_TYPE_NONE = 0xFF;
constexpr uint16_t WAKE_NONE = 0xFFFFu;  // no full movement evaluation left today  // type of a missing location; no location type may use it

enum AgentFlag : uint8_t {
    F_DIAGNOSED = 1,
    F_QUARANTINED = 2,
    F_STAYED_HOME = 4,
    F_NO_WORK = 8,
    F_HCW = 16
};

// The movement-hot fields of one agent.
struct __align__(32) AgentRec {
    uint32_t quarUntil;  // AgentStats::quarantinedUntilTimestamp
    uint32_t diagTs;     // AgentStats::diagnosedTimestamp
    uint32_t hospTs;     // AgentStats::hospitalizedTimestamp
    uint32_t hospUntil;  // AgentStats::hospitalizedUntilTimestamp
    uint32_t home;       // first home-type entry of the location list, NOLOC if none
    uint32_t csrBegin;   // start of the possible-location list
    uint8_t type;        // agent type (schedule index)
    uint8_t state;       // progression state
    uint8_t flags;       // AgentFlag bits
    uint8_t variant;
    uint8_t age;
    uint8_t wb;          // well-being class of the state (Tables::stWB[state]), written with every state
    uint16_t csrCount;   // length of the possible-location list
};
static_assert(sizeof(AgentRec) == 32, "AgentRec is one 32-byte record");

// What movement reads of a location.
struct __align__(16) LocRec {
    uint32_t quarUntil;    // quarantineUntil
    uint32_t closedUntil;
    uint32_t noWork;       // midnight child-care claim (0, NOWORK_FLAGGED or agent + 1)
    uint8_t type;
    uint8_t open;
    uint8_t essential;
    uint8_t flag;          // unused (the testing flags are the bits of Locs::flagBits)
};
static_assert(sizeof(LocRec) == 16, "LocRec is one 16-byte record");

// Occupancy, exact variant-0 infectiousness sum (2^-32 fixed point) and
// the location's infectiousness times its type's multiplier (the factor the
// infection pass uses; Locs::infRaw holds the infectiousness itself).
struct __align__(16) LocDyn {
    long long sum0;
    uint32_t count;
    float inf;
};
static_assert(sizeof(LocDyn) == 16, "LocDyn is one 16-byte record");

struct Agents {
    uint32_t n;
    // agent i of the device arrays is agent id[i] of the input: agents are
    // stored grouped by schedule (agent type), input order within a type;
    // random streams and tie-breaks are keyed by the input index, so results
    // do not depend on the storage order
    const uint32_t* id;
    AgentRec* rec;
    uint8_t* prog;
    uint8_t* sex;
    uint8_t* precond;
    int16_t* days;
    uint32_t* location;
    uint16_t* wake;  // step of the day of the next full movement evaluation, WAKE_NONE if none today
    uint32_t* lastTest;
    float* scal;   // scalingAgeSex
    float* susc;   // [V][suscStride]
    uint32_t suscStride;  // n rounded up to whole infection tiles (padding: 0)
    float* slot6;  // [V][n] scalingSymptoms[v*7+6]
    // AgentStats not in AgentRec
    uint32_t* infTs;
    uint32_t* infLoc;
    uint32_t* worstTs;
    uint32_t* worstEndTs;
    uint32_t* quarTs;
    uint32_t* daysInQ;
    uint32_t* immTs;
    uint32_t* recTs;
    uint8_t* immCount;
    uint8_t* immRound;  // vaccination category round, 255 = declines
    uint8_t* statVar;
    uint8_t* infCount;
    int8_t* worstState;
    // possible-location lists
    const uint32_t* possLoc;
    const uint8_t* possType;
    const uint32_t* types16;  // [n][4]: the type bytes of each list of at most 16 entries, zero padded
};

struct Locs {
    uint32_t n;
    uint32_t cemetery;
    LocRec* rec;
    LocDyn* dyn;
    float* infRaw;                  // [n] infectiousness (input, scaled by the closure rules)
    const uint8_t* type8;           // [n] location type (the type of LocRec, 1 B per location: cache resident)
    long long* sumX;                // [V-1][n]: sums of variants 1..V-1
    uint32_t* visited;              // [(n + 31) / 32] bit l: an infectious agent was at l today (superset of sum != 0)
    uint32_t* special;              // [(n + 31) / 32] bit l clear: l is open, not closed and not quarantined all day
    uint32_t* noWorkBits;           // [(n + 31) / 32] bit l: home l was flagged for child care by tonight's planning
    uint32_t* flagBits;             // [(n + 31) / 32] bit l: the testing flag of l (flagLocations)
    unsigned long long* pendKey;    // [n]
    const uint32_t* schools;
    const uint32_t* classrooms;
    const uint32_t* classOff;
    uint32_t nSchools;
};

// Whether movement must read location l's record to know if it is closed
// or quarantined: a clear bit means open, closedUntil and quarUntil at most
// the day's first step.  The midnight planning recomputes the bits; every
// later write that can close or quarantine a location sets its bit.  A
// missing location (NOLOC) always reads its (closed) record.
__device__ __forceinline__ bool loc_special(const Locs& L, uint32_t l) {
    return l >= L.n || ((L.special[l >> 5] >> (l & 31u)) & 1u) != 0u;
}
__device__ __forceinline__ void mark_special(const Locs& L, uint32_t l) { atomicOr(&L.special[l >> 5], 1u << (l & 31u)); }

// Marks location l as visited by an infectious agent (bits are only cleared
// by the full rebuild at midnight, so a set bit may have a zero sum again).
__device__ __forceinline__ void mark_visited(const Locs& L, uint32_t l) { atomicOr(&L.visited[l >> 5], 1u << (l & 31u)); }

// Address of the exact infectiousness sum of variant v at location l.
__device__ __forceinline__ long long* loc_sum(const Locs& L, uint32_t v, uint32_t l) {
    return v == 0 ? &L.dyn[l].sum0 : L.sumX + (size_t)(v - 1) * L.n + l;
}

struct Tables {
    uint32_t nStates, nVar, nProg;
    uint32_t firstInfected, nonCOVIDDead, deadState;
    const float* stInf;
    const uint8_t* stWB;
    const uint8_t* stSusc;
    const uint8_t* stInfected;
    const float* accPCR;
    const float* accAnt;
    const float* varMul;
    const float* progScaling;
    const float* deathScaling;
    const uint32_t* edgeOff;
    const uint32_t* badCount;
    const uint32_t* edgeTo;
    const float* edgeChance;
    const float* lenAvg;
    const float* lenMax;
    const double* lenP;
    const uint32_t* evOff;
    const pansim_event* ev;
    const int4* evWin;  // [steps per day][nOff]: active-window bounds (ab, ae) of each schedule slot, and the
                        // location type and duration (h | m << 8) of event ab
    uint32_t nOff;      // schedule slots: agent types x 28
    pansim_location_types lt;
};

// Per-step scratch shared by the movement and infection kernels.
struct StepScratch {
    uint32_t* pendList;     // locations with a pending quarantine write
    uint32_t* pendCount;    // [2]: entries of pendList, by step parity
    uint32_t pendCap;
};

}  // namespace dev
}  // namespace ps
