// Device-side structure-of-arrays layout of one population and the launch
// parameter blocks.  Everything is passed to kernels by value (plain
// pointers), so a kernel never chases a host-side object.
//
// HBM layout (bytes per agent / per location), all arrays 256-byte aligned:
//   agents: state u8, variant u8, progression u8, type u8, age u8, sex u8,
//           precond u8, flags u8 (diagnosed, quarantined, stayedHome,
//           noWorkAgent, healthcare worker), days i16, location u32,
//           wake u32 (absolute step of the next full movement evaluation),
//           lastTest u32, scalingAgeSex f32, susceptibility f32 x V,
//           symptom scaling slot f32 x V, 12 AgentStats u32 + 4 u8 counters
//   locations: type u8, open u8, essential u8, test flag u8, closedUntil u32,
//           quarantineUntil u32, noWork u32, occupancy u32, infectiousness f32,
//           infectiousness sum i64 x V (2^-32 fixed point), pending
//           quarantine key u64
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/pansim_hip.h"

namespace ps {
namespace dev {

constexpr uint32_t NEVER = 0xFFFFFFFFu;
constexpr uint32_t NOLOC = 0xFFFFFFFFu;
constexpr int MAXV = 7;

enum AgentFlag : uint8_t {
    F_DIAGNOSED = 1,
    F_QUARANTINED = 2,
    F_STAYED_HOME = 4,
    F_NO_WORK = 8,
    F_HCW = 16
};

struct Agents {
    uint32_t n;
    uint8_t* state;
    uint8_t* variant;
    uint8_t* prog;
    uint8_t* type;
    uint8_t* age;
    uint8_t* sex;
    uint8_t* precond;
    uint8_t* flags;
    int16_t* days;
    uint32_t* location;
    uint32_t* wake;
    uint32_t* lastTest;
    float* scal;   // scalingAgeSex
    float* susc;   // [V][n]
    float* slot6;  // [V][n] scalingSymptoms[v*7+6]
    // AgentStats
    uint32_t* infTs;
    uint32_t* infLoc;
    uint32_t* worstTs;
    uint32_t* worstEndTs;
    uint32_t* diagTs;
    uint32_t* quarTs;
    uint32_t* quarUntil;
    uint32_t* daysInQ;
    uint32_t* hospTs;
    uint32_t* hospUntil;
    uint32_t* immTs;
    uint32_t* recTs;
    uint8_t* immCount;
    uint8_t* statVar;
    uint8_t* infCount;
    int8_t* worstState;
    // location lists
    const uint32_t* locOffset;
    const uint32_t* possLoc;
    const uint8_t* possType;
};

struct Locs {
    uint32_t n;
    uint32_t cemetery;
    uint8_t* type;
    uint8_t* open;
    uint8_t* essential;
    uint8_t* flag;
    uint32_t* closedUntil;
    uint32_t* quarUntil;
    uint32_t* noWork;
    uint32_t* count;
    float* inf;
    long long* sum;                 // [V][n]
    unsigned long long* pendKey;    // [n]
    const uint32_t* schools;
    const uint32_t* classrooms;
    const uint32_t* classOff;
    uint32_t nSchools;
};

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
