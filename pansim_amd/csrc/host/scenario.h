// Host-side input layer: command-line options (same flag names and defaults as
// the reference's cxxopts registrations), the reference JSON input formats,
// and the population (dummy -n/-N generator, JSON agents/locations files, or
// the synthetic household generator) turned into flat structure-of-arrays
// ready for upload.  Everything here is one-time initialisation; nothing in
// this file runs per step.
//
// Reference counterparts:
//   options           include/simulation.h:16-51,135-167 and every policy's
//                     addProgramParameters
//   JSON formats      include/IO/inputFormats/*.h
//   data provider     src/IO/dataProviders/dataProvider.cpp:8-217
//   location init     include/locationList.h:78-222
//   agent init        include/agentsList.h:94-271, src/agentData/agentMeta.cpp:50-84
//   schedules         src/agentData/agentType.cpp:4-26
//   state tables      src/agentData/dynamicPPState.cpp:66-205 (table part)
//   matrices (raw)    src/tools/progressionMatrices/multiBadMatrix.cpp:43-83
#pragma once
#include <stdint.h>

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace ps {

constexpr int MAX_STRAINS = 7;
constexpr uint32_t NO_LOC = 0xFFFFFFFFu;

// Errors raised during initialisation.  They map to the reference's
// CustomErrors hierarchy (include/tools/customExceptions.h) and to the
// negative status codes of the C ABI.
struct InitError : std::runtime_error {
    int code;
    InitError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
enum ErrorCode {
    ERR_BAD_ARG = -1,
    ERR_BAD_INPUT_FILE = -2,
    ERR_BAD_JSON = -3,
    ERR_BAD_TIMESTEP = -4,
    ERR_BAD_CLOSURE = -5,
    ERR_BAD_TESTING = -6,
    ERR_BAD_DATA = -7,
    ERR_DEVICE = -8,
    ERR_UNSUPPORTED = -9
};

struct Options {
    // Simulation (include/simulation.h:16-51, 135-167)
    unsigned weeks = 12;
    unsigned deltat = 10;
    int numagents = -1;
    int numlocs = -1;
    std::string progression = "inputConfigFiles/progressions/transition_config.json";
    std::string agents = "inputRealExample/agents.json";
    std::string agentTypes = "inputConfigFiles/agentTypes.json";
    std::string locations = "inputRealExample/locations.json";
    std::string locationTypes = "inputConfigFiles/locationTypes.json";
    std::string parameters = "inputConfigFiles/parameters.json";
    std::string configRandom = "inputConfigFiles/configRandom.json";
    std::string closures = "inputConfigFiles/closureRules.json";
    bool randomStates = false;
    std::string outAgentStat;
    unsigned diags = 0;
    int otherDisease = 1;
    std::string infectiousnessMultiplier = "1.03,1.79,2.45,2.7,3.6";
    std::string diseaseProgressionScaling = "0.9,1.05,0.9,0.8,0.8";
    std::string diseaseProgressionDeathScaling = "1.0,1.03,1.3,0.6,0.6";
    unsigned startDay = 2;
    unsigned startDate = 267;
    std::string totalHospitalizations = "inputConfigFiles/dailyHospitalizationTargets.txt";
    unsigned dueWithCOVID = 0;
    // BasicInfection (include/policies/infectionPolicies.h:89-129)
    double k = 0.000347;
    unsigned dumpLocationInfections = 0;
    std::string dumpLocationInfectiousList;
    std::string d_offset = "-5";
    int d_peak_offset = 0;
    double trunc = 0.5;
    double c0 = 3.08;
    // RealMovement (include/policies/movementPolicies.h:1370-1394)
    unsigned trace = 0xFFFFFFFFu;
    unsigned quarantinePolicy = 3;
    unsigned quarantineLength = 10;
    std::string dumpLocationAgentList;
    std::string dumpLoctypeStat;
    // DetailedTesting (include/policies/testingPolicies.h:318-356)
    std::string testingProbabilities = "0.00005,0.01,0.0005,0.0005,0.005,0.05";
    unsigned testingRepeatDelay = 5;
    std::string testingMethod = "PCR";
    // RuleClosure (include/policies/closurePolicies.h:73-83)
    unsigned enableClosures = 1;
    // AgentList (include/agentsList.h:121-125)
    unsigned disableTourists = 1;
    // Immunization (include/immunization.h:91-137)
    unsigned immunizationStart = 96;
    std::string boosterStart = "315";
    std::string immunizationsPerWeek = "inputConfigFiles/vaccPerWeek.txt";
    std::string boostersPerWeek = "inputConfigFiles/boosterPerWeek.txt";
    std::string immunizationOrder = "1,2,3,4,5,6,0,0,7,8";
    std::string vaccinationGroupLevel = "0.9,0.85,0.9,0.82,0.8,0.75,0.8,0.67,0.4,0.2";
    std::string protectionInfection = "0.85,0.5,0.77,0.4,0.85,0.63,0.9,0.65,0.92,0.7";
    std::string protectionInfectionWaning = "0.003,0.006,0.01,0.02,0.005,0.02,0.002,0.004,0.002,0.004";
    std::string protectionSymptomatic = "0.9,0.55,0.8,0.6,0.95,0.69,0.92,0.7,0.94,0.77";
    std::string protectionSymptomaticWaning = "0.003,0.006,0.01,0.02,0.005,0.02,0.002,0.004,0.002,0.004";
    std::string protectionHospitalization = "0.92,0.85,0.95,0.9,0.97,0.92,0.974,0.93,0.98,0.97";
    std::string protectionHospitalizationWaning = "0.0021,0.006,0.006,0.006,0.003,0.003,0.0014,0.003,0.0014,0.003";
    std::string variantSimilarity = "0,0,0,1,1,1,1";
    std::string variantSimilarMultiplier = "0.0,0.0,0.0,0.0,0.0,0.0,0.0";
    // New in this implementation
    uint64_t seed = 0;               // --seed: key of the Philox stream
    std::string dataDir;             // --dataDir: fallback directory for relative input paths
    int synthAgents = -1;            // --synthAgents: synthetic household population of this size
    uint64_t synthSeed = 1;          // --synthSeed: generator seed for --synthAgents
    double synthInfected = 0.001;    // --synthInfected: initial fraction in the first infected state
    // --synthJitter: schedule jitter of the synthetic population in minutes:
    // every agent follows its type's schedule shifted by one of
    // -J, -J+dt, ..., +J (uniform), stored as shifted agent-type variants
    unsigned synthJitter = 30;
    // --movementPolicy real|dummy|none: RealMovement, DummyMovement or
    // NoMovement (movementPolicies.h:9-66, 1339-1752), which the reference
    // selects at build time (CMakeLists.txt:76-85)
    std::string movementPolicy = "real";
    // --testingPolicy detailed|none: DetailedTesting or NoTesting
    // (testingPolicies.h:11-20, 292-476), and --closurePolicy rule|none:
    // RuleClosure or NoClosure (closurePolicies.h:36-46, 47-481), chosen at
    // build time in the reference (CMakeLists.txt:76-85, configTypes.h.in:15)
    std::string testingPolicy = "detailed";
    std::string closurePolicy = "rule";
    bool help = false;
    bool version = false;
};

// Parses argv (argv[0] is skipped when skip_program_name).  Unknown flags and
// malformed values raise InitError(ERR_BAD_ARG).
Options parse_options(int argc, const char* const* argv, bool skip_program_name = true);
std::string options_help();

// Reference-compatible helpers (src/tools/smallTools.cpp:44-130).
std::vector<double> split_double(const std::string& s, char sep);
std::vector<float> split_float(const std::string& s, char sep);
std::vector<int> split_int(const std::string& s, char sep);

struct LocationTypeIds {
    unsigned publicSpace = 1, home = 2, hospital = 12, doctor = 14, school = 3, classroom = 33, work = 4, nurseryhome = 22;
    std::vector<std::pair<int, std::string>> types;
    unsigned cemeteryType = 0;  // largest declared type ID + 1 (include/locationList.h:78-82)
};

// Well-being classes of include/agentData/globalStates.h:9
enum WB : uint8_t { WB_W = 0, WB_N = 1, WB_M = 2, WB_S = 3, WB_D = 4 };

struct StateTable {
    std::vector<std::string> names;
    std::vector<uint8_t> wb;
    std::vector<float> infectious;
    std::vector<float> accuracyPCR;
    std::vector<float> accuracyAntigen;
    std::vector<uint8_t> susceptible;
    std::vector<uint8_t> infected;
    unsigned firstInfected = 0;
    unsigned nonCOVIDDead = 0;
    unsigned deadState = 0;  // first state whose WB is D (dynamicPPState.cpp:177-184)
    std::string header;      // state names joined by tabs, each followed by a tab
};

// One progression matrix as read from transition_illness*.json, before any
// derived quantity (geometric parameter p) is computed.
struct ProgressionMatrixRaw {
    struct Next {
        unsigned state;
        float chance;
    };
    struct StateRow {
        float avgLength[MAX_STRAINS];
        float maxLength[MAX_STRAINS];  // -1 already replaced by FLT_MAX
        std::vector<Next> bad;
        std::vector<Next> neutral;  // cleanUp: a state with no outgoing edges gets (self, 1.0)
    };
    std::vector<StateRow> rows;
};

// Schedule event with TimeDay-encoded times (src/tools/timeDay.cpp:58-62).
struct Event {
    uint32_t locType;
    float chance;
    uint8_t start_h, start_m, end_h, end_m, dur_h, dur_m;
};

struct Schedules {
    unsigned nTypes = 0;
    std::vector<uint32_t> eventOffset;  // nTypes*28 + 1, index = 28*type + 7*WB + day
    std::vector<Event> events;
};

struct LocationsSoA {
    uint32_t n = 0;        // including the cemetery (last)
    uint32_t cemetery = 0;
    std::vector<std::string> ids;
    std::vector<uint32_t> type;
    std::vector<float> infectiousness;
    std::vector<uint8_t> essential;
    std::vector<uint8_t> open;
    std::vector<uint32_t> capacity;
    std::vector<uint32_t> schools;
    std::vector<uint32_t> classrooms;
    std::vector<uint32_t> classroomOffsets;
};

struct AgentsSoA {
    uint32_t n = 0;
    std::vector<uint8_t> age;
    std::vector<uint8_t> sex;        // 0 = F, 1 = M
    std::vector<uint8_t> preCond;    // preCondIdx (stoi of the ID)
    std::vector<uint8_t> typeIdx;    // index into Schedules
    std::vector<uint8_t> progID;     // progression matrix index
    std::vector<uint8_t> state;      // initial state index
    std::vector<uint8_t> diagnosed;  // initial diagnosed flag
    std::vector<uint32_t> location;  // initial location (home, or 0)
    std::vector<float> scalingAgeSex;
    std::vector<uint32_t> locOffset;     // N + 1
    std::vector<uint32_t> possibleLocations;
    std::vector<uint8_t> possibleTypes;
    std::vector<uint8_t> healthcareWorker;  // simulation.h:667-692
};

struct ClosureRuleRaw {
    std::string name, conditionType, parameter;
    double threshold = 0, threshold2 = 0;
    int closeAfter = 0, openAfter = 0;
    std::vector<int> locationTypesToClose;
};

// Vaccination and booster rollout inputs (Immunization::initializeArgs,
// include/immunization.h:139-222).
struct VaccinationPlan {
    unsigned startDay = 96;                              // immunizationStart
    std::vector<int> boosterStart;                       // first day of each booster round
    unsigned dailyDoses = 0;                             // immunizationsPerWeek given as a number, / 7
    std::vector<float> vaccPerWeek;                      // or per week, from a file
    unsigned dailyBoosters = 0;                          // boostersPerWeek given as a number, / 7
    std::vector<std::vector<float>> boosterPerWeek;      // per round, from a file
    std::vector<std::vector<float>> boosterAgePercent;   // per round: 10 age-group shares
    std::vector<int> order;                              // immunizationOrder: 10 values
    std::vector<float> groupLevel;                       // vaccinationGroupLevel: 10 values
};

struct Scenario {
    Options opt;
    VaccinationPlan vacc;
    LocationTypeIds lt;
    StateTable st;
    std::vector<ProgressionMatrixRaw> prog;
    Schedules sched;
    LocationsSoA locs;
    AgentsSoA agents;
    std::vector<ClosureRuleRaw> rules;
    std::vector<float> variantMultiplier;
    std::vector<float> progressionScaling;
    std::vector<float> deathScaling;
    std::vector<int> totalHospitalizations;
    std::vector<int> dOffset;
    unsigned stepsPerDay = 144;
    std::string statesHeader;  // full per-day statistics header (simulation.h:754)
};

// Resolves an input path: as given, then relative to opt.dataDir.
std::string resolve_path(const Options& opt, const std::string& path);

// Builds the whole scenario (reads files, generates the population).
Scenario build_scenario(const Options& opt);

}  // namespace ps
