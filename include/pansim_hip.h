/*
 * pansim_hip.h -- C ABI of the MI355X (gfx950) per-timestep agent pipeline.
 *
 * This is the drop-in boundary between the host simulation loop (the
 * policy-templated Simulation<...> of the reference, mirrored in
 * pansim_amd/csrc/host/simulation.h) and the HIP kernels.  Plain C: opaque
 * context handle, POD parameter structs, pointers and sizes; no C++ or torch
 * types.  Every function returns 0 on success or a negative status (see
 * PANSIM_E_*); pansim_last_error() gives the message.  All work is enqueued on
 * the context's stream; functions that return data to the host synchronise.
 *
 * Each entry point names the reference routine it replaces (file:line in
 * khbence/pansim).  In the reference those routines are member functions of
 * the policy classes that launch Thrust algorithms or __global__ kernels
 * directly; there is no FFI in the reference, so this header is the binding a
 * maintainer would add for the HIP-backed policies.
 */
#ifndef PANSIM_HIP_H
#define PANSIM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PANSIM_OK = 0,
    PANSIM_E_ARG = -1,
    PANSIM_E_INPUT_FILE = -2,
    PANSIM_E_JSON = -3,
    PANSIM_E_TIMESTEP = -4,
    PANSIM_E_CLOSURE = -5,
    PANSIM_E_TESTING = -6,
    PANSIM_E_DATA = -7,
    PANSIM_E_DEVICE = -8,
    PANSIM_E_UNSUPPORTED = -9
};

typedef struct pansim_ctx pansim_ctx;

/* Sizes of one population (one replica). */
typedef struct {
    uint32_t n_agents;
    uint32_t n_locations;  /* including the cemetery, which is the last location */
    uint32_t n_csr;        /* total entries of the per-agent location lists */
    uint32_t n_states;     /* progression states (<= 127) */
    uint32_t n_variants;   /* 1..7 */
    uint32_t n_progressions;
    uint32_t n_agent_types;
    uint32_t n_events;
    uint32_t n_schools;
    uint32_t n_classrooms;
    uint32_t n_matrix_edges;
} pansim_dims;

/* Location type IDs of locationTypes.json (include/IO/inputFormats/locationTypesFormat.h). */
typedef struct {
    uint32_t public_space, home, hospital, doctor, school, classroom, work, nursery_home, cemetery_type;
} pansim_location_types;

/* Schedule event: locType, chance, TimeDay start/end/duration as (hour, minute) bytes
 * (src/agentData/agentType.cpp:4-26, src/tools/timeDay.cpp:58-62). */
typedef struct {
    uint32_t loc_type;
    float chance;
    uint8_t start_h, start_m, end_h, end_m, dur_h, dur_m, pad0, pad1;
} pansim_event;

/* Static model tables (src/agentData/dynamicPPState.cpp:66-205,
 * src/tools/progressionMatrices/multiBadMatrix.cpp:43-83,
 * basicLengthAbstract.cpp:4-44).  Matrix edges of (progression p, state s)
 * are edge_to/edge_chance[edge_off[p*n_states+s] ...); the first bad_count
 * of them are the "bad" transitions, the rest the neutral ones. */
typedef struct {
    const float* state_infectious;    /* [n_states] */
    const uint8_t* state_wb;          /* [n_states] 0=W 1=N 2=M 3=S 4=D */
    const uint8_t* state_susceptible; /* [n_states] */
    const uint8_t* state_infected;    /* [n_states] */
    const float* state_accuracy_pcr;  /* [n_states] */
    const float* state_accuracy_antigen;
    uint32_t first_infected, non_covid_dead, dead_state;
    const float* variant_multiplier;  /* [n_variants] */
    const float* progression_scaling; /* [n_variants] */
    const float* death_scaling;       /* [n_variants] */
    const uint32_t* edge_off;         /* [n_progressions*n_states + 1] */
    const uint32_t* bad_count;        /* [n_progressions*n_states] */
    const uint32_t* edge_to;          /* [n_matrix_edges] */
    const float* edge_chance;         /* [n_matrix_edges] */
    const float* length_avg;          /* [n_progressions*n_states*7] */
    const float* length_max;          /* [n_progressions*n_states*7] */
    const double* length_p;           /* [n_progressions*n_states*7] geometric parameter */
    const uint32_t* event_offset;     /* [n_agent_types*28 + 1] */
    const pansim_event* events;       /* [n_events] */
    pansim_location_types lt;
} pansim_tables;

/* Locations, structure of arrays (include/locationList.h:46-67, 90-222). */
typedef struct {
    const uint32_t* type;
    const float* infectiousness;
    const uint8_t* essential;
    const uint8_t* open;
    const uint32_t* schools;           /* [n_schools] */
    const uint32_t* classrooms;        /* [n_classrooms], grouped by school */
    const uint32_t* classroom_offsets; /* [n_schools + 1] (reference layout) */
    uint32_t cemetery;
} pansim_locations;

/* Agents, structure of arrays (include/agentsList.h:39-53, 137-271). */
typedef struct {
    const uint8_t* age;
    const uint8_t* sex;            /* 0 F, 1 M */
    const uint8_t* precond;        /* preCondIdx */
    const uint8_t* agent_type;     /* schedule index */
    const uint8_t* progression;    /* progression matrix index */
    const uint8_t* state;          /* initial state */
    const uint8_t* diagnosed;      /* initial diagnosed flag */
    const uint32_t* location;      /* initial location */
    const float* scaling_age_sex;  /* BasicAgentMeta::scalingAgeSex */
    const uint8_t* healthcare_worker;
    const uint32_t* loc_offset;    /* [n_agents + 1] */
    const uint32_t* possible_locations; /* [n_csr] */
    const uint8_t* possible_types;      /* [n_csr] */
} pansim_agents;

/* Creates a context on the current HIP device, allocates all device arrays
 * and uploads the tables, locations and agents (copied; the caller keeps
 * ownership).  On the device the agents are stored grouped by agent type
 * (input order within a type); every random stream and tie-break is keyed by
 * the agent's input index, and every agent index or per-agent array this
 * interface returns is in input order, so the storage order is invisible.  Draws every agent's initial daysBeforeNextState on the device
 * (DynamicPPState(name, progressionID), dynamicPPState.cpp:279-285) and sets
 * the initial AgentStats (agentsList.h:179-194).  quarantine_policy and
 * quarantine_length are needed for those initial stats.
 * Replaces: Simulation::Simulation init block (include/simulation.h:734-761). */
int pansim_create(pansim_ctx** out, const pansim_dims* dims, const pansim_tables* tables,
                  const pansim_locations* locations, const pansim_agents* agents, uint64_t seed,
                  uint32_t replica, uint32_t time_step, uint32_t quarantine_policy, uint32_t quarantine_length);
void pansim_destroy(pansim_ctx* ctx);
const char* pansim_last_error(const pansim_ctx* ctx);
/* Last error of a failed pansim_create (no context exists then). */
const char* pansim_create_error(void);

/* Movement knobs (RealMovement public members, movementPolicies.h:1359-1367,
 * set by Simulation::toggle* simulation.h:963-973 and the closure rules). */
typedef struct {
    uint32_t timestamp;
    uint32_t time_step;           /* minutes */
    uint32_t day_of_week;         /* 0 = Monday; the holiday-mode Sunday substitution is applied by the caller */
    uint32_t quarantine_policy;
    uint32_t quarantine_length;   /* days */
    uint32_t enable_curfew;
    uint32_t curfew_begin;        /* in steps (already divided by time_step) */
    uint32_t curfew_end;          /* in steps */
    uint32_t school_age_restriction;
    uint32_t quarantine_immune_active;
    uint32_t lockdown_nonvacc_active;
} pansim_step_params;

/* One 10-minute movement step: RealMovement::movement + the per-location
 * occupancy index.  Replaces RealMovementOps::doMovement over all agents
 * (movementPolicies.h:258-1041, launch :1717-1723) and
 * Util::updatePerLocationAgentLists (src/tools/util.cpp:55-110): occupancy and
 * per-variant infectiousness sums are kept per location and updated by the
 * agents that move. */
int pansim_movement_step(pansim_ctx* ctx, const pansim_step_params* p);

/* Infection for every variant whose bit is set in variant_mask, in ascending
 * variant order: the per-location infectiousness sum, the ratio
 * 1 - exp(-k * sum/n * infectiousness * season * dt), and the per-agent
 * Bernoulli draw.  Replaces BasicInfection::infectionsAtLocations
 * (infectionPolicies.h:481-558) with Util::reduce_by_location (util.h:47-103)
 * and LocationsList::infectAgents (locationList.h:233-310); the caller loops
 * over variants with non-zero midnight counts (simulation.h:920-921) by
 * setting their bits. */
int pansim_infection_step(pansim_ctx* ctx, uint32_t timestamp, uint32_t variant_mask, float k, float seasonality,
                          uint32_t time_step);

/* Midnight location planning: stepsUntilMove reset, stayedHome reset, school
 * and work quarantine propagation, closedUntil conversion, children kept at
 * home and the adult who stays with them.  Replaces RealMovement::planLocations
 * (movementPolicies.h:1428-1622).  Also rebuilds the per-location occupancy
 * and infectiousness sums from scratch, because the midnight passes change
 * agent states. */
int pansim_plan_locations(pansim_ctx* ctx, const pansim_step_params* p);

/* DummyMovement (movementPolicies.h:22-65): pansim_dummy_plan_locations is its
 * planLocations (every countdown to 0; also rebuilds the per-location sums
 * after the midnight passes); pansim_dummy_movement_step its movement (agents
 * whose countdown is 0 move to a uniformly drawn location and draw a new
 * countdown below 36; countdowns then decrement with the reference's unsigned
 * wrap).  NoMovement needs no engine call beyond pansim_rebuild_locations at
 * midnight. */
int pansim_dummy_plan_locations(pansim_ctx* ctx);
int pansim_dummy_movement_step(pansim_ctx* ctx, uint32_t timestamp);

/* Recomputes the per-location occupancy and infectiousness sums from the
 * agents' current locations and states (after a midnight block that is not
 * followed by pansim_plan_locations, e.g. at the end of runForDay). */
int pansim_rebuild_locations(pansim_ctx* ctx);

/* DetailedTesting::performTests (testingPolicies.h:370-475).  probs[6] =
 * random, home, work, school, hospital, nursery home.  out[3] = tests,
 * positives of this round, positives diagnosed in the last day elsewhere. */
int pansim_testing_day(pansim_ctx* ctx, uint32_t timestamp, uint32_t time_step, const double probs[6],
                       uint32_t testing_delay, uint32_t quarantine_length, uint32_t use_pcr, uint32_t out[3]);

/* Simulation::updateAgents (simulation.h:425-471) -> DynamicPPState::update
 * (dynamicPPState.cpp:246-303) -> MultiBadMatrix (multiBadMatrix.cpp:9-31,120-125). */
int pansim_progression_day(pansim_ctx* ctx, uint32_t timestamp, uint32_t time_step);

/* Immunization::update susceptibility pass for state-0 agents
 * (immunization.h:547-695).  protection[60] = protectionInfection[10],
 * protectionInfectionWaning[10], protectionSymptomatic[10],
 * protectionSymptomaticWaning[10], protectionHospitalization[10],
 * protectionHospitalizationWaning[10]; similarity[7], similar_multiplier[7]. */
int pansim_immunization_day(pansim_ctx* ctx, uint32_t timestamp, uint32_t time_step, const float protection[60],
                            const int32_t similarity[7], const float similar_multiplier[7]);

/* Immunization::initCategories (immunization.h:229-420): every agent's
 * vaccination round from its category (health worker, nursery home, 60+,
 * 18-59 with condition, essential worker, 18-59, 60+ with condition, school
 * worker, 12-17, 5-11) and one willingness draw.  order[10] =
 * immunizationOrder, group_level[10] = vaccinationGroupLevel. */
int pansim_vaccination_init(pansim_ctx* ctx, const int32_t order[10], const float group_level[10]);

/* Agents per 10-year age group (90+ last) for the booster shares
 * (Immunization::initAgeGroups, immunization.h:421-479). */
int pansim_age_group_counts(pansim_ctx* ctx, uint32_t out[10]);

/* One booster round of Immunization::update (immunization.h:696-830):
 * `available` boosters today, split over the age groups by age_group_frac;
 * eligible agents (first dose > 120 days ago, well, not diagnosed in the last
 * 120 days, from day 90) are boosted with the group's probability. */
int pansim_booster_round(pansim_ctx* ctx, uint32_t timestamp, uint32_t time_step, uint32_t round, uint32_t available,
                         const float age_group_frac[10]);

/* First doses of Immunization::update (immunization.h:836-900): category by
 * category from *current_category (kept by the caller across days) until the
 * `available` doses are used; *immunized_today = agents immunized at this
 * timestamp (boosters included). */
int pansim_vaccinate(pansim_ctx* ctx, uint32_t timestamp, uint32_t time_step, uint32_t available,
                     uint32_t* current_category, uint32_t* immunized_today);

/* Simulation::otherDisease (simulation.h:169-423).  probability_mul as
 * computed there from the hospitalisation targets (0 when none are given). */
int pansim_other_disease_day(pansim_ctx* ctx, uint32_t timestamp, uint32_t time_step, float probability_mul);

/* Number of agents in COVID hospital states 6..8 (simulation.h:178-181), or
 * with dueWithCOVID the infected agents at hospitals not discharged today
 * (simulation.h:183-193). */
int pansim_count_hospitalized_covid(pansim_ctx* ctx, uint32_t timestamp, uint32_t due_with_covid, uint32_t* out);

/* Per-day statistics vector of Simulation::refreshAndPrintStatistics
 * (simulation.h:473-665); test_stats[3] from pansim_testing_day,
 * immunized_today from the vaccination rollout.  Writes up to cap values and
 * returns the full length in *len. */
int pansim_stats_day(pansim_ctx* ctx, uint32_t timestamp, uint32_t time_step, const uint32_t test_stats[3],
                     uint32_t immunized_today, uint32_t* out, size_t cap, size_t* len);

/* Simulation::countVariantCases (simulation.h:768-777): out[n_variants]. */
int pansim_count_variants(pansim_ctx* ctx, uint32_t* out);

/* Closure-rule location toggles (closurePolicies.h:452-467, simulation.h:796-827):
 * every non-essential location whose type is in types[0..n) gets open = open. */
int pansim_set_open_by_type(pansim_ctx* ctx, const uint32_t* types, uint32_t n, uint32_t open);

/* Mask rule (closurePolicies.h:214-246, simulation.h:857-893): infectiousness of
 * every non-home location is divided (divide != 0) or multiplied by factor,
 * in double precision, then rounded to float. */
int pansim_scale_infectiousness(pansim_ctx* ctx, double factor, uint32_t divide);

/* ExposeToMutation rule (closurePolicies.h:347-391). */
int pansim_expose_mutation(pansim_ctx* ctx, uint32_t timestamp, uint32_t variant, double fraction);

/* Re-uploads the schedules after ReduceMovement changed them (agentType.cpp:42-109). */
int pansim_upload_schedules(pansim_ctx* ctx, const uint32_t* event_offset, uint32_t n_offsets, const pansim_event* events,
                            uint32_t n_events);

/* Downloads a named per-agent (in input order) or per-location field as
 * uint32 (floats as raw bits): "state", "variant", "days", "location", "stepsUntilMove",
 * "diagnosed", "quarantined", "stayedHome", "noWorkAgent", "lastTest",
 * "susceptible0", AgentStats members, "vaccinationRound", "locQuarantineUntil",
 * "locClosedUntil", "locOpen", "locInfectiousness", "locCount", and the
 * possible-location lists: "csrBegin", "csrCount" (per agent) and
 * "possibleLocations" (all entries, n >= their number).
 * "stepsUntilMove" is reported relative to `timestamp`. */
int pansim_download_field(pansim_ctx* ctx, const char* name, uint32_t timestamp, uint32_t* out, size_t n);

/* Per-location infectiousness sum of one variant, 2^-32 fixed point. */
int pansim_download_location_sums(pansim_ctx* ctx, uint32_t variant, int64_t* out, size_t n);

/* Infectious-list dump (BasicInfection::dumpLocationInfectiousList,
 * infectionPolicies.h:262-441, DumpStore :27-85).  While enabled, every
 * infection launch records the agents it infects.  pansim_infection_dump_collect
 * then lists the agents infected since the previous collection with their
 * locations, and every agent of non-zero infectiousness at those locations
 * with its location and infectiousness (both lists unordered).  It stores the
 * counts in *n_new / *n_inf; with all five arrays null it only counts and
 * keeps the record, otherwise it fills the arrays (capacities cap_new /
 * cap_inf, PANSIM_E_ARG when too small) and clears the record. */
int pansim_infection_dump_enable(pansim_ctx* ctx, uint32_t on);
int pansim_infection_dump_collect(pansim_ctx* ctx, uint32_t* new_agent, uint32_t* new_loc, size_t cap_new, size_t* n_new,
                                  uint32_t* inf_agent, uint32_t* inf_loc, float* inf_value, size_t cap_inf,
                                  size_t* n_inf);

/* Device timing with HIP events on the context stream.
 * PANSIM_TIMING_KERNELS: every pansim_movement_step and pansim_infection_step
 *   is bracketed by its own event pair; pansim_timing_summary returns the
 *   number of bracketed calls of each and their summed device time in
 *   milliseconds, pansim_last_kernel_ms the time of the last call of each.
 * PANSIM_TIMING_PIPELINE: one event pair around each uninterrupted run of
 *   movement and infection launches (from the first movement launch after any
 *   other entry point to the last launch before the next one, e.g. a day's
 *   steps between two midnight blocks), so that the events add almost no
 *   work; pansim_timing_summary returns the number of timed steps (movement
 *   launches) and their summed device time as n_move / move_ms, and zero
 *   infection calls.
 * PANSIM_TIMING_MOVEMENT: as PANSIM_TIMING_KERNELS for the movement launches
 *   only (half the events; the infection launches are not timed).
 * pansim_enable_timing resets the record. */
#define PANSIM_TIMING_OFF 0u
#define PANSIM_TIMING_KERNELS 1u
#define PANSIM_TIMING_PIPELINE 2u
#define PANSIM_TIMING_MOVEMENT 3u
int pansim_enable_timing(pansim_ctx* ctx, uint32_t mode);
int pansim_timing_summary(pansim_ctx* ctx, uint32_t* n_move, double* move_ms, uint32_t* n_infect, double* infect_ms);
int pansim_last_kernel_ms(pansim_ctx* ctx, float* move_ms, float* infect_ms);

/* Totals since creation (or the last reset): agents whose full movement
 * routine ran, and agents whose location changed. */
int pansim_movement_counters(pansim_ctx* ctx, uint64_t* evaluated, uint64_t* relocated, uint32_t reset);

/* Synchronises the context stream. */
int pansim_sync(pansim_ctx* ctx);

/* Project math on the device (for the tests): out[i] = expf(in[i]) and log. */
int pansim_device_expf(const float* in, float* out, size_t n);
int pansim_device_log(const double* in, double* out, size_t n);
int pansim_device_philox(const uint32_t* ctr4, uint32_t k0, uint32_t k1, uint32_t* out4, size_t n);

/* pansim_device_philox without any allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc); it has no counterpart in the reference.  Counters and
 * blocks are staged in a fixed device array of 2 x 16 B x
 * PANSIM_PHILOX_NOALLOC_MAX (2 MiB) that each GPU keeps until the process
 * ends; more counters return PANSIM_E_ARG (message in pansim_last_error(NULL)).
 * Calls on one GPU share that array: calls that overlap on one GPU race. */
#define PANSIM_PHILOX_NOALLOC_MAX 65536u
int pansim_device_philox_noalloc(const uint32_t* ctr4, uint32_t k0, uint32_t k1, uint32_t* out4, size_t n);

/* ---- Hazard gradient -----------------------------------------------------
 * Calibration support, with no counterpart in the reference: per location
 * type multipliers m_t of the location infectiousness (type id < 256; all 1
 * reproduce the reference bit for bit), and the expected number of
 * infections J = sum over infection steps, variants and susceptible agents
 * at a location with infectious visitors of s * (1 - exp(-x)), x = k *
 * density * infectiousness * m_type * seasonality * dt, with its gradient
 * with respect to k and every m_t along the simulated trajectory (which is
 * held fixed), accumulated in FP64 while enabled.  out[0] = J, out[1] =
 * dJ/dk, out[2 + t] = dJ/dm_t; n >= 258. */
int pansim_set_location_type_multipliers(pansim_ctx* ctx, const float* m, uint32_t n);
int pansim_hazard_enable(pansim_ctx* ctx, uint32_t on);
/* Finite-difference probes along the same trajectory (0 <= h < 1; h = 0
 * switches them off): with them, out[258..261] of pansim_hazard_read (n >=
 * 262) hold J(k (1+h)) - J, J(k (1-h)) - J, and the same for the multiplier
 * of location type `type`; (out[258] - out[259]) / (2 h k) is a central
 * difference of dJ/dk. */
int pansim_hazard_set_probe(pansim_ctx* ctx, double h, uint32_t type);
int pansim_hazard_read(pansim_ctx* ctx, double* out, uint32_t n, uint32_t reset);

/* ---- Replica batches --------------------------------------------------
 * Many contexts (replicas of uneven size, own seed, k, policies) on one
 * device stepped together: every step is one movement, one quarantine-write
 * and one infection launch for all replicas.  This has no single counterpart
 * in the reference, which runs one population per process; each replica's
 * results are those of stepping its context alone (pansim_movement_step +
 * pansim_infection_step with the same arguments).  While batched, the
 * members work on the batch's stream, and their midnight routines are called
 * on the members as usual; pansim_batch_set_day takes each replica's
 * movement knobs and infection mask / k / seasonality after its midnight
 * block, and pansim_batch_step advances every replica by one step at the
 * shared timestamp.  Plain steps of a member in between are not allowed. */
typedef struct pansim_batch pansim_batch;
int pansim_batch_create(pansim_batch** out, pansim_ctx* const* replicas, uint32_t n);
void pansim_batch_destroy(pansim_batch* batch);
const char* pansim_batch_error(const pansim_batch* batch);
int pansim_batch_set_day(pansim_batch* batch, const pansim_step_params* params, const uint32_t* variant_masks,
                         const float* k, const float* seasonality);
int pansim_batch_step(pansim_batch* batch, uint32_t timestamp);
/* The whole day that starts at `timestamp` (the first step of the day given
 * to pansim_batch_set_day) as one graph launch: the day's steps are captured
 * once as a hipGraph of their movement and infection launches and replayed
 * every day (captured again when hazard accumulation or the pending-write
 * parity changes).  Equivalent to steps_per_day calls of pansim_batch_step. */
int pansim_batch_run_day(pansim_batch* batch, uint32_t timestamp);
/* Tile sizes of the batch kernels: agents per thread of the movement tile
 * (8, 16 or 32; its LDS list of movers holds 256 times as many entries) and
 * of the infection tile (4 or 8).  By default an autotuner picks them: the
 * eager pansim_batch_step calls before the first pansim_batch_run_day cycle
 * through the six pairs, timed with HIP events, and the fastest pair on
 * average is used from then on (the results do not depend on the sizes).
 * pansim_batch_set_tiles(b, 0, 0) restarts the tuning; other values fix the
 * sizes.  pansim_batch_tiles reports the pair in use (0, 0 while tuning). */
int pansim_batch_set_tiles(pansim_batch* batch, uint32_t move_items, uint32_t infect_items);
int pansim_batch_tiles(pansim_batch* batch, uint32_t* move_items, uint32_t* infect_items);
int pansim_batch_sync(pansim_batch* batch);

#ifdef __cplusplus
}
#endif
#endif
