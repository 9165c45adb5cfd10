/*
 * pansim_sim.h -- C ABI of the whole host simulation loop (the reference's
 * Simulation<...> composition with the HIP-backed policies), for callers
 * that drive PanSim as a library: the CLI, the Python package, and the
 * calibration harness of the reference's MATLAB/pybind11 binding.
 *
 * pansim_sim_create + pansim_sim_run_for_day are the counterparts of
 * SimulatorInterface::initSimulation / runForDay (matlab/mexPanSim.cpp:75-126,
 * pybind11 module pyPanSim, :207-219).  pansim_sim_run is
 * Simulation::runSimulation (include/simulation.h:899-926).
 */
#ifndef PANSIM_SIM_H
#define PANSIM_SIM_H

#include <stddef.h>
#include <stdint.h>

#include "pansim_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pansim_sim pansim_sim;

/* argv holds the reference's command-line flags without the program name
 * ("-n", "100000", "--quarantinePolicy", "3", ...) plus --seed / --dataDir.
 * print != 0 writes the statistics header and one line per day to stdout as
 * the reference does.  On failure *out is NULL and pansim_sim_create_error()
 * holds the message. */
int pansim_sim_create(pansim_sim** out, int argc, const char* const* argv, int print);
const char* pansim_sim_create_error(void);
const char* pansim_sim_error(const pansim_sim* sim);
void pansim_sim_destroy(pansim_sim* sim);

/* n iterations of runSimulation's loop (midnight block when due, movement,
 * closure step, infection, ++time). */
int pansim_sim_steps(pansim_sim* sim, uint32_t n);
/* Simulation::runSimulation: all weeks*7*stepsPerDay steps, then the output
 * files (--outAgentStat, --dumpLocationInfectiousList) and the --diags
 * summary. */
int pansim_sim_run(pansim_sim* sim);
/* Simulation::finalize (simulation.h:958-961), which the reference's
 * SimulatorInterface calls at its destruction (matlab/mexPanSim.cpp:70-73):
 * the same output files, for a simulation driven by steps or runForDay. */
int pansim_sim_finalize(pansim_sim* sim);
/* Simulation::runForDay with processFlags strings (TPdef, PL0, CF2000-0500,
 * SO12, QU2, MA0.8, ...).  Writes the day's statistics vector. */
int pansim_sim_run_for_day(pansim_sim* sim, int nflags, const char* const* flags, uint32_t* out, size_t cap, size_t* len);

uint32_t pansim_sim_timestamp(const pansim_sim* sim);
uint32_t pansim_sim_steps_per_day(const pansim_sim* sim);
uint32_t pansim_sim_total_steps(const pansim_sim* sim);
int pansim_sim_num_days(const pansim_sim* sim);
/* Statistics vector of the d-th midnight (runSimulation) or runForDay call. */
int pansim_sim_day_stats(const pansim_sim* sim, int day, uint32_t* out, size_t cap, size_t* len);
/* Tab-separated statistics header (simulation.h:754). */
int pansim_sim_header(const pansim_sim* sim, char* out, size_t cap);
int pansim_sim_sizes(const pansim_sim* sim, uint32_t* n_agents, uint32_t* n_locations, uint32_t* n_variants);
pansim_ctx* pansim_sim_context(pansim_sim* sim);

/* Replica ensembles: n simulations, each from its own command line (own
 * population, seed, k, policies; all with the same time step), advanced in
 * lockstep with one movement and one infection launch per step for all of
 * them.  Each replica's results are those of running it alone with
 * pansim_sim_*.  No counterpart in the reference (one population per
 * process there). */
typedef struct pansim_ensemble pansim_ensemble;
int pansim_ensemble_create(pansim_ensemble** out, uint32_t n, const int* argc, const char* const* const* argv);
const char* pansim_ensemble_create_error(void);
const char* pansim_ensemble_error(const pansim_ensemble* ens);
void pansim_ensemble_destroy(pansim_ensemble* ens);
int pansim_ensemble_steps(pansim_ensemble* ens, uint32_t n);
/* n whole days from a midnight, each one replay of the captured day graph
 * (pansim_batch_run_day); the results equal those of pansim_ensemble_steps. */
int pansim_ensemble_days(pansim_ensemble* ens, uint32_t n);
/* The replica batch of the engine ABI under the ensemble (for its tile-size
 * controls, pansim_batch_set_tiles / pansim_batch_tiles). */
struct pansim_batch* pansim_ensemble_batch(pansim_ensemble* ens);
int pansim_ensemble_run(pansim_ensemble* ens);
int pansim_ensemble_sync(pansim_ensemble* ens);
uint32_t pansim_ensemble_size(const pansim_ensemble* ens);
uint32_t pansim_ensemble_timestamp(const pansim_ensemble* ens);
uint32_t pansim_ensemble_steps_per_day(const pansim_ensemble* ens);
/* replica r's engine context (field downloads through pansim_hip.h) */
pansim_ctx* pansim_ensemble_context(pansim_ensemble* ens, uint32_t r);
int pansim_ensemble_num_days(const pansim_ensemble* ens, uint32_t r);
int pansim_ensemble_day_stats(const pansim_ensemble* ens, uint32_t r, int day, uint32_t* out, size_t cap, size_t* len);
int pansim_ensemble_sizes(const pansim_ensemble* ens, uint32_t r, uint32_t* n_agents, uint32_t* n_locations,
                          uint32_t* n_variants);

#ifdef __cplusplus
}
#endif
#endif
