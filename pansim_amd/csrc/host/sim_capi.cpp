// C ABI of the host simulation loop (include/pansim_sim.h).
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/pansim_sim.h"
#include "ensemble.h"
#include "simulation.h"

struct pansim_sim {
    std::unique_ptr<ps::ConfiguredSimulation> sim;
    std::string err;
};

struct pansim_ensemble {
    std::unique_ptr<ps::Ensemble> ens;
    std::string err;
};

namespace {
std::string g_sim_create_error;
std::string g_ensemble_create_error;

template <class H, class F>
int guarded(H* s, F&& f) {
    try {
        f();
        return 0;
    } catch (const ps::InitError& e) {
        s->err = e.what();
        return e.code < 0 ? e.code : PANSIM_E_DATA;
    } catch (const std::exception& e) {
        s->err = e.what();
        return PANSIM_E_DATA;
    }
}
}  // namespace

extern "C" {

int pansim_sim_create(pansim_sim** out, int argc, const char* const* argv, int print) {
    if (!out) return PANSIM_E_ARG;
    *out = nullptr;
    ps::Options o;
    try {
        o = ps::parse_options(argc, argv, false);
    } catch (const ps::InitError& e) {
        g_sim_create_error = e.what();
        return e.code;
    }
    auto* s = new pansim_sim();
    s->sim.reset(new ps::ConfiguredSimulation(o, print != 0));
    if (!s->sim->succesfullyInitialized) {
        g_sim_create_error = s->sim->initError;
        int code = s->sim->initErrorCode < 0 ? s->sim->initErrorCode : PANSIM_E_DATA;
        delete s;
        return code;
    }
    *out = s;
    return 0;
}

const char* pansim_sim_create_error(void) { return g_sim_create_error.c_str(); }
const char* pansim_sim_error(const pansim_sim* s) { return s ? s->err.c_str() : g_sim_create_error.c_str(); }
void pansim_sim_destroy(pansim_sim* s) { delete s; }

int pansim_sim_steps(pansim_sim* s, uint32_t n) {
    if (!s) return PANSIM_E_ARG;
    return guarded(s, [&] {
        for (uint32_t i = 0; i < n; ++i) s->sim->step();
    });
}

int pansim_sim_run(pansim_sim* s) {
    if (!s) return PANSIM_E_ARG;
    return guarded(s, [&] { s->sim->runSimulation(); });
}

int pansim_sim_finalize(pansim_sim* s) {
    if (!s) return PANSIM_E_ARG;
    return guarded(s, [&] { s->sim->finalize(); });
}

int pansim_sim_run_for_day(pansim_sim* s, int nflags, const char* const* flags, uint32_t* out, size_t cap, size_t* len) {
    if (!s) return PANSIM_E_ARG;
    return guarded(s, [&] {
        std::vector<std::string> f;
        for (int i = 0; i < nflags; ++i) f.emplace_back(flags[i]);
        auto st = s->sim->runForDay(f);
        if (len) *len = st.size();
        for (size_t i = 0; out && i < cap && i < st.size(); ++i) out[i] = st[i];
    });
}

uint32_t pansim_sim_timestamp(const pansim_sim* s) { return s ? s->sim->getSimTime().getTimestamp() : 0u; }
uint32_t pansim_sim_steps_per_day(const pansim_sim* s) { return s ? s->sim->getSimTime().getStepsPerDay() : 0u; }
uint32_t pansim_sim_total_steps(const pansim_sim* s) {
    return s ? s->sim->lengthOfSimulationWeeks * 7u * s->sim->getSimTime().getStepsPerDay() : 0u;
}
int pansim_sim_num_days(const pansim_sim* s) { return s ? (int)s->sim->dayStats.size() : -1; }

int pansim_sim_day_stats(const pansim_sim* s, int d, uint32_t* out, size_t cap, size_t* len) {
    if (!s || d < 0 || d >= (int)s->sim->dayStats.size()) return PANSIM_E_ARG;
    const auto& v = s->sim->dayStats[d];
    if (len) *len = v.size();
    for (size_t i = 0; out && i < cap && i < v.size(); ++i) out[i] = v[i];
    return 0;
}

int pansim_sim_header(const pansim_sim* s, char* out, size_t cap) {
    if (!s || !out || cap == 0) return PANSIM_E_ARG;
    std::strncpy(out, s->sim->statesHeader.c_str(), cap - 1);
    out[cap - 1] = 0;
    return (int)s->sim->statesHeader.size();
}

int pansim_sim_sizes(const pansim_sim* s, uint32_t* na, uint32_t* nl, uint32_t* nv) {
    if (!s) return PANSIM_E_ARG;
    if (na) *na = s->sim->sc.agents.n;
    if (nl) *nl = s->sim->sc.locs.n;
    if (nv) *nv = (uint32_t)s->sim->sc.variantMultiplier.size();
    return 0;
}

pansim_ctx* pansim_sim_context(pansim_sim* s) { return s ? s->sim->ctx : nullptr; }

// ---- ensembles
int pansim_ensemble_create(pansim_ensemble** out, uint32_t n, const int* argc, const char* const* const* argv) {
    if (!out || n == 0 || !argc || !argv) return PANSIM_E_ARG;
    *out = nullptr;
    std::vector<ps::Options> opts;
    for (uint32_t r = 0; r < n; ++r) {
        try {
            opts.push_back(ps::parse_options(argc[r], argv[r], false));
        } catch (const ps::InitError& e) {
            g_ensemble_create_error = "replica " + std::to_string(r) + ": " + e.what();
            return e.code;
        }
    }
    auto* e = new pansim_ensemble();
    e->ens.reset(new ps::Ensemble(opts));
    if (!e->ens->ok()) {
        g_ensemble_create_error = e->ens->initError;
        int code = e->ens->initErrorCode < 0 ? e->ens->initErrorCode : PANSIM_E_DATA;
        delete e;
        return code;
    }
    *out = e;
    return 0;
}

const char* pansim_ensemble_create_error(void) { return g_ensemble_create_error.c_str(); }
const char* pansim_ensemble_error(const pansim_ensemble* e) { return e ? e->err.c_str() : g_ensemble_create_error.c_str(); }
void pansim_ensemble_destroy(pansim_ensemble* e) { delete e; }

int pansim_ensemble_steps(pansim_ensemble* e, uint32_t n) {
    if (!e) return PANSIM_E_ARG;
    return guarded(e, [&] {
        for (uint32_t i = 0; i < n; ++i) e->ens->step();
    });
}

pansim_batch* pansim_ensemble_batch(pansim_ensemble* e) { return e ? e->ens->batch : nullptr; }

int pansim_ensemble_days(pansim_ensemble* e, uint32_t n) {
    if (!e) return PANSIM_E_ARG;
    return guarded(e, [&] {
        for (uint32_t i = 0; i < n; ++i) e->ens->day();
    });
}

int pansim_ensemble_run(pansim_ensemble* e) {
    if (!e) return PANSIM_E_ARG;
    return guarded(e, [&] {
        while (e->ens->timestamp() < e->ens->totalSteps()) e->ens->step();
        for (auto& s : e->ens->sims) s->printAgentStatJSON();  // each replica's runSimulation end
    });
}

int pansim_ensemble_sync(pansim_ensemble* e) {
    if (!e) return PANSIM_E_ARG;
    return pansim_batch_sync(e->ens->batch);
}

uint32_t pansim_ensemble_size(const pansim_ensemble* e) { return e ? (uint32_t)e->ens->sims.size() : 0u; }
uint32_t pansim_ensemble_timestamp(const pansim_ensemble* e) { return e ? e->ens->timestamp() : 0u; }
uint32_t pansim_ensemble_steps_per_day(const pansim_ensemble* e) {
    return e ? e->ens->sims.front()->getSimTime().getStepsPerDay() : 0u;
}

pansim_ctx* pansim_ensemble_context(pansim_ensemble* e, uint32_t r) {
    return (e && r < e->ens->sims.size()) ? e->ens->sims[r]->ctx : nullptr;
}

int pansim_ensemble_num_days(const pansim_ensemble* e, uint32_t r) {
    return (e && r < e->ens->sims.size()) ? (int)e->ens->sims[r]->dayStats.size() : -1;
}

int pansim_ensemble_day_stats(const pansim_ensemble* e, uint32_t r, int d, uint32_t* out, size_t cap, size_t* len) {
    if (!e || r >= e->ens->sims.size()) return PANSIM_E_ARG;
    const auto& ds = e->ens->sims[r]->dayStats;
    if (d < 0 || d >= (int)ds.size()) return PANSIM_E_ARG;
    if (len) *len = ds[d].size();
    for (size_t i = 0; out && i < cap && i < ds[d].size(); ++i) out[i] = ds[d][i];
    return 0;
}

int pansim_ensemble_sizes(const pansim_ensemble* e, uint32_t r, uint32_t* na, uint32_t* nl, uint32_t* nv) {
    if (!e || r >= e->ens->sims.size()) return PANSIM_E_ARG;
    const auto& s = *e->ens->sims[r];
    if (na) *na = s.sc.agents.n;
    if (nl) *nl = s.sc.locs.n;
    if (nv) *nv = (uint32_t)s.sc.variantMultiplier.size();
    return 0;
}

}  // extern "C"
