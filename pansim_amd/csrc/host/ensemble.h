// Replica ensembles: many simulations of the reference's Simulation<...>
// (each with its own population, seed, k and policies) advanced in lockstep,
// with one movement, one quarantine-write and one infection launch per step
// for all of them (pansim_batch of include/pansim_hip.h).  The midnight block
// of every replica runs on that replica as in Simulation::step.  Every
// replica's results are those of running it alone.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "../../../include/pansim_hip.h"
#include "simulation.h"

namespace ps {

class Ensemble {
public:
    std::vector<std::unique_ptr<ConfiguredSimulation>> sims;
    pansim_batch* batch = nullptr;
    std::string initError;
    int initErrorCode = 0;

    explicit Ensemble(const std::vector<Options>& opts, bool printOutput = false) {
        try {
            if (opts.empty()) throw InitError(ERR_BAD_ARG, "an ensemble needs at least one replica");
            for (const auto& o : opts) {
                sims.emplace_back(new ConfiguredSimulation(o, printOutput));
                auto& s = *sims.back();
                if (!s.succesfullyInitialized) throw InitError(s.initErrorCode ? s.initErrorCode : ERR_BAD_DATA, s.initError);
                if (!s.realMovement())
                    throw InitError(ERR_BAD_ARG, "ensemble replicas use the real movement policy");
                if (s.dumpsInfectiousList())
                    throw InitError(ERR_BAD_ARG, "--dumpLocationInfectiousList is not available for ensemble replicas");
                if (s.timeStep != sims.front()->timeStep)
                    throw InitError(ERR_BAD_ARG, "ensemble replicas must share the time step");
            }
            std::vector<pansim_ctx*> ctxs;
            for (auto& s : sims) ctxs.push_back(s->ctx);
            int rc = pansim_batch_create(&batch, ctxs.data(), (uint32_t)ctxs.size());
            if (rc != 0) throw InitError(rc, std::string("pansim_batch_create: ") + pansim_batch_error(nullptr));
        } catch (const InitError& e) {
            initError = e.what();
            initErrorCode = e.code;
        }
    }
    ~Ensemble() {
        if (batch) pansim_batch_destroy(batch);  // before the members it borrows
        sims.clear();
    }
    Ensemble(const Ensemble&) = delete;
    Ensemble& operator=(const Ensemble&) = delete;

    bool ok() const { return batch != nullptr; }
    unsigned timestamp() const { return sims.front()->getSimTime().getTimestamp(); }

    // One step of every replica: midnight blocks where due, then the batched
    // movement + infection.
    void step() {
        const bool midnight = sims.front()->getSimTime().isMidnight();
        if (midnight || !dayReady_) setDay();
        int rc = pansim_batch_step(batch, timestamp());
        if (rc != 0) throw InitError(rc, std::string("batch step: ") + pansim_batch_error(batch));
        for (auto& s : sims) s->finishBatchedStep();
        if (sims.front()->getSimTime().isMidnight()) dayReady_ = false;
    }

    // One whole simulated day from its midnight: the midnight blocks, then
    // the day's steps as one replay of the captured day graph.
    void day() {
        if (!sims.front()->getSimTime().isMidnight())
            throw InitError(ERR_BAD_ARG, "a whole-day batch starts at a midnight");
        setDay();
        int rc = pansim_batch_run_day(batch, timestamp());
        if (rc != 0) throw InitError(rc, std::string("batch day: ") + pansim_batch_error(batch));
        const unsigned spd = sims.front()->getSimTime().getStepsPerDay();
        for (auto& s : sims)
            for (unsigned k = 0; k < spd; ++k) s->finishBatchedStep();
        dayReady_ = false;
    }

    unsigned totalSteps() const {
        return sims.front()->lengthOfSimulationWeeks * 7u * sims.front()->getSimTime().getStepsPerDay();
    }

private:
    bool dayReady_ = false;

    // midnight blocks where due, then this day's batch arguments
    void setDay() {
        {
            for (auto& s : sims) s->midnightIfDue();
            std::vector<pansim_step_params> p;
            std::vector<uint32_t> mask;
            std::vector<float> k, season;
            for (auto& s : sims) {
                p.push_back(s->batchStepParams());
                uint32_t m;
                float kk, se;
                s->batchInfectionParams(m, kk, se);
                mask.push_back(m);
                k.push_back(kk);
                season.push_back(se);
            }
            int rc = pansim_batch_set_day(batch, p.data(), mask.data(), k.data(), season.data());
            if (rc != 0) throw InitError(rc, std::string("batch day parameters: ") + pansim_batch_error(batch));
            dayReady_ = true;
        }
    }
};

}  // namespace ps
