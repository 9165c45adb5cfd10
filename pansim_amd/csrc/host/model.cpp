#include "model.h"

#include <cfloat>
#include <cmath>

namespace ps {

namespace {

// Loop count of the truncated expectation: (unsigned)round(max) as the
// reference's x86-64 build evaluates it (64-bit truncating conversion, low
// 32 bits kept; values outside the int64 range give 0).
unsigned truncation_count(float max) {
    double r = std::round((double)max);
    if (!(r < 9223372036854775808.0) || !(r >= -9223372036854775808.0)) return 0u;
    return (unsigned)(uint64_t)(int64_t)r;
}

// Mean of a geometric length (support 1..n) truncated at n draws.
double truncated_mean(double p, float max) {
    if (p < 0.0) return 0.0;
    double num = 0.0, den = 0.0;
    const unsigned n = truncation_count(max);
    for (unsigned i = 0; i < n; ++i) {
        double w = std::pow(1 - p, (double)i) * p;
        num += (i + 1) * w;
        den += w;
    }
    return num * (1.0 / den);
}

}  // namespace

double length_parameter(float avgLength, float maxLength) {
    if (avgLength == -1.0f) return 1.0;
    const double target = (double)avgLength;
    const double p0 = 1.0 / target;
    auto g = [maxLength, target](double p) { return truncated_mean(p, maxLength) - target; };
    // secant iteration from (p0, 0.9 p0) to |g| <= 1e-4 (smallTools.h:13-29)
    double cur = p0 - (p0 * 0.1), prev = p0;
    bool converged = true;
    unsigned iterations = 0;
    while (0.0001 < std::abs(g(cur))) {
        double keep = cur;
        double slope = g(cur) - g(prev);
        if (slope == 0.0) {
            converged = false;
            break;
        }
        cur -= g(cur) * (cur - prev) / slope;
        prev = keep;
        if (++iterations > 10000) {
            converged = false;
            break;
        }
    }
    double p = converged ? cur : p0;
    if ((p * 10.0 < (1.0 / target)) || p > 1.0) p = 1.0 / target;
    return p;
}

std::vector<pansim_event> to_c_events(const std::vector<Event>& ev) {
    std::vector<pansim_event> out(ev.size());
    for (size_t i = 0; i < ev.size(); ++i) {
        out[i].loc_type = ev[i].locType;
        out[i].chance = ev[i].chance;
        out[i].start_h = ev[i].start_h;
        out[i].start_m = ev[i].start_m;
        out[i].end_h = ev[i].end_h;
        out[i].end_m = ev[i].end_m;
        out[i].dur_h = ev[i].dur_h;
        out[i].dur_m = ev[i].dur_m;
        out[i].pad0 = out[i].pad1 = 0;
    }
    return out;
}

void build_model(const Scenario& sc, ModelArrays& m) {
    const unsigned NS = (unsigned)sc.st.names.size();
    const unsigned NP = (unsigned)sc.prog.size();
    // progression matrices, flattened
    m.edgeOff.assign((size_t)NP * NS + 1, 0);
    m.badCount.assign((size_t)NP * NS, 0);
    m.lenAvg.assign((size_t)NP * NS * 7, -1.0f);
    m.lenMax.assign((size_t)NP * NS * 7, 0.0f);
    m.lenP.assign((size_t)NP * NS * 7, 1.0);
    m.edgeTo.clear();
    m.edgeChance.clear();
    for (unsigned p = 0; p < NP; ++p) {
        const auto& rows = sc.prog[p].rows;
        for (unsigned s = 0; s < NS; ++s) {
            size_t k = (size_t)p * NS + s;
            m.edgeOff[k] = (uint32_t)m.edgeTo.size();
            if (rows.empty()) {  // matrix index with no file: self loop, absorbing
                m.edgeTo.push_back(s);
                m.edgeChance.push_back(1.0f);
                continue;
            }
            const auto& row = rows[s];
            m.badCount[k] = (uint32_t)row.bad.size();
            for (auto& e : row.bad) {
                m.edgeTo.push_back(e.state);
                m.edgeChance.push_back(e.chance);
            }
            for (auto& e : row.neutral) {
                m.edgeTo.push_back(e.state);
                m.edgeChance.push_back(e.chance);
            }
            for (int v = 0; v < 7; ++v) {
                m.lenAvg[k * 7 + v] = row.avgLength[v];
                m.lenMax[k * 7 + v] = row.maxLength[v];
                m.lenP[k * 7 + v] = length_parameter(row.avgLength[v], row.maxLength[v]);
            }
        }
    }
    m.edgeOff[(size_t)NP * NS] = (uint32_t)m.edgeTo.size();
    m.eventOffset = sc.sched.eventOffset;
    m.events = to_c_events(sc.sched.events);

    pansim_dims& d = m.dims;
    d.n_agents = sc.agents.n;
    d.n_locations = sc.locs.n;
    d.n_csr = (uint32_t)sc.agents.possibleLocations.size();
    d.n_states = NS;
    d.n_variants = (uint32_t)sc.variantMultiplier.size();
    d.n_progressions = NP;
    d.n_agent_types = sc.sched.nTypes;
    d.n_events = (uint32_t)m.events.size();
    d.n_schools = (uint32_t)sc.locs.schools.size();
    d.n_classrooms = (uint32_t)sc.locs.classrooms.size();
    d.n_matrix_edges = (uint32_t)m.edgeTo.size();

    pansim_tables& t = m.tables;
    t.state_infectious = sc.st.infectious.data();
    t.state_wb = sc.st.wb.data();
    t.state_susceptible = sc.st.susceptible.data();
    t.state_infected = sc.st.infected.data();
    t.state_accuracy_pcr = sc.st.accuracyPCR.data();
    t.state_accuracy_antigen = sc.st.accuracyAntigen.data();
    t.first_infected = sc.st.firstInfected;
    t.non_covid_dead = sc.st.nonCOVIDDead;
    t.dead_state = sc.st.deadState;
    t.variant_multiplier = sc.variantMultiplier.data();
    t.progression_scaling = sc.progressionScaling.data();
    t.death_scaling = sc.deathScaling.data();
    t.edge_off = m.edgeOff.data();
    t.bad_count = m.badCount.data();
    t.edge_to = m.edgeTo.data();
    t.edge_chance = m.edgeChance.data();
    t.length_avg = m.lenAvg.data();
    t.length_max = m.lenMax.data();
    t.length_p = m.lenP.data();
    t.event_offset = m.eventOffset.data();
    t.events = m.events.data();
    t.lt.public_space = sc.lt.publicSpace;
    t.lt.home = sc.lt.home;
    t.lt.hospital = sc.lt.hospital;
    t.lt.doctor = sc.lt.doctor;
    t.lt.school = sc.lt.school;
    t.lt.classroom = sc.lt.classroom;
    t.lt.work = sc.lt.work;
    t.lt.nursery_home = sc.lt.nurseryhome;
    t.lt.cemetery_type = sc.lt.cemeteryType;

    pansim_locations& l = m.locations;
    l.type = sc.locs.type.data();
    l.infectiousness = sc.locs.infectiousness.data();
    l.essential = sc.locs.essential.data();
    l.open = sc.locs.open.data();
    l.schools = sc.locs.schools.data();
    l.classrooms = sc.locs.classrooms.data();
    l.classroom_offsets = sc.locs.classroomOffsets.data();
    l.cemetery = sc.locs.cemetery;

    pansim_agents& a = m.agents;
    a.age = sc.agents.age.data();
    a.sex = sc.agents.sex.data();
    a.precond = sc.agents.preCond.data();
    a.agent_type = sc.agents.typeIdx.data();
    a.progression = sc.agents.progID.data();
    a.state = sc.agents.state.data();
    a.diagnosed = sc.agents.diagnosed.data();
    a.location = sc.agents.location.data();
    a.scaling_age_sex = sc.agents.scalingAgeSex.data();
    a.healthcare_worker = sc.agents.healthcareWorker.data();
    a.loc_offset = sc.agents.locOffset.data();
    a.possible_locations = sc.agents.possibleLocations.data();
    a.possible_types = sc.agents.possibleTypes.data();
}

}  // namespace ps
