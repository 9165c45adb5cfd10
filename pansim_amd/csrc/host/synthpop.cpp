// Synthetic household population for the large plain-GPU workloads (10M and
// 100M agents).  The reference has no such generator: its only generator is
// the uniform dummy population of src/IO/dataProviders/dataProvider.cpp:102-184.
// This one builds households, schools with classrooms, workplaces, shops,
// leisure sites, doctors and hospitals, and gives every agent a location list
// in the reference's per-agent format (one location per required type, sorted
// by type ID).  It writes the location and agent arrays of the Scenario
// directly (no string IDs), so 100M agents build in seconds per 10M.
//
// Shape (all draws from the Philox stream of philox.h, phase PH_SYNTH):
//   households   sizes 1..6 with p = {.30,.30,.17,.14,.06,.03}; the first
//                member is an adult
//   ages         10-year bins of a population pyramid, uniform inside a bin
//   agent types  0-2 infant(1), 3-5 kindergarten(2), 6-13 elementary(3),
//                14-17 high school(4), 18-22 university(5) or worker(6),
//                23-64 worker(6) with 5% afternoon shift(7) and 8% at home(8),
//                65+ at home(8); 2% of workers are health workers (their
//                workplace is a hospital or a health centre)
//   schools      300-800 pupils in classrooms of 25
//   workplaces   lognormal sizes (mean 15, sigma 1.2, capped at 2000)
//   shops / leisure / hospitals  one per 200 / 500 / 20000 agents
//   schedule jitter  each agent follows its type's schedule shifted by a
//                uniform multiple of dt in [-J, +J] (--synthJitter J,
//                default 30 min): the shifted schedules are extra agent
//                types in the reference's schedule format (every event's
//                start and end moved, durations kept), appended after the
//                file's types; the afternoon-shift type (ID 7) is not
//                shifted, because the reference keys its curfew exception
//                on that type's index (movementPolicies.h:273, 829)
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../common/philox.h"
#include "scenario.h"

namespace ps {

void build_synthetic(Scenario& sc) {
    const uint32_t N = (uint32_t)sc.opt.synthAgents;
    const uint64_t seed = sc.opt.synthSeed;
    const auto& lt = sc.lt;
    LocationsSoA& L = sc.locs;
    AgentsSoA& A = sc.agents;
    auto add = [&](uint32_t type, uint8_t ess) -> uint32_t {
        uint32_t id = (uint32_t)L.type.size();
        L.type.push_back(type);
        L.infectiousness.push_back(1.0f);
        L.essential.push_back(ess);
        L.open.push_back(1);
        L.capacity.push_back(100);
        return id;
    };
    L.type.reserve(N / 2 + 1024);
    std::vector<uint32_t> homeOf(N), schoolOf(N, NO_LOC), classOf(N, NO_LOC), workOf(N, NO_LOC);
    std::vector<uint8_t> ageOf(N), typeOf(N);
    static const double hhP[6] = {.30, .30, .17, .14, .06, .03};
    static const double agePyr[10] = {.10, .10, .11, .13, .14, .14, .13, .10, .04, .01};
    // households and ages
    for (uint32_t a = 0, hh = 0; a < N; ++hh) {
        Stream rs(seed, hh, 0, PH_SYNTH);
        double u = rs.unit(), c = 0;
        int size = 6;
        for (int i = 0; i < 6; ++i) {
            c += hhP[i];
            if (u < c) {
                size = i + 1;
                break;
            }
        }
        uint32_t home = add(lt.home, 0);
        for (int m = 0; m < size && a < N; ++m, ++a) {
            Stream ra(seed, a, 1, PH_SYNTH);
            double v = ra.unit(), cc = 0;
            int bin = 9;
            for (int i = 0; i < 10; ++i) {
                cc += agePyr[i];
                if (v < cc) {
                    bin = i;
                    break;
                }
            }
            if (m == 0 && bin < 2) bin = 2 + (int)ra.uint_below(5);
            ageOf[a] = (uint8_t)(bin * 10 + (int)ra.uint_below(10));
            homeOf[a] = home;
        }
    }
    // agent types
    for (uint32_t i = 0; i < N; ++i) {
        Stream rs(seed, i, 2, PH_SYNTH);
        unsigned age = ageOf[i];
        uint8_t t;
        if (age <= 2) t = 1;
        else if (age <= 5) t = 2;
        else if (age <= 13) t = 3;
        else if (age <= 17) t = 4;
        else if (age <= 22) t = rs.unit() < 0.5 ? 5 : 6;
        else if (age <= 64) {
            double u = rs.unit();
            t = u < 0.05 ? 7 : (u < 0.13 ? 8 : 6);
        } else t = 8;
        typeOf[i] = t;
    }
    // schools and classrooms: pupils in agent order
    {
        uint32_t sIdx = 0, school = NO_LOC, cap = 0, inSchool = 0, cls = NO_LOC, inClass = 0;
        for (uint32_t i = 0; i < N; ++i) {
            uint8_t t = typeOf[i];
            if (t < 2 || t > 4) continue;
            if (school == NO_LOC || inSchool == cap) {
                Stream rs(seed, sIdx++, 3, PH_SYNTH);
                cap = 300 + rs.uint_below(501);
                school = add(lt.school, 0);
                L.schools.push_back(school);
                L.classroomOffsets.push_back((uint32_t)L.classrooms.size());
                inSchool = 0;
                cls = NO_LOC;
            }
            if (cls == NO_LOC || inClass == 25) {
                cls = add(lt.classroom, 0);
                L.classrooms.push_back(cls);
                inClass = 0;
            }
            schoolOf[i] = school;
            classOf[i] = cls;
            ++inClass;
            ++inSchool;
        }
        L.classroomOffsets.push_back((uint32_t)L.classrooms.size());
    }
    struct Block {
        uint32_t first = 0, n = 0;
    };
    auto block = [&](uint32_t type, uint32_t count, uint8_t ess) {
        Block b;
        b.first = (uint32_t)L.type.size();
        b.n = std::max<uint32_t>(1, count);
        for (uint32_t k = 0; k < b.n; ++k) add(type, ess);
        return b;
    };
    Block shops = block(7, N / 200, 0), leisure = block(8, N / 500, 0), restaurants = block(5, N / 1000, 0),
          cinemas = block(6, N / 20000, 0), churches = block(9, N / 5000, 0), parks = block(10, N / 5000, 0),
          publics = block(lt.publicSpace, N / 2000, 0), doctors = block(lt.doctor, N / 2000, 1),
          hospitals = block(lt.hospital, N / 20000, 1), universities = block(44, N / 50000, 0);
    // workplaces: lognormal sizes, mean 15 (mu = ln 15 - sigma^2/2), cap 2000
    {
        const double sigma = 1.2, mu = std::log(15.0) - 0.5 * sigma * sigma;
        uint32_t wIdx = 0, left = 0, cur = NO_LOC;
        for (uint32_t i = 0; i < N; ++i) {
            uint8_t t = typeOf[i];
            if (t != 6 && t != 7) continue;
            Stream rw(seed, i, 4, PH_SYNTH);
            if (rw.unit() < 0.02) {  // health worker
                bool hosp = rw.unit() < 0.5;
                workOf[i] = hosp ? hospitals.first + rw.uint_below(hospitals.n) : doctors.first + rw.uint_below(doctors.n);
                continue;
            }
            if (left == 0) {
                Stream rs(seed, wIdx++, 5, PH_SYNTH);
                double u1 = 1.0 - rs.unit(), u2 = rs.unit();
                double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
                double sz = std::exp(mu + sigma * z);
                left = (uint32_t)std::min(2000.0, std::max(1.0, std::round(sz)));
                cur = add(lt.work, rs.unit() < 0.1 ? 1 : 0);
            }
            workOf[i] = cur;
            --left;
        }
    }
    // schedule jitter: agent-type variants with every event moved by a
    // multiple of dt in [-J, +J]; variant v of base type b sits at index
    // nBase + (b' * nShift + v), b' counting the shifted base types
    const unsigned dt = sc.opt.deltat ? sc.opt.deltat : 10;
    const int J = (int)(sc.opt.synthJitter / dt) * (int)dt;
    std::vector<int> shifts;
    for (int d = -J; d <= J; d += (int)dt)
        if (d != 0) shifts.push_back(d);
    const unsigned nBase = sc.sched.nTypes;
    const unsigned nightIdx = 6;  // agent type ID 7
    std::vector<int> firstVariant(nBase, -1);
    if (!shifts.empty()) {
        std::vector<uint32_t> baseOff = sc.sched.eventOffset;
        std::vector<Event> baseEv = sc.sched.events;
        unsigned nTypes = nBase;
        for (unsigned b = 0; b < nBase; ++b)
            if (b != nightIdx) {
                firstVariant[b] = (int)nTypes;
                nTypes += (unsigned)shifts.size();
            }
        if (nTypes > 255) throw InitError(ERR_BAD_DATA, "synthJitter: too many agent-type variants");
        auto moved = [](uint8_t h, uint8_t m, int d) {
            if (h == 255) return std::make_pair(h, m);  // undefined time (-1)
            int t = std::min(1440, std::max(0, (int)h * 60 + (int)m + d));
            return std::make_pair((uint8_t)(t / 60), (uint8_t)(t % 60));
        };
        std::vector<uint32_t> off(nTypes * 28 + 1, 0);
        std::vector<Event> ev = baseEv;
        for (unsigned s = 0; s < nBase * 28; ++s) off[s + 1] = baseOff[s + 1];
        for (unsigned b = 0; b < nBase; ++b) {
            if (firstVariant[b] < 0) continue;
            for (size_t v = 0; v < shifts.size(); ++v) {
                unsigned t = (unsigned)firstVariant[b] + (unsigned)v;
                for (unsigned slot = 0; slot < 28; ++slot) {
                    for (uint32_t k = baseOff[b * 28 + slot]; k < baseOff[b * 28 + slot + 1]; ++k) {
                        Event e = baseEv[k];
                        auto st = moved(e.start_h, e.start_m, shifts[v]);
                        auto en = moved(e.end_h, e.end_m, shifts[v]);
                        e.start_h = st.first;
                        e.start_m = st.second;
                        e.end_h = en.first;
                        e.end_m = en.second;
                        ev.push_back(e);
                    }
                    off[t * 28 + slot + 1] = (uint32_t)ev.size();
                }
            }
        }
        sc.sched.nTypes = nTypes;
        sc.sched.eventOffset = off;
        sc.sched.events = ev;
    }

    // cemetery
    L.cemetery = (uint32_t)L.type.size();
    L.type.push_back(lt.cemeteryType);
    L.infectiousness.push_back(0.0f);
    L.essential.push_back(1);
    L.open.push_back(1);
    L.capacity.push_back(0xFFFFFFFFu);
    L.n = (uint32_t)L.type.size();

    // agents
    static const unsigned bandBegin[7] = {0, 5, 15, 30, 60, 70, 80};
    static const double pcP[5] = {0.7924562878640882, 0.07058081624733199, 0.07118987560097449,
                                  0.046110643985943126, 0.019662376301662246};
    const unsigned nCond = 5;
    unsigned susState = 0;
    for (unsigned s = 0; s < sc.st.names.size(); ++s)
        if (sc.st.susceptible[s]) {
            susState = s;
            break;
        }
    A.n = N;
    A.age = ageOf;
    A.sex.resize(N);
    A.preCond.resize(N);
    A.typeIdx.resize(N);
    A.progID.resize(N);
    A.state.resize(N);
    A.diagnosed.assign(N, 0);
    A.scalingAgeSex.assign(N, 1.0f);
    A.location = homeOf;
    A.locOffset.assign(1, 0);
    A.locOffset.reserve((size_t)N + 1);
    A.possibleLocations.reserve((size_t)N * 12);
    A.possibleTypes.reserve((size_t)N * 12);
    std::vector<std::pair<uint8_t, uint32_t>> locs;
    for (uint32_t i = 0; i < N; ++i) {
        Stream rs(seed, i, 6, PH_SYNTH);
        A.sex[i] = rs.unit() < 0.5 ? 1 : 0;
        double u = rs.unit(), c = 0;
        unsigned pc = 0;
        for (unsigned k = 0; k < nCond; ++k) {
            c += pcP[k];
            if (u < c) {
                pc = k;
                break;
            }
        }
        A.preCond[i] = (uint8_t)pc;
        unsigned age = ageOf[i], band = 0;
        for (unsigned b = 0; b < 7; ++b)
            if (age >= bandBegin[b]) band = b;
        A.progID[i] = (uint8_t)std::min<unsigned>(band * nCond + pc, (unsigned)sc.prog.size() - 1);
        A.state[i] = (uint8_t)(rs.unit() < sc.opt.synthInfected ? sc.st.firstInfected : susState);
        int tid = typeOf[i];
        A.typeIdx[i] = (uint8_t)(tid - 1);  // agentTypes.json IDs 1..9 in file order
        if (!shifts.empty() && firstVariant[tid - 1] >= 0) {
            Stream rj(seed, i, 7, PH_SYNTH);
            uint32_t pick = rj.uint_below((uint32_t)shifts.size() + 1);  // 0 = unshifted
            if (pick > 0) A.typeIdx[i] = (uint8_t)(firstVariant[tid - 1] + (int)pick - 1);
        }
        locs.clear();
        locs.push_back({(uint8_t)lt.home, homeOf[i]});
        locs.push_back({(uint8_t)lt.hospital, hospitals.first + rs.uint_below(hospitals.n)});
        locs.push_back({(uint8_t)lt.publicSpace, publics.first + rs.uint_below(publics.n)});
        locs.push_back({(uint8_t)lt.doctor, doctors.first + rs.uint_below(doctors.n)});
        locs.push_back({7, shops.first + rs.uint_below(shops.n)});
        locs.push_back({8, leisure.first + rs.uint_below(leisure.n)});
        locs.push_back({10, parks.first + rs.uint_below(parks.n)});
        if (tid >= 4) {
            locs.push_back({5, restaurants.first + rs.uint_below(restaurants.n)});
            locs.push_back({6, cinemas.first + rs.uint_below(cinemas.n)});
            locs.push_back({9, churches.first + rs.uint_below(churches.n)});
        }
        if (schoolOf[i] != NO_LOC) {
            locs.push_back({(uint8_t)lt.school, schoolOf[i]});
            locs.push_back({(uint8_t)lt.classroom, classOf[i]});
        }
        if (tid == 5) locs.push_back({44, universities.first + rs.uint_below(universities.n)});
        if (workOf[i] != NO_LOC) locs.push_back({(uint8_t)lt.work, workOf[i]});
        std::stable_sort(locs.begin(), locs.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
        for (auto& l : locs) {
            A.possibleTypes.push_back(l.first);
            A.possibleLocations.push_back(l.second);
        }
        A.locOffset.push_back((uint32_t)A.possibleLocations.size());
    }
}

}  // namespace ps
