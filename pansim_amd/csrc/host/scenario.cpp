#include "scenario.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <set>
#include <sstream>
#include <sys/stat.h>

#include "../common/philox.h"
#include "json.h"

namespace ps {

// Defined in synthpop.cpp: fills sc.locs and sc.agents directly.
void build_synthetic(Scenario& sc);

// ---------------------------------------------------------------------------
// Option parsing
// ---------------------------------------------------------------------------
namespace {

struct OptSpec {
    const char* shortName;
    const char* longName;
    enum Kind { U, I, D, S, B, U64 } kind;
    void* target;
};

std::vector<OptSpec> option_table(Options& o) {
    return {
        {"w", "weeks", OptSpec::U, &o.weeks},
        {"t", "deltat", OptSpec::U, &o.deltat},
        {"n", "numagents", OptSpec::I, &o.numagents},
        {"N", "numlocs", OptSpec::I, &o.numlocs},
        {"P", "progression", OptSpec::S, &o.progression},
        {"a", "agents", OptSpec::S, &o.agents},
        {"A", "agentTypes", OptSpec::S, &o.agentTypes},
        {"l", "locations", OptSpec::S, &o.locations},
        {"L", "locationTypes", OptSpec::S, &o.locationTypes},
        {"p", "parameters", OptSpec::S, &o.parameters},
        {"c", "configRandom", OptSpec::S, &o.configRandom},
        {"", "closures", OptSpec::S, &o.closures},
        {"r", "randomStates", OptSpec::B, &o.randomStates},
        {"", "outAgentStat", OptSpec::S, &o.outAgentStat},
        {"", "diags", OptSpec::U, &o.diags},
        {"", "otherDisease", OptSpec::I, &o.otherDisease},
        {"", "infectiousnessMultiplier", OptSpec::S, &o.infectiousnessMultiplier},
        {"", "diseaseProgressionScaling", OptSpec::S, &o.diseaseProgressionScaling},
        {"", "diseaseProgressionDeathScaling", OptSpec::S, &o.diseaseProgressionDeathScaling},
        {"", "startDay", OptSpec::U, &o.startDay},
        {"", "startDate", OptSpec::U, &o.startDate},
        {"", "totalHospitalizations", OptSpec::S, &o.totalHospitalizations},
        {"", "dueWithCOVID", OptSpec::U, &o.dueWithCOVID},
        {"k", "infectionCoefficient", OptSpec::D, &o.k},
        {"", "dumpLocationInfections", OptSpec::U, &o.dumpLocationInfections},
        {"", "dumpLocationInfectiousList", OptSpec::S, &o.dumpLocationInfectiousList},
        {"", "d_offset", OptSpec::S, &o.d_offset},
        {"", "d_peak_offset", OptSpec::I, &o.d_peak_offset},
        {"", "trunc", OptSpec::D, &o.trunc},
        {"", "c0", OptSpec::D, &o.c0},
        {"", "trace", OptSpec::U, &o.trace},
        {"", "quarantinePolicy", OptSpec::U, &o.quarantinePolicy},
        {"", "quarantineLength", OptSpec::U, &o.quarantineLength},
        {"", "dumpLocationAgentList", OptSpec::S, &o.dumpLocationAgentList},
        {"", "dumpLoctypeStat", OptSpec::S, &o.dumpLoctypeStat},
        {"", "testingProbabilities", OptSpec::S, &o.testingProbabilities},
        {"", "testingRepeatDelay", OptSpec::U, &o.testingRepeatDelay},
        {"", "testingMethod", OptSpec::S, &o.testingMethod},
        {"", "enableClosures", OptSpec::U, &o.enableClosures},
        {"", "disableTourists", OptSpec::U, &o.disableTourists},
        {"", "immunizationStart", OptSpec::U, &o.immunizationStart},
        {"", "boosterStart", OptSpec::S, &o.boosterStart},
        {"", "immunizationsPerWeek", OptSpec::S, &o.immunizationsPerWeek},
        {"", "boostersPerWeek", OptSpec::S, &o.boostersPerWeek},
        {"", "immunizationOrder", OptSpec::S, &o.immunizationOrder},
        {"", "vaccinationGroupLevel", OptSpec::S, &o.vaccinationGroupLevel},
        {"", "protectionInfection", OptSpec::S, &o.protectionInfection},
        {"", "protectionInfectionWaning", OptSpec::S, &o.protectionInfectionWaning},
        {"", "protectionSymptomatic", OptSpec::S, &o.protectionSymptomatic},
        {"", "protectionSymptomaticWaning", OptSpec::S, &o.protectionSymptomaticWaning},
        {"", "protectionHospitalization", OptSpec::S, &o.protectionHospitalization},
        {"", "protectionHospitalizationWaning", OptSpec::S, &o.protectionHospitalizationWaning},
        {"", "variantSimilarity", OptSpec::S, &o.variantSimilarity},
        {"", "variantSimilarMultiplier", OptSpec::S, &o.variantSimilarMultiplier},
        {"", "seed", OptSpec::U64, &o.seed},
        {"", "dataDir", OptSpec::S, &o.dataDir},
        {"", "synthAgents", OptSpec::I, &o.synthAgents},
        {"", "synthSeed", OptSpec::U64, &o.synthSeed},
        {"", "synthInfected", OptSpec::D, &o.synthInfected},
        {"", "synthJitter", OptSpec::U, &o.synthJitter},
        {"", "movementPolicy", OptSpec::S, &o.movementPolicy},
        {"", "testingPolicy", OptSpec::S, &o.testingPolicy},
        {"", "closurePolicy", OptSpec::S, &o.closurePolicy},
        {"h", "help", OptSpec::B, &o.help},
        {"", "version", OptSpec::B, &o.version},
    };
}

void assign(const OptSpec& s, const std::string& name, const std::string& v) {
    char* e = nullptr;
    errno = 0;
    switch (s.kind) {
    case OptSpec::U: {
        long long x = std::strtoll(v.c_str(), &e, 10);
        if (e == v.c_str() || *e || x < 0) throw InitError(ERR_BAD_ARG, "bad value for --" + name + ": " + v);
        *(unsigned*)s.target = (unsigned)x;
        break;
    }
    case OptSpec::U64: {
        unsigned long long x = std::strtoull(v.c_str(), &e, 10);
        if (e == v.c_str() || *e) throw InitError(ERR_BAD_ARG, "bad value for --" + name + ": " + v);
        *(uint64_t*)s.target = (uint64_t)x;
        break;
    }
    case OptSpec::I: {
        long x = std::strtol(v.c_str(), &e, 10);
        if (e == v.c_str() || *e) throw InitError(ERR_BAD_ARG, "bad value for --" + name + ": " + v);
        *(int*)s.target = (int)x;
        break;
    }
    case OptSpec::D: {
        double x = std::strtod(v.c_str(), &e);
        if (e == v.c_str() || *e) throw InitError(ERR_BAD_ARG, "bad value for --" + name + ": " + v);
        *(double*)s.target = x;
        break;
    }
    case OptSpec::S: *(std::string*)s.target = v; break;
    case OptSpec::B: {
        if (v == "" || v == "true" || v == "1") *(bool*)s.target = true;
        else if (v == "false" || v == "0") *(bool*)s.target = false;
        else throw InitError(ERR_BAD_ARG, "bad value for --" + name + ": " + v);
        break;
    }
    }
}

}  // namespace

Options parse_options(int argc, const char* const* argv, bool skip_program_name) {
    Options o;
    auto table = option_table(o);
    for (int i = skip_program_name ? 1 : 0; i < argc; ++i) {
        std::string a = argv[i] ? argv[i] : "";
        if (a.size() < 2 || a[0] != '-') throw InitError(ERR_BAD_ARG, "unexpected positional argument: " + a);
        std::string name, value;
        bool hasValue = false;
        bool isLong = a[1] == '-';
        if (isLong) {
            name = a.substr(2);
            auto eq = name.find('=');
            if (eq != std::string::npos) {
                value = name.substr(eq + 1);
                name = name.substr(0, eq);
                hasValue = true;
            }
        } else {
            name = a.substr(1, 1);
            if (a.size() > 2) {
                value = a.substr(2);
                if (!value.empty() && value[0] == '=') value = value.substr(1);
                hasValue = true;
            }
        }
        const OptSpec* spec = nullptr;
        for (auto& s : table)
            if ((isLong && name == s.longName) || (!isLong && name == s.shortName && s.shortName[0])) spec = &s;
        if (!spec) throw InitError(ERR_BAD_ARG, "unknown option: " + a);
        if (spec->kind == OptSpec::B) {
            assign(*spec, spec->longName, hasValue ? value : "");
            continue;
        }
        if (!hasValue) {
            if (i + 1 >= argc) throw InitError(ERR_BAD_ARG, "missing value for option " + a);
            value = argv[++i];
        }
        assign(*spec, spec->longName, value);
    }
    if (o.movementPolicy != "real" && o.movementPolicy != "dummy" && o.movementPolicy != "none")
        throw InitError(ERR_BAD_ARG, "bad value for --movementPolicy: " + o.movementPolicy + " (real, dummy or none)");
    if (o.testingPolicy != "detailed" && o.testingPolicy != "none")
        throw InitError(ERR_BAD_ARG, "bad value for --testingPolicy: " + o.testingPolicy + " (detailed or none)");
    if (o.closurePolicy != "rule" && o.closurePolicy != "none")
        throw InitError(ERR_BAD_ARG, "bad value for --closurePolicy: " + o.closurePolicy + " (rule or none)");
    return o;
}

std::string options_help() {
    Options o;
    std::string out = "panSim - MI355X agent-based epidemic simulator\nOptions:\n";
    for (auto& s : option_table(o)) {
        out += "  ";
        out += s.shortName[0] ? std::string("-") + s.shortName + ", " : std::string("    ");
        out += std::string("--") + s.longName + "\n";
    }
    return out;
}

// ---------------------------------------------------------------------------
// smallTools.cpp list parsing (src/tools/smallTools.cpp:44-108)
// ---------------------------------------------------------------------------
namespace {
template <class T, class Conv>
std::vector<T> split_numbers(const std::string& s, char sep, Conv conv, bool stopAtNewline) {
    std::vector<T> out;
    std::string arg;
    size_t i = 0;
    auto is_ws = [](char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; };
    while (i < s.size()) {
        // operator>> on char skips whitespace
        while (i < s.size() && is_ws(s[i])) ++i;
        if (i >= s.size()) break;
        arg.push_back(s[i++]);
        char peek = i < s.size() ? s[i] : '\0';
        if (peek == sep || (stopAtNewline && peek == '\n')) {
            if (!arg.empty() && (std::isdigit((unsigned char)arg[0]) || arg[0] == '-')) {
                out.push_back(conv(arg));
                arg.clear();
            }
            ++i;  // ignore the separator
            if (stopAtNewline && i < s.size() && s[i] == '\n') break;
        }
    }
    if (!arg.empty() && (std::isdigit((unsigned char)arg[0]) || arg[0] == '-')) out.push_back(conv(arg));
    return out;
}
}  // namespace

std::vector<double> split_double(const std::string& s, char sep) {
    return split_numbers<double>(s, sep, [](const std::string& a) { return std::atof(a.c_str()); }, false);
}
std::vector<float> split_float(const std::string& s, char sep) {
    return split_numbers<float>(s, sep, [](const std::string& a) { return (float)std::atof(a.c_str()); }, true);
}
std::vector<int> split_int(const std::string& s, char sep) {
    return split_numbers<int>(s, sep, [](const std::string& a) { return std::atoi(a.c_str()); }, true);
}

// ---------------------------------------------------------------------------
// Files
// ---------------------------------------------------------------------------
static bool file_exists(const std::string& p) {
    struct stat st;
    return ::stat(p.c_str(), &st) == 0;
}

std::string resolve_path(const Options& opt, const std::string& path) {
    if (path.empty() || file_exists(path)) return path;
    if (!opt.dataDir.empty() && path[0] != '/') {
        std::string p = opt.dataDir + "/" + path;
        if (file_exists(p)) return p;
    }
    return path;
}

static json::Value load_json(const Options& opt, const std::string& path) {
    std::string p = resolve_path(opt, path);
    try {
        return json::parse_file(p);
    } catch (const json::Error& e) {
        std::string m = e.what();
        if (m.rfind("cannot open", 0) == 0) throw InitError(ERR_BAD_INPUT_FILE, m);
        throw InitError(ERR_BAD_JSON, m);
    }
}

// TimeDay(double raw) (src/tools/timeDay.cpp:58-62); raw is a float widened to double.
static void to_timeday(double raw, uint8_t& h, uint8_t& m) {
    if (raw == -1.0) {
        h = 255;
        m = 0;
        return;
    }
    if (raw < 0.0 || raw >= 256.0) throw InitError(ERR_BAD_DATA, "schedule time out of range: " + std::to_string(raw));
    h = (uint8_t)raw;
    m = (uint8_t)std::round(((raw - (double)(int)raw) / 0.6) * 60.0);
}

static uint8_t parse_wb(const std::string& s) {
    if (s.size() != 1) throw InitError(ERR_BAD_DATA, "invalid WB state: " + s);
    switch (std::toupper((unsigned char)s[0])) {
    case 'W': return WB_W;
    case 'N': return WB_N;
    case 'M': return WB_M;
    case 'S': return WB_S;
    case 'D': return WB_D;
    }
    throw InitError(ERR_BAD_DATA, "invalid WB state: " + s);
}

static std::vector<unsigned> parse_days(const std::string& raw) {
    std::string d;
    for (char c : raw) d.push_back((char)std::toupper((unsigned char)c));
    if (d == "ALL") return {0, 1, 2, 3, 4, 5, 6};
    if (d == "WEEKDAYS") return {0, 1, 2, 3, 4};
    if (d == "WEEKENDS") return {5, 6};
    const char* names[] = {"MONDAY", "TUESDAY", "WEDNESDAY", "THURSDAY", "FRIDAY", "SATURDAY", "SUNDAY"};
    for (unsigned i = 0; i < 7; ++i)
        if (d == names[i]) return {i};
    throw InitError(ERR_BAD_DATA, "invalid day in schedule: " + raw);
}

// ---------------------------------------------------------------------------
// Dummy population (src/IO/dataProviders/dataProvider.cpp:102-184) with the
// reference's unseeded RandomGenerator replaced by Philox streams keyed on
// the location / agent index.
// ---------------------------------------------------------------------------
namespace {

struct ChanceStr {
    std::string value;
    long double chance;
    long double diagnosedChance;
};

template <class V>
size_t random_select_index(Stream& rs, const V& list) {
    // dataProvider.h:285-296: long double running sum; overrunning the list
    // returns the first element
    double r = rs.unit();
    size_t i = 0;
    long double pre = list.empty() ? 0.0L : list[0].chance;
    while (pre < r && i != list.size()) {
        ++i;
        if (i < list.size()) pre += list[i].chance;
    }
    return i == list.size() ? 0 : i;
}

struct Person {
    int age;
    std::string sex, preCond, state;
    int typeID;
    bool diagnosed;
    std::vector<std::pair<int, std::string>> locations;  // (typeID, locID)
};

struct Place {
    std::string ID;
    int type;
    int essential;
    double infectious;
    std::string state;
    int capacity;
};

struct IrregularDetail {
    long double chanceForType = 0;
    std::vector<ChanceStr> switchedToWhat;
};

}  // namespace

// ---------------------------------------------------------------------------
// build_scenario
// ---------------------------------------------------------------------------
Scenario build_scenario(const Options& opt) {
    Scenario sc;
    sc.opt = opt;
    if (opt.deltat == 0 || 1440 % opt.deltat != 0)
        throw InitError(ERR_BAD_TIMESTEP, "Bad time step " + std::to_string(opt.deltat) + ", 1440 must be divisible by it");
    sc.stepsPerDay = 1440 / opt.deltat;
    sc.variantMultiplier = split_float(opt.infectiousnessMultiplier, ',');
    sc.progressionScaling = split_float(opt.diseaseProgressionScaling, ',');
    sc.deathScaling = split_float(opt.diseaseProgressionDeathScaling, ',');
    if (sc.variantMultiplier.empty() || sc.variantMultiplier.size() > (size_t)MAX_STRAINS)
        throw InitError(ERR_BAD_ARG, "infectiousnessMultiplier needs 1..7 values");
    if (sc.progressionScaling.size() < sc.variantMultiplier.size() || sc.deathScaling.size() < sc.variantMultiplier.size())
        throw InitError(ERR_BAD_ARG, "diseaseProgressionScaling/diseaseProgressionDeathScaling need one value per variant");
    sc.dOffset = split_int(opt.d_offset, ',');
    {  // Immunization::initializeArgs (immunization.h:139-222)
        VaccinationPlan& v = sc.vacc;
        v.startDay = opt.immunizationStart;
        v.boosterStart = split_int(opt.boosterStart, ',');
        auto read_all = [&](const std::string& path) {
            std::ifstream t(resolve_path(opt, path));
            std::stringstream buf;
            buf << t.rdbuf();
            return buf.str();
        };
        char* end = nullptr;
        v.dailyDoses = (unsigned)(std::strtoul(opt.immunizationsPerWeek.c_str(), &end, 10) / 7);
        if (end != opt.immunizationsPerWeek.c_str() + opt.immunizationsPerWeek.size())
            v.vaccPerWeek = split_float(read_all(opt.immunizationsPerWeek), ',');
        const size_t rounds = v.boosterStart.size();
        v.boosterPerWeek.assign(rounds, {});
        v.boosterAgePercent.assign(rounds, {});
        v.dailyBoosters = (unsigned)(std::strtoul(opt.boostersPerWeek.c_str(), &end, 10) / 7);
        if (end != opt.boostersPerWeek.c_str() + opt.boostersPerWeek.size()) {
            std::ifstream t(resolve_path(opt, opt.boostersPerWeek));
            std::string line;
            size_t r = 0;
            while (r < rounds && std::getline(t, line)) {
                if (line.empty() || line[0] == '#') continue;
                v.boosterPerWeek[r] = split_float(line, ',');
                if (std::getline(t, line)) v.boosterAgePercent[r] = split_float(line, ',');
                ++r;
            }
            if (r != rounds) throw InitError(ERR_BAD_ARG, "booster rounds size mismatch");
        }
        for (size_t r = 0; r < rounds; ++r)
            if (v.boosterAgePercent[r].size() < 10)
                throw InitError(ERR_BAD_ARG, "boostersPerWeek: every booster round needs 10 age-group percentages");
        v.order = split_int(opt.immunizationOrder, ',');
        if (v.order.size() != 10) throw InitError(ERR_BAD_ARG, "immunizationOrder mush have exactly 10 values");
        for (int x : v.order)
            if (x > 10) throw InitError(ERR_BAD_ARG, "immunizationOrder values have to be  less or equal to 10");
        v.groupLevel = split_float(opt.vaccinationGroupLevel, ',');
        if (v.groupLevel.size() < 10) throw InitError(ERR_BAD_ARG, "vaccinationGroupLevel needs 10 values");
    }
    if (opt.testingMethod != "PCR" && opt.testingMethod != "antigen")
        throw InitError(ERR_BAD_TESTING,
                        "unrecognized testingMethod " + opt.testingMethod + " must be either PCR or antigen");
    {  // setupHospitalizations (simulation.h:694-705): missing file leaves the list empty
        std::ifstream t(resolve_path(opt, opt.totalHospitalizations));
        std::string buf;
        while (std::getline(t, buf)) {
            if (buf.empty() || buf[0] == '#') continue;
            sc.totalHospitalizations = split_int(buf, ',');
        }
    }

    // ---- parameters.json -> agent meta scaling (agentMeta.cpp:74-91)
    json::Value params = load_json(opt, opt.parameters);
    std::pair<char, float> sexScaling[2];
    struct AgeInt {
        unsigned from, to;
        float symptoms;
    };
    std::vector<AgeInt> ageScaling;
    std::map<std::string, float> preCondScaling;
    std::vector<std::string> preCondIds;
    {
        auto& sx = params.at("sex");
        if (sx.size() != 2) throw InitError(ERR_BAD_DATA, "parameters: sex must have exactly two entries");
        for (int i = 0; i < 2; ++i) {
            std::string nm = sx[i].at("name").str();
            if (!(sx[0].at("name").str() == "F" || sx[0].at("name").str() == "M"))
                throw InitError(ERR_BAD_DATA, "parameters: wrong gender name");
            sexScaling[i] = std::make_pair(nm.empty() ? '?' : nm[0], (float)sx[i].at("symptoms").num());
        }
        if (sexScaling[0].first == sexScaling[1].first) throw InitError(ERR_BAD_DATA, "parameters: wrong gender name");
        if (sexScaling[0].first == 'M') std::swap(sexScaling[0], sexScaling[1]);
        for (auto& a : params.at("age").arr) {
            long from = (long)a.at("from").integer(), to = (long)a.at("to").integer();
            if (from < 0) throw InitError(ERR_BAD_DATA, "parameters: negative age 'from'");
            AgeInt ai{(unsigned)from, to < 0 ? 0xFFFFFFFFu : (unsigned)to, (float)a.at("symptoms").num()};
            if (ai.to < ai.from) throw InitError(ERR_BAD_DATA, "parameters: negative age interval");
            ageScaling.push_back(ai);
        }
        for (auto& c : params.at("preCondition").arr) {
            preCondScaling.emplace(c.at("ID").str(), (float)c.at("symptoms").num());
            preCondIds.push_back(c.at("ID").str());
        }
    }

    // ---- progression config + matrices (dataProvider.cpp:14-58, dynamicPPState.cpp:118-184)
    json::Value pconf = load_json(opt, opt.progression);
    {
        auto& si = pconf.at("stateInformation");
        sc.st.names = si.at("stateNames").strings();
        unsigned ns = (unsigned)sc.st.names.size();
        if (ns == 0 || ns > 127) throw InitError(ERR_BAD_DATA, "progression: bad number of states");
        auto idx = [&](const std::string& n) -> unsigned {
            for (unsigned i = 0; i < ns; ++i)
                if (sc.st.names[i] == n) return i;
            throw InitError(ERR_BAD_DATA, "progression: unknown state name " + n);
        };
        sc.st.wb.assign(ns, WB_W);
        sc.st.infectious.assign(ns, 0.f);
        sc.st.accuracyPCR.assign(ns, 0.f);
        sc.st.accuracyAntigen.assign(ns, 0.f);
        sc.st.susceptible.assign(ns, 0);
        sc.st.infected.assign(ns, 0);
        for (auto& n : sc.st.names) sc.st.header += n + '\t';
        for (auto& s : pconf.at("states").arr) {
            unsigned i = idx(s.at("stateName").str());
            sc.st.wb[i] = parse_wb(s.at("WB").str());
            sc.st.infectious[i] = (float)s.at("infectious").num();
            sc.st.accuracyPCR[i] = (float)s.at("accuracyPCR").num();
            sc.st.accuracyAntigen[i] = (float)s.at("accuracyAntigen").num();
        }
        for (auto& n : si.at("infectedStates").strings()) sc.st.infected[idx(n)] = 1;
        for (auto& n : si.at("susceptibleStates").strings()) sc.st.susceptible[idx(n)] = 1;
        sc.st.firstInfected = idx(si.at("firstInfectedState").str());
        sc.st.nonCOVIDDead = idx(si.at("nonCOVIDDeadState").str());
        sc.st.deadState = 0;
        for (unsigned i = 0; i < ns; ++i)
            if (sc.st.wb[i] == WB_D) {
                sc.st.deadState = i;
                break;
            }
    }
    struct MatrixEntry {
        unsigned ageBegin, ageEnd;
        std::string preCond;
        unsigned index;
    };
    std::vector<MatrixEntry> mentries;
    {
        std::string ppath = resolve_path(opt, opt.progression);
        std::string dir = ppath.substr(0, ppath.find_last_of('/'));
        if (ppath.find('/') == std::string::npos) dir = ".";
        struct Raw {
            unsigned ab, ae;
            std::string pc, file;
        };
        std::vector<Raw> raws;
        for (auto& m : pconf.at("transitionMatrices").arr) {
            auto ages = m.at("age").numbers();
            if (ages.size() != 2) throw InitError(ERR_BAD_DATA, "progression: age must have two values");
            raws.push_back({(unsigned)(int)ages[0], (unsigned)(int)ages[1], m.at("preCond").str(), m.at("fileName").str()});
        }
        // std::map<ProgressionType,...> order: (ageBegin, preCond)
        std::stable_sort(raws.begin(), raws.end(), [](const Raw& a, const Raw& b) {
            if (a.ab == b.ab) return a.pc < b.pc;
            return a.ab < b.ab;
        });
        raws.erase(std::unique(raws.begin(), raws.end(),
                               [](const Raw& a, const Raw& b) { return a.ab == b.ab && a.pc == b.pc; }),
                   raws.end());
        std::vector<unsigned> ageBegins;
        for (auto& r : raws)
            if (std::find(ageBegins.begin(), ageBegins.end(), r.ab) == ageBegins.end()) ageBegins.push_back(r.ab);
        unsigned numberCond = (unsigned)preCondIds.size();
        unsigned maxIndex = 0;
        std::vector<std::pair<unsigned, std::string>> files;
        for (auto& r : raws) {
            unsigned ageIndex = (unsigned)(std::find(ageBegins.begin(), ageBegins.end(), r.ab) - ageBegins.begin());
            unsigned condIndex = (unsigned)(std::find(preCondIds.begin(), preCondIds.end(), r.pc) - preCondIds.begin());
            unsigned index = ageIndex * numberCond + condIndex;
            mentries.push_back({r.ab, r.ae, r.pc, index});
            files.emplace_back(index, dir + "/" + r.file);
            maxIndex = std::max(maxIndex, index);
        }
        if (mentries.empty() || maxIndex >= 255) throw InitError(ERR_BAD_DATA, "progression: bad matrix list");
        sc.prog.resize(maxIndex + 1);
        unsigned ns = (unsigned)sc.st.names.size();
        for (auto& f : files) {
            json::Value mat;
            try {
                mat = json::parse_file(f.second);
            } catch (const json::Error& e) {
                throw InitError(ERR_BAD_INPUT_FILE, e.what());
            }
            auto& states = mat.at("states").arr;
            ProgressionMatrixRaw pm;
            pm.rows.resize(states.size());
            auto sidx = [&](const std::string& n) -> unsigned {
                for (unsigned i = 0; i < states.size(); ++i)
                    if (states[i].at("stateName").str() == n) return i;
                throw InitError(ERR_BAD_DATA, "progression matrix: wrong state name " + n);
            };
            if (states.size() != ns) throw InitError(ERR_BAD_DATA, "progression matrix: state count mismatch in " + f.second);
            for (unsigned i = 0; i < states.size(); ++i) {
                auto& s = states[i];
                auto avg = s.at("avgLength").numbers();
                auto mx = s.at("maxlength").numbers();
                if (avg.empty() || mx.size() != avg.size() || avg.size() > (size_t)MAX_STRAINS)
                    throw InitError(ERR_BAD_DATA, "progression matrix: bad avgLength/maxlength");
                auto& row = pm.rows[i];
                // LengthOfState ctor (basicLengthAbstract.cpp:31-44)
                for (size_t v = 0; v < avg.size(); ++v) {
                    row.avgLength[v] = (float)avg[v];
                    row.maxLength[v] = (float)mx[v];
                    if (row.maxLength[v] == -1.0f) row.maxLength[v] = FLT_MAX;
                }
                for (size_t v = avg.size(); v < (size_t)MAX_STRAINS; ++v) {
                    row.avgLength[v] = row.avgLength[avg.size() - 1];
                    row.maxLength[v] = row.maxLength[avg.size() - 1];
                }
                double sumChance = 0.0;
                auto& progs = s.at("progressions").arr;
                for (auto& t : progs) {
                    unsigned to = sidx(t.at("name").str());
                    double ch = t.at("chance").num();
                    sumChance += ch;
                    if (t.at("isBadProgression").boolean()) row.bad.push_back({to, (float)ch});
                    else row.neutral.push_back({to, (float)ch});
                }
                if (row.neutral.empty() && row.bad.empty()) row.neutral.push_back({i, 1.0f});
                if (!progs.empty() && !(0.9999 < sumChance && sumChance < 1.0001))
                    throw InitError(ERR_BAD_DATA, "progression matrix: chances of " + s.at("stateName").str() +
                                                      " do not sum to 1");
            }
            sc.prog[f.first] = std::move(pm);
        }
    }

    // ---- closure rules
    {
        json::Value cr = load_json(opt, opt.closures);
        for (auto& r : cr.at("rules").arr) {
            ClosureRuleRaw x;
            x.name = r.at("name").str();
            x.conditionType = r.at("conditionType").str();
            x.threshold = r.at("threshold").num();
            x.threshold2 = r.at("threshold2").num();
            x.parameter = r.has("parameter") ? r.at("parameter").str() : std::string();
            x.closeAfter = (int)r.at("closeAfter").integer();
            x.openAfter = (int)r.at("openAfter").integer();
            for (auto& t : r.at("locationTypesToClose").arr) x.locationTypesToClose.push_back((int)t.integer());
            sc.rules.push_back(x);
        }
    }

    // ---- location types
    json::Value ltj = load_json(opt, opt.locationTypes);
    {
        auto& lt = sc.lt;
        lt.publicSpace = (unsigned)ltj.at("publicSpace").integer();
        lt.home = (unsigned)ltj.at("home").integer();
        lt.hospital = (unsigned)ltj.at("hospital").integer();
        lt.doctor = (unsigned)ltj.at("doctor").integer();
        lt.school = (unsigned)ltj.at("school").integer();
        lt.classroom = (unsigned)ltj.at("classroom").integer();
        lt.work = (unsigned)ltj.at("work").integer();
        lt.nurseryhome = (unsigned)ltj.at("nurseryhome").integer();
        int maxId = 0;
        for (auto& t : ltj.at("types").arr) {
            lt.types.emplace_back((int)t.at("ID").integer(), t.at("name").str());
            maxId = std::max(maxId, (int)t.at("ID").integer());
        }
        if (lt.types.empty() || maxId + 1 > 255) throw InitError(ERR_BAD_DATA, "location types: IDs must be in 0..254");
        lt.cemeteryType = (unsigned)maxId + 1;
    }

    // ---- agent types + schedules (agentsList.h:94-119, agentType.cpp:4-26; dataProvider.cpp:67-79)
    json::Value atj = load_json(opt, opt.agentTypes);
    std::map<int, unsigned> typeMap;
    std::map<int, std::vector<unsigned>> aTypeToLocTypes;
    {
        auto& types = atj.at("types").arr;
        sc.sched.nTypes = (unsigned)types.size();
        if (sc.sched.nTypes == 0 || sc.sched.nTypes > 255) throw InitError(ERR_BAD_DATA, "agent types: bad count");
        sc.sched.eventOffset.assign(sc.sched.nTypes * 28 + 1, 0);
        unsigned tidx = 0;
        for (auto& t : types) {
            int id = (int)t.at("ID").integer();
            typeMap.emplace(id, tidx);
            std::set<unsigned> locs{sc.lt.hospital, sc.lt.publicSpace, sc.lt.home, sc.lt.doctor};
            for (auto& su : t.at("schedulesUnique").arr) {
                uint8_t wb = parse_wb(su.at("WB").str());
                auto days = parse_days(su.at("dayType").str());
                std::vector<Event> evs;
                for (auto& e : su.at("schedule").arr) {
                    Event ev;
                    int ty = (int)e.at("type").integer();
                    if (ty < 0 || ty > 254) throw InitError(ERR_BAD_DATA, "schedule: location type out of range");
                    ev.locType = (uint32_t)ty;
                    ev.chance = (float)e.at("chance").num();
                    to_timeday((double)(float)e.at("start").num(), ev.start_h, ev.start_m);
                    to_timeday((double)(float)e.at("end").num(), ev.end_h, ev.end_m);
                    to_timeday((double)(float)e.at("duration").num(), ev.dur_h, ev.dur_m);
                    evs.push_back(ev);
                    locs.insert((unsigned)ty);
                }
                if (wb > WB_S) throw InitError(ERR_BAD_DATA, "schedule: WB state D has no schedule slot");
                for (unsigned d : days) {
                    // addSchedule: insert at the slot start, shift later offsets
                    unsigned slot = 7 * wb + d + tidx * 28;
                    for (size_t i = slot + 1; i < sc.sched.eventOffset.size(); ++i)
                        sc.sched.eventOffset[i] += (uint32_t)evs.size();
                    sc.sched.events.insert(sc.sched.events.begin() + sc.sched.eventOffset[slot], evs.begin(), evs.end());
                }
            }
            aTypeToLocTypes.emplace(id, std::vector<unsigned>(locs.begin(), locs.end()));
            ++tidx;
        }
    }

    // ---- population
    if (opt.synthAgents > 0) {
        build_synthetic(sc);
    } else {
    bool randomAgents = opt.numagents != -1;
    bool randomLocs = opt.numlocs != -1;
    bool needConfigRandom = randomAgents || randomLocs || opt.randomStates;
    json::Value crj;
    std::vector<ChanceStr> locTypeDist, preCondDist, agentTypeDist;
    struct AgeStates {
        int ageStart, ageEnd;
        std::vector<ChanceStr> dist;
    };
    std::vector<AgeStates> stateDist;
    std::map<std::string, IrregularDetail> irregular;
    if (needConfigRandom) {
        crj = load_json(opt, opt.configRandom);
        auto rd = [](const json::Value& arr) {
            std::vector<ChanceStr> v;
            for (auto& e : arr.arr) v.push_back({e.at("value").str(), (long double)e.at("chance").num(), 0.0L});
            return v;
        };
        locTypeDist = rd(crj.at("locationTypeDistribution"));
        preCondDist = rd(crj.at("preCondDistribution"));
        agentTypeDist = rd(crj.at("agentTypeDistribution"));
        for (auto& s : crj.at("stateDistribution").arr) {
            AgeStates as{(int)s.at("ageStart").integer(), (int)s.at("ageEnd").integer(), {}};
            for (auto& d : s.at("distribution").arr)
                as.dist.push_back({d.at("value").str(), (long double)d.at("chance").num(),
                                   d.has("diagnosedChance") ? (long double)d.at("diagnosedChance").num() : 0.0L});
            stateDist.push_back(as);
        }
        for (auto& d : crj.at("irregularLocationChance").at("detailsOfChances").arr) {
            IrregularDetail det;
            det.chanceForType = (long double)d.at("chanceForType").num();
            det.switchedToWhat = rd(d.at("switchedToWhat"));
            irregular[d.at("value").str()] = det;
        }
    }

    std::vector<Place> places;
    std::map<unsigned, std::vector<std::string>> typeToLocationMapping;
    std::vector<Person> people;
    const uint64_t seed = opt.seed;

    if (!randomLocs) {
        json::Value lj = load_json(opt, opt.locations);
        for (auto& l : lj.at("places").arr) {
            Place p;
            p.ID = l.at("ID").str();
            p.type = (int)l.at("type").integer();
            p.essential = l.has("essential") ? (int)l.at("essential").integer() : 0;
            p.infectious = l.at("infectious").num();
            p.state = l.at("state").str();
            p.capacity = l.has("capacity") ? (int)l.at("capacity").integer() : 0;
            places.push_back(p);
            // readLocations builds the type mapping only when agents come from a file
            if (!randomAgents) typeToLocationMapping[(unsigned)p.type].push_back(p.ID);
        }
    } else {
        // randomLocations (dataProvider.cpp:102-133)
        for (int i = 0; i < opt.numlocs; ++i) {
            Stream rs(seed, (uint32_t)i, 0, PH_POP_LOCATION);
            Place cur{};
            cur.ID = std::to_string(i);
            cur.type = std::atoi(locTypeDist[random_select_index(rs, locTypeDist)].value.c_str());
            if (cur.type == 33) {
                auto it = typeToLocationMapping.find(3);
                if (it != typeToLocationMapping.end()) {
                    auto& schools = it->second;
                    cur.ID += "_" + schools[rs.uint_below((uint32_t)schools.size())];
                } else {
                    cur.type = 3;
                }
            }
            typeToLocationMapping[(unsigned)cur.type].push_back(cur.ID);
            cur.state = "ON";
            cur.capacity = 100;
            cur.infectious = 1.0;
            if (cur.type == 4 || cur.type == 7 || cur.type == 8) cur.essential = rs.unit() < 0.1;
            else if (cur.type == 12 || cur.type == 14) cur.essential = 1;
            else cur.essential = 0;
            places.push_back(cur);
        }
    }

    auto singleRandomState = [&](unsigned age, Stream& rs) -> std::pair<std::string, long double> {
        for (auto& s : stateDist)
            if (s.ageStart <= (int)age && (int)age < s.ageEnd) {
                size_t i = random_select_index(rs, s.dist);
                return {s.dist[i].value, s.dist[i].diagnosedChance};
            }
        throw InitError(ERR_BAD_DATA, "configRandom: no state distribution for age " + std::to_string(age));
    };

    if (!randomAgents) {
        json::Value aj = load_json(opt, opt.agents);
        for (auto& p : aj.at("people").arr) {
            Person x;
            x.age = (int)p.at("age").integer();
            x.sex = p.at("sex").str();
            x.preCond = p.at("preCond").str();
            x.state = p.at("state").str();
            x.typeID = (int)p.at("typeID").integer();
            x.diagnosed = p.has("diagnosed") ? p.at("diagnosed").boolean() : false;
            for (auto& l : p.at("locations").arr) x.locations.emplace_back((int)l.at("typeID").integer(), l.at("locID").str());
            people.push_back(x);
        }
        if (opt.randomStates) {
            for (size_t i = 0; i < people.size(); ++i) {
                Stream rs(seed, (uint32_t)i, 1, PH_POP_AGENT);
                auto sd = singleRandomState((unsigned)people[i].age, rs);
                people[i].state = sd.first;
                people[i].diagnosed = rs.unit() < sd.second;
            }
        }
    } else {
        // randomAgents (dataProvider.cpp:135-184)
        people.reserve(opt.numagents);
        for (int i = 0; i < opt.numagents; ++i) {
            Stream rs(seed, (uint32_t)i, 0, PH_POP_AGENT);
            Person cur{};
            cur.age = (int)rs.uint_below(90);
            cur.sex = rs.unit() < 0.5 ? "M" : "F";
            cur.preCond = preCondDist[random_select_index(rs, preCondDist)].value;
            auto sd = singleRandomState((unsigned)cur.age, rs);
            cur.state = sd.first;
            cur.diagnosed = rs.unit() < sd.second;
            cur.typeID = std::atoi(agentTypeDist[random_select_index(rs, agentTypeDist)].value.c_str());
            auto reqIt = aTypeToLocTypes.find(cur.typeID);
            if (reqIt == aTypeToLocTypes.end()) throw InitError(ERR_BAD_DATA, "configRandom: unknown agent type");
            for (unsigned l : reqIt->second) {
                std::pair<int, std::string> loc{(int)l, std::string()};
                IrregularDetail& irr = irregular[std::to_string(l)];
                bool found = false;
                if (rs.unit() < irr.chanceForType) {
                    std::string irrType = irr.switchedToWhat[random_select_index(rs, irr.switchedToWhat)].value;
                    auto& newPossible = typeToLocationMapping[(unsigned)std::atoi(irrType.c_str())];
                    if (newPossible.empty()) {
                        for (auto& sw : irr.switchedToWhat) {
                            newPossible = typeToLocationMapping[(unsigned)std::atoi(sw.value.c_str())];
                            if (!newPossible.empty()) {
                                found = true;
                                break;
                            }
                        }
                    }
                    if (found) loc.second = newPossible[rs.uint_below((uint32_t)newPossible.size())];
                } else {
                    auto& possible = typeToLocationMapping[l];
                    if (!possible.empty()) {
                        loc.second = possible[rs.uint_below((uint32_t)possible.size())];
                        found = true;
                    }
                }
                if (!found) loc.second = places[rs.uint_below((uint32_t)places.size())].ID;
                cur.locations.push_back(loc);
            }
            people.push_back(std::move(cur));
        }
    }

    // ---- initLocations (locationList.h:90-222)
    std::map<std::string, unsigned> idMapping;
    {
        auto& L = sc.locs;
        std::vector<std::string> classroomIds;
        unsigned idx = 0;
        for (auto& p : places) {
            idMapping.emplace(p.ID, idx);
            if (p.type < 0 || p.type > 254) throw InitError(ERR_BAD_DATA, "location type out of range");
            L.ids.push_back(p.ID);
            L.type.push_back((uint32_t)p.type);
            L.infectiousness.push_back((float)p.infectious);
            L.capacity.push_back((uint32_t)p.capacity);
            L.essential.push_back((uint8_t)p.essential);
            if ((unsigned)p.type == sc.lt.classroom) {
                L.classrooms.push_back(idx);
                classroomIds.push_back(p.ID);
            }
            if ((unsigned)p.type == sc.lt.school) L.schools.push_back(idx);
            std::string st = p.state;  // the reference's toupper lambda discards its result
            if (st == "ON" || st == "OPEN") L.open.push_back(1);
            else if (st == "OFF" || st == "CLOSED") L.open.push_back(0);
            else throw InitError(ERR_BAD_DATA, "location with wrong state: " + p.state);
            ++idx;
        }
        // cemetery
        L.cemetery = idx;
        L.ids.push_back("cemetery");
        L.type.push_back(sc.lt.cemeteryType);
        L.infectiousness.push_back(0.0f);
        L.capacity.push_back(0xFFFFFFFFu);
        L.essential.push_back(1);
        L.open.push_back(1);
        L.n = idx + 1;
        // classroom-school pairing
        std::vector<uint32_t> schoolOf(L.classrooms.size());
        for (size_t i = 0; i < L.classrooms.size(); ++i) {
            const std::string& s = classroomIds[i];
            size_t pos = s.find('_');
            if (pos == std::string::npos)
                throw InitError(ERR_BAD_DATA, "classroom id does not have class_school structure " + s);
            auto it = idMapping.find(s.substr(pos + 1));
            if (it == idMapping.end())
                throw InitError(ERR_BAD_DATA, "classroom id does not have class_school structure: " + s +
                                                  " school ID not found " + s.substr(pos + 1));
            schoolOf[i] = it->second;
        }
        std::vector<size_t> perm(L.classrooms.size());
        for (size_t i = 0; i < perm.size(); ++i) perm[i] = i;
        std::stable_sort(perm.begin(), perm.end(), [&](size_t a, size_t b) { return schoolOf[a] < schoolOf[b]; });
        std::vector<uint32_t> sortedSchoolOf(perm.size()), sortedClass(perm.size());
        for (size_t i = 0; i < perm.size(); ++i) {
            sortedSchoolOf[i] = schoolOf[perm[i]];
            sortedClass[i] = L.classrooms[perm[i]];
        }
        L.classrooms = sortedClass;
        for (size_t i = 0; i < L.schools.size(); ++i) {
            auto it = std::find(sortedSchoolOf.begin(), sortedSchoolOf.end(), L.schools[i]);
            if (it != sortedSchoolOf.end()) L.classroomOffsets.push_back((uint32_t)(it - sortedSchoolOf.begin()));
            else {
                if (L.classroomOffsets.empty()) L.classroomOffsets.push_back(0);
                L.classroomOffsets.push_back(L.classroomOffsets.back());
            }
        }
        L.classroomOffsets.push_back((uint32_t)sortedSchoolOf.size());
    }

    // ---- initAgents (agentsList.h:137-271)
    {
        auto& A = sc.agents;
        A.locOffset.push_back(0);
        for (auto& person : people) {
            if (opt.disableTourists && person.typeID == 9) continue;
            // progression lookup: std::map::find with the heterogeneous
            // comparators of progressionType.cpp:338-360
            unsigned age = (unsigned)person.age;
            const std::string& pc = person.preCond;
            auto eLess = [&](const MatrixEntry& e) {  // e < (age, pc)
                if (e.ageBegin < age) {
                    if (age <= e.ageEnd) return e.preCond < pc;
                    return true;
                }
                return false;
            };
            auto kLess = [&](const MatrixEntry& e) {  // (age, pc) < e
                if (age < e.ageEnd) {
                    if (e.ageBegin <= age) return pc < e.preCond;
                    return true;
                }
                return false;
            };
            size_t lb = 0;
            while (lb < mentries.size() && eLess(mentries[lb])) ++lb;
            if (lb == mentries.size() || kLess(mentries[lb]))
                throw InitError(ERR_BAD_DATA, "no progression matrix for age " + std::to_string(age) + " preCond " + pc);
            unsigned progID = mentries[lb].index;
            unsigned stateIdx = 0xFFFF;
            for (unsigned s = 0; s < sc.st.names.size(); ++s)
                if (sc.st.names[s] == person.state) stateIdx = s;
            if (stateIdx == 0xFFFF) throw InitError(ERR_BAD_DATA, "unknown agent state " + person.state);
            if (person.sex.size() != 1) throw InitError(ERR_BAD_DATA, "invalid gender " + person.sex);
            // BasicAgentMeta (agentMeta.cpp:95-126)
            float scaling = 1.0f;
            uint8_t sex;
            if (person.sex[0] == 'F') {
                scaling *= sexScaling[0].second;
                sex = 0;
            } else if (person.sex[0] == 'M') {
                scaling *= sexScaling[1].second;
                sex = 1;
            } else
                throw InitError(ERR_BAD_DATA, "invalid gender " + person.sex);
            bool ageFound = false;
            for (auto& ai : ageScaling)
                if (ai.from <= age && age < ai.to) {
                    scaling *= ai.symptoms;
                    ageFound = true;
                    break;
                }
            if (!ageFound) throw InitError(ERR_BAD_DATA, "age not defined in parameters: " + std::to_string(age));
            auto pcIt = preCondScaling.find(pc);
            if (pcIt == preCondScaling.end()) throw InitError(ERR_BAD_DATA, "preCondition not defined: " + pc);
            scaling *= pcIt->second;
            auto tIt = typeMap.find(person.typeID);
            if (tIt == typeMap.end()) throw InitError(ERR_BAD_DATA, "invalid agent type " + std::to_string(person.typeID));

            A.age.push_back((uint8_t)age);
            A.sex.push_back(sex);
            A.preCond.push_back((uint8_t)std::atoi(pc.c_str()));
            A.typeIdx.push_back((uint8_t)tIt->second);
            A.progID.push_back((uint8_t)progID);
            A.state.push_back((uint8_t)stateIdx);
            A.diagnosed.push_back(person.diagnosed ? 1 : 0);
            A.scalingAgeSex.push_back(scaling);
            // locations, sorted by type (std::sort in the reference; stable here)
            auto locs = person.locations;
            std::stable_sort(locs.begin(), locs.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
            uint32_t home = 0;
            bool hasHome = false;
            for (auto& l : locs) {
                auto it = idMapping.find(l.second);
                if (it == idMapping.end()) throw InitError(ERR_BAD_DATA, "invalid location ID " + l.second);
                if (l.first < 0 || l.first > 254) throw InitError(ERR_BAD_DATA, "agent location type out of range");
                A.possibleLocations.push_back(it->second);
                A.possibleTypes.push_back((uint8_t)l.first);
                if (!hasHome && (unsigned)l.first == sc.lt.home) {
                    home = it->second;
                    hasHome = true;
                }
            }
            A.locOffset.push_back((uint32_t)A.possibleLocations.size());
            A.location.push_back(hasHome ? home : 0);
        }
        A.n = (uint32_t)A.age.size();
    }
    }  // generic population
    // healthcare workers (simulation.h:667-692)
    {
        auto& A = sc.agents;
        A.healthcareWorker.assign(A.n, 0);
        for (uint32_t i = 0; i < A.n; ++i)
            for (uint32_t j = A.locOffset[i]; j < A.locOffset[i + 1]; ++j) {
                uint32_t t = sc.locs.type[A.possibleLocations[j]];
                if (A.possibleTypes[j] == 4 && (t == 12 || t == 14)) {
                    A.healthcareWorker[i] = 1;
                    break;
                }
            }
    }
    if (sc.agents.n == 0) throw InitError(ERR_BAD_DATA, "empty population");
    sc.statesHeader = sc.st.header + "H\tT\tP1\tP2\tQ\tQT\tNQ\tMUT\tHOM\tVAC\tNI\tINF\tREINF\tBSTR\tIMM\tHCI\tHCE\tINFV\tINFH\tVNI";
    return sc;
}

}  // namespace ps
