// Output files of the reference's host loop:
//  * the per-agent statistics JSON of --outAgentStat
//    (AgentStatOutput, src/IO/outputs/agentStatOutput.cpp:5-44, written by
//    JSONWriter::writeFile, src/IO/outputs/JSONWriter.cpp:15-25, a rapidjson
//    PrettyWriter with 4-space indentation), and the --diags summary printed
//    after it (AgentList::printAgentStatJSON, include/agentsList.h:280-327);
//  * the infectious-list dump of --dumpLocationInfectiousList
//    (BasicInfection::DumpStore, include/policies/infectionPolicies.h:27-85,
//    filled by dumpLocationInfectiousList :262-441).
#pragma once
#include <algorithm>
#include <cstdint>
#include <fstream>
#include <iostream>
#include <iterator>
#include <string>
#include <vector>

namespace ps {

// The per-agent statistics (AgentStats, include/agentStats.h:5-24), one
// column per field.
struct AgentStatColumns {
    std::vector<uint32_t> infectedTimestamp, infectedCount, variant, infectedLocation, diagnosedTimestamp,
        quarantinedTimestamp, quarantinedUntilTimestamp, daysInQuarantine, hospitalizedTimestamp,
        hospitalizedUntilTimestamp, immunizationTimestamp, immunizationCount, recoveredTimestamp, worstState,
        worstStateTimestamp, worstStateEndTimestamp;
};

// A JSON string as rapidjson writes it (quotation mark, backslash and
// control characters escaped).
inline void json_string(std::ostream& os, const std::string& s) {
    static const char* hex = "0123456789ABCDEF";
    os << '"';
    for (unsigned char ch : s) {
        switch (ch) {
            case '"': os << "\\\""; break;
            case '\\': os << "\\\\"; break;
            case '\b': os << "\\b"; break;
            case '\f': os << "\\f"; break;
            case '\n': os << "\\n"; break;
            case '\r': os << "\\r"; break;
            case '\t': os << "\\t"; break;
            default:
                if (ch < 0x20) os << "\\u00" << hex[ch >> 4] << hex[ch & 15];
                else os << (char)ch;
        }
    }
    os << '"';
}

// {"Statistics": [ one object per agent ]} in the layout of a 4-space
// PrettyWriter.  An empty path writes nothing (the reference opens "" and
// fails silently).
inline void write_agent_stats_json(const std::string& path, const AgentStatColumns& c,
                                   const std::vector<std::string>& stateNames) {
    if (path.empty()) return;
    std::ofstream f(path);
    if (!f) return;
    const size_t n = c.infectedTimestamp.size();
    f << "{\n    \"Statistics\": [";
    if (n == 0) {
        f << "]\n}";
        return;
    }
    const std::string in3 = "            ", in4 = "                ";
    for (size_t i = 0; i < n; ++i) {
        f << (i ? ",\n" : "\n") << "        {\n";
        auto num = [&](const char* key, uint64_t v, bool last = false) {
            f << in3 << '"' << key << "\": " << v << (last ? "\n" : ",\n");
        };
        num("ID", i);
        num("infectionTime", c.infectedTimestamp[i]);
        num("infectedCount", c.infectedCount[i]);
        num("variant", c.variant[i]);
        num("InfectionLoc", c.infectedLocation[i]);
        num("diagnosisTime", c.diagnosedTimestamp[i]);
        num("quarantinedTime", c.quarantinedTimestamp[i]);
        num("quarantinedUntilTime", c.quarantinedUntilTimestamp[i]);
        num("daysInQuarantine", c.daysInQuarantine[i]);
        num("hospitalizedTime", c.hospitalizedTimestamp[i]);
        num("hospitalizedUntilTime", c.hospitalizedUntilTimestamp[i]);
        num("immunizationTimestamp", c.immunizationTimestamp[i]);
        num("immunizationCount", c.immunizationCount[i]);
        num("recoveredTimestamp", c.recoveredTimestamp[i]);
        f << in3 << "\"worstState\": {\n" << in4 << "\"name\": ";
        const uint32_t ws = c.worstState[i];
        json_string(f, ws < stateNames.size() ? stateNames[ws] : std::string());
        f << ",\n" << in4 << "\"begin\": " << c.worstStateTimestamp[i] << ",\n" << in4 << "\"end\": ";
        if (c.worstStateEndTimestamp[i] == 0) f << -1;
        else f << c.worstStateEndTimestamp[i];
        f << "\n" << in3 << "}\n        }";
    }
    f << "\n    ]\n}";
}

// The --diags > 0 summary: days in quarantine of the agents aged 21..65,
// then the number of infections by location type (agents never infected
// count at the cemetery's type), as the unique sorted types and their counts,
// each padded with zeros to (declared types + cemetery) + 2 entries.
inline void print_agent_stat_diags(std::ostream& os, const AgentStatColumns& c, const std::vector<uint8_t>& age,
                                   const std::vector<uint32_t>& locType, uint32_t cemeteryLocation,
                                   size_t generalTypes) {
    const size_t n = c.infectedTimestamp.size();
    uint64_t days = 0;
    for (size_t i = 0; i < n; ++i)
        if (age[i] > 20 && age[i] <= 65) days += c.daysInQuarantine[i];
    os << "Total number of days in quarantine for agents 20<age<=65" << std::endl;
    os << (unsigned)days << std::endl;
    std::vector<uint32_t> keys(n);
    for (size_t i = 0; i < n; ++i)
        keys[i] = locType[c.infectedTimestamp[i] == 0xFFFFFFFFu ? cemeteryLocation : c.infectedLocation[i]];
    std::sort(keys.begin(), keys.end());
    std::vector<uint32_t> ids(generalTypes + 2, 0u), counts(generalTypes + 2, 0u);
    size_t u = 0;
    for (size_t i = 0; i < n; ++i) {
        if (i > 0 && keys[i] == keys[i - 1]) {
            counts[u - 1]++;
            continue;
        }
        if (u < ids.size()) {
            ids[u] = keys[i];
            counts[u] = 1;
        }
        ++u;
    }
    std::copy(ids.begin(), ids.end(), std::ostream_iterator<unsigned>(os, " "));
    os << std::endl;
    std::copy(counts.begin(), counts.end(), std::ostream_iterator<unsigned>(os, " "));
    os << std::endl;
}

// The infectious-list dump: one insert per infection pass (timestamp,
// variant) that infected anyone.  Per insert: the locations of the new
// infections (ascending), the newly infected agents by location, and for
// each of those locations the agents with non-zero infectiousness there and
// their infectiousness (ascending agent index within a location, the order
// of the reference's stable sort by location).
class InfectiousListDump {
    std::vector<unsigned> timestamp_, variant_, locationIds_, locationOffsets_{0}, infected_, infectedOffsets_{0},
        infectious_, infectiousOffsets_{0};
    std::vector<float> infectiousness_;

public:
    // newAgent/newLoc: the infected agents and their locations; infAgent/
    // infLoc/infValue: the agents of non-zero infectiousness at those
    // locations (any order).
    void insert(unsigned timestamp, unsigned variant, const std::vector<uint32_t>& newAgent,
                const std::vector<uint32_t>& newLoc, const std::vector<uint32_t>& infAgent,
                const std::vector<uint32_t>& infLoc, const std::vector<float>& infValue) {
        if (newAgent.empty()) return;  // no new infection: no insert
        std::vector<std::pair<uint32_t, uint32_t>> nw(newAgent.size());
        for (size_t j = 0; j < nw.size(); ++j) nw[j] = {newLoc[j], newAgent[j]};
        std::sort(nw.begin(), nw.end());
        struct Inf {
            uint32_t loc, agent;
            float value;
        };
        std::vector<Inf> inf(infAgent.size());
        for (size_t j = 0; j < inf.size(); ++j) inf[j] = {infLoc[j], infAgent[j], infValue[j]};
        std::sort(inf.begin(), inf.end(), [](const Inf& x, const Inf& y) {
            return x.loc != y.loc ? x.loc < y.loc : x.agent < y.agent;
        });
        timestamp_.push_back(timestamp);
        variant_.push_back(variant);
        const unsigned prev = (unsigned)infectious_.size();
        size_t k = 0;
        unsigned lengthSoFar = 0;
        for (size_t j = 0; j < nw.size(); ++j) {
            if (j > 0 && nw[j].first == nw[j - 1].first) continue;
            const uint32_t l = nw[j].first;
            locationIds_.push_back(l);
            while (k < inf.size() && inf[k].loc < l) ++k;
            while (k < inf.size() && inf[k].loc == l) {
                infectious_.push_back(inf[k].agent);
                infectiousness_.push_back(inf[k].value);
                ++lengthSoFar;
                ++k;
            }
            infectiousOffsets_.push_back(prev + lengthSoFar);
        }
        locationOffsets_.push_back((unsigned)locationIds_.size());
        for (const auto& p : nw) infected_.push_back(p.second);
        infectedOffsets_.push_back((unsigned)infected_.size());
    }

    void write(const std::string& path) const {
        std::ofstream f(path);
        if (!f) return;
        auto line = [&](const auto& v) {
            using T = typename std::decay_t<decltype(v)>::value_type;
            std::copy(v.begin(), v.end(), std::ostream_iterator<T>(f, " "));
            f << "\n";
        };
        f << timestamp_.size() << "\n";
        line(timestamp_);
        line(variant_);
        line(locationOffsets_);
        line(locationIds_);
        line(infectedOffsets_);
        line(infected_);
        line(infectiousOffsets_);
        line(infectious_);
        line(infectiousness_);
    }
};

}  // namespace ps
