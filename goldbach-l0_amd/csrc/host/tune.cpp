// Tuned knob defaults (contracts in tune.hpp).
#include "tune.hpp"

#include <fstream>
#include <map>
#include <mutex>
#include <set>
#include <sstream>
#include <stdexcept>

#include "env.hpp"

namespace gb {
namespace {

std::mutex g_mu;
// process lifetime: the returned value pointers stay valid
std::map<std::string, const std::string*>* g_tuned = new std::map<std::string, const std::string*>();
std::set<std::string>* g_values = new std::set<std::string>();

constexpr const char* kKnobs[] = {"GB_WIN_KB", "GB_WIN_P", "GB_WIN_P_LONG", "GB_KA", "GB_BATCH", "GB_HL_MODE",
                                  "GB_OVERLAP", "GB_HL_WHEEL_HITS", "GB_GRAPH"};

bool known(const std::string& k)
{
    for (const char* n : kKnobs)
        if (k == n) return true;
    return false;
}

uint64_t parse_bound(const std::string& t, const std::string& where)
{
    if (t.empty() || t.find_first_not_of("0123456789") != std::string::npos)
        throw std::runtime_error("tune file: " + where + ": band bound '" + t + "' is not a number");
    return env::parse_u64("band bound", t.c_str());
}

}  // namespace

const char* env::tuned(const char* name)
{
    std::lock_guard<std::mutex> lk(g_mu);
    const auto it = g_tuned->find(name);
    return it == g_tuned->end() ? nullptr : it->second->c_str();
}

std::string tune_load(const std::string& path, uint64_t limit)
{
    std::ifstream f(path);
    if (!f) throw std::runtime_error("tune file: cannot read " + path);
    std::map<std::string, std::string> pick;
    bool found = false;
    std::string line;
    for (int ln = 1; std::getline(f, line); ++ln) {
        const size_t h = line.find('#');
        if (h != std::string::npos) line.resize(h);
        std::istringstream is(line);
        std::string word;
        if (!(is >> word)) continue;
        const std::string where = path + ":" + std::to_string(ln);
        if (word != "band") throw std::runtime_error("tune file: " + where + ": expected 'band'");
        std::string lo_s, hi_s;
        if (!(is >> lo_s >> hi_s)) throw std::runtime_error("tune file: " + where + ": band needs two bounds");
        const uint64_t lo = parse_bound(lo_s, where), hi = parse_bound(hi_s, where);
        if (hi <= lo) throw std::runtime_error("tune file: " + where + ": empty band");
        std::map<std::string, std::string> kv;
        while (is >> word) {
            const size_t eq = word.find('=');
            if (eq == std::string::npos || eq == 0 || eq + 1 == word.size())
                throw std::runtime_error("tune file: " + where + ": '" + word + "' is not KNOB=value");
            const std::string k = word.substr(0, eq);
            if (!known(k)) throw std::runtime_error("tune file: " + where + ": unknown knob " + k);
            kv[k] = word.substr(eq + 1);
        }
        if (!found && limit >= lo && limit < hi) {
            pick = kv;
            found = true;
        }
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_tuned->clear();
    std::string out;
    for (const auto& [k, v] : pick) {
        (*g_tuned)[k] = &*g_values->insert(v).first;
        out += (out.empty() ? "" : " ") + k + "=" + v;
    }
    return out;
}

void tune_clear()
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_tuned->clear();
}

}  // namespace gb
