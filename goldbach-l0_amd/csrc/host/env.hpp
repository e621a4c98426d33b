#pragma once
// Strict GB_* environment knobs (the reference's include/env.hpp:88-170
// contract): a set knob must parse completely, or the run fails before any
// GPU work; nothing is silently ignored.
#include <cstdint>
#include <cstdlib>
#include <initializer_list>
#include <stdexcept>
#include <string>

namespace gb::env {

/// A tuned default of a pipeline knob (tune.cpp: the GB_TUNE_FILE band of
/// the engine's range), or nullptr. The environment always wins.
const char* tuned(const char* name);

// the knob's value: the environment's, else the tuned default
inline const char* value(const char* name)
{
    const char* v = std::getenv(name);
    return v ? v : tuned(name);
}

inline const char* raw(const char* name) { return value(name); }
inline bool present(const char* name) { return value(name) != nullptr; }
inline bool is_set(const char* name)
{
    const char* v = value(name);
    return v && *v;
}

inline uint64_t parse_u64(const char* name, const char* v)
{
    if (!v || !*v) throw std::runtime_error(std::string(name) + " is empty");
    uint64_t x = 0;
    for (const char* p = v; *p; ++p) {
        if (*p < '0' || *p > '9') throw std::runtime_error(std::string(name) + "=" + v + " is not a number");
        const uint64_t d = uint64_t(*p - '0');
        if (x > (UINT64_MAX - d) / 10) throw std::runtime_error(std::string(name) + "=" + v + " overflows");
        x = x * 10 + d;
    }
    return x;
}

inline uint64_t u64(const char* name, uint64_t dflt, uint64_t lo = 0, uint64_t hi = UINT64_MAX)
{
    const char* v = value(name);
    if (!v) return dflt;
    const uint64_t x = parse_u64(name, v);
    if (x < lo || x > hi)
        throw std::runtime_error(std::string(name) + "=" + v + " outside [" + std::to_string(lo) + ", " +
                                 std::to_string(hi) + "]");
    return x;
}

inline uint32_t u32(const char* name, uint32_t dflt, uint32_t lo = 0, uint32_t hi = UINT32_MAX)
{
    return uint32_t(u64(name, dflt, lo, hi));
}

inline bool onoff(const char* name, bool dflt)
{
    const char* v = value(name);
    if (!v) return dflt;
    const std::string s = v;
    if (s == "1" || s == "on" || s == "yes" || s == "true") return true;
    if (s == "0" || s == "off" || s == "no" || s == "false") return false;
    throw std::runtime_error(std::string(name) + "=" + v + " is not on/off");
}

inline std::string choice(const char* name, const char* dflt, std::initializer_list<const char*> ok)
{
    const char* v = value(name);
    if (!v) return dflt;
    for (const char* o : ok)
        if (std::string(v) == o) return v;
    throw std::runtime_error(std::string(name) + "=" + v + " is not a valid choice");
}

inline std::string path(const char* name, const char* dflt)
{
    const char* v = std::getenv(name);
    return (v && *v) ? v : dflt;
}

}  // namespace gb::env
