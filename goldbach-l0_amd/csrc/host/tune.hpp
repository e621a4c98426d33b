#pragma once
// Tuned knob defaults (the autotuner's output, goldbach-l0_amd/autotune.py).
//
// A tune file holds one band per line:
//     band <lo> <hi> KNOB=value ...
// for ranges whose largest limit lies in [lo, hi); '#' starts a comment.
// The knobs are the pipeline's shape choices (GB_WIN_KB, GB_WIN_P,
// GB_WIN_P_LONG, GB_KA, GB_BATCH, GB_HL_MODE, GB_OVERLAP, GB_HL_WHEEL_HITS);
// a value set in the
// environment always wins over the file, and the values pass the knobs' own strict checks when
// read. An engine loads GB_TUNE_FILE, when set, for its limit.
#include <cstdint>
#include <string>

namespace gb {

/// Installs the knobs of the first band of @p path that holds @p limit as
/// the tuned defaults (replacing any earlier ones; none when no band holds
/// it). @return the installed knobs as "KNOB=value ..." ("" for none).
/// @throws std::runtime_error on an unreadable file, a malformed line or an
///         unknown knob.
std::string tune_load(const std::string& path, uint64_t limit);

/// Drops the tuned defaults.
void tune_clear();

}  // namespace gb
