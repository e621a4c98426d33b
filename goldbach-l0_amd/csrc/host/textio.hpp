#pragma once
// Text and file helpers of the campaign files (statistics state, gap state,
// run log, records, ridge store, prime cache): whole-file reads, crash-safe
// replacement, synced appends, an advisory lock, and the decimal and
// count-list codecs the line formats use.
#include <cstdint>
#include <optional>
#include <string>
#include <string_view>
#include <vector>

namespace gb::textio {

using u128 = unsigned __int128;

/// The file's bytes; nullopt when it cannot be opened (absent).
std::optional<std::string> read_file(const std::string& path);

/// Replaces @p path by @p bytes so that a crash leaves either the old or
/// the new file: a sibling temp file (path + tmp_suffix) written and (when
/// @p durable) fsynced, renamed over path, the directory fsynced.
/// @throws std::runtime_error (the old file is kept) when any step fails.
void replace_file(const std::string& path, std::string_view bytes, const std::string& tmp_suffix = ".tmp",
                  bool durable = true);

/// Appends @p bytes to @p path and fsyncs it; false when any step failed.
bool append_file(const std::string& path, std::string_view bytes);

/// Exclusive advisory lock (flock) on a lock file, held for the object's
/// lifetime. @throws std::runtime_error when the file cannot be locked.
class FileLock {
public:
    explicit FileLock(const std::string& path);
    ~FileLock();
    FileLock(const FileLock&) = delete;
    FileLock& operator=(const FileLock&) = delete;

private:
    int fd_ = -1;
};

/// The words of a line split at single spaces (consecutive spaces give
/// empty words, as the formats write an empty list as nothing).
std::vector<std::string_view> words(std::string_view line);

/// Strict decimal parsers: digits only, no sign, no overflow.
bool parse(std::string_view s, u128& v);
bool parse(std::string_view s, uint64_t& v);
bool parse(std::string_view s, uint32_t& v);

/// Decimal text of a 128-bit value.
std::string dec(u128 v);

/// Counts as a comma list by index: zero entries left empty, trailing zeros
/// dropped ("", "5", "1,,3").
std::string count_list(const uint64_t* v, size_t n);
inline std::string count_list(const std::vector<uint64_t>& v) { return count_list(v.data(), v.size()); }

/// In-place forms for text built piece by piece (no temporaries): the
/// decimal of @p v, and count_list of v[0, n), appended to @p s.
void append_dec(std::string& s, uint64_t v);
void append_dec(std::string& s, u128 v);
void append_count_list(std::string& s, const uint64_t* v, size_t n);

/// The inverse of count_list into @p n entries; false on a malformed entry
/// or more than n of them.
bool parse_count_list(std::string_view s, uint64_t* out, size_t n);

}  // namespace gb::textio
