// Text and file helpers (contracts in textio.hpp).
#include "textio.hpp"

#include <fcntl.h>
#include <sys/file.h>
#include <unistd.h>

#include <cerrno>
#include <charconv>
#include <cstring>
#include <stdexcept>

namespace gb::textio {

std::optional<std::string> read_file(const std::string& path)
{
    const int fd = ::open(path.c_str(), O_RDONLY | O_CLOEXEC);
    if (fd < 0) return std::nullopt;
    std::string out;
    char buf[1 << 15];
    for (;;) {
        const ssize_t r = ::read(fd, buf, sizeof buf);
        if (r > 0) {
            out.append(buf, size_t(r));
        } else if (r == 0) {
            break;
        } else if (errno != EINTR) {
            ::close(fd);
            throw std::runtime_error("cannot read " + path + ": " + std::strerror(errno));
        }
    }
    ::close(fd);
    return out;
}

namespace {

// writes all of `bytes` to fd, retrying short writes
bool write_all(int fd, std::string_view bytes)
{
    const char* p = bytes.data();
    size_t left = bytes.size();
    while (left) {
        const ssize_t w = ::write(fd, p, left);
        if (w < 0) {
            if (errno == EINTR) continue;
            return false;
        }
        p += w;
        left -= size_t(w);
    }
    return true;
}

std::string parent_dir(const std::string& path)
{
    const size_t slash = path.rfind('/');
    if (slash == std::string::npos) return ".";
    return slash == 0 ? "/" : path.substr(0, slash);
}

}  // namespace

void replace_file(const std::string& path, std::string_view bytes, const std::string& tmp_suffix, bool durable)
{
    const std::string tmp = path + tmp_suffix;
    const int fd = ::open(tmp.c_str(), O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
    if (fd < 0) throw std::runtime_error("cannot create " + tmp + ": " + std::strerror(errno));
    const bool written = write_all(fd, bytes) && (!durable || ::fsync(fd) == 0);
    const bool closed = ::close(fd) == 0;
    if (!written || !closed || ::rename(tmp.c_str(), path.c_str()) != 0) {
        ::unlink(tmp.c_str());
        throw std::runtime_error("cannot write " + path + " (the previous file is unchanged)");
    }
    if (!durable) return;
    const int dfd = ::open(parent_dir(path).c_str(), O_RDONLY | O_DIRECTORY | O_CLOEXEC);
    if (dfd >= 0) {
        (void)::fsync(dfd);
        ::close(dfd);
    }
}

bool append_file(const std::string& path, std::string_view bytes)
{
    const int fd = ::open(path.c_str(), O_WRONLY | O_CREAT | O_APPEND | O_CLOEXEC, 0644);
    if (fd < 0) return false;
    const bool ok = write_all(fd, bytes) && ::fsync(fd) == 0;
    return ::close(fd) == 0 && ok;
}

FileLock::FileLock(const std::string& path)
{
    fd_ = ::open(path.c_str(), O_RDWR | O_CREAT | O_CLOEXEC, 0644);
    if (fd_ < 0) throw std::runtime_error("cannot open the lock file " + path);
    while (::flock(fd_, LOCK_EX) != 0) {
        if (errno == EINTR) continue;
        ::close(fd_);
        fd_ = -1;
        throw std::runtime_error("cannot lock " + path);
    }
}

FileLock::~FileLock()
{
    if (fd_ >= 0) ::close(fd_);   // closing releases the lock
}

std::vector<std::string_view> words(std::string_view line)
{
    std::vector<std::string_view> out;
    size_t from = 0;
    for (size_t i = 0; i <= line.size(); ++i)
        if (i == line.size() || line[i] == ' ') {
            out.push_back(line.substr(from, i - from));
            from = i + 1;
        }
    return out;
}

bool parse(std::string_view s, u128& v)
{
    if (s.empty() || s.size() > 39) return false;
    u128 acc = 0;
    for (const char ch : s) {
        if (ch < '0' || ch > '9') return false;
        const u128 next = acc * 10 + unsigned(ch - '0');
        if (next / 10 != acc) return false;   // wrapped
        acc = next;
    }
    v = acc;
    return true;
}

bool parse(std::string_view s, uint64_t& v)
{
    u128 w = 0;
    if (!parse(s, w) || (w >> 64) != 0) return false;
    v = uint64_t(w);
    return true;
}

bool parse(std::string_view s, uint32_t& v)
{
    u128 w = 0;
    if (!parse(s, w) || w > 0xFFFFFFFFu) return false;
    v = uint32_t(w);
    return true;
}

std::string dec(u128 v)
{
    char digits[40];
    size_t n = sizeof digits;
    do {
        digits[--n] = char('0' + unsigned(v % 10));
        v /= 10;
    } while (v);
    return std::string(digits + n, sizeof digits - n);
}

void append_dec(std::string& s, uint64_t v)
{
    char buf[20];
    const auto r = std::to_chars(buf, buf + sizeof buf, v);
    s.append(buf, size_t(r.ptr - buf));
}

void append_dec(std::string& s, u128 v)
{
    if (v <= UINT64_MAX) return append_dec(s, uint64_t(v));
    s += dec(v);
}

void append_count_list(std::string& s, const uint64_t* v, size_t n)
{
    while (n && !v[n - 1]) --n;
    if (!n) return;
    // at most 20 digits and a comma per entry: written in place, then cut
    const size_t at = s.size();
    s.resize(at + 21 * n);
    char* const out = s.data() + at;
    char* p = out;
    for (size_t i = 0; i < n; ++i) {
        if (i) *p++ = ',';
        if (v[i]) p = std::to_chars(p, p + 20, v[i]).ptr;
    }
    s.resize(at + size_t(p - out));
}

std::string count_list(const uint64_t* v, size_t n)
{
    std::string s;
    append_count_list(s, v, n);
    return s;
}

bool parse_count_list(std::string_view s, uint64_t* out, size_t n)
{
    for (size_t i = 0; i < n; ++i) out[i] = 0;
    if (s.empty()) return true;
    size_t idx = 0, from = 0;
    for (size_t i = 0; i <= s.size(); ++i) {
        if (i < s.size() && s[i] != ',') continue;
        if (idx >= n) return false;
        const std::string_view item = s.substr(from, i - from);
        if (!item.empty() && !parse(item, out[idx])) return false;
        ++idx;
        from = i + 1;
    }
    return true;
}

}  // namespace gb::textio
