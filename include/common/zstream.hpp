// Gzip output stream of the kept API (common/zstream.hpp:148-248).  ozstream
// opens a .gz file for writing; text streamed into it with << is collected in
// memory and compressed into the file when the stream is flushed or closed
// (the destructor closes it).  write() adds raw bytes after any pending text.
// Consumers link with -lz.
#pragma once

#include <zlib.h>

#include <cstddef>
#include <cstdio>
#include <ostream>
#include <sstream>
#include <string>

class ozstream {
    gzFile m_gz = nullptr;
    std::ostringstream m_text;

    int drain() {
        const std::string s = m_text.str();
        m_text.str(std::string());
        if (s.empty()) return 0;
        return m_gz ? ::gzwrite(m_gz, s.data(), (unsigned)s.size()) : 0;
    }

  public:
    ozstream() = default;
    // level 0..9, or Z_DEFAULT_COMPRESSION
    explicit ozstream(const char* name, int level = Z_DEFAULT_COMPRESSION) { open(name, level); }
    ozstream(FILE* fp, int level = Z_DEFAULT_COMPRESSION) { open(fp, level); }
    ~ozstream() { close(); }
    ozstream(const ozstream&) = delete;
    ozstream& operator=(const ozstream&) = delete;

    void open(const char* name, int level = Z_DEFAULT_COMPRESSION) {
        close();
        char mode[4] = {'w', 'b', 0, 0};
        if (level >= 0 && level <= 9) mode[2] = (char)('0' + level);
        m_gz = ::gzopen(name, mode);
    }
    void open(FILE* fp, int level = Z_DEFAULT_COMPRESSION) {
        close();
        char mode[4] = {'w', 'b', 0, 0};
        if (level >= 0 && level <= 9) mode[2] = (char)('0' + level);
        m_gz = ::gzdopen(fileno(fp), mode);
    }
    bool is_open() const { return m_gz != nullptr; }
    // compresses the pending text, then closes the file; returns zlib's status
    int close() {
        if (!m_gz) return Z_OK;
        drain();
        const int r = ::gzclose(m_gz);
        m_gz = nullptr;
        return r;
    }
    int write(const void* buf, size_t len) {
        drain();
        return m_gz ? ::gzwrite(m_gz, buf, (unsigned)len) : 0;
    }
    int flush(int how = Z_SYNC_FLUSH) {
        drain();
        return m_gz ? ::gzflush(m_gz, how) : Z_STREAM_ERROR;
    }
    const char* error(int* errnum) { return m_gz ? ::gzerror(m_gz, errnum) : "not open"; }
    gzFile fp() { return m_gz; }
    std::ostream& os() { return m_text; }
};

// text output; further << go to the same in-memory text
template <class T> inline std::ostream& operator<<(ozstream& zs, const T& x) { return zs.os() << x; }
