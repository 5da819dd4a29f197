// .npy v1/v2 reader and writer for the kept host API.  Same entry points as
// the reference's bundled reader (common/npyutil.hpp:162-188 save, 238-322
// load): aoba::SaveArrayAsNumpy / aoba::LoadArrayFromNumpy, little-endian,
// C order.  Malformed files, dtype or order mismatches throw
// std::runtime_error, as the reference does (npyutil.hpp:253-313).
#pragma once

#include <complex>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace aoba {

template <typename T> struct NpyDescr;
template <> struct NpyDescr<float> { static const char* str() { return "<f4"; } };
template <> struct NpyDescr<double> { static const char* str() { return "<f8"; } };
template <> struct NpyDescr<std::complex<float>> { static const char* str() { return "<c8"; } };
template <> struct NpyDescr<std::complex<double>> { static const char* str() { return "<c16"; } };
template <> struct NpyDescr<int> { static const char* str() { return "<i4"; } };
template <> struct NpyDescr<long long> { static const char* str() { return "<i8"; } };

template <typename T>
inline void SaveArrayAsNumpy(const std::string& filename, int n_dims, const int shape[], const T* data) {
    std::ostringstream dict;
    dict << "{'descr': '" << NpyDescr<T>::str() << "', 'fortran_order': False, 'shape': (";
    size_t count = 1;
    for (int i = 0; i < n_dims; ++i) {
        dict << shape[i] << (n_dims == 1 || i + 1 < n_dims ? ", " : "");
        count *= (size_t)shape[i];
    }
    dict << "), }";
    std::string h = dict.str();
    const size_t pre = 10;  // magic(6) + version(2) + header length(2)
    size_t total = pre + h.size() + 1;
    h.append((64 - total % 64) % 64, ' ');
    h.push_back('\n');
    std::ofstream f(filename, std::ios::binary);
    if (!f) throw std::runtime_error("npy: cannot open " + filename + " for writing");
    const char magic[8] = {'\x93', 'N', 'U', 'M', 'P', 'Y', 1, 0};
    f.write(magic, 8);
    const uint16_t hl = (uint16_t)h.size();
    const char hlb[2] = {(char)(hl & 0xff), (char)(hl >> 8)};
    f.write(hlb, 2);
    f.write(h.data(), (std::streamsize)h.size());
    f.write(reinterpret_cast<const char*>(data), (std::streamsize)(count * sizeof(T)));
    if (!f) throw std::runtime_error("npy: write failed for " + filename);
}

inline std::string npy_field(const std::string& h, const std::string& key) {
    const size_t k = h.find("'" + key + "'");
    if (k == std::string::npos) throw std::runtime_error("npy: header lacks " + key);
    const size_t c = h.find(':', k);
    if (c == std::string::npos) throw std::runtime_error("npy: malformed header");
    return h.substr(c + 1);
}

template <typename T>
inline void LoadArrayFromNumpy(const std::string& filename, std::vector<int>& shape, std::vector<T>& data) {
    std::ifstream f(filename, std::ios::binary);
    if (!f) throw std::runtime_error("npy: cannot open " + filename);
    char magic[8];
    f.read(magic, 8);
    if (!f || std::memcmp(magic, "\x93NUMPY", 6) != 0) throw std::runtime_error("npy: bad magic in " + filename);
    uint32_t hl = 0;
    if (magic[6] == 1) {
        unsigned char b[2];
        f.read(reinterpret_cast<char*>(b), 2);
        hl = b[0] | (b[1] << 8);
    } else if (magic[6] == 2 || magic[6] == 3) {
        unsigned char b[4];
        f.read(reinterpret_cast<char*>(b), 4);
        hl = b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24);
    } else {
        throw std::runtime_error("npy: unsupported version");
    }
    std::string h(hl, ' ');
    f.read(&h[0], hl);
    if (!f) throw std::runtime_error("npy: truncated header");
    std::string d = npy_field(h, "descr");
    const size_t q0 = d.find('\''), q1 = d.find('\'', q0 + 1);
    d = d.substr(q0 + 1, q1 - q0 - 1);
    if (d != NpyDescr<T>::str() && !(d[0] == '|' && d.substr(1) == std::string(NpyDescr<T>::str()).substr(1)))
        throw std::runtime_error("npy: dtype " + d + " does not match the requested type");
    const std::string fo = npy_field(h, "fortran_order");
    if (fo.find("True") < fo.find(',')) throw std::runtime_error("npy: Fortran order is not supported");
    std::string s = npy_field(h, "shape");
    s = s.substr(s.find('(') + 1, s.find(')') - s.find('(') - 1);
    shape.clear();
    size_t count = 1;
    std::stringstream ss(s);
    std::string tok;
    while (std::getline(ss, tok, ',')) {
        if (tok.find_first_not_of(" ") == std::string::npos) continue;
        const int v = std::stoi(tok);
        shape.push_back(v);
        count *= (size_t)v;
    }
    data.resize(count);
    f.read(reinterpret_cast<char*>(data.data()), (std::streamsize)(count * sizeof(T)));
    if (!f) throw std::runtime_error("npy: truncated data in " + filename);
}

}  // namespace aoba
