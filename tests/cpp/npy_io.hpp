// npy_io.hpp -- minimal .npy reader / writer (little-endian, C order, float64 / int32 / int64)
// for the C++ host-mirror test driver.
#pragma once
#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace npy
{
template <class T>
const char *descr();
template <>
inline const char *descr<double>() { return "<f8"; }
template <>
inline const char *descr<int32_t>() { return "<i4"; }
template <>
inline const char *descr<int64_t>() { return "<i8"; }

template <class T>
std::vector<T> load(const std::string &path, std::vector<size_t> *shape = nullptr)
{
    std::ifstream f(path, std::ios::binary);
    if (!f)
        throw std::runtime_error("cannot open " + path);
    char magic[6];
    f.read(magic, 6);
    if (std::memcmp(magic, "\x93NUMPY", 6) != 0)
        throw std::runtime_error("not a .npy file: " + path);
    unsigned char ver[2];
    f.read(reinterpret_cast<char *>(ver), 2);
    uint32_t hlen = 0;
    if (ver[0] == 1)
    {
        uint16_t h16;
        f.read(reinterpret_cast<char *>(&h16), 2);
        hlen = h16;
    }
    else
        f.read(reinterpret_cast<char *>(&hlen), 4);
    std::string hdr(hlen, ' ');
    f.read(&hdr[0], hlen);
    if (hdr.find(descr<T>()) == std::string::npos || hdr.find("'fortran_order': False") == std::string::npos)
        throw std::runtime_error("unexpected dtype or order in " + path + ": " + hdr);
    const size_t a = hdr.find('(') + 1, b = hdr.find(')');
    std::vector<size_t> shp;
    size_t n = 1;
    std::string dims = hdr.substr(a, b - a);
    size_t pos = 0;
    while (pos < dims.size())
    {
        size_t q = dims.find(',', pos);
        std::string tok = dims.substr(pos, q == std::string::npos ? std::string::npos : q - pos);
        if (tok.find_first_not_of(" ") != std::string::npos)
        {
            shp.push_back(std::stoull(tok));
            n *= shp.back();
        }
        if (q == std::string::npos)
            break;
        pos = q + 1;
    }
    std::vector<T> v(n);
    f.read(reinterpret_cast<char *>(v.data()), n * sizeof(T));
    if (shape)
        *shape = shp;
    return v;
}

template <class T>
void save(const std::string &path, const std::vector<T> &v, const std::vector<size_t> &shape)
{
    std::string dims;
    for (size_t s : shape)
        dims += std::to_string(s) + ",";
    std::string hdr = std::string("{'descr': '") + descr<T>() + "', 'fortran_order': False, 'shape': (" + dims + "), }";
    while ((10 + hdr.size() + 1) % 64 != 0)
        hdr += ' ';
    hdr += '\n';
    std::ofstream f(path, std::ios::binary);
    f.write("\x93NUMPY\x01\x00", 8);
    const uint16_t hl = (uint16_t)hdr.size();
    f.write(reinterpret_cast<const char *>(&hl), 2);
    f.write(hdr.data(), hdr.size());
    f.write(reinterpret_cast<const char *>(v.data()), v.size() * sizeof(T));
}
} // namespace npy
