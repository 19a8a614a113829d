// sigproc.h -- writers of SIGPROC header items, shared by the filterbank
// writer (sigproc.cpp) and the --savetim dumps (dump.cpp).  A string item is
// its 32-bit length then its bytes; a keyword is followed by its value.
#pragma once
#include <cstdint>
#include <string>

namespace txgpu {
namespace sig {

inline void put_str(std::string &h, const std::string &s)
{
    const int32_t n = (int32_t)s.size();
    h.append((const char *)&n, 4);
    h += s;
}
inline void put_int(std::string &h, const char *key, int32_t v)
{
    put_str(h, key);
    h.append((const char *)&v, 4);
}
inline void put_double(std::string &h, const char *key, double v)
{
    put_str(h, key);
    h.append((const char *)&v, 8);
}
inline void put_byte(std::string &h, const char *key, int8_t v)
{
    put_str(h, key);
    h.append((const char *)&v, 1);
}

} // namespace sig
} // namespace txgpu
