// Layout helpers of the kept host API (common/technicalities.hpp:3-32).
#pragma once

#include <hip/hip_runtime.h>

#include <ostream>
#include <stdexcept>

inline std::ostream& operator<<(std::ostream& o, const dim3& v) { return o << v.x << ' ' << v.y << ' ' << v.z; }
inline std::ostream& operator<<(std::ostream& o, const float3& v) { return o << v.x << ' ' << v.y << ' ' << v.z; }
inline std::ostream& operator<<(std::ostream& o, const double3& v) { return o << v.x << ' ' << v.y << ' ' << v.z; }
inline std::ostream& operator<<(std::ostream& o, const int3& v) { return o << v.x << ' ' << v.y << ' ' << v.z; }
inline std::ostream& operator<<(std::ostream& o, const float2& v) { return o << v.x << ' ' << v.y; }
inline std::ostream& operator<<(std::ostream& o, const double2& v) { return o << v.x << ' ' << v.y; }
inline std::ostream& operator<<(std::ostream& o, const int2& v) { return o << v.x << ' ' << v.y; }

inline bool operator==(const int3& a, const int3& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
inline bool operator==(const uint3& a, const uint3& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
inline bool operator==(const int2& a, const int2& b) { return a.x == b.x && a.y == b.y; }
inline bool operator==(const uint2& a, const uint2& b) { return a.x == b.x && a.y == b.y; }

// Reverses the axis order of the used dimensions: (x, y, z) extents as
// numpy's (z, y, x) shape, for npy files and row-major plans.
inline int3 shape_tr(int3 in) {
    const int dims = in.z != 1 ? 3 : (in.y != 1 ? 2 : 1);
    int3 out{1, 1, 1};
    if (dims == 1) out.x = in.x;
    else if (dims == 2) { out.x = in.y; out.y = in.x; }
    else { out.x = in.z; out.y = in.y; out.z = in.x; }
    return out;
}
