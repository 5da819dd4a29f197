// Kept host API: vector, lattice and FFT-size helpers (cuda/math.hpp:1-153).
// Same names and results; FFT sizes follow the 2,3,5-smooth policy that the
// library's plans accept.
#pragma once

#include <hip/hip_runtime.h>

#include <cassert>
#include <cmath>
#include <utility>

#ifndef PI
#define PI M_PI
#endif
#ifndef PIX2
#define PIX2 (2.0 * M_PI)
#endif

inline dim3 i2d3(const int v) { return dim3(v, 1, 1); }
inline dim3 i2d3(const int2 v) { return dim3(v.x, v.y, 1); }
inline dim3 i2d3(const int3 v) { return dim3(v.x, v.y, v.z); }

inline int norm2(int3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
inline double norm2(double3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
inline double norm2(double2 v) { return v.x * v.x + v.y * v.y; }
inline double dist2(double2 a, double2 b) { return norm2(double2{a.x - b.x, a.y - b.y}); }
inline double norm(double3 v) { return std::sqrt(norm2(v)); }
inline double norm(double2 v) { return std::sqrt(norm2(v)); }
inline double norm(int3 v) { return std::sqrt((double)norm2(v)); }
inline double3 normalized(int3 v) { const double n = norm(v); return {v.x / n, v.y / n, v.z / n}; }
inline double3 normalized(double3 v) { const double n = norm(v); return {v.x / n, v.y / n, v.z / n}; }
inline double2 normalized(double2 v) { const double n = norm(v); return {v.x / n, v.y / n}; }
inline double inner(double3 a, double3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double inner(double2 a, double2 b) { return a.x * b.x + a.y * b.y; }
inline double inner(int3 a, double3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline int3 cross_int3(int3 p, int3 q) {
    return {p.y * q.z - p.z * q.y, p.z * q.x - p.x * q.z, p.x * q.y - p.y * q.x};
}

// largest singular value of the 2x2 matrix {m0 m1; m2 m3} (the root of the
// larger eigenvalue of its column Gram matrix)
inline double condition_2x2(double* m) {
    const double a = m[0] * m[0] + m[2] * m[2];
    const double b = m[0] * m[1] + m[2] * m[3];
    const double d = m[1] * m[1] + m[3] * m[3];
    const double disc = std::sqrt((a - d) * (a - d) + 4.0 * b * b);
    const double lo = 0.5 * (a + d - disc), hi = 0.5 * (a + d + disc);
    assert(lo > 0.0 && "eigenvalue must be positive");
    assert(hi > 0.0 && "eigenvalue must be positive");
    return std::sqrt(hi > lo ? hi : lo);
}

// conditioning of the in-plane basis (pv, qv) expressed in the orthonormal
// frame {pv/|pv|, pv x (pv x qv)/|..|}
inline double crystal_condition(int3 pv, int3 qv) {
    const double3 e1 = normalized(pv);
    const double3 e2 = normalized(cross_int3(pv, cross_int3(pv, qv)));
    double m[4] = {norm(pv), inner(qv, e1), 0.0, inner(qv, e2)};
    return condition_2x2(m);
}

// nearest of the four lattice points (n1,n2) + {0,1}^2 of the triangular
// lattice {a, b} to the point p; ties go to the earlier candidate
inline std::pair<int2, double> calc_closest_lattice_indices(int2 n, double2 a, double2 b, double2 p) {
    int2 best = n;
    double bd = 0;
    const int2 cand[4] = {{n.x, n.y}, {n.x + 1, n.y}, {n.x, n.y + 1}, {n.x + 1, n.y + 1}};
    double d[4];
    for (int c = 0; c < 4; ++c)
        d[c] = dist2(double2{cand[c].x * a.x + cand[c].y * b.x, cand[c].x * a.y + cand[c].y * b.y}, p);
    if (d[0] < d[1] && d[0] < d[2] && d[0] < d[3]) { best = cand[0]; bd = d[0]; }
    else if (d[1] < d[2] && d[1] < d[3]) { best = cand[1]; bd = d[1]; }
    else if (d[2] < d[3]) { best = cand[2]; bd = d[2]; }
    else { best = cand[3]; bd = d[3]; }
    return std::make_pair(best, bd);
}

// true iff n has no prime factor above 5 (trial division from i)
inline bool factors235(int n, int i = 2) {
    for (;; ++i) {
        if (i * i > n) return n <= 5;
        if (n % i == 0) return factors235(i) && factors235(n / i);
    }
}

inline int _nextoneup(int n) {
    do { ++n; } while (!factors235(n));
    return n;
}

// factor n by trial division and replace every prime factor above 5 by the
// next 2,3,5-smooth number (so 7 -> 8, 14 -> 16, 49 -> 64, 97 -> 100)
inline int roundup235(int n, int i = 2) {
    for (;; ++i) {
        if (i * i > n) return n <= 5 ? n : _nextoneup(n);
        if (n % i == 0) return roundup235(i) * roundup235(n / i);
    }
}

inline int isgn(int v) { return (v > 0) - (v < 0); }
