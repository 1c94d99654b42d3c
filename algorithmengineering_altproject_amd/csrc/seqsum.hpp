// seqsum.hpp -- exact emulation of a sequential IEEE binary32 running sum
//     s_0 = 0, s_i = fl32(s_{i-1} + v_i)      (v_i >= 0)
// as computed by std::accumulate<float> in the reference's OC update
// (cpp/src/driver.cpp:181, 188-189, 201), evaluated in parallel.
//
// While the running sum stays inside one binade [2^(e-1), 2^e), every partial
// sum is an integer multiple m of u = ulp = 2^(e-24), 2^23 <= m < 2^24, and one
// step adds round(v/u) to m with round-half-to-even deciding ties on the
// parity of m.  So each addend is a two-state transducer: for incoming parity
// p it adds r_p in units of u and flips the parity by r_p.  Transducers compose
// associatively, which allows a block-wide scan; the first step whose result
// would leave the binade (m + r >= 2^24) is recomputed with a real float add
// and the scan restarts after it in the new binade.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

#ifndef T3D_HD
#if defined(__HIPCC__)
#define T3D_HD __host__ __device__
#else
#define T3D_HD
#endif
#endif

namespace t3d {
namespace seqsum {

constexpr uint32_t M_TOP = 1u << 24;
constexpr uint32_t R_CAP = 1u << 30;   // any increment >= 2^24 already crosses

// increment (in ulps of the binade whose ulp is 2^ulpExp) for parities 0 and 1;
// both are at most R_CAP, so a sum of two fits 32 bits
struct Tr {
    uint32_t r0, r1;
};

T3D_HD inline Tr tr_identity() { return Tr{0u, 0u}; }

// p ? x1 : x0 for p in {0, 1}, as a mask (a conditional on a struct member
// becomes an indexed scratch load on the GPU)
T3D_HD inline uint32_t sel01(uint32_t p, uint32_t x1, uint32_t x0)
{
    const uint32_t mk = 0u - p;
    return (x0 & ~mk) | (x1 & mk);
}

// v / 2^ulpExp (v >= 0) rounded half-to-even against the running multiple's
// parity, from the bits of v: v = mant 2^ev, so the quotient is mant shifted
// right by ulpExp - ev places, with the shifted-out bits deciding the rounding
T3D_HD inline Tr tr_elem(float v, int ulpExp)
{
    uint32_t b;
    memcpy(&b, &v, sizeof(b));
    b &= 0x7fffffffu;
    if (b == 0u) return Tr{0u, 0u};
    const int E = int(b >> 23);
    if (E == 255) return Tr{R_CAP, R_CAP};   // inf, NaN
    const uint32_t mant = E ? ((b & 0x7fffffu) | 0x800000u) : b;   // < 2^24
    const int sh = ulpExp - ((E ? E : 1) - 150);
    if (sh <= 0) {   // an integer multiple of the ulp
        if (sh <= -30) return Tr{R_CAP, R_CAP};
        const uint64_t q = uint64_t(mant) << (-sh);
        const uint32_t f = q >= R_CAP ? R_CAP : uint32_t(q);
        return Tr{f, f};
    }
    if (sh >= 25) return Tr{0u, 0u};   // below a quarter ulp
    const uint32_t f = mant >> sh;
    const uint32_t rem = mant & ((1u << sh) - 1u), half = 1u << (sh - 1);
    if (rem > half) return Tr{f + 1u, f + 1u};
    if (rem < half) return Tr{f, f};
    // tie: the result m + f + {0,1} must be even
    return Tr{f + (f & 1u), f + ((f + 1u) & 1u)};
}

// A followed by B
T3D_HD inline Tr tr_compose(const Tr& A, const Tr& B)
{
    // parity after A: from 0 it is A.r0 & 1, from 1 it is (1 + A.r1) & 1
    uint32_t c0 = A.r0 + sel01(A.r0 & 1u, B.r1, B.r0);
    uint32_t c1 = A.r1 + sel01((A.r1 + 1u) & 1u, B.r1, B.r0);
    if (c0 > R_CAP) c0 = R_CAP;
    if (c1 > R_CAP) c1 = R_CAP;
    return Tr{c0, c1};
}

T3D_HD inline int64_t tr_apply(const Tr& T, int64_t m) { return int64_t(sel01(uint32_t(m & 1), T.r1, T.r0)); }

// s = m * 2^ulpExp with 2^23 <= m < 2^24 for a normal positive float s
T3D_HD inline void decompose(float s, int64_t& m, int& ulpExp)
{
    int e = 0;
    const float fr = frexpf(s, &e);   // s = fr * 2^e, fr in [0.5, 1)
    m = int64_t(ldexpf(fr, 24));
    ulpExp = e - 24;
}

}  // namespace seqsum
}  // namespace t3d
