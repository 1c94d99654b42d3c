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
#include <cstdint>

#ifndef T3D_HD
#if defined(__HIPCC__)
#define T3D_HD __host__ __device__
#else
#define T3D_HD
#endif
#endif

namespace t3d {
namespace seqsum {

constexpr int64_t M_TOP = int64_t(1) << 24;
constexpr int64_t R_CAP = int64_t(1) << 30;   // any increment >= 2^24 already crosses

// increment (in ulps of the binade whose ulp is 2^ulpExp) for parities 0 and 1
struct Tr {
    int64_t r0, r1;
};

T3D_HD inline Tr tr_identity() { return Tr{0, 0}; }

// v / 2^ulpExp rounded half-to-even against the running multiple's parity
T3D_HD inline Tr tr_elem(float v, int ulpExp)
{
    // q = v * 2^-ulpExp is exact in double (power-of-two scaling of a float)
    const double q = double(v) * ldexp(1.0, -ulpExp);
    if (!(q < double(R_CAP))) return Tr{R_CAP, R_CAP};
    const double fl = floor(q);
    const double fr = q - fl;
    const int64_t f = int64_t(fl);
    if (fr > 0.5) return Tr{f + 1, f + 1};
    if (fr < 0.5) return Tr{f, f};
    // tie: the result m + f + {0,1} must be even
    return Tr{f + (f & 1), f + ((f + 1) & 1)};
}

// A followed by B
T3D_HD inline Tr tr_compose(const Tr& A, const Tr& B)
{
    const int64_t a0 = A.r0, a1 = A.r1;
    const int64_t b_after0 = (a0 & 1) ? B.r1 : B.r0;          // parity after A from 0 is a0&1
    const int64_t b_after1 = ((1 + a1) & 1) ? B.r1 : B.r0;    // from 1 it is (1+a1)&1
    int64_t c0 = a0 + b_after0, c1 = a1 + b_after1;
    if (c0 > R_CAP) c0 = R_CAP;
    if (c1 > R_CAP) c1 = R_CAP;
    return Tr{c0, c1};
}

T3D_HD inline int64_t tr_apply(const Tr& T, int64_t m) { return (m & 1) ? T.r1 : T.r0; }

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
