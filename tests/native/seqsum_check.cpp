// seqsum_check.cpp -- CPU check of the transducer arithmetic of
// algorithmengineering_altproject_amd/csrc/seqsum.hpp against real binary32
// additions (the host's IEEE float adds, no contraction).  Built and run by
// tests/test_seqsum.py; prints "ok <cases>" or the first mismatch.
//
//  1. tr_elem against the quotient v / 2^ulpExp computed in long double and
//     rounded by hand, for random and edge-case (v, ulpExp) pairs;
//  2. one addition: for a normal s = m 2^ue and v >= 0 whose transducer does
//     not leave the binade, fl(s + v) == (m + r) 2^ue;
//  3. whole sums: a chunked walk (chunk aggregates composed in the running
//     binade, a chunk that leaves it summed element by element with a real
//     float add at the crossing) equals the sequential float sum, on arrays
//     that cross many binades, with zeros, denormals and constructed ties.
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../algorithmengineering_altproject_amd/csrc/seqsum.hpp"

using namespace t3d::seqsum;

namespace {
uint64_t rs = 0x243f6a8885a308d3ull;
uint64_t next()
{
    rs ^= rs >> 12;
    rs ^= rs << 25;
    rs ^= rs >> 27;
    return rs * 2685821657736338717ull;
}
double unif() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
float bits(uint32_t b)
{
    float f;
    memcpy(&f, &b, 4);
    return f;
}

int fails = 0;
long cases = 0;

// independent restatement: q = v / 2^e in long double (exact: 64-bit mantissa)
Tr elem_ref(float v, int e)
{
    const long double q = std::ldexp((long double)v, -e);
    if (!(q < (long double)R_CAP)) return Tr{R_CAP, R_CAP};
    const long double fl = std::floor(q), fr = q - fl;
    const uint32_t f = uint32_t(fl);
    if (fr > 0.5L) return Tr{f + 1, f + 1};
    if (fr < 0.5L) return Tr{f, f};
    return Tr{f + (f & 1u), f + ((f + 1u) & 1u)};
}

void check_elem(float v, int e)
{
    ++cases;
    const Tr a = tr_elem(v, e), b = elem_ref(v, e);
    if (a.r0 != b.r0 || a.r1 != b.r1) {
        if (fails++ < 5)
            std::printf("tr_elem mismatch v=%a e=%d got {%u,%u} want {%u,%u}\n", v, e, a.r0, a.r1, b.r0, b.r1);
    }
}

void check_add(float s, float v)
{
    int64_t m;
    int ue;
    decompose(s, m, ue);
    const int64_t r = tr_apply(tr_elem(v, ue), m);
    if (m + r >= int64_t(M_TOP)) return;   // leaves the binade: the walk adds it for real
    ++cases;
    const float want = s + v;
    const float got = std::ldexp(float(m + r), ue);
    if (got != want) {
        if (fails++ < 5) std::printf("add mismatch s=%a v=%a got %a want %a\n", s, v, got, want);
    }
}

// the chunked walk over v[0..n) after the first element
float walk(const std::vector<float>& v, size_t chunk)
{
    const size_t n = v.size();
    float s = 0.0f;
    size_t i = 0;
    for (; i < n && !(s >= 1e-30f); ++i) s += v[i];
    int64_t m;
    int ue;
    decompose(s, m, ue);
    while (i < n) {
        const size_t end = i + chunk < n ? i + chunk : n;
        Tr A = tr_identity();
        for (size_t k = i; k < end; ++k) A = tr_compose(A, tr_elem(v[k], ue));
        if (m + tr_apply(A, m) < int64_t(M_TOP)) {   // the whole chunk in this binade
            m += tr_apply(A, m);
            i = end;
            continue;
        }
        for (; i < end; ++i) {   // element by element up to the crossing
            const int64_t r = tr_apply(tr_elem(v[i], ue), m);
            if (m + r >= int64_t(M_TOP)) {
                const float t = std::ldexp(float(m), ue) + v[i];
                decompose(t, m, ue);
                ++i;
                break;
            }
            m += r;
        }
    }
    return std::ldexp(float(m), ue);
}

void check_sum(const std::vector<float>& v, size_t chunk)
{
    ++cases;
    float s = 0.0f;
    for (float x : v) s += x;
    bool any = false;
    for (float x : v) any = any || x > 0.0f;
    if (!any) return;
    const float w = walk(v, chunk);
    if (w != s) {
        if (fails++ < 5) std::printf("sum mismatch n=%zu chunk=%zu got %a want %a\n", v.size(), chunk, w, s);
    }
}
}  // namespace

int main()
{
    // 1. tr_elem: edge cases and random pairs
    const float edge[] = {0.0f, bits(1), bits(0x7fffff), bits(0x800000), 1.0f, 1.5f, 2.5f, 3.0f,
                          0.75f, 1e-3f, 0.12f, 16777215.0f, 16777216.0f, 3.4e38f, bits(0x7f800000)};
    for (float v : edge)
        for (int e = -170; e <= 140; ++e) check_elem(v, e);
    for (int t = 0; t < 2000000; ++t) {
        const uint32_t b = uint32_t(next() >> 33) & 0x7f7fffffu;   // finite non-negative
        const float v = bits(b);
        int ev = 0;
        std::frexp(v, &ev);
        const int e = ev - 24 + int(next() % 60) - 30;
        check_elem(v, e);
    }
    // 2. single additions, including exact ties (v = (k + 1/2) ulp)
    for (int t = 0; t < 2000000; ++t) {
        const float s = float(std::ldexp(1.0 + unif(), int(next() % 60) - 30));
        int64_t m;
        int ue;
        decompose(s, m, ue);
        float v;
        switch (t % 4) {
        case 0: v = float(s * unif() * 1e-3); break;
        case 1: v = std::ldexp(float(int(next() % 64)) + 0.5f, ue); break;   // a tie
        case 2: v = std::ldexp(float(next() % 4096), ue - int(next() % 8)); break;
        default: v = float(s * unif()); break;
        }
        check_add(s, v);
    }
    // 3. whole sums
    for (int t = 0; t < 300; ++t) {
        const size_t n = 1 + size_t(next() % 200000);
        std::vector<float> v(n);
        const int kind = t % 5;
        for (size_t i = 0; i < n; ++i) {
            const double u = unif();
            switch (kind) {
            case 0: v[i] = float(1e-3 + u * (1.0 - 1e-3)); break;                 // OC densities
            case 1: { const float x = float(u); v[i] = x * (1.0f - x); break; }    // x (1 - x)
            case 2: v[i] = (next() % 3 == 0) ? 0.0f : float(u); break;             // zeros
            case 3: v[i] = std::ldexp(float(next() % 8) + 0.5f, -20); break;       // many ties
            default: v[i] = (next() % 50 == 0) ? bits(uint32_t(next() % 0x800000)) : float(u * 1e3); break;
            }
        }
        check_sum(v, 1 + size_t(next() % 4096));
    }
    if (fails) {
        std::printf("FAIL %d of %ld\n", fails, cases);
        return 1;
    }
    std::printf("ok %ld\n", cases);
    return 0;
}
