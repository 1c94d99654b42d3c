// topvoxel.cpp -- the TopVoxel voxel-model file of the MATLAB original.
//
// Reference: CreateVoxelFEAmodel (matlab/TOP3D_XL.m:458-552), version 1.0:
//   4 header words, a version word and number (1.0),
//   "Resolution:" nelx nely nelz,
//   2 words and a flag (1: densities follow each solid voxel),
//   2 words and the number of solid voxels, then their indices (with a density each),
//   "Passive" "elements:" count and indices,
//   "Fixations:" count and rows (node, fix-x, fix-y, fix-z),
//   "Loads:" count and rows (node, Fx, Fy, Fz),
//   2 words and the number of further load cases, each "Loads:" count case rows.
// Indices are 1-based in MATLAB order: voxel i-1 = row + nely (x + nelx z), node
// n-1 = row + (nely+1)(x + (nelx+1) z), rows counted from the top.  They become
// the solid-mask layout of t3d_problem_create_voxels (ey = nely-1-row) and node
// coordinates (x, y, z) with y counted from the bottom (y = nely - row).
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "t3d_internal.hpp"

using namespace t3d;

namespace {

struct Tokens {
    std::vector<std::string> t;
    size_t k = 0;
    std::string err;
    bool word(std::string* w = nullptr)
    {
        if (k >= t.size()) return fail("unexpected end of file");
        if (w) *w = t[k];
        ++k;
        return true;
    }
    bool expect(const char* w)
    {
        if (k >= t.size()) return fail(std::string("unexpected end of file, expected '") + w + "'");
        if (t[k] != w) return fail("expected '" + std::string(w) + "', found '" + t[k] + "'");
        ++k;
        return true;
    }
    bool integer(long long& v)
    {
        if (k >= t.size()) return fail("unexpected end of file, expected an integer");
        char* end = nullptr;
        errno = 0;
        v = std::strtoll(t[k].c_str(), &end, 10);
        if (errno || !end || *end) return fail("expected an integer, found '" + t[k] + "'");
        ++k;
        return true;
    }
    bool real(double& v)
    {
        if (k >= t.size()) return fail("unexpected end of file, expected a number");
        char* end = nullptr;
        errno = 0;
        v = std::strtod(t[k].c_str(), &end);
        if (errno || !end || *end) return fail("expected a number, found '" + t[k] + "'");
        ++k;
        return true;
    }
    bool fail(const std::string& m)
    {
        if (err.empty()) err = m;
        return false;
    }
};

void free_model(t3d_voxel_model* m)
{
    if (!m) return;
    std::free(m->solid);
    std::free(m->passive);
    std::free(m->density);
    std::free(m->fix);
    std::free(m->loadNode);
    std::free(m->loadF);
    std::free(m);
}

template <class T>
T* alloc(size_t n)
{
    return static_cast<T*>(std::calloc(n ? n : 1, sizeof(T)));
}

bool parse(Tokens& tk, t3d_voxel_model& m)
{
    for (int i = 0; i < 4; ++i)
        if (!tk.word()) return false;
    double version = 0.0;
    if (!tk.word() || !tk.real(version)) return false;
    if (version != 1.0) return tk.fail("unsupported TopVoxel version " + std::to_string(version));
    long long nx = 0, ny = 0, nz = 0;
    if (!tk.expect("Resolution:") || !tk.integer(nx) || !tk.integer(ny) || !tk.integer(nz)) return false;
    if (nx < 1 || ny < 1 || nz < 1 || nx * ny * nz > (1ll << 31)) return tk.fail("bad resolution");
    m.nelx = int(nx);
    m.nely = int(ny);
    m.nelz = int(nz);
    const long long nvox = nx * ny * nz, nnod = (nx + 1) * (ny + 1) * (nz + 1);
    long long withDensity = 0, nsolid = 0;
    if (!tk.word() || !tk.word() || !tk.integer(withDensity)) return false;
    if (!tk.word() || !tk.word() || !tk.integer(nsolid)) return false;
    if (nsolid < 0 || nsolid > nvox) return tk.fail("bad number of solid voxels");
    m.solid = alloc<uint8_t>(size_t(nvox));
    m.passive = alloc<uint8_t>(size_t(nvox));
    if (withDensity) m.density = alloc<float>(size_t(nvox));
    if (!m.solid || !m.passive || (withDensity && !m.density)) return tk.fail("out of memory");
    // 1-based MATLAB voxel index -> solid-mask position (ey = nely - 1 - row)
    auto voxel = [&](long long i, size_t& pos) {
        if (i < 1 || i > nvox) return tk.fail("voxel index " + std::to_string(i) + " out of range");
        const long long v = i - 1, row = v % ny, x = (v / ny) % nx, z = v / (ny * nx);
        pos = size_t((z * nx + x) * ny + (ny - 1 - row));
        return true;
    };
    // 1-based MATLAB node index -> (x, y, z), y = nely - row
    auto node = [&](long long i, int32_t* xyz) {
        if (i < 1 || i > nnod) return tk.fail("node index " + std::to_string(i) + " out of range");
        const long long v = i - 1, row = v % (ny + 1), x = (v / (ny + 1)) % (nx + 1), z = v / ((ny + 1) * (nx + 1));
        xyz[0] = int32_t(x);
        xyz[1] = int32_t(ny - row);
        xyz[2] = int32_t(z);
        return true;
    };
    for (long long k = 0; k < nsolid; ++k) {
        long long i = 0;
        size_t pos = 0;
        if (!tk.integer(i) || !voxel(i, pos)) return false;
        m.solid[pos] = 1;
        if (withDensity) {
            double d = 0.0;
            if (!tk.real(d)) return false;
            m.density[pos] = float(d);
        }
    }
    long long npas = 0;
    if (!tk.expect("Passive") || !tk.expect("elements:") || !tk.integer(npas)) return false;
    if (npas < 0 || npas > nvox) return tk.fail("bad number of passive elements");
    for (long long k = 0; k < npas; ++k) {
        long long i = 0;
        size_t pos = 0;
        if (!tk.integer(i) || !voxel(i, pos)) return false;
        m.passive[pos] = 1;
    }
    long long nfix = 0;
    if (!tk.expect("Fixations:") || !tk.integer(nfix)) return false;
    if (nfix < 0 || nfix > nnod * 4) return tk.fail("bad number of fixations");
    m.nFix = int(nfix);
    m.fix = alloc<int32_t>(size_t(4 * nfix));
    for (long long k = 0; k < nfix; ++k) {
        long long i = 0, f[3] = {0, 0, 0};
        if (!tk.integer(i) || !node(i, m.fix + 4 * k)) return false;
        for (int c = 0; c < 3; ++c)
            if (!tk.integer(f[c])) return false;
        m.fix[4 * k + 3] = int32_t((f[0] ? 1 : 0) | (f[1] ? 2 : 0) | (f[2] ? 4 : 0));
    }
    long long nload = 0;
    if (!tk.expect("Loads:") || !tk.integer(nload)) return false;
    if (nload < 0 || nload > nnod * 4) return tk.fail("bad number of loads");
    m.nLoad = int(nload);
    m.loadNode = alloc<int32_t>(size_t(3 * nload));
    m.loadF = alloc<double>(size_t(3 * nload));
    for (long long k = 0; k < nload; ++k) {
        long long i = 0;
        if (!tk.integer(i) || !node(i, m.loadNode + 3 * k)) return false;
        for (int c = 0; c < 3; ++c)
            if (!tk.real(m.loadF[3 * k + c])) return false;
    }
    // further load cases (matlab/TOP3D_XL.m:509-518): read through and counted
    long long nextra = 0;
    if (!tk.word() || !tk.word() || !tk.integer(nextra)) return false;
    if (nextra < 0) return tk.fail("bad number of further load cases");
    for (long long e = 0; e < nextra; ++e) {
        long long n = 0, idx = 0;
        if (!tk.expect("Loads:") || !tk.integer(n) || !tk.integer(idx)) return false;
        for (long long k = 0; k < 4 * n; ++k) {
            double v = 0.0;
            if (!tk.real(v)) return false;
        }
    }
    m.extraLoadCases = int(nextra);
    return true;
}

}  // namespace

extern "C" {

int t3d_topvoxel_read(const char* path, t3d_voxel_model** out)
{
    T3D_CHECK(path && out, T3D_ERR_ARG, "t3d_topvoxel_read: null argument");
    *out = nullptr;
    FILE* fp = std::fopen(path, "rb");
    T3D_CHECK(fp, T3D_ERR_ARG, std::string("t3d_topvoxel_read: cannot open ") + path);
    Tokens tk;
    std::string cur;
    char buf[1 << 16];
    size_t got = 0;
    while ((got = std::fread(buf, 1, sizeof(buf), fp)) > 0)
        for (size_t i = 0; i < got; ++i) {
            const char ch = buf[i];
            if (ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r' || ch == '\f' || ch == '\v') {
                if (!cur.empty()) tk.t.push_back(cur), cur.clear();
            } else {
                cur.push_back(ch);
            }
        }
    if (!cur.empty()) tk.t.push_back(cur);
    std::fclose(fp);
    t3d_voxel_model* m = static_cast<t3d_voxel_model*>(std::calloc(1, sizeof(t3d_voxel_model)));
    T3D_CHECK(m, T3D_ERR_ARG, "t3d_topvoxel_read: out of memory");
    if (!parse(tk, *m)) {
        free_model(m);
        set_error(std::string("t3d_topvoxel_read: ") + path + ": " + tk.err);
        return T3D_ERR_ARG;
    }
    *out = m;
    return T3D_OK;
}

int t3d_voxel_model_destroy(t3d_voxel_model* m)
{
    free_model(m);
    return T3D_OK;
}

}  // extern "C"
