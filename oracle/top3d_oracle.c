/*
 * top3d_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see top3d_oracle.h).
 *
 * Deterministic restatement of the reference TOP3D_XL C++ port.  Citations are
 * file:line in the reference tree (cpp/...).  Built by oracle/Makefile with
 * -O2 -fopenmp -ffp-contract=off (no FMA; SSE2 binary64/binary32 arithmetic).
 * OpenMP is used only where every output is computed by exactly one thread in
 * a fixed order, so results do not depend on the thread count.
 */
#include "top3d_oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <float.h>

#define XALLOC(T, n) ((T*)calloc((size_t)(n) > 0 ? (size_t)(n) : 1, sizeof(T)))

/* ===================================================================== */
/* canonical reduction                                                    */
/* ===================================================================== */
static double canon_tile_items(const double* items, size_t n, size_t t)
{
    double s[O_CANON_LANES];
    for (int l = 0; l < O_CANON_LANES; ++l) {
        double acc = 0.0;
        for (int k = 0; k < O_CANON_PER_LANE; ++k) {
            size_t i = t * O_CANON_TILE + (size_t)k * O_CANON_LANES + (size_t)l;
            if (i < n) acc = acc + items[i];
        }
        s[l] = acc;
    }
    for (int w = O_CANON_LANES / 2; w >= 1; w >>= 1)
        for (int l = 0; l < w; ++l) s[l] = s[l] + s[l + w];
    return s[0];
}

static double canon_tile_dot(const double* a, const double* b, size_t n, size_t t)
{
    double s[O_CANON_LANES];
    for (int l = 0; l < O_CANON_LANES; ++l) {
        double acc = 0.0;
        for (int k = 0; k < O_CANON_PER_LANE; ++k) {
            size_t i = t * O_CANON_TILE + (size_t)k * O_CANON_LANES + (size_t)l;
            if (i < n) acc = acc + a[i] * b[i];
        }
        s[l] = acc;
    }
    for (int w = O_CANON_LANES / 2; w >= 1; w >>= 1)
        for (int l = 0; l < w; ++l) s[l] = s[l] + s[l + w];
    return s[0];
}

static double canon_reduce_partials(double* P, size_t m)
{
    while (m > 1) {
        size_t m2 = (m + O_CANON_TILE - 1) / O_CANON_TILE;
        double* Q = XALLOC(double, m2);
        for (size_t t = 0; t < m2; ++t) Q[t] = canon_tile_items(P, m, t);
        free(P);
        P = Q;
        m = m2;
    }
    double r = P[0];
    free(P);
    return r;
}

double o_canon_sum(const double* items, size_t n)
{
    if (n == 0) return 0.0;
    size_t m = (n + O_CANON_TILE - 1) / O_CANON_TILE;
    double* P = XALLOC(double, m);
#pragma omp parallel for schedule(static)
    for (long t = 0; t < (long)m; ++t) P[t] = canon_tile_items(items, n, (size_t)t);
    return canon_reduce_partials(P, m);
}

double o_canon_dot(const double* a, const double* b, size_t n)
{
    if (n == 0) return 0.0;
    size_t m = (n + O_CANON_TILE - 1) / O_CANON_TILE;
    double* P = XALLOC(double, m);
#pragma omp parallel for schedule(static)
    for (long t = 0; t < (long)m; ++t) P[t] = canon_tile_dot(a, b, n, (size_t)t);
    return canon_reduce_partials(P, m);
}

static inline double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max(a,b) */
static inline float fmaxr(float a, float b) { return (a < b) ? b : a; }   /* std::max(a,b) */
static inline float fclampr(float v, float lo, float hi)                  /* std::clamp */
{
    return (v < lo) ? lo : ((hi < v) ? hi : v);
}

/* ===================================================================== */
/* padding (padding.hpp:93-123)                                          */
/* ===================================================================== */
static int base_levels(uint64_t numVoxels)
{
    int L = 0;
    while (numVoxels >= 50000u) { ++L; numVoxels = (numVoxels + 4u) / 8u; }
    return L < 3 ? 3 : L;
}
static int pad_pow2(int n, int L)
{
    long step = 1L << L;
    return (int)(((long)n + step - 1) / step * step);
}
/* padding of a region with nv solid voxels (compute_adjusted_dims,
 * padding.hpp:42-53, counts the solid voxels) */
static void padding_n(int nelx, int nely, int nelz, uint64_t nv, int* rx, int* ry, int* rz, int* L)
{
    int l = base_levels(nv);
    *rx = pad_pow2(nelx, l); *ry = pad_pow2(nely, l); *rz = pad_pow2(nelz, l);
    if (L) *L = l;
}
void o_padding(int nelx, int nely, int nelz, int* rx, int* ry, int* rz, int* L)
{
    padding_n(nelx, nely, nelz, (uint64_t)nely * (uint64_t)nelx * (uint64_t)nelz, rx, ry, rz, L);
}

/* ===================================================================== */
/* Ke (fea.cpp:117-179)                                                   */
/* ===================================================================== */
void o_voxel_ke(double nu, double cellSize, double Ke[576])
{
    double E = 1.0;
    double lambda = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    double mu = E / (2.0 * (1.0 + nu));
    double D[6][6] = {
        {lambda + 2 * mu, lambda, lambda, 0, 0, 0},
        {lambda, lambda + 2 * mu, lambda, 0, 0, 0},
        {lambda, lambda, lambda + 2 * mu, 0, 0, 0},
        {0, 0, 0, mu, 0, 0},
        {0, 0, 0, 0, mu, 0},
        {0, 0, 0, 0, 0, mu}};
    memset(Ke, 0, sizeof(double) * 576);
    const double g = 1.0 / sqrt(3.0);
    const double gp[2] = {-g, g};
    const double w[2] = {1.0, 1.0};
    double hx = cellSize, hy = cellSize, hz = cellSize;
    double detJ = (hx * hy * hz) / 8.0;
    double sx = 2.0 / hx, sy = 2.0 / hy, sz = 2.0 / hz;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++) {
                double s = gp[a], t = gp[b], p = gp[c];
                /* natural-coordinate derivatives of the 8 trilinear shapes,
                 * node order (-,-,-),(+,-,-),(+,+,-),(-,+,-),(-,-,+),(+,-,+),(+,+,+),(-,+,+) */
                double dNds[8] = {-0.125 * (1 - t) * (1 - p), 0.125 * (1 - t) * (1 - p),
                                  0.125 * (1 + t) * (1 - p), -0.125 * (1 + t) * (1 - p),
                                  -0.125 * (1 - t) * (1 + p), 0.125 * (1 - t) * (1 + p),
                                  0.125 * (1 + t) * (1 + p), -0.125 * (1 + t) * (1 + p)};
                double dNdt[8] = {-0.125 * (1 - s) * (1 - p), -0.125 * (1 + s) * (1 - p),
                                  0.125 * (1 + s) * (1 - p), 0.125 * (1 - s) * (1 - p),
                                  -0.125 * (1 - s) * (1 + p), -0.125 * (1 + s) * (1 + p),
                                  0.125 * (1 + s) * (1 + p), 0.125 * (1 - s) * (1 + p)};
                double dNdp[8] = {-0.125 * (1 - s) * (1 - t), -0.125 * (1 + s) * (1 - t),
                                  -0.125 * (1 + s) * (1 + t), -0.125 * (1 - s) * (1 + t),
                                  0.125 * (1 - s) * (1 - t), 0.125 * (1 + s) * (1 - t),
                                  0.125 * (1 + s) * (1 + t), 0.125 * (1 - s) * (1 + t)};
                double dx[8], dy[8], dz[8];
                for (int i = 0; i < 8; i++) { dx[i] = dNds[i] * sx; dy[i] = dNdt[i] * sy; dz[i] = dNdp[i] * sz; }
                double B[6][24];
                memset(B, 0, sizeof B);
                for (int i = 0; i < 8; i++) {
                    int c0 = 3 * i;
                    B[0][c0 + 0] = dx[i];
                    B[1][c0 + 1] = dy[i];
                    B[2][c0 + 2] = dz[i];
                    B[3][c0 + 1] = dz[i]; B[3][c0 + 2] = dy[i];
                    B[4][c0 + 0] = dz[i]; B[4][c0 + 2] = dx[i];
                    B[5][c0 + 0] = dy[i]; B[5][c0 + 1] = dx[i];
                }
                double DB[6][24];
                for (int i = 0; i < 6; i++)
                    for (int j = 0; j < 24; j++) {
                        double sum = 0.0;
                        for (int k = 0; k < 6; k++) sum += D[i][k] * B[k][j];
                        DB[i][j] = sum;
                    }
                double weight = w[a] * w[b] * w[c] * detJ;
                for (int i = 0; i < 24; i++)
                    for (int j = 0; j < 24; j++) {
                        double sum = 0.0;
                        for (int k = 0; k < 6; k++) sum += B[k][i] * DB[k][j];
                        Ke[i * 24 + j] += sum * weight;
                    }
            }
}

/* ===================================================================== */
/* model builder (fea.cpp:181-447)                                        */
/* ===================================================================== */
static uint64_t spread3(uint64_t v)
{
    /* bit i of v (i < 21) moves to bit 3*i */
    uint64_t r = 0;
    for (int i = 0; i < 21; ++i) r |= ((v >> i) & 1ull) << (3 * i);
    return r;
}
static uint64_t morton_key(int ex, int ey, int ez)
{
    return (spread3((uint64_t)ex) << 2) | (spread3((uint64_t)ey) << 1) | spread3((uint64_t)ez);
}

typedef struct { uint64_t key; int32_t full; } o_mk;
static int cmp_mk(const void* a, const void* b)
{
    uint64_t ka = ((const o_mk*)a)->key, kb = ((const o_mk*)b)->key;
    return (ka < kb) ? -1 : (ka > kb) ? 1 : 0;
}

/* local node a -> offset (dx,dy,dz) in the structured node grid of cell
 * (cx, ty, cz), ty = ny-1-ey (fea.cpp:257-264) */
static const int A_DX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
static const int A_DY[8] = {1, 1, 0, 0, 1, 1, 0, 0};
static const int A_DZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};

o_model* o_model_new(int nely, int nelx, int nelz)
{
    return o_model_new_voxels(nely, nelx, nelz, NULL);
}

/* solid[(ez*nelx + ex)*nely + ey] != 0: solid voxel of the original region
 * (NULL: all); void voxels carry no element, like the padding
 * (fea.cpp:203-227, generalised from the cuboid's all-solid region) */
static o_model* build_mesh(int nely, int nelx, int nelz, const uint8_t* solid)
{
    o_model* m = (o_model*)calloc(1, sizeof(o_model));
    m->nelx = nelx; m->nely = nely; m->nelz = nelz;
    uint64_t nsolid = 0;
    for (size_t i = 0; i < (size_t)nely * nelx * nelz; ++i) nsolid += (!solid || solid[i]) ? 1u : 0u;
    padding_n(nelx, nely, nelz, nsolid, &m->resX, &m->resY, &m->resZ, &m->baseL);
    const int nx = m->resX, ny = m->resY, nz = m->resZ;
    m->nnx = nx + 1; m->nny = ny + 1; m->nnz = nz + 1;
    m->nN = m->nnx * m->nny * m->nnz;
    m->nDOF = 3 * m->nN;
    int ne = (int)nsolid;
    m->ne = ne;
    o_mk* el = XALLOC(o_mk, ne > 0 ? ne : 1);
    int k = 0;
    for (int ez = 0; ez < nelz; ++ez)
        for (int ex = 0; ex < nelx; ++ex)
            for (int ey = 0; ey < nely; ++ey) {
                if (solid && !solid[((size_t)ez * nelx + ex) * nely + ey]) continue;
                el[k].key = morton_key(ex, ey, ez);
                el[k].full = ny * nx * ez + ny * ex + (ny - 1 - ey);
                ++k;
            }
    /* Morton order (fea.cpp:289-317); keys are unique so the order is total */
    qsort(el, (size_t)ne, sizeof(o_mk), cmp_mk);
    m->eleMapBack = XALLOC(int32_t, ne);
    m->eNodMat = XALLOC(int32_t, 8 * (size_t)ne);
    m->color = XALLOC(int8_t, ne);
    m->compactToCell = XALLOC(int32_t, ne);
    m->cellToCompact = XALLOC(int32_t, (size_t)nx * ny * nz);
    for (size_t c = 0; c < (size_t)nx * ny * nz; ++c) m->cellToCompact[c] = -1;
    const int nnx = m->nnx, nny = m->nny;
    for (int e = 0; e < ne; ++e) {
        int full = el[e].full;
        m->eleMapBack[e] = full;
        int cpl = nx * ny;
        int ez = full / cpl, rem = full % cpl;
        int ex = rem / ny, ty = rem % ny; /* ty = ny-1-ey */
        int ey = ny - 1 - ty;
        m->color[e] = (int8_t)((ex & 1) | ((ey & 1) << 1) | ((ez & 1) << 2));
        for (int a = 0; a < 8; ++a) {
            int ix = ex + A_DX[a], iy = ty + A_DY[a], iz = ez + A_DZ[a];
            m->eNodMat[8 * (size_t)e + a] = nnx * nny * iz + nnx * iy + ix;
        }
        int cell = (ez * ny + ty) * nx + ex;
        m->compactToCell[e] = cell;
        m->cellToCompact[cell] = e;
    }
    free(el);
    /* colour buckets, ascending compact ids (fea.cpp:318-326) */
    m->bucket = XALLOC(int32_t, ne);
    int cnt[8] = {0};
    for (int e = 0; e < ne; ++e) cnt[m->color[e]]++;
    m->bucketStart[0] = 0;
    for (int c = 0; c < 8; ++c) m->bucketStart[c + 1] = m->bucketStart[c] + cnt[c];
    int pos[8];
    for (int c = 0; c < 8; ++c) pos[c] = m->bucketStart[c];
    for (int e = 0; e < ne; ++e) m->bucket[pos[m->color[e]]++] = e;
    o_voxel_ke(0.3, 1.0, m->Ke); /* fea.cpp:345 */
    m->isFree = XALLOC(uint8_t, m->nDOF);
    m->F = XALLOC(double, m->nDOF);
    return m;
}

/* free-DOF maps after the boundary conditions: nodes touched by no element
 * fixed (fea.cpp:405-417), freeDofIndex and eDofFreeMat (fea.cpp:419-447) */
static void finish_bcs(o_model* m)
{
    const int ne = m->ne;
    {
        uint8_t* used = XALLOC(uint8_t, m->nN);
        for (size_t i = 0; i < 8 * (size_t)ne; ++i) used[m->eNodMat[i]] = 1;
        for (int n = 0; n < m->nN; ++n)
            if (!used[n]) m->isFree[3 * n] = m->isFree[3 * n + 1] = m->isFree[3 * n + 2] = 0;
        free(used);
    }
    int nF = 0;
    for (int i = 0; i < m->nDOF; ++i) nF += m->isFree[i];
    m->nF = nF;
    m->freeDofIndex = XALLOC(int32_t, nF);
    int32_t* g2f = XALLOC(int32_t, m->nDOF);
    for (int i = 0, j = 0; i < m->nDOF; ++i) {
        if (m->isFree[i]) { m->freeDofIndex[j] = i; g2f[i] = j++; }
        else g2f[i] = -1;
    }
    m->eDofFree = XALLOC(int32_t, 24 * (size_t)ne);
    for (int e = 0; e < ne; ++e)
        for (int a = 0; a < 8; ++a)
            for (int c = 0; c < 3; ++c)
                m->eDofFree[24 * (size_t)e + 3 * a + c] = g2f[3 * m->eNodMat[8 * (size_t)e + a] + c];
    free(g2f);
}

o_model* o_model_new_voxels(int nely, int nelx, int nelz, const uint8_t* solid)
{
    o_model* m = build_mesh(nely, nelx, nelz, solid);
    const int nnx = m->nnx, nny = m->nny, nnz = m->nnz, nz = m->resZ;
    /* boundary conditions (fea.cpp:360-447) */
    memset(m->isFree, 1, (size_t)m->nDOF);
    for (int iz = 0; iz < nnz; ++iz)
        for (int iy = 0; iy < nny; ++iy) {
            int n = nnx * nny * iz + nnx * iy + 0;
            m->isFree[3 * n] = m->isFree[3 * n + 1] = m->isFree[3 * n + 2] = 0;
        }
    {
        int ixLoad = nelx < nnx - 1 ? nelx : nnx - 1;
        int izMax = (nz / 6) > 1 ? (nz / 6) : 1;
        int count = (izMax + 1) * nny;
        double fz = -1.0 / (double)count;
        for (int iz = 0; iz <= izMax; ++iz)
            for (int iy = 0; iy < nny; ++iy) {
                int n = nnx * nny * iz + nnx * iy + ixLoad;
                m->F[3 * n + 2] += fz;
            }
    }
    finish_bcs(m);
    return m;
}

/* fixations and loads as lists on the original box's nodes (x, y, z), y from
 * the box's bottom (padded node row resY - y), as a TopVoxel file gives them
 * (matlab/TOP3D_XL.m:470-552) and FEA_ApplyBoundaryCondition applies them
 * (:1319-1360): the first entry of a node counts (unique), loaded nodes are not
 * fixed (setdiff), a load sets the node's force, flagged DOFs are fixed */
o_model* o_model_new_bc(int nely, int nelx, int nelz, const uint8_t* solid, int nFix, const int32_t* fix, int nLoad,
                        const int32_t* loadNode, const double* loadF)
{
    o_model* m = build_mesh(nely, nelx, nelz, solid);
    const int nnx = m->nnx, nny = m->nny;
    memset(m->isFree, 1, (size_t)m->nDOF);
    memset(m->F, 0, sizeof(double) * (size_t)m->nDOF);
    uint8_t* loaded = XALLOC(uint8_t, m->nN);
    uint8_t* seen = XALLOC(uint8_t, m->nN);
    int bad = 0;
    for (int i = 0; i < nLoad && !bad; ++i) {
        const int32_t* c = loadNode + 3 * i;
        if (c[0] < 0 || c[0] > nelx || c[1] < 0 || c[1] > nely || c[2] < 0 || c[2] > nelz) { bad = 1; break; }
        const int n = nnx * nny * c[2] + nnx * (m->resY - c[1]) + c[0];
        if (loaded[n]) continue;
        loaded[n] = 1;
        for (int d = 0; d < 3; ++d) m->F[3 * (size_t)n + d] = loadF[3 * i + d];
    }
    for (int i = 0; i < nFix && !bad; ++i) {
        const int32_t* c = fix + 4 * i;
        if (c[0] < 0 || c[0] > nelx || c[1] < 0 || c[1] > nely || c[2] < 0 || c[2] > nelz) { bad = 1; break; }
        const int n = nnx * nny * c[2] + nnx * (m->resY - c[1]) + c[0];
        if (seen[n] || loaded[n]) continue;
        seen[n] = 1;
        for (int d = 0; d < 3; ++d)
            if ((c[3] >> d) & 1) m->isFree[3 * (size_t)n + d] = 0;
    }
    free(loaded);
    free(seen);
    if (bad) {
        free(m->isFree); free(m->F); m->isFree = NULL; m->F = NULL;
        o_model_free(m);
        return NULL;
    }
    finish_bcs(m);
    return m;
}

void o_model_free(o_model* m)
{
    if (!m) return;
    free(m->eleMapBack); free(m->eNodMat); free(m->color); free(m->bucket);
    free(m->F); free(m->isFree); free(m->freeDofIndex); free(m->eDofFree);
    free(m->cellToCompact); free(m->compactToCell);
    free(m);
}

void o_free_to_full(const o_model* m, const double* vfree, double* vfull)
{
    memset(vfull, 0, sizeof(double) * (size_t)m->nDOF);
    for (int i = 0; i < m->nF; ++i) vfull[m->freeDofIndex[i]] = vfree[i];
}
void o_full_to_free(const o_model* m, const double* vfull, double* vfree)
{
    for (int i = 0; i < m->nF; ++i) vfree[i] = vfull[m->freeDofIndex[i]];
}

/* ===================================================================== */
/* K*u (fea.cpp:692-837)                                                  */
/* ===================================================================== */
void o_kxu_free(const o_model* m, const double* E, const double* uFree, double* yFree)
{
    memset(yFree, 0, sizeof(double) * (size_t)m->nF);
    const double* K = m->Ke;
    for (int col = 0; col < 8; ++col) {
        const int b0 = m->bucketStart[col], b1 = m->bucketStart[col + 1];
#pragma omp parallel for schedule(static)
        for (int bi = b0; bi < b1; ++bi) {
            const int e = m->bucket[bi];
            const int32_t* ed = m->eDofFree + 24 * (size_t)e;
            double ue[24], fe[24];
            for (int a = 0; a < 24; ++a) ue[a] = ed[a] >= 0 ? uFree[ed[a]] : 0.0;
            for (int i = 0; i < 24; ++i) {
                double s = 0.0;
                for (int j = 0; j < 24; ++j) s += K[i * 24 + j] * ue[j];
                fe[i] = s;
            }
            const double Ee = E[e];
            if (Ee <= 1.0e-16) continue;
            for (int a = 0; a < 24; ++a)
                if (ed[a] >= 0) yFree[ed[a]] += fe[a] * Ee;
        }
    }
}

void o_kxu_full(const o_model* m, const double* E, const double* u, double* y)
{
    memset(y, 0, sizeof(double) * (size_t)m->nDOF);
    const double* K = m->Ke;
    for (int col = 0; col < 8; ++col) {
        const int b0 = m->bucketStart[col], b1 = m->bucketStart[col + 1];
#pragma omp parallel for schedule(static)
        for (int bi = b0; bi < b1; ++bi) {
            const int e = m->bucket[bi];
            const int32_t* en = m->eNodMat + 8 * (size_t)e;
            double ue[24], fe[24];
            for (int a = 0; a < 8; ++a)
                for (int c = 0; c < 3; ++c) {
                    int g = 3 * en[a] + c;
                    ue[3 * a + c] = m->isFree[g] ? u[g] : 0.0;
                }
            for (int i = 0; i < 24; ++i) {
                double s = 0.0;
                for (int j = 0; j < 24; ++j) s += K[i * 24 + j] * ue[j];
                fe[i] = s;
            }
            const double Ee = E[e];
            if (Ee <= 1.0e-16) continue;
            for (int a = 0; a < 8; ++a)
                for (int c = 0; c < 3; ++c) {
                    int g = 3 * en[a] + c;
                    if (m->isFree[g]) y[g] += fe[3 * a + c] * Ee;
                }
        }
    }
}

void o_kxu_full_f32(const o_model* m, const float* E, const float* u, float* y)
{
    memset(y, 0, sizeof(float) * (size_t)m->nDOF);
    float K[576];
    for (int i = 0; i < 576; ++i) K[i] = (float)m->Ke[i];
    for (int col = 0; col < 8; ++col) {
        const int b0 = m->bucketStart[col], b1 = m->bucketStart[col + 1];
#pragma omp parallel for schedule(static)
        for (int bi = b0; bi < b1; ++bi) {
            const int e = m->bucket[bi];
            const int32_t* en = m->eNodMat + 8 * (size_t)e;
            float ue[24], fe[24];
            for (int a = 0; a < 8; ++a)
                for (int c = 0; c < 3; ++c) {
                    int g = 3 * en[a] + c;
                    ue[3 * a + c] = m->isFree[g] ? u[g] : 0.0f;
                }
            for (int i = 0; i < 24; ++i) {
                float s = 0.0f;
                for (int j = 0; j < 24; ++j) s += K[i * 24 + j] * ue[j];
                fe[i] = s;
            }
            const float Ee = E[e];
            if (Ee <= 1.0e-16f) continue;
            for (int a = 0; a < 8; ++a)
                for (int c = 0; c < 3; ++c) {
                    int g = 3 * en[a] + c;
                    if (m->isFree[g]) y[g] += fe[3 * a + c] * Ee;
                }
        }
    }
}

/* ===================================================================== */
/* compliance (fea.cpp:587-677)                                           */
/* ===================================================================== */
double o_compliance(const o_model* m, const double* E, const double* U, double* ce)
{
    const int ne = m->ne;
    const double* K = m->Ke;
    double* items = XALLOC(double, ne);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ne; ++e) {
        const double Ee = E[e];
        if (Ee <= 1.0e-16) { ce[e] = 0.0; items[e] = 0.0; continue; }
        const int32_t* en = m->eNodMat + 8 * (size_t)e;
        double ue[24];
        for (int a = 0; a < 8; ++a)
            for (int c = 0; c < 3; ++c) ue[3 * a + c] = U[3 * (size_t)en[a] + c];
        double s = 0.0;
        for (int i = 0; i < 24; ++i) {
            double rowDot = 0.0;
            for (int j = 0; j < 24; ++j) rowDot += K[i * 24 + j] * ue[j];
            s += ue[i] * rowDot;
        }
        ce[e] = s;
        items[e] = Ee * s;
    }
    double C = o_canon_sum(items, (size_t)ne);
    free(items);
    return C;
}

/* ===================================================================== */
/* multigrid                                                              */
/* ===================================================================== */
static void shape8(double xi, double eta, double zeta, double N[8])
{
    /* hierarchy.cpp:13-22 */
    N[0] = 0.125 * (1.0 - xi) * (1.0 - eta) * (1.0 - zeta);
    N[1] = 0.125 * (1.0 + xi) * (1.0 - eta) * (1.0 - zeta);
    N[2] = 0.125 * (1.0 + xi) * (1.0 + eta) * (1.0 - zeta);
    N[3] = 0.125 * (1.0 - xi) * (1.0 + eta) * (1.0 - zeta);
    N[4] = 0.125 * (1.0 - xi) * (1.0 - eta) * (1.0 + zeta);
    N[5] = 0.125 * (1.0 + xi) * (1.0 - eta) * (1.0 + zeta);
    N[6] = 0.125 * (1.0 + xi) * (1.0 + eta) * (1.0 + zeta);
    N[7] = 0.125 * (1.0 - xi) * (1.0 + eta) * (1.0 + zeta);
}
static double* weights_table(int span)
{
    /* hierarchy.cpp:26-42 */
    int g = span + 1;
    double* W = XALLOC(double, g * g * g * 8);
    for (int iz = 0; iz <= span; ++iz)
        for (int iy = 0; iy <= span; ++iy)
            for (int ix = 0; ix <= span; ++ix) {
                double xi = -1.0 + 2.0 * ((double)ix / (double)span);
                double eta = -1.0 + 2.0 * ((double)iy / (double)span);
                double zeta = -1.0 + 2.0 * ((double)iz / (double)span);
                double N[8];
                shape8(xi, eta, zeta, N);
                int v = (iz * g + iy) * g + ix;
                for (int a = 0; a < 8; ++a) W[v * 8 + a] = N[a];
            }
    return W;
}

/* restriction, clamped (2s-1)^3 hat stencil (transfers.cpp:95-191 / 449-533) */
static void restrict_core(const o_mg* g, int lc, const double* rf, double* rc, int ncomp, int divide)
{
    const int lf = lc - 1;
    const int cNnx = g->resX[lc] + 1, cNny = g->resY[lc] + 1, cNnz = g->resZ[lc] + 1;
    const int fNnx = g->resX[lf] + 1, fNny = g->resY[lf] + 1, fNnz = g->resZ[lf] + 1;
    const int span = g->span[lc];
    const double invSpan = 1.0 / span;
    const int wSize = 2 * span - 1;
    double W[16];
    for (int i = 0; i < wSize; ++i) W[i] = 1.0 - abs(i - (span - 1)) * invSpan;
#pragma omp parallel for collapse(2) schedule(static)
    for (int cz = 0; cz < cNnz; ++cz)
        for (int cy = 0; cy < cNny; ++cy) {
            int fz0 = cz * span, fy0 = cy * span;
            for (int cx = 0; cx < cNnx; ++cx) {
                int fx0 = cx * span;
                double vs[3] = {0.0, 0.0, 0.0};
                double wSum = 0.0;
                int iz0 = (fz0 - span + 1 < 0) ? (span - 1 - fz0) : 0;
                int iz1 = (fz0 + span - 1 >= fNnz) ? (wSize - (fz0 + span - fNnz)) : wSize;
                int iy0 = (fy0 - span + 1 < 0) ? (span - 1 - fy0) : 0;
                int iy1 = (fy0 + span - 1 >= fNny) ? (wSize - (fy0 + span - fNny)) : wSize;
                int ix0 = (fx0 - span + 1 < 0) ? (span - 1 - fx0) : 0;
                int ix1 = (fx0 + span - 1 >= fNnx) ? (wSize - (fx0 + span - fNnx)) : wSize;
                for (int iz = iz0; iz < iz1; ++iz) {
                    double wz = W[iz];
                    size_t zo = (size_t)(fz0 + (iz - (span - 1))) * fNny * fNnx;
                    for (int iy = iy0; iy < iy1; ++iy) {
                        double wzy = wz * W[iy];
                        size_t yo = (size_t)(fy0 + (iy - (span - 1))) * fNnx;
                        for (int ix = ix0; ix < ix1; ++ix) {
                            double w = wzy * W[ix];
                            size_t fi = (zo + yo + (size_t)(fx0 + (ix - (span - 1)))) * ncomp;
                            for (int c = 0; c < ncomp; ++c) vs[c] += rf[fi + c] * w;
                            wSum += w;
                        }
                    }
                }
                size_t ci = ((size_t)cz * cNny * cNnx + (size_t)cy * cNnx + cx) * ncomp;
                if (wSum > 1e-12) {
                    if (divide) {
                        for (int c = 0; c < ncomp; ++c) rc[ci + c] = vs[c] / wSum;
                    } else {
                        double invW = 1.0 / wSum;
                        for (int c = 0; c < ncomp; ++c) rc[ci + c] = vs[c] * invW;
                    }
                } else {
                    for (int c = 0; c < ncomp; ++c) rc[ci + c] = 0.0;
                }
            }
        }
}
void o_restrict_vec3(const o_mg* g, int lc, const double* rf, double* rc) { restrict_core(g, lc, rf, rc, 3, 0); }
void o_restrict_scalar_div(const o_mg* g, int lc, const double* rf, double* rc) { restrict_core(g, lc, rf, rc, 1, 1); }

/* trilinear prolongation (transfers.cpp:378-445) */
void o_prolong_vec3(const o_mg* g, int lc, const double* xc, double* xf)
{
    const int lf = lc - 1;
    const int fNnx = g->resX[lf] + 1, fNny = g->resY[lf] + 1, fNnz = g->resZ[lf] + 1;
    const int cNnx = g->resX[lc] + 1, cNny = g->resY[lc] + 1;
    const int span = g->span[lc];
    const double invSpan = 1.0 / span;
#pragma omp parallel for collapse(2) schedule(static)
    for (int fz = 0; fz < fNnz; ++fz)
        for (int fy = 0; fy < fNny; ++fy)
            for (int fx = 0; fx < fNnx; ++fx) {
                int cx = fx / span, cy = fy / span, cz = fz / span;
                int rx = fx % span, ry = fy % span, rz = fz % span;
                if (cx >= g->resX[lc]) { cx = g->resX[lc] - 1; rx = span; }
                if (cy >= g->resY[lc]) { cy = g->resY[lc] - 1; ry = span; }
                if (cz >= g->resZ[lc]) { cz = g->resZ[lc] - 1; rz = span; }
                double wx1 = (double)rx * invSpan, wx0 = 1.0 - wx1;
                double wy1 = (double)ry * invSpan, wy0 = 1.0 - wy1;
                double wz1 = (double)rz * invSpan, wz0 = 1.0 - wz1;
                int b00 = cNnx * cNny * cz + cNnx * cy + cx;
                int b01 = b00 + 1, b10 = b00 + cNnx, b11 = b10 + 1;
                int n00 = b00 + cNnx * cNny, n01 = n00 + 1, n10 = n00 + cNnx, n11 = n10 + 1;
                size_t fi = ((size_t)fNnx * fNny * fz + (size_t)fNnx * fy + fx) * 3;
                for (int c = 0; c < 3; ++c) {
                    double v = xc[b00 * 3 + c] * wx0 * wy0 * wz0 + xc[b01 * 3 + c] * wx1 * wy0 * wz0 +
                               xc[b10 * 3 + c] * wx0 * wy1 * wz0 + xc[b11 * 3 + c] * wx1 * wy1 * wz0 +
                               xc[n00 * 3 + c] * wx0 * wy0 * wz1 + xc[n01 * 3 + c] * wx1 * wy0 * wz1 +
                               xc[n10 * 3 + c] * wx0 * wy1 * wz1 + xc[n11 * 3 + c] * wx1 * wy1 * wz1;
                    xf[fi + c] = v;
                }
            }
}

o_mg* o_mg_new(const o_model* m, int nonDyadic, int maxLevels, float weight, float Emin, int NlimitDofs)
{
    o_mg* g = (o_mg*)calloc(1, sizeof(o_mg));
    g->weight = (double)weight;
    g->Emin = (double)Emin;
    g->NlimitDofs = NlimitDofs;
    g->cholOk = 1;
    /* hierarchy.cpp:84-138: level 0 = finest; first coarsening span 4 when
     * nonDyadic, then span 2; stop on indivisibility or when min dim <= 2 */
    int nx = m->resX, ny = m->resY, nz = m->resZ;
    g->resX[0] = nx; g->resY[0] = ny; g->resZ[0] = nz; g->span[0] = 1;
    g->nlev = 1;
    if (maxLevels > O_MAXLEV) maxLevels = O_MAXLEV;
    for (int li = 1; li < maxLevels; ++li) {
        int span = (li == 1 && nonDyadic) ? 4 : 2;
        if (nx % span != 0 || ny % span != 0 || nz % span != 0) break;
        if (nx / span < 1 || ny / span < 1 || nz / span < 1) break;
        nx /= span; ny /= span; nz /= span;
        int l = g->nlev++;
        g->resX[l] = nx; g->resY[l] = ny; g->resZ[l] = nz; g->span[l] = span;
        g->W[l] = weights_table(span);
        int mn = nx < ny ? nx : ny; mn = mn < nz ? mn : nz;
        if (mn <= 2) break;
    }
    for (int l = 0; l < g->nlev; ++l) {
        g->nN[l] = (g->resX[l] + 1) * (g->resY[l] + 1) * (g->resZ[l] + 1);
        g->mask[l] = XALLOC(uint8_t, 3 * (size_t)g->nN[l]);
        g->diag[l] = XALLOC(double, 3 * (size_t)g->nN[l]);
        g->r[l] = XALLOC(double, 3 * (size_t)g->nN[l]);
        g->x[l] = XALLOC(double, 3 * (size_t)g->nN[l]);
    }
    g->tmp = XALLOC(double, 3 * (size_t)g->nN[0]);
    /* masks (masks.cpp:11-37) */
    for (int i = 0; i < 3 * g->nN[0]; ++i) g->mask[0][i] = m->isFree[i] ? 0 : 1;
    for (int l = 0; l + 1 < g->nlev; ++l) {
        int fnn = g->nN[l], cnn = g->nN[l + 1];
        double* rf = XALLOC(double, fnn);
        double* rc = XALLOC(double, cnn);
        for (int c = 0; c < 3; ++c) {
            for (int n = 0; n < fnn; ++n) rf[n] = g->mask[l][3 * n + c] ? 1.0 : 0.0;
            o_restrict_scalar_div(g, l + 1, rf, rc);
            for (int n = 0; n < cnn; ++n) g->mask[l + 1][3 * n + c] = (rc[n] > 0.5) ? 1 : 0;
        }
        free(rf); free(rc);
    }
    return g;
}

void o_mg_free(o_mg* g)
{
    if (!g) return;
    for (int l = 0; l < O_MAXLEV; ++l) { free(g->W[l]); free(g->mask[l]); free(g->diag[l]); free(g->r[l]); free(g->x[l]); }
    free(g->tmp); free(g->Lc); free(g->Kc);
    free(g);
}

/* diagonal coarsening (diagonal.cpp:66-138) */
static void coarsen_diag(const o_mg* g, int lc)
{
    const int lf = lc - 1;
    const int span = g->span[lc], grid = span + 1;
    const int rXc = g->resX[lc], rYc = g->resY[lc], rZc = g->resZ[lc];
    const int fNx = g->resX[lf] + 1, fNy = g->resY[lf] + 1;
    const int cNx = rXc + 1, cNy = rYc + 1;
    const double* Df = g->diag[lf];
    const uint8_t* fm = g->mask[lf];
    const uint8_t* cm = g->mask[lc];
    double* Dc = g->diag[lc];
    const int nc = 3 * g->nN[lc];
    double* wsum = XALLOC(double, nc);
    for (int i = 0; i < nc; ++i) Dc[i] = 0.0;
    for (int color = 0; color < 8; ++color) {
#pragma omp parallel for collapse(3) schedule(static)
        for (int cez = 0; cez < rZc; ++cez)
            for (int cex = 0; cex < rXc; ++cex)
                for (int cey = 0; cey < rYc; ++cey) {
                    int myColor = (cex & 1) | ((cey & 1) << 1) | ((cez & 1) << 2);
                    if (myColor != color) continue;
                    int fx0 = cex * span, fy0 = (rYc - cey) * span, fz0 = cez * span;
                    int cidx[8];
                    for (int a = 0; a < 8; ++a) {
                        int x = cex + A_DX[a], y = (rYc - cey - 1) + A_DY[a], z = cez + A_DZ[a];
                        cidx[a] = cNx * cNy * z + cNx * y + x;
                    }
                    for (int iz = 0; iz <= span; ++iz)
                        for (int iy = 0; iy <= span; ++iy)
                            for (int ix = 0; ix <= span; ++ix) {
                                int fn = fNx * fNy * (fz0 + iz) + fNx * (fy0 - iy) + (fx0 + ix);
                                const double* W = &g->W[lc][((iz * grid + iy) * grid + ix) * 8];
                                for (int a = 0; a < 8; ++a) {
                                    double w2 = W[a] * W[a];
                                    for (int c = 0; c < 3; ++c) {
                                        int fd = 3 * fn + c, cd = 3 * cidx[a] + c;
                                        if (cm[cd]) continue;
                                        if (fm[fd]) continue;
                                        Dc[cd] += w2 * Df[fd];
                                        wsum[cd] += w2;
                                    }
                                }
                            }
                }
    }
    for (int i = 0; i < nc; ++i) {
        if (cm[i]) Dc[i] = 1.0;
        else if (wsum[i] > 0) {
            Dc[i] /= wsum[i];
            if (!(Dc[i] > 0.0)) Dc[i] = 1.0;
        } else Dc[i] = 1.0;
    }
    free(wsum);
}

/* T for sub-position (ix,iy,iz) of the coarsest element (coarsest.cpp:179-191) */
static void build_T(const double* Wc, int grid, int ix, int iy, int iz, double* T)
{
    memset(T, 0, sizeof(double) * 576);
    for (int v = 0; v < 8; ++v) {
        int vx = (v == 0 || v == 3 || v == 4 || v == 7) ? 0 : 1;
        int vy = (v == 0 || v == 1 || v == 4 || v == 5) ? 0 : 1;
        int vz = (v <= 3) ? 0 : 1;
        const double* W = &Wc[(((iz + vz) * grid + (iy + vy)) * grid + (ix + vx)) * 8];
        for (int a = 0; a < 8; ++a)
            for (int c = 0; c < 3; ++c) T[(3 * v + c) * 24 + (3 * a + c)] = W[a];
    }
}
/* Ksub = T^T (A T), sequential k sums (coarsest.cpp:193-212, 289-303) */
static void triple_product(const double* A, const double* T, double* out, int accumulate)
{
    double M[576];
    for (int i = 0; i < 24; i++)
        for (int j = 0; j < 24; j++) {
            double s = 0.0;
            for (int k = 0; k < 24; k++) s += A[i * 24 + k] * T[k * 24 + j];
            M[i * 24 + j] = s;
        }
    for (int i = 0; i < 24; i++)
        for (int j = 0; j < 24; j++) {
            double s = 0.0;
            for (int k = 0; k < 24; k++) s += T[k * 24 + i] * M[k * 24 + j];
            if (accumulate) out[i * 24 + j] += s;
            else out[i * 24 + j] = s;
        }
}

/* Galerkin coarsest assembly in the canonical element order (cez,cex,cey)
 * (coarsest.cpp:138-323) */
static void assemble_coarsest(const o_model* m, const o_mg* g, const double* E, double* Kc, int N)
{
    const int L = g->nlev - 1;
    const int s = g->span[L], grid = s + 1;
    const int fX = g->resX[0], fY = g->resY[0], fZ = g->resZ[0];
    const int cX = g->resX[L], cY = g->resY[L], cZ = g->resZ[L];
    const int fnnx = fX + 1, fnny = fY + 1;
    const int cnnx = cX + 1, cnny = cY + 1;
    const uint8_t* fmask = g->mask[0];
    double* emFull = XALLOC(double, (size_t)fX * fY * fZ);
    for (int e = 0; e < m->ne; ++e) emFull[m->eleMapBack[e]] = E[e];
    double* tpl = XALLOC(double, (size_t)s * s * s * 576);
    for (int iz = 0; iz < s; ++iz)
        for (int iy = 0; iy < s; ++iy)
            for (int ix = 0; ix < s; ++ix) {
                double T[576];
                build_T(g->W[L], grid, ix, iy, iz, T);
                triple_product(m->Ke, T, tpl + (size_t)((iz * s + iy) * s + ix) * 576, 0);
            }
    memset(Kc, 0, sizeof(double) * (size_t)N * N);
    for (int cez = 0; cez < cZ; ++cez)
        for (int cex = 0; cex < cX; ++cex)
            for (int cey = 0; cey < cY; ++cey) {
                int cdof[24];
                for (int a = 0; a < 8; ++a) {
                    int x = cex + A_DX[a], y = (cY - cey - 1) + A_DY[a], z = cez + A_DZ[a];
                    int n = cnnx * cnny * z + cnnx * y + x;
                    for (int c = 0; c < 3; ++c) cdof[3 * a + c] = 3 * n + c;
                }
                double Kce[576];
                memset(Kce, 0, sizeof Kce);
                int fx0 = cex * s, fy0 = (cY - cey) * s, fz0 = cez * s;
                for (int iz = 0; iz < s; ++iz)
                    for (int iy = 0; iy < s; ++iy)
                        for (int ix = 0; ix < s; ++ix) {
                            int fex = fx0 + ix, fey = fy0 - iy - 1, fez = fz0 + iz;
                            int ef = (fY * fX) * fez + fY * fex + (fY - 1 - fey);
                            double Ee = dmax(g->Emin, emFull[ef]);
                            /* mask test on the structured finest element numbered ef
                             * in build_structured_level order (ez, ex, ey): the
                             * y-mirrored element (coarsest.cpp:250-256) */
                            int mez = ef / (fX * fY), rem = ef % (fX * fY);
                            int mex = rem / fY, mey = rem % fY;
                            int nodes[8];
                            for (int v = 0; v < 8; ++v) {
                                int x = mex + A_DX[v], y = (fY - mey - 1) + A_DY[v], z = mez + A_DZ[v];
                                nodes[v] = fnnx * fnny * z + fnnx * y + x;
                            }
                            int masked = 0;
                            for (int v = 0; v < 8; ++v)
                                if (fmask[3 * nodes[v]] | fmask[3 * nodes[v] + 1] | fmask[3 * nodes[v] + 2]) { masked = 1; break; }
                            if (!masked) {
                                const double* Ks = tpl + (size_t)((iz * s + iy) * s + ix) * 576;
                                for (int i = 0; i < 576; ++i) Kce[i] += Ee * Ks[i];
                            } else {
                                double Kf[576], T[576];
                                for (int i = 0; i < 576; ++i) Kf[i] = Ee * m->Ke[i];
                                for (int v = 0; v < 8; ++v)
                                    for (int c = 0; c < 3; ++c)
                                        if (fmask[3 * nodes[v] + c]) {
                                            int d = 3 * v + c;
                                            for (int j = 0; j < 24; ++j) { Kf[d * 24 + j] = 0.0; Kf[j * 24 + d] = 0.0; }
                                            Kf[d * 24 + d] = 1.0;
                                        }
                                build_T(g->W[L], grid, ix, iy, iz, T);
                                triple_product(Kf, T, Kce, 1);
                            }
                        }
                for (int i = 0; i < 24; ++i)
                    for (int j = 0; j < 24; ++j) {
                        int gi = cdof[i], gj = cdof[j];
                        if (gi >= gj) Kc[(size_t)gi * N + gj] += Kce[i * 24 + j];
                    }
            }
    free(tpl);
    free(emFull);
    /* coarsest-level BC rows/cols (preconditioner.cpp:64-73) */
    const uint8_t* cm = g->mask[L];
    for (int d = 0; d < N; ++d)
        if (cm[d]) {
            for (int j = 0; j < N; ++j) { Kc[(size_t)d * N + j] = 0.0; Kc[(size_t)j * N + d] = 0.0; }
            Kc[(size_t)d * N + d] = 1.0;
        }
}

/* dense Cholesky (coarsest.cpp:13-90): row form if N < 256, else column form */
int o_chol_inplace(double* A, int N)
{
    if (N < 256) {
        for (int i = 0; i < N; i++) {
            for (int j = 0; j <= i; j++) {
                double sum = A[(size_t)i * N + j];
                for (int k = 0; k < j; k++) sum -= A[(size_t)i * N + k] * A[(size_t)j * N + k];
                if (i == j) {
                    if (sum <= 0.0) return 0;
                    A[(size_t)i * N + j] = sqrt(sum);
                } else {
                    A[(size_t)i * N + j] = sum / A[(size_t)j * N + j];
                }
            }
            for (int j = i + 1; j < N; j++) A[(size_t)i * N + j] = 0.0;
        }
        return 1;
    }
    for (int j = 0; j < N; j++) {
        double sum = A[(size_t)j * N + j];
        for (int k = 0; k < j; k++) sum -= A[(size_t)j * N + k] * A[(size_t)j * N + k];
        if (sum <= 0.0) return 0;
        A[(size_t)j * N + j] = sqrt(sum);
        double invDiag = 1.0 / A[(size_t)j * N + j];
        for (int i = j + 1; i < N; i++) {
            double s2 = A[(size_t)i * N + j];
            for (int k = 0; k < j; k++) s2 -= A[(size_t)i * N + k] * A[(size_t)j * N + k];
            A[(size_t)i * N + j] = s2 * invDiag;
        }
    }
    for (int i = 0; i < N; i++)
        for (int j = i + 1; j < N; j++) A[(size_t)i * N + j] = 0.0;
    return 1;
}

/* forward / saxpy-backward substitution (coarsest.cpp:92-122) */
void o_chol_solve(const double* L, const double* b, double* x, int N)
{
    for (int i = 0; i < N; ++i) x[i] = b[i];
    for (int i = 0; i < N; i++) {
        double sum = x[i];
        for (int k = 0; k < i; k++) sum -= L[(size_t)i * N + k] * x[k];
        x[i] = sum / L[(size_t)i * N + i];
    }
    for (int i = N - 1; i >= 0; i--) {
        x[i] /= L[(size_t)i * N + i];
        const double xi = x[i];
        for (int k = 0; k < i; k++) x[k] -= L[(size_t)i * N + k] * xi;
    }
}

/* make_diagonal_preconditioner_from_static (preconditioner.cpp:38-79) */
void o_mg_setup(const o_model* m, o_mg* g, const double* E)
{
    /* finest diagonal by colour (diagonal.cpp:11-63) */
    double* D0 = g->diag[0];
    const int n0 = 3 * g->nN[0];
    for (int i = 0; i < n0; ++i) D0[i] = 0.0;
    for (int col = 0; col < 8; ++col) {
        const int b0 = m->bucketStart[col], b1 = m->bucketStart[col + 1];
#pragma omp parallel for schedule(static)
        for (int bi = b0; bi < b1; ++bi) {
            int e = m->bucket[bi];
            double Ee = E[e];
            for (int a = 0; a < 8; ++a) {
                int n = m->eNodMat[8 * (size_t)e + a];
                for (int d = 0; d < 3; ++d) {
                    int l = 3 * a + d;
                    D0[3 * n + d] += m->Ke[l * 24 + l] * Ee;
                }
            }
        }
    }
    for (int i = 0; i < n0; ++i) if (!(D0[i] > 0.0)) D0[i] = 1.0;
    for (int i = 0; i < n0; ++i) if (g->mask[0][i]) D0[i] = 1.0;
    for (int l = 1; l < g->nlev; ++l) coarsen_diag(g, l);

    free(g->Lc); g->Lc = NULL;
    free(g->Kc); g->Kc = NULL;
    g->Nc = 0;
    g->cholOk = 1;
    const int L = g->nlev - 1;
    int N = 3 * g->nN[L];
    if (g->nlev == 1 || N > g->NlimitDofs) return;
    double* Kc = XALLOC(double, (size_t)N * N);
    assemble_coarsest(m, g, E, Kc, N);
    g->Kc = XALLOC(double, (size_t)N * N);
    memcpy(g->Kc, Kc, sizeof(double) * (size_t)N * N);
    if (o_chol_inplace(Kc, N)) { g->Lc = Kc; g->Nc = N; }
    else { free(Kc); g->cholOk = 0; }
}

/* V-cycle (preconditioner.cpp:88-265) on the full layout */
void o_mg_apply_full(const o_model* m, o_mg* g, const double* rin, double* z)
{
    (void)m;
    const int L = g->nlev - 1;
    const double w = g->weight;
    {
        const int n0 = 3 * g->nN[0];
        for (int i = 0; i < n0; ++i) g->r[0][i] = g->mask[0][i] ? 0.0 : rin[i];
    }
    for (int l = 0; l < L; ++l) {
        o_restrict_vec3(g, l + 1, g->r[l], g->r[l + 1]);
        const int nc = 3 * g->nN[l + 1];
        for (int i = 0; i < nc; ++i) if (g->mask[l + 1][i]) g->r[l + 1][i] = 0.0;
    }
    {
        const int nc = 3 * g->nN[L];
        if (g->Lc && g->Nc == nc) {
            o_chol_solve(g->Lc, g->r[L], g->x[L], g->Nc);
        } else {
            for (int d = 0; d < nc; ++d)
                g->x[L][d] = g->mask[L][d] ? 0.0 : g->r[L][d] / dmax(1.0e-30, g->diag[L][d]);
        }
        for (int d = 0; d < nc; ++d) if (g->mask[L][d]) g->x[L][d] = 0.0;
    }
    for (int l = L - 1; l >= 0; --l) {
        o_prolong_vec3(g, l + 1, g->x[l + 1], g->tmp);
        const int nf = 3 * g->nN[l];
        double* xo = (l == 0) ? z : g->x[l];
        for (int d = 0; d < nf; ++d) {
            double t = g->mask[l][d] ? 0.0 : w * g->r[l][d] / dmax(1.0e-30, g->diag[l][d]);
            double xv = t + g->tmp[d];
            xo[d] = g->mask[l][d] ? 0.0 : xv + t;
        }
    }
    if (L == 0) {
        const int n0 = 3 * g->nN[0];
        for (int d = 0; d < n0; ++d) z[d] = g->x[0][d];
    }
}

/* ===================================================================== */
/* PCG (solver.cpp:87-217)                                                */
/* ===================================================================== */
int o_pcg_full(const o_model* m, const double* E, const double* b, double* x,
               double tol, int maxIt, o_precond_fn M, void* Mctx)
{
    const size_t n = (size_t)m->nDOF;
    double* r = XALLOC(double, n);
    double* z = XALLOC(double, n);
    double* p = XALLOC(double, n);
    double* Ap = XALLOC(double, n);
    for (size_t i = 0; i < n; ++i) r[i] = b[i];
    o_kxu_full(m, E, x, Ap);
    for (size_t i = 0; i < n; ++i) r[i] -= Ap[i];
    const double* zp = r;
    if (M) { M(Mctx, r, z); zp = z; }
    double rz_old = M ? o_canon_dot(r, z, n) : o_canon_dot(r, r, n);
    for (size_t i = 0; i < n; ++i) p[i] = zp[i];
    const double normb = sqrt(o_canon_dot(b, b, n));
    const double stop = dmax(1.0e-30, normb) * tol;
    int ret = maxIt;
    for (int it = 0; it < maxIt; ++it) {
        o_kxu_full(m, E, p, Ap);
        double denom = o_canon_dot(p, Ap, n);
        double alpha = rz_old / dmax(1.0e-30, denom);
        for (size_t i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] = r[i] - alpha * Ap[i];
        }
        double rn2 = o_canon_dot(r, r, n);
        if (sqrt(rn2) < stop) { ret = it + 1; break; }
        if (M) {
            M(Mctx, r, z);
            double rz_new = o_canon_dot(r, z, n);
            double beta = rz_new / dmax(1.0e-30, rz_old);
            rz_old = rz_new;
            for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        } else {
            double beta = rn2 / dmax(1.0e-30, rz_old);
            rz_old = rn2;
            for (size_t i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
        }
    }
    free(r); free(z); free(p); free(Ap);
    return ret;
}

typedef struct { const o_model* m; o_mg* g; } mg_ctx;
static void mg_cb(void* c, const double* r, double* z)
{
    mg_ctx* mc = (mg_ctx*)c;
    o_mg_apply_full(mc->m, mc->g, r, z);
}
int o_pcg_mg(const o_model* m, o_mg* g, const double* E, const double* b, double* x,
             double tol, int maxIt)
{
    mg_ctx c = {m, g};
    return o_pcg_full(m, E, b, x, tol, maxIt, g ? mg_cb : NULL, &c);
}

/* ===================================================================== */
/* PDE filter (filter.cpp:9-275)                                          */
/* ===================================================================== */
o_filter* o_filter_new(const o_model* m, float rMin)
{
    o_filter* f = (o_filter*)calloc(1, sizeof(o_filter));
    const double s[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    const double t[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    const double p[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    double N[8][8], dS[24][8];
    for (int gp = 0; gp < 8; ++gp) {
        double sg = s[gp] / sqrt(3.0), tg = t[gp] / sqrt(3.0), pg = p[gp] / sqrt(3.0);
        double Ns[8];
        shape8(sg, tg, pg, Ns);
        for (int a = 0; a < 8; a++) N[gp][a] = Ns[a];
        double dNds[8] = {-0.125 * (1 - tg) * (1 - pg), 0.125 * (1 - tg) * (1 - pg),
                          0.125 * (1 + tg) * (1 - pg), -0.125 * (1 + tg) * (1 - pg),
                          -0.125 * (1 - tg) * (1 + pg), 0.125 * (1 - tg) * (1 + pg),
                          0.125 * (1 + tg) * (1 + pg), -0.125 * (1 + tg) * (1 + pg)};
        double dNdt[8] = {-0.125 * (1 - sg) * (1 - pg), -0.125 * (1 + sg) * (1 - pg),
                          0.125 * (1 + sg) * (1 - pg), 0.125 * (1 - sg) * (1 - pg),
                          -0.125 * (1 - sg) * (1 + pg), -0.125 * (1 + sg) * (1 + pg),
                          0.125 * (1 + sg) * (1 + pg), 0.125 * (1 - sg) * (1 + pg)};
        double dNdp[8] = {-0.125 * (1 - sg) * (1 - tg), -0.125 * (1 + sg) * (1 - tg),
                          -0.125 * (1 + sg) * (1 + tg), -0.125 * (1 - sg) * (1 + tg),
                          0.125 * (1 - sg) * (1 - tg), 0.125 * (1 + sg) * (1 - tg),
                          0.125 * (1 + sg) * (1 + tg), 0.125 * (1 - sg) * (1 + tg)};
        for (int a = 0; a < 8; a++) { dS[3 * gp][a] = dNds[a]; dS[3 * gp + 1][a] = dNdt[a]; dS[3 * gp + 2][a] = dNdp[a]; }
    }
    double KEF0[64], KEF1[64];
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) {
            double sum = 0.0;
            for (int k = 0; k < 24; k++) sum += dS[k][i] * dS[k][j];
            KEF0[i * 8 + j] = sum;
        }
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) {
            double sum = 0.0;
            for (int gp = 0; gp < 8; gp++) sum += N[gp][i] * (1.0 / 8.0) * N[gp][j];
            KEF1[i * 8 + j] = sum;
        }
    double iRmin = ((double)rMin * (double)1.0f) / (2.0 * sqrt(3.0));
    for (int i = 0; i < 64; i++) f->kernel[i] = iRmin * iRmin * KEF0[i] + KEF1[i];
    double rowSum[8];
    for (int a = 0; a < 8; a++) {
        double ss = 0.0;
        for (int b = 0; b < 8; b++) ss += f->kernel[a * 8 + b];
        rowSum[a] = ss;
    }
    f->diagPrecond = XALLOC(double, m->nN);
    for (int e = 0; e < m->ne; e++)
        for (int a = 0; a < 8; a++) f->diagPrecond[m->eNodMat[8 * (size_t)e + a]] += rowSum[a];
    for (int n = 0; n < m->nN; ++n) f->diagPrecond[n] = f->diagPrecond[n] > 0 ? 1.0 / f->diagPrecond[n] : 1.0;
    f->lastX = XALLOC(double, m->nN);
    f->lastRhs = XALLOC(double, m->nN);
    return f;
}

void o_filter_free(o_filter* f)
{
    if (!f) return;
    free(f->diagPrecond); free(f->lastX); free(f->lastRhs);
    free(f);
}

/* MatTimesVec_PDE (filter.cpp:77-173): colour order, sequential 8-term rows */
void o_pde_matvec(const o_model* m, const o_filter* f, const double* x, double* y)
{
    memset(y, 0, sizeof(double) * (size_t)m->nN);
    for (int col = 0; col < 8; ++col) {
        const int b0 = m->bucketStart[col], b1 = m->bucketStart[col + 1];
#pragma omp parallel for schedule(static)
        for (int bi = b0; bi < b1; ++bi) {
            int e = m->bucket[bi];
            const int32_t* en = m->eNodMat + 8 * (size_t)e;
            double u[8], fo[8];
            for (int a = 0; a < 8; ++a) u[a] = x[en[a]];
            for (int k = 0; k < 8; ++k) {
                double s = 0.0;
                for (int j = 0; j < 8; ++j) s += f->kernel[k * 8 + j] * u[j];
                fo[k] = s;
            }
            for (int a = 0; a < 8; ++a) y[en[a]] += fo[a];
        }
    }
}

void o_filter_apply(const o_model* m, o_filter* f, const float* src, float* dst)
{
    const int nN = m->nN, ne = m->ne;
    double* rhs = XALLOC(double, nN);
    double* x = XALLOC(double, nN);
    double* r = XALLOC(double, nN);
    double* z = XALLOC(double, nN);
    double* p = XALLOC(double, nN);
    double* Ap = XALLOC(double, nN);
    double* d = XALLOC(double, nN);
    for (int e = 0; e < ne; e++) {
        double val = (double)src[e] * (1.0 / 8.0);
        for (int a = 0; a < 8; a++) rhs[m->eNodMat[8 * (size_t)e + a]] += val;
    }
    for (int i = 0; i < nN; ++i) d[i] = rhs[i] - f->lastRhs[i];
    double normRhs = sqrt(o_canon_dot(rhs, rhs, (size_t)nN));
    double diff = sqrt(o_canon_dot(d, d, (size_t)nN));
    int useWarm = (normRhs > 0.0) && (diff / dmax(1.0e-30, normRhs) < 1000.0);
    f->lastWarm = useWarm;
    if (useWarm) {
        memcpy(x, f->lastX, sizeof(double) * (size_t)nN);
        o_pde_matvec(m, f, x, Ap);
        for (int i = 0; i < nN; ++i) r[i] = rhs[i] - Ap[i];
    } else {
        for (int i = 0; i < nN; ++i) r[i] = rhs[i];
    }
    for (int i = 0; i < nN; ++i) z[i] = f->diagPrecond[i] * r[i];
    double rz = o_canon_dot(r, z, (size_t)nN);
    if (rz == 0) rz = 1.0;
    memcpy(p, z, sizeof(double) * (size_t)nN);
    int iters = -1;
    double rn = sqrt(o_canon_dot(r, r, (size_t)nN));
    if (!(rn < 1.0e-2)) {
        iters = 0;
        for (int it = 0; it < 80; ++it) {
            o_pde_matvec(m, f, p, Ap);
            double denom = o_canon_dot(p, Ap, (size_t)nN);
            double alpha = rz / dmax(1.0e-30, denom);
            for (int i = 0; i < nN; ++i) x[i] += alpha * p[i];
            for (int i = 0; i < nN; ++i) r[i] -= alpha * Ap[i];
            iters = it + 1;
            double rn2 = sqrt(o_canon_dot(r, r, (size_t)nN));
            if (rn2 < 1e-3) break;
            for (int i = 0; i < nN; ++i) z[i] = f->diagPrecond[i] * r[i];
            double rz_new = o_canon_dot(r, z, (size_t)nN);
            double beta = rz_new / dmax(1.0e-30, rz);
            for (int i = 0; i < nN; ++i) p[i] = z[i] + beta * p[i];
            rz = rz_new;
        }
    }
    f->lastIters = iters;
    for (int e = 0; e < ne; e++) {
        double sum = 0.0;
        for (int a = 0; a < 8; a++) sum += x[m->eNodMat[8 * (size_t)e + a]];
        dst[e] = (float)(sum * (1.0 / 8.0));
    }
    memcpy(f->lastX, x, sizeof(double) * (size_t)nN);
    memcpy(f->lastRhs, rhs, sizeof(double) * (size_t)nN);
    free(rhs); free(x); free(r); free(z); free(p); free(Ap); free(d);
}

/* ===================================================================== */
/* SIMP / sensitivity / OC (driver.cpp:117-189)                           */
/* ===================================================================== */
static float next_toward(float f, int up)
{
    /* neighbouring float of a finite f >= 0 */
    uint32_t b;
    memcpy(&b, &f, 4);
    if (up) b = b + 1u;
    else b = (b == 0u) ? 0x80000001u : b - 1u;
    float r;
    memcpy(&r, &b, 4);
    return r;
}
/* round hi+lo (|lo| <= ulp(hi)/2, hi = RN(hi+lo)) once to binary32 */
static float round_dd_to_float(double hi, double lo)
{
    float f = (float)hi;
    double df = (double)f;
    if (lo == 0.0 || df == hi) return f;
    float g = next_toward(f, hi > df);
    double mid = 0.5 * (df + (double)g);
    if (hi != mid) return f;
    if ((lo > 0.0) == ((double)g > df)) return g;
    return f;
}
float o_cr_cube(float x)
{
    double d = (double)x;
    double p2 = d * d;            /* exact: 48 significant bits */
    double hi = p2 * d;
    double lo = fma(p2, d, -hi);  /* exact remainder */
    return round_dd_to_float(hi, lo);
}
float o_cr_square(float x)
{
    double d = (double)x;
    return (float)(d * d);        /* exact product, one rounding */
}
float o_seq_sum(const float* v, int n)
{
    float s = 0.0f;
    for (int i = 0; i < n; ++i) s = s + v[i];
    return s;
}
float o_seq_sum_x1mx(const float* v, int n)
{
    float s = 0.0f;
    for (int i = 0; i < n; ++i) s = s + v[i] * (1.0f - v[i]);
    return s;
}
void o_simp(int ne, const float* xPhys, float E0, float Emin, double* E)
{
    for (int e = 0; e < ne; ++e) {
        float rho = fclampr(xPhys[e], 0.0f, 1.0f);
        float powp = o_cr_cube(rho);
        float v = Emin + powp * (E0 - Emin);
        E[e] = (double)v;
    }
}
void o_sensitivity(int ne, const float* xPhys, const double* ce, double Csolid,
                   float E0, float Emin, float penal, float* dc)
{
    for (int e = 0; e < ne; ++e) {
        float rho = fclampr(xPhys[e], 0.0f, 1.0f);
        float dEdrho = penal * o_cr_square(rho) * (E0 - Emin);
        float ceNorm = (Csolid > 0) ? (float)(ce[e] / Csolid) : (float)ce[e];
        dc[e] = -dEdrho * ceNorm;
    }
}
int o_oc_update(int ne, const float* x, const float* dc, float V0, float move,
                float* xnew, float* change)
{
    return o_oc_update_passive(ne, x, dc, NULL, V0, move, xnew, change);
}

/* passive[e] != 0: element e is held at density 1 inside the bisection, as
 * matlab/TOP3D_XL.m:154 (xnew(passiveElements) = 1.0), so the volume of every
 * trial counts it as 1; the C++ port declares but never uses passive layers
 * (core.hpp:39-41).  passive == NULL: the C++ port's OC. */
int o_oc_update_passive(int ne, const float* x, const float* dc, const uint8_t* passive, float V0, float move,
                        float* xnew, float* change)
{
    float l1 = 0.0f, l2 = 1e9f;
    int steps = 0;
    while ((double)((l2 - l1) / (l1 + l2)) > 1e-6) {
        float lmid = 0.5f * (l1 + l2);
        for (int e = 0; e < ne; ++e) {
            if (passive && passive[e]) {
                xnew[e] = 1.0f;
                continue;
            }
            float val = sqrtf(fmaxr(1e-30f, -dc[e] / lmid));
            float xe = fclampr(x[e] * val, x[e] - move, x[e] + move);
            xe = fclampr(xe, 0.0f, 1.0f);
            xnew[e] = fmaxr(1.0e-3f, xe);
        }
        float vol = o_seq_sum(xnew, ne) / (float)ne;
        if (vol - V0 > 0) l1 = lmid; else l2 = lmid;
        ++steps;
    }
    float ch = 0.0f;
    for (int e = 0; e < ne; ++e) ch = fmaxr(ch, fabsf(xnew[e] - x[e]));
    *change = ch;
    return steps;
}

/* ===================================================================== */
/* driver (driver.cpp:23-218)                                             */
/* ===================================================================== */
static uint64_t splitmix64(uint64_t z)
{
    /* finaliser of Steele, Lea & Flood's SplitMix64 (OOPSLA 2014) */
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t o_x_check(const float* x, int64_t n)
{
    uint64_t h = 0;
    for (int64_t i = 0; i < n; ++i) {
        uint32_t b;
        memcpy(&b, &x[i], 4);
        h += splitmix64(((uint64_t)i << 32) | b);
    }
    return h;
}

void o_params_default(o_params* p)
{
    /* core.hpp:26-42, driver.cpp:54, top3d_cli.cpp:33 */
    p->E0 = 1.0f; p->Emin = 1.0e-6f; p->nu = 0.3f; p->penal = 3.0f; p->cgTol = 1.0e-3f;
    p->cgMaxIt = 800; p->V0 = 0.12f; p->nLoop = 30; p->rMin = sqrtf(3.0f);
    p->mgMaxLevels = 5; p->mgWeight = 0.6f; p->nonDyadic = 1; p->NlimitDofs = 1000;
    p->rebuildEvery = 5;
}

int o_top3d_run(int nely, int nelx, int nelz, const o_params* P,
                o_loop_rec* rec, float* x_out, double* CsolidOut, int* solidIters)
{
    return o_top3d_run_ex(nely, nelx, nelz, NULL, NULL, P, rec, x_out, CsolidOut, solidIters);
}

/* passive mask of a voxel-order mask [(ez*nelx + ex)*nely + ey] in compact element order */
static uint8_t* passive_compact(const o_model* m, const uint8_t* passive)
{
    if (!passive) return NULL;
    uint8_t* pc = XALLOC(uint8_t, m->ne);
    const int ny = m->resY, nx = m->resX;
    for (int e = 0; e < m->ne; ++e) {
        const int full = m->eleMapBack[e];
        const int ez = full / (nx * ny), rem = full % (nx * ny);
        const int ex = rem / ny, ey = ny - 1 - rem % ny;
        pc[e] = passive[((size_t)ez * m->nelx + ex) * m->nely + ey] ? 1 : 0;
    }
    return pc;
}

static int run_model(o_model* m, const uint8_t* passive, const o_params* P, o_loop_rec* rec, float* x_out,
                     double* CsolidOut, int* solidIters);

int o_top3d_run_ex(int nely, int nelx, int nelz, const uint8_t* solid, const uint8_t* passive, const o_params* P,
                   o_loop_rec* rec, float* x_out, double* CsolidOut, int* solidIters)
{
    o_model* m = solid ? o_model_new_voxels(nely, nelx, nelz, solid) : o_model_new(nely, nelx, nelz);
    return run_model(m, passive, P, rec, x_out, CsolidOut, solidIters);
}

int o_top3d_run_bc(int nely, int nelx, int nelz, const uint8_t* solid, const uint8_t* passive, int nFix,
                   const int32_t* fix, int nLoad, const int32_t* loadNode, const double* loadF, const o_params* P,
                   o_loop_rec* rec, float* x_out, double* CsolidOut, int* solidIters)
{
    o_model* m = o_model_new_bc(nely, nelx, nelz, solid, nFix, fix, nLoad, loadNode, loadF);
    if (!m) return -1;
    return run_model(m, passive, P, rec, x_out, CsolidOut, solidIters);
}

/* the loop on a built model (takes ownership of m) */
static int run_model(o_model* m, const uint8_t* passive, const o_params* P, o_loop_rec* rec, float* x_out,
                     double* CsolidOut, int* solidIters)
{
    uint8_t* pc = passive_compact(m, passive);
    o_filter* flt = o_filter_new(m, P->rMin);
    o_mg* g = o_mg_new(m, P->nonDyadic, P->mgMaxLevels, P->mgWeight, P->Emin, P->NlimitDofs);
    const int ne = m->ne;
    const size_t nD = (size_t)m->nDOF;
    float* x = XALLOC(float, ne);
    float* xnew = XALLOC(float, ne);
    float* dc = XALLOC(float, ne);
    float* xdc = XALLOC(float, ne);
    float* dcf = XALLOC(float, ne);
    double* ce = XALLOC(double, ne);
    double* E = XALLOC(double, ne);
    double* u = XALLOC(double, nD);
    double* b = XALLOC(double, nD);
    /* bFree = restrict_to_free(F) (driver.cpp:75): loads on fixed DOFs are dropped */
    for (size_t i = 0; i < nD; ++i) b[i] = m->isFree[i] ? m->F[i] : 0.0;
    for (int e = 0; e < ne; ++e) { x[e] = P->V0; E[e] = (double)P->E0; }
    const double tol = (double)P->cgTol;
    /* solid reference solve (driver.cpp:85-103) */
    o_mg_setup(m, g, E);
    int it0 = o_pcg_mg(m, g, E, b, u, tol, P->cgMaxIt);
    double Csolid = o_compliance(m, E, u, ce);
    if (CsolidOut) *CsolidOut = Csolid;
    if (solidIters) *solidIters = it0;
    const int progress = getenv("O_LOOP_PROGRESS") != NULL;   /* one stderr line per solve */
    if (progress) fprintf(stderr, "oracle loop: solid solve %d iterations\n", it0);
    int loop = 0;
    float change = 1.0f, sharp = 1.0f;
    double* items = XALLOC(double, ne);
    while (loop < P->nLoop && (double)change > 1e-4 && (double)sharp > 0.01) {
        ++loop;
        o_simp(ne, x, P->E0, P->Emin, E);
        if (loop == 1 || (loop % P->rebuildEvery) == 0) o_mg_setup(m, g, E);
        int its = o_pcg_mg(m, g, E, b, u, tol, P->cgMaxIt);
        double C = o_compliance(m, E, u, ce);
        for (int e = 0; e < ne; ++e) items[e] = (Csolid > 0) ? E[e] * (ce[e] / Csolid) : E[e] * ce[e];
        double cObj = o_canon_sum(items, (size_t)ne);
        double Cdisp = (Csolid > 0) ? cObj * Csolid : C;
        o_sensitivity(ne, x, ce, Csolid, P->E0, P->Emin, P->penal, dc);
        for (int e = 0; e < ne; ++e) xdc[e] = x[e] * dc[e];
        o_filter_apply(m, flt, xdc, dcf);
        for (int e = 0; e < ne; ++e) dc[e] = dcf[e] / fmaxr(1.0e-3f, x[e]);
        const int bsteps = o_oc_update_passive(ne, x, dc, pc, P->V0, 0.2f, xnew, &change);
        memcpy(x, xnew, sizeof(float) * (size_t)ne);
        sharp = 4.0f * o_seq_sum_x1mx(x, ne) / (float)ne;
        float volFrac = o_seq_sum(x, ne) / (float)ne;
        if (rec) {
            o_loop_rec* r = &rec[loop - 1];
            r->loop = loop; r->pcgIters = its; r->filterIters = flt->lastIters;
            r->C = C; r->Cdisp = Cdisp; r->volFrac = volFrac; r->sharpness = sharp;
            r->change = change; r->fval = volFrac - P->V0;
            r->bisectSteps = bsteps; r->xCheck = o_x_check(x, ne);
        }
        if (progress) fprintf(stderr, "oracle loop %d: %d PCG iterations, C %.17g\n", loop, its, C);
    }
    if (x_out) memcpy(x_out, x, sizeof(float) * (size_t)ne);
    free(items); free(b); free(x); free(xnew); free(dc); free(xdc); free(dcf); free(ce); free(E); free(u); free(pc);
    o_mg_free(g); o_filter_free(flt); o_model_free(m);
    return loop;
}

/* ===================================================================== */
/* voxel-surface STL (io.cpp:80-87 export_surface_stl,                    */
/* geometry_out/voxel_surface.hpp:12-93 extract_faces,                    */
/* geometry_out/export_stl.hpp:14-33 write_stl_binary)                    */
/* ===================================================================== */
/* Face d of voxel (y, x, z), d = 0..5 for -X, +X, -Y, +Y, -Z, +Z
 * (voxel_surface.hpp:52-88): present when the voxel is inside and its
 * neighbour across the face is outside or beyond the volume. */
static int stl_face_present(int ny, int nx, int nz, const float* vol, float iso, int y, int x, int z, int d)
{
    const size_t pl = (size_t)ny * nx;
    const size_t v = (size_t)y + (size_t)ny * x + pl * z;
    if (!(vol[v] >= iso)) return 0;
    const int axis = d >> 1, up = d & 1;
    const int c = axis == 0 ? x : axis == 1 ? y : z;
    const int n = axis == 0 ? nx : axis == 1 ? ny : nz;
    const size_t step = axis == 0 ? (size_t)ny : axis == 1 ? 1 : pl;
    if (up ? c == n - 1 : c == 0) return 1;
    const size_t w = up ? v + step : v - step;
    return !(vol[w] >= iso);
}

static uint8_t* stl_put_tri(uint8_t* p, const float* a, const float* b, const float* c)
{
    memset(p, 0, 12);                  /* normal (0, 0, 0) */
    memcpy(p + 12, a, 12);             /* little-endian host: the bytes of the floats */
    memcpy(p + 24, b, 12);
    memcpy(p + 36, c, 12);
    p[48] = 0; p[49] = 0;              /* attribute */
    return p + 50;
}

uint64_t o_stl_image(int ny, int nx, int nz, const float* vol, float iso, uint8_t* out, uint64_t cap)
{
    uint64_t faces = 0;
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y)
                for (int d = 0; d < 6; ++d) faces += (uint64_t)stl_face_present(ny, nx, nz, vol, iso, y, x, z, d);
    const uint64_t bytes = 84 + 100 * faces;
    if (!out || cap < bytes) return bytes;
    memset(out, 0, 80);
    memcpy(out, "Generated by TOP3D_XL", 21);
    const uint32_t ntri = (uint32_t)(2 * faces);
    out[80] = (uint8_t)(ntri & 0xffu); out[81] = (uint8_t)((ntri >> 8) & 0xffu);
    out[82] = (uint8_t)((ntri >> 16) & 0xffu); out[83] = (uint8_t)(ntri >> 24);
    uint8_t* p = out + 84;
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y)
                for (int d = 0; d < 6; ++d) {
                    if (!stl_face_present(ny, nx, nz, vol, iso, y, x, z, d)) continue;
                    /* the face's corner with the smallest coordinates, vertex order (x, y, z)
                     * (voxel_surface.hpp:19-45), and its two in-plane unit steps a, b:
                     * x-faces (y, z), y-faces (x, z), z-faces (x, y) */
                    const int axis = d >> 1, up = d & 1;
                    const float o[3] = {(float)(x + (axis == 0 && up)), (float)(y + (axis == 1 && up)),
                                        (float)(z + (axis == 2 && up))};
                    const int ia = axis == 0 ? 1 : 0, ib = axis == 2 ? 1 : 2;
                    float c0[3], c1[3], c2[3], c3[3];
                    for (int k = 0; k < 3; ++k) c0[k] = c1[k] = c2[k] = c3[k] = o[k];
                    c1[ia] += 1.0f;
                    c2[ia] += 1.0f; c2[ib] += 1.0f;
                    c3[ib] += 1.0f;
                    p = stl_put_tri(p, c0, c1, c2);
                    p = stl_put_tri(p, c0, c2, c3);
                }
    return bytes;
}
