// mg_host.cpp -- host side of the multigrid preconditioner (see mg_host.hpp).
#include "mg_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace t3d {

namespace {

// corner a of a structured cell (x, ty, z) sits at node offset (DX, DY, DZ)
const int DX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
const int DY[8] = {1, 1, 0, 0, 1, 1, 0, 0};
const int DZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};

// trilinear shape values at the (span+1)^3 embedded vertices; natural coords
// xi = -1 + 2*(i/span) (hierarchy.cpp:13-42)
std::vector<double> weights_table(int span)
{
    const int g = span + 1;
    std::vector<double> W(size_t(g) * g * g * 8);
    for (int iz = 0; iz <= span; ++iz)
        for (int iy = 0; iy <= span; ++iy)
            for (int ix = 0; ix <= span; ++ix) {
                const double xi = -1.0 + 2.0 * (double(ix) / double(span));
                const double et = -1.0 + 2.0 * (double(iy) / double(span));
                const double ze = -1.0 + 2.0 * (double(iz) / double(span));
                const double mx = 1.0 - xi, px = 1.0 + xi, my = 1.0 - et, py = 1.0 + et;
                const double mz = 1.0 - ze, pz = 1.0 + ze;
                double* w = &W[size_t((iz * g + iy) * g + ix) * 8];
                w[0] = 0.125 * mx * my * mz; w[1] = 0.125 * px * my * mz;
                w[2] = 0.125 * px * py * mz; w[3] = 0.125 * mx * py * mz;
                w[4] = 0.125 * mx * my * pz; w[5] = 0.125 * px * my * pz;
                w[6] = 0.125 * px * py * pz; w[7] = 0.125 * mx * py * pz;
            }
    return W;
}

// scalar hat-stencil restriction with a final division (transfers.cpp:95-191);
// inputs are 0/1 masks so every partial sum is an exact dyadic
void restrict_scalar(const MGLevelDims& Lf, const MGLevelDims& Lc, const std::vector<double>& rf,
                     std::vector<double>& rc)
{
    const int span = Lc.span, wSize = 2 * span - 1;
    const double inv = 1.0 / span;
    double W[16];
    for (int i = 0; i < wSize; ++i) W[i] = 1.0 - std::abs(i - (span - 1)) * inv;
    const int fx = Lf.resX + 1, fy = Lf.resY + 1, fz = Lf.resZ + 1;
    const int cx = Lc.resX + 1, cy = Lc.resY + 1, cz = Lc.resZ + 1;
    rc.assign(size_t(cx) * cy * cz, 0.0);
    auto range = [&](int f0, int n, int& lo, int& hi) {
        lo = (f0 - span + 1 < 0) ? (span - 1 - f0) : 0;
        hi = (f0 + span - 1 >= n) ? (wSize - (f0 + span - n)) : wSize;
    };
    for (int k = 0; k < cz; ++k)
        for (int j = 0; j < cy; ++j)
            for (int i = 0; i < cx; ++i) {
                int z0, z1, y0, y1, x0, x1;
                range(k * span, fz, z0, z1);
                range(j * span, fy, y0, y1);
                range(i * span, fx, x0, x1);
                double vs = 0.0, ws = 0.0;
                for (int a = z0; a < z1; ++a) {
                    const size_t zo = size_t(k * span + a - (span - 1)) * fy * fx;
                    for (int b = y0; b < y1; ++b) {
                        const double wzy = W[a] * W[b];
                        const size_t yo = size_t(j * span + b - (span - 1)) * fx;
                        for (int c = x0; c < x1; ++c) {
                            const double w = wzy * W[c];
                            vs += rf[zo + yo + size_t(i * span + c - (span - 1))] * w;
                            ws += w;
                        }
                    }
                }
                rc[(size_t(k) * cy + j) * cx + i] = (ws > 1e-12) ? vs / ws : 0.0;
            }
}

// M = A*T, out (+)= T^T*M with sequential k sums (coarsest.cpp:193-212, 289-303)
void tt_a_t(const double* A, const double* T, double* out, bool accumulate)
{
    double M[576];
    for (int i = 0; i < 24; ++i)
        for (int j = 0; j < 24; ++j) {
            double s = 0.0;
            for (int k = 0; k < 24; ++k) s += A[i * 24 + k] * T[k * 24 + j];
            M[i * 24 + j] = s;
        }
    for (int i = 0; i < 24; ++i)
        for (int j = 0; j < 24; ++j) {
            double s = 0.0;
            for (int k = 0; k < 24; ++k) s += T[k * 24 + i] * M[k * 24 + j];
            out[i * 24 + j] = accumulate ? out[i * 24 + j] + s : s;
        }
}

// interpolation matrix of fine sub-cell (ix,iy,iz) inside a coarse cell of span s
void sub_T(const std::vector<double>& W, int s, int ix, int iy, int iz, double* T)
{
    const int g = s + 1;
    std::memset(T, 0, sizeof(double) * 576);
    for (int v = 0; v < 8; ++v) {
        const int vx = DX[v], vy = 1 - DY[v], vz = DZ[v];   // natural-coordinate corner bits
        const double* w = &W[size_t(((iz + vz) * g + (iy + vy)) * g + (ix + vx)) * 8];
        for (int a = 0; a < 8; ++a)
            for (int c = 0; c < 3; ++c) T[(3 * v + c) * 24 + 3 * a + c] = w[a];
    }
}

}  // namespace

void mg_build_static(const HostModel& m, bool nonDyadic, int maxLevels, float weight, float Emin, int NlimitDofs,
                     HostMG& H)
{
    H.weight = double(weight);
    H.Emin = double(Emin);
    H.NlimitDofs = NlimitDofs;
    H.lv.clear();
    H.W.clear();
    MGLevelDims L0;
    L0.resX = m.resX; L0.resY = m.resY; L0.resZ = m.resZ; L0.span = 1;
    L0.nN = int64_t(m.resX + 1) * (m.resY + 1) * (m.resZ + 1);
    H.lv.push_back(L0);
    H.W.emplace_back();
    int nx = m.resX, ny = m.resY, nz = m.resZ;
    for (int li = 1; li < maxLevels; ++li) {
        const int span = (li == 1 && nonDyadic) ? 4 : 2;
        if (nx % span || ny % span || nz % span) break;
        if (nx / span < 1 || ny / span < 1 || nz / span < 1) break;
        nx /= span; ny /= span; nz /= span;
        MGLevelDims L;
        L.resX = nx; L.resY = ny; L.resZ = nz; L.span = span;
        L.nN = int64_t(nx + 1) * (ny + 1) * (nz + 1);
        H.lv.push_back(L);
        H.W.push_back(weights_table(span));
        if (std::min({nx, ny, nz}) <= 2) break;
    }
    const size_t nl = H.lv.size();
    H.masks.assign(nl, {});
    H.masks[0].resize(size_t(3) * H.lv[0].nN);
    for (size_t d = 0; d < H.masks[0].size(); ++d) H.masks[0][d] = m.isFreeDOF[d] ? 0 : 1;
    for (size_t l = 0; l + 1 < nl; ++l) {
        const int64_t fn = H.lv[l].nN, cn = H.lv[l + 1].nN;
        H.masks[l + 1].assign(size_t(3) * cn, 0);
        std::vector<double> rf(fn), rc;
        for (int c = 0; c < 3; ++c) {
            for (int64_t n = 0; n < fn; ++n) rf[n] = H.masks[l][3 * n + c] ? 1.0 : 0.0;
            restrict_scalar(H.lv[l], H.lv[l + 1], rf, rc);
            for (int64_t n = 0; n < cn; ++n) H.masks[l + 1][3 * n + c] = (rc[n] > 0.5) ? 1 : 0;
        }
    }
}

int mg_assemble_coarsest(const HostModel& m, const HostMG& H, const double* E, std::vector<double>& Kc)
{
    const size_t L = H.lv.size() - 1;
    const MGLevelDims& Lf = H.lv[0];
    const MGLevelDims& Lc = H.lv[L];
    const int N = int(3 * Lc.nN);
    const int s = Lc.span;
    const std::vector<double>& W = H.W[L];
    const int fX = Lf.resX, fY = Lf.resY;
    const int fnx = fX + 1, fny = fY + 1;
    const int cnx = Lc.resX + 1, cny = Lc.resY + 1;
    const std::vector<uint8_t>& fm = H.masks[0];
    // moduli in the reference's full (y-flipped) element order; void = 0
    std::vector<double> Efull(size_t(fX) * fY * Lf.resZ, 0.0);
    for (int e = 0; e < m.numElements; ++e) Efull[m.eleMapBack[e]] = E[e];
    std::vector<double> tpl(size_t(s) * s * s * 576);
    for (int iz = 0; iz < s; ++iz)
        for (int iy = 0; iy < s; ++iy)
            for (int ix = 0; ix < s; ++ix) {
                double T[576];
                sub_T(W, s, ix, iy, iz, T);
                tt_a_t(m.Ke.data(), T, &tpl[size_t((iz * s + iy) * s + ix) * 576], false);
            }
    Kc.assign(size_t(N) * N, 0.0);
    for (int cez = 0; cez < Lc.resZ; ++cez)
        for (int cex = 0; cex < Lc.resX; ++cex)
            for (int cey = 0; cey < Lc.resY; ++cey) {
                int cdof[24];
                for (int a = 0; a < 8; ++a) {
                    const int n = cnx * cny * (cez + DZ[a]) + cnx * (Lc.resY - cey - 1 + DY[a]) + (cex + DX[a]);
                    for (int c = 0; c < 3; ++c) cdof[3 * a + c] = 3 * n + c;
                }
                double Kce[576] = {};
                const int fx0 = cex * s, fy0 = (Lc.resY - cey) * s, fz0 = cez * s;
                for (int iz = 0; iz < s; ++iz)
                    for (int iy = 0; iy < s; ++iy)
                        for (int ix = 0; ix < s; ++ix) {
                            const int fex = fx0 + ix, fey = fy0 - iy - 1, fez = fz0 + iz;
                            const int ef = (fY * fX) * fez + fY * fex + (fY - 1 - fey);
                            const double Ee = (H.Emin < Efull[ef]) ? Efull[ef] : H.Emin;
                            // the fixed-DOF test reads the structured finest element
                            // numbered ef in (ez, ex, ey) order -- the y-mirrored cell
                            const int mz = ef / (fX * fY), mr = ef % (fX * fY), mx = mr / fY, my = mr % fY;
                            int nodes[8];
                            bool masked = false;
                            for (int v = 0; v < 8; ++v) {
                                nodes[v] = fnx * fny * (mz + DZ[v]) + fnx * (fY - my - 1 + DY[v]) + (mx + DX[v]);
                                if (fm[3 * nodes[v]] | fm[3 * nodes[v] + 1] | fm[3 * nodes[v] + 2]) masked = true;
                            }
                            if (!masked) {
                                const double* Ks = &tpl[size_t((iz * s + iy) * s + ix) * 576];
                                for (int i = 0; i < 576; ++i) Kce[i] += Ee * Ks[i];
                            } else {
                                double Kf[576], T[576];
                                for (int i = 0; i < 576; ++i) Kf[i] = Ee * m.Ke[i];
                                for (int v = 0; v < 8; ++v)
                                    for (int c = 0; c < 3; ++c)
                                        if (fm[3 * nodes[v] + c]) {
                                            const int d = 3 * v + c;
                                            for (int j = 0; j < 24; ++j) Kf[d * 24 + j] = Kf[j * 24 + d] = 0.0;
                                            Kf[d * 24 + d] = 1.0;
                                        }
                                sub_T(W, s, ix, iy, iz, T);
                                tt_a_t(Kf, T, Kce, true);
                            }
                        }
                for (int i = 0; i < 24; ++i)
                    for (int j = 0; j < 24; ++j)
                        if (cdof[i] >= cdof[j]) Kc[size_t(cdof[i]) * N + cdof[j]] += Kce[i * 24 + j];
            }
    // coarsest-level boundary rows/columns (preconditioner.cpp:64-73)
    const std::vector<uint8_t>& cm = H.masks[L];
    for (int d = 0; d < N; ++d)
        if (cm[d]) {
            for (int j = 0; j < N; ++j) Kc[size_t(d) * N + j] = Kc[size_t(j) * N + d] = 0.0;
            Kc[size_t(d) * N + d] = 1.0;
        }
    return N;
}

bool chol_factor(std::vector<double>& A, int N)
{
    auto at = [&](int i, int j) -> double& { return A[size_t(i) * N + j]; };
    if (N < 256) {
        for (int i = 0; i < N; ++i) {
            for (int j = 0; j <= i; ++j) {
                double s = at(i, j);
                for (int k = 0; k < j; ++k) s -= at(i, k) * at(j, k);
                if (i == j) {
                    if (s <= 0.0) return false;
                    at(i, j) = std::sqrt(s);
                } else {
                    at(i, j) = s / at(j, j);
                }
            }
            for (int j = i + 1; j < N; ++j) at(i, j) = 0.0;
        }
        return true;
    }
    for (int j = 0; j < N; ++j) {
        double s = at(j, j);
        for (int k = 0; k < j; ++k) s -= at(j, k) * at(j, k);
        if (s <= 0.0) return false;
        at(j, j) = std::sqrt(s);
        const double inv = 1.0 / at(j, j);
        for (int i = j + 1; i < N; ++i) {
            double t = at(i, j);
            for (int k = 0; k < j; ++k) t -= at(i, k) * at(j, k);
            at(i, j) = t * inv;
        }
    }
    for (int i = 0; i < N; ++i)
        for (int j = i + 1; j < N; ++j) at(i, j) = 0.0;
    return true;
}

}  // namespace t3d
