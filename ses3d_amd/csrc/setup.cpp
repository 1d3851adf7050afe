// Host-side set-up of one section, float32, reference operation order.
// Compiled with -ffp-contract=off: every product and sum is rounded on its
// own, as in the reference's x86-64 build.
#include "setup.hpp"

#include <cmath>
#include <map>

namespace ses3d {

// GLL knots for lpd=4 as the reference spells them (SOURCE/ses3d_init.f90:129-133,
// SOURCE/ses3d_miscellaneous.f90:420-424).
static const float KN[NN] = {-1.0f, -0.6546536707079772f, 0.0f, 0.6546536707079772f, 1.0f};

// Lagrange basis i of degree 4 at x (SOURCE/ses3d_miscellaneous.f90:462-472).
float lgll(int i, float x) {
    float v = 1.0f;
    for (int k = 0; k < NN; ++k)
        if (k != i) v = (v * (x - KN[k])) / (KN[i] - KN[k]);
    return v;
}

// The reference's knots for Lagrange degree n = 0..4
// (SOURCE/ses3d_miscellaneous.f90:396-424, also ses3d_output.f90:41-79).
static float knot_n(int n, int i) {
    static const float K[5][5] = {{0.0f},
                                  {-1.0f, 1.0f},
                                  {-1.0f, 0.0f, 1.0f},
                                  {-1.0f, -0.4472135954999579f, 0.4472135954999579f, 1.0f},
                                  {-1.0f, -0.6546536707079772f, 0.0f, 0.6546536707079772f, 1.0f}};
    return K[n][i];
}

// Lagrange basis i of degree n at x (SOURCE/ses3d_miscellaneous.f90:377-481).
float lgll_n(int n, int i, float x) {
    float v = 1.0f;
    if (n == 0) return v;
    for (int k = 0; k <= n; ++k)
        if (k != i) v = (v * (x - knot_n(n, k))) / (knot_n(n, i) - knot_n(n, k));
    return v;
}

void stored_degree_tables(int deg, float *hlc, float *hld) {
    for (int k = 0; k < 40; ++k) (k < 20 ? hlc[k] : hld[k - 20]) = 0.0f;
    if (deg < 0 || deg >= 4) return;
    for (int p = 0; p < NN; ++p)
        for (int i = 0; i <= deg; ++i) {
            hlc[p * 4 + i] = lgll(p, knot_n(deg, i));      // lgll(lpd, ip, cknots(i)), ses3d_output.f90:119
            hld[i * 5 + p] = lgll_n(deg, i, KN[p]);        // lgll(deg, il, knots(i)), ses3d_output.f90:254
        }
}

// L_i'(knot_j) (SOURCE/ses3d_miscellaneous.f90:575-605).
float dlgll(int i, int j) {
    float y;
    if (i == j) {
        y = 0.0f;
        for (int k = 0; k < NN; ++k)
            if (k != i) y = y + 1.0f / (KN[i] - KN[k]);
    } else {
        y = 1.0f / (KN[i] - KN[j]);
        for (int k = 0; k < NN; ++k)
            if (k != i && k != j) y = (y * (KN[j] - KN[k])) / (KN[i] - KN[k]);
    }
    return y;
}

void MeshSetup::build(int nx_, int ny_, int nz_, int pml_, float xmin_, float xmax_, float ymin_,
                      float ymax_, float zmin_, float zmax_, float zmax_global_, float dt_) {
    nx = nx_; ny = ny_; nz = nz_; pml = pml_;
    xmin = xmin_; xmax = xmax_; ymin = ymin_; ymax = ymax_; zmin = zmin_; zmax = zmax_;
    zmax_global = zmax_global_; dt = dt_;
    for (int n = 0; n < NN; ++n) knots[n] = KN[n];
    // weights as the reference's literals (SOURCE/ses3d_init.f90:135-139)
    w[0] = 0.1000000000000f; w[1] = 0.5444444444444f; w[2] = 0.7111111111111f;
    w[3] = 0.5444444444444f; w[4] = 0.1000000000000f;
    for (int i = 0; i < NN; ++i)
        for (int j = 0; j < NN; ++j) dl[i][j] = dlgll(i, j);
    // element sizes, coordinates, Jacobian (SOURCE/ses3d_init.f90:213-248)
    dx = (xmax - xmin) / (float)(nx + 1);
    dy = (ymax - ymin) / (float)(ny + 1);
    dz = (zmax - zmin) / (float)(nz + 1);
    x.assign((nx + 1) * NN, 0.f); sin_x = x; cos_x = x; cot_x = x;
    y.assign((ny + 1) * NN, 0.f);
    z.assign((nz + 1) * NN, 0.f);
    for (int i = 0; i <= nx; ++i)
        for (int n = 0; n < NN; ++n) {
            float xv = (xmin + (float)i * dx) + (0.5f * (1.0f + knots[n])) * dx;
            x[i * NN + n] = xv;
            sin_x[i * NN + n] = sinf(xv);
            cos_x[i * NN + n] = cosf(xv);
            cot_x[i * NN + n] = cosf(xv) / sinf(xv);
        }
    for (int j = 0; j <= ny; ++j)
        for (int n = 0; n < NN; ++n)
            y[j * NN + n] = (ymin + (float)j * dy) + (0.5f * (1.0f + knots[n])) * dy;
    for (int k = 0; k <= nz; ++k)
        for (int n = 0; n < NN; ++n)
            z[k * NN + n] = (zmax - (float)k * dz) - (0.5f * (1.0f + knots[n])) * dz;
    Jac = ((dx * dy) * dz) / 8.0f;
    // derivative proxies dummy_1..3 (SOURCE/ses3d_evolution.f90:66-68), [q][i]
    for (int i = 0; i < NN; ++i)
        for (int q = 0; q < NN; ++q) {
            D3[q * NN + i] = (2.0f * dl[q][i]) / dz;
            D2[q * NN + i] = (2.0f * dl[q][i]) / dy;
            D1[q * NN + i] = (2.0f * dl[q][i]) / dx;
        }
    // divergence weights (SOURCE/ses3d_evolution.f90:241-249)
    for (int a = 0; a < NN; ++a)
        for (int b = 0; b < NN; ++b) {
            WDL[a * NN + b] = w[b] * dl[a][b];   // w(q)*dl(i,q) with q=b
            DUM1[a * NN + b] = (((2.0f / dx) * Jac) * w[a]) * w[b];
            DUM2[a * NN + b] = (((2.0f / dy) * Jac) * w[a]) * w[b];
            DUM3[a * NN + b] = -((((2.0f / dz) * Jac) * w[a]) * w[b]);
        }
    // absorbing profiles (SOURCE/ses3d_init.f90:578-686); one box holds the
    // whole section, so ix_multi==1==px etc. and every lateral and the lower
    // boundary carry a layer; the upper one is disabled (:586).
    const float alpha = 1.2f;
    IX.assign((nx + 1) * NN, 0); IY.assign((ny + 1) * NN, 0); IZ.assign((nz + 1) * NN, 0);
    VX.assign(1, 0.f); VY.assign(1, 0.f); VZ.assign(1, 0.f);
    std::vector<float> px((nx + 1) * NN, 0.f), py((ny + 1) * NN, 0.f), pz((nz + 1) * NN, 0.f);
    std::vector<char> sx(px.size(), 0), sy(py.size(), 0), sz(pz.size(), 0);
    for (int k = nz - pml + 1; k <= nz; ++k)
        for (int n = 0; n < NN; ++n) {
            if (k < 0) continue;
            float d = ((zmin + (float)pml * dz) - z[k * NN + n]) / ((float)pml * dz);
            pz[k * NN + n] = (alpha * d) * d; sz[k * NN + n] = 1;
        }
    for (int i = 0; i <= pml - 1 && i <= nx; ++i)
        for (int n = 0; n < NN; ++n) {
            float d = ((xmin + (float)pml * dx) - x[i * NN + n]) / ((float)pml * dx);
            px[i * NN + n] = (alpha * d) * d; sx[i * NN + n] = 1;
        }
    for (int i = nx - pml + 1; i <= nx; ++i)
        for (int n = 0; n < NN; ++n) {
            if (i < 0) continue;
            float d = ((xmax - dx * (float)pml) - x[i * NN + n]) / ((float)pml * dx);
            px[i * NN + n] = (alpha * d) * d; sx[i * NN + n] = 1;
        }
    for (int j = 0; j <= pml - 1 && j <= ny; ++j)
        for (int n = 0; n < NN; ++n) {
            float d = ((ymin + (float)pml * dy) - y[j * NN + n]) / ((float)pml * dy);
            py[j * NN + n] = (alpha * d) * d; sy[j * NN + n] = 1;
        }
    for (int j = ny - pml + 1; j <= ny; ++j)
        for (int n = 0; n < NN; ++n) {
            if (j < 0) continue;
            float d = ((ymax - dy * (float)pml) - y[j * NN + n]) / ((float)pml * dy);
            py[j * NN + n] = (alpha * d) * d; sy[j * NN + n] = 1;
        }
    for (size_t q = 0; q < px.size(); ++q)
        if (sx[q]) { IX[q] = (uint16_t)VX.size(); VX.push_back(px[q]); }
    for (size_t q = 0; q < py.size(); ++q)
        if (sy[q]) { IY[q] = (uint16_t)VY.size(); VY.push_back(py[q]); }
    for (size_t q = 0; q < pz.size(); ++q)
        if (sz[q]) { IZ[q] = (uint16_t)VZ.size(); VZ.push_back(pz[q]); }
    KX = (int)VX.size(); KY = (int)VY.size(); KZ = (int)VZ.size();
    LUT.assign((size_t)KX * KY * KZ * 5, 0.f);
    for (int iz = 0; iz < KZ; ++iz)
        for (int iy = 0; iy < KY; ++iy)
            for (int ix = 0; ix < KX; ++ix) {
                float fx = VX[ix], fy = VY[iy], fz = VZ[iz];
                float pr = (fx + fy) + fz;
                float nzv = (2.0f * fz) / (1.0f + pr / alpha);
                float nyv = (2.0f * fy) / (1.0f + pr / alpha);
                float nxv = (2.0f * fx) / (1.0f + pr / alpha);
                float p2 = (nxv + nyv) + nzv;
                float tp = expf(-((0.1f * p2) * p2));
                float *o = &LUT[(((size_t)iz * KY + iy) * KX + ix) * 5];
                o[0] = nxv; o[1] = nyv; o[2] = nzv; o[3] = p2; o[4] = tp;
            }
}

SourceLoc locate_source(const MeshSetup &m, float xxs, float yys, float zzs) {
    SourceLoc s;
    // SOURCE/ses3d_input.f90:407-409
    float zr = m.zmax_global - zzs;
    s.is_source = (m.xmin <= xxs) && (m.xmax >= xxs) && (m.ymin <= yys) && (m.ymax >= yys) &&
                  (m.zmin <= zr) && (m.zmax >= zr);
    if (!s.is_source) return s;
    // nearest node, strict < so the first minimum wins (SOURCE/ses3d_init.f90:367-417)
    float ddx = m.dx, ddy = m.dy, ddz = m.dz;
    for (int i = 0; i <= m.nx; ++i)
        for (int n = 0; n < NN; ++n)
            if (fabsf(m.x[i * NN + n] - xxs) < ddx) { ddx = fabsf(m.x[i * NN + n] - xxs); s.isx = i; }
    for (int j = 0; j <= m.ny; ++j)
        for (int n = 0; n < NN; ++n)
            if (fabsf(m.y[j * NN + n] - yys) < ddy) { ddy = fabsf(m.y[j * NN + n] - yys); s.isy = j; }
    for (int k = 0; k <= m.nz; ++k)
        for (int n = 0; n < NN; ++n)
            if (fabsf((m.zmax_global - m.z[k * NN + n]) - zzs) < ddz) {
                ddz = fabsf((m.zmax_global - m.z[k * NN + n]) - zzs); s.isz = k;
            }
    // local coordinates (SOURCE/ses3d_init.f90:431-433)
    float z4 = m.z[s.isz * NN + 4], z0 = m.z[s.isz * NN];
    s.zl = ((z4 + z0) - 2.0f * (m.zmax_global - zzs)) / (z0 - z4);
    float y4 = m.y[s.isy * NN + 4], y0 = m.y[s.isy * NN];
    s.yl = ((2.0f * yys - y4) - y0) / (y4 - y0);
    float x4 = m.x[s.isx * NN + 4], x0 = m.x[s.isx * NN];
    s.xl = ((2.0f * xxs - x4) - x0) / (x4 - x0);
    return s;
}

void moment_weights(const MeshSetup &m, const SourceLoc &s, const float mom[6], float *md) {
    // SOURCE/ses3d_miscellaneous.f90:126-142; md = MOM*delta, so the device
    // forms src = 0 - md*so(it) exactly as src = src - MOM*delta*so(it).
    const float M9[9] = {mom[0], mom[1], mom[2], mom[3], mom[3], mom[4], mom[4], mom[5], mom[5]};
    for (int in = 0; in < NN; ++in)
        for (int jn = 0; jn < NN; ++jn)
            for (int kn = 0; kn < NN; ++kn) {
                float delta = ((lgll(in, s.xl) * lgll(jn, s.yl)) * lgll(kn, s.zl)) /
                              ((m.w[in] * m.w[jn]) * m.w[kn]);
                float zs = m.z[s.isz * NN + kn];
                delta = delta / (((zs * zs) * sinf(m.x[s.isx * NN + in])) * m.Jac);
                int nd = in + NN * (jn + NN * kn);
                for (int c = 0; c < 9; ++c) md[c * 125 + nd] = M9[c] * delta;
            }
}

void point_weights(float xl, float yl, float zl, float *w125) {
    for (int in = 0; in < NN; ++in)
        for (int jn = 0; jn < NN; ++jn)
            for (int kn = 0; kn < NN; ++kn)
                w125[in + NN * (jn + NN * kn)] = (lgll(in, xl) * lgll(jn, yl)) * lgll(kn, zl);
}

std::vector<PointLoc> locate_receivers(const MeshSetup &m, const float *rec, int nr) {
    std::vector<PointLoc> out;
    const float tol = 3.1415926535898f / (float)(180 * 100000);   // SOURCE/ses3d_init.f90:28
    for (int n = 0; n < nr; ++n) {
        float r1 = rec[3 * n], r2 = rec[3 * n + 1], r3 = rec[3 * n + 2];
        float zr = m.zmax_global - r3;
        if (!(r1 <= m.xmax && r1 > m.xmin && r2 <= m.ymax && r2 > m.ymin && zr <= m.zmax && zr > m.zmin))
            continue;
        PointLoc p;
        p.index = n;
        // last matching element wins (SOURCE/ses3d_init.f90:304-343)
        for (int i = 0; i <= m.nx; ++i)
            for (int j = 0; j <= m.ny; ++j)
                for (int k = 0; k <= m.nz; ++k)
                    if (r1 <= m.x[i * NN + 4] + tol && r1 >= m.x[i * NN] - tol &&
                        r2 <= m.y[j * NN + 4] + tol && r2 >= m.y[j * NN] - tol &&
                        zr <= m.z[k * NN] && zr >= m.z[k * NN + 4]) {
                        p.xl = 2.0f * (((r1 - m.xmin) / m.dx) - (float)i) - 1.0f;
                        p.yl = 2.0f * (((r2 - m.ymin) / m.dy) - (float)j) - 1.0f;
                        p.zl = -(2.0f * ((((m.zmax_global - r3) - m.zmax) / m.dz) + (float)k)) - 1.0f;
                        p.ex = i; p.ey = j; p.ez = k;
                    }
        out.push_back(p);
    }
    return out;
}

std::vector<PointLoc> locate_adjoint_sources(const MeshSetup &m, const float *loc, int n) {
    std::vector<PointLoc> out;
    for (int k = 0; k < n; ++k) {
        float a1 = loc[3 * k], a2 = loc[3 * k + 1], a3 = loc[3 * k + 2];
        float zr = m.zmax_global - a3;
        if (!(a1 <= m.xmax && a1 > m.xmin && a2 <= m.ymax && a2 > m.ymin && zr <= m.zmax && zr > m.zmin))
            continue;
        PointLoc p;
        p.index = k;
        float d = m.xmax - m.xmin;
        for (int i = 0; i <= m.nx; ++i)
            for (int q = 0; q < NN; ++q)
                if (fabsf(a1 - m.x[i * NN + q]) < d) { p.ex = i; d = fabsf(a1 - m.x[i * NN + q]); }
        p.xl = ((2.0f * a1 - m.x[p.ex * NN + 4]) - m.x[p.ex * NN]) / (m.x[p.ex * NN + 4] - m.x[p.ex * NN]);
        d = m.ymax - m.ymin;
        for (int j = 0; j <= m.ny; ++j)
            for (int q = 0; q < NN; ++q)
                if (fabsf(a2 - m.y[j * NN + q]) < d) { p.ey = j; d = fabsf(a2 - m.y[j * NN + q]); }
        p.yl = ((2.0f * a2 - m.y[p.ey * NN + 4]) - m.y[p.ey * NN]) / (m.y[p.ey * NN + 4] - m.y[p.ey * NN]);
        d = m.zmax - m.zmin;
        for (int kk = 0; kk <= m.nz; ++kk)
            for (int q = 0; q < NN; ++q)
                if (fabsf((m.zmax_global - a3) - m.z[kk * NN + q]) < d) {
                    p.ez = kk; d = fabsf((m.zmax_global - a3) - m.z[kk * NN + q]);
                }
        float z4 = m.z[p.ez * NN + 4], z0 = m.z[p.ez * NN];
        p.zl = ((z4 + z0) - 2.0f * (m.zmax_global - a3)) / (z0 - z4);
        out.push_back(p);
    }
    return out;
}

}  // namespace ses3d
