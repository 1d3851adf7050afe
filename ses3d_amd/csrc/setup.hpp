// Host-side set-up of one SES3D section in float32, in the reference's
// operation order (SOURCE/ses3d_init.f90:7-729).  Produces the small
// per-axis tables and source/receiver weights the device kernels consume;
// nothing here is per GLL point except what the caller asks for.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace ses3d {

constexpr int NN = 5;   // GLL points per element edge (lpd = 4)

struct MeshSetup {
    int nx = 0, ny = 0, nz = 0;      // elements - 1
    int pml = 2;
    float xmin = 0, xmax = 0, ymin = 0, ymax = 0, zmin = 0, zmax = 0, zmax_global = 0;
    float dt = 0;
    // GLL
    float knots[NN] = {}, w[NN] = {}, dl[NN][NN] = {};
    float dx = 0, dy = 0, dz = 0, Jac = 0;
    // per-axis coordinates: index e*5+n
    std::vector<float> x, y, z, sin_x, cos_x, cot_x;
    // contraction / divergence coefficient tables, [a][b] flattened a*5+b
    float D1[25], D2[25], D3[25];     // (2*dl(q,i))/dx etc., index q*5+i
    float WDL[25];                    // w(q)*dl(a,q), index a*5+q
    float DUM1[25], DUM2[25], DUM3[25];
    // absorbing profiles: per-axis ids into value tables; LUT over id triples
    std::vector<uint16_t> IX, IY, IZ; // index e*5+n -> id (0: outside the layer)
    std::vector<float> VX, VY, VZ;    // raw profile value per id (VX[0]=0)
    int KX = 1, KY = 1, KZ = 1;
    std::vector<float> LUT;           // 5 per (ix,iy,iz): prof_x', prof_y', prof_z', prof', taper

    void build(int nx_, int ny_, int nz_, int pml_, float xmin_, float xmax_, float ymin_,
               float ymax_, float zmin_, float zmax_, float zmax_global_, float dt_);
};

// One located moment/point source (SOURCE/ses3d_init.f90:363-437).
struct SourceLoc {
    bool is_source = false;
    int isx = 0, isy = 0, isz = 0;
    float xl = 0, yl = 0, zl = 0;     // local coordinates in [-1,1]
};

// One located receiver or adjoint source.
struct PointLoc {
    int index = -1;                   // row index in the caller's list
    int ex = 0, ey = 0, ez = 0;
    float xl = 0, yl = 0, zl = 0;
};

float lgll(int i, float x);
float dlgll(int i, int j);
// Lagrange basis i of degree n (0..4) through the reference's knots for
// that degree (SOURCE/ses3d_miscellaneous.f90:377-481).
float lgll_n(int n, int i, float x);
// Tables of the degree-deg stored forward fields (deg < 4; ses3d_store /
// ses3d_restore, SOURCE/ses3d_output.f90:36-130, :242-265):
// hlc[p*4+i] = lgll(4, p, knot_deg(i)), hld[i*5+p] = lgll(deg, i, knot_4(p)).
// Both are 20 floats; unused entries are 0.
void stored_degree_tables(int deg, float *hlc, float *hld);

SourceLoc locate_source(const MeshSetup &m, float xxs, float yys, float zzs);
// Moment-tensor weights: md[c*125+node] = MOM_c * delta(node), c in
// xx yy zz xy yx xz zx yz zy (SOURCE/ses3d_miscellaneous.f90:120-146).
void moment_weights(const MeshSetup &m, const SourceLoc &s, const float mom[6], float *md);
// Point-force weights delta(node) = L_l L_m L_n (misc.f90:39, :78).
void point_weights(float xl, float yl, float zl, float *w125);
// Receivers (SOURCE/ses3d_init.f90:292-347): returns located receivers in order.
std::vector<PointLoc> locate_receivers(const MeshSetup &m, const float *rec, int nr);
// Adjoint sources (SOURCE/ses3d_init.f90:457-570).
std::vector<PointLoc> locate_adjoint_sources(const MeshSetup &m, const float *loc, int n);

}  // namespace ses3d
