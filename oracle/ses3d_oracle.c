/*
 * ses3d_oracle.c -- CPU restatement of the SES3D time step, sources,
 * receivers and Frechet kernels.  TEST INFRASTRUCTURE ONLY: this file is the
 * checker that the HIP path is compared with.  Only tests/, the smoke check
 * and the CPU-baseline leg of the benchmark may load it; the product library
 * never links or calls it.
 *
 * Built twice from this one source:  -DREAL=float  -> libses3d_oracle32.so
 * (same FP32 arithmetic, same operation order as the reference Fortran) and
 * -DREAL=double -> libses3d_oracle64.so (the FP64 error yardstick).  Built
 * with -ffp-contract=off so no multiply-add is fused.
 *
 * Arrays use the reference layout: element-local fields are Fortran
 * (0:nx,0:ny,0:nz,0:4,0:4,0:4) column-major (element index along theta
 * fastest, GLL index along r slowest); global fields are
 * (0:4(nx+1),0:4(ny+1),0:4(nz+1)).
 *
 * Every expression keeps the reference's left-to-right evaluation order.
 * Function -> reference (file:line) map:
 *   o_init_coeffs     SOURCE/ses3d_init.f90:47-85 (material), :127-139 (GLL),
 *                     :201-207 (dl), :213-248 (geometry, Jac), :254-279 (mass)
 *   o_locate          SOURCE/ses3d_init.f90:285-439 (receivers, source),
 *                     :445-572 (adjoint sources); ses3d_input.f90:405-437
 *   o_init_pml        SOURCE/ses3d_init.f90:578-709
 *   lgll / dlgll      SOURCE/ses3d_miscellaneous.f90:377-481, :492-613
 *   o_store / o_grad  degree-compressed stored fields (hdeg < 4):
 *                     ses3d_store / ses3d_restore, ses3d_output.f90:36-130, :242-265
 *   o_step            SOURCE/ses3d_evolution.f90:7-373 with
 *                     make_moment_source (ses3d_miscellaneous.f90:97-150) and
 *                     add_single_force (ses3d_miscellaneous.f90:11-94),
 *                     then record_seismograms (ses3d_output.f90:456-497),
 *                     forward-field store (ses3d_main.f90:142-166,
 *                     ses3d_output.f90:5-164), kernel step ses3d_grad
 *                     (ses3d_grad.f90:25-85) and the shear-Q kernel of
 *                     ses3d_grad_Q.f90:111-128,151-153.
 *
 * Parity pins for each function are listed in the tests that use it.  The
 * reference never writes the shear-Q kernel (ses3d_output.f90:346-370 is
 * commented out); it is pinned against the reference's ses3d_grad_Q.f90
 * built with oracle/ref_grad_q_harness.f90, which writes it.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#define REAL float
#endif
typedef REAL real;

#define LPD 4
#define NN 5

typedef struct {
    int32_t nx, ny, nz;           /* elements - 1 per axis (boxfile nx_max-nx_min) */
    int32_t pml;
    int32_t is_diss;
    int32_t adjoint_flag;
    int32_t samp_ad;
    int32_t nt;
    float dt;
    float xmin, xmax, ymin, ymax, zmin, zmax;   /* rad, rad, m (boxfile) */
    float zmax_global;                           /* m (setup) */
    float tau_p[3], D_p[3];
} o_params;

typedef struct o_sim {
    o_params p;
    int nx, ny, nz, NX, NY, NZ;
    int64_t NE, NP, NGX, NGY, NGZ, NG;
    real dt, xmin, xmax, ymin, ymax, zmin, zmax, zmax_global;
    real tau_p[3], D_p[3], sum_D_p;
    real knots[NN], w[NN], dl[NN][NN];
    real dx, dy, dz, Jac;
    real *x, *y, *z, *sin_x, *cos_x;  /* (n+1)*5, index e*5+node */
    /* material */
    real *rhoinv, *mu, *lambda, *A, *B, *C, *QQ, *tau, *kappa, *mu_tau, *cs, *cp, *rho;
    /* geometry / pml */
    real *MM_global, *prof_x, *prof_y, *prof_z, *prof, *taper, *prof_global, *taper_global;
    /* dynamic fields */
    real *vxg, *vyg, *vzg, *sxg, *syg, *szg;
    real *vx, *vy, *vz;
    real *du[9];    /* dxux dyuy dzuz dxuy dyux dxuz dzux dyuz dzuy */
    real *sp[9];    /* sxx syy szz sxy syx sxz szx syz szy (split, "_pml") */
    real *Mv[6][3]; /* Mxx Myy Mzz Mxy Mxz Myz, per mechanism */
    real *e[9];     /* exx eyy ezz exy eyx exz ezx eyz ezy (rates) */
    real *sx, *sy, *sz;
    /* moment source in the source element: 9 x 125 */
    real src[9][125];
    /* source */
    int source_type, is_source, isx, isy, isz;
    real xxs, yys, zzs, xxs_loc, yys_loc, zzs_loc, MOM[6];
    real *so;
    /* receivers */
    int nr;
    int *rx, *ry, *rz;
    real *rstd;     /* 3*nr */
    real *seis;     /* 3 x nr x nt */
    /* adjoint sources */
    int nad;
    int *ad_is;     /* 3*nad */
    real *ad_loc;   /* 3*nad local coords */
    real *ad_stf;   /* 3 x nad x nt */
    /* forward history (adjoint_flag 1 writes, 2 reads) */
    int n_hist;
    int *hist_step;
    float **hist;   /* per stored step as written to disk: 9 x NP float32, or with
                       hdeg < 4 the 9 degree-hdeg fields (9 x NE x (hdeg+1)^3) */
    int hdeg;       /* polynomial degree of the stored fields (4 = uncompressed) */
    int *saving_vector;
    /* kernels */
    real *grad_rho, *grad_cp, *grad_csh, *grad_csv, *grad_Q_mu;
    real *fw[9];    /* restored forward fields */
} o_sim;

#define LIDX(o, i, j, k, l, m, n) \
    ((int64_t)(i) + (o)->NX * ((int64_t)(j) + (o)->NY * ((int64_t)(k) + (o)->NZ * \
    ((int64_t)(l) + NN * ((int64_t)(m) + NN * (int64_t)(n))))))
#define GIDX(o, a, b, c) ((int64_t)(a) + (o)->NGX * ((int64_t)(b) + (o)->NGY * (int64_t)(c)))

static real *alloc0(int64_t n) { return (real *)calloc((size_t)(n > 0 ? n : 1), sizeof(real)); }

static const real KN[NN] = {(real)-1.0, (real)-0.6546536707079772, (real)0.0,
                            (real)0.6546536707079772, (real)1.0};

/* SOURCE/ses3d_miscellaneous.f90:377-481 */
static real lgll(int i, real xv) {
    real v = (real)1.0;
    for (int k = 0; k <= LPD; ++k)
        if (k != i) v = (v * (xv - KN[k])) / (KN[i] - KN[k]);
    return v;
}

/* lgll(n, i, x) of SOURCE/ses3d_miscellaneous.f90:377-481 for the degrees
 * 0..3 that the degree-compressed store uses (knot literals of :396-416). */
static real knot_n(int n, int i) {
    static const real K[5][5] = {{(real)0.0},
                                 {(real)-1.0, (real)1.0},
                                 {(real)-1.0, (real)0.0, (real)1.0},
                                 {(real)-1.0, (real)-0.4472135954999579, (real)0.4472135954999579, (real)1.0},
                                 {(real)-1.0, (real)-0.6546536707079772, (real)0.0, (real)0.6546536707079772, (real)1.0}};
    return K[n][i];
}
static real lgll_n(int n, int i, real xv) {
    real v = (real)1.0;
    if (n == 0) return v;
    for (int k = 0; k <= n; ++k)
        if (k != i) v = (v * (xv - knot_n(n, k))) / (knot_n(n, i) - knot_n(n, k));
    return v;
}

/* SOURCE/ses3d_miscellaneous.f90:492-613 ; dl(i,j) = L_i'(knot_j) */
static real dlgll(int i, int j) {
    real y;
    if (i == j) {
        y = (real)0.0;
        for (int k = 0; k <= LPD; ++k)
            if (k != i) y = y + (real)1.0 / (KN[i] - KN[k]);
    } else {
        y = (real)1.0 / (KN[i] - KN[j]);
        for (int k = 0; k <= LPD; ++k)
            if (k != i && k != j) y = (y * (KN[j] - KN[k])) / (KN[i] - KN[k]);
    }
    return y;
}

o_sim *o_create(const o_params *p) {
    o_sim *o = (o_sim *)calloc(1, sizeof(o_sim));
    o->p = *p;
    o->nx = p->nx; o->ny = p->ny; o->nz = p->nz;
    o->NX = p->nx + 1; o->NY = p->ny + 1; o->NZ = p->nz + 1;
    o->NE = (int64_t)o->NX * o->NY * o->NZ;
    o->NP = o->NE * 125;
    o->NGX = 4 * (int64_t)o->NX + 1; o->NGY = 4 * (int64_t)o->NY + 1; o->NGZ = 4 * (int64_t)o->NZ + 1;
    o->NG = o->NGX * o->NGY * o->NGZ;
    o->dt = (real)p->dt;
    o->xmin = (real)p->xmin; o->xmax = (real)p->xmax;
    o->ymin = (real)p->ymin; o->ymax = (real)p->ymax;
    o->zmin = (real)p->zmin; o->zmax = (real)p->zmax;
    o->zmax_global = (real)p->zmax_global;
    for (int k = 0; k < 3; ++k) { o->tau_p[k] = (real)p->tau_p[k]; o->D_p[k] = (real)p->D_p[k]; }
    o->sum_D_p = (o->D_p[0] + o->D_p[1]) + o->D_p[2];
    int64_t NP = o->NP, NG = o->NG;
    real **np_arrays[] = {&o->rhoinv, &o->mu, &o->lambda, &o->A, &o->B, &o->C, &o->QQ, &o->tau,
                          &o->kappa, &o->mu_tau, &o->cs, &o->cp, &o->rho, &o->prof_x, &o->prof_y,
                          &o->prof_z, &o->prof, &o->taper, &o->vx, &o->vy, &o->vz, &o->sx, &o->sy,
                          &o->sz, &o->grad_rho, &o->grad_cp, &o->grad_csh, &o->grad_csv, &o->grad_Q_mu};
    for (size_t a = 0; a < sizeof(np_arrays) / sizeof(np_arrays[0]); ++a) *np_arrays[a] = alloc0(NP);
    for (int c = 0; c < 9; ++c) { o->du[c] = alloc0(NP); o->sp[c] = alloc0(NP); o->e[c] = alloc0(NP); o->fw[c] = alloc0(NP); }
    if (p->is_diss)
        for (int c = 0; c < 6; ++c)
            for (int k = 0; k < 3; ++k) o->Mv[c][k] = alloc0(NP);
    real **ng_arrays[] = {&o->MM_global, &o->prof_global, &o->taper_global, &o->vxg, &o->vyg,
                          &o->vzg, &o->sxg, &o->syg, &o->szg};
    for (size_t a = 0; a < sizeof(ng_arrays) / sizeof(ng_arrays[0]); ++a) *ng_arrays[a] = alloc0(NG);
    o->x = alloc0(o->NX * NN); o->y = alloc0(o->NY * NN); o->z = alloc0(o->NZ * NN);
    o->sin_x = alloc0(o->NX * NN); o->cos_x = alloc0(o->NX * NN);
    o->so = alloc0(p->nt + 1);
    o->saving_vector = (int *)calloc((size_t)p->nt + 2, sizeof(int));
    o->hdeg = LPD;
    return o;
}

/* Degree of the stored forward fields (ses3d_store/ses3d_restore argument
 * `deg`, SOURCE/ses3d_output.f90:5-267); 4 = lpd stores them unchanged. */
int o_set_history_degree(o_sim *o, int deg) {
    if (deg < 0 || deg > LPD || o->n_hist) return -1;
    o->hdeg = deg;
    return 0;
}

/* floats per stored step */
int64_t o_history_size(o_sim *o) {
    const int64_t D = o->hdeg + 1;
    return o->hdeg < LPD ? 9 * o->NE * D * D * D : 9 * o->NP;
}

void o_destroy(o_sim *o) {
    if (!o) return;
    real *arrs[] = {o->rhoinv, o->mu, o->lambda, o->A, o->B, o->C, o->QQ, o->tau, o->kappa,
                    o->mu_tau, o->cs, o->cp, o->rho, o->prof_x, o->prof_y, o->prof_z, o->prof,
                    o->taper, o->vx, o->vy, o->vz, o->sx, o->sy, o->sz, o->grad_rho, o->grad_cp,
                    o->grad_csh, o->grad_csv, o->grad_Q_mu, o->MM_global, o->prof_global,
                    o->taper_global, o->vxg, o->vyg, o->vzg, o->sxg, o->syg, o->szg, o->x, o->y,
                    o->z, o->sin_x, o->cos_x, o->so, o->rstd, o->seis, o->ad_loc, o->ad_stf};
    for (size_t a = 0; a < sizeof(arrs) / sizeof(arrs[0]); ++a) free(arrs[a]);
    for (int c = 0; c < 9; ++c) { free(o->du[c]); free(o->sp[c]); free(o->e[c]); free(o->fw[c]); }
    for (int c = 0; c < 6; ++c)
        for (int k = 0; k < 3; ++k) free(o->Mv[c][k]);
    free(o->rx); free(o->ry); free(o->rz); free(o->ad_is); free(o->saving_vector);
    for (int h = 0; h < o->n_hist; ++h) free(o->hist[h]);
    free(o->hist); free(o->hist_step);
    free(o);
}

/* Model in reference layout, float32 as read from the model files.
 * Q may be NULL when is_diss == 0.  SOURCE/ses3d_input.f90:479-499. */
void o_set_model(o_sim *o, const float *rhoinv, const float *mu, const float *lambda,
                 const float *A, const float *B, const float *C, const float *Q) {
    for (int64_t i = 0; i < o->NP; ++i) {
        o->rhoinv[i] = (real)rhoinv[i]; o->mu[i] = (real)mu[i]; o->lambda[i] = (real)lambda[i];
        o->A[i] = (real)A[i]; o->B[i] = (real)B[i]; o->C[i] = (real)C[i];
        if (Q) { o->QQ[i] = (real)Q[i]; o->tau[i] = (real)1.0 / o->QQ[i]; }
    }
}

/* Source: radians/metres as held after ses3d_input; so = nt samples. */
void o_set_source(o_sim *o, int source_type, float xxs, float yys, float zzs, const float *mom6,
                  const float *so) {
    o->source_type = source_type;
    o->xxs = (real)xxs; o->yys = (real)yys; o->zzs = (real)zzs;
    for (int k = 0; k < 6; ++k) o->MOM[k] = (real)mom6[k];
    for (int k = 0; k < o->p.nt; ++k) o->so[k] = (real)so[k];
    /* SOURCE/ses3d_input.f90:407-409 */
    real zr = o->zmax_global - o->zzs;
    o->is_source = (o->xmin <= o->xxs) && (o->xmax >= o->xxs) && (o->ymin <= o->yys) &&
                   (o->ymax >= o->yys) && (o->zmin <= zr) && (o->zmax >= zr);
}

/* Receivers: (nr,3) float32 rows (colat rad, lon rad, depth m). Located in o_init. */
static float *g_rec = NULL;
static int g_nrec = 0;
void o_set_receivers(o_sim *o, int nr, const float *rec) {
    (void)o;
    free(g_rec);
    g_rec = (float *)malloc(sizeof(float) * 3 * (nr > 0 ? nr : 1));
    memcpy(g_rec, rec, sizeof(float) * 3 * nr);
    g_nrec = nr;
}

/* Adjoint sources: loc (n,3) rad/rad/m; stf (n, nt, 3) rows as read from
 * ad_src_<k> (already time-reversed by the writer). */
static float *g_ad = NULL, *g_adstf = NULL;
static int g_nad = 0;
void o_set_adjoint_sources(o_sim *o, int n, const float *loc, const float *stf) {
    free(g_ad); free(g_adstf);
    g_ad = (float *)malloc(sizeof(float) * 3 * (n > 0 ? n : 1));
    g_adstf = (float *)malloc(sizeof(float) * 3 * (n > 0 ? n : 1) * (size_t)o->p.nt);
    memcpy(g_ad, loc, sizeof(float) * 3 * n);
    memcpy(g_adstf, stf, sizeof(float) * 3 * n * (size_t)o->p.nt);
    g_nad = n;
}

static void o_init_coeffs(o_sim *o) {
    int64_t NP = o->NP;
    /* SOURCE/ses3d_init.f90:47-58 */
    for (int64_t i = 0; i < NP; ++i) {
        o->kappa[i] = o->lambda[i] + ((real)2 * o->mu[i]) / (real)3;
        o->cs[i] = (real)sqrt((double)(o->mu[i] * o->rhoinv[i]));
        o->cp[i] = (real)sqrt((double)((o->lambda[i] + (real)2 * o->mu[i]) * o->rhoinv[i]));
        o->rho[i] = o->rhoinv[i] > (real)0 ? (real)1 / o->rhoinv[i] : (real)0;
        if (o->p.is_diss) o->mu_tau[i] = ((real)1 + o->sum_D_p * o->tau[i]) * o->mu[i];
    }
    /* GLL: SOURCE/ses3d_init.f90:129-139 */
    for (int n = 0; n < NN; ++n) o->knots[n] = KN[n];
    o->w[0] = (real)0.1000000000000; o->w[1] = (real)0.5444444444444; o->w[2] = (real)0.7111111111111;
    o->w[3] = (real)0.5444444444444; o->w[4] = (real)0.1000000000000;
    for (int i = 0; i < NN; ++i)
        for (int j = 0; j < NN; ++j) o->dl[i][j] = dlgll(i, j);
    /* :213-242 */
    o->dx = (o->xmax - o->xmin) / (real)(o->nx + 1);
    o->dy = (o->ymax - o->ymin) / (real)(o->ny + 1);
    o->dz = (o->zmax - o->zmin) / (real)(o->nz + 1);
    for (int i = 0; i <= o->nx; ++i)
        for (int n = 0; n < NN; ++n) {
            real xv = (o->xmin + (real)i * o->dx) + ((real)0.5 * ((real)1 + o->knots[n])) * o->dx;
            o->x[i * NN + n] = xv;
            if (sizeof(real) == sizeof(float)) {
                o->sin_x[i * NN + n] = (real)sinf((float)xv);
                o->cos_x[i * NN + n] = (real)cosf((float)xv);
            } else {
                o->sin_x[i * NN + n] = (real)sin((double)xv);
                o->cos_x[i * NN + n] = (real)cos((double)xv);
            }
        }
    for (int j = 0; j <= o->ny; ++j)
        for (int n = 0; n < NN; ++n)
            o->y[j * NN + n] = (o->ymin + (real)j * o->dy) + ((real)0.5 * ((real)1 + o->knots[n])) * o->dy;
    for (int k = 0; k <= o->nz; ++k)
        for (int n = 0; n < NN; ++n)
            o->z[k * NN + n] = (o->zmax - (real)k * o->dz) - ((real)0.5 * ((real)1 + o->knots[n])) * o->dz;
    /* :248 */
    o->Jac = ((o->dx * o->dy) * o->dz) / (real)8;
    /* :254-277 mass matrix, scatter in loop order i,j,k,l,m,n */
    memset(o->MM_global, 0, sizeof(real) * o->NG);
    for (int i = 0; i <= o->nx; ++i)
        for (int j = 0; j <= o->ny; ++j)
            for (int k = 0; k <= o->nz; ++k)
                for (int l = 0; l < NN; ++l)
                    for (int m = 0; m < NN; ++m)
                        for (int n = 0; n < NN; ++n) {
                            int64_t g = GIDX(o, i * LPD + l, j * LPD + m, k * LPD + n);
                            real zk = o->z[k * NN + n];
                            real t = ((((((o->w[l] * o->w[m]) * o->w[n]) * o->Jac) * zk) * zk) *
                                      o->sin_x[i * NN + l]) / o->rhoinv[LIDX(o, i, j, k, l, m, n)];
                            o->MM_global[g] = o->MM_global[g] - t;
                        }
}

static void o_locate(o_sim *o) {
    const int af = o->p.adjoint_flag;
    const real tol = (real)3.1415926535898 / (real)(180 * 100000);  /* init.f90:28 */
    if (af == 0 || af == 1) {
        /* receivers, init.f90:292-347 */
        free(o->rx); free(o->ry); free(o->rz); free(o->rstd); free(o->seis);
        o->rx = (int *)calloc(g_nrec + 1, sizeof(int));
        o->ry = (int *)calloc(g_nrec + 1, sizeof(int));
        o->rz = (int *)calloc(g_nrec + 1, sizeof(int));
        o->rstd = alloc0(3 * (g_nrec + 1));
        o->nr = 0;
        for (int n = 0; n < g_nrec; ++n) {
            real r1 = (real)g_rec[3 * n], r2 = (real)g_rec[3 * n + 1], r3 = (real)g_rec[3 * n + 2];
            real zr = o->zmax_global - r3;
            if (r1 <= o->xmax && r1 > o->xmin && r2 <= o->ymax && r2 > o->ymin && zr <= o->zmax &&
                zr > o->zmin) {
                int nr = o->nr++;
                for (int i = 0; i <= o->nx; ++i)
                    for (int j = 0; j <= o->ny; ++j)
                        for (int k = 0; k <= o->nz; ++k) {
                            if (r1 <= o->x[i * NN + LPD] + tol && r1 >= o->x[i * NN] - tol &&
                                r2 <= o->y[j * NN + LPD] + tol && r2 >= o->y[j * NN] - tol &&
                                zr <= o->z[k * NN] && zr >= o->z[k * NN + LPD]) {
                                o->rstd[3 * nr + 0] = (real)2 * (((r1 - o->xmin) / o->dx) - (real)i) - (real)1;
                                o->rstd[3 * nr + 1] = (real)2 * (((r2 - o->ymin) / o->dy) - (real)j) - (real)1;
                                o->rstd[3 * nr + 2] = -((real)2 * ((((o->zmax_global - r3) - o->zmax) / o->dz) + (real)k)) - (real)1;
                                o->rx[nr] = i; o->ry[nr] = j; o->rz[nr] = k;
                            }
                        }
            }
        }
        o->seis = alloc0(3 * (int64_t)(o->nr + 1) * o->p.nt);
        /* source, init.f90:365-437 */
        if (o->is_source) {
            real dx_ = o->dx, dy_ = o->dy, dz_ = o->dz;
            for (int i = 0; i <= o->nx; ++i)
                for (int n = 0; n < NN; ++n)
                    if ((real)fabs((double)(o->x[i * NN + n] - o->xxs)) < dx_) {
                        dx_ = (real)fabs((double)(o->x[i * NN + n] - o->xxs)); o->isx = i;
                    }
            for (int j = 0; j <= o->ny; ++j)
                for (int n = 0; n < NN; ++n)
                    if ((real)fabs((double)(o->y[j * NN + n] - o->yys)) < dy_) {
                        dy_ = (real)fabs((double)(o->y[j * NN + n] - o->yys)); o->isy = j;
                    }
            for (int k = 0; k <= o->nz; ++k)
                for (int n = 0; n < NN; ++n)
                    if ((real)fabs((double)((o->zmax_global - o->z[k * NN + n]) - o->zzs)) < dz_) {
                        dz_ = (real)fabs((double)((o->zmax_global - o->z[k * NN + n]) - o->zzs)); o->isz = k;
                    }
            real z4 = o->z[o->isz * NN + LPD], z0 = o->z[o->isz * NN];
            o->zzs_loc = ((z4 + z0) - (real)2 * (o->zmax_global - o->zzs)) / (z0 - z4);
            real y4 = o->y[o->isy * NN + LPD], y0 = o->y[o->isy * NN];
            o->yys_loc = (((real)2 * o->yys - y4) - y0) / (y4 - y0);
            real x4 = o->x[o->isx * NN + LPD], x0 = o->x[o->isx * NN];
            o->xxs_loc = (((real)2 * o->xxs - x4) - x0) / (x4 - x0);
        }
    }
    if (af == 2) {
        /* adjoint sources, init.f90:453-572 */
        free(o->ad_is); free(o->ad_loc); free(o->ad_stf);
        o->ad_is = (int *)calloc(3 * (g_nad + 1), sizeof(int));
        o->ad_loc = alloc0(3 * (g_nad + 1));
        o->ad_stf = alloc0(3 * (int64_t)(g_nad + 1) * o->p.nt);
        o->nad = 0;
        for (int k = 0; k < g_nad; ++k) {
            real a1 = (real)g_ad[3 * k], a2 = (real)g_ad[3 * k + 1], a3 = (real)g_ad[3 * k + 2];
            real zr = o->zmax_global - a3;
            if (!(a1 <= o->xmax && a1 > o->xmin && a2 <= o->ymax && a2 > o->ymin && zr <= o->zmax &&
                  zr > o->zmin))
                continue;
            int idx = o->nad++;
            for (int t = 0; t < o->p.nt; ++t)
                for (int c = 0; c < 3; ++c)
                    o->ad_stf[((int64_t)c * (g_nad + 1) + idx) * o->p.nt + t] =
                        (real)g_adstf[((int64_t)k * o->p.nt + t) * 3 + c];
            real d = o->xmax - o->xmin;
            for (int i = 0; i <= o->nx; ++i)
                for (int n = 0; n < NN; ++n)
                    if ((real)fabs((double)(a1 - o->x[i * NN + n])) < d) {
                        o->ad_is[3 * idx] = i; d = (real)fabs((double)(a1 - o->x[i * NN + n]));
                    }
            int ii = o->ad_is[3 * idx];
            o->ad_loc[3 * idx] = (((real)2 * a1 - o->x[ii * NN + LPD]) - o->x[ii * NN]) /
                                 (o->x[ii * NN + LPD] - o->x[ii * NN]);
            d = o->ymax - o->ymin;
            for (int j = 0; j <= o->ny; ++j)
                for (int n = 0; n < NN; ++n)
                    if ((real)fabs((double)(a2 - o->y[j * NN + n])) < d) {
                        o->ad_is[3 * idx + 1] = j; d = (real)fabs((double)(a2 - o->y[j * NN + n]));
                    }
            int jj = o->ad_is[3 * idx + 1];
            o->ad_loc[3 * idx + 1] = (((real)2 * a2 - o->y[jj * NN + LPD]) - o->y[jj * NN]) /
                                     (o->y[jj * NN + LPD] - o->y[jj * NN]);
            d = o->zmax - o->zmin;
            for (int k2 = 0; k2 <= o->nz; ++k2)
                for (int n = 0; n < NN; ++n)
                    if ((real)fabs((double)((o->zmax_global - a3) - o->z[k2 * NN + n])) < d) {
                        o->ad_is[3 * idx + 2] = k2; d = (real)fabs((double)((o->zmax_global - a3) - o->z[k2 * NN + n]));
                    }
            int kk = o->ad_is[3 * idx + 2];
            real z4 = o->z[kk * NN + LPD], z0 = o->z[kk * NN];
            o->ad_loc[3 * idx + 2] = ((z4 + z0) - (real)2 * (o->zmax_global - a3)) / (z0 - z4);
        }
    }
}

static void o_init_pml(o_sim *o) {
    const real alpha = (real)1.2;
    const int pml = o->p.pml;
    int64_t NP = o->NP;
    memset(o->prof_x, 0, sizeof(real) * NP);
    memset(o->prof_y, 0, sizeof(real) * NP);
    memset(o->prof_z, 0, sizeof(real) * NP);
    /* lower z boundary (iz_multi==1), init.f90:603-614; upper z disabled (:586) */
    for (int k = o->nz - pml + 1; k <= o->nz; ++k)
        for (int n = 0; n < NN; ++n) {
            real d = ((o->zmin + (real)pml * o->dz) - o->z[k * NN + n]) / ((real)pml * o->dz);
            real v = (alpha * d) * d;
            for (int i = 0; i <= o->nx; ++i) for (int j = 0; j <= o->ny; ++j)
                for (int l = 0; l < NN; ++l) for (int m = 0; m < NN; ++m)
                    o->prof_z[LIDX(o, i, j, k, l, m, n)] = v;
        }
    /* left / right x, :618-644 */
    for (int side = 0; side < 2; ++side) {
        int i0 = side == 0 ? 0 : o->nx - pml + 1, i1 = side == 0 ? pml - 1 : o->nx;
        for (int i = i0; i <= i1; ++i)
            for (int n = 0; n < NN; ++n) {
                real d = side == 0 ? ((o->xmin + (real)pml * o->dx) - o->x[i * NN + n]) / ((real)pml * o->dx)
                                   : ((o->xmax - o->dx * (real)pml) - o->x[i * NN + n]) / ((real)pml * o->dx);
                real v = (alpha * d) * d;
                for (int j = 0; j <= o->ny; ++j) for (int k = 0; k <= o->nz; ++k)
                    for (int m = 0; m < NN; ++m) for (int l = 0; l < NN; ++l)
                        o->prof_x[LIDX(o, i, j, k, n, m, l)] = v;
            }
    }
    /* left / right y, :648-674 */
    for (int side = 0; side < 2; ++side) {
        int j0 = side == 0 ? 0 : o->ny - pml + 1, j1 = side == 0 ? pml - 1 : o->ny;
        for (int j = j0; j <= j1; ++j)
            for (int n = 0; n < NN; ++n) {
                real d = side == 0 ? ((o->ymin + (real)pml * o->dy) - o->y[j * NN + n]) / ((real)pml * o->dy)
                                   : ((o->ymax - o->dy * (real)pml) - o->y[j * NN + n]) / ((real)pml * o->dy);
                real v = (alpha * d) * d;
                for (int i = 0; i <= o->nx; ++i) for (int k = 0; k <= o->nz; ++k)
                    for (int l = 0; l < NN; ++l) for (int m = 0; m < NN; ++m)
                        o->prof_y[LIDX(o, i, j, k, l, n, m)] = v;
            }
    }
    /* normalise the corners, :678-686 */
    for (int64_t q = 0; q < NP; ++q) {
        real pr = (o->prof_x[q] + o->prof_y[q]) + o->prof_z[q];
        o->prof_z[q] = ((real)2 * o->prof_z[q]) / ((real)1 + pr / alpha);
        o->prof_y[q] = ((real)2 * o->prof_y[q]) / ((real)1 + pr / alpha);
        o->prof_x[q] = ((real)2 * o->prof_x[q]) / ((real)1 + pr / alpha);
        pr = (o->prof_x[q] + o->prof_y[q]) + o->prof_z[q];
        o->prof[q] = pr;
        real arg = -(((real)0.1 * pr) * pr);
        o->taper[q] = sizeof(real) == sizeof(float) ? (real)expf((float)arg) : (real)exp((double)arg);
    }
    /* local -> global, last writer in loop order i,j,k,in,jn,kn wins (:690-709) */
    for (int i = 0; i <= o->nx; ++i) for (int j = 0; j <= o->ny; ++j) for (int k = 0; k <= o->nz; ++k)
        for (int l = 0; l < NN; ++l) for (int m = 0; m < NN; ++m) for (int n = 0; n < NN; ++n) {
            int64_t g = GIDX(o, i * LPD + l, j * LPD + m, k * LPD + n);
            int64_t q = LIDX(o, i, j, k, l, m, n);
            o->prof_global[g] = o->prof[q];
            o->taper_global[g] = o->taper[q];
        }
}

/* ses3d_init: zero state, coefficients, locations, PML. */
void o_init(o_sim *o) {
    int64_t NP = o->NP, NG = o->NG;
    for (int c = 0; c < 9; ++c) { memset(o->du[c], 0, sizeof(real) * NP); memset(o->sp[c], 0, sizeof(real) * NP); }
    if (o->p.is_diss)
        for (int c = 0; c < 6; ++c) for (int k = 0; k < 3; ++k) memset(o->Mv[c][k], 0, sizeof(real) * NP);
    memset(o->vxg, 0, sizeof(real) * NG); memset(o->vyg, 0, sizeof(real) * NG); memset(o->vzg, 0, sizeof(real) * NG);
    memset(o->grad_rho, 0, sizeof(real) * NP); memset(o->grad_cp, 0, sizeof(real) * NP);
    memset(o->grad_csh, 0, sizeof(real) * NP); memset(o->grad_csv, 0, sizeof(real) * NP);
    memset(o->grad_Q_mu, 0, sizeof(real) * NP);
    memset(o->src, 0, sizeof(o->src));
    o_init_coeffs(o);
    o_locate(o);
    o_init_pml(o);
}

/* ---------------------------------------------------------------------- */
/* time step: SOURCE/ses3d_evolution.f90:7-373                              */
/* ---------------------------------------------------------------------- */

enum { EXX, EYY, EZZ, EXY, EYX, EXZ, EZX, EYZ, EZY };
/* du index: dxux dyuy dzuz dxuy dyux dxuz dzux dyuz dzuy (same slots as e) */
/* sp index: sxx syy szz sxy syx sxz szx syz szy (same slots) */

static void o_evolution(o_sim *o, int it) {
    const int nx = o->nx, ny = o->ny, nz = o->nz;
    const real dt = o->dt, ispml = (real)1.0;
    const int diss = o->p.is_diss;
    const int af = o->p.adjoint_flag;
    /* gather, :27-51 */
    for (int n = 0; n < NN; ++n) for (int m = 0; m < NN; ++m) for (int l = 0; l < NN; ++l)
        for (int k = 0; k <= nz; ++k) for (int j = 0; j <= ny; ++j) for (int i = 0; i <= nx; ++i) {
            int64_t q = LIDX(o, i, j, k, l, m, n), g = GIDX(o, i * LPD + l, j * LPD + m, k * LPD + n);
            o->vx[q] = o->vxg[g]; o->vy[q] = o->vyg[g]; o->vz[q] = o->vzg[g];
        }
    /* velocity gradients, :57-87 */
    real d1[NN][NN], d2[NN][NN], d3[NN][NN];
    for (int i = 0; i < NN; ++i) for (int q = 0; q < NN; ++q) {
        d3[q][i] = ((real)2.0 * o->dl[q][i]) / o->dz;
        d2[q][i] = ((real)2.0 * o->dl[q][i]) / o->dy;
        d1[q][i] = ((real)2.0 * o->dl[q][i]) / o->dx;
    }
    int64_t NE = o->NE;
    #pragma omp parallel for collapse(3) schedule(static)
    for (int n = 0; n < NN; ++n) for (int m = 0; m < NN; ++m) for (int l = 0; l < NN; ++l)
        for (int64_t el = 0; el < NE; ++el) {
            int64_t q0 = el + NE * (l + NN * (m + NN * (int64_t)n));
            real ezz = 0, ezy = 0, ezx = 0, eyy = 0, eyx = 0, eyz = 0, exx = 0, exy = 0, exz = 0;
            for (int q = 0; q < NN; ++q) {
                int64_t qz = el + NE * (l + NN * (m + NN * (int64_t)q));
                int64_t qy = el + NE * (l + NN * (q + NN * (int64_t)n));
                int64_t qx = el + NE * (q + NN * (m + NN * (int64_t)n));
                ezz = ezz + o->vz[qz] * d3[q][n]; ezy = ezy + o->vy[qz] * d3[q][n]; ezx = ezx + o->vx[qz] * d3[q][n];
                eyy = eyy + o->vy[qy] * d2[q][m]; eyx = eyx + o->vx[qy] * d2[q][m]; eyz = eyz + o->vz[qy] * d2[q][m];
                exx = exx + o->vx[qx] * d1[q][l]; exy = exy + o->vy[qx] * d1[q][l]; exz = exz + o->vz[qx] * d1[q][l];
            }
            o->e[EZZ][q0] = ezz; o->e[EZY][q0] = ezy; o->e[EZX][q0] = ezx;
            o->e[EYY][q0] = eyy; o->e[EYX][q0] = eyx; o->e[EYZ][q0] = eyz;
            o->e[EXX][q0] = exx; o->e[EXY][q0] = exy; o->e[EXZ][q0] = exz;
        }
    /* pointwise: spherical terms (:93-103), integration (:109-117), APML
     * (:123-138), stress rates (:146-202), stress integration (:208-219) */
    #pragma omp parallel for collapse(3) schedule(static)
    for (int n = 0; n < NN; ++n) for (int m = 0; m < NN; ++m) for (int l = 0; l < NN; ++l)
        for (int64_t el = 0; el < NE; ++el) {
            int i = (int)(el % o->NX), k = (int)(el / ((int64_t)o->NX * o->NY));
            int64_t q = el + NE * (l + NN * (m + NN * (int64_t)n));
            real r = o->z[k * NN + n], st = o->sin_x[i * NN + l];
            real ct = o->cos_x[i * NN + l] / o->sin_x[i * NN + l];
            real rst = r * st;
            real vxq = o->vx[q], vyq = o->vy[q], vzq = o->vz[q];
            real ezz = -o->e[EZZ][q];
            real eyy = o->e[EYY][q] / rst + (vzq + vxq * ct) / r;
            real exx = (vzq + o->e[EXX][q]) / r;
            real ezx = -o->e[EZX][q];
            real exz = (o->e[EXZ][q] - vxq) / r;
            real ezy = -o->e[EZY][q];
            real eyz = -vyq / r + o->e[EYZ][q] / rst;
            real exy = o->e[EXY][q] / r;
            real eyx = -((vyq * ct) / r) + o->e[EYX][q] / rst;
            real *du = NULL; (void)du;
            o->du[EZZ][q] = o->du[EZZ][q] + dt * ezz;
            o->du[EYY][q] = o->du[EYY][q] + dt * eyy;
            o->du[EXX][q] = o->du[EXX][q] + dt * exx;
            o->du[EZX][q] = o->du[EZX][q] + dt * ezx;
            o->du[EXZ][q] = o->du[EXZ][q] + dt * exz;
            o->du[EZY][q] = o->du[EZY][q] + dt * ezy;
            o->du[EYZ][q] = o->du[EYZ][q] + dt * eyz;
            o->du[EXY][q] = o->du[EXY][q] + dt * exy;
            o->du[EYX][q] = o->du[EYX][q] + dt * eyx;
            real px = o->prof_x[q], py = o->prof_y[q], pz = o->prof_z[q];
            exx = exx + ((pz + py) * o->du[EXX][q]) * ispml;
            eyy = eyy + ((px + pz) * o->du[EYY][q]) * ispml;
            ezz = ezz + ((px + py) * o->du[EZZ][q]) * ispml;
            exy = exy + ((py + pz) * o->du[EXY][q]) * ispml;
            eyx = eyx + ((px + pz) * o->du[EYX][q]) * ispml;
            exz = exz + ((py + pz) * o->du[EXZ][q]) * ispml;
            ezx = ezx + ((px + py) * o->du[EZX][q]) * ispml;
            eyz = eyz + ((px + pz) * o->du[EYZ][q]) * ispml;
            ezy = ezy + ((px + py) * o->du[EZY][q]) * ispml;
            exy = (exy + eyx) / (real)2.0;
            exz = (exz + ezx) / (real)2.0;
            eyz = (eyz + ezy) / (real)2.0;
            real sxx, syy, szz, sxy, sxz, syz;
            real mu = o->mu[q], A = o->A[q], B = o->B[q], C = o->C[q];
            if (diss) {
                real tau = o->tau[q], mut = o->mu_tau[q], kap = o->kappa[q];
                real Mdx = (((real)0 + o->Mv[0][0][q]) + o->Mv[0][1][q]) + o->Mv[0][2][q];
                real Mdy = (((real)0 + o->Mv[1][0][q]) + o->Mv[1][1][q]) + o->Mv[1][2][q];
                real Mdz = (((real)0 + o->Mv[2][0][q]) + o->Mv[2][1][q]) + o->Mv[2][2][q];
                real Md = (Mdx + Mdy) + Mdz;
                real L = (kap - ((real)2.0 * mut) / (real)3.0) * ((exx + eyy) + ezz) -
                         ((((real)2.0 * mu) * tau) * Md) / (real)3.0;
                sxx = (((L + ((real)2.0 * mut) * exx) + (((real)2.0 * mu) * tau) * Mdx) + C * ezz) + A * (eyy + exx);
                syy = (((L + ((real)2.0 * mut) * eyy) + (((real)2.0 * mu) * tau) * Mdy) + C * ezz) + A * (eyy + exx);
                szz = ((L + ((real)2.0 * mut) * ezz) + (((real)2.0 * mu) * tau) * Mdz) + C * (eyy + exx);
                Mdx = (((real)0 + o->Mv[3][0][q]) + o->Mv[3][1][q]) + o->Mv[3][2][q];
                Mdy = (((real)0 + o->Mv[4][0][q]) + o->Mv[4][1][q]) + o->Mv[4][2][q];
                Mdz = (((real)0 + o->Mv[5][0][q]) + o->Mv[5][1][q]) + o->Mv[5][2][q];
                sxy = ((real)2.0 * mut) * exy + (((real)2.0 * mu) * tau) * Mdx;
                sxz = (((real)2.0 * mut) * exz + (((real)2.0 * mu) * tau) * Mdy) + ((real)2.0 * B) * exz;
                syz = (((real)2.0 * mut) * eyz + (((real)2.0 * mu) * tau) * Mdz) + ((real)2.0 * B) * eyz;
            } else {
                real L = o->lambda[q] * ((exx + eyy) + ezz);
                sxx = ((L + ((real)2 * mu) * exx) + C * ezz) + A * (eyy + exx);
                syy = ((L + ((real)2 * mu) * eyy) + C * ezz) + A * (eyy + exx);
                szz = (L + ((real)2 * mu) * ezz) + C * (eyy + exx);
                sxy = ((real)2.0 * mu) * exy;
                sxz = ((real)2.0 * mu) * exz + ((real)2.0 * B) * exz;
                syz = ((real)2.0 * mu) * eyz + ((real)2.0 * B) * eyz;
            }
            /* keep post-APML, symmetrised rates for the memory variables */
            o->e[EXX][q] = exx; o->e[EYY][q] = eyy; o->e[EZZ][q] = ezz;
            o->e[EXY][q] = exy; o->e[EXZ][q] = exz; o->e[EYZ][q] = eyz;
            real **s = o->sp;
            s[EXX][q] = s[EXX][q] + dt * (sxx - (ispml * px) * s[EXX][q]);
            s[EYY][q] = s[EYY][q] + dt * (syy - (ispml * py) * s[EYY][q]);
            s[EZZ][q] = s[EZZ][q] + dt * (szz - (ispml * pz) * s[EZZ][q]);
            s[EXY][q] = s[EXY][q] + dt * (sxy - (ispml * px) * s[EXY][q]);
            s[EYX][q] = s[EYX][q] + dt * (sxy - (ispml * py) * s[EYX][q]);
            s[EXZ][q] = s[EXZ][q] + dt * (sxz - (ispml * px) * s[EXZ][q]);
            s[EZX][q] = s[EZX][q] + dt * (sxz - (ispml * pz) * s[EZX][q]);
            s[EYZ][q] = s[EYZ][q] + dt * (syz - (ispml * py) * s[EYZ][q]);
            s[EZY][q] = s[EZY][q] + dt * (syz - (ispml * pz) * s[EZY][q]);
        }
    /* moment source, misc.f90:115-146 (src arrays zero except source element) */
    int have_src = 0;
    memset(o->src, 0, sizeof(o->src));
    if ((af == 0 || af == 1) && o->is_source && o->source_type == 10) {
        have_src = 1;
        real sot = o->so[it - 1];
        for (int in = 0; in < NN; ++in) for (int jn = 0; jn < NN; ++jn) for (int kn = 0; kn < NN; ++kn) {
            real delta = ((lgll(in, o->xxs_loc) * lgll(jn, o->yys_loc)) * lgll(kn, o->zzs_loc)) /
                         ((o->w[in] * o->w[jn]) * o->w[kn]);
            real zs = o->z[o->isz * NN + kn];
            real sx_ = sizeof(real) == sizeof(float) ? (real)sinf((float)o->x[o->isx * NN + in])
                                                     : (real)sin((double)o->x[o->isx * NN + in]);
            delta = delta / (((zs * zs) * sx_) * o->Jac);
            int nd = in + NN * (jn + NN * kn);
            o->src[EXX][nd] = (real)0 - (o->MOM[0] * delta) * sot;
            o->src[EYY][nd] = (real)0 - (o->MOM[1] * delta) * sot;
            o->src[EZZ][nd] = (real)0 - (o->MOM[2] * delta) * sot;
            o->src[EXY][nd] = (real)0 - (o->MOM[3] * delta) * sot;
            o->src[EYX][nd] = (real)0 - (o->MOM[3] * delta) * sot;
            o->src[EXZ][nd] = (real)0 - (o->MOM[4] * delta) * sot;
            o->src[EZX][nd] = (real)0 - (o->MOM[4] * delta) * sot;
            o->src[EYZ][nd] = (real)0 - (o->MOM[5] * delta) * sot;
            o->src[EZY][nd] = (real)0 - (o->MOM[5] * delta) * sot;
        }
    }
    /* stress divergence, :235-270 */
    real dum1[NN][NN], dum2[NN][NN], dum3[NN][NN];
    for (int a = 0; a < NN; ++a) for (int b = 0; b < NN; ++b) {
        dum1[a][b] = ((((real)2 / o->dx) * o->Jac) * o->w[a]) * o->w[b];   /* (j,k) */
        dum2[a][b] = ((((real)2 / o->dy) * o->Jac) * o->w[a]) * o->w[b];   /* (i,k) */
        dum3[a][b] = -(((((real)2 / o->dz) * o->Jac) * o->w[a]) * o->w[b]); /* (i,j) */
    }
    #pragma omp parallel for collapse(3) schedule(static)
    for (int kk = 0; kk < NN; ++kk) for (int jj = 0; jj < NN; ++jj) for (int ii = 0; ii < NN; ++ii)
        for (int64_t el = 0; el < NE; ++el) {
            int ei = (int)(el % o->NX), ej = (int)((el / o->NX) % o->NY), ek = (int)(el / ((int64_t)o->NX * o->NY));
            int is_src_el = have_src && ei == o->isx && ej == o->isy && ek == o->isz;
            int64_t qo = el + NE * (ii + NN * (jj + NN * (int64_t)kk));
            real ax = 0, ay = 0, az = 0;
            for (int q = 0; q < NN; ++q) {
                real wx = (((o->w[q] * o->dl[ii][q]) * o->z[ek * NN + kk]) * o->sin_x[ei * NN + q]) * dum1[jj][kk];
                real wy = ((o->w[q] * o->dl[jj][q]) * o->z[ek * NN + kk]) * dum2[ii][kk];
                real zq = o->z[ek * NN + q];
                real wz = ((((o->w[q] * o->dl[kk][q]) * zq) * zq) * o->sin_x[ei * NN + ii]) * dum3[ii][jj];
                int64_t pz_ = el + NE * (ii + NN * (jj + NN * (int64_t)q));
                int64_t py_ = el + NE * (ii + NN * (q + NN * (int64_t)kk));
                int64_t px_ = el + NE * (q + NN * (jj + NN * (int64_t)kk));
                int nz_ = ii + NN * (jj + NN * q), ny_ = ii + NN * (q + NN * kk), nx_ = q + NN * (jj + NN * kk);
                real szx = o->sp[EZX][pz_], syx = o->sp[EYX][py_], sxx = o->sp[EXX][px_];
                real szy = o->sp[EZY][pz_], syy = o->sp[EYY][py_], sxy = o->sp[EXY][px_];
                real szz = o->sp[EZZ][pz_], syz = o->sp[EYZ][py_], sxz = o->sp[EXZ][px_];
                if (is_src_el) {
                    szx = szx + o->src[EZX][nz_]; syx = syx + o->src[EYX][ny_]; sxx = sxx + o->src[EXX][nx_];
                    szy = szy + o->src[EZY][nz_]; syy = syy + o->src[EYY][ny_]; sxy = sxy + o->src[EXY][nx_];
                    szz = szz + o->src[EZZ][nz_]; syz = syz + o->src[EYZ][ny_]; sxz = sxz + o->src[EXZ][nx_];
                }
                ax = ((ax + szx * wz) + syx * wy) + sxx * wx;
                ay = ((ay + szy * wz) + syy * wy) + sxy * wx;
                az = ((az + szz * wz) + syz * wy) + sxz * wx;
            }
            o->sx[qo] = ax; o->sy[qo] = ay; o->sz[qo] = az;
        }
    /* single forces, misc.f90:29-92 */
    if ((af == 0 || af == 1) && o->is_source && o->source_type < 10) {
        real sot = o->so[it - 1];
        for (int in = 0; in < NN; ++in) for (int jn = 0; jn < NN; ++jn) for (int kn = 0; kn < NN; ++kn) {
            real delta = (lgll(in, o->xxs_loc) * lgll(jn, o->yys_loc)) * lgll(kn, o->zzs_loc);
            int64_t q = LIDX(o, o->isx, o->isy, o->isz, in, jn, kn);
            if (o->source_type == 1) o->sx[q] = o->sx[q] - delta * sot;
            else if (o->source_type == 2) o->sy[q] = o->sy[q] - delta * sot;
            else if (o->source_type == 3) o->sz[q] = o->sz[q] - delta * sot;
        }
    }
    if (af == 2) {
        for (int idx = 0; idx < o->nad; ++idx)
            for (int in = 0; in < NN; ++in) for (int jn = 0; jn < NN; ++jn) for (int kn = 0; kn < NN; ++kn) {
                real delta = (lgll(in, o->ad_loc[3 * idx]) * lgll(jn, o->ad_loc[3 * idx + 1])) *
                             lgll(kn, o->ad_loc[3 * idx + 2]);
                int64_t q = LIDX(o, o->ad_is[3 * idx], o->ad_is[3 * idx + 1], o->ad_is[3 * idx + 2], in, jn, kn);
                int64_t s0 = (int64_t)(g_nad + 1) * o->p.nt;
                o->sx[q] = o->sx[q] - delta * o->ad_stf[0 * s0 + (int64_t)idx * o->p.nt + (it - 1)];
                o->sy[q] = o->sy[q] - delta * o->ad_stf[1 * s0 + (int64_t)idx * o->p.nt + (it - 1)];
                o->sz[q] = o->sz[q] - delta * o->ad_stf[2 * s0 + (int64_t)idx * o->p.nt + (it - 1)];
            }
    }
    /* assembly, :282-310 (scatter in loop order i,in,j,jn,k,kn) */
    memset(o->sxg, 0, sizeof(real) * o->NG); memset(o->syg, 0, sizeof(real) * o->NG); memset(o->szg, 0, sizeof(real) * o->NG);
    for (int i = 0; i <= nx; ++i) for (int in = 0; in < NN; ++in)
        for (int j = 0; j <= ny; ++j) for (int jn = 0; jn < NN; ++jn)
            for (int k = 0; k <= nz; ++k) for (int kn = 0; kn < NN; ++kn) {
                int64_t g = GIDX(o, i * LPD + in, j * LPD + jn, k * LPD + kn);
                int64_t q = LIDX(o, i, j, k, in, jn, kn);
                o->sxg[g] = o->sxg[g] + o->sx[q];
                o->syg[g] = o->syg[g] + o->sy[q];
                o->szg[g] = o->szg[g] + o->sz[q];
            }
    /* mass, velocity update, :318-328; taper of v, :366 */
    #pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < o->NG; ++g) {
        real ax = o->sxg[g] / o->MM_global[g], ay = o->syg[g] / o->MM_global[g], az = o->szg[g] / o->MM_global[g];
        real pg = ispml * o->prof_global[g];
        real vx = o->vxg[g] + dt * (ax - pg * o->vxg[g]);
        real vy = o->vyg[g] + dt * (ay - pg * o->vyg[g]);
        real vz = o->vzg[g] + dt * (az - pg * o->vzg[g]);
        o->vxg[g] = o->taper_global[g] * vx; o->vyg[g] = o->taper_global[g] * vy; o->vzg[g] = o->taper_global[g] * vz;
    }
    /* memory variables, :334-360; stress taper, :368-371 */
    #pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < o->NP; ++q) {
        if (diss) {
            const int eidx[6] = {EXX, EYY, EZZ, EXY, EXZ, EYZ};
            for (int k = 0; k < 3; ++k) {
                real L = (dt * o->D_p[k]) / o->tau_p[k];
                for (int c = 0; c < 6; ++c) {
                    real M = o->Mv[c][k][q];
                    o->Mv[c][k][q] = (M - (dt * M) / o->tau_p[k]) - L * o->e[eidx[c]][q];
                }
            }
        }
        real t = o->taper[q];
        for (int c = 0; c < 9; ++c) o->sp[c][q] = o->sp[c][q] * t;
    }
}

/* record_seismograms, ses3d_output.f90:475-495 (local v = pre-update v) */
static void o_record(o_sim *o, int it) {
    for (int idx = 0; idx < o->nr; ++idx) {
        real *sx = &o->seis[(0 * (int64_t)(o->nr + 1) + idx) * o->p.nt + (it - 1)];
        real *sy = &o->seis[(1 * (int64_t)(o->nr + 1) + idx) * o->p.nt + (it - 1)];
        real *sz = &o->seis[(2 * (int64_t)(o->nr + 1) + idx) * o->p.nt + (it - 1)];
        for (int i = 0; i < NN; ++i) for (int j = 0; j < NN; ++j) for (int k = 0; k < NN; ++k) {
            real d = (lgll(i, o->rstd[3 * idx]) * lgll(j, o->rstd[3 * idx + 1])) * lgll(k, o->rstd[3 * idx + 2]);
            int64_t q = LIDX(o, o->rx[idx], o->ry[idx], o->rz[idx], i, j, k);
            *sx = *sx + o->vx[q] * d; *sy = *sy + o->vy[q] * d; *sz = *sz + o->vz[q] * d;
        }
    }
}

/* ses3d_store with deg < lpd (ses3d_output.f90:36-130) for one field in
 * reference layout (NE elements fastest, then the 125 nodes): the degree-4
 * interpolant sampled at the degree-deg knots, c[e + NE*(i + D(j + D k))] =
 * sum over ip (x), iq (y), ir (z), in that order from 0, of
 * f(e,ip,iq,ir) * ((L4_ip(k_i) L4_iq(k_j)) L4_ir(k_k)). */
void o_compress(int deg, int64_t NE, const real *f, real *c) {
    const int D = deg + 1;
    real Lc[NN][4];
    for (int ip = 0; ip < NN; ++ip)
        for (int i = 0; i < D; ++i) Lc[ip][i] = lgll(ip, knot_n(deg, i));
    for (int k = 0; k < D; ++k) for (int j = 0; j < D; ++j) for (int i = 0; i < D; ++i)
        for (int64_t e = 0; e < NE; ++e) {
            real acc = (real)0.0;
            for (int ip = 0; ip < NN; ++ip) for (int iq = 0; iq < NN; ++iq) for (int ir = 0; ir < NN; ++ir) {
                const real w = (Lc[ip][i] * Lc[iq][j]) * Lc[ir][k];
                acc = acc + f[e + NE * (ip + NN * (iq + NN * ir))] * w;
            }
            c[e + NE * (i + D * (j + D * k))] = acc;
        }
}

/* ses3d_restore with deg < lpd (ses3d_output.f90:242-265): back to the GLL
 * nodes, f(e,l,m,n) = sum over il (x), im (y), in (z), in that order from 0,
 * of c(e,il,im,in) * ((Ld_il(x_l) Ld_im(x_m)) Ld_in(x_n)). */
void o_restore(int deg, int64_t NE, const real *c, real *f) {
    const int D = deg + 1;
    real Ld[4][NN];
    for (int il = 0; il < D; ++il)
        for (int l = 0; l < NN; ++l) Ld[il][l] = lgll_n(deg, il, KN[l]);
    for (int n = 0; n < NN; ++n) for (int m = 0; m < NN; ++m) for (int l = 0; l < NN; ++l)
        for (int64_t e = 0; e < NE; ++e) {
            real acc = (real)0.0;
            for (int il = 0; il < D; ++il) for (int im = 0; im < D; ++im) for (int in = 0; in < D; ++in) {
                const real w = (Ld[il][l] * Ld[im][m]) * Ld[in][n];
                acc = acc + c[e + NE * (il + D * (im + D * in))] * w;
            }
            f[e + NE * (l + NN * (m + NN * n))] = acc;
        }
}

/* forward store, ses3d_main.f90:142-166: v (pre-update), d, symmetric strain */
static void o_store(o_sim *o, int it) {
    int h = o->n_hist++;
    o->hist = (float **)realloc(o->hist, sizeof(float *) * o->n_hist);
    o->hist_step = (int *)realloc(o->hist_step, sizeof(int) * o->n_hist);
    o->hist_step[h] = it;
    int64_t NP = o->NP;
    if (o->hdeg < LPD) {
        const int D = o->hdeg + 1;
        const int64_t NE = o->NE, NC = (int64_t)D * D * D;
        float *b = (float *)malloc(sizeof(float) * 9 * (size_t)(NE * NC));
        o->hist[h] = b;
        real *f = alloc0(NP), *cf = alloc0(NE * NC);
        for (int c = 0; c < 9; ++c) {
            for (int64_t q = 0; q < NP; ++q)
                f[q] = c == 0 ? o->vx[q] : c == 1 ? o->vy[q] : c == 2 ? o->vz[q]
                     : c == 3 ? o->du[EXX][q] : c == 4 ? o->du[EYY][q] : c == 5 ? o->du[EZZ][q]
                     : c == 6 ? (o->du[EXY][q] + o->du[EYX][q]) / (real)2
                     : c == 7 ? (o->du[EXZ][q] + o->du[EZX][q]) / (real)2
                     : (o->du[EYZ][q] + o->du[EZY][q]) / (real)2;
            o_compress(o->hdeg, NE, f, cf);
            for (int64_t i = 0; i < NE * NC; ++i) b[c * NE * NC + i] = (float)cf[i];
        }
        free(f);
        free(cf);
        o->saving_vector[it] = it;
        return;
    }
    float *b = (float *)malloc(sizeof(float) * 9 * (size_t)o->NP);
    o->hist[h] = b;
    for (int64_t q = 0; q < NP; ++q) {
        b[0 * NP + q] = (float)o->vx[q]; b[1 * NP + q] = (float)o->vy[q]; b[2 * NP + q] = (float)o->vz[q];
        b[3 * NP + q] = (float)o->du[EXX][q]; b[4 * NP + q] = (float)o->du[EYY][q]; b[5 * NP + q] = (float)o->du[EZZ][q];
        b[6 * NP + q] = (float)((o->du[EXY][q] + o->du[EYX][q]) / (real)2);
        b[7 * NP + q] = (float)((o->du[EXZ][q] + o->du[EZX][q]) / (real)2);
        b[8 * NP + q] = (float)((o->du[EYZ][q] + o->du[EZY][q]) / (real)2);
    }
    o->saving_vector[it] = it;
}

/* ses3d_grad (ses3d_grad.f90:25-85) + shear-Q kernel (ses3d_grad_Q.f90:111-128) */
static void o_grad(o_sim *o, o_sim *fwd, int it) {
    int fstep = o->p.nt + 1 - it;
    if (!fwd || fwd->saving_vector[fstep] <= 0) return;
    int h = -1;
    for (int t = 0; t < fwd->n_hist; ++t) if (fwd->hist_step[t] == fstep) h = t;
    if (h < 0) return;
    const float *b = fwd->hist[h];
    int64_t NP = o->NP;
    real *dec = NULL;
    if (fwd->hdeg < LPD) {
        const int D = fwd->hdeg + 1;
        const int64_t NE = o->NE, NC = (int64_t)D * D * D;
        real *cf = alloc0(NE * NC);
        dec = alloc0(9 * NP);
        for (int c = 0; c < 9; ++c) {
            for (int64_t i = 0; i < NE * NC; ++i) cf[i] = (real)b[c * NE * NC + i];
            o_restore(fwd->hdeg, NE, cf, dec + c * NP);
        }
        free(cf);
    }
#define FW(c) (dec ? dec[(int64_t)(c) * NP + q] : (real)b[(int64_t)(c) * NP + q])
    const real sdt = (real)o->p.samp_ad * o->dt, Jac = o->Jac;
    #pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < NP; ++q) {
        real vxf = FW(0), vyf = FW(1), vzf = FW(2);
        real exxf = FW(3), eyyf = FW(4), ezzf = FW(5);
        real exyf = FW(6), exzf = FW(7), eyzf = FW(8);
        real exxa = o->du[EXX][q], eyya = o->du[EYY][q], ezza = o->du[EZZ][q];
        real exya = (real)0.5 * (o->du[EXY][q] + o->du[EYX][q]);
        real exza = (real)0.5 * (o->du[EXZ][q] + o->du[EZX][q]);
        real eyza = (real)0.5 * (o->du[EYZ][q] + o->du[EZY][q]);
        real tra = (exxa + eyya) + ezza, trf = (exxf + eyyf) + ezzf;
        real ede = ((((exxa * exxf + eyya * eyyf) + ezza * ezzf) + ((real)2.0 * exya) * exyf) +
                    ((real)2.0 * exza) * exzf) + ((real)2.0 * eyza) * eyzf;
        real rho = o->rho[q], cp = o->cp[q], cs = o->cs[q];
        o->grad_cp[q] = o->grad_cp[q] + (sdt * ((((real)2.0 * rho) * cp) * trf * tra)) / Jac;
        o->grad_rho[q] = o->grad_rho[q] + (sdt * ((vxf * o->vx[q] + vyf * o->vy[q]) + vzf * o->vz[q])) / Jac;
        o->grad_rho[q] = o->grad_rho[q] +
            (sdt * (((cp * cp - ((real)2.0 * cs) * cs) * trf) * tra + (((real)2.0 * cs) * cs) * ede)) / Jac;
        o->grad_csh[q] = o->grad_csh[q] +
            (sdt * ((((real)4.0 * rho) * cs) * (((ede - tra * trf) - ((real)2.0 * exzf) * exza) - ((real)2.0 * eyzf) * eyza))) / Jac;
        o->grad_csv[q] = o->grad_csv[q] + (sdt * ((((real)8.0 * rho) * cs) * (exzf * exza + eyzf * eyza))) / Jac;
        o->fw[0][q] = exxf;  /* keep a copy for inspection */
        if (o->p.is_diss) {
            for (int k = 0; k < 3; ++k) {
                real Mxx = o->Mv[0][k][q], Myy = o->Mv[1][k][q], Mzz = o->Mv[2][k][q];
                real Mxy = o->Mv[3][k][q], Mxz = o->Mv[4][k][q], Myz = o->Mv[5][k][q];
                real md = (((((Mxx * exxf + Myy * eyyf) + Mzz * ezzf) + Mxy * exyf) + Mxz * exzf) + Myz * eyzf) -
                          (trf * ((Mxx + Myy) + Mzz)) / (real)3.0;
                o->grad_Q_mu[q] = o->grad_Q_mu[q] + ((sdt * md) * o->tau_p[k]) / Jac;
            }
        }
    }
#undef FW
    free(dec);
}

/* One iteration of the reference time loop (ses3d_main.f90:95-178).
 * fwd: the forward-for-adjoint sim whose stored history the adjoint uses. */
void o_step(o_sim *o, int it, o_sim *fwd) {
    o_evolution(o, it);
    int af = o->p.adjoint_flag;
    if (af == 0 || af == 1) o_record(o, it);
    if (af == 1 && ((it % o->p.samp_ad) == 0 || it == 1 || it == o->p.nt)) o_store(o, it);
    if (af == 2) {
        o_grad(o, fwd, it);
        if (it == o->p.nt && o->p.is_diss) {   /* ses3d_grad_Q.f90:151-153 */
            for (int64_t q = 0; q < o->NP; ++q)
                o->grad_Q_mu[q] = (((real)2.0 * o->mu[q]) * o->grad_Q_mu[q]) / o->QQ[q];
        }
    }
}

void o_run(o_sim *o, int it0, int n, o_sim *fwd) {
    for (int it = it0; it < it0 + n; ++it) o_step(o, it, fwd);
}

/* ------------------------------ accessors ------------------------------ */

int64_t o_np(o_sim *o) { return o->NP; }
int64_t o_ng(o_sim *o) { return o->NG; }
int o_nr(o_sim *o) { return o->nr; }
int o_nad(o_sim *o) { return o->nad; }

static void cp_out(double *dst, const real *src, int64_t n) {
    for (int64_t i = 0; i < n; ++i) dst[i] = (double)src[i];
}

/* Named field access as float64 copies (exact for both float and double). */
int o_get(o_sim *o, const char *name, double *out) {
    int64_t NP = o->NP, NG = o->NG;
    if (!strcmp(name, "vxg")) { cp_out(out, o->vxg, NG); return 0; }
    if (!strcmp(name, "vyg")) { cp_out(out, o->vyg, NG); return 0; }
    if (!strcmp(name, "vzg")) { cp_out(out, o->vzg, NG); return 0; }
    if (!strcmp(name, "MM_global")) { cp_out(out, o->MM_global, NG); return 0; }
    if (!strcmp(name, "prof_global")) { cp_out(out, o->prof_global, NG); return 0; }
    if (!strcmp(name, "taper_global")) { cp_out(out, o->taper_global, NG); return 0; }
    if (!strcmp(name, "vx")) { cp_out(out, o->vx, NP); return 0; }
    if (!strcmp(name, "vy")) { cp_out(out, o->vy, NP); return 0; }
    if (!strcmp(name, "vz")) { cp_out(out, o->vz, NP); return 0; }
    if (!strcmp(name, "prof")) { cp_out(out, o->prof, NP); return 0; }
    if (!strcmp(name, "taper")) { cp_out(out, o->taper, NP); return 0; }
    if (!strcmp(name, "prof_x")) { cp_out(out, o->prof_x, NP); return 0; }
    if (!strcmp(name, "prof_y")) { cp_out(out, o->prof_y, NP); return 0; }
    if (!strcmp(name, "prof_z")) { cp_out(out, o->prof_z, NP); return 0; }
    if (!strcmp(name, "grad_rho")) { cp_out(out, o->grad_rho, NP); return 0; }
    if (!strcmp(name, "grad_cp")) { cp_out(out, o->grad_cp, NP); return 0; }
    if (!strcmp(name, "grad_csh")) { cp_out(out, o->grad_csh, NP); return 0; }
    if (!strcmp(name, "grad_csv")) { cp_out(out, o->grad_csv, NP); return 0; }
    if (!strcmp(name, "grad_Q_mu")) { cp_out(out, o->grad_Q_mu, NP); return 0; }
    static const char *dun[9] = {"dxux", "dyuy", "dzuz", "dxuy", "dyux", "dxuz", "dzux", "dyuz", "dzuy"};
    static const char *spn[9] = {"sxx", "syy", "szz", "sxy", "syx", "sxz", "szx", "syz", "szy"};
    for (int c = 0; c < 9; ++c) {
        if (!strcmp(name, dun[c])) { cp_out(out, o->du[c], NP); return 0; }
        if (!strcmp(name, spn[c])) { cp_out(out, o->sp[c], NP); return 0; }
    }
    static const char *mn[6] = {"Mxx", "Myy", "Mzz", "Mxy", "Mxz", "Myz"};
    for (int c = 0; c < 6; ++c)
        if (!strcmp(name, mn[c]) && o->p.is_diss) {
            for (int k = 0; k < 3; ++k) cp_out(out + k * NP, o->Mv[c][k], NP);
            return 0;
        }
    if (!strcmp(name, "seismograms")) {   /* (3, nr, nt) */
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < o->nr; ++r)
                cp_out(out + ((int64_t)c * o->nr + r) * o->p.nt,
                       o->seis + ((int64_t)c * (o->nr + 1) + r) * o->p.nt, o->p.nt);
        return 0;
    }
    return -1;
}

/* Scalar setup values: knots, weights, dl, dx, dy, dz, Jac, source location. */
void o_get_setup(o_sim *o, double *out /* >= 64 */) {
    int p = 0;
    for (int i = 0; i < NN; ++i) out[p++] = o->w[i];
    for (int i = 0; i < NN; ++i) for (int j = 0; j < NN; ++j) out[p++] = o->dl[i][j];
    out[p++] = o->dx; out[p++] = o->dy; out[p++] = o->dz; out[p++] = o->Jac;
    out[p++] = o->isx; out[p++] = o->isy; out[p++] = o->isz;
    out[p++] = o->xxs_loc; out[p++] = o->yys_loc; out[p++] = o->zzs_loc; out[p++] = o->is_source;
}

/* Receiver placement: (nr, 6) = rx, ry, rz, xi, eta, zeta. */
void o_get_receivers(o_sim *o, double *out) {
    for (int r = 0; r < o->nr; ++r) {
        out[6 * r] = o->rx[r]; out[6 * r + 1] = o->ry[r]; out[6 * r + 2] = o->rz[r];
        out[6 * r + 3] = o->rstd[3 * r]; out[6 * r + 4] = o->rstd[3 * r + 1]; out[6 * r + 5] = o->rstd[3 * r + 2];
    }
}

/* Forward history access for a stored step, order vx vy vz exx eyy ezz exy
 * exz eyz: 9 x NP float32, or with hdeg < 4 per field the (0:nx,0:ny,0:nz,
 * 0:hdeg,0:hdeg,0:hdeg) values ses3d_store writes (o_history_size floats). */
int o_get_history(o_sim *o, int step, float *out) {
    for (int t = 0; t < o->n_hist; ++t)
        if (o->hist_step[t] == step) {
            memcpy(out, o->hist[t], sizeof(float) * (size_t)o_history_size(o));
            return 0;
        }
    return -1;
}

int o_saving_vector(o_sim *o, int it) { return (it >= 1 && it <= o->p.nt) ? o->saving_vector[it] : 0; }

double o_lgll(int i, double x) { return (double)lgll(i, (real)x); }
double o_dlgll(int i, int j) { return (double)dlgll(i, j); }
