/*
 * CPU oracle for the SWM time-stepping hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A pass-for-pass C restatement of the reference Fortran.  It is the parity
 * checker for the HIP path and the CPU baseline timed by bench.py; nothing in
 * the product links or calls it.  Every routine cites the reference routine
 * (file:line) it follows.  Expression trees are kept in the Fortran
 * left-to-right order and the file is compiled with -ffp-contract=off, so a
 * correct restatement reproduces the gfortran -O2 results bit for bit.
 *
 * Layout: Fortran (Nx,Ny) arrays as C [Ny][Nx], 0-based; periodic neighbour
 * index vectors as built with cshift in domain_module.f90:106-115.
 */
#include "swm_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#define NG 3           /* swm_vars.f90:21 */
#define TRC_NG 3       /* tracer_module.h:12-14 */

typedef struct {       /* t_interpolater2point / t_interpolater4point, calc_lib.f90:89-103 */
    int npt;
    int *iind, *jind;  /* (npt, Nx, Ny) */
    double *w;         /* weight_vec (npt, Nx, Ny) */
} interp_t;

typedef struct {
    double kappa_h;
    double *CH[2];     /* TRC_NLEVEL = 2 */
    double *G_CH[TRC_NG];
    double *C, *C0, *gamma_C, *cons, *impl, *uhc, *vhc, *diff, *forcing;
} tracer_t;

/* memoryChunk, memchunk_module.f90:37-50: a buffer of chunkSize records of a
 * dataset with its time axis.  The dataset is held in memory (file). */
typedef struct memchunk {
    const double *file;     /* nrec records of n doubles */
    size_t n;
    int64_t nrec, chunkSize, fileRec, chunkCounter, nFpC;
    int isInitialised, isPersistent, isConstant;
    double dt, tOffset;
    double *var;            /* chunkSize records */
    double *time;           /* chunkSize */
} memchunk;

struct orc_model {
    orc_desc d;
    int Nx, Ny;
    int64_t itt;
    int *ip0, *ip1, *im1, *jp0, *jp1, *jm1;
    const int8_t *ocean[4];
    double *bc[4];
    const double *cosl[4], *tanl[4], *f[4];
    double *H[4];                 /* private copies; linear variants overwrite them through D aliases */
    double ab_coeff[5][5];        /* ab_coeff(ti, tstep) stored [tstep-1][ti-1] */
    /* state, swm_vars.f90:24-39 */
    double *SWM_u[2], *SWM_v[2], *SWM_eta[2];
    double *G_u[NG], *G_v[NG], *G_eta[NG];
    double *D, *Dh, *Du, *Dv;     /* FN: own arrays; linear: aliases of H[ETA],H[H],H[U],H[V] */
    double *EDens, *Pot, *zeta, *MU, *MV;
    double *u, *v, *eta;          /* master fields (level N0 only is ever read) */
    /* lateral mixing, swm_lateralmixing_module.f90:47-54 */
    double *lat_mixing_u, *lat_mixing_v, *pll_coeff, *plt_coeff, *pll, *plt, *latmix_u, *latmix_v;
    /* forcing, swm_forcing_module.f90:52-57 */
    double *F_x, *F_y, *F_eta, *F_x_const, *F_y_const, *F_eta_const;
    const double *impl_u, *impl_v, *impl_eta, *gamma_sq_u, *gamma_sq_v;
    /* interpolators used on the path (calc_lib.f90:110-119) */
    interp_t eta2u_noland, eta2v_noland, eta2H_noland, u2eta, v2eta, H2u, H2v, v2u, u2v;
    /* tracers */
    int ntrc;
    tracer_t trc[ORC_MAX_TRACERS];
    double *TRC_coeff;            /* (12, Nx, Ny), tracer_module.f90:42 */
    /* memory chunks of the time-interpolated forcing streams */
    memchunk mc[ORC_MAX_STREAMS];
    memchunk dff_mc[3];           /* DYNFROMFILE datasets */
    double *tdf_buf;
    /* SOR, ElSolv_SOR.f90:25-32 */
    double *sor_c;                /* (Nx, Ny, 5) */
    int n_oe[2];
    int *i_oe;                    /* (2, n_odd, 2) */
};

#define IX(m, i, j) ((size_t)(j) * (size_t)(m)->Nx + (size_t)(i))

static double *zalloc(size_t n) { return (double *)calloc(n, sizeof(double)); }
static double *falloc(size_t n, double v) {
    double *p = (double *)malloc(n * sizeof(double));
    for (size_t k = 0; k < n; ++k) p[k] = v;
    return p;
}

/* getOutGrid, calc_lib.f90:1220-1310 (first derivative directions only) */
static int get_out_grid(const orc_model *m, int grid, char dir, const int **ind0, const int **indm1) {
    if (dir == 'x') {
        switch (grid) {
        case ORC_GRID_U:   *ind0 = m->ip1; *indm1 = m->ip0; return ORC_GRID_ETA;
        case ORC_GRID_H:   *ind0 = m->ip1; *indm1 = m->ip0; return ORC_GRID_V;
        case ORC_GRID_V:   *ind0 = m->ip0; *indm1 = m->im1; return ORC_GRID_H;
        default:           *ind0 = m->ip0; *indm1 = m->im1; return ORC_GRID_U;
        }
    }
    switch (grid) {
    case ORC_GRID_U:   *ind0 = m->jp0; *indm1 = m->jm1; return ORC_GRID_H;
    case ORC_GRID_H:   *ind0 = m->jp1; *indm1 = m->jp0; return ORC_GRID_U;
    case ORC_GRID_V:   *ind0 = m->jp1; *indm1 = m->jp0; return ORC_GRID_ETA;
    default:           *ind0 = m->jp0; *indm1 = m->jm1; return ORC_GRID_V;
    }
}

/* set2PointInterpolater, calc_lib.f90:204-280 */
static void set2(orc_model *m, interp_t *it, int from, char dir, int exclude_land) {
    const int *ip, *im, *jp, *jm;
    const int Nx = m->Nx, Ny = m->Ny;
    if (dir == 'x') { get_out_grid(m, from, 'x', &ip, &im); jp = m->jp0; jm = m->jp0; }
    else            { get_out_grid(m, from, 'y', &jp, &jm); ip = m->ip0; im = m->ip0; }
    it->npt = 2;
    it->iind = (int *)malloc(sizeof(int) * 2 * Nx * Ny);
    it->jind = (int *)malloc(sizeof(int) * 2 * Nx * Ny);
    it->w = zalloc((size_t)2 * Nx * Ny);
    const int8_t *mask = m->ocean[from];
    for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
            size_t b = 2 * IX(m, i, j);
            it->iind[b] = ip[i]; it->iind[b + 1] = im[i];
            it->jind[b] = jp[j]; it->jind[b + 1] = jm[j];
            if (exclude_land) {
                int m1 = mask[IX(m, it->iind[b], it->jind[b])];
                int m2 = mask[IX(m, it->iind[b + 1], it->jind[b + 1])];
                double s = (double)(float)(m1 + m2);
                double weight = 1.0 / (s > 1.0 ? s : 1.0);
                it->w[b] = (double)(float)m1 * weight;
                it->w[b + 1] = (double)(float)m2 * weight;
            } else {
                it->w[b] = 0.5; it->w[b + 1] = 0.5;
            }
        }
}

/* set4PointInterpolater, calc_lib.f90:282-352 */
static void set4(orc_model *m, interp_t *it, int from, int exclude_land) {
    const int *ip, *im, *jp, *jm;
    const int Nx = m->Nx, Ny = m->Ny;
    int inter = get_out_grid(m, from, 'x', &ip, &im);
    get_out_grid(m, inter, 'y', &jp, &jm);
    it->npt = 4;
    it->iind = (int *)malloc(sizeof(int) * 4 * Nx * Ny);
    it->jind = (int *)malloc(sizeof(int) * 4 * Nx * Ny);
    it->w = zalloc((size_t)4 * Nx * Ny);
    const int8_t *mask = m->ocean[from];
    for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
            size_t b = 4 * IX(m, i, j);
            it->iind[b] = ip[i]; it->iind[b + 1] = im[i]; it->iind[b + 2] = im[i]; it->iind[b + 3] = ip[i];
            it->jind[b] = jp[j]; it->jind[b + 1] = jm[j]; it->jind[b + 2] = jp[j]; it->jind[b + 3] = jm[j];
            if (exclude_land) {
                int mk[4], s = 0;
                for (int k = 0; k < 4; ++k) { mk[k] = mask[IX(m, it->iind[b + k], it->jind[b + k])]; s += mk[k]; }
                double fs = (double)(float)s;
                double weight = 1.0 / (fs > 1.0 ? fs : 1.0);
                for (int k = 0; k < 4; ++k) it->w[b + k] = (double)(float)mk[k] * weight;
            } else {
                for (int k = 0; k < 4; ++k) it->w[b + k] = 0.25;
            }
        }
}

static void free_interp(interp_t *it) { free(it->iind); free(it->jind); free(it->w); }

/* interp2 / interp4 macros, include/interp.inc:2-8 */
static inline double interp2(const orc_model *m, const double *v, const interp_t *it, int i, int j) {
    size_t b = 2 * IX(m, i, j);
    return it->w[b] * v[IX(m, it->iind[b], it->jind[b])]
         + it->w[b + 1] * v[IX(m, it->iind[b + 1], it->jind[b + 1])];
}
static inline double interp4(const orc_model *m, const double *v, const interp_t *it, int i, int j) {
    size_t b = 4 * IX(m, i, j);
    return it->w[b] * v[IX(m, it->iind[b], it->jind[b])]
         + it->w[b + 1] * v[IX(m, it->iind[b + 1], it->jind[b + 1])]
         + it->w[b + 2] * v[IX(m, it->iind[b + 2], it->jind[b + 2])]
         + it->w[b + 3] * v[IX(m, it->iind[b + 3], it->jind[b + 3])];
}

/* pder_zonal3D, calc_lib.f90:892-926 */
static void pder_zonal(const orc_model *m, const double *var, int grid, double *out) {
    const int *ind0, *indm1;
    int go = get_out_grid(m, grid, 'x', &ind0, &indm1);
    const int8_t *oc = m->ocean[grid];
    #pragma omp parallel for schedule(static, 5)
    for (int j = 0; j < m->Ny; ++j)
        for (int i = 0; i < m->Nx; ++i) {
            if (oc[IX(m, indm1[i], j)] + oc[IX(m, ind0[i], j)] == 0)
                out[IX(m, i, j)] = 0.0;
            else
                out[IX(m, i, j)] = m->bc[go][IX(m, i, j)] / (m->d.A * m->cosl[go][j] * m->d.dLambda)
                                   * (var[IX(m, ind0[i], j)] - var[IX(m, indm1[i], j)]);
        }
}

/* pder_meridional3D, calc_lib.f90:965-998 */
static void pder_meridional(const orc_model *m, const double *var, int grid, double *out) {
    const int *ind0, *indm1;
    int go = get_out_grid(m, grid, 'y', &ind0, &indm1);
    const int8_t *oc = m->ocean[grid];
    #pragma omp parallel for schedule(static, 5)
    for (int j = 0; j < m->Ny; ++j)
        for (int i = 0; i < m->Nx; ++i) {
            if (oc[IX(m, i, ind0[j])] + oc[IX(m, i, indm1[j])] == 0)
                out[IX(m, i, j)] = 0.0;
            else
                out[IX(m, i, j)] = m->bc[go][IX(m, i, j)] * (var[IX(m, i, ind0[j])] - var[IX(m, i, indm1[j])])
                                   / (m->d.A * m->d.dTheta);
        }
}

/* vorticityFromVelocities2D, calc_lib.f90:663-678 (spread temporaries as in the reference) */
static void vorticity(const orc_model *m, const double *u, const double *v, double *vort) {
    size_t n = (size_t)m->Nx * m->Ny;
    double *a = zalloc(n), *cu = zalloc(n), *b = zalloc(n);
    pder_zonal(m, v, ORC_GRID_V, a);
    #pragma omp parallel for schedule(static, 5)
    for (int j = 0; j < m->Ny; ++j)
        for (int i = 0; i < m->Nx; ++i) cu[IX(m, i, j)] = m->cosl[ORC_GRID_U][j] * u[IX(m, i, j)];
    pder_meridional(m, cu, ORC_GRID_U, b);
    #pragma omp parallel for schedule(static, 5)
    for (int j = 0; j < m->Ny; ++j)
        for (int i = 0; i < m->Nx; ++i)
            vort[IX(m, i, j)] = a[IX(m, i, j)] - b[IX(m, i, j)] / m->cosl[ORC_GRID_H][j];
    free(a); free(cu); free(b);
}

/* computeDivergence, calc_lib.f90:607-622 */
void orc_divergence(const orc_model *m, const double *u, const double *v, double *div) {
    size_t n = (size_t)m->Nx * m->Ny;
    double *a = zalloc(n), *cv = zalloc(n), *b = zalloc(n);
    pder_zonal(m, u, ORC_GRID_U, a);
    for (int j = 0; j < m->Ny; ++j)
        for (int i = 0; i < m->Nx; ++i) cv[IX(m, i, j)] = m->cosl[ORC_GRID_V][j] * v[IX(m, i, j)];
    pder_meridional(m, cv, ORC_GRID_V, b);
    for (int j = 0; j < m->Ny; ++j)
        for (int i = 0; i < m->Nx; ++i)
            div[IX(m, i, j)] = a[IX(m, i, j)] + b[IX(m, i, j)] / m->cosl[ORC_GRID_ETA][j];
    free(a); free(cv); free(b);
}

/* computeNonDivergentFlowField with the elliptic solver (calc_lib.f90:548-590):
 * u_nd = u - u_corr, v_nd = v - v_corr with (u_corr, v_corr) =
 * computeGradient(chi, eta_grid) (calc_lib.f90:638-648), i.e. pder_zonal and
 * pder_meridional of chi on the eta grid, given the velocity potential chi. */
void orc_nondivergent(const orc_model *m, const double *u, const double *v, const double *chi, double *u_nd,
                      double *v_nd) {
    size_t n = (size_t)m->Nx * m->Ny;
    double *uc = zalloc(n), *vc = zalloc(n);
    pder_zonal(m, chi, ORC_GRID_ETA, uc);
    pder_meridional(m, chi, ORC_GRID_ETA, vc);
    for (size_t k = 0; k < n; ++k) {
        u_nd[k] = u[k] - uc[k];
        v_nd[k] = v[k] - vc[k];
    }
    free(uc); free(vc);
}

/* ---------------- SWM_timestep_module diagnostics ---------------- */

/* computeD, swm_timestep_module.f90:182-218 */
static void computeD(orc_model *m) {
    const double *eta = m->SWM_eta[0];
    const int8_t *oe = m->ocean[ORC_GRID_ETA];
    if (m->d.variant == ORC_FULLY_NONLINEAR) {
        #pragma omp parallel for schedule(static, 5)
        for (int j = 0; j < m->Ny; ++j)
            for (int i = 0; i < m->Nx; ++i) {
                if (oe[IX(m, i, j)] != 1) continue;
                m->D[IX(m, i, j)] = eta[IX(m, i, j)] + m->H[ORC_GRID_ETA][IX(m, i, j)];
            }
    } else if (m->itt != 0) {
        return;
    }
    #pragma omp parallel for schedule(static, 5)
    for (int j = 0; j < m->Ny; ++j)
        for (int i = 0; i < m->Nx; ++i) {
            m->Dh[IX(m, i, j)] = interp4(m, m->D, &m->eta2H_noland, i, j);
            m->Du[IX(m, i, j)] = interp2(m, m->D, &m->eta2u_noland, i, j);
            m->Dv[IX(m, i, j)] = interp2(m, m->D, &m->eta2v_noland, i, j);
        }
}

/* computeRelVort, swm_timestep_module.f90:220-228 */
static void computeRelVort(orc_model *m) {
    if (m->d.variant == ORC_FULLY_NONLINEAR || m->d.variant == ORC_LINEARISED_MEAN_STATE)
        vorticity(m, m->SWM_u[0], m->SWM_v[0], m->zeta);
}

/* computePotVort, swm_timestep_module.f90:230-257 */
static void computePotVort(orc_model *m) {
    const double eps = DBL_EPSILON;
    const double *fH = m->f[ORC_GRID_H];
    #pragma omp parallel for schedule(static, 5)
    for (int j = 0; j < m->Ny; ++j)
        for (int i = 0; i < m->Nx; ++i) {
            double dh = m->Dh[IX(m, i, j)];
            double pD = dh > eps ? dh : eps;
            if (m->d.variant == ORC_FULLY_NONLINEAR)
                m->Pot[IX(m, i, j)] = (fH[j] + m->zeta[IX(m, i, j)]) / pD;
            else if (m->d.variant == ORC_LINEARISED_MEAN_STATE)
                m->Pot[IX(m, i, j)] = (fH[j] + m->d.zeta_bs[IX(m, i, j)]) / pD;
            else
                m->Pot[IX(m, i, j)] = fH[j] / pD;
        }
}

/* computeEdens, swm_timestep_module.f90:259-310 */
static void computeEdens(orc_model *m) {
    size_t n = (size_t)m->Nx * m->Ny;
    const double *u = m->SWM_u[0], *v = m->SWM_v[0], *eta = m->SWM_eta[0];
    double *u2 = NULL, *v2 = NULL;
    int var = m->d.variant;
    if (var != ORC_LINEARISED_STATE_OF_REST) {
        u2 = zalloc(n); v2 = zalloc(n);
        for (size_t k = 0; k < n; ++k) {
            if (var == ORC_FULLY_NONLINEAR) { u2[k] = u[k] * u[k]; v2[k] = v[k] * v[k]; }
            else { u2[k] = u[k] * m->d.u_bs[k]; v2[k] = v[k] * m->d.v_bs[k]; }
        }
    }
    const int8_t *oe = m->ocean[ORC_GRID_ETA];
    #pragma omp parallel for schedule(static, 5)
    for (int j = 0; j < m->Ny; ++j)
        for (int i = 0; i < m->Nx; ++i) {
            if (oe[IX(m, i, j)] == 0) continue;   /* eta_grid%land == 1 */
            size_t k = IX(m, i, j);
            if (var == ORC_FULLY_NONLINEAR)
                m->EDens[k] = (interp2(m, u2, &m->u2eta, i, j) + interp2(m, v2, &m->v2eta, i, j)) / 2.0
                              + m->d.G * eta[k];
            else if (var == ORC_LINEARISED_MEAN_STATE)
                m->EDens[k] = (interp2(m, u2, &m->u2eta, i, j) + interp2(m, v2, &m->v2eta, i, j))
                              + m->d.G * eta[k];
            else
                m->EDens[k] = m->d.G * eta[k];
        }
    free(u2); free(v2);
}

/* computeMassFluxes, swm_timestep_module.f90:312-326 */
static void computeMassFluxes(orc_model *m) {
    const int8_t *ou = m->ocean[ORC_GRID_U], *ov = m->ocean[ORC_GRID_V];
    #pragma omp parallel for schedule(static, 5)
    for (int j = 0; j < m->Ny; ++j)
        for (int i = 0; i < m->Nx; ++i) {
            size_t k = IX(m, i, j);
            if (ou[k] == 1) m->MU[k] = m->Du[k] * m->SWM_u[0][k];
            if (ov[k] == 1) m->MV[k] = m->Dv[k] * m->SWM_v[0][k];
        }
}

/* ---------------- lateral mixing ---------------- */

/* SWM_LateralMixing_init + _init_p_coefficients, swm_lateralmixing_module.f90:66-146 */
static void latmix_init(orc_model *m) {
    size_t n = (size_t)m->Nx * m->Ny;
    m->lat_mixing_u = zalloc(3 * n); m->lat_mixing_v = zalloc(3 * n);
    m->pll_coeff = zalloc(3 * n); m->plt_coeff = zalloc(3 * n);
    m->pll = zalloc(n); m->plt = zalloc(n); m->latmix_u = zalloc(n); m->latmix_v = zalloc(n);
    const double A = m->d.A, dL = m->d.dLambda, dT = m->d.dTheta, Ah = m->d.Ah;
    const int8_t *oH = m->ocean[ORC_GRID_H], *ou = m->ocean[ORC_GRID_U], *ov = m->ocean[ORC_GRID_V],
                 *oe = m->ocean[ORC_GRID_ETA];
    for (int j = 0; j < m->Ny; ++j)
        for (int i = 0; i < m->Nx; ++i) {
            size_t k = IX(m, i, j);
            int o_tmp = oH[k] + oH[IX(m, i, m->jp1[j])]; if (o_tmp < 1) o_tmp = 1;
            double bcu = m->bc[ORC_GRID_U][k], tu = m->tanl[ORC_GRID_U][j];
            m->lat_mixing_u[3 * k + 0] = -(bcu / A / m->cosl[ORC_GRID_U][j] / dL);
            m->lat_mixing_u[3 * k + 1] = -(bcu / A / dT - 2.0 * tu / A / (double)o_tmp);
            m->lat_mixing_u[3 * k + 2] = -(-(bcu / A / dT) - 2.0 * tu / A / (double)o_tmp);
            o_tmp = oe[k] + oe[IX(m, i, m->jm1[j])]; if (o_tmp < 1) o_tmp = 1;
            double bcv = m->bc[ORC_GRID_V][k], tv = m->tanl[ORC_GRID_V][j];
            m->lat_mixing_v[3 * k + 0] = -(bcv / A / m->cosl[ORC_GRID_V][j] / dL);
            m->lat_mixing_v[3 * k + 1] = -(-(bcv / A / dT) + 2.0 * tv / A / (double)o_tmp);
            m->lat_mixing_v[3 * k + 2] = -(bcv / A / dT + 2.0 * tv / A / (double)o_tmp);
        }
    for (int j = 0; j < m->Ny; ++j)
        for (int i = 0; i < m->Nx; ++i) {
            size_t k = IX(m, i, j);
            int o_tmp = ov[k] + ov[IX(m, i, m->jp1[j])]; if (o_tmp < 1) o_tmp = 1;
            double bce = m->bc[ORC_GRID_ETA][k], te = m->tanl[ORC_GRID_ETA][j];
            m->pll_coeff[3 * k + 0] = -(Ah * bce / dL / A / m->cosl[ORC_GRID_ETA][j]);
            m->pll_coeff[3 * k + 1] = Ah / A * (bce / dT + (double)ov[IX(m, i, m->jp1[j])] * te / (double)o_tmp);
            m->pll_coeff[3 * k + 2] = Ah / A * (-(bce / dT) + (double)ov[k] * te / (double)o_tmp);
            o_tmp = ou[k] + ou[IX(m, i, m->jm1[j])]; if (o_tmp < 1) o_tmp = 1;
            double bch = m->bc[ORC_GRID_H][k], th = m->tanl[ORC_GRID_H][j];
            m->plt_coeff[3 * k + 0] = -(Ah * bch / dL / A / m->cosl[ORC_GRID_H][j]);
            m->plt_coeff[3 * k + 1] = -(Ah / A * (bch / dT + (double)ou[k] * th / (double)o_tmp));
            m->plt_coeff[3 * k + 2] = -(Ah / A * (-(bch / dT) + (double)ou[IX(m, i, m->jm1[j])] * th / (double)o_tmp));
        }
}

/* SWM_LateralMixing_compute_p, swm_lateralmixing_module.f90:215-240 */
static void latmix_compute_p(orc_model *m) {
    const double *u = m->SWM_u[0], *v = m->SWM_v[0];
    const int8_t *oe = m->ocean[ORC_GRID_ETA], *oH = m->ocean[ORC_GRID_H];
    #pragma omp parallel for schedule(static, 5)
    for (int j = 0; j < m->Ny; ++j)
        for (int i = 0; i < m->Nx; ++i) {
            size_t k = IX(m, i, j);
            if (oe[k] == 1)
                m->pll[k] = m->D[k] * (m->pll_coeff[3 * k] * (u[IX(m, m->ip1[i], j)] - u[k])
                                       + m->pll_coeff[3 * k + 1] * v[IX(m, i, m->jp1[j])]
                                       + m->pll_coeff[3 * k + 2] * v[k]);
            if (oH[k] == 1)
                m->plt[k] = m->Dh[k] * (m->plt_coeff[3 * k] * (v[k] - v[IX(m, m->im1[i], j)])
                                        + m->plt_coeff[3 * k + 1] * u[k]
                                        + m->plt_coeff[3 * k + 2] * u[IX(m, i, m->jm1[j])]);
        }
}

/* SWM_LateralMixing_compute_uv, swm_lateralmixing_module.f90:162-186 */
static void latmix_compute_uv(orc_model *m) {
    const int8_t *ou = m->ocean[ORC_GRID_U], *ov = m->ocean[ORC_GRID_V];
    #pragma omp parallel for schedule(static, 5)
    for (int j = 0; j < m->Ny; ++j)
        for (int i = 0; i < m->Nx; ++i) {
            size_t k = IX(m, i, j);
            if (ou[k] == 1)
                m->latmix_u[k] = (m->lat_mixing_u[3 * k] * (m->pll[k] - m->pll[IX(m, m->im1[i], j)])
                                  + m->lat_mixing_u[3 * k + 1] * m->plt[IX(m, i, m->jp1[j])]
                                  + m->lat_mixing_u[3 * k + 2] * m->plt[k]) / m->Du[k];
            if (ov[k] == 1)
                m->latmix_v[k] = (m->lat_mixing_v[3 * k] * (m->plt[IX(m, m->ip1[i], j)] - m->plt[k])
                                  + m->lat_mixing_v[3 * k + 1] * m->pll[k]
                                  + m->lat_mixing_v[3 * k + 2] * m->pll[IX(m, i, m->jm1[j])]) / m->Dv[k];
        }
}

/* ---------------- time integration ---------------- */

/* time_integration_init, time_integration_module.f90:35-52 */
static void ab_init(orc_model *m) {
    memset(m->ab_coeff, 0, sizeof(m->ab_coeff));
    double chi = m->d.ab_chi;
    m->ab_coeff[0][0] = 1.0;
    m->ab_coeff[1][0] = -0.5 - chi; m->ab_coeff[1][1] = 1.5 + chi;
    m->ab_coeff[2][0] = 5.0 / 12.0; m->ab_coeff[2][1] = -16.0 / 12.0; m->ab_coeff[2][2] = 23.0 / 12.0;
    m->ab_coeff[3][0] = -3.0 / 8.0; m->ab_coeff[3][1] = 37.0 / 24.0; m->ab_coeff[3][2] = -59.0 / 24.0;
    m->ab_coeff[3][3] = 55.0 / 24.0;
    m->ab_coeff[4][0] = 251.0 / 720.0; m->ab_coeff[4][1] = -637.0 / 360.0; m->ab_coeff[4][2] = 109.0 / 30.0;
    m->ab_coeff[4][3] = -1387.0 / 360.0; m->ab_coeff[4][4] = 1901.0 / 720.0;
}

/* integrate_AB_vec, time_integration_module.f90:74-112 */
static void integrate_AB(orc_model *m, const double *A_N, double *const *G, const double *impl, int order,
                         double *A_NP1) {
    size_t n = (size_t)m->Nx * m->Ny;
    int64_t ts = m->itt < order ? m->itt : order;
    int tstep = (int)ts;
    for (size_t k = 0; k < n; ++k) A_NP1[k] = A_N[k];
    for (int ti = 1; ti <= tstep; ++ti) {
        const double c = m->ab_coeff[tstep - 1][ti - 1];
        const double *g = G[ti - 1];
        #pragma omp parallel for schedule(static, 5)
        for (int j = 0; j < m->Ny; ++j)
            for (int i = 0; i < m->Nx; ++i) {
                size_t k = IX(m, i, j);
                A_NP1[k] = A_NP1[k] + m->d.dt * c * g[k];
            }
    }
    #pragma omp parallel for schedule(static, 5)
    for (int j = 0; j < m->Ny; ++j)
        for (int i = 0; i < m->Nx; ++i) {
            size_t k = IX(m, i, j);
            A_NP1[k] = A_NP1[k] / impl[k];
        }
}

/* ---------------- SWM tendencies ---------------- */

/* SWM_timestep_nonlinear, swm_timestep_module.f90:328-449 */
static void timestep_nonlinear(orc_model *m) {
    const int Nx = m->Nx, Ny = m->Ny;
    const int var = m->d.variant;
    const double A = m->d.A, dL = m->d.dLambda, dT = m->d.dTheta;
    const int8_t *oe = m->ocean[ORC_GRID_ETA], *ou = m->ocean[ORC_GRID_U], *ov = m->ocean[ORC_GRID_V];
    const double *cv = m->cosl[ORC_GRID_V], *ce = m->cosl[ORC_GRID_ETA], *cu = m->cosl[ORC_GRID_U];
    const double *eta_now = m->SWM_eta[0], *u_now = m->SWM_u[0], *v_now = m->SWM_v[0];
    double *Ge = m->G_eta[NG - 1], *Gu = m->G_u[NG - 1], *Gv = m->G_v[NG - 1];
    const double *MU = m->MU, *MV = m->MV;
    #pragma omp parallel for schedule(static, 5)
    for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
            size_t k = IX(m, i, j);
            if (oe[k] != 1) continue;
            double t = -(MU[IX(m, m->ip1[i], j)] - MU[k]) / dL
                       - (MV[IX(m, i, m->jp1[j])] * cv[m->jp1[j]] - MV[k] * cv[j]) / dT;
            if (var == ORC_LINEARISED_MEAN_STATE) {
                /* the reference indexes the first interpolation with ip1(j) as its i index */
                t = t - (interp2(m, eta_now, &m->eta2u_noland, m->ip1[j], j) * m->d.u_bs[IX(m, m->ip1[i], j)]
                         - interp2(m, eta_now, &m->eta2u_noland, i, j) * m->d.u_bs[k]) / dL;
                t = t - (cv[m->jp1[j]] * interp2(m, eta_now, &m->eta2v_noland, i, m->jp1[j]) * m->d.v_bs[IX(m, i, m->jp1[j])]
                         - cv[j] * interp2(m, eta_now, &m->eta2v_noland, i, j) * m->d.v_bs[k]) / dT;
            }
            Ge[k] = (t / (A * ce[j]) + m->F_eta[k]);
        }
    #pragma omp parallel for schedule(static, 5)
    for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
            size_t k = IX(m, i, j);
            if (ou[k] != 1) continue;
            double g = interp2(m, m->Pot, &m->H2u, i, j) * interp4(m, MV, &m->v2u, i, j);
            if (var == ORC_LINEARISED_MEAN_STATE)
                g = g + interp2(m, m->zeta, &m->H2u, i, j) * interp4(m, m->d.v_bs, &m->v2u, i, j);
            g = g - ((m->EDens[k] - m->EDens[IX(m, m->im1[i], j)]) / (A * cu[j] * dL));
            if (m->d.quadratic_friction) {
                double vi = interp4(m, v_now, &m->v2u, i, j);
                double q = m->d.gamma_sq_u[k] * sqrt(u_now[k] * u_now[k] + (vi * vi)) * u_now[k];
                if (m->d.barotropic) q = q / m->H[ORC_GRID_U][k];
                g = g - q;
            }
            if (m->d.lateral_mixing) g = g + m->latmix_u[k];
            Gu[k] = (g + m->F_x[k]);
        }
    #pragma omp parallel for schedule(static, 5)
    for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
            size_t k = IX(m, i, j);
            if (ov[k] != 1) continue;
            double g = -(interp2(m, m->Pot, &m->H2v, i, j) * interp4(m, MU, &m->u2v, i, j));
            if (var == ORC_LINEARISED_MEAN_STATE)
                g = g - interp2(m, m->zeta, &m->H2v, i, j) * interp4(m, m->d.u_bs, &m->u2v, i, j);
            g = g - ((m->EDens[k] - m->EDens[IX(m, i, m->jm1[j])]) / (A * dT));
            if (m->d.quadratic_friction) {
                double ui = interp4(m, u_now, &m->u2v, i, j);
                double q = m->d.gamma_sq_v[k] * sqrt(v_now[k] * v_now[k] + (ui * ui)) * v_now[k];
                if (m->d.barotropic) q = q / m->H[ORC_GRID_V][k];
                g = g - q;
            }
            if (m->d.lateral_mixing) g = g + m->latmix_v[k];
            Gv[k] = (g + m->F_y[k]);
        }
    integrate_AB(m, m->SWM_eta[0], m->G_eta, m->impl_eta, NG, m->SWM_eta[1]);
    integrate_AB(m, m->SWM_u[0], m->G_u, m->impl_u, NG, m->SWM_u[1]);
    integrate_AB(m, m->SWM_v[0], m->G_v, m->impl_v, NG, m->SWM_v[1]);
}

/* SWM_timestep_step, swm_timestep_module.f90:164-180 */
static void timestep_step(orc_model *m) {
    if (m->d.lateral_mixing) { latmix_compute_p(m); latmix_compute_uv(m); }
    timestep_nonlinear(m);
    if (m->d.variant == ORC_FULLY_NONLINEAR) {
        double *eta = m->SWM_eta[1];
        const double *He = m->H[ORC_GRID_ETA];
        size_t n = (size_t)m->Nx * m->Ny;
        for (size_t k = 0; k < n; ++k)
            if (eta[k] < (m->d.minD - He[k])) eta[k] = m->d.minD - He[k];
    }
}

/* SWM_timestep_advance, swm_timestep_module.f90:129-155 */
static void timestep_advance(orc_model *m) {
    size_t n = (size_t)m->Nx * m->Ny;
    for (int ti = 0; ti < NG - 1; ++ti) {
        memcpy(m->G_u[ti], m->G_u[ti + 1], n * sizeof(double));
        memcpy(m->G_v[ti], m->G_v[ti + 1], n * sizeof(double));
        memcpy(m->G_eta[ti], m->G_eta[ti + 1], n * sizeof(double));
    }
    computeD(m);
    computeRelVort(m);
    computePotVort(m);
    computeEdens(m);
    computeMassFluxes(m);
}

/* getVarChunkFromDisk, memchunk_module.f90:181-207: len records from record
 * *nstart (1-based) into var, wrapping to the start of the dataset. */
static void mc_get_var_chunk(memchunk *mc, double *var, int64_t *nstart, int64_t len) {
    if (*nstart > mc->nrec) *nstart = *nstart - mc->nFpC * mc->nrec;
    if (*nstart == 0) *nstart = mc->nrec;
    int64_t nend = *nstart + len - 1 < mc->nrec ? *nstart + len - 1 : mc->nrec;
    int64_t len2 = (nend - *nstart > 0 ? nend - *nstart : 0) + 1;
    memcpy(var, mc->file + (size_t)(*nstart - 1) * mc->n, (size_t)len2 * mc->n * sizeof(double));
    if (len2 < len) {
        int64_t nstart2 = 1;
        mc_get_var_chunk(mc, var + (size_t)len2 * mc->n, &nstart2, len - len2);
    }
}

/* getChunkFromDisk, memchunk_module.f90:149-170 */
static void mc_get_chunk(memchunk *mc) {
    if (mc->isInitialised) {
        mc->fileRec = mc->fileRec + mc->chunkSize - 1;
        mc->tOffset = mc->tOffset - mc->dt;
    }
    if (!mc->isPersistent || !mc->isInitialised) mc_get_var_chunk(mc, mc->var, &mc->fileRec, mc->chunkSize);
    if (!mc->isConstant) {
        for (int64_t i = 0; i < mc->chunkSize; ++i) {
            mc->time[i] = mc->tOffset;
            mc->tOffset = mc->tOffset + mc->dt;
        }
    }
    mc->chunkCounter = 1;
}

/* initMemChunk, memchunk_module.f90:77-115 (the dataset's getDt is passed in) */
static void mc_init(memchunk *mc, const double *file, size_t n, int64_t nrec, int64_t chunkSize, double dt) {
    memset(mc, 0, sizeof(*mc));
    mc->file = file;
    mc->n = n;
    mc->nrec = nrec;
    mc->fileRec = 1;
    mc->chunkCounter = 1;
    mc->chunkSize = chunkSize;
    if (nrec <= 1) {
        mc->isPersistent = 1;
        mc->isConstant = 1;
        mc->chunkSize = 1;
    } else if (chunkSize >= nrec + 1) {
        mc->isPersistent = 1;
        mc->chunkSize = nrec + 1;
    }
    mc->dt = mc->isConstant ? 0.0 : dt;
    mc->tOffset = 0.0;
    mc->var = zalloc((size_t)mc->chunkSize * n);
    mc->time = zalloc((size_t)mc->chunkSize);
    mc_get_chunk(mc);
    int64_t q = (int64_t)floorf((float)mc->chunkSize / (float)nrec);   /* floor(REAL(chunkSize)/nrec) */
    mc->nFpC = q > 1 ? q : 1;
    mc->isInitialised = 1;
}

/* getChunkData, memchunk_module.f90:214-243, with interp1D_array2D
 * (calc_lib.f90:386-397) between the bracketing records. */
static void mc_get(memchunk *mc, double t, double *out) {
    if (mc->isConstant) { memcpy(out, mc->var, mc->n * sizeof(double)); return; }
    if (mc->chunkCounter >= mc->chunkSize) mc_get_chunk(mc);
    while (mc->time[mc->chunkCounter] < t) {          /* time(chunkCounter+1), 1-based */
        mc->chunkCounter = mc->chunkCounter + 1;
        if (mc->chunkCounter >= mc->chunkSize) mc_get_chunk(mc);
    }
    const double *y0 = mc->var + (size_t)(mc->chunkCounter - 1) * mc->n;
    const double x0 = mc->time[mc->chunkCounter - 1];
    if (t == x0) { memcpy(out, y0, mc->n * sizeof(double)); return; }
    const double *y1 = y0 + mc->n;
    const double x1 = mc->time[mc->chunkCounter];
    for (size_t k = 0; k < mc->n; ++k) out[k] = y0[k] + (t - x0) * (y1[k] - y0[k]) / (x1 - x0);
}

/* SWM_forcing_update, swm_forcing_module.f90:135-170 (+ processWindstress 274-324,
 * processOscillation 326-374) */
static void forcing_update(orc_model *m) {
    size_t n = (size_t)m->Nx * m->Ny;
    memcpy(m->F_eta, m->F_eta_const, n * sizeof(double));
    memcpy(m->F_x, m->F_x_const, n * sizeof(double));
    memcpy(m->F_y, m->F_y_const, n * sizeof(double));
    for (int s = 0; s < m->d.n_wind; ++s) {
        int comp = m->d.wind_comp[s];
        double *F = comp == ORC_COMP_X ? m->F_x : m->F_y;
        const int8_t *mask = comp == ORC_COMP_X ? m->ocean[ORC_GRID_U] : m->ocean[ORC_GRID_V];
        const double *H = comp == ORC_COMP_X ? m->Du : m->Dv;
        const double *tau = m->d.wind_tau[s];
        int scaleH = m->d.barotropic || m->d.variant == ORC_FULLY_NONLINEAR;
        for (size_t k = 0; k < n; ++k) {
            if (mask[k] != 1) continue;
            double num = m->d.has_tau_scale ? m->d.tau_scale * tau[k] : tau[k];
            F[k] = F[k] + num / (scaleH ? m->d.RHO0 * H[k] : m->d.RHO0);
        }
    }
    for (int s = 0; s < m->d.n_osc; ++s) {
        int comp = m->d.osc_comp[s];
        double *F = comp == ORC_COMP_X ? m->F_x : comp == ORC_COMP_Y ? m->F_y : m->F_eta;
        const int8_t *mask = comp == ORC_COMP_X ? m->ocean[ORC_GRID_U]
                           : comp == ORC_COMP_Y ? m->ocean[ORC_GRID_V] : m->ocean[ORC_GRID_ETA];
        double arg = m->d.osc_omega[s] * (double)m->itt * m->d.dt;
        double r_iot = cos(arg), i_iot = sin(arg);
        const double *R = m->d.osc_re[s], *I = m->d.osc_im[s];
        for (size_t k = 0; k < n; ++k)
            F[k] = F[k] + (2.0 * (R[k] * r_iot - I[k] * i_iot)) * (double)mask[k];
    }
    /* processWindstress 274-324, processCustomForcing 395-434, processHeating 535-552
     * on getChunkData(memChunk, itt*dt) */
    for (int s = 0; s < m->d.n_tdf; ++s) {
        int comp = m->d.tdf_comp[s];
        double *F = comp == ORC_COMP_X ? m->F_x : comp == ORC_COMP_Y ? m->F_y : m->F_eta;
        const int8_t *mask = comp == ORC_COMP_X ? m->ocean[ORC_GRID_U]
                           : comp == ORC_COMP_Y ? m->ocean[ORC_GRID_V] : m->ocean[ORC_GRID_ETA];
        double *data = m->tdf_buf;
        mc_get(&m->mc[s], (double)m->itt * m->d.dt, data);
        if (m->d.tdf_kind[s] == ORC_TDF_WIND) {
            const double *H = comp == ORC_COMP_X ? m->Du : m->Dv;
            int scaleH = m->d.barotropic || m->d.variant == ORC_FULLY_NONLINEAR;
            for (size_t k = 0; k < n; ++k) {
                if (mask[k] != 1) continue;
                double num = m->d.has_tau_scale ? m->d.tau_scale * data[k] : data[k];
                F[k] = F[k] + num / (scaleH ? m->d.RHO0 * H[k] : m->d.RHO0);
            }
        } else {
            for (size_t k = 0; k < n; ++k)
                if (mask[k] == 1) F[k] = F[k] + data[k];
        }
    }
}

/* ---------------- tracer ---------------- */

/* TRC_initCoeffs, tracer_module.f90:291-342 */
static void trc_init_coeffs(orc_model *m) {
    size_t n = (size_t)m->Nx * m->Ny;
    m->TRC_coeff = zalloc(12 * n);
    const double A = m->d.A, dL = m->d.dLambda, dT = m->d.dTheta;
    const int8_t *oe = m->ocean[ORC_GRID_ETA], *ou = m->ocean[ORC_GRID_U], *ov = m->ocean[ORC_GRID_V];
    const double *cu = m->cosl[ORC_GRID_U], *cv = m->cosl[ORC_GRID_V], *ce = m->cosl[ORC_GRID_ETA];
    for (int j = 0; j < m->Ny; ++j)
        for (int i = 0; i < m->Nx; ++i) {
            size_t k = IX(m, i, j);
            double *c = m->TRC_coeff + 12 * k;
            int ip = m->ip1[i], im = m->im1[i], jp = m->jp1[j], jm = m->jm1[j];
            double oe0 = oe[k];
            double oe_ip = oe[IX(m, ip, j)], oe_im = oe[IX(m, im, j)], oe_jp = oe[IX(m, i, jp)], oe_jm = oe[IX(m, i, jm)];
            double ou_ip = ou[IX(m, ip, j)], ou0 = ou[k], ov_jp = ov[IX(m, i, jp)], ov0 = ov[k];
            c[0] = -1.0 * ou_ip * oe_ip / (double)(oe[IX(m, ip, j)] + oe[k]) / A / dL / ce[j];
            c[1] = 1.0 * ou0 * oe_im / (double)(oe[IX(m, im, j)] + oe[k]) / A / dL / ce[j];
            c[2] = -1.0 * ov_jp * cv[jp] * oe_jp / (double)(oe[IX(m, i, jp)] + oe[k]) / A / dT / ce[j];
            c[3] = 1.0 * ov0 * cv[j] * oe_jm / (double)(oe[IX(m, i, jm)] + oe[k]) / A / dT / ce[j];
            c[4] = -1.0 * ou_ip * oe0 / (double)(oe[IX(m, ip, j)] + oe[k]) / A / dL / ce[j];
            c[5] = 1.0 * ou0 * oe0 / (double)(oe[IX(m, im, j)] + oe[k]) / A / dL / ce[j];
            c[6] = -1.0 * ov_jp * cv[jp] * oe0 / (double)(oe[IX(m, i, jp)] + oe[k]) / A / dT / ce[j];
            c[7] = 1.0 * ov0 * cv[j] * oe0 / (double)(oe[IX(m, i, jm)] + oe[k]) / A / dT / ce[j];
            c[8] = 1.0 * ou_ip / (dL * dL) / (A * A) / cu[j] / ce[j];
            c[9] = -1.0 * ou0 / (dL * dL) / (A * A) / cu[j] / ce[j];
            c[10] = 1.0 * ov_jp * cv[jp] / (dT * dT) / (A * A) / ce[j];
            c[11] = -1.0 * ov0 * cv[j] / (dT * dT) / (A * A) / ce[j];
        }
}

/* TRC_tracer_incr, tracer_module.f90:154-209 */
static void trc_incr(orc_model *m, tracer_t *t) {
    const double *CH = t->CH[0], *C = t->C;
    double *GCH = t->G_CH[TRC_NG - 1];
    const double kappa_h = t->kappa_h;
    const double *h = m->D, *h_u = m->Du, *h_v = m->Dv, *u = m->u, *v = m->v;
    const int8_t *oe = m->ocean[ORC_GRID_ETA];
    #pragma omp parallel for schedule(static, 5)
    for (int j = 0; j < m->Ny; ++j)
        for (int i = 0; i < m->Nx; ++i) {
            size_t k = IX(m, i, j);
            if (oe[k] == 0) continue;   /* eta_grid%land == 1 */
            const double *c = m->TRC_coeff + 12 * k;
            int ip = m->ip1[i], im = m->im1[i], jp = m->jp1[j], jm = m->jm1[j];
            t->diff[k] = kappa_h * h_u[IX(m, ip, j)] * (C[IX(m, ip, j)] - C[k]) * c[8]
                       + kappa_h * h_u[k] * (C[k] - C[IX(m, im, j)]) * c[9]
                       + kappa_h * h_v[IX(m, i, jp)] * (C[IX(m, i, jp)] - C[k]) * c[10]
                       + kappa_h * h_v[k] * (C[k] - C[IX(m, i, jm)]) * c[11];
        }
    for (int j = 0; j < m->Ny; ++j)
        for (int i = 0; i < m->Nx; ++i) { size_t k = IX(m, i, j); t->forcing[k] = C[k] * m->F_eta[k]; }
    #pragma omp parallel for schedule(static, 5)
    for (int j = 0; j < m->Ny; ++j)
        for (int i = 0; i < m->Nx; ++i) {
            size_t k = IX(m, i, j);
            const double *c = m->TRC_coeff + 12 * k;
            int ip = m->ip1[i], im = m->im1[i], jp = m->jp1[j], jm = m->jm1[j];
            GCH[k] = (CH[IX(m, ip, j)] * u[IX(m, ip, j)] * c[0]
                      + CH[IX(m, im, j)] * u[k] * c[1]
                      + CH[IX(m, i, jp)] * v[IX(m, i, jp)] * c[2]
                      + CH[IX(m, i, jm)] * v[k] * c[3]
                      + CH[k] * u[IX(m, ip, j)] * c[4]
                      + CH[k] * u[k] * c[5]
                      + CH[k] * v[IX(m, i, jp)] * c[6]
                      + CH[k] * v[k] * c[7]
                      + t->diff[k]
                      - t->gamma_C[k] * h[k] * (C[k] - t->C0[k])
                      + t->forcing[k]);
        }
}

/* TRC_tracer_advance, tracer_module.f90:237-280 */
static void trc_advance(orc_model *m, tracer_t *t) {
    size_t n = (size_t)m->Nx * m->Ny;
    for (int ti = 0; ti < TRC_NG - 1; ++ti) memcpy(t->G_CH[ti], t->G_CH[ti + 1], n * sizeof(double));
    memcpy(t->CH[0], t->CH[1], n * sizeof(double));
    for (size_t k = 0; k < n; ++k) t->C[k] = t->CH[0][k] / m->D[k];
    for (int j = 0; j < m->Ny; ++j)
        for (int i = 0; i < m->Nx; ++i) {
            size_t k = IX(m, i, j);
            t->uhc[k] = m->MU[k] * .5 * (t->C[k] + t->C[IX(m, m->im1[i], j)]);
            t->vhc[k] = m->MV[k] * .5 * (t->C[k] + t->C[IX(m, i, m->jm1[j])]);
        }
}

/* ---------------- driver ---------------- */

/* SWM_advance, swm_module.f90:86-139 (master fields reset by model_advance, model.f90:238-259) */
static void swm_advance(orc_model *m) {
    size_t n = (size_t)m->Nx * m->Ny;
    memcpy(m->SWM_eta[0], m->SWM_eta[1], n * sizeof(double));
    memcpy(m->SWM_u[0], m->SWM_u[1], n * sizeof(double));
    memcpy(m->SWM_v[0], m->SWM_v[1], n * sizeof(double));
    for (size_t k = 0; k < n; ++k) m->u[k] = m->u[k] + m->SWM_u[0][k];
    for (size_t k = 0; k < n; ++k) m->v[k] = m->v[k] + m->SWM_v[0][k];
    for (size_t k = 0; k < n; ++k) m->eta[k] = m->eta[k] + m->SWM_eta[0][k];
    timestep_advance(m);
    forcing_update(m);
}

/* DFF_advance, dynFromFile_module.f90:137-160: the master field += ocean *
 * getChunkData(chunk, itt*dt) for each dataset given (u, v, eta). */
static void dff_advance(orc_model *m) {
    size_t n = (size_t)m->Nx * m->Ny;
    const double model_time = (double)m->itt * m->d.dt;
    for (int s = 0; s < m->d.n_dff; ++s) {
        const int f = m->d.dff_field[s];
        double *dst = f == ORC_COMP_X ? m->u : (f == ORC_COMP_Y ? m->v : m->eta);
        const int8_t *oc = m->ocean[f == ORC_COMP_X ? ORC_GRID_U : (f == ORC_COMP_Y ? ORC_GRID_V : ORC_GRID_ETA)];
        mc_get(&m->dff_mc[s], model_time, m->tdf_buf);
        for (size_t k = 0; k < n; ++k) dst[k] = dst[k] + (double)oc[k] * m->tdf_buf[k];
    }
}

static void model_advance(orc_model *m) {
    size_t n = (size_t)m->Nx * m->Ny;
    for (size_t k = 0; k < n; ++k) { m->u[k] = 0.0; m->v[k] = 0.0; m->eta[k] = 0.0; }
    dff_advance(m);
    swm_advance(m);
    for (int t = 0; t < m->ntrc; ++t) trc_advance(m, &m->trc[t]);
}

orc_model *orc_create(const orc_desc *d) {
    orc_model *m = (orc_model *)calloc(1, sizeof(orc_model));
    m->d = *d;
    const int Nx = d->Nx, Ny = d->Ny;
    m->Nx = Nx; m->Ny = Ny;
    size_t n = (size_t)Nx * Ny;
    m->ip0 = malloc(sizeof(int) * Nx); m->ip1 = malloc(sizeof(int) * Nx); m->im1 = malloc(sizeof(int) * Nx);
    m->jp0 = malloc(sizeof(int) * Ny); m->jp1 = malloc(sizeof(int) * Ny); m->jm1 = malloc(sizeof(int) * Ny);
    for (int i = 0; i < Nx; ++i) { m->ip0[i] = i; m->ip1[i] = (i + 1) % Nx; m->im1[i] = (i + Nx - 1) % Nx; }
    for (int j = 0; j < Ny; ++j) { m->jp0[j] = j; m->jp1[j] = (j + 1) % Ny; m->jm1[j] = (j + Ny - 1) % Ny; }
    for (int g = 0; g < 4; ++g) {
        m->ocean[g] = d->grid[g].ocean;
        m->cosl[g] = d->grid[g].cos_lat; m->tanl[g] = d->grid[g].tan_lat; m->f[g] = d->grid[g].f;
        m->H[g] = zalloc(n);
        memcpy(m->H[g], d->grid[g].H, n * sizeof(double));
        m->bc[g] = zalloc(n);                 /* setGridEuler, grid_module.f90:167-195 */
        for (size_t k = 0; k < n; ++k) m->bc[g][k] = d->grid[g].ocean[k] == 1 ? 1.0 : d->lbc;
    }
    ab_init(m);
    /* initCalcLib, calc_lib.f90:133-176 (interpolators used on the path) */
    set2(m, &m->eta2u_noland, ORC_GRID_ETA, 'x', 1);
    set2(m, &m->eta2v_noland, ORC_GRID_ETA, 'y', 1);
    set4(m, &m->eta2H_noland, ORC_GRID_ETA, 1);
    set2(m, &m->u2eta, ORC_GRID_U, 'x', 0);
    set2(m, &m->v2eta, ORC_GRID_V, 'y', 0);
    set2(m, &m->H2u, ORC_GRID_H, 'y', 0);
    set2(m, &m->H2v, ORC_GRID_H, 'x', 0);
    set4(m, &m->v2u, ORC_GRID_V, 0);
    set4(m, &m->u2v, ORC_GRID_U, 0);
    /* swm_vars_init, swm_vars.f90:56-123 */
    for (int l = 0; l < 2; ++l) { m->SWM_u[l] = zalloc(n); m->SWM_v[l] = zalloc(n); m->SWM_eta[l] = zalloc(n); }
    for (int l = 0; l < NG; ++l) { m->G_u[l] = zalloc(n); m->G_v[l] = zalloc(n); m->G_eta[l] = zalloc(n); }
    m->EDens = zalloc(n); m->Pot = zalloc(n); m->zeta = zalloc(n); m->MU = zalloc(n); m->MV = zalloc(n);
    if (d->variant == ORC_FULLY_NONLINEAR) {
        m->D = falloc(n, 1.0); m->Dh = falloc(n, 1.0); m->Du = falloc(n, 1.0); m->Dv = falloc(n, 1.0);
    } else {
        m->D = m->H[ORC_GRID_ETA]; m->Dh = m->H[ORC_GRID_H]; m->Du = m->H[ORC_GRID_U]; m->Dv = m->H[ORC_GRID_V];
    }
    m->u = zalloc(n); m->v = zalloc(n); m->eta = zalloc(n);
    m->itt = 0;
    /* SWM_timestep_init, swm_timestep_module.f90:60-101 */
    if (d->lateral_mixing) latmix_init(m);
    computeD(m); computeRelVort(m); computePotVort(m); computeEdens(m); computeMassFluxes(m);
    /* SWM_damping_init: coefficients come from the host */
    m->impl_u = d->impl_u; m->impl_v = d->impl_v; m->impl_eta = d->impl_eta;
    m->gamma_sq_u = d->gamma_sq_u; m->gamma_sq_v = d->gamma_sq_v;
    /* SWM_forcing_init, swm_forcing_module.f90:78-127 */
    m->F_x = zalloc(n); m->F_y = zalloc(n); m->F_eta = zalloc(n);
    m->F_x_const = zalloc(n); m->F_y_const = zalloc(n); m->F_eta_const = zalloc(n);
    if (d->F_x_const) memcpy(m->F_x_const, d->F_x_const, n * sizeof(double));
    if (d->F_y_const) memcpy(m->F_y_const, d->F_y_const, n * sizeof(double));
    if (d->F_eta_const) memcpy(m->F_eta_const, d->F_eta_const, n * sizeof(double));
    for (int s = 0; s < d->n_tdf; ++s)   /* SWM_forcing_initStream -> initMemChunk, 593-607 */
        mc_init(&m->mc[s], d->tdf_rec[s], n, d->tdf_nrec[s], d->tdf_chunk[s], d->tdf_dt[s]);
    m->tdf_buf = zalloc(n);
    /* DFF_initDynFromFile (model.f90:108-113, before SWM_initSWM): initMemChunk
     * of each dataset and DFF_advance at itt = 0 */
    for (int s = 0; s < d->n_dff; ++s)
        mc_init(&m->dff_mc[s], d->dff_rec[s], n, d->dff_nrec[s], d->dff_chunk[s], d->dff_dt[s]);
    dff_advance(m);
    forcing_update(m);
    /* SWM_initialConditions, swm_module.f90:156-181 */
    for (size_t k = 0; k < n; ++k) {
        m->SWM_eta[1][k] = (double)d->grid[ORC_GRID_ETA].ocean[k] * (d->eta0 ? d->eta0[k] : 0.0);
        m->SWM_u[1][k] = (double)d->grid[ORC_GRID_U].ocean[k] * (d->u0 ? d->u0[k] : 0.0);
        m->SWM_v[1][k] = (double)d->grid[ORC_GRID_V].ocean[k] * (d->v0 ? d->v0[k] : 0.0);
    }
    swm_advance(m);
    /* TRC_initTracer, tracer_module.f90:58-80 + TRC_add_to_list, tracer_vars.f90:124-185 */
    m->ntrc = d->n_tracer;
    if (m->ntrc > 0) trc_init_coeffs(m);
    for (int t = 0; t < m->ntrc; ++t) {
        tracer_t *tr = &m->trc[t];
        tr->kappa_h = d->kappa_h[t];
        tr->C = zalloc(n); tr->C0 = zalloc(n); tr->gamma_C = zalloc(n); tr->cons = zalloc(n);
        memcpy(tr->C, d->trc_C[t], n * sizeof(double));
        memcpy(tr->C0, d->trc_C0[t], n * sizeof(double));
        memcpy(tr->gamma_C, d->trc_gamma_C[t], n * sizeof(double));
        memcpy(tr->cons, d->trc_cons[t], n * sizeof(double));
        for (int l = 0; l < 2; ++l) tr->CH[l] = zalloc(n);
        for (int l = 0; l < TRC_NG; ++l) tr->G_CH[l] = zalloc(n);
        for (size_t k = 0; k < n; ++k) tr->CH[0][k] = m->D[k] * tr->C[k];
        tr->impl = zalloc(n);
        for (size_t k = 0; k < n; ++k) tr->impl[k] = 1.0 + d->dt * tr->cons[k];
        tr->uhc = zalloc(n); tr->vhc = zalloc(n); tr->diff = zalloc(n); tr->forcing = zalloc(n);
    }
    return m;
}

void orc_step(orc_model *m, int64_t nsteps) {
    for (int64_t s = 0; s < nsteps; ++s) {
        m->itt += 1;
        timestep_step(m);
        for (int t = 0; t < m->ntrc; ++t) {
            tracer_t *tr = &m->trc[t];
            trc_incr(m, tr);
            integrate_AB(m, tr->CH[0], tr->G_CH, tr->impl, TRC_NG, tr->CH[1]);
        }
        model_advance(m);
    }
}

int64_t orc_itt(const orc_model *m) { return m->itt; }

/* computeStreamfunction, calc_lib.f90:727-780, without an elliptic solver
 * (u_nd = u, calc_lib.f90:563-564): path integrals of the (BAROTROPIC:
 * thickness-weighted) flow; every SUM runs from its first to its last element. */
void orc_psi(const orc_model *m, const double *u, const double *v, double *psi) {
    const int Nx = m->Nx, Ny = m->Ny;
    const int bt = m->d.barotropic;
    const double A = m->d.A, dL = m->d.dLambda, dT = m->d.dTheta;
    const double *Hu = m->H[ORC_GRID_U], *Hv = m->H[ORC_GRID_V], *cv = m->cosl[ORC_GRID_V];
    const int8_t *ou = m->ocean[ORC_GRID_U], *ov = m->ocean[ORC_GRID_V];
    for (size_t k = 0; k < (size_t)Nx * Ny; ++k) psi[k] = 0.0;
    for (int i = 0; i < Nx - 1; ++i) {          /* psi(i,1), i = 1..Nx-1 */
        double s = 0.0;
        for (int q = i; q < Nx - 1; ++q) {
            size_t k = IX(m, q, 0);
            s = s + (bt ? Hv[k] * (double)ov[k] * v[k] : (double)ov[k] * v[k]);
        }
        psi[IX(m, i, 0)] = -s * A * cv[0] * dL;
    }
    for (int j = 1; j < Ny; ++j) {              /* psi(Nx,j), j = 2..Ny */
        double s = 0.0;
        for (int q = 0; q < j; ++q) {
            size_t k = IX(m, Nx - 1, q);
            s = s + (bt ? Hu[k] * (double)ou[k] * u[k] : (double)ou[k] * u[k]);
        }
        psi[IX(m, Nx - 1, j)] = -s * A * dT;
    }
    for (int j = 1; j < Ny; ++j)
        for (int i = 0; i < Nx - 1; ++i) {
            double s1 = 0.0, s2 = 0.0;
            for (int q = i; q < Nx - 1; ++q) {
                size_t k = IX(m, q, j);
                s1 = s1 + (bt ? Hv[k] * v[k] : v[k]);
            }
            for (int q = 0; q < j; ++q) {
                size_t k = IX(m, i, q);
                s2 = s2 + (bt ? Hu[k] * u[k] : u[k]);
            }
            psi[IX(m, i, j)] = (-s1 * A * cv[j] * dL + psi[IX(m, Nx - 1, j)] - s2 * A * dT + psi[IX(m, i, 0)]) / 2.0;
        }
}

int orc_get(const orc_model *m, const char *name, int tracer, double *out) {
    size_t n = (size_t)m->Nx * m->Ny;
    const double *src = NULL;
    const tracer_t *t = (tracer >= 0 && tracer < m->ntrc) ? &m->trc[tracer] : NULL;
    if (!strcmp(name, "U")) src = m->u;            /* master fields (vars_module u, v, eta at N0) */
    else if (!strcmp(name, "V")) src = m->v;
    else if (!strcmp(name, "ETA")) src = m->eta;
    else if (!strcmp(name, "u")) src = m->SWM_u[0];
    else if (!strcmp(name, "v")) src = m->SWM_v[0];
    else if (!strcmp(name, "eta")) src = m->SWM_eta[0];
    else if (!strcmp(name, "G_u")) src = m->G_u[NG - 1];
    else if (!strcmp(name, "G_v")) src = m->G_v[NG - 1];
    else if (!strcmp(name, "G_eta")) src = m->G_eta[NG - 1];
    else if (!strcmp(name, "G_u_m1")) src = m->G_u[0];      /* oldest slot G(1) */
    else if (!strcmp(name, "G_v_m1")) src = m->G_v[0];
    else if (!strcmp(name, "G_eta_m1")) src = m->G_eta[0];
    else if (!strcmp(name, "D")) src = m->D;
    else if (!strcmp(name, "Dh")) src = m->Dh;
    else if (!strcmp(name, "Du")) src = m->Du;
    else if (!strcmp(name, "Dv")) src = m->Dv;
    else if (!strcmp(name, "EDens")) src = m->EDens;
    else if (!strcmp(name, "Pot")) src = m->Pot;
    else if (!strcmp(name, "zeta")) src = m->zeta;
    else if (!strcmp(name, "MU")) src = m->MU;
    else if (!strcmp(name, "MV")) src = m->MV;
    else if (!strcmp(name, "F_x")) src = m->F_x;
    else if (!strcmp(name, "F_y")) src = m->F_y;
    else if (!strcmp(name, "F_eta")) src = m->F_eta;
    else if (!strcmp(name, "latmix_u")) src = m->latmix_u;
    else if (!strcmp(name, "latmix_v")) src = m->latmix_v;
    else if (!strcmp(name, "H")) src = m->H[ORC_GRID_H];
    else if (!strcmp(name, "H_u")) src = m->H[ORC_GRID_U];
    else if (!strcmp(name, "H_v")) src = m->H[ORC_GRID_V];
    else if (!strcmp(name, "H_eta")) src = m->H[ORC_GRID_ETA];
    else if (t && !strcmp(name, "CH")) src = t->CH[0];
    else if (t && !strcmp(name, "C")) src = t->C;
    else if (t && !strcmp(name, "G_CH")) src = t->G_CH[TRC_NG - 1];
    else if (t && !strcmp(name, "G_CH_m1")) src = t->G_CH[0];
    else if (t && !strcmp(name, "uhc")) src = t->uhc;
    else if (t && !strcmp(name, "vhc")) src = t->vhc;
    if (!src) return 1;
    memcpy(out, src, n * sizeof(double));
    return 0;
}

/* init_ElSolv_SOR + init_oe_index_space, ElSolv_SOR.f90:44-102 */
static void sor_init(orc_model *m) {
    if (m->sor_c) return;
    const int Nx = m->Nx, Ny = m->Ny;
    size_t n = (size_t)Nx * Ny;
    const double A = m->d.A, dL = m->d.dLambda, dT = m->d.dTheta;
    const int8_t *ou = m->ocean[ORC_GRID_U], *ov = m->ocean[ORC_GRID_V];
    const double *ce = m->cosl[ORC_GRID_ETA], *cv = m->cosl[ORC_GRID_V];
    m->sor_c = zalloc(5 * n);
    for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
            size_t k = IX(m, i, j);
            int ip = m->ip1[i], jp = m->jp1[j];
            double x = A * ce[j] * dL; x = x * x;
            double y = A * A * ce[j] * (dT * dT);
            m->sor_c[0 * n + k] = (double)ou[IX(m, ip, j)] / x;
            m->sor_c[1 * n + k] = (double)ou[k] / x;
            m->sor_c[2 * n + k] = cv[jp] * (double)ov[IX(m, i, jp)] / y;
            m->sor_c[3 * n + k] = cv[j] * (double)ov[k] / y;
            m->sor_c[4 * n + k] = (double)(-ou[IX(m, ip, j)] - ou[k]) / x
                                  + (-(cv[jp] * (double)ov[IX(m, i, jp)]) - cv[j] * (double)ov[k]) / y;
        }
    int n_odd = (int)((n + 1) / 2);
    m->n_oe[0] = n_odd; m->n_oe[1] = (int)(n / 2);
    m->i_oe = malloc(sizeof(int) * 2 * 2 * (size_t)n_odd);
    int cnt[2] = {0, 0};
    for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
            int oe = ((i + 1) + (j + 1)) % 2 == 0 ? 0 : 1;
            int *p = m->i_oe + 2 * ((size_t)oe * n_odd + cnt[oe]);
            p[0] = i; p[1] = j; cnt[oe]++;
        }
}

static inline double sor_lhs(const orc_model *m, const double *chi, int i, int j) {
    size_t n = (size_t)m->Nx * m->Ny, k = IX(m, i, j);
    const double *c = m->sor_c;
    return c[0 * n + k] * chi[IX(m, m->ip1[i], j)] + c[1 * n + k] * chi[IX(m, m->im1[i], j)]
         + c[2 * n + k] * chi[IX(m, i, m->jp1[j])] + c[3 * n + k] * chi[IX(m, i, m->jm1[j])]
         + c[4 * n + k] * chi[k];
}

/* main_ElSolv_SOR, ElSolv_SOR.f90:150-222 (anorm summed in list order, one thread) */
int orc_sor(orc_model *m, const double *B, double *chi, double epsilon, int first_guess,
            int64_t *iters, double *anorm_out) {
    sor_init(m);
    size_t n = (size_t)m->Nx * m->Ny;
    int mx = m->Nx > m->Ny ? m->Nx : m->Ny;
    int64_t max_count = first_guess ? (int64_t)lrintf(1000.f * (float)mx) : (int64_t)lrintf(10.f * (float)mx);
    const double rJ = (double)9.9999178e-1f;
    double relax = 1.0, anorm = 0.0;
    const int8_t *oe = m->ocean[ORC_GRID_ETA];
    int64_t l;
    for (l = 1; l <= max_count; ++l) {
        anorm = 0.0;
        for (int oddeven = 0; oddeven < 2; ++oddeven) {
            const int *p = m->i_oe + 2 * ((size_t)oddeven * m->n_oe[0]);
            for (int isw = 0; isw < m->n_oe[oddeven]; ++isw) {
                int i = p[2 * isw], j = p[2 * isw + 1];
                size_t k = IX(m, i, j);
                if (oe[k] == 0) continue;
                double res = relax * (sor_lhs(m, chi, i, j) - B[k]) / m->sor_c[4 * n + k];
                chi[k] = chi[k] - res;
                anorm = anorm + res * res;
            }
            if (oddeven == 0 && l == 1) relax = 1.0 / (1.0 - .5 * (rJ * rJ));
            else relax = 1.0 / (1.0 - .25 * (rJ * rJ) * relax);
        }
        anorm = sqrt(anorm);
        if (anorm < epsilon) { *iters = l; *anorm_out = anorm; return 0; }
    }
    /* anorm already holds sqrt(sum res^2) of the last sweep (the reference's
     * warning line prints a further sqrt of it, ElSolv_SOR.f90:220) */
    *iters = max_count; *anorm_out = anorm;
    return 1;
}

/* Fixed-omega red-black SOR with a relative-residual stopping rule:
 * r = (L chi - B) at each cell when it is updated, stop when
 * sqrt(sum r^2) / sqrt(sum B^2) < rel_tol after a full (red+black) sweep. */
int orc_sor_fixed(orc_model *m, const double *B, double *chi, double omega, double rel_tol,
                  int64_t max_sweeps, int64_t *iters, double *rel_res) {
    sor_init(m);
    size_t n = (size_t)m->Nx * m->Ny;
    const int8_t *oe = m->ocean[ORC_GRID_ETA];
    double bnorm = 0.0;
    for (size_t k = 0; k < n; ++k) if (oe[k]) bnorm += B[k] * B[k];
    bnorm = sqrt(bnorm);
    double rn = 0.0;
    for (int64_t l = 1; l <= max_sweeps; ++l) {
        rn = 0.0;
        for (int oddeven = 0; oddeven < 2; ++oddeven) {
            const int *p = m->i_oe + 2 * ((size_t)oddeven * m->n_oe[0]);
            for (int isw = 0; isw < m->n_oe[oddeven]; ++isw) {
                int i = p[2 * isw], j = p[2 * isw + 1];
                size_t k = IX(m, i, j);
                if (oe[k] == 0) continue;
                double r = sor_lhs(m, chi, i, j) - B[k];
                chi[k] = chi[k] - omega * r / m->sor_c[4 * n + k];
                rn = rn + r * r;
            }
        }
        double rel = bnorm > 0.0 ? sqrt(rn) / bnorm : sqrt(rn);
        if (rel < rel_tol) { *iters = l; *rel_res = rel; return 0; }
    }
    *iters = max_sweeps; *rel_res = bnorm > 0.0 ? sqrt(rn) / bnorm : sqrt(rn);
    return 1;
}

/* The SOR stencil itself, (L chi)(i,j) on ocean eta points, 0 elsewhere. */
void orc_sor_apply(orc_model *m, const double *chi, double *out) {
    sor_init(m);
    const int8_t *oe = m->ocean[ORC_GRID_ETA];
    for (int j = 0; j < m->Ny; ++j)
        for (int i = 0; i < m->Nx; ++i)
            out[IX(m, i, j)] = oe[IX(m, i, j)] ? sor_lhs(m, chi, i, j) : 0.0;
}

void orc_destroy(orc_model *m) {
    if (!m) return;
    size_t n = 0; (void)n;
    free(m->ip0); free(m->ip1); free(m->im1); free(m->jp0); free(m->jp1); free(m->jm1);
    for (int g = 0; g < 4; ++g) { free(m->H[g]); free(m->bc[g]); }
    free_interp(&m->eta2u_noland); free_interp(&m->eta2v_noland); free_interp(&m->eta2H_noland);
    free_interp(&m->u2eta); free_interp(&m->v2eta); free_interp(&m->H2u); free_interp(&m->H2v);
    free_interp(&m->v2u); free_interp(&m->u2v);
    for (int l = 0; l < 2; ++l) { free(m->SWM_u[l]); free(m->SWM_v[l]); free(m->SWM_eta[l]); }
    for (int l = 0; l < NG; ++l) { free(m->G_u[l]); free(m->G_v[l]); free(m->G_eta[l]); }
    free(m->EDens); free(m->Pot); free(m->zeta); free(m->MU); free(m->MV);
    if (m->d.variant == ORC_FULLY_NONLINEAR) { free(m->D); free(m->Dh); free(m->Du); free(m->Dv); }
    free(m->u); free(m->v); free(m->eta);
    free(m->lat_mixing_u); free(m->lat_mixing_v); free(m->pll_coeff); free(m->plt_coeff);
    free(m->pll); free(m->plt); free(m->latmix_u); free(m->latmix_v);
    free(m->F_x); free(m->F_y); free(m->F_eta); free(m->F_x_const); free(m->F_y_const); free(m->F_eta_const);
    for (int t = 0; t < m->ntrc; ++t) {
        tracer_t *tr = &m->trc[t];
        free(tr->C); free(tr->C0); free(tr->gamma_C); free(tr->cons); free(tr->impl);
        for (int l = 0; l < 2; ++l) free(tr->CH[l]);
        for (int l = 0; l < TRC_NG; ++l) free(tr->G_CH[l]);
        free(tr->uhc); free(tr->vhc); free(tr->diff); free(tr->forcing);
    }
    free(m->TRC_coeff); free(m->sor_c); free(m->i_oe);
    for (int s = 0; s < m->d.n_tdf; ++s) { free(m->mc[s].var); free(m->mc[s].time); }
    for (int s = 0; s < m->d.n_dff; ++s) { free(m->dff_mc[s].var); free(m->dff_mc[s].time); }
    free(m->tdf_buf);
    free(m);
}
