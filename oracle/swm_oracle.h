/*
 * CPU oracle for the SWM time-stepping hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the parity checker and the CPU baseline.  It restates the
 * reference Fortran pass for pass (same loop nests, same per-cell coefficient
 * arrays and gather interpolation tables, same expression order, OpenMP
 * schedule(static,5) over j).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in the product links it.
 *
 * Arrays are the reference's Fortran (Nx,Ny) arrays viewed from C as [Ny][Nx]
 * (i fastest), 0-based.  Masks are int8 0/1 ("ocean" masks; land = 1 - ocean).
 */
#ifndef SWM_ORACLE_H
#define SWM_ORACLE_H
#include <stdint.h>

#define ORC_MAX_STREAMS 4
#define ORC_MAX_TRACERS 4

enum { ORC_GRID_H = 0, ORC_GRID_U = 1, ORC_GRID_V = 2, ORC_GRID_ETA = 3 };
enum { ORC_FULLY_NONLINEAR = 1, ORC_LINEARISED_MEAN_STATE = 2, ORC_LINEARISED_STATE_OF_REST = 3 };
enum { ORC_COMP_X = 0, ORC_COMP_Y = 1, ORC_COMP_ETA = 2 };
enum { ORC_TDF_WIND = 0, ORC_TDF_CUSTOM = 1, ORC_TDF_HEATING = 2 };

typedef struct orc_grid {
    const double *H;        /* (Nx,Ny) layer thickness on this grid, as left by initDomain/computeD */
    const int8_t *ocean;    /* (Nx,Ny) ocean mask */
    const double *cos_lat;  /* (Ny) */
    const double *tan_lat;  /* (Ny) */
    const double *f;        /* (Ny) Coriolis parameter */
} orc_grid;

typedef struct orc_desc {
    int32_t Nx, Ny;
    int32_t variant;             /* ORC_FULLY_NONLINEAR | ORC_LINEARISED_MEAN_STATE | ORC_LINEARISED_STATE_OF_REST */
    int32_t lateral_mixing;      /* model.h LATERAL_MIXING */
    int32_t quadratic_friction;  /* model.h QUADRATIC_BOTTOM_FRICTION */
    int32_t barotropic;          /* model.h BAROTROPIC */
    int32_t has_tau_scale;       /* model.h TAU_SCALE defined */
    int32_t n_wind, n_osc, n_tracer;
    double tau_scale;            /* value of the TAU_SCALE literal as the Fortran compiler sees it */
    double G, Ah, dt, A, dLambda, dTheta, RHO0, minD, lbc, ab_chi;
    orc_grid grid[4];            /* H, u, v, eta */
    const double *impl_u, *impl_v, *impl_eta;
    const double *gamma_sq_u, *gamma_sq_v;            /* only read with quadratic_friction */
    const double *F_x_const, *F_y_const, *F_eta_const;
    /* time-dependent windstress streams (constant-in-time data, re-applied every step) */
    int32_t wind_comp[ORC_MAX_STREAMS];
    const double *wind_tau[ORC_MAX_STREAMS];
    /* oscillating forcing streams */
    int32_t osc_comp[ORC_MAX_STREAMS];
    const double *osc_re[ORC_MAX_STREAMS], *osc_im[ORC_MAX_STREAMS];
    double osc_omega[ORC_MAX_STREAMS];
    /* LINEARISED_MEAN_STATE basic state (H grid), NULL otherwise */
    const double *u_bs, *v_bs, *zeta_bs;
    /* initial conditions (unmasked; masked as in SWM_initialConditions) */
    const double *u0, *v0, *eta0;
    /* tracers */
    double kappa_h[ORC_MAX_TRACERS];
    const double *trc_C[ORC_MAX_TRACERS], *trc_C0[ORC_MAX_TRACERS];
    const double *trc_gamma_C[ORC_MAX_TRACERS], *trc_cons[ORC_MAX_TRACERS];
    /* time-interpolated forcing streams on memchunk data (memchunk_module.f90),
     * applied after the wind and oscillating streams in the order given */
    int32_t n_tdf;
    int32_t tdf_kind[ORC_MAX_STREAMS];      /* ORC_TDF_WIND, ORC_TDF_CUSTOM, ORC_TDF_HEATING */
    int32_t tdf_comp[ORC_MAX_STREAMS];      /* ORC_COMP_* */
    int32_t tdf_nrec[ORC_MAX_STREAMS];      /* records in the dataset */
    int32_t tdf_chunk[ORC_MAX_STREAMS];     /* namelist chunksize */
    double tdf_dt[ORC_MAX_STREAMS];         /* record spacing in model seconds (getDt) */
    const double *tdf_rec[ORC_MAX_STREAMS]; /* nrec (Nx,Ny) records back to back */
    /* DYNFROMFILE datasets (dynFromFile_module.f90): the field each adds to
     * the master fields at model_advance (ORC_COMP_X = u, _Y = v, _ETA = eta) */
    int32_t n_dff;
    int32_t dff_field[3];
    int32_t dff_nrec[3];
    int32_t dff_chunk[3];                   /* namelist DFF_Nt_chunksize */
    double dff_dt[3];
    const double *dff_rec[3];
} orc_desc;

typedef struct orc_model orc_model;

orc_model *orc_create(const orc_desc *d);
void orc_destroy(orc_model *m);
/* advance nsteps time steps (SWM_timestep, TRC_timestep, model_advance) */
void orc_step(orc_model *m, int64_t nsteps);
int64_t orc_itt(const orc_model *m);
/* copy a named field (register-like names) into out (Nx*Ny doubles); returns 0 on success */
int orc_get(const orc_model *m, const char *name, int tracer, double *out);
/* main_ElSolv_SOR contract: reference (Chebyshev) mode */
int orc_sor(orc_model *m, const double *B, double *chi, double epsilon, int first_guess,
            int64_t *iters, double *anorm);
/* fixed-omega mode with relative-residual stopping rule */
int orc_sor_fixed(orc_model *m, const double *B, double *chi, double omega, double rel_tol,
                  int64_t max_sweeps, int64_t *iters, double *rel_res);
/* apply the SOR operator (L chi) on ocean eta points */
void orc_sor_apply(orc_model *m, const double *chi, double *out);
/* computeStreamfunction(u, v, psi) without an elliptic solver (calc_lib.f90:727-780) */
void orc_psi(const orc_model *m, const double *u, const double *v, double *psi);
/* computeDivergence(u, v, div, u_grid, v_grid) */
void orc_divergence(const orc_model *m, const double *u, const double *v, double *div);
void orc_nondivergent(const orc_model *m, const double *u, const double *v, const double *chi, double *u_nd,
                      double *v_nd);
#endif
