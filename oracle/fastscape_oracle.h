/*
 * fastscape_oracle.h -- CPU restatement of fastscapelib-fortran's per-step landscape update.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product path: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The product never links it.
 *
 * Every routine restates one reference routine (cited file:line, paths relative to the
 * fastscapelib-fortran source tree) with the reference's operation order, so that a build with
 * -ffp-contract=off reproduces the reference's FP64 results bit for bit (libm pow aside).
 * Node indices are 0-based here (the reference is 1-based); basin ids are 0-based.
 */
#ifndef FASTSCAPE_ORACLE_H
#define FASTSCAPE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fso_state fso_state;

/* Statistics of the last LocalMinima call (for workload characterisation and tests). */
typedef struct {
  int nbasins;       /* basins found by compute_basins                       */
  int nopen;         /* basins whose outlet is a fixed (bc) node             */
  int nconn;         /* edges of the basin graph, virtual ones included      */
  int ntree;         /* accepted MST edges                                   */
  int nlocmin_before;
  int nlocmin_after;
  int max_tie;       /* largest group of equal real edge weights             */
  int depth;         /* longest receiver chain (nodes) after correction      */
  int ngs;           /* Gauss-Seidel iterations of the last SPL call         */
  long newton_iters; /* total Newton iterations of the last SPL call         */
  int marine_iters;  /* silt / sand diffusion iterations of the last Marine call */
} fso_stats;

/* ---- state lifecycle (src/FastScape_ctx.f90 Init/SetUp/Destroy) ---- */
fso_state *fso_create(int nx, int ny);
void fso_destroy(fso_state *s);

void fso_set_xl_yl(fso_state *s, double xl, double yl);
void fso_set_dt(fso_state *s, double dt);
void fso_set_bc(fso_state *s, int ibc);
void fso_set_u(fso_state *s, const double *u);
void fso_set_v(fso_state *s, const double *vx, const double *vy);  /* SetV: turns on Advect */
void fso_set_precip(fso_state *s, const double *precip);
void fso_init_h(fso_state *s, const double *h);   /* also sets b = h */
void fso_set_h(fso_state *s, const double *h);
void fso_set_basement(fso_state *s, const double *b);
void fso_set_erosional_parameters(fso_state *s, const double *kf, double kfsed, double m, double n,
                                  const double *kd, double kdsed, double g1, double g2, double p);
/* Extension: the reference hard-codes the Newton tolerance to 1e-3 (src/StreamPowerLaw.f90:146,344). */
void fso_set_newton_tol(fso_state *s, double tol);
void fso_reset_cumulative_erosion(fso_state *s);
/* SetMarineParam (src/FastScape_ctx.f90:454-472): turns Marine on after Diffusion in every step. */
void fso_set_marine_parameters(fso_state *s, double sealevel, double poro1, double poro2, double zporo1,
                               double zporo2, double ratio, double layer, double kdsea1, double kdsea2);
void fso_init_f(fso_state *s, const double *f); /* InitF: silt fraction Fmix (default 0.5) */
/* Activate_Strati (src/FastScape_ctx.f90:781-805): reflectors updated every step (run_Strati), fields of
 * Strati (src/Strati.f90:31-50) recomputed every nfreq steps; 0, or -1 for a bad schedule. */
int fso_set_strati(fso_state *s, int nstep, int nreflector, int nfreq, double vex);
/* reflectors (nref x nn), fields (nref x 10 x nn, field l of reflector i at (i*10+l)*nn), the active
 * reflector (1-based), and 1 when the last step was an output step; any pointer may be NULL */
void fso_strati_state(const fso_state *s, double *refl, double *fields, int *iref, int *out);
void fso_distance_to_shore(const double *h, double *d, int nx, int ny, const int *rec, double xl, double yl);
/* SiltSandCouplingDiffusion (src/Marine.f90:171-499) on caller arrays: h and f in place, q1 / q2 the
 * silt / sand supply; returns the iteration count, or < 0 (zero pivot, no convergence in 1000). */
int fso_silt_sand(double *h, double *f, const double *q1, const double *q2, int nx, int ny, double dx, double dy,
                  double dt, double sealevel, double layer, double kdsea1, double kdsea2, int ibc);

int fso_execute_step(fso_state *s);
int fso_get_step(const fso_state *s);
void fso_get_stats(const fso_state *s, fso_stats *out);

/* field ids for fso_copy_field */
enum {
  FSO_H = 0, FSO_B, FSO_ETOT, FSO_ERATE, FSO_AREA, FSO_LAKE_DEPTH, FSO_LENGTH, FSO_HWATER, FSO_SEDFLUX, FSO_FMIX
};
enum { FSO_REC = 0, FSO_NDON, FSO_STACK, FSO_MNREC, FSO_MSTACK };
void fso_copy_field(const fso_state *s, int field, double *dst);
void fso_copy_ifield(const fso_state *s, int field, int *dst);
void fso_copy_don(const fso_state *s, int *dst);   /* (nn,8) row-major: dst[8*ij+k] */
void fso_copy_mrec(const fso_state *s, int *mrec, double *mwrec, double *mlrec);

/* read-outs after a step: FastScape_Copy_Slope / _Curvature / _Chi, FastScape_Get_Fluxes
 * (out = tectonic, erosion, boundary flux) */
void fso_copy_slope(const fso_state *s, double *out);
void fso_copy_curvature(const fso_state *s, double *out);
void fso_copy_chi(const fso_state *s, double *out);
void fso_get_fluxes(const fso_state *s, double out[3]);
/* the same on caller-owned arrays (order: receivers before donors; mrec/mwrec 8 slots per node) */
void fso_slope(const double *h, int nx, int ny, double dx, double dy, double *out);
void fso_curvature(const double *h, int nx, int ny, double dx, double dy, double *out);
void fso_chi(const int *order, const int *rec, const double *a, const double *length, int nn, double dx,
             double dy, double m, double n, double *out);
void fso_catchment(const int *rec, int nn, double *out); /* rec: receivers before LocalMinima */
void fso_fluxes_sfd(const int *order, const int *rec, const double *u, const double *erate,
                    const unsigned char *bc, int nn, double surf, double out[3]);
void fso_fluxes_mfd(const int *mstack, const int *mnrec, const int *mrec, const double *mwrec, const double *u,
                    const double *erate, const unsigned char *bc, int nn, double surf, double out[3]);

/* ---- individual routines, stateless, for unit tests ---- */
void fso_bc_decode(int ibc, int nx, int ny, int *i1, int *i2, int *j1, int *j2, int *xcyc, int *ycyc,
                   unsigned char *bc);
void fso_find_receiver(const double *h, int nx, int ny, double dx, double dy, int *rec, double *length,
                       int i1, int i2, int j1, int j2, int xcyc, int ycyc);
void fso_find_donor(const int *rec, int nn, int *ndon, int *don);
void fso_find_stack(const int *rec, const int *don, const int *ndon, int nn, int *stack);
void fso_indexx(int n, const double *arr, int *indx);   /* indx: 1-based positions, as the reference */
int fso_local_minima(int *stack, int *rec, const unsigned char *bc, int *ndon, int *don, const double *h,
                     double *length, int nx, int ny, double dx, double dy, fso_stats *st);
/* basin graph of connect_basins, for tests of the edge order: returns nconn, fills up to cap */
int fso_connect_basins_dump(const int *stack, const int *rec, const unsigned char *bc, const double *h,
                            int nx, int ny, int cap, int *basin0, int *cb, int *cn, double *cw);

#ifdef __cplusplus
}
#endif
#endif
