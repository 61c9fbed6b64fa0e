/*
 * lee_oracle.h -- CPU oracle for the LEE + porous-ground time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, single-domain restatement of the
 * reference solver's FB_p_v time step (azhang22/3D_MPI_TEST).  It is the checker for the
 * HIP product path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and never as the thing measured or shipped.
 *
 * Reference files followed (paths relative to 3D_MPI_TEST/):
 *   air.cpp 40-256 (wind), 509-937 (PML set-up), 1286-1304 (wind per frame),
 *           1476-1949 (velocity stencil), 1953-2322 (pressure stencil),
 *           2323-2439 (PML auxiliaries), 2447-2730 (PML shift per frame);
 *   calculation_2D.cpp 12-236 (grid), 292-314 (IC), 317-352 (receiver lookup),
 *           354-369 (nn->n), 387-403 (receiver sample), 405-496 (moving frame + metrics),
 *           548-613 (FB add loops), 646-1028 (BCs, interface, TDBC);
 *   calculation_2D.h 71-75 (TDBC poles); porous.cpp 40-143; df_over_var.cpp 5-9;
 *   airpore.cpp 17-316 (derived parameters), 324-398 (phase order), 400-641 (time loop),
 *           881-897 (frame order); FFT.cpp 15-121.
 */
#ifndef LEE_ORACLE_H
#define LEE_ORACLE_H

#include "../include/lee_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

/* Build the oracle from a parsed deck.  nrec/rec override the reference receiver rule
 * when nrec > 0.  fill_nn_nan: initialise the time-level n+1 buffers with NaN instead of
 * zero (checks that no stale boundary value ever leaks into the state).  Returns NULL on
 * error (message via orc_error()). */
orc_sim *orc_create(const lee_deck *d, int nrec, const double *rec, int fill_nn_nan);
/* The same with the sidecar options: receivers, and the extensions of lee_options
 * (integrator = LEE_INT_RK4, terrain hills, log wind of velocity_method 7). */
orc_sim *orc_create_opt(const lee_deck *d, const lee_options *opt, int fill_nn_nan);
void     orc_destroy(orc_sim *s);
const char *orc_error(void);

void orc_dims(const orc_sim *s, int *dims /* nx ny nz pnx pny pnz ijkmax nrec */);
int  orc_receiver_indices(const orc_sim *s, int *ijk);

int  orc_frame_setup(orc_sim *s, int mf_count);
int  orc_frame_steps(const orc_sim *s, int mf_count, int temp_time);
int  orc_frame_count(const orc_sim *s);
/* nsteps steps, sampling receivers before each step into rec (nsteps*nrec). */
void orc_step(orc_sim *s, int nsteps, double *rec);
/* Whole time loop (airpore.cpp:539-624); rec gets time_domain*nrec samples. */
int  orc_run(orc_sim *s, double *rec);

void orc_get_state(const orc_sim *s, int domain, double *u, double *v, double *w, double *p);
void orc_set_state(orc_sim *s, int domain, const double *u, const double *v,
                   const double *w, const double *p);
void orc_get_psi(const orc_sim *s, double *psi);
/* Embedded obstacle: returns curve_judge (0 = none); box = istar istop jstar jstop kstar
 * kstop; mask (may be NULL) gets the material id per box node, i fastest. */
int  orc_obstacle(const orc_sim *s, int *box, int *mask);
/* Time-series source of calculation_2D::UpdateSource; n = 0 removes it.  0 or -1. */
int  orc_set_source_series(orc_sim *s, const double *ps, long n, int period);
/* Ground impedance histories [i + nx*j][ground_poles] (nothing without a ground impedance). */
void orc_get_ground_psi(const orc_sim *s, double *psi);

/* Increments only (no add, no BC): the stencils air::cal_fp / air::cal_fuvw /
 * porous::cal_fp / porous::cal_fuvw applied to the given fields, written to out
 * (interior points; other points untouched).  Used to pin the stencils against the
 * reference's own compiled code. */
void orc_air_dp(orc_sim *s, const double *u, const double *v, const double *w,
                const double *p, double *dp);
void orc_air_duvw(orc_sim *s, const double *u, const double *v, const double *w,
                  const double *p, double *du, double *dv, double *dw);
void orc_pore_dp(orc_sim *s, const double *u, const double *v, const double *w, double *dp);
void orc_pore_duvw(orc_sim *s, const double *u, const double *v, const double *w,
                   const double *p, double *du, double *dv, double *dw);
/* Access to the PML auxiliary slabs: slab 0..5 = x1 x2 y1 y2 z1 z2, field 0..3 = u v w p.
 * Returns the element count (0 if the slab is inactive); copies into/out of buf. */
long orc_q_size(const orc_sim *s, int slab);
void orc_q_get(const orc_sim *s, int slab, int field, double *buf);
void orc_q_set(orc_sim *s, int slab, int field, const double *buf);
void orc_update_q(orc_sim *s, int which /* 0 = Qp from p_nn, 1 = Quvw from uvw_nn */,
                  const double *u, const double *v, const double *w, const double *p);
/* Advance only the y1 time-gate counter as air::cal_fp does (air.cpp:1965-1969). */
void orc_tick_gate(orc_sim *s);
void orc_wind(const orc_sim *s, double *U, double *V, double *W);
/* Inverse metrics xX xY xZ yX yY yZ zX zY zZ and Jacobian at an interior node. */
void orc_metric(const orc_sim *s, int domain, int i, int j, int k, double *out10);

/* Reference FFT (FFT.cpp:15-82) and spectrum writer (FFT.cpp:84-121). */
void orc_set_threads(int n);
int orc_get_threads(void);
void orc_fft(int dir, int m, double *x, double *y);

#ifdef __cplusplus
}
#endif

#endif
