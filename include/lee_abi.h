/*
 * lee_abi.h -- C ABI of the MI355X linearised-Euler (LEE) + porous-ground solver.
 *
 * This is the drop-in boundary for the time-stepping hot path of the CPU/MPI solver
 * azhang22/3D_MPI_TEST.  Every entry point is extern "C" with plain pointers, sizes and
 * POD structs; no C++ or torch types cross it.  Each declaration names the reference
 * interface it replaces (paths relative to the reference repository, 3D_MPI_TEST/).
 *
 * Conventions
 *   - Status: int, 0 = LEE_OK, < 0 = error.  A failed call leaves a message readable
 *     with lee_last_error(ctx) (or lee_global_error() when no context exists).  Nothing
 *     is printed and nothing "continues after an error" (the reference prints and
 *     carries on: calculation_2D.cpp:49-52, 719, 862).
 *   - Ownership: the library owns all device memory.  Host buffers passed in belong to
 *     the caller and are only read/written during the call.
 *   - Threading: one host thread per context; each context owns one HIP stream.
 *     Distinct contexts may be used from distinct threads.
 *   - Layout: 3-D fields are SoA, x (i) fastest: idx = i + nx*(j + ny*k).  Host-side
 *     state buffers are FP64 regardless of the device precision.
 */
#ifndef LEE_ABI_H
#define LEE_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LEE_ABI_VERSION 1

enum {
    LEE_OK = 0,
    LEE_ERR_ARG = -1,         /* invalid argument / inconsistent deck */
    LEE_ERR_DEVICE = -2,      /* HIP runtime error */
    LEE_ERR_IO = -3,          /* file missing or malformed */
    LEE_ERR_UNSUPPORTED = -4, /* deck selects a path the reference never implemented */
    LEE_ERR_NONFINITE = -5,   /* NaN/Inf detected in the state by the health check */
    LEE_ERR_STATE = -6        /* call out of order (e.g. step before frame setup) */
};

/* Boundary types of input.txt line 4 (airpore.cpp:48-159). */
enum { LEE_BC_RIGID = 1, LEE_BC_ABSORBING = 2, LEE_BC_POROUS = 3, LEE_BC_RADIATION = 4, LEE_BC_TDBC = 5 };
/* Face slots of input.txt line 4, in file order (airpore.cpp:48, 52-159). */
enum { LEE_FACE_N = 0, LEE_FACE_S = 1, LEE_FACE_E = 2, LEE_FACE_W = 3, LEE_FACE_F = 4, LEE_FACE_B = 5 };

/*
 * lee_deck: the parsed drop-in input contract -- input.txt (airpore.cpp:17-316) plus the
 * coordi_air{1,2,3}.dat / coordi_pore{1,2,3}.dat grid files (calculation_2D.cpp:44-68,
 * 154-204).  Field names follow the reference members they hold.
 */
typedef struct lee_deck {
    double gauss_width;                          /* L1  airpore.cpp:20 */
    int    mf_jmax, mf_lead_dj, mf_trail_dj;     /* L2  airpore.cpp:21 (MovingFrame) */
    int    mf_judge;
    double mf_ratio;
    int    scheme;                               /* L3  1 FB_v_p, 2 FB_p_v, 5 RK4 (airpore.cpp:24-43) */
    int    bc[6];                                /* L4  N S E W F B types (airpore.cpp:48) */
    int    time_domain, out_difft, fft_m;        /* L6  airpore.cpp:171 */
    double frequency;
    double velocity_coef;                        /* L8  airpore.cpp:175 */
    int    velocity_method;
    double gamma, Pav, c;                        /* L9  adiabatic_coef, Pav, sound_speed */
    double pml_alpha[3], pml_R[3];               /* L10-12 x, y, z: alpha, resistivity */
    int    pml_y_time;                           /* L11 PMLy time_step (y1 gate) */
    double pore_q, pore_porosity, pore_resistivity; /* L13 PorePara.q (sqrt tortuosity), Omega, sigma */
    double src[3], rcv[3];                       /* L14 source.xyz, receiver.xyz */
    double dt;                                   /* L16 AirStep.diff_t */
    double dx, dy, dz;                           /* L17 AirStep.diff_x/y/z */
    int    mpi_dims[3];                          /* L19 (ignored: one GPU holds the domain) */
    int    restart, restart_out;                 /* L22 (restart is not on the hot path) */
    int    hill_judge;                           /* L24 curve_judge: embedded porous obstacle 0..5 */
    double hill_coef[12];                        /* L24 curve_coefa .. curve_coefl */
    int    hill_points;                          /* L24 curve_points */
    double hill_mat[3][3];                       /* L25-27 hill, canopy, trunk: q (Cs), porosity, resistivity */
    double hill_box[6];                          /* L29 xstar xstop ystar ystop zstar zstop */
    double pore_dx, pore_dy, pore_dz;            /* L32 SouthStep (read only if South = porous) */
    /* coordi_air1.dat header: IMAX IMAX1 IMAX2 JMAX JMAX1 JMAX2 KMAX KMAX1 KMAX2 IJKMAX */
    int    air_hdr[10];
    const double *air_x, *air_y, *air_z;         /* IMAX, JMAX, KMAX coordinates */
    int    has_pore;                             /* 1 when coordi_pore*.dat were read */
    int    pore_hdr[10];
    const double *pore_x, *pore_y, *pore_z;
} lee_deck;

/*
 * lee_options: everything input.txt cannot say (the sidecar).  Zero-initialise and set
 * what you need; lee_options_default() gives the reference behaviour.
 */
typedef struct lee_options {
    int precision;          /* 64 (default, reference FP64) or 32 (FP32 state, FP64 coefficients) */
    int n_receivers;        /* 0: the reference's receiver rule (airpore.cpp:413-458) */
    const double *receivers;/* n_receivers * (x,y,z) when n_receivers > 0 */
    int health_every;       /* check the state for NaN/Inf every N steps (0 = never) */
    int use_graph;          /* 1: replay each step from a captured hipGraph */
    int device;             /* HIP device ordinal */
    int path;               /* LEE_PATH_AUTO (0), LEE_PATH_SPLIT, LEE_PATH_FUSED */
    /* ---- extensions beyond the reference (no input.txt field can express them) ---- */
    int integrator;         /* LEE_INT_DECK (0): the deck's scheme (FB_p_v); LEE_INT_RK4 */
    int n_hills;            /* terrain-following grid over n_hills Gaussian hills (0: flat) */
    double hills[32];       /* per hill: x0, y0 [m, frame-0 coordinates], height, sigma [m] */
    double wind_ustar;      /* velocity_method 7: V(z) = wind_ustar/0.41 ln(z/wind_z0) along +y, */
    double wind_z0;         /*   z = height above the local ground, V = 0 for z <= wind_z0 */
    /* ground impedance (no counterpart in the reference): the air's plane k = 0 becomes the
     * ground's interface plane -- its pressure and tangential velocities follow the porous
     * equations of the deck's top material (pore_q, pore_porosity, pore_resistivity;
     * porous.cpp:44-50, 77-143), as the porous layer's top plane does through the interface
     * copy (calculation_2D.cpp:878-886) -- and the layer below it is the impedance
     * Z(s) = ground_zinf (1 + sum_k ground_a[k] / (ground_r[k] + s ground_tau)) through the
     * East TDBC's recursion (calculation_2D.cpp:986-1028) with K poles: the normal velocity
     * half a cell below the plane follows the porous momentum equation with the impedance's
     * pressure half a cell further down.  On the terrain-following grid (n_hills > 0) the
     * plane takes the metric of a porous top plane shifted rigidly with the ground.  FB_p_v,
     * deck South face rigid, fixed frame, no obstacle; residues and poles >= 0 (passive). */
    int ground_poles;       /* 0: off, else 1..LEE_MAX_POLES pole/residue pairs */
    double ground_a[8], ground_r[8];
    double ground_zinf, ground_tau;
    /* RK4 air-stage launch shape: 0 = autotune on the first steps (times the candidates on the
     * context's own buffers), 1..11 = candidate 0..10 (64x4 plane order, 64x4 / 64x8 XCD
     * columns, 64x8 plane order, 64x16 XCD columns, 64x4 tiles split into the interior outside
     * every layer and the layer shell, the same with the shell split by slab axes, those two
     * with the interior's stage state staged through LDS, and with the shell's staged as
     * well).  All give the same bits. */
    int rk_tile;
    /* FB_p_v air-pass launch form: 0 = autotune on the first step among the forms `path`
     * allows (times each form's air and face passes on the coming step), 1 = the split
     * passes (in FP32 the interior box four points per lane, 16-byte accesses, when there is
     * no mean flow and nx is a multiple of 4; over terrain in the velocity pass only), 2 = the
     * same with each plane's tiles padded to a multiple of 8 (a tile column on one XCD),
     * 3..6 = the fused pass with pencils deep enough for about 2, 4, 8, 16 blocks per CU,
     * 7 = the split passes with one point per lane throughout.  1, 2 and 7 need path AUTO or SPLIT, 3..6 AUTO or FUSED.  All give
     * the same bits. */
    int fb_form;
} lee_options;

/* Time integrators.  LEE_INT_DECK runs the deck's scheme (the reference implements only
 * FB_p_v, calculation_2D.cpp:551-612).  LEE_INT_RK4 is classical four-stage Runge-Kutta
 * on the same semi-discrete operator (all four fields from one stage state, the PML
 * auxiliaries as ODE state, face conditions re-applied after every stage); the porous drag
 * enters through its integrating factor (Lawson's RK4: exact for the drag, no step limit
 * from sigma dt / rho_e).  The reference parses scheme 5 = RK4 but never integrates with it. */
enum { LEE_INT_DECK = 0, LEE_INT_RK4 = 1 };
#define LEE_MAX_HILLS 8
#define LEE_MAX_POLES 8

/* Kernel path of the time step.  SPLIT runs the reference's two phases as separate passes
 * over the grid; FUSED computes both phases of a plane in one pass (one read and one write
 * of u, v, w, p per point).  Both give the same bits.  AUTO lets the FB autotuner choose
 * among the forms of both (lee_options.fb_form); on MI355X it keeps the split passes.  FUSED
 * needs the porous interface copy when South is porous. */
enum { LEE_PATH_AUTO = 0, LEE_PATH_SPLIT = 1, LEE_PATH_FUSED = 2 };

typedef struct lee_info {
    int nx, ny, nz;                 /* air frame dimensions (IMAX, frame JMAX, KMAX) */
    int pnx, pny, pnz;              /* porous dimensions (0 when South is not porous) */
    int ijkmax;                     /* whole propagation length in j (moving frame) */
    int n_receivers;
    int precision;
    int mf_count;                   /* frame set up last (-1 before lee_frame_setup) */
    long long steps_done;           /* time steps advanced so far */
    long long pml_y1_counter;       /* reference PML_JMAX1 step counter */
    int path;                       /* LEE_PATH_SPLIT or LEE_PATH_FUSED, as chosen */
    int rk_tile;                    /* RK4 air-stage launch shape in use (1..11), 0 before tuning */
    int fb_form;                    /* FB_p_v air-pass launch form in use (1..7), 0 before tuning */
} lee_info;

typedef struct lee_ctx lee_ctx;

/* Parse input.txt + coordi_* in `dir` (airpore::airpore airpore.cpp:17-316 and the grid
 * reader in calculation_2D::calculation_2D calculation_2D.cpp:44-68, 154-204).  Arrays in
 * `out` are owned by the deck; release them with lee_deck_free. */
int  lee_deck_load(const char *dir, lee_deck *out);
void lee_deck_free(lee_deck *deck);
void lee_options_default(lee_options *opt);

/* Build a solver context: validates the deck, derives every coefficient in FP64 on the
 * host (Aav airpore.cpp:177, PML airpore.cpp:182-247 + air.cpp:509-937, porous
 * airpore.cpp:249-254 + porous.cpp:40-51, TDBC calculation_2D.h:71-75), allocates the
 * SoA fields in HBM and sets the Gaussian initial condition
 * (calculation_2D::Set_InitialCond calculation_2D.cpp:300-314).
 * Replaces: airpore::SetInitialCond airpore.cpp:936-964. */
int  lee_create(const lee_deck *deck, const lee_options *opt, lee_ctx **out);
int  lee_destroy(lee_ctx *ctx);
const char *lee_last_error(const lee_ctx *ctx);
const char *lee_global_error(void);
int  lee_get_info(const lee_ctx *ctx, lee_info *info);

/* Receiver grid indices (ii, jj, kk per receiver, whole-domain j), as located by
 * calculation_2D::get_position (calculation_2D.cpp:317-352) and clamped as in
 * airpore.cpp:454-457. */
int  lee_receiver_indices(const lee_ctx *ctx, int *ijk, int cap);

/* Set up moving-frame number mf_count: shift fields and PML auxiliaries, recompute
 * coordinates, metrics and wind.  mf_count 0 is the initial frame (zeroes the PML
 * auxiliaries).  Replaces airpore::SetMovingFrame airpore.cpp:881-897
 * (calculation_2D::SetMovingFrame calculation_2D.cpp:405-496, air::SetWindProfile
 * air.cpp:1286-1304, air::SetQMove air.cpp:2447-2730). */
int  lee_frame_setup(lee_ctx *ctx, int mf_count);

/* Advance nsteps time steps.  Before each step the receivers are sampled from p_n
 * (airpore.cpp:573-580, calculation_2D::get_pressure calculation_2D.cpp:387-403; |p| <
 * 1e-100 is written as 0).  rec_out (may be NULL) receives nsteps*n_receivers FP64
 * samples, step-major.  One step is airpore::Cal_pressure + airpore::Cal_velocity +
 * airpore::UpdateInitialCond (airpore.cpp:589-614). */
int  lee_step(lee_ctx *ctx, int nsteps, double *rec_out);

/* Number of steps the reference runs in frame mf_count given temp_time steps already
 * done (the MF_limit rule of airpore.cpp:542-560), and the number of frames
 * (airpore.cpp:624 loop bound). */
int  lee_frame_steps(const lee_ctx *ctx, int mf_count, int temp_time);
int  lee_frame_count(const lee_ctx *ctx);

/* The whole time loop of airpore::get_output (airpore.cpp:539-624): for each frame,
 * frame set-up then MF_limit steps, until time_domain steps.  rec_out receives
 * time_domain*n_receivers samples (cap = its capacity in doubles). */
int  lee_run(lee_ctx *ctx, double *rec_out, long long cap);

/* Copy the current state (time level n) to or from the host as FP64 SoA arrays.
 * domain 0 = air (nx*ny*nz), 1 = porous ground (pnx*pny*pnz). */
int  lee_get_state(lee_ctx *ctx, int domain, double *u, double *v, double *w, double *p);
int  lee_set_state(lee_ctx *ctx, int domain, const double *u, const double *v,
                   const double *w, const double *p);
/* TDBC history Psi on the East face, [i][k][6] (calculation_2D.cpp:128-134). */
int  lee_get_psi(lee_ctx *ctx, double *psi, long long cap);

/* Ground impedance histories [i + nx*j][ground_poles] (FP64), when lee_options.ground_poles > 0. */
int  lee_get_ground_psi(lee_ctx *ctx, double *psi, long long cap);
/* Device-side time of the last lee_step call (HIP events on the context stream), ms. */
double lee_last_step_ms(const lee_ctx *ctx);

/* Per-kernel device time of the stencil kernels, measured with HIP events around each
 * launch on the context stream while profiling is on.  No reference counterpart (the
 * reference prints clock() per step, airpore.cpp:587-595); used by bench.py for the
 * roofline.  bytes_per_launch = compulsory HBM bytes of one launch (every array element
 * the kernel must read or write counted once). */
typedef struct lee_kstat {
    char name[32];
    long long launches;
    double total_ms;
    double bytes_per_launch;
    double points_per_launch;
} lee_kstat;
/* enable: 0 off, 1 every kernel kind, 2 the air stencil launches only (the dominant kernels,
 * with fewer timing marks in the stream) */
int  lee_profile(lee_ctx *ctx, int enable);
int  lee_kernel_stats(const lee_ctx *ctx, lee_kstat *out, int cap);

/* Plain device copy dst[i] = src[i] of n 4-byte words on the context stream, timed with
 * HIP events (returns ms).  No reference counterpart: calibrates HBM counters and the
 * copy roofline for bench.py. */
double lee_copy_bench(lee_ctx *ctx, long long n_words, int repeats);
/* Many-stream reference (no counterpart in the reference): a kernel that reads `reads` and
 * writes `writes` arrays of n_words 4-byte words, one word of each per lane, all reads in
 * flight before its stores -- the bandwidth a stencil stage with that many field streams can
 * reach.  (reads, writes) one of (1,1) (4,4) (8,4) (12,8) (20,8) (28,8).  Returns ms per
 * launch (mean over `repeats` after a warm-up), -1 on error.  Allocates and frees its arrays. */
double lee_stream_bench(lee_ctx *ctx, long long n_words, int reads, int writes, int repeats);

/* Checkpoint in the reference's restart text format (airpore::Save_restart airpore.cpp:664-700,
 * airpore::save_restartfile 783-796, calculation_2D::save_restart_cal calculation_2D.cpp:497-520,
 * air::save_restart_air air.cpp:1305-1390): a "1 0 index" line, then u v w p per node (i, j, k
 * loops), the mean flow, the PML auxiliaries, the porous fields, all as 10-digit scientific
 * text.  A trailing block the reference's reader never reaches adds what the reference does
 * not save: frame number, PML y1 counter, steps done and the East TDBC histories.  The file
 * name and directory (myfile/ctNNN_rank.dat in the reference) are the caller's. */
int  lee_save_restart(lee_ctx *ctx, const char *path, int index);
/* Read such a file into the context (airpore::Read_restart airpore.cpp:799-862): the fields,
 * auxiliaries and (if present) the extension block; the mean flow is rebuilt from the deck.
 * *index receives the step index of the header.  Call lee_frame_setup for the file's frame
 * first; a mismatch is LEE_ERR_STATE. */
int  lee_load_restart(lee_ctx *ctx, const char *path, int *index);

/* Contour file of the current pressure (airpore::get_data_contour airpore.cpp:1018-1050,
 * calculation_2D::mpi_send_data calculation_2D.cpp:1030-1105, getOstream 1165-1182): every
 * 4th node plus the last on each axis, Tecplot POINT zones "X Y Z P" at scientific precision 6
 * (an air zone, and a porous zone when South is porous), written to dir/ptNNN.dat with
 * NNN = (steps done + 1) / out_difft.  The reference calls it when the steps done are a
 * multiple of out_difft (airpore.cpp:615-618) and writes to a fixed absolute path. */
int  lee_write_contour(lee_ctx *ctx, const char *dir);
/* The same contour file, written without allocating device memory: no counterpart in the
 * reference; for callers that cannot allocate (code that runs while a HIP graph is being
 * captured, tight serving loops, harnesses that forbid hipMalloc).  lee_write_contour
 * allocates a device buffer per call for the subsample; this variant stages it in a fixed
 * 4 MiB array per GPU (the library holds it until the process ends), so each zone may have
 * at most 524288 subsampled nodes (LEE_ERR_ARG beyond that, before anything is written).
 * Calls that overlap on one GPU share that array and race. */
int  lee_write_contour_noalloc(lee_ctx *ctx, const char *dir);

/* Time-series source (calculation_2D::UpdateSource calculation_2D.cpp:371-385, the series
 * of source_input.dat read at calculation_2D.cpp:220-235, the call at airpore.cpp:608-611):
 * after step N (steps counted from 1 as index_pressure, airpore.cpp:572), when
 * N % period == 0 and N / period < n, the pressure of the new level gains
 * ps[N / period] * exp(-gc r^2) at every air node, r the distance to the deck's source and
 * gc = 4 ln 2 / width^2 of the initial condition.  The reference uses period 23 and ships
 * the call switched off.  n = 0 removes the series.  Steps then run without graph replay. */
int  lee_set_source_series(lee_ctx *ctx, const double *ps, long long n, int period);
/* Reads up to cap whitespace-separated values of dir/source_input.dat into ps (the
 * reference reads 100001 values into a 100000-element array); returns the count, or a
 * negative status. */
long long lee_read_source_series(const char *dir, double *ps, long long cap);

/* Ensemble: n independent domains (decks[d] with opts[d]; opts may be NULL) stepped together,
 * each RK4 stage as one launch per kernel kind over every member -- the ensemble runner of
 * parameter sweeps.  No counterpart in the reference, which runs one domain per MPI job.
 * Members must share precision, wind kind (none, U V W, or V only) and grid kind (flat or
 * terrain), integrate with RK4 and keep a fixed frame.  Each member is a full lee_ctx
 * (lee_ensemble_member) for state, restart and contour calls between steps; do not step a
 * member on its own.  rec_out takes nsteps rows of all members' receivers, member d's at
 * columns offsets[d] .. offsets[d+1]-1 (lee_ensemble_receivers fills n+1 offsets).
 * use_graph replays captured chunks of steps. */
typedef struct lee_ensemble lee_ensemble;
int  lee_ensemble_create(int n, const lee_deck *decks, const lee_options *opts, lee_ensemble **out);
int  lee_ensemble_destroy(lee_ensemble *e);
int  lee_ensemble_size(const lee_ensemble *e);
int  lee_ensemble_member(lee_ensemble *e, int d, lee_ctx **ctx);
int  lee_ensemble_receivers(const lee_ensemble *e, int *offsets, int cap);
int  lee_ensemble_step(lee_ensemble *e, int nsteps, double *rec_out, int use_graph);
double lee_ensemble_last_ms(const lee_ensemble *e);
const char *lee_ensemble_error(const lee_ensemble *e);
/* Launch form of the members' air and porous stages (no counterpart in the reference):
 * 0 = autotune at the next step (the default: each form timed on the coming step's first
 * stage, the fastest kept), 1 = a block per 256-column tile, 2 = the same with each member
 * plane's tiles padded to a multiple of 8 (a tile and its vertical neighbours on one XCD),
 * 3 = persistent (the resident grid walks the tiles).  All forms give identical bits.
 * The autotuner picks the air stage's and the porous stage's forms separately; a pinned form
 * applies to both.  Drops the captured graphs.  lee_ensemble_launch returns the form in use
 * (1..3) of the air stage (porous = 0) or of the porous stage (porous = 1), 0 before the
 * first step. */
int  lee_ensemble_set_launch(lee_ensemble *e, int form);
int  lee_ensemble_launch(const lee_ensemble *e, int porous);

/* Learned ground model (no counterpart in the reference): an MLP 4 -> 64 -> 64 -> 64 -> 8 with
 * GELU, from normalised ground parameters (flow resistivity, porosity, tortuosity, layer
 * depth) to the residues of the ground impedance's eight poles (lee_options.ground_*; the
 * Python driver trains it on the square roots of the residues).
 * Trained and evaluated on the GPU: BF16 MFMA products with FP32 accumulation, FP32 master
 * weights, mean-squared-error loss, Adam.  Parameters are one flat FP32 array, layer by
 * layer W[out][in] then b[out] over padded shapes 16->64, 64->64, 64->64, 64->16 (padding
 * stays zero).  Rows of x have 4 values, rows of y / t have 8. */
typedef struct lee_mlp lee_mlp;
int  lee_mlp_create(int device, unsigned long long seed, lee_mlp **out);
int  lee_mlp_destroy(lee_mlp *n);
const char *lee_mlp_error(const lee_mlp *n);
long long lee_mlp_param_count(const lee_mlp *n);
int  lee_mlp_get_params(lee_mlp *n, float *w, long long cap);
int  lee_mlp_set_params(lee_mlp *n, const float *w, long long count);
int  lee_mlp_predict(lee_mlp *n, const float *x, int rows, float *y);
/* Gradient of the mean squared error over rows (into grad, nparam floats, may be NULL) and
 * the error itself; the weights are left unchanged. */
int  lee_mlp_gradient(lee_mlp *n, const float *x, const float *t, int rows, float *grad, double *mse);
/* epochs of shuffled mini-batches (the last partial batch of an epoch is dropped) with Adam
 * (beta 0.9 / 0.999, eps 1e-8) at learning rate lr. */
int  lee_mlp_train(lee_mlp *n, const float *x, const float *t, int rows, int batch, int epochs, float lr,
                   unsigned long long seed, double *final_mse);

/* Output writers of the drop-in file contract.
 * ./solution/pty%04d.dat: one %.6e sample per line (airpore.cpp:468-531, 577).
 * ./solution/p_f2_%04d.dat: zero-padded radix-2 FFT, "f |X|" at setprecision(10)
 * (airpore::get_FFT_y1 airpore.cpp:966-1011, FFT.cpp:15-121). */
int  lee_write_receivers(const char *dir, const double *rec, int nsteps, int nrec);
int  lee_write_spectra(const char *dir, const double *rec, int nsteps, int nrec,
                       int fft_m, double dt);
/* In-place radix-2 complex FFT of 2^m points, dir = 1 forward (scaled by 1/N),
 * -1 inverse.  Replaces FFT() FFT.cpp:15-82. */
int  lee_fft(int dir, int m, double *re, double *im);

#ifdef __cplusplus
}
#endif

#endif /* LEE_ABI_H */
