/*
 * tmpg.h -- C ABI of the MI355X per-step PIC library (libtmpg.so).
 *
 * Replaces the reference's argument-less Fortran module procedures on the
 * per-step hot path (called from code/tristanmainloop.F90:117-242) by plain
 * C entry points over an opaque context.  The Fortran host binds them with
 * bind(C) (tristan-mp-pu-master_densdecomp_amd/fortran/m_tmpg.F90).
 *
 * Conventions
 *  - Every entry point returns 0 on success or a negative TMPG_E* code; it
 *    never aborts.  tmpg_last_error() gives the message.
 *  - Host arrays are borrowed only for the duration of the call.
 *  - Grid arrays at the boundary use the reference layout: Fortran
 *    column-major (mx,my,mz) including ghost cells, mx = nx + nghost,
 *    nghost = 5 (order 1) or 7 (orders 2, 3); in 2D mz = 1
 *    (code/fields.F90:166-184, 221-229).  Inside the library the grid is the
 *    periodic interior only; ghost layers returned by the getters hold the
 *    periodic images the reference's bc_b1/bc_e1 would put there.
 *  - Particle coordinates at the boundary are the reference's absolute
 *    1-based coordinates (interior x in [nghost/2+1, nghost/2+1+nx)).
 *    Inside, a particle is an integer cell plus an FP32 offset in [0,1).
 *  - One HIP stream per context; calls are stream-ordered.  Getters
 *    synchronise.  One host thread per context.
 */
#ifndef TMPG_H
#define TMPG_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TMPG_OK            0
#define TMPG_EINVAL       -1   /* bad argument */
#define TMPG_ENOMEM       -2   /* device allocation failed */
#define TMPG_EHIP         -3   /* HIP runtime error */
#define TMPG_EOVERFLOW    -4   /* particle capacity exceeded */
#define TMPG_ESTATE       -5   /* call not valid in the current state */

/* reference quirks reproduced by default (compat bits) */
#define TMPG_Q_GATHER2   1  /* 2nd-order gather: primal weights over the dual range
                               (code/particles_movedeposit.F90:709-832) */
#define TMPG_Q_BXY_Z     2  /* 3D primal Bx/By not averaged in z
                               (code/particles_movedeposit.F90:988-989) */
#define TMPG_COMPAT_DEFAULT (TMPG_Q_GATHER2 | TMPG_Q_BXY_Z)

#define TMPG_SCHEME_ESIRKEPOV 0
#define TMPG_SCHEME_ZIGZAG    1

typedef struct tmpg_config {
    int32_t dim;          /* 2 or 3 (reference: -DtwoD at compile time) */
    int32_t order;        /* shape order 1, 2, 3 (reference: -Ddd1/-Ddd2/-Ddd3) */
    int32_t nx, ny, nz;   /* interior cells (<grid> mx0, my0, mz0); nz = 1 in 2D */
    int32_t ntimes;       /* filter passes (<algorithm> ntimes) */
    int32_t compat;       /* TMPG_Q_* bits */
    int32_t track_ids;    /* nonzero: carry a per-particle int32 id (like p%ind) */
    float   c;            /* Courant number (<time> c) */
    float   corr;         /* speed-of-light correction (<algorithm> Corr) */
    float   qm[2];        /* q/m of species 0 (ions) and 1 (electrons): qmi, qme */
    float   q[2];         /* charge per macro-particle: qi, qe */
    float   mass[2];      /* mi, me after normalisation (energy diagnostics) */
    int64_t capacity[2];  /* particle capacity per species (0: sized at first load) */
    int32_t tile[3];      /* cell tile for the sorted push (0: default, 8x8x4 in
                             3D; 16x16 in 2D, 16x8 on grids too small for two
                             16x16 tiles per CU) */
    int32_t sort_every;   /* tmpg_step sorts each tile's particles by cell every
                             this many steps (0: every 4, negative: never) */
    int32_t highorder;    /* nonzero: 4th-order field solver (<algorithm>
                             highorder; advance_b_halfstep_42 and
                             advance_e_fullstep_42, code/fields.F90:1039-1361) */
    int32_t pusher;       /* 0: Boris; 1: Vay (the reference's -Dvay build,
                             code/particles_movedeposit.F90:1199-1223) */
    int32_t scheme;       /* 0: Esirkepov current of the shape `order` (the
                             reference's -Ddd<n> builds); 1: the original
                             scheme of the -Dzzag build: trilinear mover
                             (code/particles_movedeposit.F90:98-347) and the
                             zigzag deposit (code/particles.F90:550-669),
                             nghost = 5; needs order = 1 */
    int32_t reserved[1];
} tmpg_config;

typedef struct tmpg_ctx tmpg_ctx;

/* life cycle --------------------------------------------------------- */
/* replaces initialize()'s allocate_fields/allocate_particles
 * (code/fields.F90:221-377, code/particles.F90:286-351) */
int tmpg_create(const tmpg_config* cfg, tmpg_ctx** out);
int tmpg_destroy(tmpg_ctx* ctx);
/* message of the last failing call (ctx may be NULL for tmpg_create) */
const char* tmpg_last_error(const tmpg_ctx* ctx);
int tmpg_sync(tmpg_ctx* ctx);

/* state transfer ----------------------------------------------------- */
/* ex..bz module arrays (code/fields.F90:81-86, 346-351) */
int tmpg_set_fields(tmpg_ctx* ctx, const float* ex, const float* ey, const float* ez,
                    const float* bx, const float* by, const float* bz);
int tmpg_get_fields(tmpg_ctx* ctx, float* ex, float* ey, float* ez,
                    float* bx, float* by, float* bz);
/* curx..curz (code/fields.F90:346) */
int tmpg_set_currents(tmpg_ctx* ctx, const float* jx, const float* jy, const float* jz);
int tmpg_get_currents(tmpg_ctx* ctx, float* jx, float* jy, float* jz);
/* p(:) of species 0 (ions, p(1:ions)) or 1 (electrons, p(maxhlf+1:)):
 * x,y,z,u,v,w,ch (code/particles.F90:51-55).  ch may be NULL (all 1);
 * id may be NULL.  set replaces the species' particles. */
int tmpg_set_particles(tmpg_ctx* ctx, int species, int64_t n,
                       const float* x, const float* y, const float* z,
                       const float* u, const float* v, const float* w,
                       const float* ch, const int32_t* id);
/* the same arguments, appended to the species' particles (the reference's
 * injections add particles, code/particles.F90:2549-2938); ids not given
 * continue the count */
int tmpg_add_particles(tmpg_ctx* ctx, int species, int64_t n,
                       const float* x, const float* y, const float* z,
                       const float* u, const float* v, const float* w,
                       const float* ch, const int32_t* id);
int tmpg_count_particles(tmpg_ctx* ctx, int species, int64_t* n);
/* [reorder_particles] code/particles.F90:394-497 (a counting sort of each
 * species by cell every 10 laps, called from code/tristanmainloop.F90:303):
 * sorts each cell tile's particles by cell now; tmpg_step does this every
 * cfg.sort_every steps.  The order sets how many LDS bank conflicts the
 * push's field gather meets.  TMPG_ESTATE between tmpg_move and
 * tmpg_deposit. */
int tmpg_sort_particles(tmpg_ctx* ctx);
/* writes min(n_capacity, count) particles in the library's order */
int tmpg_get_particles(tmpg_ctx* ctx, int species, int64_t n_capacity, int64_t* n,
                       float* x, float* y, float* z, float* u, float* v, float* w,
                       float* ch, int32_t* id);

/* per-step kernels (reference names in brackets) -------------------- */
int tmpg_advance_b_half(tmpg_ctx* ctx);   /* [advance_Bhalfstep] code/fields.F90:586-728 */
int tmpg_advance_e(tmpg_ctx* ctx);        /* [advance_Efield]    code/fields.F90:739-870 */
int tmpg_reset_currents(tmpg_ctx* ctx);   /* [reset_currents]    code/fields.F90:499-506 */
/* [move_particles] code/particles_movedeposit.F90:63-88: gather + Boris +
 * position advance of both species, no deposit */
int tmpg_move(tmpg_ctx* ctx);
/* [deposit_particles] code/particles_movedeposit.F90:1281-2051: Esirkepov
 * current of the last tmpg_move (old positions are the ones before it),
 * added to J, with periodic wrap of positions */
int tmpg_deposit(tmpg_ctx* ctx);
/* fused move + deposit (one pass over the particles): the hot path */
int tmpg_push_deposit(tmpg_ctx* ctx);
/* [exchange_current] code/fieldboundaries.F90:1768-2189.  The library
 * deposits with periodic wrap, so the fold is already done: a no-op kept
 * for call-sequence parity */
int tmpg_fold_currents(tmpg_ctx* ctx);
/* [apply_filter1_opt] code/filter.F90:8-221: npasses binomial passes */
int tmpg_filter(tmpg_ctx* ctx, int npasses);
int tmpg_add_current(tmpg_ctx* ctx);      /* [add_current] code/fields.F90:1372-1395 */
/* nsteps of the main loop (code/tristanmainloop.F90:107-274, periodic box):
 * B half, push+deposit, B half, E, filter (or denoiser), add current */
int tmpg_step(tmpg_ctx* ctx, int nsteps);

/* diagnostics (FP64 on the device) ------------------------------------ */
/* out[0] = 0.5*sum(E^2), out[1] = 0.5*sum(B^2) over the interior,
 * out[2], out[3] = kinetic energy sum(ch*m*c^2*(gamma-1)) of species 0, 1,
 * out[4] = particles of species 0, out[5] = of species 1,
 * out[6] = max |rho - rho_prev - div cur| of the last deposit (continuity,
 *          needs tmpg_track_continuity(ctx,1) before the step), else -1,
 * out[7] = L2 norm of rho used for that check */
int tmpg_diag(tmpg_ctx* ctx, double out[8]);
/* Gauss's law (no reference counterpart: the reference computes no
 * divergence diagnostic).  rho = the charge density of tmpg_charge_density,
 * F^n rho = rho after the n binomial passes the step applies to the
 * deposited current (cfg.ntimes, or the denoiser's base passes), div_b E =
 * the backward-difference divergence at the nodes of the staggered E:
 *   out[0] = max |div_b E - F^n rho| over the interior nodes,
 *   out[1] = rms of F^n rho, out[2] = max |F^n rho|, out[3] = max |div_b E|.
 * The step adds F^n cur to E with div_b(cur) = rho_new - rho_old (Esirkepov)
 * and an E update whose curl has no divergence (highorder = 0), so out[0]
 * keeps its initial value up to round-off; a neutral start (E = 0, the
 * species at the same places) keeps it near zero. */
int tmpg_gauss_residual(tmpg_ctx* ctx, double out[4]);
/* tmpg_gauss_residual without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * kernels work in the library's static per-GPU scratch (2^22 cells of two
 * FP64 grids, 64 MiB, held by each GPU the library runs on until the process
 * ends) and clear it with kernels, not memsets.  Grids of more than 2^22
 * cells: TMPG_EINVAL.  Calls on one GPU share that scratch: calls that
 * overlap in time on one GPU (several host threads, several streams) race. */
int tmpg_gauss_residual_noalloc(tmpg_ctx* ctx, double out[4]);
/* charge density at the nodes (interior nx*ny*nz, x fastest), FP64,
 * deposited with the order's own shape from the current positions */
int tmpg_charge_density(tmpg_ctx* ctx, double* rho);
int tmpg_track_continuity(tmpg_ctx* ctx, int on);

/* [inject_plasma_region] code/particles.F90:2549-2938 on the host, for one
 * periodic domain, with the reference's own random stream: maxwell_dist
 * (:2176-2273) on the tables of init_maxw_table (:2126-2163) and the
 * Park-Miller random() (code/aux.F90:82-99), Poisson counts below 10
 * expected particles (:109-134).  Fills `n` ions and n*upsamp_e electrons
 * into ions[7*cap] and lecs[7*cap] (rows x, y, z, u, v, w, ch, each of
 * stride cap; absolute coordinates as at tmpg_set_particles) and advances
 * *dseed; a region of more than cap particles returns TMPG_EOVERFLOW with
 * the count in *n and *dseed unchanged.  Seeded like the reference
 * (123457 + rank, code/communications.F90:224-231) the positions are the
 * reference's FP32 values; momenta agree to the math library's last bits.
 * Host only: no context, no GPU. */
typedef struct tmpg_plasma_region {
    int32_t dim, order;          /* ghost layout: nghost 5 (order 1) or 7 */
    int32_t nx, ny, nz;          /* interior cells (<grid> mx0, my0, mz0) */
    float   c;                   /* Courant number (the cd of maxwell_dist) */
    float   x1, x2, y1, y2, z1, z2;  /* region, the reference's global coordinates */
    float   ppc;                 /* particles per cell, both species together */
    float   gamma_drift;         /* drift gamma with its sign; |.| < 1 read as v/c */
    float   delgam_i, delgam_e;  /* kT / m c^2 of ions and of electrons */
    float   weight;              /* ch of the ions; electrons weight/upsamp_e */
    int32_t direction;           /* drift along 1: x, 2: y, 3: z */
    int32_t pcosthmult;          /* 2D: 1 three-dimensional spread, 0 in the plane */
    float   sigma;               /* <particles> sigma: nonzero draws the 2D polar angle again */
    int32_t upsamp_e;            /* electrons per ion (0 or 1: one) */
} tmpg_plasma_region;
int tmpg_inject_plasma_region(const tmpg_plasma_region* r, double* dseed, int64_t cap, int64_t* n,
                              float* ions, float* lecs);

/* field output file flds.tot.NNN ------------------------------------- */
/* [output_tot] code/output.F90:2494-3875 for one periodic domain: one HDF5
 * file of FP32 datasets over the padded nodes i = istep, 2 istep, ... <= mx
 * (and j, k; k = 1 in 2D), dimensions (mz/istep, my/istep, mx/istep) in C
 * order, in the order of :2604-2628:
 *   dens densi v3x v3xi v3y v3yi [v3z v3zi] ex ey ez bx by bz jx jy jz
 * (v3z, v3zi omitted in 2D when in_plane != 0, the reference's list for
 * sigma = 0 and pcosthmult = 0).  dens/densi and the mean 3-velocities
 * v3* follow meanq_fld_cur (:5229-5482: boxes of +-2 cells, +-0 in z in 2D,
 * folded like exchange_current); E and B are interpolfld (:4995-5127) at
 * the nodes; J is the current array.  The HDF5 C library is loaded at the
 * first call (TMPG_HDF5_LIB, else libhdf5.so from the loader's path, else
 * /opt/conda/lib/libhdf5.so); without it: TMPG_EINVAL. */
typedef struct tmpg_flds_tot_state {
    int32_t dim, order;          /* ghost layout as in tmpg_config */
    int32_t nx, ny, nz;          /* interior cells */
    int32_t istep;               /* <output> istep */
    int32_t in_plane;            /* 2D: the 15-dataset list */
    const float *ex, *ey, *ez, *bx, *by, *bz, *jx, *jy, *jz;   /* padded (mx,my,mz) */
    int64_t ni, ne;              /* particles of each species */
    const float* ions;           /* rows x, y, z, u, v, w, ch of stride ni */
    const float* lecs;           /* the same, stride ne */
} tmpg_flds_tot_state;
/* host only: writes the state given (errbuf, if not NULL, receives the
 * message of a failure) */
int tmpg_flds_tot_write(const char* path, const tmpg_flds_tot_state* st, char* errbuf, int errlen);
/* the context's fields, currents and particles (copied to the host) */
int tmpg_write_flds_tot(tmpg_ctx* ctx, const char* path, int istep, int in_plane);

/* plasma loader on the device (next to code/particles.F90:2549-2938) ---- */
/* appends ppc particles per cell of one species, uniform positions,
 * drift gamma*beta along x (sign), isotropic thermal spread with
 * temperature theta = kT/mc^2 (Maxwell-Juttner, boosted), counter-based RNG
 * keyed by seed. */
int tmpg_load_uniform(tmpg_ctx* ctx, int species, int ppc, float drift_gb,
                      float theta, uint64_t seed);

/* timing (HIP events on the context's stream): milliseconds summed over the
 * tmpg_step iterations since the previous call, per phase:
 * [0] field stencils (B half steps, E step + add current, J reset),
 * [1] push+deposit (including its field staging and particle re-sort),
 * [2] filter or denoiser, [3] number of steps summed, [4] total,
 * [5] the tiled push+deposit kernel alone (part of [1]) */
int tmpg_last_timings(tmpg_ctx* ctx, double ms[6]);

/* per-phase shader-cycle sums of the 3D cell-sum push kernel, summed over
 * its workgroups' tiles since the last reset (no reference counterpart: a
 * profiling aid).  Phases: [0] field staging, [1] push, emit and deposit
 * records, [2] y/z crosser deposits and scans, [3] record index scatter,
 * [4] per-cell current sums, [5] record reset, [6] flush to the global
 * currents; counts[] holds the number of stamps per phase.  Returns
 * TMPG_ESTATE unless the library was built with phase stamps (the package
 * build's --phase-stamps option, which writes lib/libtmpg_phases.so). */
int tmpg_phase_cycles(unsigned long long cycles[8], unsigned long long counts[8], int reset);

/* learned current denoiser (no reference counterpart: takes the place of
 * the apply_filter1_opt passes in low particles-per-cell runs) ------------
 * A periodic CNN (2D: 4 layers 3x3, width 32; 3D: 3 layers 3x3x3, width 16,
 * GELU) maps J/s to a vector potential A on the B grid; the correction is
 * J += s*curl_b(A) with the E update's backward-difference curl, so it adds
 * no divergence.  With base_passes = n > 0 the model acts on the current
 * after n binomial passes (J <- F^n J, then the correction); with n = 0 it
 * acts on the raw deposit and replaces the filter entirely.  BF16 MFMA
 * convolutions with FP32 accumulation; training (Adam) on the device.
 * Current arrays at this interface are the periodic interior,
 * [batch][3][nx*ny*nz] floats, x fastest; training pairs are given as the
 * model's input (already filtered when base_passes > 0) and its target. */
typedef struct tmpg_model tmpg_model;
int tmpg_model_create(int dim, uint64_t seed, tmpg_model** out);
int tmpg_model_destroy(tmpg_model* m);
const char* tmpg_model_last_error(const tmpg_model* m);
int tmpg_model_num_params(const tmpg_model* m, int64_t* n);
/* flat order: per layer W[cout][tap][cin] (tap x fastest, 3x3(x3)) then b[cout] */
int tmpg_model_get_params(tmpg_model* m, float* out);
int tmpg_model_set_params(tmpg_model* m, const float* in);
int tmpg_model_set_scale(tmpg_model* m, float s);
int tmpg_model_get_scale(const tmpg_model* m, float* s);
int tmpg_model_set_base_passes(tmpg_model* m, int npasses);
/* training loss = mean(d^2) + lam * <d,t>^2 / (N <t,t>) per image, d = output
 * - target, t = target (both / s): lam > 0 also penalises the error along the
 * target (a shrunk or inflated amplitude); 0 (default) is plain MSE */
int tmpg_model_set_amplitude_weight(tmpg_model* m, float lam);
int tmpg_model_get_base_passes(const tmpg_model* m, int* npasses);
/* one forward/backward on a batch: loss = mean((J_lo/s + curl A - J_t/s)^2)
 * (+ the amplitude term);
 * grad_out (may be NULL) receives the flat gradient; update != 0 applies Adam */
int tmpg_model_train_batch(tmpg_model* m, int nx, int ny, int nz, int batch, const float* jlo,
                           const float* jt, float lr, int update, double* loss, float* grad_out);
int tmpg_model_eval(tmpg_model* m, int nx, int ny, int nz, int batch, const float* jlo,
                    const float* jt, double* loss, float* jout);
/* device-resident training set and Adam loop (cosine learning-rate decay) */
int tmpg_model_add_sample(tmpg_model* m, int nx, int ny, int nz, const float* jlo, const float* jt);
int tmpg_model_train(tmpg_model* m, int iters, float lr, uint64_t seed, double* losses);
/* apply to the context's current now (base passes, then the correction) /
 * use instead of the filter in tmpg_step (NULL restores the filter) */
int tmpg_denoise(tmpg_ctx* ctx, tmpg_model* m);
int tmpg_set_denoiser(tmpg_ctx* ctx, tmpg_model* m);

/* input file: the reference's <block> / key = value # comment format
 * (code/par.c) */
typedef struct tmpg_input tmpg_input;
int tmpg_input_open(const char* path, tmpg_input** out);
int tmpg_input_close(tmpg_input* in);
int tmpg_input_get_int(const tmpg_input* in, const char* block, const char* key, int def, int* val);
int tmpg_input_get_double(const tmpg_input* in, const char* block, const char* key, double def, double* val);

#ifdef __cplusplus
}
#endif
#endif
