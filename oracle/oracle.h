/*
 * oracle.h -- CPU restatement of the reference's per-step PIC loop.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the CPU baseline; the product library never links or calls it.
 *
 * Layout follows the reference exactly: fields are Fortran column-major
 * (mx,my,mz) arrays including ghost cells, 1-based in the formulas below
 * (linear index (i-1) + mx*((j-1) + my*(k-1))); particles are absolute
 * coordinates in that 1-based frame (code/particles.F90:51-55).
 *
 * Every function exists as _f32 (the reference's real(sprec)=4 arithmetic)
 * and _f64 (the same algorithm in double precision, the accuracy yardstick).
 */
#ifndef TMPG_ORACLE_H
#define TMPG_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* compat bits: which reference quirks to reproduce */
#define ORC_Q_GATHER2   1  /* 2nd-order gather sums primal weights over the dual range */
#define ORC_Q_BXY_Z     2  /* 3D primal Bx/By are not averaged in z */
#define ORC_Q_CARRY     4  /* Esirkepov prefix accumulators carried across rows/planes */
#define ORC_COMPAT_ALL  7

typedef struct {
    int dim;        /* 2 or 3 */
    int order;      /* 1, 2, 3 */
    int nx, ny, nz; /* interior cells (nz = 1 in 2D) */
    int ntimes;     /* filter passes */
    int compat;     /* ORC_Q_* bits */
    double c;       /* Courant number */
    double corr;    /* speed-of-light correction */
} orc_cfg;

/* padded extents and ghost widths (code/fields.F90:166-184, 221-229) */
void orc_extent(const orc_cfg* cfg, int* mx, int* my, int* mz, int* nghost, int* nghostz);

/* reference normalisation (code/particles.F90:196-235) evaluated in the given
 * precision: out = {qe, qi, me, mi, qme, qmi} */
void orc_normalize_f32(float c, float c_omp, float gamma0, float ppc0,
                       float me, float mi, float out[6]);
void orc_normalize_f64(double c, double c_omp, double gamma0, double ppc0,
                       double me, double mi, double out[6]);

/* fields: 6 arrays ex,ey,ez,bx,by,bz each mx*my*mz, contiguous (f6);
 * currents: 3 arrays curx,cury,curz (j3).  particles: 7 arrays of n
 * (x,y,z,u,v,w,ch) contiguous (prt). */
#define ORC_DECL(T, S) \
void orc_ghost_fill_##S(const orc_cfg*, T* arr3);                 /* bc_b1 / bc_e1 */ \
void orc_bhalf_##S(const orc_cfg*, T* f6);                        /* advance_b_halfstep */ \
void orc_efull_##S(const orc_cfg*, T* f6);                        /* advance_e_fullstep */ \
void orc_add_current_##S(const orc_cfg*, T* f6, const T* j3);     /* add_current */ \
void orc_fold_##S(const orc_cfg*, T* j3);                         /* exchange_current */ \
void orc_filter_##S(const orc_cfg*, T* j3, int ntimes);           /* apply_filter1_opt */ \
void orc_mover_##S(const orc_cfg*, const T* f6, T* prt, int64_t n, T qm); \
void orc_deposit_##S(const orc_cfg*, T* j3, T* prt, int64_t n, T qsp); \
void orc_step_##S(const orc_cfg*, T* f6, T* j3, T* prt_i, int64_t ni, T* prt_e, int64_t ne, \
                  T qmi, T qme, T qi, T qe, int nsteps); \
void orc_rho_##S(const orc_cfg*, const T* prt, int64_t n, T qsp, double* rho_interior);

ORC_DECL(float, f32)
ORC_DECL(double, f64)
#undef ORC_DECL

#ifdef __cplusplus
}
#endif
#endif
