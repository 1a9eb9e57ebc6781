/*
 * alens_oracle.h -- FP64 CPU restatement of the aLENS/SimToolbox collision-constraint path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the timed CPU baseline.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only as the
 * checker.  The product (alens_amd/, include/alens_capi.h) never links, loads or calls it.
 *
 * The reference hot path lives in the SimToolbox submodule, which is absent from the reference tree
 * (.gitmodules:1-3, CMakeLists.txt:108-117 name the missing translation units).  What follows is a
 * restatement of the published method with every free choice stated; the call sites that fix the
 * conventions are cited on each function.
 *
 * Arithmetic contract: IEEE-754 binary64, no FMA contraction (build with -ffp-contract=off), every
 * expression evaluated in the order written.  sqrt and '/' are correctly rounded.  Transcendentals
 * (log, sin, cos) are the library's own polynomial routines below, so a GPU build that follows the
 * same operation order reproduces every bit.
 */
#ifndef ALENS_ORACLE_H
#define ALENS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double box_lo[3], box_hi[3];
    int32_t pbc[3];
    int32_t monolayer;      /* 1: motion restricted to the x-y plane */
    double col_buf;         /* absolute gap threshold (um) for recording a collision constraint */
    double viscosity;       /* pN s / um^2 */
    double kT;              /* pN um; <= 0 disables Brownian velocity */
    double dt;              /* s */
    double tol;             /* BBPGD residual tolerance (inf-norm of the projected gradient step) */
    double diam_col_ratio;  /* collision diameter = ratio * diameter */
    double len_col_ratio;   /* collision length   = ratio * length   */
    double cell_size;       /* cell-list cell edge target (um); <= 0 selects the automatic rule */
    int32_t max_ite;
    int32_t pad0;
    uint64_t seed;
    int64_t step;           /* step counter keying the Brownian stream */
} orc_cfg;

typedef struct {
    int32_t iters;          /* BBPGD loop iterations (0 if the start point already satisfied tol) */
    int32_t matvecs;        /* applications of A = dt D^T M D (iters + 2 when iters > 0) */
    int32_t converged;
    int32_t fallback_used;  /* warm start rejected, solve restarted cold */
    int32_t status;         /* 0 ok, 4 non-finite, 5 not converged */
    int32_t pad0;
    double res0;            /* residual at the start point actually used */
    double res;             /* final residual */
} orc_report;

/* ---- numerics shared by every stage ---- */
void   orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_log(double x);
void   orc_sincos2pi(double u, double *s, double *c);
double orc_sqrt(double x);
/* out = {s, u, dist, D0, D1, D2} with D = P(s) - Q(u) */
void   orc_segseg(const double cI[3], const double tI[3], double hI,
                  const double cJ[3], const double tJ[3], double hJ, double out[6]);
double orc_block_sum(const double *v, int64_t n);
double orc_chunked_sum(const double *v, int64_t n);

/* ---- per-step stages ---- */
int64_t orc_pairs_brute(int32_t n, const double *center, const double *dir, const double *length,
                        const double *radius, const orc_cfg *cfg, int64_t cap,
                        int32_t *I, int32_t *J, int32_t *img);
int64_t orc_pairs_cells(int32_t n, const double *center, const double *dir, const double *length,
                        const double *radius, const orc_cfg *cfg, int64_t cap,
                        int32_t *I, int32_t *J, int32_t *img);
void orc_mobility(int32_t n, const double *dir, const double *length, const double *radius,
                  const uint8_t *immovable, const orc_cfg *cfg, double *mob);
void orc_unc(int32_t n, const int32_t *gid, const double *mob, const double *f_ext,
             const orc_cfg *cfg, double *unc);
void orc_rows(int32_t n, const double *center, const double *dir, const double *length,
              const double *radius, const orc_cfg *cfg, int64_t m, const int32_t *I,
              const int32_t *J, const int32_t *img, const double *unc, double *delta0,
              double *nrm, double *posI, double *posJ, double *rI, double *rJ, double *b);
void orc_apply_D_M(int32_t n, int64_t m, const int32_t *I, const int32_t *J, const double *nrm,
                   const double *rI, const double *rJ, const double *mob, const orc_cfg *cfg,
                   const double *gamma, double *F, double *U);
void orc_apply_DT(int64_t m, const int32_t *I, const int32_t *J, const double *nrm,
                  const double *rI, const double *rJ, const double *U, double *du);
int  orc_bbpgd(int32_t n, int64_t m, const int32_t *I, const int32_t *J, const double *nrm,
               const double *rI, const double *rJ, const double *b, const double *lb,
               const double *invk, const double *mob, const orc_cfg *cfg, const double *x0,
               double *gamma, orc_report *rep);
void orc_rows_ext(int64_t m, const int32_t *I, const int32_t *J, const double *delta0,
                  const double *nrm, const double *posI, const double *posJ, const double *unc,
                  const orc_cfg *cfg, double *rI, double *rJ, double *b);
void orc_stress_sum(int64_t m, const double *nrm, const double *labI, const double *labJ,
                    const double *gamma, const double *stress, const uint8_t *use_given, double out[9]);
void orc_writeback(int32_t n, int64_t m, const int32_t *I, const int32_t *J, const double *nrm,
                   const double *rI, const double *rJ, const double *gamma, const double *mob,
                   const double *unc, const orc_cfg *cfg, double *fcon, double *u);
void orc_euler(int32_t n, double *center, double *dir, const double *u, const orc_cfg *cfg);
void orc_set_threads(int32_t nthreads);
int32_t orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
