/*
 * scn_oracle.h -- CPU restatement of SCONE's multigroup tracking loop (TEST INFRASTRUCTURE).
 *
 * This library is the checker for the HIP engine, nothing else: only the test suite, the smoke
 * check and the CPU-baseline leg of bench.py load it.  It follows the reference history by
 * history, in the structure of the Fortran (coordList levels, distance cache, tallyAdmin report
 * order, particleDungeon), citing each routine it restates.  Arithmetic that is not IEEE basic
 * comes from include/scn_fpmath.h, which the engine shares, so a history computed here and on
 * the GPU has identical bits.
 *
 * Parity pins (see tests/): RNG against Python integers and the reference's pow_of_gsq table;
 * geometry against the reference's own unit tests (latUniverse_test, pinUniverse_test,
 * box_test, squareCylinder_test, cylinder_test, rootUniverse_test, geometryStd_iTest);
 * MG data against baseMgNeutronDatabase_iTest; tallies against scoreMemory_test and
 * keff*Clerk_test numbers; sampling and the whole loop against brute force and closed forms
 * (infinite-medium k_inf and group fluxes).
 */
#ifndef SCN_ORACLE_H
#define SCN_ORACLE_H

#include "scone_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* per-history record for trace parity */
typedef struct orc_hist {
    int64_t n_events;      /* transport segments / tentative flights + collisions */
    int32_t n_coll;        /* real collisions */
    int32_t n_sites;
    uint64_t rng_end;      /* RNG state when the history ended */
    double w_end;
    int32_t fate;          /* 5001 absorbed, 5002 leaked */
    int32_t G_end;
    double r_end[3];
} orc_hist;

/* coordinate placement result (geometryStd%whatIsAt + the level stack) */
typedef struct orc_place_out {
    int32_t matIdx;
    int32_t uniqueID;
    int32_t nesting;
    int32_t localID[SCN_MAX_NEST];
    int32_t uniIdx[SCN_MAX_NEST];
    double r[SCN_MAX_NEST][3];
} orc_place_out;

const char* orc_last_error(void);
int orc_error_code(void);

/* math / RNG (the shared header, exported for tests) */
double orc_rng_get(uint64_t* s);
uint64_t orc_rng_skip(uint64_t s, int64_t k);
double orc_log(double x);
void orc_sincos(double x, double* s, double* c);
double orc_norm2(const double v[3]);
void orc_rotate(const double dir[3], double mu, double phi, double out[3]);
double orc_dsum(const double* x, int64_t n);

/* geometry primitives */
double orc_box_distance(const scn_box* b, const double r[3], const double u[3]);
double orc_box_evaluate(const scn_box* b, const double r[3]);
int orc_box_halfspace(const scn_box* b, const double r[3], const double u[3]);
double orc_surf_evaluate(const scn_problem* P, const scn_surface* S, const double r[3]);
double orc_surf_distance(const scn_problem* P, const scn_surface* S, const double r[3], const double u[3]);
int orc_surf_halfspace(const scn_problem* P, const scn_surface* S, const double r[3], const double u[3]);
void orc_box_explicit_bc(const scn_box* b, double r[3], double u[3]);
void orc_box_transform_bc(const scn_box* b, double r[3], double u[3]);
double orc_cyl_distance(double radius, const double r[3], const double u[3]);
int orc_uni_findcell(const scn_problem* P, int32_t uniIdx, const double r[3], const double u[3]);
double orc_uni_distance(const scn_problem* P, int32_t uniIdx, int32_t localID, const double r[3],
                        const double u[3], int32_t* surfIdx);
int orc_place(const scn_problem* P, const double r[3], const double u[3], orc_place_out* out);
/* geometryStd%move_withCache / teleport from a placed position */
int orc_move(const scn_problem* P, const double r[3], const double u[3], double max_dist,
             orc_place_out* out, double u_out[3], double* dist_out, int32_t* event);
int orc_teleport(const scn_problem* P, const double r[3], const double u[3], double dist,
                 orc_place_out* out, double u_out[3]);

/* transport */
int orc_eigen_cycle(const scn_problem* P, const scn_particle_state* bank_in, int32_t n_in,
                    int64_t rng_state, double k_eff, scn_particle_state* bank_out, int32_t cap,
                    int32_t* n_out, double* bin_scores, double* keff_scores, orc_hist* hist,
                    int32_t n_threads);
int orc_norm_size_repr(scn_particle_state* bank, int32_t* n, int32_t cap, int32_t tot_pop,
                       int64_t rng_state);
int orc_fission_source(const scn_problem* P, int32_t n, int64_t rng_state, const double bottom[3],
                       const double top[3], int32_t g_in, int32_t attempts, scn_particle_state* bank);
int orc_source_sample(const scn_problem* P, const scn_source* src, int32_t n, int64_t rng_state,
                      scn_particle_state* bank);
int orc_fixed_run(const scn_problem* P, const scn_source* src, int32_t pop, int32_t n_batches, int64_t seed,
                  int64_t stop_bin, double stop_rel, int32_t buffer_size, double* bins_result, orc_hist* hist,
                  int32_t* n_done, int32_t n_threads);
void orc_score_result(double csum, double csum2, int N, double out[2]);
void orc_close_cycle(double* bins, double* csum, double* csum2, int64_t n, double norm);
double orc_keff_implicit(const double s[4]);
/* single operations of the loop, exposed for known-answer tests */
int orc_clerk_bin(const scn_problem* P, int32_t clerk, const double r[3], int32_t mat, int32_t cell);
void orc_score_collision(const scn_problem* P, const double r[3], int32_t mat, int32_t cell, int32_t G, double w,
                         double track_xs, int32_t is_virtual, double* bins, double keff[4]);
int orc_scatter_gout(const scn_problem* P, int32_t mat, int32_t G_in, double u, double* w_mul);
int orc_chi_gout(const scn_problem* P, int32_t mat, double u);
void orc_keff_analog(double start_w, double end_w, double k_norm, double acc[2]);
int orc_legendre_p1(double c, uint64_t* rng, double* mu);
int orc_eigen_run(const scn_problem* P, const scn_eigen_settings* S,
                  const scn_particle_state* initial_bank, scn_eigen_result* res, int32_t n_threads);

#ifdef __cplusplus
}
#endif
#endif
