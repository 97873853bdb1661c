/* oracle.h -- FP64 CPU restatement of the SDPD step, used ONLY as the checker
 * of libgpucell_hip.so (tests, the engine smoke check, and the CPU baseline
 * leg of the benchmark).  Nothing in the product links or calls it.
 *
 * The reference's kernel sources (ParticleGPU.cu, ConfigDomain.cu) are not
 * part of the reference tree (src/makefile:7,14-18 name them; they are
 * absent), so this file restates the published methods the reference's
 * parameter names identify, function by function against the reference's
 * entry points (src/ParticleGPU.cuh:8-37, src/ConfigDomain.cuh:9-18) and
 * call order (src/Particle.cu:1310-1391):
 *   - Lucy kernel; number density; linear EOS P = c^2 rho
 *   - Espanol-Revenga SDPD pair force with a Philox-4x32-10 pair noise
 *   - Groot-Warren modified velocity Verlet (lambda = VVsigma)
 *   - Fedosov WLC-POW springs, dihedral bending, area and volume penalties
 *   - Morse inter-cell aggregation
 *   - counting-sort cell list keyed by (cube, id)
 * State arrays are in id order, FP64.  The cell-key functions take FP32
 * positions and use FP32 arithmetic so their output is bit-comparable.
 *
 * Pinning of each function (the checks in tests/test_oracle_*.py):
 *   philox           Random123 known-answer vectors (tests/golden)
 *   lucy W, F        closed forms: W(0), unit integral, F = -W'/r by
 *                    central differences
 *   cell list        brute-force O(N^2) pair sets; std sort of (key, id)
 *   density          lattice sums against direct summation
 *   SDPD force       two-particle closed form; sum F = 0; zero on a
 *                    lattice at rest; dissipative power <= 0;
 *                    noise covariance (fluctuation-dissipation)
 *   integrator       constant-force parabola
 *   periodic wrap    in-box, unwrapped continuity
 *   geometry         exact area/volume of an icosahedron and a cube
 *   membrane force   -grad E by FP64 central differences; sum F = 0,
 *                    sum r x F = 0
 *   Morse            -dU/dr by central differences; F(r0) = 0
 *   Bound1 repulsion closed form (parity unpinned against the reference:
 *   Bound2 velocity  the reference's law is not in its tree)
 */
#ifndef GPUCELL_ORACLE_H
#define GPUCELL_ORACLE_H

#include <stdint.h>
#include "../include/gpucell/types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- primitives ------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Three standard normals of the pair {ida, idb} at noise counter `counter`. */
void orc_pair_normals(uint64_t seed, uint64_t counter, uint32_t ida, uint32_t idb,
                      double n[3]);
double orc_lucy_w(double r, double h);
double orc_lucy_f(double r, double h); /* -W'(r)/r */

/* ---- cell list (FP32 keys, bit-comparable) --------------------------- */
int orc_config_domain(const float bbox[6], float cut, const int periodic[3],
                      int ncube[3], float outer[3], float invwid[3], int* sizecube);
void orc_preconlist(int n, float cut, const int periodic[3], const float* coor,
                    const float bbox[6], const float invwid[3], const int ncube[3],
                    int* cc, int* cubepart, const int* sortpart, int* oldpos);
void orc_sort(int n, int* keys, int* values);
void orc_begin_end(int n, int sizecube, const int* cubepart, int* beginend);

/* ---- system ---------------------------------------------------------- */
typedef struct orc_sys orc_sys;

/* Membrane arrays follow the reference's loader conventions: tri and edge
 * ids 1-based (Particle.cu:252,268-284), beginendtri inclusive. */
orc_sys* orc_create(const Structconst* c, const gc_ext_consts* ext,
                    const float bbox[6], int ncell, int ntri, int nedge,
                    const int* tri, const int* edge, const int* celltab,
                    const int* beginendtri, const float* shearmod,
                    const float* bendmod, const float* hlarea,
                    const float* hgarea, const float* hvol);
void orc_destroy(orc_sys* s);
void orc_set_threads(int nthreads);

/* Each function below works on id-ordered FP64 arrays (3 doubles per vec). */
void orc_update_coor(const orc_sys* s, double* pos, const double* vel,
                     const double* force);
void orc_periodic_bc(const orc_sys* s, double* pos, int* img);
void orc_density(orc_sys* s, const double* pos, double* rho);
void orc_predict(const orc_sys* s, double* vel, double* velhalf,
                 const double* force);
void orc_bound2_vel(orc_sys* s, const double* pos, double* vel);
void orc_eos(const orc_sys* s, const double* rho, double* press);
/* overwrite force; abs_sum (may be NULL) gets sum_j |f_ij| per particle */
void orc_sdpd_force(orc_sys* s, const double* pos, const double* vel,
                    const double* rho, const double* press, uint64_t counter,
                    double* force, double* abs_sum);
void orc_rep_force(orc_sys* s, const double* pos, double* force, double* abs_sum);
void orc_ext_force(const orc_sys* s, const double* pos, int driveon,
                   double* force, double* abs_sum);
void orc_correct(const orc_sys* s, double* vel, const double* velhalf,
                 const double* force);

/* membrane: stepcount == -1 stores the reference state inside s */
void orc_geo(orc_sys* s, long stepcount, const double* pos, const int* img,
             double* tri_nor, double* tri_area, double* tri_cen,
             double* cell_area, double* cell_vol, double* par_area);
/* per-edge parameters from the current (initial) geometry; stored in s.
 * Any output pointer may be NULL.  */
void orc_mech_para(orc_sys* s, const double* pos, const int* img,
                   double* l0, double* lmax, double* kb, double* mexp,
                   double* p, double* kp, double* theta0);
/* membrane force per vertex (defforce, nCellPar*3) and added to force */
void orc_def_force(orc_sys* s, const double* pos, const int* img,
                   double* defforce, double* force, double* abs_sum);
double orc_membrane_energy(orc_sys* s, const double* pos, const int* img);
void orc_agg_force(orc_sys* s, const double* pos, double* aggforce,
                   double* force, double* abs_sum);
double orc_morse_energy(orc_sys* s, const double* pos);

/* One Particle::UpdatePar (Particle.cu:1310-1329) on the state. */
typedef struct {
    double* pos; int* img; double* vel; double* rho; double* velhalf;
    double* force; double* press;
    double* defforce; double* aggforce;   /* nCellPar*3 or NULL if no cells */
    double* cell_area; double* cell_vol;  /* nCell or NULL */
    double* abs_sum;                      /* N, may be NULL */
    long step; uint64_t counter;
} orc_state;
void orc_step(orc_sys* s, orc_state* st, int driveon);

#ifdef __cplusplus
}
#endif
#endif
