/*
 * oracle.h -- CPU FP64 restatement of DelMaestroGroup/pimc's all-bead-pairs hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only the test suite
 * (tests/), __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load
 * it.  The product (pimc_amd/, libpimc_gpu.so) never links or calls it, and it is
 * never the thing that is measured as the GPU result.
 *
 * Every function restates the reference's own algorithm in the reference's loop
 * and accumulation order, single-threaded, in FP64, compiled with
 * -ffp-contract=off so that each expression is evaluated as written.  Each
 * declaration cites the reference file:line it follows.
 *
 * Parity pinning (what checks each function against something other than
 * itself) is listed per function in tests/test_oracle_*.py.
 */
#ifndef PIMC_ORACLE_H
#define PIMC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- MTRand (include/MersenneTwister.h:76-419) ----------------------------- */
void*    orc_mt_create(uint32_t seed);                  /* MTRand(seed)  :173-188, :206-211 */
void     orc_mt_destroy(void* mt);
uint32_t orc_mt_randint(void* mt);                      /* randInt()     :290-304 */
uint32_t orc_mt_randint_n(void* mt, uint32_t n);        /* randInt(n)    :306-323 */
double   orc_mt_rand(void* mt);                         /* rand()        :325-326 */
double   orc_mt_randexc(void* mt);                      /* randExc()     :331-332 */
double   orc_mt_randnorm(void* mt, double mean, double sd); /* randNorm  :349-363 */
void     orc_mt_save(void* mt, uint32_t* out625);       /* save()        :370-377 */
void     orc_mt_load(void* mt, const uint32_t* in625);  /* load()        :379-387 */

/* ---- Aziz potential (src/potential.cpp:1741-1909, include/potential.h:129-277,933-1018) */
void*  orc_aziz_create(int year, double max_sep, double rc, int build_derivs);
void   orc_aziz_destroy(void* az);
double orc_aziz_valueV(void* az, double r);             /* valueV      :1821-1840 */
double orc_aziz_valuedVdr(void* az, double r);          /* valuedVdr   :1847-1873 */
double orc_aziz_valued2Vdr2(void* az, double r);        /* valued2Vdr2 :1880-1909 */
double orc_aziz_V(void* az, const double* sep3);        /* V(dVec)     potential.h:985-989 */
void   orc_aziz_gradV(void* az, const double* sep3, double* out3); /* potential.h:997-1003 */
double orc_aziz_grad2V(void* az, const double* sep3);   /* potential.h:1011-1017 */
double orc_aziz_tailV(void* az);                        /* :1800-1806 */
double orc_aziz_dr(void* az);
int    orc_aziz_table_length(void* az);
const double* orc_aziz_table(void* az, int which);      /* 0=lookupV 1=dVdr 2=d2Vdr2 */
/* Bins where the valueV branches switch, replayed from the accumulated table
 * radius r_n (include/potential.h:176-183): first n with x_n >= EPS, >= 0.01, >= D. */
void   orc_aziz_branch_bins(void* az, int32_t* out3);
void   orc_aziz_params(void* az, double* out9);         /* eps rm A alpha beta D C6 C8 C10 */

/* ---- Simulation state: Prism box + Path + Worm + LookupTable + LocalAction ----- */
void* orc_state_create(double side, int M, int nalloc, double tau, double lambda,
                       double rc, int year);
void  orc_state_destroy(void* st);
/* All beads on, numBeadsAtSlice = nalloc, periodic links (src/path.cpp:28-129),
 * then LookupTable::updateGrid (src/lookuptable.cpp:314-349). beads: [M][nalloc][3]. */
void  orc_state_set_beads(void* st, const double* beads);
/* General worm state: bead_on [M][nalloc], num_beads [M], next/prev [M][nalloc][2],
 * worm head/tail/special1/special2 [2] each; rebuilds the grid. */
void  orc_state_set_topology(void* st, const uint8_t* bead_on, const int32_t* num_beads,
                             const int32_t* next, const int32_t* prev,
                             const int32_t* head, const int32_t* tail,
                             const int32_t* special1, const int32_t* special2);
void  orc_state_get_beads(void* st, double* beads);
void  orc_state_set_action(void* st, int shift, double vf0, double vf1); /* setShift, VFactor */
void* orc_state_aziz(void* st);

double orc_total_action(void* st);                      /* LocalAction::potentialAction() action.cpp:456-472 */
double orc_total_action_threads(void* st, int threads);  /* the same sum, slices over host threads */
void   orc_V_slice(void* st, int slice, double* out2, int32_t* sep_hist50); /* V(int) :902-947 */
double orc_Vnn_bead(void* st, int slice, int ptcl);     /* Vnn(bead) :1026-1060 */
double orc_Vnn_slice(void* st, int slice);              /* Vnn(int)  :1069-1123 */
double orc_action_bead(void* st, int slice, int ptcl);  /* potentialAction(bead) :487-514 */
double orc_action_range(void* st, const int32_t* start2, const int32_t* end2); /* :522-570 */
double orc_action_list(void* st, const int32_t* beads2, int n); /* potentialAction(vector) :572-620 */
void   orc_update_bead(void* st, int slice, int ptcl, const double* pos3); /* Path::updateBead path.cpp:197-200 */
/* GSF / Li-Broughton correction factors gradVFactor[2] and the estimator period
 * (src/setup.cpp:1232-1256); builds the dV/dr table when a factor is finite. */
void   orc_state_set_correction(void* st, double gvf0, double gvf1, int period);
double orc_gradV2_slice(void* st, int slice);             /* gradVSquared(slice) action.cpp:1188-1223 */
double orc_gradVnn2(void* st, int slice, int ptcl);       /* gradVnnSquared(bead) action.cpp:1282-1337 */
void   orc_bare_action_list(void* st, const int32_t* beads2, int n, double* out); /* barePotentialAction :647-689 */
double orc_action_correction(void* st, int slice, int ptcl); /* potentialActionCorrection(bead) :700-716 */
/* Virial pair terms per slice (action.cpp:1446-1786, :776-787), out [M][6] =
 * {rDOTgradUterm1, rDOTgradUterm2, deltadotgradUterm1, deltadotgradUterm2,
 *  virialKinCorrection, secondderivPotentialActionTau}; W = virial window. */
/* Path::leftPack (path.cpp:145-189) on raw arrays beads [M][N][3], next/prev [M][N][2], worm [M][N]. */
void   orc_left_pack(int M, int N, double* beads, int32_t* next, int32_t* prev, uint32_t* worm);
void   orc_virial_terms(void* st, int W, double* out);
/* VirialEnergyEstimator::accumulate (estimator.cpp:1086-1251): the 19 columns of one call. */
void   orc_virial_energy(void* st, int W, double mu, double* out19);
/* EnergyEstimator::accumulate (estimator.cpp:940-1029), primitive PIMC:
 * out = {K, V, V_ext, V_int, E, K/N, V/N, E/N} for one sample. */
void   orc_energy(void* st, double mu, double* out9);   /* K V V_ext V_int E E_mu K/N V/N E/N */
/* StaticStructureFactorEstimator::accumulate contribution sf/N (estimator.cpp:3705-3737) */
void   orc_ssf(void* st, const double* q3, int nq, double* out);
/* IntermediateScatteringFunctionEstimator::accumulate contribution isf/N, [nq][M] (:3923-3961) */
void   orc_isf(void* st, const double* q3, int nq, double* out);
/* S(q) and F(q,tau) on a subset of q-vectors only (timing baseline). */

/* ---- LookupTable dumps (src/lookuptable.cpp, include/lookuptable.h) ------------ */
void orc_grid_dims(void* st, int32_t* num_nn_grid3, int32_t* hash_cap); /* setupNNGrid :67-231 */
/* grid [M][nalloc][3], bead_label [M][nalloc], num_labels [cells][M],
 * hash [cells][M][cap] (cap as returned by orc_grid_dims at call time). */
void orc_grid_dump(void* st, int32_t* grid, int32_t* bead_label, int32_t* num_labels,
                   int32_t* hash);
void orc_grid_nn(void* st, int32_t* grid_nn);           /* gridNN [cells][27][3] after dedup */
int  orc_interaction_list(void* st, int slice, int ptcl, int32_t* list2, double* sep3); /* :464-516 */

/* ---- Wavevectors (src/estimator.cpp:439-570) ---------------------------------- */
/* type: "int" | "float" | "max_int" | "max_float"; input = whitespace separated.
 * Returns the number of vectors; writes at most max_out of them. */
int orc_qvectors(const double* side3, const char* type, const char* input,
                 double* out3, int max_out);

/* ---- Synthetic worldlines (the measurement plan's generator) ------------------- */
/* kind 0: slice 0 uniform (Prism::randPosition, container.cpp:155-160), slices
 * m>0 = randNorm(prev, sqrt(2*lambda*tau)) per axis then putInBC.
 * kind 1: closed Gaussian-bridge worldlines: slice 0 uniform, slices 1..M-1 by
 * the staging construction (move.cpp:270-291) towards slice 0's position. */
void orc_gen_worldlines(uint32_t seed, double side, int M, int N, double lambda, double tau,
                        int kind, double* beads);

/* ---- Diagonal moves (src/move.cpp), ΔV contract replay ------------------------- */
/* Each call performs one attemptMove with the reference's RNG consumption order and
 * returns 1 if accepted.  info (may be NULL) receives:
 *   [0] attempted flag, [1] last level's deltaAction, [2] number of levels / beads. */
void* orc_moves_create(void* st, int Mbar, int b, int var_updates, int max_wind,
                       double com_delta);
void  orc_moves_destroy(void* mv);
int   orc_move_bisection(void* mv, void* mt, double* info);   /* :1484-1661 */
int   orc_move_staging(void* mv, void* mt, double* info);     /* :1315-1408 */
int   orc_move_com(void* mv, void* mt, double* info);         /* :1124-1259 */

#ifdef __cplusplus
}
#endif
#endif
