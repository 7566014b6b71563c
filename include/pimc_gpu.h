/*
 * pimc_gpu.h -- C-ABI of the MI355X (gfx950) all-bead-pairs hot path.
 *
 * Drop-in boundary for DelMaestroGroup/pimc's PotentialBase / ActionBase /
 * EstimatorBase hot path.  Plain pointers and sizes only; every entry point
 * returns PIMC_OK (0) or a negative PIMC_E* status, never throws, and leaves a
 * message for pimc_last_error().  Host buffers belong to the caller; device
 * memory belongs to the context.  One context holds one worldline
 * configuration (one Markov chain) resident in HBM.  Calls on one context are
 * not reentrant (the reference's LocalAction is not either,
 * include/action.h:138-141); distinct contexts may be used from distinct
 * threads.
 *
 * Layout conventions follow the reference's Path (include/path.h:164-168):
 * beads are [M][N_alloc][3] FP64 (row-major, x fastest), bead locators are
 * {slice, ptcl} int32 pairs, links are [M][N_alloc][2] int32, XXX = -1.
 *
 * All reference citations are file:line in the pimc repository.
 */
#ifndef PIMC_GPU_H
#define PIMC_GPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------- */
#define PIMC_OK 0
#define PIMC_E_ARG -1       /* invalid argument (null pointer, bad size, bad index) */
#define PIMC_E_HIP -2       /* HIP runtime error (message has the HIP error string) */
#define PIMC_E_NOMEM -3     /* device allocation failed or HBM budget exceeded */
#define PIMC_E_STATE -4     /* call not valid in the current state */
#define PIMC_E_UNSUPPORTED -5

/* ---- pair-term evaluation modes -------------------------------------------- */
/* All modes first compute the reference's bin k = int(|r|/dr), dr = 1e-6 rm,
 * from an FP64 minimum-image separation (include/potential.h:249-260), then:   */
#define PIMC_V_MIXED 0  /* Aziz closed form at the bin radius: FP32 transcendental units
                         * (exp2, rcp) around FP64 arithmetic, FP64 sums (default) */
#define PIMC_V_FP32 PIMC_V_MIXED
#define PIMC_V_FP64 1   /* Aziz closed form at the bin radius in FP64 */
#define PIMC_V_TABLE 2  /* gather lookupV[k] from an FP64 table built exactly as the reference's */

typedef struct pimc_ctx pimc_ctx;
typedef struct pimc_aziz pimc_aziz;

/* Cube box + Aziz + chain constants.  Replaces Prism(density, N)
 * (src/container.cpp:84-109), AzizPotential(year, box) (src/potential.cpp:1741-1807),
 * LookupTable(box, M, N) (src/lookuptable.cpp:26-55) and the constants() values
 * tau, lambda, rc that LocalAction reads (src/constants.cpp:127-137). */
typedef struct {
    double side;        /* cube edge [A] */
    int num_slices;     /* M (any M >= 1; the reference itself forces even M) */
    int num_alloc;      /* N_alloc: particle columns of the bead array */
    double tau;         /* imaginary time step [1/K] */
    double lambda;      /* hbar^2/2m [A^2 K] */
    double rc;          /* potential cutoff [A] (Vnn only, src/lookuptable.cpp:504) */
    int aziz_year;      /* 1979 (HFDHE2) | 1987 | 1995 */
    int v_mode;         /* PIMC_V_MIXED | PIMC_V_FP64 | PIMC_V_TABLE */
} pimc_desc;

int  pimc_ctx_create(const pimc_desc* desc, pimc_ctx** out);
void pimc_ctx_destroy(pimc_ctx* ctx);
const char* pimc_last_error(const pimc_ctx* ctx);   /* NULL ctx: last global error */
int  pimc_set_v_mode(pimc_ctx* ctx, int v_mode);
/* LocalAction's factors (src/setup.cpp:1232-1256): VFactor[2] (used where a call
 * passes vfactor = NULL), gradVFactor[2] of the GSF / Li-Broughton correction
 * and the estimator period.  With a finite gradVFactor the total action adds
 * gradVFactor tau^3 lambda gradVSquared(slice) per slice (src/action.cpp:456-472),
 * every bead action at shift 1 adds gradVFactor tau^3 lambda gradVnnSquared(bead)
 * (:487-620, :1282-1337), and pimc_energy uses the corrected tau and lambda
 * derivatives (:751-802).  Primitive action: {1,1}, {0,0}, 1 (the default);
 * GSF: {2/3, 4/3}, {0, 2/9}, 2; Li-Broughton: {1,1}, {1/12, 1/12}, 2. */
int  pimc_set_action_factors(pimc_ctx* ctx, const double vfactor[2], const double gradvfactor[2], int period);

/* Upload a worldline configuration.  Replaces the Path constructor
 * (src/path.cpp:28-129) / loadState + leftPack + LookupTable::updateGrid
 * (src/lookuptable.cpp:314-349).  beads [M][N_alloc][3] (required);
 * bead_on [M][N_alloc] (NULL: all on); num_beads [M] (NULL: N_alloc each);
 * the device cell lists are rebuilt bit-exactly as updateGrid builds them. */
int pimc_upload_path(pimc_ctx* ctx, const double* beads, const uint8_t* bead_on,
                     const int32_t* num_beads);
/* Worm::head/tail/special1/special2 (include/worm.h:31-34), each {slice,ptcl}
 * or {-1,-1}; they set the pair weights of Worm::factor (src/worm.cpp:108-118). */
int pimc_set_worm(pimc_ctx* ctx, const int32_t head[2], const int32_t tail[2],
                  const int32_t special1[2], const int32_t special2[2]);
/* Download the device copy of the beads [M][N_alloc][3]. */
int pimc_download_path(pimc_ctx* ctx, double* beads);
/* Path::updateBead for n beads in order (src/path.cpp:197-200 with
 * LookupTable::updateBead, src/lookuptable.cpp:355-393). */
int pimc_update_beads(pimc_ctx* ctx, const int32_t* slice_ptcl, const double* pos, int n);

/* LocalAction::potentialAction() (src/action.cpp:456-472):
 * action = sum_m vfactor[m%2] * shift*tau * (V_ext(m) + V_int(m)), all i<j pairs,
 * no cutoff.  v_slices [M][2] = {V_ext, V_int} of LocalAction::V(int)
 * (src/action.cpp:902-947), may be NULL. */
int pimc_total_action(pimc_ctx* ctx, int shift, const double vfactor[2], double* action,
                      double* v_slices);
/* Launch shape of the all-pairs kernel (no reference counterpart): about
 * pairs_per_lane pair evaluations per lane and tile (0: about two tiles per
 * workgroup), grid_per_cu persistent workgroups per CU (0: 4), stage_in_lds 1 /
 * 0 / 2 / -1 (stage each slice in LDS in bead order / read partners through the
 * caches / stage slices of 512 beads and more in spatial order, mixed mode /
 * automatic: staged, in spatial order above 1024 beads).  The tiles and the
 * staging order set the summation order, so a context's results are
 * reproducible for a fixed shape and differ across shapes by rounding only. */
int pimc_set_launch_shape(pimc_ctx* ctx, int pairs_per_lane, int grid_per_cu, int stage_in_lds);
/* Times pimc_total_action's pair kernel (reps launches each, kernel events) over
 * 48 shapes (72 where the spatial staging applies) and keeps the fastest: out4 =
 * {pairs_per_lane, grid_per_cu, stage_in_lds, kernel ms} (may be NULL). */
int pimc_autotune(pimc_ctx* ctx, int reps, double* out4);
/* LocalAction::V(int) for every slice: v_slices [M][2]; sep_hist [50] (may be NULL)
 * is the separation histogram the reference leaves behind after evaluating the
 * last slice, M-1 (src/action.cpp:216-224). */
int pimc_potential_slices(pimc_ctx* ctx, double* v_slices, int32_t* sep_hist);
/* EnergyEstimator::accumulate (src/estimator.cpp:940-1029) with the context's
 * action factors: the nine columns in the reference's order (:922)
 *   out9 = {K, V, V_ext, V_int, E, E_mu, K/N, V/N, E/N},  E_mu = E - mu N.
 * next_links [M][N_alloc][2] may be NULL (periodic closed worldlines); the
 * links of the last call stay on the device, and the same links again are
 * neither checked nor uploaded. */
int pimc_energy(pimc_ctx* ctx, const int32_t* next_links, double mu, double* out9);
/* LocalAction::gradVSquared(slice) (src/action.cpp:1188-1223) for every slice:
 * g2 [M] = sum over active beads i of |sum_{j != i} gradV(r_i - r_j)|^2, the
 * values the gradient-corrected action, derivPotentialActionTau (:751-767) and
 * derivPotentialActionLambda (:796-802) weight. */
int pimc_gradv_squared(pimc_ctx* ctx, double* g2);
/* A counter of the context's state (no reference counterpart): every call that
 * changes positions, the bead mask, the worm, the pair-term mode or the action
 * factors increases it (upload, update, a committed Delta V, chain moves), so a
 * caller can cache per-slice values -- V(slice), gradVSquared(slice), the virial
 * terms -- and refresh them only when it moves. */
int pimc_path_version(const pimc_ctx* ctx, uint64_t* version);
/* The virial pair terms of LocalAction for every slice (src/action.cpp:1446-1786,
 * :776-787; ActionBase virtuals include/action.h:75-76, :183-194), with the
 * context's action factors (pimc_set_action_factors) at shift 1:
 *   out6 [M][6] = {rDOTgradUterm1, rDOTgradUterm2, deltadotgradUterm1,
 *                  deltadotgradUterm2, virialKinCorrection,
 *                  secondderivPotentialActionTau}.
 * Every pair of a slice counts (no cutoff).  The T-matrix of the second terms
 * weights its transverse part with |gradV|, as the reference does.  window is
 * virialWindow (>= 1, the reference's default 5); next_links as in pimc_energy
 * (NULL: periodic closed worldlines); every active bead must link to an active
 * bead (PIMC_E_ARG otherwise).  dV/dr and d2V/dr2 at the reference's bin:
 * PIMC_V_MIXED with FP32 exponentials of split exponents (1e-6 relative),
 * PIMC_V_FP64 in FP64 closed form, PIMC_V_TABLE gathered from lookupdVdr and
 * lookupd2Vdr2; sums in FP64.  The links of the last call stay on the device:
 * a call with the same links on the same beads per slice neither checks nor
 * uploads them again. */
int pimc_virial_terms(pimc_ctx* ctx, const int32_t* next_links, int window, double* out6);
/* VirialEnergyEstimator::accumulate (src/estimator.cpp:1086-1251): the 19 values
 * one call adds, in the reference's column order
 *   K_op K_cv V_op V_cv E E_mu K_op/N K_cv/N V_op/N V_cv/N E/N EEcv*Beta^2
 *   Ecv*Beta dEdB CvCov1 CvCov2 CvCov3 E_th P
 * mu is the chemical potential of E_mu.  As in the reference, the term the
 * estimator files under the unregistered key "cVCov2" lands on K_op (index 0)
 * and CvCov2 stays 0 (src/estimator.cpp:1244). */
int pimc_virial_energy(pimc_ctx* ctx, const int32_t* next_links, int window, double mu, double* out19);

/* LocalAction::potentialAction(std::vector<beadLocator>) (src/action.cpp:572-620):
 * sum over active listed beads of vfactor[slice%2]*shift*tau*Vnn(bead)
 * (Vnn: src/action.cpp:1044-1060, pairs inside rc and inside neighbouring cells).
 * per_bead [n] (may be NULL) receives each bead's Vnn. */
int pimc_action_list(pimc_ctx* ctx, const int32_t* slice_ptcl, int n, int shift,
                     const double vfactor[2], double* action, double* per_bead);
/* The two parts of the bead actions that the worm moves take apart:
 *   bare [n] -- LocalAction::barePotentialAction(beads, values) (src/action.cpp:647-689):
 *               vfactor[slice%2]*shift*tau*Vnn(bead), 0 for a bead that is off;
 *   cor  [n] -- LocalAction::potentialActionCorrection(bead) (src/action.cpp:696-716):
 *               gradVFactor[slice%2]*(shift*tau)^3*lambda*gradVnnSquared(bead) at
 *               ANY shift, 0 where gradVFactor is not finite (set by
 *               pimc_set_action_factors) and for a bead that is off (the reference's
 *               callers walk next links, which reach only beads that are on).
 * The range form potentialActionCorrection(start, end) (:723-738) is the sum of
 * cor over the beads from start to end.  Either output may be NULL. */
int pimc_action_parts(pimc_ctx* ctx, const int32_t* slice_ptcl, int n, int shift,
                      const double vfactor[2], double* bare, double* cor);
/* ActionBase/LocalAction::potentialAction(start, end) (src/action.cpp:384-399, :522-570):
 * walks next links from start to end inclusive; next_links as in pimc_energy. */
int pimc_action_range(pimc_ctx* ctx, const int32_t start[2], const int32_t end[2],
                      const int32_t* next_links, int shift, const double vfactor[2],
                      double* action);
/* One-pass Delta V of a move over n beads on distinct slices (the Bisection /
 * Staging / CenterOfMass contract, src/move.cpp:1124-1661): old_pos [n][3]
 * (NULL: the device copy), new_pos [n][3].  Returns
 *   out3 = {A_old, A_new, dU} with A = sum_b vfactor[s%2]*shift*tau*Vnn(b) and
 *   dU = sum of the per-pair differences (new - old), accumulated in FP64.
 * commit != 0 writes new_pos into the device path and updates the device cell
 * lists (Path::updateBead) in list order. */
int pimc_delta_action(pimc_ctx* ctx, const int32_t* slice_ptcl, const double* old_pos,
                      const double* new_pos, int n, int shift, const double vfactor[2],
                      int commit, double* out3);

/* StaticStructureFactorEstimator::accumulate (src/estimator.cpp:3705-3737):
 * sq[i] = sf(q_i)/N, sf = sum_m [N_m + 2 sum_{i<j} cos(q . r_ij)] (minimum image).
 * q [nq][3].  Lattice q (q L / 2pi integer) use the phasor factorisation
 * sum_m |rho_q(m)|^2; other q use the minimum-image pair sum. */
int pimc_ssf(pimc_ctx* ctx, const double* q, int nq, double* sq);
/* pimc_ssf without any allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate (hipMalloc, hipFree or host allocations) inside
 * their loop -- harnesses that forbid it, tight sampling loops whose q lists
 * change size.  Same arguments and results as pimc_ssf, bit for bit: the q list
 * is walked in runs of one kind (lattice / general) that fit one 32 MiB
 * device array per GPU, held by the library until the process ends, instead of
 * the context's buffers grown to the largest list seen.  Any nq is accepted
 * (M <= 512 as for pimc_ssf).  Calls that overlap on one GPU (several threads,
 * several contexts' streams) share that array and race. */
int pimc_ssf_noalloc(pimc_ctx* ctx, const double* q, int nq, double* sq);
/* IntermediateScatteringFunctionEstimator::accumulate (src/estimator.cpp:3923-3961):
 * fqt [nq][M], fqt[i][t] = isf(q_i, t)/N for all M lags, raw positions. */
int pimc_isf(pimc_ctx* ctx, const double* q, int nq, double* fqt);

/* LookupTable dumps (src/lookuptable.cpp:67-516, include/lookuptable.h:139-253).
 * dims3 = numNNGrid; grid [M][N_alloc][3], bead_label [M][N_alloc],
 * num_labels [cells][M], hash [cells][M][hash_cap] (any NULL skipped). */
int pimc_cell_dims(pimc_ctx* ctx, int32_t dims3[3], int32_t* hash_cap);
int pimc_cell_lists(pimc_ctx* ctx, int32_t* grid, int32_t* bead_label, int32_t* num_labels,
                    int32_t* hash);
/* LookupTable::updateInteractionList (src/lookuptable.cpp:464-516) for one bead:
 * list2 [N_alloc][2], sep3 [N_alloc][3]; *count receives numBeads. */
int pimc_interaction_list(pimc_ctx* ctx, int slice, int ptcl, int32_t* list2, double* sep3,
                          int* count);

/* ---- PotentialBase drop-in for the Aziz pair potential ------------------------ */
/* AzizPotential(year, box) with maxSep and the tail-correction cutoff rc
 * (src/potential.cpp:1741-1807). */
int  pimc_aziz_create(int year, double max_sep, double rc, int v_mode, pimc_aziz** out);
void pimc_aziz_destroy(pimc_aziz* az);
/* PotentialBase::V(const dVec*, double*, int) batched form (include/potential.h:86,
 * src/potential.cpp:127-132) for pair separations: out[i] = V(sep[i]). */
int pimc_aziz_V(pimc_aziz* az, const double* sep3, double* out, int n);
/* gradV / grad2V (include/potential.h:997-1018): gradV = (dV/dr)(k)/|r| * r. */
int pimc_aziz_gradV(pimc_aziz* az, const double* sep3, double* out3, int n);
int pimc_aziz_grad2V(pimc_aziz* az, const double* sep3, double* out, int n);
double pimc_aziz_tailV(const pimc_aziz* az);   /* PotentialBase::tailV */
int  pimc_aziz_table_length(const pimc_aziz* az);


/* ---- batched form: many independent walkers in one launch -------------------- */
/* LocalAction::potentialAction() (src/action.cpp:456-472) of n walkers with uneven
 * N and M (all primitive, one Aziz year).  Walkers are cut into work items of
 * pairs_per_lane * 256 pairs each; a persistent grid of workgroups walks them;
 * the launch sequence can be captured once into a hipGraph and replayed.  No
 * reference counterpart: the reference evaluates one walker per process. */
typedef struct pimc_batch pimc_batch;
int  pimc_batch_create(const pimc_desc* walkers, int n_walkers, pimc_batch** out);
void pimc_batch_destroy(pimc_batch* b);
const char* pimc_batch_last_error(const pimc_batch* b);
int  pimc_batch_upload(pimc_batch* b, int walker, const double* beads);
int  pimc_batch_set_tiles(pimc_batch* b, int pairs_per_lane, int grid);
int  pimc_batch_use_graph(pimc_batch* b, int enable);
/* actions [n_walkers]; v_slices [total slices] (walker-major, may be NULL). */
int  pimc_batch_total_action(pimc_batch* b, int shift, const double vfactor[2], double* actions,
                             double* v_slices);
int  pimc_batch_info(const pimc_batch* b, int64_t* total_pairs, int* n_items, int* n_groups,
                     double* last_ms);
/* Autotuner: times pairs_per_lane in {16..256} x grid in {2,4,8,16} per CU and keeps
 * the fastest; out4 = {pairs_per_lane, grid, ms, candidates}. */
int  pimc_batch_autotune(pimc_batch* b, int reps, double* out4);

/* ---- serving form: Delta V and S(q) requests of many chains, continuous batching */
/* A Delta V request is the move over beads on distinct slices of one chain (the
 * Staging / CenterOfMass contract, src/move.cpp:1124-1408), decided on the device
 * by rand() < exp(-(A_new - A_old)) with the caller's uniform deviate and
 * committed when accepted; an S(q) request is the static structure factor of the
 * chain (pimc_server_submit_ssf).  No reference counterpart: the reference runs
 * one chain per process. */
typedef struct pimc_server pimc_server;
typedef struct {
    size_t hbm_budget;  /* bytes of chain state kept resident (arena size) */
    int streams;        /* batches in flight, 1..4 */
    int aziz_year;
    int v_mode;
    double max_side;    /* largest box edge any chain will use [A] */
} pimc_server_desc;
int  pimc_server_create(const pimc_server_desc* desc, pimc_server** out);
void pimc_server_destroy(pimc_server* s);
const char* pimc_server_last_error(const pimc_server* s);
/* Adds a chain (beads [M][N][3]); it is resident at once when the budget allows,
 * else it waits (state 0) until pimc_server_remove_chain frees room. */
int  pimc_server_add_chain(pimc_server* s, const pimc_desc* d, const double* beads, int* chain_id);
int  pimc_server_chain_state(pimc_server* s, int chain, int* state);  /* 0 waiting, 1 resident, 2 removed */
int  pimc_server_remove_chain(pimc_server* s, int chain);
int  pimc_server_download(pimc_server* s, int chain, double* beads);
/* commit_mode: 0 evaluate only, 1 Metropolis with u, 2 always commit. */
int  pimc_server_submit(pimc_server* s, int chain, const int32_t* slice_ptcl, const double* new_pos,
                        int n, int shift, const double vfactor[2], int commit_mode, double u,
                        uint64_t* req_id);
int  pimc_server_poll(pimc_server* s, uint64_t req_id, int* done);
/* out5 = {A_old, A_new, dU, accepted, latency_us}; releases the request.
 * (A Delta V request only; S(q) requests are read with pimc_server_wait_ssf.) */
int  pimc_server_wait(pimc_server* s, uint64_t req_id, double* out5);
/* StaticStructureFactorEstimator::accumulate (src/estimator.cpp:3705-3737) on a
 * resident chain: sq[i] = sf(q_i)/N of the chain's worldlines at the time the
 * request reaches the device.  It is queued behind the chain's earlier requests
 * (so it sees every move submitted before it) and takes a batch slot of its own
 * when one is free.  q [nq][3] must be lattice vectors 2 pi n / L of the chain's
 * box with |n_d| <= 40 (PIMC_E_UNSUPPORTED otherwise); 1 <= nq <= 65536.  The
 * values equal pimc_ssf's on the same worldlines bit for bit. */
int  pimc_server_submit_ssf(pimc_server* s, int chain, const double* q, int nq, uint64_t* req_id);
/* Waits for an S(q) request: sq [nq] receives S(q); latency_us (may be NULL) the
 * time from submission to completion.  Releases the request. */
int  pimc_server_wait_ssf(pimc_server* s, uint64_t req_id, double* sq, double* latency_us);
int  pimc_server_stats(pimc_server* s, int64_t* batches, int64_t* requests, int64_t* pair_evals,
                       size_t* free_bytes);
/* Delta V batches by how they reached the GPU (no counterpart in the reference):
 * as AQL packets on the launching slot's own queue, their descriptors written
 * into device memory through the BAR (batches of up to 64 KB of descriptors),
 * or as a copy and two hipLaunchKernel calls on the slot's stream (larger
 * batches, and every batch when PIMC_DIRECT_DISPATCH=0 was set at creation). */
int  pimc_server_dispatch_stats(pimc_server* s, int64_t* packet_batches, int64_t* hip_batches);
typedef struct {
    int n_chains;       /* chains 0..n_chains-1 receive the load */
    double rate;        /* Poisson arrival rate [requests/s] over all chains; <= 0: closed loop */
    double duration_s;
    int seg_min, seg_max;  /* segment length in slices (moved beads = length - 1) */
    double p_com;       /* fraction of CenterOfMass requests (whole worldline) */
    double com_sigma;   /* Gaussian CoM shift per axis [A] */
    double lambda;      /* segment proposal width sqrt(lambda tau) */
    uint32_t seed;
    int threads;        /* client threads submitting (chains split among them); 0 = 1 */
    double p_ssf;       /* fraction of arrivals that are S(q) requests (0: none) */
    int ssf_nmax;       /* their q: every 2 pi n / L with |n| <= ssf_nmax (4: 257 q) */
} pimc_load_desc;
/* stats16 = {completed, req/s, pair evals/s, p50_us, p99_us, mean_us, accept frac,
 *            batches, requests per batch, seconds, service p50_us, service p99_us,
 *            S(q) completed, S(q) p50_us, S(q) p99_us, (q, bead) phasors/s};
 * the first twelve count the Delta V requests, the last four the S(q) ones;
 * latency counts from arrival, service from submission to the worker retiring
 * the result. */
int  pimc_server_run_synthetic(pimc_server* s, const pimc_load_desc* load, double* stats16);

/* ---- one Markov chain: the diagonal moves over a resident path ---------------- */
/* StagingMove / BisectionMove / CenterOfMassMove::attemptMove (src/move.cpp:
 * 1124-1259, 1315-1408, 1484-1661; MoveBase helpers :270-396, :467-488) on a
 * context holding closed worldlines (every bead on, no worm).  Proposals and the
 * Metropolis test draw on the chain's own MTRand(seed) stream in the reference's
 * order against a host copy of the positions; every oldAction/newAction pair is
 * one pimc_delta_action pass on the device; accepted moves stay committed on the
 * device and rejected ones are undone as the reference undoes them.  The chain
 * requires that nothing else moves the context's beads while it lives. */
typedef struct pimc_chain pimc_chain;
typedef struct pimc_chain_desc {
    uint32_t seed;        /* MTRand(seed) */
    int Mbar;             /* staging length (2 <= Mbar < M) */
    int levels;           /* bisection levels b (2^b < M) */
    int var_updates;      /* draw the length / level count per move */
    int max_wind;         /* winding sectors sampled by staging */
    double com_delta;     /* centre-of-mass shift width */
} pimc_chain_desc;
int  pimc_chain_create(pimc_ctx* ctx, const pimc_chain_desc* d, pimc_chain** out);
void pimc_chain_destroy(pimc_chain* ch);
const char* pimc_chain_last_error(const pimc_chain* ch);
/* kind 0 staging, 1 bisection, 2 centre of mass; *accepted = Metropolis outcome;
 * info3 = {attempted, last newAction - oldAction, length (stage length, levels or
 * worldline length)}. */
int  pimc_chain_move(pimc_chain* ch, int kind, int* accepted, double* info3);
/* The chain's positions, [M][N][3] (equal to the device copy). */
int  pimc_chain_positions(const pimc_chain* ch, double* beads);
/* n_moves moves of kinds[k] (0 staging, 1 bisection, 2 centre of mass) in one
 * device launch: the proposals, the Metropolis draws and the commits run on the
 * device from the chain's MTRand state, which continues the host stream and is
 * handed back afterwards (pimc_chain_move may follow).  accepted [n_moves] and
 * info3 [n_moves][3] as pimc_chain_move (either may be NULL).  The device's
 * log / sqrt / exp / pow may differ from the host's in the last ulp, so a
 * proposal can too; *near_threshold (may be NULL) counts the decisions whose
 * uniform deviate lay within a relative 1e-10 of the acceptance probability,
 * the only ones such a difference could flip.  With the gradient correction
 * (pimc_set_action_factors) every shift-1 bead action adds gradVFactor tau^3
 * lambda gradVnnSquared at the old and the new position, as pimc_delta_action
 * does.  Requires M <= 512 (and N <= 8192 with the correction;
 * PIMC_E_UNSUPPORTED otherwise). */
int  pimc_chain_run_device(pimc_chain* ch, int n_moves, const int* kinds, int* accepted, double* info3,
                           int* near_threshold);
/* The same for n_chains chains (each on its own context, one pair-term mode) in
 * one launch, one workgroup per chain: independent chains, the reference's
 * process-per-seed ensemble (src/pdrive.cpp:41,63), sharing the GPU.  kinds
 * [n_moves] is every chain's move list; accepted [n_chains][n_moves], info3
 * [n_chains][n_moves][3], near_threshold [n_chains] (each may be NULL). */
int  pimc_chain_run_device_many(pimc_chain* const* chains, int n_chains, int n_moves, const int* kinds,
                                int* accepted, double* info3, int* near_threshold);

/* ---- state files ------------------------------------------------------------- */
/* The reference's text state file (PathIntegralMonteCarlo::saveState, src/pimc.cpp:
 * 926-986, written with the stream operators of include/common.h:244-272):
 * the worldline count, the acceptance and estimator counter lines ("%16d\t%16d"),
 * then beads (16 significant digits), nextLink, prevLink and worm.beads as
 * "(0,M-1) x (0,N-1)" arrays, then the MTRand save words.  Host only.  Links are
 * {slice, ptcl} with {-1,-1} for XXX.  The HDF5 variant (useHDF5State) is not
 * read: no HDF5 library is present.  Errors: pimc_state_last_error(). */
enum { PIMC_MT_SAVE = 625 };   /* MTRand::SAVE = N + 1 words (include/MersenneTwister.h:80) */
typedef struct pimc_state pimc_state;
const char* pimc_state_last_error(void);
/* Parse a state file (the stream extractors of include/common.h:276-397 with
 * loadState's skip to the first '(' line, src/pimc.cpp:1131-1144). */
int  pimc_state_read(const char* path, pimc_state** out);
int  pimc_state_parse(const char* text, size_t len, pimc_state** out);
/* An empty state of M x N beads (links XXX, worm 0, no counters, no random words). */
int  pimc_state_create(int num_slices, int num_alloc, pimc_state** out);
void pimc_state_destroy(pimc_state* st);
/* Array extents (slices, columns) and the first line's worldline count. */
int  pimc_state_dims(const pimc_state* st, int* num_slices, int* num_alloc, int* num_worldlines);
/* Copy the arrays out / in: beads [M][N][3], next / prev [M][N][2], worm [M][N];
 * any pointer may be NULL; pimc_state_set rejects links outside the arrays. */
int  pimc_state_get(const pimc_state* st, double* beads, int32_t* next, int32_t* prev, uint32_t* worm);
int  pimc_state_set(pimc_state* st, const double* beads, const int32_t* next, const int32_t* prev,
                    const uint32_t* worm);
/* Counter lines as {a, b} pairs in file order (the totals first): *count pairs
 * exist, up to cap are copied to pairs; set_pairs (set_count pairs) replaces them. */
int  pimc_state_counters(pimc_state* st, int64_t* pairs, int cap, int* count, const int64_t* set_pairs,
                         int set_count);
/* MTRand::save words [PIMC_MT_SAVE] (uint32 is unsigned long in the reference):
 * *have = 0 when the file carried none; set_words replaces them. */
int  pimc_state_random(pimc_state* st, uint64_t* words, int* have, const uint64_t* set_words);
/* Path::leftPack (src/path.cpp:145-189): active beads moved to the left of each
 * slice, their links and their neighbours' links rewritten.  saveState calls it
 * before writing.  A worm end's missing neighbour (link XXX) is left alone. */
int  pimc_state_left_pack(pimc_state* st);
/* The file text, byte for byte as saveState writes it; *len excludes the NUL. */
int  pimc_state_format(const pimc_state* st, char* buf, size_t cap, size_t* len);
int  pimc_state_write(const pimc_state* st, const char* path);
/* loadState's path set-up (src/pimc.cpp:1150-1253) for a run of num_slices
 * slices in a cube of side `side`: the file's arrays when it has num_slices
 * slices, else the classical path from slice 0 of every active worldline
 * (loadClassicalState, :1006-1027; periodic links, inactive beads unlinked);
 * then every bead put inside the box and the active beads counted per slice.
 * Outputs have num_worldlines columns: beads [M][NW][3], bead_on [M][NW],
 * num_beads [M], next / prev [M][NW][2] -- ready for pimc_upload_path,
 * pimc_energy and pimc_virial_*.  The worm is reset (no head or tail). */
int  pimc_state_to_path(const pimc_state* st, int num_slices, double side, double* beads, uint8_t* bead_on,
                        int32_t* num_beads, int32_t* next, int32_t* prev);

/* ---- seeded synthetic worldlines -------------------------------------------- */
/* MTRand(seed) stream (include/MersenneTwister.h); slice 0 Prism::randPosition
 * (src/container.cpp:155-160); kind 0: randNorm(prev, sqrt(2 lambda tau)) + putInBC
 * per slice; kind 1: closed Gaussian bridge by the staging construction
 * (src/move.cpp:270-291).  beads [M][N][3].  Host only. */
int pimc_gen_worldlines(uint32_t seed, double side, int M, int N, double lambda, double tau, int kind,
                        double* beads);

/* ---- measurement hooks ------------------------------------------------------- */
/* When enabled, every launch of a hot kernel is bracketed by events (from a
 * pre-created pool) on the stream it runs on; pimc_kernel_time returns the summed
 * duration [ms] and the launch count for name in
 * {"pair_action","delta_action","rho","ssf","isf","gradv2","gradvnn","virial","chain_run"}.
 * "pair_action" and "delta_action" are the pair kernels alone (k_pair_action,
 * k_vnn), without their reductions; "gradv2" and "gradvnn" the gradient
 * correction's slice and three-body kernels. */
int pimc_set_timing(pimc_ctx* ctx, int enable);
/* Timing for a subset of the kernels above: bit k of `kinds` brackets the k-th
 * name in the list (0: "pair_action", 1: "delta_action", ...); 0 turns timing
 * off.  A caller that needs one kernel's duration leaves the others' launches
 * free of event records. */
int pimc_set_timing_kinds(pimc_ctx* ctx, unsigned kinds);
/* Besides the names above, "delta_packets" counts the Delta V requests sent as
 * dispatch packets straight into the device's AQL queue (total_ms 0): requests
 * of up to 32 beads on slices of up to 1024, without the gradient correction
 * or per-bead outputs, while timing of "delta_action" is off.  The environment
 * variable PIMC_DIRECT_DISPATCH=0, read when a context is created, sends them
 * through hipLaunchKernel instead.  "isf_chunks" counts the q chunks of
 * pimc_isf calls that took the chunked form (total_ms 0): lattice q only, at
 * least 2048 of them, results of 4 MB and more into page-locked memory
 * (pimc_host_alloc), while timing of "rho" and "isf" is off; lists of up to
 * 68 000 q run as two halves, longer ones in chunks of 34 000 q (more when the
 * list exceeds 64 chunks); each chunk computes its density modes and lags
 * while the previous chunk's rows are copied out on a second stream, with the
 * same values as one launch per kernel. */
int pimc_kernel_time(pimc_ctx* ctx, const char* name, double* total_ms, long* count);
int pimc_device_synchronize(void);
/* Page-locked host memory for large results (no counterpart in the reference,
 * whose estimators write into their own arrays): F(q,tau) over 100 401 q and 250
 * lags is 200 MB, which the device copies into page-locked memory at the full
 * link rate but into pageable memory through staging buffers at a fraction of
 * it.  Any entry point's output pointer may point into such a block. */
int pimc_host_alloc(size_t bytes, void** ptr);
int pimc_host_free(void* ptr);
/* A plain device-to-device streaming copy of `bytes` run `iters` times (no
 * reference counterpart): a known-size kernel against which the HBM byte
 * counters of the hot kernels are scaled.  *avg_ms receives the mean duration. */
int pimc_copy_probe(size_t bytes, int iters, double* avg_ms);

#ifdef __cplusplus
}
#endif
#endif
