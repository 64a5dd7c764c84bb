/*
 * souffle_gpu.h — C ABI of the MI355X (gfx950) force-and-integrate engine for SOUFFLE
 * capsomere models.
 *
 * The library replaces the reference's per-step hot path:
 *   forceCalculation(...)                   src/forces.h:13-15, src/forces.cpp:13-171
 *   update_LJ_ES_energies_simplified(...)   src/energies.h:10,  src/energies.cpp:24-93
 *   particle_kinetic_energy(...)            src/energies.h:12,  src/energies.cpp:13-22
 *   the body of the step loop               src/md.cpp:387-514 (Nose-Hoover MD and "Brownian")
 * behind plain-pointer entry points. Everything runs on the GPU; the caller owns host arrays,
 * the library owns device memory. A handle holds one system or a batch of independent replicas
 * (same template, own box/salt/epsilon/temperature each). Handles are not thread-safe.
 *
 * Arrays of bead vectors are AoS [N][3] doubles, beads in the reference's order
 * (subunit-major, template order inside a subunit, src/initialize.cpp:221-223); for a batch,
 * replica 0's beads come first, then replica 1's, and so on.
 *
 * Every int-returning call returns 0 on success and a negative code on failure
 * (SF_E_*); sf_last_error() gives the message of the last failure on this thread.
 */
#ifndef SOUFFLE_GPU_H
#define SOUFFLE_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SF_OK 0
#define SF_E_ARG (-1)       /* invalid argument or unsupported template */
#define SF_E_HIP (-2)       /* HIP runtime error (no device, out of memory, launch failure) */
#define SF_E_IO (-3)        /* template or file could not be read */
#define SF_E_OVERFLOW (-4)  /* a neighbour list outgrew its capacity and could not be regrown */
#define SF_E_NONFINITE (-5) /* a replica's velocities became non-finite (sf_step stops before it) */

typedef struct sf_template sf_template; /* parsed capsomere template + derived topology */
typedef struct sf_handle sf_handle;     /* device-resident system or batch of replicas */

/* The reference's command-line values that reach the hot path (src/md.cpp:53-90). */
typedef struct {
  double C_uM;    /* -C capsomere concentration, micromolar */
  double c_mM;    /* -c salt concentration, millimolar */
  double ks;      /* -s stretching constant (kT) */
  double kb;      /* -b bending constant (kT) */
  double elj_att; /* -E attractive LJ well depth */
  double ecut_c;  /* -e Yukawa cutoff in screening lengths; 0 = no cutoff */
  double dt;      /* -t timestep */
  double K;       /* -K temperature (Kelvin) */
  double Q;       /* -Q Nose-Hoover mass */
  double damp;    /* -d BD damping */
  int n_subunits; /* -S */
  int q_chain;    /* -q Nose-Hoover chain length */
  int brownian;   /* -D: 1 = 'b', 0 = 'm' */
} sf_cli;

/* Physical and numerical parameters of one system (one replica). */
typedef struct {
  double box;      /* cubic box edge, sigma units */
  double ecut;     /* LJ cutoff */
  double lb;       /* Bjerrum length */
  double kappa;    /* inverse screening length */
  double ecut_el;  /* Yukawa cutoff; 0 = none: every charged pair interacts (minimum image), as
                      the reference's -e 0 does with a list radius that covers the box */
  double ks, kb, elj_att;
  double T;        /* reduced temperature K/298 */
  double dt, damp;
  double Q;        /* mass of the first Nose-Hoover bath */
  int q_chain;     /* Nose-Hoover chain length (>= 1) */
  int brownian;    /* 1 = BD integrator, 0 = MD + Nose-Hoover */
  double skin;     /* Verlet skin of the GPU lists (sigma); the list is rebuilt on the device
                      when any bead of the replica has moved more than skin/2 */
  int pair_fp32;   /* 1 = LJ and Yukawa pair arithmetic in FP32 (positions, velocities,
                      thermostat and energy sums stay FP64) */
  uint32_t bd_seed;/* seed of the BD noise engine; the reference uses 69521 */
  double far_rf;   /* uncut Yukawa only (ecut_el = 0): centre-bead distance beyond which a
                      subunit pair is far field (its Yukawa sum is taken subunit pair by subunit
                      pair, exactly or by a far-field model); 0 = 2.5 capsomere diameters */
} sf_params;

typedef struct {
  int64_t steps;          /* loop iterations done since sf_set_state */
  int64_t rebuilds;       /* neighbour-list rebuilds since sf_set_state */
  int64_t lj_entries;     /* LJ/WCA list entries at the last rebuild (all beads) */
  int64_t q_entries;      /* Yukawa list entries at the last rebuild (charged beads) */
  int32_t lj_capacity;    /* current per-bead capacities */
  int32_t q_capacity;
  int32_t n_beads, n_subunits, n_charged, n_replicas;
  double k_integrate_ms;  /* mean kernel durations over the last timed sf_step (HIP events on the */
  double k_rebuild_ms;    /* library's stream around every 10th step's kernels) */
  double k_force_ms;
  double k_reduce_ms;
  double k_timed_steps;   /* steps the means were taken over (0 if timing was off) */
  double vcom;            /* replica 0: |centre-of-mass velocity| removed at the last step */
  double k_far_ms;        /* far field of the uncut Yukawa (exact or model), as the means above */
  int64_t graph_launches; /* HIP graph launches of the step loop since sf_create */
  int32_t graph_steps;    /* steps per graph launch (sf_set_graph_steps) */
} sf_stats;

/* ---- template and parameters ---------------------------------------------------------- */

/* Reads a template in the reference's infiles format (src/initialize.cpp:46-128) and derives
 * the per-template topology: bending hinges (src/initialize.cpp:348-404, src/face.cpp:27-34),
 * per-bead edge and hinge incidence, charged beads, centre bead. NULL on failure. */
sf_template* sf_template_load(const char* path);
void sf_template_free(sf_template* t);

/* Template sizes: any pointer may be NULL. */
int sf_template_info(const sf_template* t, int* n_beads, int* n_edges, int* n_faces,
                     int* n_hinges, int* n_charged, int* centre_bead, double* si_sigma,
                     double* si_mass);

/* Capsomere diameter (largest bead-bead distance of the template, sigma) and the largest
 * distance of any bead / any charged bead from the centre bead. */
int sf_template_extent(const sf_template* t, double* diameter, double* r_all, double* r_charged);

/* Body frame of a subunit (used by the far-field model): e1 along (frame bead 1 - centre bead),
 * e2 the part of (frame bead 2 - centre bead) normal to e1, e3 = e1 x e2. Frame bead 1 is the
 * bead farthest from the centre bead, frame bead 2 the bead farthest from that axis.
 * qbody (nullable) [n_charged][3]: charged beads' template coordinates in that frame.
 * Returns SF_E_ARG if the template's beads are collinear. */
int sf_template_frame(const sf_template* t, int* fp1, int* fp2, double* qbody);

/* Template-local hinge table [n_hinges][4] = (B0, B1, across F0, across F1). */
int sf_template_hinges(const sf_template* t, int* out, int cap);

/* Derives box, cutoffs, Bjerrum length and kappa from the reference's flags exactly as
 * src/md.cpp:135-173 does; skin = 0.3, pair_fp32 = 0, bd_seed = 69521, far_rf = 0. */
int sf_params_derive(const sf_template* t, const sf_cli* cli, sf_params* out);

/* Seeded initial state: non-overlapping subunits (bead gaps >= 1 sigma) at uniform random
 * centres and uniform random orientations in [-box/2, box/2)^3, Maxwell-Boltzmann velocities
 * at T, centre-of-mass velocity removed and kinetic energy rescaled to exactly 1.5 N T.
 * Host-only. pos, vel: [n_subunits * n_beads][3]. */
int sf_make_state(const sf_template* t, int n_subunits, double box, double T, uint64_t seed,
                  double* pos, double* vel);

/* ---- device system -------------------------------------------------------------------- */

/* One handle for n_replicas independent systems; n_subunits[r] and params[r] per replica.
 * Allocates device memory on the current HIP device. NULL on failure. */
sf_handle* sf_create(const sf_template* t, int n_replicas, const int* n_subunits,
                     const sf_params* params);
void sf_destroy(sf_handle* h);

/* Uploads positions and velocities (all replicas) and (re)starts every replica as the
 * reference's run start does: thermostat xi = eta = 0, BD engine reseeded, step counter 0,
 * neighbour lists built, forces computed (src/md.cpp:241) and the kinetic energy taken
 * (src/md.cpp:258). vel may be NULL (zero velocities). */
int sf_set_state(sf_handle* h, const double* pos, const double* vel);

/* Downloads positions and velocities (either may be NULL). */
int sf_get_state(sf_handle* h, double* pos, double* vel);

/* What a run needs besides positions and velocities to continue bit for bit: the reference keeps
 * none of it across a restart (src/md.cpp:187-196 starts the chain at zero, src/md.cpp:382 seeds
 * the BD engine afresh, src/initialize.cpp:166-191 reads positions and velocities only), so a
 * reference restart is sf_set_state; a bit-continuous resume is sf_set_state_resume. */
typedef struct {
  int64_t step;     /* loop iterations done since the start (the reference's counter a) */
  int32_t q_chain;  /* Nose-Hoover chain length (0 for BD) */
  uint32_t rng;     /* BD engine state before the next step's draws (minstd_rand0 state) */
  double ke;        /* kinetic energy the chain's next update uses (measured before the
                       centre-of-mass removal of the last step, src/md.cpp:456-514) */
  double xi[8];     /* chain variables after the last completed step (src/thermostat.h:8-54) */
  double eta[8];
} sf_resume;

/* Resume records of every replica (out[n_replicas]) and, when tforce is not NULL, the integrator
 * force [N][3] the next step's first half kick uses (for BD it holds the last step's noise and
 * drag, src/md.cpp:462-498, so a BD resume needs it; for MD it equals the force at the current
 * positions). */
int sf_get_resume(sf_handle* h, sf_resume* out, double* tforce);

/* As sf_set_state, but every replica continues from its resume record (step counter, chain,
 * kinetic energy, BD engine) instead of starting afresh, and, when given, with the integrator
 * force tforce [N][3] (required for BD). With the positions, velocities, records and force of
 * sf_get_state / sf_get_resume, the steps that follow are bit for bit those of the uninterrupted
 * run. */
int sf_set_state_resume(sf_handle* h, const double* pos, const double* vel, const sf_resume* res,
                        const double* tforce);

/* Replaces positions only, for callers that run their own integrator and call the engine where
 * the reference calls forceCalculation (src/md.cpp:241,443): velocities, thermostat and BD
 * engine are kept, and a replica's lists are marked for a build when any of its beads moved more
 * than skin/2 since the last one. */
int sf_set_positions(sf_handle* h, const double* pos);

/* Nose-Hoover chain variables of one replica (arrays of q_chain); returns q_chain. */
int sf_get_thermostat(sf_handle* h, int replica, double* xi, double* eta);

/* Forces at the current positions with a fresh list build: the forceCalculation equivalent.
 * Outputs [N][3], any may be NULL: total = stretch + bend + (Yukawa + LJ), and the four terms.
 * Does not change the integrator state. */
int sf_forces(sf_handle* h, double* f_total, double* f_stretch, double* f_bend,
              double* f_yukawa, double* f_lj);

/* sf_forces with the reference's updatePairlist argument (src/forces.h:13-15): force_rebuild = 1
 * builds the lists anew; 0 builds them only if a bead moved more than skin/2 since the last
 * build (the forces are the same either way: the lists hold every pair inside the cutoffs). */
int sf_forces_ex(sf_handle* h, int force_rebuild, double* f_total, double* f_stretch, double* f_bend,
                 double* f_yukawa, double* f_lj);

/* Energies at the current state (src/md.cpp:553-605): totals[n_replicas][7] =
 * {kinetic, stretching, bending, LJ, Yukawa, thermostat potential, thermostat kinetic};
 * per_bead (nullable) [N][5] = {se, be, ne, ce, ke} as the reference's BEAD fields. */
int sf_energies(sf_handle* h, double* totals, double* per_bead);

/* Runs n_steps iterations of the step loop body (src/md.cpp:400-514) for every replica.
 * If a replica's velocities become non-finite (the reference carries on with NaN), every replica
 * is left at the last finite step and SF_E_NONFINITE is returned; sf_last_error names the step
 * and the replica, and sf_get_state / sf_get_resume give that state. */
int sf_step(sf_handle* h, int n_steps);

/* Steps per HIP graph launch of sf_step (no reference counterpart: the reference's loop is host
 * code, src/md.cpp:387-514). sf_step captures that many loop iterations of every replica into
 * one graph (5 kernel nodes a step), launches it once, and reads the capacity, deformation and
 * finiteness flags only when it ends; a launch that overflowed a capacity is replayed from its
 * start with grown buffers, one that went non-finite up to its last finite step. 0 = the
 * default of 100; at most 20000. A run of n <= steps iterations is then a single launch with no
 * host round trip. With a far-field model attached (audited every 100 steps) or with timing on,
 * launches stay 100 steps long. Results are the same bit for bit for every length. */
int sf_set_graph_steps(sf_handle* h, int steps);

/* Per-kernel timing with HIP events on the library's stream during sf_step (off by default). */
int sf_set_timing(sf_handle* h, int enable);

int sf_get_stats(sf_handle* h, sf_stats* out);

/* The BD noise draws of one replica for loop iteration `step` (0 = first iteration after
 * sf_set_state): [n_beads_of_replica][3] uniform(-0.5, 0.5) values, the stream of the
 * reference's std::default_random_engine(seed) + uniform_real_distribution(-0.5, 0.5)
 * (src/md.cpp:382-385, 464-466), computed on the GPU by skip-ahead. */
int sf_bd_uniforms(sf_handle* h, int replica, int64_t step, double* out);

/* ---- tuning (no reference counterpart: the reference's list is rebuilt every -B steps with
 * radius -L, src/md.cpp:438-443; here the list is a Verlet list with a displacement trigger) -- */

/* Verlet skin of one replica (replica >= 0) or of all (-1). The next step rebuilds the lists.
 * Forces do not depend on the skin (lists hold every pair inside the cutoffs either way), so
 * the trajectory is unchanged bit for bit; only the rebuild frequency and list lengths are. */
int sf_set_skin(sf_handle* h, int replica, double skin);
int sf_get_skin(sf_handle* h, int replica, double* skin);

/* Pair-gather tile of the force kernel: list entries a subunit workgroup evaluates per round
 * (128, 256, 384 or 512). Results are the same bit for bit for every tile. */
int sf_set_tile(sf_handle* h, int tile);

typedef struct {
  int trial_steps;       /* timed steps per candidate; <= 0: 200 */
  int n_skins;           /* candidate skins; 0: {0.2, 0.3, 0.5, 0.8, 1.2} */
  const double* skins;
  int n_tiles;           /* candidate tiles; 0: {128, 256, 384, 512} */
  const int* tiles;
} sf_tune_opts;

typedef struct {
  double skin;           /* chosen skin (a replica whose box cannot take it gets its largest valid one) */
  int tile;              /* chosen tile */
  double ms_per_step;    /* wall time per step of the chosen pair in its trial */
  int n_trials;
} sf_tune_result;

/* Autotuner: from the current state, runs every (skin, tile) candidate for one untimed
 * 100-step chunk and then trial_steps timed steps (wall clock, HIP graphs as in sf_step),
 * restores the state after each trial, and leaves the fastest pair set. The state, the
 * step counter and the trajectory that follows are exactly as without the call.
 * opts may be NULL (defaults); trial_ms (nullable) receives [n_skins][n_tiles] ms/step. */
int sf_autotune(sf_handle* h, const sf_tune_opts* opts, sf_tune_result* out, double* trial_ms);

/* ---- far-field model of the uncut Yukawa (no reference counterpart: the reference sums every
 * list pair, src/forces.cpp:104-107). A small network (13 -> 128 -> 128 -> 128 -> 6, SiLU)
 * maps the pose of subunit B seen from subunit A (B's centre in A's body frame, relative
 * orientation, kappa) to the force on A's charged beads and its torque about A's centre bead;
 * trained on the GPU on poses scored by the exact pair sum; evaluated with BF16 MFMA and FP32
 * accumulation for subunit pairs beyond far_rf. ----------------------------------------- */

typedef struct sf_far_model sf_far_model;

/* New model for a template (1..8 charged beads), split radius far_rf (0 = 2.5 capsomere
 * diameters) and salt range [c_lo_mM, c_hi_mM]; Glorot-uniform weights from seed. */
sf_far_model* sf_far_model_create(const sf_template* t, double far_rf, double c_lo_mM, double c_hi_mM,
                                  uint64_t seed);
void sf_far_model_destroy(sf_far_model* m);
int sf_far_model_info(const sf_far_model* m, double* far_rf, double* kappa_min, double* kappa_max, int* n_params);

/* Parameters (FP32, packed W1 [128][13], b1, W2 [128][128], b2, W3, b3, W4 [6][128], b4) and
 * output standardisation (mean[6], sd[6]): any pointer may be NULL. */
int sf_far_model_params(sf_far_model* m, float* get, const float* set, float* norm_get, const float* norm_set);

/* n seeded training poses generated and scored on the GPU: features x [n][13], scaled targets
 * y [n][6] (force / S, torque / (S rq)) with S = lb (kappa + 1/d) / d exp(-kappa d)
 * sum_a exp(kappa p_a) sum_b exp(kappa p'_b), p_a and p'_b the charges' projections towards
 * the other subunit; pose [n][16] = (d_vec[3], Q[9] row-major, kappa, c_mM, S, 0); any output
 * may be NULL. */
int sf_far_model_samples(sf_far_model* m, int n, uint64_t seed, double c_lo_mM, double c_hi_mM, float* x, float* y,
                         double* pose);

/* Mean squared error of the network on (x [n][13], standardised y [n][6]) and its gradient
 * with respect to every parameter, by the GPU training kernels. */
int sf_far_model_grad(sf_far_model* m, int n, const float* x, const float* y, float* grad, double* loss);

/* Network outputs (standardised units) for x [n][13]: BF16 MFMA (bf16_mfma = 1, the engine's
 * kernel) or the FP32 training forward (0). */
int sf_far_model_infer(sf_far_model* m, int n, const float* x, float* y, int bf16_mfma);

typedef struct {
  int n_samples;       /* training poses (<= 0: 1,000,000) */
  int epochs;          /* <= 0: 20 */
  int batch;           /* <= 0: 4096 */
  double lr;           /* Adam step, cosine-decayed to lr/50; <= 0: 2e-3 */
  double val_frac;     /* held-out share; <= 0: 0.05 */
  uint64_t seed;
  double c_lo_mM, c_hi_mM;  /* salt range of the poses; <= 0: the model's range */
} sf_train_opts;

typedef struct {
  double train_loss, val_loss;   /* MSE in standardised units */
  double val_rel_err_force;      /* |F_net - F| / |F| over held-out poses (scaled units) */
  double val_rel_err_torque;
  double train_rel_err_force;
  double seconds;
  int steps;
} sf_train_report;

int sf_far_model_train(sf_far_model* m, const sf_train_opts* opts, sf_train_report* report);

/* Uses model m for the far field of an uncut-Yukawa handle (NULL: exact far field again).
 * The handle copies what it needs; m may be destroyed afterwards. After every 100-step chunk
 * of sf_step the far field is audited against the exact one, per replica relative to the far
 * field the model replaces: if sqrt(sum |F_far,model - F_far,exact|^2 / sum |F_far,exact|^2)
 * over a replica's charged beads exceeds tol, that chunk is run again from its start with the
 * exact far field, which the handle keeps for the rest of its life (sf_far_stats reports it).
 * MD only. */
int sf_set_far_model(sf_handle* h, const sf_far_model* m, double tol);

/* drop = 1: leave the far field out (the uncut Yukawa truncated at far_rf by centre-bead
 * distance), the method without a far-field model; 0: back to the exact or model far field. */
int sf_far_drop(sf_handle* h, int drop);

/* Audit now: relative error of the model's far field against the exact one, as above (the
 * largest over the replicas). */
int sf_far_audit(sf_handle* h, double* rel_err);

typedef struct {
  int64_t audits;
  double last_err, max_err;  /* audit error relative to the far field (the criterion) */
  int using_model;       /* 1 while the model supplies the far field */
  int fell_back;         /* 1 if an audit exceeded the tolerance */
  double last_err_force, max_err_force;  /* the same difference relative to the total force,
                            sqrt(sum |F_far,model - F_far,exact|^2 / sum |F|^2) over a replica */
} sf_far_stats;

int sf_get_far_stats(sf_handle* h, sf_far_stats* out);

/* ---- assembly analysis -------------------------------------------------------------------- */

/* Oligomers at the current positions, the mass-spectrum analysis of src/analysis.cpp:143-186
 * (there run on ovito.lammpstrj frames, contact 2): two subunits of one replica belong to one
 * oligomer when any of their beads are closer than `contact` under the single-shift minimum
 * image (src/functions.h:42-56), closed transitively. Outputs, both [total subunits], replica r
 * in the slots of its subunits (replica-major), either may be NULL:
 *   label[s]          replica-local index of the smallest subunit in s's oligomer;
 *   mass_spectrum[.]  at replica r's first slot + k - 1: its oligomers of k subunits (the
 *                     reference's ms_bin row of one frame).
 * Returns the total number of oligomers over all replicas, or a negative error code. */
int sf_oligomers(sf_handle* h, double contact, int* label, int* mass_spectrum);

/* The same analysis with no device allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Outputs and return value are those of sf_oligomers.
 * The union-find buffers come from a fixed array in the library's device code (one per GPU,
 * about 6 MB, held until the process ends) instead of memory the handle allocates; handles of
 * more than 262,144 subunits in total are refused with SF_E_ARG. Every call on one GPU shares
 * that GPU's array (the call clears what it uses and waits for its results before returning),
 * so calls that overlap on one GPU, from several host threads, race. */
int sf_oligomers_noalloc(sf_handle* h, double contact, int* label, int* mass_spectrum);

/* ---- encapsulation big beads ------------------------------------------------------------ */

/* Places n big beads of diameter sigma among the subunit beads pos[n_subunits * n_beads][3] of
 * a cubic box, as generate_big_beads (src/initialize.cpp:566-680) does without a restart file:
 * srand(0), then for each big bead rand() positions (x, y, z) until one lies at least
 * (sigma_j + sigma) / 2 from every subunit bead j and sigma from every big bead placed before
 * (minimum image by rounding). out[n][3]. Host code only (the C library's rand(), so the same
 * stream as the reference on the same C library). Fails with SF_E_ARG when a bead needs
 * max_attempts tries (the reference exits there). */
int sf_big_beads_place(const sf_template* t, int n_subunits, const double* pos, double box, int n,
                       double sigma, unsigned max_attempts, double* out);

/* Immobile big beads (replaces `-p` encapsulation, src/md.cpp:183-185, 242, 444): n beads at
 * xyz[n][3] of diameter sigma, charge `charge` (e) and hard-core distance shc, acting on every
 * subunit bead with the forces of forceCalculation_bigbead (src/forces.cpp:174-236) and the
 * energies of update_LJ_ES_energies_bigbeads (src/energies.cpp:96-136, added to the beads' ne and
 * ce in sf_energies). In BD their noise is drawn after the subunit beads' every step and not
 * used (src/md.cpp:482-491), so the noise stream matches the reference's. Single-replica
 * handles only; n = 0 removes them. The new set applies from the next sf_set_state (sf_step
 * refuses to run before it); sf_forces and sf_energies use it at once. */
int sf_set_big_beads(sf_handle* h, int n, const double* xyz, double sigma, double charge, double shc);

/* Energies of the big beads at the current positions: out[2] = {sum of ce, sum of ne}
 * (the reference's big_cenergy and big_ljenergy, src/md.cpp:606-613). */
int sf_big_energies(sf_handle* h, double* out);

const char* sf_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SOUFFLE_GPU_H */
