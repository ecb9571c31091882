/*
 * etoheom.h -- C ABI of the MI355X-native ETO-HEOM propagator (libetoheom.so).
 *
 * The reference exposes no FFI for this path; its library boundary is a pair of
 * C++ free functions over a mutable `param` object:
 *   void twoD_spectrum_solver(param&, gsl_matrix_complex* H, int nv1, int nv2, int nv3,
 *                             vector<gsl_complex>& polarization)   etoheom-cpu/lib/dynamics.h:23
 *   void dynamics_solver(param&, gsl_matrix_complex* H, vector<int>& sites,
 *                        vector<vector<double>>& population)       etoheom-cpu/lib/dynamics.h:21
 * plus the key-file parser param::param_2des / param::param_dynamics (param.cpp:18-348),
 * the hierarchy builder construct_ADO_set (utilize.cpp:26-36), construct_Hal
 * (dynamics.cpp:3-51) and the dipole assembly of polar.cpp:17-101.  Each entry point
 * below names the one it replaces.
 *
 * Conventions: complex numbers are interleaved (re, im) doubles; matrices are
 * row-major N*N (the gsl_matrix_complex layout); a hierarchy state is M such
 * matrices in the reference DFS ADO order.  Every function returns EH_OK (0) or a
 * negative EH_E* code and never exits the process; eh_last_error() gives the
 * message of the last failure on the calling thread.  Contexts are not shared
 * between threads; distinct contexts may be used concurrently.
 */
#ifndef ETOHEOM_H
#define ETOHEOM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EH_OK 0
#define EH_EINVAL (-1)   /* bad argument (NULL, size, |alpha| = 0, ...) */
#define EH_EPARSE (-2)   /* key file missing or malformed */
#define EH_ESHAPE (-3)   /* inconsistent sizes */
#define EH_ENAN (-4)     /* non-finite state produced */
#define EH_EHIP (-5)     /* HIP runtime error / no device */
#define EH_ENOMEM (-6)   /* allocation failed */
#define EH_ERANGE (-7)   /* index out of range, output capacity too small */
#define EH_EPULSE (-8)   /* PULSE block with a count other than 3 (param.cpp:227-230) */

#define EH_MODE_2DES 0      /* param_2des: N = n + (n-1)(n-2)/2, S_j = |j+1><j+1| */
#define EH_MODE_DYNAMICS 1  /* param_dynamics: N = n, S_j = |j><j| */

typedef struct eh_ctx eh_ctx;
typedef struct eh_problem eh_problem;
typedef struct eh_batch eh_batch;

/* Plain-data view of one propagation problem (the numbers `param` carries,
 * param.h:38-82).  Pointers are borrowed for the duration of a call only. */
typedef struct eh_system {
    int N;                 /* sys_size */
    int K;                 /* number of baths (sites) */
    int Km;                /* modes per bath */
    int L;                 /* hierarchy depth */
    int site_offset;       /* 1 (2DES) or 0 (dynamics): bath j couples |j+site_offset> */
    int reserved;
    const double* alpha;   /* K*Km complex, [j*Km + k] */
    const double* alpha_t; /* K*Km complex */
    const double* gamma;   /* K*Km complex */
    double pulse[3][4];    /* E0 [fs^-1], tau0 [fs], taup [fs], w0 [fs^-1] */
    float t_start, t_end, step_size, print_step;   /* float, as param.h:75-78 */
} eh_system;

/* ---- errors ------------------------------------------------------------ */
const char* eh_last_error(void);

/* ---- key files (param::param_2des / param_dynamics, param.cpp:18-348) ---- */
int eh_problem_from_key(const char* path, int mode, eh_problem** out);
int eh_problem_destroy(eh_problem* p);
/* sizes and scalars: out[0..] = {mode, single_size, N, K, Km, L, M, n_sample, seed} */
int eh_problem_dims(const eh_problem* p, int64_t* out9);
/* fill *sys with a view into p (valid while p lives) */
int eh_problem_system(const eh_problem* p, eh_system* sys);
/* named arrays: "Hal" (N*N complex, lambda-shifted), "Disorder" (N*N real sigma/h_bar),
 * "X_x","X_y","X_z" (N*N complex), "X" (4*N*N complex, after eh_problem_set_polar),
 * "angle" (4 radians), "lambda" (K), "beta" (1).  Returns the number of doubles
 * written, or a negative code (EH_ERANGE if cap is too small). */
int64_t eh_problem_get(const eh_problem* p, const char* name, double* out, int64_t cap);
/* polar_mat_set (polar.cpp:17-29): polar[3][4] from the POLARIZATION angles */
int eh_problem_polar_default(const eh_problem* p, double* polar_3x4);
/* compute_pulse_interaction (polar.cpp:74-101): X[i] = sum_c polar[c][i] X_c */
int eh_problem_set_polar(eh_problem* p, const double* polar_3x4);
/* random rotation of polar (polar.cpp:31-72) given the two unit directions c and a */
int eh_polar_rotate(const double* c3, const double* a3, double* polar_3x4);
/* construct_Hal (dynamics.cpp:3-51) with the disorder draw dH (single_size^2
 * real, NULL = none): H = Hal + dH, one-exciton diagonal sorted ascending,
 * double-exciton block rebuilt.  H_out is N*N complex. */
int eh_construct_hal(const eh_problem* p, const double* dH, double* H_out);

/* The reference sample loop's random draws (CPU_2DES.cpp:50-60): per sample the
 * realised Hamiltonian (construct_Hal with gsl_ran_gaussian disorder, N*N complex)
 * and dipole operators (polar_mat_set + polar_mat_ranrot + compute_pulse_interaction,
 * 4*N*N complex), from a restatement of GSL's mt19937 / gaussian / dir_3d. */
int eh_problem_samples(const eh_problem* p, int n_sample, double* H_out, double* X_out);
/* first `count` raw 32-bit outputs of that generator for `seed` */
int eh_rng_stream(unsigned long seed, int count, uint32_t* out);

/* ---- hierarchy (Build_ADO / Build_ADO_map, utilize.cpp:3-24) -------------- */
int64_t eh_hierarchy_size(int d, int L);
/* multi-indices in DFS order (M*d bytes); up[M*d] / dn[M*d]: index of n+e_p / n-e_p or -1 */
int eh_hierarchy_build(int d, int L, uint8_t* ado_out, int32_t* up_out, int32_t* dn_out);

/* ---- GPU context ---------------------------------------------------------- */
int eh_ctx_create(int device, eh_ctx** out);
int eh_ctx_destroy(eh_ctx* ctx);
/* device name / CU count / LDS per block, for logs */
int eh_ctx_info(const eh_ctx* ctx, char* name, int cap, int* n_cu, int* lds_per_block);
/* free and total device memory (hipMemGetInfo), for sizing batches with eh_batch_plan.
 * No reference counterpart: the reference holds one propagation per process. */
int eh_ctx_mem_info(eh_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes);

/* ---- one propagation (twoD_spectrum_solver, dynamics.cpp:298-432) ---------
 * X4 = X[0..3] (4*N*N complex).  P_out receives one complex per readout point
 * (t >= tau3 and i % print_step == 0, read before step i).  rho_out: final state
 * M*N*N complex or NULL. */
int eh_2des_solve(eh_ctx* ctx, const eh_system* sys, const double* H, const double* X4,
                  int nv1, int nv2, int nv3, double* P_out, int64_t cap, int64_t* n_points,
                  double* rho_out);
/* ---- one propagation that allocates no device memory ------------------------
 * No counterpart in the reference.  For callers that cannot allocate device memory
 * (serving loops, harnesses that forbid hipMalloc).  Same arguments, checks and
 * results as eh_2des_solve (bit for bit: the same kernels, launched one by one),
 * computed in one static device array of EH_SCRATCH_BYTES per GPU that the library
 * holds from eh_ctx_create until the process ends; it calls no hipMalloc or hipFree and
 * creates no stream (the null stream), and returns after the outputs are copied back.
 * EH_ERANGE when the propagation's four state buffers and tables exceed the array
 * (about 4 M N^2 complex: the dimer, trimer and tetramer hierarchies fit; the
 * 3060-ADO N = 29 heptamer, about 170 MB, needs eh_2des_solve).  Every call on one
 * GPU shares the array and overwrites it; a per-device lock held for the whole call
 * makes calls from several threads of one process run one after another instead of
 * racing on it. */
#define EH_SCRATCH_BYTES (64LL << 20)
int eh_2des_solve_noalloc(eh_ctx* ctx, const eh_system* sys, const double* H, const double* X4,
                          int nv1, int nv2, int nv3, double* P_out, int64_t cap, int64_t* n_points,
                          double* rho_out);
/* dynamics_solver (dynamics.cpp:202-295): pop_out[s*cap + t] = Re rho_0(site_s-1, site_s-1) */
int eh_dynamics_solve(eh_ctx* ctx, const eh_system* sys, const double* H, const int* sites,
                      int n_sites, double* pop_out, int64_t cap, int64_t* n_points,
                      double* rho_out);

/* ---- whole drivers ------------------------------------------------------
 * main of CPU_2DES (etoheom-cpu/src/CPU_2DES.cpp:14-96): n_sample samples, three
 * propagations (1,1,1), (1,1,0), (1,0,1) each, P = mean_s(p111 - p110 - p101);
 * all 3*n_sample propagations run as one batch. */
int eh_2des_key_run(eh_ctx* ctx, const char* path, double* P_out, int64_t cap, int64_t* n_points);
/* main of CPU_dynamics (etoheom-cpu/src/CPU_dynamics.cpp:14-49) for the given sites;
 * n_points is the reference's population length (t_end - t_start)/print_step + 1
 * evaluated in float, readouts beyond it are dropped and missing ones stay 0 */
int eh_dynamics_key_run(eh_ctx* ctx, const char* path, const int* sites, int n_sites, double* pop_out,
                        int64_t cap, int64_t* n_points);

/* ---- batches of independent propagations (replaces the reference's
 * one-process-per-(tau,T) farming, script/cpu-script/submit_jobs.sh:49-66) -- */
typedef struct eh_instance {
    const eh_system* sys;
    const double* H;       /* N*N complex */
    const double* X4;      /* 4*N*N complex (2DES) or NULL (dynamics) */
    int nv[3];
    int mode;              /* EH_MODE_2DES or EH_MODE_DYNAMICS */
    const int* sites;      /* dynamics readout sites (1-based) */
    int n_sites;
} eh_instance;

int eh_batch_create(eh_ctx* ctx, int n, const eh_instance* inst, eh_batch** out);
/* Host-only (no device call): runs every check eh_batch_create makes on the
 * instances (sizes, at most 2^31 - 1 state elements per instance, time grid) and
 * returns out6 = {total ADOs, total state elements, max steps, device bytes of the
 * forward batch, device bytes with the adjoint enabled at `checkpoint_every`
 * (checkpoint + replay), device bytes with the adjoint's whole trajectory kept}.
 * No reference counterpart (the reference allocates per ADO inside the solver,
 * utilize.cpp:30-35, dynamics.cpp:311-341). */
int eh_batch_plan(int n, const eh_instance* inst, int checkpoint_every, int64_t* out6);
int eh_batch_destroy(eh_batch* b);
/* reset every state to rho_0 = |0><0| and run all steps (asynchronous) */
int eh_batch_run(eh_batch* b);
int eh_batch_sync(eh_batch* b);
/* A propagation in pieces: the time loop of twoD_spectrum_solver / dynamics_solver
 * (dynamics.cpp:385-419, :257-289) split into segments.  eh_batch_start resets
 * every state to rho_0 = |0><0| (asynchronous); each eh_batch_advance runs the next
 * nsteps RK4 steps of every instance (a positive multiple of the graph chunk; steps
 * past an instance's last do nothing), between HIP events (eh_batch_elapsed_ms).
 * After all steps, outputs and states equal eh_batch_run's bit for bit.
 * eh_batch_set_chunk re-captures the forward graph with `chunk` steps per replay
 * (1 <= chunk <= max steps; before eh_batch_enable_adjoint). */
int eh_batch_set_chunk(eh_batch* b, int chunk);
int eh_batch_start(eh_batch* b);
int eh_batch_advance(eh_batch* b, int nsteps);
/* elapsed GPU time of the last run, from HIP events on the batch stream */
int eh_batch_elapsed_ms(eh_batch* b, double* ms);
/* readouts of instance i: 2DES -> complex per point; dynamics -> n_sites x points real.
 * EH_EINVAL before the batch has run (its readouts are undefined until then) */
int eh_batch_get_output(eh_batch* b, int i, double* out, int64_t cap, int64_t* n_points);
int eh_batch_get_rho(eh_batch* b, int i, double* rho_out);
/* sizes: out[0..] = {n_inst, total ADOs, total state elements, max steps, graph chunk} */
int eh_batch_dims(const eh_batch* b, int64_t* out5);
/* per-kernel timing pass: nsteps steps launched without the graph, HIP events
 * around each launch on the batch stream; out5 = mean ms of {prologue, stage 0..3} */
int eh_batch_profile(eh_batch* b, int nsteps, double* out5);
/* the plain per-launch path: same kernels and results as eh_batch_run, launched one
 * by one from the host step after step, no graph (the reference GPU path's
 * per-propagation launch pattern, etoheom-gpu/lib/HEOM_dynamics.cu:489-658, time
 * loop :566); for
 * comparison with the graph-replayed batch; asynchronous like eh_batch_run */
int eh_batch_run_eager(eh_batch* b);
/* autotuner: time `steps` RK4 steps of every launch candidate of every N class
 * (small-N tile sizes of 256/512/1024 state elements, MFMA / pipelined / generic
 * kernels, the small-N base one with the other state-store cache hint and the other load
 * placement), keep the fastest and
 * re-capture the graph; call before
 * eh_batch_enable_adjoint.  eh_batch_tuning_report writes one line per candidate
 * ("... ms_per_step=<t>", '*' marks the chosen one) into buf (NUL-terminated,
 * truncated to cap) and returns the full length. */
int eh_batch_autotune(eh_batch* b, int steps);
/* manual choice: group g launches its candidate c (numbered as in the report) */
int eh_batch_select(eh_batch* b, int group, int cand);
int64_t eh_batch_tuning_report(const eh_batch* b, char* buf, int64_t cap);
/* ---- gradients (discrete adjoint of the RK4 propagation) -------------------
 * No reference counterpart (the reference has no derivatives); this is the
 * "adjoint gradient" of the propagation path.  After enable, eh_batch_run stores
 * the state every `checkpoint_every` steps; eh_batch_backward takes
 * dP = dL/dRe P + i dL/dIm P for every readout (instances concatenated in
 * order, eh_batch_get_output's point count times readout operators each; n_dP
 * complex entries in all, EH_ESHAPE unless that is the batch's total), replays each checkpoint
 * segment and runs the transposed hierarchy backwards.  eh_batch_get_grad then
 * returns dL/dH (N*N), dL/dX (4*N*N, X[0..3] as passed in), dL/dalpha,
 * dL/dalpha_t, dL/dgamma (K*Km) of instance i, same convention, complex
 * interleaved; any output pointer may be NULL. */
int eh_batch_enable_adjoint(eh_batch* b, int checkpoint_every);
int eh_batch_backward(eh_batch* b, const double* dP, int64_t n_dP);
int eh_batch_get_grad(eh_batch* b, int i, double* dH, double* dX4, double* dalpha, double* dalpha_t,
                      double* dgamma);

/* ADO-updates per run (sum over instances of M * steps) */
int64_t eh_batch_ado_updates(const eh_batch* b);

#ifdef __cplusplus
}
#endif
#endif
