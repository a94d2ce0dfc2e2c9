/*
 * pfpa.h -- C-ABI of the MI355X-native KKS phase-field step (libpfpa.so).
 *
 * Plain pointers and sizes only.  Field pointers are DEVICE pointers to
 * (nx x ny) row-major arrays with the reference's one-cell ghost ring
 * (nx = Nx + 2, idx = nx*y + x), as src/type.h:58-107 lays them out.
 * `stream` is a hipStream_t (NULL = the default stream).  Every call returns
 * PF_OK or an error code; pf_last_error() describes the last failure of the
 * calling thread.  Calls are asynchronous on `stream` unless noted.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the reference repository root).
 */
#ifndef PFPA_H
#define PFPA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PF_OK 0
#define PF_ERR_ARG 1         /* invalid argument (sizes, null pointers, mode) */
#define PF_ERR_HIP 2         /* HIP runtime error */
#define PF_ERR_OVERFLOW 3    /* nucleation event list overflowed */
#define PF_ERR_NO_DEVICE 4   /* no GPU visible */

#define PF_BC_NOFLUX 0       /* ghost := adjacent interior (src/discretization.cu:360-394) */
#define PF_BC_PERIODIC 1

/* Kinetic and discretisation parameters of a step: the values init_cuda
 * uploads to __constant__ memory (src/data.cu:70-80) plus the dt and alpha
 * that device_evolution takes (src/discretization.cuh:89-93). */
typedef struct pf_params {
	double dt;
	double DCr[2];    /* d_DCr: row of the constant diffusivity matrix */
	double DNb[2];    /* d_DNb */
	double kappa[2];  /* d_Kapp: gradient coefficients (delta, Laves) */
	double omega[2];  /* d_Omeg: well heights */
	double Lmob[2];   /* d_Lmob: Allen-Cahn mobilities */
	double alpha;     /* three-phase coefficient Omega */
	double mask[9];   /* d_mask: 3x3 Laplacian mask, row-major (src/mesh.cpp:187-196) */
} pf_params;

/* Arguments of device_nucleation (src/discretization.cuh:105-111), plus the
 * counter-based random stream that replaces the per-cell cuRAND state. */
typedef struct pf_nuc_params {
	double sigma_del, sigma_lav;  /* interfacial energies (J/m^2) */
	double unit_a;                /* lattice constant (m) */
	double ifce_width;            /* accepted, unused (as in the reference) */
	double dx, dy, dz;            /* mesh spacing; the pass uses dx*nx, dy*ny (src/discretization.cu:1206) */
	double dt;                    /* nuc_interval * dt (src/main.cpp:314) */
	uint64_t seed;                /* Philox4x32-10 key */
	uint64_t call;                /* stream position: index of this nucleation pass */
	uint32_t instance;            /* simulation id folded into the counter */
	uint32_t reserved;
} pf_nuc_params;

typedef struct pf_ctx pf_ctx;

/* Context: device scratch (nucleation event list, reduction partials) and
 * tile settings.  Replaces the allocation half of init_cuda / free_cuda
 * (src/data.cu:10-154). */
int pf_ctx_create(pf_ctx** ctx);
void pf_ctx_destroy(pf_ctx* ctx);
const char* pf_last_error(void);
int pf_device_count(void);
/* Tile of the marching step kernel: output rows per block (0 = heuristic),
 * XCD-aware block order on/off, rows of loads kept in flight (2, 3 or 6;
 * 0 = default 3).  Results do not depend on the tile. */
int pf_ctx_set_tile(pf_ctx* ctx, int rows, int xcd_remap, int depth);
/* Composition offset of FP32 storage (no counterpart in the reference, whose
 * fields are FP64): on this context the conc_Cr / conc_Nb arrays of every
 * *_f32 entry point (step, multi-step, nucleation, summarize, energy) hold
 * x - x_cr and x - x_nb.  With the offsets near the mean composition the FP32
 * values are ~10x smaller, so each step's small composition increment keeps
 * 3-4 more bits.  Default 0 (arrays hold x); FP64 entry points ignore it.
 * Returns PF_ERR_ARG for a non-finite offset. */
int pf_ctx_set_storage_offset(pf_ctx* ctx, double x_cr, double x_nb);

/* One fused explicit-Euler step, old -> new.  Replaces the per-step sequence
 * device_boundaries, device_fictitious, fictitious_boundaries,
 * device_laplacian, device_laplacian_boundaries, device_evolution
 * (src/main.cpp:289-305; src/discretization.cu:128-173,396-548,780-1020).
 * nuc != NULL fuses the nucleation pass on the new fields (src/main.cpp:307-316).
 * Interior cells and the ghost ring of the outputs are written. */
int pf_step_f64(pf_ctx* ctx, int nx, int ny, int bc, const pf_params* prm,
                const double* conc_Cr_old, const double* conc_Nb_old,
                const double* phi_del_old, const double* phi_lav_old,
                double* conc_Cr_new, double* conc_Nb_new,
                double* phi_del_new, double* phi_lav_new,
                const pf_nuc_params* nuc, void* stream);

/* Precision form: fields stored in FP32, arithmetic (KKS solve included) in FP64. */
int pf_step_f32(pf_ctx* ctx, int nx, int ny, int bc, const pf_params* prm,
                const float* conc_Cr_old, const float* conc_Nb_old,
                const float* phi_del_old, const float* phi_lav_old,
                float* conc_Cr_new, float* conc_Nb_new,
                float* phi_del_new, float* phi_lav_new,
                const pf_nuc_params* nuc, void* stream);

/* Two fused steps, old -> new, without the intermediate level ever reaching
 * memory (temporal blocking).  The result equals two pf_step_* calls without
 * nucleation, the first writing a scratch level that the second reads: the
 * reference's step sequence (src/main.cpp:289-305, 318-321) run twice.
 * Symmetric five-point masks only; the interior must be at least 3x3.  Nucleation steps
 * (src/main.cpp:307-316) go through pf_step_* one at a time. */
int pf_step2_f64(pf_ctx* ctx, int nx, int ny, int bc, const pf_params* prm,
                 const double* conc_Cr_old, const double* conc_Nb_old,
                 const double* phi_del_old, const double* phi_lav_old,
                 double* conc_Cr_new, double* conc_Nb_new,
                 double* phi_del_new, double* phi_lav_new, void* stream);
int pf_step2_f32(pf_ctx* ctx, int nx, int ny, int bc, const pf_params* prm,
                 const float* conc_Cr_old, const float* conc_Nb_old,
                 const float* phi_del_old, const float* phi_lav_old,
                 float* conc_Cr_new, float* conc_Nb_new,
                 float* phi_del_new, float* phi_lav_new, void* stream);
/* Three fused steps, old -> new: the same as three pf_step_* calls without
 * nucleation (src/main.cpp:289-305, 318-321 run three times).  Same
 * restrictions as pf_step2_*. */
int pf_step3_f64(pf_ctx* ctx, int nx, int ny, int bc, const pf_params* prm,
                 const double* conc_Cr_old, const double* conc_Nb_old,
                 const double* phi_del_old, const double* phi_lav_old,
                 double* conc_Cr_new, double* conc_Nb_new,
                 double* phi_del_new, double* phi_lav_new, void* stream);
int pf_step3_f32(pf_ctx* ctx, int nx, int ny, int bc, const pf_params* prm,
                 const float* conc_Cr_old, const float* conc_Nb_old,
                 const float* phi_del_old, const float* phi_lav_old,
                 float* conc_Cr_new, float* conc_Nb_new,
                 float* phi_del_new, float* phi_lav_new, void* stream);
/* Four fused steps, old -> new: the same as four pf_step_* calls without
 * nucleation (src/main.cpp:289-305, 318-321 run four times).  Same
 * restrictions as pf_step2_*, and the interior must be at least 4x4. */
int pf_step4_f64(pf_ctx* ctx, int nx, int ny, int bc, const pf_params* prm,
                 const double* conc_Cr_old, const double* conc_Nb_old,
                 const double* phi_del_old, const double* phi_lav_old,
                 double* conc_Cr_new, double* conc_Nb_new,
                 double* phi_del_new, double* phi_lav_new, void* stream);
int pf_step4_f32(pf_ctx* ctx, int nx, int ny, int bc, const pf_params* prm,
                 const float* conc_Cr_old, const float* conc_Nb_old,
                 const float* phi_del_old, const float* phi_lav_old,
                 float* conc_Cr_new, float* conc_Nb_new,
                 float* phi_del_new, float* phi_lav_new, void* stream);

/* Standalone nucleation pass in place on the given fields.  Replaces
 * device_nucleation (src/discretization.cu:1179-1209). */
int pf_nucleation_f64(pf_ctx* ctx, int nx, int ny, double* conc_Cr, double* conc_Nb,
                      double* phi_del, double* phi_lav, const pf_nuc_params* nuc, void* stream);
int pf_nucleation_f32(pf_ctx* ctx, int nx, int ny, float* conc_Cr, float* conc_Nb,
                      float* phi_del, float* phi_lav, const pf_nuc_params* nuc, void* stream);

/* Synchronous: events recorded since the context was created that needed a
 * deferred (radius >= 1 cell) embed, and whether the event list overflowed. */
int pf_nucleation_status(pf_ctx* ctx, uint64_t* wide_events, int* overflow);

/* Means over interior cells of {x_Cr, x_Nb, gamma, delta, Laves} into the
 * device array out5[5]: deterministic compensated FP64 sums.  Replaces
 * summarize_fields (src/alloy625.cpp:811-863). */
int pf_summarize_f64(pf_ctx* ctx, int nx, int ny, const double* conc_Cr, const double* conc_Nb,
                     const double* phi_del, const double* phi_lav, double* out5, void* stream);
int pf_summarize_f32(pf_ctx* ctx, int nx, int ny, const float* conc_Cr, const float* conc_Nb,
                     const float* phi_del, const float* phi_lav, double* out5, void* stream);

/* pf_summarize_f64 without a context, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc); no counterpart in the reference.  It
 * allocates nothing: the reduction partials live in a fixed 80 KiB device
 * array the library holds on every GPU it runs on until the process ends
 * (looked up on the first call on each device).  Same result, bit for bit,
 * as pf_summarize_f64 (compositions as stored: no storage offset).  Grids up
 * to nx * ny < 2^31 cells.  Calls that overlap on one GPU (other streams or
 * threads) race on the shared array; calls in stream order do not. */
int pf_summarize_f64_noalloc(int nx, int ny, const double* conc_Cr, const double* conc_Nb,
                             const double* phi_del, const double* phi_lav, double* out5,
                             void* stream);

/* Free energy of the fields: src/alloy625.cpp:752-784 (gibbs) + :786-806
 * (maskedgradient) summed as in summarize_energy (:865-919), with the
 * per-cell densities times dV = dx*dy; compensated FP64 sums in a fixed order.
 * kappa, omega (2 each, delta then Laves) and Omega are the values of
 * src/parameters.h:36-48; bc as for the step.  out3 (DEVICE, 3 doubles) =
 * {total, chemical, gradient}; chem and grad (DEVICE nx*ny doubles, may be
 * NULL) receive the per-cell energies at interior cells, as the driver's
 * chem_nrg / grad_nrg host arrays do. */
int pf_energy_f64(pf_ctx* ctx, int nx, int ny, int bc, double dx, double dy,
                  const double* kappa, const double* omega, double Omega,
                  const double* conc_Cr, const double* conc_Nb, const double* phi_del,
                  const double* phi_lav, double* chem, double* grad, double* out3, void* stream);
int pf_energy_f32(pf_ctx* ctx, int nx, int ny, int bc, double dx, double dy,
                  const double* kappa, const double* omega, double Omega,
                  const float* conc_Cr, const float* conc_Nb, const float* phi_del,
                  const float* phi_lav, double* chem, double* grad, double* out3, void* stream);

/* x_Ni = 1 - x_Cr - x_Nb and phi = p(phi_del) + p(phi_lav) over every cell.
 * Replaces dataviz_kernel (src/discretization.cu:1211-1229). */
int pf_dataviz_f64(int nx, int ny, const double* conc_Cr, const double* conc_Nb,
                   const double* phi_del, const double* phi_lav, double* conc_Ni,
                   double* phi, void* stream);

/* Pointwise device thermodynamics over n inputs (device arrays), the
 * __device__ API of src/parabola625.cuh:13-95 and src/nucleation.cuh:8-20:
 *   what 0: in x[n]                         -> out {p, pPrime}[n][2]
 *   what 1: in {xCr,xNb,pDel,pGam,pLav}[n]  -> out fict {gam,del,lav}x{Cr,Nb}[n][6]
 *   what 2: in {xCr,xNb}[n]                 -> out {g x3, dg x6, D_CrCr,D_CrNb,D_NbCr,D_NbNb}[n][13]
 *   what 3: in nucleation_probability_sphere arguments[n][10] -> out {Rstar, P}[n][2] */
int pf_thermo_eval(int what, int n, const double* in, double* out, void* stream);

/* Philox4x32-10 over n {ctr0, ctr1, ctr2, ctr3, key0, key1} tuples -> n x 4
 * words (device arrays).  Replaces the cuRAND XORWOW state of
 * init_prng_kernel (src/discretization.cu:1022-1044). */
int pf_philox_eval(int n, const uint32_t* in, uint32_t* out, void* stream);

/* A loop of plain FP64 steps over independent grids, one launch each, issued
 * from C: the "loop of plain launches" the batched form is compared with. */
typedef struct pf_step_item {
	int nx, ny, bc;
	const pf_params* prm;
	const double* old_fields[4];   /* conc_Cr, conc_Nb, phi_del, phi_lav */
	double* new_fields[4];
} pf_step_item;
int pf_step_loop_f64(pf_ctx* ctx, int n, const pf_step_item* items, void* stream);
/* The same loop with pf_step2_f64 / pf_step3_f64 / pf_step4_f64 (two / three / four steps per item
 * and launch). */
int pf_step2_loop_f64(pf_ctx* ctx, int n, const pf_step_item* items, void* stream);
int pf_step3_loop_f64(pf_ctx* ctx, int n, const pf_step_item* items, void* stream);
int pf_step4_loop_f64(pf_ctx* ctx, int n, const pf_step_item* items, void* stream);

/* Variable-mobility Cahn-Hilliard divergence: the face-averaged
 * discrete_laplacian sums of chemical_convolution_Cr_kernel /
 * chemical_convolution_Nb_kernel (src/discretization.cu:605-778; in the
 * reference they compile but their launch is commented out at :809-835, and
 * the mobility kernels that would fill the arrays, :175-358, call d_M_*
 * functions no file defines, so the mobilities are caller inputs).  All
 * arrays are DEVICE nx x ny FP64 with the ghost ring, filled by the caller
 * (no-flux copy or periodic wrap).  For component i (Cr, Nb) and every cell,
 * ghost ring included:
 *   div_i = sum_{t<6} [D_x + D_y](mob_i[t], conc[t]) + sum_{k<2} D_x(mob_i[6+k], phi_k),
 *   D_x(M, c) = ((M[x+1] + M[x]) (c[x+1] - c[x]) - (M[x] + M[x-1]) (c[x] - c[x-1])) / (2 dx^2)
 * (D_y likewise with dy), where the neighbour beyond the array edge is the
 * cell itself (the reference's index clamp).  The phase-field terms are
 * differenced in x only, as in the reference (:688-695, :769-776).
 * div_Cr / div_Nb must not overlap any input. */
typedef struct pf_mobility_fields {
	const double* conc[6];    /* x_gam_Cr, x_gam_Nb, x_del_Cr, x_del_Nb, x_lav_Cr, x_lav_Nb */
	const double* mob_Cr[8];  /* gam_CrCr, gam_CrNb, del_CrCr, del_CrNb, lav_CrCr, lav_CrNb,
	                             phi_del_Cr, phi_lav_Cr */
	const double* mob_Nb[8];  /* gam_NbCr, gam_NbNb, del_NbCr, del_NbNb, lav_NbCr, lav_NbNb,
	                             phi_del_Nb, phi_lav_Nb */
	const double* phi_del;
	const double* phi_lav;
} pf_mobility_fields;
int pf_mobility_divergence_f64(int nx, int ny, double dx, double dy, const pf_mobility_fields* f,
                               double* div_Cr, double* div_Nb, void* stream);

/* Heuristic rows-per-block the step kernel uses for an nx x ny grid. */
int pf_default_rows(int nx, int ny);
/* Tile of the two-step (three-, four-step) kernel: rows per chunk (0 =
 * heuristic) and prefetch depth (2, 3, 6; 0 = 3).  pf_default_rows2 (3, 4) is
 * its heuristic for storage PF_STORAGE_F64 or PF_STORAGE_F32. */
int pf_ctx_set_tile2(pf_ctx* ctx, int rows, int depth);
int pf_default_rows2(int nx, int ny, int storage);
int pf_ctx_set_tile3(pf_ctx* ctx, int rows, int depth);
int pf_default_rows3(int nx, int ny, int storage);
int pf_ctx_set_tile4(pf_ctx* ctx, int rows, int depth);
int pf_default_rows4(int nx, int ny, int storage);

/* ------------------------------------------------------------------------
 * Batched form: many independent simulations of uneven size advanced by one
 * persistent launch per step (no counterpart in the reference, which runs one
 * grid per process).  Jobs live in one device pool; their fields stay
 * resident between calls.  A job's result does not depend on the other jobs
 * in the batch.
 * ------------------------------------------------------------------------ */
#define PF_STORAGE_F64 0
#define PF_STORAGE_F32 1   /* fields in FP32, arithmetic and reductions in FP64 */

typedef struct pf_job_desc {
	int Nx, Ny;            /* interior size; arrays are (Nx+2) x (Ny+2) with the ghost ring */
	int bc;                /* PF_BC_NOFLUX or PF_BC_PERIODIC */
	int nucleation;        /* 1 = nucleation pass every nuc_every steps (step % nuc_every == 0) */
	double dx;             /* mesh spacing (m), used by the nucleation pass */
	pf_params prm;         /* dt, diffusivity, kappa, omega, mobility, alpha, symmetric five-point mask */
	uint64_t steps;        /* steps to run */
	uint64_t seed;         /* Philox key */
	uint32_t instance;     /* folded into the random counter: unique per simulation */
	int nuc_every;
	double sigma_del, sigma_lav, unit_a;
	double x_offset[2];    /* FP32 storage: the job's Cr / Nb arrays hold x - x_offset
	                          (pf_ctx_set_storage_offset); 0 = absolute.  FP64: must be 0 */
} pf_job_desc;

typedef struct pf_batch pf_batch;

/* HBM a job needs (two buffer sets of four planes). */
size_t pf_batch_job_bytes(int storage, int Nx, int Ny);
int pf_batch_create(pf_batch** batch, int storage, size_t pool_bytes, int max_jobs);
void pf_batch_destroy(pf_batch* batch);
/* Largest contiguous free block of the pool, in bytes. */
size_t pf_batch_free_bytes(pf_batch* batch);
/* Rows per strip task and prefetch depth (2, 3 or 6) of the batched step. */
int pf_batch_set_tile(pf_batch* batch, int rows, int depth);
/* Steps per launch: 1, 2 or 3 (the marches of pf_step2_* / pf_step3_*).  A
 * launch advances each active job by up to that many steps: fewer where fewer
 * remain or where a nucleation step falls before the last of them (the
 * nucleation pass runs after the launch's last step). */
int pf_batch_set_launch_steps(pf_batch* batch, int steps);
/* Admit a job; fields are HOST arrays of the batch's storage type, (Nx+2)*(Ny+2) each
 * (compositions less desc->x_offset).
 * Returns PF_ERR_OVERFLOW when the pool has no room (the caller may retry later). */
int pf_batch_add(pf_batch* batch, const pf_job_desc* desc, const void* conc_Cr,
                 const void* conc_Nb, const void* phi_del, const void* phi_lav, int* job_id);
/* Retire a job and return its memory to the pool. */
int pf_batch_remove(pf_batch* batch, int job_id);
/* Publish admissions/removals to the device tables (done by pf_batch_run too). */
int pf_batch_commit(pf_batch* batch);
/* Enqueue `launches` step launches; each advances every active job by its
 * steps per launch (one, or two: pf_batch_set_launch_steps), and a job stops
 * at its target.  With launches_per_graph > 1 the launches are replayed from a
 * captured hipGraph of that many launches.  Asynchronous. */
int pf_batch_run(pf_batch* batch, int launches, int launches_per_graph);
int pf_batch_sync(pf_batch* batch);
/* Synchronous job state. */
int pf_batch_poll(pf_batch* batch, int job_id, uint64_t* steps_done, int* active);
int pf_batch_poll_all(pf_batch* batch, int capacity, int* job_ids, uint64_t* steps_done,
                      int* active, int* count);
/* Synchronous copy of a job's current fields to host arrays (NULL skips a field),
 * as stored (compositions less the job's x_offset). */
int pf_batch_read(pf_batch* batch, int job_id, void* conc_Cr, void* conc_Nb, void* phi_del,
                  void* phi_lav);
/* Synchronous per-job means {x_Cr, x_Nb, gamma, delta, Laves} (compensated FP64;
 * compositions absolute). */
int pf_batch_summarize(pf_batch* batch, int job_id, double* out5);
/* Fails if any nucleation event needed an embed wider than one cell. */
int pf_batch_status(pf_batch* batch, uint64_t* wide_events);
/* Strip tasks per step of the active, unheld jobs (for throughput accounting). */
int pf_batch_waves(pf_batch* batch);
/* Hold (hold = 1) or release (0) a job: a held job keeps its fields and step
 * count but takes no waves and does not advance until released, so the
 * launches run only the other jobs (a scheduler's lever: e.g. run the jobs
 * with the least remaining work first).  Takes effect from the next
 * pf_batch_run; a job's results do not depend on when it was held. */
int pf_batch_hold(pf_batch* batch, int job_id, int hold);
/* The HIP stream (hipStream_t) every launch, copy and reduction of the batch
 * is enqueued on, so callers can record events on it or order their own work
 * after the batch's. */
void* pf_batch_stream(pf_batch* batch);

/* Host-side evaluation of the same pointwise thermodynamics (no GPU needed).
 * Replace the host functions of thermo/parabola625.h the driver calls:
 * p (src/main.cpp:340), fict_gam_Cr/Nb (src/main.cpp:346-347) and
 * D_CrCr/D_CrNb/D_NbCr/D_NbNb (src/alloy625.cpp:91-95). */
double pf_host_p(double x);
double pf_host_pPrime(double x);
double pf_host_g(int phase, double xCr, double xNb);           /* phase 0 gamma, 1 delta, 2 Laves */
double pf_host_D(int i, int j, double xCr, double xNb);        /* i, j: 0 Cr, 1 Nb */
void pf_host_fict(double xCr, double xNb, double pDel, double pGam, double pLav, double* out6);
double pf_host_xe(int phase, int comp);
void pf_host_nucleation_probability(double xCr, double xNb, double dG, double DCrCr,
                                    double DNbNb, double sigma, double Vatom, double n_gam,
                                    double dV, double dt, double* Rstar, double* P);
/* Largest eigenvalue over the three phases of D * H_gamma^-1 * H_a: the
 * effective diffusivity that bounds the explicit step inside each phase. */
double pf_host_max_effective_diffusivity(const double* DCr, const double* DNb);
/* The per-step coefficients the step kernels run on, folded on the host from
 * the device_evolution constants (src/discretization.cu:859-863, 888-942); no
 * counterpart in the reference, which evaluates them term by term per cell.
 * out22: q_Cr = dt D_Cr.(xe_g + A_g mu) as the coefficients of (mu0, mu1, 1),
 * the same for q_Nb, then per phase (delta, Laves) the quadratic form
 * 30 dt L dOmega = g0 mu0^2 + g1 mu0 mu1 + g2 mu1^2 + g3 mu0 + g4 mu1, then
 * 2 omega dt L, 2 alpha dt L and kappa dt L per phase.  For a symmetric
 * five-point mask (zero corners, mask[1] = mask[7] != 0, mask[3] = mask[5])
 * the q coefficients and kappa dt L carry the factor mask[1], by which the
 * kernels' copy of the mask is divided; for any other mask the factor is 1. */
int pf_host_step_coefficients(const pf_params* prm, double* out22);

#ifdef __cplusplus
}
#endif

#endif /* PFPA_H */
