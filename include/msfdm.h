/*
 * msfdm.h - C ABI of the MI355X fluid-solid time stepper (libmsfdm.so).
 *
 * Drop-in replacement for the time loop of the reference's src/main.cu driven
 * by PmlFs::FluidSolid::update_RK4 (include/pml_fs_kernels.cuh:848-852).  The
 * reference has no library API: its boundaries are the process
 * (bin/main_debug <model.json> <out_root>, src/main.cu:58-77) and the
 * in-process MemManager + update_RK4 pair (include/elements.h:112-173).  Each
 * entry point below names the reference interface it replaces.
 *
 * Conventions: every function returns MSFDM_OK (0) or an error code and never
 * exits or throws; msfdm_last_error() returns the calling thread's last error
 * message.  Host buffers are caller-owned; device memory is owned by the
 * solver handle.  Arrays use the reference layout: float32, row-major
 * [rows][nx], x fastest (docs/io.md:5-12).
 */
#ifndef MSFDM_H
#define MSFDM_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MSFDM_OK = 0,
    MSFDM_EINVAL = 1,   /* invalid argument or inconsistent model */
    MSFDM_EIO = 2,      /* file cannot be opened, read or written */
    MSFDM_EPARSE = 3,   /* malformed JSON, missing key or wrong type */
    MSFDM_EDEVICE = 4,  /* HIP runtime error / no device */
    MSFDM_ESTATE = 5    /* call not valid in the current state */
};

/* Field identifiers for msfdm_get_field / msfdm_set_field.  U_fluid, U_solid and
 * the ADE memory sets of include/elements.h:6-26.  The fluid zeta memory
 * (MSFDM_FZ*) is driven on the fluid-solid CPML grid only (its top strip,
 * include/cpml_kernels.cuh:331-342, 399-413); with the pressure-release
 * surface the reference never drives it and it is not stored. */
enum {
    MSFDM_FVX = 0, MSFDM_FVZ, MSFDM_FP,
    MSFDM_FXVX, MSFDM_FXVZ, MSFDM_FXEX, MSFDM_FXEZ,
    MSFDM_SVX, MSFDM_SVZ, MSFDM_SSX, MSFDM_SSZ, MSFDM_SSXZ,
    MSFDM_SXVX, MSFDM_SXVZ, MSFDM_SXEX, MSFDM_SXEZ,
    MSFDM_SZVX, MSFDM_SZVZ, MSFDM_SZEX, MSFDM_SZEZ,
    MSFDM_FZVX, MSFDM_FZVZ, MSFDM_FZEX, MSFDM_FZEZ,
    MSFDM_NFIELD
};

/* Model arrays for msfdm_model_get_array. */
enum {
    MSFDM_A_K = 0, MSFDM_A_RHO_F, MSFDM_A_C11, MSFDM_A_C13, MSFDM_A_C33, MSFDM_A_C55,
    MSFDM_A_RHO_S, MSFDM_A_J, MSFDM_A_XI_X, MSFDM_A_XI_Z, MSFDM_A_ZE_X, MSFDM_A_ZE_Z,
    MSFDM_NARRAY
};

/* Scalar part of the model: the JSON keys read by MemManager::MemManager
 * (src/elements.cu:72-111).  Accepted: calc_type = 1 (FANDS, fluid over
 * solid) with boundary_type = 2 (PANDF: CFS-PML sides and bottom,
 * pressure-release sea surface; PmlFs::FluidSolid) or 1 (CPML: a top PML in
 * the water instead; Cpml::FluidSolid, include/cpml_kernels.cuh:237-770), and
 * calc_type = 0 (SOLID) with boundary_type = 2 (traction-free top row 3;
 * PmlFs::SolidOnly, include/pml_fs_kernels.cuh:8-307), 1 (PML on four sides;
 * Cpml::SolidOnly) or 0 (Dirichlet edges; include/dirichlet_kernels.cuh).
 * A solid-only model has nz_fluid = 0, no fluid arrays (K and rho_f may be
 * NULL) and metrics of nz_solid rows. */
typedef struct msfdm_desc {
    int nx, nz_fluid, nz_solid;
    int calc_type, boundary_type;
    float fpeak, dt;
    int nt, snapshot;
    int npd;
    float thickness;
    int npow;
    float vp_max, rc, beta0;
    int nsrc;
    const int *src_x; /* column of each shot (grid.src_pos.posx) */
    const int *src_z; /* row of each shot (grid.src_pos.posz): fluid-array row
                         (calc_type 1) or solid-array row (calc_type 0) */
} msfdm_desc;

typedef struct msfdm_model msfdm_model;
typedef struct msfdm_solver msfdm_solver;

/* Thread-local message of the last failing call ("" if none). */
const char *msfdm_last_error(void);
/* Library version string. */
const char *msfdm_version(void);

/* --- Host model (no GPU needed) ------------------------------------------ */

/* Parse <model.json> and read its 12 <f4 arrays (paths relative to the CWD).
 * Replaces JsonConfig + MemManager::MemManager's input half
 * (include/json_config.hpp:20-368, src/elements.cu:64-176). */
int msfdm_model_load_json(const char *json_path, msfdm_model **out);

/* Build a model from in-memory arrays (same shapes as the files). */
int msfdm_model_create(const msfdm_desc *desc, const float *K, const float *rho_f,
                       const float *C11, const float *C13, const float *C33, const float *C55,
                       const float *rho_s, const float *J, const float *xi_x, const float *xi_z,
                       const float *ze_x, const float *ze_z, msfdm_model **out);

/* Copy the scalars; desc->src_x/src_z point into the model (valid while it lives). */
int msfdm_model_get_desc(const msfdm_model *m, msfdm_desc *out);
/* Copy one array (MSFDM_A_*) into out ([rows][nx] floats). */
int msfdm_model_get_array(const msfdm_model *m, int which, float *out);
void msfdm_model_destroy(msfdm_model *m);

/* --- Device solver ------------------------------------------------------- */

/* Upload the model, build PML profiles, normals and the source table, and
 * allocate the wavefield on HIP device `device`.  Replaces the device half of
 * MemManager (src/elements.cu:557-905, 979-987) and grid_cp_h2d
 * (src/elements.cu:307-333). */
int msfdm_solver_create(const msfdm_model *m, int device, msfdm_solver **out);
void msfdm_solver_destroy(msfdm_solver *s);

/* Zero the wavefield, the records and the step counter and select shot isrc.
 * (The reference keeps the previous shot's wavefield, src/main.cu:94-136.) */
int msfdm_reset_shot(msfdm_solver *s, int isrc);

/* Advance nsteps RK4 steps (each = PmlFs::FluidSolid::update_RK4 with
 * rk4dire = it & 1, src/main.cu:136-176, plus the interface filter; on a
 * solid-only grid PmlFs::SolidOnly::update_RK4, src/pml_fs_kernels.cu:105-381),
 * sampling the receiver rows on the device after steps with it % gap == 0
 * (gap = (int)(1e-3/dt + 0.5), src/main.cu:92).  Asynchronous. */
int msfdm_run(msfdm_solver *s, int nsteps);
/* Wait for all queued work. */
int msfdm_synchronize(msfdm_solver *s);

/* Steps done since the last reset. */
int msfdm_step_index(const msfdm_solver *s, int *it);
/* Records available (= number of it with it % gap == 0 done so far). */
int msfdm_num_records(const msfdm_solver *s, int *n);
/* Copy records [first, first+count) of solid-row-0 vx, vz and fluid interface
 * row p (each count x nx floats, time-major): output_record, src/main.cu:245-295.
 * A solid-only grid records vz of the free-surface row 3 only (vz; vx and p
 * read as zeros), src/main.cu:297-322.  Any output pointer may be NULL. */
int msfdm_get_records(msfdm_solver *s, int first, int count, float *vx, float *vz, float *p);

/* Full-array field access in the reference layout (ADE fields expand to the
 * full [rows][nx] array, zero outside the PML strips).  U_cp_d2h,
 * src/elements.cu:335-343. */
int msfdm_get_field(msfdm_solver *s, int field, float *out);
int msfdm_set_field(msfdm_solver *s, int field, const float *in);

/* Evaluate one RK stage right-hand side k = BC(L_dir(U) + src) of the current
 * state (dir 0 = F, 1 = B; amp = Ricker amplitude added at the source) and
 * keep it for msfdm_get_rhs: linear_operate + fl_so_apply,
 * src/pml_fs_kernels.cu:533-660. */
int msfdm_eval_rhs(msfdm_solver *s, int dir, float amp);
int msfdm_get_rhs(msfdm_solver *s, int field, float *out);

/* Tile height (grid rows per 256-thread block) of the stage kernels: 8 or 16.
 * The default is picked from the grid size; the autotuner overrides it. */
int msfdm_set_tile_rows(msfdm_solver *s, int rows);
int msfdm_get_tile_rows(const msfdm_solver *s, int *rows);
/* Occupancy policy of the stage kernels: 0 (default) = the resident waves per
 * SIMD tuned per RK stage on the 2001 x 1001 seafloor grid (stage 1 at least
 * 5, stages 2-3 the compiler's choice, stage 4 at most 4); 1 = at least 5 in
 * every stage (kernels compiled for it: grids with vertical columns, not the
 * fluid-solid CPML grid; EINVAL otherwise).  Results are bit-identical under
 * both; msfdm_autotune picks one.  No reference counterpart. */
int msfdm_set_occupancy(msfdm_solver *s, int policy);
int msfdm_get_occupancy(const msfdm_solver *s, int *policy);
/* How the adjoint products of a stage are launched: 1 = one launch over every
 * source tile (boundary tiles first), 0 = three launches (interior tiles of
 * both media, the other fluid tiles, the other solid tiles), -1 (default) =
 * one launch up to 65536 tiles over the shots of a launch, three above.  The
 * gradients are bit-identical either way.  No reference counterpart. */
int msfdm_set_adjoint_launch(msfdm_solver *s, int mode);
int msfdm_get_adjoint_launch(const msfdm_solver *s, int *mode);
/* Enable (1) / disable (0) hipGraph replay of the time loop (default 1). */
int msfdm_set_graph(msfdm_solver *s, int enable);
/* Time `reps` launches of one kernel on the solver's stream with HIP events
 * (which: 0 = RK stage-2 kernel over both media, 1 = RK stage-4 kernel,
 * 2 = stage-4 interface rows + filter + record); returns the mean ms per
 * launch.  The kernels write a scratch buffer; the interface rows of the stage
 * buffers may be rewritten: reset the shot afterwards. */
int msfdm_time_kernel(msfdm_solver *s, int which, int reps, float *ms);
/* Time nsteps of msfdm_run between HIP events on the solver's stream. */
int msfdm_time_run(msfdm_solver *s, int nsteps, float *ms);
/* Step nsteps with the launch sequence of msfdm_run as direct launches; the
 * launches of the kernel classes in class_mask (bit c = class c) carry start /
 * stop events stamped by their own dispatch (hipExtLaunchKernelGGL), so their
 * durations exclude the gaps between launches.  Classes: 0..3 = RK stages
 * 1..4, 4 = the interface finish after a launch sequence, 5 = the step
 * counter.  kernel_ms[6] / launches[6] (either may be NULL): total ms and
 * count of the timed launches per class.  wall_ms: the whole run between two
 * stream events.  The state advances as msfdm_run's would.  No reference
 * counterpart (measurement). */
int msfdm_profile_run(msfdm_solver *s, int nsteps, int class_mask, float *kernel_ms, int *launches, float *wall_ms);
/* HBM streaming probe on `device`: `reads` input streams (1..4) and one
 * output stream of `mbytes` MB each, out = sum of the inputs over float4
 * elements, timed over `reps` launches after 3 warm-up launches; *tbps =
 * (reads + 1) * bytes per stream / time.  It gives the rate a plain streaming
 * kernel reaches on this part, beside the nominal HBM peak.  Allocates and
 * frees its own buffers.  No reference counterpart (measurement). */
int msfdm_stream_probe(int device, double mbytes, int reads, int reps, double *tbps);
/* Algorithmic HBM bytes of one time step (state + coefficients + PML memory). */
int msfdm_bytes_per_step(const msfdm_solver *s, double *bytes);
/* Algorithmic HBM bytes of one launch timed by msfdm_time_kernel(which): an
 * RK stage-2 launch reads the stage state and U^n and writes the next stage
 * state (3 floats per field), a stage-4 launch reads T4, U^n, T2, T3 and
 * writes U' (5 floats per field), each plus the coefficients once, per
 * updated cell. */
int msfdm_kernel_bytes(const msfdm_solver *s, int which, double *bytes);

/* Tile-shape autotuner: times nsteps steps for each candidate (tile height 8
 * or 16 rows x occupancy policy 0 or 1 where msfdm_set_occupancy offers both)
 * three times, the candidates interleaved, keeps the one with the fastest
 * single timing (msfdm_get_tile_rows / msfdm_get_occupancy read it back) and
 * resets the shot (no reference counterpart).  msfdm_batch_autotune does the
 * same for a batch (msfdm_batch_get_tile_rows / msfdm_batch_get_occupancy). */
int msfdm_autotune(msfdm_solver *s, int nsteps, int *best_rows);

/* Replace the medium arrays of an existing solver (each [rows][nx] float32 in
 * the model-file layout; NULL keeps an array).  Takes effect for the next
 * steps; the wavefield is left untouched. */
int msfdm_update_media(msfdm_solver *s, const float *K, const float *rho_f, const float *C11, const float *C13,
                       const float *C33, const float *C55, const float *rho_s);

/* --- Grid metrics (model generation) ------------------------------------ */
/* FP64 curvilinear metric terms of the grid (x, z), each [nz][nx] row-major
 * doubles in metres, computed on `device`: numpy.gradient(edge_order=2)
 * derivatives in unit computational spacing and the inverse metric
 * J = x_xi z_ze - x_ze z_xi, xi_x = z_ze/J, xi_z = -x_ze/J, ze_x = -z_xi/J,
 * ze_z = x_xi/J.  Replaces the metric block of the reference's model
 * generators (tools/run_models/model_ortho_classic_scholte.py:188-195);
 * the outputs are what they write as jacob/{J,xi_x,xi_z,ze_x,ze_z}.bin
 * after rounding to float32.  No solver is needed. */
int msfdm_grid_metrics(const double *x, const double *z, int nz, int nx, int device, double *J, double *xi_x,
                       double *xi_z, double *ze_x, double *ze_z);

/* --- Discrete adjoint (no reference counterpart) ------------------------- */
/* The adjoint, multi-shot and batched entry points, and msfdm_update_media,
 * serve fluid-solid models with the pressure-release surface (calc_type 1,
 * boundary_type 2); with any other model they return MSFDM_EINVAL. */
/* Residual contract: each residual plane is exactly nrec * nx floats, row-major
 * [nrec][nx] over every grid column (PML and frozen columns included), with
 * nrec = (nsteps - 1) / gap + 1 and gap = (int)(1e-3 / dt + 0.5), the record
 * count of an nsteps-step run.  The library reads that many floats from each
 * plane and cannot check the caller's buffer sizes; the Python binding refuses
 * any other shape (a PML-cut gather, a shorter record).
 * Gradient of J = sum_r <res_r, rec_r> (res = dJ/d(records), three [nrec][nx]
 * arrays for solid vx, solid vz, fluid p as returned by msfdm_get_records) with
 * respect to the model arrays K, rho_f ([nz_f][nx]) and C11, C13, C33, C55,
 * rho_s ([nz_s][nx]), all doubles, for shot isrc and nsteps steps from a zero
 * wavefield.  The transpose of the FP32 scheme (RHS, interface BC, filter,
 * RK4) runs backwards in time; the forward is recomputed per segment from
 * checkpoints taken every `every` steps (0: about sqrt(nsteps)), and the
 * gradients accumulate in FP64.  Any output pointer may be NULL.  Leaves the
 * records of the forward run readable; the wavefield is undefined afterwards.
 * The checkpoint and segment buffers stay allocated for the next gradient
 * (sized for the longest record seen) until msfdm_gradient_shots or
 * msfdm_solver_destroy releases them. */
int msfdm_gradient(msfdm_solver *s, int isrc, int nsteps, int every, const float *res_vx, const float *res_vz,
                   const float *res_p, double *gK, double *grho_f, double *gC11, double *gC13, double *gC33,
                   double *gC55, double *grho_s);
/* The two halves of msfdm_gradient: a checkpointed forward run (records
 * available afterwards) and the reverse sweep for given residuals. */
int msfdm_adjoint_forward(msfdm_solver *s, int isrc, int nsteps, int every);
int msfdm_adjoint_backward(msfdm_solver *s, const float *res_vx, const float *res_vz, const float *res_p,
                           double *gK, double *grho_f, double *gC11, double *gC13, double *gC33, double *gC55,
                           double *grho_s);
/* Multi-shot gradient: the sum over shots isrc[0..nshot) of the gradients of
 * msfdm_gradient, with res = [nshot][3][nrec][nx] floats (per shot the three
 * record planes, nrec = (nsteps - 1) / gap + 1).  The shots run in waves; the
 * shots of a wave step together, forward (checkpointed every `every` steps;
 * 0 picks the interval that needs the least memory) and adjoint, one launch
 * per RK stage for the whole wave.  The wave size is the number of shots whose
 * checkpoints and segment storage fit mem_bytes of device memory (0: 90% of
 * the free memory) and is returned in *wave_shots (may be NULL).  Per-shot
 * gradients are FP64 and are added in shot order, so the result equals the
 * in-order sum of the single-shot gradients.  The solver's wavefield and step
 * counter are left as they were.  The checkpoint and segment buffers that
 * msfdm_gradient keeps between calls are released first (a pending
 * msfdm_adjoint_forward has to be repeated). */
int msfdm_gradient_shots(msfdm_solver *s, int nshot, const int *isrc, int nsteps, int every, double mem_bytes,
                         const float *res, double *gK, double *grho_f, double *gC11, double *gC13, double *gC33,
                         double *gC55, double *grho_s, int *wave_shots);
/* Device bytes one shot of a msfdm_gradient_shots wave needs, and the
 * checkpoint interval it would use. */
int msfdm_gradient_shots_plan(const msfdm_solver *s, int nsteps, int every, double *bytes_per_shot, int *every_used);
/* Operator probes: set an input field, apply op (0 = L_dir^T B_dir^T at the
 * current wavefield, 1 = Filter^T), read an output field. */
int msfdm_adjoint_set_input(msfdm_solver *s, int field, const float *in);
int msfdm_adjoint_apply(msfdm_solver *s, int op, int dir);
int msfdm_adjoint_get_output(msfdm_solver *s, int field, float *out);

/* --- Batched shots ------------------------------------------------------- */
/* Independent shots, each with its own model (grid size, medium, sources) and
 * source index, stepped together: one launch per RK stage covers the tiles of
 * every shot, and the time loop is replayed from a hipGraph.  All models must
 * share dt.  Replaces running src/main.cu's shot loop (src/main.cu:94-213)
 * one shot at a time.  Results per shot are bit-identical to msfdm_run. */
typedef struct msfdm_batch msfdm_batch;
int msfdm_batch_create(const msfdm_model *const *models, const int *shot_index, int nshots, int device,
                       msfdm_batch **out);
void msfdm_batch_destroy(msfdm_batch *b);
/* Zero every shot's wavefield and records. */
int msfdm_batch_reset(msfdm_batch *b);
/* Advance every shot nsteps steps (asynchronous). */
int msfdm_batch_run(msfdm_batch *b, int nsteps);
int msfdm_batch_synchronize(msfdm_batch *b);
int msfdm_batch_size(const msfdm_batch *b, int *nshots);
/* Solver handle of one shot (for records / fields); waits for queued work.
 * The handle stays owned by the batch. */
int msfdm_batch_shot(msfdm_batch *b, int shot, msfdm_solver **out);
int msfdm_batch_set_tile_rows(msfdm_batch *b, int rows);
int msfdm_batch_get_tile_rows(const msfdm_batch *b, int *rows);
/* Occupancy policy of the batched stage kernels, as msfdm_set_occupancy (1
 * needs vertical grid columns in every shot).  The multi-shot gradient's
 * waves use the solver's own policy.  No reference counterpart. */
int msfdm_batch_set_occupancy(msfdm_batch *b, int policy);
int msfdm_batch_get_occupancy(const msfdm_batch *b, int *policy);
int msfdm_batch_set_graph(msfdm_batch *b, int enable);
int msfdm_batch_time_run(msfdm_batch *b, int nsteps, float *ms);
int msfdm_batch_autotune(msfdm_batch *b, int nsteps, int *best_rows);

/* --- Record analysis ----------------------------------------------------- */

/* Phase-shift dispersion image of a time-major gather [nt][ntr] (float32, as
 * the record files hold it after the PML columns are cut) with signed offsets
 * [ntr] (m), sampling rate fs (Hz) and test velocities [nc] (m/s), in FP64 on
 * HIP device `device`: amp[nf][nc] with nf = nt / 2 + 1, and the frequencies
 * [nf] in freqs (may be NULL).  Replaces dispersion_imaging of tools/MAS.py:50-98
 * (tools/test.py:50-99). */
int msfdm_dispersion_image(const float *gather, int nt, int ntr, const double *offsets, double fs,
                           const double *velocities, int nc, int device, double *freqs, double *amp);

/* Allocation-free variant of msfdm_dispersion_image.  No reference counterpart:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and bit-identical results, over caller-owned DEVICE buffers of the
 * current HIP device: d_gather [nt][ntr] float, d_offsets [ntr] and
 * d_velocities [nc] doubles, d_amp [nf][nc] doubles out (nf = nt / 2 + 1;
 * the frequencies are f * fs / nt).  The twiddle table and the phase spectrum
 * live in a static device buffer of MSFDM_DISP_SCRATCH_BYTES per GPU, held
 * until the process ends; 256 * ceil(16 nt / 256) + 16 nf ntr bytes beyond it
 * return MSFDM_EINVAL.  Work is queued on `stream` (a hipStream_t; NULL = the
 * null stream) and the call returns without waiting.  Calls on one GPU share
 * that buffer: calls that overlap in time on one GPU (other streams or threads)
 * race.  Offsets and velocities are not validated (they stay on the device). */
#define MSFDM_DISP_SCRATCH_BYTES (32u << 20)
int msfdm_dispersion_image_noalloc(const float *d_gather, int nt, int ntr, const double *d_offsets, double fs,
                                   const double *d_velocities, int nc, double *d_amp, void *stream);

/* --- Process boundary ---------------------------------------------------- */

/* The whole bin/main_debug run: argument guards, output-dir reset, shot loop,
 * snapshots and records in the src/main.cu:58-339 layout.  Returns the process
 * exit code (0 ok, 1 error) with the reason in msfdm_last_error(). */
int msfdm_run_cli(const char *json_path, const char *out_root);

#ifdef __cplusplus
}
#endif

#endif /* MSFDM_H */
