/*
 * rfdtd -- MI355X (gfx950) FDTD engine for rfnetwork's field solver: the C ABI.
 *
 * Plain C types only: pointers, sizes and status codes.  Every function returns 0 on success and a
 * non-zero RFDTD_E* code on failure; rfdtd_last_error() then holds a message for the calling
 * thread.  No C++ exception crosses this boundary and no global solver state exists: everything
 * lives in an rfdtd_session, and one session per stream may be used from any thread.
 *
 * Layout ("GPU packing" of the reference, rfnetwork/solver.py:1463-1523 with gpu=True): every one of
 * the 30 coefficient arrays and every field array is float[Nx][Ny][Nz], C order, flat index
 * (x*Ny + y)*Nz + z.  Cell (x,y,z) holds Ex(x,y+1,z+1), Ey(x+1,y,z+1), Ez(x+1,y+1,z), Hx(x+1,y,z),
 * Hy(x,y+1,z) and Hz(x,y,z+1) of the global Yee grid (rfnetwork/core/src/solver.cu:134).
 */
#ifndef RFDTD_H
#define RFDTD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RFDTD_NUM_COEFFS 30
#define RFDTD_NUM_FIELDS 18
#define RFDTD_MAX_MONITORS 50  /* rfnetwork/core/inc/solver.h:15 */
#define RFDTD_MAX_PROBES 5000  /* rfnetwork/core/inc/solver.h:16 */

/* status codes */
#define RFDTD_OK 0
#define RFDTD_EINVAL 1   /* bad argument (shape, index, count, dtype) */
#define RFDTD_EDEVICE 2  /* HIP runtime error */
#define RFDTD_ESTATE 3   /* call out of order (e.g. run past the probe series) */
#define RFDTD_ENOMEM 4   /* the request does not fit in device memory */

/* flags for rfdtd_create / rfdtd_solver_run */
#define RFDTD_F64 0x1u            /* fields, coefficients, probe series and monitors in double */
#define RFDTD_COEF_RAW 0x2u       /* keep the 30 coefficient arrays as given (no deduplication) */
#define RFDTD_COEF_TABLE 0x4u     /* force the deduplicated coefficient table */
#define RFDTD_NO_GRAPHS 0x8u      /* launch every step eagerly instead of replaying a captured chunk */
#define RFDTD_QUIET 0x10u         /* no progress output */
#define RFDTD_KERNEL_TIMING 0x20u /* launch eagerly and time the E/H update kernels with HIP events */
#define RFDTD_VERBOSE 0x40u       /* progress lines on stderr during long gradient passes */
#define RFDTD_HOST_TABLE 0x80u    /* sessions from host arrays: build the coefficient table on the host
                                     (default: on the device; assembled sessions always do) */
#define RFDTD_TUNE_SEPARATE 0x100u /* rfdtd_autotune tries the separate E and H kernels only (variants 0,
                                      1, 4, 10: one step per pass, every field and coefficient read once) */

/*
 * Coefficient order: the keys of the reference's coefficient dictionary (solver.py:1473-1518):
 *   0 Ca_ex_y  1 Ca_ex_z  2 Cb_ex_y  3 Cb_ex_z  4 Ca_ey_z  5 Ca_ey_x  6 Cb_ey_z  7 Cb_ey_x
 *   8 Ca_ez_x  9 Ca_ez_y 10 Cb_ez_x 11 Cb_ez_y
 *  12 Da_hx_y 13 Da_hx_z 14 Db_hx_y1 15 Db_hx_y2 16 Db_hx_z1 17 Db_hx_z2
 *  18 Da_hy_z 19 Da_hy_x 20 Db_hy_z1 21 Db_hy_z2 22 Db_hy_x1 23 Db_hy_x2
 *  24 Da_hz_x 25 Da_hz_y 26 Db_hz_x1 27 Db_hz_x2 28 Db_hz_y1 29 Db_hz_y2
 * Field order (rfdtd_read_fields): ex_y ex_z ey_z ey_x ez_x ez_y hx_y hx_z hy_z hy_x hz_x hz_y
 *                                  ex ey ez hx hy hz
 */

/* One probe or soft source (solver.py:1541-1575; solver.cpp:465-532).  idx holds GLOBAL Yee
 * indices; the callee applies the component shifts.  values: float (or double with RFDTD_F64)
 * [>= Nt]; source samples in, sampled total field out, written back by rfdtd_read_probes. */
typedef struct rfdtd_probe {
    int32_t field;    /* 0 ex, 1 ey, 2 ez, 3 hx, 4 hy, 5 hz */
    int32_t idx[3];
    int32_t is_source;
    int32_t reserved;
    void* values;
} rfdtd_probe;

/* One plane monitor (solver.py:1597-1620; solver.cpp:351-463 with gpu=1).  position is the global
 * field index along axis.  n_frequencies == 0: values is float[Nt/n_step+1][N1][N2] (time samples);
 * otherwise values is complex64[n_frequencies][N1][N2] (running DFT) and dtft_phase is
 * complex64[Nt/n_step+1][n_frequencies].  (N1,N2) = (Ny,Nz), (Nx,Nz), (Nx,Ny) for axis 0, 1, 2.
 * In a float64 session every one of these is double / complex128.  flags: 0, or
 * RFDTD_MONITOR_DFT64 for a running DFT accumulated in double in a float32 session (values and
 * dtft_phase complex128; the field samples stay float32).  The reference accumulates in complex64
 * only (solver.cpp:1194-1225), so flags = 0 is its behaviour. */
#define RFDTD_MONITOR_DFT64 0x1u
typedef struct rfdtd_monitor {
    int32_t axis;
    int32_t position;
    int32_t field;
    int32_t n_step;
    int32_t n_frequencies;
    int32_t flags;
    void* values;
    const void* dtft_phase;
} rfdtd_monitor;

typedef struct rfdtd_session rfdtd_session;

/* Message of the last failed call on this thread ("" if none). */
const char* rfdtd_last_error(void);

/* Library/ABI version, e.g. 1. */
int32_t rfdtd_version(void);

/*
 * One-shot solve with the semantics of the reference's GPU entry point
 * cuda_func.solver_run_cu(coefficients, probes, monitors, mem, Nx, Ny, Nz, Nt, n_threads,
 * update_interval) (rfnetwork/core/src/cuda_binding.cpp:20-68 -> solver.cu:546-1143): copies the 30
 * host coefficient arrays to HBM (treating the x = Nx-1 planes of Cy, Cz and Dx as zero, the PEC
 * wall of solver.cpp:294-328), runs Nt steps, and writes every probe series and monitor back into the
 * caller's host arrays.  Monitors sample after both half steps, as the CPU solver does
 * (solver.cpp:1130).  update_interval > 0 prints the reference's progress bar.
 */
int rfdtd_solver_run(const void* const coefficients[RFDTD_NUM_COEFFS], const rfdtd_probe* probes,
                     int32_t n_probes, const rfdtd_monitor* monitors, int32_t n_monitors, int32_t Nx,
                     int32_t Ny, int32_t Nz, int32_t Nt, int32_t update_interval, uint32_t flags);

/* Session API: the same solve split into resident steps (fields and coefficients stay in HBM).
 * rfdtd_set_probes / rfdtd_set_monitors / rfdtd_read_fields address instance 0. */
int rfdtd_create(rfdtd_session** out, int32_t Nx, int32_t Ny, int32_t Nz,
                 const void* const coefficients[RFDTD_NUM_COEFFS], uint32_t flags);
int rfdtd_set_probes(rfdtd_session* s, const rfdtd_probe* probes, int32_t n_probes, int32_t Nt);
int rfdtd_set_monitors(rfdtd_session* s, const rfdtd_monitor* monitors, int32_t n_monitors, int32_t Nt);
/*
 * Batched session (no counterpart in the reference, which solves one grid and one excitation per
 * call): n_grids grids of their own sizes, grid g advanced for excitations[g] excitations that share
 * its 30 coefficient arrays.  dims is int32[3*n_grids] (Nx, Ny, Nz per grid) and coefficients holds
 * 30 pointers per grid, in the order above.  Instances (grid, excitation) are numbered grid by grid,
 * excitation fastest; every phase of a time step is one launch over all of them.
 */
int rfdtd_create_batch(rfdtd_session** out, int32_t n_grids, const int32_t* dims, const int32_t* excitations,
                       const void* const* coefficients, uint32_t flags);
int32_t rfdtd_instance_count(const rfdtd_session* s);
/* Probes/sources and monitors of one instance (all instances share Nt). */
int rfdtd_set_instance_sources(rfdtd_session* s, int32_t instance, const rfdtd_probe* probes, int32_t n_probes,
                               const rfdtd_monitor* monitors, int32_t n_monitors, int32_t Nt);
/* Fields of one instance, in the layout of rfdtd_read_fields. */
int rfdtd_read_instance_fields(rfdtd_session* s, int32_t instance, void* const out[RFDTD_NUM_FIELDS]);

/* Advance n_steps time steps on stream (NULL = the session's own stream); asynchronous.  For a
 * batched session every instance advances. */
int rfdtd_run(rfdtd_session* s, int32_t n_steps, void* stream);
int rfdtd_synchronize(rfdtd_session* s);
/* Copy probe series / monitor values back into the host arrays given to set_probes / set_monitors
 * (all instances of a batched session). */
int rfdtd_read_probes(rfdtd_session* s);
int rfdtd_read_monitors(rfdtd_session* s);
/* Copy the 18 fields into global Yee-shaped host arrays: ex* (Nx,Ny+1,Nz+1), ey* (Nx+1,Ny,Nz+1),
 * ez* (Nx+1,Ny+1,Nz), hx* (Nx+1,Ny,Nz), hy* (Nx,Ny+1,Nz), hz* (Nx,Ny,Nz+1); NULL entries skipped. */
int rfdtd_read_fields(rfdtd_session* s, void* const out[RFDTD_NUM_FIELDS]);
/* Zero fields, probe series state and monitor accumulators; the step counter returns to 0. */
int rfdtd_reset(rfdtd_session* s);
/* Time steps taken since create/reset. */
int64_t rfdtd_step(const rfdtd_session* s);
/* Timing of the last rfdtd_run, read after rfdtd_synchronize (HIP events on the launch stream),
 * milliseconds: the whole run, and with RFDTD_KERNEL_TIMING the average duration of one E and one
 * H update kernel over up to 64 evenly spread steps of that run. */
int rfdtd_last_timing(const rfdtd_session* s, double* total_ms, double* e_kernel_ms, double* h_kernel_ms);
/* Device bytes held by the session, and coefficient storage kind (0 raw arrays, 1 table). */
int rfdtd_memory_info(const rfdtd_session* s, uint64_t* device_bytes, int32_t* coef_kind,
                      int64_t* n_e_entries, int64_t* n_h_entries);
/* Launch shape: threads per block (64..1024, multiple of 64) and x planes per block (>= 1, or -1 for
 * the automatic choice that gives enough workgroups to fill the chip); 0 keeps the current value. */
int rfdtd_set_launch(rfdtd_session* s, int32_t block_threads, int32_t x_planes, int32_t graph_steps);
/* Launch autotuner (no counterpart in the reference, whose CUDA path uses one fixed launch shape):
 * times every candidate {variant 1, 4, 10, 0} x {block 128, 256, 512} x {x planes: automatic, Nx/2 ..
 * Nx/128} and {variant 2, 3, 5, 6, 7, 8} x {x planes} (RFDTD_TUNE_SEPARATE: the first set only) over
 * trial_steps steps of this session's grids (with its probes and monitors applied when they are
 * installed, restored afterwards), keeps the fastest and returns it as config_out = {variant,
 * block_threads, x_planes (0 = automatic)} with its ms per step.  Only on a session that has not run
 * yet (step 0); the fields are left zero. */
int rfdtd_autotune(rfdtd_session* s, int32_t trial_steps, int32_t config_out[3], double* ms_per_step);
/*
 * Adjoint gradients (no counterpart in the reference, which has no differentiation).
 * rfdtd_set_checkpointing(s, K): subsequent rfdtd_run calls save the 12 split fields of every
 * instance before each step that is a multiple of K (one device copy per checkpoint).
 * rfdtd_gradient: after a checkpointed forward run of Nt steps (or one that recorded the box's
 * history, below), runs the discrete adjoint of those steps backwards, recomputing each K-step
 * segment from its checkpoint when there is no history, and returns dJ/dCb for the
 * six E splits of every cell of box = {x0, x1, y0, y1, z0, z1} (grid cells, upper bounds
 * exclusive) as double[6][x1-x0][y1-y0][z1-z0] in split order ex_y ex_z ey_z ey_x ez_x ez_y, where
 * Cb is the packed coefficient (reference dictionary entries 2, 3, 6, 7, 10, 11).
 * adjoint_series[p] points to dJ/dy_p[n], n < Nt, for probe p of the instance (same order and dtype
 * as its probes).  The forward fields are left untouched.
 */
int rfdtd_set_checkpointing(rfdtd_session* s, int32_t every);
/* Design-box history (no counterpart in the reference): record the H totals of box k of instance k
 * (boxes[6k .. 6k+6), as for rfdtd_gradient) at every step of later runs, for up to n_steps steps;
 * a gradient over exactly these boxes then needs neither checkpoints nor recomputation.  Only before
 * the first step (or after rfdtd_reset); n = 0 drops the history.  RFDTD_ENOMEM if it does not fit. */
int rfdtd_set_history(rfdtd_session* s, int32_t n, const int32_t* instances, const int32_t* boxes, int32_t n_steps);
/* Free and total device memory in bytes (no counterpart in the reference). */
int rfdtd_device_memory(uint64_t* free_bytes, uint64_t* total_bytes);
int rfdtd_gradient(rfdtd_session* s, int32_t instance, const int32_t box[6], const void* const* adjoint_series,
                   int32_t n_series, double* grad_out);
/* The same for n instances of a batched session at once (their reverse passes share every launch):
 * instances[k], boxes[6k .. 6k+6), adjoint_series[k][p], n_series[k], grad_out[k] as above.  An
 * instance appears at most once. */
int rfdtd_gradient_batch(rfdtd_session* s, int32_t n, const int32_t* instances, const int32_t* boxes,
                         const void* const* const* adjoint_series, const int32_t* n_series, double* const* grad_out);
/* Checkpoint interval for such a gradient after an Nt-step forward run: the K that minimises the
 * memory of ceil(Nt/K) checkpoint sets (12 split arrays of every instance) plus K per-step H slots
 * of the design boxes, with the adjoint buffers added; fails with RFDTD_ENOMEM when even that
 * exceeds budget_bytes (0: the device's free memory less 1 GiB).  Returns K and the bytes needed. */
int rfdtd_plan_checkpoints(rfdtd_session* s, int32_t Nt, int32_t n, const int32_t* instances, const int32_t* boxes,
                           uint64_t budget_bytes, int32_t* every_out, uint64_t* bytes_out);

/* Kernel variant (no counterpart in the reference): 1 separate E and H kernels that issue
 * the loads of the next x plane before the arithmetic and stores of the current one; 0 the same
 * kernels loading one plane at a time; 2 and 3 one fused pass per time step (E and H of a tile of
 * 6 or 14 rows x 62 cells, neighbours through LDS, halos recomputed) that reads one buffer of a
 * per-instance pair and writes the other (the session then holds two field buffers per
 * instance); 4 variant 1 with nontemporal stores of the fields; 5 and 6 variants 2 and 3 with the
 * next x plane's H splits loaded one plane earlier and the coefficient values issued first; 7 and 8
 * the two-step pass: steps n and n + 1 of a tile of 12 (8: 4) rows x 60 cells in one launch (two-cell halos
 * recomputed, the x planes skewed by one), reading one buffer of the pair and writing the other, so
 * that the fields are read and written once per two steps (single steps, and every step while a
 * design-box history is recorded, go through the same pass one step at a time); 9 the two-step
 * pass over 8-row tiles with the x planes skewed by three (one barrier per plane); 10 variant 4 with
 * nontemporal loads of the arrays streamed once per step (split fields, raw coefficient arrays, entry
 * indices); 11 and 12 the separate kernels over 2 and 4 consecutive z cells per thread, every array
 * moved as 8- or 16-byte vectors (Nz a multiple of 2 or 4, else RFDTD_EINVAL); 13, 14 and 15 the
 * two-step pass over tiles of 32-lane rows, 28 x 28, 12 x 28 and 26 x 28 cells.  Results are
 * identical; the choice only changes speed (the autotuner picks it).  A session starts on variant
 * 15 with automatic x chunks when it is float, on the coefficient table, and advances at least 1.5 Mi
 * cell-updates per step (cells x excitations over all grids); otherwise on variant 1. */
int rfdtd_set_variant(rfdtd_session* s, int32_t variant);

/* Workgroup numbering of the separate update kernels (variants 0, 1, 4, 10, 11, 12; no counterpart in
 * the reference): 0 (the default) gives each XCD a contiguous range of tiles, so that neighbouring
 * tiles share y-neighbour rows in its L2; 1 takes workgroups in launch order, so that every XCD
 * streams through every region of the arrays.  Launch order is faster on grids of tens of millions of
 * cells, the XCD ranges on small ones; the autotuner picks it with the launch shape (and apart for the
 * checkpointed gradient's recomputation, which this call also sets).  Results are identical either
 * way.  RFDTD_EINVAL for other values. */
int rfdtd_set_column_order(rfdtd_session* s, int32_t order);
int rfdtd_get_column_order(rfdtd_session* s, int32_t* order);

/* Launch shape of the reverse pass of rfdtd_gradient(_batch) (no counterpart in the reference):
 * variant 1 x-marching adjoint columns that issue the next plane's loads (and the coefficient keys
 * of the plane after) before the current plane's arithmetic, 0 one plane at a time, 2 and 3 variants
 * 0 and 1 with the workgroups in launch order (rfdtd_set_column_order); block_threads
 * per workgroup; x_chunks workgroups along x per (y,z) tile (0: the forward's planes per
 * workgroup).  block_threads 0 (the default) times the candidates on the zeroed adjoint state at
 * the next gradient and keeps the fastest for the session.  Results are identical either way.
 * rfdtd_get_adjoint_launch returns {variant, block_threads, x_chunks} in force. */
/* Streamed time-domain monitors (the reference keeps the whole series in host memory it writes to,
 * solver.cpp:1168-1191): with ring_samples > 0, time-domain monitors set afterwards keep only
 * ring_samples sample planes on the device; the run is cut into segments producing at most half a
 * ring each, and after every segment the new planes are copied, on a copy stream of the session,
 * into the monitor's host array (rfdtd_monitor.values), overlapping the next segment.  For the
 * copies to be asynchronous that array should be page-locked (rfdtd_host_alloc); any host array
 * works.  rfdtd_read_monitors waits for the copies.  0 (default):
 * the whole series stays on the device until rfdtd_read_monitors. */
int rfdtd_set_monitor_streaming(rfdtd_session* s, int32_t ring_samples);

/* Page-locked host memory from this library's runtime (hipHostMalloc / hipHostFree), for the host
 * arrays of streamed monitors (no counterpart in the reference). */
int rfdtd_host_alloc(uint64_t bytes, void** ptr);
int rfdtd_host_free(void* ptr);

/* Design-region coefficient assembly on the GPU (the on-device counterpart of the host assembly
 * rfnetwork/solver.py:892-961 and the packing :1463-1523, for edges of plain dielectric): the 12 E
 * coefficients of every packed cell of gbox (x0, x1, y0, y1, z0, z1, packed cells) of member
 * `member` are recomputed from eps_new (F/m) and sigma (S/m) given per cell over gbox grown by one
 * cell on the upper side of each axis (clipped to the grid), C order, float64.  dx, dy, dz: cell
 * widths (m, float64, Nx / Ny / Nz); dx1, dy1, dz1: the packing's inverse dual spacings in the
 * session's element type (Nx / Ny / Nz, last one 0); dt: the time step; e0: vacuum permittivity.
 * eps_now is the permittivity the session's coefficients were built from (or last set): the tuples
 * it gives must equal the session's, else RFDTD_EINVAL and nothing is changed (so a box touching
 * PML, conductor, lumped-element or edge-corrected edges is refused).  Every instance of the member
 * sees the new coefficients from its next step. */
int rfdtd_set_design_permittivity(rfdtd_session* s, int32_t member, const int32_t gbox[6], const double* eps_now,
                                  const double* eps_new, const double* sigma, const double* dx, const double* dy,
                                  const double* dz, const void* dx1, const void* dy1, const void* dz1, double dt,
                                  double e0);

/* The same update from permittivity held in device memory (a ROCm torch tensor, say), with no host
 * copy or synchronisation: the device-pointer extension of the reference's array boundary
 * (rfnetwork/core/src/cuda_binding.cpp:20-68 takes host arrays only).  rfdtd_design_prepare takes
 * rfdtd_set_design_permittivity's arguments without eps_new, checks the box as it does (RFDTD_EINVAL
 * if eps_now's tuples are not the session's), and keeps the box's arrays on the device; a host update
 * of the member (rfdtd_set_design_permittivity) discards them.
 * rfdtd_set_design_permittivity_device then sets the relative permittivity of the cells of cbox
 * (x0, x1, y0, y1, z0, z1, cells; inside the prepared box grown by one cell on each upper side) from
 * eps_r (device pointer, C order, float64 if f64 else float32): eps_r * e0 (in float64) replaces the
 * held permittivity there and the box's coefficients are recomputed from it, bit for bit as the host
 * update would, on `stream` (NULL: the session's stream).  It is ordered after the work on the
 * session's own stream and after the session's last rfdtd_run, whatever stream that was given; the
 * session's own stream and its next rfdtd_run, on any stream, wait for it.  eps_r must stay valid
 * until the update has run.  Later updates start from the values set last. */
int rfdtd_design_prepare(rfdtd_session* s, int32_t member, const int32_t gbox[6], const double* eps_now,
                         const double* sigma, const double* dx, const double* dy, const double* dz, const void* dx1,
                         const void* dy1, const void* dz1, double dt, double e0);
int rfdtd_set_design_permittivity_device(rfdtd_session* s, int32_t member, const int32_t cbox[6], const void* eps_r,
                                         int32_t f64, void* stream);

int rfdtd_set_adjoint_launch(rfdtd_session* s, int32_t variant, int32_t block_threads, int32_t x_chunks);
int rfdtd_get_adjoint_launch(rfdtd_session* s, int32_t config_out[3]);
void rfdtd_destroy(rfdtd_session* s);

/*
 * Far field (replaces core_func.nf2ff, rfnetwork/core/src/postprocess.cpp:130-505, called from
 * FDTD_Solver.get_farfield_rE, rfnetwork/solver.py:2870-3020): the radiation integrals of the
 * equivalent surface currents J = n x H and M = -n x E on the faces of a closed box (Balanis,
 * Advanced Engineering Electromagnetics, 2nd ed., 6.8.2), on the GPU.  out: complex64
 * [2][n_frequencies][n_theta][n_phi], polarisation theta then phi, r E with exp(-j beta r)/r removed.
 */
typedef struct rfdtd_ff_face {
    int32_t axis;       /* normal axis 0..2 */
    int32_t n0, n1;     /* cells along the two in-face axes (lower axis number first) */
    int32_t reserved;
    double position;    /* face position along the normal, metres */
    const float* r0;    /* [n0][n1] cell-centre coordinate along the first in-face axis, metres */
    const float* r1;    /* [n0][n1] along the second */
    const float* ds;    /* [n0][n1] cell areas, square metres */
    const void* J;      /* complex64 [3][n_frequencies][n0][n1] */
    const void* M;      /* complex64 [3][n_frequencies][n0][n1] */
} rfdtd_ff_face;
int rfdtd_nf2ff(int32_t n_faces, const rfdtd_ff_face* faces, int32_t n_frequencies, const float* beta,
                int32_t n_theta, const float* theta, int32_t n_phi, const float* phi, float eta0, void* out);
/*
 * rfdtd_nf2ff without any device allocation (no counterpart in the reference): for callers that
 * cannot allocate device memory -- code under graph capture, serving loops, harnesses that forbid
 * hipMalloc.  The inputs are staged in, and the result taken from, a static 64 MiB device workspace
 * (one per GPU, held until the process ends); the kernel and the result are those of rfdtd_nf2ff.
 * At most 64 faces; a request whose staged bytes (each array rounded up to 256 B: faces, r0, r1,
 * ds, J, M, beta, theta, phi and the result) exceed 64 MiB returns RFDTD_EINVAL before anything is
 * copied.  Calls of one process share the workspace and are serialised by a lock.
 */
int rfdtd_nf2ff_static(int32_t n_faces, const rfdtd_ff_face* faces, int32_t n_frequencies, const float* beta,
                       int32_t n_theta, const float* theta, int32_t n_phi, const float* phi, float eta0, void* out);

/*
 * On-device coefficient assembly of a whole grid (the counterpart of the host assembly
 * FDTD_Solver._init_coefficients / _init_PML / _init_conductors / _init_lumped_elements /
 * edge_correction and the packing, rfnetwork/solver.py:817-1325, :2534-2700, :1463-1523): a session
 * whose 30 coefficient arrays are computed on the GPU from cell permittivity and conductivity, the
 * PML profiles and the overrides below, equal bit for bit to the host assembly's, then kept in the
 * deduplicated table form (or raw, as rfdtd_create decides).  The element type is the session's
 * (RFDTD_F64 in flags).  Sizes: Nx, Ny, Nz cells.
 */
/* One PML side: n layers along axis (0, 1, 2) at its lower (upper = 0) or upper end;
 * sigma_edge[i], sigma_half[i], i = 0 .. n-1: the graded conductivity at the E edges and at the H
 * half cells, as solver.py:1255-1274 forms them (float64). */
typedef struct rfdtd_pml_side {
    int32_t axis, upper, n, reserved;
    const double* sigma_edge;
    const double* sigma_half;
} rfdtd_pml_side;
/* Ca and Cb of both splits of E edges replaced after the PML, in list order.  component 0..2
 * (Ex, Ey, Ez); cells: packed cell indices of those edges (the GPU packing's cell holding the
 * component).  kind 0, a conductor (solver.py:863-868): s = values[0] = sigma * dt as the host forms
 * it, Ca = (2 eps - s) / (2 eps + s), Cb = 2 dt / (2 eps + s).  kind 1, a lumped element
 * (solver.py:1190-1200): q = values[j] per edge, Ca = (eps / dt - q) / (eps / dt + q),
 * Cb = 1 / (eps / dt + q).  eps is the edge's permittivity in the element type; wide = 1 when s or
 * q is a float64 value, so that the arithmetic after eps / dt is float64 and rounded at the end,
 * as numpy's type promotion makes it. */
typedef struct rfdtd_edge_fix {
    int32_t component, kind, wide, reserved;
    int64_t count;
    const int64_t* cells;
    const double* values;
} rfdtd_edge_fix;
/* One factor of an edge correction (solver.py:2614-2655): the unpacked Db of coefficient `coef`
 * (14..17, 20..23, 26..29 in the coefficient order above) at packed cell `cell` is multiplied by
 * factor, in float64 (wide = 1) or in the element type, in list order. */
typedef struct rfdtd_h_scale {
    int32_t coef, wide;
    int64_t cell;
    double factor;
} rfdtd_h_scale;
typedef struct rfdtd_assembly {
    int32_t Nx, Ny, Nz, n_pml;
    const double* eps;              /* [Nx][Ny][Nz] cell permittivity, F/m */
    const double* sigma;            /* [Nx][Ny][Nz] cell conductivity, S/m */
    const double* width[3];         /* cell widths along x, y, z, metres (float32 values) */
    const void* inv_dual[3];        /* the packing's inverse dual spacings (element type, last 0) */
    const void* inv_width[3];       /* 1 / width in the element type */
    double dt;                      /* time step (a float32 value) */
    double e0, u0;                  /* vacuum permittivity and permeability */
    const rfdtd_pml_side* pml;      /* n_pml sides, in the order the host applies them */
    int32_t n_fix, n_scale;
    const rfdtd_edge_fix* fix;      /* n_fix groups */
    const rfdtd_h_scale* scale;     /* n_scale factors */
} rfdtd_assembly;
int rfdtd_create_assembled(rfdtd_session** out, const rfdtd_assembly* a, int32_t excitations, uint32_t flags);
/* A batch (as rfdtd_create_batch) of n_grids grids each assembled on the GPU from a[g], with
 * excitations[g] instances each. */
int rfdtd_create_batch_assembled(rfdtd_session** out, int32_t n_grids, const rfdtd_assembly* a,
                                 const int32_t* excitations, uint32_t flags);
/* The member's 30 packed coefficient arrays (the x = Nx-1 plane of arrays 4..17 reads zero), from
 * its table or raw arrays, into host arrays of the element type (Nx*Ny*Nz each; null skips one).
 * No counterpart in the reference (its coefficients never leave the host): the tests compare the
 * device assembly with the host's through it, and the gradient's chain rule reads a box of them. */
int rfdtd_read_coefficients(rfdtd_session* s, int32_t member, void* const out[30]);
/* The same over a box of packed cells (x0, x1, y0, y1, z0, z1): out[k] holds
 * (x1-x0)*(y1-y0)*(z1-z0) values, C order. */
int rfdtd_read_coefficient_box(rfdtd_session* s, int32_t member, const int32_t box[6], void* const out[30]);

/* Bandwidth probe (no counterpart in the reference): copy `bytes` between two device buffers `iters`
 * times with a plain streaming kernel and report read+write GB/s and ms per copy -- the measured HBM
 * ceiling, and the known-size kernel that calibrates the profiler's byte counters. */
int rfdtd_copy_bandwidth(uint64_t bytes, int32_t iters, double* gbytes_per_s, double* ms_per_copy);

/* Barrier probe (no counterpart in the reference; the measurement behind advancing time steps in
 * graph-replayed launches rather than in one persistent launch): `workgroups` workgroups of 256
 * threads in one cooperative launch meet `iters` times at a device-wide barrier; fences 0: none (the
 * synchronisation alone), 1: one wave per workgroup writes back and invalidates its XCD's L2 (what
 * makes one step's stores visible to the next across XCDs), 2: every wave does.  Reports the best of
 * three runs in microseconds per barrier, and the spacing of dependent graph-replayed launches of an
 * empty kernel of the same grid.  Every wait is bounded (one second of wall clock); a run that
 * exceeds it returns RFDTD_EDEVICE.  RFDTD_EINVAL: more workgroups than can be resident at once. */
int rfdtd_grid_barrier_cost(int32_t workgroups, int32_t fences, int32_t iters, double* us_per_barrier,
                            double* us_per_launch);

#ifdef __cplusplus
}
#endif

#endif /* RFDTD_H */
