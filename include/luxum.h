/*
 * luxum.h -- C ABI of the MI355X-native Yee FDTD timestep.
 *
 * The reference (Luxum) has no FFI layer: its boundary is the C++ class API
 * that main() drives.  Each entry point below names the reference interface
 * it replaces.  All entry points are extern "C", take plain pointers and
 * sizes, never throw and never exit(); errors are status codes with a
 * thread-local message from lx_last_error().
 *
 * Conventions shared with the reference:
 *   - nx, ny, nz are NODE counts ("X/Y/Z Length", lib/parallel/grid.cc:83-85).
 *   - Field component extents follow the Yee stagger of lib/field/vfield.cc:58-87
 *     (a staggered axis has n-1 entries).
 *   - Host-side field buffers are FP64, C order (i, j, k) with k fastest,
 *     exactly the layout lib/io/fileio.cc:143-157 writes.
 *   - Planar (2D) runs use the reference's x-z plane with (Ey, Hx, Hz),
 *     ny == 1 (lib/io/reader.cc:151-154).
 */
#ifndef LUXUM_H
#define LUXUM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum lx_status {
    LX_OK = 0,
    LX_EINVAL = 1,    /* invalid argument or parameter set              */
    LX_ENOMEM = 2,    /* device or host allocation failed                */
    LX_EDEVICE = 3,   /* HIP runtime error                               */
    LX_EIO = 4,       /* file could not be read or written               */
    LX_EBUSY = 5,     /* admission refused: not enough device memory     */
    LX_EUNSTABLE = 6  /* non-finite field value detected by the monitor  */
} lx_status;

/* Field component ids (order of vfield Vx,Vy,Vz for E then H). */
enum { LX_EX = 0, LX_EY = 1, LX_EZ = 2, LX_HX = 3, LX_HY = 4, LX_HZ = 5 };

/* Storage precision of E and H on the device (arithmetic is FP32). */
enum { LX_FP32 = 0, LX_FP16 = 1 };

/* Incident-field provider for TF/SF injection. */
enum {
    LX_INC_ANALYTIC = 0, /* reference formula exp(-xi^2/sigma), src/solvers/maxwell.cc:194-812 */
    LX_INC_AUX1D = 1     /* FP64 1D auxiliary Yee line along k (axis-aligned k only)          */
};

/* Waveform of the aux-1D line source and of the soft point source. */
enum {
    LX_WAVE_GAUSSIAN = 0,      /* exp(-((t-t0)/tau)^2)                        */
    LX_WAVE_GAUSSIAN_SINE = 1  /* exp(-((t-t0)/tau)^2) * sin(2 pi f0 (t-t0))   */
};

/* Point-source kind. */
enum {
    LX_SRC_HARD_CW = 0,  /* reference: E(src) = int(amp)*sin(int(omega)*t*dt)  (maxwell_three_dim_pml.cc:161-165) */
    LX_SRC_SOFT = 1      /* E(src) += amp * waveform(t*dt)                                                     */
};

/*
 * Simulation parameters.  Fields of documents 1-5 mirror class reader
 * (lib/io/reader.h:52-167) one to one; the "extension" block is the optional
 * sixth document of parameters.yaml.  lx_params_default() fills a vacuum,
 * PEC-walled 3D grid with every extension at its reference-compatible value.
 */
typedef struct lx_params {
    /* doc 1 */
    int32_t is_planar;
    /* doc 2 */
    int32_t nx, ny, nz;
    double dx, dy, dz;          /* metres */
    double courant;             /* S: dt = S*dx/c (maxwell.cc:67) */
    int32_t num_timesteps;
    /* doc 3 (kept for format parity; one GPU holds the whole domain) */
    int32_t npx, npy, n_threads;
    /* doc 4: sources */
    int32_t use_ps;
    int32_t src_pos[3];
    double src_amp, src_omega;
    int32_t use_pw;
    double k_pw[3];
    double e_pw[3];
    double off_pw;
    int32_t tf_lo[3];           /* x1_pw, y1_pw, z1_pw */
    int32_t tf_hi[3];           /* x2_pw, y2_pw, z2_pw */
    /* doc 5: CPML */
    int32_t use_pml;
    double pml_thickness;       /* metres */
    double sigma_ratio;         /* sigma_max / sigma_optimal */
    double sigma_m;
    double kappa_max;
    double a_max;
    double a_m;
    /* doc 6: extensions (absent from the reference) */
    int32_t precision;          /* LX_FP32 | LX_FP16 */
    int32_t incident;           /* LX_INC_ANALYTIC | LX_INC_AUX1D */
    double pw_sigma;            /* analytic Gaussian width, m^2 (reference: 2e-4, maxwell.cc:71) */
    int32_t pw_wave;            /* aux-1D line waveform */
    double pw_lambda0;          /* carrier wavelength, metres */
    double pw_fwhm;             /* envelope FWHM in carrier periods */
    double pw_delay;            /* envelope centre, in envelope widths tau */
    int32_t src_kind;           /* LX_SRC_HARD_CW | LX_SRC_SOFT */
    int32_t src_comp;           /* component of the point source (default Ez in 3D, Ey in 2D) */
    int32_t src_wave;
    double src_lambda0, src_fwhm, src_delay;
    int32_t snapshot_every;     /* 0: none */
    int32_t monitor_every;      /* 0: none; else FP64 max|E|^2 + NaN guard every n steps */
    double storage_scale;       /* fields are stored times this power of two and every output divides it
                                   out (results equal the unscaled run's wherever no value leaves the
                                   storage type's normal range); 0: automatic, FP32 1, FP16 the power of
                                   two nearest 1/(largest source amplitude), so that binary16's normal
                                   range (6.1e-5 .. 65504) holds the field */
} lx_params;

typedef struct lx_ctx lx_ctx;
typedef struct lx_sim lx_sim;
typedef struct lx_batch lx_batch;

/* Derived quantities of a created simulation (read-only report). */
typedef struct lx_sim_info {
    double dt, c0, eps0, mu0, eta0;
    int64_t cells;              /* nx*ny*nz */
    int64_t slab_cells[3];      /* cells inside the CPML slabs of each axis (both sides) */
    int64_t device_bytes;       /* device memory held by the simulation */
    int64_t bytes_per_step;     /* algorithmic HBM bytes per full timestep */
    int32_t step;               /* timesteps taken so far */
    int32_t slab_lo[3], slab_hi[3];  /* CPML slab: [0,lo) and [hi,n) per axis */
    int32_t fused_step;         /* 3D: fused H+E step over two buffer sets (see lx_sim_time_kernels) */
    int32_t kz, waves;          /* z planes and wavefronts per block of the 3D kernels (autotuned) */
    int64_t core_bytes;         /* fused step on a CPML grid: algorithmic bytes of its interior-box launch
                                   (rows and planes outside the y and z slabs); else bytes_per_step */
    int32_t cells_per_lane;     /* fused step: x cells per lane (1, 2 or 4, autotuned); else 1 */
    int32_t shell_kz;           /* fused step on a CPML grid: z planes per block of the shell launch
                                   (autotuned); else kz */
    int32_t shell_concurrent;   /* fused step on a CPML grid: 1 when the shell launch runs on a second
                                   stream beside the interior launch (autotuned); else 0 */
    int32_t shell_pipelined;    /* 1 when, in addition, the shell launch of a step may overlap the TF/SF
                                   launch before it (every face cell, probe and the point source lie two
                                   cells inside the interior box); else 0 */
    int32_t interior_rows;      /* fused step: rows per block of the interior launch that share their
                                   neighbour rows through LDS (3 or 7, autotuned); 0: one unit per
                                   wavefront, neighbour rows read and recomputed */
} lx_sim_info;

/* ---- parameters ------------------------------------------------------- */

/* Fill p with defaults (no reference counterpart: reader has no defaults). */
lx_status lx_params_default(lx_params *p);

/* Replaces reader::reader / readYAML / checkData (lib/io/reader.cc:68-162):
 * parse the five-document parameters.yaml at `path` (+ optional doc 6). */
lx_status lx_params_load_yaml(const char *path, lx_params *p);

/* Replaces reader::checkData (lib/io/reader.cc:141-162) plus the solver's own
 * preconditions (TF box strictly inside the grid, Courant limit). */
lx_status lx_params_check(const lx_params *p);

/* ---- context ------------------------------------------------------------ */

/* No reference counterpart (MPI_Init + rank setup, main.cc:55, parallel.cc:57-77,
 * are replaced by one device). */
lx_status lx_ctx_create(int32_t device, lx_ctx **out);
void lx_ctx_destroy(lx_ctx *ctx);
/* Device memory in bytes: total and currently free. */
lx_status lx_ctx_mem_info(lx_ctx *ctx, int64_t *free_bytes, int64_t *total_bytes);

/* ---- single simulation ------------------------------------------------- */

/* Replaces maxwell::maxwell (src/solvers/maxwell.cc:60-148) and the field
 * allocation at the top of each solve* (e.g. maxwell_three_dim_pml.cc:62-132). */
lx_status lx_sim_create(lx_ctx *ctx, const lx_params *p, lx_sim **out);
void lx_sim_destroy(lx_sim *sim);

/* Replaces the Cb set-up of maxwell_three_dim_pml.cc:109-132: ids[n] per node
 * (nx*ny*nz, C order, k fastest); eps_r[n_ids] relative permittivity per id.
 * Every E component at storage index (i,j,k) takes eps_r[ids[i,j,k]]. */
lx_status lx_sim_set_materials(lx_sim *sim, const uint8_t *ids, const double *eps_r,
                               int32_t n_ids);

/* Replaces the per-component Cb set-up of maxwell_three_dim_pml.cc:109-128,
 * where Cb->Vx, Cb->Vy and Cb->Vz come from three files and may differ:
 * ids as above; eps_r3[3*id + c] is the relative permittivity that E
 * component c (LX_EX, LX_EY, LX_EZ) at storage index (i,j,k) takes when
 * ids[i,j,k] == id.  Planar runs use the LX_EY column. */
lx_status lx_sim_set_materials3(lx_sim *sim, const uint8_t *ids, const double *eps_r3,
                                int32_t n_ids);

/* Probe points (no reference counterpart; the reference only prints max|E|^2):
 * comp_ijk holds n quadruples (component, i, j, k).  Values are recorded after
 * every timestep, FP64, SI units. */
lx_status lx_sim_set_probes(lx_sim *sim, const int32_t *comp_ijk, int32_t n);

/* Replaces one or more iterations of the solve* time loop
 * (maxwell_three_dim_pml.cc:158-283, maxwell_three_dim.cc:85-195,
 *  maxwell_two_dim_pml.cc:126-188, maxwell_two_dim.cc:85-144).
 * Asynchronous on the simulation's compute stream. */
lx_status lx_sim_step(lx_sim *sim, int32_t nsteps);

/* Block until all work of the simulation is done. */
lx_status lx_sim_sync(lx_sim *sim);

/* Copy component `comp` (LX_EX..LX_HZ) to host, reference layout and SI units.
 * n must equal the component's element count.  Synchronous. */
lx_status lx_sim_read_field(lx_sim *sim, int32_t comp, double *dst, size_t n);

/* No counterpart in the reference: lx_sim_read_field for callers that cannot
 * allocate device memory (tight serving loops with a fixed footprint,
 * harnesses that forbid allocation).  Same arguments and the same values; calls no hipMalloc/hipFree:
 * the component is gathered in slices of whole x indices through a fixed
 * 32 MiB device array per GPU (held from library load until the process
 * ends) and copied slice by slice into dst.  LX_EINVAL when one x slice
 * (ny*nz values of the component's extents) exceeds 4194304 values.  Calls
 * that overlap on one GPU share that array and race.  Synchronous. */
lx_status lx_sim_read_field_noalloc(lx_sim *sim, int32_t comp, double *dst, size_t n);

/* Replaces fileio::writeHDF5's data path (lib/io/fileio.cc:184-212): enqueue
 * a snapshot of `comp` on the side stream into dst (pinned or pageable host
 * memory, reference layout, FP64).  The copy reflects the state after the
 * steps enqueued so far; it is complete after lx_sim_sync. */
lx_status lx_sim_snapshot_async(lx_sim *sim, int32_t comp, double *dst, size_t n);

/* Copy recorded probe values: dst[t*n_probes + p] for t < *n_steps. */
lx_status lx_sim_read_probes(lx_sim *sim, double *dst, size_t n, int32_t *n_steps);

/* FP64 max over cells of Ex^2+Ey^2+Ez^2 at equal indices (the reference
 * monitor, maxwell_three_dim_pml.cc:266-276; 2D: max Ey^2,
 * maxwell_two_dim_pml.cc:181) and the total discrete field energy
 * sum(eps0*eps_r*E^2 + mu0*H^2)/2 * cell volume over every stored value
 * (eps_r of each E component from its material table).  Synchronous. */
lx_status lx_sim_monitor(lx_sim *sim, double *max_e2, double *energy);

/* The reference monitor alone: FP64 max over cells of Ex^2+Ey^2+Ez^2 at
 * equal indices (maxwell_three_dim_pml.cc:266-276; 2D: max Ey^2,
 * maxwell_two_dim_pml.cc:181), the same value as lx_sim_monitor's max_e2,
 * from a pass over the E components only (3*s bytes per cell instead of
 * 6*s+1).  LX_EUNSTABLE when an E value is not finite.  Synchronous. */
lx_status lx_sim_max_e2(lx_sim *sim, double *max_e2);

/* No counterpart in the reference: lx_sim_max_e2 in two halves, for time
 * loops that print the previous step's max (maxwell_three_dim_pml.cc:166-173
 * prints at the top of iteration t what :266-276 computed at the end of
 * t-1).  _start enqueues the E-only monitor of the state after the steps
 * enqueued so far on a second stream and returns; the next lx_sim_step of a
 * fused 3D simulation (one step, no point source) runs beside it, and any
 * later work that writes the buffer it reads waits for it.  _finish waits
 * and returns the value lx_sim_max_e2 would have returned at _start
 * (LX_EUNSTABLE likewise; LX_EINVAL without a started monitor).  A second
 * _start before _finish drops the first result. */
lx_status lx_sim_max_e2_start(lx_sim *sim);
lx_status lx_sim_max_e2_finish(lx_sim *sim, double *max_e2);

lx_status lx_sim_get_info(lx_sim *sim, lx_sim_info *info);

/* The simulation's compute stream (hipStream_t) for event timing. */
void *lx_sim_stream(lx_sim *sim);

/* Average device time per launch, in ms, of the H-update and E-update
 * kernels over `nsteps` extra timesteps, timed with events on the compute
 * stream around every launch (the steps are real steps: state advances).
 * With the fused 3D step, ms_h is its interior-box launch and ms_e its CPML
 * shell launch (0 without CPML). */
lx_status lx_sim_time_kernels(lx_sim *sim, int32_t nsteps, double *ms_h, double *ms_e,
                              double *ms_step);

/* No reference counterpart: choose the z-block depth of the bulk kernels by
 * timing the candidates on this grid (before the first step; the state is
 * reset afterwards).  Results are cached per (shape, precision, CPML) for the
 * process.  *kz_out receives the chosen depth (may be NULL). */
lx_status lx_sim_autotune(lx_sim *sim, int32_t *kz_out);

/* ---- batches of independent simulations ------------------------------- */

/* No reference counterpart: advance many simulations of uneven sizes
 * together.  Planar members (aux-1D or no plane wave, no point source) run
 * as one temporally tiled launch over a flattened tile list of every member
 * per chunk of 8 steps; other members step with their own kernels on their
 * own streams.  Members may still be stepped, read and snapshotted on their
 * own between batch steps; lx_sim_destroy removes a member. */
lx_status lx_batch_create(lx_ctx *ctx, lx_batch **out);
void lx_batch_destroy(lx_batch *b);
lx_status lx_batch_add(lx_batch *b, lx_sim *sim);
lx_status lx_batch_remove(lx_batch *b, lx_sim *sim);
lx_status lx_batch_step(lx_batch *b, int32_t nsteps);
lx_status lx_batch_sync(lx_batch *b);

/* ---- serving: continuous batching of simulation requests --------------- */

/* No reference counterpart.  A server takes simulation requests with arrival
 * times, admits arrived requests (first fit in arrival order) while the sum
 * of their device footprints stays within an HBM cap and the number of
 * active requests within a slot limit, and advances every active request
 * together in ticks (LX_SERVE_BATCHED: one lx_batch; a tick is the smallest
 * remaining step count of the active requests, at most 128 steps).  3D
 * requests are admitted in arrival order while the 3D cells already active
 * stay below 8 Mi cells (LX_SERVE_3D_CELLS overrides; 0: no limit), so large
 * 3D grids, each of which fills the device, run one after another.  Finished
 * requests leave at the end of their tick; their probe series (and fields,
 * when asked) stay readable.  LX_SERVE_PLAIN runs the requests one at a time
 * in arrival order with the single-simulation kernels (the per-request
 * baseline). */
typedef struct lx_server lx_server;

enum { LX_SERVE_BATCHED = 0, LX_SERVE_PLAIN = 1 };
enum { LX_REQ_QUEUED = 0, LX_REQ_ACTIVE = 1, LX_REQ_DONE = 2, LX_REQ_FAILED = 3 };

typedef struct lx_request {
    const lx_params *params;
    const uint8_t *ids;       /* material ids, C order over the grid (as lx_sim_set_materials) */
    size_t n_cells;
    const double *eps_r;      /* eps_r per id */
    int32_t n_ids;
    const int32_t *probes;    /* n_probes x (component, i, j, k), may be NULL */
    int32_t n_probes;
    int32_t steps;            /* timesteps to run */
    double arrival;           /* seconds after the start of lx_server_run */
    int32_t keep_fields;      /* keep the final fields of every component on the host */
} lx_request;

typedef struct lx_request_info {
    int32_t state;            /* LX_REQ_* */
    int32_t steps_done;
    double arrival, admitted, finished;   /* seconds after the start of lx_server_run */
    int64_t footprint;        /* device bytes reserved at admission */
    int64_t cells;
    int32_t n_probes;
} lx_request_info;

typedef struct lx_server_stats {
    int32_t submitted, completed, failed;
    int32_t peak_active;
    int64_t peak_bytes;       /* largest sum of reserved footprints */
    int32_t ticks;
    double makespan;          /* seconds from the start of the run to the last completion */
    double cell_updates;      /* sum over requests of cells x steps */
    double alg_bytes;         /* sum over requests of bytes_per_step x steps (half-step method) */
} lx_server_stats;

/* Upper bound of the device bytes a simulation with these parameters uses
 * (3D: including the second field/psi buffers of the fused step; batched:
 * including those of a tiled 2D batch member), with `steps` probe records
 * of `n_probes` probes. */
lx_status lx_params_footprint(const lx_params *p, int32_t batched, int32_t steps, int32_t n_probes,
                              int64_t *bytes);
lx_status lx_server_create(lx_ctx *ctx, int32_t mode, int64_t hbm_cap_bytes, int32_t max_active,
                           lx_server **out);
void lx_server_destroy(lx_server *srv);
/* Copies the request (parameters, ids, eps_r, probes); *id is its index. */
lx_status lx_server_submit(lx_server *srv, const lx_request *req, int64_t *id);
/* Runs until every submitted request is done or failed. */
lx_status lx_server_run(lx_server *srv);
lx_status lx_server_info(lx_server *srv, int64_t id, lx_request_info *info);
/* Probe series of a finished request: steps x n_probes doubles. */
lx_status lx_server_probes(lx_server *srv, int64_t id, double *dst, size_t n);
/* Final field of a finished request kept with keep_fields (layout as
 * lx_sim_read_field). */
lx_status lx_server_field(lx_server *srv, int64_t id, int32_t comp, double *dst, size_t n);
lx_status lx_server_get_stats(lx_server *srv, lx_server_stats *st);

/* ---- HDF5 output and input (reference layout) --------------------------- */

/* Replaces fileio (lib/io/fileio.cc:55-247) for one process: one file per
 * output, one FP64 dataset per snapshot named "t=<t>" (t formatted as the
 * reference's ostream does), shaped by the field's global extents (staggered
 * axes n-1), C order with k fastest.  LX_OUT_E2 is the 3D solvers' output:
 * |E|^2 at equal indices in an Ez-shaped array, zero outside i<Nx-1,
 * j<Ny-1, k<Nz-1 (maxwell_three_dim_pml.cc:264-280); the 2D solvers write
 * the field Ey (maxwell_two_dim_pml.cc:123,185).  libhdf5 (1.10 C API) is
 * loaded on first use from LX_HDF5_LIB or the usual sonames. */
enum { LX_OUT_FIELD = 0, LX_OUT_E2 = 1 };
typedef struct lx_writer lx_writer;

/* Create (truncate) the file; comp is used for LX_OUT_FIELD.  Close every
 * writer of a simulation before destroying the simulation. */
lx_status lx_writer_open(lx_sim *sim, const char *path, int32_t what, int32_t comp, lx_writer **out);
/* Snapshot of the state after the steps enqueued so far, as dataset "t=<t>".
 * Asynchronous: the step loop waits only for a gather kernel; the copy to
 * host and the file write run on the side stream and a writer thread (at
 * most two snapshots in flight per writer). */
lx_status lx_writer_write(lx_writer *w, double t);
/* Drain pending snapshots and close the file; reports the first write error. */
lx_status lx_writer_close(lx_writer *w);
/* Snapshots written to the file so far. */
int64_t lx_writer_count(lx_writer *w);

/* Write one FP64 dataset "t=<t>" of shape dims (create the file, or append
 * to an existing one). */
lx_status lx_h5_write(const char *path, double t, const double *data, const int64_t dims[3], int32_t append);
/* The same under any dataset name, e.g. the "data" of the material inputs
 * the 3D CPML solver reads (maxwell_three_dim_pml.cc:109-118). */
lx_status lx_h5_write_dataset(const char *path, const char *name, const double *data, const int64_t dims[3],
                              int32_t append);
/* Replaces fileio::readHDF5 (fileio.cc:224-247): read an FP64 dataset; dims
 * receives its shape (rank < 3 padded with 1).  dst may be NULL to query
 * the shape only. */
lx_status lx_h5_read(const char *path, const char *dataset, double *dst, size_t n, int64_t dims[3]);
/* The dataset name the reference gives snapshot t ("t=100", "t=1e+06"). */
lx_status lx_h5_dataset_name(double t, char *buf, size_t n);

/* ---- errors ------------------------------------------------------------- */
const char *lx_last_error(void);
const char *lx_version(void);

#ifdef __cplusplus
}
#endif
#endif /* LUXUM_H */
