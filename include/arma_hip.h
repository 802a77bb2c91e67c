/*
 * arma_hip.h -- C ABI of the MI355X-native ARMA hot path (libarma_hip.so).
 *
 * The entry points replace, one for one, the framework-specific bodies of
 * the reference generator and velocity solver:
 *
 *   arma_hip_ar_generate        AR_model<T>::do_generate()
 *                               (src/generator/ar_model_sequential.cc:3-11,
 *                                src/generator/ar_model_opencl.cc:46-121,
 *                                src/generator/ar_model_parallel.cc:76-201;
 *                                selected at src/generator/ar_model.cc:189-197)
 *   arma_hip_white_noise        Basic_ARMA_model<T>::generate_white_noise()
 *                               (src/generator/basic_arma_model.cc:36-48) and
 *                               the per-partition ::generate_white_noise
 *                               (src/generator/ar_model_wn.cc:3-18)
 *   arma_hip_ar_recurrence      ar_generate_surface over a whole grid
 *                               (src/generator/ar_model_surf.cc:3-36,
 *                                src/kernels/ar_generate_surface.cl:6-49)
 *   arma_hip_rescale            the post-rescale of Basic_ARMA_model::generate
 *                               (src/generator/basic_arma_model.cc:102-106)
 *   arma_hip_linear_potential   Velocity_potential_solver<T>::operator() with
 *                               Linear_solver<T>
 *                               (src/velocity/basic_solver.cc:39-77,
 *                                src/velocity/linear_solver.cc:20-184)
 *   arma_hip_read_mt_config     read_parallel_mts / mt_config >>
 *                               (src/parallel_mt.cc:28-59,
 *                                src/parallel_mt.hh:90-97)
 *   arma_hip_write_mt_config    mt_config << (src/parallel_mt.hh:99-104)
 *   arma_hip_dcmt_create        parallel_mt_seq<p,w>::operator()
 *                               (src/parallel_mt.hh:177-210, src/dcmt.cc:15-40)
 *   arma_hip_init / _destroy    arma_init / arma_finalise
 *                               (src/common_main.hh:48-75)
 *
 * Conventions: every function returns ARMA_HIP_OK (0) or a negative error
 * code; arma_hip_last_error() returns a thread-local message for the last
 * failure.  Arrays are row-major (t, x, y) exactly as Blitz lays out
 * Array3D<T> (src/types.hh:25-43); potentials are (t, z, x, y).  A context
 * owns one HIP stream and the device buffers it allocates; it is used by one
 * host thread at a time.
 */
#ifndef ARMA_HIP_H
#define ARMA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ARMA_HIP_OK 0
#define ARMA_HIP_EINVAL (-1)      /* std::invalid_argument in the reference */
#define ARMA_HIP_EPRNG (-2)       /* arma::PRNG_error (src/errors.hh:10-33) */
#define ARMA_HIP_ERUNTIME (-3)    /* std::runtime_error */
#define ARMA_HIP_EDEVICE (-4)     /* HIP runtime failure */
#define ARMA_HIP_ENOMEM (-5)

/* Fields of the dynamic-creator mt_struct, in file order
   (src/parallel_mt.hh:109-123); the state words are not part of it because
   every stream is re-seeded with sgenrand_mt (src/parallel_mt.hh:228-260). */
typedef struct arma_mt_params {
	uint32_t aaa;
	int32_t mm, nn, rr, ww;
	uint32_t wmask, umask, lmask;
	int32_t shift0, shift1, shiftB, shiftC;
	uint32_t maskB, maskC;
} arma_mt_params;

enum arma_precision {
	ARMA_FP64 = 0,        /* real_type=double */
	ARMA_FP32 = 1         /* FP32 storage, compensated FP32 accumulation */
};

enum arma_noise_mode {
	ARMA_NOISE_SEQUENTIAL = 0,   /* stream 0 over the whole grid (framework=none) */
	ARMA_NOISE_PARTITIONED = 1   /* stream p fills partition p (row-major) */
};

enum arma_noise_semantics {
	ARMA_NOISE_DOUBLE = 0,  /* normal_distribution<double>; FP32 output = float(FP64 noise) */
	ARMA_NOISE_FLOAT = 1    /* normal_distribution<float> (real_type=float build) */
};

enum arma_ar_algorithm {
	ARMA_AR_PLAIN = 0,      /* one work-item per partition, DAG levels as launches */
	ARMA_AR_WAVEFRONT = 1   /* persistent wavefront over (x,y) tiles marching in t */
};

typedef struct arma_ar_desc {
	int32_t shape[3];            /* nt, nx, ny (out_grid num_points) */
	const double* phi;           /* host FP64 coefficients, row-major */
	int32_t phi_shape[3];
	double varwn;                /* white-noise variance */
	int32_t noise_mode;          /* enum arma_noise_mode */
	int32_t noise_semantics;     /* enum arma_noise_semantics */
	int32_t partition[3];        /* partition shape (AR "partition" key) */
	const arma_mt_params* mts;   /* host array of n_mts parameter sets */
	int32_t n_mts;
	const uint32_t* seeds;       /* n_mts seeds; NULL = all zero (no_seed) */
	int32_t precision;           /* enum arma_precision */
	int32_t algorithm;           /* enum arma_ar_algorithm */
	int32_t tile[3];             /* tile hint for ARMA_AR_WAVEFRONT (0 = auto) */
} arma_ar_desc;

typedef struct arma_pot_desc {
	int32_t shape[3];            /* zeta extent nt, nx, ny */
	double delta[3];             /* zeta grid spacing (grid().delta()) */
	double depth;                /* water depth h */
	double lbound[2];            /* domain from (t0, z0) */
	double ubound[2];            /* domain to (t1, z1) */
	int32_t npoints[2];          /* domain npoints (nt_out, nz) */
	int32_t precision;           /* enum arma_precision */
} arma_pot_desc;

/* Target law of the non-linear inertialess transform (the "distribution"
   parameter of transform = nit, src/nonlinear/nit_transform.cc:206-223). */
enum arma_nit_distribution { ARMA_NIT_GRAM_CHARLIER = 0, ARMA_NIT_SKEW_NORMAL = 1 };

typedef struct arma_nit_desc {
	int32_t distribution;        /* enum arma_nit_distribution */
	double skewness, kurtosis;   /* gram_charlier { skewness= kurtosis= } */
	double mean, stdev, alpha;   /* skew_normal { mean= stdev= alpha= } */
	double interval[2];          /* cdf_solver.interval [a, b], a < b */
	double absolute_error;       /* cdf_solver.absolute_error (> 0; default 1e-6) */
	int32_t max_iterations;      /* cdf_solver.max_iterations (> 0; default 100) */
} arma_nit_desc;

typedef struct arma_hip_ctx arma_hip_ctx;

int arma_hip_init(int device, arma_hip_ctx** ctx);
void arma_hip_destroy(arma_hip_ctx* ctx);
const char* arma_hip_last_error(void);
/* Use a caller-owned hipStream_t instead of the context's own stream. */
int arma_hip_set_stream(arma_hip_ctx* ctx, void* stream);
void* arma_hip_get_stream(arma_hip_ctx* ctx);
int arma_hip_synchronize(arma_hip_ctx* ctx);

/* White noise eps ~ N(0, varwn) into eps (device pointer, FP64 or FP32 per
   desc->precision).  Returns ARMA_HIP_EPRNG when n_mts is smaller than the
   number of streams the mode needs, ARMA_HIP_EINVAL when varwn < 0. */
int arma_hip_white_noise(arma_hip_ctx* ctx, const arma_ar_desc* desc, void* eps);

/* The AR recurrence in place over zeta (device pointer holding the noise). */
int arma_hip_ar_recurrence(arma_hip_ctx* ctx, const arma_ar_desc* desc, void* zeta);

/* do_generate(): noise + recurrence.  zeta is a device pointer when
   zeta_on_device != 0, else a host pointer of nt*nx*ny elements. */
int arma_hip_ar_generate(arma_hip_ctx* ctx, const arma_ar_desc* desc, void* zeta,
                         int zeta_on_device);

/* No counterpart in the reference.  For callers that cannot allocate device
   memory (tight serving loops, harnesses that forbid hipMalloc):
   arma_hip_ar_generate with a device pointer zeta and algorithm
   ARMA_AR_WAVEFRONT, whose working buffers (MT records and seeds,
   coefficient tables, tile progress counters, halo mailboxes) come from a
   static device array of 32 MiB per GPU instead of allocations.  Returns
   ARMA_HIP_EINVAL, before enqueuing anything, for a surface whose buffers
   exceed that (the mailboxes grow with tiles x (nt + 30 + order) x halo
   width).  All calls on one GPU share the array: calls that overlap on a GPU
   (other threads or streams) race, and the caller serialises them.  The
   small host tables are uploaded with a stream synchronisation, so a call
   cannot be captured into a HIP graph. */
int arma_hip_ar_generate_noalloc(arma_hip_ctx* ctx, const arma_ar_desc* desc, void* zeta);

/* zeta *= sqrt(gamma0 / var(run)), where var is the sample variance of the
   contiguous run of ceil(nt/2)ceil(nx/2)ceil(ny/2) elements that starts at
   zeta(nt/2, nx/2, ny/2).  factor_out (optional, host) receives the factor. */
int arma_hip_rescale(arma_hip_ctx* ctx, void* zeta, const int32_t shape[3], int32_t precision,
                     double gamma0, double* factor_out);

/* Velocity potential phi(t, z, x, y) of zeta under the linear solver.  Both
   pointers are device pointers when on_device != 0, else host pointers.
   wnmax_out (optional, host, 2*npoints[0] doubles) receives per-slice
   (1/kmin_x, 1/kmin_y).  Returns ARMA_HIP_EINVAL when min z < -depth. */
int arma_hip_linear_potential(arma_hip_ctx* ctx, const arma_pot_desc* desc, const void* zeta,
                              void* phi, int on_device, double* wnmax_out);

/* The same under the high-amplitude solver: replaces
   High_amplitude_solver<T>::operator() (registered as "high_amplitude",
   src/register_all.hh:53; precompute src/velocity/high_amplitude_solver.cc:6-23,
   transforms src/velocity/linear_solver.cc:54-121), whose multiplier is
   zeta_t / (i((zeta_x + zeta_y)/s - zeta_x - zeta_y) - 1/s),
   s = sqrt(1 + zeta_x^2 + zeta_y^2), instead of -zeta_t.  Same arguments,
   layout and errors as arma_hip_linear_potential. */
int arma_hip_high_amplitude_potential(arma_hip_ctx* ctx, const arma_pot_desc* desc,
                                      const void* zeta, void* phi, int on_device,
                                      double* wnmax_out);

/* NIT_transform<T>::transform_realisation (src/nonlinear/nit_transform.cc:
   127-164, called at src/generator/basic_arma_model.cc:107-111 when
   transform = nit): every point x of zeta (device, nt*nx*ny elements of
   precision) becomes the root in interval of F(c) = Phi(x / sqrt(gamma0)),
   F the target CDF, by the reference's bisection (src/linalg.hh:199-213).
   Returns ARMA_HIP_EINVAL for an empty interval, a non-positive absolute
   error or iteration count, or gamma0 < 0.  The reference's transform_ACF
   works on a copy of the ACF that is then dropped
   (src/generator/basic_arma_model.cc:74-79) and changes no output; it has no
   counterpart here. */
int arma_hip_nit_realisation(arma_hip_ctx* ctx, void* zeta, const int32_t shape[3],
                             int32_t precision, double gamma0, const arma_nit_desc* desc);

/* Streaming writer of device arrays to a file in network byte order: the
   zeta.bin format of io::Binary_stream (src/io/binary_stream.hh:43-70), and
   the role of the writer thread of the OpenMP generator
   (src/generator/ar_model_parallel.cc:112-145), which writes finished
   t-slabs while the run goes on.  A copy stream byte-swaps (on the GPU) and
   copies slabs of slab_bytes (0 = 32 MiB) into pinned slots, four in
   flight, and I/O threads write them at their file offsets, while the
   context's stream goes on with later work.  element_size is 4 or 8.  Close
   every writer before destroying its context. */
typedef struct arma_hip_writer arma_hip_writer;
int arma_hip_writer_open(arma_hip_ctx* ctx, const char* path, int32_t element_size,
                         size_t slab_bytes, arma_hip_writer** out);
/* Appends nbytes at the device pointer data once the work enqueued so far on
   the context's stream has finished.  data must stay allocated and unchanged
   until arma_hip_writer_close returns.  Returns the first error of an
   earlier append, if any. */
int arma_hip_writer_append(arma_hip_writer* w, const void* data, size_t nbytes);
/* Waits for every append, closes the file and frees the writer (also on
   error).  bytes_out and seconds_out (optional) receive the bytes written and
   the time from the first appended data being ready on the device to the
   last byte written. */
int arma_hip_writer_close(arma_hip_writer* w, uint64_t* bytes_out, double* seconds_out);

/* Raw tempered uint32 draws of n_streams streams (draws_per_stream each)
   into out (device), stream-major: the bit-exact check of the generator. */
int arma_hip_mt_raw(arma_hip_ctx* ctx, const arma_mt_params* mts, const uint32_t* seeds,
                    int32_t n_streams, int32_t draws_per_stream, uint32_t* out);

/* Per-kernel timing with HIP events on the context stream.  The report has
   one line per kernel, "prfl <kernel> <total microseconds>us <launches>", in
   the style of the reference's profile output (src/profile.hh:70-101). */
int arma_hip_profile(arma_hip_ctx* ctx, int32_t enable);
int arma_hip_profile_report(arma_hip_ctx* ctx, char* buf, size_t len, int32_t reset);

/* mt.dat reading and writing (binary mt_struct dumps followed by nn state
   words; 72-byte header on LP64).  *n_read receives the number of records
   read; returns ARMA_HIP_EPRNG if fewer than n were present. */
int arma_hip_read_mt_config(const char* path, int32_t n, arma_mt_params* out,
                            int32_t* n_read);
int arma_hip_write_mt_config(const char* path, const arma_mt_params* params, int32_t n);

/* Dynamic creation of one parameter set for Mersenne exponent p and word
   size w=32, with the 16-bit id embedded in the low bits of aaa, searched
   from a generator seeded with seed. */
int arma_hip_dcmt_create(int32_t p, int32_t id, uint32_t seed, arma_mt_params* out);
/* Creates n parameter sets with ids 0..n-1 using up to nthreads host threads. */
int arma_hip_dcmt_create_many(int32_t p, int32_t n, uint32_t seed, int32_t nthreads,
                              arma_mt_params* out);

/* ---- FP64 model setup (host) ---- */

enum arma_acf_function {
	ARMA_ACF_STANDING_WAVE = 0,     /* standing_wave_ACF, src/acf.hh:18-37 */
	ARMA_ACF_PROPAGATING_WAVE = 1   /* propagating_wave_ACF, src/acf.hh:39-58 */
};

/* Plain wave profile of the ACF generator (src/generator/plain_wave.hh,
   input names "sin", "cos", "stokes", "standing_wave"). */
enum arma_wave_profile {
	ARMA_WAVE_SIN = 0, ARMA_WAVE_COS = 1, ARMA_WAVE_STOKES = 2, ARMA_WAVE_STANDING = 3
};

typedef struct arma_acf_gen_desc {
	int32_t func;            /* enum arma_wave_profile (default cos) */
	double amplitude;        /* default 1 */
	double velocity;         /* default 1 */
	double beta[2];          /* wave numbers (default (0.8, 0)) */
	double alpha[3];         /* decay (default 0.06 each) */
	double nwaves[3];        /* half extent of the wave's domain (default 1.5 each) */
	int32_t shape[3];        /* ACF shape; any zero: chosen by doubling (default 16 each) */
} arma_acf_gen_desc;

/* ACF_generator<T>::generate (src/generator/acf_generator.cc:130-215), the
   ACF the reference's AR model uses when its acf block has no grid
   (src/generator/basic_arma_model.cc:57): the auto-covariance
   real(IFFT(|FFT(w)|^2))/n^2 (src/arma.cc:81-100) of the plain wave w on the
   domain [-nwaves, nwaves] with 2*shape+1 points, multiplied by
   exp(-sum |alpha (t,x,y)|) and rescaled to the wave's variance, cut to its
   first half in every dimension.  out (capacity cap doubles) receives the
   ACF, shape_out its shape and delta_out its grid spacing
   nwaves/(shape_out-1) (Grid{shape, nwaves}).  Returns ARMA_HIP_EINVAL when
   cap is too small.  Runs on the host: the sizes are at most 129^3. */
int arma_hip_acf_generate(const arma_acf_gen_desc* desc, double* out, size_t cap,
                          int32_t shape_out[3], double delta_out[3]);

/* arma_hip_acf_generate with the auto-covariance on the GPU of ctx: the
   circular auto-covariance (1/n) sum_x w(x) w((x - d) mod s) of the decayed
   wave, which real(IFFT(|FFT(w)|^2))/n^2 equals, summed directly in FP64 for
   every lag d of the first half (src/arma.cc:81-100).  Host inputs and
   outputs; waits for the context stream. */
int arma_hip_acf_generate_device(arma_hip_ctx* ctx, const arma_acf_gen_desc* desc, double* out,
                                 size_t cap, int32_t shape_out[3], double delta_out[3]);

/* Closed-form ACF on a grid of shape points with spacing delta
   (src/acf.hh:18-62 via src/bits/acf_wrapper.cc:41-70). */
int arma_hip_acf(int32_t func, const double delta[3], const int32_t shape[3], double amplitude,
                 double velocity, const double alpha[3], const double beta[2], double* out);

/* Choi order-recursive Yule-Walker solver (Yule_walker_solver<T>::solve,
   src/yule_walker.cc:420-539) on an n^3 autocovariance cube; phi_out holds
   max_order^3 doubles and receives phi row-major with shape order_out. */
int arma_hip_yule_walker_choi(const double* acf, int32_t n, int32_t max_order, double var_eps,
                              int32_t determine_order, int32_t chop, double chop_eps,
                              double* phi_out, int32_t order_out[3], double* varwn_solver);

/* Gauss-elimination coefficients (AR_model<T>::determine_coefficients_gauss,
   src/generator/ar_model.cc:97-135); phi_out holds prod(order) doubles. */
int arma_hip_yule_walker_gauss(const double* acf, const int32_t acf_shape[3],
                               const int32_t order[3], double* phi_out);

/* AR_model<T>::white_noise_variance (src/generator/ar_model.cc:55-61). */
double arma_hip_white_noise_variance(const double* acf, const int32_t acf_shape[3],
                                     const double* phi, const int32_t phi_shape[3]);

/* The same two setup steps in FP64 on the GPU of ctx (host inputs and
   outputs).  The Choi recursion runs as one workgroup in the host solver's
   order of operations; its coefficients, order and variance equal
   arma_hip_yule_walker_choi's to the last bit. */
int arma_hip_yule_walker_choi_device(arma_hip_ctx* ctx, const double* acf, int32_t n,
                                     int32_t max_order, double var_eps, int32_t determine_order,
                                     int32_t chop, double chop_eps, double* phi_out,
                                     int32_t order_out[3], double* varwn_solver);
int arma_hip_white_noise_variance_device(arma_hip_ctx* ctx, const double* acf,
                                         const int32_t acf_shape[3], const double* phi,
                                         const int32_t phi_shape[3], double* out);
/* Many Choi solves in one launch, one workgroup per model (the models of a
   batch of surfaces; each result equals the single solve's to the last
   bit).  Model i reads the n[i]^3 cube at acf + acf_off[i] and writes at
   most max_order[i]^3 doubles at phi_out + phi_off[i]; order_out[3i..3i+2]
   receives its shape, varwn_solver[i] (optional) the solver's variance and
   status[i] ARMA_HIP_OK or the model's own error. */
int arma_hip_yule_walker_choi_device_many(arma_hip_ctx* ctx, int32_t count, const double* acf,
                                          const int64_t* acf_off, const int32_t* n,
                                          const int32_t* max_order, double var_eps,
                                          int32_t determine_order, int32_t chop,
                                          double chop_eps, double* phi_out,
                                          const int64_t* phi_off, int32_t* order_out,
                                          double* varwn_solver, int32_t* status);

/* validate_process (src/arma.cc:15-70): the roots of 1 - phi_1 x - phi_2 x^2
   - ... over the row-major coefficients (phi(0,0,0) replaced by the 1) must
   all lie outside the unit circle.  n_bad receives the number that do not;
   roots (optional) receives up to roots_len/2 (re, im) pairs.  Returns
   ARMA_HIP_ERUNTIME with the reference's message when a root is inside or on
   the circle, or when the roots cannot be found. */
int arma_hip_validate_process(const double* phi, const int32_t shape[3], int32_t* n_bad,
                              double* roots, int32_t roots_len);

/* get_partition_shape (src/generator/ar_model.cc:23-50). */
void arma_hip_partition_shape(const int32_t partition[3], const int32_t grid[3],
                              const int32_t order[3], int32_t parallelism, int32_t out[3]);

/* ---- Batching: many surfaces in one persistent launch ----
 *
 * No counterpart in the reference, whose driver makes one surface per run
 * (src/arma_driver.cc:29-87).  Every member is generated with exactly the
 * semantics of arma_hip_ar_generate followed by arma_hip_rescale on its own
 * descriptor, and its values do not depend on the rest of the batch, the
 * tile shape or the schedule.  A batch owns its device tables and scratch;
 * generate, capture and replay only enqueue kernels on the context stream.
 *
 * The recurrence launch is persistent: its workgroups wait on each other, so
 * all of them must be resident together.  Recurrence launches on one device
 * therefore run one at a time, whatever context or stream enqueues them
 * (each waits for the previous one); noise, rescale and other work still
 * overlap.  A captured graph is outside that chain: do not replay it while
 * another batch's recurrence runs on the device.  The waits inside a launch
 * are bounded (10 s without progress); a launch that gave up makes the next
 * call that waits for it (arma_hip_batch_factors, arma_hip_batch_tune, the
 * server's lanes) return ARMA_HIP_ERUNTIME. */

enum arma_batch_schedule {
	ARMA_BATCH_AUTO = 0,
	ARMA_BATCH_SURFACE_MAJOR = 1,  /* surfaces by decreasing work, tile rows in order */
	ARMA_BATCH_INTERLEAVED = 2     /* tile rows dealt round-robin over the surfaces */
};

typedef struct arma_batch_config {
	int32_t tile_x;        /* 8, 16 or 32: tiles of tile_x x (256 / tile_x) columns; 0 = auto */
	int32_t schedule;      /* enum arma_batch_schedule */
	int32_t slots_per_cu;  /* resident workgroups per CU; 0 = as many as fit */
} arma_batch_config;

typedef struct arma_batch_desc {
	int32_t n_surfaces;
	const arma_ar_desc* surfaces;  /* one generator per surface (algorithm, tile ignored) */
	const double* gamma0;          /* per-surface rescale target acf(0,0,0); NULL or <= 0 = none */
	const int32_t* history;        /* per-surface count of leading slices that the caller
	                                  writes into the arena before generate and that enter
	                                  the recurrence as given values (a realtime stream's
	                                  last slices); noise, partition and output cover
	                                  slices history..nt-1 only.  NULL = none */
	int32_t precision;             /* enum arma_precision, shared by all surfaces */
	arma_batch_config config;      /* all zero: the tuned configuration of an identical
	                                  batch on this context, else 16 x 16 surface-major */
} arma_batch_desc;

typedef struct arma_batch_trial {
	arma_batch_config config;
	int32_t grid;                  /* workgroups launched */
	double ms;                     /* best AR launch time over the repetitions */
} arma_batch_trial;

typedef struct arma_hip_batch arma_hip_batch;

int arma_hip_batch_create(arma_hip_ctx* ctx, const arma_batch_desc* desc, arma_hip_batch** out);
void arma_hip_batch_destroy(arma_hip_batch* batch);
/* offsets[n_surfaces + 1]: the first arena element of every surface, then
   the arena length in elements. */
int arma_hip_batch_layout(const arma_hip_batch* batch, int64_t* offsets);
/* Noise, recurrence and rescale of every surface into arena (device). */
int arma_hip_batch_generate(arma_hip_ctx* ctx, arma_hip_batch* batch, void* arena);
/* Records the launch sequence of generate for arena into a HIP graph held
   by the batch; replay launches it on the context stream. */
int arma_hip_batch_capture(arma_hip_ctx* ctx, arma_hip_batch* batch, void* arena);
int arma_hip_batch_replay(arma_hip_ctx* ctx, arma_hip_batch* batch);
/* Autotuner: times the recurrence launch of every candidate tile shape,
   schedule and residency on this batch (noise regenerated before each
   timing, reps timings each), keeps the fastest, and remembers it on the
   context for identical batches.  trials (optional) receives up to
   max_trials timings; n_trials the number of candidates timed. */
int arma_hip_batch_tune(arma_hip_ctx* ctx, arma_hip_batch* batch, void* arena, int32_t reps,
                        arma_batch_trial* trials, int32_t max_trials, int32_t* n_trials);
int arma_hip_batch_get_config(const arma_hip_batch* batch, arma_batch_config* config,
                              int32_t* grid);
int arma_hip_batch_set_config(arma_hip_ctx* ctx, arma_hip_batch* batch,
                              const arma_batch_config* config);
/* Post-rescale factors of the last generate or replay (0 where none); waits
   for the context stream and reports a stalled recurrence launch. */
int arma_hip_batch_factors(arma_hip_ctx* ctx, arma_hip_batch* batch, double* out);

/* ---- Serving: continuous batching of surface and potential requests ----
 *
 * No reference counterpart (the reference makes one surface per run,
 * src/arma_driver.cc:29-87).  A server owns `lanes` contexts, each with a
 * HIP stream and a host thread.  Requests queue as they arrive; a scheduler
 * admits them in arrival order while their device bytes (outputs plus the
 * working set while in flight) fit hbm_budget, and gives each admitted
 * group to a free lane as one batch (one persistent recurrence launch over
 * its surfaces, then their potentials), so work from different lanes
 * overlaps on the device.  Every result equals the one the single-surface
 * entry points give for the same descriptor.
 *
 * A realtime stream (open_stream) keeps its last phi_shape[0] - 1 slices on
 * the device; each STREAM_CHUNK request generates its next `chunk` slices.
 * Chunk k draws its noise with partition (chunk, px, py) and seeds
 * seeds[0] + k * columns + column, so the chunks concatenate to exactly the
 * surface that arma_hip_ar_generate makes with partition (chunk, px, py),
 * mts[p % columns] and seeds seeds[0] + p. */

enum arma_request_kind {
	ARMA_REQ_SURFACE = 0,       /* noise + recurrence + rescale, optionally the potential */
	ARMA_REQ_STREAM_CHUNK = 1   /* the next chunk of a realtime stream */
};

typedef struct arma_server_config {
	int32_t lanes;              /* batches in flight at once (1..16) */
	int32_t max_batch;          /* most requests per batch; 0 = 64 */
	int64_t hbm_budget;         /* device bytes admitted at once */
	int32_t precision;          /* enum arma_precision of every request */
	arma_batch_config batch;    /* tile shape / schedule of every batch (0 = defaults) */
} arma_server_config;

typedef struct arma_request {
	int32_t kind;               /* enum arma_request_kind */
	arma_ar_desc surface;       /* SURFACE: the generator (copied at submit) */
	double gamma0;              /* SURFACE: rescale target acf(0,0,0); <= 0 = none */
	int32_t want_potential;     /* SURFACE: also the linear potential described by: */
	arma_pot_desc potential;    /*   (its shape must equal the surface's) */
	int64_t stream;             /* STREAM_CHUNK: handle from arma_hip_server_open_stream */
	double arrival;             /* arrival time on arma_hip_server_clock(); 0 = now */
} arma_request;

typedef struct arma_completion {
	int64_t id;
	int32_t status;             /* ARMA_HIP_OK or the error code of its batch */
	int32_t kind;
	int64_t stream, chunk;      /* STREAM_CHUNK: its stream and chunk index */
	const void* zeta;           /* device surface (or chunk), valid until release */
	const void* phi;            /* device potential (t, z, x, y), or NULL */
	int32_t shape[3];           /* of zeta */
	double t_arrival, t_admit, t_done;   /* arma_hip_server_clock() seconds */
	int32_t batch_size;         /* requests in the batch it ran in */
	int32_t lane;
} arma_completion;

typedef struct arma_server_stats {
	int64_t in_use, peak, budget;       /* admitted device bytes */
	int64_t batches, batched_requests, completed, pending;
	int32_t busy_lanes;
} arma_server_stats;

typedef struct arma_hip_server arma_hip_server;

double arma_hip_server_clock(void);
int arma_hip_server_create(int device, const arma_server_config* config,
                           arma_hip_server** out);
/* Drains the admitted work, drops what is still queued, frees everything. */
void arma_hip_server_destroy(arma_hip_server* server);
/* desc: shape[1..2] = (nx, ny), phi, varwn, noise semantics, partition[1..2]
   = (px, py) when noise_mode is partitioned, mts (one per column of
   partitions), seeds[0] = base seed; shape[0] is ignored. */
int arma_hip_server_open_stream(arma_hip_server* server, const arma_ar_desc* desc,
                                int32_t chunk, int64_t* stream);
int arma_hip_server_close_stream(arma_hip_server* server, int64_t stream);
/* ARMA_HIP_ENOMEM when the request alone exceeds the budget. */
int arma_hip_server_submit(arma_hip_server* server, const arma_request* request, int64_t* id);
/* Up to max completions, waiting up to wait_ms (-1: until one is there). */
int arma_hip_server_poll(arma_hip_server* server, arma_completion* out, int32_t max,
                         int32_t* n, int32_t wait_ms);
/* Frees a polled request's outputs and returns its bytes to the budget. */
int arma_hip_server_release(arma_hip_server* server, int64_t id);
int arma_hip_server_stats(arma_hip_server* server, arma_server_stats* stats);
/* Copies bytes from a result pointer (device) to dst (host or device);
   complete on return. */
int arma_hip_server_copy(arma_hip_server* server, void* dst, const void* src, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif
