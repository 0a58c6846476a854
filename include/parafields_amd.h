/*
 * parafields_amd.h -- C ABI of the MI355X (gfx950) circulant-embedding
 * Gaussian-random-field engine.
 *
 * The entry points replace the arithmetic that parafields-core's
 * `parafields::Matrix` performs through FFTW-MPI (include/parafields/matrix.hh).
 * They take plain pointers, sizes and POD structs; no C++ types cross the
 * boundary and no exception escapes it. Every call returns a status code
 * (PF_OK or a negative PF_E* value); pf_last_error() returns a message for the
 * most recent failure on the calling thread.
 *
 * One plan = one covariance model on one grid (the state a reference `Matrix`
 * holds: traits, lazily built embedding eigenvalues, field work buffers).
 * One GPU holds the whole extended grid; calls on one plan are serialised by
 * the plan; distinct plans may be used from distinct threads.
 *
 * Pointers: every data argument is accompanied by a pf_mem flag. PF_MEM_HOST
 * pointers are staged through plan-owned device buffers; PF_MEM_DEVICE
 * pointers must be device allocations on the plan's device and are used in
 * place (the throughput path).
 */
#ifndef PARAFIELDS_AMD_H
#define PARAFIELDS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (reference exception types they map to) ------------- */
#define PF_OK 0
#define PF_ENEG (-1)    /* NegativeEigenvalueError, include/parafields/exceptions.hh:8-15 */
#define PF_EINVAL (-2)  /* std::runtime_error on bad configuration, matrix.hh:258-260 */
#define PF_EHIP (-3)    /* device runtime failure */
#define PF_ENOMEM (-4)  /* device allocation failure */
#define PF_ESTATE (-5)  /* call not valid in the plan's current state */

/* ---- enums ------------------------------------------------------------- */
/* stochastic.covariance, covariance.hh:296-823 */
enum pf_covariance {
  PF_COV_EXPONENTIAL = 0,
  PF_COV_GAUSSIAN = 1,
  PF_COV_SPHERICAL = 2,
  PF_COV_SEPARABLE_EXPONENTIAL = 3,
  PF_COV_MATERN = 4,   /* general nu: own Bessel K_nu (reference needs GSL) */
  PF_COV_MATERN32 = 5,
  PF_COV_MATERN52 = 6,
  PF_COV_DAMPED_OSCILLATION = 7,
  PF_COV_GAMMA_EXPONENTIAL = 8,
  PF_COV_CAUCHY = 9,
  PF_COV_GENERALIZED_CAUCHY = 10,
  PF_COV_CUBIC = 11,
  PF_COV_WHITE_NOISE = 12,
  PF_COV_CUSTOM = 13   /* host-evaluated values uploaded with pf_eigs_custom */
};

/* stochastic.anisotropy, matrix.hh:476-494 / covariance.hh:18-138 */
enum pf_anisotropy { PF_ANISO_NONE = 0, PF_ANISO_AXIPARALLEL = 1, PF_ANISO_GEOMETRIC = 2 };

/* MatrixBackend template argument of Matrix, matrix.hh:55-60; PF_MATRIX_DFT is
   DFTMatrixBackend (dftmatrixbackend.hh:19-384, the 1-D default) */
enum pf_matrix_backend { PF_MATRIX_DCT = 0, PF_MATRIX_R2C = 1, PF_MATRIX_DFT = 2 };

/* FieldBackend template argument of Matrix, matrix.hh:55-60; PF_FIELD_DFT is
   DFTFieldBackend (dftfieldbackend.hh:19-477, the 1-D default): one draw of
   2 N_ext normals gives two fields, the real part and the imaginary part (the
   reference's spare field, matrix.hh:336, :411-420) */
enum pf_field_backend { PF_FIELD_R2C = 0, PF_FIELD_DCTDST = 1, PF_FIELD_DFT = 2 };

/* arithmetic of the per-realization transforms (eigenvalues are always FP64) */
enum pf_precision { PF_FP64 = 0, PF_FP32 = 1 };

/* embedding.periodization, matrix.hh:515-580: classical mirroring
   (computeMatrixEntriesWithMirroring, :615-638) or the smooth merge embedding
   (computeMatrixEntriesWithMerge, :652-775), fold (computeMatrixEntriesWithFold,
   :790-888) or cofold (modifyMatrixEntriesWithCofold on the classical entries,
   :890-1013; not in batches: its offset is reduced on the host). */
enum pf_periodization {
  PF_PERIOD_CLASSICAL = 0,
  PF_PERIOD_MERGE = 1,
  PF_PERIOD_FOLD = 2,
  PF_PERIOD_COFOLD = 3
};
/* embedding.sigmoid: SmoothstepSigmoid (default) or SmoothSigmoid, covariance.hh:147-290 */
enum pf_sigmoid { PF_SIGMOID_SMOOTHSTEP = 0, PF_SIGMOID_SMOOTH = 1 };
/* stochastic.sigmoidCombine, matrix.hh:660-671 */
enum pf_sigmoid_combine {
  PF_COMBINE_RADIAL_PLUS = 0, /* default */
  PF_COMBINE_RADIAL = 1,
  PF_COMBINE_TENSOR = 2
};

/* randomField.transform, mutators.hh:126-163 */
enum pf_transform {
  PF_TRANSFORM_NONE = 0,
  PF_TRANSFORM_LOGNORMAL = 1,
  PF_TRANSFORM_FOLDEDNORMAL = 2,
  PF_TRANSFORM_SIGN = 3,      /* param = threshold (reference: 0) */
  PF_TRANSFORM_BOXCOX = 4     /* param = transform.boxCoxLambda */
};

enum pf_mem { PF_MEM_HOST = 0, PF_MEM_DEVICE = 1 };

/* operator applied by pf_matvec */
enum pf_matvec_op {
  PF_OP_MULTIPLY = 0,        /* Matrix::operator*, matrix.hh:142-151 (C x / N_ext) */
  PF_OP_ROOT = 1,            /* Matrix::multiplyRoot, matrix.hh:167-176 */
  PF_OP_INVERSE_EMBED = 2    /* multiplyInverseExtended, matrix.hh:1587-1646 */
};

/* ---- descriptor: the configuration keys a reference Matrix reads ------- */
typedef struct pf_desc {
  int32_t dim;                 /* 1 (DFT backends), 2 or 3 */
  int32_t cells[3];            /* grid.cells */
  double extensions[3];        /* grid.extensions */
  int32_t embedding_factor;    /* embedding.factor (default 2) */
  int32_t periodic;            /* randomField.periodic (forces factor 1) */
  int32_t covariance;          /* enum pf_covariance */
  double variance;             /* stochastic.variance */
  int32_t anisotropy;          /* enum pf_anisotropy */
  double corr_length[9];       /* stochastic.corrLength: 1, dim or dim*dim values (row-major) */
  double matern_nu;            /* stochastic.maternNu */
  double exp_gamma;            /* stochastic.expGamma */
  double cauchy_alpha;         /* stochastic.cauchyAlpha */
  double cauchy_beta;          /* stochastic.cauchyBeta */
  int32_t matrix_backend;      /* enum pf_matrix_backend */
  int32_t field_backend;       /* enum pf_field_backend */
  int32_t transposed;          /* fftw.transposed: -1 = reference default (dim > 1 with the
                                  fieldtraits.hh:229-267 exceptions), 0 or 1 */
  int32_t approximate;         /* embedding.approximate */
  double threshold;            /* embedding.threshold (default 1e-14) */
  int32_t precision;           /* enum pf_precision */
  int32_t cg_iterations;       /* randomField.cgIterations (default 100) */
  double nugget;               /* added to C(0) (custom-iso hook of the reference); 0 = none */
  int32_t r2c_check_all;       /* R2C matrix: 0 = the reference's check/clamp over the interleaved
                                  buffer (matrix.hh:270-291 reading r2cmatrixbackend.hh:378-400, which
                                  covers Re and Im of the first half of the entries only); 1 = check
                                  and clamp the real part of every entry (not reference behaviour) */
  /* Embedding (matrix.hh:515-775). A zeroed tail gives the reference defaults. */
  int32_t periodization;       /* enum pf_periodization */
  int32_t sigmoid;             /* enum pf_sigmoid (merge, fold, cofold) */
  int32_t sigmoid_combine;     /* enum pf_sigmoid_combine (merge only) */
  int32_t merge_recursions;    /* embedding.mergeRecursions (merge), embedding.foldRecursions (fold)
                                  or embedding.cofoldRecursions (cofold, default 1): 0 = the reference
                                  default (its value 99:
                                  a count per covariance family, matrix.hh:679-705), k > 0, or -1 for
                                  zero recursions */
  double sigmoid_start;        /* embedding.sigmoidStart; 0 = default 1 */
  double sigmoid_end;          /* embedding.sigmoidEnd; 0 = default factor - 1 */
  double fold_max_factor;      /* embedding.foldMaxFactor (fold, default 1) or
                                  embedding.cofoldMaxFactor (cofold, default 5); 0 = the default */
} pf_desc;

/* eigenvalue classification, matrix.hh:266-305 */
typedef struct pf_eig_stats {
  int64_t small;       /* 0 <= ... entries with threshold <= v < 1e-6 */
  int64_t zero;        /* |v| < threshold */
  int64_t negative;    /* v <= -threshold */
  double smallest;     /* minimum entry */
  int64_t entries;     /* entries classified (localMatrixSize after the transform) */
} pf_eig_stats;

/* sizes derived from the descriptor (fieldtraits.hh:217-314) */
typedef struct pf_sizes {
  int32_t dim;
  int32_t cells[3];            /* n_i (1 for unused dims) */
  int32_t extended[3];         /* N_i = factor * n_i */
  int32_t transposed;          /* effective fftw.transposed */
  int64_t domain_size;         /* prod n_i */
  int64_t extended_size;       /* prod N_i */
  int64_t eig_entries;         /* entries of the FP64 eigenvalue array */
  int64_t noise_per_field;     /* N(0,1) samples one realization consumes */
} pf_sizes;

/* conjugate-gradient modes for pf_cg */
enum pf_cg_mode {
  PF_CG_REFERENCE = 0,   /* Matrix::multiplyInverse + innerCG, matrix.hh:189-206, :1315-1432 */
  PF_CG_PLAIN = 1,       /* unpreconditioned CG, relative-residual stop */
  PF_CG_EMBED_PCG = 2,   /* CG preconditioned by the embedding inverse, relative-residual stop */
  PF_CG_LEARNED_PCG = 3  /* CG preconditioned by a learned radial spectral multiplier */
};

typedef struct pf_cg_opts {
  int32_t mode;            /* enum pf_cg_mode */
  int32_t max_iterations;  /* 0 = desc.cg_iterations */
  double rel_tol;          /* stop at ||b - A x|| / ||b|| <= rel_tol (modes 1-3) */
  const float* bin_log_multipliers; /* PF_CG_LEARNED_PCG: nbins host values */
  int32_t nbins;
} pf_cg_opts;

typedef struct pf_cg_stats {
  int32_t iterations;
  double rel_residual;     /* FP64 true residual ||b - A x|| / ||b|| at exit */
  double seconds;
} pf_cg_stats;

/* ---- plan lifecycle ------------------------------------------------------ */

/* Replaces Matrix(traits) + update(), matrix.hh:98-133 and the size logic of
 * RandomFieldTraits::update, fieldtraits.hh:217-314. Validates the
 * configuration and allocates nothing large yet (eigenvalues are lazy, as in
 * the reference). */
int pf_plan_create(const pf_desc* desc, void** plan);
int pf_plan_destroy(void* plan);
int pf_plan_sizes(void* plan, pf_sizes* sizes);
const char* pf_last_error(void);
int pf_device_sync(void* plan);

/* Replaces Matrix::fillTransformedMatrix(std::string), matrix.hh:223-318:
 * covariance evaluation on the embedding grid, forward transform divided by
 * N_ext, classification with clamping of negatives, and PF_ENEG when
 * negatives exist and desc.approximate is 0 (stats are filled either way).
 * Idempotent: a second call returns the cached statistics. */
int pf_eigs(void* plan, pf_eig_stats* stats);

/* Replaces Matrix::fillTransformedMatrix(Covariance&&) with a user callable
 * (randomfield.hh:440-451, matrix.hh:261-264): `values` holds the covariance
 * at every entry of the embedding grid the matrix backend evaluates
 * (pf_sizes.eig_entries entries for PF_MATRIX_DCT; for PF_MATRIX_R2C the
 * padded real array of 2*(N0/2+1)*N1*N2 entries), in the reference's
 * x-fastest order. */
int pf_eigs_custom(void* plan, const double* values, pf_eig_stats* stats);

/* Copies the FP64 eigenvalues in the reference's final layout (DCT: the
 * (N/2+1)^d array, 3D transposed as the reference stores it; R2C: the real
 * part array of (N0/2+1)*N1*N2 entries). n = number of entries to copy. */
int pf_eigs_copy(void* plan, double* out, int64_t n);

/* Replaces Matrix::generateField (matrix.hh:330-421) followed by
 * RandomField::bulkEvaluate's value transform (randomfield.hh:664-695,
 * mutators.hh:126-163) for `batch` realizations.
 * Noise source: if `noise` is non-NULL it holds batch * noise_per_field
 * N(0,1) samples (always host FP64), consumed in the reference's order
 * (the RNG `sample()` sequence of generateField); otherwise the device draws
 * Philox4x32-10 normals keyed by `seed`, realization first_realization + b.
 * Output: batch * domain_size values, x fastest, each realization contiguous;
 * FP64 for PF_FP64 plans, FP32 for PF_FP32 plans. */
int pf_generate(void* plan, const double* noise, uint64_t seed, uint64_t first_realization,
                int32_t batch, int32_t transform, double transform_param, void* out,
                int32_t out_mem);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): pf_generate with device Philox noise (seed, realization
 * first_realization + b) into device memory `out`, whose pass work arrays live
 * in one static library buffer of PF_NOALLOC_SCRATCH_BYTES per GPU (held
 * from the library's load until the process ends) instead of plan-owned
 * allocations. R2C field backend only. The plan must be prepared by pf_eigs
 * (PF_ESTATE otherwise); a realization whose work arrays exceed the buffer
 * gives PF_EINVAL; a batch runs in groups that fit. Results equal
 * pf_generate's bit for bit. Every call on one GPU shares that GPU's buffer:
 * calls that overlap on one GPU (from several threads or plans) race. */
#define PF_NOALLOC_SCRATCH_BYTES (64u << 20)
int pf_generate_noalloc(void* plan, uint64_t seed, uint64_t first_realization, int32_t batch,
                        int32_t transform, double transform_param, void* out);

/* No counterpart in the reference; its analogue is FFTW's planner with wisdom
 * files (fftw.useWisdom, fftw.measure: dftmatrixbackend.hh:52-70, :164-185).
 * Launch choices of a plan's generation that change no result beyond
 * rounding (tile_div, radix_order, lines) or none at all (chunks, overlap,
 * sub_batch); 0 (overlap: -1) keeps the default. */
typedef struct pf_tuning {
  int32_t tile_div;    /* R2C field, y and x passes: tiles of L / tile_div lines (1 or 2) */
  int32_t chunks;      /* R2C field: chunks of the two-stream pipeline (at most 8) */
  int32_t overlap;     /* R2C field: two-stream pipeline 0 off, 1 on, -1 default */
  int32_t lines;       /* generic kernels: at most this many lines per LDS tile */
  int32_t radix_order; /* generic kernels: 0 radix 8 first, 1 reversed, 2 radix 4 */
  int32_t sub_batch;   /* DCT/DST and DFT fields: realizations per sub-batch */
} pf_tuning;
typedef struct pf_tune_result {
  pf_tuning best;
  double ms;           /* one pf_generate of `batch` fields under `best` */
  double default_ms;   /* the same under the defaults (0 for a cache hit) */
  int32_t candidates;  /* configurations timed (0 for a cache hit) */
  int32_t from_cache;
} pf_tune_result;
/* The autotuner: times pf_generate of `batch` device-Philox realizations into
 * `out` (device memory, batch * domain_size values) under each candidate
 * choice, keeps the fastest in the plan and reports it. cache_path
 * (optional): a text file of earlier results keyed by device and field
 * shape; a hit applies the stored choice without timing, a miss appends. */
int pf_autotune(void* plan, int32_t batch, uint64_t seed, int32_t transform, double transform_param,
                void* out, const char* cache_path, pf_tune_result* result);
int pf_set_tuning(void* plan, const pf_tuning* tuning);
int pf_get_tuning(void* plan, pf_tuning* tuning);

/* Replaces Matrix::operator* / multiplyRoot / multiplyInverseExtended
 * (matrix.hh:142-176, :1443-1646) on `batch` vectors of domain_size values
 * (element type as in pf_generate; x and y may alias). */
int pf_matvec(void* plan, int32_t op, const void* x, void* y, int32_t batch, int32_t mem);

/* Replaces Matrix::multiplyInverse (matrix.hh:189-206 with innerCG
 * :1315-1432) in PF_CG_REFERENCE mode; other modes solve (C/N_ext) x = b to a
 * relative residual. FP64 vectors of domain_size values. */
int pf_cg(void* plan, const double* b, double* x, int32_t mem, const pf_cg_opts* opts,
          pf_cg_stats* stats);

/* Elementwise value transform (mutators.hh:126-163) of n values in place. */
int pf_value_transform(void* plan, int32_t transform, double param, void* data, int64_t n,
                       int32_t mem);

/* Device Philox4x32-10 stream used by pf_generate when noise is NULL:
 * writes n normals (FP64) of realization `realization` starting at sample
 * index `first`. Exposed for statistical checks and parity tests. */
int pf_philox_normals(void* plan, uint64_t seed, uint64_t realization, uint64_t first,
                      int64_t n, double* out, int32_t mem);
/* Raw Philox4x32-10 blocks: out[4*i..4*i+3] = philox(counter = (ctr0 + i,
 * ctr1, ctr2, ctr3), key). */
int pf_philox_raw(void* plan, const uint32_t ctr[4], const uint32_t key[2], int64_t nblocks,
                  uint32_t* out, int32_t mem);

/* ---- learned spectral preconditioner (no counterpart in the reference) ---
 * A small MLP maps covariance parameters to per-radial-frequency-bin
 * log-multipliers L_b; PF_CG_LEARNED_PCG applies
 * M^{-1} = R F^{-1} diag(exp(L_{bin(k)})) F E (bin(k): radius of the folded
 * spectral position |k/(N/2)|/sqrt(dim), nbins equal bins) as the
 * preconditioner of multiplyInverse; plain CG stays the fallback and the FP64
 * true residual judges convergence. */

/* Training targets from the plan's own FP64 embedding eigenvalues (DCT
 * backend): L*_b = -log(mean over bin b of max(lambda_k, floor_value)). */
int pf_spectral_targets(void* plan, int32_t nbins, double floor_value, double* log_targets);

/* FP32 MLP with layer sizes dims[0..ndims-1], exact-GELU hidden layers and a
 * linear output layer; LeCun-normal initialisation from `seed`. Parameters
 * are stored layer by layer as W (rows = outputs, row-major) then b. */
int pf_mlp_create(const int32_t* dims, int32_t ndims, uint64_t seed, void** mlp);
int pf_mlp_destroy(void* mlp);
int pf_mlp_nparams(void* mlp, int64_t* n);
/* set = 1 copies `values` into the network, set = 0 reads them out. */
int pf_mlp_params(void* mlp, float* values, int64_t n, int32_t set);
/* y[batch x dims[last]] = MLP(x[batch x dims[0]]) (host pointers). */
int pf_mlp_forward(void* mlp, const float* x, int32_t batch, float* y);
/* Gradient of the mean squared error mean((MLP(x) - y)^2) w.r.t. every
 * parameter (host pointers), and the loss. */
int pf_mlp_gradients(void* mlp, const float* x, const float* y, int32_t batch, float* grads,
                     double* loss);
/* `steps` full-batch Adam steps (beta 0.9 / 0.999, eps 1e-8) at rate lr. */
int pf_mlp_train(void* mlp, const float* x, const float* y, int32_t batch, int32_t steps, float lr,
                 double* final_loss);

/* ---- uneven batches (no counterpart in the reference) --------------------
 * n independent fields, each with its own descriptor (own grid and
 * covariance; all of one precision), one realization each: per-field FP64
 * eigenvalues followed by the generation passes, enqueued on one stream.
 * With use_graph the whole sequence is captured once into a hipGraph (keyed
 * by seed, first_realization, transform, param and the output pointer) and
 * replayed as one launch. Field i uses Philox realization first_realization
 * + i and is written at out + offsets[i] (pf_batch_offsets, n+1 entries).
 * stats (n entries, may be NULL) receives each field's eigenvalue
 * classification; PF_ENEG is returned if any field has negative eigenvalues
 * without desc.approximate (its field is still written). */
int pf_batch_create(const pf_desc* descs, int32_t n, void** batch);
int pf_batch_destroy(void* batch);
int pf_batch_offsets(void* batch, int64_t* offsets);
int pf_batch_generate(void* batch, uint64_t seed, uint64_t first_realization, int32_t transform,
                      double transform_param, void* out, int32_t out_mem, int32_t use_graph,
                      pf_eig_stats* stats);

/* Batch packing: the fields are bin-packed by algorithmic bytes onto
 * `branches` concurrent graph branches (streams forked from and joined to the
 * batch stream; never more branches than fields). pf_batch_autotune times
 * one replay for 1, 2, 4, 8 and 16 branches (up to the number of fields) on
 * the real batch and keeps the fastest (best, best_ms out). */
int pf_batch_set_branches(void* batch, int32_t branches);
int pf_batch_autotune(void* batch, uint64_t seed, uint64_t first_realization, int32_t transform,
                      double transform_param, void* out, int32_t out_mem, int32_t* best,
                      double* best_ms);

/* Per-field device times (no counterpart in the reference): with timing on,
 * an eager pf_batch_generate (use_graph = 0) brackets every field
 * (eigenvalues + realization, serially on the batch stream) by HIP events
 * (on a batch's first call, in a second pass after the one that allocates);
 * pf_batch_field_times then writes the n durations in ms (PF_ESTATE if the
 * last generate call was not such a pass). */
int pf_batch_set_field_timing(void* batch, int32_t on);
int pf_batch_field_times(void* batch, double* ms);

/* Observability (the analogue of the reference's randomField.verbose
 * prints, fieldtraits.hh:180, matrix.hh:301-305): when on, every kernel
 * launch of a call is bracketed by HIP events on the plan's stream and
 * pf_last_timings returns the per-launch device times (ms) of the last call
 * (names ';'-separated). Also switched on by the environment variable
 * PF_PROFILE. No counterpart in the reference. */
int pf_set_profiling(void* plan, int32_t on);
int pf_last_timings(void* plan, char* names, int32_t names_len, double* ms, int32_t max_n,
                    int32_t* n_out);

#ifdef __cplusplus
}
#endif

#endif /* PARAFIELDS_AMD_H */
