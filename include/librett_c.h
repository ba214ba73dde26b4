/*
 * librett_c.h -- C-ABI twin of librett.h for foreign-function callers
 * (Python ctypes, Go cgo, Java JNI ...): plain pointers and sizes, the stream
 * passed by value as an opaque pointer, no C++ types.
 *
 * Each librett_c_* entry point below names the reference interface it
 * replaces. The entry points marked "extension" have no reference
 * counterpart; they expose plan introspection and host-side planning for
 * tests, benchmarks and tuning tools.
 */
#ifndef LIBRETT_C_H
#define LIBRETT_C_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same values as librettResult (reference src/librett.h:43-50). */
enum {
  LIBRETT_C_SUCCESS = 0,
  LIBRETT_C_INVALID_PLAN = 1,
  LIBRETT_C_INVALID_PARAMETER = 2,
  LIBRETT_C_INVALID_DEVICE = 3,
  LIBRETT_C_INTERNAL_ERROR = 4,
  LIBRETT_C_UNDEFINED_ERROR = 5
};

/* Method numbers, as the reference's enum (src/plan.h:54-56). */
enum {
  LIBRETT_C_TRIVIAL = 1,
  LIBRETT_C_PACKED = 2,
  LIBRETT_C_PACKED_SPLIT = 3,
  LIBRETT_C_TILED = 4,
  LIBRETT_C_TILED_COPY = 5
};

/* Replaces librettInitialize / librettFinalize (src/librett.h:57,62). */
void librett_c_initialize(void);
void librett_c_finalize(void);

/* Replaces librettPlan (src/librett.h:78); `stream` is a hipStream_t. */
int librett_c_plan(unsigned int* handle, int rank, const int* dim, const int* permutation, size_t sizeofType,
                   void* stream);

/* Replaces librettPlanMeasure (src/librett.h:96-97). */
int librett_c_plan_measure(unsigned int* handle, int rank, const int* dim, const int* permutation,
                           size_t sizeofType, void* stream, void* idata, void* odata);

/* Replaces librettDestroy (src/librett.h:108). */
int librett_c_destroy(unsigned int handle);

/* Replaces librettExecute (src/librett.h:121). */
int librett_c_execute(unsigned int handle, const void* idata, void* odata);

/* Description of a plan or of a candidate plan. */
typedef struct librett_c_plan_info {
  int method;        /* LIBRETT_C_TRIVIAL ... LIBRETT_C_TILED_COPY */
  int elem;          /* bytes per element */
  int tile_x, tile_y, vec_x, vec_y; /* Tiled */
  int word, copy_tx;                /* TiledCopy */
  int nreg, num_split, split_rank;  /* Packed / PackedSplit */
  int size_mm, size_mk, vol_mmk, vol_mbar;
  int grid_x, grid_y, block, lds_bytes, active_blocks;
  double predicted_s;    /* gfx950 model estimate */
  double measured_s;     /* last librett_c_plan_measure time, or -1 */
  double read_lines, write_lines, partial_lines; /* modelled 128-byte lines */
  long long bytes;       /* algorithmic bytes, 2 * vol * elem */
  /* model features: [0] load transactions, [1] store transactions (128-byte
   * requests of all wave instructions), [2] items (block
   * rounds), [3] resident blocks, [4] rounds per block, [5] modelled HBM
   * time, [6] modelled latency time, [7] staged bytes per round */
  double feat[8];
  char desc[192];
} librett_c_plan_info;

/* Device description for host-only planning (extension). */
typedef struct librett_c_device {
  int cu_count;
  int lds_per_block;
  int lds_per_cu;
  int max_threads_per_block;
  int max_waves_per_cu;
  double hbm_bytes_per_s;
} librett_c_device;

/* Extension: describes a stored plan. */
int librett_c_plan_info_get(unsigned int handle, librett_c_plan_info* info);

/* Extension: number of candidates and their descriptions for a problem on
 * the current device (at most `max` are written). */
int librett_c_candidates(int rank, const int* dim, const int* permutation, size_t sizeofType,
                         librett_c_plan_info* infos, int max, int* count);

/* Extension: plan that uses candidate `index` of librett_c_candidates. */
int librett_c_plan_candidate(unsigned int* handle, int rank, const int* dim, const int* permutation,
                             size_t sizeofType, void* stream, int index);

/* Extension: host-only candidate list for a described device (no GPU). */
int librett_c_host_candidates(int rank, const int* dim, const int* permutation, size_t sizeofType,
                              const librett_c_device* dev, librett_c_plan_info* infos, int max, int* count);

/* Extension: runs candidate `index` of librett_c_host_candidates on the host
 * (the kernels' per-thread code with barrier order kept); host buffers.
 * The replay runs between guard bands and returns LIBRETT_C_UNDEFINED_ERROR
 * if any byte outside the output tensor was written. */
int librett_c_host_replay(int rank, const int* dim, const int* permutation, size_t sizeofType,
                          const librett_c_device* dev, int index, const void* idata, void* odata);

/* Extension: host-only rank reduction (reference reduceRanks,
 * src/plan.cpp:64-127) after size-1 dims are dropped. Writes the reduced
 * rank, dims and permutation. */
int librett_c_reduce_ranks(int rank, const int* dim, const int* permutation, int* red_rank, int* red_dim,
                           int* red_perm);

/* Extension: device-to-device copy kernel used as a bandwidth ceiling next
 * to hipMemcpyAsync (the reference's bench copy kernels, src/GpuMemcpy.cpp,
 * play the same role). bytes and both pointers must be multiples of 16;
 * words_per_thread is 1, 2, 4 or 8 16-byte words per round; nontemporal != 0
 * streams the stores; grid <= 0 launches one block per 256*words_per_thread
 * words. Asynchronous on `stream`. */
int librett_c_device_copy(void* dst, const void* src, size_t bytes, void* stream, int words_per_thread,
                          int nontemporal, int grid);

/* ---------------------------------------------------------------------
 * Access counters (extension). Reference: librettGpuModelKernel and its
 * MemStat counters, src/GpuModelKernel.cpp:43-64, 343-946, 1013-1224.
 * Exact counts over every work item of a candidate's launch, computed on
 * the device by a kernel that forms the same lane addresses as the
 * transpose kernel (the planner's model estimates them from sampled work
 * items). A work item is a tile (Tiled, TiledCopy) or one staged round,
 * i.e. one chunk at one outer index (Packed, PackedSplit). Trivial plans
 * (a device copy) have no kernel to count: LIBRETT_C_INVALID_PARAMETER.
 * ------------------------------------------------------------------- */
typedef struct librett_c_counters {
  unsigned long long ld_req, st_req;     /* wave load / store instructions with an active lane */
  unsigned long long ld_tran, st_tran;   /* distinct 128-byte segments per instruction, summed */
  unsigned long long ld_bytes, st_bytes; /* bytes the active lanes load / store (vol * elem each) */
  unsigned long long read_lines;         /* distinct 128-byte lines each work item reads, summed */
  unsigned long long write_lines;        /* distinct 128-byte lines each work item writes, summed */
  unsigned long long partial_lines;      /* written lines a work item covers only in part, summed */
  unsigned long long items;              /* work items */
} librett_c_counters;

/* Counts of a stored plan's kernel, on the current device (synchronises
 * the plan's stream). */
int librett_c_plan_counters(unsigned int handle, librett_c_counters* out);

/* Counts of candidate `index` of librett_c_candidates, on the current
 * device; synchronises `stream`. */
int librett_c_candidate_counters(int rank, const int* dim, const int* permutation, size_t sizeofType,
                                 void* stream, int index, librett_c_counters* out);

/* The same counts computed on the host by the same per-item code, for
 * candidate `index` of librett_c_host_candidates. */
int librett_c_host_candidate_counters(int rank, const int* dim, const int* permutation, size_t sizeofType,
                                      const librett_c_device* dev, int index, librett_c_counters* out);

/* Host-only access trace of candidate `index` of librett_c_host_candidates,
 * six values per access: work item, 0 load / 1 store, instruction, thread,
 * byte offset from the tensor start, bytes. source 0: the counters'
 * enumeration (instruction = pass); source 1: the transpose kernel's host
 * replay (instruction = call order within the thread). At most `max`
 * records are written; *count receives the total. */
int librett_c_host_access_trace(int rank, const int* dim, const int* permutation, size_t sizeofType,
                                const librett_c_device* dev, int index, int source, long long* records,
                                long long max, long long* count);

/* Extension: overrides the analytic model's constants, in this order:
 * sustained HBM fraction, written-line weight, partial-line weight, round
 * latency (s), launch overhead (s), seconds per load transaction per CU,
 * store-transaction weight, per-CU bandwidth (B/s), per-block bandwidth
 * (B/s). Negative entries keep their value. Affects plans created
 * afterwards (plans already cached keep their choice). */
void librett_c_set_model_constants(const double* values, int count);

/* ---------------------------------------------------------------------
 * Batched engine (extension; the reference has no batched mode). Many
 * independent transposes run in one persistent kernel launch, replayed from
 * a HIP graph. Problems are given flattened: problem i has ranks[i] dims at
 * dims[o_i ...] and its permutation at perms[o_i ...], o_i = sum_{j<i}
 * ranks[j]. tune = 0 picks each problem's plan with the cost model; tune = k
 * > 0 times the model's k best batch plans of each distinct problem inside
 * the persistent kernel and keeps the fastest (remembered across batches).
 * ------------------------------------------------------------------- */
typedef struct librett_c_batch_info {
  int count;            /* problems */
  int descriptors;      /* device descriptors (a flat copy may use two) */
  int total_units;      /* work units of one launch */
  int grid;             /* persistent workgroups */
  int tuned;            /* problems timed by the autotuner in this batch */
  int method_count[6];  /* problems per method (index = method number) */
  long long bytes;      /* algorithmic bytes, sum of 2 * vol * elem */
  double plan_seconds, tune_seconds;
} librett_c_batch_info;

int librett_c_batch_plan(unsigned int* handle, int count, const int* ranks, const int* dims, const int* perms,
                         const size_t* elems, void* stream, int tune);

/* Device buffers of every problem (idata[i], odata[i]); may be called again
 * to point the batch at other buffers. Synchronises the batch's stream. */
int librett_c_batch_set_buffers(unsigned int handle, const void* const* idata, void* const* odata);

/* Enqueues the whole batch on the stream given at planning: one graph
 * launch (use_graph != 0, captured on first use) or one plain launch. */
int librett_c_batch_execute(unsigned int handle, int use_graph);

int librett_c_batch_destroy(unsigned int handle);
int librett_c_batch_info_get(unsigned int handle, librett_c_batch_info* info);

/* Host replay of the batch's persistent kernel for a described device (no
 * GPU): host buffers, `grid` workgroups (<= 0: the planned grid). */
int librett_c_batch_host_replay(int count, const int* ranks, const int* dims, const int* perms, const size_t* elems,
                                const librett_c_device* dev, const void* const* idata, void* const* odata, int grid);

/* Host replay of batch-mode candidate `index` of one problem (the order of
 * the batch planner's ranking), through the persistent kernel's per-thread
 * code between guard bands. `fallback` selects the descriptor for buffers
 * not aligned to the vector width; `unit_bytes` the work-unit size (0:
 * default); `grid` the workgroups replayed (0: 4 per CU). With idata and
 * odata both NULL it only stores the number of candidates in *count.
 * Returns LIBRETT_UNDEFINED_ERROR if a store leaves the output or a unit
 * writes past the kernel's LDS. No reference counterpart: test support for
 * the autotuner, which runs candidates other than the model's choice. */
int librett_c_batch_host_candidate(int rank, const int* dim, const int* permutation, size_t sizeofType,
                                   const librett_c_device* device, int index, int fallback, long long unit_bytes,
                                   int grid, const void* idata, void* odata, int* count);

/* ---------------------------------------------------------------------
 * Serving engine (extension; the reference has no serving mode).
 * Continuous batching: requests queued from any thread are launched as
 * micro-batches of the batched engine on `slots` streams as soon as a slot
 * frees up, so requests arriving while others run overlap with them on the
 * device. Output buffers come from a device arena with a hard byte budget;
 * plans are cached in an LRU. Execution order between requests is not
 * defined; each request's result does not depend on the others.
 * ------------------------------------------------------------------- */
#define LIBRETT_C_BUSY 6 /* the arena cannot hold the allocation right now */

typedef struct librett_c_server_stats {
  long long submitted, completed, failed;
  long long bytes;              /* algorithmic bytes of completed requests */
  long long batches;            /* micro-batches launched */
  double mean_batch;            /* requests per micro-batch */
  long long admission_stalls;   /* arrivals that waited for arena space */
  double stall_seconds;
  long long arena_peak_bytes, arena_in_use_bytes;
  double p50_us, p90_us, p99_us, max_us, mean_us; /* arrival -> completion */
  double wall_seconds;          /* first arrival -> last completion */
  double gbs;                   /* bytes / wall_seconds */
  long long plan_misses;        /* requests whose shape was not in the plan cache */
  double plan_seconds;          /* time spent planning those shapes (in the submitting threads) */
} librett_c_server_stats;

typedef struct librett_c_poisson {
  double qps;                 /* mean arrival rate (all clients together) */
  int requests;               /* number of arrivals */
  unsigned long long seed;
  int catalog;                /* distinct request shapes, drawn and planned up front */
  double min_bytes, max_bytes;/* tensor bytes, log-uniform */
  int rank_lo, rank_hi;
  int elems[3];               /* element sizes drawn uniformly */
  int nelems;
  int fresh;                  /* != 0: every request draws a new shape (no catalog), planned when it arrives */
  int inputs_in_arena;        /* != 0: inputs are allocated from the arena too, so admission waits for room for
                                 input and output; their contents are left as the arena holds them */
  int clients;                /* arrival threads, each a Poisson stream of qps / clients (<= 1: one) */
} librett_c_poisson;

int librett_c_server_create(unsigned int* handle, size_t arena_bytes, int slots, int max_batch);
/* Arena allocation (256-byte aligned); LIBRETT_C_BUSY when it does not fit now. */
int librett_c_server_alloc(unsigned int handle, size_t bytes, void** ptr);
int librett_c_server_free(unsigned int handle, void* ptr);
/* Queues one request; returns at once with a ticket. */
int librett_c_server_submit(unsigned int handle, int rank, const int* dim, const int* permutation,
                            size_t sizeofType, const void* idata, void* odata, unsigned long long* ticket);
/* Blocks until the request is complete; latency = completion - submission.
 * A finished request can still be waited on until 2^20 newer requests have
 * finished or the statistics are reset. */
int librett_c_server_wait(unsigned int handle, unsigned long long ticket, double* latency_s);
int librett_c_server_drain(unsigned int handle);
int librett_c_server_stats_get(unsigned int handle, librett_c_server_stats* stats);
int librett_c_server_reset_stats(unsigned int handle);
int librett_c_server_destroy(unsigned int handle);
/* Native Poisson load generator: `requests` arrivals at `qps`, shapes from a
 * catalog planned up front or (fresh != 0) a new shape per request planned on
 * arrival by the arriving client thread, inputs read from slices of
 * `input_pool` (device memory of pool_bytes >= max_bytes) or taken from the
 * arena (inputs_in_arena != 0), outputs allocated from the arena at arrival
 * (waiting while it is full) and everything returned at completion. Blocks
 * until all requests complete. */
int librett_c_server_run_poisson(unsigned int handle, const librett_c_poisson* cfg, const void* input_pool,
                                 size_t pool_bytes);

/* Extension: library version string. */
const char* librett_c_version(void);

/* ------------------------------------------------------------------------
 * Learned plan selector (extension; the reference ranks candidates with its
 * analytic model only, src/plan.cpp:721-789). A 2 x 64 FP32 MLP scores
 * every candidate plan of a problem from LIBRETT_C_SELECTOR_FEATURES
 * features; librettPlan / librett_c_plan take its lowest score when
 * weights are loaded and the analytic model otherwise, or when the reduced
 * rank exceeds 7. Weights load on first use from the file named by
 * LIBRETT_SELECTOR_WEIGHTS (empty or "none": analytic model only), else from
 * selector_weights.bin beside the library; librett_c_selector_load / _set
 * replace them at run time.
 * ------------------------------------------------------------------------ */
#define LIBRETT_C_SELECTOR_FEATURES 50

/* Input features, hidden width and parameter count of the MLP. */
void librett_c_mlp_sizes(int* nfeat, int* hidden, int* nparams);

/* Times every candidate of a problem with librettPlanMeasure's own loop
 * (warm-up, then the median of LIBRETT_MEASURE_REPS runs, each after an
 * output memset); times[i] in seconds, negative if candidate i did not run.
 * Candidate order is librett_c_candidates'. The selector's training labels. */
int librett_c_measure_candidates(int rank, const int* dim, const int* permutation, size_t sizeofType, void* stream,
                                 const void* idata, void* odata, double* times, int max, int* count);

/* Raw features (count x LIBRETT_C_SELECTOR_FEATURES floats, candidate order
 * as above) and the analytic model's choice. `device` NULL: the current GPU;
 * otherwise host-only planning for the described device. */
int librett_c_selector_features(int rank, const int* dim, const int* permutation, size_t sizeofType,
                                const librett_c_device* device, float* feats, int max, int* count,
                                int* model_choice);

/* Candidate librettPlan's selector step picks on the current GPU: the
 * ensemble's argmin, or the analytic model's choice when the ensemble
 * scores that within librett_c_selector_margin() of its own pick; *index
 * is -1 when no weights are loaded or the problem is outside the
 * selector's domain (librettPlan then uses the analytic model). */
int librett_c_selector_choose(int rank, const int* dim, const int* permutation, size_t sizeofType, int* index);

/* Loads weights: nparams floats in the layout W1[64][50], b1[64], W2[64][64],
 * b2[64], w3[64], b3, with per-feature mean and std used to normalise the
 * inputs. nparams may be a multiple k (k <= 16) of one network's count: an
 * ensemble of k networks whose scores are averaged. params NULL unloads. */
int librett_c_selector_set(const float* params, int nparams, const float* mean, const float* std, int nfeat);

/* Loads a weights file: uint32 magic 0x4D54524C, version 1, 50, 64, then
 * float32 mean[50], std[50], params[nparams] (little-endian); version 2
 * has uint32 k (networks) after the 64 and k * nparams parameters. */
int librett_c_selector_load(const char* path);

/* Generation of the loaded weights (0: none). */
unsigned long long librett_c_selector_generation(void);

/* librettPlan keeps the analytic model's choice unless the selector scores
 * its own pick lower by more than this margin (log2 of predicted time; 0.03
 * by default, LIBRETT_SELECTOR_MARGIN overrides). */
double librett_c_selector_margin(void);

/* MLP kernels on device pointers. forward: out[n] = MLP(x[n][50]).
 * grad: workgroup b (of nblocks) writes the gradient of samples
 * idx[base + 128 b .. min(end, base + 128 (b+1))) of sum w (out - y)^2 * inv_m
 * (w: per-sample weights, NULL for 1) to partial[b][nparams] and its weighted
 * squared-error sum to loss[b].
 * adam: params -= lr * mhat / (sqrt(vhat) + eps) with the gradient summed
 * over the nparts partials; c1 = 1/(1-beta1^t), c2 = 1/(1-beta2^t).
 * train: Adam over `epochs` shuffled passes (std::mt19937(seed)) with
 * minibatches of `minibatch` samples (a multiple of 128), each normalised by
 * its weight sum; w (device) and w_host (the same weights on the host) are
 * both NULL for unweighted training; epoch_loss (host, may be NULL) gets the
 * weighted mean squared error of each epoch. */
int librett_c_mlp_forward(const float* params, const float* x, int n, float* out, void* stream);
int librett_c_mlp_grad(const float* params, const float* x, const float* y, const float* w, const int* idx, int base,
                       int end, float inv_m, int nblocks, float* partial, float* loss, void* stream);
int librett_c_mlp_adam(const float* partial, int nparts, float* params, float* m, float* v, float lr, float beta1,
                       float beta2, float c1, float c2, float eps, void* stream);
int librett_c_mlp_train(const float* x, const float* y, const float* w, const float* w_host, int n, float* params,
                        int epochs, float lr, int minibatch, unsigned int seed, float* epoch_loss, void* stream);

/* Host replays of the same kernels' per-thread code (one workgroup for
 * grad: count <= 128 samples idx[0..count)). */
void librett_c_mlp_host_forward(const float* params, const float* x, int n, float* out);
int librett_c_mlp_host_grad(const float* params, const float* x, const float* y, const float* w, const int* idx,
                            int count, float inv_m, float* grad, float* loss);
void librett_c_mlp_host_adam(const float* partial, int nparts, float* params, float* m, float* v, float lr,
                             float beta1, float beta2, float c1, float c2, float eps);

/* ------------------------------------------------------------------------
 * Index-pattern checker (extension; the reference's TensorTester,
 * src/TensorTester.cpp:34-126,201-220). check_fill sets 32-bit word k of
 * `data` to k (mod 2^32), so input element e of B bytes holds the words
 * e*B/4 + w. check_transpose verifies on the GPU that `odata` is the
 * transpose of such an input (column-major dims, output dims
 * dim[permutation[i]]), for element sizes 4, 8 and 16; it stores the number
 * of wrong elements and the first wrong output position (-1 if none) and
 * synchronises `stream`. The host_ variants run the same per-element code on
 * host memory.
 * ------------------------------------------------------------------------ */
int librett_c_check_fill(void* data, size_t nbytes, void* stream);
int librett_c_check_transpose(const void* odata, int rank, const int* dim, const int* permutation, size_t sizeofType,
                              void* stream, unsigned long long* mismatches, long long* first);
void librett_c_host_check_fill(void* data, size_t nbytes);

/* Allocation-free variant of librett_c_check_transpose. It has no
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc). It allocates nothing and does not synchronise: it
 * enqueues the check on `stream` and writes result[0] = wrong elements and
 * result[1] = first wrong output position (all ones if none) to `result`,
 * two unsigned long longs of device memory owned by the caller. Its scratch
 * is a fixed 64 KiB array in device memory per GPU (4096 per-workgroup
 * partials), kept until the process ends; problems whose check would need
 * more partials are refused with LIBRETT_C_INVALID_PARAMETER (the checker's
 * grid is capped at 4096, so none are today). Calls that overlap on one GPU
 * share that scratch and race: serialise them (one stream, or events). */
int librett_c_check_transpose_noalloc(const void* odata, int rank, const int* dim, const int* permutation,
                                      size_t sizeofType, void* stream, unsigned long long* result);
int librett_c_host_check_transpose(const void* odata, int rank, const int* dim, const int* permutation,
                                   size_t sizeofType, unsigned long long* mismatches, long long* first);

/* ------------------------------------------------------------------------
 * Device allocator hook (extension; the reference uses an Umpire allocator
 * chosen at librettInitialize, src/librett.cpp:387-395). The library's own
 * device buffers -- batch descriptors and autotuner scratch, the serving
 * arena and slot buffers, selector training and checker buffers -- come
 * from `alloc` and go back through `free_fn` when installed (both NULL:
 * hipMalloc / hipFree). A buffer always returns to the allocator that made
 * it. Plans own no device memory. The selector's small per-device inference
 * buffers are kept between plans; installing another allocator (or NULL)
 * first returns them to the allocator that made them, and the next plan
 * takes new ones from the allocator then installed.
 * allocator_live counts buffers currently held from a caller's allocator.
 * ------------------------------------------------------------------------ */
typedef void* (*librett_c_alloc_fn)(size_t bytes, void* ctx);
typedef void (*librett_c_free_fn)(void* ptr, void* ctx);
int librett_c_set_allocator(librett_c_alloc_fn alloc, librett_c_free_fn free_fn, void* ctx);
unsigned long long librett_c_allocator_live(void);

#ifdef __cplusplus
}
#endif

#endif /* LIBRETT_C_H */
