/* mclicks.h -- C ABI of the MI355X EM engine for the PBM, UBM, CCM and DBN
 * click models.
 *
 * Each entry point replaces one piece of the reference's click-model
 * interface or EM driver; the reference file:line it stands in for is given
 * beside it.  All pointers are host pointers; sizes are element counts.
 * Every call returns an mc_status; mc_last_error() returns a thread-local
 * message describing the most recent failure on the calling thread.  No call
 * exits the process or throws across the ABI (the reference aborts through
 * CUDA_CHECK/MPI_Abort, utils/utils.cuh:27-42).
 *
 * Threading: one mc_model is bound to one device and one HIP stream; calls on
 * a model must be serialised by the caller.  Different models may be used
 * from different threads.
 */
#ifndef MCLICKS_H
#define MCLICKS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MC_MAX_SERP 10 /* utils/definitions.h:50 */

/* Model ids as on the reference command line (main.cpp:178-183). */
enum { MC_PBM = 0, MC_CCM = 1, MC_DBN = 2, MC_UBM = 3 };

typedef enum {
    MC_OK = 0,
    MC_ERR_INVALID_ARG = 1,
    MC_ERR_OOM = 2,
    MC_ERR_HIP = 3,
    MC_ERR_IO = 4,
    MC_ERR_STATE = 5,
    MC_ERR_UNSUPPORTED = 6
} mc_status;

/* Options for mc_model_set_option. */
enum {
    /* 0 = deterministic segmented reduction (default), 1 = float atomics. */
    MC_OPT_REDUCE_MODE = 0,
    /* Target sessions per workgroup work item (default 8192). */
    MC_OPT_ITEM_SESSIONS = 1,
    /* Per-qd accumulator window per work item, in entries (default 1280). */
    MC_OPT_ITEM_WINDOW = 2,
    /* 1 = reorder sessions inside each work item by length, then last-click
     * rank, so that wavefronts run uniform chains (default 1). */
    MC_OPT_SORT_BY_LCR = 3,
    /* Workgroup size of the E/M kernel: 64, 128, 256 or 512 (default 256). */
    MC_OPT_BLOCK = 4,
    /* 1 = sessions that are the only session of their query (and show
     * distinct documents) update their entries directly, without a
     * reduction (default 0). */
    MC_OPT_DIRECT = 5,
    /* 1 = window items run in one persistent kernel: each resident
     * workgroup walks a fixed stride of items and prefetches its next item
     * (default 1; needs MC_OPT_BLOCK 256 and a window of at most 1280
     * entries). */
    MC_OPT_PERSISTENT = 6,
    /* Minibatches of the gradient path (mc_adam_steps): the query-sorted
     * training set is cut into this many runs of whole query groups of about
     * equal session counts (default 1 = full batch). */
    MC_OPT_BATCHES = 7,
    /* 1 = in the segmented mode, window items whose sessions each show
     * documents within 256 consecutive entries of the window store every
     * session as one 16-byte record (8-bit offsets from the session's
     * smallest index, the click bits and the length), read with one load;
     * 0 = every window item keeps 16-bit rank-major indices (default 1). */
    MC_OPT_PACKED = 8,
    /* 1 = the EM kernels skip DBN and CCM sessions without a click (default
     * 1).  By the reference's last-click convention (search.cu:160-173
     * returns MAX_SERP when nothing is clicked) every count such a session
     * contributes is parameter-independent; the counts are folded into the
     * constant bases at init, so visiting them changes no result.  The
     * sessions stay on the device for the gradient path.  0 = visit them. */
    MC_OPT_SKIP_CONSTANT = 9,
    /* 1 = when the window kernel (persistent or one workgroup per item) is
     * an iteration's only E-step launch, it also runs the M-step: the
     * workgroup that adds the last slice of a query group split over several
     * work items finalises the group's entries, and the last workgroup to
     * finish the shared tables, so no merge kernel is launched (default 1).
     * Same results. */
    MC_OPT_FUSED_MERGE = 10
};
enum { MC_REDUCE_SEGMENTED = 0, MC_REDUCE_ATOMIC = 1 };

/* ---------------------------------------------------------------------------
 * Dataset preparation (host).  Replaces Dataset::parse_dataset
 * (data/dataset.cpp:421-455), which parses the raw log (dataset.cpp:224-305),
 * splits train/test (dataset.cpp:385-396), sorts the train set by query
 * (dataset.cpp:17-43) and assigns query-document indices
 * (dataset.cpp:314-374).  One worker: the hash ring sends every session to
 * device 0.
 * ------------------------------------------------------------------------- */
typedef struct mc_dataset mc_dataset;

/* max_sessions == -1 reads every complete session (main.cpp -s flag). */
mc_status mc_dataset_parse(const char* path, int max_sessions, float test_share, mc_dataset** out);

/* Same preparation for sessions already in memory: query[n], docs[n*10],
 * clicks[n*10] (0/1), in log order. */
mc_status mc_dataset_from_sessions(int64_t n, const int32_t* query, const int32_t* docs, const uint8_t* clicks,
                                   float test_share, mc_dataset** out);

/* Same for result lists of 1..10 ranks (len[n]); ranks r >= len[s] hold no
 * result: their docs are ignored, their clicks must be 0, they get no
 * query-document index (pidx -1) and do not count as seen documents when
 * test sessions are filtered.  The reference has fixed 10-rank lists
 * (utils/definitions.h:50, dataset.cpp:266); with every len = 10 this equals
 * mc_dataset_from_sessions. */
mc_status mc_dataset_from_sessions_len(int64_t n, const int32_t* query, const int32_t* docs, const uint8_t* clicks,
                                       const uint8_t* len, float test_share, mc_dataset** out);

mc_status mc_dataset_sizes(const mc_dataset* ds, int64_t* n_parsed, int64_t* n_train, int64_t* n_test,
                           int64_t* n_qd);

/* which: 0 = train (sorted by query), 1 = test (filtered, log order).
 * Arrays sized n, n*10, n*10, n*10.  pidx is -1 for unseen test pairs. */
mc_status mc_dataset_export(const mc_dataset* ds, int which, int32_t* query, int32_t* docs, uint8_t* clicks,
                            int32_t* pidx);

/* Ranks shown per session of the train (0) or test (1) set, len[n]; all 10
 * for datasets prepared without lengths. */
mc_status mc_dataset_export_len(const mc_dataset* ds, int which, uint8_t* len);

void mc_dataset_free(mc_dataset* ds);

/* ---------------------------------------------------------------------------
 * Click model on one device.
 * ------------------------------------------------------------------------- */
typedef struct mc_model mc_model;

/* create_cm_host (click_models/base.cu:19-49).  device = HIP ordinal. */
mc_status mc_model_create(int model_type, int device, mc_model** out);

mc_status mc_model_set_option(mc_model* m, int option, int64_t value);

/* The model id given to mc_model_create (config.model_type, main.cpp:230). */
mc_status mc_get_model_type(const mc_model* m, int* model_type);

/* ClickModel_Hst::init_parameters (base.cuh:39, common.cu:31-61) together
 * with the one-time dataset upload (parallel_em.cu:248-253).
 * query[n] must group equal queries contiguously (the reference sorts the
 * train set by query); pidx[n*10] and clicks[n*10] are session-major;
 * len[n] (1..10) may be NULL for full SERPs.  Tables start at (1,2). */
mc_status mc_init_parameters(mc_model* m, int64_t n, const int32_t* query, const int32_t* pidx,
                             const uint8_t* clicks, const uint8_t* len, int64_t n_qd);

/* Convenience: init_parameters from a prepared dataset's train set. */
mc_status mc_init_from_dataset(mc_model* m, const mc_dataset* ds);

/* ClickModel_Hst::process_session (base.cuh:36), i.e. Kernel::em_training
 * (parallel_em/kernel.cu:42-64): the E-step over all training sessions with
 * the expected counts pre-reduced per work item. */
mc_status mc_process_session(mc_model* m);

/* reset_parameters + update_parameters (base.cuh:37,43; parallel_em.cu:503-547):
 * theta <- (1,2) + sum of the E-step's expected counts. */
mc_status mc_update_parameters(mc_model* m);

/* n_iter x (E-step, M-step) as in parallel_em.cu:461-609.  use_graph != 0
 * replays the whole run as one captured HIP graph. */
mc_status mc_em_iterations(mc_model* m, int n_iter, int use_graph);

/* Same as mc_em_iterations(use_graph = 0) but timed with HIP events on the
 * model's stream: total wall time and the average E-step/M-step kernel time. */
mc_status mc_em_iterations_timed(mc_model* m, int n_iter, double* ms_total, double* ms_estep_avg,
                                 double* ms_mstep_avg);

/* Times n_iter iterations replayed from one captured graph (after one warm
 * replay). */
mc_status mc_em_graph_timed(mc_model* m, int n_iter, double* ms_total);

mc_status mc_synchronize(mc_model* m);

/* get_parameter_information (base.cuh:33): table t's name, whether it is
 * public (rank-indexed, shared by all queries) and its length.  Tables are
 * numbered per-qd first: attractiveness (all), satisfaction (DBN), then the
 * public one: examination (PBM 10, UBM 55), continuation (CCM tau1..tau3,
 * DBN gamma).  Names as in get_parameter_information (pbm.cu:124-125,
 * ubm.cu:118-119, ccm.cu:118-119, dbn.cu:135-136). */
mc_status mc_get_n_tables(const mc_model* m, int* n_tables);
mc_status mc_get_table_info(const mc_model* m, int t, const char** name, int* is_public, int64_t* size);

/* get_parameters + transfer_parameters(D2H) (base.cuh:40-41, common.cu:93-102). */
mc_status mc_get_parameters(mc_model* m, int t, float* num, float* den);

/* set_parameters + transfer_parameters(H2D) (base.cuh:40,42). */
mc_status mc_set_parameters(mc_model* m, int t, const float* num, const float* den);

/* reset_parameters (base.cuh:43, common.cu:71-82): every table to (1,2). */
mc_status mc_reset_parameters(mc_model* m);

/* Held-out evaluation on the device: get_log_conditional_click_probs and
 * get_full_click_probs per session (e.g. dbn.cu:611-683), LogLikelihood and
 * Perplexity (click_models/evaluation.cpp:25-66).  pidx[n*10] (-1 = unseen
 * pair, valued 1/2), clicks[n*10], len[n] (NULL = 10).  Outputs: logp/full
 * (n*10, nullable), llh[n] per-session mean log-probability (nullable),
 * *llh_sum = sum of llh, ppl_log2_sum[10] = per-rank sums of log2(full)
 * (FP32 partials per device lane over its few sessions, FP64 beyond),
 * rank_count[10] = sessions contributing to each rank.  Without per-session
 * outputs a session's llh is formed from log2 and scaled by ln 2 (within
 * 1e-8 relative of the natural-log sum).  LL = llh_sum / n;
 * PPL_r = 2^(-ppl_log2_sum[r] / rank_count[r]). */
mc_status mc_evaluate(mc_model* m, int64_t n, const int32_t* pidx, const uint8_t* clicks, const uint8_t* len,
                      float* logp, float* full, float* llh, double* llh_sum, double* ppl_log2_sum,
                      int64_t* rank_count);

/* get_serp_probability (base.cuh:50; pbm.cu:351-365, ubm.cu:345-362,
 * dbn.cu:572-600, ccm.cu:566-593) for n sessions: prob[n*10] receives, per
 * rank, a*e for the PBM and UBM (the click probability given the earlier
 * clicks, whatever this rank's click) and, for the DBN and CCM, the
 * probability of this rank's observed outcome given the earlier clicks
 * (the value whose log mc_evaluate returns in logp).  Arguments as in
 * mc_evaluate; ranks past len[i] receive 0. */
mc_status mc_serp_probability(mc_model* m, int64_t n, const int32_t* pidx, const uint8_t* clicks, const uint8_t* len,
                              float* prob);

/* The held-out set of the evaluation in parallel_em.cu:694-724 (the test
 * set LogLikelihood::evaluate and Perplexity::evaluate are given), uploaded
 * once and kept on the device: arguments as in mc_evaluate.  The sessions
 * are stored in the order of their smallest pair index, which puts the
 * sessions of one query next to each other, with the permutation back to
 * the caller's order.  A later call replaces the set; mc_init_parameters
 * drops it. */
mc_status mc_set_test_set(mc_model* m, int64_t n, const int32_t* pidx, const uint8_t* clicks, const uint8_t* len);

/* mc_evaluate on the resident held-out set at the current parameters.
 * Per-session outputs (nullable) are in the order the set was given.
 * kernel_ms (nullable) receives the evaluation kernel's duration (HIP events
 * on the model's stream). */
mc_status mc_evaluate_test_set(mc_model* m, float* logp, float* full, float* llh, double* llh_sum,
                               double* ppl_log2_sum, int64_t* rank_count, double* kernel_ms);

/* mc_evaluate for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc); no
 * counterpart in the reference.  Every array is a DEVICE pointer: pidx[n*10],
 * clicks[n*10] (session-major as in mc_evaluate), len[n] (NULL = 10),
 * logp/full (n*10) and llh[n] (each nullable), totals[21] = llh_sum,
 * ppl_log2_sum[10], rank_count[10] (as doubles).  Pair indices outside
 * [-1, n_qd) read as unseen pairs and lengths are clamped to [1, 10], since
 * the host cannot check device arrays.  The call enqueues the evaluation on
 * the model's stream and returns without waiting (mc_synchronize waits); it
 * calls no hipMalloc, hipFree or allocating routine.  Its FP64 block
 * partials live in a fixed device array of the library, 344 KB on each GPU
 * it runs on, held until the process ends: evaluations that overlap on one
 * GPU (models with different streams) race on it.  n at most 2^31 - 1. */
mc_status mc_evaluate_device(mc_model* m, int64_t n, const int32_t* pidx, const uint8_t* clicks, const uint8_t* len,
                             float* logp, float* full, float* llh, double* totals);

/* get_memory_usage (base.cuh:31): device bytes held by the model. */
mc_status mc_get_memory_usage(const mc_model* m, size_t* bytes);

/* compute_memory_footprint (base.cuh:32) for n sessions and n_qd pairs. */
mc_status mc_compute_memory_footprint(int model_type, int64_t n, int64_t n_qd, size_t* bytes);

/* Work decomposition chosen at init: number of work items, number of per-qd
 * entries merged across items, the largest item in sessions, and the
 * sessions the EM kernels visit per iteration (MC_OPT_SKIP_CONSTANT). */
mc_status mc_get_layout_info(const mc_model* m, int64_t* n_items, int64_t* n_shared_entries,
                             int64_t* max_item_sessions, int64_t* n_em_sessions);

/* Session stream chosen at init (no counterpart in the reference, which
 * keeps a 32-bit index per rank): sessions of window items store
 * window-relative indices (16-byte packed records or 16-bit indices), the
 * others 32-bit absolute ones.  Returns the number of window-item sessions
 * and the device bytes of all indices (packed records counted whole). */
mc_status mc_get_stream_info(const mc_model* m, int64_t* n_window_sessions, int64_t* index_bytes);

/* Sessions per stored format (no counterpart in the reference): packed
 * 16-byte records, 16-bit window-relative rank-major indices, 32-bit
 * absolute rank-major indices (MC_OPT_PACKED). */
mc_status mc_get_session_formats(const mc_model* m, int64_t* n_packed, int64_t* n_window16, int64_t* n_absolute);

/* Number of mc_model_set_option options (indices 0 .. MC_N_OPTIONS - 1). */
#define MC_N_OPTIONS 11

/* Autotuner of the EM work decomposition; no counterpart in the reference
 * (whose launch shape is fixed: BLOCK_SIZE 128, one session per thread).
 * options[MC_N_OPTIONS] holds the starting option values on entry (a
 * candidate itself) and the fastest found on exit.  Each candidate is
 * measured as `iterations` EM iterations replayed from a graph on `device`
 * on this training set (arguments as mc_init_parameters): the reduction mode
 * first, then sessions per work item (2048..16384, i.e. sessions per lane),
 * the LDS window, the launch shape (persistent kernel, or one workgroup
 * of 256 or 512 lanes per item) and the session format (MC_OPT_PACKED).  *best_ms_per_iter and *n_trials report the
 * winner's time and the number of candidates measured. */
mc_status mc_autotune(int model_type, int device, int64_t n, const int32_t* query, const int32_t* pidx,
                      const uint8_t* clicks, const uint8_t* len, int64_t n_qd, int iterations, int64_t* options,
                      double* best_ms_per_iter, int* n_trials);

/* ---------------------------------------------------------------------------
 * Gradient path: direct maximum-likelihood fitting beside plain EM.  No
 * counterpart in the reference, which only runs EM.  The objective is the
 * training log-likelihood the reference's evaluation defines
 * (get_log_conditional_click_probs summed over ranks, e.g. dbn.cu:611-639)
 * as a function of logits: every table value is sigmoid(logit).  Needs the
 * segmented layout without direct items.
 * ------------------------------------------------------------------------- */

/* Logits from the current values (logit = log(v / (1 - v))), Adam moments
 * and step counts to zero. */
mc_status mc_grad_init(mc_model* m);

/* Log-likelihood of the training sessions of minibatch `batch` (-1 = all)
 * at the current values, and its gradient with respect to every logit:
 * grad[t] for t < n_tables (table sizes as in mc_get_table_info; entries
 * outside the batch are 0).  grad or any grad[t] may be NULL. */
mc_status mc_loglik_gradient(mc_model* m, int batch, double* loglik, float* const* grad);

/* Adam hyper-parameters (defaults 0.05, 0.9, 0.999, 1e-8, prior on).  With
 * use_prior the objective adds log Beta(1 + w, 1 + w) per parameter, w = 1
 * plus the temporaries the reference leaves at (1,2) (DBN satisfaction:
 * unclicked occurrences; PBM/UBM examination: sessions not writing the
 * slot), which makes EM's fixed point for those tables a stationary point. */
mc_status mc_adam_configure(mc_model* m, float lr, float beta1, float beta2, float eps, int use_prior);

/* n_steps Adam steps (ascent on the objective), one minibatch each, batches
 * in order 0, 1, ..., MC_OPT_BATCHES - 1, 0, ...; whole epochs replay one
 * captured graph when use_graph != 0.  Afterwards mc_get_parameters returns
 * (value, 1) for the per-qd tables. */
mc_status mc_adam_steps(mc_model* m, int n_steps, int use_graph);

/* mc_adam_steps(use_graph = 1) timed with HIP events on the model's stream. */
mc_status mc_adam_steps_timed(mc_model* m, int n_steps, double* ms);

/* Logits of table t (after mc_grad_init); setting them sets the values. */
mc_status mc_get_logits(mc_model* m, int t, float* logits);
mc_status mc_set_logits(mc_model* m, int t, const float* logits);

/* ---------------------------------------------------------------------------
 * Parameter output and checkpoints.
 * ------------------------------------------------------------------------- */

/* Communicate::output_parameters (parallel_em/communicator.cpp:415-516) for
 * one worker: one file <prefix>_<name>.csv per table (names as in
 * mc_get_table_info).  Public table: a "rank, <name>" header, then one
 * "<rank>, <value>" line per entry (ostream default formatting, 6
 * significant digits).  Per-qd table: one "<query>, <document>, <value>"
 * line per rank of every training session, in train order (std::to_string,
 * 6 decimals); the reference writes that file's header and then the data
 * over it from offset 0, and the resulting bytes (the data, then whatever of
 * the header extends past it) are reproduced.  value = Param::value(). */
mc_status mc_output_parameters(mc_model* m, const mc_dataset* ds, const char* prefix);

/* The same files from host tables: num[t], den[t] for each table t of
 * model_type in mc_get_table_info order (per-qd tables of ds's n_qd
 * entries, then the public table).  Needs no device. */
mc_status mc_write_parameters_csv(const mc_dataset* ds, const char* prefix, int model_type,
                                  const float* const* num, const float* const* den);

/* Binary checkpoint of every table (numerators and denominators) for
 * resuming training: continuing from a loaded checkpoint gives the same
 * iterates as an uninterrupted run.  No counterpart in the reference, whose
 * only output is the CSV above (values only).  Load requires a model of the
 * same type initialised with the same n_qd. */
mc_status mc_save_parameters(mc_model* m, const char* path);
mc_status mc_load_parameters(mc_model* m, const char* path);

/* destroy_parameters (base.cuh:44) and the model object. */
void mc_model_destroy(mc_model* m);

const char* mc_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MCLICKS_H */
