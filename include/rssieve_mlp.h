/*
 * rssieve_mlp.h -- configuration search for librssieve.so: the autotuner,
 * which times candidate BSGS configurations of a job on the device, and the
 * learned cost model, a small fp32 MLP trained and evaluated by the library's
 * own HIP kernels.  Neither has a counterpart in the reference
 * (mfl0p/RSSieve v1.3), whose only configuration rule is the fixed split of
 * build_giant_kernels (cl_sieve.cpp:885-938); that rule stays available as
 * rs_params_reference() and is the fallback of both.
 * Same conventions as rssieve.h.
 */
#ifndef RSSIEVE_MLP_H
#define RSSIEVE_MLP_H

#include <stdint.h>

#include "rssieve.h"

#ifdef __cplusplus
extern "C" {
#endif

/* --- autotuner ------------------------------------------------------------ */

/* Candidate configurations for a job shape: the reference split and the
 * merged table at several sizes, hash loads and workgroup sizes that fit the
 * device's LDS.  Writes at most cap entries; *n receives the count. */
int rs_tune_candidates(rs_ctx *ctx, uint32_t nmin, uint32_t nmax, int32_t kcount, rs_params *out, int cap, int *n);

/* Time every candidate on the same sample [pmin, pmin + width) of the job:
 * the job is reconfigured, sieved, and restored (bitmap and counters), so the
 * tuning leaves no trace.  seconds[i] = stream time of candidate i, primes[i]
 * = primes of the sample; *best = index of the fastest.  The job ends up
 * configured with the best candidate. */
int rs_autotune(rs_job *job, uint64_t pmin, uint64_t width, const rs_params *cand, int ncand, double *seconds,
                uint64_t *primes, int *best);

/* Change the BSGS configuration of an existing job (algorithm, Q, hash size,
 * workgroup size, segment target); the bitmap and counters are kept. */
int rs_job_reconfigure(rs_job *job, const rs_params *prm);

/* The batch split of a job (no reference counterpart; the reference sizes one
 * segment per round, primes_per_bsgs, cl_sieve.cpp:693-735): option i sieves
 * [pmin, pmin + width) as views[i] equal parts in one batch (rs_job_view,
 * rs_sieve_batch) with seg_primes[i] target primes per segment, so one
 * persistent BSGS launch serves views[i] * seg_primes[i] primes.
 * seconds[i] = stream time, primes[i] = primes of the sample; *best = the
 * option with the fewest seconds per prime.  Bitmap and counters are
 * restored and the job keeps its configuration. */
int rs_autotune_split(rs_job *job, uint64_t pmin, uint64_t width, const uint32_t *views, const uint32_t *seg_primes,
                      int nopt, double *seconds, uint64_t *primes, int *best);

/* --- learned cost model ----------------------------------------------------- */

typedef struct rs_mlp rs_mlp;

/* n_in -> n_hidden -> n_hidden -> n_out, ReLU, He-uniform initialisation from
 * 'seed'.  Parameters are one fp32 vector in the order W1[n_in][H], b1[H],
 * W2[H][H], b2[H], W3[H][n_out], b3[n_out]. */
int rs_mlp_create(rs_ctx *ctx, int n_in, int n_hidden, int n_out, uint64_t seed, rs_mlp **out);
void rs_mlp_destroy(rs_mlp *m);
int rs_mlp_num_params(const rs_mlp *m, uint64_t *n);
int rs_mlp_get_params(rs_mlp *m, float *out);
/* Replace the parameters and reset the optimiser state. */
int rs_mlp_set_params(rs_mlp *m, const float *in);
/* Y[n][n_out] = network(X[n][n_in]). */
int rs_mlp_forward(rs_mlp *m, const float *X, int n, float *Y);
/* Masked mean squared error / 2 of network(X) against T (M = per-entry mask,
 * may be NULL) and, if grads is not NULL, its gradient w.r.t. every parameter
 * (same order as the parameter vector). */
int rs_mlp_loss_grad(rs_mlp *m, const float *X, const float *T, const float *M, int n, float *loss, float *grads);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with the last gradient. */
int rs_mlp_adam_step(rs_mlp *m, float lr);
/* Mini-batch Adam for 'epochs' passes in a seeded order; *final_loss = loss
 * over the whole set afterwards. */
int rs_mlp_train(rs_mlp *m, const float *X, const float *T, const float *M, int n, int epochs, int batch, float lr,
                 uint64_t seed, float *final_loss);

/* --- the trained model from C ----------------------------------------------- */

/* A cost model file written by the training script (rssieve_amd/tuning.py;
 * the committed one is rssieve_amd/model/cost_model.bin): a head over the
 * BSGS candidates of rs_tune_candidates and, optionally, a head over batch
 * splits, each an MLP whose parameters are loaded into rs_mlp objects on ctx
 * and evaluated by the library's kernels.  RS_E_IO on a missing or malformed
 * file. */
typedef struct rs_model rs_model;
int rs_model_load(rs_ctx *ctx, const char *path, rs_model **out);
void rs_model_free(rs_model *m);
/* The candidate (index into cand, the list rs_tune_candidates gives for the
 * job's shape) with the lowest predicted seconds per prime for a job of
 * 'kcount' sequences of base 'base', 'nterms' candidate terms over an n-span
 * of 'span', sieved at p ~ 2^log2p; only candidates with threads != 0 (and
 * threads == threads_only, if nonzero) are eligible; 0, the reference split,
 * if none is.  RS_E_ARG if ncand differs from the model's candidate count. */
int rs_model_choose(rs_model *m, uint32_t base, int32_t kcount, uint64_t nterms, uint32_t span, double log2p,
                    const rs_params *cand, int ncand, uint32_t threads_only, int *idx);
/* The same job's eligible candidates in order of predicted seconds per prime,
 * fastest first: the first min(k, eligible) indices into idx[0..k), their
 * count in *nout (no reference counterpart: a caller times these few on a
 * sample with rs_autotune instead of all of them, and so recovers most of
 * the model's error); idx[0] = 0 and *nout = 1 if none is eligible. */
int rs_model_rank(rs_model *m, uint32_t base, int32_t kcount, uint64_t nterms, uint32_t span, double log2p,
                  const rs_params *cand, int ncand, uint32_t threads_only, int k, int *idx, int *nout);
/* The batch split with the lowest predicted seconds per prime for the same
 * job: views per batched round and target primes per segment (the options
 * rs_autotune_split times).  RS_E_ARG if the model has no split head. */
int rs_model_split(rs_model *m, uint32_t base, int32_t kcount, uint64_t nterms, uint32_t span, double log2p,
                   uint32_t *views, uint32_t *seg_primes);

#ifdef __cplusplus
}
#endif
#endif /* RSSIEVE_MLP_H */
