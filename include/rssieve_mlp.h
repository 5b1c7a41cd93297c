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

#ifdef __cplusplus
}
#endif
#endif /* RSSIEVE_MLP_H */
