/*
 * mordred_costmodel.h -- learned plan cost model (C ABI).
 *
 * No counterpart in the reference, whose optimizer scores plans with the
 * analytic model of src/gpudb/CostModel.cu:283-440 alone.  An fp32 MLP with
 * MORDRED_COST_FEATURES inputs, two hidden layers of 64 ReLU units and one
 * output predicts the natural log of a plan's time in milliseconds.  Its
 * forward (inference), backward and Adam kernels are hand-written HIP for
 * gfx950.  The engine feeds it (plan features, measured time) pairs from its
 * own kernels; each new measurement first scores the current model and the
 * analytic estimate (exponential moving averages of |log error|), and the
 * model counts as trusted only while it beats the analytic model on those
 * fresh measurements.  It only ranks plans that all give the same exact
 * answer.
 *
 * Entry points return 0 on success, a positive hipError_t, or -1 for an
 * invalid argument.
 */
#ifndef MORDRED_COSTMODEL_H
#define MORDRED_COSTMODEL_H

#ifdef __cplusplus
extern "C" {
#endif

#define MORDRED_COST_FEATURES 32

typedef struct mordred_cost_model mordred_cost_model;

typedef struct mordred_cost_model_stats {
  long long samples;                   /* measurements observed */
  int scored;                          /* measurements the model was scored on */
  double ewma_abs_log_error;           /* learned model, fresh measurements */
  double ewma_abs_log_error_analytic;  /* analytic model, same measurements */
  int trusted;                         /* the learned model may rank plans */
} mordred_cost_model_stats;

mordred_cost_model* mordred_cost_model_create(int device, unsigned long long seed);
void mordred_cost_model_destroy(mordred_cost_model* m);
const char* mordred_cost_model_last_error(const mordred_cost_model* m);

/* Number of fp32 parameters: W1[32][64], b1[64], W2[64][64], b2[64], W3[64], b3. */
int mordred_cost_model_num_params(void);
int mordred_cost_model_get_params(mordred_cost_model* m, float* out);
/* Replace the parameters and reset the Adam state. */
int mordred_cost_model_set_params(mordred_cost_model* m, const float* in);

/* y[i] = predicted log-ms of x[i*32 .. i*32+31] (inference kernel). */
int mordred_cost_model_predict(mordred_cost_model* m, const float* x, int n, float* y);

/* Gradients of the mean squared error over (x, t) w.r.t. the parameters
 * (forward + backward kernels); grads may be NULL; loss receives the MSE. */
int mordred_cost_model_gradients(mordred_cost_model* m, const float* x, const float* t, int n, float* grads,
                                 float* loss);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the minibatch (x, t). */
int mordred_cost_model_step(mordred_cost_model* m, const float* x, const float* t, int n, float lr, float* loss);

/* Record a measurement (features, measured ms, analytic estimate in ms or
 * <= 0 if none): scores both models on it, then adds it to the replay buffer. */
int mordred_cost_model_observe(mordred_cost_model* m, const float* x, double measured_ms, double analytic_ms);

/* `steps` Adam steps on minibatches of `batch` samples from the replay buffer. */
int mordred_cost_model_train(mordred_cost_model* m, int steps, int batch, float lr, float* last_loss);

int mordred_cost_model_get_stats(const mordred_cost_model* m, mordred_cost_model_stats* st);

#ifdef __cplusplus
}
#endif

#endif /* MORDRED_COSTMODEL_H */
