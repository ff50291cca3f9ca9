/* gpu_bpf_batch.h -- batched inner SV bootstrap filters and the SMC^2 layer
 * on them, for MI355X (gfx950).  Plain C ABI.
 *
 * No reference function maps one to one onto these entry points: the
 * reference steps its SMC^2 inner filters with one fixed particle count per
 * build (smc2/smc2_rbpf_batch.cu:248-398, kernel_rbpf_step_impl<N_INNER>) and
 * its parameter sweeps loop the plain filter (gpu_bpf/gpu_bpf.cu:1022-1080,
 * gpu_imm_step).  A batch runs M filters of uneven particle counts in one
 * launch per tick; inner filter m is the plain filter of gpu_bpf.h created
 * with (n_particles[m], phi[m], sigma_z[m], mu[m], 0, nu_obs, seeds[m]) and
 * the batch's ESS threshold, tick for tick.
 */
#ifndef GPU_BPF_BATCH_H
#define GPU_BPF_BATCH_H

#include "gpu_bpf.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct GpuBpfBatch GpuBpfBatch;

/* Largest particle count of an inner filter. */
#define GPU_BPF_BATCH_MAX_N 16384

/* n_filters inner filters; arrays of n_filters entries.  nu_obs > 0 selects
 * Student-t observations for every filter.  Returns NULL on failure
 * (gpu_bpf_last_error()). */
GpuBpfBatch* gpu_bpf_batch_create(int n_filters, const int* n_particles, const float* mu,
                                  const float* phi, const float* sigma_z, const int* seeds,
                                  float nu_obs, float ess_threshold);
void gpu_bpf_batch_destroy(GpuBpfBatch* batch);

/* Threads per inner filter (64, 128, 256 or 512); 0 = chosen by the
 * autotuner at the first run.  Returns 0 or -1. */
int gpu_bpf_batch_set_block(GpuBpfBatch* batch, int threads);
int gpu_bpf_batch_get_block(GpuBpfBatch* batch);
/* Times each block size over `ticks` ticks of a scratch batch of the same
 * filters on y and keeps the fastest; returns it, or -1. */
int gpu_bpf_batch_autotune(GpuBpfBatch* batch, const float* y, int ticks);

/* Advance every filter by n_ticks observations (host or device array);
 * returns 0 or -1.  Ticks run as replays of one captured HIP graph. */
int gpu_bpf_batch_run(GpuBpfBatch* batch, const float* y, int n_ticks);
int gpu_bpf_batch_run_device(GpuBpfBatch* batch, const float* y_dev, int n_ticks);
int gpu_bpf_batch_synchronize(GpuBpfBatch* batch);
int gpu_bpf_batch_get_timestep(GpuBpfBatch* batch);

/* Per-tick statistics of the last n_ticks ticks, tick-major
 * (out[t * n_filters + m]); at most the ring capacity (1024 ticks). */
int gpu_bpf_batch_copy_stats(GpuBpfBatch* batch, BpfTickStats* out, int n_ticks);
/* FP64 cumulative log-likelihood of each filter. */
int gpu_bpf_batch_get_loglik(GpuBpfBatch* batch, double* ll);
int gpu_bpf_batch_get_resample_counts(GpuBpfBatch* batch, int* counts);
/* The current particles of filter m (n_particles[m] floats). */
int gpu_bpf_batch_copy_particles(GpuBpfBatch* batch, int filter, float* x);

#ifdef __cplusplus
}
#endif
#endif /* GPU_BPF_BATCH_H */
