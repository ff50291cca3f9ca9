/* rbpf_learn.c -- a plain C caller of include/smc2_rbpf.h following the
 * learning flow of smc2/test_smc2_rbpf.cu:250-335 (alloc, seed, noise
 * capacity, fixed lag, init from prior, one update per observation,
 * posterior mean / std, counters), linked against libgpubpf.so.
 *
 * usage: rbpf_learn <y.f32> <T> <N_theta> <N_inner> <seed> <lag>
 * prints: one line of theta mean and std (%.9g), one line of counters. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "gpu_bpf.h"
#include "smc2_rbpf.h"

int main(int argc, char** argv) {
    if (argc != 7) {
        fprintf(stderr, "usage: %s y.f32 T N_theta N_inner seed lag\n", argv[0]);
        return 2;
    }
    const int T = atoi(argv[2]), nt = atoi(argv[3]), ni = atoi(argv[4]), lag = atoi(argv[6]);
    const unsigned long long seed = strtoull(argv[5], NULL, 10);
    float* y = malloc(sizeof(float) * (size_t)T);
    FILE* f = fopen(argv[1], "rb");
    if (!f || !y || fread(y, sizeof(float), (size_t)T, f) != (size_t)T) {
        fprintf(stderr, "cannot read %d observations from %s\n", T, argv[1]);
        return 2;
    }
    fclose(f);
    SMC2StateCUDA* s = smc2_cuda_alloc(nt, ni);
    if (!s) {
        fprintf(stderr, "smc2_cuda_alloc: %s\n", gpu_bpf_last_error());
        return 1;
    }
    smc2_cuda_set_seed(s, seed);
    smc2_cuda_set_noise_capacity(s, T + 128);
    smc2_cuda_set_fixed_lag(s, lag);
    smc2_cuda_init_from_prior(s);
    for (int t = 0; t < T; ++t) {
        const float ess = smc2_cuda_update(s, y[t]);
        if (isnan(ess)) {
            fprintf(stderr, "smc2_cuda_update at %d: %s\n", t, gpu_bpf_last_error());
            return 1;
        }
    }
    float m[SMC2_N_PARAMS], sd[SMC2_N_PARAMS];
    smc2_cuda_get_theta_mean(s, m);
    smc2_cuda_get_theta_std(s, sd);
    SMC2Counters c;
    if (smc2_cuda_get_counters(s, &c) != 0) return 1;
    for (int q = 0; q < SMC2_N_PARAMS; ++q) printf("%.9g %.9g ", m[q], sd[q]);
    printf("\n%d %d %d %d %d\n", c.n_resamples, c.n_rejuv_accepts, c.n_rejuv_total, c.t_current,
           c.t_checkpoint);
    smc2_cuda_free(s);
    free(y);
    return 0;
}
