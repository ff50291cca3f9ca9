/* bpf_step.c -- a plain C caller of include/gpu_bpf.h following the
 * reference's per-tick flow (tests/test_bpf_matched_dgp.cu: create, one
 * gpu_bpf_step per observation, resample count, destroy), linked against
 * libgpubpf.so.
 *
 * usage: bpf_step <y.f32> <T> <N> <seed> <out.f32>
 * writes h_mean and log_lik of every tick (2T floats) to out.f32 and prints
 * the resample count. */
#include <stdio.h>
#include <stdlib.h>

#include "gpu_bpf.h"

int main(int argc, char** argv) {
    if (argc != 6) {
        fprintf(stderr, "usage: %s y.f32 T N seed out.f32\n", argv[0]);
        return 2;
    }
    const int T = atoi(argv[2]), n = atoi(argv[3]), seed = atoi(argv[4]);
    float* y = malloc(sizeof(float) * (size_t)T);
    float* out = malloc(sizeof(float) * 2 * (size_t)T);
    FILE* f = fopen(argv[1], "rb");
    if (!f || !y || !out || fread(y, sizeof(float), (size_t)T, f) != (size_t)T) {
        fprintf(stderr, "cannot read %d observations from %s\n", T, argv[1]);
        return 2;
    }
    fclose(f);
    GpuBpfState* s = gpu_bpf_create(n, 0.97f, 0.15f, -1.0f, 0.0f, 0.0f, seed);
    if (!s) {
        fprintf(stderr, "gpu_bpf_create: %s\n", gpu_bpf_last_error());
        return 1;
    }
    for (int t = 0; t < T; ++t) {
        const BpfResult r = gpu_bpf_step(s, y[t]);
        out[2 * t] = r.h_mean;
        out[2 * t + 1] = r.log_lik;
    }
    printf("%d\n", gpu_bpf_get_resample_count(s));
    gpu_bpf_destroy(s);
    f = fopen(argv[5], "wb");
    if (!f || fwrite(out, sizeof(float), 2 * (size_t)T, f) != 2 * (size_t)T) return 2;
    fclose(f);
    free(y);
    free(out);
    return 0;
}
