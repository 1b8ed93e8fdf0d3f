/*
 * kmig_nn.h — C ABI of the learned traveltime surrogate of libkmig.so.
 *
 * The reference has no counterpart: its traveltimes always come from
 * fast-marching tables (src/eikods2d.cpp) read slice by slice
 * (src/kirchmig2d_cig.cpp:274-284) and interpolated between table positions
 * (SERECKIRCH/src/tinterpsr.cpp:142-188).  The surrogate is a coordinate
 * MLP that stands in for that table interpolation: it maps (position s,
 * depth z, lateral x) to the traveltime T and its derivatives dT/dx, dT/dz
 * and dT/ds, is trained on fast-marching tables by its own BF16-MFMA forward
 * and backward kernels (FP32 accumulation, FP32 master weights, Adam), and is
 * evaluated inside the migration's tuple preparation, where the stored
 * tables remain the per-slot fallback and the judge of its error
 * (kmig_set_surrogate in kmig.h).
 *
 * Network: inputs s, z, x scaled to [0, 1] over the configured ranges, each
 * expanded to 16 Fourier features sin(pi 2^k u), cos(pi 2^k u), k = 0..7
 * (48 features); four hidden layers of width 128 with SiLU; a linear output
 * y[4].  With d = sqrt((x - s)^2 + (z - z_src)^2) and S0 the reference
 * slowness:  T = d * S0 * (1 + y0),  dT/dx = S0 * y1,  dT/dz = S0 * y2,
 * dT/ds = S0 * y3 (the distance factor carries the source singularity).
 *
 * Parameter vector (FP32), in this order: W1[128][48], b1[128],
 * W2[128][128], b2[128], W3, b3, W4, b4, W5[4][128], b5[4];
 * KMIG_NN_PARAMS values in all.
 *
 * Training records are 7 floats: s, z, x, T, dT/dx, dT/dz, dT/ds.
 * Loss per record: ((T' - T) / sigma_t)^2 + sum over the three derivatives
 * of ((g' - g) / sigma_g)^2, averaged over the batch.
 *
 * Conventions as in kmig.h: 0 = OK, negative = KMIG_E_* class, one network
 * per host thread, all device work on the network's own HIP stream.
 */
#ifndef KMIG_NN_H
#define KMIG_NN_H

#ifdef __cplusplus
extern "C" {
#endif

#define KMIG_NN_FEATURES 48
#define KMIG_NN_WIDTH 128
#define KMIG_NN_LAYERS 4
#define KMIG_NN_OUTPUTS 4
#define KMIG_NN_PARAMS 56324

typedef struct kmig_nn kmig_nn;

typedef struct {
    float s_min, s_max;   /* position range of the scaling (km)        */
    float z_min, z_max;   /* depth range                                */
    float x_min, x_max;   /* lateral range                              */
    float z_src;          /* depth of the table positions (km)          */
    float slowness;       /* reference slowness S0 (s/km)               */
    float sigma_t;        /* loss scale of T (s)                        */
    float sigma_g;        /* loss scale of the derivatives (s/km)       */
    float lr;             /* Adam step size at the start of a train call */
    float lr_final;       /* at its end (cosine schedule)                */
    float beta1, beta2, eps;
    int batch;            /* records per training step (multiple of 64) */
    unsigned long long seed; /* weight initialisation and batch sampling */
} kmig_nn_config;

int kmig_nn_config_default(kmig_nn_config *cfg);
int kmig_nn_create(kmig_nn **out, const kmig_nn_config *cfg);
/* Training records [n][7], copied to the device (replaces earlier ones). */
int kmig_nn_set_data(kmig_nn *nn, const float *records, long n);
/* `steps` Adam steps on batches drawn uniformly from the records; the whole
   step is one captured HIP graph.  loss (nullable) receives the batch loss of
   every step. */
int kmig_nn_train(kmig_nn *nn, int steps, float *loss);
/* Forward pass at points [n][3] (s, z, x) -> out [n][4] (T, dT/dx, dT/dz,
   dT/ds). */
int kmig_nn_eval(kmig_nn *nn, const float *points, long n, float *out);
/* Loss and its gradient with respect to the parameter vector for the given
   records (one forward and backward pass, no update). */
int kmig_nn_grad(kmig_nn *nn, const float *records, long n, float *grad, float *loss);
int kmig_nn_get_params(const kmig_nn *nn, float *params);
int kmig_nn_set_params(kmig_nn *nn, const float *params);
/* Device time of the last kmig_nn_train call (ms) and its GEMM FLOPs. */
int kmig_nn_stats(const kmig_nn *nn, double *train_ms, double *train_flops);
const char *kmig_nn_last_error(const kmig_nn *nn);
void kmig_nn_destroy(kmig_nn *nn);

#ifdef __cplusplus
}
#endif

#endif
