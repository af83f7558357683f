/*
 * ymx_guide.h -- C ABI of the path-guiding network (libymx.so).
 *
 * No counterpart in the reference: libYafaRay's path tracer samples the next
 * direction from the BSDF only (src/integrator/surface/integrator_path_tracer.cc:190-240).
 * This is the learned proposal distribution the path tracer can mix in by
 * one-sample MIS (ymx_pt_set_guide in ymx_pt.h):
 *   input   a multi-resolution hash encoding of the position (16 levels of
 *           2^log2_table_size entries, 2 BF16 features each, trilinear) plus the
 *           surface normal, the outgoing direction and a constant 1;
 *   network 3 hidden ReLU layers of width 64 and a 40-wide output, BF16 weights
 *           and activations, FP32 accumulation (MFMA 16x16x32 BF16), no biases;
 *   output  an 8-lobe von Mises-Fisher mixture on the sphere;
 *   training the neural-importance-sampling loss -c log q(w) over samples
 *           (position, normal, wo, w, c) that the path tracer writes, Adam on
 *           FP32 master weights.
 * Conventions as in ymx_accel.h: int status, ymx_last_error(), device
 * buffers, hipStream_t as void *.  A handle's calls must not overlap.
 */
#ifndef YMX_GUIDE_H
#define YMX_GUIDE_H

#include <stdint.h>

#include "ymx_accel.h"

#ifdef __cplusplus
extern "C" {
#endif

#define YMX_GUIDE_LEVELS 16
#define YMX_GUIDE_FEATURES 2
#define YMX_GUIDE_LOBES 8
#define YMX_GUIDE_MLP_WEIGHTS (64 * 64 * 3 + 48 * 64) /* W1, W2, W3 (64 x 64), W4 (48 x 64), row-major [out][in] */
#define YMX_GUIDE_MIX_FLOATS 48  /* per query: 8 lobes x (mu x, y, z, kappa, weight, vMF normalisation) */

typedef struct ymx_guide_params
{
	uint32_t struct_size;
	uint32_t log2_table_size;   /* entries per level: 2^log2 (19 for the full network) */
	int32_t n_min, n_max;       /* coarsest and finest grid resolution (16, 2048) */
	float bmin[3], bmax[3];     /* scene box mapped onto the unit cube */
	uint32_t max_batch;         /* most training samples per step */
	uint32_t seed;              /* weight initialisation */
	float lr, beta1, beta2, eps; /* Adam */
} ymx_guide_params;

/* One training sample (64 B); a query uses pos, normal and wo only. */
typedef struct ymx_guide_sample
{
	float pos[3], c;            /* position; target weight c >= 0 */
	float normal[3], pad0;
	float wo[3], pad1;
	float w[3], pad2;           /* the sampled direction (the pad fields are not read by the guide) */
} ymx_guide_sample;

typedef struct ymx_guide ymx_guide;

int ymx_guide_create(const ymx_guide_params *params, ymx_guide **out);
int ymx_guide_destroy(ymx_guide *g);

/* Replace / read the FP32 master weights: grid[level][entry][feature] (16 * 2^log2 * 2
 * floats) and the MLP (YMX_GUIDE_MLP_WEIGHTS floats).  Host buffers, synchronous;
 * set also resets the Adam moments. */
int ymx_guide_set_weights(ymx_guide *g, const float *grid, const float *mlp);
int ymx_guide_get_weights(ymx_guide *g, float *grid, float *mlp);

/* Mixture of each query: mix[48 * i] (see YMX_GUIDE_MIX_FLOATS).  n_dev, if not
 * NULL, is a device counter that replaces n (n is then the capacity). */
int ymx_guide_infer(ymx_guide *g, const ymx_guide_sample *queries, float *mix, uint32_t n, const uint32_t *n_dev,
                    void *stream);

/* Gradients of the summed loss over `n` samples into the handle's gradient
 * buffers (cleared first); ymx_guide_get_gradients copies them to the host
 * (layout as the weights) with the loss sum. */
int ymx_guide_backward(ymx_guide *g, const ymx_guide_sample *samples, uint32_t n, const uint32_t *n_dev,
                       void *stream);
int ymx_guide_get_gradients(ymx_guide *g, float *grid, float *mlp, double *loss);

/* One Adam step with the current gradients; ymx_guide_train_step = backward + step. */
int ymx_guide_adam_step(ymx_guide *g, void *stream);
int ymx_guide_train_step(ymx_guide *g, const ymx_guide_sample *samples, uint32_t n, const uint32_t *n_dev,
                         void *stream);

#ifdef __cplusplus
}
#endif
#endif
