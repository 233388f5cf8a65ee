/*
 * rv_sparse_learned.h — the learned per-bin accumulator selector
 * (RVSP_SPGEMM_ALGO_HIP_LEARNED) and the network kernels behind it.
 *
 * No counterpart in the reference. Its nearest relative is the ADAPTIVE
 * accumulator (reference src/backend/rvv/accum_adaptive_f32.c:213-278), which
 * picks an accumulator per row from a fixed structural rule; here a small
 * multilayer perceptron picks, per bin of rows, which accumulator kernel
 * serves the bin (LDS hash, LDS bitmap, LDS search, workgroup LDS block,
 * global dense), with how many lanes per row and which LDS table size.
 *
 * The network only chooses among kernel shapes. Every choice passes through
 * the same veto as the fixed table: when the exact symbolic count shows that
 * the chosen table cannot hold a bin's largest row, the bin runs on the
 * global dense accumulator. Output bits never depend on the network.
 *
 * The last input feature says whether the choice is for numeric-only
 * computes that reuse an analysed structure (the first such compute of a
 * HIP_LEARNED descriptor selects its numeric bins again with it set, and may
 * then pick the position-map classes, which pay off only over repeats).
 *
 * Training data comes from the library itself: rvsp_spgemm_autotune times
 * every candidate shape of every non-empty bin and appends one sample per bin
 * (features, time per class) to a process-wide sample store. rvsp_learned_fit
 * trains the process-wide model on the device with the forward / backward /
 * Adam kernels that rvsp_mlp_train exposes directly.
 */
#ifndef RV_SPARSE_LEARNED_H
#define RV_SPARSE_LEARNED_H

#include <stddef.h>
#include <stdint.h>

#include "rv_sparse.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Network dimensions: features -> hidden (ReLU) -> hidden (ReLU) -> classes. */
#define RVSP_MLP_FEATURES 17
#define RVSP_MLP_HIDDEN 64
#define RVSP_MLP_CLASSES 56
/* Parameter count. Layout, row-major: W1[H][F], b1[H], W2[H][H], b2[H],
 * W3[K][H], b3[K]. */
#define RVSP_MLP_PARAMS                                                   \
    (RVSP_MLP_HIDDEN * RVSP_MLP_FEATURES + RVSP_MLP_HIDDEN +              \
     RVSP_MLP_HIDDEN * RVSP_MLP_HIDDEN + RVSP_MLP_HIDDEN +                \
     RVSP_MLP_CLASSES * RVSP_MLP_HIDDEN + RVSP_MLP_CLASSES)

/* What one output class of the network stands for. cap_mult scales the bin's
 * base LDS table size (hash classes: 1 or 2; others 1); cap is the tile
 * size in positions of the long-row classes (RVSP_KIND_BLOCK; variant 0:
 * workgroup tiles, variant 1: wave tiles), 0 for the others. */
typedef struct
{
    int32_t kind;     /* rvsp_kind_t */
    int32_t group;    /* lanes per row */
    int32_t cap_mult;
    int32_t variant;
    int32_t cap;
} rvsp_learned_class_t;

/* RVSP_ERROR_INVALID_ARGUMENT when cls is outside [0, RVSP_MLP_CLASSES). */
rvsp_status_t rvsp_learned_class_info(int32_t cls, rvsp_learned_class_t *out);

/* One training sample: one non-empty bin of one phase of one analysed
 * product. us[c] is the measured time of class c in microseconds, or a
 * negative number when class c was not a candidate for the bin. */
typedef struct
{
    int32_t phase;        /* rvsp_phase_t */
    int32_t bin;
    int32_t fixed_class;  /* the class the fixed table uses for this bin */
    int32_t n_candidates;
    float features[RVSP_MLP_FEATURES];
    float us[RVSP_MLP_CLASSES];
} rvsp_learned_sample_t;

int32_t rvsp_learned_sample_count(void);
rvsp_status_t rvsp_learned_sample_get(int32_t i, rvsp_learned_sample_t *out);
rvsp_status_t rvsp_learned_sample_add(const rvsp_learned_sample_t *s);
void rvsp_learned_sample_clear(void);

/* Trains the process-wide model on samples [first, first + count) with
 * `epochs` full-batch Adam steps at learning rate lr from a seeded
 * initialisation. The loss of the last step is stored in *final_loss (may be
 * NULL). RVSP_ERROR_INVALID_ARGUMENT for an empty or out-of-range slice. */
rvsp_status_t rvsp_learned_fit(int32_t first, int32_t count, int32_t epochs,
                               float lr, uint64_t seed, float *final_loss);

/* Model persistence (text). path NULL means $RVSP_LEARNED_MODEL, else
 * learned_model.txt next to the library, which HIP_LEARNED estimation also
 * reads on first use. save/load return 0 or -1. */
int32_t rvsp_learned_model_save(const char *path);
int32_t rvsp_learned_model_load(const char *path);
void rvsp_learned_model_clear(void);
int32_t rvsp_learned_model_ready(void);

/* Runs the process-wide model on n raw feature vectors (n x
 * RVSP_MLP_FEATURES) and writes the highest-scoring class among those whose
 * bit is set in mask[i] (bit c = class c allowed), or -1 when none is. */
rvsp_status_t rvsp_learned_predict(const float *features,
                                   const uint64_t *mask, int32_t n,
                                   int32_t *cls);

/*
 * The network kernels with explicit parameters (host arrays; the device work
 * is synchronous). x holds n normalised feature vectors, q n target
 * distributions over the classes, mask n allowed-class bit sets and w n
 * sample weights. The loss is sum_i w[i] * CE(q[i], softmax over allowed
 * classes of the logits of x[i]).
 */

/* Logits (n x RVSP_MLP_CLASSES, may be NULL) and the masked argmax class of
 * every sample (may be NULL). */
rvsp_status_t rvsp_mlp_forward(const float *params, const float *x,
                               const uint64_t *mask, int32_t n,
                               float *logits, int32_t *cls);

/* Loss and its gradient with respect to every parameter. */
rvsp_status_t rvsp_mlp_gradient(const float *params, const float *x,
                                const float *q, const uint64_t *mask,
                                const float *w, int32_t n, float *grad,
                                float *loss);

/* `epochs` Adam steps (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected) on
 * the full batch, updating params in place. loss_out[e] (may be NULL)
 * receives the loss at the parameters before step e. */
rvsp_status_t rvsp_mlp_train(float *params, const float *x, const float *q,
                             const uint64_t *mask, const float *w, int32_t n,
                             int32_t epochs, float lr, float *loss_out);

#ifdef __cplusplus
}
#endif

#endif /* RV_SPARSE_LEARNED_H */
