// Neural radiance cache: internal interface shared with the integrator.
#pragma once

#include "internal.hpp"

namespace pyre {

// Scene box used to normalise query positions to [0, 1]^3.
struct NrcBounds {
  float lo[3];
  float inv_ext[3];
};

int nrc_param_count();
// Forward + relative-L2 loss + backward into the cache's gradient buffer;
// feats (13 floats per query) and targets (RGB) are device arrays. Returns
// the mean loss, or, when loss_dev is given, adds it to *loss_dev on the
// stream without synchronising (and returns 0).
float nrc_gradients_dev(pyre_nrc* nrc, const float* feats, const float* targets, uint32_t n, hipStream_t st,
                        float* loss_dev = nullptr);
// One Adam step (Kingma & Ba) on a batch: the gradients as above, then the
// FP32 master weights, moments and bf16 fragments updated in one launch; the
// loss is added to *loss_dev when given, else returned (read_back, which
// synchronises) or not computed on the host.
float nrc_train_step_dev(pyre_nrc* nrc, const float* feats, const float* targets, uint32_t n, hipStream_t st,
                         float* loss_dev = nullptr, bool read_back = true);
// One Adam step queued before its batch size is known: the samples
// [off, off + batch) of the min(*count, cap) in feats/targets, skipped on the
// device when fewer than 16 remain. Adds the batch's mean loss to *loss_dev
// and counts the step in *steps_dev (either may be null). No host sync.
void nrc_train_step_counted(pyre_nrc* nrc, const float* feats, const float* targets, const uint32_t* count,
                            uint32_t cap, uint32_t off, uint32_t batch, hipStream_t st, float* loss_dev,
                            uint32_t* steps_dev);
// Records the loss reported by pyre_nrc_stats for training run elsewhere.
void nrc_set_last_loss(pyre_nrc* nrc, float loss);
// Fused BF16-MFMA inference: rgb[3n] from feats[13n] (device arrays).
void nrc_infer_dev(pyre_nrc* nrc, const float* feats, uint32_t n, float* rgb, hipStream_t st);
// Inference over a device-side query count (*n_dev, at most cap queries):
// query i reads its 13 features at row rows[i] of feats, writes rgb[3i..].
void nrc_infer_dev_count(pyre_nrc* nrc, const float* feats, const uint32_t* rows, const uint32_t* n_dev, uint32_t cap,
                         float* rgb, hipStream_t st);

}  // namespace pyre
