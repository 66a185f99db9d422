/* Neural fragment shading: a small network (srp/neural_abi.h) trained on
 * the GPU, from fragments the ordinary pipeline renders, to stand in for an
 * expensive procedural fragment shader.  Coverage, depth and stencil never
 * pass through it: a shader module's fragment shader calls it for colour
 * only (srp/device/neural.hpp), and the procedural shader stays the
 * reference it is evaluated against.
 *
 * These calls have no counterpart in the reference renderer; they extend
 * the library for the learned shader.  All work is queued on the library
 * stream (srpHipGetStream) and calls that return host data synchronise it.
 * Misuse is reported through the message callback, like the core API. */
#pragma once

#include <stddef.h>
#include <stdint.h>
#include "srp/framebuffer.h"
#include "srp/neural_abi.h"
#include "srp/detail/qualifiers.h"

SRP_EXTERN_C_BEGIN

typedef struct SRPNeuralDataset SRPNeuralDataset;
typedef struct SRPNeuralNet SRPNeuralNet;

/* ------------------------------------------------------------ training data */

/** Device storage for up to `capacity` samples (8 features, RGB target). */
SRPNeuralDataset* srpNewNeuralDataset(size_t capacity);
void srpFreeNeuralDataset(SRPNeuralDataset* ds);

/** Copy the device colour of `fb` into staging slot 0..3 (G-buffer capture
 *  pass k: features 2k, 2k+1 as 16-bit pairs) or 4 (the shaded colour). */
void srpNeuralDatasetStage(SRPNeuralDataset* ds, const SRPFramebuffer* fb, int slot);
/** Append the staged frame's pixels whose depth in `fb` is below 1, each
 *  kept when a hash of (seed, pixel) falls below `keep` (0..1).  Features
 *  are decoded from slots 0..3 (position and UV over [0,1], normal over
 *  [-1,1]), the target from slot 4.  Returns the number appended (stops at
 *  the capacity). */
size_t srpNeuralDatasetAppend(SRPNeuralDataset* ds, const SRPFramebuffer* fb, float keep, uint64_t seed);
size_t srpNeuralDatasetSize(const SRPNeuralDataset* ds);
/** Host copies: features[count][8], targets[count][3]. */
void srpNeuralDatasetRead(const SRPNeuralDataset* ds, size_t first, size_t count, float* features, float* targets);
/** Replace the contents with `count` host samples. */
void srpNeuralDatasetWrite(SRPNeuralDataset* ds, const float* features, const float* targets, size_t count);

/* ----------------------------------------------------------------- network */

/** Parameters drawn uniformly in +-1/sqrt(fan_in) from `seed`; Adam state 0. */
SRPNeuralNet* srpNewNeuralNet(uint64_t seed);
void srpFreeNeuralNet(SRPNeuralNet* net);
/** Canonical fp32 parameters (SRP_NN_PARAMS values, layout of neural_abi.h). */
void srpNeuralNetGetParams(const SRPNeuralNet* net, float* params);
/** Set the parameters (also resets the optimiser state). */
void srpNeuralNetSetParams(SRPNeuralNet* net, const float* params);

/** `steps` Adam steps (beta 0.9/0.999, eps 1e-8) on minibatches of `batch`
 *  samples (a multiple of 1024, at most 65536) drawn with replacement from
 *  the dataset by a hash of (seed, step, slot).  The learning rate follows a
 *  cosine from lr0 to lr1 over the call.  Loss = mean squared error over the
 *  batch's RGB values; lossOut[step] receives it when not NULL. */
void srpNeuralNetTrain(SRPNeuralNet* net, const SRPNeuralDataset* ds, int steps, int batch, float lr0, float lr1,
                       uint64_t seed, float* lossOut);
/** Gradient of the loss over dataset samples [0, count) (count a multiple
 *  of 1024, at most 65536) with respect to the canonical parameters; no
 *  update.  *loss receives the loss. */
void srpNeuralNetGradients(SRPNeuralNet* net, const SRPNeuralDataset* ds, size_t count, float* grad, float* loss);
/** Network output for dataset samples [0, count): rgb[count][3].
 *  lanes = 0: matrix-core path (64 samples per wave); 1: per-lane path. */
void srpNeuralNetInfer(const SRPNeuralNet* net, const SRPNeuralDataset* ds, size_t count, float* rgb, int lanes);
/** Device time of the matrix-core inference over dataset samples [0, count)
 *  (the network alone, dense: 64 samples per wave), mean ms per pass over
 *  `reps` passes after one warm-up pass. */
double srpNeuralNetTimeInference(const SRPNeuralNet* net, const SRPNeuralDataset* ds, size_t count, int reps);
/** Device address of the network's SrpNnDevice view, for a uniform. */
const void* srpNeuralNetDevice(const SRPNeuralNet* net);
/** Frequency step of each of the SRP_NN_IN features' encodings (feature d
 *  at frequency j enters as sin/cos of 2*pi*(j+1)*step[d]*x_d); every step
 *  in (0, 1024), else an error message and no change.  A new network has
 *  SRP_NN_FREQ_STEP for every feature.  Parameters are kept. */
void srpNeuralNetSetEncoding(SRPNeuralNet* net, const float* step);
void srpNeuralNetGetEncoding(const SRPNeuralNet* net, float* step);

/* -------------------------------------------------------------- evaluation */

/** Compare the RGB of two framebuffers over the pixels whose depth in
 *  `mask` is below 1: PSNR in dB (peak 255), largest channel difference
 *  (0..255), and the pixel count. */
void srpNeuralCompare(const SRPFramebuffer* a, const SRPFramebuffer* b, const SRPFramebuffer* mask,
                      double* psnr, double* maxAbs, uint64_t* pixels);
/** srpNeuralCompare without allocating anything (no hipMalloc or hipFree,
 *  no allocating library call): for callers that cannot allocate, such as
 *  tight evaluation loops and harnesses that forbid device allocation.  It
 *  has no counterpart in the reference.  Its partial sums live in a static
 *  24 KiB device array per GPU (kept until the process ends) and a static
 *  host array; calls that overlap on one GPU share them and race.  Any
 *  framebuffer size fits (the grid strides; at most 1,024 blocks of
 *  partials).  Same results as srpNeuralCompare. */
void srpNeuralCompareNoAlloc(const SRPFramebuffer* a, const SRPFramebuffer* b, const SRPFramebuffer* mask,
                             double* psnr, double* maxAbs, uint64_t* pixels);

SRP_EXTERN_C_END
