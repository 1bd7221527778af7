// Learned kernel selector: a 3-layer perceptron (ReLU, width 64, FP32) that maps the features
// of a problem (runtime.hip: features_of) to the predicted log2 slowdown of every kernel
// variant relative to the fastest one.  The planner picks the valid variant with the smallest
// prediction.  The same parameter layout is used by the GPU training kernels (mlp.hip), the
// host inference below and the oracle's FP64 restatement.
//
// Parameter vector (floats), in this order:
//   mean[F], inv_std[F]          feature standardisation  x' = (x - mean) * inv_std
//   W1[H][F], b1[H]              h1 = relu(W1 x' + b1)
//   W2[H][H], b2[H]              h2 = relu(W2 h1 + b2)
//   W3[V][H], b3[V]              y  = W3 h2 + b3
#pragma once

#include <cstddef>

namespace ttvx
{
constexpr int kMlpF = 24;   // TTVX_NUM_FEATURES
constexpr int kMlpH = 64;
constexpr int kMlpV = 15;   // number of kernel variants (V_COUNT)
constexpr std::size_t kMlpOffMean = 0;
constexpr std::size_t kMlpOffStd  = kMlpOffMean + kMlpF;
constexpr std::size_t kMlpOffW1   = kMlpOffStd + kMlpF;
constexpr std::size_t kMlpOffB1   = kMlpOffW1 + static_cast<std::size_t>(kMlpH) * kMlpF;
constexpr std::size_t kMlpOffW2   = kMlpOffB1 + kMlpH;
constexpr std::size_t kMlpOffB2   = kMlpOffW2 + static_cast<std::size_t>(kMlpH) * kMlpH;
constexpr std::size_t kMlpOffW3   = kMlpOffB2 + kMlpH;
constexpr std::size_t kMlpOffB3   = kMlpOffW3 + static_cast<std::size_t>(kMlpV) * kMlpH;
constexpr std::size_t kMlpParams  = kMlpOffB3 + kMlpV;
constexpr std::size_t kMlpTrainable = kMlpParams - kMlpOffW1;   // everything after the standardisation

// Host forward pass of one feature vector (used per call by the planner).
void mlp_forward_host(const float* params, const float* x, float* y);
} // namespace ttvx
