// learned.hpp -- shapes and layout of the light-edge threshold network.
#pragma once
#include <hip/hip_runtime.h>

namespace emst {

constexpr int kIn = 32;                 // input features
constexpr int kH = 64;                  // hidden width (two hidden layers)
// Flat fp32 parameter layout: W1[f][i], b1[i], W2[i][k], b2[k], W3[k], b3.
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kIn * kH;
constexpr int kW2 = kB1 + kH;
constexpr int kB2 = kW2 + kH * kH;
constexpr int kW3 = kB2 + kH;
constexpr int kB3 = kW3 + kH;
constexpr int kParams = kB3 + 1;        // 6337

constexpr int kWeightSamples = 1024;    // hashed weight samples (power of two)
constexpr int kVertexSamples = 4096;    // hashed vertex samples for the degree histogram
constexpr int kDegBins = 12;            // degree bins 1, 2-3, ..., 2048+
constexpr int kFeatThreads = 1024;
constexpr float kMinLog2Q = -8.0f;      // smallest quantile 2^-8
constexpr float kNoFilterQ = 0.9f;      // q at or above: unfiltered path

void launch_learned_threshold(hipStream_t s, int nodes, int edges, const int* nindex, const int* eweight,
                              const unsigned short* p16, int have_model, int* out, float* outf, float* feat,
                              int* samples, int* host_decide);

void launch_learned_batch(hipStream_t s, int graphs, const int* nindex, const int* eweight, const long long* voff,
                          const long long* eoff, const unsigned short* p16, int have_model, int* gthr, int* gflag);

}  // namespace emst
