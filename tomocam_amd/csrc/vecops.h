// Streaming helpers over device float arrays.
#pragma once

#include "tc_common.h"

namespace tc {

// sum_i a[i] b[i] accumulated in FP64 (deterministic); synchronises `stream`.
double dot(const float* a, const float* b, long n, hipStream_t stream);

// max and min, seeded with -1e20 / +1e20 (src/dist_array.h:143-152);
// synchronises `stream`.
void minmax(const float* a, long n, float* mx, float* mn, hipStream_t stream);

// out[i] = (a[i] + add) * mul
void affine(const float* a, float* out, long n, float add, float mul,
            hipStream_t stream);

// out[i] = a[i] - b[i]
void affine_sub(const float* a, const float* b, float* out, long n,
                hipStream_t stream);

// sum_i (a[i] - b[i])^2 in FP64; synchronises `stream`.
double diff_norm2(const float* a, const float* b, long n, hipStream_t stream);

}  // namespace tc
