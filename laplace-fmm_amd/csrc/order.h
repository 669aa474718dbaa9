// order.h -- per-box expansion orders: the order network's layout and the
// host side of order.hip.
#pragma once
#include <cstdint>

namespace lfmm {

// The order network: 8 leaf features -> 64 -> 64 -> 13 classes (FP32).
// Class c is order p = 4 + c, i.e. kOrderMinTerms + c terms.
constexpr int kOrderFeatures = 8;
constexpr int kOrderHidden = 64;
constexpr int kOrderClasses = 13;
constexpr int kOrderMinTerms = 5;
// Parameter block (floats): feature normalisation mu[8], sd[8] (fixed, set
// from the training data), then W1[64][8], b1[64], W2[64][64], b2[64],
// W3[13][64], b3[13] (trained).
constexpr int kOffMu = 0;
constexpr int kOffSd = kOffMu + kOrderFeatures;
constexpr int kOffW1 = kOffSd + kOrderFeatures;
constexpr int kOffB1 = kOffW1 + kOrderHidden * kOrderFeatures;
constexpr int kOffW2 = kOffB1 + kOrderHidden;
constexpr int kOffB2 = kOffW2 + kOrderHidden * kOrderHidden;
constexpr int kOffW3 = kOffB2 + kOrderHidden;
constexpr int kOffB3 = kOffW3 + kOrderClasses * kOrderHidden;
constexpr int kOrderModelSize = kOffB3 + kOrderClasses;

}  // namespace lfmm
