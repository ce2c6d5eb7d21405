// guide_net.h — device pointers and constants of the path-guiding network,
// shared by the host API (which owns the buffers) and the kernels.
#pragma once
#include <cstdint>

namespace xnh {

constexpr int kGLevels = 16;   // hash-grid levels
constexpr int kGFeat = 2;      // features per level
constexpr int kGIn = 48;       // 32 grid features + 16 SH coefficients
constexpr int kGHid = 64;      // hidden width (3 hidden layers)
constexpr int kGOut = 48;      // 8 lobes x 5 raw values, padded to a multiple of 16
constexpr int kGLobes = 8;
constexpr int kMixVecs = 12;   // float4 per stored mixture (guide.h)

// Offsets of the four row-major [in][out] weight matrices.
constexpr int kW1 = 0, kW2 = kW1 + kGIn * kGHid, kW3 = kW2 + kGHid * kGHid, kW4 = kW3 + kGHid * kGHid,
              kWTotal = kW4 + kGHid * kGOut;

struct GuideNet {
  const uint16_t* grid;   // FP16 bits [level][slot][feature]
  const uint16_t* w;      // FP16 bits, kWTotal, row-major per layer
  const uint16_t* wt;     // FP16 bits, kWTotal, as MFMA operand tiles (guide.h)
  float lo[3], inv_ext[3];   // position normalisation to the unit cube
  int32_t t_log2;            // table entries per level = 2^t_log2
  int32_t res[kGLevels];     // lattice resolution per level
};

}  // namespace xnh
