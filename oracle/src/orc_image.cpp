// orc_image.cpp — CPU oracle of the output stage: the 8-bit conversion of
// save_image (main.cpp:25-31) and the reference-free noise estimate of
// math/image.h:35-182.
//
// TEST INFRASTRUCTURE ONLY: tests/ load this library as the checker of the
// HIP image kernels; the product library never does.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../oracle_api.h"

namespace {

// image.h:127-135: Rec. 709 luminance, evaluated left to right
void luminance(const float* rgb, float* lum, int n) {
  for (int i = 0; i < n; ++i) lum[i] = 0.2126f * rgb[3 * i] + 0.7152f * rgb[3 * i + 1] + 0.0722f * rgb[3 * i + 2];
}

// image.h:152-174: 1-2-1 binomial blur with clamped borders, normalised by 16
void blur(const float* src, float* dst, int w, int h) {
  static const float k[3][3] = {{1.f, 2.f, 1.f}, {2.f, 4.f, 2.f}, {1.f, 2.f, 1.f}};
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      float s = 0.f;
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int xx = std::min(std::max(x + dx, 0), w - 1), yy = std::min(std::max(y + dy, 0), h - 1);
          s += k[dy + 1][dx + 1] * src[yy * w + xx];
        }
      dst[y * w + x] = s / 16.f;
    }
}

// image.h:176-186: k = floor(p * (n - 1)) in float, then the k-th smallest
float percentile(std::vector<float>& v, float p) {
  if (v.empty()) return 0.f;
  const size_t k = static_cast<size_t>(p * static_cast<float>(v.size() - 1));
  std::nth_element(v.begin(), v.begin() + static_cast<std::ptrdiff_t>(k), v.end());
  return v[k];
}

}  // namespace

extern "C" void orc_noise_stats(const float* rgb, int w, int h, float* o, int* oi) {
  for (int k = 0; k < 9; ++k) o[k] = 0.f;
  oi[0] = oi[1] = oi[2] = 0;
  if (!rgb || w <= 2 || h <= 2) return;   // image.h:41-42
  const int n = w * h;
  oi[1] = (w - 2) * (h - 2);
  std::vector<float> lum(n), bl(n);
  luminance(rgb, lum.data(), n);
  double sum = 0.0;   // image.h:137-146
  for (int i = 0; i < n; ++i) sum += lum[i];
  const float avg = static_cast<float>(sum / static_cast<double>(n));
  blur(lum.data(), bl.data(), w, h);
  const float edge = std::max(1e-6f, avg * 0.1f), eps = std::max(1e-8f, avg * avg * 1e-4f);
  o[6] = avg;
  o[7] = edge;
  o[8] = eps;
  std::vector<float> res;
  double s2 = 0.0, s1 = 0.0;
  float mx = 0.f;
  int cnt = 0;
  for (int y = 1; y < h - 1; ++y)   // image.h:63-98: interior pixels away from strong edges
    for (int x = 1; x < w - 1; ++x) {
      const int i = y * w + x;
      const float gx = 0.5f * (lum[i + 1] - lum[i - 1]);
      const float gy = 0.5f * (lum[i + w] - lum[i - w]);
      if (std::sqrt(gx * gx + gy * gy) >= edge) continue;
      const float r = lum[i] - bl[i];
      const float den = bl[i] * bl[i] + eps;
      const float q2 = (r * r) / den, q1 = std::fabs(r) / std::sqrt(den);
      s2 += q2;
      s1 += q1;
      res.push_back(q1);
      if (q1 > mx) mx = q1;
      ++cnt;
    }
  oi[0] = cnt;
  const float ratio = oi[1] > 0 ? static_cast<float>(cnt) / static_cast<float>(oi[1]) : 0.f;
  std::memcpy(&oi[2], &ratio, 4);
  o[5] = mx;
  if (cnt == 0) return;
  o[0] = static_cast<float>(s2 / static_cast<double>(cnt));
  o[1] = std::sqrt(o[0]);
  o[2] = static_cast<float>(s1 / static_cast<double>(cnt));
  o[3] = percentile(res, 0.50f);
  o[4] = percentile(res, 0.95f);
}

extern "C" void orc_rgb8(const float* rgb, int w, int h, uint8_t* out) {
  // main.cpp:26-31: gamma 1/2.2 of the clamped value, scaled by 255.99 and
  // truncated; the writer flips rows, so file row r is framebuffer row h-1-r
  for (int r = 0; r < h; ++r)
    for (int k = 0; k < 3 * w; ++k) {
      const float c = std::pow(std::clamp(rgb[3 * static_cast<size_t>(h - 1 - r) * w + k], 0.f, 1.f), 1.f / 2.2f);
      out[3 * static_cast<size_t>(r) * w + k] = static_cast<uint8_t>(255.99f * c);
    }
}
