// Test driver: the library's restated reference hull (cuacd-web_amd/csrc/host/qhull.cpp) as a C
// call, compiled with it by tests/test_hull_clean.py and compared there with the reference's
// own quickhull (oracle/_ref).
#include <cstdint>
#include <vector>

#include "../cuacd-web_amd/csrc/host/qhull.hpp"

extern "C" int our_qh_tris(const double* pts, uint32_t n, double eps, uint32_t* out, uint32_t cap) {
  std::vector<uint32_t> t;
  if (!cuacd::hullEps(pts, n, eps, &t)) return 0;
  const uint32_t nt = (uint32_t)(t.size() / 3);
  for (uint32_t i = 0; i < 3 * nt && i < 3 * cap; i++) out[i] = t[i];
  return (int)nt;
}
