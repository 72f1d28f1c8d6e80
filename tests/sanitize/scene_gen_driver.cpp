// Driver of the host scene generator (warpsph_amd/csrc/scene_gen.cpp,
// wsph_poisson_surface) for the AddressSanitizer / UndefinedBehaviorSanitizer
// build of tests/test_sanitizers.py: a plane and a torus sampled with and
// without an output buffer, a too-small buffer, and invalid arguments.
// Exits 0 when every call returns what it should.
#include <cstdio>
#include <vector>

#include "../../include/wsph.h"

int main() {
  const double plane[5] = {-0.2, -0.1, 0.3, 0.25, 0.0};
  const double torus[5] = {0.1, 0.05, -0.1, 0.12, 0.04};
  for (int shape = 0; shape < 2; ++shape) {
    const double *prm = shape == WSPH_SURF_PLANE ? plane : torus;
    int count = 0;
    if (wsph_poisson_surface(shape, prm, 0.01, 7, 0, nullptr, &count) != WSPH_OK || count <= 0) {
      std::fprintf(stderr, "count pass failed for shape %d\n", shape);
      return 1;
    }
    std::vector<float> out(3 * (size_t)count);
    int again = 0;
    if (wsph_poisson_surface(shape, prm, 0.01, 7, count, out.data(), &again) != WSPH_OK ||
        again != count) {
      std::fprintf(stderr, "fill pass failed for shape %d\n", shape);
      return 1;
    }
    if (wsph_poisson_surface(shape, prm, 0.01, 7, count - 1, out.data(), &again) !=
        WSPH_E_CAPACITY) {
      std::fprintf(stderr, "capacity check failed for shape %d\n", shape);
      return 1;
    }
  }
  int c = 0;
  const double bad_plane[5] = {0.3, 0.0, 0.1, 0.2, 0.0};
  const double bad_torus[5] = {0.0, 0.0, 0.0, 0.05, 0.1};
  if (wsph_poisson_surface(WSPH_SURF_PLANE, bad_plane, 0.01, 1, 0, nullptr, &c) != WSPH_E_ARG ||
      wsph_poisson_surface(WSPH_SURF_TORUS, bad_torus, 0.01, 1, 0, nullptr, &c) != WSPH_E_ARG ||
      wsph_poisson_surface(7, plane, 0.01, 1, 0, nullptr, &c) != WSPH_E_ARG ||
      wsph_poisson_surface(WSPH_SURF_PLANE, plane, 0.0, 1, 0, nullptr, &c) != WSPH_E_ARG) {
    std::fprintf(stderr, "argument checks failed\n");
    return 1;
  }
  std::printf("scene_gen driver ok\n");
  return 0;
}
