// scene_gen.cpp -- Poisson-disk sampling of boundary surfaces (host code).
//
// The reference builds its boundaries from lattice shells and OBJ meshes
// (IISPH.py:857-985, model.py:14-60); the torus-and-plane workload asks for
// surfaces sampled with a minimum spacing instead. This is Bridson's
// algorithm ("Fast Poisson disk sampling in arbitrary dimensions", SIGGRAPH
// 2007 sketch) run on a surface: from a random active sample, up to 30
// candidates at distance [d, 2d) along the tangent plane are projected onto
// the surface and accepted when no sample lies closer than d (3-D distance,
// background grid of cells d/sqrt(3) wide, so a cell holds at most one
// sample); a sample whose candidates all fail leaves the active list.
// Random numbers come from a splitmix64 stream, so the result depends only on
// the seed.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../../include/wsph.h"

namespace {

struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct V3 {
  double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// A surface: uniform (area) sampling, projection of a nearby point onto it,
// tangent basis at a surface point, bounding box.
struct Surface {
  int shape;
  double p[5];
  bool sample(Rng &g, V3 &out) const {
    if (shape == WSPH_SURF_PLANE) {
      out = {p[0] + (p[2] - p[0]) * g.uniform(), p[4], p[1] + (p[3] - p[1]) * g.uniform()};
      return true;
    }
    const double R = p[3], r = p[4];
    for (int t = 0; t < 1000; ++t) {  // area element (R + r cos phi) dtheta dphi
      const double th = 2.0 * M_PI * g.uniform(), ph = 2.0 * M_PI * g.uniform();
      if (g.uniform() * (R + r) <= R + r * cos(ph)) {
        out = torus_point(th, ph);
        return true;
      }
    }
    return false;
  }
  V3 torus_point(double th, double ph) const {
    const double R = p[3], r = p[4];
    return {p[0] + (R + r * cos(ph)) * cos(th), p[1] + r * sin(ph), p[2] + (R + r * cos(ph)) * sin(th)};
  }
  // nearest surface point (false when q is outside the surface's domain)
  bool project(V3 q, V3 &out) const {
    if (shape == WSPH_SURF_PLANE) {
      if (q.x < p[0] || q.x > p[2] || q.z < p[1] || q.z > p[3]) return false;
      out = {q.x, p[4], q.z};
      return true;
    }
    const double dx = q.x - p[0], dz = q.z - p[2];
    const double rho = sqrt(dx * dx + dz * dz);
    if (rho < 1e-12) return false;
    const V3 cc = {p[0] + p[3] * dx / rho, p[1], p[2] + p[3] * dz / rho};
    const V3 n = q - cc;
    const double ln = sqrt(dot(n, n));
    if (ln < 1e-12) return false;
    out = cc + (p[4] / ln) * n;
    return true;
  }
  void tangents(V3 s, V3 &t1, V3 &t2) const {
    if (shape == WSPH_SURF_PLANE) {
      t1 = {1, 0, 0};
      t2 = {0, 0, 1};
      return;
    }
    const double th = atan2(s.z - p[2], s.x - p[0]);
    const double rho = sqrt((s.x - p[0]) * (s.x - p[0]) + (s.z - p[2]) * (s.z - p[2]));
    const double ph = atan2(s.y - p[1], rho - p[3]);
    t1 = {-sin(th), 0.0, cos(th)};
    t2 = {-sin(ph) * cos(th), cos(ph), -sin(ph) * sin(th)};
  }
  void bounds(V3 &lo, V3 &hi) const {
    if (shape == WSPH_SURF_PLANE) {
      lo = {p[0], p[4], p[1]};
      hi = {p[2], p[4], p[3]};
      return;
    }
    const double e = p[3] + p[4];
    lo = {p[0] - e, p[1] - p[4], p[2] - e};
    hi = {p[0] + e, p[1] + p[4], p[2] + e};
  }
};

struct Grid {
  V3 lo;
  double cell;
  int nx, ny, nz;
  std::vector<int> id;
  void init(V3 l, V3 h, double c) {
    lo = l - V3{2 * c, 2 * c, 2 * c};
    cell = c;
    nx = (int)((h.x - lo.x) / c) + 5;
    ny = (int)((h.y - lo.y) / c) + 5;
    nz = (int)((h.z - lo.z) / c) + 5;
    id.assign((size_t)nx * ny * nz, -1);
  }
  bool cell_of(V3 q, int &ix, int &iy, int &iz) const {
    ix = (int)floor((q.x - lo.x) / cell);
    iy = (int)floor((q.y - lo.y) / cell);
    iz = (int)floor((q.z - lo.z) / cell);
    return ix >= 2 && iy >= 2 && iz >= 2 && ix < nx - 2 && iy < ny - 2 && iz < nz - 2;
  }
  int &at(int ix, int iy, int iz) { return id[((size_t)iz * ny + iy) * nx + ix]; }
};

}  // namespace

extern "C" int wsph_poisson_surface(int shape, const double *params, double radius, uint64_t seed,
                                    int max_points, float *out, int *count) {
  if (!params || !count || !(radius > 0) || (shape != WSPH_SURF_PLANE && shape != WSPH_SURF_TORUS))
    return WSPH_E_ARG;
  Surface S;
  S.shape = shape;
  memcpy(S.p, params, sizeof(S.p));
  if (shape == WSPH_SURF_PLANE && !(S.p[2] > S.p[0] && S.p[3] > S.p[1])) return WSPH_E_ARG;
  if (shape == WSPH_SURF_TORUS && !(S.p[3] > S.p[4] && S.p[4] > 0)) return WSPH_E_ARG;
  V3 lo, hi;
  S.bounds(lo, hi);
  Grid G;
  G.init(lo, hi, radius / sqrt(3.0));
  Rng g{seed};
  std::vector<V3> pts;
  std::vector<int> active;
  const double r2 = radius * radius;
  auto fits = [&](V3 q, int &ix, int &iy, int &iz) {
    if (!G.cell_of(q, ix, iy, iz) || G.at(ix, iy, iz) >= 0) return false;
    for (int z = iz - 2; z <= iz + 2; ++z)
      for (int y = iy - 2; y <= iy + 2; ++y)
        for (int x = ix - 2; x <= ix + 2; ++x) {
          const int j = G.at(x, y, z);
          if (j >= 0) {
            const V3 d = pts[j] - q;
            if (dot(d, d) < r2) return false;
          }
        }
    return true;
  };
  auto add = [&](V3 q, int ix, int iy, int iz) {
    G.at(ix, iy, iz) = (int)pts.size();
    active.push_back((int)pts.size());
    pts.push_back(q);
  };
  V3 first;
  int ix, iy, iz;
  if (!S.sample(g, first) || !fits(first, ix, iy, iz)) return WSPH_E_ARG;
  add(first, ix, iy, iz);
  while (!active.empty()) {
    const size_t a = (size_t)(g.uniform() * active.size()) % active.size();
    const V3 s = pts[active[a]];
    V3 t1, t2;
    S.tangents(s, t1, t2);
    bool placed = false;
    for (int k = 0; k < 30 && !placed; ++k) {
      const double ang = 2.0 * M_PI * g.uniform();
      const double len = radius * (1.0 + g.uniform());
      const V3 q0 = s + (len * cos(ang)) * t1 + (len * sin(ang)) * t2;
      V3 q;
      if (!S.project(q0, q)) continue;
      if (fits(q, ix, iy, iz)) {
        add(q, ix, iy, iz);
        placed = true;
      }
    }
    if (!placed) {
      active[a] = active.back();
      active.pop_back();
    }
  }
  *count = (int)pts.size();
  if (!out) return WSPH_OK;
  if ((int)pts.size() > max_points) return WSPH_E_CAPACITY;
  for (size_t k = 0; k < pts.size(); ++k) {
    out[3 * k] = (float)pts[k].x;
    out[3 * k + 1] = (float)pts[k].y;
    out[3 * k + 2] = (float)pts[k].z;
  }
  return WSPH_OK;
}
