/* mesh_oracle.c -- CPU oracle of the mesh containment test (TEST
 * INFRASTRUCTURE ONLY: the oracle of wsph_mesh_inside, never shipped).
 *
 * The reference fills a model with the np.arange lattice points that
 * trimesh 4.6.10 (a dependency absent here) reports inside the mesh
 * (model.py:14-47 -> Trimesh.contains). trimesh's published method is a ray
 * cast with a parity count of the triangles crossed. This restates it with
 * rays along +z and a deterministic rule for rays through an edge or vertex:
 * the rasterisation top-left rule applied to edge functions whose endpoints
 * are taken in lexicographic order (so the two triangles of a shared edge see
 * exactly opposite values). A point is inside when an odd number of
 * crossings lie strictly above it. Every operation is a separately rounded
 * double +, -, *, / (built with -ffp-contract=off).
 *
 * Loop structure: brute force, each point against every triangle, with no
 * column sharing or sorting. */
#include <math.h>
#include <stddef.h>

static double edge_value(double px, double py, double qx, double qy, double rx, double ry) {
  int swap = qx < px || (qx == px && qy < py);
  if (swap) {
    double t = px;
    px = qx;
    qx = t;
    t = py;
    py = qy;
    qy = t;
  }
  double e = (qx - px) * (ry - py) - (qy - py) * (rx - px);
  return swap ? -e : e;
}

/* edge P->Q of a counter-clockwise triangle, y up: downward (left) edges and
 * horizontal edges running towards -x (top edges) own their points */
static int owns_ties(double px, double py, double qx, double qy) {
  return qy < py || (qy == py && qx < px);
}

/* 1 and the crossing height in *z when the vertical line through (rx, ry)
 * crosses the triangle (a, b, c) */
static int vertical_crossing(const double *a, const double *b, const double *c, double rx,
                             double ry, double *z) {
  const double *v[3] = {a, b, c};
  /* the line must meet the triangle's closed xy bounding box */
  if (rx < fmin(fmin(a[0], b[0]), c[0]) || rx > fmax(fmax(a[0], b[0]), c[0]) ||
      ry < fmin(fmin(a[1], b[1]), c[1]) || ry > fmax(fmax(a[1], b[1]), c[1]))
    return 0;
  double area = edge_value(a[0], a[1], b[0], b[1], c[0], c[1]);
  int orient = area > 0.0 ? 1 : (area < 0.0 ? -1 : 0);
  if (orient == 0) return 0;
  double w[3];
  for (int e = 0; e < 3; ++e) { /* w[e]: the edge opposite vertex e, v[e+1] -> v[e+2] */
    const double *p = v[(e + 1) % 3], *q = v[(e + 2) % 3];
    w[e] = edge_value(p[0], p[1], q[0], q[1], rx, ry);
  }
  for (int e = 0; e < 3; ++e) {
    const double *p = v[(e + 1) % 3], *q = v[(e + 2) % 3];
    double we = orient > 0 ? w[e] : -w[e];
    int tie = orient > 0 ? owns_ties(p[0], p[1], q[0], q[1]) : owns_ties(q[0], q[1], p[0], p[1]);
    if (!(we > 0.0 || (we == 0.0 && tie))) return 0;
  }
  double s = (w[0] + w[1]) + w[2];
  *z = ((w[0] * a[2] + w[1] * b[2]) + w[2] * c[2]) / s;
  return 1;
}

int orc_mesh_inside(const double *verts, int nv, const int *tris, int nt, const double *ax, int nx,
                    const double *ay, int ny, const double *az, int nz, unsigned char *inside) {
  for (int t = 0; t < nt; ++t)
    for (int k = 0; k < 3; ++k)
      if (tris[3 * t + k] < 0 || tris[3 * t + k] >= nv) return -1;
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        int above = 0;
        for (int t = 0; t < nt; ++t) {
          double z;
          if (vertical_crossing(verts + 3 * (size_t)tris[3 * t], verts + 3 * (size_t)tris[3 * t + 1],
                                verts + 3 * (size_t)tris[3 * t + 2], ax[i], ay[j], &z) &&
              z > az[k])
            above++;
        }
        inside[((size_t)i * ny + j) * nz + k] = (unsigned char)(above & 1);
      }
  return 0;
}
