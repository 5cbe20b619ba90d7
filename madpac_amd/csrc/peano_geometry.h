// Builtin geometry of the reference-shaped C++ layer: hexahedra and spheres, each possibly inverted,
// intersected into one domain, and the cell flags the block store derives from them.
//
// Reference (paths relative to madpac/peano/geometry/): builtin/Hexahedron.cpp:28-99,
// builtin/Sphere.cpp:30-210, builtin/AbstractGeometry.cpp:26-66 (inversion), Intersection.cpp:22-48
// (the domain is the intersection of the configured objects), and the flag rule of
// applications/latticeboltzmann/blocklatticeboltzmann/services/GridManagementService.cpp:86-118.
// Comparisons carry the reference's absolute tolerance of 1e-14 (tarch/la/Scalar.h:13,
// tarch/la/ScalarOperations.cpph:39-70), and points are formed with the reference's arithmetic, so
// the flags agree with the reference's at cell centres and voxel corners that lie on a surface.
// Host code only.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace madpac_amd {
namespace geometry {

using Vec3 = std::array<double, 3>;

constexpr double kTolerance = 1.0e-14;
inline bool greater(double a, double b) { return a - b > kTolerance; }
inline bool smaller(double a, double b) { return a - b < -kTolerance; }
inline bool equals(double a, double b) { return std::abs(a - b) <= kTolerance; }

// One builtin object.  Hexahedron: the box [offset, offset + width].  Sphere: centre and radius.
// An inverted object is an obstacle: the domain lies outside it.
struct Object {
  enum class Kind { Hexahedron, Sphere };
  Kind kind = Kind::Hexahedron;
  bool inverted = false;
  Vec3 offset{0.0, 0.0, 0.0}, width{1.0, 1.0, 1.0};  // hexahedron
  Vec3 center{0.0, 0.0, 0.0};                          // sphere
  double radius = 0.0;

  static Object hexahedron(const Vec3& offset, const Vec3& width, bool inverted = false) {
    Object o;
    o.kind = Kind::Hexahedron;
    o.offset = offset;
    o.width = width;
    o.inverted = inverted;
    return o;
  }
  static Object sphere(const Vec3& center, double radius, bool inverted = false) {
    Object o;
    o.kind = Kind::Sphere;
    o.center = center;
    o.radius = radius;
    o.inverted = inverted;
    return o;
  }

  // AbstractGeometry: an inverted object swaps "completely inside" and "completely outside", and
  // a point is outside its closed domain when it lies in the object's open interior (a corner test
  // at zero resolution).
  bool isCompletelyOutside(const Vec3& x, const Vec3& res) const {
    return inverted ? insideNotInverted(x, res) : outsideNotInverted(x, res);
  }
  bool isCompletelyInside(const Vec3& x, const Vec3& res) const {
    return inverted ? outsideNotInverted(x, res) : insideNotInverted(x, res);
  }
  bool isOutsideClosedDomain(const Vec3& x) const {
    if (inverted) return insideNotInverted(x, Vec3{0.0, 0.0, 0.0});
    return kind == Kind::Hexahedron ? !insideClosedBox(x) : greater(distance(x), radius);
  }

 private:
  // ---- hexahedron
  bool insideOpenBox(const Vec3& x) const {
    bool r = true;
    for (int d = 0; d < 3; d++) {
      r = r && greater(x[d], offset[d]);
      r = r && smaller(x[d], offset[d] + width[d]);
    }
    return r;
  }
  bool insideClosedBox(const Vec3& x) const {
    bool r = true;
    for (int d = 0; d < 3; d++) {
      r = r && !smaller(x[d], offset[d]);
      r = r && !greater(x[d], offset[d] + width[d]);
    }
    return r;
  }
  // ---- sphere
  double distance(const Vec3& x) const {
    double s = 0.0;
    for (int d = 0; d < 3; d++) s += (x[d] - center[d]) * (x[d] - center[d]);
    return std::sqrt(s);
  }
  bool insideOpenSphere(const Vec3& x) const { return smaller(distance(x), radius); }
  bool insideClosedSphere(const Vec3& x) const { return !greater(distance(x), radius); }
  // the open circle of the sphere's radius around the axis through `base` along `dir`, between
  // base and base + height along it, contains the sphere's centre
  bool centreInOpenCylinder(int dir, const Vec3& base, double height) const {
    double s = 0.0;
    for (int d = 0; d < 3; d++)
      if (d != dir) s += (center[d] - base[d]) * (center[d] - base[d]);
    return smaller(std::sqrt(s), radius) && greater(center[dir], base[dir]) &&
           smaller(center[dir], base[dir] + height);
  }

  // the voxel x +- res lies completely inside the (not inverted) object
  bool insideNotInverted(const Vec3& x, const Vec3& res) const {
    if (kind == Kind::Hexahedron) {
      bool r = insideOpenBox(x);
      for (int c = 0; c < 8; c++) {
        Vec3 p0, p1;
        for (int d = 0; d < 3; d++) {
          const bool hi = (c >> d) & 1;
          p0[d] = hi ? x[d] + 0.5 * res[d] : x[d] - 0.5 * res[d];
          p1[d] = hi ? x[d] + 1.0 * res[d] : x[d] - 1.0 * res[d];
        }
        r = r && insideClosedBox(p0);
        r = r && insideClosedBox(p1);
      }
      return r;
    }
    bool r = insideOpenSphere(x);
    for (int c = 0; c < 8; c++) {
      Vec3 p = x;
      for (int d = 0; d < 3; d++) p[d] = ((c >> d) & 1) ? p[d] + res[d] : p[d] - res[d];
      r = r && insideClosedSphere(p);
    }
    return r;
  }

  // the voxel x +- res lies completely outside the (not inverted) object
  bool outsideNotInverted(const Vec3& x, const Vec3& res) const {
    if (kind == Kind::Hexahedron) {
      // per dimension: neither face of the voxel strictly inside the box's extent, and the voxel
      // does not span it; outside when that holds along some dimension
      bool r = false;
      for (int d = 0; d < 3; d++) {
        const double lo = x[d] - res[d], hi = x[d] + res[d], a = offset[d], b = offset[d] + width[d];
        bool dim = true;
        dim = dim && !(smaller(lo, b) && greater(lo, a));
        dim = dim && !(smaller(hi, b) && greater(hi, a));
        dim = dim && !(!greater(lo, a) && !smaller(hi, b));
        r = r || dim;
      }
      return r;
    }
    if (insideNotInverted(x, res)) return false;
    if (!greater(distance(x), radius)) return false;
    const bool point = equals(res[0], 0.0) && equals(res[1], 0.0) && equals(res[2], 0.0);
    if (point) return true;
    // a voxel corner inside the open sphere
    for (int c = 0; c < 8; c++) {
      Vec3 p = x;
      for (int d = 0; d < 3; d++) p[d] = ((c >> d) & 1) ? p[d] + res[d] : p[d] - res[d];
      if (insideOpenSphere(p)) return false;
    }
    // the centre within the voxel grown by the radius along one axis (a face of the voxel cut)
    for (int d = 0; d < 3; d++) {
      Vec3 m = res;
      m[d] += radius;
      bool in = true;
      for (int e = 0; e < 3; e++) {
        in = in && greater(center[e], x[e] - m[e]);
        in = in && smaller(center[e], x[e] + m[e]);
      }
      if (in) return false;
    }
    // the centre near one of the 12 voxel edges (an edge of the voxel cut): a cylinder of the
    // radius around each edge, the four edges along each axis
    static const int kEdges[3][4][3] = {
        {{-1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}, {-1, 1, 1}},
        {{-1, -1, -1}, {1, -1, -1}, {-1, -1, 1}, {1, -1, 1}},
        {{-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {1, 1, -1}}};
    for (int dir = 0; dir < 3; dir++) {
      const double height = 2.0 * res[dir];
      for (int k = 0; k < 4; k++) {
        Vec3 base;
        for (int d = 0; d < 3; d++) base[d] = x[d] + kEdges[dir][k][d] * res[d];
        if (centreInOpenCylinder(dir, base, height)) return false;
      }
    }
    return true;
  }
};

// The configured domain: the intersection of the objects (peano/geometry/Intersection.cpp).
struct Domain {
  std::vector<Object> objects;
  bool isCompletelyInside(const Vec3& x, const Vec3& res) const {
    bool r = true;
    for (const Object& o : objects) r = r && o.isCompletelyInside(x, res);
    return r;
  }
  bool isCompletelyOutside(const Vec3& x, const Vec3& res) const {
    bool r = false;
    for (const Object& o : objects) r = r || o.isCompletelyOutside(x, res);
    return r;
  }
  bool isOutsideClosedDomain(const Vec3& x) const {
    bool r = false;
    for (const Object& o : objects) r = r || o.isOutsideClosedDomain(x);
    return r;
  }
};

// The flags of one block-store cell (GridManagementService.cpp:86-118): inner unless the cell centre
// lies outside the closed domain (cells completely inside are inner); boundary unless the cell's
// voxel is completely inside or completely outside.
struct CellFlags {
  bool inner, boundary;
};
inline CellFlags cellFlags(const Domain& g, const Vec3& centre, const Vec3& meshSize) {
  if (g.isCompletelyInside(centre, meshSize)) return {true, false};
  return {!g.isOutsideClosedDomain(centre), !g.isCompletelyOutside(centre, meshSize)};
}

// The regular grid the block store sits on (BLB/runners/BlockLatticeBoltzmannBatchJobRunnerForRegularGrid.cpp:
// 70-92, peano/kernel/regulargrid/Grid.cpp:3-13, RegularGridVertexEnumerator.cpp:89-101): vertices
// (one 6^3 block each) cover the lattice from the domain offset, blocks(d) = n/6 + 1 (+1 if n%6);
// the Peano cell size is (dx (blocks - 1) 6) / (blocks - 1), vertex (b0, b1, b2) sits at
// offset + b h, and block slot s holds the cell centred at x - h/2 + meshSize/2 + s meshSize with
// meshSize = h/6.
struct BlockGrid {
  int n[3];
  int blocks[3];
  double dx;
  Vec3 offset;
  Vec3 h;

  BlockGrid(const int cells[3], double dx_, const Vec3& offset_) : dx(dx_), offset(offset_) {
    for (int d = 0; d < 3; d++) {
      n[d] = cells[d];
      blocks[d] = n[d] / 6 + 1 + (n[d] % 6 ? 1 : 0);
      const double size = dx * (blocks[d] - 1) * 6;
      h[d] = size / (double)(blocks[d] - 1);
    }
  }
  Vec3 vertexPosition(int b0, int b1, int b2) const {
    const int b[3] = {b0, b1, b2};
    Vec3 x = offset;
    for (int d = 0; d < 3; d++) x[d] += (double)b[d] * h[d];
    return x;
  }
  Vec3 meshSize() const {
    const double m = (1.0 / 6.0) * h[0];
    return Vec3{m, m, m};
  }
  // centre of slot s of the block at vertex position x
  Vec3 cellCentre(const Vec3& x, const int s[3]) const {
    const Vec3 m = meshSize();
    Vec3 p;
    for (int d = 0; d < 3; d++) p[d] = x[d] - 0.5 * h[d] + 0.5 * m[d] + s[d] * m[d];
    return p;
  }
  // flags of every block-store slot, [6 blocks(2)][6 blocks(1)][6 blocks(0)], x fastest
  void storeFlags(const Domain& g, std::vector<uint8_t>& inner, std::vector<uint8_t>& boundary) const {
    const int64_t sx = 6 * blocks[0], sy = 6 * blocks[1], sz = 6 * blocks[2];
    inner.assign((size_t)(sx * sy * sz), 0);
    boundary.assign(inner.size(), 0);
    const Vec3 m = meshSize();
    for (int bz = 0; bz < blocks[2]; bz++)
      for (int by = 0; by < blocks[1]; by++)
        for (int bx = 0; bx < blocks[0]; bx++) {
          const Vec3 x = vertexPosition(bx, by, bz);
          for (int z = 0; z < 6; z++)
            for (int y = 0; y < 6; y++)
              for (int xs = 0; xs < 6; xs++) {
                const int s[3] = {xs, y, z};
                const CellFlags f = cellFlags(g, cellCentre(x, s), m);
                const int64_t k = (6 * bx + xs) + sx * ((6 * by + y) + sy * (int64_t)(6 * bz + z));
                inner[(size_t)k] = f.inner ? 1 : 0;
                boundary[(size_t)k] = f.boundary ? 1 : 0;
              }
        }
  }
  // inner flags of the lattice cells (row-major x + n0 (y + n1 z)) from the block store's (storeFlags);
  // lattice cell c is block-store slot c + 3.  Throws if a slot outside the lattice is inner (the
  // domain must fit the lattice).
  std::vector<uint8_t> latticeInner(const std::vector<uint8_t>& in) const {
    const int64_t sx = 6 * blocks[0], sy = 6 * blocks[1];
    std::vector<uint8_t> out((size_t)n[0] * n[1] * n[2], 0);
    int64_t total = 0, kept = 0;
    for (size_t k = 0; k < in.size(); k++) total += in[k];
    for (int z = 0; z < n[2]; z++)
      for (int y = 0; y < n[1]; y++)
        for (int x = 0; x < n[0]; x++) {
          const uint8_t v = in[(size_t)((x + 3) + sx * ((y + 3) + sy * (int64_t)(z + 3)))];
          out[(size_t)(x + (int64_t)n[0] * (y + (int64_t)n[1] * z))] = v;
          kept += v;
        }
    if (kept != total) throw std::runtime_error("the configured geometry has inner cells outside the lattice");
    return out;
  }
};

}  // namespace geometry
}  // namespace madpac_amd
