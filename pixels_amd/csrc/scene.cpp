// scene.cpp - seeded scene builder: SDF primitives rasterised into the sparse
// two-level brick map (grid / pool / super grid) of the reference.
//
// Encoding: src/world/VoxelWorld.ts:4-13 (grid entry 0 = air, UNIFORM|v =
// uniform brick, p+1 = pool brick p of 512 u16 voxels, local index
// x|y<<3|z<<6).  Rasteriser: brick -> 4^3 sub-block -> voxel, exactly the
// tests of apply() in src/world/shapes.ts:334-414 (brick-centre SDF against
// the brick half-diagonal, then sub-block centres, then voxel centres), with
// compactCell semantics (VoxelWorld.ts:147-156) so every pool brick holds at
// least two distinct values.  SDFs follow src/world/shapes.ts:36-143.
// Super-grid counts follow GpuWorld.ts:90-100.
//
// Unlike the reference (one primitive after another over one world), bricks
// are rasterised in parallel: every brick replays, in primitive order, the
// primitives whose brick test keeps it, which gives the same result because
// apply() touches each brick independently.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr uint32_t kUniform = 0x80000000u;

struct Rng {  // splitmix64
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double uni(double a, double b) { return a + (b - a) * uni(); }
  double loguni(double a, double b) { return std::exp(uni(std::log(a), std::log(b))); }
  uint32_t below(uint32_t n) { return (uint32_t)(uni() * n); }
};

enum PrimType { kBox = 0, kRoundBox = 1, kSphere = 2, kCylinder = 3, kCapsule = 4, kTorus = 5, kAabb = 6 };

// One primitive in voxel units.  p[] meaning per type:
//  box/roundbox: c[3] h[3] r yaw          (r = corner radius)
//  sphere:       c[3] r
//  cylinder:     c[3] r halfHeight axis yaw
//  capsule:      a[3] b[3] r
//  torus:        c[3] R r axis yaw
//  aabb:         lo[3] hi[3]              (axis-aligned slab/box with the exact contains test)
struct Prim {
  int type;
  uint16_t mat;
  double p[12];
  double lo[3], hi[3];  // AABB (voxel units)
  double cy, sy;        // cos/sin of yaw
};

inline double len2(double a, double b) { return std::sqrt(a * a + b * b); }
inline double len3(double a, double b, double c) { return std::sqrt(a * a + b * b + c * c); }

// world point -> local frame of a yawed primitive centred at c
inline void to_local(const Prim& P, double x, double y, double z, double& lx, double& ly, double& lz) {
  double dx = x - P.p[0], dy = y - P.p[1], dz = z - P.p[2];
  lx = P.cy * dx + P.sy * dz;
  ly = dy;
  lz = -P.sy * dx + P.cy * dz;
}

inline void axis_split(int axis, double lx, double ly, double lz, double& a, double& b, double& h) {
  if (axis == 1) { a = lx; b = lz; h = ly; }
  else if (axis == 0) { a = ly; b = lz; h = lx; }
  else { a = lx; b = ly; h = lz; }
}

double sdf(const Prim& P, double x, double y, double z) {
  switch (P.type) {
    case kBox: {
      double lx, ly, lz;
      to_local(P, x, y, z, lx, ly, lz);
      double qx = std::fabs(lx) - P.p[3], qy = std::fabs(ly) - P.p[4], qz = std::fabs(lz) - P.p[5];
      double out = len3(std::max(qx, 0.0), std::max(qy, 0.0), std::max(qz, 0.0));
      return out + std::min(std::max(qx, std::max(qy, qz)), 0.0);
    }
    case kRoundBox: {
      double lx, ly, lz;
      to_local(P, x, y, z, lx, ly, lz);
      double r = P.p[6];
      double qx = std::fabs(lx) - (P.p[3] - r), qy = std::fabs(ly) - (P.p[4] - r), qz = std::fabs(lz) - (P.p[5] - r);
      return len3(std::max(qx, 0.0), std::max(qy, 0.0), std::max(qz, 0.0)) +
             std::min(std::max(qx, std::max(qy, qz)), 0.0) - r;
    }
    case kSphere:
      return len3(x - P.p[0], y - P.p[1], z - P.p[2]) - P.p[3];
    case kCylinder: {
      double lx, ly, lz, a, b, h;
      to_local(P, x, y, z, lx, ly, lz);
      axis_split((int)P.p[5], lx, ly, lz, a, b, h);
      double dr = len2(a, b) - P.p[3], dh = std::fabs(h) - P.p[4];
      return std::min(std::max(dr, dh), 0.0) + len2(std::max(dr, 0.0), std::max(dh, 0.0));
    }
    case kCapsule: {
      double ax = P.p[0], ay = P.p[1], az = P.p[2];
      double bax = P.p[3] - ax, bay = P.p[4] - ay, baz = P.p[5] - az;
      double bb = bax * bax + bay * bay + baz * baz;
      if (bb == 0.0) bb = 1e-12;
      double px = x - ax, py = y - ay, pz = z - az;
      double h = std::min(1.0, std::max(0.0, (px * bax + py * bay + pz * baz) / bb));
      return len3(px - bax * h, py - bay * h, pz - baz * h) - P.p[6];
    }
    case kTorus: {
      double lx, ly, lz, a, b, h;
      to_local(P, x, y, z, lx, ly, lz);
      axis_split((int)P.p[5], lx, ly, lz, a, b, h);
      return len2(len2(a, b) - P.p[3], h) - P.p[4];
    }
    case kAabb: {
      double cx = 0.5 * (P.p[0] + P.p[3]), cyy = 0.5 * (P.p[1] + P.p[4]), cz = 0.5 * (P.p[2] + P.p[5]);
      double hx = 0.5 * (P.p[3] - P.p[0]), hy = 0.5 * (P.p[4] - P.p[1]), hz = 0.5 * (P.p[5] - P.p[2]);
      double qx = std::fabs(x - cx) - hx, qy = std::fabs(y - cyy) - hy, qz = std::fabs(z - cz) - hz;
      double out = len3(std::max(qx, 0.0), std::max(qy, 0.0), std::max(qz, 0.0));
      return out + std::min(std::max(qx, std::max(qy, qz)), 0.0);
    }
  }
  return 1e30;
}

// shapes.ts:47-48: exact containment test of an axis-aligned box
bool contains(const Prim& P, double x0, double y0, double z0, double x1, double y1, double z1) {
  if (P.type == kAabb)
    return x0 >= P.p[0] && x1 <= P.p[3] && y0 >= P.p[1] && y1 <= P.p[4] && z0 >= P.p[2] && z1 <= P.p[5];
  if (P.type == kBox && P.sy == 0.0 && P.cy == 1.0)
    return x0 >= P.p[0] - P.p[3] && x1 <= P.p[0] + P.p[3] && y0 >= P.p[1] - P.p[4] && y1 <= P.p[1] + P.p[4] &&
           z0 >= P.p[2] - P.p[5] && z1 <= P.p[2] + P.p[5];
  return false;
}
bool has_contains(const Prim& P) { return P.type == kAabb || (P.type == kBox && P.sy == 0.0 && P.cy == 1.0); }

void finish_prim(Prim& P, double yaw) {
  P.cy = std::cos(yaw);
  P.sy = std::sin(yaw);
  if (yaw == 0.0) { P.cy = 1.0; P.sy = 0.0; }
  double ex[3] = {0, 0, 0};
  switch (P.type) {
    case kBox:
    case kRoundBox: ex[0] = P.p[3]; ex[1] = P.p[4]; ex[2] = P.p[5]; break;
    case kSphere: ex[0] = ex[1] = ex[2] = P.p[3]; break;
    case kCylinder: {
      int ax = (int)P.p[5];
      ex[0] = ex[1] = ex[2] = P.p[3];
      ex[ax] = P.p[4];
      break;
    }
    case kTorus: {
      int ax = (int)P.p[5];
      double e = P.p[3] + P.p[4];
      ex[0] = ex[1] = ex[2] = e;
      ex[ax] = P.p[4];
      break;
    }
    default: break;
  }
  if (P.type == kCapsule) {
    for (int a = 0; a < 3; a++) {
      P.lo[a] = std::min(P.p[a], P.p[3 + a]) - P.p[6];
      P.hi[a] = std::max(P.p[a], P.p[3 + a]) + P.p[6];
    }
  } else if (P.type == kAabb) {
    for (int a = 0; a < 3; a++) { P.lo[a] = P.p[a]; P.hi[a] = P.p[3 + a]; }
  } else {
    // yaw about y: rotate the local AABB
    double c = std::fabs(P.cy), s = std::fabs(P.sy);
    double hx = c * ex[0] + s * ex[2], hz = s * ex[0] + c * ex[2];
    P.lo[0] = P.p[0] - hx; P.hi[0] = P.p[0] + hx;
    P.lo[1] = P.p[1] - ex[1]; P.hi[1] = P.p[1] + ex[1];
    P.lo[2] = P.p[2] - hz; P.hi[2] = P.p[2] + hz;
  }
}

struct Scene {
  int32_t dims[3];
  int32_t bd[3];
  int32_t sd[3];
  std::vector<uint32_t> grid;
  std::vector<uint16_t> pool;
  std::vector<uint32_t> super_grid;
  std::vector<float> mats;
  std::vector<Prim> prims;
  uint64_t solid = 0, glass = 0;
  uint32_t pool_bricks = 0, uniform_bricks = 0;
  uint64_t seed = 0;
};

int n_threads() {
  const char* e = std::getenv("PX_SCENE_THREADS");
  int n = e ? std::atoi(e) : 0;
  if (n <= 0) n = (int)std::thread::hardware_concurrency();
  if (n <= 0) n = 4;
  return std::min(n, 64);
}

template <class F>
void parallel_for(size_t n, F f) {
  int nt = n_threads();
  std::atomic<size_t> next{0};
  const size_t chunk = 64;
  auto worker = [&]() {
    for (;;) {
      size_t b = next.fetch_add(chunk);
      if (b >= n) break;
      size_t e = std::min(n, b + chunk);
      for (size_t i = b; i < e; i++) f(i);
    }
  };
  std::vector<std::thread> th;
  for (int i = 1; i < nt; i++) th.emplace_back(worker);
  worker();
  for (auto& t : th) t.join();
}

// apply() of shapes.ts:334-414 for one primitive over one brick, Op.Add with
// a constant paint. state: 0 = empty, 1 = uniform(val), 2 = explicit voxels.
void apply_brick(const Prim& P, int32_t bx, int32_t by, int32_t bz, int& state, uint16_t& val, uint16_t* vox) {
  const double brickHalfDiag = 4.0 * std::sqrt(3.0) + 0.01;
  const double subHalfDiag = 2.0 * std::sqrt(3.0) + 0.01;
  int32_t ox = bx * 8, oy = by * 8, oz = bz * 8;
  bool full = false;
  if (has_contains(P)) full = contains(P, ox + 0.5, oy + 0.5, oz + 0.5, ox + 7.5, oy + 7.5, oz + 7.5);
  if (!full) {
    double d = sdf(P, ox + 4.0, oy + 4.0, oz + 4.0);
    if (d > brickHalfDiag) return;
    full = d < -brickHalfDiag;
  }
  uint16_t pv = P.mat;
  if (full) {  // setBrickUniform
    state = pv == 0 ? 0 : 1;
    val = pv;
    return;
  }
  uint16_t scratch[512];
  if (state == 0) std::memset(scratch, 0, sizeof scratch);
  else if (state == 1) std::fill(scratch, scratch + 512, val);
  else std::memcpy(scratch, vox, sizeof scratch);
  bool changed = false;
  for (int s = 0; s < 8; s++) {
    int32_t sx = ox + (s & 1) * 4, sy = oy + ((s >> 1) & 1) * 4, sz = oz + ((s >> 2) & 1) * 4;
    double d = sdf(P, sx + 2.0, sy + 2.0, sz + 2.0);
    if (d > subHalfDiag) continue;
    bool subFull = d < -subHalfDiag;
    for (int32_t z = sz; z < sz + 4; z++)
      for (int32_t y = sy; y < sy + 4; y++)
        for (int32_t x = sx; x < sx + 4; x++) {
          if (!subFull && sdf(P, x + 0.5, y + 0.5, z + 0.5) > 0) continue;
          int li = (x & 7) | ((y & 7) << 3) | ((z & 7) << 6);
          if (scratch[li] != pv) {
            scratch[li] = pv;
            changed = true;
          }
        }
  }
  if (!changed) return;
  // editableBrick + write + compactCell
  uint16_t first = scratch[0];
  bool same = true;
  for (int i = 1; i < 512; i++)
    if (scratch[i] != first) { same = false; break; }
  if (same) {
    state = first == 0 ? 0 : 1;
    val = first;
  } else {
    state = 2;
    std::memcpy(vox, scratch, sizeof scratch);
  }
}

void rasterise(Scene& S) {
  const int32_t BX = S.bd[0], BY = S.bd[1], BZ = S.bd[2];
  const size_t ncell = (size_t)BX * BY * BZ;
  const double brickHalfDiag = 4.0 * std::sqrt(3.0) + 0.01;
  // candidate bricks per primitive (brick-centre test), in parallel
  std::vector<std::vector<uint32_t>> cand(S.prims.size());
  parallel_for(S.prims.size(), [&](size_t i) {
    const Prim& P = S.prims[i];
    int32_t v0[3], v1[3];
    for (int a = 0; a < 3; a++) {
      v0[a] = std::max(0, (int32_t)std::floor(P.lo[a] - 0.5));
      v1[a] = std::min(S.dims[a] - 1, (int32_t)std::ceil(P.hi[a] + 0.5));
    }
    if (v0[0] > v1[0] || v0[1] > v1[1] || v0[2] > v1[2]) return;
    auto& out = cand[i];
    for (int32_t bz = v0[2] >> 3; bz <= v1[2] >> 3; bz++)
      for (int32_t by = v0[1] >> 3; by <= v1[1] >> 3; by++)
        for (int32_t bx = v0[0] >> 3; bx <= v1[0] >> 3; bx++) {
          if (!has_contains(P) && sdf(P, bx * 8 + 4.0, by * 8 + 4.0, bz * 8 + 4.0) > brickHalfDiag) continue;
          out.push_back((uint32_t)(bx + BX * (by + BY * bz)));
        }
  });
  // CSR by brick, primitives kept in order
  std::vector<uint32_t> start(ncell + 1, 0);
  for (auto& c : cand)
    for (uint32_t ci : c) start[ci + 1]++;
  for (size_t i = 0; i < ncell; i++) start[i + 1] += start[i];
  std::vector<uint32_t> fill(start.begin(), start.end() - 1);
  std::vector<uint32_t> list(start[ncell]);
  for (size_t p = 0; p < cand.size(); p++)
    for (uint32_t ci : cand[p]) list[fill[ci]++] = (uint32_t)p;
  cand.clear();
  cand.shrink_to_fit();

  std::vector<uint8_t> st(ncell, 0);
  std::vector<uint16_t> uv(ncell, 0);
  // explicit bricks are produced into per-cell slots lazily
  std::vector<std::vector<uint16_t>> buf(ncell);
  parallel_for(ncell, [&](size_t ci) {
    if (start[ci] == start[ci + 1]) return;
    int32_t bx = (int32_t)(ci % BX), by = (int32_t)((ci / BX) % BY), bz = (int32_t)(ci / ((size_t)BX * BY));
    int state = 0;
    uint16_t val = 0;
    uint16_t vox[512];
    for (uint32_t k = start[ci]; k < start[ci + 1]; k++) apply_brick(S.prims[list[k]], bx, by, bz, state, val, vox);
    st[ci] = (uint8_t)state;
    uv[ci] = val;
    if (state == 2) buf[ci].assign(vox, vox + 512);
  });
  // pool indices in cell order
  S.grid.assign(ncell, 0u);
  uint32_t np = 0;
  for (size_t ci = 0; ci < ncell; ci++) {
    if (st[ci] == 1) { S.grid[ci] = kUniform | uv[ci]; S.uniform_bricks++; }
    else if (st[ci] == 2) S.grid[ci] = ++np;
  }
  S.pool_bricks = np;
  S.pool.assign((size_t)np * 512, 0);
  parallel_for(ncell, [&](size_t ci) {
    if (st[ci] == 2) std::memcpy(&S.pool[(size_t)(S.grid[ci] - 1) * 512], buf[ci].data(), 1024);
  });
}

void finish_scene(Scene& S) {
  const int32_t BX = S.bd[0], BY = S.bd[1], BZ = S.bd[2];
  S.super_grid.assign((size_t)S.sd[0] * S.sd[1] * S.sd[2], 0u);
  uint64_t solid = 0, glass = 0;
  for (int32_t z = 0; z < BZ; z++)
    for (int32_t y = 0; y < BY; y++)
      for (int32_t x = 0; x < BX; x++) {
        uint32_t g = S.grid[x + (size_t)BX * (y + (size_t)BY * z)];
        if (g == 0) continue;
        S.super_grid[(x >> 2) + (size_t)S.sd[0] * ((y >> 2) + (size_t)S.sd[1] * (z >> 2))]++;
        if (g & kUniform) {
          solid += 512;
          if (S.mats[(g & 0xffu) * 20 + 13] > 0.0f) glass += 512;
        } else {
          const uint16_t* b = &S.pool[(size_t)(g - 1) * 512];
          for (int i = 0; i < 512; i++)
            if (b[i]) {
              solid++;
              if (S.mats[(b[i] & 0xffu) * 20 + 13] > 0.0f) glass++;
            }
        }
      }
  S.solid = solid;
  S.glass = glass;
}

// ---- material table (packMaterials layout, src/world/materials.ts:193-218)
struct MatDef {
  uint32_t color, color2;
  float pattern, scale, noise, emission, spec, shininess, reflect, glass, liquid, price;
  float softness, fragile, structural, debris;
};

float srgb_to_linear(int v) {
  double s = v / 255.0;
  return (float)(s <= 0.04045 ? s / 12.92 : std::pow((s + 0.055) / 1.055, 2.4));
}

void pack_materials(std::vector<float>& out) {
  // 16 entries drawn from the reference palette (src/world/materials.ts:115-189);
  // id 1 is the structural slab material, id 9 the window pane.
  static const MatDef defs[16] = {
      /*1 concrete   */ {0x8a8580, 0x8a8580, 0, 0.1f, 0.12f, 0, 0.1f, 16, 0, 0, 0, 0.2f, 0.6f, 0, 1, 0.6f},
      /*2 marble     */ {0xf2eee8, 0x7d7468, 1, 0.35f, 0.03f, 0, 0.5f, 60, 0.12f, 0, 0, 4, 0.6f, 0, 0, 0.6f},
      /*3 gold       */ {0xe7b549, 0xa47722, 10, 0.05f, 0.05f, 0, 0.9f, 90, 0.25f, 0, 0, 25, 0.4f, 0, 0, 0.6f},
      /*4 chrome     */ {0xc8ccd2, 0x8e949c, 10, 0.05f, 0.03f, 0, 1, 120, 0.45f, 0, 0, 5, 0.35f, 0, 0, 0.6f},
      /*5 walnut     */ {0x5a3720, 0x3b2213, 2, 0.03f, 0.06f, 0, 0.3f, 30, 0, 0, 0, 4, 1, 0, 0, 0.6f},
      /*6 velvet     */ {0x4a1a5c, 0x2e0f3a, 8, 0.02f, 0.08f, 0, 0.05f, 16, 0, 0, 0, 6, 1.2f, 0, 0, 0.6f},
      /*7 foam       */ {0xe6d9a6, 0xe6d9a6, 0, 0.1f, 0.15f, 0, 0, 16, 0, 0, 0, 0.5f, 1.4f, 0, 0, 0.6f},
      /*8 steelDark  */ {0x3b3d42, 0x3b3d42, 0, 0.1f, 0.05f, 0, 0.6f, 50, 0, 0, 0, 3, 0.4f, 0, 0, 0.6f},
      /*9 windowGlass*/ {0xc6dde8, 0xc6dde8, 0, 0.1f, 0.08f, 0, 1, 200, 0, 0.82f, 0, 1.5f, 1, 2, 0, 0.5f},
      /*10 glassClear*/ {0xe8f2f4, 0xe8f2f4, 0, 0.1f, 0.08f, 0, 1, 250, 0, 0.9f, 0, 12, 1, 1, 0, 1},
      /*11 glassGreen*/ {0x2f7a3a, 0x2f7a3a, 0, 0.1f, 0.08f, 0, 1, 200, 0, 0.7f, 0, 15, 1, 1, 0, 1},
      /*12 glassTable*/ {0xa6d4cc, 0xa6d4cc, 0, 0.1f, 0.08f, 0, 1, 250, 0, 0.75f, 0, 10, 1, 2, 0, 0.8f},
      /*13 leather   */ {0x1a1414, 0x2a201c, 9, 0.01f, 0.06f, 0, 0.35f, 25, 0, 0, 0, 8, 1, 0, 0, 0.6f},
      /*14 porcelain */ {0xf7f6f2, 0x1f3f8f, 0, 0.1f, 0.02f, 0, 0.7f, 90, 0.08f, 0, 0, 80, 1, 1, 0, 1},
      /*15 lacquer   */ {0x08080a, 0x08080a, 0, 0.1f, 0.01f, 0, 1, 200, 0.35f, 0, 0, 60, 0.8f, 0, 0, 0.6f},
      /*16 bulb      */ {0xffd9a0, 0xffd9a0, 0, 0.1f, 0, 8, 0, 16, 0, 0, 0, 5, 1, 1, 0, 1},
  };
  out.assign(256 * 20, 0.0f);
  for (int i = 0; i < 16; i++) {
    const MatDef& m = defs[i];
    int id = i + 1;
    float* o = &out[id * 20];
    float a[3] = {srgb_to_linear((m.color >> 16) & 255), srgb_to_linear((m.color >> 8) & 255), srgb_to_linear(m.color & 255)};
    float b[3] = {srgb_to_linear((m.color2 >> 16) & 255), srgb_to_linear((m.color2 >> 8) & 255), srgb_to_linear(m.color2 & 255)};
    float v[20] = {a[0], a[1], a[2], m.emission, b[0], b[1], b[2], m.pattern, m.scale, m.noise, m.spec, m.shininess,
                   m.reflect, m.glass, m.liquid, m.price, m.softness, m.fragile, m.structural, m.debris};
    std::memcpy(o, v, sizeof v);
    // rubble twin: same look, not structural, already paid for (materials.ts:210-215)
    float* r = &out[(id | 128) * 20];
    std::memcpy(r, v, sizeof v);
    r[15] = 0.0f;
    r[17] = std::min(m.fragile, 1.0f);
    r[18] = 0.0f;
  }
}

void init_dims(Scene& S, int32_t dx, int32_t dy, int32_t dz) {
  S.dims[0] = dx; S.dims[1] = dy; S.dims[2] = dz;
  for (int a = 0; a < 3; a++) {
    S.bd[a] = S.dims[a] >> 3;
    S.sd[a] = (S.bd[a] + 3) >> 2;
  }
}

// Small scene: 200 boxes/spheres, centres uniform in the grid, half-extents
// (radius) uniform 4-64 voxels, material uniform in 1..16.
void gen_small(Scene& S, Rng& R, int n_prims, bool floor_slab) {
  for (int i = 0; i < n_prims; i++) {
    Prim P{};
    bool box = R.uni() < 0.5;
    P.type = box ? kBox : kSphere;
    for (int a = 0; a < 3; a++) P.p[a] = R.uni(0.0, (double)S.dims[a]);
    if (box)
      for (int a = 0; a < 3; a++) P.p[3 + a] = R.uni(4.0, 64.0);
    else
      P.p[3] = R.uni(4.0, 64.0);
    P.mat = (uint16_t)(1 + R.below(16));
    finish_prim(P, 0.0);
    S.prims.push_back(P);
  }
  if (floor_slab) {
    Prim P{};
    P.type = kAabb;
    P.mat = 1;
    P.p[0] = 0; P.p[1] = 0; P.p[2] = 0;
    P.p[3] = S.dims[0]; P.p[4] = 16; P.p[5] = S.dims[2];
    finish_prim(P, 0.0);
    S.prims.push_back(P);
  }
}

// Full-scale scene: structural floor/ceiling slabs 16 voxels thick and 2000
// primitives (box, rounded box, cylinder, capsule, torus) with a major size
// log-uniform in [2, 400] voxels, a thickness log-uniform in [2, min(24, size)],
// random yaw; 10% glass materials.
void gen_large(Scene& S, Rng& R, int n_prims, double scale) {
  static const uint16_t opaque[] = {2, 3, 4, 5, 6, 7, 8, 13, 14, 15, 16};
  static const uint16_t glass[] = {9, 10, 11, 12};
  const double slab = 16.0 * scale;
  for (int i = 0; i < n_prims; i++) {
    Prim P{};
    P.type = (int)R.below(5);
    if (P.type == 2) P.type = kCylinder;      // 0 box 1 roundbox 3 cylinder
    else if (P.type == 3) P.type = kCapsule;  // 4 capsule
    else if (P.type == 4) P.type = kTorus;    // 5 torus
    double s = R.loguni(2.0, 400.0) * scale;
    double th = R.loguni(2.0, std::max(2.0, std::min(24.0, s / scale))) * scale;
    double c[3] = {R.uni(0.0, S.dims[0]), R.uni(slab, S.dims[1] - slab), R.uni(0.0, S.dims[2])};
    double yaw = R.uni(0.0, 2.0 * M_PI);
    P.mat = R.uni() < 0.1 ? glass[R.below(4)] : opaque[R.below(11)];
    for (int a = 0; a < 3; a++) P.p[a] = c[a];
    switch (P.type) {
      case kBox:
      case kRoundBox: {
        double h[3] = {0.5 * s, 0.5 * s * R.uni(0.25, 1.0), 0.5 * th};
        int thin = (int)R.below(3);  // which local axis is the thin one
        double hh[3];
        hh[thin] = h[2];
        hh[(thin + 1) % 3] = h[0];
        hh[(thin + 2) % 3] = h[1];
        for (int a = 0; a < 3; a++) P.p[3 + a] = std::max(hh[a], 0.5);
        if (P.type == kRoundBox) P.p[6] = 0.5 * std::min(P.p[3], std::min(P.p[4], P.p[5]));
        finish_prim(P, yaw);
        break;
      }
      case kCylinder:
        P.p[3] = std::max(0.5 * th, 0.6);
        P.p[4] = std::max(0.5 * s, 0.5);
        P.p[5] = (double)R.below(3);
        finish_prim(P, yaw);
        break;
      case kCapsule: {
        double pitch = R.uni(-0.5 * M_PI, 0.5 * M_PI);
        double d[3] = {std::cos(pitch) * std::cos(yaw), std::sin(pitch), std::cos(pitch) * std::sin(yaw)};
        for (int a = 0; a < 3; a++) {
          P.p[a] = c[a] - 0.5 * s * d[a];
          P.p[3 + a] = c[a] + 0.5 * s * d[a];
        }
        P.p[6] = std::max(0.5 * th, 0.9);
        finish_prim(P, 0.0);
        break;
      }
      case kTorus:
        P.p[3] = 0.5 * s;
        P.p[4] = std::max(0.5 * th, 0.9);
        P.p[5] = (double)R.below(3);
        finish_prim(P, yaw);
        break;
    }
    S.prims.push_back(P);
  }
  for (int k = 0; k < 2; k++) {
    Prim P{};
    P.type = kAabb;
    P.mat = 1;
    P.p[0] = 0; P.p[2] = 0; P.p[3] = S.dims[0]; P.p[5] = S.dims[2];
    P.p[1] = k == 0 ? 0.0 : S.dims[1] - slab;
    P.p[4] = k == 0 ? slab : S.dims[1];
    finish_prim(P, 0.0);
    S.prims.push_back(P);
  }
}

}  // namespace

extern "C" {

typedef struct pxs_scene pxs_scene;

// kind 0: small scene (512x256x512 by default), kind 1: small scene + 16-voxel
// structural floor slab, kind 2: full-scale scene (1280x352x960 by default),
// kind 3: full-scale generator scaled by `scale` (dims given by the caller).
pxs_scene* pxs_build(int kind, uint64_t seed, int32_t dx, int32_t dy, int32_t dz, int n_prims, double scale) {
  Scene* S = new Scene();
  S->seed = seed;
  if (kind == 0 || kind == 1) {
    init_dims(*S, dx > 0 ? dx : 512, dy > 0 ? dy : 256, dz > 0 ? dz : 512);
  } else {
    init_dims(*S, dx > 0 ? dx : 1280, dy > 0 ? dy : 352, dz > 0 ? dz : 960);
  }
  if ((S->dims[0] & 7) || (S->dims[1] & 7) || (S->dims[2] & 7) || S->dims[0] <= 0 || S->dims[1] <= 0 ||
      S->dims[2] <= 0) {
    delete S;
    return nullptr;
  }
  pack_materials(S->mats);
  Rng R(seed);
  if (kind == 0 || kind == 1) gen_small(*S, R, n_prims > 0 ? n_prims : 200, kind == 1);
  else gen_large(*S, R, n_prims > 0 ? n_prims : 2000, scale > 0 ? scale : 1.0);
  rasterise(*S);
  finish_scene(*S);
  return reinterpret_cast<pxs_scene*>(S);
}

// Primitives of the scene in the px_prim layout of include/pixels.h (type,
// mat, p[12], cy, sy, lo[3], hi[3]); returns the count.  out may be NULL.
uint32_t pxs_export_prims(const pxs_scene* s, void* out) {
  const Scene* S = reinterpret_cast<const Scene*>(s);
  if (out) {
    unsigned char* o = static_cast<unsigned char*>(out);
    for (const Prim& P : S->prims) {
      uint32_t t = (uint32_t)P.type, m = P.mat;
      std::memcpy(o, &t, 4);
      std::memcpy(o + 4, &m, 4);
      std::memcpy(o + 8, P.p, 96);
      std::memcpy(o + 104, &P.cy, 8);
      std::memcpy(o + 112, &P.sy, 8);
      std::memcpy(o + 120, P.lo, 24);
      std::memcpy(o + 144, P.hi, 24);
      o += 168;
    }
  }
  return (uint32_t)S->prims.size();
}

// Encode a dense voxel array (x fastest) into the brick map with canonical
// compaction; mats may be NULL for the built-in table.
pxs_scene* pxs_encode_dense(int32_t dx, int32_t dy, int32_t dz, const uint16_t* vox, const float* mats) {
  if ((dx & 7) || (dy & 7) || (dz & 7) || dx <= 0 || dy <= 0 || dz <= 0 || !vox) return nullptr;
  Scene* S = new Scene();
  init_dims(*S, dx, dy, dz);
  if (mats) S->mats.assign(mats, mats + 256 * 20);
  else pack_materials(S->mats);
  const int32_t BX = S->bd[0], BY = S->bd[1], BZ = S->bd[2];
  size_t ncell = (size_t)BX * BY * BZ;
  S->grid.assign(ncell, 0u);
  uint32_t np = 0;
  std::vector<uint16_t> b(512);
  for (size_t ci = 0; ci < ncell; ci++) {
    int32_t bx = (int32_t)(ci % BX), by = (int32_t)((ci / BX) % BY), bz = (int32_t)(ci / ((size_t)BX * BY));
    for (int l = 0; l < 512; l++) {
      int32_t x = bx * 8 + (l & 7), y = by * 8 + ((l >> 3) & 7), z = bz * 8 + (l >> 6);
      b[l] = vox[(size_t)x + (size_t)dx * ((size_t)y + (size_t)dy * z)];
    }
    bool same = true;
    for (int l = 1; l < 512; l++)
      if (b[l] != b[0]) { same = false; break; }
    if (same) {
      if (b[0]) { S->grid[ci] = kUniform | b[0]; S->uniform_bricks++; }
    } else {
      S->grid[ci] = ++np;
      S->pool.insert(S->pool.end(), b.begin(), b.end());
    }
  }
  S->pool_bricks = np;
  finish_scene(*S);
  return reinterpret_cast<pxs_scene*>(S);
}

void pxs_free(pxs_scene* s) { delete reinterpret_cast<Scene*>(s); }

// info[0..2] dims, [3..5] bricks, [6..8] supers, [9] pool bricks, [10] uniform
// bricks, [11] primitives; solid/glass voxel counts in counts[0..1].
void pxs_info(const pxs_scene* s, int64_t* info, uint64_t* counts) {
  const Scene* S = reinterpret_cast<const Scene*>(s);
  for (int a = 0; a < 3; a++) {
    info[a] = S->dims[a];
    info[3 + a] = S->bd[a];
    info[6 + a] = S->sd[a];
  }
  info[9] = S->pool_bricks;
  info[10] = S->uniform_bricks;
  info[11] = (int64_t)S->prims.size();
  counts[0] = S->solid;
  counts[1] = S->glass;
}

const uint32_t* pxs_grid(const pxs_scene* s) { return reinterpret_cast<const Scene*>(s)->grid.data(); }
const uint16_t* pxs_pool(const pxs_scene* s) { return reinterpret_cast<const Scene*>(s)->pool.data(); }
const uint32_t* pxs_super(const pxs_scene* s) { return reinterpret_cast<const Scene*>(s)->super_grid.data(); }
const float* pxs_mats(const pxs_scene* s) { return reinterpret_cast<const Scene*>(s)->mats.data(); }

// Decode the whole world to a dense array (x fastest).
void pxs_decode(const pxs_scene* s, uint16_t* out) {
  const Scene* S = reinterpret_cast<const Scene*>(s);
  const int32_t dx = S->dims[0], dy = S->dims[1], dz = S->dims[2];
  const int32_t BX = S->bd[0], BY = S->bd[1];
  parallel_for((size_t)dz, [&](size_t z) {
    for (int32_t y = 0; y < dy; y++)
      for (int32_t x = 0; x < dx; x++) {
        uint32_t g = S->grid[(size_t)(x >> 3) + (size_t)BX * ((size_t)(y >> 3) + (size_t)BY * (z >> 3))];
        uint16_t v = 0;
        if (g & kUniform) v = (uint16_t)(g & 0xffffu);
        else if (g) v = S->pool[(size_t)(g - 1) * 512 + ((x & 7) | ((y & 7) << 3) | (((int)z & 7) << 6))];
        out[(size_t)x + (size_t)dx * ((size_t)y + (size_t)dy * z)] = v;
      }
  });
}

static uint16_t voxel_at(const Scene* S, int32_t x, int32_t y, int32_t z) {
  if (x < 0 || y < 0 || z < 0 || x >= S->dims[0] || y >= S->dims[1] || z >= S->dims[2]) return 0;
  uint32_t g = S->grid[(size_t)(x >> 3) + (size_t)S->bd[0] * ((size_t)(y >> 3) + (size_t)S->bd[1] * (z >> 3))];
  if (g & kUniform) return (uint16_t)(g & 0xffffu);
  if (!g) return 0;
  return S->pool[(size_t)(g - 1) * 512 + ((x & 7) | ((y & 7) << 3) | ((z & 7) << 6))];
}

// Seeded air position with a clear 5^3 neighbourhood (eye, voxel units) plus
// yaw/pitch: cam[0..2] eye, cam[3] yaw, cam[4] pitch.  Returns 0 on success.
int pxs_pick_camera(const pxs_scene* s, uint64_t seed, double* cam) {
  const Scene* S = reinterpret_cast<const Scene*>(s);
  Rng R(seed ^ 0xC0FFEE1234ull);
  for (int tries = 0; tries < 200000; tries++) {
    int32_t x = (int32_t)R.below((uint32_t)S->dims[0]);
    int32_t y = (int32_t)R.below((uint32_t)S->dims[1]);
    int32_t z = (int32_t)R.below((uint32_t)S->dims[2]);
    bool ok = true;
    for (int dz = -2; dz <= 2 && ok; dz++)
      for (int dy = -2; dy <= 2 && ok; dy++)
        for (int dx = -2; dx <= 2 && ok; dx++) {
          int32_t px = x + dx, py = y + dy, pz = z + dz;
          if (px < 0 || py < 0 || pz < 0 || px >= S->dims[0] || py >= S->dims[1] || pz >= S->dims[2]) { ok = false; break; }
          if (voxel_at(S, px, py, pz)) ok = false;
        }
    if (!ok) continue;
    cam[0] = x + 0.5; cam[1] = y + 0.5; cam[2] = z + 0.5;
    cam[3] = R.uni(0.0, 2.0 * M_PI);
    cam[4] = R.uni(-0.35, 0.35);
    return 0;
  }
  return -1;
}

// Pinhole camera rays in float64 rounded to FP32 (32-byte px_ray records):
// vertical field of view fov_deg, pixel centres, ro = eye, max_t = 1e9.
// tile > 0 orders rays in tile x tile pixel blocks (row-major inside a tile).
void pxs_camera_rays(const double* cam, int32_t W, int32_t H, double fov_deg, int32_t tile, float* out) {
  double yaw = cam[3], pitch = cam[4];
  double fw[3] = {std::cos(pitch) * std::cos(yaw), std::sin(pitch), std::cos(pitch) * std::sin(yaw)};
  double right[3] = {-fw[2], 0.0, fw[0]};  // cross(fw, (0,1,0))
  double rl = std::sqrt(right[0] * right[0] + right[2] * right[2]);
  right[0] /= rl; right[2] /= rl;
  double up[3] = {right[1] * fw[2] - right[2] * fw[1], right[2] * fw[0] - right[0] * fw[2],
                  right[0] * fw[1] - right[1] * fw[0]};
  double tf = std::tan(fov_deg * M_PI / 360.0);
  double aspect = (double)W / (double)H;
  int32_t tw = tile > 0 ? tile : W, th = tile > 0 ? tile : 1;
  int32_t ntx = (W + tw - 1) / tw;
  size_t idx = 0;
  int32_t nty = (H + th - 1) / th;
  for (int32_t ty = 0; ty < nty; ty++)
    for (int32_t tx = 0; tx < ntx; tx++)
      for (int32_t j = ty * th; j < std::min(H, (ty + 1) * th); j++)
        for (int32_t i = tx * tw; i < std::min(W, (tx + 1) * tw); i++) {
          double nx = 2.0 * (i + 0.5) / W - 1.0, ny = 1.0 - 2.0 * (j + 0.5) / H;
          double d[3];
          for (int a = 0; a < 3; a++) d[a] = fw[a] + right[a] * (nx * tf * aspect) + up[a] * (ny * tf);
          double l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
          float* r = out + idx * 8;
          for (int a = 0; a < 3; a++) {
            r[a] = (float)cam[a];
            r[3 + a] = (float)(d[a] / l);
          }
          r[6] = 1e9f;
          uint32_t flags = 0;
          std::memcpy(&r[7], &flags, 4);
          idx++;
        }
}

// Uniform block (shared.wgsl:5-19) and light list (render_common.wgsl:8) for a
// frame: sun from a seeded direction, n_lights point lights at uniform
// interior positions (metres = voxels * voxel_size), all casting shadows.
void pxs_frame_inputs(const pxs_scene* s, uint64_t seed, int32_t W, int32_t H, float voxel_size, int n_lights,
                      double light_radius_vox, float* u52, float* lights) {
  const Scene* S = reinterpret_cast<const Scene*>(s);
  Rng R(seed ^ 0x5EEDF00Dull);
  std::memset(u52, 0, 52 * sizeof(float));
  u52[0] = (float)S->dims[0]; u52[1] = (float)S->dims[1]; u52[2] = (float)S->dims[2]; u52[3] = voxel_size;
  double el = R.uni(0.3, 1.2), az = R.uni(0.0, 2.0 * M_PI);
  double sd[3] = {std::cos(el) * std::cos(az), std::sin(el), std::cos(el) * std::sin(az)};
  u52[4] = (float)sd[0]; u52[5] = (float)sd[1]; u52[6] = (float)sd[2]; u52[7] = 1.0f;
  u52[8] = 1.0f; u52[9] = 0.95f; u52[10] = 0.9f;
  u52[12] = 0.3f; u52[13] = 0.3f; u52[14] = 0.35f; u52[15] = 0.5f;
  u52[16] = (float)W; u52[17] = (float)H; u52[18] = 1.0f / (float)W; u52[19] = 1.0f / (float)H;
  u52[20] = 0.0f; u52[21] = 1.0f; u52[22] = 1.0f; u52[23] = 0.0f;
  lights[0] = (float)n_lights; lights[1] = lights[2] = lights[3] = 0.0f;
  double slab = std::min(16.0, S->dims[1] / 4.0);
  for (int i = 0; i < n_lights; i++) {
    float* lp = lights + (1 + 2 * i) * 4;
    float* lc = lights + (2 + 2 * i) * 4;
    double p[3] = {R.uni(0.0, S->dims[0]), R.uni(slab, S->dims[1] - slab), R.uni(0.0, S->dims[2])};
    for (int a = 0; a < 3; a++) lp[a] = (float)(p[a] * voxel_size);
    lp[3] = (float)(light_radius_vox * voxel_size);
    lc[0] = 1.0f; lc[1] = 0.9f; lc[2] = 0.8f; lc[3] = 1.0f;
  }
}

}  // extern "C"
