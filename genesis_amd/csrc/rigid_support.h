// Support maps of the analytic convex shapes (support_field_decomp.py:124-311, mpr_decomp.py:171-197) and the pose
// type the narrowphase passes around (possibly perturbed for multi-contact, kept in registers).
#pragma once
#include "rigid_env.h"

namespace gsr {

// Geometry pose used by the narrowphase (may be perturbed for multi-contact); kept in registers.
struct GPose {
  f3 p;
  q4 q;
};

// Mesh support from the model's direction grid (support_field_decomp.py:145-186): gsr_model.support_v holds, per
// mesh geom, 180 x 180 cells (theta-major) of the vertex that maximises the projection on each grid direction, from
// the geom's first cell support_cell_start. The mesh-frame direction's angles select a cell; of the four cells at
// the floor / ceil corners (theta wrapped, phi kept off the poles) the candidate with the largest projection wins,
// the first on ties. vid: the vertex's index within the geom.
constexpr int kSupportRes = 180;

__device__ inline f3 dev_support_mesh_local(const gsr_model& m, int g, f3 dm, int& vid) {
  const float pi = 3.14159265358979323846f;
  const float theta = cratan2(dm.y, dm.x);
  // a unit direction's z can round past +-1; clamped so the pole cells are looked up instead of a NaN angle
  const float phi = cracos(fminf(fmaxf(dm.z, -1.f), 1.f));
  const float ii = (theta + pi) / pi / 2.f * (float)kSupportRes;
  const float jj = phi / pi * (float)kSupportRes;
  const int i_lo = (int)floorf(ii) % kSupportRes, i_hi = (int)ceilf(ii) % kSupportRes;
  int j_lo = min(max((int)floorf(jj), 0), kSupportRes - 1);
  int j_hi = min(max((int)ceilf(jj), 0), kSupportRes - 1);
  if (j_lo == 0) j_lo = 1;
  if (j_hi == kSupportRes - 1) j_hi = kSupportRes - 2;
  const int base = m.support_cell_start[g];
  float dot_max = -1e20f;
  f3 v = zero3();
  vid = 0;
#pragma unroll
  for (int i4 = 0; i4 < 4; i4++) {
    const int idx = base + ((i4 & 1) ? i_hi : i_lo) * kSupportRes + ((i4 >> 1) ? j_hi : j_lo);
    const f3 pos = T3(m.support_v, idx);
    const float dp = dot(pos, dm);
    if (dp > dot_max) {
      v = pos;
      dot_max = dp;
      vid = m.support_vid[idx];
    }
  }
  return v;
}

// world-frame mesh support (support_field_decomp.py:125-141)
__device__ inline f3 dev_support_mesh(const gsr_model& m, int g, const GPose& P, f3 d, int& vid) {
  return xform(dev_support_mesh_local(m, g, qrot(d, qinv(P.q)), vid), P.p, P.q);
}

__device__ inline f3 dev_support(const gsr_model& m, int g, const GPose& P, f3 d) {
  int t = m.geom_type[g];
  const float* data = m.geom_data + 7 * g;
  if (t == GEOM_MESH) {
    int vid;
    return dev_support_mesh(m, g, P, d, vid);
  }
  if (t == GEOM_SPHERE) return P.p + d * data[0];
  if (t == GEOM_CAPSULE) {
    float r = data[0], hl = 0.5f * data[1];
    f3 axis = qrot(f3{0.f, 0.f, 1.f}, P.q);
    float side = dot(d, axis) < 0.f ? -1.f : 1.f;
    f3 ep = P.p + (hl * side) * axis;
    return ep + d * r;
  }
  if (t == GEOM_BOX) {
    f3 db = qrot_inv(d, P.q);
    f3 v{(db.x < 0.f ? -1.f : 1.f) * data[0] * 0.5f, (db.y < 0.f ? -1.f : 1.f) * data[1] * 0.5f,
         (db.z < 0.f ? -1.f : 1.f) * data[2] * 0.5f};
    return xform(v, P.p, P.q);
  }
  return P.p;
}

}  // namespace gsr
