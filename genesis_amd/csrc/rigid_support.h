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

__device__ inline f3 dev_support(const gsr_model& m, int g, const GPose& P, f3 d) {
  int t = m.geom_type[g];
  const float* data = m.geom_data + 7 * g;
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
