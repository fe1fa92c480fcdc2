// Device-side vector, quaternion and spatial-algebra helpers for the gfx950 rigid step.
// Quaternions are (w, x, y, z). Operation order follows genesis/utils/geom.py:18-398 so that FP32 results track the
// reference's Taichi functions; the library is compiled with -ffp-contract=off (no implicit FMA).
#pragma once
#include <hip/hip_runtime.h>

namespace gsr {

// Transcendentals are evaluated in double and rounded once to float: the result is the correctly rounded float
// value except in vanishingly rare ties, which makes it independent of the float libm in use (device or host).
__device__ __forceinline__ float crsin(float x) { return (float)sin((double)x); }
__device__ __forceinline__ float crcos(float x) { return (float)cos((double)x); }
__device__ __forceinline__ float cratan2(float y, float x) { return (float)atan2((double)y, (double)x); }
__device__ __forceinline__ float cracos(float x) { return (float)acos((double)x); }
__device__ __forceinline__ float crpow(float x, float y) { return (float)pow((double)x, (double)y); }

struct f3 {
  float x, y, z;
};
struct q4 {
  float w, x, y, z;
};
struct m3 {
  float a[9];  // row-major
};

__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 zero3() { return f3{0.f, 0.f, 0.f}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float norm(f3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ f3 normalized(f3 a) { return a / norm(a); }
__device__ __forceinline__ f3 vmin(f3 a, f3 b) { return f3{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
__device__ __forceinline__ f3 vmax(f3 a, f3 b) { return f3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
__device__ __forceinline__ f3 vabs(f3 a) { return f3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
__device__ __forceinline__ float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

__device__ __forceinline__ q4 qid() { return q4{1.f, 0.f, 0.f, 0.f}; }
__device__ __forceinline__ q4 qinv(q4 q) { return q4{q.w, -q.x, -q.y, -q.z}; }

// u * v through the 4x4 outer product (geom.py:185-191)
__device__ __forceinline__ q4 qmul(q4 u, q4 v) {
  float vu00 = u.w * v.w, vu01 = u.w * v.x, vu02 = u.w * v.y, vu03 = u.w * v.z;
  float vu10 = u.x * v.w, vu11 = u.x * v.x, vu12 = u.x * v.y, vu13 = u.x * v.z;
  float vu20 = u.y * v.w, vu21 = u.y * v.x, vu22 = u.y * v.y, vu23 = u.y * v.z;
  float vu30 = u.z * v.w, vu31 = u.z * v.x, vu32 = u.z * v.y, vu33 = u.z * v.z;
  return q4{vu00 - vu11 - vu22 - vu33, vu01 + vu10 + vu23 - vu32, vu02 - vu13 + vu20 + vu31,
            vu03 + vu12 - vu21 + vu30};
}
__device__ __forceinline__ q4 qnormalized(q4 q) {
  float n = sqrtf(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
  return q4{q.w / n, q.x / n, q.y / n, q.z / n};
}
// rotation of quaternion v by u: normalize(u * v)
__device__ __forceinline__ q4 qrot_q(q4 v, q4 u) { return qnormalized(qmul(u, v)); }

// rotate a vector by a (not necessarily unit) quaternion (geom.py:203-216)
__device__ __forceinline__ f3 qrot(f3 v, q4 q) {
  float q_xx = q.x * q.x, q_xy = q.x * q.y, q_xz = q.x * q.z, q_wx = q.x * q.w;
  float q_yy = q.y * q.y, q_yz = q.y * q.z, q_wy = q.y * q.w;
  float q_zz = q.z * q.z, q_wz = q.z * q.w;
  float q_ww = q.w * q.w;
  f3 r = f3{v.x * (q_xx + q_ww - q_yy - q_zz) + v.y * (2.f * q_xy - 2.f * q_wz) + v.z * (2.f * q_xz + 2.f * q_wy),
            v.x * (2.f * q_wz + 2.f * q_xy) + v.y * (q_ww - q_xx + q_yy - q_zz) + v.z * (-2.f * q_wx + 2.f * q_yz),
            v.x * (-2.f * q_wy + 2.f * q_xz) + v.y * (2.f * q_wx + 2.f * q_yz) + v.z * (q_ww - q_xx - q_yy + q_zz)};
  float d = q_ww + q_xx + q_yy + q_zz;
  return r / d;
}
__device__ __forceinline__ f3 qrot_inv(f3 v, q4 q) { return qrot(v, qinv(q)); }
__device__ __forceinline__ f3 xform(f3 p, f3 t, q4 q) { return qrot(p, q) + t; }

__device__ __forceinline__ m3 quat_to_R(q4 q, float eps) {
  m3 R;
  float d = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
  if (d > eps) {
    float s = 2.f / d;
    float xs = q.x * s, ys = q.y * s, zs = q.z * s;
    float wx = q.w * xs, wy = q.w * ys, wz = q.w * zs;
    float xx = q.x * xs, xy = q.x * ys, xz = q.x * zs;
    float yy = q.y * ys, yz = q.y * zs, zz = q.z * zs;
    R.a[0] = 1.f - (yy + zz); R.a[1] = xy - wz; R.a[2] = xz + wy;
    R.a[3] = xy + wz; R.a[4] = 1.f - (xx + zz); R.a[5] = yz - wx;
    R.a[6] = xz - wy; R.a[7] = yz + wx; R.a[8] = 1.f - (xx + yy);
  } else {
    R.a[0] = 1.f; R.a[1] = 0.f; R.a[2] = 0.f; R.a[3] = 0.f; R.a[4] = 1.f; R.a[5] = 0.f;
    R.a[6] = 0.f; R.a[7] = 0.f; R.a[8] = 1.f;
  }
  return R;
}
__device__ __forceinline__ f3 mcol(const m3& R, int j) { return f3{R.a[j], R.a[3 + j], R.a[6 + j]}; }
__device__ __forceinline__ f3 mrow(const m3& R, int i) { return f3{R.a[3 * i], R.a[3 * i + 1], R.a[3 * i + 2]}; }
__device__ __forceinline__ f3 mmul(const m3& A, f3 v) {
  return f3{A.a[0] * v.x + A.a[1] * v.y + A.a[2] * v.z, A.a[3] * v.x + A.a[4] * v.y + A.a[5] * v.z,
            A.a[6] * v.x + A.a[7] * v.y + A.a[8] * v.z};
}

__device__ __forceinline__ q4 rotvec_to_quat(f3 rv, float eps) {
  float thetasq = rv.x * rv.x + rv.y * rv.y + rv.z * rv.z;
  if (thetasq > eps * eps) {
    float theta = sqrtf(thetasq);
    float th = 0.5f * theta;
    float c = crcos(th), s = crsin(th);
    f3 xyz = (s / theta) * rv;
    q4 q{c, xyz.x, xyz.y, xyz.z};
    float nsq = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
    float f = 0.5f * (3.f - nsq);
    return q4{q.w * f, q.x * f, q.y * f, q.z * f};
  }
  return qid();
}

__device__ __forceinline__ f3 quat_to_xyz(q4 q, float eps) {
  float roll = 0.f, pitch = 0.f, yaw = 0.f;
  float nsq = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
  if (nsq > eps) {
    float s = 2.f / nsq;
    float q_xs = q.x * s, q_ys = q.y * s, q_zs = q.z * s;
    float q_wx = q.w * q_xs, q_wy = q.w * q_ys, q_wz = q.w * q_zs;
    float q_xx = q.x * q_xs, q_xy = q.x * q_ys, q_xz = q.x * q_zs;
    float q_yy = q.y * q_ys, q_yz = q.y * q_zs, q_zz = q.z * q_zs;
    float sinycosp = q_wz - q_xy;
    float cosycosp = 1.f - (q_yy + q_zz);
    float cosp = sqrtf(cosycosp * cosycosp + sinycosp * sinycosp);
    pitch = cratan2(q_xz + q_wy, cosp);
    if (cosp > eps) {
      roll = cratan2(q_wx - q_yz, 1.f - (q_xx + q_yy));
      yaw = cratan2(sinycosp, cosycosp);
    } else {
      yaw = cratan2(q_wz + q_xy, 1.f - (q_xx + q_zz));
    }
  }
  return f3{roll, pitch, yaw};
}

__device__ __forceinline__ q4 xyz_to_quat(f3 xyz) {
  float ai = 0.5f * xyz.z, aj = -0.5f * xyz.y, ak = 0.5f * xyz.x;
  float si = crsin(ai), sj = crsin(aj), sk = crsin(ak);
  float ci = crcos(ai), cj = crcos(aj), ck = crcos(ak);
  float cc = ci * ck, cs = ci * sk, sc = si * ck, ss = si * sk;
  return q4{cj * cc + sj * ss, cj * cs - sj * sc, -cj * ss - sj * cc, cj * sc - sj * cs};
}

__device__ __forceinline__ f3 quat_to_rotvec(q4 q, float eps) {
  f3 rv{q.x, q.y, q.z};
  float s2 = norm(rv);
  if (s2 > eps) {
    float angle = 2.f * cratan2(s2, fabsf(q.w));
    float inv_sinc = angle / s2;
    rv = ((q.w < 0.f ? -1.f : 1.f) * inv_sinc) * rv;
  }
  return rv;
}

// spatial algebra (geom.py:338-356)
__device__ __forceinline__ void inertial_mul(f3 pos, const float* I, float mass, f3 vel, f3 ang, f3& o_ang, f3& o_vel) {
  f3 Ia = f3{I[0] * ang.x + I[1] * ang.y + I[2] * ang.z, I[3] * ang.x + I[4] * ang.y + I[5] * ang.z,
             I[6] * ang.x + I[7] * ang.y + I[8] * ang.z};
  o_ang = Ia + cross(pos, vel);
  o_vel = mass * vel - cross(pos, ang);
}
__device__ __forceinline__ void motion_cross_force(f3 m_ang, f3 m_vel, f3 f_ang, f3 f_vel, f3& o_ang, f3& o_vel) {
  o_vel = cross(m_ang, f_vel);
  o_ang = cross(m_ang, f_ang) + cross(m_vel, f_vel);
}
__device__ __forceinline__ void motion_cross_motion(f3 s_ang, f3 s_vel, f3 m_ang, f3 m_vel, f3& o_ang, f3& o_vel) {
  o_vel = cross(s_ang, m_vel) + cross(s_vel, m_ang);
  o_ang = cross(s_ang, m_ang);
}
__device__ __forceinline__ void orthogonals(f3 a, f3& b, f3& c) {
  f3 bb;
  if (fabsf(a.y) < 0.5f) {
    bb = f3{-a.x * a.y, 1.f - a.y * a.y, -a.z * a.y};
  } else {
    bb = f3{-a.x * a.z, -a.y * a.z, 1.f - a.z * a.z};
  }
  b = normalized(bb);
  c = cross(a, b);
}
// constraint impedance and reference acceleration (geom.py:377-398); the 7 solver parameters at p[k * st]
__device__ __forceinline__ void imp_aref(const float* p, size_t st, float neg_pen, float vel, float pos, float& imp,
                                         float& aref) {
  float timeconst = p[0], dampratio = p[st], dmin = p[2 * st], dmax = p[3 * st], width = p[4 * st], mid = p[5 * st],
        power = p[6 * st];
  float imp_x = fabsf(neg_pen) / width;
  float imp_a = (1.f / crpow(mid, power - 1.f)) * crpow(imp_x, power);
  float imp_b = 1.f - (1.f / crpow(1.f - mid, power - 1.f)) * crpow(1.f - imp_x, power);
  float imp_y = imp_x < mid ? imp_a : imp_b;
  imp = dmin + imp_y * (dmax - dmin);
  imp = fminf(fmaxf(imp, dmin), dmax);
  imp = imp_x > 1.f ? dmax : imp;
  float bb = 2.f / (dmax * timeconst);
  float k = 1.f / (dmax * dmax * timeconst * timeconst * dampratio * dampratio);
  aref = -bb * vel - k * imp * pos;
}
__device__ __forceinline__ void imp_aref(const float* p, float neg_pen, float vel, float pos, float& imp, float& aref) {
  imp_aref(p, 1, neg_pen, vel, pos, imp, aref);
}

}  // namespace gsr
