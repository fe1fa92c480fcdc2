// CPU oracle -- TEST INFRASTRUCTURE ONLY. Nothing in the product (genesis_amd/) may include or call this.
//
// Small vector/quaternion helpers restating genesis/utils/geom.py:18-398 (the `ti_*` Taichi functions) with the
// same operation order. Quaternions are (w, x, y, z).
#pragma once
#include <cmath>
#include <algorithm>

namespace orc {

// Transcendentals: evaluated in double, rounded once to T (the correctly rounded value for T = float in all but
// vanishingly rare ties), so the float result does not depend on the float libm.
template <class T> inline T tsin(T x) { return T(std::sin(double(x))); }
template <class T> inline T tcos(T x) { return T(std::cos(double(x))); }
template <class T> inline T tatan2(T y, T x) { return T(std::atan2(double(y), double(x))); }
template <class T> inline T tacos(T x) { return T(std::acos(double(x))); }
template <class T> inline T tpow(T x, T y) { return T(std::pow(double(x), double(y))); }

template <class T>
struct V3 {
  T x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(T a, T b, T c) : x(a), y(b), z(c) {}
  T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T norm(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> inline V3<T> normalized(V3<T> a) { return a / norm(a); }
template <class T> inline V3<T> vmin(V3<T> a, V3<T> b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
template <class T> inline V3<T> vmax(V3<T> a, V3<T> b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }
template <class T> inline V3<T> vabs(V3<T> a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }

template <class T>
struct Q4 {
  T w, x, y, z;
  Q4() : w(1), x(0), y(0), z(0) {}
  Q4(T a, T b, T c, T d) : w(a), x(b), y(c), z(d) {}
  T operator[](int i) const { return i == 0 ? w : (i == 1 ? x : (i == 2 ? y : z)); }
};

template <class T>
struct M3 {
  T m[3][3];
  M3() { for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) m[i][j] = 0; }
  static M3 eye() { M3 r; r.m[0][0] = r.m[1][1] = r.m[2][2] = 1; return r; }
  V3<T> col(int j) const { return {m[0][j], m[1][j], m[2][j]}; }
  V3<T> row(int i) const { return {m[i][0], m[i][1], m[i][2]}; }
};
template <class T> inline V3<T> mul(const M3<T>& A, V3<T> v) {
  return {A.m[0][0] * v.x + A.m[0][1] * v.y + A.m[0][2] * v.z, A.m[1][0] * v.x + A.m[1][1] * v.y + A.m[1][2] * v.z,
          A.m[2][0] * v.x + A.m[2][1] * v.y + A.m[2][2] * v.z};
}
template <class T> inline M3<T> mul(const M3<T>& A, const M3<T>& B) {
  M3<T> r;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) r.m[i][j] = A.m[i][0] * B.m[0][j] + A.m[i][1] * B.m[1][j] + A.m[i][2] * B.m[2][j];
  return r;
}
template <class T> inline M3<T> transpose(const M3<T>& A) {
  M3<T> r;
  for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) r.m[i][j] = A.m[j][i];
  return r;
}

// ti_quat_mul (geom.py:185-191): outer-product formulation
template <class T> inline Q4<T> quat_mul(Q4<T> u, Q4<T> v) {
  T vu00 = u.w * v.w, vu01 = u.w * v.x, vu02 = u.w * v.y, vu03 = u.w * v.z;
  T vu10 = u.x * v.w, vu11 = u.x * v.x, vu12 = u.x * v.y, vu13 = u.x * v.z;
  T vu20 = u.y * v.w, vu21 = u.y * v.x, vu22 = u.y * v.y, vu23 = u.y * v.z;
  T vu30 = u.z * v.w, vu31 = u.z * v.x, vu32 = u.z * v.y, vu33 = u.z * v.z;
  T w = vu00 - vu11 - vu22 - vu33;
  T x = vu01 + vu10 + vu23 - vu32;
  T y = vu02 - vu13 + vu20 + vu31;
  T z = vu03 + vu12 - vu21 + vu30;
  return {w, x, y, z};
}
template <class T> inline Q4<T> quat_normalized(Q4<T> q) {
  T n = std::sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
  return {q.w / n, q.x / n, q.y / n, q.z / n};
}
// ti_transform_quat_by_quat(v, u) = normalize(u * v) (geom.py:194-200)
template <class T> inline Q4<T> transform_quat_by_quat(Q4<T> v, Q4<T> u) { return quat_normalized(quat_mul(u, v)); }
template <class T> inline Q4<T> inv_quat(Q4<T> q) { return {q.w, -q.x, -q.y, -q.z}; }

// ti_transform_by_quat (geom.py:203-216)
template <class T> inline V3<T> transform_by_quat(V3<T> v, Q4<T> q) {
  T q_xx = q.x * q.x, q_xy = q.x * q.y, q_xz = q.x * q.z, q_wx = q.x * q.w;
  T q_yy = q.y * q.y, q_yz = q.y * q.z, q_wy = q.y * q.w;
  T q_zz = q.z * q.z, q_wz = q.z * q.w;
  T q_ww = q.w * q.w;
  T two = T(2);
  V3<T> r(v.x * (q_xx + q_ww - q_yy - q_zz) + v.y * (two * q_xy - two * q_wz) + v.z * (two * q_xz + two * q_wy),
          v.x * (two * q_wz + two * q_xy) + v.y * (q_ww - q_xx + q_yy - q_zz) + v.z * (-two * q_wx + two * q_yz),
          v.x * (-two * q_wy + two * q_xz) + v.y * (two * q_wx + two * q_yz) + v.z * (q_ww - q_xx - q_yy + q_zz));
  T d = q_ww + q_xx + q_yy + q_zz;
  return r / d;
}
template <class T> inline V3<T> inv_transform_by_quat(V3<T> v, Q4<T> q) { return transform_by_quat(v, inv_quat(q)); }
template <class T> inline V3<T> transform_by_trans_quat(V3<T> p, V3<T> t, Q4<T> q) { return transform_by_quat(p, q) + t; }

// ti_quat_to_R (geom.py:105-125)
template <class T> inline M3<T> quat_to_R(Q4<T> q, T eps) {
  M3<T> R = M3<T>::eye();
  T d = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
  if (d > eps) {
    T s = T(2) / d;
    T xs = q.x * s, ys = q.y * s, zs = q.z * s;
    T wx = q.w * xs, wy = q.w * ys, wz = q.w * zs;
    T xx = q.x * xs, xy = q.x * ys, xz = q.x * zs;
    T yy = q.y * ys, yz = q.y * zs, zz = q.z * zs;
    R.m[0][0] = T(1) - (yy + zz); R.m[0][1] = xy - wz; R.m[0][2] = xz + wy;
    R.m[1][0] = xy + wz; R.m[1][1] = T(1) - (xx + zz); R.m[1][2] = yz - wx;
    R.m[2][0] = xz - wy; R.m[2][1] = yz + wx; R.m[2][2] = T(1) - (xx + yy);
  }
  return R;
}

// ti_rotvec_to_quat (geom.py:83-102)
template <class T> inline Q4<T> rotvec_to_quat(V3<T> rv, T eps) {
  Q4<T> q(0, 0, 0, 0);
  T thetasq = rv.x * rv.x + rv.y * rv.y + rv.z * rv.z;
  if (thetasq > eps * eps) {
    T theta = std::sqrt(thetasq);
    T th = T(0.5) * theta;
    T c = tcos(th), s = tsin(th);
    q.w = c;
    V3<T> xyz = (s / theta) * rv;
    q.x = xyz.x; q.y = xyz.y; q.z = xyz.z;
    T nsq = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
    T f = T(0.5) * (T(3) - nsq);
    q.w *= f; q.x *= f; q.y *= f; q.z *= f;
  } else {
    q.w = 1;
  }
  return q;
}

// ti_quat_to_xyz (geom.py:128-153)
template <class T> inline V3<T> quat_to_xyz(Q4<T> q, T eps) {
  T roll = 0, pitch = 0, yaw = 0;
  T nsq = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
  if (nsq > eps) {
    T s = T(2) / nsq;
    T q_xs = q.x * s, q_ys = q.y * s, q_zs = q.z * s;
    T q_wx = q.w * q_xs, q_wy = q.w * q_ys, q_wz = q.w * q_zs;
    T q_xx = q.x * q_xs, q_xy = q.x * q_ys, q_xz = q.x * q_zs;
    T q_yy = q.y * q_ys, q_yz = q.y * q_zs, q_zz = q.z * q_zs;
    T sinycosp = q_wz - q_xy;
    T cosycosp = T(1) - (q_yy + q_zz);
    T cosp = std::sqrt(cosycosp * cosycosp + sinycosp * sinycosp);
    pitch = tatan2(q_xz + q_wy, cosp);
    if (cosp > eps) {
      roll = tatan2(q_wx - q_yz, T(1) - (q_xx + q_yy));
      yaw = tatan2(sinycosp, cosycosp);
    } else {
      yaw = tatan2(q_wz + q_xy, T(1) - (q_xx + q_zz));
    }
  }
  return {roll, pitch, yaw};
}

// ti_xyz_to_quat (geom.py:18-36)
template <class T> inline Q4<T> xyz_to_quat(V3<T> xyz) {
  T ai = T(0.5) * xyz.z, aj = T(-0.5) * xyz.y, ak = T(0.5) * xyz.x;
  T si = tsin(ai), sj = tsin(aj), sk = tsin(ak);
  T ci = tcos(ai), cj = tcos(aj), ck = tcos(ak);
  T cc = ci * ck, cs = ci * sk, sc = si * ck, ss = si * sk;
  return {cj * cc + sj * ss, cj * cs - sj * sc, -cj * ss - sj * cc, cj * sc - sj * cs};
}

// ti_quat_to_rotvec (geom.py:156-165)
template <class T> inline V3<T> quat_to_rotvec(Q4<T> q, T eps) {
  V3<T> rv(q.x, q.y, q.z);
  T s2 = norm(rv);
  if (s2 > eps) {
    T angle = T(2) * tatan2(s2, std::fabs(q.w));
    T inv_sinc = angle / s2;
    rv = ((q.w < 0 ? T(-1) : T(1)) * inv_sinc) * rv;
  }
  return rv;
}

// gu.inertial_mul (geom.py:338-342): returns (ang, vel)
template <class T> inline void inertial_mul(V3<T> pos, const M3<T>& I, T mass, V3<T> vel, V3<T> ang, V3<T>& o_ang, V3<T>& o_vel) {
  o_ang = mul(I, ang) + cross(pos, vel);
  o_vel = mass * vel - cross(pos, ang);
}
// motion_cross_force (geom.py:345-349)
template <class T> inline void motion_cross_force(V3<T> m_ang, V3<T> m_vel, V3<T> f_ang, V3<T> f_vel, V3<T>& o_ang, V3<T>& o_vel) {
  o_vel = cross(m_ang, f_vel);
  o_ang = cross(m_ang, f_ang) + cross(m_vel, f_vel);
}
// motion_cross_motion (geom.py:352-356)
template <class T> inline void motion_cross_motion(V3<T> s_ang, V3<T> s_vel, V3<T> m_ang, V3<T> m_vel, V3<T>& o_ang, V3<T>& o_vel) {
  o_vel = cross(s_ang, m_vel) + cross(s_vel, m_ang);
  o_ang = cross(s_ang, m_ang);
}
// ti_orthogonals (geom.py:359-374)
template <class T> inline void orthogonals(V3<T> a, V3<T>& b, V3<T>& c) {
  V3<T> bb;
  if (std::fabs(a.y) < T(0.5)) {
    bb.x = -a.x * a.y; bb.y = T(1) - a.y * a.y; bb.z = -a.z * a.y;
  } else {
    bb.x = -a.x * a.z; bb.y = -a.y * a.z; bb.z = T(1) - a.z * a.z;
  }
  b = normalized(bb);
  c = cross(a, b);
}
// imp_aref (geom.py:377-398)
template <class T> inline void imp_aref(const T* p, T neg_pen, T vel, T pos, T& imp, T& aref) {
  T timeconst = p[0], dampratio = p[1], dmin = p[2], dmax = p[3], width = p[4], mid = p[5], power = p[6];
  T imp_x = std::fabs(neg_pen) / width;
  T imp_a = (T(1) / tpow(mid, power - T(1))) * tpow(imp_x, power);
  T imp_b = T(1) - (T(1) / tpow(T(1) - mid, power - T(1))) * tpow(T(1) - imp_x, power);
  T imp_y = imp_x < mid ? imp_a : imp_b;
  imp = dmin + imp_y * (dmax - dmin);
  imp = std::min(std::max(imp, dmin), dmax);
  imp = imp_x > T(1) ? dmax : imp;
  T b = T(2) / (dmax * timeconst);
  T k = T(1) / (dmax * dmax * timeconst * timeconst * dampratio * dampratio);
  aref = -b * vel - k * imp * pos;
}

}  // namespace orc
