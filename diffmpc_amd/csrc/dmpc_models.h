// Dynamics models and integrators as device templates over the scalar type.
//
// Every model is a struct with NX, NU and a templated
//   f(const T* x, const T* u, PA prm, T* xdot)
// where T is float or Dual (forward-mode derivative with one tangent, for
// the linearization), double (value only, for the line search's merit
// evaluations) or HD (float64 hyper-dual: two tangents and their mixed
// second derivative, for the dynamics-parameter gradients).  `prm` is
// anything indexable that yields the flat parameters: a `const float*`, or
// ParamTangent, which seeds one parameter's tangent.  Constants derived from
// the parameters are computed in the scalar type S of T (float for float and
// Dual, double for double, HD for HD).  The flat parameter layouts are the
// ones documented in include/dmpc.h.
#pragma once
#include <type_traits>
#include <hip/hip_runtime.h>

namespace dmpc {

// ---------------------------------------------------------------------------
// Forward-mode dual number with one tangent direction.
struct Dual {
  float v, d;
};
__device__ __forceinline__ Dual mk(float v) { return Dual{v, 0.0f}; }
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator*(float s, Dual a) { return {s * a.v, s * a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, float s) { return {s * a.v, s * a.d}; }
__device__ __forceinline__ Dual operator+(Dual a, float s) { return {a.v + s, a.d}; }
__device__ __forceinline__ Dual operator-(Dual a, float s) { return {a.v - s, a.d}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  float inv = 1.0f / b.v;
  float q = a.v * inv;
  return {q, (a.d - q * b.d) * inv};
}
__device__ __forceinline__ Dual operator/(Dual a, float s) { return {a.v / s, a.d / s}; }
__device__ __forceinline__ Dual dsqrt(Dual a) {
  float s = sqrtf(a.v);
  return {s, a.d * (0.5f / s)};
}
__device__ __forceinline__ float dsqrt(float a) { return sqrtf(a); }
__device__ __forceinline__ double dsqrt(double a) { return sqrt(a); }
__device__ __forceinline__ Dual dsin(Dual a) { return {sinf(a.v), a.d * cosf(a.v)}; }
__device__ __forceinline__ Dual dcos(Dual a) { return {cosf(a.v), -a.d * sinf(a.v)}; }
__device__ __forceinline__ float dsin(float a) { return sinf(a); }
__device__ __forceinline__ float dcos(float a) { return cosf(a); }
__device__ __forceinline__ double dsin(double a) { return sin(a); }
__device__ __forceinline__ double dcos(double a) { return cos(a); }
__device__ __forceinline__ Dual dexp(Dual a) {
  const float e = expf(a.v);
  return {e, a.d * e};
}
__device__ __forceinline__ float dexp(float a) { return expf(a); }
__device__ __forceinline__ double dexp(double a) { return exp(a); }
__device__ __forceinline__ float val(float a) { return a; }
__device__ __forceinline__ float val(Dual a) { return a.v; }
__device__ __forceinline__ float tan_(Dual a) { return a.d; }

// ---------------------------------------------------------------------------
// Float64 hyper-dual number v + a e1 + b e2 + ab e1 e2 (e1^2 = e2^2 = 0):
// evaluating f on (x + e1 dx, p + e2 dp) gives f, its directional
// derivatives along dx (a) and dp (b) and the mixed second derivative
// d2f/dx dp [dx, dp] (ab), all exact to rounding.
struct HD {
  double v, a, b, ab;
  HD() = default;
  __host__ __device__ constexpr HD(double x) : v(x), a(0.0), b(0.0), ab(0.0) {}
  __host__ __device__ constexpr HD(double v_, double a_, double b_, double ab_) : v(v_), a(a_), b(b_), ab(ab_) {}
};
__device__ __forceinline__ HD operator+(HD x, HD y) { return {x.v + y.v, x.a + y.a, x.b + y.b, x.ab + y.ab}; }
__device__ __forceinline__ HD operator-(HD x, HD y) { return {x.v - y.v, x.a - y.a, x.b - y.b, x.ab - y.ab}; }
__device__ __forceinline__ HD operator-(HD x) { return {-x.v, -x.a, -x.b, -x.ab}; }
__device__ __forceinline__ HD operator*(HD x, HD y) {
  return {x.v * y.v, x.v * y.a + x.a * y.v, x.v * y.b + x.b * y.v,
          x.v * y.ab + x.a * y.b + x.b * y.a + x.ab * y.v};
}
// g(x) for a scalar function with g(v) = g0, g'(v) = g1, g''(v) = g2
__device__ __forceinline__ HD hd_apply(HD x, double g0, double g1, double g2) {
  return {g0, g1 * x.a, g1 * x.b, g1 * x.ab + g2 * x.a * x.b};
}
__device__ __forceinline__ HD operator/(HD x, HD y) {
  const double r = 1.0 / y.v;
  return x * hd_apply(y, r, -r * r, 2.0 * r * r * r);
}
__device__ __forceinline__ HD dsqrt(HD x) {
  const double s = sqrt(x.v);
  return hd_apply(x, s, 0.5 / s, -0.25 / (s * x.v));
}
__device__ __forceinline__ HD dsin(HD x) {
  const double s = sin(x.v), c = cos(x.v);
  return hd_apply(x, s, c, -s);
}
__device__ __forceinline__ HD dexp(HD x) {
  const double e = exp(x.v);
  return hd_apply(x, e, e, e);
}
__device__ __forceinline__ HD dcos(HD x) {
  const double s = sin(x.v), c = cos(x.v);
  return hd_apply(x, c, -s, -c);
}

// Flat parameters with the e2 tangent on parameter j (hyper-dual values).
struct ParamTangent {
  const float* p;
  int j;
  __device__ HD operator[](int q) const { return HD((double)p[q], 0.0, q == j ? 1.0 : 0.0, 0.0); }
};

// scalar type of T
template <class T>
struct ScalarOf {
  using type = float;
};
template <>
struct ScalarOf<double> {
  using type = double;
};
template <>
struct ScalarOf<HD> {
  using type = HD;
};
template <class T>
using scalar_t = typename ScalarOf<T>::type;

template <class T>
__device__ __forceinline__ T sc(scalar_t<T> v);
template <>
__device__ __forceinline__ float sc<float>(float v) { return v; }
template <>
__device__ __forceinline__ Dual sc<Dual>(float v) { return Dual{v, 0.0f}; }
template <>
__device__ __forceinline__ double sc<double>(double v) { return v; }
template <>
__device__ __forceinline__ HD sc<HD>(HD v) { return v; }

enum ModelId : int {
  MODEL_LINEAR = 0,
  MODEL_QUADROTOR = 1,
  MODEL_SPACECRAFT3 = 2,
  MODEL_SPACECRAFT7 = 3,
  MODEL_CARTPOLE = 4,
  MODEL_CHAIN = 5,
  MODEL_CUSTOM = 6,  // the user-defined model of a custom build (CustomModel)
};

// ---------------------------------------------------------------------------
// xdot = A x + B u + b   (diffmpc/dynamics/linear_dynamics.py:46-70)
// Linear in (x, u): the forward-mode Jacobian's tangent parts never involve
// the point's values, so [A_t | B_t] comes out bit for bit the same at every
// knot (kTimeInvariant).
template <int NX_, int NU_>
struct LinearModel {
  static constexpr int NX = NX_, NU = NU_;
  static constexpr int NPARAM = NX * NX + NX * NU + NX;
  static constexpr bool kTimeInvariant = true;
  template <class T, class PA>
  __device__ static void f(const T* x, const T* u, PA p, T* xd) {
#pragma unroll
    for (int i = 0; i < NX; ++i) {
      T acc = sc<T>(p[NX * NX + NX * NU + i]);
#pragma unroll
      for (int j = 0; j < NX; ++j) acc = acc + p[i * NX + j] * x[j];
#pragma unroll
      for (int k = 0; k < NU; ++k) acc = acc + p[NX * NX + i * NU + k] * u[k];
      xd[i] = acc;
    }
  }
};

// Quadrotor, 13 states [p, v, q (scalar first), w], 4 controls [f, tau].
// diffmpc/dynamics/quadrotor_dynamics.py:84-133.  Parameters: mass, J (3x3).
struct QuadrotorModel {
  static constexpr int NX = 13, NU = 4, NPARAM = 10;
  template <class T, class PA>
  __device__ static void f(const T* x, const T* u, PA p, T* xd) {
    using S = scalar_t<T>;
    const S m = p[0];
    S J[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) J[i] = p[1 + i];
    // inverse of J by the adjugate (J is constant: plain scalars)
    S c00 = J[4] * J[8] - J[5] * J[7], c01 = J[5] * J[6] - J[3] * J[8], c02 = J[3] * J[7] - J[4] * J[6];
    S det = J[0] * c00 + J[1] * c01 + J[2] * c02;
    S id = S(1) / det;
    S Ji[9] = {c00 * id, (J[2] * J[7] - J[1] * J[8]) * id, (J[1] * J[5] - J[2] * J[4]) * id,
               c01 * id, (J[0] * J[8] - J[2] * J[6]) * id, (J[2] * J[3] - J[0] * J[5]) * id,
               c02 * id, (J[1] * J[6] - J[0] * J[7]) * id, (J[0] * J[4] - J[1] * J[3]) * id};
    T qw = x[6], qx = x[7], qy = x[8], qz = x[9];
    T inv = sc<T>(S(1)) / dsqrt(qw * qw + qx * qx + qy * qy + qz * qz);
    qw = qw * inv; qx = qx * inv; qy = qy * inv; qz = qz * inv;
    T wx = x[10], wy = x[11], wz = x[12];
    T fm = u[0] * (S(1) / m);
    xd[0] = x[3]; xd[1] = x[4]; xd[2] = x[5];
    // v' = -g e3 + (f/m) R(q) e3
    xd[3] = fm * (S(2) * (qx * qz + qw * qy));
    xd[4] = fm * (S(2) * (qy * qz - qw * qx));
    xd[5] = fm * (qw * qw - qx * qx - qy * qy + qz * qz) - S(9.81);
    // q' = 1/2 L(q) [0; w]
    xd[6] = S(0.5) * (-(qx * wx) - qy * wy - qz * wz);
    xd[7] = S(0.5) * (qw * wx - qz * wy + qy * wz);
    xd[8] = S(0.5) * (qz * wx + qw * wy - qx * wz);
    xd[9] = S(0.5) * (qx * wy - qy * wx + qw * wz);
    // w' = J^-1 ((J w) x w + tau)
    T hx = J[0] * wx + J[1] * wy + J[2] * wz;
    T hy = J[3] * wx + J[4] * wy + J[5] * wz;
    T hz = J[6] * wx + J[7] * wy + J[8] * wz;
    T ex = hy * wz - hz * wy + u[1];
    T ey = hz * wx - hx * wz + u[2];
    T ez = hx * wy - hy * wx + u[3];
    xd[10] = Ji[0] * ex + Ji[1] * ey + Ji[2] * ez;
    xd[11] = Ji[3] * ex + Ji[4] * ey + Ji[5] * ez;
    xd[12] = Ji[6] * ex + Ji[7] * ey + Ji[8] * ez;
  }
};

// Rate-only spacecraft: w' = J^-1 (tau - w x J w), diagonal J.
// diffmpc/dynamics/spacecraft_dynamics.py:31-66.
struct Spacecraft3Model {
  static constexpr int NX = 3, NU = 3, NPARAM = 3;
  template <class T, class PA>
  __device__ static void f(const T* w, const T* u, PA p, T* xd) {
    using S = scalar_t<T>;
    const S J0 = p[0], J1 = p[1], J2 = p[2];
    T hx = J0 * w[0], hy = J1 * w[1], hz = J2 * w[2];
    xd[0] = (u[0] - (w[1] * hz - w[2] * hy)) * (S(1) / J0);
    xd[1] = (u[1] - (w[2] * hx - w[0] * hz)) * (S(1) / J1);
    xd[2] = (u[2] - (w[0] * hy - w[1] * hx)) * (S(1) / J2);
  }
};

// Attitude spacecraft, 7 states [q (scalar first), w], 3 torques: the
// normalised quaternion kinematics of quadrotor_dynamics.py:122,129 with
// Euler's equation of spacecraft_dynamics.py:63 (diagonal J).
struct Spacecraft7Model {
  static constexpr int NX = 7, NU = 3, NPARAM = 3;
  template <class T, class PA>
  __device__ static void f(const T* x, const T* u, PA p, T* xd) {
    using S = scalar_t<T>;
    T qw = x[0], qx = x[1], qy = x[2], qz = x[3];
    T inv = sc<T>(S(1)) / dsqrt(qw * qw + qx * qx + qy * qy + qz * qz);
    qw = qw * inv; qx = qx * inv; qy = qy * inv; qz = qz * inv;
    T wx = x[4], wy = x[5], wz = x[6];
    xd[0] = S(0.5) * (-(qx * wx) - qy * wy - qz * wz);
    xd[1] = S(0.5) * (qw * wx - qz * wy + qy * wz);
    xd[2] = S(0.5) * (qz * wx + qw * wy - qx * wz);
    xd[3] = S(0.5) * (qx * wy - qy * wx + qw * wz);
    const S J0 = p[0], J1 = p[1], J2 = p[2];
    T hx = J0 * wx, hy = J1 * wy, hz = J2 * wz;
    xd[4] = (u[0] - (wy * hz - wz * hy)) * (S(1) / J0);
    xd[5] = (u[1] - (wz * hx - wx * hz)) * (S(1) / J1);
    xd[6] = (u[2] - (wx * hy - wy * hx)) * (S(1) / J2);
  }
};

// Cart-pole (diffmpc/dynamics/cartpole_dynamics.py:36-80).
// Parameters: masscart, masspole, length, gravity.
struct CartpoleModel {
  static constexpr int NX = 4, NU = 1, NPARAM = 4;
  template <class T, class PA>
  __device__ static void f(const T* x, const T* u, PA p, T* xd) {
    using S = scalar_t<T>;
    const S mc = p[0], mp = p[1], len = p[2], g = p[3];
    const S total = mc + mp, pml = mp * len;
    T s = dsin(x[2]), c = dcos(x[2]);
    T temp = (u[0] + pml * (x[3] * x[3]) * s) * (S(1) / total);
    T th_acc = (g * s - c * temp) / (len * (sc<T>(S(4) / S(3)) - (mp / total) * (c * c)));
    T x_acc = temp - (pml / total) * (th_acc * c);
    xd[0] = x[1];
    xd[1] = x_acc;
    xd[2] = x[3];
    xd[3] = th_acc;
  }
};

// Planar chain of 12 unit masses hanging from a wall; link j (spring k_j,
// damper c_j, zero rest length) joins mass j-1 and j; u_j pushes mass j
// along x.  State [p_1x, p_1y, ..., p_12y, v_1x, ..., v_12y].
struct ChainModel {
  static constexpr int M = 12;
  static constexpr int NX = 4 * M, NU = M, NPARAM = 2 * M;
  static constexpr bool kTimeInvariant = true;  // linear, as LinearModel
  template <class T, class PA>
  __device__ static void f(const T* x, const T* u, PA p, T* xd) {
#pragma unroll
    for (int i = 0; i < 2 * M; ++i) xd[i] = x[2 * M + i];
#pragma unroll
    for (int j = 0; j < M; ++j) {
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        T dp = x[2 * j + d] - (j > 0 ? x[2 * (j - 1) + d] : sc<T>(0));
        T dv = x[2 * M + 2 * j + d] - (j > 0 ? x[2 * M + 2 * (j - 1) + d] : sc<T>(0));
        using S = scalar_t<T>;
        T fj = -(S(p[j]) * dp) - S(p[M + j]) * dv;
        T acc = fj;
        if (j + 1 < M) {
          T dp1 = x[2 * (j + 1) + d] - x[2 * j + d];
          T dv1 = x[2 * M + 2 * (j + 1) + d] - x[2 * M + 2 * j + d];
          acc = acc + S(p[j + 1]) * dp1 + S(p[M + j + 1]) * dv1;
        }
        if (d == 0) acc = acc + u[j];
        xd[2 * M + 2 * j + d] = acc;
      }
    }
  }
};

// Models whose linearization is the same at every knot declare
// kTimeInvariant = true; the others are treated as time-varying.
template <class M, class = void>
struct time_invariant : std::false_type {};
template <class M>
struct time_invariant<M, std::void_t<decltype(M::kTimeInvariant)>> : std::integral_constant<bool, M::kTimeInvariant> {};

// ---------------------------------------------------------------------------
// One integration step (diffmpc/dynamics/integrators.py:20-71):
// scheme 0 = Euler, 1 = midpoint, 2 = RK4.
template <class Model, class T, class PA>
__device__ void step(const T* x, const T* u, PA p, scalar_t<T> dt, int scheme, T* out) {
  using S = scalar_t<T>;
  constexpr int NX = Model::NX;
  T k1[NX], xt[NX];
  Model::f(x, u, p, k1);
  if (scheme == 0) {
#pragma unroll
    for (int i = 0; i < NX; ++i) out[i] = x[i] + dt * k1[i];
    return;
  }
  if (scheme == 1) {
#pragma unroll
    for (int i = 0; i < NX; ++i) xt[i] = x[i] + (S(0.5) * dt) * k1[i];
    T k2[NX];
    Model::f(xt, u, p, k2);
#pragma unroll
    for (int i = 0; i < NX; ++i) out[i] = x[i] + dt * k2[i];
    return;
  }
  T acc[NX];
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    acc[i] = k1[i];
    xt[i] = x[i] + (S(0.5) * dt) * k1[i];
  }
  T k[NX];
  Model::f(xt, u, p, k);
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    acc[i] = acc[i] + S(2) * k[i];
    xt[i] = x[i] + (S(0.5) * dt) * k[i];
  }
  Model::f(xt, u, p, k);
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    acc[i] = acc[i] + S(2) * k[i];
    xt[i] = x[i] + dt * k[i];
  }
  Model::f(xt, u, p, k);
#pragma unroll
  for (int i = 0; i < NX; ++i) out[i] = x[i] + (acc[i] + k[i]) * (dt / S(6));
}

// f_d(x_t, u_t) of instance b in float64 at the float64 iterate (trajectories
// of N + 1 knots per instance), written to out[0..NX).
template <class Model>
__device__ void fd_value(const double* x, const double* u, const float* p, double dt, int scheme, int N, int b, int t,
                         double* out) {
  constexpr int NX = Model::NX, NU = Model::NU;
  double xs[NX], us[NU], f[NX];
#pragma unroll
  for (int i = 0; i < NX; ++i) xs[i] = x[((size_t)b * (N + 1) + t) * NX + i];
#pragma unroll
  for (int k = 0; k < NU; ++k) us[k] = u[((size_t)b * (N + 1) + t) * NU + k];
  step<Model, double>(xs, us, p, dt, scheme, f);
#pragma unroll
  for (int i = 0; i < NX; ++i) out[i] = f[i];
}

// Value and column j of the Jacobian d f_d / d[x; u] at (x, u) in float.
// Mirrors value_and_jacfwd over next_state (optimal_control_problem.py:360-383)
// with one forward-mode tangent e_j.
template <class Model>
__device__ void step_jacobian_column(const float* x, const float* u, const float* p, float dt, int scheme, int j,
                                     float* fval, float* col) {
  constexpr int NX = Model::NX, NU = Model::NU;
  Dual xd[NX], ud[NU], out[NX];
#pragma unroll
  for (int i = 0; i < NX; ++i) xd[i] = Dual{x[i], (i == j) ? 1.0f : 0.0f};
#pragma unroll
  for (int k = 0; k < NU; ++k) ud[k] = Dual{u[k], (NX + k == j) ? 1.0f : 0.0f};
  step<Model, Dual>(xd, ud, p, dt, scheme, out);
#pragma unroll
  for (int i = 0; i < NX; ++i) {
    fval[i] = out[i].v;
    col[i] = out[i].d;
  }
}

}  // namespace dmpc

// A custom build compiles this library once more with -DDMPC_CUSTOM_MODEL_FILE
// naming a header that defines dmpc::CustomModel (NX, NU, NPARAM and f, in
// the form of the models above); the dispatchers then take MODEL_CUSTOM.
#ifdef DMPC_CUSTOM_MODEL_FILE
namespace dmpc {
#include DMPC_CUSTOM_MODEL_FILE
// block-group width of the custom model in the fused kernels
constexpr int kCustomG = CustomModel::NX <= 4 ? 4 : CustomModel::NX <= 8 ? 8 : CustomModel::NX <= 16 ? 16 : 64;
}  // namespace dmpc
#define DMPC_CUSTOM_CASE(...) \
  case MODEL_CUSTOM:          \
    __VA_ARGS__;
#else
#define DMPC_CUSTOM_CASE(...)
#endif
