// MI355X (gfx950) kernels for the barotropic RK step on the cubed-sphere
// duo-grid, behind the C-ABI declared in include/licom_sw.h.
//
// Layout: every field is (6, N, N), N = nx + 6, C-contiguous (j fastest);
// halo width 3; interior i, j in [3, N-3).  All six faces live on one GPU, so
// the reference's cross-face exchanges (all_gather + routing transforms,
// kernel/communication.py:355-570) become plain loads from the neighbouring
// face's memory.
//
// One RK stage (momentum/barotr_pure.py:100-168) runs as two launches:
//   k_fused         new height on the interior (celerity, C-grid velocity,
//                   LMARS correction, upwind flux, divergence;
//                   barotr_pure.py:111-120) and new covariant velocity on the
//                   deep cells, whose pressure-gradient stencil stays inside
//                   the face (D-grid upwind, vorticity / kinetic-energy
//                   advection, LMARS dissipation, pressure gradient, Coriolis;
//                   barotr_pure.py:133-150), one column march for both.
//   k_frame_edge    one block per segment of a cube edge: velocity on the
//                   3-cell frame of both faces along the segment, the vector
//                   halo fill (cube.py:137-153) and the scalar fill of the
//                   new height (cube.py:130-134) on their halo cells, and the
//                   next stage's cube-edge pass: C-grid normal velocity
//                   averaged over the two faces (remap.py:107 ->
//                   communication.py:485-564) and, for a last stage, the
//                   averaged mass flux (barotr_pure.py:116-117).
// Batched runs (the *_batch kernels) use k_frame_edge_batch when every member's
// context uses the cube-edge form.  With the frame form 0 (sw_set_frame_form),
// k_frame_edge is replaced by two launches:
//   k_frame         velocity on the 3-cell frame along the face edges, with
//                   the new height's duo-grid halo values formed from the halo
//                   plan, and the halo fill of the new height itself
//                   (cube.py:130-134) in the same launch.
//   k_halo_uv_edge  vector halo fill of the new velocity (cube.py:137-153) and
//                   the next stage's cube-edge pass: C-grid normal velocity
//                   averaged over the two faces of each cube edge (remap.py:107
//                   -> communication.py:485-564) and, for a last stage, the
//                   averaged mass flux (barotr_pure.py:116-117).
// The first stage of a call runs its edge pass as a launch of its own
// (k_edge).  With the fused form switched off (band height rows_f = 0) a
// stage runs k_march_a (height), k_halo_scalar, k_march_b (velocity on every
// interior cell) and k_halo_uv_edge instead.
// The ten diagnostic outputs of the last stage are written once per call by
// k_diag1..3 over the full padded arrays, with the reference's values on the
// halo ring.  The stage kernels march wavefronts down columns with the
// stencil windows in registers and one LDS row buffer per wave; every
// multiply and add is evaluated in the reference's order and the library is
// compiled with -ffp-contract=off, so the binary64 path reproduces the NumPy
// restatement of the reference exactly.  *_batch variants advance several
// contexts per launch.

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <type_traits>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "licom_sw.h"

namespace {

// ---------------------------------------------------------------------------
// error plumbing
// ---------------------------------------------------------------------------
thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define SW_HIP(call)                                                           \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail(SW_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------------------
// topology: directions E=0, W=1, N=2, S=3 (communication.py:32)
// ---------------------------------------------------------------------------
constexpr int DE = 0, DW = 1, DN = 2, DS = 3;
constexpr int HALO = 3;

struct Route {
  int r[6][4][3];  // (neighbour face, neighbour edge, 1 if the joint flips)
};

// Face joints of the cube: even faces attach S->(t+4):E, E->(t+2):S,
// N->(t+1):S, W->(t+5):E; odd faces E->(t+1):W, N->(t+2):W, W->(t+4):N,
// S->(t+5):N (communication.py:57-87).  A joint between parallel edges keeps
// the along-edge order; a joint between perpendicular edges reverses it
// (communication.py:40-55).
Route build_route() {
  Route R;
  std::memset(&R, 0, sizeof(R));
  for (int t = 0; t < 6; ++t) {
    int nb[4], ed[4];
    if (t % 2 == 0) {
      nb[DS] = (t + 4) % 6; ed[DS] = DE;
      nb[DE] = (t + 2) % 6; ed[DE] = DS;
      nb[DN] = (t + 1) % 6; ed[DN] = DS;
      nb[DW] = (t + 5) % 6; ed[DW] = DE;
    } else {
      nb[DE] = (t + 1) % 6; ed[DE] = DW;
      nb[DN] = (t + 2) % 6; ed[DN] = DW;
      nb[DW] = (t + 4) % 6; ed[DW] = DN;
      nb[DS] = (t + 5) % 6; ed[DS] = DN;
    }
    for (int d = 0; d < 4; ++d) {
      const bool a_ew = (d == DE || d == DW);
      const bool b_ew = (ed[d] == DE || ed[d] == DW);
      const int flip = (a_ew != b_ew) ? 1 : 0;
      // both directions of a joint are recorded, as the reference does
      R.r[t][d][0] = nb[d];
      R.r[t][d][1] = ed[d];
      R.r[t][d][2] = flip;
      R.r[nb[d]][ed[d]][0] = t;
      R.r[nb[d]][ed[d]][1] = d;
      R.r[nb[d]][ed[d]][2] = flip;
    }
  }
  return R;
}

const Route& route() {
  static const Route R = build_route();
  return R;
}

// ---------------------------------------------------------------------------
// stencil coefficients (operators/poly.py:29-37)
// ---------------------------------------------------------------------------
template <typename T>
struct PC {
  static constexpr T EP23 = T(-1.0 / 20.0);
  static constexpr T EP13 = T(9.0 / 20.0);
  static constexpr T EP03 = T(47.0 / 60.0);
  static constexpr T EM13 = T(-13.0 / 60.0);
  static constexpr T EM23 = T(1.0 / 30.0);
  static constexpr T M13 = T(37.0 / 60.0);
  static constexpr T M23 = T(-2.0 / 15.0);
  static constexpr T M33 = T(1.0 / 60.0);
};

// East/north 5-point face value from samples at k-2..k+2 (poly.py:50-56).
template <typename T>
__device__ __forceinline__ T e5(T a, T b, T c, T d, T e) {
  return PC<T>::EM23 * a + PC<T>::EM13 * b + PC<T>::EP03 * c + PC<T>::EP13 * d + PC<T>::EP23 * e;
}
// West/south 5-point face value (poly.py:57-63).
template <typename T>
__device__ __forceinline__ T w5(T a, T b, T c, T d, T e) {
  return PC<T>::EP23 * a + PC<T>::EP13 * b + PC<T>::EP03 * c + PC<T>::EM13 * d + PC<T>::EM23 * e;
}
// 6-point centred value from samples at k-3..k+2 (poly.py:120-124).
template <typename T>
__device__ __forceinline__ T c6(T m3, T m2, T m1, T p0, T p1, T p2) {
  return PC<T>::M13 * (m1 + p0) + PC<T>::M23 * (m2 + p1) + PC<T>::M33 * (m3 + p2);
}
// jnp.maximum(x, 10.0): NaN propagates.
template <typename T>
__device__ __forceinline__ T max10(T x) {
  return (x >= T(10.0) || x != x) ? x : T(10.0);
}
template <typename T>
__device__ __forceinline__ T celerity(T centred, T dz) {
  return max10<T>(sqrt(T(9.8) * (centred + dz)));
}
// FP64 square root as the general sequence evaluates it (reciprocal square
// root estimate, two Goldschmidt steps, two Newton corrections, +-0 / +inf
// passed through) without its input scaling, which is the identity for
// x >= 2^-767: the same bits there.  Below, the unscaled steps give a finite
// value under 10 for x in (0, 2^-767) and NaN for x < 0, the same results
// after max10 as the general sequence's.
__device__ __forceinline__ double sqrt_f64(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = y * 0.5;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  return __builtin_amdgcn_class(x, 0x260) ? x : g;  // 0x260: +-0, +inf
}
template <>
__device__ __forceinline__ double celerity<double>(double centred, double dz) {
  return max10<double>(sqrt_f64(9.8 * (centred + dz)));
}
// 4.9 / c of the LMARS correction (barotr_pure.py:33).
template <typename T>
__device__ __forceinline__ T k49_over(T c) {
  return T(4.9) / c;
}
// Contravariant component (ga u + gb v) sina (remap.py:35-39).
template <typename T>
__device__ __forceinline__ T contra(T ga, T gb, T u, T v, T sina) {
  return (ga * u + gb * v) * sina;
}

// FP64 quotient with the operand scaling of the general sequence left out
// where it is the identity, so the result is the same correctly rounded bits
// with fewer instructions.  a / b: reciprocal estimate, two Newton steps,
// quotient and one residual correction, then the special-value fixup.  The compiler's sequence scales
// both operands (v_div_scale) only when the exponents are far apart, b is
// denormal or a / b or 1 / b would be: never for a = 4.9 and b in [10, 2^1021]
// or b = +inf / NaN, the divisors of k49_over.
__device__ __forceinline__ double div_f64(double a, double b) {
  const double y0 = __builtin_amdgcn_rcp(b);
  double e = __builtin_fma(-b, y0, 1.0);
  double y = __builtin_fma(y0, e, y0);
  e = __builtin_fma(-b, y, 1.0);
  y = __builtin_fma(y, e, y);
  const double q = a * y;
  const double r = __builtin_fma(-b, q, a);
  return __builtin_amdgcn_div_fixup(__builtin_fma(r, y, q), b, a);
}
// c >= 10 (or NaN) and c <= sqrt(9.8 DBL_MAX) < 2^1021, or +inf
template <>
__device__ __forceinline__ double k49_over<double>(double c) {
  return div_f64(4.9, c);
}

// The FP32 form is measured against the FP64 path (relative error), not bit
// for bit, so it evaluates the same sums as fused multiply-adds in the same
// association and takes the celerity's square root and reciprocal from the
// hardware (v_sqrt_f32, v_rcp_f32, 1 ulp).  Every FP32 kernel uses these, so
// the fused and separate forms and batched runs still agree bit for bit.
template <>
__device__ __forceinline__ float e5<float>(float a, float b, float c, float d, float e) {
  return fmaf(PC<float>::EP23, e, fmaf(PC<float>::EP13, d, fmaf(PC<float>::EP03, c, fmaf(PC<float>::EM13, b, PC<float>::EM23 * a))));
}
template <>
__device__ __forceinline__ float w5<float>(float a, float b, float c, float d, float e) {
  return fmaf(PC<float>::EM23, e, fmaf(PC<float>::EM13, d, fmaf(PC<float>::EP03, c, fmaf(PC<float>::EP13, b, PC<float>::EP23 * a))));
}
template <>
__device__ __forceinline__ float c6<float>(float m3, float m2, float m1, float p0, float p1, float p2) {
  return fmaf(PC<float>::M33, m3 + p2, fmaf(PC<float>::M23, m2 + p1, PC<float>::M13 * (m1 + p0)));
}
template <>
__device__ __forceinline__ float celerity<float>(float centred, float dz) {
  return max10<float>(__builtin_amdgcn_sqrtf(9.8f * (centred + dz)));
}
template <>
__device__ __forceinline__ float k49_over<float>(float c) {
  return 4.9f * __builtin_amdgcn_rcpf(c);
}
template <>
__device__ __forceinline__ float contra<float>(float ga, float gb, float u, float v, float sina) {
  return fmaf(ga, u, gb * v) * sina;
}

// ---------------------------------------------------------------------------
// context-owned const planes (each (6, N, N))
// ---------------------------------------------------------------------------
template <typename T>
struct Planes {
  const T *dzx, *dzy, *rdx, *rdy, *af, *rda, *cdy, *ddx, *sina;
  const T *g00, *g01, *g10, *g11;
  const T *c00, *c01, *c10, *c11;
  const T *l00, *l01, *l10, *l11;
  const T *k0, *k1;
  const int *li, *lj;
  T dzx_s, dzy_s;
  int dz_uniform;
  // interleaved copies read by the stage kernels as one 2-element load per
  // point: (g00, g01), (rdx, rdy), (rda, ddx)
  const T *gp, *rxy, *rdd;
  // the constant planes of the W and E face-edge strips (the 9 columns next
  // to each edge) stored row by row, for the cube-edge frame kernel: element
  // ((((t * 2 + side) * kStripPlanes + k) * N + i) * 9 + c) is plane k at row
  // i, column c (W) or N - 9 + c (E) of face t
  const T* strip;
};

// Planes of the W / E strip copy.
constexpr int kStripPlanes = 11;
enum { SP_G00, SP_G01, SP_G10, SP_G11, SP_SINA, SP_CDY, SP_RDA, SP_DDX, SP_RDX, SP_RDY, SP_AF };


// Two consecutive elements of an interleaved pair plane: one 16-byte (FP64)
// or 8-byte (FP32) load per lane.
template <typename T>
using Pair = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
template <typename T>
__device__ __forceinline__ Pair<T> ldg2(const T* base, uint32_t i) {
  return *reinterpret_cast<const Pair<T>*>(reinterpret_cast<const char*>(base) +
                                           i * static_cast<uint32_t>(2 * sizeof(T)));
}

struct Geo {
  int nx, N;
  __host__ __device__ size_t at(int t, int i, int j) const {
    return (static_cast<size_t>(t) * N + i) * N + j;
  }
};

template <typename T>
__device__ __forceinline__ T dzx(const Planes<T>& P, size_t o) {
  return P.dz_uniform ? P.dzx_s : P.dzx[o];
}
template <typename T>
__device__ __forceinline__ T dzy(const Planes<T>& P, size_t o) {
  return P.dz_uniform ? P.dzy_s : P.dzy[o];
}
template <typename T>
__device__ __forceinline__ T uct_at(const Planes<T>& P, const T* u, const T* v, size_t o) {
  return contra(P.g00[o], P.g01[o], u[o], v[o], P.sina[o]);
}
template <typename T>
__device__ __forceinline__ T vct_at(const Planes<T>& P, const T* u, const T* v, size_t o) {
  return contra(P.g10[o], P.g11[o], u[o], v[o], P.sina[o]);
}

// ---------------------------------------------------------------------------
// edge lines
// ---------------------------------------------------------------------------
// Halo plan (built by k_plan_build): entry q holds the halo cell's offset,
// its two interpolation sources and weights and the rotation coefficients a
// vector fill reads.
constexpr int kPlanRot = 12;

template <typename T>
struct HaloPlan {
  const uint32_t *dst, *s0, *s1;
  const T *c0, *c1;
  const T* rot;  // kPlanRot x n: c2l(s0) 00 01 10 11, c2l(s1) 00 01 10 11, l2c(dst) 00 01 10 11
  int n;
};

// Plan entry of the strip halo cell (row, col) of face t (not a corner): per
// face the W/E strips (entries (d, r, layer)), then the N/S strips ((d,
// layer, r)), then the corners, as k_plan_build numbers them.
__device__ __forceinline__ int halo_entry(int nx, int N, int t, int row, int col) {
  const int e = col >= N - HALO ? 3 * (row - HALO) + (col - (N - HALO))
              : col < HALO      ? 3 * nx + 3 * (row - HALO) + (HALO - 1 - col)
              : row >= N - HALO ? 6 * nx + (row - (N - HALO)) * nx + (col - HALO)
                                : 9 * nx + (HALO - 1 - row) * nx + (col - HALO);
  return t * (12 * nx + 36) + e;
}

// The vector fill's value of a halo cell from the values (u0, v0), (u1, v1)
// at its two sources, R its plan entry's rotation coefficients: remap to
// lat-lon at the sources, interpolation, remap back (cube.py:137-153).
template <typename T>
__device__ __forceinline__ void halo_uv_combine(const T* R, T c0, T c1, T u0, T v0, T u1, T v1, T& uo, T& vo) {
  const T ul0 = R[0] * u0 + R[1] * v0;
  const T vl0 = R[2] * u0 + R[3] * v0;
  const T ul1 = R[4] * u1 + R[5] * v1;
  const T vl1 = R[6] * u1 + R[7] * v1;
  const T p0 = c0 * ul0 + c1 * ul1;
  const T p1 = c0 * vl0 + c1 * vl1;
  uo = R[8] * p0 + R[9] * p1;
  vo = R[10] * p0 + R[11] * p1;
}

// The same for plan entry q, its sources read from u, v (k_halo_vector).
template <typename T>
__device__ __forceinline__ void halo_uv_value(const HaloPlan<T>& plan, const T* u, const T* v, int q, T& uo, T& vo) {
  const int n = plan.n;
  const uint32_t s0 = plan.s0[q], s1 = plan.s1[q];
  const T c0 = plan.c0[q], c1 = plan.c1[q];
  T R[kPlanRot];
#pragma unroll
  for (int k = 0; k < kPlanRot; ++k) R[k] = plan.rot[static_cast<size_t>(k) * n + q];
  halo_uv_combine(R, c0, c1, u[s0], v[s0], u[s1], v[s1], uo, vo);
}

template <typename T>
struct EdgeArgs {
  Geo g;
  Route R;
  Planes<T> P;
  const T *h, *u, *v;
  T *evel, *eflux;
  int last;
  int ue[12];  // the 12 cube edges, each once: t * 4 + d with (t, d) < its neighbour
  // fly: u, v on halo cells are formed from the plan instead of read, so the
  // pass can share a launch with the vector halo fill of u, v
  int fly;
  HaloPlan<T> plan;
};

// Edge values of both faces meeting at a cube edge (communication.py:485-564
// averaging of the C-grid normal velocity and, on the last stage, of the
// upwind flux).  A group of 8 lanes takes one side of one edge point: lane k
// (k < 6) evaluates stencil point k across the edge, and the group gathers
// the six values with shuffles.  Groups 2m and 2m+1 are the two sides of one
// edge point and swap their one-sided values, so every cube edge is
// evaluated once and both faces' entries of evel / eflux (index
// (t * 4 + d) * nx + r) are written.  IEEE addition is commutative, so
// (own + other) / 2 equals the reference's (mine + nb) / 2 seen from either
// face.
constexpr int kEdgeLanes = 8;

template <typename T>
__device__ __forceinline__ void edge_body(const EdgeArgs<T>& a, int q) {
  const int nx = a.g.nx, N = a.g.N;
  const int unit = q / kEdgeLanes, k = q % kEdgeLanes;
  if (unit >= 24 * nx) return;  // 24 nx is even: groups and group pairs stay whole
  const int p = unit >> 1, side = unit & 1;
  const int e = p / nx, r = p - e * nx;
  const int t = a.ue[e] >> 2, d = a.ue[e] & 3;
  const int tb = a.R.r[t][d][0], db = a.R.r[t][d][1];
  const int rr = a.R.r[t][d][2] ? nx - 1 - r : r;
  const int ts = side ? tb : t, ds = side ? db : d, rs = side ? rr : r;
  const bool ns = ds == DN || ds == DS;
  const int kk = min(k, 5);
  // stencil point kk across the edge line (remap.py:42-55 at rows / columns 3, N-3)
  const int line = (ds == DS || ds == DW) ? 3 : N - 3;
  const size_t o = ns ? a.g.at(ts, line - 3 + kk, 3 + rs) : a.g.at(ts, 3 + rs, line - 3 + kk);
  const size_t oc = ns ? a.g.at(ts, line, 3 + rs) : a.g.at(ts, 3 + rs, line);
  const Planes<T>& P = a.P;
  T val;
  const int prow = ns ? line - 3 + kk : 3 + rs, pcol = ns ? 3 + rs : line - 3 + kk;
  if (a.fly && (ns ? (prow < HALO || prow >= N - HALO) : (pcol < HALO || pcol >= N - HALO))) {
    T uh, vh;
    halo_uv_value(a.plan, a.u, a.v, halo_entry(nx, N, ts, prow, pcol), uh, vh);
    val = ns ? contra(P.g00[o], P.g01[o], uh, vh, P.sina[o]) : contra(P.g10[o], P.g11[o], uh, vh, P.sina[o]);
  } else {
    val = ns ? uct_at(P, a.u, a.v, o) : vct_at(P, a.u, a.v, o);
  }
  const int base = (threadIdx.x & 63) & ~(kEdgeLanes - 1);
  T w[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) w[m] = __shfl(val, base + m);
  const T vel = T(0.5) * (e5(w[0], w[1], w[2], w[3], w[4]) + w5(w[1], w[2], w[3], w[4], w[5]));
  const T avg = (vel + __shfl_xor(vel, kEdgeLanes)) / T(2.0);
  const int oe = (ts * 4 + ds) * nx + rs;
  if (k == 0) a.evel[oe] = avg;
  if (a.last) {  // upwind mass flux through the edge (barotr_pure.py:29-49)
    const T hv = a.h[o];
    T s6[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) s6[m] = __shfl(hv, base + m);
    const T dz = ns ? dzx(P, oc) : dzy(P, oc);
    const T he = e5(s6[0], s6[1], s6[2], s6[3], s6[4]);
    const T hw = w5(s6[1], s6[2], s6[3], s6[4], s6[5]);
    const T c = celerity(c6(s6[0], s6[1], s6[2], s6[3], s6[4], s6[5]), dz);
    const T vc = avg + k49_over(c) * (he - hw);
    const T f = dz * vc + (vc > T(0) ? he : hw) * vc;
    const T fo = __shfl_xor(f, kEdgeLanes);
    if (k == 0) a.eflux[oe] = (f + fo) / T(2.0);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_edge(EdgeArgs<T> a) {
  edge_body(a, static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x));
}

// ---------------------------------------------------------------------------
// stage kernels
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Column-marching stage kernels.
//
// One wavefront owns a strip of 64 consecutive columns j = c0 + lane of one
// face and marches down a band of R interior rows.  Values along i (the
// stencil direction of the i-faces) live in per-lane register windows that
// slide one row per step; values along j come from a 69-entry LDS row buffer
// shared by the wave (its own 64 columns plus 3 columns on the left and 2 on
// the right, loaded by lanes 0..4).  A wave computes j-faces at its 64
// columns and writes cells c0 .. c0+62, so strips start every 63 columns.
// The next row's loads are issued one step ahead of their use.  Waves are
// independent (no workgroup barrier).
//
// Tile order.  The dispatcher hands workgroup b to XCD b % 8, and each XCD has
// its own L2.  decode_wave turns the hardware index into a logical one that is
// contiguous per XCD, and logical wave ids run face fastest, then strip, then
// band: the six faces of one (strip, band) - which read the same
// face-independent metric planes - and the neighbouring strips of a band -
// which share the 6 overlap columns and the 128-byte lines that straddle strip
// boundaries - are resident on one XCD at the same time and fetch those lines
// from HBM once.
// ---------------------------------------------------------------------------
constexpr int kXcds = 8;
constexpr int kStripOut = 63;  // cells written per wave and row
constexpr int kRowBuf = 72;    // LDS row buffer length (69 used)

template <typename T>
struct MarchArgs {
  Geo g;
  Planes<T> P;
  const T *h, *hp, *u, *v, *up, *vp;  // stage inputs
  const T *hn;                          // new height, halo filled (kernel B)
  const T *evel, *eflux;
  T *h_out, *u_out, *v_out;
  T dt, beta;
  int last;
  int S, R, nbands, nwaves;
  // fused kernel, split bands (R2 > 0): the first `nlong` hardware blocks
  // (multiple of 8, at most one per CU, so their waves are the oldest on
  // their SIMDs) march `nwlong` bands of R rows, the rest `nwaves` bands of
  // R2 rows; columns alternate R and R2 bands.  R2 = 0: uniform R bands.
  int R2, nlong, nwlong;
  int shared;  // rdx, rdy, rda, c_dy, d_dx identical on every face
  int gsym;    // a_gct symmetric with equal diagonal
  T* sink;     // 256 elements absorbing the stores of inactive lanes
  // null, or 2 wall-clock stamps (entry, exit) per wave of the launch, written
  // by the fused kernel's lane 0 (launch-duration measurement in graphs)
  unsigned long long* stamps;
};

// Lane k <- lane k - 1 across the whole wave (DPP wave_shr:1); lane 0
// receives 0.
__device__ __forceinline__ int lane_prev_b32(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xF, 0xF, false); }
__device__ __forceinline__ double lane_prev(double x) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
  const unsigned lo = static_cast<unsigned>(lane_prev_b32(static_cast<int>(b)));
  const unsigned hi = static_cast<unsigned>(lane_prev_b32(static_cast<int>(b >> 32)));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}
__device__ __forceinline__ float lane_prev(float x) { return __int_as_float(lane_prev_b32(__float_as_int(x))); }

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Element i of a field addressed through a 32-bit byte offset from a uniform
// base, so loads and stores use the scalar-base + 32-bit vector-offset form.
template <typename T>
__device__ __forceinline__ T ldg(const T* base, uint32_t i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * static_cast<uint32_t>(sizeof(T)));
}
template <typename T>
__device__ __forceinline__ T* at32(T* base, uint32_t i) {
  return reinterpret_cast<T*>(reinterpret_cast<char*>(base) + i * static_cast<uint32_t>(sizeof(T)));
}

struct WaveTile {
  int t, lane, wslot, c0, j, jl, r0, r1;
  bool up;  // fused march direction (alternates between neighbouring bands)
  uint32_t fo, mo;  // element offsets (fields stay below 2^29 elements)
};

// Bijection blockIdx -> logical workgroup index, contiguous per XCD.
__device__ __forceinline__ int xcd_block() {
  const int G = static_cast<int>(gridDim.x), b = static_cast<int>(blockIdx.x);
  const int x = b % kXcds, q = G / kXcds, r = G % kXcds;
  return x * q + min(x, r) + b / kXcds;
}

// Wave of logical block L (four waves per block) -> its (face, strip, band).
template <typename T>
__device__ __forceinline__ bool decode_wave(const MarchArgs<T>& a, int L, WaveTile& w) {
  const int N = a.g.N;
  w.lane = threadIdx.x & 63;
  w.wslot = threadIdx.x >> 6;
  const int wid = L * 4 + w.wslot;
  if (wid >= a.nwaves) return false;
  w.t = wid % 6;
  const int rest = wid / 6;
  const int s = rest % a.S;
  const int band = rest / a.S;
  w.c0 = HALO + kStripOut * s;
  w.j = w.c0 + w.lane;
  w.jl = min(w.j, N - 1);
  w.r0 = HALO + band * a.R;
  w.r1 = min(w.r0 + a.R, N - HALO);
  const uint32_t NN = static_cast<uint32_t>(N) * N;
  w.fo = static_cast<uint32_t>(w.t) * NN;
  w.mo = a.shared ? 0 : w.fo;
  return true;
}

// Raw per-point inputs of the contravariant velocity (remap.py:35-39).
template <typename T>
struct CtRaw {
  T u, v, g00, g01, g10, g11, sina;
};

template <typename T>
__device__ __forceinline__ T ct_u(const CtRaw<T>& r) { return contra(r.g00, r.g01, r.u, r.v, r.sina); }
template <typename T>
__device__ __forceinline__ T ct_v(const CtRaw<T>& r) { return contra(r.g10, r.g11, r.u, r.v, r.sina); }

// Kernel A: new height on the interior (barotr_pure.py:111-120).
//
// Every memory operation in the row loop is unconditional (addresses of
// inactive lanes are clamped, their stores go to a per-wave sink), so the
// compiler can count outstanding loads exactly and the loads issued one row
// ahead stay in flight across the row's arithmetic.
template <typename T>
struct APre {
  CtRaw<T> cr;   // column point of row i+2
  T h;
  T cdy, dzx;    // i-face row i
  T evi, efi;    // edge velocity / flux of the i-face (valid on cube edges)
  T rda, ddx, hp, dzy;  // cell row i-1
  T evj, efj;    // edge velocity / flux of the j-face at (i-1, j)
  CtRaw<T> hr;   // row-buffer column outside the strip, row i-1
  T hh;
};

template <typename T, bool GSYM>
__device__ __forceinline__ void load_ct2(const MarchArgs<T>& a, uint32_t o, CtRaw<T>& r) {
  const Planes<T>& P = a.P;
  r.u = ldg(a.u, o);
  r.v = ldg(a.v, o);
  r.sina = ldg(P.sina, o);
  if (GSYM) {
    const Pair<T> g = ldg2(P.gp, o);
    r.g00 = g.x;
    r.g01 = g.y;
    r.g10 = r.g01;
    r.g11 = r.g00;
  } else {
    r.g00 = ldg(P.g00, o);
    r.g01 = ldg(P.g01, o);
    r.g10 = ldg(P.g10, o);
    r.g11 = ldg(P.g11, o);
  }
}

// Inputs of one march step: the column point of row crow, the i-face of row
// i, the cell of row ii and the halo column of row hrow.  SAMEP: hp is h
// (stage 1 of a substep that starts from its own output), so hp at row ii is
// already in the lane's h window and is not loaded.  Loads whose values no
// lane of the wave uses are skipped behind wave-uniform tests: the cube-edge
// velocity and flux of the i-face off the edge rows and of the j-face in
// strips without an edge column, the depth where it is uniform (the scalar
// argument holds it).  hl: load the halo column (kernel A; the fused kernel
// has none and passes false, leaving p.hr and p.hh unset).  je: load the
// j-face's cube-edge values (the fused kernel passes false and reads them
// from the band's copy in LDS, leaving p.evj and p.efj zero).
template <typename T, bool GSYM, bool SAMEP>
__device__ __forceinline__ void a_fetch(const MarchArgs<T>& a, const WaveTile& w, int i, int crow, int ii,
                                        int hrow, int hcol, APre<T>& p, bool hl = true, bool je = true) {
  const int N = a.g.N, nx = a.g.nx;
  const Planes<T>& P = a.P;
  const uint32_t NS = static_cast<uint32_t>(N);
  const int jl = w.jl, j = w.j, t = w.t;
  const uint32_t oc = w.fo + static_cast<uint32_t>(crow) * NS + jl;
  load_ct2<T, GSYM>(a, oc, p.cr);
  p.h = ldg(a.h, oc);
  const uint32_t oi = static_cast<uint32_t>(min(i, N - 1)) * NS + jl;
  p.cdy = ldg(P.cdy, w.mo + oi);
  const uint32_t om = static_cast<uint32_t>(ii) * NS + jl;
  if (P.dz_uniform) {
    p.dzx = P.dzx_s;
    p.dzy = P.dzy_s;
  } else {
    p.dzx = ldg(P.dzx, w.fo + oi);
    p.dzy = ldg(P.dzy, w.fo + om);
  }
  p.evi = p.efi = T(0);
  if (i == HALO || i == N - HALO) {
    const bool iedge = j >= HALO && j <= N - 4;
    const int ie = iedge ? (t * 4 + (i == HALO ? DS : DN)) * nx + (j - HALO) : 0;
    p.evi = ldg(a.evel, ie);
    p.efi = ldg(a.eflux, ie);
  }
  const Pair<T> rd = ldg2(P.rdd, w.mo + om);
  p.rda = rd.x;
  p.ddx = rd.y;
  if constexpr (!SAMEP) p.hp = ldg(a.hp, w.fo + om);
  p.evj = p.efj = T(0);
  if (je && ((w.c0 <= HALO && HALO <= w.c0 + 63) || (w.c0 <= N - HALO && N - HALO <= w.c0 + 63))) {
    const bool jedge = (j == HALO || j == N - HALO) && ii >= HALO && ii < N - HALO;
    const int je = jedge ? (t * 4 + (j == HALO ? DW : DE)) * nx + (ii - HALO) : 0;
    p.evj = ldg(a.evel, je);
    p.efj = ldg(a.eflux, je);
  }
  if (hl) {
    const uint32_t oh = w.fo + static_cast<uint32_t>(hrow) * NS + hcol;
    load_ct2<T, GSYM>(a, oh, p.hr);
    p.hh = ldg(a.h, oh);
  }
}

// Kernel A's step at face i (marching down): column point row i+2, cell row i-1.
template <typename T, bool GSYM, bool SAMEP>
__device__ __forceinline__ void a_prefetch(const MarchArgs<T>& a, const WaveTile& w, int i, int hcol,
                                           APre<T>& p) {
  a_fetch<T, GSYM, SAMEP>(a, w, i, min(i + 2, a.g.N - 1), max(i - 1, 0), max(i - 1, 0), hcol, p);
}

// Mass flux through one face (barotr_pure.py:29-49) given the averaged normal
// velocity on cube edges; s[] = h and w[] = contravariant velocity samples
// k-3 .. k+2 across the face.
template <typename T>
__device__ __forceinline__ T face_flux2(const T* s, const T* w, bool edge, T upre_edge, T dz) {
  const T he = e5(s[0], s[1], s[2], s[3], s[4]);
  const T hw = w5(s[1], s[2], s[3], s[4], s[5]);
  const T inner = T(0.5) * (e5(w[0], w[1], w[2], w[3], w[4]) + w5(w[1], w[2], w[3], w[4], w[5]));
  const T upre = edge ? upre_edge : inner;
  const T c = celerity(c6(s[0], s[1], s[2], s[3], s[4], s[5]), dz);
  const T uc = upre + k49_over(c) * (he - hw);
  return dz * uc + (uc > T(0) ? he : hw) * uc;
}

template <typename T, bool GSYM, bool SAMEP>
__device__ __forceinline__ void march_a_body(const MarchArgs<T>& a, int L, T (*s_v)[kRowBuf],
                                             T (*s_h)[kRowBuf]) {
  WaveTile w;
  if (!decode_wave(a, L, w)) return;
  const int N = a.g.N;
  T* bv = s_v[w.wslot];
  T* bh = s_h[w.wslot];
  const int lane = w.lane, j = w.j, c0 = w.c0;
  const uint32_t NS = static_cast<uint32_t>(N);

  int hcol = lane < 3 ? c0 - 3 + lane : c0 + 61 + lane;
  const bool halo_lane = lane < 5;
  hcol = halo_lane ? min(hcol, N - 1) : w.jl;
  const int hslot = halo_lane ? hcol - c0 + 3 : lane + 3;  // others rewrite their own slot
  const bool out_lane = lane < kStripOut && j <= N - 4;
  T* sink = a.sink + threadIdx.x;

  T wu[6], wvc[6], wh[6];
  for (int m = 0; m < 5; ++m) {
    const uint32_t o = w.fo + static_cast<uint32_t>(w.r0 - 3 + m) * NS + w.jl;
    CtRaw<T> r;
    load_ct2<T, GSYM>(a, o, r);
    wu[m] = ct_u(r);
    wvc[m] = ct_v(r);
    wh[m] = ldg(a.h, o);
  }
  APre<T> pre;
  a_prefetch<T, GSYM, SAMEP>(a, w, w.r0, hcol, pre);

  // step 0: the i-face at row r0 only
  T Fprev, cdy_prev;
  {
    const APre<T> cur = pre;
    a_prefetch<T, GSYM, SAMEP>(a, w, w.r0 + 1, hcol, pre);
    wu[5] = ct_u(cur.cr);
    wvc[5] = ct_v(cur.cr);
    wh[5] = cur.h;
    const int i = w.r0;
    const bool edge = (i == HALO) || (i == N - HALO);
    const T F = face_flux2(wh, wu, edge, cur.evi, cur.dzx);
    Fprev = (a.last && edge) ? cur.efi : F;
    cdy_prev = cur.cdy;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      wu[m] = wu[m + 1];
      wvc[m] = wvc[m + 1];
      wh[m] = wh[m + 1];
    }
  }

  for (int i = w.r0 + 1; i <= w.r1; ++i) {
    const APre<T> cur = pre;
    a_prefetch<T, GSYM, SAMEP>(a, w, i + 1, hcol, pre);
    wu[5] = ct_u(cur.cr);
    wvc[5] = ct_v(cur.cr);
    wh[5] = cur.h;

    // mass flux through the i-face at row i
    const bool iedge = (i == HALO) || (i == N - HALO);
    T F = face_flux2(wh, wu, iedge, cur.evi, cur.dzx);
    F = (a.last && iedge) ? cur.efi : F;

    // row i-1 of vct and h through the row buffer
    const int ii = i - 1;
    bv[lane + 3] = wvc[2];
    bh[lane + 3] = wh[2];
    bv[hslot] = ct_v(cur.hr);
    bh[hslot] = cur.hh;
    wave_lds_sync();
    T sv[6], sh[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      sv[m] = bv[lane + m];
      sh[m] = bh[lane + m];
    }
    wave_lds_sync();
    const bool jedge = (j == HALO) || (j == N - HALO);
    T G = face_flux2(sh, sv, jedge, cur.evj, cur.dzy);
    G = (a.last && jedge) ? cur.efj : G;
    const T Gn = __shfl_down(G, 1);
    const T ddxn = __shfl_down(cur.ddx, 1);
    // divergence (agrid.py:44-53) and height update (barotr_pure.py:120)
    const T D = cur.rda * (((F * cur.cdy - Fprev * cdy_prev) + Gn * ddxn) - G * cur.ddx);
    T* dst = out_lane ? at32(a.h_out, w.fo + static_cast<uint32_t>(ii) * NS + j) : sink;
    *dst = (SAMEP ? wh[2] : cur.hp) - D * a.dt;  // wh[2]: h at row ii

    Fprev = F;
    cdy_prev = cur.cdy;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      wu[m] = wu[m + 1];
      wvc[m] = wvc[m + 1];
      wh[m] = wh[m + 1];
    }
  }
}

// Kernel B: new covariant velocity on the interior (barotr_pure.py:133-150).
// The per-row inputs are fetched one row ahead in four groups, each issued
// right after the current values of that group are consumed, so the current
// and the next values of a group are never live together.
template <typename T>
struct BCol {  // column point of row i+2
  T u, v, g00, g01, g10, g11, sina, h, hn;
};
template <typename T>
struct BFace {  // i-face row i
  T cdy, dzx, evi;
};
template <typename T>
struct BHalo {  // row-buffer column outside the strip, row i-1
  T u, v, g10, g11, sina, h, hn;
};
template <typename T>
struct BCell {  // cell row i-1
  T rda, ddx, rdx, rdy, af, up, vp, dzy, evj;
};

template <typename T, bool GSYM>
__device__ __forceinline__ void b_fetch_col(const MarchArgs<T>& a, const WaveTile& w, int i, BCol<T>& p) {
  const int N = a.g.N;
  const Planes<T>& P = a.P;
  const uint32_t o = w.fo + static_cast<uint32_t>(min(i + 2, N - 1)) * N + w.jl;
  p.u = ldg(a.u, o);
  p.v = ldg(a.v, o);
  p.g00 = ldg(P.g00, o);
  p.g01 = ldg(P.g01, o);
  if (GSYM) {
    p.g10 = p.g01;
    p.g11 = p.g00;
  } else {
    p.g10 = ldg(P.g10, o);
    p.g11 = ldg(P.g11, o);
  }
  p.sina = ldg(P.sina, o);
  p.h = ldg(a.h, o);
  p.hn = ldg(a.hn, o);
}
template <typename T>
__device__ __forceinline__ void b_fetch_face(const MarchArgs<T>& a, const WaveTile& w, int i, BFace<T>& p) {
  const int N = a.g.N, nx = a.g.nx;
  const Planes<T>& P = a.P;
  const uint32_t oi = static_cast<uint32_t>(min(i, N - 1)) * N + w.jl;
  p.cdy = ldg(P.cdy, w.mo + oi);
  // loads no lane uses are skipped behind wave-uniform tests (as in a_fetch)
  p.dzx = P.dz_uniform ? P.dzx_s : ldg(P.dzx, w.fo + oi);
  p.evi = T(0);
  if (i == HALO || i == N - HALO) {
    const bool iedge = w.j <= N - 4;
    p.evi = ldg(a.evel, iedge ? (w.t * 4 + (i == HALO ? DS : DN)) * nx + (w.j - HALO) : 0);
  }
}
template <typename T, bool GSYM>
__device__ __forceinline__ void b_fetch_halo(const MarchArgs<T>& a, const WaveTile& w, int ii, int hcol,
                                             BHalo<T>& p) {
  const int N = a.g.N;
  const Planes<T>& P = a.P;
  const uint32_t o = w.fo + static_cast<uint32_t>(min(ii, N - 1)) * N + hcol;
  p.u = ldg(a.u, o);
  p.v = ldg(a.v, o);
  if (GSYM) {
    p.g10 = ldg(P.g01, o);
    p.g11 = ldg(P.g00, o);
  } else {
    p.g10 = ldg(P.g10, o);
    p.g11 = ldg(P.g11, o);
  }
  p.sina = ldg(P.sina, o);
  p.h = ldg(a.h, o);
  p.hn = ldg(a.hn, o);
}
// SAMEP: up, vp are u, v, whose row ii is already in the lane's windows.
template <typename T, bool SAMEP>
__device__ __forceinline__ void b_fetch_cell(const MarchArgs<T>& a, const WaveTile& w, int ii, BCell<T>& p) {
  const int N = a.g.N, nx = a.g.nx;
  const Planes<T>& P = a.P;
  const uint32_t om = static_cast<uint32_t>(min(ii, N - 1)) * N + w.jl;
  p.rda = ldg(P.rda, w.mo + om);
  p.ddx = ldg(P.ddx, w.mo + om);
  p.rdx = ldg(P.rdx, w.mo + om);
  p.rdy = ldg(P.rdy, w.mo + om);
  p.af = ldg(P.af, w.fo + om);
  if constexpr (!SAMEP) {
    p.up = ldg(a.up, w.fo + om);
    p.vp = ldg(a.vp, w.fo + om);
  }
  p.dzy = P.dz_uniform ? P.dzy_s : ldg(P.dzy, w.fo + om);
  p.evj = T(0);
  if ((w.c0 <= HALO && HALO <= w.c0 + 63) || (w.c0 <= N - HALO && N - HALO <= w.c0 + 63)) {
    const bool jedge = (w.j == HALO || w.j == N - HALO) && ii >= HALO && ii < N - HALO;
    p.evj = ldg(a.evel, jedge ? (w.t * 4 + (w.j == HALO ? DW : DE)) * nx + (ii - HALO) : 0);
  }
}

// C-grid quantities on the i-face at row i from the column windows
// (remap.py:42-95, barotr_pure.py:29-34, 63-75).
template <typename T>
__device__ __forceinline__ void iface_b(const T* wu, const T* wv, const T* wh, bool edge, T evi, T dzx,
                                        T& ucx, T& vcx, T& Cx) {
  const T ue = e5(wu[0], wu[1], wu[2], wu[3], wu[4]);
  const T uw = w5(wu[1], wu[2], wu[3], wu[4], wu[5]);
  const T inner = T(0.5) * (ue + uw);
  const T upre = edge ? evi : inner;
  const T ve = e5(wv[0], wv[1], wv[2], wv[3], wv[4]);
  const T vw = w5(wv[1], wv[2], wv[3], wv[4], wv[5]);
  vcx = upre > T(0) ? ve : vw;
  const T he = e5(wh[0], wh[1], wh[2], wh[3], wh[4]);
  const T hw = w5(wh[1], wh[2], wh[3], wh[4], wh[5]);
  const T cx = celerity(c6(wh[0], wh[1], wh[2], wh[3], wh[4], wh[5]), dzx);
  ucx = upre + k49_over(cx) * (he - hw);
  Cx = T(0.5) * cx * (ue - uw);
}

template <typename T, bool GSYM, bool SAMEP>
__device__ __forceinline__ void march_b_body(const MarchArgs<T>& a, int L, T (*s_vct)[kRowBuf],
                                             T (*s_u)[kRowBuf], T (*s_h)[kRowBuf], T (*s_n)[kRowBuf]) {
  WaveTile w;
  if (!decode_wave(a, L, w)) return;
  const int N = a.g.N;
  T* bvct = s_vct[w.wslot];
  T* bu = s_u[w.wslot];
  T* bh = s_h[w.wslot];
  T* bn = s_n[w.wslot];
  const int lane = w.lane, j = w.j, c0 = w.c0;
  const uint32_t NS = static_cast<uint32_t>(N);
  const T wn = T(1.0) - a.beta, wo = a.beta;

  int hcol = lane < 3 ? c0 - 3 + lane : c0 + 61 + lane;
  const bool halo_lane = lane < 5;
  hcol = halo_lane ? min(hcol, N - 1) : w.jl;
  const int hslot = halo_lane ? hcol - c0 + 3 : lane + 3;  // others rewrite their own slot
  const bool out_lane = lane < kStripOut && j <= N - 4;
  T* sink = a.sink + threadIdx.x;

  // column windows over rows i-3 .. i+2 (slot 5 newest); contravariant v and
  // covariant u only at rows i-1 .. i+2; blended new height over i-4 .. i+2
  T wu[6], wv[6], wh[6], wn7[7], wvc[4], wuu[4];
  for (int m = 0; m < 5; ++m) {
    BCol<T> r;
    b_fetch_col<T, GSYM>(a, w, w.r0 - 5 + m, r);
    wu[m] = contra(r.g00, r.g01, r.u, r.v, r.sina);
    wv[m] = r.v;
    wh[m] = r.h;
    wn7[m + 1] = wn * r.hn + wo * r.h;
    if (m >= 2) {
      wvc[m - 2] = contra(r.g10, r.g11, r.u, r.v, r.sina);
      wuu[m - 2] = r.u;
    }
  }
  wn7[0] = T(0);
  BCol<T> pcol;
  BFace<T> pface;
  BHalo<T> phalo;
  BCell<T> pcell;
  b_fetch_col<T, GSYM>(a, w, w.r0, pcol);
  b_fetch_face(a, w, w.r0, pface);

  T ucx_prev, vcx_prev, Cx_prev, cdy_prev;
  {
    const BCol<T> c = pcol;
    b_fetch_col<T, GSYM>(a, w, w.r0 + 1, pcol);
    wu[5] = contra(c.g00, c.g01, c.u, c.v, c.sina);
    wv[5] = c.v;
    wh[5] = c.h;
    wn7[6] = wn * c.hn + wo * c.h;
    wvc[3] = contra(c.g10, c.g11, c.u, c.v, c.sina);
    wuu[3] = c.u;
    const BFace<T> f = pface;
    b_fetch_face(a, w, w.r0 + 1, pface);
    const int i = w.r0;
    iface_b(wu, wv, wh, (i == HALO) || (i == N - HALO), f.evi, f.dzx, ucx_prev, vcx_prev, Cx_prev);
    cdy_prev = f.cdy;
    b_fetch_halo<T, GSYM>(a, w, w.r0, hcol, phalo);
    b_fetch_cell<T, SAMEP>(a, w, w.r0, pcell);
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      wu[m] = wu[m + 1];
      wv[m] = wv[m + 1];
      wh[m] = wh[m + 1];
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      wvc[m] = wvc[m + 1];
      wuu[m] = wuu[m + 1];
    }
#pragma unroll
    for (int m = 0; m < 6; ++m) wn7[m] = wn7[m + 1];
  }

  for (int i = w.r0 + 1; i <= w.r1; ++i) {
    const int ii = i - 1;
    {
      const BCol<T> c = pcol;
      b_fetch_col<T, GSYM>(a, w, i + 1, pcol);
      wu[5] = contra(c.g00, c.g01, c.u, c.v, c.sina);
      wv[5] = c.v;
      wh[5] = c.h;
      wn7[6] = wn * c.hn + wo * c.h;
      wvc[3] = contra(c.g10, c.g11, c.u, c.v, c.sina);
      wuu[3] = c.u;
    }
    T ucx, vcx, Cx, cdy_i;
    {
      const BFace<T> f = pface;
      b_fetch_face(a, w, i + 1, pface);
      iface_b(wu, wv, wh, (i == HALO) || (i == N - HALO), f.evi, f.dzx, ucx, vcx, Cx);
      cdy_i = f.cdy;
    }
    // row ii of vct, u, h and the blended new height through LDS
    {
      const BHalo<T> hh = phalo;
      bvct[lane + 3] = wvc[0];
      bu[lane + 3] = wuu[0];
      bh[lane + 3] = wh[2];
      bn[lane + 3] = wn7[3];
      bvct[hslot] = contra(hh.g10, hh.g11, hh.u, hh.v, hh.sina);
      bu[hslot] = hh.u;
      bh[hslot] = hh.h;
      bn[hslot] = wn * hh.hn + wo * hh.h;
      b_fetch_halo<T, GSYM>(a, w, i, hcol, phalo);
    }
    wave_lds_sync();
    T vcy, ucy, Cy, hy0, hy1;
    const BCell<T> cl = pcell;
    {
      T sv[6], su[6], sh[6];
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        sv[m] = bvct[lane + m];
        su[m] = bu[lane + m];
        sh[m] = bh[lane + m];
      }
      // C-grid quantities on the j-face at column j, row ii
      iface_b(sv, su, sh, (j == HALO) || (j == N - HALO), cl.evj, cl.dzy, vcy, ucy, Cy);
      T sn[7];
#pragma unroll
      for (int m = 0; m < 7; ++m) sn[m] = bn[min(lane + m, kRowBuf - 1)];
      hy0 = c6(sn[0], sn[1], sn[2], sn[3], sn[4], sn[5]);
      hy1 = c6(sn[1], sn[2], sn[3], sn[4], sn[5], sn[6]);
    }
    wave_lds_sync();
    const T ucy_n = __shfl_down(ucy, 1);
    const T vcy_n = __shfl_down(vcy, 1);
    const T Cy_n = __shfl_down(Cy, 1);
    const T ddx_n = __shfl_down(cl.ddx, 1);
    // vorticity (agrid.py:32-41) with vv = vb_cx, uu = ub_cy
    const T zeta = cl.rda * (((ucy * cl.ddx + vcx * cdy_i) - ucy_n * ddx_n) - vcx_prev * cdy_prev);
    const T ku1 = T(0.5) * (ucx * ucx + vcx * vcx);
    const T ku0 = T(0.5) * (ucx_prev * ucx_prev + vcx_prev * vcx_prev);
    const T kv1 = T(0.5) * (vcy_n * vcy_n + ucy_n * ucy_n);
    const T kv0 = T(0.5) * (vcy * vcy + ucy * ucy);
    const T vct_c = wvc[0];
    const T uct_c = wu[2];
    const T advx = zeta * vct_c - (ku1 - ku0) * cl.rdx;
    const T advy = -zeta * uct_c - (kv1 - kv0) * cl.rdy;
    const T visx = -(cl.rdx * (Cx - Cx_prev));
    const T visy = -(cl.rdy * (Cy_n - Cy));
    const T hx0 = c6(wn7[0], wn7[1], wn7[2], wn7[3], wn7[4], wn7[5]);
    const T hx1 = c6(wn7[1], wn7[2], wn7[3], wn7[4], wn7[5], wn7[6]);
    const T pgfu = T(-9.80) * (cl.rdx * (hx1 - hx0));
    const T pgfv = T(-9.80) * (cl.rdy * (hy1 - hy0));
    const uint32_t oc = w.fo + static_cast<uint32_t>(ii) * NS + j;
    T* du = out_lane ? at32(a.u_out, oc) : sink;
    T* dv = out_lane ? at32(a.v_out, oc) : sink;
    // wuu[0], wv[2]: u, v at row ii
    *du = (SAMEP ? wuu[0] : cl.up) + (((pgfu + advx) + cl.af * vct_c) + visx) * a.dt;
    *dv = (SAMEP ? wv[2] : cl.vp) + (((pgfv + advy) - cl.af * uct_c) + visy) * a.dt;
    b_fetch_cell<T, SAMEP>(a, w, i, pcell);

    ucx_prev = ucx;
    vcx_prev = vcx;
    Cx_prev = Cx;
    cdy_prev = cdy_i;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
      wu[m] = wu[m + 1];
      wv[m] = wv[m + 1];
      wh[m] = wh[m + 1];
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      wvc[m] = wvc[m + 1];
      wuu[m] = wuu[m + 1];
    }
#pragma unroll
    for (int m = 0; m < 6; ++m) wn7[m] = wn7[m + 1];
  }
}

template <typename T, bool GSYM, bool SAMEP>
__global__ __launch_bounds__(256) void k_march_a(MarchArgs<T> a) {
  __shared__ T s_v[4][kRowBuf];
  __shared__ T s_h[4][kRowBuf];
  march_a_body<T, GSYM, SAMEP>(a, xcd_block(), s_v, s_h);
}

template <typename T, bool GSYM, bool SAMEP>
__global__ __launch_bounds__(256) void k_march_b(MarchArgs<T> a) {
  __shared__ T s_vct[4][kRowBuf];
  __shared__ T s_u[4][kRowBuf];
  __shared__ T s_h[4][kRowBuf];
  __shared__ T s_n[4][kRowBuf];
  march_b_body<T, GSYM, SAMEP>(a, xcd_block(), s_vct, s_u, s_h, s_n);
}

// ---------------------------------------------------------------------------
// Fused stage kernel: kernel A's height march and kernel B's velocity march in
// one pass over the column loads.
//
// A wave of strip s loads the 64 columns c0 = 52 s - 3 .. c0 + 63 of a band
// (clamped to the face) and nothing outside them: a j-face reads the six
// columns around it, so the new height (kernel A's expressions) is valid on
// lanes 3..60, and kernel B's pressure gradient needs the new height at +-3
// rows and columns, so the velocity is valid on lanes 6..57.  The wave writes
// both on lanes 6..57 (columns 52 s + 3 .. 52 s + 54); the neighbouring strips
// recompute the three columns on either side.  The velocity is written at
// cells whose 7 x 7 new-height stencil lies inside the face ("deep" cells:
// rows and columns 6 .. N-7); the 3-cell frame along the face edges needs the
// new height's halo from the neighbouring faces and is done by k_frame after
// the halo fill.  Rows of the band: the march covers r0-3 .. r1+2 (clipped to
// the interior) and writes the height on r0 .. r1-1.
//
// Kernel B evaluates, on each i-face and j-face, the same upwind velocity,
// celerity and face values of h and of the normal velocity as kernel A's mass
// flux (the same IEEE operations on the same values), so one evaluation per
// face serves both (face_q).  Every term of the velocity update is formed at
// the height's row ia.  The v update completes there: its pressure gradient
// along j reads the new height of row ia only.  The u update's pressure
// gradient along i reads the new height at rows ib-3 .. ib+3, so its other
// terms wait in an LDS FIFO for three rows and the update of row ib (ia - 3
// marching down, ia + 3 up) is completed in the reference's order of
// operations.
//
// Column windows (registers) hold rows i-3 .. i+2 of the current i-face i:
// contravariant u, h and covariant v; and the blended new height of the
// lane's column over the seven rows the pressure gradient of row ib reads.
// LDS per wave: row buffers of contravariant v, covariant u and h (4 rows,
// written when the row is loaded, three rows ahead of its cell row), the
// blended new height of row ia and the FIFO (3 rows).
// ---------------------------------------------------------------------------
constexpr int kFStride = 52;  // columns a wave writes per row (lanes 6..57)
constexpr int kRing = 4;
constexpr int kFifo = 3;
constexpr int kFTerms = 5;    // advx cor_u visx u rdx of the u update

template <typename T>
struct FRing {
  T v[kRing][kRowBuf];  // contravariant v
  T u[kRing][kRowBuf];  // covariant u
  T h[kRing][kRowBuf];
  T b[kRowBuf];         // blended new height of row ia (own columns only)
  T q[kFifo][kFTerms][64];
  T e[2][2][64];        // cube-edge velocity / flux of the W (0) and E (1) j-faces, rows eb .. eb + 63
};

// Kernel A's mass flux through a face (barotr_pure.py:29-49) together with
// what kernel B takes from the same face (remap.py:42-95): s[] = h and w[] =
// normal velocity samples k-3 .. k+2 across the face, t[] the tangential
// velocity samples.  F mass flux, uc the upwind-corrected normal velocity, vt
// the upwinded tangential velocity, C the dissipation coefficient term.
template <typename T>
struct FaceQ {
  T F, uc, vt, C;
};

template <typename T>
__device__ __forceinline__ FaceQ<T> face_q(const T* s, const T* w, const T* t, bool edge, T upre_edge, T dz) {
  const T he = e5(s[0], s[1], s[2], s[3], s[4]);
  const T hw = w5(s[1], s[2], s[3], s[4], s[5]);
  const T ue = e5(w[0], w[1], w[2], w[3], w[4]);
  const T uw = w5(w[1], w[2], w[3], w[4], w[5]);
  const T inner = T(0.5) * (ue + uw);
  const T upre = edge ? upre_edge : inner;
  const T c = celerity(c6(s[0], s[1], s[2], s[3], s[4], s[5]), dz);
  FaceQ<T> q;
  q.uc = upre + k49_over(c) * (he - hw);
  q.F = dz * q.uc + (q.uc > T(0) ? he : hw) * q.uc;
  const T te = e5(t[0], t[1], t[2], t[3], t[4]);
  const T tw = w5(t[1], t[2], t[3], t[4], t[5]);
  q.vt = upre > T(0) ? te : tw;
  q.C = T(0.5) * c * (ue - uw);
  return q;
}

// Wave of logical block L -> (face, strip, band) of the fused kernel.
template <typename T>
__device__ __forceinline__ bool decode_fwave(const MarchArgs<T>& a, int L, WaveTile& w) {
  const int N = a.g.N;
  w.lane = threadIdx.x & 63;
  w.wslot = threadIdx.x >> 6;
  int wid, span, off, height, nw;
  if (a.R2 > 0) {
    // the long set is blocks [0, nlong), the short set the rest; each set in
    // its own XCD-contiguous order (block b runs on XCD b % 8; nlong is a
    // multiple of 8)
    const int b = static_cast<int>(blockIdx.x), G = static_cast<int>(gridDim.x);
    const bool lng = b < a.nlong;
    const int bb = lng ? b : b - a.nlong, cnt = lng ? a.nlong : G - a.nlong;
    const int x = bb % kXcds, q = cnt / kXcds, r = cnt % kXcds;
    wid = (x * q + min(x, r) + bb / kXcds) * 4 + w.wslot;
    span = a.R + a.R2;
    off = lng ? 0 : a.R;
    height = lng ? a.R : a.R2;
    nw = lng ? a.nwlong : a.nwaves;
    w.up = !lng;
  } else {
    wid = L * 4 + w.wslot;
    span = height = a.R;
    off = 0;
    nw = a.nwaves;
    w.up = false;
  }
  if (wid >= nw) return false;
  w.t = wid % 6;
  const int rest = wid / 6;
  const int s = rest % a.S;
  const int band = rest / a.S;
  w.c0 = kFStride * s - HALO;
  w.j = w.c0 + w.lane;
  w.jl = min(max(w.j, 0), N - 1);
  w.r0 = HALO + band * span + off;
  w.r1 = min(w.r0 + height, N - HALO);
  if (a.R2 <= 0) w.up = (band & 1) != 0;
  const uint32_t NN = static_cast<uint32_t>(N) * N;
  w.fo = static_cast<uint32_t>(w.t) * NN;
  w.mo = a.shared ? 0 : w.fo;
  return true;
}

// One wave's march over its band.  UP marches from the band's last row to its
// first, so that neighbouring bands (odd ones up, even ones down) read the
// rows they share near the same time, while the lines are in L2.  Both
// directions evaluate every expression on the same operands in the same order
// (the bottom face of a cell is the current face marching down and the
// previous one marching up).
template <typename T, bool GSYM, bool SAMEP, bool UP>
__device__ __forceinline__ void fused_march(const MarchArgs<T>& a, const WaveTile& w, FRing<T>& rg) {
  const int N = a.g.N;
  const Planes<T>& P = a.P;
  const int lane = w.lane, j = w.j, c0 = w.c0;
  const uint32_t NS = static_cast<uint32_t>(N);
  const T wn = T(1.0) - a.beta, wo = a.beta;
  // window slots: rows i-3 .. i+2 of the current face i; the newest row
  // (i+2 down, i-3 up: three rows ahead of the cell row) enters at NEW, the
  // cell row ia (i-1 down, i up) sits at SA
  constexpr int NEW = UP ? 0 : 5;
  constexpr int SA = UP ? 3 : 2;
  constexpr int STEP = UP ? -1 : 1;

  // the lanes load columns c0 .. c0+63 (clamped to the face); the new height
  // is valid on lanes 3..60 and the velocity on lanes 6..57, the columns the
  // wave writes
  const bool own = lane >= 2 * HALO && lane < 2 * HALO + kFStride && j >= HALO && j <= N - 1 - HALO;
  const bool deep = own && j >= 2 * HALO && j <= N - 1 - 2 * HALO;
  const int a0 = max(w.r0 - 3, HALO), a1 = min(w.r1 + 3, N - HALO);
  const int b0 = max(w.r0, 2 * HALO), b1 = min(w.r1, N - 2 * HALO);
  T* sink = a.sink + threadIdx.x;
  // j-faces on cube edges: lane j = HALO (W) or N - HALO (E) needs the
  // averaged edge velocity / flux of its cell row.  The band's values sit in
  // LDS, 64 rows from eb, lane k loading row eb + k: one wave-wide load per
  // 64 rows instead of two per row.
  const bool ew = w.c0 <= HALO && HALO <= w.c0 + 63, ee = w.c0 <= N - HALO && N - HALO <= w.c0 + 63;
  const int es = j == HALO ? 0 : 1;  // the side lane j reads (its value is used only on an edge lane)
  int eb = -1000;
  auto edge_rows = [&](int base) {
    const int nx = a.g.nx;
    const int row = min(base + lane, N - HALO - 1);
    const int o = row - HALO;
    if (ew) {
      rg.e[0][0][lane] = ldg(a.evel, (w.t * 4 + DW) * nx + o);
      rg.e[0][1][lane] = ldg(a.eflux, (w.t * 4 + DW) * nx + o);
    }
    if (ee) {
      rg.e[1][0][lane] = ldg(a.evel, (w.t * 4 + DE) * nx + o);
      rg.e[1][1][lane] = ldg(a.eflux, (w.t * 4 + DE) * nx + o);
    }
    eb = base;
  };

  // the step's inputs at face i: the newest row's column point, i-face i,
  // cell row ia (no halo column: hl = false)
  auto fetch = [&](int i, APre<T>& p) {
    const int rn = UP ? max(i - 3, 0) : min(i + 2, N - 1);
    a_fetch<T, GSYM, SAMEP>(a, w, i, rn, UP ? min(i, N - 1) : max(i - 1, 0), rn, w.jl, p, false, false);
  };
  auto shift = [](T* x) {
    if (UP) {
#pragma unroll
      for (int m = 5; m > 0; --m) x[m] = x[m - 1];
    } else {
#pragma unroll
      for (int m = 0; m < 5; ++m) x[m] = x[m + 1];
    }
  };
  // row buffers of the newest row: contravariant v, covariant u and h
  auto put_row = [&](int row, const CtRaw<T>& c, T hc) {
    const int s = row & (kRing - 1);
    rg.v[s][lane + 3] = ct_v(c);
    rg.u[s][lane + 3] = c.u;
    rg.h[s][lane + 3] = hc;
  };

  T cu[6], hh[6], cv[6], wn7[7];
  // the first face f0 = a0 (down) / a1 (up) reads rows f0-3 .. f0+2; all but
  // the newest are loaded here, with the row buffers of the first cells
  const int f0 = UP ? a1 : a0;
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    const int slot = UP ? m + 1 : m;
    const int row = f0 - 3 + slot;
    const uint32_t o = w.fo + static_cast<uint32_t>(row) * NS;
    CtRaw<T> r;
    load_ct2<T, GSYM>(a, o + w.jl, r);
    cu[slot] = ct_u(r);
    cv[slot] = r.v;
    hh[slot] = ldg(a.h, o + w.jl);
    if (UP ? m < 2 : m >= 2) put_row(row, r, hh[slot]);
  }
#pragma unroll
  for (int m = 0; m < 7; ++m) wn7[m] = T(0);
  APre<T> pre;
  fetch(f0, pre);

  // face f0 only: its flux and the velocity terms of the face
  T Fprev, cdy_prev, ucx_prev, vcx_prev, Cx_prev, ku_prev;
  {
    const APre<T> cur = pre;
    fetch(f0 + STEP, pre);
    cu[NEW] = ct_u(cur.cr);
    cv[NEW] = cur.cr.v;
    hh[NEW] = cur.h;
    put_row(UP ? f0 - 3 : f0 + 2, cur.cr, cur.h);
    const bool edge = (f0 == HALO) || (f0 == N - HALO);
    const FaceQ<T> q = face_q(hh, cu, cv, edge, cur.evi, cur.dzx);
    Fprev = (a.last && edge) ? cur.efi : q.F;
    cdy_prev = cur.cdy;
    ucx_prev = q.uc;
    vcx_prev = q.vt;
    Cx_prev = q.C;
    ku_prev = T(0.5) * (q.uc * q.uc + q.vt * q.vt);
    shift(cu);
    shift(cv);
    shift(hh);
  }

  int sq = 0;  // FIFO slot of rows ia and ib (read before written)
  // c6 of the u update's new-height window that the next step reuses: the
  // window slides one row, so one of its two 6-point values is the other
  // one of the step before (a window of zeros before the first step)
  T hx_prev = T(0);
  // step n reads the inputs fetched by the step before (cur) and fetches the
  // next step's (nxt); the loop alternates two buffers, so no copy between
  // them is needed
  auto step = [&](int n, const APre<T>& cur, APre<T>& nxt) __attribute__((always_inline)) {
    const int i = f0 + STEP * n;
    const int ia = UP ? i : i - 1;
    const int ib = UP ? ia + 3 : ia - 3;
    const int sa = ia & (kRing - 1);
    const uint32_t om = static_cast<uint32_t>(ia) * NS + w.jl;
    T evj = T(0), efj = T(0);
    if (ew || ee) {  // wave-uniform: strips holding a cube-edge column
      if (ia < eb || ia >= eb + 64) {
        wave_lds_sync();  // earlier reads of the previous rows are done
        edge_rows(UP ? max(ia - 63, HALO) : ia);
        wave_lds_sync();
      }
      evj = rg.e[es][0][ia - eb];
      efj = rg.e[es][1][ia - eb];
    }
    // cell inputs of the velocity terms at row ia
    const Pair<T> rxy = ldg2(P.rxy, w.mo + om);
    const T rdx = rxy.x, rdy = rxy.y, af = ldg(P.af, w.fo + om);
    T upl = T(0), vpl = T(0);
    if constexpr (!SAMEP) {
      upl = ldg(a.up, w.fo + om);
      vpl = ldg(a.vp, w.fo + om);
    }
    // row ia's buffers (written three steps ago) and the FIFO entry of row
    // ib, read before this step's LDS writes so their latency overlaps the
    // i-face arithmetic
    T sv[6], su[6], sh[6], tq0[kFTerms];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      sv[m] = rg.v[sa][lane + m];
      su[m] = rg.u[sa][lane + m];
      sh[m] = rg.h[sa][lane + m];
    }
#pragma unroll
    for (int k = 0; k < kFTerms; ++k) tq0[k] = rg.q[sq][k][lane];
    fetch(i + STEP, nxt);
    cu[NEW] = ct_u(cur.cr);
    cv[NEW] = cur.cr.v;
    hh[NEW] = cur.h;
    put_row(UP ? i - 3 : i + 2, cur.cr, cur.h);

    // i-face at row i: mass flux and velocity terms
    const bool iedge = (i == HALO) || (i == N - HALO);
    const FaceQ<T> qi = face_q(hh, cu, cv, iedge, cur.evi, cur.dzx);
    const T F = (a.last && iedge) ? cur.efi : qi.F;

    // j-face at column j, row ia
    const T vct_c = sv[3], ucov_c = su[3];
    const bool jedge = (j == HALO) || (j == N - HALO);
    const FaceQ<T> qj = face_q(sh, sv, su, jedge, evj, cur.dzy);
    const T G = (a.last && jedge) ? efj : qj.F;

    // the cell's bottom (row ia+1) and top (row ia) i-faces
    const T Fb = UP ? Fprev : F, Ft = UP ? F : Fprev;
    const T cdyb = UP ? cdy_prev : cur.cdy, cdyt = UP ? cur.cdy : cdy_prev;

    // height at row ia: divergence (agrid.py:44-53), barotr_pure.py:120
    // lane j+1's G * ddx is the j-face term at j+1 (the same product)
    const T Gd = G * cur.ddx;
    const T Gdn = __shfl_down(Gd, 1);
    const T D = cur.rda * (((Fb * cdyb - Ft * cdyt) + Gdn) - Gd);
    const T hnew = (SAMEP ? hh[SA] : cur.hp) - D * a.dt;  // hh[SA]: h at row ia
    const bool arow = ia >= w.r0 && ia < w.r1;
    T* dh = (own && arow) ? at32(a.h_out, w.fo + static_cast<uint32_t>(ia) * NS + j) : sink;
    *dh = hnew;
    const T hb = wn * hnew + wo * hh[SA];
    rg.b[lane + 3] = hb;
    if (UP) {
#pragma unroll
      for (int m = 6; m > 0; --m) wn7[m] = wn7[m - 1];
      wn7[0] = hb;
    } else {
#pragma unroll
      for (int m = 0; m < 6; ++m) wn7[m] = wn7[m + 1];
      wn7[6] = hb;
    }

    // velocity terms at row ia (barotr_pure.py:133-150, agrid.py:32-41):
    // vorticity with vv = vb_cx, uu = ub_cy, kinetic-energy gradient,
    // dissipation and Coriolis
    const T ucy = qj.vt, vcy = qj.uc, Cy = qj.C;
    // lane j+1's products at the j-face j+1: ucy * ddx and the kinetic energy
    const T Uy = ucy * cur.ddx;
    const T kv0 = T(0.5) * (vcy * vcy + ucy * ucy);
    const T Uyn = __shfl_down(Uy, 1);
    const T kv1 = __shfl_down(kv0, 1);
    const T Cy_n = __shfl_down(Cy, 1);
    const T ucxb = UP ? ucx_prev : qi.uc, ucxt = UP ? qi.uc : ucx_prev;
    const T vcxb = UP ? vcx_prev : qi.vt, vcxt = UP ? qi.vt : vcx_prev;
    const T Cxb = UP ? Cx_prev : qi.C, Cxt = UP ? qi.C : Cx_prev;
    const T zeta = cur.rda * (((Uy + vcxb * cdyb) - Uyn) - vcxt * cdyt);
    // kinetic energy of the i-faces: this step's face and the previous one's
    const T kuf = T(0.5) * (qi.uc * qi.uc + qi.vt * qi.vt);
    const T ku1 = UP ? ku_prev : kuf, ku0 = UP ? kuf : ku_prev;
    const T uct_c = cu[SA];  // contravariant u at row ia
    const T advx = zeta * vct_c - (ku1 - ku0) * rdx;
    const T advy = -zeta * uct_c - (kv1 - kv0) * rdy;
    const T visx = -(rdx * (Cxb - Cxt));
    const T visy = -(rdy * (Cy_n - Cy));
    wave_lds_sync();

    // v at row ia: the pressure gradient along j needs the new height of row
    // ia only (columns j-3 .. j+3)
    {
      T sn[7];
#pragma unroll
      for (int m = 0; m < 7; ++m) sn[m] = rg.b[min(lane + m, kRowBuf - 1)];
      // lane j-1's hy1 is this lane's hy0 (the same six values)
      const T hy1 = c6(sn[1], sn[2], sn[3], sn[4], sn[5], sn[6]);
      const T hy0 = lane_prev(hy1);
      const T pgfv = T(-9.80) * (rdy * (hy1 - hy0));
      const bool vrow = ia >= b0 && ia < b1;
      T* dv = (deep && vrow) ? at32(a.v_out, w.fo + static_cast<uint32_t>(ia) * NS + j) : sink;
      *dv = (SAMEP ? cv[SA] : vpl) + (((pgfv + advy) - af * uct_c) + visy) * a.dt;
    }

    // u at row ib: the pressure gradient along i needs the new height of rows
    // ib-3 .. ib+3; the other terms were formed three steps ago
    {
      T tq[kFTerms];
#pragma unroll
      for (int k = 0; k < kFTerms; ++k) tq[k] = tq0[k];
      const T hx0 = UP ? c6(wn7[0], wn7[1], wn7[2], wn7[3], wn7[4], wn7[5]) : hx_prev;
      const T hx1 = UP ? hx_prev : c6(wn7[1], wn7[2], wn7[3], wn7[4], wn7[5], wn7[6]);
      hx_prev = UP ? hx0 : hx1;
      const T pgfu = T(-9.80) * (tq[4] * (hx1 - hx0));
      const bool urow = ib >= b0 && ib < b1;
      const uint32_t oc = w.fo + static_cast<uint32_t>(min(max(ib, 0), N - 1)) * NS + j;
      T* du = (deep && urow) ? at32(a.u_out, oc) : sink;
      // barotr_pure.py:133-150 with the terms advx, cor_u, visx, u of row ib
      *du = tq[3] + (((pgfu + tq[0]) + tq[1]) + tq[2]) * a.dt;
      tq[0] = advx;
      tq[1] = af * vct_c;
      tq[2] = visx;
      tq[3] = SAMEP ? ucov_c : upl;  // u at row ia
      tq[4] = rdx;
#pragma unroll
      for (int k = 0; k < kFTerms; ++k) rg.q[sq][k][lane] = tq[k];
    }
    wave_lds_sync();

    Fprev = F;
    cdy_prev = cur.cdy;
    ucx_prev = qi.uc;
    vcx_prev = qi.vt;
    Cx_prev = qi.C;
    ku_prev = kuf;
    sq = sq == kFifo - 1 ? 0 : sq + 1;
    shift(cu);
    shift(cv);
    shift(hh);
  };
  const int cnt = a1 - a0;
  APre<T> alt;
  int n = 1;
  for (; n < cnt; n += 2) {
    step(n, pre, alt);
    step(n + 1, alt, pre);
  }
  if (n == cnt) step(n, pre, alt);
}

// Launch stamps (sw_profile_graph): lane 0 of each wave writes the device
// wall clock to slot 2 w (entry) or 2 w + 1 (exit) of a.stamps, w the
// hardware wave index of the launch; nothing is held across the march.
template <typename T>
__device__ __forceinline__ void stamp(const MarchArgs<T>& a, int which) {
  if (a.stamps && (threadIdx.x & 63) == 0)
    a.stamps[2 * (static_cast<size_t>(blockIdx.x) * 4 + (threadIdx.x >> 6)) + which] = wall_clock64();
}

template <typename T, bool GSYM, bool SAMEP>
__device__ __forceinline__ void fused_body(const MarchArgs<T>& a, int L, FRing<T>* rings) {
  stamp(a, 0);
  WaveTile w;
  if (!decode_fwave(a, L, w)) {
    stamp(a, 1);
    return;
  }
  FRing<T>& rg = rings[w.wslot];
  if (w.up)
    fused_march<T, GSYM, SAMEP, true>(a, w, rg);
  else
    fused_march<T, GSYM, SAMEP, false>(a, w, rg);
  stamp(a, 1);
}

template <typename T, bool GSYM, bool SAMEP>
__global__ __launch_bounds__(256, 2) void k_fused(MarchArgs<T> a) {
  __shared__ FRing<T> rings[4];
  fused_body<T, GSYM, SAMEP>(a, xcd_block(), rings);
}

// ---------------------------------------------------------------------------
// duo-grid halo fill (update_domain pass 1 + cube_rmp, closed index form)
// ---------------------------------------------------------------------------
// Interior cell at distance layer (1..3) inside edge d, along-edge index r.
__device__ __forceinline__ void interior_cell(int N, int d, int layer, int r, int& i, int& j) {
  if (d == DE) { i = 3 + r; j = N - 3 - layer; }
  else if (d == DW) { i = 3 + r; j = 2 + layer; }
  else if (d == DN) { i = N - 3 - layer; j = 3 + r; }
  else { i = 2 + layer; j = 3 + r; }
}

// Source offset of the value copied into the halo of face t, edge d, layer,
// at along-edge index r (communication.py:404-430).
__device__ __forceinline__ size_t copy_source(const Geo& g, const Route& R, int t, int d, int layer, int r) {
  const int tb = R.r[t][d][0], db = R.r[t][d][1];
  const int rr = R.r[t][d][2] ? g.nx - 1 - r : r;
  int i, j;
  interior_cell(g.N, db, layer, rr, i, j);
  return g.at(tb, i, j);
}

// Strip cell (i, j) of face t: its edge, layer and the index along the strip.
// Returns the two interpolation sources (cube.py:64-94).
template <typename T>
__device__ __forceinline__ void strip_sources(const Geo& g, const Route& R, const Planes<T>& P, int t,
                                              int i, int j, size_t& s0, size_t& s1, T& c0, T& c1) {
  const int N = g.N;
  const size_t o = g.at(t, i, j);
  c0 = P.k0[o];
  c1 = P.k1[o];
  if (j < 3 || j >= N - 3) {            // W / E strips interpolate along i
    const int d = j < 3 ? DW : DE;
    const int layer = j < 3 ? 3 - j : j - (N - 4);
    const int L = P.li[o];
    s0 = copy_source(g, R, t, d, layer, L - 3);
    s1 = copy_source(g, R, t, d, layer, L + 1 - 3);
  } else {                              // S / N strips interpolate along j
    const int d = i < 3 ? DS : DN;
    const int layer = i < 3 ? 3 - i : i - (N - 4);
    const int L = P.lj[o];
    s0 = copy_source(g, R, t, d, layer, L - 3);
    s1 = copy_source(g, R, t, d, layer, L + 1 - 3);
  }
}

// Halo cell q of a face (strips then corner blocks) -> (i, j) of the cell and
// (si, sj) of the strip cell whose remapped value it takes (itself for strip
// cells; corners copy row 3 / N-4 of the W/E strips, cube.py:96-125).
__device__ __forceinline__ void halo_cell(int N, int nx, int q, int& i, int& j, int& si, int& sj) {
  const int strip = 12 * nx;
  if (q < strip) {
    const int d = q / (3 * nx);
    const int rem = q - d * 3 * nx;
    const int layer = 1 + rem / nx;
    const int r = rem % nx;
    if (d == DE) { i = 3 + r; j = N - 4 + layer; }
    else if (d == DW) { i = 3 + r; j = 3 - layer; }
    else if (d == DN) { i = N - 4 + layer; j = 3 + r; }
    else { i = 3 - layer; j = 3 + r; }
    si = i;
    sj = j;
  } else {
    const int c = q - strip;          // 0..35
    const int corner = c / 9, w = c % 9;
    const int a = w / 3, b = w % 3;
    const bool north = corner >= 2, east = (corner & 1) != 0;
    i = north ? N - 3 + a : a;
    j = east ? N - 3 + b : b;
    si = north ? N - 4 : 3;
    sj = j;
  }
}

// Halo plan.  Every halo cell's gathers are resolved once, at context
// creation: entry q holds the cell's offset, its two interpolation sources and
// weights, and - for vector fields - the rotation coefficients the remap
// reads at the two sources (cube to lat-lon) and at the cell itself (lat-lon
// to cube).  They are copies of const entries, so the planned fill evaluates
// exactly the expressions of the direct one while reading a few contiguous
// arrays instead of scattered lines of ten (6, N, N) planes.  Entries of the
// W/E strips run layer fastest: three neighbouring lanes write three adjacent
// cells of one row.

template <typename T>
struct PlanBuildArgs {
  Geo g;
  Route R;
  Planes<T> P;
  uint32_t *dst, *s0, *s1;
  T *c0, *c1, *rot;
  int n;
  int* far;  // counts entries whose sources lie more than 3 positions from the cell
};

template <typename T>
__global__ __launch_bounds__(256) void k_plan_build(PlanBuildArgs<T> a) {
  const Geo g = a.g;
  const int nx = g.nx, per_face = 12 * nx + 36;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.n) return;
  const int t = q / per_face;
  int qf = q - t * per_face;
  if (qf < 6 * nx) {  // W/E strips: entry order (d, r, layer) -> halo_cell order (d, layer, r)
    const int d = qf / (3 * nx), rem = qf - d * 3 * nx;
    qf = d * 3 * nx + (rem % 3) * nx + rem / 3;
  }
  int i, j, si, sj;
  halo_cell(g.N, nx, qf, i, j, si, sj);
  size_t s0, s1;
  T c0, c1;
  strip_sources(g, a.R, a.P, t, si, sj, s0, s1, c0, c1);
  const size_t o = g.at(t, i, j);
  a.dst[q] = static_cast<uint32_t>(o);
  a.s0[q] = static_cast<uint32_t>(s0);
  a.s1[q] = static_cast<uint32_t>(s1);
  a.c0[q] = c0;
  a.c1[q] = c1;
  const Planes<T>& P = a.P;
  const T r[kPlanRot] = {P.c00[s0], P.c01[s0], P.c10[s0], P.c11[s0], P.c00[s1], P.c01[s1],
                         P.c10[s1], P.c11[s1], P.l00[o], P.l01[o], P.l10[o], P.l11[o]};
#pragma unroll
  for (int k = 0; k < kPlanRot; ++k) a.rot[static_cast<size_t>(k) * a.n + q] = r[k];
  // k_frame_edge holds the neighbour's frame cells within 3 along-edge
  // positions of a halo cell's own position: the sources, at along-edge
  // indices L-3 and L-2 (strip_sources), must fall there
  const bool w_e = sj < HALO || sj >= g.N - HALO;
  const size_t os = g.at(t, si, sj);
  const int L = w_e ? P.li[os] : P.lj[os];
  const int r0 = w_e ? si - HALO : sj - HALO;
  if (L - 3 < max(r0 - HALO, 0) || L - 2 > min(r0 + HALO, nx - 1)) atomicAdd(a.far, 1);
}

template <typename T>
struct HaloArgs {
  HaloPlan<T> plan;
  uint32_t NN;  // N * N
  T *f;         // scalar field base (6, ncomp, N, N)
  int ncomp;
  T *u, *v;     // vector pair
};

// Scalar halo fill (update_domain pass 1 + cube_rmp, communication.py:404-430,
// cube.py:64-125) from the plan.
template <typename T>
__device__ __forceinline__ void halo_scalar_body(const HaloArgs<T>& a, int q) {
  if (q >= a.plan.n) return;
  const uint32_t d = a.plan.dst[q], s0 = a.plan.s0[q], s1 = a.plan.s1[q];
  const T c0 = a.plan.c0[q], c1 = a.plan.c1[q];
  if (a.ncomp == 1) {
    a.f[d] = c0 * a.f[s0] + c1 * a.f[s1];
    return;
  }
  // face-major stacking: face f, component comp at (f * ncomp + comp) * NN
  const uint32_t NN = a.NN;
  auto off = [&](uint32_t o, int comp) {
    const size_t f = o / NN;
    return (f * a.ncomp + comp) * NN + (o - f * NN);
  };
  for (int comp = 0; comp < a.ncomp; ++comp) {
    const T val = c0 * a.f[off(s0, comp)] + c1 * a.f[off(s1, comp)];
    a.f[off(d, comp)] = val;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_halo_scalar(HaloArgs<T> a) {
  halo_scalar_body(a, static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x));
}

// Vector halo fill from the plan: the copied interior values rotated to
// lat-lon (cube.py:139-140; inner = 1), interpolated, and rotated back with
// the cell's own rotation (cube.py:147-151; outer = 1).
template <typename T>
__device__ __forceinline__ void halo_vector_body(const HaloArgs<T>& a, int q) {
  const int n = a.plan.n;
  if (q >= n) return;
  const uint32_t d = a.plan.dst[q];
  T uo, vo;
  halo_uv_value(a.plan, a.u, a.v, q, uo, vo);
  a.u[d] = uo;
  a.v[d] = vo;
}

template <typename T>
__global__ __launch_bounds__(256) void k_halo_vector(HaloArgs<T> a) {
  halo_vector_body(a, static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x));
}

// The velocity halo fill of a stage (blocks 0 .. nh-1) and the next stage's
// edge pass on the same fields (the remaining blocks), which forms the halo
// values it reads from the plan.
template <typename T>
__global__ __launch_bounds__(256) void k_halo_uv_edge(HaloArgs<T> huv, EdgeArgs<T> en, int nh) {
  const int blk = static_cast<int>(blockIdx.x);
  if (blk < nh)
    halo_vector_body(huv, blk * 256 + static_cast<int>(threadIdx.x));
  else
    edge_body(en, (blk - nh) * 256 + static_cast<int>(threadIdx.x));
}

// Frame tiles of a face: the two groups of frame rows (3..5, N-6..N-4) in
// chunks of kTile columns, then the two groups of frame columns (3..5,
// N-6..N-4) of rows 6..N-7 in chunks of kTile rows (nx >= 7).  A block loads
// the points its cells' stencils read (the cells and 3 around them) into LDS
// with row-contiguous loads.  Then a group of 4 lanes takes each cell: lane k
// evaluates face k of the cell (0 bottom i-face, 1 top i-face, 2 the j-face
// at j, 3 the j-face at j+1; kernel B's expressions) and even / odd lanes the
// pressure gradient along i / j, and lane 0 gathers them and completes the
// update in the reference's order of operations.
// Value of x in lane Q (0..3) of this lane's quad (lanes 4k .. 4k + 3): a DPP
// quad permutation, a register move with no LDS round trip.  Every lane of
// the quad must be active.
template <int Q>
__device__ __forceinline__ int quad_bcast_b32(int x) {
  return __builtin_amdgcn_update_dpp(0, x, Q | (Q << 2) | (Q << 4) | (Q << 6), 0xF, 0xF, false);
}
template <int Q, typename T>
__device__ __forceinline__ T quad_bcast(T x) {
  if constexpr (sizeof(T) == 8) {
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = static_cast<uint32_t>(quad_bcast_b32<Q>(static_cast<int>(static_cast<uint32_t>(b))));
    const uint32_t hi = static_cast<uint32_t>(quad_bcast_b32<Q>(static_cast<int>(static_cast<uint32_t>(b >> 32))));
    return __builtin_bit_cast(T, (static_cast<uint64_t>(hi) << 32) | lo);
  } else {
    return __builtin_bit_cast(T, quad_bcast_b32<Q>(__builtin_bit_cast(int, x)));
  }
}

constexpr int kTile = 21;                    // cells along a tile's long side
constexpr int kTilePts = (kTile + 6) * 9;    // points of a tile's region
constexpr int kFrameLanes = 4;

__host__ __device__ inline int frame_tiles(int nx) {
  return 2 * ((nx + kTile - 1) / kTile) + 2 * ((nx - 6 + kTile - 1) / kTile);
}

// A rectangle of frame cells: crows x ccols cells of face t from (cr0, cc0),
// one side at most kTile long and the other 3.
struct FTile {
  int t, cr0, crows, cc0, ccols;
  int strip = -1;  // W (0) / E (1) face-edge tile reading the strip copy of the constant planes, or -1
};

// Element (row i, strip column c) of strip plane k, face t, side 0 (W) / 1 (E).
template <typename T>
__device__ __forceinline__ T strip_ld(const Planes<T>& P, int N, int t, int side, int k, int i, int c) {
  return ldg(P.strip, static_cast<uint32_t>((((t * 2 + side) * kStripPlanes + k) * N + i) * 9 + c));
}

// The frame cells of tile ft, evaluated by a group of 256 threads (ltid its
// thread index in the group): the group loads the region its cells' stencils
// read (the cells and 3 points around them) into s, then 4 lanes take each
// cell.  Holds one __syncthreads(), which every thread of the block reaches.
// Returns true on the lane that completes a cell, with its new velocity
// (un, vn), its position (ci, cj) and its tile-local index cq = (ci - cr0) *
// ccols + (cj - cc0).
template <typename T, bool GSYM>
__device__ __forceinline__ bool frame_cell(const MarchArgs<T>& a, const HaloPlan<T>& plan, const FTile& ft,
                                           T (*s)[kTilePts], int ltid, T& un, T& vn, int& ci, int& cj, int& cq) {
  const int nx = a.g.nx, N = a.g.N;
  const int t = ft.t, cr0 = ft.cr0, crows = ft.crows, cc0 = ft.cc0, ccols = ft.ccols;
  const int ra = cr0 - 3, ca = cc0 - 3, rcols = ccols + 6, npts = (crows + 6) * rcols;
  const Planes<T>& P = a.P;
  const uint32_t NS = static_cast<uint32_t>(N);
  const uint32_t fo = static_cast<uint32_t>(t) * NS * NS, mo = a.shared ? 0 : fo;
  // the cell of this lane's group (the last cell for lanes past the tile's
  // cells, which only help fill the tile)
  const int q0 = ltid / kFrameLanes, kl = ltid % kFrameLanes;
  const int q = min(q0, crows * ccols - 1);
  const int ii = cr0 + q / ccols, j = cc0 + q % ccols;
  const int li = ii - ra, lj = j - ca;
  const bool iface = kl < 2;
  const int fr = kl == 0 ? ii + 1 : ii, fc = kl == 3 ? j + 1 : j;
  const bool pi = (kl & 1) == 0;
  const uint32_t oc = static_cast<uint32_t>(ii) * NS + static_cast<uint32_t>(j);
  // global inputs besides the tile, issued before the tile loads
  const int evi_e = (fr == HALO || fr == N - HALO) && j <= N - 4 ? (t * 4 + (fr == HALO ? DS : DN)) * nx + (j - HALO) : 0;
  const int evj_e = (fc == HALO || fc == N - HALO) && ii >= HALO && ii < N - HALO
                        ? (t * 4 + (fc == HALO ? DW : DE)) * nx + (ii - HALO) : 0;
  const T fev = ldg(a.evel, iface ? evi_e : evj_e);
  const uint32_t fd = static_cast<uint32_t>(fr) * NS + static_cast<uint32_t>(fc);
  // (choosing among the argument struct's pointers per lane would make the
  // compiler copy the struct to scratch memory: the loads stay in branches)
  const T fdz = iface ? ldg(P.dzx, P.dz_uniform ? 0 : fo + fd) : ldg(P.dzy, P.dz_uniform ? 0 : fo + fd);
  const T rpg = pi ? ldg(P.rdx, mo + oc) : ldg(P.rdy, mo + oc);
  // cell inputs of lane 0's update (lane kl loads the kl-th group)
  const uint32_t on = oc + NS;
  T cin[4];
  if (kl == 0) {
    cin[0] = ldg(P.cdy, mo + on); cin[1] = ldg(P.cdy, mo + oc); cin[2] = ldg(P.rda, mo + oc); cin[3] = ldg(P.ddx, mo + oc);
  } else if (kl == 1) {
    cin[0] = ldg(P.ddx, mo + oc + 1); cin[1] = ldg(P.rdx, mo + oc); cin[2] = ldg(P.rdy, mo + oc); cin[3] = ldg(P.af, fo + oc);
  } else {
    cin[0] = ldg(a.up, fo + oc); cin[1] = ldg(a.vp, fo + oc); cin[2] = T(0); cin[3] = T(0);
  }
  for (int e = ltid; e < npts; e += 256) {
    // the region is kTile + 6 or 9 points wide: constant divisors
    const int rr = rcols == 9 ? e / 9 : rcols == kTile + 6 ? e / (kTile + 6) : e / rcols;
    const int cc = e - rr * rcols;
    const uint32_t o = fo + static_cast<uint32_t>(ra + rr) * NS + static_cast<uint32_t>(ca + cc);
    s[0][e] = ldg(a.u, o);
    s[1][e] = ldg(a.v, o);
    s[2][e] = ldg(P.g00, o);
    s[3][e] = ldg(P.g01, o);
    if (!GSYM) {
      s[4][e] = ldg(P.g10, o);
      s[5][e] = ldg(P.g11, o);
    }
    s[6][e] = ldg(P.sina, o);
    s[7][e] = ldg(a.h, o);
    // the new height on a strip halo cell is the scalar fill's value, formed
    // here from the halo plan (its sources are interior cells), so the fill
    // may run in the same launch; corner cells are never read
    const int row = ra + rr, col = ca + cc;
    const bool hrow = row < HALO || row >= N - HALO, hcol = col < HALO || col >= N - HALO;
    if (hrow != hcol) {
      const int pq = halo_entry(nx, N, t, row, col);
      s[8][e] = plan.c0[pq] * ldg(a.hn, plan.s0[pq]) + plan.c1[pq] * ldg(a.hn, plan.s1[pq]);
    } else {
      s[8][e] = ldg(a.hn, o);
    }
  }
  __syncthreads();
  if (q0 >= crows * ccols) return false;  // whole groups: 4 divides 64
  const T wn = T(1.0) - a.beta, wo = a.beta;

  // face kl from its 6 points: normal velocity (contravariant), tangential
  // velocity (covariant) and h
  const int e0 = iface ? (fr - 3 - ra) * rcols + lj : li * rcols + (fc - 3 - ca);
  const int es = iface ? rcols : 1;
  T wf[6], tf[6], hf[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    const int e = e0 + m * es;
    const T u = s[0][e], v = s[1][e], sina = s[6][e];
    const T g00 = s[2][e], g01 = s[3][e];
    const T g10 = GSYM ? g01 : s[4][e], g11 = GSYM ? g00 : s[5][e];
    // the normal component's coefficients chosen per lane, one expression
    const T ga = iface ? g00 : g10, gb = iface ? g01 : g11;
    wf[m] = contra(ga, gb, u, v, sina);
    tf[m] = iface ? v : u;
    hf[m] = s[7][e];
  }
  const bool fedge = iface ? (fr == HALO) || (fr == N - HALO) : (fc == HALO) || (fc == N - HALO);
  T uc, vt, C;
  iface_b(wf, tf, hf, fedge, fev, fdz, uc, vt, C);

  // pressure gradient along i (even lanes) or j (odd lanes) at (ii, j)
  const int p0 = pi ? (li - 3) * rcols + lj : li * rcols + lj - 3;
  const int ps = pi ? rcols : 1;
  T hb[7];
#pragma unroll
  for (int m = 0; m < 7; ++m) hb[m] = wn * s[8][p0 + m * ps] + wo * s[7][p0 + m * ps];
  const T h0 = c6(hb[0], hb[1], hb[2], hb[3], hb[4], hb[5]);
  const T h1 = c6(hb[1], hb[2], hb[3], hb[4], hb[5], hb[6]);
  const T pgf = T(-9.80) * (rpg * (h1 - h0));

  // gather in lane 0 of the group (the group is a quad: DPP moves)
  const T ucx_prev = quad_bcast<1>(uc), vcx_prev = quad_bcast<1>(vt), Cx_prev = quad_bcast<1>(C);
  const T vcy = quad_bcast<2>(uc), ucy = quad_bcast<2>(vt), Cy = quad_bcast<2>(C);
  const T vcy_n = quad_bcast<3>(uc), ucy_n = quad_bcast<3>(vt), Cy_n = quad_bcast<3>(C);
  const T pgfv = quad_bcast<1>(pgf);
  T cv1[4], cv2[2];
#pragma unroll
  for (int m = 0; m < 4; ++m) cv1[m] = quad_bcast<1>(cin[m]);
  cv2[0] = quad_bcast<2>(cin[0]);
  cv2[1] = quad_bcast<2>(cin[1]);
  if (kl != 0) return false;
  const T ucx = uc, vcx = vt, Cx = C, pgfu = pgf;
  const int ec = li * rcols + lj;
  const T uc0 = s[0][ec], vc0 = s[1][ec], sc0 = s[6][ec];
  const T g00c = s[2][ec], g01c = s[3][ec];
  const T g10c = GSYM ? g01c : s[4][ec], g11c = GSYM ? g00c : s[5][ec];
  const T vct_c = contra(g10c, g11c, uc0, vc0, sc0);
  const T uct_c = contra(g00c, g01c, uc0, vc0, sc0);
  const T cdy_i = cin[0], cdy_prev = cin[1], rda = cin[2], ddx = cin[3];
  const T ddx_n = cv1[0], rdx = cv1[1], rdy = cv1[2], af = cv1[3];
  const T up = cv2[0], vp = cv2[1];
  const T zeta = rda * (((ucy * ddx + vcx * cdy_i) - ucy_n * ddx_n) - vcx_prev * cdy_prev);
  const T ku1 = T(0.5) * (ucx * ucx + vcx * vcx);
  const T ku0 = T(0.5) * (ucx_prev * ucx_prev + vcx_prev * vcx_prev);
  const T kv1 = T(0.5) * (vcy_n * vcy_n + ucy_n * ucy_n);
  const T kv0 = T(0.5) * (vcy * vcy + ucy * ucy);
  const T advx = zeta * vct_c - (ku1 - ku0) * rdx;
  const T advy = -zeta * uct_c - (kv1 - kv0) * rdy;
  const T visx = -(rdx * (Cx - Cx_prev));
  const T visy = -(rdy * (Cy_n - Cy));
  un = up + (((pgfu + advx) + af * vct_c) + visx) * a.dt;
  vn = vp + (((pgfv + advy) - af * uct_c) + visy) * a.dt;
  ci = ii;
  cj = j;
  cq = q;
  return true;
}

// Lane 1's value of x in lane 0 of each lane pair (a DPP quad permutation
// [1, 1, 3, 3]); both lanes of the pair must be active.
template <typename T>
__device__ __forceinline__ T pair_hi(T x) {
  constexpr int sel = 1 | (1 << 2) | (3 << 4) | (3 << 6);
  if constexpr (sizeof(T) == 8) {
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(b)), sel, 0xF, 0xF, false));
    const uint32_t hi = static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(b >> 32)), sel, 0xF, 0xF, false));
    return __builtin_bit_cast(T, (static_cast<uint64_t>(hi) << 32) | lo);
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), sel, 0xF, 0xF, false));
  }
}

// The other lane's value of x in each lane pair (a DPP quad permutation
// [1, 0, 3, 2]); both lanes of the pair must be active.
template <typename T>
__device__ __forceinline__ T pair_swap(T x) {
  constexpr int sel = 1 | (0 << 2) | (3 << 4) | (2 << 6);
  if constexpr (sizeof(T) == 8) {
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(b)), sel, 0xF, 0xF, false));
    const uint32_t hi = static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(b >> 32)), sel, 0xF, 0xF, false));
    return __builtin_bit_cast(T, (static_cast<uint64_t>(hi) << 32) | lo);
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), sel, 0xF, 0xF, false));
  }
}

// frame_cell with two lanes per cell, for groups of G::group threads and
// tiles of up to G::tile positions: lane 0 of a
// pair evaluates the cell's bottom i-face (row ii+1), its j-face at j and the
// pressure gradient along i, lane 1 the top i-face (row ii), the j-face at
// j+1 and the gradient along j; lane 0 completes the update.  The same
// expressions on the same operands as frame_cell.
// k_frame_edge's geometry for SEG owned positions per block and side: tiles
// of up to tile = SEG + 6 positions (3 rows), regions of pts points, group
// threads per side (two per cell), halo_lanes halo-entry lanes per side (3 SEG
// strip cells and up to two 9-cell corners), threads per block.
template <int SEG>
struct EdgeGeo {
  static constexpr int seg = SEG;
  static constexpr int tile = SEG + 6;
  static constexpr int pts = (tile + 6) * 9;
  static constexpr int group = (6 * tile + 63) / 64 * 64;
  static constexpr int halo_lanes = (3 * SEG + 18 + 63) / 64 * 64;
  static constexpr int halo_max = 3 * SEG + 18;  // halo entries per side
  static constexpr int threads = 2 * group;
  static_assert(pts <= 2 * group, "two region points per thread");
  static_assert(2 * halo_lanes <= threads, "halo entries fit the block");
  static_assert(2 * SEG <= threads, "one lane pair per owned position in the edge pass");
};
// Segment lengths: single-context launches use 42 positions (228 blocks at
// C768, one round on the device, where the block's latency sets the time).
// Batched launches run many rounds, where the blocks resident per CU set the
// throughput: blocks of 28 positions (512 threads; LDS 77 KB FP64, 39 KB
// FP32) fit two (FP64) or three (FP32) per CU, where blocks of 42 (106 KB /
// 53 KB) fit one or two.
constexpr int kSeg = 42;
constexpr int kBatchSeg = 28;

// Planes of frame_cell2's LDS region: u v g00 g01 [g10 g11] sina h hn (the
// a_gct planes g10, g11 only when the metric is not symmetric).
template <bool GSYM>
struct RegionPlanes {
  static constexpr int u = 0, v = 1, g00 = 2, g01 = 3, g10 = 4, g11 = 5;
  static constexpr int sina = GSYM ? 4 : 6, h = GSYM ? 5 : 7, hn = GSYM ? 6 : 8;
  // contravariant u and v of the point (remap.py:35-39), formed once when the
  // point is stored: every face stencil through the point reads them
  static constexpr int cu = GSYM ? 7 : 9, cv = GSYM ? 8 : 10, n = GSYM ? 9 : 11;
};

// `pre` runs after the region loads are issued and before the barrier.
template <typename T, bool GSYM, typename G, typename Pre>
__device__ __forceinline__ bool frame_cell2(const MarchArgs<T>& a, const HaloPlan<T>& plan, const FTile& ft,
                                            T (*s)[G::pts], int ltid, T& un, T& vn, int& ci, int& cj, int& cq,
                                            Pre&& pre) {
  using RP = RegionPlanes<GSYM>;
  const int nx = a.g.nx, N = a.g.N;
  const int t = ft.t, cr0 = ft.cr0, crows = ft.crows, cc0 = ft.cc0, ccols = ft.ccols;
  const int ra = cr0 - 3, ca = cc0 - 3, rcols = ccols + 6, npts = (crows + 6) * rcols;
  const Planes<T>& P = a.P;
  const uint32_t NS = static_cast<uint32_t>(N);
  const uint32_t fo = static_cast<uint32_t>(t) * NS * NS, mo = a.shared ? 0 : fo;
  const int q0 = ltid >> 1, kl = ltid & 1;
  const int q = min(q0, crows * ccols - 1);
  const int ii = cr0 + q / ccols, j = cc0 + q % ccols;
  const int li = ii - ra, lj = j - ca;
  const int fr = kl == 0 ? ii + 1 : ii, fc = kl == 0 ? j : j + 1;
  const uint32_t oc = static_cast<uint32_t>(ii) * NS + static_cast<uint32_t>(j);
  // global inputs besides the tile, issued before the tile loads
  const int evi_e = (fr == HALO || fr == N - HALO) && j <= N - 4 ? (t * 4 + (fr == HALO ? DS : DN)) * nx + (j - HALO) : 0;
  const int evj_e = (fc == HALO || fc == N - HALO) && ii >= HALO && ii < N - HALO
                        ? (t * 4 + (fc == HALO ? DW : DE)) * nx + (ii - HALO) : 0;
  const T fevi = ldg(a.evel, evi_e), fevj = ldg(a.evel, evj_e);
  const uint32_t fdi = static_cast<uint32_t>(fr) * NS + static_cast<uint32_t>(j);
  const uint32_t fdj = static_cast<uint32_t>(ii) * NS + static_cast<uint32_t>(fc);
  const T fdzi = ldg(P.dzx, P.dz_uniform ? 0 : fo + fdi), fdzj = ldg(P.dzy, P.dz_uniform ? 0 : fo + fdj);
  // cell inputs of the update: lane 0 loads its own, lane 1 the rest; a W /
  // E tile takes the constant ones from the strip copy (its 3 columns are
  // strip columns 3..5: a 128-byte line holds 5 of its rows, not 1)
  const uint32_t on = oc + NS;
  const int sd = ft.strip, sc = lj;  // strip column of j: the region starts at strip column 0
  auto cst = [&](const T* plane, uint32_t o, int k, int di, int dc) {
    return sd >= 0 ? strip_ld(P, N, t, sd, k, ii + di, sc + dc) : ldg(plane, o);
  };
  const T rpg = kl == 0 ? cst(P.rdx, mo + oc, SP_RDX, 0, 0) : cst(P.rdy, mo + oc, SP_RDY, 0, 0);
  T cin[6];
  if (kl == 0) {
    cin[0] = cst(P.cdy, mo + on, SP_CDY, 1, 0);
    cin[1] = cst(P.cdy, mo + oc, SP_CDY, 0, 0);
    cin[2] = cst(P.rda, mo + oc, SP_RDA, 0, 0);
    cin[3] = cst(P.ddx, mo + oc, SP_DDX, 0, 0);
    cin[4] = T(0); cin[5] = T(0);
  } else {
    cin[0] = cst(P.ddx, mo + oc + 1, SP_DDX, 0, 1);
    cin[1] = cst(P.rdx, mo + oc, SP_RDX, 0, 0);
    cin[2] = cst(P.rdy, mo + oc, SP_RDY, 0, 0);
    cin[3] = cst(P.af, fo + oc, SP_AF, 0, 0);
    cin[4] = ldg(a.up, fo + oc); cin[5] = ldg(a.vp, fo + oc);
  }
  // the region (at most G::pts <= 2 G::group points): each thread loads
  // points ltid and ltid + G::group, every load issued before the first
  // LDS write
  struct Pt {
    T u, v, g00, g01, g10, g11, sina, h, hn;
  };
  // the stencils read the tile's rows across the region and its columns
  // across the region, not the region's four 3 x 3 corner blocks
  auto read = [&](int e) {
    const int rr = rcols == 9 ? e / 9 : rcols == G::tile + 6 ? e / (G::tile + 6) : e / rcols;
    const int cc = e - rr * rcols;
    return (rr >= 3 && rr < 3 + crows) || (cc >= 3 && cc < 3 + ccols);
  };
  auto fetch = [&](int e, Pt& p) {
    const int rr = rcols == 9 ? e / 9 : rcols == G::tile + 6 ? e / (G::tile + 6) : e / rcols;
    const int cc = e - rr * rcols;
    const uint32_t o = fo + static_cast<uint32_t>(ra + rr) * NS + static_cast<uint32_t>(ca + cc);
    p.u = ldg(a.u, o);
    p.v = ldg(a.v, o);
    if (sd >= 0) {
      p.g00 = strip_ld(P, N, t, sd, SP_G00, ra + rr, cc);
      p.g01 = strip_ld(P, N, t, sd, SP_G01, ra + rr, cc);
      if (!GSYM) {
        p.g10 = strip_ld(P, N, t, sd, SP_G10, ra + rr, cc);
        p.g11 = strip_ld(P, N, t, sd, SP_G11, ra + rr, cc);
      }
      p.sina = strip_ld(P, N, t, sd, SP_SINA, ra + rr, cc);
    } else {
      p.g00 = ldg(P.g00, o);
      p.g01 = ldg(P.g01, o);
      if (!GSYM) {
        p.g10 = ldg(P.g10, o);
        p.g11 = ldg(P.g11, o);
      }
      p.sina = ldg(P.sina, o);
    }
    p.h = ldg(a.h, o);
    const int row = ra + rr, col = ca + cc;
    const bool hrow = row < HALO || row >= N - HALO, hcol = col < HALO || col >= N - HALO;
    if (hrow != hcol) {
      const int pq = halo_entry(nx, N, t, row, col);
      p.hn = plan.c0[pq] * ldg(a.hn, plan.s0[pq]) + plan.c1[pq] * ldg(a.hn, plan.s1[pq]);
    } else {
      p.hn = ldg(a.hn, o);
    }
  };
  auto put = [&](int e, const Pt& p) {
    s[RP::u][e] = p.u;
    s[RP::v][e] = p.v;
    s[RP::g00][e] = p.g00;
    s[RP::g01][e] = p.g01;
    if (!GSYM) {
      s[RP::g10][e] = p.g10;
      s[RP::g11][e] = p.g11;
    }
    s[RP::sina][e] = p.sina;
    s[RP::h][e] = p.h;
    s[RP::hn][e] = p.hn;
    s[RP::cu][e] = contra(p.g00, p.g01, p.u, p.v, p.sina);
    s[RP::cv][e] = GSYM ? contra(p.g01, p.g00, p.u, p.v, p.sina) : contra(p.g10, p.g11, p.u, p.v, p.sina);
  };
  const int e1 = ltid + G::group;
  const bool r0 = ltid < npts && read(ltid), r1 = e1 < npts && read(e1);
  Pt pt0, pt1;
  if (r0) fetch(ltid, pt0);
  if (r1) fetch(e1, pt1);
  if (r0) put(ltid, pt0);
  if (r1) put(e1, pt1);
  pre();
  __syncthreads();
  if (q0 >= crows * ccols) return false;  // whole pairs
  const T wn = T(1.0) - a.beta, wo = a.beta;

  // one face from its 6 points (frame_cell's expressions)
  auto face = [&](bool iface, int r, int c, T fev, T fdz, T& uc, T& vt, T& C) {
    const int e0 = iface ? (r - 3 - ra) * rcols + lj : li * rcols + (c - 3 - ca);
    const int es = iface ? rcols : 1;
    T wf[6], tf[6], hf[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const int e = e0 + m * es;
      wf[m] = s[iface ? RP::cu : RP::cv][e];
      tf[m] = s[iface ? RP::v : RP::u][e];
      hf[m] = s[RP::h][e];
    }
    const bool fedge = iface ? (r == HALO) || (r == N - HALO) : (c == HALO) || (c == N - HALO);
    iface_b(wf, tf, hf, fedge, fev, fdz, uc, vt, C);
  };
  // the i-face, then the j-face: a loop that is not unrolled, so that one
  // face's operands are live at a time
  T uci = T(0), vti = T(0), Ci = T(0), ucj = T(0), vtj = T(0), Cj = T(0);
#pragma unroll 1
  for (int f = 0; f < 2; ++f) {
    T uc, vt, C;
    if (f == 0)
      face(true, fr, j, fevi, fdzi, uc, vt, C);
    else
      face(false, ii, fc, fevj, fdzj, uc, vt, C);
    if (f == 0) {
      uci = uc; vti = vt; Ci = C;
    } else {
      ucj = uc; vtj = vt; Cj = C;
    }
  }

  // pressure gradient along i (lane 0) or j (lane 1) at (ii, j)
  const bool pi = kl == 0;
  const int p0 = pi ? (li - 3) * rcols + lj : li * rcols + lj - 3;
  const int ps = pi ? rcols : 1;
  T hb[7];
#pragma unroll
  for (int m = 0; m < 7; ++m) hb[m] = wn * s[RP::hn][p0 + m * ps] + wo * s[RP::h][p0 + m * ps];
  const T h0 = c6(hb[0], hb[1], hb[2], hb[3], hb[4], hb[5]);
  const T h1 = c6(hb[1], hb[2], hb[3], hb[4], hb[5], hb[6]);
  const T pgf = T(-9.80) * (rpg * (h1 - h0));

  // lane 1's faces (top i-face, j-face at j+1), gradient and inputs
  const T ucx_prev = pair_hi(uci), vcx_prev = pair_hi(vti), Cx_prev = pair_hi(Ci);
  const T vcy_n = pair_hi(ucj), ucy_n = pair_hi(vtj), Cy_n = pair_hi(Cj);
  const T pgfv = pair_hi(pgf);
  T cv[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) cv[m] = pair_hi(cin[m]);
  if (kl != 0) return false;
  const T ucx = uci, vcx = vti, Cx = Ci, pgfu = pgf;
  const T vcy = ucj, ucy = vtj, Cy = Cj;
  const int ec = li * rcols + lj;
  const T vct_c = s[RP::cv][ec];
  const T uct_c = s[RP::cu][ec];
  const T cdy_i = cin[0], cdy_prev = cin[1], rda = cin[2], ddx = cin[3];
  const T ddx_n = cv[0], rdx = cv[1], rdy = cv[2], af = cv[3];
  const T up = cv[4], vp = cv[5];
  const T zeta = rda * (((ucy * ddx + vcx * cdy_i) - ucy_n * ddx_n) - vcx_prev * cdy_prev);
  const T ku1 = T(0.5) * (ucx * ucx + vcx * vcx);
  const T ku0 = T(0.5) * (ucx_prev * ucx_prev + vcx_prev * vcx_prev);
  const T kv1 = T(0.5) * (vcy_n * vcy_n + ucy_n * ucy_n);
  const T kv0 = T(0.5) * (vcy * vcy + ucy * ucy);
  const T advx = zeta * vct_c - (ku1 - ku0) * rdx;
  const T advy = -zeta * uct_c - (kv1 - kv0) * rdy;
  const T visx = -(rdx * (Cx - Cx_prev));
  const T visy = -(rdy * (Cy_n - Cy));
  un = up + (((pgfu + advx) + af * vct_c) + visx) * a.dt;
  vn = vp + (((pgfv + advy) - af * uct_c) + visy) * a.dt;
  ci = ii;
  cj = j;
  cq = q;
  return true;
}

template <typename T, bool GSYM>
__device__ __forceinline__ void frame_body(const MarchArgs<T>& a, const HaloPlan<T>& plan, int blk,
                                           T (*s)[kTilePts]) {
  const int nx = a.g.nx, N = a.g.N;
  const int nr = (nx + kTile - 1) / kTile, ns = (nx - 6 + kTile - 1) / kTile;
  const int per = 2 * nr + 2 * ns;
  FTile ft;
  ft.t = blk / per;
  int k = blk - ft.t * per;
  if (k < 2 * nr) {
    const int g = k / nr, c = k - g * nr;
    ft.cr0 = g ? N - 2 * HALO : HALO;
    ft.crows = 3;
    ft.cc0 = HALO + kTile * c;
    ft.ccols = min(kTile, nx - kTile * c);
  } else {
    k -= 2 * nr;
    const int g = k / ns, c = k - g * ns;
    ft.cr0 = 2 * HALO + kTile * c;
    ft.crows = min(kTile, nx - 6 - kTile * c);
    ft.cc0 = g ? N - 2 * HALO : HALO;
    ft.ccols = 3;
  }
  T un, vn;
  int ci, cj, cq;
  if (frame_cell<T, GSYM>(a, plan, ft, s, static_cast<int>(threadIdx.x), un, vn, ci, cj, cq)) {
    const uint32_t o = (static_cast<uint32_t>(ft.t) * static_cast<uint32_t>(N) + static_cast<uint32_t>(ci)) *
                           static_cast<uint32_t>(N) + static_cast<uint32_t>(cj);
    *at32(a.u_out, o) = un;
    *at32(a.v_out, o) = vn;
  }
}

// The frame tiles (blocks 0 .. nft-1) and the new height's halo fill (the
// remaining blocks) in one launch.
template <typename T, bool GSYM>
__device__ __forceinline__ void frame_h_body(const MarchArgs<T>& a, const HaloArgs<T>& hh, int nft, int blk,
                                             T (*s)[kTilePts]) {
  if (blk < nft)
    frame_body<T, GSYM>(a, hh.plan, blk, s);
  else
    halo_scalar_body(hh, (blk - nft) * 256 + static_cast<int>(threadIdx.x));
}

template <typename T, bool GSYM>
__global__ __launch_bounds__(256) void k_frame(MarchArgs<T> a, HaloArgs<T> hh, int nft) {
  __shared__ T s[9][kTilePts];  // u v g00 g01 g10 g11 sina h hn
  frame_h_body<T, GSYM>(a, hh, nft, static_cast<int>(blockIdx.x), s);
}

// ---------------------------------------------------------------------------
// Cube-edge frame kernel: the frame cells, both halo fills and the next
// stage's edge pass in one launch.
//
// Block b takes cube edge e = b / nseg and the G::seg along-edge positions
// [p0, p1) of its first face edge (A = (t, d)); the second face edge (B,
// across the joint) owns the mirrored positions.  Each half of the block
// (G::group threads, two per cell: frame_cell2) evaluates the frame cells
// of one side over its owned positions widened by 3 on either end and keeps
// their new velocity in LDS.  A halo cell at an owned position takes its two
// interpolation sources from the other side's frame cells, at most 3
// positions away (checked when the plan is built), so the block forms the
// vector fill of the owned halo cells (cube.py:137-153) from LDS and writes
// the new height's fill from its LDS region, where frame_cell2 formed it
// from the plan.  Then it forms the next stage's averaged edge velocity /
// flux of the owned positions (communication.py:485-564), whose six stencil
// points across the edge are three frame cells and three halo cells it
// holds.  The halo corners are filled by the blocks holding positions 0 and
// nx-1 of the W / E face edges, whose strip cells they copy
// (cube.py:96-125).  Every value is the expression the separate kernels
// evaluate, on the same operands.
// ---------------------------------------------------------------------------

// One side of a cube-edge block: face edge (t, d), owned positions [lo, hi),
// evaluated positions [pa, pa + m) and the frame tile holding them.
struct ESide {
  int t, d, lo, hi, pa, m;
  FTile ft;
};

__device__ __forceinline__ ESide eside(int nx, int N, int t, int d, int lo, int hi) {
  ESide S;
  S.t = t;
  S.d = d;
  S.lo = lo;
  S.hi = hi;
  S.pa = max(lo - HALO, 0);
  S.m = min(hi + HALO, nx) - S.pa;
  S.ft.t = t;
  if (d == DS || d == DN) {
    S.ft.cr0 = d == DS ? HALO : N - 2 * HALO;
    S.ft.crows = 3;
    S.ft.cc0 = HALO + S.pa;
    S.ft.ccols = S.m;
  } else {
    S.ft.cr0 = HALO + S.pa;
    S.ft.crows = S.m;
    S.ft.cc0 = d == DW ? HALO : N - 2 * HALO;
    S.ft.ccols = 3;
    S.ft.strip = d == DW ? 0 : 1;
  }
  return S;
}

// Tile-local index of cell (i, j) of the side's tile.
__device__ __forceinline__ int eside_cell(const ESide& S, int i, int j) {
  return (i - S.ft.cr0) * S.ft.ccols + (j - S.ft.cc0);
}

// Region-local index (frame_cell's LDS layout) of point (i, j) of the side.
__device__ __forceinline__ int eside_point(const ESide& S, int i, int j) {
  return (i - (S.ft.cr0 - 3)) * (S.ft.ccols + 6) + (j - (S.ft.cc0 - 3));
}

template <typename T, bool GSYM, typename G>
struct FrameEdgeLds {
  T s[2][RegionPlanes<GSYM>::n][G::pts];  // the two sides' regions (frame_cell2's planes)
  T ur[2][3 * G::tile];  // new velocity of the evaluated frame cells
  T vr[2][3 * G::tile];
  T uh[2][3][G::seg];    // halo values at the owned positions, layers 1..3
  T vh[2][3][G::seg];
  T rot[kPlanRot][2 * G::halo_max];  // rotation coefficients of the halo entries, side by side
};

// Contravariant normal velocity (edge_body's stencil value) and new height of
// point (i, j) of side X: halo points take the block's fill values, interior
// points its frame values.
template <typename T, bool GSYM, typename G>
__device__ __forceinline__ void eside_sample(const FrameEdgeLds<T, GSYM, G>& L, int X, const ESide& S, int N, int i,
                                             int j, int pos, T& w, T& h) {
  using RP = RegionPlanes<GSYM>;
  const bool ns = S.d == DS || S.d == DN;
  const int x = ns ? i : j;
  T u, v;
  if (x < HALO || x >= N - HALO) {
    const int layer = x < HALO ? HALO - x : x - (N - HALO - 1);
    u = L.uh[X][layer - 1][pos - S.lo];
    v = L.vh[X][layer - 1][pos - S.lo];
  } else {
    const int c = eside_cell(S, i, j);
    u = L.ur[X][c];
    v = L.vr[X][c];
  }
  const int e = eside_point(S, i, j);
  const T g00 = L.s[X][RP::g00][e], g01 = L.s[X][RP::g01][e], sina = L.s[X][RP::sina][e];
  const T g10 = GSYM ? g01 : L.s[X][RP::g10][e], g11 = GSYM ? g00 : L.s[X][RP::g11][e];
  w = ns ? contra(g00, g01, u, v, sina) : contra(g10, g11, u, v, sina);
  h = L.s[X][RP::hn][e];
}

// Edge velocity (and the six heights) of side X at position pos: the six
// stencil points across the edge line (remap.py:42-55 at rows / columns 3,
// N-3), as edge_body evaluates them.
template <typename T, bool GSYM, typename G>
__device__ __forceinline__ T eside_vel(const FrameEdgeLds<T, GSYM, G>& L, int X, const ESide& S, int N, int pos, T* hv) {
  const bool ns = S.d == DS || S.d == DN;
  const int line = (S.d == DS || S.d == DW) ? HALO : N - HALO;
  T w[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int i = ns ? line - 3 + k : HALO + pos, j = ns ? HALO + pos : line - 3 + k;
    eside_sample<T, GSYM, G>(L, X, S, N, i, j, pos, w[k], hv[k]);
  }
  return T(0.5) * (e5(w[0], w[1], w[2], w[3], w[4]) + w5(w[1], w[2], w[3], w[4], w[5]));
}

// Upwind mass flux of side X through the edge (barotr_pure.py:29-49) with the
// averaged edge velocity avg, as edge_body evaluates it.
template <typename T>
__device__ __forceinline__ T eside_flux(const T* hv, T avg, T dz) {
  const T he = e5(hv[0], hv[1], hv[2], hv[3], hv[4]);
  const T hw = w5(hv[1], hv[2], hv[3], hv[4], hv[5]);
  const T c = celerity(c6(hv[0], hv[1], hv[2], hv[3], hv[4], hv[5]), dz);
  const T vc = avg + k49_over(c) * (he - hw);
  return dz * vc + (vc > T(0) ? he : hw) * vc;
}

// dz of side S's edge-line cell at position pos (edge_body's dz).
template <typename T>
__device__ __forceinline__ T eside_dz(const EdgeArgs<T>& en, const ESide& S, int pos) {
  const int N = en.g.N;
  const bool ns = S.d == DS || S.d == DN;
  const int line = (S.d == DS || S.d == DW) ? HALO : N - HALO;
  const size_t oc = ns ? en.g.at(S.t, line, HALO + pos) : en.g.at(S.t, HALO + pos, line);
  return ns ? dzx(en.P, oc) : dzy(en.P, oc);
}

template <typename T, bool GSYM, typename G>
__device__ __forceinline__ void frame_edge_body(const MarchArgs<T>& a, const EdgeArgs<T>& en, int blk,
                                                FrameEdgeLds<T, GSYM, G>& L) {
  using RP = RegionPlanes<GSYM>;
  const int nx = a.g.nx, N = a.g.N;
  const int tid = static_cast<int>(threadIdx.x);
  const int nseg = (nx + G::seg - 1) / G::seg;
  const int e = blk / nseg, p0 = (blk - e * nseg) * G::seg, p1 = min(p0 + G::seg, nx);
  const int ta = en.ue[e] >> 2, da = en.ue[e] & 3;
  const int tb = en.R.r[ta][da][0], db = en.R.r[ta][da][1], flip = en.R.r[ta][da][2];
  const ESide A = eside(nx, N, ta, da, p0, p1);
  const ESide B = eside(nx, N, tb, db, flip ? nx - p1 : p0, flip ? nx - p0 : p1);
  const HaloPlan<T>& plan = en.plan;
  const uint32_t NS = static_cast<uint32_t>(N), NN = NS * NS;

  // the halo entries of the owned positions and of the corners they carry
  // (threads 0 .. 2 G::halo_lanes - 1, side tid / G::halo_lanes): their plan data
  // is fetched with the frame loads, the rotation coefficients staged in LDS
  const int hs = tid >= G::halo_lanes ? 1 : 0, hk = tid - hs * G::halo_lanes;
  const ESide H = hs ? B : A;
  const int nown = H.hi - H.lo;
  const bool we = H.d == DW || H.d == DE;
  const int c_lo = (we && H.lo == 0) ? 1 : 0, c_hi = (we && H.hi == nx) ? 1 : 0;
  const bool hact = tid < 2 * G::halo_lanes && hk < 3 * nown + 9 * (c_lo + c_hi);
  int hq = 0, hlayer = 0, hpos = 0, hrow = 0, hcol = 0;
  if (hact) {
    if (hk < 3 * nown) {
      hlayer = hk / nown + 1;
      hpos = H.lo + hk % nown;
      hrow = H.d == DS ? HALO - hlayer : H.d == DN ? N - HALO - 1 + hlayer : HALO + hpos;
      hcol = H.d == DW ? HALO - hlayer : H.d == DE ? N - HALO - 1 + hlayer : HALO + hpos;
      hq = halo_entry(nx, N, H.t, hrow, hcol);
    } else {
      // corners 0 (S W), 1 (S E), 2 (N W), 3 (N E): 9 cells each
      const int kc = hk - 3 * nown;
      const bool north = (kc >= 9) || !c_lo;
      const int corner = (north ? 2 : 0) + (H.d == DE ? 1 : 0);
      hq = H.t * (12 * nx + 36) + 12 * nx + corner * 9 + (kc % 9);
    }
  }
  uint32_t hdst = 0, hs0 = 0, hs1 = 0;
  T hc0 = T(0), hc1 = T(0), R[kPlanRot];
  if (hact) {
    hdst = plan.dst[hq];
    hs0 = plan.s0[hq];
    hs1 = plan.s1[hq];
    hc0 = plan.c0[hq];
    hc1 = plan.c1[hq];
#pragma unroll
    for (int k = 0; k < kPlanRot; ++k) R[k] = plan.rot[static_cast<size_t>(k) * plan.n + hq];
  }
  auto stage_rot = [&]() {
    if (hact) {
#pragma unroll
      for (int k = 0; k < kPlanRot; ++k) L.rot[k][hs * G::halo_max + hk] = R[k];
    }
  };

  // frame cells of both sides
  {
    const int side = tid >= G::group ? 1 : 0;
    const ESide S = side ? B : A;
    T un, vn;
    int ci, cj, cq;
    if (frame_cell2<T, GSYM, G>(a, plan, S.ft, L.s[side], tid - side * G::group, un, vn, ci, cj, cq, stage_rot)) {
      L.ur[side][cq] = un;
      L.vr[side][cq] = vn;
      const bool ns = S.d == DS || S.d == DN;
      const int pos = ns ? cj - HALO : ci - HALO;
      // W / E face edges leave the corner frame blocks to the S / N ones
      if (pos >= S.lo && pos < S.hi && (ns || (pos >= HALO && pos < nx - HALO))) {
        const uint32_t o = static_cast<uint32_t>(S.t) * NN + static_cast<uint32_t>(ci) * NS + static_cast<uint32_t>(cj);
        *at32(a.u_out, o) = un;
        *at32(a.v_out, o) = vn;
      }
    }
  }
  __syncthreads();

  // halo fill: the velocity from the other side's frame cells in LDS, the new
  // height from the region (where frame_cell2 formed it from the plan) or,
  // on corner cells, from the other side's region values at the sources
  // edge lanes: lane pair k takes owned position p0 + k, its even lane side
  // A, its odd lane side B
  const bool eact = tid < 2 * (p1 - p0);
  const int ex = tid & 1, epa = p0 + (tid >> 1);
  const int epos = ex ? (flip ? nx - 1 - epa : epa) : epa;
  const ESide E = ex ? B : A;
  T edz = T(0);
  if (eact && en.last) edz = eside_dz(en, E, epos);
  if (hact) {
    const ESide Y = hs ? A : B;  // the sources lie on the other side
    const int y = hs ? 0 : 1;
    auto src = [&](uint32_t o, int& c, int& pnt) {
      const uint32_t r = o - static_cast<uint32_t>(Y.t) * NN;
      const int i = static_cast<int>(r / NS), j = static_cast<int>(r - (r / NS) * NS);
      c = min(max(eside_cell(Y, i, j), 0), 3 * Y.m - 1);
      pnt = min(max(eside_point(Y, i, j), 0), (Y.ft.crows + 6) * (Y.ft.ccols + 6) - 1);
    };
    int q0, q1, e0, e1;
    src(hs0, q0, e0);
    src(hs1, q1, e1);
    T Rl[kPlanRot];
#pragma unroll
    for (int k = 0; k < kPlanRot; ++k) Rl[k] = L.rot[k][hs * G::halo_max + hk];
    T uo, vo;
    halo_uv_combine(Rl, hc0, hc1, L.ur[y][q0], L.vr[y][q0], L.ur[y][q1], L.vr[y][q1], uo, vo);
    const T hh = hlayer ? L.s[hs][RP::hn][eside_point(H, hrow, hcol)]
                        : hc0 * L.s[y][RP::hn][e0] + hc1 * L.s[y][RP::hn][e1];
    *at32(a.u_out, hdst) = uo;
    *at32(a.v_out, hdst) = vo;
    *at32(a.h_out, hdst) = hh;
    if (hlayer) {
      L.uh[hs][hlayer - 1][hpos - H.lo] = uo;
      L.vh[hs][hlayer - 1][hpos - H.lo] = vo;
    }
  }
  __syncthreads();

  // the next stage's edge values of the owned positions: each lane its
  // side's one-sided value, the pair's sum (IEEE addition commutes, so both
  // lanes hold the same bits) halved
  if (eact) {
    T hv[6];
    const T vel = eside_vel<T, GSYM, G>(L, ex, E, N, epos, hv);
    const T avg = (vel + pair_swap(vel)) / T(2.0);
    const int o = (E.t * 4 + E.d) * nx + epos;
    en.evel[o] = avg;
    if (en.last) {
      const T f = eside_flux(hv, avg, edz);
      en.eflux[o] = (f + pair_swap(f)) / T(2.0);
    }
  }
}

template <typename T, bool GSYM>
__global__ __launch_bounds__(EdgeGeo<kSeg>::threads, 3) void k_frame_edge(MarchArgs<T> a, EdgeArgs<T> en) {
  __shared__ FrameEdgeLds<T, GSYM, EdgeGeo<kSeg>> L;
  frame_edge_body<T, GSYM, EdgeGeo<kSeg>>(a, en, static_cast<int>(blockIdx.x), L);
}

// Kernel arguments of one RK stage of one context.
template <typename T>
struct StageArgs {
  EdgeArgs<T> e, en;  // this stage's edge pass and the next stage's
  MarchArgs<T> ma, mb, mf;  // kernel A, kernel B, fused kernel and frame
  HaloArgs<T> hh, huv;
};

// ---------------------------------------------------------------------------
// Batched launches: several contexts, of any resolutions, advanced by one
// launch per kernel class.  Each member's blocks are contiguous; a block finds
// its member in a block -> member table and reads the member's arguments from
// a device table of StageArgs.
// ---------------------------------------------------------------------------
enum { BK_EDGE = 0, BK_MARCH_A = 1, BK_HALO = 2, BK_MARCH_B = 3, BK_FUSED = 4, BK_FRAME = 5, BK_HUVE = 6,
       BK_FEDGE = 7, BK_CLASSES = 8 };

template <typename T>
struct BatchLaunch {
  const StageArgs<T>* __restrict__ tab;  // [stage][variant][member]
  const int* __restrict__ nbb;           // substeps of each member
  const int* __restrict__ blk2m;         // block -> member, this kernel class
  const int* __restrict__ first;         // first block of each member, this class
  int count;                             // members in the table
  int stage, first_sub, parity;          // stage index; substep 0?; substep parity
};

// Variant 0 / 1 read the caller's input and write the output / the ping-pong
// field X; variant 2 / 3 read X / the output and write the output / X.  The
// choice by (nbb - 1 - k) parity makes every member finish its own substep
// count in its output buffers.
template <typename T>
__device__ __forceinline__ const StageArgs<T>& batch_entry(const BatchLaunch<T>& b, int m) {
  const int e = (b.first_sub ? 0 : 2) + (((b.nbb[m] - 1) & 1) ^ b.parity);
  return b.tab[(static_cast<size_t>(b.stage) * 4 + e) * b.count + m];
}

template <typename T>
__global__ __launch_bounds__(256) void k_edge_batch(BatchLaunch<T> b) {
  const int blk = static_cast<int>(blockIdx.x);
  const int m = b.blk2m[blk];
  // by reference: the routing table is indexed per lane, a private copy would
  // live in scratch memory
  const EdgeArgs<T>& a = batch_entry(b, m).e;
  edge_body(a, (blk - b.first[m]) * 256 + static_cast<int>(threadIdx.x));
}

template <typename T, bool GSYM, bool SAMEP>
__global__ __launch_bounds__(256) void k_march_a_batch(BatchLaunch<T> b) {
  __shared__ T s_v[4][kRowBuf];
  __shared__ T s_h[4][kRowBuf];
  const int L = xcd_block();
  const int m = b.blk2m[L];
  const MarchArgs<T> a = batch_entry(b, m).ma;
  march_a_body<T, GSYM, SAMEP>(a, L - b.first[m], s_v, s_h);
}

template <typename T, bool GSYM, bool SAMEP>
__global__ __launch_bounds__(256) void k_march_b_batch(BatchLaunch<T> b) {
  __shared__ T s_vct[4][kRowBuf];
  __shared__ T s_u[4][kRowBuf];
  __shared__ T s_h[4][kRowBuf];
  __shared__ T s_n[4][kRowBuf];
  const int L = xcd_block();
  const int m = b.blk2m[L];
  const MarchArgs<T> a = batch_entry(b, m).mb;
  march_b_body<T, GSYM, SAMEP>(a, L - b.first[m], s_vct, s_u, s_h, s_n);
}

template <typename T, bool GSYM, bool SAMEP>
__global__ __launch_bounds__(256, 2) void k_fused_batch(BatchLaunch<T> b) {
  __shared__ FRing<T> rings[4];
  const int L = xcd_block();
  const int m = b.blk2m[L];
  const MarchArgs<T> a = batch_entry(b, m).mf;
  fused_body<T, GSYM, SAMEP>(a, L - b.first[m], rings);
}

template <typename T, bool GSYM>
__global__ __launch_bounds__(256) void k_frame_batch(BatchLaunch<T> b) {
  __shared__ T s[9][kTilePts];
  const int blk = static_cast<int>(blockIdx.x);
  const int m = b.blk2m[blk];
  const StageArgs<T>& e = batch_entry(b, m);
  frame_h_body<T, GSYM>(e.mf, e.hh, 6 * frame_tiles(e.mf.g.nx), blk - b.first[m], s);
}

template <typename T>
__global__ __launch_bounds__(256) void k_halo_scalar_batch(BatchLaunch<T> b) {
  const int blk = static_cast<int>(blockIdx.x);
  const int m = b.blk2m[blk];
  const HaloArgs<T> a = batch_entry(b, m).hh;
  halo_scalar_body(a, (blk - b.first[m]) * 256 + static_cast<int>(threadIdx.x));
}

// Each member's velocity halo fill and its next stage's edge pass (as
// k_halo_uv_edge).
template <typename T>
__global__ __launch_bounds__(256) void k_halo_uv_edge_batch(BatchLaunch<T> b) {
  const int blk = static_cast<int>(blockIdx.x);
  const int m = b.blk2m[blk];
  const StageArgs<T>& e = batch_entry(b, m);
  const int local = blk - b.first[m], nh = (e.huv.plan.n + 255) / 256;
  if (local < nh)
    halo_vector_body(e.huv, local * 256 + static_cast<int>(threadIdx.x));
  else
    edge_body(e.en, (local - nh) * 256 + static_cast<int>(threadIdx.x));
}

// Each member's cube-edge frame launch (as k_frame_edge).
// Waves per SIMD of the batched frame launch: its blocks per CU (LDS-limited:
// two FP64 ones, one with the four a_gct planes; three FP32 ones, two with
// them) times their 8 waves, over 4 SIMDs; the bound keeps the registers
// within that occupancy.
template <typename T, bool GSYM>
constexpr int frame_batch_waves() {
  return sizeof(T) == 8 ? (GSYM ? 4 : 2) : (GSYM ? 6 : 4);
}

template <typename T, bool GSYM>
__global__ __launch_bounds__(EdgeGeo<kBatchSeg>::threads, (frame_batch_waves<T, GSYM>()))
void k_frame_edge_batch(BatchLaunch<T> b) {
  using G = EdgeGeo<kBatchSeg>;
  __shared__ FrameEdgeLds<T, GSYM, G> L;
  const int blk = static_cast<int>(blockIdx.x);
  const int m = b.blk2m[blk];
  const StageArgs<T>& e = batch_entry(b, m);
  frame_edge_body<T, GSYM, G>(e.mf, e.en, blk - b.first[m], L);
}

// ---------------------------------------------------------------------------
// standalone edge averaging (communication.py:485-564), two phases
// ---------------------------------------------------------------------------
template <typename T>
struct EdgeAvgArgs {
  Geo g;
  Route R;
  T *u, *v, *buf;
};

__device__ __forceinline__ void edge_line_cell(int N, int d, int r, int& i, int& j) {
  if (d == DE) { i = 3 + r; j = N - 3; }
  else if (d == DW) { i = 3 + r; j = 3; }
  else if (d == DN) { i = N - 3; j = 3 + r; }
  else { i = 3; j = 3 + r; }
}

template <typename T>
__global__ __launch_bounds__(256) void k_edge_avg_read(EdgeAvgArgs<T> a) {
  const Geo g = a.g;
  const int nx = g.nx;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 24 * nx) return;
  const int t = q / (4 * nx), d = (q / nx) % 4, r = q % nx;
  const int tb = a.R.r[t][d][0], db = a.R.r[t][d][1];
  const int rr = a.R.r[t][d][2] ? nx - 1 - r : r;
  int i, j, bi, bj;
  edge_line_cell(g.N, d, r, i, j);
  edge_line_cell(g.N, db, rr, bi, bj);
  const T mine = (d == DE || d == DW) ? a.v[g.at(t, i, j)] : a.u[g.at(t, i, j)];
  const T nb = (db == DE || db == DW) ? a.v[g.at(tb, bi, bj)] : a.u[g.at(tb, bi, bj)];
  a.buf[q] = (mine + nb) / T(2.0);
}

template <typename T>
__global__ __launch_bounds__(256) void k_edge_avg_write(EdgeAvgArgs<T> a) {
  const Geo g = a.g;
  const int nx = g.nx;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 24 * nx) return;
  const int t = q / (4 * nx), d = (q / nx) % 4, r = q % nx;
  int i, j;
  edge_line_cell(g.N, d, r, i, j);
  if (d == DE || d == DW) a.v[g.at(t, i, j)] = a.buf[q];
  else a.u[g.at(t, i, j)] = a.buf[q];
}

// ---------------------------------------------------------------------------
// diagnostics over the full padded arrays (the reference's returned fields)
// ---------------------------------------------------------------------------
template <typename T>
struct DiagArgs {
  Geo g;
  Planes<T> P;
  const T *h, *u, *v, *evel;
  T *cx, *cy, *uct, *vct, *ucx, *ucy, *vcx, *vcy, *advx, *advy;
};

template <typename T>
struct DiagCell {
  const Geo& g;
  const Planes<T>& P;
  const T *h, *u, *v, *evel;
  int t;

  __device__ T U(int i, int j) const { return uct_at(P, u, v, g.at(t, i, j)); }
  __device__ T V(int i, int j) const { return vct_at(P, u, v, g.at(t, i, j)); }
  __device__ T Hh(int i, int j) const { return h[g.at(t, i, j)]; }

  __device__ T cx(int i, int j) const {
    const int N = g.N;
    T s = T(0);
    if (i >= 3 && i < N - 2 && j >= 3 && j < N - 2)
      s = c6(Hh(i - 3, j), Hh(i - 2, j), Hh(i - 1, j), Hh(i, j), Hh(i + 1, j), Hh(i + 2, j));
    return celerity(s, dzx(P, g.at(t, i, j)));
  }
  __device__ T cy(int i, int j) const {
    const int N = g.N;
    T s = T(0);
    if (i >= 3 && i < N - 2 && j >= 3 && j < N - 2)
      s = c6(Hh(i, j - 3), Hh(i, j - 2), Hh(i, j - 1), Hh(i, j), Hh(i, j + 1), Hh(i, j + 2));
    return celerity(s, dzy(P, g.at(t, i, j)));
  }
};

// The diagnostics in three passes over the padded arrays.  Each pass stores
// outputs that the next reads instead of recomputing them at every
// neighbour; every value is the reference's expression in its order.
// Pass 1: celerity and contravariant velocity.
template <typename T>
__global__ __launch_bounds__(256) void k_diag1(DiagArgs<T> a) {
  const Geo g = a.g;
  const int N = g.N;
  const size_t NN = static_cast<size_t>(N) * N;
  const size_t q = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= 6 * NN) return;
  const int t = static_cast<int>(q / NN);
  const int i = static_cast<int>((q % NN) / N);
  const int j = static_cast<int>(q % N);
  const DiagCell<T> C{g, a.P, a.h, a.u, a.v, a.evel, t};
  a.cx[q] = C.cx(i, j);
  a.cy[q] = C.cy(i, j);
  a.uct[q] = C.U(i, j);
  a.vct[q] = C.V(i, j);
}

// Pass 2: corrected C-grid velocities (remap.py:42-55, barotr_pure.py:29-34)
// and upwind D-grid velocities (remap.py:74-95) from the stored uct, vct, cx, cy.
template <typename T>
__global__ __launch_bounds__(256) void k_diag2(DiagArgs<T> a) {
  const Geo g = a.g;
  const int N = g.N;
  const size_t NN = static_cast<size_t>(N) * N;
  const size_t q = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= 6 * NN) return;
  const int t = static_cast<int>(q / NN);
  const int i = static_cast<int>((q % NN) / N);
  const int j = static_cast<int>(q % N);
  const size_t f0 = static_cast<size_t>(t) * NN;
  auto U = [&](int ii, int jj) { return a.uct[f0 + static_cast<size_t>(ii) * N + jj]; };
  auto V = [&](int ii, int jj) { return a.vct[f0 + static_cast<size_t>(ii) * N + jj]; };
  auto Hh = [&](int ii, int jj) { return a.h[f0 + static_cast<size_t>(ii) * N + jj]; };
  T ucx = T(0), vcx = T(0), vcy = T(0), ucy = T(0);
  if (i >= 3 && i < N - 2) {
    T pre;
    if ((i == 3 || i == N - 3) && j >= 3 && j < N - 3)
      pre = a.evel[(t * 4 + (i == 3 ? DS : DN)) * g.nx + (j - 3)];
    else
      pre = T(0.5) * (e5(U(i - 3, j), U(i - 2, j), U(i - 1, j), U(i, j), U(i + 1, j)) +
                      w5(U(i - 2, j), U(i - 1, j), U(i, j), U(i + 1, j), U(i + 2, j)));
    const T he = e5(Hh(i - 3, j), Hh(i - 2, j), Hh(i - 1, j), Hh(i, j), Hh(i + 1, j));
    const T hw = w5(Hh(i - 2, j), Hh(i - 1, j), Hh(i, j), Hh(i + 1, j), Hh(i + 2, j));
    ucx = pre + k49_over(a.cx[q]) * (he - hw);
    const size_t b = f0 + j;
    auto Vv = [&](int ii) { return a.v[b + static_cast<size_t>(ii) * N]; };
    vcx = pre > T(0) ? e5(Vv(i - 3), Vv(i - 2), Vv(i - 1), Vv(i), Vv(i + 1))
                     : w5(Vv(i - 2), Vv(i - 1), Vv(i), Vv(i + 1), Vv(i + 2));
  }
  if (j >= 3 && j < N - 2) {
    T pre;
    if ((j == 3 || j == N - 3) && i >= 3 && i < N - 3)
      pre = a.evel[(t * 4 + (j == 3 ? DW : DE)) * g.nx + (i - 3)];
    else
      pre = T(0.5) * (e5(V(i, j - 3), V(i, j - 2), V(i, j - 1), V(i, j), V(i, j + 1)) +
                      w5(V(i, j - 2), V(i, j - 1), V(i, j), V(i, j + 1), V(i, j + 2)));
    const T hn = e5(Hh(i, j - 3), Hh(i, j - 2), Hh(i, j - 1), Hh(i, j), Hh(i, j + 1));
    const T hs = w5(Hh(i, j - 2), Hh(i, j - 1), Hh(i, j), Hh(i, j + 1), Hh(i, j + 2));
    vcy = pre + k49_over(a.cy[q]) * (hn - hs);
    const size_t b = f0 + static_cast<size_t>(i) * N;
    auto Uu = [&](int jj) { return a.u[b + jj]; };
    ucy = pre > T(0) ? e5(Uu(j - 3), Uu(j - 2), Uu(j - 1), Uu(j), Uu(j + 1))
                     : w5(Uu(j - 2), Uu(j - 1), Uu(j), Uu(j + 1), Uu(j + 2));
  }
  a.ucx[q] = ucx;
  a.vcx[q] = vcx;
  a.vcy[q] = vcy;
  a.ucy[q] = ucy;
}

// Pass 3: vorticity and kinetic-energy advection (barotr_pure.py:78-97,
// agrid.py:32-41) from the stored velocities.
template <typename T>
__global__ __launch_bounds__(256) void k_diag3(DiagArgs<T> a) {
  const Geo g = a.g;
  const int N = g.N;
  const size_t NN = static_cast<size_t>(N) * N;
  const size_t q = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= 6 * NN) return;
  const int i = static_cast<int>((q % NN) / N);
  const int j = static_cast<int>(q % N);
  const Planes<T>& P = a.P;
  auto ku = [&](size_t o) { return T(0.5) * (a.ucx[o] * a.ucx[o] + a.vcx[o] * a.vcx[o]); };
  auto kv = [&](size_t o) { return T(0.5) * (a.vcy[o] * a.vcy[o] + a.ucy[o] * a.ucy[o]); };
  T ax = T(0), ay = T(0);
  if (i < N - 1 || j < N - 1) {
    T z = T(0);
    if (i < N - 1 && j < N - 1)
      z = P.rda[q] * (((a.ucy[q] * P.ddx[q] + a.vcx[q + N] * P.cdy[q + N]) - a.ucy[q + 1] * P.ddx[q + 1]) -
                      a.vcx[q] * P.cdy[q]);
    if (i < N - 1) ax = z * a.vct[q] - (ku(q + N) - ku(q)) * P.rdx[q];
    if (j < N - 1) ay = -z * a.uct[q] - (kv(q + 1) - kv(q)) * P.rdy[q];
  }
  a.advx[q] = ax;
  a.advy[q] = ay;
}

// ---------------------------------------------------------------------------
// global integrals in binary64 with compensated (Neumaier) summation
//   mass   = sum_interior (dzph + h) * da
//   energy = sum_interior da * (1/2 (dzph + h) |u|^2 + 1/2 g h^2)
// with da = 1 / rda and |u|^2 = u_i G^{ij} u_j (G^{ij} = a_gct), whatever
// the precision of the state.
// ---------------------------------------------------------------------------
struct Comp {
  double s, c;
};

__device__ __forceinline__ void comp_add(Comp& a, double x) {
  const double t = a.s + x;
  if (fabs(a.s) >= fabs(x)) a.c += (a.s - t) + x;
  else a.c += (x - t) + a.s;
  a.s = t;
}

__device__ __forceinline__ Comp comp_merge(Comp a, Comp b) {
  comp_add(a, b.s);
  a.c += b.c;
  return a;
}

template <typename T>
struct IntegralArgs {
  Geo g;
  Planes<T> P;
  const T *h, *u, *v;
  double* partial;  // 4 doubles per block: mass (s, c), energy (s, c)
};

template <typename T>
__global__ __launch_bounds__(256) void k_integrals(IntegralArgs<T> a) {
  const Geo g = a.g;
  const int nx = g.nx, N = g.N;
  const size_t cells = 6 * static_cast<size_t>(nx) * nx;
  Comp m{0.0, 0.0}, e{0.0, 0.0};
  for (size_t q = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; q < cells;
       q += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int t = static_cast<int>(q / (static_cast<size_t>(nx) * nx));
    const int r = static_cast<int>(q % (static_cast<size_t>(nx) * nx));
    const int i = 3 + r / nx, j = 3 + r % nx;
    const size_t o = g.at(t, i, j);
    const double da = 1.0 / static_cast<double>(a.P.rda[o]);
    const double depth = static_cast<double>(a.P.dz_uniform ? a.P.dzx_s : a.P.dzx[o]) + static_cast<double>(a.h[o]);
    const double uu = a.u[o], vv = a.v[o];
    const double q2 = uu * (static_cast<double>(a.P.g00[o]) * uu + static_cast<double>(a.P.g01[o]) * vv) +
                      vv * (static_cast<double>(a.P.g10[o]) * uu + static_cast<double>(a.P.g11[o]) * vv);
    const double hh = a.h[o];
    comp_add(m, depth * da);
    comp_add(e, da * (0.5 * depth * q2 + 0.5 * 9.8 * hh * hh));
  }
  __shared__ Comp sm[256], se[256];
  sm[threadIdx.x] = m;
  se[threadIdx.x] = e;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      sm[threadIdx.x] = comp_merge(sm[threadIdx.x], sm[threadIdx.x + w]);
      se[threadIdx.x] = comp_merge(se[threadIdx.x], se[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a.partial[4 * blockIdx.x + 0] = sm[0].s;
    a.partial[4 * blockIdx.x + 1] = sm[0].c;
    a.partial[4 * blockIdx.x + 2] = se[0].s;
    a.partial[4 * blockIdx.x + 3] = se[0].c;
  }
}

__global__ __launch_bounds__(256) void k_integrals_final(const double* partial, int nblocks, double* out) {
  Comp m{0.0, 0.0}, e{0.0, 0.0};
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
    m = comp_merge(m, Comp{partial[4 * b + 0], partial[4 * b + 1]});
    e = comp_merge(e, Comp{partial[4 * b + 2], partial[4 * b + 3]});
  }
  __shared__ Comp sm[256], se[256];
  sm[threadIdx.x] = m;
  se[threadIdx.x] = e;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      sm[threadIdx.x] = comp_merge(sm[threadIdx.x], sm[threadIdx.x + w]);
      se[threadIdx.x] = comp_merge(se[threadIdx.x], se[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sm[0].s + sm[0].c;
    out[1] = se[0].s + se[0].c;
  }
}

// ---------------------------------------------------------------------------
// const repacking (binary64 reference layouts -> context planes of type T)
// ---------------------------------------------------------------------------
template <typename T>
struct RepackArgs {
  size_t n;  // cells
  const double *dzx, *dzy, *rdx, *rdy, *af, *coef, *c2l, *l2c, *gct, *sina, *rda, *cdy, *ddx;
  const int *li, *lj;
  T *o_dzx, *o_dzy, *o_rdx, *o_rdy, *o_af, *o_rda, *o_cdy, *o_ddx, *o_sina;
  T *o_g00, *o_g01, *o_g10, *o_g11, *o_c00, *o_c01, *o_c10, *o_c11, *o_l00, *o_l01, *o_l10, *o_l11;
  T *o_k0, *o_k1;
  T *o_gp, *o_rxy, *o_rdd;  // pair planes, 2 n elements each
  int *o_li, *o_lj;
};

template <typename T>
__global__ __launch_bounds__(256) void k_repack(RepackArgs<T> a) {
  const size_t q = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= a.n) return;
  a.o_dzx[q] = T(a.dzx[q]);
  a.o_dzy[q] = T(a.dzy[q]);
  a.o_rdx[q] = T(a.rdx[q]);
  a.o_rdy[q] = T(a.rdy[q]);
  a.o_af[q] = T(a.af[q]);
  a.o_rda[q] = T(a.rda[q]);
  a.o_cdy[q] = T(a.cdy[q]);
  a.o_ddx[q] = T(a.ddx[q]);
  a.o_sina[q] = T(a.sina[q]);
  a.o_g00[q] = T(a.gct[4 * q + 0]);
  a.o_g01[q] = T(a.gct[4 * q + 1]);
  a.o_g10[q] = T(a.gct[4 * q + 2]);
  a.o_g11[q] = T(a.gct[4 * q + 3]);
  a.o_c00[q] = T(a.c2l[4 * q + 0]);
  a.o_c01[q] = T(a.c2l[4 * q + 1]);
  a.o_c10[q] = T(a.c2l[4 * q + 2]);
  a.o_c11[q] = T(a.c2l[4 * q + 3]);
  a.o_l00[q] = T(a.l2c[4 * q + 0]);
  a.o_l01[q] = T(a.l2c[4 * q + 1]);
  a.o_l10[q] = T(a.l2c[4 * q + 2]);
  a.o_l11[q] = T(a.l2c[4 * q + 3]);
  a.o_k0[q] = T(a.coef[2 * q + 0]);
  a.o_k1[q] = T(a.coef[2 * q + 1]);
  a.o_gp[2 * q] = T(a.gct[4 * q + 0]);
  a.o_gp[2 * q + 1] = T(a.gct[4 * q + 1]);
  a.o_rxy[2 * q] = T(a.rdx[q]);
  a.o_rxy[2 * q + 1] = T(a.rdy[q]);
  a.o_rdd[2 * q] = T(a.rda[q]);
  a.o_rdd[2 * q + 1] = T(a.ddx[q]);
  a.o_li[q] = a.li[q];
  a.o_lj[q] = a.lj[q];
}

constexpr int kNumPlanes = 23;  // T planes in the context (see Planes)
constexpr int kPairPlanes = 3;  // interleaved pair planes after them, 2 n elements each

// Elements of the W / E strip copy (after the pair planes).
__host__ __device__ inline size_t strip_elems(int N) { return static_cast<size_t>(12) * kStripPlanes * N * 9; }

// The W / E strip copy of the constant planes (Planes::strip).
template <typename T>
__global__ __launch_bounds__(256) void k_pack_strips(Planes<T> P, int N, T* out) {
  const size_t q = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= strip_elems(N)) return;
  const int c = static_cast<int>(q % 9);
  const size_t r1 = q / 9;
  const int i = static_cast<int>(r1 % N);
  const size_t r2 = r1 / N;
  const int k = static_cast<int>(r2 % kStripPlanes);
  const int ts = static_cast<int>(r2 / kStripPlanes);
  const int t = ts >> 1, side = ts & 1;
  const T* const src[kStripPlanes] = {P.g00, P.g01, P.g10, P.g11, P.sina, P.cdy, P.rda, P.ddx, P.rdx, P.rdy, P.af};
  const size_t o = (static_cast<size_t>(t) * N + i) * N + (side ? N - 9 + c : c);
  out[q] = src[k][o];
}

inline unsigned blocks_for(size_t n, unsigned per = 256) {
  return static_cast<unsigned>((n + per - 1) / per);
}

}  // namespace

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
// Device storage of one resolution's consts (planes, index tables, halo
// plan).  Contexts created with sw_create_like share it; it is freed with the
// last of them.
struct GridMem {
  void* planes = nullptr;
  int* locs = nullptr;
  void* plan = nullptr;
  ~GridMem() {
    if (planes) (void)hipFree(planes);
    if (locs) (void)hipFree(locs);
    if (plan) (void)hipFree(plan);
  }
};

struct sw_ctx {
  int nx = 0, N = 0, dtype = SW_F64;
  int device = 0;                // device current at sw_create
  double* noalloc = nullptr;     // this device's address of the file-scope `scratch` array
  hipStream_t stream = nullptr;
  size_t cells = 0;        // 6 * N * N
  std::shared_ptr<GridMem> grid;  // owner of planes, locs and plan below
  void* planes = nullptr;  // kNumPlanes T planes, then kPairPlanes pair planes
  int* locs = nullptr;     // loc_i, loc_j planes
  void* scratch = nullptr; // 12 T fields: S1, S2 (stage scratch), X, Y (substep ping-pong)
  void* edges = nullptr;   // evel, eflux, edge-average buffer, store sink
  double* ibuf = nullptr;  // integral partials (4 per block) and the result
  double dzx_s = 0, dzy_s = 0;
  int dz_uniform = 0;
  int shared = 0;          // rdx, rdy, rda, c_dy, d_dx identical on every face
  int gsym = 0;            // a_gct symmetric with equal diagonal everywhere
  void* plan = nullptr;    // halo plan (uint32 dst, s0, s1; T c0, c1, rot[12]), plan_n entries
  int plan_n = 0;
  int frame_edge_ok = 0;   // the plan's sources fit k_frame_edge (within 3 positions)
  int frame_edge = 0;      // fused path: k_frame_edge (1) or k_frame + k_halo_uv_edge (0)
  int rows_a = 32;         // rows per wave of k_march_a
  int rows_b = 32;         // rows per wave of k_march_b
  int rows_f = 32;         // rows per wave of k_fused (sw_create sets default_rows_f); 0 selects k_march_a + k_march_b
  int rows_f2 = 0;         // > 0: split bands, the short height (the long one is rows_f)
  hipStream_t capture = nullptr;  // private stream used only to record graphs
  struct GraphEntry {
    int rk, pairs;
    double dtb;
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> graphs;  // recorded substep-pair sequences
  struct ProfRec {
    int cls, reps;
    hipEvent_t a, b;
  };
  std::vector<ProfRec>* prof = nullptr;  // set only by sw_profile
  int prof_reps = 1;                     // back-to-back launches per timed launch
  // launch-duration stamps of the fused kernel (sw_profile_graph only): the
  // buffer, its next free slot and end, and (first slot, waves) per launch
  mutable unsigned long long *stamp_base = nullptr, *stamp_next = nullptr, *stamp_end = nullptr;
  mutable std::vector<std::pair<size_t, int>> stamp_launches;
};

// Reusable state of sw_batch_run: the device copy of the member tables and
// two pinned host staging buffers used in turn, so a run never waits for the
// previous run's staging copy (only for the one before it).
struct sw_batch {
  void* dev = nullptr;
  size_t dev_cap = 0;
  void* host[2] = {nullptr, nullptr};
  size_t host_cap[2] = {0, 0};
  hipEvent_t staged[2] = {nullptr, nullptr};  // recorded after each buffer's last staging copy
  int slot = 0;                  // staging buffer of the next run
  hipStream_t stream = nullptr;  // stream of the last run
  bool used = false;
  hipStream_t capture = nullptr;  // private stream used only to record graphs
  struct Graph {
    std::vector<long long> sig;  // table layout, member count, grids
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;     // recorded substep pairs
  struct Retired {
    hipGraphExec_t exec;
    hipEvent_t done;  // recorded when the graph left the cache
  };
  std::vector<Retired> retired;  // evicted graphs, destroyed once `done` completes
  void reap(bool wait) {
    std::vector<Retired> keep;
    for (auto& r : retired) {
      if (wait) (void)hipEventSynchronize(r.done);
      if (hipEventQuery(r.done) == hipSuccess) {
        (void)hipGraphExecDestroy(r.exec);
        (void)hipEventDestroy(r.done);
      } else {
        keep.push_back(r);
      }
    }
    retired.swap(keep);
  }
};

namespace {

template <typename T>
Planes<T> planes_of(const sw_ctx* c) {
  const T* p = static_cast<const T*>(c->planes);
  const size_t n = c->cells;
  Planes<T> P;
  const T** slots[kNumPlanes] = {&P.dzx, &P.dzy, &P.rdx, &P.rdy, &P.af, &P.rda, &P.cdy, &P.ddx,
                                 &P.sina, &P.g00, &P.g01, &P.g10, &P.g11, &P.c00, &P.c01, &P.c10,
                                 &P.c11, &P.l00, &P.l01, &P.l10, &P.l11, &P.k0, &P.k1};
  for (int k = 0; k < kNumPlanes; ++k) *slots[k] = p + k * n;
  P.gp = p + kNumPlanes * n;
  P.rxy = P.gp + 2 * n;
  P.rdd = P.rxy + 2 * n;
  P.strip = P.rdd + 2 * n;
  P.li = c->locs;
  P.lj = c->locs + n;
  P.dzx_s = T(c->dzx_s);
  P.dzy_s = T(c->dzy_s);
  P.dz_uniform = c->dz_uniform;
  return P;
}

// Kernel classes timed by sw_profile.
enum { K_EDGE = 0, K_MARCH_A = 1, K_HALO_H = 2, K_MARCH_B = 3, K_HALO_UV = 4, K_DIAG = 5, K_FUSED = 6,
       K_FRAME = 7, K_CLASSES = 8 };

// Launch `kernel` on the context's stream.  Under sw_profile the launch is
// repeated prof_reps times (every launch of the step is idempotent) and each
// launch carries a start / stop hipEvent on its own dispatch
// (hipExtLaunchKernelGGL), so the events time the kernel itself with no
// marker packets between kernels.
template <typename K, typename... A>
int launch_n(const sw_ctx* c, int cls, K kernel, dim3 grid, int threads, const A&... args) {
  if (!c->prof) {
    hipLaunchKernelGGL(kernel, grid, dim3(threads), 0, c->stream, args...);
    SW_HIP(hipGetLastError());
    return SW_OK;
  }
  for (int r = 0; r < c->prof_reps; ++r) {
    sw_ctx::ProfRec rec{cls, 1, nullptr, nullptr};
    SW_HIP(hipEventCreate(&rec.a));
    SW_HIP(hipEventCreate(&rec.b));
    c->prof->push_back(rec);
    hipExtLaunchKernelGGL(kernel, grid, dim3(threads), 0, c->stream, rec.a, rec.b, 0, args...);
    SW_HIP(hipGetLastError());
  }
  return SW_OK;
}

template <typename K, typename... A>
int launch(const sw_ctx* c, int cls, K kernel, dim3 grid, const A&... args) {
  return launch_n(c, cls, kernel, grid, 256, args...);
}

template <typename T>
T* scratch_field(const sw_ctx* c, int k) {  // k in [0, 12)
  return static_cast<T*>(c->scratch) + static_cast<size_t>(k) * c->cells;
}

// Edge buffers (each >= 256 elements): the averaged edge velocity and flux of
// RK stage index s at 2 s and 2 s + 1 (s < 3: the velocity halo launch of a
// stage forms the next stage's while the stage kernels of this one may still
// be read by the diagnostics), 6 scratch of sw_edge_avg, 7 store sink of the
// stage kernels.
constexpr int kEdgeBufs = 8;
constexpr int kEdgeScratch = 6, kEdgeSink = 7;
size_t edge_stride(const sw_ctx* c) {
  const size_t n = 24 * static_cast<size_t>(c->nx);
  return n < 256 ? 256 : n;
}

template <typename T>
T* edge_buf(const sw_ctx* c, int k) {
  return static_cast<T*>(c->edges) + static_cast<size_t>(k) * edge_stride(c);
}

Geo geo_of(const sw_ctx* c) { return Geo{c->nx, c->N}; }

// Plan arrays inside the one allocation c->plan: three uint32 arrays, then
// (2 + kPlanRot) arrays of T, each starting on a 256-byte boundary.
size_t plan_stride(size_t n, size_t esize) { return ((n * esize + 255) / 256) * 256; }

template <typename T>
HaloPlan<T> plan_of(const sw_ctx* c) {
  const size_t n = static_cast<size_t>(c->plan_n);
  char* b = static_cast<char*>(c->plan);
  const size_t su = plan_stride(n, 4), st = plan_stride(n, sizeof(T));
  HaloPlan<T> p;
  p.dst = reinterpret_cast<const uint32_t*>(b);
  p.s0 = reinterpret_cast<const uint32_t*>(b + su);
  p.s1 = reinterpret_cast<const uint32_t*>(b + 2 * su);
  p.c0 = reinterpret_cast<const T*>(b + 3 * su);
  p.c1 = reinterpret_cast<const T*>(b + 3 * su + st);
  p.rot = reinterpret_cast<const T*>(b + 3 * su + 2 * st);
  p.n = c->plan_n;
  return p;
}

template <typename T>
int build_plan(sw_ctx* c) {
  const size_t n = 6 * (12 * static_cast<size_t>(c->nx) + 36);
  const size_t su = plan_stride(n, 4), st = plan_stride(n, sizeof(T));
  // the rot arrays are addressed k * n + q from p.rot: keep them n apart
  const size_t bytes = 3 * su + 2 * st + kPlanRot * n * sizeof(T) + 256;
  SW_HIP(hipMalloc(&c->plan, bytes));
  c->grid->plan = c->plan;
  c->plan_n = static_cast<int>(n);
  const HaloPlan<T> p = plan_of<T>(c);
  int* far = nullptr;
  SW_HIP(hipMalloc(&far, sizeof(int)));
  PlanBuildArgs<T> a{geo_of(c), route(), planes_of<T>(c), const_cast<uint32_t*>(p.dst),
                     const_cast<uint32_t*>(p.s0), const_cast<uint32_t*>(p.s1), const_cast<T*>(p.c0),
                     const_cast<T*>(p.c1), const_cast<T*>(p.rot), c->plan_n, far};
  int hfar = -1;
  hipError_t e = hipMemsetAsync(far, 0, sizeof(int), c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_plan_build<T>, dim3(blocks_for(n)), dim3(256), 0, c->stream, a);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&hfar, far, sizeof(int), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(far);
  if (e != hipSuccess) return fail(SW_E_HIP, std::string("halo plan build: ") + hipGetErrorString(e));
  c->frame_edge_ok = c->frame_edge = hfar == 0 ? 1 : 0;
  return SW_OK;
}

template <typename T>
int launch_halo_scalar(const sw_ctx* c, T* f, int ncomp) {
  const uint32_t NN = static_cast<uint32_t>(c->N) * static_cast<uint32_t>(c->N);
  HaloArgs<T> a{plan_of<T>(c), NN, f, ncomp, nullptr, nullptr};
  hipLaunchKernelGGL(k_halo_scalar<T>, dim3(blocks_for(c->plan_n)), dim3(256), 0, c->stream, a);
  SW_HIP(hipGetLastError());
  return SW_OK;
}

template <typename T>
int launch_halo_vector(const sw_ctx* c, T* u, T* v) {
  const uint32_t NN = static_cast<uint32_t>(c->N) * static_cast<uint32_t>(c->N);
  HaloArgs<T> a{plan_of<T>(c), NN, nullptr, 1, u, v};
  hipLaunchKernelGGL(k_halo_vector<T>, dim3(blocks_for(c->plan_n)), dim3(256), 0, c->stream, a);
  SW_HIP(hipGetLastError());
  return SW_OK;
}

// The 12 cube edges, each once, as t * 4 + d with (t, d) before its neighbour.
int edge_pairs(int ue[12]) {
  const Route& R = route();
  int k = 0;
  for (int t = 0; t < 6; ++t)
    for (int d = 0; d < 4; ++d)
      if (t * 4 + d < R.r[t][d][0] * 4 + R.r[t][d][1] && k < 12) ue[k++] = t * 4 + d;
  if (k != 12) return fail(SW_E_STATE, "routing table does not pair the 24 face edges");
  return SW_OK;
}

template <typename T>
int stage_args(const sw_ctx* c, const T* h, const T* hp, const T* u, const T* v, const T* up,
               const T* vp, T* h_out, T* u_out, T* v_out, double dt, double beta, int last, int st,
               int rk, StageArgs<T>& A, bool split = true) {
  const Geo g = geo_of(c);
  const Planes<T> P = planes_of<T>(c);
  T* evel = edge_buf<T>(c, 2 * st);
  T* eflux = edge_buf<T>(c, 2 * st + 1);
  A.e = EdgeArgs<T>{g, route(), P, h, u, v, evel, eflux, last, {}, 0, {}};
  int rc = edge_pairs(A.e.ue);
  if (rc) return rc;
  // the next stage's edge pass, on this stage's outputs, shares the launch of
  // their velocity halo fill
  const int nst = (st + 1) % rk;
  A.en = EdgeArgs<T>{g, route(), P, h_out, u_out, v_out, edge_buf<T>(c, 2 * nst), edge_buf<T>(c, 2 * nst + 1),
                     nst == rk - 1 ? 1 : 0, {}, 1, plan_of<T>(c)};
  if ((rc = edge_pairs(A.en.ue))) return rc;
  MarchArgs<T> ma{};
  ma.g = g;
  ma.P = P;
  ma.h = h; ma.hp = hp; ma.u = u; ma.v = v; ma.up = up; ma.vp = vp;
  ma.hn = h_out;
  ma.evel = evel; ma.eflux = eflux;
  ma.h_out = h_out; ma.u_out = u_out; ma.v_out = v_out;
  ma.dt = T(dt);
  ma.beta = T(beta);
  ma.last = last;
  ma.S = (c->nx + kStripOut - 1) / kStripOut;
  ma.shared = c->shared;
  ma.gsym = c->gsym;
  ma.sink = edge_buf<T>(c, kEdgeSink);
  A.ma = ma;
  A.ma.R = c->rows_a;
  A.ma.nbands = (c->nx + A.ma.R - 1) / A.ma.R;
  A.ma.nwaves = 6 * ma.S * A.ma.nbands;
  A.mb = ma;
  A.mb.R = c->rows_b;
  A.mb.nbands = (c->nx + A.mb.R - 1) / A.mb.R;
  A.mb.nwaves = 6 * ma.S * A.mb.nbands;
  A.mf = ma;
  A.mf.S = (c->nx + kFStride - 1) / kFStride;
  // without the split (batched runs) a split context's pair of bands counts
  // as two uniform ones
  A.mf.R = c->rows_f <= 0 ? 32 : (!split && c->rows_f2 > 0) ? (c->rows_f + c->rows_f2) / 2 : c->rows_f;
  A.mf.nbands = (c->nx + A.mf.R - 1) / A.mf.R;
  A.mf.nwaves = 6 * A.mf.S * A.mf.nbands;
  A.mf.R2 = A.mf.nlong = A.mf.nwlong = 0;
  if (split && c->rows_f > 0 && c->rows_f2 > 0 && c->rows_f < c->nx) {
    const int span = c->rows_f + c->rows_f2;
    const int pairs = (c->nx + span - 1) / span;                   // long bands per column
    const int nshort = (c->nx - c->rows_f + span - 1) / span;      // non-empty short bands per column
    A.mf.R2 = c->rows_f2;
    A.mf.nwlong = 6 * A.mf.S * pairs;
    A.mf.nlong = ((A.mf.nwlong + 3) / 4 + kXcds - 1) / kXcds * kXcds;
    A.mf.nwaves = 6 * A.mf.S * nshort;
  }
  const uint32_t NN = static_cast<uint32_t>(c->N) * static_cast<uint32_t>(c->N);
  A.hh = HaloArgs<T>{plan_of<T>(c), NN, h_out, 1, nullptr, nullptr};
  A.huv = HaloArgs<T>{plan_of<T>(c), NN, nullptr, 1, u_out, v_out};
  return SW_OK;
}

// Blocks of a fused-kernel launch: the long set then the short one when the
// bands are split, else one per four waves.
template <typename T>
int fused_blocks(const MarchArgs<T>& m) {
  return m.R2 > 0 ? m.nlong + (m.nwaves + 3) / 4 : (m.nwaves + 3) / 4;
}

// One RK stage (stage index st of an rk-stage substep) from (h, hp, u, v, up,
// vp) into (h_out, u_out, v_out), halos filled.  The stage's averaged edge
// values come from the previous stage's last launch, or from a launch of
// their own when `own_edge` (the first stage of a run).  The last launch
// also forms the next stage's edge values from this stage's outputs.
template <typename T>
int run_stage(const sw_ctx* c, const T* h, const T* hp, const T* u, const T* v, const T* up,
              const T* vp, T* h_out, T* u_out, T* v_out, double dt, double beta, int last, int st,
              int rk, bool own_edge) {
  StageArgs<T> A;
  int rc = stage_args<T>(c, h, hp, u, v, up, vp, h_out, u_out, v_out, dt, beta, last, st, rk, A);
  if (rc) return rc;
  const int ne = static_cast<int>(blocks_for(24 * kEdgeLanes * static_cast<size_t>(c->nx)));
  if (own_edge && (rc = launch(c, K_EDGE, k_edge<T>, dim3(ne), A.e))) return rc;
  const dim3 ga((A.ma.nwaves + 3) / 4), gb((A.mb.nwaves + 3) / 4), gh(blocks_for(c->plan_n));
  const int nh = static_cast<int>(blocks_for(c->plan_n));
  auto tail = [&]() { return launch(c, K_HALO_UV, k_halo_uv_edge<T>, dim3(nh + ne), A.huv, A.en, nh); };
  if (c->rows_f > 0) {
    // fused height + deep-cell velocity, new-height halo, frame velocity
    const bool sf = A.mf.hp == A.mf.h && A.mf.up == A.mf.u && A.mf.vp == A.mf.v;
    const int nft = 6 * frame_tiles(c->nx);
    const dim3 gf(fused_blocks(A.mf)), gr(nft + static_cast<int>(blocks_for(c->plan_n)));
    if (c->stamp_next) {  // launch-duration stamps (sw_profile_graph)
      const int waves = 4 * static_cast<int>(gf.x);
      if (c->stamp_next + 2 * waves > c->stamp_end) return fail(SW_E_STATE, "stamp buffer too small");
      A.mf.stamps = c->stamp_next;
      c->stamp_launches.push_back({static_cast<size_t>(c->stamp_next - c->stamp_base), waves});
      c->stamp_next += 2 * waves;
    }
    auto lf = [&](auto k) { return launch(c, K_FUSED, k, gf, A.mf); };
    if (c->gsym)
      rc = sf ? lf(k_fused<T, true, true>) : lf(k_fused<T, true, false>);
    else
      rc = sf ? lf(k_fused<T, false, true>) : lf(k_fused<T, false, false>);
    if (rc) return rc;
    if (c->frame_edge) {
      // frame cells, velocity and height halo fill and the next edge pass in
      // one launch
      const dim3 ge(12 * ((c->nx + kSeg - 1) / kSeg));
      constexpr int nt = EdgeGeo<kSeg>::threads;
      return c->gsym ? launch_n(c, K_FRAME, k_frame_edge<T, true>, ge, nt, A.mf, A.en)
                     : launch_n(c, K_FRAME, k_frame_edge<T, false>, ge, nt, A.mf, A.en);
    }
    // frame cells and the new height's halo fill in one launch
    rc = c->gsym ? launch(c, K_FRAME, k_frame<T, true>, gr, A.mf, A.hh, nft)
                 : launch(c, K_FRAME, k_frame<T, false>, gr, A.mf, A.hh, nft);
    if (rc) return rc;
    return tail();
  }
  const bool sa = A.ma.hp == A.ma.h, sb = A.mb.up == A.mb.u && A.mb.vp == A.mb.v;
  auto la = [&](auto k) { return launch(c, K_MARCH_A, k, ga, A.ma); };
  auto lb = [&](auto k) { return launch(c, K_MARCH_B, k, gb, A.mb); };
  if (c->gsym)
    rc = sa ? la(k_march_a<T, true, true>) : la(k_march_a<T, true, false>);
  else
    rc = sa ? la(k_march_a<T, false, true>) : la(k_march_a<T, false, false>);
  if (rc) return rc;
  if ((rc = launch(c, K_HALO_H, k_halo_scalar<T>, gh, A.hh))) return rc;
  if (c->gsym)
    rc = sb ? lb(k_march_b<T, true, true>) : lb(k_march_b<T, true, false>);
  else
    rc = sb ? lb(k_march_b<T, false, true>) : lb(k_march_b<T, false, false>);
  if (rc) return rc;
  return tail();
}

template <typename T>
int run_diag(const sw_ctx* c, const T* h, const T* u, const T* v, void* const out[13], int st) {
  DiagArgs<T> d{};
  d.g = geo_of(c);
  d.P = planes_of<T>(c);
  d.h = h; d.u = u; d.v = v;
  d.evel = edge_buf<T>(c, 2 * st);  // the averaged edge velocity of that stage
  T** dst[10] = {&d.cx, &d.cy, &d.uct, &d.vct, &d.ucx, &d.ucy, &d.vcx, &d.vcy, &d.advx, &d.advy};
  for (int k = 0; k < 10; ++k) *dst[k] = static_cast<T*>(out[3 + k]);
  int rc = launch(c, K_DIAG, k_diag1<T>, dim3(blocks_for(c->cells)), d);
  if (!rc) rc = launch(c, K_DIAG, k_diag2<T>, dim3(blocks_for(c->cells)), d);
  if (!rc) rc = launch(c, K_DIAG, k_diag3<T>, dim3(blocks_for(c->cells)), d);
  return rc;
}

int check_ptrs(const void* const* p, int n, const char* what) {
  if (!p) return fail(SW_E_ARG, std::string(what) + " is null");
  for (int k = 0; k < n; ++k)
    if (!p[k]) return fail(SW_E_ARG, std::string(what) + "[" + std::to_string(k) + "] is null");
  return SW_OK;
}

// out[k] may alias an input for k < first_checked (SW_INPLACE: h, u, v).
int check_disjoint(const sw_ctx* c, const void* const in[6], void* const out[13], int first_checked = 0) {
  for (int k = first_checked; k < 13; ++k)
    for (int m = 0; m < 6; ++m)
      if (out[k] && out[k] == in[m]) return fail(SW_E_ARG, "output " + std::to_string(k) + " aliases an input");
  for (int k = 0; k < 13; ++k)
    for (int m = k + 1; m < 13; ++m)
      if (out[k] && out[k] == out[m]) return fail(SW_E_ARG, "outputs alias each other");
  (void)c;
  return SW_OK;
}

// One substep (RK2: 2 stages, RK3: 3 stages, barotr.py:47-59 / 70-85) from
// (h, hp, u, v, up, vp) into dst.  Stage scratch S1 / S2 holds the inputs of
// the last stage afterwards (the diagnostics are computed from them).
template <typename T>
int run_substep(const sw_ctx* c, int rk, double dtb, const T* h, const T* hp, const T* u,
                const T* v, const T* up, const T* vp, T* const dst[3], bool first = false) {
  T* s1[3] = {scratch_field<T>(c, 0), scratch_field<T>(c, 1), scratch_field<T>(c, 2)};
  T* s2[3] = {scratch_field<T>(c, 3), scratch_field<T>(c, 4), scratch_field<T>(c, 5)};
  int rc;
  if (rk == 2) {
    if ((rc = run_stage<T>(c, h, hp, u, v, up, vp, s1[0], s1[1], s1[2], dtb / 2.0, 0.0, 0, 0, 2, first)))
      return rc;
    return run_stage<T>(c, s1[0], hp, s1[1], s1[2], up, vp, dst[0], dst[1], dst[2], dtb, 0.0, 1, 1, 2, false);
  }
  if ((rc = run_stage<T>(c, h, hp, u, v, up, vp, s1[0], s1[1], s1[2], dtb / 3.0, 1.0, 0, 0, 3, first)))
    return rc;
  if ((rc = run_stage<T>(c, s1[0], hp, s1[1], s1[2], up, vp, s2[0], s2[1], s2[2], dtb / 2.0, 1.0, 0, 1, 3, false)))
    return rc;
  return run_stage<T>(c, s2[0], hp, s2[1], s2[2], up, vp, dst[0], dst[1], dst[2], dtb, 1.0, 1, 2, 3, false);
}

constexpr int kGraphPairs = 16;  // substep pairs recorded in one graph

// Graph executing `pairs` times the two substeps X -> Y -> X on the internal
// ping-pong buffers; recorded once per (rk, dtb, pairs) on the private stream.
template <typename T>
int pair_graph(sw_ctx* c, int rk, double dtb, int pairs, hipGraphExec_t* exec) {
  for (const auto& e : c->graphs)
    if (e.rk == rk && e.pairs == pairs && e.dtb == dtb) {
      *exec = e.exec;
      return SW_OK;
    }
  if (!c->capture) SW_HIP(hipStreamCreateWithFlags(&c->capture, hipStreamNonBlocking));
  T* X[3] = {scratch_field<T>(c, 6), scratch_field<T>(c, 7), scratch_field<T>(c, 8)};
  T* Y[3] = {scratch_field<T>(c, 9), scratch_field<T>(c, 10), scratch_field<T>(c, 11)};
  hipStream_t user = c->stream;
  c->stream = c->capture;
  SW_HIP(hipStreamBeginCapture(c->capture, hipStreamCaptureModeThreadLocal));
  int rc = SW_OK;
  for (int p = 0; p < pairs && !rc; ++p) {
    rc = run_substep<T>(c, rk, dtb, X[0], X[0], X[1], X[2], X[1], X[2], Y);
    if (!rc) rc = run_substep<T>(c, rk, dtb, Y[0], Y[0], Y[1], Y[2], Y[1], Y[2], X);
  }
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(c->capture, &graph);
  c->stream = user;
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return fail(SW_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(SW_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  if (c->graphs.size() >= 8) {
    (void)hipGraphExecDestroy(c->graphs.front().exec);
    c->graphs.erase(c->graphs.begin());
  }
  c->graphs.push_back({rk, pairs, dtb, x});
  *exec = x;
  return SW_OK;
}

// RK driver (barotr.py:41-87).  Eager: every substep is launched from the
// host.  Graph mode (SW_GRAPH): the first substep reads the caller's fields
// into the internal buffer X; the remaining ones replay recorded X -> Y -> X
// pairs; the final h, u, v are copied to out.
template <typename T>
int run_rk(sw_ctx* c, const void* const in[6], void* const out[13], int nbb, double dtb, int flags,
           int rk) {
  int rc = check_ptrs(in, 6, "in");
  if (rc) return rc;
  // the 10 diagnostic outputs may be null when they are not written
  rc = check_ptrs(const_cast<const void* const*>(out), (flags & SW_SKIP_DIAG) ? 3 : 13, "out");
  if (rc) return rc;
  const bool inplace = (flags & SW_INPLACE) != 0;
  rc = check_disjoint(c, in, out, inplace ? 3 : 0);
  if (rc) return rc;
  if (nbb < 0) return fail(SW_E_ARG, "nbb < 0");
  if (!(dtb == dtb) || dtb > 1e300 || dtb < -1e300) return fail(SW_E_ARG, "dtb must be finite");
  const size_t bytes = c->cells * sizeof(T);
  if (nbb == 0) {
    // the reference's empty loop returns the input's diagnostics, which the
    // six input fields do not carry: refuse rather than leave them unwritten
    if (!(flags & SW_SKIP_DIAG))
      return fail(SW_E_ARG, "nbb == 0 computes no diagnostics: pass SW_SKIP_DIAG (the caller keeps its own)");
    const int from[3] = {0, 2, 3};
    for (int k = 0; k < 3; ++k)
      if (out[k] != in[from[k]])
        SW_HIP(hipMemcpyAsync(out[k], in[from[k]], bytes, hipMemcpyDeviceToDevice, c->stream));
    return SW_OK;
  }
  const T* h = static_cast<const T*>(in[0]);
  const T* hp = static_cast<const T*>(in[1]);
  const T* u = static_cast<const T*>(in[2]);
  const T* v = static_cast<const T*>(in[3]);
  const T* up = static_cast<const T*>(in[4]);
  const T* vp = static_cast<const T*>(in[5]);
  T* fin[3] = {static_cast<T*>(out[0]), static_cast<T*>(out[1]), static_cast<T*>(out[2])};
  T* X[3] = {scratch_field<T>(c, 6), scratch_field<T>(c, 7), scratch_field<T>(c, 8)};
  T* Y[3] = {scratch_field<T>(c, 9), scratch_field<T>(c, 10), scratch_field<T>(c, 11)};
  const int last_stage_scratch = rk == 2 ? 0 : 3;
  const T* dh = scratch_field<T>(c, last_stage_scratch);
  const T* du = scratch_field<T>(c, last_stage_scratch + 1);
  const T* dv = scratch_field<T>(c, last_stage_scratch + 2);

  if (inplace && !((flags & SW_GRAPH) && nbb >= 3 && !c->prof)) {
    // outputs may alias inputs: the inputs are read by the first substep only,
    // so later substeps may write the outputs (the last one does)
    if ((rc = run_substep<T>(c, rk, dtb, h, hp, u, v, up, vp, X, true))) return rc;
    T** cur = X;
    for (int k = 1; k < nbb; ++k) {
      T** nxt = (k == nbb - 1) ? fin : (cur == X) ? Y : X;
      if ((rc = run_substep<T>(c, rk, dtb, cur[0], cur[0], cur[1], cur[2], cur[1], cur[2], nxt))) return rc;
      cur = nxt;
    }
    if (cur != fin)
      for (int k = 0; k < 3; ++k)
        SW_HIP(hipMemcpyAsync(fin[k], cur[k], bytes, hipMemcpyDeviceToDevice, c->stream));
  } else if ((flags & SW_GRAPH) && nbb >= 3 && !c->prof) {
    if ((rc = run_substep<T>(c, rk, dtb, h, hp, u, v, up, vp, X, true))) return rc;
    const int pairs = (nbb - 1) / 2;
    int done = 0;
    while (done < pairs) {
      const int p = (pairs - done >= kGraphPairs) ? kGraphPairs : pairs - done;
      hipGraphExec_t ex = nullptr;
      if ((rc = pair_graph<T>(c, rk, dtb, p, &ex))) return rc;
      SW_HIP(hipGraphLaunch(ex, c->stream));
      done += p;
    }
    T** cur = X;
    if ((nbb - 1) % 2) {
      if ((rc = run_substep<T>(c, rk, dtb, X[0], X[0], X[1], X[2], X[1], X[2], Y))) return rc;
      cur = Y;
    }
    for (int k = 0; k < 3; ++k)
      SW_HIP(hipMemcpyAsync(fin[k], cur[k], bytes, hipMemcpyDeviceToDevice, c->stream));
  } else {
    for (int k = 0; k < nbb; ++k) {
      T** dst = ((nbb - 1 - k) % 2 == 0) ? fin : X;
      if ((rc = run_substep<T>(c, rk, dtb, h, hp, u, v, up, vp, dst, k == 0))) return rc;
      h = hp = dst[0];
      u = up = dst[1];
      v = vp = dst[2];
    }
  }
  if (!(flags & SW_SKIP_DIAG)) return run_diag<T>(c, dh, du, dv, out, rk - 1);
  return SW_OK;
}

// The substep-pair graph of run_rk's SW_GRAPH path, recorded with the fused
// kernel's launch-duration stamps on (each wave's lane 0 writes the device
// wall clock at entry and exit), replayed `replays` times.  After each replay
// the host reads the stamps: a launch lasts from its first wave's entry to
// its last wave's exit.  ms[K_FUSED] / counts[K_FUSED] sum these over every
// replay; *graph_ms sums the replays themselves (events around each graph
// launch).  Each measured replay follows kLeadReplays unmeasured ones queued
// back to back.  The first substep, from the caller's fields into X, runs
// eagerly and untimed; out[0..2] receive the state after
// 1 + 2 pairs (1 + kLeadReplays) replays substeps.
constexpr int kLeadReplays = 3;  // unmeasured replays ahead of each measured one

template <typename T>
int run_profile_graph(sw_ctx* c, const void* const in[6], void* const out[13], int pairs, int replays,
                      double dtb, double* ms, int* counts, double* graph_ms) {
  int rc = check_ptrs(in, 6, "in");
  if (!rc) rc = check_ptrs(const_cast<const void* const*>(out), 3, "out");
  if (!rc) rc = check_disjoint(c, in, out);
  if (rc) return rc;
  if (c->rows_f <= 0) return fail(SW_E_STATE, "graph timing covers the fused stage kernel (rows_f > 0)");
  for (int k = 0; k < K_CLASSES; ++k) {
    ms[k] = 0.0;
    counts[k] = 0;
  }
  *graph_ms = 0.0;
  int rate_khz = 0;
  SW_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, c->device));
  if (rate_khz <= 0) return fail(SW_E_HIP, "device reports no wall-clock rate");
  const T* p[6];
  for (int k = 0; k < 6; ++k) p[k] = static_cast<const T*>(in[k]);
  T* X[3] = {scratch_field<T>(c, 6), scratch_field<T>(c, 7), scratch_field<T>(c, 8)};
  T* Y[3] = {scratch_field<T>(c, 9), scratch_field<T>(c, 10), scratch_field<T>(c, 11)};
  if ((rc = run_substep<T>(c, 2, dtb, p[0], p[1], p[2], p[3], p[4], p[5], X, true))) return rc;
  // stamp slots: 2 per wave of each of the graph's 4 pairs fused launches
  StageArgs<T> shape;  // launch shape of the graph's fused launches
  if ((rc = stage_args<T>(c, X[0], X[0], X[1], X[2], X[1], X[2], Y[0], Y[1], Y[2], dtb, 0.0, 0, 0, 2, shape)))
    return rc;
  const size_t waves = 4 * static_cast<size_t>(fused_blocks(shape.mf));
  const size_t nslots = 2 * waves * 4 * static_cast<size_t>(pairs);
  unsigned long long* dstamps = nullptr;
  SW_HIP(hipMalloc(&dstamps, nslots * sizeof(unsigned long long)));
  std::vector<unsigned long long> hstamps(nslots);
  if (!c->capture) rc = hipStreamCreateWithFlags(&c->capture, hipStreamNonBlocking) == hipSuccess
                            ? SW_OK : fail(SW_E_HIP, "hipStreamCreateWithFlags failed");
  hipStream_t user = c->stream;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ea = nullptr, eb = nullptr;
  if (!rc && hipStreamSynchronize(user) != hipSuccess) rc = fail(SW_E_HIP, "hipStreamSynchronize failed");
  if (!rc) {
    c->stream = c->capture;
    c->stamp_base = c->stamp_next = dstamps;
    c->stamp_end = dstamps + nslots;
    c->stamp_launches.clear();
    hipError_t e = hipStreamBeginCapture(c->capture, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) rc = fail(SW_E_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
    for (int q = 0; q < pairs && !rc; ++q) {
      rc = run_substep<T>(c, 2, dtb, X[0], X[0], X[1], X[2], X[1], X[2], Y);
      if (!rc) rc = run_substep<T>(c, 2, dtb, Y[0], Y[0], Y[1], Y[2], Y[1], Y[2], X);
    }
    if (e == hipSuccess) {
      e = hipStreamEndCapture(c->capture, &graph);
      if (!rc && e != hipSuccess) rc = fail(SW_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    c->stream = user;
    c->stamp_base = c->stamp_next = c->stamp_end = nullptr;
  }
  if (!rc && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
    rc = fail(SW_E_HIP, "hipGraphInstantiate failed");
  if (!rc && (hipEventCreate(&ea) != hipSuccess || hipEventCreate(&eb) != hipSuccess))
    rc = fail(SW_E_HIP, "hipEventCreate failed");
  for (int r = 0; r < replays && !rc; ++r) {
    // kLeadReplays replays back to back before the measured one, so the
    // device runs it at the clocks of a sustained sequence of replays (the
    // timed steps of a benchmark), not those of a device that just idled
    hipError_t e = hipSuccess;
    for (int q = 0; q < kLeadReplays && e == hipSuccess; ++q) e = hipGraphLaunch(exec, c->stream);
    if (e == hipSuccess) e = hipEventRecord(ea, c->stream);
    if (e == hipSuccess) e = hipGraphLaunch(exec, c->stream);
    if (e == hipSuccess) e = hipEventRecord(eb, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, ea, eb);
    if (e == hipSuccess)
      e = hipMemcpy(hstamps.data(), dstamps, nslots * sizeof(unsigned long long), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      rc = fail(SW_E_HIP, std::string("timed graph replay: ") + hipGetErrorString(e));
      break;
    }
    *graph_ms += t;
    for (const auto& L : c->stamp_launches) {
      unsigned long long lo = ~0ull, hi = 0;
      for (int w = 0; w < L.second; ++w) {
        lo = std::min(lo, hstamps[L.first + 2 * w]);
        hi = std::max(hi, hstamps[L.first + 2 * w + 1]);
      }
      ms[K_FUSED] += static_cast<double>(hi - lo) / rate_khz;
      counts[K_FUSED] += 1;
    }
  }
  c->stamp_launches.clear();
  if (ea) (void)hipEventDestroy(ea);
  if (eb) (void)hipEventDestroy(eb);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipFree(dstamps);
  if (rc) return rc;
  for (int k = 0; k < 3; ++k)
    SW_HIP(hipMemcpyAsync(out[k], X[k], c->cells * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
  SW_HIP(hipStreamSynchronize(c->stream));
  return SW_OK;
}

// Recorded graphs bake in the launch shape; drop them when it changes.
int drop_graphs(sw_ctx* c) {
  if (c->graphs.empty()) return SW_OK;
  SW_HIP(hipStreamSynchronize(c->stream));
  for (auto& e : c->graphs) (void)hipGraphExecDestroy(e.exec);
  c->graphs.clear();
  return SW_OK;
}

// Per-context workspace: stage scratch, substep ping-pong fields, edge lines.
template <typename T>
int alloc_workspace(sw_ctx* c) {
  const size_t n = c->cells;
  SW_HIP(hipMalloc(&c->scratch, 12 * n * sizeof(T)));
  SW_HIP(hipMalloc(&c->edges, kEdgeBufs * edge_stride(c) * sizeof(T)));
  SW_HIP(hipMemsetAsync(c->scratch, 0, 12 * n * sizeof(T), c->stream));
  SW_HIP(hipMemsetAsync(c->edges, 0, kEdgeBufs * edge_stride(c) * sizeof(T), c->stream));
  return SW_OK;
}

template <typename T>
int do_create_planes(sw_ctx* c, const void* const consts[SW_N_CONSTS]) {
  const size_t n = c->cells;
  c->grid = std::make_shared<GridMem>();
  SW_HIP(hipMalloc(&c->planes, ((kNumPlanes + 2 * kPairPlanes) * n + strip_elems(c->N)) * sizeof(T)));
  c->grid->planes = c->planes;
  SW_HIP(hipMalloc(&c->locs, 2 * n * sizeof(int)));
  c->grid->locs = c->locs;
  int rc = alloc_workspace<T>(c);
  if (rc) return rc;
  T* p = static_cast<T*>(c->planes);
  RepackArgs<T> a{};
  a.n = n;
  a.dzx = static_cast<const double*>(consts[0]);
  a.dzy = static_cast<const double*>(consts[1]);
  a.rdx = static_cast<const double*>(consts[2]);
  a.rdy = static_cast<const double*>(consts[3]);
  a.af = static_cast<const double*>(consts[4]);
  a.coef = static_cast<const double*>(consts[5]);
  a.li = static_cast<const int*>(consts[6]);
  a.lj = static_cast<const int*>(consts[7]);
  a.c2l = static_cast<const double*>(consts[8]);
  a.l2c = static_cast<const double*>(consts[9]);
  a.gct = static_cast<const double*>(consts[12]);
  a.sina = static_cast<const double*>(consts[13]);
  a.rda = static_cast<const double*>(consts[14]);
  a.cdy = static_cast<const double*>(consts[15]);
  a.ddx = static_cast<const double*>(consts[16]);
  T** outs[kNumPlanes] = {&a.o_dzx, &a.o_dzy, &a.o_rdx, &a.o_rdy, &a.o_af, &a.o_rda, &a.o_cdy, &a.o_ddx,
                          &a.o_sina, &a.o_g00, &a.o_g01, &a.o_g10, &a.o_g11, &a.o_c00, &a.o_c01, &a.o_c10,
                          &a.o_c11, &a.o_l00, &a.o_l01, &a.o_l10, &a.o_l11, &a.o_k0, &a.o_k1};
  for (int k = 0; k < kNumPlanes; ++k) *outs[k] = p + k * n;
  a.o_gp = p + kNumPlanes * n;
  a.o_rxy = a.o_gp + 2 * n;
  a.o_rdd = a.o_rxy + 2 * n;
  a.o_li = c->locs;
  a.o_lj = c->locs + n;
  hipLaunchKernelGGL(k_repack<T>, dim3(blocks_for(n)), dim3(256), 0, c->stream, a);
  SW_HIP(hipGetLastError());
  const Planes<T> P = planes_of<T>(c);
  hipLaunchKernelGGL(k_pack_strips<T>, dim3(blocks_for(strip_elems(c->N))), dim3(256), 0, c->stream, P, c->N,
                     const_cast<T*>(P.strip));
  SW_HIP(hipGetLastError());
  return build_plan<T>(c);
}

// Host-side validation of the tables the kernels index with.
int validate_consts(sw_ctx* c, const void* const consts[SW_N_CONSTS]) {
  const size_t n = c->cells;
  const int N = c->N;
  std::vector<int> li(n), lj(n);
  std::vector<double> inner(n), outer(n), dzx(n), dzy(n);
  SW_HIP(hipMemcpy(li.data(), consts[6], n * sizeof(int), hipMemcpyDeviceToHost));
  SW_HIP(hipMemcpy(lj.data(), consts[7], n * sizeof(int), hipMemcpyDeviceToHost));
  SW_HIP(hipMemcpy(inner.data(), consts[10], n * sizeof(double), hipMemcpyDeviceToHost));
  SW_HIP(hipMemcpy(outer.data(), consts[11], n * sizeof(double), hipMemcpyDeviceToHost));
  SW_HIP(hipMemcpy(dzx.data(), consts[0], n * sizeof(double), hipMemcpyDeviceToHost));
  SW_HIP(hipMemcpy(dzy.data(), consts[1], n * sizeof(double), hipMemcpyDeviceToHost));
  for (int t = 0; t < 6; ++t) {
    for (int i = 0; i < N; ++i) {
      for (int j = 0; j < N; ++j) {
        const size_t o = (static_cast<size_t>(t) * N + i) * N + j;
        const bool in = i >= 3 && i < N - 3 && j >= 3 && j < N - 3;
        if (inner[o] != (in ? 1.0 : 0.0) || outer[o] != (in ? 0.0 : 1.0))
          return fail(SW_E_CONST, "inner/outer are not the standard interior/halo masks");
        const bool ew = (j < 3 || j >= N - 3) && i >= 3 && i < N - 3;
        const bool ns = (i < 3 || i >= N - 3) && j >= 3 && j < N - 3;
        if (ew && (li[o] < 3 || li[o] > N - 5))
          return fail(SW_E_CONST, "loc_i_local out of range on a W/E halo strip");
        if (ns && (lj[o] < 3 || lj[o] > N - 5))
          return fail(SW_E_CONST, "loc_j_local out of range on a S/N halo strip");
      }
    }
  }
  bool uni = true;
  for (size_t o = 0; o < n && uni; ++o) uni = (dzx[o] == dzx[0]) && (dzy[o] == dzy[0]);
  // Planes that are the same on every face can be read from face 0, and a
  // symmetric metric with equal diagonal needs two of its four entries.
  const size_t NN = static_cast<size_t>(N) * N;
  std::vector<double> buf(n), gct(4 * n);
  bool shared = true;
  for (int k : {2, 3, 14, 15, 16}) {
    SW_HIP(hipMemcpy(buf.data(), consts[k], n * sizeof(double), hipMemcpyDeviceToHost));
    for (int t = 1; t < 6 && shared; ++t)
      shared = std::memcmp(buf.data(), buf.data() + t * NN, NN * sizeof(double)) == 0;
  }
  SW_HIP(hipMemcpy(gct.data(), consts[12], 4 * n * sizeof(double), hipMemcpyDeviceToHost));
  bool gsym = true;
  for (size_t o = 0; o < n && gsym; ++o) {
    uint64_t b[4];
    std::memcpy(b, &gct[4 * o], sizeof(b));
    gsym = (b[1] == b[2]) && (b[0] == b[3]);
  }
  c->shared = shared ? 1 : 0;
  c->gsym = gsym ? 1 : 0;
  c->dz_uniform = uni ? 1 : 0;
  c->dzx_s = dzx[0];
  c->dzy_s = dzy[0];
  return SW_OK;
}

void free_ctx(sw_ctx* c) {
  if (!c) return;
  for (auto& e : c->graphs) (void)hipGraphExecDestroy(e.exec);
  c->graphs.clear();
  if (c->capture) (void)hipStreamDestroy(c->capture);
  if (c->scratch) (void)hipFree(c->scratch);
  if (c->edges) (void)hipFree(c->edges);
  if (c->ibuf) (void)hipFree(c->ibuf);
  c->grid.reset();
  delete c;
}

constexpr int kIntegralBlocks = 1024;

template <typename T>
int run_integrals(sw_ctx* c, const void* h, const void* u, const void* v, double out[2]) {
  if (!c->ibuf) SW_HIP(hipMalloc(&c->ibuf, (4 * kIntegralBlocks + 2) * sizeof(double)));
  IntegralArgs<T> a{geo_of(c), planes_of<T>(c), static_cast<const T*>(h), static_cast<const T*>(u),
                    static_cast<const T*>(v), c->ibuf};
  hipLaunchKernelGGL(k_integrals<T>, dim3(kIntegralBlocks), dim3(256), 0, c->stream, a);
  SW_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_integrals_final, dim3(1), dim3(256), 0, c->stream, c->ibuf, kIntegralBlocks,
                     c->ibuf + 4 * kIntegralBlocks);
  SW_HIP(hipGetLastError());
  SW_HIP(hipMemcpyAsync(out, c->ibuf + 4 * kIntegralBlocks, 2 * sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  SW_HIP(hipStreamSynchronize(c->stream));
  return SW_OK;
}

// Workspace of sw_integrals_noalloc: the per-block compensated partials of
// k_integrals (4 doubles per block, kIntegralBlocks blocks).  Each GPU holds
// its own copy until the process ends.  In host code the name `scratch` is
// only a stand-in for the symbol: kernels get the device address that
// hipGetSymbolAddress returns for the current device (scratch_address).
// Calls on one GPU share the array, so calls that overlap on one GPU (two
// streams, two threads) race; every slot a call reads is written by that
// call's first kernel, so nothing needs clearing between calls.
constexpr size_t kScratchDoubles = 4 * static_cast<size_t>(kIntegralBlocks);
static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_scratch_mu;
std::vector<double*> g_scratch_addr;  // per device ordinal

// Device address of `scratch` on device `dev`, which must be current; looked
// up once per device and kept.
int scratch_address(int dev, double** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (dev < 0) return fail(SW_E_ARG, "bad device ordinal");
  if (static_cast<size_t>(dev) >= g_scratch_addr.size()) g_scratch_addr.resize(dev + 1, nullptr);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(SW_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_addr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_addr[dev];
  return SW_OK;
}

// The integrals with the block partials in `scratch` and the result written
// to device memory: no allocation, no host synchronisation.
template <typename T>
int run_integrals_noalloc(sw_ctx* c, const void* h, const void* u, const void* v, double* out) {
  if (!c->noalloc) return fail(SW_E_STATE, "context has no scratch address (sw_create did not resolve it)");
  const size_t need = 4 * static_cast<size_t>(kIntegralBlocks);
  if (need > kScratchDoubles)
    return fail(SW_E_ARG, "block partials (" + std::to_string(need) + " doubles) exceed the scratch array (" +
                              std::to_string(kScratchDoubles) + ")");
  IntegralArgs<T> a{geo_of(c), planes_of<T>(c), static_cast<const T*>(h), static_cast<const T*>(u),
                    static_cast<const T*>(v), c->noalloc};
  hipLaunchKernelGGL(k_integrals<T>, dim3(kIntegralBlocks), dim3(256), 0, c->stream, a);
  SW_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_integrals_final, dim3(1), dim3(256), 0, c->stream, c->noalloc, kIntegralBlocks, out);
  SW_HIP(hipGetLastError());
  return SW_OK;
}

// ---------------------------------------------------------------------------
// batched runs
// ---------------------------------------------------------------------------
constexpr int kBatchMax = 4096;
constexpr size_t kBatchGraphs = 16;  // recorded substep-pair graphs kept per batch object

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// nbb substeps for every member, all members advanced by the same launches.
// Members are ordered by decreasing nbb, so the members still stepping at
// substep k are a prefix of the tables and the grid of a launch covers
// exactly their blocks.
template <typename T>
int run_batch(sw_batch* b, const sw_member* mem, int count, int rk, int flags, hipStream_t stream) {
  // ---- validation
  std::vector<const void*> outs;
  std::vector<const void*> ins;
  bool gsym = true;
  bool fused = true;  // the fused stage kernel when every member's context uses it
  bool fedge = true;  // and the cube-edge frame launch when every member's uses that
  for (int i = 0; i < count; ++i) {
    const sw_member& m = mem[i];
    if (!m.ctx) return fail(SW_E_ARG, "member " + std::to_string(i) + ": ctx is null");
    if (m.ctx->dtype != (std::is_same<T, double>::value ? SW_F64 : SW_F32))
      return fail(SW_E_ARG, "members must share one dtype");
    int rc = check_ptrs(m.in, 6, "in");
    if (!rc) rc = check_ptrs(const_cast<const void* const*>(m.out), (flags & SW_SKIP_DIAG) ? 3 : 13, "out");
    if (!rc) rc = check_disjoint(m.ctx, m.in, m.out);
    if (rc) return fail(rc, "member " + std::to_string(i) + ": " + sw_last_error());
    if (m.nbb < 0) return fail(SW_E_ARG, "member " + std::to_string(i) + ": nbb < 0");
    if (m.nbb == 0 && !(flags & SW_SKIP_DIAG))
      return fail(SW_E_ARG, "member " + std::to_string(i) +
                                ": nbb == 0 computes no diagnostics: pass SW_SKIP_DIAG (the caller keeps its own)");
    if (!(m.dtb == m.dtb) || m.dtb > 1e300 || m.dtb < -1e300)
      return fail(SW_E_ARG, "member " + std::to_string(i) + ": dtb must be finite");
    for (int k = 0; k < 13; ++k)
      if (m.out[k]) outs.push_back(m.out[k]);
    for (int k = 0; k < 6; ++k) ins.push_back(m.in[k]);
    gsym = gsym && m.ctx->gsym;
    fused = fused && m.ctx->rows_f > 0;
    fedge = fedge && m.ctx->frame_edge;
  }
  {
    std::vector<const sw_ctx*> ctxs;
    for (int i = 0; i < count; ++i) ctxs.push_back(mem[i].ctx);
    std::sort(ctxs.begin(), ctxs.end());
    if (std::adjacent_find(ctxs.begin(), ctxs.end()) != ctxs.end())
      return fail(SW_E_ARG, "each member needs its own context (sw_create_like shares a grid)");
    std::sort(outs.begin(), outs.end());
    if (std::adjacent_find(outs.begin(), outs.end()) != outs.end())
      return fail(SW_E_ARG, "output buffers of the members alias each other");
    for (const void* p : ins)
      if (std::binary_search(outs.begin(), outs.end(), p))
        return fail(SW_E_ARG, "an output buffer aliases an input of another member");
  }

  std::vector<int> order(count);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return mem[x].nbb > mem[y].nbb; });
  int K = 0;
  while (K < count && mem[order[K]].nbb > 0) ++K;
  bool in_alias = true;  // every stepping member's input has h0p = h0, ubp = ub, vbp = vb
  for (int q = 0; q < K; ++q) {
    const sw_member& m = mem[order[q]];
    in_alias = in_alias && m.in[1] == m.in[0] && m.in[4] == m.in[2] && m.in[5] == m.in[3];
  }
  for (int i = K; i < count; ++i) {  // nbb = 0: the reference's empty loop returns the input
    const sw_member& m = mem[order[i]];
    const size_t bytes = m.ctx->cells * sizeof(T);
    SW_HIP(hipMemcpyAsync(m.out[0], m.in[0], bytes, hipMemcpyDeviceToDevice, stream));
    SW_HIP(hipMemcpyAsync(m.out[1], m.in[2], bytes, hipMemcpyDeviceToDevice, stream));
    SW_HIP(hipMemcpyAsync(m.out[2], m.in[3], bytes, hipMemcpyDeviceToDevice, stream));
  }
  if (K == 0) return SW_OK;

  // ---- tables: StageArgs [stage][variant][member], nbb, first / blk2m per class
  const size_t ntab = static_cast<size_t>(rk) * 4 * K;
  std::vector<StageArgs<T>> tab(ntab);
  std::vector<int> nbb(K);
  std::vector<int> first[BK_CLASSES];
  for (auto& f : first) f.assign(K + 1, 0);
  for (int q = 0; q < K; ++q) {
    const sw_member& m = mem[order[q]];
    const sw_ctx* c = m.ctx;
    nbb[q] = m.nbb;
    const T* in[6];
    for (int k = 0; k < 6; ++k) in[k] = static_cast<const T*>(m.in[k]);
    T* fin[3] = {static_cast<T*>(m.out[0]), static_cast<T*>(m.out[1]), static_cast<T*>(m.out[2])};
    T* X[3] = {scratch_field<T>(c, 6), scratch_field<T>(c, 7), scratch_field<T>(c, 8)};
    T* S1[3] = {scratch_field<T>(c, 0), scratch_field<T>(c, 1), scratch_field<T>(c, 2)};
    T* S2[3] = {scratch_field<T>(c, 3), scratch_field<T>(c, 4), scratch_field<T>(c, 5)};
    for (int e = 0; e < 4; ++e) {
      // source (h, hp, u, v, up, vp) and destination of the substep
      const T* src[6];
      if (e < 2) {
        for (int k = 0; k < 6; ++k) src[k] = in[k];
      } else {
        T* const* a = (e == 2) ? X : fin;
        const T* f6[6] = {a[0], a[0], a[1], a[2], a[1], a[2]};
        for (int k = 0; k < 6; ++k) src[k] = f6[k];
      }
      T* const* dst = (e == 0 || e == 2) ? fin : X;
      const double dtb = m.dtb;
      int rc;
      StageArgs<T>* row = &tab[0];
      auto at = [&](int st) -> StageArgs<T>& { return row[(static_cast<size_t>(st) * 4 + e) * K + q]; };
      if (rk == 2) {
        rc = stage_args<T>(c, src[0], src[1], src[2], src[3], src[4], src[5], S1[0], S1[1], S1[2],
                           dtb / 2.0, 0.0, 0, 0, 2, at(0), false);
        if (!rc) rc = stage_args<T>(c, S1[0], src[1], S1[1], S1[2], src[4], src[5], dst[0], dst[1],
                                    dst[2], dtb, 0.0, 1, 1, 2, at(1), false);
      } else {
        rc = stage_args<T>(c, src[0], src[1], src[2], src[3], src[4], src[5], S1[0], S1[1], S1[2],
                           dtb / 3.0, 1.0, 0, 0, 3, at(0), false);
        if (!rc) rc = stage_args<T>(c, S1[0], src[1], S1[1], S1[2], src[4], src[5], S2[0], S2[1],
                                    S2[2], dtb / 2.0, 1.0, 0, 1, 3, at(1), false);
        if (!rc) rc = stage_args<T>(c, S2[0], src[1], S2[1], S2[2], src[4], src[5], dst[0], dst[1],
                                    dst[2], dtb, 1.0, 1, 2, 3, at(2), false);
      }
      if (rc) return rc;
    }
    const StageArgs<T>& A = tab[q];
    const int blocks[BK_CLASSES] = {static_cast<int>(blocks_for(24 * kEdgeLanes * static_cast<size_t>(c->nx))),
                                    (A.ma.nwaves + 3) / 4, static_cast<int>(blocks_for(c->plan_n)),
                                    (A.mb.nwaves + 3) / 4, (A.mf.nwaves + 3) / 4,
                                    6 * frame_tiles(c->nx) + static_cast<int>(blocks_for(c->plan_n)),
                                    static_cast<int>(blocks_for(c->plan_n) +
                                                     blocks_for(24 * kEdgeLanes * static_cast<size_t>(c->nx))),
                                    12 * ((c->nx + kBatchSeg - 1) / kBatchSeg)};
    for (int k = 0; k < BK_CLASSES; ++k) first[k][q + 1] = first[k][q] + blocks[k];
  }
  // byte layout of the device copy
  size_t off = 0;
  const size_t o_tab = off;
  off = align256(off + ntab * sizeof(StageArgs<T>));
  const size_t o_nbb = off;
  off = align256(off + K * sizeof(int));
  size_t o_first[BK_CLASSES], o_blk[BK_CLASSES];
  for (int k = 0; k < BK_CLASSES; ++k) {
    o_first[k] = off;
    off = align256(off + (K + 1) * sizeof(int));
    o_blk[k] = off;
    off = align256(off + static_cast<size_t>(first[k][K]) * sizeof(int));
  }
  const size_t bytes = off;

  b->reap(false);
  // ---- stage into pinned memory and copy (the copy that last read this
  // staging buffer, two runs ago, must be done)
  const int sl = b->slot;
  b->slot ^= 1;
  if (b->staged[sl]) SW_HIP(hipEventSynchronize(b->staged[sl]));
  if (b->host_cap[sl] < bytes) {
    if (b->host[sl]) SW_HIP(hipHostFree(b->host[sl]));
    b->host[sl] = nullptr;
    SW_HIP(hipHostMalloc(&b->host[sl], bytes, hipHostMallocDefault));
    b->host_cap[sl] = bytes;
  }
  if (b->dev_cap < bytes) {
    if (b->used && b->stream) SW_HIP(hipStreamSynchronize(b->stream));  // kernels may still read it
    if (b->used && b->stream != stream) SW_HIP(hipStreamSynchronize(stream));
    if (b->dev) SW_HIP(hipFree(b->dev));
    b->dev = nullptr;
    SW_HIP(hipMalloc(&b->dev, bytes));
    b->dev_cap = bytes;
  } else if (b->used && b->stream && b->stream != stream) {
    // the previous run's launches on another stream still read the tables
    SW_HIP(hipStreamSynchronize(b->stream));
  }
  char* h = static_cast<char*>(b->host[sl]);
  std::memcpy(h + o_tab, tab.data(), ntab * sizeof(StageArgs<T>));
  std::memcpy(h + o_nbb, nbb.data(), K * sizeof(int));
  for (int k = 0; k < BK_CLASSES; ++k) {
    std::memcpy(h + o_first[k], first[k].data(), (K + 1) * sizeof(int));
    int* blk = reinterpret_cast<int*>(h + o_blk[k]);
    for (int q = 0; q < K; ++q)
      for (int x = first[k][q]; x < first[k][q + 1]; ++x) blk[x] = q;
  }
  SW_HIP(hipMemcpyAsync(b->dev, b->host[sl], bytes, hipMemcpyHostToDevice, stream));
  if (!b->staged[sl]) SW_HIP(hipEventCreateWithFlags(&b->staged[sl], hipEventDisableTiming));
  SW_HIP(hipEventRecord(b->staged[sl], stream));
  b->stream = stream;
  b->used = true;

  // ---- launches
  const char* d = static_cast<const char*>(b->dev);
  BatchLaunch<T> L[BK_CLASSES];
  for (int k = 0; k < BK_CLASSES; ++k) {
    L[k].tab = reinterpret_cast<const StageArgs<T>*>(d + o_tab);
    L[k].nbb = reinterpret_cast<const int*>(d + o_nbb);
    L[k].first = reinterpret_cast<const int*>(d + o_first[k]);
    L[k].blk2m = reinterpret_cast<const int*>(d + o_blk[k]);
    L[k].count = K;
  }
  // one substep k of the members [0, active) on stream s
  auto substep = [&](int k, int active, hipStream_t s) -> int {
    for (int st = 0; st < rk; ++st) {
      for (auto& l : L) {
        l.stage = st;
        l.first_sub = (k == 0);
        l.parity = k & 1;
      }
      // edge values: the first stage's own pass; later stages' come from the
      // previous stage's last launch
      if (k == 0 && st == 0)
        hipLaunchKernelGGL(k_edge_batch<T>, dim3(first[BK_EDGE][active]), dim3(256), 0, s, L[BK_EDGE]);
      // stage 1 of a substep reads its *p fields at the addresses of h, u, v
      // for every member, except on the first substep unless every member's
      // input aliases them
      const bool samep = st == 0 && (k > 0 || in_alias);
      const dim3 ga(first[BK_MARCH_A][active]), gb(first[BK_MARCH_B][active]);
      if (fused) {
        const dim3 gf(first[BK_FUSED][active]), gr(first[BK_FRAME][active]);
        if (gsym && samep)
          hipLaunchKernelGGL((k_fused_batch<T, true, true>), gf, dim3(256), 0, s, L[BK_FUSED]);
        else if (gsym)
          hipLaunchKernelGGL((k_fused_batch<T, true, false>), gf, dim3(256), 0, s, L[BK_FUSED]);
        else if (samep)
          hipLaunchKernelGGL((k_fused_batch<T, false, true>), gf, dim3(256), 0, s, L[BK_FUSED]);
        else
          hipLaunchKernelGGL((k_fused_batch<T, false, false>), gf, dim3(256), 0, s, L[BK_FUSED]);
        if (fedge) {
          // frame cells, both halo fills and the next edge pass per cube edge
          const dim3 ge(first[BK_FEDGE][active]);
          constexpr int nt = EdgeGeo<kBatchSeg>::threads;
          if (gsym)
            hipLaunchKernelGGL((k_frame_edge_batch<T, true>), ge, dim3(nt), 0, s, L[BK_FEDGE]);
          else
            hipLaunchKernelGGL((k_frame_edge_batch<T, false>), ge, dim3(nt), 0, s, L[BK_FEDGE]);
          SW_HIP(hipGetLastError());
          continue;
        }
        // frame cells and the new height's halo fill in one launch
        if (gsym)
          hipLaunchKernelGGL((k_frame_batch<T, true>), gr, dim3(256), 0, s, L[BK_FRAME]);
        else
          hipLaunchKernelGGL((k_frame_batch<T, false>), gr, dim3(256), 0, s, L[BK_FRAME]);
        hipLaunchKernelGGL(k_halo_uv_edge_batch<T>, dim3(first[BK_HUVE][active]), dim3(256), 0, s, L[BK_HUVE]);
        SW_HIP(hipGetLastError());
        continue;
      }
      if (gsym && samep)
        hipLaunchKernelGGL((k_march_a_batch<T, true, true>), ga, dim3(256), 0, s, L[BK_MARCH_A]);
      else if (gsym)
        hipLaunchKernelGGL((k_march_a_batch<T, true, false>), ga, dim3(256), 0, s, L[BK_MARCH_A]);
      else if (samep)
        hipLaunchKernelGGL((k_march_a_batch<T, false, true>), ga, dim3(256), 0, s, L[BK_MARCH_A]);
      else
        hipLaunchKernelGGL((k_march_a_batch<T, false, false>), ga, dim3(256), 0, s, L[BK_MARCH_A]);
      hipLaunchKernelGGL(k_halo_scalar_batch<T>, dim3(first[BK_HALO][active]), dim3(256), 0, s, L[BK_HALO]);
      if (gsym && samep)
        hipLaunchKernelGGL((k_march_b_batch<T, true, true>), gb, dim3(256), 0, s, L[BK_MARCH_B]);
      else if (gsym)
        hipLaunchKernelGGL((k_march_b_batch<T, true, false>), gb, dim3(256), 0, s, L[BK_MARCH_B]);
      else if (samep)
        hipLaunchKernelGGL((k_march_b_batch<T, false, true>), gb, dim3(256), 0, s, L[BK_MARCH_B]);
      else
        hipLaunchKernelGGL((k_march_b_batch<T, false, false>), gb, dim3(256), 0, s, L[BK_MARCH_B]);
      hipLaunchKernelGGL(k_halo_uv_edge_batch<T>, dim3(first[BK_HUVE][active]), dim3(256), 0, s, L[BK_HUVE]);
      SW_HIP(hipGetLastError());
    }
    return SW_OK;
  };
  const int nmax = nbb[0];
  if ((flags & SW_GRAPH) && nmax >= 3) {
    // Substep 0 eagerly, then segments in which the members still stepping
    // (a prefix of the table, members sorted by decreasing nbb) stay the
    // same: each segment replays a recorded (odd, even) substep pair for its
    // prefix length, with an eager substep to reach an odd start or finish an
    // odd count.  A graph reads the tables from b->dev, so it stays valid for
    // any later run with the same table layout, grids and prefix.
    int rc = substep(0, K, stream);
    if (rc) return rc;
    auto pair_graph = [&](int active, hipGraphExec_t* out) -> int {
      std::vector<long long> sig = {reinterpret_cast<long long>(b->dev), static_cast<long long>(o_tab),
                                    static_cast<long long>(o_nbb), K, active, rk, gsym ? 1 : 0, fused ? 1 : 0,
                                    fedge ? 1 : 0, static_cast<long long>(sizeof(T))};
      for (int k = 0; k < BK_CLASSES; ++k) {
        sig.push_back(static_cast<long long>(o_first[k]));
        sig.push_back(static_cast<long long>(o_blk[k]));
        sig.push_back(first[k][active]);
      }
      for (const auto& g : b->graphs)
        if (g.sig == sig) {
          *out = g.exec;
          return SW_OK;
        }
      if (!b->capture) SW_HIP(hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking));
      SW_HIP(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal));
      int rc2 = substep(1, active, b->capture);
      if (!rc2) rc2 = substep(2, active, b->capture);
      hipGraph_t graph = nullptr;
      const hipError_t e = hipStreamEndCapture(b->capture, &graph);
      if (rc2) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc2;
      }
      if (e != hipSuccess) return fail(SW_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      hipGraphExec_t exec = nullptr;
      const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) return fail(SW_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
      if (b->graphs.size() >= kBatchGraphs) {
        // replays of the oldest graph may still be queued: destroy it once
        // the work launched so far on `stream` has completed
        hipEvent_t ev = nullptr;
        SW_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        SW_HIP(hipEventRecord(ev, stream));
        b->retired.push_back({b->graphs.front().exec, ev});
        b->graphs.erase(b->graphs.begin());
      }
      b->graphs.push_back({sig, exec});
      *out = exec;
      return SW_OK;
    };
    int k = 1, active = K;
    while (k < nmax) {
      while (active > 0 && nbb[active - 1] <= k) --active;
      const int kend = nbb[active - 1];  // every member of the prefix steps through kend - 1
      if ((k & 1) == 0 || kend - k < 2) {
        if ((rc = substep(k, active, stream))) return rc;
        ++k;
        continue;
      }
      const int pairs = (kend - k) / 2;
      hipGraphExec_t exec = nullptr;
      if ((rc = pair_graph(active, &exec))) return rc;
      for (int p = 0; p < pairs; ++p) SW_HIP(hipGraphLaunch(exec, stream));
      k += 2 * pairs;
    }
  } else {
    int active = K;
    for (int k = 0; k < nmax; ++k) {
      while (active > 0 && nbb[active - 1] <= k) --active;
      const int rc = substep(k, active, stream);
      if (rc) return rc;
    }
  }
  // ---- diagnostics of each member from the inputs of its last stage
  if (!(flags & SW_SKIP_DIAG)) {
    for (int q = 0; q < K; ++q) {
      const sw_member& m = mem[order[q]];
      sw_ctx* c = m.ctx;
      const int ls = rk == 2 ? 0 : 3;
      hipStream_t keep = c->stream;
      c->stream = stream;
      const int rc = run_diag<T>(c, scratch_field<T>(c, ls), scratch_field<T>(c, ls + 1),
                                 scratch_field<T>(c, ls + 2), m.out, rk - 1);
      c->stream = keep;
      if (rc) return rc;
    }
  }
  return SW_OK;
}

}  // namespace

extern "C" {

// Default band height of k_fused: the height that minimises (rounds of
// resident waves) x (rows one wave marches, its band plus the 6 rows of
// overlap).  Resident waves per CU: 8 in FP64 (two 4-wave blocks, 2 waves per
// SIMD by registers), 12 in FP32 (3 per SIMD).
int default_rows_f(int nx, int cus, int dtype) {
  const long S = (nx + kFStride - 1) / kFStride;
  const long slots = (dtype == SW_F64 ? 8L : 12L) * (cus > 0 ? cus : 256);
  // beyond the list, the lowest height whose waves all fit one round (C1536:
  // 140 rows FP64, 91 FP32, measured fastest there among 28 ... 200 rows)
  const long nb1 = slots / (6 * S);
  const int r1 = nb1 >= 1 && (nx + nb1 - 1) / nb1 > 64 ? static_cast<int>((nx + nb1 - 1) / nb1) : 0;
  int best = 32;
  long best_cost = -1;
  for (int R : {8, 12, 16, 20, 24, 28, 32, 36, 40, 48, 64, r1}) {
    if (R <= 0) continue;
    const long waves = 6 * S * ((nx + R - 1) / R);
    const long cost = ((waves + slots - 1) / slots) * (std::min(R, nx) + 6);
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = R;
    }
  }
  return best;
}

// Default split of the fused kernel's bands (sw_set_band_split) for the
// uniform height *rows_f: when a launch of uniform bands is one round of
// resident blocks with more blocks than CUs, the first block on each CU
// finishes far ahead of the others, so pairs of 2 rows_f rows are split into
// a long band (69 % in FP64, 60 % in FP32: the measured optima at C384 and
// C768) for the first blocks and a short one for the rest.  The long bands'
// blocks must be the older ones on their CUs (FP64: at most one per CU);
// bands taller than 64 rows stay uniform (at C1536 the splits measured no
// faster than the uniform one-round height).  Returns the short height (0:
// uniform) and sets *rows_f to the long one.
int default_rows_f2(int nx, int cus, int dtype, int* rows_f) {
  const long S = (nx + kFStride - 1) / kFStride;
  const long blocks = (6 * S * ((nx + *rows_f - 1) / *rows_f) + 3) / 4;
  const long per_cu = dtype == SW_F64 ? 2 : 3;  // resident 4-wave blocks per CU
  if (cus <= 0 || blocks <= cus || blocks > per_cu * cus || *rows_f > 64) return 0;
  const int span = 2 * *rows_f;
  const int lng = static_cast<int>(std::lround(span * (dtype == SW_F64 ? 0.69 : 0.60)));
  if (lng < 4 || span - lng < 4 || lng >= nx) return 0;
  const long long_blocks = (6 * S * ((nx + span - 1) / span) + 3) / 4;
  if (long_blocks > (per_cu - 1) * cus) return 0;
  *rows_f = lng;
  return span - lng;
}

int sw_create(sw_ctx** out, int nx, int dtype, const void* const consts[SW_N_CONSTS], void* stream) {
  if (!out) return fail(SW_E_ARG, "out is null");
  *out = nullptr;
  if (nx < 7 || nx > 8192) return fail(SW_E_ARG, "nx must be in [7, 8192]");
  if (dtype != SW_F64 && dtype != SW_F32) return fail(SW_E_ARG, "dtype must be SW_F64 or SW_F32");
  int rc = check_ptrs(consts, SW_N_CONSTS, "consts");
  if (rc) return rc;
  sw_ctx* c = new sw_ctx();
  c->nx = nx;
  c->N = nx + 2 * HALO;
  c->dtype = dtype;
  c->stream = static_cast<hipStream_t>(stream);
  c->cells = 6 * static_cast<size_t>(c->N) * c->N;
  {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 0;
    c->rows_f = default_rows_f(nx, cus, dtype);
    c->rows_f2 = default_rows_f2(nx, cus, dtype, &c->rows_f);
    c->device = dev;
  }
  // this device's address of the no-allocation workspace (once per device)
  rc = scratch_address(c->device, &c->noalloc);
  if (!rc) rc = validate_consts(c, consts);
  if (!rc) rc = (dtype == SW_F64) ? do_create_planes<double>(c, consts) : do_create_planes<float>(c, consts);
  if (!rc) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = fail(SW_E_HIP, std::string("sw_create: ") + hipGetErrorString(e));
  }
  if (rc) {
    free_ctx(c);
    return rc;
  }
  *out = c;
  return SW_OK;
}

void sw_destroy(sw_ctx* ctx) {
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  free_ctx(ctx);
}

int sw_create_like(sw_ctx** out, const sw_ctx* base) {
  if (!out) return fail(SW_E_ARG, "out is null");
  *out = nullptr;
  if (!base) return fail(SW_E_ARG, "base is null");
  sw_ctx* c = new sw_ctx();
  c->nx = base->nx;
  c->N = base->N;
  c->dtype = base->dtype;
  c->stream = base->stream;
  c->cells = base->cells;
  c->device = base->device;
  c->noalloc = base->noalloc;
  c->grid = base->grid;
  c->planes = base->planes;
  c->locs = base->locs;
  c->plan = base->plan;
  c->plan_n = base->plan_n;
  c->frame_edge_ok = base->frame_edge_ok;
  c->frame_edge = base->frame_edge;
  c->dzx_s = base->dzx_s;
  c->dzy_s = base->dzy_s;
  c->dz_uniform = base->dz_uniform;
  c->shared = base->shared;
  c->gsym = base->gsym;
  c->rows_a = base->rows_a;
  c->rows_b = base->rows_b;
  c->rows_f = base->rows_f;
  c->rows_f2 = base->rows_f2;
  int rc = c->dtype == SW_F64 ? alloc_workspace<double>(c) : alloc_workspace<float>(c);
  if (!rc) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = fail(SW_E_HIP, std::string("sw_create_like: ") + hipGetErrorString(e));
  }
  if (rc) {
    free_ctx(c);
    return rc;
  }
  *out = c;
  return SW_OK;
}

int sw_set_stream(sw_ctx* ctx, void* stream) {
  if (!ctx) return fail(SW_E_ARG, "ctx is null");
  ctx->stream = static_cast<hipStream_t>(stream);
  return SW_OK;
}

int sw_rk2(sw_ctx* ctx, const void* const in[6], void* const out[13], int nbb, double dtb, int flags) {
  if (!ctx) return fail(SW_E_ARG, "ctx is null");
  return ctx->dtype == SW_F64 ? run_rk<double>(ctx, in, out, nbb, dtb, flags, 2)
                              : run_rk<float>(ctx, in, out, nbb, dtb, flags, 2);
}

int sw_rk3(sw_ctx* ctx, const void* const in[6], void* const out[13], int nbb, double dtb, int flags) {
  if (!ctx) return fail(SW_E_ARG, "ctx is null");
  return ctx->dtype == SW_F64 ? run_rk<double>(ctx, in, out, nbb, dtb, flags, 3)
                              : run_rk<float>(ctx, in, out, nbb, dtb, flags, 3);
}

int sw_stage(sw_ctx* ctx, const void* const in[6], void* const out[13], double dt, double beta, int is_last) {
  if (!ctx) return fail(SW_E_ARG, "ctx is null");
  int rc = check_ptrs(in, 6, "in");
  if (rc) return rc;
  rc = check_ptrs(const_cast<const void* const*>(out), 13, "out");
  if (rc) return rc;
  rc = check_disjoint(ctx, in, out);
  if (rc) return rc;
  if (ctx->dtype == SW_F64) {
    auto p = [&](int k) { return static_cast<const double*>(in[k]); };
    rc = run_stage<double>(ctx, p(0), p(1), p(2), p(3), p(4), p(5), static_cast<double*>(out[0]),
                           static_cast<double*>(out[1]), static_cast<double*>(out[2]), dt, beta, is_last ? 1 : 0,
                           0, 2, true);
    if (rc) return rc;
    return run_diag<double>(ctx, p(0), p(2), p(3), out, 0);
  }
  auto p = [&](int k) { return static_cast<const float*>(in[k]); };
  rc = run_stage<float>(ctx, p(0), p(1), p(2), p(3), p(4), p(5), static_cast<float*>(out[0]),
                        static_cast<float*>(out[1]), static_cast<float*>(out[2]), dt, beta, is_last ? 1 : 0,
                        0, 2, true);
  if (rc) return rc;
  return run_diag<float>(ctx, p(0), p(2), p(3), out, 0);
}

int sw_halo_scalar(sw_ctx* ctx, void* field, int ncomp) {
  if (!ctx) return fail(SW_E_ARG, "ctx is null");
  if (!field) return fail(SW_E_ARG, "field is null");
  if (ncomp < 1 || ncomp > 64) return fail(SW_E_ARG, "ncomp must be in [1, 64]");
  return ctx->dtype == SW_F64 ? launch_halo_scalar<double>(ctx, static_cast<double*>(field), ncomp)
                              : launch_halo_scalar<float>(ctx, static_cast<float*>(field), ncomp);
}

int sw_halo_vector(sw_ctx* ctx, void* u, void* v) {
  if (!ctx) return fail(SW_E_ARG, "ctx is null");
  if (!u || !v || u == v) return fail(SW_E_ARG, "u, v must be distinct non-null fields");
  return ctx->dtype == SW_F64 ? launch_halo_vector<double>(ctx, static_cast<double*>(u), static_cast<double*>(v))
                              : launch_halo_vector<float>(ctx, static_cast<float*>(u), static_cast<float*>(v));
}

int sw_edge_avg(sw_ctx* ctx, void* u, void* v) {
  if (!ctx) return fail(SW_E_ARG, "ctx is null");
  if (!u || !v || u == v) return fail(SW_E_ARG, "u, v must be distinct non-null fields");
  const size_t n = 24 * static_cast<size_t>(ctx->nx);
  if (ctx->dtype == SW_F64) {
    EdgeAvgArgs<double> a{geo_of(ctx), route(), static_cast<double*>(u), static_cast<double*>(v),
                          edge_buf<double>(ctx, kEdgeScratch)};
    hipLaunchKernelGGL(k_edge_avg_read<double>, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, a);
    hipLaunchKernelGGL(k_edge_avg_write<double>, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, a);
  } else {
    EdgeAvgArgs<float> a{geo_of(ctx), route(), static_cast<float*>(u), static_cast<float*>(v),
                         edge_buf<float>(ctx, kEdgeScratch)};
    hipLaunchKernelGGL(k_edge_avg_read<float>, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, a);
    hipLaunchKernelGGL(k_edge_avg_write<float>, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, a);
  }
  SW_HIP(hipGetLastError());
  return SW_OK;
}

int sw_sync(sw_ctx* ctx) {
  if (!ctx) return fail(SW_E_ARG, "ctx is null");
  SW_HIP(hipStreamSynchronize(ctx->stream));
  return SW_OK;
}

int sw_routing_table(int32_t out[6 * 4 * 3]) {
  if (!out) return fail(SW_E_ARG, "out is null");
  const Route& R = route();
  for (int t = 0; t < 6; ++t)
    for (int d = 0; d < 4; ++d)
      for (int k = 0; k < 3; ++k) out[(t * 4 + d) * 3 + k] = R.r[t][d][k];
  return SW_OK;
}

int sw_profile(sw_ctx* ctx, const void* const in[6], void* const out[13], int nbb, double dtb,
               int reps, double ms[SW_N_KERNEL_CLASSES], int counts[SW_N_KERNEL_CLASSES]) {
  if (!ctx || !ms || !counts) return fail(SW_E_ARG, "null argument");
  if (reps < 1 || reps > 1000) return fail(SW_E_ARG, "reps must be in [1, 1000]");
  std::vector<sw_ctx::ProfRec> recs;
  ctx->prof = &recs;
  ctx->prof_reps = reps;
  int rc = ctx->dtype == SW_F64 ? run_rk<double>(ctx, in, out, nbb, dtb, 0, 2)
                                : run_rk<float>(ctx, in, out, nbb, dtb, 0, 2);
  ctx->prof = nullptr;
  hipError_t e = hipStreamSynchronize(ctx->stream);
  for (int k = 0; k < K_CLASSES; ++k) {
    ms[k] = 0.0;
    counts[k] = 0;
  }
  for (auto& r : recs) {
    float t = 0.0f;
    if (!rc && e == hipSuccess && hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
      ms[r.cls] += t;
      counts[r.cls] += r.reps;
    }
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  if (rc) return rc;
  if (e != hipSuccess) return fail(SW_E_HIP, std::string("sw_profile: ") + hipGetErrorString(e));
  return SW_OK;
}

int sw_profile_graph(sw_ctx* ctx, const void* const in[6], void* const out[13], int pairs, int replays,
                     double dtb, double ms[SW_N_KERNEL_CLASSES], int counts[SW_N_KERNEL_CLASSES],
                     double* graph_ms) {
  if (!ctx || !ms || !counts || !graph_ms) return fail(SW_E_ARG, "null argument");
  if (pairs < 1 || pairs > 64) return fail(SW_E_ARG, "pairs must be in [1, 64]");
  if (replays < 1 || replays > 10000) return fail(SW_E_ARG, "replays must be in [1, 10000]");
  if (!(dtb == dtb) || dtb > 1e300 || dtb < -1e300) return fail(SW_E_ARG, "dtb must be finite");
  return ctx->dtype == SW_F64 ? run_profile_graph<double>(ctx, in, out, pairs, replays, dtb, ms, counts, graph_ms)
                              : run_profile_graph<float>(ctx, in, out, pairs, replays, dtb, ms, counts, graph_ms);
}

int sw_integrals(sw_ctx* ctx, const void* h, const void* u, const void* v, double out[2]) {
  if (!ctx || !h || !u || !v || !out) return fail(SW_E_ARG, "null argument");
  return ctx->dtype == SW_F64 ? run_integrals<double>(ctx, h, u, v, out)
                              : run_integrals<float>(ctx, h, u, v, out);
}

int sw_integrals_noalloc(sw_ctx* ctx, const void* h, const void* u, const void* v, double* out) {
  if (!ctx || !h || !u || !v || !out) return fail(SW_E_ARG, "null argument");
  return ctx->dtype == SW_F64 ? run_integrals_noalloc<double>(ctx, h, u, v, out)
                              : run_integrals_noalloc<float>(ctx, h, u, v, out);
}

int sw_get_tuning(const sw_ctx* ctx, int* rows_a, int* rows_b, int* rows_f) {
  if (!ctx || !rows_a || !rows_b || !rows_f) return fail(SW_E_ARG, "null argument");
  *rows_a = ctx->rows_a;
  *rows_b = ctx->rows_b;
  *rows_f = ctx->rows_f;
  return SW_OK;
}

int sw_set_tuning(sw_ctx* ctx, int rows_a, int rows_b, int rows_f) {
  if (!ctx) return fail(SW_E_ARG, "ctx is null");
  if (rows_a < 4 || rows_a > 4096 || rows_b < 4 || rows_b > 4096 || (rows_f != 0 && (rows_f < 4 || rows_f > 4096)))
    return fail(SW_E_ARG, "rows per wave must be in [4, 4096] (rows_f may be 0)");
  if (rows_a == ctx->rows_a && rows_b == ctx->rows_b && rows_f == ctx->rows_f) return SW_OK;
  ctx->rows_a = rows_a;
  ctx->rows_b = rows_b;
  ctx->rows_f = rows_f;
  return drop_graphs(ctx);
}

int sw_set_band_split(sw_ctx* ctx, int rows_short) {
  if (!ctx) return fail(SW_E_ARG, "ctx is null");
  if (rows_short != 0 && (rows_short < 4 || rows_short > 4096))
    return fail(SW_E_ARG, "rows_short must be 0 or in [4, 4096]");
  if (rows_short == ctx->rows_f2) return SW_OK;
  ctx->rows_f2 = rows_short;
  return drop_graphs(ctx);
}

int sw_get_band_split(const sw_ctx* ctx, int* rows_short) {
  if (!ctx || !rows_short) return fail(SW_E_ARG, "null argument");
  *rows_short = ctx->rows_f2;
  return SW_OK;
}

int sw_set_frame_form(sw_ctx* ctx, int cube_edge) {
  if (!ctx) return fail(SW_E_ARG, "ctx is null");
  if (cube_edge != 0 && cube_edge != 1) return fail(SW_E_ARG, "cube_edge must be 0 or 1");
  if (cube_edge && !ctx->frame_edge_ok)
    return fail(SW_E_ARG, "the grid's halo sources lie farther than 3 positions: cube-edge frame form unavailable");
  if (cube_edge == ctx->frame_edge) return SW_OK;
  ctx->frame_edge = cube_edge;
  return drop_graphs(ctx);
}

int sw_get_frame_form(const sw_ctx* ctx, int* cube_edge) {
  if (!ctx || !cube_edge) return fail(SW_E_ARG, "null argument");
  *cube_edge = ctx->frame_edge;
  return SW_OK;
}

int sw_layout_flags(const sw_ctx* ctx, int* flags) {
  if (!ctx || !flags) return fail(SW_E_ARG, "null argument");
  *flags = (ctx->dz_uniform ? 1 : 0) | (ctx->shared ? 2 : 0) | (ctx->gsym ? 4 : 0);
  return SW_OK;
}

int sw_info(const sw_ctx* ctx, int* nx, int* dtype) {
  if (!ctx) return fail(SW_E_ARG, "ctx is null");
  if (nx) *nx = ctx->nx;
  if (dtype) *dtype = ctx->dtype;
  return SW_OK;
}

int sw_batch_create(sw_batch** out) {
  if (!out) return fail(SW_E_ARG, "out is null");
  *out = new sw_batch();
  return SW_OK;
}

void sw_batch_destroy(sw_batch* b) {
  if (!b) return;
  if (b->used && b->stream) (void)hipStreamSynchronize(b->stream);
  for (int k = 0; k < 2; ++k)
    if (b->staged[k]) {
      (void)hipEventSynchronize(b->staged[k]);
      (void)hipEventDestroy(b->staged[k]);
    }
  b->reap(true);
  for (auto& g : b->graphs) (void)hipGraphExecDestroy(g.exec);
  if (b->capture) (void)hipStreamDestroy(b->capture);
  if (b->dev) (void)hipFree(b->dev);
  for (int k = 0; k < 2; ++k)
    if (b->host[k]) (void)hipHostFree(b->host[k]);
  delete b;
}

int sw_batch_run(sw_batch* b, const sw_member* members, int count, int rk, int flags, void* stream) {
  if (!b) return fail(SW_E_ARG, "batch is null");
  if (!members) return fail(SW_E_ARG, "members is null");
  if (count < 1 || count > kBatchMax) return fail(SW_E_ARG, "count must be in [1, 4096]");
  if (rk != 2 && rk != 3) return fail(SW_E_ARG, "rk must be 2 or 3");
  if (!members[0].ctx) return fail(SW_E_ARG, "member 0: ctx is null");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return members[0].ctx->dtype == SW_F64 ? run_batch<double>(b, members, count, rk, flags, s)
                                         : run_batch<float>(b, members, count, rk, flags, s);
}

const char* sw_last_error(void) { return g_err.c_str(); }

}  // extern "C"
