// ns-scheme kernels, templated on the compute type T of slopes, limiter,
// reconstruction and face fluxes (double: FP64 step; float: FP32-mixed step).
// Geometry is read as FP64, the face-flux sum, the update and the p/T
// recovery are FP64 in both forms.
//
// Three launches per step, gathers without atomics on the state
// (deterministic, numbering independent):
//   k_ns_grad    per cell: least-squares gradient of (rho,u,v,p) +
//                Venkatakrishnan limiter, stored as the cell's record
//   k_ns_face    one flux per face (owner = lower caller id): reconstruction of
//                both sides, AUSM+/Roe, viscous stress/heat flux, wall force
//   k_ns_update  per cell: sum of +-face fluxes in slot order (FP64), explicit
//                Euler update, primitive recovery, next-dt partials
// Everything read across a face is stored as one record per cell so that a
// neighbour gather is one contiguous load: primitives [nc][4] (16 B in FP32)
// and the face-reconstruction record [nc][16] = (q, d/dx d/dy of each of
// rho,u,v,p, limiter of each) (64 B in FP32).  Per-cell geometry is
// face-slot-major SoA, read unit-stride.
#pragma once
#include <hip/hip_runtime.h>

#include "gd_fold.hpp"
#include "gd_internal.hpp"

namespace gd {
namespace ns {

constexpr int kBlock = kThreads;

template <class T>
struct NsKArgs {
  MeshView m;
  gd_params prm;
  const double* Uo;
  double* Un;
  const T* Q;       // current primitives (rho,u,v,p) [nc][4]
  T* Qw;            // next primitives [nc][4]
  T* rec;           // [nc][16]: q[4], (d/dx, d/dy) of rho,u,v,p [8], limiter [4]
  T* fflux;         // [n_faces][4]: flux leaving the face's owner
  Acc* acc;
  double* part;     // per-block partials of the next dt [3 x blocks]
  int32_t step;
  int32_t dt_policy;
  double dt, cfl;
  double* wall_f;
  int32_t wall_code;
  int32_t fold_prev;  // see NsStepArgs
  int64_t nparts;
  double* drag;
};

template <class T>
NsKArgs<T> kargs(const NsStepArgs& a) {
  NsKArgs<T> k;
  k.m = a.m;
  k.prm = a.prm;
  k.Uo = a.Uo;
  k.Un = a.Un;
  k.Q = static_cast<const T*>(a.Q);
  k.Qw = static_cast<T*>(a.Qw);
  k.rec = static_cast<T*>(a.rec);
  k.fflux = static_cast<T*>(a.fflux);
  k.acc = a.acc;
  k.part = a.part;
  k.step = a.step;
  k.dt_policy = a.dt_policy;
  k.dt = a.dt;
  k.cfl = a.cfl;
  k.wall_f = a.wall_f;
  k.wall_code = a.wall_code;
  k.fold_prev = a.fold_prev;
  k.nparts = a.nparts;
  k.drag = a.drag;
  return k;
}

// 4 consecutive values: one 16 B load (float) or two (double)
__device__ __forceinline__ void ld4(const float* p, float* o) {
  const float4 v = *reinterpret_cast<const float4*>(p);
  o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
}
__device__ __forceinline__ void ld4(const double* p, double* o) {
  const double2 a = *reinterpret_cast<const double2*>(p), b = *reinterpret_cast<const double2*>(p + 2);
  o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
}
__device__ __forceinline__ void st4(float* p, float a, float b, float c, float d) {
  *reinterpret_cast<float4*>(p) = make_float4(a, b, c, d);
}
__device__ __forceinline__ void st4(double* p, double a, double b, double c, double d) {
  *reinterpret_cast<double2*>(p) = make_double2(a, b);
  *reinterpret_cast<double2*>(p + 2) = make_double2(c, d);
}
template <class T>
__device__ __forceinline__ void ld16(const T* p, T* o) {
  ld4(p, o); ld4(p + 4, o + 4); ld4(p + 8, o + 8); ld4(p + 12, o + 12);
}

// Sutherland's law mu_ref (T/T_ref)^(3/2) (T_ref + S)/(T + S), with
// x^(3/2) = x sqrt(x)
template <class T>
__device__ __forceinline__ T sutherland(T Tt, const gd_params& p) {
  const T Tr = (T)p.T_ref, S = (T)p.S_suth;
  const T t = Tt / Tr;
  return (T)p.mu_ref * (t * sqrt(t)) * (Tr + S) / (Tt + S);
}

// ghost primitive state of a boundary face (first order)
template <class T>
__device__ __forceinline__ void ghost(const gd_params& p, int code, const T* q, T nx, T ny, T* g) {
  const int idx = -code;
  const int kind = (idx >= 0 && idx <= GD_MAX_BC) ? p.bc[idx].kind : GD_BC_EXTRAPOLATE;
  if (kind == GD_BC_FIXED) {
#pragma unroll
    for (int v = 0; v < 4; ++v) g[v] = (T)p.bc[idx].data[v];
  } else if (kind == GD_BC_SLIP_WALL || kind == GD_BC_ISOTHERMAL_WALL) {
    const T un = q[1] * nx + q[2] * ny;
    g[0] = q[0]; g[1] = q[1] - (T)2 * un * nx; g[2] = q[2] - (T)2 * un * ny; g[3] = q[3];
  } else {
#pragma unroll
    for (int v = 0; v < 4; ++v) g[v] = q[v];
  }
}

// AUSM+ per unit length (Liou 1996): neighbour side "L", inward normal; -(F.n)
template <class T>
__device__ __forceinline__ void ausm(const T* L, const T* R, T nx, T ny, T g, T* out) {
  const T rL = L[0], uL = L[1], vL = L[2], pL = L[3];
  const T rR = R[0], uR = R[1], vR = R[2], pR = R[3];
  const T htL = (T)0.5 * (uL * uL + vL * vL) + g / (g - 1) * pL / rL;
  const T htR = (T)0.5 * (uR * uR + vR * vR) + g / (g - 1) * pR / rR;
  const T c = 2 * (g - 1) / (g + 1);
  const T unL = uL * nx + vL * ny, unR = uR * nx + vR * ny;
  const T asL = sqrt(c * htL), asR = sqrt(c * htR);
  const T aL = asL * fmin((T)1, asL / fabs(unL));
  const T aR = asR * fmin((T)1, asR / fabs(unR));
  const T a12 = fmin(aL, aR);
  const T ML = unL / a12, MR = unR / a12;
  T mp, mm, pp, pm;
  {
    const T d = ML * ML - 1;
    mp = fabs(ML) >= 1 ? (T)0.5 * (ML + fabs(ML)) : (T)0.25 * (ML + 1) * (ML + 1) + (T)0.125 * d * d;
    pp = fabs(ML) >= 1 ? (ML > 0 ? (T)1 : (T)0) : (T)0.25 * (ML + 1) * (ML + 1) * (2 - ML) + (T)0.1875 * ML * d * d;
  }
  {
    const T d = MR * MR - 1;
    mm = fabs(MR) >= 1 ? (T)0.5 * (MR - fabs(MR)) : -(T)0.25 * (MR - 1) * (MR - 1) - (T)0.125 * d * d;
    pm = fabs(MR) >= 1 ? (MR < 0 ? (T)1 : (T)0) : (T)0.25 * (MR - 1) * (MR - 1) * (2 + MR) - (T)0.1875 * MR * d * d;
  }
  const T m = mp + mm;
  const T p12 = pp * pL + pm * pR;
  const T mpl = (T)0.5 * (m + fabs(m)), mmi = (T)0.5 * (m - fabs(m));
  out[0] = -(a12 * (mpl * rL + mmi * rR));
  out[1] = -(a12 * (mpl * rL * uL + mmi * rR * uR) + p12 * nx);
  out[2] = -(a12 * (mpl * rL * vL + mmi * rR * vR) + p12 * ny);
  out[3] = -(a12 * (mpl * rL * htL + mmi * rR * htR));
}

// Roe per unit length, Harten fix of width 0.2 a on the acoustic waves
template <class T>
__device__ __forceinline__ void roe(const T* L, const T* R, T nx, T ny, T g, T* out) {
  const T rL = L[0], uL = L[1], vL = L[2], pL = L[3];
  const T rR = R[0], uR = R[1], vR = R[2], pR = R[3];
  const T mx = -ny, my = nx;
  const T unL = uL * nx + vL * ny, umL = uL * mx + vL * my;
  const T unR = uR * nx + vR * ny, umR = uR * mx + vR * my;
  const T HL = g / (g - 1) * pL / rL + (T)0.5 * (uL * uL + vL * vL);
  const T HR = g / (g - 1) * pR / rR + (T)0.5 * (uR * uR + vR * vR);
  const T RT = sqrt(rR / rL);
  const T rho = RT * rL;
  const T u = (uL + RT * uR) / (1 + RT), v = (vL + RT * vR) / (1 + RT), H = (HL + RT * HR) / (1 + RT);
  const T a = sqrt((g - 1) * (H - (T)0.5 * (u * u + v * v)));
  const T un = u * nx + v * ny, um = u * mx + v * my;
  const T dp = pR - pL, dun = unR - unL, dum = umR - umL, drho = rR - rL;
  const T w1 = (dp - rho * a * dun) / (2 * a * a), w2 = rho * dum, w3 = drho - dp / (a * a),
          w4 = (dp + rho * a * dun) / (2 * a * a);
  T s1 = fabs(un - a), s4 = fabs(un + a);
  const T s2 = fabs(un);
  const T dw = (T)0.2 * a;
  if (s1 < dw) s1 = (T)0.5 * (s1 * s1 / dw + dw);
  if (s4 < dw) s4 = (T)0.5 * (s4 * s4 / dw + dw);
  const T d1 = s1 * w1 + s2 * w3 + s4 * w4;
  const T d2 = s1 * w1 * (u - a * nx) + s2 * w2 * mx + s2 * w3 * u + s4 * w4 * (u + a * nx);
  const T d3 = s1 * w1 * (v - a * ny) + s2 * w2 * my + s2 * w3 * v + s4 * w4 * (v + a * ny);
  const T d4 = s1 * w1 * (H - un * a) + s2 * w2 * um + s2 * w3 * (T)0.5 * (u * u + v * v) + s4 * w4 * (H + un * a);
  out[0] = -(T)0.5 * (rL * unL + rR * unR - d1);
  out[1] = -(T)0.5 * (rL * unL * uL + pL * nx + rR * unR * uR + pR * nx - d2);
  out[2] = -(T)0.5 * (rL * unL * vL + pL * ny + rR * unR * vR + pR * ny - d3);
  out[3] = -(T)0.5 * (rL * unL * HL + rR * unR * HR - d4);
}

// primitives (rho, u, v, p) from conservative, in FP64
__device__ __forceinline__ void prims64(const double* U, int64_t nc, int64_t i, double g, double* q) {
  const double r = U[i], mx = U[nc + i], my = U[2 * nc + i], E = U[3 * nc + i];
  q[0] = r;
  q[1] = mx / r;
  q[2] = my / r;
  q[3] = (g - 1.0) * (E - 0.5 * (mx * mx + my * my) / r);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_ns_prims(int64_t nc, double g, const double* U, T* Q) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= nc) return;
  double q[4];
  prims64(U, nc, i, g, q);
  st4(Q + 4 * i, (T)q[0], (T)q[1], (T)q[2], (T)q[3]);
}

// face geometry at the compute type: FP64 arrays for the FP64 step, the float
// copies for the FP32-mixed step.  Geometry is read once per step, so it is
// loaded non-temporally and leaves the caches to the gathered cell records;
// so are the neighbour and face ids, the old state and 1/A (FP64 step
// 0.2981 -> 0.2929 ms, FP32-mixed 0.1489 -> 0.1430 ms on one box).
template <class U>
__device__ __forceinline__ U ntl(const U* p) { return __builtin_nontemporal_load(p); }

template <class T>
struct FaceGeo {
  static constexpr bool F = sizeof(T) == 4;
  __device__ __forceinline__ static T nx(const MeshView& m, int64_t s) { return F ? ntl(&m.fnx32[s]) : ntl(&m.fnx[s]); }
  __device__ __forceinline__ static T ny(const MeshView& m, int64_t s) { return F ? ntl(&m.fny32[s]) : ntl(&m.fny[s]); }
  // face midpoint - own centroid, component c of slot k
  __device__ __forceinline__ static T rf(const MeshView& m, int k, int c, int64_t i) {
    const int64_t o = (2 * k + c) * m.nc + i;
    return F ? ntl(&m.rf32[o]) : ntl(&m.rf[o]);
  }
  __device__ __forceinline__ static T lsq(const MeshView& m, int k, int c, int64_t i) {
    const int64_t o = (2 * k + c) * m.nc + i;
    return F ? ntl(&m.lsq32[o]) : ntl(&m.lsq[o]);
  }
  __device__ __forceinline__ static T hmin(const MeshView& m, int64_t i) { return F ? ntl(&m.hmin32[i]) : ntl(&m.hmin[i]); }
  // face-list geometry: the owner's inward normal, length, offsets
  __device__ __forceinline__ static T fnx(const MeshView& m, int64_t f) { return F ? ntl(&m.f_nx32[f]) : ntl(&m.f_nx[f]); }
  __device__ __forceinline__ static T fny(const MeshView& m, int64_t f) { return F ? ntl(&m.f_ny32[f]) : ntl(&m.f_ny[f]); }
  __device__ __forceinline__ static T flen(const MeshView& m, int64_t f) {
    return F ? ntl(&m.f_len32[f]) : ntl(&m.f_len[f]);
  }
  // face midpoint - owner centroid (fr0) and - neighbour centroid (fr1), component c
  __device__ __forceinline__ static T fr0(const MeshView& m, int c, int64_t f) {
    return F ? ntl(&m.f_r032[c * m.n_faces + f]) : ntl(&m.f_r0[c * m.n_faces + f]);
  }
  __device__ __forceinline__ static T fr1(const MeshView& m, int c, int64_t f) {
    return F ? ntl(&m.f_r132[c * m.n_faces + f]) : ntl(&m.f_r1[c * m.n_faces + f]);
  }
};

// The face-reconstruction record of cell i: (q, d/dx d/dy of rho,u,v,p,
// limiter of each), as 4 groups of 4 values in the order stored.  Every load
// first: neighbour ids, then all neighbour states (boundary and absent faces
// read the own cell, overwritten below), then geometry.
template <class T, bool QUADS>
__device__ __forceinline__ void cell_record(const NsKArgs<T>& a, int64_t i, T* r) {
  constexpr int NF = QUADS ? 4 : 3;
  using G = FaceGeo<T>;
  const MeshView& m = a.m;
  const int64_t nc = m.nc;
  const int nf = QUADS ? m.nv[i] : 3;
  int32_t jn[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) jn[k] = ntl(&m.nb[k * nc + i]);
  T qi[4], qk[NF][4];
  ld4(a.Q + 4 * i, qi);
#pragma unroll
  for (int k = 0; k < NF; ++k) ld4(a.Q + 4 * (int64_t)(jn[k] >= 0 ? jn[k] : i), qk[k]);
  T wx[NF], wy[NF], rx[NF], ry[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    wx[k] = G::lsq(m, k, 0, i); wy[k] = G::lsq(m, k, 1, i);
    rx[k] = G::rf(m, k, 0, i); ry[k] = G::rf(m, k, 1, i);
  }
  const T kh = (T)a.prm.venkat_K * G::hmin(m, i);
  T gx[4] = {0, 0, 0, 0}, gy[4] = {0, 0, 0, 0}, qmx[4], qmn[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) { qmx[v] = qi[v]; qmn[v] = qi[v]; }
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    if (k < nf) {
      if (jn[k] < 0) {
        const int64_t s = k * nc + i;
        ghost<T>(a.prm, jn[k] + 1, qi, G::nx(m, s), G::ny(m, s), qk[k]);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const T d = qk[k][v] - qi[v];
        gx[v] += wx[k] * d;
        gy[v] += wy[k] * d;
        qmx[v] = fmax(qmx[v], qk[k][v]);
        qmn[v] = fmin(qmn[v], qk[k][v]);
      }
    }
  }
  const T kh3 = kh * kh * kh;
  T ph[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const T e2 = kh3 * (T)(a.prm.venkat_scale[v] * a.prm.venkat_scale[v]);
    T phi = 1;
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      if (k < nf) {
        const T d2 = gx[v] * rx[k] + gy[v] * ry[k];
        if (d2 != 0) {
          const T d1 = d2 > 0 ? qmx[v] - qi[v] : qmn[v] - qi[v];
          const T psi = ((d1 * d1 + e2) * d2 + 2 * d2 * d2 * d1) / (d1 * d1 + 2 * d2 * d2 + d1 * d2 + e2) / d2;
          phi = fmin(phi, psi);
        }
      }
    }
    ph[v] = fmax(phi, (T)0);
  }
  r[0] = qi[0]; r[1] = qi[1]; r[2] = qi[2]; r[3] = qi[3];
  r[4] = gx[0]; r[5] = gy[0]; r[6] = gx[1]; r[7] = gy[1];
  r[8] = gx[2]; r[9] = gy[2]; r[10] = gx[3]; r[11] = gy[3];
  r[12] = ph[0]; r[13] = ph[1]; r[14] = ph[2]; r[15] = ph[3];
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// The record array is stored non-temporally when it is larger than this: the
// face pass reads it right after, and while it fits in the 256 MB Infinity
// Cache beside the step's other arrays it is read from there (2.08M cells,
// FP32: 133 MB, cached stores are faster); larger, its dirty lines are
// written back while the face pass runs and compete with its reads (FP64:
// 266 MB, streamed stores take the face pass from 123.9 to 106.8 us and the
// step from 0.2857 to 0.2753 ms).
constexpr int64_t kRecStreamBytes = int64_t(192) << 20;

// blocks 0 and 1 of a folding launch close the previous step's reductions
template <class T>
__device__ __forceinline__ void fold_blocks(const NsKArgs<T>& a) {
  if (blockIdx.x == 0) fold_acc_block(a.part, a.nparts, &a.acc[a.step % 3]);
  else if (a.wall_f) wall_sum_block(a.wall_f, a.m.n_wall, a.drag);
}

// ---- pass G: gradients + limiter
// STREAM: the records are stored non-temporally (a template parameter: a
// run-time choice between two stores of the same value to the same address
// is merged by the compiler into one plain store)
template <class T, bool QUADS, bool STREAM>
__global__ __launch_bounds__(kBlock) void k_ns_grad(NsKArgs<T> a) {
  const MeshView& m = a.m;
  const int64_t nc = m.nc;
  const unsigned nx = a.fold_prev ? 2u : 0u;
  if (blockIdx.x < nx) {
    fold_blocks(a);
    return;
  }
  const int64_t base = (int64_t)xcd_block(blockIdx.x - nx, gridDim.x - nx) * kBlock;
  const int64_t i = base + threadIdx.x;
  // the block's records are staged in LDS and written out as one contiguous
  // range, 16 B per lane and 4 KB per store instruction, so that every
  // instruction writes whole lines (a 64 B / 128 B record per lane would
  // leave every store scattered over 64 partial lines)
  __shared__ __align__(16) T sRec[kBlock * 16];
  if (i < nc) {
    T r[16];
    cell_record<T, QUADS>(a, i, r);
    T* o = sRec + 16 * threadIdx.x;
    st4(o, r[0], r[1], r[2], r[3]);
    st4(o + 4, r[4], r[5], r[6], r[7]);
    st4(o + 8, r[8], r[9], r[10], r[11]);
    st4(o + 12, r[12], r[13], r[14], r[15]);
  }
  __syncthreads();
  const int64_t nvalid = nc - base < kBlock ? nc - base : kBlock;
  const u32x4* src = reinterpret_cast<const u32x4*>(sRec);
  u32x4* dst = reinterpret_cast<u32x4*>(a.rec + 16 * base);
  for (int64_t p = threadIdx.x; p < nvalid * (int64_t)sizeof(T); p += kBlock) {
    if constexpr (STREAM) __builtin_nontemporal_store(src[p], dst + p);
    else dst[p] = src[p];
  }
}


// viscous flux per unit length through a face with inward normal n, from the
// face gradients of u, v, T (gf[var][x|y]) and face values u, v, T:
// (0, tau.n, (tau.u).n + kappa grad T.n)
template <class T>
__device__ __forceinline__ void viscous_face(const gd_params& P, T cp, T uf, T vf, T Tf, const T (*gf)[2], T nx,
                                             T ny, T* gv) {
  const T mu = sutherland<T>(Tf, P);
  const T kap = mu * cp / (T)P.Pr;
  const T div = gf[0][0] + gf[1][1];
  const T txx = mu * (2 * gf[0][0] - (T)(2.0 / 3.0) * div);
  const T tyy = mu * (2 * gf[1][1] - (T)(2.0 / 3.0) * div);
  const T txy = mu * (gf[0][1] + gf[1][0]);
  gv[0] = 0;
  gv[1] = txx * nx + txy * ny;
  gv[2] = txy * nx + tyy * ny;
  gv[3] = uf * gv[1] + vf * gv[2] + kap * (gf[2][0] * nx + gf[2][1] * ny);
}

// LDS ordering between lanes of one wave (a wave's LDS operations complete in
// issue order; the fences keep the compiler from moving them across)
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// R records per lane of a wave (R x 64 records), gathered cooperatively: a
// record is kP = sizeof(T) pieces of 16 B, and kP adjacent lanes load one
// record's pieces in the same instruction, so an instruction touches 64 / kP
// cache lines instead of 64 (the per-line cost of scattered 16 B lane loads
// is what bounds a one-record-per-lane gather).  Every load is issued first;
// the pieces then pass through a padded per-wave LDS slab of 64 records (one
// piece of padding per record keeps the read-back free of bank conflicts),
// one round per record of the lane.
template <class T, int R>
struct RecGather {
  static constexpr int kP = (int)sizeof(T);  // 16 B pieces per 16-value record
  static constexpr int kRPI = 64 / kP;       // records per load instruction
  static constexpr int kStride = kP + 1;     // LDS pieces per record
  static constexpr int kWaves = kBlock / 64;
  struct Shared {
    u32x4 rec[kWaves][64 * kStride];
    int32_t cell[kWaves][R * 64];
  };
  __device__ __forceinline__ static void unpack(u32x4 x, T* o) {
    if constexpr (sizeof(T) == 8) {
      typedef double f64x2 __attribute__((ext_vector_type(2)));
      const f64x2 d = __builtin_bit_cast(f64x2, x);
      o[0] = d.x; o[1] = d.y;
    } else {
      typedef float f32x4 __attribute__((ext_vector_type(4)));
      const f32x4 d = __builtin_bit_cast(f32x4, x);
      o[0] = d.x; o[1] = d.y; o[2] = d.z; o[3] = d.w;
    }
  }
  // cells[r]: the cell whose record this lane receives; fn(r, record) is
  // called for r = 0 .. R-1 as soon as record r has arrived, so that the
  // caller can reduce it before the next one is unpacked.  Every lane of the
  // wave must take part.
  template <class F>
  __device__ __forceinline__ static void gather_each(Shared& sh, const T* rec, const int32_t (&cells)[R], F&& fn) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < R; ++r) sh.cell[w][r * 64 + lane] = cells[r];
    wave_lds_sync();
    const u32x4* G = reinterpret_cast<const u32x4*>(rec);
    const int sub = lane / kP, piece = lane % kP;
    u32x4 v[R][kP];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int k = 0; k < kP; ++k) v[r][k] = G[(int64_t)sh.cell[w][r * 64 + k * kRPI + sub] * kP + piece];
    u32x4* S = sh.rec[w];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (r > 0) wave_lds_sync();
#pragma unroll
      for (int k = 0; k < kP; ++k) S[(k * kRPI + sub) * kStride + piece] = v[r][k];
      wave_lds_sync();
      T out[16];
#pragma unroll
      for (int p = 0; p < kP; ++p) unpack(S[lane * kStride + p], out + p * (16 / kP));
      fn(r, out);
    }
  }
};

// One side of a face, reduced from that side's record as soon as it arrives:
// the limited reconstruction at the face (offset ox, oy from the centroid)
// and, for the viscous flux, the cell's u, v, T and their gradients
// (grad T = (grad p - T R grad rho) / (rho R)).
template <class T, bool VISC>
struct FaceSide {
  T q[4];
  T u, v, t;
  T gr[3][2];
};

template <class T, bool VISC>
__device__ __forceinline__ void face_side(const T (&r)[16], T ox, T oy, T Rg, FaceSide<T, VISC>& s) {
  const T* gq = r + 4;
#pragma unroll
  for (int v = 0; v < 4; ++v) s.q[v] = r[v] + r[12 + v] * (gq[2 * v] * ox + gq[2 * v + 1] * oy);
  if (!(s.q[0] > 0) || !(s.q[3] > 0)) {
#pragma unroll
    for (int v = 0; v < 4; ++v) s.q[v] = r[v];
  }
  if (VISC) {
    const T rR = 1 / (r[0] * Rg);
    const T t = r[3] * rR;
    s.u = r[1]; s.v = r[2]; s.t = t;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      s.gr[0][c] = gq[2 + c];
      s.gr[1][c] = gq[4 + c];
      s.gr[2][c] = (gq[6 + c] - t * Rg * gq[c]) * rR;
    }
  }
}

// A face's geometry: the owner's inward normal, the length, and the face
// midpoint minus the owner's (rx, ry) and the neighbour's (rxj, ryj) centroid.
template <class T>
struct FaceGeoVals {
  T nx, ny, len, rx, ry, rxj, ryj;
};

// The flux of face f from its owner's side s0 and neighbour's side s1 (c1 =
// the neighbour, or code - 1 at a boundary, where s1's state is the ghost of
// s0's): the viscous flux, AUSM+ or Roe, returned in ofl; the wall force is
// stored here.
template <class T, bool VISC>
__device__ __forceinline__ void face_flux(const NsKArgs<T>& a, int64_t f, int32_t c1, const FaceGeoVals<T>& fg,
                                          const FaceSide<T, VISC>& s0, FaceSide<T, VISC>& s1, T* ofl) {
  const MeshView& m = a.m;
  const gd_params& P = a.prm;
  const T nx = fg.nx, ny = fg.ny, len = fg.len;
  const T g = (T)P.Gamma, Rg = (T)P.RGAS;
  if (c1 < 0) ghost<T>(P, c1 + 1, s0.q, nx, ny, s1.q);
  T gv[4] = {0, 0, 0, 0};
  if (VISC) {
    const T cp = g * Rg / (g - 1);
    const T rx = fg.rx, ry = fg.ry;
    if (c1 >= 0) {
      // edge-corrected average of the two cell gradients of u, v, T
      const T ddx = rx - fg.rxj, ddy = ry - fg.ryj;  // centroid j - centroid i
      const T rdl = 1 / sqrt(ddx * ddx + ddy * ddy);
      const T tx = ddx * rdl, ty = ddy * rdl;
      const T dphi[3] = {s1.u - s0.u, s1.v - s0.v, s1.t - s0.t};
      T gf[3][2];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T ax = (T)0.5 * (s0.gr[c][0] + s1.gr[c][0]), ay = (T)0.5 * (s0.gr[c][1] + s1.gr[c][1]);
        const T corr = dphi[c] * rdl - (ax * tx + ay * ty);
        gf[c][0] = ax + corr * tx; gf[c][1] = ay + corr * ty;
      }
      viscous_face<T>(P, cp, (T)0.5 * (s0.u + s1.u), (T)0.5 * (s0.v + s1.v), (T)0.5 * (s0.t + s1.t), gf, nx, ny,
                      gv);
    } else {
      const int idx = -(c1 + 1);
      if (idx <= GD_MAX_BC && P.bc[idx].kind == GD_BC_ISOTHERMAL_WALL) {
        // one-sided normal gradient to the wall (u = 0, T = T_w); the fluid's
        // outward normal is -n
        const T Tw = (T)P.bc[idx].data[0];
        const T rdp = -1 / (rx * nx + ry * ny);
        const T dudn = (0 - s0.u) * rdp, dvdn = (0 - s0.v) * rdp, dTdn = (Tw - s0.t) * rdp;
        const T ox = -nx, oy = -ny;
        const T gf[3][2] = {{dudn * ox, dudn * oy}, {dvdn * ox, dvdn * oy}, {dTdn * ox, dTdn * oy}};
        viscous_face<T>(P, cp, (T)0, (T)0, Tw, gf, nx, ny, gv);
      }
    }
  }
  T e[4];
  if (P.flux == GD_FLUX_ROE) roe<T>(s1.q, s0.q, nx, ny, g, e);
  else ausm<T>(s1.q, s0.q, nx, ny, g, e);
  double fk[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) fk[v] = (double)((e[v] + gv[v]) * len);
  ofl[0] = (T)fk[0]; ofl[1] = (T)fk[1]; ofl[2] = (T)fk[2]; ofl[3] = (T)fk[3];
  if (c1 < 0 && a.wall_f && a.wall_code == c1 + 1) {
    const int32_t wid = m.fwall[f];
    if (wid >= 0) { a.wall_f[2 * wid] = fk[1]; a.wall_f[2 * wid + 1] = fk[2]; }
  }
}

// ---- pass F1: one flux per face
template <class T, bool VISC>
__global__ __launch_bounds__(kBlock) void k_ns_face(NsKArgs<T> a) {
  using G = FaceGeo<T>;
  const MeshView& m = a.m;
  const int64_t nF = m.n_faces;
  const int64_t f0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  // lanes past the end gather the last face's records and store nothing, so
  // that every lane of a wave takes part in the cooperative gather and in the
  // staging of the fluxes
  const bool live = f0 < nF;
  const int64_t f = live ? f0 : nF - 1;
  const int32_t c0 = ntl(&m.fc0[f]), c1 = ntl(&m.fc1[f]);
  const T Rg = (T)a.prm.RGAS;
  // the face geometry is loaded first, so that it arrives during the gather
  // (the f_r1 rows span every face; a boundary face's are not used)
  const FaceGeoVals<T> fg = {G::fnx(m, f),    G::fny(m, f),    G::flen(m, f),  G::fr0(m, 0, f),
                             G::fr0(m, 1, f), G::fr1(m, 0, f), G::fr1(m, 1, f)};
  // both records (a boundary face reads the owner's twice and uses it once),
  // each reduced to its side of the face as soon as it has arrived
  __shared__ typename RecGather<T, 2>::Shared sh;
  FaceSide<T, VISC> s[2];
  RecGather<T, 2>::gather_each(sh, a.rec, {c0, c1 >= 0 ? c1 : c0}, [&](int r, const T (&rec)[16]) {
    if (r == 0) face_side<T, VISC>(rec, fg.rx, fg.ry, Rg, s[0]);
    else if (c1 >= 0) face_side<T, VISC>(rec, fg.rxj, fg.ryj, Rg, s[1]);
  });
  T ofl[4] = {0, 0, 0, 0};
  if (live) face_flux<T, VISC>(a, f, c1, fg, s[0], s[1], ofl);
  if constexpr (sizeof(T) == 4) {
    if (live) st4(a.fflux + 4 * f, ofl[0], ofl[1], ofl[2], ofl[3]);
  } else {
    // FP64: a face's flux is two 16 B pieces, which the wave passes through its
    // LDS slab so that each store instruction writes 1 KB of whole lines, and
    // stores non-temporally: the update reads each flux twice within its own
    // pass, and the lines left dirty in the Infinity Cache would otherwise be
    // written back during the next step's passes (face pass 106.8 -> 101.8 us,
    // update 71.1 -> 74.3 us, step 0.2753 -> 0.2738 ms)
    typedef double f64x2 __attribute__((ext_vector_type(2)));
    const int lane = threadIdx.x & 63;
    u32x4* S = sh.rec[threadIdx.x >> 6];
    wave_lds_sync();
    S[2 * lane] = __builtin_bit_cast(u32x4, f64x2{(double)ofl[0], (double)ofl[1]});
    S[2 * lane + 1] = __builtin_bit_cast(u32x4, f64x2{(double)ofl[2], (double)ofl[3]});
    wave_lds_sync();
    const int64_t wf0 = f0 - lane;  // the wave's first face
    u32x4* dst = reinterpret_cast<u32x4*>(a.fflux) + 2 * wf0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int idx = k * 64 + lane;
      if (wf0 + idx / 2 < nF) __builtin_nontemporal_store(S[idx], dst + idx);
    }
  }
}

// ---- pass F2: per-cell flux sum, update, primitives, dt inputs
template <class T, bool QUADS, bool VISC>
__global__ __launch_bounds__(kBlock) void k_ns_update(NsKArgs<T> a) {
  constexpr int NF = QUADS ? 4 : 3;
  const MeshView& m = a.m;
  const gd_params& P = a.prm;
  const int64_t nc = m.nc;
  const int64_t i = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  const Acc r = a.acc[a.step % 3];
  double dt = a.dt;
  if (a.dt_policy == GD_DT_LOCAL) dt = a.cfl * __longlong_as_double((long long)r.hdt_bits);
  else if (a.dt_policy == GD_DT_REF_TAU) {
    const double vm = __longlong_as_double((long long)r.vmax_bits);
    dt = a.cfl * m.max_hx / (vm > 1.0e-6 ? vm : 1.0e-6);
  }
  double vmx = 0.0, hdt = __longlong_as_double(0x7ff0000000000000ll);
  unsigned bad = 0;
  double qw[4] = {0.0, 0.0, 0.0, 0.0};
  if (i < nc) {
    const int nf = QUADS ? m.nv[i] : 3;
    int32_t fc[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) fc[k] = ntl(&m.cface[k * nc + i]);
    double Uo[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) Uo[v] = ntl(&a.Uo[v * nc + i]);
    double F[4] = {0.0, 0.0, 0.0, 0.0};
    T fl[NF][4];
#pragma unroll
    for (int k = 0; k < NF; ++k)
      if (k < nf) ld4(a.fflux + 4 * (int64_t)(fc[k] >> 1), fl[k]);
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      if (k < nf) {
        const bool neg = (fc[k] & 1) != 0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const double x = (double)fl[k][v];
          F[v] += neg ? -x : x;
        }
      }
    }
    const double ra = ntl(&m.rarea[i]);
    // read before the stores below, which the compiler may not move a load past
    const double h = (double)FaceGeo<T>::hmin(m, i);
    double Un[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      Un[v] = Uo[v] - F[v] * dt * ra;
      // U' is read again only by the next step's update, after every other
      // array of the step: a non-temporal store keeps it out of the caches
      // the next gradient and face passes use (FP64 face pass 125.0 -> 121.4 us)
      __builtin_nontemporal_store(Un[v], &a.Un[v * nc + i]);
    }
    const double Gm = P.Gamma;
    const double rho = Un[0], ux = Un[1] / rho, uy = Un[2] / rho;
    const double p = (Gm - 1.0) * (Un[3] - 0.5 * (Un[1] * Un[1] + Un[2] * Un[2]) / rho);
    if constexpr (sizeof(T) == 4) st4(a.Qw + 4 * i, (T)rho, (T)ux, (T)uy, (T)p);
    else { qw[0] = rho; qw[1] = ux; qw[2] = uy; qw[3] = p; }
    const double as = sqrt(Gm * p / rho);
    const double vm = sqrt(ux * ux + uy * uy) + as;
    double lam = vm;
    if (VISC) {
      const double Tc = p / (rho * P.RGAS);
      const double t = Tc / P.T_ref;
      const double mu = P.mu_ref * (t * sqrt(t)) * (P.T_ref + P.S_suth) / (Tc + P.S_suth);
      lam = vm + 2.0 * Gm * mu / (P.Pr * rho * h);
    }
    vmx = vm;
    hdt = h / lam;
    bad = (!(rho > 0.0) || !(p > 0.0) || vm != vm) ? 1u : 0u;
    if (vm != vm || lam != lam) { vmx = 0.0; hdt = __longlong_as_double(0x7ff0000000000000ll); }
  }
  // block reduction of the next step's dt inputs into per-block partials
  // FP64: the block's next primitives (32 B per cell) pass through LDS so
  // that each store instruction writes 4 KB of whole lines rather than half
  // of 64 lines (update 74.1 -> 70.8 us)
  if constexpr (sizeof(T) == 8) {
    __shared__ __align__(16) double sQ[kBlock * 4];
    sQ[4 * threadIdx.x] = qw[0]; sQ[4 * threadIdx.x + 1] = qw[1];
    sQ[4 * threadIdx.x + 2] = qw[2]; sQ[4 * threadIdx.x + 3] = qw[3];
    __syncthreads();
    const int64_t base = i - threadIdx.x;
    const u32x4* src = reinterpret_cast<const u32x4*>(sQ);
    u32x4* dst = reinterpret_cast<u32x4*>(a.Qw + 4 * base);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int p = threadIdx.x + k * kBlock;
      if (base + p / 2 < nc) dst[p] = src[p];
    }
  }
  __shared__ double s_max[kBlock / 64], s_min[kBlock / 64];
  __shared__ unsigned s_bad[kBlock / 64];
  vmx = wmax64(vmx);
  hdt = wmin64(hdt);
  bad = wsum_u32(bad);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { s_max[wid] = vmx; s_min[wid] = hdt; s_bad[wid] = bad; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < kBlock / 64; ++q) { vmx = fmax(vmx, s_max[q]); hdt = fmin(hdt, s_min[q]); bad += s_bad[q]; }
    a.part[3 * (int64_t)blockIdx.x] = vmx;
    a.part[3 * (int64_t)blockIdx.x + 1] = hdt;
    a.part[3 * (int64_t)blockIdx.x + 2] = (double)bad;
    if (blockIdx.x == 0) {
      Acc* w = &a.acc[(a.step + 1) % 3];
      w->time = r.time + dt;
      w->dt = dt;
    }
  }
}

inline unsigned nblocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

template <class T>
hipError_t launch_grad(const NsKArgs<T>& a, hipStream_t s) {
  const dim3 g(nblocks(a.m.nc) + (a.fold_prev ? 2 : 0));
  const bool stream = (int64_t)16 * (int64_t)sizeof(T) * a.m.nc > kRecStreamBytes;
  if (a.m.has_quads) {
    if (stream) hipLaunchKernelGGL((k_ns_grad<T, true, true>), g, dim3(kBlock), 0, s, a);
    else hipLaunchKernelGGL((k_ns_grad<T, true, false>), g, dim3(kBlock), 0, s, a);
  } else {
    if (stream) hipLaunchKernelGGL((k_ns_grad<T, false, true>), g, dim3(kBlock), 0, s, a);
    else hipLaunchKernelGGL((k_ns_grad<T, false, false>), g, dim3(kBlock), 0, s, a);
  }
  return hipGetLastError();
}

template <class T>
hipError_t launch_face(const NsKArgs<T>& a, hipStream_t s) {
  const dim3 g(nblocks(a.m.n_faces)), b(kBlock);
  if (a.prm.viscous) hipLaunchKernelGGL((k_ns_face<T, true>), g, b, 0, s, a);
  else hipLaunchKernelGGL((k_ns_face<T, false>), g, b, 0, s, a);
  return hipGetLastError();
}

template <class T>
hipError_t launch_update(const NsKArgs<T>& a, hipStream_t s) {
  const dim3 g(nblocks(a.m.nc)), b(kBlock);
  const bool visc = a.prm.viscous != 0;
  if (a.m.has_quads) {
    if (visc) hipLaunchKernelGGL((k_ns_update<T, true, true>), g, b, 0, s, a);
    else hipLaunchKernelGGL((k_ns_update<T, true, false>), g, b, 0, s, a);
  } else {
    if (visc) hipLaunchKernelGGL((k_ns_update<T, false, true>), g, b, 0, s, a);
    else hipLaunchKernelGGL((k_ns_update<T, false, false>), g, b, 0, s, a);
  }
  return hipGetLastError();
}

template <class T>
hipError_t launch_prims(int64_t nc, double g, const double* U, T* Q, hipStream_t s) {
  hipLaunchKernelGGL((k_ns_prims<T>), dim3(nblocks(nc)), dim3(kBlock), 0, s, nc, g, U, Q);
  return hipGetLastError();
}

}  // namespace ns
}  // namespace gd
