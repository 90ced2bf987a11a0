// Reference-scheme kernels (FP64), compiled with -ffp-contract=off.
//
// Cell-centric and deterministic: one work-item per cell gathers its
// neighbours' states and sums its own face fluxes in face order, so there are
// no atomics on the state and results are bit-reproducible run to run and
// independent of the cell numbering.  Per step of the device-resident loop:
//   k_ref_step    flux sum + explicit Euler update + primitive recovery
//                 (+ the calcDiffTerm Laplacian from node means of gradients)
//   k_ref_nodes   cell -> node weighted means (next step's MUSCL stencil)
//                 and unweighted cluster means of rho*u, rho*v, E
//   k_ref_gg1     Green-Gauss gradients of rho*u, rho*v, E   (viscous only)
//   k_ref_gg_nodes  cluster means of those gradients         (viscous only)
#include "gd_fold.hpp"
#include "gd_internal.hpp"
#include "gd_ref_math.hpp"

namespace gd {
namespace {

constexpr int kBlock = kThreads;


// Block-reduce the step's dt inputs and non-physical count and store the
// block's partial (max |u|+a, min local dt, count) in part[3 b .. 3 b + 2];
// k_acc_fold combines the partials in one block.  (Same-address device-scope
// atomics from every block serialise across the XCDs.)
__device__ __forceinline__ void reduce_acc(double vmax, double hdt, unsigned bad, double* part) {
  __shared__ double s_max[kBlock / 64], s_min[kBlock / 64];
  __shared__ unsigned s_bad[kBlock / 64];
  vmax = wmax64(vmax);
  hdt = wmin64(hdt);
  bad = wsum_u32(bad);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { s_max[wid] = vmax; s_min[wid] = hdt; s_bad[wid] = bad; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < kBlock / 64; ++q) {
      vmax = fmax(vmax, s_max[q]); hdt = fmin(hdt, s_min[q]); bad += s_bad[q];
    }
    part[3 * (int64_t)blockIdx.x] = vmax;
    part[3 * (int64_t)blockIdx.x + 1] = hdt;
    part[3 * (int64_t)blockIdx.x + 2] = (double)bad;
  }
}

__device__ __forceinline__ double step_dt(const Acc& r, int policy, double dt_fixed, double cfl, double max_hx) {
  if (policy == GD_DT_REF_TAU) return cfl * max_hx / refm::jmax(__longlong_as_double((long long)r.vmax_bits), 1.0e-6);
  if (policy == GD_DT_LOCAL) return cfl * __longlong_as_double((long long)r.hdt_bits);
  return dt_fixed;
}

__device__ __forceinline__ void rotate_acc(Acc* acc, int step, double t, double dt) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    Acc* w = &acc[(step + 1) % 3];
    w->time = t + dt;
    w->dt = dt;
  }
}

// 2K consecutive doubles as K 16 B loads / stores (p 16 B aligned)
template <int K>
__device__ __forceinline__ void ld2k(const double* p, double* o) {
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const double2 v = *reinterpret_cast<const double2*>(p + 2 * q);
    o[2 * q] = v.x; o[2 * q + 1] = v.y;
  }
}

template <int K>
__device__ __forceinline__ void st2k(double* p, const double* o) {
#pragma unroll
  for (int q = 0; q < K; ++q) *reinterpret_cast<double2*>(p + 2 * q) = make_double2(o[2 * q], o[2 * q + 1]);
}

// GG face sums of the first W fields of node records of S doubles (S even):
// per field, sum_k ((0.5 (phi_a + phi_b)) (-n_k)) len_k in face order,
// calcGradients.jl:116-129 (the x/y swap of the source is applied by the
// caller).  A node's fields are one record, so each face end is S/2 16 B
// loads within one or two cache lines rather than one scattered load per
// field.  The edge ids and geometry are read once
// by each of the two Green-Gauss levels of a step, too far apart to meet in a
// cache: non-temporal (k_ref_gg1 72.0 -> 69.2 us, k_ref_step 123.6 -> 113.4 us).
template <bool QUADS, int S, int W>
__device__ __forceinline__ void gg_sums(const MeshView& m, int64_t i, int nf, const double* rec, double* ax,
                                        double* ay) {
  const int64_t nc = m.nc;
#pragma unroll
  for (int k = 0; k < (QUADS ? 4 : 3); ++k) {
    if (k < nf) {
      double pa[S], pb[S];
      ld2k<S / 2>(rec + (int64_t)S * __builtin_nontemporal_load(&m.ea[k * nc + i]), pa);
      ld2k<S / 2>(rec + (int64_t)S * __builtin_nontemporal_load(&m.eb[k * nc + i]), pb);
      const double mny = -__builtin_nontemporal_load(&m.fny[k * nc + i]),
                   mnx = -__builtin_nontemporal_load(&m.fnx[k * nc + i]),
                   len = __builtin_nontemporal_load(&m.flen[k * nc + i]);
#pragma unroll
      for (int f = 0; f < W; ++f) {
        const double hm = 0.5 * (pa[f] + pb[f]);
        const double fx = hm * mny * len;
        const double fy = hm * mnx * len;
        if (k == 0) { ax[f] = fx; ay[f] = fy; } else { ax[f] = ax[f] + fx; ay[f] = ay[f] + fy; }
      }
    }
  }
}

// calcDiffTerm's per-cell tail (calcDiffTerm.jl:98-119): Laplacians from the
// second Green-Gauss level, mu = mu_const, energy row scaled by k(T).
template <bool QUADS>
__device__ __forceinline__ void ref_udiff(const MeshView& m, const gd_params& p, const double* g6N, int64_t i, int nf,
                                          const double* q, double* ud) {
  const double A = __builtin_nontemporal_load(&m.area[i]);
  double ax[6], ay[6];
  gg_sums<QUADS, 6, 6>(m, i, nf, g6N, ax, ay);
  double lap[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) lap[c] = ax[2 * c] / A + ay[2 * c + 1] / A;  // "gradX" of "gradX" + "gradY" of "gradY"
  const double T = q[3] / q[0] / p.RGAS;
  const double mu = p.mu_const;
  const double k = p.energy_law == GD_ENERGY_CONST ? p.k_const : refm::air_k(T);
  ud[1] = mu * lap[0];
  ud[2] = mu * lap[1];
  ud[3] = mu * k * lap[2];
}

// The two reconstructed states of face k of cell i, computeInterfaceSlope
// (computeslope2d.jl:44-126): fL = UpLeft (own side), fR = UpRight (the
// neighbour's side, or the boundary ghost's with a first-order slope).
// Returns the neighbour (>= 0) or code - 1.
template <bool QUADS>
__device__ __forceinline__ int32_t ref_face_states(const RefStepArgs& a, int64_t i, int k, int nf, const double* qL,
                                                   double t, double* fL, double* fR) {
  const MeshView& m = a.m;
  const int64_t nc = m.nc;
  const int64_t s = k * nc + i;
  const int32_t j = m.nb[s];
  double up[4], dn[4], qR[4];
  if (j >= 0) {
    const int32_t d1 = m.dn1[s], u1 = m.up1[s];
    if (QUADS && nf == 4) {
      const int32_t d2 = m.dn2[s], u2 = m.up2[s];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        up[v] = 0.5 * (a.Qn[v * m.nn + u1] + a.Qn[v * m.nn + u2]);
        dn[v] = 0.5 * (a.Qn[v * m.nn + d1] + a.Qn[v * m.nn + d2]);
      }
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v) { up[v] = a.Qn[v * m.nn + u1]; dn[v] = a.Qn[v * m.nn + d1]; }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) qR[v] = a.Q[v * nc + j];
  } else {
    const int32_t na = m.ea[s], nbn = m.eb[s];
    refm::bc_ghost(a.prm.bc, j + 1, qL, m.fnx[s], m.fny[s], m.xc[i], m.yc[i], m.xn[na], m.yn[na], m.xn[nbn],
                   m.yn[nbn], t, a.prm.RGAS, qR);
#pragma unroll
    for (int v = 0; v < 4; ++v) { dn[v] = qL[v]; up[v] = qR[v]; }
  }
  const double eps = a.prm.limiter_eps;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    fL[v] = qL[v] + 0.5 * refm::limiter(a.prm.limiter, qL[v] - dn[v], qR[v] - qL[v], eps);
    fR[v] = qR[v] - 0.5 * refm::limiter(a.prm.limiter, qR[v] - qL[v], up[v] - qR[v], eps);
  }
  return j;
}

// One cell of computeInterfaceSlope (computeslope2d.jl:3-136) summed over its
// faces as in SecondOrderUpwindM2 (SOUscheme.jl:56-92).
template <bool QUADS>
__device__ __forceinline__ void ref_cell_flux(const RefStepArgs& a, int64_t i, int nf, const double* qL, double t,
                                              double* F, double* wallF) {
  const MeshView& m = a.m;
  const int64_t nc = m.nc;
  const double g = a.prm.Gamma;
#pragma unroll
  for (int k = 0; k < (QUADS ? 4 : 3); ++k) {
    if (k < nf) {
      const int64_t s = k * nc + i;
      double e[4], fL[4], fR[4];
      const int32_t j = ref_face_states<QUADS>(a, i, k, nf, qL, t, fL, fR);
      refm::face_flux(a.prm.flux, fR, fL, m.fnx[s], m.fny[s], m.flen[s], g, e);  // neighbour side is "L"
      if (j < 0 && wallF && m.wall_id && a.wall_code == j + 1) {
        const int32_t wid = m.wall_id[s];
        if (wid >= 0) { wallF[2 * wid] = e[1]; wallF[2 * wid + 1] = e[2]; }
      }
      if (k == 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) F[v] = e[v];
      } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) F[v] = F[v] + e[v];
      }
    }
  }
}

// ref_cell_flux from the face pass's fluxes (MeshView::rcf covers every face
// slot of every cell): a mirrored face's flux evaluated once from its owner's
// side enters the other cell negated — that side's own evaluation is its exact
// negative (mirrored stencils, odd and symmetric limiter, AUSM+ antisymmetric
// under L <-> R, n -> -n) — and the others as evaluated from this cell's side.
// All slots' fluxes are loaded before the sum, which keeps the slot order.
template <bool QUADS>
__device__ __forceinline__ void ref_cell_flux_gathered(const RefStepArgs& a, int64_t i, int nf, double* F) {
  constexpr int NS = QUADS ? 4 : 3;
  const int64_t nc = a.m.nc;
  int32_t cf[NS];
#pragma unroll
  // the slot ids, U, 1/A and h are read once per step, U' is read again only
  // by the next step's update: non-temporal (k_ref_step 126.6 -> 122.5 us)
  for (int k = 0; k < NS; ++k) cf[k] = (k < nf) ? __builtin_nontemporal_load(&a.m.rcf[k * nc + i]) : 0;
  double x[NS][4];
#pragma unroll
  for (int k = 0; k < NS; ++k) ld2k<2>(a.rflux + 4 * (int64_t)(cf[k] >> 1), x[k]);
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    if (k < nf) {
      double e[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) e[v] = (cf[k] & 1) ? -x[k][v] : x[k][v];
      if (k == 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) F[v] = e[v];
      } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) F[v] = F[v] + e[v];
      }
    }
  }
}

// The face pass of the reference loop (the accumulator slot of the step gives
// its flow time, as in k_ref_step).  Entries below n_rmirror: the flux of an
// interior face between two triangles with mirrored stencils from its owner's
// side, from face-ordered copies of the owner slot's data; the rest: the flux
// of one cell's face slot from that cell's side, with its boundary ghost and
// wall force.  Both are computed exactly as ref_cell_flux computes them there.
template <bool QUADS>
__global__ __launch_bounds__(kBlock) void k_ref_faces(RefStepArgs a, double* rflux) {
  const MeshView& m = a.m;
  const int64_t f0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  const bool live = f0 < m.n_rfaces;
  const int64_t f = live ? f0 : m.n_rfaces - 1;
  const int64_t nc = m.nc, nn = m.nn;
  // the face-ordered ids and geometry are read once per step: non-temporal,
  // leaving the caches to the gathered cell and node primitives
  const int64_t i = __builtin_nontemporal_load(&m.rf_cell[f]);
  double e[4];
  if (f < m.n_rmirror) {
    const int32_t j = __builtin_nontemporal_load(&m.rf_nb[f]), d1 = __builtin_nontemporal_load(&m.rf_dn1[f]),
                  u1 = __builtin_nontemporal_load(&m.rf_up1[f]);
    const double nx = __builtin_nontemporal_load(&m.rf_nx[f]), ny = __builtin_nontemporal_load(&m.rf_ny[f]),
                 len = __builtin_nontemporal_load(&m.rf_len[f]);
    double qL[4], qR[4], up[4], dn[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      qL[v] = a.Q[v * nc + i];
      qR[v] = a.Q[v * nc + j];
      up[v] = a.Qn[v * nn + u1];
      dn[v] = a.Qn[v * nn + d1];
    }
    const double eps = a.prm.limiter_eps;
    double fL[4], fR[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {  // ref_face_states on an interior face
      fL[v] = qL[v] + 0.5 * refm::limiter(a.prm.limiter, qL[v] - dn[v], qR[v] - qL[v], eps);
      fR[v] = qR[v] - 0.5 * refm::limiter(a.prm.limiter, qR[v] - qL[v], up[v] - qR[v], eps);
    }
    refm::face_flux(a.prm.flux, fR, fL, nx, ny, len, a.prm.Gamma, e);
  } else {
    const int k = m.rf_slot[f];
    const double t = a.acc[a.step % 3].time;
    const double qL[4] = {a.Q[i], a.Q[nc + i], a.Q[2 * nc + i], a.Q[3 * nc + i]};
    const int nf = QUADS ? m.nv[i] : 3;
    double fL[4], fR[4];
    const int32_t j = ref_face_states<QUADS>(a, i, k, nf, qL, t, fL, fR);
    const int64_t s = (int64_t)k * nc + i;
    refm::face_flux(a.prm.flux, fR, fL, m.fnx[s], m.fny[s], m.flen[s], a.prm.Gamma, e);  // neighbour side is "L"
    if (live && j < 0 && a.wall_f && m.wall_id && a.wall_code == j + 1) {
      const int32_t wid = m.wall_id[s];
      if (wid >= 0) { a.wall_f[2 * wid] = e[1]; a.wall_f[2 * wid + 1] = e[2]; }
    }
  }
  __shared__ WaveStage<2> stg;
  stg.store(rflux, f0, m.n_rfaces, e);
}

// The reference's GPU stages, calcOneStageCUDA (calcOneStageCUDA.jl:4-135)
// and calcOneStageCUDAx32 (calcOneStageCUDAx32.jl:5-155): kernel_AUSM2d of
// AUSMflux2dCUDA.jl:56-98 / AUSMflux2dCUDAx32.jl:56-98 on face slots 1..3
// (the 4th slot's launch is commented out in the source), accumulated per
// cell in the kernel's type, then the Float64 update without Udiff.  The
// x32 kernel's arrays are Float32, but the literals of its split functions
// (AUSMflux2dCUDAx32.jl:19-48) are Float64, so Julia evaluates M + 1.0 and
// the products with them in Float64 (M*M and AUSM_ALFA*M stay Float32) and
// each function returns Float32.  Built without contraction, like the source
// expressions.
struct CudaSplitF64 {
  __device__ static double M_p(double M, double B) {
    return fabs(M) >= 1.0 ? 0.5 * (M + fabs(M)) : 0.25 * (M + 1.0) * (M + 1.0) + B * (M * M - 1.0) * (M * M - 1.0);
  }
  __device__ static double M_m(double M, double B) {
    return fabs(M) >= 1.0 ? 0.5 * (M - fabs(M)) : -0.25 * (M - 1.0) * (M - 1.0) - B * (M * M - 1.0) * (M * M - 1.0);
  }
  __device__ static double P_p(double M, double A) {
    return fabs(M) >= 1.0 ? 0.5 * (1.0 + refm::jsign(M))
                          : 0.25 * (M + 1.0) * (M + 1.0) * (2.0 - M) + A * M * (M * M - 1.0) * (M * M - 1.0);
  }
  __device__ static double P_m(double M, double A) {
    return fabs(M) >= 1.0 ? 0.5 * (1.0 - refm::jsign(M))
                          : 0.25 * (M - 1.0) * (M - 1.0) * (2.0 + M) - A * M * (M * M - 1.0) * (M * M - 1.0);
  }
};
__device__ __forceinline__ float fsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : x); }
struct CudaSplitX32 {
  __device__ static float M_p(float M, float B) {
    if (fabsf(M) >= 1.0) return (float)(0.5 * (double)(M + fabsf(M)));
    const double a = (double)M + 1.0, d = (double)(M * M) - 1.0;
    return (float)(0.25 * a * a + (double)B * d * d);
  }
  __device__ static float M_m(float M, float B) {
    if (fabsf(M) >= 1.0) return (float)(0.5 * (double)(M - fabsf(M)));
    const double a = (double)M - 1.0, d = (double)(M * M) - 1.0;
    return (float)(-0.25 * a * a - (double)B * d * d);
  }
  __device__ static float P_p(float M, float A) {
    if (fabsf(M) >= 1.0) return (float)(0.5 * (1.0 + (double)fsign(M)));
    const double a = (double)M + 1.0, d = (double)(M * M) - 1.0;
    return (float)(0.25 * a * a * (2.0 - (double)M) + (double)(A * M) * d * d);
  }
  __device__ static float P_m(float M, float A) {
    if (fabsf(M) >= 1.0) return (float)(0.5 * (1.0 - (double)fsign(M)));
    const double a = (double)M - 1.0, d = (double)(M * M) - 1.0;
    return (float)(0.25 * a * a * (2.0 + (double)M) - (double)(A * M) * d * d);
  }
};

template <class T>
__device__ __forceinline__ T tmin(T x, T y) { return y < x ? y : x; }

template <class T, class Split>
__device__ __forceinline__ void cuda_ausm(const T* L, const T* R, T nx, T ny, T side, T g, T* acc) {
  const T rL = L[0], uL = L[1], vL = L[2], pL = L[3];
  const T rR = R[0], uR = R[1], vR = R[2], pR = R[3];
  const T one = (T)1.0, two = (T)2.0, half = (T)0.5;
  const T htL = pL / rL / (g - one) + half * (uL * uL + vL * vL) + pL / rL;
  const T htR = pR / rR / (g - one) + half * (uR * uR + vR * vR) + pR / rR;
  const T aL = sqrt(two * (g - one) / (g + one) * htL) *
               tmin(one, sqrt(two * (g - one) / (g + one) * htL) / fabs(uL * nx + vL * ny));
  const T aR = sqrt(two * (g - one) / (g + one) * htR) *
               tmin(one, sqrt(two * (g - one) / (g + one) * htR) / fabs(uR * nx + vR * ny));
  const T a12 = tmin(aL, aR);
  const T ML = (uL * nx + vL * ny) / a12, MR = (uR * nx + vR * ny) / a12;
  const T B = (T)(1.0 / 8.0), A = (T)(3.0 / 16.0);
  const T m = Split::M_p(ML, B) + Split::M_m(MR, B);
  const T p12 = Split::P_p(ML, A) * pL + Split::P_m(MR, A) * pR;
  const T mp = half * (m + fabs(m)), mm = half * (m - fabs(m));
  acc[0] = acc[0] - (a12 * (mp * rL + mm * rR)) * side;
  acc[1] = acc[1] - (a12 * (mp * rL * uL + mm * rR * uR) + p12 * nx) * side;
  acc[2] = acc[2] - (a12 * (mp * rL * vL + mm * rR * vR) + p12 * ny) * side;
  acc[3] = acc[3] - (a12 * (mp * rL * htL + mm * rR * htR)) * side;
}

// UL/UR: the face states, [slot][4][nc] (slots 0..2 read); F: iFlux [4][nc]
template <class T, class Split>
__global__ __launch_bounds__(kBlock) void k_ref_cuda_stage(MeshView m, double gamma, const double* UL,
                                                           const double* UR, const double* Uo, double betta,
                                                           double dt, double* F, double* Un) {
  const int64_t nc = m.nc;
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= nc) return;
  const T g = (T)gamma;
  T acc[4] = {0, 0, 0, 0};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    T L[4], R[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      L[v] = (T)UR[((int64_t)k * 4 + v) * nc + i];  // the neighbour side is "L"
      R[v] = (T)UL[((int64_t)k * 4 + v) * nc + i];
    }
    const int64_t s = (int64_t)k * nc + i;
    cuda_ausm<T, Split>(L, R, (T)m.fnx[s], (T)m.fny[s], (T)m.flen[s], g, acc);
  }
  const double Rarea = m.rarea[i];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const double fl = (double)acc[v];
    F[v * nc + i] = fl;
    Un[v * nc + i] = Uo[v * nc + i] - fl * betta * dt * Rarea;
  }
}

// Phase one of calcOneStageMUSCL (computeMUSCLStencilsCUDA.jl:3-28 with
// computeInterfaceSlopeCUDA, computeslope2d.jl:139-269): every face slot's
// two states, UL/UR = [slot][4][nc] (slots past a cell's faces untouched).
template <bool QUADS>
__global__ __launch_bounds__(kBlock) void k_ref_face_states(RefStepArgs a, double* UL, double* UR) {
  const MeshView& m = a.m;
  const int64_t nc = m.nc;
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= nc) return;
  const double qL[4] = {a.Q[i], a.Q[nc + i], a.Q[2 * nc + i], a.Q[3 * nc + i]};
  const int nf = QUADS ? m.nv[i] : 3;
  for (int k = 0; k < nf; ++k) {
    double fL[4], fR[4];
    (void)ref_face_states<QUADS>(a, i, k, nf, qL, a.flow_time, fL, fR);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      UL[((int64_t)k * 4 + v) * nc + i] = fL[v];
      UR[((int64_t)k * 4 + v) * nc + i] = fR[v];
    }
  }
}

// FP: the loop with the face pass (a.rflux holds every face slot's flux)
template <bool QUADS, bool VISC, bool FUSED, bool FP>
__global__ __launch_bounds__(kBlock) void k_ref_step(RefStepArgs a) {
  const MeshView& m = a.m;
  const int64_t nc = m.nc;
  const int64_t i = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  double dt = a.dt, t = a.flow_time;
  if (FUSED) {
    const Acc r = a.acc[a.step % 3];
    t = r.time;
    dt = step_dt(r, a.dt_policy, a.dt, a.cfl, m.max_hx);
  }
  double vmx = 0.0, hdt = __longlong_as_double(0x7ff0000000000000ll);
  unsigned bad = 0;
  if (i < nc) {
    const double qL[4] = {a.Q[i], a.Q[nc + i], a.Q[2 * nc + i], a.Q[3 * nc + i]};
    // everything of the cell itself is read before the first store (the
    // compiler does not move a load past a store it cannot prove disjoint)
    const int nf = QUADS ? m.nv[i] : 3;
    double F[4];
    if (FP) ref_cell_flux_gathered<QUADS>(a, i, nf, F);
    else ref_cell_flux<QUADS>(a, i, nf, qL, t, F, FUSED ? a.wall_f : nullptr);
    double ud[4] = {0.0, 0.0, 0.0, 0.0};
    bool with_ud = FUSED ? true : (a.include_udiff && a.Ud);
    if (VISC) ref_udiff<QUADS>(m, a.prm, a.g6N, i, nf, qL, ud);
    else if (!FUSED && with_ud) {
#pragma unroll
      for (int v = 0; v < 4; ++v) ud[v] = a.Ud[v * nc + i];
    }
    const double Rarea = __builtin_nontemporal_load(&m.rarea[i]);
    const double b = a.betta;
    double Un[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double uo = __builtin_nontemporal_load(&a.Uo[v * nc + i]);
      Un[v] = with_ud ? (uo - F[v] * b * dt * Rarea + b * dt * ud[v]) : (uo - F[v] * b * dt * Rarea);
      __builtin_nontemporal_store(Un[v], &a.Un[v * nc + i]);
      if (a.F) a.F[v * nc + i] = F[v];
    }
    if (FUSED) {
      // updateVariablesSA, evaluate2d.jl:87-93
      const double G = a.prm.Gamma;
      const double rho = Un[0];
      const double ux = Un[1] / Un[0];
      const double uy = Un[2] / Un[0];
      const double p = (G - 1.0) * (Un[3] - 0.5 * (Un[1] * Un[1] + Un[2] * Un[2]) / Un[0]);
      const double as = sqrt(G * p / rho);
      const double vm = sqrt(ux * ux + uy * uy) + as;
      a.Qw[i] = rho; a.Qw[nc + i] = ux; a.Qw[2 * nc + i] = uy; a.Qw[3 * nc + i] = p;
      if (a.aS) a.aS[i] = as;
      if (a.vmax) a.vmax[i] = vm;
      vmx = vm;
      hdt = __builtin_nontemporal_load(&m.hmin[i]) / vm;
      bad = (!(rho > 0.0) || !(p > 0.0) || vm != vm) ? 1u : 0u;
      if (vm != vm) { vmx = 0.0; hdt = __longlong_as_double(0x7ff0000000000000ll); }
    }
  }
  if (FUSED) {
    reduce_acc(vmx, hdt, bad, a.part);
    rotate_acc(a.acc, a.step, t, dt);
  }
}

// cells2nodesSolutionReconstructionWithStencils (utilsFVM2dp.jl:299-365) plus,
// when PHI, the unweighted cluster means of calcPhiInCellCluster
// (calcGradients.jl:3-18) for rho*u, rho*v, E; QN / UN / PHI select the outputs
// (Qn, Un, phiN).  With fa.slot set, two leading blocks close the step that ran
// before: the dt partials into fa.slot and the wall faces' forces into fa.drag.
//
// NCL > 0 is the cluster width of the mesh, known at compile time: every
// slot's cell id and weight are loaded first (non-temporally: they are read
// once per step), then every slot's primitives, which are summed before every
// slot's conservative values are loaded and summed.  A node waits on three
// rounds of loads instead of two per slot, with half the values in flight of
// loading both at once (87 VGPRs instead of 110 at width 6: five waves per
// SIMD; 78.6 -> 64.5 us on the 2M TGV).  Each sum runs in slot order, skipping
// empty slots, exactly as the one-slot-at-a-time loop (NCL == 0, any width)
// does, so both give the same bits.
template <int NCL, bool QN, bool UN, bool PHI>
__global__ __launch_bounds__(kBlock) void k_ref_nodes(MeshView m, const double* Q, const double* U, double* Qn,
                                                      double* Un, double* phiN, FoldArgs fa) {
  const unsigned nfold = fa.slot ? 2u : 0u;
  if (blockIdx.x < nfold) {
    if (blockIdx.x == 0) fold_acc_block(fa.part, fa.nparts, fa.slot);
    else if (fa.wall_f) wall_sum_block(fa.wall_f, fa.n_wall, fa.drag);
    return;
  }
  const int64_t J0 = (int64_t)xcd_block(blockIdx.x - nfold, gridDim.x - nfold) * kBlock + threadIdx.x;
  const bool live = J0 < m.nn;
  const int64_t J = live ? J0 : m.nn - 1;
  const int64_t nc = m.nc, nn = m.nn;
  double det = 0.0, r = 0.0, u = 0.0, v = 0.0, p = 0.0;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  double f1 = 0.0, f2 = 0.0, f3 = 0.0;
  int cnt = 0;
  // one slot's contribution, in the source's order of operations
  auto add = [&](double wi, const double* qc, const double* uc) {
    if (QN) { r = r + qc[0] * wi; u = u + qc[1] * wi; v = v + qc[2] * wi; p = p + qc[3] * wi; }
    if (UN) { s0 = s0 + uc[0] * wi; s1 = s1 + uc[1] * wi; s2 = s2 + uc[2] * wi; s3 = s3 + uc[3] * wi; }
    if (PHI) { f1 = f1 + uc[1]; f2 = f2 + uc[2]; f3 = f3 + uc[3]; }
    det = det + wi;
    ++cnt;
  };
  if constexpr (NCL > 0) {
    int32_t c[NCL];
    double wv[NCL];
#pragma unroll
    for (int q = 0; q < NCL; ++q) {
      c[q] = __builtin_nontemporal_load(&m.cl[q * nn + J]);
      wv[q] = __builtin_nontemporal_load(&m.w[q * nn + J]);
    }
    // an empty slot loads cell 0 and is skipped in the sums
    if (QN) {
      double qv[NCL][4];
#pragma unroll
      for (int q = 0; q < NCL; ++q) {
        const int64_t k = c[q] >= 0 ? c[q] : 0;
#pragma unroll
        for (int t = 0; t < 4; ++t) qv[q][t] = Q[t * nc + k];
      }
#pragma unroll
      for (int q = 0; q < NCL; ++q)
        if (c[q] >= 0) {
          const double wi = wv[q];
          r = r + qv[q][0] * wi; u = u + qv[q][1] * wi; v = v + qv[q][2] * wi; p = p + qv[q][3] * wi;
        }
    }
    if (UN || PHI) {
      double uv[NCL][4];
#pragma unroll
      for (int q = 0; q < NCL; ++q) {
        const int64_t k = c[q] >= 0 ? c[q] : 0;
#pragma unroll
        for (int t = UN ? 0 : 1; t < 4; ++t) uv[q][t] = U[t * nc + k];
      }
#pragma unroll
      for (int q = 0; q < NCL; ++q)
        if (c[q] >= 0) {
          const double wi = wv[q];
          if (UN) { s0 = s0 + uv[q][0] * wi; s1 = s1 + uv[q][1] * wi; s2 = s2 + uv[q][2] * wi; s3 = s3 + uv[q][3] * wi; }
          if (PHI) { f1 = f1 + uv[q][1]; f2 = f2 + uv[q][2]; f3 = f3 + uv[q][3]; }
        }
    }
#pragma unroll
    for (int q = 0; q < NCL; ++q)
      if (c[q] >= 0) { det = det + wv[q]; ++cnt; }
  } else {
    for (int q = 0; q < m.ncl; ++q) {
      const int32_t c = m.cl[q * nn + J];
      if (c >= 0) {
        double qc[4] = {0, 0, 0, 0}, uc[4] = {0, 0, 0, 0};
        if (QN) { qc[0] = Q[c]; qc[1] = Q[nc + c]; qc[2] = Q[2 * nc + c]; qc[3] = Q[3 * nc + c]; }
        if (UN) uc[0] = U[c];
        if (UN || PHI) { uc[1] = U[nc + c]; uc[2] = U[2 * nc + c]; uc[3] = U[3 * nc + c]; }
        add(m.w[q * nn + J], qc, uc);
      }
    }
  }
  if (det != 0) {
    r = r / det; u = u / det; v = v / det; p = p / det;
    s0 = s0 / det; s1 = s1 / det; s2 = s2 / det; s3 = s3 / det;
  }
  // the node primitives are read three launches later, by the face pass, after
  // the viscous passes have moved several hundred MB: stored non-temporally
  if (QN && live) {
    __builtin_nontemporal_store(r, &Qn[J]);
    __builtin_nontemporal_store(u, &Qn[nn + J]);
    __builtin_nontemporal_store(v, &Qn[2 * nn + J]);
    __builtin_nontemporal_store(p, &Qn[3 * nn + J]);
  }
  if (UN && live) { Un[J] = s0; Un[nn + J] = s1; Un[2 * nn + J] = s2; Un[3 * nn + J] = s3; }
  if (PHI) {
    const double dc = (double)cnt;
    const double o[4] = {cnt ? f1 / dc : 0.0, cnt ? f2 / dc : 0.0, cnt ? f3 / dc : 0.0, 0.0};
    __shared__ WaveStage<2> stg;
    stg.store(phiN, J0, nn, o);
  }
}

// First Green-Gauss level of calcDiffTerm (calcDiffTerm.jl:65-67): for
// rho*u, rho*v, E the pair ("gradX", "gradY") = (d/dy, d/dx).
template <bool QUADS>
__global__ __launch_bounds__(kBlock) void k_ref_gg1(MeshView m, const double* phiN, double* g6) {
  const int64_t i0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  const int64_t i = i0 < m.nc ? i0 : m.nc - 1;
  const int nf = QUADS ? m.nv[i] : 3;
  const double A = __builtin_nontemporal_load(&m.area[i]);
  double ax[3], ay[3], o[6];
  gg_sums<QUADS, 4, 3>(m, i, nf, phiN, ax, ay);
#pragma unroll
  for (int c = 0; c < 3; ++c) { o[2 * c] = ax[c] / A; o[2 * c + 1] = ay[c] / A; }
  __shared__ WaveStage<3> stg;
  stg.store(g6, i0, m.nc, o);
}

// Cluster means of the six gradient fields (calcPhiInCellCluster applied
// inside the second gradient level, calcDiffTerm.jl:88-95); NCL as in
// k_ref_nodes (every slot's record loaded before the sums, which keep the
// slot order).
template <int NCL>
__global__ __launch_bounds__(kBlock) void k_ref_gg_nodes(MeshView m, const double* g6, double* g6N) {
  const int64_t J0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  const int64_t J = J0 < m.nn ? J0 : m.nn - 1;
  const int64_t nn = m.nn;
  double s[6] = {0, 0, 0, 0, 0, 0};
  int cnt = 0;
  if constexpr (NCL > 0) {
    int32_t c[NCL];
#pragma unroll
    for (int q = 0; q < NCL; ++q) c[q] = __builtin_nontemporal_load(&m.cl[q * nn + J]);
    double x[NCL][6];
#pragma unroll
    for (int q = 0; q < NCL; ++q) ld2k<3>(g6 + 6 * (int64_t)(c[q] >= 0 ? c[q] : 0), x[q]);
#pragma unroll
    for (int q = 0; q < NCL; ++q)
      if (c[q] >= 0) {
#pragma unroll
        for (int t = 0; t < 6; ++t) s[t] = s[t] + x[q][t];
        ++cnt;
      }
  } else {
    for (int q = 0; q < m.ncl; ++q) {
      const int32_t c = m.cl[q * nn + J];
      if (c >= 0) {
        double x[6];
        ld2k<3>(g6 + 6 * (int64_t)c, x);
#pragma unroll
        for (int t = 0; t < 6; ++t) s[t] = s[t] + x[t];
        ++cnt;
      }
    }
  }
  const double dc = (double)cnt;
  double o[6];
#pragma unroll
  for (int t = 0; t < 6; ++t) o[t] = cnt ? s[t] / dc : 0.0;
  __shared__ WaveStage<3> stg;
  stg.store(g6N, J0, nn, o);
}

template <bool QUADS>
__global__ __launch_bounds__(kBlock) void k_ref_diff_term(MeshView m, gd_params p, const double* Q, const double* g6N,
                                                          double* Ud) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= m.nc) return;
  const int64_t nc = m.nc;
  const double q[4] = {Q[i], Q[nc + i], Q[2 * nc + i], Q[3 * nc + i]};
  double ud[4] = {0, 0, 0, 0};
  ref_udiff<QUADS>(m, p, g6N, i, QUADS ? m.nv[i] : 3, q, ud);
  Ud[nc + i] = ud[1];
  Ud[2 * nc + i] = ud[2];
  Ud[3 * nc + i] = ud[3];
}

// updateVariablesSA, evaluate2d.jl:83-108
__global__ __launch_bounds__(kBlock) void k_update_variables(int64_t nc, double G, const double* Un, double* Uo,
                                                             double* Q, double* aS, double* vmax, double* Delta) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= nc) return;
  const double u0 = Un[i], u1 = Un[nc + i], u2 = Un[2 * nc + i], u3 = Un[3 * nc + i];
  const double rho = u0, ux = u1 / u0, uy = u2 / u0;
  const double p = (G - 1.0) * (u3 - 0.5 * (u1 * u1 + u2 * u2) / u0);
  const double as = sqrt(G * p / rho);
  Q[i] = rho; Q[nc + i] = ux; Q[2 * nc + i] = uy; Q[3 * nc + i] = p;
  aS[i] = as;
  vmax[i] = sqrt(ux * ux + uy * uy) + as;
  const double un[4] = {u0, u1, u2, u3};
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    if (Delta) Delta[v * nc + i] = un[v] - Uo[v * nc + i];
    Uo[v * nc + i] = un[v];
  }
}

__global__ void k_acc_reset(Acc* acc, double t) {
  if (threadIdx.x < 3) {
    Acc* z = &acc[threadIdx.x];
    z->vmax_bits = 0ull;
    z->hdt_bits = 0x7ff0000000000000ull;
    z->nonphysical = 0u;
    z->time = t;
    z->dt = 0.0;
    z->force[0] = z->force[1] = 0.0;
  }
}

// Fills accumulator slot 0 from the current state (the reference takes
// max(VMAXCells) at the top of each step, testViscousFluxesTGV2d.jl:420).
__global__ __launch_bounds__(kBlock) void k_prime_acc(MeshView m, gd_params p, const double* Q, const double* vmax,
                                                      double* part, int32_t viscous_ns) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double vm = 0.0, hdt = __longlong_as_double(0x7ff0000000000000ll);
  unsigned bad = 0;
  if (i < m.nc) {
    const int64_t nc = m.nc;
    const double rho = Q[i], pr = Q[3 * nc + i];
    vm = vmax[i];
    double lam = vm;
    if (viscous_ns) {
      const double T = pr / (rho * p.RGAS);
      const double t = T / p.T_ref;
      const double mu = p.mu_ref * (t * sqrt(t)) * (p.T_ref + p.S_suth) / (T + p.S_suth);
      lam = vm + 2.0 * p.Gamma * mu / (p.Pr * rho * m.hmin[i]);
    }
    hdt = m.hmin[i] / lam;
    bad = (!(rho > 0.0) || !(pr > 0.0) || vm != vm) ? 1u : 0u;
    if (vm != vm) { vm = 0.0; hdt = __longlong_as_double(0x7ff0000000000000ll); }
  }
  reduce_acc(vm, hdt, bad, part);
}

// The start of the ns loop from U alone: the primitives in the compute type
// (the gradient pass's input) and the first step's dt inputs, computed as
// updateVariablesSA's primitives and k_prime_acc's reduction would be
// (evaluate2d.jl:87-93), without storing the FP64 fields.
template <class T>
__device__ __forceinline__ void st_prims(T* q, double r, double u, double v, double p);
template <>
__device__ __forceinline__ void st_prims<double>(double* q, double r, double u, double v, double p) {
  const double o[4] = {r, u, v, p};
  st2k<2>(q, o);
}
template <>
__device__ __forceinline__ void st_prims<float>(float* q, double r, double u, double v, double p) {
  *reinterpret_cast<float4*>(q) = make_float4((float)r, (float)u, (float)v, (float)p);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_ns_prime(MeshView m, gd_params p, const double* U, T* Qt, double* part,
                                                     int32_t viscous_ns) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double vm = 0.0, hdt = __longlong_as_double(0x7ff0000000000000ll);
  unsigned bad = 0;
  if (i < m.nc) {
    const int64_t nc = m.nc;
    const double G = p.Gamma;
    const double u0 = U[i], u1 = U[nc + i], u2 = U[2 * nc + i], u3 = U[3 * nc + i];
    const double rho = u0, ux = u1 / u0, uy = u2 / u0;
    const double pr = (G - 1.0) * (u3 - 0.5 * (u1 * u1 + u2 * u2) / u0);
    st_prims<T>(Qt + 4 * i, rho, ux, uy, pr);
    const double as = sqrt(G * pr / rho);
    vm = sqrt(ux * ux + uy * uy) + as;
    double lam = vm;
    if (viscous_ns) {
      const double Tc = pr / (rho * p.RGAS);
      const double t = Tc / p.T_ref;
      const double mu = p.mu_ref * (t * sqrt(t)) * (p.T_ref + p.S_suth) / (Tc + p.S_suth);
      lam = vm + 2.0 * p.Gamma * mu / (p.Pr * rho * m.hmin[i]);
    }
    hdt = m.hmin[i] / lam;
    bad = (!(rho > 0.0) || !(pr > 0.0) || vm != vm) ? 1u : 0u;
    if (vm != vm) { vm = 0.0; hdt = __longlong_as_double(0x7ff0000000000000ll); }
  }
  reduce_acc(vm, hdt, bad, part);
}

// Folds the per-block partials of reduce_acc into one accumulator slot, in a
// fixed order (max and min are exact; the count is an integer sum).
__global__ __launch_bounds__(1024) void k_acc_fold(const double* part, int64_t nparts, Acc* w) {
  fold_acc_block(part, nparts, w);
}

// Deterministic sums of (Unew - Uold)^2 per variable: per-block partials, then
// one block folds them in a fixed order.
__global__ __launch_bounds__(kBlock) void k_resid_partial(int64_t nc, const double* Un, const double* Uo,
                                                          double* part) {
  __shared__ double s[4][kBlock];
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    double d = 0.0;
    if (i < nc) { d = Un[v * nc + i] - Uo[v * nc + i]; d = d * d; }
    s[v][threadIdx.x] = d;
  }
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o)
#pragma unroll
      for (int v = 0; v < 4; ++v) s[v][threadIdx.x] += s[v][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 4) part[(int64_t)blockIdx.x * 4 + threadIdx.x] = s[threadIdx.x][0];
}

__global__ __launch_bounds__(kBlock) void k_fold(const double* part, int64_t nparts, int width, double* out) {
  __shared__ double s[kBlock];
  for (int v = 0; v < width; ++v) {
    double a = 0.0;
    for (int64_t b = threadIdx.x; b < nparts; b += kBlock) a += part[b * width + v];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
      if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[v] = s[0];
    __syncthreads();
  }
}

// wall force of one step: fold the per-face contributions in a fixed order;
// out = {fx, fy, running sum fx, running sum fy}
__global__ __launch_bounds__(kBlock) void k_wall_sum(const double* wf, int32_t n, double* out) {
  wall_sum_block(wf, n, out);
}

// start of a run: the three accumulator slots at flow_time, no wall force yet
__global__ __launch_bounds__(kBlock) void k_run_begin(Acc* acc, double t, double* wall_f, int32_t n_wall,
                                                      double* drag) {
  if (threadIdx.x < 3) {
    Acc* z = &acc[threadIdx.x];
    z->vmax_bits = 0ull;
    z->hdt_bits = 0x7ff0000000000000ull;
    z->nonphysical = 0u;
    z->time = t;
    z->dt = 0.0;
    z->force[0] = z->force[1] = 0.0;
  }
  if (threadIdx.x < 4) drag[threadIdx.x] = 0.0;
  if (wall_f)
    for (int64_t k = threadIdx.x; k < 2 * (int64_t)n_wall; k += kBlock) wall_f[k] = 0.0;
}

// end of the ns loop: block 0 folds the last update's dt partials, block 1
// sums the last face pass's wall forces (the same block size as the gradient
// launch's fold blocks, so the sums are the same as theirs)
__global__ __launch_bounds__(kBlock) void k_ns_close(const double* part, int64_t nparts, Acc* slot,
                                                     const double* wall_f, int32_t n_wall, double* drag) {
  if (blockIdx.x == 0) fold_acc_block(part, nparts, slot);
  else wall_sum_block(wall_f, n_wall, drag);
}

// aSoundCells and VMAXCells from the primitives (initfields_TGV2d.jl:47-48,
// evaluate2d.jl:92-93)
__global__ __launch_bounds__(kBlock) void k_fields_derived(int64_t nc, double G, const double* Q, double* aS,
                                                           double* vmax) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= nc) return;
  const double rho = Q[i], ux = Q[nc + i], uy = Q[2 * nc + i], p = Q[3 * nc + i];
  const double as = sqrt(G * p / rho);
  aS[i] = as;
  vmax[i] = sqrt(ux * ux + uy * uy) + as;
}

__global__ __launch_bounds__(kBlock) void k_fill(double* p, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_copy(double* d, const double* s, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) d[i] = s[i];
}

// column-wise gather dst[q][o] = src[q][idx[o]] (host <-> device numbering)
__global__ __launch_bounds__(kBlock) void k_permute(double* d, const double* s, const int32_t* idx, int64_t n,
                                                    int32_t ncols) {
  const int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (o >= n) return;
  const int64_t j = idx[o];
  for (int32_t q = 0; q < ncols; ++q) d[q * n + o] = s[q * n + j];
}

__global__ __launch_bounds__(kBlock) void k_delta(int64_t nc, const double* Un, const double* Uo, double* D) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < 4 * nc) D[i] = Un[i] - Uo[i];
}

inline unsigned nblocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

hipError_t launch_fields_derived(int64_t nc, double gamma, const double* Q, double* aS, double* vmax, hipStream_t s) {
  hipLaunchKernelGGL(k_fields_derived, dim3(nblocks(nc)), dim3(kBlock), 0, s, nc, gamma, Q, aS, vmax);
  return hipGetLastError();
}

hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3(nblocks(n)), dim3(kBlock), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_copy(double* dst, const double* src, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_copy, dim3(nblocks(n)), dim3(kBlock), 0, s, dst, src, n);
  return hipGetLastError();
}

hipError_t launch_permute(double* dst, const double* src, const int32_t* idx, int64_t n, int32_t ncols,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_permute, dim3(nblocks(n)), dim3(kBlock), 0, s, dst, src, idx, n, ncols);
  return hipGetLastError();
}

hipError_t launch_delta(int64_t nc, const double* Un, const double* Uo, double* D, hipStream_t s) {
  hipLaunchKernelGGL(k_delta, dim3(nblocks(4 * nc)), dim3(kBlock), 0, s, nc, Un, Uo, D);
  return hipGetLastError();
}

hipError_t launch_ref_step(const RefStepArgs& a, hipStream_t s) {
  const bool q = a.m.has_quads != 0;
  const bool fused = a.acc != nullptr;
  const bool visc = a.g6N != nullptr;
  const dim3 g(nblocks(a.m.nc)), b(kBlock);
#define GD_LAUNCH(Q_, V_, F_, P_) hipLaunchKernelGGL((k_ref_step<Q_, V_, F_, P_>), g, b, 0, s, a)
  if (fused && a.rflux) {
    if (q) { if (visc) GD_LAUNCH(true, true, true, true); else GD_LAUNCH(true, false, true, true); }
    else { if (visc) GD_LAUNCH(false, true, true, true); else GD_LAUNCH(false, false, true, true); }
  } else if (fused) {
    if (q) { if (visc) GD_LAUNCH(true, true, true, false); else GD_LAUNCH(true, false, true, false); }
    else { if (visc) GD_LAUNCH(false, true, true, false); else GD_LAUNCH(false, false, true, false); }
  } else {
    if (q) GD_LAUNCH(true, false, false, false); else GD_LAUNCH(false, false, false, false);
  }
#undef GD_LAUNCH
  return hipGetLastError();
}

hipError_t launch_ref_face_states(const RefStepArgs& a, double* UL, double* UR, hipStream_t s) {
  if (a.m.has_quads)
    hipLaunchKernelGGL(k_ref_face_states<true>, dim3(nblocks(a.m.nc)), dim3(kBlock), 0, s, a, UL, UR);
  else
    hipLaunchKernelGGL(k_ref_face_states<false>, dim3(nblocks(a.m.nc)), dim3(kBlock), 0, s, a, UL, UR);
  return hipGetLastError();
}

hipError_t launch_ref_faces(const RefStepArgs& a, double* rflux, hipStream_t s) {
  if (a.m.n_rfaces == 0) return hipSuccess;
  if (a.m.has_quads)
    hipLaunchKernelGGL(k_ref_faces<true>, dim3(nblocks(a.m.n_rfaces)), dim3(kBlock), 0, s, a, rflux);
  else
    hipLaunchKernelGGL(k_ref_faces<false>, dim3(nblocks(a.m.n_rfaces)), dim3(kBlock), 0, s, a, rflux);
  return hipGetLastError();
}

hipError_t launch_ref_cuda_stage(const MeshView& m, double gamma, bool x32, const double* UL, const double* UR,
                                 const double* Uo, double betta, double dt, double* F, double* Un, hipStream_t s) {
  if (x32)
    hipLaunchKernelGGL((k_ref_cuda_stage<float, CudaSplitX32>), dim3(nblocks(m.nc)), dim3(kBlock), 0, s, m, gamma, UL,
                       UR, Uo, betta, dt, F, Un);
  else
    hipLaunchKernelGGL((k_ref_cuda_stage<double, CudaSplitF64>), dim3(nblocks(m.nc)), dim3(kBlock), 0, s, m, gamma,
                       UL, UR, Uo, betta, dt, F, Un);
  return hipGetLastError();
}

// the cluster widths with a kernel of their own; others take the slot loop
template <template <int> class L, class... A>
static void by_width(int ncl, A&&... args) {
  switch (ncl) {
    case 3: L<3>::run(args...); break;
    case 4: L<4>::run(args...); break;
    case 5: L<5>::run(args...); break;
    case 6: L<6>::run(args...); break;
    case 7: L<7>::run(args...); break;
    case 8: L<8>::run(args...); break;
    default: L<0>::run(args...); break;
  }
}
template <bool QN, bool UN, bool PHI>
struct NodesLaunch {
  template <int NCL>
  struct W {
    static void run(const MeshView& m, dim3 g, hipStream_t s, const double* Q, const double* U, double* Qn, double* Un,
                    double* phiN, const FoldArgs& fa) {
      hipLaunchKernelGGL((k_ref_nodes<NCL, QN, UN, PHI>), g, dim3(kBlock), 0, s, m, Q, U, Qn, Un, phiN, fa);
    }
  };
};
template <int NCL>
struct GgNodesLaunch {
  static void run(const MeshView& m, hipStream_t s, const double* g6, double* g6N) {
    hipLaunchKernelGGL(k_ref_gg_nodes<NCL>, dim3(nblocks(m.nn)), dim3(kBlock), 0, s, m, g6, g6N);
  }
};

hipError_t launch_ref_nodes(const MeshView& m, const double* Q, const double* U, double* Qn, double* Un, double* phiN,
                            hipStream_t s, const FoldArgs& fold) {
  const dim3 g(nblocks(m.nn) + (fold.slot ? 2u : 0u));
  // the loop's node pass, gd_cells2nodes and the cluster means of calcDiffTerm
  // have kernels per cluster width; other output sets take the slot loop
  switch ((Qn ? 4 : 0) | (Un ? 2 : 0) | (phiN ? 1 : 0)) {
    case 5: by_width<NodesLaunch<true, false, true>::W>(m.ncl, m, g, s, Q, U, Qn, Un, phiN, fold); break;
    case 6: by_width<NodesLaunch<true, true, false>::W>(m.ncl, m, g, s, Q, U, Qn, Un, phiN, fold); break;
    case 1: by_width<NodesLaunch<false, false, true>::W>(m.ncl, m, g, s, Q, U, Qn, Un, phiN, fold); break;
    case 4: NodesLaunch<true, false, false>::W<0>::run(m, g, s, Q, U, Qn, Un, phiN, fold); break;
    case 2: NodesLaunch<false, true, false>::W<0>::run(m, g, s, Q, U, Qn, Un, phiN, fold); break;
    case 3: NodesLaunch<false, true, true>::W<0>::run(m, g, s, Q, U, Qn, Un, phiN, fold); break;
    case 7: NodesLaunch<true, true, true>::W<0>::run(m, g, s, Q, U, Qn, Un, phiN, fold); break;
    default: NodesLaunch<false, false, false>::W<0>::run(m, g, s, Q, U, Qn, Un, phiN, fold); break;
  }
  return hipGetLastError();
}

hipError_t launch_ref_phi(const MeshView& m, const double* U, double* phiN, hipStream_t s) {
  return launch_ref_nodes(m, nullptr, U, nullptr, nullptr, phiN, s, FoldArgs{});
}

hipError_t launch_ns_prime(const MeshView& m, const gd_params& prm, const double* U, void* Qt, bool f32, Acc* acc,
                           double* part, double flow_time, int32_t viscous_ns, hipStream_t s, bool reset_acc) {
  if (reset_acc) hipLaunchKernelGGL(k_acc_reset, dim3(1), dim3(64), 0, s, acc, flow_time);
  if (f32)
    hipLaunchKernelGGL(k_ns_prime<float>, dim3(nblocks(m.nc)), dim3(kBlock), 0, s, m, prm, U, static_cast<float*>(Qt),
                       part, viscous_ns);
  else
    hipLaunchKernelGGL(k_ns_prime<double>, dim3(nblocks(m.nc)), dim3(kBlock), 0, s, m, prm, U,
                       static_cast<double*>(Qt), part, viscous_ns);
  return hipGetLastError();
}

hipError_t launch_ref_gg1(const MeshView& m, const double* phiN, double* g6, hipStream_t s) {
  if (m.has_quads) hipLaunchKernelGGL(k_ref_gg1<true>, dim3(nblocks(m.nc)), dim3(kBlock), 0, s, m, phiN, g6);
  else hipLaunchKernelGGL(k_ref_gg1<false>, dim3(nblocks(m.nc)), dim3(kBlock), 0, s, m, phiN, g6);
  return hipGetLastError();
}

hipError_t launch_ref_gg_nodes(const MeshView& m, const double* g6, double* g6N, hipStream_t s) {
  by_width<GgNodesLaunch>(m.ncl, m, s, g6, g6N);
  return hipGetLastError();
}

hipError_t launch_ref_diff_term(const MeshView& m, const gd_params& prm, const double* Q, const double* g6N,
                                double* Ud, hipStream_t s) {
  if (m.has_quads)
    hipLaunchKernelGGL(k_ref_diff_term<true>, dim3(nblocks(m.nc)), dim3(kBlock), 0, s, m, prm, Q, g6N, Ud);
  else
    hipLaunchKernelGGL(k_ref_diff_term<false>, dim3(nblocks(m.nc)), dim3(kBlock), 0, s, m, prm, Q, g6N, Ud);
  return hipGetLastError();
}

hipError_t launch_update_variables(const MeshView& m, double gamma, const double* Unew, double* Uold, double* Q,
                                   double* aS, double* vmax, double* Delta, hipStream_t s) {
  hipLaunchKernelGGL(k_update_variables, dim3(nblocks(m.nc)), dim3(kBlock), 0, s, m.nc, gamma, Unew, Uold, Q, aS,
                     vmax, Delta);
  return hipGetLastError();
}

hipError_t launch_prime_acc(const MeshView& m, const gd_params& prm, const double* Q, const double* vmax, Acc* acc,
                            double* part, double flow_time, int32_t dt_policy, int32_t viscous_ns, hipStream_t s) {
  (void)dt_policy;
  hipLaunchKernelGGL(k_acc_reset, dim3(1), dim3(64), 0, s, acc, flow_time);
  hipLaunchKernelGGL(k_prime_acc, dim3(nblocks(m.nc)), dim3(kBlock), 0, s, m, prm, Q, vmax, part, viscous_ns);
  hipLaunchKernelGGL(k_acc_fold, dim3(1), dim3(1024), 0, s, part, (int64_t)nblocks(m.nc), acc);
  return hipGetLastError();
}

hipError_t launch_acc_fold(const double* part, int64_t nparts, Acc* slot, hipStream_t s) {
  hipLaunchKernelGGL(k_acc_fold, dim3(1), dim3(1024), 0, s, part, nparts, slot);
  return hipGetLastError();
}

hipError_t launch_residuals(const MeshView& m, const double* Unew, const double* Uold, double* scratch, double* out,
                            hipStream_t s) {
  const unsigned nb = nblocks(m.nc);
  hipLaunchKernelGGL(k_resid_partial, dim3(nb), dim3(kBlock), 0, s, m.nc, Unew, Uold, scratch);
  hipLaunchKernelGGL(k_fold, dim3(1), dim3(kBlock), 0, s, scratch, (int64_t)nb, 4, out);
  return hipGetLastError();
}

hipError_t launch_run_begin(Acc* acc, double flow_time, double* wall_f, int32_t n_wall, double* drag, hipStream_t s) {
  hipLaunchKernelGGL(k_run_begin, dim3(1), dim3(kBlock), 0, s, acc, flow_time, wall_f, n_wall, drag);
  return hipGetLastError();
}

hipError_t launch_ns_close(const double* part, int64_t nparts, Acc* slot, const double* wall_f, int32_t n_wall,
                           double* drag, hipStream_t s) {
  hipLaunchKernelGGL(k_ns_close, dim3(wall_f ? 2 : 1), dim3(kBlock), 0, s, part, nparts, slot, wall_f, n_wall, drag);
  return hipGetLastError();
}

hipError_t launch_wall_sum(const double* wall_f, int32_t n, double* out4, hipStream_t s) {
  hipLaunchKernelGGL(k_wall_sum, dim3(1), dim3(kBlock), 0, s, wall_f, n, out4);
  return hipGetLastError();
}

}  // namespace gd
