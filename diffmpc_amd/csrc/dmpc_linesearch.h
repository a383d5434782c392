// Merit line search of the SQP step (diffmpc/solvers/sqp.py:301-411), as a
// workgroup-cooperative device function shared by the fused kernels.
#pragma once
#include "dmpc_device.h"
#include "dmpc_internal.h"
#include "dmpc_models.h"

namespace dmpc {

// Scratch doubles the line search needs for one instance: merit parts of
// (n_alphas + 1) x (N + 1) items (cost, constraint) and their sums.
__host__ __device__ inline int linesearch_doubles(int linesearch, int n_alphas, int N) {
  return linesearch ? 2 * (n_alphas + 1) * (N + 1) + 2 * (n_alphas + 1) : 0;
}

namespace {

// Nocedal & Wright, Alg. 18.3, for instance b with horizon Nb:
//   phi(z; mu) = cost(z) + mu ||g(z)||_1,
//   mu = max(0.1, (dcost/dz . dz) / (0.5 ||g||_1))   (NaN propagates, as jnp.maximum),
//   decrease_k = phi(z + alpha_k dz) - (phi(z) + eta alpha_k D phi),
// then the candidates are scanned from the largest with the reference's
// clamped indexing (alphas[min(i, n-1)]).  z = (xb, ub) rows 0..Nb (row
// strides nx, nu), dz = (DX, DU) likewise.  Everything is float64, including
// the dynamics evaluations of the constraints g = [x_0 - x_init;
// f_d(x_t, u_t) - x_{t+1}] (optimal_control_problem.py:104-151).  Items
// (candidate k, knot t) are spread over the workgroup and summed in a fixed
// order, so every thread returns the same alpha.  LS holds
// linesearch_doubles(1, n_alphas, Nb) doubles; red is a 16-double LDS
// reduction buffer (TEAMS: the team's 64-double block_sum_rows buffer, zero
// past the team's rows).  Problem arrays are read with the padded knot
// stride a.N + 1.  Must be called by every thread of the workgroup; tid and
// nt are the thread's index and the thread count of its instance's team
// (the workgroup's with one instance per workgroup).
template <class Model, bool TEAMS = false>
__device__ double line_search(const SolveArgs& a, int b, int Nb, const double* xb, const double* ub,
                              const double* DX, const double* DU, double* LS, double* red, int tid, int nt) {
  constexpr int NX = Model::NX, NU = Model::NU;
  const int NB = Nb + 1, NBs = a.N + 1, nal = a.n_alphas, items = (nal + 1) * NB;
  double* LC = LS;               // cost part of each item
  double* LG = LS + items;       // constraint part of each item
  double* SUM = LS + 2 * items;  // (nal + 1) x {cost, constraint}
  const double* wx = a.wx + (size_t)b * NX;
  const double* wu = a.wu + (size_t)b * NU;
  const double* wf = a.wf + (size_t)b * NX;
  const double* xr = a.xref + (size_t)b * NBs * NX;
  const double* ur = a.uref + (size_t)b * NBs * NU;
  const double* xf = a.xf + (size_t)b * NX;
  const double* x0 = a.x0 + (size_t)b * NX;
  const float* prm = a.dyn + (size_t)b * a.P;
  // directional derivative of the cost (the last control row has none)
  double part = 0.0;
  for (int k = tid; k < NB * NX; k += nt) {
    const int t = k / NX, i = k - t * NX;
    const double gx = (t < Nb) ? 2.0 * wx[i] * (xb[k] - xr[k]) : 2.0 * (wx[i] + wf[i]) * (xb[k] - xf[i]);
    part += gx * DX[k];
  }
  for (int k = tid; k < Nb * NU; k += nt) {
    const int j = k % NU;
    const double ref = a.penalize ? ur[k] : 0.0;
    part += 2.0 * wu[j] * (ub[k] - ref) * DU[k];
  }
  double dcdz;
  if constexpr (TEAMS) {
    dcdz = block_sum_rows(part, red, tid >> 4);
  } else {
    dcdz = block_sum(part, red);
  }
  // merit parts: item (kc, t) holds the cost of knot t and the constraint
  // block t at alpha = 0 (kc = 0) or alphas[kc - 1]
  for (int j = tid; j < items; j += nt) {
    const int kc = j / NB, t = j - kc * NB;
    const double al = kc == 0 ? 0.0 : a.alphas[kc - 1];
    double c = 0.0, g = 0.0;
    if (t < Nb) {
#pragma unroll
      for (int i = 0; i < NX; ++i) {
        const double d = xb[t * NX + i] + al * DX[t * NX + i] - xr[t * NX + i];
        c += wx[i] * d * d;
      }
#pragma unroll
      for (int i = 0; i < NU; ++i) {
        const double d = ub[t * NU + i] + al * DU[t * NU + i] - (a.penalize ? ur[t * NU + i] : 0.0);
        c += wu[i] * d * d;
      }
    } else {
#pragma unroll
      for (int i = 0; i < NX; ++i) {
        const double d = xb[t * NX + i] + al * DX[t * NX + i] - xf[i];
        c += (wx[i] + wf[i]) * d * d;
      }
    }
    if (t == 0) {
#pragma unroll
      for (int i = 0; i < NX; ++i) g += fabs(xb[i] + al * DX[i] - x0[i]);
    } else {
      double xp[NX], up[NU], fx[NX];
#pragma unroll
      for (int i = 0; i < NX; ++i) xp[i] = xb[(t - 1) * NX + i] + al * DX[(t - 1) * NX + i];
#pragma unroll
      for (int i = 0; i < NU; ++i) up[i] = ub[(t - 1) * NU + i] + al * DU[(t - 1) * NU + i];
      step<Model, double>(xp, up, prm, a.dt64, a.scheme, fx);
#pragma unroll
      for (int i = 0; i < NX; ++i) g += fabs(fx[i] - (xb[t * NX + i] + al * DX[t * NX + i]));
    }
    LC[j] = c;
    LG[j] = g;
  }
  __syncthreads();
  if (tid <= nal) {
    double c = 0.0, g = 0.0;
    for (int t = 0; t < NB; ++t) {
      c += LC[tid * NB + t];
      g += LG[tid * NB + t];
    }
    SUM[2 * tid] = c;
    SUM[2 * tid + 1] = g;
  }
  __syncthreads();
  const double c0 = SUM[0], g0 = SUM[1];
  const double mu_min = dcdz / (0.5 * g0);
  const double mu = (mu_min != mu_min) ? mu_min : fmax(0.1, mu_min);
  const double merit = c0 + mu * g0;
  const double dmerit = dcdz + -mu * g0;
  double amax = a.alphas[0];
  for (int k = 1; k < nal; ++k) amax = (a.alphas[k] != a.alphas[k] || amax != amax) ? NAN : fmax(amax, a.alphas[k]);
  // decrease of candidate k (0-based in alphas)
  auto decrease = [&](int k) {
    const double al = a.alphas[k];
    return (SUM[2 * (k + 1)] + mu * SUM[2 * (k + 1) + 1]) - (merit + a.ls_eta * al * dmerit);
  };
  int i = nal + 1;
  double alpha = amax;
  while (!(decrease(i < nal - 1 ? i : nal - 1) < 0.0 || i == 0)) {
    alpha = a.alphas[(i - 1) < nal - 1 ? (i - 1) : nal - 1];
    --i;
  }
  return alpha;
}

}  // namespace
}  // namespace dmpc
