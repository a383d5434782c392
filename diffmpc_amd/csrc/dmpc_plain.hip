// Per-launch ("plain") GPU path: the reference's op graph with one kernel per
// operation and host-driven loops.  Each SQP iteration materialises the dense
// QP blocks of diffmpc/solvers/sqp.py:230-260, assembles S and Phiinv as
// pcg.py:160-243 lays them out (N+1, nx, 3nx), and the PCG while-loop of
// pcg.py:397-436 reads its predicate back to the host on every trip, as a
// vmapped lax.while_loop does.  All QP arithmetic is float64; linearization
// shares the float32 dual-number code of the fused path.
//
// The dense kernels here are also the PCGOptimalControl-level entry points
// of the C ABI (dmpc_pcg_assemble / _gamma / _solve / _recover) and accept
// general As_next, Qinv, Rinv blocks.
#include <vector>

#include "dmpc_internal.h"
#include "dmpc_models.h"

namespace dmpc {
namespace {

#define HIPCHK(x)                              \
  do {                                         \
    if ((x) != hipSuccess) return DMPC_ERR_HIP; \
  } while (0)

inline int nblk(long long n, int t = 256) { return (int)((n + t - 1) / t); }

// ---------------------------------------------------------------------------
// linearization: one thread per (b, t, j), column j of [A_t B_t] at (x_t, u_t)
template <class Model>
__global__ void lin_kernel(int B, int N, int P, float dt, double dt64, int scheme, const float* dyn, const double* x,
                           const double* u, double* fd, double* As, double* Bs) {
  constexpr int NX = Model::NX, NU = Model::NU, NZ = NX + NU;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)B * N * NZ) return;
  const int j = idx % NZ;
  const int t = (idx / NZ) % N;
  const int b = idx / ((long long)NZ * N);
  float xs[NX], us[NU], prm[Model::NPARAM];
  for (int i = 0; i < NX; ++i) xs[i] = (float)x[((size_t)b * (N + 1) + t) * NX + i];
  for (int k = 0; k < NU; ++k) us[k] = (float)u[((size_t)b * (N + 1) + t) * NU + k];
  for (int k = 0; k < Model::NPARAM; ++k) prm[k] = dyn[(size_t)b * P + k];
  float fv[NX], col[NX];
  step_jacobian_column<Model>(xs, us, prm, dt, scheme, j, fv, col);
  const size_t bt = (size_t)b * N + t;
  if (j < NX) {
    for (int i = 0; i < NX; ++i) As[(bt * NX + i) * NX + j] = col[i];
  } else {
    for (int i = 0; i < NX; ++i) Bs[(bt * NX + i) * NU + (j - NX)] = col[i];
  }
  // the value in float64 (the Jacobian in float32), as the fused linearization
  if (j == 0) fd_value<Model>(x, u, prm, dt64, scheme, N, b, t, fd + bt * NX);
}

// C_0 = x0; C_{t+1} = -f_d + A_t x_t + B_t u_t; As_next = -I
// (optimal_control_problem.py:384-397)
__global__ void c_kernel(int B, int N, int nx, int nu, const double* x0, const double* x, const double* u,
                         const double* fd, const double* As, const double* Bs, double* Cs, double* As_next) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)B * (N + 1) * nx) return;
  const int i = idx % nx;
  const int t = (idx / nx) % (N + 1);
  const int b = idx / ((long long)nx * (N + 1));
  double c;
  if (t == 0) {
    c = x0[b * nx + i];
  } else {
    const size_t bt = (size_t)b * N + (t - 1);
    c = -fd[bt * nx + i];
    for (int k = 0; k < nx; ++k) c += As[(bt * nx + i) * nx + k] * x[((size_t)b * (N + 1) + t - 1) * nx + k];
    for (int k = 0; k < nu; ++k) c += Bs[(bt * nx + i) * nu + k] * u[((size_t)b * (N + 1) + t - 1) * nu + k];
    for (int k = 0; k < nx; ++k) As_next[(bt * nx + i) * nx + k] = (k == i) ? -1.0 : 0.0;
  }
  Cs[idx] = c;
}

// dense cost blocks of the diagonal tracking cost (sqp.py:160-197, 243-244)
__global__ void cost_kernel(int B, int N, int nx, int nu, int penalize, const double* wx, const double* wu,
                            const double* wf, const double* xref, const double* uref, const double* xf,
                            double* Qinv, double* qv, double* Rinv, double* rv) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nq = (long long)B * (N + 1) * nx * nx;
  const long long nr = (long long)B * N * nu * nu;
  if (idx < nq) {
    const int j = idx % nx;
    const int i = (idx / nx) % nx;
    const int t = (idx / ((long long)nx * nx)) % (N + 1);
    const int b = idx / ((long long)nx * nx * (N + 1));
    double w = wx[b * nx + i] + (t == N ? wf[b * nx + i] : 0.0);
    Qinv[idx] = (i == j) ? 1.0 / fmax(2.0 * w, kPsdMinEig) : 0.0;
    if (j == 0) {
      const double ref = (t < N) ? xref[((size_t)b * (N + 1) + t) * nx + i] : xf[b * nx + i];
      qv[((size_t)b * (N + 1) + t) * nx + i] = -2.0 * w * ref;
    }
  } else if (idx < nq + nr) {
    const long long k2 = idx - nq;
    const int j = k2 % nu;
    const int i = (k2 / nu) % nu;
    const int t = (k2 / ((long long)nu * nu)) % N;
    const int b = k2 / ((long long)nu * nu * N);
    Rinv[k2] = (i == j) ? 1.0 / fmax(2.0 * wu[b * nu + i], kPsdMinEig) : 0.0;
    if (j == 0)
      rv[((size_t)b * N + t) * nu + i] = penalize ? -2.0 * wu[b * nu + i] * uref[((size_t)b * (N + 1) + t) * nu + i] : 0.0;
  }
}

// premultiplied products (sqp.py:247-251): AQ = A_t Qinv_t, AnQ = An_t Qinv_{t+1}, BR = B_t Rinv_t
__global__ void premult_kernel(int B, int N, int nx, int nu, const double* As_next, const double* As, const double* Bs,
                               const double* Qinv, const double* Rinv, double* AQ, double* AnQ, double* BR) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n1 = (long long)B * N * nx * nx;
  if (idx < n1) {
    const int j = idx % nx;
    const int i = (idx / nx) % nx;
    const long long bt = idx / ((long long)nx * nx);
    const int t = bt % N;
    const int b = bt / N;
    const double* Qt = Qinv + ((size_t)b * (N + 1) + t) * nx * nx;
    const double* Qn = Qt + nx * nx;
    double a = 0.0, c = 0.0;
    for (int k = 0; k < nx; ++k) {
      a += As[(bt * nx + i) * nx + k] * Qt[k * nx + j];
      c += As_next[(bt * nx + i) * nx + k] * Qn[k * nx + j];
    }
    AQ[idx] = a;
    AnQ[idx] = c;
  } else if (idx < n1 + (long long)B * N * nx * nu) {
    const long long k2 = idx - n1;
    const int j = k2 % nu;
    const int i = (k2 / nu) % nx;
    const long long bt = k2 / ((long long)nx * nu);
    double a = 0.0;
    for (int k = 0; k < nu; ++k) a += Bs[(bt * nx + i) * nu + k] * Rinv[(bt * nu + k) * nu + j];
    BR[k2] = a;
  }
}

// theta_t and phi_t (pcg.py:182-195): theta_0 = Qinv_0,
// theta_{t+1} = AQ A^T + BR B^T + AnQ An^T, phi_{t+1} = AQ An_{t-1}^T (An_{-1} = I)
__global__ void theta_kernel(int B, int N, int nx, int nu, const double* As_next, const double* As, const double* Bs,
                             const double* Qinv, const double* AQ, const double* AnQ, const double* BR, double* theta,
                             double* phi) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)B * (N + 1) * nx * nx) return;
  const int j = idx % nx;
  const int i = (idx / nx) % nx;
  const int t = (idx / ((long long)nx * nx)) % (N + 1);
  const int b = idx / ((long long)nx * nx * (N + 1));
  if (t == 0) {
    theta[idx] = Qinv[((size_t)b * (N + 1)) * nx * nx + i * nx + j];
    phi[idx] = 0.0;
    return;
  }
  const size_t bt = (size_t)b * N + (t - 1);
  double th = 0.0, ph = 0.0;
  for (int k = 0; k < nx; ++k) {
    th += AQ[(bt * nx + i) * nx + k] * As[(bt * nx + j) * nx + k];
    th += AnQ[(bt * nx + i) * nx + k] * As_next[(bt * nx + j) * nx + k];
    const double anp = (t - 1 == 0) ? (j == k ? 1.0 : 0.0) : As_next[((bt - 1) * nx + j) * nx + k];
    ph += AQ[(bt * nx + i) * nx + k] * anp;
  }
  for (int k = 0; k < nu; ++k) th += BR[(bt * nx + i) * nu + k] * Bs[(bt * nx + j) * nu + k];
  theta[idx] = th;
  phi[idx] = ph;  // phi_t; phi_0 = 0
}

// in-place Gauss-Jordan inverse of each SPD block (one thread per block)
__global__ void inv_kernel(long long nblocks, int n, double* M, double* W) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nblocks) return;
  double* A = M + idx * n * n;
  double* I = W + idx * n * n;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) I[i * n + j] = (i == j) ? 1.0 : 0.0;
  for (int k = 0; k < n; ++k) {
    const double pinv = 1.0 / A[k * n + k];
    for (int j = 0; j < n; ++j) {
      A[k * n + j] *= pinv;
      I[k * n + j] *= pinv;
    }
    for (int i = 0; i < n; ++i) {
      if (i == k) continue;
      const double f = A[i * n + k];
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) {
        A[i * n + j] -= f * A[k * n + j];
        I[i * n + j] -= f * I[k * n + j];
      }
    }
  }
}

// M1 = theta_t^-1 phi_t, M2 = theta_t^-1 phi_{t+1}^T (phi_{N+1} = 0)
__global__ void m12_kernel(int B, int N, int nx, const double* phi, const double* thinv, double* M1, double* M2) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)B * (N + 1) * nx * nx) return;
  const int j = idx % nx;
  const int i = (idx / nx) % nx;
  const long long blk = idx / ((long long)nx * nx);
  const int t = blk % (N + 1);
  const size_t sq = (size_t)nx * nx;
  double m1 = 0.0, m2 = 0.0;
  for (int k = 0; k < nx; ++k) {
    m1 += thinv[blk * sq + i * nx + k] * phi[blk * sq + k * nx + j];
    if (t < N) m2 += thinv[blk * sq + i * nx + k] * phi[(blk + 1) * sq + j * nx + k];
  }
  M1[idx] = m1;
  M2[idx] = m2;
}

// S and Phiinv rows (pcg.py:198-239); thinv = theta^-1 (B, N+1, nx, nx)
__global__ void schur_kernel(int B, int N, int nx, const double* theta, const double* phi, const double* thinv,
                             const double* M1, const double* M2, double* S, double* Phiinv) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int W = 3 * nx;
  if (idx >= (long long)B * (N + 1) * nx * W) return;
  const int c = idx % W;
  const int i = (idx / W) % nx;
  const int t = (idx / ((long long)W * nx)) % (N + 1);
  const int b = idx / ((long long)W * nx * (N + 1));
  const size_t blk = (size_t)b * (N + 1) + t;
  const int part = c / nx, j = c % nx;
  const size_t sq = (size_t)nx * nx;
  double s = 0.0, p = 0.0;
  if (part == 0) {
    if (t >= 1) {
      s = -phi[blk * sq + i * nx + j];
      // -theta_t^-1 phi_t theta_{t-1}^-1
      for (int a = 0; a < nx; ++a) p -= M1[blk * sq + i * nx + a] * thinv[(blk - 1) * sq + a * nx + j];
    }
  } else if (part == 1) {
    s = -theta[blk * sq + i * nx + j];
    p = thinv[blk * sq + i * nx + j];
  } else {
    if (t < N) {
      s = -phi[(blk + 1) * sq + j * nx + i];
      // -theta_t^-1 phi_{t+1}^T theta_{t+1}^-1
      for (int a = 0; a < nx; ++a) p -= M2[blk * sq + i * nx + a] * thinv[(blk + 1) * sq + a * nx + j];
    }
  }
  S[idx] = s;
  Phiinv[idx] = p;
}

// gamma (pcg.py:275-288)
__global__ void gamma_kernel(int B, int N, int nx, int nu, const double* AQ, const double* AnQ, const double* BR,
                             const double* Cs, const double* Qinv, const double* qs, const double* rs, double* gamma) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)B * (N + 1) * nx) return;
  const int i = idx % nx;
  const int t = (idx / nx) % (N + 1);
  const int b = idx / ((long long)nx * (N + 1));
  const double* q = qs + (size_t)b * (N + 1) * nx;
  double g = Cs[idx];
  if (t == 0) {
    const double* Q0 = Qinv + (size_t)b * (N + 1) * nx * nx;
    for (int k = 0; k < nx; ++k) g += Q0[i * nx + k] * q[k];
  } else {
    const size_t bt = (size_t)b * N + (t - 1);
    for (int k = 0; k < nx; ++k) {
      g += AQ[(bt * nx + i) * nx + k] * q[(t - 1) * nx + k];
      g += AnQ[(bt * nx + i) * nx + k] * q[t * nx + k];
    }
    for (int k = 0; k < nu; ++k) g += BR[(bt * nx + i) * nu + k] * rs[bt * nu + k];
  }
  gamma[idx] = g;
}

// y_t = M_t [v_{t-1}; v_t; v_{t+1}] (pcg.py:370-392, 405-428).
// One workgroup per block row (b, t): the 3nx-entry vector window is staged
// in LDS, each row of the block is read by 16 consecutive lanes (128-byte
// coalesced segments) and reduced with shuffles.  Columns outside 0..N are
// skipped.  Launch with matvec_threads(nx) threads, B*(N+1) workgroups.
constexpr int kMatvecMaxNx = 64;
inline int matvec_threads(int nx) { return nx * 16 >= 256 ? 256 : ((nx * 16 + 63) / 64) * 64; }

__global__ void matvec_kernel(int B, int N, int nx, const double* M, const double* v, double* y, const int* active) {
  __shared__ double vs[3 * kMatvecMaxNx];
  const long long bt = blockIdx.x;
  const int b = (int)(bt / (N + 1)), t = (int)(bt - (long long)b * (N + 1));
  if (active && !active[b]) return;
  const int w = 3 * nx;
  const double* vb = v + (size_t)b * (N + 1) * nx;
  for (int c = threadIdx.x; c < w; c += blockDim.x) {
    const int tt = t - 1 + c / nx;
    vs[c] = (tt >= 0 && tt <= N) ? vb[tt * nx + c % nx] : 0.0;
  }
  __syncthreads();
  const int c_lo = t == 0 ? nx : 0, c_hi = t == N ? 2 * nx : w;
  const int g = threadIdx.x >> 4, j = threadIdx.x & 15, ng = blockDim.x >> 4;
  for (int i = g; i < nx; i += ng) {
    const double* row = M + ((size_t)bt * nx + i) * w;
    double acc = 0.0;
    for (int c = c_lo + j; c < c_hi; c += 16) acc += row[c] * vs[c];
    acc += __shfl_xor(acc, 8, 16);
    acc += __shfl_xor(acc, 4, 16);
    acc += __shfl_xor(acc, 2, 16);
    acc += __shfl_xor(acc, 1, 16);
    if (j == 0) y[(size_t)bt * nx + i] = acc;
  }
}

// per-instance dot product, one workgroup per instance, fixed order
__global__ void dot_kernel(int n, const double* a, const double* b, double* out, const int* active) {
  __shared__ double red[256];
  const int inst = blockIdx.x;
  if (active && !active[inst]) return;
  double acc = 0.0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) acc += a[(size_t)inst * n + k] * b[(size_t)inst * n + k];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[inst] = red[0];
}

__global__ void sub_kernel(long long n, const double* a, const double* b, double* out) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) out[idx] = a[idx] - b[idx];
}

__global__ void copy_kernel(long long n, const double* a, double* out) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) out[idx] = a[idx];
}

// active[b] = (|eta| > eps && it <= max_iter - 1) || it < 1 (pcg.py:397-403)
__global__ void pcg_active_kernel(int B, const double* eta, const int* it, double eps, int max_iter, int* active,
                                  int* any) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int a = ((fabs(eta[b]) > eps && it[b] <= max_iter - 1) || it[b] < 1) ? 1 : 0;
  active[b] = a;
  if (a) atomicOr(any, 1);
}

// lam += alpha p; r -= alpha ups, alpha = eta / v
__global__ void pcg_step_kernel(int B, int n, const double* eta, const double* v, const double* p, const double* ups,
                                double* lam, double* r, const int* active) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)B * n) return;
  const int b = idx / n;
  if (!active[b]) return;
  const double alpha = eta[b] / v[b];
  lam[idx] += alpha * p[idx];
  r[idx] -= alpha * ups[idx];
}

// p = rt + beta p with beta = eta' / eta; then eta = eta', it += 1
__global__ void pcg_dir_kernel(int B, int n, const double* etan, const double* eta, const double* rt, double* p,
                               const int* active) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)B * n) return;
  const int b = idx / n;
  if (!active[b]) return;
  const double beta = etan[b] / eta[b];
  p[idx] = rt[idx] + beta * p[idx];
}

__global__ void pcg_scalar_kernel(int B, double* eta, const double* etan, int* it, const int* active) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B || !active[b]) return;
  eta[b] = etan[b];
  it[b] += 1;
}

// x, u from lam (pcg.py:316-348)
__global__ void recover_kernel(int B, int N, int nx, int nu, const double* lam, const double* As_next, const double* As,
                               const double* Bs, const double* Qinv, const double* qs, const double* Rinv,
                               const double* rs, double* x, double* u) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nxs = (long long)B * (N + 1) * nx;
  const long long nus = (long long)B * (N + 1) * nu;
  if (idx < nxs) {
    const int i = idx % nx;
    const int t = (idx / nx) % (N + 1);
    const int b = idx / ((long long)nx * (N + 1));
    const double* lb = lam + (size_t)b * (N + 1) * nx;
    double w[64];
    for (int k = 0; k < nx; ++k) {
      double acc = qs[((size_t)b * (N + 1) + t) * nx + k];
      // An_{t-1}^T lam_t (An_{-1} = I)
      if (t == 0) {
        acc += lb[k];
      } else {
        const size_t bp = (size_t)b * N + (t - 1);
        for (int m = 0; m < nx; ++m) acc += As_next[(bp * nx + m) * nx + k] * lb[t * nx + m];
      }
      if (t < N) {
        const size_t bt = (size_t)b * N + t;
        for (int m = 0; m < nx; ++m) acc += As[(bt * nx + m) * nx + k] * lb[(t + 1) * nx + m];
      }
      w[k] = acc;
    }
    const double* Q = Qinv + ((size_t)b * (N + 1) + t) * nx * nx;
    double s = 0.0;
    for (int k = 0; k < nx; ++k) s += Q[i * nx + k] * w[k];
    x[idx] = -s;
  } else if (idx < nxs + nus) {
    const long long k2 = idx - nxs;
    const int i = k2 % nu;
    const int t = (k2 / nu) % (N + 1);
    const int b = k2 / ((long long)nu * (N + 1));
    if (t == N) {
      u[k2] = 0.0;
      return;
    }
    const size_t bt = (size_t)b * N + t;
    const double* lb = lam + ((size_t)b * (N + 1) + t + 1) * nx;
    double s = 0.0;
    for (int k = 0; k < nu; ++k) {
      double acc = rs[bt * nu + k];
      for (int m = 0; m < nx; ++m) acc += Bs[(bt * nx + m) * nu + k] * lb[m];
      s += Rinv[(bt * nu + i) * nu + k] * acc;
    }
    u[k2] = -s;
  }
}

// per-instance SQP step bookkeeping (sqp.py:506-528): err = max|dx|,|du|;
// accept the new iterate for active instances
// Merit line search of sqp.py:301-411, per-launch form.  ls_items_kernel:
// one thread per (instance, candidate kc, knot t); kc = 0 is the current
// iterate, kc >= 1 the step alphas[kc - 1].  It writes the cost of knot t and
// the l1 norm of constraint block t (g_0 = x_0 - x_init,
// g_t = f_d(x_{t-1}, u_{t-1}) - x_t), dynamics evaluated in float64.
template <class Model>
__global__ void ls_items_kernel(SolveArgs a, const double* xn, const double* un, double* lsC, double* lsG,
                                const int* sqp_active) {
  constexpr int NX = Model::NX, NU = Model::NU;
  const int N = a.N, NB = N + 1, nal = a.n_alphas;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)a.B * (nal + 1) * NB) return;
  const int b = (int)(idx / ((nal + 1) * NB));
  const int rem = (int)(idx - (long long)b * (nal + 1) * NB);
  const int kc = rem / NB, t = rem - kc * NB;
  if (!sqp_active[b]) return;
  const double al = kc == 0 ? 0.0 : a.alphas[kc - 1];
  const double* x = a.x + (size_t)b * NB * NX;
  const double* u = a.u + (size_t)b * NB * NU;
  const double* xq = xn + (size_t)b * NB * NX;
  const double* uq = un + (size_t)b * NB * NU;
  const double* wx = a.wx + (size_t)b * NX;
  const double* wu = a.wu + (size_t)b * NU;
  const double* wf = a.wf + (size_t)b * NX;
  double c = 0.0, g = 0.0;
  if (t < N) {
    for (int i = 0; i < NX; ++i) {
      const double d = x[t * NX + i] + al * (xq[t * NX + i] - x[t * NX + i]) - a.xref[((size_t)b * NB + t) * NX + i];
      c += wx[i] * d * d;
    }
    for (int i = 0; i < NU; ++i) {
      const double ref = a.penalize ? a.uref[((size_t)b * NB + t) * NU + i] : 0.0;
      const double d = u[t * NU + i] + al * (uq[t * NU + i] - u[t * NU + i]) - ref;
      c += wu[i] * d * d;
    }
  } else {
    for (int i = 0; i < NX; ++i) {
      const double d = x[t * NX + i] + al * (xq[t * NX + i] - x[t * NX + i]) - a.xf[(size_t)b * NX + i];
      c += (wx[i] + wf[i]) * d * d;
    }
  }
  if (t == 0) {
    for (int i = 0; i < NX; ++i) g += fabs(x[i] + al * (xq[i] - x[i]) - a.x0[(size_t)b * NX + i]);
  } else {
    double xp[NX], up[NU], fx[NX];
    for (int i = 0; i < NX; ++i) xp[i] = x[(t - 1) * NX + i] + al * (xq[(t - 1) * NX + i] - x[(t - 1) * NX + i]);
    for (int i = 0; i < NU; ++i) up[i] = u[(t - 1) * NU + i] + al * (uq[(t - 1) * NU + i] - u[(t - 1) * NU + i]);
    step<Model, double>(xp, up, a.dyn + (size_t)b * a.P, a.dt64, a.scheme, fx);
    for (int i = 0; i < NX; ++i) g += fabs(fx[i] - (x[t * NX + i] + al * (xq[t * NX + i] - x[t * NX + i])));
  }
  lsC[(size_t)b * (kMaxAlphas + 1) * NB + rem] = c;
  lsG[(size_t)b * (kMaxAlphas + 1) * NB + rem] = g;
}

// One block per instance: mu, merit, decreases and the reference's clamped
// scan over the candidates (thread 0, fixed summation order), then the step
// x + alpha (x_qp - x) is applied to the QP solution in place.
__global__ void ls_select_kernel(SolveArgs a, int nx, int nu, double* xn, double* un, const double* lsC,
                                 const double* lsG, double* alpha_out, const int* sqp_active) {
  __shared__ double alpha_s;
  const int b = blockIdx.x;
  if (!sqp_active[b]) return;
  const int N = a.N, NB = N + 1, nal = a.n_alphas;
  const double* x = a.x + (size_t)b * NB * nx;
  const double* u = a.u + (size_t)b * NB * nu;
  double* xq = xn + (size_t)b * NB * nx;
  double* uq = un + (size_t)b * NB * nu;
  if (threadIdx.x == 0) {
    double dcdz = 0.0;
    for (int k = 0; k < NB * nx; ++k) {
      const int t = k / nx, i = k - t * nx;
      const double wxi = a.wx[(size_t)b * nx + i];
      const double gx = (t < N) ? 2.0 * wxi * (x[k] - a.xref[(size_t)b * NB * nx + k])
                                : 2.0 * (wxi + a.wf[(size_t)b * nx + i]) * (x[k] - a.xf[(size_t)b * nx + i]);
      dcdz += gx * (xq[k] - x[k]);
    }
    for (int k = 0; k < N * nu; ++k) {
      const int j = k % nu;
      const double ref = a.penalize ? a.uref[(size_t)b * NB * nu + k] : 0.0;
      dcdz += 2.0 * a.wu[(size_t)b * nu + j] * (u[k] - ref) * (uq[k] - u[k]);
    }
    double sc[kMaxAlphas + 1], sg[kMaxAlphas + 1];
    for (int kc = 0; kc <= nal; ++kc) {
      double c = 0.0, g = 0.0;
      for (int t = 0; t < NB; ++t) {
        c += lsC[(size_t)b * (kMaxAlphas + 1) * NB + kc * NB + t];
        g += lsG[(size_t)b * (kMaxAlphas + 1) * NB + kc * NB + t];
      }
      sc[kc] = c;
      sg[kc] = g;
    }
    const double mu_min = dcdz / (0.5 * sg[0]);
    const double mu = (mu_min != mu_min) ? mu_min : fmax(0.1, mu_min);
    const double merit = sc[0] + mu * sg[0];
    const double dmerit = dcdz + -mu * sg[0];
    double amax = a.alphas[0];
    for (int k = 1; k < nal; ++k) amax = (a.alphas[k] != a.alphas[k] || amax != amax) ? NAN : fmax(amax, a.alphas[k]);
    int i = nal + 1;
    double alpha = amax;
    while (true) {
      const int ic = i < nal - 1 ? i : nal - 1;
      const double dec = (sc[ic + 1] + mu * sg[ic + 1]) - (merit + a.ls_eta * a.alphas[ic] * dmerit);
      if (dec < 0.0 || i == 0) break;
      alpha = a.alphas[(i - 1) < nal - 1 ? (i - 1) : nal - 1];
      --i;
    }
    alpha_s = alpha;
    alpha_out[b] = alpha;
  }
  __syncthreads();
  const double alpha = alpha_s;
  for (int k = threadIdx.x; k < NB * nx; k += blockDim.x) xq[k] = x[k] + alpha * (xq[k] - x[k]);
  for (int k = threadIdx.x; k < NB * nu; k += blockDim.x) uq[k] = u[k] + alpha * (uq[k] - u[k]);
}

__global__ void sqp_accept_kernel(int B, int N, int nx, int nu, int maxit, int sqp_it, const int* sqp_active,
                                  double* x, double* u, double* lam, const double* xn, const double* un,
                                  const double* lamn, double* conv_err, double* conv_errs, int32_t* pcg_hist,
                                  const int* pcg_it, double* ls_hist, int linesearch, int ls_eval,
                                  const double* alpha, int32_t* num_iter) {
  __shared__ double red[256];
  const int b = blockIdx.x;
  if (!sqp_active[b]) return;
  const int nxs = (N + 1) * nx, nus = (N + 1) * nu;
  double e = 0.0;
  for (int k = threadIdx.x; k < nxs; k += blockDim.x) {
    const double d = fabs(xn[(size_t)b * nxs + k] - x[(size_t)b * nxs + k]);
    e = (d != d || e != e) ? d + e : fmax(e, d);
  }
  for (int k = threadIdx.x; k < nus; k += blockDim.x) {
    const double d = fabs(un[(size_t)b * nus + k] - u[(size_t)b * nus + k]);
    e = (d != d || e != e) ? d + e : fmax(e, d);
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const double a = red[threadIdx.x], c = red[threadIdx.x + s];
      red[threadIdx.x] = (a != a || c != c) ? a + c : fmax(a, c);
    }
    __syncthreads();
  }
  const double err = red[0];
  for (int k = threadIdx.x; k < nxs; k += blockDim.x) {
    x[(size_t)b * nxs + k] = xn[(size_t)b * nxs + k];
    lam[(size_t)b * nxs + k] = lamn[(size_t)b * nxs + k];
  }
  for (int k = threadIdx.x; k < nus; k += blockDim.x) u[(size_t)b * nus + k] = un[(size_t)b * nus + k];
  if (threadIdx.x == 0) {
    conv_err[b] = err;
    conv_errs[(size_t)b * maxit + sqp_it] = err;
    pcg_hist[(size_t)b * maxit + sqp_it] = pcg_it[b];
    if (linesearch) ls_hist[(size_t)b * maxit + sqp_it] = ls_eval ? alpha[b] : 1.0;
    num_iter[b] = sqp_it + 1;
  }
}

// SQP continuation: (num_iter < maxit && err > tol) || num_iter < 1
__global__ void sqp_active_kernel(int B, int maxit, double tol, const int32_t* num_iter, const double* err,
                                  int* active, int* any) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const int a = ((num_iter[b] < maxit && err[b] > tol) || num_iter[b] < 1) ? 1 : 0;
  active[b] = a;
  if (a) atomicOr(any, 1);
}

// status and nan_to_num (sqp.py:535-559)
__global__ void finish_kernel(int B, int N, int nx, int nu, double tol, double* x, double* u, double* lam,
                              const double* err, int32_t* status) {
  __shared__ int bad;
  const int b = blockIdx.x;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  const int nxs = (N + 1) * nx, nus = (N + 1) * nu;
  int mybad = 0;
  for (int k = threadIdx.x; k < nxs; k += blockDim.x) {
    const double v = x[(size_t)b * nxs + k];
    if (v != v) mybad = 1;
  }
  for (int k = threadIdx.x; k < nus; k += blockDim.x) {
    const double v = u[(size_t)b * nus + k];
    if (v != v) mybad = 1;
  }
  if (mybad) atomicOr(&bad, 1);
  __syncthreads();
  auto fix = [](double v) {
    if (v != v) return 0.0;
    if (v > 1.7976931348623157e308) return 1.7976931348623157e308;
    if (v < -1.7976931348623157e308) return -1.7976931348623157e308;
    return v;
  };
  for (int k = threadIdx.x; k < nxs; k += blockDim.x) {
    x[(size_t)b * nxs + k] = fix(x[(size_t)b * nxs + k]);
    lam[(size_t)b * nxs + k] = fix(lam[(size_t)b * nxs + k]);
  }
  for (int k = threadIdx.x; k < nus; k += blockDim.x) u[(size_t)b * nus + k] = fix(u[(size_t)b * nus + k]);
  if (threadIdx.x == 0) status[b] = (bad || err[b] > tol) ? -100 : 0;
}

// saved linearization for the backward: AB (float), diag Qinv, diag Rinv
__global__ void save_kernel(int B, int N, int nx, int nu, const double* As, const double* Bs, const double* Qinv,
                            const double* Rinv, const double* fd, const int* active, float* AB, double* fdo,
                            double* qinv, double* rinv) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nz = nx + nu;
  const long long nab = (long long)B * N * nx * nz;
  if (idx < nab) {
    const int k = idx % nz;
    const long long row = idx / nz;  // (b, t, i)
    if (active[row / ((long long)N * nx)]) AB[idx] = (float)(k < nx ? As[row * nx + k] : Bs[row * nu + (k - nx)]);
  }
  if (idx < (long long)B * (N + 1) * nx) {
    const long long blk = idx / nx;
    const int i = idx % nx;
    if (active[blk / (N + 1)]) qinv[idx] = Qinv[blk * nx * nx + i * nx + i];
  }
  if (idx < (long long)B * N * nx) {
    if (active[idx / ((long long)N * nx)]) fdo[idx] = fd[idx];
  }
  if (idx < (long long)B * nu) {
    const int b = idx / nu, i = idx % nu;
    if (active[b]) rinv[idx] = Rinv[((size_t)b * N) * nu * nu + i * nu + i];
  }
}

// dense QP blocks from the saved linearization (backward, plain mode)
__global__ void unsave_kernel(int B, int N, int nx, int nu, const float* AB, const double* qinv, const double* rinv,
                              double* As_next, double* As, double* Bs, double* Qinv, double* Rinv) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nz = nx + nu;
  if (idx < (long long)B * N * nx * nz) {
    const int k = idx % nz;
    const long long row = idx / nz;
    if (k < nx) {
      As[row * nx + k] = AB[idx];
      As_next[row * nx + k] = (k == row % nx) ? -1.0 : 0.0;
    } else {
      Bs[row * nu + (k - nx)] = AB[idx];
    }
  }
  if (idx < (long long)B * (N + 1) * nx * nx) {
    const int j = idx % nx;
    const int i = (idx / nx) % nx;
    const long long blk = idx / ((long long)nx * nx);
    Qinv[idx] = (i == j) ? qinv[blk * nx + i] : 0.0;
  }
  if (idx < (long long)B * N * nu * nu) {
    const int j = idx % nu;
    const int i = (idx / nu) % nu;
    const int b = idx / ((long long)nu * nu * N);
    Rinv[idx] = (i == j) ? rinv[(size_t)b * nu + i] : 0.0;
  }
}

// backward contraction with the cost's mixed derivatives (sqp.py:605-666,
// 724-739), one workgroup per instance
__global__ void bwd_contract_kernel(int B, int N, int nx, int nu, int penalize, const double* x, const double* u,
                                    const double* xref, const double* uref, const double* xf, const double* wx,
                                    const double* wu, const double* wf, const double* xt, const double* ut,
                                    const double* lamt, double* g_wx, double* g_wu, double* g_wf, double* g_x0,
                                    double* g_xref, double* g_uref, double* g_xf) {
  const int b = blockIdx.x;
  const int NB = N + 1;
  for (int k = threadIdx.x; k < NB * nx; k += blockDim.x) {
    const int t = k / nx, i = k % nx;
    g_xref[(size_t)b * NB * nx + k] = (t < N) ? 2.0 * wx[b * nx + i] * xt[(size_t)b * NB * nx + k] : 0.0;
  }
  for (int k = threadIdx.x; k < NB * nu; k += blockDim.x) {
    const int t = k / nu, i = k % nu;
    g_uref[(size_t)b * NB * nu + k] = (t < N && penalize) ? 2.0 * wu[b * nu + i] * ut[(size_t)b * NB * nu + k] : 0.0;
  }
  for (int i = threadIdx.x; i < nx + nu; i += blockDim.x) {
    if (i < nx) {
      double acc = 0.0;
      for (int t = 0; t <= N; ++t) {
        const double ref = (t < N) ? xref[((size_t)b * NB + t) * nx + i] : xf[b * nx + i];
        acc += 2.0 * (x[((size_t)b * NB + t) * nx + i] - ref) * xt[((size_t)b * NB + t) * nx + i];
      }
      g_wx[b * nx + i] = -acc;
      const double eN = x[((size_t)b * NB + N) * nx + i] - xf[b * nx + i];
      const double xtN = xt[((size_t)b * NB + N) * nx + i];
      g_wf[b * nx + i] = -2.0 * eN * xtN;
      g_xf[b * nx + i] = 2.0 * (wx[b * nx + i] + wf[b * nx + i]) * xtN;
      g_x0[b * nx + i] = lamt[(size_t)b * NB * nx + i];
    } else {
      const int k = i - nx;
      double acc = 0.0;
      for (int t = 0; t < N; ++t) {
        const double ur = penalize ? uref[((size_t)b * NB + t) * nu + k] : 0.0;
        acc += 2.0 * (u[((size_t)b * NB + t) * nu + k] - ur) * ut[((size_t)b * NB + t) * nu + k];
      }
      g_wu[b * nu + k] = -acc;
    }
  }
}

__global__ void neg_kernel(int B, int N, int nu, int rows_in, const double* in, double* out, int drop_last) {
  // out (B, rows_out, nu) = -in (B, rows_in, nu) with rows_out = rows_in - drop_last
  const int rows_out = rows_in - drop_last;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)B * rows_out * nu) return;
  const int k = idx % nu;
  const int t = (idx / nu) % rows_out;
  const int b = idx / ((long long)nu * rows_out);
  out[idx] = -in[((size_t)b * rows_in + t) * nu + k];
}

// ---------------------------------------------------------------------------
// device scratch, grown on demand and kept for the process (per device)
struct Scratch {
  void* ptr = nullptr;
  size_t bytes = 0;
};
Scratch g_scratch[16];

int get_scratch(size_t bytes, void** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 16) return DMPC_ERR_HIP;
  Scratch& s = g_scratch[dev];
  if (s.bytes < bytes) {
    if (s.ptr) HIPCHK(hipFree(s.ptr));
    s.ptr = nullptr;
    s.bytes = 0;
    HIPCHK(hipMalloc(&s.ptr, bytes));
    s.bytes = bytes;
  }
  *out = s.ptr;
  return DMPC_OK;
}

struct Carver {
  char* p;
  template <class T>
  T* take(size_t n) {
    T* r = reinterpret_cast<T*>(p);
    p += ((n * sizeof(T) + 255) / 256) * 256;
    return r;
  }
};

struct QpBufs {
  double *As_next, *As, *Bs, *Cs, *fd, *Qinv, *qv, *Rinv, *rv, *AQ, *AnQ, *BR, *theta, *phi, *thinv, *tmp, *M2, *S, *Phi;
  double *gamma, *lam, *r, *rt, *p, *ups, *eta, *etan, *v, *xn, *un;
  double *lsC, *lsG, *alpha;  // line search: merit parts per (instance, candidate, knot), chosen step
  int *active, *iters, *any, *sqp_active;
};

size_t qp_bytes(int B, int N, int nx, int nu) {
  const size_t NB = N + 1;
  size_t d = 0;
  d += 2 * (size_t)B * N * nx * nx;            // As_next, As
  d += (size_t)B * N * nx * nu;                // Bs
  d += 2 * (size_t)B * NB * nx;                // Cs, qv
  d += (size_t)B * N * nx;                     // fd
  d += (size_t)B * NB * nx * nx;               // Qinv
  d += (size_t)B * N * nu * nu + (size_t)B * N * nu;  // Rinv, rv
  d += 2 * (size_t)B * N * nx * nx + (size_t)B * N * nx * nu;  // AQ, AnQ, BR
  d += 5 * (size_t)B * NB * nx * nx;           // theta, phi, thinv, tmp, M2
  d += 2 * (size_t)B * NB * nx * 3 * nx;       // S, Phi
  d += 6 * (size_t)B * NB * nx;                // gamma lam r rt p ups
  d += 3 * (size_t)B + 2 * (size_t)B * NB * nx + (size_t)B * NB * nu;  // eta etan v xn un(+)
  d += 2 * (size_t)B * (kMaxAlphas + 1) * NB + (size_t)B;  // lsC lsG alpha
  return d * 8 + 4 * (size_t)B * 4 + 64 * 256;
}

QpBufs carve_qp(void* base, int B, int N, int nx, int nu) {
  const size_t NB = N + 1;
  Carver c{reinterpret_cast<char*>(base)};
  QpBufs q;
  q.As_next = c.take<double>((size_t)B * N * nx * nx);
  q.As = c.take<double>((size_t)B * N * nx * nx);
  q.Bs = c.take<double>((size_t)B * N * nx * nu);
  q.Cs = c.take<double>((size_t)B * NB * nx);
  q.qv = c.take<double>((size_t)B * NB * nx);
  q.fd = c.take<double>((size_t)B * N * nx);
  q.Qinv = c.take<double>((size_t)B * NB * nx * nx);
  q.Rinv = c.take<double>((size_t)B * N * nu * nu);
  q.rv = c.take<double>((size_t)B * N * nu);
  q.AQ = c.take<double>((size_t)B * N * nx * nx);
  q.AnQ = c.take<double>((size_t)B * N * nx * nx);
  q.BR = c.take<double>((size_t)B * N * nx * nu);
  q.theta = c.take<double>((size_t)B * NB * nx * nx);
  q.phi = c.take<double>((size_t)B * NB * nx * nx);
  q.thinv = c.take<double>((size_t)B * NB * nx * nx);
  q.tmp = c.take<double>((size_t)B * NB * nx * nx);
  q.M2 = c.take<double>((size_t)B * NB * nx * nx);
  q.S = c.take<double>((size_t)B * NB * nx * 3 * nx);
  q.Phi = c.take<double>((size_t)B * NB * nx * 3 * nx);
  q.gamma = c.take<double>((size_t)B * NB * nx);
  q.lam = c.take<double>((size_t)B * NB * nx);
  q.r = c.take<double>((size_t)B * NB * nx);
  q.rt = c.take<double>((size_t)B * NB * nx);
  q.p = c.take<double>((size_t)B * NB * nx);
  q.ups = c.take<double>((size_t)B * NB * nx);
  q.eta = c.take<double>(B);
  q.etan = c.take<double>(B);
  q.v = c.take<double>(B);
  q.xn = c.take<double>((size_t)B * NB * nx);
  q.un = c.take<double>((size_t)B * NB * nu);
  q.lsC = c.take<double>((size_t)B * (kMaxAlphas + 1) * NB);
  q.lsG = c.take<double>((size_t)B * (kMaxAlphas + 1) * NB);
  q.alpha = c.take<double>(B);
  q.active = c.take<int>(B);
  q.iters = c.take<int>(B);
  q.sqp_active = c.take<int>(B);
  q.any = c.take<int>(64);
  return q;
}

// S, Phiinv from dense blocks (pcg.py:160-243), into q.S / q.Phi
int assemble(const QpBufs& q, int B, int N, int nx, int nu, hipStream_t st) {
  const long long n1 = (long long)B * N * nx * nx + (long long)B * N * nx * nu;
  hipLaunchKernelGGL(premult_kernel, dim3(nblk(n1)), dim3(256), 0, st, B, N, nx, nu, q.As_next, q.As, q.Bs, q.Qinv,
                     q.Rinv, q.AQ, q.AnQ, q.BR);
  const long long n2 = (long long)B * (N + 1) * nx * nx;
  hipLaunchKernelGGL(theta_kernel, dim3(nblk(n2)), dim3(256), 0, st, B, N, nx, nu, q.As_next, q.As, q.Bs, q.Qinv,
                     q.AQ, q.AnQ, q.BR, q.theta, q.phi);
  hipLaunchKernelGGL(copy_kernel, dim3(nblk(n2)), dim3(256), 0, st, n2, q.theta, q.tmp);
  hipLaunchKernelGGL(inv_kernel, dim3(nblk((long long)B * (N + 1), 64)), dim3(64), 0, st, (long long)B * (N + 1), nx,
                     q.tmp, q.thinv);
  hipLaunchKernelGGL(m12_kernel, dim3(nblk(n2)), dim3(256), 0, st, B, N, nx, q.phi, q.thinv, q.tmp, q.M2);
  const long long n3 = (long long)B * (N + 1) * nx * 3 * nx;
  hipLaunchKernelGGL(schur_kernel, dim3(nblk(n3)), dim3(256), 0, st, B, N, nx, q.theta, q.phi, q.thinv, q.tmp, q.M2,
                     q.S, q.Phi);
  HIPCHK(hipGetLastError());
  return DMPC_OK;
}

// host-driven PCG (pcg.py:350-436); q.lam holds the guess on entry
int pcg_loop(const QpBufs& q, int B, int N, int nx, int max_iter, double eps, hipStream_t st) {
  const int n = (N + 1) * nx;
  const long long nn = (long long)B * n;
  hipLaunchKernelGGL(matvec_kernel, dim3((unsigned)((long long)B * (N + 1))), dim3(matvec_threads(nx)), 0, st, B, N, nx, q.S, q.lam, q.ups, nullptr);
  hipLaunchKernelGGL(sub_kernel, dim3(nblk(nn)), dim3(256), 0, st, nn, q.gamma, q.ups, q.r);
  hipLaunchKernelGGL(matvec_kernel, dim3((unsigned)((long long)B * (N + 1))), dim3(matvec_threads(nx)), 0, st, B, N, nx, q.Phi, q.r, q.rt, nullptr);
  hipLaunchKernelGGL(copy_kernel, dim3(nblk(nn)), dim3(256), 0, st, nn, q.rt, q.p);
  hipLaunchKernelGGL(dot_kernel, dim3(B), dim3(256), 0, st, n, q.r, q.rt, q.eta, nullptr);
  HIPCHK(hipMemsetAsync(q.iters, 0, sizeof(int) * B, st));
  int host_any = 0;
  while (true) {
    HIPCHK(hipMemsetAsync(q.any, 0, sizeof(int), st));
    hipLaunchKernelGGL(pcg_active_kernel, dim3(nblk(B)), dim3(256), 0, st, B, q.eta, q.iters, eps, max_iter, q.active,
                       q.any);
    HIPCHK(hipMemcpyAsync(&host_any, q.any, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (!host_any) break;
    hipLaunchKernelGGL(matvec_kernel, dim3((unsigned)((long long)B * (N + 1))), dim3(matvec_threads(nx)), 0, st, B, N, nx, q.S, q.p, q.ups, q.active);
    hipLaunchKernelGGL(dot_kernel, dim3(B), dim3(256), 0, st, n, q.p, q.ups, q.v, q.active);
    hipLaunchKernelGGL(pcg_step_kernel, dim3(nblk(nn)), dim3(256), 0, st, B, n, q.eta, q.v, q.p, q.ups, q.lam, q.r,
                       q.active);
    hipLaunchKernelGGL(matvec_kernel, dim3((unsigned)((long long)B * (N + 1))), dim3(matvec_threads(nx)), 0, st, B, N, nx, q.Phi, q.r, q.rt, q.active);
    hipLaunchKernelGGL(dot_kernel, dim3(B), dim3(256), 0, st, n, q.r, q.rt, q.etan, q.active);
    hipLaunchKernelGGL(pcg_dir_kernel, dim3(nblk(nn)), dim3(256), 0, st, B, n, q.etan, q.eta, q.rt, q.p, q.active);
    hipLaunchKernelGGL(pcg_scalar_kernel, dim3(nblk(B)), dim3(256), 0, st, B, q.eta, q.etan, q.iters, q.active);
    HIPCHK(hipGetLastError());
  }
  return DMPC_OK;
}

template <class Model>
int plain_forward_t(const SolveArgs& a, hipStream_t st) {
  constexpr int nx = Model::NX, nu = Model::NU;
  const int B = a.B, N = a.N;
  void* base = nullptr;
  int rc = get_scratch(qp_bytes(B, N, nx, nu), &base);
  if (rc) return rc;
  QpBufs q = carve_qp(base, B, N, nx, nu);
  const long long nst = (long long)B * (N + 1) * nx;
  // cost blocks do not depend on the iterate
  hipLaunchKernelGGL(cost_kernel, dim3(nblk((long long)B * (N + 1) * nx * nx + (long long)B * N * nu * nu)), dim3(256),
                     0, st, B, N, nx, nu, a.penalize, a.wx, a.wu, a.wf, a.xref, a.uref, a.xf, q.Qinv, q.qv, q.Rinv,
                     q.rv);
  HIPCHK(hipMemsetAsync(a.num_iter, 0, sizeof(int32_t) * B, st));
  int sqp_it = 0;
  int host_any = 0;
  while (true) {
    HIPCHK(hipMemsetAsync(q.any, 0, sizeof(int), st));
    hipLaunchKernelGGL(sqp_active_kernel, dim3(nblk(B)), dim3(256), 0, st, B, a.maxit, a.tol, a.num_iter, a.conv_err,
                       q.sqp_active, q.any);
    HIPCHK(hipMemcpyAsync(&host_any, q.any, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (!host_any || sqp_it >= a.maxit) break;
    // linearize (optimal_control_problem.py:325-398)
    hipLaunchKernelGGL(lin_kernel<Model>, dim3(nblk((long long)B * N * (nx + nu))), dim3(256), 0, st, B, N, a.P, a.dt,
                       a.dt64, a.scheme, a.dyn, a.x, a.u, q.fd, q.As, q.Bs);
    hipLaunchKernelGGL(c_kernel, dim3(nblk(nst)), dim3(256), 0, st, B, N, nx, nu, a.x0, a.x, a.u, q.fd, q.As, q.Bs,
                       q.Cs, q.As_next);
    rc = assemble(q, B, N, nx, nu, st);
    if (rc) return rc;
    hipLaunchKernelGGL(gamma_kernel, dim3(nblk(nst)), dim3(256), 0, st, B, N, nx, nu, q.AQ, q.AnQ, q.BR, q.Cs, q.Qinv,
                       q.qv, q.rv, q.gamma);
    hipLaunchKernelGGL(copy_kernel, dim3(nblk(nst)), dim3(256), 0, st, nst, a.lam, q.lam);
    rc = pcg_loop(q, B, N, nx, a.pcg_max, a.pcg_eps, st);
    if (rc) return rc;
    hipLaunchKernelGGL(recover_kernel, dim3(nblk(nst + (long long)B * (N + 1) * nu)), dim3(256), 0, st, B, N, nx, nu,
                       q.lam, q.As_next, q.As, q.Bs, q.Qinv, q.qv, q.Rinv, q.rv, q.xn, q.un);
    if (a.ls_eval) {
      const long long nit = (long long)B * (a.n_alphas + 1) * (N + 1);
      hipLaunchKernelGGL(ls_items_kernel<Model>, dim3(nblk(nit)), dim3(256), 0, st, a, q.xn, q.un, q.lsC, q.lsG,
                         q.sqp_active);
      hipLaunchKernelGGL(ls_select_kernel, dim3(B), dim3(256), 0, st, a, nx, nu, q.xn, q.un, q.lsC, q.lsG, q.alpha,
                         q.sqp_active);
    }
    hipLaunchKernelGGL(save_kernel, dim3(nblk((long long)B * N * nx * (nx + nu) + nst)), dim3(256), 0, st, B, N, nx,
                       nu, q.As, q.Bs, q.Qinv, q.Rinv, q.fd, q.sqp_active, a.AB, a.fd, a.qinv, a.rinv);
    hipLaunchKernelGGL(sqp_accept_kernel, dim3(B), dim3(256), 0, st, B, N, nx, nu, a.maxit, sqp_it, q.sqp_active, a.x,
                       a.u, a.lam, q.xn, q.un, q.lam, a.conv_err, a.conv_errs, a.pcg_iters, q.iters, a.ls_alphas,
                       a.linesearch, a.ls_eval, q.alpha, a.num_iter);
    HIPCHK(hipGetLastError());
    ++sqp_it;
  }
  hipLaunchKernelGGL(finish_kernel, dim3(B), dim3(256), 0, st, B, N, nx, nu, a.tol, a.x, a.u, a.lam, a.conv_err,
                     a.status);
  HIPCHK(hipGetLastError());
  return DMPC_OK;
}

}  // namespace

int plain_forward(int model, int nx, int nu, const SolveArgs& a, hipStream_t st) {
  switch (model) {
    case MODEL_LINEAR:
      if (nx == 4 && nu == 2) return plain_forward_t<LinearModel<4, 2>>(a, st);
      if (nx == 8 && nu == 4) return plain_forward_t<LinearModel<8, 4>>(a, st);
      return DMPC_ERR_UNSUPPORTED;
    case MODEL_QUADROTOR: return plain_forward_t<QuadrotorModel>(a, st);
    case MODEL_SPACECRAFT3: return plain_forward_t<Spacecraft3Model>(a, st);
    case MODEL_SPACECRAFT7: return plain_forward_t<Spacecraft7Model>(a, st);
    case MODEL_CARTPOLE: return plain_forward_t<CartpoleModel>(a, st);
    case MODEL_CHAIN: return plain_forward_t<ChainModel>(a, st);
    DMPC_CUSTOM_CASE(return plain_forward_t<CustomModel>(a, st))
    default: return DMPC_ERR_UNSUPPORTED;
  }
}

int plain_backward(int model, int nx, int nu, const BackwardArgs& a, hipStream_t st) {
  (void)model;
  const int B = a.B, N = a.N;
  void* base = nullptr;
  int rc = get_scratch(qp_bytes(B, N, nx, nu), &base);
  if (rc) return rc;
  QpBufs q = carve_qp(base, B, N, nx, nu);
  const long long nst = (long long)B * (N + 1) * nx;
  const long long big = (long long)B * (N + 1) * nx * (nx + nu + nx);
  hipLaunchKernelGGL(unsave_kernel, dim3(nblk(big)), dim3(256), 0, st, B, N, nx, nu, a.AB, a.qinv, a.rinv, q.As_next,
                     q.As, q.Bs, q.Qinv, q.Rinv);
  rc = assemble(q, B, N, nx, nu, st);
  if (rc) return rc;
  // q~ = -dl/dx (B,N+1,nx); r~ = -dl/du[:-1] (B,N,nu); C = 0
  hipLaunchKernelGGL(neg_kernel, dim3(nblk(nst)), dim3(256), 0, st, B, N, nx, N + 1, a.dl_dx, q.qv, 0);
  hipLaunchKernelGGL(neg_kernel, dim3(nblk((long long)B * N * nu)), dim3(256), 0, st, B, N, nu, N + 1, a.dl_du, q.rv,
                     1);
  HIPCHK(hipMemsetAsync(q.Cs, 0, sizeof(double) * nst, st));
  hipLaunchKernelGGL(gamma_kernel, dim3(nblk(nst)), dim3(256), 0, st, B, N, nx, nu, q.AQ, q.AnQ, q.BR, q.Cs, q.Qinv,
                     q.qv, q.rv, q.gamma);
  hipLaunchKernelGGL(copy_kernel, dim3(nblk(nst)), dim3(256), 0, st, nst, a.lam_bwd, q.lam);
  rc = pcg_loop(q, B, N, nx, a.pcg_max, a.pcg_eps, st);
  if (rc) return rc;
  hipLaunchKernelGGL(recover_kernel, dim3(nblk(nst + (long long)B * (N + 1) * nu)), dim3(256), 0, st, B, N, nx, nu,
                     q.lam, q.As_next, q.As, q.Bs, q.Qinv, q.qv, q.Rinv, q.rv, q.xn, q.un);
  hipLaunchKernelGGL(bwd_contract_kernel, dim3(B), dim3(64), 0, st, B, N, nx, nu, a.penalize, a.x, a.u, a.xref,
                     a.uref, a.xf, a.wx, a.wu, a.wf, q.xn, q.un, q.lam, a.g_wx, a.g_wu, a.g_wf, a.g_x0, a.g_xref,
                     a.g_uref, a.g_xf);
  hipLaunchKernelGGL(copy_kernel, dim3(nblk(nst)), dim3(256), 0, st, nst, q.lam, a.lam_bwd);
  if (a.adj_x) hipLaunchKernelGGL(copy_kernel, dim3(nblk(nst)), dim3(256), 0, st, nst, q.xn, a.adj_x);
  if (a.adj_u) {
    const long long nu_all = (long long)B * (N + 1) * nu;
    hipLaunchKernelGGL(copy_kernel, dim3(nblk(nu_all)), dim3(256), 0, st, nu_all, q.un, a.adj_u);
  }
  HIPCHK(hipMemcpyAsync(a.pcg_iters, q.iters, sizeof(int) * B, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipGetLastError());
  return DMPC_OK;
}

// ---------------------------------------------------------------------------
// PCGOptimalControl-level entry points (dense, general blocks)
int pcg_layer_assemble(int B, int N, int nx, int nu, const double* As_next, const double* As, const double* Bs,
                       const double* Qinv, const double* Rinv, double* S, double* Phiinv, hipStream_t st) {
  void* base = nullptr;
  int rc = get_scratch(qp_bytes(B, N, nx, nu), &base);
  if (rc) return rc;
  QpBufs q = carve_qp(base, B, N, nx, nu);
  q.As_next = const_cast<double*>(As_next);
  q.As = const_cast<double*>(As);
  q.Bs = const_cast<double*>(Bs);
  q.Qinv = const_cast<double*>(Qinv);
  q.Rinv = const_cast<double*>(Rinv);
  q.S = S;
  q.Phi = Phiinv;
  return assemble(q, B, N, nx, nu, st);
}

int pcg_layer_gamma(int B, int N, int nx, int nu, const double* As_next, const double* As, const double* Bs,
                    const double* Qinv, const double* Rinv, const double* Cs, const double* qs, const double* rs,
                    double* gamma, hipStream_t st) {
  void* base = nullptr;
  int rc = get_scratch(qp_bytes(B, N, nx, nu), &base);
  if (rc) return rc;
  QpBufs q = carve_qp(base, B, N, nx, nu);
  const long long n1 = (long long)B * N * nx * nx + (long long)B * N * nx * nu;
  hipLaunchKernelGGL(premult_kernel, dim3(nblk(n1)), dim3(256), 0, st, B, N, nx, nu, As_next, As, Bs, Qinv, Rinv, q.AQ,
                     q.AnQ, q.BR);
  const long long nst = (long long)B * (N + 1) * nx;
  hipLaunchKernelGGL(gamma_kernel, dim3(nblk(nst)), dim3(256), 0, st, B, N, nx, nu, q.AQ, q.AnQ, q.BR, Cs, Qinv, qs, rs,
                     gamma);
  HIPCHK(hipGetLastError());
  return DMPC_OK;
}

int pcg_layer_solve(int B, int N, int nx, const double* S, const double* Phiinv, const double* gamma, double* lam,
                    int max_iter, double eps, int32_t* iters, double* eta, hipStream_t st) {
  void* base = nullptr;
  int rc = get_scratch(qp_bytes(B, N, nx, 1), &base);
  if (rc) return rc;
  QpBufs q = carve_qp(base, B, N, nx, 1);
  q.S = const_cast<double*>(S);
  q.Phi = const_cast<double*>(Phiinv);
  q.gamma = const_cast<double*>(gamma);
  q.lam = lam;
  rc = pcg_loop(q, B, N, nx, max_iter, eps, st);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(iters, q.iters, sizeof(int) * B, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(eta, q.eta, sizeof(double) * B, hipMemcpyDeviceToDevice, st));
  return DMPC_OK;
}

int pcg_layer_recover(int B, int N, int nx, int nu, const double* lam, const double* As_next, const double* As,
                      const double* Bs, const double* Qinv, const double* qs, const double* Rinv, const double* rs,
                      double* x, double* u, hipStream_t st) {
  const long long n = (long long)B * (N + 1) * (nx + nu);
  hipLaunchKernelGGL(recover_kernel, dim3(nblk(n)), dim3(256), 0, st, B, N, nx, nu, lam, As_next, As, Bs, Qinv, qs,
                     Rinv, rs, x, u);
  HIPCHK(hipGetLastError());
  return DMPC_OK;
}

template <class Model>
int linearize_t(const dmpc_problem* p, const dmpc_problem_data* d, const double* x, const double* u, double* fd,
                double* A, double* Bm, double* C, hipStream_t st) {
  constexpr int nx = Model::NX, nu = Model::NU;
  const int B = p->batch, N = p->horizon;
  hipLaunchKernelGGL(lin_kernel<Model>, dim3(nblk((long long)B * N * (nx + nu))), dim3(256), 0, st, B, N, p->n_params,
                     (float)p->dt, p->dt, p->scheme, d->dyn_params, x, u, fd, A, Bm);
  void* base = nullptr;
  int rc = get_scratch((size_t)B * N * nx * nx * 8 + 256, &base);
  if (rc) return rc;
  hipLaunchKernelGGL(c_kernel, dim3(nblk((long long)B * (N + 1) * nx)), dim3(256), 0, st, B, N, nx, nu, d->x0, x, u, fd,
                     A, Bm, C, reinterpret_cast<double*>(base));
  HIPCHK(hipGetLastError());
  return DMPC_OK;
}

int linearize_any(const dmpc_problem* p, const dmpc_problem_data* d, const double* x, const double* u, double* fd,
                  double* A, double* Bm, double* C, hipStream_t st) {
  switch (p->model) {
    case MODEL_LINEAR:
      if (p->nx == 4 && p->nu == 2) return linearize_t<LinearModel<4, 2>>(p, d, x, u, fd, A, Bm, C, st);
      if (p->nx == 8 && p->nu == 4) return linearize_t<LinearModel<8, 4>>(p, d, x, u, fd, A, Bm, C, st);
      return DMPC_ERR_UNSUPPORTED;
    case MODEL_QUADROTOR: return linearize_t<QuadrotorModel>(p, d, x, u, fd, A, Bm, C, st);
    case MODEL_SPACECRAFT3: return linearize_t<Spacecraft3Model>(p, d, x, u, fd, A, Bm, C, st);
    case MODEL_SPACECRAFT7: return linearize_t<Spacecraft7Model>(p, d, x, u, fd, A, Bm, C, st);
    case MODEL_CARTPOLE: return linearize_t<CartpoleModel>(p, d, x, u, fd, A, Bm, C, st);
    case MODEL_CHAIN: return linearize_t<ChainModel>(p, d, x, u, fd, A, Bm, C, st);
    DMPC_CUSTOM_CASE(return linearize_t<CustomModel>(p, d, x, u, fd, A, Bm, C, st))
    default: return DMPC_ERR_UNSUPPORTED;
  }
}

}  // namespace dmpc
