// poisson.hip -- MI355X (gfx950) pressure-Poisson solver: HIP kernels and
// the C ABI of include/cup2d_poisson.h.
//
// Method (reference source/Poisson/BiCGSTAB.cu:412-620): right-preconditioned
// BiCGSTAB with the dense 64x64 block-Jacobi inverse P_inv = -K^-1
// (ExpAMRSolver.cpp:25-95), restart on breakdown, best iterate x_opt, then
// removal of the h^2-weighted mean (ExpAMRSolver.cpp:383-415).
//
// MI355X design: one wavefront per 8x8 block (lane = cell); every vector
// pass of an iteration is fused into five kernels that keep all scalars on
// the device; each reduction ends in the last-arriving workgroup, which also
// runs the scalar recurrence (alpha, omega, beta, restart, convergence) so
// that no iteration needs a host round trip.  Iterations are replayed in
// fixed-size chunks from a hipGraph; kernels of a finished solve exit at
// launch on the device-side `done` flag.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cup2d_poisson.h"
#include "assembly.h"
#include "kernels.h"

using namespace cup2d;

namespace {

constexpr int MAXG = 512;    // workgroups of the vector kernels (grid-stride beyond)

// ----------------------------------------------------------------------------
// kernels
// ----------------------------------------------------------------------------

__device__ __forceinline__ void load_P(double *Ps, const double *__restrict__ P) {
  for (int i = threadIdx.x; i < 64 * 64; i += WG) Ps[i] = P[i];
  __syncthreads();
}

// State at the top of iteration i from the state of iteration i-1 and the
// totals o = (|r|_inf, |r|^2, (rhat,r), |rhat|^2) of iteration i-1's update
// (BiCGSTAB.cu:471-509 and :572-600).  prev.iter < 0 marks the state after
// r0 = b - A x0 (BiCGSTAB.cu:427-465).  Deterministic: recomputing it from
// the same inputs gives the same state.
__device__ Scal evaluate(const Scal &prev, const double (&o)[4]) {
  Scal s = prev;
  const double rr = o[1], rhr = o[2], hh = o[3];
  const double err = prev.norm ? sqrt(rr) : o[0];
  if (prev.iter < 0) {
    s.err = s.err0 = s.err_opt = err;
    s.iter = 0;
    s.alpha = s.omega = s.rho_prev = 1.0;
    s.rho = rhr;
    s.beta = (s.rho / (s.rho_prev + s.eps)) * (s.alpha / (s.omega + s.eps));
    s.restarts = s.converged = s.improved = s.restart = 0;
    s.done = (err == 0.0) ? 1 : 0;
    return s;
  }
  if (prev.done) {
    s.restart = 0;
    return s;
  }
  s.err = err;
  s.iter = prev.iter + 1;
  s.improved = 0;
  if (err < prev.err_opt) {
    s.improved = 1;
    s.err_opt = err;
    if (err <= s.tol_abs || err / s.err0 <= s.tol_rel) {
      s.converged = 1;
      s.done = 1;
    }
  }
  if (!s.done && s.iter >= s.max_iter) s.done = 1;
  s.rho_prev = prev.rho;
  s.rho = rhr;
  const bool breakdown = rhr * rhr < s.bd_tol * rr * hh;
  s.beta = (s.rho / (s.rho_prev + s.eps)) * (s.alpha / (s.omega + s.eps));
  s.restart = 0;
  if (!s.done && breakdown && s.max_restarts > 0) {
    s.restarts += 1;
    if (s.restarts >= s.max_restarts) {
      s.done = 1;
    } else {
      s.restart = 1;
      s.rho = rr;  // rhat = r
      s.rho_prev = s.alpha = s.omega = 1.0;
      s.beta = (s.rho / (s.rho_prev + s.eps)) * (s.alpha / (s.omega + s.eps));
    }
  }
  s.red[5] = rr;
  s.red[6] = rhr;
  s.red[7] = hh;
  return s;
}

// Partial sums of h^2 x0 (mean row of A x0 when the mean constraint is on).
__global__ __launch_bounds__(WG) void k_init_mean(int nb, const double *__restrict__ x0,
                                                  const double *__restrict__ h2, double *pm) {
  const long n = (long)nb * 64;
  double v[1] = {0.0};
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG)
    v[0] += h2[i >> 6] * x0[i];
  store_partials<1, 0u>(v, pm);
}

// r = b - A x0, rhat = r, p = nu = 0, x = x_opt = x0 (BiCGSTAB.cu:431-465);
// partials of (|r|_inf, |r|^2, (rhat,r), |rhat|^2) for the first state.
__global__ __launch_bounds__(WG) void k_init(OpView A, const double *__restrict__ b,
                                             const double *__restrict__ x0, double *__restrict__ x,
                                             double *__restrict__ xopt, double *__restrict__ r,
                                             double *__restrict__ rhat, double *__restrict__ p,
                                             double *__restrict__ nu, const double *pm, int Gm,
                                             double *p5) {
  __shared__ double xs[WAVES][64], ys[WAVES][64];
  double m0[1] = {0.0};
  if (A.mean_row >= 0) load_totals<1, 0u>(pm, Gm, m0);
  const int c = lane_id(), w = wave_id();
  const int nw = gridDim.x * WAVES;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int bb = blockIdx.x * WAVES + w; bb < A.nb; bb += nw) {
    const long i = (long)bb * 64 + c;
    const double xc = x0[i];
    xs[w][c] = xc;
    wave_sync();
    const double y = apply_block(A, x0, bb, c, xs[w], ys[w], xc);
    wave_sync();
    // mean row: b = 0 (ExpAMRSolver.cpp:340-345) and (A x)_row = sum h^2 x
    const double ri = (i == A.mean_row) ? 0.0 - m0[0] : b[i] - y;
    r[i] = ri;
    rhat[i] = ri;
    p[i] = 0.0;
    nu[i] = 0.0;
    x[i] = xc;
    xopt[i] = xc;
    v[0] = fmax(v[0], fabs(ri));
    v[1] += ri * ri;
  }
  v[2] = v[1];
  v[3] = v[1];
  store_partials<4, 1u>(v, p5);
}

// Top of iteration: evaluate the state (every workgroup, identically; the
// first one stores it), then p = r + beta (p - omega nu) (or p = rhat = r
// after a restart), x_opt = x when the last iteration improved, z1 = P p
// (BiCGSTAB.cu:471-520).
template <int NB>
__global__ __launch_bounds__(WG) void k_pupdate(int nb, const double *__restrict__ P,
                                                const double *__restrict__ r,
                                                double *__restrict__ rhat, double *__restrict__ p,
                                                const double *__restrict__ nu,
                                                const double *__restrict__ x,
                                                double *__restrict__ xopt, double *__restrict__ z1,
                                                Scal *st, int par, const double *p5, int G5) {
  __shared__ Scal cur;
  __shared__ double Ps[64 * 64];
  __shared__ double vs[WAVES][NB * 64];
  double o[4];
  load_totals<4, 1u>(p5, G5, o);
  if (threadIdx.x == 0) cur = evaluate(st[1 - par], o);
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x == 0) st[par] = cur;
  if (cur.done) return;
  const double beta = cur.beta, omega = cur.omega;
  const int restart = cur.restart, improved = cur.improved;
  load_P(Ps, P);
  const int c = lane_id(), w = wave_id();
  const int ng = (nb + NB - 1) / NB, nw = gridDim.x * WAVES;
  for (int g = blockIdx.x * WAVES + w; g < ng; g += nw) {
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const int bb = g * NB + j;
      double pi = 0.0;
      if (bb < nb) {
        const long i = (long)bb * 64 + c;
        const double ri = r[i];
        if (restart) {
          rhat[i] = ri;
          pi = ri;
        } else {
          pi = ri + beta * (p[i] - omega * nu[i]);
        }
        p[i] = pi;
        if (improved) xopt[i] = x[i];
      }
      vs[w][j * 64 + c] = pi;
    }
    wave_sync();
    double z[NB];
    prec_group<NB>(Ps, vs[w], c, z);
    wave_sync();
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const int bb = g * NB + j;
      if (bb < nb) z1[(long)bb * 64 + c] = z[j];
    }
  }
}

// State evaluation alone (end of an iteration chunk, one workgroup), so the
// host sees convergence of the chunk's last iteration.
__global__ __launch_bounds__(WG) void k_eval(Scal *st, int par, const double *p5, int G5) {
  double o[4];
  load_totals<4, 1u>(p5, G5, o);
  if (threadIdx.x == 0) st[par] = evaluate(st[1 - par], o);
}

// nu = A z1; partials of (rhat, nu) and sum h^2 z1   (BiCGSTAB.cu:523-527)
__global__ __launch_bounds__(WG) void k_spmv_nu(OpView A, const double *__restrict__ z1,
                                                double *__restrict__ nu,
                                                const double *__restrict__ rhat, const Scal *st,
                                                int par, double *p2) {
  if (st[par].done) return;
  __shared__ double xs[WAVES][64], ys[WAVES][64];
  const int c = lane_id(), w = wave_id();
  const int nw = gridDim.x * WAVES;
  double v[2] = {0.0, 0.0};
  for (int bb = blockIdx.x * WAVES + w; bb < A.nb; bb += nw) {
    const long i = (long)bb * 64 + c;
    const double zc = z1[i];
    xs[w][c] = zc;
    wave_sync();
    double y = apply_block(A, z1, bb, c, xs[w], ys[w], zc);
    wave_sync();
    if (i == A.mean_row) y = 0.0;  // replaced by the consumer
    nu[i] = y;
    v[0] += rhat[i] * y;
    v[1] += A.h2[bb] * zc;
  }
  store_partials<2, 0u>(v, p2);
}

// alpha = rho / ((rhat, nu) + eps); s = r - alpha nu (stored in r); z2 = P s
// (BiCGSTAB.cu:527-545)
template <int NB>
__global__ __launch_bounds__(WG) void k_supdate(int nb, const double *__restrict__ P,
                                                double *__restrict__ r, double *__restrict__ nu,
                                                const double *__restrict__ rhat,
                                                double *__restrict__ z2, Scal *st, int par,
                                                const double *p2, int G2, long mean_row) {
  if (st[par].done) return;
  __shared__ double Ps[64 * 64];
  __shared__ double vs[WAVES][NB * 64];
  double o[2];
  load_totals<2, 0u>(p2, G2, o);
  double rn = o[0];
  if (mean_row >= 0) rn += rhat[mean_row] * o[1];
  const double alpha = st[par].rho / (rn + st[par].eps);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st[par].alpha = alpha;
    st[par].red[2] = rn;
  }
  load_P(Ps, P);
  const int c = lane_id(), w = wave_id();
  const int ng = (nb + NB - 1) / NB, nw = gridDim.x * WAVES;
  for (int g = blockIdx.x * WAVES + w; g < ng; g += nw) {
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const int bb = g * NB + j;
      double si = 0.0;
      if (bb < nb) {
        const long i = (long)bb * 64 + c;
        double ni = nu[i];
        if (i == mean_row) {
          ni = o[1];
          nu[i] = ni;
        }
        si = r[i] - alpha * ni;
        r[i] = si;
      }
      vs[w][j * 64 + c] = si;
    }
    wave_sync();
    double z[NB];
    prec_group<NB>(Ps, vs[w], c, z);
    wave_sync();
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const int bb = g * NB + j;
      if (bb < nb) z2[(long)bb * 64 + c] = z[j];
    }
  }
}

// t = A z2; partials of (t, s), (t, t) and sum h^2 z2   (BiCGSTAB.cu:548-553)
__global__ __launch_bounds__(WG) void k_spmv_t(OpView A, const double *__restrict__ z2,
                                               double *__restrict__ t,
                                               const double *__restrict__ sv, const Scal *st,
                                               int par, double *p4) {
  if (st[par].done) return;
  __shared__ double xs[WAVES][64], ys[WAVES][64];
  const int c = lane_id(), w = wave_id();
  const int nw = gridDim.x * WAVES;
  double v[3] = {0.0, 0.0, 0.0};
  for (int bb = blockIdx.x * WAVES + w; bb < A.nb; bb += nw) {
    const long i = (long)bb * 64 + c;
    const double zc = z2[i];
    xs[w][c] = zc;
    wave_sync();
    double y = apply_block(A, z2, bb, c, xs[w], ys[w], zc);
    wave_sync();
    if (i == A.mean_row) y = 0.0;
    t[i] = y;
    v[0] += y * sv[i];
    v[1] += y * y;
    v[2] += A.h2[bb] * zc;
  }
  store_partials<3, 0u>(v, p4);
}

// omega = (t, s) / ((t, t) + eps); x += alpha z1 + omega z2; r = s - omega t;
// partials of (|r|_inf, |r|^2, (rhat, r), |rhat|^2)   (BiCGSTAB.cu:552-572)
__global__ __launch_bounds__(WG) void k_xrupdate(int nb, double *__restrict__ x,
                                                 const double *__restrict__ z1,
                                                 const double *__restrict__ z2,
                                                 double *__restrict__ r,
                                                 const double *__restrict__ t,
                                                 const double *__restrict__ rhat, Scal *st,
                                                 int par, const double *p4, int G4, double *p5,
                                                 long mean_row) {
  if (st[par].done) return;
  double o[3];
  load_totals<3, 0u>(p4, G4, o);
  double ts = o[0], tt = o[1];
  if (mean_row >= 0) {
    ts += o[2] * r[mean_row];
    tt += o[2] * o[2];
  }
  const double omega = ts / (tt + st[par].eps);
  const double alpha = st[par].alpha;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st[par].omega = omega;
    st[par].red[3] = ts;
    st[par].red[4] = tt;
  }
  const long n = (long)nb * 64;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) {
    x[i] = x[i] + alpha * z1[i] + omega * z2[i];
    const double ti = (i == mean_row) ? o[2] : t[i];
    const double ri = r[i] - omega * ti;
    r[i] = ri;
    const double rh = rhat[i];
    v[0] = fmax(v[0], fabs(ri));
    v[1] += ri * ri;
    v[2] += rh * ri;
    v[3] += rh * rh;
  }
  store_partials<4, 1u>(v, p5);
}

// mean of the returned iterate (x if the last iteration improved, else x_opt)
__global__ __launch_bounds__(WG) void k_final_mean(int nb, const double *__restrict__ x,
                                                   const double *__restrict__ xopt,
                                                   const double *__restrict__ h2, const Scal *st,
                                                   int par, double *pf) {
  const double *src = st[par].improved ? x : xopt;
  const long n = (long)nb * 64;
  double v[2] = {0.0, 0.0};
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) {
    const double hh = h2[i >> 6];
    v[0] += src[i] * hh;
    v[1] += hh;
  }
  store_partials<2, 0u>(v, pf);
}

__global__ __launch_bounds__(WG) void k_final_write(int nb, const double *__restrict__ x,
                                                    const double *__restrict__ xopt,
                                                    double *__restrict__ out, Scal *st, int par,
                                                    const double *pf, int Gf) {
  double o[2];
  load_totals<2, 0u>(pf, Gf, o);
  const double mean = o[0] / o[1];
  const double *src = st[par].improved ? x : xopt;
  if (blockIdx.x == 0 && threadIdx.x == 0) st[par].mean = mean;
  const long n = (long)nb * 64;
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG)
    out[i] = src[i] - mean;
}

// y = A x (stencil form); partials of sum h^2 x for the mean row
__global__ __launch_bounds__(WG) void k_spmv(OpView A, const double *__restrict__ x,
                                             double *__restrict__ y, double *pm) {
  __shared__ double xs[WAVES][64], ys[WAVES][64];
  const int c = lane_id(), w = wave_id();
  const int nw = gridDim.x * WAVES;
  double v[1] = {0.0};
  for (int bb = blockIdx.x * WAVES + w; bb < A.nb; bb += nw) {
    const long i = (long)bb * 64 + c;
    const double xc = x[i];
    xs[w][c] = xc;
    wave_sync();
    const double yi = apply_block(A, x, bb, c, xs[w], ys[w], xc);
    wave_sync();
    y[i] = yi;
    v[0] += A.h2[bb] * xc;
  }
  if (A.mean_row >= 0) store_partials<1, 0u>(v, pm);
}

// y = A x, plain CSR, one row per thread
__global__ __launch_bounds__(WG) void k_spmv_csr(long n, const int *__restrict__ rp,
                                                 const int *__restrict__ col,
                                                 const double *__restrict__ val,
                                                 const double *__restrict__ x,
                                                 double *__restrict__ y) {
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) {
    double acc = 0.0;
    const int k1 = rp[i + 1];
    for (int k = rp[i]; k < k1; k++) acc += val[k] * x[col[k]];
    y[i] = acc;
  }
}

// y[mean_row] = sum h^2 x from the partials (one workgroup)
__global__ __launch_bounds__(WG) void k_meanrow_fix(const double *pm, int G, long mean_row,
                                                    double *__restrict__ y) {
  double o[1];
  load_totals<1, 0u>(pm, G, o);
  if (threadIdx.x == 0) y[mean_row] = o[0];
}

// z = P r on every block
template <int NB>
__global__ __launch_bounds__(WG) void k_prec(int nb, const double *__restrict__ P,
                                             const double *__restrict__ r,
                                             double *__restrict__ z) {
  __shared__ double Ps[64 * 64];
  __shared__ double vs[WAVES][NB * 64];
  load_P(Ps, P);
  const int c = lane_id(), w = wave_id();
  const int ng = (nb + NB - 1) / NB, nw = gridDim.x * WAVES;
  for (int g = blockIdx.x * WAVES + w; g < ng; g += nw) {
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const int bb = g * NB + j;
      vs[w][j * 64 + c] = bb < nb ? r[(long)bb * 64 + c] : 0.0;
    }
    wave_sync();
    double zz[NB];
    prec_group<NB>(Ps, vs[w], c, zz);
    wave_sync();
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const int bb = g * NB + j;
      if (bb < nb) z[(long)bb * 64 + c] = zz[j];
    }
  }
}

// partials of |b - y|^2, |b - y|_inf, |b|^2, |b|_inf; y holds A x
__global__ __launch_bounds__(WG) void k_residual(long n, const double *__restrict__ b,
                                                 const double *__restrict__ y, long mean_row,
                                                 double *pr) {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) {
    const double bi = (i == mean_row) ? 0.0 : b[i];
    const double ri = bi - y[i];
    v[0] += ri * ri;
    v[1] = fmax(v[1], fabs(ri));
    v[2] += bi * bi;
    v[3] = fmax(v[3], fabs(bi));
  }
  store_partials<4, 0xAu>(v, pr);
}

// ----------------------------------------------------------------------------
// host side
// ----------------------------------------------------------------------------

std::string g_create_err;

template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
};

}  // namespace

struct cup2d_ps {
  cup2d_ps_config cfg{};
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool have_mesh = false;
  Operator op;
  // device operator
  int4 *d_nbr = nullptr;
  unsigned long long *d_mask = nullptr;
  int *d_bptr = nullptr, *d_cell = nullptr, *d_rptr = nullptr, *d_col = nullptr;
  double *d_val = nullptr, *d_h2 = nullptr;
  int *d_csr_rp = nullptr, *d_csr_col = nullptr;
  double *d_csr_val = nullptr;
  double *d_P = nullptr;
  // work vectors
  double *d_x = nullptr, *d_xopt = nullptr, *d_r = nullptr, *d_rhat = nullptr, *d_p = nullptr,
         *d_nu = nullptr, *d_z1 = nullptr, *d_z2 = nullptr, *d_t = nullptr, *d_b = nullptr,
         *d_xio = nullptr;
  Scal *d_s = nullptr;                 // two state slots (iteration parity)
  double *d_p2 = nullptr, *d_p4 = nullptr, *d_p5 = nullptr, *d_pm = nullptr, *d_pf = nullptr;
  Scal *h_s = nullptr;                 // pinned copy of the state slots
  int G = 1;                           // workgroups of the SpMV / vector kernels
  int Gp = 1;                          // workgroups of the preconditioner kernels
  int nbp = 2;                         // blocks per wave in the preconditioner kernels
  // graph of one chunk of iterations
  hipGraphExec_t gexec = nullptr;
  int gchunk = 0;
  std::vector<void *> allocs;
};

namespace {

int fail(cup2d_ps *h, int code, const std::string &msg) {
  if (h) h->err = msg;
  else g_create_err = msg;
  return code;
}

#define HIPCHK(h, expr)                                                                      \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess)                                                                    \
      return fail((h), _e == hipErrorOutOfMemory ? CUP2D_PS_ENOMEM : CUP2D_PS_EHIP,          \
                  std::string(#expr) + ": " + hipGetErrorString(_e));                        \
  } while (0)

template <class T>
int dalloc(cup2d_ps *h, T **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  HIPCHK(h, hipMalloc((void **)p, count * sizeof(T)));
  h->allocs.push_back(*p);
  return CUP2D_PS_OK;
}

template <class T>
int upload(cup2d_ps *h, T **p, const std::vector<T> &v) {
  int rc = dalloc(h, p, v.size());
  if (rc) return rc;
  if (!v.empty()) HIPCHK(h, hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return CUP2D_PS_OK;
}

void free_mesh(cup2d_ps *h) {
  if (h->gexec) {
    (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
  }
  for (void *p : h->allocs) (void)hipFree(p);
  h->allocs.clear();
  h->have_mesh = false;
}

OpView view(const cup2d_ps *h) {
  OpView A;
  A.nb = (int)h->op.nb;
  A.mean_row = (int)h->op.mean_row;
  A.nbr = h->d_nbr;
  A.mask = h->d_mask;
  A.bptr = h->d_bptr;
  A.cell = h->d_cell;
  A.rptr = h->d_rptr;
  A.col = h->d_col;
  A.val = h->d_val;
  A.h2 = h->d_h2;
  return A;
}

// -K^-1 for the 8x8 block (Cholesky-Crout of K then L^-1 by Gauss-Jordan and
// P = -(L^-T L^-1), the construction of ExpAMRSolver.cpp:37-95).
std::vector<double> make_pinv() {
  const int N = 64;
  auto K = [](int a, int b) {
    const int ya = a / 8, xa = a % 8, yb = b / 8, xb = b % 8;
    if (a == b) return 4.0;
    return (std::abs(xa - xb) + std::abs(ya - yb) == 1) ? -1.0 : 0.0;
  };
  std::vector<double> L(N * N, 0.0), Li(N * N, 0.0), P(N * N);
  for (int i = 0; i < N; i++) {
    double s1 = 0;
    for (int k = 0; k < i; k++) s1 += L[i * N + k] * L[i * N + k];
    L[i * N + i] = std::sqrt(K(i, i) - s1);
    for (int j = i + 1; j < N; j++) {
      double s2 = 0;
      for (int k = 0; k < i; k++) s2 += L[i * N + k] * L[j * N + k];
      L[j * N + i] = (K(j, i) - s2) / L[i * N + i];
    }
  }
  for (int i = 0; i < N; i++) Li[i * N + i] = 1.0;
  for (int br = 0; br < N; br++) {
    const double f = 1.0 / L[br * N + br];
    for (int c = 0; c <= br; c++) Li[br * N + c] *= f;
    for (int wr = br + 1; wr < N; wr++) {
      const double wf = L[wr * N + br];
      for (int c = 0; c <= br; c++) Li[wr * N + c] -= wf * Li[br * N + c];
    }
  }
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      double a = 0;
      for (int k = std::max(i, j); k < N; k++) a += Li[k * N + i] * Li[k * N + j];
      P[i * N + j] = -a;
    }
  return P;
}

// Launches one BiCGSTAB iteration (parity `par` selects the state slot).
void launch_iteration(cup2d_ps *h, int par) {
  const OpView A = view(h);
  const int G = h->G, Gp = h->Gp;
  const int nb = (int)h->op.nb;
  const long mr = h->op.mean_row;
  hipStream_t st = h->stream;
  if (h->nbp == 4)
    k_pupdate<4><<<Gp, WG, 0, st>>>(nb, h->d_P, h->d_r, h->d_rhat, h->d_p, h->d_nu, h->d_x,
                                    h->d_xopt, h->d_z1, h->d_s, par, h->d_p5, G);
  else if (h->nbp == 2)
    k_pupdate<2><<<Gp, WG, 0, st>>>(nb, h->d_P, h->d_r, h->d_rhat, h->d_p, h->d_nu, h->d_x,
                                    h->d_xopt, h->d_z1, h->d_s, par, h->d_p5, G);
  else
    k_pupdate<1><<<Gp, WG, 0, st>>>(nb, h->d_P, h->d_r, h->d_rhat, h->d_p, h->d_nu, h->d_x,
                                    h->d_xopt, h->d_z1, h->d_s, par, h->d_p5, G);
  k_spmv_nu<<<G, WG, 0, st>>>(A, h->d_z1, h->d_nu, h->d_rhat, h->d_s, par, h->d_p2);
  if (h->nbp == 4)
    k_supdate<4><<<Gp, WG, 0, st>>>(nb, h->d_P, h->d_r, h->d_nu, h->d_rhat, h->d_z2, h->d_s, par,
                                    h->d_p2, G, mr);
  else if (h->nbp == 2)
    k_supdate<2><<<Gp, WG, 0, st>>>(nb, h->d_P, h->d_r, h->d_nu, h->d_rhat, h->d_z2, h->d_s, par,
                                    h->d_p2, G, mr);
  else
    k_supdate<1><<<Gp, WG, 0, st>>>(nb, h->d_P, h->d_r, h->d_nu, h->d_rhat, h->d_z2, h->d_s, par,
                                    h->d_p2, G, mr);
  k_spmv_t<<<G, WG, 0, st>>>(A, h->d_z2, h->d_t, h->d_r, h->d_s, par, h->d_p4);
  k_xrupdate<<<G, WG, 0, st>>>(nb, h->d_x, h->d_z1, h->d_z2, h->d_r, h->d_t, h->d_rhat, h->d_s,
                               par, h->d_p4, G, h->d_p5, mr);
}

// Iterations per chunk: even, so that every chunk starts on state slot 0.
int chunk_of(const cup2d_ps *h) {
  int c = h->cfg.chunk > 0 ? h->cfg.chunk : 16;
  return c + (c & 1);
}

// One chunk: C iterations, then the evaluation of the next state into slot 0.
void launch_chunk(cup2d_ps *h) {
  const int C = chunk_of(h);
  for (int k = 0; k < C; k++) launch_iteration(h, k & 1);
  k_eval<<<1, WG, 0, h->stream>>>(h->d_s, 0, h->d_p5, h->G);
}

int ensure_graph(cup2d_ps *h) {
  const int C = chunk_of(h);
  if (h->gexec && h->gchunk == C) return CUP2D_PS_OK;
  if (h->gexec) {
    (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
  }
  hipGraph_t g = nullptr;
  HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  launch_chunk(h);
  HIPCHK(h, hipStreamEndCapture(h->stream, &g));
  hipError_t e = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess)
    return fail(h, CUP2D_PS_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  h->gchunk = C;
  return CUP2D_PS_OK;
}

int solve_impl(cup2d_ps *h, const double *d_b, const double *d_x0, double *d_out, double tol_abs,
               double tol_rel, int max_restarts, cup2d_ps_stats *st) {
  if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "solve before cup2d_ps_set_mesh");
  HIPCHK(h, hipSetDevice(h->device));
  const int G = h->G;
  const int nb = (int)h->op.nb;
  // slot 1 holds the marker state from which the first evaluation starts
  Scal init{};
  init.tol_abs = tol_abs;
  init.tol_rel = tol_rel;
  init.eps = h->cfg.eps;
  init.bd_tol = h->cfg.breakdown_tol;
  init.max_iter = h->cfg.max_iter;
  init.max_restarts = max_restarts;
  init.norm = h->cfg.norm;
  init.iter = -1;
  h->h_s[0] = init;
  h->h_s[1] = init;
  HIPCHK(h, hipEventRecord(h->ev0, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->d_s, h->h_s, 2 * sizeof(Scal), hipMemcpyHostToDevice, h->stream));
  const OpView A = view(h);
  if (h->op.mean_row >= 0) k_init_mean<<<G, WG, 0, h->stream>>>(nb, d_x0, h->d_h2, h->d_pm);
  k_init<<<G, WG, 0, h->stream>>>(A, d_b, d_x0, h->d_x, h->d_xopt, h->d_r, h->d_rhat, h->d_p,
                                  h->d_nu, h->d_pm, G, h->d_p5);
  HIPCHK(h, hipGetLastError());
  const int C = chunk_of(h);
  if (h->cfg.use_graph) {
    int rc = ensure_graph(h);
    if (rc) return rc;
  }
  int launched = 0;
  for (;;) {
    if (h->cfg.use_graph) {
      HIPCHK(h, hipGraphLaunch(h->gexec, h->stream));
    } else {
      launch_chunk(h);
      HIPCHK(h, hipGetLastError());
    }
    launched += C;
    HIPCHK(h, hipMemcpyAsync(&h->h_s[0].done, &h->d_s[0].done, sizeof(int),
                             hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (h->h_s[0].done || launched > h->cfg.max_iter + C) break;
  }
  k_final_mean<<<G, WG, 0, h->stream>>>(nb, h->d_x, h->d_xopt, h->d_h2, h->d_s, 0, h->d_pf);
  k_final_write<<<G, WG, 0, h->stream>>>(nb, h->d_x, h->d_xopt, d_out, h->d_s, 0, h->d_pf, G);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipEventRecord(h->ev1, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->h_s, h->d_s, 2 * sizeof(Scal), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (st) {
    const Scal &f = h->h_s[0];
    std::memset(st, 0, sizeof(*st));
    st->iterations = f.iter;
    st->restarts = f.restarts;
    st->converged = f.converged;
    st->refinements = 0;
    st->err_init = f.err0;
    st->err_final = f.err_opt;
    float ms = 0;
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    st->ms = ms;
  }
  return CUP2D_PS_OK;
}

}  // namespace

extern "C" {

void cup2d_ps_config_default(cup2d_ps_config *cfg) {
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->device = 0;
  cfg->precision = CUP2D_PS_FP64;
  cfg->mean_constraint = 0;
  cfg->norm = CUP2D_PS_NORM_INF;
  cfg->max_iter = 1000;
  cfg->chunk = 0;
  cfg->use_graph = 1;
  cfg->eps = 1e-100;
  cfg->breakdown_tol = 1e-16;
  cfg->inner_tol = 1e-3;
}

int cup2d_ps_create(const cup2d_ps_config *cfg, cup2d_ps **out) {
  if (!out) return fail(nullptr, CUP2D_PS_EINVAL, "out is NULL");
  *out = nullptr;
  cup2d_ps_config c;
  if (cfg) c = *cfg;
  else cup2d_ps_config_default(&c);
  if (c.max_iter <= 0 || c.chunk < 0 || c.eps < 0 || c.breakdown_tol < 0 ||
      (c.precision != CUP2D_PS_FP64 && c.precision != CUP2D_PS_MIXED) ||
      (c.norm != CUP2D_PS_NORM_INF && c.norm != CUP2D_PS_NORM_L2))
    return fail(nullptr, CUP2D_PS_EINVAL, "invalid configuration");
  if (c.precision == CUP2D_PS_MIXED)
    return fail(nullptr, CUP2D_PS_EINVAL, "mixed precision is not available in this build");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(nullptr, CUP2D_PS_ENODEV, "no HIP device");
  if (c.device < 0 || c.device >= ndev) return fail(nullptr, CUP2D_PS_EINVAL, "bad device ordinal");
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c.device) != hipSuccess)
    return fail(nullptr, CUP2D_PS_ENODEV, "hipGetDeviceProperties failed");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(nullptr, CUP2D_PS_ENODEV, std::string("built for gfx950, device is ") + prop.gcnArchName);
  auto h = std::make_unique<cup2d_ps>();
  h->cfg = c;
  h->device = c.device;
  if (hipSetDevice(c.device) != hipSuccess) return fail(nullptr, CUP2D_PS_EHIP, "hipSetDevice failed");
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess ||
      hipHostMalloc((void **)&h->h_s, 2 * sizeof(Scal), hipHostMallocDefault) != hipSuccess)
    return fail(nullptr, CUP2D_PS_EHIP, "stream/event/pinned allocation failed");
  *out = h.release();
  return CUP2D_PS_OK;
}

int cup2d_ps_destroy(cup2d_ps *h) {
  if (!h) return CUP2D_PS_OK;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  free_mesh(h);
  if (h->h_s) (void)hipHostFree(h->h_s);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return CUP2D_PS_OK;
}

int cup2d_ps_set_mesh(cup2d_ps *h, const cup2d_mesh_desc *m) {
  if (!h || !m) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  HIPCHK(h, hipSetDevice(h->device));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  free_mesh(h);
  MeshDesc md{m->bpdx, m->bpdy, m->levels, m->periodic_x, m->periodic_y, m->extent, m->n_blocks,
              m->level, m->bx, m->by};
  std::string err;
  if (!build_operator(md, h->cfg.mean_constraint, h->op, err)) return fail(h, CUP2D_PS_EINVAL, err);
  const Operator &op = h->op;
  int rc = 0;
  std::vector<int4> nbr(op.nb);
  for (long long b = 0; b < op.nb; b++)
    nbr[b] = make_int4(op.nbr[4 * b], op.nbr[4 * b + 1], op.nbr[4 * b + 2], op.nbr[4 * b + 3]);
  std::vector<unsigned long long> mask(op.exc_mask.begin(), op.exc_mask.end());
  if ((rc = upload(h, &h->d_nbr, nbr)) || (rc = upload(h, &h->d_mask, mask)) ||
      (rc = upload(h, &h->d_bptr, op.exc_bptr)) || (rc = upload(h, &h->d_cell, op.exc_cell)) ||
      (rc = upload(h, &h->d_rptr, op.exc_rptr)) || (rc = upload(h, &h->d_col, op.exc_col)) ||
      (rc = upload(h, &h->d_val, op.exc_val)) || (rc = upload(h, &h->d_h2, op.h2)) ||
      (rc = upload(h, &h->d_csr_rp, op.rp)) || (rc = upload(h, &h->d_csr_col, op.col)) ||
      (rc = upload(h, &h->d_csr_val, op.val)) || (rc = upload(h, &h->d_P, make_pinv())))
    return rc;
  const size_t n = op.n;
  double **vecs[] = {&h->d_x, &h->d_xopt, &h->d_r, &h->d_rhat, &h->d_p, &h->d_nu,
                     &h->d_z1, &h->d_z2, &h->d_t, &h->d_b, &h->d_xio};
  for (double **v : vecs)
    if ((rc = dalloc(h, v, n))) return rc;
  if ((rc = dalloc(h, &h->d_s, 2)) || (rc = dalloc(h, &h->d_p2, (size_t)2 * MAXG)) ||
      (rc = dalloc(h, &h->d_p4, (size_t)3 * MAXG)) || (rc = dalloc(h, &h->d_p5, (size_t)4 * MAXG)) ||
      (rc = dalloc(h, &h->d_pm, (size_t)MAXG)) || (rc = dalloc(h, &h->d_pf, (size_t)4 * MAXG)))
    return rc;
  HIPCHK(h, hipMemset(h->d_s, 0, 2 * sizeof(Scal)));
  h->G = (int)std::max<long long>(1, std::min<long long>(MAXG, (op.nb + WAVES - 1) / WAVES));
  h->nbp = h->cfg.reserved_i[0] > 0 ? h->cfg.reserved_i[0] : (op.nb >= 8192 ? 4 : 2);
  if (h->nbp != 1 && h->nbp != 2 && h->nbp != 4) h->nbp = 2;
  const long long groups = (op.nb + h->nbp - 1) / h->nbp;
  h->Gp = (int)std::max<long long>(1, std::min<long long>(MAXG, (groups + WAVES - 1) / WAVES));
  h->have_mesh = true;
  return CUP2D_PS_OK;
}

int cup2d_ps_solve_device(cup2d_ps *h, const double *d_b, double *d_x, double tol_abs,
                          double tol_rel, int max_restarts, cup2d_ps_stats *st) {
  if (!h || !d_b || !d_x) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  if (tol_abs < 0 || tol_rel < 0 || max_restarts < 0) return fail(h, CUP2D_PS_EINVAL, "negative tolerance");
  return solve_impl(h, d_b, d_x, d_x, tol_abs, tol_rel, max_restarts, st);
}

int cup2d_ps_solve(cup2d_ps *h, const double *b, double *x, double tol_abs, double tol_rel,
                   int max_restarts, cup2d_ps_stats *st) {
  if (!h || !b || !x) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "solve before cup2d_ps_set_mesh");
  if (tol_abs < 0 || tol_rel < 0 || max_restarts < 0) return fail(h, CUP2D_PS_EINVAL, "negative tolerance");
  HIPCHK(h, hipSetDevice(h->device));
  const size_t bytes = h->op.n * sizeof(double);
  HIPCHK(h, hipMemcpyAsync(h->d_b, b, bytes, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->d_xio, x, bytes, hipMemcpyHostToDevice, h->stream));
  int rc = solve_impl(h, h->d_b, h->d_xio, h->d_xio, tol_abs, tol_rel, max_restarts, st);
  if (rc) return rc;
  HIPCHK(h, hipMemcpyAsync(x, h->d_xio, bytes, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return CUP2D_PS_OK;
}

int cup2d_ps_spmv_device(cup2d_ps *h, const double *d_x, double *d_y, int kind) {
  if (!h || !d_x || !d_y) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "no mesh");
  if (kind != 0 && kind != 1) return fail(h, CUP2D_PS_EINVAL, "kind must be 0 or 1");
  HIPCHK(h, hipSetDevice(h->device));
  const long n = (long)h->op.n;
  if (kind == 0) {
    k_spmv<<<h->G, WG, 0, h->stream>>>(view(h), d_x, d_y, h->d_pm);
  } else {
    const int G = (int)std::min<long>(4096, (n + WG - 1) / WG);
    k_spmv_csr<<<G, WG, 0, h->stream>>>(n, h->d_csr_rp, h->d_csr_col, h->d_csr_val, d_x, d_y);
    if (h->op.mean_row >= 0) k_init_mean<<<h->G, WG, 0, h->stream>>>((int)h->op.nb, d_x, h->d_h2, h->d_pm);
  }
  if (h->op.mean_row >= 0) k_meanrow_fix<<<1, WG, 0, h->stream>>>(h->d_pm, h->G, h->op.mean_row, d_y);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return CUP2D_PS_OK;
}

int cup2d_ps_spmv_time(cup2d_ps *h, const double *d_x, double *d_y, int kind, int reps,
                       double *ms_per_launch) {
  if (!h || !d_x || !d_y || !ms_per_launch || reps <= 0) return fail(h, CUP2D_PS_EINVAL, "bad argument");
  if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "no mesh");
  if (kind != 0 && kind != 1) return fail(h, CUP2D_PS_EINVAL, "kind must be 0 or 1");
  HIPCHK(h, hipSetDevice(h->device));
  const long n = (long)h->op.n;
  const int Gc = (int)std::min<long>(4096, (n + WG - 1) / WG);
  const OpView A = view(h);
  HIPCHK(h, hipEventRecord(h->ev0, h->stream));
  for (int k = 0; k < reps; k++) {
    if (kind == 0) k_spmv<<<h->G, WG, 0, h->stream>>>(A, d_x, d_y, h->d_pm);
    else k_spmv_csr<<<Gc, WG, 0, h->stream>>>(n, h->d_csr_rp, h->d_csr_col, h->d_csr_val, d_x, d_y);
  }
  HIPCHK(h, hipEventRecord(h->ev1, h->stream));
  HIPCHK(h, hipEventSynchronize(h->ev1));
  HIPCHK(h, hipGetLastError());
  float ms = 0;
  HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  *ms_per_launch = ms / reps;
  return CUP2D_PS_OK;
}

int cup2d_ps_precond_device(cup2d_ps *h, const double *d_r, double *d_z) {
  if (!h || !d_r || !d_z) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "no mesh");
  HIPCHK(h, hipSetDevice(h->device));
  const int nb = (int)h->op.nb;
  if (h->nbp == 4) k_prec<4><<<h->Gp, WG, 0, h->stream>>>(nb, h->d_P, d_r, d_z);
  else if (h->nbp == 2) k_prec<2><<<h->Gp, WG, 0, h->stream>>>(nb, h->d_P, d_r, d_z);
  else k_prec<1><<<h->Gp, WG, 0, h->stream>>>(nb, h->d_P, d_r, d_z);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return CUP2D_PS_OK;
}

int cup2d_ps_residual_device(cup2d_ps *h, const double *d_b, const double *d_x, double out[4]) {
  if (!h || !d_b || !d_x || !out) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  if (!h->have_mesh) return fail(h, CUP2D_PS_ENOMESH, "no mesh");
  HIPCHK(h, hipSetDevice(h->device));
  k_spmv<<<h->G, WG, 0, h->stream>>>(view(h), d_x, h->d_t, h->d_pm);
  if (h->op.mean_row >= 0) k_meanrow_fix<<<1, WG, 0, h->stream>>>(h->d_pm, h->G, h->op.mean_row, h->d_t);
  k_residual<<<h->G, WG, 0, h->stream>>>((long)h->op.n, d_b, h->d_t, h->op.mean_row, h->d_pf);
  HIPCHK(h, hipGetLastError());
  std::vector<double> part(4 * (size_t)h->G);
  HIPCHK(h, hipMemcpyAsync(part.data(), h->d_pf, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  double r2 = 0, rinf = 0, b2 = 0, binf = 0;
  for (int i = 0; i < h->G; i++) {
    r2 += part[i];
    rinf = std::max(rinf, part[h->G + i]);
    b2 += part[2 * h->G + i];
    binf = std::max(binf, part[3 * h->G + i]);
  }
  out[0] = std::sqrt(r2);
  out[1] = rinf;
  out[2] = std::sqrt(b2);
  out[3] = binf;
  return CUP2D_PS_OK;
}

int cup2d_ps_sizes(const cup2d_ps *h, long long *n, long long *nnz, long long *n_explicit,
                   long long *mean_row) {
  if (!h) return CUP2D_PS_EINVAL;
  if (!h->have_mesh) return CUP2D_PS_ENOMESH;
  if (n) *n = h->op.n;
  if (nnz) *nnz = h->op.nnz();
  if (n_explicit) *n_explicit = h->op.n_exc();
  if (mean_row) *mean_row = h->op.mean_row;
  return CUP2D_PS_OK;
}

int cup2d_ps_get_csr(const cup2d_ps *h, long long *row_ptr, int *col, double *val) {
  if (!h || !row_ptr || !col || !val) return CUP2D_PS_EINVAL;
  if (!h->have_mesh) return CUP2D_PS_ENOMESH;
  for (size_t i = 0; i < h->op.rp.size(); i++) row_ptr[i] = h->op.rp[i];
  std::memcpy(col, h->op.col.data(), h->op.col.size() * sizeof(int));
  std::memcpy(val, h->op.val.data(), h->op.val.size() * sizeof(double));
  return CUP2D_PS_OK;
}

long long cup2d_op_csr(const cup2d_mesh_desc *m, int mean_constraint, long long *row_ptr, int *col,
                       double *val, long long cap, long long *mean_row) {
  if (!m) {
    g_create_err = "NULL mesh";
    return -1;
  }
  MeshDesc md{m->bpdx, m->bpdy, m->levels, m->periodic_x, m->periodic_y, m->extent, m->n_blocks,
              m->level, m->bx, m->by};
  Operator op;
  std::string err;
  if (!build_operator(md, mean_constraint, op, err)) {
    g_create_err = err;
    return -1;
  }
  if (mean_row) *mean_row = op.mean_row;
  const long long nnz = op.nnz();
  if (nnz > cap || !row_ptr || !col || !val) return -nnz;
  for (size_t i = 0; i < op.rp.size(); i++) row_ptr[i] = op.rp[i];
  std::memcpy(col, op.col.data(), nnz * sizeof(int));
  std::memcpy(val, op.val.data(), nnz * sizeof(double));
  return nnz;
}

int cup2d_ps_dmalloc(cup2d_ps *h, size_t bytes, void **ptr) {
  if (!h || !ptr) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  HIPCHK(h, hipSetDevice(h->device));
  HIPCHK(h, hipMalloc(ptr, bytes ? bytes : 1));
  return CUP2D_PS_OK;
}
int cup2d_ps_dfree(cup2d_ps *h, void *ptr) {
  if (!h) return CUP2D_PS_EINVAL;
  HIPCHK(h, hipSetDevice(h->device));
  HIPCHK(h, hipFree(ptr));
  return CUP2D_PS_OK;
}
int cup2d_ps_h2d(cup2d_ps *h, void *dst, const void *src, size_t bytes) {
  if (!h || !dst || !src) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  HIPCHK(h, hipSetDevice(h->device));
  HIPCHK(h, hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return CUP2D_PS_OK;
}
int cup2d_ps_d2h(cup2d_ps *h, void *dst, const void *src, size_t bytes) {
  if (!h || !dst || !src) return fail(h, CUP2D_PS_EINVAL, "NULL argument");
  HIPCHK(h, hipSetDevice(h->device));
  HIPCHK(h, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return CUP2D_PS_OK;
}

const char *cup2d_ps_last_error(const cup2d_ps *h) {
  return h ? h->err.c_str() : g_create_err.c_str();
}

}  // extern "C"
