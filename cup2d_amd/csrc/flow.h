// flow.h -- the operators of the projection time step besides the pressure
// step, on the block mesh (one thread per cell, a wavefront per 8x8 block):
//
//  * advection-diffusion: KernelAdvectDiffuse (source/Operators/advDiff.cpp:
//    140-236), the upwinded WENO5 differences of advDiff.cpp:18-82 and the
//    5-point diffusion, with the two-stage update of advDiff::operator()
//    (advDiff.cpp:238-282);
//  * Brinkman penalization toward a rigid body velocity
//    (PressureSingle::penalize, PressureSingle.cpp:268-322, no deformation);
//  * a disk's characteristic function: signed distance (FillBlocks_Cylinder,
//    ShapeLibrary.cpp:19-34) and Towers' mollified Heaviside
//    (PutObjectsOnGrid.cpp:70-96);
//  * the largest velocity component (findMaxU, Helpers.cpp:155-207).
//
// Ghost cells (three deep across each block face; the stencil is a cross,
// so no corners): a same-level neighbour's cells; at a non-periodic domain
// face the boundary cell for all three layers with the face-normal
// component negated (the velocity lab of Definitions.h:84-117); across a
// coarser face the coarse cell containing the ghost's centre plus its
// gradient (central, one-sided at the coarse block's edges) times the
// offset; across a finer face the average of the four fine cells the ghost
// covers.  The diffusive term is flux-corrected at coarse-fine faces (the
// face values of advDiff.cpp:163-233): a coarse cell's term across a finer
// face is minus the sum of the two fine cells' terms across it.  The
// reference fills coarse-fine ghosts from Cubism's AMR lab, absent here.
// Floating-point contraction is off: every value is the CPU restatement's
// sequence of IEEE operations.
#pragma once

#include "kernels.h"

// (floating-point contraction is switched off inside each function of this
// header, not for the file: it is included by the solver's translation unit)

namespace cup2d {
namespace {

constexpr int LH = 3;                 // ghost layers
constexpr int LW = BS + 2 * LH;       // lab side (14)
constexpr int FWB = WG / 64;          // blocks per workgroup (a wavefront each)

// What the flow kernels read of the mesh.
struct FlowMesh {
  int nb;
  const int4 *nbr;   // same-level neighbour per face (Xm, Xp, Ym, Yp), -1 wall, -2 coarse-fine
  const int2 *fid;   // coarse-fine neighbours per face: (coarse, -1 - half) or (fine low, fine high)
  const double *h;   // cell size per block
  const int2 *bxy;   // block index (bx, by) at its level
};

__device__ __forceinline__ double sqd(double x) { return x * x; }

// advDiff.cpp:18-40 (pow(., 2) as a product)
__device__ __forceinline__ double weno5_plus(double um2, double um1, double u, double up1, double up2) {
#pragma clang fp contract(off)
  const double e = 1e-6;
  const double b1 = 13.0 / 12.0 * sqd((um2 + u) - 2 * um1) + 0.25 * sqd((um2 + 3 * u) - 4 * um1);
  const double b2 = 13.0 / 12.0 * sqd((um1 + up1) - 2 * u) + 0.25 * sqd(um1 - up1);
  const double b3 = 13.0 / 12.0 * sqd((u + up2) - 2 * up1) + 0.25 * sqd((3 * u + up2) - 4 * up1);
  const double what1 = 0.1 / sqd(b1 + e);
  const double what2 = 0.6 / sqd(b2 + e);
  const double what3 = 0.3 / sqd(b3 + e);
  const double aux = 1.0 / ((what1 + what3) + what2);
  const double w1 = what1 * aux, w2 = what2 * aux, w3 = what3 * aux;
  const double f1 = (11.0 / 6.0) * u + ((1.0 / 3.0) * um2 - (7.0 / 6.0) * um1);
  const double f2 = (5.0 / 6.0) * u + ((-1.0 / 6.0) * um1 + (1.0 / 3.0) * up1);
  const double f3 = (1.0 / 3.0) * u + ((+5.0 / 6.0) * up1 - (1.0 / 6.0) * up2);
  return (w1 * f1 + w3 * f3) + w2 * f2;
}
// advDiff.cpp:42-62
__device__ __forceinline__ double weno5_minus(double um2, double um1, double u, double up1, double up2) {
#pragma clang fp contract(off)
  const double e = 1e-6;
  const double b1 = 13.0 / 12.0 * sqd((um2 + u) - 2 * um1) + 0.25 * sqd((um2 + 3 * u) - 4 * um1);
  const double b2 = 13.0 / 12.0 * sqd((um1 + up1) - 2 * u) + 0.25 * sqd(um1 - up1);
  const double b3 = 13.0 / 12.0 * sqd((u + up2) - 2 * up1) + 0.25 * sqd((3 * u + up2) - 4 * up1);
  const double what1 = 0.3 / sqd(b1 + e);
  const double what2 = 0.6 / sqd(b2 + e);
  const double what3 = 0.1 / sqd(b3 + e);
  const double aux = 1.0 / ((what1 + what3) + what2);
  const double w1 = what1 * aux, w2 = what2 * aux, w3 = what3 * aux;
  const double f1 = (1.0 / 3.0) * u + ((-1.0 / 6.0) * um2 + (5.0 / 6.0) * um1);
  const double f2 = (5.0 / 6.0) * u + ((1.0 / 3.0) * um1 - (1.0 / 6.0) * up1);
  const double f3 = (11.0 / 6.0) * u + ((-7.0 / 6.0) * up1 + (1.0 / 3.0) * up2);
  return (w1 * f1 + w3 * f3) + w2 * f2;
}
// fp - fm, upwinded on the sign of U (advDiff.cpp:64-82)
__device__ __forceinline__ double weno_diff(double U, const double (&s)[7]) {
#pragma clang fp contract(off)
  if (U > 0) return weno5_plus(s[1], s[2], s[3], s[4], s[5]) - weno5_plus(s[0], s[1], s[2], s[3], s[4]);
  return weno5_minus(s[2], s[3], s[4], s[5], s[6]) - weno5_minus(s[1], s[2], s[3], s[4], s[5]);
}

// Value at the fine-frame point (X, Y) (fine units, the fine block's cell
// (i, j) centred at (i + 0.5, j + 0.5)) from the coarse block whose cell
// (cx, cy) is centred at (X0 + 2 cx + 1, Y0 + 2 cy + 1): the containing
// cell plus its gradient times the offset.  c(cy, cx) reads the coarse block.
template <class Get>
__device__ __forceinline__ double prolong(Get c, double X, double Y, double X0, double Y0) {
#pragma clang fp contract(off)
  const int cx = (int)floor((X - X0) / 2.0), cy = (int)floor((Y - Y0) / 2.0);
  const double v = c(cy, cx);
  double gx, gy;
  if (cx == 0) gx = (c(cy, 1) - v) / 2.0;
  else if (cx == BS - 1) gx = (v - c(cy, BS - 2)) / 2.0;
  else gx = (c(cy, cx + 1) - c(cy, cx - 1)) / 4.0;
  if (cy == 0) gy = (c(1, cx) - v) / 2.0;
  else if (cy == BS - 1) gy = (v - c(BS - 2, cx)) / 2.0;
  else gy = (c(cy + 1, cx) - c(cy - 1, cx)) / 4.0;
  const double dx = X - (X0 + 2 * cx + 1), dy = Y - (Y0 + 2 * cy + 1);
  return (v + gx * dx) + gy * dy;
}

// Ghost g (0..95: face g / 24, layer d = (g % 24) / 8 + 1, tangential t =
// g % 8) of block b for field f; s = the factor of the wall ghost on this
// face.  Returns the lab position in (gy, gx).
__device__ __forceinline__ double ghost3(const FlowMesh &M, const double *__restrict__ f, long b,
                                         int g, double sx, double sy, int &gy, int &gx) {
#pragma clang fp contract(off)
  const int fc = g / 24, d = (g % 24) / 8 + 1, t = g % 8;
  const bool xf = fc < 2;
  gy = fc == 0 ? LH + t : fc == 1 ? LH + t : fc == 2 ? LH - d : LH + BS - 1 + d;
  gx = fc == 0 ? LH - d : fc == 1 ? LH + BS - 1 + d : LH + t;
  const int4 n4 = M.nbr[b];
  const int nv = fc == 0 ? n4.x : fc == 1 ? n4.y : fc == 2 ? n4.z : n4.w;
  const double *fb = f + b * 64;
  if (nv == -1) {  // domain face
    const double s = xf ? sx : sy;
    const double bv = fc == 0 ? fb[t * 8] : fc == 1 ? fb[t * 8 + 7] : fc == 2 ? fb[t] : fb[56 + t];
    return s * bv;
  }
  if (nv >= 0) {  // same level
    const double *nbk = f + (long)nv * 64;
    return fc == 0 ? nbk[t * 8 + BS - d] : fc == 1 ? nbk[t * 8 + d - 1]
         : fc == 2 ? nbk[(BS - d) * 8 + t] : nbk[(d - 1) * 8 + t];
  }
  const int2 fd = M.fid[4 * b + fc];
  if (fd.y < 0) {  // coarser neighbour: prolongation
    const int half = -1 - fd.y;
    const double *cb = f + (long)fd.x * 64;
    double X, Y, X0, Y0;
    if (xf) {
      X = fc == 0 ? (-d + 0.5) : (BS - 1 + d + 0.5);
      Y = t + 0.5;
      X0 = fc == 0 ? -2 * BS : BS;
      Y0 = -BS * half;
    } else {
      Y = fc == 2 ? (-d + 0.5) : (BS - 1 + d + 0.5);
      X = t + 0.5;
      Y0 = fc == 2 ? -2 * BS : BS;
      X0 = -BS * half;
    }
    return prolong([&](int cy, int cx) { return cb[cy * 8 + cx]; }, X, Y, X0, Y0);
  }
  // finer neighbours: the average of the four fine cells under the ghost
  const double *fk = f + (long)(t >= BS / 2 ? fd.y : fd.x) * 64;
  const int tt = 2 * (t % (BS / 2));
  int r0, c0;
  if (fc == 0) { r0 = tt; c0 = BS - 2 * d; }
  else if (fc == 1) { r0 = tt; c0 = 2 * (d - 1); }
  else if (fc == 2) { r0 = BS - 2 * d; c0 = tt; }
  else { r0 = 2 * (d - 1); c0 = tt; }
  return 0.25 * ((fk[r0 * 8 + c0] + fk[r0 * 8 + c0 + 1]) + (fk[(r0 + 1) * 8 + c0] + fk[(r0 + 1) * 8 + c0 + 1]));
}

// tmp = the RHS of KernelAdvectDiffuse for every cell (u, v in; tu, tv out),
// afac = -dt h, dfac = nu dt; the diffusive flux correction included.
__global__ __launch_bounds__(WG) void kf_advdiff_rhs(FlowMesh M, const double *__restrict__ u,
                                                     const double *__restrict__ v, double uinf_x,
                                                     double uinf_y, double nu, double dt,
                                                     double *__restrict__ tu, double *__restrict__ tv) {
#pragma clang fp contract(off)
  __shared__ double lab[FWB][2][LW * LW];
  const int w = threadIdx.x >> 6, c = threadIdx.x & 63, ix = c & 7, iy = c >> 3;
  for (long g0 = (long)blockIdx.x * FWB; g0 < M.nb; g0 += (long)gridDim.x * FWB) {
    const long b = g0 + w;
    const bool ok = b < M.nb;
    double *LU = lab[w][0], *LV = lab[w][1];
    if (ok) {
      LU[(LH + iy) * LW + LH + ix] = u[b * 64 + c];
      LV[(LH + iy) * LW + LH + ix] = v[b * 64 + c];
      for (int g = c; g < 96; g += 64) {
        int gy, gx;
        const double gu = ghost3(M, u, b, g, -1.0, 1.0, gy, gx);
        const double gv = ghost3(M, v, b, g, 1.0, -1.0, gy, gx);
        LU[gy * LW + gx] = gu;
        LV[gy * LW + gx] = gv;
      }
    }
    __syncthreads();
    if (ok) {
      const double h = M.h[b];
      const double afac = -dt * h, dfac = nu * dt;
      const int o = (LH + iy) * LW + LH + ix;
      const double UU = LU[o] + uinf_x, VV = LV[o] + uinf_y;
      double res[2];
#pragma unroll
      for (int comp = 0; comp < 2; comp++) {
        const double *L = comp ? LV : LU;
        double sx[7], sy[7];
#pragma unroll
        for (int k = 0; k < 7; k++) {
          sx[k] = L[o + k - 3];
          sy[k] = L[o + (k - 3) * LW];
        }
        const double ddx = weno_diff(UU, sx), ddy = weno_diff(VV, sy);
        const double lap = dfac * (((sx[4] + sx[2]) + (sy[4] + sy[2])) - 4 * sx[3]);
        double t = afac * (UU * ddx + VV * ddy) + lap;
        // diffusive flux correction on the faces with finer neighbours
        const int4 n4 = M.nbr[b];
#pragma unroll 1
        for (int fc = 0; fc < 4; fc++) {
          const bool edge = fc == 0 ? ix == 0 : fc == 1 ? ix == BS - 1 : fc == 2 ? iy == 0 : iy == BS - 1;
          const int nv = fc == 0 ? n4.x : fc == 1 ? n4.y : fc == 2 ? n4.z : n4.w;
          if (!edge || nv != -2) continue;
          const int2 fd = M.fid[4 * b + fc];
          if (fd.y < 0) continue;  // coarser neighbour: nothing to correct here
          const double own = L[o];
          const double nbv = fc == 0 ? L[o - 1] : fc == 1 ? L[o + 1] : fc == 2 ? L[o - LW] : L[o + LW];
          double corr = -(dfac * (nbv - own));
          const int tc = fc < 2 ? iy : ix;
          const long fbk = tc >= BS / 2 ? fd.y : fd.x;
          const int t0 = (tc % (BS / 2)) * 2;
          const double *fld = comp ? v : u;
          for (int hh = 0; hh < 2; hh++) {
            const int tf = t0 + hh;
            int fx, fy;
            double X, Y, X0, Y0;
            // the fine cell next to the face, and the point of its ghost
            // across it in the fine block's frame (this block is the coarse one)
            const double half = tc >= BS / 2 ? 1.0 : 0.0;
            if (fc == 0) { fx = BS - 1; fy = tf; X = BS + 0.5; Y = tf + 0.5; X0 = BS; Y0 = -BS * half; }
            else if (fc == 1) { fx = 0; fy = tf; X = -0.5; Y = tf + 0.5; X0 = -2 * BS; Y0 = -BS * half; }
            else if (fc == 2) { fx = tf; fy = BS - 1; Y = BS + 0.5; X = tf + 0.5; Y0 = BS; X0 = -BS * half; }
            else { fx = tf; fy = 0; Y = -0.5; X = tf + 0.5; Y0 = -2 * BS; X0 = -BS * half; }
            const double fown = fld[fbk * 64 + fy * 8 + fx];
            const double gv = prolong([&](int cy, int cx) { return L[(LH + cy) * LW + LH + cx]; }, X, Y, X0, Y0);
            corr = corr - dfac * (gv - fown);
          }
          t = t + corr;
        }
        res[comp] = t;
      }
      tu[b * 64 + c] = res[0];
      tv[b * 64 + c] = res[1];
    }
    __syncthreads();  // the labs are refilled by the next pass
  }
}

// out = base + (fac t) / h^2 (stage 1: fac = 0.5, (0.5 t) ih2; stage 2:
// fac = 1, t ih2), for u and v.
// (uo may be u0: each element is read before it is written)
__global__ __launch_bounds__(WG) void kf_stage(FlowMesh M, const double *u0, const double *v0,
                                               const double *__restrict__ tu,
                                               const double *__restrict__ tv, int half, double *uo,
                                               double *vo) {
#pragma clang fp contract(off)
  const long n = (long)M.nb * 64;
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) {
    const double h = M.h[i >> 6];
    const double ih2 = 1.0 / (h * h);
    if (half) {
      uo[i] = u0[i] + (0.5 * tu[i]) * ih2;
      vo[i] = v0[i] + (0.5 * tv[i]) * ih2;
    } else {
      uo[i] = u0[i] + tu[i] * ih2;
      vo[i] = v0[i] + tv[i] * ih2;
    }
  }
}

// u = alpha u + (1 - alpha) u_s where chi > 0, alpha = 1 / (1 + lambda dt)
// if chi > 0.5 else 1 (PressureSingle.cpp:296-320, one obstacle, no u_def)
__global__ __launch_bounds__(WG) void kf_penalize(long n, const double *__restrict__ chi, double us,
                                                  double vs, double lam, double dt,
                                                  double *__restrict__ u, double *__restrict__ v) {
#pragma clang fp contract(off)
  const double a1 = 1.0 / (1.0 + lam * dt);
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) {
    const double X = chi[i];
    if (!(X > 0)) continue;
    const double alpha = X > 0.5 ? a1 : 1.0;
    u[i] = alpha * u[i] + (1.0 - alpha) * us;
    v[i] = alpha * v[i] + (1.0 - alpha) * vs;
  }
}

// chi of the disk (cx, cy, r): Towers' mollified Heaviside of the signed
// distance within one cell of the circle, 1 / 0 beyond
__global__ __launch_bounds__(WG) void kf_chi_disk(FlowMesh M, double cx, double cy, double r,
                                                  double *__restrict__ chi) {
#pragma clang fp contract(off)
  const long n = (long)M.nb * 64;
  const double EPS = 2.220446049250313e-16;  // DBL_EPSILON
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) {
    const long b = i >> 6;
    const int c = (int)(i & 63);
    const double h = M.h[b];
    const int2 bi = M.bxy[b];
    const double x = ((double)(bi.x * 8 + (c & 7)) + 0.5) * h;
    const double y = ((double)(bi.y * 8 + (c >> 3)) + 0.5) * h;
    auto dist = [&](double px, double py) {
      return r - sqrt((px - cx) * (px - cx) + (py - cy) * (py - cy));
    };
    const double d = dist(x, y);
    double X;
    if (d > h || d < -h) {
      X = d > 0 ? 1.0 : 0.0;
    } else {
      const double dpx = dist(x + h, y), dmx = dist(x - h, y);
      const double dpy = dist(x, y + h), dmy = dist(x, y - h);
      const double gIx = fmax(0.0, dpx) - fmax(0.0, dmx);
      const double gIy = fmax(0.0, dpy) - fmax(0.0, dmy);
      const double gUx = dpx - dmx, gUy = dpy - dmy;
      const double gsq = (gUx * gUx + gUy * gUy) + EPS;
      X = (gIx * gUx + gIy * gUy) / gsq;
    }
    chi[i] = X;
  }
}

// partials of max |u + uinf_x|, |v + uinf_y|, |u|, |v| (findMaxU)
__global__ __launch_bounds__(WG) void kf_max_speed(long n, const double *__restrict__ u,
                                                   const double *__restrict__ v, double uinf_x,
                                                   double uinf_y, double *part) {
#pragma clang fp contract(off)
  double m[4] = {0.0, 0.0, 0.0, 0.0};
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) {
    m[0] = fmax(m[0], fabs(u[i] + uinf_x));
    m[1] = fmax(m[1], fabs(v[i] + uinf_y));
    m[2] = fmax(m[2], fabs(u[i]));
    m[3] = fmax(m[3], fabs(v[i]));
  }
  store_partials<4, 0xFu>(m, part);
}

// ---------------------------------------------------------------------------
// Mesh adaptation (AdaptTheMesh.cpp:11-122): the refinement score of every
// block and the transfer of fields to the adapted mesh.
// ---------------------------------------------------------------------------

// score[b] = max over the block's cells of |omega|, omega = (0.5 / h)
// ((u(x, y-1) - u(x, y+1)) + (v(x+1, y) - v(x-1, y))) (KernelVorticity,
// Helpers.h:103-116, the velocity lab's ghosts as kf_advdiff_rhs); or 2 rtol
// when a cell of the block, or of its lab within `off` cells of it (2, 3 on
// the finest level), has 0 < chi < 0.9 (GradChiOnTmp, AdaptTheMesh.cpp:
// 11-71; chi clamped to [0, 1]; the lab is a cross, so the corner regions
// the reference also scans are not).
__global__ __launch_bounds__(WG) void kf_block_score(FlowMesh M, const double *__restrict__ u,
                                                     const double *__restrict__ v,
                                                     const double *__restrict__ chi, double rtol,
                                                     double h_min, double *__restrict__ score) {
#pragma clang fp contract(off)
  __shared__ double lab[FWB][3][LW * LW];
  const int w = threadIdx.x >> 6, c = threadIdx.x & 63, ix = c & 7, iy = c >> 3;
  for (long g0 = (long)blockIdx.x * FWB; g0 < M.nb; g0 += (long)gridDim.x * FWB) {
    const long b = g0 + w;
    const bool ok = b < M.nb;
    double *LU = lab[w][0], *LV = lab[w][1], *LC = lab[w][2];
    if (ok) {
      const int o = (LH + iy) * LW + LH + ix;
      LU[o] = u[b * 64 + c];
      LV[o] = v[b * 64 + c];
      LC[o] = chi[b * 64 + c];
      for (int g = c; g < 96; g += 64) {
        int gy, gx;
        const double gu = ghost3(M, u, b, g, -1.0, 1.0, gy, gx);
        const double gv = ghost3(M, v, b, g, 1.0, -1.0, gy, gx);
        const double gc = ghost3(M, chi, b, g, 1.0, 1.0, gy, gx);
        LU[gy * LW + gx] = gu;
        LV[gy * LW + gx] = gv;
        LC[gy * LW + gx] = gc;
      }
    }
    __syncthreads();
    if (ok) {
      const double h = M.h[b];
      const double i2h = 0.5 / h;
      const int o = (LH + iy) * LW + LH + ix;
      const double om = i2h * ((LU[o - LW] - LU[o + LW]) + (LV[o + 1] - LV[o - 1]));
      double m = fabs(om);
      // chi interface within off cells (the cross of the lab)
      const int off = h == h_min ? 3 : 2;
      bool hit = false;
      auto test = [&](double x) {
        x = fmin(fmax(x, 0.0), 1.0);
        return x > 0.0 && x < 0.9;
      };
      hit = test(LC[o]);
      if (ix == 0)
        for (int k = 1; k <= off; k++) hit = hit || test(LC[o - k]);
      if (ix == BS - 1)
        for (int k = 1; k <= off; k++) hit = hit || test(LC[o + k]);
      if (iy == 0)
        for (int k = 1; k <= off; k++) hit = hit || test(LC[o - k * LW]);
      if (iy == BS - 1)
        for (int k = 1; k <= off; k++) hit = hit || test(LC[o + k * LW]);
      m = wave_max(m);
      const bool any = __any(hit);
      if (c == 0) score[b] = any ? 2 * rtol : m;
    }
    __syncthreads();
  }
}

// Per new block: src_kind, the level and index of the new block and of its
// source blocks (old mesh), as the transfer kernel reads them.
struct Transfer {
  int nb;                   // new blocks
  const int *kind;          // [nb] 0 copy, 1 refine from an ancestor, 2 compress four children
  const long long *src;     // [4 nb]
  const int *lvl_new;       // [nb]
  const int2 *ij_new;       // [nb]
  const int *lvl_old;       // [old nb]
  const int2 *ij_old;       // [old nb]
};

// out (new mesh) from in (old mesh): a copy; a refinement -- the cell of the
// ancestor containing the new cell's centre plus its gradient (central,
// one-sided at the ancestor's edges) times the offset, exact for linear
// fields and, one level down, keeping each parent cell's mean; or a
// compression -- the average of the four fine cells.
__global__ __launch_bounds__(WG) void kf_transfer(Transfer T, const double *__restrict__ in,
                                                  double *__restrict__ out) {
#pragma clang fp contract(off)
  const long n = (long)T.nb * 64;
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < n; i += (long)gridDim.x * WG) {
    const long b = i >> 6;
    const int c = (int)(i & 63), ix = c & 7, iy = c >> 3;
    const int k = T.kind[b];
    double val;
    if (k == 0) {
      val = in[T.src[4 * b] * 64 + c];
    } else if (k == 1) {
      const long a = T.src[4 * b];
      const int d = T.lvl_new[b] - T.lvl_old[a];
      const int2 nij = T.ij_new[b], oij = T.ij_old[a];
      const double sc = (double)(1 << d);
      // the new cell's centre in the ancestor's cell units
      const double X = ((double)((nij.x - (oij.x << d)) * BS + ix) + 0.5) / sc;
      const double Y = ((double)((nij.y - (oij.y << d)) * BS + iy) + 0.5) / sc;
      const int cx = min(max((int)floor(X), 0), BS - 1), cy = min(max((int)floor(Y), 0), BS - 1);
      const double *A = in + a * 64;
      const double vc = A[cy * 8 + cx];
      double gx, gy;
      if (cx == 0) gx = A[cy * 8 + 1] - vc;
      else if (cx == BS - 1) gx = vc - A[cy * 8 + BS - 2];
      else gx = (A[cy * 8 + cx + 1] - A[cy * 8 + cx - 1]) / 2.0;
      if (cy == 0) gy = A[8 + cx] - vc;
      else if (cy == BS - 1) gy = vc - A[(BS - 2) * 8 + cx];
      else gy = (A[(cy + 1) * 8 + cx] - A[(cy - 1) * 8 + cx]) / 2.0;
      val = (vc + gx * (X - (cx + 0.5))) + gy * (Y - (cy + 0.5));
    } else {
      const int q = (ix >= BS / 2 ? 1 : 0) + (iy >= BS / 2 ? 2 : 0);
      const double *F = in + T.src[4 * b + q] * 64;
      const int fx = 2 * (ix % (BS / 2)), fy = 2 * (iy % (BS / 2));
      val = 0.25 * ((F[fy * 8 + fx] + F[fy * 8 + fx + 1]) + (F[(fy + 1) * 8 + fx] + F[(fy + 1) * 8 + fx + 1]));
    }
    out[i] = val;
  }
}

}  // namespace
}  // namespace cup2d
