// srmhd.h -- special-relativistic adiabatic MHD for the gfx950 kernels.
//
// The equation system of equations.F90:673-760 (nv = 9, GLM psi, c = 1):
// prim2cons_srmhd_adi (equations.F90:4590-4653), the Mignone & Bodo (2006)
// fluxes and characteristic speeds of fluxspeed_srmhd_adi (4821-5101), the
// 1D(W) primitive solver of Noble et al. (2006) with the brackets of
// nr_initial_brackets_srmhd_adi and the iteration of nr_iterate_srmhd_adi_1dw
// (5192-5788), and the polynomial root finders it relies on (algebra.F90:
// 86-1085).  Branches and the order of the operations follow the reference;
// the compiler contracts a*b+c into FMAs, a**(1/3) is cbrt and divisions by
// a shared divisor use the refined reciprocal, which is the whole of the
// rounding difference.
//
// Fluxes run in T (double, or float in the mixed mode); the characteristic
// speeds always run in double, because the quartic of the fast and slow
// speeds loses about half of its digits near multiple roots, and the
// primitive recovery (a Newton iteration to a 1e-10..1e-12 tolerance) is
// double by construction.
#pragma once

#include "physics.h"

namespace amun {

struct SRMHD {
  static constexpr int nv = 9;
  static constexpr int idn = 0, ivx = 1, ivy = 2, ivz = 3;
  static constexpr int imx = 1, imy = 2, imz = 3;
  static constexpr int ibx = 4, iby = 5, ibz = 6, ibp = 7;
  static constexpr int ipr = 8, ien = 8;
  static constexpr bool mhd = true;
  static constexpr bool sr = true;
  static constexpr bool iso = false;
};

// the special-relativistic HD system (equations.F90:537-600): nv = 5, no
// field; its equation of state, primitive solvers and HLLC are the !mhd
// branches below (equations.F90:3478-4566, schemes.F90:4656-4968)
struct SRHD {
  static constexpr int nv = 5;
  static constexpr int idn = 0, ivx = 1, ivy = 2, ivz = 3;
  static constexpr int imx = 1, imy = 2, imz = 3;
  static constexpr int ibx = -1, iby = -1, ibz = -1, ibp = -1;
  static constexpr int ipr = 4, ien = 4;
  static constexpr bool mhd = false;
  static constexpr bool sr = true;
  static constexpr bool iso = false;
};

// constants of the primitive solver ("tolerance", "nr_maxit", "nr_extra",
// "pmin"; equations.F90:184, 196-201)
struct SrConst {
  double gamma, gammaxi, rgammaxi;   // gammaxi = (gamma-1)/gamma
  double tol, pmin;
  int nrmax, nrext;
  int solver;                        // AMUN_C2P_1DW, _2DWV or _2DWU
};

// status bits of one primitive recovery
// (C2P_LOWER: no bracket, the two-variable solvers go on from the lower one)
enum { C2P_FAIL = 1, C2P_PNEG = 2, C2P_NOCONV = 4, C2P_LOWER = 8 };

// ---------------------------------------------------------------------------
// polynomial roots, algebra.F90:86-1085 (coefficients in increasing powers)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double sgn1(double b) { return copysign(1.0, b); }

// sort (algebra.F90:1040-1085): a stable ascending sort of x[0..n-1].  The
// reference's insertion sort is written here as a bubble sort over static
// indices (both are stable, so the result is the same) so that x stays in
// registers.
template <int N>
__device__ __forceinline__ void alg_sort(int n, double* x) {
#pragma unroll
  for (int i = 0; i < N - 1; i++) {
#pragma unroll
    for (int j = N - 1; j > i; j--) {
      if (j < n && x[j - 1] > x[j]) {
        const double t = x[j - 1];
        x[j - 1] = x[j];
        x[j] = t;
      }
    }
  }
}

// x[k] of a root array for a run-time k < N, without dynamic indexing
template <int N>
__device__ __forceinline__ double pick(const double* x, int k) {
  double r = x[0];
#pragma unroll
  for (int i = 1; i < N; i++) r = (k == i) ? x[i] : r;
  return r;
}

// quadratic, algebra.F90:86-242
__device__ __forceinline__ int alg_quadratic(const double* a, double* x) {
  const double mx = fmax(fmax(fabs(a[0]), fabs(a[1])), fabs(a[2]));
  const double b0 = a[0] / mx, b1 = a[1] / mx, b2 = a[2] / mx;
  if (b2 != 0.0) {
    if (b0 != 0.0) {
      const double bh = 0.5 * b1;
      const double dl = bh * bh - b2 * b0;
      if (dl > 0.0) {
        const double dr = sqrt(dl);
        if (b1 > 0.0) {
          const double tm = -bh - dr;
          x[0] = tm / b2;
          x[1] = b0 / tm;
        } else if (b1 < 0.0) {
          const double tm = -bh + dr;
          x[0] = b0 / tm;
          x[1] = tm / b2;
        } else {
          x[1] = dr / b2;
          x[0] = -x[1];
        }
        return 2;
      } else if (dl == 0.0) {
        x[0] = -0.5 * b1 / b2;
        x[1] = x[0];
        return 2;
      }
      x[0] = 0.0;
      x[1] = 0.0;
      return 0;
    }
    const double tm = -b1 / b2;
    if (tm >= 0.0) { x[0] = 0.0; x[1] = tm; }
    else           { x[0] = tm;  x[1] = 0.0; }
    return 2;
  }
  x[1] = 0.0;
  if (b1 != 0.0) { x[0] = -b0 / b1; return 1; }
  x[0] = 0.0;
  return 0;
}

// quadratic_normalized, algebra.F90:262-372
__device__ __forceinline__ int alg_quadratic_normalized(const double* a, double* x) {
  if (a[0] != 0.0) {
    const double bh = 0.5 * a[1];
    const double dl = bh * bh - a[0];
    if (dl > 0.0) {
      const double dr = sqrt(dl);
      if (a[1] > 0.0) {
        const double tm = bh + dr;
        x[0] = -a[0] / tm;
        x[1] = -tm;
      } else if (a[1] < 0.0) {
        const double tm = bh - dr;
        x[0] = -tm;
        x[1] = -a[0] / tm;
      } else {
        x[0] = dr;
        x[1] = -dr;
      }
      return 2;
    } else if (dl == 0.0) {
      x[0] = -bh;
      x[1] = -bh;
      return 2;
    }
    x[0] = 0.0;
    x[1] = 0.0;
    return 0;
  }
  x[0] = 0.0;
  x[1] = -a[1];
  return 2;
}

constexpr double kPi2 = 6.28318530717958647692528676655900;   // constants.F90:40
constexpr double kC0 = 1.0 / 3.0;

// cubic, algebra.F90:398-614.  The three-real-root branch takes
// sqrt(qc) = sqrt(bc)/(3|a3|), as the reference does, which is right for the
// positive leading coefficient of its only caller (quartic: b3 = 8).
__device__ __forceinline__ int alg_cubic(const double* a, double* x) {
  int nr;
  if (a[3] != 0.0) {
    if (a[0] != 0.0) {
      double ac = (2.0 * (a[2] * a[2] * a[2]) - 9.0 * a[3] * a[2] * a[1]) +
                  27.0 * (a[3] * a[3]) * a[0];
      double bc = a[2] * a[2] - 3.0 * a[3] * a[1];
      const double a2 = ac * ac;
      const double b3 = bc * bc * bc;
      double qc = a2 - 4.0 * b3;
      if (qc > 0.0) {
        const double cc = 0.5 * (sqrt(qc) + ac);
        const double dc = cbrt(cc);
        double ec;
        if (cc == 0.0) {
          ec = b3 / a2;
          ec = 2.0 * ec * (1.0 / sqrt(1.0 - 4.0 * ec) + 1.0 / ac);
        } else {
          ec = b3 / cc;
        }
        ec = cbrt(ec);
        x[0] = -(kC0 * ((a[2] + dc) + ec)) / a[3];
        x[1] = 0.0;
        x[2] = 0.0;
        nr = 1;
      } else if (qc == 0.0) {
        if (bc != 0.0) {
          x[0] = 0.5 * (9.0 * a[3] * a[0] - a[2] * a[1]) / bc;
          x[1] = x[0];
          x[2] = ((-(9.0 * a[3] * a[0]) + 4.0 * a[2] * a[1]) - (a[2] * a[2] * a[2]) / a[3]) / bc;
        } else {
          x[0] = x[1] = x[2] = -(kC0 * a[2]) / a[3];
        }
        nr = 3;
      } else {
        double th = 3.0 * a[3];
        qc = bc / (th * th);
        const double rc = 0.5 * ac / (th * th * th);
        const double fb = fabs(bc);
        const double cs = copysign(0.5, ac * bc) * sqrt((ac * ac) / (fb * fb * fb));
        if (cs <= 1.0) {
          th = acos(cs);
          ac = -2.0 * sqrt(qc);
          bc = -(kC0 * (a[2] / a[3]));
          x[0] = ac * cos(kC0 * th) + bc;
          x[1] = ac * cos(kC0 * (th + kPi2)) + bc;
          x[2] = ac * cos(kC0 * (th - kPi2)) + bc;
          nr = 3;
        } else {
          ac = -cbrt(copysign(fabs(rc) + sqrt(rc * rc - qc * qc * qc), rc));
          bc = (ac != 0.0) ? qc / ac : 0.0;
          x[0] = (ac + bc) - kC0 * (a[2] / a[3]);
          x[1] = 0.0;
          x[2] = 0.0;
          nr = 1;
        }
      }
    } else {
      x[2] = 0.0;
      nr = alg_quadratic(a + 1, x) + 1;
    }
  } else {
    x[2] = 0.0;
    nr = alg_quadratic(a, x);
  }
  alg_sort<3>(nr, x);
  return nr;
}

// cubic_normalized, algebra.F90:636-796
__device__ __forceinline__ int alg_cubic_normalized(const double* a, double* x) {
  constexpr double c1 = 1.0 / 2.7e+01, c2 = 1.0 / 4.0;
  int nr;
  if (a[0] != 0.0) {
    const double ac = kC0 * (a[2] * a[2] - 3.0 * a[1]);
    const double bc = c1 * ((2.0 * (a[2] * a[2] * a[2]) - 9.0 * a[2] * a[1]) + 2.7e+01 * a[0]);
    const double a3 = c1 * (ac * ac * ac);
    const double b2 = c2 * (bc * bc);
    const double qc = b2 - a3;
    const double sh = kC0 * a[2];
    if (qc > 0.0) {
      const double sq = sqrt(qc);
      const double aa = cbrt(-0.5 * bc + sq);
      const double bb = cbrt(-0.5 * bc - sq);
      x[0] = (aa + bb) - sh;
      x[1] = 0.0;
      x[2] = 0.0;
      nr = 1;
    } else if (qc < 0.0) {
      double cs;
      if (bc > 0.0)      cs = -sqrt(b2 / a3);
      else if (bc < 0.0) cs = sqrt(b2 / a3);
      else               cs = 0.0;
      const double ph = acos(cs);
      const double aa = 2.0 * sqrt(kC0 * ac);
      x[0] = aa * cos(kC0 * (ph - kPi2)) - sh;
      x[1] = aa * cos(kC0 * ph) - sh;
      x[2] = aa * cos(kC0 * (ph + kPi2)) - sh;
      nr = 3;
    } else {
      if (bc > 0.0) {
        const double aa = sqrt(kC0 * ac);
        x[0] = -2.0 * aa; x[1] = aa; x[2] = aa;
      } else if (bc < 0.0) {
        const double aa = sqrt(kC0 * ac);
        x[0] = 2.0 * aa; x[1] = -aa; x[2] = -aa;
      } else {
        x[0] = x[1] = x[2] = 0.0;
      }
      x[0] -= sh; x[1] -= sh; x[2] -= sh;
      nr = 3;
    }
  } else {
    x[2] = 0.0;
    nr = alg_quadratic_normalized(a + 1, x) + 1;
  }
  alg_sort<3>(nr, x);
  return nr;
}

// quartic, algebra.F90:820-1017: Hacke's method through the resolvent cubic,
// taking its smallest root z(1)
__device__ __forceinline__ int alg_quartic(const double* a, double* x) {
  int nr;
  if (a[4] != 0.0) {
    if (a[0] != 0.0) {
      const double a42 = a[3] * a[3], a52 = a[4] * a[4], a53 = a52 * a[4];
      const double pc = (-(3.0 * a42) + 8.0 * a[4] * a[2]) / (8.0 * a52);
      const double qc = ((a42 * a[3] - 4.0 * a[4] * a[3] * a[2]) + 8.0 * a52 * a[1]) / (8.0 * a53);
      const double rc = (((-(3.0 * (a42 * a[3] * a[3])) + 16.0 * a[4] * a42 * a[2]) -
                          64.0 * a52 * a[3] * a[1]) + 256.0 * a53 * a[0]) /
                        (256.0 * (a53 * a[4]));
      double b[4], z[3];
      b[3] = 8.0;
      b[2] = -4.0 * pc;
      b[1] = -8.0 * rc;
      b[0] = 4.0 * pc * rc - qc * qc;
      nr = alg_cubic(b, z);
      x[0] = x[1] = x[2] = x[3] = 0.0;
      if (nr > 0) {
        const double zm = z[0];
        const double k2 = 2.0 * zm - pc;
        const double l2 = zm * zm - rc;
        if (k2 > 0.0 || l2 > 0.0) {
          double kc, lc;
          if (k2 > l2) {
            kc = sqrt(k2);
            lc = 0.5 * qc / kc;
          } else if (l2 > k2) {
            lc = sqrt(l2);
            kc = 0.5 * qc / lc;
          } else {
            kc = sqrt(k2);
            lc = sqrt(l2);
          }
          const double sh = 0.25 * a[3] / a[4];
          double c[3], y[2], y2[2];
          c[2] = 1.0; c[1] = -kc; c[0] = zm + lc;
          const int m1 = alg_quadratic(c, y);
          c[2] = 1.0; c[1] = kc; c[0] = zm - lc;
          const int m2 = alg_quadratic(c, y2);
          // roots of the first factor, then those of the second, then zeros
          const double r2 = (m2 > 0) ? y2[0] - sh : 0.0;
          const double r3 = (m2 > 1) ? y2[1] - sh : 0.0;
          if (m1 == 2) {
            x[0] = y[0] - sh; x[1] = y[1] - sh; x[2] = r2; x[3] = r3;
          } else if (m1 == 1) {
            x[0] = y[0] - sh; x[1] = r2; x[2] = r3;
          } else {
            x[0] = r2; x[1] = r3;
          }
          nr = m1 + m2;
        } else {
          nr = 0;
        }
      }
    } else {
      x[3] = 0.0;
      nr = alg_cubic(a + 1, x) + 1;
    }
  } else {
    x[3] = 0.0;
    nr = alg_cubic(a, x);
  }
  alg_sort<4>(nr, x);
  return nr;
}

// ---------------------------------------------------------------------------
// equation of state
// ---------------------------------------------------------------------------
// prim2cons_srmhd_adi, equations.F90:4590-4653
template <class P, typename T>
__device__ __forceinline__ void prim2cons_sr(T rgammaxi, const T* q, T* u) {
  if constexpr (!P::mhd) {
    // prim2cons_srhd_adi, equations.F90:3478-3532
    const T vv = (q[P::ivx] * q[P::ivx] + q[P::ivy] * q[P::ivy]) + q[P::ivz] * q[P::ivz];
    const T vm = T(1) - vv;
    const T vs = tsqrt(vm);
    const T ww = (q[P::idn] + q[P::ipr] * rgammaxi) * rcp(vm);
    u[P::idn] = q[P::idn] * rcp(vs);
    u[P::imx] = ww * q[P::ivx];
    u[P::imy] = ww * q[P::ivy];
    u[P::imz] = ww * q[P::ivz];
    u[P::ien] = (ww - q[P::ipr]) - u[P::idn];
    return;
  } else {
  const T vv = (q[P::ivx] * q[P::ivx] + q[P::ivy] * q[P::ivy]) + q[P::ivz] * q[P::ivz];
  const T bb = (q[P::ibx] * q[P::ibx] + q[P::iby] * q[P::iby]) + q[P::ibz] * q[P::ibz];
  const T vb = (q[P::ivx] * q[P::ibx] + q[P::ivy] * q[P::iby]) + q[P::ivz] * q[P::ibz];
  const T vm = T(1) - vv;
  const T vs = tsqrt(vm);
  const T ww = (q[P::idn] + q[P::ipr] * rgammaxi) * rcp(vm);
  const T wt = ww + bb;
  u[P::idn] = q[P::idn] * rcp(vs);
  u[P::imx] = wt * q[P::ivx] - vb * q[P::ibx];
  u[P::imy] = wt * q[P::ivy] - vb * q[P::iby];
  u[P::imz] = wt * q[P::ivz] - vb * q[P::ibz];
  u[P::ibx] = q[P::ibx];
  u[P::iby] = q[P::iby];
  u[P::ibz] = q[P::ibz];
  u[P::ibp] = q[P::ibp];
  u[P::ien] = ((wt - q[P::ipr]) - u[P::idn]) - T(0.5) * (vm * bb + vb * vb);
  }
}

// characteristic speeds of fluxspeed_srmhd_adi (equations.F90:4895-5083), in
// double: the quartic (Bx != 0), the quadratic (Bx = 0) or the bi-quadratic
// (v = 0) of the fast/slow magnetosonic waves; -1/+1 when fewer than two real
// roots come out
__device__ __forceinline__ void sr_speeds(const SrConst& s, double rho, double vx,
                                          double vv, double vm, double vb, double b2,
                                          double bxn, double pr, double& cm, double& cp) {
  double a[5], x[4] = {0.0, 0.0, 0.0, 0.0};
  int nr;
  const double rh = rho + pr * s.rgammaxi;
  const double vs = sqrt(vm);
  const double c2 = s.gamma * pr / rh;
  if (vv > 0.0) {
    if (bxn != 0.0) {
      const double v1 = fabs(vx);
      const double v2 = v1 * v1;
      const double fa = rh * (1.0 - c2);
      const double fb = c2 * (rh - vb * vb) + b2;
      const double fc = sgn1(vx) * bxn * vs;
      const double fd = c2 * fc * fc;
      const double fe = 1.0 - v2;
      const double ff = c2 * vb * fc;
      const double fg = v1 * vs;
      a[4] = fa + fb * vm;
      a[3] = 2.0 * (ff * vm + fb * fg);
      a[2] = (-(fd * vm) + 4.0 * ff * fg) - fb * fe;
      a[1] = -(2.0 * (fd * fg + fe * ff));
      a[0] = fd * fe;
      nr = alg_quartic(a, x);
      const double sv = sgn1(vx);
#pragma unroll
      for (int n = 0; n < 4; n++)
        if (n < nr) x[n] = sv * (v1 + x[n] * vs);
    } else {
      const double cc = (1.0 - c2) / vm;
      const double gn = b2 - c2 * vb * vb;
      a[2] = rh * (c2 + cc) + gn;
      a[1] = -(2.0 * rh * cc * vx);
      a[0] = rh * (cc * (vx * vx) - c2) - gn;
      nr = alg_quadratic(a, x);
    }
  } else {
    const double rt = rh + b2;
    double ca = bxn * vs + vb * vx / vs;
    ca = ca * ca;
    a[2] = 1.0;
    a[1] = -(((rh + ca) * c2 + b2) / rt);
    a[0] = c2 * ca / rt;
    nr = alg_quadratic(a, x);
    if (nr > 0) {
      double zm = fmin(x[0], x[1]);
      double zp = fmax(x[0], x[1]);
      if (zm >= 0.0) {
        zm = sqrt(zm);
        zp = sqrt(zp);
        x[0] = zp; x[1] = zm; x[2] = -zm; x[3] = -zp;
        nr = 4;
      } else if (zp >= 0.0) {
        zp = sqrt(zp);
        x[0] = zp; x[1] = -zp;
        nr = 2;
      } else {
        nr = 0;
      }
    }
  }
  if (nr > 1) {
    double mn = x[0], mx = x[0];
#pragma unroll
    for (int n = 1; n < 4; n++)
      if (n < nr) { mn = fmin(mn, x[n]); mx = fmax(mx, x[n]); }
    cm = mn;
    cp = mx;
  } else {
    cm = -1.0;
    cp = 1.0;
  }
}

// fluxspeed_srmhd_adi, equations.F90:4821-5101: fluxes (eq. 13 of Mignone &
// Bodo 2006) in T, speeds in double
template <class P, typename T>
__device__ __forceinline__ void fluxspeed_sr(const SrConst& s, T cmax2, const T* q, const T* u,
                                             T* f, T& cm, T& cp) {
  if constexpr (!P::mhd) {
    // fluxspeed_srhd_adi, equations.F90:3665-3731 (speeds in double)
    f[P::idn] = u[P::idn] * q[P::ivx];
    f[P::imx] = u[P::imx] * q[P::ivx] + q[P::ipr];
    f[P::imy] = u[P::imy] * q[P::ivx];
    f[P::imz] = u[P::imz] * q[P::ivx];
    f[P::ien] = u[P::imx] - f[P::idn];
    const double dn = q[P::idn], pr = q[P::ipr];
    const double vx = q[P::ivx], vy = q[P::ivy], vz = q[P::ivz];
    const double ww = dn + pr * s.rgammaxi;
    const double c2 = s.gamma * pr / ww;
    const double vv = (vx * vx + vy * vy) + vz * vz;
    const double ss = c2 * (1.0 - vv) / (1.0 - c2);
    const double fc = 1.0 + ss;
    const double cc = sqrt(ss * (fc - vx * vx));
    cm = T((vx - cc) / fc);
    cp = T((vx + cc) / fc);
    return;
  } else {
  const T vv = (q[P::ivx] * q[P::ivx] + q[P::ivy] * q[P::ivy]) + q[P::ivz] * q[P::ivz];
  const T bb = (q[P::ibx] * q[P::ibx] + q[P::iby] * q[P::iby]) + q[P::ibz] * q[P::ibz];
  const T vb = (q[P::ivx] * q[P::ibx] + q[P::ivy] * q[P::iby]) + q[P::ivz] * q[P::ibz];
  const T vm = T(1) - vv;
  const T bx = q[P::ibx] * vm + vb * q[P::ivx];
  const T by = q[P::iby] * vm + vb * q[P::ivy];
  const T bz = q[P::ibz] * vm + vb * q[P::ivz];
  const T b2 = bb * vm + vb * vb;
  const T pt = q[P::ipr] + T(0.5) * b2;
  f[P::idn] = u[P::idn] * q[P::ivx];
  f[P::imx] = (u[P::imx] * q[P::ivx] - q[P::ibx] * bx) + pt;
  f[P::imy] = u[P::imy] * q[P::ivx] - q[P::ibx] * by;
  f[P::imz] = u[P::imz] * q[P::ivx] - q[P::ibx] * bz;
  f[P::ibx] = q[P::ibp];
  f[P::ibp] = cmax2 * q[P::ibx];
  f[P::iby] = q[P::ivx] * q[P::iby] - q[P::ibx] * q[P::ivy];
  f[P::ibz] = q[P::ivx] * q[P::ibz] - q[P::ibx] * q[P::ivz];
  f[P::ien] = u[P::imx] - f[P::idn];
  double dcm, dcp;
  if constexpr (sizeof(T) == 8) {
    sr_speeds(s, q[P::idn], q[P::ivx], vv, vm, vb, b2, q[P::ibx], q[P::ipr], dcm, dcp);
  } else {
    // the speed polynomial from the state promoted to double
    const double dvx = q[P::ivx], dvy = q[P::ivy], dvz = q[P::ivz];
    const double dbx = q[P::ibx], dby = q[P::iby], dbz = q[P::ibz];
    const double dvv = (dvx * dvx + dvy * dvy) + dvz * dvz;
    const double dbb = (dbx * dbx + dby * dby) + dbz * dbz;
    const double dvb = (dvx * dbx + dvy * dby) + dvz * dbz;
    const double dvm = 1.0 - dvv;
    sr_speeds(s, q[P::idn], dvx, dvv, dvm, dvb, dbb * dvm + dvb * dvb, dbx, q[P::ipr], dcm, dcp);
  }
  cm = T(dcm);
  cp = T(dcp);
  }
}

// ---------------------------------------------------------------------------
// primitive recovery, 1D(W) (equations.F90:5192-5788)
// ---------------------------------------------------------------------------
// nr_function_srmhd_adi_1d: F(W) and dF/dW
__device__ __forceinline__ void nr_function(const SrConst& s, double mm, double bb, double mb,
                                            double en, double dn, double w, double& f,
                                            double* df) {
  const double ww = w * w;
  const double wt = w + bb;
  const double wd = wt * wt;
  const double ss = mb * mb;
  const double sw = ss / ww;
  const double ws = (w + wt) * sw;
  const double fn = (mm * bb - ss) / wd;
  const double dvv = (wd - mm) - ws;
  const double ds = sqrt(dvv);
  const double gm2 = wd / dvv;
  const double gm = wt / ds;
  const double pr = s.gammaxi * (w - gm * dn) / gm2;
  f = ((w - pr) - en) + 0.5 * (bb + fn);
  if (df) {
    const double dg = (1.0 - wt * ((wt - sw) + ws / w) / dvv) / ds;
    const double dp = s.gammaxi * (1.0 - (2.0 * w / gm - dn) * dg) / gm2;
    *df = (1.0 - dp) - fn / wt;
  }
}

// nr_velocity_srmhd_adi_1d (5249-5290): v^2 of W
__device__ __forceinline__ double nr_velocity(double mm, double bb, double mb, double w) {
  const double ss = mb * mb, ww = w * w, wt = w + bb;
  return (mm * ww + (w + wt) * ss) / (ww * (wt * wt));
}

// nr_initial_brackets_srmhd_adi (5459-5635): the bracket [wl, wu] of the
// root of F(W), F at its ends and the secant start w; false without one
// (wl then holds the last lower bracket tried)
__device__ __forceinline__ bool nr_brackets(const SrConst& s, double mm, double bb, double mb,
                                            double en, double dn, double& wl, double& wu,
                                            double& w, double& fl, double& fu) {
  const double dd = dn * dn;
  const double ss = mb * mb;
  const double ec = en + s.pmin;
  wu = en + s.pmin;
  // lower bracket: the largest roots of the cubic (pressure) and quartic
  // (1 - v^2) positivity conditions
  {
    double a[5], x[4];
    a[2] = 2.5 * bb - ec;
    a[1] = 2.0 * (bb - ec) * bb;
    a[0] = 0.5 * (((bb - 2.0 * ec) * bb + mm) * bb - ss);
    int nr = alg_cubic_normalized(a, x);
    if (nr > 0) {
      wl = pick<3>(x, nr - 1);
      a[4] = 1.0;
      a[3] = 2.0 * bb;
      a[2] = (bb * bb - dd) - mm;
      a[1] = -(2.0 * (ss + dd * bb));
      a[0] = -(ss + dd * bb) * bb;
      nr = alg_quartic(a, x);
      if (nr > 0) wl = tmax(wl, pick<4>(x, nr - 1));
    } else {
      // Newton on the pressure from the upper bound (never reached: the
      // normalised cubic always returns a root)
      int it = s.nrmax;
      bool keep = true;
      wl = wu;
      while (keep) {
        const double wt = wl + bb, fn = (mm * bb - ss) / (wt * wt);
        const double f = (wl - ec) + 0.5 * (bb + fn), df = 1.0 - fn / wt;
        const double dw = f / df;
        wl = wl - dw;
        const double err = fabs(dw / wl);
        it = it - 1;
        keep = (err > s.tol) && it > 0;
      }
    }
  }
  nr_function(s, mm, bb, mb, en, dn, wl, fl, nullptr);
  if (!(fl <= 0.0)) return false;
  if (wu <= wl) wu = 2.0 * wl;
  nr_function(s, mm, bb, mb, en, dn, wu, fu, nullptr);
  {
    int it = s.nrmax;
    bool keep = fl * fu > 0.0;
    while (keep) {
      it = it - 1;
      wl = wu;
      fl = fu;
      wu = 2.0 * wu;
      nr_function(s, mm, bb, mb, en, dn, wu, fu, nullptr);
      keep = (fl * fu > 0.0) && it > 0;
    }
    if (!(it > 0 && fu >= 0.0)) return false;
  }
  w = wl - fl * (wu - wl) / (fu - fl);
  return true;
}

// nr_iterate_srmhd_adi_1dw (5665-5788).  Returns the status bits; on
// success w and vv hold W and v^2.
__device__ __forceinline__ int nr_iterate_1dw(const SrConst& s, double mm, double bb,
                                              double mb, double en, double dn, double& w,
                                              double& vv) {
  double wl, wu, fl, fu;
  if (!nr_brackets(s, mm, bb, mb, en, dn, wl, wu, w, fl, fu)) return C2P_FAIL;
  // Newton-Raphson with bracket updates and bisection when outside
  int it = s.nrmax, cn = s.nrext;
  bool keep = true;
  double err = 0.0;
  while (keep) {
    double f, df;
    nr_function(s, mm, bb, mb, en, dn, w, f, &df);
    if (f > fl && f < 0.0) { wl = w; fl = f; }
    if (f < fu && f > 0.0) { wu = w; fu = f; }
    const double dw = f / df;
    w = w - dw;
    err = fabs(dw / w);
    if (err < s.tol) {
      keep = cn > 0;
      cn = cn - 1;
    } else {
      keep = it > 0;
    }
    if (w < wl || w > wu) w = 0.5 * (wl + wu);
    it = it - 1;
  }
  vv = nr_velocity(mm, bb, mb, w);
  return err >= s.tol ? C2P_NOCONV : 0;
}

// nr_iterate_srmhd_adi_2dwv (5819-6000) and _2dwu (6020-6210): Newton on
// (W, v^2), or on (W, u^2 = v^2 / (1 - v^2)), from the bracketed start of
// 1D(W), with the 2x2 system solved by its inverse.  Without a bracket the
// lower one is the solution (C2P_LOWER); an iterate below it or with an
// unphysical speed fails.
template <bool WU>
__device__ __forceinline__ int nr_iterate_2d(const SrConst& s, double mm, double bb, double mb,
                                             double en, double dn, double& w, double& vv) {
  const double gx = s.gammaxi;
  double wl, wu, fl, fu;
  if (!nr_brackets(s, mm, bb, mb, en, dn, wl, wu, w, fl, fu)) {
    w = wl;
    vv = nr_velocity(mm, bb, mb, w);
    return C2P_LOWER;
  }
  vv = nr_velocity(mm, bb, mb, w);
  double x = WU ? vv / (1.0 - vv) : vv, err = 0.0;
  int it = s.nrmax, cn = s.nrext;
  bool keep = true;
  while (keep) {
    double f, g, dfw, dfx, dgw, dgx;
    if constexpr (!WU) {
      const double mw = (mb / w) * (mb / w), wt = w + bb, wt2 = wt * wt;
      const double vm = 1.0 - x, vs = sqrt(vm);
      f = ((w - en) - gx * (w * vm - dn * vs)) + 0.5 * ((1.0 + x) * bb - mw);
      g = (x * wt2 - (wt + w) * mw) - mm;
      dfw = (1.0 - gx * vm) + mw / w;
      dfx = -(gx * (0.5 * dn / vs - w)) + 0.5 * bb;
      dgw = 2.0 * wt * (x + mw / w);
      dgx = wt2;
    } else {
      const double up = 1.0 + x, gm = sqrt(up), ss = mb * mb, ww = w * w;
      const double wt = w + bb, wd = w - gm * dn, wp = wt / up;
      f = ((w - en) - gx * wd / up) + 0.5 * (bb * (1.0 + x / up) - ss / ww);
      g = (wp * wt * x - (w + wt) * ss / ww) - mm;
      dfw = (1.0 - gx / up) + ss / ww / w;
      dfx = 0.5 * (gx * (w + wd) + bb) / (up * up);
      dgw = 2.0 * wt * (x / up + ss / ww / w);
      dgx = wp * wp;
    }
    const double det = dfw * dgx - dfx * dgw;
    const double jfw = dgx / det, jgw = -dfx / det, jfx = -dgw / det, jgx = dfw / det;
    const double dw = f * jfw + g * jgw, dx = f * jfx + g * jgx;
    w = w - dw;
    x = x - dx;
    if (w < wl) return C2P_FAIL;
    if (WU ? x < 0.0 : (x < 0.0 || x >= 1.0)) return C2P_FAIL;
    err = tmax(fabs(dw / w), fabs(dx));
    if (err < s.tol) {
      keep = cn > 0;
      cn = cn - 1;
    } else {
      keep = it > 0;
    }
    it = it - 1;
  }
  vv = WU ? x / (1.0 + x) : x;
  return err >= s.tol ? C2P_NOCONV : 0;
}


// the face states of the four-velocity reconstruction turned back into rho
// and v (schemes.F90:4505-4517, 5064-5076): W = sqrt(1 + |W v|^2)
template <class P, typename T>
__device__ __forceinline__ void states_from_4vec(T* f) {
  const T vm = tsqrt(T(1) + ((f[P::ivx] * f[P::ivx] + f[P::ivy] * f[P::ivy]) +
                             f[P::ivz] * f[P::ivz]));
  f[P::idn] = f[P::idn] / vm;
  f[P::ivx] = f[P::ivx] / vm;
  f[P::ivy] = f[P::ivy] / vm;
  f[P::ivz] = f[P::ivz] / vm;
}

// the four-velocity states of the primitive pool (schemes.F90:4481-4496,
// 5040-5056): rho W and W v with W = 1 / sqrt(1 - v^2), the rest copied
template <class P, typename T>
static __global__ void __launch_bounds__(256)
k_to_4vec(const T* __restrict__ Q, T* __restrict__ Q4, long long ncells, int ncell) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ncells) return;
  const long long b = g / ncell, c = g % ncell;
  const size_t base = (size_t)b * P::nv * ncell + c;
  T q[P::nv];
#pragma unroll
  for (int v = 0; v < P::nv; v++) q[v] = Q[base + (size_t)v * ncell];
  const T vm = tsqrt(T(1) - ((q[P::ivx] * q[P::ivx] + q[P::ivy] * q[P::ivy]) +
                             q[P::ivz] * q[P::ivz]));
  q[P::idn] = q[P::idn] / vm;
  q[P::ivx] = q[P::ivx] / vm;
  q[P::ivy] = q[P::ivy] / vm;
  q[P::ivz] = q[P::ivz] / vm;
#pragma unroll
  for (int v = 0; v < P::nv; v++) Q4[base + (size_t)v * ncell] = q[v];
}

// ---------------------------------------------------------------------------
// primitive recovery of the relativistic HD system (equations.F90:3550-4566)
// ---------------------------------------------------------------------------
// nr_function_srhd_adi_1d (3853-3881): F(W) and dF/dW
__device__ __forceinline__ void nr_function_srhd(const SrConst& s, double mm, double en,
                                                 double dn, double w, double& f, double* df) {
  const double gx = s.gammaxi;
  const double vv = mm / (w * w), vm = 1.0 - vv, vs = sqrt(vm);
  f = ((1.0 - gx * vm) * w + gx * dn * vs) - en;
  if (df) *df = 1.0 - gx * (1.0 + (1.0 - dn / (vs * w)) * vv);
}

// nr_iterate_srhd_adi_1dw (3911-4091): brackets from sqrt(|m|^2 + D^2) +
// gammaxi pmin and E + pmin (moved up while F keeps its sign), a secant
// start, bracketed Newton.  Where the lower bracket is an exact root the
// reference leaves v^2 undefined; here it is |m|^2 / W^2 of that root.
__device__ __forceinline__ int nr_iterate_srhd_1dw(const SrConst& s, double mm, double en,
                                                   double dn, double& w, double& vv) {
  double wl = sqrt(mm + dn * dn), wu, fl, fu;
  if (!(en > wl)) return C2P_FAIL;
  wl = wl + s.gammaxi * s.pmin;
  wu = en + s.pmin;
  nr_function_srhd(s, mm, en, dn, wl, fl, nullptr);
  if (!(fl < 0.0)) {
    if (fl == 0.0) {
      w = wl;
      vv = mm / (wl * wl);
      return 0;
    }
    return C2P_FAIL;
  }
  if (wu <= wl) wu = 2.0 * wl;
  nr_function_srhd(s, mm, en, dn, wu, fu, nullptr);
  int it = s.nrmax;
  bool keep = fl * fu > 0.0;
  while (keep) {
    it = it - 1;
    wl = wu;
    fl = fu;
    wu = 2.0 * wu;
    nr_function_srhd(s, mm, en, dn, wu, fu, nullptr);
    keep = (fl * fu > 0.0) && it > 0;
  }
  if (!(it > 0 && fu >= 0.0)) return C2P_FAIL;
  w = wl - fl * (wu - wl) / (fu - fl);
  it = s.nrmax;
  int cn = s.nrext;
  keep = true;
  double err = 0.0;
  while (keep) {
    double f, df;
    nr_function_srhd(s, mm, en, dn, w, f, &df);
    if (f > fl && f < 0.0) { wl = w; fl = f; }
    if (f < fu && f > 0.0) { wu = w; fu = f; }
    const double dw = f / df;
    w = w - dw;
    err = fabs(dw / w);
    if (err < s.tol) {
      keep = cn > 0;
      cn = cn - 1;
    } else {
      keep = it > 0;
    }
    if (w < wl || w > wu) w = 0.5 * (wl + wu);
    it = it - 1;
  }
  vv = mm / (w * w);
  return err >= s.tol ? C2P_NOCONV : 0;
}

// nr_iterate_srhd_adi_2dwv (4128-4330) and _2dwu (4360-4566): the 1D(W)
// brackets, a secant start, then Newton on (W, v^2) or (W, u^2) with the
// 2x2 system solved by its inverse; an iterate below the lower bracket or
// with an unphysical speed fails
template <bool WU>
__device__ __forceinline__ int nr_iterate_srhd_2d(const SrConst& s, double mm, double en,
                                                  double dn, double& w, double& vv) {
  const double gx = s.gammaxi;
  double wl = sqrt(mm + dn * dn) + gx * s.pmin, wu = en + s.pmin, fl, fu;
  int it = s.nrmax;
  bool keep = wl >= wu;
  while (keep) {
    wu = 2.0 * wu;
    it = it - 1;
    keep = (wl >= wu) && it > 0;
  }
  if (it <= 0) return C2P_FAIL;
  nr_function_srhd(s, mm, en, dn, wl, fl, nullptr);
  nr_function_srhd(s, mm, en, dn, wu, fu, nullptr);
  keep = fl * fu > 0.0;
  it = s.nrmax;
  while (keep) {
    wl = wu;
    fl = fu;
    wu = 2.0 * wu;
    nr_function_srhd(s, mm, en, dn, wu, fu, nullptr);
    it = it - 1;
    keep = (fl * fu > 0.0) && it > 0;
  }
  if (it <= 0) return C2P_FAIL;
  w = wl - fl * (wu - wl) / (fu - fl);
  double x = WU ? mm / (w * w - mm) : mm / (w * w), err = 0.0;
  it = s.nrmax;
  int cn = s.nrext;
  keep = true;
  while (keep) {
    const double gd = gx * dn, ww = w * w;
    double f, g, dfw, dfx, dgw, dgx;
    if constexpr (!WU) {
      const double vm = 1.0 - x, vs = sqrt(vm), gv = 1.0 - gx * vm;
      f = (gv * w - en) + gd * vs;
      g = x * ww - mm;
      dfw = gv;
      dfx = gx * w - 0.5 * gd / vs;
      dgw = 2.0 * x * w;
      dgx = ww;
    } else {
      const double up = 1.0 + x, gm = sqrt(up);
      f = ((up - gx) * w - up * en) + gm * gd;
      g = x * ww - up * mm;
      dfw = up - gx;
      dfx = (w - en) + 0.5 * gd / gm;
      dgw = 2.0 * x * w;
      dgx = ww - mm;
    }
    const double det = dfw * dgx - dfx * dgw;
    const double jfw = dgx / det, jgw = -dfx / det, jfx = -dgw / det, jgx = dfw / det;
    const double dw = f * jfw + g * jgw, dx = f * jfx + g * jgx;
    w = w - dw;
    x = x - dx;
    if (w < wl) return C2P_FAIL;
    if (WU ? x < 0.0 : (x < 0.0 || x >= 1.0)) return C2P_FAIL;
    err = tmax(fabs(dw / w), fabs(dx));
    if (err < s.tol) {
      keep = cn > 0;
      cn = cn - 1;
    } else {
      keep = it > 0;
    }
    it = it - 1;
  }
  vv = WU ? x / (1.0 + x) : x;
  return err >= s.tol ? C2P_NOCONV : 0;
}

// cons2prim_srhd_adi (3550-3639): the pressure is w - E as recovered (no
// floor); a failed recovery writes p = 0 and keeps the rest of q
template <class P>
__device__ __forceinline__ int cons2prim_srhd(const SrConst& s, const double* u, double* q) {
  const double mm = (u[P::imx] * u[P::imx] + u[P::imy] * u[P::imy]) + u[P::imz] * u[P::imz];
  const double en = u[P::ien] + u[P::idn];
  const double dn = u[P::idn];
  double w = 0.0, vv = 0.0;
  const int st = s.solver == AMUN_C2P_2DWV   ? nr_iterate_srhd_2d<false>(s, mm, en, dn, w, vv)
                 : s.solver == AMUN_C2P_2DWU ? nr_iterate_srhd_2d<true>(s, mm, en, dn, w, vv)
                                             : nr_iterate_srhd_1dw(s, mm, en, dn, w, vv);
  if (st & C2P_FAIL) {
    q[P::ipr] = 0.0;
    return st;
  }
  const double vs = sqrt(1.0 - vv);
  q[P::idn] = dn * vs;
  q[P::ivx] = u[P::imx] / w;
  q[P::ivy] = u[P::imy] / w;
  q[P::ivz] = u[P::imz] / w;
  q[P::ipr] = w - en;
  return st;
}

// cons2prim_srmhd_adi, equations.F90:4671-4790.  q must hold the previous
// primitive state on entry: a cell whose bracketing fails gets p = 0 and
// keeps the rest, as the reference leaves its q array.
template <class P>
__device__ __forceinline__ int cons2prim_sr(const SrConst& s, const double* u, double* q) {
  if constexpr (!P::mhd) return cons2prim_srhd<P>(s, u, q);
  else {
  const double mm = (u[P::imx] * u[P::imx] + u[P::imy] * u[P::imy]) + u[P::imz] * u[P::imz];
  const double mb = (u[P::imx] * u[P::ibx] + u[P::imy] * u[P::iby]) + u[P::imz] * u[P::ibz];
  const double bb = (u[P::ibx] * u[P::ibx] + u[P::iby] * u[P::iby]) + u[P::ibz] * u[P::ibz];
  const double en = u[P::ien] + u[P::idn];
  const double dn = u[P::idn];
  double w = 0.0, vv = 0.0;
  int st = s.solver == AMUN_C2P_2DWV   ? nr_iterate_2d<false>(s, mm, bb, mb, en, dn, w, vv)
           : s.solver == AMUN_C2P_2DWU ? nr_iterate_2d<true>(s, mm, bb, mb, en, dn, w, vv)
                                       : nr_iterate_1dw(s, mm, bb, mb, en, dn, w, vv);
  if (st & C2P_FAIL) {
    q[P::ipr] = 0.0;
    return st;
  }
  const double vm = 1.0 - vv;
  const double vs = sqrt(vm);
  const double wt = w + bb;
  const double fc = mb / w;
  const double rwt = rcp(wt);
  q[P::idn] = dn * vs;
  q[P::ivx] = (u[P::imx] + fc * u[P::ibx]) * rwt;
  q[P::ivy] = (u[P::imy] + fc * u[P::iby]) * rwt;
  q[P::ivz] = (u[P::imz] + fc * u[P::ibz]) * rwt;
  q[P::ibx] = u[P::ibx];
  q[P::iby] = u[P::iby];
  q[P::ibz] = u[P::ibz];
  q[P::ibp] = u[P::ibp];
  q[P::ipr] = (w - en) + 0.5 * (bb + (bb * mm - mb * mb) * (rwt * rwt));
  if (q[P::ipr] <= 0.0) {
    q[P::ipr] = 0.0;
    st |= C2P_PNEG;
  }
  return st;
  }
}

// riemann_srmhd_adi_hll (schemes.F90:5244-5355): the GLM pre-step, then HLL
// with the relativistic fluxes and speeds
template <class P, typename T>
__device__ __forceinline__ void riemann_sr_hll(const SrConst& s, T cmax, T cmax2, T* ql,
                                               T* qr, T* f) {
  constexpr int nv = P::nv;
  T ul[nv], ur[nv], fl[nv], fr[nv];
  T clm, clp, crm, crp;
  glm_upwind<P>(ql, qr, cmax);
  const T rgx = T(s.rgammaxi);
  prim2cons_sr<P>(rgx, ql, ul);
  prim2cons_sr<P>(rgx, qr, ur);
  fluxspeed_sr<P>(s, cmax2, ql, ul, fl, clm, clp);
  fluxspeed_sr<P>(s, cmax2, qr, ur, fr, crm, crp);
  const T sl = tmin(clm, crm), sr = tmax(clp, crp);
  if (sl >= T(0)) {
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = fl[p];
  } else if (sr <= T(0)) {
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = fr[p];
  } else {
    const T rsrml = rcp(sr - sl);
#pragma unroll
    for (int p = 0; p < nv; p++) {
      const T wl = sl * ul[p] - fl[p];
      const T wr = sr * ur[p] - fr[p];
      f[p] = (sl * wr - sr * wl) * rsrml;
    }
  }
}

// riemann_srhd_adi_hllc (schemes.F90:4775-4968, Mignone & Bodo 2005): the
// HLL average state and flux, the contact speed as the smaller root of
// F_E x^2 - (E + F_m) x + m = 0 of that state (in FP64), the star state on
// the side of the contact; no root, a root outside (sl, sr) or a
// non-positive star pressure keep the HLL flux
template <class P, typename T>
__device__ __forceinline__ void riemann_srhd_hllc(const SrConst& s, T* ql, T* qr, T* f) {
  constexpr int nv = P::nv;
  constexpr int idn = P::idn, imx = P::imx, imy = P::imy, imz = P::imz, ien = P::ien;
  T ul[nv], ur[nv], fl[nv], fr[nv], wl[nv], wr[nv], uh[nv], fh[nv];
  T clm, clp, crm, crp;
  const T rgx = T(s.rgammaxi);
  prim2cons_sr<P>(rgx, ql, ul);
  prim2cons_sr<P>(rgx, qr, ur);
  fluxspeed_sr<P>(s, T(0), ql, ul, fl, clm, clp);
  fluxspeed_sr<P>(s, T(0), qr, ur, fr, crm, crp);
  const T sl = tmin(clm, crm), sr = tmax(clp, crp);
  if (sl >= T(0)) {
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = fl[p];
    return;
  }
  if (sr <= T(0)) {
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = fr[p];
    return;
  }
  const T srml = sr - sl;
#pragma unroll
  for (int p = 0; p < nv; p++) {
    wl[p] = sl * ul[p] - fl[p];
    wr[p] = sr * ur[p] - fr[p];
    uh[p] = (wr[p] - wl[p]) / srml;
    fh[p] = (sl * wr[p] - sr * wl[p]) / srml;
    f[p] = fh[p];                           // the flux of every fallback
  }
  wl[ien] = wl[ien] + wl[idn];
  wr[ien] = wr[ien] + wr[idn];
  double a[3], x[2];
  a[0] = double(uh[imx]);
  a[1] = double(-((fh[imx] + uh[ien]) + uh[idn]));
  a[2] = double(fh[ien] + fh[idn]);
  if (alg_quadratic(a, x) < 1) return;
  const T sm = T(x[0]);
  if (sm <= sl || sm >= sr) return;
  const T pr = fh[imx] - (fh[ien] + fh[idn]) * sm;
  if (pr <= T(0)) return;
  if (sm == T(0)) {
    f[idn] = T(0);
    f[imx] = pr;
    f[imy] = T(0);
    f[imz] = T(0);
    f[ien] = T(0);
    return;
  }
  const bool left = sm > T(0);
  const T sx = left ? sl : sr;
  const T* w = left ? wl : wr;
  const T* uq = left ? ul : ur;
  const T* fq = left ? fl : fr;
  const T dv = sx - sm;
  T us[nv];
  us[idn] = w[idn] / dv;
  us[imy] = w[imy] / dv;
  us[imz] = w[imz] / dv;
  us[ien] = (w[ien] + pr * sm) / dv;
  us[imx] = (us[ien] + pr) * sm;
  us[ien] = us[ien] - us[idn];
#pragma unroll
  for (int p = 0; p < nv; p++) f[p] = fq[p] + sx * (us[p] - uq[p]);
}

// riemann_srmhd_adi_hllc (schemes.F90:5382-5700, Mignone & Bodo 2006): the
// HLL average state and flux, the contact speed as the smaller root of the
// quadratic of the average state (solved in FP64), the star state of the
// Bx != 0 or the Bx = 0 case; every degenerate case keeps the HLL flux
template <class P, typename T>
__device__ __forceinline__ void riemann_sr_hllc(const SrConst& s, T cmax, T cmax2, T* ql,
                                                T* qr, T* f) {
  if constexpr (!P::mhd) {
    riemann_srhd_hllc<P, T>(s, ql, qr, f);
    return;
  } else {
  constexpr int nv = P::nv;
  constexpr int idn = P::idn, imx = P::imx, imy = P::imy, imz = P::imz, ien = P::ien;
  constexpr int ibx = P::ibx, iby = P::iby, ibz = P::ibz, ibp = P::ibp, ivx = P::ivx;
  T ul[nv], ur[nv], fl[nv], fr[nv], wl[nv], wr[nv], uh[nv], fh[nv];
  T clm, clp, crm, crp;
  glm_upwind<P>(ql, qr, cmax);
  const T rgx = T(s.rgammaxi);
  prim2cons_sr<P>(rgx, ql, ul);
  prim2cons_sr<P>(rgx, qr, ur);
  fluxspeed_sr<P>(s, cmax2, ql, ul, fl, clm, clp);
  fluxspeed_sr<P>(s, cmax2, qr, ur, fr, crm, crp);
  const T sl = tmin(clm, crm), sr = tmax(clp, crp);
  if (sl >= T(0)) {
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = fl[p];
    return;
  }
  if (sr <= T(0)) {
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = fr[p];
    return;
  }
  const T srml = sr - sl;
#pragma unroll
  for (int p = 0; p < nv; p++) {
    wl[p] = sl * ul[p] - fl[p];
    wr[p] = sr * ur[p] - fr[p];
    uh[p] = (wr[p] - wl[p]) / srml;
    fh[p] = (sl * wr[p] - sr * wl[p]) / srml;
    f[p] = fh[p];                           // the flux of every fallback
  }
  wl[ien] = wl[ien] + wl[idn];
  wr[ien] = wr[ien] + wr[idn];
  const T bx2 = ql[ibx] * qr[ibx];
  double a[3], x[2];
  T us[nv];
  if (bx2 >= T(1.0e-16)) {
    const T uu = uh[iby] * uh[iby] + uh[ibz] * uh[ibz];
    const T ff = fh[iby] * fh[iby] + fh[ibz] * fh[ibz];
    const T uf = uh[iby] * fh[iby] + uh[ibz] * fh[ibz];
    a[0] = double(uh[imx] - uf);
    a[1] = double((uu + ff) - ((fh[imx] + uh[ien]) + uh[idn]));
    a[2] = double((fh[ien] + fh[idn]) - uf);
    if (alg_quadratic(a, x) < 1) return;
    const T sm = T(x[0]);
    if (sm <= sl || sm >= sr) return;
    us[ibx] = ql[ibx];
    us[iby] = uh[iby];
    us[ibz] = uh[ibz];
    us[ibp] = ql[ibp];
    const T vs1 = (us[iby] * sm - fh[iby]) / us[ibx];
    const T vs2 = (us[ibz] * sm - fh[ibz]) / us[ibx];
    const T vv = (sm * sm + vs1 * vs1) + vs2 * vs2;
    const T vb = (sm * us[ibx] + vs1 * us[iby]) + vs2 * us[ibz];
    const T gi = T(1) - vv;
    const T pt = (fh[imx] + gi * bx2) - ((fh[ien] + fh[idn]) - us[ibx] * vb) * sm;
    if (pt <= T(0) || sm == T(0)) return;
    const bool left = sm > T(0);
    const T sx = left ? sl : sr;
    const T* qq = left ? ql : qr;
    const T* uq = left ? ul : ur;
    const T* fq = left ? fl : fr;
    const T dv = sx - sm;
    us[idn] = (sx - qq[ivx]) / dv * uq[idn];
    us[imy] = ((sx * uq[imy] - fq[imy]) - us[ibx] * (gi * us[iby] + vs1 * vb)) / dv;
    us[imz] = ((sx * uq[imz] - fq[imz]) - us[ibx] * (gi * us[ibz] + vs2 * vb)) / dv;
    us[ien] = ((((sx * (uq[ien] + uq[idn])) - (fq[ien] + fq[idn])) + pt * sm) - us[ibx] * vb) / dv -
              us[idn];
    us[imx] = ((us[ien] + us[idn]) + pt) * sm - us[ibx] * vb;
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = fq[p] + sx * (us[p] - uq[p]);
  } else {
    a[0] = double(uh[imx]);
    a[1] = double(-((fh[imx] + uh[ien]) + uh[idn]));
    a[2] = double(fh[ien] + fh[idn]);
    if (alg_quadratic(a, x) < 1) return;
    const T sm = T(x[0]);
    if (sm <= sl || sm >= sr) return;
    const T pt = fh[imx] - (fh[ien] + fh[idn]) * sm;
    if (pt <= T(0) || sm == T(0)) return;
    const bool left = sm > T(0);
    const T sx = left ? sl : sr;
    const T* w = left ? wl : wr;
    const T* qq = left ? ql : qr;
    const T* uq = left ? ul : ur;
    const T* fq = left ? fl : fr;
    const T dv = sx - sm;
    us[idn] = w[idn] / dv;
    us[imy] = w[imy] / dv;
    us[imz] = w[imz] / dv;
    us[ien] = (w[ien] + pt * sm) / dv;
    us[imx] = (us[ien] + pt) * sm;
    us[ien] = us[ien] - us[idn];
    us[ibx] = T(0);
    us[iby] = w[iby] / dv;
    us[ibz] = w[ibz] / dv;
    us[ibp] = qq[ibp];
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = fq[p] + sx * (us[p] - uq[p]);
  }
  }
}

}  // namespace amun
