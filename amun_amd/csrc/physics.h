// physics.h -- device functions of the AMUN per-block update for gfx950.
//
// Pointwise equation-of-state routines, TVD limiters, the WENO5-Z cell
// reconstruction and the HLL / HLLC / HLLD Riemann solvers, written for
// register-resident state vectors (fully unrolled over the variable count).
// Each routine follows the reference routine cited next to it: the branch
// structure and the order of the sums are kept exactly.  Round-off differs
// from the CPU in two controlled ways: the compiler contracts a*b+c into
// FMAs, and divisions by a shared divisor use a refined hardware reciprocal
// (rcp below); WENO5-Z forms its normalised weights with fewer divisions.
//
// T is the arithmetic type of reconstruction and fluxes (double, or float in
// the mixed-precision mode).
#pragma once

#include <hip/hip_runtime.h>

namespace amun {

// variable layout of the Newtonian equation systems (equations.F90:335-485);
// the relativistic one is in srmhd.h
struct HD {
  static constexpr int nv = 5;
  static constexpr int idn = 0, ivx = 1, ivy = 2, ivz = 3;
  static constexpr int imx = 1, imy = 2, imz = 3;
  static constexpr int ipr = 4, ien = 4;
  static constexpr int ibx = -1, iby = -1, ibz = -1, ibp = -1;
  static constexpr bool mhd = false;
  static constexpr bool sr = false;
  static constexpr bool iso = false;
};

struct MHD {
  static constexpr int nv = 9;
  static constexpr int idn = 0, ivx = 1, ivy = 2, ivz = 3;
  static constexpr int imx = 1, imy = 2, imz = 3;
  static constexpr int ibx = 4, iby = 5, ibz = 6, ibp = 7;
  static constexpr int ipr = 8, ien = 8;
  static constexpr bool mhd = true;
  static constexpr bool sr = false;
  static constexpr bool iso = false;
};

// the isothermal systems (equations.F90:399-830): no energy equation, the
// pressure csnd^2 rho
struct HDI {
  static constexpr int nv = 4;
  static constexpr int idn = 0, ivx = 1, ivy = 2, ivz = 3;
  static constexpr int imx = 1, imy = 2, imz = 3;
  static constexpr int ipr = -1, ien = -1;
  static constexpr int ibx = -1, iby = -1, ibz = -1, ibp = -1;
  static constexpr bool mhd = false;
  static constexpr bool sr = false;
  static constexpr bool iso = true;
};

struct MHDI {
  static constexpr int nv = 8;
  static constexpr int idn = 0, ivx = 1, ivy = 2, ivz = 3;
  static constexpr int imx = 1, imy = 2, imz = 3;
  static constexpr int ibx = 4, iby = 5, ibz = 6, ibp = 7;
  static constexpr int ipr = -1, ien = -1;
  static constexpr bool mhd = true;
  static constexpr bool sr = false;
  static constexpr bool iso = true;
};

// constants that every kernel needs, passed by value
template <typename T>
struct EosConst {
  T gamma, gammam1, gammam1i;
  T csnd, csnd2;                  // isothermal sound speed and its square
};

template <typename T> __device__ __forceinline__ T tsqrt(T x);
template <> __device__ __forceinline__ double tsqrt<double>(double x) { return sqrt(x); }
template <> __device__ __forceinline__ float tsqrt<float>(float x) { return sqrtf(x); }

template <typename T> __device__ __forceinline__ T tabs(T x) { return x < T(0) ? -x : x; }
// MIN and MAX as the reference's compiler evaluates them (a < b ? a : b,
// a > b ? a : b): a NaN first operand yields the second, a NaN second
// operand propagates -- the wave speeds of a negative reconstructed pressure
// take the same path as in the reference
template <typename T> __device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }
template <typename T> __device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }
template <typename T> __device__ __forceinline__ T tsign(T x, T a) {
  return copysign(x, a);
}

// Reciprocal of the hot loops.  FP64: the hardware estimate v_rcp_f64
// refined by two Newton-Raphson steps (e = 1 - d r, r += r e), which is
// within an ulp of 1/d for the normal, finite operands met here; FP32: the
// v_rcp_f32 estimate (1 ulp).  a * rcp(d) replaces a / d where the divisor is
// shared, which is the only intended rounding difference against the
// reference's divisions.
template <typename T> __device__ __forceinline__ T rcp(T d);
template <> __device__ __forceinline__ double rcp<double>(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}
template <> __device__ __forceinline__ float rcp<float>(float d) {
  return __builtin_amdgcn_rcpf(d);
}
template <typename T> __device__ __forceinline__ T dv(T a, T b) { return a * rcp(b); }

// ---------------------------------------------------------------------------
// TVD limiters, interpolations.F90:4684-4877 (x is the scaling factor 1/2)
// ---------------------------------------------------------------------------
enum { LIM_ZERO = 0, LIM_MINMOD = 1, LIM_MC = 2, LIM_SUPERBEE = 3,
       LIM_VANLEER = 4, LIM_VANALBADA = 5 };

template <typename T>
__device__ __forceinline__ T limiter(int kind, T x, T a, T b) {
  switch (kind) {
  case LIM_MINMOD:
    return T(0.5) * (tsign(x, a) + tsign(x, b)) * tmin(tabs(a), tabs(b));
  case LIM_MC:
    return (tsign(x, a) + tsign(x, b)) *
           tmin(tmin(tabs(a), tabs(b)), T(0.25) * tabs(a + b));
  case LIM_SUPERBEE:
    return T(0.5) * (tsign(x, a) + tsign(x, b)) *
           tmax(tmin(T(2) * tabs(a), tabs(b)), tmin(tabs(a), T(2) * tabs(b)));
  case LIM_VANLEER: {
    T c = a * b;
    return c > T(0) ? T(2) * x * c / (a + b) : T(0);
  }
  case LIM_VANALBADA:
    return x * a * b * (a + b) / tmax(T(2.220446049250313e-16), a * a + b * b);
  default:
    return T(0);
  }
}

// ---------------------------------------------------------------------------
// WENO5-Z of one cell (reconstruct_weno5z, interpolations.F90:1986-2065):
// from f[0..4] = f(i-2..i+2) the state at the right face (fl(i), i+1/2) and
// the state at the left face (fr(i-1), i-1/2).  The smoothness indicators are
// shared by both faces exactly as in the reference loop body.
// ---------------------------------------------------------------------------
// YC = true: reconstruct_weno5yc (interpolations.F90:2107-2254), whose tau
// is the squared fourth difference (6 f(i) - 4 (f(i-1) + f(i+1)) + f(i-2) +
// f(i+2))^2; everything else is WENO5-Z's
template <typename T, bool YC = false>
__device__ __forceinline__ void weno5z_cell(const T f[5], T eps, T& fl, T& fr) {
  const T c1 = T(13.0 / 12.0), c2 = T(0.25);
  const T dl = T(0.1), dc = T(0.6), dr = T(0.3);
  const T a11 = T(2.0 / 6.0), a12 = T(-7.0 / 6.0), a13 = T(11.0 / 6.0);
  const T a21 = T(-1.0 / 6.0), a22 = T(5.0 / 6.0), a23 = T(2.0 / 6.0);
  const T a31 = T(2.0 / 6.0), a32 = T(5.0 / 6.0), a33 = T(-1.0 / 6.0);
  // dfm(i) = f(i)-f(i-1), dfp(i) = f(i+1)-f(i); df2(i) = c1*(dfp-dfm)^2
  const T dfm_m1 = f[1] - f[0];
  const T dfm_0 = f[2] - f[1];
  const T dfp_0 = f[3] - f[2];
  const T dfp_p1 = f[4] - f[3];
  const T dfp_m1 = dfm_0, dfm_p1 = dfp_0;
  T t;
  t = dfp_m1 - dfm_m1; const T df2_m1 = c1 * (t * t);
  t = dfp_0 - dfm_0;   const T df2_0 = c1 * (t * t);
  t = dfp_p1 - dfm_p1; const T df2_p1 = c1 * (t * t);
  t = T(3) * dfm_0 - dfm_m1; const T bl = df2_m1 + c2 * (t * t);
  t = dfp_0 + dfm_0;         const T bc = df2_0 + c2 * (t * t);
  t = T(3) * dfp_0 - dfp_p1; const T br = df2_p1 + c2 * (t * t);
  T tt;
  if constexpr (YC) {
    const T d4 = (T(6) * f[2] - T(4) * (f[1] + f[3])) + (f[0] + f[4]);
    tt = d4 * d4;
  } else {
    tt = tabs(br - bl);
  }
  // the three interpolations of each face
  const T ql = a11 * f[0] + a12 * f[1] + a13 * f[2];
  const T qc = a21 * f[1] + a22 * f[2] + a23 * f[3];
  const T qr = a31 * f[2] + a32 * f[3] + a33 * f[4];
  const T pl = a11 * f[4] + a12 * f[3] + a13 * f[2];
  const T pc = a21 * f[3] + a22 * f[2] + a23 * f[1];
  const T pr = a31 * f[2] + a32 * f[1] + a33 * f[0];
  // alpha_k = 1 + tau/(beta_k + eps) (eq. 28 of Borges et al. 2008, p = 1);
  // the normalised weights w_k = d_k alpha_k / sum(d alpha) are formed
  // without the reference's three alpha divisions: in FP64 all alphas are
  // scaled by (e_l e_c e_r), e_k = beta_k + eps, which leaves one division
  // per face; in FP32 the alphas use the hardware reciprocal
  T gl, gc, gr;
  const T el = bl + eps, ec = bc + eps, er = br + eps;
  if constexpr (sizeof(T) == 8) {
    gl = (el + tt) * (ec * er);
    gc = (ec + tt) * (el * er);
    gr = (er + tt) * (el * ec);
  } else {
    gl = T(1) + tt * rcp(el);
    gc = T(1) + tt * rcp(ec);
    gr = T(1) + tt * rcp(er);
  }
  // left state of the upper face (weights dl, dc, dr on gl, gc, gr) and
  // right state of the lower face (mirrored stencils, gl and gr swapped)
  const T wl = dl * gl, wc = dc * gc, wr = dr * gr;
  const T vl = dl * gr, vr = dr * gl;
  fl = dv((wl * ql + wr * qr) + wc * qc, (wl + wr) + wc);
  fr = dv((vl * pl + vr * pr) + wc * pc, (vl + vr) + wc);
}

// ---------------------------------------------------------------------------
// Gaussian-process reconstruction of one cell (reconstruct_gp,
// interpolations.F90:4561-4668) with the monotonicity-preserving limiter of
// Suresh & Huynh (mp_limiting, 5209-5284): from f[0..4] = f(i-2..i+2) and the
// weights cgp of prepare_gp, the state at the upper face (fl) and, from the
// mirrored stencil, at the lower face (fr)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T mmod(T a, T b) {   // minmod, interpolations.F90:4892
  return (tsign(T(0.5), a) + tsign(T(0.5), b)) * tmin(tabs(a), tabs(b));
}

// mp_limiting (interpolations.F90:5209-5284) of the unlimited face value u
// of the cell whose neighbourhood is f[0..4] = f(i-2..i+2)
template <typename T>
__device__ __forceinline__ T mp_limit(const T f[5], T u, T kappa, T kbeta, T eps) {
  const T fi = f[2];
  // dm(k) = f(k) - f(k-1) around the cell
  const T dmm = f[1] - f[0], dm0 = f[2] - f[1], dmp = f[3] - f[2], dmpp = f[4] - f[3];
  const T df = (dm0 * dmp >= T(0)) ? kappa * dm0 : kbeta * dm0;
  const T fmp = fi + mmod(dmp, df);
  const T ds = (u - fi) * (u - fmp);
  if (ds > eps) {
    const T dm1 = dm0 - dmm, dc0 = dmp - dm0, dp1 = dmpp - dmp, dc4 = T(4) * dc0;
    const T dml = T(0.5) * mmod(mmod(dc4 - dm1, T(4) * dm1 - dc0), mmod(dc0, dm1));
    const T dmr = T(0.5) * mmod(mmod(dc4 - dp1, T(4) * dp1 - dc0), mmod(dc0, dp1));
    const T fmd = (fi + T(0.5) * dmp) - dmr;
    const T ful = fi + df;
    const T flc = (fi + T(0.5) * df) + dml;
    const T fmx = tmax(tmin(tmin(fi, f[3]), fmd), tmin(tmin(fi, ful), flc));
    const T fmn = tmin(tmax(tmax(fi, f[3]), fmd), tmax(tmax(fi, ful), flc));
    u = u + mmod(fmn - u, fmx - u);               // median(u, fmn, fmx)
  }
  return u;
}

template <typename T>
__device__ __forceinline__ T gp_face(const T f[5], const T cgp[5], T kappa, T kbeta, T eps) {
  const T fi = f[2];
  T u = cgp[0] * (f[0] - fi);
#pragma unroll
  for (int m = 1; m < 5; m++) u = u + cgp[m] * (f[m] - fi);
  return mp_limit<T>(f, u + fi, kappa, kbeta, eps);
}

// reconstruct_mp5 / reconstruct_mp7 (interpolations.F90:3675-3770,
// 3799-3901): the fifth- or seventh-order polynomial face value, summed in
// stencil order, limited by mp_limiting; f holds 2 M + 1 cells centred on
// the cell
template <typename T, int M>
__device__ __forceinline__ T mp_face(const T* f, T kappa, T kbeta, T eps) {
  T u;
  if constexpr (M == 2) {
    const T ce[5] = {T(2.0 / 6.0e+01), T(-1.3e+01 / 6.0e+01), T(4.7e+01 / 6.0e+01),
                     T(2.7e+01 / 6.0e+01), T(-3.0 / 6.0e+01)};
    u = ce[0] * f[0];
#pragma unroll
    for (int m = 1; m < 5; m++) u = u + ce[m] * f[m];
  } else {
    const T ce[7] = {T(-3.0 / 4.2e+02), T(2.5e+01 / 4.2e+02), T(-1.01e+02 / 4.2e+02),
                     T(3.19e+02 / 4.2e+02), T(2.14e+02 / 4.2e+02), T(-3.8e+01 / 4.2e+02),
                     T(4.0 / 4.2e+02)};
    u = ce[0] * f[0];
#pragma unroll
    for (int m = 1; m < 7; m++) u = u + ce[m] * f[m];
  }
  return mp_limit<T>(f + M - 2, u, kappa, kbeta, eps);
}

template <typename T, int M>
__device__ __forceinline__ void mp_cell(const T* f, T kappa, T kbeta, T eps, T& fl, T& fr) {
  fl = mp_face<T, M>(f, kappa, kbeta, eps);
  T g[2 * M + 1];
#pragma unroll
  for (int m = 0; m <= 2 * M; m++) g[m] = f[2 * M - m];
  fr = mp_face<T, M>(g, kappa, kbeta, eps);
}

// reconstruct_weno5ns (interpolations.F90:2278-2441): WENO5 with the
// smoothness measures and weights of the "NS" variant (absolute first
// differences, xi = 0.4, zeta from the cubed slope)
template <typename T>
__device__ __forceinline__ void weno5ns_cell(const T f[5], T eps, T& fl, T& fr) {
  const T dl = T(0.1), dc = T(0.6), dr = T(0.3), xi = T(4.0e-01);
  const T a11 = T(2.0 / 6.0), a12 = T(-7.0 / 6.0), a13 = T(11.0 / 6.0);
  const T a21 = T(-1.0 / 6.0), a22 = T(5.0 / 6.0), a23 = T(2.0 / 6.0);
  const T a31 = T(2.0 / 6.0), a32 = T(5.0 / 6.0), a33 = T(-1.0 / 6.0);
  const T dfm_m1 = f[1] - f[0], dfm_0 = f[2] - f[1], dfp_0 = f[3] - f[2], dfp_p1 = f[4] - f[3];
  const T dfp_m1 = dfm_0, dfm_p1 = dfp_0;
  const T d2m = T(0.5) * tabs(dfp_m1 - dfm_m1);
  const T d20 = T(0.5) * tabs(dfp_0 - dfm_0);
  const T d2p = T(0.5) * tabs(dfp_p1 - dfm_p1);
  auto weights = [&](T df, T b2l, T b2r, T s, T& wl, T& wc, T& wr) {
    const T lq = xi * df;
    const T bl = b2l + xi * s, bc = d20 + lq, br = b2r + lq;
    const T l3 = (df * df) * df;
    const T t = bl - br, r = l3 / (T(1) + l3);
    const T zt = T(0.5) * (t * t + r * r);
    const T el = bl + eps, ec = bc + eps, er = br + eps;
    const T al = dl * (T(1) + zt / (el * el));
    const T ac = dc * (T(1) + zt / (ec * ec));
    const T ar = dr * (T(1) + zt / (er * er));
    const T aa = (al + ar) + ac;
    wl = al / aa;
    wr = ar / aa;
    wc = T(1) - (wl + wr);
  };
  T wl, wc, wr;
  weights(tabs(dfp_0), d2m, d2p, tabs(T(2) * dfm_0 - dfm_m1), wl, wc, wr);
  const T ql = a11 * f[0] + a12 * f[1] + a13 * f[2];
  const T qc = a21 * f[1] + a22 * f[2] + a23 * f[3];
  const T qr = a31 * f[2] + a32 * f[3] + a33 * f[4];
  fl = (wl * ql + wr * qr) + wc * qc;
  weights(tabs(dfm_0), d2p, d2m, tabs(T(2) * dfp_0 - dfp_p1), wl, wc, wr);
  const T pl = a11 * f[4] + a12 * f[3] + a13 * f[2];
  const T pc = a21 * f[3] + a22 * f[2] + a23 * f[1];
  const T pr = a31 * f[2] + a32 * f[1] + a33 * f[0];
  fr = (wl * pl + wr * pr) + wc * pc;
}

// reconstruct_weno3 (interpolations.F90:1469-1574): three-cell window
// f[0..2] around the cell f[1]; h2 is the squared cell size of the sweep
// direction (it enters the smoothness weights)
template <typename T>
__device__ __forceinline__ void weno3_cell(const T f[3], T h2, T& fl, T& fr) {
  const T dp = T(2.0 / 3.0), dm = T(1.0 / 3.0);
  const T dfl = f[1] - f[0], dfr = f[2] - f[1];
  T ww = (dfr - dfl) * (dfr - dfl);
  const T ap = T(1) + ww / (dfr * dfr + h2);
  const T am = T(1) + ww / (dfl * dfl + h2);
  T wp = dp * ap, wm = dm * am;
  ww = T(2) * (wp + wm);
  fl = (wp * (f[1] + f[2]) + wm * (-f[0] + T(3) * f[1])) / ww;
  wp = dp * am;
  wm = dm * ap;
  ww = T(2) * (wp + wm);
  fr = (wp * (f[1] + f[0]) + wm * (-f[2] + T(3) * f[1])) / ww;
}

// reconstruct_limo3 (interpolations.F90:1598-1719): LimO3 (Cada & Torrilhon
// 2009) blended with the plain third-order slope by the indicator over a
// radius of limo3_rad cells; rdx2 = (limo3_rad h)^2
template <typename T>
__device__ __forceinline__ T limo3_slope(T th, T xl, T xi) {
  const T f1 = (T(2) + th) / T(6);
  const T f2 = th >= T(0) ? tmax(T(0), tmin(tmin(f1, th), T(0.8)))
                          : tmax(T(0), tmin(f1, T(-0.25) * th));
  return xl * f1 + xi * f2;
}

template <typename T>
__device__ __forceinline__ void limo3_cell(const T f[3], T eps, T rdx2, T& fl, T& fr) {
  const T dfl = f[1] - f[0], dfr = f[2] - f[1];
  const T et = (dfl * dfl + dfr * dfr) / rdx2;
  const T xi = tmax(T(0), T(0.5) * tmin(T(2), T(1) + (et - T(1)) / eps));
  const T xl = T(1) - xi;
  fl = tabs(dfr) > eps ? f[1] + dfr * limo3_slope(dfl / dfr, xl, xi) : f[1];
  fr = tabs(dfl) > eps ? f[1] - dfl * limo3_slope(dfr / dfl, xl, xi) : f[1];
}

// reconstruct_ppm (interpolations.F90:1746-1913): PPM face values with the
// Colella-Sekora limiters, on the window f[0..6] of the cell f[3] (cells
// -3..3; the cell states use -2..3).  fi(j) is the limited value at the upper
// face of window cell j, d2(j) the second difference at window cell j.
template <typename T>
__device__ __forceinline__ T ppm_d2(const T* f, int j) {
  return (f[j + 1] + f[j - 1]) - T(2) * f[j];
}

template <typename T>
__device__ __forceinline__ T ppm_face(const T* f, int j, T pc) {
  T fi = (T(7) * (f[j] + f[j + 1]) - (f[j - 1] + f[j + 2])) / T(12);
  if ((f[j + 1] - fi) * (fi - f[j]) < T(0)) {
    const T c = T(3) * ((f[j + 1] + f[j]) - T(2) * fi);
    const T m = ppm_d2(f, j), p = ppm_d2(f, j + 1);
    T l = T(0);
    if (tmin(tmin(c, m), p) * tmax(tmax(c, m), p) > T(0))
      l = tsign(tmin(pc * tmin(tabs(m), tabs(p)), tabs(c)), c);
    fi = T(0.5) * (f[j] + f[j + 1]) - l / T(6);
  }
  return fi;
}

template <typename T>
__device__ __forceinline__ void ppm_cell(const T f[7], T pc, T& fl, T& fr) {
  const T fim = ppm_face(f, 2, pc), fic = ppm_face(f, 3, pc), fip = ppm_face(f, 4, pc);
  const T fc = f[3];
  T alm = fim - fc, alp = fic - fc;
  const bool cm = tabs(alm) >= T(2) * tabs(alp);
  const bool cp = tabs(alp) >= T(2) * tabs(alm);
  bool ext = false;
  if (alm * alp > T(0)) {
    ext = true;
  } else if (cm || cp) {
    // the reference differences the lower face with itself (im2 = im1)
    const T dfm = fim - fim, dfp = fip - fic;
    const T dcm = fc - f[2], dcp = f[4] - fc;
    if (tmin(tabs(dfm), tabs(dfp)) >= tmin(tabs(dcm), tabs(dcp))) ext = dfm * dfp < T(0);
    else ext = dcm * dcp < T(0);
  }
  if (ext) {
    const T s = T(6) * (alm + alp);
    const T m = ppm_d2(f, 2), c = ppm_d2(f, 3), p = ppm_d2(f, 4);
    if (tmin(tmin(tmin(s, c), m), p) * tmax(tmax(tmax(s, c), m), p) > T(0)) {
      const T l = tsign(tmin(pc * tmin(tmin(tabs(m), tabs(c)), tabs(p)), tabs(s)), s);
      if (tabs(l) > T(0)) {
        alm = alm * l / s;
        alp = alp * l / s;
      } else {
        alm = T(0);
        alp = T(0);
      }
    } else {
      alm = T(0);
      alp = T(0);
    }
  } else {
    if (cp) {
      const T sgn = tsign(T(1), alp);
      const T amx = -(T(0.25) * (alp * alp) / (alp + alm));
      const T dfp = f[4] - fc;
      if (sgn * amx >= sgn * dfp) {
        const T del = dfp * (dfp - alm);
        alp = del >= T(0) ? T(-2) * (dfp + sgn * tsqrt(del)) : T(-2) * alm;
      }
    }
    if (cm) {
      const T sgn = tsign(T(1), alm);
      const T amx = -(T(0.25) * (alm * alm) / (alp + alm));
      const T dfm = f[2] - fc;
      if (sgn * amx >= sgn * dfm) {
        const T del = dfm * (dfm - alp);
        alm = del >= T(0) ? T(-2) * (dfm + sgn * tsqrt(del)) : T(-2) * alp;
      }
    }
  }
  fl = fc + alp;
  fr = fc + alm;
}

template <typename T>
__device__ __forceinline__ void gp_cell(const T f[5], const T cgp[5], T kappa, T kbeta, T eps,
                                        T& fl, T& fr) {
  fl = gp_face<T>(f, cgp, kappa, kbeta, eps);
  const T g[5] = {f[4], f[3], f[2], f[1], f[0]};
  fr = gp_face<T>(g, cgp, kappa, kbeta, eps);
}

// ---------------------------------------------------------------------------
// equation of state: prim2cons / cons2prim (equations.F90:1762-1876 HD,
// 2815-2939 MHD); fluxes and fast speeds (1896-1958, 2959-3045)
// ---------------------------------------------------------------------------
template <class P, typename T>
__device__ __forceinline__ void prim2cons(const EosConst<T>& e, const T* q, T* u) {
  u[P::idn] = q[P::idn];
  u[P::imx] = q[P::idn] * q[P::ivx];
  u[P::imy] = q[P::idn] * q[P::ivy];
  u[P::imz] = q[P::idn] * q[P::ivz];
  if constexpr (P::iso) {
    // prim2cons_hd_iso / prim2cons_mhd_iso (equations.F90:1360, 2220)
    if constexpr (P::mhd) {
      u[P::ibx] = q[P::ibx];
      u[P::iby] = q[P::iby];
      u[P::ibz] = q[P::ibz];
      u[P::ibp] = q[P::ibp];
    }
  } else if constexpr (!P::mhd) {
    T ek = T(0.5) * (u[P::imx] * q[P::ivx] + u[P::imy] * q[P::ivy] + u[P::imz] * q[P::ivz]);
    T ei = e.gammam1i * q[P::ipr];
    u[P::ien] = ei + ek;
  } else {
    u[P::ibx] = q[P::ibx];
    u[P::iby] = q[P::iby];
    u[P::ibz] = q[P::ibz];
    u[P::ibp] = q[P::ibp];
    T ei = e.gammam1i * q[P::ipr];
    T ek = T(0.5) * (u[P::imx] * q[P::ivx] + u[P::imy] * q[P::ivy] + u[P::imz] * q[P::ivz]);
    T em = T(0.5) * ((q[P::ibx] * q[P::ibx] + q[P::iby] * q[P::iby]) + q[P::ibz] * q[P::ibz]);
    u[P::ien] = (ei + ek) + em;
  }
}

template <class P, typename T>
__device__ __forceinline__ void cons2prim(const EosConst<T>& e, const T* u, T* q) {
  const T rdn = rcp(u[P::idn]);
  q[P::idn] = u[P::idn];
  q[P::ivx] = u[P::imx] * rdn;
  q[P::ivy] = u[P::imy] * rdn;
  q[P::ivz] = u[P::imz] * rdn;
  if constexpr (P::iso) {
    // cons2prim_hd_iso / cons2prim_mhd_iso (equations.F90:1421, 2285)
    if constexpr (P::mhd) {
      q[P::ibx] = u[P::ibx];
      q[P::iby] = u[P::iby];
      q[P::ibz] = u[P::ibz];
      q[P::ibp] = u[P::ibp];
    }
  } else if constexpr (!P::mhd) {
    T ek = T(0.5) * (u[P::imx] * q[P::ivx] + u[P::imy] * q[P::ivy] + u[P::imz] * q[P::ivz]);
    T ei = u[P::ien] - ek;
    q[P::ipr] = e.gammam1 * ei;
  } else {
    q[P::ibx] = u[P::ibx];
    q[P::iby] = u[P::iby];
    q[P::ibz] = u[P::ibz];
    q[P::ibp] = u[P::ibp];
    T ek = T(0.5) * (u[P::imx] * q[P::ivx] + u[P::imy] * q[P::ivy] + u[P::imz] * q[P::ivz]);
    T em = T(0.5) * ((q[P::ibx] * q[P::ibx] + q[P::iby] * q[P::iby]) + q[P::ibz] * q[P::ibz]);
    T ei = u[P::ien] - (ek + em);
    q[P::ipr] = e.gammam1 * ei;
  }
}

// maxspeed_hd_adi (equations.F90:2022-2035) / maxspeed_mhd_adi (3108-3122);
// maxspeed_srmhd_adi (5116-5163) is the speed of light
template <class P, typename T>
__device__ __forceinline__ T maxspeed(const EosConst<T>& e, const T* q) {
  if constexpr (P::sr && P::mhd) return T(1);
  T vv = (q[P::ivx] * q[P::ivx] + q[P::ivy] * q[P::ivy]) + q[P::ivz] * q[P::ivz];
  if constexpr (P::sr) {
    // maxspeed_srhd_adi (equations.F90:3746-3820): (|v| + c) / (1 + s)
    const T gx = e.gammam1 / e.gamma;
    const T ww = q[P::idn] + q[P::ipr] / gx;
    const T c2 = e.gamma * q[P::ipr] / ww;
    const T ss = c2 * (T(1) - vv) / (T(1) - c2);
    const T fc = T(1) + ss;
    const T cc = tsqrt(ss * (fc - vv));
    return (tsqrt(vv) + cc) / fc;
  }
  T v = tsqrt(vv);
  T c;
  if constexpr (P::iso && !P::mhd) {
    c = e.csnd;                     // maxspeed_hd_iso (equations.F90:1546-1570)
  } else if constexpr (P::iso) {
    // maxspeed_mhd_iso (equations.F90:2438-2510)
    T bb = (q[P::ibx] * q[P::ibx] + q[P::iby] * q[P::iby]) + q[P::ibz] * q[P::ibz];
    c = tsqrt(e.csnd2 + bb / q[P::idn]);
  } else if constexpr (!P::mhd) {
    c = tsqrt(e.gamma * q[P::ipr] / q[P::idn]);
  } else {
    T bb = (q[P::ibx] * q[P::ibx] + q[P::iby] * q[P::iby]) + q[P::ibz] * q[P::ibz];
    c = tsqrt((e.gamma * q[P::ipr] + bb) / q[P::idn]);
  }
  return v + c;
}

template <class P, typename T>
__device__ __forceinline__ void fluxspeed(const EosConst<T>& e, T cmax2, const T* q,
                                          const T* u, T* f, T& cm, T& cp) {
  if constexpr (P::iso && !P::mhd) {
    // fluxspeed_hd_iso (equations.F90:1484-1542)
    f[P::idn] = u[P::imx];
    f[P::imx] = q[P::ivx] * u[P::imx];
    f[P::imy] = q[P::ivx] * u[P::imy];
    f[P::imz] = q[P::ivx] * u[P::imz];
    f[P::imx] = f[P::imx] + e.csnd2 * q[P::idn];
    cm = q[P::ivx] - e.csnd;
    cp = q[P::ivx] + e.csnd;
  } else if constexpr (P::iso) {
    // fluxspeed_mhd_iso (equations.F90:2352-2435)
    T bx2 = q[P::ibx] * q[P::ibx];
    T by2 = q[P::iby] * q[P::iby];
    T bz2 = q[P::ibz] * q[P::ibz];
    T bb = (bx2 + by2) + bz2;
    T pt = e.csnd2 * q[P::idn] + T(0.5) * bb;
    f[P::idn] = u[P::imx];
    f[P::imx] = q[P::ivx] * u[P::imx] - bx2;
    f[P::imy] = q[P::ivx] * u[P::imy] - q[P::ibx] * q[P::iby];
    f[P::imz] = q[P::ivx] * u[P::imz] - q[P::ibx] * q[P::ibz];
    f[P::imx] = f[P::imx] + pt;
    f[P::ibx] = q[P::ibp];
    f[P::iby] = q[P::ivx] * q[P::iby] - q[P::ibx] * q[P::ivy];
    f[P::ibz] = q[P::ivx] * q[P::ibz] - q[P::ibx] * q[P::ivz];
    f[P::ibp] = cmax2 * q[P::ibx];
    T fa = e.csnd2 * q[P::idn];
    T fb = fa + bb;
    T fc = fb * fb - T(4) * fa * bx2;
    T cf;
    const T rdn = rcp(q[P::idn]);
    if (fc > T(0)) cf = tsqrt(T(0.5) * (fb + tsqrt(fc)) * rdn);
    else           cf = tsqrt(T(0.5) * fb * rdn);
    cm = q[P::ivx] - cf;
    cp = q[P::ivx] + cf;
  } else if constexpr (!P::mhd) {
    f[P::idn] = u[P::imx];
    f[P::imx] = q[P::ivx] * u[P::imx];
    f[P::imy] = q[P::ivx] * u[P::imy];
    f[P::imz] = q[P::ivx] * u[P::imz];
    f[P::imx] = f[P::imx] + q[P::ipr];
    f[P::ien] = q[P::ivx] * (u[P::ien] + q[P::ipr]);
    T cs = tsqrt(dv(e.gamma * q[P::ipr], q[P::idn]));
    cm = q[P::ivx] - cs;
    cp = q[P::ivx] + cs;
  } else {
    T bx2 = q[P::ibx] * q[P::ibx];
    T by2 = q[P::iby] * q[P::iby];
    T bz2 = q[P::ibz] * q[P::ibz];
    T bb = (bx2 + by2) + bz2;
    T vb = (q[P::ivx] * q[P::ibx] + q[P::ivy] * q[P::iby]) + q[P::ivz] * q[P::ibz];
    T pt = q[P::ipr] + T(0.5) * bb;
    f[P::idn] = u[P::imx];
    f[P::imx] = q[P::ivx] * u[P::imx] - bx2;
    f[P::imy] = q[P::ivx] * u[P::imy] - q[P::ibx] * q[P::iby];
    f[P::imz] = q[P::ivx] * u[P::imz] - q[P::ibx] * q[P::ibz];
    f[P::imx] = f[P::imx] + pt;
    f[P::ibx] = q[P::ibp];
    f[P::iby] = q[P::ivx] * q[P::iby] - q[P::ibx] * q[P::ivy];
    f[P::ibz] = q[P::ivx] * q[P::ibz] - q[P::ibx] * q[P::ivz];
    f[P::ibp] = cmax2 * q[P::ibx];
    f[P::ien] = q[P::ivx] * (u[P::ien] + pt) - q[P::ibx] * vb;
    T fa = e.gamma * q[P::ipr];
    T fb = fa + bb;
    T fc = fb * fb - T(4) * fa * bx2;
    T cf;
    const T rdn = rcp(q[P::idn]);
    if (fc > T(0)) cf = tsqrt(T(0.5) * (fb + tsqrt(fc)) * rdn);
    else           cf = tsqrt(T(0.5) * fb * rdn);
    cm = q[P::ivx] - cf;
    cp = q[P::ivx] + cf;
  }
}

// GLM pre-step of every MHD solver (schemes.F90:3640-3652): Bx and psi of
// both states replaced by their upwinded interface values
template <class P, typename T>
__device__ __forceinline__ void glm_upwind(T* ql, T* qr, T cmax) {
  if constexpr (P::mhd) {
    T bx = T(0.5) * ((qr[P::ibx] + ql[P::ibx]) - dv(qr[P::ibp] - ql[P::ibp], cmax));
    T bp = T(0.5) * ((qr[P::ibp] + ql[P::ibp]) - (qr[P::ibx] - ql[P::ibx]) * cmax);
    ql[P::ibx] = bx; qr[P::ibx] = bx;
    ql[P::ibp] = bp; qr[P::ibp] = bp;
  }
}

// ---------------------------------------------------------------------------
// Riemann solvers of one interface.  ql/qr are overwritten (intent(inout) in
// the reference).
// ---------------------------------------------------------------------------
// riemann_hd_adi_hll (schemes.F90:1248-1341) / riemann_mhd_adi_hll (3163-3280)
template <class P, typename T>
__device__ __forceinline__ void riemann_hll(const EosConst<T>& e, T cmax, T cmax2,
                                            T* ql, T* qr, T* f) {
  constexpr int nv = P::nv;
  T ul[nv], ur[nv], fl[nv], fr[nv];
  T clm, clp, crm, crp;
  glm_upwind<P>(ql, qr, cmax);
  prim2cons<P>(e, ql, ul);
  prim2cons<P>(e, qr, ur);
  fluxspeed<P>(e, cmax2, ql, ul, fl, clm, clp);
  fluxspeed<P>(e, cmax2, qr, ur, fr, crm, crp);
  T sl = tmin(clm, crm), sr = tmax(clp, crp);
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
      T wl = sl * ul[p] - fl[p];
      T wr = sr * ur[p] - fr[p];
      f[p] = (sl * wr - sr * wl) * rsrml;
    }
  }
}

// riemann_hd_adi_hllc, schemes.F90:1366-1513
template <class P, typename T>
__device__ __forceinline__ void riemann_hllc(const EosConst<T>& e, T cmax, T cmax2,
                                             T* ql, T* qr, T* f) {
  static_assert(!P::mhd, "HLLC is the hydrodynamic solver");
  constexpr int nv = P::nv;
  T ul[nv], ur[nv], fl[nv], fr[nv], wl[nv], wr[nv], ui[nv];
  T clm, clp, crm, crp;
  prim2cons<P>(e, ql, ul);
  prim2cons<P>(e, qr, ur);
  fluxspeed<P>(e, cmax2, ql, ul, fl, clm, clp);
  fluxspeed<P>(e, cmax2, qr, ur, fr, crm, crp);
  T sl = tmin(clm, crm), sr = tmax(clp, crp);
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
#pragma unroll
  for (int p = 0; p < nv; p++) {
    wl[p] = sl * ul[p] - fl[p];
    wr[p] = sr * ur[p] - fr[p];
  }
  const T rdn = rcp(wr[P::idn] - wl[P::idn]);
  T sm = (wr[P::imx] - wl[P::imx]) * rdn;
  T pr = (wl[P::idn] * wr[P::imx] - wr[P::idn] * wl[P::imx]) * rdn;
  if (sm > T(0)) {
    const T rslmm = rcp(sl - sm);
    ui[P::idn] = wl[P::idn] * rslmm;
    ui[P::imx] = ui[P::idn] * sm;
    ui[P::imy] = ui[P::idn] * ql[P::ivy];
    ui[P::imz] = ui[P::idn] * ql[P::ivz];
    ui[P::ien] = (wl[P::ien] + sm * pr) * rslmm;
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
  } else if (sm < T(0)) {
    const T rsrmm = rcp(sr - sm);
    ui[P::idn] = wr[P::idn] * rsrmm;
    ui[P::imx] = ui[P::idn] * sm;
    ui[P::imy] = ui[P::idn] * qr[P::ivy];
    ui[P::imz] = ui[P::idn] * qr[P::ivz];
    ui[P::ien] = (wr[P::ien] + sm * pr) * rsrmm;
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
  } else {
    f[P::idn] = T(0);
    f[P::imx] = -wl[P::imx];
    f[P::imy] = T(0);
    f[P::imz] = T(0);
    f[P::ien] = T(0);
  }
}

// riemann_hd_adi_roe (schemes.F90:1541-1661) with the eigenvectors of
// esystem_roe_hd_adi (equations.F90:2075-2196) at the Roe average; the wave
// contributions are added from the left wave for a non-negative Roe velocity
// and from the right one otherwise, as the reference does
template <class P, typename T>
__device__ __forceinline__ void riemann_roe_hd(const EosConst<T>& e, T cmax, T cmax2,
                                               T* ql, T* qr, T* f) {
  static_assert(!P::mhd, "the Roe solver on the device is the hydrodynamic one");
  constexpr int nv = P::nv;
  T ul[nv], ur[nv], fl[nv], fr[nv];
  T cm, cp;
  prim2cons<P>(e, ql, ul);
  prim2cons<P>(e, qr, ur);
  fluxspeed<P>(e, cmax2, ql, ul, fl, cm, cp);
  fluxspeed<P>(e, cmax2, qr, ur, fr, cm, cp);
  const T sdl = tsqrt(ql[P::idn]), sdr = tsqrt(qr[P::idn]), sds = sdl + sdr;
  const T rsds = rcp(sds);
  const T vx = (sdl * ql[P::ivx] + sdr * qr[P::ivx]) * rsds;
  const T vy = (sdl * ql[P::ivy] + sdr * qr[P::ivy]) * rsds;
  const T vz = (sdl * ql[P::ivz] + sdr * qr[P::ivz]) * rsds;
  const T hh = (dv(ul[P::ien] + ql[P::ipr], sdl) + dv(ur[P::ien] + qr[P::ipr], sdr)) * rsds;
  const T vv = (vx * vx + vy * vy) + vz * vz, vh = T(0.5) * vv;
  const T c2 = e.gammam1 * (hh - vh), na = T(0.5) * rcp(c2), cc = tsqrt(c2);
  const T vc = vx * cc, ng = na * e.gammam1, nd = T(2) * ng, nw = na * vc;
  const T nh = (na * e.gammam1) * vh, nc = na * cc;
  const T c[5] = {vx - cc, vx, vx, vx, vx + cc};
  if (c[0] >= T(0)) {
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = fl[p];
    return;
  }
  if (c[4] <= T(0)) {
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = fr[p];
    return;
  }
  T du[nv];
#pragma unroll
  for (int p = 0; p < nv; p++) du[p] = ur[p] - ul[p];
  // wave strengths: al_k = sum_p l(p, k) du(p)
  T al[5];
  al[0] = (nh + nw) * du[P::idn] + (-ng * vx - nc) * du[P::imx] + (-ng * vy) * du[P::imy] +
          (-ng * vz) * du[P::imz] + ng * du[P::ien];
  al[1] = -vy * du[P::idn] + du[P::imy];
  al[2] = -vz * du[P::idn] + du[P::imz];
  al[3] = (T(1) - ng * vv) * du[P::idn] + nd * vx * du[P::imx] + nd * vy * du[P::imy] +
          nd * vz * du[P::imz] - nd * du[P::ien];
  al[4] = (nh - nw) * du[P::idn] + (-ng * vx + nc) * du[P::imx] + (-ng * vy) * du[P::imy] +
          (-ng * vz) * du[P::imz] + ng * du[P::ien];
  // right eigenvectors r(k, :)
  const T r[5][5] = {{T(1), vx - cc, vy, vz, hh - vc},
                     {T(0), T(0), T(1), T(0), vy},
                     {T(0), T(0), T(0), T(1), vz},
                     {T(1), vx, vy, vz, vh},
                     {T(1), vx + cc, vy, vz, hh + vc}};
#pragma unroll
  for (int p = 0; p < nv; p++) f[p] = T(0.5) * (fl[p] + fr[p]);
  const bool fwd = vx >= T(0);
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const int k = fwd ? m : 4 - m;
    const T xx = -(T(0.5) * al[k] * tabs(c[k]));
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = f[p] + xx * r[k][p];
  }
}

// riemann_hd_iso_roe (schemes.F90:939-1057) with esystem_roe_hd_iso
// (equations.F90:1655-1738): the Roe-averaged velocity, the waves v -+ csnd
// and v, summed from the left wave for a non-negative velocity
template <class P, typename T>
__device__ __forceinline__ void riemann_roe_hd_iso(const EosConst<T>& e, T cmax, T cmax2,
                                                   T* ql, T* qr, T* f) {
  static_assert(P::iso && !P::mhd, "the isothermal hydrodynamic Roe solver");
  constexpr int nv = P::nv;
  T ul[nv], ur[nv], fl[nv], fr[nv];
  T cm, cp;
  prim2cons<P>(e, ql, ul);
  prim2cons<P>(e, qr, ur);
  fluxspeed<P>(e, cmax2, ql, ul, fl, cm, cp);
  fluxspeed<P>(e, cmax2, qr, ur, fr, cm, cp);
  const T sdl = tsqrt(ql[P::idn]), sdr = tsqrt(qr[P::idn]), sds = sdl + sdr;
  const T rsds = rcp(sds);
  const T vx = (sdl * ql[P::ivx] + sdr * qr[P::ivx]) * rsds;
  const T vy = (sdl * ql[P::ivy] + sdr * qr[P::ivy]) * rsds;
  const T vz = (sdl * ql[P::ivz] + sdr * qr[P::ivz]) * rsds;
  const T ch = T(0.5) * rcp(e.csnd);
  const T c[4] = {vx - e.csnd, vx, vx, vx + e.csnd};
  if (c[0] >= T(0)) {
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = fl[p];
    return;
  }
  if (c[3] <= T(0)) {
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = fr[p];
    return;
  }
  T du[nv];
#pragma unroll
  for (int p = 0; p < nv; p++) du[p] = ur[p] - ul[p];
  // wave strengths al_k = sum_p l(p, k) du(p)
  T al[4];
  al[0] = (ch * c[3]) * du[P::idn] + (-ch) * du[P::imx];
  al[1] = (-vy) * du[P::idn] + du[P::imy];
  al[2] = (-vz) * du[P::idn] + du[P::imz];
  al[3] = (-(ch * c[0])) * du[P::idn] + ch * du[P::imx];
  const T r[4][4] = {{T(1), c[0], vy, vz},
                     {T(0), T(0), T(1), T(0)},
                     {T(0), T(0), T(0), T(1)},
                     {T(1), c[3], vy, vz}};
#pragma unroll
  for (int p = 0; p < nv; p++) f[p] = T(0.5) * (fl[p] + fr[p]);
  const bool fwd = vx >= T(0);
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const int k = fwd ? m : 3 - m;
    const T xx = -(T(0.5) * al[k] * tabs(c[k]));
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = f[p] + xx * r[k][p];
  }
}

// riemann_mhd_iso_hlld (schemes.F90:2007-2364, M = false) and
// riemann_mhd_iso_hlldm (2388-2777, M = true): the GLM pre-step, the HLL
// bounds, the contact speed from the HLL state's momentum and the branch
// tree over Bx^2 (against 1e-8 of the outer waves' scale, > 0, = 0) and the
// degeneracies.  HLLD: star densities wl(idn) / (sl - sm) and wr(idn) /
// (sr - sm), Alfven speeds sm -+ sqrt(Bx^2 / rho*); HLLD-M: the HLL
// state's density on both sides
template <class P, bool M, typename T>
__device__ __forceinline__ void riemann_hlld_iso(const EosConst<T>& e, T cmax, T cmax2,
                                                 T* ql, T* qr, T* f) {
  static_assert(P::iso && P::mhd, "the isothermal MHD HLLD solvers");
  constexpr int nv = P::nv;
  T ul[nv], ur[nv], fl[nv], fr[nv], wl[nv], wr[nv], ui[nv];
  T clm, clp, crm, crp;
  glm_upwind<P>(ql, qr, cmax);
  prim2cons<P>(e, ql, ul);
  prim2cons<P>(e, qr, ur);
  fluxspeed<P>(e, cmax2, ql, ul, fl, clm, clp);
  fluxspeed<P>(e, cmax2, qr, ur, fr, crm, crp);
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
  }
  const T dn = wr[P::idn] - wl[P::idn];
  const T sm = dv(wr[P::imx] - wl[P::imx], dn);
  const T bx = ql[P::ibx];
  const T b2 = ql[P::ibx] * qr[P::ibx];
  const T slmm = sl - sm, srmm = sr - sm;
  // star densities and the coefficients of wl/wr(iby, ibz) in the star field
  T dnl, dnr, cl, cr;
  bool alfven;
  if constexpr (M) {
    dnl = dnr = dv(dn, srml);
    cl = slmm * dnl;
    cr = srmm * dnr;
    alfven = dv(b2, dnl) > T(1.0e-08) * tmax(sl * sl, sr * sr);
  } else {
    dnl = dv(wl[P::idn], slmm);
    dnr = dv(wr[P::idn], srmm);
    cl = wl[P::idn];
    cr = wr[P::idn];
    alfven = b2 > T(1.0e-08) * tmax(dnl * (sl * sl), dnr * (sr * sr));
  }
  auto hll = [&]() {
    const T r = rcp(srml);
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = (sl * wr[p] - sr * wl[p]) * r;
  };
  if (alfven) {
    T sml, smr, dvl, dvr;
    if constexpr (M) {
      const T ca = tsqrt(dv(b2, dnl));
      sml = sm - ca;
      smr = sm + ca;
      dvl = slmm * slmm * dnl - b2;
      dvr = srmm * srmm * dnr - b2;
    } else {
      sml = sm - tsqrt(dv(b2, dnl));
      smr = sm + tsqrt(dv(b2, dnr));
      dvl = slmm * wl[P::idn] - b2;
      dvr = srmm * wr[P::idn] - b2;
    }
    auto star_l = [&]() {
      const T r = rcp(dvl);
      ui[P::idn] = dnl;
      ui[P::imx] = dnl * sm;
      ui[P::imy] = dnl * (slmm * wl[P::imy] - bx * wl[P::iby]) * r;
      ui[P::imz] = dnl * (slmm * wl[P::imz] - bx * wl[P::ibz]) * r;
      ui[P::ibx] = bx;
      ui[P::iby] = (cl * wl[P::iby] - bx * wl[P::imy]) * r;
      ui[P::ibz] = (cl * wl[P::ibz] - bx * wl[P::imz]) * r;
      ui[P::ibp] = ul[P::ibp];
    };
    auto star_r = [&]() {
      const T r = rcp(dvr);
      ui[P::idn] = dnr;
      ui[P::imx] = dnr * sm;
      ui[P::imy] = dnr * (srmm * wr[P::imy] - bx * wr[P::iby]) * r;
      ui[P::imz] = dnr * (srmm * wr[P::imz] - bx * wr[P::ibz]) * r;
      ui[P::ibx] = bx;
      ui[P::iby] = (cr * wr[P::iby] - bx * wr[P::imy]) * r;
      ui[P::ibz] = (cr * wr[P::ibz] - bx * wr[P::imz]) * r;
      ui[P::ibp] = ur[P::ibp];
    };
    if (tmin(dvl, dvr) > T(0)) {
      if (sml >= T(0)) {
        star_l();
#pragma unroll
        for (int p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
      } else if (smr <= T(0)) {
        star_r();
#pragma unroll
        for (int p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
      } else {
        T wcl[nv];
        star_l();
#pragma unroll
        for (int p = 0; p < nv; p++) wcl[p] = (sml - sl) * ui[p] + wl[p];
        star_r();
        const T r = rcp(smr - sml);
#pragma unroll
        for (int p = 0; p < nv; p++) {
          const T wcr = (smr - sr) * ui[p] + wr[p];
          f[p] = (sml * wcr - smr * wcl[p]) * r;
        }
      }
    } else if (dvl > T(0)) {
      star_l();
      if (sml >= T(0)) {
#pragma unroll
        for (int p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
      } else {
        const T r = rcp(sr - sml);
#pragma unroll
        for (int p = 0; p < nv; p++) {
          const T wcl = (sml - sl) * ui[p] + wl[p];
          f[p] = (sml * wr[p] - sr * wcl) * r;
        }
      }
    } else if (dvr > T(0)) {
      star_r();
      if (smr <= T(0)) {
#pragma unroll
        for (int p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
      } else {
        const T r = rcp(smr - sl);
#pragma unroll
        for (int p = 0; p < nv; p++) {
          const T wcr = (smr - sr) * ui[p] + wr[p];
          f[p] = (sl * wcr - smr * wl[p]) * r;
        }
      }
    } else {
      hll();
    }
  } else if (b2 > T(0)) {
    hll();
  } else if (sm > T(0)) {
    const T r = rcp(slmm);
    ui[P::idn] = dnl;
    ui[P::imx] = dnl * sm;
    ui[P::imy] = wl[P::imy] * r;
    ui[P::imz] = wl[P::imz] * r;
    ui[P::ibx] = T(0);
    ui[P::iby] = wl[P::iby] * r;
    ui[P::ibz] = wl[P::ibz] * r;
    ui[P::ibp] = ul[P::ibp];
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
  } else if (sm < T(0)) {
    const T r = rcp(srmm);
    ui[P::idn] = dnr;
    ui[P::imx] = dnr * sm;
    ui[P::imy] = wr[P::imy] * r;
    ui[P::imz] = wr[P::imz] * r;
    ui[P::ibx] = T(0);
    ui[P::iby] = wr[P::iby] * r;
    ui[P::ibz] = wr[P::ibz] * r;
    ui[P::ibp] = ur[P::ibp];
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
  } else {
    hll();
  }
}

// riemann_mhd_iso_roe (schemes.F90:2778-2906) with the eigensystem of
// esystem_roe_mhd_iso (equations.F90:2552-2791) at the Roe average with the
// magnetic corrections x, y: six waves (the fourth and the eighth carry
// nothing), wave strengths from the left eigenvectors applied to the
// conserved jump, summed from the left wave for a non-negative Roe velocity
// and from the right one otherwise; the Bx and psi fluxes are the left ones
template <class P, typename T>
__device__ __forceinline__ void riemann_roe_mhd_iso(const EosConst<T>& e, T cmax, T cmax2,
                                                    T* ql, T* qr, T* f) {
  static_assert(P::iso && P::mhd, "the isothermal MHD Roe solver");
  constexpr int nv = P::nv;
  constexpr int idn = P::idn, ivx = P::ivx, ivy = P::ivy, ivz = P::ivz;
  constexpr int ibx = P::ibx, iby = P::iby, ibz = P::ibz;
  constexpr int imx = P::imx, imy = P::imy, imz = P::imz;
  T ul[nv], ur[nv], fl[nv], fr[nv];
  T cm, cp;
  glm_upwind<P>(ql, qr, cmax);
  prim2cons<P>(e, ql, ul);
  prim2cons<P>(e, qr, ur);
  fluxspeed<P>(e, cmax2, ql, ul, fl, cm, cp);
  fluxspeed<P>(e, cmax2, qr, ur, fr, cm, cp);
  const T sdl = tsqrt(ql[idn]), sdr = tsqrt(qr[idn]), sds = sdl + sdr;
  const T dn = sdl * sdr;
  const T vx = (sdl * ql[ivx] + sdr * qr[ivx]) / sds;
  const T vy = (sdl * ql[ivy] + sdr * qr[ivy]) / sds;
  const T vz = (sdl * ql[ivz] + sdr * qr[ivz]) / sds;
  const T bx = ql[ibx];
  const T by = (sdr * ql[iby] + sdl * qr[iby]) / sds;
  const T bz = (sdr * ql[ibz] + sdl * qr[ibz]) / sds;
  const T dby = ql[iby] - qr[iby], dbz = ql[ibz] - qr[ibz];
  const T x = T(0.5) * (dby * dby + dbz * dbz) / (sds * sds);
  const T y = T(0.5) * (ql[idn] + qr[idn]) / dn;
  // esystem_roe_mhd_iso
  const T di = T(1) / dn;
  const T casq = bx * bx * di, ca = tsqrt(casq);
  const T btsq = by * by + bz * bz;
  const T bt_starsq = btsq * y;
  const T twid_csq = e.csnd2 + x;
  const T ct2 = bt_starsq * di;
  const T tsum = (casq + ct2) + twid_csq, tdif = (casq + ct2) - twid_csq;
  const T cfsq = T(0.5) * (tsum + tsqrt(tdif * tdif + T(4) * twid_csq * ct2));
  const T cf = tsqrt(cfsq);
  const T cssq = twid_csq * casq / cfsq, cs = tsqrt(cssq);
  const T c[8] = {vx - cf, vx - ca, vx - cs, vx, vx + cs, vx + ca, vx + cf, vx + cf};
  if (c[0] >= T(0)) {
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = fl[p];
    return;
  }
  if (c[6] <= T(0)) {
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = fr[p];
    return;
  }
  const T bt = tsqrt(btsq), bt_star = tsqrt(bt_starsq);
  T bet2 = T(1), bet3 = T(0);
  if (!(bt == T(0))) {
    bet2 = by / bt;
    bet3 = bz / bt;
  }
  const T bet2_star = bet2 / tsqrt(y), bet3_star = bet3 / tsqrt(y);
  const T bet_starsq = bet2_star * bet2_star + bet3_star * bet3_star;
  T alpha_f, alpha_s;
  if ((cfsq - cssq) == T(0)) { alpha_f = T(1); alpha_s = T(0); }
  else if ((twid_csq - cssq) <= T(0)) { alpha_f = T(0); alpha_s = T(1); }
  else if ((cfsq - twid_csq) <= T(0)) { alpha_f = T(1); alpha_s = T(0); }
  else {
    alpha_f = tsqrt((twid_csq - cssq) / (cfsq - cssq));
    alpha_s = tsqrt((cfsq - twid_csq) / (cfsq - cssq));
  }
  const T sqrtd = tsqrt(dn);
  const T s = tsign(T(1), bx);
  const T twid_c = tsqrt(twid_csq);
  T qf = cf * alpha_f * s, qs = cs * alpha_s * s;
  const T af_prime = twid_c * alpha_f / sqrtd, as_prime = twid_c * alpha_s / sqrtd;
  // right eigenvectors R[wave][m] over the flux components m: idn, imx, imy,
  // imz, iby, ibz
  T R[8][6];
#pragma unroll
  for (int k = 0; k < 8; k++)
#pragma unroll
    for (int m = 0; m < 6; m++) R[k][m] = T(0);
  R[0][0] = alpha_f; R[2][0] = alpha_s; R[4][0] = alpha_s; R[6][0] = alpha_f;
  R[0][1] = alpha_f * c[0]; R[2][1] = alpha_s * c[2];
  R[4][1] = alpha_s * c[4]; R[6][1] = alpha_f * c[6];
  R[0][2] = alpha_f * vy + qs * bet2_star; R[1][2] = -bet3;
  R[2][2] = alpha_s * vy - qf * bet2_star; R[4][2] = alpha_s * vy + qf * bet2_star;
  R[5][2] = bet3; R[6][2] = alpha_f * vy - qs * bet2_star;
  R[0][3] = alpha_f * vz + qs * bet3_star; R[1][3] = bet2;
  R[2][3] = alpha_s * vz - qf * bet3_star; R[4][3] = alpha_s * vz + qf * bet3_star;
  R[5][3] = -bet2; R[6][3] = alpha_f * vz - qs * bet3_star;
  R[0][4] = as_prime * bet2_star; R[1][4] = -(bet3 * s / sqrtd);
  R[2][4] = -(af_prime * bet2_star);
  R[4][4] = R[2][4]; R[5][4] = R[1][4]; R[6][4] = R[0][4];
  R[0][5] = as_prime * bet3_star; R[1][5] = bet2 * s / sqrtd;
  R[2][5] = -(af_prime * bet3_star);
  R[4][5] = R[2][5]; R[5][5] = R[1][5]; R[6][5] = R[0][5];
  // left eigenvectors applied to the conserved jump: al[wave]
  const T norm = T(0.5) / twid_csq;
  const T cff = norm * alpha_f * cf, css = norm * alpha_s * cs;
  qf = qf * norm;
  qs = qs * norm;
  const T af = norm * af_prime * dn, as = norm * as_prime * dn;
  const T afpb = norm * af_prime * bt_star, aspb = norm * as_prime * bt_star;
  const T q2_star = bet2_star / bet_starsq, q3_star = bet3_star / bet_starsq;
  const T vqstr = vy * q2_star + vz * q3_star;
  const T d0 = ur[idn] - ul[idn], d1 = ur[imx] - ul[imx], d2 = ur[imy] - ul[imy];
  const T d3 = ur[imz] - ul[imz], d5 = ur[iby] - ul[iby], d6 = ur[ibz] - ul[ibz];
  auto lsum = [&](T l0, T l1, T l2, T l3, T l5, T l6) {
    return ((((l0 * d0 + l1 * d1) + l2 * d2) + l3 * d3) + l5 * d5) + l6 * d6;
  };
  T al[8];
  al[0] = lsum((cff * c[6] - qs * vqstr) - aspb, -cff, qs * q2_star, qs * q3_star,
               as * q2_star, as * q3_star);
  const T l20 = T(0.5) * (vy * bet3 - vz * bet2), l22 = -(T(0.5) * bet3), l23 = T(0.5) * bet2;
  const T l25 = -(T(0.5) * sqrtd * bet3 * s), l26 = T(0.5) * sqrtd * bet2 * s;
  al[1] = lsum(l20, T(0), l22, l23, l25, l26);
  const T l32 = -(qf * q2_star), l33 = -(qf * q3_star);
  const T l35 = -(af * q2_star), l36 = -(af * q3_star);
  al[2] = lsum((css * c[4] + qf * vqstr) + afpb, -css, l32, l33, l35, l36);
  al[3] = T(0);
  al[4] = lsum((-(css * c[2]) - qf * vqstr) + afpb, css, -l32, -l33, l35, l36);
  al[5] = lsum(-l20, T(0), -l22, -l23, l25, l26);
  al[6] = lsum((-(cff * c[0]) + qs * vqstr) - aspb, cff, -(qs * q2_star), -(qs * q3_star),
               as * q2_star, as * q3_star);
  al[7] = T(0);
  constexpr int fv[6] = {idn, imx, imy, imz, iby, ibz};
#pragma unroll
  for (int m = 0; m < 6; m++) f[fv[m]] = T(0.5) * (fl[fv[m]] + fr[fv[m]]);
  f[ibx] = fl[ibx];
  f[P::ibp] = fl[P::ibp];
  const bool fwd = vx >= T(0);
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int w = fwd ? k : 7 - k;
    const T xx = -(T(0.5) * al[w] * tabs(c[w]));
#pragma unroll
    for (int m = 0; m < 6; m++) f[fv[m]] = f[fv[m]] + xx * R[w][m];
  }
}

// riemann_mhd_adi_roe (schemes.F90:4230-4410) with the eigensystem of
// esystem_roe_mhd_adi (equations.F90:3162-3454) at the Roe average with the
// magnetic corrections x, y.  The wave strengths are the left eigenvectors
// applied to the conserved jump, the wave sum runs from the left wave for a
// non-negative Roe velocity and from the right one otherwise; the Bx and psi
// fluxes are the left state's.  Waves 5 and 9 carry nothing.
template <class P, typename T>
__device__ __forceinline__ void riemann_roe_mhd(const EosConst<T>& e, T cmax, T cmax2, T* ql,
                                                T* qr, T* f) {
  static_assert(P::mhd, "the MHD Roe solver");
  constexpr int nv = P::nv;
  constexpr int idn = P::idn, ivx = P::ivx, ivy = P::ivy, ivz = P::ivz, ipr = P::ipr;
  constexpr int ibx = P::ibx, iby = P::iby, ibz = P::ibz, ien = P::ien;
  constexpr int imx = P::imx, imy = P::imy, imz = P::imz;
  T ul[nv], ur[nv], fl[nv], fr[nv];
  T cm, cp;
  glm_upwind<P>(ql, qr, cmax);
  prim2cons<P>(e, ql, ul);
  prim2cons<P>(e, qr, ur);
  fluxspeed<P>(e, cmax2, ql, ul, fl, cm, cp);
  fluxspeed<P>(e, cmax2, qr, ur, fr, cm, cp);
  const T sdl = tsqrt(ql[idn]), sdr = tsqrt(qr[idn]), sds = sdl + sdr;
  const T pml = T(0.5) * ((ql[ibx] * ql[ibx] + ql[iby] * ql[iby]) + ql[ibz] * ql[ibz]);
  const T pmr = T(0.5) * ((qr[ibx] * qr[ibx] + qr[iby] * qr[iby]) + qr[ibz] * qr[ibz]);
  const T dn = sdl * sdr;
  const T vx = (sdl * ql[ivx] + sdr * qr[ivx]) / sds;
  const T vy = (sdl * ql[ivy] + sdr * qr[ivy]) / sds;
  const T vz = (sdl * ql[ivz] + sdr * qr[ivz]) / sds;
  const T hh = (((ul[ien] + ql[ipr]) + pml) / sdl + ((ur[ien] + qr[ipr]) + pmr) / sdr) / sds;
  const T bx = ql[ibx];
  const T by = (sdr * ql[iby] + sdl * qr[iby]) / sds;
  const T bz = (sdr * ql[ibz] + sdl * qr[ibz]) / sds;
  const T dby = ql[iby] - qr[iby], dbz = ql[ibz] - qr[ibz];
  const T x = T(0.5) * (dby * dby + dbz * dbz) / (sds * sds);
  const T y = T(0.5) * (ql[idn] + qr[idn]) / dn;
  // esystem_roe_mhd_adi
  const T gm1 = e.gammam1, gm2 = e.gamma - T(2);
  const T di = T(1) / dn;
  const T casq = bx * bx * di, ca = tsqrt(casq);
  const T vsq = (vx * vx + vy * vy) + vz * vz;
  const T btsq = by * by + bz * bz;
  const T bt_starsq = (gm1 - gm2 * y) * btsq;
  const T hp = hh - (casq + btsq * di);
  const T twid_asq = tmax(T(2.220446049250313e-16), gm1 * (hp - T(0.5) * vsq) - gm2 * x);
  const T ct2 = bt_starsq * di;
  const T tsum = (casq + ct2) + twid_asq, tdif = (casq + ct2) - twid_asq;
  const T cfsq = T(0.5) * (tsum + tsqrt(tdif * tdif + T(4) * twid_asq * ct2));
  const T cf = tsqrt(cfsq);
  const T cssq = twid_asq * casq / cfsq, cs = tsqrt(cssq);
  const T c[9] = {vx - cf, vx - ca, vx - cs, vx, vx, vx + cs, vx + ca, vx + cf, vx + cf};
  if (c[0] >= T(0)) {
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = fl[p];
    return;
  }
  if (c[7] <= T(0)) {
#pragma unroll
    for (int p = 0; p < nv; p++) f[p] = fr[p];
    return;
  }
  const T bt = tsqrt(btsq), bt_star = tsqrt(bt_starsq);
  T bet2 = T(1), bet3 = T(0);
  if (!(bt == T(0))) {
    bet2 = by / bt;
    bet3 = bz / bt;
  }
  const T bet2_star = bet2 / tsqrt(gm1 - gm2 * y), bet3_star = bet3 / tsqrt(gm1 - gm2 * y);
  const T bet_starsq = bet2_star * bet2_star + bet3_star * bet3_star;
  const T vbet = vy * bet2_star + vz * bet3_star;
  T alpha_f, alpha_s;
  if ((cfsq - cssq) == T(0)) { alpha_f = T(1); alpha_s = T(0); }
  else if ((twid_asq - cssq) <= T(0)) { alpha_f = T(0); alpha_s = T(1); }
  else if ((cfsq - twid_asq) <= T(0)) { alpha_f = T(1); alpha_s = T(0); }
  else {
    alpha_f = tsqrt((twid_asq - cssq) / (cfsq - cssq));
    alpha_s = tsqrt((cfsq - twid_asq) / (cfsq - cssq));
  }
  const T sqrtd = tsqrt(dn), isqrtd = T(1) / sqrtd;
  const T s = tsign(T(1), bx);
  const T twid_a = tsqrt(twid_asq);
  T qf = cf * alpha_f * s, qs = cs * alpha_s * s;
  const T af_prime = twid_a * alpha_f * isqrtd, as_prime = twid_a * alpha_s * isqrtd;
  const T afpbb = af_prime * bt_star * bet_starsq, aspbb = as_prime * bt_star * bet_starsq;
  // right eigenvectors R[wave][m] over the flux components m: idn, imx, imy,
  // imz, ien, iby, ibz
  T R[9][7];
#pragma unroll
  for (int k = 0; k < 9; k++)
#pragma unroll
    for (int m = 0; m < 7; m++) R[k][m] = T(0);
  R[3][0] = T(1);
  R[0][0] = alpha_f; R[2][0] = alpha_s; R[5][0] = alpha_s; R[7][0] = alpha_f;
  R[0][1] = alpha_f * c[0]; R[2][1] = alpha_s * c[2]; R[3][1] = vx;
  R[5][1] = alpha_s * c[5]; R[7][1] = alpha_f * c[7];
  {
    const T qa = alpha_f * vy, qb = alpha_s * vy, qc = qs * bet2_star, qd = qf * bet2_star;
    R[0][2] = qa + qc; R[1][2] = -bet3; R[2][2] = qb - qd; R[3][2] = vy;
    R[5][2] = qb + qd; R[6][2] = bet3; R[7][2] = qa - qc;
  }
  {
    const T qa = alpha_f * vz, qb = alpha_s * vz, qc = qs * bet3_star, qd = qf * bet3_star;
    R[0][3] = qa + qc; R[1][3] = bet2; R[2][3] = qb - qd; R[3][3] = vz;
    R[5][3] = qb + qd; R[6][3] = -bet2; R[7][3] = qa - qc;
  }
  R[0][4] = (alpha_f * (hp - vx * cf) + qs * vbet) + aspbb;
  R[1][4] = -(vy * bet3 - vz * bet2);
  R[2][4] = (alpha_s * (hp - vx * cs) - qf * vbet) - afpbb;
  R[3][4] = T(0.5) * vsq + gm2 * x / gm1;
  R[5][4] = (alpha_s * (hp + vx * cs) + qf * vbet) - afpbb;
  R[6][4] = -R[1][4];
  R[7][4] = (alpha_f * (hp + vx * cf) - qs * vbet) + aspbb;
  R[0][5] = as_prime * bet2_star; R[1][5] = -(bet3 * s * isqrtd);
  R[2][5] = -(af_prime * bet2_star);
  R[5][5] = R[2][5]; R[6][5] = R[1][5]; R[7][5] = R[0][5];
  R[0][6] = as_prime * bet3_star; R[1][6] = bet2 * s * isqrtd;
  R[2][6] = -(af_prime * bet3_star);
  R[5][6] = R[2][6]; R[6][6] = R[1][6]; R[7][6] = R[0][6];
  // left eigenvectors applied to the conserved jump: al[wave]
  T norm = T(0.5) / twid_asq;
  const T cff = norm * alpha_f * cf, css = norm * alpha_s * cs;
  qf = qf * norm;
  qs = qs * norm;
  const T af = norm * af_prime * dn, as = norm * as_prime * dn;
  const T afpb = norm * af_prime * bt_star, aspb = norm * as_prime * bt_star;
  norm = norm * gm1;
  const T af_n = alpha_f * norm, as_n = alpha_s * norm;
  const T q2_star = bet2_star / bet_starsq, q3_star = bet3_star / bet_starsq;
  const T vqstr = vy * q2_star + vz * q3_star;
  norm = T(2) * norm;
  const T d0 = ur[idn] - ul[idn], d1 = ur[imx] - ul[imx], d2 = ur[imy] - ul[imy];
  const T d3 = ur[imz] - ul[imz], d5 = ur[iby] - ul[iby], d6 = ur[ibz] - ul[ibz];
  const T d8 = ur[ien] - ul[ien];
  auto lsum = [&](T l0, T l1, T l2, T l3, T l5, T l6, T l8) {
    return (((((l0 * d0 + l1 * d1) + l2 * d2) + l3 * d3) + l5 * d5) + l6 * d6) + l8 * d8;
  };
  T al[9];
  al[0] = lsum(((af_n * (vsq - hp) + cff * (cf + vx)) - qs * vqstr) - aspb, -(af_n * vx) - cff,
               -(af_n * vy) + qs * q2_star, -(af_n * vz) + qs * q3_star,
               as * q2_star - af_n * by, as * q3_star - af_n * bz, af_n);
  const T l20 = T(0.5) * (vy * bet3 - vz * bet2), l22 = -(T(0.5) * bet3), l23 = T(0.5) * bet2;
  const T l25 = -(T(0.5) * sqrtd * bet3 * s), l26 = T(0.5) * sqrtd * bet2 * s;
  al[1] = lsum(l20, T(0), l22, l23, l25, l26, T(0));
  const T l35 = -(af * q2_star) - as_n * by, l36 = -(af * q3_star) - as_n * bz;
  al[2] = lsum(((as_n * (vsq - hp) + css * (cs + vx)) + qf * vqstr) + afpb, -(as_n * vx) - css,
               -(as_n * vy) - qf * q2_star, -(as_n * vz) - qf * q3_star, l35, l36, as_n);
  al[3] = lsum(T(1) - norm * (T(0.5) * vsq - gm2 * x / gm1), norm * vx, norm * vy, norm * vz,
               norm * by, norm * bz, -norm);
  al[4] = T(0);
  al[5] = lsum(((as_n * (vsq - hp) + css * (cs - vx)) - qf * vqstr) + afpb, -(as_n * vx) + css,
               -(as_n * vy) + qf * q2_star, -(as_n * vz) + qf * q3_star, l35, l36, as_n);
  al[6] = lsum(-l20, T(0), -l22, -l23, l25, l26, T(0));
  const T l15 = as * q2_star - af_n * by, l16 = as * q3_star - af_n * bz;
  al[7] = lsum(((af_n * (vsq - hp) + cff * (cf - vx)) + qs * vqstr) - aspb, -(af_n * vx) + cff,
               -(af_n * vy) - qs * q2_star, -(af_n * vz) - qs * q3_star, l15, l16, af_n);
  al[8] = T(0);
  constexpr int fv[7] = {idn, imx, imy, imz, ien, iby, ibz};
#pragma unroll
  for (int m = 0; m < 7; m++) f[fv[m]] = T(0.5) * (fl[fv[m]] + fr[fv[m]]);
  f[ibx] = fl[ibx];
  f[P::ibp] = fl[P::ibp];
  const bool fwd = vx >= T(0);
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const int w = fwd ? k : 8 - k;
    const T xx = -(T(0.5) * al[w] * tabs(c[w]));
#pragma unroll
    for (int m = 0; m < 7; m++) f[fv[m]] = f[fv[m]] + xx * R[w][m];
  }
}

// riemann_mhd_adi_hllc, schemes.F90:3305-3571: one intermediate state on
// each side of the contact (GLM pre-step as HLL; its own Bx = 0 branch)
template <class P, typename T>
__device__ __forceinline__ void riemann_hllc_mhd(const EosConst<T>& e, T cmax, T cmax2,
                                                 T* ql, T* qr, T* f) {
  constexpr int nv = P::nv;
  T ul[nv], ur[nv], fl[nv], fr[nv], wl[nv], wr[nv], ui[nv];
  T clm, clp, crm, crp;
  glm_upwind<P>(ql, qr, cmax);
  prim2cons<P>(e, ql, ul);
  prim2cons<P>(e, qr, ur);
  fluxspeed<P>(e, cmax2, ql, ul, fl, clm, clp);
  fluxspeed<P>(e, cmax2, qr, ur, fr, crm, crp);
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
  const T srml = sr - sl, rsrml = rcp(srml);
#pragma unroll
  for (int p = 0; p < nv; p++) {
    wl[p] = sl * ul[p] - fl[p];
    wr[p] = sr * ur[p] - fr[p];
  }
  const T rdn = rcp(wr[P::idn] - wl[P::idn]);
  const T sm = (wr[P::imx] - wl[P::imx]) * rdn;
  const T bx = ql[P::ibx], b2 = ql[P::ibx] * qr[P::ibx];
  if (b2 > T(0)) {
    const T pt = T(0.5) * ((wl[P::idn] * sm - wl[P::imx]) + (wr[P::idn] * sm - wr[P::imx])) + b2;
    const T vy = (wr[P::imy] - wl[P::imy]) * rdn;
    const T vz = (wr[P::imz] - wl[P::imz]) * rdn;
    const T by = (wr[P::iby] - wl[P::iby]) * rsrml;
    const T bz = (wr[P::ibz] - wl[P::ibz]) * rsrml;
    const T vb = (sm * bx + vy * by) + vz * bz;
    if (sm > T(0)) {
      const T rslmm = rcp(sl - sm);
      ui[P::idn] = wl[P::idn] * rslmm;
      ui[P::imx] = ui[P::idn] * sm;
      ui[P::imy] = ui[P::idn] * vy;
      ui[P::imz] = ui[P::idn] * vz;
      ui[P::ibx] = bx; ui[P::iby] = by; ui[P::ibz] = bz; ui[P::ibp] = ql[P::ibp];
      ui[P::ien] = ((wl[P::ien] + sm * pt) - bx * vb) * rslmm;
#pragma unroll
      for (int p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
    } else if (sm < T(0)) {
      const T rsrmm = rcp(sr - sm);
      ui[P::idn] = wr[P::idn] * rsrmm;
      ui[P::imx] = ui[P::idn] * sm;
      ui[P::imy] = ui[P::idn] * vy;
      ui[P::imz] = ui[P::idn] * vz;
      ui[P::ibx] = bx; ui[P::iby] = by; ui[P::ibz] = bz; ui[P::ibp] = qr[P::ibp];
      ui[P::ien] = ((wr[P::ien] + sm * pt) - bx * vb) * rsrmm;
#pragma unroll
      for (int p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
    } else {
#pragma unroll
      for (int p = 0; p < nv; p++) f[p] = (sl * wr[p] - sr * wl[p]) * rsrml;
    }
  } else {
    const T pt = T(0.5) * ((wl[P::idn] * sm - wl[P::imx]) + (wr[P::idn] * sm - wr[P::imx]));
    if (sm > T(0)) {
      const T rslmm = rcp(sl - sm);
      ui[P::idn] = wl[P::idn] * rslmm;
      ui[P::imx] = ui[P::idn] * sm;
      ui[P::imy] = ui[P::idn] * ql[P::ivy];
      ui[P::imz] = ui[P::idn] * ql[P::ivz];
      ui[P::ibx] = T(0); ui[P::iby] = wl[P::iby] * rslmm; ui[P::ibz] = wl[P::ibz] * rslmm;
      ui[P::ibp] = ql[P::ibp];
      ui[P::ien] = (wl[P::ien] + sm * pt) * rslmm;
#pragma unroll
      for (int p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
    } else if (sm < T(0)) {
      const T rsrmm = rcp(sr - sm);
      ui[P::idn] = wr[P::idn] * rsrmm;
      ui[P::imx] = ui[P::idn] * sm;
      ui[P::imy] = ui[P::idn] * qr[P::ivy];
      ui[P::imz] = ui[P::idn] * qr[P::ivz];
      ui[P::ibx] = T(0); ui[P::iby] = wr[P::iby] * rsrmm; ui[P::ibz] = wr[P::ibz] * rsrmm;
      ui[P::ibp] = qr[P::ibp];
      ui[P::ien] = (wr[P::ien] + sm * pt) * rsrmm;
#pragma unroll
      for (int p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
    } else {
      f[P::idn] = T(0);
      f[P::imx] = T(-0.5) * (wl[P::imx] + wr[P::imx]);
      f[P::imy] = T(0);
      f[P::imz] = T(0);
      f[P::ibx] = fl[P::ibx];
      f[P::iby] = T(0);
      f[P::ibz] = T(0);
      f[P::ibp] = fl[P::ibp];
      f[P::ien] = T(0);
    }
  }
}

// star state of the HLLD fan from a wave vector w (outer state, speed diff
// smm = s - sm, divisor dv = smm*w_rho - Bx^2), schemes.F90:3740-3760
template <class P, typename T>
__device__ __forceinline__ void hlld_outer(const T* w, T smm, T rsmm, T dvd, T dn, T sm,
                                           T bx, T pt, T psi, T* ui) {
  const T rdv = rcp(dvd);
  T vy = (smm * w[P::imy] - bx * w[P::iby]) * rdv;
  T vz = (smm * w[P::imz] - bx * w[P::ibz]) * rdv;
  T by = (w[P::idn] * w[P::iby] - bx * w[P::imy]) * rdv;
  T bz = (w[P::idn] * w[P::ibz] - bx * w[P::imz]) * rdv;
  T vb = (sm * bx + vy * by) + vz * bz;
  ui[P::idn] = dn;
  ui[P::imx] = dn * sm;
  ui[P::imy] = dn * vy;
  ui[P::imz] = dn * vz;
  ui[P::ibx] = bx;
  ui[P::iby] = by;
  ui[P::ibz] = bz;
  ui[P::ibp] = psi;
  ui[P::ien] = ((w[P::ien] + sm * pt) - bx * vb) * rsmm;
}

template <class P, typename T>
__device__ __forceinline__ void hlld_inner_state(T dn, T sm, T vy, T vz, T bx, T by,
                                                 T bz, T psi, T* ui) {
  ui[P::idn] = dn;
  ui[P::imx] = dn * sm;
  ui[P::imy] = dn * vy;
  ui[P::imz] = dn * vz;
  ui[P::ibx] = bx;
  ui[P::iby] = by;
  ui[P::ibz] = bz;
  ui[P::ibp] = psi;
}

// riemann_mhd_adi_hlld, schemes.F90:3596-4203
template <class P, typename T>
__device__ __forceinline__ void riemann_hlld(const EosConst<T>& e, T cmax, T cmax2,
                                             T* ql, T* qr, T* f) {
  static_assert(P::mhd, "HLLD is the MHD solver");
  constexpr int nv = P::nv;
  constexpr int idn = P::idn, imx = P::imx, imy = P::imy, imz = P::imz;
  constexpr int ibx = P::ibx, iby = P::iby, ibz = P::ibz, ibp = P::ibp, ien = P::ien;
  T ul[nv], ur[nv], fl[nv], fr[nv], wl[nv], wr[nv], wcl[nv], wcr[nv], ui[nv];
  T clm, clp, crm, crp;
  glm_upwind<P>(ql, qr, cmax);
  prim2cons<P>(e, ql, ul);
  prim2cons<P>(e, qr, ur);
  fluxspeed<P>(e, cmax2, ql, ul, fl, clm, clp);
  fluxspeed<P>(e, cmax2, qr, ur, fr, crm, crp);
  T sl = tmin(clm, crm), sr = tmax(clp, crp);
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
  const T rsrml = rcp(sr - sl);
#pragma unroll
  for (int p = 0; p < nv; p++) {
    wl[p] = sl * ul[p] - fl[p];
    wr[p] = sr * ur[p] - fr[p];
  }
  const T rdn = rcp(wr[idn] - wl[idn]);
  T sm = (wr[imx] - wl[imx]) * rdn;
  T slmm = sl - sm;
  T srmm = sr - sm;
  const T rslmm = rcp(slmm), rsrmm = rcp(srmm);
  T dnl = wl[idn] * rslmm;
  T dnr = wr[idn] * rsrmm;
  T bx = ql[ibx];
  T b2 = ql[ibx] * qr[ibx];
  T pt = (wl[idn] * wr[imx] - wr[idn] * wl[imx]) * rdn + b2;
  T psil = ul[ibp], psir = ur[ibp];

  if (b2 > T(1.0e-08) * tmax(dnl * (sl * sl), dnr * (sr * sr))) {
    T cal = tsqrt(b2 * (slmm * rcp(wl[idn])));
    T car = tsqrt(b2 * (srmm * rcp(wr[idn])));
    T sml = sm - cal;
    T smr = sm + car;
    T dvl = slmm * wl[idn] - b2;
    T dvr = srmm * wr[idn] - b2;
    if (dvl > T(0) && dvr > T(0)) {
      if (sml >= T(0)) {
        hlld_outer<P>(wl, slmm, rslmm, dvl, dnl, sm, bx, pt, psil, ui);
#pragma unroll
        for (int p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
      } else if (smr <= T(0)) {
        hlld_outer<P>(wr, srmm, rsrmm, dvr, dnr, sm, bx, pt, psir, ui);
#pragma unroll
        for (int p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
      } else {
        hlld_outer<P>(wl, slmm, rslmm, dvl, dnl, sm, bx, pt, psil, ui);
#pragma unroll
        for (int p = 0; p < nv; p++) wcl[p] = (sml - sl) * ui[p] + wl[p];
        hlld_outer<P>(wr, srmm, rsrmm, dvr, dnr, sm, bx, pt, psir, ui);
#pragma unroll
        for (int p = 0; p < nv; p++) wcr[p] = (smr - sr) * ui[p] + wr[p];
        T rdv = rcp(car * dnr + cal * dnl);
        T vy = (wcr[imy] - wcl[imy]) * rdv;
        T vz = (wcr[imz] - wcl[imz]) * rdv;
        rdv = rcp(car + cal);
        T by = (wcr[iby] - wcl[iby]) * rdv;
        T bz = (wcr[ibz] - wcl[ibz]) * rdv;
        T vb = (sm * bx + vy * by) + vz * bz;
        if (sm > T(0)) {
          hlld_inner_state<P>(dnl, sm, vy, vz, bx, by, bz, psil, ui);
          ui[ien] = -dv((wcl[ien] + sm * pt) - bx * vb, cal);
#pragma unroll
          for (int p = 0; p < nv; p++) f[p] = sml * ui[p] - wcl[p];
        } else if (sm < T(0)) {
          hlld_inner_state<P>(dnr, sm, vy, vz, bx, by, bz, psir, ui);
          ui[ien] = dv((wcr[ien] + sm * pt) - bx * vb, car);
#pragma unroll
          for (int p = 0; p < nv; p++) f[p] = smr * ui[p] - wcr[p];
        } else {
          const T rd = rcp(smr - sml);
#pragma unroll
          for (int p = 0; p < nv; p++) f[p] = (sml * wcr[p] - smr * wcl[p]) * rd;
        }
      }
    } else if (dvl > T(0)) {
      hlld_outer<P>(wl, slmm, rslmm, dvl, dnl, sm, bx, pt, psil, ui);
      if (sml >= T(0)) {
#pragma unroll
        for (int p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
      } else {
#pragma unroll
        for (int p = 0; p < nv; p++) wcl[p] = (sml - sl) * ui[p] + wl[p];
        T rdv = rcp(srmm * dnr + cal * dnl);
        T vy = (wr[imy] - wcl[imy]) * rdv;
        T vz = (wr[imz] - wcl[imz]) * rdv;
        rdv = rcp(sr - sml);
        T by = (wr[iby] - wcl[iby]) * rdv;
        T bz = (wr[ibz] - wcl[ibz]) * rdv;
        T vb = (sm * bx + vy * by) + vz * bz;
        hlld_inner_state<P>(dnl, sm, vy, vz, bx, by, bz, psil, ui);
        ui[ien] = -dv((wcl[ien] + sm * pt) - bx * vb, cal);
        if (sm >= T(0)) {
#pragma unroll
          for (int p = 0; p < nv; p++) f[p] = sml * ui[p] - wcl[p];
        } else {
#pragma unroll
          for (int p = 0; p < nv; p++) wcr[p] = (sm - sml) * ui[p] + wcl[p];
#pragma unroll
          for (int p = 0; p < nv; p++) f[p] = (sm * wr[p] - sr * wcr[p]) * rsrmm;
        }
      }
    } else if (dvr > T(0)) {
      hlld_outer<P>(wr, srmm, rsrmm, dvr, dnr, sm, bx, pt, psir, ui);
      if (smr <= T(0)) {
#pragma unroll
        for (int p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
      } else {
#pragma unroll
        for (int p = 0; p < nv; p++) wcr[p] = (smr - sr) * ui[p] + wr[p];
        T rdv = rcp(slmm * dnl - car * dnr);
        T vy = (wl[imy] - wcr[imy]) * rdv;
        T vz = (wl[imz] - wcr[imz]) * rdv;
        rdv = rcp(sl - smr);
        T by = (wl[iby] - wcr[iby]) * rdv;
        T bz = (wl[ibz] - wcr[ibz]) * rdv;
        T vb = (sm * bx + vy * by) + vz * bz;
        hlld_inner_state<P>(dnr, sm, vy, vz, bx, by, bz, psir, ui);
        ui[ien] = dv((wcr[ien] + sm * pt) - bx * vb, car);
        if (sm <= T(0)) {
#pragma unroll
          for (int p = 0; p < nv; p++) f[p] = smr * ui[p] - wcr[p];
        } else {
#pragma unroll
          for (int p = 0; p < nv; p++) wcl[p] = (sm - smr) * ui[p] + wcr[p];
#pragma unroll
          for (int p = 0; p < nv; p++) f[p] = (sm * wl[p] - sl * wcl[p]) * rslmm;
        }
      }
    } else {
#pragma unroll
      for (int p = 0; p < nv; p++) f[p] = (sl * wr[p] - sr * wl[p]) * rsrml;
    }
  } else if (b2 > T(0)) {
    T vy = (wr[imy] - wl[imy]) * rdn;
    T vz = (wr[imz] - wl[imz]) * rdn;
    T by = (wr[iby] - wl[iby]) * rsrml;
    T bz = (wr[ibz] - wl[ibz]) * rsrml;
    T vb = (sm * bx + vy * by) + vz * bz;
    if (sm > T(0)) {
      hlld_inner_state<P>(dnl, sm, vy, vz, bx, by, bz, psil, ui);
      ui[ien] = ((wl[ien] + sm * pt) - bx * vb) * rslmm;
#pragma unroll
      for (int p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
    } else if (sm < T(0)) {
      hlld_inner_state<P>(dnr, sm, vy, vz, bx, by, bz, psir, ui);
      ui[ien] = ((wr[ien] + sm * pt) - bx * vb) * rsrmm;
#pragma unroll
      for (int p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
    } else {
#pragma unroll
      for (int p = 0; p < nv; p++) f[p] = (sl * wr[p] - sr * wl[p]) * rsrml;
    }
  } else {
    if (sm > T(0)) {
      ui[idn] = dnl; ui[imx] = dnl * sm;
      ui[imy] = wl[imy] * rslmm; ui[imz] = wl[imz] * rslmm;
      ui[ibx] = T(0); ui[iby] = wl[iby] * rslmm; ui[ibz] = wl[ibz] * rslmm;
      ui[ibp] = psil;
      ui[ien] = (wl[ien] + sm * pt) * rslmm;
#pragma unroll
      for (int p = 0; p < nv; p++) f[p] = sl * ui[p] - wl[p];
    } else if (sm < T(0)) {
      ui[idn] = dnr; ui[imx] = dnr * sm;
      ui[imy] = wr[imy] * rsrmm; ui[imz] = wr[imz] * rsrmm;
      ui[ibx] = T(0); ui[iby] = wr[iby] * rsrmm; ui[ibz] = wr[ibz] * rsrmm;
      ui[ibp] = psir;
      ui[ien] = (wr[ien] + sm * pt) * rsrmm;
#pragma unroll
      for (int p = 0; p < nv; p++) f[p] = sr * ui[p] - wr[p];
    } else {
#pragma unroll
      for (int p = 0; p < nv; p++) f[p] = (sl * wr[p] - sr * wl[p]) * rsrml;
    }
  }
}

}  // namespace amun
