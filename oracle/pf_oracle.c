/*
 * C/OpenMP restatement of the reference phase-field explicit step.
 * TEST INFRASTRUCTURE ONLY: the CPU baseline timed by bench.py and a second
 * implementation the tests compare with oracle/pf_oracle.py.  Never linked into,
 * or called by, the product library.
 *
 * Same algorithm and operation order as pf_oracle.py (which cites the reference
 * per step): allen_cahn.py:53-85 state_rhs (closed-form local gradient, 7-point
 * edge-replicated Laplacian ((p-c)+(n-c))/h^2, arccos anisotropy path :28-39,
 * Arrhenius mobility), :106-114 explicit Euler, :117-128 liquid reset; laser
 * temperatures of simple/example.py:21-37 (Rosenthal), a moving Gaussian spot and
 * solidification/example.py:21-34 (quench).
 *
 * Built twice from this file: REAL=double (pfo_step_f64) and REAL=float
 * (pfo_step_f32); temperature, zeta and mobility are FP64 in both.
 */
#include <math.h>
#include <stddef.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REAL
#define REAL double
#define SUFFIX f64
#endif

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)

typedef struct pfo_params {
  int nz, ny, nx, k;
  double h[3];
  double m_g, gamma, kappa_g, ad_hoc, L0, Qg, gas_const, T_liquidus, T_clip, zeta_sharpness;
  double anisotropy, liquid_shift;
  int force_liquid;
  int model; /* 0 buffer, 1 Rosenthal, 2 Gaussian, 3 quench */
  double T_ambient, pos0[3], dir[3], vel, Q, kappa_T, alpha, T_cap, dT_peak, sigma_xy, sigma_z;
  double T_ref, grad_T, cool_rate;
} pfo_params;

static double temperature(const pfo_params* p, const double* T, long c, int ix, int iy, int iz, double t) {
  if (p->model == 0) return T[c];
  const double cx = ((double)ix + 0.5) * p->h[0], cy = ((double)iy + 0.5) * p->h[1], cz = ((double)iz + 0.5) * p->h[2];
  if (p->model == 3) return p->T_ref + p->grad_T * cz - p->cool_rate * t;
  const double vt = p->vel * t;
  const double X = cx - p->pos0[0] - vt * p->dir[0];
  const double Y = cy - p->pos0[1] - vt * p->dir[1];
  const double Z = cz - p->pos0[2] - vt * p->dir[2];
  if (p->model == 1) {
    const double R = sqrt(X * X + Y * Y + Z * Z);
    const double proj = X * p->dir[0] + Y * p->dir[1] + Z * p->dir[2];
    double v = p->T_ambient + p->Q / (2 * M_PI * p->kappa_T) / R * exp(-p->vel / (2 * p->alpha) * (R + proj));
    if (p->T_cap > 0.0 && v > p->T_cap) v = p->T_cap;
    return v;
  }
  const double gxy = 1.0 / (2.0 * p->sigma_xy * p->sigma_xy), gz = 1.0 / (2.0 * p->sigma_z * p->sigma_z);
  return p->T_ambient + p->dT_peak * exp(-((X * X + Y * Y) * gxy + Z * Z * gz));
}

/* eta_out = force_eta(eta_in + (t_crt - t_prev) * state_rhs(eta_in, t_prev, T)); axes [3][K][3] */
int CAT(pfo_step, SUFFIX)(const REAL* in, REAL* out, const pfo_params* p, const double* axes, const double* T,
                          double t_prev, double t_crt, int nthreads) {
  const int nz = p->nz, ny = p->ny, nx = p->nx, K = p->k;
  if (K > 256) return -1;
  const long n = (long)nz * ny * nx;
  const double hstep = t_crt - t_prev;
  const double h2[3] = {p->h[0] * p->h[0], p->h[1] * p->h[1], p->h[2] * p->h[2]};
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
  for (long c = 0; c < n; ++c) {
    const int ix = (int)(c % nx), iy = (int)((c / nx) % ny), iz = (int)(c / ((long)nx * ny));
    const REAL* e = in + c * K;
    const long sx = K, sy = (long)nx * K, sz = (long)nx * ny * K;
    const REAL* xm = e - (ix > 0 ? sx : 0);
    const REAL* xp = e + (ix < nx - 1 ? sx : 0);
    const REAL* ym = e - (iy > 0 ? sy : 0);
    const REAL* yp = e + (iy < ny - 1 ? sy : 0);
    const REAL* zm = e - (iz > 0 ? sz : 0);
    const REAL* zp = e + (iz < nz - 1 ? sz : 0);
    const double Traw = temperature(p, T, c, ix, iy, iz, t_prev);
    const double Tc = Traw > p->T_clip ? p->T_clip : Traw;
    const double zeta = 0.5 * (1 - tanh(p->zeta_sharpness * (Tc / p->T_liquidus - 1)));
    const double Lg = p->L0 * exp(-p->Qg / (Tc * p->gas_const));
    const REAL omz2 = (REAL)((1.0 - zeta) * (1.0 - zeta));
    REAL S = 0;
    for (int i = 0; i < K; ++i) S += e[i] * e[i];
    const int liquid = p->force_liquid && (Traw > p->T_liquidus + p->liquid_shift);
    for (int i = 0; i < K; ++i) {
      const REAL v = e[i];
      const REAL grad = (REAL)p->m_g * (v * v * v - v + (REAL)(4.0 * p->gamma) * v * (S - v * v) + 2 * omz2 * v) /
                        (REAL)p->ad_hoc;
      REAL a = 1;
      if (p->anisotropy > 0.0) {
        const REAL ed[3] = {xp[i] - xm[i], yp[i] - ym[i], zp[i] - zm[i]};
        const REAL nrm = (REAL)sqrt((double)(ed[0] * ed[0] + ed[1] * ed[1] + ed[2] * ed[2]));
        double amin = 1e300;
        for (int kk = 0; kk < 3; ++kk) {
          const double* d = axes + ((long)kk * K + i) * 3;
          const REAL cs = ((REAL)d[0] * ed[0] + (REAL)d[1] * ed[1] + (REAL)d[2] * ed[2]) / nrm;
          double ang = acos((double)cs);
          if (!isfinite(ang)) ang = 0.0;
          if (!(ang < M_PI / 2.0)) ang = M_PI - ang;
          if (ang < amin) amin = ang;
        }
        const double cc = cos(amin), ss = sin(amin);
        a = (REAL)(1.0 + p->anisotropy * (cc * cc * cc * cc + ss * ss * ss * ss));
      }
      const REAL lap = -((((xp[i] - v) + (xm[i] - v)) / (REAL)h2[0] + ((yp[i] - v) + (ym[i] - v)) / (REAL)h2[1] +
                          ((zp[i] - v) + (zm[i] - v)) / (REAL)h2[2]) *
                         a * (REAL)(p->kappa_g * p->ad_hoc));
      const REAL rhs = -(REAL)Lg * (grad + lap);
      REAL y = v + (REAL)hstep * rhs;
      if (liquid) y = 0;
      out[c * K + i] = y;
    }
  }
  return 0;
}

#if defined(PFO_FIRST_UNIT)
int pfo_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
#endif
