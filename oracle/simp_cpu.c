/*
 * simp_cpu.c -- multi-core FP64 CPU restatement of the SIMP loop.
 * TEST / BASELINE INFRASTRUCTURE ONLY: used by tests/ (checked against the
 * NumPy oracle, simp_oracle.py) and by the cpu_baseline leg of bench.py. The
 * product never links it.
 *
 * Algorithm, step for step, from genpipeline/topology/simp_solver.py:
 *   run loop                     :56-79   (x0 = volfrac, xPhys0 = x0)
 *   K u (matrix-free, KE given)  :201-205 (K = sum_e E_e KE^T scattered)
 *   solve                        Jacobi-PCG FP64 instead of SuperLU (:213),
 *                                ||r|| <= tol ||f_free||, warm start
 *   sensitivity                  :215-218
 *   filter chain rule            :221-223
 *   OC bisection                 :225-249 (filter + mean every step, literal)
 *   final compliance             :78,177-193
 * Parallelism: OpenMP over nodes (gather form, no atomics) and elements.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const int CX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
static const int CY[8] = {0, 0, 1, 1, 0, 0, 1, 1};
static const int CZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};

typedef struct {
  int nx, ny, nz, sJ, sI, eJ, eI;
  long nn, ne, nd;
} grid_t;

static grid_t mk(int nx, int ny, int nz) {
  grid_t g;
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.sJ = nz + 1; g.sI = (ny + 1) * (nz + 1);
  g.eJ = nz; g.eI = ny * nz;
  g.nn = (long)(nx + 1) * g.sI; g.ne = (long)nx * ny * nz; g.nd = 3 * g.nn;
  return g;
}

static int corner_of(int a, int b, int c) { /* corner index of offset (a,b,c) */
  for (int m = 0; m < 8; ++m)
    if (CX[m] == a && CY[m] == b && CZ[m] == c) return m;
  return -1;
}

/* y = K(E) u with K = sum_e E_e P_e KE^T P_e^T (reference CPU assembly);
 * fixed DOFs of y zeroed when fixmask != NULL (uint8 per DOF). */
void ocpu_ku(int nx, int ny, int nz, const double* ke, const double* E, const double* u,
             double* y, const uint8_t* fixmask) {
  const grid_t g = mk(nx, ny, nz);
#pragma omp parallel for schedule(static)
  for (long n = 0; n < g.nn; ++n) {
    const int i = (int)(n / g.sI), rem = (int)(n % g.sI), j = rem / g.sJ, k = rem % g.sJ;
    double acc[3] = {0, 0, 0};
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int c = 0; c < 2; ++c) {
          const int ei = i - 1 + a, ej = j - 1 + b, ek = k - 1 + c;
          if (ei < 0 || ei >= nx || ej < 0 || ej >= ny || ek < 0 || ek >= nz) continue;
          const double Ee = E[(long)ei * g.eI + ej * g.eJ + ek];
          const int L = corner_of(1 - a, 1 - b, 1 - c);
          const long n0 = (long)ei * g.sI + (long)ej * g.sJ + ek;
          for (int r = 0; r < 3; ++r) {
            double s = 0;
            for (int m = 0; m < 8; ++m) {
              const long nm = n0 + CX[m] * g.sI + CY[m] * g.sJ + CZ[m];
              for (int q = 0; q < 3; ++q) s += ke[(3 * m + q) * 24 + 3 * L + r] * u[3 * nm + q];
            }
            acc[r] += Ee * s;
          }
        }
    for (int r = 0; r < 3; ++r) y[3 * n + r] = (fixmask && fixmask[3 * n + r]) ? 0.0 : acc[r];
  }
}

static void diag_inv(const grid_t* g, const double* ke, const double* E, const uint8_t* fm,
                     double* minv) {
#pragma omp parallel for schedule(static)
  for (long n = 0; n < g->nn; ++n) {
    const int i = (int)(n / g->sI), rem = (int)(n % g->sI), j = rem / g->sJ, k = rem % g->sJ;
    double d[3] = {0, 0, 0};
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int c = 0; c < 2; ++c) {
          const int ei = i - 1 + a, ej = j - 1 + b, ek = k - 1 + c;
          if (ei < 0 || ei >= g->nx || ej < 0 || ej >= g->ny || ek < 0 || ek >= g->nz) continue;
          const double Ee = E[(long)ei * g->eI + ej * g->eJ + ek];
          const int L = corner_of(1 - a, 1 - b, 1 - c);
          for (int r = 0; r < 3; ++r) d[r] += Ee * ke[(3 * L + r) * 24 + 3 * L + r];
        }
    for (int r = 0; r < 3; ++r)
      minv[3 * n + r] = (fm[3 * n + r] || !(d[r] > 0)) ? 0.0 : 1.0 / d[r];
  }
}

/* Jacobi-PCG, FP64. u: in guess (if use_x0) / out. Returns iterations; sets
 * *status 0 ok, 4 breakdown, 5 max_iter. */
int ocpu_pcg(int nx, int ny, int nz, const double* ke, const double* E, const double* f,
             const uint8_t* fm, double* u, int use_x0, double tol, int max_iter, int* status,
             double* relres) {
  const grid_t g = mk(nx, ny, nz);
  const long nd = g.nd;
  double* r = malloc(nd * sizeof(double));
  double* z = malloc(nd * sizeof(double));
  double* p = malloc(nd * sizeof(double));
  double* Ap = malloc(nd * sizeof(double));
  double* mi = malloc(nd * sizeof(double));
  diag_inv(&g, ke, E, fm, mi);
  if (use_x0) {
    for (long d = 0; d < nd; ++d)
      if (fm[d]) u[d] = 0.0;
    ocpu_ku(nx, ny, nz, ke, E, u, Ap, fm);
  } else {
    memset(u, 0, nd * sizeof(double));
    memset(Ap, 0, nd * sizeof(double));
  }
  double rz = 0, rr = 0, bb = 0;
#pragma omp parallel for reduction(+ : rz, rr, bb) schedule(static)
  for (long d = 0; d < nd; ++d) {
    const double b = fm[d] ? 0.0 : f[d];
    r[d] = fm[d] ? 0.0 : b - Ap[d];
    z[d] = mi[d] * r[d];
    p[d] = z[d];
    rz += r[d] * z[d];
    rr += r[d] * r[d];
    bb += b * b;
  }
  int it = 0;
  *status = 0;
  const double tol2 = tol * tol * bb;
  if (bb == 0.0) {
    memset(u, 0, nd * sizeof(double));
    rr = 0;
  } else {
    while (rr > tol2) {
      if (it >= max_iter) {
        *status = 5;
        break;
      }
      ocpu_ku(nx, ny, nz, ke, E, p, Ap, fm);
      double pAp = 0;
#pragma omp parallel for reduction(+ : pAp) schedule(static)
      for (long d = 0; d < nd; ++d) pAp += p[d] * Ap[d];
      if (!(pAp > 0)) {
        *status = 4;
        break;
      }
      const double a = rz / pAp;
      double rz2 = 0, rr2 = 0;
#pragma omp parallel for reduction(+ : rz2, rr2) schedule(static)
      for (long d = 0; d < nd; ++d) {
        u[d] += a * p[d];
        r[d] -= a * Ap[d];
        z[d] = mi[d] * r[d];
        rz2 += r[d] * z[d];
        rr2 += r[d] * r[d];
      }
      ++it;
      rr = rr2;
      if (rr <= tol2) break;
      const double beta = rz2 / rz;
      rz = rz2;
#pragma omp parallel for schedule(static)
      for (long d = 0; d < nd; ++d) p[d] = z[d] + beta * p[d];
    }
  }
  if (relres) *relres = bb > 0 ? sqrt(rr / bb) : 0.0;
  free(r); free(z); free(p); free(Ap); free(mi);
  return it;
}

/* -------------------------------------------------------------- filter */
typedef struct {
  int n, di[343], dj[343], dk[343];
  double w[343];
} stencil_t;

static void mk_stencil(double rmin, stencil_t* s) {
  const int r = (int)ceil(rmin);
  s->n = 0;
  for (int a = -r; a <= r; ++a)
    for (int b = -r; b <= r; ++b)
      for (int c = -r; c <= r; ++c) {
        const double d = sqrt((double)(a * a + b * b + c * c));
        if (d > rmin || !(rmin - d > 0)) continue;
        s->di[s->n] = a; s->dj[s->n] = b; s->dk[s->n] = c; s->w[s->n] = rmin - d; s->n++;
      }
}

/* mode 0: out = (H in)/Hs ; mode 1: out = H (in/Hs) */
static void filt(const grid_t* g, const stencil_t* s, const double* Hs, const double* in,
                 double* out, int mode) {
#pragma omp parallel for schedule(static)
  for (long e = 0; e < g->ne; ++e) {
    const int i = (int)(e / g->eI), rem = (int)(e % g->eI), j = rem / g->eJ, k = rem % g->eJ;
    double acc = 0;
    for (int o = 0; o < s->n; ++o) {
      const int ii = i + s->di[o], jj = j + s->dj[o], kk = k + s->dk[o];
      if (ii < 0 || ii >= g->nx || jj < 0 || jj >= g->ny || kk < 0 || kk >= g->nz) continue;
      const long f = (long)ii * g->eI + jj * g->eJ + kk;
      acc += s->w[o] * (mode ? in[f] / Hs[f] : in[f]);
    }
    out[e] = mode ? acc : acc / Hs[e];
  }
}

static double lawE(double x, int law, double p, double E0, double Emin) {
  return law == 0 ? pow(fmax(x, 1e-3), p) : Emin + pow(x, p) * (E0 - Emin);
}
static double lawdE(double x, int law, double p, double E0, double Emin) {
  return law == 0 ? p * pow(fmax(x, 1e-3), p - 1) : p * pow(x, p - 1) * (E0 - Emin);
}

/* Whole SIMP loop; x, xPhys in/out (n_elem), u in/out warm state (n_dof).
 * Returns total PCG iterations; compliance of the final solve in *comp. */
long ocpu_simp_run(int nx, int ny, int nz, const double* ke, int law, double penal, double E0,
                   double Emin, double rmin, double volfrac, double move, const double* f,
                   const uint8_t* fm, int n_iters, double tol, int max_iter, double* x,
                   double* xPhys, double* u, double* comp, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  const grid_t g = mk(nx, ny, nz);
  stencil_t s;
  mk_stencil(rmin, &s);
  double* Hs = malloc(g.ne * sizeof(double));
  double* ones = malloc(g.ne * sizeof(double));
  double* E = malloc(g.ne * sizeof(double));
  double* dc = malloc(g.ne * sizeof(double));
  double* dcf = malloc(g.ne * sizeof(double));
  double* xn = malloc(g.ne * sizeof(double));
  double* xp = malloc(g.ne * sizeof(double));
  for (long e = 0; e < g.ne; ++e) ones[e] = 1.0;
  for (long e = 0; e < g.ne; ++e) Hs[e] = 1.0;
  filt(&g, &s, Hs, ones, Hs, 0); /* Hs = H 1 (divide by 1) */
  long total = 0;
  int st;
  for (int it = 0; it < n_iters; ++it) {
    for (long e = 0; e < g.ne; ++e) E[e] = lawE(xPhys[e], law, penal, E0, Emin);
    total += ocpu_pcg(nx, ny, nz, ke, E, f, fm, u, it > 0, tol, max_iter, &st, NULL);
#pragma omp parallel for schedule(static)
    for (long e = 0; e < g.ne; ++e) {
      const int i = (int)(e / g.eI), rem = (int)(e % g.eI), j = rem / g.eJ, k = rem % g.eJ;
      const long n0 = (long)i * g.sI + (long)j * g.sJ + k;
      double ue[24];
      for (int m = 0; m < 8; ++m)
        for (int q = 0; q < 3; ++q) ue[3 * m + q] = u[3 * (n0 + CX[m] * g.sI + CY[m] * g.sJ + CZ[m]) + q];
      double ce = 0;
      for (int b = 0; b < 24; ++b) {
        double t = 0;
        for (int a = 0; a < 24; ++a) t += ue[a] * ke[a * 24 + b];
        ce += t * ue[b];
      }
      dc[e] = -lawdE(xPhys[e], law, penal, E0, Emin) * ce;
    }
    filt(&g, &s, Hs, dc, dcf, 1);
    double l1 = 0, l2 = 1e9;
    for (int b = 0; b < 50; ++b) {
      const double lmid = 0.5 * (l2 + l1);
#pragma omp parallel for schedule(static)
      for (long e = 0; e < g.ne; ++e) {
        double v = x[e] * sqrt(fmax(-dcf[e] / lmid, 1e-12));
        v = fmin(fmax(v, 0.0), 1.0);
        xn[e] = fmin(fmax(v, x[e] - move), x[e] + move);
      }
      filt(&g, &s, Hs, xn, xp, 0);
      if ((l2 - l1) / (l1 + l2 + 1e-12) < 1e-4) break;
      double sum = 0;
#pragma omp parallel for reduction(+ : sum) schedule(static)
      for (long e = 0; e < g.ne; ++e) sum += xp[e];
      if (sum / (double)g.ne > volfrac) l1 = lmid;
      else l2 = lmid;
    }
    memcpy(x, xn, g.ne * sizeof(double));
    memcpy(xPhys, xp, g.ne * sizeof(double));
  }
  for (long e = 0; e < g.ne; ++e) E[e] = lawE(xPhys[e], law, penal, E0, Emin);
  total += ocpu_pcg(nx, ny, nz, ke, E, f, fm, u, n_iters > 0, tol, max_iter, &st, NULL);
  double c = 0;
  for (long d = 0; d < g.nd; ++d) c += f[d] * u[d];
  *comp = c;
  free(Hs); free(ones); free(E); free(dc); free(dcf); free(xn); free(xp);
  return total;
}

int ocpu_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
