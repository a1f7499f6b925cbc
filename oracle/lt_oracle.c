/*
 * lt_oracle.c -- FP64 CPU restatement of the Lenstool-HPC lens-map path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library (see lt_oracle.h). It is written from
 * the mathematics of each reference function named below; the operation order inside a halo term
 * follows the reference's expressions so that agreement is at the rounding level.
 *
 * Pins (what fixes each function independently of this file; the tests hold them):
 *   gradient type 81          KAT-1 (Benchmarks/GradientBenchmark/setup.cpp:19-44), circular limit,
 *                             finite differences of the potential
 *   gradient types 5, 8       finite differences of the potential; type 8 = type-81 rcut -> inf
 *   Hessian type 81           golden grids from the compiled reference (src/gradient2.cpp),
 *                             finite differences of the gradient
 *   Hessian types 5, 8, 14    finite differences of the gradient; type 14 trace/determinant
 *   Kmap                      equals dlsds-weighted sum of per-run Hessians
 *   potential 5, 8, 81        gradient of the potential equals the gradient (finite differences)
 *   epilogues                 closed forms of kappa/gamma/lambda/mu (eigenvalues of the Jacobian)
 *   delensing, chi^2          golden outputs of the compiled reference (src/delense_CPU.cpp,
 *                             src/chi_computation.cpp)
 *   cosmology                 quadrature in the tests (scipy) and the Mattig closed form
 */
#include "lt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 0;
void orc_set_threads(int n) { g_threads = n; }
int orc_get_threads(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

/* ============================================================================================ */
/* per-halo terms in double (reference precision) and long double (extended)                    */
/* ============================================================================================ */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(n) CAT(n, SFX)

#define REAL double
#define SFX _d
#define ORACLE_EXTENDED 0
#define SQRT sqrt
#define LOG log
#define ATAN2 atan2
#define EXP exp
#define SIN sin
#define COS cos
#define ASINH asinh
#define ASIN asin
#include "lt_oracle_terms.inc"
#undef REAL
#undef SFX
#undef ORACLE_EXTENDED
#undef SQRT
#undef LOG
#undef ATAN2
#undef EXP
#undef SIN
#undef COS
#undef ASINH
#undef ASIN

#define REAL long double
#define SFX _ld
#define ORACLE_EXTENDED 1
#define SQRT sqrtl
#define LOG logl
#define ATAN2 atan2l
#define EXP expl
#define SIN sinl
#define COS cosl
#define ASINH asinhl
#define ASIN asinl
#include "lt_oracle_terms.inc"

int orc_point_gradient(lt_point* out, const lt_point* p, const lt_potential_soa* L, int nhalos)
{
    return point_gradient_d(out, p, L, nhalos);
}
int orc_point_gradient_ld(lt_point* out, const lt_point* p, const lt_potential_soa* L, int nhalos)
{
    return point_gradient_ld(out, p, L, nhalos);
}
int orc_point_gradient2(lt_matrix* out, const lt_point* p, const lt_potential_soa* L, int nhalos)
{
    return point_hessian_d(out, p, L, nhalos, 0);
}
int orc_point_gradient2_ld(lt_matrix* out, const lt_point* p, const lt_potential_soa* L, int nhalos)
{
    return point_hessian_ld(out, p, L, nhalos, 0);
}
int orc_point_kmap(lt_matrix* out, const lt_point* p, const lt_potential_soa* L, int nhalos)
{
    if (!L->dlsds) return LT_ERR_INVALID;
    return point_hessian_d(out, p, L, nhalos, 1);
}
int orc_point_potential(double* out, const lt_point* p, const lt_potential_soa* L, int nhalos)
{
    return point_potential_d(out, p, L, nhalos);
}
int orc_point_potential_ld(double* out, const lt_point* p, const lt_potential_soa* L, int nhalos)
{
    return point_potential_ld(out, p, L, nhalos);
}

int orc_point_gradient_abs(lt_point* out, const lt_point* p, const lt_potential_soa* L, int nhalos)
{
    double ax = 0., ay = 0.;
    for (int i = 0; i < nhalos; ++i) {
        double gx, gy;
        int st = grad_term_d(p, L, i, &gx, &gy);
        if (st) return st;
        ax += fabs(gx);
        ay += fabs(gy);
    }
    out->x = ax;
    out->y = ay;
    return LT_OK;
}

static int point_hessian_abs(lt_matrix* out, const lt_point* p, const lt_potential_soa* L, int nhalos, int kmap)
{
    mat_d g = {0., 0., 0., 0.};
    for (int i = 0; i < nhalos; ++i) {
        mat_d h;
        int st = hess_term_d(p, L, i, &h);
        if (st) return st;
        double w = kmap ? fabs(L->dlsds[i]) : 1.;
        g.a += fabs(h.a) * w;
        g.b += fabs(h.b) * w;
        g.c += fabs(h.c) * w;
        g.d += fabs(h.d) * w;
    }
    out->a = g.a;
    out->b = g.b;
    out->c = g.c;
    out->d = g.d;
    return LT_OK;
}

int orc_point_gradient2_abs(lt_matrix* out, const lt_point* p, const lt_potential_soa* L, int nhalos)
{
    return point_hessian_abs(out, p, L, nhalos, 0);
}

/* ============================================================================================ */
/* grid drivers                                                                                  */
/* ============================================================================================ */

static int check_grid(const lt_grid_param* frame, const lt_potential_soa* L, int nhalos, int nx, int ny)
{
    if (!frame || !L || nhalos < 0 || nx <= 0 || ny <= 0) return LT_ERR_INVALID;
    if (nhalos > 0 && (!L->type || !L->position_x || !L->position_y || !L->anglecos || !L->anglesin))
        return LT_ERR_INVALID;
    return LT_OK;
}

/* scan the types once so the parallel loops cannot fail half way */
static int check_types(const lt_potential_soa* L, int nhalos, int want)
{
    for (int i = 0; i < nhalos; ++i) {
        int t = L->type[i];
        int ok = (t == 5 || t == 8 || t == 81) || (want == 2 && t == 14);
        if (!ok) return LT_ERR_TYPE;
    }
    return LT_OK;
}

#define PIXEL(frame, col, row, istart, jstart, dx, dy, pt)           \
    do {                                                             \
        (pt).x = (frame)->xmin + (double)((col) + (istart)) * (dx);   \
        (pt).y = (frame)->ymin + (double)((row) + (jstart)) * (dy);   \
    } while (0)

int orc_gradient_grid_p(double* gx, double* gy, const lt_grid_param* frame, const lt_potential_soa* L,
                        int nhalos, double dx, double dy, int nx, int ny, int istart, int jstart, int prec)
{
    int st = check_grid(frame, L, nhalos, nx, ny);
    if (st) return st;
    if (!gx || !gy) return LT_ERR_INVALID;
    if ((st = check_types(L, nhalos, 1))) return st;
#pragma omp parallel for schedule(dynamic, 1) num_threads(orc_get_threads())
    for (int row = 0; row < ny; ++row)
        for (int col = 0; col < nx; ++col) {
            lt_point pt, g = {0., 0.};
            PIXEL(frame, col, row, istart, jstart, dx, dy, pt);
            if (prec) point_gradient_ld(&g, &pt, L, nhalos);
            else point_gradient_d(&g, &pt, L, nhalos);
            long idx = (long)row * nx + col;
            gx[idx] = g.x;
            gy[idx] = g.y;
        }
    return LT_OK;
}

int orc_gradient_grid(double* gx, double* gy, const lt_grid_param* frame, const lt_potential_soa* L,
                      int nhalos, double dx, double dy, int nx, int ny, int istart, int jstart)
{
    return orc_gradient_grid_p(gx, gy, frame, L, nhalos, dx, dy, nx, ny, istart, jstart, 0);
}

int orc_gradient_abs_planes(double* ax, double* ay, const lt_grid_param* frame, const lt_potential_soa* L,
                            int nhalos, double dx, double dy, int nx, int ny, int istart, int jstart)
{
    int st = check_grid(frame, L, nhalos, nx, ny);
    if (st) return st;
    if ((st = check_types(L, nhalos, 1))) return st;
#pragma omp parallel for schedule(dynamic, 1) num_threads(orc_get_threads())
    for (int row = 0; row < ny; ++row)
        for (int col = 0; col < nx; ++col) {
            lt_point pt, g;
            PIXEL(frame, col, row, istart, jstart, dx, dy, pt);
            orc_point_gradient_abs(&g, &pt, L, nhalos);
            long idx = (long)row * nx + col;
            if (ax) ax[idx] = g.x;
            if (ay) ay[idx] = g.y;
        }
    return LT_OK;
}

int orc_gradient2_grid(lt_matrix* out, const lt_grid_param* frame, const lt_potential_soa* L,
                       int nhalos, double dx, double dy, int nx, int ny, int istart, int jstart)
{
    int st = check_grid(frame, L, nhalos, nx, ny);
    if (st) return st;
    if (!out) return LT_ERR_INVALID;
    if ((st = check_types(L, nhalos, 2))) return st;
#pragma omp parallel for schedule(dynamic, 1) num_threads(orc_get_threads())
    for (int row = 0; row < ny; ++row)
        for (int col = 0; col < nx; ++col) {
            lt_point pt;
            PIXEL(frame, col, row, istart, jstart, dx, dy, pt);
            point_hessian_d(&out[(long)row * nx + col], &pt, L, nhalos, 0);
        }
    return LT_OK;
}

int orc_gradient2_planes(double* a, double* b, double* c, double* d, double* abs_a, double* abs_b,
                         double* abs_c, const lt_grid_param* frame, const lt_potential_soa* L, int nhalos,
                         double dx, double dy, int nx, int ny, int istart, int jstart, int kmap,
                         int prec)
{
    int st = check_grid(frame, L, nhalos, nx, ny);
    if (st) return st;
    if ((st = check_types(L, nhalos, 2))) return st;
    if (kmap && !L->dlsds) return LT_ERR_INVALID;
#pragma omp parallel for schedule(dynamic, 1) num_threads(orc_get_threads())
    for (int row = 0; row < ny; ++row)
        for (int col = 0; col < nx; ++col) {
            lt_point pt;
            lt_matrix h = {0., 0., 0., 0.};
            PIXEL(frame, col, row, istart, jstart, dx, dy, pt);
            if (prec) point_hessian_ld(&h, &pt, L, nhalos, kmap);
            else point_hessian_d(&h, &pt, L, nhalos, kmap);
            long idx = (long)row * nx + col;
            if (a) a[idx] = h.a;
            if (b) b[idx] = h.b;
            if (c) c[idx] = h.c;
            if (d) d[idx] = h.d;
            if (abs_a || abs_b || abs_c) {
                point_hessian_abs(&h, &pt, L, nhalos, kmap);
                if (abs_a) abs_a[idx] = h.a;
                if (abs_b) abs_b[idx] = h.b;
                if (abs_c) abs_c[idx] = h.c;
            }
        }
    return LT_OK;
}

int orc_potential_grid_p(double* phi, const lt_grid_param* frame, const lt_potential_soa* L, int nhalos,
                         double dx, double dy, int nx, int ny, int istart, int jstart, int prec)
{
    int st = check_grid(frame, L, nhalos, nx, ny);
    if (st) return st;
    if (!phi) return LT_ERR_INVALID;
    if ((st = check_types(L, nhalos, 1))) return st;
#pragma omp parallel for schedule(dynamic, 1) num_threads(orc_get_threads())
    for (int row = 0; row < ny; ++row)
        for (int col = 0; col < nx; ++col) {
            lt_point pt;
            PIXEL(frame, col, row, istart, jstart, dx, dy, pt);
            if (prec) point_potential_ld(&phi[(long)row * nx + col], &pt, L, nhalos);
            else point_potential_d(&phi[(long)row * nx + col], &pt, L, nhalos);
        }
    return LT_OK;
}

int orc_potential_grid(double* phi, const lt_grid_param* frame, const lt_potential_soa* L, int nhalos,
                       double dx, double dy, int nx, int ny, int istart, int jstart)
{
    return orc_potential_grid_p(phi, frame, L, nhalos, dx, dy, nx, ny, istart, jstart, 0);
}


/* ---- FP64 rounding-error scale of the PIEMD closed forms ------------------------------------
 * The complex ratio Z = N/D of the PIEMD term (src/gradient.hpp:64-76) is 0/0 at (x, y) =
 * (0, 2 r sqrt(eps)) in the halo frame (N and D vanish together), so near that point the closed
 * forms of the gradient and of the Hessian (src/gradient2.cpp:54-82) cancel catastrophically
 * although the field itself is smooth. First-order error analysis of the two differences
 * t = 2 sqrt(eps) w - y/cx1 (in N and den1) and n2 = 2 r sqrt(eps) - y (in D and den2) gives the
 * absolute rounding error of one evaluation in units of the unit roundoff u = 2^-53:
 *   gradient: b0' ci (M/|N| + M2/|D|),
 *   Hessian:  b0' ci ((|A| 2|t| M/den1 + cx1 M)/den1 + 3 M2/den2),
 * with M = 2 sqrt(eps) w + |y|/cx1, M2 = 2 r sqrt(eps) + |y|, A the den1 numerator and b0' the
 * halo amplitude. Away from the point these reduce to a few ulp of the term. The sum over halos
 * and terms, times u, is the scale the tests use for "within FP64 rounding of the formula". */
static void cond_piemd(double x, double y, double eps, double r, double amp, double* cg, double* ch)
{
    double sqe = sqrt(eps);
    double cx1 = (1. - eps) / (1. + eps);
    double ci = 0.5 * (1. - eps * eps) / sqe;
    double w = sqrt(r * r + x * x / ((1. + eps) * (1. + eps)) + y * y / ((1. - eps) * (1. - eps)));
    double t = 2. * sqe * w - y / cx1;
    double M = 2. * sqe * w + fabs(y) / cx1;
    double n2 = 2. * r * sqe - y;
    double M2 = 2. * r * sqe + fabs(y);
    double nN = sqrt(cx1 * cx1 * x * x + t * t);
    double nD = sqrt(x * x + n2 * n2);
    double den1 = nN * nN, den2 = nD * nD;
    double A = cx1 * (2. * sqe * x * x / ((1. + eps) * (1. + eps)) / w + M) + fabs(x) + fabs(y) / (cx1 * cx1);
    double k = fabs(amp) * ci;
    *cg = k * (M / nN + M2 / nD + 4.);
    *ch = k * ((A * 2. * fabs(t) * M / den1 + cx1 * M + A) / den1 + 3. * M2 / den2);
}

int orc_point_cond(double* cg, double* ch, const lt_point* p, const lt_potential_soa* L, int nhalos)
{
    const double u = 1.1102230246251565e-16;
    double sg = 0., sh = 0.;
    for (int i = 0; i < nhalos; ++i) {
        double x, y;
        to_halo_frame_d(p, L, i, &x, &y);
        int t = L->type[i];
        double g1 = 0., h1 = 0., g2 = 0., h2 = 0.;
        if (t == 8 || t == 81) {
            double rc = L->rcore[i], rcut = L->rcut[i], b0 = L->b0[i];
            double amp = (t == 81) ? b0 * rcut / (rcut - rc) : b0;
            cond_piemd(x, y, L->ellipticity_potential[i], rc, amp, &g1, &h1);
            if (t == 81) cond_piemd(x, y, L->ellipticity_potential[i], rcut, amp, &g2, &h2);
        } else if (t == 5) {
            double R = sqrt(x * x + y * y);
            g1 = 8. * L->b0[i];
            h1 = 8. * L->b0[i] / R;
        }
        sg += g1 + g2;
        sh += h1 + h2;
    }
    if (cg) *cg = u * sg;
    if (ch) *ch = u * sh;
    return LT_OK;
}

int orc_cond_planes(double* cg, double* ch, const lt_grid_param* frame, const lt_potential_soa* L, int nhalos,
                    double dx, double dy, int nx, int ny, int istart, int jstart)
{
    int st = check_grid(frame, L, nhalos, nx, ny);
    if (st) return st;
#pragma omp parallel for schedule(dynamic, 1) num_threads(orc_get_threads())
    for (int row = 0; row < ny; ++row)
        for (int col = 0; col < nx; ++col) {
            lt_point pt;
            PIXEL(frame, col, row, istart, jstart, dx, dy, pt);
            long idx = (long)row * nx + col;
            double g, h;
            orc_point_cond(&g, &h, &pt, L, nhalos);
            if (cg) cg[idx] = g;
            if (ch) ch[idx] = h;
        }
    return LT_OK;
}

/* src/grid_amplif_CPU.cpp:38-84: only mode 5 (kappa = (a + c)/2, no dlsds) exists on the CPU. */
int orc_amplif_grid(double* out, const lt_grid_param* frame, const lt_potential_soa* L, int nhalos,
                    int nbgridcells, int mode_amp)
{
    if (mode_amp != 5) return LT_ERR_INVALID;
    if (nbgridcells < 2) return LT_ERR_INVALID;
    int st = check_grid(frame, L, nhalos, nbgridcells, nbgridcells);
    if (st) return st;
    if ((st = check_types(L, nhalos, 2))) return st;
    double dx = (frame->xmax - frame->xmin) / (nbgridcells - 1);
    double dy = (frame->ymax - frame->ymin) / (nbgridcells - 1);
#pragma omp parallel for schedule(dynamic, 1) num_threads(orc_get_threads())
    for (int row = 0; row < nbgridcells; ++row)
        for (int col = 0; col < nbgridcells; ++col) {
            lt_point pt;
            lt_matrix h = {0., 0., 0., 0.};
            PIXEL(frame, col, row, 0, 0, dx, dy, pt);
            orc_point_gradient2(&h, &pt, L, nhalos);
            out[(long)row * nbgridcells + col] = (h.a + h.c) / 2.;
        }
    return LT_OK;
}

/* ============================================================================================ */
/* map epilogues                                                                                 */
/* ============================================================================================ */

/* eigenvalues of the Jacobian, formeli_ampli (src/grid_map_ampli_GPU.cu:549-572) */
static inline void formeli(double a, double b, double c, double* lambda, double* mu)
{
    double delta = (a - c) * (a - c) + 4 * b * b;
    double e = sqrt(delta);
    *lambda = .5 * (a + c + e);
    *mu = .5 * (a + c - e);
}

int orc_map_epilogue(int kind, int mode, double* o0, double* o1, double* o2, double* o3, double* o4,
                     const double* a, const double* b, const double* c, long n, double dl0s, double dos,
                     double dol, double h, double dx, double dy)
{
    double r = dl0s / dos;
    double mult = 1.;
    if (kind == LT_MAP_MASS) {
        switch (mode) {
        case 1: mult = 1.; break;                                            /* :233 */
        case 2: mult = LT_CH4PIG * h / dol / dl0s * dos; break;               /* :246 */
        case 3: mult = LT_MCRIT12 / h * dx * dy * dol / dl0s * dos; break;    /* :270 */
        case 4: mult = LT_CH0_4PIG * h / dol / dl0s * dos; break;             /* :291 */
        default: return LT_ERR_INVALID;
        }
    } else if (kind == LT_MAP_AMPLI) {
        if (mode < 1 || mode > 6) return LT_ERR_INVALID;
    } else if (kind == LT_MAP_SHEAR) {
        if (mode < 1 || mode > 4) return LT_ERR_INVALID;
    } else if (kind != LT_MAP_FIELDS) {
        return LT_ERR_INVALID;
    }
#pragma omp parallel for schedule(static) num_threads(orc_get_threads())
    for (long i = 0; i < n; ++i) {
        double A = 1. - a[i] * r, B = -b[i] * r, C = 1. - c[i] * r;
        double lam, mu;
        if (kind == LT_MAP_AMPLI) {
            double v;
            switch (mode) {
            case 1: formeli(A, B, C, &lam, &mu); v = 1. / (lam * mu); break;
            case 2: formeli(A, B, C, &lam, &mu); v = 1. / fabs(lam * mu); break;
            case 3: formeli(A, B, C, &lam, &mu); v = -2.5 * log10(fabs(lam * mu)); break;
            case 4: {
                double kappa = (a[i] + c[i]) * r / 2.;
                double ga1 = (a[i] - c[i]) * r / 2.;
                double ga2 = b[i] * r;
                double gam = sqrt(ga1 * ga1 + ga2 * ga2);
                double gp = gam / (1 - kappa);
                v = (1 - kappa) * (1 + gp * gp) / (1 - gp * gp);
                break;
            }
            case 5: v = (a[i] + c[i]) / 2.; break;
            default: {
                double ga1 = (a[i] - c[i]) / 2.;
                double ga2 = b[i];
                v = sqrt(ga1 * ga1 + ga2 * ga2);
                break;
            }
            }
            o0[i] = v;
        } else if (kind == LT_MAP_SHEAR) {
            double v;
            switch (mode) {
            case 1: formeli(A, B, C, &lam, &mu); v = (lam - mu) / 2.; break;
            case 2: {
                formeli(A, B, C, &lam, &mu);
                double q = lam / mu;
                v = fabs((q * q - 1.) / (q * q + 1.));
                break;
            }
            case 3: v = (a[i] - c[i]) * r / 2.; break;
            default: v = b[i] * r; break;
            }
            o0[i] = v;
        } else if (kind == LT_MAP_MASS) {
            o0[i] = (a[i] + c[i]) * 0.5 * mult;
        } else {
            /* kappa, gamma1, gamma2, det A = (1-kappa)^2 - gamma^2, mu = 1/det A */
            double kappa = (a[i] + c[i]) * r / 2.;
            double g1 = (a[i] - c[i]) * r / 2.;
            double g2 = b[i] * r;
            double omk = 1. - kappa;
            double det = omk * omk - (g1 * g1 + g2 * g2);
            if (o0) o0[i] = kappa;
            if (o1) o1[i] = g1;
            if (o2) o2[i] = g2;
            if (o3) o3[i] = det;
            if (o4) o4[i] = 1. / det;
        }
    }
    return LT_OK;
}

/* ============================================================================================ */
/* cosmology                                                                                     */
/* ============================================================================================ */

/* 1/E(z) for the CPL model 1 (src/module_cosmodistances.cpp:217-264) */
static double inv_e(double z, const lt_cosmo_param* C)
{
    double x = 1 + z;
    double e2 = -x * x * C->curvature + x * x * x * C->omegaM +
                C->omegaX * pow(x, 3 * (1 + C->wX + C->wa)) * exp(-3 * C->wa * z / x);
    return 1. / sqrt(e2);
}

/* composite Gauss-Legendre (10 points) on 512 panels: far below the reference's GSL QAG
 * tolerance (epsrel 1e-6, src/module_cosmodistances.cpp:329-337). */
static double integral_inv_e(double a, double b, const lt_cosmo_param* C)
{
    static const double xg[5] = {0.1488743389816312, 0.4333953941292472, 0.6794095682990244,
                                 0.8650633666889845, 0.9739065285171717};
    static const double wg[5] = {0.2955242247147529, 0.2692667193099963, 0.2190863625159820,
                                 0.1494513491505806, 0.0666713443086881};
    const int panels = 512;
    double h = (b - a) / panels, sum = 0.;
    for (int k = 0; k < panels; ++k) {
        double m = a + (k + 0.5) * h, r = 0.5 * h;
        for (int j = 0; j < 5; ++j) sum += wg[j] * (inv_e(m - r * xg[j], C) + inv_e(m + r * xg[j], C)) * r;
    }
    return sum;
}

static double sk(double x, double k)
{
    if (k > 0) return sin(x);
    if (k < 0) return sinh(x);
    return x;
}

double orc_cosmo_observer_object(double z, const lt_cosmo_param* C)
{
    if (C->omegaX == 0.) {
        double g = sqrt(1. + C->omegaM * z);
        return 2. * ((1. - C->omegaM - g) * (1. - g)) / C->omegaM / C->omegaM / (1. + z) / (1. + z);
    }
    double chi = integral_inv_e(0., z, C);
    if (C->curvature != 0.)
        return sk(chi * sqrt(fabs(C->curvature)), C->curvature) / (1 + z) / sqrt(fabs(C->curvature));
    return chi / (1 + z);
}

double orc_cosmo_object_object(double z1, double z2, const lt_cosmo_param* C)
{
    if (z1 >= z2) return 0.;
    if (C->omegaX == 0.) {
        double g1 = sqrt(1. + C->omegaM * z1), g2 = sqrt(1. + C->omegaM * z2);
        return 2. * ((1. - C->omegaM - g1 * g2) * (g1 - g2)) / C->omegaM / C->omegaM / (1. + z1) /
               (1. + z2) / (1. + z2);
    }
    double chi = integral_inv_e(z1, z2, C);
    if (C->curvature != 0.)
        return sk(chi * sqrt(fabs(C->curvature)), C->curvature) / (1 + z2) / sqrt(fabs(C->curvature));
    return chi / (1 + z2);
}

double orc_cosmo_dlsds(double zl, double zs, const lt_cosmo_param* C)
{
    if (zl >= zs) return 0.;
    return orc_cosmo_object_object(zl, zs, C) / orc_cosmo_observer_object(zs, C);
}

/* ============================================================================================ */
/* delensing and chi^2                                                                           */
/* ============================================================================================ */

/* scalar triple product of (A,1), (B,1), (C,1) (src/delense_CPU_utils.hpp:110-117) */
static inline double tri_det(const lt_point* A, const lt_point* B, const lt_point* C)
{
    return (B->x * C->y - B->y * C->x + A->x * B->y - A->y * B->x + A->y * C->x - A->x * C->y);
}

/* inside test with the border included; the reference's mychi_inside and mychi_insideborder
 * (src/delense_CPU_utils.hpp:125-142, 180-193) are the same test. */
static inline int tri_inside(const lt_point* P, const lt_triplet* T)
{
    double d = tri_det(&T->a, &T->b, &T->c);
    double s = tri_det(&T->a, &T->b, P) * d;
    if (s < 0.) return 0;
    double s1 = tri_det(&T->b, &T->c, P) * d;
    if (s1 < 0.) return 0;
    double s2 = tri_det(&T->c, &T->a, P) * d;
    if (s2 < 0.) return 0;
    return 1;
}

int orc_delense_barycenter(lt_point* image_pos, int* locimagesfound, int* numimg, int nsets, int n,
                           const lt_grid_param* frame, const lt_galaxy* sources, const double* gx,
                           const double* gy)
{
    if (!image_pos || !locimagesfound || !numimg || !frame || !sources || !gx || !gy || n < 2 || nsets < 0)
        return LT_ERR_INVALID;
    const double dx = (frame->xmax - frame->xmin) / (n - 1);
    const double dy = (frame->ymax - frame->ymin) / (n - 1);
    int status = LT_OK;
    for (int sid = 0; sid < nsets; ++sid) {
        int found = 0;
        const double dr = sources[sid].dr;
        const lt_point P = sources[sid].center;
        for (int yid = 0; yid < n - 1; ++yid)
            for (int xid = 0; xid < n - 1; ++xid) {
                double xp = frame->xmin + xid * dx;
                double yp = frame->ymin + yid * dy;
                long g = (long)yid * n + xid;
                lt_triplet Ts, Ti, S;
                /* upper triangle a=(x,y) b=(x,y+dy) c=(x+dx,y), gradient indices g, g+n, g+1 */
                Ts.a.x = xp;       Ts.a.y = yp;
                Ts.b.x = xp;       Ts.b.y = yp + dy;
                Ts.c.x = xp + dx;  Ts.c.y = yp;
                /* lower triangle a=(x+dx,y+dy) b=(x+dx,y) c=(x,y+dy), indices g+n+1, g+1, g+n */
                Ti.a.x = xp + dx;  Ti.a.y = yp + dy;
                Ti.b.x = xp + dx;  Ti.b.y = yp;
                Ti.c.x = xp;       Ti.c.y = yp + dy;
                S.a.x = Ts.a.x - dr * gx[g];         S.a.y = Ts.a.y - dr * gy[g];
                S.b.x = Ts.b.x - dr * gx[g + n];     S.b.y = Ts.b.y - dr * gy[g + n];
                S.c.x = Ts.c.x - dr * gx[g + 1];     S.c.y = Ts.c.y - dr * gy[g + 1];
                const lt_triplet* hit = 0;
                if (tri_inside(&P, &S)) {
                    hit = &Ts;
                } else {
                    S.a.x = Ti.a.x - dr * gx[g + n + 1]; S.a.y = Ti.a.y - dr * gy[g + n + 1];
                    S.b.x = Ti.b.x - dr * gx[g + 1];     S.b.y = Ti.b.y - dr * gy[g + 1];
                    S.c.x = Ti.c.x - dr * gx[g + n];     S.c.y = Ti.c.y - dr * gy[g + n];
                    if (tri_inside(&P, &S)) hit = &Ti;
                }
                if (hit) {
                    if (found < LT_MAXIMPERSOURCE) {
                        lt_point bc;
                        bc.x = (hit->a.x + hit->b.x + hit->c.x) / 3.;
                        bc.y = (hit->a.y + hit->b.y + hit->c.y) / 3.;
                        image_pos[sid * LT_MAXIMPERSOURCE + found] = bc;
                    } else {
                        status = LT_ERR_OVERFLOW;
                    }
                    found++;
                    *numimg = *numimg + 1;
                }
            }
        locimagesfound[sid] = found < LT_MAXIMPERSOURCE ? found : LT_MAXIMPERSOURCE;
    }
    return status;
}

int orc_source_barycenters(lt_galaxy* sources, int nsets, const int* nimages, const lt_galaxy* images,
                           const lt_potential_soa* L, int nhalos)
{
    int index = 0;
    for (int sid = 0; sid < nsets; ++sid) {
        int ni = nimages[sid];
        if (ni <= 0) return LT_ERR_INVALID;
        double sx = 0., sy = 0.;
        for (int k = 0; k < ni; ++k) {
            lt_point g;
            int st = orc_point_gradient(&g, &images[index + k].center, L, nhalos);
            if (st) return st;
            sx += images[index + k].center.x - images[index + k].dr * g.x;
            sy += images[index + k].center.y - images[index + k].dr * g.y;
        }
        memset(&sources[sid], 0, sizeof(lt_galaxy));
        sources[sid].center.x = sx / ni;
        sources[sid].center.y = sy / ni;
        sources[sid].redshift = images[index].redshift;
        sources[sid].dr = images[index].dr;
        sources[sid].dls = images[index].dls;
        sources[sid].dos = images[index].dos;
        index += ni;
    }
    return LT_OK;
}

static inline double pdist(lt_point A, lt_point B)
{
    double x = A.x - B.x, y = A.y - B.y;
    return sqrt(x * x + y * y);
}

/* src/chi_computation.cpp:10-195, including its two loop conditions that stop early
 * (:110 `i != image_index` and :118 `nimagesfound[i] == 0`). */
int orc_chi_computation(double* chi, int* images_found, const int* numimagesfound,
                        const lt_point* imagesposition, const int* nimages, const lt_galaxy* images, int nsets)
{
    lt_point tim[LT_MAXIMPERSOURCE];
    memset(tim, 0, sizeof(tim));
    int index = 0;
    for (int sid = 0; sid < nsets; ++sid) {
        int ni = nimages[sid];
        if (ni <= 0 || ni > LT_MAXIMPERSOURCE) return LT_ERR_INVALID;
        int nfound[LT_MAXIMPERSOURCE];
        memset(nfound, 0, sizeof(nfound));
        double image_dist[LT_MAXIMPERSOURCE];
        int nf = numimagesfound[sid];
        if (nf > LT_MAXIMPERSOURCE) nf = LT_MAXIMPERSOURCE;
        for (int ii = 0; ii < nf; ++ii) {
            int image_index = 0;
            lt_point pos = imagesposition[sid * LT_MAXIMPERSOURCE + ii];
            image_dist[0] = pdist(pos, images[index + 0].center);
            for (int i = 1; i < ni; i++) {
                image_dist[i] = pdist(pos, images[index + i].center);
                if (image_dist[i] < image_dist[image_index]) image_index = i;
            }
            (*images_found)++;
            if (nfound[image_index] == 0) {
                tim[image_index] = pos;
                nfound[image_index]++;
            } else if (nfound[image_index] > 0) {
                lt_point temp;
                if (image_dist[image_index] < pdist(images[index + image_index].center, tim[image_index])) {
                    temp = tim[image_index];
                    tim[image_index] = pos;
                } else {
                    temp = pos;
                }
                int second = 0;
                for (int i = 1; i < ni && (i != image_index); i++)
                    if (image_dist[i] > image_dist[second]) second = i;
                for (int i = 0; i < ni && nfound[i] == 0; i++) {
                    if (image_dist[i] < image_dist[second]) {
                        second = i;
                        image_index = i;
                        tim[image_index] = temp;
                    }
                }
                nfound[image_index]++;
            }
        }
        for (int i = 0; i < ni; i++) {
            if (nfound[i] > 0) {
                double p1 = images[index + i].center.x - tim[i].x;
                double p2 = images[index + i].center.y - tim[i].y;
                *chi += p1 * p1 + p2 * p2;
            } else if (nfound[i] == 0) {
                *chi += 1000. * ni;
            }
        }
        index += ni;
    }
    return LT_OK;
}

int orc_chi_barycentersource(double* chi, int* error, int n, int nsets, int nhalos, const lt_potential_soa* L,
                             const lt_grid_param* frame, const int* nimages, const lt_galaxy* images)
{
    if (!chi || !frame || !L || !nimages || !images || n < 2 || nsets <= 0) return LT_ERR_INVALID;
    double dx = (frame->xmax - frame->xmin) / (n - 1);
    double dy = (frame->ymax - frame->ymin) / (n - 1);
    long N = (long)n * n;
    double* gx = (double*)malloc(N * sizeof(double));
    double* gy = (double*)malloc(N * sizeof(double));
    lt_galaxy* src = (lt_galaxy*)calloc(nsets, sizeof(lt_galaxy));
    lt_point* pos = (lt_point*)calloc((size_t)nsets * LT_MAXIMPERSOURCE, sizeof(lt_point));
    int* loc = (int*)calloc(nsets, sizeof(int));
    int st = LT_ERR_NOMEM;
    if (gx && gy && src && pos && loc) {
        st = orc_gradient_grid(gx, gy, frame, L, nhalos, dx, dy, n, n, 0, 0);
        if (!st) st = orc_source_barycenters(src, nsets, nimages, images, L, nhalos);
        int numimg = 0;
        int dst = LT_OK;
        if (!st) dst = orc_delense_barycenter(pos, loc, &numimg, nsets, n, frame, src, gx, gy);
        if (!st) {
            int found = 0;
            *chi = 0.;
            st = orc_chi_computation(chi, &found, loc, pos, nimages, images, nsets);
            if (!st && dst) st = dst;
        }
    }
    if (error) *error = st;
    free(gx);
    free(gy);
    free(src);
    free(pos);
    free(loc);
    return st;
}

/* ---- critical lines (marching squares; see lt_oracle.h) ----------------------------------------- */

static double orc_coord(double min, int i, double d)
{
    double t = (double)i * d;
    return min + t;
}

/* crossing on edge e (0 bottom, 1 right, 2 top, 3 left) of cell (gi, gj), corners c[4] */
static void orc_edge_point(int e, const double* c, int gi, int gj, double xmin, double ymin, double dx, double dy,
                           double* x, double* y, double* t)
{
    if (e == 0) {
        *t = c[0] / (c[0] - c[1]);
        *x = orc_coord(xmin, gi, dx) + *t * dx;
        *y = orc_coord(ymin, gj, dy);
    } else if (e == 1) {
        *t = c[1] / (c[1] - c[2]);
        *x = orc_coord(xmin, gi + 1, dx);
        *y = orc_coord(ymin, gj, dy) + *t * dy;
    } else if (e == 2) {
        *t = c[3] / (c[3] - c[2]);
        *x = orc_coord(xmin, gi, dx) + *t * dx;
        *y = orc_coord(ymin, gj + 1, dy);
    } else {
        *t = c[0] / (c[0] - c[3]);
        *x = orc_coord(xmin, gi, dx);
        *y = orc_coord(ymin, gj, dy) + *t * dy;
    }
}

static double orc_edge_value(int e, double t, const double* k)
{
    static const int a[4] = {0, 1, 3, 0}, b[4] = {1, 2, 2, 3};
    return k[a[e]] + t * (k[b[e]] - k[a[e]]);
}

/* edge pairs per corner mask: each entry lists up to two (from, to) segments; saddles resolved below */
static int orc_cell_table(const double* c, int* e)
{
    int m = (c[0] > 0.) | ((c[1] > 0.) << 1) | ((c[2] > 0.) << 2) | ((c[3] > 0.) << 3);
    /* single segments: from the edge where the positive side begins, positive on the left */
    static const int from[16] = {-1, 0, 1, 1, 2, -1, 2, 2, 3, 0, -1, 1, 3, 0, 3, -1};
    static const int to[16] = {-1, 3, 0, 3, 1, -1, 0, 3, 2, 2, -1, 2, 1, 1, 0, -1};
    double sum = (c[0] + c[1]) + (c[2] + c[3]);
    if (m == 5) {
        if (sum > 0.) { e[0] = 0; e[1] = 1; e[2] = 2; e[3] = 3; }
        else { e[0] = 0; e[1] = 3; e[2] = 2; e[3] = 1; }
        return 2;
    }
    if (m == 10) {
        if (sum > 0.) { e[0] = 3; e[1] = 0; e[2] = 1; e[3] = 2; }
        else { e[0] = 1; e[1] = 0; e[2] = 3; e[3] = 2; }
        return 2;
    }
    if (from[m] < 0) return 0;
    e[0] = from[m];
    e[1] = to[m];
    return 1;
}

int orc_critic_lines(double* segs, int* kinds, long max_seg, long* nseg, const double* det, const double* kappa,
                     int nx, int ny, long pitch, double xmin, double ymin, double dx, double dy, int istart,
                     int jstart)
{
    if (!det || !nseg || nx < 2 || ny < 2 || pitch < nx) return LT_ERR_INVALID;
    if (kinds && !kappa) return LT_ERR_INVALID;
    long n = 0;
    for (int j = 0; j < ny - 1; ++j) {
        for (int i = 0; i < nx - 1; ++i) {
            const double* r0 = det + (long)j * pitch;
            const double* r1 = r0 + pitch;
            double c[4] = {r0[i], r0[i + 1], r1[i + 1], r1[i]};
            int e[4];
            int ns = orc_cell_table(c, e);
            for (int q = 0; q < ns; ++q, ++n) {
                if (!segs || n >= max_seg) continue;
                double xa, ya, ta, xb, yb, tb;
                orc_edge_point(e[2 * q], c, istart + i, jstart + j, xmin, ymin, dx, dy, &xa, &ya, &ta);
                orc_edge_point(e[2 * q + 1], c, istart + i, jstart + j, xmin, ymin, dx, dy, &xb, &yb, &tb);
                segs[4 * n] = xa;
                segs[4 * n + 1] = ya;
                segs[4 * n + 2] = xb;
                segs[4 * n + 3] = yb;
                if (kinds) {
                    const double* k0 = kappa + (long)j * pitch;
                    const double* k1 = k0 + pitch;
                    double k[4] = {k0[i], k0[i + 1], k1[i + 1], k1[i]};
                    double ka = orc_edge_value(e[2 * q], ta, k), kb = orc_edge_value(e[2 * q + 1], tb, k);
                    kinds[n] = 1. - 0.5 * (ka + kb) > 0. ? 0 : 1;
                }
            }
        }
    }
    *nseg = n;
    return (segs && n > max_seg) ? LT_ERR_OVERFLOW : LT_OK;
}
