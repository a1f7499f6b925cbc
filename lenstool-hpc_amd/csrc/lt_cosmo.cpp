// lt_cosmo.cpp -- angular-diameter distances for the map epilogues (D_ls/D_s, D_l), in units of
// c/H0, for the reference's cosmology model 1 (CPL dark energy with curvature):
//   E(z)^2 = -(1+z)^2 Ok + (1+z)^3 Om + OX (1+z)^(3(1+wX+wa)) exp(-3 wa z/(1+z))
//   D_A(z1, z2) = S_k(int_{z1}^{z2} dz/E(z)) / (1 + z2)      (src/module_cosmodistances.cpp:94-174)
// and the Mattig closed form when OX = 0 (src/module_cosmodistances.cpp:99-105, 135-142).
// The reference integrates with GSL QAG (epsrel 1e-6, 15-point rule, src/module_cosmodistances.cpp:
// 313-338); GSL is not a dependency here: the integral is done with an adaptive Gauss-Kronrod
// (7/15-point) rule to a relative tolerance of 1e-13.
#include <cmath>

#include "lt_kernels.h"

namespace lt {

static double inv_e(double z, const lt_cosmo_param& C)
{
    const double x = 1.0 + z;
    const double e2 = -x * x * C.curvature + x * x * x * C.omegaM +
                      C.omegaX * std::pow(x, 3.0 * (1.0 + C.wX + C.wa)) * std::exp(-3.0 * C.wa * z / x);
    return 1.0 / std::sqrt(e2);
}

// Gauss-Kronrod 7-15 nodes and weights (Kronrod extension of the 7-point Gauss rule).
static const double XK[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                             0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                             0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                             0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
static const double WK[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                             0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                             0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                             0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
static const double WG[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                             0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

static double gk15(double a, double b, const lt_cosmo_param& C, double* err)
{
    const double c = 0.5 * (a + b), h = 0.5 * (b - a);
    const double fc = inv_e(c, C);
    double rk = fc * WK[7], rg = fc * WG[3];
    for (int j = 0; j < 7; ++j) {
        const double f1 = inv_e(c - h * XK[j], C), f2 = inv_e(c + h * XK[j], C);
        rk += WK[j] * (f1 + f2);
        if (j % 2 == 1) rg += WG[j / 2] * (f1 + f2);
    }
    *err = std::fabs((rk - rg) * h);
    return rk * h;
}

static double adapt(double a, double b, const lt_cosmo_param& C, double whole, double err, int depth)
{
    if (err <= 1e-13 * std::fabs(whole) || depth > 40) return whole;
    const double m = 0.5 * (a + b);
    double e1, e2;
    const double l = gk15(a, m, C, &e1), r = gk15(m, b, C, &e2);
    return adapt(a, m, C, l, e1, depth + 1) + adapt(m, b, C, r, e2, depth + 1);
}

static double integral(double a, double b, const lt_cosmo_param& C)
{
    double e;
    const double w = gk15(a, b, C, &e);
    return adapt(a, b, C, w, e, 0);
}

static double sk(double x, double k)
{
    if (k > 0) return std::sin(x);
    if (k < 0) return std::sinh(x);
    return x;
}

double cosmo_observer_object(double z, const lt_cosmo_param& C)
{
    if (C.omegaX == 0.) {
        const double g = std::sqrt(1. + C.omegaM * z);
        return 2. * ((1. - C.omegaM - g) * (1. - g)) / C.omegaM / C.omegaM / (1. + z) / (1. + z);
    }
    const double chi = integral(0.0, z, C);
    if (C.curvature != 0.)
        return sk(chi * std::sqrt(std::fabs(C.curvature)), C.curvature) / (1 + z) / std::sqrt(std::fabs(C.curvature));
    return chi / (1 + z);
}

double cosmo_object_object(double z1, double z2, const lt_cosmo_param& C)
{
    if (z1 >= z2) return 0.;
    if (C.omegaX == 0.) {
        const double g1 = std::sqrt(1. + C.omegaM * z1), g2 = std::sqrt(1. + C.omegaM * z2);
        return 2. * ((1. - C.omegaM - g1 * g2) * (g1 - g2)) / C.omegaM / C.omegaM / (1. + z1) / (1. + z2) /
               (1. + z2);
    }
    const double chi = integral(z1, z2, C);
    if (C.curvature != 0.)
        return sk(chi * std::sqrt(std::fabs(C.curvature)), C.curvature) / (1 + z2) / std::sqrt(std::fabs(C.curvature));
    return chi / (1 + z2);
}

}  // namespace lt
