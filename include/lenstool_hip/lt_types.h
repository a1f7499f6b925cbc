/*
 * lt_types.h -- plain-C restatement of the Lenstool-HPC data model used on the lens-map path.
 *
 * Field order and field types are identical to the reference's double-precision build
 * (type_t = double, `-D_double`, src/type.h:4-8) so that a caller's own structures can be passed
 * through the C-ABI unchanged:
 *   lt_point          <-> struct point          (src/structure_hpc.hpp:84-88)
 *   lt_matrix         <-> struct matrix         (src/structure_hpc.hpp:139-145)
 *   lt_ellipse        <-> struct ellipse        (src/structure_hpc.hpp:152-157)
 *   lt_triplet        <-> struct triplet        (src/structure_hpc.hpp:110-115)
 *   lt_galaxy         <-> struct galaxy         (src/structure_hpc.hpp:172-182)
 *   lt_grid_param     <-> struct grid_param     (src/structure_hpc.hpp:445-454)
 *   lt_cosmo_param    <-> struct cosmo_param    (src/structure_hpc.hpp:468-478)
 *   lt_potential_soa  <-> struct Potential_SOA  (src/structure_hpc.hpp:257-287)
 * The names carry an `lt_` prefix so this header can be included next to the reference's own
 * structure header; include/lenstool_hip/structure_hpc.hpp provides the reference spellings.
 */
#ifndef LENSTOOL_HIP_LT_TYPES_H
#define LENSTOOL_HIP_LT_TYPES_H

#ifdef __cplusplus
extern "C" {
#endif

/* Maximum number of delensed images kept per source (src/structure_hpc.hpp:41). */
#define LT_MAXIMPERSOURCE 20

/* Physical constants of the reference (src/structure_hpc.hpp:62-68). */
#define LT_PI_C2    7.209970e-06   /* pi in arcsec / c^2                          */
#define LT_CH2PIG   0.11585881     /* c H0 / 2 pi G in g/cm^2 (H0 = 50)           */
#define LT_CH4PIG   0.057929405    /* c H0 / 4 pi G in g/cm^2 (H0 = 50)           */
#define LT_CH0_4PIG 2.7730112e-4   /* c H0 / 4 pi G in 1e12 Msun/kpc^2 (H0 = 50)  */
#define LT_D0       29.068701      /* c/H0 in h^-1 kpc/arcsec (h0 = 50)          */
#define LT_MCRIT12  .2343165       /* c^3/4piGh0/RTA/RTA in 1e12 Msun/arcsec^2   */

typedef struct lt_point   { double x, y; } lt_point;
typedef struct lt_matrix  { double a, b, c, d; } lt_matrix;
typedef struct lt_ellipse { double a, b, theta; } lt_ellipse;
typedef struct lt_triplet { lt_point a, b, c; } lt_triplet;

typedef struct lt_galaxy {
    lt_point   center;
    lt_ellipse shape;
    double     mag;
    double     redshift;
    double     dls;
    double     dos;
    double     dr;
} lt_galaxy;

typedef struct lt_grid_param {
    double xmin, xmax, ymin, ymax, lmin, lmax, rmax;
} lt_grid_param;

typedef struct lt_cosmo_param {
    int    model;
    double omegaM, omegaX, curvature, wX, wa, H0, h;
} lt_cosmo_param;

/* Structure-of-arrays lens model. Only the fields marked [used] are read by this library; the
 * others exist so the layout matches the reference and may be NULL. Halos must be grouped in runs
 * of equal `type` (the reference slots them by ascending type); the run-length scan never reads
 * type[nhalos]. */
typedef struct lt_potential_soa {
    int*    type;                   /* [used] 5 = SIS/SIE, 8 = PIEMD, 81 = dPIE, 14 = ext. shear */
    char    type_name[10];
    double* name;
    double* position_x;             /* [used] */
    double* position_y;             /* [used] */
    double* weight;
    double* b0;                     /* [used] */
    double* vdisp;
    double* ellipticity_angle;      /* [used] by the potential only via anglecos/anglesin */
    double* ellipticity;
    double* ellipticity_potential;  /* [used] */
    double* rcore;                  /* [used] */
    double* rcut;                   /* [used] */
    double* rscale;
    double* exponent;
    double* alpha;
    double* einasto_kappacritic;
    double* z;                      /* [used] by the map entry points: z[0] is the lens plane */
    double* mag;
    double* lum;
    double* theta;
    double* anglecos;               /* [used] */
    double* anglesin;               /* [used] */
    double* sigma;
    double* dlsds;                  /* [used] by the Kmap (mass-map) Hessian */
    int*    SOA_index;
    int*    N_types;
} lt_potential_soa;

#ifdef __cplusplus
}
#endif

#endif /* LENSTOOL_HIP_LT_TYPES_H */
