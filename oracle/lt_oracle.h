/*
 * lt_oracle.h -- FP64 CPU restatement of the Lenstool-HPC lens-map path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker the HIP library is compared against: only the
 * test-suite, the smoke check and the CPU-baseline leg of the benchmark load it. The product
 * library never links, loads or calls it.
 *
 * Every function restates the reference CPU algorithm it names (file:line into the reference
 * sources) from the mathematics, in plain C99 + OpenMP over grid rows, in double precision.
 * Where the reference has no CPU version (Kmap, potential, type-14 Hessian, epilogue maps) the
 * restatement follows the reference GPU semantics it names; where the reference has no version at
 * all (SIS / PIEMD-8 Hessians, SIS / PIEMD-8 potentials) the closed form is derived here and is
 * pinned by finite differences of the gradient in the tests.
 *
 * Return codes follow the library: 0 = ok, LT_ERR_* (lt_capi.h) otherwise.
 */
#ifndef LT_ORACLE_H
#define LT_ORACLE_H

#include "lenstool_hip/lt_types.h"
#include "lenstool_hip/lt_status.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- point evaluators ---------------------------------------------------------------------- */
/* alpha(p) = sum over halos of grad phi            (src/gradient.cpp:899-930)                    */
int orc_point_gradient(lt_point* out, const lt_point* p, const lt_potential_soa* lens, int nhalos);
/* H(p) = sum over halos of grad grad phi           (src/gradient2.cpp:226-254, + types 5/8/14)   */
int orc_point_gradient2(lt_matrix* out, const lt_point* p, const lt_potential_soa* lens, int nhalos);
/* Kmap: each (type, dlsds) run weighted by dlsds  (src/gradient2_GPU.cu:453-522)               */
int orc_point_kmap(lt_matrix* out, const lt_point* p, const lt_potential_soa* lens, int nhalos);
/* phi(p) = sum over halos of the lensing potential (src/potential_GPU.cu:75-116, + types 5/8)   */
int orc_point_potential(double* out, const lt_point* p, const lt_potential_soa* lens, int nhalos);
/* Sum over halos of |per-halo term| for each Hessian component (a, b, c): the conditioning scale
 * the parity tests use as the denominator of the relative error. */
int orc_point_gradient2_abs(lt_matrix* out, const lt_point* p, const lt_potential_soa* lens, int nhalos);
int orc_point_gradient_abs(lt_point* out, const lt_point* p, const lt_potential_soa* lens, int nhalos);
/* The same sums evaluated in 80-bit extended precision (same formulas, same order): the tests use
 * them as the exact value next to which the FP64 rounding of a formula is measured. */
int orc_point_gradient_ld(lt_point* out, const lt_point* p, const lt_potential_soa* lens, int nhalos);
int orc_point_gradient2_ld(lt_matrix* out, const lt_point* p, const lt_potential_soa* lens, int nhalos);
int orc_point_potential_ld(double* out, const lt_point* p, const lt_potential_soa* lens, int nhalos);

/* ---- grid drivers (row-major, index = row*nx + col, pixel (col,row) at
 *      (xmin + (col+istart)*dx, ymin + (row+jstart)*dy))  src/grid_gradient_CPU.cpp:67-104 ---- */
int orc_gradient_grid(double* gx, double* gy, const lt_grid_param* frame, const lt_potential_soa* lens,
                      int nhalos, double dx, double dy, int nx, int ny, int istart, int jstart);
/* prec = 0: double, 1: 80-bit extended (results rounded to double) */
int orc_gradient_grid_p(double* gx, double* gy, const lt_grid_param* frame, const lt_potential_soa* lens,
                        int nhalos, double dx, double dy, int nx, int ny, int istart, int jstart, int prec);
/* AoS matrix output like src/grid_gradient2_CPU.cpp:56-83 */
int orc_gradient2_grid(lt_matrix* out, const lt_grid_param* frame, const lt_potential_soa* lens,
                       int nhalos, double dx, double dy, int nx, int ny, int istart, int jstart);
/* planar a,b,c,d output (any plane may be NULL), optional |term| sums for a,b,c (may be NULL) */
int orc_gradient2_planes(double* a, double* b, double* c, double* d,
                         double* abs_a, double* abs_b, double* abs_c,
                         const lt_grid_param* frame, const lt_potential_soa* lens, int nhalos,
                         double dx, double dy, int nx, int ny, int istart, int jstart, int kmap,
                         int prec);
int orc_gradient_abs_planes(double* abs_x, double* abs_y, const lt_grid_param* frame,
                            const lt_potential_soa* lens, int nhalos, double dx, double dy,
                            int nx, int ny, int istart, int jstart);
int orc_potential_grid(double* phi, const lt_grid_param* frame, const lt_potential_soa* lens, int nhalos,
                       double dx, double dy, int nx, int ny, int istart, int jstart);
int orc_potential_grid_p(double* phi, const lt_grid_param* frame, const lt_potential_soa* lens, int nhalos,
                         double dx, double dy, int nx, int ny, int istart, int jstart, int prec);
/* FP64 rounding-error scale of the closed forms (absolute, per pixel) for the gradient (cg) and the
 * Hessian (ch); large only next to the removable 0/0 points of the PIEMD terms. */
int orc_point_cond(double* cg, double* ch, const lt_point* p, const lt_potential_soa* lens, int nhalos);
int orc_cond_planes(double* cg, double* ch, const lt_grid_param* frame, const lt_potential_soa* lens, int nhalos,
                    double dx, double dy, int nx, int ny, int istart, int jstart);
/* kappa map, mode 5 only, no dlsds  (src/grid_amplif_CPU.cpp:38-84) */
int orc_amplif_grid(double* out, const lt_grid_param* frame, const lt_potential_soa* lens, int nhalos,
                    int nbgridcells, int mode_amp);

/* ---- map epilogues from Hessian planes (src/grid_map_*_GPU.cu kernels) ---------------------- */
/* kind: LT_MAP_AMPLI / SHEAR / MASS (modes as the reference runmode), LT_MAP_FIELDS for the
 * multi-plane request (out0..out4 = kappa, gamma1, gamma2, detA, mu, each may be NULL). */
int orc_map_epilogue(int kind, int mode, double* out0, double* out1, double* out2, double* out3,
                     double* out4, const double* a, const double* b, const double* c, long n,
                     double dl0s, double dos, double dol, double h, double dx, double dy);

/* ---- cosmology (src/module_cosmodistances.cpp:94-174, 217-338); flat/curved CPL model 1 ----- */
double orc_cosmo_observer_object(double z, const lt_cosmo_param* c);
double orc_cosmo_object_object(double z1, double z2, const lt_cosmo_param* c);
double orc_cosmo_dlsds(double zl, double zs, const lt_cosmo_param* c);

/* ---- delensing and chi^2 --------------------------------------------------------------------- */
/* src/delense_CPU.cpp:6-138 with one thread (row-major scan order). image_pos has
 * nsets*LT_MAXIMPERSOURCE slots. Returns LT_ERR_OVERFLOW (after clamping) when a source has more
 * than LT_MAXIMPERSOURCE images, which the reference does not guard against. */
int orc_delense_barycenter(lt_point* image_pos, int* locimagesfound, int* numimg, int nsets,
                           int nbgridcells, const lt_grid_param* frame, const lt_galaxy* sources,
                           const double* gx, const double* gy);
/* src/chi_CPU_barycenter.cpp:173-210: source barycentres from point gradients */
int orc_source_barycenters(lt_galaxy* sources, int nsets, const int* nimages, const lt_galaxy* images,
                           const lt_potential_soa* lens, int nhalos);
/* src/chi_computation.cpp:10-195 verbatim semantics (chi and images_found are accumulated) */
int orc_chi_computation(double* chi, int* images_found, const int* numimagesfound,
                        const lt_point* imagesposition, const int* nimages, const lt_galaxy* images,
                        int nsets);
/* src/chi_CPU_barycenter.cpp:58-287 on one rank: gradient grid, barycentres, delensing, chi^2 */
int orc_chi_barycentersource(double* chi, int* error, int nbgridcells, int nsets, int nhalos,
                             const lt_potential_soa* lens, const lt_grid_param* frame,
                             const int* nimages, const lt_galaxy* images);

/* ---- critical lines -------------------------------------------------------------------------- */
/* det A = 0 contour of a det-A plane (ny rows of `pitch` values, pixel (i, j) at
 * (xmin + (istart+i) dx, ymin + (jstart+j) dy)) by marching squares: the library's rules (no
 * reference counterpart; the reference has no critical-line code): per cell, corner bit k set when
 * c_k > 0 (c0 = (i,j), c1 = (i+1,j), c2 = (i+1,j+1), c3 = (i,j+1)); crossings at t = va/(va - vb)
 * from the first end of an edge (bottom/top from the left, left/right from below); segments
 * oriented with positive det on the left; saddles joined on the positive side when c0+c1+c2+c3 > 0;
 * row-major cell order. kinds (if kappa given): 0 when 1 - mean kappa of the ends > 0, else 1.
 * Writes min(total, max_seg) segments (x0, y0, x1, y1), *nseg = total. Single-threaded. */
int orc_critic_lines(double* segs, int* kinds, long max_seg, long* nseg, const double* det, const double* kappa,
                     int nx, int ny, long pitch, double xmin, double ymin, double dx, double dy, int istart,
                     int jstart);

/* number of OpenMP threads the grid drivers use (0 = library default) */
void orc_set_threads(int n);
int  orc_get_threads(void);

#ifdef __cplusplus
}
#endif

#endif
