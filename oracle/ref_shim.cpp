// ref_shim.cpp -- C entry points into the reference's own CPU functions, built by oracle/Makefile
// from the reference sources where they lie (gradient2.cpp, grid_gradient2_CPU.cpp, delense_CPU.cpp,
// chi_computation.cpp). TEST INFRASTRUCTURE ONLY: it is used to write golden vectors and to
// validate the oracle restatement; nothing in the product library links it.
//
// gradient.cpp (the reference gradient) is not compiled into ref_shim: it does not compile unmodified
// (a `switch` jumps past an initialised variable in the unused AOS grad_halo, gradient.cpp:152-157).
#include <cstring>
#include <string>
#include <omp.h>

#include "structure_hpc.hpp"
#include "gradient2.hpp"
#include "grid_gradient2_CPU.hpp"
#include "delense_CPU.hpp"
#include "chi_computation.hpp"

#include "lenstool_hip/lt_types.h"

static_assert(sizeof(lt_potential_soa) == sizeof(Potential_SOA), "SOA layout");
static_assert(sizeof(lt_grid_param) == sizeof(grid_param), "grid layout");
static_assert(sizeof(lt_galaxy) == sizeof(galaxy), "galaxy layout");
static_assert(sizeof(lt_matrix) == sizeof(matrix), "matrix layout");

extern "C" {

void ref_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

// gradient2_grid_CPU, (dx, dy, nx, ny) overload (src/grid_gradient2_CPU.cpp:49-52)
int ref_gradient2_grid(lt_matrix* out, const lt_grid_param* frame, const lt_potential_soa* lens, int nhalos,
                       double dx, double dy, int nx, int ny, int istart, int jstart)
{
    gradient2_grid_CPU(reinterpret_cast<matrix*>(out), reinterpret_cast<const grid_param*>(frame),
                       reinterpret_cast<const Potential_SOA*>(lens), nhalos, dx, dy, nx, ny, istart, jstart);
    return 0;
}

// module_potentialDerivatives_totalGradient2_SOA (src/gradient2.cpp:226)
int ref_point_gradient2(lt_matrix* out, const lt_point* p, const lt_potential_soa* lens, int nhalos)
{
    matrix m = module_potentialDerivatives_totalGradient2_SOA(reinterpret_cast<const point*>(p),
                                                              reinterpret_cast<const Potential_SOA*>(lens), nhalos);
    std::memcpy(out, &m, sizeof(m));
    return 0;
}

// delense_barycenter (src/delense_CPU.cpp:6)
int ref_delense_barycenter(lt_point* image_pos, int* locimagesfound, int* numimg, int nsets, int nbgridcells,
                           const lt_grid_param* frame, const lt_galaxy* sources, double* gx, double* gy)
{
    runmode_param rm;
    rm.nsets = nsets;
    rm.nbgridcells = nbgridcells;
    int* nimages = new int[nsets];
    for (int i = 0; i < nsets; ++i) nimages[i] = 0;
    delense_barycenter(reinterpret_cast<point*>(image_pos), locimagesfound, numimg, &rm, nullptr,
                       reinterpret_cast<const grid_param*>(frame), nimages,
                       reinterpret_cast<const galaxy*>(sources), gx, gy);
    delete[] nimages;
    return 0;
}

// chi_computation (src/chi_computation.cpp:10)
int ref_chi_computation(double* chi, int* images_found, int* numimagesfound, lt_point* imagesposition,
                        const int* nimages, const lt_galaxy* images, int nsets)
{
    chi_computation(chi, images_found, numimagesfound, reinterpret_cast<point*>(imagesposition), nimages,
                    reinterpret_cast<const galaxy*>(images), nsets);
    return 0;
}

}  // extern "C"
