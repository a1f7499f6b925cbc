// lenstool_gpu.hpp -- the reference's GPU call signatures (C++ linkage, default arguments and
// overloads as in the reference headers), implemented in liblenstool_hip.so on top of the C-ABI
// (lt_capi.h). A program written against the reference's grid_gradient_GPU.cuh,
// grid_gradient2_GPU.cuh, grid_potential_GPU.cuh, grid_map_{ampli,shear,mass,pot,dpl}_GPU.cuh,
// delense_GPU.cuh and chi_barycenter_GPU.cuh includes this header instead and links this library.
//
// Semantics kept: host arrays in and out, synchronous, same argument meaning and pixel convention.
// Differences: device buffers are cached between calls instead of allocated and freed every time;
// an unsupported halo type or a GPU error does not exit the process (the reference's cudasafe
// calls exit(-1), src/cudafunctions.cu:33-39): the message goes to stderr and
// lenstool_gpu_last_status() returns the LT_* code. Set LT_EXIT_ON_ERROR=1 to exit(-1) instead.
#pragma once

#include "structure_hpc.hpp"

int lenstool_gpu_last_status();

// ---- src/grid_gradient_GPU.cuh:34-38 ------------------------------------------------------------
void gradient_grid_GPU(type_t* grid_grad_x, type_t* grid_grad_y, const struct grid_param* frame,
                       const struct Potential_SOA* lens, int Nlens, int nbgridcells);
void gradient_grid_GPU(type_t* grid_grad_x, type_t* grid_grad_y, const struct grid_param* frame,
                       const struct Potential_SOA* lens, int nhalos, type_t dx, type_t dy, int nbgridcells_x,
                       int nbgridcells_y, int istart = 0, int jstart = 0);
void gradient_grid_GPU_UM(type_t* grid_grad_x, type_t* grid_grad_y, const struct grid_param* frame,
                          const struct Potential_SOA* lens, int nhalos, type_t dx, type_t dy, int nbgridcells_x,
                          int nbgridcells_y, int istart, int jstart);

// ---- src/grid_gradient2_GPU.cuh:32-34 -----------------------------------------------------------
void gradient2_grid_GPU(type_t* grid_grad2_a, type_t* grid_grad2_b, type_t* grid_grad2_c, type_t* grid_grad2_d,
                        const struct grid_param* frame, const struct Potential_SOA* lens, int Nlens,
                        int nbgridcells);
void gradient2_grid_GPU(type_t* grid_grad2_a, type_t* grid_grad2_b, type_t* grid_grad2_c, type_t* grid_grad2_d,
                        const struct grid_param* frame, const struct Potential_SOA* lens, int nhalos, type_t dx,
                        type_t dy, int nbgridcells_x, int nbgridcells_y, int istart = 0, int jstart = 0);

// ---- the launchers on device data (src/grid_gradient_GPU.cuh:40-41, src/grid_gradient2_GPU.cuh:36-43) --
// Output planes, frame and the Potential_SOA descriptor are device pointers; the descriptor's arrays
// are device arrays (the subset type, position_x/y, b0, ellipticity_angle, ellipticity_potential,
// rcore, rcut, anglecos/sin, and dlsds for the Kmap). Synchronous.
void module_potentialDerivatives_totalGradient_SOA_CPU_GPU(type_t* grid_grad_x, type_t* grid_grad_y,
                                                           const struct grid_param* frame,
                                                           const struct Potential_SOA* lens_gpu, int nbgridcells,
                                                           int nhalos);
void module_potentialDerivatives_totalGradient_SOA_CPU_GPU(type_t* grid_grad_x, type_t* grid_grad_y,
                                                           const struct grid_param* frame,
                                                           const struct Potential_SOA* lens, int nhalos, type_t dx,
                                                           type_t dy, int nbgridcells_x, int nbgridcells_y,
                                                           int istart, int jstart);
void module_potentialDerivatives_totalGradient2_SOA_CPU_GPU(type_t* grid_grad2_a, type_t* grid_grad2_b,
                                                            type_t* grid_grad2_c, type_t* grid_grad2_d,
                                                            const struct grid_param* frame,
                                                            const struct Potential_SOA* lens_gpu, int nbgridcells,
                                                            int nhalos);
void module_potentialDerivatives_totalGradient2_SOA_CPU_GPU(type_t* grid_grad2_a, type_t* grid_grad_b,
                                                            type_t* grid_grad2_c, type_t* grid_grad_d,
                                                            const struct grid_param* frame,
                                                            const struct Potential_SOA* lens, int nhalos, type_t dx,
                                                            type_t dy, int nbgridcells_x, int nbgridcells_y,
                                                            int istart, int jstart);
void module_potentialDerivatives_Kmap_SOA_CPU_GPU(type_t* grid_grad2_a, type_t* grid_grad2_b, type_t* grid_grad2_c,
                                                  type_t* grid_grad2_d, const struct grid_param* frame,
                                                  const struct Potential_SOA* lens_gpu, int nhalos, type_t dx,
                                                  type_t dy, int nbgridcells_x, int nbgridcells_y, int istart,
                                                  int jstart);

// ---- src/grid_potential_GPU.cuh:31-32 -----------------------------------------------------------
void potential_grid_GPU(type_t* grid_potential, const struct grid_param* frame, const struct Potential_SOA* lens,
                        int nhalos, type_t dx, type_t dy, int nbgridcells_x, int nbgridcells_y, int istart,
                        int jstart);
void potential_grid_GPU(type_t* grid_potential, const struct grid_param* frame, const struct Potential_SOA* lens,
                        int nhalos, int nbgridcells);

// ---- map functions: device planes in, device map out (src/grid_map_ampli_GPU.cuh:35-47 & co) ----
typedef void (*map_gpu_function_t)(type_t* map, type_t* grid_grad2_a, type_t* grid_grad2_b, type_t* grid_grad2_c,
                                   type_t* grid_grad2_d, type_t dl0s, type_t dos, type_t dl, type_t h, type_t z,
                                   int nbgridcells_x, int nbgridcells_y, const struct grid_param* frame);
typedef void (*map_pot_function_t)(type_t* map, type_t dl0s, type_t dos, type_t dl, type_t h, type_t z,
                                   int nbgridcells_x, int nbgridcells_y, const struct grid_param* frame);

map_gpu_function_t select_map_ampli_function(const struct runmode_param* runmode);
map_gpu_function_t select_map_shear_function(const struct runmode_param* runmode);
map_gpu_function_t select_map_mass_function(const struct runmode_param* runmode);
map_pot_function_t select_map_potential_function(const struct runmode_param* runmode);
// src/grid_map_dpl_GPU.cuh:39 (mode 1: dpl_1_grid_CPU_GPU, the in-place D_ls/D_s scaling)
map_gpu_function_t select_map_dpl_function(const struct runmode_param* runmode);

void amplif_1_grid_CPU_GPU(type_t*, type_t*, type_t*, type_t*, type_t*, type_t, type_t, type_t, type_t, type_t, int,
                           int, const struct grid_param*);
void amplif_2_grid_CPU_GPU(type_t*, type_t*, type_t*, type_t*, type_t*, type_t, type_t, type_t, type_t, type_t, int,
                           int, const struct grid_param*);
void amplif_3_grid_CPU_GPU(type_t*, type_t*, type_t*, type_t*, type_t*, type_t, type_t, type_t, type_t, type_t, int,
                           int, const struct grid_param*);
void amplif_4_grid_CPU_GPU(type_t*, type_t*, type_t*, type_t*, type_t*, type_t, type_t, type_t, type_t, type_t, int,
                           int, const struct grid_param*);
void amplif_5_grid_CPU_GPU(type_t*, type_t*, type_t*, type_t*, type_t*, type_t, type_t, type_t, type_t, type_t, int,
                           int, const struct grid_param*);
void amplif_6_grid_CPU_GPU(type_t*, type_t*, type_t*, type_t*, type_t*, type_t, type_t, type_t, type_t, type_t, int,
                           int, const struct grid_param*);
void shear_1_grid_CPU_GPU(type_t*, type_t*, type_t*, type_t*, type_t*, type_t, type_t, type_t, type_t, type_t, int,
                          int, const struct grid_param*);
void shear_2_grid_CPU_GPU(type_t*, type_t*, type_t*, type_t*, type_t*, type_t, type_t, type_t, type_t, type_t, int,
                          int, const struct grid_param*);
void shear_3_grid_CPU_GPU(type_t*, type_t*, type_t*, type_t*, type_t*, type_t, type_t, type_t, type_t, type_t, int,
                          int, const struct grid_param*);
void shear_4_grid_CPU_GPU(type_t*, type_t*, type_t*, type_t*, type_t*, type_t, type_t, type_t, type_t, type_t, int,
                          int, const struct grid_param*);
void mass_1_grid_CPU_GPU(type_t*, type_t*, type_t*, type_t*, type_t*, type_t, type_t, type_t, type_t, type_t, int,
                         int, const struct grid_param*);
void mass_2_grid_CPU_GPU(type_t*, type_t*, type_t*, type_t*, type_t*, type_t, type_t, type_t, type_t, type_t, int,
                         int, const struct grid_param*);
void mass_3_grid_CPU_GPU(type_t*, type_t*, type_t*, type_t*, type_t*, type_t, type_t, type_t, type_t, type_t, int,
                         int, const struct grid_param*);
void mass_4_grid_CPU_GPU(type_t*, type_t*, type_t*, type_t*, type_t*, type_t, type_t, type_t, type_t, type_t, int,
                         int, const struct grid_param*);
void potential_1_grid_CPU_GPU(type_t*, type_t, type_t, type_t, type_t, type_t, int, int, const struct grid_param*);
void potential_2_grid_CPU_GPU(type_t*, type_t, type_t, type_t, type_t, type_t, int, int, const struct grid_param*);
void potential_3_grid_CPU_GPU(type_t*, type_t, type_t, type_t, type_t, type_t, int, int, const struct grid_param*);
void dpl_1_grid_CPU_GPU(type_t*, type_t*, type_t*, type_t*, type_t*, type_t, type_t, type_t, type_t, type_t, int, int,
                        const struct grid_param*);
void dpl_grid_CPU_GPU(type_t* map, type_t dl0s, type_t ds, type_t dl, type_t h, type_t z, int nbgridcells_x,
                      int nbgridcells_y, const struct grid_param* frame);

// ---- whole maps, host in / host out (src/grid_map_*_GPU.cuh) -------------------------------------
void map_grid_ampli_GPU(map_gpu_function_t mapfunction, type_t* amplif, const struct cosmo_param* cosmo,
                        const struct grid_param* frame, const struct Potential_SOA* lens, int Nlens, int nbgridcells,
                        int mode_amp, type_t z);
void map_grid_ampli_GPU(map_gpu_function_t mapfunction, type_t* amplif, const struct cosmo_param* cosmo,
                        const struct grid_param* frame, const struct Potential_SOA* lens, int Nlens, int mode_amp,
                        type_t z, type_t dx, type_t dy, int nbgridcells_x, int nbgridcells_y, int istart = 0,
                        int jstart = 0);
void map_grid_shear_GPU(map_gpu_function_t mapfunction, type_t* amplif, const struct cosmo_param* cosmo,
                        const struct grid_param* frame, const struct Potential_SOA* lens, int Nlens, int nbgridcells,
                        int mode_amp, type_t z);
void map_grid_shear_GPU(map_gpu_function_t mapfunction, type_t* amplif, const struct cosmo_param* cosmo,
                        const struct grid_param* frame, const struct Potential_SOA* lens, int Nlens, int mode_amp,
                        type_t z, type_t dx, type_t dy, int nbgridcells_x, int nbgridcells_y, int istart = 0,
                        int jstart = 0);
void map_grid_mass_GPU(map_gpu_function_t mapfunction, type_t* map, const struct cosmo_param* cosmo,
                       const struct grid_param* frame, const struct Potential_SOA* lens, int nhalos, int nbgridcells,
                       int mode_amp, type_t zm, type_t zs);
void map_grid_mass_GPU(map_gpu_function_t mapfunction, type_t* map, const struct cosmo_param* cosmo,
                       const struct grid_param* frame, const struct Potential_SOA* lens, int nhalos, int mode_amp,
                       type_t zm, type_t zs, type_t dx, type_t dy, int nbgridcells_x, int nbgridcells_y, int istart,
                       int jstart);
void map_grid_potential_GPU(map_pot_function_t mapfunction, type_t* amplif, const struct cosmo_param* cosmo,
                            const struct grid_param* frame, const struct Potential_SOA* lens, int Nlens,
                            int nbgridcells, int mode_amp, type_t z);
void map_grid_potential_GPU(map_pot_function_t mapfunction, type_t* amplif, const struct cosmo_param* cosmo,
                            const struct grid_param* frame, const struct Potential_SOA* lens, int Nlens, int mode_amp,
                            type_t z, type_t dx, type_t dy, int nbgridcells_x, int nbgridcells_y, int istart = 0,
                            int jstart = 0);
void map_grid_dpl_GPU(map_gpu_function_t mapfunction, type_t* dplx, type_t* dply, const struct cosmo_param* cosmo,
                      const struct grid_param* frame, const struct Potential_SOA* lens, int Nlens, int nbgridcells,
                      int mode_amp, type_t z);
void map_grid_dpl_GPU(map_gpu_function_t mapfunction, type_t* dplx, type_t* dply, const struct cosmo_param* cosmo,
                      const struct grid_param* frame, const struct Potential_SOA* lens, int Nlens, int mode_amp,
                      type_t z, type_t dx, type_t dy, int nbgridcells_x, int nbgridcells_y, int istart = 0,
                      int jstart = 0);

// ---- delensing and chi^2 (src/delense_GPU.cuh:27, src/chi_barycenter_GPU.cuh:13) ----------------
void delense_barycenter_GPU(struct point* image_pos, int* locimagesfound, int* numimg, const runmode_param* runmode,
                            const struct Potential_SOA* lens, const struct grid_param* frame,
                            const int* nimages_strongLensing, const struct galaxy* sources, double* grid_gradient_x,
                            double* grid_gradient_y);
void mychi_bruteforce_SOA_GPU_grid_gradient_barycentersource(double* chi, int* error, runmode_param* runmode,
                                                             const struct Potential_SOA* lens,
                                                             const struct grid_param* frame,
                                                             const int* nimages_strongLensing, galaxy* images);

// ---- the reference's CPU entry points of the same path, served by the GPU ------------------------
// Same declarations as src/grid_gradient_CPU.hpp:38-39, src/grid_gradient2_CPU.hpp:42-44,
// src/grid_amplif_CPU.hpp:40, src/gradient.hpp:43, src/gradient2.hpp:29, src/delense_CPU.hpp:20,
// src/chi_computation.hpp:3-4, src/chi_CPU_barycenter.hpp:13, so that callers of the CPU path
// (e.g. src/chi_CPU_barycenter.cpp, utils/exec/main.cpp) link this library instead. Reference
// behaviour kept: gradient_grid_CPU's square overload ignores istart/jstart
// (src/grid_gradient_CPU.cpp:38-41); amplif_grid_CPU mode 5 is kappa = (a + c) / 2 without dlsds
// (src/grid_amplif_CPU.cpp:50-83), its other modes (absent on the reference CPU) are the GPU maps'.
void gradient_grid_CPU(type_t* grid_grad_x, type_t* grid_grad_y, const struct grid_param* frame,
                       const struct Potential_SOA* lens, int Nlens, int nbgridcells, int istart = 0, int jstart = 0);
void gradient_grid_CPU(type_t* grid_grad_x, type_t* grid_grad_y, const struct grid_param* frame,
                       const struct Potential_SOA* lens, int Nlens, type_t dx, type_t dy, int nbgridcells_x,
                       int nbgridcells_y, int istart = 0, int jstart = 0);
void gradient2_grid_CPU(struct matrix* grid_grad2, const struct grid_param* frame, const struct Potential_SOA* lens,
                        int Nlens, int nbgridcells, int istart = 0, int jstart = 0);
void gradient2_grid_CPU(struct matrix* grid_grad2, const struct grid_param* frame, const struct Potential_SOA* lens,
                        int Nlens, int nbgridcells_x, int nbgridcells_y, int istart = 0, int jstart = 0);
void gradient2_grid_CPU(struct matrix* grid_grad2, const struct grid_param* frame, const struct Potential_SOA* lens,
                        int Nlens, type_t dx, type_t dy, int nbgridcells_x, int nbgridcells_y, int istart = 0,
                        int jstart = 0);
void amplif_grid_CPU(type_t* amplif, const struct cosmo_param* cosmo, const struct grid_param* frame,
                     const struct Potential_SOA* lens, int nhalos, int nbgridcells, int mode_amp, type_t z,
                     int istart = 0, int jstart = 0);
struct point module_potentialDerivatives_totalGradient_SOA(const struct point* pImage, const struct Potential_SOA* lens,
                                                           int nhalos);
struct matrix module_potentialDerivatives_totalGradient2_SOA(const struct point* pImage,
                                                             const struct Potential_SOA* lens, int nhalos);
void delense_barycenter(struct point* image_pos, int* locimagesfound, int* numimg, const runmode_param* runmode,
                        const struct Potential_SOA* lens, const struct grid_param* frame,
                        const int* nimages_strongLensing, const struct galaxy* sources, double* grid_gradient_x,
                        double* grid_gradient_y);
void chi_computation(double* chi, int* images_found, int* numimagesfound, struct point* imagesposition,
                     const int* nimages_strongLensing, const galaxy* images, const int nsets);
void mychi_bruteforce_SOA_CPU_grid_gradient_barycentersource(double* chi, int* error, runmode_param* runmode,
                                                             const struct Potential_SOA* lens,
                                                             const struct grid_param* frame,
                                                             const int* nimages_strongLensing, galaxy* images);

// ---- FITS maps (src/module_writeFits.hpp): float32 images with the frame's world coordinates ----
int module_writeFits_Image(char* filename, double* ima, int nx, int ny, double xmin, double xmax, double ymin,
                           double ymax);
int module_writeFits_ImageAbsoluteCoordinates(char* filename, double* ima, int nx, int ny, double xmin, double xmax,
                                              double ymin, double ymax, double ra, double dec);
void module_writeFits(std::string path, std::string filename, int ii, type_t* map, const struct runmode_param* runmode,
                      int size, const struct grid_param* frame, type_t ra, type_t dec);
void module_writeFits(std::string path, std::string filename, type_t* map, const struct runmode_param* runmode,
                      int size, const struct grid_param* frame, type_t ra, type_t dec);
