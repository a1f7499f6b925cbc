// structure_hpc.hpp -- the reference's data model under its own names (C++), for the host shims that
// keep the reference's call signatures (lenstool_gpu.hpp). Field order and types follow the
// reference's double build (src/structure_hpc.hpp, src/type.h:4-8) so objects built by a reference
// caller are passed through unchanged; the C-ABI spells the same layouts with an lt_ prefix
// (lt_types.h), and static_asserts in the shim check that both agree.
#pragma once

#include <string>

#include "lt_types.h"

typedef double type_t;

#define MAXIMPERSOURCE LT_MAXIMPERSOURCE
#define N_MAX_POTFILE 10
#define NTYPES 2

struct point {
    type_t x;
    type_t y;
};

struct complex {
    type_t re;
    type_t im;
};

struct triplet {
    struct point a;
    struct point b;
    struct point c;
};

struct matrix {
    type_t a;
    type_t b;
    type_t c;
    type_t d;
};

struct ellipse {
    type_t a;
    type_t b;
    type_t theta;
};

struct galaxy {
    struct point center;
    struct ellipse shape;
    type_t mag;
    type_t redshift;
    type_t dls;
    type_t dos;
    type_t dr;
};

struct grid_param {
    type_t xmin;
    type_t xmax;
    type_t ymin;
    type_t ymax;
    type_t lmin;
    type_t lmax;
    type_t rmax;
};

struct cosmo_param {
    int model;
    type_t omegaM;
    type_t omegaX;
    type_t curvature;
    type_t wX;
    type_t wa;
    type_t H0;
    type_t h;
};

struct Potential_SOA {
    int* type;
    char type_name[10];
    type_t* name;
    type_t* position_x;
    type_t* position_y;
    type_t* weight;
    type_t* b0;
    type_t* vdisp;
    type_t* ellipticity_angle;
    type_t* ellipticity;
    type_t* ellipticity_potential;
    type_t* rcore;
    type_t* rcut;
    type_t* rscale;
    type_t* exponent;
    type_t* alpha;
    type_t* einasto_kappacritic;
    type_t* z;
    type_t* mag;
    type_t* lum;
    type_t* theta;
    type_t* anglecos;
    type_t* anglesin;
    type_t* sigma;
    type_t* dlsds;
    int* SOA_index;
    int* N_types;
};

// Run-mode block of a .par file (src/structure_hpc.hpp:354-424); the shims read nbgridcells, nsets,
// nhalos and the map modes (amplif, shear, mass, potential, dpl).
struct runmode_param {
    int nbgridcells;
    int source;
    int nsets;
    int image;
    int N_z_param;
    int nimagestot;
    int multi;
    type_t ref_ra;
    type_t ref_dec;
    int mass;
    int mass_gridcells;
    type_t z_mass;
    type_t z_mass_s;
    std::string mass_name;
    int potential;
    int pot_gridcells;
    type_t z_pot;
    std::string pot_name;
    int nhalos;
    int n_tot_halos;
    int potfile;
    int npotfile[N_MAX_POTFILE];
    int Nb_potfile;
    int dpl;
    int dpl_gridcells;
    type_t z_dpl;
    std::string dpl_name1;
    std::string dpl_name2;
    int inverse;
    int arclet;
    int debug;
    int grid;
    int gridcells;
    type_t zgrid;
    int cline;
    int amplif;
    int amplif_gridcells;
    type_t z_amplif;
    std::string amplif_name;
    int shear;
    int shear_gridcells;
    type_t z_shear;
    std::string shear_name;
    int time;
    int Nlens[NTYPES];
    std::string imagefile;
    std::string potfilename[N_MAX_POTFILE];
    std::string sourfile;
};
