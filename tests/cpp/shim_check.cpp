// shim_check.cpp -- drives the reference-signature entry points (lenstool_gpu.hpp) the way a
// reference caller would: reference structures, host arrays, map-function selectors.
//   shim_check <input.bin> <output.bin>
// input : int n, nx, ny; double xmin, xmax, ymin, ymax, zs; int type[n];
//         double x, y, b0, angle, epot, rcore, rcut, z, dlsds [n each]
// output: doubles, in order: gx, gy (square overload), a, b, c, d (dx/dy overload, istart=1, jstart=2),
//         ampli mode 1, ampli mode 1 through a foreign map function, shear mode 3, mass mode 2,
//         potential mode 1, dplx, dply; the CPU-named entry points: gradient2_grid_CPU (AoS, dx/dy
//         overload, istart=1, jstart=2) a and c, amplif_grid_CPU mode 5, gradient_grid_CPU (square) x;
//         a plane whose first 16 values are module_potentialDerivatives_totalGradient2_SOA (a,b,c,d)
//         at the first 4 pixels of the diagonal; the launchers on device data (device planes, frame and
//         Potential_SOA descriptor): totalGradient x, y (dx/dy overload, istart=1, jstart=2),
//         totalGradient2 a, b, c, d (same), Kmap a, c, totalGradient x (square overload),
//         totalGradient2 a (square overload), dpl mode 1 from select_map_dpl_function applied to
//         the square gradient's x plane; then the int status of lenstool_gpu_last_status().
#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "lenstool_hip/lenstool_gpu.hpp"
#include "lenstool_hip/lt_capi.h"

// a host array copied to a new device array
template <class T>
static T* to_device(const std::vector<T>& v)
{
    T* d = nullptr;
    if (lt_dev_alloc((void**)&d, (long long)(v.size() * sizeof(T)))) std::exit(6);
    if (lt_memcpy_h2d(d, v.data(), (long long)(v.size() * sizeof(T)), nullptr)) std::exit(6);
    return d;
}

static void foreign_ampli1(type_t* map, type_t* a, type_t* b, type_t* c, type_t* d, type_t dl0s, type_t dos, type_t dl,
                           type_t h, type_t z, int nx, int ny, const grid_param* frame)
{
    amplif_1_grid_CPU_GPU(map, a, b, c, d, dl0s, dos, dl, h, z, nx, ny, frame);
}

int main(int argc, char** argv)
{
    if (argc != 3) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 3;
    int n, nx, ny;
    double xmin, xmax, ymin, ymax, zs;
    if (std::fread(&n, 4, 1, f) != 1 || std::fread(&nx, 4, 1, f) != 1 || std::fread(&ny, 4, 1, f) != 1) return 4;
    if (std::fread(&xmin, 8, 1, f) != 1 || std::fread(&xmax, 8, 1, f) != 1 || std::fread(&ymin, 8, 1, f) != 1 ||
        std::fread(&ymax, 8, 1, f) != 1 || std::fread(&zs, 8, 1, f) != 1)
        return 4;
    std::vector<int> type(n);
    std::vector<double> arr[9];
    if (std::fread(type.data(), 4, n, f) != (size_t)n) return 4;
    for (auto& a : arr) {
        a.resize(n);
        if (std::fread(a.data(), 8, n, f) != (size_t)n) return 4;
    }
    std::fclose(f);
    std::vector<double> cs(n), sn(n);
    for (int i = 0; i < n; ++i) {
        cs[i] = std::cos(arr[3][i]);
        sn[i] = std::sin(arr[3][i]);
    }
    Potential_SOA lens{};
    lens.type = type.data();
    lens.position_x = arr[0].data();
    lens.position_y = arr[1].data();
    lens.b0 = arr[2].data();
    lens.ellipticity_angle = arr[3].data();
    lens.ellipticity_potential = arr[4].data();
    lens.rcore = arr[5].data();
    lens.rcut = arr[6].data();
    lens.z = arr[7].data();
    lens.dlsds = arr[8].data();
    lens.anglecos = cs.data();
    lens.anglesin = sn.data();
    grid_param frame{xmin, xmax, ymin, ymax, 0, 0, 0};
    cosmo_param cosmo{1, 0.3, 0.7, 0.0, -1.0, 0.0, 70.0, 1.4};
    runmode_param rm{};
    const size_t N = (size_t)nx * ny;
    std::vector<double> out;
    std::vector<double> p0(N), p1(N), p2(N), p3(N);
    auto put = [&](const std::vector<double>& v) { out.insert(out.end(), v.begin(), v.end()); };
    int worst = 0;
    auto st = [&]() { int s = lenstool_gpu_last_status(); if (s) worst = s; };

    gradient_grid_GPU(p0.data(), p1.data(), &frame, &lens, n, nx);  // square overload (nx == ny)
    st();
    put(p0);
    put(p1);
    const double dx = (xmax - xmin) / (nx - 1), dy = (ymax - ymin) / (ny - 1);
    gradient2_grid_GPU(p0.data(), p1.data(), p2.data(), p3.data(), &frame, &lens, n, dx, dy, nx, ny, 1, 2);
    st();
    put(p0);
    put(p1);
    put(p2);
    put(p3);
    rm.amplif = 1;
    map_grid_ampli_GPU(select_map_ampli_function(&rm), p0.data(), &cosmo, &frame, &lens, n, nx, 1, zs);
    st();
    put(p0);
    map_grid_ampli_GPU(foreign_ampli1, p0.data(), &cosmo, &frame, &lens, n, nx, 1, zs);
    st();
    put(p0);
    rm.shear = 3;
    map_grid_shear_GPU(select_map_shear_function(&rm), p0.data(), &cosmo, &frame, &lens, n, nx, 3, zs);
    st();
    put(p0);
    rm.mass = 2;
    map_grid_mass_GPU(select_map_mass_function(&rm), p0.data(), &cosmo, &frame, &lens, n, nx, 2, 0.0, zs);
    st();
    put(p0);
    rm.potential = 1;
    map_grid_potential_GPU(select_map_potential_function(&rm), p0.data(), &cosmo, &frame, &lens, n, nx, 1, zs);
    st();
    put(p0);
    map_grid_dpl_GPU(nullptr, p0.data(), p1.data(), &cosmo, &frame, &lens, n, nx, 1, zs);
    st();
    put(p0);
    put(p1);
    std::vector<matrix> aos(N);
    gradient2_grid_CPU(aos.data(), &frame, &lens, n, dx, dy, nx, ny, 1, 2);
    st();
    for (size_t i = 0; i < N; ++i) {
        p0[i] = aos[i].a;
        p1[i] = aos[i].c;
    }
    put(p0);
    put(p1);
    amplif_grid_CPU(p0.data(), &cosmo, &frame, &lens, n, nx, 5, zs);
    st();
    put(p0);
    gradient_grid_CPU(p0.data(), p1.data(), &frame, &lens, n, nx);
    st();
    put(p0);
    std::fill(p0.begin(), p0.end(), 0.0);
    for (int k = 0; k < 4; ++k) {
        point q{xmin + k * dx, ymin + k * dy};
        matrix m = module_potentialDerivatives_totalGradient2_SOA(&q, &lens, n);
        st();
        p0[4 * k + 0] = m.a;
        p0[4 * k + 1] = m.b;
        p0[4 * k + 2] = m.c;
        p0[4 * k + 3] = m.d;
    }
    put(p0);

    // the reference's launchers on device data: a Potential_SOA in device memory whose arrays are
    // device arrays, a device frame and device planes (src/grid_gradient2_GPU.cu:101-157)
    Potential_SOA dl{};
    dl.type = to_device(type);
    dl.position_x = to_device(arr[0]);
    dl.position_y = to_device(arr[1]);
    dl.b0 = to_device(arr[2]);
    dl.ellipticity_angle = to_device(arr[3]);
    dl.ellipticity_potential = to_device(arr[4]);
    dl.rcore = to_device(arr[5]);
    dl.rcut = to_device(arr[6]);
    dl.dlsds = to_device(arr[8]);
    dl.anglecos = to_device(cs);
    dl.anglesin = to_device(sn);
    Potential_SOA* lens_gpu = to_device(std::vector<Potential_SOA>{dl});
    grid_param* frame_gpu = to_device(std::vector<grid_param>{frame});
    double* dp[4];
    for (auto& q : dp)
        if (lt_dev_alloc((void**)&q, (long long)(N * sizeof(double)))) return 6;
    auto get = [&](int k) {
        if (lt_memcpy_d2h(p0.data(), dp[k], (long long)(N * sizeof(double)), nullptr)) std::exit(6);
        put(p0);
    };
    module_potentialDerivatives_totalGradient_SOA_CPU_GPU(dp[0], dp[1], frame_gpu, lens_gpu, n, dx, dy, nx, ny, 1, 2);
    st();
    get(0);
    get(1);
    module_potentialDerivatives_totalGradient2_SOA_CPU_GPU(dp[0], dp[1], dp[2], dp[3], frame_gpu, lens_gpu, n, dx, dy,
                                                           nx, ny, 1, 2);
    st();
    for (int k = 0; k < 4; ++k) get(k);
    module_potentialDerivatives_Kmap_SOA_CPU_GPU(dp[0], dp[1], dp[2], dp[3], frame_gpu, lens_gpu, n, dx, dy, nx, ny, 0,
                                                 0);
    st();
    get(0);
    get(2);
    module_potentialDerivatives_totalGradient_SOA_CPU_GPU(dp[0], dp[1], frame_gpu, lens_gpu, nx, n);
    st();
    get(0);
    module_potentialDerivatives_totalGradient2_SOA_CPU_GPU(dp[0], dp[1], dp[2], dp[3], frame_gpu, lens_gpu, nx, n);
    st();
    get(0);
    module_potentialDerivatives_totalGradient_SOA_CPU_GPU(dp[0], dp[1], frame_gpu, lens_gpu, nx, n);
    st();
    rm.dpl = 1;
    map_gpu_function_t dpl = select_map_dpl_function(&rm);
    st();
    if (!dpl) return 7;
    const double zl = arr[7][0];
    dpl(dp[0], nullptr, nullptr, nullptr, nullptr, lt_cosmo_object_object(zl, zs, (const lt_cosmo_param*)&cosmo),
        lt_cosmo_observer_object(zs, (const lt_cosmo_param*)&cosmo), 0.0, cosmo.h, zs, nx, ny, &frame);
    st();
    get(0);

    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 5;
    std::fwrite(out.data(), 8, out.size(), o);
    std::fwrite(&worst, 4, 1, o);
    std::fclose(o);
    return 0;
}
