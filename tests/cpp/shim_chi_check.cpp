// shim_chi_check.cpp -- drives the reference-signature entry points of the potential, the managed-memory
// gradient, the point gradient, delensing and chi^2 (lenstool_gpu.hpp) the way a reference caller does:
// reference structures (Potential_SOA, grid_param, runmode_param, galaxy, point) in host memory.
//   shim_chi_check <input.bin> <output.bin>
// input : int nh, n, nsets, nimg; double xmin, xmax, ymin, ymax; int type[nh];
//         double x, y, b0, angle, epot, rcore, rcut, z, dlsds [nh each]; int nimages[nsets];
//         galaxy images[nimg]; galaxy sources[nsets]; double gx[n*n], gy[n*n]; double px[16], py[16]
// output: doubles, in order: potential_grid_GPU (square overload) [n*n]; potential_grid_GPU (dx/dy
//         overload, istart=1, jstart=2) [n*n]; gradient_grid_GPU_UM (same) x, y [n*n each];
//         module_potentialDerivatives_totalGradient_SOA at the 16 points, x then y [16 each];
//         delense_barycenter_GPU on (sources, gx, gy): image_pos [nsets*20*2], locimagesfound [nsets],
//         numimg [1]; the same from delense_barycenter; chi_computation on the GPU-named delensing's
//         images: chi, images_found; mychi_bruteforce_SOA_GPU_grid_gradient_barycentersource: chi, error;
//         mychi_bruteforce_SOA_CPU_grid_gradient_barycentersource: chi, error; then the int worst status
//         of lenstool_gpu_last_status() over the calls.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "lenstool_hip/lenstool_gpu.hpp"

static const int kMaxPerSource = 20;  // MAXIMPERSOURCE of the reference

template <class T>
static bool rd(FILE* f, T* p, size_t n)
{
    return std::fread(p, sizeof(T), n, f) == n;
}

int main(int argc, char** argv)
{
    if (argc != 3) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 3;
    int nh, n, nsets, nimg;
    double xmin, xmax, ymin, ymax;
    if (!rd(f, &nh, 1) || !rd(f, &n, 1) || !rd(f, &nsets, 1) || !rd(f, &nimg, 1)) return 4;
    if (!rd(f, &xmin, 1) || !rd(f, &xmax, 1) || !rd(f, &ymin, 1) || !rd(f, &ymax, 1)) return 4;
    std::vector<int> type(nh), nimages(nsets);
    std::vector<double> arr[9];
    if (!rd(f, type.data(), nh)) return 4;
    for (auto& a : arr) {
        a.resize(nh);
        if (!rd(f, a.data(), nh)) return 4;
    }
    std::vector<galaxy> images(nimg), sources(nsets);
    const size_t N = (size_t)n * n;
    std::vector<double> gx(N), gy(N), px(16), py(16);
    if (!rd(f, nimages.data(), nsets) || !rd(f, images.data(), nimg) || !rd(f, sources.data(), nsets) ||
        !rd(f, gx.data(), N) || !rd(f, gy.data(), N) || !rd(f, px.data(), 16) || !rd(f, py.data(), 16))
        return 4;
    std::fclose(f);

    std::vector<double> cs(nh), sn(nh);
    for (int i = 0; i < nh; ++i) {
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
    runmode_param rm{};
    rm.nbgridcells = n;
    rm.nsets = nsets;
    rm.nhalos = nh;

    std::vector<double> out;
    auto put = [&](const std::vector<double>& v) { out.insert(out.end(), v.begin(), v.end()); };
    int worst = 0;
    auto st = [&]() {
        int s = lenstool_gpu_last_status();
        if (s) worst = s;
    };
    std::vector<double> p0(N), p1(N);

    potential_grid_GPU(p0.data(), &frame, &lens, nh, n);
    st();
    put(p0);
    const double dx = (xmax - xmin) / (n - 1), dy = (ymax - ymin) / (n - 1);
    potential_grid_GPU(p0.data(), &frame, &lens, nh, dx, dy, n, n, 1, 2);
    st();
    put(p0);
    gradient_grid_GPU_UM(p0.data(), p1.data(), &frame, &lens, nh, dx, dy, n, n, 1, 2);
    st();
    put(p0);
    put(p1);
    std::vector<double> qx(16), qy(16);
    for (int k = 0; k < 16; ++k) {
        point q{px[k], py[k]};
        point g = module_potentialDerivatives_totalGradient_SOA(&q, &lens, nh);
        st();
        qx[k] = g.x;
        qy[k] = g.y;
    }
    put(qx);
    put(qy);

    // delensing of the given sources on the given gradient planes, through both names
    std::vector<point> pos((size_t)nsets * kMaxPerSource);
    std::vector<int> loc(nsets);
    int numimg = 0;
    auto put_delense = [&]() {
        for (const point& p : pos) {
            out.push_back(p.x);
            out.push_back(p.y);
        }
        for (int v : loc) out.push_back(v);
        out.push_back(numimg);
    };
    delense_barycenter_GPU(pos.data(), loc.data(), &numimg, &rm, &lens, &frame, nimages.data(), sources.data(),
                           gx.data(), gy.data());
    st();
    put_delense();
    const std::vector<point> pos_gpu = pos;
    const std::vector<int> loc_gpu = loc;
    std::fill(pos.begin(), pos.end(), point{0., 0.});
    std::fill(loc.begin(), loc.end(), 0);
    numimg = 0;
    delense_barycenter(pos.data(), loc.data(), &numimg, &rm, &lens, &frame, nimages.data(), sources.data(), gx.data(),
                       gy.data());
    st();
    put_delense();

    double chi = 0;
    int found = 0;
    std::vector<point> pos_in = pos_gpu;
    std::vector<int> loc_in = loc_gpu;
    chi_computation(&chi, &found, loc_in.data(), pos_in.data(), nimages.data(), images.data(), nsets);
    st();
    out.push_back(chi);
    out.push_back(found);

    int err = 0;
    chi = 0;
    mychi_bruteforce_SOA_GPU_grid_gradient_barycentersource(&chi, &err, &rm, &lens, &frame, nimages.data(),
                                                            images.data());
    st();
    out.push_back(chi);
    out.push_back(err);
    err = 0;
    chi = 0;
    mychi_bruteforce_SOA_CPU_grid_gradient_barycentersource(&chi, &err, &rm, &lens, &frame, nimages.data(),
                                                            images.data());
    st();
    out.push_back(chi);
    out.push_back(err);

    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 5;
    std::fwrite(out.data(), 8, out.size(), o);
    std::fwrite(&worst, 4, 1, o);
    std::fclose(o);
    return 0;
}
