// Calls the reference's FITS writer signatures (module_writeFits ×2, module_writeFits_Image,
// module_writeFits_ImageAbsoluteCoordinates) as a C++ caller of the reference would, on a small
// ramp map: argv[1] = output directory. Host only (no GPU needed).
#include <cstdio>
#include <string>
#include <vector>

#include "lenstool_hip/lenstool_gpu.hpp"

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    const int n = 7;
    std::vector<double> map(n * n);
    for (int k = 0; k < n * n; ++k) map[k] = 0.25 * k - 3.0;
    runmode_param runmode{};
    grid_param frame{};
    frame.xmin = -30.;
    frame.xmax = 30.;
    frame.ymin = -20.;
    frame.ymax = 40.;
    module_writeFits(dir, "rel", 3, map.data(), &runmode, n, &frame, 0., 0.);         // <dir>/rel_3.fits
    module_writeFits(dir, "abs", map.data(), &runmode, n, &frame, 3.586259, -30.4);   // <dir>/abs.fits
    std::string p1 = dir + "/img.fits", p2 = dir + "/imgabs.fits";
    std::vector<char> b1(p1.begin(), p1.end()), b2(p2.begin(), p2.end());
    b1.push_back(0);
    b2.push_back(0);
    int s1 = module_writeFits_Image(b1.data(), map.data(), n, n, -30., 30., -20., 40.);
    int s2 = module_writeFits_ImageAbsoluteCoordinates(b2.data(), map.data(), n, n, -30., 30., -20., 40., 0., 0.);
    std::printf("%d %d %d\n", s1, s2, lenstool_gpu_last_status());
    return s1 || s2;
}
