// lt_fits.cpp -- FITS map writer: a primary image of 32-bit IEEE floats with the linear (arcsec)
// or tangent-plane (RA/DEC, deg) world coordinates of the map's frame, as the reference writes
// its maps with cfitsio (src/module_writeFits.cpp:113-335: module_writeFits_Image and
// module_writeFits_ImageAbsoluteCoordinates; the keyword set, their order and the value types,
// float for the relative frame and double for the absolute one, are the same, so tools that read
// the reference's maps read these). The file layout is the FITS standard's: 2880-byte records,
// 80-character header cards, big-endian data, the last record zero-padded. No cfitsio: the writer
// is the few rules the image needs.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "lenstool_hip/lt_capi.h"
#include "lt_internal.h"

using namespace lt;

namespace {

constexpr size_t kRecord = 2880;
constexpr size_t kCard = 80;

// a real value the way cfitsio formats keyword values: %.<digits>G, with a decimal point added when
// the result has none ("2000" -> "2000.", "1E-20" -> "1.E-20")
std::string real_value(double v, int digits)
{
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.*G", digits, v);
    std::string s = buf;
    if (s.find('.') == std::string::npos && s.find('N') == std::string::npos && s.find('I') == std::string::npos) {
        const size_t e = s.find('E');
        if (e == std::string::npos) s += ".";
        else s.insert(e, ".");
    }
    return s;
}

struct Header {
    std::string text;

    void card(const std::string& key, const std::string& value, const std::string& comment)
    {
        char buf[kCard + 64];
        std::snprintf(buf, sizeof buf, "%-8.8s= %20s / %s", key.c_str(), value.c_str(), comment.c_str());
        put(buf);
    }
    // a string value: quoted, padded to 8 characters inside the quotes, left-justified in 20
    void string_card(const std::string& key, const std::string& value, const std::string& comment)
    {
        char q[96];
        std::snprintf(q, sizeof q, "'%-8s'", value.c_str());
        char buf[kCard + 96];
        std::snprintf(buf, sizeof buf, "%-8.8s= %-20s / %s", key.c_str(), q, comment.c_str());
        put(buf);
    }
    void raw(const std::string& s) { put(s); }
    void put(std::string s)
    {
        s.resize(kCard, ' ');
        text += s;
    }
    void finish()
    {
        put("END");
        text.resize((text.size() + kRecord - 1) / kRecord * kRecord, ' ');
    }
};

std::string utc_now()
{
    const std::time_t t = std::time(nullptr);
    std::tm u{};
    gmtime_r(&t, &u);
    char buf[32];
    std::strftime(buf, sizeof buf, "%Y-%m-%dT%H:%M:%S", &u);
    return buf;
}

}  // namespace

extern "C" {

int lt_fits_write(const char* path, const double* map, int nx, int ny, double xmin, double xmax, double ymin,
                  double ymax, int absolute, double ra, double dec)
{
    if (!path || !map) return set_error(LT_ERR_INVALID, "NULL argument");
    if (nx < 2 || ny < 2) return set_error(LT_ERR_INVALID, "a FITS map needs at least 2 x 2 pixels");
    if (!(xmax != xmin) || !(ymax != ymin)) return set_error(LT_ERR_INVALID, "empty frame");

    Header h;
    h.card("SIMPLE", "T", "file does conform to FITS standard");
    h.card("BITPIX", "-32", "number of bits per data pixel");
    h.card("NAXIS", "2", "number of data axes");
    h.card("NAXIS1", std::to_string(nx), "length of data axis 1");
    h.card("NAXIS2", std::to_string(ny), "length of data axis 2");
    h.card("EXTEND", "T", "FITS dataset may contain extensions");
    h.raw("COMMENT   FITS (Flexible Image Transport System) format is defined in 'Astronomy");
    h.raw("COMMENT   and Astrophysics', volume 376, page 359; bibcode: 2001A&A...376..359H");
    if (!absolute) {
        // relative frame, float keywords: pixel 1 at (xmin, ymin), steps over n - 1 intervals
        auto f = [](double v) { return real_value((double)(float)v, 7); };
        h.card("BZERO", f(0.), "BZERO");
        h.card("BSCALE", f(1.), "BSCALE");
        h.card("CRVAL1", f(xmin), "CRVAL1");
        h.card("CRPIX1", f(1.), "CRPIX1");
        h.card("CDELT1", f((xmax - xmin) / (nx - 1)), "CDELT1");
        h.string_card("CUNIT1", "arcsec", "CUNIT1");
        h.string_card("CTYPE1", "LINEAR", "CTYPE1");
        h.card("CRVAL2", f(ymin), "CRVAL2");
        h.card("CRPIX2", f(1.), "CRPIX2");
        h.card("CDELT2", f((ymax - ymin) / (ny - 1)), "CDELT2");
        h.string_card("CUNIT2", "arcsec", "CUNIT2");
        h.string_card("CTYPE2", "LINEAR", "CTYPE2");
    } else {
        // absolute frame, double keywords: the reference point (ra, dec) at the pixel of relative
        // (0, 0); RA grows to the east, so its step is negative
        auto d = [](double v) { return real_value(v, 15); };
        h.card("BZERO", d(0.), "BZERO");
        h.card("BSCALE", d(1.), "BSCALE");
        h.card("EQUINOX", d(2000.), "EQUINOX");
        h.string_card("RADECSYS", "FK5", "RADECSYS");
        h.string_card("CTYPE1", "RA---TAN", "CTYPE1");
        h.card("CRVAL1", d(ra), "CRVAL1");
        h.card("CRPIX1", d(-(nx - 1) * xmin / (xmax - xmin) + 1), "CRPIX1");
        h.card("CDELT1", d(-(xmax - xmin) / (nx - 1) / 3600), "CDELT1");
        h.string_card("CUNIT1", "deg", "CUNIT1");
        h.string_card("CTYPE2", "DEC--TAN", "CTYPE2");
        h.card("CRVAL2", d(dec), "CRVAL2");
        h.card("CRPIX2", d(-(ny - 1) * ymin / (ymax - ymin) + 1), "CRPIX2");
        h.card("CDELT2", d((ymax - ymin) / (ny - 1) / 3600), "CDELT2");
        h.string_card("CUNIT2", "deg", "CUNIT2");
    }
    h.string_card("ORIGIN", "JPK-SOFT", "ORIGIN");
    {
        char buf[kCard + 1];
        std::snprintf(buf, sizeof buf, "DATE    = '%s' / file creation date (YYYY-MM-DDThh:mm:ss UT)", utc_now().c_str());
        h.raw(buf);
    }
    h.finish();

    // data: rows of nx values, float32 big-endian, zero-padded to whole records
    const size_t n = (size_t)nx * (size_t)ny;
    std::vector<unsigned char> data((n * 4 + kRecord - 1) / kRecord * kRecord, 0);
    for (size_t k = 0; k < n; ++k) {
        const float v = (float)map[k];
        uint32_t u;
        std::memcpy(&u, &v, 4);
        data[4 * k + 0] = (unsigned char)(u >> 24);
        data[4 * k + 1] = (unsigned char)(u >> 16);
        data[4 * k + 2] = (unsigned char)(u >> 8);
        data[4 * k + 3] = (unsigned char)u;
    }
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return set_error(LT_ERR_NOT_FOUND, std::string("cannot write ") + path);
    const bool ok = std::fwrite(h.text.data(), 1, h.text.size(), f) == h.text.size() &&
                    std::fwrite(data.data(), 1, data.size(), f) == data.size();
    const bool closed = std::fclose(f) == 0;
    if (!ok || !closed) return set_error(LT_ERR_INVALID, std::string("write failed: ") + path);
    return LT_OK;
}

}  // extern "C"
