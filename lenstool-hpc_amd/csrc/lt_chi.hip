// lt_chi.hip -- host side of delensing and chi^2: a batch of lens-model samples against one set of
// observed multiple-image families (src/chi_CPU_barycenter.cpp:58-287, one rank), plus the
// single-model entry points with the reference's argument meaning.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "lenstool_hip/lt_capi.h"
#include "lt_delense.h"
#include "lt_internal.h"
#include "lt_kernels.h"

using namespace lt;

struct lt_chi_batch_s {
    int device = 0;
    int S = 0, H = 0, n = 0, nsets = 0, nimg = 0;
    lt_grid_param frame{};
    double dx = 0, dy = 0;
    std::vector<int> types;
    HaloRec* staging = nullptr;  // pinned host image of the samples' records (one async upload per run)
    hipEvent_t staged = nullptr;  // recorded after that upload: the next run refills staging after it
    std::vector<Run> runs;
    void* block = nullptr;
    HaloRec* recs = nullptr;
    Run* druns = nullptr;
    double* gx = nullptr;
    double* gy = nullptr;
    double* img_gx = nullptr;
    double* img_gy = nullptr;
    DelenseDesc d{};
};

namespace {

// carve aligned sub-buffers out of one allocation
struct Carver {
    size_t off = 0;
    template <class T>
    size_t take(size_t count)
    {
        size_t at = (off + 255) / 256 * 256;
        off = at + count * sizeof(T);
        return at;
    }
};

int validate_families(int nsets, const int* nimages, const lt_galaxy* images, int* nimg)
{
    if (nsets <= 0 || !nimages || !images) return set_error(LT_ERR_INVALID, "image families: bad arguments");
    int tot = 0;
    for (int k = 0; k < nsets; ++k) {
        if (nimages[k] < 1 || nimages[k] > LT_MAXIMPERSOURCE)
            return set_error(LT_ERR_INVALID, "image family " + std::to_string(k) + ": image count must be 1.." +
                                                 std::to_string(LT_MAXIMPERSOURCE));
        tot += nimages[k];
    }
    *nimg = tot;
    return LT_OK;
}

int batch_alloc(lt_chi_batch_s* b, const int* nimages, const lt_galaxy* images)
{
    const size_t S = b->S, K = b->nsets, I = b->nimg, N2 = (size_t)b->n * b->n;
    Carver c;
    size_t o_recs = c.take<HaloRec>(S * std::max(b->H, 1));
    size_t o_runs = c.take<Run>(std::max<size_t>(b->runs.size(), 1));
    size_t o_gx = c.take<double>(S * N2), o_gy = c.take<double>(S * N2);
    size_t o_nim = c.take<int32_t>(K), o_first = c.take<int32_t>(K);
    size_t o_ix = c.take<double>(I), o_iy = c.take<double>(I), o_idr = c.take<double>(I);
    size_t o_igx = c.take<double>(S * I), o_igy = c.take<double>(S * I);
    size_t o_sx = c.take<double>(S * K), o_sy = c.take<double>(S * K), o_sdr = c.take<double>(S * K);
    size_t o_sb = c.take<double>(S * 3);
    size_t o_cnt = c.take<int32_t>(S * K), o_hits = c.take<int32_t>(S * K * DELENSE_CAP);
    size_t o_px = c.take<double>(S * K * LT_MAXIMPERSOURCE), o_py = c.take<double>(S * K * LT_MAXIMPERSOURCE);
    size_t o_loc = c.take<int32_t>(S * K), o_st = c.take<int32_t>(S), o_terms = c.take<double>(S * I);
    size_t o_chi = c.take<double>(S), o_found = c.take<int32_t>(S);
    hipError_t e = hipMalloc(&b->block, c.off);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        b->block = nullptr;
        return set_error(LT_ERR_NOMEM, "chi batch: device allocation of " + std::to_string(c.off) + " bytes failed");
    }
    char* p = (char*)b->block;
    b->recs = (HaloRec*)(p + o_recs);
    b->druns = (Run*)(p + o_runs);
    b->gx = (double*)(p + o_gx);
    b->gy = (double*)(p + o_gy);
    b->img_gx = (double*)(p + o_igx);
    b->img_gy = (double*)(p + o_igy);
    DelenseDesc& d = b->d;
    d.n = b->n;
    d.nsets = b->nsets;
    d.nsamples = b->S;
    d.nimg_total = b->nimg;
    d.xmin = b->frame.xmin;
    d.ymin = b->frame.ymin;
    d.dx = b->dx;
    d.dy = b->dy;
    d.grid_stride = (int64_t)N2;
    d.nimages = (int32_t*)(p + o_nim);
    d.first_image = (int32_t*)(p + o_first);
    d.img_x = (double*)(p + o_ix);
    d.img_y = (double*)(p + o_iy);
    d.img_dr = (double*)(p + o_idr);
    d.src_x = (double*)(p + o_sx);
    d.src_y = (double*)(p + o_sy);
    d.src_dr = (double*)(p + o_sdr);
    d.src_bound = (double*)(p + o_sb);
    d.count = (int32_t*)(p + o_cnt);
    d.hits = (int32_t*)(p + o_hits);
    d.pos_x = (double*)(p + o_px);
    d.pos_y = (double*)(p + o_py);
    d.loc = (int32_t*)(p + o_loc);
    d.status = (int32_t*)(p + o_st);
    d.terms = (double*)(p + o_terms);
    d.chi = (double*)(p + o_chi);
    d.found = (int32_t*)(p + o_found);
    if (nimages && images) {
        std::vector<int32_t> first(K);
        std::vector<double> ix(I), iy(I), idr(I);
        int at = 0;
        for (size_t k = 0; k < K; ++k) {
            first[k] = at;
            at += nimages[k];
        }
        for (size_t i = 0; i < I; ++i) {
            ix[i] = images[i].center.x;
            iy[i] = images[i].center.y;
            idr[i] = images[i].dr;
        }
        LT_HIP_TRY(hipMemcpy((void*)d.nimages, nimages, K * sizeof(int32_t), hipMemcpyHostToDevice));
        LT_HIP_TRY(hipMemcpy((void*)d.first_image, first.data(), K * sizeof(int32_t), hipMemcpyHostToDevice));
        LT_HIP_TRY(hipMemcpy((void*)d.img_x, ix.data(), I * sizeof(double), hipMemcpyHostToDevice));
        LT_HIP_TRY(hipMemcpy((void*)d.img_y, iy.data(), I * sizeof(double), hipMemcpyHostToDevice));
        LT_HIP_TRY(hipMemcpy((void*)d.img_dr, idr.data(), I * sizeof(double), hipMemcpyHostToDevice));
    }
    return LT_OK;
}

int reset_counters(lt_chi_batch_s* b, hipStream_t s)
{
    LT_HIP_TRY(launch_fill_i32(b->d.count, 0, (size_t)b->S * b->nsets, s));
    LT_HIP_TRY(launch_fill_i32(b->d.status, 0, (size_t)b->S, s));
    return LT_OK;
}

}  // namespace

extern "C" {

int lt_chi_batch_create(lt_chi_batch* out, int nsamples, int nhalos, const int* types, int nbgridcells,
                        const lt_grid_param* frame, int nsets, const int* nimages, const lt_galaxy* images)
{
    if (!out) return set_error(LT_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (nsamples <= 0 || nhalos < 0 || (nhalos && !types) || nbgridcells < 2 || !frame)
        return set_error(LT_ERR_INVALID, "chi batch: bad sizes");
    int nimg, st;
    if ((st = validate_families(nsets, nimages, images, &nimg))) return st;
    int dev;
    if ((st = current_device(&dev))) return st;
    lt_chi_batch_s* b = new lt_chi_batch_s();
    b->device = dev;
    b->S = nsamples;
    b->H = nhalos;
    b->n = nbgridcells;
    b->nsets = nsets;
    b->nimg = nimg;
    b->frame = *frame;
    b->dx = (frame->xmax - frame->xmin) / (nbgridcells - 1);
    b->dy = (frame->ymax - frame->ymin) / (nbgridcells - 1);
    b->types.assign(types, types + nhalos);
    // run structure from the types (identical for every sample)
    for (int s0 = 0; s0 < nhalos;) {
        int c = 1;
        while (s0 + c < nhalos && types[s0 + c] == types[s0]) ++c;
        Run r{};
        r.type = types[s0];
        r.start = s0;
        r.count = c;
        r.weight = 1.0;
        b->runs.push_back(r);
        s0 += c;
    }
    for (int t : b->types)
        if (t != 5 && t != 8 && t != 81) {
            delete b;
            return set_error(LT_ERR_TYPE, "chi batch: halo type " + std::to_string(t) + " has no gradient");
        }
    if ((st = batch_alloc(b, nimages, images))) {
        delete b;
        return st;
    }
    if (!b->runs.empty() &&
        hipMemcpy(b->druns, b->runs.data(), b->runs.size() * sizeof(Run), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(b->block);
        delete b;
        return set_error(LT_ERR_HIP, "chi batch: run upload failed");
    }
    if (hipHostMalloc((void**)&b->staging, (size_t)nsamples * std::max(nhalos, 1) * sizeof(HaloRec),
                      hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&b->staged, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        if (b->staging) (void)hipHostFree(b->staging);
        (void)hipFree(b->block);
        delete b;
        return set_error(LT_ERR_NOMEM, "chi batch: pinned staging allocation failed");
    }
    *out = b;
    return LT_OK;
}

int lt_chi_batch_destroy(lt_chi_batch b)
{
    if (!b) return LT_OK;
    (void)hipSetDevice(b->device);
    hipError_t e = b->block ? hipFree(b->block) : hipSuccess;
    if (b->staged) {
        (void)hipEventSynchronize(b->staged);
        (void)hipEventDestroy(b->staged);
    }
    const hipError_t eh = b->staging ? hipHostFree(b->staging) : hipSuccess;
    if (e == hipSuccess) e = eh;
    delete b;
    if (e != hipSuccess) return hip_error(e, "chi batch free");
    return LT_OK;
}

int lt_chi_batch_run(lt_chi_batch b, const lt_potential_soa* lenses, double* chi, int* found, int* status, void* stream)
{
    if (!b) return set_error(LT_ERR_NOT_FOUND, "chi batch is NULL");
    if (!lenses || !chi) return set_error(LT_ERR_INVALID, "chi batch run: NULL argument");
    LT_HIP_TRY(hipSetDevice(b->device));
    hipStream_t s = (hipStream_t)stream;
    std::vector<HaloRec> recs;
    std::vector<HaloRec32> recs32;
    std::vector<Run> runs;
    std::string err;
    // the previous run's upload may still read the staging area if that run stopped on an error
    LT_HIP_TRY(hipEventSynchronize(b->staged));
    // halo types of every sample, and the circular-SIS flag of each run: set when every halo of the run
    // is a circular SIS in every sample (the packer's per-sample rule, ANDed), known before any launch
    for (Run& r : b->runs) r.circ = r.type == TYPE_SIS ? 1 : 0;
    for (int i = 0; i < b->S; ++i) {
        const lt_potential_soa& L = lenses[i];
        for (int h = 0; h < b->H; ++h)
            if (!L.type || L.type[h] != b->types[h])
                return set_error(LT_ERR_INVALID, "sample " + std::to_string(i) + ": halo types differ from the batch's");
        for (Run& r : b->runs)
            for (int h = r.start; h < r.start + r.count && r.circ; ++h)
                r.circ = L.ellipticity_potential && L.ellipticity_potential[h] == 0.0;
    }
    if (!b->runs.empty())
        LT_HIP_TRY(hipMemcpyAsync(b->druns, b->runs.data(), b->runs.size() * sizeof(Run), hipMemcpyHostToDevice, s));
    int st = reset_counters(b, s);
    if (st) return st;
    // gradient grids of every sample (src/chi_CPU_barycenter.cpp:125-137), launched per chunk of samples
    // as soon as the chunk is packed, so that the host packs the next chunk while the device computes
    GridDesc g{};
    g.xmin = b->frame.xmin;
    g.ymin = b->frame.ymin;
    g.dx = b->dx;
    g.dy = b->dy;
    g.nx = b->n;
    g.ny = b->n;
    g.pitch = b->n;
    g.halo_stride = b->H;
    g.plane_stride = (int64_t)b->n * b->n;
    OutDesc o{};
    o.kind = LT_MAP_GRADIENT;
    o.r = 1.0;
    o.mult = 1.0;
    const int chunk = std::max(64, (b->S + 7) / 8);
    for (int c0 = 0; c0 < b->S; c0 += chunk) {
        const int cn = std::min(chunk, b->S - c0);
        for (int i = c0; i < c0 + cn; ++i) {
            st = pack_model(&lenses[i], b->H, Use::Gradient, recs, recs32, runs, err);
            if (!st && runs.size() != b->runs.size()) {
                st = LT_ERR_INVALID;
                err = "run structure differs";
            }
            if (st) {
                // uploads of earlier chunks may still read the staging area: the next run waits for them
                const hipError_t e = hipEventRecord(b->staged, s);
                if (e != hipSuccess) return hip_error(e, "chi batch: event record");
                return set_error(st, "sample " + std::to_string(i) + ": " + err);
            }
            if (b->H) std::memcpy(&b->staging[(size_t)i * b->H], recs.data(), (size_t)b->H * sizeof(HaloRec));
        }
        if (b->H)
            LT_HIP_TRY(hipMemcpyAsync(b->recs + (size_t)c0 * b->H, b->staging + (size_t)c0 * b->H,
                                      (size_t)cn * b->H * sizeof(HaloRec), hipMemcpyHostToDevice, s));
        g.nsamples = cn;
        o.p0 = b->gx + (size_t)c0 * g.plane_stride;
        o.p1 = b->gy + (size_t)c0 * g.plane_stride;
        // 4 pixels per thread: with few halos per sample each halo record's scalar load is shared by
        // more pixels (1024 samples of 11 halos on 512^2: 9.9 ms per batch step against 10.4 at 2; the
        // planes are bit-identical for every pixels-per-thread)
        LT_HIP_TRY(launch_gradient64(b->recs + (size_t)c0 * b->H, b->druns, (int)b->runs.size(), g, o, 4, s));
    }
    LT_HIP_TRY(hipEventRecord(b->staged, s));
    LT_HIP_TRY(launch_points_gradient64(b->recs, b->druns, (int)b->runs.size(), b->d.img_x, b->d.img_y, b->nimg,
                                        b->S, b->H, b->img_gx, b->img_gy, s));
    LT_HIP_TRY(launch_barycentres(b->img_gx, b->img_gy, b->d, s));
    LT_HIP_TRY(launch_delense(b->gx, b->gy, b->d, s));
    LT_HIP_TRY(launch_chi(b->d, s));
    LT_HIP_TRY(hipMemcpyAsync(chi, b->d.chi, (size_t)b->S * sizeof(double), hipMemcpyDeviceToHost, s));
    if (found) LT_HIP_TRY(hipMemcpyAsync(found, b->d.found, (size_t)b->S * sizeof(int), hipMemcpyDeviceToHost, s));
    if (status) LT_HIP_TRY(hipMemcpyAsync(status, b->d.status, (size_t)b->S * sizeof(int), hipMemcpyDeviceToHost, s));
    LT_HIP_TRY(hipStreamSynchronize(s));
    return LT_OK;
}

int lt_chi_batch_images(lt_chi_batch b, int sample, lt_point* image_pos, int* locimagesfound)
{
    if (!b) return set_error(LT_ERR_NOT_FOUND, "chi batch is NULL");
    if (sample < 0 || sample >= b->S || !image_pos || !locimagesfound) return set_error(LT_ERR_INVALID, "bad arguments");
    const size_t K = b->nsets, M = K * LT_MAXIMPERSOURCE;
    std::vector<double> px(M), py(M);
    LT_HIP_TRY(hipSetDevice(b->device));
    LT_HIP_TRY(hipMemcpy(px.data(), b->d.pos_x + sample * M, M * sizeof(double), hipMemcpyDeviceToHost));
    LT_HIP_TRY(hipMemcpy(py.data(), b->d.pos_y + sample * M, M * sizeof(double), hipMemcpyDeviceToHost));
    LT_HIP_TRY(hipMemcpy(locimagesfound, b->d.loc + sample * K, K * sizeof(int), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < K; ++k)
        for (int j = 0; j < LT_MAXIMPERSOURCE; ++j) {
            image_pos[k * LT_MAXIMPERSOURCE + j].x = j < locimagesfound[k] ? px[k * LT_MAXIMPERSOURCE + j] : 0.0;
            image_pos[k * LT_MAXIMPERSOURCE + j].y = j < locimagesfound[k] ? py[k * LT_MAXIMPERSOURCE + j] : 0.0;
        }
    return LT_OK;
}

int lt_delense_barycenter(lt_point* image_pos, int* locimagesfound, int* numimg, int nsets, int nbgridcells,
                          const lt_grid_param* frame, const lt_galaxy* sources, const double* gx, const double* gy)
{
    if (!image_pos || !locimagesfound || !numimg || !frame || !sources || !gx || !gy || nsets <= 0 || nbgridcells < 2)
        return set_error(LT_ERR_INVALID, "delense_barycenter: bad arguments");
    int dev, st;
    if ((st = current_device(&dev))) return st;
    lt_chi_batch_s b;
    b.device = dev;
    b.S = 1;
    b.H = 0;
    b.n = nbgridcells;
    b.nsets = nsets;
    b.nimg = 1;
    b.frame = *frame;
    b.dx = (frame->xmax - frame->xmin) / (nbgridcells - 1);
    b.dy = (frame->ymax - frame->ymin) / (nbgridcells - 1);
    if ((st = batch_alloc(&b, nullptr, nullptr))) return st;
    const size_t N2 = (size_t)nbgridcells * nbgridcells;
    std::vector<double> sx(nsets), sy(nsets), sdr(nsets);
    for (int k = 0; k < nsets; ++k) {
        sx[k] = sources[k].center.x;
        sy[k] = sources[k].center.y;
        sdr[k] = sources[k].dr;
    }
    std::vector<int> cnt(nsets), stv(1);
    hipError_t e = hipSuccess;
    e = hipMemcpy(b.gx, gx, N2 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b.gy, gy, N2 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b.d.src_x, sx.data(), nsets * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b.d.src_y, sy.data(), nsets * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b.d.src_dr, sdr.data(), nsets * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_fill_i32(b.d.count, 0, nsets, nullptr);
    if (e == hipSuccess) e = launch_fill_i32(b.d.status, 0, 1, nullptr);
    if (e == hipSuccess) e = launch_delense(b.gx, b.gy, b.d, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e == hipSuccess) e = hipMemcpy(cnt.data(), b.d.count, nsets * sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(stv.data(), b.d.status, sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        (void)hipFree(b.block);
        return hip_error(e, "delense_barycenter");
    }
    st = lt_chi_batch_images(&b, 0, image_pos, locimagesfound);
    (void)hipFree(b.block);
    if (st) return st;
    for (int k = 0; k < nsets; ++k) *numimg += cnt[k];
    if (stv[0] != LT_OK) return set_error(stv[0], "delense_barycenter: more than 20 images for one source (kept the first 20 in scan order)");
    return LT_OK;
}

int lt_chi_barycentersource(double* chi, int* error, int nbgridcells, int nsets, int nhalos, const lt_potential_soa* lens,
                            const lt_grid_param* frame, const int* nimages, const lt_galaxy* images)
{
    if (!chi || !lens) return set_error(LT_ERR_INVALID, "chi: NULL argument");
    lt_chi_batch b = nullptr;
    int st = lt_chi_batch_create(&b, 1, nhalos, lens->type, nbgridcells, frame, nsets, nimages, images);
    int found = 0, sst = LT_OK;
    if (st == LT_OK) st = lt_chi_batch_run(b, lens, chi, &found, &sst, nullptr);
    lt_chi_batch_destroy(b);
    if (st == LT_OK && sst != LT_OK) st = set_error(sst, "chi: more than 20 images for one source (kept the first 20)");
    if (error) *error = st;
    return st;
}

int lt_chi_computation(double* chi, int* images_found, const int* numimagesfound, const lt_point* imagesposition,
                       const int* nimages, const lt_galaxy* images, int nsets)
{
    if (!chi || !images_found || !numimagesfound || !imagesposition) return set_error(LT_ERR_INVALID, "chi: NULL argument");
    int nimg, st;
    if ((st = validate_families(nsets, nimages, images, &nimg))) return st;
    int dev;
    if ((st = current_device(&dev))) return st;
    lt_chi_batch_s b;
    b.device = dev;
    b.S = 1;
    b.n = 2;
    b.nsets = nsets;
    b.nimg = nimg;
    if ((st = batch_alloc(&b, nimages, images))) return st;
    const size_t M = (size_t)nsets * LT_MAXIMPERSOURCE;
    std::vector<double> px(M), py(M);
    std::vector<int> loc(nsets);
    for (int k = 0; k < nsets; ++k) loc[k] = std::min(std::max(numimagesfound[k], 0), LT_MAXIMPERSOURCE);
    for (size_t i = 0; i < M; ++i) {
        px[i] = imagesposition[i].x;
        py[i] = imagesposition[i].y;
    }
    double out = 0.;
    int found = 0;
    hipError_t e = hipMemcpy(b.d.pos_x, px.data(), M * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b.d.pos_y, py.data(), M * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b.d.loc, loc.data(), nsets * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_chi(b.d, nullptr);
    if (e == hipSuccess) e = hipMemcpy(&out, b.d.chi, sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&found, b.d.found, sizeof(int), hipMemcpyDeviceToHost);
    (void)hipFree(b.block);
    if (e != hipSuccess) return hip_error(e, "chi computation");
    *chi += out;
    *images_found += found;
    return LT_OK;
}

int lt_point_gradients(double* gx, double* gy, const double* px, const double* py, int npts,
                       const lt_potential_soa* lens, int nhalos)
{
    if (npts < 0 || (npts && (!gx || !gy || !px || !py))) return set_error(LT_ERR_INVALID, "bad point arrays");
    if (!npts) return LT_OK;
    std::vector<HaloRec> recs;
    std::vector<HaloRec32> recs32;
    std::vector<Run> runs;
    std::string err;
    int st = pack_model(lens, nhalos, Use::Gradient, recs, recs32, runs, err);
    if (st) return set_error(st, err);
    int dev;
    if ((st = current_device(&dev))) return st;
    Carver c;
    size_t o_r = c.take<HaloRec>(std::max<size_t>(recs.size(), 1)), o_u = c.take<Run>(std::max<size_t>(runs.size(), 1));
    size_t o_x = c.take<double>(npts), o_y = c.take<double>(npts), o_gx = c.take<double>(npts), o_gy = c.take<double>(npts);
    void* blk = nullptr;
    LT_HIP_TRY(hipMalloc(&blk, c.off));
    char* p = (char*)blk;
    hipError_t e = hipSuccess;
    if (!recs.empty()) e = hipMemcpy(p + o_r, recs.data(), recs.size() * sizeof(HaloRec), hipMemcpyHostToDevice);
    if (e == hipSuccess && !runs.empty())
        e = hipMemcpy(p + o_u, runs.data(), runs.size() * sizeof(Run), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p + o_x, px, npts * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p + o_y, py, npts * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = launch_points_gradient64((HaloRec*)(p + o_r), (Run*)(p + o_u), (int)runs.size(), (double*)(p + o_x),
                                     (double*)(p + o_y), npts, 1, 0, (double*)(p + o_gx), (double*)(p + o_gy), nullptr);
    if (e == hipSuccess) e = hipMemcpy(gx, p + o_gx, npts * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(gy, p + o_gy, npts * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(blk);
    if (e != hipSuccess) return hip_error(e, "point gradients");
    return LT_OK;
}

}  // extern "C"
