/*
 * lt_capi.h -- C-ABI of the MI355X lens-map library (liblenstool_hip.so).
 *
 * Plain C: pointers, sizes and the structures of lt_types.h; no C++ or HIP types in any signature
 * (streams are passed as void*, a hipStream_t or NULL for the library's own stream). Every function
 * returns an LT_* status (lt_status.h); lt_last_error() gives the message of the calling thread's
 * last failure.
 *
 * Two families of entry points:
 *  1. Synchronous, host-array entry points with the argument meaning of the reference's GPU
 *     wrappers (each cites the function it replaces). Inputs are host SOA arrays, outputs are
 *     caller-allocated host arrays of nx*ny doubles, row-major (index = row*nx + col), pixel (col, row)
 *     at (xmin + (col+istart) dx, ymin + (row+jstart) dy). They keep device buffers cached between
 *     calls instead of allocating and freeing every time, and are serialised by an internal lock.
 *  2. Resident models and device-pointer entry points: a lens model is packed and uploaded once
 *     (lt_model_create) and maps are computed asynchronously on a caller stream into device memory.
 *
 * precision: LT_PREC_FP64 or LT_PREC_FP32C (lt_status.h). FP32C (FP32 halo terms with FP64 offsets
 * and FP64 chunk sums; FP64 terms next to the cores for the Hessian, gradient and displacement) applies
 * to every map kind.
 */
#ifndef LENSTOOL_HIP_LT_CAPI_H
#define LENSTOOL_HIP_LT_CAPI_H

#include "lt_status.h"
#include "lt_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- library ----------------------------------------------------------------------------------- */
const char* lt_last_error(void);
int lt_version(void);                       /* 10000*major + 100*minor + patch                  */
int lt_device_count(int* count);            /* number of visible GPUs                           */
int lt_init(int device);                    /* select the GPU used by the calling thread        */
int lt_sync(void);                          /* wait for the library stream                      */

/* ---- device memory, streams and events (for callers without their own HIP runtime) ------------- */
int lt_dev_alloc(void** ptr, long long bytes);
int lt_dev_free(void* ptr);
int lt_memcpy_h2d(void* dst, const void* src, long long bytes, void* stream);  /* waits when stream == NULL */
int lt_memcpy_d2h(void* dst, const void* src, long long bytes, void* stream);
/* bytes of dst set to (unsigned char)value by a fill kernel (no memset node inside a captured graph) */
int lt_memset_d(void* dst, int value, long long bytes, void* stream);
/* device-to-device copy done by a kernel (a streaming copy of known size, also used to calibrate
 * hardware traffic counters) */
int lt_memcpy_d2d(void* dst, const void* src, long long bytes, void* stream);
int lt_stream_create(void** stream);
int lt_stream_destroy(void* stream);
int lt_stream_sync(void* stream);
int lt_event_create(void** event);
int lt_event_destroy(void* event);
int lt_event_record(void* event, void* stream);
int lt_event_sync(void* event);
int lt_event_elapsed(float* ms, void* start, void* stop);

/* ---- reference-compatible synchronous entry points --------------------------------------------- */

/* alpha = sum of halo gradients. Replaces gradient_grid_GPU(..., dx, dy, nx, ny, istart, jstart)
 * (src/grid_gradient_GPU.cuh:36, src/grid_gradient_GPU.cu:87-179). */
int lt_gradient_grid(double* grid_grad_x, double* grid_grad_y, const lt_grid_param* frame,
                     const lt_potential_soa* lens, int nhalos, double dx, double dy,
                     int nbgridcells_x, int nbgridcells_y, int istart, int jstart, int precision);

/* Hessian planes a = phi_xx, b = d = phi_xy, c = phi_yy (any plane may be NULL). Replaces
 * gradient2_grid_GPU(a, b, c, d, ..., dx, dy, nx, ny, istart, jstart)
 * (src/grid_gradient2_GPU.cuh:34, src/grid_gradient2_GPU.cu:67-183). */
int lt_gradient2_grid(double* a, double* b, double* c, double* d, const lt_grid_param* frame,
                      const lt_potential_soa* lens, int nhalos, double dx, double dy,
                      int nbgridcells_x, int nbgridcells_y, int istart, int jstart, int precision);

/* Hessian with every (type, dlsds) run weighted by lens->dlsds. Replaces
 * module_potentialDerivatives_Kmap_SOA_CPU_GPU (src/grid_gradient2_GPU.cuh:43,
 * src/grid_gradient2_GPU.cu:212-236). */
int lt_kmap_grid(double* a, double* b, double* c, double* d, const lt_grid_param* frame,
                 const lt_potential_soa* lens, int nhalos, double dx, double dy,
                 int nbgridcells_x, int nbgridcells_y, int istart, int jstart, int precision);

/* Lensing potential. Replaces potential_grid_GPU(..., dx, dy, nx, ny, istart, jstart)
 * (src/grid_potential_GPU.cuh:31); types 5, 8 and 81 (the reference has 81 only). FP64 (lt_map_device
 * with LT_MAP_POT also has an FP32C kernel). */
int lt_potential_grid(double* grid_potential, const lt_grid_param* frame, const lt_potential_soa* lens,
                      int nhalos, double dx, double dy, int nbgridcells_x, int nbgridcells_y,
                      int istart, int jstart);

/* One map of the reference runmode families, Hessian (or gradient / potential) and epilogue fused
 * in one kernel. Replaces
 *   map_grid_ampli_GPU (src/grid_map_ampli_GPU.cuh:51)   kind LT_MAP_AMPLI, modes 1..6
 *   map_grid_shear_GPU (src/grid_map_shear_GPU.cuh:49)   kind LT_MAP_SHEAR, modes 1..4
 *   map_grid_mass_GPU  (src/grid_map_mass_GPU.cuh:48)    kind LT_MAP_MASS,  modes 1..4 (Kmap)
 *   map_grid_potential_GPU (src/grid_map_pot_GPU.cuh:48) kind LT_MAP_POT,   modes 1..3
 *   map_grid_dpl_GPU   (src/grid_map_dpl_GPU.cuh:45)     kind LT_MAP_DPL,   map0 = x, map1 = y
 * z_lens is the lens plane (the reference passes lens->z[0]; for LT_MAP_MASS 0 means lens->z[0]),
 * z_source the source plane. map1 is used by LT_MAP_DPL only. */
int lt_map_grid(int kind, int mode, double* map0, double* map1, const lt_cosmo_param* cosmo,
                const lt_grid_param* frame, const lt_potential_soa* lens, int nhalos, double z_lens,
                double z_source, double dx, double dy, int nbgridcells_x, int nbgridcells_y, int istart,
                int jstart, int precision);

/* kappa, gamma1, gamma2, det A = (1-kappa)^2 - gamma^2 and mu = 1/det A in one pass (any output may
 * be NULL), for the source plane z_source behind a lens at z_lens. No reference counterpart: the
 * reference computes each of these with its own Hessian launch (src/grid_map_ampli_GPU.cu:93-192). */
int lt_map_fields(double* kappa, double* gamma1, double* gamma2, double* det_a, double* mu,
                  const lt_cosmo_param* cosmo, const lt_grid_param* frame, const lt_potential_soa* lens,
                  int nhalos, double z_lens, double z_source, double dx, double dy, int nbgridcells_x,
                  int nbgridcells_y, int istart, int jstart, int precision);

/* ---- cosmology (src/module_cosmodistances.cpp:94-174), distances in units of c/H0 -------------- */
double lt_cosmo_observer_object(double z, const lt_cosmo_param* cosmo);
double lt_cosmo_object_object(double z1, double z2, const lt_cosmo_param* cosmo);
double lt_cosmo_dlsds(double z_lens, double z_source, const lt_cosmo_param* cosmo);

/* ---- resident models and device-pointer entry points ------------------------------------------- */
typedef struct lt_model_s* lt_model;

/* Pack and upload a lens model; use = LT_MAP_GRADIENT, LT_MAP_HESSIAN (also maps of kinds AMPLI,
 * SHEAR, FIELDS), LT_MAP_MASS (Kmap runs) or LT_MAP_POT. */
int lt_model_create(lt_model* out, const lt_potential_soa* lens, int nhalos, int use);
/* Replace the parameters of a model in place (same halo count and types), e.g. per Bayes sample;
 * ordered on `stream` after the work already queued there. Returns once the parameters are packed
 * into the model's pinned staging buffer; the upload itself is asynchronous. */
int lt_model_update(lt_model model, const lt_potential_soa* lens, int nhalos, void* stream);
/* A batch of nsamples lens models with the same halo types (and, for LT_MAP_MASS, the same dlsds
 * runs), e.g. Bayes samples of one model (the reference's per-sample loop in
 * utils/maps/main.cpp:428-587). One lt_map_device call then computes every sample's map in one
 * launch; sample s writes its planes at out + s * pitch * ny. */
int lt_model_create_batch(lt_model* out, const lt_potential_soa* lenses, int nsamples, int nhalos, int use);
int lt_model_update_batch(lt_model model, const lt_potential_soa* lenses, int nsamples, void* stream);
/* The two halves of an update, for captured launch sequences: lt_model_stage packs new parameters
 * into the pinned staging buffer (host only; it first waits for the previous upload to have read
 * it), lt_model_upload queues the staging -> device copy on a stream and may be captured into a
 * graph, so that each replay uploads whatever was staged last. A caller replaying such a graph
 * waits for the replay to finish before staging again. */
int lt_model_stage(lt_model model, const lt_potential_soa* lenses, int nsamples);
int lt_model_upload(lt_model model, void* stream);
int lt_model_destroy(lt_model model);
/* device bytes held by a model */
int lt_model_bytes(lt_model model, long long* bytes);

/* Grid of a device map. pitch = elements between rows (>= nx). */
typedef struct lt_grid_desc {
    double xmin, ymin, dx, dy;
    int nx, ny, istart, jstart;
    long long pitch;
} lt_grid_desc;

/* The r (D_ls/D_s) and mult arguments of lt_map_device for a reference map kind and mode, formed
 * as the reference's map wrappers form them (src/grid_map_ampli_GPU.cu:110-112,
 * src/grid_map_mass_GPU.cu:104-106, 219-305, src/grid_map_pot_GPU.cu:100-102, 178-270): mass modes 2, 3, 4 scale by
 * the critical density in g/cm^2, 1e12 Msun/pixel, 1e12 Msun/kpc^2; potential mode 1 by r,
 * mode 2 by c^2 / (2 pi G) h / D_ol. dx, dy (arcsec) matter for mass mode 3 only. */
int lt_map_factors(int kind, int mode, const lt_cosmo_param* cosmo, double z_lens, double z_source, double dx,
                   double dy, double* r, double* mult);

/* Compute one map into device arrays on `stream` (asynchronous). kind/mode as lt_map_grid plus
 * LT_MAP_HESSIAN (out0..out3 = a, b, c, d), LT_MAP_GRADIENT (out0, out1 = x, y) and LT_MAP_FIELDS
 * (out0..out4 = kappa, gamma1, gamma2, det A, mu). r = D_ls/D_s used by the map formulas, mult = the
 * mass / potential multiplier of the mode. px = launch shape: 1, 2 or 4 pixels per thread (results
 * identical bit for bit), 0 = the tuned default, or LT_PX_HALO_LANES (FP32C Hessian-derived maps
 * only): lanes over halos, a wavefront per 16-pixel row segment; its FP32 terms are summed in
 * another order and its near-core test is another region, so its results differ from the other
 * shapes' within the FP32C error; it is never the tuned default (measured slower than px 2 on a
 * 503-halo map: 12.9 against 8.95 ms at 4096^2). */
#define LT_PX_HALO_LANES 16
int lt_map_device(lt_model model, int kind, int mode, double* out0, double* out1, double* out2, double* out3,
                  double* out4, double r, double mult, const lt_grid_desc* grid, int precision, int px,
                  void* stream);

/* Time the map with 1, 2 and 4 pixels per thread (one warm-up and two timed launches each, into
 * scratch planes) and make the fastest the default (px = 0) for maps of this family, precision,
 * halo-count octave and pixel-count octave in this process. ms[3] (optional) gets the per-launch
 * times. Results do not depend on these shapes: only speed does. */
int lt_autotune(lt_model model, int kind, int mode, const lt_grid_desc* grid, int precision, void* stream,
                int* best_px, float* ms);

/* ---- captured launch sequences (HIP graphs) ------------------------------------------------------
 * Work issued on a caller stream between lt_graph_begin and lt_graph_end (lt_model_upload,
 * lt_map_device, ...) is captured, not run; lt_graph_launch replays it on any stream with one
 * submission. The captured calls issue no memset nodes. */
typedef struct lt_graph_s* lt_graph;
int lt_graph_begin(void* stream);
int lt_graph_end(void* stream, lt_graph* out);
int lt_graph_launch(lt_graph graph, void* stream);
int lt_graph_destroy(lt_graph graph);

/* ---- map-serving loop ---------------------------------------------------------------------------
 * The reference has no counterpart (its map drivers, utils/exec/main.cpp:397-545, run one
 * synchronous map after another). A server keeps lens models resident on the calling thread's GPU
 * and admits queued requests as HBM frees up (hbm_budget_bytes, reserved at creation; the output
 * planes of admitted requests are carved out of it and held until fetched). Admitted requests are
 * cut into row tiles, which the nslots streams take in turn from every admitted request, so work
 * of different requests overlaps on the device and a small request does not wait for the whole of
 * a large one. A later request is admitted ahead of an earlier one that does not fit only when that
 * cannot keep the earlier one out for good: a caller that fetches its results in submission order
 * never waits forever. Results stay in device memory until fetched. */
typedef struct lt_server_s* lt_server;
typedef struct lt_request {
    int model;            /* id from lt_server_add_model                                          */
    int kind, mode;       /* LT_MAP_* and mode, as lt_map_device                                  */
    int precision;        /* LT_PREC_FP64 or LT_PREC_FP32C                                        */
    lt_grid_desc grid;    /* pitch is ignored (planes are nx * ny)                                */
    double r, mult;       /* map-formula scalars, as lt_map_device                                */
} lt_request;
typedef struct lt_request_times {
    double submit_ms, start_ms, done_ms, fetched_ms;  /* since server creation; done = observed on device */
} lt_request_times;
typedef struct lt_server_stats {
    long long submitted, completed, failed, queued, running;
    double p50_ms, p99_ms;              /* submit -> result ready on the device                   */
    double p50_fetch_ms, p99_fetch_ms;  /* submit -> result copied to the host by lt_server_fetch */
    double halo_px;                     /* sum over completed requests of halos x pixels          */
    double elapsed_ms;                  /* since creation or the last reset                       */
    long long bytes_in_use, bytes_peak;
} lt_server_stats;
/* nslots concurrent requests; hbm_budget_bytes of device memory are reserved at creation (the
 * result planes of admitted requests are carved out of it, so serving allocates nothing) */
int lt_server_create(lt_server* out, int nslots, long long hbm_budget_bytes);
/* Pack the model for every map kind it supports (gradient, Hessian, Kmap when dlsds is given,
 * potential); *model_id indexes it in requests. */
int lt_server_add_model(lt_server server, const lt_potential_soa* lens, int nhalos, int* model_id);
/* Queue a request (non-blocking). LT_ERR_CAPACITY if its planes alone exceed the budget. */
int lt_server_submit(lt_server server, const lt_request* req, long long* ticket);
/* Wait for a request (timeout_ms < 0: no time limit); the request's status if it failed. LT_ERR_BUSY
 * on timeout, and also, whatever the timeout, when the request cannot be admitted until results
 * submitted before it are fetched and no work or fetch is under way (the wait would never end).
 * LT_ERR_NOT_FOUND if the ticket is unknown or is fetched by another thread during the wait. */
int lt_server_wait(lt_server server, long long ticket, double timeout_ms);
/* Wait, copy the request's planes to the given host arrays (NULL = skip a plane), release the
 * ticket and its device memory. Plane order as lt_map_device. LT_ERR_BUSY as lt_server_wait (the
 * ticket stays valid). */
int lt_server_fetch(lt_server server, long long ticket, double* out0, double* out1, double* out2, double* out3,
                    double* out4, lt_request_times* times);
int lt_server_get_stats(lt_server server, lt_server_stats* stats);
int lt_server_reset_stats(lt_server server);
/* Stops the dispatcher, waits for the device and frees everything (unfetched results included). */
int lt_server_destroy(lt_server server);

/* ---- Lenstool parameter files ---------------------------------------------------------------------
 * Reads a .par file and the catalogues it names (potfiles, multiple-image file) the way the
 * reference's readers do (src/module_readParameters.cpp: cosmology :278-384, runmode :385-851,
 * potentials :2234-2489, potfiles :1102-1656, frame :2954-3030, images :852-935): potentials slotted by
 * ascending type, then potfile clumps with their scaling relations; kpc radii converted with the
 * file's cosmology. Catalogue paths are taken as written, else relative to the .par's directory. */
typedef struct lt_par_s* lt_par;
typedef struct lt_par_info {
    int nhalos;               /* potentials + potfile clumps                                  */
    int nhalos_main;          /* potentiel blocks                                             */
    int npotfiles;
    int nbgridcells;          /* runmode nbgridcells (default 1000)                            */
    int ampli, ampli_cells;  double ampli_z;   /* runmode ampli  <mode> <cells> <z>          */
    int mass, mass_cells;    double mass_z;    /* runmode mass                                */
    int shear, shear_cells;  double shear_z;   /* runmode shear                               */
    int poten, poten_cells;  double poten_z;   /* runmode poten                               */
    int dpl, dpl_cells;      double dpl_z;     /* runmode dpl                                 */
    int inverse;
    int ref_mode;  double ref_ra, ref_dec;     /* runmode reference                           */
    double z_source;          /* source block z_source (default 0.8); the dlsds of the model    */
    int nsets, nimagestot;    /* multiple-image families and images                           */
} lt_par_info;
int lt_par_load(lt_par* out, const char* path);
/* Run settings, the lens model (arrays owned by the parameter set, valid until lt_par_destroy;
 * type, position, b0, vdisp, angle, ellipticity, epot, rcore, rcut, z, mag, lum, theta, cos/sin,
 * dlsds filled), the frame and the cosmology. Any pointer may be NULL. */
int lt_par_get(lt_par par, lt_par_info* info, lt_potential_soa* lens, lt_grid_param* frame, lt_cosmo_param* cosmo);
/* The multiple images (arrays of nsets and nimagestot elements, caller-allocated; NULL skips). */
int lt_par_images(lt_par par, int* nimages_per_set, lt_galaxy* images);
/* Reads a bayes.dat file of MCMC samples for this model (module_readParameters_preparebayes and
 * _bayesmodels, src/module_readParameters.cpp:53-108; a path instead of ./bayes.dat): one column per
 * '#' line, one sample per line. A line with more values than columns is LT_ERR_INVALID. */
int lt_par_bayes_load(lt_par par, const char* path, int* nsamples, int* ncolumns);
/* Sets the lens model to bayes sample `sample` (module_readParameters_setbayesmapmodels,
 * src/module_readParameters.cpp:132-227): columns from index 2 fill the parameters each `limit`
 * block frees (block >= 1), potential by potential in file order, in the order x, y, epot,
 * ellipticity, angle [deg], rcore, vdisp, rcut, z; b0 and the ellipticity pair are recomputed; with
 * potfiles, the image-redshift (z_m_limit) columns are skipped, then each potfile's cut [arcsec] and
 * sigma (if freed) are read and its scaling relations re-applied. In place, like the reference
 * (lt_par_get then returns the sample's model); D_ls/D_s is left as read. A sample with fewer
 * values than the free parameters need is LT_ERR_INVALID and leaves the model unchanged. */
int lt_par_bayes_apply(lt_par par, int sample);
/* File name of a runmode map line (`ampli`, `shear`, `mass`, `poten`; `dpl` has two: which = 0 for
 * x, 1 for y), as read by read_runmode (src/module_readParameters.cpp:415-465); "" if absent. */
int lt_par_map_name(lt_par par, int kind, int which, char* buf, int cap);
/* The `cline` block (cline_param, src/structure_hpc.hpp:563-578; its reader
 * src/module_readParameters.cpp:1008-1095 is compiled out in the reference): source planes
 * (nplan clamped to LT_CLINE_MAXPLANES), the half-width dmax of the square frame, the step
 * (`pas` / `step` / `limitLow`), limitHigh and nbgridcells (0 if absent). Defaults as the
 * reference's reader: dmax 1, step 1, limitHigh 10. */
#define LT_CLINE_MAXPLANES 9
typedef struct lt_cline_info {
    int nplan;
    double z[LT_CLINE_MAXPLANES];
    double dmax, step, limit_high;
    int nbgridcells;
} lt_cline_info;
int lt_par_cline(lt_par par, lt_cline_info* cline);
int lt_par_destroy(lt_par par);

/* ---- critical lines and caustics ---------------------------------------------------------------------
 * No reference counterpart: the reference parses a `cline` block (src/structure_hpc.hpp:563-578,
 * src/module_readParameters.cpp:1008-1095, compiled out) but computes no critical line.
 *
 * The det A = 0 contour of a device det-A plane (grid as lt_map_device: ny rows of pitch doubles)
 * by marching squares: per cell at most two segments (x0, y0, x1, y1), crossings linearly
 * interpolated on the cell edges (cells sharing an edge share the point exactly), oriented with
 * det A > 0 on the left, saddles joined on the positive side when the corner sum is positive,
 * written to `segs` (device, 4 doubles per segment) in row-major cell order. kinds (device,
 * optional, needs the kappa plane): 0 tangential (1 - kappa > 0 on the segment), 1 radial.
 * *nsegments = number of segments of the whole contour; segs == NULL only counts. With more
 * segments than max_segments the first max_segments are written and LT_ERR_OVERFLOW is returned.
 * Waits for the stream. */
int lt_critic_device(const double* det, const double* kappa, const lt_grid_desc* grid, double* segs, int* kinds,
                     long long max_segments, long long* nsegments, void* stream);
/* The same contour without allocating: no counterpart in the reference. For callers that cannot
 * call hipMalloc (code under HIP graph capture, tight serving loops, harnesses that forbid device
 * allocation). The per-row counts and offsets live in a fixed scratch array of the library's code
 * object (about 192 KB on each GPU the library runs on, held until the process ends), so grids are
 * limited to ny - 1 <= LT_CRITIC_NOALLOC_MAX_ROWS cell rows (LT_ERR_CAPACITY beyond). Same
 * arguments, results and status as lt_critic_device. Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (several host threads or streams) race. */
#define LT_CRITIC_NOALLOC_MAX_ROWS 16384
int lt_critic_device_noalloc(const double* det, const double* kappa, const lt_grid_desc* grid, double* segs,
                             int* kinds, long long max_segments, long long* nsegments, void* stream);
/* Critical lines of a lens for the source plane z_source on an nx x ny grid over `frame` (det A
 * and kappa from the fields map in `precision`, then lt_critic_device), and optionally the
 * caustics: each segment end theta mapped to beta = theta - (D_ls/D_s) alpha(theta), alpha the
 * FP64 deflection at that point (4 doubles per segment, like `critic`). Host arrays; any of
 * critic, caustic, kinds may be NULL. A model with external-shear (type 14) halos has no deflection
 * kernel: critic, kinds and *nsegments are filled, caustic is set to NaN and the call returns
 * LT_ERR_TYPE. */
int lt_critic_lines(const lt_cosmo_param* cosmo, const lt_grid_param* frame, const lt_potential_soa* lens, int nhalos,
                    double z_lens, double z_source, int nx, int ny, int precision, double* critic, double* caustic,
                    int* kinds, long long max_segments, long long* nsegments);

/* ---- FITS maps -------------------------------------------------------------------------------------
 * Writes a host map (ny rows of nx doubles) as a FITS primary image of float32 values with the
 * frame's world coordinates: absolute == 0 gives module_writeFits_Image's linear arcsec axes
 * (pixel 1 at xmin / ymin, step (max - min) / (n - 1)), absolute != 0
 * module_writeFits_ImageAbsoluteCoordinates' RA---TAN / DEC--TAN axes in degrees around (ra, dec)
 * (src/module_writeFits.cpp:113-335; module_writeFits picks absolute unless ra == dec == 0).
 * Overwrites `path`. nx, ny >= 2. */
int lt_fits_write(const char* path, const double* map, int nx, int ny, double xmin, double xmax, double ymin,
                  double ymax, int absolute, double ra, double dec);

/* ---- point evaluations, delensing and chi^2 ------------------------------------------------------ */

/* Gradient at npts points (host arrays). The batched counterpart of the reference's point evaluator
 * module_potentialDerivatives_totalGradient_SOA (src/gradient.cpp:899-930). */
int lt_point_gradients(double* grad_x, double* grad_y, const double* px, const double* py, int npts,
                       const lt_potential_soa* lens, int nhalos);

/* Image-plane triangles of an nbgridcells^2 gradient grid mapped to the source plane, images found
 * for each source. Replaces delense_barycenter_GPU (src/delense_GPU.cuh:27,
 * src/delense_GPU.cu:227-350): image_pos has nsets*LT_MAXIMPERSOURCE slots, locimagesfound nsets
 * entries, numimg is incremented by the number of hits. Images come in the reference CPU path's
 * single-thread scan order (row by row, upper triangle before lower). A source with more than
 * LT_MAXIMPERSOURCE images keeps the first ones and the call returns LT_ERR_OVERFLOW (the
 * reference writes past its buffer). gx, gy: host gradient planes, sources[k].center and .dr used. */
int lt_delense_barycenter(lt_point* image_pos, int* locimagesfound, int* numimg, int nsets, int nbgridcells,
                          const lt_grid_param* frame, const lt_galaxy* sources, const double* gx,
                          const double* gy);

/* chi^2 of one lens model against nsets families of observed images (nimages[k] each, images in
 * family order, .center and .dr used): gradient grid, source barycentres, delensing and matching as
 * mychi_bruteforce_SOA_GPU_grid_gradient_barycentersource (src/chi_barycenter_GPU.cuh:13,
 * src/chi_CPU_barycenter.cpp:58-287) on one rank; *error receives the status. */
int lt_chi_barycentersource(double* chi, int* error, int nbgridcells, int nsets, int nhalos,
                            const lt_potential_soa* lens, const lt_grid_param* frame, const int* nimages,
                            const lt_galaxy* images);

/* chi^2 of found images against observed families, the matching and sum of chi_computation
 * (src/chi_computation.hpp:3-4, src/chi_computation.cpp:10-195) on the GPU: *chi and *images_found are
 * incremented as the reference does. imagesposition has nsets*LT_MAXIMPERSOURCE slots. */
int lt_chi_computation(double* chi, int* images_found, const int* numimagesfound, const lt_point* imagesposition,
                       const int* nimages, const lt_galaxy* images, int nsets);

/* The same for a batch of lens-model samples sharing halo types (e.g. MCMC samples): created once
 * for the observed families and the delensing grid, then run for nsamples SOA models at a time on
 * `stream`. chi, found (images found) and status are host arrays of nsamples; the call returns when
 * they are filled. No reference counterpart (the reference loops over samples on the host). */
typedef struct lt_chi_batch_s* lt_chi_batch;
int lt_chi_batch_create(lt_chi_batch* out, int nsamples, int nhalos, const int* types, int nbgridcells,
                        const lt_grid_param* frame, int nsets, const int* nimages, const lt_galaxy* images);
int lt_chi_batch_run(lt_chi_batch batch, const lt_potential_soa* lenses, double* chi, int* found, int* status,
                     void* stream);
/* images found for one sample of the last run (nsets*LT_MAXIMPERSOURCE slots, zero past the count) */
int lt_chi_batch_images(lt_chi_batch batch, int sample, lt_point* image_pos, int* locimagesfound);
int lt_chi_batch_destroy(lt_chi_batch batch);

/* x[i] *= s for n doubles in device memory (the reference's in-place potential and displacement
 * scalings, src/grid_map_pot_GPU.cu:201-250, src/grid_map_dpl_GPU.cu:194-208). */
int lt_scale_device(double* x, double s, long long n, void* stream);

/* The map epilogue alone, on Hessian planes a, b, c already in device memory (n pixels), for kinds
 * LT_MAP_AMPLI, LT_MAP_SHEAR, LT_MAP_MASS (one output plane out0) and LT_MAP_FIELDS (out0..out4).
 * This is what the reference's map functions (amplif_k_grid_CPU_GPU, shear_k..., mass_k...,
 * src/grid_map_ampli_GPU.cu:305-547) do after their separate Hessian launch. */
int lt_map_epilogue_device(int kind, int mode, double* out0, double* out1, double* out2, double* out3,
                           double* out4, const double* a, const double* b, const double* c, long long n,
                           double r, double mult, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* LENSTOOL_HIP_LT_CAPI_H */
