/*
 * C ABI of the MI355X PETL projector library (libpetl.so).
 *
 * Part 1 keeps every export of the reference's src/petl_c_interface.h:14-97 with the same name, C
 * signature and bool conventions, so the reference's ctypes bindings (src/petl_server.py) load this
 * library unchanged.  Each declaration cites the reference line it replaces.  Exports of subsystems
 * this library does not provide (voxelisation, scatter simulation, bad-pixel correction, TV denoising
 * and diffusion) are present so symbol lookup succeeds, but they print an error and return
 * false / 0.
 *
 * Part 2 adds entry points the reference does not have: device-resident projection on caller-owned
 * device buffers (torch tensors), back projector mode selection, list-mode binning of event arrays, and
 * device-resident MLEM/OSEM and RWLS reconstructions.  Host-pointer calls are synchronous, like the
 * reference's.
 *
 * Threading: like the reference, calls act on the model selected by set_model / create_new_model,
 * which is process-global state; do not call concurrently from several threads.
 */
#ifndef PETL_C_INTERFACE_H
#define PETL_C_INTERFACE_H

#ifdef __cplusplus
#include <cstdint>
#define PETL_EXTERN extern "C"
#else
#include <stdbool.h>
#include <stdint.h>
#define PETL_EXTERN
#endif

/* ------------------------------------------------------------------------------------------------
 * Part 1: reference ABI (src/petl_c_interface.h)
 * ---------------------------------------------------------------------------------------------- */
PETL_EXTERN bool set_model(int);                 /* src/petl_c_interface.h:14 */
PETL_EXTERN int create_new_model();              /* :15 */
PETL_EXTERN void about();                        /* :16 */
PETL_EXTERN bool copy_parameters(int);           /* :18 */
PETL_EXTERN void clearAll();                     /* :20 */
PETL_EXTERN void print_parameters();             /* :22 */
PETL_EXTERN bool volume_defined();               /* :23 */
PETL_EXTERN bool geometry_defined();             /* :24 */

PETL_EXTERN int get_numX();                      /* :26 */
PETL_EXTERN bool set_numX(int);                  /* :27 */
PETL_EXTERN int get_numY();                      /* :28 */
PETL_EXTERN bool set_numY(int);                  /* :29 */
PETL_EXTERN int get_numZ();                      /* :30 */
PETL_EXTERN bool set_numZ(int);                  /* :31 */
PETL_EXTERN float get_voxelWidth();              /* :33 */
PETL_EXTERN bool set_voxelWidth(float);          /* :34 */
PETL_EXTERN float get_voxelHeight();             /* :35 */
PETL_EXTERN bool set_voxelHeight(float);         /* :36 */
PETL_EXTERN float get_offsetX();                 /* :38 */
PETL_EXTERN bool set_offsetX(float);             /* :39  (rejects values <= 0, as the reference) */
PETL_EXTERN float get_offsetY();                 /* :40 */
PETL_EXTERN bool set_offsetY(float);             /* :41 */
PETL_EXTERN float get_offsetZ();                 /* :42 */
PETL_EXTERN bool set_offsetZ(float);             /* :43 */

PETL_EXTERN int get_numPlanograms();                                        /* :45 */
PETL_EXTERN bool get_planogramSize(int which, int* shape);                  /* :46  shape = {N_v1,N_v0,N_u1,N_u0} */
/* psi in degrees, full panel separation D, full length L, full height H (the reference halves them) */
PETL_EXTERN bool add_planogram(float psi, float R, float L, float H, float v_m0, float v_m1, float T); /* :48 */
PETL_EXTERN bool remove_planogram(int);                                     /* :49 */
PETL_EXTERN bool keep_only_planogram(int);                                  /* :50 */
PETL_EXTERN bool set_default_volume(float scale);                           /* :52 */
PETL_EXTERN bool set_volume(int numX, int numY, int numZ, float voxelWidth, float voxelHeight, float offsetX,
                            float offsetY, float offsetZ);                  /* :53 */

/* Forward projection: g[i] = A_i f for every planogram i, overwriting all bins (masked bins -> 0).
 * g: array of host pointers to C-contiguous float32 [N_v1][N_v0][N_u1][N_u0]; f: [numZ][numY][numX]. */
PETL_EXTERN bool project(float** g, float* f);                              /* :59 */
/* Back projection: f = sum_i A_i^T g[i], overwriting f.  Semantics follow the selected back projector
 * mode (set_backprojector_mode), PETL (reference) mode by default. */
PETL_EXTERN bool backproject(float** g, float* f);                          /* :60 */

/* elementwise host operations with the reference's zero conventions (src/data_cube.cpp:97-167) */
PETL_EXTERN bool multiply3D(float* x, float* y, int N_1, int N_2, int N_3);               /* :75 */
PETL_EXTERN bool multiply4D(float* x, float* y, int N_1, int N_2, int N_3, int N_4);      /* :76 */
PETL_EXTERN bool divide3D(float* x, float* y, int N_1, int N_2, int N_3);                 /* :78  y==0 -> 0 */
PETL_EXTERN bool divide4D(float* x, float* y, int N_1, int N_2, int N_3, int N_4);        /* :79 */
PETL_EXTERN bool rdivide3D(float* num, float* denom, int N_1, int N_2, int N_3);          /* :81  denom = num/denom */
PETL_EXTERN bool rdivide4D(float* num, float* denom, int N_1, int N_2, int N_3, int N_4); /* :82 */
PETL_EXTERN bool reciprocal3D(float* x, int N_1, int N_2, int N_3, float divide_by_zero_value);          /* :84 */
PETL_EXTERN bool reciprocal4D(float* x, int N_1, int N_2, int N_3, int N_4, float divide_by_zero_value); /* :85 */
PETL_EXTERN float inner_product3D(float* x, float* y, float* w, int N_1, int N_2, int N_3);             /* :87 */
PETL_EXTERN float inner_product4D(float* x, float* y, float* w, int N_1, int N_2, int N_3, int N_4);    /* :88 */

/* prior gradients (3D arrays [N_1][N_2][N_3]) */
PETL_EXTERN bool TV_gradient(float* f, float* Df, int N_1, int N_2, int N_3, float delta, float beta, float p,
                             bool data_on_cpu);                                                   /* :92 */
PETL_EXTERN bool relativeDifferences_gradient(float* f, float* Df, int N_1, int N_2, int N_3, float delta,
                                              float beta);                                        /* :96 */
/* p-Huber TV cost and quadratic form <d, R''(f) d> (the RWLS step size); -1 for a null input.  With
 * data_on_cpu false, f and d are device arrays. */
PETL_EXTERN float TV_cost(float* f, int N_1, int N_2, int N_3, float delta, float beta, float p,
                          bool data_on_cpu);                                               /* :91 */
PETL_EXTERN float TV_quadForm(float* f, float* d, int N_1, int N_2, int N_3, float delta, float beta, float p,
                              bool data_on_cpu);                                           /* :93 */
/* Gaussian smoothing of a host array in place: numDims 3 all axes, 2 the last two, 1 the first. */
PETL_EXTERN bool gaussian_filter(float* f, int N_1, int N_2, int N_3, float FWHM, int numDims); /* :97 */

/* analytic reconstruction: PFDR rebins host planograms to one v1 row each (g_reb: [N_v0][N_u1][N_u0]) and
 * leaves the model's planograms at N_v1 = 1, v1_0 = 0; FBP = PFDR (on a copy of the geometry, when N_v1 > 1)
 * + ramp filter + planogram weight + back projection into f (already rebinned data are filtered in place). */
PETL_EXTERN bool FBP(float** g, float* f);                                                 /* :61 */
PETL_EXTERN bool PFDR(float** g, float** g_reb);                                           /* :62 */

/* response model: geometric solids (type 0 ellipsoid, 1 parallelepiped, 2-4 cylinder x/y/z, 5-7 cone x/y/z;
 * c, r: 3 floats; A: 3x3 row-major rotation or NULL; clip: 3 floats or NULL), a ray-tracing phantom and the
 * scintillator modules; ray_trace / stopping_power fill host planograms (oversampling: odd, 1..11). */
PETL_EXTERN bool add_object(int type, float* c, float* r, float val, float* A, float* clip); /* :64 */
PETL_EXTERN bool add_module(int type, float* c, float* r, float val, float* A, float* clip); /* :65 */
PETL_EXTERN bool clear_phantom();                                                          /* :66 */
PETL_EXTERN bool clear_modules();                                                          /* :67 */
PETL_EXTERN bool scale_phantom(float, float, float);                                       /* :68 */
PETL_EXTERN bool ray_trace(float** g, int oversampling);                                   /* :70 */
PETL_EXTERN bool stopping_power(float** g, int oversampling);                              /* :71 */
PETL_EXTERN bool set_solid_angle_correction(float** g, bool do_inverse);                   /* :72 */

/* exports outside the projector / MLEM path: present, return false / 0 with a message */

PETL_EXTERN bool simulate_scatter(float** g, float* mu);                                   /* :57 */
PETL_EXTERN bool voxelize(float* f, int oversampling);                                     /* :69 */
PETL_EXTERN bool apply_corrections(float** g, float** r, float threshold);                 /* :73 */
PETL_EXTERN bool diffuse(float* f, int N_1, int N_2, int N_3, float delta, float p, int numIter,
                         bool data_on_cpu);                                                /* :94 */
PETL_EXTERN bool TV_denoise(float* f, int N_1, int N_2, int N_3, float delta, float beta, float p, int numIter,
                            bool doMean, bool data_on_cpu);                                /* :95 */

/* ------------------------------------------------------------------------------------------------
 * Part 2: additions (no counterpart in the reference)
 * ---------------------------------------------------------------------------------------------- */
/* Back projector semantics: 0 = PETL (the reference's backproject_planogram: voxel-centre u1 test, no u0
 * mask; the default), 1 = matched (exact transpose of project, needed for <Ax,y> = <x,A^T y> and for
 * gradients). */
PETL_EXTERN bool set_backprojector_mode(int mode);
PETL_EXTERN int get_backprojector_mode();
/* Kernel family: 0 = automatic (fast separable kernels when the geometry allows, else generic),
 * 1 = generic kernels only. */
PETL_EXTERN bool set_projector_path(int path);
PETL_EXTERN int get_projector_path();
/* 1 when planogram `which` runs on the fast separable kernels under the current settings. */
PETL_EXTERN int planogram_uses_fast_path(int which);

/* Device-resident forward / back projection on caller-owned device buffers (e.g. torch tensors).
 * d_g: host array of get_numPlanograms() device pointers; d_f: device pointer; stream: the hipStream_t
 * to enqueue on (NULL = the default stream).  Work is enqueued on the stream; the call returns after enqueueing
 * unless sync is true.  backproject_gpu overwrites d_f. */
PETL_EXTERN bool project_gpu(float** d_g, float* d_f, void* stream, bool sync);
PETL_EXTERN bool backproject_gpu(float** d_g, float* d_f, void* stream, bool sync);

/* List-mode binning (src/petl_c_interface.h:55, src/petl.cpp:65-79, src/binning.cpp).  bin() zeroes the
 * planograms and bins the coincidences of the file with the reference's defaults: millimetres, list-mode
 * z origin at the chest wall (u1 -= H), nearest v bins and clamped bilinear u bins.  ".lst" files hold
 * records of two {float pos[4]; uint64 time} events; any other name is read in the acquisition layout
 * (int64 count; per end point int16 interactions, float pos[3], float energy, 18 bytes per further
 * interaction).  Returns false when the file cannot be opened.  The binned values are bit-identical to
 * the reference's: contributions are added to each bin in event order. */
PETL_EXTERN bool bin(float** g, char* file_name);                                          /* :55 */
/* Bins n coincidences given as arrays of x, y, z (mm) of the blue and pink end points, ADDING to the
 * planograms (no reset): host arrays and planograms / device arrays and planograms on a stream.
 * nearest: 1 = nearest u bins, 0 = clamped bilinear (the reference default); chest_wall: 1 = z origin at
 * the chest wall (the reference default), 0 = at the centre of the field of view.  accepted (optional,
 * host): coincidences each planogram accepted. */
PETL_EXTERN bool bin_events(const float* blue, const float* pink, long long n, float** g, int nearest,
                            int chest_wall, long long* accepted);
PETL_EXTERN bool bin_events_gpu(const float* d_blue, const float* d_pink, long long n, float** d_g, int nearest,
                                int chest_wall, void* stream, bool sync);

/* Storage formats of device volumes / planograms for the *_ex and *_storage calls: 0 = fp32, 1 = bfloat16,
 * 2 = IEEE binary16 with a per-tensor scale (value = stored * scale).  Arithmetic is fp32 throughout;
 * 16-bit formats run on the fast kernels only (psi 0 or 90 degrees, voxel pitch = crystal pitch). */
#define PETL_STORE_F32 0
#define PETL_STORE_BF16 1
#define PETL_STORE_F16 2
/* project_gpu with the volume d_f in `volume_storage` and the planograms d_g in `planogram_storage` (equal, or
 * an fp32 volume with 16-bit planograms); f_scale, g_scale: binary16 scales, ignored otherwise. */
PETL_EXTERN bool project_gpu_ex(void** d_g, const void* d_f, int volume_storage, int planogram_storage,
                                float f_scale, float g_scale, void* stream, bool sync);
/* backproject_gpu with the planograms in `storage`; the volume d_f is fp32 (overwritten). */
PETL_EXTERN bool backproject_gpu_ex(void** d_g, float* d_f, int storage, float g_scale, void* stream, bool sync);
/* For callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid it); no reference counterpart.  prepare_gpu builds the current model's kernel tables
 * (it allocates; call it once beforehand).  backproject_gpu_noalloc is backproject_gpu with the fast back
 * projector's workspace in a fixed 128 MiB device array that each GPU holds until the process ends (set up on
 * the first call on a device): it allocates nothing, and returns false when the tables are not built or the
 * geometry's workspace (N_v1 x N_u1 x in-plane voxels x 4 bytes for the largest planogram) exceeds 128 MiB.
 * Every call on one GPU shares that array: calls that overlap on one GPU race. */
PETL_EXTERN bool prepare_gpu(void);
PETL_EXTERN bool backproject_gpu_noalloc(float** d_g, float* d_f, void* stream, bool sync);
/* Batches for differentiable callers (no reference counterpart; the reference projects one volume per
 * call): d_f holds `batch` contiguous fp32 volumes and d_g[i] `batch` contiguous copies of planogram i.
 * project_gpu_batch overwrites every planogram; backproject_gpu_batch overwrites every volume with the sum
 * of its planograms' back projections in `mode` (0 PETL, 1 matched = exact transpose of the forward,
 * -1 = the model's mode), without changing the model's mode.  The fast kernels take a whole batch in one
 * launch per planogram (the back projector's workspace holds up to 4 GiB of batch elements). */
PETL_EXTERN bool project_gpu_batch(float** d_g, float* d_f, int batch, void* stream, bool sync);
PETL_EXTERN bool backproject_gpu_batch(float** d_g, float* d_f, int batch, int mode, void* stream, bool sync);

/* Device-resident MLEM / OSEM (the loop of src/petl_server.py:1534-1606 with every step on the GPU).
 * g: host planograms (data); f: host volume, used as the initial estimate (all zeros -> ones, negatives
 * clipped) and overwritten with the result.  prior: 0 none, 1 relative differences (OSL, as the
 * reference's beta > 0 branch), 2 p-Huber TV (OSL).  mask / response: optional host planogram lists
 * (NULL for none), used as PETLserver.MLEM uses them.  numSubsets > 1 runs OSEM over interleaved v1
 * subsets.  loglik: optional host array of numIter doubles receiving sum(y log Af - Af) over bins with
 * Af > 0, accumulated in FP64 over the iteration's forward projections (for OSEM: each subset's views,
 * projected from the estimate in force when that subset is processed). */
PETL_EXTERN bool MLEM_gpu(float** g, float* f, int numIter, float delta, float beta, int prior, float p,
                          int numSubsets, float** mask, float** response, double* loglik);
/* MLEM_gpu with the planograms (data and ratio) held in `planogram_storage` and the image in
 * `image_storage` (equal, or an fp32 image) on the device; the sensitivities and back projections stay
 * fp32.  binary16 scales are chosen per tensor (data: from its largest count; image: from
 * sum(data) / sum(sensitivity)).  No mask, response or log-likelihood. */
PETL_EXTERN bool MLEM_gpu_storage(float** g, float* f, int numIter, float delta, float beta, int prior, float p,
                                  int numSubsets, int planogram_storage, int image_storage);

/* ray_trace (stopping = 0) or stopping_power (stopping = 1) into device planograms d_g on a stream
 * (NULL = the default stream); without modules, stopping_power writes 1. */
PETL_EXTERN bool ray_trace_gpu(float** d_g, int oversampling, int stopping, void* stream, bool sync);

/* Device-resident RWLS (PETLserver.RWLS, src/petl_server.py:1691-1877): preconditioned conjugate gradients
 * on 0.5 (Pf - g)^T W (Pf - g) + TV(f) (p-Huber TV with p = 1, delta, beta; none when beta <= 0), every
 * array on the GPU.  g: host planograms; f: host image, the start (negatives clipped when constrained,
 * all zeros -> Pf = 0) and overwritten with the result.  weight_mode: 0 = the reference's default
 * 1 / max(gaussian_filter(g, 3), 7), 1 = the host planogram weights W, 2 = the scalar w_scalar (gradient
 * unweighted, <Pd, Pd> scaled, as the reference treats a scalar).  preconditioner: 0 none, 1 SQS.
 * response: optional host planograms the projectors include (the reference's
 * include_response_with_projectors = True).  cost (optional, numIter doubles): 0.5 <r, r>_W + TV cost after
 * each iteration run; iterations (optional): iterations run before a step <= 0 ended the loop. */
PETL_EXTERN bool RWLS_gpu(float** g, float* f, int numIter, float delta, float beta, int weight_mode, float** W,
                          float w_scalar, int preconditioner, bool nonnegativityConstraint, float** response,
                          double* cost, int* iterations);

/* ------------------------------------------------------------------------------------------------
 * Part 3: reconstruction service (no counterpart in the reference).  Continuous batching of MLEM/OSEM
 * requests of different panel and volume sizes on one GPU: a submitted request waits until a slot is
 * free and its device footprint fits the HBM budget, then gets its own device buffers and stream; each
 * MLEM iteration is captured once as a hipGraph and replayed.  Requests that arrive while others run
 * are admitted as soon as slots and memory free up, and their work overlaps on the device.  A request
 * gives the same image as MLEM_gpu on the same inputs, alone or in any mix.
 * ---------------------------------------------------------------------------------------------- */
#define PETL_MAX_PLANOGRAMS 8
typedef struct petl_recon_request {
    int num_planograms;                /* 1 .. PETL_MAX_PLANOGRAMS */
    float psi[PETL_MAX_PLANOGRAMS];    /* panel angles, degrees */
    float D, L, H, T;                  /* add_planogram geometry (mm), shared by the planograms */
    float v_m0, v_m1;                  /* add_planogram's v_m0, v_m1 */
    int numX, numY, numZ;              /* volume */
    float voxelWidth, voxelHeight;
    float** y;                         /* host data, one planogram array each; kept alive until wait returns */
    int numIter;
    int numSubsets;                    /* OSEM subsets (1 = MLEM) */
    int prior;                         /* 0 none, 1 relative differences, 2 p-Huber TV (one-step-late) */
    float delta, beta, p;
    int storage;                       /* planograms on the device: PETL_STORE_F32 / _BF16 / _F16 */
    int image_storage;                 /* image on the device: equal to storage, or PETL_STORE_F32 */
} petl_recon_request;

/* hbm_budget_bytes: device memory the service may hold for requests; max_slots: requests on the device
 * at once.  Returns NULL on failure. */
PETL_EXTERN void* petl_service_create(double hbm_budget_bytes, int max_slots);
/* petl_service_create with flags: PETL_SERVICE_BATCHED runs the requests that allow it (fp32 MLEM, one
 * subset, at most two planograms on the fast kernels with up to 255 u0 bins and 128-voxel in-plane and
 * slice axes, PETL-mode back projection) in batched iterations: one launch per phase of an iteration
 * (forward with the ratio, per planogram the two back projector stages, the update) advances every such
 * request on the device by one iteration, whatever its size, its grid covering all their work items; the
 * others run their own captured iteration as without the flag.
 * No reference counterpart (the reference has no service). */
#define PETL_SERVICE_BATCHED 1
PETL_EXTERN void* petl_service_create_ex(double hbm_budget_bytes, int max_slots, int flags);
/* Queues a request (the struct is copied; the data arrays are read when the request is admitted).
 * Returns its id (>= 0), or -1 when the request is invalid or can never fit the budget. */
PETL_EXTERN int petl_service_submit(void* service, const petl_recon_request* request);
/* Device bytes a request would hold (its admission footprint), or 0 when it is invalid. */
PETL_EXTERN double petl_service_request_bytes(const petl_recon_request* request);
/* Checks the admission footprint: out[0] = the bytes petl_service_request_bytes counts, out[1] = the bytes
 * the request's device buffers hold once allocated (the buffers are allocated and freed again).  Added for
 * tests; no reference counterpart.  Returns false when the request is invalid or allocation fails. */
PETL_EXTERN bool petl_service_request_footprint(const petl_recon_request* request, double* out);
/* Blocks until request `id` is done and copies its image (numZ x numY x numX) to f.  times (optional,
 * 3 doubles): seconds from submission to admission, from admission to completion, and in total.
 * Returns false for an unknown id or a failed request. */
PETL_EXTERN bool petl_service_wait(void* service, int id, float* f, double* times);
/* stats (15 doubles): completed, failed, peak requests on the device, peak reserved bytes, budget,
 * requests currently queued, the host seconds spent so far admitting requests (allocation, data upload
 * with sensitivities, graph capture) and freeing resources, how many requests reused the pooled
 * resources (model, buffers, captured iteration) of a finished request with the same geometry and MLEM
 * parameters, and how many of those captured the iteration again because their binary16 data or image
 * scale differed, how many requests ran in the batched iterations, how many batched iterations were
 * launched, and the most requests one batched iteration advanced.  Pooled memory counts against the budget until a new request needs it. */
PETL_EXTERN bool petl_service_stats(void* service, double* stats);
/* Finishes every queued request, then frees the service. */
PETL_EXTERN void petl_service_destroy(void* service);

#endif /* PETL_C_INTERFACE_H */
