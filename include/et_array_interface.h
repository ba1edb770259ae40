/*
 * et_array_interface.h -- C ABI of lib_et_array_interface.so, the MI355X (gfx950)
 * SPECT projector.
 *
 * Part 1 keeps NiftyRec's array interface so existing callers (the ctypes
 * package Python/NiftyRec/NiftyRec/NiftyRec.py and the MATLAB mex files) load
 * this library unchanged.  Each declaration cites the reference prototype it
 * replaces (paths relative to the NiftyRec source tree).
 *
 * Conventions shared by every entry point:
 *   - all arrays are float32 with x fastest: idx = x + nx*(y + ny*z);
 *   - activity / attenuation [N, m, N], sinogram [N, m, n_cameras] (camera
 *     slowest), psf [h, h, N] (one plane per depth), cameras [n_cameras x {1,3,6}]
 *     stored cameras[axis*n_cameras + cam] in radians (one column: rotation
 *     about axis 1); size arrays {0,0,0} mark an absent input;
 *   - the caller owns every buffer; outputs are fully overwritten;
 *   - return value 0 on success, otherwise a niftyrec status code (below);
 *     the library never terminates the process.
 *   - computation always runs on the HIP device selected by et_array_set_gpu
 *     (default 0); the GPU / enable_gpu flags are accepted for compatibility.
 *     Without a usable device the call returns niftyrec_error_nogpuinstalled.
 *
 * Part 2 adds device-pointer entry points (plans) for callers that keep data
 * in GPU memory (the torch binding, the OSEM loop, the benchmark).
 */
#ifndef ET_ARRAY_INTERFACE_H
#define ET_ARRAY_INTERFACE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes, reference emission-lib/_et_common.h:40-54. */
#define niftyrec_success 0
#define niftyrec_error_unspecified 1
#define niftyrec_error_parameters 2
#define niftyrec_error_kernel 3
#define niftyrec_error_allocgpu 4
#define niftyrec_error_alloccpu 5
#define niftyrec_error_transfergpu 6
#define niftyrec_error_nogpubuilt 7
#define niftyrec_error_nogpuinstalled 8
#define STATUS_SUCCESS 0
#define STATUS_IO_ERROR 101
#define STATUS_INITIALISATION_ERROR 103
#define STATUS_PARAMETER_ERROR 104
#define STATUS_UNHANDLED_ERROR 105
#define STATUS_MEMORY_ERROR 106

/* ---- Part 1: NiftyRec array interface ------------------------------------- */

/* SimpleWrap status helpers, emission-lib/_et_array_interface.h:22-27. */
int status_success(int *status);
int status_io_error(int *status);
int status_initialisation_error(int *status);
int status_parameter_error(int *status);
int status_unhandled_error(int *status);
int echo(int *in, int *out);

/* Message for a status code; thread-local buffer.
 * Replaces niftyrec_error_msg, emission-lib/_et_common.h:56 (_et_common.cpp:269-306). */
char *niftyrec_error_msg(int status);

/* Device management, emission-lib/_et_array_interface.h:36-41
 * (_et_array_interface.cpp:230-279).  info per device: {id, gflops, CUs, clock kHz, MB}. */
int et_array_list_gpus(int *gpu_count, int *gpus_info_array);
int et_array_set_gpu(int id);
int et_array_set_gpu_pointer(int *id);
int et_array_isinstalled(void);
int et_array_reset_gpu(void);
/* The kernels accept any volume size: both return values report that
 * (the reference required multiples of 64, emission-lib/_et.cpp:27-36). */
unsigned short int et_array_is_block_multiple(unsigned short int size);
unsigned short int et_array_get_block_size(void);

/* Projection, emission-lib/_et_array_interface.h:44 (_et_array_interface.cpp:337-541).
 * Activity absent ({0,0,0}) -> background-only projection through the attenuation;
 * as in the reference, activity_size is then overwritten with attenuation_size. */
int SPECT_project_parallelholes(float *activity, int *activity_size, float *sinogram,
                                int *sinogram_size, float *cameras, int *cameras_size,
                                float *psf, int *psf_size, float *attenuation,
                                int *attenuation_size, float *Background,
                                float *Background_attenuation, int *Gpu,
                                int *Truncate_negative_values);

/* Backprojection (inverse-rotation pull, the reference operator),
 * emission-lib/_et_array_interface.h:45 (_et_array_interface.cpp:547-710). */
int SPECT_backproject_parallelholes(float *sino, int *sino_size, float *bkpr, int *bkpr_size,
                                    float *cameras, int *cameras_size, float *psf,
                                    int *psf_size, float *attenuation, int *attenuation_size,
                                    float *Background, float *Background_attenuation, int *Gpu,
                                    int *Truncate_negative_values);

/* By-value spellings bound by the Python package,
 * Python/NiftyRec/NiftyRec/NiftyRec.py:83-84 and :145-146. */
int et_array_project(float *activity, int *activity_size, float *sinogram, int *sinogram_size,
                     float *cameras, int *cameras_size, float *psf, int *psf_size,
                     float *attenuation, int *attenuation_size, float background,
                     float background_attenuation, int GPU, int truncate_negative_values);
int et_array_backproject(float *sino, int *sino_size, float *bkpr, int *bkpr_size,
                         float *cameras, int *cameras_size, float *psf, int *psf_size,
                         float *attenuation, int *attenuation_size, float background,
                         float background_attenuation, int GPU, int truncate_negative_values);

/* Partial projection: sinogram plus, for every camera, the partial line
 * integrals (what each ray has collected beyond each plane), in the image
 * frame when do_rotate_partial, else in the rotated frame; partialsum is
 * [N, m, N, n_cameras].  Reference _et_array_interface.h:47,
 * _et_array_interface.cpp:1561-1787 (CUDA path _et.cpp:2106-2440; the
 * reference CPU path is a stub returning 1). */
int et_array_project_partial(float *activity, int *activity_size, float *sinogram,
                             int *sinogram_size, float *partialsum, int *partialsum_size,
                             float *cameras, int *cameras_size, float *psf, int *psf_size,
                             float *attenuation, int *attenuation_size, float background,
                             float *background_image, float background_attenuation, int GPU,
                             int truncate_negative_values, int do_rotate_partial);

/* Reference attenuation gradient, emission-lib/_et_array_interface.h:48
 * (_et_array_interface.cpp:714-892; CUDA implementation emission-lib/_et.cpp:1679-1924). */
int et_array_gradient_attenuation(float *sino, int *sino_size, float *activity,
                                  int *activity_size, float *gradient, int *gradient_size,
                                  float *cameras, int *cameras_size, float *psf, int *psf_size,
                                  float *attenuation, int *attenuation_size, float background,
                                  float background_attenuation, int GPU,
                                  int truncate_negative_values);

/* Standalone rotation resampler, emission-lib/_et_array_interface.h:32
 * (_et_array_interface.cpp:112-156); axis_order 0..5 = XYZ,XZY,YXZ,YZX,ZXY,ZYX. */
int et_array_rotate(float *image, int *size, float *rotated_image, float *angles,
                    float *centers, float background, int axis_order, int GPU);
/* Affine resampler, emission-lib/_et_array_interface.h:31 (_et_array_interface.cpp:56-108).
 * affine is a row-major 4x4 matrix mapping output voxel to input voxel (the
 * reference ignores it and resamples with a zero matrix, marked FIXME there). */
int et_array_affine(float *image, int *image_size, float *transformed_image,
                    int *transformed_image_size, float *affine, int *affine_size,
                    float background, int GPU);
/* Depth-dependent separable 2-D convolution of every z plane with the matching
 * psf plane, emission-lib/_et_array_interface.h:33 (_et_array_interface.cpp:161-225). */
int et_array_convolve(float *image, int *image_size, float *out, int *out_size, float *psf,
                      int *psf_size, int GPU);

/* Fisher information of a grid of points, emission-lib/_et_array_interface.h:50-51
 * (_et_array_interface.cpp:1135-1559; algorithm emission-lib/_et.cpp:531-746).
 * grid [N, m, N] holds 0 or the 1-based index of a point; fisher (and
 * fisher_prior, may be NULL) are [n, n] with n = fisher_size[0]:
 *   F[i][j] = sum over cameras and the overlap of the two points' PSF boxes on
 *             the detector of PSF_j * PSF_i * projection
 * (the projection itself, as in the reference, whose inversion by epsilon is
 * commented out; epsilon is unused).  The first form projects the activity
 * (truncated); the second takes the projection [N, m, n_cameras] (bkpr_size
 * gives the grid size).  GPU != 0 selects the reference CUDA prior rule (1
 * where |d| <= 3 on every axis, i != j), GPU == 0 its CPU rule (1 where
 * |d| <= 1, the diagonal included).  Status 1 for the reference's parameter
 * checks, 2 (ET_ERROR_BADGRID) for a repeated or out-of-range grid index,
 * niftyrec_error_parameters for a missing PSF (the reference dereferences
 * it), a PSF with fewer planes than the grid's depth, or (first form) a
 * non-square PSF. */
int et_array_fisher_grid(float *activity_ptr, int *activity_size, float *cameras, int *cameras_size,
                         float *psf, int *psf_size, float *grid_ptr, float *fisher_ptr,
                         float *fisher_prior_ptr, int *fisher_size, float *attenuation,
                         int *attenuation_size, float epsilon, float background,
                         float background_attenuation, int GPU);
int et_array_fisher_grid_projection(float *sinogram_ptr, int *sinogram_size, int *bkpr_size,
                                    float *cameras, int *cameras_size, float *psf, int *psf_size,
                                    float *grid_ptr, float *fisher_ptr, float *fisher_prior_ptr,
                                    int *fisher_size, float *attenuation, int *attenuation_size,
                                    float epsilon, float background, float background_attenuation,
                                    int GPU);

/* PSF helpers kept with the reference's (stub) behaviour,
 * emission-lib/_et_array_interface.h:53-54 (_et_array_interface.cpp:896-920). */
int et_array_calculate_size_psf(unsigned int *psf_size_x, unsigned int *psf_size_y,
                                float fwhm_pixels_dist0, float sensitivity0, float dist0,
                                float fwhm_pixels_dist1, float sensitivity1, float dist1);
int et_array_make_psf(float *psf_data, unsigned int psf_size_x, unsigned int psf_size_y,
                      float fwhm_pixels_dist0, float sensitivity0, float dist0,
                      float fwhm_pixels_dist1, float sensitivity1, float dist1,
                      unsigned int N_psf_planes);

/* Cameras and EM drivers, emission-lib/_et_array_interface.h:55-59
 * (_et_array_interface.cpp:923-1131). */
int et_array_make_cameras(float *cameras_data, float firstcamera_deg, float lastcamera_deg,
                          unsigned int n_cameras, unsigned int rotation_axis);
int et_array_mlem_step(float *activity_data, int size_x, int size_y, float *sinogram_data,
                       int n_cameras, float *cameras_data, int use_attenuation,
                       float *attenuation_data, int use_psf, float *psf_data,
                       unsigned int psf_size_x, unsigned int psf_size_y, float background,
                       float background_attenuation, float epsilon, int use_gpu);
int et_array_osem_step(float *activity_data, int size_x, int size_y, unsigned int subset_order,
                       float *sinogram_data, int n_cameras, float *cameras_data,
                       int use_attenuation, float *attenuation_data, int use_psf,
                       float *psf_data, unsigned int psf_size_x, unsigned int psf_size_y,
                       float background, float background_attenuation, float epsilon,
                       int use_gpu);
int et_array_mlem(float *activity_data, unsigned int size_x, unsigned int size_y,
                  float *sinogram_data, int n_cameras, float firstcamera, float lastcamera,
                  unsigned int rotation_axis, int iterations, int use_attenuation,
                  float *attenuation_data, int use_psf, float *psf_data,
                  unsigned int psf_size_x, unsigned int psf_size_y, float background,
                  float background_attenuation, float epsilon, int use_gpu);
int et_array_osem(float *activity_data, unsigned int size_x, unsigned int size_y,
                  unsigned int subset_order, float *sinogram_data, int n_cameras,
                  float firstcamera, float lastcamera, unsigned int rotation_axis,
                  int iterations, int use_attenuation, float *attenuation_data, int use_psf,
                  float *psf_data, unsigned int psf_size_x, unsigned int psf_size_y,
                  float background, float background_attenuation, float epsilon,
                  int use_gpu);
/* The drivers et_array_mlem and et_array_osem keep the reference's results
 * (activity reset to 1, cameras from et_array_make_cameras, et_array_mlem:
 * `iterations` MLEM updates; et_array_osem: `iterations` et_array_osem_step
 * updates on subsets drawn from rand()) but run device-resident: data,
 * attenuation and iterate are uploaded once, one plan serves every step and
 * the MLEM sensitivity is computed once.  et_array_osem draws all its
 * subsets from rand() before any device work (the GPU runtime's start-up may
 * call rand() itself); the reference CPU path draws the same sequence step by
 * step.  They return the first non-zero status (the reference adds the
 * statuses up).
 *
 * Deterministic ordered subsets (no counterpart in the reference; the
 * reconstruction-driver row of the reference's rec_spect tool,
 * apps/command_line_tools/rec_spect.cpp:553-555, calls et_array_osem):
 * `iterations` passes over n_subsets subsets, subset s holding cameras s,
 * s + n_subsets, s + 2 n_subsets, ... in that order; each subset update is
 * the reference MLEM update restricted to its cameras (the update of
 * et_plan_osem).  Other arguments as et_array_osem. */
int et_array_osem_ordered(float *activity_data, unsigned int size_x, unsigned int size_y,
                          unsigned int n_subsets, float *sinogram_data, int n_cameras,
                          float firstcamera, float lastcamera, unsigned int rotation_axis,
                          int iterations, int use_attenuation, float *attenuation_data,
                          int use_psf, float *psf_data, unsigned int psf_size_x,
                          unsigned int psf_size_y, float background,
                          float background_attenuation, float epsilon, int use_gpu);

/* ---- Part 2: device-resident plans (no counterpart in the reference) ------- */
/* A plan holds the geometry (cameras, centres, PSF, backgrounds) on the device
 * and its own workspaces.  Array arguments below are DEVICE pointers in the
 * layouts of Part 1; `stream` is a hipStream_t (NULL: the default stream).
 * Calls on one plan must not overlap; distinct plans are independent. */
typedef struct et_plan et_plan;

/* Volumes up to 512 voxels wide (nx; the projection kernels hold one thread
 * per rotated column); wider ones are refused with status 2 by the plan and
 * by the projection, backprojection, gradient and MLEM/OSEM entry points.
 * cameras as in Part 1 (host memory, camera_cols 1, 3 or 6); psf host memory
 * [psf_h, psf_h, nz] or NULL.  For camera_cols != 6 the rotation centres are
 * ((nx-1)/2, (ny-1)/2, (nz-1)/2). */
int et_plan_create(et_plan **plan, int nx, int ny, int nz, const float *cameras,
                   int n_cameras, int camera_cols, const float *psf, int psf_h,
                   float background, float background_attenuation);
int et_plan_destroy(et_plan *plan);
/* act or mu may be NULL (not both). */
int et_plan_project(et_plan *plan, const float *act, const float *mu, float *sino,
                    int truncate, void *stream);
/* et_plan_project without allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory during the call (code under HIP
 * graph capture, tight serving loops).  act and sino are required, mu may be
 * NULL; the plan must be axial with zero backgrounds.  The temporaries (the
 * padded copies of act and mu) live in a static device array of 32 MiB per
 * GPU, held until the process ends; plans whose copies exceed it return
 * niftyrec_error_parameters (each copy holds (nx+2)(nz+2) padded columns of
 * about ny + 16 floats: 128^3 and 160x96x160 fit with mu, 256^3 does not).
 * One depth chunk per view: the same kernels and results as et_plan_project
 * after et_plan_set_tuning(plan, 1, 0).  Calls share the array: calls that
 * overlap on one GPU race. */
int et_plan_project_noalloc(et_plan *plan, const float *act, const float *mu, float *sino,
                            int truncate, void *stream);
/* adjoint = 0: the reference backprojector; 1: exact transpose of et_plan_project. */
int et_plan_backproject(et_plan *plan, const float *sino, const float *mu, float *out,
                        int adjoint, int truncate, void *stream);
/* exact = 0: the reference attenuation gradient; 1: gradient of <sino, project(act, mu)>
 * with respect to mu (sino holds dL/dp). */
int et_plan_gradient_attenuation(et_plan *plan, const float *sino, const float *act,
                                 const float *mu, float *out, int exact, int truncate,
                                 void *stream);
/* Both gradients of <sino, project(act, mu)> (sino holds dL/dp): out_act as
 * et_plan_backproject with adjoint 1, out_mu as et_plan_gradient_attenuation
 * with exact 1, one truncate flag for both.  No counterpart in the reference:
 * the backward pass of a differentiable projection.  Right after
 * et_plan_project of the same act and mu, plans on the row-lane kernels run
 * both rotated-frame sweeps and one exact adjoint pull whose candidate weights
 * serve both outputs (rotated-frame volumes of all views twice: at most
 * 4 GiB); otherwise the two operations in turn.  Outputs must not overlap
 * each other or an input: an output pointer equal to an input's or to the
 * other output's is rejected (niftyrec_error_parameters); partial overlaps
 * are not detected. */
int et_plan_project_backward(et_plan *plan, const float *sino, const float *act,
                             const float *mu, float *out_act, float *out_mu, int truncate,
                             void *stream);
/* Ordered-subsets EM on the device: subset s holds views s, s+S, s+2S, ...;
 * each subset update is the reference MLEM update restricted to its views.
 * act is updated in place.  The sensitivity of every subset, the attenuation
 * factors e^{-A} of every view (4 bytes per voxel and view of device memory)
 * and the depth chunking (timed on the first call unless fixed with
 * et_plan_set_tuning) are cached in the plan, keyed on (subsets, eps, the
 * attenuation map).  By default the map is keyed on a fingerprint of its
 * content, computed on the device at every call (one pass over mu and one
 * stream synchronisation), so a map rewritten in place, or another map at the
 * same address, rebuilds the caches. */
int et_plan_osem(et_plan *plan, float *act, const float *sino, const float *mu, int subsets,
                 int iterations, float epsilon, void *stream);
/* check != 0 (default): key the OSEM caches on the attenuation map's content
 * fingerprint.  check == 0: key them on its address only; the caller then
 * calls et_plan_invalidate whenever the map's content changes (this saves the
 * fingerprint pass and the synchronisation per call). */
int et_plan_set_mu_check(et_plan *plan, int check);
/* Drops the cached OSEM data (rebuilt by the next et_plan_osem). */
int et_plan_invalidate(et_plan *plan);
/* HIP device the plan's buffers live on (the device selected when it was created). */
int et_plan_get_device(const et_plan *plan, int *device);
/* Tuning: depth chunks per ray (0 = automatic), views per backprojection batch (0 = automatic). */
int et_plan_set_tuning(et_plan *plan, int depth_chunks, int view_batch);
/* Depth chunks per ray of the cached OSEM data: projection and rotated-frame
 * backprojection sweeps (0 and 0 before the first et_plan_osem). */
int et_plan_get_tuning(const et_plan *plan, int *depth_chunks_project, int *depth_chunks_backproject);
/* Projector of the cached OSEM data: *tiles = the number of 64-column tiles of
 * the tiled row-lane projector, 1 for the full-width row-lane projector, 0 for
 * the slab kernels or before the first et_plan_osem.  The choice is timed
 * with the depth chunking where both apply (an extension: no counterpart in
 * the reference). */
int et_plan_get_projector(const et_plan *plan, int *tiles);
/* Kernel timing with HIP events on the plan's launches (category ids below):
 * et_plan_profile times every category (enable != 0) or none; _select times
 * the categories whose bits are set in `categories` (bit ET_PROF_*), so that
 * a caller can time one kernel with fewer event records in the stream.  Both
 * reset the accumulated figures. */
int et_plan_profile(et_plan *plan, int enable);
int et_plan_profile_select(et_plan *plan, unsigned int categories);
/* Time only every `every`-th launch of each selected category (1 = all; the
 * first launch after et_plan_profile/_select is timed): a caller measuring a
 * kernel's mean launch time over a long run puts fewer events in the stream. */
int et_plan_profile_sample(et_plan *plan, int every);
int et_plan_profile_read(et_plan *plan, int category, int *launches, double *total_ms,
                         double *bytes);
#define ET_PROF_PROJECT 0
#define ET_PROF_BPU 1
#define ET_PROF_PULL 2
#define ET_PROF_MUSUM 3
#define ET_PROF_OTHER 4
#define ET_PROF_GRAD 5
#define ET_PROF_COUNT 6

/* ---- Part 3: batches of plans (no counterpart in the reference) ----------- */
/* Many independent reconstructions of uneven sizes, angle counts and PSFs run
 * as one OSEM iteration of grouped launches (plans with the same PSF radius,
 * volume width class and attenuation presence share each launch), captured
 * once into a hipGraph and replayed per iteration.  Every plan must take the
 * fast path (axial cameras, PSF radius <= 10, zero backgrounds) and live on
 * the same device; plans keep their own caches and work buffers, and must not
 * be used concurrently with the batch. */
typedef struct et_batch et_batch;

int et_batch_create(et_batch **batch, et_plan *const *plans, int n_plans);
int et_batch_destroy(et_batch *batch);
/* act[i], sino[i] (and mu[i], or mu == NULL / mu[i] == NULL for none): DEVICE
 * pointers of plan i in the layouts of Part 1; host arrays of n_plans pointers.
 * Runs `iterations` OSEM iterations of `subsets` ordered subsets on every plan
 * (the update of et_plan_osem); use_graph = 0 issues the launches directly.
 * With stream == NULL the call orders after the null stream and returns when
 * done; otherwise it is asynchronous on `stream`. */
int et_batch_osem(et_batch *batch, float *const *act, const float *const *sino,
                  const float *const *mu, int subsets, int iterations, float epsilon,
                  int use_graph, void *stream);

/* Library version string. */
const char *et_library_version(void);

#ifdef __cplusplus
}
#endif
#endif
