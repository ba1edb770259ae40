/*
 * et_oracle.h -- CPU restatement of NiftyRec's SPECT projector path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (niftyrec_amd/,
 * lib_et_array_interface.so) links, loads or calls this code.  It is the
 * checker used by tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py.
 *
 * Every routine exists twice: *_f32 restates the reference CPU path in
 * single precision with the reference's order of operations (so it can be
 * compared bit-for-bit with the reference build in oracle/_ref), *_f64 is the
 * same algorithm in double precision and is the accuracy oracle.
 *
 * Conventions (reference emission-lib/_et_array_interface.cpp:448-493):
 *   volumes  [nx, ny, nz], x fastest: idx = x + nx*(y + ny*z)
 *   sinogram [nx, ny, V],  camera slowest
 *   cameras  cams[axis*V + v] in radians, centers[axis*V + v] in voxels
 *   psf      [h, h, nz] one h*h plane per depth (h odd)
 *
 * Pinning (see tests/test_oracle_*.py): golden vectors from oracle/_ref,
 * closed forms (uniform columns, 90-degree rotations), dense-matrix brute
 * force on tiny grids, the dot-product identity for the adjoint and FP64
 * central finite differences for the attenuation gradient.
 */
#ifndef ET_ORACLE_H
#define ET_ORACLE_H

#define ORC_DECLARE(REAL, S)                                                              \
    void orc_rotation_matrix_##S(REAL M[16], REAL ax, REAL ay, REAL az, REAL cx, REAL cy, \
                                 REAL cz, int order);                                     \
    void orc_resample_##S(const REAL *src, int nx, int ny, int nz, const REAL M[16],      \
                          REAL bg, REAL *out);                                            \
    void orc_resample_adj_##S(const REAL *g, int nx, int ny, int nz, const REAL M[16],    \
                              REAL *acc);                                                 \
    void orc_psf_separate_##S(const REAL *psf, int h, int nplanes, REAL *sep);            \
    void orc_convolve_planes_##S(REAL *vol, int nx, int ny, int nz, const REAL *sep,      \
                                 int h, int transpose);                                   \
    int orc_project_##S(const REAL *act, const REAL *mu, int nx, int ny, int nz,          \
                        const REAL *cams, const REAL *centers, int V, const REAL *psf,    \
                        int h, REAL bg, REAL bga, int truncate, REAL *sino);              \
    int orc_backproject_##S(const REAL *sino, const REAL *mu, int nx, int ny, int nz,     \
                            const REAL *cams, const REAL *centers, int V,                 \
                            const REAL *psf, int h, REAL bg, REAL bga, int truncate,      \
                            int mode, REAL *bkpr);                                        \
    int orc_gradient_ref_##S(const REAL *sino, const REAL *act, const REAL *mu, int nx,   \
                             int ny, int nz, const REAL *cams, const REAL *centers,       \
                             int V, const REAL *psf, int h, REAL bg, REAL bga,            \
                             int truncate, REAL *grad);                                   \
    int orc_gradient_mu_##S(const REAL *sino, const REAL *act, const REAL *mu, int nx,    \
                            int ny, int nz, const REAL *cams, const REAL *centers, int V, \
                            const REAL *psf, int h, REAL bg, REAL bga, REAL *grad);       \
    int orc_mlem_step_##S(REAL *act, const REAL *sino, const REAL *mu, int nx, int ny,    \
                          int nz, const REAL *cams, const REAL *centers, int V,           \
                          const REAL *psf, int h, REAL bg, REAL bga, REAL eps);           \
    int orc_project_partial_##S(const REAL *act, const REAL *mu, int nx, int ny, int nz,  \
                                const REAL *cams, const REAL *centers, int V,             \
                                const REAL *psf, int h, REAL bg, const REAL *bgimg,       \
                                REAL bga, int truncate, int rotate_partial, REAL *sino,   \
                                REAL *partial);                                           \
    int orc_fisher_grid_##S(const REAL *sino, int snx, int sny, const REAL *grid, int nx,   \
                            int ny, int nz, int n, const REAL *cams, const REAL *centers,  \
                            int V, const REAL *psf, int px, int py, int prior_rule,        \
                            REAL *fisher, REAL *prior);

ORC_DECLARE(float, f32)
ORC_DECLARE(double, f64)

/* Backprojection modes. */
#define ORC_BP_REFERENCE 0 /* inverse-rotation pull, reference _et.cpp:474-494 */
#define ORC_BP_ADJOINT 1   /* exact transpose of orc_project */

/* glibc rand()-driven camera subset of et_array_osem_step
 * (_et_array_interface.cpp:990-1009); writes n_sub indices clamped to V-1. */
int orc_osem_subset(int V, int n_sub, int *idx);

#endif
