// nifti_io.h -- NIfTI-1 single-file volumes (.nii, .nii.gz) for the command
// line tools: read any scalar datatype into float32 [nx, ny, nz] (x fastest),
// write float32.  Written from the published NIfTI-1 header layout (348-byte
// header, 4 extension bytes, data at vox_offset); the reference tools read and
// write through nifti/nifti1_io.c.
#pragma once
#include <string>
#include <vector>

struct NiftiVolume {
    int nx = 0, ny = 0, nz = 0;        // dim[1..3]; further dimensions must be 1
    float dx = 1.f, dy = 1.f, dz = 1.f;  // pixdim[1..3]
    std::vector<float> data;           // nx * ny * nz values, x fastest
};

// Reads `path` (.nii or .nii.gz).  Float32 data are taken as stored (as the
// reference tools take them); other scalar types are converted, with the
// header's scl_slope / scl_inter applied when the slope is non-zero.  Returns
// false and sets `err` on failure.
bool nifti_read(const std::string &path, NiftiVolume &v, std::string &err);

// Writes a float32 volume to `path` (.nii, or gzip-compressed .nii.gz).
bool nifti_write(const std::string &path, const NiftiVolume &v, std::string &err);

// True for names ending in .nii or .nii.gz.
bool nifti_name(const std::string &path);
