// rec_spect -- MLEM / OSEM reconstruction of a SPECT sinogram from the
// command line: NIfTI sinogram in, et_array_mlem / et_array_osem(_ordered)
// through lib_et_array_interface.so, NIfTI activity out.  Flags, defaults,
// printed parameters and the PSF of the reference tool
// apps/command_line_tools/rec_spect.cpp:361-633 (PSF from
// et_array_calculate_size_psf / et_array_make_psf, :534-538; background 0 and
// epsilon 1e-7, :27-29).  Differences, each documented in the usage text:
//   - OSEM runs deterministic ordered subsets (et_array_osem_ordered); the
//     reference's random subsets drawn with replacement are --legacy-subsets;
//   - an attenuation map given with -a is used (the reference reads the flag
//     and reconstructs without attenuation, :509, :548);
//   - NIfTI volumes only (.nii, .nii.gz): the reference's optional NRRD input
//     needs the teem library;
//   - the legacy subsets draw from rand() seeded with --seed (default 1, the
//     C library's own initial seed, so the default draws are the reference
//     tool's).
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "../../include/et_array_interface.h"
#include "nifti_io.h"

namespace {

const float BACKGROUND_ACTIVITY = 0.f;
const float BACKGROUND_ATTENUATION = 0.f;
const float EPSILON = 1e-7f;

struct Flag {
    const char *shrt, *lng, *help, *def;
    bool is_switch;
};

// (short, long, help, default, switch)
const Flag FLAGS[] = {
    {"s", "sinogram", "Sinogram data file to reconstruct (.nii, .nii.gz), size [N x m x n_cameras].", nullptr, false},
    {"o", "output", "Output file for the reconstructed activity (.nii, .nii.gz), size [N x m x N].", nullptr, false},
    {"a", "attenuation", "Attenuation map (.nii, .nii.gz), size [N x m x N].", "none", false},
    {"c", "firstcamera", "Position of the first camera in degrees.", "0.0", false},
    {"C", "lastcamera", "Position of the last camera in degrees.", "180.0", false},
    {"x", "axis", "Axis of rotation of the gamma camera [0,1,2].", "0", false},
    {"f", "fwhm0", "FWHM of the PSF at distance dist0 in pixels.", "3.0", false},
    {"F", "fwhm1", "FWHM of the PSF at distance dist1 in pixels.", "3.0", false},
    {"e", "efficiency0", "Efficiency of the collimator/detector response at distance dist0.", "0.9", false},
    {"E", "efficiency1", "Efficiency of the collimator/detector response at distance dist1.", "0.9", false},
    {"d", "dist0", "Distance of PSF characterisation 0 in pixels.", "10.0", false},
    {"D", "dist1", "Distance of PSF characterisation 1 in pixels.", "10.0", false},
    {"r", "radius", "Radius of rotation of the gamma camera in pixels.", "20.0", false},
    {"i", "iterations", "Iterations (MLEM updates, or OSEM passes over the subsets).", "20", false},
    {"u", "subsets", "OSEM subsets; 0 or 1 for MLEM.", "16", false},
    {"L", "legacy-subsets", "OSEM on random camera subsets drawn with replacement (the reference tool's et_array_osem).", nullptr, true},
    {"S", "seed", "Seed of the legacy subsets' rand() draws (1: the C library's initial seed).", "1", false},
    {"g", "gpu_off", "Accepted for compatibility; the reconstruction always runs on the GPU.", nullptr, true},
    {"v", "verbose", "Print verbose information.", nullptr, true},
};

void usage()
{
    std::cout << "rec_spect: MLEM / OSEM reconstruction of SPECT sinogram data on the GPU.\n\nUsage: rec_spect -s "
                 "<sinogram> -o <output> [options]\n       rec_spect --xml\n\n";
    for (const Flag &f : FLAGS) {
        std::cout << "  -" << f.shrt << ", --" << f.lng << (f.is_switch ? "" : " <value>") << "\n      " << f.help;
        if (f.def) std::cout << " Default " << f.def << ".";
        std::cout << "\n";
    }
}

// Module description for hosts that run command line modules (the reference
// tool's --xml, rec_spect.cpp:268-287).
void stream_xml()
{
    std::cout << "<?xml version=\"1.0\" encoding=\"utf-8\"?>\n<executable>\n  <category>Reconstruction</category>\n"
                 "  <title>NiftyRec SPECT</title>\n  <description>MLEM and OSEM reconstruction of SPECT sinogram "
                 "data</description>\n  <parameters>\n";
    for (const Flag &f : FLAGS) {
        const bool file = !std::strcmp(f.lng, "sinogram") || !std::strcmp(f.lng, "output") ||
                          !std::strcmp(f.lng, "attenuation");
        const char *tag = f.is_switch ? "boolean" : file ? "image" : !std::strcmp(f.lng, "axis") ||
                                                                     !std::strcmp(f.lng, "iterations") ||
                                                                     !std::strcmp(f.lng, "subsets")
                                                                     ? "integer"
                                                                     : "float";
        std::cout << "    <" << tag << ">\n      <name>" << f.lng << "</name>\n      <flag>" << f.shrt
                  << "</flag>\n      <longflag>" << f.lng << "</longflag>\n      <description>" << f.help
                  << "</description>\n";
        if (f.def && !file) std::cout << "      <default>" << f.def << "</default>\n";
        if (!std::strcmp(f.lng, "output")) std::cout << "      <channel>output</channel>\n";
        std::cout << "    </" << tag << ">\n";
    }
    std::cout << "  </parameters>\n</executable>\n";
}

bool parse(int argc, char **argv, std::map<std::string, std::string> &val, std::string &err)
{
    for (const Flag &f : FLAGS)
        if (f.def) val[f.lng] = f.def;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        const Flag *hit = nullptr;
        for (const Flag &f : FLAGS)
            if (a == std::string("-") + f.shrt || a == std::string("--") + f.lng) hit = &f;
        if (!hit) {
            err = "unknown argument " + a;
            return false;
        }
        if (hit->is_switch) {
            val[hit->lng] = "1";
            continue;
        }
        if (i + 1 >= argc) {
            err = "missing value for " + a;
            return false;
        }
        val[hit->lng] = argv[++i];
    }
    for (const char *req : {"sinogram", "output"})
        if (!val.count(req)) {
            err = std::string("required argument missing: --") + req;
            return false;
        }
    return true;
}

bool to_float(const std::string &s, float &out)
{
    char *end = nullptr;
    errno = 0;
    out = std::strtof(s.c_str(), &end);
    return errno == 0 && end && *end == 0 && !s.empty();
}

bool to_int(const std::string &s, int &out)
{
    char *end = nullptr;
    errno = 0;
    const long v = std::strtol(s.c_str(), &end, 10);
    out = (int)v;
    return errno == 0 && end && *end == 0 && !s.empty();
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc == 2 && !std::strcmp(argv[1], "--xml")) {
        stream_xml();
        return 0;
    }
    if (argc == 2 && (!std::strcmp(argv[1], "-h") || !std::strcmp(argv[1], "--help"))) {
        usage();
        return 0;
    }
    std::map<std::string, std::string> val;
    std::string err;
    if (!parse(argc, argv, val, err)) {
        std::cerr << "rec_spect: " << err << "\n\n";
        usage();
        return 1;
    }
    float first, last, fwhm0, fwhm1, eff0, eff1, dist0, dist1, radius;
    int axis, iterations, subsets, seed;
    const bool ok = to_float(val["firstcamera"], first) && to_float(val["lastcamera"], last) &&
                    to_float(val["fwhm0"], fwhm0) && to_float(val["fwhm1"], fwhm1) &&
                    to_float(val["efficiency0"], eff0) && to_float(val["efficiency1"], eff1) &&
                    to_float(val["dist0"], dist0) && to_float(val["dist1"], dist1) &&
                    to_float(val["radius"], radius) && to_int(val["axis"], axis) &&
                    to_int(val["iterations"], iterations) && to_int(val["subsets"], subsets) &&
                    to_int(val["seed"], seed);
    if (!ok || axis < 0 || axis > 2 || iterations < 0) {
        std::cerr << "rec_spect: bad numeric argument\n";
        return 1;
    }
    const std::string sino_name = val["sinogram"], out_name = val["output"], mu_name = val["attenuation"];
    const bool legacy = val.count("legacy-subsets") > 0;
    for (const std::string &n : {sino_name, out_name})
        if (!nifti_name(n)) {
            std::cerr << "rec_spect: unknown file format (NIfTI .nii or .nii.gz): " << n << "\n";
            return 1;
        }
    NiftiVolume sino;
    if (!nifti_read(sino_name, sino, err)) {
        std::cerr << "rec_spect: couldn't load the sinogram: " << err << "\n";
        return 1;
    }
    const unsigned size_x = sino.nx, size_y = sino.ny;
    const int n_cameras = sino.nz;
    NiftiVolume mu;
    const bool use_mu = mu_name != "none";
    if (use_mu) {
        if (!nifti_read(mu_name, mu, err)) {
            std::cerr << "rec_spect: couldn't load the attenuation map: " << err << "\n";
            return 1;
        }
        if ((unsigned)mu.nx != size_x || (unsigned)mu.ny != size_y || (unsigned)mu.nz != size_x) {
            std::cerr << "rec_spect: the attenuation map must be [N x m x N] for a sinogram of [N x m x n_cameras]\n";
            return 1;
        }
    }
    std::cout << "=============== NiftyRec SPECT Parameters ===============\n"
              << "Sinogram file name:       " << sino_name << "\n"
              << "Output file name:         " << out_name << "\n"
              << "Attenuation file name:    " << mu_name << "\n"
              << "First camera:             " << first << "\n"
              << "Last camera:              " << last << "\n"
              << "Rotation axis:            " << axis << "\n"
              << "PSF distance 0 [pixels]:  " << dist0 << "\n"
              << "PSF FWHM 0 [pixels]:      " << fwhm0 << "\n"
              << "PSF efficiency 0:         " << eff0 << "\n"
              << "PSF distance 1 [pixels]:  " << dist1 << "\n"
              << "PSF FWHM 1 [pixels]:      " << fwhm1 << "\n"
              << "PSF efficiency 1:         " << eff1 << "\n"
              << "Rotation radius [pixels]: " << radius << "\n"
              << "Iterations:               " << iterations << "\n"
              << "OSEM subsets:             " << subsets << (subsets > 1 ? (legacy ? " (random, with replacement)"
                                                                                    : " (ordered)")
                                                                         : "")
              << "\n"
              << "Use GPU:                  1\n"
              << "Sinogram size X [pixels]: " << size_x << "\n"
              << "Sinogram size Y [pixels]: " << size_y << "\n"
              << "Number of projections:    " << n_cameras << "\n"
              << "Activity size X [pixels]: " << size_x << "\n"
              << "Activity size Y [pixels]: " << size_y << "\n"
              << "Activity size Z [pixels]: " << size_x << "\n"
              << "=========================================================" << std::endl;

    unsigned psf_x = 0, psf_y = 0;
    int status = et_array_calculate_size_psf(&psf_x, &psf_y, fwhm0, eff0, dist0, fwhm1, eff1, dist1);
    if (status != 0) {
        std::cerr << "rec_spect: error creating the PSF\n";
        return status;
    }
    std::vector<float> psf((size_t)psf_x * psf_y * size_x);
    status = et_array_make_psf(psf.data(), psf_x, psf_y, fwhm0, eff0, dist0, fwhm1, eff1, dist1, size_x);
    if (status != 0) {
        std::cerr << "rec_spect: error creating the PSF\n";
        return status;
    }
    NiftiVolume act;
    act.nx = size_x;
    act.ny = size_y;
    act.nz = size_x;
    act.dx = sino.dx;
    act.dy = sino.dy;
    act.dz = sino.dx;
    act.data.assign((size_t)size_x * size_y * size_x, 0.f);
    float *mu_data = use_mu ? mu.data.data() : nullptr;
    if (subsets <= 1)
        status = et_array_mlem(act.data.data(), size_x, size_y, sino.data.data(), n_cameras, first, last, axis,
                               iterations, use_mu, mu_data, 1, psf.data(), psf_x, psf_y, BACKGROUND_ACTIVITY,
                               BACKGROUND_ATTENUATION, EPSILON, 1);
    else if (legacy) {
        std::srand((unsigned)seed);
        status = et_array_osem(act.data.data(), size_x, size_y, subsets, sino.data.data(), n_cameras, first, last,
                               axis, iterations, use_mu, mu_data, 1, psf.data(), psf_x, psf_y, BACKGROUND_ACTIVITY,
                               BACKGROUND_ATTENUATION, EPSILON, 1);
    } else
        status = et_array_osem_ordered(act.data.data(), size_x, size_y, subsets, sino.data.data(), n_cameras, first,
                                       last, axis, iterations, use_mu, mu_data, 1, psf.data(), psf_x, psf_y,
                                       BACKGROUND_ACTIVITY, BACKGROUND_ATTENUATION, EPSILON, 1);
    if (status != 0) {
        std::cerr << "rec_spect: reconstruction failed: " << niftyrec_error_msg(status) << "\n";
        return 1;
    }
    std::cout << "Nifti output file: " << out_name << std::endl;
    if (!nifti_write(out_name, act, err)) {
        std::cerr << "rec_spect: " << err << "\n";
        return 1;
    }
    return 0;
}
