// nifti_io.cpp -- see nifti_io.h.
#include "nifti_io.h"

#include <zlib.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <utility>

namespace {

constexpr int HDR = 348;       // header bytes
constexpr int VOX_OFFSET = 352;  // header + 4 extension bytes

bool ends_with(const std::string &s, const std::string &t)
{
    return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

bool slurp(const std::string &path, std::vector<unsigned char> &buf, std::string &err)
{
    gzFile f = gzopen(path.c_str(), "rb");  // reads plain and gzip files alike
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    unsigned char chunk[1 << 16];
    int n;
    while ((n = gzread(f, chunk, sizeof(chunk))) > 0) buf.insert(buf.end(), chunk, chunk + n);
    const bool ok = n == 0;
    gzclose(f);
    if (!ok) err = "read error in " + path;
    return ok;
}

template <class T>
T get(const unsigned char *p, bool swap)
{
    T v;
    std::memcpy(&v, p, sizeof(T));
    if (swap) {
        unsigned char b[sizeof(T)];
        std::memcpy(b, &v, sizeof(T));
        for (size_t i = 0; i < sizeof(T) / 2; i++) std::swap(b[i], b[sizeof(T) - 1 - i]);
        std::memcpy(&v, b, sizeof(T));
    }
    return v;
}

template <class T>
void put(unsigned char *p, T v)
{
    std::memcpy(p, &v, sizeof(T));
}

}  // namespace

bool nifti_name(const std::string &path) { return ends_with(path, ".nii") || ends_with(path, ".nii.gz"); }

bool nifti_read(const std::string &path, NiftiVolume &v, std::string &err)
{
    std::vector<unsigned char> buf;
    if (!slurp(path, buf, err)) return false;
    if (buf.size() < (size_t)HDR) {
        err = path + ": shorter than a NIfTI-1 header";
        return false;
    }
    const unsigned char *h = buf.data();
    bool swap = false;
    if (get<int32_t>(h, false) != HDR) {
        if (get<int32_t>(h, true) != HDR) {
            err = path + ": not a NIfTI-1 file";
            return false;
        }
        swap = true;
    }
    if (std::memcmp(h + 344, "n+1", 4) != 0) {
        err = path + ": not a single-file NIfTI-1 volume (magic n+1)";
        return false;
    }
    int16_t dim[8];
    for (int i = 0; i < 8; i++) dim[i] = get<int16_t>(h + 40 + 2 * i, swap);
    if (dim[0] < 1 || dim[0] > 7) {
        err = path + ": bad dim[0]";
        return false;
    }
    int n[7] = {1, 1, 1, 1, 1, 1, 1};
    for (int i = 1; i <= dim[0]; i++) n[i - 1] = dim[i] > 0 ? dim[i] : 1;
    for (int i = 3; i < 7; i++)
        if (n[i] != 1) {
            err = path + ": only volumes of up to three dimensions are read";
            return false;
        }
    const int16_t dt = get<int16_t>(h + 70, swap);
    const float slope = get<float>(h + 112, swap), inter = get<float>(h + 116, swap);
    const size_t off = (size_t)get<float>(h + 108, swap);
    v.nx = n[0];
    v.ny = n[1];
    v.nz = n[2];
    v.dx = get<float>(h + 80, swap);
    v.dy = get<float>(h + 84, swap);
    v.dz = get<float>(h + 88, swap);
    const size_t count = (size_t)v.nx * v.ny * v.nz;
    size_t bytes;
    switch (dt) {
    case 2: case 256: bytes = 1; break;          // uint8, int8
    case 4: case 512: bytes = 2; break;          // int16, uint16
    case 8: case 16: case 768: bytes = 4; break; // int32, float32, uint32
    case 64: bytes = 8; break;                   // float64
    default:
        err = path + ": unsupported datatype " + std::to_string(dt);
        return false;
    }
    if (off < (size_t)HDR || buf.size() < off + count * bytes) {
        err = path + ": data shorter than its dimensions";
        return false;
    }
    const unsigned char *d = buf.data() + off;
    v.data.resize(count);
    const bool scale = dt != 16 && slope != 0.f;
    for (size_t i = 0; i < count; i++) {
        const unsigned char *p = d + i * bytes;
        double x;
        switch (dt) {
        case 2: x = *p; break;
        case 256: x = (int8_t)*p; break;
        case 4: x = get<int16_t>(p, swap); break;
        case 512: x = get<uint16_t>(p, swap); break;
        case 8: x = get<int32_t>(p, swap); break;
        case 768: x = get<uint32_t>(p, swap); break;
        case 16: x = get<float>(p, swap); break;
        default: x = get<double>(p, swap); break;
        }
        v.data[i] = scale ? (float)(x * slope + inter) : (float)x;
    }
    return true;
}

bool nifti_write(const std::string &path, const NiftiVolume &v, std::string &err)
{
    const size_t count = (size_t)v.nx * v.ny * v.nz;
    if (v.data.size() != count) {
        err = "volume data do not match its dimensions";
        return false;
    }
    std::vector<unsigned char> buf(VOX_OFFSET + count * sizeof(float), 0);
    unsigned char *h = buf.data();
    put<int32_t>(h, HDR);                       // sizeof_hdr
    h[38] = 'r';                                // regular
    const int16_t dim[8] = {3, (int16_t)v.nx, (int16_t)v.ny, (int16_t)v.nz, 1, 1, 1, 1};
    for (int i = 0; i < 8; i++) put<int16_t>(h + 40 + 2 * i, dim[i]);
    put<int16_t>(h + 70, 16);                   // datatype float32
    put<int16_t>(h + 72, 32);                   // bitpix
    const float pixdim[8] = {1.f, v.dx, v.dy, v.dz, 1.f, 1.f, 1.f, 1.f};
    for (int i = 0; i < 8; i++) put<float>(h + 76 + 4 * i, pixdim[i]);
    put<float>(h + 108, (float)VOX_OFFSET);     // vox_offset
    put<float>(h + 112, 1.f);                   // scl_slope
    std::memcpy(h + 148, "niftyrec-mi355x", 15); // descrip
    std::memcpy(h + 344, "n+1", 4);             // magic (extension bytes 348..351 stay 0)
    std::memcpy(h + VOX_OFFSET, v.data.data(), count * sizeof(float));
    const bool gz = path.size() >= 3 && path.compare(path.size() - 3, 3, ".gz") == 0;
    if (gz) {
        gzFile f = gzopen(path.c_str(), "wb");
        if (!f) {
            err = "cannot create " + path;
            return false;
        }
        size_t done = 0;
        bool ok = true;
        while (done < buf.size() && ok) {
            const unsigned n = (unsigned)std::min<size_t>(buf.size() - done, 1u << 30);
            ok = gzwrite(f, buf.data() + done, n) == (int)n;
            done += n;
        }
        ok = (gzclose(f) == Z_OK) && ok;
        if (!ok) err = "write error in " + path;
        return ok;
    }
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) {
        err = "cannot create " + path;
        return false;
    }
    const bool ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    if (std::fclose(f) != 0 || !ok) {
        err = "write error in " + path;
        return false;
    }
    return true;
}
