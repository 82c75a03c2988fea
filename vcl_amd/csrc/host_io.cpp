// Host-only C ABI for the reference's on-disk formats and a headless image output:
//   OBJ meshes     -- Attribute::loadOBJ (buffer.cu:203-306);
//   BMP textures   -- Texture::loadBMP's decode (buffer.cu:486-541), and a BMP writer that is
//                     its exact inverse;
//   SG lobe text   -- SGBuffer::loadTXT (buffer.cu:706-732);
//   PFM images     -- float output for the headless replacement of GLbuffer display.
// No device code.  Arrays returned to the caller are malloc'd and freed by the matching
// vcl_*_free call.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "vcl.h"

namespace {

template <class T>
T* dup_array(const std::vector<T>& v) {
    if (v.empty()) return nullptr;
    T* p = (T*)std::malloc(v.size() * sizeof(T));
    if (p) std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

struct FileCloser {
    FILE* f;
    ~FileCloser() { if (f) std::fclose(f); }
};

void put_u16(unsigned char* p, uint32_t v) { p[0] = (unsigned char)(v & 0xff); p[1] = (unsigned char)((v >> 8) & 0xff); }
void put_u32(unsigned char* p, uint32_t v) { put_u16(p, v & 0xffff); put_u16(p + 2, v >> 16); }
uint32_t get_u32(const unsigned char* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24); }

}  // namespace

// The reader is token driven like the reference's: a whitespace-delimited token is a line
// header; "v" (3 floats), "vt" (2 floats) and "vn" (3 floats) append a record; "f" (only once a
// position has been read) reads three corners in the form fixed by the records read so far --
// v/vt/vn when both texels and normals exist, v/vt, v//vn, or plain v -- with 1-based indices;
// every other token is skipped.  A face that does not match the form is an error (the
// reference prints and returns without filling the attributes).
extern "C" int vcl_load_obj(const char* path, vcl_obj_mesh* out) {
    if (!path || !out) return VCL_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    FileCloser fc{std::fopen(path, "r")};
    if (!fc.f) return VCL_ERR_INVALID;
    std::vector<float> pos, tex, nrm;
    std::vector<uint32_t> pi, ti, ni;
    char tok[128];
    while (std::fscanf(fc.f, "%127s", tok) == 1) {
        if (!std::strcmp(tok, "v")) {
            float a, b, c;
            if (std::fscanf(fc.f, "%f %f %f", &a, &b, &c) != 3) return VCL_ERR_INVALID;
            pos.insert(pos.end(), {a, b, c});
        } else if (!std::strcmp(tok, "vt")) {
            float a, b;
            if (std::fscanf(fc.f, "%f %f", &a, &b) != 2) return VCL_ERR_INVALID;
            tex.insert(tex.end(), {a, b});
        } else if (!std::strcmp(tok, "vn")) {
            float a, b, c;
            if (std::fscanf(fc.f, "%f %f %f", &a, &b, &c) != 3) return VCL_ERR_INVALID;
            nrm.insert(nrm.end(), {a, b, c});
        } else if (!std::strcmp(tok, "f") && !pos.empty()) {
            int v[3], t[3], n[3];
            bool has_t = !tex.empty(), has_n = !nrm.empty();
            int got;
            if (has_t && has_n) {
                got = std::fscanf(fc.f, "%d/%d/%d %d/%d/%d %d/%d/%d", &v[0], &t[0], &n[0], &v[1], &t[1], &n[1], &v[2], &t[2], &n[2]);
                if (got != 9) return VCL_ERR_INVALID;
            } else if (has_t) {
                got = std::fscanf(fc.f, "%d/%d %d/%d %d/%d", &v[0], &t[0], &v[1], &t[1], &v[2], &t[2]);
                if (got != 6) return VCL_ERR_INVALID;
            } else if (has_n) {
                got = std::fscanf(fc.f, "%d//%d %d//%d %d//%d", &v[0], &n[0], &v[1], &n[1], &v[2], &n[2]);
                if (got != 6) return VCL_ERR_INVALID;
            } else {
                got = std::fscanf(fc.f, "%d %d %d", &v[0], &v[1], &v[2]);
                if (got != 3) return VCL_ERR_INVALID;
            }
            for (int k = 0; k < 3; k++) {
                pi.push_back((uint32_t)(v[k] - 1));
                if (has_t) ti.push_back((uint32_t)(t[k] - 1));
                if (has_n) ni.push_back((uint32_t)(n[k] - 1));
            }
        }
    }
    out->posNum = (int)(pos.size() / 3);
    out->texelNum = (int)(tex.size() / 2);
    out->normalNum = (int)(nrm.size() / 3);
    out->indexNum = (int)(pi.size() / 3);
    out->pos = dup_array(pos);
    out->texel = dup_array(tex);
    out->normal = dup_array(nrm);
    out->posIndex = dup_array(pi);
    // texel / normal indices exist only for faces read after texels / normals appeared; the
    // reference leaves the missing ones uninitialised, here they are 0
    if (!ti.empty()) { ti.resize(pi.size(), 0u); out->texelIndex = dup_array(ti); }
    if (!ni.empty()) { ni.resize(pi.size(), 0u); out->normalIndex = dup_array(ni); }
    return VCL_OK;
}

extern "C" void vcl_obj_free(vcl_obj_mesh* m) {
    if (!m) return;
    std::free(m->pos); std::free(m->texel); std::free(m->normal);
    std::free(m->posIndex); std::free(m->texelIndex); std::free(m->normalIndex);
    std::memset(m, 0, sizeof(*m));
}

// Texture::loadBMP's decode (buffer.cu:498-541): the 54-byte header gives the data offset
// (0x0A, 54 when 0), image size (0x22, width * height * channel when 0), width (0x12), height
// (0x16) and channels = bits per pixel (0x1C) / 8; the pixel bytes are read as one contiguous
// block (no row padding) with row 0 at the bottom, and value i of pixel p is byte
// (p + 1) * channel - (i + 1) / 255 (the channel order reversed per pixel, BGR -> RGB).
extern "C" int vcl_load_bmp(const char* path, vcl_image* out) {
    if (!path || !out) return VCL_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    FileCloser fc{std::fopen(path, "rb")};
    if (!fc.f) return VCL_ERR_INVALID;
    unsigned char h[54];
    if (std::fread(h, 1, 54, fc.f) != 54 || h[0] != 'B' || h[1] != 'M') return VCL_ERR_INVALID;
    uint32_t dataPos = get_u32(h + 0x0A), imageSize = get_u32(h + 0x22);
    int32_t w = (int32_t)get_u32(h + 0x12), hh = (int32_t)get_u32(h + 0x16);
    int c = (int)(get_u32(h + 0x1C) / 8);
    if (w <= 0 || hh <= 0 || c <= 0 || c > 4) return VCL_ERR_INVALID;
    size_t need = (size_t)w * hh * c;
    if (imageSize == 0) imageSize = (uint32_t)need;
    if (dataPos == 0) dataPos = 54;
    if (std::fseek(fc.f, (long)dataPos, SEEK_SET) != 0) return VCL_ERR_INVALID;
    std::vector<unsigned char> data(std::max<size_t>(imageSize, need), 0);
    size_t got = std::fread(data.data(), 1, imageSize, fc.f);
    if (got < need) return VCL_ERR_INVALID;
    std::vector<float> img(need);
    for (size_t p = 0; p < (size_t)w * hh; p++)
        for (int i = 0; i < c; i++) img[p * c + i] = (float)data[(p + 1) * c - (i + 1)] / 255.f;
    out->width = w; out->height = hh; out->channel = c;
    out->data = dup_array(img);
    return out->data ? VCL_OK : VCL_ERR_OOM;
}

extern "C" void vcl_image_free(vcl_image* im) {
    if (!im) return;
    std::free(im->data);
    std::memset(im, 0, sizeof(*im));
}

// The inverse of the decode above: byte (p + 1) * channel - (i + 1) = round(clamp(v, 0, 1) * 255)
// of value i of pixel p, rows bottom first, no row padding (so that the reader above, like the
// reference's, reads it back; widths with width * channel a multiple of 4 also give a standard
// BMP).  8-bit images carry a grey palette.
extern "C" int vcl_save_bmp(const char* path, const float* data, int width, int height, int channel) {
    if (!path || !data || width <= 0 || height <= 0 || channel <= 0 || channel > 4 || channel == 2) return VCL_ERR_INVALID;
    FileCloser fc{std::fopen(path, "wb")};
    if (!fc.f) return VCL_ERR_INVALID;
    size_t n = (size_t)width * height * channel;
    uint32_t pal = channel == 1 ? 256 * 4 : 0;
    unsigned char h[54];
    std::memset(h, 0, sizeof(h));
    h[0] = 'B'; h[1] = 'M';
    put_u32(h + 0x02, (uint32_t)(54 + pal + n));
    put_u32(h + 0x0A, 54 + pal);
    put_u32(h + 0x0E, 40);
    put_u32(h + 0x12, (uint32_t)width);
    put_u32(h + 0x16, (uint32_t)height);
    put_u16(h + 0x1A, 1);
    put_u16(h + 0x1C, (uint32_t)(8 * channel));
    put_u32(h + 0x22, (uint32_t)n);
    put_u32(h + 0x26, 2835);
    put_u32(h + 0x2A, 2835);
    if (channel == 1) put_u32(h + 0x2E, 256);
    if (std::fwrite(h, 1, 54, fc.f) != 54) return VCL_ERR_INVALID;
    if (channel == 1)
        for (int i = 0; i < 256; i++) {
            unsigned char e[4] = {(unsigned char)i, (unsigned char)i, (unsigned char)i, 0};
            if (std::fwrite(e, 1, 4, fc.f) != 4) return VCL_ERR_INVALID;
        }
    std::vector<unsigned char> bytes(n);
    for (size_t p = 0; p < (size_t)width * height; p++)
        for (int i = 0; i < channel; i++) {
            float v = data[p * channel + i];
            v = v < 0.f ? 0.f : (v > 1.f ? 1.f : v);
            if (v != v) v = 0.f;
            bytes[(p + 1) * channel - (i + 1)] = (unsigned char)(v * 255.f + 0.5f);
        }
    return std::fwrite(bytes.data(), 1, n, fc.f) == n ? VCL_OK : VCL_ERR_INVALID;
}

// Portable float map: "Pf" (1 channel) or "PF" (3 channels; 2 channels pad the third with 0,
// 4 drop the fourth), little-endian scale -1, rows bottom first as the images are stored.
extern "C" int vcl_save_pfm(const char* path, const float* data, int width, int height, int channel) {
    if (!path || !data || width <= 0 || height <= 0 || channel <= 0) return VCL_ERR_INVALID;
    FileCloser fc{std::fopen(path, "wb")};
    if (!fc.f) return VCL_ERR_INVALID;
    int oc = channel == 1 ? 1 : 3;
    std::fprintf(fc.f, "%s\n%d %d\n-1.0\n", oc == 1 ? "Pf" : "PF", width, height);
    std::vector<float> row((size_t)width * oc);
    for (int y = 0; y < height; y++) {
        for (int x = 0; x < width; x++)
            for (int i = 0; i < oc; i++) row[(size_t)x * oc + i] = i < channel ? data[((size_t)y * width + x) * channel + i] : 0.f;
        if (std::fwrite(row.data(), sizeof(float), row.size(), fc.f) != row.size()) return VCL_ERR_INVALID;
    }
    return VCL_OK;
}

// SGBuffer::loadTXT (buffer.cu:706-732): `num` lines of seven floats -- axis (3), sharpness,
// amplitude (3); the reference reads 256 lobes of 3 channels.
extern "C" int vcl_load_sg_txt(const char* path, int num, float* axis, float* sharpness, float* amplitude) {
    if (!path || !axis || !sharpness || !amplitude || num <= 0) return VCL_ERR_INVALID;
    FileCloser fc{std::fopen(path, "r")};
    if (!fc.f) return VCL_ERR_INVALID;
    for (int i = 0; i < num; i++) {
        float v[7];
        if (std::fscanf(fc.f, "%f %f %f %f %f %f %f", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6]) != 7) return VCL_ERR_INVALID;
        axis[3 * i] = v[0]; axis[3 * i + 1] = v[1]; axis[3 * i + 2] = v[2];
        sharpness[i] = v[3];
        amplitude[3 * i] = v[4]; amplitude[3 * i + 1] = v[5]; amplitude[3 * i + 2] = v[6];
    }
    return VCL_OK;
}
