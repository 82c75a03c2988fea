// vcl_layers.hpp -- header-only C++ drop-in for the reference's layer classes, over the C ABI
// of vcl.h.  Class names, Params/GradParams structs and the init/forward/backward overload
// sets follow the reference headers (236088/vcl nvdiffrast/nvdiffrast/*.h) so preset-style
// host code (preset_primitives.cpp / preset_pose.cpp call order) compiles against it; there
// are no GL/CUDA/glm/OpenCV includes.  Semantics kept from the reference:
//   * init() borrows upstream device pointers and allocates its own outputs;
//   * forward(XGradParams&) zeroes that op's output grads, then runs the forward;
//   * errors print and exit(1) (common.h:36-47 ERROR_STRING / CUDA_ERROR_CHECK).
// Differences: every call takes an optional stream (default: the null stream); scratch for
// rasterize / SG backward / loss is allocated once in init(); Loss::loss() copies the
// device scalar down only when asked; the random helpers draw their seeds with rand() as the
// reference does; GLbuffer::draw composites into a headless window canvas (vcl::Window) that
// Window::save writes to a BMP or PFM file instead of an OpenGL window.
// Where the reference's backward formula differs from the derivative (Rotation, Camera, SSIM,
// ComputeNormal angle weights), the reference formula is the default and the exact one is
// opt-in (the `exact` flags below).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "vcl.h"
#include "vcl_matrix.hpp"

// OpenGL format names used by GLbuffer::draw calls in the reference presets (values of the GL
// specification), so that those calls compile without GL headers.
#ifndef GL_RED
#define GL_RED 0x1903
#endif
#ifndef GL_RGB
#define GL_RGB 0x1907
#endif
#ifndef GL_RGBA
#define GL_RGBA 0x1908
#endif
#ifndef GL_LUMINANCE
#define GL_LUMINANCE 0x1909
#endif
#ifndef GL_RGBA32F
#define GL_RGBA32F 0x8814
#endif
#ifndef GL_RGB32F
#define GL_RGB32F 0x8815
#endif

namespace vcl {

#define VCL_LAYER_CHECK(call)                                                                         \
    do {                                                                                              \
        int st_ = (call);                                                                             \
        if (st_ != VCL_OK) {                                                                          \
            std::fprintf(stderr, "%s in file %s at line %d\n", vcl_status_string(st_), __FILE__, __LINE__); \
            std::exit(1);                                                                             \
        }                                                                                             \
    } while (0)

#define VCL_HIP_LAYER_CHECK(call)                                                                     \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess) {                                                                       \
            std::fprintf(stderr, "%s in file %s at line %d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
            std::exit(1);                                                                             \
        }                                                                                             \
    } while (0)

inline vcl_stream_t& default_stream() {
    static vcl_stream_t s = nullptr;
    return s;
}
inline hipStream_t hs() { return (hipStream_t)default_stream(); }

template <class T>
inline T* dmalloc(size_t n) {
    void* p = nullptr;
    VCL_HIP_LAYER_CHECK(hipMalloc(&p, n ? n * sizeof(T) : sizeof(T)));
    return (T*)p;
}
inline void dzero(void* p, size_t bytes) {
    if (p && bytes) VCL_HIP_LAYER_CHECK(hipMemsetAsync(p, 0, bytes, hs()));
}

constexpr int TEX_MAX_MIP_LEVEL = VCL_TEX_MAX_MIP_LEVEL;

// The reference draws every helper's random seed with rand() (e.g. buffer.cu:148).
inline uint32_t draw_seed() { return (uint32_t)std::rand(); }

// ------------------------------------------------------------------ buffers (buffer.h)
struct Buffer {  // buffer.h:6-20
    float* buffer = nullptr;
    int num = 0, dimention = 0;
    size_t Size() const { return (size_t)num * dimention * sizeof(float); }
    static void init(Buffer& b, int num, int dimention) {
        b.num = num; b.dimention = dimention;
        b.buffer = dmalloc<float>((size_t)num * dimention);
    }
    static void init(Buffer& b, Buffer& src, int dimention) { init(b, src.num, dimention); }
    static void copy(Buffer& dst, Buffer& src) {
        VCL_HIP_LAYER_CHECK(hipMemcpy(dst.buffer, src.buffer, dst.Size(), hipMemcpyDeviceToDevice));
    }
    static void copy(Buffer& dst, const float* host) { VCL_HIP_LAYER_CHECK(hipMemcpy(dst.buffer, host, dst.Size(), hipMemcpyHostToDevice)); }
    static void copy(float* host, Buffer& src) { VCL_HIP_LAYER_CHECK(hipMemcpy(host, src.buffer, src.Size(), hipMemcpyDeviceToHost)); }
    static void liner(Buffer& b, float w, float c) {
        VCL_LAYER_CHECK(vcl_buffer_linear(b.buffer, b.buffer, (size_t)b.num * b.dimention, w, c, default_stream()));
    }
    static void addRandom(Buffer& b, float min, float max) {
        VCL_LAYER_CHECK(vcl_buffer_add_random(b.buffer, (size_t)b.num * b.dimention, min, max, draw_seed(), default_stream()));
    }
    static void clamp(Buffer& b, float min, float max) {
        VCL_LAYER_CHECK(vcl_buffer_clamp(b.buffer, b.buffer, (size_t)b.num * b.dimention, min, max, default_stream()));
    }
    static void normalize(Buffer& b) { VCL_LAYER_CHECK(vcl_buffer_normalize(b.buffer, b.buffer, b.num, b.dimention, default_stream())); }
};

struct BufferGrad : Buffer {  // buffer.h:22-27
    float* grad = nullptr;
    static void init(BufferGrad& b, int num, int dimention) {
        Buffer::init(b, num, dimention);
        b.grad = dmalloc<float>((size_t)num * dimention);
        clear(b);
    }
    static void init(BufferGrad& b, Buffer& src, int dimention) { init(b, src.num, dimention); }
    static void clear(BufferGrad& b) { dzero(b.grad, b.Size()); }
};

struct Attribute {  // buffer.h:29-45
    float* vbo = nullptr;
    unsigned int* vao = nullptr;
    int vboNum = 0, vaoNum = 0, dimention = 0;
    size_t vboSize() const { return (size_t)vboNum * dimention * sizeof(float); }
    size_t vaoSize() const { return (size_t)vaoNum * 3 * sizeof(unsigned int); }
    static void init(Attribute& a, int vboNum, int vaoNum, int dimention) {
        a.vboNum = vboNum; a.vaoNum = vaoNum; a.dimention = dimention;
        a.vbo = dmalloc<float>((size_t)vboNum * dimention);
        a.vao = dmalloc<unsigned int>((size_t)vaoNum * 3);
    }
    static void init(Attribute& a, Attribute& src, int dimention) {  // shares the index buffer
        a.vboNum = src.vboNum; a.vaoNum = src.vaoNum; a.dimention = dimention;
        a.vbo = dmalloc<float>((size_t)a.vboNum * dimention);
        a.vao = src.vao;
    }
    static void copy(Attribute& dst, const float* host) {
        VCL_HIP_LAYER_CHECK(hipMemcpy(dst.vbo, host, dst.vboSize(), hipMemcpyHostToDevice));
    }
    static void copyIndex(Attribute& dst, const unsigned int* host) {
        VCL_HIP_LAYER_CHECK(hipMemcpy(dst.vao, host, dst.vaoSize(), hipMemcpyHostToDevice));
    }
    static void copy(Attribute& dst, Attribute& src) {  // buffer.cu:308-311
        VCL_HIP_LAYER_CHECK(hipMemcpy(dst.vbo, src.vbo, dst.vboSize(), hipMemcpyDeviceToDevice));
        if (dst.vao != src.vao) VCL_HIP_LAYER_CHECK(hipMemcpy(dst.vao, src.vao, dst.vaoSize(), hipMemcpyDeviceToDevice));
    }
    // buffer.cu:203-306: each non-null attribute gets the file's records and its own index list
    static void loadOBJ(const char* path, Attribute* pos, Attribute* texel, Attribute* normal) {
        vcl_obj_mesh m;
        int st = vcl_load_obj(path, &m);
        if (st != VCL_OK) { std::printf("Impossible to open or parse the file %s\n", path); return; }
        auto put = [&](Attribute* a, int n, int dim, const float* v, const uint32_t* idx) {
            if (!a || n <= 0 || !v) return;
            init(*a, n, m.indexNum, dim);
            copy(*a, v);
            if (idx && m.indexNum) copyIndex(*a, idx);
        };
        put(pos, m.posNum, 3, m.pos, m.posIndex);
        put(texel, m.texelNum, 2, m.texel, m.texelIndex);
        put(normal, m.normalNum, 3, m.normal, m.normalIndex);
        vcl_obj_free(&m);
    }
    static void liner(Attribute& a, float w, float b) {  // buffer.cu:313-319
        VCL_LAYER_CHECK(vcl_buffer_linear(a.vbo, a.vbo, (size_t)a.vboNum * a.dimention, w, b, default_stream()));
    }
    static float distanceError(Attribute& predict, Attribute& target) {  // buffer.cu:321-344
        float* d = dmalloc<float>(1);
        VCL_LAYER_CHECK(vcl_attribute_distance_error(predict.vbo, target.vbo, predict.vboNum, predict.dimention, d, default_stream()));
        float h = 0.f;
        VCL_HIP_LAYER_CHECK(hipMemcpyAsync(&h, d, sizeof(float), hipMemcpyDeviceToHost, hs()));
        VCL_HIP_LAYER_CHECK(hipStreamSynchronize(hs()));
        VCL_HIP_LAYER_CHECK(hipFree(d));
        return h;
    }
    static void addRandom(Attribute& a, float min, float max) {  // buffer.cu:346-353
        VCL_LAYER_CHECK(vcl_buffer_add_random(a.vbo, (size_t)a.vboNum * a.dimention, min, max, draw_seed(), default_stream()));
    }
    static void step(Attribute& a, float threshold) {  // buffer.cu:355-361
        VCL_LAYER_CHECK(vcl_buffer_step(a.vbo, a.vbo, (size_t)a.vboNum * a.dimention, threshold, default_stream()));
    }
};

struct AttributeGrad : Attribute {  // buffer.h:47-52
    float* grad = nullptr;
    static void init(AttributeGrad& a, int vboNum, int vaoNum, int dimention) {
        Attribute::init(a, vboNum, vaoNum, dimention);
        a.grad = dmalloc<float>((size_t)vboNum * dimention);
        clear(a);
    }
    static void init(AttributeGrad& a, Attribute& src, int dimention) {
        Attribute::init(a, src, dimention);
        a.grad = dmalloc<float>((size_t)a.vboNum * dimention);
        clear(a);
    }
    static void clear(AttributeGrad& a) { dzero(a.grad, a.vboSize()); }
};

struct Texture {  // buffer.h:54-71
    int width = 0, height = 0, channel = 0, miplevel = 0;
    float* texture[TEX_MAX_MIP_LEVEL] = {};
    size_t Size() const { return (size_t)width * height * channel * sizeof(float); }
    vcl_texture c() const {
        vcl_texture t;
        t.width = width; t.height = height; t.channel = channel; t.miplevel = miplevel;
        for (int i = 0; i < TEX_MAX_MIP_LEVEL; i++) t.texture[i] = texture[i];
        return t;
    }
    static void init(Texture& t, int width, int height, int channel, int miplevel) {  // buffer.cu:379-391
        t.width = width; t.height = height; t.channel = channel;
        t.miplevel = vcl_texture_mip_levels(width, height, miplevel);
        for (int i = 0; i < t.miplevel; i++) t.texture[i] = dmalloc<float>((size_t)(width >> i) * (height >> i) * channel);
    }
    static void init(Texture& t, float* original, int width, int height, int channel, int miplevel) {  // buffer.cu:393-407
        t.width = width; t.height = height; t.channel = channel;
        t.miplevel = vcl_texture_mip_levels(width, height, miplevel);
        t.texture[0] = original;
        for (int i = 1; i < t.miplevel; i++) t.texture[i] = dmalloc<float>((size_t)(width >> i) * (height >> i) * channel);
        buildMIP(t);
    }
    static void buildMIP(Texture& t) {
        vcl_texture c = t.c();
        VCL_LAYER_CHECK(vcl_texture_build_mip(&c, default_stream()));
    }
    static void clamp(Texture& t, float min, float max) {
        vcl_texture c = t.c();
        VCL_LAYER_CHECK(vcl_texture_clamp(&c, min, max, default_stream()));
    }
    static void setLevel0(Texture& t, const float* host) {
        VCL_HIP_LAYER_CHECK(hipMemcpy(t.texture[0], host, t.Size(), hipMemcpyHostToDevice));
        buildMIP(t);
    }
    static void bilinearDownsampling(Texture& t) {  // buffer.cu:437-447
        vcl_texture c = t.c();
        VCL_LAYER_CHECK(vcl_texture_bilinear_downsampling(&c, default_stream()));
    }
    // buffer.cu:498-541: a texture sized from the file, its level 0 decoded, mips rebuilt
    static void loadBMP(const char* path, Texture& t, int miplevel) {
        vcl_image im;
        if (vcl_load_bmp(path, &im) != VCL_OK) { std::printf("Image could not be opened or is not a correct BMP file: %s\n", path); std::exit(1); }
        init(t, im.width, im.height, im.channel, miplevel);
        VCL_HIP_LAYER_CHECK(hipMemcpy(t.texture[0], im.data, t.Size(), hipMemcpyHostToDevice));
        vcl_image_free(&im);
        buildMIP(t);
    }
    static void setColor(Texture& t, const float* color) {  // buffer.cu:543-556, every level
        int w = t.width, h = t.height;
        for (int i = 0; i < t.miplevel; i++, w >>= 1, h >>= 1)
            VCL_LAYER_CHECK(vcl_buffer_set_value(t.texture[i], w * h, t.channel, color, default_stream()));
    }
    static void liner(Texture& t, float w, float b) {  // buffer.cu:558-566, every level
        int tw = t.width, th = t.height;
        for (int i = 0; i < t.miplevel; i++, tw >>= 1, th >>= 1)
            VCL_LAYER_CHECK(vcl_buffer_linear(t.texture[i], t.texture[i], (size_t)tw * th * t.channel, w, b, default_stream()));
    }
    static void normalize(Texture& t) {  // buffer.cu:568-576, every level, rows of `channel`
        int tw = t.width, th = t.height;
        for (int i = 0; i < t.miplevel; i++, tw >>= 1, th >>= 1)
            VCL_LAYER_CHECK(vcl_buffer_normalize(t.texture[i], t.texture[i], tw * th, t.channel, default_stream()));
    }
    // buffer.cu:578-586: the reference's parameters are named (max, min) but passed as the
    // kernel's (min, max), and its launch covers width x height values of level 0 (one layer of
    // the grid); both kept.  Then the mips are rebuilt.
    static void addRandom(Texture& t, float max, float min) {
        VCL_LAYER_CHECK(vcl_buffer_add_random(t.texture[0], (size_t)t.width * t.height, max, min, draw_seed(), default_stream()));
        buildMIP(t);
    }
};

struct TextureGrad : Texture {  // buffer.h:73-78
    float* grad[TEX_MAX_MIP_LEVEL] = {};
    vcl_texture g() const {
        vcl_texture t;
        t.width = width; t.height = height; t.channel = channel; t.miplevel = miplevel;
        for (int i = 0; i < TEX_MAX_MIP_LEVEL; i++) t.texture[i] = grad[i];
        return t;
    }
    static void init(TextureGrad& t, int width, int height, int channel, int miplevel) {
        Texture::init(t, width, height, channel, miplevel);
        for (int i = 0; i < t.miplevel; i++) t.grad[i] = dmalloc<float>((size_t)(width >> i) * (height >> i) * channel);
        clear(t);  // the reference leaves them uninitialised (buffer.cu:598-605)
    }
    static void clear(TextureGrad& t) {
        vcl_texture c = t.g();
        VCL_LAYER_CHECK(vcl_texture_grad_clear(&c, default_stream()));
    }
    // gradSumup (buffer.cu:638-648): each coarse level's grad added to its four children as the
    // reference does (x1); exact = true scales each folded level by 1/4, the adjoint of buildMIP
    static void gradSumup(TextureGrad& t, bool exact = false) {
        vcl_texture c = t.g();
        VCL_LAYER_CHECK(vcl_texture_grad_sumup(&c, exact ? 1 : 0, default_stream()));
    }
};

struct SGBuffer {  // buffer.h:82-94
    float* axis = nullptr;
    float* sharpness = nullptr;
    float* amplitude = nullptr;
    int num = 0, channel = 0;
    static void init(SGBuffer& s, int num, int channel) {
        s.num = num; s.channel = channel;
        s.axis = dmalloc<float>((size_t)num * 3);
        s.sharpness = dmalloc<float>(num);
        s.amplitude = dmalloc<float>((size_t)num * channel);
        dzero(s.axis, (size_t)num * 3 * sizeof(float));  // the reference leaves them uninitialised
        dzero(s.sharpness, (size_t)num * sizeof(float));
        dzero(s.amplitude, (size_t)num * channel * sizeof(float));
    }
    static void copy(SGBuffer& dst, const float* axis, const float* sharpness, const float* amplitude) {  // buffer.cu:660-667
        if (axis) VCL_HIP_LAYER_CHECK(hipMemcpy(dst.axis, axis, (size_t)dst.num * 3 * sizeof(float), hipMemcpyHostToDevice));
        if (sharpness) VCL_HIP_LAYER_CHECK(hipMemcpy(dst.sharpness, sharpness, (size_t)dst.num * sizeof(float), hipMemcpyHostToDevice));
        if (amplitude) VCL_HIP_LAYER_CHECK(hipMemcpy(dst.amplitude, amplitude, (size_t)dst.num * dst.channel * sizeof(float), hipMemcpyHostToDevice));
    }
    // buffer.cu:669-683: random unit axes; sharpness += U[1, 100] (the reference adds to
    // uninitialised memory; here the buffer starts from zero)
    static void randomize(SGBuffer& s) {
        VCL_LAYER_CHECK(vcl_buffer_spherical_random(s.axis, s.num, draw_seed(), default_stream()));
        VCL_LAYER_CHECK(vcl_buffer_add_random(s.sharpness, (size_t)s.num, 1.f, 100.f, draw_seed(), default_stream()));
    }
    static void normalize(SGBuffer& s) {  // buffer.cu:685-704
        VCL_LAYER_CHECK(vcl_sg_normalize(s.axis, s.sharpness, s.amplitude, s.num, s.channel, default_stream()));
    }
    static void loadTXT(const char* path, SGBuffer* s) {  // buffer.cu:706-732: 256 lobes x 3 channels
        init(*s, 256, 3);
        std::vector<float> a(256 * 3), sh(256), am(256 * 3);
        if (vcl_load_sg_txt(path, 256, a.data(), sh.data(), am.data()) != VCL_OK) { std::printf("Impossible to read the file %s\n", path); return; }
        copy(*s, a.data(), sh.data(), am.data());
    }
    static void bake(SGBuffer& s, Texture& t) {  // buffer.cu:755-761: level 0 <- the lobes over equirectangular directions
        vcl_texture c = t.c();
        VCL_LAYER_CHECK(vcl_sg_bake(s.axis, s.sharpness, s.amplitude, s.num, s.channel, &c, default_stream()));
    }
};

// SGBufferGrad redeclares axis/sharpness/amplitude: base = parameters, derived = gradients
// (buffer.h:96-100, A-16); ((SGBuffer&)g).axis is the parameter, g.axis the gradient.
struct SGBufferGrad : SGBuffer {
    float* axis = nullptr;
    float* sharpness = nullptr;
    float* amplitude = nullptr;
    float* buffer = nullptr;  // num x 3 (disperse)
    static void init(SGBufferGrad& s, int num, int channel) {
        SGBuffer::init(s, num, channel);
        s.axis = dmalloc<float>((size_t)num * 3);
        s.sharpness = dmalloc<float>(num);
        s.amplitude = dmalloc<float>((size_t)num * channel);
        s.buffer = dmalloc<float>((size_t)num * 3);
        clear(s);
    }
    static void clear(SGBufferGrad& s) {  // buffer.cu:806-811
        dzero(s.axis, (size_t)s.num * 3 * sizeof(float));
        dzero(s.sharpness, (size_t)s.num * sizeof(float));
        dzero(s.amplitude, (size_t)s.num * s.channel * sizeof(float));
        dzero(s.buffer, (size_t)s.num * 3 * sizeof(float));
    }
    // buffer.cu:798-804: the parameter axes pushed apart (amplitude differences from the
    // amplitudes; the reference indexes sharpness there)
    static void disperse(SGBufferGrad& s) {
        SGBuffer& p = s;
        VCL_LAYER_CHECK(vcl_sg_disperse(p.axis, p.sharpness, p.amplitude, s.num, s.channel, s.buffer, default_stream()));
    }
};

// ------------------------------------------------------------------ headless display (GLbuffer)
// The reference shows device images in a GLUT window (GLbuffer, buffer.cu:814-851).  Here
// GLbuffer::draw copies the image down and composites it, nearest-texel and clamped to the
// edge as the reference's GL_NEAREST / GL_CLAMP_TO_EDGE texture, into the normalised-device
// rectangle it names of a window-sized RGB canvas; Window::save writes the canvas (row 0 at
// the bottom, as GL's window origin) to a .bmp or .pfm file.
struct Window {
    int width = 0, height = 0;
    std::vector<float> rgb;  // width x height x 3
    static Window& get() { static Window w; return w; }
    static void init(int width, int height) {
        Window& w = get();
        w.width = width; w.height = height;
        w.rgb.assign((size_t)width * height * 3, 0.f);
    }
    static void clear() { std::fill(get().rgb.begin(), get().rgb.end(), 0.f); }
    static bool save(const char* path) {
        Window& w = get();
        std::string p(path);
        bool pfm = p.size() >= 4 && p.compare(p.size() - 4, 4, ".pfm") == 0;
        int st = pfm ? vcl_save_pfm(path, w.rgb.data(), w.width, w.height, 3) : vcl_save_bmp(path, w.rgb.data(), w.width, w.height, 3);
        return st == VCL_OK;
    }
};

struct GLbuffer {  // buffer.h:108-119
    unsigned int id = 0;
    float* gl_buffer = nullptr;  // host copy
    float* buffer = nullptr;     // device image
    int width = 0, height = 0, channel = 0;
    size_t Size() const { return (size_t)width * height * channel * sizeof(float); }
    static void init(GLbuffer& rb, float* buffer, int width, int height, int channel) {
        rb.width = width; rb.height = height; rb.channel = channel; rb.buffer = buffer;
        VCL_HIP_LAYER_CHECK(hipHostMalloc((void**)&rb.gl_buffer, rb.Size() ? rb.Size() : sizeof(float)));
    }
    static void draw(GLbuffer& rb, int internalformat, int format, float minX, float minY, float maxX, float maxY) {
        draw(rb, internalformat, format, 0.f, 0.f, 1.f, 1.f, minX, minY, maxX, maxY);
    }
    static void draw(GLbuffer& rb, int internalformat, int format, float texminX, float texminY, float texmaxX, float texmaxY, float minX,
                     float minY, float maxX, float maxY) {
        (void)internalformat;
        VCL_HIP_LAYER_CHECK(hipMemcpy(rb.gl_buffer, rb.buffer, rb.Size(), hipMemcpyDeviceToHost));
        Window& w = Window::get();
        if (w.width <= 0 || w.height <= 0 || rb.width <= 0 || rb.height <= 0) return;
        for (int py = 0; py < w.height; py++)
            for (int px = 0; px < w.width; px++) {
                float X = (px + 0.5f) * 2.f / w.width - 1.f, Y = (py + 0.5f) * 2.f / w.height - 1.f;
                if (X < minX || X >= maxX || Y < minY || Y >= maxY) continue;
                float s = texminX + (texmaxX - texminX) * (X - minX) / (maxX - minX);
                float t = texminY + (texmaxY - texminY) * (Y - minY) / (maxY - minY);
                int tx = (int)std::floor(s * rb.width), ty = (int)std::floor(t * rb.height);
                tx = tx < 0 ? 0 : tx >= rb.width ? rb.width - 1 : tx;
                ty = ty < 0 ? 0 : ty >= rb.height ? rb.height - 1 : ty;
                const float* v = rb.gl_buffer + ((size_t)ty * rb.width + tx) * rb.channel;
                float* o = w.rgb.data() + ((size_t)py * w.width + px) * 3;
                if (format == GL_RED || format == GL_LUMINANCE || rb.channel == 1) o[0] = o[1] = o[2] = v[0];
                else for (int i = 0; i < 3; i++) o[i] = i < rb.channel ? v[i] : 0.f;
            }
    }
};

// ------------------------------------------------------------------ Rotation / Camera (matrix.h)
// Host classes of matrix.h:1-96 over include/vcl_matrix.hpp (glm layout: vcl::quat stores
// x, y, z, w; vcl::mat4 is column-major).  Matrices live in host memory and are copied to the
// device `out` on forward, as in the reference (cudaMallocHost there, pinned host memory here).
struct RotationKernelParams {
    quat q;
    mat4* rotation = nullptr;
    float* out = nullptr;  // device 16 floats
};
struct RotationKernelGradParams {
    quat q;
    mat4* rotation = nullptr;
    float* out = nullptr;
};
struct RotationParams {
    RotationKernelParams kernel;
};
struct RotationGradParams : RotationParams {
    int it = 0;
    double rhom = 0, rhov = 0, rhomt = 0, rhovt = 0, eta = 0, eps = 0;
    quat m, v;
    RotationKernelGradParams grad;
    bool exact = false;  // true: d loss / d q of getRotation; false: the reference's half (matrix.cu:85-99)
};

inline mat4* host_mat4() {
    mat4* m = nullptr;
    VCL_HIP_LAYER_CHECK(hipHostMalloc((void**)&m, sizeof(mat4)));
    new (m) mat4();
    return m;
}

class Rotation {
public:
    // Rotation::step prints the gradient's length and its alignment with q every step
    // (matrix.cu:115-125); off unless set.
    static bool& verbose() { static bool v = false; return v; }
    static void init(RotationParams& r) {  // matrix.cu:3-7
        r.kernel.q = quat(1.f, 0.f, 0.f, 0.f);
        r.kernel.rotation = host_mat4();
        r.kernel.out = dmalloc<float>(16);
    }
    static void init(RotationGradParams& r) {  // matrix.cu:9-14
        init((RotationParams&)r);
        r.grad.q = quat(1.f, 0.f, 0.f, 0.f);
        r.grad.rotation = host_mat4();
        r.grad.out = dmalloc<float>(16);
    }
    static void setRandom(RotationParams& r) {  // matrix.cu:16-24
        float angle = ((float)std::rand() / (float)RAND_MAX * 2.f - 1.f) * 3.14159265f;
        float theta = ((float)std::rand() / (float)RAND_MAX * 2.f - 1.f) * 3.14159265f;
        float z = (float)std::rand() / (float)RAND_MAX * 2.f - 1.f;
        float c = std::sqrt(1.f - z * z);
        r.kernel.q = angle_axis(angle, normalize(vec3(c * std::cos(theta), c * std::sin(theta), z)));
    }
    static void setRotation(RotationParams& r, float angle, float x, float y, float z) {  // matrix.cu:27-30
        float s = std::sin(angle * .5f) / std::sqrt(x * x + y * y + z * z);
        r.kernel.q = quat(std::cos(angle * .5f), s * x, s * y, s * z);
    }
    static void addRotation(RotationParams& r, float angle, float x, float y, float z) {  // matrix.cu:32-35
        float s = std::sin(angle * .5f) / std::sqrt(x * x + y * y + z * z);
        r.kernel.q = quat(std::cos(angle * .5f), s * x, s * y, s * z) * r.kernel.q;
    }
    static void forward(RotationParams& r) {  // matrix.cu:55-58
        *r.kernel.rotation = rotation_matrix(r.kernel.q);
        VCL_HIP_LAYER_CHECK(hipMemcpy(r.kernel.out, r.kernel.rotation->data(), 16 * sizeof(float), hipMemcpyHostToDevice));
    }
    static void forward(RotationGradParams& r) {  // matrix.cu:60-65
        r.grad.q = quat();
        *r.grad.rotation = mat4::zero();
        dzero(r.grad.out, 16 * sizeof(float));
        forward((RotationParams&)r);
    }
    static void backward(RotationGradParams& r) {  // matrix.cu:85-99
        if (*r.grad.rotation == mat4::zero()) {
            VCL_HIP_LAYER_CHECK(hipStreamSynchronize(hs()));
            VCL_HIP_LAYER_CHECK(hipMemcpy(r.grad.rotation->data(), r.grad.out, 16 * sizeof(float), hipMemcpyDeviceToHost));
        }
        r.grad.q = rotation_backward(r.kernel.q, *r.grad.rotation, r.exact);
    }
    static void init(RotationGradParams& r, double eta, double rhom, double rhov, double eps) {  // matrix.cu:101-109
        r.it = 0; r.rhom = rhom; r.rhov = rhov; r.rhomt = rhom; r.rhovt = rhov; r.eta = eta; r.eps = eps;
    }
    static void step(RotationGradParams& r) {  // matrix.cu:111-137
        if (verbose()) {
            float d = std::sqrt(dot(r.grad.q, r.grad.q));
            if (d > 0) std::cout << "distance :" << d << " similar :" << dot(r.kernel.q, r.grad.q) / d << std::endl;
            else std::cout << "distance :" << d << std::endl;
        }
        RotationAdam a;
        a.it = r.it; a.rhom = r.rhom; a.rhov = r.rhov; a.rhomt = r.rhomt; a.rhovt = r.rhovt; a.eta = r.eta; a.eps = r.eps;
        a.m = r.m; a.v = r.v;
        r.kernel.q = rotation_adam_step(a, r.kernel.q, r.grad.q);
        r.it = a.it; r.rhomt = a.rhomt; r.rhovt = a.rhovt; r.m = a.m; r.v = a.v;
    }
    static void reset(RotationGradParams& r) {  // matrix.cu:139-144
        r.rhomt = r.rhom; r.rhovt = r.rhov;
        r.m = quat(); r.v = quat();
    }
};

struct CameraKernelParams {
    vec3 eye;
    quat q;
    float size = 0, aspect = 0, znear = 0, zfar = 0;
    mat4* rotation = nullptr;
    mat4* view = nullptr;
    mat4* projection = nullptr;
    mat4* mat = nullptr;
    float* out = nullptr;  // device P*V*R, 16 floats
};
struct CameraKernelGradParams {
    vec3 eye;
    quat q;
    float size = 0;
    mat4* rotation = nullptr;
    mat4* view = nullptr;
    mat4* projection = nullptr;
    mat4* mat = nullptr;
    float* out = nullptr;  // device dL/d mat, 16 floats
};
struct CameraParams {
    CameraKernelParams kernel;
};
struct CameraGradParams : CameraParams {
    CameraKernelGradParams grad;
    bool exact = false;  // true: the derivatives of P V R; false: the reference's products (matrix.cu:236-238)
};

class Camera {
public:
    static void init(CameraParams& c, RotationParams& rot, vec3 eye, vec3 center, vec3 up, float size, float aspect, float znear,
                     float zfar) {  // matrix.cu:147-158
        c.kernel.size = size; c.kernel.aspect = aspect; c.kernel.znear = znear; c.kernel.zfar = zfar;
        c.kernel.rotation = rot.kernel.rotation;
        c.kernel.view = host_mat4();
        c.kernel.projection = host_mat4();
        c.kernel.mat = host_mat4();
        c.kernel.out = dmalloc<float>(16);
        setCam(c, eye, center, up);
    }
    static void init(CameraGradParams& c, RotationParams& rot, vec3 eye, vec3 center, vec3 up, float size, float aspect, float znear,
                     float zfar) {  // matrix.cu:160-168
        init((CameraParams&)c, rot, eye, center, up, size, aspect, znear, zfar);
        c.grad.eye = vec3(); c.grad.size = 0.f;
        c.grad.view = host_mat4();
        c.grad.projection = host_mat4();
        c.grad.mat = host_mat4();
        c.grad.rotation = host_mat4();  // the reference leaves it unset for a plain rotation
        c.grad.out = dmalloc<float>(16);
        dzero(c.grad.out, 16 * sizeof(float));
    }
    static void init(CameraGradParams& c, RotationGradParams& rot, vec3 eye, vec3 center, vec3 up, float size, float aspect, float znear,
                     float zfar) {  // matrix.cu:170-173
        init(c, (RotationParams&)rot, eye, center, up, size, aspect, znear, zfar);
        c.grad.rotation = rot.grad.rotation;
    }
    static void setCam(CameraParams& c, vec3 eye, vec3 center, vec3 up) {  // matrix.cu:182-192
        c.kernel.eye = eye;
        c.kernel.q = look_at_quat(eye, center, up);
    }
    static void forward(CameraParams& c) {  // matrix.cu:209-222
        *c.kernel.view = view_matrix(c.kernel.q, c.kernel.eye);
        *c.kernel.projection = projection_matrix(c.kernel.size, c.kernel.aspect, c.kernel.znear, c.kernel.zfar);
        *c.kernel.mat = *c.kernel.projection * *c.kernel.view * *c.kernel.rotation;
        VCL_HIP_LAYER_CHECK(hipMemcpy(c.kernel.out, c.kernel.mat->data(), 16 * sizeof(float), hipMemcpyHostToDevice));
    }
    static void forward(CameraGradParams& c) {  // matrix.cu:224-232
        c.grad.eye = vec3(); c.grad.size = 0.f;
        *c.grad.view = mat4::zero(); *c.grad.projection = mat4::zero(); *c.grad.mat = mat4::zero();
        dzero(c.grad.out, 16 * sizeof(float));
        forward((CameraParams&)c);
    }
    static void backward(CameraGradParams& c) {  // matrix.cu:234-239
        VCL_HIP_LAYER_CHECK(hipStreamSynchronize(hs()));
        VCL_HIP_LAYER_CHECK(hipMemcpy(c.grad.mat->data(), c.grad.out, 16 * sizeof(float), hipMemcpyDeviceToHost));
        camera_backward(*c.grad.mat, *c.kernel.rotation, *c.kernel.view, *c.kernel.projection, c.exact, *c.grad.rotation, *c.grad.view,
                        *c.grad.projection);
    }
};

// ------------------------------------------------------------------ Project (project.h)
struct ProjectParams {
    vcl_project_params kernel{};
    unsigned int* vao = nullptr;
    int vaoNum = 0;
    size_t vecSize() const { return (size_t)kernel.vecNum * kernel.dimention * sizeof(float); }
    size_t vaoSize() const { return (size_t)vaoNum * 3 * sizeof(unsigned int); }
};
struct ProjectGradParams : ProjectParams {
    vcl_project_grad_params grad{};
};

class Project {
public:
    static void init(ProjectParams& p, float* mat, Attribute& vec, bool homogeneous) {  // project.cu:3-12
        if (vec.dimention != 3) { std::fprintf(stderr, "dimention is not 3\n"); std::exit(1); }
        p.kernel.vecNum = vec.vboNum;
        p.kernel.dimention = homogeneous ? 4 : 3;
        p.kernel.vec = vec.vbo;
        p.kernel.mat = mat;
        p.kernel.views = 1;
        p.vao = vec.vao;
        p.vaoNum = vec.vaoNum;
        p.kernel.out = dmalloc<float>((size_t)vec.vboNum * p.kernel.dimention);
    }
    static void init(ProjectGradParams& p, float* mat, AttributeGrad& vec, bool homogeneous) {
        init((ProjectParams&)p, mat, vec, homogeneous);
        float* g = dmalloc<float>((size_t)p.kernel.vecNum * p.kernel.dimention);
        p.grad.out = g;
        p.grad.vec = vec.grad;
        p.grad.mat = nullptr;
    }
    static void init(ProjectGradParams& p, float* mat, float* grad, AttributeGrad& vec, bool homogeneous) {
        init(p, mat, vec, homogeneous);
        p.grad.mat = grad;
    }
    static void init(ProjectGradParams& p, float* mat, float* grad, Attribute& vec, bool homogeneous) {
        init((ProjectParams&)p, mat, vec, homogeneous);
        p.grad.out = dmalloc<float>((size_t)p.kernel.vecNum * p.kernel.dimention);
        p.grad.vec = nullptr;
        p.grad.mat = grad;
    }
    // project.cu:14-23: the output is an Attribute sharing vec's index buffer
    static void init(ProjectParams& p, float* mat, Attribute& vec, Attribute& out, bool homogeneous) {
        if (vec.dimention != 3) { std::fprintf(stderr, "dimention is not 3\n"); std::exit(1); }
        Attribute::init(out, vec, homogeneous ? 4 : 3);
        p.kernel.vecNum = vec.vboNum;
        p.kernel.dimention = out.dimention;
        p.kernel.vec = vec.vbo;
        p.kernel.mat = mat;
        p.kernel.views = 1;
        p.vao = vec.vao;
        p.vaoNum = vec.vaoNum;
        p.kernel.out = out.vbo;
    }
    // project.cu:65-69: output grads in out.grad, positions' grads in vec.grad
    static void init(ProjectGradParams& p, float* mat, AttributeGrad& vec, AttributeGrad& out, bool homogeneous) {
        init((ProjectParams&)p, mat, vec, (Attribute&)out, homogeneous);
        out.grad = dmalloc<float>((size_t)out.vboNum * out.dimention);
        AttributeGrad::clear(out);
        p.grad.out = out.grad;
        p.grad.vec = vec.grad;
        p.grad.mat = nullptr;
    }
    static void forward(ProjectParams& p) { VCL_LAYER_CHECK(vcl_project_forward(&p.kernel, default_stream())); }
    static void forward(ProjectGradParams& p) {
        dzero((void*)p.grad.out, p.vecSize());
        forward((ProjectParams&)p);
    }
    static void backward(ProjectGradParams& p) { VCL_LAYER_CHECK(vcl_project_backward(&p.kernel, &p.grad, default_stream())); }
};

// ------------------------------------------------------------------ Rasterize (rasterize.h)
struct RasterizeParams {
    vcl_rasterize_params kernel{};
    int projNum = 0, idxNum = 0;
    // debug shading of wireframeinit / idhashinit: kernel.out then holds the colour image and
    // the raster's own output goes to these buffers
    int debugMode = 0;  // 0 none, 1 wireframe, 2 id hash
    float* debugRast = nullptr;
    float* debugDB = nullptr;
    size_t Size() const { return (size_t)kernel.width * kernel.height * kernel.depth * 4 * sizeof(float); }
};
struct RasterizeGradParams : RasterizeParams {
    vcl_rasterize_grad_params grad{};
    float* gradOut = nullptr;  // the P x 4 allocation behind grad.out (float2 per pixel used)
};

class Rasterize {
public:
    static void init(RasterizeParams& r, ProjectParams& proj, int width, int height, int depth, bool enableDB) {
        if (proj.kernel.dimention != 4) { std::fprintf(stderr, "dimention is not 4\n"); std::exit(1); }
        vcl_rasterize_params& k = r.kernel;
        k.width = width; k.height = height; k.depth = depth; k.enableDB = enableDB ? 1 : 0;
        k.xs = 2.f / (float)width; k.ys = 2.f / (float)height;
        k.proj = proj.kernel.out; k.idx = proj.vao;
        k.projNum = proj.kernel.vecNum; k.idxNum = proj.vaoNum;
        r.projNum = k.projNum; r.idxNum = k.idxNum;
        k.out = dmalloc<float>((size_t)width * height * depth * 4);
        k.outDB = enableDB ? dmalloc<float>((size_t)width * height * depth * 4) : nullptr;
        k.workspace_size = vcl_rasterize_workspace_size(width, height, depth, k.projNum, k.idxNum);
        k.workspace = dmalloc<char>(k.workspace_size);
    }
    static void init(RasterizeGradParams& r, ProjectGradParams& proj, int width, int height, int depth, bool enableDB) {
        init((RasterizeParams&)r, proj, width, height, depth, enableDB);
        r.gradOut = dmalloc<float>((size_t)width * height * depth * 4);
        r.grad.out = r.gradOut;
        r.grad.outDB = nullptr;
        r.grad.proj = const_cast<float*>(proj.grad.out);
    }
    static void forward(RasterizeParams& r) { VCL_LAYER_CHECK(vcl_rasterize_forward(&r.kernel, default_stream())); }
    static void forward(RasterizeGradParams& r) {
        dzero(r.gradOut, r.Size());
        forward((RasterizeParams&)r);
    }
    static void backward(RasterizeGradParams& r) { VCL_LAYER_CHECK(vcl_rasterize_backward(&r.kernel, &r.grad, default_stream())); }
    // rasterize.cpp:292-387 (wireframe) and :388-492 (id hash): kernel.out receives the shading
    static void wireframeinit(RasterizeParams& r, ProjectParams& proj, int width, int height) { debuginit(r, proj, width, height, 1); }
    static void idhashinit(RasterizeParams& r, ProjectParams& proj, int width, int height) { debuginit(r, proj, width, height, 2); }
    static void drawforward(RasterizeParams& r) {  // rasterize.cpp:494-512
        vcl_rasterize_params k = r.kernel;
        k.out = r.debugRast;
        k.outDB = r.debugDB;
        VCL_LAYER_CHECK(vcl_rasterize_forward(&k, default_stream()));
        if (r.debugMode == 1)
            VCL_LAYER_CHECK(vcl_rasterize_wireframe(k.width, k.height, k.depth, r.debugRast, r.debugDB, r.kernel.out, default_stream()));
        else
            VCL_LAYER_CHECK(vcl_rasterize_idhash(k.width, k.height, k.depth, r.debugRast, r.kernel.out, default_stream()));
    }

private:
    static void debuginit(RasterizeParams& r, ProjectParams& proj, int width, int height, int mode) {
        init(r, proj, width, height, 1, mode == 1);
        r.debugMode = mode;
        r.debugRast = r.kernel.out;
        r.debugDB = r.kernel.outDB;
        r.kernel.out = dmalloc<float>((size_t)width * height * 4);
    }
};

// ------------------------------------------------------------------ Interpolate (interpolate.h)
struct InterpolateParams {
    vcl_interpolate_params kernel{};
    size_t Size() const { return (size_t)kernel.width * kernel.height * kernel.depth * kernel.dimention * sizeof(float); }
};
struct InterpolateGradParams : InterpolateParams {
    vcl_interpolate_grad_params grad{};
    float* gradOut = nullptr;
    float* gradOutDA = nullptr;
};

class Interpolate {
public:
    static void init(InterpolateParams& i, RasterizeParams& r, Attribute& attr) {  // interpolate.cu:3-18
        vcl_interpolate_params& k = i.kernel;
        k.width = r.kernel.width; k.height = r.kernel.height; k.depth = r.kernel.depth;
        k.enableDA = r.kernel.enableDB;
        k.rast = r.kernel.out; k.rastDB = r.kernel.outDB;
        k.dimention = attr.dimention; k.attr = attr.vbo; k.idx = attr.vao; k.attrNum = attr.vboNum;
        k.out = dmalloc<float>(i.Size() / sizeof(float));
        k.outDA = k.enableDA ? dmalloc<float>(2 * i.Size() / sizeof(float)) : nullptr;
    }
    static void initGrad(InterpolateGradParams& i) {
        i.gradOut = dmalloc<float>(i.Size() / sizeof(float));
        i.gradOutDA = i.kernel.enableDA ? dmalloc<float>(2 * i.Size() / sizeof(float)) : nullptr;
        i.grad.out = i.gradOut;
        i.grad.outDA = i.gradOutDA;
    }
    static void init(InterpolateGradParams& i, RasterizeGradParams& r, AttributeGrad& attr) {
        init((InterpolateParams&)i, r, attr);
        initGrad(i);
        i.grad.attr = attr.grad;
        i.grad.rast = r.gradOut;
    }
    static void init(InterpolateGradParams& i, RasterizeGradParams& r, Attribute& attr) {
        init((InterpolateParams&)i, r, attr);
        initGrad(i);
        i.grad.attr = nullptr;
        i.grad.rast = r.gradOut;
    }
    static void init(InterpolateGradParams& i, RasterizeParams& r, AttributeGrad& attr) {
        init((InterpolateParams&)i, r, attr);
        initGrad(i);
        i.grad.attr = attr.grad;
        i.grad.rast = nullptr;
    }
    static void forward(InterpolateParams& i) { VCL_LAYER_CHECK(vcl_interpolate_forward(&i.kernel, default_stream())); }
    static void forward(InterpolateGradParams& i) {
        dzero(i.gradOut, i.Size());
        if (i.gradOutDA) dzero(i.gradOutDA, 2 * i.Size());
        forward((InterpolateParams&)i);
    }
    static void backward(InterpolateGradParams& i) { VCL_LAYER_CHECK(vcl_interpolate_backward(&i.kernel, &i.grad, default_stream())); }
};

// ------------------------------------------------------------------ Texturemap (texturemap.h)
struct TexturemapParams {
    vcl_texturemap_params kernel{};
    size_t Size() const { return (size_t)kernel.width * kernel.height * kernel.depth * kernel.channel * sizeof(float); }
};
struct TexturemapGradParams : TexturemapParams {
    vcl_texturemap_grad_params grad{};
    float* gradOut = nullptr;
};

class Texturemap {
public:
    static void init(TexturemapParams& t, RasterizeParams& r, InterpolateParams& intr, Texture& tex) {  // texturemap.cu:4-20
        vcl_texturemap_params& k = t.kernel;
        k.width = r.kernel.width; k.height = r.kernel.height; k.depth = r.kernel.depth;
        k.texwidth = tex.width; k.texheight = tex.height; k.channel = tex.channel; k.miplevel = tex.miplevel;
        k.rast = r.kernel.out; k.uv = intr.kernel.out; k.uvDA = intr.kernel.outDA;
        for (int l = 0; l < TEX_MAX_MIP_LEVEL; l++) k.texture[l] = l < tex.miplevel ? tex.texture[l] : nullptr;
        k.out = dmalloc<float>(t.Size() / sizeof(float));
    }
    static void initGrad(TexturemapGradParams& t) {
        t.gradOut = dmalloc<float>(t.Size() / sizeof(float));
        t.grad.out = t.gradOut;
        t.grad.daMode = VCL_TEX_GRAD_DA_REF;
    }
    static void init(TexturemapGradParams& t, RasterizeParams& r, InterpolateGradParams& intr, TextureGrad& tex) {
        init((TexturemapParams&)t, r, intr, tex);
        initGrad(t);
        t.grad.uv = intr.gradOut; t.grad.uvDA = intr.gradOutDA;
        for (int l = 0; l < TEX_MAX_MIP_LEVEL; l++) t.grad.texture[l] = l < tex.miplevel ? tex.grad[l] : nullptr;
    }
    static void init(TexturemapGradParams& t, RasterizeParams& r, InterpolateParams& intr, TextureGrad& tex) {
        init((TexturemapParams&)t, r, intr, tex);
        initGrad(t);
        t.grad.uv = nullptr; t.grad.uvDA = nullptr;
        for (int l = 0; l < TEX_MAX_MIP_LEVEL; l++) t.grad.texture[l] = l < tex.miplevel ? tex.grad[l] : nullptr;
    }
    static void init(TexturemapGradParams& t, RasterizeParams& r, InterpolateGradParams& intr, Texture& tex) {
        init((TexturemapParams&)t, r, intr, tex);
        initGrad(t);
        t.grad.uv = intr.gradOut; t.grad.uvDA = intr.gradOutDA;
        for (int l = 0; l < TEX_MAX_MIP_LEVEL; l++) t.grad.texture[l] = nullptr;
    }
    static void forward(TexturemapParams& t) { VCL_LAYER_CHECK(vcl_texturemap_forward(&t.kernel, default_stream())); }
    static void forward(TexturemapGradParams& t) {
        dzero(t.gradOut, t.Size());
        forward((TexturemapParams&)t);
    }
    static void backward(TexturemapGradParams& t) { VCL_LAYER_CHECK(vcl_texturemap_backward(&t.kernel, &t.grad, default_stream())); }
};

// ------------------------------------------------------------------ Antialias (antialias.h)
struct AntialiasParams {
    vcl_antialias_params kernel{};
    int projNum = 0;
    size_t Size() const { return (size_t)kernel.width * kernel.height * kernel.depth * kernel.channel * sizeof(float); }
};
struct AntialiasGradParams : AntialiasParams {
    vcl_antialias_grad_params grad{};
    float* gradOut = nullptr;
};

class Antialias {
public:
    static void init(AntialiasParams& a, RasterizeParams& r, ProjectParams& proj, float* in, int channel) {  // antialias.cu:3-15
        vcl_antialias_params& k = a.kernel;
        k.width = r.kernel.width; k.height = r.kernel.height; k.depth = r.kernel.depth; k.channel = channel;
        k.s[0] = r.kernel.width / 2.f; k.s[1] = r.kernel.height / 2.f;
        k.proj = proj.kernel.out; k.idx = proj.vao; k.rast = r.kernel.out; k.in = in;
        k.projNum = proj.kernel.vecNum;
        a.projNum = k.projNum;
        k.out = dmalloc<float>(a.Size() / sizeof(float));
    }
    static void init(AntialiasGradParams& a, RasterizeParams& r, ProjectGradParams& proj, float* in, float* grad, int channel) {
        init((AntialiasParams&)a, r, proj, in, channel);
        a.gradOut = dmalloc<float>(a.Size() / sizeof(float));
        a.grad.out = a.gradOut; a.grad.proj = const_cast<float*>(proj.grad.out); a.grad.in = grad;
    }
    static void init(AntialiasGradParams& a, RasterizeParams& r, ProjectGradParams& proj, float* in, int channel) {
        init(a, r, proj, in, nullptr, channel);
    }
    static void init(AntialiasGradParams& a, RasterizeParams& r, ProjectParams& proj, float* in, float* grad, int channel) {
        init((AntialiasParams&)a, r, proj, in, channel);
        a.gradOut = dmalloc<float>(a.Size() / sizeof(float));
        a.grad.out = a.gradOut; a.grad.proj = nullptr; a.grad.in = grad;
    }
    static void forward(AntialiasParams& a) { VCL_LAYER_CHECK(vcl_antialias_forward(&a.kernel, default_stream())); }
    static void forward(AntialiasGradParams& a) {
        dzero(a.gradOut, a.Size());
        forward((AntialiasParams&)a);
    }
    static void backward(AntialiasGradParams& a) {
        if (!a.grad.in) return;  // the reference writes grad.in unconditionally (antialias.cu:267-269)
        VCL_LAYER_CHECK(vcl_antialias_backward(&a.kernel, &a.grad, default_stream()));
    }
};

// ------------------------------------------------------------------ material (material.h)
struct NormalAxisParams {
    vcl_normal_axis_params kernel{};
    size_t Size() const { return (size_t)kernel.width * kernel.height * kernel.depth * 3 * sizeof(float); }
};

class NormalAxis {
public:
    static void init(NormalAxisParams& n, RotationParams& rot, RasterizeParams& r, Attribute& normal) {  // material.cu:3-12
        vcl_normal_axis_params& k = n.kernel;
        k.width = r.kernel.width; k.height = r.kernel.height; k.depth = r.kernel.depth;
        k.rast = r.kernel.out; k.rot = rot.kernel.out; k.normal = normal.vbo; k.normalidx = normal.vao;
        k.out = dmalloc<float>(n.Size() / sizeof(float));
    }
    // material.cu:14-23: the normal-map inputs are recorded; the reference's normal-map branch
    // is inverted (material.cu:40-48) and its effective output, rot * normalize(barycentric
    // normal), does not read them
    static void init(NormalAxisParams& n, RotationParams& rot, RasterizeParams& r, Attribute& normal, Attribute& pos, Attribute& texel,
                     TexturemapParams& normalmap);
    static void forward(NormalAxisParams& n) { VCL_LAYER_CHECK(vcl_normal_axis_forward(&n.kernel, default_stream())); }
};

inline void NormalAxis::init(NormalAxisParams& n, RotationParams& rot, RasterizeParams& r, Attribute& normal, Attribute& pos,
                             Attribute& texel, TexturemapParams& normalmap) {
    init(n, rot, r, normal);
    if (pos.dimention != 3) { std::fprintf(stderr, "dimention is not 3\n"); std::exit(1); }
    n.kernel.pos = pos.vbo; n.kernel.posidx = pos.vao;
    n.kernel.texel = texel.vbo; n.kernel.texelidx = texel.vao;
    if (normalmap.kernel.channel != 3) { std::fprintf(stderr, "dimention is not 3\n"); std::exit(1); }
    n.kernel.normalmap = normalmap.kernel.out;
}

struct ViewAxisParams {
    vcl_view_axis_params kernel{};
    size_t Size() const { return (size_t)kernel.width * kernel.height * kernel.depth * 3 * sizeof(float); }
};

class ViewAxis {
public:
    static void init(ViewAxisParams& v, RotationParams& rot, CameraParams& cam, RasterizeParams& r) {  // material.cu:81-90
        vcl_view_axis_params& k = v.kernel;
        k.width = r.kernel.width; k.height = r.kernel.height; k.depth = 1;
        k.rot = rot.kernel.out; k.view = cam.kernel.view->data(); k.projection = cam.kernel.projection->data();
        k.out = dmalloc<float>(v.Size() / sizeof(float));
    }
    static void forward(ViewAxisParams& v) { VCL_LAYER_CHECK(vcl_view_axis_forward(&v.kernel, default_stream())); }
};

struct SGSpecularParams {
    vcl_sg_specular_params kernel{};
    size_t Size(int d) const { return (size_t)kernel.width * kernel.height * kernel.depth * d * sizeof(float); }
};
struct SGSpecularGradParams : SGSpecularParams {
    vcl_sg_specular_grad_params grad{};
    float *gAxis = nullptr, *gSharp = nullptr, *gAmp = nullptr;
};

class SGSpecular {
public:
    static void init(SGSpecularParams& s, RasterizeParams& r, NormalAxisParams& n, ViewAxisParams& v, TexturemapParams& rough, float ior) {
        vcl_sg_specular_params& k = s.kernel;
        k.width = r.kernel.width; k.height = r.kernel.height; k.depth = r.kernel.depth;
        k.rast = r.kernel.out; k.normal = n.kernel.out; k.view = v.kernel.out; k.roughness = rough.kernel.out; k.ior = ior;
        k.outAxis = dmalloc<float>(s.Size(3) / sizeof(float));
        k.outAmplitude = dmalloc<float>(s.Size(1) / sizeof(float));
        k.outSharpness = dmalloc<float>(s.Size(1) / sizeof(float));
    }
    static void init(SGSpecularGradParams& s, RasterizeParams& r, NormalAxisParams& n, ViewAxisParams& v, TexturemapGradParams& rough, float ior) {
        init((SGSpecularParams&)s, r, n, v, rough, ior);
        s.gAxis = dmalloc<float>(s.Size(3) / sizeof(float));
        s.gAmp = dmalloc<float>(s.Size(1) / sizeof(float));
        s.gSharp = dmalloc<float>(s.Size(1) / sizeof(float));
        s.grad.outAxis = s.gAxis; s.grad.outAmplitude = s.gAmp; s.grad.outSharpness = s.gSharp;
        s.grad.roughness = rough.gradOut; s.grad.ior = ior;
    }
    static void forward(SGSpecularParams& s) { VCL_LAYER_CHECK(vcl_sg_specular_forward(&s.kernel, default_stream())); }
    static void forward(SGSpecularGradParams& s) {
        dzero(s.gAxis, s.Size(3)); dzero(s.gSharp, s.Size(1)); dzero(s.gAmp, s.Size(1));
        forward((SGSpecularParams&)s);
    }
    static void backward(SGSpecularGradParams& s) { VCL_LAYER_CHECK(vcl_sg_specular_backward(&s.kernel, &s.grad, default_stream())); }
};

struct SphericalGaussianParams {
    vcl_spherical_gaussian_params kernel{};
    size_t Size() const { return (size_t)kernel.width * kernel.height * kernel.depth * kernel.channel * sizeof(float); }
};
struct SphericalGaussianGradParams : SphericalGaussianParams {
    vcl_spherical_gaussian_grad_params grad{};
    float* gradOut = nullptr;
};

class SphericalGaussian {
public:
    static void init(SphericalGaussianParams& s, RasterizeParams& r, NormalAxisParams& n, TexturemapParams& diffuse, SGSpecularParams& spec,
                     SGBuffer& sg) {  // material.cu:260-283
        vcl_spherical_gaussian_params& k = s.kernel;
        k.width = r.kernel.width; k.height = r.kernel.height; k.depth = r.kernel.depth; k.channel = sg.channel;
        k.rast = r.kernel.out; k.sgnum = sg.num; k.axis = sg.axis; k.sharpness = sg.sharpness; k.amplitude = sg.amplitude;
        k.normal = n.kernel.out; k.diffuse = diffuse.kernel.out;
        k.specAxis = spec.kernel.outAxis; k.specSharpness = spec.kernel.outSharpness; k.specAmplitude = spec.kernel.outAmplitude;
        k.out = dmalloc<float>(s.Size() / sizeof(float));
        k.outDiffenv = dmalloc<float>(s.Size() / sizeof(float));
        k.outSpecenv = dmalloc<float>(s.Size() / sizeof(float));
    }
    static void initGrad(SphericalGaussianGradParams& s) {
        s.gradOut = dmalloc<float>(s.Size() / sizeof(float));
        s.grad.out = s.gradOut;
        s.grad.workspace_size = vcl_spherical_gaussian_workspace_size(s.kernel.width, s.kernel.height, s.kernel.depth, s.kernel.channel, s.kernel.sgnum);
        s.grad.workspace = dmalloc<char>(s.grad.workspace_size);
    }
    static void init(SphericalGaussianGradParams& s, RasterizeParams& r, NormalAxisParams& n, TexturemapParams& diffuse, SGSpecularParams& spec,
                     SGBufferGrad& sg) {
        init((SphericalGaussianParams&)s, r, n, diffuse, spec, (SGBuffer&)sg);
        initGrad(s);
        s.grad.axis = sg.axis; s.grad.sharpness = sg.sharpness; s.grad.amplitude = sg.amplitude;
    }
    static void init(SphericalGaussianGradParams& s, RasterizeParams& r, NormalAxisParams& n, TexturemapGradParams& diffuse, SGSpecularParams& spec,
                     SGBufferGrad& sg) {
        init(s, r, n, (TexturemapParams&)diffuse, spec, sg);
        s.grad.diffuse = diffuse.gradOut;
    }
    static void init(SphericalGaussianGradParams& s, RasterizeParams& r, NormalAxisParams& n, TexturemapGradParams& diffuse, SGSpecularParams& spec,
                     SGBuffer& sg) {
        init((SphericalGaussianParams&)s, r, n, diffuse, spec, sg);
        initGrad(s);
        s.grad.diffuse = diffuse.gradOut;
    }
    static void init(SphericalGaussianGradParams& s, RasterizeParams& r, NormalAxisParams& n, TexturemapGradParams& diffuse,
                     SGSpecularGradParams& spec, SGBuffer& sg) {
        init((SphericalGaussianParams&)s, r, n, diffuse, spec, sg);
        initGrad(s);
        s.grad.diffuse = diffuse.gradOut;
        s.grad.specAxis = spec.gAxis; s.grad.specSharpness = spec.gSharp; s.grad.specAmplitude = spec.gAmp;
    }
    static void forward(SphericalGaussianParams& s) { VCL_LAYER_CHECK(vcl_spherical_gaussian_forward(&s.kernel, default_stream())); }
    static void forward(SphericalGaussianGradParams& s) {
        dzero(s.gradOut, s.Size());
        forward((SphericalGaussianParams&)s);
    }
    static void backward(SphericalGaussianGradParams& s) {
        VCL_LAYER_CHECK(vcl_spherical_gaussian_backward(&s.kernel, &s.grad, default_stream()));
    }
};

// ------------------------------------------------------------------ Loss (loss.h)
struct LossParams {
    vcl_loss_params kernel{};
    float loss = 0.f;
    size_t Size() const { return (size_t)kernel.width * kernel.height * kernel.depth * sizeof(float); }
};

class Loss {
public:
    // loss.cu:3-20; the scratch `buffer` is sized for every loss of this header (MSE partials,
    // the SSIM maps, the texture-loss map)
    static void init(LossParams& l, float* target, float* predict, float* grad, int width, int height, int depth) {
        vcl_loss_params& k = l.kernel;
        k.target = target; k.predict = predict; k.grad = grad;
        k.width = width; k.height = height; k.depth = depth; k.size = width * height * depth;
        size_t ws = std::max({vcl_loss_workspace_size(k.size), vcl_ssim_workspace_size(width, height, depth),
                              vcl_texture_loss_workspace_size(k.size)});
        k.buffer = (float*)dmalloc<char>(ws);
        k.loss = dmalloc<float>(1);
    }
    // The reference returns the host copy made every step (loss.cu:55); here it is copied on demand.
    static float loss(LossParams& l) {
        VCL_HIP_LAYER_CHECK(hipMemcpyAsync(&l.loss, l.kernel.loss, sizeof(float), hipMemcpyDeviceToHost, hs()));
        VCL_HIP_LAYER_CHECK(hipStreamSynchronize(hs()));
        return l.loss;
    }
};

class MSELoss : Loss {
public:
    static void backward(LossParams& l) { VCL_LAYER_CHECK(vcl_mse_loss_backward(&l.kernel, default_stream())); }
    static void textureloss(LossParams& l) { VCL_LAYER_CHECK(vcl_mse_texture_loss(&l.kernel, default_stream())); }  // loss.cu:75-91
};

class SSIMLoss : Loss {
public:
    // VCL_SSIM_GRAD_REF: the reference's gradient formula (loss.cu:146-149); VCL_SSIM_GRAD_EXACT:
    // the derivative of sum(1 - SSIM)
    static int& gradMode() { static int m = VCL_SSIM_GRAD_REF; return m; }
    static void backward(LossParams& l) { VCL_LAYER_CHECK(vcl_ssim_loss_backward(&l.kernel, gradMode(), default_stream())); }
};

// ------------------------------------------------------------------ Optimizers (optimize.h)
struct OptimizeParams {
    vcl_optimize_params kernel{};
    int it = 1;
    size_t Size() const { return (size_t)kernel.size * sizeof(float); }
};
struct AdamParams : OptimizeParams {
    vcl_adam_hyper_params hyper{};
};
struct SGDParams : OptimizeParams {
    vcl_sgd_hyper_params hyper{};
};
struct NadamParams : OptimizeParams {
    vcl_nadam_hyper_params hyper{};
};

class Optimizer {
public:
    static void init(OptimizeParams& o, float* param, float* grad, int size, int width, int height, int depth) {  // optimize.cu:3-13
        o.kernel.param = param; o.kernel.grad = grad; o.kernel.size = size;
        o.kernel.width = width; o.kernel.height = height; o.kernel.depth = depth;
        o.it = 1;
    }
    static void init(OptimizeParams& o, BufferGrad& b) { init(o, b.buffer, b.grad, b.num * b.dimention, b.num, b.dimention, 1); }  // optimize.cu:15-18
    static void init(OptimizeParams& o, AttributeGrad& a) { init(o, a.vbo, a.grad, a.vboNum * a.dimention, a.vboNum, a.dimention, 1); }
    static void init(OptimizeParams& o, TextureGrad& t) {  // level 0 only (optimize.cu:25-28)
        init(o, t.texture[0], t.grad[0], t.width * t.height * t.channel, t.width, t.height, t.channel);
    }
    static void randomParams(OptimizeParams& o, float min, float max) {  // optimize.cu:40-44
        VCL_LAYER_CHECK(vcl_optimize_random_params(&o.kernel, draw_seed(), min, max, default_stream()));
    }
    static void clampParams(OptimizeParams& o, float min, float max) {  // optimize.cu:56-60
        VCL_LAYER_CHECK(vcl_optimize_clamp_params(&o.kernel, min, max, default_stream()));
    }
};

class SGD : Optimizer {
public:
    static void setHyperParams(SGDParams& s, double eta) { s.hyper.eta = eta; }  // optimize.cu:64-66
    static void step(SGDParams& s) {  // optimize.cu:78-82
        s.it++;
        VCL_LAYER_CHECK(vcl_sgd_step(&s.kernel, &s.hyper, default_stream()));
    }
};

class Adam : Optimizer {
public:
    static void setHyperParams(AdamParams& a, double eta, double rhom, double rhov, double eps) {  // optimize.cu:84-93
        a.hyper.rhom = rhom; a.hyper.rhov = rhov; a.hyper.rhomt = rhom; a.hyper.rhovt = rhov;
        a.hyper.eta = eta; a.hyper.eps = eps;
        a.hyper.m = dmalloc<float>(a.kernel.size);
        a.hyper.v = dmalloc<float>(a.kernel.size);
        dzero(a.hyper.m, a.Size());  // never zeroed in the reference (A-10)
        dzero(a.hyper.v, a.Size());
    }
    static void step(AdamParams& a) {  // optimize.cu:109-115
        a.it++;
        a.hyper.rhomt *= a.hyper.rhom;
        a.hyper.rhovt *= a.hyper.rhov;
        VCL_LAYER_CHECK(vcl_adam_step(&a.kernel, &a.hyper, default_stream()));
    }
};

class Nadam : Optimizer {
public:
    static void init(NadamParams& n, float* param, float* grad, int size, int width, int height, int depth) {  // optimize.cu:118-120
        Optimizer::init(n, param, grad, size, width, height, depth);
    }
    static void init(NadamParams& n, Attribute& attr, float* grad) {  // optimize.cu:122-125
        Optimizer::init(n, attr.vbo, grad, attr.vboNum * attr.dimention, attr.vboNum, attr.dimention, 1);
    }
    // optimize.cu:126-141: the momentum schedule mu_t = mu (1 - 0.96^(0.004 t) / 2) starts at
    // t = 1, 2 here and step() advances it before each launch
    static void setHyperParams(NadamParams& n, double alpha, double mu, double rho, double eps) {
        vcl_nadam_hyper_params& h = n.hyper;
        h.mupow = std::pow(0.96, 0.004);
        h.mupowt = h.mupow;
        h.mu = mu;
        h.mu0 = mu * (1 - .5 * h.mupowt);
        h.mupowt *= h.mupow;
        h.mu1 = mu * (1 - .5 * h.mupowt);
        h.rho = rho;
        h.alpha = alpha;
        h.mut0 = h.mu0;
        h.mut1 = h.mu0 * h.mu1;
        h.rhot = rho;
        h.eps = eps;
        h.m = dmalloc<float>(n.kernel.size);
        h.v = dmalloc<float>(n.kernel.size);
        dzero(h.m, n.Size());  // never zeroed in the reference
        dzero(h.v, n.Size());
    }
    static void step(NadamParams& n) {  // optimize.cu:157-167
        vcl_nadam_hyper_params& h = n.hyper;
        n.it++;
        h.mu0 = h.mu1;
        h.mut0 = h.mut1;
        h.mupowt *= h.mupow;
        h.mu1 = h.mu * (1 - .5 * h.mupowt);
        h.mut1 *= h.mu1;
        h.rhot *= h.rho;
        VCL_LAYER_CHECK(vcl_nadam_step(&n.kernel, &n.hyper, default_stream()));
    }
};

// ------------------------------------------------------------------ ComputeNormal (computenormal.h)
struct ComputeNormalParams {
    vcl_compute_normal_params kernel{};
    size_t posSize() const { return (size_t)kernel.posNum * 3 * sizeof(float); }
    size_t vaoSize() const { return (size_t)kernel.vaoNum * 3 * sizeof(unsigned int); }
};
struct ComputeNormalGradParams : ComputeNormalParams {
    vcl_compute_normal_grad_params grad{};
};

class ComputeNormal {
public:
    // VCL_NORMAL_ANGLE_REF: the reference's corner weight acos(e1 . e2) of the raw edges;
    // VCL_NORMAL_ANGLE_EXACT: the corner's interior angle
    static int& angleMode() { static int m = VCL_NORMAL_ANGLE_REF; return m; }
    static void init(ComputeNormalParams& n, Attribute& pos, Attribute& normal) {  // computenormal.cu:3-15
        if (pos.dimention != 3) { std::fprintf(stderr, "dimention is not 3\n"); std::exit(1); }
        Attribute::init(normal, pos, 3);
        vcl_compute_normal_params& k = n.kernel;
        k.posNum = pos.vboNum; k.pos = pos.vbo; k.vao = pos.vao; k.vaoNum = pos.vaoNum;
        k.mat = nullptr; k.out = normal.vbo;
        k.sum = dmalloc<float>((size_t)pos.vboNum * 3);
        k.angleMode = angleMode();
    }
    static void init(ComputeNormalGradParams& n, AttributeGrad& pos, AttributeGrad& normal) {  // computenormal.cu:50-54
        init((ComputeNormalParams&)n, pos, (Attribute&)normal);
        normal.grad = dmalloc<float>((size_t)normal.vboNum * 3);
        AttributeGrad::clear(normal);
        n.grad.out = normal.grad;
        n.grad.pos = pos.grad;
    }
    static void forward(ComputeNormalParams& n) { VCL_LAYER_CHECK(vcl_compute_normal_forward(&n.kernel, default_stream())); }
    static void forward(ComputeNormalGradParams& n) {  // computenormal.cu:45-48
        dzero((void*)n.grad.out, n.posSize());
        forward((ComputeNormalParams&)n);
    }
    // the derivative w.r.t. the positions (a stub in the reference, computenormal.cu:56-65)
    static void backward(ComputeNormalGradParams& n) { VCL_LAYER_CHECK(vcl_compute_normal_backward(&n.kernel, &n.grad, default_stream())); }
};

// ------------------------------------------------------------------ Filter (filter.h)
struct FilterParams {
    float h_sig = 0.f;
    vcl_filter_params kernel{};
    size_t Size() const { return (size_t)kernel.width * kernel.height * kernel.depth * kernel.channel * sizeof(float); }
};
struct FilterGradParams : FilterParams {
    vcl_filter_grad_params grad{};
};

class Filter {
public:
    static void init(FilterParams& f, RasterizeParams& r, float* in, int channel, float sigma) {  // filter.cu:4-16
        vcl_filter_params& k = f.kernel;
        k.width = r.kernel.width; k.height = r.kernel.height; k.depth = r.kernel.depth; k.channel = channel;
        k.in = in;
        f.h_sig = sigma;
        float* sd = dmalloc<float>(1);
        VCL_HIP_LAYER_CHECK(hipMemcpy(sd, &sigma, sizeof(float), hipMemcpyHostToDevice));
        k.sigma = sd;
        k.filter = dmalloc<float>(256);
        k.out = dmalloc<float>(f.Size() / sizeof(float));
        k.buf = dmalloc<float>(f.Size() / sizeof(float));
    }
    static void init(FilterGradParams& f, RasterizeParams& r, float* in, float* grad, int channel, float sigma) {  // filter.cu:95-104
        init((FilterParams&)f, r, in, channel, sigma);
        f.grad.sigma = dmalloc<float>(1);
        f.grad.filter = dmalloc<float>(256);
        f.grad.out = dmalloc<float>(f.Size() / sizeof(float));
        f.grad.buf = dmalloc<float>(f.Size() / sizeof(float));
        f.grad.bufx = dmalloc<float>(f.Size() / sizeof(float));
        f.grad.bufy = dmalloc<float>(f.Size() / sizeof(float));
        f.grad.in = grad;
    }
    static void forward(FilterParams& f) { VCL_LAYER_CHECK(vcl_filter_forward(&f.kernel, default_stream())); }
    static void forward(FilterGradParams& f) {  // filter.cu:90-93
        dzero((void*)f.grad.out, f.Size());
        forward((FilterParams&)f);
    }
    // grad.in stored, *grad.sigma = d loss / d sigma (exact)
    static void backward(FilterGradParams& f) { VCL_LAYER_CHECK(vcl_filter_backward(&f.kernel, &f.grad, default_stream())); }
};

// ------------------------------------------------------------------ Noise (extention.h)
struct NoiseParams {
    vcl_noise_params kernel{};
    size_t Size() const { return (size_t)kernel.width * kernel.height * kernel.depth * kernel.channel * sizeof(float); }
};

class Noise {
public:
    static void init(NoiseParams& n, RasterizeParams& r, float* in, int channel, float intensity) {  // extention.cu:3-11
        vcl_noise_params& k = n.kernel;
        k.width = r.kernel.width; k.height = r.kernel.height; k.depth = r.kernel.depth; k.channel = channel;
        k.intensity = intensity; k.in = in;
        k.out = dmalloc<float>(n.Size() / sizeof(float));
    }
    static void forward(NoiseParams& n) { VCL_LAYER_CHECK(vcl_noise_forward(&n.kernel, draw_seed(), default_stream())); }  // extention.cu:24-30
};

}  // namespace vcl
