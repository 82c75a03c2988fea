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
// device scalar down only when asked.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "vcl.h"

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

// ------------------------------------------------------------------ buffers (buffer.h)
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
    // gradSumup (buffer.cu:638-648); exact = true scales each folded level by 1/4
    static void gradSumup(TextureGrad& t, bool exact = true) {
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
    }
    static void copy(SGBuffer& dst, const float* axis, const float* sharpness, const float* amplitude) {  // buffer.cu:660-667
        if (axis) VCL_HIP_LAYER_CHECK(hipMemcpy(dst.axis, axis, (size_t)dst.num * 3 * sizeof(float), hipMemcpyHostToDevice));
        if (sharpness) VCL_HIP_LAYER_CHECK(hipMemcpy(dst.sharpness, sharpness, (size_t)dst.num * sizeof(float), hipMemcpyHostToDevice));
        if (amplitude) VCL_HIP_LAYER_CHECK(hipMemcpy(dst.amplitude, amplitude, (size_t)dst.num * dst.channel * sizeof(float), hipMemcpyHostToDevice));
    }
};

// SGBufferGrad redeclares axis/sharpness/amplitude: base = parameters, derived = gradients
// (buffer.h:96-100, A-16); ((SGBuffer&)g).axis is the parameter, g.axis the gradient.
struct SGBufferGrad : SGBuffer {
    float* axis = nullptr;
    float* sharpness = nullptr;
    float* amplitude = nullptr;
    static void init(SGBufferGrad& s, int num, int channel) {
        SGBuffer::init(s, num, channel);
        s.axis = dmalloc<float>((size_t)num * 3);
        s.sharpness = dmalloc<float>(num);
        s.amplitude = dmalloc<float>((size_t)num * channel);
        clear(s);
    }
    static void clear(SGBufferGrad& s) {
        dzero(s.axis, (size_t)s.num * 3 * sizeof(float));
        dzero(s.sharpness, (size_t)s.num * sizeof(float));
        dzero(s.amplitude, (size_t)s.num * s.channel * sizeof(float));
    }
};

// ------------------------------------------------------------------ host matrices (matrix.h)
// Minimal stand-ins: the ops need the model rotation on the device (NormalAxis) and the
// camera's view/projection on the host (ViewAxis); column-major float[16] as glm stores them.
struct RotationParams {
    struct { float* out = nullptr; } kernel;  // device 16 floats
    static void init(RotationParams& r, const float* host16) {
        r.kernel.out = dmalloc<float>(16);
        VCL_HIP_LAYER_CHECK(hipMemcpy(r.kernel.out, host16, 16 * sizeof(float), hipMemcpyHostToDevice));
    }
};

struct CameraParams {
    struct {
        float view[16];
        float projection[16];
        float* out = nullptr;  // device P*V*R, 16 floats
    } kernel;
    static void init(CameraParams& c, const float* view16, const float* proj16, const float* mat16) {
        std::memcpy(c.kernel.view, view16, sizeof(c.kernel.view));
        std::memcpy(c.kernel.projection, proj16, sizeof(c.kernel.projection));
        c.kernel.out = dmalloc<float>(16);
        VCL_HIP_LAYER_CHECK(hipMemcpy(c.kernel.out, mat16, 16 * sizeof(float), hipMemcpyHostToDevice));
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
    static void forward(NormalAxisParams& n) { VCL_LAYER_CHECK(vcl_normal_axis_forward(&n.kernel, default_stream())); }
};

struct ViewAxisParams {
    vcl_view_axis_params kernel{};
    size_t Size() const { return (size_t)kernel.width * kernel.height * kernel.depth * 3 * sizeof(float); }
};

class ViewAxis {
public:
    static void init(ViewAxisParams& v, RotationParams& rot, CameraParams& cam, RasterizeParams& r) {  // material.cu:81-90
        vcl_view_axis_params& k = v.kernel;
        k.width = r.kernel.width; k.height = r.kernel.height; k.depth = 1;
        k.rot = rot.kernel.out; k.view = cam.kernel.view; k.projection = cam.kernel.projection;
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
    static void init(LossParams& l, float* target, float* predict, float* grad, int width, int height, int depth) {  // loss.cu:3-20
        vcl_loss_params& k = l.kernel;
        k.target = target; k.predict = predict; k.grad = grad;
        k.width = width; k.height = height; k.depth = depth; k.size = width * height * depth;
        k.buffer = (float*)dmalloc<char>(vcl_loss_workspace_size(k.size));
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
};

// ------------------------------------------------------------------ Optimizer / Adam (optimize.h)
struct OptimizeParams {
    vcl_optimize_params kernel{};
    int it = 1;
    size_t Size() const { return (size_t)kernel.size * sizeof(float); }
};
struct AdamParams : OptimizeParams {
    vcl_adam_hyper_params hyper{};
};

class Optimizer {
public:
    static void init(OptimizeParams& o, float* param, float* grad, int size, int width, int height, int depth) {  // optimize.cu:3-13
        o.kernel.param = param; o.kernel.grad = grad; o.kernel.size = size;
        o.kernel.width = width; o.kernel.height = height; o.kernel.depth = depth;
        o.it = 1;
    }
    static void init(OptimizeParams& o, AttributeGrad& a) { init(o, a.vbo, a.grad, a.vboNum * a.dimention, a.vboNum, a.dimention, 1); }
    static void init(OptimizeParams& o, TextureGrad& t) {  // level 0 only (optimize.cu:25-28)
        init(o, t.texture[0], t.grad[0], t.width * t.height * t.channel, t.width, t.height, t.channel);
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

}  // namespace vcl
