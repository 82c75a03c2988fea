// preset_fit.cpp -- host program written against include/vcl_layers.hpp in the call order of
// the reference preset (preset_primitives.cpp:56-116 init, :160-207 update), extended with a
// trainable mesh / SG lobes: proves the drop-in layer surface end to end.
//
// usage: preset_fit <scene.bin> <out.bin> <steps>
// scene.bin (little endian): int32 W,H,V,T,tw,th,C,L,K; float32 pos[V*3], normals[V*3], uv[V*2];
//   uint32 idx[T*3]; float32 tex0[tw*th*C], axis[K*3], sharp[K], amp[K*C], camera[10] (eye 3,
//   center 3, size, aspect, znear, zfar), target[H*W*C]; float32 rough, ior
// out.bin: float32 loss per step, then from the first step: image[H*W*C], the gradients before
//   the updates -- positions [V*3], texture level 0 after gradSumup [tw*th*C], lobe axis [K*3],
//   sharpness [K], amplitude [K*C] -- then pos[V*3] after the last step
#include <cstdio>
#include <vector>

#include "../../include/vcl_layers.hpp"

using namespace vcl;

template <class T>
static std::vector<T> rd(FILE* f, size_t n) {
    std::vector<T> v(n);
    if (n && std::fread(v.data(), sizeof(T), n, f) != n) { std::fprintf(stderr, "short read\n"); std::exit(2); }
    return v;
}

int main(int argc, char** argv) {
    if (argc < 4) { std::fprintf(stderr, "usage: %s scene.bin out.bin steps\n", argv[0]); return 2; }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) { std::perror("scene"); return 2; }
    auto hdr = rd<int>(f, 9);
    int W = hdr[0], H = hdr[1], V = hdr[2], T = hdr[3], tw = hdr[4], th = hdr[5], C = hdr[6], L = hdr[7], K = hdr[8];
    auto pos = rd<float>(f, (size_t)V * 3), normals = rd<float>(f, (size_t)V * 3), uv = rd<float>(f, (size_t)V * 2);
    auto idx = rd<unsigned>(f, (size_t)T * 3);
    auto tex0 = rd<float>(f, (size_t)tw * th * C), axis = rd<float>(f, (size_t)K * 3), sharp = rd<float>(f, K), amp = rd<float>(f, (size_t)K * C);
    auto camera = rd<float>(f, 10);
    auto target = rd<float>(f, (size_t)W * H * C);
    auto tail = rd<float>(f, 2);
    std::fclose(f);
    int steps = std::atoi(argv[3]);

    // ---- init (preset_primitives.cpp:7-116 pattern)
    AttributeGrad pos_a;
    AttributeGrad::init(pos_a, V, T, 3);
    Attribute::copy(pos_a, pos.data());
    Attribute::copyIndex(pos_a, idx.data());
    Attribute normal_a, texel_a;
    Attribute::init(normal_a, pos_a, 3);
    Attribute::copy(normal_a, normals.data());
    Attribute::init(texel_a, pos_a, 2);
    Attribute::copy(texel_a, uv.data());
    TextureGrad diffusemap;
    TextureGrad::init(diffusemap, tw, th, C, L);
    Texture::setLevel0(diffusemap, tex0.data());
    Texture roughness;
    Texture::init(roughness, 1, 1, 1, 1);
    float r0 = tail[0];
    VCL_HIP_LAYER_CHECK(hipMemcpy(roughness.texture[0], &r0, sizeof(float), hipMemcpyHostToDevice));
    SGBufferGrad sgbuf;
    SGBufferGrad::init(sgbuf, K, C);
    SGBuffer::copy(sgbuf, axis.data(), sharp.data(), amp.data());

    // model rotation and camera (matrix.h), as preset_primitives.cpp:56-58 sets them up
    RotationParams rot;
    Rotation::init(rot);
    CameraParams cam;
    Camera::init(cam, rot, vec3(camera[0], camera[1], camera[2]), vec3(camera[3], camera[4], camera[5]), vec3(0.f, 1.f, 0.f), camera[6],
                 camera[7], camera[8], camera[9]);
    Rotation::forward(rot);
    Camera::forward(cam);

    ProjectGradParams proj_p;
    Project::init(proj_p, cam.kernel.out, pos_a, true);
    RasterizeGradParams rast;
    Rasterize::init(rast, proj_p, W, H, 1, true);
    InterpolateGradParams intr;
    Interpolate::init(intr, rast, texel_a);
    NormalAxisParams normal_axis;
    NormalAxis::init(normal_axis, rot, rast, normal_a);
    ViewAxisParams view_axis;
    ViewAxis::init(view_axis, rot, cam, rast);
    TexturemapGradParams diff;
    Texturemap::init(diff, rast, intr, diffusemap);
    // constant roughness map: a per-pixel buffer filled once (the configs give no roughness texture)
    TexturemapParams rough_map;
    rough_map.kernel.width = W; rough_map.kernel.height = H; rough_map.kernel.depth = 1; rough_map.kernel.channel = 1;
    {
        std::vector<float> rm((size_t)W * H, r0);
        rough_map.kernel.out = dmalloc<float>(rm.size());
        VCL_HIP_LAYER_CHECK(hipMemcpy(rough_map.kernel.out, rm.data(), rm.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    SGSpecularParams spec;
    SGSpecular::init(spec, rast, normal_axis, view_axis, rough_map, tail[1]);
    SphericalGaussianGradParams sg;
    SphericalGaussian::init(sg, rast, normal_axis, diff, spec, sgbuf);
    AntialiasGradParams aa;
    Antialias::init(aa, rast, proj_p, sg.kernel.out, sg.gradOut, C);
    float* target_d = dmalloc<float>(target.size());
    VCL_HIP_LAYER_CHECK(hipMemcpy(target_d, target.data(), target.size() * sizeof(float), hipMemcpyHostToDevice));
    LossParams loss;
    Loss::init(loss, target_d, aa.kernel.out, aa.gradOut, W, H, C);

    AdamParams adam_tex, adam_pos, adam_axis, adam_sharp, adam_amp;
    Optimizer::init(adam_tex, diffusemap);
    Adam::setHyperParams(adam_tex, 1e-2, .9, .999, 1e-8);
    Optimizer::init(adam_pos, pos_a);
    Adam::setHyperParams(adam_pos, 1e-2, .9, .999, 1e-8);
    Optimizer::init(adam_axis, ((SGBuffer&)sgbuf).axis, sgbuf.axis, K * 3, K, 3, 1);
    Adam::setHyperParams(adam_axis, 1e-2, .9, .999, 1e-8);
    Optimizer::init(adam_sharp, ((SGBuffer&)sgbuf).sharpness, sgbuf.sharpness, K, K, 1, 1);
    Adam::setHyperParams(adam_sharp, 1e-2, .9, .999, 1e-8);
    Optimizer::init(adam_amp, ((SGBuffer&)sgbuf).amplitude, sgbuf.amplitude, K * C, K, C, 1);
    Adam::setHyperParams(adam_amp, 1e-2, .9, .999, 1e-8);

    std::vector<float> losses, first_img((size_t)W * H * C), g_pos((size_t)V * 3), g_tex((size_t)tw * th * C), g_axis((size_t)K * 3),
        g_sharp(K), g_amp((size_t)K * C);
    auto down = [](std::vector<float>& h, const float* d) {
        VCL_HIP_LAYER_CHECK(hipMemcpy(h.data(), d, h.size() * sizeof(float), hipMemcpyDeviceToHost));
    };
    for (int s = 0; s < steps; s++) {
        // ---- update (preset_primitives.cpp:167-207 call order)
        Project::forward(proj_p);
        Rasterize::forward(rast);
        Interpolate::forward(intr);
        NormalAxis::forward(normal_axis);
        ViewAxis::forward(view_axis);
        Texturemap::forward(diff);
        SGSpecular::forward(spec);
        SphericalGaussian::forward(sg);
        Antialias::forward(aa);
        MSELoss::backward(loss);
        Antialias::backward(aa);
        SphericalGaussian::backward(sg);
        Texturemap::backward(diff);
        Interpolate::backward(intr);
        Rasterize::backward(rast);
        Project::backward(proj_p);
        TextureGrad::gradSumup(diffusemap);
        if (s == 0) {
            down(first_img, aa.kernel.out); down(g_pos, pos_a.grad); down(g_tex, diffusemap.grad[0]);
            down(g_axis, sgbuf.axis); down(g_sharp, sgbuf.sharpness); down(g_amp, sgbuf.amplitude);
        }
        Adam::step(adam_tex);
        TextureGrad::clear(diffusemap);
        Texture::clamp(diffusemap, 0.f, 1.f);
        Adam::step(adam_pos);
        AttributeGrad::clear(pos_a);
        Adam::step(adam_amp);
        Adam::step(adam_sharp);
        Adam::step(adam_axis);
        SGBufferGrad::clear(sgbuf);
        losses.push_back(Loss::loss(loss));
    }
    std::vector<float> posout((size_t)V * 3);
    VCL_HIP_LAYER_CHECK(hipMemcpy(posout.data(), pos_a.vbo, posout.size() * sizeof(float), hipMemcpyDeviceToHost));
    FILE* o = std::fopen(argv[2], "wb");
    if (!o) { std::perror("out"); return 2; }
    std::fwrite(losses.data(), sizeof(float), losses.size(), o);
    for (auto* v : {&first_img, &g_pos, &g_tex, &g_axis, &g_sharp, &g_amp, &posout}) std::fwrite(v->data(), sizeof(float), v->size(), o);
    std::fclose(o);
    std::printf("preset_fit: %d steps, first loss %.6e last loss %.6e\n", steps, losses.empty() ? 0.f : losses.front(),
                losses.empty() ? 0.f : losses.back());
    return 0;
}
