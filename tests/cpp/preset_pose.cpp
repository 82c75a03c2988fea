// preset_pose.cpp -- host program on include/vcl_layers.hpp that runs the reference's pose preset
// (preset_pose.cpp:7-49 init, :63-97 update): a mesh loaded from OBJ, coloured by its positions,
// rendered from a camera whose model rotation is fitted to a target rotation through the SSIM
// loss, the rotation's gradient flowing through Project -> Camera -> Rotation.  The GLUT window
// becomes the headless vcl::Window (GLbuffer::draw composites, Window::save writes a BMP).
//
// usage: preset_pose <mesh.obj> <out.bin> <steps> <exact 0|1> <image.bmp> [offset [seed]]
// exact 1: the exact Camera / Rotation derivatives and the exact SSIM gradient (the derivative of
//   the loss the program reports) instead of the reference's formulas.
// offset (radians, optional): the fitted rotation starts at the target's rotated by this angle
//   about a fixed axis instead of at a second random draw.
// seed (optional): the rand() seed of the random rotations (default 12345).
// out.bin: float32 loss per step; then the first step's quaternions (x, y, z, w) of the target
//   and of the fitted rotation; then the last rendered image [512*512*3] and the target image;
//   then the fitted rotation after the last step; then the rotation's gradient (x, y, z, w) of
//   the first step.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/vcl_layers.hpp"

using namespace vcl;

int main(int argc, char** argv) {
    if (argc < 6) { std::fprintf(stderr, "usage: %s mesh.obj out.bin steps exact image.bmp [offset [seed]]\n", argv[0]); return 2; }
    const char* mesh = argv[1];
    int steps = std::atoi(argv[3]);
    bool exact = std::atoi(argv[4]) != 0;
    const float offset = argc > 6 ? (float)std::atof(argv[6]) : 0.f;
    // the reference seeds with the clock (main.cpp:21); fixed here
    std::srand(argc > 7 ? (unsigned)std::atoi(argv[7]) : 12345u);

    const int resolution = 512, INIT_INTERVAL = 200;
    // ---- PresetPose::init (preset_pose.cpp:7-49)
    Attribute target_pos, target_color;
    Attribute::loadOBJ(mesh, &target_pos, nullptr, nullptr);
    if (!target_pos.vbo) { std::fprintf(stderr, "no mesh\n"); return 2; }
    Attribute::init(target_color, target_pos, 3);
    Attribute::copy(target_color, target_pos);
    Attribute::liner(target_color, 1.f, .5f);

    RotationParams target_rot;
    CameraParams target_cam;
    Rotation::init(target_rot);
    Camera::init(target_cam, target_rot, vec3(0.f, 0.f, 3.5f), vec3(0.f, 0.f, -1.f), vec3(0.f, 1.f, 0.f), .5f, 1.f, 1.f, 10.f);
    ProjectParams target_proj;
    Project::init(target_proj, target_cam.kernel.out, target_pos, true);
    RasterizeParams target_rast;
    Rasterize::init(target_rast, target_proj, resolution, resolution, 1, false);
    InterpolateParams target_intr;
    Interpolate::init(target_intr, target_rast, target_color);
    AntialiasParams target_aa;
    Antialias::init(target_aa, target_rast, target_proj, target_intr.kernel.out, 3);

    AttributeGrad pos, color;
    AttributeGrad::init(pos, target_pos.vboNum, target_pos.vaoNum, 3);
    Attribute::copy(pos, target_pos);
    AttributeGrad::init(color, target_color.vboNum, target_color.vaoNum, 3);
    Attribute::copy(color, target_color);

    RotationGradParams rot;
    rot.exact = exact;
    CameraGradParams cam;
    cam.exact = exact;
    if (exact) SSIMLoss::gradMode() = VCL_SSIM_GRAD_EXACT;
    Rotation::init(rot);
    Camera::init(cam, rot, vec3(0.f, 0.f, 3.5f), vec3(0.f, 0.f, -1.f), vec3(0.f, 1.f, 0.f), .5f, 1.f, 1.f, 10.f);
    ProjectGradParams proj;
    Project::init(proj, cam.kernel.out, cam.grad.out, pos, true);
    RasterizeGradParams rast;
    Rasterize::init(rast, proj, resolution, resolution, 1, false);
    InterpolateGradParams intr;
    Interpolate::init(intr, rast, color);
    AntialiasGradParams aa;
    Antialias::init(aa, rast, proj, intr.kernel.out, intr.gradOut, 3);
    LossParams loss;
    Loss::init(loss, target_aa.kernel.out, aa.kernel.out, aa.gradOut, resolution, resolution, 3);
    Rotation::init(rot, 1e-2, .9f, .999f, 1e-8);
    Window::init(1024, 512);
    GLbuffer gl_target, gl;
    GLbuffer::init(gl_target, target_aa.kernel.out, resolution, resolution, 3);
    GLbuffer::init(gl, aa.kernel.out, resolution, resolution, 3);
    Rotation::setRandom(rot);

    // ---- PresetPose::update (preset_pose.cpp:63-97) for steps 0 .. steps-1
    std::vector<float> losses, quats(8), grad0(4);
    for (int step = 0; step < steps; step++) {
        if (step % INIT_INTERVAL == 0) {
            Rotation::setRandom(target_rot);
            Rotation::forward(target_rot);
            Camera::forward(target_cam);
            Project::forward(target_proj);
            Rasterize::forward(target_rast);
            Interpolate::forward(target_intr);
            Antialias::forward(target_aa);
            Rotation::reset(rot);
            Rotation::setRandom(rot);
            if (offset != 0.f) {
                rot.kernel.q = target_rot.kernel.q;
                Rotation::addRotation(rot, offset, .3f, 1.f, .2f);
            }
        }
        if (step == 0) {
            const quat& a = target_rot.kernel.q;
            const quat& b = rot.kernel.q;
            float q[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
            std::copy(q, q + 8, quats.begin());
        }
        Rotation::forward(rot);
        Camera::forward(cam);
        Project::forward(proj);
        Rasterize::forward(rast);
        Interpolate::forward(intr);
        Antialias::forward(aa);
        SSIMLoss::backward(loss);
        Antialias::backward(aa);
        Interpolate::backward(intr);
        Rasterize::backward(rast);
        Project::backward(proj);
        Camera::backward(cam);
        Rotation::backward(rot);
        if (step == 0) {
            const quat& g = rot.grad.q;
            grad0 = {g.x, g.y, g.z, g.w};
        }
        Rotation::step(rot);
        losses.push_back(Loss::loss(loss));
    }

    // ---- PresetPose::display (preset_pose.cpp:51-61), headless
    Window::clear();
    GLbuffer::draw(gl, GL_RGB32F, GL_RGB, -1.f, -1.f, 0.f, 1.f);
    GLbuffer::draw(gl_target, GL_RGB32F, GL_RGB, 0.f, -1.f, 1.f, 1.f);
    if (!Window::save(argv[5])) { std::fprintf(stderr, "cannot write %s\n", argv[5]); return 2; }

    std::vector<float> img((size_t)resolution * resolution * 3), timg(img.size());
    VCL_HIP_LAYER_CHECK(hipMemcpy(img.data(), aa.kernel.out, img.size() * sizeof(float), hipMemcpyDeviceToHost));
    VCL_HIP_LAYER_CHECK(hipMemcpy(timg.data(), target_aa.kernel.out, timg.size() * sizeof(float), hipMemcpyDeviceToHost));
    FILE* o = std::fopen(argv[2], "wb");
    if (!o) { std::perror("out"); return 2; }
    std::fwrite(losses.data(), sizeof(float), losses.size(), o);
    std::fwrite(quats.data(), sizeof(float), quats.size(), o);
    std::fwrite(img.data(), sizeof(float), img.size(), o);
    std::fwrite(timg.data(), sizeof(float), timg.size(), o);
    const quat& f = rot.kernel.q;
    const float qf[4] = {f.x, f.y, f.z, f.w};
    std::fwrite(qf, sizeof(float), 4, o);
    std::fwrite(grad0.data(), sizeof(float), 4, o);
    std::fclose(o);
    std::printf("preset_pose: %d steps, loss %.6e -> %.6e\n", steps, losses.front(), losses.back());
    return 0;
}
