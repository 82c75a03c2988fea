// engine.hip -- device-resident inverse-rendering fit request (the call chain of
// preset_primitives.cpp:160-207, generalised to vertex positions + mip texture + SG lobes).
// vcl_fit_create allocates every buffer once; vcl_fit_step only launches kernels (no host
// sync, no allocation), so a step can be captured in a HIP graph and replayed.
#include "internal.hpp"
#include <cstring>
#include <string>
#include <vector>
#include <map>

// Per-op HIP-event timing (enabled between vcl_fit_profile_begin/end).
struct OpTimer {
    std::vector<std::string> names;
    std::vector<std::vector<std::pair<hipEvent_t, hipEvent_t>>> ev;  // per op, per launch
    std::vector<size_t> used;
    bool on = false;
    size_t cap = 0;
    std::string only;  // empty: every op
};

struct vcl_fit {
    vcl_fit_desc d;
    OpTimer prof;
    size_t P1;          // pixels per view
    size_t P;           // pixels over all views
    size_t ntex;        // floats over all texture levels
    std::vector<size_t> lvl_off;
    std::map<std::string, void*> buf;
    std::map<std::string, size_t> bytes;
    size_t total = 0;
    void* arena = nullptr;   // every buffer of the request
    bool own_arena = false;  // allocated by vcl_fit_create (else the caller's memory)
    vcl_tuning tune{};       // kernel schedule of this request's steps
    bool has_tune = false;   // false: the process-wide tuning
    int it = 0;         // steps taken (eager or graph)
    size_t ws_rast = 0, ws_sg = 0, ws_loss = 0;
    // HIP graph of one step (vcl_fit_step_graph): captured once on a private stream
    hipStream_t gstream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;

    float* f(const char* n) { return (float*)buf.at(n); }
    template <class T> T* t(const char* n) { return (T*)buf.at(n); }
};

// Buffers live in one device arena (one allocation per request, or caller-provided memory
// from a pool): plan() lays them out, 256-byte aligned.
struct Plan {
    std::vector<std::pair<std::string, size_t>> items;   // name, bytes
    size_t total = 0;
    void add(const char* n, size_t bytes) {
        size_t nb = bytes ? bytes : 4;
        items.push_back({n, nb});
        total += (nb + 255) & ~(size_t)255;
    }
};

static Plan plan(const vcl_fit_desc* d) {
    Plan L;
    size_t P = (size_t)d->width * d->height * d->views, f4 = sizeof(float);
    int C = d->channel, K = d->sgnum, V = d->V, Tn = d->T, nv = d->views;
    size_t ntex = 0;
    for (int l = 0; l < d->miplevel; l++) ntex += (size_t)(d->texwidth >> l) * (d->texheight >> l) * C;
    L.add("pos", (size_t)V * 3 * f4); L.add("normals", (size_t)V * 3 * f4); L.add("uv", (size_t)V * 2 * f4);
    L.add("idx", (size_t)Tn * 3 * 4); L.add("texture", ntex * f4);
    L.add("axis", (size_t)K * 3 * f4); L.add("sharpness", (size_t)K * f4); L.add("amplitude", (size_t)K * C * f4);
    L.add("mats", (size_t)nv * 16 * f4); L.add("pvinv", (size_t)nv * 9 * f4); L.add("rot", 16 * f4);
    L.add("target", P * C * f4); L.add("rough", P * f4);
    // Adam state
    L.add("m_pos", (size_t)V * 3 * f4); L.add("v_pos", (size_t)V * 3 * f4);
    L.add("m_tex", (size_t)d->texwidth * d->texheight * C * f4); L.add("v_tex", (size_t)d->texwidth * d->texheight * C * f4);
    L.add("m_axis", (size_t)K * 3 * f4); L.add("v_axis", (size_t)K * 3 * f4);
    L.add("m_sharp", (size_t)K * f4); L.add("v_sharp", (size_t)K * f4);
    L.add("m_amp", (size_t)K * C * f4); L.add("v_amp", (size_t)K * C * f4);
    L.add("adam_bt", 2 * sizeof(double));
    // intermediates
    L.add("clip", (size_t)nv * V * 4 * f4);
    L.add("rast", P * 4 * f4); L.add("db", P * 4 * f4);
    L.add("uvp", P * 2 * f4); L.add("uvDA", P * 4 * f4);
    L.add("nrm", P * 3 * f4); L.add("view", P * 3 * f4);
    L.add("diff", P * C * f4);
    L.add("sAxis", P * 3 * f4); L.add("sSharp", P * f4); L.add("sAmp", P * f4);
    L.add("color", P * C * f4); L.add("denv", P * C * f4); L.add("senv", P * C * f4);
    L.add("image", P * C * f4); L.add("gimage", P * C * f4); L.add("gcolor", P * C * f4);
    L.add("gdiff", P * C * f4);
    L.add("gtex", ntex * f4);
    L.add("guv", P * 2 * f4);
    L.add("grast", P * 4 * f4);
    L.add("gclip", (size_t)nv * V * 4 * f4);
    L.add("grad_pos", (size_t)V * 3 * f4);
    L.add("grad_axis", (size_t)K * 3 * f4); L.add("grad_sharpness", (size_t)K * f4); L.add("grad_amplitude", (size_t)K * C * f4);
    L.add("loss", 4 * f4);
    L.add("ws_rast", vcl_rasterize_workspace_size(d->width, d->height, nv, V, Tn));
    L.add("ws_sg", vcl_spherical_gaussian_workspace_size(d->width, d->height, nv, C, K));
    // the loss partials: one per 1,024 elements (the MSE pass) or one per antialias tile (the MSE
    // fused into the antialias forward)
    L.add("ws_loss", std::max(vcl_loss_workspace_size((int)(P * C)), vcl_antialias_loss_parts(d->width, d->height, nv) * sizeof(float) + 256));
    return L;
}

static void prof_release(OpTimer& T) {
    for (auto& v : T.ev)
        for (auto& e : v) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    T.ev.clear(); T.names.clear(); T.used.clear(); T.on = false;
}

// Record the start/stop events of op `name` around a launch sequence.
struct ProfScope {
    OpTimer* T; int slot = -1; hipStream_t s;
    ProfScope(vcl_fit* F, const char* name, hipStream_t st) : T(&F->prof), s(st) {
        if (!T->on) return;
        if (!T->only.empty() && T->only != name) return;
        for (size_t i = 0; i < T->names.size(); i++) if (T->names[i] == name) { slot = (int)i; break; }
        if (slot < 0) {
            if (T->names.size() >= 64) return;
            T->names.push_back(name);
            T->ev.emplace_back();
            T->used.push_back(0);
            slot = (int)T->names.size() - 1;
            for (size_t k = 0; k < T->cap; k++) {
                hipEvent_t a, b;
                if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) break;
                T->ev[slot].push_back({a, b});
            }
        }
        if (T->used[slot] >= T->ev[slot].size()) { slot = -1; return; }
        (void)hipEventRecord(T->ev[slot][T->used[slot]].first, s);
    }
    ~ProfScope() {
        if (slot < 0) return;
        (void)hipEventRecord(T->ev[slot][T->used[slot]].second, s);
        T->used[slot]++;
    }
};

extern "C" int vcl_fit_profile_begin(vcl_fit* F, int max_launches, const char* only) {
    VCL_REQUIRE(F && max_launches > 0);
    prof_release(F->prof);
    F->prof.only = only ? only : "";
    F->prof.cap = (size_t)max_launches;
    F->prof.on = true;
    return VCL_OK;
}

extern "C" int vcl_fit_profile_end(vcl_fit* F, int max_ops, int* nops, char* names, int name_len, double* total_ms, int* count) {
    VCL_REQUIRE(F && nops);
    OpTimer& T = F->prof;
    T.on = false;
    int n = (int)T.names.size();
    if (n > max_ops) n = max_ops;
    *nops = n;
    for (int i = 0; i < n; i++) {
        double tot = 0;
        for (size_t k = 0; k < T.used[i]; k++) {
            VCL_HIP_CHECK(hipEventSynchronize(T.ev[i][k].second));
            float ms = 0;
            VCL_HIP_CHECK(hipEventElapsedTime(&ms, T.ev[i][k].first, T.ev[i][k].second));
            tot += ms;
        }
        if (total_ms) total_ms[i] = tot;
        if (count) count[i] = (int)T.used[i];
        if (names && name_len > 0) {
            std::strncpy(names + (size_t)i * name_len, T.names[i].c_str(), name_len - 1);
            names[(size_t)i * name_len + name_len - 1] = 0;
        }
    }
    prof_release(T);
    return VCL_OK;
}

extern "C" int vcl_fit_destroy(vcl_fit* F) {
    if (!F) return VCL_OK;
    prof_release(F->prof);
    if (F->gstream) (void)hipStreamSynchronize(F->gstream);
    if (F->gexec) (void)hipGraphExecDestroy(F->gexec);
    if (F->graph) (void)hipGraphDestroy(F->graph);
    if (F->ev_in) (void)hipEventDestroy(F->ev_in);
    if (F->ev_out) (void)hipEventDestroy(F->ev_out);
    if (F->gstream) (void)hipStreamDestroy(F->gstream);
    if (F->own_arena && F->arena) (void)hipFree(F->arena);
    delete F;
    return VCL_OK;
}

static bool desc_ok(const vcl_fit_desc* d) {
    return d && d->width > 0 && d->height > 0 && d->views >= 1 && d->V > 0 && d->T > 0 && d->channel >= 1 && d->channel <= 4 &&
           d->sgnum >= 0 && d->miplevel >= 1 && d->miplevel <= VCL_TEX_MAX_MIP_LEVEL &&
           d->miplevel == vcl_texture_mip_levels(d->texwidth, d->texheight, d->miplevel);
}

extern "C" size_t vcl_fit_arena_bytes(const vcl_fit_desc* d) { return desc_ok(d) ? plan(d).total : 0; }

extern "C" int vcl_fit_create_in(const vcl_fit_desc* d, void* arena, size_t arena_bytes, vcl_stream_t stream, vcl_fit** out) {
    VCL_REQUIRE(out && desc_ok(d));
    VCL_REQUIRE(d->pos && d->normals && d->uv && d->idx && d->texture && d->mats && d->pvinv && d->rot);
    Plan L = plan(d);
    VCL_REQUIRE(arena && arena_bytes >= L.total);
    hipStream_t hs = as_hip(stream);
    vcl_fit* F = new vcl_fit();
    F->d = *d;
    F->d.target = nullptr;
    F->arena = arena;
    F->P1 = (size_t)d->width * d->height;
    F->P = F->P1 * d->views;
    const int C = d->channel;
    size_t off = 0;
    for (int l = 0; l < d->miplevel; l++) {
        F->lvl_off.push_back(off);
        off += (size_t)(d->texwidth >> l) * (d->texheight >> l) * C;
    }
    F->ntex = off;
    char* base = (char*)arena;
    size_t o = 0;
    for (auto& it : L.items) {
        F->buf[it.first] = base + o;
        F->bytes[it.first] = it.second;
        o += (it.second + 255) & ~(size_t)255;
    }
    F->total = L.total;
    F->ws_rast = F->bytes["ws_rast"]; F->ws_sg = F->bytes["ws_sg"]; F->ws_loss = F->bytes["ws_loss"];
    size_t P = F->P, f4 = sizeof(float);
    int st = vcl_fill_async(arena, L.total, hs);
    auto up = [&](const char* name, const void* host, size_t nbytes) {
        if (!st && host && nbytes && hipMemcpyAsync(F->buf[name], host, nbytes, hipMemcpyHostToDevice, hs) != hipSuccess) st = VCL_ERR_HIP;
    };
    up("pos", d->pos, (size_t)d->V * 3 * f4);
    up("normals", d->normals, (size_t)d->V * 3 * f4);
    up("uv", d->uv, (size_t)d->V * 2 * f4);
    up("idx", d->idx, (size_t)d->T * 3 * 4);
    up("texture", d->texture, F->ntex * f4);
    up("axis", d->axis, (size_t)d->sgnum * 3 * f4);
    up("sharpness", d->sharpness, (size_t)d->sgnum * f4);
    up("amplitude", d->amplitude, (size_t)d->sgnum * C * f4);
    up("mats", d->mats, (size_t)d->views * 16 * f4);
    up("pvinv", d->pvinv, (size_t)d->views * 9 * f4);
    up("rot", d->rot, 16 * f4);
    up("target", d->target, P * C * f4);
    std::vector<float> r(P, d->roughness);
    up("rough", r.data(), P * f4);
    // Adam bias-correction powers (rhomt, rhovt of AdamHyperParams) live on the device so that a
    // captured step advances them on every replay
    double bt[2] = {d->beta1, d->beta2};
    up("adam_bt", bt, sizeof(bt));
    // the host staging (r, bt and the caller's arrays) must outlive the copies
    if (!st && hipStreamSynchronize(hs) != hipSuccess) st = VCL_ERR_HIP;
    if (st) { vcl_fit_destroy(F); return st; }
    *out = F;
    return VCL_OK;
}

extern "C" int vcl_fit_create(const vcl_fit_desc* d, vcl_fit** out) {
    VCL_REQUIRE(out && desc_ok(d));
    size_t nb = plan(d).total;
    void* arena = nullptr;
    if (hipMalloc(&arena, nb) != hipSuccess) return VCL_ERR_OOM;
    vcl_fit* F = nullptr;
    int st = vcl_fit_create_in(d, arena, nb, nullptr, &F);
    if (st) { (void)hipFree(arena); return st; }
    F->own_arena = true;
    *out = F;
    return VCL_OK;
}

extern "C" size_t vcl_fit_device_bytes(const vcl_fit* F) { return F ? F->total : 0; }

extern "C" void* vcl_fit_buffer(vcl_fit* F, const char* name) {
    if (!F || !name) return nullptr;
    if (!std::strcmp(name, "grad_texture")) name = "gtex";
    auto it = F->buf.find(name);
    return it == F->buf.end() ? nullptr : it->second;
}

extern "C" int vcl_fit_read(vcl_fit* F, const char* name, void* host, size_t nbytes, vcl_stream_t s) {
    VCL_REQUIRE(F && name && host);
    void* d = vcl_fit_buffer(F, name);
    VCL_REQUIRE(d);
    if (!std::strcmp(name, "grad_texture")) name = "gtex";
    VCL_REQUIRE(nbytes <= F->bytes[name]);
    VCL_HIP_CHECK(hipMemcpyAsync(host, d, nbytes, hipMemcpyDeviceToHost, as_hip(s)));
    VCL_HIP_CHECK(hipStreamSynchronize(as_hip(s)));
    return VCL_OK;
}

extern "C" int vcl_fit_write(vcl_fit* F, const char* name, const void* host, size_t nbytes, vcl_stream_t s) {
    VCL_REQUIRE(F && name && host);
    void* d = vcl_fit_buffer(F, name);
    VCL_REQUIRE(d);
    if (!std::strcmp(name, "grad_texture")) name = "gtex";
    VCL_REQUIRE(nbytes <= F->bytes[name]);
    VCL_HIP_CHECK(hipMemcpyAsync(d, host, nbytes, hipMemcpyHostToDevice, as_hip(s)));
    VCL_HIP_CHECK(hipStreamSynchronize(as_hip(s)));
    return VCL_OK;
}

static void fill_texture(vcl_fit* F, const char* name, vcl_texture& t) {
    t.width = F->d.texwidth; t.height = F->d.texheight; t.channel = F->d.channel; t.miplevel = F->d.miplevel;
    for (int l = 0; l < VCL_TEX_MAX_MIP_LEVEL; l++) t.texture[l] = l < t.miplevel ? F->f(name) + F->lvl_off[l] : nullptr;
}

// Forward chain up to the antialiased image.
// step: also the step's start work, folded into the forward's first passes: the clip-space grad
// clear of the forward(XGradParams&) overloads, and the Adam bias-correction bump (Adam::step
// bumps rhomt/rhovt before its launch, optimize.cu:109-115), which the updates at the end of
// the step read.  The texture-output grad needs no clear in the fused step: the SG backward
// writes it for every covered pixel, the only pixels the texture backward reads (the op-by-op
// sequence clears it, as the reference does).
static int forward(vcl_fit* F, vcl_stream_t s, bool step = false) {
    const vcl_fit_desc& d = F->d;
    int W = d.width, H = d.height, nv = d.views, C = d.channel;
    int st;
    vcl_rasterize_params rp;
    rp.width = W; rp.height = H; rp.depth = nv; rp.enableDB = 1;
    rp.xs = 2.f / W; rp.ys = 2.f / H;
    rp.proj = F->f("clip"); rp.idx = F->t<uint32_t>("idx");
    rp.out = F->f("rast"); rp.outDB = F->f("db");
    rp.projNum = d.V; rp.idxNum = d.T;
    rp.workspace = F->buf["ws_rast"]; rp.workspace_size = F->ws_rast;
    // in a step, the projection pass also does the rasterizer's workspace reset
    WsClear rc;
    if (step && (st = vcl_rasterize_clear_spec(&rp, &rc))) return st;
    vcl_project_params pj = {d.V, 4, F->f("pos"), F->f("mats"), F->f("clip"), nv};
    { ProfScope ps_(F, "project", as_hip(s));
      st = step ? vcl_project_forward_dev(&pj, F->f("gclip"), F->t<double>("adam_bt"), d.beta1, d.beta2, &rc, s) : vcl_project_forward(&pj, s); }
    if (st) return st;
    // the fused C == 4 path resolves the winning keys in its per-pixel pass (pixel_prep)
    const bool fused = C == 4 && !d.op_by_op;
    const unsigned long long* keys = nullptr;
    { ProfScope ps_(F, "rasterize", as_hip(s)); st = vcl_rasterize_forward_dev(&rp, step ? 1 : 0, s, fused ? &keys : nullptr); }
    if (st) return st;
    vcl_normal_axis_params np;
    std::memset(&np, 0, sizeof(np));
    np.width = W; np.height = H; np.depth = nv; np.rast = F->f("rast"); np.rot = F->f("rot");
    np.normal = F->f("normals"); np.normalidx = F->t<uint32_t>("idx"); np.out = F->f("nrm");
    vcl_sg_specular_params sp = {W, H, nv, F->f("rast"), F->f("nrm"), F->f("view"), F->f("rough"), d.ior, F->f("sAxis"), F->f("sSharp"), F->f("sAmp")};
    vcl_texturemap_params tp;
    tp.width = W; tp.height = H; tp.depth = nv; tp.texwidth = d.texwidth; tp.texheight = d.texheight; tp.channel = C; tp.miplevel = d.miplevel;
    tp.rast = F->f("rast"); tp.uv = F->f("uvp"); tp.uvDA = F->f("uvDA");
    for (int l = 0; l < VCL_TEX_MAX_MIP_LEVEL; l++) tp.texture[l] = l < d.miplevel ? F->f("texture") + F->lvl_off[l] : nullptr;
    tp.out = F->f("diff");
    vcl_interpolate_params ip = {W, H, nv, 2, 1, F->f("rast"), F->f("db"), F->f("uv"), F->t<uint32_t>("idx"), F->f("uvp"), F->f("uvDA"), d.V};
    if (d.op_by_op) {
        // the reference's sequence: NormalAxis, ViewAxis, SGSpecular as separate launches
        { ProfScope ps_(F, "normal_axis", as_hip(s)); st = vcl_normal_axis_forward(&np, s); }
        if (!st) { ProfScope ps_(F, "view_axis", as_hip(s)); st = vcl_view_axis_forward_dev(W, H, nv, F->f("pvinv"), F->f("view"), s); }
        if (!st) { ProfScope ps_(F, "sg_specular", as_hip(s)); st = vcl_sg_specular_forward(&sp, s); }
    } else if (fused) {
        // the raster's resolve, NormalAxis -> ViewAxis -> SGSpecular, the uv interpolation and the
        // texture sample, fused per pixel
        ProfScope ps_(F, "pixel_prep", as_hip(s));
        st = vcl_pixel_prep_dev(&rp, keys, &ip, &tp, &np, F->f("pvinv"), &sp, s);
    } else {
        // NormalAxis -> ViewAxis -> SGSpecular, fused per pixel
        ProfScope ps_(F, "shade_prep", as_hip(s));
        st = vcl_shade_prep_dev(&np, F->f("pvinv"), &sp, s);
    }
    if (st) return st;
    if (!fused) {
        // (the fused C == 4 path did these in the pixel_prep pass)
        { ProfScope ps_(F, "interpolate", as_hip(s)); st = vcl_interpolate_forward(&ip, s); }
        if (st) return st;
        { ProfScope ps_(F, "texture", as_hip(s)); st = vcl_texturemap_forward(&tp, s); }
        if (!st && step) st = vcl_fill_async(F->buf["gdiff"], F->bytes["gdiff"], as_hip(s));
    }
    if (st) return st;
    vcl_spherical_gaussian_params gp;
    gp.width = W; gp.height = H; gp.depth = nv; gp.channel = C; gp.rast = F->f("rast");
    gp.sgnum = d.sgnum; gp.axis = F->f("axis"); gp.sharpness = F->f("sharpness"); gp.amplitude = F->f("amplitude");
    gp.normal = F->f("nrm"); gp.view = F->f("view"); gp.diffuse = F->f("diff");
    gp.specAxis = F->f("sAxis"); gp.specSharpness = F->f("sSharp"); gp.specAmplitude = F->f("sAmp");
    gp.out = F->f("color"); gp.outDiffenv = F->f("denv"); gp.outSpecenv = F->f("senv");
    { ProfScope ps_(F, "spherical_gaussian", as_hip(s)); st = vcl_spherical_gaussian_forward(&gp, s); }
    if (st) return st;
    vcl_antialias_params ap = {W, H, nv, C, {W / 2.f, H / 2.f}, F->f("clip"), F->t<uint32_t>("idx"), F->f("rast"), F->f("color"), F->f("image"), d.V};
    // the fused step's L2 loss grad and tile partials come from the antialias pass (the op-by-op
    // sequence runs the loss as its own pass)
    const bool loss_fused = step && !d.op_by_op;
    { ProfScope ps_(F, "antialias", as_hip(s));
      st = loss_fused ? vcl_antialias_forward_dev(&ap, F->f("target"), F->f("gimage"), F->f("ws_loss"), s) : vcl_antialias_forward(&ap, s); }
    if (st) return st;
    return VCL_OK;
}

extern "C" int vcl_fit_set_target_from_render(vcl_fit* F, vcl_stream_t s) {
    VCL_REQUIRE(F);
    int st = forward(F, s);
    if (st) return st;
    VCL_HIP_CHECK(hipMemcpyAsync(F->buf["target"], F->buf["image"], F->P * F->d.channel * sizeof(float), hipMemcpyDeviceToDevice, as_hip(s)));
    return VCL_OK;
}

// Applies the request's tuning to the launches of one step (restored afterwards).
struct TuneScope {
    vcl_tuning saved;
    bool on;
    explicit TuneScope(const vcl_fit* F) : saved(g_vcl_tuning), on(F->has_tune) { if (on) g_vcl_tuning = F->tune; }
    ~TuneScope() { if (on) g_vcl_tuning = saved; }
};

static int step_body_tuned(vcl_fit* F, vcl_stream_t s);

static int step_body(vcl_fit* F, vcl_stream_t s) {
    TuneScope ts(F);
    return step_body_tuned(F, s);
}

static int step_body_tuned(vcl_fit* F, vcl_stream_t s) {
    const vcl_fit_desc& d = F->d;
    int W = d.width, H = d.height, nv = d.views, C = d.channel, K = d.sgnum;
    int st;
    // the step's output-grad clears and the Adam bias-correction bump run inside the forward's
    // first passes (forward(..., true))
    double* bt = F->t<double>("adam_bt");
    if ((st = forward(F, s, true))) return st;
    vcl_loss_params lp = {F->f("image"), F->f("target"), F->f("gimage"), F->f("ws_loss"), W, H, nv, (int)(F->P * C), F->f("loss")};
    // the loss's final sum runs in the Adam launch at the end of the step, over the antialias
    // pass's tile partials (fused step) or the MSE pass's
    const long long loss_parts = d.op_by_op ? -1 : (long long)vcl_antialias_loss_parts(W, H, nv);
    if (d.op_by_op) {
        { ProfScope ps_(F, "mse_loss", as_hip(s)); st = vcl_mse_loss_grad_dev(&lp, s); }
        if (st) return st;
    }
    // backward
    vcl_antialias_params ap = {W, H, nv, C, {W / 2.f, H / 2.f}, F->f("clip"), F->t<uint32_t>("idx"), F->f("rast"), F->f("color"), F->f("image"), d.V};
    vcl_antialias_grad_params ag = {F->f("gclip"), F->f("gcolor"), F->f("gimage")};
    { ProfScope ps_(F, "antialias_bwd", as_hip(s)); st = vcl_antialias_backward(&ap, &ag, s); }
    if (st) return st;
    vcl_spherical_gaussian_params gp;
    gp.width = W; gp.height = H; gp.depth = nv; gp.channel = C; gp.rast = F->f("rast");
    gp.sgnum = K; gp.axis = F->f("axis"); gp.sharpness = F->f("sharpness"); gp.amplitude = F->f("amplitude");
    gp.normal = F->f("nrm"); gp.view = F->f("view"); gp.diffuse = F->f("diff");
    gp.specAxis = F->f("sAxis"); gp.specSharpness = F->f("sSharp"); gp.specAmplitude = F->f("sAmp");
    gp.out = F->f("color"); gp.outDiffenv = F->f("denv"); gp.outSpecenv = F->f("senv");
    vcl_spherical_gaussian_grad_params gg;
    std::memset(&gg, 0, sizeof(gg));
    gg.diffuse = F->f("gdiff"); gg.axis = F->f("grad_axis"); gg.sharpness = F->f("grad_sharpness"); gg.amplitude = F->f("grad_amplitude");
    gg.out = F->f("gcolor"); gg.workspace = F->buf["ws_sg"]; gg.workspace_size = F->ws_sg;
    // the lobe grads' final sums run in the Adam launch at the end of the step
    // the fused step also leaves the sums' segment stage to the texture backward's launch
    vcl_sg_lobe_sums lobes = {nullptr, 0, 0, 0, d.op_by_op ? 0 : 1, 0, {nullptr, 0, 0, 0, nullptr}};
    // the fused forward leaves the texture-output grad uncleared: the SG backward writes it
    // (the op-by-op forward clears it, as the reference's forward(GradParams&) does)
    { ProfScope ps_(F, "spherical_gaussian_bwd", as_hip(s));
      st = vcl_spherical_gaussian_backward_dev(&gp, &gg, &lobes, (C == 4 && !d.op_by_op) ? 1 : 0, s); }
    if (st) return st;
    vcl_texturemap_params tp;
    tp.width = W; tp.height = H; tp.depth = nv; tp.texwidth = d.texwidth; tp.texheight = d.texheight; tp.channel = C; tp.miplevel = d.miplevel;
    tp.rast = F->f("rast"); tp.uv = F->f("uvp"); tp.uvDA = F->f("uvDA");
    for (int l = 0; l < VCL_TEX_MAX_MIP_LEVEL; l++) tp.texture[l] = l < d.miplevel ? F->f("texture") + F->lvl_off[l] : nullptr;
    tp.out = F->f("diff");
    vcl_texturemap_grad_params tg;
    // the uvDA grad is not requested: the interpolate backward ignores it, as the reference's does
    // (interpolate.cu:114-143), so nothing downstream would read it
    tg.out = F->f("gdiff"); tg.uv = F->f("guv"); tg.uvDA = nullptr; tg.daMode = VCL_TEX_GRAD_DA_REF;
    for (int l = 0; l < VCL_TEX_MAX_MIP_LEVEL; l++) tg.texture[l] = l < d.miplevel ? F->f("gtex") + F->lvl_off[l] : nullptr;
    vcl_interpolate_params ip = {W, H, nv, 2, 1, F->f("rast"), F->f("db"), F->f("uv"), F->t<uint32_t>("idx"), F->f("uvp"), F->f("uvDA"), d.V};
    vcl_interpolate_grad_params ig = {F->f("guv"), nullptr, F->f("grast"), nullptr, nullptr};
    if (d.op_by_op) {
        // the reference's sequence: the texture backward stores the uv grad, the interpolate
        // backward turns it into the rast grad
        { ProfScope ps_(F, "texture_bwd", as_hip(s)); st = vcl_texturemap_backward(&tp, &tg, s); }
        if (!st) { ProfScope ps_(F, "interpolate_bwd", as_hip(s)); st = vcl_interpolate_backward(&ip, &ig, s); }
    } else {
        ProfScope ps_(F, "texture_interp_bwd", as_hip(s));
        st = vcl_texture_interp_backward_dev(&tp, &tg, &ip, &ig, s, &lobes);
    }
    if (st) return st;
    vcl_rasterize_params rp;
    std::memset(&rp, 0, sizeof(rp));
    rp.width = W; rp.height = H; rp.depth = nv; rp.enableDB = 1; rp.xs = 2.f / W; rp.ys = 2.f / H;
    rp.proj = F->f("clip"); rp.idx = F->t<uint32_t>("idx"); rp.out = F->f("rast"); rp.outDB = F->f("db");
    rp.projNum = d.V; rp.idxNum = d.T;
    vcl_rasterize_grad_params rg = {F->f("gclip"), F->f("grast"), nullptr};
    { ProfScope ps_(F, "rasterize_bwd", as_hip(s)); st = vcl_rasterize_backward(&rp, &rg, s); }
    if (st) return st;
    // the project backward (positions' grad from gclip) runs inside the Adam launch below for a
    // few views; with many views each position element's loop over the views is better spread
    // over the project backward's one thread per vertex
    constexpr int FOLD_PROJECT_MAX_VIEWS = 4;
    vcl_project_params pj = {d.V, 4, F->f("pos"), F->f("mats"), F->f("gclip"), nv};
    const bool fold_project = nv <= FOLD_PROJECT_MAX_VIEWS;
    if (!fold_project) {
        vcl_project_params pf = {d.V, 4, F->f("pos"), F->f("mats"), F->f("clip"), nv};
        vcl_project_grad_params pg = {F->f("grad_pos"), nullptr, F->f("gclip")};
        { ProfScope ps_(F, "project_bwd", as_hip(s)); st = vcl_project_backward(&pf, &pg, s); }
        if (st) return st;
    }
    // update
    vcl_adam_hyper_params ht = {d.beta1, d.beta2, 0.0, 0.0, d.lr, d.eps, F->f("m_tex"), F->f("v_tex")};
    vcl_texture tex, gtex;
    fill_texture(F, "texture", tex);
    fill_texture(F, "gtex", gtex);
    { ProfScope ps_(F, "texture_adam", as_hip(s)); st = vcl_texture_adam_update_parts(&tex, &gtex, &ht, 0.f, 1.f, d.fold_quarter, 1 | 4, bt, s); }
    if (st) return st;
    // the mip rebuild's last pass, when it is one workgroup, runs in the Adam launch below
    VclMipTail mtail;
    { ProfScope ps_(F, "texture_mip", as_hip(s)); st = vcl_texture_adam_update_parts(&tex, &gtex, &ht, 0.f, 1.f, d.fold_quarter, 2 | 4, bt, s, &mtail); }
    if (st) return st;
    vcl_optimize_params op[4] = {
        {F->f("pos"), F->f("grad_pos"), d.V * 3, d.V, 3, 1},
        {F->f("axis"), F->f("grad_axis"), K * 3, K, 3, 1},
        {F->f("sharpness"), F->f("grad_sharpness"), K, K, 1, 1},
        {F->f("amplitude"), F->f("grad_amplitude"), K * C, K, C, 1}};
    vcl_adam_hyper_params hp[4] = {
        {d.beta1, d.beta2, 0.0, 0.0, d.lr, d.eps, F->f("m_pos"), F->f("v_pos")},
        {d.beta1, d.beta2, 0.0, 0.0, d.lr, d.eps, F->f("m_axis"), F->f("v_axis")},
        {d.beta1, d.beta2, 0.0, 0.0, d.lr, d.eps, F->f("m_sharp"), F->f("v_sharp")},
        {d.beta1, d.beta2, 0.0, 0.0, d.lr, d.eps, F->f("m_amp"), F->f("v_amp")}};
    // the parameter-grad clears (AttributeGrad::clear, SGBufferGrad::clear) happen inside Adam
    // pj.out: the clip-space grads the folded project backward takes
    { ProfScope ps_(F, "adam", as_hip(s));
      st = vcl_adam_step_multi_proj_dev(K > 0 ? 4 : 1, op, hp, bt, 1, fold_project ? &pj : nullptr, &lp, lobes.seg ? &lobes : nullptr, s,
                                        loss_parts, &mtail); }
    if (st) return st;
    return VCL_OK;
}

extern "C" int vcl_fit_step(vcl_fit* F, vcl_stream_t s) {
    VCL_REQUIRE(F);
    int st = step_body(F, s);
    if (!st) F->it++;
    return st;
}

// Records the launches of one step into rec without running anything.  The step runs with
// the recorder installed, inside a capture of a private stream: groupable launches go to the
// recorder, and any other device work would land in the captured graph, which must stay empty.
int vcl_fit_record_step(vcl_fit* F, VRecorder* rec) {
    VCL_REQUIRE(F && rec && !g_vcl_recorder);
    if (F->prof.on) return VCL_ERR_INVALID;
    thread_local hipStream_t cs = nullptr;
    if (!cs) VCL_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    VCL_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    g_vcl_recorder = rec;
    int st = step_body(F, (vcl_stream_t)cs);
    g_vcl_recorder = nullptr;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(cs, &g);
    size_t nodes = 0;
    if (e == hipSuccess && g) e = hipGraphGetNodes(g, nullptr, &nodes);
    if (g) (void)hipGraphDestroy(g);
    if (st) return st;
    VCL_HIP_CHECK(e);
    return nodes ? VCL_ERR_INVALID : VCL_OK;
}

void vcl_fit_count_steps(vcl_fit* F, int k) { F->it += k; }

// Capture one step (same launches as vcl_fit_step) into a HIP graph on a private stream the
// first time, then replay it.  The caller's stream and the replay are ordered with events.
extern "C" int vcl_fit_step_graph(vcl_fit* F, vcl_stream_t s) {
    VCL_REQUIRE(F);
    if (!F->gexec) {
        VCL_HIP_CHECK(hipStreamCreateWithFlags(&F->gstream, hipStreamNonBlocking));
        VCL_HIP_CHECK(hipEventCreateWithFlags(&F->ev_in, hipEventDisableTiming));
        VCL_HIP_CHECK(hipEventCreateWithFlags(&F->ev_out, hipEventDisableTiming));
        bool prof = F->prof.on;
        F->prof.on = false;
        VCL_HIP_CHECK(hipStreamBeginCapture(F->gstream, hipStreamCaptureModeThreadLocal));
        int st = step_body(F, (vcl_stream_t)F->gstream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(F->gstream, &g);
        F->prof.on = prof;
        if (st) { if (g) (void)hipGraphDestroy(g); return st; }
        VCL_HIP_CHECK(e);
        F->graph = g;
        VCL_HIP_CHECK(hipGraphInstantiate(&F->gexec, F->graph, nullptr, nullptr, 0));
    }
    hipStream_t hs = as_hip(s);
    VCL_HIP_CHECK(hipEventRecord(F->ev_in, hs));
    VCL_HIP_CHECK(hipStreamWaitEvent(F->gstream, F->ev_in, 0));
    VCL_HIP_CHECK(hipGraphLaunch(F->gexec, F->gstream));
    VCL_HIP_CHECK(hipEventRecord(F->ev_out, F->gstream));
    VCL_HIP_CHECK(hipStreamWaitEvent(hs, F->ev_out, 0));
    F->it++;
    return VCL_OK;
}

static void drop_graph(vcl_fit* F) {
    if (F->gstream) (void)hipStreamSynchronize(F->gstream);
    if (F->gexec) { (void)hipGraphExecDestroy(F->gexec); F->gexec = nullptr; }
    if (F->graph) { (void)hipGraphDestroy(F->graph); F->graph = nullptr; }
}

extern "C" int vcl_fit_set_tuning(vcl_fit* F, const vcl_tuning* t) {
    VCL_REQUIRE(F);
    if (t) {
        vcl_tuning saved = g_vcl_tuning;
        int st = vcl_set_tuning(t);   // validates
        g_vcl_tuning = saved;
        if (st) return st;
        F->tune = *t;
    }
    F->has_tune = t != nullptr;
    drop_graph(F);                 // a captured step holds the old launch shapes
    return VCL_OK;
}

extern "C" int vcl_fit_get_tuning(const vcl_fit* F, vcl_tuning* t) {
    VCL_REQUIRE(F && t);
    *t = F->has_tune ? F->tune : g_vcl_tuning;
    return VCL_OK;
}

// Times `steps` steps of each candidate tuning, every candidate starting from the same
// snapshot of the request's persistent state (parameters, Adam moments and bias powers), then
// restores the snapshot and keeps the fastest candidate.  Coordinate descent: raster tile x bin
// chunk first, then the texel-splat rows.
static const char* const kTuneKeep[] = {"pos", "texture", "axis", "sharpness", "amplitude", "m_pos", "v_pos", "m_tex", "v_tex",
                                        "m_axis", "v_axis", "m_sharp", "v_sharp", "m_amp", "v_amp", "adam_bt"};

extern "C" size_t vcl_fit_autotune_scratch_bytes(const vcl_fit* F) {
    if (!F) return 0;
    size_t total = 0;
    for (const char* n : kTuneKeep) total += (F->bytes.at(n) + 255) & ~(size_t)255;
    return total;
}

extern "C" int vcl_fit_autotune_in(vcl_fit* F, int steps, void* scratch, size_t scratch_bytes, vcl_stream_t stream, vcl_tuning* best,
                                   double* best_ms) {
    VCL_REQUIRE(F && steps >= 1 && scratch);
    if (scratch_bytes < vcl_fit_autotune_scratch_bytes(F)) return VCL_ERR_OOM;
    char* snap = (char*)scratch;
    hipStream_t s = as_hip(stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int st = VCL_OK;
    auto copy = [&](bool save) {
        size_t o = 0;
        for (const char* n : kTuneKeep) {
            size_t b = F->bytes.at(n);
            void* dev = F->buf.at(n);
            if (hipMemcpyAsync(save ? (void*)(snap + o) : dev, save ? dev : (void*)(snap + o), b, hipMemcpyDeviceToDevice, s) != hipSuccess)
                st = VCL_ERR_HIP;
            o += (b + 255) & ~(size_t)255;
        }
    };
    const vcl_tuning start = F->has_tune ? F->tune : g_vcl_tuning;
    const bool had = F->has_tune;
    vcl_tuning best_t = start;
    double best_v = 1e30;
    // a candidate replaces the incumbent only when it is at least 3 % faster, so timing noise
    // does not move the schedule away from the starting one
    constexpr double MARGIN = 0.97;
    auto same = [](const vcl_tuning& a, const vcl_tuning& b) {
        return a.raster_tile == b.raster_tile && a.bin_chunk == b.bin_chunk && a.splat_rows == b.splat_rows;
    };
    auto trial = [&](vcl_tuning cand) {
        if (st) return;
        F->tune = cand;
        F->has_tune = true;
        copy(false);
        if (!st) st = step_body(F, (vcl_stream_t)s);   // warm-up (first launch of a variant)
        copy(false);
        if (!st && hipEventRecord(e0, s) != hipSuccess) st = VCL_ERR_HIP;
        for (int i = 0; i < steps && !st; i++) st = step_body(F, (vcl_stream_t)s);
        if (!st && hipEventRecord(e1, s) != hipSuccess) st = VCL_ERR_HIP;
        float ms = 0.f;
        if (!st && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)) st = VCL_ERR_HIP;
        if (!st && (best_v >= 1e30 || ms / steps < MARGIN * best_v)) { best_v = ms / steps; best_t = cand; }
    };
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) st = VCL_ERR_HIP;
    copy(true);
    static const int tiles[] = {8, 16, 32}, chunks[] = {64, 128, 256}, rows[] = {16, 32, 64};
    trial(start);   // the incumbent is timed first
    for (int ti : tiles)
        for (int ch : chunks)
            if (!same(vcl_tuning{ti, ch, start.splat_rows}, start)) trial(vcl_tuning{ti, ch, start.splat_rows});
    const vcl_tuning r1 = best_t;
    for (int r : rows)
        if (r != r1.splat_rows) trial(vcl_tuning{r1.raster_tile, r1.bin_chunk, r});
    copy(false);
    if (!st && hipStreamSynchronize(s) != hipSuccess) st = VCL_ERR_HIP;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) { F->tune = start; F->has_tune = had; return st; }
    F->tune = best_t;
    F->has_tune = true;
    drop_graph(F);
    if (best) *best = best_t;
    if (best_ms) *best_ms = best_v;
    return VCL_OK;
}

// Standalone form: its own snapshot memory (hipMalloc / hipFree) and the null stream after a
// device synchronisation; vcl_fit_autotune_in is the allocation-free form on a caller stream.
extern "C" int vcl_fit_autotune(vcl_fit* F, int steps, vcl_tuning* best, double* best_ms) {
    VCL_REQUIRE(F && steps >= 1);
    size_t bytes = vcl_fit_autotune_scratch_bytes(F);
    void* snap = nullptr;
    if (hipMalloc(&snap, bytes) != hipSuccess) return VCL_ERR_OOM;
    int st = hipDeviceSynchronize() == hipSuccess ? VCL_OK : VCL_ERR_HIP;
    if (!st) st = vcl_fit_autotune_in(F, steps, snap, bytes, nullptr, best, best_ms);
    (void)hipFree(snap);
    return st;
}

